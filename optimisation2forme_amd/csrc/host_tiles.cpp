// Cell tiles of the fused stage kernel: partition and per-tile tables
// (see host_tiles.hpp).
#include "host_tiles.hpp"

#include <algorithm>
#include <queue>
#include <tuple>
#include <unordered_map>

namespace o2f {

namespace {

bool tagged_bc(char b) { return b == 'i' || b == 'o' || b == 'f' || b == 'w'; }

// Cells of vertex v's interpolation stencil that its gather reads
// (stencil kinds of compute_stencils).
int stencil_used(const HostMesh &hm, int v) {
    const int n = hm.vptr[v + 1] - hm.vptr[v];
    switch (hm.skind[v]) {
        case kStencilWsum: return n;
        case kStencilAvg2: return 2;
        case kStencilCopy: return 1;
        default: return 0;
    }
}

// Greedy growth of compact tiles (see build_tiles).
std::vector<int> grow_tiles(const HostMesh &hm, int target, int &ntiles) {
    const int nc = hm.nc;
    std::vector<int> tile(nc, -1), score(nc, 0), depth(nc, 0), stamp(nc, -1), vstamp(hm.nv, -1);
    // candidates: most shared vertices first, then the smallest depth, then the lowest id
    typedef std::tuple<int, int, int> Key;  // (score, -depth, -id)
    std::priority_queue<Key> heap;
    int t = 0, seed = 0;
    for (;;) {
        while (seed < nc && tile[seed] >= 0) ++seed;
        if (seed == nc) break;
        heap = std::priority_queue<Key>();
        stamp[seed] = t;
        score[seed] = 0;
        depth[seed] = 0;
        heap.emplace(0, 0, -seed);
        int count = 0;
        while (count < target && !heap.empty()) {
            const Key k = heap.top();
            heap.pop();
            const int c = -std::get<2>(k);
            if (tile[c] >= 0 || std::get<0>(k) != score[c] || -std::get<1>(k) != depth[c]) continue;
            tile[c] = t;
            ++count;
            for (int i = 0; i < hm.cnv[c]; ++i) {
                const int v = hm.cvtx[4 * c + i];
                if (vstamp[v] == t) continue;
                vstamp[v] = t;
                for (int j = hm.vptr[v]; j < hm.vptr[v + 1]; ++j) {
                    const int d = hm.vcell_asc[j];
                    if (tile[d] >= 0) continue;
                    if (stamp[d] != t) {
                        stamp[d] = t;
                        score[d] = 0;
                        depth[d] = depth[c] + 1;
                    }
                    ++score[d];
                    depth[d] = std::min(depth[d], depth[c] + 1);
                    heap.emplace(score[d], -depth[d], -d);
                }
            }
        }
        ++t;
    }
    ntiles = t;
    return tile;
}

TileVertex vertex_record(const HostMesh &hm, int v, int ls) {
    const int b = hm.vptr[v], n = hm.vptr[v + 1] - b;
    return TileVertex{v, b, n | (int)hm.skind[v] << 16 | (int)(uint8_t)hm.vbc[v] << 24, ls};
}

// Sorted unique ids of a list.
void uniq(std::vector<int> &v) {
    std::sort(v.begin(), v.end());
    v.erase(std::unique(v.begin(), v.end()), v.end());
}

}  // namespace

size_t TileSet::lds_bytes() const {
    const size_t vert = std::max<size_t>(32 * (size_t)max_vg, 40 * (size_t)max_vo);
    return vert + 40 * (size_t)max_gd + 64 * (size_t)max_e;
}

bool build_tiles(const HostMesh &hm, int target, TileSet &ts, std::string &err) {
    ts = TileSet();
    ts.target = target;
    if (target < 1) {
        err = "tile size must be positive";
        return false;
    }
    for (int e = 0; e < hm.ne; ++e)
        if (hm.ecr[e] < 0 && !tagged_bc(hm.ebc[e])) {
            err = "boundary edge " + std::to_string(e) + " has no boundary tag";
            return false;
        }
    int ntiles = 0;
    const std::vector<int> tile = grow_tiles(hm, target, ntiles);
    std::vector<std::vector<int>> owned(ntiles);
    for (int c = 0; c < hm.nc; ++c) owned[tile[c]].push_back(c);  // ascending id

    std::vector<int> lg(hm.nc, -1), lvg(hm.nv, -1), lvo(hm.nv, -1), lle(hm.ne, -1);
    for (int t = 0; t < ntiles; ++t) {
        const std::vector<int> &O = owned[t];
        std::vector<int> Fx, Vo, Q, E;
        for (int c : O) {
            for (int j = 0; j < 4; ++j) {
                const int nb = hm.cnbr[4 * c + j];
                if (nb >= 0 && tile[nb] != t) Fx.push_back(nb);
                const int e = hm.cedge[4 * c + j];
                if (e >= 0) E.push_back(e);
            }
            for (int j = 0; j < 8; ++j)
                if (hm.cevx[8 * c + j] >= 0) Vo.push_back(hm.cevx[8 * c + j]);
        }
        uniq(Fx);
        uniq(Vo);
        uniq(E);
        for (int v : Vo) {
            const int b = hm.vptr[v], n = stencil_used(hm, v);
            for (int k = 0; k < n; ++k) Q.push_back(hm.scell[b + k]);
        }
        uniq(Q);
        // Gd = [O | F \ O | Q \ F]
        std::vector<int> Gd(O);
        Gd.insert(Gd.end(), Fx.begin(), Fx.end());
        const int nF = (int)Gd.size();
        for (int i = 0; i < nF; ++i) lg[Gd[i]] = i;
        for (int c : Q)
            if (lg[c] < 0) {
                lg[c] = (int)Gd.size();
                Gd.push_back(c);
            }
        std::vector<int> Vg;
        for (int c : Gd)
            for (int j = 0; j < 8; ++j)
                if (hm.cevx[8 * c + j] >= 0) Vg.push_back(hm.cevx[8 * c + j]);
        uniq(Vg);
        if (Gd.size() > 0x7fff || Vg.size() >= 0xffff || Vo.size() >= 0xffff || E.size() > 0x7fff) {
            err = "tile " + std::to_string(t) + " does not fit the 16-bit local indices";
            return false;
        }
        for (size_t i = 0; i < Vg.size(); ++i) lvg[Vg[i]] = (int)i;
        for (size_t i = 0; i < Vo.size(); ++i) lvo[Vo[i]] = (int)i;
        for (size_t i = 0; i < E.size(); ++i) lle[E[i]] = (int)i;

        TileHead h{};
        h.cell0 = (int)ts.cell.size();
        h.nO = (int)O.size();
        h.nF = nF;
        h.nGd = (int)Gd.size();
        h.vg0 = (int)ts.vg.size();
        h.nVg = (int)Vg.size();
        h.e0 = (int)ts.edge.size();
        h.nE = (int)E.size();
        h.vo0 = (int)ts.vo.size();
        h.nVo = (int)Vo.size();
        for (size_t i = 0; i < Gd.size(); ++i) {
            const int c = Gd[i];
            TileCell r{};
            r.c = c;
            r.info = hm.cinfo[c];
            for (int j = 0; j < 4; ++j) {
                r.edge[j] = hm.cedge[4 * c + j];
                r.nbr[j] = hm.cnbr[4 * c + j];
                r.le[j] = (int16_t)(r.edge[j] >= 0 ? lle[r.edge[j]] : -1);
            }
            for (int j = 0; j < 8; ++j) {
                const int v = hm.cevx[8 * c + j];
                r.lv[j] = (uint16_t)(v >= 0 ? lvg[v] : 0xffff);
            }
            TileOwn o{};
            for (int j = 0; j < 8; ++j) {
                const int v = hm.cevx[8 * c + j];
                o.lo[j] = (uint16_t)((i < O.size() && v >= 0) ? lvo[v] : 0xffff);
            }
            ts.cell.push_back(r);
            ts.own.push_back(o);
        }
        for (int v : Vg) ts.vg.push_back(vertex_record(hm, v, -1));
        ts.edge.insert(ts.edge.end(), E.begin(), E.end());
        for (int v : Vo) {
            ts.vo.push_back(vertex_record(hm, v, (int)ts.lsten.size()));
            const int b = hm.vptr[v], n = stencil_used(hm, v);
            for (int k = 0; k < n; ++k) ts.lsten.push_back((uint16_t)lg[hm.scell[b + k]]);
        }
        ts.head.push_back(h);
        ts.max_o = std::max(ts.max_o, h.nO);
        ts.max_gd = std::max(ts.max_gd, h.nGd);
        ts.max_vg = std::max(ts.max_vg, h.nVg);
        ts.max_vo = std::max(ts.max_vo, h.nVo);
        ts.max_e = std::max(ts.max_e, h.nE);
        ts.sum_gd += h.nGd;
        ts.sum_f += h.nF;
        ts.sum_vg += h.nVg;
        ts.sum_vo += h.nVo;
        ts.sum_e += h.nE;
        for (int c : Gd) lg[c] = -1;
        for (int v : Vg) lvg[v] = -1;
        for (int v : Vo) lvo[v] = -1;
        for (int e : E) lle[e] = -1;
    }
    return check_tiles(hm, ts, err);
}

bool check_tiles(const HostMesh &hm, const TileSet &ts, std::string &err) {
    auto fail = [&](int t, const std::string &what) {
        err = "tile " + std::to_string(t) + ": " + what;
        return false;
    };
    std::vector<int> owner(hm.nc, -1);
    const int ntiles = (int)ts.head.size();
    for (int t = 0; t < ntiles; ++t) {
        const TileHead &h = ts.head[t];
        if (h.nO < 1 || h.nO > h.nF || h.nF > h.nGd || h.cell0 < 0 || h.cell0 + h.nGd > (int)ts.cell.size() ||
            h.vg0 < 0 || h.vg0 + h.nVg > (int)ts.vg.size() || h.e0 < 0 || h.e0 + h.nE > (int)ts.edge.size() ||
            h.vo0 < 0 || h.vo0 + h.nVo > (int)ts.vo.size() || (int)ts.own.size() != (int)ts.cell.size())
            return fail(t, "counts or offsets out of range");
        if (h.nGd > ts.max_gd || h.nVg > ts.max_vg || h.nVo > ts.max_vo || h.nE > ts.max_e || h.nO > ts.max_o)
            return fail(t, "a count exceeds the LDS maxima");
        const TileCell *C = &ts.cell[h.cell0];
        const TileOwn *Ow = &ts.own[h.cell0];
        const TileVertex *Vg = &ts.vg[h.vg0], *Vo = &ts.vo[h.vo0];
        const int *E = &ts.edge[h.e0];
        // a vertex record carries the vertex's own stencil and tag
        auto vertex_ok = [&](const TileVertex &r) {
            const TileVertex want = vertex_record(hm, r.v, r.ls);
            return r.v >= 0 && r.v < hm.nv && r.b == want.b && r.nkb == want.nkb;
        };
        for (int i = 0; i < h.nVg; ++i)
            if (!vertex_ok(Vg[i]) || Vg[i].ls != -1) return fail(t, "state vertex record differs");
        for (int i = 0; i < h.nVo; ++i)
            if (!vertex_ok(Vo[i])) return fail(t, "stress vertex record differs");
        std::unordered_map<int, int> gd_of, e_of;
        for (int i = 0; i < h.nGd; ++i) {
            const int c = C[i].c;
            if (c < 0 || c >= hm.nc) return fail(t, "cell id out of range");
            if (!gd_of.emplace(c, i).second) return fail(t, "cell listed twice");
            if (i < h.nO) {
                if (owner[c] >= 0) return fail(t, "cell owned by two tiles");
                owner[c] = t;
            }
            if (C[i].info != hm.cinfo[c]) return fail(t, "cell info differs");
            const int nk = C[i].info >> 4;
            for (int j = 0; j < 4; ++j) {
                if (C[i].edge[j] != hm.cedge[4 * c + j] || C[i].nbr[j] != hm.cnbr[4 * c + j])
                    return fail(t, "edge or neighbour list differs");
                if (j < nk) {
                    for (int s = 0; s < 2; ++s) {
                        const int lv = C[i].lv[2 * j + s];
                        if (lv >= h.nVg || Vg[lv].v != hm.cevx[8 * c + 2 * j + s])
                            return fail(t, "state vertex of a gradient cell missing");
                    }
                } else if (C[i].lv[2 * j] != 0xffff || C[i].lv[2 * j + 1] != 0xffff || C[i].le[j] != -1) {
                    return fail(t, "padding slot not marked");
                }
            }
        }
        for (int i = 0; i < h.nE; ++i) {
            const int e = E[i];
            if (e < 0 || e >= hm.ne || !e_of.emplace(e, i).second) return fail(t, "edge id invalid or repeated");
        }
        for (int i = 0; i < h.nGd; ++i) {
            const int c = C[i].c, nk = C[i].info >> 4;
            for (int j = 0; j < nk; ++j) {
                const auto it = e_of.find(hm.cedge[4 * c + j]);
                const int want = (it == e_of.end()) ? -1 : it->second;
                if (C[i].le[j] != want) return fail(t, "local edge of a slot differs");
                if (i < h.nO && want < 0) return fail(t, "edge of an owned cell missing");
                if (want >= 0 && i >= h.nF) return fail(t, "a tile edge's cell is not a face cell");
            }
        }
        // both sides of every tile edge are face cells (or a tagged boundary ghost)
        for (int i = 0; i < h.nE; ++i) {
            const int e = E[i];
            const auto l = gd_of.find(hm.ecl[e]);
            if (l == gd_of.end() || l->second >= h.nF) return fail(t, "left cell of a tile edge missing");
            if (hm.ecr[e] >= 0) {
                const auto r = gd_of.find(hm.ecr[e]);
                if (r == gd_of.end() || r->second >= h.nF) return fail(t, "right cell of a tile edge missing");
            } else if (!tagged_bc(hm.ebc[e])) {
                return fail(t, "untagged boundary edge");
            }
        }
        // stress vertices: the owned cells' vertices, with their stencils local
        std::unordered_map<int, int> vo_of;
        for (int i = 0; i < h.nVo; ++i) {
            const int v = Vo[i].v, off = Vo[i].ls;
            if (!vo_of.emplace(v, i).second) return fail(t, "stress vertex listed twice");
            const int b = hm.vptr[v], n = stencil_used(hm, v);
            if (off < 0 || off + n > (int)ts.lsten.size()) return fail(t, "stencil offset out of range");
            for (int k = 0; k < n; ++k) {
                const int li = ts.lsten[off + k];
                if (li >= h.nGd || C[li].c != hm.scell[b + k]) return fail(t, "stencil cell not local");
            }
        }
        for (int i = 0; i < h.nO; ++i) {
            const int c = C[i].c, nk = C[i].info >> 4;
            for (int j = 0; j < 2 * nk; ++j) {
                const int lo = Ow[i].lo[j];
                if (lo >= h.nVo || Vo[lo].v != hm.cevx[8 * c + j]) return fail(t, "stress vertex of a slot missing");
            }
        }
    }
    for (int c = 0; c < hm.nc; ++c)
        if (owner[c] < 0) return fail(-1, "cell " + std::to_string(c) + " owned by no tile");
    return true;
}

}  // namespace o2f
