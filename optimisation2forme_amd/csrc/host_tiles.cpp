// Cell tiles of the fused stage kernel: partition and per-tile tables
// (see host_tiles.hpp).
#include "host_tiles.hpp"

#include <algorithm>
#include <cstring>
#include <queue>
#include <tuple>
#include <unordered_map>

namespace o2f {

namespace {

bool tagged_bc(char b) { return b == 'i' || b == 'o' || b == 'f' || b == 'w'; }

// Cells of vertex v's interpolation stencil that its gather reads
// (stencil kinds of compute_stencils).
int stencil_used(int kind, int n) {
    switch (kind) {
        case kStencilWsum: return n;
        case kStencilAvg2: return 2;
        case kStencilCopy: return 1;
        default: return 0;
    }
}
int stencil_used(const HostMesh &hm, int v) { return stencil_used(hm.skind[v], hm.vptr[v + 1] - hm.vptr[v]); }

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

TileGather gather_record(const HostMesh &hm, int v) {
    const TileVertex r = vertex_record(hm, v, -1);
    const int n = r.nkb & 0xffff;
    TileGather g{};
    for (int k = 0; k < kTileSlots; ++k) g.cid[k] = k < n ? hm.scell[r.b + k] : -1;
    g.b = r.b;
    g.nkb = r.nkb;
    return g;
}

// lsten: the vertex's local stencil (its used entries).
TileStress stress_record(const HostMesh &hm, int v, int ls, const uint16_t *lsten) {
    const TileVertex r = vertex_record(hm, v, ls);
    const int used = stencil_used(hm, v);
    TileStress s{};
    for (int k = 0; k < kTileSlots; ++k) {
        s.w[k] = k < used ? hm.sw[r.b + k] : 0.0;
        s.lc[k] = k < used ? lsten[k] : (uint16_t)0xffff;
    }
    s.nkb = r.nkb;
    s.b = r.b;
    s.ls = ls;
    return s;
}

// Sorted unique ids of a list.
void uniq(std::vector<int> &v) {
    std::sort(v.begin(), v.end());
    v.erase(std::unique(v.begin(), v.end()), v.end());
}

}  // namespace

size_t TileSet::lds_layout(int *tq, int *fs) const {
    // four doubles per state vertex, later five per stress vertex; five per
    // gradient cell; eight per edge
    const int vert = std::max(4 * max_vg, 5 * max_vo);
    const int q = (vert + 3) & ~3, f = q + ((5 * max_gd + 3) & ~3);
    if (tq) *tq = q;
    if (fs) *fs = f;
    return sizeof(double) * ((size_t)f + 8 * (size_t)max_e);
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
        h.slot = (int)ts.head.size();
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
        for (int v : Vg) {
            ts.vg.push_back(vertex_record(hm, v, -1));
            ts.vgf.push_back(gather_record(hm, v));
        }
        ts.edge.insert(ts.edge.end(), E.begin(), E.end());
        for (int e : E) {
            ts.en.push_back(hm.enx[e]);
            ts.en.push_back(hm.eny[e]);
        }
        for (int v : Vo) {
            const int ls = (int)ts.lsten.size();
            ts.vo.push_back(vertex_record(hm, v, ls));
            const int b = hm.vptr[v], n = stencil_used(hm, v);
            for (int k = 0; k < n; ++k) ts.lsten.push_back((uint16_t)lg[hm.scell[b + k]]);
            ts.vof.push_back(stress_record(hm, v, ls, ts.lsten.data() + ls));
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
    if (ts.vgf.size() != ts.vg.size() || ts.vof.size() != ts.vo.size() || ts.en.size() != 2 * ts.edge.size())
        return fail(-1, "read records do not match the vertex and edge lists");
    auto same = [](const void *a, const void *b, size_t bytes) { return std::memcmp(a, b, bytes) == 0; };
    for (int t = 0; t < ntiles; ++t) {
        const TileHead &h = ts.head[t];
        if (h.nO < 1 || h.nO > h.nF || h.nF > h.nGd || h.cell0 < 0 || h.cell0 + h.nGd > (int)ts.cell.size() ||
            h.vg0 < 0 || h.vg0 + h.nVg > (int)ts.vg.size() || h.e0 < 0 || h.e0 + h.nE > (int)ts.edge.size() ||
            h.vo0 < 0 || h.vo0 + h.nVo > (int)ts.vo.size() || (int)ts.own.size() != (int)ts.cell.size())
            return fail(t, "counts or offsets out of range");
        if (h.member != 0 || h.slot != t) return fail(t, "member or slot differs");
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
        for (int i = 0; i < h.nVg; ++i) {
            if (!vertex_ok(Vg[i]) || Vg[i].ls != -1) return fail(t, "state vertex record differs");
            const TileGather want = gather_record(hm, Vg[i].v);
            if (!same(&want, &ts.vgf[h.vg0 + i], sizeof want)) return fail(t, "state vertex gather record differs");
        }
        for (int i = 0; i < h.nVo; ++i)
            if (!vertex_ok(Vo[i])) return fail(t, "stress vertex record differs");
        for (int i = 0; i < h.nE; ++i) {
            const int e = ts.edge[h.e0 + i];
            if (e < 0 || e >= hm.ne || !same(&ts.en[2 * (h.e0 + i)], &hm.enx[e], sizeof(double)) ||
                !same(&ts.en[2 * (h.e0 + i) + 1], &hm.eny[e], sizeof(double)))
                return fail(t, "edge normal differs");
        }
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
            const TileStress want = stress_record(hm, v, off, ts.lsten.data() + off);
            if (!same(&want, &ts.vof[h.vo0 + i], sizeof want)) return fail(t, "stress vertex read record differs");
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

void concat_tiles(const std::vector<TileSet> &sets, TileSet &all) {
    all = TileSet();
    for (int m = 0; m < (int)sets.size(); ++m) {
        const TileSet &ts = sets[m];
        const int cb = (int)all.cell.size(), vgb = (int)all.vg.size(), eb = (int)all.edge.size(),
                  vob = (int)all.vo.size(), lsb = (int)all.lsten.size();
        for (TileHead h : ts.head) {
            h.cell0 += cb;
            h.vg0 += vgb;
            h.e0 += eb;
            h.vo0 += vob;
            h.member = m;
            all.head.push_back(h);
        }
        all.cell.insert(all.cell.end(), ts.cell.begin(), ts.cell.end());
        all.own.insert(all.own.end(), ts.own.begin(), ts.own.end());
        all.vg.insert(all.vg.end(), ts.vg.begin(), ts.vg.end());
        all.vgf.insert(all.vgf.end(), ts.vgf.begin(), ts.vgf.end());
        all.edge.insert(all.edge.end(), ts.edge.begin(), ts.edge.end());
        all.en.insert(all.en.end(), ts.en.begin(), ts.en.end());
        for (TileVertex r : ts.vo) {
            r.ls += lsb;
            all.vo.push_back(r);
        }
        for (TileStress r : ts.vof) {
            r.ls += lsb;
            all.vof.push_back(r);
        }
        all.lsten.insert(all.lsten.end(), ts.lsten.begin(), ts.lsten.end());
        all.target = std::max(all.target, ts.target);
        all.max_vg = std::max(all.max_vg, ts.max_vg);
        all.max_vo = std::max(all.max_vo, ts.max_vo);
        all.max_gd = std::max(all.max_gd, ts.max_gd);
        all.max_e = std::max(all.max_e, ts.max_e);
        all.max_o = std::max(all.max_o, ts.max_o);
        all.sum_gd += ts.sum_gd;
        all.sum_f += ts.sum_f;
        all.sum_vg += ts.sum_vg;
        all.sum_vo += ts.sum_vo;
        all.sum_e += ts.sum_e;
    }
}

bool check_concat(const std::vector<TileSet> &sets, const TileSet &all, std::string &err) {
    auto fail = [&](int k, const std::string &what) {
        err = "batch tile " + std::to_string(k) + ": " + what;
        return false;
    };
    auto same = [](const void *a, const void *b, size_t bytes) { return bytes == 0 || std::memcmp(a, b, bytes) == 0; };
    size_t k = 0;
    for (int m = 0; m < (int)sets.size(); ++m) {
        const TileSet &ts = sets[m];
        if (ts.max_vg > all.max_vg || ts.max_vo > all.max_vo || ts.max_gd > all.max_gd || ts.max_e > all.max_e)
            return fail((int)k, "LDS maxima below a member's");
        for (int t = 0; t < (int)ts.head.size(); ++t, ++k) {
            if (k >= all.head.size()) return fail((int)k, "missing");
            const TileHead &a = all.head[k], &h = ts.head[t];
            if (a.member != m || a.slot != t || a.nO != h.nO || a.nF != h.nF || a.nGd != h.nGd ||
                a.nVg != h.nVg || a.nE != h.nE || a.nVo != h.nVo)
                return fail((int)k, "head differs from its member's tile");
            if (a.cell0 < 0 || a.cell0 + a.nGd > (int)all.cell.size() || a.vg0 < 0 ||
                a.vg0 + a.nVg > (int)all.vg.size() || a.e0 < 0 || a.e0 + a.nE > (int)all.edge.size() ||
                a.vo0 < 0 || a.vo0 + a.nVo > (int)all.vo.size() || all.own.size() != all.cell.size())
                return fail((int)k, "offsets out of range");
            if (all.vgf.size() != all.vg.size() || all.vof.size() != all.vo.size() ||
                all.en.size() != 2 * all.edge.size())
                return fail((int)k, "read records do not match the vertex and edge lists");
            if (!same(&all.cell[a.cell0], &ts.cell[h.cell0], h.nGd * sizeof(TileCell)) ||
                !same(&all.own[a.cell0], &ts.own[h.cell0], h.nGd * sizeof(TileOwn)) ||
                !same(&all.vg[a.vg0], &ts.vg[h.vg0], h.nVg * sizeof(TileVertex)) ||
                !same(&all.vgf[a.vg0], &ts.vgf[h.vg0], h.nVg * sizeof(TileGather)) ||
                !same(&all.edge[a.e0], &ts.edge[h.e0], h.nE * sizeof(int)) ||
                !same(&all.en[2 * a.e0], &ts.en[2 * h.e0], 2 * h.nE * sizeof(double)))
                return fail((int)k, "records differ from its member's tile");
            for (int i = 0; i < h.nVo; ++i) {
                const TileVertex &x = all.vo[a.vo0 + i], &y = ts.vo[h.vo0 + i];
                if (x.v != y.v || x.b != y.b || x.nkb != y.nkb) return fail((int)k, "stress vertex differs");
                TileStress fy = ts.vof[h.vo0 + i];
                fy.ls = x.ls;
                if (!same(&all.vof[a.vo0 + i], &fy, sizeof fy))
                    return fail((int)k, "stress vertex read record differs");
                const int n = stencil_used((y.nkb >> 16) & 0xff, y.nkb & 0xffff);
                if (x.ls < 0 || x.ls + n > (int)all.lsten.size() ||
                    !same(&all.lsten[x.ls], &ts.lsten[y.ls], n * sizeof(uint16_t)))
                    return fail((int)k, "local stencil differs");
            }
        }
    }
    if (k != all.head.size()) return fail((int)k, "extra tiles");
    return true;
}

}  // namespace o2f
