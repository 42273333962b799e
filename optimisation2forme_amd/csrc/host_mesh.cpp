// Host-side mesh preparation: topology, geometry, stencils, msh2 reader.
// Compiled with -ffp-contract=off: every expression rounds exactly as the
// reference's C++ does, so the device sees bit-identical geometry.
#include "host_mesh.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <fstream>
#include <map>
#include <numeric>
#include <sstream>
#include <utility>

namespace o2f {

namespace {

inline bool tagged(char b) { return b == 'i' || b == 'o' || b == 'f' || b == 'w'; }

// Area and centre of one cell (triangle: src/mesh.cpp:123-130; quad with the
// two-triangle area of :158-163, exact only for parallelograms), from the
// vertex positions vx / vy.
void cell_geom_xy(const HostMesh &hm, const double *vx, const double *vy, int c, double &vol, double &cx,
                  double &cy) {
    const int *v = &hm.cvtx[4 * c];
    const double x0 = vx[v[0]], y0 = vy[v[0]];
    const double x1 = vx[v[1]], y1 = vy[v[1]];
    const double x2 = vx[v[2]], y2 = vy[v[2]];
    if (hm.cnv[c] == 3) {
        const double e1x = x1 - x0, e1y = y1 - y0, e2x = x2 - x0, e2y = y2 - y0;
        vol = std::abs(0.5 * (e1x * e2y - e2x * e1y));
        cx = 1.0 / 3.0 * (x0 + x1 + x2);
        cy = 1.0 / 3.0 * (y0 + y1 + y2);
    } else {
        const double x3 = vx[v[3]], y3 = vy[v[3]];
        const double e1x = x1 - x0, e1y = y1 - y0, e2x = x2 - x0, e2y = y2 - y0;
        const double f1x = x1 - x2, f1y = y1 - y2, f2x = x3 - x2, f2y = y3 - y2;
        vol = 0.5 * (std::abs(e1x * e2y - e2x * e1y) + std::abs(f1x * f2y - f2x * f1y));
        cx = 0.25 * (x0 + x1 + x2 + x3);
        cy = 0.25 * (y0 + y1 + y2 + y3);
    }
}
void cell_geom(const HostMesh &hm, int c, double &vol, double &cx, double &cy) {
    cell_geom_xy(hm, hm.vx.data(), hm.vy.data(), c, vol, cx, cy);
}

// Length, unit normal turned towards the left cell centre, midpoint
// (src/mesh.cpp:199-211), from the vertex positions and cell centres given:
// out = (length, n_x, n_y, mid_x, mid_y).
void edge_geom_xy(const HostMesh &hm, const double *vx, const double *vy, const double *ccx, const double *ccy,
                  int e, double *out) {
    const int a = hm.eva[e], b = hm.evb[e];
    const double tx = vx[a] - vx[b], ty = vy[a] - vy[b], tz = 0.0;
    const double len = std::sqrt(tx * tx + ty * ty + tz * tz);
    double nx = ty / len, ny = -tx / len;
    const double mx = 0.5 * (vx[a] + vx[b]), my = 0.5 * (vy[a] + vy[b]);
    const int cl = hm.ecl[e];
    const double to_cell_x = ccx[cl] - mx, to_cell_y = ccy[cl] - my;
    if (to_cell_x * nx + to_cell_y * ny < 0) {
        nx = -nx;
        ny = -ny;
    }
    out[0] = len;
    out[1] = nx;
    out[2] = ny;
    out[3] = mx;
    out[4] = my;
}
void edge_geom(HostMesh &hm, int e) {
    double o[5];
    edge_geom_xy(hm, hm.vx.data(), hm.vy.data(), hm.ccx.data(), hm.ccy.data(), e, o);
    hm.elen[e] = o[0];
    hm.enx[e] = o[1];
    hm.eny[e] = o[2];
    hm.emx[e] = o[3];
    hm.emy[e] = o[4];
}

// The weights of vertex v's weighted stencil (n = 3: barycentric; n = 4,
// or n >= 5 in mvc mode: mean value) from the vertex position and the cell
// centres given, into w[0, n) in the stencil's slot order; false when the
// vertex has no weighted stencil.  tmp: scratch of 4 n doubles.
bool stencil_weights(const HostMesh &hm, int interp_mode, int v, double px, double py, const double *ccx,
                     const double *ccy, double *w, std::vector<double> &tmp) {
    const int b = hm.vptr[v], n = hm.vptr[v + 1] - b;
    if (n == 3) {
        // barycentric weights of the vertex in the triangle of the three
        // cell centres (src/advection.cpp:29-50)
        const int i0 = hm.vcell_asc[b], i1 = hm.vcell_asc[b + 1], i2 = hm.vcell_asc[b + 2];
        const double ax = ccx[i0], ay = ccy[i0], bx = ccx[i1], by = ccy[i1];
        const double qx = ccx[i2], qy = ccy[i2];
        const double num0 = (bx - px) * (qy - py) - (qx - px) * (by - py);
        const double den = (bx - ax) * (qy - ay) - (qx - ax) * (by - ay);
        const double w0 = (num0 / den);
        const double num1 = (qx - px) * (ay - py) - (ax - px) * (qy - py);
        const double w1 = (num1 / den);
        w[0] = w0;
        w[1] = w1;
        w[2] = 1 - w0 - w1;
        return true;
    }
    if (n == 4 || (n >= 5 && interp_mode == O2F_INTERP_MVC)) {
        // mean-value weights over the angularly ordered fan
        // (src/advection.cpp:52-104); n >= 5 is the generalisation
        tmp.assign(4 * static_cast<size_t>(n), 0.0);
        double *ww = tmp.data(), *dist = ww + n, *ux = dist + n, *uy = ux + n;
        for (int k = 0; k < n; ++k) {
            const int c = hm.vcell_ang[b + k];
            const double dx = ccx[c] - px, dy = ccy[c] - py;
            dist[k] = std::hypot(dx, dy);
            ux[k] = dx / dist[k];
            uy[k] = dy / dist[k];
        }
        for (int k = 0; k < n; ++k) {
            const int prev = (k + n - 1) % n, next = (k + 1) % n;
            const double cp = std::clamp(ux[prev] * ux[k] + uy[prev] * uy[k], -1.0, 1.0);
            const double cn = std::clamp(ux[k] * ux[next] + uy[k] * uy[next], -1.0, 1.0);
            ww[k] = (std::tan(0.5 * std::acos(cp)) + std::tan(0.5 * std::acos(cn))) / dist[k];
        }
        const double wsum = std::accumulate(ww, ww + n, 0.0);
        for (int k = 0; k < n; ++k) w[k] = ww[k] / wsum;
        return true;
    }
    return false;
}

inline uint64_t pair_key(int a, int b) {
    const uint32_t lo = static_cast<uint32_t>(std::min(a, b));
    const uint32_t hi = static_cast<uint32_t>(std::max(a, b));
    return (static_cast<uint64_t>(lo) << 32) | hi;
}

}  // namespace

bool build_host_mesh(const o2f_mesh_desc &d, int interp_mode, HostMesh &hm, std::string &err) {
    if (d.n_vertices <= 0 || d.n_cells <= 0 || !d.xy || !d.cell_nv || !d.cell_vtx ||
        (d.n_bnd > 0 && (!d.bnd_vtx || !d.bnd_tag))) {
        err = "o2f_ctx_create: empty or incomplete mesh description";
        return false;
    }
    hm = HostMesh();
    hm.nv = d.n_vertices;
    hm.nc = d.n_cells;
    hm.vx.resize(hm.nv);
    hm.vy.resize(hm.nv);
    hm.vbc.assign(hm.nv, '\0');
    for (int i = 0; i < hm.nv; ++i) {
        hm.vx[i] = d.xy[2 * i];
        hm.vy[i] = d.xy[2 * i + 1];
    }
    // Boundary tags: the last line element wins for vertices and edges.
    std::map<uint64_t, char> edge_tag;
    for (int j = 0; j < d.n_bnd; ++j) {
        const int a = d.bnd_vtx[2 * j], b = d.bnd_vtx[2 * j + 1];
        if (a < 0 || b < 0 || a >= hm.nv || b >= hm.nv) {
            err = "o2f_ctx_create: boundary line references a missing vertex";
            return false;
        }
        hm.vbc[a] = d.bnd_tag[j];
        hm.vbc[b] = d.bnd_tag[j];
        edge_tag[pair_key(a, b)] = d.bnd_tag[j];
    }
    hm.cnv.resize(hm.nc);
    hm.cvtx.assign(4 * hm.nc, -1);
    hm.cvol.resize(hm.nc);
    hm.ccx.resize(hm.nc);
    hm.ccy.resize(hm.nc);
    for (int c = 0; c < hm.nc; ++c) {
        const int n = d.cell_nv[c];
        if (n != 3 && n != 4) {
            err = "o2f_ctx_create: cells must be triangles or quadrilaterals";
            return false;
        }
        hm.cnv[c] = n;
        for (int k = 0; k < n; ++k) {
            const int v = d.cell_vtx[4 * c + k];
            if (v < 0 || v >= hm.nv) {
                err = "o2f_ctx_create: cell references a missing vertex";
                return false;
            }
            hm.cvtx[4 * c + k] = v;
        }
        cell_geom(hm, c, hm.cvol[c], hm.ccx[c], hm.ccy[c]);
    }
    // vertex -> cells
    hm.vptr.assign(hm.nv + 1, 0);
    for (int c = 0; c < hm.nc; ++c)
        for (int k = 0; k < hm.cnv[c]; ++k) ++hm.vptr[hm.cvtx[4 * c + k] + 1];
    std::partial_sum(hm.vptr.begin(), hm.vptr.end(), hm.vptr.begin());
    hm.vcell_asc.resize(hm.vptr[hm.nv]);
    {
        std::vector<int> pos(hm.vptr.begin(), hm.vptr.end() - 1);
        for (int c = 0; c < hm.nc; ++c)
            for (int k = 0; k < hm.cnv[c]; ++k) hm.vcell_asc[pos[hm.cvtx[4 * c + k]]++] = c;
    }
    // Edges: one record per cell side, ordered by (sorted vertex pair,
    // recording order); the first recorder is the left cell, a later one the
    // right cell (recordEdge, src/mesh.cpp:288-297).
    struct Side {
        uint64_t key;
        int cell;
    };
    std::vector<Side> sides;
    sides.reserve(4 * static_cast<size_t>(hm.nc));
    for (int c = 0; c < hm.nc; ++c) {
        const int n = hm.cnv[c];
        for (int k = 0; k < n; ++k)
            sides.push_back({pair_key(hm.cvtx[4 * c + k], hm.cvtx[4 * c + (k + 1) % n]), c});
    }
    std::stable_sort(sides.begin(), sides.end(),
                     [](const Side &x, const Side &y) { return x.key < y.key; });
    for (size_t i = 0; i < sides.size(); ++i) {
        if (i == 0 || sides[i].key != sides[i - 1].key) {
            hm.eva.push_back(static_cast<int>(sides[i].key >> 32));
            hm.evb.push_back(static_cast<int>(sides[i].key & 0xffffffffu));
            hm.ecl.push_back(sides[i].cell);
            hm.ecr.push_back(-1);
        } else {
            hm.ecr.back() = sides[i].cell;
        }
    }
    hm.ne = static_cast<int>(hm.eva.size());
    hm.ebc.assign(hm.ne, '\0');
    hm.elen.resize(hm.ne);
    hm.enx.resize(hm.ne);
    hm.eny.resize(hm.ne);
    hm.emx.resize(hm.ne);
    hm.emy.resize(hm.ne);
    hm.cedge.assign(4 * hm.nc, -1);
    hm.cnbr.assign(4 * hm.nc, -1);
    hm.cinfo.assign(hm.nc, 0);
    std::vector<int> n_edges(hm.nc, 0), n_nbrs(hm.nc, 0);
    auto add_nbr = [&](int c, int nb) {
        for (int k = 0; k < n_nbrs[c]; ++k)
            if (hm.cnbr[4 * c + k] == nb) return;
        hm.cnbr[4 * c + n_nbrs[c]++] = nb;
    };
    for (int e = 0; e < hm.ne; ++e) {
        auto it = edge_tag.find(pair_key(hm.eva[e], hm.evb[e]));
        if (it != edge_tag.end()) hm.ebc[e] = it->second;
        edge_geom(hm, e);
        const int cl = hm.ecl[e], cr = hm.ecr[e];
        // edges are visited in ascending id: each cell's list stays sorted
        hm.cinfo[cl] |= static_cast<uint8_t>(1u << n_edges[cl]);
        hm.cedge[4 * cl + n_edges[cl]++] = e;
        if (cr >= 0) {
            hm.cedge[4 * cr + n_edges[cr]++] = e;
            add_nbr(cl, cr);
            add_nbr(cr, cl);
        }
    }
    for (int c = 0; c < hm.nc; ++c) hm.cinfo[c] |= static_cast<uint8_t>(n_edges[c] << 4);
    hm.cevx.assign(8 * static_cast<size_t>(hm.nc), -1);
    for (int c = 0; c < hm.nc; ++c)
        for (int k = 0; k < n_edges[c]; ++k) {
            const int e = hm.cedge[4 * c + k];
            hm.cevx[8 * c + 2 * k] = hm.eva[e];
            hm.cevx[8 * c + 2 * k + 1] = hm.evb[e];
        }
    // Vertex::cellIDsOrdered: cells by increasing atan2 angle around the
    // vertex (src/mesh.cpp:231-248); equal angles keep ascending cell id.
    hm.vcell_ang = hm.vcell_asc;
    for (int v = 0; v < hm.nv; ++v) {
        std::vector<std::pair<double, int>> fan;
        for (int k = hm.vptr[v]; k < hm.vptr[v + 1]; ++k) {
            const int c = hm.vcell_asc[k];
            fan.emplace_back(std::atan2(hm.ccy[c] - hm.vy[v], hm.ccx[c] - hm.vx[v]), c);
        }
        std::stable_sort(fan.begin(), fan.end(),
                         [](const std::pair<double, int> &a, const std::pair<double, int> &b) {
                             return a.first < b.first;
                         });
        for (size_t k = 0; k < fan.size(); ++k) hm.vcell_ang[hm.vptr[v] + k] = fan[k].second;
    }
    for (int e = 0; e < hm.ne; ++e)
        if (hm.ebc[e] == 'w') hm.wall_edges.push_back(e);
    compute_stencils(hm, interp_mode);
    hm.min_h = compute_min_h(hm);
    return true;
}

void compute_stencils(HostMesh &hm, int interp_mode) {
    const size_t nnz = hm.vptr[hm.nv];
    hm.skind.assign(hm.nv, kStencilZero);
    hm.scell.assign(nnz, 0);
    hm.sw.assign(nnz, 0.0);
    std::vector<double> tmp;
    for (int v = 0; v < hm.nv; ++v) {
        const int b = hm.vptr[v], n = hm.vptr[v + 1] - b;
        if (stencil_weights(hm, interp_mode, v, hm.vx[v], hm.vy[v], hm.ccx.data(), hm.ccy.data(), &hm.sw[b], tmp)) {
            hm.skind[v] = kStencilWsum;
            const std::vector<int> &order = (n == 3) ? hm.vcell_asc : hm.vcell_ang;
            for (int k = 0; k < n; ++k) hm.scell[b + k] = order[b + k];
        } else if (n == 2) {
            hm.skind[v] = kStencilAvg2;
            hm.scell[b] = hm.vcell_asc[b];
            hm.scell[b + 1] = hm.vcell_asc[b + 1];
        } else if (n == 1) {
            hm.skind[v] = kStencilCopy;
            hm.scell[b] = hm.vcell_asc[b];
            hm.sw[b] = 1.0;
        }
    }
}

double compute_min_h(const HostMesh &hm) {
    double min_h = INFINITY;
    for (int c = 0; c < hm.nc; ++c) {
        double perimeter = 0.0;
        for (int k = 0; k < 4; ++k) {
            const int e = hm.cedge[4 * c + k];
            if (e < 0) break;
            perimeter += hm.elen[e];
        }
        const double h = 2.0 * hm.cvol[c] / perimeter;
        min_h = std::min(min_h, h);
    }
    return min_h;
}

namespace {
// NACA 4-digit mean line and thickness (src/shape_optimization.cpp:807-826).
double naca_camber(double m, double p, double xr) {
    xr = std::clamp(xr, 0.0, 1.0);
    if (m == 0.0) return 0.0;
    if (xr <= p) return m / (p * p) * (2.0 * p * xr - xr * xr);
    return m / ((1.0 - p) * (1.0 - p)) * ((1.0 - 2.0 * p) + 2.0 * p * xr - xr * xr);
}
double naca_thickness(double t, double c, double xr) {
    xr = std::clamp(xr, 0.0, 1.0);
    return 5.0 * t * c *
           (0.2969 * std::sqrt(xr) - 0.1260 * xr - 0.3516 * xr * xr + 0.2843 * xr * xr * xr -
            0.1015 * xr * xr * xr * xr);
}
// The wall vertices applyShapePerturbationNACA12 moves, from hm's positions
// into vx / vy (which may be hm's own arrays); moved[v] set for each.
void move_wall_vertices(const HostMesh &hm, double pos_x, double c, double m0, double p0, double m1, double p1,
                        double t, double *vx, double *vy, char *moved) {
    const double freeze_start = 0.9;  // the last tenth of the chord stays fixed
    for (int v = 0; v < hm.nv; ++v) {
        if (hm.vbc[v] != 'w') continue;
        const double xr = (hm.vx[v] - pos_x) / c;
        if (xr >= freeze_start) continue;
        const double yc0 = naca_camber(m0, p0, xr);
        const double yt0 = naca_thickness(t, c, xr);
        const double side = (hm.vy[v] >= yc0) ? +1.0 : -1.0;
        vx[v] = pos_x + xr * c;
        vy[v] = naca_camber(m1, p1, xr) + side * yt0;
        if (moved) moved[v] = 1;
    }
}
}  // namespace

bool perturb_naca(HostMesh &hm, int which, double pos_x, double c, double m0, double p0,
                  double t, double eps) {
    double m1 = m0, p1 = p0;
    if (which == 0)
        m1 = m0 + eps;
    else if (which == 1)
        p1 = p0 + eps;
    else
        return false;
    move_wall_vertices(hm, pos_x, c, m0, p0, m1, p1, t, hm.vx.data(), hm.vy.data(), nullptr);
    // updateDerivedGeometry: only edges touching a wall vertex, and only their
    // left cells, are refreshed (src/shape_optimization.cpp:748-805).
    for (int e = 0; e < hm.ne; ++e) {
        if (hm.ebc[e] != 'w' && hm.vbc[hm.eva[e]] != 'w' && hm.vbc[hm.evb[e]] != 'w') continue;
        const int cl = hm.ecl[e];
        cell_geom(hm, cl, hm.cvol[cl], hm.ccx[cl], hm.ccy[cl]);
        edge_geom(hm, e);
    }
    return true;
}

bool perturb_naca_sparse(const HostMesh &hm, int interp_mode, int which, double pos_x, double c, double m0,
                         double p0, double t, double eps, GeomSparse &out) {
    double m1 = m0, p1 = p0;
    if (which == 0)
        m1 = m0 + eps;
    else if (which == 1)
        p1 = p0 + eps;
    else
        return false;
    out = GeomSparse();
    // the wall vertices, moved as perturb_naca moves them
    std::vector<double> vx(hm.vx), vy(hm.vy);
    std::vector<char> vmoved(hm.nv, 0);
    move_wall_vertices(hm, pos_x, c, m0, p0, m1, p1, t, vx.data(), vy.data(), vmoved.data());
    // updateDerivedGeometry: the edges touching a wall vertex and their left
    // cells (each cell from its own vertices, each edge from the updated
    // centre of its left cell, as the in-place update leaves them)
    std::vector<double> ccx(hm.ccx), ccy(hm.ccy);
    std::vector<char> cell_on(hm.nc, 0);
    std::vector<int> edges;
    for (int e = 0; e < hm.ne; ++e)
        if (hm.ebc[e] == 'w' || hm.vbc[hm.eva[e]] == 'w' || hm.vbc[hm.evb[e]] == 'w') {
            edges.push_back(e);
            cell_on[hm.ecl[e]] = 1;
        }
    for (int cl = 0; cl < hm.nc; ++cl) {
        if (!cell_on[cl]) continue;
        double vol, cx, cy;
        cell_geom_xy(hm, vx.data(), vy.data(), cl, vol, cx, cy);
        ccx[cl] = cx;
        ccy[cl] = cy;
        out.cell.push_back(cl);
        out.cval.insert(out.cval.end(), {vol, cx, cy});
    }
    for (int e : edges) {
        double o[5];
        edge_geom_xy(hm, vx.data(), vy.data(), ccx.data(), ccy.data(), e, o);
        out.edge.push_back(e);
        out.eval.insert(out.eval.end(), o, o + 5);
    }
    // compute_stencils: the weights of the vertices that moved or whose
    // cells did (every other vertex's weights are hm's, recomputed from the
    // same inputs)
    std::vector<double> tmp, w;
    for (int v = 0; v < hm.nv; ++v) {
        const int b = hm.vptr[v], n = hm.vptr[v + 1] - b;
        bool any = vmoved[v];
        for (int k = 0; k < n && !any; ++k) any = cell_on[hm.vcell_asc[b + k]];
        if (!any) continue;
        w.resize(n);
        if (!stencil_weights(hm, interp_mode, v, vx[v], vy[v], ccx.data(), ccy.data(), w.data(), tmp)) continue;
        for (int k = 0; k < n; ++k) {
            out.swi.push_back(b + k);
            out.swv.push_back(w[k]);
        }
    }
    return true;
}

bool read_msh2(const std::string &path, std::vector<double> &xy, std::vector<int> &cell_nv,
               std::vector<int> &cell_vtx, std::vector<int> &bnd_vtx, std::vector<char> &bnd_tag,
               std::string &err) {
    std::ifstream in(path);
    if (!in) {
        err = "cannot open mesh file " + path;
        return false;
    }
    std::map<std::pair<int, int>, std::string> names;
    std::map<long, int> node_index;
    struct Elem {
        int type;
        std::vector<long> nodes;
    };
    // (dim, physical tag) -> elementary tag -> elements in file order
    std::map<std::pair<int, int>, std::map<int, std::vector<Elem>>> groups;
    std::string line;
    xy.clear();
    while (std::getline(in, line)) {
        if (line.rfind("$PhysicalNames", 0) == 0) {
            int n = 0;
            in >> n;
            std::getline(in, line);
            for (int k = 0; k < n; ++k) {
                std::getline(in, line);
                std::istringstream ss(line);
                int dim, tag;
                ss >> dim >> tag;
                std::string rest;
                std::getline(ss, rest);
                const size_t q0 = rest.find('"'), q1 = rest.rfind('"');
                names[{dim, tag}] =
                    (q0 != std::string::npos && q1 > q0) ? rest.substr(q0 + 1, q1 - q0 - 1) : rest;
            }
        } else if (line.rfind("$Nodes", 0) == 0) {
            long n = 0;
            in >> n;
            xy.resize(2 * n);
            for (long k = 0; k < n; ++k) {
                long tag;
                double x, y, z;
                in >> tag >> x >> y >> z;
                node_index[tag] = static_cast<int>(k);
                xy[2 * k] = x;
                xy[2 * k + 1] = y;
            }
        } else if (line.rfind("$Elements", 0) == 0) {
            long n = 0;
            in >> n;
            for (long k = 0; k < n; ++k) {
                long num;
                int type, ntags;
                in >> num >> type >> ntags;
                std::vector<int> tags(ntags);
                for (int q = 0; q < ntags; ++q) in >> tags[q];
                int nn = type == 1 ? 2 : type == 2 ? 3 : type == 3 ? 4 : type == 15 ? 1 : -1;
                if (nn < 0) {
                    err = "unsupported element type in " + path;
                    return false;
                }
                Elem el{type, std::vector<long>(nn)};
                for (int q = 0; q < nn; ++q) in >> el.nodes[q];
                if (type == 15 || ntags < 2) continue;
                const int dim = type == 1 ? 1 : 2;
                groups[{dim, tags[0]}][tags[1]].push_back(std::move(el));
            }
        }
        if (!in) break;
    }
    if (xy.empty()) {
        err = "no $Nodes section in " + path;
        return false;
    }
    cell_nv.clear();
    cell_vtx.clear();
    bnd_vtx.clear();
    bnd_tag.clear();
    bool unknown_node = false;
    auto idx = [&](long tag) -> int {
        auto it = node_index.find(tag);
        if (it == node_index.end()) {
            unknown_node = true;
            return 0;
        }
        return it->second;
    };
    for (auto &g : groups) {
        std::string name;
        auto nit = names.find(g.first);
        if (nit != names.end()) name = nit->second;
        char tag = name == "Inlet" ? 'i' : name == "Outlet" ? 'o' : name == "FreeStream" ? 'f'
                 : (name == "Wall" || name == "Airfoil") ? 'w' : '\0';
        for (auto &ent : g.second)
            for (auto &el : ent.second) {
                if (el.type == 1) {
                    if (!tag) continue;
                    bnd_vtx.push_back(idx(el.nodes[0]));
                    bnd_vtx.push_back(idx(el.nodes[1]));
                    bnd_tag.push_back(tag);
                } else {
                    const int nn = el.type == 2 ? 3 : 4;
                    cell_nv.push_back(nn);
                    for (int q = 0; q < 4; ++q) cell_vtx.push_back(q < nn ? idx(el.nodes[q]) : -1);
                }
            }
    }
    if (unknown_node) {
        err = "element references an unknown node in " + path;
        return false;
    }
    return true;
}

}  // namespace o2f
