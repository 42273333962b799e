// Host-side forward-mode derivatives of the mesh geometry along a vertex
// displacement: the tangent of the derived geometry (cell areas and centres,
// edge lengths, normals and midpoints, interpolation weights) that the
// device tangent residual consumes, and the displacement of the wall
// vertices per NACA 4-digit parameter.
//
// The reference differentiates the shape by central finite differences
// over copies of the mesh (computeShapeGradient, src/shape_optimization.cpp:
// 942-1016, perturbing with applyShapePerturbationNACA12 and
// updateDerivedGeometry, :748-890).  The functions here give the limit of
// those differences, eps -> 0, exactly: every formula below is the
// derivative of the corresponding one in host_mesh.cpp.
#include <algorithm>
#include <cmath>
#include <string>

#include "host_mesh.hpp"

namespace o2f {

namespace {

// (value, derivative) pair with the arithmetic the formulas need.
struct HD {
    double v, d;
};
inline HD operator+(HD a, HD b) { return {a.v + b.v, a.d + b.d}; }
inline HD operator-(HD a, HD b) { return {a.v - b.v, a.d - b.d}; }
inline HD operator-(HD a) { return {-a.v, -a.d}; }
inline HD operator*(HD a, HD b) { return {a.v * b.v, a.d * b.v + a.v * b.d}; }
inline HD operator/(HD a, HD b) {
    const double q = a.v / b.v;
    return {q, (a.d - q * b.d) / b.v};
}
inline HD operator*(double a, HD b) { return {a * b.v, a * b.d}; }
inline HD operator+(HD a, double b) { return {a.v + b, a.d}; }
inline HD operator-(double a, HD b) { return {a - b.v, -b.d}; }
inline HD hd_abs(HD a) { return a.v < 0.0 ? -a : a; }
inline HD hd_sqrt(HD a) {
    const double r = std::sqrt(a.v);
    return {r, a.d / (2.0 * r)};
}
inline HD hd_hypot(HD x, HD y) {
    const double r = std::hypot(x.v, y.v);
    return {r, (x.v * x.d + y.v * y.d) / r};
}
// std::clamp(x, -1, 1): constant outside the interval
inline HD hd_clamp1(HD x) { return x.v < -1.0 ? HD{-1.0, 0.0} : (1.0 < x.v ? HD{1.0, 0.0} : x); }
inline HD hd_acos(HD x) {
    const double s = std::sqrt(std::max(0.0, 1.0 - x.v * x.v));
    return {std::acos(x.v), s > 0.0 ? -x.d / s : 0.0};
}
inline HD hd_tan(HD x) {
    const double t = std::tan(x.v);
    return {t, x.d * (1.0 + t * t)};
}

inline HD vxd(const HostMesh &hm, const double *dxy, int v) { return {hm.vx[v], dxy[2 * v]}; }
inline HD vyd(const HostMesh &hm, const double *dxy, int v) { return {hm.vy[v], dxy[2 * v + 1]}; }

// cell_geom (host_mesh.cpp) differentiated
void cell_geom_d(const HostMesh &hm, const double *dxy, int c, HD &vol, HD &cx, HD &cy) {
    const int *v = &hm.cvtx[4 * c];
    const HD x0 = vxd(hm, dxy, v[0]), y0 = vyd(hm, dxy, v[0]);
    const HD x1 = vxd(hm, dxy, v[1]), y1 = vyd(hm, dxy, v[1]);
    const HD x2 = vxd(hm, dxy, v[2]), y2 = vyd(hm, dxy, v[2]);
    if (hm.cnv[c] == 3) {
        const HD e1x = x1 - x0, e1y = y1 - y0, e2x = x2 - x0, e2y = y2 - y0;
        vol = hd_abs(0.5 * (e1x * e2y - e2x * e1y));
        cx = (1.0 / 3.0) * (x0 + x1 + x2);
        cy = (1.0 / 3.0) * (y0 + y1 + y2);
    } else {
        const HD x3 = vxd(hm, dxy, v[3]), y3 = vyd(hm, dxy, v[3]);
        const HD e1x = x1 - x0, e1y = y1 - y0, e2x = x2 - x0, e2y = y2 - y0;
        const HD f1x = x1 - x2, f1y = y1 - y2, f2x = x3 - x2, f2y = y3 - y2;
        vol = 0.5 * (hd_abs(e1x * e2y - e2x * e1y) + hd_abs(f1x * f2y - f2x * f1y));
        cx = 0.25 * (x0 + x1 + x2 + x3);
        cy = 0.25 * (y0 + y1 + y2 + y3);
    }
}

// edge_geom (host_mesh.cpp) differentiated; the orientation test keeps the
// sign of the mesh's own normal (a sign does not change under an
// infinitesimal displacement).  d(length), d(normal), d(midpoint).
void edge_geom_d(const HostMesh &hm, const double *dxy, int e, double &dlen, double &dnx, double &dny,
                 double &dmx, double &dmy) {
    const int a = hm.eva[e], b = hm.evb[e];
    const HD tx = vxd(hm, dxy, a) - vxd(hm, dxy, b), ty = vyd(hm, dxy, a) - vyd(hm, dxy, b);
    const HD len = hd_sqrt(tx * tx + ty * ty);
    HD nx = ty / len, ny = -tx / len;
    if (nx.v * hm.enx[e] + ny.v * hm.eny[e] < 0) {
        nx = -nx;
        ny = -ny;
    }
    dlen = len.d;
    dnx = nx.d;
    dny = ny.d;
    dmx = 0.5 * (dxy[2 * a] + dxy[2 * b]);
    dmy = 0.5 * (dxy[2 * a + 1] + dxy[2 * b + 1]);
}

// The derivative of every cell, edge and stencil weight that can follow
// the displacement, handed to `out` (out.cell, out.edge, out.sw); every
// other entry's derivative is exactly zero.
template <class Out>
void tangent_entities(const HostMesh &hm, const double *dxy, int geometry, Out &out) {
    // which vertices move; which cells and edges follow their vertices
    // (entities none of whose vertices move have a zero derivative and are
    // skipped)
    std::vector<char> moved(hm.nv, 0);
    for (int v = 0; v < hm.nv; ++v) moved[v] = dxy[2 * v] != 0.0 || dxy[2 * v + 1] != 0.0;
    std::vector<char> cell_on(hm.nc, 0), edge_on(hm.ne, 0);
    if (geometry == O2F_GEOM_FULL) {
        for (int c = 0; c < hm.nc; ++c)
            for (int k = 0; k < hm.cnv[c]; ++k) cell_on[c] |= moved[hm.cvtx[4 * c + k]];
        for (int e = 0; e < hm.ne; ++e) edge_on[e] = moved[hm.eva[e]] || moved[hm.evb[e]];
    } else {
        // updateDerivedGeometry (src/shape_optimization.cpp:748-805): edges
        // touching a wall vertex and their left cells
        for (int e = 0; e < hm.ne; ++e)
            if (hm.ebc[e] == 'w' || hm.vbc[hm.eva[e]] == 'w' || hm.vbc[hm.evb[e]] == 'w') {
                edge_on[e] = 1;
                cell_on[hm.ecl[e]] = 1;
            }
    }
    std::vector<HD> cxd(hm.nc), cyd(hm.nc);
    std::vector<char> centre_moved(hm.nc, 0);
    for (int c = 0; c < hm.nc; ++c) {
        cxd[c] = {hm.ccx[c], 0.0};
        cyd[c] = {hm.ccy[c], 0.0};
        if (!cell_on[c]) continue;
        HD vol, cx, cy;
        cell_geom_d(hm, dxy, c, vol, cx, cy);
        out.cell(c, vol.d, cx.d, cy.d);
        cxd[c].d = cx.d;
        cyd[c].d = cy.d;
        centre_moved[c] = cx.d != 0.0 || cy.d != 0.0;
    }
    for (int e = 0; e < hm.ne; ++e)
        if (edge_on[e]) {
            double dl, dnx, dny, dmx, dmy;
            edge_geom_d(hm, dxy, e, dl, dnx, dny, dmx, dmy);
            out.edge(e, dl, dnx, dny, dmx, dmy);
        }
    // compute_stencils (host_mesh.cpp) differentiated: the weights follow the
    // cell centres and the vertex's own position
    std::vector<HD> w, dist, ux, uy;
    for (int v = 0; v < hm.nv; ++v) {
        if (hm.skind[v] != kStencilWsum) continue;
        const int b = hm.vptr[v], n = hm.vptr[v + 1] - b;
        bool any = moved[v];
        for (int k = 0; k < n && !any; ++k) any = centre_moved[hm.scell[b + k]];
        if (!any) continue;  // neither the vertex nor a stencil centre moves
        const HD px = vxd(hm, dxy, v), py = vyd(hm, dxy, v);
        if (n == 3) {
            const int i0 = hm.scell[b], i1 = hm.scell[b + 1], i2 = hm.scell[b + 2];
            const HD ax = cxd[i0], ay = cyd[i0], bx = cxd[i1], by = cyd[i1], qx = cxd[i2], qy = cyd[i2];
            const HD num0 = (bx - px) * (qy - py) - (qx - px) * (by - py);
            const HD den = (bx - ax) * (qy - ay) - (qx - ax) * (by - ay);
            const HD w0 = num0 / den;
            const HD num1 = (qx - px) * (ay - py) - (ax - px) * (qy - py);
            const HD w1 = num1 / den;
            out.sw(b, w0.d);
            out.sw(b + 1, w1.d);
            out.sw(b + 2, -w0.d - w1.d);
        } else {
            // a weighted stencil of n != 3 cells is the mean-value form
            w.assign(n, HD{0.0, 0.0});
            dist.assign(n, HD{0.0, 0.0});
            ux.assign(n, HD{0.0, 0.0});
            uy.assign(n, HD{0.0, 0.0});
            for (int k = 0; k < n; ++k) {
                const int c = hm.scell[b + k];
                const HD dx = cxd[c] - px, dy = cyd[c] - py;
                dist[k] = hd_hypot(dx, dy);
                ux[k] = dx / dist[k];
                uy[k] = dy / dist[k];
            }
            HD wsum{0.0, 0.0};
            for (int k = 0; k < n; ++k) {
                const int prev = (k + n - 1) % n, next = (k + 1) % n;
                const HD cp = hd_clamp1(ux[prev] * ux[k] + uy[prev] * uy[k]);
                const HD cn = hd_clamp1(ux[k] * ux[next] + uy[k] * uy[next]);
                w[k] = (hd_tan(0.5 * hd_acos(cp)) + hd_tan(0.5 * hd_acos(cn))) / dist[k];
                wsum = wsum + w[k];
            }
            for (int k = 0; k < n; ++k) out.sw(b + k, (w[k] / wsum).d);
        }
    }
}

struct DenseOut {
    GeomTangent &t;
    void cell(int c, double v, double x, double y) {
        t.cvol[c] = v;
        t.ccx[c] = x;
        t.ccy[c] = y;
    }
    void edge(int e, double l, double nx, double ny, double mx, double my) {
        t.elen[e] = l;
        t.enx[e] = nx;
        t.eny[e] = ny;
        t.emx[e] = mx;
        t.emy[e] = my;
    }
    void sw(int i, double v) { t.sw[i] = v; }
};

struct SparseOut {
    GeomSparse &t;
    void cell(int c, double v, double x, double y) {
        t.cell.push_back(c);
        t.cval.insert(t.cval.end(), {v, x, y});
    }
    void edge(int e, double l, double nx, double ny, double mx, double my) {
        t.edge.push_back(e);
        t.eval.insert(t.eval.end(), {l, nx, ny, mx, my});
    }
    void sw(int i, double v) {
        t.swi.push_back(i);
        t.swv.push_back(v);
    }
};

}  // namespace

void geometry_tangent(const HostMesh &hm, const double *dxy, int geometry, GeomTangent &t) {
    t.cvol.assign(hm.nc, 0.0);
    t.ccx.assign(hm.nc, 0.0);
    t.ccy.assign(hm.nc, 0.0);
    t.elen.assign(hm.ne, 0.0);
    t.enx.assign(hm.ne, 0.0);
    t.eny.assign(hm.ne, 0.0);
    t.emx.assign(hm.ne, 0.0);
    t.emy.assign(hm.ne, 0.0);
    t.sw.assign(hm.sw.size(), 0.0);
    DenseOut out{t};
    tangent_entities(hm, dxy, geometry, out);
}

void geometry_tangent_sparse(const HostMesh &hm, const double *dxy, int geometry, GeomSparse &t) {
    t = GeomSparse();
    SparseOut out{t};
    tangent_entities(hm, dxy, geometry, out);
}

namespace {
// NACA 4-digit mean line and thickness (src/shape_optimization.cpp:807-826)
// with a differentiated parameter.
HD camber_d(HD m, HD p, double xr) {
    xr = std::clamp(xr, 0.0, 1.0);
    if (m.v == 0.0 && m.d == 0.0) return {0.0, 0.0};
    if (xr <= p.v) return m / (p * p) * (2.0 * p * HD{xr, 0.0} - HD{xr * xr, 0.0});
    const HD q = 1.0 - p;
    return m / (q * q) * (((1.0 - 2.0 * p) + 2.0 * p * HD{xr, 0.0}) - HD{xr * xr, 0.0});
}
double thickness_unit(double c, double xr) {  // naca_thickness / t
    xr = std::clamp(xr, 0.0, 1.0);
    return 5.0 * c *
           (0.2969 * std::sqrt(xr) - 0.1260 * xr - 0.3516 * xr * xr + 0.2843 * xr * xr * xr -
            0.1015 * xr * xr * xr * xr);
}
double camber(double m, double p, double xr) { return camber_d({m, 0.0}, {p, 0.0}, xr).v; }
}  // namespace

bool naca_displacement(const HostMesh &hm, int which, double pos_x, double c, double m0, double p0,
                       std::vector<double> &dxy) {
    dxy.assign(2 * static_cast<size_t>(hm.nv), 0.0);
    if (which < 0 || which > 2) return false;
    const double freeze_start = 0.9;  // applyShapePerturbationNACA12 leaves x_r >= 0.9 fixed
    for (int v = 0; v < hm.nv; ++v) {
        if (hm.vbc[v] != 'w') continue;
        const double xr = (hm.vx[v] - pos_x) / c;
        if (xr >= freeze_start) continue;
        const double side = (hm.vy[v] >= camber(m0, p0, xr)) ? +1.0 : -1.0;
        double dy;
        if (which == 2)
            dy = side * thickness_unit(c, xr);  // y = yc + side * t * (thickness per unit t)
        else
            dy = camber_d({m0, which == 0 ? 1.0 : 0.0}, {p0, which == 1 ? 1.0 : 0.0}, xr).d;
        dxy[2 * v + 1] = dy;
    }
    return true;
}

}  // namespace o2f
