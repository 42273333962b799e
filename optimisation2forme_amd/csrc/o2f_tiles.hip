// Fused residual stage over compact cell tiles (tables: host_tiles.hpp).
//
// One workgroup runs one whole stage -- vertex states, Green-Gauss
// gradients, limiters, face states, HLLC, cell and vertex stresses, the
// convective and viscous sums and the Shu-Osher update -- for the cells of
// one tile, with every intermediate in LDS.  The four-kernel stage
// (o2f_kernels.hip) writes about 88 MB of intermediates to memory and reads
// them back per stage on the 200k-triangle workload and spreads a cell over
// four lanes; here each cell, vertex and edge is one lane and only the state
// in, the state out and the read-only mesh cross the memory system.
//
// Every entity is formed with the arithmetic of the four-kernel stage (same
// expressions, same summation orders), so the results are bit-identical to
// it and to the CPU restatement.  A tile recomputes the entities on its rim
// that its neighbours also form (about 1.5x the gradients at 192-cell
// tiles): that is the price of keeping the rest on chip.
//
// The state is read from U and written to Uout (a different buffer): the
// tiles of one launch read the old state of their rim cells while others
// write their own cells.  Between stages the state is AoS (four doubles,
// one 32-byte run per cell): a cell's state is then one cache line where the
// SoA layout of the C ABI spreads it over four, and the stage's time follows
// the number of distinct lines its gathers touch.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "host_tiles.hpp"
#include "o2f_device.hpp"
#include "o2f_tiles.hpp"

namespace o2f {

namespace {

// LDS rows of four doubles as two 16-byte accesses.
__device__ __forceinline__ void lds_ld4(const double *p, double (&v)[4]) {
    const double2 a = reinterpret_cast<const double2 *>(p)[0], b = reinterpret_cast<const double2 *>(p)[1];
    v[0] = a.x;
    v[1] = a.y;
    v[2] = b.x;
    v[3] = b.y;
}
__device__ __forceinline__ void lds_st4(double *p, const double (&v)[4]) {
    reinterpret_cast<double2 *>(p)[0] = make_double2(v[0], v[1]);
    reinterpret_cast<double2 *>(p)[1] = make_double2(v[2], v[3]);
}

// One gradient-cell record (TileCell) in four 16-byte loads.
struct Rec {
    int c, info;
    int edge[4], nbr[4];
    int lv[8], le[4];
};
__device__ __forceinline__ Rec load_rec(const TileCell *p) {
    const int4 *q = reinterpret_cast<const int4 *>(p);
    const int4 w0 = q[0], w1 = q[1], w2 = q[2], w3 = q[3];
    Rec r;
    r.c = w0.x;
    r.edge[0] = w0.y;
    r.edge[1] = w0.z;
    r.edge[2] = w0.w;
    r.edge[3] = w1.x;
    r.nbr[0] = w1.y;
    r.nbr[1] = w1.z;
    r.nbr[2] = w1.w;
    r.nbr[3] = w2.x;
    const unsigned lvw[4] = {(unsigned)w2.y, (unsigned)w2.z, (unsigned)w2.w, (unsigned)w3.x};
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        r.lv[2 * j] = (int)(lvw[j] & 0xffffu);
        r.lv[2 * j + 1] = (int)(lvw[j] >> 16);
    }
    const unsigned lew[2] = {(unsigned)w3.y, (unsigned)w3.z};
#pragma unroll
    for (int j = 0; j < 2; ++j) {
        r.le[2 * j] = (int)(short)(lew[j] & 0xffffu);
        r.le[2 * j + 1] = (int)(short)(lew[j] >> 16);
    }
    r.info = w3.w & 0xff;
    return r;
}

static_assert(kTileSlots == kStencilSlots, "a vertex record holds the gathers' stencil slots");

// A state vertex's record in three 16-byte loads: its first stencil cells,
// the stencil offset, length, kind and tag.
struct Gat {
    int cid[kStencilSlots];
    int b, n;
    uint8_t kind;
    char bc;
};
__device__ __forceinline__ Gat load_gat(const TileGather *p) {
    const int4 *q = reinterpret_cast<const int4 *>(p);
    const int4 w0 = q[0], w1 = q[1], w2 = q[2];
    Gat g;
    g.cid[0] = w0.x;
    g.cid[1] = w0.y;
    g.cid[2] = w0.z;
    g.cid[3] = w0.w;
    g.cid[4] = w1.x;
    g.cid[5] = w1.y;
    g.cid[6] = w1.z;
    g.cid[7] = w1.w;
    g.b = w2.x;
    g.n = w2.y & 0xffff;
    g.kind = (uint8_t)((w2.y >> 16) & 0xff);
    g.bc = (char)(w2.y >> 24);
    return g;
}

// A stress vertex's record in six 16-byte loads: its first stencil weights
// and local cells, then length, kind, tag, stencil offset and local stencil
// offset.
struct Str {
    double w[kStencilSlots];
    int lc[kStencilSlots];
    int n, b, ls;
    uint8_t kind;
    char bc;
};
__device__ __forceinline__ Str load_str(const TileStress *p) {
    const double2 *d = reinterpret_cast<const double2 *>(p);
    Str r;
#pragma unroll
    for (int h = 0; h < kStencilSlots / 2; ++h) {
        const double2 t = d[h];
        r.w[2 * h] = t.x;
        r.w[2 * h + 1] = t.y;
    }
    const int4 *q = reinterpret_cast<const int4 *>(p->lc);
    const int4 l = q[0], m = q[1];
    const unsigned lw[4] = {(unsigned)l.x, (unsigned)l.y, (unsigned)l.z, (unsigned)l.w};
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        r.lc[2 * j] = (int)(lw[j] & 0xffffu);
        r.lc[2 * j + 1] = (int)(lw[j] >> 16);
    }
    r.n = m.x & 0xffff;
    r.kind = (uint8_t)((m.x >> 16) & 0xff);
    r.bc = (char)(m.x >> 24);
    r.b = m.y;
    r.ls = m.z;
    return r;
}

// LDS layout (in doubles; every region a multiple of four):
//   [0, vert)            state vertices (4 each), later stress vertices (4 each, then q_y of each)
//   [vert, vert + 5 gd)  cell stresses of the gradient cells (4 each, then q_y of each)
//   [.., + 8 e)          face states of the tile's edges: L (4) then R (4); the flux replaces L
struct Lds {
    double *vert, *tq, *fs;
    int max_vo, max_gd;
};

// Gradient cell i (local): edge values + Green-Gauss (src/advection.cpp:136-188)
// and the cell stress / heat flux (src/diffusion.cpp:223-241); for a face
// cell also the neighbour min / max, the limiter and its face states on the
// tile's edges with the boundary ghosts (src/advection.cpp:232-321,
// src/boundary.cpp:48-80).  The arithmetic of reconstruct_body, one lane
// per cell.
template <int NK>
__device__ __forceinline__ void tile_cell(const Rec &r, bool face, const Geom &G, const double2 *__restrict__ EP,
                                          const Phys &P, int nc, const double *__restrict__ U, const Lds &L, int i) {
    const int c = r.c;
    const int nk = r.info >> 4;
    double u[4];
    ld4(U, c, u);
    const double vol = G.cvol[c];
    double gx[4] = {0.0, 0.0, 0.0, 0.0}, gy[4] = {0.0, 0.0, 0.0, 0.0};
    double dx[4] = {0.0, 0.0, 0.0, 0.0}, dy[4] = {0.0, 0.0, 0.0, 0.0};
    int bcs[4] = {0, 0, 0, 0};
    const double cx = face ? G.ccx[c] : 0.0, cy = face ? G.ccy[c] : 0.0;
#pragma unroll
    for (int j = 0; j < NK; ++j) {
        if (j >= nk) break;
        const double2 *pe = EP + 4 * (long)r.edge[j];
        const double2 n2 = pe[0], lb = pe[1];
        const double el = lb.x;
        double nx = n2.x, ny = n2.y;
        const char ebc = (char)__double2loint(lb.y);
        bcs[j] = ebc;
        if (face) {
            const double2 m2 = pe[2];
            dx[j] = m2.x - cx;
            dy[j] = m2.y - cy;
        }
        double xa[4], xb[4], val[4];
        lds_ld4(L.vert + 4 * r.lv[2 * j], xa);
        lds_ld4(L.vert + 4 * r.lv[2 * j + 1], xb);
#pragma unroll
        for (int q = 0; q < 4; ++q) val[q] = 0.5 * (xa[q] + xb[q]);
        if (tagged(ebc)) bc_state(P, ebc, val[0], val[1], val[2], val[3]);
        if ((r.info >> j) & 1) {
            nx = -nx;
            ny = -ny;
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            gx[q] += val[q] * el * nx / vol;
            gy[q] += val[q] * el * ny / vol;
        }
    }
    double t5[5];
    cell_stress_t(u, gx, gy, P.mu, P.gamma, P.kappa, t5);
    {
        const double t4[4] = {t5[0], t5[1], t5[2], t5[3]};
        lds_st4(L.tq + 4 * i, t4);
        L.tq[4 * L.max_gd + i] = t5[4];
    }
    if (!face) return;
    double nbv[4][4];
#pragma unroll
    for (int j = 0; j < NK; ++j) {
        const int nb = r.nbr[j];
        if (nb >= 0) {
            ld4(U, nb, nbv[j]);
        } else {
#pragma unroll
            for (int q = 0; q < 4; ++q) nbv[j][q] = u[q];
        }
    }
    const double veps = (P.limiter == 1) ? 0.0 : venkat_eps(vol, P.venkat_k);
    const double fsq[4] = {P.fs_rho, P.fs_ru, P.fs_rv, P.fs_E};
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        double mn = u[q], mx = u[q];
#pragma unroll
        for (int j = 0; j < NK; ++j) {
            const double v = nbv[j][q];
            if (v < mn) mn = v;
            if (v > mx) mx = v;
        }
        double al = 1.0;
#pragma unroll
        for (int j = 0; j < NK; ++j) {
            if (j >= nk) break;
            const double pred = u[q] + gx[q] * dx[j] + gy[q] * dy[j];
            const double a = (P.limiter == 1) ? alpha_bj(u[q], pred, mn, mx)
                                              : alpha_venkat_eps(u[q], pred, mn, mx, veps);
            al = smin(al, a);
        }
#pragma unroll
        for (int j = 0; j < NK; ++j) {
            if (j >= nk) break;
            const int le = r.le[j];
            if (le < 0) continue;
            const double f = u[q] + al * (gx[q] * dx[j] + gy[q] * dy[j]);
            double *row = L.fs + 8 * le;
            if ((r.info >> j) & 1) {
                row[q] = f;
                const char bcj = (char)bcs[j];
                if (tagged(bcj)) row[4 + q] = (bcj == 'i' || bcj == 'f') ? fsq[q] : (bcj == 'w' && (q == 1 || q == 2)) ? -f : f;
            } else {
                row[4 + q] = f;
            }
        }
    }
}

// Stress vertex: the viscous gather of vertex v from the tile's cell
// stresses (src/diffusion.cpp:28-140, apply_diffusion_bc), the arithmetic of
// vertex_gather_vals<5, true>, with the stencil's cells given by their
// local indices and the first weights in the record (a longer stencil reads
// the rest from lsten and the stencil weights).
__device__ __forceinline__ void tile_vertex_stress(const Str &r, const uint16_t *__restrict__ lsten, const Geom &G,
                                                   const Lds &L, int j) {
    auto val = [&](int li, double (&t)[5]) {
        double t4[4];
        lds_ld4(L.tq + 4 * li, t4);
#pragma unroll
        for (int q = 0; q < 4; ++q) t[q] = t4[q];
        t[4] = L.tq[4 * L.max_gd + li];
    };
    const int n = r.n;
    double out[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    if (r.kind == kStencilWsum) {
        double t[5];
        val(r.lc[0], t);
#pragma unroll
        for (int q = 0; q < 5; ++q) out[q] = r.w[0] * t[q];
#pragma unroll
        for (int k = 1; k < kStencilSlots; ++k)
            if (k < n) {
                val(r.lc[k], t);
#pragma unroll
                for (int q = 0; q < 5; ++q) out[q] = out[q] + r.w[k] * t[q];
            }
        for (int k = kStencilSlots; k < n; ++k) {
            val(lsten[r.ls + k], t);
            const double wk = G.sw[r.b + k];
#pragma unroll
            for (int q = 0; q < 5; ++q) out[q] = out[q] + wk * t[q];
        }
    } else if (r.kind == kStencilAvg2) {
        double t0[5], t1[5];
        val(r.lc[0], t0);
        val(r.lc[1], t1);
#pragma unroll
        for (int q = 0; q < 5; ++q) out[q] = 0.5 * (t0[q] + t1[q]);
    } else if (r.kind == kStencilCopy) {
        val(r.lc[0], out);
    }
    if (r.bc == 'f' || r.bc == 'w') {
#pragma unroll
        for (int q = 0; q < 5; ++q) out[q] = 0.0;
    }
    const double o4[4] = {out[0], out[1], out[2], out[3]};
    lds_st4(L.vert + 4 * j, o4);
    L.vert[4 * L.max_vo + j] = out[4];
}

// Owned cell i: the convective sum of its edges' fluxes
// (src/advection.cpp:433-464), the viscous divergence
// (src/diffusion.cpp:145-201) and the Shu-Osher update
// (src/simulation.cpp:159-167): the arithmetic of flux_update_body, one
// lane per cell.  Returns |u| + c of the new state (kFlagSpeed), else 0.
template <bool OUT_AOS, int NK>
__device__ __forceinline__ double tile_update(const Rec &r, const TileOwn &o, const Topo &T, const Geom &G,
                                              const double2 *__restrict__ EP, const Phys &P,
                                              const double *__restrict__ U,
                                              double *__restrict__ Uout, double *__restrict__ Uold,
                                              double *__restrict__ R, const double *__restrict__ dtp, double c0,
                                              double c1, int flags, const Lds &L) {
    const int nc = T.nc, c = r.c;
    const int nk = r.info >> 4;
    const double vol = G.cvol[c];
    double u[4];
    ld4(U, c, u);
    // one division each: u, v, 1 / V
    const double uu = u[1] / u[0], vv = u[2] / u[0], invV = 1.0 / vol;
    double rk[4] = {0.0, 0.0, 0.0, 0.0};
    double w1[4] = {0.0, 0.0, 0.0, 0.0}, w2[4] = {0.0, 0.0, 0.0, 0.0}, w3[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int j = 0; j < NK; ++j) {
        if (j >= nk) break;
        const double2 *pe = EP + 4 * (long)r.edge[j];
        const double2 n2 = pe[0], lt = pe[1];
        const double el = lt.x;
        double nx = n2.x, ny = n2.y;
        const char ebc = (char)__double2loint(lt.y);
        const bool is_left = (r.info >> j) & 1;
        double F[4];
        lds_ld4(L.fs + 8 * r.le[j], F);
        const double sign = is_left ? +1.0 : -1.0;
        // convective term of this edge, src/advection.cpp:453-456, summed in slot order
#pragma unroll
        for (int q = 0; q < 4; ++q) rk[q] += sign * F[q] * el;
        // viscous terms of this edge, src/diffusion.cpp:183-200
        if (is_left) {
            nx = -nx;
            ny = -ny;
        }
        double ta[4], tb[4], t[5];
        const int la = o.lo[2 * j], lb = o.lo[2 * j + 1];
        lds_ld4(L.vert + 4 * la, ta);
        lds_ld4(L.vert + 4 * lb, tb);
#pragma unroll
        for (int q = 0; q < 4; ++q) t[q] = 0.5 * (ta[q] + tb[q]);
        t[4] = 0.5 * (L.vert[4 * L.max_vo + la] + L.vert[4 * L.max_vo + lb]);
        if (ebc == 'f' || ebc == 'w') {
#pragma unroll
            for (int q = 0; q < 5; ++q) t[q] = 0.0;
        }
        const double txx = t[0], tyy = t[1], txy = t[2], qx = t[3], qy = t[4];
        w1[j] = txx * el * nx / vol + txy * el * ny / vol;
        w2[j] = txy * el * nx / vol + tyy * el * ny / vol;
        w3[j] = (uu * txx + vv * txy - qx) * el * nx / vol + (uu * txy + vv * tyy - qy) * el * ny / vol;
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) rk[q] = rk[q] * -invV;
#pragma unroll
    for (int j = 0; j < NK; ++j) {
        if (j >= nk) break;
        rk[1] = rk[1] + w1[j];
        rk[2] = rk[2] + w2[j];
        rk[3] = rk[3] + w3[j];
    }
    if (flags & kFlagStoreR) {
#pragma unroll
        for (int q = 0; q < 4; ++q) R[(long)q * nc + c] = rk[q];
    }
    const double dt = *dtp;
    double unew[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        double old;
        if (flags & kFlagSaveOld) {
            old = u[q];
            Uold[(long)q * nc + c] = u[q];
        } else {
            old = Uold[(long)q * nc + c];
        }
        unew[q] = c0 * old + c1 * (u[q] + dt * rk[q]);
        if (!OUT_AOS) Uout[(long)q * nc + c] = unew[q];
    }
    if (OUT_AOS) st4(Uout, c, unew);
    return (flags & kFlagSpeed) ? wave_speed(unew[0], unew[1], unew[2], unew[3], P.gamma) : 0.0;
}

// The same two cell phases with a cell's work spread over a group of four
// lanes (k = 0..3), for tiles that run one per CU: a lane then carries a
// quarter of a cell's dependent arithmetic.  The lanes first own edge slot
// k (its loads, edge value and Green-Gauss numerators), then conservative
// variable k; the arithmetic is that of the four-kernel stage's
// reconstruct_body and flux_update_body, with the tile's local vertex,
// edge and cell indices, so every value is the one-lane form's bit for bit.
// All four lanes of a group run together (the cross-lane reads need them).
__device__ __forceinline__ int sel4(const int (&a)[4], int k) {
    return (k == 0) ? a[0] : (k == 1) ? a[1] : (k == 2) ? a[2] : a[3];
}
__device__ __forceinline__ double sel4(const double (&a)[4], int k) {
    return (k == 0) ? a[0] : (k == 1) ? a[1] : (k == 2) ? a[2] : a[3];
}

template <int NK>
__device__ __forceinline__ void tile_cell_quad(const Rec &r, bool face, const Geom &G,
                                               const double2 *__restrict__ EP, const Phys &P,
                                               const double *__restrict__ U, const Lds &L, int i, int k) {
    const int c = r.c;
    const unsigned info = (unsigned)r.info;
    const int nk = info >> 4;
    const bool has_edge = k < nk;
    const bool is_left = (info >> k) & 1u;
    double u[4];
    ld4(U, c, u);
    const double vol = G.cvol[c], cx = G.ccx[c], cy = G.ccy[c];
    // this lane's edge slot
    double el = 0.0, dx = 0.0, dy = 0.0, tx[4] = {0.0, 0.0, 0.0, 0.0}, ty[4] = {0.0, 0.0, 0.0, 0.0};
    int ebc = 0;
    if (has_edge) {
        const int e = sel4(r.edge, k);
        const double2 *pe = EP + 4 * (long)e;
        const double2 n2 = pe[0], lb = pe[1], m2 = pe[2];
        el = lb.x;
        ebc = (char)__double2loint(lb.y);
        dx = m2.x - cx;
        dy = m2.y - cy;
        const int la = (k == 0) ? r.lv[0] : (k == 1) ? r.lv[2] : (k == 2) ? r.lv[4] : r.lv[6];
        const int lb2 = (k == 0) ? r.lv[1] : (k == 1) ? r.lv[3] : (k == 2) ? r.lv[5] : r.lv[7];
        double xa[4], xb[4], val[4];
        lds_ld4(L.vert + 4 * la, xa);
        lds_ld4(L.vert + 4 * lb2, xb);
#pragma unroll
        for (int q = 0; q < 4; ++q) val[q] = 0.5 * (xa[q] + xb[q]);
        if (tagged((char)ebc)) bc_state(P, (char)ebc, val[0], val[1], val[2], val[3]);
        double nx = n2.x, ny = n2.y;
        if (is_left) {
            nx = -nx;
            ny = -ny;
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            tx[q] = val[q] * el * nx;
            ty[q] = val[q] * el * ny;
        }
    }
    // lane k: variable k's terms of the edges in slot order, divided by the
    // volume and summed from +0.0
    quad_transpose(tx, k);
    quad_transpose(ty, k);
    double gxk = 0.0, gyk = 0.0;
#pragma unroll
    for (int j = 0; j < NK; ++j) {
        if (j >= nk) break;
        gxk += tx[j] / vol;
        gyk += ty[j] / vol;
    }
    if (face) {
        const double uk = sel4(u, k);
        // neighbour min / max of variable k in neighbour-list order
        double mn = uk, mx = uk;
#pragma unroll
        for (int j = 0; j < NK; ++j) {
            const int nb = r.nbr[j];
            const double v = (nb >= 0) ? U[4 * (long)nb + k] : uk;
            if (v < mn) mn = v;
            if (v > mx) mx = v;
        }
        const double dxs[4] = {from_lane<0>(dx), from_lane<1>(dx), from_lane<2>(dx), from_lane<3>(dx)};
        const double dys[4] = {from_lane<0>(dy), from_lane<1>(dy), from_lane<2>(dy), from_lane<3>(dy)};
        const int bcs[4] = {from_lane_i<0>(ebc), from_lane_i<1>(ebc), from_lane_i<2>(ebc), from_lane_i<3>(ebc)};
        const double veps = (P.limiter == 1) ? 0.0 : venkat_eps(vol, P.venkat_k);
        double al = 1.0;
#pragma unroll
        for (int j = 0; j < NK; ++j) {
            if (j >= nk) break;
            const double pred = uk + gxk * dxs[j] + gyk * dys[j];
            const double a = (P.limiter == 1) ? alpha_bj(uk, pred, mn, mx) : alpha_venkat_eps(uk, pred, mn, mx, veps);
            al = smin(al, a);
        }
        const double fsk = (k == 0) ? P.fs_rho : (k == 1) ? P.fs_ru : (k == 2) ? P.fs_rv : P.fs_E;
#pragma unroll
        for (int j = 0; j < NK; ++j) {
            if (j >= nk) break;
            const int le = r.le[j];
            if (le < 0) continue;
            const double f = uk + al * (gxk * dxs[j] + gyk * dys[j]);
            double *row = L.fs + 8 * le;
            if ((info >> j) & 1u) {
                row[k] = f;
                const char bcj = (char)bcs[j];
                if (tagged(bcj)) row[4 + k] = (bcj == 'i' || bcj == 'f') ? fsk : (bcj == 'w' && (k == 1 || k == 2)) ? -f : f;
            } else {
                row[4 + k] = f;
            }
        }
    }
    // viscous stress and heat flux (src/diffusion.cpp:227-240), the
    // divisions spread over the four lanes as in reconstruct_body
    const double gx0 = from_lane<0>(gxk), gy0 = from_lane<0>(gyk);
    const double gx3 = from_lane<3>(gxk), gy3 = from_lane<3>(gyk);
    const double px = quad_perm<0x99>(gxk), py = quad_perm<0x99>(gyk);
    const double mu = P.mu, gamma = P.gamma, kappa = P.kappa;
    const double rho = u[0], E = u[3];
    const double qn = (k == 0)   ? u[1]
                      : (k == 1) ? u[2]
                      : (k == 2) ? (gx3 * rho - E * gx0)
                                 : (gy3 * rho - E * gy0);
    const double qv = qn / ((k < 2) ? rho : rho * rho);
    const double uu = from_lane<0>(qv), vvel = from_lane<1>(qv), ex = from_lane<2>(qv), ey = from_lane<3>(qv);
    const bool xk = (k == 0 || k == 3);
    const double num = (xk ? px : py) - ((k & 1) ? vvel : uu) * (xk ? gx0 : gy0);
    const double d = num / rho;
    const double u_x = from_lane<0>(d), v_y = from_lane<1>(d), u_y = from_lane<2>(d), v_x = from_lane<3>(d);
    const double tk = (k & 1) ? (uu * u_y + vvel * v_y) : (uu * u_x + vvel * v_x);
    const double Tk = (gamma - 1.0) * (((k & 1) ? ey : ex) - tk);
    const double T_x = from_lane<0>(Tk), T_y = from_lane<1>(Tk);
    if (k != 0) return;
    const double t4[4] = {2 * mu * u_x - 2.0 / 3.0 * mu * (u_x + v_y), 2 * mu * v_y - 2.0 / 3.0 * mu * (u_x + v_y),
                          mu * (u_y + v_x), -kappa * T_x};
    lds_st4(L.tq + 4 * i, t4);
    L.tq[4 * L.max_gd + i] = -kappa * T_y;
}

// Owned cell i by its quad (flux_update_body's arithmetic); returns |u| + c
// of the new state (kFlagSpeed), else 0, in every lane of the group.
template <bool OUT_AOS, int NK>
__device__ __forceinline__ double tile_update_quad(const Rec &r, const TileOwn &o, const Topo &T, const Geom &G,
                                                   const double2 *__restrict__ EP, const Phys &P,
                                                   const double *__restrict__ U, double *__restrict__ Uout,
                                                   double *__restrict__ Uold, double *__restrict__ R,
                                                   const double *__restrict__ dtp, double c0, double c1, int flags,
                                                   const Lds &L, int k) {
    const int nc = T.nc, c = r.c;
    const unsigned info = (unsigned)r.info;
    const int nk = info >> 4;
    const bool has_edge = k < nk;
    const bool is_left = (info >> k) & 1u;
    const double vol = G.cvol[c];
    double u[4];
    ld4(U, c, u);
    double adv[4] = {0.0, 0.0, 0.0, 0.0}, w1 = 0.0, w2 = 0.0, w3 = 0.0;
    // one division for three quotients: lanes 0 / 1 / 2 form u, v, 1 / V
    const double qv = ((k == 0) ? u[1] : (k == 1) ? u[2] : 1.0) / ((k < 2) ? u[0] : vol);
    const double uu = from_lane<0>(qv), vv = from_lane<1>(qv), invV = from_lane<2>(qv);
    if (has_edge) {
        const int e = sel4(r.edge, k);
        const double2 *pe = EP + 4 * (long)e;
        const double2 n2 = pe[0], lt = pe[1];
        const double el = lt.x;
        const char ebc = (char)__double2loint(lt.y);
        const int le = (k == 0) ? r.le[0] : (k == 1) ? r.le[1] : (k == 2) ? r.le[2] : r.le[3];
        double F[4];
        lds_ld4(L.fs + 8 * le, F);
        // convective term of this edge, src/advection.cpp:453-456
        const double sign = is_left ? +1.0 : -1.0;
#pragma unroll
        for (int q = 0; q < 4; ++q) adv[q] = sign * F[q] * el;
        // viscous terms of this edge, src/diffusion.cpp:183-200
        double nx = n2.x, ny = n2.y;
        if (is_left) {
            nx = -nx;
            ny = -ny;
        }
        const int la = (k == 0) ? o.lo[0] : (k == 1) ? o.lo[2] : (k == 2) ? o.lo[4] : o.lo[6];
        const int lb = (k == 0) ? o.lo[1] : (k == 1) ? o.lo[3] : (k == 2) ? o.lo[5] : o.lo[7];
        double ta[4], tb[4], t[5];
        lds_ld4(L.vert + 4 * la, ta);
        lds_ld4(L.vert + 4 * lb, tb);
#pragma unroll
        for (int q = 0; q < 4; ++q) t[q] = 0.5 * (ta[q] + tb[q]);
        t[4] = 0.5 * (L.vert[4 * L.max_vo + la] + L.vert[4 * L.max_vo + lb]);
        if (ebc == 'f' || ebc == 'w') {
#pragma unroll
            for (int q = 0; q < 5; ++q) t[q] = 0.0;
        }
        const double txx = t[0], tyy = t[1], txy = t[2], qx = t[3], qy = t[4];
        w1 = txx * el * nx / vol + txy * el * ny / vol;
        w2 = txy * el * nx / vol + tyy * el * ny / vol;
        w3 = (uu * txx + vv * txy - qx) * el * nx / vol + (uu * txy + vv * tyy - qy) * el * ny / vol;
    }
    // lane k: variable k's terms in slot order (+0.0 from lanes without an
    // edge; the sum starts at +0.0, so it is never -0.0)
    quad_transpose(adv, k);
    double rk = 0.0;
#pragma unroll
    for (int j = 0; j < 4; ++j) rk += adv[j];
    rk = rk * -invV;
    double wv[4] = {0.0, w1, w2, w3};
    quad_transpose(wv, k);
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const double tt = rk + wv[j];
        rk = (k > 0 && j < nk) ? tt : rk;
    }
    if (flags & kFlagStoreR) R[(long)k * nc + c] = rk;
    const double dt = *dtp;
    const double uk = sel4(u, k);
    double old;
    if (flags & kFlagSaveOld) {
        old = uk;
        Uold[(long)k * nc + c] = uk;
    } else {
        old = Uold[(long)k * nc + c];
    }
    const double unew = c0 * old + c1 * (uk + dt * rk);
    if (OUT_AOS)
        Uout[4 * (long)c + k] = unew;
    else
        Uout[(long)k * nc + c] = unew;
    return (flags & kFlagSpeed)
               ? wave_speed(from_lane<0>(unew), from_lane<1>(unew), from_lane<2>(unew), from_lane<3>(unew), P.gamma)
               : 0.0;
}

// The largest of every lane's m (all >= 1e-8 and NaN-free, so the order
// of this tree does not matter) into *dst, by thread 0.
__device__ __forceinline__ void block_max_to(double m, double *dst) {
    __shared__ double wmax[kTileQuadThreads / 64];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) m = smax(m, __shfl_xor(m, off));
    if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        double b = wmax[0];
        for (int w = 1; w < (int)(blockDim.x >> 6); ++w) b = smax(b, wmax[w]);
        *dst = b;
    }
}

// One stage over the tile with head h, on the mesh (T, G, EP) with the
// physics P: reads U, writes Uout (and Uold / R as the flags say) and, with
// kFlagSpeed, the tile's max(|u| + c) into *speed_slot.
template <bool OUT_AOS, int NK, bool QUAD>
__device__ __forceinline__ void stage_tile(const Topo &T, const Geom &G, const Phys &P, const TileView &A,
                                           const TileHead &h, const double2 *__restrict__ EP,
                                           const double *__restrict__ U, double *__restrict__ Uout,
                                           double *__restrict__ Uold, double *__restrict__ R,
                                           const double *__restrict__ dtp, double c0, double c1, int flags,
                                           double *__restrict__ speed_slot) {
    extern __shared__ double4 lds_raw[];
    double *lds = reinterpret_cast<double *>(lds_raw);
    const Lds L{lds, lds + A.lds_tq, lds + A.lds_fs, A.max_vo, A.max_gd};
    const int tid = threadIdx.x, nth = blockDim.x;

    // 1. states of the tile's state vertices (the four-kernel stage's k_vertex_gather)
    for (int i = tid; i < h.nVg; i += nth) {
        const Gat r = load_gat(A.vg + h.vg0 + i);
        vertex_gather_ids<4, true>(
            T, G, U, P, r.kind, r.b, r.n, r.cid, [&] { return r.bc; },
            [&](const double(&out)[4]) { lds_st4(L.vert + 4 * i, out); });
    }
    __syncthreads();
    // 2. gradients and stresses of every gradient cell; limiter and face states of the face cells
    if (QUAD) {
        for (int i = tid >> 2; i < h.nGd; i += nth >> 2) {
            const Rec r = load_rec(A.cell + h.cell0 + i);
            tile_cell_quad<NK>(r, i < h.nF, G, EP, P, U, L, i, tid & 3);
        }
    } else {
        for (int i = tid; i < h.nGd; i += nth) {
            const Rec r = load_rec(A.cell + h.cell0 + i);
            tile_cell<NK>(r, i < h.nF, G, EP, P, T.nc, U, L, i);
        }
    }
    __syncthreads();
    // 3. HLLC once per tile edge (in place of its left face state), and the
    //    stresses of the owned cells' vertices (over the state vertices, no
    //    longer read)
    for (int i = tid; i < h.nE + h.nVo; i += nth) {
        if (i < h.nE) {
            const double2 n2 = A.en[h.e0 + i];
            double qL[4], qR[4], f[4];
            lds_ld4(L.fs + 8 * i, qL);
            lds_ld4(L.fs + 8 * i + 4, qR);
            hllc(qL, qR, n2.x, n2.y, P.gamma, f);
            lds_st4(L.fs + 8 * i, f);
        } else {
            const int j = i - h.nE;
            tile_vertex_stress(load_str(A.vo + h.vo0 + j), A.lsten, G, L, j);
        }
    }
    __syncthreads();
    // 4. sums and stage update of the owned cells
    double m = 1e-8;
    if (QUAD) {
        for (int i = tid >> 2; i < h.nO; i += nth >> 2) {
            const Rec r = load_rec(A.cell + h.cell0 + i);
            const double s = tile_update_quad<OUT_AOS, NK>(r, A.own[h.cell0 + i], T, G, EP, P, U, Uout, Uold, R,
                                                           dtp, c0, c1, flags, L, tid & 3);
            if (flags & kFlagSpeed) m = smax(m, s);
        }
    } else {
        for (int i = tid; i < h.nO; i += nth) {
            const Rec r = load_rec(A.cell + h.cell0 + i);
            const double s =
                tile_update<OUT_AOS, NK>(r, A.own[h.cell0 + i], T, G, EP, P, U, Uout, Uold, R, dtp, c0, c1, flags, L);
            if (flags & kFlagSpeed) m = smax(m, s);
        }
    }
    if (flags & kFlagSpeed) block_max_to(m, speed_slot);
}

// Workgroups go to the eight XCDs in turn; workgroup b runs tile
// x q + min(x, r) + b / 8 (x = b mod 8, ntiles = 8 q + r), so each XCD
// takes one contiguous run of the tile order, whose neighbouring tiles
// share rim cells, vertices and edges in that XCD's L2.
__device__ __forceinline__ int xcd_tile(int nt) {
    const int xcd = blockIdx.x & 7, q8 = nt >> 3, r8 = nt & 7;
    return xcd * q8 + min(xcd, r8) + (int)(blockIdx.x >> 3);
}

// At most 128 VGPRs (four waves per SIMD): the kernel keeps every tile of
// the mesh resident in one round, so occupancy is what hides its loads.
template <bool OUT_AOS, int NK>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void k_tile_stage(Topo T, Geom G, Phys P, TileView A,
                                                    const double *__restrict__ U, double *__restrict__ Uout,
                                                    double *__restrict__ Uold, double *__restrict__ R,
                                                    const double *__restrict__ dtp, double c0, double c1, int flags,
                                                    double *__restrict__ speed, StepBegin B) {
    if (P.gate && *P.gate == 0.0) return;  // uniform
    __shared__ double s_dt;
    if (flags & kFlagBeginDt) {
        // the step's start from the max(|u| + c) the previous step's last
        // stage folded, as k_step_begin forms it (the stop test's inputs and
        // the max do not change during this launch)
        if (P.gate && (!(B.scal[kScalT] < P.final_t) || B.scal[kScalConv] != 0.0)) {
            if (blockIdx.x == 0 && threadIdx.x == 0) B.scal[kScalActive] = 0.0;
            return;  // uniform
        }
        if (threadIdx.x == 0) {
            const double max_wave = B.scal[kScalSpeedMax];
            const double dt_adv = P.cfl * B.min_h / max_wave;
            const double dt_diff = 0.5 * (B.min_h * B.min_h) / P.mu;
            const double dt = smin(dt_adv, dt_diff);
            s_dt = dt;
            if (blockIdx.x == 0) {
                B.scal[kScalDt] = dt;
                if (B.hist) *B.hist = dt;
            }
        }
        __syncthreads();
        dtp = &s_dt;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        if (flags & kFlagAdvance) {
            // t is read by the next step's start only
            B.scal[kScalTPrev] = B.scal[kScalT];
            B.scal[kScalT] = B.scal[kScalT] + B.scal[kScalDt];
            if (P.gate) B.scal[kScalSteps] = B.scal[kScalSteps] + 1.0;
        }
        // the fold's start value, as k_step_begin's (read by the next step)
        if (flags & kFlagSpeedReset) B.scal[kScalSpeedMax] = 1e-8;
    }
    const int tile = xcd_tile(A.ntiles);
    const TileHead h = A.head[tile];
    stage_tile<OUT_AOS, NK, false>(T, G, P, A, h, reinterpret_cast<const double2 *>(A.epk), U, Uout, Uold, R, dtp, c0,
                                   c1, flags, speed + tile);
    if ((flags & kFlagSpeedMax) && threadIdx.x == 0) {
        // the tile's partial into the step's max: positive doubles order as
        // their bit patterns, so one 64-bit atomic max is exact and
        // order-free; smax(1e-8, .) skips a NaN partial as k_step_begin does
        const double m = smax(1e-8, speed[tile]);
        __hip_atomic_fetch_max(reinterpret_cast<unsigned long long *>(B.scal + kScalSpeedMax),
                               static_cast<unsigned long long>(__double_as_longlong(m)), __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
    }
}

// The same stage for meshes whose tiles run one per CU: 1,024 threads per
// tile, the cell phases four lanes per cell (tile_cell_quad,
// tile_update_quad), so that a CU's sixteen waves share the tile's cells.
template <bool OUT_AOS, int NK>
__global__ __launch_bounds__(kTileQuadThreads) void k_tile_stage_q(Topo T, Geom G, Phys P, TileView A,
                                                                   const double *__restrict__ U,
                                                                   double *__restrict__ Uout,
                                                                   double *__restrict__ Uold, double *__restrict__ R,
                                                                   const double *__restrict__ dtp, double c0,
                                                                   double c1, int flags,
                                                                   double *__restrict__ speed, StepBegin B) {
    if (P.gate && *P.gate == 0.0) return;  // uniform
    __shared__ double s_dt;
    if (flags & kFlagBegin) {
        // the step's start, as k_step_begin forms it: the max of the speed
        // partials (exact and order-free), the stop test on the time or the
        // residual flag (neither changes during this launch), dt
        double m = 1e-8;
        for (int i = threadIdx.x; i < B.nparts; i += blockDim.x) m = smax(m, B.parts[i]);
        __shared__ double s_max;
        block_max_to(m, &s_max);
        __syncthreads();
        const bool stop = P.gate && (!(B.scal[kScalT] < P.final_t) || B.scal[kScalConv] != 0.0);
        if (stop) {
            if (blockIdx.x == 0 && threadIdx.x == 0) B.scal[kScalActive] = 0.0;
            return;  // uniform
        }
        const double dt_adv = P.cfl * B.min_h / s_max;
        const double dt_diff = 0.5 * (B.min_h * B.min_h) / P.mu;
        const double dt = smin(dt_adv, dt_diff);
        if (threadIdx.x == 0) {
            s_dt = dt;
            if (blockIdx.x == 0) {
                B.scal[kScalDt] = dt;
                if (B.hist) *B.hist = dt;
            }
        }
        __syncthreads();
        dtp = &s_dt;
    }
    if ((flags & kFlagAdvance) && blockIdx.x == 0 && threadIdx.x == 0) {
        // t is read by the next step's start only
        B.scal[kScalTPrev] = B.scal[kScalT];
        B.scal[kScalT] = B.scal[kScalT] + B.scal[kScalDt];
        if (P.gate) B.scal[kScalSteps] = B.scal[kScalSteps] + 1.0;
    }
    const int tile = xcd_tile(A.ntiles);
    const TileHead h = A.head[tile];
    stage_tile<OUT_AOS, NK, true>(T, G, P, A, h, reinterpret_cast<const double2 *>(A.epk), U, Uout, Uold, R, dtp, c0,
                                  c1, flags, speed + tile);
}

// The same stage over the tiles of every member of a batch (tables
// concatenated, each head naming its member): member B's tiles read its
// state buffer B.Us[stage] and write B.Us[stage + 1 mod 3]; a member whose
// gate is closed skips its tiles.
template <int NK>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void kb_tile_stage(
    const Member *__restrict__ M, TileView A, int stage, double c0, double c1, int flags) {
    const TileHead h = A.head[xcd_tile(A.ntiles)];
    const Member &B = M[h.member];
    if (*B.P.gate == 0.0) return;  // uniform
    const Topo T = B.T;
    const Geom G = B.G;
    const Phys P = B.P;
    stage_tile<true, NK, false>(T, G, P, A, h, reinterpret_cast<const double2 *>(B.epk), B.Us[stage],
                         B.Us[stage == 2 ? 0 : stage + 1], B.Uold, B.R, B.scal + kScalDt, c0, c1, flags,
                         B.speed + h.slot);
}

// Speed partials of the state B.U (SoA) over the owned cells of every tile,
// in the slots the last stage's kFlagSpeed fills: the first dt of a batched
// tiled run.
__global__ __launch_bounds__(256) void kb_tile_speed(const Member *__restrict__ M, TileView A) {
    const TileHead h = A.head[blockIdx.x];
    const Member &B = M[h.member];
    const int nc = B.T.nc;
    const double *__restrict__ U = B.U;
    double m = 1e-8;
    for (int i = threadIdx.x; i < h.nO; i += blockDim.x) {
        const int c = A.cell[h.cell0 + i].c;
        m = smax(m, wave_speed(U[c], U[nc + c], U[2 * nc + c], U[3 * nc + c], B.P.gamma));
    }
    block_max_to(m, B.speed + h.slot);
}

// Packed edge geometry: (n_x, n_y), (L, tag), (mid_x, mid_y), pad -- the
// values of the geometry arrays, three 16-byte loads per edge instead of six
// scattered words.
__global__ __launch_bounds__(256) void k_pack_edges(Geom G, const char *__restrict__ ebc, int ne,
                                                    double *__restrict__ epk) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= ne) return;
    double2 *p = reinterpret_cast<double2 *>(epk) + 4 * (long)e;
    p[0] = make_double2(G.enx[e], G.eny[e]);
    p[1] = make_double2(G.elen[e], __hiloint2double(0, (int)ebc[e]));
    p[2] = make_double2(G.emx[e], G.emy[e]);
    p[3] = make_double2(0.0, 0.0);
}

__global__ __launch_bounds__(256) void k_state_to_aos(const double *__restrict__ U, int nc, double *__restrict__ A) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nc) return;
    const double u[4] = {U[c], U[(long)nc + c], U[2 * (long)nc + c], U[3 * (long)nc + c]};
    st4(A, c, u);
}

__global__ __launch_bounds__(256) void k_state_to_soa(const double *__restrict__ A, int nc, double *__restrict__ U) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nc) return;
    double u[4];
    ld4(A, c, u);
#pragma unroll
    for (int q = 0; q < 4; ++q) U[(long)q * nc + c] = u[q];
}

}  // namespace

// The stage kernels of cells of at most nk_max edges: the 256-thread ones
// (occupancy counted on them) and the 1,024-thread four-lane ones.
void tile_kernels(int nk_max, const void *(&f)[3], const void *(&fq)[2]) {
    if (nk_max <= 3) {
        f[0] = reinterpret_cast<const void *>(&k_tile_stage<true, 3>);
        f[1] = reinterpret_cast<const void *>(&k_tile_stage<false, 3>);
        f[2] = reinterpret_cast<const void *>(&kb_tile_stage<3>);
        fq[0] = reinterpret_cast<const void *>(&k_tile_stage_q<true, 3>);
        fq[1] = reinterpret_cast<const void *>(&k_tile_stage_q<false, 3>);
    } else {
        f[0] = reinterpret_cast<const void *>(&k_tile_stage<true, 4>);
        f[1] = reinterpret_cast<const void *>(&k_tile_stage<false, 4>);
        f[2] = reinterpret_cast<const void *>(&kb_tile_stage<4>);
        fq[0] = reinterpret_cast<const void *>(&k_tile_stage_q<true, 4>);
        fq[1] = reinterpret_cast<const void *>(&k_tile_stage_q<false, 4>);
    }
}

hipError_t launch_tile_stage(const Topo &T, const Geom &G, const Phys &P, const TileView &A, const double *U,
                             double *Uout, bool out_aos, double *Uold, double *R, const double *dtp, double c0,
                             double c1, int flags, double *speed, hipStream_t s, const StepBegin &B) {
    const dim3 grid(A.ntiles), block(kTileThreads);
    constexpr int one_lane_only = kFlagBeginDt | kFlagSpeedReset | kFlagSpeedMax;
    if (A.quad ? (flags & one_lane_only) != 0 : (flags & kFlagBegin) != 0) return hipErrorInvalidValue;
    if ((flags & (kFlagBegin | kFlagAdvance | one_lane_only)) && !B.scal) return hipErrorInvalidValue;
    if ((flags & kFlagSpeedMax) && (!(flags & kFlagSpeed) || !speed)) return hipErrorInvalidValue;
    if (A.quad) {
        const dim3 qblock(kTileQuadThreads);
        if (A.nk_max <= 3) {
            if (out_aos)
                hipLaunchKernelGGL((k_tile_stage_q<true, 3>), grid, qblock, A.lds_bytes, s, T, G, P, A, U, Uout, Uold,
                                   R, dtp, c0, c1, flags, speed, B);
            else
                hipLaunchKernelGGL((k_tile_stage_q<false, 3>), grid, qblock, A.lds_bytes, s, T, G, P, A, U, Uout,
                                   Uold, R, dtp, c0, c1, flags, speed, B);
        } else {
            if (out_aos)
                hipLaunchKernelGGL((k_tile_stage_q<true, 4>), grid, qblock, A.lds_bytes, s, T, G, P, A, U, Uout, Uold,
                                   R, dtp, c0, c1, flags, speed, B);
            else
                hipLaunchKernelGGL((k_tile_stage_q<false, 4>), grid, qblock, A.lds_bytes, s, T, G, P, A, U, Uout,
                                   Uold, R, dtp, c0, c1, flags, speed, B);
        }
        return hipGetLastError();
    }
    if (A.nk_max <= 3) {
        if (out_aos)
            hipLaunchKernelGGL((k_tile_stage<true, 3>), grid, block, A.lds_bytes, s, T, G, P, A, U, Uout, Uold, R, dtp,
                               c0, c1, flags, speed, B);
        else
            hipLaunchKernelGGL((k_tile_stage<false, 3>), grid, block, A.lds_bytes, s, T, G, P, A, U, Uout, Uold, R,
                               dtp, c0, c1, flags, speed, B);
    } else {
        if (out_aos)
            hipLaunchKernelGGL((k_tile_stage<true, 4>), grid, block, A.lds_bytes, s, T, G, P, A, U, Uout, Uold, R, dtp,
                               c0, c1, flags, speed, B);
        else
            hipLaunchKernelGGL((k_tile_stage<false, 4>), grid, block, A.lds_bytes, s, T, G, P, A, U, Uout, Uold, R,
                               dtp, c0, c1, flags, speed, B);
    }
    return hipGetLastError();
}

hipError_t launch_b_tile_stage(const Member *M, const TileView &A, int stage, double c0, double c1, int flags,
                               hipStream_t s) {
    const dim3 grid(A.ntiles), block(kTileThreads);
    if (A.nk_max <= 3)
        hipLaunchKernelGGL((kb_tile_stage<3>), grid, block, A.lds_bytes, s, M, A, stage, c0, c1, flags);
    else
        hipLaunchKernelGGL((kb_tile_stage<4>), grid, block, A.lds_bytes, s, M, A, stage, c0, c1, flags);
    return hipGetLastError();
}

hipError_t launch_b_tile_speed(const Member *M, const TileView &A, hipStream_t s) {
    hipLaunchKernelGGL(kb_tile_speed, dim3(A.ntiles), dim3(kTileThreads), 0, s, M, A);
    return hipGetLastError();
}

hipError_t launch_pack_edges(const Geom &G, const char *ebc, int ne, double *epk, hipStream_t s) {
    hipLaunchKernelGGL(k_pack_edges, dim3((ne + 255) / 256), dim3(256), 0, s, G, ebc, ne, epk);
    return hipGetLastError();
}

hipError_t launch_state_to_aos(const double *U, int nc, double *A, hipStream_t s) {
    hipLaunchKernelGGL(k_state_to_aos, dim3((nc + 255) / 256), dim3(256), 0, s, U, nc, A);
    return hipGetLastError();
}

hipError_t launch_state_to_soa(const double *A, int nc, double *U, hipStream_t s) {
    hipLaunchKernelGGL(k_state_to_soa, dim3((nc + 255) / 256), dim3(256), 0, s, A, nc, U);
    return hipGetLastError();
}

hipError_t tile_stage_attributes(size_t lds_bytes, int nk_max, int *blocks_per_cu, int *vgprs, int *quad_per_cu) {
    // every context may launch up to the device's LDS per workgroup, less
    // the kernel's static LDS; a failed query leaves no error behind
    const void *f[3], *fq[2];
    tile_kernels(nk_max, f, fq);
    if (quad_per_cu) *quad_per_cu = 0;
    int dev = 0, lds_max = 0;
    *blocks_per_cu = 0;
    if (vgprs) *vgprs = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, dev);
    int bpc = 1 << 30;
    for (const void *k : f) {
        hipFuncAttributes fa{};
        if (e == hipSuccess) e = hipFuncGetAttributes(&fa, k);
        const int dyn_max = lds_max - (int)fa.sharedSizeBytes;
        if (e != hipSuccess || lds_bytes > (size_t)dyn_max) {
            bpc = 0;
            break;
        }
        if (vgprs) *vgprs = std::max(*vgprs, fa.numRegs);
        int b = 0;
        e = hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, dyn_max);
        if (e == hipSuccess) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, k, kTileThreads, lds_bytes);
        bpc = std::min(bpc, b);
    }
    int qpc = bpc > 0 ? 1 << 30 : 0;
    for (const void *k : fq) {
        if (e != hipSuccess || qpc == 0) break;
        hipFuncAttributes fa{};
        e = hipFuncGetAttributes(&fa, k);
        const int dyn_max = lds_max - (int)fa.sharedSizeBytes;
        if (e != hipSuccess || lds_bytes > (size_t)dyn_max) {
            qpc = 0;
            break;
        }
        int b = 0;
        e = hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, dyn_max);
        if (e == hipSuccess) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, k, kTileQuadThreads, lds_bytes);
        qpc = std::min(qpc, b);
    }
    if (e != hipSuccess) {
        (void)hipGetLastError();
        bpc = 0;
        qpc = 0;
    }
    *blocks_per_cu = bpc;
    if (quad_per_cu) *quad_per_cu = qpc;
    return e;
}

}  // namespace o2f
