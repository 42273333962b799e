// Tangent of the residual on gfx950: R(U, G) and its directional
// derivative dR = (dR/dG) dG + (dR/dU) dU along a geometry tangent dG (cell
// areas and centres, edge lengths, normals and midpoints, stencil weights,
// derived on the host from a vertex displacement) and a state tangent dU
// (either may be absent: null pointers are zero).
//
// Every quantity of the residual (computeResidual,
// src/shape_optimization.cpp:171-189: advection src/advection.cpp:26-466,
// diffusion src/diffusion.cpp:6-243) is carried as a dual number (value,
// derivative) through the same expressions, in the same order, as the
// forward kernels, so the value parts are the forward residual bit for bit
// and the derivative parts are exact (branches follow the values, as the
// forward takes them).  One work-item per vertex, cell or edge, simple
// loops: this runs once per shape parameter, not per step.
//
//   kt_vertex_state  per vertex: stencil gather of U with moving weights, BC
//   kt_reconstruct   per cell:   edge values, Green-Gauss, limiter, face
//                                states and ghosts, stresses / heat flux
//   kt_vertex_visc   per vertex: stencil gather of the stresses
//   kt_edge_flux     per edge:   HLLC with a moving normal
//   kt_flux_update   per cell:   convective and viscous sums -> R, dR
//   kt_objective     one block:  wall forces -> L, D and their derivatives
//                                (terms in parallel, summed in wall order)
#include "o2f_kernels.hpp"

#include "o2f_device.hpp"

namespace o2f {

namespace {

__device__ __forceinline__ Dual gd(const double *a, const double *da, long i) {
    return da ? Dual(a[i], da[i]) : Dual(a[i]);
}

// Stencil gather of NV quantities at vertex v (gather_vertex of the forward
// kernels): weights w0 a0 + w1 a1 + ... left to right, or the two-cell
// average, the copy, or zero; then the vertex boundary condition.
template <int NV, class Src>
__device__ __forceinline__ void gather_t(const Topo &T, const Geom &G, const Geom &dG, int v, const Src &val,
                                         Dual (&out)[NV]) {
    const uint8_t kind = T.skind[v];
    const int b = T.sptr[v], n = T.sptr[v + 1] - b;
    if (kind == kStencilWsum) {
        for (int k = 0; k < n; ++k) {
            const int c = T.scell[b + k];
            const Dual w = gd(G.sw, dG.sw, b + k);
            for (int q = 0; q < NV; ++q) out[q] = (k == 0) ? w * val(c, q) : out[q] + w * val(c, q);
        }
    } else if (kind == kStencilAvg2) {
        const int c0 = T.scell[b], c1 = T.scell[b + 1];
        for (int q = 0; q < NV; ++q) out[q] = 0.5 * (val(c0, q) + val(c1, q));
    } else if (kind == kStencilCopy) {
        const int c0 = T.scell[b];
        for (int q = 0; q < NV; ++q) out[q] = val(c0, q);
    } else {
        for (int q = 0; q < NV; ++q) out[q] = Dual(0.0);
    }
}

// the cell state (SoA [q][nc]) with its tangent (null: zero)
struct StateSrc {
    const double *U, *dU;
    int nc;
    __device__ Dual operator()(int c, int q) const { return gd(U, dU, (long)q * nc + c); }
};
// the cell stresses (AoS 4 + trailing block) with their derivatives
struct StressSrc {
    const double *tq, *dtq;
    int nc;
    __device__ Dual operator()(int c, int q) const {
        const long i = q < 4 ? 4 * (long)c + q : 4 * (long)nc + c;
        return Dual(tq[i], dtq[i]);
    }
};

__device__ __forceinline__ void st4d(double *v, double *d, long i, const Dual (&x)[4]) {
    for (int q = 0; q < 4; ++q) {
        v[4 * i + q] = x[q].v;
        d[4 * i + q] = x[q].d;
    }
}
__device__ __forceinline__ void ld4d(const double *v, const double *d, long i, Dual (&x)[4]) {
    for (int q = 0; q < 4; ++q) x[q] = Dual(v[4 * i + q], d[4 * i + q]);
}

__global__ __launch_bounds__(256) void kt_vertex_state(Topo T, Geom G, Geom dG, Phys P, const double *__restrict__ U,
                                                       const double *__restrict__ dU, double *__restrict__ vU,
                                                       double *__restrict__ dvU) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= T.nv) return;
    const StateSrc src{U, dU, T.nc};
    Dual out[4];
    gather_t<4>(T, G, dG, v, src, out);
    const char bc = T.vbc[v];
    if (tagged(bc)) bc_state_g(P, bc, out);
    st4d(vU, dvU, v, out);
}

__global__ __launch_bounds__(256) void kt_vertex_visc(Topo T, Geom G, Geom dG, Phys P, const double *__restrict__ tq,
                                                      const double *__restrict__ dtq, double *__restrict__ vtq,
                                                      double *__restrict__ dvtq) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= T.nv) return;
    const StressSrc src{tq, dtq, T.nc};
    Dual out[5];
    gather_t<5>(T, G, dG, v, src, out);
    const char bc = T.vbc[v];
    if (bc == 'f' || bc == 'w')  // apply_diffusion_bc (src/boundary.cpp:101-127)
        for (int q = 0; q < 5; ++q) out[q] = Dual(0.0);
    Dual o4[4] = {out[0], out[1], out[2], out[3]};
    st4d(vtq, dvtq, v, o4);
    vtq[4 * (long)T.nv + v] = out[4].v;
    dvtq[4 * (long)T.nv + v] = out[4].d;
}

// Edge values + Green-Gauss (src/advection.cpp:136-188), min / max,
// limiter, face states and ghosts (:232-321, src/boundary.cpp:48-80), cell
// stress / heat flux (src/diffusion.cpp:223-241).
__global__ __launch_bounds__(256) void kt_reconstruct(Topo T, Geom G, Geom dG, Phys P, const double *__restrict__ U,
                                                      const double *__restrict__ dU, const double *__restrict__ vU,
                                                      const double *__restrict__ dvU,
                                                      double *__restrict__ fL, double *__restrict__ dfL,
                                                      double *__restrict__ fR, double *__restrict__ dfR,
                                                      double *__restrict__ tq, double *__restrict__ dtq) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= T.nc) return;
    const int nc = T.nc;
    const unsigned info = T.cinfo[c];
    const int nk = info >> 4;
    Dual u[4];
    for (int q = 0; q < 4; ++q) u[q] = gd(U, dU, (long)q * nc + c);
    const Dual vol = gd(G.cvol, dG.cvol, c), cx = gd(G.ccx, dG.ccx, c), cy = gd(G.ccy, dG.ccy, c);
    Dual gx[4], gy[4];
    for (int q = 0; q < 4; ++q) gx[q] = gy[q] = Dual(0.0);
    for (int k = 0; k < nk; ++k) {
        const int e = T.cedge[4 * c + k];
        const int va = T.cevx[8 * c + 2 * k], vb = T.cevx[8 * c + 2 * k + 1];
        Dual xa[4], xb[4], val[4];
        ld4d(vU, dvU, va, xa);
        ld4d(vU, dvU, vb, xb);
        for (int q = 0; q < 4; ++q) val[q] = 0.5 * (xa[q] + xb[q]);
        const char ebc = T.ebc[e];
        if (tagged(ebc)) bc_state_g(P, ebc, val);
        const Dual el = gd(G.elen, dG.elen, e);
        Dual nx = gd(G.enx, dG.enx, e), ny = gd(G.eny, dG.eny, e);
        if ((info >> k) & 1u) {
            nx = -nx;
            ny = -ny;
        }
        for (int q = 0; q < 4; ++q) {
            gx[q] = gx[q] + val[q] * el * nx / vol;
            gy[q] = gy[q] + val[q] * el * ny / vol;
        }
    }
    Dual mn[4], mx[4];
    for (int q = 0; q < 4; ++q) mn[q] = mx[q] = u[q];
    for (int j = 0; j < 4; ++j) {
        const int nb = T.cnbr[4 * c + j];
        for (int q = 0; q < 4; ++q) {
            const Dual v = (nb >= 0) ? gd(U, dU, (long)q * nc + nb) : u[q];
            if (v < mn[q]) mn[q] = v;
            if (v > mx[q]) mx[q] = v;
        }
    }
    Dual al[4];
    for (int q = 0; q < 4; ++q) al[q] = Dual(1.0);
    for (int k = 0; k < nk; ++k) {
        const int e = T.cedge[4 * c + k];
        const Dual dx = gd(G.emx, dG.emx, e) - cx, dy = gd(G.emy, dG.emy, e) - cy;
        for (int q = 0; q < 4; ++q) {
            const Dual pred = u[q] + gx[q] * dx + gy[q] * dy;
            const Dual a = (P.limiter == 1) ? alpha_bj_g(u[q], pred, mn[q], mx[q])
                                            : alpha_venkat_g(u[q], pred, mn[q], mx[q], vol, P.venkat_k);
            al[q] = smin_t(al[q], a);
        }
    }
    for (int k = 0; k < nk; ++k) {
        const int e = T.cedge[4 * c + k];
        const Dual dx = gd(G.emx, dG.emx, e) - cx, dy = gd(G.emy, dG.emy, e) - cy;
        Dual f[4];
        for (int q = 0; q < 4; ++q) f[q] = u[q] + al[q] * (gx[q] * dx + gy[q] * dy);
        if ((info >> k) & 1u) {
            st4d(fL, dfL, e, f);
            const char ebc = T.ebc[e];
            if (tagged(ebc)) {
                Dual g[4];
                bc_ghost_g(P, ebc, f, g);
                st4d(fR, dfR, e, g);
            }
        } else {
            st4d(fR, dfR, e, f);
        }
    }
    Dual t5[5];
    cell_stress_t<Dual>(u, gx, gy, P.mu, P.gamma, P.kappa, t5);
    const Dual t4[4] = {t5[0], t5[1], t5[2], t5[3]};
    st4d(tq, dtq, c, t4);
    tq[4 * (long)nc + c] = t5[4].v;
    dtq[4 * (long)nc + c] = t5[4].d;
}

__global__ __launch_bounds__(256) void kt_edge_flux(Topo T, Geom G, Geom dG, Phys P, const double *__restrict__ fL,
                                                    const double *__restrict__ dfL, const double *__restrict__ fR,
                                                    const double *__restrict__ dfR, double *__restrict__ Fe,
                                                    double *__restrict__ dFe) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= T.ne) return;
    Dual qL[4], qR[4], F[4];
    ld4d(fL, dfL, e, qL);
    ld4d(fR, dfR, e, qR);
    hllc_t<Dual, Dual>(qL, qR, gd(G.enx, dG.enx, e), gd(G.eny, dG.eny, e), P.gamma, F);
    st4d(Fe, dFe, e, F);
}

// Convective sum R = -(1/V) sum_e s F_e L_e (src/advection.cpp:439-464),
// then the viscous divergence (src/diffusion.cpp:145-201), edges in slot
// (ascending id) order.
__global__ __launch_bounds__(256) void kt_flux_update(Topo T, Geom G, Geom dG, Phys P, const double *__restrict__ U,
                                                      const double *__restrict__ dU, const double *__restrict__ Fe,
                                                      const double *__restrict__ dFe,
                                                      const double *__restrict__ vtq, const double *__restrict__ dvtq,
                                                      double *__restrict__ R, double *__restrict__ dR) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= T.nc) return;
    const int nc = T.nc, nv = T.nv;
    const unsigned info = T.cinfo[c];
    const int nk = info >> 4;
    Dual u[4];
    for (int q = 0; q < 4; ++q) u[q] = gd(U, dU, (long)q * nc + c);
    const Dual vol = gd(G.cvol, dG.cvol, c);
    Dual r[4];
    for (int q = 0; q < 4; ++q) r[q] = Dual(0.0);
    for (int k = 0; k < nk; ++k) {
        const int e = T.cedge[4 * c + k];
        Dual F[4];
        ld4d(Fe, dFe, e, F);
        const Dual el = gd(G.elen, dG.elen, e);
        const double sign = ((info >> k) & 1u) ? +1.0 : -1.0;
        for (int q = 0; q < 4; ++q) r[q] = r[q] + sign * F[q] * el;
    }
    const Dual invV = 1.0 / vol;
    for (int q = 0; q < 4; ++q) r[q] = r[q] * -invV;
    const Dual uu = u[1] / u[0], vv = u[2] / u[0];
    for (int k = 0; k < nk; ++k) {
        const int e = T.cedge[4 * c + k];
        const int va = T.cevx[8 * c + 2 * k], vb = T.cevx[8 * c + 2 * k + 1];
        const Dual el = gd(G.elen, dG.elen, e);
        Dual nx = gd(G.enx, dG.enx, e), ny = gd(G.eny, dG.eny, e);
        if ((info >> k) & 1u) {
            nx = -nx;
            ny = -ny;
        }
        Dual t[5];
        for (int q = 0; q < 4; ++q)
            t[q] = 0.5 * (Dual(vtq[4 * (long)va + q], dvtq[4 * (long)va + q]) +
                          Dual(vtq[4 * (long)vb + q], dvtq[4 * (long)vb + q]));
        t[4] = 0.5 * (Dual(vtq[4 * (long)nv + va], dvtq[4 * (long)nv + va]) +
                      Dual(vtq[4 * (long)nv + vb], dvtq[4 * (long)nv + vb]));
        const char ebc = T.ebc[e];
        if (ebc == 'f' || ebc == 'w')
            for (int q = 0; q < 5; ++q) t[q] = Dual(0.0);
        const Dual txx = t[0], tyy = t[1], txy = t[2], qx = t[3], qy = t[4];
        const Dual w1 = txx * el * nx / vol + txy * el * ny / vol;
        const Dual w2 = txy * el * nx / vol + tyy * el * ny / vol;
        const Dual w3 = (uu * txx + vv * txy - qx) * el * nx / vol + (uu * txy + vv * tyy - qy) * el * ny / vol;
        r[1] = r[1] + w1;
        r[2] = r[2] + w2;
        r[3] = r[3] + w3;
    }
    for (int q = 0; q < 4; ++q) {
        R[(long)q * nc + c] = r[q].v;
        dR[(long)q * nc + c] = r[q].d;
    }
}

// Wall forces of computeLiftDragCoefficients / computeDirectObjectiveTerm
// (src/compute.cpp:92-132, src/shape_optimization.cpp:893-939): each
// work-item forms one wall edge's terms, one work-item adds them in
// wall-list order (the order of k_wall_forces): out = L, D, dL, dD.
__global__ __launch_bounds__(kRedBlock) void kt_objective(Geom G, Geom dG, const int *__restrict__ wall, int nw,
                                                          const double *__restrict__ fL,
                                                          const double *__restrict__ dfL, double gamma,
                                                          int skip_x, double *__restrict__ out) {
    __shared__ double tl[kRedBlock], td[kRedBlock], tdl[kRedBlock], tdd[kRedBlock];
    Dual L(0.0), D(0.0);
    for (int base = 0; base < nw; base += kRedBlock) {
        const int k = base + threadIdx.x;
        Dual l(0.0), d(0.0);
        bool use = false;
        if (k < nw) {
            const int e = wall[k];
            use = !(skip_x && G.emx[e] > 1.3);
            const Dual nx = gd(G.enx, dG.enx, e), ny = gd(G.eny, dG.eny, e), A = gd(G.elen, dG.elen, e);
            Dual q[4];
            ld4d(fL, dfL, e, q);
            const Dual p = (gamma - 1.0) * (q[3] - 0.5 * (q[1] * q[1] + q[2] * q[2]) / q[0]);
            const double txx = 0.0, txy = 0.0, tyy = 0.0;
            l = (-p * ny + txy * nx + tyy * ny) * A;
            d = (-p * nx + txx * nx + txy * ny) * A;
        }
        tl[threadIdx.x] = use ? l.v : 0.0;
        td[threadIdx.x] = use ? d.v : 0.0;
        tdl[threadIdx.x] = use ? l.d : 0.0;
        tdd[threadIdx.x] = use ? d.d : 0.0;
        __syncthreads();
        // the serial sum in wall order; a skipped edge adds +0.0, which
        // changes nothing (the sums start at +0.0 and so are never -0.0),
        // so the loop has no branch and its LDS reads run ahead
        if (threadIdx.x == 0) {
            const int m = min(kRedBlock, nw - base);
#pragma unroll 8
            for (int j = 0; j < m; ++j) {
                L = L + Dual(tl[j], tdl[j]);
                D = D + Dual(td[j], tdd[j]);
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[0] = L.v;
        out[1] = D.v;
        out[2] = L.d;
        out[3] = D.d;
    }
}

// One thread per listed entry: cells [0, ncell), edges next, stencil
// weights last.
__global__ __launch_bounds__(256) void kt_scatter_geometry(GeomW g, int ncell, const int *__restrict__ cell,
                                                           const double *__restrict__ cval, int nedge,
                                                           const int *__restrict__ edge,
                                                           const double *__restrict__ eval, int nsw,
                                                           const int *__restrict__ swi,
                                                           const double *__restrict__ swv) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < ncell) {
        const int c = cell[i];
        g.cvol[c] = cval[3 * i];
        g.ccx[c] = cval[3 * i + 1];
        g.ccy[c] = cval[3 * i + 2];
        return;
    }
    i -= ncell;
    if (i < nedge) {
        const int e = edge[i];
        g.elen[e] = eval[5 * i];
        g.enx[e] = eval[5 * i + 1];
        g.eny[e] = eval[5 * i + 2];
        g.emx[e] = eval[5 * i + 3];
        g.emy[e] = eval[5 * i + 4];
        return;
    }
    i -= nedge;
    if (i < nsw) g.sw[swi[i]] = swv[i];
}

inline int nblk(long n) { return static_cast<int>((n + 255) / 256); }

}  // namespace

hipError_t launch_scatter_geometry(const GeomW &g, int ncell, const int *cell, const double *cval, int nedge,
                                   const int *edge, const double *eval, int nsw, const int *swi, const double *swv,
                                   hipStream_t s) {
    const long n = (long)ncell + nedge + nsw;
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(kt_scatter_geometry, dim3(nblk(n)), dim3(256), 0, s, g, ncell, cell, cval, nedge, edge, eval,
                       nsw, swi, swv);
    return hipGetLastError();
}

hipError_t launch_residual_tangent(const Topo &T, const Geom &G, const Geom &dG, const Phys &P, const double *U,
                                   const double *dU, double *vU, double *fL, double *fR, double *tq, double *vtq,
                                   double *Fe, double *R, const TanWork &W, hipStream_t s) {
    hipLaunchKernelGGL(kt_vertex_state, dim3(nblk(T.nv)), dim3(256), 0, s, T, G, dG, P, U, dU, vU, W.dvU);
    hipLaunchKernelGGL(kt_reconstruct, dim3(nblk(T.nc)), dim3(256), 0, s, T, G, dG, P, U, dU, vU, W.dvU, fL, W.dfL,
                       fR, W.dfR, tq, W.dtq);
    hipLaunchKernelGGL(kt_vertex_visc, dim3(nblk(T.nv)), dim3(256), 0, s, T, G, dG, P, tq, W.dtq, vtq, W.dvtq);
    hipLaunchKernelGGL(kt_edge_flux, dim3(nblk(T.ne)), dim3(256), 0, s, T, G, dG, P, fL, W.dfL, fR, W.dfR, Fe, W.dFe);
    hipLaunchKernelGGL(kt_flux_update, dim3(nblk(T.nc)), dim3(256), 0, s, T, G, dG, P, U, dU, Fe, W.dFe, vtq, W.dvtq,
                       R, W.dR);
    return hipGetLastError();
}

hipError_t launch_objective_tangent(const Geom &G, const Geom &dG, const int *wall, int nw, const double *fL,
                                    const double *dfL, double gamma, int skip_x, double *out, hipStream_t s) {
    hipLaunchKernelGGL(kt_objective, dim3(1), dim3(kRedBlock), 0, s, G, dG, wall, nw, fL, dfL, gamma, skip_x, out);
    return hipGetLastError();
}

}  // namespace o2f
