// gfx950 kernels of the residual / SSP-RK3 / adjoint path.
//
// FP64 throughout, one work-item per vertex, cell or edge, no MFMA (the
// flux path is gather-bound, HBM / Infinity-Cache bound).  Compiled with
// -ffp-contract=off: every expression rounds exactly as the reference's C++,
// so results match the CPU restatement bit for bit wherever the summation
// order is the same.  Scatter-free: every cell gathers its own sums in
// ascending edge id, which makes the results deterministic and independent
// of scheduling (the reference's `omp atomic` scatter is not).
#include "o2f_kernels.hpp"

#include <cfloat>

namespace o2f {

// std::min / std::max / std::clamp semantics (NaN handling included).
__device__ __forceinline__ double smin(double a, double b) { return (b < a) ? b : a; }
__device__ __forceinline__ double smax(double a, double b) { return (a < b) ? b : a; }
__device__ __forceinline__ double sclamp(double v, double lo, double hi) {
    return (v < lo) ? lo : (hi < v) ? hi : v;
}
__device__ __forceinline__ bool tagged(char b) {
    return b == 'i' || b == 'o' || b == 'f' || b == 'w';
}

// apply_inflow_outflow_bc, src/boundary.cpp:7-45 (freestream state precomputed
// on the host with the same expression order).
__device__ __forceinline__ void bc_state(const Phys &P, char bc, double &r, double &ru, double &rv,
                                         double &E) {
    if (bc == 'i' || bc == 'f') {
        r = P.fs_rho;
        ru = P.fs_ru;
        rv = P.fs_rv;
        E = P.fs_E;
    } else if (bc == 'w') {
        const double kinetic_old = 0.5 * (ru * ru + rv * rv) / r;
        const double e_internal = E - kinetic_old;
        ru = 0.0;
        rv = 0.0;
        E = e_internal;
    }
}

// apply_InterfaceStates_boundary_conditions, src/boundary.cpp:48-80.
__device__ __forceinline__ void bc_ghost(const Phys &P, char bc, const double *L, double *R) {
    R[0] = L[0];
    R[1] = L[1];
    R[2] = L[2];
    R[3] = L[3];
    if (bc == 'i' || bc == 'o' || bc == 'f') {
        bc_state(P, bc, R[0], R[1], R[2], R[3]);
    } else if (bc == 'w') {
        R[1] = -L[1];
        R[2] = -L[2];
    }
}

// compute_hllc_flux, src/advection.cpp:325-412.  (enx, eny) is the stored
// normal (into the left cell); the flux is taken along -n.
__device__ __forceinline__ void hllc(const double *qL, const double *qR, double enx, double eny,
                                     double gamma, double *F) {
    const double nx = -enx, ny = -eny;
    const double rho_L = qL[0];
    const double u_L = (qL[1] * (nx) + qL[2] * ny) / rho_L;
    const double v_L = (qL[1] * (-ny) + qL[2] * nx) / rho_L;
    const double E_L = qL[3];
    const double p_L = (gamma - 1.0) * (E_L - 0.5 * rho_L * (u_L * u_L + v_L * v_L));
    const double c_L = sqrt(gamma * p_L / rho_L);
    const double rho_R = qR[0];
    const double u_R = (qR[1] * (nx) + qR[2] * ny) / rho_R;
    const double v_R = (qR[1] * (-ny) + qR[2] * nx) / rho_R;
    const double E_R = qR[3];
    const double p_R = (gamma - 1.0) * (E_R - 0.5 * rho_R * (u_R * u_R + v_R * v_R));
    const double c_R = sqrt(gamma * p_R / rho_R);

    const double S_L = smin(u_L - c_L, u_R - c_R);
    const double S_R = smax(u_L + c_L, u_R + c_R);
    const double numer = p_R - p_L + rho_L * u_L * (S_L - u_L) - rho_R * u_R * (S_R - u_R);
    const double denom = rho_L * (S_L - u_L) - rho_R * (S_R - u_R);
    const double S_star = numer / denom;
    const double p_star = p_L + rho_L * (S_L - u_L) * (S_star - u_L);
    double f0, fu, fv, fE;
    if (S_L >= 0.0) {
        f0 = rho_L * u_L;
        fu = rho_L * u_L * u_L + p_L;
        fv = rho_L * u_L * v_L;
        fE = (E_L + p_L) * u_L;
    } else if (S_L <= 0.0 && S_star >= 0.0) {
        const double rs = rho_L * (S_L - u_L) / (S_L - S_star);
        const double rvs = rs * v_L;
        const double Es = ((S_L - u_L) * E_L - p_L * u_L + p_star * S_star) / (S_L - S_star);
        const double rus = rs * S_star;
        f0 = rus;
        fu = rus * rus / rs + p_star;
        fv = rvs * rus / rs;
        fE = (Es + p_star) * rus / rs;
    } else if (S_star <= 0.0 && S_R >= 0.0) {
        const double rs = rho_R * (S_R - u_R) / (S_R - S_star);
        const double rvs = rs * v_R;
        const double Es = ((S_R - u_R) * E_R - p_R * u_R + p_star * S_star) / (S_R - S_star);
        const double rus = rs * S_star;
        f0 = rus;
        fu = rus * rus / rs + p_star;
        fv = rvs * rus / rs;
        fE = (Es + p_star) * rus / rs;
    } else {
        f0 = rho_R * u_R;
        fu = rho_R * u_R * u_R + p_R;
        fv = rho_R * u_R * v_R;
        fE = (E_R + p_R) * u_R;
    }
    F[0] = f0;
    F[1] = (fu * nx + fv * (-ny));
    F[2] = (fu * ny + fv * (nx));
    F[3] = fE;
}

// ------------------------------------------------------------------------
// Cell -> vertex gather with precomputed stencils
// (src/advection.cpp:26-131 for the state, src/diffusion.cpp:28-140 for the
// viscous quantities).  NV = 4 state variables with the state BC, or NV = 5
// viscous quantities with apply_diffusion_bc (src/boundary.cpp:101-127).
// ------------------------------------------------------------------------
constexpr int kStencilSlots = 8;  // stencil entries loaded together (longer ones loop)

template <int NV>
__global__ __launch_bounds__(256) void k_vertex_gather(Topo T, Geom G, const double *__restrict__ src,
                                                       double *__restrict__ dst, Phys P) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= T.nv || (P.gate && *P.gate == 0.0)) return;
    const int nc = T.nc;
    const uint8_t kind = T.skind[v];
    const int b = T.sptr[v], n = T.sptr[v + 1] - b;
    // All stencil ids and weights first, then all values: two dependent
    // load levels instead of one per stencil entry.
    int cid[kStencilSlots];
    double w[kStencilSlots], val[kStencilSlots][NV];
#pragma unroll
    for (int k = 0; k < kStencilSlots; ++k) {
        cid[k] = (k < n) ? T.scell[b + k] : 0;
        w[k] = (k < n) ? G.sw[b + k] : 0.0;
    }
#pragma unroll
    for (int k = 0; k < kStencilSlots; ++k)
        if (k < n) {
#pragma unroll
            for (int q = 0; q < NV; ++q) val[k][q] = src[q * nc + cid[k]];
        }
    double out[NV];
    if (kind == kStencilWsum) {
        // w0*a0 + w1*a1 + ... left to right (src/advection.cpp:47-50, :101-104)
#pragma unroll
        for (int q = 0; q < NV; ++q) out[q] = w[0] * val[0][q];
#pragma unroll
        for (int k = 1; k < kStencilSlots; ++k)
            if (k < n) {
#pragma unroll
                for (int q = 0; q < NV; ++q) out[q] = out[q] + w[k] * val[k][q];
            }
        for (int k = kStencilSlots; k < n; ++k) {
            const int c = T.scell[b + k];
            const double wk = G.sw[b + k];
#pragma unroll
            for (int q = 0; q < NV; ++q) out[q] = out[q] + wk * src[q * nc + c];
        }
    } else if (kind == kStencilAvg2) {
#pragma unroll
        for (int q = 0; q < NV; ++q) out[q] = 0.5 * (val[0][q] + val[1][q]);
    } else if (kind == kStencilCopy) {
#pragma unroll
        for (int q = 0; q < NV; ++q) out[q] = val[0][q];
    } else {
#pragma unroll
        for (int q = 0; q < NV; ++q) out[q] = 0.0;
    }
    const char bc = T.vbc[v];
    if (NV == 4) {
        if (tagged(bc)) bc_state(P, bc, out[0], out[1], out[2], out[3]);
    } else {
        if (bc == 'f' || bc == 'w') {
#pragma unroll
            for (int q = 0; q < NV; ++q) out[q] = 0.0;
        }
    }
#pragma unroll
    for (int q = 0; q < NV; ++q) dst[q * T.nv + v] = out[q];
}

// ------------------------------------------------------------------------
// Cell reconstruction: edge values + Green-Gauss gradient
// (src/advection.cpp:136-188), neighbour min/max, limiter and face states
// with boundary ghosts (src/advection.cpp:232-321), and the cell stress /
// heat flux of computeDiffusionTerm (src/diffusion.cpp:223-241), fused.
// ------------------------------------------------------------------------
__device__ __forceinline__ double alpha_venkat(double phi_i, double phi_pred, double phi_min,
                                               double phi_max, double cellArea, double k_eps) {
    const double delta = phi_pred - phi_i;
    if (fabs(delta) < DBL_EPSILON) return 1.0;
    const double h = sqrt(cellArea / M_PI);
    const double eps = k_eps * k_eps * k_eps * h * h * h;
    const double a = (delta > 0.0) ? (phi_max - phi_i) : (phi_min - phi_i);
    const double alpha = (a * a + eps + 2.0 * a * delta) / (a * a + eps + 2.0 * delta * delta);
    return sclamp(alpha, 0.0, 1.0);
}

__device__ __forceinline__ double alpha_bj(double phi_i, double phi_pred, double phi_min,
                                           double phi_max) {
    double alpha = 1.0;
    if (fabs(phi_pred - phi_i) < DBL_EPSILON) return alpha;
    if (phi_pred > phi_i)
        alpha = smin(1.0, (phi_max - phi_i) / (phi_pred - phi_i));
    else if (phi_pred < phi_i)
        alpha = smin(1.0, (phi_min - phi_i) / (phi_pred - phi_i));
    return alpha;
}

__global__ __launch_bounds__(256) void k_cell_reconstruct(Topo T, Geom G, const double *__restrict__ U,
                                                          const double *__restrict__ vU,
                                                          double *__restrict__ fL, double *__restrict__ fR,
                                                          double *__restrict__ tq, double *__restrict__ Gout,
                                                          Phys P) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= T.nc || (P.gate && *P.gate == 0.0)) return;
    const int nc = T.nc, nv = T.nv, ne = T.ne;
    // level 1: the cell's own rows
    const unsigned info = T.cinfo[c];
    const int nk = info >> 4;
    const int4 ed4 = reinterpret_cast<const int4 *>(T.cedge)[c];
    const int4 nb4 = reinterpret_cast<const int4 *>(T.cnbr)[c];
    const int4 ea4 = reinterpret_cast<const int4 *>(T.cevx)[2 * c];
    const int4 eb4 = reinterpret_cast<const int4 *>(T.cevx)[2 * c + 1];
    const int ed[4] = {ed4.x, ed4.y, ed4.z, ed4.w};
    const int nb[4] = {nb4.x, nb4.y, nb4.z, nb4.w};
    const int va[4] = {ea4.x, ea4.z, eb4.x, eb4.z}, vb[4] = {ea4.y, ea4.w, eb4.y, eb4.w};
    double u[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) u[q] = U[q * nc + c];
    const double vol = G.cvol[c], cx = G.ccx[c], cy = G.ccy[c];
    // level 2: edge end-point values, edge geometry, neighbour states, all at once
    double xa[4][4], xb[4][4];
#pragma unroll
    for (int k = 0; k < 4; ++k)
        if (k < nk) {
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                xa[k][q] = vU[q * nv + va[k]];
                xb[k][q] = vU[q * nv + vb[k]];
            }
        }
    double el[4], enx[4], eny[4], dxs[4], dys[4];
    char ebc[4];
#pragma unroll
    for (int k = 0; k < 4; ++k)
        if (k < nk) {
            const int e = ed[k];
            el[k] = G.elen[e];
            enx[k] = G.enx[e];
            eny[k] = G.eny[e];
            dxs[k] = G.emx[e] - cx;
            dys[k] = G.emy[e] - cy;
            ebc[k] = T.ebc[e];
        }
    double un[4][4];
#pragma unroll
    for (int k = 0; k < 4; ++k)
        if (nb[k] >= 0) {
#pragma unroll
            for (int q = 0; q < 4; ++q) un[k][q] = U[q * nc + nb[k]];
        }
    // edge values (+ edge BC) and the Green-Gauss sum in ascending edge id
    double gx[4] = {0.0, 0.0, 0.0, 0.0}, gy[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int k = 0; k < 4; ++k)
        if (k < nk) {
            double val[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) val[q] = 0.5 * (xa[k][q] + xb[k][q]);
            if (tagged(ebc[k])) bc_state(P, ebc[k], val[0], val[1], val[2], val[3]);
            double nx = enx[k], ny = eny[k];
            if ((info >> k) & 1u) {
                nx = -nx;
                ny = -ny;
            }
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                gx[q] += val[q] * el[k] * nx / vol;
                gy[q] += val[q] * el[k] * ny / vol;
            }
        }
    if (Gout) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            Gout[q * nc + c] = gx[q];
            Gout[(4 + q) * nc + c] = gy[q];
        }
    }
    double mn[4], mx[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) mn[q] = mx[q] = u[q];
#pragma unroll
    for (int k = 0; k < 4; ++k)
        if (nb[k] >= 0) {
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                if (un[k][q] < mn[q]) mn[q] = un[k][q];
                if (un[k][q] > mx[q]) mx[q] = un[k][q];
            }
        }
    double al[4] = {1.0, 1.0, 1.0, 1.0};
#pragma unroll
    for (int k = 0; k < 4; ++k)
        if (k < nk) {
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const double pred = u[q] + gx[q] * dxs[k] + gy[q] * dys[k];
                const double a = (P.limiter == 1) ? alpha_bj(u[q], pred, mn[q], mx[q])
                                                  : alpha_venkat(u[q], pred, mn[q], mx[q], vol, P.venkat_k);
                al[q] = smin(al[q], a);
            }
        }
#pragma unroll
    for (int k = 0; k < 4; ++k)
        if (k < nk) {
            const int e = ed[k];
            double f[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) f[q] = u[q] + al[q] * (gx[q] * dxs[k] + gy[q] * dys[k]);
            if ((info >> k) & 1u) {
#pragma unroll
                for (int q = 0; q < 4; ++q) fL[q * ne + e] = f[q];
                if (tagged(ebc[k])) {
                    double g[4];
                    bc_ghost(P, ebc[k], f, g);
#pragma unroll
                    for (int q = 0; q < 4; ++q) fR[q * ne + e] = g[q];
                }
            } else {
#pragma unroll
                for (int q = 0; q < 4; ++q) fR[q * ne + e] = f[q];
            }
        }
    // viscous stress and heat flux of the cell (src/diffusion.cpp:227-240)
    const double mu = P.mu, gamma = P.gamma, kappa = P.kappa;
    const double rho = u[0], E = u[3];
    const double uu = u[1] / rho, vvel = u[2] / rho;
    const double u_x = (gx[1] - uu * gx[0]) / rho;
    const double v_y = (gy[2] - vvel * gy[0]) / rho;
    const double u_y = (gy[1] - uu * gy[0]) / rho;
    const double v_x = (gx[2] - vvel * gx[0]) / rho;
    tq[c] = 2 * mu * u_x - 2.0 / 3.0 * mu * (u_x + v_y);
    tq[nc + c] = 2 * mu * v_y - 2.0 / 3.0 * mu * (u_x + v_y);
    tq[2 * nc + c] = mu * (u_y + v_x);
    const double T_x = (gamma - 1.0) * ((gx[3] * rho - E * gx[0]) / (rho * rho) - (uu * u_x + vvel * v_x));
    const double T_y = (gamma - 1.0) * ((gy[3] * rho - E * gy[0]) / (rho * rho) - (uu * u_y + vvel * v_y));
    tq[3 * nc + c] = -kappa * T_x;
    tq[4 * nc + c] = -kappa * T_y;
}

// Stress / heat flux from stored gradients (computeDiffusionTerm called on
// its own, src/diffusion.cpp:223-241).
__global__ __launch_bounds__(256) void k_cell_stress(Topo T, const double *__restrict__ U,
                                                     const double *__restrict__ Gs,
                                                     double *__restrict__ tq, Phys P) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= T.nc) return;
    const int nc = T.nc;
    const double mu = P.mu, gamma = P.gamma, kappa = P.kappa;
    const double rho = U[c], E = U[3 * nc + c];
    const double rx = Gs[c], rux = Gs[nc + c], rvx = Gs[2 * nc + c], Ex = Gs[3 * nc + c];
    const double ry = Gs[4 * nc + c], ruy = Gs[5 * nc + c], rvy = Gs[6 * nc + c], Ey = Gs[7 * nc + c];
    const double uu = U[nc + c] / rho, vv = U[2 * nc + c] / rho;
    const double u_x = (rux - uu * rx) / rho;
    const double v_y = (rvy - vv * ry) / rho;
    const double u_y = (ruy - uu * ry) / rho;
    const double v_x = (rvx - vv * rx) / rho;
    tq[c] = 2 * mu * u_x - 2.0 / 3.0 * mu * (u_x + v_y);
    tq[nc + c] = 2 * mu * v_y - 2.0 / 3.0 * mu * (u_x + v_y);
    tq[2 * nc + c] = mu * (u_y + v_x);
    const double T_x = (gamma - 1.0) * ((Ex * rho - E * rx) / (rho * rho) - (uu * u_x + vv * v_x));
    const double T_y = (gamma - 1.0) * ((Ey * rho - E * ry) / (rho * rho) - (uu * u_y + vv * v_y));
    tq[3 * nc + c] = -kappa * T_x;
    tq[4 * nc + c] = -kappa * T_y;
}

// ------------------------------------------------------------------------
// Flux + update: HLLC on each of the cell's edges and the convective sum
// (src/advection.cpp:433-464), the viscous divergence
// (src/diffusion.cpp:145-201), and the Shu-Osher stage update
// (src/simulation.cpp:159-167), fused.  Each interior edge's HLLC is
// evaluated by both of its cells (bit-identical), trading 2x flux flops for
// no flux array and no scatter.
// ------------------------------------------------------------------------
// |u| + c of one state (src/simulation.cpp:84-95)
__device__ __forceinline__ double wave_speed(double rho, double ru, double rv, double E, double gamma) {
    const double u = ru / rho, v = rv / rho;
    const double kinetic = 0.5 * (ru * ru + rv * rv) / rho;
    const double p = (gamma - 1.0) * (E - kinetic);
    const double cs = sqrt(gamma * p / rho);
    return hypot(u, v) + cs;
}

// Block max of m (NaN ignored like std::max) folded into the speed word.
// Positive doubles order like their bit patterns, so an integer atomic max is
// exact and independent of arrival order.
__device__ __forceinline__ void speed_fold(double m, unsigned long long *word) {
    __shared__ double sm[256];
    sm[threadIdx.x] = m;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) sm[threadIdx.x] = smax(sm[threadIdx.x], sm[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0)
        __hip_atomic_fetch_max(word, static_cast<unsigned long long>(__double_as_longlong(sm[0])), __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
}

__global__ __launch_bounds__(256) void k_cell_flux_update(Topo T, Geom G, double *__restrict__ U,
                                                          double *__restrict__ Uold,
                                                          const double *__restrict__ fL,
                                                          const double *__restrict__ fR,
                                                          const double *__restrict__ vtq,
                                                          double *__restrict__ R,
                                                          const double *__restrict__ dtp, double c0,
                                                          double c1, int flags, Phys P,
                                                          unsigned long long *__restrict__ speed) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    const bool gated = P.gate && *P.gate == 0.0;
    double m = 1e-8;
    if (c < T.nc && !gated) {
        const int nc = T.nc, nv = T.nv, ne = T.ne;
        const unsigned info = T.cinfo[c];
        const int nk = info >> 4;
        const int4 ed4 = reinterpret_cast<const int4 *>(T.cedge)[c];
        const int ed[4] = {ed4.x, ed4.y, ed4.z, ed4.w};
        const double vol = G.cvol[c];
        double u[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) u[q] = U[q * nc + c];
        // level 2: face states, edge geometry, vertex stresses, all at once
        double qL[4][4], qR[4][4], el[4], enx[4], eny[4];
        char ebc[4];
#pragma unroll
        for (int k = 0; k < 4; ++k)
            if (k < nk) {
                const int e = ed[k];
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    qL[k][q] = fL[q * ne + e];
                    qR[k][q] = fR[q * ne + e];
                }
                el[k] = G.elen[e];
                enx[k] = G.enx[e];
                eny[k] = G.eny[e];
                ebc[k] = T.ebc[e];
            }
        double ta[4][5], tb[4][5];
        if (flags & kFlagVisc) {
            const int4 ea4 = reinterpret_cast<const int4 *>(T.cevx)[2 * c];
            const int4 eb4 = reinterpret_cast<const int4 *>(T.cevx)[2 * c + 1];
            const int va[4] = {ea4.x, ea4.z, eb4.x, eb4.z}, vb[4] = {ea4.y, ea4.w, eb4.y, eb4.w};
#pragma unroll
            for (int k = 0; k < 4; ++k)
                if (k < nk) {
#pragma unroll
                    for (int q = 0; q < 5; ++q) {
                        ta[k][q] = vtq[q * nv + va[k]];
                        tb[k][q] = vtq[q * nv + vb[k]];
                    }
                }
        }
        double r[4];
        if (flags & kFlagAdv) {
            r[0] = r[1] = r[2] = r[3] = 0.0;
#pragma unroll
            for (int k = 0; k < 4; ++k)
                if (k < nk) {
                    double F[4];
                    hllc(qL[k], qR[k], enx[k], eny[k], P.gamma, F);
                    const double sign = ((info >> k) & 1u) ? +1.0 : -1.0;
#pragma unroll
                    for (int q = 0; q < 4; ++q) r[q] += sign * F[q] * el[k];
                }
            const double invV = 1.0 / vol;
#pragma unroll
            for (int q = 0; q < 4; ++q) r[q] = r[q] * -invV;
        } else {
#pragma unroll
            for (int q = 0; q < 4; ++q) r[q] = R[q * nc + c];
        }
        if (flags & kFlagVisc) {
            const double uu = u[1] / u[0];
            const double vv = u[2] / u[0];
#pragma unroll
            for (int k = 0; k < 4; ++k)
                if (k < nk) {
                    double nx = enx[k], ny = eny[k];
                    if ((info >> k) & 1u) {
                        nx = -nx;
                        ny = -ny;
                    }
                    double t[5];
#pragma unroll
                    for (int q = 0; q < 5; ++q) t[q] = 0.5 * (ta[k][q] + tb[k][q]);
                    if (ebc[k] == 'f' || ebc[k] == 'w') {
#pragma unroll
                        for (int q = 0; q < 5; ++q) t[q] = 0.0;
                    }
                    const double txx = t[0], tyy = t[1], txy = t[2], qx = t[3], qy = t[4];
                    const double L = el[k];
                    r[1] += txx * L * nx / vol + txy * L * ny / vol;
                    r[2] += txy * L * nx / vol + tyy * L * ny / vol;
                    r[3] += (uu * txx + vv * txy - qx) * L * nx / vol + (uu * txy + vv * tyy - qy) * L * ny / vol;
                }
        }
        if (flags & kFlagStoreR) {
#pragma unroll
            for (int q = 0; q < 4; ++q) R[q * nc + c] = r[q];
        }
        if (flags & kFlagRK) {
            const double dt = *dtp;
            double un[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                double old;
                if (flags & kFlagSaveOld) {
                    old = u[q];
                    Uold[q * nc + c] = u[q];
                } else {
                    old = Uold[q * nc + c];
                }
                un[q] = c0 * old + c1 * (u[q] + dt * (r[q]));
                U[q * nc + c] = un[q];
            }
            if (flags & kFlagSpeed) m = smax(m, wave_speed(un[0], un[1], un[2], un[3], P.gamma));
        }
    }
    if ((flags & kFlagSpeed) && !gated) speed_fold(m, speed);
}

// Speed word: max_c(|u|+c) of the current state (src/simulation.cpp:80-96),
// for the first step of a run or after the state was set from the host.
__global__ __launch_bounds__(256) void k_speed_reset(unsigned long long *word) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *word = static_cast<unsigned long long>(__double_as_longlong(1e-8));
}

__global__ __launch_bounds__(256) void k_speed_max(const double *__restrict__ U, int nc, Phys P,
                                                   unsigned long long *word) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    double m = 1e-8;
    if (c < nc) m = smax(m, wave_speed(U[c], U[nc + c], U[2 * nc + c], U[3 * nc + c], P.gamma));
    speed_fold(m, word);
}

// Start of a step (one thread): dt = min(CFL min_h / max_wave, 0.5 min_h^2 / mu)
// (src/simulation.cpp:108-110; min_h is static, precomputed on the host),
// the time advance of :121, the while (t < final_time) test of :75 when
// gated, and the re-arm of the speed word for the step's last stage.
__global__ __launch_bounds__(64) void k_step_begin(unsigned long long *word, Phys P, double min_h,
                                                   double *__restrict__ scal, double *__restrict__ hist,
                                                   int advance_t) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    if (P.gate && !(scal[kScalT] < P.final_t)) {
        scal[kScalActive] = 0.0;
        return;
    }
    const double max_wave = __longlong_as_double(static_cast<long long>(*word));
    const double dt_adv = P.cfl * min_h / max_wave;
    const double dt_diff = 0.5 * (min_h * min_h) / P.mu;
    const double dt = smin(dt_adv, dt_diff);
    scal[kScalDt] = dt;
    if (advance_t) {
        scal[kScalTPrev] = scal[kScalT];
        scal[kScalT] = scal[kScalT] + dt;
    }
    if (P.gate) scal[kScalSteps] = scal[kScalSteps] + 1.0;
    if (hist) *hist = dt;
    *word = static_cast<unsigned long long>(__double_as_longlong(1e-8));
}

// Sums of products sum_i a_i b_i (and optionally sum_i c_i d_i): per-block
// partials in a fixed tree, then one block combines them in block order and
// applies `op` (two launches: the kernel boundary orders the partials, no
// device-wide fence per block).
__global__ __launch_bounds__(kRedBlock) void k_dot2_partial(const double *__restrict__ a,
                                                           const double *__restrict__ b,
                                                           const double *__restrict__ c,
                                                           const double *__restrict__ d, long n,
                                                           double *__restrict__ partials) {
    __shared__ double s1[kRedBlock], s2[kRedBlock];
    double x = 0.0, y = 0.0;
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        x += a[i] * b[i];
        if (c) y += c[i] * d[i];
    }
    s1[threadIdx.x] = x;
    s2[threadIdx.x] = y;
    __syncthreads();
    for (int s = kRedBlock / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            s1[threadIdx.x] += s1[threadIdx.x + s];
            s2[threadIdx.x] += s2[threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        partials[2 * blockIdx.x] = s1[0];
        partials[2 * blockIdx.x + 1] = s2[0];
    }
}

__global__ __launch_bounds__(kRedBlock) void k_dot2_final(const double *__restrict__ partials, int nparts,
                                                         double *__restrict__ scal, int op) {
    __shared__ double s1[kRedBlock], s2[kRedBlock];
    double x = 0.0, y = 0.0;
    for (int bb = threadIdx.x; bb < nparts; bb += blockDim.x) {
        x += partials[2 * bb];
        y += partials[2 * bb + 1];
    }
    s1[threadIdx.x] = x;
    s2[threadIdx.x] = y;
    __syncthreads();
    for (int s = kRedBlock / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            s1[threadIdx.x] += s1[threadIdx.x + s];
            s2[threadIdx.x] += s2[threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x != 0) return;
    const double ab = s1[0], cd = s2[0];
    switch (op) {
        case kOpStore:  // raw sums
            scal[kScalDotA] = ab;
            scal[kScalDotB] = cd;
            break;
        case kOpRho:  // (r0.r, r.r): rho_old = rho; rho = r0.r
            scal[kScalRhoOld] = scal[kScalRho];
            scal[kScalRho] = ab;
            scal[kScalRR] = cd;
            break;
        case kOpRestart:  // (r.r): rho = r0_sqnorm = r.r
            scal[kScalRho] = ab;
            scal[kScalR0Sq] = ab;
            scal[kScalRR] = ab;
            break;
        case kOpAlpha:  // (r0.v): alpha = rho / r0.v
            scal[kScalAlpha] = scal[kScalRho] / ab;
            break;
        case kOpOmega:  // (t.t, t.s): w = t.s / t.t or 0
            scal[kScalW] = (ab > 0) ? cd / ab : 0.0;
            break;
    }
}

// BiCGSTAB vector updates (Eigen's internal::bicgstab recurrence).
__global__ __launch_bounds__(256) void k_bicg_p(long n, const double *__restrict__ r, double *__restrict__ p,
                                                const double *__restrict__ v, const double *__restrict__ scal) {
    const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double beta = (scal[kScalRho] / scal[kScalRhoOld]) * (scal[kScalAlpha] / scal[kScalW]);
    const double w = scal[kScalW];
    p[i] = r[i] + beta * (p[i] - w * v[i]);
}

__global__ __launch_bounds__(256) void k_bicg_s(long n, const double *__restrict__ r, const double *__restrict__ v,
                                                double *__restrict__ s, const double *__restrict__ scal) {
    const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i >= n) return;
    s[i] = r[i] - scal[kScalAlpha] * v[i];
}

__global__ __launch_bounds__(256) void k_bicg_xr(long n, double *__restrict__ x, double *__restrict__ r,
                                                 const double *__restrict__ p, const double *__restrict__ s,
                                                 const double *__restrict__ t, const double *__restrict__ scal) {
    const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double alpha = scal[kScalAlpha], w = scal[kScalW];
    x[i] += alpha * p[i] + w * s[i];
    r[i] = s[i] - w * t[i];
}

// out = a - b (r = b - A x); also used to copy (b == nullptr).
__global__ __launch_bounds__(256) void k_sub(long n, const double *__restrict__ a, const double *__restrict__ b,
                                             double *__restrict__ out) {
    const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i >= n) return;
    out[i] = b ? a[i] - b[i] : a[i];
}

__global__ __launch_bounds__(256) void k_fill(long n, double *__restrict__ out, double v) {
    const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i < n) out[i] = v;
}

// (R+ - R-) / (2 eps), the shape derivative of the residual
// (src/shape_optimization.cpp:1004-1006).
__global__ __launch_bounds__(256) void k_fd_diff(long n, const double *__restrict__ Rp,
                                                 const double *__restrict__ Rm, double eps,
                                                 double *__restrict__ out) {
    const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i < n) out[i] = (Rp[i] - Rm[i]) / (2.0 * eps);
}

// ------------------------------------------------------------------------
// Wall integrals.  One block: the per-edge terms are formed in parallel and
// summed by one lane in ascending edge order (the order of the reference's
// serial loops), so C_L, C_D and the adjoint RHS match it bit for bit.
// ------------------------------------------------------------------------
// computeLiftDragCoefficients (src/compute.cpp:92-132) and, with
// skip_x = 1, computeDirectObjectiveTerm (src/shape_optimization.cpp:893-939).
// The wall stresses the reference stores are identically zero
// (apply_diffusion_bc zeroes them, src/diffusion.cpp:155-162); the terms are
// kept so the sums round the same way.
__global__ __launch_bounds__(kRedBlock) void k_wall_forces(Geom G, const int *__restrict__ wall, int nw,
                                                          const double *__restrict__ fL, int ne,
                                                          double gamma, int skip_x,
                                                          double *__restrict__ out) {
    __shared__ double tl[kRedBlock], td[kRedBlock];
    __shared__ int flag[kRedBlock];
    double L = 0.0, D = 0.0;
    for (int base = 0; base < nw; base += kRedBlock) {
        const int k = base + threadIdx.x;
        double l = 0.0, d = 0.0;
        bool use = false;
        if (k < nw) {
            const int e = wall[k];
            use = !(skip_x && G.emx[e] > 1.3);
            const double nx = G.enx[e], ny = G.eny[e], A = G.elen[e];
            const double rl = fL[e], rul = fL[ne + e], rvl = fL[2 * ne + e], El = fL[3 * ne + e];
            const double p = (gamma - 1.0) * (El - 0.5 * (rul * rul + rvl * rvl) / rl);
            const double txx = 0.0, txy = 0.0, tyy = 0.0;
            l = (-p * ny + txy * nx + tyy * ny) * A;
            d = (-p * nx + txx * nx + txy * ny) * A;
        }
        tl[threadIdx.x] = use ? l : 0.0;
        td[threadIdx.x] = use ? d : 0.0;
        flag[threadIdx.x] = use;
        __syncthreads();
        if (threadIdx.x == 0) {
            const int m = min(kRedBlock, nw - base);
            for (int j = 0; j < m; ++j)
                if (flag[j]) {
                    L += tl[j];
                    D += td[j];
                }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[0] = L;
        out[1] = D;
    }
}

// assembleAdjointRHS (src/shape_optimization.cpp:491-554); b must be zero.
__global__ __launch_bounds__(64) void k_adjoint_rhs(Topo T, Geom G, const int *__restrict__ wall, int nw,
                                                   const double *__restrict__ fL, double gamma, double q_inf,
                                                   double beta, double *__restrict__ b) {
    if (threadIdx.x != 0) return;
    const int nc = T.nc, ne = T.ne;
    const double corde = 1;
    for (int k = 0; k < nw; ++k) {
        const int e = wall[k];
        const int iL = T.ecl[e];
        const double A_f = G.elen[e], nx = G.enx[e], ny = G.eny[e];
        const double rL = fL[e], ruL = fL[ne + e], rvL = fL[2 * ne + e];
        double dp[4];
        dp[0] = 0.5 * (gamma - 1.0) * ((ruL * ruL + rvL * rvL) / (rL * rL));
        dp[1] = -(gamma - 1.0) * (ruL / rL);
        dp[2] = -(gamma - 1.0) * (rvL / rL);
        dp[3] = (gamma - 1.0);
        const double coeff_CL = +ny * A_f / (q_inf * corde);
        for (int q = 0; q < 4; ++q) b[q * nc + iL] += coeff_CL * dp[q];
        const double coeff_CD = -nx * A_f / (q_inf * corde) * beta;
        for (int q = 0; q < 4; ++q) b[q * nc + iL] += coeff_CD * dp[q];
    }
}

// ------------------------------------------------------------------------
// Adjoint J^T x (src/shape_optimization.cpp:237-484).
// ------------------------------------------------------------------------
// FD flux-Jacobian blocks per edge at the frozen face states:
// blk[(r*4+c)*ne + e] = dF_r/dU_L,c, blk[(16 + r*4+c)*ne + e] = dF_r/dU_R,c.
// A boundary edge rebuilds its ghost from each perturbed left state; its
// right-state block is never read by the scatter and is stored as zero.
__global__ __launch_bounds__(256) void k_jt_blocks(Topo T, Geom G, const double *__restrict__ fL,
                                                   const double *__restrict__ fR, double eps, Phys P,
                                                   double *__restrict__ blk) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= T.ne) return;
    const int ne = T.ne;
    const char bc = T.ebc[e];
    const double enx = G.enx[e], eny = G.eny[e];
    double bL[4], bR[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        bL[q] = fL[q * ne + e];
        bR[q] = fR[q * ne + e];
    }
    for (int c = 0; c < 4; ++c) {
        double Lp[4], Lm[4], Rp[4], Rm[4], Fp[4], Fm[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            Lp[q] = Lm[q] = bL[q];
            Rp[q] = Rm[q] = bR[q];
        }
        Lp[c] = bL[c] + eps;
        Lm[c] = bL[c] - eps;
        if (bc != '\0') {
            bc_ghost(P, bc, Lp, Rp);
            bc_ghost(P, bc, Lm, Rm);
        }
        hllc(Lp, Rp, enx, eny, P.gamma, Fp);
        hllc(Lm, Rm, enx, eny, P.gamma, Fm);
#pragma unroll
        for (int r = 0; r < 4; ++r) blk[(r * 4 + c) * ne + e] = (Fp[r] - Fm[r]) / (2 * eps);
    }
    for (int c = 0; c < 4; ++c) {
        if (bc != '\0') {
#pragma unroll
            for (int r = 0; r < 4; ++r) blk[(16 + r * 4 + c) * ne + e] = 0.0;
            continue;
        }
        double Rp[4], Rm[4], Fp[4], Fm[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) Rp[q] = Rm[q] = bR[q];
        Rp[c] = bR[c] + eps;
        Rm[c] = bR[c] - eps;
        hllc(bL, Rp, enx, eny, P.gamma, Fp);
        hllc(bL, Rm, enx, eny, P.gamma, Fm);
#pragma unroll
        for (int r = 0; r < 4; ++r) blk[(16 + r * 4 + c) * ne + e] = (Fp[r] - Fm[r]) / (2 * eps);
    }
}

// y = J^T x as a per-cell gather in ascending edge id: the serialisation of
// the reference's atomic scatter in edge order, without atomics.
__global__ __launch_bounds__(256) void k_jt_apply(Topo T, Geom G, const double *__restrict__ blk,
                                                  const double *__restrict__ x, double *__restrict__ y) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= T.nc) return;
    const int nc = T.nc, ne = T.ne;
    const int nk = T.cinfo[c] >> 4;
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    for (int k = 0; k < nk; ++k) {
        const int e = T.cedge[4 * c + k];
        const double L = G.elen[e];
        const int iL = T.ecl[e];
        const double invV_L = 1.0 / G.cvol[iL];
        double D[4];
        int off;
        if (T.ebc[e] != '\0') {
            if (c != iL) continue;
#pragma unroll
            for (int q = 0; q < 4; ++q) D[q] = (-x[q * nc + iL] * invV_L) * L;
            off = 0;
        } else {
            const int iR = T.ecr[e];
            const double invV_R = 1.0 / G.cvol[iR];
#pragma unroll
            for (int q = 0; q < 4; ++q) D[q] = (-x[q * nc + iL] * invV_L + x[q * nc + iR] * invV_R) * L;
            off = (c == iL) ? 0 : 16;
        }
#pragma unroll
        for (int kk = 0; kk < 4; ++kk) {
            const double *F = blk + (size_t)(off + kk) * ne + e;  // column kk: rows r at stride 4*ne
            acc[kk] += (F[0] * D[0] + F[(size_t)4 * ne] * D[1] + F[(size_t)8 * ne] * D[2] +
                        F[(size_t)12 * ne] * D[3]);
        }
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) y[q * nc + c] = acc[q];
}

// Standalone HLLC on a batch of faces (known-answer surface).
__global__ __launch_bounds__(256) void k_hllc_batch(const double *__restrict__ qL, const double *__restrict__ qR,
                                                    const double *__restrict__ nrm, double gamma,
                                                    double *__restrict__ F, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double a[4], b[4], f[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        a[q] = qL[q * n + i];
        b[q] = qR[q * n + i];
    }
    hllc(a, b, nrm[i], nrm[n + i], gamma, f);
#pragma unroll
    for (int q = 0; q < 4; ++q) F[q * n + i] = f[q];
}

// ------------------------------------------------------------------------
// Launch helpers
// ------------------------------------------------------------------------
static inline unsigned nblk(long n, int b = 256) { return (unsigned)((n + b - 1) / b); }

hipError_t launch_vertex_state(const Topo &T, const Geom &G, const double *U, double *vU, const Phys &P,
                               hipStream_t s) {
    hipLaunchKernelGGL(k_vertex_gather<4>, dim3(nblk(T.nv)), dim3(256), 0, s, T, G, U, vU, P);
    return hipGetLastError();
}
hipError_t launch_vertex_visc(const Topo &T, const Geom &G, const double *tq, double *vtq, const Phys &P,
                              hipStream_t s) {
    hipLaunchKernelGGL(k_vertex_gather<5>, dim3(nblk(T.nv)), dim3(256), 0, s, T, G, tq, vtq, P);
    return hipGetLastError();
}
hipError_t launch_reconstruct(const Topo &T, const Geom &G, const double *U, const double *vU, double *fL,
                              double *fR, double *tq, double *Gout, const Phys &P, hipStream_t s) {
    hipLaunchKernelGGL(k_cell_reconstruct, dim3(nblk(T.nc)), dim3(256), 0, s, T, G, U, vU, fL, fR, tq, Gout, P);
    return hipGetLastError();
}
hipError_t launch_stress(const Topo &T, const double *U, const double *Gs, double *tq, const Phys &P,
                         hipStream_t s) {
    hipLaunchKernelGGL(k_cell_stress, dim3(nblk(T.nc)), dim3(256), 0, s, T, U, Gs, tq, P);
    return hipGetLastError();
}
hipError_t launch_flux_update(const Topo &T, const Geom &G, double *U, double *Uold, const double *fL,
                              const double *fR, const double *vtq, double *R, const double *dtp, double c0,
                              double c1, int flags, const Phys &P, hipStream_t s, unsigned long long *speed) {
    hipLaunchKernelGGL(k_cell_flux_update, dim3(nblk(T.nc)), dim3(256), 0, s, T, G, U, Uold, fL, fR, vtq, R,
                       dtp, c0, c1, flags, P, speed);
    return hipGetLastError();
}
hipError_t launch_speed_reset(unsigned long long *word, hipStream_t s) {
    hipLaunchKernelGGL(k_speed_reset, dim3(1), dim3(64), 0, s, word);
    return hipGetLastError();
}
hipError_t launch_speed_max(const double *U, int nc, const Phys &P, unsigned long long *word, hipStream_t s) {
    hipLaunchKernelGGL(k_speed_max, dim3(nblk(nc)), dim3(256), 0, s, U, nc, P, word);
    return hipGetLastError();
}
hipError_t launch_step_begin(unsigned long long *word, const Phys &P, double min_h, double *scal, double *hist,
                             int advance_t, hipStream_t s) {
    hipLaunchKernelGGL(k_step_begin, dim3(1), dim3(64), 0, s, word, P, min_h, scal, hist, advance_t);
    return hipGetLastError();
}
hipError_t launch_dot2(const double *a, const double *b, const double *c, const double *d, long n,
                       double *partials, double *scal, int op, hipStream_t s) {
    const unsigned g = (unsigned)std::min<long>(kRedGrid, std::max<long>(1, (long)nblk(n, kRedBlock)));
    hipLaunchKernelGGL(k_dot2_partial, dim3(g), dim3(kRedBlock), 0, s, a, b, c, d, n, partials);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_dot2_final, dim3(1), dim3(kRedBlock), 0, s, partials, (int)g, scal, op);
    return hipGetLastError();
}
hipError_t launch_bicg_p(long n, const double *r, double *p, const double *v, const double *scal, hipStream_t s) {
    hipLaunchKernelGGL(k_bicg_p, dim3(nblk(n)), dim3(256), 0, s, n, r, p, v, scal);
    return hipGetLastError();
}
hipError_t launch_bicg_s(long n, const double *r, const double *v, double *sv, const double *scal, hipStream_t s) {
    hipLaunchKernelGGL(k_bicg_s, dim3(nblk(n)), dim3(256), 0, s, n, r, v, sv, scal);
    return hipGetLastError();
}
hipError_t launch_bicg_xr(long n, double *x, double *r, const double *p, const double *sv, const double *t,
                          const double *scal, hipStream_t s) {
    hipLaunchKernelGGL(k_bicg_xr, dim3(nblk(n)), dim3(256), 0, s, n, x, r, p, sv, t, scal);
    return hipGetLastError();
}
hipError_t launch_sub(long n, const double *a, const double *b, double *out, hipStream_t s) {
    hipLaunchKernelGGL(k_sub, dim3(nblk(n)), dim3(256), 0, s, n, a, b, out);
    return hipGetLastError();
}
hipError_t launch_fill(long n, double *out, double v, hipStream_t s) {
    hipLaunchKernelGGL(k_fill, dim3(nblk(n)), dim3(256), 0, s, n, out, v);
    return hipGetLastError();
}
hipError_t launch_fd_diff(long n, const double *Rp, const double *Rm, double eps, double *out, hipStream_t s) {
    hipLaunchKernelGGL(k_fd_diff, dim3(nblk(n)), dim3(256), 0, s, n, Rp, Rm, eps, out);
    return hipGetLastError();
}
hipError_t launch_wall_forces(const Geom &G, const int *wall, int nw, const double *fL, int ne, double gamma,
                              int skip_x, double *out, hipStream_t s) {
    hipLaunchKernelGGL(k_wall_forces, dim3(1), dim3(kRedBlock), 0, s, G, wall, nw, fL, ne, gamma, skip_x, out);
    return hipGetLastError();
}
hipError_t launch_adjoint_rhs(const Topo &T, const Geom &G, const int *wall, int nw, const double *fL,
                              double gamma, double q_inf, double beta, double *b, hipStream_t s) {
    hipLaunchKernelGGL(k_adjoint_rhs, dim3(1), dim3(64), 0, s, T, G, wall, nw, fL, gamma, q_inf, beta, b);
    return hipGetLastError();
}
hipError_t launch_jt_blocks(const Topo &T, const Geom &G, const double *fL, const double *fR, double eps,
                            const Phys &P, double *blk, hipStream_t s) {
    hipLaunchKernelGGL(k_jt_blocks, dim3(nblk(T.ne)), dim3(256), 0, s, T, G, fL, fR, eps, P, blk);
    return hipGetLastError();
}
hipError_t launch_jt_apply(const Topo &T, const Geom &G, const double *blk, const double *x, double *y,
                           hipStream_t s) {
    hipLaunchKernelGGL(k_jt_apply, dim3(nblk(T.nc)), dim3(256), 0, s, T, G, blk, x, y);
    return hipGetLastError();
}
hipError_t launch_hllc_batch(const double *qL, const double *qR, const double *nrm, double gamma, double *F,
                             int n, hipStream_t s) {
    hipLaunchKernelGGL(k_hllc_batch, dim3(nblk(n)), dim3(256), 0, s, qL, qR, nrm, gamma, F, n);
    return hipGetLastError();
}

}  // namespace o2f
