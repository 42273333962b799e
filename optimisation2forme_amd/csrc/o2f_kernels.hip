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

#include "o2f_device.hpp"

namespace o2f {

// ------------------------------------------------------------------------
// Cell -> vertex gather with precomputed stencils
// (src/advection.cpp:26-131 for the state, src/diffusion.cpp:28-140 for the
// viscous quantities).  NV = 4 state variables with the state BC, or NV = 5
// viscous quantities with apply_diffusion_bc (src/boundary.cpp:101-127).
// ------------------------------------------------------------------------
// The destination is always AoS (4 + trailing block).
template <int NV, bool SRC_AOS>
__device__ __forceinline__ void gather_vertex(const Topo &T, const Geom &G, const double *__restrict__ src,
                                              double *__restrict__ dst, const Phys &P, int v) {
    const uint8_t kind = T.skind[v];
    const int b = T.sptr[v], n = T.sptr[v + 1] - b;
    vertex_gather_vals<NV, SRC_AOS>(T, G, src, P, kind, b, n, [&] { return T.vbc[v]; }, [&](const double(&out)[NV]) {
        double o4[4] = {out[0], out[1], out[2], out[3]};
        st4(dst, v, o4);
        if (NV == 5) dst[4 * (long)T.nv + v] = out[NV - 1];
    });
}

template <int NV, bool SRC_AOS>
__global__ __launch_bounds__(256) void k_vertex_gather(Topo T, Geom G, const double *__restrict__ src,
                                                       double *__restrict__ dst, Phys P) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= T.nv || (P.gate && *P.gate == 0.0)) return;
    gather_vertex<NV, SRC_AOS>(T, G, src, dst, P, v);
}

// HLLC flux of one edge from its face states (src/advection.cpp:433-437).
__device__ __forceinline__ void edge_flux(const Topo &T, const Geom &G, const double *__restrict__ fL,
                                          const double *__restrict__ fR, double *__restrict__ F, const Phys &P,
                                          int e) {
    double qL[4], qR[4], f[4];
    ld4(fL, e, qL);
    ld4(fR, e, qR);
    hllc(qL, qR, G.enx[e], G.eny[e], P.gamma, f);
    st4(F, e, f);
}

// One launch, two independent jobs that both follow the reconstruction:
// blocks [0, nbv) gather the viscous quantities at the vertices
// (src/diffusion.cpp:28-140), the other blocks evaluate HLLC once per edge.
// `b` is the block's index in the mesh's own grid (blockIdx.x for a
// single-mesh launch, the tile's block for a batched one).
__device__ __forceinline__ void visc_edge_body(int b, const Topo &T, const Geom &G, const double *__restrict__ tq,
                                               double *__restrict__ vtq, const double *__restrict__ fL,
                                               const double *__restrict__ fR, double *__restrict__ F,
                                               const Phys &P, int nbv, int do_visc) {
    if (P.gate && *P.gate == 0.0) return;
    if (b < nbv) {
        const int v = b * blockDim.x + threadIdx.x;
        if (do_visc && v < T.nv) gather_vertex<5, true>(T, G, tq, vtq, P, v);
    } else {
        const int e = (b - nbv) * blockDim.x + threadIdx.x;
        if (e < T.ne) edge_flux(T, G, fL, fR, F, P, e);
    }
}

__global__ __launch_bounds__(256) void k_viscous_gather_edge_flux(Topo T, Geom G, const double *__restrict__ tq,
                                                                  double *__restrict__ vtq,
                                                                  const double *__restrict__ fL,
                                                                  const double *__restrict__ fR,
                                                                  double *__restrict__ F, Phys P, int nbv,
                                                                  int do_visc) {
    visc_edge_body(blockIdx.x, T, G, tq, vtq, fL, fR, F, P, nbv, do_visc);
}

// ------------------------------------------------------------------------
// Cell reconstruction: edge values + Green-Gauss gradient
// (src/advection.cpp:136-188), neighbour min/max, limiter and face states
// with boundary ghosts (src/advection.cpp:232-321), and the cell stress /
// heat flux of computeDiffusionTerm (src/diffusion.cpp:223-241), fused.
// ------------------------------------------------------------------------
__device__ __forceinline__ void reconstruct_body(int blk, const Topo &T, const Geom &G,
                                                 const double *__restrict__ U, const double *__restrict__ vU,
                                                 double *__restrict__ fL, double *__restrict__ fR,
                                                 double *__restrict__ tq, double *__restrict__ Gout,
                                                 const Phys &P) {
    const int gid = blk * blockDim.x + threadIdx.x;
    const int c = gid / kLanesPerCell, k = gid % kLanesPerCell;
    if (c >= T.nc || (P.gate && *P.gate == 0.0)) return;  // whole lane groups only
    const int nc = T.nc, nv = T.nv, ne = T.ne;
    const unsigned info = T.cinfo[c];
    const int nk = info >> 4;
    const bool has_edge = k < nk;
    const bool is_left = (info >> k) & 1u;
    double u[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) u[q] = U[q * nc + c];
    const double vol = G.cvol[c], cx = G.ccx[c], cy = G.ccy[c];
    // this lane's edge and face neighbour
    const int e = has_edge ? T.cedge[4 * c + k] : 0;
    const int nb = T.cnbr[4 * c + k];
    double el = 0.0, enx = 0.0, eny = 0.0, dx = 0.0, dy = 0.0, xa[4], xb[4];
    char ebc = 0;
    if (has_edge) {
        const int va = T.cevx[8 * c + 2 * k], vb = T.cevx[8 * c + 2 * k + 1];
        el = G.elen[e];
        enx = G.enx[e];
        eny = G.eny[e];
        dx = G.emx[e] - cx;
        dy = G.emy[e] - cy;
        ebc = T.ebc[e];
        ld4(vU, va, xa);
        ld4(vU, vb, xb);
    }
    // edge value (+ edge BC) and this edge's Green-Gauss terms
    double tx[4] = {0.0, 0.0, 0.0, 0.0}, ty[4] = {0.0, 0.0, 0.0, 0.0};
    if (has_edge) {
        double val[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) val[q] = 0.5 * (xa[q] + xb[q]);
        if (tagged(ebc)) bc_state(P, ebc, val[0], val[1], val[2], val[3]);
        double nx = enx, ny = eny;
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
    // From here lane k owns conservative variable k.  The 4 x 4 in-quad
    // transpose gives lane k the numerators of variable k's terms of the
    // edges in slot order; it divides them by the volume (the same
    // val * el * n / vol), so a triangle costs 3 division pairs per lane
    // rather than 4.  The sums start at +0.0 and take the nk terms in slot
    // order.
    quad_transpose(tx, k);
    quad_transpose(ty, k);
    double gxk = 0.0, gyk = 0.0;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        if (j >= nk) break;
        gxk += tx[j] / vol;
        gyk += ty[j] / vol;
    }
    const double uk = (k == 0) ? u[0] : (k == 1) ? u[1] : (k == 2) ? u[2] : u[3];
    // neighbour min / max of variable k in neighbour-list order (no
    // neighbour: the cell's own value, which changes nothing)
    const int nbs[4] = {from_lane_i<0>(nb), from_lane_i<1>(nb), from_lane_i<2>(nb), from_lane_i<3>(nb)};
    double mn = uk, mx = uk;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const double v = (nbs[j] >= 0) ? U[(long)k * nc + nbs[j]] : uk;
        if (v < mn) mn = v;
        if (v > mx) mx = v;
    }
    // limiter of variable k: the min over the edges in slot order (edges
    // beyond nk would offer 1.0, which changes nothing)
    const double dxs[4] = {from_lane<0>(dx), from_lane<1>(dx), from_lane<2>(dx), from_lane<3>(dx)};
    const double dys[4] = {from_lane<0>(dy), from_lane<1>(dy), from_lane<2>(dy), from_lane<3>(dy)};
    const double veps = (P.limiter == 1) ? 0.0 : venkat_eps(vol, P.venkat_k);
    double al = 1.0;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        if (j >= nk) break;
        const double pred = uk + gxk * dxs[j] + gyk * dys[j];
        const double a = (P.limiter == 1) ? alpha_bj(uk, pred, mn, mx) : alpha_venkat_eps(uk, pred, mn, mx, veps);
        al = smin(al, a);
    }
    // face states of variable k on every edge (+ boundary ghost,
    // src/boundary.cpp:48-80)
    const int es[4] = {from_lane_i<0>(e), from_lane_i<1>(e), from_lane_i<2>(e), from_lane_i<3>(e)};
    const int bcs[4] = {from_lane_i<0>(ebc), from_lane_i<1>(ebc), from_lane_i<2>(ebc), from_lane_i<3>(ebc)};
    const double fsk = (k == 0) ? P.fs_rho : (k == 1) ? P.fs_ru : (k == 2) ? P.fs_rv : P.fs_E;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        if (j >= nk) break;
        const double f = uk + al * (gxk * dxs[j] + gyk * dys[j]);
        const long o = 4 * (long)es[j] + k;
        if ((info >> j) & 1u) {
            fL[o] = f;
            const char bcj = (char)bcs[j];
            if (tagged(bcj)) fR[o] = (bcj == 'i' || bcj == 'f') ? fsk : (bcj == 'w' && (k == 1 || k == 2)) ? -f : f;
        } else {
            fR[o] = f;
        }
    }
    if (Gout) {
        Gout[(long)k * nc + c] = gxk;
        Gout[(long)(4 + k) * nc + c] = gyk;
    }
    // viscous stress and heat flux of the cell (src/diffusion.cpp:227-240),
    // the divisions spread over the four lanes: the velocities and the
    // heat-flux quotients share one division, then lane k forms one velocity
    // derivative (u_x, v_y, u_y, v_x) and lanes 0 / 1 one temperature
    // derivative (T_x, T_y); every value is the same expression as in one
    // lane, so the results are unchanged
    // Each lane reads only the gradient components it needs: the density
    // gradient (lane 0), the energy gradient (lane 3) and, through one
    // quad_perm [1, 2, 1, 2], (d rho u / dx, d rho v / dx) or
    // (d rho u / dy, d rho v / dy).
    const double gx0 = from_lane<0>(gxk), gy0 = from_lane<0>(gyk);
    const double gx3 = from_lane<3>(gxk), gy3 = from_lane<3>(gyk);
    const double px = quad_perm<0x99>(gxk), py = quad_perm<0x99>(gyk);
    const double mu = P.mu, gamma = P.gamma, kappa = P.kappa;
    const double rho = u[0], E = u[3];
    // one division for four quotients: lanes 0 / 1 the velocities u, v,
    // lanes 2 / 3 the heat-flux terms (E_x rho - E rho_x) / rho^2 in x, y
    const double qn = (k == 0)   ? u[1]
                      : (k == 1) ? u[2]
                      : (k == 2) ? (gx3 * rho - E * gx0)
                                 : (gy3 * rho - E * gy0);
    const double qv = qn / ((k < 2) ? rho : rho * rho);
    const double uu = from_lane<0>(qv), vvel = from_lane<1>(qv), ex = from_lane<2>(qv), ey = from_lane<3>(qv);
    // lane 0: u_x, 1: v_y, 2: u_y, 3: v_x
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
    st4(tq, c, t4);
    tq[4 * (long)nc + c] = -kappa * T_y;
}

__global__ __launch_bounds__(256) void k_cell_reconstruct(Topo T, Geom G, const double *__restrict__ U,
                                                          const double *__restrict__ vU,
                                                          double *__restrict__ fL, double *__restrict__ fR,
                                                          double *__restrict__ tq, double *__restrict__ Gout,
                                                          Phys P) {
    reconstruct_body(blockIdx.x, T, G, U, vU, fL, fR, tq, Gout, P);
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
    const double T_x = (gamma - 1.0) * ((Ex * rho - E * rx) / (rho * rho) - (uu * u_x + vv * v_x));
    const double T_y = (gamma - 1.0) * ((Ey * rho - E * ry) / (rho * rho) - (uu * u_y + vv * v_y));
    const double t4[4] = {2 * mu * u_x - 2.0 / 3.0 * mu * (u_x + v_y), 2 * mu * v_y - 2.0 / 3.0 * mu * (u_x + v_y),
                          mu * (u_y + v_x), -kappa * T_x};
    st4(tq, c, t4);
    tq[4 * (long)nc + c] = -kappa * T_y;
}

// ------------------------------------------------------------------------
// Flux + update: HLLC on each of the cell's edges and the convective sum
// (src/advection.cpp:433-464), the viscous divergence
// (src/diffusion.cpp:145-201), and the Shu-Osher stage update
// (src/simulation.cpp:159-167), fused.  The edge fluxes come from the
// per-edge HLLC pass; each cell gathers its edges' fluxes (no scatter, no
// atomics) and sums them in ascending edge order.
// ------------------------------------------------------------------------
// Block max of m (NaN ignored like std::max) stored as this block's partial;
// the step-begin block takes the max of the partials (exact, order-free).
__device__ __forceinline__ void speed_fold(double m, double *parts, int blk) {
    __shared__ double sm[256];
    sm[threadIdx.x] = m;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) sm[threadIdx.x] = smax(sm[threadIdx.x], sm[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) parts[blk] = sm[0];
    __syncthreads();  // sm is reused when one workgroup runs several tiles
}

__device__ __forceinline__ void flux_update_body(int blk, const Topo &T, const Geom &G, double *__restrict__ U,
                                                 double *__restrict__ Uold, const double *__restrict__ Fe,
                                                 const double *__restrict__ vtq, double *__restrict__ R,
                                                 const double *__restrict__ dtp, double c0, double c1, int flags,
                                                 const Phys &P, double *__restrict__ speed) {
    const int gid = blk * blockDim.x + threadIdx.x;
    const int c = gid / kLanesPerCell, k = gid % kLanesPerCell;
    const bool gated = P.gate && *P.gate == 0.0;
    double m = 1e-8;
    if (c < T.nc && !gated) {
        const int nc = T.nc, nv = T.nv, ne = T.ne;
        const unsigned info = T.cinfo[c];
        const int nk = info >> 4;
        const bool has_edge = k < nk;
        const double vol = G.cvol[c];
        double u[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) u[q] = U[q * nc + c];
        // this lane's edge: flux, geometry, end-point stresses
        double Fk[4] = {0.0, 0.0, 0.0, 0.0}, el = 0.0, enx = 0.0, eny = 0.0, ta[5], tb[5];
        char ebc = 0;
        if (has_edge) {
            const int e = T.cedge[4 * c + k];
            if (flags & kFlagAdv) ld4(Fe, e, Fk);
            el = G.elen[e];
            enx = G.enx[e];
            eny = G.eny[e];
            ebc = T.ebc[e];
            if (flags & kFlagVisc) {
                const int va = T.cevx[8 * c + 2 * k], vb = T.cevx[8 * c + 2 * k + 1];
                double a4[4], b4[4];
                ld4(vtq, va, a4);
                ld4(vtq, vb, b4);
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    ta[q] = a4[q];
                    tb[q] = b4[q];
                }
                ta[4] = vtq[4 * (long)nv + va];
                tb[4] = vtq[4 * (long)nv + vb];
            }
        }
        const bool is_left = (info >> k) & 1u;
        // one division for three quotients: lanes 0 / 1 / 2 form u, v, 1 / V
        const double qv = ((k == 0) ? u[1] : (k == 1) ? u[2] : 1.0) / ((k < 2) ? u[0] : vol);
        const double uu = from_lane<0>(qv), vv = from_lane<1>(qv), invV = from_lane<2>(qv);
        // convective term of this edge, src/advection.cpp:453-456
        double adv[4] = {0.0, 0.0, 0.0, 0.0};
        if (has_edge) {
            const double sign = is_left ? +1.0 : -1.0;
#pragma unroll
            for (int q = 0; q < 4; ++q) adv[q] = sign * Fk[q] * el;
        }
        // viscous terms of this edge, src/diffusion.cpp:183-200
        double w1 = 0.0, w2 = 0.0, w3 = 0.0;
        if ((flags & kFlagVisc) && has_edge) {
            double nx = enx, ny = eny;
            if (is_left) {
                nx = -nx;
                ny = -ny;
            }
            double t[5];
#pragma unroll
            for (int q = 0; q < 5; ++q) t[q] = 0.5 * (ta[q] + tb[q]);
            if (ebc == 'f' || ebc == 'w') {
#pragma unroll
                for (int q = 0; q < 5; ++q) t[q] = 0.0;
            }
            const double txx = t[0], tyy = t[1], txy = t[2], qx = t[3], qy = t[4];
            w1 = txx * el * nx / vol + txy * el * ny / vol;
            w2 = txy * el * nx / vol + tyy * el * ny / vol;
            w3 = (uu * txx + vv * txy - qx) * el * nx / vol + (uu * txy + vv * tyy - qy) * el * ny / vol;
        }
        // From here lane k owns conservative variable k.  The in-quad
        // transposes give lane k variable k's terms of the edges in slot
        // order; it sums them, updates and stores variable k.
        double rk;
        if (flags & kFlagAdv) {
            // all four slots in order: +0.0 from lanes without an edge (the
            // sum starts at +0.0, so it is never -0.0 and +0.0 changes nothing)
            quad_transpose(adv, k);
            rk = 0.0;
#pragma unroll
            for (int j = 0; j < 4; ++j) rk += adv[j];
            rk = rk * -invV;
        } else {
            rk = R[k * nc + c];
        }
        if (flags & kFlagVisc) {
            // per edge, in slot order: R_rho_u, R_rho_v, R_E each += that
            // edge's term (slots j < nk only: rk may be -0.0 here)
            double wv[4] = {0.0, w1, w2, w3};
            quad_transpose(wv, k);
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const double t = rk + wv[j];
                rk = (k > 0 && j < nk) ? t : rk;
            }
        }
        if (flags & kFlagStoreR) R[k * nc + c] = rk;
        if (flags & kFlagRK) {
            const double dt = *dtp;
            const double uk = (k == 0) ? u[0] : (k == 1) ? u[1] : (k == 2) ? u[2] : u[3];
            double old;
            if (flags & kFlagSaveOld) {
                old = uk;
                Uold[k * nc + c] = uk;
            } else {
                old = Uold[k * nc + c];
            }
            const double unew = c0 * old + c1 * (uk + dt * rk);
            U[k * nc + c] = unew;
            if (flags & kFlagSpeed)
                m = smax(m, wave_speed(from_lane<0>(unew), from_lane<1>(unew), from_lane<2>(unew),
                                       from_lane<3>(unew), P.gamma));
        }
    }
    if ((flags & kFlagSpeed) && !gated) speed_fold(m, speed, blk);
}

__global__ __launch_bounds__(256) void k_cell_flux_update(Topo T, Geom G, double *__restrict__ U,
                                                          double *__restrict__ Uold,
                                                          const double *__restrict__ Fe,
                                                          const double *__restrict__ vtq,
                                                          double *__restrict__ R,
                                                          const double *__restrict__ dtp, double c0,
                                                          double c1, int flags, Phys P,
                                                          double *__restrict__ speed) {
    flux_update_body(blockIdx.x, T, G, U, Uold, Fe, vtq, R, dtp, c0, c1, flags, P, speed);
}

// Speed partials: per-block max_c(|u|+c) of the current state
// (src/simulation.cpp:80-96), for the first step of a run or after the state
// was set from the host (later steps get them from the last stage).
__device__ __forceinline__ void speed_max_body(int blk, const double *__restrict__ U, int nc, const Phys &P,
                                               double *__restrict__ parts) {
    // same grid as the 4-lane cell kernels, so both leave the same partials
    const int gid = blk * blockDim.x + threadIdx.x;
    const int c = gid / kLanesPerCell;
    double m = 1e-8;
    if (c < nc && gid % kLanesPerCell == 0)
        m = smax(m, wave_speed(U[c], U[nc + c], U[2 * nc + c], U[3 * nc + c], P.gamma));
    speed_fold(m, parts, blk);
}

__global__ __launch_bounds__(256) void k_speed_max(const double *__restrict__ U, int nc, Phys P,
                                                   double *__restrict__ parts) {
    speed_max_body(blockIdx.x, U, nc, P, parts);
}

// Start of a step (one block): max_wave = max of the speed partials, then
// dt = min(CFL min_h / max_wave, 0.5 min_h^2 / mu) (src/simulation.cpp:108-110;
// min_h is static, precomputed on the host), the time advance of :121 and,
// when gated, the while (t < final_time) test of :75.
__device__ __forceinline__ void step_begin_body(const double *__restrict__ parts, int nparts, const Phys &P,
                                                double min_h, double *__restrict__ scal,
                                                double *__restrict__ hist, int advance_t) {
    __shared__ double sm[256];
    if (P.gate && *P.gate == 0.0) return;  // uniform: read before any thread writes it
    // The max of the partials is exact and order-free (m starts at 1e-8, so
    // a NaN partial is skipped wherever it comes): sixteen loads per round,
    // indices past the end clamped to the last partial.
    double m = 1e-8;
    constexpr int kRound = 16;
    for (int b0 = 0; b0 < nparts; b0 += kRound * (int)blockDim.x) {
        double v[kRound];
#pragma unroll
        for (int u = 0; u < kRound; ++u) v[u] = parts[min(b0 + u * (int)blockDim.x + (int)threadIdx.x, nparts - 1)];
#pragma unroll
        for (int u = 0; u < kRound; ++u) m = smax(m, v[u]);
    }
    sm[threadIdx.x] = m;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) sm[threadIdx.x] = smax(sm[threadIdx.x], sm[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x != 0) return;
    if (P.gate && (!(scal[kScalT] < P.final_t) || scal[kScalConv] != 0.0)) {
        scal[kScalActive] = 0.0;
        return;
    }
    const double max_wave = sm[0];
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
}

__global__ __launch_bounds__(256) void k_step_begin(const double *__restrict__ parts, int nparts, Phys P,
                                                    double min_h, double *__restrict__ scal,
                                                    double *__restrict__ hist, int advance_t) {
    step_begin_body(parts, nparts, P, min_h, scal, hist, advance_t);
}

// Start of a step fused with the first stage's vertex gather: workgroup 0
// runs the step-begin reduction (dt is first read by the stage's update
// kernel, two launches later), concurrently with the gathers of the others
// (dispatched first, it does not extend the launch's tail).  A gather block
// that still sees the gate open while workgroup 0 closes it only writes
// vertex values nothing reads once the gate is shut.
__global__ __launch_bounds__(256) void k_step_gather(Topo T, Geom G, const double *__restrict__ U,
                                                     double *__restrict__ vU, Phys P,
                                                     const double *__restrict__ parts, int nparts, double min_h,
                                                     double *__restrict__ scal, double *__restrict__ hist,
                                                     int advance_t) {
    if (blockIdx.x == 0) {
        step_begin_body(parts, nparts, P, min_h, scal, hist, advance_t);
        return;
    }
    const int v = (blockIdx.x - 1) * blockDim.x + threadIdx.x;
    if (v >= T.nv || (P.gate && *P.gate == 0.0)) return;
    gather_vertex<4, false>(T, G, U, vU, P, v);
}

// The block's two partial sums of a reduction (each thread's running sums x,
// y, then a fixed tree) into partials[2 blockIdx.x + 0/1].  Every reduction
// of the adjoint solve uses this one tree, so fused kernels reproduce the
// plain reduction's partials bit for bit.
__device__ __forceinline__ void block_partials2(double x, double y, double *__restrict__ partials) {
    __shared__ double s1[kRedBlock], s2[kRedBlock];
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

// The two totals of nparts block partials, in block order then a fixed tree;
// every thread of the block gets them.  Any block computes the same bits,
// so a consumer kernel may redo this in each of its blocks instead of
// reading a result a separate one-block launch stored.
__device__ __forceinline__ void final_sums2(const double *__restrict__ partials, int nparts, double &ab,
                                            double &cd) {
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
    ab = s1[0];
    cd = s2[0];
}

// Sums of products sum_i a_i b_i (and optionally sum_i c_i d_i): per-block
// partials in a fixed tree, then a later launch combines them in block order
// (k_dot2_final, or every block of the kernel that uses the result): the
// kernel boundary orders the partials, no device-wide fence per block.
__global__ __launch_bounds__(kRedBlock) void k_dot2_partial(const double *__restrict__ a,
                                                           const double *__restrict__ b,
                                                           const double *__restrict__ c,
                                                           const double *__restrict__ d, long n,
                                                           double *__restrict__ partials,
                                                           const double *__restrict__ gate) {
    if (gate && *gate != 1.0) return;  // a stopped BiCGSTAB loop
    double x = 0.0, y = 0.0;
    // unrolled (and the optional second product tested once, outside the
    // loop): the loads of four elements are in flight together, the sums
    // still add them one after another in element order
    const long i0 = blockIdx.x * (long)blockDim.x + threadIdx.x, stride = (long)gridDim.x * blockDim.x;
    if (c) {
#pragma unroll 4
        for (long i = i0; i < n; i += stride) {
            x += a[i] * b[i];
            y += c[i] * d[i];
        }
    } else {
#pragma unroll 4
        for (long i = i0; i < n; i += stride) x += a[i] * b[i];
    }
    block_partials2(x, y, partials);
}

// The last update of a BiCGSTAB iteration fused with the reductions on
// either side of it.  Each block first combines the (t.t, t.s) partials
// (wparts, nparts of them) in k_dot2_final's order into w = t.s / t.t, or 0
// when t.t is not positive (block 0 also stores it); then x += alpha px + w sx;
// r = s - w t; and partials of (r0.r, r.r), with the elements of
// k_dot2_partial's grid-stride loop in its order (the same grid), so the
// partials are those of k_dot2_partial(r0, r, r, r).  wparts and partials
// are distinct: other blocks still read wparts while this one writes.
__global__ __launch_bounds__(kRedBlock) void k_bicg_xr_dot(long n, double *__restrict__ x, double *__restrict__ r,
                                                          const double *__restrict__ px,
                                                          const double *__restrict__ sx,
                                                          const double *__restrict__ s,
                                                          const double *__restrict__ t,
                                                          const double *__restrict__ r0,
                                                          double *__restrict__ scal,
                                                          const double *__restrict__ wparts, int nparts,
                                                          double *__restrict__ partials) {
    if (scal[kScalGo] != 1.0) return;  // a stopped BiCGSTAB loop
    double tt, ts;
    final_sums2(wparts, nparts, tt, ts);
    const double w = (tt > 0) ? ts / tt : 0.0;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        scal[kScalW] = w;
        scal[kScalIter] += 1.0;  // the iteration ends here; no block reads the count
    }
    const double alpha = scal[kScalAlpha];
    double a = 0.0, b = 0.0;
#pragma unroll 4
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        x[i] += alpha * px[i] + w * sx[i];
        const double ri = s[i] - w * t[i];
        r[i] = ri;
        a += r0[i] * ri;
        b += ri * ri;
    }
    block_partials2(a, b, partials);
}

__global__ __launch_bounds__(kRedBlock) void k_dot2_final(const double *__restrict__ partials, int nparts,
                                                         double *__restrict__ scal, int op,
                                                         const double *__restrict__ gate, BicgTest test) {
    if (gate && *gate != 1.0) return;  // a stopped BiCGSTAB loop
    double ab, cd;
    final_sums2(partials, nparts, ab, cd);
    if (threadIdx.x != 0) return;
    switch (op) {
        case kOpStore:  // raw sums
            scal[kScalDotA] = ab;
            scal[kScalDotB] = cd;
            break;
        case kOpRho: {  // (r0.r, r.r): rho_old = rho; rho = r0.r; then the loop test
            scal[kScalRhoOld] = scal[kScalRho];
            scal[kScalRho] = ab;
            scal[kScalRR] = cd;
            // while (r.r > tol2 && i < maxit), then the restart test of the
            // loop body, in the host loop's order
            const bool more = cd > test.tol2 && scal[kScalIter] < test.maxit;
            scal[kScalGo] = !more ? 0.0 : (fabs(ab) < test.eps2 * scal[kScalR0Sq]) ? 2.0 : 1.0;
            break;
        }
        case kOpRestart:  // (r.r): rho = r0_sqnorm = r.r
            scal[kScalRho] = ab;
            scal[kScalR0Sq] = ab;
            scal[kScalRR] = ab;
            break;
    }
}

// BiCGSTAB vector updates (Eigen's internal::bicgstab recurrence); each
// runs only while the device loop flag scal[kScalGo] is 1.

// The top of a BiCGSTAB iteration, in every block: rho_old = rho, rho = r0.r
// and r.r from the (r0.r, r.r) partials, the loop test as k_dot2_final's
// kOpRho applies it (the iteration was counted by k_bicg_xr_dot), then
// p = r + beta (p - w v) while the loop goes on.  Block 0 writes kScalGo, which the other blocks read at their
// start: a block that sees the new value sees the value it would compute
// itself, so either way it updates p only if the loop goes on.  The flag is
// read once per block (one thread, then shared) so that the whole block
// leaves or stays together.
__global__ __launch_bounds__(kRedBlock) void k_bicg_p_top(long n, const double *__restrict__ r,
                                                         double *__restrict__ p, const double *__restrict__ v,
                                                         double *__restrict__ scal,
                                                         const double *__restrict__ parts, int nparts, int odd,
                                                         BicgTest test) {
    // the flag is the one slot that this launch both reads and writes:
    // relaxed atomic accesses (either value read leads to the same outcome)
    __shared__ double go_in;
    if (threadIdx.x == 0) go_in = __hip_atomic_load(scal + kScalGo, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __syncthreads();
    if (go_in != 1.0) return;  // a stopped BiCGSTAB loop
    double rho, rr;
    final_sums2(parts, nparts, rho, rr);
    const double rho_old = scal[odd ? kScalRhoAlt : kScalRho];
    // while (r.r > tol2 && i < maxit), then the restart test of the loop body
    const bool more = rr > test.tol2 && scal[kScalIter] < test.maxit;
    const double go = !more ? 0.0 : (fabs(rho) < test.eps2 * scal[kScalR0Sq]) ? 2.0 : 1.0;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        scal[kScalRhoOld] = rho_old;
        scal[odd ? kScalRho : kScalRhoAlt] = rho;
        scal[kScalRR] = rr;
        __hip_atomic_store(scal + kScalGo, go, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    if (go != 1.0) return;
    const double beta = (rho / rho_old) * (scal[kScalAlpha] / scal[kScalW]);
    const double w = scal[kScalW];
#pragma unroll 4
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        p[i] = r[i] + beta * (p[i] - w * v[i]);
}
__global__ __launch_bounds__(256) void k_bicg_p(long n, const double *__restrict__ r, double *__restrict__ p,
                                                const double *__restrict__ v, const double *__restrict__ scal) {
    const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i >= n || scal[kScalGo] != 1.0) return;
    const double beta = (scal[kScalRho] / scal[kScalRhoOld]) * (scal[kScalAlpha] / scal[kScalW]);
    const double w = scal[kScalW];
    p[i] = r[i] + beta * (p[i] - w * v[i]);
}

// s = r - alpha v, each block first forming alpha = rho / (r0.v) from the
// nparts partials of r0.v in k_dot2_final's order (block 0 also stores it;
// no block of this kernel reads the stored value).  rho is in kScalRho for
// even iterations of a group, kScalRhoAlt for odd ones (k_bicg_p_top).
__global__ __launch_bounds__(kRedBlock) void k_bicg_s(long n, const double *__restrict__ r,
                                                     const double *__restrict__ v, double *__restrict__ s,
                                                     double *__restrict__ scal,
                                                     const double *__restrict__ aparts, int nparts, int odd) {
    if (scal[kScalGo] != 1.0) return;  // a stopped BiCGSTAB loop
    double rv, unused;
    final_sums2(aparts, nparts, rv, unused);
    const double alpha = scal[odd ? kScalRhoAlt : kScalRho] / rv;
    if (blockIdx.x == 0 && threadIdx.x == 0) scal[kScalAlpha] = alpha;
#pragma unroll 4
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        s[i] = r[i] - alpha * v[i];
}

// Block-Jacobi preconditioner of the adjoint solves.  Probing: x = 1 on
// variable q of every cell of colour k (cells of one colour are more than
// two vertex rings apart, so none lies in another's residual stencil), y =
// J^T x; then y_(c,r) = dR_(c,q)/dU_(c,r) for every cell c of that colour.
__global__ __launch_bounds__(256) void k_probe_seed(int nc, const int *__restrict__ color, int k, int q,
                                                    double *__restrict__ x) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nc) return;
    for (int r = 0; r < 4; ++r) x[(long)r * nc + c] = (r == q && color[c] == k) ? 1.0 : 0.0;
}
__global__ __launch_bounds__(256) void k_probe_take(int nc, const int *__restrict__ color, int k, int q,
                                                    const double *__restrict__ y, double *__restrict__ D) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nc || color[c] != k) return;
    for (int r = 0; r < 4; ++r) D[16 * (long)c + 4 * q + r] = y[(long)r * nc + c];
}
// P_c = (D_c^T)^{-1} (the diagonal block of J^T is D_c^T), Gauss-Jordan
// with partial pivoting; a singular block keeps the identity.
__global__ __launch_bounds__(256) void k_invert_blocks(int nc, const double *__restrict__ D, double *__restrict__ P) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nc) return;
    double a[4][8];
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) {
            a[i][j] = D[16 * (long)c + 4 * j + i];  // transpose
            a[i][4 + j] = (i == j) ? 1.0 : 0.0;
        }
    bool ok = true;
    for (int col = 0; col < 4 && ok; ++col) {
        int piv = col;
        for (int i = col + 1; i < 4; ++i)
            if (fabs(a[i][col]) > fabs(a[piv][col])) piv = i;
        if (!(fabs(a[piv][col]) > 0.0)) {
            ok = false;
            break;
        }
        if (piv != col)
            for (int j = 0; j < 8; ++j) {
                const double t = a[col][j];
                a[col][j] = a[piv][j];
                a[piv][j] = t;
            }
        const double inv = 1.0 / a[col][col];
        for (int j = 0; j < 8; ++j) a[col][j] *= inv;
        for (int i = 0; i < 4; ++i)
            if (i != col) {
                const double f = a[i][col];
                for (int j = 0; j < 8; ++j) a[i][j] -= f * a[col][j];
            }
    }
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) P[16 * (long)c + 4 * i + j] = ok ? a[i][4 + j] : (i == j ? 1.0 : 0.0);
}
// out_c = P_c in_c (SoA vectors of 4 nc)
__global__ __launch_bounds__(256) void k_apply_blocks(int nc, const double *__restrict__ P,
                                                      const double *__restrict__ in, double *__restrict__ out) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nc) return;
    double v[4];
    for (int r = 0; r < 4; ++r) v[r] = in[(long)r * nc + c];
    for (int i = 0; i < 4; ++i) {
        const double *row = P + 16 * (long)c + 4 * i;
        out[(long)i * nc + c] = row[0] * v[0] + row[1] * v[1] + row[2] * v[2] + row[3] * v[3];
    }
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
    double L = 0.0, D = 0.0;
    for (int base = 0; base < nw; base += kRedBlock) {
        const int k = base + threadIdx.x;
        double l = 0.0, d = 0.0;
        bool use = false;
        if (k < nw) {
            const int e = wall[k];
            use = !(skip_x && G.emx[e] > 1.3);
            const double nx = G.enx[e], ny = G.eny[e], A = G.elen[e];
            const double rl = fL[4 * e], rul = fL[4 * e + 1], rvl = fL[4 * e + 2], El = fL[4 * e + 3];
            const double p = (gamma - 1.0) * (El - 0.5 * (rul * rul + rvl * rvl) / rl);
            const double txx = 0.0, txy = 0.0, tyy = 0.0;
            l = (-p * ny + txy * nx + tyy * ny) * A;
            d = (-p * nx + txx * nx + txy * ny) * A;
        }
        tl[threadIdx.x] = use ? l : 0.0;
        td[threadIdx.x] = use ? d : 0.0;
        __syncthreads();
        // the serial sum in wall order; a skipped edge adds +0.0, which
        // changes nothing (the sums start at +0.0 and so are never -0.0),
        // so the loop has no branch and its LDS reads run ahead
        if (threadIdx.x == 0) {
            const int m = min(kRedBlock, nw - base);
#pragma unroll 8
            for (int j = 0; j < m; ++j) {
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
__global__ __launch_bounds__(256) void k_adjoint_rhs(Topo T, Geom G, const int *__restrict__ wall_by_cell,
                                                    const int *__restrict__ grp, int ngrp,
                                                    const double *__restrict__ fL, double gamma, double q_inf,
                                                    double beta, double *__restrict__ b) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= ngrp) return;
    const int nc = T.nc;
    const double corde = 1;
    // one wall cell: its wall edges in ascending edge id, the order in
    // which the serial loop adds them into this cell's entries
    for (int j = grp[g]; j < grp[g + 1]; ++j) {
        const int e = wall_by_cell[j];
        const int iL = T.ecl[e];
        const double A_f = G.elen[e], nx = G.enx[e], ny = G.eny[e];
        const double rL = fL[4 * e], ruL = fL[4 * e + 1], rvL = fL[4 * e + 2];
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
// FD flux-Jacobian blocks per edge at the frozen face states, one 256-byte
// record per edge: blk[32 e + r*4+c] = dF_r/dU_L,c and
// blk[32 e + 16 + r*4+c] = dF_r/dU_R,c, so each side's block is one
// 128-byte line.  A boundary edge rebuilds its ghost from each perturbed
// left state; its right-state block is never read by the scatter and is
// stored as zero.
__global__ __launch_bounds__(256) void k_jt_blocks(Topo T, Geom G, const double *__restrict__ fL,
                                                   const double *__restrict__ fR, double eps, Phys P,
                                                   double *__restrict__ blk) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= T.ne) return;
    double *rec = blk + 32 * (size_t)e;
    const char bc = T.ebc[e];
    const double enx = G.enx[e], eny = G.eny[e];
    double bL[4], bR[4];
    ld4(fL, e, bL);
    ld4(fR, e, bR);
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
        for (int r = 0; r < 4; ++r) rec[r * 4 + c] = (Fp[r] - Fm[r]) / (2 * eps);
    }
    for (int c = 0; c < 4; ++c) {
        if (bc != '\0') {
#pragma unroll
            for (int r = 0; r < 4; ++r) rec[16 + r * 4 + c] = 0.0;
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
        for (int r = 0; r < 4; ++r) rec[16 + r * 4 + c] = (Fp[r] - Fm[r]) / (2 * eps);
    }
}

// y = J^T x as a per-cell gather in ascending edge id: the serialisation of
// the reference's atomic scatter in edge order, without atomics.  Four
// lanes per cell: lane k forms edge slot k's contribution F_sideᵀ Delta
// (its side's block read as one 128-byte record), the in-quad transpose
// gives lane q variable q of every slot, and lane q sums them in slot order
// (+0.0 from a slot without a contribution: the sum starts at +0.0, so it
// is never -0.0 and adding +0.0 changes nothing).
__global__ __launch_bounds__(256) void k_jt_apply(Topo T, Geom G, const double *__restrict__ blk,
                                                  const double *__restrict__ x, double *__restrict__ y) {
    const int gid = blockIdx.x * blockDim.x + threadIdx.x;
    const int c = gid / kLanesPerCell, k = gid % kLanesPerCell;
    if (c >= T.nc) return;  // whole lane groups only
    const int nc = T.nc;
    const unsigned info = T.cinfo[c];
    const int e = T.cedge[4 * c + k];  // -1 past the cell's edges
    double g[4] = {0.0, 0.0, 0.0, 0.0};
    if (e >= 0) {
        const bool left = (info >> k) & 1u;  // c is the edge's left cell
        const double L = G.elen[e];
        const int iL = T.ecl[e];
        const char bc = T.ebc[e];
        typedef double v2d __attribute__((ext_vector_type(2)));
        const v2d *rp = reinterpret_cast<const v2d *>(blk + 32 * (size_t)e + (left ? 0 : 16));
        double F[16];
#pragma unroll
        for (int h = 0; h < 8; ++h) {
            const v2d t = rp[h];
            F[2 * h] = t.x;
            F[2 * h + 1] = t.y;
        }
        const double invV_L = 1.0 / G.cvol[iL];
        double D[4];
        bool use = true;
        if (bc != '\0') {
            use = left;
#pragma unroll
            for (int q = 0; q < 4; ++q) D[q] = (-x[q * nc + iL] * invV_L) * L;
        } else {
            const int iR = T.ecr[e];
            const double invV_R = 1.0 / G.cvol[iR];
#pragma unroll
            for (int q = 0; q < 4; ++q) D[q] = (-x[q * nc + iL] * invV_L + x[q * nc + iR] * invV_R) * L;
        }
        if (use) {
#pragma unroll
            for (int kk = 0; kk < 4; ++kk)  // column kk of the block: rows r at stride 4
                g[kk] = (F[kk] * D[0] + F[4 + kk] * D[1] + F[8 + kk] * D[2] + F[12 + kk] * D[3]);
        }
    }
    quad_transpose(g, k);
    double acc = 0.0;
#pragma unroll
    for (int j = 0; j < 4; ++j) acc += g[j];
    y[k * nc + c] = acc;
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
// Batched stage kernels (one launch for every member of a batch)
// ------------------------------------------------------------------------
template <class F>
__device__ __forceinline__ void each_tile(const int2 *__restrict__ tiles, int ntiles, F &&f) {
    for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const int2 tl = tiles[t];
        f(tl.x, tl.y);
    }
}

__global__ __launch_bounds__(256) void kb_speed_max(const Member *__restrict__ M, const int2 *__restrict__ tiles,
                                                    int ntiles) {
    each_tile(tiles, ntiles, [&](int m, int b) {
        const Member &A = M[m];
        speed_max_body(b, A.U, A.T.nc, A.P, A.speed);
    });
}

__global__ __launch_bounds__(256) void kb_step_begin(const Member *__restrict__ M) {
    const Member &A = M[blockIdx.x];
    step_begin_body(A.speed, A.nparts, A.P, A.min_h, A.scal, nullptr, 1);
}

__global__ __launch_bounds__(256) void kb_vertex_gather(const Member *__restrict__ M,
                                                        const int2 *__restrict__ tiles, int ntiles) {
    each_tile(tiles, ntiles, [&](int m, int b) {
        const Member &A = M[m];
        const int v = b * blockDim.x + threadIdx.x;
        if (v < A.T.nv && *A.P.gate != 0.0) gather_vertex<4, false>(A.T, A.G, A.U, A.vU, A.P, v);
    });
}

__global__ __launch_bounds__(256) void kb_reconstruct(const Member *__restrict__ M, const int2 *__restrict__ tiles,
                                                      int ntiles) {
    each_tile(tiles, ntiles, [&](int m, int b) {
        const Member &A = M[m];
        reconstruct_body(b, A.T, A.G, A.U, A.vU, A.fL, A.fR, A.tq, nullptr, A.P);
    });
}

__global__ __launch_bounds__(256) void kb_visc_edge(const Member *__restrict__ M, const int2 *__restrict__ tiles,
                                                    int ntiles) {
    each_tile(tiles, ntiles, [&](int m, int b) {
        const Member &A = M[m];
        const int nbv = (A.T.nv + (int)blockDim.x - 1) / (int)blockDim.x;
        visc_edge_body(b, A.T, A.G, A.tq, A.vtq, A.fL, A.fR, A.Fe, A.P, nbv, 1);
    });
}

__global__ __launch_bounds__(256) void kb_flux_update(const Member *__restrict__ M, const int2 *__restrict__ tiles,
                                                      int ntiles, double c0, double c1, int flags) {
    each_tile(tiles, ntiles, [&](int m, int b) {
        const Member &A = M[m];
        flux_update_body(b, A.T, A.G, A.U, A.Uold, A.Fe, A.vtq, A.R, A.scal + kScalDt, c0, c1, flags, A.P,
                         A.speed);
    });
}

// Sum of R_rho^2 over cells [chunk kResChunk, +kResChunk): thread t adds
// cells t, t + 256, ... in order, then a fixed tree.
// Sum of R_rho^2 over cells [chunk kResChunk, +kResChunk) in a fixed
// 256-wide tree: partial `chunk` of the density residual.
__device__ __forceinline__ void residual_partial_body(const double *__restrict__ R, int nc, int chunk,
                                                      double *__restrict__ rparts) {
    __shared__ double sm[256];
    const int c0 = chunk * kResChunk, c1 = min(nc, c0 + kResChunk);
    double acc = 0.0;
    for (int c = c0 + threadIdx.x; c < c1; c += 256) acc += R[c] * R[c];
    sm[threadIdx.x] = acc;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if (threadIdx.x < st) sm[threadIdx.x] += sm[threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x == 0) rparts[chunk] = sm[0];
    __syncthreads();
}

// r = sqrt(mean_c R_rho^2) / (rho_inf U_inf / c) from the partials in order
// (one block); sets the converged flag once r <= tol.
__device__ __forceinline__ void residual_final_body(const double *__restrict__ rparts, int nrparts, int nc,
                                                    double res_scale, double tol, double *__restrict__ scal) {
    __shared__ double sm[256];
    double acc = 0.0;
    for (int k = threadIdx.x; k < nrparts; k += 256) acc += rparts[k];
    sm[threadIdx.x] = acc;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if (threadIdx.x < st) sm[threadIdx.x] += sm[threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const double r = sqrt(sm[0] / nc) / res_scale;
        scal[kScalRes] = r;
        if (r <= tol) scal[kScalConv] = 1.0;
    }
}

__global__ __launch_bounds__(256) void kb_residual_partial(const Member *__restrict__ M,
                                                           const int2 *__restrict__ tiles, int ntiles) {
    each_tile(tiles, ntiles, [&](int m, int chunk) {
        const Member &A = M[m];
        if (*A.P.gate == 0.0) return;  // uniform per tile
        residual_partial_body(A.R, A.T.nc, chunk, A.rparts);
    });
}

__global__ __launch_bounds__(256) void kb_residual_final(const Member *__restrict__ M) {
    const Member &A = M[blockIdx.x];
    if (*A.P.gate == 0.0) return;
    residual_final_body(A.rparts, A.nrparts, A.T.nc, A.res_scale, A.res_tol, A.scal);
}

// The same for one mesh (its gate, one block per chunk, then one block).
__global__ __launch_bounds__(256) void k_residual_partial(const double *__restrict__ R, int nc,
                                                          double *__restrict__ rparts, const double *gate) {
    if (gate && *gate == 0.0) return;
    residual_partial_body(R, nc, blockIdx.x, rparts);
}
__global__ __launch_bounds__(256) void k_residual_final(const double *__restrict__ rparts, int nrparts, int nc,
                                                        double res_scale, double tol, double *__restrict__ scal,
                                                        const double *gate) {
    if (gate && *gate == 0.0) return;
    residual_final_body(rparts, nrparts, nc, res_scale, tol, scal);
}

__global__ __launch_bounds__(64) void kb_collect(const Member *__restrict__ M, int n, double *__restrict__ flags) {
    for (int i = threadIdx.x; i < n; i += blockDim.x) flags[i] = M[i].scal[kScalActive];
}

// ------------------------------------------------------------------------
// Launch helpers
// ------------------------------------------------------------------------
static inline unsigned nblk(long n, int b = 256) { return (unsigned)((n + b - 1) / b); }

hipError_t launch_vertex_state(const Topo &T, const Geom &G, const double *U, double *vU, const Phys &P,
                               hipStream_t s) {
    hipLaunchKernelGGL((k_vertex_gather<4, false>), dim3(nblk(T.nv)), dim3(256), 0, s, T, G, U, vU, P);
    return hipGetLastError();
}
hipError_t launch_vertex_visc(const Topo &T, const Geom &G, const double *tq, double *vtq, const Phys &P,
                              hipStream_t s) {
    hipLaunchKernelGGL((k_vertex_gather<5, true>), dim3(nblk(T.nv)), dim3(256), 0, s, T, G, tq, vtq, P);
    return hipGetLastError();
}
hipError_t launch_reconstruct(const Topo &T, const Geom &G, const double *U, const double *vU, double *fL,
                              double *fR, double *tq, double *Gout, const Phys &P, hipStream_t s) {
    hipLaunchKernelGGL(k_cell_reconstruct, dim3(nblk((long)kLanesPerCell * T.nc)), dim3(256), 0, s, T, G, U, vU, fL,
                       fR, tq, Gout, P);
    return hipGetLastError();
}
hipError_t launch_stress(const Topo &T, const double *U, const double *Gs, double *tq, const Phys &P,
                         hipStream_t s) {
    hipLaunchKernelGGL(k_cell_stress, dim3(nblk(T.nc)), dim3(256), 0, s, T, U, Gs, tq, P);
    return hipGetLastError();
}
hipError_t launch_flux_update(const Topo &T, const Geom &G, double *U, double *Uold, const double *Fe,
                              const double *vtq, double *R, const double *dtp, double c0, double c1, int flags,
                              const Phys &P, hipStream_t s, double *speed) {
    hipLaunchKernelGGL(k_cell_flux_update, dim3(nblk((long)kLanesPerCell * T.nc)), dim3(256), 0, s, T, G, U, Uold,
                       Fe, vtq, R, dtp, c0, c1, flags, P, speed);
    return hipGetLastError();
}
hipError_t launch_visc_and_flux(const Topo &T, const Geom &G, const double *tq, double *vtq, const double *fL,
                                const double *fR, double *Fe, const Phys &P, int do_visc, hipStream_t s) {
    const int nbv = (int)nblk(T.nv);
    hipLaunchKernelGGL(k_viscous_gather_edge_flux, dim3(nbv + nblk(T.ne)), dim3(256), 0, s, T, G, tq, vtq, fL, fR,
                       Fe, P, nbv, do_visc);
    return hipGetLastError();
}
int speed_parts(int nc) { return (int)nblk((long)kLanesPerCell * nc); }
hipError_t launch_speed_max(const double *U, int nc, const Phys &P, double *parts, hipStream_t s) {
    hipLaunchKernelGGL(k_speed_max, dim3(speed_parts(nc)), dim3(256), 0, s, U, nc, P, parts);
    return hipGetLastError();
}
hipError_t launch_step_gather(const Topo &T, const Geom &G, const double *U, double *vU, const Phys &P,
                              const double *parts, int nparts, double min_h, double *scal, double *hist,
                              int advance_t, hipStream_t s) {
    hipLaunchKernelGGL(k_step_gather, dim3(nblk(T.nv) + 1), dim3(256), 0, s, T, G, U, vU, P, parts, nparts, min_h,
                       scal, hist, advance_t);
    return hipGetLastError();
}
hipError_t launch_step_begin(const double *parts, int nparts, const Phys &P, double min_h, double *scal,
                             double *hist, int advance_t, hipStream_t s) {
    hipLaunchKernelGGL(k_step_begin, dim3(1), dim3(256), 0, s, parts, nparts, P, min_h, scal, hist, advance_t);
    return hipGetLastError();
}
// Blocks of a reduction over n elements (fixed, so the order is fixed).
static unsigned red_grid(long n) {
    return (unsigned)std::min<long>(kRedGrid, std::max<long>(1, (long)nblk(n, kRedBlock)));
}
hipError_t launch_bicg_xr_dot(long n, double *x, double *r, const double *px, const double *sx, const double *sv,
                              const double *t, const double *r0, const double *wparts, double *partials,
                              double *scal, hipStream_t s) {
    const unsigned g = red_grid(n);
    hipLaunchKernelGGL(k_bicg_xr_dot, dim3(g), dim3(kRedBlock), 0, s, n, x, r, px, sx, sv, t, r0, scal, wparts,
                       (int)g, partials);
    return hipGetLastError();
}
hipError_t launch_bicg_p_top(long n, const double *r, double *p, const double *v, double *scal,
                             const double *parts, int odd, const BicgTest &test, hipStream_t s) {
    hipLaunchKernelGGL(k_bicg_p_top, dim3(red_grid(n)), dim3(kRedBlock), 0, s, n, r, p, v, scal, parts,
                       (int)red_grid(n), odd, test);
    return hipGetLastError();
}
hipError_t launch_dot2_partial(const double *a, const double *b, const double *c, const double *d, long n,
                               double *partials, const double *gate, hipStream_t s) {
    hipLaunchKernelGGL(k_dot2_partial, dim3(red_grid(n)), dim3(kRedBlock), 0, s, a, b, c, d, n, partials, gate);
    return hipGetLastError();
}
hipError_t launch_dot2(const double *a, const double *b, const double *c, const double *d, long n,
                       double *partials, double *scal, int op, hipStream_t s, const double *gate,
                       const BicgTest &test) {
    const unsigned g = red_grid(n);
    hipLaunchKernelGGL(k_dot2_partial, dim3(g), dim3(kRedBlock), 0, s, a, b, c, d, n, partials, gate);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_dot2_final, dim3(1), dim3(kRedBlock), 0, s, partials, (int)g, scal, op, gate, test);
    return hipGetLastError();
}
hipError_t launch_bicg_p(long n, const double *r, double *p, const double *v, const double *scal, hipStream_t s) {
    hipLaunchKernelGGL(k_bicg_p, dim3(nblk(n)), dim3(256), 0, s, n, r, p, v, scal);
    return hipGetLastError();
}
hipError_t launch_bicg_s(long n, const double *r, const double *v, double *sv, double *scal, const double *aparts,
                         int odd, hipStream_t s) {
    hipLaunchKernelGGL(k_bicg_s, dim3(red_grid(n)), dim3(kRedBlock), 0, s, n, r, v, sv, scal, aparts,
                       (int)red_grid(n), odd);
    return hipGetLastError();
}
hipError_t launch_probe_seed(int nc, const int *color, int k, int q, double *x, hipStream_t s) {
    hipLaunchKernelGGL(k_probe_seed, dim3(nblk(nc)), dim3(256), 0, s, nc, color, k, q, x);
    return hipGetLastError();
}
hipError_t launch_probe_take(int nc, const int *color, int k, int q, const double *y, double *D, hipStream_t s) {
    hipLaunchKernelGGL(k_probe_take, dim3(nblk(nc)), dim3(256), 0, s, nc, color, k, q, y, D);
    return hipGetLastError();
}
hipError_t launch_invert_blocks(int nc, const double *D, double *P, hipStream_t s) {
    hipLaunchKernelGGL(k_invert_blocks, dim3(nblk(nc)), dim3(256), 0, s, nc, D, P);
    return hipGetLastError();
}
hipError_t launch_apply_blocks(int nc, const double *P, const double *in, double *out, hipStream_t s) {
    hipLaunchKernelGGL(k_apply_blocks, dim3(nblk(nc)), dim3(256), 0, s, nc, P, in, out);
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
hipError_t launch_adjoint_rhs(const Topo &T, const Geom &G, const int *wall_by_cell, const int *grp, int ngrp,
                              const double *fL, double gamma, double q_inf, double beta, double *b,
                              hipStream_t s) {
    if (ngrp == 0) return hipSuccess;
    hipLaunchKernelGGL(k_adjoint_rhs, dim3(nblk(ngrp)), dim3(256), 0, s, T, G, wall_by_cell, grp, ngrp, fL, gamma,
                       q_inf, beta, b);
    return hipGetLastError();
}
hipError_t launch_jt_blocks(const Topo &T, const Geom &G, const double *fL, const double *fR, double eps,
                            const Phys &P, double *blk, hipStream_t s) {
    hipLaunchKernelGGL(k_jt_blocks, dim3(nblk(T.ne)), dim3(256), 0, s, T, G, fL, fR, eps, P, blk);
    return hipGetLastError();
}
hipError_t launch_jt_apply(const Topo &T, const Geom &G, const double *blk, const double *x, double *y,
                           hipStream_t s) {
    hipLaunchKernelGGL(k_jt_apply, dim3(nblk((long)kLanesPerCell * T.nc)), dim3(256), 0, s, T, G, blk, x, y);
    return hipGetLastError();
}
hipError_t launch_hllc_batch(const double *qL, const double *qR, const double *nrm, double gamma, double *F,
                             int n, hipStream_t s) {
    hipLaunchKernelGGL(k_hllc_batch, dim3(nblk(n)), dim3(256), 0, s, qL, qR, nrm, gamma, F, n);
    return hipGetLastError();
}

static inline dim3 tile_grid(int ntiles, TileShape sh) {
    return dim3((unsigned)std::max(1, sh.grid > 0 ? std::min(sh.grid, ntiles) : ntiles));
}
hipError_t launch_b_speed_max(const Member *M, const int2 *tiles, int ntiles, TileShape sh, hipStream_t s) {
    hipLaunchKernelGGL(kb_speed_max, tile_grid(ntiles, sh), dim3(sh.block), 0, s, M, tiles, ntiles);
    return hipGetLastError();
}
hipError_t launch_b_step_begin(const Member *M, int nmembers, hipStream_t s) {
    hipLaunchKernelGGL(kb_step_begin, dim3(nmembers), dim3(256), 0, s, M);
    return hipGetLastError();
}
hipError_t launch_b_vertex_gather(const Member *M, const int2 *tiles, int ntiles, TileShape sh, hipStream_t s) {
    hipLaunchKernelGGL(kb_vertex_gather, tile_grid(ntiles, sh), dim3(sh.block), 0, s, M, tiles, ntiles);
    return hipGetLastError();
}
hipError_t launch_b_reconstruct(const Member *M, const int2 *tiles, int ntiles, TileShape sh, hipStream_t s) {
    hipLaunchKernelGGL(kb_reconstruct, tile_grid(ntiles, sh), dim3(sh.block), 0, s, M, tiles, ntiles);
    return hipGetLastError();
}
hipError_t launch_b_visc_edge(const Member *M, const int2 *tiles, int ntiles, TileShape sh, hipStream_t s) {
    hipLaunchKernelGGL(kb_visc_edge, tile_grid(ntiles, sh), dim3(sh.block), 0, s, M, tiles, ntiles);
    return hipGetLastError();
}
hipError_t launch_b_flux_update(const Member *M, const int2 *tiles, int ntiles, TileShape sh, double c0, double c1,
                                int flags, hipStream_t s) {
    hipLaunchKernelGGL(kb_flux_update, tile_grid(ntiles, sh), dim3(sh.block), 0, s, M, tiles, ntiles, c0, c1,
                       flags);
    return hipGetLastError();
}
hipError_t launch_b_residual(const Member *M, const int2 *chunk_tiles, int nchunks, int nmembers, hipStream_t s) {
    hipLaunchKernelGGL(kb_residual_partial, dim3((unsigned)std::max(1, nchunks)), dim3(256), 0, s, M, chunk_tiles,
                       nchunks);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(kb_residual_final, dim3(nmembers), dim3(256), 0, s, M);
    return hipGetLastError();
}
hipError_t launch_residual_check(const double *R, int nc, double *rparts, double res_scale, double tol,
                                 double *scal, const double *gate, hipStream_t s) {
    const int nparts = (nc + kResChunk - 1) / kResChunk;
    hipLaunchKernelGGL(k_residual_partial, dim3((unsigned)std::max(1, nparts)), dim3(256), 0, s, R, nc, rparts, gate);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_residual_final, dim3(1), dim3(256), 0, s, rparts, nparts, nc, res_scale, tol, scal, gate);
    return hipGetLastError();
}
hipError_t launch_b_collect(const Member *M, int nmembers, double *flags, hipStream_t s) {
    hipLaunchKernelGGL(kb_collect, dim3(1), dim3(64), 0, s, M, nmembers, flags);
    return hipGetLastError();
}

}  // namespace o2f

