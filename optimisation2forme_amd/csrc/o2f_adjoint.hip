// Exact discrete adjoint of the residual on gfx950: y = (dR/dU)^T x.
//
// R(U) is the full residual of computeResidual (advection then diffusion,
// src/shape_optimization.cpp:171-189) as the forward kernels evaluate it.
// The reference's applyJacobianTranspose (src/shape_optimization.cpp:237-484)
// freezes gradients and limiters and differentiates HLLC by finite
// differences; here every dependence is differentiated exactly: the
// cell -> vertex interpolation and its boundary states, the edge values,
// the Green-Gauss gradients, the neighbour min / max and limiter (through the
// active branch), the face states and ghosts, HLLC, the stresses and heat
// flux, and the viscous divergence.  The forward quantities each reverse
// step needs (gradients, min / max, limiter and its active edge) are
// recomputed per cell, not stored; only the forward vertex values, face
// states and vertex stresses of the linearisation point are read.
//
// Six gather-only launches (no atomics, deterministic):
//   ka_edge        per edge:   dR/dF of both cells, HLLC vector-Jacobian
//                              product (forward-mode duals, 8 directions),
//                              ghost-state transpose; viscous edge terms
//   ka_vertex_visc per vertex: adjoint of the vertex stress gather
//   ka_cell        per cell (four lanes): adjoint of the face states,
//                              limiter, min / max, stresses and the
//                              cell-centre velocity of the viscous work
//   ka_edge_value  per edge:   adjoint of the Green-Gauss edge values
//   ka_vertex_state per vertex: adjoint of the vertex state boundary states
//   ka_cell_final  per cell:   transpose of the interpolation stencil plus
//                              the neighbour min / max contributions
#include "o2f_kernels.hpp"

#include "o2f_device.hpp"

namespace o2f {

namespace {

__device__ __forceinline__ int from_lane_int(int v, int j) {
    double r = 0.0;
    for_lanes(static_cast<double>(v), [&](int i, double a) {
        if (i == j) r = a;
    });
    return static_cast<int>(r);
}

// Transpose of the wall state map (r, ru, rv, E) -> (r, 0, 0, E - (ru^2 +
// rv^2) / (2 r)) of apply_inflow_outflow_bc (src/boundary.cpp:30-40) at w.
__device__ __forceinline__ void wall_state_t(const double (&w)[4], const double (&s)[4], double (&o)[4]) {
    const double r = w[0], ru = w[1], rv = w[2];
    o[0] = s[0] + s[3] * (0.5 * (ru * ru + rv * rv) / (r * r));
    o[1] = s[3] * (-ru / r);
    o[2] = s[3] * (-rv / r);
    o[3] = s[3];
}

// Transpose of bc_state (vertex and edge state boundary conditions).
__device__ __forceinline__ void bc_state_t(char bc, const double (&w)[4], const double (&s)[4], double (&o)[4]) {
    if (bc == 'i' || bc == 'f') {
        o[0] = o[1] = o[2] = o[3] = 0.0;
    } else if (bc == 'w') {
        wall_state_t(w, s, o);
    } else {
        o[0] = s[0];
        o[1] = s[1];
        o[2] = s[2];
        o[3] = s[3];
    }
}

}  // namespace

// ---------------------------------------------------------------- edges ---
__global__ __launch_bounds__(256) void ka_edge(Topo T, Geom G, Phys P, const double *__restrict__ x,
                                               const double *__restrict__ U, const double *__restrict__ fL,
                                               const double *__restrict__ fR, AdjWork W) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= T.ne) return;
    const int nc = T.nc, ne = T.ne;
    const int cl = T.ecl[e], cr = T.ecr[e];
    const double el = G.elen[e], enx = G.enx[e], eny = G.eny[e];
    const char ebc = T.ebc[e];
    // R_c = -(1/V_c) sum_e s F_e L_e, s = +1 for the left cell
    double Fb[4];
    const double vl = G.cvol[cl];
#pragma unroll
    for (int q = 0; q < 4; ++q) Fb[q] = -el * x[q * nc + cl] / vl;
    if (cr >= 0) {
        const double vr = G.cvol[cr];
#pragma unroll
        for (int q = 0; q < 4; ++q) Fb[q] += el * x[q * nc + cr] / vr;
    }
    double qL[4], qR[4], bL[4], bR[4];
    ld4(fL, e, qL);
    ld4(fR, e, qR);
    // Fb^T dF/dq, two directions per HLLC evaluation: pass j takes
    // components 2 (j % 2), 2 (j % 2) + 1 of the left (j < 2) or right state
    double bq[8];
#pragma unroll 1
    for (int j = 0; j < 4; ++j) {
        DualN<2> dl[4], dr[4], F[4];
        const int side = j >> 1, q0 = 2 * (j & 1);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            dl[q] = DualN<2>::seed(qL[q], side == 0 ? q - q0 : -1);
            dr[q] = DualN<2>::seed(qR[q], side == 1 ? q - q0 : -1);
        }
        hllc_t<DualN<2>>(dl, dr, enx, eny, P.gamma, F);
#pragma unroll
        for (int i = 0; i < 2; ++i)
            bq[2 * j + i] = Fb[0] * F[0].d[i] + Fb[1] * F[1].d[i] + Fb[2] * F[2].d[i] + Fb[3] * F[3].d[i];
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        bL[q] = bq[q];
        bR[q] = bq[4 + q];
    }
    if (cr < 0) {
        // ghost R state = function of the L state (src/boundary.cpp:48-80)
        if (ebc == 'o') {
#pragma unroll
            for (int q = 0; q < 4; ++q) bL[q] += bR[q];
        } else if (ebc == 'w') {
            bL[0] += bR[0];
            bL[1] -= bR[1];
            bL[2] -= bR[2];
            bL[3] += bR[3];
        }
        bR[0] = bR[1] = bR[2] = bR[3] = 0.0;
    }
    st4(W.fbL, e, bL);
    st4(W.fbR, e, bR);
    // viscous divergence (src/diffusion.cpp:183-200): adjoint of the edge's
    // (txx, tyy, txy, qx, qy) from each cell's three momentum / energy rows
    double tb[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    auto add_cell = [&](int c, double nx, double ny) {
        const double V = G.cvol[c];
        const double x1 = x[nc + c], x2 = x[2 * nc + c], x3 = x[3 * nc + c];
        const double rho = U[c], uu = U[nc + c] / rho, vv = U[2 * nc + c] / rho;
        const double a = el * nx / V, b = el * ny / V;
        tb[0] += x1 * a + x3 * uu * a;
        tb[1] += x2 * b + x3 * vv * b;
        tb[2] += x1 * b + x2 * a + x3 * (vv * a + uu * b);
        tb[3] += -x3 * a;
        tb[4] += -x3 * b;
    };
    add_cell(cl, -enx, -eny);
    if (cr >= 0) add_cell(cr, enx, eny);
    const bool zero = (ebc == 'f' || ebc == 'w');
    const double t4[4] = {zero ? 0.0 : 0.5 * tb[0], zero ? 0.0 : 0.5 * tb[1], zero ? 0.0 : 0.5 * tb[2],
                          zero ? 0.0 : 0.5 * tb[3]};
    st4(W.th, e, t4);
    W.th[4 * (long)ne + e] = zero ? 0.0 : 0.5 * tb[4];
}

// Objective seed: dJ/d(face L state) of J = -C_L + beta C_D over the wall
// edges computeDirectObjectiveTerm counts (src/shape_optimization.cpp:
// 893-939: p from the L state, edges with centre x > 1.3 skipped, wall
// stresses zero), added to the adjoint face states.  With x = 0 the rest of
// the chain then carries it to (dJ/dU)^T through the reconstruction.
__global__ __launch_bounds__(256) void ka_seed_objective(Geom G, const int *__restrict__ wall, int nw,
                                                         const double *__restrict__ fL, double gamma,
                                                         double q_inf, double beta, AdjWork W) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nw) return;
    const int e = wall[k];
    if (G.emx[e] > 1.3) return;
    const double nx = G.enx[e], ny = G.eny[e], A = G.elen[e];
    double q[4], b[4];
    ld4(fL, e, q);
    ld4(W.fbL, e, b);
    const double r = q[0], ru = q[1], rv = q[2];
    // J = sum (p ny A - beta p nx A) / (q_inf c), c = 1
    const double dJdp = (ny * A - beta * nx * A) / q_inf;
    const double g1 = gamma - 1.0;
    b[0] += dJdp * g1 * 0.5 * (ru * ru + rv * rv) / (r * r);
    b[1] += dJdp * (-g1 * ru / r);
    b[2] += dJdp * (-g1 * rv / r);
    b[3] += dJdp * g1;
    st4(W.fbL, e, b);
}

// ------------------------------------------------------------- vertices ---
__global__ __launch_bounds__(256) void ka_vertex_visc(Topo T, AdjTopo A, AdjWork W) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= T.nv) return;
    const long ne = T.ne;
    double s[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    const char bc = T.vbc[v];
    if (!(bc == 'f' || bc == 'w')) {
        for (int k = A.veptr[v]; k < A.veptr[v + 1]; ++k) {
            const int e = A.vedge[k];
            double t4[4];
            ld4(W.th, e, t4);
#pragma unroll
            for (int q = 0; q < 4; ++q) s[q] += t4[q];
            s[4] += W.th[4 * ne + e];
        }
    }
    const double o4[4] = {s[0], s[1], s[2], s[3]};
    st4(W.vtb, v, o4);
    W.vtb[4 * (long)T.nv + v] = s[4];
}

// ---------------------------------------------------------------- cells ---
__global__ __launch_bounds__(256) void ka_cell(Topo T, Geom G, Phys P, AdjTopo A, const double *__restrict__ x,
                                               const double *__restrict__ U, const double *__restrict__ vU,
                                               const double *__restrict__ vtq, AdjWork W) {
    const int gid = blockIdx.x * blockDim.x + threadIdx.x;
    const int c = gid / kLanesPerCell, k = gid % kLanesPerCell;
    if (c >= T.nc) return;  // whole lane groups only
    const int nc = T.nc, nv = T.nv;
    const unsigned info = T.cinfo[c];
    const int nk = info >> 4;
    const bool has_edge = k < nk;
    const bool is_left = (info >> k) & 1u;
    double u[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) u[q] = U[q * nc + c];
    const double vol = G.cvol[c], cx = G.ccx[c], cy = G.ccy[c];
    const int e = has_edge ? T.cedge[4 * c + k] : 0;
    const int nb = T.cnbr[4 * c + k];
    double el = 0.0, enx = 0.0, eny = 0.0, dx = 0.0, dy = 0.0, xa[4], xb[4], un[4];
    int va = 0, vb = 0;
    char ebc = 0;
    if (has_edge) {
        va = T.cevx[8 * c + 2 * k];
        vb = T.cevx[8 * c + 2 * k + 1];
        el = G.elen[e];
        enx = G.enx[e];
        eny = G.eny[e];
        dx = G.emx[e] - cx;
        dy = G.emy[e] - cy;
        ebc = T.ebc[e];
        ld4(vU, va, xa);
        ld4(vU, vb, xb);
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) un[q] = (nb >= 0) ? U[q * nc + nb] : u[q];
    double nx = enx, ny = eny;
    if (is_left) {
        nx = -nx;
        ny = -ny;
    }
    // ---- forward recomputation (the arithmetic of the reconstruction kernel)
    double tx[4] = {0.0, 0.0, 0.0, 0.0}, ty[4] = {0.0, 0.0, 0.0, 0.0};
    if (has_edge) {
        double val[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) val[q] = 0.5 * (xa[q] + xb[q]);
        if (tagged(ebc)) bc_state(P, ebc, val[0], val[1], val[2], val[3]);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            tx[q] = val[q] * el * nx / vol;
            ty[q] = val[q] * el * ny / vol;
        }
    }
    double gx[4] = {0.0, 0.0, 0.0, 0.0}, gy[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        for_lanes(tx[q], [&](int j, double a) {
            if (j < nk) gx[q] += a;
        });
        for_lanes(ty[q], [&](int j, double b) {
            if (j < nk) gy[q] += b;
        });
    }
    double mn[4], mx[4];
    int imn[4], imx[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        mn[q] = mx[q] = u[q];
        imn[q] = imx[q] = -1;
        for_lanes(un[q], [&](int j, double v) {
            if (v < mn[q]) {
                mn[q] = v;
                imn[q] = j;
            }
            if (v > mx[q]) {
                mx[q] = v;
                imx[q] = j;
            }
        });
    }
    double a_k[4] = {1.0, 1.0, 1.0, 1.0};
    if (has_edge) {
        const double veps = (P.limiter == 1) ? 0.0 : venkat_eps(vol, P.venkat_k);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const double pred = u[q] + gx[q] * dx + gy[q] * dy;
            a_k[q] = (P.limiter == 1) ? alpha_bj(u[q], pred, mn[q], mx[q])
                                      : alpha_venkat_eps(u[q], pred, mn[q], mx[q], veps);
        }
    }
    double al[4] = {1.0, 1.0, 1.0, 1.0};
    int kact[4] = {-1, -1, -1, -1};
#pragma unroll
    for (int q = 0; q < 4; ++q)
        for_lanes(a_k[q], [&](int j, double a) {
            if (j < nk && a < al[q]) {
                al[q] = a;
                kact[q] = j;
            }
        });
    // ---- reverse
    double fb[4] = {0.0, 0.0, 0.0, 0.0};
    if (has_edge) ld4(is_left ? W.fbL : W.fbR, e, fb);
    double ub[4], gxb[4], gyb[4], mb[4];
    double sk[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        sk[q] = gx[q] * dx + gy[q] * dy;
        ub[q] = fb[q];  // f = u + alpha (g . d)
        gxb[q] = fb[q] * al[q] * dx;
        gyb[q] = fb[q] * al[q] * dy;
        mb[q] = 0.0;
    }
    double abar[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int q = 0; q < 4; ++q)
        for_lanes(fb[q] * sk[q], [&](int, double v) { abar[q] += v; });
    // limiter through its active edge (the min over edges picks one psi)
    int mtarget[4] = {-1, -1, -1, -1};
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        if (!(has_edge && k == kact[q])) continue;
        const double pred = u[q] + gx[q] * dx + gy[q] * dy;
        const double delta = pred - u[q];
        if (fabs(delta) < DBL_EPSILON) continue;
        const bool up = delta > 0.0;
        const double a = up ? (mx[q] - u[q]) : (mn[q] - u[q]);
        double psi_a = 0.0, psi_d = 0.0;
        if (P.limiter == 1) {
            const double r = a / delta;  // active only below 1
            psi_a = 1.0 / delta;
            psi_d = -r / delta;
        } else {
            const double h = sqrt(vol / M_PI);
            const double eps = P.venkat_k * P.venkat_k * P.venkat_k * h * h * h;
            const double N = a * a + eps + 2.0 * a * delta, D = a * a + eps + 2.0 * delta * delta;
            const double raw = N / D;
            if (raw > 0.0 && raw < 1.0) {
                psi_a = (2.0 * a + 2.0 * delta) / D - N * (2.0 * a) / (D * D);
                psi_d = 2.0 * a / D - N * (4.0 * delta) / (D * D);
            }
        }
        gxb[q] += abar[q] * psi_d * dx;
        gyb[q] += abar[q] * psi_d * dy;
        ub[q] -= abar[q] * psi_a;  // a = M - u
        const int idx = up ? imx[q] : imn[q];
        if (idx < 0)
            ub[q] += abar[q] * psi_a;  // M = u (no neighbour exceeded the cell)
        else {
            mb[q] = abar[q] * psi_a;
            mtarget[q] = idx;
        }
    }
    // cell-centre velocity of the viscous work term (src/diffusion.cpp:190-198)
    double duu = 0.0, dvv = 0.0;
    if (has_edge && !(ebc == 'f' || ebc == 'w')) {
        double ta[4], tb4[4];
        ld4(vtq, va, ta);
        ld4(vtq, vb, tb4);
        const double txx = 0.5 * (ta[0] + tb4[0]), tyy = 0.5 * (ta[1] + tb4[1]), txy = 0.5 * (ta[2] + tb4[2]);
        const double a = el * nx / vol, b = el * ny / vol;
        duu = txx * a + txy * b;
        dvv = txy * a + tyy * b;
    }
    {
        const double x3 = x[3 * nc + c], rho = u[0];
        ub[0] += x3 * (duu * (-u[1] / (rho * rho)) + dvv * (-u[2] / (rho * rho)));
        ub[1] += x3 * duu / rho;
        ub[2] += x3 * dvv / rho;
    }
    // stresses and heat flux of the cell: adjoint of the vertex gather, then
    // forward-mode duals over the 12 inputs (three per lane)
    double tqb[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    {
        double lane_t[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
        const int v = A.cvtx[4 * c + k];  // this lane's vertex of the cell (-1: none)
        if (v >= 0) {
            const double w = A.cvw[4 * c + k];
            double t4[4];
            ld4(W.vtb, v, t4);
#pragma unroll
            for (int q = 0; q < 4; ++q) lane_t[q] = w * t4[q];
            lane_t[4] = w * W.vtb[4 * (long)nv + v];
        }
#pragma unroll
        for (int o = 0; o < 5; ++o)
            for_lanes(lane_t[o], [&](int, double v) { tqb[o] += v; });
    }
    for (int i = 3 * k; i < 3 * k + 3; ++i) {
        Dual du[4], dgx[4], dgy[4], out[5];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            du[q] = Dual{u[q], i == q ? 1.0 : 0.0};
            dgx[q] = Dual{gx[q], i == 4 + q ? 1.0 : 0.0};
            dgy[q] = Dual{gy[q], i == 8 + q ? 1.0 : 0.0};
        }
        cell_stress_t<Dual>(du, dgx, dgy, P.mu, P.gamma, P.kappa, out);
        const double s = tqb[0] * out[0].d + tqb[1] * out[1].d + tqb[2] * out[2].d + tqb[3] * out[3].d +
                         tqb[4] * out[4].d;
        if (i < 4)
            ub[i] += s;
        else if (i < 8)
            gxb[i - 4] += s;
        else
            gyb[i - 8] += s;
    }
    // ---- sums over the lanes, lane 0 stores
    double Ub[4] = {0.0, 0.0, 0.0, 0.0}, Gxb[4] = {0.0, 0.0, 0.0, 0.0}, Gyb[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        for_lanes(ub[q], [&](int, double v) { Ub[q] += v; });
        for_lanes(gxb[q], [&](int, double v) { Gxb[q] += v; });
        for_lanes(gyb[q], [&](int, double v) { Gyb[q] += v; });
    }
    double Mb[4];
    int Mt[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        // the active lane's contribution and the neighbour it targets
        Mb[q] = 0.0;
        for_lanes(mb[q], [&](int j, double v) {
            if (j == kact[q]) Mb[q] = v;
        });
        const int tl = from_lane_int(mtarget[q], kact[q] < 0 ? 0 : kact[q]);
        Mt[q] = (kact[q] < 0 || tl < 0) ? -1 : from_lane_int(nb, tl);
    }
    if (k != 0) return;
    st4(W.ub, c, Ub);
    st4(W.gb, 2L * c, Gxb);
    st4(W.gb, 2L * c + 1, Gyb);
    st4(W.mval, c, Mb);
#pragma unroll
    for (int q = 0; q < 4; ++q) W.mtgt[4 * c + q] = Mt[q];
}

// ------------------------------------------------- edge values (GG) ----
__global__ __launch_bounds__(256) void ka_edge_value(Topo T, Geom G, const double *__restrict__ vU, AdjWork W) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= T.ne) return;
    const int cl = T.ecl[e], cr = T.ecr[e];
    const double el = G.elen[e], enx = G.enx[e], eny = G.eny[e];
    double gxL[4], gyL[4], s[4];
    ld4(W.gb, 2L * cl, gxL);
    ld4(W.gb, 2L * cl + 1, gyL);
    const double vl = G.cvol[cl];
#pragma unroll
    for (int q = 0; q < 4; ++q) s[q] = gxL[q] * el * (-enx) / vl + gyL[q] * el * (-eny) / vl;
    if (cr >= 0) {
        double gxR[4], gyR[4];
        ld4(W.gb, 2L * cr, gxR);
        ld4(W.gb, 2L * cr + 1, gyR);
        const double vr = G.cvol[cr];
#pragma unroll
        for (int q = 0; q < 4; ++q) s[q] += gxR[q] * el * enx / vr + gyR[q] * el * eny / vr;
    }
    const char ebc = T.ebc[e];
    double o[4] = {s[0], s[1], s[2], s[3]};
    if (tagged(ebc)) {
        double xa[4], xb[4], w[4];
        ld4(vU, T.eva[e], xa);
        ld4(vU, T.evb[e], xb);
#pragma unroll
        for (int q = 0; q < 4; ++q) w[q] = 0.5 * (xa[q] + xb[q]);
        bc_state_t(ebc, w, s, o);
    }
    const double h[4] = {0.5 * o[0], 0.5 * o[1], 0.5 * o[2], 0.5 * o[3]};
    st4(W.ha, e, h);
}

// ------------------------------------------ vertex state boundary (T) ---
__global__ __launch_bounds__(256) void ka_vertex_state(Topo T, Geom G, AdjTopo A, const double *__restrict__ U,
                                                       AdjWork W) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= T.nv) return;
    const int nc = T.nc;
    double s[4] = {0.0, 0.0, 0.0, 0.0};
    for (int k = A.veptr[v]; k < A.veptr[v + 1]; ++k) {
        double h[4];
        ld4(W.ha, A.vedge[k], h);
#pragma unroll
        for (int q = 0; q < 4; ++q) s[q] += h[q];
    }
    const char bc = T.vbc[v];
    double o[4] = {s[0], s[1], s[2], s[3]};
    if (tagged(bc)) {
        // the interpolated state before the boundary state (wall: its Jacobian)
        double w[4] = {0.0, 0.0, 0.0, 0.0};
        if (bc == 'w') {
            const int b = T.sptr[v], n = T.sptr[v + 1] - b;
            const uint8_t kind = T.skind[v];
            for (int q = 0; q < 4; ++q) {
                double acc = 0.0;
                if (kind == kStencilWsum) {
                    acc = G.sw[b] * U[q * nc + T.scell[b]];
                    for (int j = 1; j < n; ++j) acc = acc + G.sw[b + j] * U[q * nc + T.scell[b + j]];
                } else if (kind == kStencilAvg2) {
                    acc = 0.5 * (U[q * nc + T.scell[b]] + U[q * nc + T.scell[b + 1]]);
                } else if (kind == kStencilCopy) {
                    acc = U[q * nc + T.scell[b]];
                }
                w[q] = acc;
            }
        }
        bc_state_t(bc, w, s, o);
    }
    st4(W.vi, v, o);
}

// --------------------------------------------------------- cell output ---
__global__ __launch_bounds__(256) void ka_cell_final(Topo T, AdjTopo A, AdjWork W, double *__restrict__ y) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= T.nc) return;
    const int nc = T.nc;
    double acc[4];
    ld4(W.ub, c, acc);
    for (int j = 0; j < 4; ++j) {
        const int v = A.cvtx[4 * c + j];
        if (v < 0) break;
        const double w = A.cvw[4 * c + j];
        double h[4];
        ld4(W.vi, v, h);
#pragma unroll
        for (int q = 0; q < 4; ++q) acc[q] += w * h[q];
    }
    for (int j = 0; j < 4; ++j) {
        const int nb = T.cnbr[4 * c + j];
        if (nb < 0) continue;
#pragma unroll
        for (int q = 0; q < 4; ++q)
            if (W.mtgt[4 * nb + q] == c) acc[q] += W.mval[4 * nb + q];
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) y[q * nc + c] = acc[q];
}

static inline unsigned nb256(long n) { return (unsigned)((n + 255) / 256); }

hipError_t launch_adjoint_exact(const Topo &T, const Geom &G, const Phys &P, const AdjTopo &A, const AdjWork &W,
                                const double *x, const double *U, const double *vU, const double *fL,
                                const double *fR, const double *vtq, double *y, hipStream_t s,
                                const ObjSeed *seed) {
    hipLaunchKernelGGL(ka_edge, dim3(nb256(T.ne)), dim3(256), 0, s, T, G, P, x, U, fL, fR, W);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    if (seed && seed->nw > 0) {
        hipLaunchKernelGGL(ka_seed_objective, dim3(nb256(seed->nw)), dim3(256), 0, s, G, seed->wall, seed->nw, fL,
                           P.gamma, seed->q_inf, seed->beta, W);
        if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    hipLaunchKernelGGL(ka_vertex_visc, dim3(nb256(T.nv)), dim3(256), 0, s, T, A, W);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    hipLaunchKernelGGL(ka_cell, dim3(nb256((long)kLanesPerCell * T.nc)), dim3(256), 0, s, T, G, P, A, x, U, vU, vtq,
                       W);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    hipLaunchKernelGGL(ka_edge_value, dim3(nb256(T.ne)), dim3(256), 0, s, T, G, vU, W);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    hipLaunchKernelGGL(ka_vertex_state, dim3(nb256(T.nv)), dim3(256), 0, s, T, G, A, U, W);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    hipLaunchKernelGGL(ka_cell_final, dim3(nb256(T.nc)), dim3(256), 0, s, T, A, W, y);
    return hipGetLastError();
}

}  // namespace o2f
