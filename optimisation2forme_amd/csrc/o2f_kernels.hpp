// Device data views and kernel launchers of the gfx950 residual / adjoint
// path (kernels in o2f_kernels.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

#include "host_mesh.hpp"

namespace o2f {

constexpr int kRedBlock = 256;  // reduction block size
constexpr int kRedGrid = 512;   // max blocks of a reduction (fixed => deterministic order)

// Flags of k_cell_flux_update.
enum : int {
    kFlagAdv = 1,       // convective HLLC sum (else start from the stored R)
    kFlagVisc = 2,      // add the viscous divergence
    kFlagStoreR = 4,    // write R
    kFlagRK = 8,        // stage update U = c0 Uold + c1 (U + dt R)
    kFlagSaveOld = 16,  // first stage: Uold = U before the update
    kFlagSpeed = 32,    // after the update: max_c(|u|+c) of the new state into the speed word
    kFlagBegin = 64,    // tile stage: the step's start (dt, the time loop's stop test) in every workgroup
    kFlagAdvance = 128, // tile stage: workgroup 0 advances the time bookkeeping by the step's dt
    kFlagBeginDt = 256,     // tile stage: the step's start from the max(|u|+c) the previous step folded
    kFlagSpeedReset = 512,  // tile stage: workgroup 0 starts the fold of the step's max(|u|+c)
    kFlagSpeedMax = 1024,   // tile stage: each tile's speed partial folded into that max (atomic)
};

// Slots of the device scalar block.
enum : int {
    kScalDt = 0,
    kScalT = 1,
    kScalRho = 2,
    kScalRhoOld = 3,
    kScalAlpha = 4,
    kScalW = 5,
    kScalRR = 6,
    kScalR0Sq = 7,
    kScalDotA = 8,
    kScalDotB = 9,
    kScalActive = 10,  // 1 while t < final_time (gated time loop), else 0
    kScalSteps = 11,   // steps taken by the gated time loop
    kScalTPrev = 12,   // t at the start of the last step
    kScalConv = 13,    // set once the density residual met the tolerance: the next step stops
    kScalRes = 14,     // last density residual norm evaluated on the device
    kScalGo = 15,      // BiCGSTAB loop: 1 run the next iteration, 0 stop, 2 restart first (host)
    kScalIter = 16,    // BiCGSTAB iterations taken in the current chunk
    kScalRhoAlt = 17,  // BiCGSTAB: rho of odd iterations of a group (kScalRho: even)
    kScalSpeedMax = 18,  // tile stages: max(1e-8, max_c(|u|+c)) of the new state, folded by the last stage
    kScalCount = 20,
};

// Epilogues of k_dot2.  kOpRho also evaluates the BiCGSTAB loop test into
// kScalGo.  (alpha, w and the later tops are formed by the vector-update
// kernels that use them.)
enum : int { kOpStore = 0, kOpRho = 1, kOpRestart = 2 };

// The loop test of a BiCGSTAB chunk, evaluated on the device.
struct BicgTest {
    double tol2;   // stop once r.r <= tol2
    double maxit;  // or once this many iterations ran
    double eps2;   // restart when |rho| < eps2 * r0_sqnorm
};

// Constants every kernel needs (host-computed with the reference's
// expression order, see Context::make_phys).
struct Phys {
    double gamma, mu, kappa, cfl;
    double fs_rho, fs_ru, fs_rv, fs_E;  // apply_inflow_outflow_bc freestream state
    double venkat_k;
    int limiter;
    // Gated time loop: when non-null, every stage kernel returns at once
    // while *gate == 0, and the dt kernel clears it once t >= final_t.
    const double *gate;
    double final_t;
};

// Stencil entries a vertex gather loads together (longer stencils loop
// over the rest); the stencil id and weight arrays carry this many entries
// of padding past their end.
constexpr int kStencilSlots = 8;

// Read-only topology (device pointers).
struct Topo {
    int nv, ne, nc;
    const int *cedge, *cnbr;  // 4 per cell, 16-byte aligned rows
    const int *cevx;          // 8 per cell: edge slot endpoints (min, max vertex)
    const uint8_t *cinfo;
    const int *eva, *evb, *ecl, *ecr;
    const char *ebc;
    const int *sptr, *scell;
    const uint8_t *skind;
    const char *vbc;
};

// Geometry (device pointers); several sets exist for shape perturbations.
struct Geom {
    const double *cvol, *ccx, *ccy;
    const double *elen, *enx, *eny, *emx, *emy;
    const double *sw;
};

// ------------------------------------------------------------------------
// Batched launches: one launch of a stage kernel covers every member of a
// batch (meshes of any sizes).  Each member's kernels read its own entry of
// a device-resident Member table; a launch walks a tile list of (member,
// block of that member's own grid) pairs, so every member sees exactly the
// per-entity arithmetic of its single-mesh launch.
// ------------------------------------------------------------------------
struct Member {
    Topo T;
    Geom G;
    Phys P;  // P.gate = scal + kScalActive; P.final_t = +inf
    double *U, *Uold, *R, *tq, *vU, *vtq, *fL, *fR, *Fe;
    double *scal;    // the member's scalar block
    double *speed;   // speed partials (one per flux-update block)
    double *rparts;  // density-residual partials (one per kResChunk cells)
    double min_h;
    double res_scale;  // rho_inf U_inf / chord
    double res_tol;    // stop once r <= res_tol (r from the last checked step)
    int nparts;        // speed partials in use (flux-update grid of the member, or its tiles)
    int nrparts;       // residual partials
    // fused tile stages: packed edge geometry and the three AoS state
    // buffers a step cycles through (stage s reads Us[s], writes Us[s+1 mod 3])
    const double *epk;
    double *Us[3];
};

constexpr int kResChunk = 2048;  // cells per residual partial (fixed: results do not depend on launch shapes)

// Launch shape of one batched kernel: threads per workgroup (64, 128 or
// 256) and grid (0: one workgroup per tile; G > 0: G persistent workgroups
// striding over the tiles).
struct TileShape {
    int block;
    int grid;
};

// Tile lists are int2 (member, block).  Stage flags as for the single-mesh
// kernels; the dt of each member is its own scal[kScalDt].
hipError_t launch_b_speed_max(const Member *M, const int2 *tiles, int ntiles, TileShape sh, hipStream_t s);
hipError_t launch_b_step_begin(const Member *M, int nmembers, hipStream_t s);
hipError_t launch_b_vertex_gather(const Member *M, const int2 *tiles, int ntiles, TileShape sh, hipStream_t s);
hipError_t launch_b_reconstruct(const Member *M, const int2 *tiles, int ntiles, TileShape sh, hipStream_t s);
hipError_t launch_b_visc_edge(const Member *M, const int2 *tiles, int ntiles, TileShape sh, hipStream_t s);
hipError_t launch_b_flux_update(const Member *M, const int2 *tiles, int ntiles, TileShape sh, double c0, double c1,
                                int flags, hipStream_t s);
// Density residual of the R stored by the last StoreR stage: partials over
// fixed chunks of kResChunk cells, then one block per member sums them in
// order, stores r in scal[kScalRes] and sets scal[kScalConv] when r <= tol.
hipError_t launch_b_residual(const Member *M, const int2 *chunk_tiles, int nchunks, int nmembers, hipStream_t s);
// The same density-residual check for one mesh: the R_rho stored by the
// last StoreR stage, partials over kResChunk cells (rparts, ceil(nc /
// kResChunk) of them), r into scal[kScalRes], scal[kScalConv] = 1 when
// r <= tol; gated on *gate when given.
hipError_t launch_residual_check(const double *R, int nc, double *rparts, double res_scale, double tol,
                                 double *scal, const double *gate, hipStream_t s);
// flags[i] = member i still active
hipError_t launch_b_collect(const Member *M, int nmembers, double *flags, hipStream_t s);

// ------------------------------------------------------------------------
// Exact adjoint (o2f_adjoint.hip): y = (dR/dU)^T x at the state whose
// forward vertex values vU, face states fL / fR and vertex stresses vtq the
// last residual evaluation left.
// ------------------------------------------------------------------------
struct AdjTopo {
    const int *veptr, *vedge;  // vertex -> incident edges (CSR, ascending edge id)
    const int *cvtx;           // 4 per cell, -1 pad
    const double *cvw;         // 4 per cell: the cell's weight in the stencil of its j-th vertex
};
struct AdjWork {
    double *fbL, *fbR;  // adjoint face states (AoS 4 per edge)
    double *th;         // half the adjoint edge stresses (AoS 4 + trailing block)
    double *vtb;        // adjoint vertex stresses (AoS 4 + trailing block)
    double *ub;         // adjoint cell state, direct part (AoS 4)
    double *gb;         // adjoint cell gradients (AoS 8: x then y)
    double *mval;       // limiter min / max adjoint per variable (AoS 4) ...
    int *mtgt;          // ... and the neighbour cell it belongs to (-1: none)
    double *ha;         // half the adjoint edge values (AoS 4)
    double *vi;         // adjoint interpolated vertex states (AoS 4)
};
// seed: when given (with x = 0), the objective's face-state derivative is
// added after the edge step, and y = (dJ/dU)^T exactly.
struct ObjSeed {
    const int *wall;
    int nw;
    double q_inf, beta;
};
hipError_t launch_adjoint_exact(const Topo &T, const Geom &G, const Phys &P, const AdjTopo &A, const AdjWork &W,
                                const double *x, const double *U, const double *vU, const double *fL,
                                const double *fR, const double *vtq, double *y, hipStream_t s,
                                const ObjSeed *seed = nullptr);

// ------------------------------------------------------------------------
// Tangent (o2f_tangent.hip): the residual and its directional derivative
// along a geometry tangent dG (same layout as the geometry) and a state
// tangent dU.  Values go to the forward arrays (bit-identical to the
// forward kernels), derivatives to the TanWork arrays of the same layouts.
// ------------------------------------------------------------------------
struct TanWork {
    double *dvU, *dvtq, *dfL, *dfR, *dFe, *dtq, *dR;
};
// dU (4 nc, SoA) and the arrays of dG may be null (zero tangent).
hipError_t launch_residual_tangent(const Topo &T, const Geom &G, const Geom &dG, const Phys &P, const double *U,
                                   const double *dU, double *vU, double *fL, double *fR, double *tq, double *vtq,
                                   double *Fe, double *R, const TanWork &W, hipStream_t s);
// Writable pointers of a geometry set, in Geom's layout.
struct GeomW {
    double *cvol, *ccx, *ccy, *elen, *enx, *eny, *emx, *emy, *sw;
};
// Listed entries of a geometry set (GeomSparse: cells with 3 values each,
// edges with 5, stencil weights with 1) written over its arrays.
hipError_t launch_scatter_geometry(const GeomW &g, int ncell, const int *cell, const double *cval, int nedge,
                                   const int *edge, const double *eval, int nsw, const int *swi, const double *swv,
                                   hipStream_t s);
// out = L, D, dL, dD over the wall edges (skip_x: computeDirectObjectiveTerm's x > 1.3 rule)
hipError_t launch_objective_tangent(const Geom &G, const Geom &dG, const int *wall, int nw, const double *fL,
                                    const double *dfL, double gamma, int skip_x, double *out, hipStream_t s);

hipError_t launch_vertex_state(const Topo &T, const Geom &G, const double *U, double *vU, const Phys &P,
                               hipStream_t s);
hipError_t launch_vertex_visc(const Topo &T, const Geom &G, const double *tq, double *vtq, const Phys &P,
                              hipStream_t s);
hipError_t launch_reconstruct(const Topo &T, const Geom &G, const double *U, const double *vU, double *fL,
                              double *fR, double *tq, double *Gout, const Phys &P, hipStream_t s);
hipError_t launch_stress(const Topo &T, const double *U, const double *Gs, double *tq, const Phys &P,
                         hipStream_t s);
// Cell sum of the edge fluxes, viscous divergence, stage update.
hipError_t launch_flux_update(const Topo &T, const Geom &G, double *U, double *Uold, const double *Fe,
                              const double *vtq, double *R, const double *dtp, double c0, double c1, int flags,
                              const Phys &P, hipStream_t s, double *speed = nullptr);
// Vertex gather of the viscous quantities (do_visc) and HLLC once per edge.
hipError_t launch_visc_and_flux(const Topo &T, const Geom &G, const double *tq, double *vtq, const double *fL,
                                const double *fR, double *Fe, const Phys &P, int do_visc, hipStream_t s);
// Per-block max_c(|u|+c) partials (speed_parts(nc) of them).
int speed_parts(int nc);
hipError_t launch_speed_max(const double *U, int nc, const Phys &P, double *parts, hipStream_t s);
// dt from nparts speed partials, time advance, gate (one block).  The
// partials come from launch_speed_max or the last stage of a step
// (speed_parts(nc) of them), or from the last tile stage (one per tile).
hipError_t launch_step_begin(const double *parts, int nparts, const Phys &P, double min_h, double *scal,
                             double *hist, int advance_t, hipStream_t s);
// The step begin and the first stage's vertex state gather in one launch.
hipError_t launch_step_gather(const Topo &T, const Geom &G, const double *U, double *vU, const Phys &P,
                              const double *parts, int nparts, double min_h, double *scal, double *hist,
                              int advance_t, hipStream_t s);
hipError_t launch_dot2(const double *a, const double *b, const double *c, const double *d, long n,
                       double *partials, double *scal, int op, hipStream_t s, const double *gate = nullptr,
                       const BicgTest &test = BicgTest{});
// The block partials of a reduction only (2 per block, red_grid(n) blocks),
// for a consumer kernel that combines them itself.
hipError_t launch_dot2_partial(const double *a, const double *b, const double *c, const double *d, long n,
                               double *partials, const double *gate, hipStream_t s);
// The iteration's last update (x += alpha px + w sx; r = s - w t) fused with
// w = t.s / t.t (or 0) from the (t.t, t.s) partials wparts before it and
// the (r0.r, r.r) partials after it (identical to launch_dot2(r0, r, r, r,
// ...)'s); counts the iteration in scal[kScalIter].  Gated on scal[kScalGo].
// wparts and partials must not overlap.
hipError_t launch_bicg_xr_dot(long n, double *x, double *r, const double *px, const double *sx, const double *sv,
                              const double *t, const double *r0, const double *wparts, double *partials,
                              double *scal, hipStream_t s);
hipError_t launch_bicg_p(long n, const double *r, double *p, const double *v, const double *scal, hipStream_t s);
// The top of a BiCGSTAB iteration and its direction update in one launch:
// rho = r0.r and r.r from the partials of launch_bicg_xr_dot, the
// loop test into scal[kScalGo] (as kOpRho, with the iteration already
// counted), then, if the loop goes on, p = r + beta (p - w v).  Iteration
// `odd` of a group reads rho_old from kScalRho (odd: kScalRhoAlt) and
// writes the new rho to the other slot, so no block reads a slot that
// block 0 writes; kScalRhoOld gets rho_old for a host restart.
hipError_t launch_bicg_p_top(long n, const double *r, double *p, const double *v, double *scal,
                             const double *parts, int odd, const BicgTest &test, hipStream_t s);
// s = r - alpha v with alpha = rho / (r0.v) formed from the r0.v partials
// aparts (stored to scal too); rho from kScalRho, or kScalRhoAlt in the odd
// iterations of a group (see launch_bicg_p_top).
hipError_t launch_bicg_s(long n, const double *r, const double *v, double *sv, double *scal, const double *aparts,
                         int odd, hipStream_t s);
// Block-Jacobi preconditioner: probing seeds / extraction by colour,
// P_c = (D_c^T)^{-1} (16 per cell, row-major), out_c = P_c in_c.
hipError_t launch_probe_seed(int nc, const int *color, int k, int q, double *x, hipStream_t s);
hipError_t launch_probe_take(int nc, const int *color, int k, int q, const double *y, double *D, hipStream_t s);
hipError_t launch_invert_blocks(int nc, const double *D, double *P, hipStream_t s);
hipError_t launch_apply_blocks(int nc, const double *P, const double *in, double *out, hipStream_t s);
hipError_t launch_sub(long n, const double *a, const double *b, double *out, hipStream_t s);
hipError_t launch_fill(long n, double *out, double v, hipStream_t s);
hipError_t launch_fd_diff(long n, const double *Rp, const double *Rm, double eps, double *out, hipStream_t s);
hipError_t launch_wall_forces(const Geom &G, const int *wall, int nw, const double *fL, int ne, double gamma,
                              int skip_x, double *out, hipStream_t s);
hipError_t launch_adjoint_rhs(const Topo &T, const Geom &G, const int *wall_by_cell, const int *grp, int ngrp,
                              const double *fL, double gamma, double q_inf, double beta, double *b, hipStream_t s);
hipError_t launch_jt_blocks(const Topo &T, const Geom &G, const double *fL, const double *fR, double eps,
                            const Phys &P, double *blk, hipStream_t s);
hipError_t launch_jt_apply(const Topo &T, const Geom &G, const double *blk, const double *x, double *y,
                           hipStream_t s);
hipError_t launch_hllc_batch(const double *qL, const double *qR, const double *nrm, double gamma, double *F,
                             int n, hipStream_t s);

}  // namespace o2f
