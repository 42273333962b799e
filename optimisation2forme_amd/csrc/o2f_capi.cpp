// C ABI of the gfx950 residual / SSP-RK3 / adjoint path (include/o2f.h).
//
// One context owns one mesh: its topology and geometry (uploaded once), the
// SoA FP64 state, the per-stage intermediates, the adjoint work vectors and
// one HIP stream.  Everything between o2f_set_state and o2f_get_state stays
// on the device; host round trips happen only where the API hands results
// back.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <functional>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

#include "host_mesh.hpp"
#include "o2f.h"
#include "o2f_kernels.hpp"
#include "o2f_ctx.hpp"

using namespace o2f;

namespace o2f {
// the message of the last failed call without a context (o2f_global_error)
thread_local std::string g_error;
}  // namespace o2f

namespace {

// Freestream constants of apply_inflow_outflow_bc (src/boundary.cpp:10-17)
// and the other per-run scalars, with the reference's expression order.
Phys make_phys(const o2f_params &p, double *q_inf) {
    Phys P{};
    P.gamma = p.gamma;
    P.mu = p.viscosity;
    P.kappa = p.viscosity * p.cp / p.Pr;
    if (p.heat_flux == O2F_HEAT_FLUX_FOURIER) P.kappa = P.kappa / (p.cp - p.cp / p.gamma);
    P.cfl = p.cfl_number;
    const double rho0 = p.rho_ref, p0 = p.p_ref, M = p.inflow_velocity;
    const double a0 = std::sqrt(p.gamma * p0 / rho0);
    const double u0 = M * a0, v0 = 0.0;
    const double kinetic = 0.5 * rho0 * (u0 * u0 + v0 * v0);
    P.fs_rho = rho0;
    P.fs_ru = rho0 * u0;
    P.fs_rv = rho0 * v0;
    P.fs_E = p.p_ref / (p.gamma - 1.0) + kinetic;
    P.venkat_k = p.venkat_k;
    P.limiter = p.limiter;
    const double a_inf = std::sqrt(p.gamma * p.p_ref / p.rho_ref);
    const double U_inf = M * a_inf;
    *q_inf = 0.5 * (p.rho_ref * U_inf * U_inf);
    return P;
}

bool any_tagged(const HostMesh &hm) {
    for (char b : hm.vbc)
        if (b == 'i' || b == 'o' || b == 'f' || b == 'w') return true;
    for (char b : hm.ebc)
        if (b == 'i' || b == 'o' || b == 'f' || b == 'w') return true;
    return false;
}

// apply_boundary_conditions throws for any boundary_conditions other than
// "inflow_outflow_x" on the first tagged vertex (src/boundary.cpp:84-92).
int check_bc(o2f_ctx *c) {
    if (!c->prm.bc_inflow_outflow_x && any_tagged(c->hm))
        return c->fail(O2F_ERR_BC, "Unknown boundary condition: boundary_conditions is not inflow_outflow_x");
    return O2F_OK;
}

const long kNone = -1;

// One residual evaluation (optionally fused with a stage update) on
// geometry set g.  Four kernels: vertex gather, cell reconstruction,
// viscous vertex gather, flux + update.
// begin_step: the step's dt reduction rides in the vertex gather launch.
// Uin: the state to evaluate instead of c->U (flags without kFlagRK).
int run_stage(o2f_ctx *c, int g, int flags, const double *dtp, double c0, double c1, bool store_grad,
              bool begin_step = false, double *hist = nullptr, double *Uin = nullptr) {
    const Topo T = c->topo();
    const Geom G = c->geo[g].view();
    hipStream_t s = c->stream;
    double *U = Uin ? Uin : c->U;
    if (begin_step)
        HIPCHK(c, launch_step_gather(T, G, U, c->vU, c->phys, c->speed, c->speed_nparts, c->hm.min_h, c->scal, hist,
                                     1, s));
    else
        HIPCHK(c, launch_vertex_state(T, G, U, c->vU, c->phys, s));
    HIPCHK(c, launch_reconstruct(T, G, U, c->vU, c->fL, c->fR, c->tq, store_grad ? c->Gr : nullptr, c->phys, s));
    HIPCHK(c, launch_visc_and_flux(T, G, c->tq, c->vtq, c->fL, c->fR, c->Fe, c->phys, 1, s));
    HIPCHK(c, launch_flux_update(T, G, U, c->Uold, c->Fe, c->vtq, c->R, dtp, c0, c1, flags | kFlagAdv | kFlagVisc,
                                 c->phys, s, c->speed));
    if (flags & kFlagSpeed) c->speed_nparts = speed_parts(c->hm.nc);
    ++c->faces_version;
    if (store_grad) c->have_grad = true;
    return O2F_OK;
}

int residual_on(o2f_ctx *c, int g) {
    int st = check_bc(c);
    if (st) return st;
    return run_stage(c, g, kFlagStoreR, nullptr, 0.0, 0.0, true);
}

int copy_out(o2f_ctx *c, double *host, const double *dev, size_t n) {
    if (!host) return O2F_OK;
    HIPCHK(c, hipMemcpyAsync(host, dev, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return O2F_OK;
}

// One SSP-RK3 step (src/simulation.cpp:130-142).  The dt in scal[kScalDt]
// is read by the stage kernels on the device.
int rk3_stages(o2f_ctx *c, bool last_outputs, bool next_speed, bool begin_step = false, double *hist = nullptr) {
    const double *dtp = c->scal + kScalDt;
    int st;
    if ((st = run_stage(c, 0, kFlagRK | kFlagSaveOld, dtp, 0.0, 1.0, false, begin_step, hist))) return st;
    if ((st = run_stage(c, 0, kFlagRK, dtp, 3.0 / 4.0, 1.0 / 4.0, false))) return st;
    return run_stage(c, 0, kFlagRK | (last_outputs ? kFlagStoreR : 0) | (next_speed ? kFlagSpeed : 0), dtp,
                     1.0 / 3.0, 2.0 / 3.0, last_outputs);
}

void release_tiles(o2f_ctx *c) {
    for (void *p : {c->tile_mem, (void *)c->Ua, (void *)c->Ub, (void *)c->Uc, (void *)c->epk, (void *)c->rparts})
        if (p) (void)hipFree(p);
    c->tile_mem = nullptr;
    c->Ua = c->Ub = c->Uc = c->epk = c->rparts = nullptr;
    c->tile_aos = false;
    c->tv = TileView{};
    c->tiles_ready = false;
}

// The tile tables of the current mesh and stencils on the device (built on
// first use).  Returns false when the stages stay the four-kernel ones.
// Automatic size: the fused kernel holds tiles resident on every CU at once
// (blocks_per_cu of them, set by its VGPRs and its LDS), so the tiles are
// made just large enough that one round covers the mesh: the rim each tile
// recomputes shrinks as the tiles grow, and no CU waits on a second round.
bool tiles_on(o2f_ctx *c) {
    if (c->tile_cells == 0 || c->tiles_unusable) return false;
    if (c->tiles_ready) return true;
    release_tiles(c);
    auto unusable = [&](const std::string &why) {
        c->tiles_unusable = true;
        c->tiles_why = why;
        return false;
    };
    int dev = 0, ncu = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || ncu <= 0)
        return unusable("no device attributes");
    const int nc = c->hm.nc;
    TileSet ts;
    std::string err;
    int bpc = 0, vgprs = 0, qpc = 0, target = c->tile_cells;
    const int nk_max = *std::max_element(c->hm.cnv.begin(), c->hm.cnv.end());
    auto build = [&](int t) -> bool {
        if (!build_tiles(c->hm, t, ts, err)) return false;
        const hipError_t e = tile_stage_attributes(ts.lds_bytes(), nk_max, &bpc, &vgprs, &qpc);
        if (e != hipSuccess) err = std::string("tile kernel attributes: ") + hipGetErrorString(e);
        return e == hipSuccess && bpc > 0;
    };
    // the smallest tiles (of at least kMinCells; kMinQuadCells for one tile
    // per CU, which runs the four-lane kernels) of which every CU holds at
    // most m at once
    constexpr int kMinCells = 96, kMinQuadCells = 32, kSmallCells = 128;
    auto fits = [&](int m) -> bool {
        target = std::max(m == 1 ? kMinQuadCells : kMinCells, (int)std::ceil(nc / (0.985 * m * ncu)));
        bool ok = build(target);
        for (int it = 0; ok && it < 12 && (int)ts.head.size() > m * ncu; ++it)
            ok = build(target = (int)std::ceil(target * 1.02));
        return ok && bpc >= m && (int)ts.head.size() <= m * ncu;
    };
    if (target > 0) {
        if (!build(target)) return unusable(err.empty() ? "tile LDS too large" : err);
    } else {
        if (!build(192)) return unusable(err.empty() ? "tile LDS too large" : err);
        const int resident = bpc;
        // A small mesh: the fewest tiles per CU whose tiles stay within
        // kSmallCells (a stage then takes about one tile's time, and larger
        // tiles recompute less rim; measured on 10k triangles and the
        // reference's 25k quads); with one per CU, every CU gets one (the
        // four-lane kernels).  Otherwise: as many resident tiles per CU as
        // the kernel allows, which hides the phases' load latency.
        bool done = false;
        for (int m = 1; m <= resident && !done; ++m)
            if (nc / (0.985 * m * ncu) <= kSmallCells) done = fits(m) && target <= kSmallCells;
        for (int m = resident; m >= 1 && !done; --m) done = fits(m);
        if (!done) return unusable("no tile size fits");
    }
    // one allocation: heads, cells, owned-cell slots, vertices, edges, stencils
    size_t off = 0;
    auto slot = [&](size_t bytes) {
        const size_t o = off;
        off += (bytes + 255) & ~size_t(255);
        return o;
    };
    const size_t o_head = slot(ts.head.size() * sizeof(TileHead)), o_cell = slot(ts.cell.size() * sizeof(TileCell)),
                 o_own = slot(ts.own.size() * sizeof(TileOwn)), o_vg = slot(ts.vgf.size() * sizeof(TileGather)),
                 o_edge = slot(ts.en.size() * sizeof(double)), o_vo = slot(ts.vof.size() * sizeof(TileStress)),
                 o_ls = slot(ts.lsten.size() * sizeof(uint16_t) + 16);
    char *m = nullptr;
    hipError_t e = hipMalloc(reinterpret_cast<void **>(&m), off);
    if (e == hipSuccess) c->tile_mem = m;
    auto put = [&](size_t o, const void *src, size_t bytes) {
        if (e == hipSuccess && bytes) e = hipMemcpy(m + o, src, bytes, hipMemcpyHostToDevice);
    };
    put(o_head, ts.head.data(), ts.head.size() * sizeof(TileHead));
    put(o_cell, ts.cell.data(), ts.cell.size() * sizeof(TileCell));
    put(o_own, ts.own.data(), ts.own.size() * sizeof(TileOwn));
    put(o_vg, ts.vgf.data(), ts.vgf.size() * sizeof(TileGather));
    put(o_edge, ts.en.data(), ts.en.size() * sizeof(double));
    put(o_vo, ts.vof.data(), ts.vof.size() * sizeof(TileStress));
    put(o_ls, ts.lsten.data(), ts.lsten.size() * sizeof(uint16_t));
    if (e == hipSuccess) e = dalloc(c->epk, 8 * (size_t)c->hm.ne);
    if (e == hipSuccess) e = launch_pack_edges(c->geo[0].view(), c->ebc, c->hm.ne, c->epk, c->stream);
    if (e == hipSuccess) e = dalloc(c->Ua, 4 * (size_t)nc);
    if (e == hipSuccess) e = dalloc(c->Ub, 4 * (size_t)nc);
    if (e == hipSuccess) e = dalloc(c->Uc, 4 * (size_t)nc);
    // one speed partial per tile: small tiles outnumber the four-lane blocks
    const int nparts = std::max((int)ts.head.size(), speed_parts(nc));
    if (e == hipSuccess && nparts > speed_parts(nc)) {
        (void)hipFree(c->speed);
        c->speed = nullptr;
        c->speed_valid = false;
        e = dalloc(c->speed, nparts);
    }
    if (e != hipSuccess) {
        release_tiles(c);
        return unusable(std::string("tile tables: ") + hipGetErrorString(e));
    }
    TileView &v = c->tv;
    v.head = reinterpret_cast<const TileHead *>(m + o_head);
    v.cell = reinterpret_cast<const TileCell *>(m + o_cell);
    v.own = reinterpret_cast<const TileOwn *>(m + o_own);
    v.vg = reinterpret_cast<const TileGather *>(m + o_vg);
    v.en = reinterpret_cast<const double2 *>(m + o_edge);
    v.vo = reinterpret_cast<const TileStress *>(m + o_vo);
    v.lsten = reinterpret_cast<const uint16_t *>(m + o_ls);
    v.epk = c->epk;
    v.ntiles = (int)ts.head.size();
    v.nk_max = nk_max;
    v.max_vo = ts.max_vo;
    v.max_gd = ts.max_gd;
    v.lds_bytes = ts.lds_layout(&v.lds_tq, &v.lds_fs);
    // tiles that run one per CU (a small mesh): the four-lane kernels, whose
    // sixteen waves share one tile's cells
    v.quad = (qpc >= 1 && v.ntiles <= ncu) ? 1 : 0;
    c->tile_target = ts.target;
    c->tile_bpc = bpc;
    c->tile_vgprs = vgprs;
    c->tiles_ready = true;
    return true;
}

// The state of a tiled run lives in Ua (AoS) from the first tiled step on;
// tiles_leave writes it back to c->U (SoA) before the entry point returns.
int tiles_enter(o2f_ctx *c) {
    if (!c->tile_aos) {
        HIPCHK(c, launch_state_to_aos(c->U, c->hm.nc, c->Ua, c->stream));
        c->tile_aos = true;
    }
    return O2F_OK;
}
int tiles_leave(o2f_ctx *c) {
    if (c->tile_aos) {
        c->tile_aos = false;
        HIPCHK(c, launch_state_to_soa(c->Ua, c->hm.nc, c->U, c->stream));
    }
    return O2F_OK;
}

// One step on the tiles: dt, then three fused stages, Ua -> Ub -> Uc -> Ua
// (each stage reads one buffer and writes another, so no tile overwrites a
// state its neighbours still read).  last_outputs: the second stage writes
// the C ABI's c->U, and the third stage runs on the four kernels, in place,
// leaving the intermediates the adjoint and the force coefficients read
// (and the state in c->U).
// check: the first stage also stores R_rho of the step's start state and
// the density-residual check follows it (o2f_run_to_residual).
// chain (256-thread tiles; a loop's steps after its first): kChainUse starts
// the step from the speed max the previous step's last stage folded, in
// place of the one-block step start; kChainForm has this step's last stage
// fold it (the same max of the same partials).
enum : int { kChainUse = 1, kChainForm = 2 };
int tiled_step(o2f_ctx *c, bool last_outputs, double *hist, bool check, double res_tol, int chain) {
    hipStream_t s = c->stream;
    const Topo T = c->topo();
    const Geom G = c->geo[0].view();
    if (!c->speed_valid) {
        HIPCHK(c, launch_speed_max(c->U, c->hm.nc, c->phys, c->speed, s));
        c->speed_nparts = speed_parts(c->hm.nc);
    }
    int st = tiles_enter(c);
    if (st) return st;
    const TileView &A = c->tv;
    // the four-lane stages start the step themselves (no one-block launch)
    const StepBegin B{c->speed, c->speed_nparts, c->hm.min_h, c->scal, hist};
    const bool use = !A.quad && (chain & kChainUse), form = !A.quad && !last_outputs && (chain & kChainForm);
    const int begin = A.quad ? kFlagBegin : (use ? kFlagBeginDt : 0);
    const int advance = (A.quad || use) ? kFlagAdvance : 0;
    if (!A.quad && !use)
        HIPCHK(c, launch_step_begin(c->speed, c->speed_nparts, c->phys, c->hm.min_h, c->scal, hist, 1, s));
    const double *dtp = c->scal + kScalDt;
    HIPCHK(c, launch_tile_stage(T, G, c->phys, A, c->Ua, c->Ub, true, c->Uold, c->R, dtp, 0.0, 1.0,
                                kFlagRK | kFlagSaveOld | (check ? kFlagStoreR : 0) | begin, nullptr, s, B));
    if (check) {
        const double a_inf = std::sqrt(c->prm.gamma * c->prm.p_ref / c->prm.rho_ref);
        const double res_scale = c->prm.rho_ref * (c->prm.inflow_velocity * a_inf) / 1.0;
        HIPCHK(c, launch_residual_check(c->R, c->hm.nc, c->rparts, res_scale, res_tol, c->scal, c->phys.gate, s));
    }
    if (last_outputs) {
        HIPCHK(c, launch_tile_stage(T, G, c->phys, A, c->Ub, c->U, false, c->Uold, c->R, dtp, 3.0 / 4.0, 1.0 / 4.0,
                                    kFlagRK | advance, nullptr, s, B));
        c->tile_aos = false;
        return run_stage(c, 0, kFlagRK | kFlagStoreR | kFlagSpeed, dtp, 1.0 / 3.0, 2.0 / 3.0, true);
    }
    HIPCHK(c, launch_tile_stage(T, G, c->phys, A, c->Ub, c->Uc, true, c->Uold, c->R, dtp, 3.0 / 4.0, 1.0 / 4.0,
                                kFlagRK | advance | (form ? kFlagSpeedReset : 0), nullptr, s, B));
    HIPCHK(c, launch_tile_stage(T, G, c->phys, A, c->Uc, c->Ua, true, c->Uold, c->R, dtp, 1.0 / 3.0, 2.0 / 3.0,
                                kFlagRK | kFlagSpeed | (form ? kFlagSpeedMax : 0), c->speed, s, B));
    c->speed_nparts = A.ntiles;
    return O2F_OK;
}

// One step of the time loop body: dt from the speed word, the three stages,
// and (fused into the last stage) the speed of the new state for the next
// step's dt.
int one_step(o2f_ctx *c, bool last_outputs, double *hist, int chain = 0) {
    hipStream_t s = c->stream;
    int st;
    if (tiles_on(c)) {
        st = tiled_step(c, last_outputs, hist, false, -1.0, chain);
    } else {
        if (!c->speed_valid) {
            HIPCHK(c, launch_speed_max(c->U, c->hm.nc, c->phys, c->speed, s));
            c->speed_nparts = speed_parts(c->hm.nc);
        }
        st = rk3_stages(c, last_outputs, true, true, hist);
    }
    c->speed_valid = (st == O2F_OK);
    return st;
}

std::string fmt_double(double v) {
    char b[64];
    std::snprintf(b, sizeof b, "%g", v);
    return b;
}

}  // namespace

int o2f_ctx_check_bc(o2f_ctx *c) { return check_bc(c); }
int o2f_ctx_form_outputs(o2f_ctx *c, double *Uin) {
    return run_stage(c, 0, 0, nullptr, 0.0, 0.0, false, false, nullptr, Uin);
}

extern "C" {

void o2f_params_default(o2f_params *p) {
    p->gamma = 1.4;
    p->viscosity = 1e-6;
    p->Pr = 0.71;
    p->cp = 1005.0;
    p->cfl_number = 0.1;
    p->rho_ref = 1.225;
    p->p_ref = 101325.0;
    p->inflow_velocity = 1.0;
    p->venkat_k = 1e-2;
    p->limiter = O2F_LIMITER_VENKAT;
    p->heat_flux = O2F_HEAT_FLUX_REF;
    p->interp_mode = O2F_INTERP_REF;
    p->bc_inflow_outflow_x = 1;
}

const char *o2f_global_error(void) { return g_error.c_str(); }
const char *o2f_last_error(const o2f_ctx *ctx) { return ctx ? ctx->err.c_str() : g_error.c_str(); }

int o2f_msh2_read(const char *path, o2f_mesh_desc *out) {
    if (!path || !out) {
        g_error = "o2f_msh2_read: null argument";
        return O2F_ERR_ARG;
    }
    std::vector<double> xy;
    std::vector<int> cnv, cvtx, bvtx;
    std::vector<char> btag;
    std::string err;
    if (!read_msh2(path, xy, cnv, cvtx, bvtx, btag, err)) {
        g_error = err;
        return O2F_ERR_IO;
    }
    auto *pxy = new double[xy.size()];
    auto *pcnv = new int[cnv.size() + 1];
    auto *pcv = new int[cvtx.size() + 1];
    auto *pbv = new int[bvtx.size() + 1];
    auto *pbt = new char[btag.size() + 1];
    std::copy(xy.begin(), xy.end(), pxy);
    std::copy(cnv.begin(), cnv.end(), pcnv);
    std::copy(cvtx.begin(), cvtx.end(), pcv);
    std::copy(bvtx.begin(), bvtx.end(), pbv);
    std::copy(btag.begin(), btag.end(), pbt);
    out->n_vertices = static_cast<int>(xy.size() / 2);
    out->xy = pxy;
    out->n_cells = static_cast<int>(cnv.size());
    out->cell_nv = pcnv;
    out->cell_vtx = pcv;
    out->n_bnd = static_cast<int>(btag.size());
    out->bnd_vtx = pbv;
    out->bnd_tag = pbt;
    return O2F_OK;
}

void o2f_msh2_free(o2f_mesh_desc *d) {
    if (!d) return;
    delete[] d->xy;
    delete[] d->cell_nv;
    delete[] d->cell_vtx;
    delete[] d->bnd_vtx;
    delete[] d->bnd_tag;
    std::memset(d, 0, sizeof(*d));
}

int o2f_mesh_topology(const o2f_mesh_desc *mesh, int *n_edges, int *edge_vtx, int *edge_cells, char *edge_bc,
                      double *edge_geom, double *cell_geom, int *cell_edges) {
    if (!mesh) {
        g_error = "o2f_mesh_topology: null mesh";
        return O2F_ERR_ARG;
    }
    HostMesh hm;
    std::string err;
    if (!build_host_mesh(*mesh, O2F_INTERP_REF, hm, err)) {
        g_error = err;
        return O2F_ERR_ARG;
    }
    const int ne = hm.ne, nc = hm.nc;
    if (n_edges) *n_edges = ne;
    for (int e = 0; e < ne; ++e) {
        if (edge_vtx) {
            edge_vtx[2 * e] = hm.eva[e];
            edge_vtx[2 * e + 1] = hm.evb[e];
        }
        if (edge_cells) {
            edge_cells[2 * e] = hm.ecl[e];
            edge_cells[2 * e + 1] = hm.ecr[e];
        }
        if (edge_bc) edge_bc[e] = hm.ebc[e];
        if (edge_geom) {
            edge_geom[e] = hm.elen[e];
            edge_geom[ne + e] = hm.enx[e];
            edge_geom[2 * ne + e] = hm.eny[e];
            edge_geom[3 * ne + e] = hm.emx[e];
            edge_geom[4 * ne + e] = hm.emy[e];
        }
    }
    for (int c = 0; c < nc; ++c) {
        if (cell_geom) {
            cell_geom[c] = hm.cvol[c];
            cell_geom[nc + c] = hm.ccx[c];
            cell_geom[2 * nc + c] = hm.ccy[c];
        }
        if (cell_edges)
            for (int k = 0; k < 4; ++k) cell_edges[4 * c + k] = hm.cedge[4 * c + k];
    }
    return O2F_OK;
}

int o2f_ctx_create(const o2f_mesh_desc *mesh, const o2f_params *params, o2f_ctx **out) {
    if (!mesh || !params || !out) {
        g_error = "o2f_ctx_create: null argument";
        return O2F_ERR_ARG;
    }
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
        g_error = "o2f_ctx_create: no HIP device (this library runs only on a GPU)";
        return O2F_ERR_HIP;
    }
    std::unique_ptr<o2f_ctx> c(new o2f_ctx());
    c->prm = *params;
    std::string err;
    if (!build_host_mesh(*mesh, params->interp_mode, c->hm, err)) {
        g_error = err;
        return O2F_ERR_ARG;
    }
    c->phys = make_phys(c->prm, &c->q_inf);
    const HostMesh &hm = c->hm;
    const size_t nc = hm.nc, ne = hm.ne, nv = hm.nv;
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    std::vector<int> sptr(hm.vptr.begin(), hm.vptr.end());
    if (!e) e = upload(c->cedge, hm.cedge);
    if (!e) e = upload(c->cnbr, hm.cnbr);
    if (!e) e = upload(c->cevx, hm.cevx);
    if (!e) e = upload(c->cinfo, hm.cinfo);
    if (!e) e = upload(c->eva, hm.eva);
    if (!e) e = upload(c->evb, hm.evb);
    if (!e) e = upload(c->ecl, hm.ecl);
    if (!e) e = upload(c->ecr, hm.ecr);
    if (!e) e = upload(c->ebc, hm.ebc);
    if (!e) e = upload(c->vbc, hm.vbc);
    if (!e) e = upload(c->sptr, sptr);
    {
        // kStencilSlots entries of padding: the vertex gathers read the
        // first slots of a stencil in 16-byte loads
        std::vector<int> padded(hm.scell);
        padded.resize(hm.scell.size() + kStencilSlots, 0);
        if (!e) e = upload(c->scell, padded);
    }
    if (!e) e = upload(c->skind, hm.skind);
    if (!e) e = upload(c->wall, hm.wall_edges);
    {
        const int nw = (int)hm.wall_edges.size();
        std::vector<int> order(nw);
        for (int k = 0; k < nw; ++k) order[k] = k;
        std::stable_sort(order.begin(), order.end(),
                         [&](int a, int b) { return hm.ecl[hm.wall_edges[a]] < hm.ecl[hm.wall_edges[b]]; });
        std::vector<int> by_cell(nw), grp;
        for (int j = 0; j < nw; ++j) {
            by_cell[j] = hm.wall_edges[order[j]];
            if (j == 0 || hm.ecl[by_cell[j]] != hm.ecl[by_cell[j - 1]]) grp.push_back(j);
        }
        grp.push_back(nw);
        c->wall_ngrp = (int)grp.size() - 1;
        if (!e) e = upload(c->wall_by_cell, by_cell);
        if (!e) e = upload(c->wall_grp, grp);
    }
    if (!e) e = c->geo[0].alloc(hm);
    if (!e) e = c->geo[0].fill(hm, c->stream);
    if (!e) e = dalloc(c->U, 4 * nc);
    if (!e) e = dalloc(c->Uold, 4 * nc);
    if (!e) e = dalloc(c->R, 4 * nc);
    if (!e) e = dalloc(c->Gr, 8 * nc);
    if (!e) e = dalloc(c->tq, 5 * nc);
    if (!e) e = dalloc(c->vU, 4 * nv);
    if (!e) e = dalloc(c->vtq, 5 * nv);
    if (!e) e = dalloc(c->fL, 4 * ne);
    if (!e) e = dalloc(c->fR, 4 * ne);
    if (!e) e = dalloc(c->Fe, 4 * ne);
    if (!e) e = dalloc(c->scal, kScalCount);
    if (!e) e = dalloc(c->partials, 4 * kRedGrid);  // two sets of block partials
    if (!e) e = dalloc(c->wall_out, 2);
    if (!e) e = dalloc(c->speed, speed_parts(hm.nc));
    if (!e) e = hipMemset(c->scal, 0, kScalCount * sizeof(double));
    if (!e) e = hipMemset(c->U, 0, 4 * nc * sizeof(double));
    if (!e) e = hipMemset(c->fL, 0, 4 * ne * sizeof(double));
    if (!e) e = hipMemset(c->fR, 0, 4 * ne * sizeof(double));
    if (!e) e = hipDeviceSynchronize();
    if (e) {
        g_error = std::string("o2f_ctx_create: ") + hipGetErrorString(e);
        o2f_ctx_destroy(c.release());
        return O2F_ERR_HIP;
    }
    *out = c.release();
    return O2F_OK;
}

void o2f_ctx_destroy(o2f_ctx *c) {
    if (!c) return;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    void *ptrs[] = {c->cedge, c->cnbr, c->eva, c->evb, c->ecl, c->ecr, c->sptr, c->scell, c->wall, c->wall_by_cell, c->wall_grp, c->cinfo,
                    c->skind, c->ebc, c->vbc, c->U, c->Uold, c->R, c->Gr, c->tq, c->vU, c->vtq, c->fL, c->fR,
                    c->scal, c->partials, c->wall_out, c->speed, c->blk, c->Rp, c->Rm, c->cevx, c->Fe,
                    c->a_veptr, c->a_vedge, c->a_cvtx, c->a_mtgt, c->a_cvw, c->a_work, c->t_work, c->t_out, c->ts_i, c->ts_d,
                    c->pc_color, c->pc_D, c->pc_P};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    for (double *p : c->kv)
        if (p) (void)hipFree(p);
    for (auto &g : c->geo) g.release();
    release_tiles(c);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

int o2f_counts(const o2f_ctx *c, int *nv, int *ne, int *nc) {
    if (!c) return O2F_ERR_ARG;
    if (nv) *nv = c->hm.nv;
    if (ne) *ne = c->hm.ne;
    if (nc) *nc = c->hm.nc;
    return O2F_OK;
}

int o2f_set_params(o2f_ctx *c, const o2f_params *p) {
    if (!c || !p) return O2F_ERR_ARG;
    const bool restencil = p->interp_mode != c->prm.interp_mode;
    c->prm = *p;
    c->phys = make_phys(c->prm, &c->q_inf);
    if (restencil) {
        compute_stencils(c->hm, p->interp_mode);
        HIPCHK(c, hipMemcpyAsync(c->skind, c->hm.skind.data(), c->hm.skind.size(), hipMemcpyHostToDevice, c->stream));
        HIPCHK(c, hipMemcpyAsync(c->scell, c->hm.scell.data(), c->hm.scell.size() * sizeof(int),
                                 hipMemcpyHostToDevice, c->stream));
        HIPCHK(c, c->geo[0].fill(c->hm, c->stream));
        release_tiles(c);  // the stress vertices' stencils changed
        ++c->stencil_version;
    }
    return O2F_OK;
}

int o2f_set_state(o2f_ctx *c, const double *U) {
    if (!c || !U) return O2F_ERR_ARG;
    HIPCHK(c, hipMemcpyAsync(c->U, U, 4 * (size_t)c->hm.nc * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->speed_valid = false;
    return O2F_OK;
}

int o2f_get_state(o2f_ctx *c, double *U) {
    if (!c || !U) return O2F_ERR_ARG;
    return copy_out(c, U, c->U, 4 * (size_t)c->hm.nc);
}

int o2f_init_freestream(o2f_ctx *c) {
    if (!c) return O2F_ERR_ARG;
    // init_one_x + primitives_to_conserved (src/init.cpp:8-16, :59-81)
    const o2f_params &p = c->prm;
    const double rho0 = p.rho_ref, p0 = p.p_ref;
    const double a0 = std::sqrt(p.gamma * p0 / rho0);
    const double u0 = p.inflow_velocity * a0, v0 = 0.0;
    const double vals[4] = {rho0, rho0 * u0, rho0 * v0, p0 / (p.gamma - 1.0) + 0.5 * rho0 * (u0 * u0 + v0 * v0)};
    const long nc = c->hm.nc;
    for (int q = 0; q < 4; ++q) HIPCHK(c, launch_fill(nc, c->U + q * nc, vals[q], c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->speed_valid = false;
    return O2F_OK;
}

int o2f_load_state_file(o2f_ctx *c, const char *path) {
    if (!c || !path) return O2F_ERR_ARG;
    // loadCellStates, src/utils.cpp:29-67
    std::ifstream in(path);
    if (!in) return c->fail(O2F_ERR_IO, std::string("Impossible d'ouvrir ") + path + " en lecture.");
    size_t n = 0;
    in >> n;
    if (!in)
        return c->fail(O2F_ERR_IO, std::string("Fichier corrompu (") + path +
                                       "): impossible de lire le nombre de cellules.");
    const size_t nc = c->hm.nc;
    if (n != nc) {
        std::ostringstream oss;
        oss << "Taille incohérente (" << path << "): fichier=" << n << " vs mesh=" << nc
            << " (vérifier l'ordre/la topologie).";
        return c->fail(O2F_ERR_IO, oss.str());
    }
    std::vector<double> U(4 * nc);
    for (size_t i = 0; i < nc; ++i) {
        double v[4];
        in >> v[0] >> v[1] >> v[2] >> v[3];
        if (!in) {
            std::ostringstream oss;
            oss << "Lecture interrompue à la cellule " << i << " dans " << path << ".";
            return c->fail(O2F_ERR_IO, oss.str());
        }
        for (int q = 0; q < 4; ++q) U[q * nc + i] = v[q];
    }
    return o2f_set_state(c, U.data());
}

int o2f_save_state_file(o2f_ctx *c, const char *path) {
    if (!c || !path) return O2F_ERR_ARG;
    // saveCellStates, src/utils.cpp:70-107
    const size_t nc = c->hm.nc;
    std::vector<double> U(4 * nc);
    int st = o2f_get_state(c, U.data());
    if (st) return st;
    FILE *f = std::fopen(path, "w");
    if (!f) return c->fail(O2F_ERR_IO, std::string("Impossible d'ouvrir ") + path + " en écriture.");
    std::fprintf(f, "%zu\n", nc);
    for (size_t i = 0; i < nc; ++i)
        std::fprintf(f, "%.17e %.17e %.17e %.17e\n", U[i], U[nc + i], U[2 * nc + i], U[3 * nc + i]);
    if (std::fclose(f) != 0) return c->fail(O2F_ERR_IO, std::string("Erreur d'écriture dans ") + path);
    return O2F_OK;
}

int o2f_advection(o2f_ctx *c, double *R_out) {
    if (!c) return O2F_ERR_ARG;
    int st = check_bc(c);
    if (st) return st;
    const Topo T = c->topo();
    const Geom G = c->geo[0].view();
    HIPCHK(c, launch_vertex_state(T, G, c->U, c->vU, c->phys, c->stream));
    HIPCHK(c, launch_reconstruct(T, G, c->U, c->vU, c->fL, c->fR, c->tq, c->Gr, c->phys, c->stream));
    HIPCHK(c, launch_visc_and_flux(T, G, c->tq, c->vtq, c->fL, c->fR, c->Fe, c->phys, 0, c->stream));
    HIPCHK(c, launch_flux_update(T, G, c->U, c->Uold, c->Fe, c->vtq, c->R, nullptr, 0.0, 0.0,
                                 kFlagAdv | kFlagStoreR, c->phys, c->stream));
    ++c->faces_version;
    c->have_grad = true;
    if (!R_out) {
        HIPCHK(c, hipStreamSynchronize(c->stream));
        return O2F_OK;
    }
    return copy_out(c, R_out, c->R, 4 * (size_t)c->hm.nc);
}

int o2f_diffusion(o2f_ctx *c, double *R_out) {
    if (!c) return O2F_ERR_ARG;
    if (!c->have_grad) return c->fail(O2F_ERR_ARG, "o2f_diffusion: no cell gradients yet (call o2f_advection first)");
    int st = check_bc(c);
    if (st) return st;
    const Topo T = c->topo();
    const Geom G = c->geo[0].view();
    HIPCHK(c, launch_stress(T, c->U, c->Gr, c->tq, c->phys, c->stream));
    HIPCHK(c, launch_vertex_visc(T, G, c->tq, c->vtq, c->phys, c->stream));
    HIPCHK(c, launch_flux_update(T, G, c->U, c->Uold, c->Fe, c->vtq, c->R, nullptr, 0.0, 0.0,
                                 kFlagVisc | kFlagStoreR, c->phys, c->stream));
    if (!R_out) {
        HIPCHK(c, hipStreamSynchronize(c->stream));
        return O2F_OK;
    }
    return copy_out(c, R_out, c->R, 4 * (size_t)c->hm.nc);
}

int o2f_residual(o2f_ctx *c, double *R_out) {
    if (!c) return O2F_ERR_ARG;
    int st = residual_on(c, 0);
    if (st) return st;
    if (!R_out) {
        HIPCHK(c, hipStreamSynchronize(c->stream));
        return O2F_OK;
    }
    return copy_out(c, R_out, c->R, 4 * (size_t)c->hm.nc);
}

int o2f_get_gradients(o2f_ctx *c, double *G) {
    if (!c || !G) return O2F_ERR_ARG;
    return copy_out(c, G, c->Gr, 8 * (size_t)c->hm.nc);
}

// Face states live AoS on the device (4 doubles per edge); the ABI is SoA.
static int faces_to_host(o2f_ctx *c, const double *dev, double *host) {
    if (!host) return O2F_OK;
    const size_t ne = c->hm.ne;
    std::vector<double> a(4 * ne);
    int st = copy_out(c, a.data(), dev, 4 * ne);
    if (st) return st;
    for (size_t e = 0; e < ne; ++e)
        for (int q = 0; q < 4; ++q) host[q * ne + e] = a[4 * e + q];
    return O2F_OK;
}

int o2f_get_face_states(o2f_ctx *c, double *L, double *R) {
    if (!c) return O2F_ERR_ARG;
    int st = faces_to_host(c, c->fL, L);
    if (st) return st;
    return faces_to_host(c, c->fR, R);
}

int o2f_set_face_states(o2f_ctx *c, const double *L, const double *R) {
    if (!c || !L || !R) return O2F_ERR_ARG;
    const size_t ne = c->hm.ne;
    std::vector<double> a(4 * ne), b(4 * ne);
    for (size_t e = 0; e < ne; ++e)
        for (int q = 0; q < 4; ++q) {
            a[4 * e + q] = L[q * ne + e];
            b[4 * e + q] = R[q * ne + e];
        }
    HIPCHK(c, hipMemcpyAsync(c->fL, a.data(), 4 * ne * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->fR, b.data(), 4 * ne * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    ++c->faces_version;
    return O2F_OK;
}

int o2f_hllc_flux_batch(const double *qL, const double *qR, const double *normal, double gamma, double *F, int n) {
    if (!qL || !qR || !normal || !F || n < 0) {
        g_error = "o2f_hllc_flux_batch: bad argument";
        return O2F_ERR_ARG;
    }
    if (n == 0) return O2F_OK;
    double *d = nullptr;
    const size_t bytes = (size_t)n * sizeof(double);
    hipError_t e = hipMalloc(reinterpret_cast<void **>(&d), 14 * bytes);
    if (!e) e = hipMemcpy(d, qL, 4 * bytes, hipMemcpyHostToDevice);
    if (!e) e = hipMemcpy(d + 4 * (size_t)n, qR, 4 * bytes, hipMemcpyHostToDevice);
    if (!e) e = hipMemcpy(d + 8 * (size_t)n, normal, 2 * bytes, hipMemcpyHostToDevice);
    if (!e) e = launch_hllc_batch(d, d + 4 * (size_t)n, d + 8 * (size_t)n, gamma, d + 10 * (size_t)n, n, nullptr);
    if (!e) e = hipMemcpy(F, d + 10 * (size_t)n, 4 * bytes, hipMemcpyDeviceToHost);
    if (d) (void)hipFree(d);
    if (e) {
        g_error = std::string("o2f_hllc_flux_batch: ") + hipGetErrorString(e);
        return O2F_ERR_HIP;
    }
    return O2F_OK;
}

int o2f_compute_dt(o2f_ctx *c, double *dt) {
    if (!c || !dt) return O2F_ERR_ARG;
    HIPCHK(c, launch_speed_max(c->U, c->hm.nc, c->phys, c->speed, c->stream));
    HIPCHK(c, launch_step_begin(c->speed, speed_parts(c->hm.nc), c->phys, c->hm.min_h, c->scal, nullptr, 0,
                                c->stream));
    HIPCHK(c, hipMemcpyAsync(dt, c->scal + kScalDt, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->speed_valid = false;
    return O2F_OK;
}

int o2f_rk3_step(o2f_ctx *c, double dt) {
    if (!c) return O2F_ERR_ARG;
    int st = check_bc(c);
    if (st) return st;
    HIPCHK(c, hipMemcpyAsync(c->scal + kScalDt, &dt, sizeof(double), hipMemcpyHostToDevice, c->stream));
    c->speed_valid = false;
    if ((st = rk3_stages(c, true, false))) return st;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return O2F_OK;
}

int o2f_rk3_save_old(o2f_ctx *c) {
    if (!c) return O2F_ERR_ARG;
    HIPCHK(c, hipMemcpyAsync(c->Uold, c->U, 4 * sizeof(double) * c->hm.nc, hipMemcpyDeviceToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return O2F_OK;
}

// One stage on its own: the residual (gradients, face states and R stored
// as computeAdvectionTerm / computeDiffusionTerm leave them) and the update
// U = c0 U_old + c1 (U + dt R) in the same kernels as the fused step, so
// three substeps after o2f_rk3_save_old are bitwise o2f_rk3_step.
int o2f_rk3_substep(o2f_ctx *c, double c0, double c1, double dt) {
    if (!c) return O2F_ERR_ARG;
    int st = check_bc(c);
    if (st) return st;
    HIPCHK(c, hipMemcpyAsync(c->scal + kScalDt, &dt, sizeof(double), hipMemcpyHostToDevice, c->stream));
    c->speed_valid = false;
    if ((st = run_stage(c, 0, kFlagRK | kFlagStoreR, c->scal + kScalDt, c0, c1, true))) return st;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return O2F_OK;
}

int o2f_run_steps(o2f_ctx *c, int nsteps, double *dts, double *t_inout) {
    if (!c || nsteps < 0) return O2F_ERR_ARG;
    int st = check_bc(c);
    if (st) return st;
    double t0 = t_inout ? *t_inout : 0.0;
    HIPCHK(c, hipMemcpyAsync(c->scal + kScalT, &t0, sizeof(double), hipMemcpyHostToDevice, c->stream));
    double *hist = nullptr;
    if (dts && nsteps > 0) HIPCHK(c, hipMalloc(reinterpret_cast<void **>(&hist), nsteps * sizeof(double)));
    for (int s = 0; s < nsteps && st == 0; ++s)
        st = one_step(c, s == nsteps - 1, hist ? hist + s : nullptr, (s > 0 ? kChainUse : 0) | kChainForm);
    const int st_leave = tiles_leave(c);
    if (st) return st;
    if (st_leave) return st_leave;
    if (hist) {
        HIPCHK(c, hipMemcpyAsync(dts, hist, nsteps * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        HIPCHK(c, hipFree(hist));
    }
    double t1 = t0;
    HIPCHK(c, hipMemcpyAsync(&t1, c->scal + kScalT, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (t_inout) *t_inout = t1;
    return O2F_OK;
}

int o2f_run_until(o2f_ctx *c, double final_time, int max_steps, int *steps, double *t_inout, double *t_last_start,
                  double *dt_last) {
    if (!c) return O2F_ERR_ARG;
    int st = check_bc(c);
    if (st) return st;
    hipStream_t s = c->stream;
    const double t0 = t_inout ? *t_inout : 0.0;
    const double init[4] = {1.0, 0.0, t0, 0.0};  // active, steps, t_prev, (unused)
    HIPCHK(c, hipMemcpyAsync(c->scal + kScalT, &t0, sizeof(double), hipMemcpyHostToDevice, s));
    HIPCHK(c, hipMemcpyAsync(c->scal + kScalActive, init, 3 * sizeof(double), hipMemcpyHostToDevice, s));
    const Phys saved = c->phys;
    c->phys.gate = c->scal + kScalActive;
    c->phys.final_t = final_time;
    double h[kScalCount];
    int launched = 0;
    // Any step may be the last: the four-kernel stages store the
    // intermediates of every third stage; tiled steps store none, and the
    // last one's are formed once the loop has stopped.
    const bool tiled = tiles_on(c);
    for (;;) {
        int batch = 64;
        if (max_steps >= 0) batch = std::min(batch, max_steps - launched);
        if (batch <= 0) break;
        for (int k = 0; k < batch && st == 0; ++k)
            st = one_step(c, !tiled, nullptr, (launched + k > 0 ? kChainUse : 0) | kChainForm);
        launched += batch;
        if (st) break;
        hipError_t e = hipMemcpyAsync(h, c->scal, sizeof(h), hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) {
            st = c->fail(O2F_ERR_HIP, std::string("o2f_run_until: ") + hipGetErrorString(e));
            break;
        }
        if (h[kScalActive] == 0.0) break;
    }
    c->phys = saved;
    c->speed_valid = false;  // gated steps may have skipped the speed fold
    const int st_leave = tiles_leave(c);
    if (st) return st;
    if (st_leave) return st_leave;
    HIPCHK(c, hipMemcpyAsync(h, c->scal, sizeof(h), hipMemcpyDeviceToHost, s));
    HIPCHK(c, hipStreamSynchronize(s));
    // the last step's third stage read its input from Uc: its residual,
    // gradients, face states and stresses, as that stage's four kernels
    // would have left them (from an SoA copy of Uc in Ub)
    if (tiled && h[kScalSteps] > 0.0) {
        HIPCHK(c, launch_state_to_soa(c->Uc, c->hm.nc, c->Ub, s));
        if ((st = run_stage(c, 0, kFlagStoreR, nullptr, 0.0, 0.0, true, false, nullptr, c->Ub))) return st;
    }
    if (steps) *steps = static_cast<int>(h[kScalSteps]);
    if (t_inout) *t_inout = h[kScalT];
    if (t_last_start) *t_last_start = h[kScalTPrev];
    if (dt_last) *dt_last = h[kScalDt];
    return O2F_OK;
}

int o2f_time_loop(o2f_ctx *c, double final_time, int max_steps, int *steps, double *t) {
    double tt = 0.0;
    int st = o2f_run_until(c, final_time, max_steps, steps, &tt, nullptr, nullptr);
    if (st == O2F_OK && t) *t = tt;
    return st;
}

// Tiled steps to a density-residual tolerance: the batch path's semantics
// (steps 0, k, 2k, ... checked at their start state; the next step's start
// stops the loop; steps and r from the device) on the tile stages, one host
// poll per 64 steps; the last stage's intermediates formed after the loop.
int o2f_run_to_residual(o2f_ctx *c, double tol, int max_steps, int check_every, int use_graph, int *steps,
                        double *res) {
    if (!c) return O2F_ERR_ARG;
    if (max_steps < 0 || check_every < 1) return c->fail(O2F_ERR_ARG, "batch run: max_steps >= 0, check_every >= 1");
    int st = check_bc(c);
    if (st) return st;
    if (!tiles_on(c)) return run_to_residual_batch(c, tol, max_steps, check_every, use_graph, steps, res);
    hipStream_t s = c->stream;
    const long nrp = (c->hm.nc + kResChunk - 1) / kResChunk;
    if (!c->rparts) HIPCHK(c, dalloc(c->rparts, nrp));
    double init[kScalCount] = {};
    init[kScalActive] = 1.0;
    init[kScalRes] = -1.0;
    HIPCHK(c, hipMemcpyAsync(c->scal + kScalT, &init[kScalT], sizeof(double), hipMemcpyHostToDevice, s));
    HIPCHK(c, hipMemcpyAsync(c->scal + kScalActive, &init[kScalActive], (kScalRes - kScalActive + 1) * sizeof(double),
                             hipMemcpyHostToDevice, s));
    const Phys saved = c->phys;
    c->phys.gate = c->scal + kScalActive;
    c->phys.final_t = HUGE_VAL;
    c->speed_valid = false;
    const double res_tol = tol > 0.0 ? tol : -1.0;
    double h[kScalCount];
    for (int launched = 0; launched < max_steps && st == 0;) {
        const int n = std::min(64, max_steps - launched);
        for (int k = 0; k < n && st == 0; ++k, ++launched) {
            // after a step the speed partials are the new state's (c->U is
            // stale while the state lives in Ua)
            st = tiled_step(c, false, nullptr, tol > 0.0 && launched % check_every == 0, res_tol,
                            (launched > 0 ? kChainUse : 0) | kChainForm);
            c->speed_valid = (st == O2F_OK);
        }
        if (st) break;
        hipError_t e = hipMemcpyAsync(h, c->scal, sizeof(h), hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) st = c->fail(O2F_ERR_HIP, std::string("o2f_run_to_residual: ") + hipGetErrorString(e));
        else if (h[kScalActive] == 0.0) break;
    }
    c->phys = saved;
    c->speed_valid = false;
    const int st_leave = tiles_leave(c);
    if (st) return st;
    if (st_leave) return st_leave;
    HIPCHK(c, hipMemcpyAsync(h, c->scal, sizeof(h), hipMemcpyDeviceToHost, s));
    HIPCHK(c, hipStreamSynchronize(s));
    // the last step's third stage read Uc: its face states, stresses and
    // vertex values as those four kernels leave them (R keeps the last
    // checked step's, as on the batched stages)
    if (h[kScalSteps] > 0.0) {
        HIPCHK(c, launch_state_to_soa(c->Uc, c->hm.nc, c->Ub, s));
        if ((st = run_stage(c, 0, 0, nullptr, 0.0, 0.0, false, false, nullptr, c->Ub))) return st;
    }
    if (steps) *steps = static_cast<int>(h[kScalSteps]);
    if (res) *res = h[kScalRes];
    return O2F_OK;
}

int o2f_residual_norm(o2f_ctx *c, double *rnorm) {
    if (!c || !rnorm) return O2F_ERR_ARG;
    const long nc = c->hm.nc;
    HIPCHK(c, launch_dot2(c->R, c->R, nullptr, nullptr, nc, c->partials, c->scal, kOpStore, c->stream));
    double s = 0.0;
    HIPCHK(c, hipMemcpyAsync(&s, c->scal + kScalDotA, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    const double a_inf = std::sqrt(c->prm.gamma * c->prm.p_ref / c->prm.rho_ref);
    const double scale = c->prm.rho_ref * (c->prm.inflow_velocity * a_inf) / 1.0;
    *rnorm = std::sqrt(s / nc) / scale;
    return O2F_OK;
}

static int wall_sums(o2f_ctx *c, int g, int skip_x, double *L, double *D) {
    const int nw = static_cast<int>(c->hm.wall_edges.size());
    HIPCHK(c, launch_wall_forces(c->geo[g].view(), c->wall, nw, c->fL, c->hm.ne, c->prm.gamma, skip_x, c->wall_out,
                                 c->stream));
    double out[2];
    HIPCHK(c, hipMemcpyAsync(out, c->wall_out, 2 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    *L = out[0];
    *D = out[1];
    return O2F_OK;
}

int o2f_lift_drag(o2f_ctx *c, double *C_L, double *C_D) {
    if (!c || !C_L || !C_D) return O2F_ERR_ARG;
    double L, D;
    int st = wall_sums(c, 0, 0, &L, &D);
    if (st) return st;
    const double corde = 1;
    *C_L = L / (c->q_inf * corde);
    *C_D = D / (c->q_inf * corde);
    return O2F_OK;
}

static int objective_on(o2f_ctx *c, int g, double beta, double *J) {
    double L, D;
    int st = wall_sums(c, g, 1, &L, &D);
    if (st) return st;
    const double corde = 1;
    const double CL = L / (c->q_inf * corde), CD = D / (c->q_inf * corde);
    *J = -CL + beta * CD;
    return O2F_OK;
}

int o2f_objective(o2f_ctx *c, double beta, double *J) {
    if (!c || !J) return O2F_ERR_ARG;
    return objective_on(c, 0, beta, J);
}

static int ensure_work(o2f_ctx *c) {
    const size_t n = 4 * (size_t)c->hm.nc;
    for (double *&p : c->kv)
        if (!p) HIPCHK(c, dalloc(p, n));
    if (!c->blk) HIPCHK(c, dalloc(c->blk, 32 * (size_t)c->hm.ne));
    return O2F_OK;
}

static int rhs_device(o2f_ctx *c, double beta, double *b) {
    const long n = 4L * c->hm.nc;
    HIPCHK(c, launch_fill(n, b, 0.0, c->stream));
    HIPCHK(c, launch_adjoint_rhs(c->topo(), c->geo[0].view(), c->wall_by_cell, c->wall_grp, c->wall_ngrp, c->fL,
                                 c->prm.gamma, c->q_inf, beta, b, c->stream));
    return O2F_OK;
}

int o2f_adjoint_rhs(o2f_ctx *c, double beta, double *b) {
    if (!c || !b) return O2F_ERR_ARG;
    int st = ensure_work(c);
    if (st) return st;
    if ((st = rhs_device(c, beta, c->kv[0]))) return st;
    return copy_out(c, b, c->kv[0], 4 * (size_t)c->hm.nc);
}

// FD blocks at the current face states, rebuilt only when the face states
// or eps changed.
static int ensure_blocks(o2f_ctx *c, double eps) {
    int st = ensure_work(c);
    if (st) return st;
    if (c->blk_version == c->faces_version && c->blk_eps == eps) return O2F_OK;
    HIPCHK(c, launch_jt_blocks(c->topo(), c->geo[0].view(), c->fL, c->fR, eps, c->phys, c->blk, c->stream));
    c->blk_version = c->faces_version;
    c->blk_eps = eps;
    return O2F_OK;
}

int o2f_apply_jt(o2f_ctx *c, const double *x, double *y, double eps) {
    if (!c || !x || !y) return O2F_ERR_ARG;
    int st = ensure_blocks(c, eps);
    if (st) return st;
    const size_t n = 4 * (size_t)c->hm.nc;
    HIPCHK(c, hipMemcpyAsync(c->kv[0], x, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, launch_jt_apply(c->topo(), c->geo[0].view(), c->blk, c->kv[0], c->kv[1], c->stream));
    return copy_out(c, y, c->kv[1], n);
}

namespace {
struct HostScal {
    double v[kScalCount];
};
int read_scal(o2f_ctx *c, HostScal &h) {
    HIPCHK(c, hipMemcpyAsync(h.v, c->scal, sizeof(h.v), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return O2F_OK;
}
int set_scal(o2f_ctx *c, int slot, double v) {
    HIPCHK(c, hipMemcpyAsync(c->scal + slot, &v, sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return O2F_OK;
}

// One call of Eigen's internal::bicgstab (identity preconditioner) starting
// from x: returns the number of iterations it ran.
// y = A x on the device, A = J^T (either form).
using MatVec = std::function<hipError_t(const double *, double *)>;

// pc: a right preconditioner y = M^{-1} p (Eigen's preconditioned
// recurrence: v = A M^{-1} p, t = A M^{-1} s, x += alpha M^{-1} p + w M^{-1} s);
// empty: the identity, as solveAdjointSystem's IdentityPreconditioner.
int bicgstab_chunk(o2f_ctx *c, const double *b, double *x, int maxit, double tol, int *iters, const MatVec &mv,
                   const MatVec &pc = nullptr) {
    hipStream_t s = c->stream;
    const long n = 4L * c->hm.nc;
    double *r = c->kv[1], *r0 = c->kv[2], *v = c->kv[3], *p = c->kv[4], *sv = c->kv[5], *t = c->kv[6];
    double *tmp = c->kv[7], *py = c->kv[10], *sz = c->kv[11];
    double *wparts = c->partials + 2 * kRedGrid;  // the (t.t, t.s) partials
    HostScal h;
    int st;
    // r = b - A x; r0 = r
    HIPCHK(c, mv(x, tmp));
    HIPCHK(c, launch_sub(n, b, tmp, r, s));
    HIPCHK(c, launch_sub(n, r, nullptr, r0, s));
    HIPCHK(c, launch_dot2(r0, r0, b, b, n, c->partials, c->scal, kOpStore, s));
    if ((st = read_scal(c, h))) return st;
    const double r0_sq = h.v[kScalDotA], rhs_sq = h.v[kScalDotB];
    if (rhs_sq == 0) {
        HIPCHK(c, launch_fill(n, x, 0.0, s));
        *iters = 0;
        return O2F_OK;
    }
    // rho = alpha = w = 1; v = p = 0; r0_sqnorm = r0.r0; no iterations yet
    const double init[4] = {1.0, 1.0, 1.0, 1.0};
    HIPCHK(c, hipMemcpyAsync(c->scal + kScalRho, init, 4 * sizeof(double), hipMemcpyHostToDevice, s));
    HIPCHK(c, launch_fill(n, v, 0.0, s));
    HIPCHK(c, launch_fill(n, p, 0.0, s));
    if ((st = set_scal(c, kScalR0Sq, r0_sq)) || (st = set_scal(c, kScalIter, 0.0))) return st;
    // The loop test (r.r > tol2 && i < maxit, then the restart test) runs on
    // the device, in the last reduction of each iteration: it sets the flag
    // scal[kScalGo] that every vector update and reduction of the body
    // checks.  The host enqueues kGroup iterations and reads the flag once
    // per group; iterations enqueued after a stop change nothing (only
    // their matrix-vector products run).  A restart (flag 2) is done here.
    const BicgTest test{tol * tol * rhs_sq, (double)maxit, c->bicg_restart_eps * c->bicg_restart_eps};
    const double *go = c->scal + kScalGo;
    constexpr int kGroup = 8;
    // The loop body rotated to start after the direction update: v = A p,
    // s, t = A s, the last update, then the next iteration's top (rho, the
    // loop test) with its p update.  alpha, w and rho are formed by the
    // kernels that use them, each block from the partials (no one-block
    // launch in between).  The p update after the first top and after a
    // restart runs on its own (launch_bicg_p).
    const auto enqueue_group = [&]() -> hipError_t {
        hipError_t e = hipSuccess;
        for (int g = 0; g < kGroup && e == hipSuccess; ++g) {
            if (pc && (e = pc(p, py)) != hipSuccess) break;
            if ((e = mv(pc ? py : p, v)) != hipSuccess) break;
            if ((e = launch_dot2_partial(r0, v, nullptr, nullptr, n, c->partials, go, s)) != hipSuccess) break;
            if ((e = launch_bicg_s(n, r, v, sv, c->scal, c->partials, g & 1, s)) != hipSuccess) break;
            if (pc && (e = pc(sv, sz)) != hipSuccess) break;
            if ((e = mv(pc ? sz : sv, t)) != hipSuccess) break;
            if ((e = launch_dot2_partial(t, t, t, sv, n, wparts, go, s)) != hipSuccess) break;
            if ((e = launch_bicg_xr_dot(n, x, r, pc ? py : p, pc ? sz : sv, sv, t, r0, wparts, c->partials,
                                        c->scal, s)) != hipSuccess)
                break;
            e = launch_bicg_p_top(n, r, p, v, c->scal, c->partials, g & 1, test, s);
        }
        return e;
    };
    // the replayed group; destroyed on every way out of this call
    hipGraphExec_t group = nullptr;
    struct GroupGuard {
        hipGraphExec_t &x;
        ~GroupGuard() {
            if (x) (void)hipGraphExecDestroy(x);
        }
    } guard{group};
    int restarts = 0;
    bool owe_p = true;  // the first top (below) and a restarted one still owe their p update
    // rho_old = rho; rho = r0.r, together with r.r for the loop test
    HIPCHK(c, launch_dot2(r0, r, r, r, n, c->partials, c->scal, kOpRho, s, nullptr, test));
    if ((st = read_scal(c, h))) return st;
    while (h.v[kScalGo] != 0.0) {
        if (h.v[kScalGo] == 2.0) {
            // r became too orthogonal to r0: restart with r0 = r = b - A x
            HIPCHK(c, mv(x, tmp));
            HIPCHK(c, launch_sub(n, b, tmp, r, s));
            HIPCHK(c, launch_sub(n, r, nullptr, r0, s));
            HIPCHK(c, launch_dot2(r, r, nullptr, nullptr, n, c->partials, c->scal, kOpRestart, s));
            if (restarts++ == 0 && (st = set_scal(c, kScalIter, 0.0))) return st;
            if ((st = set_scal(c, kScalGo, 1.0))) return st;
            owe_p = true;
        }
        if (owe_p) {
            // the p update of the top just evaluated, on its own
            HIPCHK(c, launch_bicg_p(n, r, p, v, c->scal, s));
            owe_p = false;
        }
        if (!group) {
            // the group is captured once per call and replayed: its kernels
            // take only the call's fixed buffers and read every scalar from
            // the device, so each replay is one eager group
            HIPCHK(c, hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
            const hipError_t eq = enqueue_group();
            hipGraph_t g = nullptr;
            const hipError_t ec = hipStreamEndCapture(s, &g);
            if (eq != hipSuccess || ec != hipSuccess) {
                if (g) (void)hipGraphDestroy(g);
                HIPCHK(c, eq != hipSuccess ? eq : ec);
            }
            const hipError_t ei = hipGraphInstantiate(&group, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            HIPCHK(c, ei);
        }
        HIPCHK(c, hipGraphLaunch(group, s));
        if ((st = read_scal(c, h))) return st;
    }
    const int i = (int)h.v[kScalIter];
    *iters = i;
    return O2F_OK;
}
}  // namespace

namespace {

// solveAdjointSystem's loop (src/shape_optimization.cpp:690-736): chunks of
// `chunk` iterations restarted from the current x, the true residual checked
// after each; the RHS is assembled at the current face states.
// The right-hand side of a solve: b (4 nc, device) for beta.
using RhsFn = std::function<int(double beta, double *b)>;

int build_block_jacobi(o2f_ctx *c, const MatVec &mv);

int bicgstab_solve(o2f_ctx *c, double beta, double tol, int chunk, int max_total, double *lambda, int *iters,
                   double *relres, const MatVec &mv, const RhsFn &rhs = nullptr, const MatVec &pc = nullptr) {
    int st;
    const long n = 4L * c->hm.nc;
    double *b = c->kv[0], *x = c->kv[8], *tmp = c->kv[9];
    if ((st = rhs ? rhs(beta, b) : rhs_device(c, beta, b))) return st;
    HIPCHK(c, launch_fill(n, x, 0.0, c->stream));
    HostScal h;
    HIPCHK(c, launch_dot2(b, b, nullptr, nullptr, n, c->partials, c->scal, kOpStore, c->stream));
    if ((st = read_scal(c, h))) return st;
    const double bnorm = std::sqrt(h.v[kScalDotA]);
    double rr = INFINITY;
    int total = 0, status = O2F_OK;
    while (total < max_total) {
        int it = 0;
        if ((st = bicgstab_chunk(c, b, x, chunk, tol, &it, mv, pc))) return st;
        total += it;
        // true residual b - A x (src/shape_optimization.cpp:715-717)
        HIPCHK(c, mv(x, tmp));
        HIPCHK(c, launch_sub(n, b, tmp, tmp, c->stream));
        HIPCHK(c, launch_dot2(tmp, tmp, nullptr, nullptr, n, c->partials, c->scal, kOpStore, c->stream));
        if ((st = read_scal(c, h))) return st;
        rr = std::sqrt(h.v[kScalDotA]) / (bnorm > 0 ? bnorm : 1.0);
        if (rr <= tol) break;
        if (!std::isfinite(rr)) {
            status = c->fail(O2F_ERR_DIVERGED, "BiCGSTAB diverged");
            break;
        }
        if (it == 0) break;  // no progress possible from this x
    }
    if (status == O2F_OK && rr > tol) status = c->fail(O2F_ERR_NOCONV, "BiCGSTAB did not converge");
    if (iters) *iters = total;
    if (relres) *relres = rr;
    st = copy_out(c, lambda, x, n);
    return st ? st : status;
}

}  // namespace

int o2f_bicgstab_set_restart_eps(o2f_ctx *c, double restart_eps) {
    if (!c || !(restart_eps >= 0.0)) return O2F_ERR_ARG;
    c->bicg_restart_eps = restart_eps;
    return O2F_OK;
}

int o2f_bicgstab(o2f_ctx *c, double beta, double eps, double tol, int chunk, int max_total, double *lambda,
                 int *iters, double *relres) {
    if (!c || !lambda || chunk <= 0) return O2F_ERR_ARG;
    int st = ensure_blocks(c, eps);
    if (st) return st;
    const Topo T = c->topo();
    const Geom G = c->geo[0].view();
    const MatVec mv = [c, T, G](const double *x, double *y) {
        return launch_jt_apply(T, G, c->blk, x, y, c->stream);
    };
    return bicgstab_solve(c, beta, tol, chunk, max_total, lambda, iters, relres, mv);
}

int o2f_bicgstab_options(o2f_ctx *c, double beta, double eps, double tol, int chunk, int max_total, int options,
                         double *lambda, int *iters, double *relres) {
    if (!c || !lambda || chunk <= 0 || (options & ~O2F_ADJ_BLOCK_JACOBI)) return O2F_ERR_ARG;
    int st = ensure_blocks(c, eps);
    if (st) return st;
    const Topo T = c->topo();
    const Geom G = c->geo[0].view();
    const MatVec mv = [c, T, G](const double *x, double *y) {
        return launch_jt_apply(T, G, c->blk, x, y, c->stream);
    };
    MatVec pc = nullptr;
    if (options & O2F_ADJ_BLOCK_JACOBI) {
        if ((st = build_block_jacobi(c, mv))) return st;
        pc = [c](const double *in, double *out) {
            return launch_apply_blocks(c->hm.nc, c->pc_P, in, out, c->stream);
        };
    }
    return bicgstab_solve(c, beta, tol, chunk, max_total, lambda, iters, relres, mv, nullptr, pc);
}

// ---- exact adjoint ---------------------------------------------------------

namespace {

// Sub-array sizes (doubles) of the adjoint work block, each rounded up to a
// multiple of 32 doubles so every array starts 256-byte aligned (the AoS
// arrays are read as 16-byte pairs).
size_t adj_round(size_t n) { return (n + 31) / 32 * 32; }

AdjWork adj_work(o2f_ctx *c) {
    const size_t nc = c->hm.nc, ne = c->hm.ne, nv = c->hm.nv;
    double *p = c->a_work;
    AdjWork W{};
    auto take = [&](size_t n) {
        double *q = p;
        p += adj_round(n);
        return q;
    };
    W.fbL = take(4 * ne);
    W.fbR = take(4 * ne);
    W.th = take(5 * ne);
    W.vtb = take(5 * nv);
    W.ub = take(4 * nc);
    W.gb = take(8 * nc);
    W.mval = take(4 * nc);
    W.ha = take(4 * ne);
    W.vi = take(4 * nv);
    W.mtgt = c->a_mtgt;
    return W;
}

size_t adj_work_size(const HostMesh &hm) {
    const size_t nc = hm.nc, ne = hm.ne, nv = hm.nv;
    return 3 * adj_round(4 * ne) + adj_round(5 * ne) + adj_round(5 * nv) + 2 * adj_round(4 * nc) +
           adj_round(8 * nc) + adj_round(4 * nv);
}

// Vertex -> edge CSR, cell vertices, and the weight of each cell in the
// interpolation stencil of each of its vertices (the transpose of the
// cell -> vertex gather).
int ensure_adjoint(o2f_ctx *c) {
    int st = ensure_work(c);
    if (st) return st;
    const HostMesh &hm = c->hm;
    if (!c->a_veptr) {
        std::vector<int> ptr(hm.nv + 1, 0), ed(2 * (size_t)hm.ne);
        for (int e = 0; e < hm.ne; ++e) {
            ++ptr[hm.eva[e] + 1];
            ++ptr[hm.evb[e] + 1];
        }
        for (int v = 0; v < hm.nv; ++v) ptr[v + 1] += ptr[v];
        std::vector<int> fill(ptr.begin(), ptr.end() - 1);
        for (int e = 0; e < hm.ne; ++e) {
            ed[fill[hm.eva[e]]++] = e;
            ed[fill[hm.evb[e]]++] = e;
        }
        HIPCHK(c, upload(c->a_veptr, ptr));
        HIPCHK(c, upload(c->a_vedge, ed));
        HIPCHK(c, upload(c->a_cvtx, hm.cvtx));
        HIPCHK(c, dalloc(c->a_work, adj_work_size(hm)));
        HIPCHK(c, dalloc(c->a_mtgt, 4 * (size_t)hm.nc));
        HIPCHK(c, dalloc(c->a_cvw, 4 * (size_t)hm.nc));
    }
    if (c->a_interp != c->prm.interp_mode) {
        std::vector<double> w(4 * (size_t)hm.nc, 0.0);
        for (int cc = 0; cc < hm.nc; ++cc)
            for (int j = 0; j < 4; ++j) {
                const int v = hm.cvtx[4 * cc + j];
                if (v < 0) continue;
                const int b = hm.vptr[v], n = hm.vptr[v + 1] - b;
                const uint8_t kind = hm.skind[v];
                for (int k = 0; k < n; ++k)
                    if (hm.scell[b + k] == cc) {
                        w[4 * cc + j] = kind == kStencilWsum ? hm.sw[b + k]
                                        : kind == kStencilAvg2 ? 0.5
                                        : kind == kStencilCopy ? 1.0
                                                               : 0.0;
                        break;
                    }
            }
        HIPCHK(c, hipMemcpy(c->a_cvw, w.data(), w.size() * sizeof(double), hipMemcpyHostToDevice));
        c->a_interp = c->prm.interp_mode;
    }
    return O2F_OK;
}

// Forward state of the linearisation point, then y = (dR/dU)^T x on the device.
MatVec exact_matvec(o2f_ctx *c) {
    const Topo T = c->topo();
    const Geom G = c->geo[0].view();
    const AdjTopo A{c->a_veptr, c->a_vedge, c->a_cvtx, c->a_cvw};
    const AdjWork W = adj_work(c);
    return [c, T, G, A, W](const double *x, double *y) {
        return launch_adjoint_exact(T, G, c->phys, A, W, x, c->U, c->vU, c->fL, c->fR, c->vtq, y, c->stream);
    };
}

}  // namespace

int o2f_apply_jt_exact(o2f_ctx *c, const double *x, double *y) {
    if (!c || !x || !y) return O2F_ERR_ARG;
    int st = ensure_adjoint(c);
    if (st) return st;
    if ((st = residual_on(c, 0))) return st;
    const size_t n = 4 * (size_t)c->hm.nc;
    HIPCHK(c, hipMemcpyAsync(c->kv[0], x, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, exact_matvec(c)(c->kv[0], c->kv[1]));
    return copy_out(c, y, c->kv[1], n);
}

namespace {

// Greedy colouring of the cells such that two cells of one colour are more
// than two vertex rings apart: no cell of a colour lies in the residual
// stencil of another (R_c depends on U only within two vertex rings of c:
// the face neighbours' reconstructions and the vertex stresses of c's
// vertices).
int distance2_colors(const HostMesh &hm, std::vector<int> &color) {
    color.assign(hm.nc, -1);
    std::vector<int> seen;
    int ncol = 0;
    for (int c = 0; c < hm.nc; ++c) {
        for (int k = 0; k < hm.cnv[c]; ++k) {
            const int v = hm.cvtx[4 * c + k];
            for (int i = hm.vptr[v]; i < hm.vptr[v + 1]; ++i) {
                const int c1 = hm.vcell_asc[i];
                for (int k1 = 0; k1 < hm.cnv[c1]; ++k1) {
                    const int v1 = hm.cvtx[4 * c1 + k1];
                    for (int j = hm.vptr[v1]; j < hm.vptr[v1 + 1]; ++j) {
                        const int col = color[hm.vcell_asc[j]];
                        if (col >= 0) {
                            if (col >= (int)seen.size()) seen.resize(col + 1, -1);
                            seen[col] = c;
                        }
                    }
                }
            }
        }
        int col = 0;
        while (col < (int)seen.size() && seen[col] == c) ++col;
        color[c] = col;
        ncol = std::max(ncol, col + 1);
    }
    return ncol;
}

// The diagonal 4x4 blocks D_c = dR_c/dU_c of the operator behind mv (y =
// J^T x) by probing, one colour and variable at a time, and P_c =
// (D_c^T)^{-1}.  Uses work vectors kv[10] (x) and kv[11] (y).
int build_block_jacobi(o2f_ctx *c, const MatVec &mv) {
    const int nc = c->hm.nc;
    if (!c->pc_color) {
        std::vector<int> color;
        c->pc_ncolors = distance2_colors(c->hm, color);
        HIPCHK(c, upload(c->pc_color, color));
    }
    if (!c->pc_D) HIPCHK(c, dalloc(c->pc_D, 16 * (size_t)nc));
    if (!c->pc_P) HIPCHK(c, dalloc(c->pc_P, 16 * (size_t)nc));
    double *x = c->kv[10], *y = c->kv[11];
    for (int k = 0; k < c->pc_ncolors; ++k)
        for (int q = 0; q < 4; ++q) {
            HIPCHK(c, launch_probe_seed(nc, c->pc_color, k, q, x, c->stream));
            HIPCHK(c, mv(x, y));
            HIPCHK(c, launch_probe_take(nc, c->pc_color, k, q, y, c->pc_D, c->stream));
        }
    HIPCHK(c, launch_invert_blocks(nc, c->pc_D, c->pc_P, c->stream));
    return O2F_OK;
}

// (dJ/dU)^T through the reconstruction: the exact adjoint chain with x = 0
// and the objective's face-state derivative as its seed.  Work vector
// kv[8] is used as the zero x (the solve fills it with its zero start
// afterwards anyway).
int rhs_exact_device(o2f_ctx *c, double beta, double *b) {
    const long n = 4L * c->hm.nc;
    double *zero = c->kv[8];
    HIPCHK(c, launch_fill(n, zero, 0.0, c->stream));
    const AdjTopo A{c->a_veptr, c->a_vedge, c->a_cvtx, c->a_cvw};
    const ObjSeed seed{c->wall, (int)c->hm.wall_edges.size(), c->q_inf, beta};
    HIPCHK(c, launch_adjoint_exact(c->topo(), c->geo[0].view(), c->phys, A, adj_work(c), zero, c->U, c->vU, c->fL,
                                   c->fR, c->vtq, b, c->stream, &seed));
    return O2F_OK;
}

}  // namespace

int o2f_adjoint_rhs_exact(o2f_ctx *c, double beta, double *b) {
    if (!c || !b) return O2F_ERR_ARG;
    int st = ensure_adjoint(c);
    if (st) return st;
    if ((st = residual_on(c, 0))) return st;
    if ((st = rhs_exact_device(c, beta, c->kv[0]))) return st;
    return copy_out(c, b, c->kv[0], 4 * (size_t)c->hm.nc);
}

int o2f_bicgstab_exact(o2f_ctx *c, double beta, double tol, int chunk, int max_total, double *lambda, int *iters,
                       double *relres, int options) {
    if (!c || !lambda || chunk <= 0 || (options & ~(O2F_ADJ_EXACT_RHS | O2F_ADJ_BLOCK_JACOBI))) return O2F_ERR_ARG;
    int st = ensure_adjoint(c);
    if (st) return st;
    if ((st = residual_on(c, 0))) return st;
    RhsFn rhs = nullptr;
    if (options & O2F_ADJ_EXACT_RHS) rhs = [c](double beta_, double *b) { return rhs_exact_device(c, beta_, b); };
    const MatVec mv = exact_matvec(c);
    MatVec pc = nullptr;
    if (options & O2F_ADJ_BLOCK_JACOBI) {
        if ((st = build_block_jacobi(c, mv))) return st;
        pc = [c](const double *in, double *out) {
            return launch_apply_blocks(c->hm.nc, c->pc_P, in, out, c->stream);
        };
    }
    return bicgstab_solve(c, beta, tol, chunk, max_total, lambda, iters, relres, mv, rhs, pc);
}

int o2f_adjoint_block_diagonal(o2f_ctx *c, int exact, double *blocks) {
    if (!c || !blocks) return O2F_ERR_ARG;
    int st = exact ? ensure_adjoint(c) : ensure_work(c);
    if (st) return st;
    if ((st = residual_on(c, 0))) return st;
    MatVec mv;
    if (exact) {
        mv = exact_matvec(c);
    } else {
        if ((st = ensure_blocks(c, 1e-6))) return st;
        const Topo T = c->topo();
        const Geom G = c->geo[0].view();
        mv = [c, T, G](const double *a, double *b) { return launch_jt_apply(T, G, c->blk, a, b, c->stream); };
    }
    if ((st = build_block_jacobi(c, mv))) return st;
    return copy_out(c, blocks, c->pc_D, 16 * (size_t)c->hm.nc);
}

namespace {
int scatter_entries(o2f_ctx *c, GeomSet &gs, const GeomSparse &t);
// dst = src, every array of the geometry set (device to device)
int copy_set(o2f_ctx *c, const GeomSet &src, GeomSet &dst) {
    const size_t nc = c->hm.nc, ne = c->hm.ne, nsw = c->hm.sw.size() + kStencilSlots;
    const GeomW a = src.writable(), b = dst.writable();
    const std::pair<double *, double *> cells[] = {{a.cvol, b.cvol}, {a.ccx, b.ccx}, {a.ccy, b.ccy}};
    const std::pair<double *, double *> edges[] = {
        {a.elen, b.elen}, {a.enx, b.enx}, {a.eny, b.eny}, {a.emx, b.emx}, {a.emy, b.emy}};
    for (const auto &pr : cells)
        HIPCHK(c, hipMemcpyAsync(pr.second, pr.first, nc * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    for (const auto &pr : edges)
        HIPCHK(c, hipMemcpyAsync(pr.second, pr.first, ne * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(b.sw, a.sw, nsw * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    return O2F_OK;
}
}  // namespace

int o2f_shape_gradient(o2f_ctx *c, const char *const *names_all, const double *values_all, int n_all,
                       const char *const *opt_names, int n_opt, const double *lambda, double beta, double eps,
                       double *grad) {
    if (!c || (n_all > 0 && (!names_all || !values_all)) || (n_opt > 0 && (!opt_names || !grad)) || !lambda)
        return O2F_ERR_ARG;
    // getParamOrDefault (src/shape_optimization.cpp:101-109, defaults :830-834)
    auto param = [&](const char *key, double def) {
        for (int i = 0; i < n_all; ++i)
            if (std::strcmp(names_all[i], key) == 0) return values_all[i];
        return def;
    };
    const double pos_x = param("pos_x", 0.355), chord = param("c", 1.0), m0 = param("m", 0.02);
    const double p0 = param("p", 0.4), t = param("t", 0.12);
    int st = ensure_work(c);
    if (st) return st;
    const long n = 4L * c->hm.nc;
    double *lam = c->kv[0], *dR = c->kv[1];
    HIPCHK(c, hipMemcpyAsync(lam, lambda, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
    if (!c->Rp) HIPCHK(c, dalloc(c->Rp, n));
    if (!c->Rm) HIPCHK(c, dalloc(c->Rm, n));
    // computeResidual(mesh): face states at the current state
    if ((st = residual_on(c, 0))) return st;
    for (int k = 0; k < n_opt; ++k) {
        const std::string name = opt_names[k];
        const int which = (name == "m" || name == "cambrure") ? 0 : (name == "p" || name == "position_cambrure") ? 1 : -1;
        if (which < 0) {  // applyShapePerturbationNACA12 leaves the mesh unchanged
            grad[k] = 0.0;
            continue;
        }
        double J[2];
        double *Rpm[2] = {c->Rp, c->Rm};
        for (int side = 0; side < 2; ++side) {
            // the perturbed mesh (applyShapePerturbationNACA12 on a copy, then
            // the stencils): the base geometry with the entries the
            // perturbation changes written over it, on the device
            GeomSparse pert;
            perturb_naca_sparse(c->hm, c->prm.interp_mode, which, pos_x, chord, m0, p0, t, side == 0 ? eps : -eps,
                                pert);
            GeomSet &gs = c->geo[1 + side];
            if (!gs.cvol) HIPCHK(c, gs.alloc(c->hm));
            if ((st = copy_set(c, c->geo[0], gs))) return st;
            if ((st = scatter_entries(c, gs, pert))) return st;
            if ((st = residual_on(c, 1 + side))) return st;
            HIPCHK(c, hipMemcpyAsync(Rpm[side], c->R, n * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
            if ((st = objective_on(c, 1 + side, beta, &J[side]))) return st;
        }
        HIPCHK(c, launch_fd_diff(n, c->Rp, c->Rm, eps, dR, c->stream));
        HIPCHK(c, launch_dot2(lam, dR, nullptr, nullptr, n, c->partials, c->scal, kOpStore, c->stream));
        HostScal h;
        if ((st = read_scal(c, h))) return st;
        const double dJ_direct = (J[0] - J[1]) / (2.0 * eps);
        grad[k] = dJ_direct - h.v[kScalDotA];
    }
    // leave the context as computeResidual(mesh) left the reference's mesh
    if ((st = residual_on(c, 0))) return st;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return O2F_OK;
}

// ---- geometric derivatives ------------------------------------------------

namespace {

// getParamOrDefault (src/shape_optimization.cpp:101-109, defaults :830-834)
struct NacaParams {
    double pos_x, chord, m, p, t;
};
NacaParams naca_params(const char *const *names_all, const double *values_all, int n_all) {
    auto param = [&](const char *key, double def) {
        for (int i = 0; i < n_all; ++i)
            if (std::strcmp(names_all[i], key) == 0) return values_all[i];
        return def;
    };
    return {param("pos_x", 0.355), param("c", 1.0), param("m", 0.02), param("p", 0.4), param("t", 0.12)};
}

// 0 camber, 1 camber position, 2 thickness, -1 anything else
int naca_which(const char *name) {
    const std::string n = name ? name : "";
    if (n == "m" || n == "cambrure") return 0;
    if (n == "p" || n == "position_cambrure") return 1;
    if (n == "t") return 2;
    return -1;
}

// The listed entries of a geometry set written over its device arrays
// (uploaded as index / value lists, then scattered).
int scatter_entries(o2f_ctx *c, GeomSet &gs, const GeomSparse &t) {
    std::vector<int> ip;
    ip.reserve(t.cell.size() + t.edge.size() + t.swi.size());
    ip.insert(ip.end(), t.cell.begin(), t.cell.end());
    ip.insert(ip.end(), t.edge.begin(), t.edge.end());
    ip.insert(ip.end(), t.swi.begin(), t.swi.end());
    std::vector<double> dp;
    dp.reserve(t.cval.size() + t.eval.size() + t.swv.size());
    dp.insert(dp.end(), t.cval.begin(), t.cval.end());
    dp.insert(dp.end(), t.eval.begin(), t.eval.end());
    dp.insert(dp.end(), t.swv.begin(), t.swv.end());
    if (ip.size() > c->ts_ni) {
        if (c->ts_i) (void)hipFree(c->ts_i);
        c->ts_i = nullptr;
        c->ts_ni = 0;
        HIPCHK(c, dalloc(c->ts_i, ip.size()));
        c->ts_ni = ip.size();
    }
    if (dp.size() > c->ts_nd) {
        if (c->ts_d) (void)hipFree(c->ts_d);
        c->ts_d = nullptr;
        c->ts_nd = 0;
        HIPCHK(c, dalloc(c->ts_d, dp.size()));
        c->ts_nd = dp.size();
    }
    if (!ip.empty()) {
        HIPCHK(c, hipMemcpyAsync(c->ts_i, ip.data(), ip.size() * sizeof(int), hipMemcpyHostToDevice, c->stream));
        HIPCHK(c, hipMemcpyAsync(c->ts_d, dp.data(), dp.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
    }
    const int ncell = (int)t.cell.size(), nedge = (int)t.edge.size(), nsw = (int)t.swi.size();
    HIPCHK(c, launch_scatter_geometry(gs.writable(), ncell, c->ts_i, c->ts_d, nedge, c->ts_i + ncell,
                                      c->ts_d + 3 * ncell, nsw, c->ts_i + ncell + nedge, c->ts_d + 3 * ncell + 5 * nedge,
                                      c->stream));
    // the host vectors are read by the copies above
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return O2F_OK;
}

int ensure_tangent(o2f_ctx *c) {
    if (c->t_work) return O2F_OK;
    const size_t nv = c->hm.nv, ne = c->hm.ne, nc = c->hm.nc;
    HIPCHK(c, dalloc(c->t_work, 9 * nv + 12 * ne + 9 * nc));
    HIPCHK(c, dalloc(c->t_out, 4));
    return O2F_OK;
}

TanWork tan_work(o2f_ctx *c) {
    const size_t nv = c->hm.nv, ne = c->hm.ne, nc = c->hm.nc;
    double *p = c->t_work;
    TanWork W;
    W.dvU = p;
    p += 4 * nv;
    W.dvtq = p;
    p += 5 * nv;
    W.dfL = p;
    p += 4 * ne;
    W.dfR = p;
    p += 4 * ne;
    W.dFe = p;
    p += 4 * ne;
    W.dtq = p;
    p += 5 * nc;
    W.dR = p;
    return W;
}

// The residual and its derivative along dxy at the current state: R in
// c->R, dR in the tangent work; *dJ the objective's derivative.
int tangent_on(o2f_ctx *c, const double *dxy, int geometry, double beta, double *dJ) {
    if (geometry != O2F_GEOM_REFERENCE && geometry != O2F_GEOM_FULL)
        return c->fail(O2F_ERR_ARG, "geometry must be O2F_GEOM_REFERENCE or O2F_GEOM_FULL");
    int st = check_bc(c);
    if (st) return st;
    if ((st = ensure_tangent(c))) return st;
    // the derivative set: zeros, then the entries that can be non-zero
    GeomSparse t;
    geometry_tangent_sparse(c->hm, dxy, geometry, t);
    GeomSet &gs = c->geo[2];
    if (!gs.cvol) HIPCHK(c, gs.alloc(c->hm));
    {
        const size_t nc = c->hm.nc, ne = c->hm.ne, nsw = c->hm.sw.size() + kStencilSlots;
        const GeomW g = gs.writable();
        for (double *p : {g.cvol, g.ccx, g.ccy}) HIPCHK(c, launch_fill((long)nc, p, 0.0, c->stream));
        for (double *p : {g.elen, g.enx, g.eny, g.emx, g.emy}) HIPCHK(c, launch_fill((long)ne, p, 0.0, c->stream));
        HIPCHK(c, launch_fill((long)nsw, g.sw, 0.0, c->stream));
    }
    if ((st = scatter_entries(c, gs, t))) return st;
    const Topo T = c->topo();
    const Geom G = c->geo[0].view(), dG = gs.view();
    const TanWork W = tan_work(c);
    HIPCHK(c, launch_residual_tangent(T, G, dG, c->phys, c->U, nullptr, c->vU, c->fL, c->fR, c->tq, c->vtq, c->Fe,
                                      c->R, W, c->stream));
    ++c->faces_version;
    HIPCHK(c, launch_objective_tangent(G, dG, c->wall, (int)c->hm.wall_edges.size(), c->fL, W.dfL, c->prm.gamma, 1,
                                       c->t_out, c->stream));
    double out[4];
    HIPCHK(c, hipMemcpyAsync(out, c->t_out, sizeof out, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    // J = -C_L + beta C_D with C = force / (q_inf c), c = 1 (objective_on)
    const double corde = 1;
    *dJ = -(out[2] / (c->q_inf * corde)) + beta * (out[3] / (c->q_inf * corde));
    return O2F_OK;
}

}  // namespace

int o2f_naca_displacement(o2f_ctx *c, const char *name, const char *const *names_all, const double *values_all,
                          int n_all, double *dxy) {
    if (!c || !name || !dxy || (n_all > 0 && (!names_all || !values_all))) return O2F_ERR_ARG;
    const int which = naca_which(name);
    if (which < 0) return c->fail(O2F_ERR_ARG, std::string("not a NACA shape parameter: ") + name);
    const NacaParams np = naca_params(names_all, values_all, n_all);
    std::vector<double> d;
    naca_displacement(c->hm, which, np.pos_x, np.chord, np.m, np.p, d);
    std::copy(d.begin(), d.end(), dxy);
    return O2F_OK;
}

int o2f_residual_tangent(o2f_ctx *c, const double *dxy, int geometry, double beta, double *R, double *dR,
                         double *dJ) {
    if (!c || !dxy) return O2F_ERR_ARG;
    double dj = 0.0;
    int st = tangent_on(c, dxy, geometry, beta, &dj);
    if (st) return st;
    const size_t n = 4 * (size_t)c->hm.nc;
    if ((st = copy_out(c, R, c->R, n))) return st;
    if ((st = copy_out(c, dR, tan_work(c).dR, n))) return st;
    if (dJ) *dJ = dj;
    // leave the context as computeResidual(mesh) leaves it
    if ((st = residual_on(c, 0))) return st;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return O2F_OK;
}

int o2f_residual_jvp(o2f_ctx *c, const double *dU, double *R, double *dR) {
    if (!c || !dU) return O2F_ERR_ARG;
    int st = check_bc(c);
    if (st) return st;
    if ((st = ensure_work(c)) || (st = ensure_tangent(c))) return st;
    const size_t n = 4 * (size_t)c->hm.nc;
    double *du = c->kv[10];
    HIPCHK(c, hipMemcpyAsync(du, dU, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
    const TanWork W = tan_work(c);
    HIPCHK(c, launch_residual_tangent(c->topo(), c->geo[0].view(), Geom{}, c->phys, c->U, du, c->vU, c->fL, c->fR,
                                      c->tq, c->vtq, c->Fe, c->R, W, c->stream));
    ++c->faces_version;
    if ((st = copy_out(c, R, c->R, n))) return st;
    if ((st = copy_out(c, dR, W.dR, n))) return st;
    if ((st = residual_on(c, 0))) return st;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return O2F_OK;
}

int o2f_shape_gradient_tangent(o2f_ctx *c, const char *const *names_all, const double *values_all, int n_all,
                               const char *const *opt_names, int n_opt, const double *lambda, double beta,
                               int geometry, double *grad) {
    if (!c || (n_all > 0 && (!names_all || !values_all)) || (n_opt > 0 && (!opt_names || !grad)) || !lambda)
        return O2F_ERR_ARG;
    const NacaParams np = naca_params(names_all, values_all, n_all);
    int st = ensure_work(c);
    if (st) return st;
    const long n = 4L * c->hm.nc;
    double *lam = c->kv[0];
    HIPCHK(c, hipMemcpyAsync(lam, lambda, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
    std::vector<double> dxy;
    for (int k = 0; k < n_opt; ++k) {
        const int which = naca_which(opt_names[k]);
        if (which < 0) {
            grad[k] = 0.0;
            continue;
        }
        naca_displacement(c->hm, which, np.pos_x, np.chord, np.m, np.p, dxy);
        double dJ = 0.0;
        if ((st = tangent_on(c, dxy.data(), geometry, beta, &dJ))) return st;
        HIPCHK(c, launch_dot2(lam, tan_work(c).dR, nullptr, nullptr, n, c->partials, c->scal, kOpStore, c->stream));
        HostScal h;
        if ((st = read_scal(c, h))) return st;
        grad[k] = dJ - h.v[kScalDotA];
    }
    if ((st = residual_on(c, 0))) return st;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return O2F_OK;
}

int o2f_time_adjoint(o2f_ctx *c, int nprod, int exact, double *total_ms) {
    if (!c || nprod <= 0 || !total_ms) return O2F_ERR_ARG;
    int st = exact ? ensure_adjoint(c) : ensure_work(c);
    if (st) return st;
    if ((st = residual_on(c, 0))) return st;
    if (!exact && (st = ensure_blocks(c, 1e-6))) return st;
    const long n = 4L * c->hm.nc;
    double *x = c->kv[0], *y = c->kv[1];
    HIPCHK(c, launch_fill(n, x, 1.0, c->stream));
    const Topo T = c->topo();
    const Geom G = c->geo[0].view();
    const MatVec mv = exact ? exact_matvec(c) : MatVec([c, T, G](const double *a, double *b) {
        return launch_jt_apply(T, G, c->blk, a, b, c->stream);
    });
    HIPCHK(c, mv(x, y));  // warm-up
    hipEvent_t e0, e1;
    HIPCHK(c, hipEventCreate(&e0));
    HIPCHK(c, hipEventCreate(&e1));
    HIPCHK(c, hipEventRecord(e0, c->stream));
    for (int k = 0; k < nprod; ++k) HIPCHK(c, mv(k & 1 ? y : x, k & 1 ? x : y));
    HIPCHK(c, hipEventRecord(e1, c->stream));
    HIPCHK(c, hipEventSynchronize(e1));
    float ms = 0.f;
    HIPCHK(c, hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *total_ms = ms;
    return O2F_OK;
}

int o2f_time_kernel(o2f_ctx *c, int kernel, int launches, double *avg_us) {
    if (!c || kernel < 0 || kernel > 5 || launches <= 0 || !avg_us) return O2F_ERR_ARG;
    int st = check_bc(c);
    if (st) return st;
    if (kernel == 5) {
        if (!tiles_on(c)) return c->fail(O2F_ERR_ARG, "o2f_time_kernel: no tile stages (" + c->tiles_why + ")");
        HIPCHK(c, launch_state_to_aos(c->U, c->hm.nc, c->Ua, c->stream));
    }
    // every intermediate the kernels read, at the current state
    if ((st = residual_on(c, 0))) return st;
    hipStream_t s = c->stream;
    const Topo T = c->topo();
    const Geom G = c->geo[0].view();
    auto launch = [&]() -> hipError_t {
        switch (kernel) {
            case 0: return launch_vertex_state(T, G, c->U, c->vU, c->phys, s);
            case 1: return launch_reconstruct(T, G, c->U, c->vU, c->fL, c->fR, c->tq, nullptr, c->phys, s);
            case 2: return launch_visc_and_flux(T, G, c->tq, c->vtq, c->fL, c->fR, c->Fe, c->phys, 1, s);
            case 3:
                return launch_flux_update(T, G, c->U, c->Uold, c->Fe, c->vtq, c->R, nullptr, 0.0, 0.0,
                                          kFlagAdv | kFlagVisc | kFlagStoreR, c->phys, s);
            case 5:  // a second stage's work: U_old read, the new state to Ub
                return launch_tile_stage(T, G, c->phys, c->tv, c->Ua, c->Ub, true, c->Uold, c->R, c->scal + kScalDt,
                                         3.0 / 4.0, 1.0 / 4.0, kFlagRK, nullptr, s);
            default: {  // the whole residual, its four launches in order
                hipError_t e = launch_vertex_state(T, G, c->U, c->vU, c->phys, s);
                if (!e) e = launch_reconstruct(T, G, c->U, c->vU, c->fL, c->fR, c->tq, nullptr, c->phys, s);
                if (!e) e = launch_visc_and_flux(T, G, c->tq, c->vtq, c->fL, c->fR, c->Fe, c->phys, 1, s);
                if (!e)
                    e = launch_flux_update(T, G, c->U, c->Uold, c->Fe, c->vtq, c->R, nullptr, 0.0, 0.0,
                                           kFlagAdv | kFlagVisc | kFlagStoreR, c->phys, s);
                return e;
            }
        }
    };
    HIPCHK(c, launch());  // warm-up
    hipEvent_t e0, e1;
    HIPCHK(c, hipEventCreate(&e0));
    HIPCHK(c, hipEventCreate(&e1));
    HIPCHK(c, hipEventRecord(e0, s));
    for (int k = 0; k < launches; ++k) HIPCHK(c, launch());
    HIPCHK(c, hipEventRecord(e1, s));
    HIPCHK(c, hipEventSynchronize(e1));
    float ms = 0.f;
    HIPCHK(c, hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *avg_us = 1e3 * ms / launches;
    ++c->faces_version;
    return O2F_OK;
}

int o2f_set_tile_cells(o2f_ctx *c, int cells) {
    if (!c || cells < -1) return O2F_ERR_ARG;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    release_tiles(c);
    c->tile_cells = cells;
    c->tiles_unusable = false;
    c->tiles_why.clear();
    return O2F_OK;
}

int o2f_tile_info(o2f_ctx *c, int *n_tiles, int *cells, int *lds_bytes, int *tiles_per_cu, int *vgprs) {
    if (!c) return O2F_ERR_ARG;
    const bool on = tiles_on(c);
    if (!on && c->tile_cells != 0) c->err = "four-kernel stages: " + c->tiles_why;
    if (n_tiles) *n_tiles = on ? c->tv.ntiles : 0;
    if (cells) *cells = on ? c->tile_target : 0;
    if (lds_bytes) *lds_bytes = on ? (int)c->tv.lds_bytes : 0;
    if (tiles_per_cu) *tiles_per_cu = on ? c->tile_bpc : 0;
    if (vgprs) *vgprs = on ? c->tile_vgprs : 0;
    return O2F_OK;
}

int o2f_time_steps(o2f_ctx *c, int nsteps, double *total_ms, double *kernel_ms) {
    if (!c || nsteps <= 0 || !total_ms) return O2F_ERR_ARG;
    int st = check_bc(c);
    if (st) return st;
    hipStream_t s = c->stream;
    hipEvent_t e0, e1;
    HIPCHK(c, hipEventCreate(&e0));
    HIPCHK(c, hipEventCreate(&e1));
    HIPCHK(c, hipStreamSynchronize(s));
    HIPCHK(c, hipEventRecord(e0, s));
    for (int k = 0; k < nsteps && st == 0; ++k)
        st = one_step(c, k == nsteps - 1, nullptr, (k > 0 ? kChainUse : 0) | kChainForm);
    const int st_leave = tiles_leave(c);
    if (st) return st;
    if (st_leave) return st_leave;
    HIPCHK(c, hipEventRecord(e1, s));
    HIPCHK(c, hipEventSynchronize(e1));
    float ms = 0.f;
    HIPCHK(c, hipEventElapsedTime(&ms, e0, e1));
    *total_ms = ms;
    if (kernel_ms) {
        // Second pass: every stage kernel bracketed by its own events.
        const int nk = 4;
        std::vector<hipEvent_t> ev(2 * nk * 3 * nsteps);
        for (auto &e : ev) HIPCHK(c, hipEventCreate(&e));
        const Topo T = c->topo();
        const Geom G = c->geo[0].view();
        const double *dtp = c->scal + kScalDt;
        const double cs[3][2] = {{0.0, 1.0}, {3.0 / 4.0, 1.0 / 4.0}, {1.0 / 3.0, 2.0 / 3.0}};
        int q = 0;
        for (int k = 0; k < nsteps; ++k) {
            if (!c->speed_valid) {
                HIPCHK(c, launch_speed_max(c->U, c->hm.nc, c->phys, c->speed, s));
                c->speed_nparts = speed_parts(c->hm.nc);
                c->speed_valid = true;
            }
            HIPCHK(c, launch_step_begin(c->speed, c->speed_nparts, c->phys, c->hm.min_h, c->scal, nullptr, 1, s));
            for (int stg = 0; stg < 3; ++stg) {
                const int fl = kFlagAdv | kFlagVisc | kFlagRK | (stg == 0 ? kFlagSaveOld : 0) |
                               (stg == 2 ? kFlagSpeed : 0);
                HIPCHK(c, hipEventRecord(ev[q++], s));
                HIPCHK(c, launch_vertex_state(T, G, c->U, c->vU, c->phys, s));
                HIPCHK(c, hipEventRecord(ev[q++], s));
                HIPCHK(c, hipEventRecord(ev[q++], s));
                HIPCHK(c, launch_reconstruct(T, G, c->U, c->vU, c->fL, c->fR, c->tq, nullptr, c->phys, s));
                HIPCHK(c, hipEventRecord(ev[q++], s));
                HIPCHK(c, hipEventRecord(ev[q++], s));
                HIPCHK(c, launch_visc_and_flux(T, G, c->tq, c->vtq, c->fL, c->fR, c->Fe, c->phys, 1, s));
                HIPCHK(c, hipEventRecord(ev[q++], s));
                HIPCHK(c, hipEventRecord(ev[q++], s));
                HIPCHK(c, launch_flux_update(T, G, c->U, c->Uold, c->Fe, c->vtq, c->R, dtp, cs[stg][0],
                                             cs[stg][1], fl, c->phys, s, c->speed));
                HIPCHK(c, hipEventRecord(ev[q++], s));
            }
            c->speed_nparts = speed_parts(c->hm.nc);
        }
        HIPCHK(c, hipStreamSynchronize(s));
        for (int j = 0; j < nk; ++j) kernel_ms[j] = 0.0;
        for (int j = 0; j < q; j += 2) {
            float m = 0.f;
            HIPCHK(c, hipEventElapsedTime(&m, ev[j], ev[j + 1]));
            kernel_ms[(j / 2) % nk] += m;
        }
        for (auto &e : ev) (void)hipEventDestroy(e);
        ++c->faces_version;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return O2F_OK;
}

}  // extern "C"

// The device memory a BiCGSTAB call allocates on first use (work vectors,
// FD blocks; for the exact operator its tables and work arrays), allocated
// now: o2f_batch_bicgstab prepares every member on one thread before the
// members' solves run concurrently.
int o2f_ctx_adjoint_prepare(o2f_ctx *c, int exact) { return exact ? ensure_adjoint(c) : ensure_work(c); }
