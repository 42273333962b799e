// Internal: the context of one mesh (device arrays, stream, host mesh),
// shared by the C ABI (o2f_capi.cpp) and the batch (o2f_batch.cpp).
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "host_mesh.hpp"
#include "o2f.h"
#include "o2f_kernels.hpp"
#include "o2f_tiles.hpp"

namespace o2f {

template <class T>
hipError_t dalloc(T *&p, size_t n) {
    p = nullptr;
    return hipMalloc(reinterpret_cast<void **>(&p), (n ? n : 1) * sizeof(T));
}

template <class T>
hipError_t upload(T *&p, const std::vector<T> &v) {
    hipError_t e = dalloc(p, v.size());
    if (e != hipSuccess) return e;
    if (v.empty()) return hipSuccess;
    return hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice);
}

struct GeomSet {
    double *cvol = nullptr, *ccx = nullptr, *ccy = nullptr;
    double *elen = nullptr, *enx = nullptr, *eny = nullptr, *emx = nullptr, *emy = nullptr;
    double *sw = nullptr;
    Geom view() const { return Geom{cvol, ccx, ccy, elen, enx, eny, emx, emy, sw}; }
    GeomW writable() const { return GeomW{cvol, ccx, ccy, elen, enx, eny, emx, emy, sw}; }
    hipError_t alloc(const HostMesh &hm) {
        hipError_t e;
        if ((e = dalloc(cvol, hm.nc)) || (e = dalloc(ccx, hm.nc)) || (e = dalloc(ccy, hm.nc)) ||
            (e = dalloc(elen, hm.ne)) || (e = dalloc(enx, hm.ne)) || (e = dalloc(eny, hm.ne)) ||
            (e = dalloc(emx, hm.ne)) || (e = dalloc(emy, hm.ne)) || (e = dalloc(sw, hm.sw.size() + kStencilSlots)))
            return e;
        return hipSuccess;
    }
    // M: a HostMesh, or a GeomTangent (the same arrays holding derivatives)
    template <class M>
    hipError_t fill(const M &hm, hipStream_t s) {
        auto cp = [&](double *d, const std::vector<double> &v) {
            return v.empty() ? hipSuccess
                             : hipMemcpyAsync(d, v.data(), v.size() * sizeof(double), hipMemcpyHostToDevice, s);
        };
        hipError_t e;
        if ((e = cp(cvol, hm.cvol)) || (e = cp(ccx, hm.ccx)) || (e = cp(ccy, hm.ccy)) || (e = cp(elen, hm.elen)) ||
            (e = cp(enx, hm.enx)) || (e = cp(eny, hm.eny)) || (e = cp(emx, hm.emx)) || (e = cp(emy, hm.emy)) ||
            (e = cp(sw, hm.sw)))
            return e;
        return hipStreamSynchronize(s);
    }
    void release() {
        for (double *p : {cvol, ccx, ccy, elen, enx, eny, emx, emy, sw})
            if (p) (void)hipFree(p);
        *this = GeomSet();
    }
};


}  // namespace o2f

struct o2f_ctx {
    o2f::HostMesh hm;
    o2f_params prm{};
    o2f::Phys phys{};
    double q_inf = 0.0;  // 0.5 rho_inf U_inf^2 (src/compute.cpp:96-100)
    hipStream_t stream = nullptr;
    std::string err;
    // topology
    int *cedge = nullptr, *cnbr = nullptr, *cevx = nullptr, *eva = nullptr, *evb = nullptr, *ecl = nullptr,
        *ecr = nullptr;
    int *sptr = nullptr, *scell = nullptr, *wall = nullptr;
    // wall edges grouped by their (left) cell, ascending edge id within a
    // group, and the groups' offsets: the adjoint RHS adds each cell's wall
    // edges in wall order with one thread per cell
    int *wall_by_cell = nullptr, *wall_grp = nullptr;
    int wall_ngrp = 0;
    uint8_t *cinfo = nullptr, *skind = nullptr;
    char *ebc = nullptr, *vbc = nullptr;
    o2f::GeomSet geo[3];  // 0: the mesh, 1/2: +/- shape perturbations
    // state and per-stage intermediates
    double *U = nullptr, *Uold = nullptr, *R = nullptr, *Gr = nullptr, *tq = nullptr;
    double *vU = nullptr, *vtq = nullptr, *fL = nullptr, *fR = nullptr, *Fe = nullptr;
    double *scal = nullptr, *partials = nullptr, *wall_out = nullptr;
    double *speed = nullptr;              // per-block max_c(|u|+c) partials of the next step's dt
    bool speed_valid = false;             // it holds the speed of the current U
    int speed_nparts = 0;                 // partials the last speed fold wrote
    // fused tile stages (host_tiles.hpp, o2f_tiles.hip), built on first use:
    // tile_cells -1 sizes the tiles from the kernel's occupancy, 0 keeps the
    // four-kernel stages, > 0 caps a tile at that many cells
    int tile_cells = -1;
    bool tiles_ready = false;     // tables built for the current mesh and stencils
    bool tiles_unusable = false;  // the mesh cannot be tiled (tiles_why): four-kernel stages
    std::string tiles_why;
    o2f::TileView tv{};
    void *tile_mem = nullptr;  // every table, one allocation
    // AoS states of a tiled run: Ua the state (while tile_aos), Ub / Uc the
    // outputs of a step's first two stages; c->U is brought up to date
    // before any entry point returns
    double *Ua = nullptr, *Ub = nullptr, *Uc = nullptr;
    bool tile_aos = false;
    double *epk = nullptr;  // packed edge geometry of geo[0] for the tile stages
    double *rparts = nullptr;  // density-residual partials of a tiled o2f_run_to_residual
    int tile_target = 0, tile_bpc = 0, tile_vgprs = 0;
    long stencil_version = 0;  // counts changes of the interpolation stencils (tiles built on them are stale)
    bool have_grad = false;
    // adjoint
    double *blk = nullptr;
    long faces_version = 0, blk_version = -1;
    double blk_eps = 0.0;
    // epsilon of the BiCGSTAB restart test |r0.r| < eps^2 |r0|^2 (Eigen's
    // NumTraits<double>::epsilon(), settable for tests)
    double bicg_restart_eps = 2.220446049250313e-16;
    double *kv[12] = {};  // BiCGSTAB / gradient work vectors (4 nc each)
    // block-Jacobi preconditioner: cell colours (distance two in vertex
    // rings), probed diagonal blocks and their inverses (16 per cell)
    int *pc_color = nullptr;
    int pc_ncolors = 0;
    double *pc_D = nullptr, *pc_P = nullptr;
    double *Rp = nullptr, *Rm = nullptr;
    // exact adjoint: vertex -> edge CSR, cell vertices and their stencil
    // weights, work arrays (allocated on first use)
    int *a_veptr = nullptr, *a_vedge = nullptr, *a_cvtx = nullptr, *a_mtgt = nullptr;
    double *a_cvw = nullptr, *a_work = nullptr;
    int a_interp = -1;  // interp mode a_cvw was built for
    // geometric tangent: derivative arrays of the residual's intermediates
    // (allocated on first use) and its 4-double objective output
    double *t_work = nullptr, *t_out = nullptr;
    // staging of the sparse geometry tangent (indices, values; capacities)
    int *ts_i = nullptr;
    double *ts_d = nullptr;
    size_t ts_ni = 0, ts_nd = 0;

    int fail(int code, const std::string &msg) {
        err = msg;
        return code;
    }
    o2f::Topo topo() const {
        return o2f::Topo{hm.nv, hm.ne, hm.nc, cedge, cnbr, cevx, cinfo, eva, evb, ecl, ecr, ebc, sptr, scell, skind, vbc};
    }
};

#define HIPCHK(ctx, expr)                                                                       \
    do {                                                                                        \
        hipError_t e_ = (expr);                                                                 \
        if (e_ != hipSuccess) return (ctx)->fail(O2F_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)


// apply_boundary_conditions' check (src/boundary.cpp:84-92) for context c.
int o2f_ctx_check_bc(o2f_ctx *c);
// The four kernels of a residual evaluation of the SoA state Uin on the
// context's stream, leaving the intermediates a stage of the four-kernel
// path leaves (face states, stresses, vertex values; R and U untouched).
int o2f_ctx_form_outputs(o2f_ctx *c, double *Uin);
// The device allocations of an adjoint solve (exact: the exact operator's
// too), made now rather than inside the solve.
int o2f_ctx_adjoint_prepare(o2f_ctx *c, int exact);
// o2f_run_to_residual on the batched four-kernel stages (o2f_batch.cpp).
int run_to_residual_batch(o2f_ctx *c, double tol, int max_steps, int check_every, int use_graph, int *steps,
                          double *res);
