// Device view and launcher of the fused tile stage (o2f_tiles.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "host_tiles.hpp"
#include "o2f_kernels.hpp"

namespace o2f {

constexpr int kTileThreads = 256;       // one workgroup per tile
constexpr int kTileQuadThreads = 1024;  // the four-lane form, for tiles that run one per CU

// A TileSet uploaded to the device, with its LDS layout (offsets in doubles).
struct TileView {
    const TileHead *head;
    const TileCell *cell;
    const TileOwn *own;
    const TileGather *vg;  // state vertices
    const TileStress *vo;  // stress vertices
    const double2 *en;     // unit normal of each tile edge
    const uint16_t *lsten;
    const double *epk;  // packed edge geometry (launch_pack_edges), 8 doubles per edge
    int ntiles;
    int nk_max;  // most edges of a cell: 3 (triangles only) selects the three-slot kernels
    int max_vo, max_gd;
    int lds_tq, lds_fs;
    size_t lds_bytes;
    int quad;  // 1: the 1,024-thread four-lane kernels (tiles one per CU)
};

// The start of a step as a tile stage forms it (kFlagBegin): the speed
// partials of the step's start state, min_h, the scalar block and the dt
// history slot (may be NULL).
struct StepBegin {
    const double *parts;
    int nparts;
    double min_h;
    double *scal;
    double *hist;
};

// One residual stage fused with its Shu-Osher update over every tile:
// Uout = c0 Uold + c1 (U + dt R(U)) for each owned cell (flags as for
// k_cell_flux_update; kFlagAdv and kFlagVisc are implied).  U is AoS (four
// doubles per cell); Uout is AoS (out_aos) or SoA [q][nc]; Uold and R are
// SoA.  U and Uout must be different buffers.  kFlagSpeed writes one
// max(|u| + c) partial per tile.
// With the 256-thread kernels a step may carry the next step's speed max:
// its second stage starts the fold (kFlagSpeedReset), its last stage folds
// every tile's partial into scal[kScalSpeedMax] (kFlagSpeedMax, with
// kFlagSpeed), and the next step's first stage starts that step from it
// (kFlagBeginDt: the stop test, dt as k_step_begin forms it), its second
// stage advancing the time (kFlagAdvance).
// With the four-lane kernels (A.quad) the first stage may also start the
// step (kFlagBegin: every workgroup forms the step's dt from B's partials
// as k_step_begin does, workgroup 0 stores it, and the time loop's stop
// test closes the gate instead of the stage) and the second stage advance
// the time (kFlagAdvance: t_prev = t, t += dt, steps + 1 under a gate), in
// place of the one-block k_step_begin launch.
hipError_t launch_tile_stage(const Topo &T, const Geom &G, const Phys &P, const TileView &A, const double *U,
                             double *Uout, bool out_aos, double *Uold, double *R, const double *dtp, double c0,
                             double c1, int flags, double *speed, hipStream_t s, const StepBegin &B = StepBegin{});
// One stage over the concatenated tiles of a batch's members (each head
// names its member): member m's tiles read M[m].Us[stage] and write
// M[m].Us[stage + 1 mod 3] (AoS), with M[m]'s Uold, R, dt and speed
// partials (one per tile of the member, in its slot) as launch_tile_stage.
hipError_t launch_b_tile_stage(const Member *M, const TileView &A, int stage, double c0, double c1, int flags,
                               hipStream_t s);
// The speed partials of every member's SoA state M[m].U, one per tile in
// the slots the stages' kFlagSpeed fills.
hipError_t launch_b_tile_speed(const Member *M, const TileView &A, hipStream_t s);
// Packed edge geometry of the tile stage: per edge (n_x, n_y, L, tag bits,
// mid_x, mid_y, 0, 0), 8 doubles, from the geometry set and the edge tags.
hipError_t launch_pack_edges(const Geom &G, const char *ebc, int ne, double *epk, hipStream_t s);
// The state between the C ABI's SoA [q][nc] and the stages' AoS.
hipError_t launch_state_to_aos(const double *U, int nc, double *A, hipStream_t s);
hipError_t launch_state_to_soa(const double *A, int nc, double *U, hipStream_t s);
// Allows the device's full LDS for the stage kernels of cells of at most
// nk_max edges and reports how many of their 256-thread workgroups fit on
// one CU with lds_bytes each (0: they do not fit), their VGPRs and how many
// workgroups of the 1,024-thread four-lane form fit (quad_per_cu, may be
// NULL).
hipError_t tile_stage_attributes(size_t lds_bytes, int nk_max, int *blocks_per_cu, int *vgprs,
                                 int *quad_per_cu = nullptr);

}  // namespace o2f
