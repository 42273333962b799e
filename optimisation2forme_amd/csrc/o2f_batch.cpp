// Batches of meshes advanced together, and the run-to-residual loop.
//
// A batch ties several contexts.  Each stage kernel is one launch over the
// tiles of every member (batched kernels in o2f_kernels.hip); a group of
// check_every steps can be captured once as a hipGraph and replayed.  The
// launch shape of each kernel (threads per workgroup, persistent grid) is a
// free choice: every member's per-entity arithmetic and sum orders are those
// of its single-mesh launch, the dt maximum is order-free and the residual
// partials use fixed cell chunks, so results do not depend on it.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <exception>
#include <map>
#include <thread>
#include <utility>
#include <string>
#include <vector>

#include "host_tiles.hpp"
#include "o2f.h"
#include "o2f_ctx.hpp"
#include "o2f_kernels.hpp"
#include "o2f_tiles.hpp"

using namespace o2f;

namespace {

enum { kKVertex = 0, kKRecon = 1, kKViscEdge = 2, kKFlux = 3, kKinds = 4 };
// host threads of o2f_batch_bicgstab: each drives one member's solve at a
// time (its loop waits on its stream once per group of iterations); eight
// keep the GPU's queues busy (measured: 4 threads 2.56x, 8 threads 2.74x,
// 16 threads 2.72x the sequential solves of the 64-candidate batch)
constexpr int kAdjointThreads = 8;
constexpr int kMinBlock = 64;

long ceil_div(long a, long b) { return (a + b - 1) / b; }

long blocks_of(int kind, const HostMesh &hm, int block) {
    switch (kind) {
        case kKVertex: return ceil_div(hm.nv, block);
        case kKRecon:
        case kKFlux: return ceil_div(4L * hm.nc, block);
        default: return ceil_div(hm.nv, block) + ceil_div(hm.ne, block);
    }
}

bool valid_shape(const TileShape &t) {
    return (t.block == 64 || t.block == 128 || t.block == 256) && t.grid >= 0;
}

}  // namespace

struct o2f_batch {
    std::vector<o2f_ctx *> ctx;
    std::string err;
    hipStream_t stream = nullptr;
    std::vector<Member> hM;
    Member *dM = nullptr;
    std::vector<double *> speed, rparts;
    TileShape shape[kKinds] = {{256, 0}, {256, 0}, {256, 0}, {256, 0}};
    int2 *tiles[kKinds] = {};
    int ntiles[kKinds] = {};
    int2 *rtiles = nullptr;
    int nrtiles = 0;
    double *dflags = nullptr, *hflags = nullptr;  // device / pinned host copies of the active flags
    std::vector<double> scal_init;                // host source of the per-run scalar reset
    // captured step chunks, keyed by (steps, check on the first step)
    std::map<std::pair<int, bool>, std::pair<hipGraph_t, hipGraphExec_t>> graphs;
    // fused tile stages over every member's tiles (tile_cells: -1 automatic,
    // 0 the four-kernel stages, n > 0 at most n cells per tile)
    int tile_cells = -1;
    int tile_tuned = 0;  // the automatic size o2f_batch_autotune timed fastest (0: the rule's)
    bool tiles_ready = false, tiles_unusable = false;
    std::string tiles_why;
    TileView tv{};
    void *tile_mem = nullptr;
    std::vector<double *> epk, Us;  // per member: packed edges; three AoS states (3 per member)
    std::vector<int> mtiles;        // tiles of each member
    std::vector<long> mversion;     // each member's stencil_version the tiles were built on
    int tile_target = 0, tile_bpc = 0, tile_vgprs = 0;

    int n() const { return static_cast<int>(ctx.size()); }
    int fail(int code, const std::string &m) {
        err = m;
        return code;
    }
    void drop_graph() {
        for (auto &kv : graphs) {
            (void)hipGraphExecDestroy(kv.second.second);
            (void)hipGraphDestroy(kv.second.first);
        }
        graphs.clear();
    }
};

#define BCHK(b, expr)                                                                           \
    do {                                                                                        \
        hipError_t e_ = (expr);                                                                 \
        if (e_ != hipSuccess) return (b)->fail(O2F_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

namespace {

void release_tiles(o2f_batch *b) {
    if (b->tile_mem) (void)hipFree(b->tile_mem);
    b->tile_mem = nullptr;
    for (double *p : b->epk) (void)hipFree(p);
    for (double *p : b->Us) (void)hipFree(p);
    b->epk.clear();
    b->Us.clear();
    b->mtiles.clear();
    b->mversion.clear();
    b->tv = TileView{};
    b->tiles_ready = false;
}

// Whether the steps run on the fused tile stages (built on first use).  One
// launch covers the tiles of every member, in many rounds of workgroups, so
// the automatic size is the largest whose LDS lets a CU hold as many tiles
// as the kernel's registers allow (about 200 cells); members' tables are
// concatenated.  Any member that cannot be tiled keeps the whole batch on
// the four-kernel stages (tiles_why says why).
bool tiles_on(o2f_batch *b) {
    if (b->tile_cells == 0 || b->tiles_unusable) return false;
    if (b->tiles_ready) {
        bool current = true;
        for (int m = 0; m < b->n(); ++m) current = current && b->mversion[m] == b->ctx[m]->stencil_version;
        if (current) return true;
    }
    b->drop_graph();
    release_tiles(b);
    auto unusable = [&](const std::string &why) {
        release_tiles(b);
        b->tiles_unusable = true;
        b->tiles_why = why;
        return false;
    };
    int nk_max = 0;
    for (o2f_ctx *c : b->ctx) nk_max = std::max(nk_max, *std::max_element(c->hm.cnv.begin(), c->hm.cnv.end()));
    std::vector<TileSet> sets(b->n());
    TileSet all;
    std::string err;
    int bpc = 0, vgprs = 0, most = 0;
    // tiles one CU can hold at the kernel's registers (a tiny LDS)
    if (tile_stage_attributes(1024, nk_max, &most, nullptr) != hipSuccess || most < 1)
        return unusable("tile kernel attributes");
    // the members' tables are independent: built on up to 16 host threads
    const int nthreads = std::max(1, std::min({16, b->n(), (int)std::thread::hardware_concurrency()}));
    auto build = [&](int t) -> bool {
        std::vector<std::string> errs(b->n());
        std::vector<char> ok(b->n(), 0);
        std::atomic<int> next{0};
        auto work = [&] {
            for (int m; (m = next++) < b->n();) ok[m] = build_tiles(b->ctx[m]->hm, t, sets[m], errs[m]);
        };
        std::vector<std::thread> pool;
        try {
            for (int i = 1; i < nthreads; ++i) pool.emplace_back(work);
        } catch (const std::exception &) {  // no more threads: the rest on this one
        }
        work();
        for (std::thread &th : pool) th.join();
        for (int m = 0; m < b->n(); ++m)
            if (!ok[m]) {
                err = errs[m];
                return false;
            }
        concat_tiles(sets, all);
        if (!check_concat(sets, all, err)) return false;
        const hipError_t e = tile_stage_attributes(all.lds_bytes(), nk_max, &bpc, &vgprs);
        if (e != hipSuccess) err = std::string("tile kernel attributes: ") + hipGetErrorString(e);
        return e == hipSuccess && bpc > 0;
    };
    int target = (b->tile_cells < 0 && b->tile_tuned > 0) ? b->tile_tuned : b->tile_cells;
    if (target > 0) {
        if (!build(target)) return unusable(err.empty() ? "tile LDS too large" : err);
    } else {
        constexpr int kFirst = 208, kMin = 64;
        bool ok = false;
        for (target = kFirst; target >= kMin; target = target * 24 / 25) {
            ok = build(target);
            if (!ok && !err.empty()) return unusable(err);
            if (ok && bpc >= most) break;
        }
        if (!ok) return unusable("no tile size fits");
    }
    // one allocation for the tables
    size_t off = 0;
    auto slot = [&](size_t bytes) {
        const size_t o = off;
        off += (bytes + 255) & ~size_t(255);
        return o;
    };
    const size_t o_head = slot(all.head.size() * sizeof(TileHead)), o_cell = slot(all.cell.size() * sizeof(TileCell)),
                 o_own = slot(all.own.size() * sizeof(TileOwn)), o_vg = slot(all.vgf.size() * sizeof(TileGather)),
                 o_edge = slot(all.en.size() * sizeof(double)), o_vo = slot(all.vof.size() * sizeof(TileStress)),
                 o_ls = slot(all.lsten.size() * sizeof(uint16_t) + 16);
    char *mem = nullptr;
    hipError_t e = hipMalloc(reinterpret_cast<void **>(&mem), off);
    if (e == hipSuccess) b->tile_mem = mem;
    auto put = [&](size_t o, const void *src, size_t bytes) {
        if (e == hipSuccess && bytes) e = hipMemcpy(mem + o, src, bytes, hipMemcpyHostToDevice);
    };
    put(o_head, all.head.data(), all.head.size() * sizeof(TileHead));
    put(o_cell, all.cell.data(), all.cell.size() * sizeof(TileCell));
    put(o_own, all.own.data(), all.own.size() * sizeof(TileOwn));
    put(o_vg, all.vgf.data(), all.vgf.size() * sizeof(TileGather));
    put(o_edge, all.en.data(), all.en.size() * sizeof(double));
    put(o_vo, all.vof.data(), all.vof.size() * sizeof(TileStress));
    put(o_ls, all.lsten.data(), all.lsten.size() * sizeof(uint16_t));
    for (int m = 0; m < b->n() && e == hipSuccess; ++m) {
        o2f_ctx *c = b->ctx[m];
        const int nc = c->hm.nc;
        b->mtiles.push_back((int)sets[m].head.size());
        b->mversion.push_back(c->stencil_version);
        double *p = nullptr;
        e = dalloc(p, 8 * (size_t)c->hm.ne);
        b->epk.push_back(p);
        if (e == hipSuccess) e = launch_pack_edges(c->geo[0].view(), c->ebc, c->hm.ne, p, b->stream);
        for (int k = 0; k < 3 && e == hipSuccess; ++k) {
            p = nullptr;
            e = dalloc(p, 4 * (size_t)nc);
            b->Us.push_back(p);
        }
        // one speed partial per tile of the member
        if (e == hipSuccess && b->mtiles[m] > blocks_of(kKFlux, c->hm, kMinBlock)) {
            (void)hipFree(b->speed[m]);
            b->speed[m] = nullptr;
            e = dalloc(b->speed[m], b->mtiles[m]);
        }
    }
    if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
    if (e != hipSuccess) return unusable(std::string("tile tables: ") + hipGetErrorString(e));
    TileView &v = b->tv;
    v.head = reinterpret_cast<const TileHead *>(mem + o_head);
    v.cell = reinterpret_cast<const TileCell *>(mem + o_cell);
    v.own = reinterpret_cast<const TileOwn *>(mem + o_own);
    v.vg = reinterpret_cast<const TileGather *>(mem + o_vg);
    v.en = reinterpret_cast<const double2 *>(mem + o_edge);
    v.vo = reinterpret_cast<const TileStress *>(mem + o_vo);
    v.lsten = reinterpret_cast<const uint16_t *>(mem + o_ls);
    v.epk = nullptr;  // per member (Member::epk)
    v.ntiles = (int)all.head.size();
    v.nk_max = nk_max;
    v.max_vo = all.max_vo;
    v.max_gd = all.max_gd;
    v.lds_bytes = all.lds_layout(&v.lds_tq, &v.lds_fs);
    v.quad = 0;  // many rounds of tiles: the 256-thread kernel
    b->tile_target = target;
    b->tile_bpc = bpc;
    b->tile_vgprs = vgprs;
    b->tiles_ready = true;
    b->drop_graph();
    return true;
}

int build_tiles(o2f_batch *b, int kind) {
    std::vector<int2> t;
    for (int m = 0; m < b->n(); ++m) {
        const long nb = blocks_of(kind, b->ctx[m]->hm, b->shape[kind].block);
        for (long k = 0; k < nb; ++k) t.push_back(make_int2(m, static_cast<int>(k)));
    }
    if (b->tiles[kind]) (void)hipFree(b->tiles[kind]);
    b->tiles[kind] = nullptr;
    BCHK(b, dalloc(b->tiles[kind], t.size()));
    BCHK(b, hipMemcpy(b->tiles[kind], t.data(), t.size() * sizeof(int2), hipMemcpyHostToDevice));
    b->ntiles[kind] = static_cast<int>(t.size());
    return O2F_OK;
}

// Member table from the contexts' current arrays and parameters.
// final_t: every member's time loop stops once its t >= final_t
// (while (t < final_time), src/simulation.cpp:75); HUGE_VAL for none.
int upload_members(o2f_batch *b, double tol, double final_t = HUGE_VAL) {
    b->hM.resize(b->n());
    for (int m = 0; m < b->n(); ++m) {
        o2f_ctx *c = b->ctx[m];
        Member &A = b->hM[m];
        A.T = c->topo();
        A.G = c->geo[0].view();
        A.P = c->phys;
        A.P.gate = c->scal + kScalActive;
        A.P.final_t = final_t;
        A.U = c->U;
        A.Uold = c->Uold;
        A.R = c->R;
        A.tq = c->tq;
        A.vU = c->vU;
        A.vtq = c->vtq;
        A.fL = c->fL;
        A.fR = c->fR;
        A.Fe = c->Fe;
        A.scal = c->scal;
        A.speed = b->speed[m];
        A.rparts = b->rparts[m];
        A.min_h = c->hm.min_h;
        const double a_inf = std::sqrt(c->prm.gamma * c->prm.p_ref / c->prm.rho_ref);
        A.res_scale = c->prm.rho_ref * (c->prm.inflow_velocity * a_inf) / 1.0;
        A.res_tol = tol > 0.0 ? tol : -1.0;
        A.nparts = static_cast<int>(blocks_of(kKFlux, c->hm, b->shape[kKFlux].block));
        A.nrparts = static_cast<int>(ceil_div(c->hm.nc, kResChunk));
        A.epk = nullptr;
        A.Us[0] = A.Us[1] = A.Us[2] = nullptr;
        if (b->tiles_ready) {
            A.nparts = b->mtiles[m];
            A.epk = b->epk[m];
            for (int k = 0; k < 3; ++k) A.Us[k] = b->Us[3 * m + k];
        }
    }
    BCHK(b, hipMemcpyAsync(b->dM, b->hM.data(), b->hM.size() * sizeof(Member), hipMemcpyHostToDevice, b->stream));
    BCHK(b, hipStreamSynchronize(b->stream));
    return O2F_OK;
}

// Active = 1, Steps = 0, t = 0, Conv = 0, Res = -1 for every member.
int reset_scalars(o2f_batch *b) {
    b->scal_init.assign(kScalCount, 0.0);
    b->scal_init[kScalActive] = 1.0;
    b->scal_init[kScalRes] = -1.0;
    for (o2f_ctx *c : b->ctx) {
        BCHK(b, hipMemcpyAsync(c->scal + kScalT, &b->scal_init[kScalT], sizeof(double), hipMemcpyHostToDevice,
                               b->stream));
        BCHK(b, hipMemcpyAsync(c->scal + kScalActive, &b->scal_init[kScalActive],
                               (kScalRes - kScalActive + 1) * sizeof(double), hipMemcpyHostToDevice, b->stream));
    }
    BCHK(b, hipStreamSynchronize(b->stream));
    return O2F_OK;
}

const double kC0[3] = {0.0, 3.0 / 4.0, 1.0 / 3.0};
const double kC1[3] = {1.0, 1.0 / 4.0, 2.0 / 3.0};

// Residual-only pass (no state change) or one SSP-RK3 stage.
int enqueue_stage(o2f_batch *b, int flags, double c0, double c1, hipEvent_t *ev = nullptr, int timed_kind = -1) {
    hipStream_t s = b->stream;
    auto mark = [&](int kind, int which) -> hipError_t {
        return (ev && kind == timed_kind) ? hipEventRecord(ev[which], s) : hipSuccess;
    };
    BCHK(b, mark(kKVertex, 0));
    BCHK(b, launch_b_vertex_gather(b->dM, b->tiles[kKVertex], b->ntiles[kKVertex], b->shape[kKVertex], s));
    BCHK(b, mark(kKVertex, 1));
    BCHK(b, mark(kKRecon, 0));
    BCHK(b, launch_b_reconstruct(b->dM, b->tiles[kKRecon], b->ntiles[kKRecon], b->shape[kKRecon], s));
    BCHK(b, mark(kKRecon, 1));
    BCHK(b, mark(kKViscEdge, 0));
    BCHK(b, launch_b_visc_edge(b->dM, b->tiles[kKViscEdge], b->ntiles[kKViscEdge], b->shape[kKViscEdge], s));
    BCHK(b, mark(kKViscEdge, 1));
    BCHK(b, mark(kKFlux, 0));
    BCHK(b, launch_b_flux_update(b->dM, b->tiles[kKFlux], b->ntiles[kKFlux], b->shape[kKFlux], c0, c1,
                                 flags | kFlagAdv | kFlagVisc, s));
    BCHK(b, mark(kKFlux, 1));
    return O2F_OK;
}

// One step: every member's dt, then the three stages (on the tiles, the
// state cycles Us[0] -> Us[1] -> Us[2] -> Us[0]).
int enqueue_step(o2f_batch *b, bool check) {
    int st;
    BCHK(b, launch_b_step_begin(b->dM, b->n(), b->stream));
    for (int stg = 0; stg < 3; ++stg) {
        const int flags = kFlagRK | (stg == 0 ? kFlagSaveOld : 0) | (stg == 0 && check ? kFlagStoreR : 0) |
                          (stg == 2 ? kFlagSpeed : 0);
        if (b->tiles_ready)
            BCHK(b, launch_b_tile_stage(b->dM, b->tv, stg, kC0[stg], kC1[stg], flags, b->stream));
        else if ((st = enqueue_stage(b, flags, kC0[stg], kC1[stg])))
            return st;
        if (stg == 0 && check) BCHK(b, launch_b_residual(b->dM, b->rtiles, b->nrtiles, b->n(), b->stream));
    }
    return O2F_OK;
}

int enqueue_group(o2f_batch *b, int k, bool check) {
    int st;
    for (int j = 0; j < k; ++j)
        if ((st = enqueue_step(b, check && j == 0))) return st;
    return O2F_OK;
}

// At most this many steps are captured in one graph (a step is 14 kernel
// nodes); longer groups replay several chunks.
constexpr int kGraphSteps = 16;

int get_graph(o2f_batch *b, int k, bool check, hipGraphExec_t *exec) {
    const auto key = std::make_pair(k, check);
    auto it = b->graphs.find(key);
    if (it != b->graphs.end()) {
        *exec = it->second.second;
        return O2F_OK;
    }
    BCHK(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    const int st = enqueue_group(b, k, check);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(b->stream, &g);
    if (st) {
        if (g) (void)hipGraphDestroy(g);
        return st;
    }
    BCHK(b, e);
    hipGraphExec_t x = nullptr;
    const hipError_t ei = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
    if (ei != hipSuccess) {
        (void)hipGraphDestroy(g);
        BCHK(b, ei);
    }
    b->graphs[key] = std::make_pair(g, x);
    *exec = x;
    return O2F_OK;
}

// A group of k steps, the first one checked: chunks of at most kGraphSteps.
std::vector<std::pair<int, bool>> chunks_of(int k, bool check) {
    std::vector<std::pair<int, bool>> v;
    int left = k;
    bool first = true;
    while (left > 0) {
        const int n = std::min(left, kGraphSteps);
        v.emplace_back(n, check && first);
        first = false;
        left -= n;
    }
    return v;
}

int enqueue_chunks(o2f_batch *b, const std::vector<std::pair<int, bool>> &plan, bool use_graph) {
    int st;
    for (const auto &c : plan) {
        if (use_graph) {
            hipGraphExec_t x = nullptr;
            if ((st = get_graph(b, c.first, c.second, &x))) return st;
            BCHK(b, hipGraphLaunch(x, b->stream));
        } else if ((st = enqueue_group(b, c.first, c.second))) {
            return st;
        }
    }
    return O2F_OK;
}

bool any_active(o2f_batch *b, int &st) {
    st = O2F_OK;
    hipError_t e = launch_b_collect(b->dM, b->n(), b->dflags, b->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(b->hflags, b->dflags, b->n() * sizeof(double), hipMemcpyDeviceToHost,
                                            b->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
    if (e != hipSuccess) {
        st = b->fail(O2F_ERR_HIP, std::string("batch: ") + hipGetErrorString(e));
        return false;
    }
    for (int m = 0; m < b->n(); ++m)
        if (b->hflags[m] != 0.0) return true;
    return false;
}

int sync_members(o2f_batch *b) {
    for (o2f_ctx *c : b->ctx) {
        int st = o2f_ctx_check_bc(c);
        if (st) return b->fail(st, c->err);
        BCHK(b, hipStreamSynchronize(c->stream));
    }
    return O2F_OK;
}

int run(o2f_batch *b, double tol, int max_steps, int check_every, int use_graph, int *steps, double *res,
        double *elapsed_ms, double final_t = HUGE_VAL) {
    if (max_steps < 0 || check_every < 1) return b->fail(O2F_ERR_ARG, "batch run: max_steps >= 0, check_every >= 1");
    int st;
    if ((st = sync_members(b))) return st;
    const bool tiled = tiles_on(b);
    if ((st = upload_members(b, tol, final_t))) return st;
    if ((st = reset_scalars(b))) return st;
    const bool check = tol > 0.0;
    const bool poll = check || final_t < HUGE_VAL;  // members may stop before max_steps
    const int k = check_every;
    const int groups = max_steps / k, tail = max_steps % k;
    const auto plan = chunks_of(k, check), tail_plan = chunks_of(tail, check);
    if (use_graph) {  // capture before any work is queued on the stream
        hipGraphExec_t x;
        for (const auto *p : {&plan, &tail_plan})
            for (const auto &c : *p)
                if ((st = get_graph(b, c.first, c.second, &x))) return st;
    }
    if (tiled)  // the tile stages' AoS state
        for (int m = 0; m < b->n(); ++m)
            BCHK(b, launch_state_to_aos(b->ctx[m]->U, b->ctx[m]->hm.nc, b->Us[3 * m], b->stream));
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (elapsed_ms) {
        BCHK(b, hipEventCreate(&e0));
        BCHK(b, hipEventCreate(&e1));
        BCHK(b, hipEventRecord(e0, b->stream));
    }
    if (tiled)
        BCHK(b, launch_b_tile_speed(b->dM, b->tv, b->stream));
    else
        BCHK(b, launch_b_speed_max(b->dM, b->tiles[kKFlux], b->ntiles[kKFlux], b->shape[kKFlux], b->stream));
    // one host check per ~64 steps: finished members cost only early-exit launches meanwhile
    const int per_sync = std::max(1, 64 / k);
    bool live = true;
    for (int g = 0; g < groups && live; ++g) {
        if ((st = enqueue_chunks(b, plan, use_graph))) return st;
        if (poll && (g + 1) % per_sync == 0) {
            live = any_active(b, st);
            if (st) return st;
        }
    }
    if (live && tail > 0 && (st = enqueue_chunks(b, tail_plan, use_graph))) return st;
    if (elapsed_ms) {
        BCHK(b, hipEventRecord(e1, b->stream));
        BCHK(b, hipEventSynchronize(e1));
        float ms = 0.f;
        BCHK(b, hipEventElapsedTime(&ms, e0, e1));
        *elapsed_ms = ms;
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
    }
    // the state back to the contexts; the last stage's input (Us[2]) as SoA
    // in Us[1] for its intermediates below
    if (tiled)
        for (int m = 0; m < b->n(); ++m) {
            const int nc = b->ctx[m]->hm.nc;
            BCHK(b, launch_state_to_soa(b->Us[3 * m], nc, b->ctx[m]->U, b->stream));
            BCHK(b, launch_state_to_soa(b->Us[3 * m + 2], nc, b->Us[3 * m + 1], b->stream));
        }
    BCHK(b, hipStreamSynchronize(b->stream));
    for (int m = 0; m < b->n(); ++m) {
        o2f_ctx *c = b->ctx[m];
        double h[kScalCount];
        BCHK(b, hipMemcpy(h, c->scal, sizeof(h), hipMemcpyDeviceToHost));
        if (steps) steps[m] = static_cast<int>(h[kScalSteps]);
        if (res) res[m] = h[kScalRes];
        c->speed_valid = false;  // the batch kept its own speed partials
        ++c->faces_version;
        // the tile stages store no intermediates: those of the member's last
        // third stage, formed from its input as the four kernels form them
        // (R keeps the last checked step's, as on the four-kernel stages)
        if (tiled && h[kScalSteps] > 0.0) {
            if ((st = o2f_ctx_form_outputs(c, b->Us[3 * m + 1]))) return b->fail(st, c->err);
            BCHK(b, hipStreamSynchronize(c->stream));
        }
    }
    return O2F_OK;
}

}  // namespace

extern "C" {

int o2f_batch_create(o2f_ctx *const *members, int n, o2f_batch **out) {
    if (!out) return O2F_ERR_ARG;
    *out = nullptr;
    if (!members || n <= 0) return O2F_ERR_ARG;
    for (int m = 0; m < n; ++m)
        if (!members[m]) return O2F_ERR_ARG;
    o2f_batch *b = new o2f_batch();
    b->ctx.assign(members, members + n);
    bool ok = hipStreamCreate(&b->stream) == hipSuccess && dalloc(b->dM, n) == hipSuccess &&
              dalloc(b->dflags, n) == hipSuccess &&
              hipHostMalloc(reinterpret_cast<void **>(&b->hflags), n * sizeof(double), hipHostMallocDefault) ==
                  hipSuccess;
    std::vector<int2> rt;
    for (int m = 0; ok && m < n; ++m) {
        const HostMesh &hm = members[m]->hm;
        double *sp = nullptr, *rp = nullptr;
        ok = dalloc(sp, blocks_of(kKFlux, hm, kMinBlock)) == hipSuccess &&
             dalloc(rp, ceil_div(hm.nc, kResChunk)) == hipSuccess;
        b->speed.push_back(sp);
        b->rparts.push_back(rp);
        for (long k = 0; k < ceil_div(hm.nc, kResChunk); ++k) rt.push_back(make_int2(m, static_cast<int>(k)));
    }
    ok = ok && dalloc(b->rtiles, rt.size()) == hipSuccess &&
         hipMemcpy(b->rtiles, rt.data(), rt.size() * sizeof(int2), hipMemcpyHostToDevice) == hipSuccess;
    b->nrtiles = static_cast<int>(rt.size());
    for (int k = 0; ok && k < kKinds; ++k) ok = build_tiles(b, k) == O2F_OK;
    if (!ok) {
        o2f_batch_destroy(b);
        return O2F_ERR_HIP;
    }
    *out = b;
    return O2F_OK;
}

void o2f_batch_destroy(o2f_batch *b) {
    if (!b) return;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    b->drop_graph();
    release_tiles(b);
    for (int k = 0; k < kKinds; ++k)
        if (b->tiles[k]) (void)hipFree(b->tiles[k]);
    for (double *p : b->speed) (void)hipFree(p);
    for (double *p : b->rparts) (void)hipFree(p);
    if (b->rtiles) (void)hipFree(b->rtiles);
    if (b->dM) (void)hipFree(b->dM);
    if (b->dflags) (void)hipFree(b->dflags);
    if (b->hflags) (void)hipHostFree(b->hflags);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
}

const char *o2f_batch_last_error(const o2f_batch *b) { return b ? b->err.c_str() : "null batch"; }

int o2f_batch_set_shapes(o2f_batch *b, const int *shapes) {
    if (!b || !shapes) return O2F_ERR_ARG;
    for (int k = 0; k < kKinds; ++k) {
        const TileShape t{shapes[2 * k], shapes[2 * k + 1]};
        if (!valid_shape(t))
            return b->fail(O2F_ERR_ARG, "o2f_batch_set_shapes: block must be 64, 128 or 256 and grid >= 0");
    }
    for (int k = 0; k < kKinds; ++k) {
        const TileShape t{shapes[2 * k], shapes[2 * k + 1]};
        if (t.block != b->shape[k].block || t.grid != b->shape[k].grid) {
            b->shape[k] = t;
            int st = build_tiles(b, k);
            if (st) return st;
            b->drop_graph();
        }
    }
    return O2F_OK;
}

int o2f_batch_set_tile_cells(o2f_batch *b, int cells) {
    if (!b || cells < -1) return O2F_ERR_ARG;
    if (cells != b->tile_cells || b->tiles_unusable) {
        if (b->stream) (void)hipStreamSynchronize(b->stream);
        b->drop_graph();
        release_tiles(b);
        b->tiles_unusable = false;
        b->tiles_why.clear();
        b->tile_cells = cells;
        b->tile_tuned = 0;
    }
    return O2F_OK;
}

int o2f_batch_tile_info(o2f_batch *b, int *n_tiles, int *cells, int *lds_bytes, int *tiles_per_cu, int *vgprs) {
    if (!b) return O2F_ERR_ARG;
    const bool on = tiles_on(b);
    if (!on && b->tile_cells != 0) b->err = "four-kernel stages: " + b->tiles_why;
    if (n_tiles) *n_tiles = on ? b->tv.ntiles : 0;
    if (cells) *cells = on ? b->tile_target : 0;
    if (lds_bytes) *lds_bytes = on ? (int)b->tv.lds_bytes : 0;
    if (tiles_per_cu) *tiles_per_cu = on ? b->tile_bpc : 0;
    if (vgprs) *vgprs = on ? b->tile_vgprs : 0;
    return O2F_OK;
}

int o2f_batch_bicgstab(o2f_batch *b, double beta, double eps, double tol, int chunk, int max_total, int exact,
                       double *const *lambdas, int *iters, double *relres, int *status) {
    if (!b || chunk <= 0 || max_total < 0 || (exact != 0 && exact != 1)) return O2F_ERR_ARG;
    int st;
    if ((st = sync_members(b))) return st;
    BCHK(b, hipStreamSynchronize(b->stream));
    // allocations first, on this thread; then each member's own solve on
    // its own stream, several members at once
    for (o2f_ctx *c : b->ctx)
        if ((st = o2f_ctx_adjoint_prepare(c, exact))) return b->fail(st, c->err);
    std::vector<std::vector<double>> scratch(b->n());
    std::vector<int> it(b->n(), 0), code(b->n(), O2F_OK);
    std::vector<double> rr(b->n(), 0.0);
    std::atomic<int> next{0};
    auto work = [&] {
        for (int m; (m = next++) < b->n();) {
            o2f_ctx *c = b->ctx[m];
            double *lam = lambdas ? lambdas[m] : nullptr;
            if (!lam) {
                scratch[m].assign(4 * (size_t)c->hm.nc, 0.0);
                lam = scratch[m].data();
            }
            code[m] = exact ? o2f_bicgstab_exact(c, beta, tol, chunk, max_total, lam, &it[m], &rr[m], 0)
                            : o2f_bicgstab(c, beta, eps, tol, chunk, max_total, lam, &it[m], &rr[m]);
        }
    };
    const int nthreads = std::max(1, std::min({kAdjointThreads, b->n(), (int)std::thread::hardware_concurrency()}));
    std::vector<std::thread> pool;
    try {
        for (int i = 1; i < nthreads; ++i) pool.emplace_back(work);
    } catch (const std::exception &) {  // no more threads: the rest on this one
    }
    work();
    for (std::thread &th : pool) th.join();
    for (int m = 0; m < b->n(); ++m) {
        if (iters) iters[m] = it[m];
        if (relres) relres[m] = rr[m];
        if (status) status[m] = code[m];
    }
    // a member whose solve failed on the device (not a convergence
    // outcome) fails the call
    for (int m = 0; m < b->n(); ++m)
        if (code[m] == O2F_ERR_HIP || code[m] == O2F_ERR_ARG) return b->fail(code[m], b->ctx[m]->err);
    return O2F_OK;
}

int o2f_batch_get_shapes(const o2f_batch *b, int *shapes) {
    if (!b || !shapes) return O2F_ERR_ARG;
    for (int k = 0; k < kKinds; ++k) {
        shapes[2 * k] = b->shape[k].block;
        shapes[2 * k + 1] = b->shape[k].grid;
    }
    return O2F_OK;
}

int o2f_batch_autotune(o2f_batch *b, int reps, int *shapes) {
    if (!b || reps < 1) return O2F_ERR_ARG;
    int st;
    if ((st = sync_members(b))) return st;
    if ((st = upload_members(b, -1.0))) return st;
    if ((st = reset_scalars(b))) return st;  // all members active for the timing passes
    int cu = 256;
    int dev = 0;
    if (hipGetDevice(&dev) == hipSuccess) {
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, dev) == hipSuccess && prop.multiProcessorCount > 0)
            cu = prop.multiProcessorCount;
    }
    const int blocks[3] = {64, 128, 256};
    const int grids[3] = {0, 4 * cu, 8 * cu};
    hipEvent_t ev[2];
    BCHK(b, hipEventCreate(&ev[0]));
    BCHK(b, hipEventCreate(&ev[1]));
    for (int kind = 0; kind < kKinds; ++kind) {
        TileShape best = b->shape[kind];
        float best_ms = -1.f;
        for (int bi = 0; bi < 3; ++bi)
            for (int gi = 0; gi < 3; ++gi) {
                b->shape[kind] = TileShape{blocks[bi], grids[gi]};
                if ((st = build_tiles(b, kind))) return st;
                if ((st = upload_members(b, -1.0))) return st;  // nparts follows the flux block
                if ((st = enqueue_stage(b, 0, 0.0, 0.0))) return st;  // warm-up
                float total = 0.f;
                for (int r = 0; r < reps; ++r) {
                    if ((st = enqueue_stage(b, 0, 0.0, 0.0, ev, kind))) return st;
                    BCHK(b, hipEventSynchronize(ev[1]));
                    float ms = 0.f;
                    BCHK(b, hipEventElapsedTime(&ms, ev[0], ev[1]));
                    total += ms;
                }
                if (best_ms < 0.f || total < best_ms) {
                    best_ms = total;
                    best = b->shape[kind];
                }
            }
        b->shape[kind] = best;
        if ((st = build_tiles(b, kind))) return st;
    }
    // Automatic tile stages: the rule's size and two smaller ones (more
    // tiles per round, less LDS), each timed on first-stage launches from
    // the members' current states (written to the stage's scratch buffers
    // and U_old only); the fastest is kept.
    if (b->tile_cells < 0 && !b->tiles_unusable) {
        BCHK(b, hipStreamSynchronize(b->stream));
        b->drop_graph();
        release_tiles(b);
        b->tile_tuned = 0;
        if (tiles_on(b)) {
            const int t0 = b->tile_target;
            const int cand[3] = {t0, t0 * 9 / 10, t0 * 4 / 5};
            int best_t = t0;
            float best_ms = -1.f;
            for (int t : cand) {
                b->drop_graph();
                release_tiles(b);
                b->tile_tuned = t;
                if (!tiles_on(b)) {
                    b->tiles_unusable = false;  // this size only: the others may fit
                    b->tiles_why.clear();
                    continue;
                }
                if ((st = upload_members(b, -1.0))) return st;
                for (int m = 0; m < b->n(); ++m)
                    BCHK(b, launch_state_to_aos(b->ctx[m]->U, b->ctx[m]->hm.nc, b->Us[3 * m], b->stream));
                const int flags = kFlagRK | kFlagSaveOld;
                BCHK(b, launch_b_tile_stage(b->dM, b->tv, 0, 0.0, 1.0, flags, b->stream));  // warm-up
                BCHK(b, hipEventRecord(ev[0], b->stream));
                for (int r = 0; r < reps; ++r) BCHK(b, launch_b_tile_stage(b->dM, b->tv, 0, 0.0, 1.0, flags, b->stream));
                BCHK(b, hipEventRecord(ev[1], b->stream));
                BCHK(b, hipEventSynchronize(ev[1]));
                float ms = 0.f;
                BCHK(b, hipEventElapsedTime(&ms, ev[0], ev[1]));
                if (best_ms < 0.f || ms < best_ms) {
                    best_ms = ms;
                    best_t = t;
                }
            }
            b->drop_graph();
            release_tiles(b);
            b->tile_tuned = best_t;
            (void)tiles_on(b);
        }
    }
    (void)hipEventDestroy(ev[0]);
    (void)hipEventDestroy(ev[1]);
    if ((st = upload_members(b, -1.0))) return st;
    for (o2f_ctx *c : b->ctx) ++c->faces_version;
    b->drop_graph();
    return o2f_batch_get_shapes(b, shapes ? shapes : std::vector<int>(8).data());
}

int o2f_batch_run(o2f_batch *b, double tol, int max_steps, int check_every, int use_graph, int *steps,
                  double *res) {
    if (!b) return O2F_ERR_ARG;
    return run(b, tol, max_steps, check_every, use_graph, steps, res, nullptr);
}

int o2f_batch_run_until(o2f_batch *b, double final_time, int max_steps, int use_graph, int *steps, double *t) {
    if (!b || !(final_time > 0.0)) return O2F_ERR_ARG;
    int st = run(b, -1.0, max_steps, 50, use_graph, steps, nullptr, nullptr, final_time);
    if (st || !t) return st;
    for (int m = 0; m < b->n(); ++m) {
        double tm = 0.0;
        BCHK(b, hipMemcpy(&tm, b->ctx[m]->scal + kScalT, sizeof(double), hipMemcpyDeviceToHost));
        t[m] = tm;
    }
    return O2F_OK;
}

int o2f_batch_time_steps(o2f_batch *b, int nsteps, int use_graph, double *total_ms) {
    if (!b || !total_ms || nsteps < 1) return O2F_ERR_ARG;
    const int k = std::min(nsteps, 10);
    return run(b, -1.0, nsteps, k, use_graph, nullptr, nullptr, total_ms);
}

}  // extern "C"

// o2f_run_to_residual on the batched four-kernel stages (a batch of one).
int run_to_residual_batch(o2f_ctx *c, double tol, int max_steps, int check_every, int use_graph, int *steps,
                          double *res) {
    o2f_batch *b = nullptr;
    o2f_ctx *const one[1] = {c};
    int st = o2f_batch_create(one, 1, &b);
    if (st) return c->fail(st, "o2f_run_to_residual: batch set-up failed");
    st = o2f_batch_set_tile_cells(b, 0);  // the context's own stages are the four kernels
    if (st == O2F_OK) st = o2f_batch_run(b, tol, max_steps, check_every, use_graph, steps, res);
    if (st) c->fail(st, b->err);
    o2f_batch_destroy(b);
    return st;
}
