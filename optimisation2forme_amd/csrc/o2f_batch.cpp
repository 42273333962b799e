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
#include <cmath>
#include <map>
#include <utility>
#include <string>
#include <vector>

#include "o2f.h"
#include "o2f_ctx.hpp"
#include "o2f_kernels.hpp"

using namespace o2f;

namespace {

enum { kKVertex = 0, kKRecon = 1, kKViscEdge = 2, kKFlux = 3, kKinds = 4 };
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

int enqueue_step(o2f_batch *b, bool check) {
    int st;
    BCHK(b, launch_b_step_begin(b->dM, b->n(), b->stream));
    for (int stg = 0; stg < 3; ++stg) {
        const int flags = kFlagRK | (stg == 0 ? kFlagSaveOld : 0) | (stg == 0 && check ? kFlagStoreR : 0) |
                          (stg == 2 ? kFlagSpeed : 0);
        if ((st = enqueue_stage(b, flags, kC0[stg], kC1[stg]))) return st;
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
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (elapsed_ms) {
        BCHK(b, hipEventCreate(&e0));
        BCHK(b, hipEventCreate(&e1));
        BCHK(b, hipEventRecord(e0, b->stream));
    }
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
    BCHK(b, hipStreamSynchronize(b->stream));
    for (int m = 0; m < b->n(); ++m) {
        o2f_ctx *c = b->ctx[m];
        double h[kScalCount];
        BCHK(b, hipMemcpy(h, c->scal, sizeof(h), hipMemcpyDeviceToHost));
        if (steps) steps[m] = static_cast<int>(h[kScalSteps]);
        if (res) res[m] = h[kScalRes];
        c->speed_valid = false;  // the batch kept its own speed partials
        ++c->faces_version;
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
    st = o2f_batch_run(b, tol, max_steps, check_every, use_graph, steps, res);
    if (st) c->fail(st, b->err);
    o2f_batch_destroy(b);
    return st;
}
