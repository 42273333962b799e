// Learned warm start: a per-cell MLP that maps (x, y, wall distance,
// m, p, t, Mach, AoA) plus Fourier features of (x, y) to the conservative
// state, trained on fields the solver produced.  BF16 operands on the
// gfx950 matrix cores (v_mfma_f32_16x16x32_bf16) with FP32 accumulation,
// FP32 master weights and Adam, all hand-written: feature build, GEMM with
// fused bias / SiLU / SiLU' epilogues, split-K weight and bias gradients
// with a fixed-order reduction, loss, Adam.  Deterministic: no atomics.
//
// Layout: activations row-major [samples x width] in bf16 (stored as
// uint16), pre-activations in fp32; weights W_l row-major [K_l x N_l].
// Widths: input 64 (40 features, zero padded), 4 hidden layers of 256,
// output padded to 64 (the first 4 columns are the state).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <random>
#include <tuple>
#include <string>
#include <vector>

#include "o2f.h"
#include "o2f_ctx.hpp"

namespace o2f {
namespace mlp {

constexpr int kIn = 64, kHid = 256, kOut = 64, kLayers = 5, kFourier = 16, kRaw = 8;
constexpr int kTile = 64, kTK = 32, kLdsRow = kTK + 8;  // LDS rows padded by 16 bytes
constexpr int kSplitChunk = 1024;                          // samples per split-K partial of dW

typedef unsigned short u16;
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
typedef u16 v8u __attribute__((ext_vector_type(8)));
typedef float v4f __attribute__((ext_vector_type(4)));

// fp32 -> bf16, round to nearest even (NaN kept quiet)
__host__ __device__ __forceinline__ u16 to_bf16(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (u16)((u >> 16) | 0x40);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (u16)(u >> 16);
}
__host__ __device__ __forceinline__ float from_bf16(u16 h) {
    const uint32_t u = (uint32_t)h << 16;
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}
__device__ __forceinline__ float sigm(float z) { return 1.0f / (1.0f + expf(-z)); }

enum : int { kEpiHidden = 0, kEpiOut = 1, kEpiBack = 2, kEpiPartial = 3 };

struct Epi {
    int mode;
    const float *bias;   // hidden / out
    float *Z;            // hidden: pre-activation out
    u16 *act;            // hidden: silu(z) out; back: dZ out
    const float *Zprev;  // back: pre-activation of the layer below
    float *C;            // out / partial
    long ldc;
};

// C[M x N] = A[M x K] B[K x N]; A(m, k) = A[m sam + k sak], B(k, n) = B[k sbk + n sbn].
// A TM x TN workgroup tile: 256 threads = 4 waves as 2 x 2, each wave
// (TM/2) x (TN/2) = (TM/32) x (TN/32) MFMA tiles of 16 x 16 x 32.  gridDim.z
// splits K into chunks of kchunk (partial outputs at C + z M ldc).
template <int TM, int TN>
__global__ __launch_bounds__(256) void k_gemm(const u16 *__restrict__ A, long sam, long sak,
                                              const u16 *__restrict__ B, long sbk, long sbn, int M, int N,
                                              int K, int kchunk, Epi epi) {
    constexpr int WM = TM / 2, WN = TN / 2, FM = WM / 16, FN = WN / 16;
    __shared__ u16 As[TM][kLdsRow];
    __shared__ u16 Bt[TN][kLdsRow];  // transposed: Bt[n][k]
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int wm = w >> 1, wn = w & 1;
    const int m0 = blockIdx.y * TM, n0 = blockIdx.x * TN;
    const int kb = blockIdx.z * kchunk, ke = min(K, kb + kchunk);
    v4f acc[FM][FN];
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j) acc[i][j] = v4f{0.f, 0.f, 0.f, 0.f};
    for (int k0 = kb; k0 < ke; k0 += kTK) {
        // global -> LDS with 16-byte loads along whichever index is contiguous
        // (every index is inside the M x K and K x N operands: M, N are
        // multiples of the tile and K of 32, checked by the host)
#pragma unroll
        for (int it = 0; it < TM / 64; ++it) {
            const int idx = t + 256 * it;
            if (sak == 1) {  // A rows contiguous in k
                const int r = idx >> 2, k = (idx & 3) * 8;
                *reinterpret_cast<v8u *>(&As[r][k]) =
                    *reinterpret_cast<const v8u *>(A + (long)(m0 + r) * sam + (k0 + k));
            } else {  // A contiguous in m: 8 rows of one k, transposed into LDS
                const int k = idx / (TM / 8), r = (idx % (TM / 8)) * 8;
                const v8u va = *reinterpret_cast<const v8u *>(A + (long)(k0 + k) * sak + (m0 + r));
#pragma unroll
                for (int i = 0; i < 8; ++i) As[r + i][k] = va[i];
            }
        }
#pragma unroll
        for (int it = 0; it < TN / 64; ++it) {
            const int idx = t + 256 * it;
            if (sbk == 1) {  // B columns contiguous in k
                const int n = idx >> 2, k = (idx & 3) * 8;
                *reinterpret_cast<v8u *>(&Bt[n][k]) =
                    *reinterpret_cast<const v8u *>(B + (long)(n0 + n) * sbn + (k0 + k));
            } else {  // B contiguous in n
                const int k = idx / (TN / 8), n = (idx % (TN / 8)) * 8;
                const v8u vb = *reinterpret_cast<const v8u *>(B + (long)(k0 + k) * sbk + (n0 + n));
#pragma unroll
                for (int i = 0; i < 8; ++i) Bt[n + i][k] = vb[i];
            }
        }
        __syncthreads();
        v8u bv[FN];
#pragma unroll
        for (int j = 0; j < FN; ++j)
            bv[j] = *reinterpret_cast<const v8u *>(&Bt[wn * WN + j * 16 + (lane & 15)][8 * (lane >> 4)]);
#pragma unroll
        for (int i = 0; i < FM; ++i) {
            const v8u a = *reinterpret_cast<const v8u *>(&As[wm * WM + i * 16 + (lane & 15)][8 * (lane >> 4)]);
#pragma unroll
            for (int j = 0; j < FN; ++j)
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(v8bf, a),
                                                                     __builtin_bit_cast(v8bf, bv[j]), acc[i][j], 0,
                                                                     0, 0);
        }
        __syncthreads();
    }
    // D[row = 4 (lane / 16) + r][col = lane % 16] of each 16 x 16 tile
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const long m = m0 + wm * WM + i * 16 + 4 * (lane >> 4) + r;
                const long n = n0 + wn * WN + j * 16 + (lane & 15);
                const float v = acc[i][j][r];
                const long o = m * epi.ldc + n;
                if (epi.mode == kEpiHidden) {
                    const float z = v + epi.bias[n];
                    epi.Z[o] = z;
                    epi.act[o] = to_bf16(z * sigm(z));
                } else if (epi.mode == kEpiOut) {
                    epi.C[o] = v + epi.bias[n];
                } else if (epi.mode == kEpiBack) {
                    const float z = epi.Zprev[o], s = sigm(z);
                    epi.act[o] = to_bf16(v * (s * (1.0f + z * (1.0f - s))));
                } else {
                    epi.C[(long)blockIdx.z * M * epi.ldc + o] = v;
                }
            }
}

// The same GEMM with K stages of 64 double-buffered in LDS: the next
// stage's global loads are in flight while the matrix cores work on the
// current one, and one barrier per stage separates the two buffers (each
// wave issues FM x FN x 2 MFMAs per barrier instead of FM x FN).  The k
// order of every accumulation (32-wide MFMA steps in increasing k) is that
// of k_gemm, so both kernels give bit-identical results.
constexpr int kTK2 = 64, kLdsRow2 = kTK2 + 8;

template <int TM, int TN>
__global__ __launch_bounds__(256) void k_gemm2(const u16 *__restrict__ A, long sam, long sak,
                                               const u16 *__restrict__ B, long sbk, long sbn, int M, int N,
                                               int K, int kchunk, Epi epi) {
    constexpr int WM = TM / 2, WN = TN / 2, FM = WM / 16, FN = WN / 16;
    constexpr int LA = TM * kTK2 / (256 * 8), LB = TN * kTK2 / (256 * 8);  // 16-byte loads per thread
    __shared__ u16 As[2][TM][kLdsRow2];
    __shared__ u16 Bt[2][TN][kLdsRow2];  // transposed: Bt[n][k]
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int wm = w >> 1, wn = w & 1;
    const int m0 = blockIdx.y * TM, n0 = blockIdx.x * TN;
    const int kb = blockIdx.z * kchunk, ke = min(K, kb + kchunk);
    v8u ra[LA], rb[LB];
    // global -> registers for the stage at k0 (every index inside the
    // operands: M, N multiples of the tile, K and kchunk of 64, checked by
    // the host)
    auto load = [&](int k0) {
#pragma unroll
        for (int it = 0; it < LA; ++it) {
            const int idx = t + 256 * it;
            if (sak == 1) {
                const int r = idx >> 3, k = (idx & 7) * 8;
                ra[it] = *reinterpret_cast<const v8u *>(A + (long)(m0 + r) * sam + (k0 + k));
            } else {
                const int k = idx / (TM / 8), r = (idx % (TM / 8)) * 8;
                ra[it] = *reinterpret_cast<const v8u *>(A + (long)(k0 + k) * sak + (m0 + r));
            }
        }
#pragma unroll
        for (int it = 0; it < LB; ++it) {
            const int idx = t + 256 * it;
            if (sbk == 1) {
                const int n = idx >> 3, k = (idx & 7) * 8;
                rb[it] = *reinterpret_cast<const v8u *>(B + (long)(n0 + n) * sbn + (k0 + k));
            } else {
                const int k = idx / (TN / 8), n = (idx % (TN / 8)) * 8;
                rb[it] = *reinterpret_cast<const v8u *>(B + (long)(k0 + k) * sbk + (n0 + n));
            }
        }
    };
    // registers -> LDS buffer `buf`
    auto store = [&](int buf) {
#pragma unroll
        for (int it = 0; it < LA; ++it) {
            const int idx = t + 256 * it;
            if (sak == 1) {
                const int r = idx >> 3, k = (idx & 7) * 8;
                *reinterpret_cast<v8u *>(&As[buf][r][k]) = ra[it];
            } else {
                const int k = idx / (TM / 8), r = (idx % (TM / 8)) * 8;
#pragma unroll
                for (int i = 0; i < 8; ++i) As[buf][r + i][k] = ra[it][i];
            }
        }
#pragma unroll
        for (int it = 0; it < LB; ++it) {
            const int idx = t + 256 * it;
            if (sbk == 1) {
                const int n = idx >> 3, k = (idx & 7) * 8;
                *reinterpret_cast<v8u *>(&Bt[buf][n][k]) = rb[it];
            } else {
                const int k = idx / (TN / 8), n = (idx % (TN / 8)) * 8;
#pragma unroll
                for (int i = 0; i < 8; ++i) Bt[buf][n + i][k] = rb[it][i];
            }
        }
    };
    v4f acc[FM][FN];
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j) acc[i][j] = v4f{0.f, 0.f, 0.f, 0.f};
    load(kb);
    store(0);
    __syncthreads();
    int buf = 0;
    for (int k0 = kb; k0 < ke; k0 += kTK2) {
        const bool more = k0 + kTK2 < ke;
        if (more) load(k0 + kTK2);
#pragma unroll
        for (int kk = 0; kk < kTK2; kk += 32) {
            v8u bv[FN];
#pragma unroll
            for (int j = 0; j < FN; ++j)
                bv[j] = *reinterpret_cast<const v8u *>(&Bt[buf][wn * WN + j * 16 + (lane & 15)][kk + 8 * (lane >> 4)]);
#pragma unroll
            for (int i = 0; i < FM; ++i) {
                const v8u a =
                    *reinterpret_cast<const v8u *>(&As[buf][wm * WM + i * 16 + (lane & 15)][kk + 8 * (lane >> 4)]);
#pragma unroll
                for (int j = 0; j < FN; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(v8bf, a),
                                                                         __builtin_bit_cast(v8bf, bv[j]), acc[i][j],
                                                                         0, 0, 0);
            }
        }
        // the other buffer was last read before the previous barrier
        if (more) store(buf ^ 1);
        __syncthreads();
        buf ^= 1;
    }
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const long m = m0 + wm * WM + i * 16 + 4 * (lane >> 4) + r;
                const long n = n0 + wn * WN + j * 16 + (lane & 15);
                const float v = acc[i][j][r];
                const long o = m * epi.ldc + n;
                if (epi.mode == kEpiHidden) {
                    const float z = v + epi.bias[n];
                    epi.Z[o] = z;
                    epi.act[o] = to_bf16(z * sigm(z));
                } else if (epi.mode == kEpiOut) {
                    epi.C[o] = v + epi.bias[n];
                } else if (epi.mode == kEpiBack) {
                    const float z = epi.Zprev[o], s = sigm(z);
                    epi.act[o] = to_bf16(v * (s * (1.0f + z * (1.0f - s))));
                } else {
                    epi.C[(long)blockIdx.z * M * epi.ldc + o] = v;
                }
            }
}

// dW[i] = sum over split chunks s = 0, 1, ... in order of the partials.
__global__ __launch_bounds__(256) void k_sum_splits(const float *__restrict__ P, int splits, long n,
                                                    float *__restrict__ out) {
    const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i >= n) return;
    float s = 0.f;
    for (int k = 0; k < splits; ++k) s += P[(long)k * n + i];
    out[i] = s;
}

// Input features of one sample from its 8 raw inputs
// (x, y, wall distance, m, p, t, Mach, AoA in degrees) and the Fourier
// frequencies B (2 x kFourier): scaled raw inputs, sin / cos (2 pi B^T (x, y)).
__device__ __forceinline__ void features(const float *raw, const float *__restrict__ fb, u16 *row) {
    const float x = raw[0] - 0.855f, y = raw[1];
    float f[kIn];
    f[0] = x / 20.f;
    f[1] = y / 20.f;
    f[2] = log10f(raw[2] + 1e-4f) / 4.f;
    f[3] = raw[3] / 0.06f;
    f[4] = (raw[4] - 0.4f) / 0.2f;
    f[5] = (raw[5] - 0.115f) / 0.035f;
    f[6] = (raw[6] - 0.675f) / 0.175f;
    f[7] = raw[7] / 4.f;
#pragma unroll
    for (int k = 0; k < kFourier; ++k) {
        const float a = 6.283185307179586f * (fb[k] * x + fb[kFourier + k] * y);
        f[kRaw + 2 * k] = sinf(a);
        f[kRaw + 2 * k + 1] = cosf(a);
    }
#pragma unroll
    for (int k = kRaw + 2 * kFourier; k < kIn; ++k) f[k] = 0.f;
#pragma unroll
    for (int k = 0; k < kIn; ++k) row[k] = to_bf16(f[k]);
}

// Sample index of batch row i at step `step`: a fixed hash (splitmix64).
__device__ __forceinline__ long pick(unsigned long long seed, long step, long i, long n) {
    unsigned long long z = seed + 0x9e3779b97f4a7c15ull * (unsigned long long)(step * 1000003LL + i + 1);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    z ^= z >> 31;
    return (long)(z % (unsigned long long)n);
}

// X0 and targets of a batch: rows [0, n) from samples idx (or a random pick
// when step >= 0), rows [n, rows) zero (masked in the loss).
__global__ __launch_bounds__(256) void k_batch(const float *__restrict__ raw, const float *__restrict__ tgt,
                                               long ndata, long first, int n, int rows, long step,
                                               unsigned long long seed, const float *__restrict__ fb,
                                               u16 *__restrict__ X0, float *__restrict__ T) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= rows) return;
    u16 *row = X0 + (long)i * kIn;
    if (i >= n) {
        for (int k = 0; k < kIn; ++k) row[k] = 0;
        for (int q = 0; q < 4; ++q) T[4L * i + q] = 0.f;
        return;
    }
    const long s = step >= 0 ? pick(seed, step, i, ndata) : first + i;
    features(raw + kRaw * s, fb, row);
    if (tgt)
        for (int q = 0; q < 4; ++q) T[4L * i + q] = tgt[4 * s + q];
}

// Mean squared error over the n real rows and 4 outputs; dY = 2 (y - t) / (4 n)
// (bf16, zero on padding rows and columns), per-block loss partials.
__global__ __launch_bounds__(256) void k_loss(const float *__restrict__ Y, const float *__restrict__ T, int n,
                                              int rows, u16 *__restrict__ dY, float *__restrict__ parts) {
    __shared__ float sm[256];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    float l = 0.f;
    if (i < rows) {
        const float sc = 2.0f / (4.0f * n);
        for (int c = 0; c < kOut; ++c) {
            float g = 0.f;
            if (i < n && c < 4) {
                const float d = Y[(long)i * kOut + c] - T[4L * i + c];
                l += d * d;
                g = sc * d;
            }
            dY[(long)i * kOut + c] = to_bf16(g);
        }
    }
    sm[threadIdx.x] = l;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) sm[threadIdx.x] += sm[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) parts[blockIdx.x] = sm[0];
}

__global__ void k_loss_final(const float *__restrict__ parts, int np, int n, float *__restrict__ out) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    float s = 0.f;
    for (int k = 0; k < np; ++k) s += parts[k];
    *out = s / (4.0f * n);
}

// Adam (bias-corrected) on the fp32 masters, then bf16 copies of the weights.
__global__ __launch_bounds__(256) void k_adam(float *__restrict__ w, float *__restrict__ m, float *__restrict__ v,
                                              const float *__restrict__ g, const unsigned char *__restrict__ frozen,
                                              long n, float lr, float b1, float b2, float eps, float c1, float c2) {
    const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i >= n || frozen[i]) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    w[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

__global__ __launch_bounds__(256) void k_to_bf16(const float *__restrict__ a, long n, u16 *__restrict__ b) {
    const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i < n) b[i] = to_bf16(a[i]);
}

// Warm-start state of the cells: y (normalised) -> conservative variables.
__global__ __launch_bounds__(256) void k_denorm(const float *__restrict__ Y, int nc, double rho, double a, double p,
                                                double mach, double gamma, double *__restrict__ U) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nc) return;
    const double y0 = Y[(long)c * kOut], y1 = Y[(long)c * kOut + 1], y2 = Y[(long)c * kOut + 2];
    const double y3 = Y[(long)c * kOut + 3];
    const double u = mach * a;
    const double Einf = p / (gamma - 1.0) + 0.5 * rho * u * u;
    U[c] = rho * (1.0 + y0);
    U[nc + c] = rho * a * (mach + y1);
    U[2 * nc + c] = rho * a * y2;
    U[3 * nc + c] = Einf + p * y3;
}

}  // namespace mlp
}  // namespace o2f

using namespace o2f;
using namespace o2f::mlp;

struct o2f_mlp {
    int K[kLayers] = {kIn, kHid, kHid, kHid, kHid};
    int N[kLayers] = {kHid, kHid, kHid, kHid, kOut};
    long woff[kLayers] = {}, boff[kLayers] = {};
    long nparam = 0;
    float *w = nullptr, *m = nullptr, *v = nullptr, *g = nullptr, *fb = nullptr;
    unsigned char *frozen = nullptr;
    u16 *w16 = nullptr;
    int cap = 0;  // rows of the activation buffers
    u16 *X[kLayers] = {}, *dA = nullptr, *dB = nullptr, *dY = nullptr;
    float *Z[kLayers - 1] = {}, *Y = nullptr, *T = nullptr, *split = nullptr, *lparts = nullptr, *lout = nullptr;
    u16 *ones = nullptr;     // [cap x 64] bf16, column 0 = 1: bias gradients as a GEMM
    float *dbtmp = nullptr;  // [64 x 256] sums of that GEMM (row 0 is the bias gradient)
    float *raw = nullptr, *tgt = nullptr;
    long ndata = 0;
    long step = 0;
    std::map<std::tuple<int, int, int, int, bool, bool, int>, int> tile_choice;
    int gemm_variant = -1;  // >= 0: that GEMM kernel for every shape (else the per-shape choice)
    hipStream_t stream = nullptr;
    std::string err;
    int fail(int code, const std::string &s) {
        err = s;
        return code;
    }
};

#define MCHK(o, expr)                                                                         \
    do {                                                                                      \
        hipError_t e_ = (expr);                                                               \
        if (e_ != hipSuccess) return (o)->fail(O2F_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

namespace {

unsigned nblk(long n, int b = 256) { return (unsigned)((n + b - 1) / b); }

int ensure_cap(o2f_mlp *o, int rows) {
    if (rows <= o->cap) return O2F_OK;
    for (auto *p : o->X) (void)hipFree(p);
    for (auto *p : o->Z) (void)hipFree(p);
    for (void *p : {(void *)o->dA, (void *)o->dB, (void *)o->dY, (void *)o->Y, (void *)o->T, (void *)o->split,
                    (void *)o->lparts, (void *)o->ones, (void *)o->dbtmp})
        (void)hipFree(p);
    const long r = rows;
    MCHK(o, dalloc(o->X[0], r * kIn));
    for (int l = 1; l < kLayers; ++l) MCHK(o, dalloc(o->X[l], r * kHid));
    for (int l = 0; l < kLayers - 1; ++l) MCHK(o, dalloc(o->Z[l], r * kHid));
    MCHK(o, dalloc(o->dA, r * kHid));
    MCHK(o, dalloc(o->dB, r * kHid));
    MCHK(o, dalloc(o->dY, r * kOut));
    MCHK(o, dalloc(o->Y, r * kOut));
    MCHK(o, dalloc(o->T, r * 4));
    const long splits = (r + kSplitChunk - 1) / kSplitChunk;
    MCHK(o, dalloc(o->split, splits * kHid * kHid));
    MCHK(o, dalloc(o->lparts, (long)nblk(r)));
    MCHK(o, dalloc(o->ones, r * kIn));
    MCHK(o, dalloc(o->dbtmp, (long)kIn * kHid));
    {
        std::vector<u16> h((size_t)r * kIn, 0);
        for (long i = 0; i < r; ++i) h[(size_t)i * kIn] = to_bf16(1.0f);
        MCHK(o, hipMemcpy(o->ones, h.data(), h.size() * sizeof(u16), hipMemcpyHostToDevice));
    }
    o->cap = rows;
    return O2F_OK;
}

// GEMM kernels: 0-2 k_gemm <64,64>, <128,64>, <128,128>; 3-6 k_gemm2
// <128,128>, <128,64>, <64,128>, <64,64>.
constexpr int kGemmVariants = 7;
constexpr int kVarTM[kGemmVariants] = {64, 128, 128, 128, 128, 64, 64};
constexpr int kVarTN[kGemmVariants] = {64, 64, 128, 128, 64, 128, 64};
constexpr int kVarTK[kGemmVariants] = {kTK, kTK, kTK, kTK2, kTK2, kTK2, kTK2};

bool gemm_fits(int v, int M, int N, int K, int kchunk) {
    return M % kVarTM[v] == 0 && N % kVarTN[v] == 0 && K % kVarTK[v] == 0 && kchunk % kVarTK[v] == 0;
}

hipError_t launch_gemm(int variant, hipStream_t st, const u16 *A, long sam, long sak, const u16 *B, long sbk,
                       long sbn, int M, int N, int K, int kchunk, const Epi &epi) {
    const int splits = (K + kchunk - 1) / kchunk;
    if (variant == 3)
        hipLaunchKernelGGL((k_gemm2<128, 128>), dim3(N / 128, M / 128, splits), dim3(256), 0, st, A, sam, sak, B,
                           sbk, sbn, M, N, K, kchunk, epi);
    else if (variant == 4)
        hipLaunchKernelGGL((k_gemm2<128, 64>), dim3(N / 64, M / 128, splits), dim3(256), 0, st, A, sam, sak, B, sbk,
                           sbn, M, N, K, kchunk, epi);
    else if (variant == 5)
        hipLaunchKernelGGL((k_gemm2<64, 128>), dim3(N / 128, M / 64, splits), dim3(256), 0, st, A, sam, sak, B, sbk,
                           sbn, M, N, K, kchunk, epi);
    else if (variant == 6)
        hipLaunchKernelGGL((k_gemm2<64, 64>), dim3(N / 64, M / 64, splits), dim3(256), 0, st, A, sam, sak, B, sbk,
                           sbn, M, N, K, kchunk, epi);
    else if (variant == 2)
        hipLaunchKernelGGL((k_gemm<128, 128>), dim3(N / 128, M / 128, splits), dim3(256), 0, st, A, sam, sak, B, sbk,
                           sbn, M, N, K, kchunk, epi);
    else if (variant == 1)
        hipLaunchKernelGGL((k_gemm<128, 64>), dim3(N / 64, M / 128, splits), dim3(256), 0, st, A, sam, sak, B, sbk,
                           sbn, M, N, K, kchunk, epi);
    else
        hipLaunchKernelGGL((k_gemm<64, 64>), dim3(N / 64, M / 64, splits), dim3(256), 0, st, A, sam, sak, B, sbk,
                           sbn, M, N, K, kchunk, epi);
    return hipGetLastError();
}

// Workgroup tile per GEMM shape, chosen by timing the admissible tiles once
// (each output element sums its k in the same order under every tile, so
// the choice changes speed only).
hipError_t gemm(o2f_mlp *o, const u16 *A, long sam, long sak, const u16 *B, long sbk, long sbn, int M, int N, int K,
                int kchunk, const Epi &epi) {
    // the tile loaders need one unit stride per operand and whole tiles
    if (!((sak == 1 || sam == 1) && (sbk == 1 || sbn == 1)) || !gemm_fits(0, M, N, K, kchunk))
        return hipErrorInvalidValue;
    if (o->gemm_variant >= 0) {  // forced: that kernel where its tile fits, else the 64 x 64 one
        const int v = gemm_fits(o->gemm_variant, M, N, K, kchunk) ? o->gemm_variant : 0;
        return launch_gemm(v, o->stream, A, sam, sak, B, sbk, sbn, M, N, K, kchunk, epi);
    }
    const auto key = std::make_tuple(M, N, K, kchunk, sak == 1, sbk == 1, epi.mode);
    auto it = o->tile_choice.find(key);
    if (it == o->tile_choice.end()) {
        int best = 0;
        float best_ms = -1.f;
        hipEvent_t e0, e1;
        hipError_t e;
        if ((e = hipEventCreate(&e0)) != hipSuccess || (e = hipEventCreate(&e1)) != hipSuccess) return e;
        for (int v = 0; v < kGemmVariants; ++v) {
            if (!gemm_fits(v, M, N, K, kchunk)) continue;
            if ((e = launch_gemm(v, o->stream, A, sam, sak, B, sbk, sbn, M, N, K, kchunk, epi)) != hipSuccess) return e;
            (void)hipEventRecord(e0, o->stream);
            for (int r = 0; r < 3; ++r)
                if ((e = launch_gemm(v, o->stream, A, sam, sak, B, sbk, sbn, M, N, K, kchunk, epi)) != hipSuccess)
                    return e;
            (void)hipEventRecord(e1, o->stream);
            if ((e = hipEventSynchronize(e1)) != hipSuccess) return e;
            float ms = 0.f;
            (void)hipEventElapsedTime(&ms, e0, e1);
            if (best_ms < 0.f || ms < best_ms) {
                best_ms = ms;
                best = v;
            }
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        it = o->tile_choice.emplace(key, best).first;
    }
    return launch_gemm(it->second, o->stream, A, sam, sak, B, sbk, sbn, M, N, K, kchunk, epi);
}

// Forward pass over `rows` rows of X[0] (a multiple of 64).
int forward(o2f_mlp *o, int rows) {
    for (int l = 0; l < kLayers; ++l) {
        Epi e{};
        e.ldc = o->N[l];
        e.bias = o->w + o->boff[l];
        if (l < kLayers - 1) {
            e.mode = kEpiHidden;
            e.Z = o->Z[l];
            e.act = o->X[l + 1];
        } else {
            e.mode = kEpiOut;
            e.C = o->Y;
        }
        MCHK(o, gemm(o, o->X[l], o->K[l], 1, o->w16 + o->woff[l], o->N[l], 1, rows, o->N[l], o->K[l], o->K[l], e));
    }
    return O2F_OK;
}

// Loss and gradient of the `n` real rows of a forward pass over `rows` rows.
int backward(o2f_mlp *o, int n, int rows) {
    hipStream_t s = o->stream;
    const int nb = (int)nblk(rows);
    hipLaunchKernelGGL(k_loss, dim3(nb), dim3(256), 0, s, o->Y, o->T, n, rows, o->dY, o->lparts);
    MCHK(o, hipGetLastError());
    hipLaunchKernelGGL(k_loss_final, dim3(1), dim3(64), 0, s, o->lparts, nb, n, o->lout);
    MCHK(o, hipGetLastError());
    const u16 *dZ = o->dY;  // gradient at the output of layer l
    u16 *bufs[2] = {o->dA, o->dB};
    for (int l = kLayers - 1; l >= 0; --l) {
        const int Kl = o->K[l], Nl = o->N[l];
        // dW_l = X_l^T dZ (split over samples, fixed-order sum), db_l = column sums
        const int splits = (rows + kSplitChunk - 1) / kSplitChunk;
        Epi ep{};
        ep.mode = kEpiPartial;
        ep.C = o->split;
        ep.ldc = Nl;
        MCHK(o, gemm(o, o->X[l], 1, Kl, dZ, Nl, 1, Kl, Nl, rows, kSplitChunk, ep));
        hipLaunchKernelGGL(k_sum_splits, dim3(nblk((long)Kl * Nl)), dim3(256), 0, s, o->split, splits, (long)Kl * Nl,
                           o->g + o->woff[l]);
        MCHK(o, hipGetLastError());
        // db_l = 1^T dZ: the same split GEMM with a ones column as the left operand
        MCHK(o, gemm(o, o->ones, 1, kIn, dZ, Nl, 1, kIn, Nl, rows, kSplitChunk, ep));
        hipLaunchKernelGGL(k_sum_splits, dim3(nblk((long)kIn * Nl)), dim3(256), 0, s, o->split, splits,
                           (long)kIn * Nl, o->dbtmp);
        MCHK(o, hipGetLastError());
        MCHK(o, hipMemcpyAsync(o->g + o->boff[l], o->dbtmp, Nl * sizeof(float), hipMemcpyDeviceToDevice, s));
        if (l == 0) break;
        // dZ_{l-1} = (dZ W_l^T) * silu'(Z_{l-1})
        Epi eb{};
        eb.mode = kEpiBack;
        eb.Zprev = o->Z[l - 1];
        eb.act = bufs[l & 1];
        eb.ldc = Kl;
        MCHK(o, gemm(o, dZ, Nl, 1, o->w16 + o->woff[l], 1, Nl, rows, Kl, Nl, Nl, eb));
        dZ = bufs[l & 1];
    }
    return O2F_OK;
}

int refresh_bf16(o2f_mlp *o) {
    hipLaunchKernelGGL(k_to_bf16, dim3(nblk(o->nparam)), dim3(256), 0, o->stream, o->w, o->nparam, o->w16);
    MCHK(o, hipGetLastError());
    return O2F_OK;
}

int load_rows(o2f_mlp *o, const float *raw_dev, const float *tgt_dev, long ndata, long first, int n, int rows,
              long step, unsigned long long seed) {
    hipLaunchKernelGGL(k_batch, dim3(nblk(rows)), dim3(256), 0, o->stream, raw_dev, tgt_dev, ndata, first, n, rows,
                       step, seed, o->fb, o->X[0], o->T);
    MCHK(o, hipGetLastError());
    return O2F_OK;
}

int round64(long n) { return (int)((n + 63) / 64 * 64); }

}  // namespace

extern "C" {

int o2f_mlp_set_gemm_variant(o2f_mlp *o, int variant) {
    if (!o || variant < -1 || variant >= kGemmVariants) return O2F_ERR_ARG;
    o->gemm_variant = variant;
    return O2F_OK;
}

int o2f_mlp_create(unsigned long long seed, o2f_mlp **out) {
    if (!out) return O2F_ERR_ARG;
    *out = nullptr;
    o2f_mlp *o = new o2f_mlp();
    long off = 0;
    for (int l = 0; l < kLayers; ++l) {
        o->woff[l] = off;
        off += (long)o->K[l] * o->N[l];
    }
    for (int l = 0; l < kLayers; ++l) {
        o->boff[l] = off;
        off += o->N[l];
    }
    o->nparam = off;
    // He-style uniform init from a fixed seed; the padded output columns stay 0
    std::mt19937_64 rng(seed);
    std::vector<float> w(off, 0.f), fb(2 * kFourier);
    std::vector<unsigned char> frozen(off, 0);
    for (int l = 0; l < kLayers; ++l) {
        const float a = std::sqrt(6.0f / o->K[l]) * (l == kLayers - 1 ? 0.1f : 1.0f);
        std::uniform_real_distribution<float> u(-a, a);
        for (long i = 0; i < (long)o->K[l] * o->N[l]; ++i) {
            const int col = (int)(i % o->N[l]);
            if (l == kLayers - 1 && col >= 4) {
                frozen[o->woff[l] + i] = 1;
                continue;
            }
            w[o->woff[l] + i] = u(rng);
        }
    }
    for (int c = 4; c < kOut; ++c) frozen[o->boff[kLayers - 1] + c] = 1;
    std::normal_distribution<float> nd(0.f, 1.f);
    for (auto &f : fb) f = nd(rng);
    bool ok = hipStreamCreate(&o->stream) == hipSuccess && upload(o->w, w) == hipSuccess &&
              upload(o->frozen, frozen) == hipSuccess && upload(o->fb, fb) == hipSuccess &&
              dalloc(o->m, off) == hipSuccess && dalloc(o->v, off) == hipSuccess &&
              dalloc(o->g, off) == hipSuccess && dalloc(o->w16, off) == hipSuccess &&
              dalloc(o->lout, 1) == hipSuccess && hipMemset(o->m, 0, off * sizeof(float)) == hipSuccess &&
              hipMemset(o->v, 0, off * sizeof(float)) == hipSuccess &&
              hipMemset(o->g, 0, off * sizeof(float)) == hipSuccess && refresh_bf16(o) == O2F_OK &&
              hipStreamSynchronize(o->stream) == hipSuccess;
    if (!ok) {
        o2f_mlp_destroy(o);
        return O2F_ERR_HIP;
    }
    *out = o;
    return O2F_OK;
}

void o2f_mlp_destroy(o2f_mlp *o) {
    if (!o) return;
    if (o->stream) (void)hipStreamSynchronize(o->stream);
    for (auto *p : o->X) (void)hipFree(p);
    for (auto *p : o->Z) (void)hipFree(p);
    for (void *p : {(void *)o->w, (void *)o->m, (void *)o->v, (void *)o->g, (void *)o->fb, (void *)o->frozen,
                    (void *)o->w16, (void *)o->dA, (void *)o->dB, (void *)o->dY, (void *)o->Y, (void *)o->T,
                    (void *)o->split, (void *)o->lparts, (void *)o->lout, (void *)o->raw, (void *)o->tgt})
        (void)hipFree(p);
    if (o->stream) (void)hipStreamDestroy(o->stream);
    delete o;
}

const char *o2f_mlp_last_error(const o2f_mlp *o) { return o ? o->err.c_str() : "null network"; }

int o2f_mlp_num_params(const o2f_mlp *o, long *n, int *layout) {
    if (!o || !n) return O2F_ERR_ARG;
    *n = o->nparam;
    if (layout)
        for (int l = 0; l < kLayers; ++l) {
            layout[4 * l] = o->K[l];
            layout[4 * l + 1] = o->N[l];
            layout[4 * l + 2] = (int)o->woff[l];
            layout[4 * l + 3] = (int)o->boff[l];
        }
    return O2F_OK;
}

int o2f_mlp_get(o2f_mlp *o, float *weights, float *fourier) {
    if (!o) return O2F_ERR_ARG;
    if (weights) MCHK(o, hipMemcpy(weights, o->w, o->nparam * sizeof(float), hipMemcpyDeviceToHost));
    if (fourier) MCHK(o, hipMemcpy(fourier, o->fb, 2 * kFourier * sizeof(float), hipMemcpyDeviceToHost));
    return O2F_OK;
}

int o2f_mlp_set(o2f_mlp *o, const float *weights) {
    if (!o || !weights) return O2F_ERR_ARG;
    MCHK(o, hipMemcpy(o->w, weights, o->nparam * sizeof(float), hipMemcpyHostToDevice));
    int st = refresh_bf16(o);
    if (st) return st;
    MCHK(o, hipStreamSynchronize(o->stream));
    return O2F_OK;
}

int o2f_mlp_set_data(o2f_mlp *o, const float *raw, const float *targets, long n) {
    if (!o || !raw || !targets || n <= 0) return O2F_ERR_ARG;
    (void)hipFree(o->raw);
    (void)hipFree(o->tgt);
    o->raw = o->tgt = nullptr;
    MCHK(o, dalloc(o->raw, kRaw * n));
    MCHK(o, dalloc(o->tgt, 4 * n));
    MCHK(o, hipMemcpy(o->raw, raw, kRaw * n * sizeof(float), hipMemcpyHostToDevice));
    MCHK(o, hipMemcpy(o->tgt, targets, 4 * n * sizeof(float), hipMemcpyHostToDevice));
    o->ndata = n;
    return O2F_OK;
}

int o2f_mlp_forward(o2f_mlp *o, const float *raw, long n, float *out) {
    if (!o || !raw || !out || n <= 0) return O2F_ERR_ARG;
    const int rows = round64(n);
    int st = ensure_cap(o, rows);
    if (st) return st;
    float *d = nullptr;
    MCHK(o, dalloc(d, kRaw * n));
    MCHK(o, hipMemcpy(d, raw, kRaw * n * sizeof(float), hipMemcpyHostToDevice));
    st = load_rows(o, d, nullptr, n, 0, (int)n, rows, -1, 0);
    if (!st) st = forward(o, rows);
    std::vector<float> y((size_t)rows * kOut);
    if (!st) MCHK(o, hipMemcpyAsync(y.data(), o->Y, y.size() * sizeof(float), hipMemcpyDeviceToHost, o->stream));
    MCHK(o, hipStreamSynchronize(o->stream));
    (void)hipFree(d);
    if (st) return st;
    for (long i = 0; i < n; ++i)
        for (int q = 0; q < 4; ++q) out[4 * i + q] = y[i * kOut + q];
    return O2F_OK;
}

int o2f_mlp_loss_grad(o2f_mlp *o, const float *raw, const float *targets, long n, float *loss, float *grad) {
    if (!o || !raw || !targets || n <= 0) return O2F_ERR_ARG;
    const int rows = round64(n);
    int st = ensure_cap(o, rows);
    if (st) return st;
    float *d = nullptr, *t = nullptr;
    MCHK(o, dalloc(d, kRaw * n));
    MCHK(o, dalloc(t, 4 * n));
    MCHK(o, hipMemcpy(d, raw, kRaw * n * sizeof(float), hipMemcpyHostToDevice));
    MCHK(o, hipMemcpy(t, targets, 4 * n * sizeof(float), hipMemcpyHostToDevice));
    st = load_rows(o, d, t, n, 0, (int)n, rows, -1, 0);
    if (!st) st = forward(o, rows);
    if (!st) st = backward(o, (int)n, rows);
    if (!st && loss) MCHK(o, hipMemcpyAsync(loss, o->lout, sizeof(float), hipMemcpyDeviceToHost, o->stream));
    if (!st && grad)
        MCHK(o, hipMemcpyAsync(grad, o->g, o->nparam * sizeof(float), hipMemcpyDeviceToHost, o->stream));
    MCHK(o, hipStreamSynchronize(o->stream));
    (void)hipFree(d);
    (void)hipFree(t);
    return st;
}

int o2f_mlp_train(o2f_mlp *o, int steps, int batch, float lr, unsigned long long seed, float *losses) {
    if (!o || steps < 0 || batch <= 0 || batch % 64 || o->ndata <= 0) return O2F_ERR_ARG;
    int st = ensure_cap(o, batch);
    if (st) return st;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    for (int k = 0; k < steps; ++k) {
        ++o->step;
        if ((st = load_rows(o, o->raw, o->tgt, o->ndata, 0, batch, batch, o->step, seed))) return st;
        if ((st = forward(o, batch))) return st;
        if ((st = backward(o, batch, batch))) return st;
        const float c1 = 1.f - std::pow(b1, (float)o->step), c2 = 1.f - std::pow(b2, (float)o->step);
        hipLaunchKernelGGL(k_adam, dim3(nblk(o->nparam)), dim3(256), 0, o->stream, o->w, o->m, o->v, o->g,
                           o->frozen, o->nparam, lr, b1, b2, eps, c1, c2);
        MCHK(o, hipGetLastError());
        if ((st = refresh_bf16(o))) return st;
        if (losses) MCHK(o, hipMemcpyAsync(losses + k, o->lout, sizeof(float), hipMemcpyDeviceToHost, o->stream));
    }
    MCHK(o, hipStreamSynchronize(o->stream));
    return O2F_OK;
}

int o2f_mlp_warm_start(o2f_mlp *o, o2f_ctx *c, double m, double p, double t, double aoa_deg) {
    if (!o || !c) return O2F_ERR_ARG;
    const HostMesh &hm = c->hm;
    const long nc = hm.nc;
    // wall distance of each cell centre: nearest wall edge segment
    std::vector<float> raw(kRaw * nc);
    for (long i = 0; i < nc; ++i) {
        const double x = hm.ccx[i], y = hm.ccy[i];
        double best = INFINITY;
        for (int e : hm.wall_edges) {
            const double ax = hm.vx[hm.eva[e]], ay = hm.vy[hm.eva[e]], bx = hm.vx[hm.evb[e]], by = hm.vy[hm.evb[e]];
            const double dx = bx - ax, dy = by - ay, l2 = dx * dx + dy * dy;
            double s = l2 > 0 ? ((x - ax) * dx + (y - ay) * dy) / l2 : 0.0;
            s = s < 0 ? 0 : s > 1 ? 1 : s;
            const double qx = ax + s * dx - x, qy = ay + s * dy - y;
            best = std::min(best, qx * qx + qy * qy);
        }
        const float f[kRaw] = {(float)x, (float)y, (float)std::sqrt(best), (float)m, (float)p, (float)t,
                               (float)c->prm.inflow_velocity, (float)aoa_deg};
        for (int k = 0; k < kRaw; ++k) raw[kRaw * i + k] = f[k];
    }
    const int rows = round64(nc);
    int st = ensure_cap(o, rows);
    if (st) return st;
    float *d = nullptr;
    MCHK(o, dalloc(d, kRaw * nc));
    MCHK(o, hipMemcpy(d, raw.data(), raw.size() * sizeof(float), hipMemcpyHostToDevice));
    st = load_rows(o, d, nullptr, nc, 0, (int)nc, rows, -1, 0);
    if (!st) st = forward(o, rows);
    const double rho = c->prm.rho_ref, pr = c->prm.p_ref, g = c->prm.gamma;
    const double a = std::sqrt(g * pr / rho);
    if (!st) {
        MCHK(o, hipStreamSynchronize(o->stream));
        MCHK(o, hipStreamSynchronize(c->stream));
        hipLaunchKernelGGL(k_denorm, dim3(nblk(nc)), dim3(256), 0, o->stream, o->Y, (int)nc, rho, a, pr,
                           c->prm.inflow_velocity, g, c->U);
        MCHK(o, hipGetLastError());
    }
    MCHK(o, hipStreamSynchronize(o->stream));
    (void)hipFree(d);
    c->speed_valid = false;
    ++c->faces_version;
    return st;
}

}  // extern "C"
