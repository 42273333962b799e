// Inline device helpers shared by the forward kernels (o2f_kernels.hip)
// and the exact adjoint (o2f_adjoint.hip): min/max semantics, AoS loads,
// boundary states, HLLC.
#pragma once

#include <hip/hip_runtime.h>

#include <cfloat>

#include "o2f_kernels.hpp"

namespace o2f {

// Forward-mode dual number (value, one directional derivative): the HLLC
// and stress Jacobians of the exact adjoint are evaluated with it, from the
// same templated code as the forward values.
struct Dual {
    double v, d;
    Dual() = default;
    __host__ __device__ constexpr Dual(double v_, double d_) : v(v_), d(d_) {}
    __host__ __device__ constexpr explicit Dual(double v_) : v(v_), d(0.0) {}
};
__device__ __forceinline__ Dual operator+(Dual a, Dual b) { return {a.v + b.v, a.d + b.d}; }
__device__ __forceinline__ Dual operator-(Dual a, Dual b) { return {a.v - b.v, a.d - b.d}; }
__device__ __forceinline__ Dual operator-(Dual a) { return {-a.v, -a.d}; }
__device__ __forceinline__ Dual operator*(Dual a, Dual b) { return {a.v * b.v, a.d * b.v + a.v * b.d}; }
__device__ __forceinline__ Dual operator/(Dual a, Dual b) {
    const double q = a.v / b.v;
    return {q, (a.d - q * b.d) / b.v};
}
__device__ __forceinline__ Dual operator+(Dual a, double b) { return {a.v + b, a.d}; }
__device__ __forceinline__ Dual operator+(double a, Dual b) { return {a + b.v, b.d}; }
__device__ __forceinline__ Dual operator-(Dual a, double b) { return {a.v - b, a.d}; }
__device__ __forceinline__ Dual operator-(double a, Dual b) { return {a - b.v, -b.d}; }
__device__ __forceinline__ Dual operator*(Dual a, double b) { return {a.v * b, a.d * b}; }
__device__ __forceinline__ Dual operator*(double a, Dual b) { return {a * b.v, a * b.d}; }
__device__ __forceinline__ Dual operator/(Dual a, double b) { return {a.v / b, a.d / b}; }
__device__ __forceinline__ Dual operator/(double a, Dual b) {
    const double q = a / b.v;
    return {q, -q * b.d / b.v};
}
__device__ __forceinline__ bool operator<(Dual a, Dual b) { return a.v < b.v; }
__device__ __forceinline__ bool operator<(Dual a, double b) { return a.v < b; }
__device__ __forceinline__ bool operator<=(Dual a, double b) { return a.v <= b; }
__device__ __forceinline__ bool operator>=(Dual a, double b) { return a.v >= b; }
__device__ __forceinline__ bool operator>(Dual a, double b) { return a.v > b; }
__device__ __forceinline__ bool operator>(Dual a, Dual b) { return a.v > b.v; }
using ::sqrt;  // the double overloads stay visible next to sqrt(Dual)
__device__ __forceinline__ Dual sqrt(Dual a) {
    const double r = ::sqrt(a.v);
    return {r, a.d / (2.0 * r)};
}

// Several tangent directions carried together: each component follows
// exactly the arithmetic of a Dual (same operations in the same order), so
// DualN<N> gives the N tangents of N separate Dual evaluations bit for
// bit, with the value computed once.
template <int N>
struct DualN {
    double v, d[N];
    // value a, tangent 1 in direction `dir` (no direction: dir < 0)
    __device__ static DualN seed(double a, int dir) {
        DualN r;
        r.v = a;
#pragma unroll
        for (int i = 0; i < N; ++i) r.d[i] = (i == dir) ? 1.0 : 0.0;
        return r;
    }
};
#define O2F_DUALN_OP(EXPR_V, EXPR_D)          \
    DualN<N> r;                               \
    r.v = EXPR_V;                             \
    _Pragma("unroll") for (int i = 0; i < N; ++i) r.d[i] = EXPR_D; \
    return r;
template <int N>
__device__ __forceinline__ DualN<N> operator+(const DualN<N> &a, const DualN<N> &b) {
    O2F_DUALN_OP(a.v + b.v, a.d[i] + b.d[i])
}
template <int N>
__device__ __forceinline__ DualN<N> operator-(const DualN<N> &a, const DualN<N> &b) {
    O2F_DUALN_OP(a.v - b.v, a.d[i] - b.d[i])
}
template <int N>
__device__ __forceinline__ DualN<N> operator-(const DualN<N> &a) {
    O2F_DUALN_OP(-a.v, -a.d[i])
}
template <int N>
__device__ __forceinline__ DualN<N> operator*(const DualN<N> &a, const DualN<N> &b) {
    O2F_DUALN_OP(a.v * b.v, a.d[i] * b.v + a.v * b.d[i])
}
template <int N>
__device__ __forceinline__ DualN<N> operator/(const DualN<N> &a, const DualN<N> &b) {
    const double q = a.v / b.v;
    O2F_DUALN_OP(q, (a.d[i] - q * b.d[i]) / b.v)
}
template <int N>
__device__ __forceinline__ DualN<N> operator+(const DualN<N> &a, double b) {
    O2F_DUALN_OP(a.v + b, a.d[i])
}
template <int N>
__device__ __forceinline__ DualN<N> operator+(double a, const DualN<N> &b) {
    O2F_DUALN_OP(a + b.v, b.d[i])
}
template <int N>
__device__ __forceinline__ DualN<N> operator-(const DualN<N> &a, double b) {
    O2F_DUALN_OP(a.v - b, a.d[i])
}
template <int N>
__device__ __forceinline__ DualN<N> operator-(double a, const DualN<N> &b) {
    O2F_DUALN_OP(a - b.v, -b.d[i])
}
template <int N>
__device__ __forceinline__ DualN<N> operator*(const DualN<N> &a, double b) {
    O2F_DUALN_OP(a.v * b, a.d[i] * b)
}
template <int N>
__device__ __forceinline__ DualN<N> operator*(double a, const DualN<N> &b) {
    O2F_DUALN_OP(a * b.v, a * b.d[i])
}
template <int N>
__device__ __forceinline__ DualN<N> operator/(const DualN<N> &a, double b) {
    O2F_DUALN_OP(a.v / b, a.d[i] / b)
}
template <int N>
__device__ __forceinline__ DualN<N> operator/(double a, const DualN<N> &b) {
    const double q = a / b.v;
    O2F_DUALN_OP(q, -q * b.d[i] / b.v)
}
#undef O2F_DUALN_OP
template <int N>
__device__ __forceinline__ bool operator<(const DualN<N> &a, const DualN<N> &b) { return a.v < b.v; }
template <int N>
__device__ __forceinline__ bool operator<=(const DualN<N> &a, double b) { return a.v <= b; }
template <int N>
__device__ __forceinline__ bool operator>=(const DualN<N> &a, double b) { return a.v >= b; }
template <int N>
__device__ __forceinline__ DualN<N> sqrt(const DualN<N> &a) {
    const double r = ::sqrt(a.v);
    DualN<N> o;
    o.v = r;
#pragma unroll
    for (int i = 0; i < N; ++i) o.d[i] = a.d[i] / (2.0 * r);
    return o;
}

// std::min / std::max / std::clamp semantics (NaN handling included).
__device__ __forceinline__ double smin(double a, double b) { return (b < a) ? b : a; }
__device__ __forceinline__ double smax(double a, double b) { return (a < b) ? b : a; }
__device__ __forceinline__ double sclamp(double v, double lo, double hi) {
    return (v < lo) ? lo : (hi < v) ? hi : v;
}
template <class T>
__device__ __forceinline__ T smin_t(T a, T b) {
    return (b < a) ? b : a;
}
template <class T>
__device__ __forceinline__ T smax_t(T a, T b) {
    return (a < b) ? b : a;
}
__device__ __forceinline__ bool tagged(char b) {
    return b == 'i' || b == 'o' || b == 'f' || b == 'w';
}

// Per-stage intermediates (vertex values, face states, edge fluxes, cell and
// vertex stresses) are stored AoS, four doubles per entity (a fifth
// component, q_y, sits in a trailing SoA block): one gathered entity is one
// 32-byte run of a cache line instead of four scattered 8-byte words.
__device__ __forceinline__ void ld4(const double *__restrict__ p, long i, double (&v)[4]) {
    const double2 *q = reinterpret_cast<const double2 *>(p) + 2 * i;
    const double2 a = q[0], b = q[1];
    v[0] = a.x;
    v[1] = a.y;
    v[2] = b.x;
    v[3] = b.y;
}
__device__ __forceinline__ void st4(double *__restrict__ p, long i, const double (&v)[4]) {
    double2 *q = reinterpret_cast<double2 *>(p) + 2 * i;
    q[0] = make_double2(v[0], v[1]);
    q[1] = make_double2(v[2], v[3]);
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
// T: the state type; N: the normal's type (double, or Dual when the
// geometry itself is differentiated).
template <class T, class N = double>
__device__ __forceinline__ void hllc_t(const T *qL, const T *qR, N enx, N eny, double gamma, T *F) {
    const N nx = -enx, ny = -eny;
    const T rho_L = qL[0];
    const T u_L = (qL[1] * (nx) + qL[2] * ny) / rho_L;
    const T v_L = (qL[1] * (-ny) + qL[2] * nx) / rho_L;
    const T E_L = qL[3];
    const T p_L = (gamma - 1.0) * (E_L - 0.5 * rho_L * (u_L * u_L + v_L * v_L));
    const T c_L = sqrt(gamma * p_L / rho_L);
    const T rho_R = qR[0];
    const T u_R = (qR[1] * (nx) + qR[2] * ny) / rho_R;
    const T v_R = (qR[1] * (-ny) + qR[2] * nx) / rho_R;
    const T E_R = qR[3];
    const T p_R = (gamma - 1.0) * (E_R - 0.5 * rho_R * (u_R * u_R + v_R * v_R));
    const T c_R = sqrt(gamma * p_R / rho_R);

    const T S_L = smin_t(u_L - c_L, u_R - c_R);
    const T S_R = smax_t(u_L + c_L, u_R + c_R);
    const T numer = p_R - p_L + rho_L * u_L * (S_L - u_L) - rho_R * u_R * (S_R - u_R);
    const T denom = rho_L * (S_L - u_L) - rho_R * (S_R - u_R);
    const T S_star = numer / denom;
    const T p_star = p_L + rho_L * (S_L - u_L) * (S_star - u_L);
    T f0, fu, fv, fE;
    if (S_L >= 0.0) {
        f0 = rho_L * u_L;
        fu = rho_L * u_L * u_L + p_L;
        fv = rho_L * u_L * v_L;
        fE = (E_L + p_L) * u_L;
    } else if (S_L <= 0.0 && S_star >= 0.0) {
        const T rs = rho_L * (S_L - u_L) / (S_L - S_star);
        const T rvs = rs * v_L;
        const T Es = ((S_L - u_L) * E_L - p_L * u_L + p_star * S_star) / (S_L - S_star);
        const T rus = rs * S_star;
        f0 = rus;
        fu = rus * rus / rs + p_star;
        fv = rvs * rus / rs;
        fE = (Es + p_star) * rus / rs;
    } else if (S_star <= 0.0 && S_R >= 0.0) {
        const T rs = rho_R * (S_R - u_R) / (S_R - S_star);
        const T rvs = rs * v_R;
        const T Es = ((S_R - u_R) * E_R - p_R * u_R + p_star * S_star) / (S_R - S_star);
        const T rus = rs * S_star;
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


__device__ __forceinline__ void hllc(const double *qL, const double *qR, double enx, double eny, double gamma,
                                     double *F) {
    hllc_t<double>(qL, qR, enx, eny, gamma, F);
}

// Limiters of computeInterfaceStates (src/advection.cpp:195-230).
// The limiter's eps = K^3 h^3, h = sqrt(V / pi): a function of the cell
// alone, so a cell's edges can share one evaluation (same expression, same
// rounding as computing it per edge).
__device__ __forceinline__ double venkat_eps(double cellArea, double k_eps) {
    const double h = sqrt(cellArea / M_PI);
    return k_eps * k_eps * k_eps * h * h * h;
}

__device__ __forceinline__ double alpha_venkat_eps(double phi_i, double phi_pred, double phi_min,
                                                   double phi_max, double eps) {
    const double delta = phi_pred - phi_i;
    if (fabs(delta) < DBL_EPSILON) return 1.0;
    const double a = (delta > 0.0) ? (phi_max - phi_i) : (phi_min - phi_i);
    const double alpha = (a * a + eps + 2.0 * a * delta) / (a * a + eps + 2.0 * delta * delta);
    return sclamp(alpha, 0.0, 1.0);
}

__device__ __forceinline__ double alpha_venkat(double phi_i, double phi_pred, double phi_min,
                                               double phi_max, double cellArea, double k_eps) {
    return alpha_venkat_eps(phi_i, phi_pred, phi_min, phi_max, venkat_eps(cellArea, k_eps));
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

// Viscous stress and heat flux of one cell (src/diffusion.cpp:227-240), the
// arithmetic of the forward reconstruction kernel, on any scalar type.
template <class T>
__device__ __forceinline__ void cell_stress_t(const T (&u)[4], const T (&gx)[4], const T (&gy)[4], double mu,
                                              double gamma, double kappa, T (&out)[5]) {
    const T rho = u[0], E = u[3];
    const T uu = u[1] / rho, vvel = u[2] / rho;
    const T u_x = (gx[1] - uu * gx[0]) / rho;
    const T v_y = (gy[2] - vvel * gy[0]) / rho;
    const T u_y = (gy[1] - uu * gy[0]) / rho;
    const T v_x = (gx[2] - vvel * gx[0]) / rho;
    const T T_x = (gamma - 1.0) * ((gx[3] * rho - E * gx[0]) / (rho * rho) - (uu * u_x + vvel * v_x));
    const T T_y = (gamma - 1.0) * ((gy[3] * rho - E * gy[0]) / (rho * rho) - (uu * u_y + vvel * v_y));
    out[0] = 2 * mu * u_x - 2.0 / 3.0 * mu * (u_x + v_y);
    out[1] = 2 * mu * v_y - 2.0 / 3.0 * mu * (u_x + v_y);
    out[2] = mu * (u_y + v_x);
    out[3] = -kappa * T_x;
    out[4] = -kappa * T_y;
}

// Value part of a double or a Dual (branch decisions follow the value).
__device__ __forceinline__ double vpart(double a) { return a; }
__device__ __forceinline__ double vpart(Dual a) { return a.v; }

// The limiters, the state boundary conditions and the ghost state on any
// scalar type (the double versions above are what the forward kernels use;
// these are the same expressions for the geometry tangent).
template <class T>
__device__ __forceinline__ T alpha_venkat_g(T phi_i, T phi_pred, T phi_min, T phi_max, T cellArea, double k_eps) {
    const T delta = phi_pred - phi_i;
    if (fabs(vpart(delta)) < DBL_EPSILON) return T(1.0);
    const T h = sqrt(cellArea / M_PI);
    const T eps = k_eps * k_eps * k_eps * h * h * h;
    const T a = (vpart(delta) > 0.0) ? (phi_max - phi_i) : (phi_min - phi_i);
    const T alpha = (a * a + eps + 2.0 * a * delta) / (a * a + eps + 2.0 * delta * delta);
    return (vpart(alpha) < 0.0) ? T(0.0) : (1.0 < vpart(alpha)) ? T(1.0) : alpha;
}
template <class T>
__device__ __forceinline__ T alpha_bj_g(T phi_i, T phi_pred, T phi_min, T phi_max) {
    T alpha = T(1.0);
    if (fabs(vpart(phi_pred - phi_i)) < DBL_EPSILON) return alpha;
    if (vpart(phi_pred) > vpart(phi_i))
        alpha = smin_t(T(1.0), (phi_max - phi_i) / (phi_pred - phi_i));
    else if (vpart(phi_pred) < vpart(phi_i))
        alpha = smin_t(T(1.0), (phi_min - phi_i) / (phi_pred - phi_i));
    return alpha;
}
template <class T>
__device__ __forceinline__ void bc_state_g(const Phys &P, char bc, T (&q)[4]) {
    if (bc == 'i' || bc == 'f') {
        q[0] = T(P.fs_rho);
        q[1] = T(P.fs_ru);
        q[2] = T(P.fs_rv);
        q[3] = T(P.fs_E);
    } else if (bc == 'w') {
        const T kinetic_old = 0.5 * (q[1] * q[1] + q[2] * q[2]) / q[0];
        const T e_internal = q[3] - kinetic_old;
        q[1] = T(0.0);
        q[2] = T(0.0);
        q[3] = e_internal;
    }
}
template <class T>
__device__ __forceinline__ void bc_ghost_g(const Phys &P, char bc, const T (&L)[4], T (&R)[4]) {
    for (int q = 0; q < 4; ++q) R[q] = L[q];
    if (bc == 'i' || bc == 'o' || bc == 'f') {
        bc_state_g(P, bc, R);
    } else if (bc == 'w') {
        R[1] = -L[1];
        R[2] = -L[2];
    }
}

// Cell kernels run four lanes per cell, one per edge slot (triangles leave
// one idle): a cell's sums over its edges are formed by every lane of the
// group from the lanes' terms in ascending slot order, i.e. exactly the
// sequential order of the reference loops, so results stay bit-identical.
constexpr int kLanesPerCell = 4;

// Value of lane J of this lane's group of four: a DPP quad_perm broadcast
// (register to register, no LDS round trip).
template <int J>
__device__ __forceinline__ double from_lane(double v) {
    const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), J * 0x55, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), J * 0x55, 0xF, 0xF, false);
    return __hiloint2double(hi, lo);
}

template <int J>
__device__ __forceinline__ int from_lane_i(int v) {
    return __builtin_amdgcn_mov_dpp(v, J * 0x55, 0xF, 0xF, false);
}

// Value of register v in the lane the quad_perm control CTRL selects.
template <int CTRL>
__device__ __forceinline__ double quad_perm(double v) {
    const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), CTRL, 0xF, 0xF, false);
    return __hiloint2double(hi, lo);
}

// 4 x 4 transpose within a group of four lanes: lane k holds row k on
// entry (m[j] = M[k][j]) and column k on exit (m[j] = M[j][k]).  Two
// butterfly stages: exchange with lane k ^ 1 the entries whose index
// differs from k in bit 0 (quad_perm [1,0,3,2]), then with lane k ^ 2 those
// differing in bit 1 (quad_perm [2,3,0,1]).
__device__ __forceinline__ void quad_transpose(double (&m)[4], int k) {
    constexpr int kXor1 = 0xB1, kXor2 = 0x4E;
    const bool odd = k & 1, upper = k & 2;
    double t0 = quad_perm<kXor1>(m[1]), t1 = quad_perm<kXor1>(m[0]);
    double t2 = quad_perm<kXor1>(m[3]), t3 = quad_perm<kXor1>(m[2]);
    m[0] = odd ? t0 : m[0];
    m[1] = odd ? m[1] : t1;
    m[2] = odd ? t2 : m[2];
    m[3] = odd ? m[3] : t3;
    t0 = quad_perm<kXor2>(m[2]);
    t1 = quad_perm<kXor2>(m[3]);
    t2 = quad_perm<kXor2>(m[0]);
    t3 = quad_perm<kXor2>(m[1]);
    m[0] = upper ? t0 : m[0];
    m[1] = upper ? t1 : m[1];
    m[2] = upper ? m[2] : t2;
    m[3] = upper ? m[3] : t3;
}

// Cell -> vertex gather with precomputed stencils
// (src/advection.cpp:26-131 for the state, src/diffusion.cpp:28-140 for the
// viscous quantities): the value of vertex v, NV = 4 state variables with
// the state BC, or NV = 5 viscous quantities with apply_diffusion_bc
// (src/boundary.cpp:101-127).
// SRC_AOS: the source is AoS (4 + trailing block), else SoA [q][nc].
template <int NV, bool SRC_AOS>
__device__ __forceinline__ double src_val(const double *__restrict__ src, int nc, int c, int q) {
    return SRC_AOS ? (q < 4 ? src[4 * (long)c + q] : src[4 * (long)nc + c]) : src[(long)q * nc + c];
}

// The vertex's stencil is (kind, offset b into scell / sw, length n), its
// first kStencilSlots cell ids are cid (entries past the stencil are never
// used) and bc_of() gives its boundary tag; store(out) receives the NV
// values.  The weights load beside the values.
template <int NV, bool SRC_AOS, class BcOf, class Store>
__device__ __forceinline__ void vertex_gather_ids(const Topo &T, const Geom &G, const double *__restrict__ src,
                                                  const Phys &P, uint8_t kind, int b, int n,
                                                  const int (&cid)[kStencilSlots], BcOf &&bc_of, Store &&store) {
    const int nc = T.nc;
    double w[kStencilSlots], val[kStencilSlots][NV];
    {
        typedef double v2d __attribute__((ext_vector_type(2)));
        typedef v2d v2d_a8 __attribute__((aligned(8)));
        const v2d_a8 *wp = reinterpret_cast<const v2d_a8 *>(G.sw + b);
#pragma unroll
        for (int h = 0; h < kStencilSlots / 2; ++h) {
            const v2d t = wp[h];
            w[2 * h] = t.x;
            w[2 * h + 1] = t.y;
        }
    }
#pragma unroll
    for (int k = 0; k < kStencilSlots; ++k)
        if (k < n) {
            if (SRC_AOS) {
                double t4[4];
                ld4(src, cid[k], t4);
#pragma unroll
                for (int q = 0; q < 4; ++q) val[k][q] = t4[q];
                if (NV == 5) val[k][NV - 1] = src[4 * (long)nc + cid[k]];
            } else {
#pragma unroll
                for (int q = 0; q < NV; ++q) val[k][q] = src[(long)q * nc + cid[k]];
            }
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
            for (int q = 0; q < NV; ++q) out[q] = out[q] + wk * src_val<NV, SRC_AOS>(src, nc, c, q);
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
    const char bc = bc_of();
    if (NV == 4) {
        if (tagged(bc)) bc_state(P, bc, out[0], out[1], out[2], out[3]);
    } else {
        if (bc == 'f' || bc == 'w') {
#pragma unroll
            for (int q = 0; q < NV; ++q) out[q] = 0.0;
        }
    }
    store(out);
}

// The same with the ids read from the stencil arrays: the first
// kStencilSlots ids and weights in 16-byte loads (both arrays carry
// kStencilSlots entries of padding), then the values of the slots k < n:
// two dependent load levels, and few address computations.  (Per-entry
// loads under k < n let the compiler finish each id's address arithmetic
// inside its branch, which waits for the ids one at a time.)
template <int NV, bool SRC_AOS, class BcOf, class Store>
__device__ __forceinline__ void vertex_gather_vals(const Topo &T, const Geom &G, const double *__restrict__ src,
                                                   const Phys &P, uint8_t kind, int b, int n, BcOf &&bc_of,
                                                   Store &&store) {
    int cid[kStencilSlots];
    typedef int v4i __attribute__((ext_vector_type(4)));
    typedef v4i v4i_a4 __attribute__((aligned(4)));
    const v4i_a4 *ip = reinterpret_cast<const v4i_a4 *>(T.scell + b);
#pragma unroll
    for (int h = 0; h < kStencilSlots / 4; ++h) {
        const v4i t = ip[h];
        cid[4 * h] = t.x;
        cid[4 * h + 1] = t.y;
        cid[4 * h + 2] = t.z;
        cid[4 * h + 3] = t.w;
    }
    vertex_gather_ids<NV, SRC_AOS>(T, G, src, P, kind, b, n, cid, bc_of, store);
}


// |u| + c of one state (src/simulation.cpp:84-95)
__device__ __forceinline__ double wave_speed(double rho, double ru, double rv, double E, double gamma) {
    const double u = ru / rho, v = rv / rho;
    const double kinetic = 0.5 * (ru * ru + rv * rv) / rho;
    const double p = (gamma - 1.0) * (E - kinetic);
    const double cs = sqrt(gamma * p / rho);
    return hypot(u, v) + cs;
}

// f(j, value of lane j) for j = 0, 1, 2, 3 in order
template <class F>
__device__ __forceinline__ void for_lanes(double v, F &&f) {
    f(0, from_lane<0>(v));
    f(1, from_lane<1>(v));
    f(2, from_lane<2>(v));
    f(3, from_lane<3>(v));
}

}  // namespace o2f
