// device_common.hpp -- device helpers shared by the residual kernels (real type T = double or float).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "mstcfd.h"

namespace mstk {

constexpr int BLOCK = 256;

// physical and reference constants in the arithmetic type of the kernel
template <typename T>
struct Params {
    T gamma, gm1, eor, delta, dd, ihd;  // dd = delta^2, ihd = 0.5/delta (Harten fix)
    T astar_c;                          // 2 (gamma-1) / (gamma+1)   (SolverAusm.cpp:6-7)
    int compat;
    int noslip;  // viscous configuration: walls are no-slip (RhoSolver.cpp:160-163, 300-303)
    T qin[5];
};

template <typename T>
inline Params<T> make_params(const mst_config &c) {
    Params<T> p;
    p.gamma = (T)c.gamma;
    p.gm1 = (T)(c.gamma - 1.0);
    p.eor = (T)c.eor;
    p.delta = (T)c.entropy_delta;
    p.dd = (T)(c.entropy_delta * c.entropy_delta);
    p.ihd = (T)(0.5 / c.entropy_delta);
    p.astar_c = (T)(2.0 * (c.gamma - 1.0) / (c.gamma + 1.0));
    p.compat = c.compat;
    p.noslip = (c.mu > 0 || c.kappa > 0) ? 1 : 0;
    for (int k = 0; k < 5; k++) p.qin[k] = (T)c.q_inlet[k];
    return p;
}

// Each XCD receives every 8th workgroup; remap so that XCD x processes one contiguous slice of
// the (Hilbert- or RCM-ordered) index space and its L2 sees neighbouring cells.
__device__ __forceinline__ int64_t block_id() {
    const uint32_t G = gridDim.x, b = blockIdx.x;
    const uint32_t q = G >> 3, r = G & 7u, x = b & 7u, s = b >> 3;
    return (int64_t)(x < r ? x * (q + 1) : r * (q + 1) + (x - r) * q) + s;
}

// nv-record loads/stores: 2 x 16 B for a 2D double state, one 16 B load for a 2D float state
template <int NV, typename S, typename T>
__device__ __forceinline__ void ldq(const S *__restrict__ p, T *q) {
    if constexpr (NV == 4 && sizeof(S) == 8) {
        const double2 a = reinterpret_cast<const double2 *>(p)[0];
        const double2 b = reinterpret_cast<const double2 *>(p)[1];
        q[0] = (T)a.x; q[1] = (T)a.y; q[2] = (T)b.x; q[3] = (T)b.y;
    } else if constexpr (NV == 4 && sizeof(S) == 4) {
        const float4 a = reinterpret_cast<const float4 *>(p)[0];
        q[0] = (T)a.x; q[1] = (T)a.y; q[2] = (T)a.z; q[3] = (T)a.w;
    } else {
#pragma unroll
        for (int k = 0; k < NV; k++) q[k] = (T)p[k];
    }
}
template <int NV, typename S, typename T>
__device__ __forceinline__ void stq(S *__restrict__ p, const T *q) {
    if constexpr (NV == 4 && sizeof(S) == 8) {
        reinterpret_cast<double2 *>(p)[0] = make_double2((double)q[0], (double)q[1]);
        reinterpret_cast<double2 *>(p)[1] = make_double2((double)q[2], (double)q[3]);
    } else if constexpr (NV == 4 && sizeof(S) == 4) {
        reinterpret_cast<float4 *>(p)[0] = make_float4((float)q[0], (float)q[1], (float)q[2], (float)q[3]);
    } else {
#pragma unroll
        for (int k = 0; k < NV; k++) p[k] = (S)q[k];
    }
}

// 1/x.  double: the hardware reciprocal estimate refined by two Newton steps (the seed the
// compiler's IEEE division expands from, without the scale/fixup stages), within 1 ulp for the
// positive normal arguments it is used on; float: the hardware reciprocal (1 ulp).
__device__ __forceinline__ double frcp(double x) {
    double r = __builtin_amdgcn_rcp(x);
    double e = fma(-x, r, 1.0);
    r = fma(r, e, r);
    e = fma(-x, r, 1.0);
    return fma(r, e, r);
}
__device__ __forceinline__ float frcp(float x) { return __builtin_amdgcn_rcpf(x); }

__device__ __forceinline__ double fsqrt(double x) { return sqrt(x); }
__device__ __forceinline__ float fsqrt(float x) { return __builtin_amdgcn_sqrtf(x); }

// a / b: exact division in double, reciprocal product in float
__device__ __forceinline__ double fdiv(double a, double b) { return a / b; }
__device__ __forceinline__ float fdiv(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }

// Q + lg . dx  (RhoSolver.cpp:250,254 with the limiter factor folded into lg)
template <int DIM, typename T, typename G>
__device__ __forceinline__ void recon(T *q, const G *__restrict__ lg, const T *dx) {
    constexpr int NV = DIM + 2;
#pragma unroll
    for (int k = 0; k < NV; k++) {
        T s = (T)lg[k * DIM + 0] * dx[0] + (T)lg[k * DIM + 1] * dx[1];
        if constexpr (DIM == 3) s += (T)lg[k * DIM + 2] * dx[2];
        q[k] = q[k] + s;
    }
}

// wall/symmetry ghost momentum m - 2 n^ (n^ . m) from src, n^ = n/|n| (RhoSolver.cpp:150-158)
template <int DIM, typename T>
__device__ __forceinline__ void reflect(T *dst, const T *src, const T *n) {
    T nn = n[0] * n[0] + n[1] * n[1];
    if constexpr (DIM == 3) nn += n[2] * n[2];
    const T nrm = fsqrt(nn);
    T nh[DIM];
#pragma unroll
    for (int d = 0; d < DIM; d++) nh[d] = fdiv(n[d], nrm);
    T s = nh[0] * src[1] + nh[1] * src[2];
    if constexpr (DIM == 3) s += nh[2] * src[3];
#pragma unroll
    for (int d = 0; d < DIM; d++) dst[1 + d] = src[1 + d] - T(2) * nh[d] * s;
}

template <typename T>
__device__ __forceinline__ T hfix(T x, const Params<T> &P) {  // SolverRoe.cpp:114-123
    return x > P.delta ? x : (x * x + P.dd) * P.ihd;
}

}  // namespace mstk
