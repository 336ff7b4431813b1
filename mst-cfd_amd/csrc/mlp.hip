// mlp.hip -- a small fully connected network trained and evaluated by hand-written kernels (learned
// warm start of steady runs; no reference counterpart).
//
// Network: n_in -> width (tanh) -> ... -> width (tanh) -> n_out (linear), FP32.  Parameters are one
// flat array: for each layer l the weights W_l [out x in] (row-major) then the biases b_l [out].
// Features are normalised as (f - in_shift) * in_scale and outputs de-normalised as
// y * out_scale + out_shift inside the kernels.
//
// Kernels: an LDS-tiled FP32 GEMM on the matrix cores (64x64 block tile, one 32x32 v_mfma_f32_32x32x2_f32
// accumulator per wave) with strided operands and fused
// epilogues (bias, tanh, tanh-derivative), split along K with a fixed-order second pass for the
// weight gradients (deterministic); column sums for the bias gradients; the squared-error gradient;
// Adam.  A training step (batch gather by a counter-based hash, forward, backward, Adam, step
// counter) reads its step number from device memory, so it is captured once into a hipGraph and
// replayed.
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "mlp.hpp"
#include "mstcfd.h"

namespace {

#define HIPCHK(expr)                                                                                     \
    do {                                                                                                 \
        hipError_t _e = (expr);                                                                          \
        if (_e != hipSuccess)                                                                            \
            return mst::fail(MST_E_HIP, std::string(#expr) + ": " + hipGetErrorString(_e));              \
    } while (0)

constexpr int TM = 64, TN = 64, TK = 16, GT = 256;
typedef float floatx16 __attribute__((ext_vector_type(16)));
enum { EPI_NONE = 0, EPI_BIAS = 1, EPI_BIAS_TANH = 2, EPI_DTANH = 3 };

struct GemmArgs {
    const float *A, *B;  // A(m,k) = A[m*sam + k*sak], B(k,n) = B[k*sbk + n*sbn]
    float *C;            // C(m,n) = C[m*ldc + n] (+ z * cz for split K)
    int64_t M, N, K;
    int64_t sam, sak, sbk, sbn, ldc;
    const float *bias;   // EPI_BIAS*: bias[n]
    const float *P;      // EPI_DTANH: C *= 1 - P(m,n)^2, P(m,n) = P[m*ldc + n]
    int64_t kchunk, cz;  // split K: block z covers k in [z*kchunk, (z+1)*kchunk)
};

template <int EPI>
__global__ __launch_bounds__(GT) void k_gemm(const GemmArgs g) {
    __shared__ float As[TK][TM + 4];
    __shared__ float Bs[TK][TN + 4];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int wm = (wv >> 1) * 32, wn = (wv & 1) * 32;  // the wave's 32 x 32 quarter of the 64 x 64 tile
    const int64_t m0 = (int64_t)blockIdx.y * TM, n0 = (int64_t)blockIdx.x * TN;
    const int64_t k0 = (int64_t)blockIdx.z * g.kchunk;
    const int64_t k1 = k0 + g.kchunk < g.K ? k0 + g.kchunk : g.K;
    floatx16 acc;
#pragma unroll
    for (int r = 0; r < 16; r++) acc[r] = 0.f;
    for (int64_t kt = k0; kt < k1; kt += TK) {
        // 1024 elements per tile, 4 per lane, zero outside the matrices
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int e = tid + r * GT;
            const int am = e % TM, ak = e / TM;  // A tile: TK x TM, m fastest
            const int64_t gm = m0 + am, gk = kt + ak;
            As[ak][am] = (gm < g.M && gk < k1) ? g.A[gm * g.sam + gk * g.sak] : 0.f;
            const int bn = e % TN, bk = e / TN;
            const int64_t gn = n0 + bn, gk2 = kt + bk;
            Bs[bk][bn] = (gn < g.N && gk2 < k1) ? g.B[gk2 * g.sbk + gn * g.sbn] : 0.f;
        }
        __syncthreads();
        // v_mfma_f32_32x32x2_f32: A(i, k) from lane i + 32k, B(k, j) from lane j + 32k
#pragma unroll
        for (int k = 0; k < TK; k += 2) {
            const float av = As[k + (lane >> 5)][wm + (lane & 31)];
            const float bv = Bs[k + (lane >> 5)][wn + (lane & 31)];
            acc = __builtin_amdgcn_mfma_f32_32x32x2f32(av, bv, acc, 0, 0, 0);
        }
        __syncthreads();
    }
    float *C = g.C + (int64_t)blockIdx.z * g.cz;
    // D(i, j) in lane j + 32 ((i / 4) % 2), register (i % 4) + 4 (i / 8)
    const int64_t n = n0 + wn + (lane & 31);
#pragma unroll
    for (int r = 0; r < 16; r++) {
        const int64_t m = m0 + wm + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
        if (m >= g.M || n >= g.N) continue;
        float v = acc[r];
        if constexpr (EPI == EPI_BIAS || EPI == EPI_BIAS_TANH) v += g.bias[n];
        if constexpr (EPI == EPI_BIAS_TANH) v = tanhf(v);
        if constexpr (EPI == EPI_DTANH) {
            const float p = g.P[m * g.ldc + n];
            v *= 1.f - p * p;
        }
        C[m * g.ldc + n] = v;
    }
}

// out[i] = sum_z part[z*n + i] in z order (second pass of a split-K product)
__global__ __launch_bounds__(GT) void k_sum_splits(int64_t n, int nz, const float *__restrict__ part,
                                                  float *__restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * GT + threadIdx.x;
    if (i >= n) return;
    float s = 0.f;
    for (int z = 0; z < nz; z++) s += part[(int64_t)z * n + i];
    out[i] = s;
}

// column sums of X [rows x cols] (bias gradients): one block per column, fixed-order tree
__global__ __launch_bounds__(GT) void k_colsum(int64_t rows, int64_t cols, const float *__restrict__ X,
                                              float *__restrict__ out) {
    __shared__ float s[GT];
    const int64_t c = blockIdx.x;
    float a = 0.f;
    for (int64_t r = threadIdx.x; r < rows; r += GT) a += X[r * cols + c];
    s[threadIdx.x] = a;
    __syncthreads();
    for (int w = GT / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[c] = s[0];
}

// counter-based hash (splitmix64 finaliser) for the batch sample indices
__device__ __forceinline__ uint64_t mix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// batch rows: normalised features and targets of sample idx(step, i) (all samples in order when the
// batch covers the data set)
__global__ __launch_bounds__(GT) void k_gather(int64_t B, int64_t n, int n_in, int n_out, const float *__restrict__ X,
                                              const float *__restrict__ T, const float *__restrict__ in_shift,
                                              const float *__restrict__ in_scale, const int64_t *__restrict__ step,
                                              uint64_t seed, float *__restrict__ Xb, float *__restrict__ Tb) {
    const int64_t i = (int64_t)blockIdx.x * GT + threadIdx.x;
    if (i >= B) return;
    const int64_t idx = B >= n ? i : (int64_t)(mix64(seed ^ mix64((uint64_t)step[0] * 0x100000001B3ull + i)) % (uint64_t)n);
    for (int k = 0; k < n_in; k++) Xb[i * n_in + k] = (X[idx * n_in + k] - in_shift[k]) * in_scale[k];
    for (int k = 0; k < n_out; k++) Tb[i * n_out + k] = T[idx * n_out + k];
}

// dY = 2 (Y*out_scale + out_shift - T) * out_scale / (B n_out), sq = (Y' - T)^2 per element
__global__ __launch_bounds__(GT) void k_loss_grad(int64_t n, int n_out, const float *__restrict__ Y,
                                                 const float *__restrict__ T, const float *__restrict__ out_shift,
                                                 const float *__restrict__ out_scale, float inv, float *__restrict__ dY,
                                                 float *__restrict__ sq) {
    const int64_t i = (int64_t)blockIdx.x * GT + threadIdx.x;
    if (i >= n) return;
    const int k = (int)(i % n_out);
    const float d = Y[i] * out_scale[k] + out_shift[k] - T[i];
    dY[i] = 2.f * d * out_scale[k] * inv;
    if (sq) sq[i] = d * d;
}

// Adam (Kingma & Ba): the step number comes from the device counter, which lane 0 advances
__global__ __launch_bounds__(GT) void k_adam(int64_t n, float *__restrict__ p, const float *__restrict__ g,
                                            float *__restrict__ m, float *__restrict__ v, int64_t *__restrict__ step,
                                            float lr, float b1, float b2, float eps) {
    const int64_t i = (int64_t)blockIdx.x * GT + threadIdx.x;
    const double t = (double)(step[0] + 1);
    if (i < n) {
        const float mi = b1 * m[i] + (1.f - b1) * g[i];
        const float vi = b2 * v[i] + (1.f - b2) * g[i] * g[i];
        m[i] = mi;
        v[i] = vi;
        const float mh = mi / (float)(1.0 - pow((double)b1, t));
        const float vh = vi / (float)(1.0 - pow((double)b2, t));
        p[i] -= lr * mh / (sqrtf(vh) + eps);
    }
}

__global__ void k_step_advance(int64_t *step) {
    if (blockIdx.x == 0 && threadIdx.x == 0) step[0] += 1;
}

// de-normalised outputs
__global__ __launch_bounds__(GT) void k_denorm(int64_t n, int n_out, const float *__restrict__ Y,
                                              const float *__restrict__ shift, const float *__restrict__ scale,
                                              float *__restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * GT + threadIdx.x;
    if (i >= n) return;
    const int k = (int)(i % n_out);
    out[i] = Y[i] * scale[k] + shift[k];
}

__global__ __launch_bounds__(GT) void k_normalise(int64_t n, int n_in, const float *__restrict__ X,
                                                 const float *__restrict__ shift, const float *__restrict__ scale,
                                                 float *__restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * GT + threadIdx.x;
    if (i >= n) return;
    const int k = (int)(i % n_in);
    out[i] = (X[i] - shift[k]) * scale[k];
}

inline unsigned nblk(int64_t n) { return (unsigned)((n + GT - 1) / GT); }

}  // namespace

struct mst_mlp {
    int n_in = 0, width = 0, hidden = 0, n_out = 0, L = 0;
    std::vector<int> din, dout;        // per layer
    std::vector<int64_t> woff, boff;   // offsets into the parameter array
    int64_t np = 0;
    hipStream_t stream = nullptr;
    float *p = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;  // params, gradient, Adam moments
    float *scal = nullptr;  // in_shift[n_in], in_scale[n_in], out_shift[n_out], out_scale[n_out]
    int64_t *step = nullptr;
    // training set
    float *X = nullptr, *T = nullptr;
    int64_t n_data = 0;
    uint64_t seed = 0;
    // work buffers for batches of up to cap rows
    int64_t cap = 0;
    std::vector<float *> H;  // H[0] = normalised input, H[l] = activations after layer l-1, H[L] = output
    std::vector<float *> D;  // gradients w.r.t. H[l]
    float *Tb = nullptr, *sq = nullptr, *split = nullptr;
    int64_t split_cap = 0;
    // graph of one training step
    hipGraphExec_t gexec = nullptr;
    int g_batch = 0;
    float g_lr = 0;
    // per-dispatch timing (mst_mlp_timing_*): start/stop events carried by each launch of a class,
    // and the class's floating-point operations per launch (GEMMs: 2 M N K)
    bool timing = false;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> ev[MST_MLP_K_COUNT];
    std::vector<double> flops[MST_MLP_K_COUNT];
    std::vector<hipEvent_t> pool;
};

namespace {

float *fptr(float *base, int64_t off) { return base + off; }

hipEvent_t take_event(mst_mlp *q) {
    if (!q->pool.empty()) {
        hipEvent_t e = q->pool.back();
        q->pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    return hipEventCreate(&e) == hipSuccess ? e : nullptr;
}

// one launch on the network's stream; with timing on, the dispatch carries start/stop events of
// class kclass (no marker packets between the kernels)
template <typename K, typename... Args>
int mlp_launch(mst_mlp *q, int kclass, double flop, K kernel, dim3 grid, dim3 block, Args... args) {
    if (q->timing) {
        hipEvent_t e0 = take_event(q), e1 = take_event(q);
        if (!e0 || !e1) return mst::fail(MST_E_HIP, "hipEventCreate failed");
        hipExtLaunchKernelGGL(kernel, grid, block, 0, q->stream, e0, e1, 0, args...);
        HIPCHK(hipGetLastError());
        q->ev[kclass].push_back({e0, e1});
        q->flops[kclass].push_back(flop);
        return MST_OK;
    }
    hipLaunchKernelGGL(kernel, grid, block, 0, q->stream, args...);
    HIPCHK(hipGetLastError());
    return MST_OK;
}

// The captured training step holds raw pointers to the training set (X, T), the split buffer and
// the work buffers, and the row count by value: whatever frees or replaces one of them drops it.
void drop_graph(mst_mlp *q) {
    if (q->gexec) (void)hipGraphExecDestroy(q->gexec);
    q->gexec = nullptr;
}

// Whether a rocprofiler-sdk tool is loaded into this process.  Under rocprofv3 --kernel-trace the
// first capture/instantiate/replay of the training step ended in a host segmentation fault inside
// the HIP runtime, before any kernel ran (cause not found); while a tool is present the training
// steps are launched directly instead (same kernels, same results).
bool profiler_tool_loaded() {
    FILE *f = fopen("/proc/self/maps", "r");
    if (!f) return false;
    char line[4096];
    bool found = false;
    while (!found && fgets(line, sizeof(line), f)) found = strstr(line, "librocprofiler-sdk.so") != nullptr;
    fclose(f);
    return found;
}

void free_work(mst_mlp *q) {
    for (auto *p : q->H)
        if (p) (void)hipFree(p);
    for (auto *p : q->D)
        if (p) (void)hipFree(p);
    q->H.assign(q->L + 1, nullptr);
    q->D.assign(q->L + 1, nullptr);
    if (q->Tb) (void)hipFree(q->Tb);
    if (q->sq) (void)hipFree(q->sq);
    q->Tb = q->sq = nullptr;
    q->cap = 0;
    drop_graph(q);
}

int ensure_work(mst_mlp *q, int64_t rows) {
    if (rows <= q->cap) return MST_OK;
    free_work(q);
    for (int l = 0; l <= q->L; l++) {
        const int w = l == 0 ? q->n_in : (l == q->L ? q->n_out : q->width);
        HIPCHK(hipMalloc((void **)&q->H[l], sizeof(float) * rows * w));
        HIPCHK(hipMalloc((void **)&q->D[l], sizeof(float) * rows * w));
    }
    HIPCHK(hipMalloc((void **)&q->Tb, sizeof(float) * rows * q->n_out));
    HIPCHK(hipMalloc((void **)&q->sq, sizeof(float) * rows * q->n_out));
    q->cap = rows;
    return MST_OK;
}

// C = A B (+ epilogue); with kchunk < K the K range is split over gridDim.z blocks writing
// C + z*cz (returned in *nz)
template <int EPI>
int gemm(mst_mlp *q, int kclass, GemmArgs a, int64_t kchunk = 0, int *nz = nullptr) {
    a.kchunk = kchunk > 0 ? kchunk : (a.K > 0 ? a.K : 1);
    const int64_t z = (a.K + a.kchunk - 1) / a.kchunk;
    dim3 grid((unsigned)((a.N + TN - 1) / TN), (unsigned)((a.M + TM - 1) / TM), (unsigned)(z > 0 ? z : 1));
    int rc = mlp_launch(q, kclass, 2.0 * (double)a.M * (double)a.N * (double)a.K, k_gemm<EPI>, grid, dim3(GT), a);
    if (rc) return rc;
    if (nz) *nz = (int)(z > 0 ? z : 1);
    return MST_OK;
}

// forward of rows already normalised in H[0]; H[L] = raw (normalised-space) output
int forward(mst_mlp *q, int64_t rows) {
    for (int l = 0; l < q->L; l++) {
        GemmArgs a;
        memset(&a, 0, sizeof(a));
        a.A = q->H[l];
        a.B = q->p + q->woff[l];
        a.C = q->H[l + 1];
        a.M = rows;
        a.N = q->dout[l];
        a.K = q->din[l];
        a.sam = q->din[l];
        a.sak = 1;
        a.sbk = 1;
        a.sbn = q->din[l];
        a.ldc = q->dout[l];
        a.bias = q->p + q->boff[l];
        int rc = l + 1 < q->L ? gemm<EPI_BIAS_TANH>(q, MST_MLP_K_GEMM_FWD, a) : gemm<EPI_BIAS>(q, MST_MLP_K_GEMM_FWD, a);
        if (rc) return rc;
    }
    return MST_OK;
}

// the batch dimension of the weight gradients is split into chunks of >= 256 rows (at most 64)
int64_t split_chunk(int64_t rows) { return std::max<int64_t>(256, ((rows + 63) / 64 + TK - 1) / TK * TK); }
int64_t split_need(const mst_mlp *q, int64_t rows) {
    const int64_t nsplit = (rows + split_chunk(rows) - 1) / split_chunk(rows);
    return nsplit * q->width * std::max(q->width, std::max(q->n_in, q->n_out));
}
// (outside graph capture) the split buffer for batches of `rows`
int ensure_split(mst_mlp *q, int64_t rows) {
    const int64_t need = split_need(q, rows);
    if (need <= q->split_cap) return MST_OK;
    drop_graph(q);
    if (q->split) (void)hipFree(q->split);
    q->split = nullptr;
    q->split_cap = 0;
    HIPCHK(hipMalloc((void **)&q->split, sizeof(float) * need));
    q->split_cap = need;
    return MST_OK;
}

// gradient of the mean squared error over rows (targets in Tb) into q->g; per-element squares in sq
int backward(mst_mlp *q, int64_t rows) {
    const int64_t ny = rows * q->n_out;
    int rc0 = mlp_launch(q, MST_MLP_K_OTHER, 0.0, k_loss_grad, dim3(nblk(ny)), dim3(GT), ny, q->n_out,
                         (const float *)q->H[q->L], (const float *)q->Tb, (const float *)(q->scal + 2 * q->n_in),
                         (const float *)(q->scal + 2 * q->n_in + q->n_out), 1.f / (float)ny, q->D[q->L], q->sq);
    if (rc0) return rc0;
    const int64_t kchunk = split_chunk(rows);
    if (split_need(q, rows) > q->split_cap) return mst::fail(MST_E_ARG, "mlp: split buffer too small");
    for (int l = q->L - 1; l >= 0; l--) {
        // D[l+1] holds dZ_l (the output layer is linear; hidden layers got the tanh factor below)
        GemmArgs a;
        memset(&a, 0, sizeof(a));
        a.A = q->D[l + 1];  // dZ^T : A(m,k) = dZ[k*out + m]
        a.B = q->H[l];      // H[l] : B(k,n) = H[k*in + n]
        a.C = q->split;
        a.M = q->dout[l];
        a.N = q->din[l];
        a.K = rows;
        a.sam = 1;
        a.sak = q->dout[l];
        a.sbk = q->din[l];
        a.sbn = 1;
        a.ldc = q->din[l];
        a.cz = (int64_t)q->dout[l] * q->din[l];
        int nz = 1;
        int rc = gemm<EPI_NONE>(q, MST_MLP_K_GEMM_WGRAD, a, kchunk, &nz);
        if (rc) return rc;
        const int64_t nw = (int64_t)q->dout[l] * q->din[l];
        rc = mlp_launch(q, MST_MLP_K_OTHER, 0.0, k_sum_splits, dim3(nblk(nw)), dim3(GT), nw, nz,
                        (const float *)q->split, q->g + q->woff[l]);
        if (rc) return rc;
        rc = mlp_launch(q, MST_MLP_K_OTHER, 0.0, k_colsum, dim3(q->dout[l]), dim3(GT), rows, (int64_t)q->dout[l],
                        (const float *)q->D[l + 1], q->g + q->boff[l]);
        if (rc) return rc;
        if (l == 0) break;
        // dZ_{l-1} = (dZ_l W_l) * (1 - H[l]^2)
        GemmArgs b;
        memset(&b, 0, sizeof(b));
        b.A = q->D[l + 1];
        b.B = q->p + q->woff[l];  // W(k=out, n=in) = W[k*in + n]
        b.C = q->D[l];
        b.M = rows;
        b.N = q->din[l];
        b.K = q->dout[l];
        b.sam = q->dout[l];
        b.sak = 1;
        b.sbk = q->din[l];
        b.sbn = 1;
        b.ldc = q->din[l];
        b.P = q->H[l];
        rc = gemm<EPI_DTANH>(q, MST_MLP_K_GEMM_DGRAD, b);
        if (rc) return rc;
    }
    return MST_OK;
}

int train_step_enqueue(mst_mlp *q, int batch, float lr) {
    int rc = mlp_launch(q, MST_MLP_K_OTHER, 0.0, k_gather, dim3(nblk(batch)), dim3(GT), (int64_t)batch, q->n_data,
                        q->n_in, q->n_out, (const float *)q->X, (const float *)q->T, (const float *)q->scal,
                        (const float *)(q->scal + q->n_in), (const int64_t *)q->step, q->seed, q->H[0], q->Tb);
    if (!rc) rc = forward(q, batch);
    if (!rc) rc = backward(q, batch);
    if (rc) return rc;
    rc = mlp_launch(q, MST_MLP_K_OTHER, 0.0, k_adam, dim3(nblk(q->np)), dim3(GT), q->np, q->p, (const float *)q->g,
                    q->m, q->v, q->step, lr, 0.9f, 0.999f, 1e-8f);
    if (!rc) rc = mlp_launch(q, MST_MLP_K_OTHER, 0.0, k_step_advance, dim3(1), dim3(64), q->step);
    return rc;
}

double sum_host(const float *dptr, int64_t n) {
    std::vector<float> h(n);
    if (hipMemcpy(h.data(), dptr, sizeof(float) * n, hipMemcpyDeviceToHost) != hipSuccess) return NAN;
    double s = 0;
    for (float x : h) s += x;
    return s;
}

}  // namespace

namespace mst {

int mlp_forward_device(mst_mlp *q, const float *dX, int64_t rows, float *dY, hipStream_t stream) {
    // rows in chunks of at most 1 << 20 through the work buffers
    const int64_t chunk = 1 << 20;
    int rc = ensure_work(q, std::min(rows, chunk));
    if (rc) return rc;
    (void)stream;  // the network runs on its own stream; the caller's stream is synchronised around it
    HIPCHK(hipStreamSynchronize(stream));
    for (int64_t r0 = 0; r0 < rows; r0 += chunk) {
        const int64_t nr = std::min(chunk, rows - r0);
        hipLaunchKernelGGL(k_normalise, dim3(nblk(nr * q->n_in)), dim3(GT), 0, q->stream, nr * q->n_in, q->n_in,
                           dX + r0 * q->n_in, (const float *)q->scal, (const float *)(q->scal + q->n_in), q->H[0]);
        HIPCHK(hipGetLastError());
        rc = forward(q, nr);
        if (rc) return rc;
        hipLaunchKernelGGL(k_denorm, dim3(nblk(nr * q->n_out)), dim3(GT), 0, q->stream, nr * q->n_out, q->n_out,
                           (const float *)q->H[q->L], (const float *)(q->scal + 2 * q->n_in),
                           (const float *)(q->scal + 2 * q->n_in + q->n_out), dY + r0 * q->n_out);
        HIPCHK(hipGetLastError());
    }
    HIPCHK(hipStreamSynchronize(q->stream));
    return MST_OK;
}

int mlp_dims(const mst_mlp *q, int *n_in, int *n_out) {
    *n_in = q->n_in;
    *n_out = q->n_out;
    return MST_OK;
}

}  // namespace mst

extern "C" {

int mst_mlp_create(int n_in, int width, int hidden_layers, int n_out, uint64_t seed, mst_mlp **out) {
    if (!out || n_in < 1 || n_in > 64 || width < 1 || width > 4096 || hidden_layers < 1 || hidden_layers > 16 ||
        n_out < 1 || n_out > 64)
        return mst::fail(MST_E_ARG, "mlp: bad sizes");
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return mst::fail(MST_E_HIP, "no HIP device available");
    mst_mlp *q = new (std::nothrow) mst_mlp();
    if (!q) return mst::fail(MST_E_ARG, "out of host memory");
    q->n_in = n_in;
    q->width = width;
    q->hidden = hidden_layers;
    q->n_out = n_out;
    q->L = hidden_layers + 1;
    q->seed = seed;
    int64_t off = 0;
    for (int l = 0; l < q->L; l++) {
        const int di = l == 0 ? n_in : width, dd = l + 1 == q->L ? n_out : width;
        q->din.push_back(di);
        q->dout.push_back(dd);
        q->woff.push_back(off);
        off += (int64_t)di * dd;
        q->boff.push_back(off);
        off += dd;
    }
    q->np = off;
    q->H.assign(q->L + 1, nullptr);
    q->D.assign(q->L + 1, nullptr);
    // Glorot-uniform weights, zero biases, from a seeded mt19937_64
    std::vector<float> p(q->np, 0.f);
    std::mt19937_64 eng(seed);
    for (int l = 0; l < q->L; l++) {
        const double a = std::sqrt(6.0 / (q->din[l] + q->dout[l]));
        for (int64_t i = 0; i < (int64_t)q->din[l] * q->dout[l]; i++)
            p[q->woff[l] + i] = (float)(a * (2.0 * ((eng() >> 11) * (1.0 / 9007199254740992.0)) - 1.0));
    }
    std::vector<float> sc(2 * n_in + 2 * n_out, 0.f);
    for (int k = 0; k < n_in; k++) sc[n_in + k] = 1.f;
    for (int k = 0; k < n_out; k++) sc[2 * n_in + n_out + k] = 1.f;
    bool ok = hipStreamCreate(&q->stream) == hipSuccess && hipMalloc((void **)&q->p, sizeof(float) * q->np) == hipSuccess &&
              hipMalloc((void **)&q->g, sizeof(float) * q->np) == hipSuccess &&
              hipMalloc((void **)&q->m, sizeof(float) * q->np) == hipSuccess &&
              hipMalloc((void **)&q->v, sizeof(float) * q->np) == hipSuccess &&
              hipMalloc((void **)&q->scal, sizeof(float) * sc.size()) == hipSuccess &&
              hipMalloc((void **)&q->step, sizeof(int64_t)) == hipSuccess &&
              hipMemcpy(q->p, p.data(), sizeof(float) * q->np, hipMemcpyHostToDevice) == hipSuccess &&
              hipMemset(q->m, 0, sizeof(float) * q->np) == hipSuccess &&
              hipMemset(q->v, 0, sizeof(float) * q->np) == hipSuccess &&
              hipMemset(q->g, 0, sizeof(float) * q->np) == hipSuccess &&
              hipMemset(q->step, 0, sizeof(int64_t)) == hipSuccess &&
              hipMemcpy(q->scal, sc.data(), sizeof(float) * sc.size(), hipMemcpyHostToDevice) == hipSuccess;
    if (!ok) {
        mst_mlp_destroy(q);
        return mst::fail(MST_E_HIP, "mlp: device allocation failed");
    }
    *out = q;
    return MST_OK;
}

int mst_mlp_destroy(mst_mlp *q) {
    if (!q) return MST_OK;
    if (q->stream) (void)hipStreamSynchronize(q->stream);
    free_work(q);
    for (auto &v : q->ev)
        for (auto &pr : v) {
            (void)hipEventDestroy(pr.first);
            (void)hipEventDestroy(pr.second);
        }
    for (auto e : q->pool) (void)hipEventDestroy(e);
    for (void *b : {(void *)q->p, (void *)q->g, (void *)q->m, (void *)q->v, (void *)q->scal, (void *)q->step,
                    (void *)q->X, (void *)q->T, (void *)q->split})
        if (b) (void)hipFree(b);
    if (q->stream) (void)hipStreamDestroy(q->stream);
    delete q;
    return MST_OK;
}

int mst_mlp_num_params(const mst_mlp *q, int64_t *n) {
    if (!q || !n) return mst::fail(MST_E_ARG, "null argument");
    *n = q->np;
    return MST_OK;
}

int mst_mlp_get_params(mst_mlp *q, float *p, int64_t n) {
    if (!q || !p || n != q->np) return mst::fail(MST_E_ARG, "parameter count mismatch");
    HIPCHK(hipStreamSynchronize(q->stream));
    HIPCHK(hipMemcpy(p, q->p, sizeof(float) * n, hipMemcpyDeviceToHost));
    return MST_OK;
}

int mst_mlp_set_params(mst_mlp *q, const float *p, int64_t n) {
    if (!q || !p || n != q->np) return mst::fail(MST_E_ARG, "parameter count mismatch");
    HIPCHK(hipStreamSynchronize(q->stream));
    HIPCHK(hipMemcpy(q->p, p, sizeof(float) * n, hipMemcpyHostToDevice));
    return MST_OK;
}

int mst_mlp_set_scaling(mst_mlp *q, const float *in_shift, const float *in_scale, const float *out_shift,
                        const float *out_scale) {
    if (!q || !in_shift || !in_scale || !out_shift || !out_scale) return mst::fail(MST_E_ARG, "null argument");
    std::vector<float> sc;
    sc.insert(sc.end(), in_shift, in_shift + q->n_in);
    sc.insert(sc.end(), in_scale, in_scale + q->n_in);
    sc.insert(sc.end(), out_shift, out_shift + q->n_out);
    sc.insert(sc.end(), out_scale, out_scale + q->n_out);
    HIPCHK(hipStreamSynchronize(q->stream));
    HIPCHK(hipMemcpy(q->scal, sc.data(), sizeof(float) * sc.size(), hipMemcpyHostToDevice));
    return MST_OK;
}

int mst_mlp_set_data(mst_mlp *q, const float *X, const float *T, int64_t n) {
    if (!q || !X || !T || n < 1) return mst::fail(MST_E_ARG, "bad training set");
    HIPCHK(hipStreamSynchronize(q->stream));
    drop_graph(q);
    if (q->X) (void)hipFree(q->X);
    if (q->T) (void)hipFree(q->T);
    q->X = q->T = nullptr;
    HIPCHK(hipMalloc((void **)&q->X, sizeof(float) * n * q->n_in));
    HIPCHK(hipMalloc((void **)&q->T, sizeof(float) * n * q->n_out));
    HIPCHK(hipMemcpy(q->X, X, sizeof(float) * n * q->n_in, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(q->T, T, sizeof(float) * n * q->n_out, hipMemcpyHostToDevice));
    q->n_data = n;
    return MST_OK;
}

int mst_mlp_reset_optimizer(mst_mlp *q) {
    if (!q) return mst::fail(MST_E_ARG, "null argument");
    HIPCHK(hipStreamSynchronize(q->stream));
    HIPCHK(hipMemset(q->m, 0, sizeof(float) * q->np));
    HIPCHK(hipMemset(q->v, 0, sizeof(float) * q->np));
    HIPCHK(hipMemset(q->step, 0, sizeof(int64_t)));
    return MST_OK;
}

int mst_mlp_train(mst_mlp *q, int steps, int batch, double lr, int use_graph, double *loss) {
    if (!q || steps < 0 || batch < 1 || !(lr > 0)) return mst::fail(MST_E_ARG, "bad arguments");
    if (!q->X) return mst::fail(MST_E_ARG, "mlp: no training set");
    if (batch > q->n_data) batch = (int)q->n_data;
    int rc = ensure_work(q, batch);
    if (!rc) rc = ensure_split(q, batch);
    if (rc) return rc;
    if (use_graph && (q->timing || profiler_tool_loaded())) use_graph = 0;  // per-dispatch events need direct launches
    if (use_graph && (!q->gexec || q->g_batch != batch || q->g_lr != (float)lr)) {
        drop_graph(q);
        hipGraph_t g = nullptr;
        HIPCHK(hipStreamBeginCapture(q->stream, hipStreamCaptureModeThreadLocal));
        rc = train_step_enqueue(q, batch, (float)lr);
        hipError_t e = hipStreamEndCapture(q->stream, &g);
        if (rc) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        if (e != hipSuccess) return mst::fail(MST_E_HIP, std::string("graph capture: ") + hipGetErrorString(e));
        e = hipGraphInstantiate(&q->gexec, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e != hipSuccess) return mst::fail(MST_E_HIP, std::string("graph instantiate: ") + hipGetErrorString(e));
        q->g_batch = batch;
        q->g_lr = (float)lr;
    }
    for (int s = 0; s < steps; s++) {
        if (use_graph) {
            HIPCHK(hipGraphLaunch(q->gexec, q->stream));
        } else {
            rc = train_step_enqueue(q, batch, (float)lr);
            if (rc) return rc;
        }
    }
    HIPCHK(hipStreamSynchronize(q->stream));
    if (loss) *loss = steps > 0 ? sum_host(q->sq, (int64_t)batch * q->n_out) / ((double)batch * q->n_out) : NAN;
    return MST_OK;
}

int mst_mlp_forward(mst_mlp *q, const float *X, int64_t n, float *Y) {
    if (!q || !X || !Y || n < 1) return mst::fail(MST_E_ARG, "bad arguments");
    float *dX = nullptr, *dY = nullptr;
    HIPCHK(hipMalloc((void **)&dX, sizeof(float) * n * q->n_in));
    HIPCHK(hipMalloc((void **)&dY, sizeof(float) * n * q->n_out));
    int rc = MST_OK;
    if (hipMemcpy(dX, X, sizeof(float) * n * q->n_in, hipMemcpyHostToDevice) != hipSuccess)
        rc = mst::fail(MST_E_HIP, "copy failed");
    if (!rc) rc = mst::mlp_forward_device(q, dX, n, dY, q->stream);
    if (!rc && hipMemcpy(Y, dY, sizeof(float) * n * q->n_out, hipMemcpyDeviceToHost) != hipSuccess)
        rc = mst::fail(MST_E_HIP, "copy failed");
    (void)hipFree(dX);
    (void)hipFree(dY);
    return rc;
}

// full-batch gradient of the mean squared error of (X, T) (de-normalised outputs against T)
int mst_mlp_gradient(mst_mlp *q, const float *X, const float *T, int64_t n, float *grad, double *loss) {
    if (!q || !X || !T || n < 1) return mst::fail(MST_E_ARG, "bad arguments");
    int rc = ensure_work(q, n);
    if (!rc) rc = ensure_split(q, n);
    if (rc) return rc;
    float *dX = nullptr;
    HIPCHK(hipMalloc((void **)&dX, sizeof(float) * n * q->n_in));
    if (hipMemcpy(dX, X, sizeof(float) * n * q->n_in, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(q->Tb, T, sizeof(float) * n * q->n_out, hipMemcpyHostToDevice) != hipSuccess) {
        (void)hipFree(dX);
        return mst::fail(MST_E_HIP, "copy failed");
    }
    hipLaunchKernelGGL(k_normalise, dim3(nblk(n * q->n_in)), dim3(GT), 0, q->stream, n * q->n_in, q->n_in,
                       (const float *)dX, (const float *)q->scal, (const float *)(q->scal + q->n_in), q->H[0]);
    rc = forward(q, n);
    if (!rc) rc = backward(q, n);
    hipError_t e = hipStreamSynchronize(q->stream);
    (void)hipFree(dX);
    if (rc) return rc;
    if (e != hipSuccess) return mst::fail(MST_E_HIP, hipGetErrorString(e));
    if (grad) HIPCHK(hipMemcpy(grad, q->g, sizeof(float) * q->np, hipMemcpyDeviceToHost));
    if (loss) *loss = sum_host(q->sq, n * q->n_out) / ((double)n * q->n_out);
    return MST_OK;
}

int mst_mlp_timing_enable(mst_mlp *q, int on) {
    if (!q) return mst::fail(MST_E_ARG, "null argument");
    HIPCHK(hipStreamSynchronize(q->stream));
    q->timing = on != 0;
    return MST_OK;
}

int mst_mlp_timing_get(mst_mlp *q, int kclass, int64_t *count, double *total_ms, double *total_flop) {
    if (!q || kclass < 0 || kclass >= MST_MLP_K_COUNT || !count || !total_ms)
        return mst::fail(MST_E_ARG, "bad kernel class");
    HIPCHK(hipStreamSynchronize(q->stream));
    double tot = 0, fl = 0;
    for (size_t i = 0; i < q->ev[kclass].size(); i++) {
        float ms = 0;
        HIPCHK(hipEventElapsedTime(&ms, q->ev[kclass][i].first, q->ev[kclass][i].second));
        tot += ms;
        fl += q->flops[kclass][i];
    }
    *count = (int64_t)q->ev[kclass].size();
    *total_ms = tot;
    if (total_flop) *total_flop = fl;
    return MST_OK;
}

int mst_mlp_timing_reset(mst_mlp *q) {
    if (!q) return mst::fail(MST_E_ARG, "null argument");
    HIPCHK(hipStreamSynchronize(q->stream));
    for (int k = 0; k < MST_MLP_K_COUNT; k++) {
        for (auto &pr : q->ev[k]) {
            q->pool.push_back(pr.first);
            q->pool.push_back(pr.second);
        }
        q->ev[k].clear();
        q->flops[k].clear();
    }
    return MST_OK;
}

}  // extern "C"
