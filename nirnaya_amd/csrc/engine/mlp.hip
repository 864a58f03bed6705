// Learned warm start: a small MLP (in -> H -> H -> out, GELU) trained and
// evaluated on the device with bf16 MFMA GEMMs and FP32 accumulation, and the
// FP64 kernels of the proper orthogonal decomposition (POD) of solution
// snapshots.
//
// GEMM: one 256-lane workgroup per 64x64 output tile, four wave64s each own a
// 32x32 quarter and issue v_mfma_f32_32x32x16_bf16 (gfx950) over K in steps
// of 16; 64x32 bf16 tiles of both operands are staged through LDS (rows
// K-contiguous, padded by 8 elements), so either operand may be given
// K-contiguous or transposed. Epilogues fuse the bias, GELU (forward) and the
// GELU derivative (backward) and write FP32 and/or bf16 copies.
//
// Rounding points (restated by the CPU oracle): inputs, weights, hidden
// activations and back-propagated errors enter the GEMMs as bf16
// (round-to-nearest-even); every sum, bias, GELU and loss term is FP32.
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <numeric>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../../include/nirnaya/pdhg.h"
#include "ctx.hpp"
#include "devmat.hpp"
#include "trace.hpp"

namespace nir {
namespace gpu {
namespace mlp {

typedef __bf16 bf16;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int BM = 64, BN = 64, BK = 32, LDP = BK + 8;   // LDS row pitch (bf16)

// GELU (erf form) and its derivative.
__device__ __forceinline__ float gelu(float z) { return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f)); }
__device__ __forceinline__ float gelu_grad(float z) {
    const float cdf = 0.5f * (1.0f + erff(z * 0.70710678118654752f));
    const float pdf = 0.39894228040143268f * expf(-0.5f * z * z);
    return cdf + z * pdf;
}

enum { ACT_NONE = 0, ACT_GELU = 1, ACT_GELU_BWD = 2 };

struct Epi {
    float* out;         // FP32 result (ld = ldo), may be null
    bf16* out_bf;       // bf16 copy of the activated result (ld = ldo), may be null
    const float* bias;  // per column, may be null
    const float* z;     // pre-activation for ACT_GELU_BWD (ld = ldo)
    int ldo;
    int act;
    int rows_valid;     // rows >= this are written as zero (batch padding)
};

// Epilogue of one 32 x 32 accumulator block at (r0, c0): bias, activation,
// padding rows, FP32 and bf16 stores. Output layout of the 32x32 MFMA: lane
// holds column lane % 32 and rows 8 (r / 4) + 4 (lane / 32) + r % 4 for its
// 16 accumulators r.
__device__ __forceinline__ void epi_block(const f32x16& acc, int r0, int c0, int lane, const Epi& e) {
    const int col = c0 + (lane & 31);
    const float bias = e.bias ? e.bias[col] : 0.0f;
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        const int row = r0 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
        const size_t o = static_cast<size_t>(row) * e.ldo + col;
        float v = acc[r] + bias;
        float a = v;
        if (e.act == ACT_GELU) a = gelu(v);
        else if (e.act == ACT_GELU_BWD) a = v * gelu_grad(e.z[o]);
        if (row >= e.rows_valid) {
            v = 0.0f;
            a = 0.0f;
        }
        if (e.out) e.out[o] = e.act == ACT_GELU_BWD ? a : v;
        if (e.out_bf) e.out_bf[o] = static_cast<bf16>(a);
    }
}

// C[M x N] = sum_k A(m, k) B(n, k). A is [M][K] (lda) or, with TA, [K][M];
// B is [N][K] (ldb) or, with TB, [K][N]. M, N multiples of 64, K of 32.
template <bool TA, bool TB>
__global__ __launch_bounds__(256) void k_gemm(int M, int N, int K, const bf16* __restrict__ A, int lda,
                                              const bf16* __restrict__ B, int ldb, Epi e) {
    __shared__ bf16 As[BM * LDP];
    __shared__ bf16 Bs[BN * LDP];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int m0 = blockIdx.y * BM, n0 = blockIdx.x * BN;
    const int wr = w >> 1, wc = w & 1;
    f32x16 acc = {};
    for (int k0 = 0; k0 < K; k0 += BK) {
        // Stage A: 64 x 32 values, 8 per thread.
        if (!TA) {
            const int r = tid >> 2, c = (tid & 3) * 8;
            const bf16x8 v = *reinterpret_cast<const bf16x8*>(A + static_cast<size_t>(m0 + r) * lda + k0 + c);
            *reinterpret_cast<bf16x8*>(&As[r * LDP + c]) = v;
        } else {
            const int k = tid >> 3, c = (tid & 7) * 8;
            const bf16x8 v = *reinterpret_cast<const bf16x8*>(A + static_cast<size_t>(k0 + k) * lda + m0 + c);
#pragma unroll
            for (int i = 0; i < 8; ++i) As[(c + i) * LDP + k] = v[i];
        }
        if (!TB) {
            const int r = tid >> 2, c = (tid & 3) * 8;
            const bf16x8 v = *reinterpret_cast<const bf16x8*>(B + static_cast<size_t>(n0 + r) * ldb + k0 + c);
            *reinterpret_cast<bf16x8*>(&Bs[r * LDP + c]) = v;
        } else {
            const int k = tid >> 3, c = (tid & 7) * 8;
            const bf16x8 v = *reinterpret_cast<const bf16x8*>(B + static_cast<size_t>(k0 + k) * ldb + n0 + c);
#pragma unroll
            for (int i = 0; i < 8; ++i) Bs[(c + i) * LDP + k] = v[i];
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < BK; kk += 16) {
            const bf16x8 a = *reinterpret_cast<const bf16x8*>(&As[(wr * 32 + (lane & 31)) * LDP + kk + 8 * (lane >> 5)]);
            const bf16x8 b = *reinterpret_cast<const bf16x8*>(&Bs[(wc * 32 + (lane & 31)) * LDP + kk + 8 * (lane >> 5)]);
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
        }
        __syncthreads();
    }
    epi_block(acc, m0 + wr * 32, n0 + wc * 32, lane, e);
}

// The same product for large shapes: 128 x 128 tiles, each of the four waves
// a 64 x 64 quarter as 2 x 2 blocks of 32x32x16 MFMAs (four MFMAs per 16-wide
// k step instead of one) over 64-wide k slices, the next slice loaded from
// global memory into registers and stored into the second LDS buffer while
// the current one is multiplied out of the first. M, N multiples of 128, K
// of 64.
constexpr int GM = 128, GN = 128, GK = 64, GLP = GK + 8;   // k slice and LDS row pitch (bf16)

template <bool TP>
__device__ __forceinline__ void big_load(const bf16* __restrict__ P, int ld, int p0, int k0, int tid, bf16x8 (&v)[4]) {
    if (!TP) {   // P is [rows][K]: thread -> row tid / 2, 32 k values
        const int r = tid >> 1, c = (tid & 1) * 32;
        const bf16* q = P + static_cast<size_t>(p0 + r) * ld + k0 + c;
#pragma unroll
        for (int i = 0; i < 4; ++i) v[i] = *reinterpret_cast<const bf16x8*>(q + 8 * i);
    } else {     // P is [K][rows]: thread -> k tid / 4, 32 row values
        const int k = tid >> 2, c = (tid & 3) * 32;
        const bf16* q = P + static_cast<size_t>(k0 + k) * ld + p0 + c;
#pragma unroll
        for (int i = 0; i < 4; ++i) v[i] = *reinterpret_cast<const bf16x8*>(q + 8 * i);
    }
}

template <bool TP>
__device__ __forceinline__ void big_store(bf16* S, int tid, const bf16x8 (&v)[4]) {
    if (!TP) {
        const int r = tid >> 1, c = (tid & 1) * 32;
#pragma unroll
        for (int i = 0; i < 4; ++i) *reinterpret_cast<bf16x8*>(&S[r * GLP + c + 8 * i]) = v[i];
    } else {
        const int k = tid >> 2, c = (tid & 3) * 32;
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 8; ++j) S[(c + 8 * i + j) * GLP + k] = v[i][j];
    }
}

template <bool TA, bool TB>
__global__ __launch_bounds__(256) void k_gemm_big(int M, int N, int K, const bf16* __restrict__ A, int lda,
                                                  const bf16* __restrict__ B, int ldb, Epi e) {
    // Two LDS buffers: slice s + 1 is stored into one while slice s is read
    // from the other, one barrier per slice.
    __shared__ bf16 As[2][GM * GLP];
    __shared__ bf16 Bs[2][GN * GLP];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int m0 = blockIdx.y * GM, n0 = blockIdx.x * GN;
    const int wr = w >> 1, wc = w & 1;
    f32x16 acc[2][2] = {};
    bf16x8 va[4], vb[4];
    big_load<TA>(A, lda, m0, 0, tid, va);
    big_load<TB>(B, ldb, n0, 0, tid, vb);
    big_store<TA>(As[0], tid, va);
    big_store<TB>(Bs[0], tid, vb);
    __syncthreads();
    int cur = 0;
    for (int k0 = 0; k0 < K; k0 += GK) {
        const bool more = k0 + GK < K;
        if (more) {   // next slice in flight during the MFMAs
            big_load<TA>(A, lda, m0, k0 + GK, tid, va);
            big_load<TB>(B, ldb, n0, k0 + GK, tid, vb);
        }
        const bf16* as = As[cur];
        const bf16* bs = Bs[cur];
#pragma unroll
        for (int kk = 0; kk < GK; kk += 16) {
            bf16x8 a[2], b[2];
#pragma unroll
            for (int i = 0; i < 2; ++i) {
                a[i] = *reinterpret_cast<const bf16x8*>(
                    &as[(wr * 64 + i * 32 + (lane & 31)) * GLP + kk + 8 * (lane >> 5)]);
                b[i] = *reinterpret_cast<const bf16x8*>(
                    &bs[(wc * 64 + i * 32 + (lane & 31)) * GLP + kk + 8 * (lane >> 5)]);
            }
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
        }
        if (more) {
            big_store<TA>(As[cur ^ 1], tid, va);
            big_store<TB>(Bs[cur ^ 1], tid, vb);
        }
        __syncthreads();
        cur ^= 1;
    }
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) epi_block(acc[i][j], m0 + wr * 64 + i * 32, n0 + wc * 64 + j * 32, lane, e);
}

// Mean squared error over rows_valid x n: loss partials per block and the
// output error 2 (z - t) / (rows_valid n) in FP32 and bf16 (zero on padding).
__global__ void k_mse(const float* z, const float* t, int rows, int rows_valid, int n, float* d, bf16* d_bf,
                      double* part) {
    __shared__ double sh[256];
    const long long total = static_cast<long long>(rows) * n;
    const double scale = 2.0 / (static_cast<double>(rows_valid) * n);
    double s = 0.0;
    for (long long i = blockIdx.x * 256ll + threadIdx.x; i < total; i += 256ll * gridDim.x) {
        const int row = static_cast<int>(i / n);
        float g = 0.0f;
        if (row < rows_valid) {
            const float diff = z[i] - t[i];
            s += static_cast<double>(diff) * diff;
            g = static_cast<float>(scale * diff);
        }
        d[i] = g;
        d_bf[i] = static_cast<bf16>(g);
    }
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

// Column sums (bias gradients): g[c] = sum over rows of d[r][c].
__global__ void k_colsum(const float* d, int rows, int n, float* g) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    float s = 0.0f;
    for (int r = 0; r < rows; ++r) s += d[static_cast<size_t>(r) * n + c];
    g[c] = s;
}

// Adam step on FP32 master weights, with the bf16 copy the GEMMs read.
__global__ void k_adam(float* w, bf16* w_bf, const float* g, float* m, float* v, long long n, float lr, float b1,
                       float b2, float eps, float c1, float c2) {
    for (long long i = blockIdx.x * 256ll + threadIdx.x; i < n; i += 256ll * gridDim.x) {
        const float gi = g[i];
        const float mi = b1 * m[i] + (1.0f - b1) * gi;
        const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
        m[i] = mi;
        v[i] = vi;
        const float wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
        w[i] = wi;
        if (w_bf) w_bf[i] = static_cast<bf16>(wi);
    }
}

__global__ void k_to_bf16(const float* in, bf16* out, long long n) {
    for (long long i = blockIdx.x * 256ll + threadIdx.x; i < n; i += 256ll * gridDim.x) out[i] = static_cast<bf16>(in[i]);
}

// Mini-batch gather: x_bf[b] = bf16(x[idx[b]]), t[b] = target[idx[b]]; rows
// past `count` are zero.
__global__ void k_gather(const float* x, const float* target, const int* idx, int count, int rows, int n_in, int n_out,
                         bf16* x_bf, float* t) {
    const int b = blockIdx.x;
    if (b >= rows) return;
    const int src = b < count ? idx[b] : -1;
    for (int j = threadIdx.x; j < n_in; j += blockDim.x)
        x_bf[static_cast<size_t>(b) * n_in + j] = static_cast<bf16>(src >= 0 ? x[static_cast<size_t>(src) * n_in + j] : 0.0f);
    for (int j = threadIdx.x; j < n_out; j += blockDim.x)
        t[static_cast<size_t>(b) * n_out + j] = src >= 0 ? target[static_cast<size_t>(src) * n_out + j] : 0.0f;
}

// ------------------------------------------------------------------ POD (FP64)

// Mirror the upper triangle of a column-major N x N matrix into its lower one.
__global__ void k_mirror_upper(double* G, int N) {
    const long long t = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    const long long stride = static_cast<long long>(gridDim.x) * blockDim.x;
    const long long total = static_cast<long long>(N) * N;
    for (long long e = t; e < total; e += stride) {
        const int i = static_cast<int>(e % N), j = static_cast<int>(e / N);
        if (i > j) G[e] = G[static_cast<size_t>(j) + static_cast<size_t>(i) * N];
    }
}

// rows[r][j] = v[j] (r < count), or rows[r][j] -= v[j] with SUB.
template <bool SUB>
__global__ void k_rows_mean(double* rows, const double* __restrict__ v, int count, int len) {
    const long long t = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    const long long stride = static_cast<long long>(gridDim.x) * blockDim.x;
    const long long total = static_cast<long long>(count) * len;
    for (long long e = t; e < total; e += stride) {
        const double m = v[e % len];
        if (SUB) rows[e] -= m;
        else rows[e] = m;
    }
}

}  // namespace mlp
}  // namespace gpu
}  // namespace nir

using namespace nir::gpu;

#define ROCBLAS_CHECK(expr)                                                                      \
    do {                                                                                         \
        rocblas_status st_ = (expr);                                                             \
        if (st_ != rocblas_status_success)                                                       \
            throw std::runtime_error(std::string(#expr) + ": " + rocblas_status_to_string(st_)); \
    } while (0)
using namespace nir::gpu::mlp;

struct nir_mlp {
    nir_ctx* ctx = nullptr;
    hipStream_t s = nullptr;
    int n_in = 0, hid = 0, n_out = 0, max_rows = 0;
    // parameters (FP32 master, bf16 copies), Adam moments, gradients
    DevArray<float> w, g, am, av;
    DevArray<bf16> w_bf;
    size_t off[6] = {0, 0, 0, 0, 0, 0};   // W1 b1 W2 b2 W3 b3
    size_t count = 0;
    long long adam_t = 0;
    // activations for max_rows rows
    DevArray<bf16> x_bf, h1_bf, h2_bf, d3_bf, d2_bf, d1_bf;
    DevArray<float> z1, z2, z3, t, d3, d2, d1;
    DevArray<double> part;
    // data set
    DevArray<float> data_x, data_t;
    DevArray<int> idx;
    int data_rows = 0;
    double flops_per_row = 0.0;   // forward + backward GEMM flops per training row
    double train_ms = 0.0;
    long long train_rows = 0;

    float* W(int k) { return w.p + off[k]; }
    bf16* Wb(int k) { return w_bf.p + off[k]; }
    float* G(int k) { return g.p + off[k]; }
};

namespace {

int pad64(int r) { return (r + 63) / 64 * 64; }

// 128 x 128 tiles when the shape has them and they fill the chip (256 CUs);
// 64 x 64 tiles otherwise (the network's training shapes: more workgroups
// for a latency-bound product).
template <bool TA, bool TB>
void gemm(hipStream_t s, int M, int N, int K, const bf16* A, int lda, const bf16* B, int ldb, const Epi& e) {
    if (M % 64 || N % 64 || K % 32) throw std::invalid_argument("gemm shape not a multiple of the tile");
    if (M % GM == 0 && N % GN == 0 && K % GK == 0 && (M / GM) * (N / GN) >= 256) {
        dim3 grid(N / GN, M / GM);
        k_gemm_big<TA, TB><<<grid, 256, 0, s>>>(M, N, K, A, lda, B, ldb, e);
        return;
    }
    dim3 grid(N / 64, M / 64);
    k_gemm<TA, TB><<<grid, 256, 0, s>>>(M, N, K, A, lda, B, ldb, e);
}

Epi epi(float* out, bf16* out_bf, const float* bias, const float* z, int ldo, int act, int rows_valid) {
    return Epi{out, out_bf, bias, z, ldo, act, rows_valid};
}

// Forward of `rows` (multiple of 64) rows already in x_bf; `valid` real rows.
void forward(nir_mlp* P, int rows, int valid) {
    gemm<false, false>(P->s, rows, P->hid, P->n_in, P->x_bf.p, P->n_in, P->Wb(0), P->n_in,
                       epi(P->z1.p, P->h1_bf.p, P->W(1), nullptr, P->hid, ACT_GELU, valid));
    gemm<false, false>(P->s, rows, P->hid, P->hid, P->h1_bf.p, P->hid, P->Wb(2), P->hid,
                       epi(P->z2.p, P->h2_bf.p, P->W(3), nullptr, P->hid, ACT_GELU, valid));
    gemm<false, false>(P->s, rows, P->n_out, P->hid, P->h2_bf.p, P->hid, P->Wb(4), P->hid,
                       epi(P->z3.p, nullptr, P->W(5), nullptr, P->n_out, ACT_NONE, valid));
}

// Loss and gradients of the mean squared error over the rows in x_bf / t.
double backward(nir_mlp* P, int rows, int valid) {
    const int nb = 64;
    k_mse<<<nb, 256, 0, P->s>>>(P->z3.p, P->t.p, rows, valid, P->n_out, P->d3.p, P->d3_bf.p, P->part.p);
    // layer 3
    gemm<true, true>(P->s, P->n_out, P->hid, rows, P->d3_bf.p, P->n_out, P->h2_bf.p, P->hid,
                     epi(P->G(4), nullptr, nullptr, nullptr, P->hid, ACT_NONE, P->n_out));
    k_colsum<<<(P->n_out + 255) / 256, 256, 0, P->s>>>(P->d3.p, rows, P->n_out, P->G(5));
    gemm<false, true>(P->s, rows, P->hid, P->n_out, P->d3_bf.p, P->n_out, P->Wb(4), P->hid,
                      epi(P->d2.p, P->d2_bf.p, nullptr, P->z2.p, P->hid, ACT_GELU_BWD, valid));
    // layer 2
    gemm<true, true>(P->s, P->hid, P->hid, rows, P->d2_bf.p, P->hid, P->h1_bf.p, P->hid,
                     epi(P->G(2), nullptr, nullptr, nullptr, P->hid, ACT_NONE, P->hid));
    k_colsum<<<(P->hid + 255) / 256, 256, 0, P->s>>>(P->d2.p, rows, P->hid, P->G(3));
    gemm<false, true>(P->s, rows, P->hid, P->hid, P->d2_bf.p, P->hid, P->Wb(2), P->hid,
                      epi(P->d1.p, P->d1_bf.p, nullptr, P->z1.p, P->hid, ACT_GELU_BWD, valid));
    // layer 1
    gemm<true, true>(P->s, P->hid, P->n_in, rows, P->d1_bf.p, P->hid, P->x_bf.p, P->n_in,
                     epi(P->G(0), nullptr, nullptr, nullptr, P->n_in, ACT_NONE, P->hid));
    k_colsum<<<(P->hid + 255) / 256, 256, 0, P->s>>>(P->d1.p, rows, P->hid, P->G(1));
    HIP_CHECK(hipGetLastError());
    std::vector<double> h(nb);
    HIP_CHECK(hipMemcpyAsync(h.data(), P->part.p, nb * sizeof(double), hipMemcpyDeviceToHost, P->s));
    HIP_CHECK(hipStreamSynchronize(P->s));
    double sse = 0.0;
    for (double v : h) sse += v;
    return sse / (static_cast<double>(valid) * P->n_out);
}

void upload_rows_bf16(nir_mlp* P, const float* x, int valid, int rows) {
    std::vector<float> buf(static_cast<size_t>(rows) * P->n_in, 0.0f);
    std::copy(x, x + static_cast<size_t>(valid) * P->n_in, buf.begin());
    DevArray<float> tmp;
    tmp.alloc(buf.size());
    tmp.upload(buf.data(), buf.size(), P->s);
    k_to_bf16<<<256, 256, 0, P->s>>>(tmp.p, P->x_bf.p, static_cast<long long>(buf.size()));
    HIP_CHECK(hipStreamSynchronize(P->s));
}

void check_rows(nir_mlp* P, int rows) {
    if (rows <= 0 || rows > P->max_rows) throw std::invalid_argument("row count outside 1..max_rows");
}

}  // namespace

// A POD basis resident on the device: modes [len][R] row-major (the
// column-major R x len matrix M), the mean, and the rows of the last decode
// ([count][len] row-major), which the family engine can take as warm starts
// without a trip through the host.
struct nir_pod {
    nir_ctx* ctx = nullptr;
    int len = 0, R = 0;
    DevArray<double> modes, mean, rows;
    int rows_count = 0;
    rocblas_handle h = nullptr;
    ~nir_pod() {
        if (h) (void)rocblas_destroy_handle(h);
    }
};

namespace nir {
namespace gpu {
const double* pod_rows(const nir_pod* P, int* count, int* len, int* device) {
    if (!P) return nullptr;
    *count = P->rows_count;
    *len = P->len;
    *device = P->ctx->device;
    return P->rows_count > 0 ? P->rows.p : nullptr;
}
}  // namespace gpu
}  // namespace nir

extern "C" {

NIR_API int nir_mlp_create(nir_ctx* ctx, int32_t n_in, int32_t hidden, int32_t n_out, int32_t max_rows,
                           uint64_t seed, nir_mlp** out) {
    if (!ctx || !out) return -1;
    *out = nullptr;
    try {
        if (n_in <= 0 || hidden <= 0 || n_out <= 0 || n_in % 64 || hidden % 64 || n_out % 64)
            throw std::invalid_argument("layer widths must be positive multiples of 64");
        if (max_rows <= 0) throw std::invalid_argument("max_rows must be positive");
        HIP_CHECK(hipSetDevice(ctx->device));
        auto P = std::make_unique<nir_mlp>();
        P->ctx = ctx;
        P->s = ctx->stream;
        P->n_in = n_in;
        P->hid = hidden;
        P->n_out = n_out;
        P->max_rows = pad64(max_rows);
        const size_t sizes[6] = {static_cast<size_t>(hidden) * n_in, static_cast<size_t>(hidden),
                                 static_cast<size_t>(hidden) * hidden, static_cast<size_t>(hidden),
                                 static_cast<size_t>(n_out) * hidden, static_cast<size_t>(n_out)};
        size_t o = 0;
        for (int k = 0; k < 6; ++k) {
            P->off[k] = o;
            o += sizes[k];
        }
        P->count = o;
        // Uniform +-sqrt(3 / fan_in) weights (unit-variance pre-activations
        // for unit-variance inputs), zero biases; seeded.
        std::vector<float> h(o, 0.0f);
        std::mt19937_64 rng(seed);
        const int fan[3] = {n_in, hidden, hidden};
        for (int l = 0; l < 3; ++l) {
            const float a = std::sqrt(3.0f / static_cast<float>(fan[l]));
            std::uniform_real_distribution<float> U(-a, a);
            for (size_t i = 0; i < sizes[2 * l]; ++i) h[P->off[2 * l] + i] = U(rng);
        }
        P->w.alloc(o);
        P->w.upload(h.data(), o, P->s);
        P->w_bf.alloc(o);
        P->g.alloc(o);
        P->am.alloc(o);
        P->av.alloc(o);
        std::vector<float> zeros(o, 0.0f);
        P->am.upload(zeros.data(), o, P->s);
        P->av.upload(zeros.data(), o, P->s);
        k_to_bf16<<<256, 256, 0, P->s>>>(P->w.p, P->w_bf.p, static_cast<long long>(o));
        const size_t R = P->max_rows;
        P->x_bf.alloc(R * n_in);
        P->h1_bf.alloc(R * hidden);
        P->h2_bf.alloc(R * hidden);
        P->d3_bf.alloc(R * n_out);
        P->d2_bf.alloc(R * hidden);
        P->d1_bf.alloc(R * hidden);
        P->z1.alloc(R * hidden);
        P->z2.alloc(R * hidden);
        P->z3.alloc(R * n_out);
        P->t.alloc(R * n_out);
        P->d3.alloc(R * n_out);
        P->d2.alloc(R * hidden);
        P->d1.alloc(R * hidden);
        P->part.alloc(64);
        P->flops_per_row = 2.0 * 3.0 * (static_cast<double>(n_in) * hidden + static_cast<double>(hidden) * hidden +
                                        static_cast<double>(hidden) * n_out) -
                           2.0 * static_cast<double>(n_in) * hidden;   // no input gradient
        HIP_CHECK(hipStreamSynchronize(P->s));
        *out = P.release();
        return 0;
    } catch (const std::exception& ex) {
        ctx->error = ex.what();
        return -1;
    }
}

NIR_API int64_t nir_mlp_num_params(const nir_mlp* P) { return P ? static_cast<int64_t>(P->count) : -1; }

NIR_API int nir_mlp_get_params(nir_mlp* P, float* out) {
    if (!P || !out) return -1;
    try {
        P->w.download(out, P->count, P->s);
        HIP_CHECK(hipStreamSynchronize(P->s));
        return 0;
    } catch (const std::exception& ex) {
        P->ctx->error = ex.what();
        return -1;
    }
}

NIR_API int nir_mlp_set_params(nir_mlp* P, const float* in) {
    if (!P || !in) return -1;
    try {
        P->w.upload(in, P->count, P->s);
        k_to_bf16<<<256, 256, 0, P->s>>>(P->w.p, P->w_bf.p, static_cast<long long>(P->count));
        HIP_CHECK(hipStreamSynchronize(P->s));
        return 0;
    } catch (const std::exception& ex) {
        P->ctx->error = ex.what();
        return -1;
    }
}

NIR_API int nir_mlp_forward(nir_mlp* P, const float* x, int32_t rows, float* out) {
    if (!P || !x || !out) return -1;
    try {
        HIP_CHECK(hipSetDevice(P->ctx->device));
        for (int r0 = 0; r0 < rows; r0 += P->max_rows) {
            const int valid = std::min(P->max_rows, rows - r0);
            const int R = pad64(valid);
            upload_rows_bf16(P, x + static_cast<size_t>(r0) * P->n_in, valid, R);
            forward(P, R, valid);
            HIP_CHECK(hipGetLastError());
            P->z3.download(out + static_cast<size_t>(r0) * P->n_out, static_cast<size_t>(valid) * P->n_out, P->s);
            HIP_CHECK(hipStreamSynchronize(P->s));
        }
        return 0;
    } catch (const std::exception& ex) {
        P->ctx->error = ex.what();
        return -1;
    }
}

NIR_API int nir_mlp_gradient(nir_mlp* P, const float* x, const float* target, int32_t rows, float* grad,
                             double* loss) {
    if (!P || !x || !target) return -1;
    try {
        HIP_CHECK(hipSetDevice(P->ctx->device));
        check_rows(P, rows);
        const int R = pad64(rows);
        upload_rows_bf16(P, x, rows, R);
        std::vector<float> t(static_cast<size_t>(R) * P->n_out, 0.0f);
        std::copy(target, target + static_cast<size_t>(rows) * P->n_out, t.begin());
        P->t.upload(t.data(), t.size(), P->s);
        forward(P, R, rows);
        const double l = backward(P, R, rows);
        if (loss) *loss = l;
        if (grad) {
            P->g.download(grad, P->count, P->s);
            HIP_CHECK(hipStreamSynchronize(P->s));
        }
        return 0;
    } catch (const std::exception& ex) {
        P->ctx->error = ex.what();
        return -1;
    }
}

NIR_API int nir_mlp_set_data(nir_mlp* P, const float* x, const float* target, int32_t rows) {
    if (!P || !x || !target || rows <= 0) return -1;
    try {
        HIP_CHECK(hipSetDevice(P->ctx->device));
        P->data_x.alloc(static_cast<size_t>(rows) * P->n_in);
        P->data_t.alloc(static_cast<size_t>(rows) * P->n_out);
        P->data_x.upload(x, static_cast<size_t>(rows) * P->n_in, P->s);
        P->data_t.upload(target, static_cast<size_t>(rows) * P->n_out, P->s);
        P->idx.alloc(static_cast<size_t>(rows));
        P->data_rows = rows;
        HIP_CHECK(hipStreamSynchronize(P->s));
        return 0;
    } catch (const std::exception& ex) {
        P->ctx->error = ex.what();
        return -1;
    }
}

// Adam (beta 0.9 / 0.999, eps 1e-8) over shuffled mini-batches; epoch_loss
// (optional, `epochs` entries) receives the mean training loss per epoch.
NIR_API int nir_mlp_train(nir_mlp* P, int32_t epochs, int32_t batch, float lr, uint64_t seed, float* epoch_loss) {
    const nir::TraceRange trace("nir_mlp_train");
    if (!P || epochs < 0) return -1;
    try {
        HIP_CHECK(hipSetDevice(P->ctx->device));
        if (P->data_rows <= 0) throw std::invalid_argument("no training data");
        if (batch <= 0 || pad64(batch) > P->max_rows) throw std::invalid_argument("batch outside 1..max_rows");
        std::vector<int> order(P->data_rows);
        std::iota(order.begin(), order.end(), 0);
        std::mt19937_64 rng(seed);
        hipEvent_t e0, e1;
        HIP_CHECK(hipEventCreate(&e0));
        HIP_CHECK(hipEventCreate(&e1));
        for (int ep = 0; ep < epochs; ++ep) {
            std::shuffle(order.begin(), order.end(), rng);
            P->idx.upload(order.data(), order.size(), P->s);
            double lsum = 0.0;
            int nbat = 0;
            HIP_CHECK(hipEventRecord(e0, P->s));
            for (int b0 = 0; b0 < P->data_rows; b0 += batch) {
                const int valid = std::min(batch, P->data_rows - b0);
                const int R = pad64(valid);
                k_gather<<<R, 64, 0, P->s>>>(P->data_x.p, P->data_t.p, P->idx.p + b0, valid, R, P->n_in, P->n_out,
                                             P->x_bf.p, P->t.p);
                forward(P, R, valid);
                lsum += backward(P, R, valid);
                ++nbat;
                ++P->adam_t;
                const float c1 = 1.0f - std::pow(0.9f, static_cast<float>(P->adam_t));
                const float c2 = 1.0f - std::pow(0.999f, static_cast<float>(P->adam_t));
                k_adam<<<512, 256, 0, P->s>>>(P->w.p, P->w_bf.p, P->g.p, P->am.p, P->av.p,
                                              static_cast<long long>(P->count), lr, 0.9f, 0.999f, 1e-8f, c1, c2);
                P->train_rows += R;
            }
            HIP_CHECK(hipEventRecord(e1, P->s));
            HIP_CHECK(hipEventSynchronize(e1));
            float ms = 0.f;
            HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
            P->train_ms += ms;
            if (epoch_loss) epoch_loss[ep] = static_cast<float>(lsum / std::max(1, nbat));
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        return 0;
    } catch (const std::exception& ex) {
        P->ctx->error = ex.what();
        return -1;
    }
}

// Training totals: rows processed, milliseconds (events around each epoch,
// including the host synchronisation for the loss), GEMM flops.
NIR_API int nir_mlp_train_stats(nir_mlp* P, int64_t* rows, double* ms, double* gemm_flops) {
    if (!P) return -1;
    if (rows) *rows = P->train_rows;
    if (ms) *ms = P->train_ms;
    if (gemm_flops) *gemm_flops = P->flops_per_row * static_cast<double>(P->train_rows);
    return 0;
}

NIR_API void nir_mlp_destroy(nir_mlp* P) {
    if (!P) return;
    (void)hipSetDevice(P->ctx->device);
    delete P;
}

// One bf16 GEMM C = A B^T (A [M][K], B [N][K], row-major FP32 inputs rounded
// to bf16 on the device), for layout and accuracy checks of the kernel.
NIR_API int nir_mlp_gemm_check(nir_ctx* ctx, int32_t M, int32_t N, int32_t K, const float* A, const float* B,
                               int32_t transpose_a, int32_t transpose_b, float* C) {
    if (!ctx || !A || !B || !C) return -1;
    try {
        HIP_CHECK(hipSetDevice(ctx->device));
        hipStream_t s = ctx->stream;
        DevArray<float> fa, fb, fc;
        DevArray<bf16> ba, bb;
        const size_t na = static_cast<size_t>(M) * K, nb = static_cast<size_t>(N) * K;
        fa.alloc(na);
        fb.alloc(nb);
        fc.alloc(static_cast<size_t>(M) * N);
        ba.alloc(na);
        bb.alloc(nb);
        fa.upload(A, na, s);
        fb.upload(B, nb, s);
        k_to_bf16<<<256, 256, 0, s>>>(fa.p, ba.p, static_cast<long long>(na));
        k_to_bf16<<<256, 256, 0, s>>>(fb.p, bb.p, static_cast<long long>(nb));
        const Epi e = epi(fc.p, nullptr, nullptr, nullptr, N, ACT_NONE, M);
        // A: [M][K] (lda K) or [K][M] (lda M); B: [N][K] (ldb K) or [K][N] (ldb N).
        if (!transpose_a && !transpose_b) gemm<false, false>(s, M, N, K, ba.p, K, bb.p, K, e);
        else if (!transpose_a && transpose_b) gemm<false, true>(s, M, N, K, ba.p, K, bb.p, N, e);
        else if (transpose_a && !transpose_b) gemm<true, false>(s, M, N, K, ba.p, M, bb.p, K, e);
        else gemm<true, true>(s, M, N, K, ba.p, M, bb.p, N, e);
        HIP_CHECK(hipGetLastError());
        fc.download(C, static_cast<size_t>(M) * N, s);
        HIP_CHECK(hipStreamSynchronize(s));
        return 0;
    } catch (const std::exception& ex) {
        ctx->error = ex.what();
        return -1;
    }
}

// POD helpers. Snapshots X are [len][N] row-major (column a = snapshot a,
// already mean-centred by the caller). nir_pod_gram: G = X^T X (N x N).
// nir_pod_project: Y = X V for V [N][R] (Y [len][R]).
NIR_API int nir_pod_gram(nir_ctx* ctx, const double* X, int32_t len, int32_t N, double* G) {
    const nir::TraceRange trace("nir_pod_gram");
    if (!ctx || !X || !G || len < 0 || N <= 0) return -1;
    rocblas_handle h = nullptr;
    try {
        HIP_CHECK(hipSetDevice(ctx->device));
        hipStream_t s = ctx->stream;
        DevArray<double> dx, dg;
        dx.alloc(static_cast<size_t>(len) * N);
        dg.alloc(static_cast<size_t>(N) * N);
        dx.upload(X, static_cast<size_t>(len) * N, s);
        // Row-major X [len][N] is the column-major N x len matrix A, so
        // G = X^T X = A A^T: one rocBLAS DSYRK (upper triangle), mirrored.
        ROCBLAS_CHECK(rocblas_create_handle(&h));
        ROCBLAS_CHECK(rocblas_set_stream(h, s));
        const double one = 1.0, zero = 0.0;
        ROCBLAS_CHECK(rocblas_dsyrk(h, rocblas_fill_upper, rocblas_operation_none, N, len, &one, dx.p, N, &zero,
                                    dg.p, N));
        k_mirror_upper<<<2048, 256, 0, s>>>(dg.p, N);
        HIP_CHECK(hipGetLastError());
        dg.download(G, static_cast<size_t>(N) * N, s);
        HIP_CHECK(hipStreamSynchronize(s));
        (void)rocblas_destroy_handle(h);
        return 0;
    } catch (const std::exception& ex) {
        if (h) {
            (void)hipStreamSynchronize(ctx->stream);
            (void)rocblas_destroy_handle(h);
        }
        ctx->error = ex.what();
        return -1;
    }
}

NIR_API int nir_pod_project(nir_ctx* ctx, const double* X, int32_t len, int32_t N, const double* V, int32_t R,
                            double* Y) {
    if (!ctx || !X || !V || !Y || len < 0 || N <= 0 || R <= 0) return -1;
    rocblas_handle h = nullptr;
    try {
        HIP_CHECK(hipSetDevice(ctx->device));
        hipStream_t s = ctx->stream;
        DevArray<double> dx, dv, dy;
        dx.alloc(static_cast<size_t>(len) * N);
        dv.alloc(static_cast<size_t>(N) * R);
        dy.alloc(static_cast<size_t>(len) * R);
        dx.upload(X, static_cast<size_t>(len) * N, s);
        dv.upload(V, static_cast<size_t>(N) * R, s);
        // Column-major: Y^T (R x len) = V^T (R x N) X^T (N x len), one DGEMM.
        ROCBLAS_CHECK(rocblas_create_handle(&h));
        ROCBLAS_CHECK(rocblas_set_stream(h, s));
        const double one = 1.0, zero = 0.0;
        if (len > 0)
            ROCBLAS_CHECK(rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_none, R, len, N, &one, dv.p, R,
                                        dx.p, N, &zero, dy.p, R));
        dy.download(Y, static_cast<size_t>(len) * R, s);
        HIP_CHECK(hipStreamSynchronize(s));
        (void)rocblas_destroy_handle(h);
        return 0;
    } catch (const std::exception& ex) {
        if (h) {
            (void)hipStreamSynchronize(ctx->stream);
            (void)rocblas_destroy_handle(h);
        }
        ctx->error = ex.what();
        return -1;
    }
}

NIR_API int nir_pod_create(nir_ctx* ctx, const double* X, int32_t len, int32_t N, const double* V, int32_t R,
                           const double* mean, nir_pod** out) {
    const nir::TraceRange trace("nir_pod_create");
    if (!ctx || !X || !V || !mean || !out || len <= 0 || N <= 0 || R <= 0) return -1;
    *out = nullptr;
    auto P = std::make_unique<nir_pod>();
    P->ctx = ctx;
    P->len = len;
    P->R = R;
    try {
        HIP_CHECK(hipSetDevice(ctx->device));
        hipStream_t s = ctx->stream;
        ROCBLAS_CHECK(rocblas_create_handle(&P->h));
        ROCBLAS_CHECK(rocblas_set_stream(P->h, s));
        DevArray<double> dx, dv;
        dx.alloc(static_cast<size_t>(len) * N);
        dv.alloc(static_cast<size_t>(N) * R);
        P->modes.alloc(static_cast<size_t>(len) * R);
        P->mean.alloc(static_cast<size_t>(len));
        dx.upload(X, static_cast<size_t>(len) * N, s);
        dv.upload(V, static_cast<size_t>(N) * R, s);
        P->mean.upload(mean, static_cast<size_t>(len), s);
        // The same DGEMM as nir_pod_project: M (R x len) = V^T X^T.
        const double one = 1.0, zero = 0.0;
        ROCBLAS_CHECK(rocblas_dgemm(P->h, rocblas_operation_none, rocblas_operation_none, R, len, N, &one, dv.p, R,
                                    dx.p, N, &zero, P->modes.p, R));
        HIP_CHECK(hipStreamSynchronize(s));
        *out = P.release();
        return 0;
    } catch (const std::exception& ex) {
        (void)hipStreamSynchronize(ctx->stream);
        ctx->error = ex.what();
        return -1;
    }
}

NIR_API int nir_pod_modes(nir_pod* P, double* modes) {
    if (!P || !modes) return -1;
    try {
        HIP_CHECK(hipSetDevice(P->ctx->device));
        P->modes.download(modes, static_cast<size_t>(P->len) * P->R, P->ctx->stream);
        HIP_CHECK(hipStreamSynchronize(P->ctx->stream));
        return 0;
    } catch (const std::exception& ex) {
        P->ctx->error = ex.what();
        return -1;
    }
}

// rows [count][len] = mean + alpha [count][R] M: column-major, rows^T
// (len x count) = M^T (len x R) alpha^T (R x count), one DGEMM onto the
// broadcast mean.
NIR_API int nir_pod_decode(nir_pod* P, const double* alpha, int32_t count, double* out) {
    const nir::TraceRange trace("nir_pod_decode");
    if (!P || !alpha || count <= 0) return -1;
    try {
        HIP_CHECK(hipSetDevice(P->ctx->device));
        hipStream_t s = P->ctx->stream;
        const size_t total = static_cast<size_t>(count) * P->len;
        if (P->rows.n < total) P->rows.alloc(total);
        P->rows_count = 0;
        DevArray<double> da;
        da.alloc(static_cast<size_t>(count) * P->R);
        da.upload(alpha, static_cast<size_t>(count) * P->R, s);
        k_rows_mean<false><<<1024, 256, 0, s>>>(P->rows.p, P->mean.p, count, P->len);
        HIP_CHECK(hipGetLastError());
        const double one = 1.0;
        ROCBLAS_CHECK(rocblas_dgemm(P->h, rocblas_operation_transpose, rocblas_operation_none, P->len, count, P->R,
                                    &one, P->modes.p, P->R, da.p, P->R, &one, P->rows.p, P->len));
        if (out) P->rows.download(out, total, s);
        HIP_CHECK(hipStreamSynchronize(s));
        P->rows_count = count;
        return 0;
    } catch (const std::exception& ex) {
        (void)hipStreamSynchronize(P->ctx->stream);
        P->ctx->error = ex.what();
        return -1;
    }
}

// coef [count][R] = (s - mean) M^T: column-major, coef^T (R x count) =
// M (R x len) (s - mean)^T (len x count).
NIR_API int nir_pod_encode(nir_pod* P, const double* snap, int32_t count, double* coef) {
    if (!P || !snap || !coef || count <= 0) return -1;
    try {
        HIP_CHECK(hipSetDevice(P->ctx->device));
        hipStream_t s = P->ctx->stream;
        DevArray<double> ds, dc;
        ds.alloc(static_cast<size_t>(count) * P->len);
        dc.alloc(static_cast<size_t>(count) * P->R);
        ds.upload(snap, static_cast<size_t>(count) * P->len, s);
        k_rows_mean<true><<<1024, 256, 0, s>>>(ds.p, P->mean.p, count, P->len);
        HIP_CHECK(hipGetLastError());
        const double one = 1.0, zero = 0.0;
        ROCBLAS_CHECK(rocblas_dgemm(P->h, rocblas_operation_none, rocblas_operation_none, P->R, count, P->len, &one,
                                    P->modes.p, P->R, ds.p, P->len, &zero, dc.p, P->R));
        dc.download(coef, static_cast<size_t>(count) * P->R, s);
        HIP_CHECK(hipStreamSynchronize(s));
        return 0;
    } catch (const std::exception& ex) {
        (void)hipStreamSynchronize(P->ctx->stream);
        P->ctx->error = ex.what();
        return -1;
    }
}

NIR_API void nir_pod_destroy(nir_pod* P) {
    if (!P) return;
    (void)hipSetDevice(P->ctx->device);
    (void)hipStreamSynchronize(P->ctx->stream);
    delete P;
}

}  // extern "C"
