// Kernels of the surrogate network: a BF16 GEMM on the matrix cores (FP32 accumulation) with a
// fused bias / GELU epilogue and split-K, transposes, GELU backward, bias gradients, the MSE
// gradient and the Adam update.
//
// GEMM layout ("NT"): C[m][n] = sum_k A[m][k] B[n][k], both operands K-contiguous bf16 rows.  A
// block of 256 threads (4 waves, 2 x 2) owns a 128 x 128 tile of C; each wave computes 64 x 64
// with 2 x 2 v_mfma_f32_32x32x16_bf16.  Operand tiles of 128 rows x 32 k go through LDS, double
// buffered, one barrier per k tile.  MFMA lane layout (gfx950, 32x32x16 bf16): lane l supplies
// row l % 32 of A and column l % 32 of B with 8 k values (the same 8 for both; here k = (l / 32) * 8
// + j of the 16-wide k step), and element i of its 16 results is C[(i / 4) * 8 + (l / 32) * 4 + i % 4]
// [l % 32].
#include <hip/hip_runtime.h>

#include <cstdint>

#include "mlp_kernels.h"

namespace poe2d {
namespace mlp {

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef uint16_t u16x8 __attribute__((ext_vector_type(8)));

constexpr int BM = 128, BN = 128, BK = 32;
constexpr int LDS_ROW = BK + 8;  // bf16 per LDS row (16-byte pad against bank conflicts)

__device__ __forceinline__ float gelu(float x) { return 0.5f * x * (1.0f + erff(x * 0.70710678118654752f)); }

__device__ __forceinline__ float gelu_grad(float x)
{
    const float cdf = 0.5f * (1.0f + erff(x * 0.70710678118654752f));
    const float pdf = 0.39894228040143268f * expf(-0.5f * x * x);
    return cdf + x * pdf;
}

// C (+)= A B^T over k in [k0, k1) for the 128 x 128 tile (tm, tn); k0, k1 multiples of BK.
template <int EPI>
__global__ void __launch_bounds__(256) gemm_nt(const uint16_t *__restrict__ A, int lda, const uint16_t *__restrict__ B,
                                               int ldb, int M, int N, int K, int ksplit, GemmOut out)
{
    __shared__ uint16_t la[2][BM * LDS_ROW];
    __shared__ uint16_t lb[2][BN * LDS_ROW];
    const int tiles_n = (N + BN - 1) / BN;
    const int tile = blockIdx.x;
    const int tm = tile / tiles_n, tn = tile % tiles_n;
    const int split = blockIdx.y;
    const int kper = ((K / BK + ksplit - 1) / ksplit) * BK;
    const int k0 = split * kper, k1 = min(K, k0 + kper);
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int wm = wave >> 1, wn = wave & 1;

    // each thread loads 2 x 16 bytes of each operand tile: rows r0 and r0 + 64, k chunk kc * 8
    const int r0 = t >> 2, kc = t & 3;
    auto load_tile = [&](const uint16_t *P, int ld, int rows, int row0, int k, u16x8 *v) {
        for (int h = 0; h < 2; ++h) {
            const int r = row0 + r0 + h * 64;
            v[h] = r < rows ? *(const u16x8 *)(P + (size_t)r * ld + k + kc * 8) : u16x8{0, 0, 0, 0, 0, 0, 0, 0};
        }
    };
    auto store_tile = [&](uint16_t *L, const u16x8 *v) {
        for (int h = 0; h < 2; ++h) *(u16x8 *)(L + (r0 + h * 64) * LDS_ROW + kc * 8) = v[h];
    };

    f32x16 acc[2][2];
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j)
            for (int e = 0; e < 16; ++e) acc[i][j][e] = 0.f;

    u16x8 va[2], vb[2];
    int buf = 0;
    if (k0 < k1) {
        load_tile(A, lda, M, tm * BM, k0, va);
        load_tile(B, ldb, N, tn * BN, k0, vb);
        store_tile(la[0], va);
        store_tile(lb[0], vb);
    }
    __syncthreads();
    for (int k = k0; k < k1; k += BK) {
        const bool more = k + BK < k1;
        if (more) {
            load_tile(A, lda, M, tm * BM, k + BK, va);
            load_tile(B, ldb, N, tn * BN, k + BK, vb);
        }
        const uint16_t *LA = la[buf], *LB = lb[buf];
#pragma unroll
        for (int kk = 0; kk < BK; kk += 16) {
            bf16x8 fa[2], fb[2];
#pragma unroll
            for (int i = 0; i < 2; ++i) {
                const int ra = wm * 64 + i * 32 + (lane & 31);
                const int rb = wn * 64 + i * 32 + (lane & 31);
                fa[i] = *(const bf16x8 *)(LA + ra * LDS_ROW + kk + (lane >> 5) * 8);
                fb[i] = *(const bf16x8 *)(LB + rb * LDS_ROW + kk + (lane >> 5) * 8);
            }
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fa[i], fb[j], acc[i][j], 0, 0, 0);
        }
        if (more) {
            store_tile(la[buf ^ 1], va);
            store_tile(lb[buf ^ 1], vb);
        }
        __syncthreads();
        buf ^= 1;
    }

    // epilogue
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            const int n = tn * BN + wn * 64 + j * 32 + (lane & 31);
            if (n >= N) continue;
#pragma unroll
            for (int e = 0; e < 16; ++e) {
                const int m = tm * BM + wm * 64 + i * 32 + (e >> 2) * 8 + (lane >> 5) * 4 + (e & 3);
                if (m >= M) continue;
                float v = acc[i][j][e];
                if (EPI == EPI_PARTIAL) {
                    out.f32[(size_t)split * M * out.ldf + (size_t)m * out.ldf + n] = v;
                    continue;
                }
                if (out.bias) v += out.bias[n];
                if (out.pre) out.pre[(size_t)m * out.ldp + n] = v;
                if (EPI == EPI_GELU) v = gelu(v);
                if (out.f32) out.f32[(size_t)m * out.ldf + n] = v;
                if (out.b16) out.b16[(size_t)m * out.ldb + n] = bf16_rne(v);
            }
        }
}

// out[m][n] = sum over splits of part[s][m][n] (+ bias), then as the GEMM epilogue.
__global__ void reduce_splits(const float *__restrict__ part, int nsplit, int M, int N, int ldf, GemmOut out, int gelu_on)
{
    const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i >= (size_t)M * N) return;
    const int m = (int)(i / N), n = (int)(i % N);
    float v = 0.f;
    for (int s = 0; s < nsplit; ++s) v += part[(size_t)s * M * ldf + (size_t)m * ldf + n];
    if (out.bias) v += out.bias[n];
    if (out.pre) out.pre[(size_t)m * out.ldp + n] = v;
    if (gelu_on) v = gelu(v);
    if (out.f32) out.f32[(size_t)m * out.ldf + n] = v;
    if (out.b16) out.b16[(size_t)m * out.ldb + n] = bf16_rne(v);
}

// out[c][r] = in[r][c] for an R x C bf16 matrix (leading dimensions ldi, ldo), 32 x 32 tiles.
__global__ void transpose16(const uint16_t *__restrict__ in, int ldi, uint16_t *__restrict__ out, int ldo, int R, int C)
{
    __shared__ uint16_t tile[32][33];
    const int c0 = blockIdx.x * 32, r0 = blockIdx.y * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 256 threads: 8 rows per pass
    for (int y = ty; y < 32; y += 8) {
        const int r = r0 + y, c = c0 + tx;
        tile[y][tx] = (r < R && c < C) ? in[(size_t)r * ldi + c] : 0;
    }
    __syncthreads();
    for (int y = ty; y < 32; y += 8) {
        const int c = c0 + y, r = r0 + tx;
        if (c < C && r < R) out[(size_t)c * ldo + r] = tile[tx][y];
    }
}

// g[m][n] = dh[m][n] * gelu'(pre[m][n]) as bf16 (the gradient at the pre-activation).
__global__ void gelu_backward(const float *__restrict__ dh, const float *__restrict__ pre, int M, int N, int ld,
                              uint16_t *__restrict__ g, int ldg)
{
    const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i >= (size_t)M * N) return;
    const int m = (int)(i / N), n = (int)(i % N);
    g[(size_t)m * ldg + n] = bf16_rne(dh[(size_t)m * ld + n] * gelu_grad(pre[(size_t)m * ld + n]));
}

// The kernels below work on 64-column strips with 256 threads: thread (group q = t / 64, lane)
// owns column c0 + lane and rows q, q + 4, q + 8, ... of each 64-row chunk, so a wave reads 64
// consecutive elements of a row; 64 x 64 tiles pass through LDS where a transposed copy is written.

// MSE over the valid M x N outputs: g = bf16(2 (y - t) / (M N)) into g and its transpose gt
// ([N][ldgt]); db[c] = sum over the rows of the bf16 g (the bias gradient); partial[block] = the
// block's sum of (y - t)^2.
__global__ void __launch_bounds__(256) mse_grad(const float *__restrict__ y, int ldy, const float *__restrict__ tgt,
                                                int ldt, int M, int N, uint16_t *__restrict__ g, int ldg,
                                                uint16_t *__restrict__ gt, int ldgt, float *__restrict__ db,
                                                float *__restrict__ partial)
{
    __shared__ uint16_t tile[64][66];
    __shared__ float red[256];
    const int lane = threadIdx.x & 63, q = threadIdx.x >> 6;
    const int c0 = blockIdx.x * 64, c = c0 + lane;
    const float scale = 2.0f / ((float)M * (float)N);
    float loss = 0.f, col = 0.f;
    for (int r0 = 0; r0 < M; r0 += 64) {
        for (int k = 0; k < 16; ++k) {
            const int rr = q + 4 * k, r = r0 + rr;
            uint16_t v = 0;
            if (r < M && c < N) {
                const float d = y[(size_t)r * ldy + c] - tgt[(size_t)r * ldt + c];
                loss += d * d;
                v = bf16_rne(scale * d);
                g[(size_t)r * ldg + c] = v;
                col += bf16_to_f32(v);
            }
            tile[rr][lane] = v;
        }
        __syncthreads();
        for (int k = 0; k < 16; ++k) {
            const int cc = q + 4 * k, r = r0 + lane;
            if (c0 + cc < N && r < M) gt[(size_t)(c0 + cc) * ldgt + r] = tile[lane][cc];
        }
        __syncthreads();
    }
    red[threadIdx.x] = col;
    __syncthreads();
    if (q == 0 && c < N) db[c] = (red[lane] + red[lane + 64]) + (red[lane + 128] + red[lane + 192]);
    __syncthreads();
    red[threadIdx.x] = loss;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

// db[c] = sum over the M rows of a bf16 gradient g.
__global__ void __launch_bounds__(256) bias_grad(const uint16_t *__restrict__ g, int ldg, int M, int N,
                                                 float *__restrict__ db)
{
    __shared__ float red[256];
    const int lane = threadIdx.x & 63, q = threadIdx.x >> 6;
    const int c = blockIdx.x * 64 + lane;
    float s = 0.f;
    if (c < N)
        for (int r = q; r < M; r += 4) s += bf16_to_f32(g[(size_t)r * ldg + c]);
    red[threadIdx.x] = s;
    __syncthreads();
    if (q == 0 && c < N) db[c] = (red[lane] + red[lane + 64]) + (red[lane + 128] + red[lane + 192]);
}

__device__ __forceinline__ float adam_update(float *w, float *mom, float *vel, float g, size_t i, float lr, float b1,
                                             float b2, float eps, float c1, float c2)
{
    const float m = b1 * mom[i] + (1.f - b1) * g;
    const float v = b2 * vel[i] + (1.f - b2) * g * g;
    mom[i] = m;
    vel[i] = v;
    const float x = w[i] - lr * (m / c1) / (sqrtf(v / c2) + eps);
    w[i] = x;
    return x;
}

// One Adam step on n master weights (FP32), refreshing their bf16 copy (when given).
__global__ void adam(float *__restrict__ w, float *__restrict__ mom, float *__restrict__ vel, const float *__restrict__ grad,
                     size_t n, float lr, float b1, float b2, float eps, float c1, float c2, uint16_t *__restrict__ w16)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const float x = adam_update(w, mom, vel, grad[i], i, lr, b1, b2, eps, c1, c2);
        if (w16) w16[i] = bf16_rne(x);
    }
}

// Adam on an R x C weight matrix (row-major, leading dimension C) refreshing its bf16 copy w16
// and the transposed copy w16t ([C][R]); one 64 x 64 tile per block (R, C multiples of 64).
__global__ void __launch_bounds__(256) adam_matrix(float *__restrict__ w, float *__restrict__ mom,
                                                   float *__restrict__ vel, const float *__restrict__ grad, int R, int C,
                                                   float lr, float b1, float b2, float eps, float c1, float c2,
                                                   uint16_t *__restrict__ w16, uint16_t *__restrict__ w16t)
{
    __shared__ uint16_t tile[64][66];
    const int lane = threadIdx.x & 63, q = threadIdx.x >> 6;
    const int tiles_c = C / 64;
    const int r0 = (blockIdx.x / tiles_c) * 64, c0 = (blockIdx.x % tiles_c) * 64;
    for (int k = 0; k < 16; ++k) {
        const int rr = q + 4 * k;
        const size_t i = (size_t)(r0 + rr) * C + c0 + lane;
        const uint16_t v = bf16_rne(adam_update(w, mom, vel, grad[i], i, lr, b1, b2, eps, c1, c2));
        w16[i] = v;
        tile[rr][lane] = v;
    }
    __syncthreads();
    for (int k = 0; k < 16; ++k) {
        const int cc = q + 4 * k;
        w16t[(size_t)(c0 + cc) * R + r0 + lane] = tile[lane][cc];
    }
}

__global__ void to_bf16(const float *__restrict__ in, uint16_t *__restrict__ out, size_t n)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        out[i] = bf16_rne(in[i]);
}

int grid_for(size_t n, int bs = 256) { return (int)std::min<size_t>((n + bs - 1) / bs, 65536); }

}  // namespace

hipError_t gemm_bf16(const uint16_t *A, int lda, const uint16_t *B, int ldb, int M, int N, int K, int epi,
                     const GemmOut &out, float *work, size_t work_floats, int ksplit, hipStream_t stream)
{
    if (M < 1 || N < 1 || K < BK || K % BK || lda % 8 || ldb % 8 || lda < K || ldb < K) return hipErrorInvalidValue;
    const int tiles = ((M + BM - 1) / BM) * ((N + BN - 1) / BN);
    if (ksplit < 1) ksplit = 1;
    ksplit = std::min(ksplit, K / BK);
    if (ksplit == 1) {
        const dim3 grid(tiles, 1);
        if (epi == EPI_GELU) gemm_nt<EPI_GELU><<<grid, 256, 0, stream>>>(A, lda, B, ldb, M, N, K, 1, out);
        else gemm_nt<EPI_LINEAR><<<grid, 256, 0, stream>>>(A, lda, B, ldb, M, N, K, 1, out);
        return hipGetLastError();
    }
    if (!work || work_floats < (size_t)ksplit * M * N) return hipErrorInvalidValue;
    GemmOut part{};
    part.f32 = work;
    part.ldf = N;
    gemm_nt<EPI_PARTIAL><<<dim3(tiles, ksplit), 256, 0, stream>>>(A, lda, B, ldb, M, N, K, ksplit, part);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    reduce_splits<<<grid_for((size_t)M * N), 256, 0, stream>>>(work, ksplit, M, N, N, out, epi == EPI_GELU);
    return hipGetLastError();
}

hipError_t transpose_bf16(const uint16_t *in, int ldi, uint16_t *out, int ldo, int R, int C, hipStream_t stream)
{
    transpose16<<<dim3((C + 31) / 32, (R + 31) / 32), 256, 0, stream>>>(in, ldi, out, ldo, R, C);
    return hipGetLastError();
}

hipError_t gelu_bwd(const float *dh, const float *pre, int M, int N, int ld, uint16_t *g, int ldg, hipStream_t stream)
{
    gelu_backward<<<grid_for((size_t)M * N), 256, 0, stream>>>(dh, pre, M, N, ld, g, ldg);
    return hipGetLastError();
}

int mse_blocks(int N) { return (N + 63) / 64; }

hipError_t mse(const float *y, int ldy, const float *t, int ldt, int M, int N, uint16_t *g, int ldg, uint16_t *gt,
               int ldgt, float *db, float *partial, hipStream_t stream)
{
    mse_grad<<<mse_blocks(N), 256, 0, stream>>>(y, ldy, t, ldt, M, N, g, ldg, gt, ldgt, db, partial);
    return hipGetLastError();
}

hipError_t bias_gradient(const uint16_t *g, int ldg, int M, int N, float *db, hipStream_t stream)
{
    bias_grad<<<(N + 63) / 64, 256, 0, stream>>>(g, ldg, M, N, db);
    return hipGetLastError();
}

hipError_t adam_step(float *w, float *m, float *v, const float *g, size_t n, float lr, float b1, float b2, float eps,
                     long long t, uint16_t *w16, hipStream_t stream)
{
    const float c1 = 1.f - powf(b1, (float)t), c2 = 1.f - powf(b2, (float)t);
    adam<<<grid_for(n), 256, 0, stream>>>(w, m, v, g, n, lr, b1, b2, eps, c1, c2, w16);
    return hipGetLastError();
}

hipError_t adam_matrix_step(float *w, float *m, float *v, const float *g, int R, int C, float lr, float b1, float b2,
                            float eps, long long t, uint16_t *w16, uint16_t *w16t, hipStream_t stream)
{
    if (R % 64 || C % 64) return hipErrorInvalidValue;
    const float c1 = 1.f - powf(b1, (float)t), c2 = 1.f - powf(b2, (float)t);
    adam_matrix<<<(R / 64) * (C / 64), 256, 0, stream>>>(w, m, v, g, R, C, lr, b1, b2, eps, c1, c2, w16, w16t);
    return hipGetLastError();
}

hipError_t convert_bf16(const float *in, uint16_t *out, size_t n, hipStream_t stream)
{
    to_bf16<<<grid_for(n), 256, 0, stream>>>(in, out, n);
    return hipGetLastError();
}

}  // namespace mlp
}  // namespace poe2d
