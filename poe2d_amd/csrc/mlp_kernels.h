// Launchers of the surrogate network's kernels (mlp_kernels.hip) and the bf16 helpers they share
// with the host code.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstring>

namespace poe2d {
namespace mlp {

// bf16 bits of x, round to nearest even (NaN stays NaN).
__host__ __device__ inline uint16_t bf16_rne(float x)
{
    uint32_t u;
    __builtin_memcpy(&u, &x, 4);
    if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x007fffffu)) return (uint16_t)((u >> 16) | 0x40);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

__host__ __device__ inline float bf16_to_f32(uint16_t h)
{
    const uint32_t u = (uint32_t)h << 16;
    float f;
    __builtin_memcpy(&f, &u, 4);
    return f;
}

// EPI_MSE: instead of the output, the mean-squared-error gradient of output (product + bias)
// against a target, bf16, row-major and transposed, with per-block sums of squared errors.
// EPI_ADAM: the product is a weight gradient; the epilogue takes the Adam step of those master
// weights and writes their bf16 copy and its transpose.
enum { EPI_LINEAR = 0, EPI_GELU = 1, EPI_PARTIAL = 2, EPI_MSE = 3, EPI_ADAM = 4 };

// Where a GEMM's results go: bias added first (when given); `pre` receives the value before the
// activation; f32 / b16 the value after it.  Any pointer may be null.
struct GemmOut {
    float *f32 = nullptr;
    int ldf = 0;
    uint16_t *b16 = nullptr;
    int ldb = 0;
    float *pre = nullptr;
    int ldp = 0;
    const float *bias = nullptr;
    // EPI_MSE: g[m][n] = bf16(scale (out - tgt)), gt[n][m] its transpose, partial[block] = the
    // block's sum of (out - tgt)^2, dbp[tile row][n] = the sum of g over the tile's rows (lddb
    // apart; sum_rows adds them up into the bias gradient).  Target row m is tgt row m, or row
    // tgt_rows[m] when tgt_rows is given (a batch read in place from a dataset).
    const float *tgt = nullptr;
    int ldt = 0;
    const int *tgt_rows = nullptr;
    float scale = 0.f;
    uint16_t *g = nullptr, *gt = nullptr;
    int ldg = 0, ldgt = 0;
    float *partial = nullptr;
    float *dbp = nullptr;
    int lddb = 0;
    // EPI_ADAM: master weights w (leading dimension ldw; the product's [m][n] is weight [m][n]), Adam
    // moments, the step's constants, the bf16 copy w16 (ldw) and its transpose w16t (ld16t)
    float *w = nullptr, *mom = nullptr, *vel = nullptr;
    int ldw = 0;
    float lr = 0.f, b1 = 0.f, b2 = 0.f, eps = 0.f, c1 = 1.f, c2 = 1.f;
    uint16_t *w16 = nullptr, *w16t = nullptr;
    int ld16t = 0;
};

// C = A B^T (M x N, reduction over K), A [M][lda] and B [N][ldb] bf16 with K contiguous; K and the
// leading dimensions multiples of 32 / 8.  epi: EPI_LINEAR, EPI_GELU, EPI_MSE (partial holds one
// entry per 128 x 128 tile) or EPI_ADAM.  ksplit > 1 splits K over blocks into `work` (ksplit * M *
// N floats) and reduces (EPI_LINEAR and EPI_GELU only).
hipError_t gemm_bf16(const uint16_t *A, int lda, const uint16_t *B, int ldb, int M, int N, int K, int epi,
                     const GemmOut &out, float *work, size_t work_floats, int ksplit, hipStream_t stream);

// out[c][r] = in[r][c], R x C.
hipError_t transpose_bf16(const uint16_t *in, int ldi, uint16_t *out, int ldo, int R, int C, hipStream_t stream);

// g = dh * gelu'(pre) (bf16), M x N with leading dimension ld (ldg for g).
hipError_t gelu_bwd(const float *dh, const float *pre, int M, int N, int ld, uint16_t *g, int ldg, hipStream_t stream);

// Mean squared error over M x N: gradient 2 (y - t) / (M N) as bf16 into g and its transpose gt
// ([N][ldgt]), its column sums (the bias gradient) into db[N], and per-block sums of squares into
// partial[mse_blocks(N)].
int mse_blocks(int N);
// Entries of `partial` the EPI_MSE product writes for an M x N output.
inline int gemm_tiles(int M, int N) { return ((M + 127) / 128) * ((N + 127) / 128); }
hipError_t mse(const float *y, int ldy, const float *t, int ldt, int M, int N, uint16_t *g, int ldg, uint16_t *gt,
               int ldgt, float *db, float *partial, hipStream_t stream);

// db[n] = sum over r < rows of part[r * ld + n], n < N (the bias gradient from EPI_MSE's dbp).
hipError_t sum_rows(const float *part, int rows, int ld, int N, float *db, hipStream_t stream);

// db[r] = sum_k gt[r][k] over the K = batch entries of the transposed gradient gt [R][ldgt].
hipError_t bias_gradient(const uint16_t *gt, int ldgt, int R, int K, float *db, hipStream_t stream);

// Adam step t (1-based) on n FP32 master weights; their bf16 copy w16 (may be null) follows.
hipError_t adam_step(float *w, float *m, float *v, const float *g, size_t n, float lr, float b1, float b2, float eps,
                     long long t, uint16_t *w16, hipStream_t stream);

// Adam step t on an R x C weight matrix (R, C multiples of 64) refreshing its bf16 copy w16 and
// the transposed copy w16t [C][R].
hipError_t adam_matrix_step(float *w, float *m, float *v, const float *g, int R, int C, float lr, float b1, float b2,
                            float eps, long long t, uint16_t *w16, uint16_t *w16t, hipStream_t stream);

hipError_t convert_bf16(const float *in, uint16_t *out, size_t n, hipStream_t stream);

}  // namespace mlp
}  // namespace poe2d
