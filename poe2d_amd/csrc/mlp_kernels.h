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

enum { EPI_LINEAR = 0, EPI_GELU = 1, EPI_PARTIAL = 2 };

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
};

// C = A B^T (M x N, reduction over K), A [M][lda] and B [N][ldb] bf16 with K contiguous; K and the
// leading dimensions multiples of 32 / 8.  epi: EPI_LINEAR or EPI_GELU.  ksplit > 1 splits K over
// blocks into `work` (ksplit * M * N floats) and reduces.
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
hipError_t mse(const float *y, int ldy, const float *t, int ldt, int M, int N, uint16_t *g, int ldg, uint16_t *gt,
               int ldgt, float *db, float *partial, hipStream_t stream);

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
