// Kernels of the surrogate network: a BF16 GEMM on the matrix cores (FP32 accumulation) with a
// fused bias / GELU epilogue and split-K, transposes, GELU backward, bias gradients, the MSE
// gradient and the Adam update.
//
// GEMM layout ("NT"): C[m][n] = sum_k A[m][k] B[n][k], both operands K-contiguous bf16 rows.  A
// block of 256 threads (4 waves, 2 x 2) owns a 128 x 128 tile of C; each wave computes 64 x 64
// with 2 x 2 v_mfma_f32_32x32x16_bf16.  Operand tiles of 128 rows x 32 k go through LDS, double
// buffered, one barrier per k tile; loads run two to four k tiles ahead of the MFMAs, through
// register sets sized so that every variant runs three waves per SIMD without spilling.  MFMA lane
// layout (gfx950, 32x32x16 bf16): lane l supplies
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

// One Adam update of weight i (bias-corrected: c1 = 1 - b1^t, c2 = 1 - b2^t) from its gradient g.
__device__ __forceinline__ float adam_value(float w, float &m, float &v, float g, float lr, float b1, float b2,
                                            float eps, float c1, float c2)
{
    m = b1 * m + (1.f - b1) * g;
    v = b2 * v + (1.f - b2) * g * g;
    return w - lr * (m / c1) / (sqrtf(v / c2) + eps);
}

__device__ __forceinline__ float adam_update(float *w, float *mom, float *vel, float g, size_t i, float lr, float b1,
                                             float b2, float eps, float c1, float c2)
{
    float m = mom[i], v = vel[i];
    const float x = adam_value(w[i], m, v, g, lr, b1, b2, eps, c1, c2);
    mom[i] = m;
    vel[i] = v;
    w[i] = x;
    return x;
}

// C (+)= A B^T over k in [k0, k1) for the 128 x 128 tile (tm, tn); k0, k1 multiples of BK.
template <int EPI>
__global__ void __launch_bounds__(256, 3) gemm_nt(const uint16_t *__restrict__ A, int lda, const uint16_t *__restrict__ B,
                                               int ldb, int M, int N, int K, int ksplit, GemmOut out)
{
    // operand tiles (la, lb: two k tiles each), reused by the EPI_MSE / EPI_ADAM epilogues for the
    // transposed bf16 tile and the EPI_MSE sums (40 KiB)
    __shared__ uint16_t lds[2 * BM * LDS_ROW + 2 * BN * LDS_ROW];
    uint16_t(*la)[BM * LDS_ROW] = (uint16_t(*)[BM * LDS_ROW])lds;
    uint16_t(*lb)[BN * LDS_ROW] = (uint16_t(*)[BN * LDS_ROW])(lds + 2 * BM * LDS_ROW);
    // Tile order.  Row-major outputs: block b takes tile b, n fastest, so the blocks in flight write
    // long runs of the same rows.  EPI_MSE and split-K: blocks go to the 8 XCDs round-robin (block
    // b, counted x fastest, to XCD b % 8), and XCD x takes the x-th run of consecutive items
    // instead.  EPI_MSE walks m fastest, so the m tiles of one n tile follow each other on one XCD:
    // they share its operand tile in that XCD's L2 and finish the rows of the transposed gradient
    // together (225 -> 200 us on the 256 x 151552 x 512 output layer; the same order slowed the
    // plain epilogue from 89 to 104 us).  Split-K walks the tiles of one k range first, so every
    // tile of a k range runs on one XCD and each operand slice is read from HBM once.
    const int tiles_n = (N + BN - 1) / BN;
    int tm, tn, split = 0;
    if constexpr (EPI == EPI_MSE || EPI == EPI_PARTIAL) {
        const int nb = (int)(gridDim.x * gridDim.y), q8 = nb >> 3, r8 = nb & 7;
        const int b = (int)(blockIdx.y * gridDim.x + blockIdx.x);
        const int xcd = b & 7, slot = b >> 3;
        const int item = xcd < r8 ? xcd * (q8 + 1) + slot : r8 * (q8 + 1) + (xcd - r8) * q8 + slot;
        const int tile = item % (int)gridDim.x;
        if (EPI == EPI_PARTIAL) split = item / (int)gridDim.x;
        const int tiles_m = (M + BM - 1) / BM;
        tm = EPI == EPI_MSE ? tile % tiles_m : tile / tiles_n;
        tn = EPI == EPI_MSE ? tile / tiles_m : tile % tiles_n;
    } else {
        tm = (int)blockIdx.x / tiles_n;
        tn = (int)blockIdx.x % tiles_n;
    }
    const int kper = ((K / BK + ksplit - 1) / ksplit) * BK;
    const int k0 = split * kper, k1 = min(K, k0 + kper);
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int wm = wave >> 1, wn = wave & 1;
    // the bias of this lane's two result columns (wn * 64 + j * 32 + lane % 32)
    float bj[2];
    for (int j = 0; j < 2; ++j) {
        const int n = tn * BN + wn * 64 + j * 32 + (lane & 31);
        bj[j] = (EPI != EPI_PARTIAL && EPI != EPI_ADAM && out.bias && n < N) ? out.bias[n] : 0.f;
    }

    // each thread loads 2 x 16 bytes of each operand tile: rows r0 and r0 + 64, k chunk kc * 8
    const int r0 = t >> 2, kc = t & 3;
    // Loads are unconditional (rows past the edge read the last row, and are zeroed when stored to
    // LDS), so every path keeps the same number of loads in flight and the waits before the LDS
    // stores can leave the younger set's loads outstanding.
    bool a_ok[2], b_ok[2];
    for (int h = 0; h < 2; ++h) {
        a_ok[h] = tm * BM + r0 + h * 64 < M;
        b_ok[h] = tn * BN + r0 + h * 64 < N;
    }
    auto load_tile = [&](const uint16_t *P, int ld, int rows, int row0, int k, u16x8 *v) {
        for (int h = 0; h < 2; ++h) {
            const int r = min(row0 + r0 + h * 64, rows - 1);
            v[h] = *(const u16x8 *)(P + (size_t)r * ld + k + kc * 8);
        }
    };
    auto store_tile = [&](uint16_t *L, const u16x8 *v, const bool *ok) {
        for (int h = 0; h < 2; ++h)
            *(u16x8 *)(L + (r0 + h * 64) * LDS_ROW + kc * 8) = ok[h] ? v[h] : u16x8{0, 0, 0, 0, 0, 0, 0, 0};
    };

    f32x16 acc[2][2];
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j)
            for (int e = 0; e < 16; ++e) acc[i][j][e] = 0.f;

    // PF register sets carry operand tiles from memory: tile i lives in set i % PF, loaded at the
    // start of k step i - PF and stored to LDS at the end of step i - 1, so each load has about PF - 1
    // steps of MFMA work to land.  As many sets as fit beside each epilogue at three waves per SIMD
    // (168 registers) without spilling.
    constexpr int PF = (EPI == EPI_ADAM || EPI == EPI_MSE) ? 2 : EPI == EPI_PARTIAL ? 4 : 3;
    u16x8 ra[PF][2], rb[PF][2];
    const int nk = k1 > k0 ? (k1 - k0) / BK : 0;
    auto compute = [&](const uint16_t *LA, const uint16_t *LB) {
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
    };
    const int last = k0 + (nk - 1) * BK;  // tiles past the end re-read the last one (never used)
    auto kof = [&](int i) { return min(k0 + i * BK, last); };
    if (nk > 0) {
#pragma unroll
        for (int u = 0; u < PF; ++u) {
            load_tile(A, lda, M, tm * BM, kof(u), ra[u]);
            load_tile(B, ldb, N, tn * BN, kof(u), rb[u]);
        }
        store_tile(la[0], ra[0], a_ok);
        store_tile(lb[0], rb[0], b_ok);
    }
    __syncthreads();
    for (int it0 = 0; it0 < nk; it0 += PF) {
#pragma unroll
        for (int u = 0; u < PF; ++u) {
            const int it = it0 + u;
            if (it < nk) {
                // LDS buffer it & 1 holds tile it, so set u is free: refill it with tile it + PF
                load_tile(A, lda, M, tm * BM, kof(it + PF), ra[u]);
                load_tile(B, ldb, N, tn * BN, kof(it + PF), rb[u]);
                compute(la[it & 1], lb[it & 1]);
                if (it + 1 < nk) {
                    store_tile(la[(it + 1) & 1], ra[(u + 1) % PF], a_ok);
                    store_tile(lb[(it + 1) & 1], rb[(u + 1) % PF], b_ok);
                }
                __syncthreads();
            }
        }
    }

    if constexpr (EPI == EPI_MSE || EPI == EPI_ADAM) {
        // bf16 results of the tile go to memory row-major (g, or the updated weights' copy w16) and,
        // through a transposed tile in LDS, column-major (gt, or w16t) in 256-byte rows
        constexpr int TP = BM + 8;  // bf16 per transposed row
        uint16_t *tt = lds;
        float lsum = 0.f;
        __syncthreads();  // every wave is done with the operand tiles
        // Per 32 x 32 sub-tile: every operand load is issued before any store (the compiler cannot
        // tell the arrays apart and would otherwise hold each load behind the store before it).
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                const int nl = wn * 64 + j * 32 + (lane & 31), n = tn * BN + nl;
                auto row = [&](int e) { return wm * 64 + i * 32 + (e >> 2) * 8 + (lane >> 5) * 4 + (e & 3); };
                float va[16], vb[16], vc[16];
                if constexpr (EPI == EPI_MSE) {
                    // target rows: the batch's own, or looked up four at a time (the lane's rows come in
                    // runs of four, all below M or none: M % 4 == 0 with tgt_rows), lookups first
                    int tr[16];
#pragma unroll
                    for (int q = 0; q < 4; ++q) {
                        const int mg = tm * BM + row(4 * q);
                        if (out.tgt_rows) {
                            const int4 r4 = mg < M ? *(const int4 *)(out.tgt_rows + mg) : make_int4(0, 0, 0, 0);
                            tr[4 * q] = r4.x;
                            tr[4 * q + 1] = r4.y;
                            tr[4 * q + 2] = r4.z;
                            tr[4 * q + 3] = r4.w;
                        } else {
#pragma unroll
                            for (int u = 0; u < 4; ++u) tr[4 * q + u] = mg + u;
                        }
                    }
#pragma unroll
                    for (int e = 0; e < 16; ++e)
                        va[e] = (tm * BM + row(e) < M && n < N) ? out.tgt[(size_t)tr[e] * out.ldt + n] : 0.f;
                }
#pragma unroll
                for (int e = 0; e < 16; ++e) {
                    const int m = tm * BM + row(e);
                    const bool ok = m < M && n < N;
                    if constexpr (EPI == EPI_ADAM) {
                        // the optimiser state streams through once per step: non-temporal
                        const size_t idx = (size_t)m * out.ldw + n;
                        va[e] = ok ? __builtin_nontemporal_load(out.w + idx) : 0.f;
                        vb[e] = ok ? __builtin_nontemporal_load(out.mom + idx) : 0.f;
                        vc[e] = ok ? __builtin_nontemporal_load(out.vel + idx) : 0.f;
                    }
                }
#pragma unroll
                for (int e = 0; e < 16; ++e) {
                    const int ml = row(e), m = tm * BM + ml;
                    uint16_t h = 0;
                    if (m < M && n < N) {
                        if constexpr (EPI == EPI_MSE) {
                            // the output (product + bias) minus its target: loss and bf16 gradient
                            const float d = (acc[i][j][e] + bj[j]) - va[e];
                            lsum += d * d;
                            h = bf16_rne(out.scale * d);
                        } else {
                            // the product is this weight's gradient: Adam on the master weight
                            const size_t idx = (size_t)m * out.ldw + n;
                            float mo = vb[e], ve = vc[e];
                            const float x = adam_value(va[e], mo, ve, acc[i][j][e], out.lr, out.b1, out.b2, out.eps,
                                                       out.c1, out.c2);
                            __builtin_nontemporal_store(mo, out.mom + idx);
                            __builtin_nontemporal_store(ve, out.vel + idx);
                            __builtin_nontemporal_store(x, out.w + idx);
                            h = bf16_rne(x);
                        }
                    }
                    tt[nl * TP + ml] = h;
                }
            }
        __syncthreads();
        {
            const int nl = t >> 1, mh = (t & 1) * 64, n = tn * BN + nl;
            uint16_t *dst = EPI == EPI_MSE ? out.gt : out.w16t;
            const size_t ld = EPI == EPI_MSE ? (size_t)out.ldgt : (size_t)out.ld16t;
            if (n < N)
#pragma unroll
                for (int q = 0; q < 8; ++q) {
                    const int m0 = tm * BM + mh + q * 8;
                    const uint16_t *src = tt + nl * TP + mh + q * 8;
                    if (m0 + 8 <= M) {
                        *(u16x8 *)(dst + (size_t)n * ld + m0) = *(const u16x8 *)src;
                    } else {
                        for (int r = 0; r < 8 && m0 + r < M; ++r) dst[(size_t)n * ld + m0 + r] = src[r];
                    }
                }
        }
        {
            // row-major from the same tile (g, or w16): thread t writes row t / 2, columns
            // (t & 1) * 64 .. + 63, in 16-byte pieces
            const int ml = t >> 1, nh = (t & 1) * 64, m = tm * BM + ml;
            uint16_t *dst = EPI == EPI_MSE ? out.g : out.w16;
            const size_t ld = EPI == EPI_MSE ? (size_t)out.ldg : (size_t)out.ldw;
            if (m < M)
#pragma unroll
                for (int q = 0; q < 8; ++q) {
                    const int nl0 = nh + q * 8, n0 = tn * BN + nl0;
                    u16x8 v;
#pragma unroll
                    for (int r = 0; r < 8; ++r) v[r] = tt[(nl0 + r) * TP + ml];
                    if (n0 + 8 <= N) {
                        *(u16x8 *)(dst + (size_t)m * ld + n0) = v;
                    } else {
                        for (int r = 0; r < 8 && n0 + r < N; ++r) dst[(size_t)m * ld + n0 + r] = v[r];
                    }
                }
        }
        if constexpr (EPI == EPI_MSE) {
            // the bias gradient's partial sum over this tile's rows, per column, from the bf16
            // gradients in the transposed tile (fixed order: two halves of 64 rows, then the pair),
            // and the block's sum of squared errors (per wave by shuffles, then the four waves)
            const int nl = t >> 1, mh = (t & 1) * 64;
            float cs = 0.f;
            for (int q = 0; q < 64; ++q) cs += bf16_to_f32(tt[nl * TP + mh + q]);
            for (int o = 32; o > 0; o >>= 1) lsum += __shfl_xor(lsum, o);
            float *red = (float *)(lds + BN * TP);  // past the transposed tile
            red[t] = cs;
            if (lane == 0) red[256 + wave] = lsum;
            __syncthreads();
            const int n = tn * BN + nl;
            if ((t & 1) == 0 && n < N) out.dbp[(size_t)tm * out.lddb + n] = red[t] + red[t + 1];
            if (t == 0) out.partial[blockIdx.x] = (red[256] + red[257]) + (red[258] + red[259]);
        }
        return;
    }

    // epilogue: full tiles store without per-element bounds checks; the bias sits in registers
    // (loaded before the k loop), so no global load is pending here and the stores need no waits
    const bool full = (tm + 1) * BM <= M && (tn + 1) * BN <= N;
    auto emit = [&](int m, int n, float v, float b) {
        if (EPI == EPI_PARTIAL) {
            out.f32[(size_t)split * M * out.ldf + (size_t)m * out.ldf + n] = v;
            return;
        }
        v += b;
        if (out.pre) out.pre[(size_t)m * out.ldp + n] = v;
        if (EPI == EPI_GELU) v = gelu(v);
        if (out.f32) out.f32[(size_t)m * out.ldf + n] = v;
        if (out.b16) out.b16[(size_t)m * out.ldb + n] = bf16_rne(v);
    };
    if (full) {
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                const int n = tn * BN + wn * 64 + j * 32 + (lane & 31);
#pragma unroll
                for (int e = 0; e < 16; ++e)
                    emit(tm * BM + wm * 64 + i * 32 + (e >> 2) * 8 + (lane >> 5) * 4 + (e & 3), n, acc[i][j][e], bj[j]);
            }
    } else {
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                const int n = tn * BN + wn * 64 + j * 32 + (lane & 31);
#pragma unroll
                for (int e = 0; e < 16; ++e) {
                    const int m = tm * BM + wm * 64 + i * 32 + (e >> 2) * 8 + (lane >> 5) * 4 + (e & 3);
                    if (m < M && n < N) emit(m, n, acc[i][j][e], bj[j]);
                }
            }
    }
}

// db[n] = sum over r < rows of part[r][n] (n < N), in order.
__global__ void column_sums(const float *__restrict__ part, int rows, int ld, int N, float *__restrict__ db)
{
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    float v = 0.f;
    for (int r = 0; r < rows; ++r) v += part[(size_t)r * ld + n];
    db[n] = v;
}

// out[m][n] = sum over splits of part[s][m][n] (+ bias), then as the GEMM epilogue.
__global__ void reduce_splits(const float *__restrict__ part, int nsplit, int M, int N, int ldf, GemmOut out, int gelu_on)
{
    const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i >= (size_t)M * N) return;
    const int m = (int)(i / N), n = (int)(i % N);
    // eight partials in flight at a time, summed in split order
    const float *p = part + (size_t)m * ldf + n;
    const size_t stride = (size_t)M * ldf;
    float v = 0.f;
    int s = 0;
    for (; s + 8 <= nsplit; s += 8) {
        float q[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) q[u] = p[(size_t)(s + u) * stride];
#pragma unroll
        for (int u = 0; u < 8; ++u) v += q[u];
    }
    for (; s < nsplit; ++s) v += p[(size_t)s * stride];
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

// db[r] = sum of row r (K entries) of the transposed bf16 gradient gt [R][ldgt]: one wave per row.
__global__ void __launch_bounds__(256) bias_grad(const uint16_t *__restrict__ gt, int ldgt, int R, int K,
                                                 float *__restrict__ db)
{
    const int r = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (r >= R) return;
    float s = 0.f;
    for (int k = lane; k < K; k += 64) s += bf16_to_f32(gt[(size_t)r * ldgt + k]);
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
    if (lane == 0) db[r] = s;
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
    if (epi == EPI_MSE || epi == EPI_ADAM) {
        // (no split: the epilogue needs the whole sum)
        if (out.ldg % 8 && epi == EPI_MSE) return hipErrorInvalidValue;
        const dim3 grid(tiles, 1);
        if (epi == EPI_MSE) {
            if (!out.tgt || !out.g || !out.gt || !out.partial || !out.dbp || out.ldgt % 8 || (uintptr_t)out.g % 16 ||
                (uintptr_t)out.gt % 16 || (out.tgt_rows && (M % 4 || (uintptr_t)out.tgt_rows % 16)))
                return hipErrorInvalidValue;
            gemm_nt<EPI_MSE><<<grid, 256, 0, stream>>>(A, lda, B, ldb, M, N, K, 1, out);
        } else {
            if (!out.w || !out.mom || !out.vel || !out.w16 || !out.w16t || out.ld16t % 8 || out.ldw % 8 ||
                (uintptr_t)out.w16 % 16 || (uintptr_t)out.w16t % 16)
                return hipErrorInvalidValue;
            gemm_nt<EPI_ADAM><<<grid, 256, 0, stream>>>(A, lda, B, ldb, M, N, K, 1, out);
        }
        return hipGetLastError();
    }
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

hipError_t sum_rows(const float *part, int rows, int ld, int N, float *db, hipStream_t stream)
{
    column_sums<<<(N + 255) / 256, 256, 0, stream>>>(part, rows, ld, N, db);
    return hipGetLastError();
}

hipError_t bias_gradient(const uint16_t *gt, int ldgt, int R, int K, float *db, hipStream_t stream)
{
    bias_grad<<<(R + 3) / 4, 256, 0, stream>>>(gt, ldgt, R, K, db);
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
