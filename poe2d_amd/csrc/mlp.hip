// The surrogate network behind the C ABI: an ensemble of fully connected networks
// sizes[0] -> sizes[1] -> ... -> sizes[L] with GELU between layers (none after the last), bf16
// weights and activations, FP32 accumulation, FP32 master weights, Adam.  Every GEMM runs on the
// matrix cores (mlp_kernels.hip); the training step is
//   forward:   Z_l = H_l W_l^T + b_l,  H_{l+1} = gelu(Z_l)  (Y = Z_{L-1})
//   loss:      mean over the batch and the outputs of (Y - T)^2,  G_{L-1} = dLoss/dY
//   backward:  dW_l = G_l^T H_l,  db_l = sum_rows G_l,  G_{l-1} = (G_l W_l) * gelu'(Z_{l-1})
//   update:    Adam on every member's W_l, b_l (bias-corrected, step count per member).
// Widths are padded to multiples of 128 in storage (zero rows / columns that stay zero), so every
// GEMM is an NT product of K-contiguous bf16 rows; G^T and H^T come from transposes (the output
// gradient's from the MSE pass itself), and each layer keeps a transposed bf16 copy of W_l, which
// the Adam pass writes, for the G_l W_l product.  Members run one after the
// other on the handle's stream and share the activation buffers.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "../../include/poe2d.h"
#include "common.h"
#include "mlp_kernels.h"

using namespace poe2d;
using namespace poe2d::mlp;

namespace {

#define HIP_TRY(expr)                                                                              \
    do {                                                                                           \
        hipError_t e_ = (expr);                                                                    \
        if (e_ != hipSuccess)                                                                      \
            return fail(POE2D_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));           \
    } while (0)

int pad128(int n) { return (n + 127) / 128 * 128; }

// SplitMix64: deterministic initial weights from the seed.
uint64_t splitmix(uint64_t &s)
{
    uint64_t z = (s += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

struct Layer {
    int in = 0, out = 0, inp = 0, outp = 0;  // widths, padded widths
    float *w = nullptr, *b = nullptr;        // master [outp][inp], [outp]
    float *mw = nullptr, *vw = nullptr, *mb = nullptr, *vb = nullptr;
    float *gw = nullptr, *gb = nullptr;      // gradients
    uint16_t *w16 = nullptr;                 // [outp][inp]
    uint16_t *w16t = nullptr;                // [inp][outp]
};

struct Member {
    std::vector<Layer> layers;
    long long t = 0;  // Adam steps taken
};

template <class P>
poe2d_status dmalloc(P **p, size_t count)
{
    void *q = nullptr;
    hipError_t e = hipMalloc(&q, std::max<size_t>(count * sizeof(P), 1));
    if (e != hipSuccess)
        return fail(e == hipErrorOutOfMemory ? POE2D_E_NOMEM : POE2D_E_HIP,
                    std::string("hipMalloc: ") + hipGetErrorString(e));
    HIP_TRY(hipMemset(q, 0, std::max<size_t>(count * sizeof(P), 1)));
    *p = (P *)q;
    return POE2D_OK;
}

template <class P>
void dfree(P *&p)
{
    if (p) (void)hipFree((void *)p);
    p = nullptr;
}

}  // namespace

struct poe2d_mlp {
    int device = 0;
    hipStream_t stream = nullptr;
    std::vector<int> sizes;
    int L = 0;
    int max_batch = 0;
    std::vector<Member> members;
    // activations (shared by the members): per layer boundary l = 0..L, [max_batch][pad(sizes[l])]
    std::vector<uint16_t *> h16;   // bf16 inputs of each layer (h16[0] = network input)
    std::vector<uint16_t *> h16t;  // their transposes [pad][max_batch]
    std::vector<float *> pre;      // pre-activations Z_l, l = 0..L-1
    std::vector<uint16_t *> g16;   // gradients at Z_l (bf16)
    std::vector<uint16_t *> g16t;  // their transposes
    float *dh = nullptr;           // G_l W_l (FP32), largest hidden width
    float *y = nullptr;            // outputs Z_{L-1} (FP32)
    float *tgt = nullptr;          // targets of the batch
    float *xin = nullptr;          // network input of the batch (FP32)
    float *work = nullptr;         // split-K partials
    size_t work_floats = 0;
    float *partial = nullptr;      // loss partial sums
    int nred = 512;
    // prediction accumulators
    float *acc_mean = nullptr, *acc_m2 = nullptr;
    float *oscale = nullptr;  // per-output factor of the predicted mean (null: none)
    // resident dataset
    float *data_x = nullptr, *data_t = nullptr;
    int data_n = 0;
    int *idx = nullptr;      // batch row indices of up to kIndexSteps training steps
    float *dloss = nullptr;  // their losses
    // poe2d_mlp_train_many: a second set of activations, gradients and partials with its own stream
    // (same sizes, no members or dataset of its own), the step indices and losses of every member
    // trained, and one event per member ordering its steps across the two streams
    poe2d_mlp *twin = nullptr;
    int *idx_all = nullptr;
    size_t idx_all_cap = 0;
    float *loss_all = nullptr;
    size_t loss_all_cap = 0;
    std::vector<hipEvent_t> step_done;
};

// Training steps whose batch indices go to the device in one copy: a copy from pageable host
// memory holds the host until the stream reaches it, so one per step would leave the device idle
// while the host issues each step's launches.
constexpr int kIndexSteps = 1024;

namespace {

poe2d_status check_mlp(poe2d_mlp *m)
{
    if (!m) return fail(POE2D_E_INVALID, "null network handle");
    HIP_TRY(hipSetDevice(m->device));
    return POE2D_OK;
}

void free_mlp(poe2d_mlp *m)
{
    for (Member &e : m->members)
        for (Layer &l : e.layers) {
            dfree(l.w); dfree(l.b); dfree(l.mw); dfree(l.vw); dfree(l.mb); dfree(l.vb);
            dfree(l.gw); dfree(l.gb); dfree(l.w16); dfree(l.w16t);
        }
    for (auto &p : m->h16) dfree(p);
    for (auto &p : m->h16t) dfree(p);
    for (auto &p : m->pre) dfree(p);
    for (auto &p : m->g16) dfree(p);
    for (auto &p : m->g16t) dfree(p);
    dfree(m->dh); dfree(m->y); dfree(m->tgt); dfree(m->xin); dfree(m->work); dfree(m->partial);
    dfree(m->acc_mean); dfree(m->acc_m2); dfree(m->oscale); dfree(m->data_x); dfree(m->data_t); dfree(m->idx); dfree(m->dloss);
    dfree(m->idx_all); dfree(m->loss_all);
    for (hipEvent_t ev : m->step_done) (void)hipEventDestroy(ev);
    m->step_done.clear();
    if (m->twin) {
        if (m->twin->stream) (void)hipStreamSynchronize(m->twin->stream);
        free_mlp(m->twin);
        delete m->twin;
        m->twin = nullptr;
    }
    if (m->stream) (void)hipStreamDestroy(m->stream);
    m->stream = nullptr;
}

// Refresh the bf16 copy and its transpose of layer l's weights from the master weights.
poe2d_status sync_copies(poe2d_mlp *m, Layer &l)
{
    HIP_TRY(convert_bf16(l.w, l.w16, (size_t)l.outp * l.inp, m->stream));
    HIP_TRY(transpose_bf16(l.w16, l.inp, l.w16t, l.outp, l.outp, l.inp, m->stream));
    return POE2D_OK;
}

__global__ void pack_input(const float *__restrict__ x, int n, int in, uint16_t *__restrict__ h, int ldh)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n * in) return;
    h[(size_t)(i / in) * ldh + i % in] = bf16_rne(x[i]);
}

__global__ void gather_rows(const float *__restrict__ src, int ld_src, const int *__restrict__ idx, int n, int cols,
                            float *__restrict__ dst, int ld_dst)
{
    const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i >= (size_t)n * cols) return;
    const int r = (int)(i / cols), c = (int)(i % cols);
    dst[(size_t)r * ld_dst + c] = src[(size_t)idx[r] * ld_src + c];
}

// Running mean / sum of squared deviations over members (Welford), k = member index.
__global__ void welford(const float *__restrict__ y, int ldy, int n, int cols, int k, float *__restrict__ mean,
                        float *__restrict__ m2)
{
    const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i >= (size_t)n * cols) return;
    const int r = (int)(i / cols), c = (int)(i % cols);
    const float v = y[(size_t)r * ldy + c];
    const size_t j = (size_t)r * cols + c;
    if (k == 0) {
        mean[j] = v;
        m2[j] = 0.f;
        return;
    }
    const float d = v - mean[j];
    const float mu = mean[j] + d / (float)(k + 1);
    m2[j] += d * (v - mu);
    mean[j] = mu;
}

// spread[r] = ||std over members|| / ||mean|| over the outputs of row r (one block per row).
__global__ void row_spread(const float *__restrict__ mean, const float *__restrict__ m2, int cols, int members,
                           float *__restrict__ spread)
{
    __shared__ double sv[256], sn[256];
    const int r = blockIdx.x;
    double var = 0, nrm = 0;
    for (int c = threadIdx.x; c < cols; c += blockDim.x) {
        const float mu = mean[(size_t)r * cols + c];
        var += m2[(size_t)r * cols + c];
        nrm += (double)mu * mu;
    }
    sv[threadIdx.x] = var;
    sn[threadIdx.x] = nrm;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            sv[threadIdx.x] += sv[threadIdx.x + s];
            sn[threadIdx.x] += sn[threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) spread[r] = (float)(sqrt(sv[0] / (members > 1 ? members - 1 : 1)) / fmax(sqrt(sn[0]), 1e-30));
}

// mean[r][c] *= scale[c] for n rows of `cols` values.
__global__ void scale_cols(float *__restrict__ mean, int n, int cols, const float *__restrict__ scale)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < (size_t)n * cols;
         i += (size_t)gridDim.x * blockDim.x)
        mean[i] *= scale[i % cols];
}

int blocks(size_t n) { return (int)std::min<size_t>((n + 255) / 256, 1u << 20); }

// *loss = scale x the sum of count partials (FP64 accumulation: thread t sums partials t, t + 256,
// ... in order, then a fixed tree), one block of 256 threads.
__global__ void loss_sum(const float *__restrict__ partial, int count, double scale, float *__restrict__ loss)
{
    __shared__ double red[256];
    double v = 0;
    for (int i = threadIdx.x; i < count; i += 256) v += partial[i];
    red[threadIdx.x] = v;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) *loss = (float)(red[0] * scale);
}

// rows x K FP32 values into rows x Kp bf16 values (round to nearest even), zeros in columns K..Kp-1.
__global__ void pack_bf16_rows(const float *__restrict__ src, int rows, int K, int Kp, uint16_t *__restrict__ dst)
{
    const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i >= (size_t)rows * Kp) return;
    const int r = (int)(i / Kp), k = (int)(i % Kp);
    dst[i] = k < K ? bf16_rne(src[(size_t)r * K + k]) : (uint16_t)0;
}

// Device scratch of poe2d_gemm_bf16_noalloc: operand staging (FP32 as copied from the host, then
// bf16), the product and the split-K partials.  Each device the library runs on holds its own copy
// from module load until the process ends.  Host code must not pass `scratch` itself to a kernel or
// copy (in host code the name is a stand-in, not a device address): it uses the address
// hipGetSymbolAddress returns for the current device, kept in g_scratch_addr.
constexpr size_t kGemmScratchBytes = size_t(32) << 20;
__device__ __align__(256) unsigned char scratch[kGemmScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
unsigned char *g_scratch_addr[kMaxDevices] = {};

// Device address of `scratch` on the current device, looked up once per device.
poe2d_status scratch_address(unsigned char **out)
{
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return fail(POE2D_E_INVALID, "gemm_bf16_noalloc: device index out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_addr[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(POE2D_E_HIP, std::string("gemm_bf16_noalloc: hipGetSymbolAddress(scratch): ") +
                                         hipGetErrorString(e));
        g_scratch_addr[dev] = (unsigned char *)p;
    }
    *out = g_scratch_addr[dev];
    return POE2D_OK;
}

// k splits of a product with a plain epilogue (M x N over k = K): doubled while the tiles times
// the splits stay under 1024 blocks (about four per CU) and each split keeps at least 2 k tiles of
// 32, within the partials buffer.  The 512-wide hidden layers' products (8 or 16 tiles) run on 64
// blocks instead of 8; the output layer's input gradient (K = 151 552) on 1024.
int product_splits(const poe2d_mlp *m, int M, int N, int K)
{
    int ksplit = 1;
    const int tiles = ((M + 127) / 128) * ((N + 127) / 128);
    while (tiles * ksplit < 1024 && K / (2 * ksplit) >= 64 && (size_t)ksplit * 2 * M * N <= m->work_floats)
        ksplit *= 2;
    return ksplit;
}

// Forward pass of member e on the first n rows of m->xin (already packed into h16[0]).
poe2d_status forward(poe2d_mlp *m, Member &e, int n)
{
    for (int li = 0; li < m->L; ++li) {
        Layer &l = e.layers[li];
        const bool last = li == m->L - 1;
        GemmOut o;
        o.bias = l.b;
        o.pre = last ? nullptr : m->pre[li];
        o.ldp = l.outp;
        if (last) {
            o.f32 = m->y;
            o.ldf = l.outp;
        } else {
            o.b16 = m->h16[li + 1];
            o.ldb = l.outp;
        }
        HIP_TRY(gemm_bf16(m->h16[li], l.inp, l.w16, l.inp, n, l.out, l.inp, last ? EPI_LINEAR : EPI_GELU, o, m->work,
                          m->work_floats, product_splits(m, n, l.out, l.inp), m->stream));
    }
    return POE2D_OK;
}


// Gradients of the mean squared error for member e on the batch in h16[0] / tgt (n rows, n a
// multiple of 32); loss (mean) to *loss when non-null.
poe2d_status backward(poe2d_mlp *m, Member &e, int n, float *loss)
{
    const int L = m->L;
    Layer &top = e.layers[L - 1];
    // the output gradient, its transpose and the last bias gradient in one pass
    HIP_TRY(mse(m->y, top.outp, m->tgt, top.outp, n, top.out, m->g16[L - 1], top.outp, m->g16t[L - 1], m->max_batch,
                top.gb, m->partial, m->stream));
    if (loss) {
        std::vector<float> part(mse_blocks(top.out));
        HIP_TRY(hipMemcpyAsync(part.data(), m->partial, part.size() * sizeof(float), hipMemcpyDeviceToHost, m->stream));
        HIP_TRY(hipStreamSynchronize(m->stream));
        double s = 0;
        for (float v : part) s += v;
        *loss = (float)(s / ((double)n * top.out));
    }
    for (int li = L - 1; li >= 0; --li) {
        Layer &l = e.layers[li];
        // dW = G^T H: A = G^T [outp][n], B = H^T [inp][n]
        if (li < L - 1) HIP_TRY(transpose_bf16(m->g16[li], l.outp, m->g16t[li], m->max_batch, n, l.outp, m->stream));
        HIP_TRY(transpose_bf16(m->h16[li], l.inp, m->h16t[li], m->max_batch, n, l.inp, m->stream));
        GemmOut o;
        o.f32 = l.gw;
        o.ldf = l.inp;
        HIP_TRY(gemm_bf16(m->g16t[li], m->max_batch, m->h16t[li], m->max_batch, l.outp, l.inp, n, EPI_LINEAR, o,
                          m->work, m->work_floats, product_splits(m, l.outp, l.inp, n), m->stream));
        if (li < L - 1) HIP_TRY(bias_gradient(m->g16t[li], m->max_batch, l.outp, n, l.gb, m->stream));
        if (li == 0) break;
        // G_{l-1} = (G_l W_l) * gelu'(Z_{l-1}): A = G_l [n][outp], B = W_l^T [inp][outp]
        const int K = l.outp;
        const int ksplit = product_splits(m, n, l.in, K);
        GemmOut od;
        od.f32 = m->dh;
        od.ldf = l.inp;
        HIP_TRY(gemm_bf16(m->g16[li], l.outp, l.w16t, l.outp, n, l.in, K, EPI_LINEAR, od, m->work, m->work_floats,
                          ksplit, m->stream));
        HIP_TRY(gelu_bwd(m->dh, m->pre[li - 1], n, l.in, l.inp, m->g16[li - 1], l.inp, m->stream));
    }
    return POE2D_OK;
}

// One training step of member e on the batch in h16[0] (n rows, a multiple of 32) against targets
// tgt (row r of the batch: row tgt_rows[r], or r when tgt_rows is null; ldt apart), the mean loss
// to device memory *loss_dev when non-null.  The output layer's forward product ends in the loss gradient
// (EPI_MSE: no output array is written), and its weight-gradient product ends in the Adam step
// (EPI_ADAM: the gradient never reaches memory); each layer's input gradient is formed from its
// weights before they are updated.  Every value is computed as forward() and backward() compute
// it, followed by the Adam step of each layer, except for the summation order of the output bias
// gradient and of the loss.
poe2d_status train_step(poe2d_mlp *m, Member &e, int n, const float *tgt, int ldt, const int *tgt_rows, float lr,
                        float *loss_dev)
{
    const int L = m->L;
    for (int li = 0; li < L - 1; ++li) {
        Layer &l = e.layers[li];
        GemmOut o;
        o.bias = l.b;
        o.pre = m->pre[li];
        o.ldp = l.outp;
        o.b16 = m->h16[li + 1];
        o.ldb = l.outp;
        HIP_TRY(gemm_bf16(m->h16[li], l.inp, l.w16, l.inp, n, l.out, l.inp, EPI_GELU, o, m->work, m->work_floats,
                          product_splits(m, n, l.out, l.inp), m->stream));
    }
    Layer &top = e.layers[L - 1];
    {
        GemmOut o;
        o.bias = top.b;
        o.tgt = tgt;
        o.ldt = ldt;
        o.tgt_rows = tgt_rows;
        o.scale = 2.0f / ((float)n * (float)top.out);
        o.g = m->g16[L - 1];
        o.ldg = top.outp;
        o.gt = m->g16t[L - 1];
        o.ldgt = m->max_batch;
        o.partial = m->partial;
        o.dbp = m->work;  // (the split-K partials come later)
        o.lddb = top.outp;
        HIP_TRY(gemm_bf16(m->h16[L - 1], top.inp, top.w16, top.inp, n, top.out, top.inp, EPI_MSE, o, m->work,
                          m->work_floats, 1, m->stream));
    }
    HIP_TRY(sum_rows(m->work, (n + 127) / 128, top.outp, top.out, top.gb, m->stream));
    if (loss_dev) {
        loss_sum<<<1, 256, 0, m->stream>>>(m->partial, gemm_tiles(n, top.out), 1.0 / ((double)n * top.out), loss_dev);
        HIP_TRY(hipGetLastError());
    }
    e.t += 1;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const float c1 = 1.f - powf(b1, (float)e.t), c2 = 1.f - powf(b2, (float)e.t);
    for (int li = L - 1; li >= 0; --li) {
        Layer &l = e.layers[li];
        if (li < L - 1) HIP_TRY(transpose_bf16(m->g16[li], l.outp, m->g16t[li], m->max_batch, n, l.outp, m->stream));
        HIP_TRY(transpose_bf16(m->h16[li], l.inp, m->h16t[li], m->max_batch, n, l.inp, m->stream));
        if (li > 0) {
            // G_{l-1} = (G_l W_l) * gelu'(Z_{l-1}) with the weights of this step
            const int K = l.outp;
            const int ksplit = product_splits(m, n, l.in, K);
            GemmOut od;
            od.f32 = m->dh;
            od.ldf = l.inp;
            HIP_TRY(gemm_bf16(m->g16[li], l.outp, l.w16t, l.outp, n, l.in, K, EPI_LINEAR, od, m->work, m->work_floats,
                              ksplit, m->stream));
            HIP_TRY(gelu_bwd(m->dh, m->pre[li - 1], n, l.in, l.inp, m->g16[li - 1], l.inp, m->stream));
        }
        if (li == L - 1) {
            // dW = G^T H and the Adam step in one product: A = G^T [outp][n], B = H^T [inp][n]
            GemmOut ow;
            ow.w = l.w;
            ow.mom = l.mw;
            ow.vel = l.vw;
            ow.ldw = l.inp;
            ow.lr = lr;
            ow.b1 = b1;
            ow.b2 = b2;
            ow.eps = eps;
            ow.c1 = c1;
            ow.c2 = c2;
            ow.w16 = l.w16;
            ow.w16t = l.w16t;
            ow.ld16t = l.outp;
            HIP_TRY(gemm_bf16(m->g16t[li], m->max_batch, m->h16t[li], m->max_batch, l.outp, l.inp, n, EPI_ADAM, ow,
                              m->work, m->work_floats, 1, m->stream));
        } else {
            // (the hidden layers' 512 x 512 products make 16 tiles: too few blocks to stream the
            // optimiser state from the product's epilogue; a separate Adam pass over 64 tiles is faster)
            GemmOut o;
            o.f32 = l.gw;
            o.ldf = l.inp;
            HIP_TRY(gemm_bf16(m->g16t[li], m->max_batch, m->h16t[li], m->max_batch, l.outp, l.inp, n, EPI_LINEAR, o,
                              m->work, m->work_floats, product_splits(m, l.outp, l.inp, n), m->stream));
            HIP_TRY(bias_gradient(m->g16t[li], m->max_batch, l.outp, n, l.gb, m->stream));
            HIP_TRY(adam_matrix_step(l.w, l.mw, l.vw, l.gw, l.outp, l.inp, lr, b1, b2, eps, e.t, l.w16, l.w16t,
                                     m->stream));
        }
        HIP_TRY(adam_step(l.b, l.mb, l.vb, l.gb, (size_t)l.outp, lr, b1, b2, eps, e.t, nullptr, m->stream));
    }
    return POE2D_OK;
}

// Host rows x[n][sizes[0]] into h16[0].
poe2d_status load_input(poe2d_mlp *m, const float *x, int n)
{
    HIP_TRY(hipMemcpyAsync(m->xin, x, (size_t)n * m->sizes[0] * sizeof(float), hipMemcpyHostToDevice, m->stream));
    pack_input<<<blocks((size_t)n * m->sizes[0]), 256, 0, m->stream>>>(m->xin, n, m->sizes[0], m->h16[0],
                                                                      pad128(m->sizes[0]));
    HIP_TRY(hipGetLastError());
    return POE2D_OK;
}

poe2d_status check_member(poe2d_mlp *m, int member)
{
    if (member < 0 || member >= (int)m->members.size()) return fail(POE2D_E_INVALID, "network: no such member");
    return POE2D_OK;
}

}  // namespace

// One training step of member e on lane (m itself or its twin: buffers and stream) over the
// dataset rows idx[0 .. batch) of `data` (the handle that holds the dataset): the batch's inputs
// gathered and packed, the targets read in place through idx.
static poe2d_status dataset_step(poe2d_mlp *lane, const poe2d_mlp *data, Member &e, int batch, const int *idx, float lr,
                                 float *loss_dev)
{
    const int in = lane->sizes[0], out = lane->sizes[lane->L];
    gather_rows<<<blocks((size_t)batch * in), 256, 0, lane->stream>>>(data->data_x, in, idx, batch, in, lane->xin, in);
    pack_input<<<blocks((size_t)batch * in), 256, 0, lane->stream>>>(lane->xin, batch, in, lane->h16[0], pad128(in));
    HIP_TRY(hipGetLastError());
    return train_step(lane, e, batch, data->data_t, out, idx, lr, loss_dev);
}

// Activations, gradients and partials of a network of m's sizes, and with `predict` the output
// and target rows and the prediction accumulators.
static poe2d_status alloc_scratch(poe2d_mlp *m, bool predict)
{
    const int B = m->max_batch, nlayers = m->L;
    const int *sizes = m->sizes.data();
    poe2d_status st;
    m->h16.assign(nlayers + 1, nullptr);
    m->h16t.assign(nlayers + 1, nullptr);
    m->pre.assign(nlayers, nullptr);
    m->g16.assign(nlayers, nullptr);
    m->g16t.assign(nlayers, nullptr);
    int hidden_max = 1;
    for (int li = 0; li <= nlayers; ++li) {
        const int w = pad128(sizes[li]);
        if ((st = dmalloc(&m->h16[li], (size_t)B * w)) || (st = dmalloc(&m->h16t[li], (size_t)B * w))) return st;
        if (li < nlayers) {
            const int wo = pad128(sizes[li + 1]);
            if ((st = dmalloc(&m->pre[li], (size_t)B * wo)) || (st = dmalloc(&m->g16[li], (size_t)B * wo)) ||
                (st = dmalloc(&m->g16t[li], (size_t)B * wo)))
                return st;
        }
        if (li > 0 && li < nlayers) hidden_max = std::max(hidden_max, w);
    }
    const int outp = pad128(sizes[nlayers]);
    // split-K partials of the input-gradient products, and the output bias gradient's per-tile sums
    m->work_floats = std::max((size_t)128 * B * hidden_max, (size_t)((B + 127) / 128) * pad128(sizes[nlayers]));
    m->nred = std::max({m->nred, B, mse_blocks(sizes[nlayers]), gemm_tiles(B, sizes[nlayers])});
    if ((st = dmalloc(&m->dh, (size_t)B * std::max(hidden_max, pad128(sizes[0])))) ||
        (st = dmalloc(&m->xin, (size_t)B * sizes[0])) || (st = dmalloc(&m->work, m->work_floats)) ||
        (st = dmalloc(&m->partial, m->nred)) || (st = dmalloc(&m->idx, (size_t)kIndexSteps * B)) ||
        (st = dmalloc(&m->dloss, kIndexSteps)))
        return st;
    if (predict &&
        ((st = dmalloc(&m->y, (size_t)B * outp)) || (st = dmalloc(&m->tgt, (size_t)B * outp)) ||
         (st = dmalloc(&m->acc_mean, (size_t)B * sizes[nlayers])) || (st = dmalloc(&m->acc_m2, (size_t)B * sizes[nlayers]))))
        return st;
    return POE2D_OK;
}

// The second set of training buffers and stream of poe2d_mlp_train_many, made on first use.
static poe2d_status ensure_twin(poe2d_mlp *m)
{
    if (m->twin) return POE2D_OK;
    poe2d_mlp *t = new (std::nothrow) poe2d_mlp;
    if (!t) return fail(POE2D_E_NOMEM, "mlp_train_many: out of host memory");
    t->device = m->device;
    t->sizes = m->sizes;
    t->L = m->L;
    t->max_batch = m->max_batch;
    poe2d_status st = POE2D_OK;
    if (hipStreamCreateWithFlags(&t->stream, hipStreamNonBlocking) != hipSuccess)
        st = fail(POE2D_E_HIP, "mlp_train_many: no stream");
    if (!st) st = alloc_scratch(t, false);
    if (st) {
        free_mlp(t);
        delete t;
        return st;
    }
    m->twin = t;
    return POE2D_OK;
}

extern "C" {

poe2d_status poe2d_mlp_create(int nlayers, const int *sizes, int members, int max_batch, uint64_t seed, poe2d_mlp **out)
{
    if (!out) return fail(POE2D_E_INVALID, "mlp_create: null output pointer");
    *out = nullptr;
    if (nlayers < 1 || !sizes || members < 1 || max_batch < 32 || max_batch % 32)
        return fail(POE2D_E_INVALID, "mlp_create: need >= 1 layer, >= 1 member and a batch that is a multiple of 32");
    for (int i = 0; i <= nlayers; ++i)
        if (sizes[i] < 1) return fail(POE2D_E_INVALID, "mlp_create: layer widths must be >= 1");
    poe2d_mlp *m = new (std::nothrow) poe2d_mlp;
    if (!m) return fail(POE2D_E_NOMEM, "mlp_create: out of host memory");
    auto bail = [&](poe2d_status st) {
        free_mlp(m);
        delete m;
        return st;
    };
    if (hipGetDevice(&m->device) != hipSuccess || hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking) != hipSuccess)
        return bail(fail(POE2D_E_HIP, "mlp_create: no device / stream"));
    m->sizes.assign(sizes, sizes + nlayers + 1);
    m->L = nlayers;
    m->max_batch = max_batch;
    poe2d_status st;
    uint64_t rng = seed;
    m->members.resize(members);
    for (Member &e : m->members) {
        e.layers.resize(nlayers);
        for (int li = 0; li < nlayers; ++li) {
            Layer &l = e.layers[li];
            l.in = sizes[li];
            l.out = sizes[li + 1];
            l.inp = pad128(l.in);
            l.outp = pad128(l.out);
            const size_t nw = (size_t)l.outp * l.inp;
            if ((st = dmalloc(&l.w, nw)) || (st = dmalloc(&l.b, l.outp)) || (st = dmalloc(&l.mw, nw)) ||
                (st = dmalloc(&l.vw, nw)) || (st = dmalloc(&l.mb, l.outp)) || (st = dmalloc(&l.vb, l.outp)) ||
                (st = dmalloc(&l.gw, nw)) || (st = dmalloc(&l.gb, l.outp)) || (st = dmalloc(&l.w16, nw)) ||
                (st = dmalloc(&l.w16t, nw)))
                return bail(st);
            // uniform in +-sqrt(6 / (in + out)) on the valid block, zero padding, zero bias
            std::vector<float> h(nw, 0.f);
            const double a = std::sqrt(6.0 / (l.in + l.out));
            for (int o = 0; o < l.out; ++o)
                for (int i = 0; i < l.in; ++i)
                    h[(size_t)o * l.inp + i] = (float)(a * (2.0 * ((splitmix(rng) >> 11) * 0x1.0p-53) - 1.0));
            HIP_TRY(hipMemcpy(l.w, h.data(), nw * sizeof(float), hipMemcpyHostToDevice));
            if ((st = sync_copies(m, l))) return bail(st);
        }
    }
    if ((st = alloc_scratch(m, true))) return bail(st);
    if (hipStreamSynchronize(m->stream) != hipSuccess) return bail(fail(POE2D_E_HIP, "mlp_create: initialisation failed"));
    *out = m;
    return POE2D_OK;
}

poe2d_status poe2d_mlp_get_params(poe2d_mlp *m, int member, int layer, float *w, float *b)
{
    poe2d_status st = check_mlp(m);
    if (st || (st = check_member(m, member))) return st;
    if (layer < 0 || layer >= m->L) return fail(POE2D_E_INVALID, "mlp_get_params: no such layer");
    Layer &l = m->members[member].layers[layer];
    HIP_TRY(hipStreamSynchronize(m->stream));
    if (w) HIP_TRY(hipMemcpy2D(w, l.in * sizeof(float), l.w, l.inp * sizeof(float), l.in * sizeof(float), l.out,
                               hipMemcpyDeviceToHost));
    if (b) HIP_TRY(hipMemcpy(b, l.b, l.out * sizeof(float), hipMemcpyDeviceToHost));
    return POE2D_OK;
}

poe2d_status poe2d_mlp_get_bf16(poe2d_mlp *m, int member, int layer, uint16_t *w16, uint16_t *w16t)
{
    poe2d_status st = check_mlp(m);
    if (st || (st = check_member(m, member))) return st;
    if (layer < 0 || layer >= m->L) return fail(POE2D_E_INVALID, "mlp_get_bf16: no such layer");
    Layer &l = m->members[member].layers[layer];
    HIP_TRY(hipStreamSynchronize(m->stream));
    if (w16) HIP_TRY(hipMemcpy2D(w16, l.in * sizeof(uint16_t), l.w16, l.inp * sizeof(uint16_t), l.in * sizeof(uint16_t),
                                 l.out, hipMemcpyDeviceToHost));
    if (w16t) HIP_TRY(hipMemcpy2D(w16t, l.out * sizeof(uint16_t), l.w16t, l.outp * sizeof(uint16_t),
                                  l.out * sizeof(uint16_t), l.in, hipMemcpyDeviceToHost));
    return POE2D_OK;
}

poe2d_status poe2d_mlp_set_params(poe2d_mlp *m, int member, int layer, const float *w, const float *b)
{
    poe2d_status st = check_mlp(m);
    if (st || (st = check_member(m, member))) return st;
    if (layer < 0 || layer >= m->L) return fail(POE2D_E_INVALID, "mlp_set_params: no such layer");
    Layer &l = m->members[member].layers[layer];
    HIP_TRY(hipStreamSynchronize(m->stream));
    if (w) HIP_TRY(hipMemcpy2D(l.w, l.inp * sizeof(float), w, l.in * sizeof(float), l.in * sizeof(float), l.out,
                               hipMemcpyHostToDevice));
    if (b) HIP_TRY(hipMemcpy(l.b, b, l.out * sizeof(float), hipMemcpyHostToDevice));
    return sync_copies(m, l);
}

poe2d_status poe2d_mlp_forward(poe2d_mlp *m, int member, int n, const float *x, float *y)
{
    poe2d_status st = check_mlp(m);
    if (st || (st = check_member(m, member))) return st;
    if (n < 1 || n > m->max_batch || !x || !y) return fail(POE2D_E_INVALID, "mlp_forward: bad batch");
    if ((st = load_input(m, x, n)) || (st = forward(m, m->members[member], n))) return st;
    const int out = m->sizes[m->L], outp = pad128(out);
    HIP_TRY(hipMemcpy2DAsync(y, out * sizeof(float), m->y, outp * sizeof(float), out * sizeof(float), n,
                             hipMemcpyDeviceToHost, m->stream));
    HIP_TRY(hipStreamSynchronize(m->stream));
    return POE2D_OK;
}

poe2d_status poe2d_mlp_gradients(poe2d_mlp *m, int member, int n, const float *x, const float *t, float *loss,
                                 float *const *gw, float *const *gb)
{
    poe2d_status st = check_mlp(m);
    if (st || (st = check_member(m, member))) return st;
    if (n < 32 || n > m->max_batch || n % 32 || !x || !t) return fail(POE2D_E_INVALID, "mlp_gradients: bad batch");
    const int out = m->sizes[m->L], outp = pad128(out);
    Member &e = m->members[member];
    if ((st = load_input(m, x, n))) return st;
    HIP_TRY(hipMemcpy2DAsync(m->tgt, outp * sizeof(float), t, out * sizeof(float), out * sizeof(float), n,
                             hipMemcpyHostToDevice, m->stream));
    if ((st = forward(m, e, n)) || (st = backward(m, e, n, loss))) return st;
    for (int li = 0; li < m->L; ++li) {
        Layer &l = e.layers[li];
        if (gw && gw[li])
            HIP_TRY(hipMemcpy2DAsync(gw[li], l.in * sizeof(float), l.gw, l.inp * sizeof(float), l.in * sizeof(float),
                                     l.out, hipMemcpyDeviceToHost, m->stream));
        if (gb && gb[li])
            HIP_TRY(hipMemcpyAsync(gb[li], l.gb, l.out * sizeof(float), hipMemcpyDeviceToHost, m->stream));
    }
    HIP_TRY(hipStreamSynchronize(m->stream));
    return POE2D_OK;
}

poe2d_status poe2d_mlp_set_dataset(poe2d_mlp *m, int n, const float *x, const float *t)
{
    poe2d_status st = check_mlp(m);
    if (st) return st;
    if (n < 1 || !x || !t) return fail(POE2D_E_INVALID, "mlp_set_dataset: bad data");
    HIP_TRY(hipStreamSynchronize(m->stream));
    dfree(m->data_x);
    dfree(m->data_t);
    m->data_n = 0;
    const int in = m->sizes[0], out = m->sizes[m->L];
    if ((st = dmalloc(&m->data_x, (size_t)n * in)) || (st = dmalloc(&m->data_t, (size_t)n * out))) return st;
    HIP_TRY(hipMemcpy(m->data_x, x, (size_t)n * in * sizeof(float), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(m->data_t, t, (size_t)n * out * sizeof(float), hipMemcpyHostToDevice));
    m->data_n = n;
    return POE2D_OK;
}

poe2d_status poe2d_mlp_train(poe2d_mlp *m, int member, int nsteps, int batch, const int *indices, float lr, float *losses)
{
    poe2d_status st = check_mlp(m);
    if (st || (st = check_member(m, member))) return st;
    if (!m->data_n) return fail(POE2D_E_STATE, "mlp_train: no dataset (call poe2d_mlp_set_dataset)");
    if (batch < 32 || batch > m->max_batch || batch % 32 || nsteps < 0 || (nsteps > 0 && !indices) || !(lr > 0))
        return fail(POE2D_E_INVALID, "mlp_train: bad arguments");
    for (long long i = 0; i < (long long)nsteps * batch; ++i)
        if (indices[i] < 0 || indices[i] >= m->data_n) return fail(POE2D_E_INVALID, "mlp_train: index outside the dataset");
    Member &e = m->members[member];
    for (int s = 0; s < nsteps; ++s) {
        const int c = s % kIndexSteps;
        if (c == 0)
            HIP_TRY(hipMemcpyAsync(m->idx, indices + (size_t)s * batch,
                                   (size_t)std::min(kIndexSteps, nsteps - s) * batch * sizeof(int),
                                   hipMemcpyHostToDevice, m->stream));
        if ((st = dataset_step(m, m, e, batch, m->idx + (size_t)c * batch, lr, losses ? m->dloss + c : nullptr)))
            return st;
        if (losses && (c == kIndexSteps - 1 || s == nsteps - 1))
            HIP_TRY(hipMemcpyAsync(losses + (s - c), m->dloss, (size_t)(c + 1) * sizeof(float), hipMemcpyDeviceToHost,
                                   m->stream));
    }
    HIP_TRY(hipStreamSynchronize(m->stream));
    return POE2D_OK;
}

poe2d_status poe2d_mlp_train_many(poe2d_mlp *m, int count, const int *members, int nsteps, int batch,
                                  const int *indices, float lr, float *losses)
{
    poe2d_status st = check_mlp(m);
    if (st) return st;
    if (!m->data_n) return fail(POE2D_E_STATE, "mlp_train_many: no dataset (call poe2d_mlp_set_dataset)");
    if (count < 1 || !members || batch < 32 || batch > m->max_batch || batch % 32 || nsteps < 0 ||
        (nsteps > 0 && !indices) || !(lr > 0))
        return fail(POE2D_E_INVALID, "mlp_train_many: bad arguments");
    for (int j = 0; j < count; ++j) {
        if ((st = check_member(m, members[j]))) return st;
        for (int q = 0; q < j; ++q)
            if (members[q] == members[j]) return fail(POE2D_E_INVALID, "mlp_train_many: a member is listed twice");
    }
    const size_t ni = (size_t)count * nsteps * batch;
    for (size_t i = 0; i < ni; ++i)
        if (indices[i] < 0 || indices[i] >= m->data_n)
            return fail(POE2D_E_INVALID, "mlp_train_many: index outside the dataset");
    if (nsteps == 0) return POE2D_OK;
    if (count > 1 && (st = ensure_twin(m))) return st;
    if (ni > m->idx_all_cap) {
        dfree(m->idx_all);
        m->idx_all_cap = 0;
        if ((st = dmalloc(&m->idx_all, ni))) return st;
        m->idx_all_cap = ni;
    }
    const size_t nl = (size_t)count * nsteps;
    if (nl > m->loss_all_cap) {
        dfree(m->loss_all);
        m->loss_all_cap = 0;
        if ((st = dmalloc(&m->loss_all, nl))) return st;
        m->loss_all_cap = nl;
    }
    while ((int)m->step_done.size() < count) {
        hipEvent_t ev;
        HIP_TRY(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
        m->step_done.push_back(ev);
    }
    HIP_TRY(hipMemcpyAsync(m->idx_all, indices, ni * sizeof(int), hipMemcpyHostToDevice, m->stream));
    // (everything earlier on the handle's stream, these indices included, is done before the
    // second stream starts)
    HIP_TRY(hipStreamSynchronize(m->stream));
    // Member-steps k = s * count + j in that order, on alternating streams: with an odd count a
    // member's consecutive steps alternate too, and each waits for its previous one.
    for (int s = 0; s < nsteps; ++s)
        for (int j = 0; j < count; ++j) {
            const long long k = (long long)s * count + j;
            poe2d_mlp *lane = (count == 1 || k % 2 == 0) ? m : m->twin;
            if (count > 1 && count % 2 == 1 && s > 0) HIP_TRY(hipStreamWaitEvent(lane->stream, m->step_done[j], 0));
            if ((st = dataset_step(lane, m, m->members[members[j]], batch, m->idx_all + ((size_t)j * nsteps + s) * batch,
                                   lr, losses ? m->loss_all + (size_t)j * nsteps + s : nullptr)))
                return st;
            if (count > 1) HIP_TRY(hipEventRecord(m->step_done[j], lane->stream));
        }
    if (m->twin) HIP_TRY(hipStreamSynchronize(m->twin->stream));
    if (losses) HIP_TRY(hipMemcpyAsync(losses, m->loss_all, nl * sizeof(float), hipMemcpyDeviceToHost, m->stream));
    HIP_TRY(hipStreamSynchronize(m->stream));
    return POE2D_OK;
}

poe2d_status poe2d_mlp_predict(poe2d_mlp *m, int n, const float *x, float *mean, float *spread)
{
    poe2d_status st = check_mlp(m);
    if (st) return st;
    if (n < 1 || !x) return fail(POE2D_E_INVALID, "mlp_predict: bad input");
    const int in = m->sizes[0], out = m->sizes[m->L], outp = pad128(out);
    const int E = (int)m->members.size();
    for (int r0 = 0; r0 < n; r0 += m->max_batch) {
        const int nb = std::min(m->max_batch, n - r0);
        if ((st = load_input(m, x + (size_t)r0 * in, nb))) return st;
        for (int k = 0; k < E; ++k) {
            if ((st = forward(m, m->members[k], nb))) return st;
            welford<<<blocks((size_t)nb * out), 256, 0, m->stream>>>(m->y, outp, nb, out, k, m->acc_mean, m->acc_m2);
            HIP_TRY(hipGetLastError());
        }
        if (spread) {
            row_spread<<<nb, 256, 0, m->stream>>>(m->acc_mean, m->acc_m2, out, E, m->partial);
            HIP_TRY(hipGetLastError());
            HIP_TRY(hipMemcpyAsync(spread + r0, m->partial, nb * sizeof(float), hipMemcpyDefault, m->stream));
        }
        if (mean) {
            // (the spread above is of the unscaled outputs)
            if (m->oscale) {
                scale_cols<<<blocks((size_t)nb * out), 256, 0, m->stream>>>(m->acc_mean, nb, out, m->oscale);
                HIP_TRY(hipGetLastError());
            }
            // (host or device memory: the direction comes from the pointers)
            HIP_TRY(hipMemcpyAsync(mean + (size_t)r0 * out, m->acc_mean, (size_t)nb * out * sizeof(float),
                                   hipMemcpyDefault, m->stream));
        }
        HIP_TRY(hipStreamSynchronize(m->stream));
    }
    return POE2D_OK;
}

poe2d_status poe2d_mlp_set_output_scale(poe2d_mlp *m, const float *scale)
{
    poe2d_status st = check_mlp(m);
    if (st) return st;
    const int out = m->sizes[m->L];
    if (!scale) {
        dfree(m->oscale);
        m->oscale = nullptr;
        return POE2D_OK;
    }
    if (!m->oscale && (st = dmalloc(&m->oscale, out))) return st;
    HIP_TRY(hipMemcpyAsync(m->oscale, scale, out * sizeof(float), hipMemcpyHostToDevice, m->stream));
    HIP_TRY(hipStreamSynchronize(m->stream));
    return POE2D_OK;
}

poe2d_status poe2d_mlp_time(poe2d_mlp *m, int batch, int iters, double *forward_ms, double *train_ms)
{
    poe2d_status st = check_mlp(m);
    if (st) return st;
    if (batch < 32 || batch > m->max_batch || batch % 32 || iters < 1 || !forward_ms || !train_ms)
        return fail(POE2D_E_INVALID, "mlp_time: bad arguments");
    Member &e = m->members[0];
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    // inputs and targets as left in the buffers (zeros at first): timing only
    if ((st = forward(m, e, batch))) return st;
    HIP_TRY(hipEventRecord(e0, m->stream));
    for (int i = 0; i < iters && !st; ++i) st = forward(m, e, batch);
    HIP_TRY(hipEventRecord(e1, m->stream));
    HIP_TRY(hipEventSynchronize(e1));
    float ms = 0;
    HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
    *forward_ms = ms / iters;
    HIP_TRY(hipEventRecord(e0, m->stream));
    for (int i = 0; i < iters && !st; ++i) st = train_step(m, e, batch, m->tgt, pad128(m->sizes[m->L]), nullptr, 1e-12f, nullptr);
    HIP_TRY(hipEventRecord(e1, m->stream));
    HIP_TRY(hipEventSynchronize(e1));
    HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
    *train_ms = ms / iters;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return st;
}

poe2d_status poe2d_gemm_bf16(int M, int N, int K, const float *A, const float *B, float *C, int ksplit)
{
    if (M < 1 || N < 1 || K < 1 || !A || !B || !C || ksplit < 1) return fail(POE2D_E_INVALID, "gemm_bf16: bad arguments");
    const int Kp = (K + 31) / 32 * 32;
    std::vector<uint16_t> a((size_t)M * Kp, 0), b((size_t)N * Kp, 0);
    for (int i = 0; i < M; ++i)
        for (int k = 0; k < K; ++k) a[(size_t)i * Kp + k] = bf16_rne(A[(size_t)i * K + k]);
    for (int j = 0; j < N; ++j)
        for (int k = 0; k < K; ++k) b[(size_t)j * Kp + k] = bf16_rne(B[(size_t)j * K + k]);
    uint16_t *da = nullptr, *db = nullptr;
    float *dc = nullptr, *work = nullptr;
    const size_t wf = (size_t)ksplit * M * N;
    poe2d_status st = POE2D_OK;
    if ((st = dmalloc(&da, a.size())) || (st = dmalloc(&db, b.size())) || (st = dmalloc(&dc, (size_t)M * N)) ||
        (st = dmalloc(&work, ksplit > 1 ? wf : 1))) {
        dfree(da); dfree(db); dfree(dc); dfree(work);
        return st;
    }
    hipError_t e = hipMemcpy(da, a.data(), a.size() * 2, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(db, b.data(), b.size() * 2, hipMemcpyHostToDevice);
    GemmOut o;
    o.f32 = dc;
    o.ldf = N;
    if (e == hipSuccess) e = gemm_bf16(da, Kp, db, Kp, M, N, Kp, EPI_LINEAR, o, work, wf, ksplit, nullptr);
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e == hipSuccess) e = hipMemcpy(C, dc, (size_t)M * N * sizeof(float), hipMemcpyDeviceToHost);
    dfree(da); dfree(db); dfree(dc); dfree(work);
    if (e != hipSuccess) return fail(POE2D_E_HIP, std::string("gemm_bf16: ") + hipGetErrorString(e));
    return POE2D_OK;
}

poe2d_status poe2d_gemm_bf16_noalloc(int M, int N, int K, const float *A, const float *B, float *C, int ksplit)
{
    if (M < 1 || N < 1 || K < 1 || !A || !B || !C || ksplit < 1)
        return fail(POE2D_E_INVALID, "gemm_bf16_noalloc: bad arguments");
    const size_t Kp = ((size_t)K + 31) / 32 * 32;
    auto up = [](size_t b) { return (b + 255) / 256 * 256; };
    // scratch regions: A and B as FP32, A and B as bf16 (K padded to Kp), C, split-K partials
    const size_t sa32 = up((size_t)M * K * 4), sb32 = up((size_t)N * K * 4);
    const size_t sa16 = up((size_t)M * Kp * 2), sb16 = up((size_t)N * Kp * 2), sc = up((size_t)M * N * 4);
    const size_t wf = (size_t)ksplit * M * N, sw = ksplit > 1 ? up(wf * 4) : 0;
    const size_t need = sa32 + sb32 + sa16 + sb16 + sc + sw;
    if (need > kGemmScratchBytes)
        return fail(POE2D_E_INVALID, "gemm_bf16_noalloc: these operands need " + std::to_string(need) +
                                         " bytes of device scratch, more than its " +
                                         std::to_string(kGemmScratchBytes) + " (use poe2d_gemm_bf16)");
    unsigned char *base = nullptr;
    poe2d_status st = scratch_address(&base);
    if (st) return st;
    float *a32 = (float *)base, *b32 = (float *)(base + sa32);
    uint16_t *a16 = (uint16_t *)(base + sa32 + sb32), *b16 = (uint16_t *)(base + sa32 + sb32 + sa16);
    float *c = (float *)(base + sa32 + sb32 + sa16 + sb16);
    float *work = sw ? (float *)(base + sa32 + sb32 + sa16 + sb16 + sc) : nullptr;
    // Nothing here has to start at zero: the packing kernel writes every bf16 element (zeros in the
    // padding), and the product writes every element of C and of each split's partial sums.
    HIP_TRY(hipMemcpy(a32, A, (size_t)M * K * 4, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(b32, B, (size_t)N * K * 4, hipMemcpyHostToDevice));
    pack_bf16_rows<<<blocks((size_t)M * Kp), 256, 0, nullptr>>>(a32, M, K, (int)Kp, a16);
    pack_bf16_rows<<<blocks((size_t)N * Kp), 256, 0, nullptr>>>(b32, N, K, (int)Kp, b16);
    HIP_TRY(hipGetLastError());
    GemmOut o;
    o.f32 = c;
    o.ldf = N;
    HIP_TRY(gemm_bf16(a16, (int)Kp, b16, (int)Kp, M, N, (int)Kp, EPI_LINEAR, o, work, sw ? wf : 0, sw ? ksplit : 1,
                      nullptr));
    HIP_TRY(hipDeviceSynchronize());
    HIP_TRY(hipMemcpy(C, c, (size_t)M * N * sizeof(float), hipMemcpyDeviceToHost));
    return POE2D_OK;
}

void poe2d_mlp_destroy(poe2d_mlp *m)
{
    if (!m) return;
    (void)hipSetDevice(m->device);
    if (m->stream) (void)hipStreamSynchronize(m->stream);
    free_mlp(m);
    delete m;
}

poe2d_status poe2d_host_alloc(size_t bytes, void **out)
{
    if (!out || bytes == 0) return fail(POE2D_E_INVALID, "host_alloc: bad arguments");
    *out = nullptr;
    if (hipHostMalloc(out, bytes, hipHostMallocDefault) != hipSuccess) {
        *out = nullptr;
        return fail(POE2D_E_NOMEM, "host_alloc: hipHostMalloc failed");
    }
    return POE2D_OK;
}

void poe2d_host_free(void *p)
{
    if (p) (void)hipHostFree(p);
}

poe2d_status poe2d_device_alloc(size_t bytes, void **out)
{
    if (!out || bytes == 0) return fail(POE2D_E_INVALID, "device_alloc: bad arguments");
    *out = nullptr;
    if (hipMalloc(out, bytes) != hipSuccess) {
        *out = nullptr;
        return fail(POE2D_E_NOMEM, "device_alloc: hipMalloc failed");
    }
    return POE2D_OK;
}

void poe2d_device_free(void *p)
{
    if (p) (void)hipFree(p);
}

poe2d_status poe2d_copy(void *dst, const void *src, size_t bytes)
{
    if (bytes == 0) return POE2D_OK;
    if (!dst || !src) return fail(POE2D_E_INVALID, "copy: null pointer");
    HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyDefault));
    return POE2D_OK;
}

}  // extern "C"
