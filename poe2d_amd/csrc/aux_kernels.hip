// Small kernels around the time step: receiver gather, derived-field recompute for snapshots,
// and the stability reduction.
#include "aux_kernels.h"

namespace poe2d {

// Gather of step n's slot: velocities from the state before the step (v^n) and pore pressure
// recomputed from the strains / fluid content the step produced (p^(n+1/2)), plus the pressure
// source when it sits on a receiver.
template <class T>
__global__ void gather_kernel(GatherArgs<T> G)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= G.nrec) return;
    const int x = G.rix[r], z = G.riz[r];
    const ptrdiff_t o = (ptrdiff_t)z * G.pitch + x;
    int c = 0;
    if (G.mask & POE2D_VX) G.out[(size_t)(c++) * G.nrec + r] = G.vx[o];
    if (G.mask & POE2D_VY) G.out[(size_t)(c++) * G.nrec + r] = G.vz[o];
    if (G.mask & POE2D_P) {
        T p = 0;
        if (x >= 1 && x < G.nx && z >= 2 && z < G.nz - 1) {
            p = pore_pressure(G.exx[o], G.eyy[o], G.zeta[o], G.al[o], G.M[o]);
            if (G.src_kind == POE2D_SRC_PRESSURE_DGAUSS && x == G.src_x && z == G.src_z) p += G.src_val;
        }
        G.out[(size_t)c * G.nrec + r] = p;
    }
}

// p, sxx or syy of the last stress update, recomputed from the state (what = 0, 1, 2), into a
// dense (nz+1) x (nx+1) array.
template <class T>
__global__ void derived_kernel(GatherArgs<T> G, const T *lc, const T *mu, int what, T *out)
{
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int z = blockIdx.y;
    if (x > G.nx || z > G.nz) return;
    const ptrdiff_t o = (ptrdiff_t)z * G.pitch + x;
    T v = 0;
    if (x >= 1 && x < G.nx && z >= 2 && z < G.nz - 1) {
        const T al = G.al[o];
        T p = pore_pressure(G.exx[o], G.eyy[o], G.zeta[o], al, G.M[o]);
        if (G.src_kind == POE2D_SRC_PRESSURE_DGAUSS && x == G.src_x && z == G.src_z) p += G.src_val;
        if (what == 0) {
            v = p;
        } else if (x < G.nx - 1) {
            const T l = lc[o], l2m = l + T(2) * mu[o];
            v = what == 1 ? l2m * G.exx[o] + l * G.eyy[o] - al * p : l * G.exx[o] + l2m * G.eyy[o] - al * p;
        }
    }
    out[(size_t)z * (G.nx + 1) + x] = v;
}

// Largest |v| per block (non-finite values count as +inf), then one block reduces the partials.
template <class T>
__global__ void __launch_bounds__(256) absmax_blocks(const T *a, const T *b, size_t n, float *out)
{
    __shared__ float red[256];
    float m = 0.f;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        float va = fabsf((float)a[i]), vb = fabsf((float)b[i]);
        if (!(va <= 3.0e38f)) va = INFINITY;
        if (!(vb <= 3.0e38f)) vb = INFINITY;
        m = fmaxf(m, fmaxf(va, vb));
    }
    red[threadIdx.x] = m;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] = fmaxf(red[threadIdx.x], red[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) out[blockIdx.x] = red[0];
}

__global__ void __launch_bounds__(256) absmax_final(const float *in, int n, float *out)
{
    __shared__ float red[256];
    float m = 0.f;
    for (int i = threadIdx.x; i < n; i += blockDim.x) m = fmaxf(m, in[i]);
    red[threadIdx.x] = m;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] = fmaxf(red[threadIdx.x], red[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = red[0];
}

template <class T>
void launch_gather(const GatherArgs<T> &G, hipStream_t stream)
{
    gather_kernel<T><<<(G.nrec + 255) / 256, 256, 0, stream>>>(G);
}

template <class T>
void launch_derived(const GatherArgs<T> &G, const T *lc, const T *mu, int what, T *out, hipStream_t stream)
{
    derived_kernel<T><<<dim3((G.nx + 1 + 255) / 256, G.nz + 1), 256, 0, stream>>>(G, lc, mu, what, out);
}

template <class T>
void launch_absmax(const T *a, const T *b, size_t n, float *partials, int nblocks, float *result, hipStream_t stream)
{
    absmax_blocks<T><<<nblocks, 256, 0, stream>>>(a, b, n, partials);
    absmax_final<<<1, 256, 0, stream>>>(partials, nblocks, result);
}

template void launch_gather<float>(const GatherArgs<float> &, hipStream_t);
template void launch_gather<double>(const GatherArgs<double> &, hipStream_t);
template void launch_derived<float>(const GatherArgs<float> &, const float *, const float *, int, float *, hipStream_t);
template void launch_derived<double>(const GatherArgs<double> &, const double *, const double *, int, double *, hipStream_t);
template void launch_absmax<float>(const float *, const float *, size_t, float *, int, float *, hipStream_t);
template void launch_absmax<double>(const double *, const double *, size_t, float *, int, float *, hipStream_t);

}  // namespace poe2d
