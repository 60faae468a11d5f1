// Small kernels around the time step: the receiver gather of the two-pass schedule, derived-field
// recompute and component extraction for snapshots, and the stability reduction.
#include "aux_kernels.h"

namespace poe2d {

namespace {

template <class T>
__device__ T pressure_at(const StepArgs<T> &A, const Vec4<T> *S, const Vec4<T> *MS, int x, int z, long long o)
{
    if (!(x >= 1 && x < A.nx && z >= 2 && z < A.nz - 1)) return T(0);
    const Vec4<T> s = S[o], m = MS[o];
    T p = pore_pressure(s.x, s.y, s.z, m.x, m.y);
    if (A.src_kind == POE2D_SRC_PRESSURE_DGAUSS && x == A.src_x && z == A.src_z) p += A.src_val;
    return p;
}

template <class T>
__global__ void gather_kernel(StepArgs<T> A, int vbuf, int sbuf, Receivers R, T *out)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= R.n) return;
    const int x = R.ix[r], z = R.iz[r];
    const long long o = (long long)z * A.pitch + x;
    const Vec4<T> v = state_group(A, group_v(vbuf))[o];
    int c = 0;
    if (R.mask & POE2D_VX) out[(size_t)(c++) * R.n + r] = v.x;
    if (R.mask & POE2D_VY) out[(size_t)(c++) * R.n + r] = v.y;
    if (R.mask & POE2D_P) out[(size_t)c * R.n + r] = pressure_at(A, state_group(A, group_s(sbuf)), mat_group(A, G_MS), x, z, o);
}

template <class T>
__global__ void derived_kernel(StepArgs<T> A, int buf, int what, T *out)
{
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int z = blockIdx.y;
    if (x > A.nx || z > A.nz) return;
    const long long o = (long long)z * A.pitch + x;
    const Vec4<T> *S = state_group(A, group_s(buf));
    const Vec4<T> *MS = mat_group(A, G_MS);
    const T p = pressure_at(A, S, MS, x, z, o);
    T v = p;
    if (what != 0) {
        v = 0;
        if (x >= 1 && x < A.nx - 1 && z >= 2 && z < A.nz - 1) {
            const Vec4<T> s = S[o], m = MS[o];
            const T l2m = m.z + T(2) * m.w;
            v = what == 1 ? l2m * s.x + m.z * s.y - m.x * p : m.z * s.x + l2m * s.y - m.x * p;
        }
    }
    out[(size_t)z * (A.nx + 1) + x] = v;
}

template <class T>
__global__ void extract_kernel(StepArgs<T> A, int g, int comp, T *out)
{
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int z = blockIdx.y;
    if (x > A.nx || z > A.nz) return;
    const Vec4<T> v = state_group(A, g)[(long long)z * A.pitch + x];
    out[(size_t)z * (A.nx + 1) + x] = comp == 0 ? v.x : comp == 1 ? v.y : comp == 2 ? v.z : v.w;
}

template <class T>
__global__ void insert_kernel(StepArgs<T> A, int g, int comp, const T *in)
{
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int z = blockIdx.y;
    if (x > A.nx || z > A.nz) return;
    Vec4<T> &v = state_group(A, g)[(long long)z * A.pitch + x];
    const T a = in[(size_t)z * (A.nx + 1) + x];
    if (comp == 0) v.x = a;
    else if (comp == 1) v.y = a;
    else if (comp == 2) v.z = a;
    else v.w = a;
}

template <class T>
__global__ void __launch_bounds__(256) absmax_blocks(StepArgs<T> A, int buf, float *partial)
{
    __shared__ float red[256];
    const Vec4<T> *V = state_group(A, group_v(buf));
    const long long npts = (long long)(A.nz + 1) * A.pitch;
    float m = 0.f;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < npts; i += (long long)gridDim.x * blockDim.x) {
        const Vec4<T> v = V[i];  // rows 0..nz, pad columns included (zero)
        float a = fabsf((float)v.x), b = fabsf((float)v.y);
        if (!(a <= 3.0e38f)) a = INFINITY;
        if (!(b <= 3.0e38f)) b = INFINITY;
        m = fmaxf(m, fmaxf(a, b));
    }
    red[threadIdx.x] = m;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] = fmaxf(red[threadIdx.x], red[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
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

}  // namespace

template <class T>
void launch_gather(const StepArgs<T> &A, int vbuf, int sbuf, const Receivers &R, T *out, hipStream_t stream)
{
    gather_kernel<T><<<(R.n + 255) / 256, 256, 0, stream>>>(A, vbuf, sbuf, R, out);
}

template <class T>
void launch_derived(const StepArgs<T> &A, int buf, int what, T *out, hipStream_t stream)
{
    derived_kernel<T><<<dim3((A.nx + 1 + 255) / 256, A.nz + 1), 256, 0, stream>>>(A, buf, what, out);
}

template <class T>
void launch_extract(const StepArgs<T> &A, int g, int comp, T *out, hipStream_t stream)
{
    extract_kernel<T><<<dim3((A.nx + 1 + 255) / 256, A.nz + 1), 256, 0, stream>>>(A, g, comp, out);
}

template <class T>
void launch_insert(const StepArgs<T> &A, int g, int comp, const T *in, hipStream_t stream)
{
    insert_kernel<T><<<dim3((A.nx + 1 + 255) / 256, A.nz + 1), 256, 0, stream>>>(A, g, comp, in);
}

template <class T>
void launch_absmax(const StepArgs<T> &A, int buf, float *partials, int nblocks, float *result, hipStream_t stream)
{
    absmax_blocks<T><<<nblocks, 256, 0, stream>>>(A, buf, partials);
    absmax_final<<<1, 256, 0, stream>>>(partials, nblocks, result);
}

__global__ void __launch_bounds__(256) cn_palette_kernel(Vec4<float> *pal, int npal, float inv_dt)
{
    const int e = (int)threadIdx.x;
    if (e >= npal) return;
    const Vec4<float> mv = pal[PAL_MV * kMaxPalette + e];
    float co, c1, c2;
    cn_coeffs(mv.x, mv.y, mv.z, mv.w, inv_dt, co, c1, c2);
    const float rhofb = mv.y * mv.w;
    pal[PAL_CN * kMaxPalette + e] = Vec4<float>{c2, __builtin_amdgcn_rcpf(c1), rhofb, __builtin_amdgcn_rcpf(co)};
}

void launch_cn_palette(Vec4<float> *pal, int npal, float inv_dt, hipStream_t stream)
{
    static_assert(kMaxPalette <= 256, "one thread per palette entry");
    cn_palette_kernel<<<1, 256, 0, stream>>>(pal, npal, inv_dt);
}

#define POE2D_AUX_INSTANTIATE(T)                                                                      \
    template void launch_gather<T>(const StepArgs<T> &, int, int, const Receivers &, T *, hipStream_t); \
    template void launch_derived<T>(const StepArgs<T> &, int, int, T *, hipStream_t);                 \
    template void launch_extract<T>(const StepArgs<T> &, int, int, T *, hipStream_t);                 \
    template void launch_insert<T>(const StepArgs<T> &, int, int, const T *, hipStream_t);            \
    template void launch_absmax<T>(const StepArgs<T> &, int, float *, int, float *, hipStream_t);
POE2D_AUX_INSTANTIATE(float)
POE2D_AUX_INSTANTIATE(double)
#undef POE2D_AUX_INSTANTIATE

}  // namespace poe2d
