// Launchers of the auxiliary kernels (receiver gather, derived fields, component extraction,
// stability reduction).
#pragma once

#include "step_common.h"

namespace poe2d {

struct Receivers {
    int n;
    unsigned mask;
    const int *ix, *iz;
};

// Gather of one slot: velocities from state buffer `vbuf` (the state before the step, v^n) and
// pore pressure recomputed from the strains / fluid content of buffer `sbuf` (after the step's
// stress update) plus the pressure source of the step.  out = [ncomp][nrec].
template <class T>
void launch_gather(const StepArgs<T> &A, int vbuf, int sbuf, const Receivers &R, T *out, hipStream_t stream);

// p, sxx or syy (what = 0, 1, 2) of buffer `buf`'s state into a dense (nz+1) x (nx+1) array.
template <class T>
void launch_derived(const StepArgs<T> &A, int buf, int what, T *out, hipStream_t stream);

// Component `comp` of group `g` into a dense (nz+1) x (nx+1) array.
template <class T>
void launch_extract(const StepArgs<T> &A, int g, int comp, T *out, hipStream_t stream);

// A dense (nz+1) x (nx+1) array into component `comp` of group `g` (pads untouched).
template <class T>
void launch_insert(const StepArgs<T> &A, int g, int comp, const T *in, hipStream_t stream);

// Velocity-update constants (biot_velocity_cn) of the npal MV entries of an FP32 palette into its
// PAL_CN section.
void launch_cn_palette(Vec4<float> *pal, int npal, float inv_dt, hipStream_t stream);

// Largest |vx|, |vy| of buffer `buf` (non-finite counts as +inf) into result[0].
template <class T>
void launch_absmax(const StepArgs<T> &A, int buf, float *partials, int nblocks, float *result,
                   hipStream_t stream);

}  // namespace poe2d
