// Launchers of the auxiliary kernels (receiver gather, derived fields, stability reduction).
#pragma once

#include "step_common.h"

namespace poe2d {

template <class T>
struct GatherArgs {
    int nx, nz, pitch, nrec;
    unsigned mask;
    const int *rix, *riz;
    const T *vx, *vz;            // velocities at the start of the step
    const T *exx, *eyy, *zeta;   // strains / fluid content after the step's stress update
    const T *al, *M;
    int src_kind, src_x, src_z;
    T src_val;
    T *out;                      // [ncomp][nrec] of this slot
};

template <class T>
void launch_gather(const GatherArgs<T> &G, hipStream_t stream);

template <class T>
void launch_derived(const GatherArgs<T> &G, const T *lc, const T *mu, int what, T *out, hipStream_t stream);

template <class T>
void launch_absmax(const T *a, const T *b, size_t n, float *partials, int nblocks, float *result,
                   hipStream_t stream);

}  // namespace poe2d
