// Solver handle behind the C ABI of include/poe2d.h: device memory, the time loop, gather,
// snapshots, checkpoint/restore and the stability check.
//
// Device layout.  Every field is SoA with row pitch P (a multiple of 64 elements) and zero pad
// rows/columns around the (nz+1) x (nx+1) grid: 4 pad rows above and below, 4 pad columns on the
// left and at least one on the right (the reference's stress1 reads column nx+1 at x = nx-1, which
// in its flat array is column 0 of the next row, always zero; the pad reproduces that).  Pointers
// handed to kernels address point (0, 0).  The wavefield state (vx, vy, wx, wy, e_xx, e_yy, zeta,
// sxy) and the C-PML memory variables are double-buffered: a step reads buffer `cur` and writes
// buffer 1-cur.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "aux_kernels.h"
#include "common.h"
#include "step_common.h"

using namespace poe2d;

namespace {

const float kRefPiF = 3.1415926535897932f;  // the reference's float PI (seismic_CPML_2d_poroelastic.cpp:9)
const double kStabilityThreshold = 1.0e25;  // STABILITY_THRESHOLD of the reference (declared, unused there)
const int kPadRows = 4, kPadCols = 4;

#define HIP_TRY(expr)                                                                              \
    do {                                                                                           \
        hipError_t e_ = (expr);                                                                    \
        if (e_ != hipSuccess)                                                                      \
            return fail(POE2D_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));           \
    } while (0)

enum { ST_VX, ST_VZ, ST_WX, ST_WZ, ST_EXX, ST_EYY, ST_ZETA, ST_SXY, NSTATE };
enum { IN_P, IN_SXX, IN_SYY, NINTER };
enum { MT_RHO, MT_RHOF, MT_M, MT_MU, MT_LC, MT_M_BIOT, MT_ALPHA, MT_B, NMAT };
enum { SCHED_FUSED = 0, SCHED_TWOPASS = 1 };

size_t real_size(int precision) { return precision == POE2D_F32 ? 4 : 8; }

}  // namespace

struct poe2d_solver {
    int precision = POE2D_F32;
    size_t rs = 4;  // sizeof(real)
    poe2d_grid grid{};
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    size_t pitch = 0, alloc_rows = 0;  // padded row length and allocated rows (nz+1+2*pad)
    int wxs = 0, wzs = 0;              // x-strip columns, z-strip rows
    double cp_max = 0;
    int cur = 0;
    void *st[2][NSTATE] = {};
    void *inter[NINTER] = {};
    void *mat[NMAT] = {};
    void *psx[2][NPSI_X] = {};
    void *psz[2][NPSI_Z] = {};
    void *tabx[6] = {};
    void *tabz[6] = {};
    std::vector<double> htabx[6], htabz[6];
    int *xmap = nullptr, *zmap = nullptr;
    float *red = nullptr;  // reduction scratch: [0] result, [1..] per-block partials
    int red_blocks = 1024;
    void *dense = nullptr;  // (nz+1) x (nx+1) scratch for derived-field snapshots, made on first use
    // schedule
    int schedule = SCHED_FUSED;
    int block_w = 256, tile_rows = 64;
    // source
    bool have_src = false;
    poe2d_source src{};
    double t0 = 0;
    std::vector<double> user_w;
    bool have_user_w = false;
    // receivers
    int nrec = 0, dec = 1, nsamp = 0, ncomp = 0;
    unsigned mask = 0;
    int *rix = nullptr, *riz = nullptr;
    void *traces = nullptr;  // [slot][comp][rec]
    int step = 0;
    double last_ms = 0;
    int last_steps = 0;

    size_t field_elems() const { return pitch * alloc_rows; }
    size_t field_bytes() const { return field_elems() * rs; }
    size_t origin() const { return (size_t)kPadRows * pitch + kPadCols; }  // elements to point (0,0)
    size_t psx_bytes() const { return (size_t)std::max(wxs, 1) * (grid.nz + 1) * rs; }
    size_t psz_bytes() const { return (size_t)std::max(wzs, 1) * pitch * rs; }
    size_t trace_bytes() const { return (size_t)nsamp * ncomp * nrec * rs; }
    template <class T> T *at0(void *p) const { return (T *)p + origin(); }

    double source_value(int n) const
    {
        if (!have_src) return 0.0;
        const double t = n * grid.dt;
        if (src.t_cut > 0 && t >= src.t_cut) return 0.0;
        if (have_user_w) return n < (int)user_w.size() ? src.amp * user_w[n] : 0.0;
        return src.amp * wavelet_shape(src.kind, grid.f0, t0, t);
    }
};

namespace {

void free_all(poe2d_solver *s)
{
    auto f = [](void *&p) {
        if (p) (void)hipFree(p);
        p = nullptr;
    };
    for (auto &b : s->st)
        for (auto &p : b) f(p);
    for (auto &p : s->inter) f(p);
    for (auto &p : s->mat) f(p);
    for (auto &b : s->psx)
        for (auto &p : b) f(p);
    for (auto &b : s->psz)
        for (auto &p : b) f(p);
    for (auto &p : s->tabx) f(p);
    for (auto &p : s->tabz) f(p);
    void *q = s->xmap; f(q); s->xmap = nullptr;
    q = s->zmap; f(q); s->zmap = nullptr;
    q = s->red; f(q); s->red = nullptr;
    q = s->rix; f(q); s->rix = nullptr;
    q = s->riz; f(q); s->riz = nullptr;
    f(s->traces);
    f(s->dense);
    if (s->ev0) (void)hipEventDestroy(s->ev0);
    if (s->ev1) (void)hipEventDestroy(s->ev1);
    if (s->stream) (void)hipStreamDestroy(s->stream);
    s->ev0 = s->ev1 = nullptr;
    s->stream = nullptr;
}

poe2d_status alloc_zero(poe2d_solver *s, void **p, size_t bytes)
{
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) {
        *p = nullptr;
        return fail(e == hipErrorOutOfMemory ? POE2D_E_NOMEM : POE2D_E_HIP,
                    std::string("hipMalloc(") + std::to_string(bytes) + "): " + hipGetErrorString(e));
    }
    HIP_TRY(hipMemsetAsync(*p, 0, bytes, s->stream));
    return POE2D_OK;
}

template <class T>
poe2d_status upload_coeffs(poe2d_solver *s, const poe2d_coeffs *c)
{
    const poe2d_grid &g = s->grid;
    const size_t cpitch = c->pitch ? c->pitch : (size_t)g.nx + 1;
    const void *srcs[NMAT] = {c->rho, c->rhof, c->m, c->mu, c->lambdac, c->M, c->alpha, c->b};
    auto at = [&](int k, size_t i) {
        return c->dtype == POE2D_F64 ? ((const double *)srcs[k])[i] : (double)((const float *)srcs[k])[i];
    };
    // the velocity update divides by det = rho m - rho_f^2 (the 2x2 mass matrix) and needs b >= 0
    for (int z = 0; z <= g.nz; ++z)
        for (int x = 0; x <= g.nx; ++x) {
            const size_t i = (size_t)z * cpitch + x;
            const double det = at(MT_RHO, i) * at(MT_M, i) - at(MT_RHOF, i) * at(MT_RHOF, i);
            if (!(det > 0) || !(at(MT_B, i) >= 0) || !(at(MT_RHO, i) > 0))
                return fail(POE2D_E_INVALID, "create: coefficients need rho > 0, rho m > rho_f^2 and b >= 0 (point " +
                                                 std::to_string(x) + ", " + std::to_string(z) + ")");
        }
    std::vector<T> host(s->field_elems(), T(0));
    for (int k = 0; k < NMAT; ++k) {
        for (int z = 0; z <= g.nz; ++z)
            for (int x = 0; x <= g.nx; ++x) {
                const double v = at(k, (size_t)z * cpitch + x);
                if (!std::isfinite(v)) return fail(POE2D_E_INVALID, "create: non-finite coefficient");
                host[s->origin() + (size_t)z * s->pitch + x] = (T)v;
            }
        HIP_TRY(hipMemcpy(s->mat[k], host.data(), s->field_bytes(), hipMemcpyHostToDevice));
    }
    return POE2D_OK;
}

template <class T>
poe2d_status upload_tables(poe2d_solver *s)
{
    std::vector<T> h;
    for (int k = 0; k < 6; ++k) {
        h.assign(s->htabx[k].begin(), s->htabx[k].end());
        HIP_TRY(hipMemcpy(s->tabx[k], h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
        h.assign(s->htabz[k].begin(), s->htabz[k].end());
        HIP_TRY(hipMemcpy(s->tabz[k], h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
    }
    return POE2D_OK;
}

// Strip map of one axis: index into the strip arrays where any profile deviates from the
// no-absorption values (a = 0, K = 1), else -1.
std::vector<int> strip_map(const std::vector<double> *t, int n, int *count)
{
    std::vector<int> map(n, -1);
    int c = 0;
    for (int i = 0; i < n; ++i) {
        const bool active = t[0][i] != 0.0 || t[3][i] != 0.0 || t[2][i] != 1.0 || t[5][i] != 1.0;
        if (active) map[i] = c++;
    }
    *count = c;
    return map;
}

template <class T>
void set_tables(poe2d_solver *s, const T **x6, const T **z6)
{
    for (int k = 0; k < 6; ++k) {
        x6[k] = (const T *)s->tabx[k];
        z6[k] = (const T *)s->tabz[k];
    }
}

template <class T>
StepArgs<T> twopass_args(poe2d_solver *s)
{
    StepArgs<T> A{};
    const poe2d_grid &g = s->grid;
    const int b = s->cur;
    A.nx = g.nx;
    A.nz = g.nz;
    A.pitch = (int)s->pitch;
    A.wxs = s->wxs;
    A.rdx = T(1.0 / (24.0 * g.dx));
    A.rdz = T(1.0 / (24.0 * g.dz));
    A.dt = T(g.dt);
    A.inv_dt = T(1.0 / g.dt);
    T **st[NSTATE] = {&A.vx, &A.vz, &A.wx, &A.wz, &A.exx, &A.eyy, &A.zeta, &A.sxy};
    for (int k = 0; k < NSTATE; ++k) *st[k] = s->at0<T>(s->st[b][k]);
    A.p = s->at0<T>(s->inter[IN_P]);
    A.sxx = s->at0<T>(s->inter[IN_SXX]);
    A.syy = s->at0<T>(s->inter[IN_SYY]);
    const T **mt[NMAT] = {&A.rho, &A.rhof, &A.m, &A.mu, &A.lc, &A.M, &A.al, &A.b};
    for (int k = 0; k < NMAT; ++k) *mt[k] = s->at0<T>(s->mat[k]);
    for (int k = 0; k < NPSI_X; ++k) A.psx[k] = (T *)s->psx[b][k];
    for (int k = 0; k < NPSI_Z; ++k) A.psz[k] = (T *)s->psz[b][k];
    const T *x6[6], *z6[6];
    set_tables<T>(s, x6, z6);
    A.ax = x6[0]; A.bx = x6[1]; A.Kx = x6[2]; A.axh = x6[3]; A.bxh = x6[4]; A.Kxh = x6[5];
    A.az = z6[0]; A.bz = z6[1]; A.Kz = z6[2]; A.azh = z6[3]; A.bzh = z6[4]; A.Kzh = z6[5];
    A.xmap = s->xmap;
    A.zmap = s->zmap;
    A.src_kind = s->have_src ? s->src.kind : -1;
    A.src_x = s->have_src ? s->src.ix : -1;
    A.src_z = s->have_src ? s->src.iz : -1;
    A.src_val = T(0);
    return A;
}

template <class T>
FusedArgs<T> fused_args(poe2d_solver *s, int in, int out)
{
    FusedArgs<T> A{};
    const poe2d_grid &g = s->grid;
    A.nx = g.nx;
    A.nz = g.nz;
    A.pitch = (int)s->pitch;
    A.wxs = s->wxs;
    const int W = fused_tile_width(s->block_w);
    A.th = s->tile_rows;
    A.ntx = (g.nx + 1 + W - 1) / W;
    const int ntz = (g.nz + 1 + A.th - 1) / A.th;
    A.ntiles = A.ntx * ntz;
    A.per_xcd = (A.ntiles + 7) / 8;
    A.rdx = T(1.0 / (24.0 * g.dx));
    A.rdz = T(1.0 / (24.0 * g.dz));
    A.dt = T(g.dt);
    A.inv_dt = T(1.0 / g.dt);
    const T **so[NSTATE] = {&A.vx_o, &A.vz_o, &A.wx_o, &A.wz_o, &A.exx_o, &A.eyy_o, &A.zeta_o, &A.sxy_o};
    T **sn[NSTATE] = {&A.vx_n, &A.vz_n, &A.wx_n, &A.wz_n, &A.exx_n, &A.eyy_n, &A.zeta_n, &A.sxy_n};
    for (int k = 0; k < NSTATE; ++k) {
        *so[k] = s->at0<T>(s->st[in][k]);
        *sn[k] = s->at0<T>(s->st[out][k]);
    }
    const T **mt[NMAT] = {&A.rho, &A.rhof, &A.m, &A.mu, &A.lc, &A.M, &A.al, &A.b};
    for (int k = 0; k < NMAT; ++k) *mt[k] = s->at0<T>(s->mat[k]);
    for (int k = 0; k < NPSI_X; ++k) {
        A.psx_o[k] = (const T *)s->psx[in][k];
        A.psx_n[k] = (T *)s->psx[out][k];
    }
    for (int k = 0; k < NPSI_Z; ++k) {
        A.psz_o[k] = (const T *)s->psz[in][k];
        A.psz_n[k] = (T *)s->psz[out][k];
    }
    const T *x6[6], *z6[6];
    set_tables<T>(s, x6, z6);
    A.ax = x6[0]; A.bx = x6[1]; A.Kx = x6[2]; A.axh = x6[3]; A.bxh = x6[4]; A.Kxh = x6[5];
    A.az = z6[0]; A.bz = z6[1]; A.Kz = z6[2]; A.azh = z6[3]; A.bzh = z6[4]; A.Kzh = z6[5];
    A.xmap = s->xmap;
    A.zmap = s->zmap;
    A.src_kind = s->have_src ? s->src.kind : -1;
    A.src_x = s->have_src ? s->src.ix : -1;
    A.src_z = s->have_src ? s->src.iz : -1;
    A.src_val = T(0);
    return A;
}

// Gather/derived-field arguments: velocities from buffer `vbuf`, strains from buffer `sbuf`.
template <class T>
GatherArgs<T> gather_args(poe2d_solver *s, int vbuf, int sbuf, T src_val)
{
    GatherArgs<T> G{};
    G.nx = s->grid.nx;
    G.nz = s->grid.nz;
    G.pitch = (int)s->pitch;
    G.nrec = s->nrec;
    G.mask = s->mask;
    G.rix = s->rix;
    G.riz = s->riz;
    G.vx = s->at0<T>(s->st[vbuf][ST_VX]);
    G.vz = s->at0<T>(s->st[vbuf][ST_VZ]);
    G.exx = s->at0<T>(s->st[sbuf][ST_EXX]);
    G.eyy = s->at0<T>(s->st[sbuf][ST_EYY]);
    G.zeta = s->at0<T>(s->st[sbuf][ST_ZETA]);
    G.al = s->at0<T>(s->mat[MT_ALPHA]);
    G.M = s->at0<T>(s->mat[MT_M_BIOT]);
    G.src_kind = s->have_src ? s->src.kind : -1;
    G.src_x = s->have_src ? s->src.ix : -1;
    G.src_z = s->have_src ? s->src.iz : -1;
    G.src_val = src_val;
    G.out = nullptr;
    return G;
}

template <class T>
poe2d_status run_steps(poe2d_solver *s, int nsteps)
{
    HIP_TRY(hipEventRecord(s->ev0, s->stream));
    if (s->schedule == SCHED_TWOPASS) {
        StepArgs<T> A = twopass_args<T>(s);
        for (int k = 0; k < nsteps; ++k) {
            const int n = s->step + k;
            A.src_val = (T)s->source_value(n);
            launch_twopass<T>(A, s->stream, 0);
            if (s->nrec > 0 && n % s->dec == 0 && n / s->dec < s->nsamp) {
                // in place: v is still v^n, strains are already those of step n
                GatherArgs<T> G = gather_args<T>(s, s->cur, s->cur, A.src_val);
                G.out = (T *)s->traces + (size_t)(n / s->dec) * s->ncomp * s->nrec;
                launch_gather<T>(G, s->stream);
            }
            launch_twopass<T>(A, s->stream, 1);
        }
    } else {
        FusedArgs<T> F[2] = {fused_args<T>(s, 0, 1), fused_args<T>(s, 1, 0)};
        for (int k = 0; k < nsteps; ++k) {
            const int n = s->step + k;
            FusedArgs<T> &A = F[s->cur];
            A.src_val = (T)s->source_value(n);
            HIP_TRY(launch_fused<T>(A, s->block_w, s->stream));
            if (s->nrec > 0 && n % s->dec == 0 && n / s->dec < s->nsamp) {
                GatherArgs<T> G = gather_args<T>(s, s->cur, 1 - s->cur, A.src_val);
                G.out = (T *)s->traces + (size_t)(n / s->dec) * s->ncomp * s->nrec;
                launch_gather<T>(G, s->stream);
            }
            s->cur = 1 - s->cur;
        }
    }
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventRecord(s->ev1, s->stream));
    HIP_TRY(hipEventSynchronize(s->ev1));
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, s->ev0, s->ev1));
    s->last_ms = ms;
    s->last_steps = nsteps;
    s->step += nsteps;
    // stability check: largest |v| over the grid (pads are zero)
    launch_absmax<T>((const T *)s->st[s->cur][ST_VX], (const T *)s->st[s->cur][ST_VZ], s->field_elems(),
                     s->red + 1, s->red_blocks, s->red, s->stream);
    float vmax = 0.f;
    HIP_TRY(hipMemcpyAsync(&vmax, s->red, sizeof(float), hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    if (!(vmax <= kStabilityThreshold))
        return fail(POE2D_E_UNSTABLE, "run: wavefield became unstable (max |v| = " + std::to_string(vmax) +
                                          ") by step " + std::to_string(s->step));
    return POE2D_OK;
}

template <class T>
poe2d_status derived_snapshot(poe2d_solver *s, int what, void *out)
{
    const double sv = s->step > 0 ? s->source_value(s->step - 1) : 0.0;
    GatherArgs<T> G = gather_args<T>(s, s->cur, s->cur, (T)sv);
    const size_t bytes = (size_t)(s->grid.nx + 1) * (s->grid.nz + 1) * sizeof(T);
    if (!s->dense) {
        poe2d_status st = alloc_zero(s, &s->dense, bytes);
        if (st) return st;
    }
    T *dense = (T *)s->dense;
    launch_derived<T>(G, s->at0<T>(s->mat[MT_LC]), s->at0<T>(s->mat[MT_MU]), what, dense, s->stream);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(s->stream));
    HIP_TRY(hipMemcpy(out, dense, bytes, hipMemcpyDeviceToHost));
    return POE2D_OK;
}

poe2d_status check_handle(poe2d_solver *s)
{
    if (!s) return fail(POE2D_E_INVALID, "null solver handle");
    HIP_TRY(hipSetDevice(s->device));
    return POE2D_OK;
}

struct CkptHeader {
    char magic[8];
    int version, precision, nx, nz, step, nrec, ncomp, nsamp, wxs, wzs;
    unsigned long long pitch;
};

template <class F>
void for_each_buffer(poe2d_solver *s, int b, F f)
{
    for (auto p : s->st[b]) f(p, s->field_bytes());
    for (auto p : s->psx[b]) f(p, s->psx_bytes());
    for (auto p : s->psz[b]) f(p, s->psz_bytes());
    if (s->traces) f(s->traces, s->trace_bytes());
}

}  // namespace

extern "C" {

poe2d_status poe2d_device_info(int *device_count, char *arch, size_t arch_len)
{
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (device_count) *device_count = e == hipSuccess ? n : 0;
    if (e != hipSuccess || n == 0) return fail(POE2D_E_HIP, "no HIP device available");
    hipDeviceProp_t prop;
    HIP_TRY(hipGetDeviceProperties(&prop, 0));
    if (arch && arch_len) {
        std::strncpy(arch, prop.gcnArchName, arch_len - 1);
        arch[arch_len - 1] = 0;
    }
    return POE2D_OK;
}

poe2d_status poe2d_create(const poe2d_grid *g, const poe2d_coeffs *c, poe2d_solver **out)
{
    if (!out) return fail(POE2D_E_INVALID, "create: null output pointer");
    *out = nullptr;
    if (!g || !c) return fail(POE2D_E_INVALID, "create: null grid or coefficients");
    if (g->nx < 8 || g->nz < 8) return fail(POE2D_E_INVALID, "create: grid needs nx, nz >= 8");
    if (!(g->dx > 0) || !(g->dz > 0) || !(g->dt > 0)) return fail(POE2D_E_INVALID, "create: dx, dz, dt must be > 0");
    if (g->npml < 1 || 2 * g->npml + 8 > std::min(g->nx, g->nz))
        return fail(POE2D_E_INVALID, "create: npml must be >= 1 and leave an interior");
    if (g->npower < 1) return fail(POE2D_E_INVALID, "create: npower must be >= 1");
    if (!(g->rcoef > 0 && g->rcoef < 1)) return fail(POE2D_E_INVALID, "create: rcoef must be in (0, 1)");
    if (g->precision != POE2D_F32 && g->precision != POE2D_F64)
        return fail(POE2D_E_INVALID, "create: precision must be POE2D_F32 or POE2D_F64");
    if (c->dtype != POE2D_F32 && c->dtype != POE2D_F64)
        return fail(POE2D_E_INVALID, "create: coefficient dtype must be POE2D_F32 or POE2D_F64");
    if (!c->rho || !c->rhof || !c->m || !c->mu || !c->lambdac || !c->M || !c->alpha || !c->b)
        return fail(POE2D_E_INVALID, "create: null coefficient array");
    if (c->pitch && c->pitch < (size_t)g->nx + 1) return fail(POE2D_E_INVALID, "create: coefficient pitch < nx+1");
    if (!(g->f0 > 0)) return fail(POE2D_E_INVALID, "create: f0 must be > 0");

    poe2d_solver *s = new (std::nothrow) poe2d_solver;
    if (!s) return fail(POE2D_E_NOMEM, "create: out of host memory");
    s->grid = *g;
    s->precision = g->precision;
    s->rs = real_size(g->precision);
    auto bail = [&](poe2d_status st) {
        free_all(s);
        delete s;
        return st;
    };
    hipError_t e = hipGetDevice(&s->device);
    if (e != hipSuccess) {
        fail(POE2D_E_HIP, std::string("create: no HIP device: ") + hipGetErrorString(e));
        return bail(POE2D_E_HIP);
    }
    if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&s->ev0) != hipSuccess || hipEventCreate(&s->ev1) != hipSuccess) {
        fail(POE2D_E_HIP, "create: stream/event creation failed");
        return bail(POE2D_E_HIP);
    }

    // stability and C-PML set-up
    s->cp_max = g->cp_max > 0 ? g->cp_max : cp_max_from_coeffs(*g, *c);
    if (!(s->cp_max > 0) || !std::isfinite(s->cp_max)) {
        fail(POE2D_E_INVALID, "create: coefficients give no positive P velocity");
        return bail(POE2D_E_INVALID);
    }
    const double courant = s->cp_max * g->dt / std::min(g->dx, g->dz);
    if (g->check_cfl && courant > 6.0 / 7.0 / std::sqrt(2.0)) {
        fail(POE2D_E_CFL, "create: time step is too large, simulation will be unstable (Courant " +
                              std::to_string(courant) + " > 6/7/sqrt(2))");
        return bail(POE2D_E_CFL);
    }
    double amax = g->alpha_max;
    if (amax <= 0)
        amax = g->precision == POE2D_F32 ? (double)(float)(2.e0 * kRefPiF * ((float)g->f0 / 2.e0))
                                         : 3.14159265358979323846 * g->f0;
    for (int axis = 0; axis < 2; ++axis) {
        const int n = (axis == 0 ? g->nx : g->nz) + 1;
        std::vector<double> *t = axis == 0 ? s->htabx : s->htabz;
        for (int k = 0; k < 6; ++k) t[k].assign(n, 0.0);
        poe2d_status st = poe2d_cpml_profile(axis, n, axis == 0 ? g->dx : g->dz, g->dt, g->npml,
                                             g->npower, g->rcoef, g->kmax, amax, s->cp_max,
                                             g->precision, t[0].data(), t[1].data(), t[2].data(),
                                             t[3].data(), t[4].data(), t[5].data());
        if (st != POE2D_OK) return bail(st);
    }
    std::vector<int> xm = strip_map(s->htabx, g->nx + 1, &s->wxs);
    std::vector<int> zm = strip_map(s->htabz, g->nz + 1, &s->wzs);

    s->pitch = (((size_t)kPadCols + g->nx + 2 + 63) / 64) * 64;
    s->alloc_rows = (size_t)g->nz + 1 + 2 * kPadRows;
    poe2d_status st = POE2D_OK;
    for (auto &b : s->st)
        for (auto &p : b)
            if ((st = alloc_zero(s, &p, s->field_bytes()))) return bail(st);
    for (auto &p : s->inter)
        if ((st = alloc_zero(s, &p, s->field_bytes()))) return bail(st);
    for (auto &p : s->mat)
        if ((st = alloc_zero(s, &p, s->field_bytes()))) return bail(st);
    for (auto &b : s->psx)
        for (auto &p : b)
            if ((st = alloc_zero(s, &p, s->psx_bytes()))) return bail(st);
    for (auto &b : s->psz)
        for (auto &p : b)
            if ((st = alloc_zero(s, &p, s->psz_bytes()))) return bail(st);
    for (int k = 0; k < 6; ++k) {
        if ((st = alloc_zero(s, &s->tabx[k], (size_t)(g->nx + 1) * s->rs))) return bail(st);
        if ((st = alloc_zero(s, &s->tabz[k], (size_t)(g->nz + 1) * s->rs))) return bail(st);
    }
    void *q = nullptr;
    if ((st = alloc_zero(s, &q, xm.size() * sizeof(int)))) return bail(st);
    s->xmap = (int *)q;
    if ((st = alloc_zero(s, &q, zm.size() * sizeof(int)))) return bail(st);
    s->zmap = (int *)q;
    if ((st = alloc_zero(s, &q, (s->red_blocks + 1) * sizeof(float)))) return bail(st);
    s->red = (float *)q;
    if (hipStreamSynchronize(s->stream) != hipSuccess ||
        hipMemcpy(s->xmap, xm.data(), xm.size() * sizeof(int), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(s->zmap, zm.data(), zm.size() * sizeof(int), hipMemcpyHostToDevice) != hipSuccess) {
        fail(POE2D_E_HIP, "create: map upload failed");
        return bail(POE2D_E_HIP);
    }
    st = s->precision == POE2D_F32 ? upload_tables<float>(s) : upload_tables<double>(s);
    if (st) return bail(st);
    st = s->precision == POE2D_F32 ? upload_coeffs<float>(s, c) : upload_coeffs<double>(s, c);
    if (st) return bail(st);
    *out = s;
    return POE2D_OK;
}

poe2d_status poe2d_set_schedule(poe2d_solver *s, int schedule, int block_w, int tile_rows)
{
    poe2d_status st = check_handle(s);
    if (st) return st;
    if (schedule != SCHED_FUSED && schedule != SCHED_TWOPASS)
        return fail(POE2D_E_INVALID, "set_schedule: schedule must be 0 (fused) or 1 (two-pass)");
    if (block_w != 128 && block_w != 256 && block_w != 512)
        return fail(POE2D_E_INVALID, "set_schedule: block width must be 128, 256 or 512");
    if (tile_rows < 4 || tile_rows > 1 << 20) return fail(POE2D_E_INVALID, "set_schedule: tile rows out of range");
    s->schedule = schedule;
    s->block_w = block_w;
    s->tile_rows = tile_rows;
    return POE2D_OK;
}

poe2d_status poe2d_set_source(poe2d_solver *s, const poe2d_source *src)
{
    poe2d_status st = check_handle(s);
    if (st) return st;
    if (!src) return fail(POE2D_E_INVALID, "set_source: null source");
    const poe2d_grid &g = s->grid;
    if (src->kind == POE2D_SRC_PRESSURE_DGAUSS) {
        // p is only formed inside the strain-update region (x in [1,NX), z in [2,NZ-1))
        if (src->ix < 1 || src->ix >= g.nx || src->iz < 2 || src->iz >= g.nz - 1)
            return fail(POE2D_E_INVALID, "set_source: pressure source outside x in [1,NX), z in [2,NZ-1)");
    } else if (src->kind == POE2D_SRC_FORCE_Y_RICKER) {
        if (src->ix < 1 || src->ix >= g.nx - 1 || src->iz < 1 || src->iz >= g.nz - 1)
            return fail(POE2D_E_INVALID, "set_source: force source outside x in [1,NX-1), z in [1,NZ-1)");
    } else {
        return fail(POE2D_E_INVALID, "set_source: unknown source kind");
    }
    if (!std::isfinite(src->amp)) return fail(POE2D_E_INVALID, "set_source: non-finite amplitude");
    s->src = *src;
    s->t0 = src->t0 > 0 ? src->t0 : 1.0 / g.f0;
    s->have_src = true;
    return POE2D_OK;
}

poe2d_status poe2d_set_wavelet(poe2d_solver *s, const double *w, int n)
{
    poe2d_status st = check_handle(s);
    if (st) return st;
    if (n < 0 || (n > 0 && !w)) return fail(POE2D_E_INVALID, "set_wavelet: bad samples");
    s->user_w.assign(w, w + n);
    s->have_user_w = true;
    return POE2D_OK;
}

poe2d_status poe2d_set_receivers(poe2d_solver *s, const poe2d_receivers *r)
{
    poe2d_status st = check_handle(s);
    if (st) return st;
    if (!r || r->n < 0 || r->decimate < 1 || r->nsamp < 0 || (r->comp_mask & ~7u) ||
        (r->n > 0 && (!r->ix || !r->iz)))
        return fail(POE2D_E_INVALID, "set_receivers: bad receiver description");
    for (int k = 0; k < r->n; ++k)
        if (r->ix[k] < 0 || r->ix[k] > s->grid.nx || r->iz[k] < 0 || r->iz[k] > s->grid.nz)
            return fail(POE2D_E_INVALID, "set_receivers: receiver " + std::to_string(k) + " outside the grid");
    HIP_TRY(hipStreamSynchronize(s->stream));
    void *q = s->rix; if (q) (void)hipFree(q); s->rix = nullptr;
    q = s->riz; if (q) (void)hipFree(q); s->riz = nullptr;
    if (s->traces) (void)hipFree(s->traces);
    s->traces = nullptr;
    s->nrec = r->n;
    s->mask = r->comp_mask;
    s->ncomp = ((r->comp_mask & 1) != 0) + ((r->comp_mask & 2) != 0) + ((r->comp_mask & 4) != 0);
    s->dec = r->decimate;
    s->nsamp = r->nsamp;
    if (s->nrec == 0) return POE2D_OK;
    if ((st = alloc_zero(s, &q, (size_t)r->n * sizeof(int)))) return st;
    s->rix = (int *)q;
    if ((st = alloc_zero(s, &q, (size_t)r->n * sizeof(int)))) return st;
    s->riz = (int *)q;
    if ((st = alloc_zero(s, &s->traces, std::max<size_t>(s->trace_bytes(), 1)))) return st;
    HIP_TRY(hipStreamSynchronize(s->stream));
    HIP_TRY(hipMemcpy(s->rix, r->ix, (size_t)r->n * sizeof(int), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(s->riz, r->iz, (size_t)r->n * sizeof(int), hipMemcpyHostToDevice));
    return POE2D_OK;
}

poe2d_status poe2d_run(poe2d_solver *s, int nsteps)
{
    poe2d_status st = check_handle(s);
    if (st) return st;
    if (nsteps < 0) return fail(POE2D_E_INVALID, "run: negative step count");
    if (nsteps == 0) return POE2D_OK;
    return s->precision == POE2D_F32 ? run_steps<float>(s, nsteps) : run_steps<double>(s, nsteps);
}

poe2d_status poe2d_last_run_ms(poe2d_solver *s, double *ms, int *nsteps)
{
    if (!s || !ms || !nsteps) return fail(POE2D_E_INVALID, "last_run_ms: null argument");
    *ms = s->last_ms;
    *nsteps = s->last_steps;
    return POE2D_OK;
}

poe2d_status poe2d_step_count(poe2d_solver *s, int *n)
{
    if (!s || !n) return fail(POE2D_E_INVALID, "step_count: null argument");
    *n = s->step;
    return POE2D_OK;
}

poe2d_status poe2d_seismograms(poe2d_solver *s, void *out, size_t bytes)
{
    poe2d_status st = check_handle(s);
    if (st) return st;
    if (!out) return fail(POE2D_E_INVALID, "seismograms: null output");
    const size_t need = s->trace_bytes();
    if (bytes < need) return fail(POE2D_E_INVALID, "seismograms: output buffer too small");
    if (need == 0) return POE2D_OK;
    std::vector<unsigned char> tmp(need);
    HIP_TRY(hipStreamSynchronize(s->stream));
    HIP_TRY(hipMemcpy(tmp.data(), s->traces, need, hipMemcpyDeviceToHost));
    // device layout [slot][comp][rec] -> caller layout [comp][rec][slot]
    const size_t rs = s->rs;
    unsigned char *o = (unsigned char *)out;
    for (int k = 0; k < s->nsamp; ++k)
        for (int c = 0; c < s->ncomp; ++c)
            for (int r = 0; r < s->nrec; ++r)
                std::memcpy(o + (((size_t)c * s->nrec + r) * s->nsamp + k) * rs,
                            tmp.data() + (((size_t)k * s->ncomp + c) * s->nrec + r) * rs, rs);
    return POE2D_OK;
}

poe2d_status poe2d_snapshot(poe2d_solver *s, int field, void *out, size_t bytes)
{
    poe2d_status st = check_handle(s);
    if (st) return st;
    if (!out) return fail(POE2D_E_INVALID, "snapshot: null output");
    const size_t w = (size_t)s->grid.nx + 1;
    const size_t need = w * (s->grid.nz + 1) * s->rs;
    if (bytes < need) return fail(POE2D_E_INVALID, "snapshot: output buffer too small");
    int idx = -1;
    switch (field) {
    case POE2D_FIELD_VX: idx = ST_VX; break;
    case POE2D_FIELD_VY: idx = ST_VZ; break;
    case POE2D_FIELD_WX: idx = ST_WX; break;
    case POE2D_FIELD_WY: idx = ST_WZ; break;
    case POE2D_FIELD_EXX: idx = ST_EXX; break;
    case POE2D_FIELD_EYY: idx = ST_EYY; break;
    case POE2D_FIELD_ZETA: idx = ST_ZETA; break;
    case POE2D_FIELD_SXY: idx = ST_SXY; break;
    case POE2D_FIELD_P:
    case POE2D_FIELD_SXX:
    case POE2D_FIELD_SYY: {
        const int what = field - POE2D_FIELD_P;
        return s->precision == POE2D_F32 ? derived_snapshot<float>(s, what, out)
                                         : derived_snapshot<double>(s, what, out);
    }
    default: return fail(POE2D_E_INVALID, "snapshot: unknown field");
    }
    HIP_TRY(hipStreamSynchronize(s->stream));
    const unsigned char *src = (const unsigned char *)s->st[s->cur][idx] + s->origin() * s->rs;
    HIP_TRY(hipMemcpy2D(out, w * s->rs, src, s->pitch * s->rs, w * s->rs, s->grid.nz + 1, hipMemcpyDeviceToHost));
    return POE2D_OK;
}

poe2d_status poe2d_checkpoint_size(poe2d_solver *s, size_t *bytes)
{
    if (!s || !bytes) return fail(POE2D_E_INVALID, "checkpoint_size: null argument");
    size_t n = sizeof(CkptHeader);
    for_each_buffer(s, 0, [&](void *, size_t b) { n += b; });
    *bytes = n;
    return POE2D_OK;
}

poe2d_status poe2d_checkpoint(poe2d_solver *s, void *blob, size_t bytes)
{
    poe2d_status st = check_handle(s);
    if (st) return st;
    size_t need = 0;
    poe2d_checkpoint_size(s, &need);
    if (!blob || bytes < need) return fail(POE2D_E_INVALID, "checkpoint: buffer too small");
    CkptHeader h{};
    std::memcpy(h.magic, "POE2DCK", 8);
    h.version = 1; h.precision = s->precision; h.nx = s->grid.nx; h.nz = s->grid.nz;
    h.step = s->step; h.nrec = s->nrec; h.ncomp = s->ncomp; h.nsamp = s->nsamp;
    h.wxs = s->wxs; h.wzs = s->wzs; h.pitch = s->pitch;
    std::memcpy(blob, &h, sizeof h);
    HIP_TRY(hipStreamSynchronize(s->stream));
    unsigned char *p = (unsigned char *)blob + sizeof h;
    hipError_t err = hipSuccess;
    for_each_buffer(s, s->cur, [&](void *d, size_t b) {
        if (err == hipSuccess) err = hipMemcpy(p, d, b, hipMemcpyDeviceToHost);
        p += b;
    });
    HIP_TRY(err);
    return POE2D_OK;
}

poe2d_status poe2d_restore(poe2d_solver *s, const void *blob, size_t bytes)
{
    poe2d_status st = check_handle(s);
    if (st) return st;
    size_t need = 0;
    poe2d_checkpoint_size(s, &need);
    if (!blob || bytes < need) return fail(POE2D_E_INVALID, "restore: blob too small");
    CkptHeader h;
    std::memcpy(&h, blob, sizeof h);
    if (std::memcmp(h.magic, "POE2DCK", 8) || h.version != 1 || h.precision != s->precision ||
        h.nx != s->grid.nx || h.nz != s->grid.nz || h.nrec != s->nrec || h.ncomp != s->ncomp ||
        h.nsamp != s->nsamp || h.wxs != s->wxs || h.wzs != s->wzs || h.pitch != s->pitch)
        return fail(POE2D_E_INVALID, "restore: checkpoint does not match this solver's geometry");
    HIP_TRY(hipStreamSynchronize(s->stream));
    const unsigned char *p = (const unsigned char *)blob + sizeof h;
    hipError_t err = hipSuccess;
    for_each_buffer(s, s->cur, [&](void *d, size_t b) {
        if (err == hipSuccess) err = hipMemcpy(d, p, b, hipMemcpyHostToDevice);
        p += b;
    });
    HIP_TRY(err);
    s->step = h.step;
    return POE2D_OK;
}

poe2d_status poe2d_reset(poe2d_solver *s)
{
    poe2d_status st = check_handle(s);
    if (st) return st;
    for (auto &b : s->st)
        for (auto p : b) HIP_TRY(hipMemsetAsync(p, 0, s->field_bytes(), s->stream));
    for (auto p : s->inter) HIP_TRY(hipMemsetAsync(p, 0, s->field_bytes(), s->stream));
    for (auto &b : s->psx)
        for (auto p : b) HIP_TRY(hipMemsetAsync(p, 0, s->psx_bytes(), s->stream));
    for (auto &b : s->psz)
        for (auto p : b) HIP_TRY(hipMemsetAsync(p, 0, s->psz_bytes(), s->stream));
    if (s->traces) HIP_TRY(hipMemsetAsync(s->traces, 0, s->trace_bytes(), s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    s->step = 0;
    s->cur = 0;
    return POE2D_OK;
}

void poe2d_destroy(poe2d_solver *s)
{
    if (!s) return;
    (void)hipSetDevice(s->device);
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    free_all(s);
    delete s;
}

}  // extern "C"
