/*
 * poe2d_oracle.cpp -- TEST INFRASTRUCTURE ONLY (the oracle).
 *
 * A CPU restatement of the reference's 2-D poroelastic (Biot) staggered-grid C-PML time loop,
 * the answer tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg compare the library with.  The
 * product (libpoe2d.so) neither links nor calls it.
 *
 * Followed files (reference, read as text):
 *   seismic_CPML_2d_poroelastic.cpp   (CPP) -- loop order, index ranges, arithmetic: :614-839
 *   seismic_CPML_2D_poroelastic_withCUDA.cu (CU) -- addressing (:53, :1597-1600), fluid memory
 *                                                  variables (:79-80), source gate (:107),
 *                                                  recording slot (:183-199 with :1627)
 *
 * "Canonical" semantics restated here (each choice reads the CU file where the CPP file is
 * self-inconsistent):
 *   - every offset is x + z*(NX+1) (CU:53 with the (NX+1)-wide launch of CU:1597-1600); the CPP
 *     centre offset x + y*NX (CPP:620) shears the grid one column per row and is not followed;
 *   - the fluid-divergence memory variables psi3/psi4 are updated (CU:79-80; CPP never assigns
 *     them);
 *   - exactly n steps, t_n = n*dt; trace slot k holds the state at the start of step k
 *     (CU:1627 increments i before save1 records slot i-1);
 *   - storage in T; for T = float the literals 27.0, 24.0, 2.0, 0.50 and the "/1.0" promote the
 *     same sub-expressions to double as in the reference (CPP:627-836), and every other
 *     operation is float, so the float oracle rounds where the reference rounds.
 *
 * Source kinds: the reference's pressure wavelet added to p (CU:90-111), and a vertical body
 * force added to the v_y-row stress divergence (no reference counterpart).  The wavelet is
 * tabulated in FP64 (the reference evaluates it with float a, t, t0 on the device; the
 * difference is a relative 1e-7 in the wavelet).
 *
 * How each function is pinned (tests/test_oracle_pins.py):
 *   cpml_axis        -- reflection integral exp(-2/c * int d) = Rcoef; steady-state memory
 *                       variable closed form; inner edges; FP64 recomputation of a/b;
 *   material_biot    -- closed-form values of the brine and gas media; Gassmann identity
 *                       (K_fr + alpha^2 M equals Gassmann's saturated modulus);
 *   material_ref     -- reference-route identities alpha = phi, M = K_f/phi, lambda_c formula;
 *   wavelet          -- closed-form values, zero crossings and peaks;
 *   Solver::step     -- one-step polynomial exactness of every stencil and of the Biot
 *                       inverse-mass coupling; reciprocity in a random heterogeneous medium;
 *                       mirror symmetry; fast-P / S / slow-P wave speeds against Biot's
 *                       closed-form dispersion relation; linearity; zero-in zero-out;
 *   record           -- gather equals snapshot at the receiver point.
 */
#include "poe2d_oracle.h"

#include <cmath>
#include <cstring>
#include <new>
#include <string>
#include <type_traits>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

/* The reference's PI macro is a float literal (CPP:9). */
const float kRefPiF = 3.1415926535897932f;

/* ---------------------------------------------------------------------------------------------
 * C-PML profile for one axis.  Restates CPP:351-529 (= CU:985-1159).  For R = float every
 * variable is float and the double literals promote exactly the sub-expressions they promote
 * in the reference; std::pow/std::exp/std::log resolve to the float overloads as there.
 * ------------------------------------------------------------------------------------------- */
template <class R>
void cpml_axis(int axis, int n, R delta, R dt, int npml, R npower, R rcoef, R kmax, R alpha_max,
               R cpmax, R *a, R *b, R *K, R *ah, R *bh, R *Kh)
{
    const R thickness = npml * delta;
    /* d0 = -(N+1) c ln(R) / (2L): CPP:368-369 */
    const R d0 = -(npower + 1.0) * cpmax * std::log(rcoef) / (2.0 * thickness);
    const int NX = n - 1;
    /* layer origins: CPP:410-411 (x) and CPP:469-470 (z) */
    const R lo = axis == 0 ? R(thickness + (3 * delta)) : thickness;
    const R hi = (NX - 2) * delta - thickness;
    for (int i = 0; i < n; ++i) {
        R dg = 0, Kg = 1, alg = 0, dh = 0, Khv = 1, alh = 0;
        const R val = delta * (R)i;
        R s = lo - val; /* low layer (left / top), grid points */
        if (s >= 0.0) {
            R nrm = s / thickness;
            dg = d0 * std::pow(nrm, npower);
            Kg = 1.0 + (kmax - 1.0) * std::pow(nrm, npower);
            alg = axis == 0 ? R(alpha_max * (1.0 - nrm))
                            : R(alpha_max * (1.0 - nrm) + 0.10 * alpha_max);
        }
        s = lo - (val + delta / 2.0); /* low layer, half-grid points */
        if (s >= 0.0) {
            R nrm = s / thickness;
            dh = d0 * std::pow(nrm, npower);
            Khv = 1.0 + (kmax - 1.0) * std::pow(nrm, npower);
            alh = axis == 0 ? R(alpha_max * (1.0 - nrm))
                            : R(alpha_max * (1.0 - nrm) + 0.10 * alpha_max);
        }
        s = val - hi; /* high layer (right / bottom) overrides, as it is assigned second */
        if (s >= 0.0) {
            R nrm = s / thickness;
            dg = d0 * std::pow(nrm, npower);
            Kg = 1.0 + (kmax - 1.0) * std::pow(nrm, npower);
            alg = alpha_max * (1.0 - nrm);
        }
        s = (val + delta / 2.0) - hi;
        if (s >= 0.0) {
            R nrm = s / thickness;
            dh = d0 * std::pow(nrm, npower);
            Khv = 1.0 + (kmax - 1.0) * std::pow(nrm, npower);
            alh = alpha_max * (1.0 - nrm);
        }
        if (alg < 0.0) alg = 0;
        if (alh < 0.0) alh = 0;
        const R bg = std::exp(-(dg / Kg + alg) * dt);
        const R bhv = std::exp(-(dh / Khv + alh) * dt);
        R ag = 0, ahv = 0;
        if (std::fabs(dg) > 1.e-6) ag = dg * (bg - 1.0) / (Kg * (dg + Kg * alg));
        if (std::fabs(dh) > 1.e-6) ahv = dh * (bhv - 1.0) / (Khv * (dh + Khv * alh));
        a[i] = ag; b[i] = bg; K[i] = Kg;
        ah[i] = ahv; bh[i] = bhv; Kh[i] = Khv;
    }
}

/* 4th-order staggered difference (27 f1 - 27 f0 - f2 + fm)/h/24 in the reference's order, with
 * f0..f1 the pair straddling the output point, f2 the far point past f1 and fm the far point
 * before f0 (CPP:627 and every other derivative of the loop). */
template <class T>
inline T d4(T f1, T f0, T f2, T fm, T h)
{
    return (T)((27.0 * f1 - 27.0 * f0 - f2 + fm) / h / 24.0);
}

double wavelet_shape(int kind, double f0, double t0, double t)
{
    const double pi = 3.14159265358979323846;
    const double a = pi * pi * f0 * f0;
    const double tau = t - t0;
    if (kind == POE2D_SRC_PRESSURE_DGAUSS) return 2.0 * a * tau * std::exp(-a * tau * tau);
    return (1.0 - 2.0 * a * tau * tau) * std::exp(-a * tau * tau);
}

struct Base {
    virtual ~Base() {}
};

template <class T>
struct Solver : Base {
    typedef T Real;
    int NX = 0, NZ = 0;
    size_t P = 0;
    T DX = 0, DZ = 0, DT = 0;
    int nth = 1;
    bool avg = false; /* POE2D_MATERIAL_AVERAGED */
    std::vector<T> rho, rhof, rsm, rmu, lc, rbM, alph, eok;
    std::vector<T> vx, vz, wx, wz, g11, g22, xi1, xi2, sig2, sxx, szz, sxz;
    std::vector<T> psi[12];
    std::vector<T> cx[6], cz[6]; /* a, b, K, a_half, b_half, K_half */
    /* source */
    bool have_src = false;
    int src_kind = 0, is = 0, js = 0;
    double amp = 0, t0 = 0, tcut = 0, f0 = 0;
    std::vector<double> user_w;
    bool have_user_w = false;
    /* receivers */
    int nrec = 0, dec = 1, nsamp = 0;
    unsigned mask = 0;
    std::vector<int> rix, riz;
    std::vector<T> traces; /* [comp][rec][slot] */
    /* reference-structured scans */
    bool scans = false;
    int panj = 0, panj2 = 0, a0 = 0, a1 = 0;
    std::vector<T> vsp_vz, vsp_vx, hor_vz, hor_vx, pt0, pt2, pt3, pt4;
    int step = 0;

    size_t npts() const { return P * (size_t)(NZ + 1); }

    double source_value(int n) const
    {
        if (!have_src) return 0.0;
        const double t = n * (double)DT;
        if (tcut > 0 && t >= tcut) return 0.0;
        if (have_user_w) return n < (int)user_w.size() ? amp * user_w[n] : 0.0;
        return amp * wavelet_shape(src_kind, f0, t0, t);
    }

    int ncomp() const { return ((mask & 1) != 0) + ((mask & 2) != 0) + ((mask & 4) != 0); }

    void record(int n)
    {
        if (nrec == 0 || n % dec != 0) return;
        const int slot = n / dec;
        if (slot >= nsamp) return;
        int c = 0;
        const std::vector<T> *src[3] = {&vx, &vz, &sig2};
        for (int k = 0; k < 3; ++k) {
            if (!(mask & (1u << k))) continue;
            for (int r = 0; r < nrec; ++r)
                traces[((size_t)c * nrec + r) * nsamp + slot] =
                    (*src[k])[(size_t)rix[r] + (size_t)riz[r] * P];
            ++c;
        }
    }

    /* The three full-grid receiver scans of CPP:657-739 (reference-structured baseline only). */
    void reference_scans(int istep)
    {
        const int length2 = (NZ + 1) / panj, length = (NX + 1) / panj2;
        const int nslot = (int)pt0.size();
        for (int y = 0; y < NZ; y++)
            for (int x = 0; x < NX; x++) {
                size_t o = x + (size_t)y * P;
                if (x == is && istep > 0 && (y + 1) % length2 == 0 && istep <= nslot) {
                    size_t k = (size_t)(y / length2) + (size_t)(istep - 1) * panj;
                    vsp_vz[k] = vz[o]; vsp_vx[k] = vx[o];
                }
            }
        for (int y = 0; y < NZ; y++)
            for (int x = 0; x < NX; x++) {
                size_t o = x + (size_t)y * P;
                if (y == 10 && istep > 0 && (x + 1) % length == 0 && istep <= nslot) {
                    size_t k = (size_t)(x / length) + (size_t)(istep - 1) * panj2;
                    hor_vz[k] = vz[o]; hor_vx[k] = vx[o];
                }
            }
        for (int y = 0; y < NZ; y++)
            for (int x = 0; x < NX; x++) {
                size_t o = x + (size_t)y * P;
                if (istep < 1 || istep > nslot) continue;
                if (y == 50 && x == (NX + 1) / 2 - 1) pt0[istep - 1] = vz[o];
                if (y == a0 && x == (NX + 1) / 2 - 1) pt2[istep - 1] = vz[o];
                if (y == a1 && x == (NX + 1) / 2 - 1) pt3[istep - 1] = vz[o];
                if (y == js && x == (NX + 1) / 2 - 1) pt4[istep - 1] = vz[o];
            }
    }

    void step_once()
    {
        const int n = step;
        const T DXl = DX, DZl = DZ, DTl = DT;
        const size_t Pl = P;
        const int nx = NX, nz = NZ;
        const bool psrc = have_src && src_kind == POE2D_SRC_PRESSURE_DGAUSS;
        const bool fsrc = have_src && src_kind == POE2D_SRC_FORCE_Y_RICKER;
        const T S = (T)source_value(n);
        const int sx = is, sz = js;
        T *p1 = psi[0].data(), *p2 = psi[1].data(), *p3 = psi[2].data(), *p4 = psi[3].data();
        T *p5 = psi[4].data(), *p6 = psi[5].data(), *p7 = psi[6].data(), *p8 = psi[7].data();
        T *p9 = psi[8].data(), *p10 = psi[9].data(), *p11 = psi[10].data(), *p12 = psi[11].data();
        const T *ax = cx[0].data(), *bx = cx[1].data(), *Kx = cx[2].data();
        const T *axh = cx[3].data(), *bxh = cx[4].data(), *Kxh = cx[5].data();
        const T *az = cz[0].data(), *bz = cz[1].data(), *Kz = cz[2].data();
        const T *azh = cz[3].data(), *bzh = cz[4].data(), *Kzh = cz[5].data();
        T *VX = vx.data(), *VZ = vz.data(), *WX = wx.data(), *WZ = wz.data();
        T *G11 = g11.data(), *G22 = g22.data(), *XI1 = xi1.data(), *XI2 = xi2.data();
        T *SIG2 = sig2.data(), *SXX = sxx.data(), *SZZ = szz.data(), *SXZ = sxz.data();
        const T *RHO = rho.data(), *RHOF = rhof.data(), *RSM = rsm.data(), *RMU = rmu.data();
        const T *LC = lc.data(), *RBM = rbM.data(), *AL = alph.data(), *EOK = eok.data();
        const int nt = nth;

        /* stress1: strains, fluid content, pore pressure (+ pressure source).  x in [1,NX),
         * z in [2,NZ-1): CPP:618-655 / CU:49-112. */
#pragma omp parallel for schedule(static) num_threads(nt) if (nt > 1)
        for (int z = 2; z < nz - 1; z++)
            for (int x = 1; x < nx; x++) {
                const size_t o = x + z * Pl;
                T dvx = d4(VX[o + 1], VX[o], VX[o + 2], VX[o - 1], DXl);
                T dvz = d4(VZ[o], VZ[o - Pl], VZ[o + Pl], VZ[o - 2 * Pl], DZl);
                p1[o] = bxh[x] * p1[o] + axh[x] * dvx;
                p2[o] = bz[z] * p2[o] + az[z] * dvz;
                G11[o] = G11[o] + DTl * (dvx / Kxh[x] + p1[o]);
                G22[o] = G22[o] + DTl * (dvz / Kz[z] + p2[o]);
                T dwx = d4(WX[o + 1], WX[o], WX[o + 2], WX[o - 1], DXl);
                T dwz = d4(WZ[o], WZ[o - Pl], WZ[o + Pl], WZ[o - 2 * Pl], DZl);
                p3[o] = bxh[x] * p3[o] + axh[x] * dwx;
                p4[o] = bz[z] * p4[o] + az[z] * dwz;
                XI1[o] = XI1[o] - (dwx / Kxh[x] + p3[o]) * DTl;
                XI2[o] = XI2[o] - (dwz / Kz[z] + p4[o]) * DTl;
                SIG2[o] = -AL[o] * RBM[o] * (G11[o] + G22[o]) + RBM[o] * (XI1[o] + XI2[o]);
                if (psrc && x == sx && z == sz) SIG2[o] = SIG2[o] + S;
            }

        /* gather of v^n and p^(n+1/2) (CU:114-201 slot convention) */
        record(n);
        if (scans) reference_scans(n + 1);

        /* stress2: shear stress.  x in [2,NX), z in [1,NZ-1): CPP:741-761 / CU:202-229.
         * mu is taken at (x, z+1) and the z derivative is divided by K_z, both as written. */
#pragma omp parallel for schedule(static) num_threads(nt) if (nt > 1)
        for (int z = 1; z < nz - 1; z++)
            for (int x = 2; x < nx; x++) {
                const size_t o = x + z * Pl;
                /* averaged: the harmonic mean the reference computes at CU:217 (in its double
                 * arithmetic) before overwriting it with mu of (x, z+1) at CU:218 */
                const T c33 = avg ? (T)(2.0 / (1.0 / (double)RMU[o] + 1.0 / (double)RMU[o + Pl])) : RMU[o + Pl];
                T ddx = d4(VZ[o], VZ[o - 1], VZ[o + 1], VZ[o - 2], DXl);
                T ddz = d4(VX[o + Pl], VX[o], VX[o + 2 * Pl], VX[o - Pl], DZl);
                p5[o] = bx[x] * p5[o] + ax[x] * ddx;
                p6[o] = bzh[z] * p6[o] + azh[z] * ddz;
                SXZ[o] = SXZ[o] + c33 / 1.0 * (ddx / Kx[x] + p5[o] + ddz / Kz[z] + p6[o]) * DTl;
            }

        /* stress3: normal stresses.  x in [1,NX-1), z in [2,NZ): CPP:763-771 / CU:231-244. */
#pragma omp parallel for schedule(static) num_threads(nt) if (nt > 1)
        for (int z = 2; z < nz; z++)
            for (int x = 1; x < nx - 1; x++) {
                const size_t o = x + z * Pl;
                SXX[o] = (LC[o] + 2.0 * RMU[o]) * G11[o] + LC[o] * G22[o] - AL[o] * SIG2[o];
                SZZ[o] = LC[o] * G11[o] + (LC[o] + 2.0 * RMU[o]) * G22[o] - AL[o] * SIG2[o];
            }

        /* vx: solid and relative-fluid x velocities, Crank-Nicolson viscous coupling.
         * x in [2,NX), z in [2,NZ): CPP:773-802 / CU:279-314. */
#pragma omp parallel for schedule(static) num_threads(nt) if (nt > 1)
        for (int z = 2; z < nz; z++)
            for (int x = 2; x < nx; x++) {
                const size_t o = x + z * Pl;
                const T co = (RHO[o] * RSM[o] - RHOF[o] * RHOF[o]) / DTl;
                const T c1 = co + RHO[o] * EOK[o] * 0.50;
                const T c2 = co - RHO[o] * EOK[o] * 0.50;
                T wold = WX[o];
                T e1 = d4(SXX[o], SXX[o - 1], SXX[o + 1], SXX[o - 2], DXl);
                T e2 = d4(SIG2[o], SIG2[o - 1], SIG2[o + 1], SIG2[o - 2], DXl);
                T e3 = d4(SXZ[o], SXZ[o - Pl], SXZ[o + Pl], SXZ[o - 2 * Pl], DZl);
                p7[o] = bx[x] * p7[o] + ax[x] * e1;
                p8[o] = bx[x] * p8[o] + ax[x] * e2;
                p9[o] = bz[z] * p9[o] + az[z] * e3;
                const T divs = e1 / Kx[x] + p7[o] + e3 / Kz[z] + p9[o];
                const T gp = e2 / Kx[x] + p8[o];
                WX[o] = (c2 * WX[o] + (-RHOF[o] * divs - RHO[o] * gp)) / c1;
                T wavg = (wold + WX[o]) * 0.50;
                VX[o] = VX[o] + (RSM[o] * divs + RHOF[o] * gp + RHOF[o] * EOK[o] * wavg) / co;
            }

        /* vz: as vx with the material of point (x, z+1), x in [1,NX-1), z in [1,NZ-1):
         * CPP:804-838 / CU:316-355.  The vertical body force joins the stress divergence. */
#pragma omp parallel for schedule(static) num_threads(nt) if (nt > 1)
        for (int z = 1; z < nz - 1; z++)
            for (int x = 1; x < nx - 1; x++) {
                const size_t o = x + z * Pl;
                const size_t t = o + Pl;
                /* material of (x, z+1) as written (CU:330-333); averaged: means of (x, z) and (x, z+1) */
                const T r = avg ? (T)(0.50 * (RHO[o] + RHO[t])) : RHO[t];
                const T sm = avg ? (T)(0.50 * (RSM[o] + RSM[t])) : RSM[t];
                const T rf = avg ? (T)(0.50 * (RHOF[o] + RHOF[t])) : RHOF[t];
                const T ek = avg ? (T)(0.50 * (EOK[o] + EOK[t])) : EOK[t];
                const T co = (r * sm - rf * rf) / DTl;
                const T c1 = co + r * ek * 0.50;
                const T c2 = co - r * ek * 0.50;
                T wold = WZ[o];
                T e1 = d4(SXZ[o + 1], SXZ[o], SXZ[o + 2], SXZ[o - 1], DXl);
                T e2 = d4(SZZ[o + Pl], SZZ[o], SZZ[o + 2 * Pl], SZZ[o - Pl], DZl);
                T e3 = d4(SIG2[o + Pl], SIG2[o], SIG2[o + 2 * Pl], SIG2[o - Pl], DZl);
                p10[o] = bxh[x] * p10[o] + axh[x] * e1;
                p11[o] = bzh[z] * p11[o] + azh[z] * e2;
                p12[o] = bzh[z] * p12[o] + azh[z] * e3;
                T divs = e1 / Kxh[x] + p10[o] + e2 / Kzh[z] + p11[o];
                if (fsrc && x == sx && z == sz) divs = divs + S;
                const T gp = e3 / Kzh[z] + p12[o];
                WZ[o] = (c2 * WZ[o] + (-rf * divs - r * gp)) / c1;
                T wavg = (wold + WZ[o]) * 0.50;
                VZ[o] = VZ[o] + (sm * divs + rf * gp + rf * ek * wavg) / co;
            }
        ++step;
    }
};

thread_local std::string g_err;

int fail(int code, const std::string &msg)
{
    g_err = msg;
    return code;
}

double coeff_at(const poe2d_coeffs *c, const void *arr, size_t idx)
{
    if (c->dtype == POE2D_F64) return ((const double *)arr)[idx];
    return ((const float *)arr)[idx];
}

template <class T>
int build(const poe2d_grid *g, const poe2d_coeffs *c, Solver<T> *s)
{
    s->NX = g->nx; s->NZ = g->nz;
    s->P = (size_t)g->nx + 1;
    s->DX = (T)g->dx; s->DZ = (T)g->dz; s->DT = (T)g->dt;
    s->f0 = g->f0;
    const size_t N = s->npts();
    const size_t pitch = c->pitch ? c->pitch : (size_t)g->nx + 1;
    std::vector<T> *mats[8] = {&s->rho, &s->rhof, &s->rsm, &s->rmu, &s->lc, &s->rbM, &s->alph, &s->eok};
    const void *srcs[8] = {c->rho, c->rhof, c->m, c->mu, c->lambdac, c->M, c->alpha, c->b};
    for (int k = 0; k < 8; ++k) {
        if (!srcs[k]) return fail(POE2D_E_INVALID, "oracle: null coefficient array");
        mats[k]->assign(N, 0);
        for (int z = 0; z <= g->nz; ++z)
            for (int x = 0; x <= g->nx; ++x)
                (*mats[k])[x + z * s->P] = (T)coeff_at(c, srcs[k], x + z * pitch);
    }
    /* stress law: drained replaces lambda_c by lambda_c - alpha^2 M (computed in double) */
    auto lambda_at = [&](size_t i) {
        const double lc = coeff_at(c, c->lambdac, i), al = coeff_at(c, c->alpha, i), M = coeff_at(c, c->M, i);
        return g->stress_law == POE2D_STRESS_DRAINED ? lc - al * al * M : lc;
    };
    for (int z = 0; z <= g->nz; ++z)
        for (int x = 0; x <= g->nx; ++x) s->lc[x + z * s->P] = (T)lambda_at(x + z * pitch);
    s->avg = g->staggered_material == POE2D_MATERIAL_AVERAGED;
    double cpmax = g->cp_max;
    if (cpmax <= 0) {
        cpmax = 0;
        for (int z = 0; z <= g->nz; ++z)
            for (int x = 0; x <= g->nx; ++x) {
                size_t i = x + z * pitch;
                double al = coeff_at(c, c->alpha, i), M = coeff_at(c, c->M, i);
                double H = lambda_at(i) + 2 * coeff_at(c, c->mu, i) + al * al * M;
                double v = std::sqrt(H / coeff_at(c, c->rho, i));
                if (v > cpmax) cpmax = v;
            }
    }
    if (g->check_cfl && cpmax * g->dt / g->dx > 6.0 / 7.0 / std::sqrt(2.0))
        return fail(POE2D_E_CFL, "oracle: time step is too large, simulation will be unstable");
    std::vector<T> *st[12] = {&s->vx, &s->vz, &s->wx, &s->wz, &s->g11, &s->g22,
                              &s->xi1, &s->xi2, &s->sig2, &s->sxx, &s->szz, &s->sxz};
    for (auto *v : st) v->assign(N, 0);
    for (auto &v : s->psi) v.assign(N, 0);
    for (int k = 0; k < 6; ++k) {
        s->cx[k].assign(g->nx + 1, 0);
        s->cz[k].assign(g->nz + 1, 0);
    }
    double amax = g->alpha_max;
    T amaxT;
    if (amax > 0) amaxT = (T)amax;
    else if (sizeof(T) == 4) amaxT = (T)(2.e0 * kRefPiF * ((float)g->f0 / 2.e0)); /* CPP:298 */
    else amaxT = (T)(3.14159265358979323846 * g->f0);
    for (int axis = 0; axis < 2; ++axis) {
        std::vector<T> *t = axis == 0 ? s->cx : s->cz;
        int n = axis == 0 ? g->nx + 1 : g->nz + 1;
        T d = (T)(axis == 0 ? g->dx : g->dz);
        cpml_axis<T>(axis, n, d, (T)g->dt, g->npml, (T)g->npower, (T)g->rcoef, (T)g->kmax, amaxT,
                     (T)cpmax, t[0].data(), t[1].data(), t[2].data(), t[3].data(), t[4].data(),
                     t[5].data());
    }
    return 0;
}

int check_grid(const poe2d_grid *g)
{
    if (!g) return fail(POE2D_E_INVALID, "oracle: null grid");
    if (g->nx < 8 || g->nz < 8) return fail(POE2D_E_INVALID, "oracle: grid too small");
    if (!(g->dx > 0) || !(g->dz > 0) || !(g->dt > 0)) return fail(POE2D_E_INVALID, "oracle: bad spacing");
    if (g->npml < 1) return fail(POE2D_E_INVALID, "oracle: npml must be >= 1");
    if (g->npower < 1) return fail(POE2D_E_INVALID, "oracle: npower must be >= 1");
    if (g->precision != POE2D_F32 && g->precision != POE2D_F64)
        return fail(POE2D_E_INVALID, "oracle: precision must be 32 or 64");
    if (g->stress_law != POE2D_STRESS_AS_WRITTEN && g->stress_law != POE2D_STRESS_DRAINED)
        return fail(POE2D_E_INVALID, "oracle: unknown stress law");
    if (g->staggered_material != POE2D_MATERIAL_AS_WRITTEN && g->staggered_material != POE2D_MATERIAL_AVERAGED)
        return fail(POE2D_E_INVALID, "oracle: unknown staggered material setting");
    return 0;
}

} // namespace

struct poe2d_oracle {
    int precision;
    Base *impl;
};

#define DISPATCH(o, CALL)                                                                        \
    do {                                                                                         \
        if ((o)->precision == POE2D_F32) {                                                      \
            Solver<float> *s = static_cast<Solver<float> *>((o)->impl);                          \
            CALL;                                                                                \
        } else {                                                                                 \
            Solver<double> *s = static_cast<Solver<double> *>((o)->impl);                        \
            CALL;                                                                                \
        }                                                                                        \
    } while (0)

template <class T>
static int field_ptr(Solver<T> *s, int field, std::vector<T> **v)
{
    switch (field) {
    case POE2D_FIELD_VX: *v = &s->vx; return 0;
    case POE2D_FIELD_VY: *v = &s->vz; return 0;
    case POE2D_FIELD_WX: *v = &s->wx; return 0;
    case POE2D_FIELD_WY: *v = &s->wz; return 0;
    case POE2D_FIELD_EXX: *v = &s->g11; return 0;
    case POE2D_FIELD_EYY: *v = &s->g22; return 0;
    case POE2D_FIELD_ZETA: *v = &s->xi1; return 0;
    case POE2D_FIELD_SXY: *v = &s->sxz; return 0;
    case POE2D_FIELD_P: *v = &s->sig2; return 0;
    case POE2D_FIELD_SXX: *v = &s->sxx; return 0;
    case POE2D_FIELD_SYY: *v = &s->szz; return 0;
    default: return fail(POE2D_E_INVALID, "oracle: unknown field");
    }
}

extern "C" {

int poe2d_oracle_create(const poe2d_grid *g, const poe2d_coeffs *c, poe2d_oracle **out)
{
    if (!out || !c) return fail(POE2D_E_INVALID, "oracle: null argument");
    *out = nullptr;
    int rc = check_grid(g);
    if (rc) return rc;
    poe2d_oracle *o = new (std::nothrow) poe2d_oracle;
    if (!o) return fail(POE2D_E_NOMEM, "oracle: out of memory");
    o->precision = g->precision;
    try {
        if (g->precision == POE2D_F32) {
            auto *s = new Solver<float>;
            o->impl = s;
            rc = build<float>(g, c, s);
        } else {
            auto *s = new Solver<double>;
            o->impl = s;
            rc = build<double>(g, c, s);
        }
    } catch (const std::bad_alloc &) {
        rc = fail(POE2D_E_NOMEM, "oracle: out of memory");
    }
    if (rc) {
        poe2d_oracle_destroy(o);
        return rc;
    }
    *out = o;
    return 0;
}

int poe2d_oracle_set_source(poe2d_oracle *o, const poe2d_source *src)
{
    if (!o || !src) return fail(POE2D_E_INVALID, "oracle: null argument");
    if (src->kind != POE2D_SRC_PRESSURE_DGAUSS && src->kind != POE2D_SRC_FORCE_Y_RICKER)
        return fail(POE2D_E_INVALID, "oracle: unknown source kind");
    DISPATCH(o, {
        bool ok = src->kind == POE2D_SRC_PRESSURE_DGAUSS
                      ? (src->ix >= 1 && src->ix < s->NX && src->iz >= 2 && src->iz < s->NZ - 1)
                      : (src->ix >= 1 && src->ix < s->NX - 1 && src->iz >= 1 && src->iz < s->NZ - 1);
        if (!ok) return fail(POE2D_E_INVALID, "oracle: source outside the updated region");
        s->have_src = true;
        s->src_kind = src->kind;
        s->is = src->ix;
        s->js = src->iz;
        s->amp = src->amp;
        s->t0 = src->t0 > 0 ? src->t0 : 1.0 / s->f0;
        s->tcut = src->t_cut;
    });
    return 0;
}

int poe2d_oracle_set_wavelet(poe2d_oracle *o, const double *w, int n)
{
    if (!o || (n > 0 && !w) || n < 0) return fail(POE2D_E_INVALID, "oracle: bad wavelet");
    DISPATCH(o, {
        s->user_w.assign(w, w + n);
        s->have_user_w = true;
    });
    return 0;
}

int poe2d_oracle_set_receivers(poe2d_oracle *o, const poe2d_receivers *r)
{
    if (!o || !r || r->n < 0 || r->decimate < 1 || r->nsamp < 0 || (r->comp_mask & ~7u))
        return fail(POE2D_E_INVALID, "oracle: bad receivers");
    DISPATCH(o, {
        for (int k = 0; k < r->n; ++k)
            if (r->ix[k] < 0 || r->ix[k] > s->NX || r->iz[k] < 0 || r->iz[k] > s->NZ)
                return fail(POE2D_E_INVALID, "oracle: receiver outside the grid");
        s->nrec = r->n;
        s->rix.assign(r->ix, r->ix + r->n);
        s->riz.assign(r->iz, r->iz + r->n);
        s->mask = r->comp_mask;
        s->dec = r->decimate;
        s->nsamp = r->nsamp;
        s->traces.assign((size_t)s->ncomp() * r->n * r->nsamp, 0);
    });
    return 0;
}

int poe2d_oracle_set_threads(poe2d_oracle *o, int nthreads)
{
    if (!o || nthreads < 1) return fail(POE2D_E_INVALID, "oracle: bad thread count");
    DISPATCH(o, { s->nth = nthreads; });
    return 0;
}

int poe2d_oracle_set_reference_scans(poe2d_oracle *o, int panj, int panj2, int a0_pos, int a1_pos)
{
    if (!o || panj < 1 || panj2 < 1) return fail(POE2D_E_INVALID, "oracle: bad scan layout");
    DISPATCH(o, {
        if ((s->NZ + 1) / panj < 1 || (s->NX + 1) / panj2 < 1)
            return fail(POE2D_E_INVALID, "oracle: scan layout larger than the grid");
        s->scans = true;
        s->panj = panj; s->panj2 = panj2; s->a0 = a0_pos; s->a1 = a1_pos;
        const size_t nslot = 1 << 16;
        s->vsp_vz.assign(panj * nslot, 0); s->vsp_vx.assign(panj * nslot, 0);
        s->hor_vz.assign(panj2 * nslot, 0); s->hor_vx.assign(panj2 * nslot, 0);
        s->pt0.assign(nslot, 0); s->pt2.assign(nslot, 0); s->pt3.assign(nslot, 0); s->pt4.assign(nslot, 0);
    });
    return 0;
}

int poe2d_oracle_reference_scan(poe2d_oracle *o, int which, int index, double *out, int n)
{
    if (!o || !out || which < 0 || which > 7 || index < 0 || n < 0) return fail(POE2D_E_INVALID, "oracle: bad scan query");
    DISPATCH(o, {
        if (!s->scans) return fail(POE2D_E_STATE, "oracle: reference scans are off");
        typedef typename std::remove_pointer<decltype(s)>::type::Real R;
        const std::vector<R> *src[8];
        src[0] = &s->vsp_vz; src[1] = &s->vsp_vx; src[2] = &s->hor_vz; src[3] = &s->hor_vx;
        src[4] = &s->pt0; src[5] = &s->pt2; src[6] = &s->pt3; src[7] = &s->pt4;
        const int width = which < 2 ? s->panj : which < 4 ? s->panj2 : 1;
        if (index >= width || (size_t)n * width > src[which]->size())
            return fail(POE2D_E_INVALID, "oracle: scan index or sample count out of range");
        for (int k = 0; k < n; ++k) out[k] = (double)(*src[which])[(size_t)k * width + index];
    });
    return 0;
}

int poe2d_oracle_run(poe2d_oracle *o, int nsteps)
{
    if (!o || nsteps < 0) return fail(POE2D_E_INVALID, "oracle: bad step count");
    DISPATCH(o, {
        for (int k = 0; k < nsteps; ++k) s->step_once();
    });
    return 0;
}

int poe2d_oracle_step_count(poe2d_oracle *o, int *n)
{
    if (!o || !n) return fail(POE2D_E_INVALID, "oracle: null argument");
    DISPATCH(o, { *n = s->step; });
    return 0;
}

int poe2d_oracle_seismograms(poe2d_oracle *o, void *out, size_t bytes)
{
    if (!o || !out) return fail(POE2D_E_INVALID, "oracle: null argument");
    DISPATCH(o, {
        size_t need = s->traces.size() * sizeof(s->traces[0]);
        if (bytes < need) return fail(POE2D_E_INVALID, "oracle: output buffer too small");
        std::memcpy(out, s->traces.data(), need);
    });
    return 0;
}

int poe2d_oracle_snapshot(poe2d_oracle *o, int field, void *out, size_t bytes)
{
    if (!o || !out) return fail(POE2D_E_INVALID, "oracle: null argument");
    DISPATCH(o, {
        std::vector<decltype(s->DX)> *v = nullptr;
        if (field_ptr(s, field, &v)) return POE2D_E_INVALID;
        size_t need = v->size() * sizeof((*v)[0]);
        if (bytes < need) return fail(POE2D_E_INVALID, "oracle: output buffer too small");
        if (field == POE2D_FIELD_ZETA) {
            auto *dst = (decltype(s->DX) *)out;
            for (size_t i = 0; i < v->size(); ++i) dst[i] = s->xi1[i] + s->xi2[i];
        } else {
            std::memcpy(out, v->data(), need);
        }
    });
    return 0;
}

int poe2d_oracle_set_field(poe2d_oracle *o, int field, const void *in, size_t bytes)
{
    if (!o || !in) return fail(POE2D_E_INVALID, "oracle: null argument");
    if (field > POE2D_FIELD_SXY) return fail(POE2D_E_INVALID, "oracle: only state fields can be set");
    DISPATCH(o, {
        std::vector<decltype(s->DX)> *v = nullptr;
        if (field_ptr(s, field, &v)) return POE2D_E_INVALID;
        size_t need = v->size() * sizeof((*v)[0]);
        if (bytes < need) return fail(POE2D_E_INVALID, "oracle: input buffer too small");
        std::memcpy(v->data(), in, need);
        if (field == POE2D_FIELD_ZETA) std::fill(s->xi2.begin(), s->xi2.end(), 0);
    });
    return 0;
}

int poe2d_oracle_cpml_table(poe2d_oracle *o, int axis, int which, double *out, int n)
{
    if (!o || !out || axis < 0 || axis > 1 || which < 0 || which > 5)
        return fail(POE2D_E_INVALID, "oracle: bad table request");
    DISPATCH(o, {
        auto &t = axis == 0 ? s->cx[which] : s->cz[which];
        if (n < (int)t.size()) return fail(POE2D_E_INVALID, "oracle: output too small");
        for (size_t i = 0; i < t.size(); ++i) out[i] = (double)t[i];
    });
    return 0;
}

void poe2d_oracle_destroy(poe2d_oracle *o)
{
    if (!o) return;
    delete o->impl;
    delete o;
}

int poe2d_oracle_cpml_profile(int axis, int n, double delta, double dt, int npml, double npower,
                              double rcoef, double kmax, double alpha_max, double cp_max,
                              int precision, double *a, double *b, double *K, double *ah,
                              double *bh, double *Kh)
{
    if (n < 4 || npml < 1 || axis < 0 || axis > 1 || !a || !b || !K || !ah || !bh || !Kh)
        return fail(POE2D_E_INVALID, "oracle: bad profile request");
    if (precision == POE2D_F32) {
        std::vector<float> t[6];
        for (auto &v : t) v.assign(n, 0);
        cpml_axis<float>(axis, n, (float)delta, (float)dt, npml, (float)npower, (float)rcoef,
                         (float)kmax, (float)alpha_max, (float)cp_max, t[0].data(), t[1].data(),
                         t[2].data(), t[3].data(), t[4].data(), t[5].data());
        double *outs[6] = {a, b, K, ah, bh, Kh};
        for (int k = 0; k < 6; ++k)
            for (int i = 0; i < n; ++i) outs[k][i] = t[k][i];
    } else {
        cpml_axis<double>(axis, n, delta, dt, npml, npower, rcoef, kmax, alpha_max, cp_max, a, b,
                          K, ah, bh, Kh);
    }
    return 0;
}

int poe2d_oracle_material_biot(size_t n, const double *rho_s, const double *rho_f,
                               const double *phi, const double *kappa, const double *eta,
                               const double *tort, const double *K_s, const double *K_f,
                               const double *K_fr, const double *mu_fr, double *rho, double *rhof,
                               double *m, double *mu, double *lambdac, double *M, double *alpha,
                               double *b)
{
    for (size_t i = 0; i < n; ++i) {
        const double ph = phi[i], Kb = K_fr[i], Ks = K_s[i];
        rho[i] = (1.0 - ph) * rho_s[i] + ph * rho_f[i];
        rhof[i] = rho_f[i];
        mu[i] = mu_fr[i];
        m[i] = tort[i] * rho_f[i] / ph;
        b[i] = (eta[i] == 0.0 || kappa[i] == 0.0) ? 0.0 : eta[i] / kappa[i];
        alpha[i] = 1.0 - Kb / Ks;
        M[i] = 1.0 / (ph / K_f[i] + (alpha[i] - ph) / Ks);
        lambdac[i] = (Kb - 2.0 / 3.0 * mu[i]) + alpha[i] * alpha[i] * M[i];
    }
    return 0;
}

int poe2d_oracle_material_reference(size_t n, const float *cp, const float *cs,
                                    const float *rho_sat, const float *medium, int nmed,
                                    const float *kf_tab, const float *eta_tab,
                                    const float *kappa_tab, const float *phi_tab,
                                    const float *a_tab, const float *rhof_tab, float *rho,
                                    float *rhof, float *m, float *mu, float *lambdac, float *M,
                                    float *alpha, float *b)
{
    int bad = 0;
    for (size_t j = 0; j < n; ++j) {
        /* medium-table lookup: CPP:162-177 */
        float Kf = 0, eta = 0, kappa = 0, phi = 0, a = 0, rf = 0;
        bool found = false;
        for (int i = 0; i < nmed; ++i)
            if (medium[j] == i) {
                Kf = kf_tab[i]; eta = eta_tab[i]; kappa = kappa_tab[i];
                phi = phi_tab[i]; a = a_tab[i]; rf = rhof_tab[i];
                found = true;
            }
        if (!found) ++bad;
        /* derivation: CPP:182-205, float variables with double literals */
        const float crho = rho_sat[j];
        const float crmu = cs[j] * cs[j] * crho;
        const float Ksat = rho_sat[j] * (cp[j] * cp[j] - (4.0 / 3.0) * (cs[j] * cs[j]));
        const float nu = (3.0 * Ksat - 2.0 * crmu) / (2.0 * (3.0 * Ksat + crmu));
        const float Kb = ((2.0 * nu / (1.0 - 2.0 * nu)) + (2.0 / 3.0)) * crmu;
        const float Ks = Kb / (1.0 - phi);
        const float csm = a * rf / phi;
        float ceok = eta / kappa;
        const float cal = 1.0 - Kb / Ks;
        const float cM = 1.0 / (phi / Kf + (cal - phi) / Ks);
        if ((eta == 0.0) || (kappa == 0.0)) ceok = 0.0;
        const float lo = Kb - (2.0 / 3.0) * crmu;
        const float lcv = lo + cal * cal * cM;
        rho[j] = crho; rhof[j] = rf; m[j] = csm; mu[j] = crmu;
        lambdac[j] = lcv; M[j] = cM; alpha[j] = cal; b[j] = ceok;
    }
    if (bad) return fail(POE2D_E_INVALID, "oracle: medium index without a table entry");
    return 0;
}

int poe2d_oracle_wavelet(int kind, double f0, double t0, double amp, double t_cut, double dt,
                         int n, double *out)
{
    if (!out || n < 0 || !(f0 > 0)) return fail(POE2D_E_INVALID, "oracle: bad wavelet request");
    if (t0 <= 0) t0 = 1.0 / f0;
    for (int k = 0; k < n; ++k) {
        const double t = k * dt;
        out[k] = (t_cut > 0 && t >= t_cut) ? 0.0 : amp * wavelet_shape(kind, f0, t0, t);
    }
    return 0;
}

const char *poe2d_oracle_last_error(void) { return g_err.c_str(); }

} /* extern "C" */
