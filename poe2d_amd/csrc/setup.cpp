// Host set-up of the solver: C-PML profiles, material front ends, wavelets, error reporting.
//
// Built with g++ -ffp-contract=off: the FP32 C-PML profile must come out exactly as the
// reference's float code computes it (seismic_CPML_2D_poroelastic_withCUDA.cu:985-1159), so every
// multiply and add stays separately rounded and the float libm overloads are used where the
// reference calls pow/exp/log on float arguments.
#include "common.h"

#include <cmath>
#include <cstring>

namespace poe2d {

static thread_local std::string g_last_error;

poe2d_status fail(poe2d_status code, const std::string &msg)
{
    g_last_error = msg;
    return code;
}

double wavelet_shape(int kind, double f0, double t0, double t)
{
    const double pi = 3.14159265358979323846;
    const double a = pi * pi * f0 * f0;
    const double tau = t - t0;
    const double g = std::exp(-a * tau * tau);
    if (kind == POE2D_SRC_PRESSURE_DGAUSS) return 2.0 * a * tau * g;  // first derivative of a Gaussian
    return (1.0 - 2.0 * a * tau * tau) * g;                            // Ricker
}

double cp_max_from_coeffs(const poe2d_grid &g, const poe2d_coeffs &c)
{
    const size_t pitch = c.pitch ? c.pitch : (size_t)g.nx + 1;
    auto at = [&](const void *p, size_t i) -> double {
        return c.dtype == POE2D_F64 ? ((const double *)p)[i] : (double)((const float *)p)[i];
    };
    double vmax = 0.0;
    for (int z = 0; z <= g.nz; ++z)
        for (int x = 0; x <= g.nx; ++x) {
            const size_t i = (size_t)z * pitch + x;
            const double al = at(c.alpha, i), M = at(c.M, i);
            const double modulus = effective_lambda(g, at(c.lambdac, i), al, M) + 2.0 * at(c.mu, i) + al * al * M;
            const double v = std::sqrt(modulus / at(c.rho, i));
            if (v > vmax) vmax = v;
        }
    return vmax;
}

// One axis of C-PML profiles.  The layer sits between its origin and the grid edge; d rises as
// d0 (s/L)^N with d0 = -(N+1) c ln(R) / (2L), K = 1 + (Kmax-1)(s/L)^N, alpha falls linearly from
// alpha_max at the inner edge (the top layer adds 0.1 alpha_max), and the recursive-convolution
// coefficients are b = exp(-(d/K + alpha) dt), a = d (b-1) / (K (d + K alpha)) where |d| > 1e-6.
// Layer origins follow the reference: x low L+3dx, x high (NX-2)dx-L, z low L, z high (NZ-2)dz-L
// (seismic_CPML_2D_poroelastic_withCUDA.cu:1042-1043, 1101-1102); where both layers claim a point
// (tiny grids) the high layer wins, as it is assigned second there.
template <class R>
static void profile(int axis, int n, R delta, R dt, int npml, R npower, R rcoef, R kmax,
                    R alpha_max, R cpmax, double *oa, double *ob, double *oK, double *oah,
                    double *obh, double *oKh)
{
    const R L = npml * delta;
    const R d0 = -(npower + 1.0) * cpmax * std::log(rcoef) / (2.0 * L);
    const int last = n - 1;
    const R origin_low = axis == 0 ? R(L + (3 * delta)) : L;
    const R origin_high = (last - 2) * delta - L;
    struct Point { R d = 0, K = 1, alpha = 0; };
    auto layer = [&](R s, bool top, Point &pt) {
        if (!(s >= 0.0)) return;
        const R frac = s / L;
        pt.d = d0 * std::pow(frac, npower);
        pt.K = 1.0 + (kmax - 1.0) * std::pow(frac, npower);
        pt.alpha = top ? R(alpha_max * (1.0 - frac) + 0.10 * alpha_max) : R(alpha_max * (1.0 - frac));
    };
    for (int i = 0; i < n; ++i) {
        const R pos = delta * (R)i;
        Point g, h;
        layer(origin_low - pos, axis == 1, g);
        layer(origin_low - (pos + delta / 2.0), axis == 1, h);
        layer(pos - origin_high, false, g);
        layer((pos + delta / 2.0) - origin_high, false, h);
        for (Point *pt : {&g, &h})
            if (pt->alpha < 0.0) pt->alpha = 0;
        const R bg = std::exp(-(g.d / g.K + g.alpha) * dt);
        const R bh = std::exp(-(h.d / h.K + h.alpha) * dt);
        R ag = 0, ah = 0;
        if (std::fabs(g.d) > 1.e-6) ag = g.d * (bg - 1.0) / (g.K * (g.d + g.K * g.alpha));
        if (std::fabs(h.d) > 1.e-6) ah = h.d * (bh - 1.0) / (h.K * (h.d + h.K * h.alpha));
        oa[i] = ag; ob[i] = bg; oK[i] = g.K;
        oah[i] = ah; obh[i] = bh; oKh[i] = h.K;
    }
}

}  // namespace poe2d

using namespace poe2d;

extern "C" {

const char *poe2d_last_error(void) { return g_last_error.c_str(); }

int poe2d_abi_version(void) { return POE2D_ABI_VERSION; }

poe2d_status poe2d_cpml_profile(int axis, int n, double delta, double dt, int npml, double npower,
                                double rcoef, double kmax, double alpha_max, double cp_max,
                                int precision, double *a, double *b, double *K, double *ah,
                                double *bh, double *Kh)
{
    if (axis < 0 || axis > 1 || n < 4 || npml < 1 || !(delta > 0) || !(dt > 0) || !(cp_max > 0))
        return fail(POE2D_E_INVALID, "cpml_profile: bad geometry");
    if (npower < 1) return fail(POE2D_E_INVALID, "cpml_profile: npower must be >= 1");
    if (!a || !b || !K || !ah || !bh || !Kh) return fail(POE2D_E_INVALID, "cpml_profile: null output");
    if (precision == POE2D_F32)
        profile<float>(axis, n, (float)delta, (float)dt, npml, (float)npower, (float)rcoef,
                       (float)kmax, (float)alpha_max, (float)cp_max, a, b, K, ah, bh, Kh);
    else if (precision == POE2D_F64)
        profile<double>(axis, n, delta, dt, npml, npower, rcoef, kmax, alpha_max, cp_max, a, b, K,
                        ah, bh, Kh);
    else
        return fail(POE2D_E_INVALID, "cpml_profile: precision must be 32 or 64");
    return POE2D_OK;
}

poe2d_status poe2d_material_biot(size_t n, const double *rho_s, const double *rho_f,
                                 const double *phi, const double *kappa, const double *eta,
                                 const double *tort, const double *K_s, const double *K_f,
                                 const double *K_fr, const double *mu_fr, double *rho, double *rhof,
                                 double *m, double *mu, double *lambdac, double *M, double *alpha,
                                 double *b)
{
    const double *in[10] = {rho_s, rho_f, phi, kappa, eta, tort, K_s, K_f, K_fr, mu_fr};
    double *out[8] = {rho, rhof, m, mu, lambdac, M, alpha, b};
    for (auto p : in)
        if (!p) return fail(POE2D_E_INVALID, "material_biot: null input");
    for (auto p : out)
        if (!p) return fail(POE2D_E_INVALID, "material_biot: null output");
    for (size_t i = 0; i < n; ++i) {
        const double ph = phi[i];
        if (!(ph > 0 && ph < 1) || !(K_s[i] > 0) || !(K_f[i] > 0) || !(rho_f[i] > 0))
            return fail(POE2D_E_INVALID, "material_biot: need 0 < phi < 1 and positive K_s, K_f, rho_f");
        // Biot-Willis coefficient and Biot modulus with the frame modulus as K_b.
        const double aw = 1.0 - K_fr[i] / K_s[i];
        const double Mb = 1.0 / (ph / K_f[i] + (aw - ph) / K_s[i]);
        rho[i] = (1.0 - ph) * rho_s[i] + ph * rho_f[i];
        rhof[i] = rho_f[i];
        m[i] = tort[i] * rho_f[i] / ph;
        mu[i] = mu_fr[i];
        alpha[i] = aw;
        M[i] = Mb;
        lambdac[i] = (K_fr[i] - 2.0 / 3.0 * mu_fr[i]) + aw * aw * Mb;
        b[i] = (eta[i] == 0.0 || kappa[i] == 0.0) ? 0.0 : eta[i] / kappa[i];
    }
    return POE2D_OK;
}

// The reference's own medium-table front end, in its float arithmetic
// (seismic_CPML_2d_poroelastic.cpp:162-211).
poe2d_status poe2d_material_reference(size_t n, const float *cp, const float *cs,
                                      const float *rho_sat, const float *medium, int nmed,
                                      const float *kf_tab, const float *eta_tab,
                                      const float *kappa_tab, const float *phi_tab,
                                      const float *a_tab, const float *rhof_tab, float *rho,
                                      float *rhof, float *m, float *mu, float *lambdac, float *M,
                                      float *alpha, float *b)
{
    if (!cp || !cs || !rho_sat || !medium || nmed < 1 || !kf_tab || !eta_tab || !kappa_tab ||
        !phi_tab || !a_tab || !rhof_tab || !rho || !rhof || !m || !mu || !lambdac || !M || !alpha || !b)
        return fail(POE2D_E_INVALID, "material_reference: null argument");
    size_t unmatched = 0;
    for (size_t j = 0; j < n; ++j) {
        int hit = -1;
        for (int i = 0; i < nmed; ++i)
            if (medium[j] == i) hit = i;  // last matching table entry wins, as in the reference loop
        float Kf = 0, eta = 0, kappa = 0, phi = 0, tort = 0, rf = 0;
        if (hit >= 0) {
            Kf = kf_tab[hit]; eta = eta_tab[hit]; kappa = kappa_tab[hit];
            phi = phi_tab[hit]; tort = a_tab[hit]; rf = rhof_tab[hit];
        } else {
            ++unmatched;
        }
        const float r = rho_sat[j];
        const float shear = cs[j] * cs[j] * r;
        const float ksat = rho_sat[j] * (cp[j] * cp[j] - (4.0 / 3.0) * (cs[j] * cs[j]));
        const float poisson = (3.0 * ksat - 2.0 * shear) / (2.0 * (3.0 * ksat + shear));
        const float kb = ((2.0 * poisson / (1.0 - 2.0 * poisson)) + (2.0 / 3.0)) * shear;
        const float ks = kb / (1.0 - phi);
        const float aw = 1.0 - kb / ks;
        const float Mb = 1.0 / (phi / Kf + (aw - phi) / ks);
        float eok = eta / kappa;
        if ((eta == 0.0) || (kappa == 0.0)) eok = 0.0;
        const float lam_o = kb - (2.0 / 3.0) * shear;
        rho[j] = r;
        rhof[j] = rf;
        m[j] = tort * rf / phi;
        mu[j] = shear;
        alpha[j] = aw;
        M[j] = Mb;
        lambdac[j] = lam_o + aw * aw * Mb;
        b[j] = eok;
    }
    if (unmatched)
        return fail(POE2D_E_INVALID, "material_reference: " + std::to_string(unmatched) +
                                         " points have a medium index without a table entry");
    return POE2D_OK;
}

poe2d_status poe2d_wavelet(int kind, double f0, double t0, double amp, double t_cut, double dt,
                           int n, double *out)
{
    if (!out || n < 0 || !(f0 > 0) || !(dt > 0)) return fail(POE2D_E_INVALID, "wavelet: bad arguments");
    if (kind != POE2D_SRC_PRESSURE_DGAUSS && kind != POE2D_SRC_FORCE_Y_RICKER)
        return fail(POE2D_E_INVALID, "wavelet: unknown kind");
    if (t0 <= 0) t0 = 1.0 / f0;
    for (int k = 0; k < n; ++k) {
        const double t = k * dt;
        out[k] = (t_cut > 0 && t >= t_cut) ? 0.0 : amp * wavelet_shape(kind, f0, t0, t);
    }
    return POE2D_OK;
}

}  // extern "C"
