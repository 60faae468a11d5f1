// Random patchy-saturation model generator.
//
// The reference builds its patchy media in MATLAB (patchymodel.m): radius-8 disks on six
// deterministic lattices, painted into a medium-index map.  The workloads here need seeded
// random media instead: disks with uniform random centres and log-uniform radii are painted
// until the painted fraction of grid points reaches a target.  The lattice layouts of
// patchymodel.m are kept as a separate mode (poe2d_patchy_lattice).
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

#include "common.h"

namespace {

// SplitMix64 (Steele, Lea, Flood 2014): a 64-bit counter-based generator, used directly.
struct SplitMix64 {
    uint64_t s;
    uint64_t next()
    {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }  // [0, 1)
};

// Paints the points within distance r of (cx, cz) (metres); returns how many were newly painted.
// A point (x, z) is inside when (x dx - cx)^2 + (z dz - cz)^2 <= r^2, evaluated in double precision
// as written.  In a row the inside points form one run of x (the left-hand side falls, then rises,
// with x also after rounding), close to |x dx - cx| <= h = sqrt(r^2 - (z dz - cz)^2): points more
// than three columns inside that estimate (its ends truncated to integers, so within a column of
// the real ends) are inside without a doubt (the rounding moves the run's ends by far less than a
// column), and only the few columns around each end are tested.
size_t paint_disk(uint8_t *mask, int nx, int nz, double dx, double dz, double cx, double cz, double r)
{
    const int x0 = std::max(0, (int)std::ceil((cx - r) / dx));
    const int x1 = std::min(nx, (int)std::floor((cx + r) / dx));
    const int z0 = std::max(0, (int)std::ceil((cz - r) / dz));
    const int z1 = std::min(nz, (int)std::floor((cz + r) / dz));
    const double r2 = r * r, rdx = 1.0 / dx;
    size_t added = 0;
    for (int z = z0; z <= z1; ++z) {
        const double ddz = z * dz - cz;
        const double dz2 = ddz * ddz;
        uint8_t *row = mask + (size_t)z * (nx + 1);
        auto test = [&](int x) {
            const double ddx = x * dx - cx;
            const uint8_t in = ddx * ddx + dz2 <= r2;
            added += in & (row[x] ^ 1);
            row[x] |= in;
        };
        const double h = r2 > dz2 ? std::sqrt(r2 - dz2) : 0.0;
        const int ea = (int)((cx - h) * rdx), eb = (int)((cx + h) * rdx);
        const int lo = std::max(x0, ea - 3), hi = std::min(x1, eb + 3);  // every inside point
        const int ca = std::max(lo, ea + 3), cb = std::min(hi, eb - 3);  // surely inside
        if (ca > cb) {
            for (int x = lo; x <= hi; ++x) test(x);
            continue;
        }
        for (int x = lo; x < ca; ++x) test(x);
        size_t set = 0;
        for (int x = ca; x <= cb; ++x) set += row[x];
        added += (size_t)(cb - ca + 1) - set;
        std::memset(row + ca, 1, (size_t)(cb - ca + 1));
        for (int x = cb + 1; x <= hi; ++x) test(x);
    }
    return added;
}

// The mask of poe2d_patchy_mask (arguments already checked); false when the target cannot be
// reached within 10^8 disks.
bool paint_mask(int nx, int nz, double dx, double dz, double gas_fraction, uint64_t seed, double rmin, double rmax,
                uint8_t *mask, size_t *painted_out)
{
    const size_t npts = (size_t)(nx + 1) * (nz + 1);
    std::memset(mask, 0, npts);
    const size_t target = (size_t)std::ceil(gas_fraction * (double)npts);
    SplitMix64 rng{seed};
    size_t painted = 0;
    const double lr0 = std::log(rmin), lr1 = std::log(rmax);
    const double W = nx * dx, H = nz * dz;
    size_t draws = 0;
    while (painted < target) {
        const double cx = rng.uniform() * W;
        const double cz = rng.uniform() * H;
        const double r = std::exp(lr0 + (lr1 - lr0) * rng.uniform());
        painted += paint_disk(mask, nx, nz, dx, dz, cx, cz, r);
        if (++draws > 100000000ull) return false;
    }
    *painted_out = painted;
    return true;
}

}  // namespace

using poe2d::fail;

extern "C" {

// Fills mask[(nz+1)*(nx+1)] (x fastest) with 1 for patch points.  Disks are drawn with centres
// uniform over [0, nx dx] x [0, nz dz] and radii log-uniform in [rmin, rmax] metres until at least
// gas_fraction of all points are painted.  Same seed, same mask.
poe2d_status poe2d_patchy_mask(int nx, int nz, double dx, double dz, double gas_fraction,
                               uint64_t seed, double rmin, double rmax, uint8_t *mask,
                               double *realised)
{
    if (nx < 1 || nz < 1 || !(dx > 0) || !(dz > 0) || !mask)
        return fail(POE2D_E_INVALID, "patchy_mask: bad grid");
    if (!(gas_fraction >= 0 && gas_fraction < 1))
        return fail(POE2D_E_INVALID, "patchy_mask: gas_fraction must be in [0, 1)");
    if (!(rmin > 0) || !(rmax >= rmin)) return fail(POE2D_E_INVALID, "patchy_mask: need 0 < rmin <= rmax");
    size_t painted = 0;
    if (!paint_mask(nx, nz, dx, dz, gas_fraction, seed, rmin, rmax, mask, &painted))
        return fail(POE2D_E_INVALID, "patchy_mask: target fraction not reachable");
    if (realised) *realised = (double)painted / ((double)(nx + 1) * (nz + 1));
    return POE2D_OK;
}

// Many models' masks read at a few points (the surrogate's layout inputs): model k's mask is
// poe2d_patchy_mask(nx, nz, dx, dz, gas[k], seed, radius[k], radius[k]) (all zero when gas[k] is
// 0), and out[k * nidx + j] its value at flat index idx[j].  The models are dealt round-robin to
// `threads` host threads (0: one per hardware thread, at most 16), each painting into a mask of
// its own.
poe2d_status poe2d_patchy_samples(int nx, int nz, double dx, double dz, uint64_t seed, int nmodels,
                                  const double *gas, const double *radius, int nidx, const int64_t *idx,
                                  uint8_t *out, int threads)
{
    if (nx < 1 || nz < 1 || !(dx > 0) || !(dz > 0) || nmodels < 0 || nidx < 0 ||
        (nmodels > 0 && (!gas || !radius || !out)) || (nidx > 0 && !idx))
        return fail(POE2D_E_INVALID, "patchy_samples: bad arguments");
    const int64_t npts = (int64_t)(nx + 1) * (nz + 1);
    for (int j = 0; j < nidx; ++j)
        if (idx[j] < 0 || idx[j] >= npts) return fail(POE2D_E_INVALID, "patchy_samples: index outside the grid");
    for (int k = 0; k < nmodels; ++k)
        if (!(gas[k] >= 0 && gas[k] < 1) || !(radius[k] > 0))
            return fail(POE2D_E_INVALID, "patchy_samples: need gas in [0, 1) and radius > 0");
    int nt = threads > 0 ? threads : std::min(16, (int)std::thread::hardware_concurrency());
    nt = std::max(1, std::min(nt, std::max(nmodels, 1)));
    std::atomic<bool> failed{false};
    auto work = [&](int t0) {
        std::vector<uint8_t> mask((size_t)npts);
        for (int k = t0; k < nmodels; k += nt) {
            uint8_t *row = out + (size_t)k * nidx;
            if (gas[k] <= 0) {
                std::memset(row, 0, (size_t)nidx);
                continue;
            }
            size_t painted = 0;
            if (!paint_mask(nx, nz, dx, dz, gas[k], seed, radius[k], radius[k], mask.data(), &painted)) {
                failed = true;
                return;
            }
            for (int j = 0; j < nidx; ++j) row[j] = mask[(size_t)idx[j]];
        }
    };
    std::vector<std::thread> pool;
    for (int t = 1; t < nt; ++t) pool.emplace_back(work, t);
    work(0);
    for (std::thread &th : pool) th.join();
    if (failed) return fail(POE2D_E_INVALID, "patchy_samples: target fraction not reachable");
    return POE2D_OK;
}

// The six lattice layouts of patchymodel.m (types 1..6): disks of radius r points on a
// (rows x cols) layer, value 1 inside a disk and 0 outside.  Lattice steps for each type (first
// centre, step along x, first centre, step along z), with dia = 2r, 1-based centres as in the
// MATLAB loops; types 4-6 superpose two lattices:
//   1: (r+1, dia)   x (r+1, dia)        2: (r+1, 2dia)  x (r+1, 2dia)
//   3: (r+1, 4dia)  x (r+1, 2dia)       4: type 2 + (3r+1, 2dia) x (3r+1, 2dia)
//   5: (r+1, 2dia) x (3r+1, 2dia) + (r+1, dia) x (r+1, 2dia)
//   6: (r+1, 4dia) x (r+1, 2dia) + (5r+1, 4dia) x (3r+1, 2dia)
poe2d_status poe2d_patchy_lattice(int type, int rows, int cols, int r, uint8_t *layer)
{
    if (type < 1 || type > 6 || rows < 1 || cols < 1 || r < 1 || !layer)
        return fail(POE2D_E_INVALID, "patchy_lattice: bad arguments");
    std::memset(layer, 0, (size_t)rows * cols);
    const int dia = 2 * r;
    struct Lat { int x0, sx, z0, sz; };
    Lat lats[2];
    int nl = 1;
    switch (type) {
    case 1: lats[0] = {r + 1, dia, r + 1, dia}; break;
    case 2: lats[0] = {r + 1, 2 * dia, r + 1, 2 * dia}; break;
    case 3: lats[0] = {r + 1, 4 * dia, r + 1, 2 * dia}; break;
    case 4: lats[0] = {r + 1, 2 * dia, r + 1, 2 * dia}; lats[1] = {3 * r + 1, 2 * dia, 3 * r + 1, 2 * dia}; nl = 2; break;
    case 5: lats[0] = {r + 1, 2 * dia, 3 * r + 1, 2 * dia}; lats[1] = {r + 1, dia, r + 1, 2 * dia}; nl = 2; break;
    default: lats[0] = {r + 1, 4 * dia, r + 1, 2 * dia}; lats[1] = {5 * r + 1, 4 * dia, 3 * r + 1, 2 * dia}; nl = 2; break;
    }
    for (int l = 0; l < nl; ++l)
        for (int ci = lats[l].x0; ci <= cols; ci += lats[l].sx)
            for (int cj = lats[l].z0; cj <= rows; cj += lats[l].sz)
                for (int j = std::max(1, cj - r); j <= std::min(rows, cj + r); ++j)
                    for (int i = std::max(1, ci - r); i <= std::min(cols, ci + r); ++i) {
                        const double d = std::sqrt((double)(i - ci) * (i - ci) + (double)(j - cj) * (j - cj));
                        if (d <= r) layer[(size_t)(j - 1) * cols + (i - 1)] = 1;
                    }
    return POE2D_OK;
}

}  // extern "C"
