// Reference-format input and output behind the C ABI: the two text inputs, float32 model grids,
// float32 snapshots, ASCII trace tables, and the reference's receiver layout.  Host code only.
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include "../../include/poe2d.h"
#include "common.h"

using poe2d::fail;

namespace {

poe2d_status io_fail(const char *what, const char *path)
{
    return fail(POE2D_E_IO, std::string(what) + " '" + (path ? path : "(null)") + "'" +
                                (errno ? std::string(": ") + std::strerror(errno) : std::string()));
}

}  // namespace

extern "C" {

poe2d_status poe2d_read_input1(const char *path, int *nstep, int *isource, int *jsource, int *a0_pos, int *a1_pos)
{
    if (!path || !nstep || !isource || !jsource || !a0_pos || !a1_pos)
        return fail(POE2D_E_INVALID, "read_input1: null argument");
    errno = 0;
    std::ifstream in(path);
    if (!in.is_open()) return io_fail("read_input1: cannot open", path);
    int v[5];
    for (int k = 0; k < 5; ++k)
        if (!(in >> v[k]))
            return fail(POE2D_E_IO, std::string("read_input1: '") + path + "' needs five integers "
                                    "(NSTEP ISOURCE JSOURCE a0_pos a1_pos), value " + std::to_string(k + 1) + " missing");
    *nstep = v[0];
    *isource = v[1];
    *jsource = v[2];
    *a0_pos = v[3];
    *a1_pos = v[4];
    return POE2D_OK;
}

poe2d_status poe2d_read_input2(const char *path, double *f0)
{
    if (!path || !f0) return fail(POE2D_E_INVALID, "read_input2: null argument");
    errno = 0;
    std::ifstream in(path);
    if (!in.is_open()) return io_fail("read_input2: cannot open", path);
    float f = 0.f;  // the reference reads f0 into a float
    if (!(in >> f)) return fail(POE2D_E_IO, std::string("read_input2: '") + path + "' holds no frequency");
    *f0 = f;
    return POE2D_OK;
}

poe2d_status poe2d_read_grid_f32(const char *path, int nx1, int nz1, int layout, float *out)
{
    if (!path || !out || nx1 < 1 || nz1 < 1 || (layout != POE2D_X_FASTEST && layout != POE2D_Z_FASTEST))
        return fail(POE2D_E_INVALID, "read_grid_f32: bad arguments");
    const size_t n = (size_t)nx1 * nz1;
    errno = 0;
    FILE *f = std::fopen(path, "rb");
    if (!f) return io_fail("read_grid_f32: cannot open", path);
    std::vector<float> buf(n);
    const size_t got = std::fread(buf.data(), sizeof(float), n, f);
    const bool extra = std::fgetc(f) != EOF;
    std::fclose(f);
    if (got != n || extra)
        return fail(POE2D_E_IO, std::string("read_grid_f32: '") + path + "' does not hold exactly " +
                                    std::to_string(n) + " float32 values");
    if (layout == POE2D_X_FASTEST) {
        std::memcpy(out, buf.data(), n * sizeof(float));
    } else {
        // value of (x, z) at x * nz1 + z (a column-major flattening of an nz1 x nx1 matrix)
        for (int z = 0; z < nz1; ++z)
            for (int x = 0; x < nx1; ++x) out[(size_t)z * nx1 + x] = buf[(size_t)x * nz1 + z];
    }
    return POE2D_OK;
}

poe2d_status poe2d_write_grid_f32(const char *path, int nx1, int nz1, const float *data)
{
    if (!path || !data || nx1 < 1 || nz1 < 1) return fail(POE2D_E_INVALID, "write_grid_f32: bad arguments");
    errno = 0;
    FILE *f = std::fopen(path, "wb");
    if (!f) return io_fail("write_grid_f32: cannot create", path);
    const size_t n = (size_t)nx1 * nz1;
    const size_t put = std::fwrite(data, sizeof(float), n, f);
    if (std::fclose(f) != 0 || put != n) return io_fail("write_grid_f32: write failed", path);
    return POE2D_OK;
}

poe2d_status poe2d_write_traces(const char *path, int nrows, int ncols, const float *data, int one_line)
{
    if (!path || nrows < 0 || ncols < 0 || (!data && nrows * ncols > 0))
        return fail(POE2D_E_INVALID, "write_traces: bad arguments");
    errno = 0;
    std::ofstream out(path);
    if (!out.is_open()) return io_fail("write_traces: cannot create", path);
    // default stream formatting (6 significant digits), each value followed by a space; a table
    // ends every row with a newline, a single trace is one line without one
    for (int r = 0; r < nrows; ++r) {
        for (int c = 0; c < ncols; ++c) out << data[(size_t)r * ncols + c] << " ";
        if (!one_line) out << "\n";
    }
    out.flush();
    if (!out) return io_fail("write_traces: write failed", path);
    return POE2D_OK;
}

poe2d_status poe2d_reference_receivers(int nx, int nz, int isource, int jsource, int a0_pos, int a1_pos,
                                       int panj, int panj2, int cap, int *ix, int *iz, int *n_vsp,
                                       int *n_hor, int *n_pts)
{
    if (!n_vsp || !n_hor || !n_pts || nx < 1 || nz < 1 || panj < 1 || panj2 < 1 || cap < 0 ||
        (cap > 0 && (!ix || !iz)))
        return fail(POE2D_E_INVALID, "reference_receivers: bad arguments");
    const int len_v = (nz + 1) / panj;   // VSP spacing in rows
    const int len_h = (nx + 1) / panj2;  // horizontal-line spacing in columns
    if (len_v < 1 || len_h < 1)
        return fail(POE2D_E_INVALID, "reference_receivers: panj must be <= NZ+1 and panj2 <= NX+1 "
                                     "(the receiver spacing (N+1)/panj would be zero)");
    if (isource < 0 || isource > nx || nz < 10)
        return fail(POE2D_E_INVALID, "reference_receivers: ISOURCE outside the grid or NZ < 10");
    const int xp = (nx + 1) / 2 - 1;  // column of the point traces
    const int pz[3] = {a0_pos, a1_pos, jsource};
    for (int k = 0; k < 3; ++k)
        if (pz[k] < 0 || pz[k] > nz)
            return fail(POE2D_E_INVALID, "reference_receivers: a0_pos, a1_pos and JSOURCE must lie in [0, NZ]");
    int nv = 0, nh = 0;
    for (int m = 1; m <= panj && m * len_v - 1 <= nz; ++m) ++nv;
    for (int m = 1; m <= panj2 && m * len_h - 1 <= nx; ++m) ++nh;
    *n_vsp = nv;
    *n_hor = nh;
    *n_pts = 3;
    const int total = nv + nh + 3;
    if (cap == 0) return POE2D_OK;  // size query
    if (cap < total) return fail(POE2D_E_INVALID, "reference_receivers: capacity below " + std::to_string(total));
    int k = 0;
    for (int m = 1; m <= nv; ++m, ++k) {  // VSP: column ISOURCE, rows y with (y+1) % len_v == 0
        ix[k] = isource;
        iz[k] = m * len_v - 1;
    }
    for (int m = 1; m <= nh; ++m, ++k) {  // horizontal line: row 10, columns x with (x+1) % len_h == 0
        ix[k] = m * len_h - 1;
        iz[k] = 10;
    }
    for (int q = 0; q < 3; ++q, ++k) {
        ix[k] = xp;
        iz[k] = pz[q];
    }
    return POE2D_OK;
}

}  // extern "C"
