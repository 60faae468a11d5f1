// poe2d-run: the reference program's run, on the C ABI of libpoe2d.
//
// Reads input1.txt (NSTEP ISOURCE JSOURCE a0_pos a1_pos) and input2.txt (f0) from the working
// directory, the four float32 model grids (medium index, c_p, c_s, saturated density), derives
// the eight coefficients with the per-medium fluid tables, checks the Courant number against the
// largest input c_p, injects the pressure source at (ISOURCE, JSOURCE), records the VSP line, the
// horizontal line and the three point traces, writes vz%05i.vzo / vx%05i.vxo snapshots every
// `samp` steps and after the last step, and finally the ASCII seismograms recvspvz / recvspvx
// (.vsp), rechorvz / rechorvx (.hor), recats (.ats), recbwh (.bwh) and recsrc (.src), named by
// int(f0) as the reference names them (seismic_CPML_2D_poroelastic_withCUDA.cu:472-1905).
//
// The grid constants the reference hard-codes (NX = 13599, NZ = 2439, dx = 1.25 m, dt = 95e-6 s,
// 10 C-PML points, panj = 122, panj2 = 680, samp = 5000) are the defaults of the options below.
// Time stepping follows the canonical semantics: exactly NSTEP steps, sample k of every trace is
// the state at the start of step k.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include "../../include/poe2d.h"

namespace {

struct Options {
    int nx = 13599, nz = 2439, npml = 10, panj = 122, panj2 = 680, samp = 5000;
    double dx = 1.25, dt = 95e-6;
    std::string medium = "mar_poro_2440x13600.bin", cp = "mar_cp_2440x13600.bin",
                cs = "mar_cs_2440x13600.bin", rho = "mar_rho_2440x13600.bin";
    int layout = POE2D_X_FASTEST;
    std::string media;  // optional per-medium table file
    std::string input1 = "input1.txt", input2 = "input2.txt", outdir = ".";
    int precision = 32;
    std::string tile = "auto";
    int schedule = -1;  // -1: the handle's default (3 for FP32, 0 for FP64)
    bool quiet = false;
    int stress_law = POE2D_STRESS_AS_WRITTEN;
    int material = POE2D_MATERIAL_AS_WRITTEN;
};

// Per-medium fluid tables of the reference (seismic_CPML_2D_poroelastic_withCUDA.cu:585-598, the
// "elastic" porosity and fluid-density rows being the active ones there).
struct Media {
    std::vector<float> kf = {2.4e09f, 2.4e09f, 0.7e09f, 0.02e09f};
    std::vector<float> eta = {0.003f, 0.003f, 0.03f, 1.0e-5f};
    std::vector<float> kappa = {1.5e-17f, 500.0e-15f, 500.0e-15f, 500.0e-15f};
    std::vector<float> phi = {1.e-6f, 1.e-6f, 1.e-6f, 1.e-6f};
    std::vector<float> a = {2.0f, 2.0f, 2.0f, 2.0f};
    std::vector<float> rhof = {10.0f, 10.0f, 10.0f, 10.0f};
};

void usage()
{
    std::printf(
        "usage: poe2d-run [options]\n"
        "  --nx N --nz N        grid is (NX+1) x (NZ+1) points (13599, 2439)\n"
        "  --dx M --dt S        spacing in metres (1.25, also dz) and time step in seconds (95e-6)\n"
        "  --npml N             C-PML thickness in points (10)\n"
        "  --medium F --cp F --cs F --rho F   float32 model grids (mar_*_2440x13600.bin)\n"
        "  --layout x|z         grid files x fastest (arr2bin) or z fastest (patchymodel) (x)\n"
        "  --media F            per-medium table: one line per medium index 0,1,..:\n"
        "                       K_f eta kappa phi tortuosity rho_f (default: the reference's 4 media)\n"
        "  --input1 F --input2 F   (input1.txt, input2.txt)\n"
        "  --panj N --panj2 N   VSP and horizontal receiver counts (122, 680)\n"
        "  --samp N             snapshot interval in steps (5000)\n"
        "  --outdir D           where the output files go (.)\n"
        "  --precision 32|64    arithmetic of the time loop (32)\n"
        "  --schedule 0|1|2|3   step schedule (poe2d_set_schedule; default 3 in FP32, 0 in FP64)\n"
        "  --tile auto|BWxTH    tile shape of the schedule's kernel (auto: autotune)\n"
        "  --stress-law written|drained   lambda_c as the reference writes it, or Biot's drained\n"
        "                       lambda in the normal stresses (written)\n"
        "  --material written|averaged    staggered-point material as written, or harmonic mu for\n"
        "                       sigma_xy and mean material for v_y (written)\n"
        "  --quiet\n");
}

bool parse(int argc, char **argv, Options &o)
{
    for (int i = 1; i < argc; ++i) {
        const std::string k = argv[i];
        auto val = [&]() -> const char * {
            if (i + 1 >= argc) {
                std::fprintf(stderr, "poe2d-run: %s needs a value\n", k.c_str());
                std::exit(2);
            }
            return argv[++i];
        };
        if (k == "--nx") o.nx = std::atoi(val());
        else if (k == "--nz") o.nz = std::atoi(val());
        else if (k == "--dx") o.dx = std::atof(val());
        else if (k == "--dt") o.dt = std::atof(val());
        else if (k == "--npml") o.npml = std::atoi(val());
        else if (k == "--medium") o.medium = val();
        else if (k == "--cp") o.cp = val();
        else if (k == "--cs") o.cs = val();
        else if (k == "--rho") o.rho = val();
        else if (k == "--layout") {
            const std::string v = val();
            if (v != "x" && v != "z") return false;
            o.layout = v == "x" ? POE2D_X_FASTEST : POE2D_Z_FASTEST;
        } else if (k == "--media") o.media = val();
        else if (k == "--input1") o.input1 = val();
        else if (k == "--input2") o.input2 = val();
        else if (k == "--panj") o.panj = std::atoi(val());
        else if (k == "--panj2") o.panj2 = std::atoi(val());
        else if (k == "--samp") o.samp = std::atoi(val());
        else if (k == "--outdir") o.outdir = val();
        else if (k == "--precision") o.precision = std::atoi(val());
        else if (k == "--tile") o.tile = val();
        else if (k == "--schedule") o.schedule = std::atoi(val());
        else if (k == "--quiet") o.quiet = true;
        else if (k == "--stress-law") {
            const std::string v = val();
            if (v != "written" && v != "drained") return false;
            o.stress_law = v == "drained" ? POE2D_STRESS_DRAINED : POE2D_STRESS_AS_WRITTEN;
        } else if (k == "--material") {
            const std::string v = val();
            if (v != "written" && v != "averaged") return false;
            o.material = v == "averaged" ? POE2D_MATERIAL_AVERAGED : POE2D_MATERIAL_AS_WRITTEN;
        }
        else if (k == "-h" || k == "--help") {
            usage();
            std::exit(0);
        } else {
            std::fprintf(stderr, "poe2d-run: unknown option %s\n", k.c_str());
            return false;
        }
    }
    return o.nx >= 8 && o.nz >= 10 && o.dx > 0 && o.dt > 0 && o.samp >= 1 &&
           (o.precision == 32 || o.precision == 64);
}

bool read_media(const std::string &path, Media &m)
{
    std::ifstream in(path);
    if (!in.is_open()) return false;
    Media t;
    for (auto *v : {&t.kf, &t.eta, &t.kappa, &t.phi, &t.a, &t.rhof}) v->clear();
    float r[6];
    while (in >> r[0] >> r[1] >> r[2] >> r[3] >> r[4] >> r[5]) {
        t.kf.push_back(r[0]);
        t.eta.push_back(r[1]);
        t.kappa.push_back(r[2]);
        t.phi.push_back(r[3]);
        t.a.push_back(r[4]);
        t.rhof.push_back(r[5]);
    }
    if (t.kf.empty() || !in.eof()) return false;
    m = t;
    return true;
}

int die(const char *what)
{
    std::fprintf(stderr, "poe2d-run: %s: %s\n", what, poe2d_last_error());
    return 1;
}

std::string out_path(const Options &o, const char *name) { return o.outdir + "/" + name; }

}  // namespace

int main(int argc, char **argv)
{
    Options o;
    if (!parse(argc, argv, o)) {
        usage();
        return 2;
    }
    int nstep = 0, is = 0, js = 0, a0 = 0, a1 = 0;
    double f0 = 0;
    // the reference's answer to a missing input file (:472-484)
    if (poe2d_read_input1(o.input1.c_str(), &nstep, &is, &js, &a0, &a1) != POE2D_OK ||
        poe2d_read_input2(o.input2.c_str(), &f0) != POE2D_OK) {
        std::printf("File not opened...");
        std::fprintf(stderr, "\npoe2d-run: %s\n", poe2d_last_error());
        return 1;
    }
    if (nstep < 1 || !(f0 > 0)) {
        std::fprintf(stderr, "poe2d-run: NSTEP must be >= 1 and f0 > 0\n");
        return 1;
    }
    Media media;
    if (!o.media.empty() && !read_media(o.media, media)) {
        std::fprintf(stderr, "poe2d-run: cannot read the media table '%s'\n", o.media.c_str());
        return 1;
    }

    const int nx1 = o.nx + 1, nz1 = o.nz + 1;
    const size_t npts = (size_t)nx1 * nz1;
    std::vector<float> medium(npts), cp(npts), cs(npts), rhos(npts);
    if (poe2d_read_grid_f32(o.medium.c_str(), nx1, nz1, o.layout, medium.data()) ||
        poe2d_read_grid_f32(o.cp.c_str(), nx1, nz1, o.layout, cp.data()) ||
        poe2d_read_grid_f32(o.cs.c_str(), nx1, nz1, o.layout, cs.data()) ||
        poe2d_read_grid_f32(o.rho.c_str(), nx1, nz1, o.layout, rhos.data()))
        return die("model grid");

    std::vector<float> coef[8];
    for (auto &c : coef) c.assign(npts, 0.f);
    const poe2d_status mst = poe2d_material_reference(
        npts, cp.data(), cs.data(), rhos.data(), medium.data(), (int)media.kf.size(), media.kf.data(),
        media.eta.data(), media.kappa.data(), media.phi.data(), media.a.data(), media.rhof.data(),
        coef[0].data(), coef[1].data(), coef[2].data(), coef[3].data(), coef[4].data(), coef[5].data(),
        coef[6].data(), coef[7].data());
    if (mst != POE2D_OK) return die("material");

    // Courant number against the largest input c_p (:716, 1211)
    const float maxcp = *std::max_element(cp.begin(), cp.end());
    if (maxcp * o.dt / o.dx > 6.0 / 7.0 / std::sqrt(2.0)) {
        std::printf("'time step is too large, simulation will be unstable'");
        return 0;
    }

    poe2d_grid g{};
    g.nx = o.nx;
    g.nz = o.nz;
    g.dx = g.dz = o.dx;
    g.dt = o.dt;
    g.npml = o.npml;
    g.npower = 2.0;
    g.rcoef = 0.001;
    g.kmax = 1.0;
    g.alpha_max = 0.0;  // pi f0, evaluated in the handle's precision
    g.f0 = f0;
    g.cp_max = maxcp;
    g.precision = o.precision == 32 ? POE2D_F32 : POE2D_F64;
    g.check_cfl = 0;  // checked above, the reference's way
    g.stress_law = o.stress_law;
    g.staggered_material = o.material;
    poe2d_coeffs c{};
    c.dtype = POE2D_F32;
    c.rho = coef[0].data();
    c.rhof = coef[1].data();
    c.m = coef[2].data();
    c.mu = coef[3].data();
    c.lambdac = coef[4].data();
    c.M = coef[5].data();
    c.alpha = coef[6].data();
    c.b = coef[7].data();
    c.pitch = 0;
    poe2d_solver *s = nullptr;
    if (poe2d_create(&g, &c, &s) != POE2D_OK) return die("create");

    poe2d_source src{};
    src.kind = POE2D_SRC_PRESSURE_DGAUSS;
    src.ix = is;
    src.iz = js;
    src.amp = 1.0e10;  // faktor (:45)
    src.t0 = 1.0 / f0;
    src.t_cut = 0.4;   // the source gate of the CUDA program (:107)
    if (poe2d_set_source(s, &src) != POE2D_OK) return die("source");

    int nv = 0, nh = 0, np = 0;
    if (poe2d_reference_receivers(o.nx, o.nz, is, js, a0, a1, o.panj, o.panj2, 0, nullptr, nullptr, &nv, &nh, &np))
        return die("receivers");
    const int nrec = nv + nh + np;
    std::vector<int> rix(nrec), riz(nrec);
    if (poe2d_reference_receivers(o.nx, o.nz, is, js, a0, a1, o.panj, o.panj2, nrec, rix.data(), riz.data(), &nv,
                                  &nh, &np))
        return die("receivers");
    poe2d_receivers rec{};
    rec.n = nrec;
    rec.ix = rix.data();
    rec.iz = riz.data();
    rec.comp_mask = POE2D_VX | POE2D_VY;
    rec.decimate = 1;
    rec.nsamp = nstep;
    if (poe2d_set_receivers(s, &rec) != POE2D_OK) return die("receivers");

    const int sched = o.schedule >= 0 ? o.schedule : (o.precision == 64 ? 0 : 3);
    if (o.tile == "auto") {
        int bw = 0, th = 0;
        if (poe2d_set_schedule(s, sched, 256, sched == 3 ? 24 : 14) != POE2D_OK) return die("schedule");
        if (sched == 1) {
            if (!o.quiet) std::printf("two-pass schedule (no tile shape)\n");
        } else {
            if (poe2d_autotune(s, 5, &bw, &th) != POE2D_OK) return die("autotune");
            if (!o.quiet) std::printf("schedule %d, tile %dx%d (autotuned)\n", sched, bw, th);
        }
    } else {
        int bw = 0, th = 0;
        if (std::sscanf(o.tile.c_str(), "%dx%d", &bw, &th) != 2 || poe2d_set_schedule(s, sched, bw, th) != POE2D_OK)
            return die("tile");
    }

    const size_t vbytes = npts * (o.precision == 32 ? 4 : 8);
    std::vector<unsigned char> fbuf(vbytes);
    std::vector<float> f32(npts);
    auto snapshot = [&](int field, const char *fmt, int step) -> bool {
        if (poe2d_snapshot(s, field, fbuf.data(), vbytes) != POE2D_OK) return false;
        for (size_t i = 0; i < npts; ++i)
            f32[i] = o.precision == 32 ? ((const float *)fbuf.data())[i] : (float)((const double *)fbuf.data())[i];
        char name[64];
        std::snprintf(name, sizeof name, fmt, step);
        return poe2d_write_grid_f32(out_path(o, name).c_str(), nx1, nz1, f32.data()) == POE2D_OK;
    };

    const auto t_start = std::chrono::steady_clock::now();
    int done = 0;
    while (done < nstep) {
        const int k = std::min(o.samp, nstep - done);
        if (poe2d_run(s, k) != POE2D_OK) return die("run");
        done += k;
        if (!snapshot(POE2D_FIELD_VY, "vz%05i.vzo", done) || !snapshot(POE2D_FIELD_VX, "vx%05i.vxo", done))
            return die("snapshot");
        if (!o.quiet) {
            const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count();
            std::printf("f0 = %g Hz; step %5d of %d; %.2f s\n", f0, done, nstep, secs);
        }
    }

    // traces [comp][receiver][sample], comp 0 = v_x, 1 = v_y (= v_z of the reference)
    const size_t nsv = (size_t)2 * nrec * nstep;
    std::vector<unsigned char> tb(nsv * (o.precision == 32 ? 4 : 8));
    if (poe2d_seismograms(s, tb.data(), tb.size()) != POE2D_OK) return die("seismograms");
    auto tr = [&](int comp, int r, int k) -> float {
        const size_t i = ((size_t)comp * nrec + r) * nstep + k;
        return o.precision == 32 ? ((const float *)tb.data())[i] : (float)((const double *)tb.data())[i];
    };
    // a line table: one row per sample, `cols` columns (positions the layout leaves out stay 0)
    auto table = [&](int comp, int first, int count, int cols) {
        std::vector<float> t((size_t)nstep * cols, 0.f);
        for (int k = 0; k < nstep; ++k)
            for (int r = 0; r < count; ++r) t[(size_t)k * cols + r] = tr(comp, first + r, k);
        return t;
    };
    auto point = [&](int r) {
        std::vector<float> t(nstep);
        for (int k = 0; k < nstep; ++k) t[k] = tr(1, r, k);
        return t;
    };
    const int fi = (int)f0;
    char n_vspz[64], n_vspx[64], n_horz[64], n_horx[64], n_src[64], n_ats[64], n_bwh[64];
    std::snprintf(n_vspz, sizeof n_vspz, "recvspvz%08i.vsp", fi);
    std::snprintf(n_vspx, sizeof n_vspx, "recvspvx%08i.vsp", fi);
    std::snprintf(n_horz, sizeof n_horz, "rechorvz%08i.hor", fi);
    std::snprintf(n_horx, sizeof n_horx, "rechorvx%08i.hor", fi);
    if (f0 < 1.0) {
        std::snprintf(n_src, sizeof n_src, "recsrc%08f.src", f0);
        std::snprintf(n_ats, sizeof n_ats, "recats%08f.ats", f0);
        std::snprintf(n_bwh, sizeof n_bwh, "recbwh%08f.bwh", f0);
    } else {
        std::snprintf(n_src, sizeof n_src, "recsrc%08i.src", fi);
        std::snprintf(n_ats, sizeof n_ats, "recats%08i.ats", fi);
        std::snprintf(n_bwh, sizeof n_bwh, "recbwh%08i.bwh", fi);
    }
    struct Out {
        const char *name;
        std::vector<float> data;
        int cols;
        int one_line;
    };
    const Out outs[] = {
        {n_vspz, table(1, 0, nv, o.panj), o.panj, 0},      {n_vspx, table(0, 0, nv, o.panj), o.panj, 0},
        {n_horz, table(1, nv, nh, o.panj2), o.panj2, 0},   {n_horx, table(0, nv, nh, o.panj2), o.panj2, 0},
        {n_ats, point(nv + nh + 0), nstep, 1},            {n_bwh, point(nv + nh + 1), nstep, 1},
        {n_src, point(nv + nh + 2), nstep, 1},
    };
    for (const Out &w : outs) {
        const int rows = w.one_line ? 1 : nstep;
        if (poe2d_write_traces(out_path(o, w.name).c_str(), rows, w.cols, w.data.data(), w.one_line) != POE2D_OK)
            return die("write traces");
    }
    if (!o.quiet) std::printf("posisi shot : X= %i Y=%i\n", is, js);
    poe2d_destroy(s);
    return 0;
}
