"""Benchmark: FP32 Biot C-PML stepping on configuration 1 (4001 x 4001 points, patchy gas 0.15
over brine rock, 6000 steps, 801 receivers x (vx, vy, p)) on one MI355X.

A "step" of this benchmark is one full solve (nt = 6000 time steps, receiver gather included)
with the inputs resident in HBM; set-up, uploads, the tile autotuning and the trace download are
outside the timed region.  value = Gcell-updates/s = 4001^2 * 6000 * K / (device time of K solves,
hipEvents on the solver's stream).

Roofline: the dominant kernel is the step kernel: the single-pass kernel (one launch per time
step) or the two-step kernel (one launch per two steps, the intermediate state on chip), whichever
autotunes faster on the grid; either records the receivers, and the device step counter advances
once per 32 steps.  The algorithmic traffic of a time step is 104 B per point-update (9 state
fields read + written, 8 coefficients read, FP32) plus 48 B per C-PML strip point and 96 B per
corner point; `achieved` divides that by the time per step of the timed solves (hipEvents on the
solver's stream around the whole run, so launch gaps count against the kernel).  The figure is
fixed by the algorithm, not by the schedule: the two-step kernel performs two cell-updates per
point per launch while reading and writing the state once, so its `achieved` (and `frac`) can pass
the HBM peak; `hbm_gbs` beside it is the measured traffic over the same time.  `traffic` is the
measured HBM bytes per time step (rocprofv3 FETCH_SIZE and WRITE_SIZE per launch, each scaled by
what the same counter reads for a plain copy of known size, halved for the two-step kernel), read
from the committed profile summary when it was taken with the same kernel and tile shape, else
null.

The CPU baseline is the oracle in its reference-structured form (five sweeps plus three full-grid
receiver scans per step, one thread) timed on a bounded sample of the same model.
"""
import argparse
import json
import os
import platform
import statistics
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
BYTES_PER_POINT = 104.0
BYTES_PER_STRIP_POINT = 48.0


def algorithmic_bytes_per_step(nx, nz, npml_cols, npml_rows):
    """Per time step: 104 B per point + 48 B per strip point (6 memory variables read + written)
    in the x strips and again in the z strips (corners count twice: 96 B)."""
    npts = (nx + 1) * (nz + 1)
    return BYTES_PER_POINT * npts + BYTES_PER_STRIP_POINT * (npml_cols * (nz + 1) + npml_rows * (nx + 1))


def strip_counts(prob):
    from poe2d_amd.material import cpml_profile
    from poe2d_amd import F32
    cols = rows = 0
    for axis, n, d in ((0, prob.nx + 1, prob.dx), (1, prob.nz + 1, prob.dz)):
        t = cpml_profile(axis, n, d, prob.dt, prob.npml, prob.npower, prob.rcoef, prob.kmax,
                         np.pi * prob.f0, 3605.0, F32)
        active = (t["a"] != 0) | (t["a_half"] != 0) | (t["K"] != 1) | (t["K_half"] != 1)
        if axis == 0:
            cols = int(active.sum())
        else:
            rows = int(active.sum())
    return cols, rows


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "fused_step_4001x4001_traffic.txt")


def recorded_traffic(tile, n, schedule=0):
    """HBM bytes per time step from the committed PMC summary ("grid_n: N", "tile BWxTH: bytes"
    for the one-step kernel and "two-step BWxTH: bytes" for the two-step kernel, bytes per
    launch), if it was measured on this grid with this kernel and tile shape."""
    try:
        lines = [ln.strip() for ln in open(TRAFFIC_FILE) if ln.strip() and not ln.startswith("#")]
    except OSError:
        return None, None
    kv = dict(ln.split(":", 1) for ln in lines if ":" in ln)
    kv = {k.strip(): v.strip() for k, v in kv.items()}
    key, per = (f"tile {tile[0]}x{tile[1]}", 1) if schedule == 0 else (f"two-step {tile[0]}x{tile[1]}", 2)
    try:
        if int(kv.get("grid_n", -1)) != n + 1:
            return None, None
        return float(kv[key]) / per, os.path.relpath(TRAFFIC_FILE, ROOT)
    except (KeyError, ValueError):
        return None, None


def components(mask):
    from poe2d_amd import _abi
    return ", ".join(n for bit, n in ((_abi.VX, "vx"), (_abi.VY, "vy"), (_abi.P, "p")) if mask & bit)


def trace_rel_l2(g, o, floor=1e-3):
    """Per-trace relative L2 distance ||g - o|| / max(||o||, floor x the largest trace norm of the
    component) over the last axis; g, o shaped [ncomp, ntrace, nsamp]."""
    g = np.asarray(g, np.float64)
    o = np.asarray(o, np.float64)
    num = np.linalg.norm(g - o, axis=-1)
    den = np.linalg.norm(o, axis=-1)
    return num / np.maximum(np.maximum(den, floor * den.max(axis=-1, keepdims=True)), 1e-300)


def compulsory_bytes_per_step(npts, strip_points, schedule, palette):
    """Bytes per time step that the schedule cannot avoid: each launch reads and writes the V and S
    groups (64 B per point, FP32) and reads the one-byte material index (palette) or both coefficient
    groups (32 B); the two-step kernel does this once per two steps.  Every step reads and writes the
    six memory variables of each strip record (48 B per strip point and strip)."""
    per_point = 64.0 + (1.0 if palette else 32.0)
    per_launch_steps = 2 if schedule == 3 else 1
    return per_point * npts / per_launch_steps + 48.0 * strip_points


def cpu_leg(prob, make, steps, gpu_traces):
    """The oracle on the host: (1) the CPU baseline, its reference-structured form (five sweeps plus
    three full-grid receiver scans per step, one thread) timed over `steps` steps of the same model;
    (2) as the checker, the first steps of the GPU's source run (gpu_traces, [3, nrec, k]) against the
    FP32 oracle run for k steps on all host threads.  Returns (baseline, parity or None)."""
    from oracle import pyoracle
    from oracle.pyoracle import OracleSolver

    cprob = make(steps + 1, pyoracle)
    s = OracleSolver(cprob, threads=1)
    s.set_reference_scans(122, 680, cprob.nz // 2 + 100, cprob.nz // 2 + 200)
    s.run(1)  # untimed first step (page-in)
    t = time.perf_counter()
    s.run(steps)
    dt = time.perf_counter() - t
    s.close()
    npts = (cprob.nx + 1) * (cprob.nz + 1)
    cpu_model = ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                cpu_model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    baseline = {"value": npts * steps / dt / 1e9, "unit": "Gcell-updates/s", "cores": 1, "kind": "port",
                "sample": f"{steps} time steps of the same {cprob.nx + 1}x{cprob.nz + 1} model "
                          f"({dt:.1f} s), reference-structured oracle (5 sweeps + 3 full-grid receiver scans), "
                          f"1 thread; host {cpu_model or platform.processor()}, nproc {os.cpu_count()}"}
    if gpu_traces is None:
        return baseline, None
    k = gpu_traces.shape[-1]
    threads = min(16, os.cpu_count() or 1)
    t = time.perf_counter()
    o = OracleSolver(prob, threads=threads)  # the GPU's own problem: identical coefficients
    o.run(k)
    o32 = o.seismograms()[..., :k]
    o.close()
    err = trace_rel_l2(gpu_traces, o32)
    reached = int((np.linalg.norm(o32, axis=-1) > 1e-3 * np.linalg.norm(o32, axis=-1).max(axis=-1, keepdims=True)).sum())
    print(f"[bench] oracle {k} steps on {threads} threads in {time.perf_counter() - t:.1f} s: max per-trace "
          f"rel L2 {err.max():.3e} ({reached} traces above 1e-3 of their component's largest)",
          file=sys.stderr, flush=True)
    return baseline, {"max": float(err.max()),
                      "what": f"max over the {gpu_traces.shape[1]} x {gpu_traces.shape[0]} traces of the per-trace relative L2 distance between the GPU "
                              f"and the FP32 oracle (the reference's CPU loop restated) over the first {k} steps "
                              f"of the source run ({reached} traces carry signal by then); oracle on {threads} "
                              f"host threads"}


def copy_probe(mib, reps=20):
    import torch
    n = mib * (1 << 20) // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        dst.copy_(src)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    moved = 2 * n * 4
    del src, dst
    print(json.dumps({"copy_probe_mib": mib, "bytes_read_per_copy": n * 4, "bytes_written_per_copy": n * 4,
                      "ms_per_copy": round(ms, 4), "GBps_read_plus_write": round(moved / ms / 1e6, 1)}), flush=True)


def build_batch(which, nt):
    """Configuration 2's batch A (64 Ricker shots over one 1201^2 patchy model, clones sharing the
    model) or batch B (32 patchy models, sides uniform in [300, 1500] points): (members, workload,
    algorithmic bytes per point-update, points)."""
    from poe2d_amd.configs import config2a, config2b
    from poe2d_amd.solver import Solver

    if which == "batchA":
        base, shots, rec = config2a(nt=nt)
        first = Solver(base)
        members = [first] + [first.clone(source=sh, receivers=rec, nt=nt) for sh in shots[1:]]
        workload = base.meta["workload"] + f", {nt} steps, 101 receivers x (vx, vy, p) per shot"
        # material amortised over the lock-step shots: 4 B x (2 x 9 + 8 / shots)
        bytes_per_point = 4.0 * (18 + 8.0 / len(members))
    else:
        probs = config2b(nt=nt)
        members = [Solver(p) for p in probs]
        workload = f"biot-cpml {len(probs)} patchy models, sides U[300,1500], gas U[0,0.3], {nt} steps, 101 receivers each"
        bytes_per_point = BYTES_PER_POINT
    npts = sum((m.problem.nx + 1) * (m.problem.nz + 1) for m in members)
    return members, workload, bytes_per_point, npts


def time_batch(members, nt, warmup, reps, schedules=(0, 3), label="batch", fixed=None):
    """Autotune the batch on each schedule (one and two steps per launch), keep the faster (or take
    fixed = (schedule, (block width, tile rows))), then time `reps` runs of nt steps after `warmup`
    untimed ones (device time, hipEvents on the batch's stream).  Returns (median ms per run,
    schedule, tile, traces finite)."""
    from poe2d_amd.solver import Batch

    with Batch(members) as b:
        best = None if fixed is None else (0.0, fixed[0], fixed[1])
        for sched in (schedules if fixed is None else ()):
            for m in members:
                m.reset()
                m.set_schedule(sched, 256, 12 if sched == 3 else 8)
            b.set_schedule(256, 12 if sched == 3 else 8)
            t = b.autotune(6)
            b.set_schedule(*t)
            b.run(20)
            ms, n = b.last_run_ms()
            print(f"[bench] {label} schedule {sched}: autotuned {t[0]}x{t[1]}, {ms / n * 1e3:.1f} us per step",
                  file=sys.stderr, flush=True)
            if best is None or ms / n < best[0]:
                best = (ms / n, sched, t)
        _, schedule, tile = best
        for m in members:
            m.set_schedule(schedule, *tile)
        b.set_schedule(*tile)

        def one_run():
            for m in members:
                m.reset()
            b.run(nt)
            return b.last_run_ms()[0]

        for _ in range(warmup):
            one_run()
        dev_ms = [one_run() for _ in range(reps)]
        finite = all(bool(np.isfinite(m.seismograms()).all()) for m in members)
    return statistics.median(dev_ms), schedule, tile, finite


SCHEDULE_NAMES = {0: "one step per launch", 3: "two steps per launch"}


def secondary_continuous(n, nt):
    """Configuration 1's geometry over a medium whose porosity and frame moduli vary continuously."""
    from poe2d_amd.configs import config1_continuous
    return secondary_no_palette(config1_continuous(nt=nt, n=n), nt, "continuous model")


def secondary_reference_grid(nt):
    """The reference program's own grid (13600 x 2440 points, 1.25 m, 10 C-PML points) over a
    synthetic layered model whose c_p varies at every point."""
    from poe2d_amd.configs import reference_grid
    return secondary_no_palette(reference_grid(nt=nt), nt, "reference grid")


def secondary_no_palette(prob, nt, label):
    """A grid with no material palette (both kernels read the 32-byte coefficient groups), each
    schedule autotuned and timed over one nt-step solve after a warm-up solve."""
    from poe2d_amd.solver import Solver

    npts = prob.npoints
    cols, rows = strip_counts(prob)
    strips = cols * (prob.nz + 1) + rows * (prob.nx + 1)
    out = {"workload": prob.meta["workload"] + f", {nt} steps", "nt": nt}
    with Solver(prob) as s:
        out["material_palette"] = s.material_palette()
        for sched in (0, 3):
            s.reset()
            s.set_schedule(sched, 256, 12 if sched == 3 else 8)
            t = s.autotune(10)
            s.set_schedule(sched, *t)
            for _ in range(2):
                s.reset()
                s.run(nt)
            ms, k = s.last_run_ms()
            step_s = ms / 1e3 / k
            out[f"schedule_{sched}"] = {"kernel": SCHEDULE_NAMES[sched], "tile": f"{t[0]}x{t[1]} (autotuned)",
                                        "us_per_step": round(step_s * 1e6, 1),
                                        "gcell_updates_per_s": round(npts / step_s / 1e9, 2),
                                        "compulsory_frac": round(compulsory_bytes_per_step(npts, strips, sched, False)
                                                                 / step_s / 1e9 / HBM_PEAK_GBS, 4)}
            print(f"[bench] {label} schedule {sched}: {t[0]}x{t[1]}, {step_s * 1e6:.1f} us per step",
                  file=sys.stderr, flush=True)
    out["two_step_speedup"] = round(out["schedule_0"]["us_per_step"] / out["schedule_3"]["us_per_step"], 3)
    return out


def batch_bench(args):
    """Configuration 2: batch A or batch B, FP32, advanced in lock-step by one work-list launch per
    time step (per two steps on the two-step kernel) with the loop in HIP graphs; tile shape and
    kernel autotuned.  value = sum over members of points x nt / device time of one batch run."""
    nt = args.nt if args.nt_given else 3000
    members, workload, bytes_per_point, npts = build_batch(args.workload, nt)
    fixed = None
    if args.tile != "auto":
        fixed = (3 if args.schedule == "auto" else int(args.schedule), tuple(int(v) for v in args.tile.lower().split("x")))
    ms, schedule, tile, finite = time_batch(members, nt, args.warmup, args.steps, label=args.workload, fixed=fixed)
    for m in members:
        m.close()
    gcell = npts * nt / (ms / 1e3) / 1e9
    achieved = npts * bytes_per_point / (ms / 1e3 / nt) / 1e9
    print(json.dumps({
        "metric": f"Gcell-updates/s (FP32 Biot C-PML, configuration 2 {args.workload}, batched)",
        "value": round(gcell, 3), "unit": "Gcell-updates/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": workload, "members": len(members), "points": npts, "nt": nt,
                   "tile": f"{tile[0]}x{tile[1]} ({'autotuned' if fixed is None else 'fixed'})",
                   "step": "one batch run of nt steps",
                   "schedule": SCHEDULE_NAMES[schedule],
                   "launch": "one work-list launch per time step (per two steps with the two-step kernel), "
                             "32-step HIP graphs", "traces_finite": finite},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "per_launch": f"one time step of the batch: {bytes_per_point:.1f} B x {npts} points "
                                   "(algorithmic, strips not counted)"},
    }), flush=True)


def secondary_batches(nt, reps=2):
    """Configuration 2's batches, timed briefly for the default line: nt steps per run, one warm-up
    run, median of `reps`."""
    out = {}
    for which in ("batchA", "batchB"):
        members, workload, bytes_per_point, npts = build_batch(which, nt)
        ms, schedule, tile, finite = time_batch(members, nt, 1, reps, label=which)
        for m in members:
            m.close()
        step_s = ms / 1e3 / nt
        out[which] = {"workload": workload, "members": len(members), "points": npts, "nt": nt,
                      "gcell_updates_per_s": round(npts / step_s / 1e9, 2), "us_per_step": round(step_s * 1e6, 1),
                      "schedule": SCHEDULE_NAMES[schedule], "tile": f"{tile[0]}x{tile[1]} (autotuned)",
                      "algorithmic_B_per_point": round(bytes_per_point, 2),
                      "frac": round(npts * bytes_per_point / step_s / 1e9 / HBM_PEAK_GBS, 4),
                      "traces_finite": finite}
    return out


MFMA_BF16_PEAK_TFLOPS = 2500.0


def network_timing(receivers=True):
    """One ensemble member of the screening network at its full shape (207 -> 512 -> 512 -> 512 ->
    151 500 with the layout inputs, 4 -> ... without; bf16 on the matrix cores), timed with hipEvents:
    one forward pass over an inference batch of 1024 and one training step (forward, backward,
    Adam) over a batch of 256; then the 5-member training loop's wall time per member-step.  FLOPs
    are 2 x the MACs per sample forward and 3 x that per training sample.  Returns (roofline dict,
    layer sizes)."""
    from poe2d_amd import surrogate as S
    from poe2d_amd.mlp import Network

    geom = S.CONFIG3
    nin = 4 + (2 * geom["nrec"] + 1 if receivers else 0)
    sizes = [nin, 512, 512, 512, int(np.prod(S.gather_shape(geom)))]
    macs = sum(a * b for a, b in zip(sizes[:-1], sizes[1:]))
    with Network(sizes, members=1, max_batch=1024) as net:
        net.time(256, 2)
        fwd_ms, _ = net.time(1024, 10)
        _, train_ms = net.time(256, 10)
    # the training loop of the 5-member ensemble (train_many: members side by side on two streams,
    # batches gathered from a resident dataset of 2048 synthetic rows): wall time per member-step
    rng = np.random.default_rng(0)
    x = rng.standard_normal((2048, sizes[0])).astype(np.float32)
    t = (0.01 * rng.standard_normal((2048, sizes[-1]))).astype(np.float32)
    with Network(sizes, members=5, max_batch=256) as net:
        net.set_dataset(x, t)
        idx = rng.integers(0, 2048, size=(5, 220 * 256))
        net.train_many(idx[:, :20 * 256], 256, 1e-4, list(range(5)))
        t0 = time.perf_counter()
        net.train_many(idx[:, 20 * 256:], 256, 1e-4, list(range(5)))
        loop_ms = (time.perf_counter() - t0) * 1e3 / (5 * 200)
    del x, t
    fwd_tflops = 2 * macs * 1024 / (fwd_ms / 1e3) / 1e12
    train_tflops = 3 * 2 * macs * 256 / (train_ms / 1e3) / 1e12
    print(f"[bench] network {'-'.join(map(str, sizes))}: forward(1024) {fwd_ms:.3f} ms = {fwd_tflops:.0f} TFLOP/s; "
          f"train step(256) {train_ms:.3f} ms = {train_tflops:.0f} TFLOP/s; 5-member training loop "
          f"{loop_ms:.3f} ms per member-step", file=sys.stderr, flush=True)
    return ({"bound": "mfma", "achieved": round(fwd_tflops, 1), "peak": MFMA_BF16_PEAK_TFLOPS,
             "unit": "TFLOP/s", "frac": round(fwd_tflops / MFMA_BF16_PEAK_TFLOPS, 4), "traffic": None,
             "per_launch": f"one member's forward pass over 1024 samples: 2 x {macs / 1e6:.2f} M MACs each",
             "train_step_tflops": round(train_tflops, 1),
             "train_step_frac": round(train_tflops / MFMA_BF16_PEAK_TFLOPS, 4),
             "forward_ms": round(fwd_ms, 4), "train_step_ms": round(train_ms, 4),
             "train_loop_ms_per_member_step": round(loop_ms, 4),
             "train_loop": "5 members side by side on two streams (poe2d_mlp_train_many), batch 256 gathered "
                           "from a resident dataset, wall time over 200 steps each"}, sizes)


def surrogate_bench(args):
    """Configurations 3 and 4 on the geometry of configuration 3 (401^2 points, 3000 steps decimated
    to 500 samples, 101 receivers x (vx, vy, p) = 151 500 outputs), with the model counts cut to fit
    a few minutes (the options say how many).

    The network part times the full-shape ensemble member of the screening network (207 -> 512 ->
    512 -> 512 -> 151 500 with the layout inputs, 4 -> ... without; bf16 on the matrix cores) with
    hipEvents: one forward pass over an inference batch of 1024 and one training step (forward,
    backward, Adam) over a batch of 256; FLOPs are 2 x the MACs per sample forward (78.2 M with
    the layout inputs) and 3 x that per training sample.  The pipeline part draws models, solves them
    with the batched solver (solver alone: models/s over every candidate), trains the 5-member
    ensemble, calibrates the spread threshold on held-out models, screens candidates (prediction +
    spread-triggered re-solves, the models/s of the surrogate; the gathers stay in HBM) and
    re-solves a random 5 % of the accepted ones to judge it (timed apart); a last screen at the
    first bound delivers the gathers to page-locked host memory instead."""
    from poe2d_amd import surrogate as S

    geom = S.CONFIG3
    net_roofline, sizes = network_timing(bool(args.sur_receivers))
    fwd_ms, fwd_tflops = net_roofline["forward_ms"], net_roofline["achieved"]
    if args.sur_train == 0:
        print(json.dumps({
            "metric": "surrogate network TFLOP/s (bf16 forward pass, batch 1024)", "value": round(fwd_tflops, 1),
            "unit": "TFLOP/s", "n_gpus": 1, "steps": 10, "warmup": 1, "ms_per_step": round(fwd_ms, 4),
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "bf16 (FP32 accumulate)", "data": "synthetic",
            "config": {"workload": f"one ensemble member {sizes[0]}-512-512-512-{sizes[-1]} GELU; forward over "
                                   "1024 samples, training step over 256",
                       "step": "one forward pass"},
            "roofline": net_roofline}), flush=True)
        return

    ntr, nval, ncand = args.sur_train, args.sur_val, args.sur_cand
    params = S.draw_params(ntr + nval, seed=3)
    gathers, solve = S.simulate(params, geom=geom, progress=lambda d, n: d % 1024 == 0 and print(
        f"[surrogate] solved {d} of {n}", file=sys.stderr, flush=True))
    print(f"[surrogate] solver alone: {solve.n} models in {solve.wall_seconds:.1f} s "
          f"({solve.models_per_second:.1f}/s; GPU {solve.gpu_models_per_second:.1f}/s)", file=sys.stderr, flush=True)
    sur = S.Surrogate(geom, members=5, max_batch=1024, receivers=bool(args.sur_receivers))
    t0 = time.perf_counter()
    sur.fit(params[:ntr], gathers[:ntr], args.sur_steps,
            log=lambda k, l: print(f"[surrogate] member {k}: loss {l[-20:].mean():.5f}", file=sys.stderr, flush=True),
            progress=lambda c, n: print(f"[surrogate] training chunk {c + 1} of {n}", file=sys.stderr, flush=True))
    train_s = time.perf_counter() - t0
    quantiles = [float(q) for q in args.sur_quantiles.split(",")]
    _, verr, vspread = sur.calibrate(params[ntr:], gathers[ntr:], bound=0.05, quantile=quantiles[0])
    del gathers
    dump = {"val_params": params[ntr:], "val_err": verr, "val_spread": vspread}
    cand = S.draw_params(ncand, seed=4)
    # the screens' output: the 16 384 gathers stay in HBM (device memory allocated before the
    # screens and reused by each, a serving loop's output buffer), as the solver's inputs are
    # resident in HBM before its timed runs; one more screen delivers them to page-locked host
    # memory, for the rate that includes the copy over the host link
    from poe2d_amd.mlp import DeviceArray, pinned_empty
    out = DeviceArray((ncand,) + sur.shape)
    screens = []
    specs = [(q, None) for q in quantiles]
    if args.sur_confidence > 0:
        specs.append((quantiles[-1], args.sur_confidence))
    for q, conf in specs:
        thr = S.spread_threshold(verr, vspread, 0.05, q, confidence=conf)
        res = sur.screen(cand, thr, out=out, progress=lambda d, n: d % 1024 == 0 and print(
            f"[surrogate] re-solved {d} of {n}", file=sys.stderr, flush=True))
        je = res.judge_errors
        ffb = len(res.fallback) / res.n
        tag = f"p{round(q * 100)}" + ("" if conf is None else f"_ucb{round(conf * 100)}")
        dump.update({f"{tag}_threshold": thr, f"{tag}_fallback": res.fallback, f"{tag}_judged": res.judged,
                     f"{tag}_judge_err": je})
        value = res.models_per_second
        how = ("estimated by" if conf is None else
               f"its {round(conf * 100)} % upper confidence bound (an order statistic) over")
        print(f"[surrogate] bound {tag} <= 5 %: {value:.1f} models/s, fallback {ffb:.4f}", file=sys.stderr, flush=True)
        screens.append({
            "bound": f"p{round(q * 100)} of the error after fallback <= 5 %, {how} {nval} held-out models",
            "models_per_s": round(value, 2),
            "speedup_vs_solver": round(value / solve.models_per_second, 2),
            "speedup_vs_solver_gpu_time": round(value / solve.gpu_models_per_second, 2),
            "fallback_fraction": round(ffb, 4), "screen_s": round(res.screen_seconds, 3),
            "predict_s": round(res.predict_seconds, 3), "fallback_s": round(res.fallback_timing.wall_seconds, 2),
            "judged": len(res.judged), "judge_s": round(res.judge_timing.wall_seconds, 2),
            "judged_rel_l2_p50": round(float(np.median(je)), 4) if len(je) else None,
            "judged_rel_l2_p99": round(float(np.quantile(je, 0.99)), 4) if len(je) else None,
            "after_fallback_rel_l2_p50": round(S.mixture_quantile(je, ffb, 0.5), 4),
            "after_fallback_rel_l2_p99": round(S.mixture_quantile(je, ffb, 0.99), 4)})
    del out
    host_out = pinned_empty((ncand,) + sur.shape)
    thr = S.spread_threshold(verr, vspread, 0.05, quantiles[0])
    res = sur.screen(cand, thr, out=host_out, judge_frac=0.0)
    to_host = {"bound": screens[0]["bound"], "models_per_s": round(res.models_per_second, 2),
               "screen_s": round(res.screen_seconds, 3), "predict_s": round(res.predict_seconds, 3),
               "gathers": f"{ncand * sur.nout * 4 / 1e9:.2f} GB copied to page-locked host memory"}
    print(f"[surrogate] gathers to the host: {res.models_per_second:.1f} models/s", file=sys.stderr, flush=True)
    del host_out
    sur.close()
    if args.sur_dump:
        # the held-out errors and spreads and the screening outcome, for study off the GPU
        np.savez_compressed(args.sur_dump, cand_params=cand, **dump)
    first = screens[0]
    print(json.dumps({
        "metric": "surrogate-screened models/s (configuration 4 screening, configuration 3 geometry)",
        "value": first["models_per_s"], "unit": "models/s", "n_gpus": 1, "steps": 1, "warmup": 0,
        "ms_per_step": round(first["screen_s"] * 1e3, 1), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "bf16 network (FP32 accumulate), f32 solver", "data": "synthetic",
        "config": {"workload": f"{ncand} candidates screened by a 5-member "
                               f"{sur.net_sizes[0]}-512-512-512-{sizes[-1]} GELU "
                               f"ensemble trained {args.sur_steps} steps (batch 256) on {ntr} solver gathers "
                               f"(401^2, 3000 steps / 6, 101 receivers x 3)"
                               + (", inputs: 4 parameters + gas at each receiver, on its path from the source and at the source" if sur.receivers else ""),
                   "step": "one screening pass: prediction + re-solves of the high-spread candidates, "
                           "every gather left in HBM",
                   "solver_alone_models_per_s": round(solve.models_per_second, 2),
                   "solver_alone_gpu_models_per_s": round(solve.gpu_models_per_second, 2),
                   "screens": screens,
                   "screen_gathers_to_host": to_host,
                   "validation_rel_l2_p50": round(float(np.median(verr)), 4),
                   "validation_rel_l2_p99": round(float(np.quantile(verr, 0.99)), 4),
                   "train_s": round(train_s, 1)},
        "roofline": net_roofline,
    }), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3, help="timed full solves")
    ap.add_argument("--warmup", type=int, default=1, help="untimed full solves")
    ap.add_argument("--n", type=int, default=4000, help="NX = NZ (grid has n+1 points per side)")
    ap.add_argument("--nt", type=int, default=None, help="time steps per solve (6000; 3000 for the batches)")
    ap.add_argument("--workload", default="config1",
                    choices=("config1", "continuous", "reference", "batchA", "batchB", "surrogate"),
                    help="config1: the 4001^2 single shot (the headline line); continuous: the same geometry "
                         "over a continuously varying medium (no material palette); reference: the reference "
                         "program's 13600x2440 grid over a synthetic layered model; batchA / batchB: "
                         "configuration 2's batches; surrogate: configurations 3 and 4")
    ap.add_argument("--sur-train", type=int, default=32768,
                    help="surrogate: training models (0: time the network alone)")
    ap.add_argument("--sur-val", type=int, default=2048, help="surrogate: held-out models for the threshold")
    ap.add_argument("--sur-cand", type=int, default=16384,
                    help="surrogate: candidates screened (configuration 4: 16 384)")
    ap.add_argument("--sur-steps", type=int, default=60000, help="surrogate: training steps per member")
    ap.add_argument("--sur-receivers", type=int, default=1,
                    help="surrogate: 1 = per-receiver gas inputs besides the four parameters, 0 = the "
                         "four parameters alone")
    ap.add_argument("--sur-dump", default="", metavar="NPZ",
                    help="surrogate: write the held-out errors and spreads and each screen's fallback, judged "
                         "models and their errors to this file")
    ap.add_argument("--sur-quantiles", default="0.5,0.99",
                    help="surrogate: error quantiles after fallback held within 5 %% by the threshold, one "
                         "screening pass each (the first gives the line's value)")
    ap.add_argument("--sur-confidence", type=float, default=0.95,
                    help="surrogate: one more screen for the last quantile, its threshold set so that the "
                         "quantile's upper confidence bound at this level stays within 5 %% (0: none)")
    ap.add_argument("--cpu-steps", type=int, default=12,
                    help="time steps of the CPU baseline sample (0: skip the CPU leg, the oracle comparison "
                         "included)")
    ap.add_argument("--parity-steps", type=int, default=200,
                    help="steps of the source run compared with the oracle on all host threads (0: skip)")
    ap.add_argument("--fp64", type=int, default=1, help="1: solve again in FP64 and report the FP32 error")
    ap.add_argument("--secondary", type=int, default=1,
                    help="1: add configuration 2's batches (briefly) and the network timing to the line")
    ap.add_argument("--batch-nt", type=int, default=300, help="steps per run of the secondary batch timings")
    ap.add_argument("--continuous-nt", type=int, default=1000,
                    help="steps per run of the secondary continuous-model timing")
    ap.add_argument("--reference-nt", type=int, default=1000,
                    help="steps per run of the secondary timing on the reference program's grid")
    ap.add_argument("--tile", default="auto", metavar="BWxTH",
                    help="tile shape of the step kernel (block width x tile rows); 'auto' runs the "
                         "library's autotuner first")
    ap.add_argument("--schedule", default="auto", choices=("auto", "0", "3"),
                    help="step kernel: 0 = one step per launch, 3 = two steps per launch, auto = the "
                         "faster after autotuning both")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="instead of the benchmark: time a plain device copy of MIB MiB (torch; the copy "
                         "that scales the FETCH_SIZE / WRITE_SIZE counters)")
    args = ap.parse_args()
    args.nt_given = args.nt is not None
    if args.nt is None:
        args.nt = 6000
    if args.copy_probe:
        copy_probe(args.copy_probe)
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (--gpus {args.gpus} requested): every box here has one "
              "GPU; the multi-GPU slab decomposition is a design, not a measured path.", file=sys.stderr)
        sys.exit(2)
    if args.workload == "surrogate":
        surrogate_bench(args)
        return
    if args.workload not in ("config1", "continuous", "reference"):
        batch_bench(args)
        return

    from poe2d_amd import F32, F64
    from poe2d_amd.configs import config1, config1_continuous, reference_grid
    from poe2d_amd.solver import Solver

    if args.workload == "config1":
        def make(nt, lib=None):
            return config1(nt=nt, n=args.n, material_fn=None if lib is None else lib.material_biot)
    elif args.workload == "continuous":
        def make(nt, lib=None):
            return config1_continuous(nt=nt, n=args.n, material_fn=None if lib is None else lib.material_biot)
    else:
        def make(nt, lib=None):
            return reference_grid(nt=nt, material_reference_fn=None if lib is None else lib.material_reference)
    prob = make(args.nt)
    npts = (prob.nx + 1) * (prob.nz + 1)
    cols, rows = strip_counts(prob)
    bytes_step = algorithmic_bytes_per_step(prob.nx, prob.nz, cols, rows)

    solver = Solver(prob, precision=F32)
    palette = solver.material_palette()
    schedules = [0, 3] if args.schedule == "auto" else [int(args.schedule)]
    if args.tile == "auto":
        # autotune each schedule (one and two steps per launch), keep the faster
        best = None
        for sched in schedules:
            solver.reset()
            solver.set_schedule(sched, 256, 12 if sched == 3 else 8)
            t = solver.autotune(10)
            solver.reset()
            solver.set_schedule(sched, *t)
            solver.run(60)
            ms, n = solver.last_run_ms()
            print(f"[bench] schedule {sched}: autotuned {t[0]}x{t[1]}, {ms / n * 1e3:.1f} us per step",
                  file=sys.stderr, flush=True)
            if best is None or ms / n < best[0]:
                best = (ms / n, sched, t)
        _, schedule, tile = best
        solver.set_schedule(schedule, *tile)
        tile_note = "autotuned"
    else:
        schedule = schedules[0]
        tile = tuple(int(v) for v in args.tile.lower().split("x"))
        solver.set_schedule(schedule, *tile)
        tile_note = "fixed"
    for _ in range(args.warmup):
        solver.reset()
        solver.run(args.nt)
    dev_ms, wall_ms = [], []
    for _ in range(args.steps):
        solver.reset()
        t = time.perf_counter()
        solver.run(args.nt)
        wall_ms.append((time.perf_counter() - t) * 1e3)
        ms, n = solver.last_run_ms()
        assert n == args.nt
        dev_ms.append(ms)
    seis = solver.seismograms()
    finite = bool(np.isfinite(seis).all())
    parity_traces = None
    if args.parity_steps > 0:
        # the first parity_steps steps again, for the comparison with the oracle below
        solver.reset()
        solver.run(args.parity_steps)
        parity_traces = solver.seismograms()[..., :args.parity_steps]
    solver.close()

    # the same solve in FP64 (the FP64 build of the same kernels): the error the FP32 result carries
    fp64_err = None
    if args.fp64:
        with Solver(prob, precision=F64) as s64:
            s64.run(args.nt)
            seis64 = s64.seismograms()
            fp64_ms = s64.last_run_ms()[0]
        fp64_err = float(trace_rel_l2(seis, seis64).max())
        print(f"[bench] FP64 solve {fp64_ms:.1f} ms; FP32 vs FP64 max per-trace rel L2 {fp64_err:.3e}",
              file=sys.stderr, flush=True)
        del seis64

    ms_solve = statistics.median(dev_ms)
    gcell = npts * args.nt / (ms_solve / 1e3) / 1e9
    step_ms = ms_solve / args.nt
    achieved = bytes_step / (step_ms / 1e3) / 1e9
    # the committed counter readings are of configuration 1's patchy model
    traffic, traffic_src = recorded_traffic(tile, args.n, schedule) if args.workload == "config1" else (None, None)
    compulsory = compulsory_bytes_per_step(npts, cols * (prob.nz + 1) + rows * (prob.nx + 1), schedule, palette > 0)
    out = {
        "metric": ("Gcell-updates/s (FP32 Biot C-PML time loop, 4001x4001, 6000 steps)" if args.workload == "config1"
                   else f"Gcell-updates/s (FP32 Biot C-PML time loop, {prob.meta['workload']}, {args.nt} steps)"),
        "value": round(gcell * 1.0, 3),
        "unit": "Gcell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_solve, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "fp64_rel_l2_max": None if fp64_err is None else float(f"{fp64_err:.4g}"),
        "parity_rel_l2_max": None,
        "config": {"workload": prob.meta["workload"] + f", {args.nt} steps, "
                               f"{prob.receivers.n} receivers x ({components(prob.receivers.comp_mask)})",
                   "grid": [prob.nx + 1, prob.nz + 1], "nt": args.nt, "npml": prob.npml,
                   "gas_fraction": prob.meta.get("gas_fraction_realised"),
                   "step": "one full solve", "wall_ms_median": round(statistics.median(wall_ms), 3),
                   "traces_finite": finite, "tile": f"{tile[0]}x{tile[1]} ({tile_note})",
                   "schedule": SCHEDULE_NAMES[schedule], "material_palette": palette,
                   "fp64_rel_l2_max": f"max over the {seis.shape[1]} x {seis.shape[0]} traces of the per-trace "
                                      "relative L2 distance between this FP32 solve and the FP64 build's solve of the "
                                      f"same model, all {args.nt} steps"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": None if traffic is None else round(traffic),
                     "frac_measured": None if traffic is None else round(traffic / (step_ms / 1e3) / 1e9 / HBM_PEAK_GBS, 4),
                     "frac_compulsory": round(compulsory / (step_ms / 1e3) / 1e9 / HBM_PEAK_GBS, 4),
                     "compulsory_bytes_per_step": round(compulsory),
                     "kernel": ("fused_step (one launch per time step, 32-step HIP graphs)" if schedule == 0 else
                                "fused_step2 (one launch per two time steps, 32-step HIP graphs); per step: "
                                "half a launch"),
                     "step_us": round(step_ms * 1e3, 2),
                     "per_launch": (("one time step" if schedule == 0 else "two time steps, 2 x") +
                                    f" {bytes_step / 1e9:.4f} GB algorithmic per step "
                                    f"(104 B x {npts} points + 48 B x strip points, {cols} strip columns, "
                                    f"{rows} strip rows)"),
                     "hbm_gbs": None if traffic is None else round(traffic / (step_ms / 1e3) / 1e9, 1),
                     "traffic_source": traffic_src,
                     "note": ("frac: algorithmic bytes, 104 B per cell-update (+ C-PML strips) whatever the "
                              "schedule (the two-step kernel performs two cell-updates per point for each pass "
                              "over HBM, so it can pass 1); frac_measured: the HBM traffic the counters read per "
                              "step (traffic) over the same time; frac_compulsory: the bytes this schedule cannot "
                              "avoid per step (state read and written once per launch, the one-byte material "
                              "index or the coefficients, C-PML records) over the same time")},
    }
    if args.secondary and args.workload == "config1":
        out["secondary"] = {"continuous_model": secondary_continuous(args.n, args.continuous_nt),
                            "reference_grid": secondary_reference_grid(args.reference_nt),
                            "configuration_2": secondary_batches(args.batch_nt),
                            "surrogate_network": network_timing(True)[0]}
    if args.cpu_steps > 0:
        baseline, parity = cpu_leg(prob, make, args.cpu_steps, parity_traces)
        out["cpu_baseline"] = baseline
        if parity is not None:
            out["parity_rel_l2_max"] = float(f"{parity['max']:.4g}")
            out["config"]["parity_rel_l2_max"] = parity["what"]
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
