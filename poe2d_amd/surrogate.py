"""Surrogate screening of patchy models (configurations 3 and 4).

A model is four parameters: gas fraction U[0, 0.3], patch radius log-U[2, 20] m (one radius per
model), porosity U[0.15, 0.35] and permeability log-U[1e-14, 1e-11] m^2.  Its gather is what the
solver records for it: the brine rock of configuration 0 with that porosity and permeability, gas
patches of that radius painted at seeded random centres up to that fraction, a Ricker vertical
force at the centre and a line of receivers below it recording v_x, v_y and p, decimated in time.

The patch centres come from one fixed seed (LAYOUT_SEED), so a model is a function of its four
parameters.  With a fresh seed per model the gather would also depend on where the patches fell,
which the four inputs do not say: two layouts with the same parameters differ by about 50 %
relative L2 (pore pressure at a receiver drops about 40-fold when a gas patch covers it), far
beyond what any surrogate of the parameters could reach.

The network (inputs -> 512 -> 512 -> 512 -> outputs, GELU, bf16 on the matrix cores, 5 members)
takes the four parameters and, by default, inputs read off the model's gas mask (layout_inputs:
gas at each receiver, on its path from the source and at the source), and is trained on gathers
the solver itself produced, each component scaled by its RMS over the training set.  Screening
predicts every candidate; a candidate whose ensemble spread exceeds the threshold is re-solved
with the batched solver (fallback), and a random 5 % of the accepted ones are re-solved as well to
judge the surrogate (timed apart from the screening).  Errors are relative L2
over the whole gather after the per-component scaling; after fallback, a re-solved model's error
is zero.

CONFIG3 is the geometry of configuration 3: 401 x 401 points at 0.5 m, 3000 steps of 5e-5 s
decimated by 6 to 500 samples, 101 receivers at x = 50 + 3k, 40 points below the source (3 x 101 x
500 = 151 500 outputs).  SMALL is a reduced geometry with the same layout for quick tests.
"""
import ctypes
import math
import time
from concurrent.futures import ThreadPoolExecutor
from dataclasses import dataclass

import numpy as np

from . import _abi
from .configs import BRINE, GAS, biot_fields
from .material import material_biot
from .mlp import DeviceArray, Network
from .patchy import patchy_mask
from .solver import Batch, Problem, ReceiverSpec, Solver, SourceSpec

CONFIG3 = dict(n=400, dx=0.5, dt=5e-5, nt=3000, dec=6, npml=20, nrec=101, rec_dx=3, rec_dz=40, f0=80.0)
LAYOUT_SEED = 1
SMALL = dict(n=120, dx=0.5, dt=5e-5, nt=600, dec=6, npml=20, nrec=21, rec_dx=3, rec_dz=20, f0=80.0)
RANGES = dict(gas=(0.0, 0.3), radius=(2.0, 20.0), phi=(0.15, 0.35), kappa=(1e-14, 1e-11))


def draw_params(n, seed):
    """n parameter sets (gas, radius [m], phi, kappa [m^2])."""
    rng = np.random.default_rng(seed)
    gas = rng.uniform(*RANGES["gas"], n)
    radius = np.exp(rng.uniform(np.log(RANGES["radius"][0]), np.log(RANGES["radius"][1]), n))
    phi = rng.uniform(*RANGES["phi"], n)
    kappa = np.exp(rng.uniform(np.log(RANGES["kappa"][0]), np.log(RANGES["kappa"][1]), n))
    return np.stack([gas, radius, phi, kappa], axis=1)


INPUT_SCALE = 4.0


def patch_count(params, geom=CONFIG3):
    """Nominal number of patches, gas fraction x model area / (pi radius^2)."""
    p = np.asarray(params, np.float64)
    area = (geom["n"] * geom["dx"]) ** 2
    return p[:, 0] * area / (np.pi * p[:, 1] ** 2)


PATH_SAMPLES = 64


def layout_inputs(params, geom=CONFIG3, workers=0):
    """[n, 2 nrec + 1] float32 from the model's gas mask (the fixed layout, painted for its gas
    fraction and radius; the mask the solver's model is built from): 1 where gas covers receiver
    k, else 0 (columns [0, nrec)); the share of gas along the straight segment from the source to
    receiver k, PATH_SAMPLES points (columns [nrec, 2 nrec)); 1 where gas covers the source (last
    column).  Functions of the parameters."""
    ix, iz = receiver_points(geom)
    n, dx = geom["n"], geom["dx"]
    c = n // 2
    nrec = len(ix)
    t = (np.arange(PATH_SAMPLES) + 0.5) / PATH_SAMPLES
    px = np.rint(c + np.outer(ix - c, t)).astype(np.intp)
    pz = np.rint(c + np.outer(iz - c, t)).astype(np.intp)
    # flat indices into the (n+1) x (n+1) mask: receivers, path samples, source
    flat = np.concatenate([iz * (n + 1) + ix, (pz * (n + 1) + px).ravel(), [c * (n + 1) + c]])
    p = np.asarray(params, np.float64).reshape(-1, 4)
    samples = np.zeros((len(p), len(flat)), np.uint8)  # the mask at the flat indices, per model
    lib = _abi.load_library()
    gas = np.ascontiguousarray(p[:, 0])
    radius = np.ascontiguousarray(p[:, 1])
    flat = np.ascontiguousarray(flat, np.int64)
    # every model's mask painted and sampled in the library, on `workers` host threads
    _abi.check(lib, lib.poe2d_patchy_samples(
        n, n, dx, dx, ctypes.c_uint64(LAYOUT_SEED), len(p), gas.ctypes.data_as(_abi.c_double_p),
        radius.ctypes.data_as(_abi.c_double_p), len(flat), flat.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
        samples.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)), workers), "patchy_samples")
    out = np.empty((len(p), 2 * nrec + 1), np.float32)
    out[:, :nrec] = samples[:, :nrec]
    out[:, nrec:2 * nrec] = samples[:, nrec:nrec * (PATH_SAMPLES + 1)].reshape(len(p), nrec, PATH_SAMPLES).mean(
        axis=2, dtype=np.float32)
    out[:, 2 * nrec] = samples[:, -1]
    return out


def features(params, geom=CONFIG3, receivers=False):
    """Network inputs: log(1 + nominal patch count), log radius, porosity and log permeability,
    each mapped linearly onto [-1, 1] over the parameter ranges, times INPUT_SCALE.  The patch
    count replaces the gas fraction (the two determine each other given the radius): patch k of
    the fixed layout is painted once the count passes about k and covers a receiver once the
    radius passes its distance, so the jumps a receiver sees sit along the input axes.  Wider
    inputs let the first layer form those jumps sharply.  With `receivers`, the layout_inputs: per
    receiver +INPUT_SCALE where gas covers it, -INPUT_SCALE where brine does (pore pressure there
    differs about 40-fold between the two), then per receiver the gas share of its path from the
    source and whether gas covers the source, mapped from [0, 1] onto [-INPUT_SCALE, INPUT_SCALE]."""
    p = np.asarray(params, np.float64)
    lo_r, hi_r = np.log(RANGES["radius"][0]), np.log(RANGES["radius"][1])
    lo_k, hi_k = np.log(RANGES["kappa"][0]), np.log(RANGES["kappa"][1])
    nmax = patch_count([[RANGES["gas"][1], RANGES["radius"][0], 0, 0]], geom)[0]
    f = np.stack([2.0 * np.log1p(patch_count(p, geom)) / np.log1p(nmax) - 1.0,
                  2.0 * (np.log(p[:, 1]) - lo_r) / (hi_r - lo_r) - 1.0,
                  (p[:, 2] - 0.25) / 0.1,
                  2.0 * (np.log(p[:, 3]) - lo_k) / (hi_k - lo_k) - 1.0], axis=1)
    if receivers:
        f = np.concatenate([f, 2.0 * layout_inputs(p, geom) - 1.0], axis=1)
    return (INPUT_SCALE * f).astype(np.float32)


def model_problem(p, seed=LAYOUT_SEED, geom=CONFIG3, dtype=np.float64):
    """The problem of parameter set p: its two materials' coefficients placed by the gas mask, as
    `dtype` arrays (float32 is what an FP32 handle rounds them to itself)."""
    gas, radius, phi, kappa = (float(v) for v in p)
    n, dx = geom["n"], geom["dx"]
    mask = patchy_mask(n, n, dx, dx, gas, seed, radius, radius) if gas > 0 else None
    rock = dict(BRINE, phi=phi, kappa=kappa)
    fluid = dict(GAS, phi=phi, kappa=kappa)
    # the model has two materials: their coefficients (per point, so the same bits as over the
    # whole grid), then placed by the mask
    pair = material_biot(**biot_fields(1, 0, np.array([[False, True]]), rock, fluid))
    shape = (n + 1, n + 1)
    coeffs = {}
    for k, v in pair.items():
        a = np.full(shape, v[0, 0], dtype=dtype)
        if mask is not None:
            a[mask.astype(bool, copy=False)] = v[0, 1]
        coeffs[k] = a
    c = n // 2
    ix, iz = receiver_points(geom)
    rec = ReceiverSpec(ix=ix, iz=iz, decimate=geom["dec"])
    return Problem(nx=n, nz=n, dx=dx, dz=dx, dt=geom["dt"], nt=geom["nt"], coeffs=coeffs,
                   source=SourceSpec(kind=_abi.SRC_FORCE_Y_RICKER, ix=c, iz=c, amp=1.0e10),
                   receivers=rec, npml=geom["npml"], f0=geom["f0"])


def receiver_points(geom=CONFIG3):
    c = geom["n"] // 2
    k = np.arange(geom["nrec"])
    ix = (c - geom["rec_dx"] * (geom["nrec"] // 2) + geom["rec_dx"] * k).astype(np.int32)
    return ix, np.full(geom["nrec"], c + geom["rec_dz"], np.int32)


TILE = (128, 24)  # starting tile shape of the batched two-step kernel (then autotuned)


def gather_shape(geom=CONFIG3):
    return (3, geom["nrec"], -(-geom["nt"] // geom["dec"]))


@dataclass
class SolveTiming:
    n: int
    wall_seconds: float     # model building on the host (overlapped) + GPU solves + read-back
    gpu_seconds: float      # time-step launches only, as the batches report it

    @property
    def models_per_second(self):
        return self.n / self.wall_seconds

    @property
    def gpu_models_per_second(self):
        return self.n / self.gpu_seconds


class SolverPool:
    """Solver handles of one geometry and the batches that advance them, reused from model to model:
    a model enters a handle through poe2d_set_coeffs (its coefficients replaced, the state reset), so
    after the first batch no device memory is allocated per model."""

    def __init__(self, problems):
        self.solvers = []
        self.batches = {}
        try:
            for p in problems:
                s = Solver(p)
                s.set_schedule(3, *TILE)  # two steps per launch (bitwise equal to one)
                self.solvers.append(s)
        except Exception:
            self.close()
            raise

    @property
    def size(self):
        return len(self.solvers)

    def load(self, problems, executor=None):
        """The first len(problems) members take these models (same geometry), on the executor's
        threads when one is given (each handle's set-up is independent of the others')."""
        pairs = list(zip(self.solvers, problems))
        if executor is None:
            for s, p in pairs:
                s.set_coeffs(p.coeffs)
        else:
            list(executor.map(lambda sp: sp[0].set_coeffs(sp[1].coeffs), pairs))

    def has_batch(self, k):
        return k in self.batches

    def batch(self, k):
        """The batch of the first k members.  On first use it is made, its tile shape autotuned and
        its step graph captured by a warm-up run (the members are reset afterwards): nothing else
        may use the GPU from another thread meanwhile, as HIP refuses other threads' calls while a
        stream captures.  Later runs replay the graph."""
        if k not in self.batches:
            b = Batch(self.solvers[:k])
            self.batches[k] = b
            b.set_schedule(*TILE)
            b.autotune(4)  # timed once per batch class in the process, then cached
            b.run(32)      # one graph of 16 step pairs: captured here
            for s in self.solvers[:k]:
                s.reset()
        return self.batches[k]

    def close(self):
        for b in self.batches.values():
            b.close()
        self.batches = {}
        for s in self.solvers:
            s.close()
        self.solvers = []


class SimulatePools:
    """The two pools simulate() alternates between (one solving on the GPU while the host loads the
    other), kept across calls of one geometry."""

    def __init__(self, batch=64):
        self.batch_size = batch
        self.pools = [None, None]

    def get(self, i, problems, executor=None):
        """Pool i holding `problems` in its first members: made for them, or loaded with them."""
        p = self.pools[i]
        if p is None or p.size < len(problems):
            if p is not None:
                p.close()
            p = self.pools[i] = SolverPool(problems)
        else:
            p.load(problems, executor)
        return p

    def close(self):
        for p in self.pools:
            if p is not None:
                p.close()
        self.pools = [None, None]


def simulate(params, seeds=None, geom=CONFIG3, batch=64, workers=8, progress=None, pools=None):
    """Gathers [n, 3, nrec, nsamp] (v_x, v_y, p) of the models, solved in batches of `batch`
    handles advanced together; returns (gathers, SolveTiming).  Host threads build the models'
    coefficients ahead, and two pools of handles alternate: while one batch runs on the GPU the host
    loads the next models into the other pool (poe2d_set_coeffs) and reads back the batch before.
    seeds: patch-layout seed per model (LAYOUT_SEED for all when None); progress: called with
    (models done, n) after each batch; pools: a SimulatePools of this geometry to reuse (made and
    closed here when None)."""
    t0 = time.perf_counter()
    n = len(params)
    seeds = np.full(n, LAYOUT_SEED) if seeds is None else np.asarray(seeds)
    out = np.zeros((n,) + gather_shape(geom), np.float32)
    own = pools is None
    if own:
        pools = SimulatePools(batch)
    batch = pools.batch_size
    gpu_ms = 0.0
    build = lambda k: model_problem(params[k], seeds[k], geom, np.float32)
    chunks = [range(i, min(i + batch, n)) for i in range(0, n, batch)]

    def run(b):
        b.run(geom["nt"])
        return b.last_run_ms()[0]

    try:
        with ThreadPoolExecutor(workers) as builders, ThreadPoolExecutor(1) as runner:
            pending = [builders.submit(build, k) for k in chunks[0]] if chunks else []
            running = None  # (future of the batch run, pool, chunk)
            done = 0
            for ci, chunk in enumerate(chunks):
                problems = [f.result() for f in pending]
                pool = pools.get(ci % 2, problems, builders)
                del problems
                pending = [builders.submit(build, k) for k in chunks[ci + 1]] if ci + 1 < len(chunks) else []
                if not pool.has_batch(len(chunk)) and running is not None:
                    # a new batch captures its step graph: the GPU work before it finishes first
                    gpu_ms, done = _collect(running, out, gpu_ms, done, n, progress, builders)
                    running = None
                b = pool.batch(len(chunk))
                fut = runner.submit(run, b)  # queued behind the batch before, if it still runs
                if running is not None:
                    gpu_ms, done = _collect(running, out, gpu_ms, done, n, progress, builders)
                running = (fut, pool, chunk)
            if running is not None:
                gpu_ms, done = _collect(running, out, gpu_ms, done, n, progress, builders)
    finally:
        if own:
            pools.close()
    return out, SolveTiming(n, time.perf_counter() - t0, gpu_ms * 1e-3)


def _collect(running, out, gpu_ms, done, n, progress, executor):
    """Wait for a batch run and read its gathers back (per handle, on the executor's threads)."""
    fut, pool, chunk = running
    gpu_ms += fut.result()

    def read(k):
        out[chunk[k]] = pool.solvers[k].seismograms()

    list(executor.map(read, range(len(chunk))))
    done += len(chunk)
    if progress:
        progress(done, n)
    return gpu_ms, done


def after_fallback(err, spread, threshold):
    """Error of each model once screening is done: the surrogate's error where its spread is at most
    `threshold`, zero where the solver re-solved it."""
    return np.where(np.asarray(spread) <= threshold, err, 0.0)


def mixture_quantile(judge_errors, fallback_fraction, q):
    """Quantile q of the errors after fallback over all candidates, from the judged sample of the
    accepted ones: a share `fallback_fraction` is exact (zero error), the rest distributed like the
    judged errors."""
    if q <= fallback_fraction or len(judge_errors) == 0:
        return 0.0
    return float(np.quantile(judge_errors, (q - fallback_fraction) / (1.0 - fallback_fraction)))


def quantile_upper_rank(n, q, confidence):
    """Rank k (1-based) of the order statistic X_(k) of n samples that bounds the q-quantile of their
    distribution from above with probability >= confidence: the smallest k with P(Binomial(n, q) <=
    k - 1) >= confidence (the number of samples below the quantile is Binomial(n, q))."""
    cdf = 0.0
    logq, log1mq = math.log(q), math.log1p(-q)
    for j in range(n + 1):
        cdf += math.exp(math.lgamma(n + 1) - math.lgamma(j + 1) - math.lgamma(n - j + 1) + j * logq + (n - j) * log1mq)
        if cdf >= confidence:
            return j + 1
    return n


def spread_threshold(err, spread, bound=0.05, quantile=0.99, confidence=None):
    """Largest spread threshold (one of the given spreads) at which the `quantile` of the errors after
    fallback stays within `bound`; -inf when even the smallest spread exceeds it (everything falls
    back).  With `confidence`, the quantile is bounded from above at that confidence by an order
    statistic of the errors after fallback (quantile_upper_rank) instead of estimated by it."""
    err = np.asarray(err, np.float64)
    spread = np.asarray(spread, np.float64)
    rank = None if confidence is None else quantile_upper_rank(len(err), quantile, confidence)
    best = -np.inf
    for t in np.unique(spread):
        a = after_fallback(err, spread, t)
        v = np.quantile(a, quantile) if rank is None else np.partition(a, rank - 1)[rank - 1]
        if v <= bound:
            best = t
    return float(best)


@dataclass
class ScreenResult:
    n: int
    fallback: np.ndarray        # indices re-solved because of their spread
    judged: np.ndarray          # indices re-solved to judge the surrogate
    judge_errors: np.ndarray    # surrogate error on the judged models
    screen_seconds: float       # prediction + fallback re-solves (wall)
    predict_seconds: float
    fallback_timing: SolveTiming
    judge_timing: SolveTiming
    gathers: np.ndarray = None  # [n, 3, nrec, nsamp]: the surrogate's prediction where accepted,
                                # the solver's gather where re-solved for its spread (a numpy
                                # array, or the DeviceArray the screen was given)

    @property
    def models_per_second(self):
        return self.n / self.screen_seconds


class Surrogate:
    """Ensemble network from features(params, geom, receivers) to a model's gather.  The network's
    targets are the gathers with each component divided by its RMS over the training set.
    receivers: add the per-receiver gas inputs (features)."""

    def __init__(self, geom=CONFIG3, members=5, hidden=512, max_batch=1024, seed=1, receivers=True):
        self.geom = dict(geom)
        self.receivers = bool(receivers)
        self.shape = gather_shape(geom)
        self.nout = int(np.prod(self.shape))
        self.net_sizes = [4 + (2 * geom["nrec"] + 1 if self.receivers else 0), hidden, hidden, hidden, self.nout]
        self.net = Network(self.net_sizes, members=members, max_batch=max_batch, seed=seed)
        self.pools = SimulatePools()  # solver handles for the re-solves, kept across screens
        self.scale = np.ones(self.shape[0], np.float32)    # RMS of each component of the gathers

    def _scaled(self, g):
        return (g / self.scale[None, :, None, None]).reshape(len(g), -1)

    def fit(self, params, gathers, steps, batch=256, lr=1e-3, lr_final=1e-5, chunks=20, seed=1, log=None,
            progress=None):
        """Train every member on its own shuffles of the data, the learning rate falling from `lr`
        to `lr_final` on a cosine over `chunks` pieces, the members side by side (train_many);
        returns the loss curves.  log(member, losses) is called for each member at the end,
        progress(chunk, chunks) after each chunk."""
        g = np.asarray(gathers, np.float32)
        rms = np.sqrt(np.mean(g.astype(np.float64) ** 2, axis=(0, 2, 3)))
        self.scale = np.where(rms > 0, rms, 1.0).astype(np.float32)
        # predictions come back in the gathers' units: the device multiplies by the scale
        self.net.set_output_scale(np.repeat(self.scale, self.nout // 3))
        self.net.set_dataset(features(params, self.geom, self.receivers), self._scaled(g))
        rng = np.random.default_rng(seed)
        n = len(params)
        cuts = np.linspace(0, steps, chunks + 1).round().astype(int)
        need = steps * batch
        members = list(range(self.net.members))
        idx = np.stack([np.concatenate([rng.permutation(n) for _ in range(need // n + 1)])[:need] for _ in members])
        losses = []
        for c in range(chunks):
            if cuts[c + 1] == cuts[c]:
                continue
            rate = lr_final + 0.5 * (lr - lr_final) * (1 + np.cos(np.pi * c / chunks))
            losses.append(self.net.train_many(idx[:, cuts[c] * batch:cuts[c + 1] * batch], batch, rate, members))
            if progress:
                progress(c, chunks)
        curves = list(np.concatenate(losses, axis=1)) if losses else [np.zeros(0, np.float32) for _ in members]
        if log:
            for member in members:
                log(member, curves[member])
        return curves

    def predict(self, params, out=None, chunk=2048):
        """(gathers [n, 3, nrec, nsamp], ensemble spread [n] in units of the scaled targets).  out: a
        float32 array of that shape to fill (page-locked memory, mlp.pinned_empty, copies fastest),
        or an mlp.DeviceArray of that shape, where the gathers stay in HBM.  The network inputs of
        the next `chunk` candidates are formed on a host thread while the device predicts the
        current ones."""
        params = np.asarray(params, np.float64).reshape(-1, 4)
        n = len(params)
        g = np.empty((n,) + self.shape, np.float32) if out is None else out
        on_device = isinstance(g, DeviceArray)
        if g.shape != (n,) + self.shape or g.dtype != np.float32 or not (on_device or g.flags.c_contiguous):
            raise ValueError(f"out must be a C-contiguous float32 array of shape {(n,) + self.shape}")
        flat = g if on_device else g.reshape(n, self.nout)
        spread = np.empty(n, np.float32)
        with ThreadPoolExecutor(1) as pool:
            nxt = pool.submit(features, params[:chunk], self.geom, self.receivers)
            for lo in range(0, n, chunk):
                hi = min(n, lo + chunk)
                x = nxt.result()
                if hi < n:
                    nxt = pool.submit(features, params[hi:hi + chunk], self.geom, self.receivers)
                _, spread[lo:hi] = self.net.predict(x, out=flat.rows(lo, hi) if on_device else flat[lo:hi])
        return g, spread

    def errors(self, pred, true):
        """Relative L2 per model over the whole gather, each component divided by its RMS over the
        training set."""
        a, b = self._scaled(pred), self._scaled(true)
        return np.linalg.norm(a - b, axis=1) / np.maximum(np.linalg.norm(b, axis=1), 1e-30)

    def calibrate(self, params, gathers, bound=0.05, quantile=0.99):
        """(threshold, errors, spreads) on a validation set: spread_threshold of its errors."""
        pred, spread = self.predict(params)
        err = self.errors(pred, gathers)
        return spread_threshold(err, spread, bound, quantile), err, spread

    def screen(self, params, threshold, seeds=None, judge_frac=0.05, seed=7, progress=None, out=None):
        """Predict every candidate, re-solve those whose spread exceeds `threshold`, then re-solve
        a random `judge_frac` of the others to measure the surrogate's error (timed apart).  The
        result's gathers hold, per candidate, the prediction or (re-solved) the solver's gather;
        out: the array to hold them (as for predict; a DeviceArray keeps them in HBM, the re-solved
        gathers copied in), reused by a caller that screens repeatedly.  judge_frac 0 judges
        nothing.  progress: passed to simulate for both sets of re-solves."""
        on_device = isinstance(out, DeviceArray)
        t0 = time.perf_counter()
        pred, spread = self.predict(params, out=out)
        t_pred = time.perf_counter() - t0
        fallback = np.nonzero(spread > threshold)[0]
        fb = SolveTiming(0, 0.0, 0.0)
        gathers = pred  # re-solved rows replaced in place; the judged rows are accepted ones
        if len(fallback):
            solved, fb = simulate(params[fallback], None if seeds is None else seeds[fallback],
                                  self.geom, progress=progress, pools=self.pools)
            if on_device:
                gathers.set_rows(fallback, solved)
            else:
                gathers[fallback] = solved
        t_screen = time.perf_counter() - t0
        accepted = np.setdiff1d(np.arange(len(params)), fallback)
        rng = np.random.default_rng(seed)
        nj = max(1, int(round(judge_frac * len(accepted)))) if judge_frac > 0 and len(accepted) else 0
        judged = np.sort(rng.choice(accepted, nj, replace=False)) if nj else np.zeros(0, int)
        err = np.zeros(0)
        jt = SolveTiming(0, 0.0, 0.0)
        if nj:
            true, jt = simulate(params[judged], None if seeds is None else seeds[judged], self.geom,
                                progress=progress, pools=self.pools)
            err = self.errors(pred.to_host(judged) if on_device else pred[judged], true)
        return ScreenResult(len(params), fallback, judged, err, t_screen, t_pred, fb, jt, gathers)

    def close(self):
        self.pools.close()
        self.net.close()
