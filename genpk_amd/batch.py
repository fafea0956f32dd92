"""Batched forward over many realisations, captured into one hipGraph.

``Batch`` wraps the genpk_batch_* entry points: each ``add`` is one
forward (clear + deposit + fused r2c/bin) of its own dims^3 plan;
``build`` captures all of them into one graph whose branches run side by
side; ``launch`` replays it with one call.  ``tune`` is the autotuner, in
two stages: per grid size and precision it times a plain forward of the
largest job of that size under every deposit tile shape and FFT radix
ordering (``tune_plan_options``); with those options it then builds the
batch under several configurations (number of parallel branches x
job-to-branch policy) and times real graph launches.  It keeps the fastest
and remembers both choices in a JSON cache (per grid size, precision and
device; per batch shape and device).  No counterpart in the reference (one
realisation per process).
"""
import ctypes
import hashlib
import json
import math
import os
import time

from ._lib import check, lib
from . import _precision, _ptr

__all__ = ["Batch", "tune", "tune_plan_options", "cache_path", "TILE_CANDIDATES", "FFT_VARIANTS"]


class Batch:
    def __init__(self, branches=4, device=None):
        import torch
        self.torch = torch
        self.device = torch.device("cuda", torch.cuda.current_device() if device is None else device)
        self.branches = int(branches)
        self._h = ctypes.c_void_p()
        self.jobs = []       # (kwargs, outputs, keep-alive tensors)
        self.built = False
        with torch.cuda.device(self.device):
            stream = torch.cuda.current_stream(self.device).cuda_stream
            check(lib().genpk_batch_create(ctypes.byref(self._h), self.branches, ctypes.c_void_p(stream)))

    def add(self, pos, boxsize, dims, nrbins=None, masses=None, mass=1.0, total_mass=None, precision="f32"):
        """Add one forward job; returns its index.  pos: float32 [n, 3] on the
        device (kept alive by the batch; its contents are read at every launch)."""
        t = self.torch
        assert not self.built, "batch already built"
        assert pos.is_cuda and pos.dtype == t.float32 and pos.is_contiguous() and pos.dim() == 2 and pos.shape[1] == 3
        if masses is not None:
            assert masses.is_cuda and masses.dtype == t.float32 and masses.is_contiguous() and masses.shape == (pos.shape[0],)
        nrbins = int(dims if nrbins is None else nrbins)
        n = pos.shape[0]
        if total_mass is None:
            total_mass = float(masses.double().sum()) if masses is not None else float(mass) * n
        f64 = dict(dtype=t.float64, device=self.device)
        out = (t.zeros(nrbins, **f64), t.zeros(nrbins, dtype=t.int32, device=self.device), t.zeros(nrbins, **f64))
        check(lib().genpk_batch_add(self._h, int(dims), _precision(precision), float(boxsize), n, _ptr(pos),
                                    _ptr(masses), float(mass), nrbins, float(total_mass), _ptr(out[0]),
                                    _ptr(out[1]), _ptr(out[2])))
        self.jobs.append((dict(dims=int(dims), n=n, precision=precision, nrbins=nrbins), out, (pos, masses)))
        return len(self.jobs) - 1

    def plan_options(self, dims, precision, tiles, fft_variant):
        """Deposit tile edges and FFT radix variant of the batch's plans of one
        grid size and precision (Plan.set_options), before build()."""
        tx, ty, tz = (int(t) for t in tiles)
        check(lib().genpk_batch_plan_options(self._h, int(dims), _precision(precision), tx, ty, tz,
                                             int(fft_variant)))

    def build(self, assignment=None):
        """Capture the graph.  assignment: branch of each job (list of ints) or
        None for longest-processing-time-first balancing."""
        arr = None
        if assignment is not None:
            assert len(assignment) == len(self.jobs)
            arr = (ctypes.c_int * len(assignment))(*[int(a) for a in assignment])
        check(lib().genpk_batch_build(self._h, arr))
        self.built = True
        return self

    def launch(self):
        check(lib().genpk_batch_launch(self._h))

    def branch_of(self, job):
        b = ctypes.c_int()
        check(lib().genpk_batch_branch_of(self._h, int(job), ctypes.byref(b)))
        return b.value

    def results(self, job=None):
        """(power, count, keffs) of one job, or the list of all."""
        if job is not None:
            return self.jobs[job][1]
        return [j[1] for j in self.jobs]

    def close(self):
        if self._h:
            lib().genpk_batch_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------- autotuner

def _cost(job):
    c = 4 if job["precision"] in ("f32", 4) else 8
    d = job["dims"]
    return 60.0 * job["n"] + 6.0 * c * d * d * (d + 2) + 2e6


def _grouped_assignment(jobs, branches):
    """Jobs of one grid size on as few branches as their cost allows: sort by
    (dims, cost) and cut the sequence into `branches` runs of equal cost."""
    order = sorted(range(len(jobs)), key=lambda i: (jobs[i]["dims"], -_cost(jobs[i])))
    total = sum(_cost(j) for j in jobs)
    out = [0] * len(jobs)
    acc, k = 0.0, 0
    for i in order:
        c = _cost(jobs[i])
        if acc + c / 2 > total * (k + 1) / branches and k < branches - 1:
            k += 1
        out[i] = k
        acc += c
    return out


def cache_path():
    p = os.environ.get("GENPK_AUTOTUNE_CACHE")
    if p:
        return p
    return os.path.join(os.path.expanduser("~"), ".cache", "genpk_amd", "autotune.json")


def _signature(jobs, torch):
    shape = sorted((j["dims"], str(j["precision"]), j["nrbins"], int(math.ceil(math.log2(max(j["n"], 1)))))
                   for j in jobs)
    dev = torch.cuda.get_device_name(torch.cuda.current_device())
    return hashlib.sha1(json.dumps([dev, shape]).encode()).hexdigest()


def _load_cache(path):
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _store_cache(path, cache):
    try:
        os.makedirs(os.path.dirname(path), exist_ok=True)
        tmp = path + ".tmp"
        with open(tmp, "w") as f:
            json.dump(cache, f, indent=1)
        os.replace(tmp, path)
    except OSError:
        pass  # a read-only home only loses the cache


def _time_batch(b, torch, reps):
    b.launch()
    torch.cuda.synchronize()
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        b.launch()
        torch.cuda.synchronize()
        ts.append(time.perf_counter() - t0)
    ts.sort()
    return ts[len(ts) // 2]


# deposit tile shapes tried per grid size (x, y, z edges; clamped to dims)
TILE_CANDIDATES = ((16, 32, 32), (8, 32, 32), (16, 16, 32), (8, 16, 32))
FFT_VARIANTS = (0, 1, 2, 3)


def _plan_key(dims, precision, torch):
    dev = torch.cuda.get_device_name(torch.cuda.current_device())
    return "plan:" + json.dumps([dev, int(dims), "f32" if _precision(precision) == 4 else "f64"])


def tune_plan_options(add_args, tiles=TILE_CANDIDATES, fft_variants=FFT_VARIANTS, reps=3, use_cache=True, log=None):
    """Stage 1 of the autotuner: for every (dims, precision) of the jobs,
    time a plain forward (clear + deposit + fused r2c/bin, synchronised) of
    the largest job of that size under each deposit tile shape and FFT radix
    variant (median of `reps` after one warm-up), and keep the fastest.
    Variants that give the same radices as an earlier one are skipped.
    Returns {(dims, precision): {"tiles", "fft_variant", "radices", ...}}."""
    import torch
    from . import Plan, radix_plan
    cache = _load_cache(cache_path()) if use_cache else {}
    groups = {}
    for a in add_args:
        key = (int(a["dims"]), _precision(a.get("precision", "f32")))
        if key not in groups or a["pos"].shape[0] > groups[key]["pos"].shape[0]:
            groups[key] = a
    out = {}
    for (d, prec), a in sorted(groups.items()):
        ck = _plan_key(d, prec, torch)
        if use_cache and ck in cache:
            out[(d, prec)] = dict(cache[ck], cached=True)
            continue
        n = a["pos"].shape[0]
        nrbins = int(a.get("nrbins") or d)
        mass = float(a.get("mass", 1.0))
        tm = a.get("total_mass") or mass * n
        seen, timings, best = set(), {}, None
        plan = Plan(d, prec)
        for v in fft_variants:
            # the z passes (d / 2) run ordering 0 under variant 3
            rad = (tuple(radix_plan(d, v)), tuple(radix_plan(d // 2, 0 if v == 3 else v)))
            if rad in seen:
                continue
            seen.add(rad)
            for t in tiles:
                plan.set_options(t, v)
                eff = plan.options()["tiles"]

                def fwd():
                    plan.clear()
                    plan.deposit(a["pos"], a["boxsize"], a.get("masses"), mass)
                    plan.power(nrbins, mass=tm)
                    torch.cuda.synchronize()

                fwd()
                ts = []
                for _ in range(reps):
                    t0 = time.perf_counter()
                    fwd()
                    ts.append(time.perf_counter() - t0)
                ms = sorted(ts)[reps // 2] * 1e3
                name = f"{eff[0]}x{eff[1]}x{eff[2]}/v{v}"
                timings[name] = round(ms, 4)
                if log:
                    log(f"autotune d={d} {'f32' if prec == 4 else 'f64'}: tiles {name}: {ms:.3f} ms")
                if best is None or ms < best[0]:
                    best = (ms, list(eff), v, [list(rad[0]), list(rad[1])])
        del plan
        torch.cuda.empty_cache()
        choice = {"tiles": best[1], "fft_variant": best[2], "radices_d": best[3][0], "radices_half": best[3][1],
                  "ms": round(best[0], 4), "timings_ms": timings, "representative_n": int(n)}
        out[(d, prec)] = dict(choice, cached=False)
        if use_cache:
            cache[ck] = choice
    if use_cache:
        _store_cache(cache_path(), cache)
    return out


def tune(add_args, candidates=(1, 2, 4, 8), policies=("lpt", "grouped"), reps=3, use_cache=True, log=None,
         tiles=TILE_CANDIDATES, fft_variants=FFT_VARIANTS):
    """Build the fastest batch for the given jobs.

    add_args: list of dicts of Batch.add keyword arguments.  Stage 1 picks
    each grid size's deposit tile shape and FFT radix variant
    (tune_plan_options); stage 2 builds every (branches, policy) candidate
    with those options and times it over `reps` graph launches (median).
    The winner is returned built, with a report {"per_dims": ..., "choice":
    ..., "timings_ms": ..., "cached": bool}.  Both stages cache their choices
    (per grid size, precision and device; per batch shape and device).
    """
    import torch
    meta = [dict(dims=int(a["dims"]), n=int(a["pos"].shape[0]), precision=a.get("precision", "f32"),
                 nrbins=int(a.get("nrbins") or a["dims"])) for a in add_args]
    per_dims = tune_plan_options(add_args, tiles, fft_variants, reps, use_cache, log)
    sig = _signature(meta, torch)
    path = cache_path()
    cache = _load_cache(path) if use_cache else {}

    def make(branches, policy):
        b = Batch(branches)
        for a in add_args:
            b.add(**a)
        for (d, prec), ch in per_dims.items():
            b.plan_options(d, prec, ch["tiles"], ch["fft_variant"])
        b.build(None if policy == "lpt" else _grouped_assignment(meta, branches))
        return b

    report_dims = {f"{d}/{'f32' if p == 4 else 'f64'}": ch for (d, p), ch in per_dims.items()}
    if use_cache and sig in cache:
        ch = cache[sig]
        return make(ch["branches"], ch["policy"]), {"per_dims": report_dims, "choice": ch, "timings_ms": None,
                                                    "cached": True}

    best, best_t, timings = None, None, {}
    for br in candidates:
        for pol in policies:
            if br == 1 and pol != "lpt":
                continue  # one branch: the policies coincide
            b = make(br, pol)
            t = _time_batch(b, torch, reps)
            timings[f"{br}x{pol}"] = round(t * 1e3, 3)
            if log:
                log(f"autotune {br} branches, {pol}: {t * 1e3:.3f} ms")
            if best_t is None or t < best_t:
                if best is not None:
                    best.close()
                best, best_t, choice = b, t, {"branches": br, "policy": pol}
            else:
                b.close()
            torch.cuda.empty_cache()
    if use_cache:
        cache = _load_cache(path)
        cache[sig] = choice
        _store_cache(path, cache)
    return best, {"per_dims": report_dims, "choice": choice, "timings_ms": timings, "cached": False}
