"""Batched forward over many realisations, captured into one hipGraph.

``Batch`` wraps the genpk_batch_* entry points: each ``add`` is one
forward (clear + deposit + fused r2c/bin) of its own dims^3 plan;
``build`` captures all of them into one graph whose branches run side by
side; ``launch`` replays it with one call.  ``tune`` is the autotuner: it
builds the batch under several configurations (number of parallel
branches x job-to-branch policy), times real graph launches, keeps the
fastest and remembers the choice in a JSON cache keyed by the batch's
shape.  No counterpart in the reference (one realisation per process).
"""
import ctypes
import hashlib
import json
import math
import os
import time

from ._lib import check, lib
from . import _precision, _ptr

__all__ = ["Batch", "tune", "cache_path"]


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


def tune(add_args, candidates=(1, 2, 4, 8), policies=("lpt", "grouped"), reps=3, use_cache=True, log=None):
    """Build the fastest batch for the given jobs.

    add_args: list of dicts of Batch.add keyword arguments.  Every
    (branches, policy) candidate is built and timed over `reps` graph
    launches (median); the winner is returned built, with a report
    {"choice": ..., "timings_ms": ..., "cached": bool}.  The choice is cached
    under the batch's shape (grid sizes, precisions, bins, particle-count
    octaves, device) so later batches of that shape build it directly.
    """
    import torch
    meta = [dict(dims=int(a["dims"]), n=int(a["pos"].shape[0]), precision=a.get("precision", "f32"),
                 nrbins=int(a.get("nrbins") or a["dims"])) for a in add_args]
    sig = _signature(meta, torch)
    path = cache_path()
    cache = _load_cache(path) if use_cache else {}

    def make(branches, policy):
        b = Batch(branches)
        for a in add_args:
            b.add(**a)
        b.build(None if policy == "lpt" else _grouped_assignment(meta, branches))
        return b

    if use_cache and sig in cache:
        ch = cache[sig]
        return make(ch["branches"], ch["policy"]), {"choice": ch, "timings_ms": None, "cached": True}

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
        cache[sig] = choice
        _store_cache(path, cache)
    return best, {"choice": choice, "timings_ms": timings, "cached": False}
