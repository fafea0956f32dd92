"""Batched forward (one hipGraph over many uneven jobs) against plain plan
runs of the same jobs, through the C ABI.

Bars: counts and keffs equal (integer mode counts, per-bin sums in a fixed
order); P within 1e-12 relative (the binning adds FP64 partial sums with LDS
atomics, whose order varies from run to run, as the reference's
thread-private bin merge does).  Every graph replay is checked, with eager
work on the device between replays.
"""
import numpy as np
import pytest

import genpk_amd
from genpk_amd import batch as gb
from genpk_amd import synth

pytestmark = pytest.mark.gpu


def announce(*a):
    print("CASE", *a, flush=True)


def _jobs(torch):
    """Uneven small jobs: dims 32/48/64, f32 and f64, with and without
    masses, one empty particle set."""
    specs = [(20, 32, "f32", False), (30, 64, "f64", False), (12, 48, "f32", True), (0, 32, "f64", False),
             (24, 64, "f32", False), (16, 48, "f64", True), (28, 32, "f32", False)]
    jobs = []
    for i, (ns, d, prec, with_m) in enumerate(specs):
        box = 50.0 + 10 * i
        if ns:
            pos = synth.gaussian_lattice_torch(ns, box, d, 0.3, seed=200 + i)
        else:
            pos = torch.zeros((0, 3), dtype=torch.float32, device="cuda")
        m = None
        if with_m:
            g = torch.Generator(device="cuda")
            g.manual_seed(300 + i)
            m = (torch.rand(pos.shape[0], generator=g, device="cuda") + 0.5).contiguous()
        tm = float(m.double().sum()) if m is not None else float(max(pos.shape[0], 1))
        jobs.append(dict(pos=pos, boxsize=box, dims=d, nrbins=d, masses=m, total_mass=tm, precision=prec))
    return jobs


def _plain(job, options=None):
    p = genpk_amd.Plan(job["dims"], job["precision"])
    if options is not None:
        p.set_options(options["tiles"], options["fft_variant"])
    p.clear()
    p.deposit(job["pos"], job["boxsize"], job["masses"])
    P, c, k = p.power(job["nrbins"], mass=job["total_mass"])
    return P.cpu().numpy(), c.cpu().numpy(), k.cpu().numpy()


@pytest.mark.parametrize("branches", [1, 3])
def test_batch_matches_plain(gpu, branches):
    import torch
    announce("batch_vs_plain", f"branches={branches}")
    jobs = _jobs(torch)
    ref = [_plain(j) for j in jobs]
    b = gb.Batch(branches)
    for j in jobs:
        b.add(**j)
    b.build()
    assert {b.branch_of(i) for i in range(len(jobs))} <= set(range(branches))
    for rep in range(3):
        for out in b.results():
            out[0].fill_(-1.0)
        b.launch()
        torch.cuda.synchronize()
        for i, (j, r) in enumerate(zip(jobs, ref)):
            P, c, k = (t.cpu().numpy() for t in b.results(i))
            assert np.array_equal(c, r[1]), (rep, i)
            assert np.array_equal(k, r[2]), (rep, i)
            assert np.allclose(P, r[0], rtol=1e-12, atol=0), (rep, i)
        # eager device work between replays
        _plain(jobs[1])
    b.close()


def test_batch_new_positions_between_launches(gpu):
    """A launch reads the positions' current contents."""
    import torch
    announce("batch_new_positions")
    d, box = 32, 40.0
    pos = synth.gaussian_lattice_torch(16, box, d, 0.3, seed=1)
    b = gb.Batch(2)
    b.add(pos, box, d, total_mass=float(pos.shape[0]))
    b.build()
    b.launch()
    torch.cuda.synchronize()
    first = b.results(0)[0].clone()
    pos.copy_(synth.gaussian_lattice_torch(16, box, d, 0.3, seed=2))
    b.launch()
    torch.cuda.synchronize()
    job = dict(pos=pos, boxsize=box, dims=d, nrbins=d, masses=None, total_mass=float(pos.shape[0]), precision="f32")
    assert np.allclose(b.results(0)[0].cpu().numpy(), _plain(job)[0], rtol=1e-12, atol=0)
    assert not torch.equal(first, b.results(0)[0])
    b.close()


def test_batch_explicit_assignment_and_autotune(gpu, tmp_path, monkeypatch):
    import torch
    announce("batch_autotune")
    monkeypatch.setenv("GENPK_AUTOTUNE_CACHE", str(tmp_path / "tune.json"))
    jobs = _jobs(torch)
    ref = [_plain(j) for j in jobs]
    b = gb.Batch(2)
    for j in jobs:
        b.add(**j)
    b.build([i % 2 for i in range(len(jobs))])
    assert [b.branch_of(i) for i in range(len(jobs))] == [i % 2 for i in range(len(jobs))]
    b.launch()
    torch.cuda.synchronize()
    for i, j in enumerate(jobs):
        assert np.array_equal(b.results(i)[1].cpu().numpy(), ref[i][1])
    b.close()
    best, rep = gb.tune(jobs, candidates=(1, 2), reps=2)
    assert rep["cached"] is False and rep["choice"]["branches"] in (1, 2)
    assert set(rep["per_dims"]) == {f"{j['dims']}/{j['precision']}" for j in jobs}
    best.launch()
    torch.cuda.synchronize()
    for i, j in enumerate(jobs):
        # the plain path with the options the tuner chose for this grid size
        r = _plain(j, rep["per_dims"][f"{j['dims']}/{j['precision']}"])
        P = best.results(i)[0].cpu().numpy()
        assert np.array_equal(best.results(i)[1].cpu().numpy(), r[1])
        assert np.allclose(P, r[0], rtol=1e-12, atol=0)
    best.close()
    again, rep2 = gb.tune(jobs, candidates=(1, 2), reps=2)
    assert rep2["cached"] is True and rep2["choice"] == rep["choice"]
    again.close()


def test_batch_errors(gpu):
    import torch
    announce("batch_errors")
    b = gb.Batch(2)
    with pytest.raises(genpk_amd.GenPKError):
        b.build()                                       # no jobs
    pos = torch.zeros((4, 3), dtype=torch.float32, device="cuda")
    with pytest.raises(genpk_amd.GenPKError):
        b.add(pos, 10.0, 14)                            # 14/2 = 7 is not 2^a 3^b 5^c
    with pytest.raises(genpk_amd.GenPKError):
        b.add(pos, -1.0, 16)
    b.close()


@pytest.mark.timeout(600)
def test_uneven_64_tuned_graph_vs_oracle(gpu, tmp_path, monkeypatch):
    """The batched workload as the benchmark runs it: 64 realisations
    (n_side log-uniform in [64, 256], dims uniform over {128, 256, 384, 512},
    0.3-cell Gaussian lattices in a 100 Mpc/h box, FP32) through the
    autotuned graph (deposit tile shape and FFT radix ordering per grid
    size, then branches x policy).  Every job against the plain path with the
    chosen options (counts and keffs equal, P within 1e-12); the smallest
    job of each grid size against the FP64 CPU oracle: counts bit-exact, the
    FP32 graph's P within 1e-3 (these jobs hold as few as 0.003 particles per
    cell, and FP32 rounding of the sparse field reaches 2e-4 in the weakest
    bins), and the same job through an FP64 plan with the chosen options
    within 1e-8."""
    import torch
    import oracle
    announce("uneven 64 realisations, tuned graph")
    monkeypatch.setenv("GENPK_AUTOTUNE_CACHE", str(tmp_path / "tune.json"))
    box = 100.0
    jobs = []
    for ns, d, sd in synth.uneven_realisations(64):
        pos = synth.gaussian_lattice_torch(ns, box, d, 0.3, seed=sd)
        jobs.append(dict(pos=pos, boxsize=box, dims=d, nrbins=d, masses=None, total_mass=float(pos.shape[0]),
                         precision="f32"))
    assert {j["dims"] for j in jobs} == {128, 256, 384, 512}
    best, rep = gb.tune(jobs, reps=2, log=lambda m: print(m, flush=True))
    print("per grid size:", {k: (v["tiles"], v["radices_d"]) for k, v in rep["per_dims"].items()},
          "graph:", rep["choice"], flush=True)
    for rep_i in range(2):
        for out in best.results():
            out[0].fill_(-1.0)
        best.launch()
        torch.cuda.synchronize()
        for i, j in enumerate(jobs):
            r = _plain(j, rep["per_dims"][f"{j['dims']}/f32"])
            P, c, k = (t.cpu().numpy() for t in best.results(i))
            assert np.array_equal(c, r[1]) and np.array_equal(k, r[2]), (rep_i, i)
            assert np.allclose(P, r[0], rtol=1e-12, atol=0), (rep_i, i)
    for d in (128, 256, 384, 512):
        i = min((i for i, j in enumerate(jobs) if j["dims"] == d), key=lambda i: jobs[i]["pos"].shape[0])
        j = jobs[i]
        n = j["pos"].shape[0]
        po, co, ko = oracle.pk_from_particles(box, d, j["pos"].cpu().numpy(), None, 1.0, d, float(n))
        occ = co > 0
        P, c, k = (t.cpu().numpy() for t in best.results(i))
        assert np.array_equal(c, co), d
        assert np.abs(k[occ] / ko[occ] - 1).max() <= 1e-12
        e32 = np.abs(P[occ] / po[occ] - 1).max()
        opts = rep["per_dims"][f"{d}/f32"]
        P64 = _plain(dict(j, precision="f64"), opts)[0]
        e64 = np.abs(P64[occ] / po[occ] - 1).max()
        print(f"d={d} n={n} tiles {opts['tiles']} radices {opts['radices_d']}: f32 {e32:.2e} f64 {e64:.2e}",
              flush=True)
        assert e32 <= 1e-3 and e64 <= 1e-8, d
    best.close()
