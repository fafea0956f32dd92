"""Plan tuning options through the C ABI: deposit tile shapes and FFT radix
orderings give the same results as the oracle (and as numpy's FP64 FFT),
whichever the autotuner picks.

Bars: as the parity tests -- counts bit-exact, FP64 within 1e-8 of the
oracle (deposit 1e-9 of the largest cell), FP32 within the float
tolerance; transforms against numpy's rfftn within 1e-13 (FP64) and 3e-6
(FP32) of the largest coefficient.
"""
import numpy as np
import pytest

import genpk_amd
import oracle

pytestmark = pytest.mark.gpu

TILES = [(16, 32, 32), (8, 32, 32), (16, 16, 32), (8, 16, 32), (4, 8, 16), (4, 4, 4)]
_ORACLE = {}


def announce(*a):
    print("CASE", *a, flush=True)


def to_dev(torch, a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


@pytest.mark.parametrize("variant", [0, 1, 2, 3])
@pytest.mark.parametrize("prec,tol", [("f64", 1e-13), ("f32", 3e-6)])
@pytest.mark.parametrize("d", [48, 128, 384, 512])
def test_fft_variants_vs_numpy(gpu, d, prec, tol, variant):
    import torch
    announce("fft variant", d, prec, variant)
    rng = np.random.default_rng(11 + d)
    x = rng.standard_normal((d, d, d))
    grid = np.zeros((d, d, 2 * (d // 2 + 1)))
    grid[:, :, :d] = x
    p = genpk_amd.Plan(d, prec)
    p.set_options((16, 32, 32), variant)
    assert p.options()["fft_variant"] == variant
    p.import_density(to_dev(torch, grid.astype(np.float32 if prec == "f32" else np.float64)))
    F = p.r2c().cpu().numpy()
    R = np.fft.rfftn(x)
    assert np.abs(F - R).max() <= tol * np.abs(R).max()
    back = p.c2r().cpu().numpy()[:, :, :d] / d ** 3
    assert np.abs(back - x).max() <= 10 * tol * np.abs(x).max()


@pytest.mark.parametrize("tiles", TILES)
@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("d,with_masses", [(64, False), (100, True), (384, False)])
def test_deposit_tiles_vs_oracle(gpu, d, with_masses, prec, tiles):
    import torch
    announce("deposit tiles", tiles, d, with_masses, prec)
    rng = np.random.default_rng(d + sum(tiles))
    box, n = 50.0, 200000
    pos = (rng.random((n, 3)) * 1.2 * box - 0.1 * box).astype(np.float32)
    masses = (rng.random(n) + 0.5).astype(np.float32) if with_masses else None
    p = genpk_amd.Plan(d, prec)
    p.set_options(tiles, 0)
    eff = p.options()["tiles"]
    assert all(e <= t and e <= d for e, t in zip(eff, tiles))
    p.deposit(to_dev(torch, pos), box, None if masses is None else to_dev(torch, masses), 1.3)
    got = p.export_density().cpu().numpy().astype(np.float64)[:, :, :d]
    ref = oracle.new_field(d)
    oracle.fieldize(box, d, ref, pos, masses, 1.3, 1)
    ref = ref[:, :, :d]
    tol = 1e-9 if prec == "f64" else 2e-7
    assert np.abs(got - ref).max() <= tol * np.abs(ref).max()


@pytest.mark.parametrize("variant", [0, 1, 2, 3])
@pytest.mark.parametrize("tiles", [(16, 32, 32), (8, 16, 32), (4, 8, 16)])
@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("d", [128, 384])
def test_power_options_vs_oracle(gpu, d, prec, tiles, variant):
    import torch
    announce("power options", d, prec, tiles, variant)
    box, n = 25.0, 400000
    if d not in _ORACLE:
        pos = (np.random.default_rng(7 * d).random((n, 3)) * box).astype(np.float32)
        _ORACLE[d] = (pos, oracle.pk_from_particles(box, d, pos, None, 1.0, d, float(n)))
    pos, (po, co, ko) = _ORACLE[d]
    p = genpk_amd.Plan(d, prec)
    p.set_options(tiles, variant)
    p.deposit(to_dev(torch, pos), box, mass=1.0)
    pw, c, k = (t.cpu().numpy() for t in p.power(d, mass=float(n)))
    assert np.array_equal(c, co)
    occ = co > 0
    assert np.abs(k[occ] / ko[occ] - 1).max() <= 1e-12
    rel = np.abs(pw[occ] / po[occ] - 1).max()
    assert rel <= (1e-8 if prec == "f64" else 1e-4), rel


def test_options_1024_fp32(gpu):
    """d = 1024, FP32, where the default plan's z pass reads first-stage
    butterflies straight from the tile store and the x pass bins in
    registers: other tile shapes and radix orderings (which take the general
    forms of those passes) give the default plan's P within 2e-5 and equal
    counts."""
    import torch
    from genpk_amd import synth
    announce("options d=1024 f32")
    d, box, n = 1024, 1000.0, 1 << 21
    pos = synth.clumped_torch(n, box, d, n_clumps=200, seed=61, device="cuda")
    res = {}
    for tiles, variant in [((16, 32, 32), 0), ((8, 16, 32), 0), ((16, 32, 32), 2), ((8, 32, 32), 1),
                           ((16, 32, 32), 3)]:
        p = genpk_amd.Plan(d, "f32")
        p.set_options(tiles, variant)
        p.deposit(pos, box)
        P, c, _ = p.power(d, mass=float(n))
        res[(tiles, variant)] = (P.cpu().numpy(), c.cpu().numpy())
        del p
        torch.cuda.empty_cache()
    P0, c0 = res[((16, 32, 32), 0)]
    occ = c0 > 0
    for key, (P, c) in res.items():
        assert np.array_equal(c, c0), key
        assert np.abs(P[occ] / P0[occ] - 1).max() <= 2e-5, key


def test_options_1024_fp64(gpu):
    """d = 1024, FP64: the default plan (radix ordering 3, one radix 4 first
    in the column passes, the z pass with compile-time stages) against
    ordering 0 (compile-time column stages, the fused x pass without the
    register first stage) and another tile shape: equal counts, P within
    1e-10."""
    import torch
    from genpk_amd import synth
    announce("options d=1024 f64")
    d, box, n = 1024, 1000.0, 1 << 21
    pos = synth.clumped_torch(n, box, d, n_clumps=200, seed=62, device="cuda")
    res = {}
    for tiles, variant in [((16, 32, 32), None), ((16, 32, 32), 0), ((8, 32, 32), 3), ((16, 16, 32), 1)]:
        p = genpk_amd.Plan(d, "f64")
        if variant is not None:
            p.set_options(tiles, variant)
        assert p.options()["fft_variant"] == (3 if variant is None else variant)
        p.deposit(pos, box)
        P, c, _ = p.power(d, mass=float(n))
        res[(tiles, variant)] = (P.cpu().numpy(), c.cpu().numpy())
        del p
        torch.cuda.empty_cache()
    P0, c0 = res[((16, 32, 32), None)]
    occ = c0 > 0
    for key, (P, c) in res.items():
        assert np.array_equal(c, c0), key
        assert np.abs(P[occ] / P0[occ] - 1).max() <= 1e-10, key


@pytest.mark.parametrize("tiles,variant", [((8, 16, 32), 2), ((4, 8, 16), 1)])
def test_adjoint_with_options(gpu, tiles, variant):
    """The adjoint (x pass with the weight, inverse passes, tile gather) with
    another tile shape and radix ordering gives the default plan's gradients
    (FP64, to 1e-10 of the largest component)."""
    import torch
    from genpk_amd.autograd import power_spectrum
    announce("adjoint options", tiles, variant)
    d, box = 64, 100.0
    rng = np.random.default_rng(3)
    pos = to_dev(torch, (rng.random((30000, 3)) * box).astype(np.float32))
    m = to_dev(torch, (rng.random(30000) + 0.5).astype(np.float32))
    out = []
    for opts in (None, (tiles, variant)):
        plan = genpk_amd.Plan(d, "f64")
        if opts:
            plan.set_options(*opts)
        a = pos.clone().requires_grad_(True)
        b = m.clone().requires_grad_(True)
        P = power_spectrum(a, box, d, d, masses=b, plan=plan, recompute="always")
        L = (P * torch.linspace(0.5, 1.5, d, dtype=torch.float64, device="cuda")).sum()
        out.append([g.double() for g in torch.autograd.grad(L, (a, b))])
    for g0, g1 in zip(*out):
        assert (g0 - g1).abs().max().item() <= 1e-10 * g0.abs().max().item()


def test_options_errors(gpu):
    announce("options errors")
    p = genpk_amd.Plan(64, "f64")
    for bad in [((16, 32, 3), 0), ((16, 32, 1), 0), ((32, 32, 32), 0), ((16, 32, 32), 4), ((16, 32, 32), -1)]:
        with pytest.raises(genpk_amd.GenPKError):
            p.set_options(*bad)
    p.set_options((64, 4, 2), 0)          # 65 x 5 x 3 cells: fits
    assert p.options()["tiles"] == (64, 4, 2)
    q = genpk_amd.Plan(16, "f64")
    q.set_options((16, 32, 32), 0)         # clamped to the grid
    assert q.options()["tiles"] == (16, 16, 16)
