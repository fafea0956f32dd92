"""Parity of the HIP path (through the C ABI) with the CPU oracle.

Bars: mode counts bit-exact; FP64 plans within 1e-8 relative of the oracle
(the deposit accumulates weights in 2^-36 fixed point); FP32 plans within
the stated float tolerance.  Inputs are seeded; every case prints its name
and sizes before touching the GPU.
"""
import numpy as np
import pytest

import genpk_amd
import oracle
from genpk_amd import synth

pytestmark = pytest.mark.gpu


def announce(*a):
    print("CASE", *a, flush=True)


def near(x, y):
    return abs(x - y) <= max(abs(x), abs(y)) / 1e5


def to_dev(torch, a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


# ------------------------------------------------------- reference KATs via shims

def test_kat_fieldize_shim(gpu):
    announce("kat_fieldize d=5")
    dims = 5
    field = np.zeros(2 * dims * dims * (dims // 2 + 1))
    pos = np.array([i / 3.0 for i in range(30)], dtype=np.float32)
    masses = np.full(30, 10.0, dtype=np.float32)
    genpk_amd.fieldize(10, dims, field, pos, masses, 10.0, 1)
    assert near(field[0], 8.61111)
    assert field[3] == 0 and field[20] == 0 and field[125] == 0
    assert near(field[124], 1.66666)


def test_kat_powerspectrum_shim(gpu):
    announce("kat_powerspectrum d=4")
    field = np.zeros((4, 4, 6))
    f = field.reshape(-1)
    for i in range(32):
        ii = i // 4
        f[6 * ii + i % 4] = 1
    f[0] = 2
    F = genpk_amd.fft_r2c_3d_inplace(field)
    p, c, k = genpk_amd.powerspectrum(4, F, F, 10, 64.0, 64.0)
    assert near(k[2], np.sqrt(2))
    assert c[2] == 12 and c[1] == 0 and c[0] == 6
    assert near(p[0], 0.0677526)
    assert p[1] == 0
    assert near(p[2], 0.000565561)
    assert near(p[9], 0.0550908)


# ----------------------------------------------------------------- FFT

@pytest.mark.parametrize("d", [2, 4, 6, 8, 10, 12, 16, 30, 48, 64, 96, 128])
def test_fft_shim_vs_oracle(gpu, d):
    announce("fft_shim", d)
    rng = np.random.default_rng(d)
    x = rng.standard_normal((d, d, d))
    a = np.zeros((d, d, 2 * (d // 2 + 1)))
    a[:, :, :d] = x
    b = a.copy()
    Fo = oracle.fft_r2c_3d(a)
    Fg = genpk_amd.fft_r2c_3d_inplace(b)
    assert np.abs(Fg - Fo).max() <= 1e-13 * np.abs(Fo).max()


@pytest.mark.parametrize("prec,tol", [("f64", 1e-13), ("f32", 3e-6)])
@pytest.mark.parametrize("d", [16, 64, 128, 256, 384])
def test_r2c_c2r_plan(gpu, d, prec, tol):
    import torch
    announce("r2c_c2r", d, prec)
    rng = np.random.default_rng(7 + d)
    x = rng.standard_normal((d, d, d))
    grid = np.zeros((d, d, 2 * (d // 2 + 1)))
    grid[:, :, :d] = x
    p = genpk_amd.Plan(d, prec)
    p.import_density(to_dev(torch, grid.astype(np.float32 if prec == "f32" else np.float64)))
    F = p.r2c().cpu().numpy()
    R = np.fft.rfftn(x)
    assert np.abs(F - R).max() <= tol * np.abs(R).max()
    back = p.c2r().cpu().numpy()[:, :, :d] / d ** 3
    assert np.abs(back - x).max() <= 10 * tol * np.abs(x).max()


@pytest.mark.parametrize("prec,tol", [("f64", 1e-12), ("f32", 2e-5)])
def test_r2c_c2r_plan_1024(gpu, prec, tol):
    """d = 1024 (a z row is one element per thread in both z passes), checked
    on the GPU: a plane wave's single spike, Parseval on a random field, and
    c2r(r2c(x)) = d^3 x."""
    import torch
    announce("r2c_c2r 1024", prec)
    d = 1024
    dt = torch.float32 if prec == "f32" else torch.float64
    p = genpk_amd.Plan(d, prec)
    grid = torch.zeros((d, d, 2 * (d // 2 + 1)), dtype=dt, device="cuda")
    ax = torch.arange(d, device="cuda", dtype=torch.float64) * (2 * np.pi / d)
    k0 = (3, 5, 7)
    grid[:, :, :d] = torch.cos(k0[0] * ax[:, None, None] + k0[1] * ax[None, :, None] + k0[2] * ax[None, None, :]).to(dt)
    p.import_density(grid)
    F = p.r2c()
    half = d ** 3 / 2
    spike = complex(F[k0].cpu())
    assert abs(spike - half) <= tol * half
    F[k0] = 0
    assert float(F.abs().max().cpu()) <= tol * half
    del F
    g = torch.Generator(device="cuda")
    g.manual_seed(11)
    x = torch.randn((d, d, d), generator=g, device="cuda", dtype=dt)
    grid[:, :, :d] = x
    p.import_density(grid)
    F = p.r2c()
    h = torch.full((d // 2 + 1,), 2.0, dtype=torch.float64, device="cuda")
    h[0] = h[d // 2] = 1.0
    lhs = float(x.double().pow(2).sum().cpu())
    rhs = float(((F.abs().double() ** 2).sum(dim=(0, 1)) * h).sum().cpu()) / d ** 3
    assert abs(rhs - lhs) <= tol * lhs
    del F
    back = p.c2r()[:, :, :d]
    err = float((back.double() / d ** 3 - x.double()).abs().max().cpu())
    assert err <= 10 * tol * float(x.abs().max().cpu())


# ----------------------------------------------------------------- deposit

def deposit_case(torch, d, box, pos, masses, mass, prec):
    p = genpk_amd.Plan(d, prec)
    p.deposit(to_dev(torch, pos), box, None if masses is None else to_dev(torch, masses), mass)
    got = p.export_density().cpu().numpy().astype(np.float64)
    ref = oracle.new_field(d)
    oracle.fieldize(box, d, ref, pos, masses, mass, 1)
    return got[:, :, :d], ref[:, :, :d]


@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("d,n,with_masses", [(5, 10, True), (8, 1000, False), (16, 20000, True), (64, 300000, False),
                                             (100, 200000, True), (256, 2000000, False), (384, 1000000, True)])
def test_deposit_vs_oracle(gpu, d, n, with_masses, prec):
    import torch
    announce("deposit", d, n, with_masses, prec)
    rng = np.random.default_rng(d * 1000 + n)
    box = 50.0
    pos = (rng.random((n, 3)) * 1.2 * box - 0.1 * box).astype(np.float32)  # some outside [0, box)
    masses = (rng.random(n) + 0.5).astype(np.float32) if with_masses else None
    got, ref = deposit_case(torch, d, box, pos, masses, 1.7, prec)
    scale = np.abs(ref).max()
    if prec == "f64":
        assert np.abs(got - ref).max() <= 1e-9 * scale
    else:
        assert np.abs(got - ref).max() <= 2e-7 * scale
    total = masses.astype(np.float64).sum() if with_masses else 1.7 * n
    assert abs(got.sum() - total) <= 1e-6 * total


@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("kind", ["narrow", "wide", "nan", "zero"])
def test_deposit_mass_ranges_vs_oracle(gpu, kind, prec):
    """Per-particle masses whose magnitudes span at most 2^10 accumulate in
    fixed point, wider ranges and non-finite masses in FP64; both match the
    oracle (a NaN mass makes its eight cells NaN there too)."""
    import torch
    announce("deposit masses", kind, prec)
    d, box, n = 64, 40.0, 200000
    rng = np.random.default_rng({"narrow": 1, "wide": 2, "nan": 3, "zero": 4}[kind])
    pos = (rng.random((n, 3)) * box).astype(np.float32)
    if kind == "narrow":
        masses = (rng.random(n) * 900 + 1.5).astype(np.float32) * np.where(rng.random(n) < 0.1, -1, 1).astype(np.float32)
    elif kind == "wide":
        masses = np.exp(rng.uniform(np.log(1e-5), np.log(10.0), n)).astype(np.float32)
    elif kind == "nan":
        masses = (rng.random(n) + 0.5).astype(np.float32)
        masses[n // 3] = np.nan
    else:
        masses = np.zeros(n, dtype=np.float32)
        masses[::1000] = 2.0
    got, ref = deposit_case(torch, d, box, pos, masses, 1.0, prec)
    assert np.array_equal(np.isnan(got), np.isnan(ref))
    ok = ~np.isnan(ref)
    scale = np.abs(ref[ok]).max()
    tol = 1e-9 if prec == "f64" else 2e-7
    assert np.abs(got[ok] - ref[ok]).max() <= tol * scale
    if kind == "wide":
        # the light particles' cells keep their relative accuracy
        light = ok & (np.abs(ref) > 0) & (np.abs(ref) < 1e-3)
        assert light.any()
        rel = np.abs(got[light] - ref[light]) / np.abs(ref[light])
        assert rel.max() <= (1e-9 if prec == "f64" else 1e-6)


def test_deposit_masses_fixed_point_is_deterministic(gpu):
    """Masses within the fixed-point range give bit-identical grids on
    repeated deposits (integer sums do not depend on the atomics' order)."""
    import torch
    announce("deposit masses determinism")
    d, box, n = 64, 64.0, 400000
    rng = np.random.default_rng(9)
    pos = np.concatenate([rng.normal(20.3, 0.4, (300000, 3)), rng.random((n - 300000, 3)) * box]).astype(np.float32)
    masses = (rng.random(n) + 0.5).astype(np.float32)
    grids = []
    for _ in range(3):
        p = genpk_amd.Plan(d, "f64")
        p.deposit(to_dev(torch, pos), box, to_dev(torch, masses), 1.0)
        grids.append(p.export_density().cpu().numpy())
    assert np.array_equal(grids[0], grids[1]) and np.array_equal(grids[0], grids[2])


@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("with_masses", [False, True])
def test_deposit_split_tiles_independent_of_particle_order(gpu, with_masses, prec):
    """A tile of more than 2^17 particles is split into work items whose
    partial tiles are summed exactly (fixed point), so the grid is
    bit-identical whatever the particles' order, which decides the items'
    membership (here a clump of 3e5 particles in a few cells, whose sums pass
    2^53 in fixed point)."""
    import torch
    announce("deposit split tiles order", with_masses, prec)
    d, box, n = 64, 64.0, 400000
    rng = np.random.default_rng(19)
    pos = np.concatenate([rng.normal(20.3, 0.4, (300000, 3)), rng.random((n - 300000, 3)) * box]).astype(np.float32)
    masses = (rng.random(n) + 0.5).astype(np.float32) if with_masses else None
    grids = []
    for order in (np.arange(n), rng.permutation(n)):
        p = genpk_amd.Plan(d, prec)
        p.deposit(to_dev(torch, pos[order]), box, None if masses is None else to_dev(torch, masses[order]), 1.0)
        grids.append(p.export_density().cpu().numpy())
    assert np.array_equal(grids[0], grids[1])


@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("with_masses", [False, True])
def test_deposit_wide_key_vs_oracle(gpu, with_masses, prec):
    """d = 1280: 80 x 40 x 40 = 128000 tiles, a 17-bit key whose 9-bit low
    digit takes the second sort pass's histogram from the records (not from
    one-byte digits).  The particles lie in cells [0, 100) of each axis, so
    no weight wraps and the oracle deposits them on a 128^3 grid of the same
    cell size (box scaled by 128 / 1280: the same units d / box)."""
    import torch
    announce("deposit wide key d=1280", with_masses, prec)
    d, box, dsmall, n = 1280, 50.0, 128, 300000
    rng = np.random.default_rng(1280 + with_masses)
    cells = rng.random((n, 3)) * 100.0
    pos = (cells * (box / d)).astype(np.float32)
    masses = (rng.random(n) + 0.5).astype(np.float32) if with_masses else None
    box_small = box * dsmall / d
    assert (np.float64(d) / box) == (np.float64(dsmall) / box_small)
    p = genpk_amd.Plan(d, prec)
    p.deposit(to_dev(torch, pos), box, None if masses is None else to_dev(torch, masses), 1.3)
    dens = p.export_density()
    got = dens[:dsmall, :dsmall, :dsmall].cpu().numpy().astype(np.float64)
    full_sum = float(dens[:, :, :d].sum(dtype=torch.float64).cpu())
    del dens, p
    torch.cuda.empty_cache()
    ref = oracle.new_field(dsmall)
    oracle.fieldize(box_small, dsmall, ref, pos, masses, 1.3, 1)
    ref = ref[:, :, :dsmall]
    tol = 1e-9 if prec == "f64" else 2e-7
    assert np.abs(got - ref).max() <= tol * np.abs(ref).max()
    total = masses.astype(np.float64).sum() if with_masses else 1.3 * n
    assert abs(full_sum - total) <= 1e-6 * total
    assert abs(got.sum() - total) <= 1e-6 * total  # nothing outside the compared corner


def test_deposit_heavy_tile_split(gpu):
    import torch
    announce("deposit heavy tile d=64 n=400000 clumped")
    rng = np.random.default_rng(5)
    d, box, n = 64, 64.0, 400000
    # 300k particles inside one tile (forces several work items), rest uniform
    pos = np.concatenate([rng.normal(20.3, 0.4, (300000, 3)), rng.random((n - 300000, 3)) * box]).astype(np.float32)
    for prec in ("f64", "f32"):
        got, ref = deposit_case(torch, d, box, pos, None, 1.0, prec)
        tol = 1e-9 if prec == "f64" else 2e-7
        assert np.abs(got - ref).max() <= tol * np.abs(ref).max(), prec


@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("kind", ["signed", "net_negative"])
def test_deposit_heavy_tile_split_signed_masses(gpu, kind, prec):
    """A tile of 3e5 particles (split into work items whose partial tiles
    are summed as signed 128-bit integers) with per-particle masses: about
    10% negative ("signed"), or mostly negative so that the heavy tile's
    cell sums are net negative ("net_negative"); against the oracle."""
    import torch
    announce("deposit heavy tile signed masses", kind, prec)
    rng = np.random.default_rng(6 if kind == "signed" else 7)
    d, box, n = 64, 64.0, 400000
    pos = np.concatenate([rng.normal(20.3, 0.4, (300000, 3)), rng.random((n - 300000, 3)) * box]).astype(np.float32)
    mag = (rng.random(n) * 900 + 1.5).astype(np.float32)
    neg = rng.random(n) < (0.1 if kind == "signed" else 0.8)
    masses = np.where(neg, -mag, mag).astype(np.float32)
    got, ref = deposit_case(torch, d, box, pos, masses, 1.0, prec)
    if kind == "net_negative":
        assert ref.min() < 0 and ref.sum() < 0
    tol = 1e-9 if prec == "f64" else 2e-7
    assert np.abs(got - ref).max() <= tol * np.abs(ref).max()


@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("masses_kind", [None, "fixed", "wide"])
def test_deposit_nonfinite_positions(gpu, masses_kind, prec):
    """A NaN, a +inf and a -inf coordinate (three particles, three axes)
    make each particle's eight cells non-finite, as the reference's weights
    are (fieldize.cpp:67-84), at the cells the oracle puts them (the floor
    of a non-finite value converts to INT_MIN on x86-64); every other cell
    matches the oracle, and P(k) is NaN in every occupied bin.  With masses
    in the fixed-point range and beyond it (FP64 accumulation)."""
    import torch
    announce("deposit non-finite positions", masses_kind, prec)
    d, box, n = 24, 10.0, 5000
    rng = np.random.default_rng(23)
    pos = (rng.random((n, 3)) * box).astype(np.float32)
    pos[10, 0], pos[20, 1], pos[30, 2] = np.nan, np.inf, -np.inf
    masses = None
    if masses_kind == "fixed":
        masses = (rng.random(n) + 0.5).astype(np.float32)
    elif masses_kind == "wide":
        masses = np.exp(rng.uniform(np.log(1e-5), np.log(10.0), n)).astype(np.float32)
    got, ref = deposit_case(torch, d, box, pos, masses, 1.0, prec)
    bad = ~np.isfinite(ref)
    assert bad.sum() >= 8
    assert np.array_equal(~np.isfinite(got), bad)
    tol = 1e-9 if prec == "f64" else 2e-7
    assert np.abs(got[~bad] - ref[~bad]).max() <= tol * np.abs(ref[~bad]).max()
    p = genpk_amd.Plan(d, prec)
    p.deposit(to_dev(torch, pos), box, None if masses is None else to_dev(torch, masses), 1.0)
    P, c, _ = p.power(d, mass=float(n))
    P, c = P.cpu().numpy(), c.cpu().numpy()
    assert np.isnan(P[c > 0]).all()


def test_deposit_accumulates_and_exports_into(gpu):
    import torch
    announce("deposit accumulate d=32")
    rng = np.random.default_rng(9)
    d, box = 32, 10.0
    a = (rng.random((5000, 3)) * box).astype(np.float32)
    b = (rng.random((7000, 3)) * box).astype(np.float32)
    p = genpk_amd.Plan(d, "f64")
    p.deposit(to_dev(torch, a), box, mass=2.0)
    p.deposit(to_dev(torch, b), box, mass=0.5)
    base = torch.ones((d, d, d + 2), dtype=torch.float64, device="cuda")
    got = p.export_density(out=base, accumulate=True).cpu().numpy()
    ref = np.ones((d, d, d + 2))
    oracle.fieldize(box, d, ref, a, None, 2.0, 1)
    oracle.fieldize(box, d, ref, b, None, 0.5, 1)
    assert np.abs(got - ref).max() <= 1e-9 * np.abs(ref).max()


def test_deposit_empty_and_zero_particles(gpu):
    import torch
    announce("deposit empty d=16")
    p = genpk_amd.Plan(16, "f64")
    p.deposit(torch.zeros((0, 3), dtype=torch.float32, device="cuda"), 1.0)
    assert float(p.export_density().abs().max()) == 0.0
    pw, c, k = p.power()
    assert float(pw.abs().max()) == 0.0
    assert int(c.sum()) == 16 ** 3 - 1


def test_fieldize_shim_extra0_and_accumulate(gpu):
    announce("fieldize shim extra=0 d=8")
    rng = np.random.default_rng(1)
    d, box = 8, 3.0
    pos = (rng.random((500, 3)) * box).astype(np.float32)
    out = np.full(d * d * d, 0.25)
    genpk_amd.fieldize(box, d, out, pos, None, 1.0, 0)
    ref = np.full(d * d * d, 0.25)
    oracle.fieldize(box, d, ref, pos, None, 1.0, 0)
    assert np.abs(out - ref).max() <= 1e-9 * np.abs(ref).max()


# ----------------------------------------------------------------- power spectrum

@pytest.mark.parametrize("fused", [True, False])
@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("d,n", [(16, 5000), (32, 40000), (64, 262144), (128, 262144), (96, 100000)])
def test_power_vs_oracle(gpu, d, n, prec, fused):
    import torch
    announce("power", d, n, prec, "fused" if fused else "stored")
    rng = np.random.default_rng(d + n)
    box = 25.0
    pos = (rng.random((n, 3)) * box).astype(np.float32)
    po, co, ko = oracle.pk_from_particles(box, d, pos, None, 1.0, d, float(n))
    p = genpk_amd.Plan(d, prec)
    p.deposit(to_dev(torch, pos), box, mass=1.0)
    if fused:
        pw, c, k = p.power(d, mass=float(n))
    else:
        p.r2c()
        pw, c, k = p.bin(d, mass=float(n))
    pw, c, k = pw.cpu().numpy(), c.cpu().numpy(), k.cpu().numpy()
    assert np.array_equal(c, co)
    occ = co > 0
    assert np.abs(k[occ] / ko[occ] - 1).max() <= 1e-12
    rel = np.abs(pw[occ] / po[occ] - 1).max()
    print("max rel P error", prec, rel)
    assert rel <= (1e-8 if prec == "f64" else 1e-4)


def test_power_c0_lattice(gpu):
    """Config C0: 64^3 sigma=0.3-cell Gaussian lattice, d=128, 1e-5 bar."""
    import torch
    announce("power C0 lattice 64^3 d=128")
    box, d = 25000.0, 128
    pos = synth.gaussian_lattice(64, box, d, 0.3, seed=0)
    n = pos.shape[0]
    po, co, ko = oracle.pk_from_particles(box, d, pos, None, 1.0, d, float(n))
    occ = co > 0
    for prec, bar in (("f64", 1e-8), ("f32", 2e-4)):
        p = genpk_amd.Plan(d, prec)
        p.deposit(to_dev(torch, pos), box, mass=1.0)
        pw, c, k = p.power(d, mass=float(n))
        pw = pw.cpu().numpy()
        assert np.array_equal(c.cpu().numpy(), co)
        rel = np.abs(pw[occ] / po[occ] - 1)
        print("C0", prec, "max rel", rel.max(), "median", np.median(rel))
        assert rel.max() <= bar


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_cross_vs_oracle(gpu, prec):
    import torch
    announce("cross d=64", prec)
    rng = np.random.default_rng(21)
    d, box, n = 64, 100.0, 100000
    a = (rng.random((n, 3)) * box).astype(np.float32)
    b = np.mod(a + rng.normal(0, 1.0, (n, 3)), box).astype(np.float32)
    fa = oracle.new_field(d)
    oracle.fieldize(box, d, fa, a, None, 1.0, 1)
    fb = oracle.new_field(d)
    oracle.fieldize(box, d, fb, b, None, 0.2, 1)
    Fa, Fb = oracle.fft_r2c_3d(fa), oracle.fft_r2c_3d(fb)
    ma, mb = float(n), 0.2 * n
    paa, c, k = oracle.powerspectrum(d, Fa, Fa, d, ma, ma)
    pbb, _, _ = oracle.powerspectrum(d, Fb, Fb, d, mb, mb)
    pab, _, _ = oracle.powerspectrum(d, Fa, Fb, d, ma, mb)
    A = genpk_amd.Plan(d, prec)
    B = genpk_amd.Plan(d, prec)
    A.deposit(to_dev(torch, a), box, mass=1.0)
    B.deposit(to_dev(torch, b), box, mass=0.2)
    (gaa, gbb, gab), gc, gk = A.power(d, other=B, mass=ma, other_mass=mb)
    assert np.array_equal(gc.cpu().numpy(), c)
    occ = c > 0
    tol = 1e-8 if prec == "f64" else 1e-4
    for g, o in ((gaa, paa), (gbb, pbb), (gab, pab)):
        g = g.cpu().numpy()
        assert np.abs(g[occ] - o[occ]).max() <= tol * np.abs(o[occ]).max()


def test_powerspectrum_shim_cross_and_distinct_masses(gpu):
    announce("powerspectrum shim cross d=8")
    rng = np.random.default_rng(2)
    d = 8
    f1 = np.zeros((d, d, d + 2)); f1[:, :, :d] = rng.random((d, d, d))
    f2 = np.zeros((d, d, d + 2)); f2[:, :, :d] = rng.random((d, d, d))
    F1, F2 = oracle.fft_r2c_3d(f1), oracle.fft_r2c_3d(f2)
    for (x, y, m1, m2) in ((F1, F1, 3.0, 5.0), (F1, F2, 3.0, 5.0)):
        po, co, ko = oracle.powerspectrum(d, x, y, d, m1, m2)
        pg, cg, kg = genpk_amd.powerspectrum(d, x, y, d, m1, m2)
        assert np.array_equal(co, cg)
        assert np.allclose(pg, po, rtol=1e-12, atol=1e-18)
        assert np.allclose(kg, ko, rtol=1e-13)


# ----------------------------------------------------------------- bins

@pytest.mark.parametrize("d", [4, 8, 128, 256, 384])
def test_counts_bitexact(gpu, d):
    announce("counts", d)
    F = np.zeros((d, d, d // 2 + 1), np.complex128)
    po, co, ko = oracle.powerspectrum(d, F, F, d, 1.0, 1.0)
    pg, cg, kg = genpk_amd.powerspectrum(d, F, F, d, 1.0, 1.0)
    assert np.array_equal(co, cg)
    occ = co > 0
    assert np.abs(kg[occ] / ko[occ] - 1).max() <= 1e-12


def test_counts_1024_reference_facts(gpu):
    """Mode-count facts of the compiled reference at d = 1024."""
    import torch
    announce("counts d=1024")
    p = genpk_amd.Plan(1024, "f32")
    p.deposit(torch.zeros((0, 3), dtype=torch.float32, device="cuda"), 1.0)
    _, c, k = p.power(1024)
    c = c.cpu().numpy().astype(np.int64)
    assert int((c > 0).sum()) == 757
    assert c[1023] == 0 and c[1022] >= 1
    assert int(c.max()) == 11191384
    assert int(c.sum()) == 1024 ** 3 - 1


def test_r2c_from_tiles_1024_matches_density_path(gpu):
    """d = 1024, FP32: the z pass that reads the tile store (the eight inputs
    of a first-stage butterfly per thread straight from the store, z halos
    added on load) gives the half spectrum of the element-wise z pass on the
    exported density, the same arithmetic in another order of work: within
    1e-6 of the largest |F| (a wrong index or halo would differ at O(1))."""
    import torch
    announce("r2c from tiles 1024")
    d, box, n = 1024, 1000.0, 1 << 20
    pos = synth.clumped_torch(n, box, d, n_clumps=64, seed=41, device="cuda")
    p = genpk_amd.Plan(d, "f32")
    p.deposit(pos, box)
    dens = p.export_density()
    F1 = p.r2c().clone()
    del p
    q = genpk_amd.Plan(d, "f32")
    q.import_density(dens)
    F2 = q.r2c()
    scale = float(F2.abs().max())
    diff = float((F1 - F2).abs().max())
    print(f"max |dF| {diff:.3e} of {scale:.3e}", flush=True)
    assert diff <= 1e-6 * scale
    del q, F1, F2, dens
    torch.cuda.empty_cache()


@pytest.mark.parametrize("with_masses", [False, True])
def test_power_1024_fp32_vs_fp64(gpu, with_masses):
    """d = 1024, where the FP32 forward reads a z row one element per thread
    or a first-stage butterfly per thread straight from the tile store, and
    bins in the fused x pass: P against the FP64 plan's (element-wise z pass)
    on the same clustered particles within 2e-5 per occupied bin, and equal
    counts."""
    import torch
    announce("power d=1024 f32 vs f64", with_masses)
    d, box, n = 1024, 1000.0, 1 << 22
    pos = synth.clumped_torch(n, box, d, n_clumps=400, seed=31, device="cuda")
    masses = (torch.rand(n, generator=torch.Generator(device="cuda").manual_seed(31), device="cuda") + 0.5) \
        if with_masses else None
    M = float(masses.double().sum()) if with_masses else float(n)
    out = {}
    for prec in ("f32", "f64"):
        p = genpk_amd.Plan(d, prec)
        p.deposit(pos, box, masses)
        P, c, _ = p.power(d, mass=M)
        out[prec] = (P.cpu().numpy(), c.cpu().numpy())
        del p
        torch.cuda.empty_cache()
    (p32, c32), (p64, c64) = out["f32"], out["f64"]
    assert np.array_equal(c32, c64)
    occ = c64 > 0
    rel = np.abs(p32[occ] / p64[occ] - 1)
    print(f"max rel {rel.max():.3e}", flush=True)
    assert rel.max() <= 2e-5


# ----------------------------------------------------------------- errors

def test_errors_are_reported(gpu):
    import torch
    announce("errors")
    with pytest.raises(genpk_amd.GenPKError):
        genpk_amd.Plan(5, "f64").r2c()  # odd dims: no r2c
    with pytest.raises(genpk_amd.GenPKError):
        genpk_amd.Plan(14, "f64").r2c()  # 7 is not a supported radix
    p = genpk_amd.Plan(16, "f64")
    p.deposit(torch.rand((100, 3), device="cuda"), 1.0)
    with pytest.raises(genpk_amd.GenPKError):
        p.power(0)  # no bins
    with pytest.raises(genpk_amd.GenPKError):
        genpk_amd.Plan(0, "f64")
    with pytest.raises(genpk_amd.GenPKError):
        p.deposit(torch.rand((10, 3), device="cuda"), -1.0)
