"""Gradients of the differentiable power spectrum (hand-written adjoint on
the GPU, through the C ABI) against an FP64 dense PyTorch reference on the
CPU (autograd through index_add CIC + rfftn + binning, itself pinned to the
oracle's forward) and against central finite differences of the GPU forward
with perturbations that are exactly representable in float32.

Bars (written per test): FP64 plans within 1e-6 of the largest gradient
component (the gradients are returned as float32); FP32 plans within 2e-3.
"""
import numpy as np
import pytest

import genpk_amd
from genpk_amd import synth
from genpk_amd.autograd import power_spectrum
from dense_ref import c3_loss, dense_power, mode_tables

pytestmark = pytest.mark.gpu


def announce(*a):
    print("CASE", *a, flush=True)


def _instance(n_side, d, box, seed):
    """Clustered particles (half uniform, half in clumps) and masses in [0.5, 1.5)."""
    n = n_side ** 3
    pos = synth.clumped(n, box, d, n_clumps=max(8, n // 512), seed=seed)
    rng = np.random.default_rng(seed + 100)
    m = (rng.random(n) + 0.5).astype(np.float32)
    return pos.astype(np.float32), m


def _target(P):
    b = np.arange(P.shape[0])
    return P * (1.0 + 0.3 * np.sin(0.37 * b))


def _dense_grads(pos, m, box, d, nrbins, target, tables, with_masses=True):
    import torch
    tp = torch.from_numpy(pos.astype(np.float64)).requires_grad_(True)
    tm = torch.from_numpy(m.astype(np.float64)).requires_grad_(with_masses)
    M = tm.sum()
    P, count = dense_power(tp, tm, box, d, nrbins, M, tables)
    L = c3_loss(P, torch.from_numpy(target), count)
    if with_masses:
        gp, gm = torch.autograd.grad(L, (tp, tm))
        return P.detach().numpy(), L.item(), gp.numpy(), gm.numpy()
    (gp,) = torch.autograd.grad(L, (tp,))
    return P.detach().numpy(), L.item(), gp.numpy(), None


def _gpu_loss_and_grads(torch, pos, m, box, d, nrbins, target, prec, plan, recompute="auto", with_masses=True):
    tp = torch.from_numpy(pos).cuda().requires_grad_(True)
    tm = torch.from_numpy(m).cuda().requires_grad_(with_masses) if with_masses else None
    if with_masses:
        P = power_spectrum(tp, box, d, nrbins, masses=tm, total_mass=tm.double().sum(), precision=prec, plan=plan,
                           recompute=recompute)
    else:
        P = power_spectrum(tp, box, d, nrbins, mass=float(m[0]), precision=prec, plan=plan, recompute=recompute)
    cnt = torch.from_numpy(mode_tables(d, nrbins)[2].numpy()).cuda()
    L = c3_loss(P, torch.from_numpy(target).cuda(), cnt)
    grads = torch.autograd.grad(L, (tp, tm) if with_masses else (tp,))
    gp = grads[0].cpu().numpy().astype(np.float64)
    gm = grads[1].cpu().numpy().astype(np.float64) if with_masses else None
    return P.detach().cpu().numpy(), L.item(), gp, gm


def _close(g, ref, frac):
    scale = np.abs(ref).max()
    err = np.abs(g - ref).max()
    return err <= frac * scale, err / scale


@pytest.mark.parametrize("prec,bar,policy", [("f64", 1e-6, "auto"), ("f64", 1e-6, "always"), ("f64", 1e-6, "never"),
                                             ("f32", 2e-3, "auto")])
def test_adjoint_vs_dense_fp64(gpu, prec, bar, policy):
    import torch
    d, box, n_side, nrbins = 32, 100.0, 32, 32
    announce("adjoint_vs_dense", prec, policy, f"d={d} n={n_side}^3")
    pos, m = _instance(n_side, d, box, seed=11)
    tables = mode_tables(d, nrbins)
    Pd, _, _, _ = _dense_grads(pos, m, box, d, nrbins, np.ones(nrbins), tables)
    target = _target(Pd)
    Pd, Ld, gpd, gmd = _dense_grads(pos, m, box, d, nrbins, target, tables)
    plan = genpk_amd.Plan(d, prec)
    Pg, Lg, gpg, gmg = _gpu_loss_and_grads(torch, pos, m, box, d, nrbins, target, prec, plan, recompute=policy)
    occ = tables[2].numpy() > 0
    assert np.allclose(Pg[occ], Pd[occ], rtol=1e-8 if prec == "f64" else 1e-4)
    okp, ep = _close(gpg, gpd, bar)
    okm, em = _close(gmg, gmd, bar)
    print(f"max rel grad err pos {ep:.3e} mass {em:.3e}", flush=True)
    assert okp and okm


def test_adjoint_scalar_mass(gpu):
    """Particles of one mass (masses=None): gradient w.r.t. positions; M = n*mass
    is a constant.  d = 48 leaves partial deposit tiles along y and z."""
    import torch
    d, box, n_side, nrbins = 48, 50.0, 24, 48
    announce("adjoint_scalar_mass", f"d={d} n={n_side}^3")
    pos, _ = _instance(n_side, d, box, seed=12)
    m = np.full(pos.shape[0], 2.5, dtype=np.float32)
    tables = mode_tables(d, nrbins)
    Pd, _, _, _ = _dense_grads(pos, m, box, d, nrbins, np.ones(nrbins), tables, with_masses=False)
    target = _target(Pd)
    _, _, gpd, _ = _dense_grads(pos, m, box, d, nrbins, target, tables, with_masses=False)
    plan = genpk_amd.Plan(d, "f64")
    _, _, gpg, _ = _gpu_loss_and_grads(torch, pos, m, box, d, nrbins, target, "f64", plan, with_masses=False)
    ok, e = _close(gpg, gpd, 1e-6)
    print(f"max rel grad err {e:.3e}", flush=True)
    assert ok


def test_adjoint_policies_agree(gpu):
    """The three recompute policies give the same gradients to rounding
    (1e-12 for P, 1e-10 of the largest gradient component): the bins' FP64
    LDS atomics run in a varying order, and the "never" forward bins the
    stored spectrum while the others use the fused x-pass binning, whose
    sums run in another order."""
    import torch
    d, box, n_side, nrbins = 48, 100.0, 32, 48
    announce("adjoint_policies", f"d={d} n={n_side}^3")
    pos, m = _instance(n_side, d, box, seed=13)
    target = np.ones(nrbins)
    plan = genpk_amd.Plan(d, "f64")
    r1 = _gpu_loss_and_grads(torch, pos, m, box, d, nrbins, target, "f64", plan, recompute="never")
    r2 = _gpu_loss_and_grads(torch, pos, m, box, d, nrbins, target, "f64", plan, recompute="always")
    r3 = _gpu_loss_and_grads(torch, pos, m, box, d, nrbins, target, "f64", plan, recompute="auto")
    assert np.allclose(r1[0], r2[0], rtol=1e-12, atol=0)
    assert _close(r1[2], r2[2], 1e-10)[0] and _close(r1[3], r2[3], 1e-10)[0]
    assert np.allclose(r2[0], r3[0], rtol=1e-12, atol=0)
    assert _close(r2[2], r3[2], 1e-10)[0] and _close(r2[3], r3[3], 1e-10)[0]


@pytest.mark.parametrize("d,n", [(128, 1 << 18), (1024, 1 << 21)])
def test_adjoint_x_fused_matches_unfused(gpu, d, n):
    """d = 1024 and 128 (8- and 64-column blocks), FP32, where the adjoint x
    pass runs its first stage from the loaded registers, alternates two LDS
    buffers and reads the weight's per-bin factors, window factors and bin
    thresholds from LDS: gradients with recompute "always" (fused forward FFT
    -> weight -> inverse FFT per column block) equal those of "never" (the
    weight applied to the kept spectrum, then the unfused inverse passes) to
    FP32 rounding (2e-4 of the largest component)."""
    import torch
    from genpk_amd.autograd import chi2_loss
    box = 1000.0 * d / 1024
    announce("adjoint_x_fused", f"d={d} n={n}")
    pos = synth.clumped(n, box, d, n_clumps=256, seed=21).astype(np.float32)
    m = (np.random.default_rng(21).random(n) + 0.5).astype(np.float32)
    tp = torch.from_numpy(pos).cuda()
    tm = torch.from_numpy(m).cuda()
    M = float(m.astype(np.float64).sum())
    plan = genpk_amd.Plan(d, "f32")
    plan.clear()
    plan.deposit(tp, box, tm)
    P0, cnt, _ = plan.power(d, mass=M)
    target = P0 * 1.05

    def grads(policy):
        a = tp.clone().requires_grad_(True)
        b = tm.clone().requires_grad_(True)
        P = power_spectrum(a, box, d, d, masses=b, total_mass=M, plan=plan, recompute=policy)
        return [g.double().cpu().numpy() for g in torch.autograd.grad(chi2_loss(P, target, cnt), (a, b))]

    g_never = grads("never")
    g_always = grads("always")
    for ga, gn in zip(g_always, g_never):
        ok, e = _close(ga, gn, 2e-4)
        print(f"max rel grad diff {e:.3e}", flush=True)
        assert ok


def test_adjoint_plan_flows(gpu):
    """C-level flows of the adjoint: (a) deposit -> power -> adjoint from the
    state the fused power leaves, gather with the order kept by the
    retaining deposit; (b) the same density imported as a grid, adjoint
    redoing the transform, gather partitioning the particles itself.  Both
    give the same dL/drho and gradients bit for bit."""
    import torch
    d, box, n_side, nrbins = 32, 100.0, 20, 32
    announce("adjoint_plan_flows", f"d={d} n={n_side}^3")
    pos, m = _instance(n_side, d, box, seed=16)
    tp = torch.from_numpy(pos).cuda()
    tm = torch.from_numpy(m).cuda()
    M = float(m.astype(np.float64).sum())
    lam = torch.linspace(-1.0, 2.0, nrbins, dtype=torch.float64, device="cuda")
    a = genpk_amd.Plan(d, "f64")
    a.clear()
    a.deposit(tp, box, tm, retain=True)
    grid = a.export_density()
    a.power(nrbins, mass=M)
    G1 = a.adjoint_density(lam, M, nrbins).clone()
    g1 = a.adjoint_gather(tp, box, tm)
    b = genpk_amd.Plan(d, "f64")
    b.import_density(grid)
    G2 = b.adjoint_density(lam, M, nrbins).clone()
    g2 = b.adjoint_gather(tp, box, tm)
    torch.cuda.synchronize()
    assert torch.equal(G1[:, :, :d], G2[:, :, :d])
    assert torch.equal(g1[0], g2[0]) and torch.equal(g1[1], g2[1])
    # the adjoint of the kept spectrum (weight kernel + full inverse) agrees to rounding
    c = genpk_amd.Plan(d, "f64")
    c.import_density(grid)
    c.r2c()
    G3 = c.adjoint_density(lam, M, nrbins)
    err = (G3[:, :, :d] - G1[:, :, :d]).abs().max().item()
    assert err <= 1e-12 * G1[:, :, :d].abs().max().item()


def test_adjoint_finite_differences(gpu):
    """Central differences of the GPU forward loss (FP64 plan) on 64
    position coordinates at least 0.01 cell from a cell face (CIC kinks) and
    on 64 masses, with h = 1e-4 cell (positions) or 1e-4 m (masses), the
    step taken as the float32 difference actually applied to the stored
    value.  The total mass is the sum of the (perturbed) masses."""
    import torch
    d, box, n_side, nrbins = 32, 32.0, 16, 32
    announce("adjoint_fd", f"d={d} n={n_side}^3")
    pos, m = _instance(n_side, d, box, seed=14)
    cnt = torch.from_numpy(mode_tables(d, nrbins)[2].numpy()).cuda()
    plan = genpk_amd.Plan(d, "f64")

    def loss_of(p_np, m_np, target_t):
        tp = torch.from_numpy(p_np).cuda()
        tm = torch.from_numpy(m_np).cuda()
        P = power_spectrum(tp, box, d, nrbins, masses=tm, total_mass=float(m_np.astype(np.float64).sum()), plan=plan)
        return c3_loss(P, target_t, cnt).item(), P

    with torch.no_grad():
        _, P0 = loss_of(pos, m, torch.ones(nrbins, dtype=torch.float64, device="cuda"))
    target_t = torch.from_numpy(_target(P0.cpu().numpy())).cuda()
    tp = torch.from_numpy(pos).cuda().requires_grad_(True)
    tm = torch.from_numpy(m).cuda().requires_grad_(True)
    P = power_spectrum(tp, box, d, nrbins, masses=tm, total_mass=tm.double().sum(), plan=plan)
    gp, gm = torch.autograd.grad(c3_loss(P, target_t, cnt), (tp, tm))
    gp = gp.cpu().numpy().astype(np.float64)
    gm = gm.cpu().numpy().astype(np.float64)
    rng = np.random.default_rng(5)
    errs = []
    with torch.no_grad():
        checked = 0
        for j in rng.permutation(pos.shape[0]):
            a = int(rng.integers(3))
            u = float(pos[j, a]) * d / box
            frac = u - np.floor(u)
            if min(frac, 1 - frac) < 0.01 + 2e-4:
                continue
            h = np.float32(1e-4 * box / d)
            pp, pm = pos.copy(), pos.copy()
            pp[j, a] = np.float32(pos[j, a] + h)
            pm[j, a] = np.float32(pos[j, a] - h)
            step = float(pp[j, a]) - float(pm[j, a])
            fd = (loss_of(pp, m, target_t)[0] - loss_of(pm, m, target_t)[0]) / step
            errs.append(abs(fd - gp[j, a]) / np.abs(gp).max())
            assert abs(fd - gp[j, a]) <= 1e-4 * np.abs(gp).max() + 1e-3 * abs(gp[j, a]), ("pos", j, a, fd, gp[j, a])
            checked += 1
            if checked == 64:
                break
        assert checked == 64
        for j in rng.choice(pos.shape[0], 64, replace=False):
            mp, mm = m.copy(), m.copy()
            h = np.float32(1e-4 * m[j])
            mp[j] = np.float32(m[j] + h)
            mm[j] = np.float32(m[j] - h)
            step = float(mp[j]) - float(mm[j])
            fd = (loss_of(pos, mp, target_t)[0] - loss_of(pos, mm, target_t)[0]) / step
            errs.append(abs(fd - gm[j]) / np.abs(gm).max())
            assert abs(fd - gm[j]) <= 1e-4 * np.abs(gm).max() + 1e-3 * abs(gm[j]), ("mass", j, fd, gm[j])
    print(f"max FD deviation / max|g|: {max(errs):.3e}", flush=True)


def test_total_mass_gradient(gpu):
    """dL/dM = -2 sum_b lambda_b P_b / M for a total mass passed as its own tensor."""
    import torch
    d, box, n_side, nrbins = 32, 100.0, 16, 32
    announce("adjoint_total_mass", f"d={d}")
    pos, m = _instance(n_side, d, box, seed=15)
    tp = torch.from_numpy(pos).cuda()
    tm = torch.from_numpy(m).cuda()
    M = torch.tensor(float(m.astype(np.float64).sum()), dtype=torch.float64, requires_grad=True)
    P = power_spectrum(tp, box, d, nrbins, masses=tm, total_mass=M, precision="f64")
    w = torch.arange(nrbins, dtype=torch.float64, device="cuda")
    (gM,) = torch.autograd.grad((P * w).sum(), (M,))
    expect = -2.0 * (P.detach() * w).sum().item() / M.item()
    assert abs(gM.item() - expect) <= 1e-12 * abs(expect)


@pytest.mark.parametrize("d", [48, 512])
def test_redeposit_and_stale_order(gpu, d):
    """d = 48 (one-pass tile sort) and 512 (two passes).  (a) redeposit()
    from the tile order a retaining deposit into an empty plan kept, and
    from the order a gather made (work items only on occupied tiles, so the
    store is zeroed first), gives that deposit's density bit for bit (fixed
    point: the sums do not depend on which records form which item);
    (b) positions changed in place after the retaining deposit: the gather
    partitions them again instead of reusing the stale order, so the
    gradients equal a fresh plan's on the same dL/drho."""
    import torch
    box = 100.0
    announce("redeposit and stale order", f"d={d}")
    pos, m = _instance(24, d, box, seed=17)
    tp = torch.from_numpy(pos).cuda()
    tm = torch.from_numpy(m).cuda()
    a = genpk_amd.Plan(d, "f64")
    a.clear()
    a.deposit(tp, box, tm, retain=True)
    ref = a.export_density().clone()
    a.clear()
    a.redeposit()
    assert torch.equal(a.export_density(), ref)
    # an order made by the gather of an imported dL/drho
    G = torch.randn(ref.shape, dtype=torch.float64, device="cuda", generator=torch.Generator(device="cuda").manual_seed(3))
    b = genpk_amd.Plan(d, "f64")
    b.import_density(G)
    b.adjoint_gather(tp, box, tm)
    b.clear()
    b.redeposit()
    assert torch.equal(b.export_density(), ref)
    # (b) stale order
    a.clear()
    a.deposit(tp, box, tm, retain=True)
    tp.add_(0.37 * box / d)
    tp.remainder_(box)
    assert not a.keeps_order_of(tp, box, tm)
    a.import_density(G)
    g_a = a.adjoint_gather(tp, box, tm)
    c = genpk_amd.Plan(d, "f64")
    c.import_density(G)
    g_c = c.adjoint_gather(tp, box, tm)
    assert torch.equal(g_a[0], g_c[0]) and torch.equal(g_a[1], g_c[1])


@pytest.mark.parametrize("prec,bar", [("f64", 1e-6), ("f32", 2e-3)])
def test_adjoint_zeldovich_band_power_loss(gpu, prec, bar):
    """The differentiable workload's own check on a 32^3 instance: 32^3
    Zel'dovich particles (k^-2 displacement, 1-cell rms, seed 1) on a 32^3
    grid with the cell size of 256^3 in 250 Mpc/h, masses 1 (a tensor), the
    band-power loss against the P of a second seed (seed 2), sigma_b =
    P_target,b sqrt(2/count_b).  Gradients w.r.t. positions and masses (recompute
    "always") against the FP64 dense PyTorch reference, and, for the FP64 plan,
    against central differences of the GPU loss on 32 coordinates at least
    0.01 cell from a cell face."""
    import torch
    n_side, d = 32, 32
    box = 250.0 * d / 256
    announce("adjoint zeldovich band-power loss", prec, f"d={d} n={n_side}^3")
    pos = synth.zeldovich(n_side, box, d, rms_cells=1.0, seed=1)
    m = np.ones(pos.shape[0], dtype=np.float32)
    tables = mode_tables(d, d)
    ref_pos = synth.zeldovich(n_side, box, d, rms_cells=1.0, seed=2)
    target, _, _, _ = _dense_grads(ref_pos, m, box, d, d, np.ones(d), tables)
    Pd, Ld, gpd, gmd = _dense_grads(pos, m, box, d, d, target, tables)
    plan = genpk_amd.Plan(d, prec)
    Pg, Lg, gpg, gmg = _gpu_loss_and_grads(torch, pos, m, box, d, d, target, prec, plan, recompute="always")
    assert abs(Lg / Ld - 1) <= (1e-8 if prec == "f64" else 1e-3)
    okp, ep = _close(gpg, gpd, bar)
    okm, em = _close(gmg, gmd, bar)
    print(f"loss {Lg:.6e}; max rel grad err pos {ep:.3e} mass {em:.3e}", flush=True)
    assert okp and okm
    if prec != "f64":
        return
    cnt = torch.from_numpy(tables[2].numpy()).cuda()
    tt = torch.from_numpy(target).cuda()
    rng = np.random.default_rng(8)
    checked = 0
    with torch.no_grad():
        for j in rng.permutation(pos.shape[0]):
            a = int(rng.integers(3))
            u = float(pos[j, a]) * d / box
            if min(u - np.floor(u), np.ceil(u) - u) < 0.01 + 2e-4:
                continue
            h = np.float32(1e-4 * box / d)
            ls = []
            for sgn in (1, -1):
                p2 = pos.copy()
                p2[j, a] = np.float32(pos[j, a] + sgn * h)
                P = power_spectrum(torch.from_numpy(p2).cuda(), box, d, d, masses=torch.from_numpy(m).cuda(),
                                   total_mass=float(m.sum()), plan=plan)
                ls.append((c3_loss(P, tt, cnt).item(), float(p2[j, a])))
            fd = (ls[0][0] - ls[1][0]) / (ls[0][1] - ls[1][1])
            assert abs(fd - gpg[j, a]) <= 1e-4 * np.abs(gpg).max() + 1e-3 * abs(gpg[j, a]), (j, a, fd, gpg[j, a])
            checked += 1
            if checked == 32:
                break
    assert checked == 32
