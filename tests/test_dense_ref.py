"""CPU pins of the FP64 dense PyTorch reference used for the adjoint checks:
its forward equals the CPU oracle, and its autograd gradient matches
central finite differences of itself."""
import numpy as np
import torch

import oracle
from dense_ref import cic_density, dense_power, mode_tables


def test_dense_forward_matches_oracle():
    rng = np.random.default_rng(1)
    d, box, n = 16, 10.0, 3000
    pos = (rng.random((n, 3)) * box).astype(np.float32)
    m = (rng.random(n) + 0.5).astype(np.float32)
    M = float(m.astype(np.float64).sum())
    po, co, ko = oracle.pk_from_particles(box, d, pos, m, 0.0, d, M)
    P, count = dense_power(torch.from_numpy(pos.astype(np.float64)), torch.from_numpy(m.astype(np.float64)), box,
                           d, d, M)
    assert np.array_equal(count.numpy(), co)
    occ = co > 0
    assert np.allclose(P.numpy()[occ], po[occ], rtol=1e-10, atol=0)


def test_dense_density_matches_oracle():
    rng = np.random.default_rng(2)
    d, box, n = 8, 4.0, 500
    pos = (rng.random((n, 3)) * box).astype(np.float32)
    ref = oracle.new_field(d)
    oracle.fieldize(box, d, ref, pos, None, 1.0, 1)
    rho = cic_density(torch.from_numpy(pos.astype(np.float64)), torch.ones(n, dtype=torch.float64), box, d)
    assert np.allclose(rho.numpy().reshape(d, d, d), ref[:, :, :d], rtol=0, atol=1e-12)


def test_dense_gradient_fd():
    rng = np.random.default_rng(3)
    d, box, n = 8, 8.0, 200
    pos = torch.from_numpy((rng.random((n, 3)) * box)).requires_grad_(True)
    m = torch.from_numpy(rng.random(n) + 0.5).requires_grad_(True)
    tables = mode_tables(d, d)
    P, _ = dense_power(pos, m, box, d, d, 1.0, tables)
    L = (P * torch.arange(d, dtype=torch.float64)).sum()
    gp, gm = torch.autograd.grad(L, (pos, m))
    h = 1e-6
    for j, a in [(3, 0), (17, 1), (101, 2)]:
        frac = (pos[j, a].item() * d / box) % 1.0
        if min(frac, 1 - frac) < 0.01:
            continue
        pp = pos.detach().clone(); pp[j, a] += h
        pm = pos.detach().clone(); pm[j, a] -= h
        Lp = (dense_power(pp, m.detach(), box, d, d, 1.0, tables)[0] * torch.arange(d, dtype=torch.float64)).sum()
        Lm = (dense_power(pm, m.detach(), box, d, d, 1.0, tables)[0] * torch.arange(d, dtype=torch.float64)).sum()
        assert abs(((Lp - Lm) / (2 * h)).item() - gp[j, a].item()) <= 1e-6 * max(1.0, abs(gp[j, a].item()))
