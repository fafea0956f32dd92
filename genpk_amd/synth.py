"""Seeded synthetic particle sets of the benchmark workloads.

All generators return float32 positions [n, 3] in [0, box).  The numpy
versions are used for the sizes the CPU oracle can check; ``clumped_torch``
builds the 1024^3-particle clustered workload directly on the GPU.

* ``gaussian_lattice``: n_side^3 lattice, spacing box/n_side, iid Gaussian
  offsets of ``sigma_cells`` cells of a ``dims`` grid, wrapped periodically.
* ``zeldovich``: lattice displaced by a Zel'dovich field built from seeded
  white noise, psi_hat_a = i k_a / |k|^2 w_hat (zero at k = 0), scaled to
  ``rms_cells`` cells rms per component.
* ``clumped``: half uniform, half in ``n_clumps`` Gaussian clumps with
  power-law (pdf ~ n^-1.9) member counts and radii U[0.5, 4] cells.

The ``*_torch`` versions build the same recipes on the device (torch's
generators and FFT are used for input generation only).
"""
import hashlib

import numpy as np


def sha256_of(pos):
    return hashlib.sha256(np.ascontiguousarray(pos, dtype=np.float32).tobytes()).hexdigest()


def _lattice(n_side, box, offset_cells=0.0):
    g = (np.arange(n_side, dtype=np.float64) + offset_cells) * (box / n_side)
    x, y, z = np.meshgrid(g, g, g, indexing="ij")
    return np.stack([x.ravel(), y.ravel(), z.ravel()], axis=1)


def gaussian_lattice(n_side, box, dims, sigma_cells, seed=0):
    rng = np.random.default_rng(seed)
    p = _lattice(n_side, box)
    p += rng.standard_normal(p.shape) * (sigma_cells * box / dims)
    return np.mod(p, box).astype(np.float32)


def zeldovich_displacement(n_side, rms_cells, box, dims, seed=0, noise=None):
    """Displacement field [n_side^3, 3] (float64, box units) from seeded
    white noise (or the given ``noise`` [n_side]^3)."""
    w = np.random.default_rng(seed).standard_normal((n_side, n_side, n_side)) if noise is None else noise
    wk = np.fft.rfftn(w)
    k = np.fft.fftfreq(n_side) * n_side
    kz = np.fft.rfftfreq(n_side) * n_side
    KX, KY, KZ = np.meshgrid(k, k, kz, indexing="ij")
    k2 = KX ** 2 + KY ** 2 + KZ ** 2
    k2[0, 0, 0] = 1.0
    out = []
    for ka in (KX, KY, KZ):
        ph = 1j * ka / k2 * wk
        ph[0, 0, 0] = 0
        out.append(np.fft.irfftn(ph, s=(n_side,) * 3, axes=(0, 1, 2)).ravel())
    psi = np.stack(out, axis=1)
    psi *= (rms_cells * box / dims) / np.sqrt(np.mean(psi ** 2))
    return psi


def zeldovich(n_side, box, dims, rms_cells=1.0, seed=1, offset_cells=0.0, psi=None):
    if psi is None:
        psi = zeldovich_displacement(n_side, rms_cells, box, dims, seed)
    p = _lattice(n_side, box, offset_cells) + psi
    return np.mod(p, box).astype(np.float32)


def clump_counts(n_clumped, n_clumps, rng, slope=1.9):
    """Member counts ~ pdf n^-slope, scaled to n_clumped with largest-remainder rounding."""
    u = rng.random(n_clumps)
    raw = (1.0 - u) ** (-1.0 / (slope - 1.0))  # Pareto, x_min = 1
    share = raw / raw.sum() * n_clumped
    base = np.floor(share).astype(np.int64)
    rem = n_clumped - base.sum()
    if rem > 0:
        base[np.argsort(-(share - base))[:rem]] += 1
    return base


def clumped(n, box, dims, n_clumps=100000, seed=4, r_min=0.5, r_max=4.0):
    rng = np.random.default_rng(seed)
    nu = n // 2
    nc = n - nu
    uni = rng.random((nu, 3)) * box
    counts = clump_counts(nc, n_clumps, rng)
    centres = rng.random((n_clumps, 3)) * box
    radii = rng.uniform(r_min, r_max, n_clumps) * (box / dims)
    owner = np.repeat(np.arange(n_clumps), counts)
    cl = centres[owner] + rng.standard_normal((nc, 3)) * radii[owner, None]
    p = np.concatenate([uni, cl], axis=0)
    return np.mod(p, box).astype(np.float32)


def clumped_torch(n, box, dims, n_clumps=100000, seed=4, r_min=0.5, r_max=4.0, device="cuda"):
    """Same recipe as ``clumped`` generated on the GPU (torch RNG, seeded)."""
    import torch
    g = torch.Generator(device=device)
    g.manual_seed(seed)
    nu = n // 2
    nc = n - nu
    rng = np.random.default_rng(seed)
    counts = torch.from_numpy(clump_counts(nc, n_clumps, rng)).to(device)
    centres = torch.rand((n_clumps, 3), generator=g, device=device, dtype=torch.float64) * box
    radii = (torch.rand(n_clumps, generator=g, device=device, dtype=torch.float64) * (r_max - r_min) + r_min) * (box / dims)
    out = torch.empty((n, 3), dtype=torch.float32, device=device)
    step = 1 << 26
    for s in range(0, nu, step):
        e = min(nu, s + step)
        out[s:e] = (torch.rand((e - s, 3), generator=g, device=device, dtype=torch.float64) * box).float()
    owner = torch.repeat_interleave(torch.arange(n_clumps, device=device), counts)
    for s in range(0, nc, step):
        e = min(nc, s + step)
        o = owner[s:e]
        p = centres[o] + torch.randn((e - s, 3), generator=g, device=device, dtype=torch.float64) * radii[o, None]
        out[nu + s:nu + e] = torch.remainder(p, box).float()
    del owner
    return out


def zeldovich_displacement_torch(n_side, rms_cells, box, dims, seed=1, device="cuda"):
    """``zeldovich_displacement`` computed on the device in FP64 (torch's
    Philox white noise, torch.fft): psi_hat_a = i k_a / |k|^2 w_hat, zero at
    k = 0, scaled to ``rms_cells`` cells (of a ``dims`` grid) rms per
    component.  Returns float64 [n_side^3, 3] in box units."""
    import torch
    g = torch.Generator(device=device)
    g.manual_seed(int(seed))
    n = n_side
    w = torch.randn((n, n, n), generator=g, dtype=torch.float64, device=device)
    wk = torch.fft.rfftn(w)
    del w
    k = torch.fft.fftfreq(n, d=1.0 / n, dtype=torch.float64, device=device)
    kz = torch.fft.rfftfreq(n, d=1.0 / n, dtype=torch.float64, device=device)
    k2 = k[:, None, None] ** 2 + k[None, :, None] ** 2 + kz[None, None, :] ** 2
    k2[0, 0, 0] = 1.0
    base = wk / k2
    base[0, 0, 0] = 0
    del wk, k2
    psi = torch.empty((n * n * n, 3), dtype=torch.float64, device=device)
    shapes = ((n, 1, 1), (1, n, 1), (1, 1, n // 2 + 1))
    for a, ka in enumerate((k, k, kz)):
        ph = base * (1j * ka.reshape(shapes[a]))
        psi[:, a] = torch.fft.irfftn(ph, s=(n, n, n)).reshape(-1)
        del ph
    psi *= (rms_cells * box / dims) / torch.sqrt(torch.mean(psi * psi))
    return psi


def zeldovich_torch(n_side, box, dims, rms_cells=1.0, seed=1, offset_cells=0.0, psi=None, device="cuda"):
    """``zeldovich`` on the device: an n_side^3 lattice (offset by
    ``offset_cells`` lattice spacings) displaced by psi (built from ``seed``
    unless given), wrapped into [0, box), float32 [n_side^3, 3]."""
    import torch
    if psi is None:
        psi = zeldovich_displacement_torch(n_side, rms_cells, box, dims, seed, device)
    ar = (torch.arange(n_side, dtype=torch.float64, device=device) + offset_cells) * (box / n_side)
    out = torch.empty((n_side ** 3, 3), dtype=torch.float32, device=device)
    per = n_side * n_side  # one x-plane of the lattice at a time
    yy, zz = torch.meshgrid(ar, ar, indexing="ij")
    yz = torch.stack([yy.reshape(-1), zz.reshape(-1)], 1)
    for i in range(n_side):
        s = slice(i * per, (i + 1) * per)
        p = psi[s].clone()
        p[:, 0] += ar[i]
        p[:, 1:] += yz
        out[s] = torch.remainder(p, box).float()
    return out


def gaussian_lattice_torch(n_side, box, dims, sigma_cells, seed=0, device="cuda"):
    """gaussian_lattice generated on the device (torch's Philox generator): a
    cubic lattice of n_side^3 points with iid N(0, (sigma_cells*box/dims)^2)
    offsets per axis, wrapped into [0, box), float32 [n_side^3, 3]."""
    import torch
    g = torch.Generator(device=device)
    g.manual_seed(int(seed))
    ar = torch.arange(n_side, dtype=torch.float64, device=device) * (box / n_side)
    x, y, z = torch.meshgrid(ar, ar, ar, indexing="ij")
    p = torch.stack([x.reshape(-1), y.reshape(-1), z.reshape(-1)], 1)
    p += torch.randn(p.shape, generator=g, dtype=torch.float64, device=device) * (sigma_cells * box / dims)
    return torch.remainder(p, box).to(torch.float32).contiguous()


def uneven_realisations(count=64, seed=100, n_side_range=(64, 256), dims_choices=(128, 256, 384, 512)):
    """Shapes of a batch of uneven realisations: for r < count, with its own
    seed (seed + r), n_side = round(exp(U[ln lo, ln hi])) (log-uniform
    particle counts n_side^3) and dims uniform over dims_choices.
    Returns [(n_side, dims, seed_r)]."""
    out = []
    lo, hi = n_side_range
    for r in range(count):
        rng = np.random.default_rng(seed + r)
        n_side = int(round(float(np.exp(rng.uniform(np.log(lo), np.log(hi))))))
        d = int(dims_choices[int(rng.integers(len(dims_choices)))])
        out.append((n_side, d, seed + r))
    return out
