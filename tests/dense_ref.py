"""FP64 dense PyTorch reference of the forward power spectrum (CPU), with
autograd: the gradient reference for the hand-written adjoint.

Differentiable CIC (index_add of mass x weight at the 8 corners,
fieldize.cpp:63-92), torch.fft.rfftn, and binning by index_add with the
bin index, Hermitian weight and deconvolution factor of every half-spectrum
mode taken from the CPU oracle (powerspectrum.c:35-110).
"""
import numpy as np
import torch

import oracle


def mode_tables(d, nrbins):
    """Per half-spectrum mode: bin index, h * invwindow^2 (0 for DC), and the bin counts."""
    hz = d // 2 + 1
    kv = lambda n: n if n <= d // 2 else n - d
    k = np.array([kv(n) for n in range(d)])
    kz = np.arange(hz)
    KX, KY, KZ = np.meshgrid(k, k, kz, indexing="ij")
    q = (KX ** 2 + KY ** 2 + KZ ** 2).reshape(-1)
    qs = np.unique(q)
    qbin = {int(v): (oracle.bin_of_q(d, nrbins, int(v)) if v > 0 else 0) for v in qs}
    binidx = np.array([qbin[int(v)] for v in q], dtype=np.int64)
    h = np.where((KZ == 0) | (KZ == d // 2), 1.0, 2.0).reshape(-1)
    ww = {}
    dec = np.empty(q.size)
    flat = np.stack([KX.reshape(-1), KY.reshape(-1), KZ.reshape(-1)], 1)
    for i, (a, b, c) in enumerate(flat):
        key = (abs(a), abs(b), c)
        if key not in ww:
            w = oracle.invwindow(a, b, c, d)
            ww[key] = w * w
        dec[i] = ww[key]
    hdec = h * dec
    hdec[q == 0] = 0.0
    count = np.zeros(nrbins, np.int64)
    np.add.at(count, binidx[q > 0], h[q > 0].astype(np.int64))
    return torch.from_numpy(binidx), torch.from_numpy(hdec), torch.from_numpy(count)


def cic_density(pos, masses, box, d):
    """Differentiable CIC deposit, FP64, dense [d^3]."""
    units = d / box
    u = pos * units
    fl = torch.floor(u).detach()
    f = u - fl
    i0 = torch.remainder(fl.long(), d)
    i1 = torch.remainder(i0 + 1, d)
    rho = torch.zeros(d * d * d, dtype=torch.float64)
    for c in range(8):
        bx, by, bz = c & 1, (c >> 1) & 1, (c >> 2) & 1
        w = (f[:, 0] if bx else 1 - f[:, 0]) * (f[:, 1] if by else 1 - f[:, 1]) * (f[:, 2] if bz else 1 - f[:, 2])
        ix = i1[:, 0] if bx else i0[:, 0]
        iy = i1[:, 1] if by else i0[:, 1]
        iz = i1[:, 2] if bz else i0[:, 2]
        rho = rho.index_add(0, (ix * d + iy) * d + iz, masses * w)
    return rho


def dense_power(pos, masses, box, d, nrbins, M, tables=None):
    binidx, hdec, count = tables if tables is not None else mode_tables(d, nrbins)
    rho = cic_density(pos, masses, box, d)
    F = torch.fft.rfftn(rho.reshape(d, d, d))
    p2 = (F.real ** 2 + F.imag ** 2).reshape(-1) * hdec
    S = torch.zeros(nrbins, dtype=torch.float64).index_add(0, binidx, p2)
    cnt = count.to(torch.float64)
    P = torch.where(cnt > 0, S / (M * M) / torch.clamp(cnt, min=1), torch.zeros_like(S))
    return P, count


def c3_loss(P, target, count):
    """sum_b w_b (P_b - T_b)^2 / sigma_b^2, sigma_b = T_b sqrt(2/count_b), w_b = 1 on occupied bins."""
    cnt = count.to(torch.float64)
    occ = (cnt > 0) & (target > 0)
    sig2 = torch.where(occ, target ** 2 * 2.0 / torch.clamp(cnt, min=1), torch.ones_like(target))
    return torch.where(occ, (P - target) ** 2 / sig2, torch.zeros_like(P)).sum()
