"""The device generators of the synthetic workloads (run here on the CPU
device) against the numpy recipes: the Zel'dovich displacement from the
same white noise, the lattice placement and wrapping, and the
displacement's properties (curl-free k^-2 spectrum, rms, zero mean)."""
import numpy as np
import torch

from genpk_amd import synth


def test_zeldovich_torch_matches_numpy_recipe():
    n, box, d, seed = 16, 50.0, 16, 3
    psi_t = synth.zeldovich_displacement_torch(n, 1.0, box, d, seed=seed, device="cpu").numpy()
    g = torch.Generator(device="cpu")
    g.manual_seed(seed)
    w = torch.randn((n, n, n), generator=g, dtype=torch.float64).numpy()
    psi_n = synth.zeldovich_displacement(n, 1.0, box, d, noise=w)
    assert np.abs(psi_t - psi_n).max() <= 1e-12 * np.abs(psi_n).max()
    # 1-cell rms over the components, zero mean
    assert abs(np.sqrt(np.mean(psi_n ** 2)) - box / d) <= 1e-12 * box / d
    assert np.abs(psi_n.mean(axis=0)).max() <= 1e-12 * box / d
    # positions: lattice (+ offset) + psi, wrapped, as the numpy recipe
    for off in (0.0, 0.5):
        pt = synth.zeldovich_torch(n, box, d, offset_cells=off, psi=torch.from_numpy(psi_n), device="cpu").numpy()
        pn = synth.zeldovich(n, box, d, offset_cells=off, psi=psi_n)
        assert pt.dtype == np.float32 and pt.shape == (n ** 3, 3)
        assert np.array_equal(pt, pn)


def test_zeldovich_displacement_is_a_gradient_field():
    """psi_hat = i k w_hat / k^2: k x psi_hat = 0 (curl free) and
    k . psi_hat = i w_hat up to the rms scale (the divergence is the noise)."""
    n = 12
    rng = np.random.default_rng(5)
    w = rng.standard_normal((n, n, n))
    psi = synth.zeldovich_displacement(n, 1.0, 1.0, n, noise=w)
    ph = [np.fft.rfftn(psi[:, a].reshape(n, n, n)) for a in range(3)]
    k = np.fft.fftfreq(n) * n
    kz = np.fft.rfftfreq(n) * n
    KX, KY, KZ = np.meshgrid(k, k, kz, indexing="ij")
    curl = np.abs(KY * ph[2] - KZ * ph[1]) + np.abs(KZ * ph[0] - KX * ph[2]) + np.abs(KX * ph[1] - KY * ph[0])
    div = (KX * ph[0] + KY * ph[1] + KZ * ph[2]) / 1j
    wk = np.fft.rfftn(w)
    wk[0, 0, 0] = 0
    # the Nyquist planes of an even grid are not curl free after irfftn's symmetrisation; check the others
    inner = (np.abs(KX) < n / 2) & (np.abs(KY) < n / 2) & (KZ < n / 2)
    scale = np.vdot(wk[inner], div[inner]).real / np.vdot(wk[inner], wk[inner]).real
    assert scale > 0
    assert curl[inner].max() <= 1e-9 * np.abs(div).max()
    assert np.abs(div[inner] - scale * wk[inner]).max() <= 1e-9 * np.abs(div).max()
