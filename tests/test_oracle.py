"""Pins the CPU oracle (oracle/genpk_oracle.c) before it is trusted as the checker.

* Known answers of the reference's own tests (test.cpp:31-86).
* Facts recorded from the compiled reference (mode-count table): the
  bin of the Nyquist corner mode and the number of occupied bins.
* The DFT definition (brute force) and numpy's pocketfft for the FFT.
* Closed forms and invariants of the CIC deposit and of the binning.
"""
import numpy as np
import pytest

import oracle


def near(x, y):
    # FLOATS_NEAR_TO of test.cpp:25-26
    return abs(x - y) <= max(abs(x), abs(y)) / 1e5


# ---------------------------------------------------------------- test.cpp KATs

def test_kat_fieldize():
    # test.cpp:31-50
    dims = 5
    field = np.zeros(2 * dims * dims * (dims // 2 + 1))
    pos = np.array([i / 3.0 for i in range(30)], dtype=np.float32)
    masses = np.full(30, 10.0, dtype=np.float32)
    oracle.fieldize(10, dims, field, pos, masses, 10.0, 1)
    assert near(field[0], 8.61111)
    assert field[3] == 0 and field[20] == 0 and field[125] == 0
    assert near(field[124], 1.66666)


def test_kat_invwindow():
    # test.cpp:52-57
    assert near(oracle.invwindow(0, 3, 4, 5), 71.8177719)
    assert near(oracle.invwindow(4, 4, 4, 5), 6111.20801)
    assert oracle.invwindow(1, 1, 1, 0) == 0


def kat_powerspectrum_field():
    # test.cpp:61-72: half-box of ones in the padded layout, field[0] = 2
    field = np.zeros((4, 4, 6))
    f = field.reshape(-1)
    for i in range(32):
        ii = i // 4
        f[6 * ii + i % 4] = 1
    f[0] = 2
    return field


def test_kat_powerspectrum():
    # test.cpp:59-86
    F = oracle.fft_r2c_3d(kat_powerspectrum_field())
    p, c, k = oracle.powerspectrum(4, F, F, 10, 64.0, 64.0)
    assert near(k[2], np.sqrt(2))
    assert c[2] == 12 and c[1] == 0 and c[0] == 6
    assert near(p[0], 0.0677526)
    assert p[1] == 0
    assert near(p[2], 0.000565561)
    assert near(p[9], 0.0550908)


# ------------------------------------------------- compiled-reference bin facts

# Recorded from the reference's powerspectrum() compiled with its Makefile
# flags: the Nyquist corner (d/2,d/2,d/2) lands in bin nrbins-1 except at
# d = 1024, where it lands in 1022 and bin 1023 stays empty; occupied bins
# 104/199/295/389/757 at d = 128/256/384/512/1024 (nrbins = d).
CORNER_BIN = {128: 127, 256: 255, 384: 383, 512: 511, 1024: 1022}
OCCUPIED = {128: 104, 256: 199, 384: 295, 512: 389}


@pytest.mark.parametrize("d", sorted(CORNER_BIN))
def test_corner_bin(d):
    assert oracle.bin_of_q(d, d, 3 * (d // 2) ** 2) == CORNER_BIN[d]


@pytest.mark.parametrize("d", sorted(OCCUPIED))
def test_counts_occupied_and_total(d):
    F = np.zeros((d, d, d // 2 + 1), np.complex128)
    p, c, k = oracle.powerspectrum(d, F, F, d, 1.0, 1.0)
    assert int((c > 0).sum()) == OCCUPIED[d]
    assert int(c.sum(dtype=np.int64)) == d ** 3 - 1
    assert c[CORNER_BIN[d]] >= 1
    # keff of an occupied bin lies inside the bin's k range
    bpu = (d - 1) / np.log(np.sqrt(3) * d / 2.0)
    occ = np.nonzero(c)[0]
    assert np.all(np.floor(bpu * np.log(k[occ]) + 1e-9) <= occ + 1e-9)


def test_counts_1024_corner_layout():
    # bin 1023 is empty at d = 1024: no q maps there
    d = 1024
    qs = np.arange(1, 3 * 512 ** 2 + 1)
    bins = {oracle.bin_of_q(d, d, int(q)) for q in qs[-4000:]}
    assert max(bins) == 1022


# ------------------------------------------------------------------------ FFT

def brute_rdft3(x):
    d = x.shape[0]
    n = np.arange(d)
    W = np.exp(-2j * np.pi * np.outer(n, n) / d)
    X = np.einsum("ia,jb,kc,abc->ijk", W, W, W, x)
    return X[:, :, : d // 2 + 1]


@pytest.mark.parametrize("d", [2, 4, 6, 8, 10])
def test_fft_bruteforce(d):
    rng = np.random.default_rng(d)
    x = rng.standard_normal((d, d, d))
    fld = np.zeros((d, d, 2 * (d // 2 + 1)))
    fld[:, :, :d] = x
    F = oracle.fft_r2c_3d(fld)
    R = brute_rdft3(x)
    assert np.abs(F - R).max() <= 1e-12 * max(1.0, np.abs(R).max())


@pytest.mark.parametrize("d", [12, 16, 30, 48, 64, 96])
def test_fft_pocketfft(d):
    rng = np.random.default_rng(100 + d)
    x = rng.standard_normal((d, d, d))
    fld = np.zeros((d, d, 2 * (d // 2 + 1)))
    fld[:, :, :d] = x
    F = oracle.fft_r2c_3d(fld)
    R = np.fft.rfftn(x)
    assert np.abs(F - R).max() <= 1e-13 * np.abs(R).max()


@pytest.mark.parametrize("d", [4, 8, 12, 30])
def test_fft_roundtrip(d):
    rng = np.random.default_rng(200 + d)
    x = rng.standard_normal((d, d, d))
    fld = np.zeros((d, d, 2 * (d // 2 + 1)))
    fld[:, :, :d] = x
    oracle.fft_r2c_3d(fld)
    oracle.fft_c2r_3d(fld)
    assert np.abs(fld[:, :, :d] / d ** 3 - x).max() < 1e-12


def test_fft_plane_wave_spike():
    d = 16
    z = np.arange(d)
    x = np.zeros((d, d, d))
    x[:] = np.cos(2 * np.pi * 3 * z / d)[None, None, :]
    fld = np.zeros((d, d, d + 2))
    fld[:, :, :d] = x
    F = oracle.fft_r2c_3d(fld)
    A = np.abs(F)
    assert abs(A[0, 0, 3] - d ** 3 / 2) < 1e-9
    A[0, 0, 3] = 0
    assert A.max() < 1e-9


# ---------------------------------------------------------------- CIC closed forms

def test_cic_on_grid_point_and_centre():
    d, box = 8, 8.0
    f = oracle.new_field(d)
    oracle.fieldize(box, d, f, np.array([[2.0, 3.0, 4.0]], np.float32), None, 1.5)
    assert f[2, 3, 4] == 1.5 and np.count_nonzero(f) == 1
    f = oracle.new_field(d)
    oracle.fieldize(box, d, f, np.array([[2.5, 3.5, 4.5]], np.float32), None, 1.0)
    assert np.allclose(f[2:4, 3:5, 4:6], 0.125) and abs(f.sum() - 1) < 1e-15


def test_cic_weights_general_point():
    d, box = 8, 8.0
    x = np.array([[1.25, 6.75, 7.5]], np.float32)  # wraps in z
    f = oracle.new_field(d)
    oracle.fieldize(box, d, f, x, None, 2.0)
    wx = {1: 0.75, 2: 0.25}
    wy = {6: 0.25, 7: 0.75}
    wz = {7: 0.5, 0: 0.5}
    for i, a in wx.items():
        for j, b in wy.items():
            for k, c in wz.items():
                assert abs(f[i, j, k] - 2.0 * a * b * c) < 1e-15
    assert abs(f.sum() - 2.0) < 1e-14


def test_cic_invariants():
    rng = np.random.default_rng(7)
    d, box, n = 12, 30.0, 4000
    pos = (rng.random((n, 3)) * box).astype(np.float32)
    m = (rng.random(n) + 0.5).astype(np.float32)
    f = oracle.new_field(d)
    oracle.fieldize(box, d, f, pos, m, 0.0)
    # mass conservation, pads untouched
    assert abs(f[:, :, :d].sum() - m.astype(np.float64).sum()) < 1e-9
    assert np.all(f[:, :, d:] == 0)
    # linearity in mass
    g = oracle.new_field(d)
    oracle.fieldize(box, d, g, pos, None, 3.0)
    h = oracle.new_field(d)
    oracle.fieldize(box, d, h, pos, None, 1.0)
    assert np.allclose(g, 3 * h, rtol=1e-14, atol=1e-14)
    # translation by whole cells = roll of the grid
    cell = box / d
    shifted = np.mod(pos.astype(np.float64) + np.array([2, 5, 11]) * cell, box).astype(np.float32)
    s = oracle.new_field(d)
    oracle.fieldize(box, d, s, shifted, m, 0.0)
    rolled = np.roll(f[:, :, :d], shift=(2, 5, 11), axis=(0, 1, 2))
    assert np.abs(s[:, :, :d] - rolled).max() < 1e-4  # float32 re-rounding of shifted positions


def test_cic_periodic_wrap_and_accumulate():
    d, box = 6, 6.0
    p = np.array([[0.25, 0.5, 5.75]], np.float32)
    a = oracle.new_field(d)
    oracle.fieldize(box, d, a, p, None, 1.0)
    b = oracle.new_field(d)
    oracle.fieldize(box, d, b, p + np.float32(box), None, 1.0)   # one box further
    c = oracle.new_field(d)
    oracle.fieldize(box, d, c, p - np.float32(box), None, 1.0)   # negative positions
    assert np.allclose(a, b, atol=1e-6) and np.allclose(a, c, atol=1e-6)
    oracle.fieldize(box, d, a, p, None, 1.0)                     # accumulates
    assert abs(a.sum() - 2.0) < 1e-12


@pytest.mark.parametrize("bad", [np.nan, np.inf, -np.inf, 3e9])
def test_cic_nonfinite_coordinate_cells(bad):
    """fieldize.cpp:67 stores floor(x) in an int: on x86-64 a NaN, an infinity
    or a floor beyond the int range converts to INT_MIN, and the cells are
    then INT_MIN and INT_MIN + 1 taken modulo d with C's truncating %, made
    non-negative (:70-75).  For d = 12 that is cells 4 and 5 (INT_MIN % 12 =
    -8).  A non-finite coordinate makes all eight weights non-finite; a
    huge finite one gives finite weights at those cells."""
    d, box = 12, 12.0
    imin = -2 ** 31

    def cmod(a, m):  # C's % (truncation toward zero), then the reference's fix-up
        r = a - m * int(a / m)
        return r + m if r < 0 else r

    assert (cmod(imin, d), cmod(imin + 1, d)) == (4, 5)
    p = np.array([[bad, 2.5, 7.25]], np.float32)
    f = oracle.new_field(d)
    oracle.fieldize(box, d, f, p, None, 1.0)
    touched = np.zeros_like(f, dtype=bool)
    for x in (4, 5):
        for y in (2, 3):
            for z in (7, 8):
                touched[x, y, z] = True
    if np.isfinite(bad):
        assert np.array_equal(f != 0, touched & (f != 0)) and (f[touched] != 0).any()
    else:
        assert np.array_equal(~np.isfinite(f), touched)


# ---------------------------------------------------------------- binning closed forms

def test_plane_wave_power_closed_form():
    d = 16
    k0 = (1, 2, 3)
    x = np.arange(d)
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    A = 0.5
    rho = A * np.cos(2 * np.pi * (k0[0] * X + k0[1] * Y + k0[2] * Z) / d)
    fld = np.zeros((d, d, d + 2))
    fld[:, :, :d] = rho
    F = oracle.fft_r2c_3d(fld)
    M = 7.0
    p, c, k = oracle.powerspectrum(d, F, F, d, M, M)
    q = sum(v * v for v in k0)
    b = oracle.bin_of_q(d, d, q)
    # one mode of H (kz = 3 interior, weight 2) carries |F| = A d^3 / 2
    expect = 2 * (A * d ** 3 / 2) ** 2 * oracle.invwindow(*k0, d) ** 2 / (M * M) / c[b]
    assert abs(p[b] - expect) <= 1e-10 * expect
    others = np.delete(p, b)
    assert np.abs(others).max() <= 1e-18 * expect


def test_power_translation_invariant():
    rng = np.random.default_rng(11)
    d, box, n = 16, 16.0, 3000
    pos = (rng.random((n, 3)) * box).astype(np.float32)
    p1, c1, _ = oracle.pk_from_particles(box, d, pos)
    shifted = np.mod(pos + np.float32(3.0), np.float32(box)).astype(np.float32)
    p2, c2, _ = oracle.pk_from_particles(box, d, shifted)
    assert np.array_equal(c1, c2)
    occ = c1 > 0
    assert np.allclose(p1[occ], p2[occ], rtol=1e-4)


def test_cross_equals_auto_for_identical_fields():
    rng = np.random.default_rng(3)
    d = 8
    fld = np.zeros((d, d, d + 2))
    fld[:, :, :d] = rng.random((d, d, d))
    F = oracle.fft_r2c_3d(fld)
    G = F.copy()
    pa, ca, ka = oracle.powerspectrum(d, F, F, d, 2.0, 2.0)
    px, cx, kx = oracle.powerspectrum(d, F, G, d, 2.0, 2.0)
    # bins merge thread-private sums in arbitrary order: equal to rounding
    assert np.array_equal(ca, cx)
    assert np.allclose(pa, px, rtol=1e-14, atol=0) and np.allclose(ka, kx, rtol=1e-14, atol=0)
