"""The drop-in shims of include/genpk_compat.h driven as the reference's
gen-pk driver drives them (gen-pk.cpp:208-237, read_fieldize.cpp:43-94),
in eager mode and in resident mode (the grid kept on the GPU between the
calls on one host array).

* tests/helpers/compat_driver, a C++ program compiled against the header
  alone, deposits 2^26 + 2^24 particles on a 256^3 grid in read_fieldize's
  chunks of 2^26, adds 1 to the total mass per read, runs the in-place FFT
  and powerspectrum; its output is checked against the FP64 oracle (counts
  bit-exact, keffs within 1e-12, P within 1e-8) in both modes and its
  timings are printed.
* Resident-mode semantics, in process: the host array lacks the deposits
  until the in-place FFT or a flush, which give eager mode's results; a
  call on another array flushes first; turning the mode off flushes.
"""
import ctypes
import json
import os
import subprocess

import numpy as np
import pytest

import genpk_amd
import oracle
from genpk_amd._lib import check, lib

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "tests", "helpers", "compat_driver")


def announce(*a):
    print("CASE", *a, flush=True)


@pytest.mark.timeout(600)
def test_compat_driver_chunks_vs_oracle(gpu, tmp_path):
    announce("compat driver d=256, 2^26 + 2^24 particles in chunks of 2^26, eager and resident")
    assert os.path.exists(DRIVER), "build the test helpers first (__graft_entry__.build())"
    d, box, chunk = 256, 250.0, 1 << 26
    n = chunk + (1 << 24)
    rng = np.random.default_rng(77)
    pos = (rng.random((n, 3), dtype=np.float32) * np.float32(box)).astype(np.float32)
    pos[: n // 4] = np.mod(pos[: n // 4] * np.float32(0.1) + np.float32(40.0), np.float32(box))  # a dense region
    pfile = tmp_path / "pos.f32"
    pos.tofile(pfile)
    M = float(n) + 1.0  # read_fieldize.cpp:94
    po, co, ko = oracle.pk_from_particles(box, d, pos, None, 1.0, d, M)
    del pos
    occ = co > 0
    for mode in ("eager", "resident"):
        out = tmp_path / f"pk_{mode}.bin"
        args = [DRIVER, str(pfile), str(n), str(d), str(box), str(chunk), str(out)] + (["resident"] if mode == "resident" else [])
        r = subprocess.run(args, capture_output=True, text=True, timeout=300)
        assert r.returncode == 0, r.stderr
        t = json.loads(r.stdout.strip().splitlines()[-1])
        print(mode, t, flush=True)
        raw = open(out, "rb").read()
        P = np.frombuffer(raw, np.float64, d, 0)
        c = np.frombuffer(raw, np.int32, d, 8 * d)
        k = np.frombuffer(raw, np.float64, d, 12 * d)
        m = np.frombuffer(raw, np.float64, 1, 20 * d)[0]
        assert m == M
        assert np.array_equal(c, co), mode
        assert np.abs(k[occ] / ko[occ] - 1).max() <= 1e-12, mode
        assert np.abs(P[occ] / po[occ] - 1).max() <= 1e-8, mode


def test_resident_mode_semantics(gpu):
    announce("compat resident mode semantics d=32")
    d, box = 32, 10.0
    rng = np.random.default_rng(5)
    a1 = (rng.random((3000, 3)) * box).astype(np.float32)
    a2 = (rng.random((2000, 3)) * box).astype(np.float32)
    b1 = (rng.random((1000, 3)) * box).astype(np.float32)
    shape = (d, d, d + 2)

    def eager():
        check(lib().genpk_compat_resident(0))
        g = np.zeros(shape)
        genpk_amd.fieldize(box, d, g, a1, None, 1.0, 1)
        genpk_amd.fieldize(box, d, g, a2, None, 2.0, 1)
        dens = g.copy()
        F = genpk_amd.fft_r2c_3d_inplace(g)
        return dens, F.copy(), genpk_amd.powerspectrum(d, F, F, d, 7000.0, 7000.0)

    dens_e, F_e, pk_e = eager()
    try:
        check(lib().genpk_compat_resident(1))
        g = np.full(shape, 0.5)
        g[:, :, d:] = 0
        genpk_amd.fieldize(box, d, g, a1, None, 1.0, 1)
        assert np.all(g[:, :, :d] == 0.5)           # deposits kept on the GPU
        genpk_amd.fieldize(box, d, g, a2, None, 2.0, 1)
        check(lib().genpk_compat_flush())           # host array brought up to date
        assert np.allclose(g[:, :, :d] - 0.5, dens_e[:, :, :d], rtol=0, atol=1e-12)
        # a fresh grid through the driver's sequence gives eager mode's spectrum and power
        h = np.zeros(shape)
        genpk_amd.fieldize(box, d, h, a1, None, 1.0, 1)
        genpk_amd.fieldize(box, d, h, a2, None, 2.0, 1)
        F = genpk_amd.fft_r2c_3d_inplace(h)
        assert np.abs(F - F_e).max() <= 1e-12 * np.abs(F_e).max()
        pk = genpk_amd.powerspectrum(d, F, F, d, 7000.0, 7000.0)
        assert np.array_equal(pk[1], pk_e[1]) and np.allclose(pk[0], pk_e[0], rtol=1e-12, atol=0)
        # a call on another array flushes the resident one first
        u = np.zeros(shape)
        genpk_amd.fieldize(box, d, u, b1, None, 1.0, 1)
        w = np.zeros(shape)
        genpk_amd.fieldize(box, d, w, a1, None, 1.0, 1)
        ref = np.zeros(shape)
        oracle.fieldize(box, d, ref, b1, None, 1.0, 1)
        assert np.abs(u - ref).max() <= 1e-9 * np.abs(ref).max()
    finally:
        check(lib().genpk_compat_resident(0))      # flushes w
    ref = np.zeros(shape)
    oracle.fieldize(box, d, ref, a1, None, 1.0, 1)
    assert np.abs(w - ref).max() <= 1e-9 * np.abs(ref).max()
