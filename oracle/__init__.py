"""CPU oracle for the GenPK power-spectrum path -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s
``cpu_baseline`` leg may import this module, and only as the checker (or as
the timed CPU baseline).  The GPU product path never imports it.

The numerics live in ``genpk_oracle.c`` (built into ``liboracle.so`` by the
Makefile beside it).  Each function there cites the reference file:line it
restates:

* ``fieldize``        -- fieldize.cpp:46-114
* ``invwindow``       -- fieldize.cpp:117-133
* ``powerspectrum``   -- powerspectrum.c:35-110
* ``fft_r2c_3d``      -- FP64 DFT standing in for the in-place FFTW r2c the
  reference driver runs between the two (gen-pk.cpp:193,233)

Pins (tests/test_oracle.py): the known answers of test.cpp:31-86, facts of
the reference compiled from its sources (the bins of the Nyquist corner mode,
127/255/383/511/1022 at d = 128/256/384/512/1024, and the occupied-bin
counts), brute-force DFT and numpy's pocketfft, closed forms and
invariants.  ``bin_of_q`` shares the per-mode log formula of
``powerspectrum`` and is pinned by the same facts.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None


def build():
    """Compile liboracle.so with the Makefile beside this file."""
    import subprocess
    subprocess.run(["make", "-s", "-C", _HERE, "-B"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        dp = ctypes.POINTER(ctypes.c_double)
        fp = ctypes.POINTER(ctypes.c_float)
        ip = ctypes.POINTER(ctypes.c_int)
        L.orc_fieldize.argtypes = [ctypes.c_double, ctypes.c_int, dp, ctypes.c_int64, fp, fp,
                                   ctypes.c_double, ctypes.c_int]
        L.orc_fieldize.restype = ctypes.c_int
        L.orc_onedinvwindow.argtypes = [ctypes.c_int64, ctypes.c_int64]
        L.orc_onedinvwindow.restype = ctypes.c_double
        L.orc_invwindow.argtypes = [ctypes.c_int64] * 4
        L.orc_invwindow.restype = ctypes.c_double
        L.orc_fft_r2c_3d.argtypes = [ctypes.c_int, dp]
        L.orc_fft_r2c_3d.restype = ctypes.c_int
        L.orc_fft_c2r_3d.argtypes = [ctypes.c_int, dp]
        L.orc_fft_c2r_3d.restype = ctypes.c_int
        L.orc_powerspectrum.argtypes = [ctypes.c_int64, dp, dp, ctypes.c_int, dp, ip, dp,
                                        ctypes.c_double, ctypes.c_double]
        L.orc_powerspectrum.restype = ctypes.c_int
        L.orc_bin_of_q.argtypes = [ctypes.c_int64, ctypes.c_int, ctypes.c_int64]
        L.orc_bin_of_q.restype = ctypes.c_int64
        L.orc_num_threads.argtypes = []
        L.orc_num_threads.restype = ctypes.c_int
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _fp(a):
    return None if a is None else a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def padded_len(dims, extra=1):
    """Length of the last axis of the padded grid: 2*(dims/2+extra) (fieldize.cpp:48)."""
    return 2 * (dims // 2 + extra)


def new_field(dims, extra=1):
    """Zeroed FP64 padded grid [d][d][2(d/2+extra)] (gen-pk.cpp:176-181,208)."""
    return np.zeros((dims, dims, padded_len(dims, extra)), dtype=np.float64)


def fieldize(boxsize, dims, out, positions, masses=None, mass=1.0, extra=1):
    """CIC deposit accumulating into ``out`` (fieldize.cpp:46-114)."""
    pos = np.ascontiguousarray(positions, dtype=np.float32).reshape(-1, 3)
    assert out.dtype == np.float64 and out.flags.c_contiguous
    assert out.size >= dims * dims * padded_len(dims, extra)
    m = None if masses is None else np.ascontiguousarray(masses, dtype=np.float32)
    if m is not None:
        assert m.shape[0] >= pos.shape[0]
    return lib().orc_fieldize(float(boxsize), int(dims), _dp(out), pos.shape[0], _fp(pos), _fp(m),
                              float(mass), int(extra))


def onedinvwindow(k, n):
    return lib().orc_onedinvwindow(int(k), int(n))


def invwindow(kx, ky, kz, n):
    return lib().orc_invwindow(int(kx), int(ky), int(kz), int(n))


def fft_r2c_3d(field):
    """In-place FP64 r2c on a padded [d][d][d+2] grid; returns the complex view [d][d][d/2+1]."""
    d = field.shape[0]
    assert field.dtype == np.float64 and field.flags.c_contiguous and field.shape == (d, d, 2 * (d // 2 + 1))
    rc = lib().orc_fft_r2c_3d(d, _dp(field))
    assert rc == 0
    return field.view(np.complex128)


def fft_c2r_3d(field):
    """In-place unnormalised FP64 c2r (inverse of fft_r2c_3d up to a factor d^3)."""
    d = field.shape[0]
    assert field.dtype == np.float64 and field.flags.c_contiguous and field.shape == (d, d, 2 * (d // 2 + 1))
    rc = lib().orc_fft_c2r_3d(d, _dp(field))
    assert rc == 0
    return field


def powerspectrum(dims, f1, f2, nrbins, m1, m2):
    """Binned P(k), mode count and keff (powerspectrum.c:35-110).

    f1, f2: complex128 arrays [d][d][d/2+1] (may be the same object).
    Returns (power, count, keffs).
    """
    a = np.ascontiguousarray(f1).view(np.float64)
    b = np.ascontiguousarray(f2).view(np.float64)
    assert a.size == 2 * dims * dims * (dims // 2 + 1) and b.size == a.size
    power = np.zeros(nrbins, np.float64)
    count = np.zeros(nrbins, np.int32)
    keffs = np.zeros(nrbins, np.float64)
    rc = lib().orc_powerspectrum(int(dims), _dp(a), _dp(b), int(nrbins), _dp(power),
                                 count.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), _dp(keffs),
                                 float(m1), float(m2))
    assert rc == 0, "mode fell outside [0, nrbins)"
    return power, count, keffs


def bin_of_q(dims, nrbins, q):
    return lib().orc_bin_of_q(int(dims), int(nrbins), int(q))


def num_threads():
    return lib().orc_num_threads()


def pk_from_particles(boxsize, dims, positions, masses=None, mass=1.0, nrbins=None, total_mass=None):
    """fieldize -> r2c -> powerspectrum in FP64, the reference's forward path
    (gen-pk.cpp:208-234).  total_mass defaults to the deposited mass."""
    nrbins = dims if nrbins is None else nrbins
    field = new_field(dims)
    fieldize(boxsize, dims, field, positions, masses, mass, 1)
    if total_mass is None:
        n = np.asarray(positions).reshape(-1, 3).shape[0]
        total_mass = float(np.sum(masses, dtype=np.float64)) if masses is not None else mass * n
    F = fft_r2c_3d(field)
    return powerspectrum(dims, F, F, nrbins, total_mass, total_mass)
