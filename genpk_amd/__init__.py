"""genpk_amd: GenPK's power-spectrum path on AMD MI355X (gfx950).

CIC mass assignment (the reference's ``fieldize``), in-place 3-D r2c FFT and
CIC-deconvolved log-binned P(k) (``powerspectrum``), implemented as hand
written HIP kernels in ``libgenpk_amd.so`` and called through its C ABI.

Two interfaces:

* ``Plan`` -- device-resident, stream-ordered API over torch CUDA(HIP)
  tensors: deposit, r2c, bin, fused power, adjoint.
* ``fieldize`` / ``invwindow`` / ``powerspectrum`` / ``fft_r2c_3d_inplace``
  -- numpy mirrors of the reference's host call surface (gen-pk.h:93-126),
  same argument meaning, accumulation and normalisation.
"""
import ctypes

import numpy as np

from ._lib import GenPKError, check, lib

FLOAT32 = 4
FLOAT64 = 8

__all__ = ["Plan", "GenPKError", "fieldize", "invwindow", "powerspectrum", "fft_r2c_3d_inplace", "nexttwo",
           "type_str", "print_pk", "radix_plan", "FLOAT32", "FLOAT64", "lib"]


def _precision(p):
    if p in (FLOAT32, "f32", "float32", "fp32"):
        return FLOAT32
    if p in (FLOAT64, "f64", "float64", "fp64"):
        return FLOAT64
    raise ValueError(f"unknown precision {p!r}")


def _ptr(t):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


class Plan:
    """A dims^3 power-spectrum plan on the current HIP device.

    The padded grid [d][d][2(d/2+1)] lives in a torch tensor owned by the
    plan (``self.grid``); ``density()`` and ``spectrum()`` are views of it.
    """

    def __init__(self, dims, precision="f64", device=None):
        import torch
        self.torch = torch
        self.dims = int(dims)
        self.prec = _precision(precision)
        self.generation = 0  # bumped by every call that changes the plan's content
        self._kept = None    # _particles_key of the kept tile order
        self.device = torch.device("cuda", torch.cuda.current_device() if device is None else device)
        self._h = ctypes.c_void_p()
        with torch.cuda.device(self.device):
            stream = torch.cuda.current_stream(self.device).cuda_stream
            check(lib().genpk_plan_create(ctypes.byref(self._h), self.dims, self.prec, ctypes.c_void_p(stream)))
            nbytes = ctypes.c_size_t()
            check(lib().genpk_plan_field_bytes(self._h, ctypes.byref(nbytes)))
            rdt = torch.float32 if self.prec == FLOAT32 else torch.float64
            self.hz = self.dims // 2 + 1
            self.rowlen = 2 * self.hz
            self.grid = torch.empty((self.dims, self.dims, self.rowlen), dtype=rdt, device=self.device)
            assert self.grid.numel() * self.grid.element_size() == nbytes.value
            check(lib().genpk_plan_attach_field(self._h, ctypes.c_void_p(self.grid.data_ptr())))

    # -- tuning options -------------------------------------------------
    def set_options(self, tiles=(16, 32, 32), fft_variant=None):
        """Deposit tile edges (powers of two, clamped to dims) and the FFT
        radix ordering (0: radix 8 first, 1: leftover radices first, 2: radix
        4 first, 3: one radix 4 first in the y and x passes; None: the
        precision's default, 3 for f64 and 0 for f32).  Drops the plan's
        deposits."""
        self._sync_stream()
        self._kept = None
        self.generation += 1
        if fft_variant is None:
            fft_variant = 3 if self.prec == FLOAT64 else 0
        tx, ty, tz = (int(t) for t in tiles)
        check(lib().genpk_plan_set_options(self._h, tx, ty, tz, int(fft_variant)))
        return self

    def options(self):
        """{"tiles": (tx, ty, tz) in effect, "fft_variant": v}."""
        v = [ctypes.c_int() for _ in range(4)]
        check(lib().genpk_plan_get_options(self._h, *[ctypes.byref(x) for x in v]))
        return {"tiles": (v[0].value, v[1].value, v[2].value), "fft_variant": v[3].value}

    # -- plumbing -------------------------------------------------------
    def _sync_stream(self):
        s = self.torch.cuda.current_stream(self.device).cuda_stream
        check(lib().genpk_plan_set_stream(self._h, ctypes.c_void_p(s)))

    def close(self):
        if self._h:
            lib().genpk_plan_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def real_dtype(self):
        return self.torch.float32 if self.prec == FLOAT32 else self.torch.float64

    def density(self):
        """View of the grid as the padded real density [d][d][2(d/2+1)]."""
        return self.grid

    def spectrum(self):
        """View of the grid as the complex half spectrum [d][d][d/2+1]."""
        return self.grid.view(self.torch.complex64 if self.prec == FLOAT32 else self.torch.complex128)

    def _check_particles(self, pos, masses):
        t = self.torch
        assert pos.is_cuda and pos.dtype == t.float32 and pos.is_contiguous() and pos.dim() == 2 and pos.shape[1] == 3, \
            "positions must be a contiguous float32 [n, 3] device tensor"
        if masses is not None:
            assert masses.is_cuda and masses.dtype == t.float32 and masses.is_contiguous() and \
                masses.shape == (pos.shape[0],), "masses must be a contiguous float32 [n] device tensor"

    # -- forward --------------------------------------------------------
    def clear(self):
        self.generation += 1
        check(lib().genpk_clear(self._h))

    @staticmethod
    def _particles_key(pos, boxsize, masses, mass):
        """Identity and in-place version of the particle tensors (the library
        compares pointers only)."""
        mk = None if masses is None else (masses.data_ptr(), masses._version)
        return (pos.data_ptr(), pos._version, pos.shape[0], float(boxsize), mk, float(mass))

    def deposit(self, pos, boxsize, masses=None, mass=1.0, retain=False):
        """Accumulating CIC deposit (fieldize.cpp:46-114).  retain: keep the
        particles' tile order for a following adjoint_gather of the same
        tensors, or for redeposit()."""
        self._check_particles(pos, masses)
        self._sync_stream()
        self.generation += 1
        fn = lib().genpk_deposit_retain if retain else lib().genpk_deposit
        self._kept = None
        check(fn(self._h, float(boxsize), pos.shape[0], _ptr(pos), _ptr(masses), float(mass)))
        if retain and pos.shape[0] <= (1 << 31):
            self._kept = self._particles_key(pos, boxsize, masses, mass)

    def keeps_order_of(self, pos, boxsize, masses=None, mass=1.0):
        """Whether the kept tile order is that of these tensors, unchanged
        since it was made."""
        return self._kept is not None and self._kept == self._particles_key(pos, boxsize, masses, mass)

    def redeposit(self):
        """Deposit again the particles of the kept tile order (from its
        sorted copy), skipping the partition."""
        assert self._kept is not None, "no kept tile order: deposit with retain=True first"
        self._sync_stream()
        self.generation += 1
        check(lib().genpk_redeposit(self._h))

    def drop_order(self):
        self._kept = None
        check(lib().genpk_drop_order(self._h))

    def export_density(self, out=None, rowlen=None, accumulate=False):
        """Real-space density as a padded grid [d][d][rowlen] of the plan's precision."""
        rowlen = self.rowlen if rowlen is None else int(rowlen)
        if out is None:
            out = self.torch.zeros((self.dims, self.dims, rowlen), dtype=self.real_dtype, device=self.device)
        assert out.is_contiguous() and out.dtype == self.real_dtype and out.shape == (self.dims, self.dims, rowlen)
        self._sync_stream()
        check(lib().genpk_export_density(self._h, _ptr(out), rowlen, 1 if accumulate else 0))
        return out

    def import_density(self, grid):
        assert grid.is_contiguous() and grid.dtype == self.real_dtype and grid.shape == self.grid.shape
        self._sync_stream()
        self.generation += 1
        check(lib().genpk_import_density(self._h, _ptr(grid)))

    def import_spectrum(self, spec):
        """Load a half spectrum (complex tensor [d][d][d/2+1] of the plan's precision)."""
        self._sync_stream()
        self.generation += 1
        check(lib().genpk_import_spectrum(self._h, _ptr(spec)))

    def r2c(self):
        self._sync_stream()
        self.generation += 1
        check(lib().genpk_r2c(self._h))
        return self.spectrum()

    def c2r(self):
        self._sync_stream()
        self.generation += 1
        check(lib().genpk_c2r(self._h))
        return self.density()

    def _outputs(self, nrbins, cross):
        t = self.torch
        f64 = dict(dtype=t.float64, device=self.device)
        pw = [t.zeros(nrbins, **f64) for _ in range(3 if cross else 1)]
        count = t.zeros(nrbins, dtype=t.int32, device=self.device)
        keffs = t.zeros(nrbins, **f64)
        return pw, count, keffs

    def bin(self, nrbins=None, other=None, mass=1.0, other_mass=None):
        """Bin stored spectra (powerspectrum.c:35-110).  Returns
        (power, count, keffs) for the auto spectrum, or
        ((P_aa, P_bb, P_ab), count, keffs) with ``other``."""
        nrbins = self.dims if nrbins is None else int(nrbins)
        cross = other is not None
        pw, count, keffs = self._outputs(nrbins, cross)
        self._sync_stream()
        mb = mass if other_mass is None else other_mass
        check(lib().genpk_bin(self._h, other._h if cross else None, nrbins, _ptr(pw[0]),
                              _ptr(pw[1]) if cross else None, _ptr(pw[2]) if cross else None,
                              _ptr(count), _ptr(keffs), float(mass), float(mb)))
        return (tuple(pw) if cross else pw[0]), count, keffs

    def power(self, nrbins=None, other=None, mass=1.0, other_mass=None):
        """Fused r2c + bin; the spectrum is not kept.  Same outputs as bin()."""
        nrbins = self.dims if nrbins is None else int(nrbins)
        cross = other is not None
        pw, count, keffs = self._outputs(nrbins, cross)
        self._sync_stream()
        self.generation += 1
        if cross:
            other._sync_stream()
            other.generation += 1
        mb = mass if other_mass is None else other_mass
        check(lib().genpk_power(self._h, other._h if cross else None, nrbins, _ptr(pw[0]),
                                _ptr(pw[1]) if cross else None, _ptr(pw[2]) if cross else None,
                                _ptr(count), _ptr(keffs), float(mass), float(mb)))
        return (tuple(pw) if cross else pw[0]), count, keffs

    # -- timing ---------------------------------------------------------
    def timing(self, enable=True):
        check(lib().genpk_timing_enable(self._h, 1 if enable else 0))

    def timing_read(self):
        """{stage: (total_ms, launches)} accumulated since the last read."""
        names = ctypes.create_string_buffer(4096)
        ms = (ctypes.c_double * 64)()
        cnt = (ctypes.c_int * 64)()
        k = ctypes.c_int()
        check(lib().genpk_timing_read(self._h, names, 4096, ms, cnt, 64, ctypes.byref(k)))
        toks = names.raw.split(b"\0")
        return {toks[i].decode(): (ms[i], cnt[i]) for i in range(k.value)}

    # -- adjoint --------------------------------------------------------
    def adjoint_density(self, lam, total_mass, nrbins=None):
        """dL/drho from dL/dP_b.  The plan holds the forward spectrum, the
        state left by power() (nothing deposited since), or the forward
        density (the transform is then redone)."""
        nrbins = self.dims if nrbins is None else int(nrbins)
        lam = lam.to(dtype=self.torch.float64, device=self.device).contiguous()
        assert lam.shape == (nrbins,)
        self._sync_stream()
        self.generation += 1
        check(lib().genpk_adjoint_density(self._h, nrbins, _ptr(lam), float(total_mass)))
        return self.density()

    def adjoint_gather(self, pos, boxsize, masses=None, mass=1.0, want_mass_grad=True):
        """dL/dpos (and dL/dm) of the particles from dL/drho in the plan.  A
        kept tile order of these tensors is reused only if they were not
        modified in place since it was made; the order used is kept."""
        self._check_particles(pos, masses)
        t = self.torch
        gpos = t.empty_like(pos)
        gm = t.empty(pos.shape[0], dtype=t.float32, device=self.device) if want_mass_grad else None
        self._sync_stream()
        if not self.keeps_order_of(pos, boxsize, masses, mass):
            self.drop_order()
        check(lib().genpk_adjoint_gather(self._h, float(boxsize), pos.shape[0], _ptr(pos), _ptr(masses), float(mass),
                                         _ptr(gpos), _ptr(gm)))
        self._kept = self._particles_key(pos, boxsize, masses, mass) if pos.shape[0] <= (1 << 31) else None
        return gpos, gm


# ---------------------------------------------------------------------------
# Mirrors of the reference's host call surface (numpy, host memory)
# ---------------------------------------------------------------------------

def _np_dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def fieldize(boxsize, dims, out, positions, masses=None, mass=1.0, extra=1):
    """fieldize(boxsize, dims, out, segment_particles, positions, masses, mass,
    extra) of fieldize.cpp:46: CIC deposit accumulating into the FP64 host
    grid ``out`` ([dims][dims][2(dims/2+extra)], C order).  Returns 0."""
    assert out.dtype == np.float64 and out.flags.c_contiguous
    assert out.size >= dims * dims * 2 * (dims // 2 + extra)
    pos = np.ascontiguousarray(positions, dtype=np.float32).reshape(-1, 3)
    m = None if masses is None else np.ascontiguousarray(masses, dtype=np.float32)
    fp = ctypes.POINTER(ctypes.c_float)
    rc = lib()._Z8fieldizediPdlPfS0_di(float(boxsize), int(dims), _np_dp(out), pos.shape[0],
                                       pos.ctypes.data_as(fp), None if m is None else m.ctypes.data_as(fp),
                                       float(mass), int(extra))
    return check(rc)


def invwindow(kx, ky, kz, n):
    """fieldize.cpp:125-133."""
    return lib().invwindow(int(kx), int(ky), int(kz), int(n))


def powerspectrum(dims, outfield, outfield2, nrbins, total_mass, total_mass2):
    """powerspectrum.c:35-110 on host complex128 half spectra [d][d][d/2+1]
    (outfield2 may be outfield).  Returns (power, count, keffs)."""
    a = np.ascontiguousarray(outfield, dtype=np.complex128)
    b = a if outfield2 is outfield else np.ascontiguousarray(outfield2, dtype=np.complex128)
    assert a.size == dims * dims * (dims // 2 + 1) and b.size == a.size
    power = np.zeros(nrbins, np.float64)
    count = np.zeros(nrbins, np.int32)
    keffs = np.zeros(nrbins, np.float64)
    av, bv = a.view(np.float64), b.view(np.float64)
    check(lib().powerspectrum(int(dims), _np_dp(av), _np_dp(bv), int(nrbins), _np_dp(power),
                              count.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), _np_dp(keffs),
                              float(total_mass), float(total_mass2)))
    return power, count, keffs


def fft_r2c_3d_inplace(field):
    """In-place FP64 r2c of a host padded grid [d][d][2(d/2+1)] (the FFTW step
    of gen-pk.cpp:193,233).  Returns the complex view [d][d][d/2+1]."""
    d = field.shape[0]
    assert field.dtype == np.float64 and field.flags.c_contiguous and field.shape == (d, d, 2 * (d // 2 + 1))
    check(lib().genpk_fft_r2c_3d_inplace(d, _np_dp(field)))
    return field.view(np.complex128)


def radix_plan(n, fft_variant=0):
    """Radices, in stage order, of the library's length-n transform under an
    FFT variant (host only); variant -1: the radix-8-first, radix-2-last
    ordering of an even number of stages that the FP32 fused x pass takes
    where the plan's own ordering is not of that form."""
    buf = (ctypes.c_int * 32)()
    cnt = ctypes.c_int()
    check(lib().genpk_radix_plan(int(n), int(fft_variant), buf, 32, ctypes.byref(cnt)))
    return [buf[i] for i in range(cnt.value)]


def nexttwo(n):
    """utils.cpp:35-42."""
    return lib().genpk_nexttwo(int(n))


def type_str(t):
    """utils.cpp:57-72."""
    buf = ctypes.create_string_buffer(4)
    check(lib().genpk_type_str(int(t), buf, 4))
    return buf.value.decode()


def print_pk(filename, keffs, power, count):
    """utils.cpp:7-21."""
    k = np.ascontiguousarray(keffs, dtype=np.float64)
    p = np.ascontiguousarray(power, dtype=np.float64)
    c = np.ascontiguousarray(count, dtype=np.int32)
    check(lib().genpk_print_pk(str(filename).encode(), len(c), _np_dp(k), _np_dp(p),
                               c.ctypes.data_as(ctypes.POINTER(ctypes.c_int))))
