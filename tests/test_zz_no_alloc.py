"""genpk_power_noalloc, called through the C ABI: the genpk_power variant
whose per-workgroup partial bins live in the library's fixed device scratch
(no device allocation).  Checked against genpk_power on the same deposits:
mode counts and keffs identical, P within 1e-12 (the FP64 LDS bin sums run
in a varying order), on two calls in a row, the second on a smaller grid
(a two-plan cross spectrum on 128^3, then an auto spectrum on 64^3), and
its errors: bin tables not built, partial bins beyond the scratch (the
plans are left untouched by a rejected call)."""
import ctypes

import numpy as np
import pytest

import genpk_amd
from genpk_amd._lib import GenPKError, check, lib
from genpk_amd import synth

pytestmark = pytest.mark.gpu


def _fn():
    f = lib().genpk_power_noalloc
    f.restype = ctypes.c_int
    f.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int] + [ctypes.c_void_p] * 5 + [ctypes.c_double] * 2
    return f


def _outs(torch, nrbins, cross):
    f64 = dict(dtype=torch.float64, device="cuda")
    pw = [torch.full((nrbins,), -1.0, **f64) for _ in range(3 if cross else 1)]
    return pw, torch.full((nrbins,), -1, dtype=torch.int32, device="cuda"), torch.full((nrbins,), -1.0, **f64)


def _ptr(t):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def _case(torch, d, prec, cross, seed):
    box = 100.0
    a = genpk_amd.Plan(d, prec)
    b = genpk_amd.Plan(d, prec) if cross else None
    pa = synth.gaussian_lattice_torch(d // 2, box, d, 0.4, seed=seed)
    pb = synth.gaussian_lattice_torch(d // 2, box, d, 0.7, seed=seed + 1) if cross else None
    n = pa.shape[0]

    def deposit():
        a.clear()
        a.deposit(pa, box, mass=1.0)
        if cross:
            b.clear()
            b.deposit(pb, box, mass=0.5)

    deposit()
    ref = a.power(d, other=b, mass=float(n), other_mass=0.5 * n)  # builds the bin tables
    deposit()
    pw, cnt, ke = _outs(torch, d, cross)
    a._sync_stream()
    if cross:
        b._sync_stream()
    rc = _fn()(a._h, b._h if cross else None, d, _ptr(pw[0]), _ptr(pw[1]) if cross else None,
               _ptr(pw[2]) if cross else None, _ptr(cnt), _ptr(ke), float(n), 0.5 * n)
    check(rc)
    torch.cuda.synchronize()
    (rp, rc_, rk) = ref
    rp = list(rp) if cross else [rp]
    assert torch.equal(cnt, rc_)
    assert torch.equal(ke, rk)
    occ = rc_ > 0
    for g, r in zip(pw, rp):
        rel = ((g[occ] - r[occ]).abs() / r[occ].abs().clamp(min=1e-300)).max().item()
        assert rel <= 1e-12, rel
    return a, b


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_power_noalloc_matches_power(gpu, prec):
    import torch
    print("CASE power_noalloc", prec, "cross 128^3 then auto 64^3", flush=True)
    _case(torch, 128, prec, True, seed=3)
    _case(torch, 64, prec, False, seed=5)


def test_power_noalloc_errors(gpu):
    import torch
    print("CASE power_noalloc errors", flush=True)
    d, box = 256, 100.0
    a = genpk_amd.Plan(d, "f64")
    b = genpk_amd.Plan(d, "f64")
    pos = synth.gaussian_lattice_torch(32, box, d, 0.4, seed=9)
    for p in (a, b):
        p.deposit(pos, box)
    nb = 4096
    pw, cnt, ke = _outs(torch, nb, True)
    args = (_ptr(pw[0]), _ptr(pw[1]), _ptr(pw[2]), _ptr(cnt), _ptr(ke), 1.0, 1.0)
    # no bin tables of this nrbins on the plan yet
    rc = _fn()(a._h, b._h, nb, *args)
    assert rc != 0 and b"not built" in lib().genpk_last_error()
    # 4096 bins x 3 spectra x 256 resident workgroups (one per CU: the bins
    # fill the LDS) exceed the 2 Mi-double scratch
    (ref, _, _), _, _ = a.power(nb, other=b)
    for p in (a, b):
        p.clear()
        p.deposit(pos, box)
    a._sync_stream()
    rc = _fn()(a._h, b._h, nb, *args)
    assert rc != 0 and b"exceed the scratch" in lib().genpk_last_error()
    with pytest.raises(GenPKError):
        check(rc)
    # the rejected call left the deposits in place
    (again, _, _), c, _ = a.power(nb, other=b)
    occ = c > 0
    assert ((again[occ] - ref[occ]).abs() / ref[occ].abs()).max().item() <= 1e-12
