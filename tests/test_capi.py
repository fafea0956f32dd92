"""CPU-side checks of the C-ABI library: it loads, exports every symbol the
public headers declare, and its host-only utilities match the reference's
known answers (test.cpp:52-57, 102-113; utils.cpp:7-21).  No GPU calls."""
import os
import re
import subprocess

import numpy as np
import pytest

import genpk_amd
from genpk_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADERS = [os.path.join(ROOT, "include", h) for h in ("genpk_amd.h", "genpk_compat.h")]


def declared_functions():
    names = set()
    for h in HEADERS:
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b(\w+)\s*\(", text, flags=re.M):
            name = m.group(1)
            if name not in ("if", "defined"):
                names.add(name)
    return names


def exported_symbols():
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True, text=True, check=True)
    return {line.split()[-1] for line in out.stdout.splitlines() if line.strip()}


def test_library_loads():
    assert genpk_amd.lib() is not None


def test_every_declared_symbol_exported():
    decl = declared_functions()
    assert {"genpk_plan_create", "genpk_deposit", "genpk_power", "powerspectrum", "invwindow", "fieldize"} <= decl
    syms = exported_symbols()
    for name in decl:
        if name == "fieldize":
            # C++ linkage, as gen-pk.h:93 declares it
            assert "_Z8fieldizediPdlPfS0_di" in syms
        else:
            assert name in syms, name
    assert set(_lib.SIGNATURES) <= syms


def near(x, y):
    return abs(x - y) <= max(abs(x), abs(y)) / 1e5


def test_kat_invwindow():
    # test.cpp:52-57
    assert near(genpk_amd.invwindow(0, 3, 4, 5), 71.8177719)
    assert near(genpk_amd.invwindow(4, 4, 4, 5), 6111.20801)
    assert genpk_amd.invwindow(1, 1, 1, 0) == 0


def test_invwindow_matches_oracle():
    import oracle
    for args in [(1, 2, 3, 16), (-7, 0, 8, 16), (100, 200, 300, 1024), (512, 512, 512, 1024), (0, 0, 0, 8)]:
        assert genpk_amd.invwindow(*args) == oracle.invwindow(*args)


def test_kat_type_str_and_nexttwo():
    # test.cpp:102-113
    assert genpk_amd.type_str(0) == "by"
    assert genpk_amd.type_str(1) == "DM"
    assert genpk_amd.type_str(2) == "nu" and genpk_amd.type_str(4) == "st" and genpk_amd.type_str(3) == "xx"
    assert genpk_amd.nexttwo(12) == 16
    assert genpk_amd.nexttwo(8) == 8
    assert genpk_amd.nexttwo(1) == 1


def test_print_pk_format(tmp_path):
    k = np.array([0.5, 1.25, 2.0, 3.0])
    p = np.array([1.0e-3, 0.0, 123456.789, 7.0])
    c = np.array([6, 0, 12, 1], dtype=np.int32)
    f = tmp_path / "PK-DM-test"
    genpk_amd.print_pk(str(f), k, p, c)
    expect = "".join("%e\t%e\t%d\n" % (k[i], p[i], c[i]) for i in range(4) if c[i])
    assert f.read_text() == expect


def test_radix_orderings():
    """The radix orderings the FFT variants use (host only)."""
    import genpk_amd
    assert genpk_amd.radix_plan(1024, 0) == [8, 8, 8, 2]
    assert genpk_amd.radix_plan(1024, 1) == [2, 8, 8, 8]
    assert genpk_amd.radix_plan(1024, 2) == [4, 4, 4, 4, 4]
    assert genpk_amd.radix_plan(384, 1) == [2, 3, 8, 8]
    assert genpk_amd.radix_plan(1024, 3) == [4, 8, 8, 4]
    assert genpk_amd.radix_plan(512, 3) == [4, 8, 8, 2]
    for n in (2, 6, 10, 48, 100, 192, 240, 384, 512, 1000):
        for v in range(4):
            r = genpk_amd.radix_plan(n, v)
            assert int(np.prod(r)) == n and set(r) <= {2, 3, 4, 5, 8}
            if v == 1:
                assert sorted(r) == sorted(genpk_amd.radix_plan(n, 0))
            if v == 3 and n % 4 == 0:
                assert r[0] == 4
    with pytest.raises(genpk_amd.GenPKError):
        genpk_amd.radix_plan(14, 0)


def test_fused_x_orderings():
    """The orderings the FP32 fused x pass takes (variant -1): a first radix 8
    (the stage run from registers), a last radix 2 (the stage run in the
    binning) and an even number of stages (the result back in the first
    buffer), the fewest stages for powers of two."""
    import genpk_amd
    known = {16: [8, 2], 64: [8, 2, 2, 2], 96: [8, 2, 3, 2], 128: [8, 4, 2, 2], 256: [8, 4, 4, 2],
             384: [8, 8, 3, 2], 512: [8, 8, 4, 2], 1024: [8, 8, 8, 2], 2048: [8, 4, 4, 4, 2, 2]}
    for n, r in known.items():
        assert genpk_amd.radix_plan(n, -1) == r
    made = 0
    for n in range(16, 8193, 16):
        try:
            r = genpk_amd.radix_plan(n, -1)
        except genpk_amd.GenPKError:
            continue
        made += 1
        assert int(np.prod(r)) == n and set(r) <= {2, 3, 4, 5, 8}
        assert r[0] == 8 and r[-1] == 2 and len(r) % 2 == 0
        if n & (n - 1) == 0:
            # the fewest stages of radix <= 8 between a first 8 and a last 2, in an even count
            assert len(r) == 2 + 2 * -(-(n.bit_length() - 5) // 6)
    assert made >= 50  # 53 of the 512 multiples of 16 up to 8192
    # no such ordering: 32 = 8 4, 48 = 16 * 3, 80 = 16 * 5, 480 = 16 * 2 * 3 * 5
    # (an odd number of middle radices with no 8 or 4 to split), 40 (not a
    # multiple of 16), 112 (a factor 7)
    for n in (32, 48, 80, 480, 40, 112):
        with pytest.raises(genpk_amd.GenPKError):
            genpk_amd.radix_plan(n, -1)
