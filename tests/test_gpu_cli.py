"""gen-pk command line tool on the GPU against the CPU oracle.

Expected output: for each species the reference's read_fieldize + FFT +
powerspectrum (read_fieldize.cpp:18-97, gen-pk.cpp:209-237), restated with
the oracle on particles read by the independent Python reader: total mass =
sum of the float masses in double (per read chunk) + header mass x count,
plus 1 per read; output lines "%e\\t%e\\t%d" for bins with count > 0
(utils.cpp:7-21).  Bars: the same set of lines, counts equal, keff and P
within 1e-5 relative (the files hold 6 significant digits; FP64 plans agree
with the oracle to 1e-8).
"""
import os
import subprocess

import numpy as np
import pytest

import oracle
from genpk_amd import gadget

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GENPK = os.path.join(ROOT, "genpk_amd", "gen-pk")
FIXTURE = os.path.join(ROOT, "tests", "golden", "test_g2_snap")


def announce(*a):
    print("CASE", *a, flush=True)


def run(args, ok=True):
    r = subprocess.run([GENPK, *args], capture_output=True, text=True, timeout=120)
    if ok:
        assert r.returncode == 0, r.stderr
    return r


def read_pk(path):
    rows = [l.split("\t") for l in open(path).read().splitlines()]
    return np.array([float(r[0]) for r in rows]), np.array([float(r[1]) for r in rows]), \
        np.array([int(r[2]) for r in rows])


def species_field(box, d, pos, mas, hmass):
    """Oracle grid and total mass of one read_fieldize call."""
    f = oracle.new_field(d)
    n = pos.shape[0]
    M = 0.0
    if n:
        if hmass == 0:
            oracle.fieldize(box, d, f, pos, mas, 0.0, 1)
            M += float(np.sum(mas.astype(np.float64)))
        else:
            oracle.fieldize(box, d, f, pos, None, hmass, 1)
        M += hmass * n
    return f, M + 1


def expected(box, d, fields, cross=False):
    """fields: list of (grid, M); auto of the first or cross of the first two."""
    if not cross:
        f, M = fields[0]
        F = oracle.fft_r2c_3d(f)
        P, c, k = oracle.powerspectrum(d, F, F, d, M, M)
    else:
        (f1, M1), (f2, M2) = fields
        F1 = oracle.fft_r2c_3d(f1)
        F2 = oracle.fft_r2c_3d(f2)
        P, c, k = oracle.powerspectrum(d, F1, F2, d, M1, M2)
    occ = c > 0
    return k[occ], P[occ], c[occ]


def compare(path, exp, rtol=1e-5):
    k, P, c = read_pk(path)
    ke, Pe, ce = exp
    assert np.array_equal(c, ce)
    assert np.allclose(k, ke, rtol=rtol, atol=0)
    assert np.allclose(P, Pe, rtol=rtol, atol=0)


def grid_dims(npart):
    d = 0
    for n in npart:
        c = int(np.cbrt(float(n)))
        t = 1
        while t < c:
            t *= 2
        d = max(d, min(2 * (t if c else 0), 3072))
    return d


def test_cli_fixture_auto(gpu, tmp_path):
    """The reference's format-II fixture (two files, types 0/1/4)."""
    announce("cli_fixture_auto")
    r = run(["-i", FIXTURE, "-o", str(tmp_path)])
    assert "FFT grid dimension: 32" in r.stdout
    h, pos, mas = gadget.read_snapshot(FIXTURE)
    box, d = h["box"], 32
    assert grid_dims(h["npart"]) == d
    for t, name in [(0, "by"), (1, "DM"), (4, "st")]:
        f, M = species_field(box, d, pos[t], mas[t], h["mass"][t])
        assert f"total_mass in type {t} = " in r.stdout
        compare(tmp_path / f"PK-{name}-test_g2_snap", expected(box, d, [(f, M)]))
    assert sorted(os.listdir(tmp_path)) == ["PK-DM-test_g2_snap", "PK-by-test_g2_snap", "PK-st-test_g2_snap"]


def _synthetic(tmp_path, name, fmt, nfiles, seed, shift=0.0):
    rng = np.random.default_rng(seed)
    box = 50.0
    P = [(rng.random((3000, 3)) * box).astype(np.float32), None, None, None,
         (rng.random((400, 3)) * box).astype(np.float32), None]
    # dark matter: a perturbed 16^3 lattice
    g = (np.arange(16) + 0.5) * (box / 16)
    lat = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    P[1] = np.mod(lat + rng.normal(0, 0.8, lat.shape) + shift, box).astype(np.float32)
    M = [(rng.random(3000) * 0.2 + 0.1).astype(np.float32), 0.75, 0.0, 0.0,
         (rng.random(400) * 0.3 + 0.05).astype(np.float32), 0.0]
    stem = str(tmp_path / name)
    gadget.write_snapshot(stem, P, M, box, nfiles=nfiles, fmt=fmt, redshift=1.0)
    return stem


@pytest.mark.parametrize("fmt,nfiles", [(2, 3), (1, 1)])
def test_cli_auto_and_stars(gpu, tmp_path, fmt, nfiles):
    announce("cli_auto", f"fmt={fmt} nfiles={nfiles}")
    stem = _synthetic(tmp_path, "snap", fmt, nfiles, seed=7)
    out = tmp_path / "out"
    out.mkdir()
    run(["-i", stem, "-o", str(out)])
    h, pos, mas = gadget.read_snapshot(stem)
    box = h["box"]
    d = grid_dims(h["npart"])
    for t, name in [(0, "by"), (1, "DM"), (4, "st")]:
        compare(out / f"PK-{name}-snap", expected(box, d, [species_field(box, d, pos[t], mas[t], h["mass"][t])]))
    # -s 1: the stars are read into the baryon field (a second read: +1 again)
    out2 = tmp_path / "out2"
    out2.mkdir()
    run(["-i", stem, "-o", str(out2), "-s", "1"])
    f0, M0 = species_field(box, d, pos[0], mas[0], h["mass"][0])
    f4, M4 = species_field(box, d, pos[4], mas[4], h["mass"][4])
    compare(out2 / "PK-by-snap", expected(box, d, [(f0 + f4, M0 + M4)]))


def test_cli_cross_modes(gpu, tmp_path):
    announce("cli_cross")
    stem = _synthetic(tmp_path, "snap", 2, 2, seed=8)
    other = _synthetic(tmp_path, "other", 2, 1, seed=9, shift=0.3)
    h, pos, mas = gadget.read_snapshot(stem)
    h2, pos2, mas2 = gadget.read_snapshot(other)
    box = h["box"]
    d = grid_dims(h["npart"])
    # -c 4: dark matter x stars inside one snapshot
    out = tmp_path / "c"
    out.mkdir()
    run(["-i", stem, "-o", str(out), "-c", "4"])
    a = species_field(box, d, pos[1], mas[1], h["mass"][1])
    b = species_field(box, d, pos[4], mas[4], h["mass"][4])
    compare(out / "PK-DMxst-snap", expected(box, d, [a, b], cross=True))
    # -j: each type of one snapshot against the same type of the other
    out = tmp_path / "j"
    out.mkdir()
    run(["-i", stem, "-j", other, "-o", str(out)])
    for t, name in [(0, "by"), (1, "DM"), (4, "st")]:
        a = species_field(box, d, pos[t], mas[t], h["mass"][t])
        b = species_field(box, d, pos2[t], mas2[t], h2["mass"][t])
        compare(out / f"PX-{name}-snap", expected(box, d, [a, b], cross=True))
    # a type the snapshot lacks
    r = run(["-i", stem, "-o", str(out), "-c", "3"], ok=False)
    assert r.returncode == 1 and "Can't cross-correlate" in r.stderr


def test_cli_fp32_and_usage(gpu, tmp_path):
    announce("cli_fp32")
    stem = _synthetic(tmp_path, "snap", 2, 1, seed=10)
    out = tmp_path / "o"
    out.mkdir()
    run(["-i", stem, "-o", str(out), "-p", "32"])
    h, pos, mas = gadget.read_snapshot(stem)
    d = grid_dims(h["npart"])
    compare(out / "PK-DM-snap", expected(h["box"], d, [species_field(h["box"], d, pos[1], mas[1], h["mass"][1])]),
            rtol=1e-4)
    r = run(["-h"])
    assert "Usage: ./gen-pk" in r.stderr
    r = run(["-i", stem])        # no output directory: usage, status 0 (gen-pk.cpp:153-155)
    assert "Usage" in r.stderr
    r = run(["-i", str(tmp_path / "missing"), "-o", str(out)], ok=False)
    assert r.returncode == 1


# ---- Gadget-HDF5 sets (read_fieldize.cpp:99-340; gen-pk.cpp:117-134, 206-216,
# 271-288, 316-328): no +1 in the total mass; -j with an HDF5 second set
# crosses type 1 of both sets under every type's name.

import hdf5_util as h5u  # noqa: E402


def _hdf5_field(box, d, pos, mas, table_mass, nfiles):
    f = oracle.new_field(d)
    if pos is not None and pos.shape[0]:
        if table_mass == 0:
            oracle.fieldize(box, d, f, pos, mas, 0.0, 1)
        else:
            oracle.fieldize(box, d, f, pos, None, table_mass, 1)
    n = 0 if pos is None else pos.shape[0]
    return f, h5u.expected_total_mass(mas, table_mass, n, nfiles)


def _hdf5_particles(seed, shift=0.0):
    rng = np.random.default_rng(seed)
    box = 50.0
    g = (np.arange(16) + 0.5) * (box / 16)
    lat = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    pos = [(rng.random((2500, 3)) * box).astype(np.float32),
           np.mod(lat + rng.normal(0, 0.8, lat.shape) + shift, box).astype(np.float32), None, None,
           (rng.random((300, 3)) * box).astype(np.float32), None]
    mas = [(rng.random(2500) * 0.2 + 0.1).astype(np.float32), None, None, None,
           (rng.random(300) * 0.3 + 0.05).astype(np.float32), None]
    return box, pos, mas, [0.0, 0.75, 0.0, 0.0, 0.0, 0.0]


hdf5_only = pytest.mark.skipif(not h5u.available(), reason="built without the HDF5 library")


@hdf5_only
@pytest.mark.parametrize("nfiles", [3, 0])
def test_cli_hdf5_auto_and_stars(gpu, tmp_path, nfiles):
    announce("cli_hdf5_auto", f"nfiles={nfiles}")
    box, pos, mas, table = _hdf5_particles(21)
    stem = str(tmp_path / "snap")
    h5u.write_set(stem, pos, mas, table, box, nfiles)
    d = grid_dims([0 if p is None else p.shape[0] for p in pos])
    out = tmp_path / "out"
    out.mkdir()
    r = run(["-i", stem, "-o", str(out)])
    assert f"FFT grid dimension: {d}" in r.stdout
    nf = max(nfiles, 1)
    assert f"Found {nf} hdf5 files in snapshot" in r.stderr
    for t, name in [(0, "by"), (1, "DM"), (4, "st")]:
        f, M = _hdf5_field(box, d, pos[t], mas[t], table[t], nfiles)
        compare(out / f"PK-{name}-snap", expected(box, d, [(f, M)]))
    assert sorted(os.listdir(out)) == ["PK-DM-snap", "PK-by-snap", "PK-st-snap"]
    # -s 1: the stars join the baryons; the masses add, with no +1
    out2 = tmp_path / "out2"
    out2.mkdir()
    run(["-i", stem, "-o", str(out2), "-s", "1"])
    f0, M0 = _hdf5_field(box, d, pos[0], mas[0], table[0], nfiles)
    f4, M4 = _hdf5_field(box, d, pos[4], mas[4], table[4], nfiles)
    compare(out2 / "PK-by-snap", expected(box, d, [(f0 + f4, M0 + M4)]))


@hdf5_only
def test_cli_hdf5_cross_modes(gpu, tmp_path):
    announce("cli_hdf5_cross")
    box, pos, mas, table = _hdf5_particles(22)
    _, pos2, mas2, table2 = _hdf5_particles(23, shift=0.4)
    stem, other = str(tmp_path / "snap"), str(tmp_path / "other")
    h5u.write_set(stem, pos, mas, table, box, 2)
    h5u.write_set(other, pos2, mas2, table2, box, 3)
    d = grid_dims([0 if p is None else p.shape[0] for p in pos])
    # -c 4: dark matter x stars of one set
    out = tmp_path / "c"
    out.mkdir()
    run(["-i", stem, "-o", str(out), "-c", "4"])
    a = _hdf5_field(box, d, pos[1], mas[1], table[1], 2)
    b = _hdf5_field(box, d, pos[4], mas[4], table[4], 2)
    compare(out / "PK-DMxst-snap", expected(box, d, [a, b], cross=True))
    # -j with an HDF5 second set: type 1 of both sets, under every type's name
    out = tmp_path / "j"
    out.mkdir()
    run(["-i", stem, "-j", other, "-o", str(out)])
    a = _hdf5_field(box, d, pos[1], mas[1], table[1], 2)
    b = _hdf5_field(box, d, pos2[1], mas2[1], table2[1], 3)
    exp = expected(box, d, [a, b], cross=True)
    for name in ("by", "DM", "st"):
        compare(out / f"PX-{name}-snap", exp)
    # a set with a header that cannot be read: status 1
    h5u.write_set(str(tmp_path / "bad"), pos, mas, table, box, 0)
    with open(str(tmp_path / "bad.hdf5"), "r+b") as fh:
        fh.seek(0)
        fh.write(b"\0" * 8)  # no longer an HDF5 file: read as a (missing) Gadget set
    r = run(["-i", str(tmp_path / "bad"), "-o", str(out)], ok=False)
    assert r.returncode == 1


# ---- MP-Gadget bigfile snapshots (read_fieldize_bigfile.cpp:7-125; gen-pk.cpp:135-145,
# 217-221, 320-330): positions read as float64 and rounded to float32, no +1.

import bigfile_util as bfu  # noqa: E402


def _bigfile_particles(seed, shift=0.0):
    rng = np.random.default_rng(seed)
    box = 50.0
    g = (np.arange(16) + 0.5) * (box / 16)
    lat = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    pos = [rng.random((2500, 3)) * box, np.mod(lat + rng.normal(0, 0.8, lat.shape) + shift, box), None, None,
           rng.random((300, 3)) * box, None]
    mas = [(rng.random(2500) * 0.2 + 0.1).astype(np.float32), None, None, None,
           (rng.random(300) * 0.3 + 0.05).astype(np.float32), None]
    return box, pos, mas, [0.0, 0.75, 0.0, 0.0, 0.0, 0.0]


def _bigfile_field(box, d, pos, mas, table_mass):
    f = oracle.new_field(d)
    p32 = np.asarray(pos, dtype=np.float64).astype(np.float32)
    if table_mass == 0:
        oracle.fieldize(box, d, f, p32, mas, 0.0, 1)
    else:
        oracle.fieldize(box, d, f, p32, None, table_mass, 1)
    return f, bfu.expected_total_mass(mas, table_mass, p32.shape[0])


def test_cli_bigfile(gpu, tmp_path):
    announce("cli_bigfile")
    box, pos, mas, table = _bigfile_particles(31)
    snap = str(tmp_path / "PART_010")
    bfu.write_snapshot(snap, pos, mas, table, box, nfiles=3)
    d = grid_dims([0 if p is None else len(p) for p in pos])
    out = tmp_path / "out"
    out.mkdir()
    r = run(["-i", snap, "-o", str(out)])
    assert f"FFT grid dimension: {d}" in r.stdout and "Expansion factor" in r.stdout
    for t, name in [(0, "by"), (1, "DM"), (4, "st")]:
        compare(out / f"PK-{name}-PART_010", expected(box, d, [_bigfile_field(box, d, pos[t], mas[t], table[t])]))
    # -c 4 and -s 1
    out2 = tmp_path / "c"
    out2.mkdir()
    run(["-i", snap, "-o", str(out2), "-c", "4"])
    a = _bigfile_field(box, d, pos[1], mas[1], table[1])
    b = _bigfile_field(box, d, pos[4], mas[4], table[4])
    compare(out2 / "PK-DMxst-PART_010", expected(box, d, [a, b], cross=True))
    out3 = tmp_path / "s"
    out3.mkdir()
    run(["-i", snap, "-o", str(out3), "-s", "1"])
    f0, M0 = _bigfile_field(box, d, pos[0], mas[0], table[0])
    f4, M4 = _bigfile_field(box, d, pos[4], mas[4], table[4])
    compare(out3 / "PK-by-PART_010", expected(box, d, [(f0 + f4, M0 + M4)]))
    # a missing block ends the run with status 1
    import shutil
    shutil.rmtree(os.path.join(snap, "4", "Position"))
    r = run(["-i", snap, "-o", str(out3)], ok=False)
    assert r.returncode == 1 and "Failed to read from block" in r.stderr
