"""Gadget-2 snapshot reading (CPU): the gen-pk C++ reader against the
independent Python reader and writer, on written snapshots (format I and
II, one and several files, MASS blocks) and on the reference's format-II
fixture test_g2_snap.{0,1} (a data file of the reference's tests, copied
under tests/golden/).  The fixture's known facts: types 0/1/4 hold
4039/4096/57 particles over two files, header mass 0.0406... for type 1
only, box 3000."""
import os
import subprocess

import numpy as np
import pytest

from genpk_amd import gadget

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURE = os.path.join(HERE, "golden", "test_g2_snap")


@pytest.fixture(scope="module")
def dump_tool(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("bin") / "gadget_dump")
    subprocess.run(["g++", "-O1", "-std=c++17", os.path.join(HERE, "helpers", "gadget_dump.cpp"), "-o", exe],
                   check=True)
    return exe


def cpp_read(exe, name, chunk):
    out = subprocess.run([exe, name, str(chunk)], capture_output=True, text=True, check=True).stdout.splitlines()
    head = out[0].split()
    info = dict(files=int(head[1]), box=float(head[3]), posbytes=int(head[9]))
    types, rows, i = [], [], 1
    while i < len(out):
        tok = out[i].split()
        n = int(tok[3])
        vals = np.array([[float.fromhex(v) for v in out[i + 1 + j].split()] for j in range(n)]).reshape(n, 4)
        types.append((n, float(tok[5])))
        rows.append(vals)
        i += 1 + n
    return info, types, rows


def check_same(exe, name, chunk):
    h, pos, mas = gadget.read_snapshot(name)
    info, types, rows = cpp_read(exe, name, chunk)
    assert info["box"] == h["box"] and info["posbytes"] == 12
    for t in range(gadget.N_TYPE):
        assert types[t][0] == pos[t].shape[0]
        if pos[t].shape[0]:
            assert np.array_equal(rows[t][:, :3].astype(np.float32), pos[t])
            assert np.array_equal(rows[t][:, 3].astype(np.float32), mas[t])
    return h


def test_fixture_header_and_blocks(dump_tool):
    h = check_same(dump_tool, FIXTURE, 1000)
    assert h["npart"] == [4039, 4096, 0, 0, 57, 0]
    assert h["num_files"] == 2 and h["box"] == 3000.0
    assert h["mass"][1] == pytest.approx(0.0406160778174934, rel=1e-15)
    assert h["mass"][0] == 0 and h["mass"][4] == 0
    # chunks that straddle the file boundary
    check_same(dump_tool, FIXTURE, 7)
    check_same(dump_tool, FIXTURE + ".0", 2000)


@pytest.mark.parametrize("fmt,nfiles", [(2, 1), (2, 3), (1, 1), (1, 2)])
def test_written_snapshots(dump_tool, tmp_path, fmt, nfiles):
    rng = np.random.default_rng(fmt * 10 + nfiles)
    box = 20.0
    P = [rng.random((101, 3)).astype(np.float32) * box, rng.random((64, 3)).astype(np.float32) * box, None, None,
         rng.random((13, 3)).astype(np.float32) * box, None]
    M = [(rng.random(101) + 0.5).astype(np.float32), 0.25, 0.0, 0.0, (rng.random(13) + 0.1).astype(np.float32), 0.0]
    name = str(tmp_path / "snap")
    gadget.write_snapshot(name, P, M, box, nfiles=nfiles, fmt=fmt)
    for chunk in (5, 64, 1000):
        check_same(dump_tool, name, chunk)


def test_reader_errors(dump_tool, tmp_path):
    bad = tmp_path / "bad"
    bad.write_bytes(b"\x01\x02\x03\x04" * 10)
    r = subprocess.run([dump_tool, str(bad), "10"], capture_output=True, text=True)
    assert r.returncode == 1 and "Gadget" in r.stderr
    r = subprocess.run([dump_tool, str(tmp_path / "missing"), "10"], capture_output=True, text=True)
    assert r.returncode == 1 and "cannot open" in r.stderr
