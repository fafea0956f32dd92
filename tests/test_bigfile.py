"""MP-Gadget bigfile snapshots as the gen-pk tool reads them
(read_fieldize_bigfile.cpp:7-125, through genpk_amd/cli/bigfile.h).

Parity with the bigfile library is unpinned: the library is an empty
submodule of the reference and no bigfile data ships with it.  What is
pinned: the reader returns exactly what tests/bigfile_util.py writes in the
restated layout (multi-file blocks, big- and little-endian, float64 and
float32, integer attributes), with the reference's conversions (positions
read as float64 and rounded to float32, masses read as float32, redshift =
1/Time - 1), and it fails on missing or short blocks.
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

import bigfile_util as bfu

HERE = os.path.dirname(os.path.abspath(__file__))
CLI = os.path.join(os.path.dirname(HERE), "genpk_amd", "cli")


@pytest.fixture(scope="module")
def dump_tool(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("bin") / "bigfile_dump")
    subprocess.run(["g++", "-O1", "-std=c++17", "-I", CLI, os.path.join(HERE, "helpers", "bigfile_dump.cpp"), "-o",
                    exe], check=True)
    return exe


def dump(exe, snap, out):
    r = subprocess.run([exe, snap, out], capture_output=True, text=True)
    if r.returncode:
        return None, r.stderr
    info = open(out + ".info").read().splitlines()
    pos = [np.fromfile(f"{out}.pos{t}", dtype=np.float32).reshape(-1, 3) for t in range(6)]
    mas = [np.fromfile(f"{out}.mass{t}", dtype=np.float32) for t in range(6)]
    return (info, pos, mas), ""


def particles(seed, box=25.0):
    rng = np.random.default_rng(seed)
    pos = [rng.random((1500, 3)) * box, rng.random((2000, 3)) * box, None, None, rng.random((37, 3)) * box, None]
    mas = [(rng.random(1500) * 0.3 + 0.05).astype(np.float32), None, None, None, (rng.random(37) + 0.5).astype(np.float32), None]
    return pos, mas, [0.0, 0.25, 0.0, 0.0, 0.0, 0.0]


@pytest.mark.parametrize("pos_dtype,mass_dtype,nfiles,npart_dtype",
                         [("<f8", "<f4", 1, "<u8"), ("<f8", "<f4", 3, "<i8"), (">f8", ">f4", 2, ">u8"),
                          ("<f4", "<f8", 4, "<i4")])
def test_roundtrip(dump_tool, tmp_path, pos_dtype, mass_dtype, nfiles, npart_dtype):
    pos, mas, table = particles(nfiles)
    snap = str(tmp_path / "snap")
    bfu.write_snapshot(snap, pos, mas, table, 25.0, time=0.25, pos_dtype=pos_dtype, mass_dtype=mass_dtype,
                       nfiles=nfiles, npart_dtype=npart_dtype)
    got, err = dump(dump_tool, snap, str(tmp_path / "d"))
    assert got is not None, err
    info, rpos, rmas = got
    assert "box 25" in info and "time 0.25" in info and "redshift 3" in info
    for t in range(6):
        if pos[t] is None:
            assert rpos[t].size == 0
            continue
        # the reference reads float64 and rounds to float32
        want = np.asarray(pos[t], dtype=np.dtype(pos_dtype)).astype(np.float64).astype(np.float32)
        assert np.array_equal(rpos[t], want)
        if table[t] == 0:
            assert np.array_equal(rmas[t], np.asarray(mas[t], dtype=np.dtype(mass_dtype)).astype(np.float32))
        line = [l for l in info if l.startswith(f"type {t} ")][0].split()
        assert int(line[2]) == len(pos[t]) and float(line[3]) == table[t]


def test_errors(dump_tool, tmp_path):
    pos, mas, table = particles(7)
    snap = str(tmp_path / "snap")
    bfu.write_snapshot(snap, pos, mas, table, 25.0)
    # no Header block: not a bigfile
    got, err = dump(dump_tool, str(tmp_path), str(tmp_path / "d0"))
    assert got is None and "not a bigfile" in err
    # a missing Mass block for a type with MassTable 0
    shutil.rmtree(os.path.join(snap, "4", "Mass"))
    got, err = dump(dump_tool, snap, str(tmp_path / "d1"))
    assert got is None and "4/Mass" in err
    # a Position block shorter than TotNumPart
    bfu.write_snapshot(snap, pos, mas, table, 25.0)
    hdr = os.path.join(snap, "1", "Position", "header")
    text = open(hdr).read().replace("NFILE: 2", "NFILE: 1")
    open(hdr, "w").write(text)
    got, err = dump(dump_tool, snap, str(tmp_path / "d2"))
    assert got is None and "fewer rows" in err
