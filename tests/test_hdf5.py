"""Gadget-HDF5 snapshot reading of the gen-pk tool (read_fieldize.cpp:99-340),
through the reader it is built with (tests/helpers/hdf5_snap dump).

Pinned by construction: sets written from known arrays (one file, several
files, per-particle masses and header masses) must read back bit-exact, in
file order, with the total mass accumulated as the reference does (per file a
double sum of the float32 masses, then + MassTable x count, no +1), and the
file-set discovery must follow the reference's rules (name, name.0.hdf5,
name.hdf5; name.N.hdf5 while present).
"""
import os

import numpy as np
import pytest

import hdf5_util as h5u

pytestmark = pytest.mark.skipif(not h5u.available(), reason="built without the HDF5 library")


def _particles(seed, box=40.0):
    rng = np.random.default_rng(seed)
    pos = [(rng.random((3001, 3)) * box).astype(np.float32), (rng.random((4096, 3)) * box).astype(np.float32),
           None, None, (rng.random((17, 3)) * box).astype(np.float32), None]
    mas = [(rng.random(3001) * 0.3 + 0.05).astype(np.float32), None, None, None,
           (rng.random(17) + 0.5).astype(np.float32), None]
    table = [0.0, 0.625, 0.0, 0.0, 0.0, 0.0]
    return pos, mas, table


@pytest.mark.parametrize("nfiles", [0, 1, 3, 11])
def test_roundtrip(tmp_path, nfiles):
    pos, mas, table = _particles(nfiles + 1)
    stem = str(tmp_path / "snap")
    h5u.write_set(stem, pos, mas, table, 40.0, nfiles)
    info, rpos, rmas, totals = h5u.dump_set(stem, str(tmp_path / "dump"))
    nf = max(nfiles, 1)
    assert info[0] == f"files {nf}"
    names = [l.split(" ", 1)[1] for l in info if l.startswith("file ")]
    if nfiles == 0:
        assert names == [stem + ".hdf5"]
    else:
        assert names == [f"{stem}.{i}.hdf5" for i in range(nfiles)]
    assert f"box {40.0!r}" in info or any(l.startswith("box 40") for l in info)
    for t in range(6):
        if pos[t] is None:
            assert totals[t] == (0, 0.0)
            continue
        n = pos[t].shape[0]
        assert np.array_equal(rpos[t], pos[t])
        if table[t] == 0:
            assert np.array_equal(rmas[t], mas[t])
        assert totals[t][0] == n
        assert totals[t][1] == h5u.expected_total_mass(mas[t], table[t], n, nfiles)
        line = [l for l in info if l.startswith(f"total {t} ")][0].split()
        assert int(line[2]) == n and float(line[3]) == table[t]


def test_discovery(tmp_path):
    pos, mas, table = _particles(5)
    stem = str(tmp_path / "set")
    h5u.write_set(stem, pos, mas, table, 40.0, 3)
    # the first file named directly finds the whole set
    info, *_ = h5u.dump_set(stem + ".0.hdf5", str(tmp_path / "d1"))
    assert info[0] == "files 3"
    # a gap ends the set
    os.remove(stem + ".2.hdf5")
    info, *_ = h5u.dump_set(stem, str(tmp_path / "d2"))
    assert info[0] == "files 2"
    # a name that is neither HDF5 nor a set
    open(str(tmp_path / "plain"), "wb").write(b"not hdf5")
    info, *_ = h5u.dump_set(str(tmp_path / "plain"), str(tmp_path / "d3"))
    assert info[0] == "files 0"
    # name.hdf5 found from the stem, a set of one
    h5u.write_set(str(tmp_path / "one"), pos, mas, table, 40.0, 0)
    info, *_ = h5u.dump_set(str(tmp_path / "one"), str(tmp_path / "d4"))
    assert info[0] == "files 1" and info[1].endswith("one.hdf5")


def test_header_totals_and_empty_files(tmp_path):
    # more files than particles of a type: some files hold none of it
    pos, mas, table = _particles(9)
    stem = str(tmp_path / "sparse")
    h5u.write_set(stem, pos, mas, table, 40.0, 20)
    info, rpos, rmas, totals = h5u.dump_set(stem, str(tmp_path / "d"))
    assert np.array_equal(rpos[4], pos[4]) and np.array_equal(rmas[4], mas[4])
    assert totals[4][1] == h5u.expected_total_mass(mas[4], 0.0, 17, 20)
    assert [l for l in info if l.startswith("total 1 ")][0].split()[2] == "4096"
