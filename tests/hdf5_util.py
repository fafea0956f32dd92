"""Gadget-HDF5 snapshot sets for tests, written and read back with the
tests/helpers/hdf5_snap tool (built by tests/helpers/build.py, from __graft_entry__.build(), when
the HDF5 C library is present)."""
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TOOL = os.path.join(ROOT, "tests", "helpers", "hdf5_snap")


def available():
    return os.path.exists(TOOL)


def write_set(out, positions, masses, mass_table, box, nfiles, redshift=1.0):
    """positions[t]: float32 [n][3] or None; masses[t]: float32 [n] for types
    with mass_table[t] == 0.  nfiles = 0 writes the single file out.hdf5,
    else out.<i>.hdf5, each type split into contiguous ranges over the files."""
    inp = out + "_in"
    np.asarray(mass_table, dtype=np.float64).tofile(inp + ".table")
    for t in range(6):
        if positions[t] is None:
            continue
        np.asarray(positions[t], dtype=np.float32).reshape(-1, 3).tofile(f"{inp}.pos{t}")
        if mass_table[t] == 0:
            np.asarray(masses[t], dtype=np.float32).tofile(f"{inp}.mass{t}")
    subprocess.run([TOOL, "write", out, str(nfiles), repr(float(box)), repr(float(redshift)), inp], check=True)


def dump_set(name, out):
    """The reader's view of a set: (info lines, positions[t], masses[t], totals)
    with totals[t] = (particles read, total mass as gen-pk accumulates it)."""
    subprocess.run([TOOL, "dump", name, out], check=True)
    info = open(out + ".info").read().splitlines()
    pos = [np.fromfile(f"{out}.pos{t}", dtype=np.float32).reshape(-1, 3) for t in range(6)]
    mas = [np.fromfile(f"{out}.mass{t}", dtype=np.float32) for t in range(6)]
    totals = {}
    for line in info:
        if line.startswith("type "):
            _, t, n, m = line.split()
            totals[int(t)] = (int(n), float(m))
    return info, pos, mas, totals


def file_ranges(n, nfiles):
    """Contiguous ranges of the writer's split of n particles over nfiles."""
    nf = max(nfiles, 1)
    return [(n * i // nf, n * (i + 1) // nf) for i in range(nf)]


def expected_total_mass(masses, mass_table, n, nfiles):
    """gen-pk's HDF5 total mass of a type (read_fieldize.cpp:320-331): per
    file the float32 masses summed in double in file order, then + the
    MassTable entry x the file's count; no +1."""
    total = 0.0
    for lo, hi in file_ranges(n, nfiles):
        if hi == lo:
            continue
        if mass_table == 0:
            s = 0.0
            for m in np.asarray(masses[lo:hi], dtype=np.float32):
                s += float(m)
            total += s
        total += mass_table * (hi - lo)
    return total
