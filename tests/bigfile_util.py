"""Writer of MP-Gadget bigfile snapshots for tests, in the on-disk layout the
gen-pk reader restates (genpk_amd/cli/bigfile.h): one directory per block with
a text header (DTYPE, NMEMB, NFILE, per data file "%06X: rows : cksum : sysv"),
raw %06X data files and an attr-v2 text file of hex-encoded attributes.  No
bigfile data or library ships with the reference, so this writer and the
reader are checked against each other and the reference's reading semantics
only (parity with the bigfile library is unpinned)."""
import os

import numpy as np


def _write_block(path, name, arr, dtype, nfiles=1, attrs=None):
    d = os.path.join(path, name)
    os.makedirs(d, exist_ok=True)
    arr = np.asarray(arr)
    rows = arr.shape[0] if arr.ndim else 0
    nmemb = 1 if arr.ndim <= 1 else arr.shape[1]
    with open(os.path.join(d, "header"), "w") as h:
        h.write(f"DTYPE: {dtype}\nNMEMB: {nmemb}\nNFILE: {nfiles}\n")
        bounds = [rows * i // max(nfiles, 1) for i in range(nfiles + 1)]
        for i in range(nfiles):
            h.write(f"{i:06X}: {bounds[i + 1] - bounds[i]} : 0 : 0\n")
    for i in range(nfiles):
        part = arr[bounds[i]:bounds[i + 1]].astype(np.dtype(dtype))
        part.tofile(os.path.join(d, f"{i:06X}"))
    with open(os.path.join(d, "attr-v2"), "w") as f:
        for k, (dt, vals) in (attrs or {}).items():
            raw = np.asarray(vals, dtype=np.dtype(dt)).tobytes()
            f.write(f"{k} {dt} {len(np.atleast_1d(vals))} {raw.hex().upper()} #HUMANE [ {vals} ]\n")


def write_snapshot(path, positions, masses, mass_table, box, time=0.5, pos_dtype="<f8", mass_dtype="<f4",
                   nfiles=2, npart_dtype="<u8"):
    """positions[t]: [n][3] (or None), masses[t]: [n] for types with
    mass_table[t] == 0."""
    npart = [0 if p is None else len(p) for p in positions]
    _write_block(path, "Header", np.zeros(0), "<f8", 0, {
        "TotNumPart": (npart_dtype, npart), "MassTable": ("<f8", list(mass_table)), "Time": ("<f8", [time]),
        "HubbleParam": ("<f8", [0.7]), "Omega0": ("<f8", [0.3]), "BoxSize": ("<f8", [box]),
        "UnitLength_in_cm": ("<f8", [3.085678e21])})
    for t in range(6):
        if positions[t] is None:
            continue
        _write_block(path, f"{t}/Position", np.asarray(positions[t], dtype=np.float64), pos_dtype, nfiles)
        if mass_table[t] == 0:
            _write_block(path, f"{t}/Mass", np.asarray(masses[t]), mass_dtype, nfiles)


def expected_total_mass(masses, mass_table, n):
    """read_fieldize_bigfile.cpp:94-114: the float32 masses summed in double
    in block order, then + MassTable x count; no +1."""
    total = 0.0
    if mass_table == 0:
        s = 0.0
        for m in np.asarray(masses, dtype=np.float32):
            s += float(m)
        total += s
    return total + mass_table * n
