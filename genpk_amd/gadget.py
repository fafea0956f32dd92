"""Gadget-2 snapshot files in Python: a writer (format I or II, one or more
files) and an independent reader of positions, masses and the header.

The gen-pk command line tool reads snapshots with its own C++ reader; this
module makes snapshots for it and reads them back for comparison.  Layout:
every record framed by 4-byte length markers; format II puts an 8-byte
record (4-character label, framed size of the next block) before each block;
256-byte header; particle blocks in type order 0..5; the MASS block holds only
the types whose header mass is 0.
"""
import os
import struct

import numpy as np

N_TYPE = 6
_HEAD = struct.Struct("<6i6d2d2i6I2i4d2i6I")


def _record(f, payload, label=None):
    if label is not None:
        f.write(struct.pack("<i4sii", 8, label.encode().ljust(4)[:4], len(payload) + 8, 8))
    f.write(struct.pack("<i", len(payload)))
    f.write(payload)
    f.write(struct.pack("<i", len(payload)))


def write_snapshot(path, positions, masses, box, nfiles=1, fmt=2, redshift=0.0, omega0=0.3, omega_lambda=0.7,
                   hubble=0.7):
    """positions: 6 float32 arrays [n_t, 3] (or None); masses: per type a
    float (header mass) or a float32 array [n_t] (MASS block, header mass 0).
    nfiles > 1 writes path.0 ... path.{nfiles-1}, splitting each type's
    particles into contiguous ranges.  Returns the list of files written."""
    pos = [np.zeros((0, 3), np.float32) if p is None else np.ascontiguousarray(p, np.float32) for p in positions]
    ntot = [p.shape[0] for p in pos]
    hmass = [0.0 if isinstance(m, np.ndarray) else float(m) for m in masses]
    files = [path] if nfiles == 1 else [f"{path}.{i}" for i in range(nfiles)]
    ids = 0
    for fi, fn in enumerate(files):
        rng = [(t * 0 + (ntot[t] * fi) // nfiles, (ntot[t] * (fi + 1)) // nfiles) for t in range(N_TYPE)]
        npart = [b - a for a, b in rng]
        head = _HEAD.pack(*npart, *hmass, 1.0 / (1.0 + redshift), redshift, 0, 0, *[n & 0xFFFFFFFF for n in ntot],
                          0, nfiles, float(box), omega0, omega_lambda, hubble, 0, 0, *[n >> 32 for n in ntot])
        head = head + b"\0" * (256 - len(head))
        lab = (lambda s: s) if fmt == 2 else (lambda s: None)
        with open(fn, "wb") as f:
            _record(f, head, lab("HEAD"))
            _record(f, b"".join(pos[t][a:b].tobytes() for t, (a, b) in enumerate(rng)), lab("POS "))
            _record(f, np.zeros((sum(npart), 3), np.float32).tobytes(), lab("VEL "))
            _record(f, np.arange(ids, ids + sum(npart), dtype=np.uint32).tobytes(), lab("ID  "))
            ids += sum(npart)
            mb = b"".join(np.ascontiguousarray(masses[t][a:b], np.float32).tobytes()
                          for t, (a, b) in enumerate(rng) if isinstance(masses[t], np.ndarray) and b > a)
            if mb:
                _record(f, mb, lab("MASS"))
    return files


def _blocks(fn):
    out = {}
    with open(fn, "rb") as f:
        data = f.read()
    (m,) = struct.unpack_from("<i", data, 0)
    fmt2 = m == 8
    order = ["HEAD", "POS ", "VEL ", "ID  ", "MASS"]
    off, k = 0, 0
    while off < len(data):
        if fmt2:
            lab = data[off + 4:off + 8].decode()
            off += 16
        else:
            if k >= len(order):
                break
            lab = order[k]
        (n,) = struct.unpack_from("<i", data, off)
        out[lab] = data[off + 4:off + 4 + n]
        off += n + 8
        k += 1
    return out


def read_snapshot(name):
    """Header of the first file, and per type the positions (float32 [n, 3])
    and masses (float32 [n], from the header mass or the MASS block) of all
    files, in file order."""
    if os.path.exists(name):
        first = name
    elif os.path.exists(name + ".0"):
        first = name + ".0"
    else:
        raise FileNotFoundError(name)
    b0 = _blocks(first)
    h = _HEAD.unpack(b0["HEAD"][:_HEAD.size])
    nfiles = h[23]
    files = [first]
    if nfiles > 1 and first.endswith(".0"):
        files = [first[:-2] + f".{i}" for i in range(nfiles)]
    header = dict(mass=list(h[6:12]), time=h[12], redshift=h[13], npart_total=list(h[16:22]), num_files=nfiles,
                  box=h[24], omega0=h[25], omega_lambda=h[26], hubble=h[27])
    pos = [[] for _ in range(N_TYPE)]
    mas = [[] for _ in range(N_TYPE)]
    for fn in files:
        bl = _blocks(fn)
        hh = _HEAD.unpack(bl["HEAD"][:_HEAD.size])
        npart = list(hh[0:6])
        hm = list(hh[6:12])
        p = np.frombuffer(bl["POS "], np.float32).reshape(-1, 3)
        mblk = np.frombuffer(bl["MASS"], np.float32) if "MASS" in bl else np.zeros(0, np.float32)
        o = mo = 0
        for t in range(N_TYPE):
            pos[t].append(p[o:o + npart[t]])
            if npart[t] > 0 and hm[t] == 0:
                mas[t].append(mblk[mo:mo + npart[t]])
                mo += npart[t]
            else:
                mas[t].append(np.full(npart[t], hm[t], np.float32))
            o += npart[t]
    header["npart"] = [int(sum(x.shape[0] for x in pos[t])) for t in range(N_TYPE)]
    return header, [np.concatenate(x) for x in pos], [np.concatenate(x) for x in mas]
