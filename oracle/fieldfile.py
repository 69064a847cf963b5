"""Restatement of semtex's file formats for checking the library's readers
and writers.

TEST INFRASTRUCTURE ONLY.

Field files (binary nekton format):
  writeField       src/misc.cpp:106-180   (header lines, field data order)
  Header >>        src/data2df.cpp:454-470 (how the header is parsed)
  AuxField <<, >>  src/auxfield.cpp:1021-1126 (nz planes of nel np^2 values)
  format string    veclib/xbrev.c:16-50

session.num:
  printup          utility/enumerate.cpp:441-552
"""
from __future__ import annotations

import numpy as np

HDR_FMT = ["%-25s Session\n", "%-25s Created\n", "%-25s Nr, Ns, Nz, Elements\n",
           "%-25d Step\n", "%-25.6g Time\n", "%-25.6g Time step\n", "%-25.6g Kinvis\n",
           "%-25.6g Beta\n", "%-25s Fields written\n", "%-25s Format\n"]


def header_text(session, created, np_, nz, nel, step, time, dt, kinvis, beta, fields,
                fmt="binary IEEE little-endian"):
    vals = [session, created, f"{np_} {np_} {nz} {nel}", step, time, dt, kinvis, beta, fields,
            fmt]
    return "".join(f % v for f, v in zip(HDR_FMT, vals))


def write_field_file(path, fields, data, np_, nz, nel, step=0, time=0.0, dt=0.0, kinvis=0.0,
                     beta=0.0, session="session", created="Thu Jan 01 00:00:00 1970",
                     big_endian=False, append=False):
    """data: {name: [nz][>= nel np^2] array}, written in `fields` order."""
    fmt = "binary IEEE big-endian" if big_endian else "binary IEEE little-endian"
    n = nel * np_ * np_
    with open(path, "ab" if append else "wb") as f:
        f.write(header_text(session, created, np_, nz, nel, step, time, dt, kinvis, beta,
                            fields, fmt).encode())
        dt_ = ">f8" if big_endian else "<f8"
        for c in fields:
            f.write(np.ascontiguousarray(np.asarray(data[c])[:, :n], dtype=dt_).tobytes())


def read_field_file(path):
    """All dumps of a field file: list of (header dict, {name: [nz][n]})."""
    raw = open(path, "rb").read()
    out, pos = [], 0
    while pos < len(raw):
        lines = []
        for _ in range(10):
            end = raw.index(b"\n", pos)
            lines.append(raw[pos:end].decode())
            pos = end + 1
        h = {"session": lines[0][:24].rstrip(), "created": lines[1][:24].rstrip()}
        nr, ns, nz, nel = (int(v) for v in lines[2].split()[:4])
        h.update(nr=nr, ns=ns, nz=nz, nel=nel, step=int(lines[3].split()[0]),
                 time=float(lines[4].split()[0]), dt=float(lines[5].split()[0]),
                 kinvis=float(lines[6].split()[0]), beta=float(lines[7].split()[0]),
                 fields=lines[8].split()[0], format=lines[9][:25].rstrip())
        dt_ = ">f8" if "big" in h["format"] else "<f8"
        n = nel * nr * ns
        fl = {}
        for c in h["fields"]:
            fl[c] = np.frombuffer(raw, dtype=dt_, count=nz * n, offset=pos).reshape(nz, n)
            pos += 8 * nz * n
        out.append((h, fl))
    return out


def num_text(fields, set_fields, nel, np_, btog, bmask, nglobal, nsolve, optlev, nband):
    """enumerate's printup of nset numbering systems."""
    nset = len(set_fields)
    next_, nint = 4 * (np_ - 1), (np_ - 2) ** 2
    L = [f"# FIELDS         :  {fields}\n"]
    sep = "# ----------------" + "  -------------" * nset + "\n"
    L.append(sep)
    L.append(f"# {nset} NUMBER SETS  :" + "".join(f"{s:>15}" for s in set_fields) + "\n")
    for label, vals in (("NEL", [nel] * nset), ("NP_MAX", [np_] * nset),
                        ("NEXT_MAX", [next_] * nset), ("NINT_MAX", [nint] * nset),
                        ("NTOTAL", [nel * np_ * np_] * nset),
                        ("NBOUNDARY", [nel * next_] * nset), ("NGLOBAL", nglobal),
                        ("NSOLVE", nsolve), ("OPTIMIZATION", optlev), ("BANDWIDTH", nband)):
        L.append(f"# {label:<15}:" + "".join(f"{v:>15}" for v in vals) + "\n")
    L.append(sep)
    L.append("# elmt  side offst" + "     bmap  mask" * nset + "\n")
    i = 0
    for k in range(1, nel + 1):
        for side in range(1, 5):
            for soff in range(np_ - 1):
                row = f"{k:>6}{side:>6}{soff:>6}"
                for j in range(nset):
                    row += f"{btog[j][i]:>9}{bmask[j][i]:>6}"
                L.append(row + "\n")
                i += 1
    return "".join(L)
