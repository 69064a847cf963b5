"""Field-file and session.num I/O of the library (host code, no GPU).

* Field files written by sx_field_write are read by the reference's own
  format converter (utility/convert.c, compiled from its source into
  oracle/_ref by oracle/Makefile): its ASCII output holds the same header
  values and every datum to the printed precision.
* The converter's byte-swapped output (-s, "binary IEEE big-endian") and
  its binary output are read back by sx_field_read bit-exactly (swab path).
* The header text follows writeField's formats line by line
  (oracle/fieldfile.py); appended dumps are addressed by number; the plane
  padding is zeroed on read.
* session.num: sx_num_write produces enumerate's printout
  (oracle/fieldfile.num_text) and sx_num_read returns the numbering,
  masks and sizes of a mesh's velocity and pressure systems exactly;
  mismatched element counts are refused like BCmgr::buildnum refuses them.
"""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

from semtex_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CONVERT = os.path.join(ROOT, "oracle", "_ref", "convert")


class Header(C.Structure):
    _fields_ = [("session", C.c_char * 64), ("created", C.c_char * 64), ("nr", C.c_int),
                ("ns", C.c_int), ("nz", C.c_int), ("nel", C.c_int), ("step", C.c_int),
                ("time", C.c_double), ("dt", C.c_double), ("kinvis", C.c_double),
                ("beta", C.c_double), ("fields", C.c_char * 32), ("format", C.c_char * 64)]


def _write(path, fields, data, np_, nz, nel, stride, append=0, **kw):
    L = _lib.load()
    h = Header()
    h.session = kw.get("session", "chan").encode()
    h.created = kw.get("created", "").encode()
    h.nr = h.ns = np_
    h.nz, h.nel, h.step = nz, nel, kw.get("step", 0)
    h.time, h.dt = kw.get("time", 0.0), kw.get("dt", 0.0)
    h.kinvis, h.beta = kw.get("kinvis", 0.0), kw.get("beta", 0.0)
    h.fields = fields.encode()
    arrs = [np.ascontiguousarray(data[c]) for c in fields]
    ptrs = (C.c_void_p * len(arrs))(*[a.ctypes.data for a in arrs])
    rc = L.sx_field_write(path.encode(), C.byref(h), ptrs, stride, append)
    _lib.check(rc, "sx_field_write")


def _read(path, dump, name, nz, stride):
    out = np.full((nz, stride), 7.0)
    rc = _lib.load().sx_field_read(path.encode(), dump, name.encode(), out.ctypes.data_as(C.c_void_p),
                                   stride)
    _lib.check(rc, "sx_field_read")
    return out


def _fields(np_=5, nz=4, nel=6, seed=0):
    rng = np.random.default_rng(seed)
    n = nel * np_ * np_
    stride = n + 3
    data = {c: rng.standard_normal((nz, stride)) for c in "uvwp"}
    return data, n, stride


def test_field_file_header_and_data_layout(tmp_path):
    from oracle.fieldfile import header_text, read_field_file
    data, n, stride = _fields()
    path = str(tmp_path / "a.fld")
    _write(path, "uvwp", data, 5, 4, 6, stride, step=12, time=0.012, dt=1e-3,
           kinvis=5e-4, beta=1.0, created="Thu Dec 03 10:00:00 2026")
    raw = open(path, "rb").read()
    want = header_text("chan", "Thu Dec 03 10:00:00 2026", 5, 4, 6, 12, 0.012, 1e-3, 5e-4, 1.0,
                       "uvwp")
    assert raw[:len(want)] == want.encode()
    (h, f), = read_field_file(path)
    assert (h["nr"], h["nz"], h["nel"], h["step"], h["fields"]) == (5, 4, 6, 12, "uvwp")
    for c in "uvwp":
        assert np.array_equal(f[c], data[c][:, :n])


def test_appended_dumps_and_padding(tmp_path):
    data, n, stride = _fields(seed=1)
    data2, _, _ = _fields(seed=2)
    path = str(tmp_path / "b.fld")
    _write(path, "uvp", data, 5, 4, 6, stride, step=1)
    _write(path, "uvp", data2, 5, 4, 6, stride, step=2, append=1)
    h = Header()
    _lib.check(_lib.load().sx_field_read_header(path.encode(), 1, C.byref(h)), "read_header")
    assert h.step == 2 and h.fields == b"uvp"
    got = _read(path, 1, "p", 4, stride)
    assert np.array_equal(got[:, :n], data2["p"][:, :n])
    assert np.all(got[:, n:] == 0.0)
    assert _lib.load().sx_field_read(path.encode(), 0, b"w", got.ctypes.data_as(C.c_void_p),
                                     stride) == _lib.SX_ERR_ARG


@pytest.mark.skipif(not os.path.exists(CONVERT), reason="reference convert not built")
def test_reference_convert_reads_and_writes_our_files(tmp_path):
    data, n, stride = _fields(np_=4, nz=2, nel=3, seed=3)
    path = str(tmp_path / "c.fld")
    _write(path, "uv", data, 4, 2, 3, stride, step=7, time=0.5, dt=0.01, kinvis=0.02, beta=2.0)
    # binary -> ASCII by the reference's converter
    asc = subprocess.run([CONVERT, "-a", path], capture_output=True, text=True, check=True).stdout
    lines = asc.splitlines()
    assert lines[2].split()[:4] == ["4", "4", "2", "3"]
    assert lines[3].split()[0] == "7" and float(lines[4].split()[0]) == 0.5
    assert lines[8].split()[0] == "uv"
    vals = np.array([float(t) for ln in lines[10:] for t in ln.split()])
    # ASCII data: for every node, the fields side by side
    got = vals.reshape(-1, 2)
    want = np.stack([data["u"][:, :n].ravel(), data["v"][:, :n].ravel()], axis=1)
    assert np.allclose(got, want, rtol=1e-6, atol=1e-12)
    # ASCII -> byte-swapped binary by the converter; our reader swaps back
    apath = str(tmp_path / "c.asc")
    open(apath, "w").write(asc)
    swapped = str(tmp_path / "c_be.fld")
    subprocess.run([CONVERT, "-s", "-o", swapped, apath], check=True, capture_output=True)
    assert b"big-endian" in open(swapped, "rb").read(2000)
    vb = _read(swapped, 0, "v", 2, stride)
    assert np.allclose(vb[:, :n], data["v"][:, :n], rtol=1e-6, atol=1e-12)
    # binary -> binary (machine order) by the converter: bit-exact
    same = str(tmp_path / "c_le.fld")
    subprocess.run([CONVERT, "-b", "-o", same, path], check=True, capture_output=True)
    assert np.array_equal(_read(same, 0, "u", 2, stride)[:, :n], data["u"][:, :n])


def test_big_endian_file_is_swapped_on_read(tmp_path):
    from oracle.fieldfile import write_field_file
    data, n, stride = _fields(seed=4)
    path = str(tmp_path / "d.fld")
    write_field_file(path, "uw", data, 5, 4, 6, big_endian=True)
    assert np.array_equal(_read(path, 0, "w", 4, stride)[:, :n], data["w"][:, :n])


def test_session_num_write_and_read(tmp_path):
    from oracle.fieldfile import num_text
    from semtex_amd.mesh import rect_mesh
    mu = rect_mesh(3, 4, 6, 0.0, 2 * np.pi, -1.0, 1.0, bc=("D", "N", "D", "N"), periodic_x=True)
    mp = rect_mesh(3, 4, 6, 0.0, 2 * np.pi, -1.0, 1.0, bc=("N", "N", "N", "N"), periodic_x=True)
    L = _lib.load()
    path = str(tmp_path / "chan.num")
    sets = [b"uvw", b"p"]
    bt = [np.ascontiguousarray(m.btog.ravel(), np.int32) for m in (mu, mp)]
    bm = [np.ascontiguousarray(m.bmask.ravel(), np.int32) for m in (mu, mp)]
    iv = lambda v: (C.c_int * 2)(*v)  # noqa: E731
    rc = L.sx_num_write(path.encode(), b"uvwp", 2, (C.c_char_p * 2)(*sets), mu.nel, mu.np,
                        (C.c_void_p * 2)(*[a.ctypes.data for a in bt]),
                        (C.c_void_p * 2)(*[a.ctypes.data for a in bm]),
                        iv([mu.nglobal, mp.nglobal]), iv([mu.nsolve, mp.nsolve]), iv([1, 1]),
                        iv([mu.nband, mp.nband]))
    _lib.check(rc, "sx_num_write")
    want = num_text("uvwp", ["uvw", "p"], mu.nel, mu.np, bt, bm, [mu.nglobal, mp.nglobal],
                    [mu.nsolve, mp.nsolve], [1, 1], [mu.nband, mp.nband])
    assert open(path).read() == want
    nb = mu.nel * mu.next
    nset = C.c_int(0)
    names = C.create_string_buffer(2 * 16)
    btog = np.zeros(2 * nb, np.int32)
    bmask = np.zeros(2 * nb, np.int32)
    ng, ns, ol, nbw = (np.zeros(2, np.int32) for _ in range(4))
    p = lambda a: a.ctypes.data_as(C.c_void_p)  # noqa: E731
    rc = L.sx_num_read(path.encode(), mu.nel, mu.np, 2, C.byref(nset), names, 16, p(btog),
                       p(bmask), p(ng), p(ns), p(ol), p(nbw))
    _lib.check(rc, "sx_num_read")
    assert nset.value == 2 and names.raw[:3] == b"uvw" and names.raw[16:17] == b"p"
    assert np.array_equal(btog[:nb], bt[0]) and np.array_equal(btog[nb:], bt[1])
    assert np.array_equal(bmask[:nb], bm[0]) and np.array_equal(bmask[nb:], bm[1])
    assert list(ng) == [mu.nglobal, mp.nglobal] and list(ns) == [mu.nsolve, mp.nsolve]
    assert list(nbw) == [mu.nband, mp.nband]
    rc = L.sx_num_read(path.encode(), mu.nel + 1, mu.np, 2, C.byref(nset), names, 16, p(btog),
                       p(bmask), p(ng), p(ns), p(ol), p(nbw))
    assert rc == _lib.SX_ERR_ARG and b"number of elements" in L.sx_last_error()
