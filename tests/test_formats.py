"""On-disk formats of the reference (host only, no GPU).

The library's sequential unformatted / list-directed readers and writers are
checked both ways against the reference's own compiled saveData / loadData
(src/libutil/serial/io/io.f90:9-33, 98-126); the direct-access format against a
record written and read by the Fortran runtime (oracle/fio_direct.f90:
iodirect.f90 itself does not compile with this toolchain).  scalit_layout is
checked against the closed forms of calOSB (src/iterate/module/osb.init.h:2-49).
"""
import ctypes as C
import os

import numpy as np
import pytest

from scalit_amd import _lib as L


def dptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def ours_save(path, a, binary):
    return L.lib().scalit_save_data(path.encode(), a.size, dptr(a), 1 if binary else 0)


def ours_load(path, n, binary):
    out = np.zeros(n)
    rc = L.lib().scalit_load_data(path.encode(), n, dptr(out), 1 if binary else 0)
    return rc, out


def sample(n, seed):
    g = np.random.default_rng(seed)
    a = g.standard_normal(n) * np.exp(g.uniform(-30, 30, n))
    a[: min(n, 4)] = [0.0, -0.0, 1.0 / 3.0, 1e-300][: min(n, 4)]
    return a


def ref_lib():
    from oracle import oracle as O

    R = O.ref()
    for f in ("savedata_", "loaddata_", "fio_save_dir_", "fio_load_dir_", "fio_save_grid_"):
        getattr(R, f).restype = C.c_int
    return R


def ref_save(path, a, binary):
    R = ref_lib()
    b = path.encode()
    return R.savedata_(C.byref(C.c_int(a.size)), dptr(a), C.byref(C.c_int(1 if binary else 0)), b, C.c_long(len(b)))


def ref_load(path, n, binary):
    R = ref_lib()
    out = np.zeros(n)
    b = path.encode()
    ok = R.loaddata_(C.byref(C.c_int(n)), dptr(out), C.byref(C.c_int(1 if binary else 0)), b, C.c_long(len(b)))
    return ok, out


@pytest.mark.ref
@pytest.mark.parametrize("binary", [True, False])
@pytest.mark.parametrize("n", [1, 7, 1000])
def test_sequential_matches_reference_both_ways(tmp_path, binary, n):
    a = sample(n, n)
    p1, p2 = str(tmp_path / "ref.dat"), str(tmp_path / "ours.dat")
    assert ref_save(p1, a, binary)
    rc, got = ours_load(p1, n, binary)
    assert rc == L.SCALIT_OK and np.array_equal(got, a) and np.array_equal(np.signbit(got), np.signbit(a))
    assert ours_save(p2, a, binary) == L.SCALIT_OK
    ok, back = ref_load(p2, n, binary)
    assert ok and np.array_equal(back, a)
    if binary:   # one record, identical bytes
        assert open(p1, "rb").read() == open(p2, "rb").read()


@pytest.mark.ref
def test_sequential_reads_prefix_of_longer_record(tmp_path):
    a = sample(50, 1)
    p = str(tmp_path / "long.dat")
    assert ref_save(p, a, True)
    rc, got = ours_load(p, 20, True)
    assert rc == L.SCALIT_OK and np.array_equal(got, a[:20])
    # a record shorter than the request is an error (the reference's READ has
    # no IOSTAT there, so its runtime stops the program instead)
    rc, _ = ours_load(p, 51, True)
    assert rc == L.SCALIT_EIO


@pytest.mark.ref
def test_direct_access_matches_fortran_runtime(tmp_path):
    R = ref_lib()
    a = sample(333, 2)
    p1, p2 = str(tmp_path / "f.dir"), str(tmp_path / "o.dir")
    b1 = p1.encode()
    assert R.fio_save_dir_(C.byref(C.c_long(a.size)), dptr(a), b1, C.c_long(len(b1)))
    out = np.zeros(a.size)
    assert L.lib().scalit_load_data_dir(p1.encode(), a.size, dptr(out)) == L.SCALIT_OK
    assert np.array_equal(out, a)
    assert L.lib().scalit_save_data_dir(p2.encode(), a.size, dptr(a)) == L.SCALIT_OK
    assert os.path.getsize(p2) == 8 * a.size and open(p1, "rb").read() == open(p2, "rb").read()
    back = np.zeros(a.size)
    b2 = p2.encode()
    assert R.fio_load_dir_(C.byref(C.c_long(a.size)), dptr(back), b2, C.c_long(len(b2)))
    assert np.array_equal(back, a)


@pytest.mark.ref
def test_per_point_grid_records_match_fortran_runtime(tmp_path):
    """The OUTH / H0GM file of P2SaveDataGrid (one direct record of NG x NG
    per grid point) written by the Fortran runtime reads back into the
    operator's point-fastest layout, and the library's writer gives the same
    bytes."""
    R = ref_lib()
    n, ng = 37, 9
    a = np.asfortranarray(sample(n * ng * ng, 5).reshape(n, ng, ng, order="F"))
    p1, p2 = str(tmp_path / "outh_1.dat"), str(tmp_path / "outh_2.dat")
    b1 = p1.encode()
    assert R.fio_save_grid_(C.byref(C.c_long(n)), C.byref(C.c_long(ng)), a.ctypes.data_as(C.POINTER(C.c_double)), b1,
                            C.c_long(len(b1)))
    out = np.zeros(n * ng * ng)
    assert L.lib().scalit_load_data_grid(b1, n, ng, dptr(out)) == L.SCALIT_OK
    assert np.array_equal(out, a.reshape(-1, order="F"))
    assert L.lib().scalit_save_data_grid(p2.encode(), n, ng, dptr(out)) == L.SCALIT_OK
    assert open(p1, "rb").read() == open(p2, "rb").read()
    # a file shorter than npts records is an error
    assert L.lib().scalit_load_data_grid(b1, n + 1, ng, dptr(np.zeros((n + 1) * ng * ng))) == L.SCALIT_EIO


def test_list_directed_syntax(tmp_path):
    p = tmp_path / "t.txt"
    p.write_text(" 1.5D+00, 2*3.0\n -4.25E-1   5\n")
    rc, got = ours_load(str(p), 5, False)
    assert rc == L.SCALIT_OK and np.array_equal(got, [1.5, 3.0, 3.0, -0.425, 5.0])
    p.write_text("1.0 2.0 x\n")
    assert ours_load(str(p), 3, False)[0] == L.SCALIT_EIO
    assert ours_load(str(p), 2, False)[0] == L.SCALIT_OK   # stops after the requested items


def test_missing_and_bad_arguments(tmp_path):
    assert ours_load(str(tmp_path / "none.dat"), 3, True)[0] == L.SCALIT_EIO
    assert L.lib().scalit_load_data_dir(str(tmp_path / "none").encode(), 3, dptr(np.zeros(3))) == L.SCALIT_EIO
    assert L.lib().scalit_save_data(None, 3, dptr(np.zeros(3)), 1) == L.SCALIT_EINVAL
    p = tmp_path / "short.dir"
    p.write_bytes(b"\0" * 16)
    assert L.lib().scalit_load_data_dir(str(p).encode(), 3, dptr(np.zeros(3))) == L.SCALIT_EIO


@pytest.mark.parametrize("sN,sDEP,ndvr", [((64, 96, 80), (0, 0, 1), 0), ((5, 7, 3, 4), (0, 1, 1, 0), 0),
                                          ((6, 5, 4), (0, 0, 0), 1)])
def test_layout_closed_form(sN, sDEP, ndvr):
    sF = len(sN)
    h0s, deps, vs, tot = (C.c_long * sF)(), (C.c_long * sF)(), (C.c_long * sF)(), (C.c_long * 5)()
    rc = L.lib().scalit_layout(sF, (C.c_int * sF)(*sN), (C.c_int * sF)(*sDEP), ndvr, h0s, deps, vs, tot)
    assert rc == L.SCALIT_OK
    N = int(np.prod(sN))
    dim = [int(np.prod(sN[:l])) for l in range(sF)]
    blk = [int(np.prod(sN[l + 1:])) for l in range(sF)]
    h0 = [0 if (ndvr and l == sF - 1) else sN[l] ** 2 for l in range(sF)]
    dep = [dim[l] if sDEP[l] else 0 for l in range(sF)]
    vos = [blk[l] * sN[l] ** 2 for l in range(sF)]
    assert list(h0s) == list(np.cumsum([0] + h0[:-1]))
    assert list(deps) == list(np.cumsum([0] + dep[:-1]))
    assert list(vs) == list(np.cumsum([0] + vos[:-1]))
    assert list(tot) == [N, sum(h0), sum(dep), sum(vos), sN[-1] * N if ndvr else 0]
    assert L.lib().scalit_layout(0, (C.c_int * 1)(4), (C.c_int * 1)(0), 0, None, None, None, tot) == L.SCALIT_EINVAL


@pytest.mark.ref
@pytest.mark.parametrize("mode", ["seq", "text", "dir"])
def test_complex_files_read_as_the_reference_reads_them(tmp_path, mode):
    """H0CX / DEPCX / OUTHCX files (sCX = T): loadData_CX
    (src/libutil/serial/io/io_cx.f90:102-126) reads one sequential record of
    N double complex values (binary) or N list-directed complex constants
    "(re, im)" (text); loadDataDir_CX one direct record.  A binary record of
    N complex values is a record of their 2N (re, im) doubles, so the real
    writers make them: the compiled loadData_CX reads what this library
    writes bit for bit, and Context.load_file_cx reads it back.  (The
    reference's saveData_CX declares its data double precision and writes N
    doubles: it cannot make these files, so none is written with it.)"""
    from scalit_amd.context import Context

    n = 123
    z = sample(n, 3) + 1j * sample(n, 4)
    p = str(tmp_path / "cx.dat")
    pairs = np.ascontiguousarray(np.stack([z.real, z.imag], axis=1).reshape(-1))
    if mode == "text":
        with open(p, "w") as f:
            for v in z:
                f.write(f" ({float(v.real)!r},{float(v.imag)!r})\n")
    else:
        Context.save_file(p, pairs, mode)
    got = Context.load_file_cx(p, n, mode)
    assert np.array_equal(got, z)
    if mode == "dir":
        return
    R = ref_lib()
    R.loaddata_cx_.restype = C.c_int
    out = np.zeros(2 * n)
    b = p.encode()
    ok = R.loaddata_cx_(C.byref(C.c_int(n)), dptr(out), C.byref(C.c_int(1 if mode == "seq" else 0)), b,
                        C.c_long(len(b)))
    assert ok and np.array_equal(out.view(np.complex128), z)
