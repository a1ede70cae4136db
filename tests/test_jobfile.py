"""The iterate job's parameter file (host only, no GPU).

scalit_amd.jobfile.read_iterate_input is compared field by field with what the
reference's own reader (myReadOSB of the serial OSB module,
src/iterate/module/osb.print.h:192-255, compiled into oracle/_ref) reads from
the same files.  The reference ships no example input; the files below cover
the reader's branches (dependent levels, the sNDVR layer, VOSB save/load names,
list-directed syntax across lines, repeat counts, commas, D exponents).
"""
import ctypes as C

import pytest

from scalit_amd import jobfile as J

INPUTS = {
    "bound_dep": """3  64 96 80
F F T
1 1
F F T F
1000 1.0D-5   500 1.0e-6
-0.02 1e-10 30 10 200 20 2
-0.02 0.004 0.3 5
0 0 0 0 0
data/ja3.h0
data/ja3.res
data/ja3.dep
""",
    "ndvr_vosb": """4, 24, 32, 2*24
T,T,F,T
1 3
.false. .true. .true. .false.
 200 1.5D-04
 300 1.0D-07
 -1.25D-2  1.0D-11  40  5  400  10  1
 -1.25D-2 -2.0D-3 0.5 7
 0 -101 0 0 0
   inputs/h0 with spaces.dat
inputs/outh.dat
inputs/dep.dat
ckpt/vosb.dat
ckpt/eig0.dat
""",
    "no_dep_hosb": """2 10 12
F F
1 2
F F F F
50 1e-3 60 1e-4
0.0 1e-9 10 5 100 8 0
0.0 0.01 0.0 0
3 120 0 0 0
h0.bin
res.bin
hosb.bin
vosb.bin
eig.bin
""",
}


def ref_read(path, tmp_path):
    from oracle import oracle as O

    R = O.ref()
    out = str(tmp_path / "dump.txt")
    a, b = path.encode(), out.encode()
    R.refin_dump_(a, b, C.c_long(len(a)), C.c_long(len(b)))
    return [l.rstrip("\n") for l in open(out)]


def ours_as_dump(job):
    d = [str(job.sF)] + [str(n) for n in job.sN] + ["T" if v else "F" for v in job.sDEP]
    d += [str(job.sJOB), str(job.sOSB)] + ["T" if v else "F" for v in (job.sCX, job.sNDVR, job.sST, job.sAP)]
    d += [str(job.bj_max), repr(job.bj_tol), str(job.qmr_max), repr(job.qmr_tol), repr(job.conv_e0),
          repr(job.conv_tol)]
    d += [str(v) for v in (job.conv_start, job.conv_step, job.conv_max, job.conv_num, job.conv_gap)]
    d += [repr(job.osbw_e0), repr(job.osbw_de), repr(job.osbw_beta), str(job.osbw_cnt)]
    d += [str(v) for v in (job.sHOSB, job.sVOSB, job.sHW, job.sVX, job.sPT)]
    f = job.files
    d += [f["H0"], f["OUTH"] if job.sNDVR else f["RES"]]
    if any(job.sDEP):
        d.append(f["DEP"])
    if job.sHOSB:
        d.append(f["HOSB"])
    if job.sVOSB:
        d += [f["VOSB"], f["EIG"]]
    return d


def normalise(dump):
    """Numbers as numbers, logicals as T/F, names verbatim."""
    out = []
    for s in dump:
        t = s.strip()
        try:
            out.append(float(t))
            continue
        except ValueError:
            pass
        out.append(t if t in ("T", "F") else s.rstrip())
    return out


@pytest.mark.ref
@pytest.mark.parametrize("name", sorted(INPUTS))
def test_reader_matches_reference(tmp_path, name):
    p = tmp_path / "iterate.in"
    p.write_text(INPUTS[name])
    ref = normalise(ref_read(str(p), tmp_path))
    ours = normalise(ours_as_dump(J.read_iterate_input(str(p))))
    assert ours == ref


def test_reader_rules():
    job = J.read_iterate_input(INPUTS["ndvr_vosb"])
    assert job.sN == [24, 32, 24, 24]
    assert job.sDEP == [False, True, False, False]   # sDEP(1) and, with sNDVR, sDEP(sF) forced false
    assert job.osbw_de == 2.0e-3                      # |DE|
    assert job.files["H0"] == "   inputs/h0 with spaces.dat"
    assert J.pist_params(job, 10) == (-1.25e-2, 1.0e-11, 40, 5, 1, 400)
    job2 = J.read_iterate_input(INPUTS["no_dep_hosb"])
    # setPISTConvParam defaults: start < M0 -> 2 M0; gap < 1 -> 1; max < start -> start + M0
    assert J.pist_params(job2, 8) == (0.0, 1e-9, 10, 5, 1, 100)
    job2.conv_start, job2.conv_step, job2.conv_max = 3, 0, 5
    assert J.pist_params(job2, 8) == (0.0, 1e-9, 16, 5, 1, 24)
