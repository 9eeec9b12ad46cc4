"""Failure handling through the C-ABI (SURVEY's failure table: osemrecon.c:416-425,
drftabio.c:176-208, krnlio.c:221-253, IrlSetup.c:136-139, 158-162, 215-218,
651-653, krnlprj3d.c:559-562, irlosem.c:118-124, 134-139): IrlOsem returns
nonzero and pchIrlErrorString() holds the reference's message.  Every case
fails in host-side setup, before any GPU work, so these run on the CPU."""
import ctypes
import os
import struct

import numpy as np
import pytest

import libirl_amd as L
from libirl_amd import synth
from tests.cases import Problem

pytestmark = pytest.mark.skipif(not os.path.exists(L.LIB_PATH), reason="libirl_amd.so not built")

_libc = ctypes.CDLL(None)
_libc.malloc.restype = ctypes.c_void_p
_libc.malloc.argtypes = [ctypes.c_size_t]


@pytest.fixture(scope="module")
def pb():
    return Problem(n=16, nz=4, nviews=8, per=2, iters=1, drf=True)


def _run(pb, parms=None, opts=None, drf="default", krnl=None, views=None, log=None, msg=None):
    rc = L.irl_osem(parms or pb.parms(), opts or pb.options(), views or pb.views,
                    pb.drf_file if drf == "default" else drf, krnl, None, pb.mu.copy(),
                    pb.prj.copy(), pb.x0.copy(), log_file=log, msg_file=msg)
    return rc, L.error_string()


def _esse_parms(pb):
    p = pb.parms()
    p.iNumSrfIterations = 1
    return p


def _krnl(pb, tmp_path, noffs=5, ndepths=9):
    k, km = synth.gaussian_mixture_krnl(noffs, ndepths, pb.pix)
    f = str(tmp_path / "k.krnl")
    synth.write_krnl(f, k, km, 0.15, pb.pix)
    return f


def _patched(src, dst, offset, data):
    b = bytearray(open(src, "rb").read())
    b[offset:offset + len(data)] = data
    open(dst, "wb").write(bytes(b))
    return dst


def test_subset_count_errors(pb):
    p = pb.parms()
    p.NumAngPerSubset = 3
    rc, msg = _run(pb, p)
    assert rc != 0
    assert msg == ("No. of angles not an integer multiple of the no. of angles per subset\n"
                   " NumAngles=8, NumAngPerSet=3")
    p = pb.parms()
    p.NumViews, p.NumAngPerSubset = 6, 2
    rc, msg = _run(pb, p, views=pb.views[:6])
    assert rc != 0 and msg == "Number of subsets (3) for num_ang_per_set=2 is not even\n"


def test_drf_file_errors(pb, tmp_path):
    rc, msg = _run(pb, drf=str(tmp_path / "missing.drft"))
    assert rc != 0 and msg == "Error opening DRF Table file: %s" % (tmp_path / "missing.drft")
    bad = _patched(pb.drf_file, str(tmp_path / "sig.drft"), 0, b"DRFX")
    rc, msg = _run(pb, drf=bad)
    assert rc != 0 and msg == "Signature in Drf File %s is DRFX, should be DRFT" % bad
    bad = _patched(pb.drf_file, str(tmp_path / "ver.drft"), 4, struct.pack("<i", 7))
    rc, msg = _run(pb, drf=bad)
    assert rc != 0 and msg == "Version in Drf File %s is  7 hex, should be 1" % bad
    # a byte-swapped version is accepted as a big-endian file
    p = pb.parms()
    p.BinWidth = pb.pix * 1.05
    p.fAtnScaleFac = p.BinWidth
    rc, msg = _run(pb, p)
    assert rc != 0
    assert msg == ("The pixel size in Drf Table %s is 0.5000, and not 0.5250 as needed for "
                   "recon." % pb.drf_file)


def test_krnl_file_errors(pb, tmp_path):
    good = _krnl(pb, tmp_path)
    missing = str(tmp_path / "missing.krnl")
    rc, msg = _run(pb, _esse_parms(pb), krnl=missing)
    assert rc != 0 and msg == "Error opening Krnl Table file: %s" % missing
    bad = _patched(good, str(tmp_path / "sig.krnl"), 0, b"KRNX")
    rc, msg = _run(pb, _esse_parms(pb), krnl=bad)
    assert rc != 0 and msg == "Signature in Krnl File %s is KRNX, should be KRNL" % bad
    bad = _patched(good, str(tmp_path / "ver.krnl"), 4, struct.pack("<i", 0x33))
    rc, msg = _run(pb, _esse_parms(pb), krnl=bad)
    assert rc != 0 and msg == "Version in Krnl File %s is  33 hex, should be 1" % bad


def test_scatter_collapse_errors(pb, tmp_path):
    good = _krnl(pb, tmp_path)
    for fac in (0, 5):
        p = _esse_parms(pb)
        p.SrfCollapseFac = fac
        rc, msg = _run(pb, p, krnl=good)
        assert rc != 0 and msg == "srf_collapse_fac = %d: must be 1 (no collapse) to 4" % fac
    # 3 does not divide the 16 x 4 x 16 volume (krnlprj3d.c:559-562)
    p = _esse_parms(pb)
    p.SrfCollapseFac = 3
    rc, msg = _run(pb, p, krnl=good)
    assert rc != 0
    assert msg.startswith("factor does not evenly divide image dimensions")
    assert "iFac=3" in msg


def _malloc_array(ctype, values):
    n = len(values)
    ptr = _libc.malloc(ctypes.sizeof(ctype) * max(n, 1))
    arr = (ctype * n).from_address(ptr)
    for i, v in enumerate(values):
        arr[i] = v
    return ptr


def test_per_view_data_errors(pb):
    lib = L.load()
    # acquisition times for the wrong number of views (IrlSetup.c:136-139);
    # the arrays pass to the library, which frees them at the end of the call
    times = _malloc_array(ctypes.c_float, [1.0] * 5)
    assert lib.iSetupExtraPrjViewData(5, times, None) == 0
    rc, msg = _run(pb)
    assert rc != 0
    assert msg == "Number of Prj Views passed to SetupExtrPrjViewData is incorrect"
    # illegal detector index (IrlSetup.c:651-653)
    nums = _malloc_array(ctypes.c_int, [0, 0, 0, 1, 0, 0, 0, 0])
    assert lib.iSetupExtraPrjViewData(8, None, nums) == 0
    rc, msg = _run(pb)
    assert rc != 0 and msg == "Illegal detector specified for view 3 (1)"
    # return codes of the setters themselves (IrlSetup.c:376-401, 455-461)
    assert lib.iSetupExtraPrjViewData(-1, None, None) & 1
    assert lib.iSetupExtraPrjViewData(0, None, None) == 0
    neg = _malloc_array(ctypes.c_float, [1.0, -1.0])
    assert lib.iSetupExtraPrjViewData(2, neg, None) == 2
    assert lib.iSetupExtraPrjViewData(0, None, None) == 0
    assert lib.iSetupDetectors(-1, None) == 1
    assert lib.iSetupDetectors(1, None) == 2
    assert lib.iSetupDetectors(0, None) == 0
    # the times array stays registered (a NULL argument keeps the stored
    # pointer, IrlSetup.c:366-372); the next IrlOsem reports the mismatch and
    # frees it, as every IrlOsem does on return
    rc, msg = _run(pb)
    assert rc != 0 and msg == "Number of Prj Views passed to SetupExtrPrjViewData is incorrect"
    p = pb.parms()
    p.NumAngPerSubset = 3
    rc, msg = _run(pb, p)
    assert rc != 0 and msg.startswith("No. of angles not an integer multiple")


def test_log_and_message_file_errors(pb, tmp_path):
    bad = str(tmp_path / "no" / "such" / "dir" / "f")
    rc, msg = _run(pb, log=bad)
    assert rc != 0 and msg == "Error opening log file %s" % bad
    rc, msg = _run(pb, msg=bad)
    assert rc != 0 and msg == "Error opening messages file %s" % bad
