"""The C-ABI library: builds, loads, exports every function the headers in
include/ declare, and its structs match the ctypes mirrors.  No GPU work."""
import ctypes
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INCLUDE = os.path.join(ROOT, "include")


def declared_functions():
    names = set()
    for h in os.listdir(INCLUDE):
        if not h.endswith(".h"):
            continue
        text = open(os.path.join(INCLUDE, h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        text = re.sub(r"typedef[^;]*;", "", text, flags=re.S)
        text = re.sub(r"\{[^{}]*\}", "", text, flags=re.S)
        for m in re.finditer(r"\b([A-Za-z_]\w*)\s*\([^;]*\)\s*;", text):
            if m.group(1) not in ("if", "while", "for", "sizeof", "return"):
                names.add(m.group(1))
    return names


def test_headers_declare_the_reference_entry_points():
    names = declared_functions()
    for fn in ("IrlOsem", "IrlGenprj", "pchIrlErrorString", "OsemSetAdvOptions",
               "OsemSetPrevIterHook", "OsemSetPostSubsetHook", "iSetupExtraPrjViewData",
               "iSetupDetectors", "pvStoreDetectorData", "InitPrj", "Project", "BackProject",
               "FreePrj", "rot3dpar", "ConvolveDrfYpad", "sReadDrfTab", "vWriteDrfTab"):
        assert fn in names, fn


def test_library_exports_every_declared_function():
    import libirl_amd as L
    if not os.path.exists(L.LIB_PATH):
        pytest.skip("libirl_amd.so not built")
    out = subprocess.run(["nm", "-D", "--defined-only", L.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = {ln.split()[-1] for ln in out.splitlines() if ln.strip()}
    missing = sorted(declared_functions() - exported)
    assert not missing, missing
    L.load()  # loads without a GPU


def test_ctypes_struct_layouts_match_reference():
    import libirl_amd as L
    assert ctypes.sizeof(L.IrlParms) == 72
    assert L.IrlParms.pchNormImageBase.offset == 40
    assert L.IrlParms.fIntrinsicFWHM.offset == 64
    assert ctypes.sizeof(L.Options) == 36
    assert ctypes.sizeof(L.PrjView) == 16
    assert ctypes.sizeof(L.OsemData) == 48 and L.OsemData.pfUpdate.offset == 40
    assert ctypes.sizeof(L.Drf) == 24 and L.Drf.Psf.offset == 16
    assert ctypes.sizeof(L.DrfTab) == 72 and L.DrfTab.psDrfs.offset == 56


def test_drft_file_round_trip_through_the_library(tmp_path):
    import numpy as np
    import libirl_amd as L
    from libirl_amd import synth
    offs, psfs = synth.gaussian_drf_table(0.4, 0.2, 5, 0.4, 3)
    f = str(tmp_path / "t.drft")
    synth.write_drft(f, 0.4, 0.4, 0.2, offs, psfs)
    lib = L.load()
    lib.sReadDrfTab.restype = L.DrfTab
    lib.sReadDrfTab.argtypes = [ctypes.c_char_p, ctypes.c_void_p, ctypes.c_void_p]
    t = lib.sReadDrfTab(f.encode(), None, None)
    assert t.iNumDistances == 5 and abs(t.fStartDistance - 0.2) < 1e-7
    for i in range(5):
        d = t.psDrfs[i]
        a = np.ctypeslib.as_array(d.Psf, shape=(2 * d.MaxOffsY - 1, 2 * d.MaxOffsX - 1))
        assert np.array_equal(a, psfs[i])
    # a byte-swapped file is detected from the version field (drftabio.c:192-209):
    # swap every 4-byte word except the signature and the two text fields
    raw = open(f, "rb").read()
    words = raw[104 + 4:]
    sw = raw[:4] + raw[4:8][::-1] + raw[8:108] + b"".join(
        words[k:k + 4][::-1] for k in range(0, len(words), 4))
    g = str(tmp_path / "s.drft")
    open(g, "wb").write(sw)
    t2 = lib.sReadDrfTab(g.encode(), None, None)
    assert t2.iNumDistances == 5
    d = t2.psDrfs[2]
    a = np.ctypeslib.as_array(d.Psf, shape=(2 * d.MaxOffsY - 1, 2 * d.MaxOffsX - 1))
    assert np.array_equal(a, psfs[2])


def test_mode_setters_return_the_previous_setting():
    # host-side switches for reconstructions created afterwards: no GPU work
    import libirl_amd as L
    if not os.path.exists(L.LIB_PATH):
        pytest.skip("library not built")
    lib = L.load()
    fused = L.set_esse_fused(False)
    assert L.set_esse_fused(fused) is False
    prev = L.set_drf_fft("fft")
    assert prev in ("reference", "fft", "auto", "spatial")
    assert L.set_drf_fft("auto") == "fft"
    # an unknown mode is refused and leaves the setting as it was
    assert lib.irlamd_set_drf_fft(5) == -2
    assert L.set_drf_fft(prev) == "auto"
    with pytest.raises(ValueError):
        L.set_drf_fft("fastest")


def test_write_drft_round_trip_is_byte_identical(tmp_path):
    # vWriteDrfTab (drftabio.c:57-148) of a table read with sReadDrfTab gives
    # back the file the generator wrote field by field in the DRFT layout
    import numpy as np
    import libirl_amd as L
    from libirl_amd import synth
    if not os.path.exists(L.LIB_PATH):
        pytest.skip("library not built")
    offs, psfs = synth.gaussian_drf_table(0.33, 0.165, 7, 0.33, 4)
    f = str(tmp_path / "in.drft")
    synth.write_drft(f, 0.33, 0.33, 0.165, offs, psfs, title="round trip", col="col-x")
    lib = L.load()
    lib.sReadDrfTab.restype = L.DrfTab
    lib.sReadDrfTab.argtypes = [ctypes.c_char_p, ctypes.c_void_p, ctypes.c_void_p]
    lib.vWriteDrfTab.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_char_p,
                                 ctypes.c_char_p]
    lib.vWriteDrfTab.restype = None
    t = lib.sReadDrfTab(f.encode(), None, None)
    g = str(tmp_path / "out.drft")
    lib.vWriteDrfTab(ctypes.byref(t), g.encode(), b"round trip", b"col-x")
    assert open(g, "rb").read() == open(f, "rb").read()
    # a table built in memory, written and read back
    tab = L.make_drftab(0.165, 0.33, offs, psfs, pixsize=0.33)
    h = str(tmp_path / "mem.drft")
    lib.vWriteDrfTab(ctypes.cast(ctypes.pointer(tab), ctypes.c_void_p), h.encode(), b"t", b"c")
    back = lib.sReadDrfTab(h.encode(), None, None)
    assert back.iNumDistances == 7 and back.fPixSize == np.float32(0.33)
    for i in range(7):
        d = back.psDrfs[i]
        a = np.ctypeslib.as_array(d.Psf, shape=(2 * d.MaxOffsY - 1, 2 * d.MaxOffsX - 1))
        assert np.array_equal(a, psfs[i])
    # an unwritable path is reported with the reference's message
    bad = str(tmp_path / "no" / "dir" / "x.drft")
    lib.vWriteDrfTab(ctypes.byref(t), bad.encode(), b"t", b"c")
    assert L.error_string() == "Error opening DRF Table file: " + bad
