"""Build recipe for libirl_amd.so (gfx950).

HIP sources are compiled with hipcc for gfx950 only; host C++ with g++.
Everything is linked into libirl_amd/libirl_amd.so, in-tree.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libirl_amd.so")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.path.join(ROCM, "bin", "hipcc")

HIP_SOURCES = ["engine.hip", "esse.hip", "cnn.hip", "drf_fft.hip"]
CPP_SOURCES = ["host_util.cpp", "irlosem.cpp", "prjbck_api.cpp", "abi_check.cpp", "esse_host.cpp",
               "orbit.cpp", "grf.cpp"]

# -ffp-contract=off: every float multiply/add in the geometry and
# interpolation code is rounded separately, as in the reference; the
# convolution kernels request fused multiply-adds explicitly.
HIP_FLAGS = ["-O3", "-fPIC", "-std=c++17", "--offload-arch=gfx950", "-ffp-contract=off",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wno-unused-result"]
CPP_FLAGS = ["-O2", "-fPIC", "-std=c++17", "-ffp-contract=off", "-Wall",
             "-I" + os.path.join(ROCM, "include")]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout)
        raise RuntimeError("build step failed: " + " ".join(cmd))
    return r.stdout


def build(jobs=None):
    os.makedirs(OBJ, exist_ok=True)
    objs = []
    procs = []
    for src in HIP_SOURCES:
        o = os.path.join(OBJ, src + ".o")
        procs.append(([HIPCC] + HIP_FLAGS + ["-c", os.path.join(CSRC, src), "-o", o], o))
    for src in CPP_SOURCES:
        o = os.path.join(OBJ, src + ".o")
        procs.append((["g++"] + CPP_FLAGS + ["-c", os.path.join(CSRC, src), "-o", o], o))
    running = [(subprocess.Popen(c, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True), c, o)
               for c, o in procs]
    failed = False
    for p, c, o in running:
        out, _ = p.communicate()
        if p.returncode != 0:
            sys.stderr.write(out)
            failed = True
        objs.append(o)
    if failed:
        raise RuntimeError("libirl_amd compilation failed")
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", "--offload-arch=gfx950", "-o", tmp] + objs +
         ["-L" + os.path.join(ROCM, "lib"), "-lamdhip64"])
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build())
