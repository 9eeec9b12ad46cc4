"""Build recipe for the CPU oracle (test infrastructure only).

Compiles oracle/irl_oracle.c twice with gcc into oracle/_build/:
  liborc_f32.so  -DREAL=float   (reference-semantics CPU path)
  liborc_f64.so  -DREAL=double  (FP64 oracle)
"""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "irl_oracle.c")
OUT = os.path.join(HERE, "_build")

# -ffp-contract=off keeps every float multiply and add separately rounded,
# as the reference's plain C compiles on x86-64 (no FMA in the base ISA).
CFLAGS = ["-O2", "-fPIC", "-shared", "-ffp-contract=off", "-fno-fast-math",
          "-std=gnu99", "-Wall", "-Wno-unused-function"]


def lib_path(real):
    return os.path.join(OUT, "liborc_%s.so" % real)


def build(force=True):
    os.makedirs(OUT, exist_ok=True)
    for real, ctype in (("f32", "float"), ("f64", "double")):
        out = lib_path(real)
        if not force and os.path.exists(out) and os.path.getmtime(out) >= os.path.getmtime(SRC):
            continue
        tmp = out + ".tmp"
        subprocess.run(["gcc"] + CFLAGS + ["-DREAL=%s" % ctype, SRC, "-o", tmp, "-lm"],
                       check=True)
        os.replace(tmp, out)


if __name__ == "__main__":
    build()
