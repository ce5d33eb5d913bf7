"""Build recipe for the oracle (test infrastructure): oracle/liblbdem_oracle.so.

The C restatement is compiled twice, REAL=double (suffix _f64, the parity oracle) and
REAL=float (suffix _f32, the CPU timing baseline), with OpenMP and without floating-point
contraction so that integer keys derived from floats match the device code bit for bit.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liblbdem_oracle.so")
SRCS = ["lbm_oracle.c"]          # compiled for REAL = double and REAL = float
SRCS_F64 = ["dem_oracle.c"]      # FP64 only
CFLAGS = ["-O2", "-fPIC", "-fopenmp", "-ffp-contract=off", "-fno-fast-math", "-march=x86-64-v2",
          "-std=c11", "-Wall", "-Wno-unused-function"]


def build(verbose=True):
    objs = []
    bdir = os.path.join(HERE, "_build")
    os.makedirs(bdir, exist_ok=True)
    for src in SRCS:
        path = os.path.join(HERE, src)
        if not os.path.exists(path):
            continue
        for real, sfx in (("double", "_f64"), ("float", "_f32")):
            obj = os.path.join(bdir, f"{src}{sfx}.o")
            defs = [f"-DREAL={real}", f"-DSFX(name)=name##{sfx}"]
            r = subprocess.run(["gcc", *CFLAGS, *defs, "-c", path, "-o", obj], capture_output=True, text=True)
            if r.returncode != 0:
                raise RuntimeError(f"gcc failed for {src} ({real}):\n{r.stderr}")
            objs.append(obj)
    for src in SRCS_F64:
        obj = os.path.join(bdir, f"{src}.o")
        r = subprocess.run(["gcc", *CFLAGS, "-D_GNU_SOURCE", "-c", os.path.join(HERE, src), "-o", obj],
                           capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"gcc failed for {src}:\n{r.stderr}")
        objs.append(obj)
    r = subprocess.run(["gcc", "-shared", "-fopenmp", "-o", LIB + ".tmp", *objs, "-lm"], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"oracle link failed:\n{r.stderr}")
    os.replace(LIB + ".tmp", LIB)
    if verbose:
        print(f"built {LIB}", file=sys.stderr)
    return LIB


if __name__ == "__main__":
    build()
