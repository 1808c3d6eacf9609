"""Builds the HIP library librdsphp.so for gfx950, in-tree.

Each csrc/*.hip file is compiled on its own (in parallel) and the objects are
linked into redondito_amd/_lib/librdsphp.so.  The packed-FP32 feature is
switched off for every device compile.
"""
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT_DIR = os.path.join(HERE, "_lib")
LIB = os.path.join(OUT_DIR, "librdsphp.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

CFLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result", "-Wno-unused-parameter",
]


def sources():
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hip"))


def _compile(src):
    obj = os.path.join(OUT_DIR, os.path.basename(src)[:-4] + ".o")
    cmd = [HIPCC] + CFLAGS + ["-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr}")
    return obj


FDIR = os.path.join(HERE, "fortran")
FLIB = os.path.join(OUT_DIR, "librdsphp_f.so")
FTEST = os.path.join(OUT_DIR, "prueba_rd")
FC = os.environ.get("FC_AMD", "/opt/rocm/bin/amdflang")


def _run(cmd, what):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"{what} failed:\n{' '.join(cmd)}\n{r.stderr}")


def build_fortran(verbose=True):
    """The Fortran drop-in modules (librdsphp_f.so, .mod files in _lib/mod)
    and the known-answer program prueba_rd, linked to librdsphp.so."""
    if not os.path.exists(FC):
        if verbose:
            print("amdflang not found: Fortran drop-in layer not built")
        return None
    mod = os.path.join(OUT_DIR, "mod")
    os.makedirs(mod, exist_ok=True)
    objs = []
    for src in ("rd_fortran_mods.f90", "rd_sphp.f90"):
        obj = os.path.join(OUT_DIR, src[:-4] + ".o")
        _run([FC, "-O2", "-fPIC", "-module-dir", mod, "-c", os.path.join(FDIR, src), "-o", obj], f"amdflang {src}")
        objs.append(obj)
    _run([FC, "-shared", "-fPIC", "-o", FLIB] + objs + ["-L" + OUT_DIR, "-lrdsphp", "-Wl,-rpath,$ORIGIN"],
         "link librdsphp_f.so")
    _run([FC, "-O2", "-I", mod, os.path.join(FDIR, "prueba_rd.f90"), "-o", FTEST,
          "-L" + OUT_DIR, "-lrdsphp_f", "-lrdsphp", "-Wl,-rpath,$ORIGIN"], "build prueba_rd")
    for o in objs:
        os.remove(o)
    if verbose:
        print(f"built {FLIB} and {FTEST}")
    return FLIB


def build(verbose=True):
    os.makedirs(OUT_DIR, exist_ok=True)
    srcs = sources()
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", LIB + ".tmp"] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr}")
    os.replace(LIB + ".tmp", LIB)
    for o in objs:
        os.remove(o)
    if verbose:
        print(f"built {LIB}")
    build_fortran(verbose)
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
