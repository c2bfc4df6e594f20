"""Build libmumag_hip.so for gfx950 in-tree with hipcc.

``python -m mammos_mumag_amd.build`` (or ``build()``) always recompiles.
"""
import os
import pathlib
import shutil
import subprocess
import sys

PKG = pathlib.Path(__file__).resolve().parent
CSRC = PKG.parent / "csrc"
INCLUDE = PKG.parent.parent / "include"
LIBDIR = PKG / "_lib"
LIB = LIBDIR / "libmumag_hip.so"

SOURCES = ["kernels.hip", "pcg_fused.hip", "mlp.hip", "mg.hip", "amg.hip", "assemble.cpp", "structured.cpp",
           "amg.cpp", "capi.cpp"]
ARCH = os.environ.get("MUMAG_OFFLOAD_ARCH", "gfx950")


def hipcc():
    for cand in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found: the HIP library cannot be built")


def flags():
    return [
        f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17", "-fopenmp",
        "-munsafe-fp-atomics", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        f"-I{INCLUDE}", "-Wall", "-Wno-unused-result",
    ]


def build(verbose=False, jobs=None):
    LIBDIR.mkdir(exist_ok=True)
    objdir = LIBDIR / "obj"
    objdir.mkdir(exist_ok=True)
    cc = hipcc()
    procs = []
    objs = []
    for src in SOURCES:
        obj = objdir / (src + ".o")
        lang = ["-x", "hip"] if src.endswith(".cpp") else []
        cmd = [cc, *flags(), *lang, "-c", str(CSRC / src), "-o", str(obj)]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
        objs.append(str(obj))
    failed = []
    for src, p in procs:
        out, _ = p.communicate()
        if p.returncode != 0:
            failed.append((src, out.decode(errors="replace")))
        elif verbose and out:
            print(out.decode(errors="replace"))
    if failed:
        raise RuntimeError("\n".join(f"--- {s}\n{o}" for s, o in failed))
    rpath = "/opt/rocm/lib/llvm/lib"
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [cc, f"--offload-arch={ARCH}", "-shared", "-fopenmp", *objs, "-o", str(tmp),
           f"-Wl,-rpath,{rpath}", "-Wl,-rpath,/opt/rocm/lib"]
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
