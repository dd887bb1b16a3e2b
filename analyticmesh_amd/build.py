"""In-tree build of the native engine.

    python -m analyticmesh_amd.build

compiles every ``csrc/*.hip`` for gfx950 with hipcc, links ``libam.so`` (the
C-ABI of ``include/am.h``) and builds the pybind11 module ``_am`` against it.
Outputs land next to this file so they travel with the source tree.
"""
import os
import subprocess
import sys
import sysconfig

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(CSRC, "_build")
LIB = os.path.join(HERE, "libam.so")
ROOT = os.path.dirname(HERE)

HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    "-ffp-contract=off",            # FP64 kernels must keep the written operation order
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]
HIP_SOURCES = ["am_region_fp64.hip", "am_chain64.hip", "am_hash.hip", "am_combine.hip", "am_fast.hip", "am_newton.hip",
               "am_adjoint.hip", "am_layout.hip", "am_engine.hip"]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    return subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)


def _wait(procs):
    failed = []
    for cmd, p in procs:
        out, _ = p.communicate()
        lines = [l for l in out.splitlines() if "packed-fp32-ops" not in l and l.strip()]
        if p.returncode != 0:
            failed.append((cmd, "\n".join(lines)))
        elif lines:
            print("\n".join(lines[-20:]))
    if failed:
        msg = "\n\n".join(f"$ {' '.join(c)}\n{o[-4000:]}" for c, o in failed)
        raise RuntimeError("native build failed:\n" + msg)


def build(verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    procs = []
    objs = []
    for src in HIP_SOURCES:
        obj = os.path.join(OBJ, src.replace(".hip", ".o"))
        objs.append(obj)
        cmd = ["hipcc"] + HIP_FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj]
        procs.append((cmd, _run(cmd, verbose)))
    _wait(procs)
    tmp = LIB + ".tmp"
    _wait([(c, _run(c, verbose)) for c in [["hipcc", "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs]])
    os.replace(tmp, LIB)

    import pybind11
    ext = sysconfig.get_config_var("EXT_SUFFIX") or ".so"
    mod = os.path.join(HERE, "_am" + ext)
    cmd = ["g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-Wall",
           "-I" + pybind11.get_include(), "-I" + sysconfig.get_paths()["include"],
           os.path.join(CSRC, "bind.cpp"), "-o", mod + ".tmp",
           "-L" + HERE, "-lam", "-Wl,-rpath,$ORIGIN"]
    # host-only polygon-mesh tools (no GPU code)
    pmod = os.path.join(HERE, "_polylib" + ext)
    pcmd = ["g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-Wall",
            "-I" + pybind11.get_include(), "-I" + sysconfig.get_paths()["include"],
            os.path.join(CSRC, "polylib.cpp"), "-o", pmod + ".tmp"]
    _wait([(cmd, _run(cmd, verbose)), (pcmd, _run(pcmd, verbose))])
    os.replace(mod + ".tmp", mod)
    os.replace(pmod + ".tmp", pmod)
    return LIB, mod, pmod


if __name__ == "__main__":
    for p in build(verbose="-v" in sys.argv):
        print(p)
