"""Build the gfx950 library amsc_amd/_lib/libamsc_mpm.so (and the amsc_run
driver) with hipcc, in-tree.

    python -m amsc_amd.build            # or amsc_amd.build.build()

FP64 parity needs IEEE arithmetic without FMA contraction
(-ffp-contract=off, no fast-math).  Packed FP32 instructions are disabled
(-packed-fp32-ops).
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
INCLUDE = os.path.join(ROOT, "include")
OUT_DIR = os.path.join(HERE, "_lib")
OBJ_DIR = os.path.join(OUT_DIR, "obj")
LIB = os.path.join(OUT_DIR, "libamsc_mpm.so")
RUN = os.path.join(OUT_DIR, "amsc_run")
OFFLOAD_LOOP = os.path.join(OUT_DIR, "amsc_offload_loop")

HIP_SOURCES = ["runtime.hip", "dropin.hip", "binning.hip", "merge_split.hip",
               "sim.hip"]
CXX_SOURCES = ["data_json.cpp", "output.cpp"]
ARCH = os.environ.get("AMSC_OFFLOAD_ARCH", "gfx950")

HIPCC_FLAGS = [
    "-O3", f"--offload-arch={ARCH}", "-fPIC", "-std=c++17",
    "-ffp-contract=off", "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-result",
    "-I", INCLUDE, "-I", CSRC,
]
CXX_FLAGS = ["-O2", "-fPIC", "-std=c++17", "-Wall", "-ffp-contract=off",
             "-I", INCLUDE]


def _hipcc():
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found: the gfx950 library cannot be built")


def build(verbose: bool = False) -> str:
    """Compile every source (always) and link the library; returns its path."""
    os.makedirs(OBJ_DIR, exist_ok=True)
    hipcc = _hipcc()
    procs = []
    objs = []
    for src in HIP_SOURCES:
        obj = os.path.join(OBJ_DIR, src.replace(".hip", ".o"))
        cmd = [hipcc, *HIPCC_FLAGS, "-x", "hip", "-c", os.path.join(CSRC, src), "-o", obj]
        procs.append((src, cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE,
                                                  stderr=subprocess.STDOUT)))
        objs.append(obj)
    for src in CXX_SOURCES:
        obj = os.path.join(OBJ_DIR, src.replace(".cpp", ".o"))
        cmd = ["g++", *CXX_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
        procs.append((src, cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE,
                                                  stderr=subprocess.STDOUT)))
        objs.append(obj)
    failed = []
    for src, cmd, p in procs:
        out, _ = p.communicate()
        text = out.decode(errors="replace")
        # hipcc's host pass does not know the packed-fp32 feature; harmless
        text = "\n".join(l for l in text.splitlines()
                         if "packed-fp32-ops" not in l and l.strip())
        if p.returncode != 0:
            failed.append(f"{src}: {' '.join(cmd)}\n{text}")
        elif verbose and text:
            print(f"[{src}]\n{text}")
    if failed:
        raise RuntimeError("build failed:\n" + "\n".join(failed))
    tmp = LIB + ".tmp"
    subprocess.run([hipcc, f"--offload-arch={ARCH}", "-shared", "-o", tmp, *objs],
                   check=True)
    os.replace(tmp, LIB)
    for src, exe in (("amsc_run.cpp", RUN), ("offload_loop.cpp", OFFLOAD_LOOP)):
        subprocess.run([hipcc, f"--offload-arch={ARCH}", "-O2", "-std=c++17", "-I", INCLUDE,
                        os.path.join(CSRC, src), "-o", exe,
                        f"-L{OUT_DIR}", "-lamsc_mpm", f"-Wl,-rpath,{OUT_DIR}",
                        "-Wl,-rpath,$ORIGIN"],
                       check=True)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
