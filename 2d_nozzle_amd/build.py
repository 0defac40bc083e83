"""Build ``libnozzle_hip.so`` in-tree for gfx950 (``hipcc``, no CMake).

    python 2d_nozzle_amd/build.py

The library is written to ``2d_nozzle_amd/_lib/libnozzle_hip.so``.  Every
call recompiles.
"""

from __future__ import annotations

import os
import re
import shutil
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SOURCES = [os.path.join(HERE, "csrc", n) for n in ("nozzle_hip.hip", "mesh_build.hip", "initial.hip",
                                                        "observables.hip")]
OUT = os.path.join(HERE, "_lib", "libnozzle_hip.so")

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    # No packed-FP32 instructions in device code.  The device pass honours
    # this feature; the host (x86) pass does not know it and prints "not a
    # recognized feature ... (ignoring feature)", which is harmless.
    # check_no_packed_fp32() verifies the built code object.
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]

LLVM = "/opt/rocm/llvm/bin"
PACKED_FP32 = re.compile(r"\bv_pk_(add|mul|fma)_f32\b")


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found (ROCm is required to build the HIP backend)")


def build(verbose: bool = False) -> str:
    """Compile every source to an object (in parallel), link the shared
    library, and check its device code for packed-FP32 instructions."""
    os.makedirs(os.path.dirname(OUT), exist_ok=True)
    inc = ["-I", os.path.join(ROOT, "include"), "-I", os.path.join(HERE, "csrc")]
    with tempfile.TemporaryDirectory() as td:
        objs, procs = [], []
        for src in SOURCES:
            obj = os.path.join(td, os.path.basename(src) + ".o")
            cmd = [hipcc(), *HIPCC_FLAGS, *inc, "-c", "-o", obj, src]
            if verbose:
                print(" ".join(cmd), flush=True)
            procs.append((cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE,
                                                stderr=subprocess.STDOUT, text=True)))
            objs.append(obj)
        failed = []
        for cmd, pr in procs:
            out, _ = pr.communicate()
            if pr.returncode != 0:
                sys.stderr.write(out)
                failed.append(cmd[-1])
        if failed:
            raise RuntimeError(f"hipcc failed on {failed}")
        tmp = OUT + ".tmp"
        cmd = [hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs]
        if verbose:
            print(" ".join(cmd), flush=True)
        res = subprocess.run(cmd, capture_output=True, text=True)
        if res.returncode != 0:
            sys.stderr.write(res.stdout + res.stderr)
            raise RuntimeError(f"link failed with status {res.returncode}")
    check_no_packed_fp32(tmp)
    os.replace(tmp, OUT)
    return OUT


def check_no_packed_fp32(lib: str) -> int:
    """Disassemble the gfx950 code object of ``lib`` (this library's own
    build) and raise if it holds any packed-FP32 add, mul or FMA.  Returns the
    number of instructions scanned."""
    with tempfile.TemporaryDirectory() as td:
        fat, co = os.path.join(td, "fatbin"), os.path.join(td, "gfx950.co")
        subprocess.run([os.path.join(LLVM, "llvm-objcopy"), f"--dump-section=.hip_fatbin={fat}",
                        lib, os.path.join(td, "stripped")], check=True, capture_output=True)
        subprocess.run([os.path.join(LLVM, "clang-offload-bundler"), "--unbundle", "--type=o",
                        "--targets=hipv4-amdgcn-amd-amdhsa--gfx950", f"--input={fat}",
                        f"--output={co}"], check=True, capture_output=True)
        dis = subprocess.run([os.path.join(LLVM, "llvm-objdump"), "-d", co], check=True,
                             capture_output=True, text=True).stdout
    bad = [ln.strip() for ln in dis.splitlines() if PACKED_FP32.search(ln)]
    if bad:
        raise RuntimeError(f"{len(bad)} packed-FP32 instructions in the device code, e.g. {bad[0]}")
    return sum(1 for ln in dis.splitlines() if ln.startswith("\t") or ln.startswith("  "))


if __name__ == "__main__":
    print(build(verbose=True))
