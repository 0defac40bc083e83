"""The build's device flags keep packed-FP32 instructions out of the code.

hipcc's host (x86) pass reports the feature ``-packed-fp32-ops`` as unknown
and ignores it; the gfx950 device pass honours it.  A probe kernel doing
float2 arithmetic compiles, device-only, to ``v_pk_fma_f32`` without the
flag and to scalar FMAs with the build's flags.  (The built library is also
scanned after every build, ``build.check_no_packed_fp32``.)  CPU only: hipcc
cross-compiles without a GPU.
"""

from __future__ import annotations

import re
import shutil
import subprocess

import pytest

from conftest import nz_module

PROBE = r"""
#include <hip/hip_runtime.h>
__global__ void probe(const float2* __restrict__ a, const float2* __restrict__ b,
                      float2* __restrict__ c, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        float2 x = a[i], y = b[i];
        c[i] = make_float2(x.x * y.x + 1.5f, x.y * y.y + 1.5f);
    }
}
"""

PACKED = re.compile(r"\bv_pk_(add|mul|fma)_f32\b")


def _device_asm(tmp_path, flags):
    src = tmp_path / "probe.hip"
    src.write_text(PROBE)
    out = tmp_path / "probe.s"
    subprocess.run([nz_module("build").hipcc(), *flags, "--offload-device-only", "-S", "-o",
                    str(out), str(src)], check=True, capture_output=True)
    return out.read_text()


@pytest.mark.skipif(shutil.which("hipcc") is None and not __import__("os").path.exists(
    "/opt/rocm/bin/hipcc"), reason="hipcc is not installed")
def test_device_flags_exclude_packed_fp32(tmp_path):
    b = nz_module("build")
    plain = [f for f in b.HIPCC_FLAGS if f not in ("-Xclang", "-target-feature", "-packed-fp32-ops")]
    assert PACKED.search(_device_asm(tmp_path, plain)), "the probe no longer exercises packed FP32"
    assert not PACKED.search(_device_asm(tmp_path, b.HIPCC_FLAGS))
