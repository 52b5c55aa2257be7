"""The HIP plugin driven through the reference library's own device interfaces: the adapter
(libgail_amd/adapter/hip_host.cpp) compiled with the reference's sources into
oracle/_ref/adapter_check. A 2K flat X'Y'Z' frame handed over as the reference's API does
(RGB48, 16-bit widened codes) gives, through HipEncoderSlot, the same Tier-1 output block for
block as the reference's own CPU encoder slot (cpu_backend.cpp), with early termination off
and on; decoded through HipDecoderSlot it gives the same picture as the reference's CPU
decode slot, at full and quarter size."""
import os
import subprocess

import pytest

import libgail_amd as g

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CHECK = os.path.join(ROOT, "oracle", "_ref", "adapter_check")


@pytest.mark.parametrize("mode", ["load", "decode"])
def test_reference_interfaces_over_the_plugin(mode):
    if not os.path.exists(CHECK):
        pytest.skip("oracle/_ref/adapter_check is not built (make -C oracle ref, with the reference present)")
    print(f"[adapter] {mode}", flush=True)
    p = subprocess.run([CHECK, g.LIB_DIR, mode], capture_output=True, text=True, timeout=300)
    print(p.stdout + p.stderr, flush=True)
    assert p.returncode == 0, p.stdout + p.stderr
