"""The reference library's device interfaces (gail::Device, gail::EncoderSlot,
gail::DecoderSlot, src/core/backend.h:47-93) over the HIP plugin table, through the adapter
in libgail_amd/adapter, compiled with the reference's own sources (oracle/_ref/adapter_check):
on the CPU with stand-in plugins, the loader's refusal of another interface, slots made and a
frame coded with the layout's records, and DeviceOutOfMemory from a plugin out of memory.
The GPU test (test_gpu_adapter.py) checks the real plugin against the reference's CPU slots."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CHECK = os.path.join(ROOT, "oracle", "_ref", "adapter_check")
STUB = os.path.join(ROOT, "tests", "native", "stub")


def _run(variant, mode, env=None):
    if not os.path.exists(CHECK):
        pytest.skip("oracle/_ref/adapter_check is not built (make -C oracle ref, with the reference present)")
    e = dict(os.environ)
    for k in ("GAIL_STUB_FAIL_AT", "GAIL_STUB_SLOTS", "GAIL_STUB_DELAY_MS"):
        e.pop(k, None)
    e.update(env or {})
    p = subprocess.run([CHECK, os.path.join(STUB, variant), mode], env=e, capture_output=True, text=True, timeout=300)
    print(p.stdout + p.stderr)
    return p.returncode, p.stdout + p.stderr


def test_adapter_loads_and_names_the_device():
    rc, out = _run("ok", "load")
    assert rc == 0 and 'tag "hip0.0"' in out and "(HIP)" in out, out


@pytest.mark.parametrize("variant", ["other_abi", "declines", "bad_size", "bad_layout", "no_entry"])
def test_adapter_refuses_other_interfaces(variant):
    rc, out = _run(variant, "load")
    assert rc == 1 and "no device" in out, out


def test_adapter_slot_codes_a_frame():
    rc, out = _run("ok", "shape")
    assert rc == 0 and "6708 records" in out, out


def test_adapter_out_of_memory_is_soft():
    rc, out = _run("ok", "nomem", env={"GAIL_STUB_SLOTS": "0"})
    assert rc == 0 and "DeviceOutOfMemory" in out, out
