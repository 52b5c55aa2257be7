"""libgail.so's plugin loader and frame loop on the CPU, with stand-in plugins
(tests/native/stub_plugin.cpp) in place of libgail_hip.so: the loader's refusals (another
ABI, a declined ABI, a wrong table size or layout hash, no entry point, no plugin, no
driver), no device memory, fewer frame slots than asked, a device lost mid-stream, and the
C API's pull / callback / blocking / cancel cases (tests/native/test_api.c). The same cases
run under ThreadSanitizer and AddressSanitizer + UBSan builds of libgail.so. Follows the
reference's loader and fault tests (tests/test_main.cpp:130-314, 1157-1361) and its C API
test (tests/test_api.c:106-350)."""
import os
import subprocess

import pytest

NATIVE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "native")
GAIL_ERROR_UNSUPPORTED, GAIL_ERROR_NOMEM = -2, -3


def _run(variant, args, env=None, timeout=240, exe=None):
    """Runs test_api against stub/<variant> (libgail.so + stand-in plugin); returns (rc, output)."""
    exe = exe or os.path.join(NATIVE, "test_api")
    if not os.path.exists(exe):
        pytest.fail(f"{exe} is not built: run make -C tests/native")
    e = dict(os.environ)
    e["LD_LIBRARY_PATH"] = os.path.join(NATIVE, "stub", variant)
    e["TSAN_OPTIONS"] = "halt_on_error=1 second_deadlock_stack=1"
    e["ASAN_OPTIONS"] = "detect_leaks=1"
    e["UBSAN_OPTIONS"] = "halt_on_error=1 print_stacktrace=1"
    for k in ("GAIL_STUB_FAIL_AT", "GAIL_STUB_SLOTS", "GAIL_STUB_DELAY_MS"):
        e.pop(k, None)
    e.update(env or {})
    p = subprocess.run([exe] + [str(a) for a in args], env=e, capture_output=True, text=True, timeout=timeout)
    out = p.stdout + p.stderr
    print(out)
    return p.returncode, out


@pytest.mark.parametrize("variant,message", [
    ("other_abi", "plugin ABI 3, this library 2"),
    ("declines", "does not support plugin ABI"),
    ("bad_size", "plugin table of"),
    ("bad_layout", "structure layout differs"),
    ("no_entry", "no gail_hip_plugin_get entry point"),
    ("missing", "HIP backend plugin not loaded"),
    ("no_driver", "HIP runtime unavailable: no AMD GPU driver"),
])
def test_loader_refuses(variant, message):
    rc, out = _run(variant, ["create", GAIL_ERROR_UNSUPPORTED, message])
    assert rc == 0, out
    assert "0 device(s)" in out or variant == "no_driver"


def test_no_device_memory_for_a_slot():
    rc, out = _run("ok", ["create", GAIL_ERROR_NOMEM, "out of device memory"], env={"GAIL_STUB_SLOTS": "0"})
    assert rc == 0, out


def test_fewer_slots_than_asked_still_codes():
    rc, out = _run("ok", ["create", 0, "", 3], env={"GAIL_STUB_SLOTS": "1"})
    assert rc == 0, out


@pytest.mark.parametrize("args", [["pull", "xyz12", 6, "-"], ["pull", "yuv420p10", 3, "-", 1920, 800],
                                  ["pull", "gbrp16", 2, "-"], ["callback", 6], ["threads", 2], ["threads", 1],
                                  ["cancel"], ["config"]])
def test_api_cases(args):
    rc, out = _run("ok", args)
    assert rc == 0, out


@pytest.mark.parametrize("fail_at,delay", [(1, 0), (4, 0), (7, 20)])
def test_device_lost_wakes_every_caller(fail_at, delay):
    """The device fails at the n-th frame: every blocked encode, send and receive returns
    GAIL_ERROR_FAILED with the device's message; none hangs (the watchdog would end it)."""
    rc, out = _run("ok", ["lost"], env={"GAIL_STUB_FAIL_AT": str(fail_at), "GAIL_STUB_DELAY_MS": str(delay)})
    assert rc == 0, out


@pytest.mark.parametrize("san", ["tsan", "asan"])
@pytest.mark.parametrize("args,env", [
    (["pull", "xyz12", 4, "-"], {}),
    (["callback", 5], {}),
    (["threads", 3], {"GAIL_STUB_DELAY_MS": "5"}),
    (["cancel"], {"GAIL_STUB_DELAY_MS": "10"}),
    (["lost"], {"GAIL_STUB_FAIL_AT": "6", "GAIL_STUB_DELAY_MS": "20"}),
    (["config"], {}),
])
def test_sanitized(san, args, env):
    rc, out = _run(san, args, env=env, timeout=600, exe=os.path.join(NATIVE, "stub", san, "test_api"))
    assert "WARNING: ThreadSanitizer" not in out, out
    assert "ERROR: AddressSanitizer" not in out and "runtime error:" not in out, out
    assert rc == 0, out
