"""The served form (gail_server_*) on the CPU with the stand-in plugin (its batches code on the
CPU and it models the device's memory): jobs of 2K and 4K slot classes made as frames of
their class wait, idle jobs retired to make room for the other class and, one at a time, when
another process takes the device's memory; frames that no job could ever hold failed with
GAIL_ERROR_NOMEM; a device lost mid-stream; all of it again under ThreadSanitizer and
AddressSanitizer. Then the frames-in-flight rule against the reference's known answers
(tests/test_main.cpp:1067-1080), and the Python wrapper's lifetime: a Server collected without
close() while frames are in flight, and a failing submit whose report reads the frame's
pinned planes, must not crash (the callback thunk lives until the server's threads are
gone)."""
import ctypes as C
import os
import subprocess
import sys

import pytest

import libgail_amd as g

NATIVE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "native")
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _run_server_case(variant, case):
    exe = os.path.join(NATIVE, "test_api") if variant == "ok" else os.path.join(NATIVE, "stub", variant, "test_api")
    env = dict(os.environ, LD_LIBRARY_PATH=os.path.join(NATIVE, "stub", variant),
               TSAN_OPTIONS="halt_on_error=1 second_deadlock_stack=1", ASAN_OPTIONS="detect_leaks=1")
    p = subprocess.run([exe, "server", case], env=env, capture_output=True, text=True, timeout=600)
    out = p.stdout + p.stderr
    print(out)
    return p.returncode, out


@pytest.mark.parametrize("variant", ["ok", "tsan", "asan"])
@pytest.mark.parametrize("case", ["mixed", "classes", "pressure", "nomem", "lost"])
def test_server_admission(variant, case):
    rc, out = _run_server_case(variant, case)
    assert "WARNING: ThreadSanitizer" not in out and "ERROR: AddressSanitizer" not in out, out
    assert rc == 0, out


def test_frames_in_flight_rule():
    """memoryHeadroom / fitSlots of pipeline.cpp:19-34 with the reference's own cases."""
    h = g.host()
    h.gail_memory_headroom.restype = C.c_uint64
    h.gail_memory_headroom.argtypes = [C.c_uint64]
    h.gail_fit_frames.restype = C.c_int
    h.gail_fit_frames.argtypes = [C.c_uint64, C.c_uint64, C.c_uint64, C.c_int, C.c_uint64]
    G, M = 1 << 30, 1 << 20
    assert h.gail_memory_headroom(8 * G) == 512 * M and h.gail_memory_headroom(24 * G) == 24 * G * 6 // 100
    fit = h.gail_fit_frames
    assert fit(0, 0, 700 * M, 4, 0) == 4              # unknown free memory does not limit
    assert fit(6 * G, 8 * G, 0, 4, 0) == 4            # unknown slot size does not limit
    assert fit(6 * G, 8 * G, 700 * M, 4, 0) == 4      # 6 GB free holds four 4K frames
    assert fit(2 * G, 8 * G, 700 * M, 4, 0) == 1      # 2 GB less 512 MB headroom holds one
    assert fit(2700 * M, 8 * G, 700 * M, 4, 0) == 2   # headroom and slack
    assert fit(300 * M, 8 * G, 700 * M, 4, 0) == 1    # at least one is attempted
    assert fit(20 * G, 24 * G, 700 * M, 4, 1600 * M) == 2  # an explicit limit caps
    assert fit(0, 0, 700 * M, 4, 1600 * M) == 2       # a limit works without a budget
    assert fit(1 * G, 24 * G, 700 * M, 4, 16 * G) == 1     # free memory caps below the limit
    assert fit(6 * G, 8 * G, 700 * M, 0, 0) == 1      # zero requested means one
    assert h.gail_memory_headroom(288 * G) == 288 * G * 6 // 100


WRAPPER_SCRIPT = r'''
import faulthandler, gc, sys, time, traceback
faulthandler.enable()
sys.path.insert(0, sys.argv[1])
import numpy as np
import libgail_amd as g
cfg = g.stream_config("2k_flat", "xyz12_planar")
x = np.zeros((3, 1080, 1998), np.uint16)

# a server collected without close() while its frames are still on the "device"
srv = g.Server(slots_per_job=2, max_jobs=2, check=True, host_rate_control=True)
for i in range(8):
    srv.submit(i, 1, cfg, [x[0], x[1], x[2]])
del srv
gc.collect()

# a submit that fails while others are in flight, reported the way pytest reports a failure:
# the frame's arguments repr'd, here planes over a pinned buffer whose name is gone
srv = g.Server(slots_per_job=2, max_jobs=2, check=True, keep=True, host_rate_control=True)
buf = g.pinned_empty(x.nbytes)
planes = [buf[i * x[0].nbytes:(i + 1) * x[0].nbytes].view(np.uint16).reshape(1080, 1998) for i in range(3)]
del buf
for i in range(6):
    srv.submit(10 + i, 1, cfg, planes)
try:
    srv.submit(99, 3, g.stream_config("2k_scope", "yuv420p10", picture=(1920, 1080)), planes)
except g.ServerError:
    text = traceback.format_exc()
    assert "picture" in text
    reprs = [repr(p) for p in planes]
srv.close()
srv.close()  # twice is fine
assert all(srv.results[10 + i]["status"] == 0 for i in range(6))
gc.collect()
print("wrapper ok", len(srv.results))
'''


def test_wrapper_lifetimes():
    env = dict(os.environ, GAIL_AMD_LIB_DIR=os.path.join(NATIVE, "stub", "ok"), GAIL_STUB_DELAY_MS="40")
    p = subprocess.run([sys.executable, "-c", WRAPPER_SCRIPT, ROOT], env=env, capture_output=True, text=True,
                       timeout=300)
    print(p.stdout + p.stderr)
    assert p.returncode == 0 and "wrapper ok 6" in p.stdout, p.stdout + p.stderr
