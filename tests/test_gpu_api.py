"""The library API (include/gail_amd.h) on the GPU, driven by tests/native/test_api (C compiled
against the public header) with the real HIP plugin beside libgail.so: pull (send / receive),
callback, blocking encodes from several threads, finish, cancel and the end of the stream, as
the reference's tests/test_api.c:106-350 drives gail.h. Every input format the API takes goes
through gail_frame -> fill -> the GPU colour kernel, letterboxed pictures and every chroma
siting included, and each codestream is checked end to end: the oracle's colour conversion
and Tier-1 of the same pixels through the host rate control and Tier-2 give the same bytes.

With one frame in flight the early-termination threshold of frame k is 0.25 x the lambda of
frame k-1 (gail_api.cpp:607, pipeline.cpp:194-218), which the check replays."""
import os
import subprocess

import numpy as np
import pytest

import libgail_amd as g

pytestmark = pytest.mark.gpu

NATIVE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "native")
EXE = os.path.join(NATIVE, "test_api")
W2K, H2K = 1998, 1080
KR = {0: (0.2126, 0.0722), 1: (0.2627, 0.0593), 2: (0.299, 0.114)}
# name: (bits, sx, sy, rgb); bits 0 = interleaved 16-bit samples
FORMATS = {"rgb48": (0, 0, 0, 0), "xyz12": (0, 0, 0, 0), "xyz12_msb": (0, 0, 0, 0), "xyz12_planar": (12, 0, 0, 2),
           "yuv420p": (8, 1, 1, 0), "yuv420p10": (10, 1, 1, 0), "yuv420p12": (12, 1, 1, 0),
           "yuv422p": (8, 1, 0, 0), "yuv422p10": (10, 1, 0, 0), "yuv422p12": (12, 1, 0, 0),
           "yuv444p": (8, 0, 0, 0), "yuv444p10": (10, 0, 0, 0), "yuv444p12": (12, 0, 0, 0),
           "gbrp": (8, 0, 0, 1), "gbrp10": (10, 0, 0, 1), "gbrp12": (12, 0, 0, 1), "gbrp16": (16, 0, 0, 1)}


def _run(args, timeout=300):
    env = dict(os.environ)
    env["LD_LIBRARY_PATH"] = g.LIB_DIR
    print("[api] test_api", " ".join(str(a) for a in args), flush=True)
    p = subprocess.run([EXE] + [str(a) for a in args], env=env, capture_output=True, text=True, timeout=timeout)
    out = p.stdout + p.stderr
    print(out, flush=True)
    return p.returncode, out


def _codes(oracle, fmt, raw, pic, chroma, full_range, matrix):
    """The container's X'Y'Z' codes (3, H, W) the reference's CPU path makes of one input frame."""
    pw, ph = pic
    ox, oy = (W2K - pw) // 2, (H2K - ph) // 2
    bits, sx, sy, rgb = FORMATS[fmt]
    if fmt in ("xyz12", "xyz12_msb", "xyz12_planar"):
        if fmt == "xyz12_planar":
            v = np.frombuffer(raw, np.uint16).reshape(3, ph, pw).transpose(1, 2, 0)
        else:
            v = np.frombuffer(raw, np.uint16).reshape(ph, pw, 3)
            v = v >> 4 if fmt == "xyz12_msb" else v
        codes = np.zeros((3, H2K, W2K), np.uint16)
        codes[:, oy:oy + ph, ox:ox + pw] = np.minimum(v, 4095).transpose(2, 0, 1)
        return codes
    if fmt == "rgb48":
        rgb48 = np.zeros((H2K, W2K, 3), np.uint16)
        rgb48[oy:oy + ph, ox:ox + pw] = np.frombuffer(raw, np.uint16).reshape(ph, pw, 3)
    else:
        src = np.frombuffer(raw, np.uint8 if bits == 8 else np.uint16)
        kr, kb = KR[matrix]
        rgb48 = oracle.planar_to_rgb48(src, bits, pw, ph, sx, sy, rgb, 1 if rgb else full_range, chroma, kr, kb,
                                       W2K, H2K, ox, oy)
    lut, cp = oracle.color_tables(0, 0)  # gail_color_preset(SDR BT.709): BT.1886, BT.709 primaries
    _, xyz = oracle.convert_frame(lut, cp, rgb48)
    return np.ascontiguousarray(xyz.transpose(2, 0, 1))


CASES = [  # format, picture size, chroma siting, full range, matrix
    ("xyz12", (W2K, H2K), 0, 0, 0),
    ("xyz12_msb", (1920, 1080), 0, 0, 0),
    ("xyz12_planar", (1920, 800), 0, 0, 0),
    ("rgb48", (W2K, H2K), 0, 0, 0),
    ("rgb48", (1917, 803), 0, 0, 0),
    ("yuv420p", (1920, 1080), 0, 0, 0),
    ("yuv420p10", (1917, 803), 1, 0, 0),
    ("yuv420p12", (1920, 800), 2, 1, 1),
    ("yuv422p10", (1920, 1080), 1, 0, 2),
    ("yuv422p", (1917, 803), 0, 1, 0),
    ("yuv422p12", (W2K, H2K), 2, 0, 1),
    ("yuv444p", (1920, 1080), 0, 0, 0),
    ("yuv444p10", (1917, 803), 0, 1, 2),
    ("yuv444p12", (W2K, H2K), 0, 0, 0),
    ("gbrp", (1920, 1080), 0, 0, 0),
    ("gbrp10", (1917, 803), 0, 0, 0),
    ("gbrp16", (W2K, H2K), 0, 0, 0),
]


@pytest.mark.parametrize("fmt,pic,chroma,full_range,matrix", CASES,
                         ids=[f"{c[0]}-{c[1][0]}x{c[1][1]}-c{c[2]}f{c[3]}m{c[4]}" for c in CASES])
def test_pull_matches_oracle(tmp_path, oracle, fmt, pic, chroma, full_range, matrix):
    frames = 2
    rc, out = _run(["pull", fmt, frames, tmp_path, pic[0], pic[1], 1, chroma, full_range, matrix])
    assert rc == 0, out
    layout = g.FrameLayout(W2K, H2K, 0)
    lam_prev = 0.0
    for k in range(frames):
        raw = (tmp_path / f"in_{k:03d}.bin").read_bytes()
        got = (tmp_path / f"cs_{k:03d}.bin").read_bytes()
        codes = _codes(oracle, fmt, raw, pic, chroma, full_range, matrix)
        thr = 0.25 * lam_prev
        t1 = oracle.encode_frame_xyz(codes, 0, thr=thr, threads=16)
        cs, lam, _ = g.encode_codestream(layout, t1, fps=24.0, mbps=250.0, tag="hip")
        if thr > 0 and lam < thr:  # coded again without early termination
            t1 = oracle.encode_frame_xyz(codes, 0, thr=0.0, threads=16)
            cs, lam, _ = g.encode_codestream(layout, t1, fps=24.0, mbps=250.0, tag="hip")
        print(f"[api] {fmt} frame {k}: threshold {thr:.2f}, lambda {lam:.2f}, {len(got)} bytes, "
              f"oracle path {len(cs)} bytes, equal {cs == got}", flush=True)
        assert g.check_codestream(got, 24.0, False) == ""
        assert cs == got, f"{fmt} frame {k}: the codestream differs from the oracle path"
        lam_prev = lam


@pytest.mark.parametrize("args", [["callback", 6], ["threads", 2], ["threads", 1], ["cancel"], ["config"],
                                  ["pull", "xyz12", 8, "-", W2K, H2K, 3]])
def test_api_cases(args):
    rc, out = _run(args)
    assert rc == 0, out


def test_create_reports_the_device():
    rc, out = _run(["create", 0])
    assert rc == 0, out
    assert "1 device(s)" in out or "device(s):" in out
