"""Rate control on the GPU: its chosen passes, lambda and sizes equal the host rate control
run on the oracle's Tier-1 output of the same frame, so the codestreams are byte-identical;
frames with and without device rate control share a batch."""
import numpy as np
import pytest

import libgail_amd as g
from t1_compare import compare_tier1

pytestmark = pytest.mark.gpu

TAG = "hip"
COM_RESERVE = 9 - len(TAG)  # rate control sizes the COM marker at its longest device tag


def _check_against_host(out, rate, L, ref, fps, label):
    cs_dev = g.build_codestream(L, out, tag=TAG)
    cs_host, lam_host, _ = g.encode_codestream(L, ref, fps=fps, mbps=250.0, tag=TAG)
    print(f"[rc {label}] device {len(cs_dev)} B, lambda {rate['lambda']}, {rate['iterations']} measurements; "
          f"host {len(cs_host)} B, lambda {lam_host}", flush=True)
    assert len(cs_dev) == rate["total"] - COM_RESERVE
    assert min(rate["lambda"]) == lam_host
    assert cs_dev == cs_host, f"{label}: codestreams differ"
    assert g.check_codestream(cs_dev, fps, L.fourk) == ""


@pytest.mark.parametrize("name,fps,thr", [("2k_flat", 24.0, 0.0), ("4k_flat", 24.0, 1500.0), ("4k_scope", 48.0, 0.0)])
def test_device_rate_control_matches_host(device, oracle, name, fps, thr):
    W, H, k = g.CONTAINERS[name]
    L = g.FrameLayout(W, H, k)
    x = g.synth_frame(900 + int(fps), W, H)
    b = g.Batch(device, [L], [g.with_rate_control(g.xyz_settings(L), fps)])
    try:
        b.upload(0, x)
        b.set_slope(0, thr)
        outs, _ = b.encode()
        rate = b.rate(0)
    finally:
        b.close()
    ref = oracle.encode_frame_xyz(x, k, thr=thr, threads=16)
    _check_against_host(outs[0], rate, L, ref, fps, name)


def test_rate_control_per_frame_in_a_batch(device, oracle):
    """One job: a 2K frame at 24 fps, a 4K scope frame at 48 fps (both rate-controlled on the
    device) and a 2K scope frame without: each gets its own result."""
    specs = [("2k_flat", 24.0), ("4k_scope", 48.0), ("2k_scope", None)]
    layouts, settings, frames = [], [], []
    for i, (n, fps) in enumerate(specs):
        W, H, k = g.CONTAINERS[n]
        L = g.FrameLayout(W, H, k)
        layouts.append(L)
        s = g.xyz_settings(L)
        settings.append(g.with_rate_control(s, fps) if fps else s)
        frames.append(g.synth_frame(950 + i, W, H))
    b = g.Batch(device, layouts, settings)
    try:
        for i, f in enumerate(frames):
            b.upload(i, f)
            b.set_slope(i, 0.0)
        outs, _ = b.encode()
        rates = [b.rate(i) if fps else None for i, (_, fps) in enumerate(specs)]
        with pytest.raises(g.GailError):
            b.rate(2)
    finally:
        b.close()
    for i, (n, fps) in enumerate(specs):
        k = g.CONTAINERS[n][2]
        ref = oracle.encode_frame_xyz(frames[i], k, threads=16)
        if fps:
            _check_against_host(outs[i], rates[i], layouts[i], ref, fps, f"batch {n}")
        else:
            bad, lines = compare_tier1(outs[i], ref)
            for ln in lines:
                print("[rc batch]", ln, flush=True)
            assert bad == 0
