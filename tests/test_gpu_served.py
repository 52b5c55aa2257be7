"""Served form (gail_server_*): frames of several streams with different DCI sizes and input
formats, admitted into concurrent device jobs. Each frame's codestream is the same coded in
the mix and coded alone, and equals what the oracle's Tier-1 gives through the same rate
control and Tier-2 (bit-exact)."""
import numpy as np
import pytest

import libgail_amd as g
from t1_compare import planar_from_synth

pytestmark = pytest.mark.gpu

def _pic(container):
    """The R'G'B' / Y'CbCr picture size: 1920x1080 where it fits, centred in its container."""
    W, H, _ = g.CONTAINERS[container]
    return min(1920, W), min(1080, H)


def _requests():
    """(stream id, container, format, slope, seed) of the frames the test serves."""
    return [
        (1, "2k_flat", "xyz12_planar", 0.0, 11),
        (2, "4k_scope", "xyz12_planar", 80.0, 12),
        (3, "2k_scope", "gbrp12", 0.0, 13),
        (4, "4k_flat", "yuv420p10", 0.0, 14),
        (5, "2k_full", "xyz12_planar", 30.0, 15),
        (6, "4k_full", "xyz12_planar", 0.0, 16),
        (1, "2k_flat", "xyz12_planar", 20.0, 17),
        (3, "2k_scope", "gbrp12", 0.0, 18),
        (7, "2k_flat", "yuv420p10", 0.0, 19),
    ]


def _frame(container, fmt, seed, pinned):
    """The planes of one request: X'Y'Z' at the container size, or a 1920x1080 planar picture;
    with pinned, in one server-allocated pinned buffer (the DMA path), else separate arrays."""
    W, H, _ = g.CONTAINERS[container]
    if fmt == "xyz12_planar":
        x = g.synth_frame(seed, W, H)
        planes = [x[0], x[1], x[2]]
    else:
        rgb = fmt.startswith("gbrp")
        bits, sx, sy = (12, 0, 0) if rgb else (10, 1, 1)
        PIC = _pic(container)
        flat = planar_from_synth(g.synth_frame(seed, *PIC), bits, sx, sy, rgb)
        cw, ch = (PIC[0] + sx) >> sx, (PIC[1] + sy) >> sy
        n0 = PIC[0] * PIC[1]
        planes = [flat[:n0].reshape(PIC[1], PIC[0]), flat[n0:n0 + cw * ch].reshape(ch, cw),
                  flat[n0 + cw * ch:].reshape(ch, cw)]
    if not pinned:
        return [p.copy() for p in planes]  # separate allocations: the host-copy path
    total = sum(p.nbytes for p in planes)
    buf = g.pinned_empty(total)
    out, off = [], 0
    for p in planes:
        v = buf[off:off + p.nbytes].view(np.uint16).reshape(p.shape)
        v[...] = p
        out.append(v)
        off += p.nbytes
    return out


def _config(container, fmt):
    if fmt == "xyz12_planar":
        return g.stream_config(container, fmt)
    full = 1 if fmt.startswith("gbrp") else 0
    return g.stream_config(container, fmt, picture=_pic(container), matrix=0, full_range=full, chroma=0, preset=0)


def test_served_mix_equals_alone_and_oracle(device, oracle):
    # a 50 ms batching window, so that frames submitted together are coded together
    srv = g.Server(slots_per_job=3, max_jobs=3, check=True, keep=True, batch_window_us=50000)
    try:
        reqs = _requests()
        frames = [_frame(c, f, seed, pinned=(i % 2 == 0)) for i, (_, c, f, _, seed) in enumerate(reqs)]
        cfgs = [_config(c, f) for _, c, f, _, _ in reqs]
        print("[served] mix of", len(reqs), "frames", flush=True)
        for i, (sid, c, f, slope, _) in enumerate(reqs):
            srv.submit(i, sid, cfgs[i], frames[i], slope)
        srv.drain()
        mix = {i: srv.results[i] for i in range(len(reqs))}
        st = srv.stats()
        print("[served] stats", {k: st[k] for k in ("frames", "p50_ms", "p99_ms", "mean_batch", "jobs")}, flush=True)
        for i, r in mix.items():
            print(f"[served] frame {i} {reqs[i][1]} {reqs[i][2]}: status {r['status']} compliant {r.get('compliant')} "
                  f"size {r.get('size')} batch {r['batch']} latency {r['latency_ms']:.1f} ms", flush=True)
        for i, r in mix.items():
            assert r["status"] == 0, f"frame {i} failed"
            assert r["compliant"] == 1, f"frame {i} not compliant"
        assert st["mean_batch"] > 1.0, "frames were not coded together"
        print("[served] each frame alone", flush=True)
        for i, (sid, c, f, slope, _) in enumerate(reqs):
            srv.submit(100 + i, sid, cfgs[i], frames[i], slope)
            srv.drain()
            alone = srv.results[100 + i]
            same = alone.get("codestream") == mix[i].get("codestream")
            print(f"[served] frame {i} alone: status {alone['status']} batch {alone['batch']} same {same}", flush=True)
            assert alone["batch"] == 1
            assert same, f"frame {i} differs alone and in the mix"
    finally:
        srv.close()

    # end to end against the oracle: its Tier-1 through the same rate control and Tier-2
    for i in (0, 1, 3):
        sid, c, f, slope, seed = reqs[i]
        W, H, k = g.CONTAINERS[c]
        if f == "xyz12_planar":
            codes = g.synth_frame(seed, W, H)
        else:
            PIC = _pic(c)
            flat = planar_from_synth(g.synth_frame(seed, *PIC), 10, 1, 1, 0)
            ox, oy = (W - PIC[0]) // 2, (H - PIC[1]) // 2
            rgb48 = oracle.planar_to_rgb48(flat, 10, PIC[0], PIC[1], 1, 1, 0, 0, 0, 0.2126, 0.0722, W, H, ox, oy)
            lut, cp = oracle.color_tables(0, 0)
            _, xyz = oracle.convert_frame(lut, cp, rgb48)
            codes = np.ascontiguousarray(xyz.transpose(2, 0, 1))
        t1 = oracle.encode_frame_xyz(codes, k, thr=slope, threads=16)
        cs, _, _ = g.encode_codestream(g.FrameLayout(W, H, k), t1, fps=24.0, mbps=250.0, tag="hip")
        got = mix[i]["codestream"]
        print(f"[served] frame {i} vs oracle path: sizes {len(cs)} / {len(got)}, equal {cs == got}", flush=True)
        assert cs == got, f"frame {i} ({c}, {f}) differs from the oracle path"


def test_served_stream_feedback(device):
    """Frames of one stream with automatic early termination (0.25 x the stream's last
    lambda): every frame is delivered, compliant and under the byte cap."""
    srv = g.Server(slots_per_job=2, max_jobs=2, check=True)
    try:
        W, H, _ = g.CONTAINERS["2k_flat"]
        cfg = g.stream_config("2k_flat", "xyz12_planar")
        frames = [g.synth_frame(40 + i, W, H) for i in range(6)]
        for i, x in enumerate(frames):
            srv.submit(i, 9, cfg, [x[0], x[1], x[2]], -1.0)
        srv.drain()
        for i in range(6):
            r = srv.results[i]
            assert r["status"] == 0 and r["compliant"] == 1, (i, r)
            assert r["size"] <= 1302067
    finally:
        srv.close()


def test_served_device_resident_frames(device):
    """A frame resident in HBM (copied within HBM into its slot) codes to the same codestream
    as the same frame handed over from host memory; a device frame that is not in the device
    input layout is refused at submit (the host cannot rearrange device memory)."""
    W, H, _ = g.CONTAINERS["2k_flat"]
    x = g.synth_frame(88, W, H)
    cfg = g.stream_config("2k_flat", "xyz12_planar")
    dev_frame = g.DeviceFrame([x[0], x[1], x[2]])
    srv = g.Server(slots_per_job=2, max_jobs=1, check=True, keep=True)
    try:
        srv.submit(1, 1, cfg, [x[0].copy(), x[1].copy(), x[2].copy()], 35.0)
        srv.submit(2, 1, cfg, dev_frame, 35.0)
        srv.drain()
        a, b = srv.results[1], srv.results[2]
        print(f"[served] host frame {a['status']} {a.get('size')}, device frame {b['status']} {b.get('size')}", flush=True)
        assert a["status"] == 0 and b["status"] == 0
        assert a["codestream"] == b["codestream"]
        # planes not back to back: rejected
        bad = g.DeviceFrame([x[0], x[1], x[2]])
        bad.ptrs = [bad.ptrs[0], bad.ptrs[2], bad.ptrs[1]]
        with pytest.raises(g.ServerError):
            srv.submit(3, 1, cfg, bad, 35.0)
        bad.close()
    finally:
        srv.close()
        dev_frame.close()


def test_stress_4k_48fps_matches_oracle(device, oracle, ref):
    """The high-frame-rate stress stream: 4K flat 3996x2160 frames at 48 fps rate control (cap
    651,041 bytes, budget 651,025), PCRD truncation on every frame, twelve frames in flight in
    up to three device jobs with rate control on the GPU. Every codestream passes the reference's own
    DCI checker at 48 fps within the budget, and equals the oracle's Tier-1 of the same frame
    (at the frame's early-termination threshold) through the host rate control and Tier-2."""
    W, H, k = g.CONTAINERS["4k_flat"]
    L = g.FrameLayout(W, H, k)
    budget, comp_budget = g.dci_limits(48.0, 250.0, 200.0)
    assert budget == 651025
    cfg = g.stream_config("4k_flat", "xyz12_planar", fps=48.0, mbps=250.0)
    slopes = [0.0, 1800.0, 900.0, 2500.0] * 3
    frames = [g.synth_frame(7000 + i, W, H) for i in range(len(slopes))]
    srv = g.Server(slots_per_job=4, max_jobs=3, check=True, keep=True)
    try:
        print(f"[stress] {len(frames)} 4K flat frames at 48 fps", flush=True)
        for i, x in enumerate(frames):
            srv.submit(i, 48, cfg, [x[0], x[1], x[2]], slopes[i])
        srv.drain()
        st = srv.stats()
        results = {i: srv.results[i] for i in range(len(frames))}
    finally:
        srv.close()
    print("[stress] stats", {k2: st[k2] for k2 in ("frames", "fps", "p50_ms", "p99_ms", "max_frames_in_flight",
                                                   "jobs")}, flush=True)
    assert st["max_frames_in_flight"] > 4  # several jobs' worth of frames at once
    for i, x in enumerate(frames):
        r = results[i]
        assert r["status"] == 0 and r["compliant"] == 1, (i, r)
        cs = r["codestream"]
        ok, errors, facts = ref.check_codestream(cs, 48.0)
        assert ok, (i, errors)
        assert facts["total"] == len(cs) <= budget
        assert max(facts["comp0"], facts["comp1"], facts["comp2"]) <= comp_budget
        t1 = oracle.encode_frame_xyz(x, k, thr=slopes[i], threads=16)
        want, lam, _ = g.encode_codestream(L, t1, fps=48.0, mbps=250.0, tag="hip")
        print(f"[stress] frame {i}: threshold {slopes[i]:g}, {len(cs)} bytes (budget {budget}), lambda "
              f"{r['lambda']:.1f} / oracle path {lam:.1f}, equal {want == cs}", flush=True)
        assert want == cs, f"frame {i} differs from the oracle path"
