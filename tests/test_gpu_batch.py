"""Batched form, colour inputs, graph replays and the overflow re-run, against the oracle.
Every comparison is bit-exact (integer results and the float32 DWT planes alike)."""
import numpy as np
import pytest

import libgail_amd as g
from t1_compare import compare_tier1, planar_from_synth as _planar_from_synth

pytestmark = pytest.mark.gpu


def _check_frame(res, ref, label):
    bad, lines = compare_tier1(res, ref)
    for ln in lines:
        print(f"[{label}]", ln, flush=True)
    assert bad == 0, f"{label}: {bad} code-blocks differ"


def test_mixed_geometry_batch(device, oracle):
    """The mixed batched workload in small: 2K flat, 2K scope, 4K scope and 4K full frames
    with their own seeds and early-termination thresholds in one batch (one DWT launch per
    level, one Tier-1 launch for all code-blocks)."""
    names = ["2k_flat", "2k_scope", "4k_scope", "4k_full"]
    thr = [0.0, 40.0, 0.0, 400.0]
    layouts, frames = [], []
    for i, n in enumerate(names):
        W, H, k = g.CONTAINERS[n]
        layouts.append(g.FrameLayout(W, H, k))
        frames.append(g.synth_frame(100 + i, W, H))
    b = g.Batch(device, layouts, [g.xyz_settings(L) for L in layouts])
    try:
        for i, f in enumerate(frames):
            b.upload(i, f)
            b.set_slope(i, thr[i])
        print("[batch] encode 4 mixed frames", flush=True)
        outs, _ = b.encode()
    finally:
        b.close()
    for i, n in enumerate(names):
        k = g.CONTAINERS[n][2]
        ref = oracle.encode_frame_xyz(frames[i], k, thr=thr[i], threads=16)
        _check_frame(outs[i], ref, n)


def test_large_launch_lane_coder(device, oracle):
    """Three 4K flat frames in one launch: 77,958 code-blocks, more than the 65,536 up to which
    a launch codes its longest blocks one per wave, so here MQ codes every block one per lane.
    Each frame equals the oracle, as the smaller launches of the other tests (which use both
    coders) do."""
    W, H, k = g.CONTAINERS["4k_flat"]
    L = g.FrameLayout(W, H, k)
    frames = [g.synth_frame(610 + i, W, H) for i in range(3)]
    thr = [0.0, 1500.0, 300.0]
    b = g.Batch(device, [L] * 3, [g.xyz_settings(L)] * 3)
    try:
        for i, f in enumerate(frames):
            b.upload(i, f)
            b.set_slope(i, thr[i])
        print("[batch] encode 3 4K frames in one launch", flush=True)
        outs, _ = b.encode()
    finally:
        b.close()
    for i in range(3):
        ref = oracle.encode_frame_xyz(frames[i], k, thr=thr[i], threads=16)
        _check_frame(outs[i], ref, f"4k_flat #{i}")


def test_same_alone_and_in_the_mix(device):
    """A frame's Tier-1 output is the same encoded alone and inside a batch of others."""
    W, H, k = g.CONTAINERS["2k_flat"]
    L = g.FrameLayout(W, H, k)
    x = g.synth_frame(77, W, H)
    enc = g.Encoder(device, L, g.xyz_settings(L))
    try:
        enc.set_slope(25.0)
        alone = enc.encode(x)
    finally:
        enc.close()
    W4, H4, k4 = g.CONTAINERS["4k_flat"]
    L4 = g.FrameLayout(W4, H4, k4)
    b = g.Batch(device, [L4, L, L4], [g.xyz_settings(L4), g.xyz_settings(L), g.xyz_settings(L4)])
    try:
        b.upload(0, g.synth_frame(1, W4, H4))
        b.upload(1, x)
        b.upload(2, g.synth_frame(2, W4, H4))
        b.set_slope(0, 300.0)
        b.set_slope(1, 25.0)
        outs, _ = b.encode()
    finally:
        b.close()
    mix = outs[1]
    assert np.array_equal(alone.info, mix.info)
    assert np.array_equal(alone.pass_len, mix.pass_len)
    assert np.array_equal(alone.pass_dist, mix.pass_dist)
    assert np.array_equal(alone.data, mix.data)


def test_graph_replays_match_eager(device):
    """The captured chain replayed several times, with eager work and syncs in between,
    gives the eager result every time."""
    W, H, k = g.CONTAINERS["2k_scope"]
    L = g.FrameLayout(W, H, k)
    frames = [g.synth_frame(300 + i, W, H) for i in range(2)]
    eager = g.Batch(device, [L, L], [g.xyz_settings(L)] * 2)
    graph = g.Batch(device, [L, L], [g.xyz_settings(L)] * 2, use_graph=True)
    try:
        for i, f in enumerate(frames):
            eager.upload(i, f)
            graph.upload(i, f)
            eager.set_slope(i, 30.0)
            graph.set_slope(i, 30.0)
        want, _ = eager.encode()
        for rep in range(4):
            print(f"[graph] replay {rep}", flush=True)
            got, _ = graph.encode()
            for i in range(2):
                assert np.array_equal(got[i].data, want[i].data), f"replay {rep} frame {i}"
                assert np.array_equal(got[i].pass_len, want[i].pass_len)
                assert np.array_equal(got[i].info, want[i].info)
            eager.encode()  # eager work between replays
    finally:
        eager.close()
        graph.close()


def test_pool_overflow_reruns_on_gpu(device, oracle):
    """A symbol pool far too small is detected on the device; the job re-runs Tier-1 on the
    GPU with larger buffers and the result still matches the oracle."""
    W, H, k = g.CONTAINERS["2k_flat"]
    L = g.FrameLayout(W, H, k)
    x = g.synth_frame(5, W, H)
    b = g.Batch(device, [L], [g.xyz_settings(L)], pool_bytes_per_coef=0.02)
    try:
        b.upload(0, x)
        outs, _ = b.encode()
        st = b.stage_ms()
    finally:
        b.close()
    assert st["reruns"] >= 1
    _check_frame(outs[0], oracle.encode_frame_xyz(x, k, threads=16), "overflow")


@pytest.mark.parametrize("kind", ["gbrp12_bt709", "yuv420p10_bt709"])
def test_colour_inputs(device, oracle, kind):
    """The served workload's inputs: 12-bit G'B'R' and 10-bit Y'CbCr 4:2:0 (BT.709, limited range, chroma
    left), a 1920x1080 picture centred in the 2K flat container."""
    W, H, k = g.CONTAINERS["2k_flat"]
    aw, ah, ox, oy = 1920, 1080, 39, 0
    bits, sx, sy, rgb, fr = (12, 0, 0, 1, 1) if kind.startswith("gbrp") else (10, 1, 1, 0, 0)
    src = _planar_from_synth(g.synth_frame(55, aw, ah), bits, sx, sy, rgb)
    L = g.FrameLayout(W, H, k)
    s = g.planar_settings(L, bits, sx, sy, rgb, fr, active=(aw, ah), offset=(ox, oy))
    enc = g.Encoder(device, L, s)
    try:
        res = enc.encode(src, debug=True)
    finally:
        enc.close()
    rgb48 = oracle.planar_to_rgb48(src, bits, aw, ah, sx, sy, rgb, fr, 0, 0.2126, 0.0722, W, H, ox, oy)
    lut, cp = oracle.color_tables(0, 0)
    _, xyz = oracle.convert_frame(lut, cp, rgb48)
    codes = np.ascontiguousarray(xyz.transpose(2, 0, 1))
    diff = np.count_nonzero(res.debug["codes"] != codes)
    print(f"[colour {kind}] codes differing from the CPU path: {diff}", flush=True)
    assert diff == 0
    _check_frame(res, oracle.encode_frame_xyz(codes, k, threads=16), kind)


def test_slot_job_rebinds(device, oracle):
    """One job with three 4K-sized slots, bound to three different mixes in a row (2K and 4K
    containers, X'Y'Z' and Y'CbCr inputs, fewer frames than slots): every frame's Tier-1
    output equals the oracle's, and a frame bound again later gives the same bytes."""
    b = g.Batch(device, slots=[(4096, 2160, True)] * 3)
    W2, H2, k2 = g.CONTAINERS["2k_flat"]
    L2 = g.FrameLayout(W2, H2, k2)
    W4, H4, k4 = g.CONTAINERS["4k_scope"]
    L4 = g.FrameLayout(W4, H4, k4)
    Ls, _, ks = g.CONTAINERS["2k_scope"]
    L2s = g.FrameLayout(Ls, g.CONTAINERS["2k_scope"][1], ks)
    x2 = g.synth_frame(501, W2, H2)
    x4 = g.synth_frame(502, W4, H4)
    x2s = g.synth_frame(503, L2s.width, L2s.height)
    try:
        print("[slots] run 1: 2k_flat + 4k_scope", flush=True)
        b.bind([L2, L4], [g.xyz_settings(L2), g.xyz_settings(L4)])
        b.stage(0, x2)
        b.stage(1, x4)
        b.set_slope(0, 0.0)
        b.set_slope(1, 60.0)
        run1, _ = b.encode(from_staging=True)
        _check_frame(run1[0], oracle.encode_frame_xyz(x2, k2, threads=16), "slots run1 2k_flat")
        _check_frame(run1[1], oracle.encode_frame_xyz(x4, k4, thr=60.0, threads=16), "slots run1 4k_scope")

        print("[slots] run 2: 2k_scope + yuv420p10 2k_flat + 2k_flat", flush=True)
        aw, ah, ox, oy = 1920, 1080, 39, 0
        src = _planar_from_synth(g.synth_frame(504, aw, ah), 10, 1, 1, 0)
        sp = g.planar_settings(L2, 10, 1, 1, 0, 0, active=(aw, ah), offset=(ox, oy))
        b.bind([L2s, L2, L2], [g.xyz_settings(L2s), sp, g.xyz_settings(L2)])
        b.stage(0, x2s)
        b.stage(1, src)
        b.stage(2, x2)
        for i in range(3):
            b.set_slope(i, 0.0)
        run2, _ = b.encode(from_staging=True)
        _check_frame(run2[0], oracle.encode_frame_xyz(x2s, ks, threads=16), "slots run2 2k_scope")
        rgb48 = oracle.planar_to_rgb48(src, 10, aw, ah, 1, 1, 0, 0, 0, 0.2126, 0.0722, W2, H2, ox, oy)
        lut, cp = oracle.color_tables(0, 0)
        _, xyz = oracle.convert_frame(lut, cp, rgb48)
        _check_frame(run2[1], oracle.encode_frame_xyz(np.ascontiguousarray(xyz.transpose(2, 0, 1)), k2, threads=16),
                     "slots run2 yuv420p10")
        for name in ("info", "pass_len", "pass_dist", "data"):
            assert np.array_equal(getattr(run2[2], name), getattr(run1[0], name)), name
    finally:
        b.close()


@pytest.mark.parametrize("opts", [dict(dwt_tile=0), dict(dwt_tile=1), dict(dwt_tile=2), dict(dwt_tile=4),
                                  dict(dwt_tile=5), dict(dwt_tile=6), dict(dwt_tile=7), dict(dwt_tile=9),
                                  dict(dwt_tile=8), dict(dwt_tile=9), dict(dwt_tile=3),
                                  dict(cm_waves=1),
                                  dict(cm_waves=2), dict(cm_waves=8), dict(mq_group=64), dict(mq_group=128),
                                  dict(mq_read=16), dict(mq_group=64, mq_read=16), dict(mq_group=128, mq_read=16)])
def test_kernel_variants_bit_exact(device, oracle, opts):
    """Every DWT tile shape, context-modelling and MQ workgroup size the autotuner can pick
    gives the oracle's Tier-1 output (2K flat, early termination on)."""
    W, H, k = g.CONTAINERS["2k_flat"]
    L = g.FrameLayout(W, H, k)
    x = g.synth_frame(61, W, H)
    b = g.Batch(device, [L], [g.xyz_settings(L)], **opts)
    try:
        b.upload(0, x)
        b.set_slope(0, 45.0)
        outs, _ = b.encode()
    finally:
        b.close()
    _check_frame(outs[0], oracle.encode_frame_xyz(x, k, thr=45.0, threads=16), f"variant {opts}")


def test_autotune_then_encode(device, oracle):
    """After the autotuner has run every variant on a batch, the batch's output still equals
    the oracle's."""
    W, H, k = g.CONTAINERS["2k_scope"]
    L = g.FrameLayout(W, H, k)
    frames = [g.synth_frame(70 + i, W, H) for i in range(2)]
    b = g.Batch(device, [L, L], [g.xyz_settings(L)] * 2)
    try:
        for i, f in enumerate(frames):
            b.upload(i, f)
            b.set_slope(i, 0.0)
        chosen, times = b.autotune(reps=1)
        print("[autotune]", chosen, times, flush=True)
        outs, _ = b.encode()
    finally:
        b.close()
    for i in range(2):
        _check_frame(outs[i], oracle.encode_frame_xyz(frames[i], k, threads=16), f"autotuned frame {i}")


def test_full_mixed_batch_64(device, oracle):
    """The batched workload at its full size: 64 frames, 16 each of 2K flat, 2K scope, 4K scope
    and 4K full, 950,400 code-blocks in one Tier-1 launch, the DWT chain replayed as a HIP
    graph with autotuned tile shapes and workgroup sizes, and an early-termination threshold
    of its own per frame. Every frame equals the oracle's Tier-1 of the same frame."""
    from t1_compare import tier1_mismatch
    names = ["2k_flat", "2k_scope", "4k_scope", "4k_full"]
    thr_choices = [0.0, 150.0, 800.0, 1600.0]
    layouts, specs = [], []
    for i in range(64):
        n = names[i % 4]
        W, H, k = g.CONTAINERS[n]
        layouts.append(g.FrameLayout(W, H, k))
        specs.append((n, 5000 + i, thr_choices[(i // 4) % 4], k))
    b = g.Batch(device, layouts, [g.xyz_settings(L) for L in layouts], use_graph=True)
    try:
        for i, (n, seed, thr, _) in enumerate(specs):
            W, H, _ = g.CONTAINERS[n]
            b.upload(i, g.synth_frame(seed, W, H))
            b.set_slope(i, thr)
        print("[batch64] autotune", flush=True)
        chosen = b.autotune()
        print(f"[batch64] chose {chosen}; encode 64 frames", flush=True)
        outs, _ = b.encode()
    finally:
        b.close()
    blocks = sum(len(o.info) for o in outs)
    print(f"[batch64] {blocks} code-blocks in one launch", flush=True)
    assert blocks == 950400
    bad_frames = []
    for i, (n, seed, thr, k) in enumerate(specs):
        W, H, _ = g.CONTAINERS[n]
        ref = oracle.encode_frame_xyz(g.synth_frame(seed, W, H), k, thr=thr, threads=16)
        bad = tier1_mismatch(outs[i], ref)
        if bad:
            bad_frames.append((i, n, thr, bad))
        if i % 8 == 7:
            print(f"[batch64] frames up to {i} checked, {len(bad_frames)} differ", flush=True)
    assert not bad_frames, f"frames differing from the oracle (index, size, threshold, blocks): {bad_frames}"


@pytest.mark.parametrize("shape", [3, 4, 5, 6, 7, 8, 9, 10])
def test_dwt_planes_every_shape(device, oracle, shape):
    """The wavelet planes of every DWT tile shape, LDS tiles and register-streaming strips,
    equal the oracle's float32 DWT bit for bit on frames whose sizes leave ragged strips and
    tiles at both edges (odd widths and heights, LL regions down to a few samples; a frame
    67 rows high, whose second tall strip lies mostly below the frame's last row)."""
    sizes = [(257, 131, 0), (1998, 1080, 0), (1001, 515, 1), (2048, 858, 0), (301, 67, 0)]
    layouts = [g.FrameLayout(W, H, k) for W, H, k in sizes]
    frames = [g.synth_frame(900 + i, W, H) for i, (W, H, _) in enumerate(sizes)]
    b = g.Batch(device, layouts, [g.xyz_settings(L) for L in layouts], dwt_tile=shape)
    try:
        for i, f in enumerate(frames):
            b.upload(i, f)
        b.encode()
        for i, (W, H, k) in enumerate(sizes):
            got = b.debug(i)["wavelet"]
            want = oracle.encode_frame_xyz(frames[i], k, threads=16, want_planes=True)["planes"]
            bad = int(np.count_nonzero(got.view(np.uint32) != want.view(np.uint32)))
            print(f"[dwt] shape {shape} {W}x{H}: {bad} samples differ", flush=True)
            assert bad == 0
    finally:
        b.close()
