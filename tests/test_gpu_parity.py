"""Parity of the HIP path with the CPU oracle on the same seeded inputs.

Bar: bit-exact for integer and byte results (codes, quantised words, Tier-1 info, pass
lengths, pass distortions, codewords), and bit-exact float32 DWT planes as well: the
kernels evaluate the reference's operations in the reference's order with one rounding
each (no FMA contraction), so the float tolerance stated for this path is zero ulp."""
import numpy as np
import pytest

import libgail_amd as g
from t1_compare import compare_tier1

pytestmark = pytest.mark.gpu


def _frame_parity(device, oracle, name, seed, thr=0.0):
    W, H, k = g.CONTAINERS[name]
    print(f"[parity] {name} seed {seed} thr {thr}", flush=True)
    L = g.FrameLayout(W, H, k)
    x = g.synth_frame(seed, W, H)
    enc = g.Encoder(device, L, g.xyz_settings(L))
    try:
        enc.set_slope(thr)
        res = enc.encode(x, debug=True)
    finally:
        enc.close()
    ref = oracle.encode_frame_xyz(x, k, thr=thr, threads=16, want_planes=True)
    assert np.array_equal(res.debug["codes"], x), "X'Y'Z' codes differ from the input"
    wav_bad = int(np.count_nonzero(res.debug["wavelet"].view(np.uint32) != ref["planes"].view(np.uint32)))
    maxdiff = float(np.max(np.abs(res.debug["wavelet"] - ref["planes"])))
    print(f"[parity] wavelet samples differing: {wav_bad}, max |diff| {maxdiff}", flush=True)
    assert wav_bad == 0
    q = res.debug["coeffs"].reshape(-1)
    qd = q.astype(np.int64) - ref["coeffs"].astype(np.int64)
    mism = int(np.count_nonzero(qd))
    print(f"[parity] quantised words differing: {mism}", flush=True)
    assert mism == 0
    bad, lines = compare_tier1(res, ref)
    for ln in lines:
        print("[parity]", ln, flush=True)
    assert bad == 0, f"{bad} code-blocks differ"
    return res


def test_plain_2k_flat(device, oracle):
    _frame_parity(device, oracle, "2k_flat", 1)


def test_plain_2k_flat_early_stop(device, oracle):
    _frame_parity(device, oracle, "2k_flat", 2, thr=50.0)


def test_plain_4k_flat(device, oracle):
    """The plain workload: one 4K flat 3996x2160 frame, 6 levels, ragged edge blocks."""
    res = _frame_parity(device, oracle, "4k_flat", 3)
    assert len(res.info) == 25986
