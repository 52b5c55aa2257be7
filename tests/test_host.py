"""Host-side product code against the oracle and the mathematics: the layout libgail.so
builds, the synthetic generator, and rate control + Tier-2 on oracle Tier-1 output."""
import numpy as np
import pytest

import libgail_amd as g


@pytest.mark.parametrize("name", list(g.CONTAINERS))
def test_layout_matches_oracle(oracle, name):
    W, H, k = g.CONTAINERS[name]
    L = g.FrameLayout(W, H, k)
    bands, cbl = oracle.layout(W, H, k)
    assert L.num_bands == len(bands) and L.num_cblks == len(cbl)
    for a, b in zip(L.bands[:L.num_bands], bands):
        assert (a.comp, a.res, a.level, a.orient, a.px0, a.py0, a.w, a.h, a.ncbx, a.ncby, a.cblkBase, a.eps, a.mu,
                a.Mb) == (b.comp, b.res, b.level, b.orient, b.px0, b.py0, b.w, b.h, b.ncbx, b.ncby, b.cblk_base,
                          b.eps, b.mu, b.mb)
        assert a.delta == b.delta and a.weight == b.weight
    cb = np.array([[c.band, c.x0, c.y0, c.w, c.h] for c in cbl], np.int32)
    assert np.array_equal(L.cblks, cb)


def test_synth_deterministic_and_in_range():
    a = g.synth_frame(7, 512, 256, threads=1)
    b = g.synth_frame(7, 512, 256, threads=8)
    assert np.array_equal(a, b)
    assert a.dtype == np.uint16 and a.max() <= 4095
    assert abs(float(a.mean()) - 2048) < 8
    assert not np.array_equal(a, g.synth_frame(8, 512, 256))


@pytest.fixture(scope="module")
def frame_2k(oracle):
    W, H, k = g.CONTAINERS["2k_flat"]
    x = g.synth_frame(1, W, H)
    return g.FrameLayout(W, H, k), oracle.encode_frame_xyz(x, k, threads=8)


def test_codestream_under_cap(frame_2k):
    L, t1 = frame_2k
    cs, lam, mse = g.encode_codestream(L, t1, fps=24.0, mbps=250.0)
    assert len(cs) <= 1302083 - 16
    assert g.check_codestream(cs, 24.0, 0) == ""
    assert lam > 0 and 0 < mse < 4095 ** 2


def test_codestream_rates_monotone(frame_2k):
    L, t1 = frame_2k
    sizes = [len(g.encode_codestream(L, t1, fps=fps, mbps=250.0)[0]) for fps in (24.0, 48.0, 96.0)]
    assert sizes[0] > sizes[1] > sizes[2]
    assert sizes[1] <= int(250e6 / (8 * 48)) - 16


def test_check_catches_corruption(frame_2k):
    L, t1 = frame_2k
    cs = bytearray(g.encode_codestream(L, t1)[0])
    cs[-1] = 0
    assert g.check_codestream(bytes(cs), 24.0, 0) != ""
