"""The oracle against the reference compiled from its own sources (oracle/_ref), on fresh
seeded inputs. Skipped where oracle/_ref was not built."""
import numpy as np
import pytest

from oracle_bind import color_cfg, random_block, reference


@pytest.fixture(scope="module")
def ref():
    r = reference()
    if r is None:
        pytest.skip("oracle/_ref/libgailref.so not built")
    return r


def test_t1_fresh_blocks(oracle, ref):
    rng = np.random.default_rng(2024)
    for trial in range(300):
        c, w, h, orient = random_block(rng, trial)
        thr = (0.0, 1.0, 20.0)[trial % 3]
        assert oracle.t1_encode(c, w, h, orient, thr=thr, scale=0.9, trace=True) == \
            ref.t1_encode(c, w, h, orient, thr=thr, scale=0.9, trace=True), trial


def test_dwt_fresh(oracle, ref):
    rng = np.random.default_rng(3)
    for (W, H, NL) in [(157, 301, 5), (8, 8, 2), (5, 300, 3)]:
        x = rng.uniform(-4000, 4000, size=(H, W)).astype(np.float32)
        assert np.array_equal(oracle.fdwt(x, NL).view(np.uint32), ref.fdwt(x, NL).view(np.uint32))


def test_frame_scope_2k(oracle, ref):
    import libgail_amd as g
    x = g.synth_frame(9, 2048, 858)
    a = oracle.encode_frame_xyz(x, 0, thr=10.0, threads=8)
    b = ref.encode_frame_xyz(x, 0, thr=10.0, threads=8)
    for k in ("coeffs", "info", "pass_len", "pass_dist", "data"):
        assert np.array_equal(a[k], b[k]), k


# HDR and managed-colour variants (color.cpp:296-372, color_math.h:27-179): tone maps,
# tone norms, target gamuts with compression, and the 3D LUT route. ColorParams'
# enumerations: tone_map 0 signal, 1 shoulder, 2 BT.2390, 3 none; tone_norm 0 channel,
# 1 max(R,G,B), 2 luminance; gamut 0 P3-D65, 1 DCI-P3, 2 none.
COLOR_CASES = [
    dict(transfer=0, primaries=0, gamut=0),
    dict(transfer=1, primaries=1, gamut=1, gamut_threshold=0.5),
    dict(transfer=2, primaries=1, tone_map=1),
    dict(transfer=2, primaries=1, tone_map=1, tone_norm=1, gamut=0),
    dict(transfer=2, primaries=1, tone_map=1, tone_norm=2, hdr_peak=4000.0, knee=0.5),
    dict(transfer=3, primaries=1, tone_map=1, gamut=1),
    dict(transfer=2, primaries=1, tone_map=2),
    dict(transfer=2, primaries=1, tone_map=2, tone_norm=1, gamut=0),
    dict(transfer=2, primaries=2, tone_map=2, tone_norm=2, hdr_peak=600.0),
    dict(transfer=3, primaries=1, tone_map=2, gamut=1),
    dict(transfer=2, primaries=1, tone_map=3, gamut=0),
    dict(transfer=2, primaries=1, tone_map=1, hdr_peak=150.0),  # source fits: no shoulder
]


def _color_pixels(rng, n):
    px = rng.integers(0, 65536, size=(n, 3), dtype=np.uint16)
    ramp = np.linspace(0, 65535, 257).astype(np.uint16)
    grey = np.stack([ramp, ramp, ramp], 1)
    prim = np.array([[65535, 0, 0], [0, 65535, 0], [0, 0, 65535], [65535, 65535, 0], [0, 0, 0],
                     [65535, 65535, 65535], [40000, 65535, 1000]], np.uint16)
    return np.concatenate([px, grey, prim])


@pytest.mark.parametrize("case", range(len(COLOR_CASES)))
def test_color_hdr_variants(oracle, ref, case):
    cfg = color_cfg(**COLOR_CASES[case])
    lut_o, cp_o = oracle.color_setup(cfg)
    lut_r, cp_r = ref.color_setup(cfg)
    assert np.array_equal(lut_o.view(np.uint32), lut_r.view(np.uint32))
    assert np.array_equal(cp_o.view(np.uint32), cp_r.view(np.uint32))
    px = _color_pixels(np.random.default_rng(700 + case), 20000)
    a = oracle.color_pixels(lut_o, cp_o, None, px)
    b = ref.color_pixels(cfg, None, px)
    assert np.array_equal(a, b)


@pytest.mark.parametrize("n", [2, 17])
def test_color_lut3d(oracle, ref, n):
    rng = np.random.default_rng(n)
    lut3 = rng.uniform(-0.1, 1.1, size=n * n * n * 3).astype(np.float32)
    cfg = color_cfg(transfer=2, primaries=1, tone_map=1, lut3d_size=n)
    lut_o, cp_o = oracle.color_setup(cfg)
    _, cp_r = ref.color_setup(cfg, lut3)
    assert np.array_equal(cp_o.view(np.uint32), cp_r.view(np.uint32))
    px = _color_pixels(rng, 20000)
    assert np.array_equal(oracle.color_pixels(lut_o, cp_o, lut3, px), ref.color_pixels(cfg, lut3, px))
