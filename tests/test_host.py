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


def _encoder_create_status(cfg):
    import ctypes as C
    h = g._lib_api()
    h.gail_encoder_create.restype = C.c_int
    h.gail_encoder_create.argtypes = [C.POINTER(g.GailConfig), C.POINTER(C.c_void_p)]
    h.gail_encoder_destroy.argtypes = [C.c_void_p]
    enc = C.c_void_p()
    st = h.gail_encoder_create(C.byref(cfg), C.byref(enc))
    if enc:
        h.gail_encoder_destroy(enc)
    return st, h.gail_last_error().decode()


def test_stream_configs_validate():
    """The stream descriptions the served form takes are valid for every DCI size (the
    configuration is checked before any device is looked for), and a picture larger than
    its container is refused."""
    for name, (W, H, _) in g.CONTAINERS.items():
        for fmt in ("xyz12_planar", "gbrp12", "yuv420p10"):
            pic = None if fmt == "xyz12_planar" else (min(1920, W), min(1080, H))
            st, msg = _encoder_create_status(g.stream_config(name, fmt, picture=pic))
            assert st in (0, -2), (name, fmt, st, msg)  # OK on a GPU, UNSUPPORTED (no device) here
            assert "picture" not in msg and "invalid" not in msg, (name, fmt, msg)
    st, msg = _encoder_create_status(g.stream_config("2k_scope", "yuv420p10", picture=(1920, 1080)))
    assert st == -1 and "picture" in msg


def test_color_settings_validate():
    """HDR colour (tone maps, tone norms, target gamut, an app 3D LUT) is accepted for every
    preset, and out-of-range colour parameters are refused with the reference API's checks
    (gail_api.cpp:158-198), before any device is looked for."""
    import ctypes as C
    ok_codes = (0, -2)  # OK on a GPU, UNSUPPORTED (no device) here
    lut = np.linspace(0, 1, 17 ** 3 * 3, dtype=np.float32)
    for preset in range(4):
        for tone_map in range(4):
            for gamut in range(3):
                c = g.stream_config("2k_flat", "gbrp12", picture=(1920, 1080), preset=preset)
                c.color.tone_map, c.color.gamut, c.color.tone_norm = tone_map, gamut, tone_map % 3
                st, msg = _encoder_create_status(c)
                assert st in ok_codes, (preset, tone_map, gamut, st, msg)
    c = g.stream_config("2k_flat", "gbrp12", picture=(1920, 1080))
    c.color.mode, c.color.lut3d, c.color.lut3d_size = 1, lut.ctypes.data, 17
    st, msg = _encoder_create_status(c)
    assert st in ok_codes, msg
    bad = [("knee", 1.0), ("knee", 0.0), ("gamut_threshold", 1.0), ("cinema_peak", 60.0), ("hdr_peak", 0.0),
           ("gamut", 3), ("tone_map", 4), ("lut3d_size", 1)]
    for field, value in bad:
        c = g.stream_config("2k_flat", "gbrp12", picture=(1920, 1080))
        if field == "lut3d_size":
            c.color.mode, c.color.lut3d = 1, lut.ctypes.data
        setattr(c.color, field, value)
        st, msg = _encoder_create_status(c)
        assert st == -1, (field, value, st, msg)


def test_server_api_without_device():
    """The served form's entry points fail cleanly, with a message, where there is no GPU
    (or refuse bad arguments before touching one)."""
    import ctypes as C
    h = g._lib_api()
    out = C.c_void_p()
    cfg = g.ServerConfig()
    h.gail_server_config_init(C.byref(cfg), C.sizeof(cfg))
    assert cfg.slots_per_job == 8 and cfg.max_jobs == 4
    assert h.gail_server_create(None, C.byref(out)) == -1  # GAIL_ERROR_INVALID
    cfg.slots_per_job = 100000
    assert h.gail_server_create(C.byref(cfg), C.byref(out)) == -1
    assert b"too large" in h.gail_last_error()
    assert h.gail_server_submit(None, 0, 0, None, None, 0.0) == -1
    assert h.gail_server_drain(None) == -1


def test_slope_grid():
    """Rate control bins slopes on floor(1024 log2 s) with integer arithmetic only: the bin
    agrees with the logarithm away from bin edges, and every bin's edge lies in its bin with
    the double just below it in the bin before (so no slope falls between bins)."""
    import ctypes as C
    import math
    h = g.host()
    h.gail_slope_bin.restype = C.c_int64
    h.gail_slope_bin.argtypes = [C.c_double]
    h.gail_slope_bin_edge.restype = C.c_double
    h.gail_slope_bin_edge.argtypes = [C.c_int64]
    rng = np.random.default_rng(3)
    for s in np.exp2(rng.uniform(-40, 60, 20000)):
        f = 1024 * math.log2(s)
        if abs(f - round(f)) > 1e-6:
            assert h.gail_slope_bin(float(s)) == math.floor(f), s
    for k in list(range(-5000, 5000, 7)) + [0, 1, 1023, 1024, -1, -1024, -1025]:
        e = h.gail_slope_bin_edge(k)
        assert h.gail_slope_bin(e) == k, k
        assert h.gail_slope_bin(np.nextafter(e, 0.0)) == k - 1, k
        assert abs(math.log2(e) * 1024 - k) < 1e-6


@pytest.mark.parametrize("name", ["2k_flat", "2k_scope", "4k_scope"])
def test_device_packet_sizes_match_tier2(oracle, name):
    """The packet-size code the GPU rate control runs (tag trees, pass counts, Lblock, byte
    stuffing of B.10) gives exactly the sizes of the host's Tier-2 builder, per component
    and tile-part, for all passes, none, random truncations and sparse inclusion of an
    oracle frame."""
    import ctypes as C
    W, H, k = g.CONTAINERS[name]
    t1 = oracle.encode_frame_xyz(g.synth_frame(12, W, H), k, thr=0.0, threads=8)
    info, pl, pd, data = g._t1_arrays(t1)
    h = g.host()
    h.gail_rc_measure_check.restype = C.c_int
    h.gail_rc_measure_check.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, C.c_void_p, C.c_void_p, C.c_void_p,
                                        C.c_void_p, C.c_uint64, C.c_int, C.c_uint64]
    bad = h.gail_rc_measure_check(W, H, k, 1.0, info.ctypes.data, pl.ctypes.data, pd.ctypes.data, data.ctypes.data,
                                  len(info), 12, 5)
    assert bad == 0
