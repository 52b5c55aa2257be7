"""Checks of the oracle that the mathematics fixes, independent of the reference: the
X'Y'Z' pass-through identity, perfect reconstruction of the 9/7 transform, the DWT's DC
response, the quantiser's closed form, code-block counts of every DCI container, and a
Tier-1 round trip through an independent decoder at every truncation point."""
import numpy as np
import pytest

from oracle_bind import random_block


def test_xyz_passthrough_identity(oracle):
    lut, cp = oracle.color_tables(5, 4)  # gamma 2.6, X'Y'Z' primaries
    for v in range(4096):
        s = oracle.lib.or_xyz12_to_16(v)
        assert oracle.color_pixel(lut, cp, s, s, s) == (v, v, v), v


def test_dwt_perfect_reconstruction(oracle):
    rng = np.random.default_rng(5)
    for (W, H, NL) in [(301, 157, 5), (64, 64, 3), (33, 17, 2), (2, 9, 1)]:
        x = rng.uniform(-2048, 2047, size=(H, W))
        y = oracle.fdwt_f64(x, NL)
        assert np.max(np.abs(oracle.idwt_f64(y, NL) - x)) < 1e-9
        # float32 lifting against the fp64 oracle: the reference's own bound is 0.05
        f = oracle.fdwt(x.astype(np.float32), NL)
        assert np.max(np.abs(oracle.idwt_f64(f.astype(np.float64), NL) - x)) < 0.05


def test_dwt_dc_response(oracle):
    # A constant signal puts everything into the LL band: every high-pass coefficient of a
    # whole-sample symmetric transform of a constant is zero, and LL scales by K^-2 ... per
    # level with the lifting normalisation (the low-pass DC gain of the 9/7 analysis is 1/K
    # per dimension times the predict/update DC gains).
    x = np.full((64, 96), 7.0)
    y = oracle.fdwt_f64(x, 1)
    high = np.concatenate([y[:32, 48:].ravel(), y[32:, :].ravel()])
    assert np.max(np.abs(high)) < 1e-12
    ll = y[:32, :48]
    assert np.allclose(ll, ll[0, 0]) and ll[0, 0] != 0


def test_quantizer_closed_form(oracle):
    rng = np.random.default_rng(9)
    for _ in range(2000):
        v = np.float32(rng.normal(0, 500))
        inv = np.float32(1.0) / np.float32(rng.uniform(0.05, 8))
        mag = (1 << int(rng.integers(1, 23))) * 64 - 1
        a = np.float32(abs(v)) * inv * np.float32(64.0)
        m = mag if a >= np.float32(mag) else int(a)
        want = m | (0x80000000 if (v < 0 and m >= 64) else 0)
        assert oracle.quantize_coeff(float(v), float(inv), mag) == want


@pytest.mark.parametrize("name,count", [((1998, 1080, 0), 6708), ((2048, 858, 0), 5424), ((2048, 1080, 0), 6708),
                                        ((3996, 2160, 1), 25986), ((4096, 1716, 1), 20976),
                                        ((4096, 2160, 1), 26292)])
def test_block_counts(oracle, name, count):
    bands, cbl = oracle.layout(*name)
    assert len(cbl) == count
    # every sample of each component plane is covered by exactly one code-block
    W, H, _ = name
    for c in range(3):
        area = sum(cb.w * cb.h for cb in cbl if bands[cb.band].comp == c)
        assert area == W * H


def test_t1_round_trip_every_truncation(oracle):
    rng = np.random.default_rng(123)
    for trial in range(120):
        c, w, h, orient = random_block(rng, trial)
        r = oracle.t1_encode(c, w, h, orient)
        if not r["nbps"]:
            continue
        mag, sign, sig = oracle.t1_decode(r["data"], w, h, orient, r["nbps"], r["num_passes"])
        m = (c[:h, :w] & 0x7FFFFFFF) >> 6
        assert np.array_equal(mag, m)
        assert np.array_equal(sign[m > 0], (c[:h, :w] >> 31)[m > 0])
        for k in range(1, r["num_passes"] + 1):
            L = r["pass_len"][k - 1]
            assert L == len(r["data"]) or r["data"][L - 1] != 0xFF
            a = oracle.t1_decode(r["data"], w, h, orient, r["nbps"], k)
            b = oracle.t1_decode(r["data"][:L], w, h, orient, r["nbps"], k)
            assert np.array_equal(a[0], b[0]) and np.array_equal(a[2], b[2])


def test_t1_planted_mistake_is_caught(oracle):
    """A wrong byte in a codeword must change the decode (the round trip pins the coder)."""
    rng = np.random.default_rng(77)
    c, w, h, orient = random_block(rng, 3)
    r = oracle.t1_encode(c, w, h, orient)
    bad = bytearray(r["data"])
    bad[len(bad) // 2] ^= 0x10
    mag, _, _ = oracle.t1_decode(bytes(bad), w, h, orient, r["nbps"], r["num_passes"])
    assert not np.array_equal(mag, (c[:h, :w] & 0x7FFFFFFF) >> 6)


def test_tier1_mismatch_counts_planted_differences(oracle):
    """The vectorised Tier-1 comparison the large GPU tests use finds one block changed in
    each of its fields, and nothing in an identical frame."""
    import libgail_amd as g
    from t1_compare import compare_tier1, tier1_mismatch

    class Res:
        def __init__(self, d):
            self.info, self.pass_len, self.pass_dist, self.data = (np.array(d[k]) for k in
                                                                  ("info", "pass_len", "pass_dist", "data"))

        def block(self, i):
            nbps, np_, off, ln, pb = (int(v) for v in self.info[i])
            return dict(nbps=nbps, num_passes=np_, data=bytes(self.data[off:off + ln]),
                        pass_len=[int(v) for v in self.pass_len[pb:pb + np_]],
                        pass_dist=[int(v) for v in self.pass_dist[pb:pb + np_]])

    ref = oracle.encode_frame_xyz(g.synth_frame(3, 256, 128), 0, threads=4)
    assert tier1_mismatch(Res(ref), ref) == 0
    coded = np.flatnonzero(ref["info"][:, 3] > 2)
    for field in ("nbps", "pass_len", "pass_dist", "data"):
        r = Res(ref)
        i = int(coded[len(coded) // 2])
        nbps, np_, off, ln, pb = (int(v) for v in r.info[i])
        if field == "nbps":
            r.info[i, 0] += 1
        elif field == "pass_len":
            r.pass_len[pb + np_ - 1] += 1
        elif field == "pass_dist":
            r.pass_dist[pb] -= 1
        else:
            r.data[off + ln - 1] ^= 1
        assert tier1_mismatch(r, ref) == 1, field
        assert compare_tier1(r, ref)[0] == 1, field
