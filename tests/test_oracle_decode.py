"""The oracle's decoder (oracle.c: or_t1_decode_words, or_idwt97_2d_f32, or_decode_frame)
pinned on the CPU, and the host build of the GPU decoder's per-block code checked against it:

* against the compiled reference (oracle/_ref: t1_dec.cpp, dwt_cpu.cpp, cpu_decode_backend.cpp,
  codestream_parse.cpp, t2_parse.cpp), bit for bit, where it was built;
* against closed forms that hold without the reference: a full decode gives every sample
  2 * magnitude + 1 with its sign (the encoder's own coefficients); a truncated one gives
  2 * (decoded magnitude) + 2^(last plane coded) (or_t1_decode's independent decoder);
  the float inverse DWT undoes the forward one to float accuracy; a frame coded with small
  steps decodes to within a few codes of its source.
Seeds and sizes are fixed so every run checks the same cases."""
import ctypes as C
import os

import numpy as np
import pytest

import libgail_amd as g
from oracle_bind import _p, block_codes, random_block, reference

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOST_DEC = os.path.join(ROOT, "tests", "native", "libt1dec_host.so")


def _host_decoder():
    if not os.path.exists(HOST_DEC):
        pytest.skip("tests/native/libt1dec_host.so not built (make -C tests/native)")
    L = C.CDLL(HOST_DEC)
    L.host_t1_decode_words.restype = C.c_int
    L.host_t1_decode_words.argtypes = [C.c_void_p, C.c_uint32, C.c_uint32] + [C.c_int] * 5 + [C.c_void_p]
    return L


def _blocks(oracle, n, seed):
    rng = np.random.default_rng(seed)
    out = []
    for trial in range(n):
        c, w, h, orient = random_block(rng, trial)
        out.append((c, w, h, orient, oracle.t1_encode(c, w, h, orient)))
    return out


def test_t1_words_full_decode_closed_form(oracle):
    """All passes decoded: a significant sample's word is its sign and 2 * magnitude + 1 (the
    magnitude bits the encoder coded, i.e. the coefficient without its 6 fractional bits)."""
    for c, w, h, orient, r in _blocks(oracle, 120, 3):
        words = oracle.t1_decode_words(r["data"], w, h, orient, r["nbps"], r["num_passes"])
        mag = (c[:h, :w] & 0x7FFFFFFF) >> 6
        sign = c[:h, :w] >> 31
        want = np.where(mag > 0, (sign << 31) | (2 * mag + 1), 0).astype(np.uint32)
        assert np.array_equal(words, want)


def test_t1_words_truncated_closed_form(oracle):
    """At every truncation point the word of a significant sample is its sign and
    2 * (decoded magnitude) + 2^q, q the lowest plane the decoded passes coded it in: the
    last pass's plane p, unless that plane got no further than its significance pass and the
    sample was significant before it (then p + 1). The magnitudes come from or_t1_decode,
    an independent decoder that keeps magnitude bits instead of words."""
    for c, w, h, orient, r in _blocks(oracle, 40, 5):
        nbps = r["nbps"]
        for k in range(1, r["num_passes"] + 1):
            L = r["pass_len"][k - 1]
            words = oracle.t1_decode_words(r["data"][:L], w, h, orient, nbps, k)
            mag, sign, sig = oracle.t1_decode(r["data"][:L], w, h, orient, nbps, k)
            # pass 1 is the cleanup of plane nbps - 1, then (SPP, MRP, CUP) per plane below
            p = nbps - 1 - (k + 1) // 3
            last_spp = (k + 1) % 3 == 0
            mag = mag.astype(np.int64)
            q = np.where(last_spp & ((mag >> p) != 1), p + 1, p)
            want = np.where(sig.astype(bool), (sign.astype(np.int64) << 31) | (2 * mag + (1 << q)), 0)
            assert np.array_equal(words, want.astype(np.uint32)), (w, h, orient, k)


def test_host_build_of_device_decoder_matches_oracle(oracle):
    """The GPU kernel's per-block code (t1_decode_block.h) compiled for the host: every
    truncation point, codewords at every alignment of the 8-byte window."""
    lib = _host_decoder()
    rng = np.random.default_rng(11)
    checked = 0
    for c, w, h, orient, r in _blocks(oracle, 160, 42):
        for k in range(1, r["num_passes"] + 1):
            L = r["pass_len"][k - 1]
            want = oracle.t1_decode_words(r["data"][:L], w, h, orient, r["nbps"], k)
            buf = np.frombuffer(bytes(r["data"][:L]) or b"\0", np.uint8).copy()
            got = np.zeros(w * h, np.uint32)
            start = int(rng.integers(0, 16))
            assert lib.host_t1_decode_words(_p(buf), L, start, w, h, orient, r["nbps"], k, _p(got)) == 0
            assert np.array_equal(got.reshape(h, w), want), (w, h, orient, k, start)
            checked += 1
    assert checked > 3000


def test_host_build_of_device_decoder_edge_cases(oracle):
    """Empty and damaged codewords (zero length, random bytes, a codeword cut short) decode to
    what the oracle gives, and write nothing outside the block; more passes than the planes
    allow stop at the last plane."""
    lib = _host_decoder()
    rng = np.random.default_rng(13)
    for trial in range(60):
        w, h, orient = 1 + int(rng.integers(32)), 1 + int(rng.integers(32)), trial % 4
        nbps = 1 + int(rng.integers(20))
        n = int(rng.integers(0, 200))
        data = rng.integers(0, 256, size=n, dtype=np.uint8)
        if trial % 3 == 0:
            data[rng.random(n) < 0.3] = 0xFF  # markers and stuffing
        passes = 1 + int(rng.integers(3 * nbps + 4))
        want = oracle.t1_decode_words(bytes(data), w, h, orient, nbps, min(passes, 3 * nbps - 2))
        buf = data.copy() if n else np.zeros(1, np.uint8)
        got = np.zeros(w * h, np.uint32)
        assert lib.host_t1_decode_words(_p(buf), n, trial % 8, w, h, orient, nbps, passes, _p(got)) == 0
        assert np.array_equal(got.reshape(h, w), want), (trial, w, h, nbps, passes)


@pytest.mark.parametrize("shape", [(64, 48, 3), (1998, 1080, 5), (37, 5, 3), (2, 9, 2), (1, 17, 3)])
def test_idwt_perfect_reconstruction(oracle, shape):
    """Forward then inverse float DWT returns the signal to float accuracy (any size,
    including one-sample-wide and odd ones)."""
    W, H, levels = shape
    rng = np.random.default_rng(W * 7 + H)
    x = rng.uniform(-2048, 2048, size=(1, H, W)).astype(np.float32)
    y = oracle.fdwt(x[0], levels)[None]
    z = oracle.idwt(y, levels)
    assert np.max(np.abs(z - x)) < 0.02


def test_decode_frame_close_to_source(oracle):
    """A 2K frame coded with base step 1 and every pass decodes to within 3 codes of its
    source (the round trip of the quantiser and the two transforms)."""
    W, H = 1998, 1080
    x = g.synth_frame(17, W, H)
    enc = oracle.encode_frame_xyz(x, 0, threads=8)
    pic, _ = oracle.decode_frame(W, H, 0, enc["data"], block_codes(enc["info"], enc["pass_len"]), high_bits=False)
    d = pic.reshape(H, W, 3).transpose(2, 0, 1).astype(np.int32) - x.astype(np.int32)
    assert np.abs(d).max() <= 3 and np.abs(d).mean() < 0.5


# ---------------------------------------------------------------- against the compiled reference
@pytest.fixture(scope="module")
def ref():
    r = reference()
    if r is None:
        pytest.skip("compiled reference not built (oracle/_ref)")
    return r


def test_t1_words_match_reference(oracle, ref):
    """t1DecodeBlock at every truncation point, and on damaged codewords."""
    rng = np.random.default_rng(21)
    for c, w, h, orient, r in _blocks(oracle, 160, 7):
        for k in range(1, r["num_passes"] + 1):
            L = r["pass_len"][k - 1]
            a = oracle.t1_decode_words(r["data"][:L], w, h, orient, r["nbps"], k)
            b = ref.t1_decode_words(r["data"][:L], w, h, orient, r["nbps"], k)
            assert np.array_equal(a, b), (w, h, orient, k)
    for trial in range(60):
        w, h, orient, nbps = 1 + int(rng.integers(32)), 1 + int(rng.integers(32)), trial % 4, 1 + int(rng.integers(20))
        data = bytes(rng.integers(0, 256, size=int(rng.integers(0, 120)), dtype=np.uint8))
        passes = 1 + int(rng.integers(3 * nbps - 2))
        assert np.array_equal(oracle.t1_decode_words(data, w, h, orient, nbps, passes),
                              ref.t1_decode_words(data, w, h, orient, nbps, passes))


@pytest.mark.parametrize("shape", [(1998, 1080, 5, 0), (1998, 1080, 5, 2), (3996, 2160, 6, 0), (37, 5, 3, 1),
                                   (1, 17, 3, 0)])
def test_idwt_matches_reference(oracle, ref, shape):
    W, H, levels, skip = shape
    rng = np.random.default_rng(W + H)
    x = rng.normal(0, 300, size=(3, H, W)).astype(np.float32)
    a = oracle.idwt(x, levels, skip)
    b = ref.idwt(x, levels, skip)
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


def _codes_with_damage(enc, rng):
    """All-pass codes, then random truncations, blocks without passes and blocks whose bytes
    lie past the codestream."""
    info = enc["info"]
    n = len(info)
    keep = np.where(rng.random(n) < 0.5, info[:, 1], rng.integers(0, 64, size=n)).astype(np.int64)
    codes = block_codes(info, enc["pass_len"], keep)
    bad = rng.random(n) < 0.01
    codes["offset"][bad] = len(enc["data"]) - 3
    codes["len"][bad] = np.maximum(codes["len"][bad], 10)
    return codes


@pytest.mark.parametrize("case", [(0, 0, True, 0), (0, 1, False, 6), (0, 5, True, 0), (1, 0, True, 0),
                                  (1, 2, True, 4), (1, 6, False, 0)])
def test_decode_frame_matches_reference(oracle, ref, case):
    """CpuDecoderSlot::decode on a 2K or 4K frame: pictures and wavelet planes equal, for
    reductions 0..levels, both word alignments, padded rows and damaged block tables."""
    fourk, reduce, high, pad = case
    W, H = (3996, 2160) if fourk else (1998, 1080)
    x = g.synth_frame(30 + reduce, W, H)
    enc = oracle.encode_frame_xyz(x, fourk, thr=0.0 if fourk else 0.5, threads=8)
    codes = _codes_with_damage(enc, np.random.default_rng(reduce))
    a, wa = oracle.decode_frame(W, H, fourk, enc["data"], codes, reduce, high, pad_row=pad, want_wavelet=True)
    b, wb = ref.decode_frame(W, H, fourk, enc["data"], codes, reduce, high, pad_row=pad, want_wavelet=True)
    assert np.array_equal(a, b)
    assert np.array_equal(wa.view(np.uint32), wb.view(np.uint32))


@pytest.mark.parametrize("container", ["2k_scope", "4k_full"])
def test_decode_frame_other_containers_match_reference(oracle, ref, container):
    W, H, fourk = g.CONTAINERS[container]
    x = g.synth_frame(91, W, H)
    enc = oracle.encode_frame_xyz(x, fourk, thr=0.5, threads=8)
    codes = _codes_with_damage(enc, np.random.default_rng(5))
    a, _ = oracle.decode_frame(W, H, fourk, enc["data"], codes, 1, True)
    b, _ = ref.decode_frame(W, H, fourk, enc["data"], codes, 1, True)
    assert np.array_equal(a, b)


def test_codestream_blocks_match_tier1(oracle, ref):
    """The codestream written by libgail.so (rate control + Tier-2) read back by the
    reference's own header and packet parser: every code-block's bytes are the first
    bytes of its Tier-1 codeword, its bit-planes are the encoder's, and the reference decodes
    that codestream to the oracle's picture."""
    W, H = 1998, 1080
    L = g.FrameLayout(W, H, 0)
    x = g.synth_frame(44, W, H)
    enc = oracle.encode_frame_xyz(x, 0, threads=8)
    cs, _, _ = g.encode_codestream(L, enc, fps=24.0, mbps=125.0)
    parsed = ref.parse_blocks(cs)
    assert parsed is not None
    pW, pH, fk, codes, bands = parsed
    assert (pW, pH, fk) == (W, H, False) and len(codes) == L.num_cblks
    info = enc["info"]
    data = np.frombuffer(cs, np.uint8)
    coded = codes["passes"] > 0
    assert coded.sum() > 0.5 * len(codes)
    for i in np.flatnonzero(coded)[::7]:
        off, ln = int(info[i, 2]), int(codes["len"][i])
        assert int(codes["nbps"][i]) == int(info[i, 0])
        assert bytes(data[codes["offset"][i]:codes["offset"][i] + ln]) == bytes(enc["data"][off:off + ln])
    a, _ = oracle.decode_frame(W, H, 0, cs, codes)
    b, _ = ref.decode_frame(W, H, 0, cs, codes)
    assert np.array_equal(a, b)
