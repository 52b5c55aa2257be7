"""The HIP decoder (plugin decoder entries) against the oracle's CPU decoder, which equals the
reference's CPU decode slot bit for bit (tests/test_oracle_decode.py). Bar: the pictures and
the dequantised wavelet planes are identical (integer Tier-1 words; float dequantisation,
inverse DWT and inverse ICT in the reference's operation order, one rounding per operation).

Inputs are frames coded by the oracle's encoder (all passes, or truncated as rate control
would leave them), with damaged block tables mixed in: blocks without passes, blocks whose
bytes lie past the codestream, more passes than the bit-planes allow."""
import ctypes as C
import threading

import numpy as np
import pytest

import libgail_amd as g
from oracle_bind import block_codes

pytestmark = pytest.mark.gpu

_ENC = {}


def _encoded(oracle, fourk, seed, thr):
    key = (fourk, seed, thr)
    if key not in _ENC:
        W, H = (3996, 2160) if fourk else (1998, 1080)
        x = g.synth_frame(seed, W, H)
        _ENC[key] = (W, H, oracle.encode_frame_xyz(x, fourk, thr=thr, threads=8))
    return _ENC[key]


def _damaged(enc, seed):
    rng = np.random.default_rng(seed)
    info = enc["info"]
    n = len(info)
    keep = np.where(rng.random(n) < 0.5, info[:, 1], rng.integers(0, 64, size=n)).astype(np.int64)
    codes = block_codes(info, enc["pass_len"], keep)
    bad = rng.random(n) < 0.01
    codes["offset"][bad] = len(enc["data"]) - 3
    codes["len"][bad] = np.maximum(codes["len"][bad], 10)
    many = rng.random(n) < 0.01
    codes["passes"][many] = 255
    return codes


def _compare(label, got, want):
    d = np.count_nonzero(got != want)
    print(f"[decode {label}] samples differing: {d} of {got.size}", flush=True)
    assert d == 0


@pytest.mark.parametrize("case", [(0, 0, True, 0, "all"), (0, 1, False, 8, "damaged"), (0, 5, True, 0, "damaged"),
                                  (1, 0, True, 0, "damaged"), (1, 2, False, 0, "all"), (1, 6, True, 2, "all")])
def test_decoder_matches_oracle(device, oracle, case):
    fourk, reduce, high, pad, kind = case
    W, H, enc = _encoded(oracle, fourk, 60 + fourk, 0.0 if fourk else 0.5)
    codes = block_codes(enc["info"], enc["pass_len"]) if kind == "all" else _damaged(enc, reduce)
    print(f"[decode case] {W}x{H} reduce {reduce} high_bits {high} pad {pad} {kind}", flush=True)
    want, wwant = oracle.decode_frame(W, H, fourk, enc["data"], codes, reduce, high, pad_row=pad, want_wavelet=True)
    L = g.FrameLayout(W, H, fourk)
    dec = g.Decoder(device, L, reduce, high)
    try:
        ow, oh = g.decoded_size(L, reduce)
        out = np.zeros((oh, ow * 3 + pad), np.uint16)
        _, tm, wav = dec.decode(enc["data"], codes, out=out, wavelet=True)
    finally:
        dec.close()
    print(f"[decode case] times {tm}", flush=True)
    _compare("wavelet", wav.view(np.uint32), wwant.view(np.uint32))
    _compare("picture", out[:, :ow * 3], want[:, :ow * 3])
    assert np.all(out[:, ow * 3:] == 0)  # bytes of a row past the picture are not written
    assert tm["t1"] > 0 and tm["dwt"] >= 0 and tm["color"] > 0 and tm["total"] > 0


@pytest.mark.parametrize("mbps", [40.0, 125.0])
def test_decoder_on_rate_controlled_frames(device, oracle, mbps):
    """The encoder's own output at a DCI rate (device rate control keeps each block's first
    passes) through the decoder: equal to the oracle's decode of the same blocks. At the
    lower rate few blocks are long enough to take a wave of their own."""
    W, H = 1998, 1080
    L = g.FrameLayout(W, H, 0)
    frames = [g.synth_frame(70 + i, W, H) for i in range(2)]
    b = g.Batch(device, [L] * 2, [g.with_rate_control(g.xyz_settings(L), fps=24.0, mbps=mbps)] * 2)
    try:
        for i, f in enumerate(frames):
            b.upload(i, f)
        res, _ = b.encode()
    finally:
        b.close()
    dec = g.Decoder(device, L, 0, True)
    try:
        for r in res:
            data = np.ascontiguousarray(r.data)
            codes = g.block_table(r)
            print(f"[decode rate {mbps}] {data.size} bytes, {int((codes['passes'] > 0).sum())} coded blocks, "
                  f"longest {int(codes['len'].max())}", flush=True)
            got, _, _ = dec.decode(data, codes)
            want, _ = oracle.decode_frame(W, H, 0, data, codes, 0, True)
            _compare(f"rate {mbps}", got, want)
    finally:
        dec.close()


@pytest.mark.parametrize("container", ["2k_scope", "2k_full", "4k_scope", "4k_full"])
def test_decoder_other_containers(device, oracle, container):
    """Scope and full DCI sizes (other band and edge-block shapes), reduce 1, damaged table."""
    W, H, fourk = g.CONTAINERS[container]
    x = g.synth_frame(90, W, H)
    enc = oracle.encode_frame_xyz(x, fourk, thr=0.5, threads=8)
    codes = _damaged(enc, 3)
    print(f"[decode {container}] {W}x{H}", flush=True)
    want, _ = oracle.decode_frame(W, H, fourk, enc["data"], codes, 1, True)
    dec = g.Decoder(device, g.FrameLayout(W, H, fourk), 1, True)
    try:
        got, _, _ = dec.decode(enc["data"], codes)
    finally:
        dec.close()
    _compare(container, got, want)


def test_decoder_device_memory_and_reuse(device, oracle):
    """The codestream and the picture in device memory (no host copies), and one decoder used
    for frames of other step sizes in turn (the frame's own layout)."""
    W, H, enc = _encoded(oracle, 0, 61, 0.5)
    p = g.plugin()
    x2 = g.synth_frame(62, W, H)
    enc2 = oracle.encode_frame_xyz(x2, 0, thr=0.5, threads=8, base_step=1.7)
    L = g.FrameLayout(W, H, 0)
    L2 = g.FrameLayout(W, H, 0, base_step=1.7)
    dec = g.Decoder(device, L, 0, True)
    dbuf = p.device_alloc(device.handle, 8 << 20)
    dpic = p.device_alloc(device.handle, W * H * 6)
    assert dbuf and dpic
    try:
        for e, lay, step in ((enc, L, 1.0), (enc2, L2, 1.7), (enc, L, 1.0)):
            data = np.ascontiguousarray(e["data"])
            assert data.size <= 8 << 20
            assert p.copy(C.c_void_p(dbuf), data.ctypes.data_as(C.c_void_p), data.size) == 0
            codes = block_codes(e["info"], e["pass_len"])
            dec.decode((dbuf, data.size), codes, frame_layout=lay, out=(dpic, W * 6))
            got = np.zeros((H, W * 3), np.uint16)
            assert p.copy(got.ctypes.data_as(C.c_void_p), C.c_void_p(dpic), got.nbytes) == 0
            want, _ = oracle.decode_frame(W, H, 0, data, codes, 0, True, base_step=step)
            _compare(f"device memory, step {step}", got, want)
    finally:
        p.device_free(C.c_void_p(dpic))
        p.device_free(C.c_void_p(dbuf))
        dec.close()


def test_decoder_pinned_memory(device, oracle):
    """A codestream in pinned host memory is read by DMA and the picture written into pinned
    host memory by DMA, with padded rows; the same as the oracle."""
    W, H, enc = _encoded(oracle, 0, 61, 0.5)
    p = g.plugin()
    data = np.ascontiguousarray(enc["data"])
    pitch = W * 6 + 64
    hin = p.host_alloc(data.size)
    hout = p.host_alloc(pitch * H)
    assert hin and hout
    dec = g.Decoder(device, g.FrameLayout(W, H, 0), 0, False)
    try:
        C.memmove(hin, data.ctypes.data, data.size)
        C.memset(hout, 0xEE, pitch * H)
        codes = block_codes(enc["info"], enc["pass_len"])
        dec.decode((hin, data.size), codes, out=(hout, pitch))
        got = np.ctypeslib.as_array((C.c_uint8 * (pitch * H)).from_address(hout)).reshape(H, pitch)
        want, _ = oracle.decode_frame(W, H, 0, data, codes, 0, False)
        _compare("pinned", got[:, :W * 6].copy().view(np.uint16), want)
        assert np.all(got[:, W * 6:] == 0xEE)
    finally:
        dec.close()
        p.host_free(C.c_void_p(hin))
        p.host_free(C.c_void_p(hout))


def test_decoders_on_threads(device, oracle):
    """Decoders on two threads at once (each on its own stream) give the oracle's pictures."""
    W, H, enc = _encoded(oracle, 0, 61, 0.5)
    codes = block_codes(enc["info"], enc["pass_len"])
    want, _ = oracle.decode_frame(W, H, 0, enc["data"], codes, 1, True)
    L = g.FrameLayout(W, H, 0)
    results, errors = [], []

    def run():
        try:
            d = g.Decoder(device, L, 1, True)
            try:
                for _ in range(3):
                    pic, _, _ = d.decode(enc["data"], codes)
                    results.append(pic)
            finally:
                d.close()
        except Exception as e:  # noqa: BLE001
            errors.append(e)

    ts = [threading.Thread(target=run) for _ in range(2)]
    for t in ts:
        t.start()
    for t in ts:
        t.join()
    assert not errors, errors
    assert len(results) == 6
    for pic in results:
        _compare("threads", pic, want)


def test_decoder_refuses_mismatched_frames(device, oracle):
    W, H, enc = _encoded(oracle, 0, 61, 0.5)
    L = g.FrameLayout(W, H, 0)
    dec = g.Decoder(device, L, 0, True)
    try:
        codes = block_codes(enc["info"], enc["pass_len"])
        with pytest.raises(g.GailError) as e:
            dec.decode(enc["data"], codes[:-1])
        assert e.value.status == 4
        with pytest.raises(g.GailError) as e:
            dec.decode(enc["data"], codes, frame_layout=g.FrameLayout(2048, 1080, 0))
        assert e.value.status == 4
        # an empty codestream with no coded blocks decodes to mid-grey (all coefficients zero)
        empty = np.zeros(len(codes), g.CBLK_CODE)
        pic, _, _ = dec.decode(np.zeros(0, np.uint8), empty)
        assert np.all(pic == (2048 << 4))
    finally:
        dec.close()
    with pytest.raises(g.GailError):
        g.Decoder(device, L, 6, True)
