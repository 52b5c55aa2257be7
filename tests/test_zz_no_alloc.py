"""gail_hip_decode_noalloc (include/gail_hip_plugin.h): the decoder entry that allocates no
device memory, its intermediates in the plugin's fixed device array. Two calls in a row, the
second with a smaller frame, each against decoder_decode of a decoder object on the same
codestream and block table (bit-exact: the same kernels); the codestream from host memory
and from device memory; and the refusals of a frame or codestream beyond the array's limits."""
import ctypes as C

import numpy as np
import pytest

import libgail_amd as g

pytestmark = pytest.mark.gpu

MISUSE = 4


def _entry():
    lib = C.CDLL(g.PLUGIN_PATH)
    f = lib.gail_hip_decode_noalloc
    f.restype = C.c_int
    f.argtypes = [C.c_int, C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_size_t, C.c_void_p, C.c_size_t, C.c_void_p,
                  C.c_ssize_t, C.c_void_p]
    return f


def _encode(device, W, H, fourk, seed):
    """A frame coded at the DCI rate of 250 Mbit/s at 24 fps: rate control on the device
    keeps each block's chosen passes, so the data is what a codestream carries."""
    L = g.FrameLayout(W, H, fourk)
    b = g.Batch(device, [L], [g.with_rate_control(g.xyz_settings(L), fps=24.0, mbps=250.0)])
    try:
        b.upload(0, g.synth_frame(seed, W, H))
        outs, _ = b.encode()
    finally:
        b.close()
    return L, outs[0]


def _device_buffer(device, nbytes):
    p = g.plugin().device_alloc(device.handle, nbytes)
    assert p, "device allocation failed"
    return p


def _noalloc_decode(device, entry, L, reduce, data, blocks, from_device):
    """Decodes with the no-allocation entry into device memory; returns the picture on the host."""
    P = g.plugin()
    w, h = g.decoded_size(L, reduce)
    pitch = w * 6
    dst = _device_buffer(device, pitch * h)
    src = None
    try:
        data = np.ascontiguousarray(data, np.uint8)
        ptr = data.ctypes.data
        if from_device:
            src = _device_buffer(device, max(1, data.size))
            assert P.copy(C.c_void_p(src), data.ctypes.data_as(C.c_void_p), data.size) == 0
            ptr = src
        blocks = np.ascontiguousarray(blocks, dtype=g.CBLK_CODE)
        pl = L.plain()
        st = entry(0, C.addressof(pl), reduce, 1, ptr, data.size, blocks.ctypes.data, len(blocks), dst, pitch, None)
        assert st == 0, P.last_error().decode()
        out = np.zeros((h, w * 3), np.uint16)
        assert P.copy(out.ctypes.data_as(C.c_void_p), C.c_void_p(dst), pitch * h) == 0
        return out
    finally:
        P.device_free(C.c_void_p(dst))
        if src:
            P.device_free(C.c_void_p(src))


def test_no_alloc_decode_matches_decoder(device):
    entry = _entry()
    cases = [((2048, 1080, 0), 0, 21, False),   # 2K full: the largest frame the array takes
             ((1024, 512, 0), 1, 22, True)]     # then a smaller frame, half resolution, from HBM
    for (W, H, k), reduce, seed, from_device in cases:
        L, res = _encode(device, W, H, k, seed)
        blocks = g.block_table(res)
        dec = g.Decoder(device, L, reduce, True)
        try:
            want, _, _ = dec.decode(res.data, blocks)
        finally:
            dec.close()
        print(f"[no-alloc] {W}x{H} reduce {reduce}, codestream from {'HBM' if from_device else 'host memory'}: "
              f"{len(res.data)} bytes, {len(blocks)} blocks", flush=True)
        got = _noalloc_decode(device, entry, L, reduce, res.data, blocks, from_device)
        assert np.array_equal(got, want), f"{W}x{H}: the no-allocation decode differs from decoder_decode"


def test_no_alloc_refuses_beyond_limits(device):
    entry = _entry()
    P = g.plugin()
    L4 = g.FrameLayout(3996, 2160, 1)
    pl = L4.plain()
    blocks = np.zeros(L4.num_cblks, g.CBLK_CODE)
    st = entry(0, C.addressof(pl), 0, 1, None, 0, blocks.ctypes.data, len(blocks), C.c_void_p(16), 3996 * 6, None)
    assert st == MISUSE and b"2048x1080" in P.last_error()
    L2 = g.FrameLayout(1998, 1080, 0)
    pl = L2.plain()
    big = np.zeros(2 << 20, np.uint8)
    blocks = np.zeros(L2.num_cblks, g.CBLK_CODE)
    st = entry(0, C.addressof(pl), 0, 1, big.ctypes.data, big.size, blocks.ctypes.data, len(blocks), C.c_void_p(16),
               1998 * 6, None)
    assert st == MISUSE and b"byte limit" in P.last_error()
