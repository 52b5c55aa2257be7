"""ctypes bindings of the CPU oracle (oracle/liboracle.so) and, where it was built, of the
compiled reference hot path (oracle/_ref/libgailref.so). Test infrastructure: imported by
tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg only."""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libgailref.so")

MAX_PASSES = 64
BLOCK_BUF = 1024 * 4 * 24 + 64


class Band(C.Structure):
    _fields_ = [(n, C.c_int32) for n in ("comp", "res", "level", "orient", "px0", "py0", "w", "h", "ncbx",
                                          "ncby", "cblk_base", "eps", "mu", "mb")] + [
        ("delta", C.c_float), ("pad", C.c_float), ("weight", C.c_double)]


class Cblk(C.Structure):
    _fields_ = [(n, C.c_int32) for n in ("band", "x0", "y0", "w", "h")]


class T1Result(C.Structure):
    _fields_ = [("nbps", C.c_int32), ("num_passes", C.c_int32), ("len", C.c_uint32),
                ("pass_len", C.c_uint32 * MAX_PASSES), ("pass_dist", C.c_int32 * MAX_PASSES)]


class CblkInfo(C.Structure):
    _fields_ = [(n, C.c_uint32) for n in ("nbps", "num_passes", "data_offset", "data_len", "pass_base")]


class ColorCfg(C.Structure):
    """ColorParams of the managed route, in ColorParams' enumerations (or_color_cfg)."""
    _fields_ = [(n, C.c_int32) for n in ("transfer", "primaries", "tone_map", "tone_norm", "gamut",
                                          "lut3d_size")] + [
        (n, C.c_double) for n in ("sdr_white", "hdr_peak", "cinema_peak", "hdr_ref_white", "cinema_ref_white",
                                  "knee", "gamut_threshold")]


def color_cfg(transfer=0, primaries=0, tone_map=0, tone_norm=0, gamut=2, lut3d_size=0, sdr_white=48.0,
              hdr_peak=1000.0, cinema_peak=48.0, hdr_ref_white=203.0, cinema_ref_white=16.0, knee=0.25,
              gamut_threshold=0.8):
    """A ColorCfg with ColorParams' defaults (color.h:84-100) for the fields not given."""
    return ColorCfg(transfer, primaries, tone_map, tone_norm, gamut, lut3d_size, sdr_white, hdr_peak,
                    cinema_peak, hdr_ref_white, cinema_ref_white, knee, gamut_threshold)


def _p(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


# gail::CblkCode (t2_parse.h:17-22) as a numpy record: 12 bytes, as the C structure.
CBLK_CODE = np.dtype([("offset", "<u4"), ("len", "<u4"), ("passes", "u1"), ("nbps", "u1"), ("pad", "u1", 2)])


def block_codes(info, pass_len, passes=None):
    """The code-block table a Tier-2 reader would hand the decoder for a frame's Tier-1
    output coded into one codestream equal to its data buffer: each block's bytes from its
    data offset, keeping `passes[i]` passes (all when None) and their truncation length."""
    info = np.asarray(info, np.uint32).reshape(-1, 5)
    n = len(info)
    out = np.zeros(n, CBLK_CODE)
    out["offset"] = info[:, 2]
    out["nbps"] = info[:, 0]
    keep = info[:, 1] if passes is None else np.minimum(np.asarray(passes), info[:, 1])
    out["passes"] = keep
    pl = np.asarray(pass_len, np.uint32)
    base = info[:, 4].astype(np.int64)
    lens = np.where(keep > 0, pl[np.maximum(base + keep.astype(np.int64) - 1, 0)], 0)
    out["len"] = lens
    return out


class _Lib:
    """The oracle and the reference shim share one calling convention (prefix or_/ref_)."""

    def __init__(self, path, prefix):
        self.lib = C.CDLL(path)
        self.prefix = prefix
        L = self.lib
        f = getattr(L, prefix + "layout" if prefix == "ref_" else "or_layout_make")
        f.restype = C.c_int
        f.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, C.POINTER(Band), C.POINTER(C.c_int),
                      C.POINTER(Cblk), C.c_int]
        self._layout = f
        f = getattr(L, prefix + "t1_encode")
        f.restype = C.c_int
        f.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_float, C.c_float, C.c_void_p,
                      C.c_uint32, C.POINTER(T1Result), C.c_void_p, C.c_uint32, C.POINTER(C.c_uint32),
                      C.c_void_p]
        self._t1 = f
        f = getattr(L, prefix + "fdwt97_2d" if prefix == "ref_" else "or_fdwt97_2d_f32")
        f.restype = None
        f.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int]
        self._dwt = f
        f = getattr(L, prefix + "color_tables")
        f.restype = None
        f.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, C.c_void_p, C.c_void_p]
        self._tables = f
        f = getattr(L, prefix + "planar_to_rgb48")
        f.restype = None
        f.argtypes = [C.c_void_p] + [C.c_int] * 8 + [C.c_double, C.c_double] + [C.c_int] * 4 + [C.c_void_p]
        self._planar = f
        f = getattr(L, prefix + "encode_frame_xyz")
        f.restype = C.c_int64
        f.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_double, C.c_float, C.c_int, C.c_void_p,
                      C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64, C.c_void_p, C.c_void_p]
        self._frame = f
        f = getattr(L, prefix + "t1_decode_words")
        f.restype = C.c_int
        f.argtypes = [C.c_void_p, C.c_uint32] + [C.c_int] * 5 + [C.c_void_p]
        self._t1dw = f
        f = getattr(L, prefix + "decode_frame")
        f.restype = C.c_int
        f.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, C.c_void_p, C.c_size_t, C.c_void_p, C.c_int, C.c_int,
                      C.c_int, C.c_int, C.c_void_p, C.c_ssize_t, C.c_void_p]
        self._decf = f

    def t1_decode_words(self, data, w, h, orient, nbps, passes, length=None):
        """Reconstruction words of a w x h block (sign in bit 31, twice the magnitude plus
        the mid-point of the planes not decoded)."""
        buf = np.frombuffer(bytes(data), np.uint8).copy() if len(data) else np.zeros(1, np.uint8)
        n = len(data) if length is None else length
        out = np.zeros(w * h, np.uint32)
        if self._t1dw(_p(buf), n, w, h, orient, nbps, passes, _p(out)):
            raise ValueError("block size out of range")
        return out.reshape(h, w)

    def decode_frame(self, W, H, fourk, data, blocks, reduce=0, high_bits=True, base_step=1.0, threads=8,
                     pad_row=0, want_wavelet=False):
        """The CPU decode of one frame: (picture as rows of 3 x 16-bit words, wavelet planes
        before the inverse DWT or None)."""
        data = np.ascontiguousarray(np.frombuffer(bytes(data), np.uint8) if not isinstance(data, np.ndarray)
                                    else data, dtype=np.uint8)
        if data.size == 0:
            data = np.zeros(1, np.uint8)
        blocks = np.ascontiguousarray(blocks, dtype=CBLK_CODE)
        ow, oh = W, H
        for _ in range(reduce):
            ow, oh = (ow + 1) // 2, (oh + 1) // 2
        row = ow * 3 + pad_row
        out = np.zeros((oh, row), np.uint16)
        wav = np.zeros((3, H, W), np.float32) if want_wavelet else None
        rc = self._decf(W, H, int(fourk), base_step, _p(data), data.size, _p(blocks), len(blocks), reduce,
                        int(high_bits), threads, _p(out), row * 2, _p(wav))
        if rc:
            raise ValueError("decode refused the frame")
        return out, wav

    def layout(self, W, H, fourk, base_step=1.0):
        bands = (Band * 64)()
        nb = C.c_int(0)
        n = self._layout(W, H, int(fourk), base_step, bands, C.byref(nb), None, 0)
        if n < 0:
            raise ValueError("invalid layout")
        cbl = (Cblk * n)()
        self._layout(W, H, int(fourk), base_step, bands, C.byref(nb), cbl, n)
        return list(bands[:nb.value]), cbl

    def t1_encode(self, coeffs, w, h, orient, max_passes=MAX_PASSES, thr=0.0, scale=1.0, trace=False):
        coeffs = np.ascontiguousarray(coeffs, dtype=np.uint32)
        out = np.zeros(BLOCK_BUF, np.uint8)
        res = T1Result()
        syms = np.zeros(1 << 17, np.uint8) if trace else None
        pe = np.zeros(MAX_PASSES, np.uint32) if trace else None
        ns = C.c_uint32(0)
        rc = self._t1(_p(coeffs), w, h, orient, max_passes, thr, scale, _p(out), BLOCK_BUF, C.byref(res),
                      _p(syms), 1 << 17 if trace else 0, C.byref(ns), _p(pe))
        if rc:
            raise RuntimeError("t1 encode failed")
        np_ = res.num_passes
        r = dict(nbps=res.nbps, num_passes=np_, data=bytes(out[:res.len]),
                 pass_len=list(res.pass_len[:np_]), pass_dist=list(res.pass_dist[:np_]))
        if trace:
            r["symbols"] = bytes(syms[:ns.value])
            r["pass_end"] = list(pe[:np_])
        return r

    def fdwt(self, plane, levels):
        a = np.array(plane, dtype=np.float32, copy=True, order="C")
        self._dwt(_p(a), a.shape[1], a.shape[0], levels)
        return a

    def color_tables(self, transfer, primaries, sdr_white=48.0, cinema_peak=48.0):
        lut = np.zeros(65536, np.float32)
        cp = np.zeros(48, np.float32)
        self._tables(transfer, primaries, sdr_white, cinema_peak, _p(lut), _p(cp))
        return lut, cp

    def planar_to_rgb48(self, src, bits, aw, ah, sx, sy, rgb, full_range, chroma, kr, kb, W, H, ox, oy):
        src = np.ascontiguousarray(src)
        out = np.zeros((H, W, 3), np.uint16)
        self._planar(_p(src), bits, aw, ah, sx, sy, rgb, full_range, chroma, kr, kb, W, H, ox, oy, _p(out))
        return out

    def encode_frame_xyz(self, xyz, fourk, thr=0.0, threads=8, base_step=1.0, want_planes=False):
        xyz = np.ascontiguousarray(xyz, dtype=np.uint16)
        _, H, W = xyz.shape
        bands, cbl = self.layout(W, H, fourk, base_step)
        n = len(cbl)
        info = np.zeros((n, 5), np.uint32)
        pl = np.zeros(n * MAX_PASSES, np.uint32)
        pd = np.zeros(n * MAX_PASSES, np.int32)
        cap = n * 4096 + (64 << 20)
        data = np.zeros(cap, np.uint8)
        coeffs = np.zeros(n * 1024, np.uint32)
        planes = np.zeros((3, H, W), np.float32) if want_planes else None
        total = self._frame(_p(xyz), W, H, int(fourk), base_step, thr, threads, _p(info), _p(pl), _p(pd),
                            _p(data), cap, _p(coeffs), _p(planes))
        if total < 0:
            raise RuntimeError("frame encode failed")
        npass = int(info[:, 1].sum())
        return dict(info=info, pass_len=pl[:npass], pass_dist=pd[:npass], data=data[:total], coeffs=coeffs,
                    planes=planes)


class Oracle(_Lib):
    def __init__(self, path=ORACLE_SO):
        super().__init__(path, "or_")
        L = self.lib
        L.or_t1_decode.restype = C.c_int
        L.or_t1_decode.argtypes = [C.c_void_p, C.c_uint32] + [C.c_int] * 5 + [C.c_void_p] * 3
        L.or_fdwt97_2d_f64.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int]
        L.or_idwt97_2d_f64.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int]
        L.or_quantize_coeff.restype = C.c_uint32
        L.or_quantize_coeff.argtypes = [C.c_float, C.c_float, C.c_uint32]
        L.or_quantize_block.argtypes = [C.c_void_p] + [C.c_int] * 5 + [C.c_float, C.c_int, C.c_void_p]
        L.or_synthesis_norm.restype = C.c_double
        L.or_synthesis_norm.argtypes = [C.c_int, C.c_int]
        L.or_convert_frame.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_void_p,
                                       C.c_void_p]
        L.or_xyz_planar_to_ict.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_void_p]
        L.or_color_pixel.argtypes = [C.c_void_p, C.c_void_p, C.c_uint16, C.c_uint16, C.c_uint16, C.c_void_p]
        L.or_color_setup.argtypes = [C.POINTER(ColorCfg), C.c_void_p, C.c_void_p]
        L.or_color_pixels.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p]
        L.or_xyz12_to_16.restype = C.c_uint16
        L.or_xyz12_to_16.argtypes = [C.c_uint32]
        L.or_idwt97_2d_f32.restype = None
        L.or_idwt97_2d_f32.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]

    def idwt(self, planes, levels, skip=0):
        """Float inverse DWT of a (n, H, W) stack, levels down to skip + 1."""
        a = np.array(planes, dtype=np.float32, copy=True, order="C")
        n, H, W = a.shape
        ptrs = (C.c_void_p * n)(*[a[i].ctypes.data for i in range(n)])
        self.lib.or_idwt97_2d_f32(ptrs, n, W, H, levels, skip)
        return a

    def t1_decode(self, data, w, h, orient, nbps, passes, length=None):
        buf = np.frombuffer(bytes(data), np.uint8).copy() if len(data) else np.zeros(1, np.uint8)
        n = len(data) if length is None else length
        mag = np.zeros(w * h, np.uint32)
        sign = np.zeros(w * h, np.uint8)
        sig = np.zeros(w * h, np.uint8)
        self.lib.or_t1_decode(_p(buf), n, w, h, orient, nbps, passes, _p(mag), _p(sign), _p(sig))
        return mag.reshape(h, w), sign.reshape(h, w), sig.reshape(h, w)

    def fdwt_f64(self, plane, levels):
        a = np.array(plane, dtype=np.float64, copy=True, order="C")
        self.lib.or_fdwt97_2d_f64(_p(a), a.shape[1], a.shape[0], levels)
        return a

    def idwt_f64(self, plane, levels):
        a = np.array(plane, dtype=np.float64, copy=True, order="C")
        self.lib.or_idwt97_2d_f64(_p(a), a.shape[1], a.shape[0], levels)
        return a

    def quantize_coeff(self, v, inv_delta, max_mag):
        return self.lib.or_quantize_coeff(v, inv_delta, max_mag)

    def quantize_block(self, plane, x0, y0, w, h, delta, mb):
        plane = np.ascontiguousarray(plane, dtype=np.float32)
        out = np.zeros(1024, np.uint32)
        self.lib.or_quantize_block(_p(plane), plane.shape[1], x0, y0, w, h, delta, mb, _p(out))
        return out

    def synthesis_norm(self, level, orient):
        return self.lib.or_synthesis_norm(level, orient)

    def convert_frame(self, lut, cp, rgb48):
        rgb48 = np.ascontiguousarray(rgb48, dtype=np.uint16)
        H, W, _ = rgb48.shape
        planes = np.zeros((3, H, W), np.float32)
        xyz = np.zeros((H, W, 3), np.uint16)
        self.lib.or_convert_frame(_p(lut), _p(cp), _p(rgb48), W, H, _p(planes), _p(xyz))
        return planes, xyz

    def xyz_to_ict(self, xyz):
        xyz = np.ascontiguousarray(xyz, dtype=np.uint16)
        _, H, W = xyz.shape
        planes = np.zeros((3, H, W), np.float32)
        self.lib.or_xyz_planar_to_ict(_p(xyz), W, H, _p(planes))
        return planes

    def color_setup(self, cfg):
        lut = np.zeros(65536, np.float32)
        cp = np.zeros(48, np.float32)
        self.lib.or_color_setup(C.byref(cfg), _p(lut), _p(cp))
        return lut, cp

    def color_pixels(self, lut, cp, lut3, rgb48):
        rgb48 = np.ascontiguousarray(rgb48, dtype=np.uint16).reshape(-1, 3)
        out = np.zeros((len(rgb48), 3), np.int32)
        self.lib.or_color_pixels(_p(lut), _p(cp), _p(lut3), len(rgb48), _p(rgb48), _p(out))
        return out

    def color_pixel(self, lut, cp, r, g, b):
        out = np.zeros(3, np.int32)
        self.lib.or_color_pixel(_p(lut), _p(cp), r, g, b, _p(out))
        return tuple(int(v) for v in out)


class Reference(_Lib):
    def __init__(self, path=REF_SO):
        super().__init__(path, "ref_")
        L = self.lib
        L.ref_synthesis_norm.restype = C.c_double
        L.ref_synthesis_norm.argtypes = [C.c_int, C.c_int]
        L.ref_quantize_coeff.restype = C.c_uint32
        L.ref_quantize_coeff.argtypes = [C.c_float, C.c_float, C.c_uint32]
        L.ref_color_setup.restype = C.c_int
        L.ref_color_setup.argtypes = [C.POINTER(ColorCfg), C.c_void_p, C.c_void_p, C.c_void_p]
        L.ref_color_pixels.restype = C.c_int
        L.ref_color_pixels.argtypes = [C.POINTER(ColorCfg), C.c_void_p, C.c_int, C.c_void_p, C.c_void_p]
        L.ref_convert_frame.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, C.c_void_p, C.c_int, C.c_int,
                                        C.c_void_p, C.c_void_p]
        L.ref_idwt97_2d.restype = None
        L.ref_idwt97_2d.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]
        L.ref_check_codestream.restype = C.c_int
        L.ref_check_codestream.argtypes = [C.c_void_p, C.c_size_t, C.c_double, C.c_char_p, C.c_int, C.c_void_p]
        L.ref_parse_blocks.restype = C.c_int
        L.ref_parse_blocks.argtypes = [C.c_void_p, C.c_size_t, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                       C.POINTER(C.c_int), C.c_void_p, C.c_int, C.POINTER(Band),
                                       C.POINTER(C.c_int)]

    def check_codestream(self, cs, fps):
        """The reference's checkDciCodestream (codestream_check.cpp:43-280): (ok, errors, facts),
        facts a dict of total / comp bytes, tile-parts, rsiz, width, height and levels."""
        buf = np.frombuffer(cs, np.uint8)
        msg = C.create_string_buffer(4096)
        f = np.zeros(9, np.uint64)
        r = self.lib.ref_check_codestream(buf.ctypes.data, len(buf), fps, msg, 4096, f.ctypes.data)
        if r < 0:
            raise RuntimeError("the reference's checker raised an exception")
        keys = ("total", "comp0", "comp1", "comp2", "tile_parts", "rsiz", "width", "height", "levels")
        return r == 1, msg.value.decode(), {k: int(v) for k, v in zip(keys, f)}

    def idwt(self, planes, levels, skip=0):
        a = np.array(planes, dtype=np.float32, copy=True, order="C")
        n, H, W = a.shape
        self.lib.ref_idwt97_2d(_p(a), n, W, H, levels, skip)
        return a

    def parse_blocks(self, codestream):
        """The reference's header and packet parsing of a codestream: (W, H, fourk, block
        table, bands with the signalled steps), or None when it refuses the stream."""
        data = np.frombuffer(bytes(codestream), np.uint8).copy()
        W, H, fk, nb = C.c_int(0), C.c_int(0), C.c_int(0), C.c_int(0)
        bands = (Band * 64)()
        n = self.lib.ref_parse_blocks(_p(data), data.size, C.byref(W), C.byref(H), C.byref(fk), None, 0, bands,
                                      C.byref(nb))
        if n < 0:
            return None
        blocks = np.zeros(n, CBLK_CODE)
        self.lib.ref_parse_blocks(_p(data), data.size, C.byref(W), C.byref(H), C.byref(fk), _p(blocks), n, bands,
                                  C.byref(nb))
        return W.value, H.value, bool(fk.value), blocks, list(bands[:nb.value])

    def synthesis_norm(self, level, orient):
        return self.lib.ref_synthesis_norm(level, orient)

    def color_setup(self, cfg, lut3=None):
        lut = np.zeros(65536, np.float32)
        cp = np.zeros(48, np.float32)
        if self.lib.ref_color_setup(C.byref(cfg), _p(lut3), _p(lut), _p(cp)):
            raise ValueError("reference rejected the colour parameters")
        return lut, cp

    def color_pixels(self, cfg, lut3, rgb48):
        rgb48 = np.ascontiguousarray(rgb48, dtype=np.uint16).reshape(-1, 3)
        out = np.zeros((len(rgb48), 3), np.int32)
        if self.lib.ref_color_pixels(C.byref(cfg), _p(lut3), len(rgb48), _p(rgb48), _p(out)):
            raise ValueError("reference rejected the colour parameters")
        return out

    def quantize_coeff(self, v, inv_delta, max_mag):
        return self.lib.ref_quantize_coeff(v, inv_delta, max_mag)

    def convert_frame(self, transfer, primaries, rgb48, sdr_white=48.0, cinema_peak=48.0):
        rgb48 = np.ascontiguousarray(rgb48, dtype=np.uint16)
        H, W, _ = rgb48.shape
        planes = np.zeros((3, H, W), np.float32)
        xyz = np.zeros((H, W, 3), np.uint16)
        self.lib.ref_convert_frame(transfer, primaries, sdr_white, cinema_peak, _p(rgb48), W, H, _p(planes),
                                   _p(xyz))
        return planes, xyz


def oracle():
    return Oracle()


def reference():
    """The compiled reference, or None where it was not built."""
    return Reference() if os.path.exists(REF_SO) else None


def random_block(rng, trial):
    """A seeded test block in the manner of the reference's Tier-1 test (ragged sizes, four
    orientations, three magnitude regimes: small, sparse and large)."""
    w = 1 + int(rng.integers(32)) if trial % 7 == 0 else 32
    h = 1 + int(rng.integers(32)) if trial % 5 == 0 else 32
    orient = trial % 4
    mode = trial % 3
    sd = (20.0, 300.0, 3000.0)[mode]
    v = rng.normal(0, sd, size=(h, w))
    if mode == 1:
        v = np.where(rng.random((h, w)) < 0.7, v * 0.05, v)
    v = v.astype(np.float32)
    c = np.zeros((32, 32), np.uint32)
    a = np.abs(v) * np.float32(64.0)
    m = np.where(a >= np.float32((1 << 26) - 1), (1 << 26) - 1, a.astype(np.uint32))
    m = m | np.where((v < 0) & (m >= 64), np.uint32(0x80000000), np.uint32(0))
    c[:h, :w] = m
    return c, w, h, orient
