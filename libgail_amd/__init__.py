"""Python access to the HIP encoder backend of GAIL (MI355X / gfx950).

The compute path is native: lib/libgail_hip.so is the backend plugin (HIP kernels behind the
plain-C table of include/gail_hip_plugin.h) and lib/libgail.so the host library. This module
only binds them with ctypes, for tests and the benchmark; it raises if they are missing and
never falls back to anything else.
"""
import ctypes as C
import os
import threading
import weakref

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# libgail.so and the plugin beside it; GAIL_AMD_LIB_DIR names another build of the pair
LIB_DIR = os.environ.get("GAIL_AMD_LIB_DIR") or os.path.join(HERE, "lib")
PLUGIN_PATH = os.path.join(LIB_DIR, "libgail_hip.so")
HOST_PATH = os.path.join(LIB_DIR, "libgail.so")
ABI = 2
MAX_PASSES = 64


class GailError(RuntimeError):
    def __init__(self, status, message):
        super().__init__(f"{message} (status {status})")
        self.status = status


class Subband(C.Structure):
    _fields_ = [(n, C.c_int32) for n in ("comp", "res", "level", "orient", "px0", "py0", "w", "h", "ncbx", "ncby",
                                          "cblkBase", "eps", "mu", "Mb")] + [("delta", C.c_float),
                                                                             ("weight", C.c_double)]


class Layout(C.Structure):
    _fields_ = [(n, C.c_int32) for n in ("profile", "width", "height", "levels", "guardBits", "numBands",
                                          "numCblks")] + [("baseStep", C.c_double), ("bands", C.POINTER(Subband))]


class FrameFormat(C.Structure):
    _fields_ = [(n, C.c_int32) for n in ("kind", "chroma", "width", "height", "activeW", "activeH", "ox", "oy",
                                          "sx", "sy", "rgb", "bits", "fullRange", "pad")] + [("kr", C.c_double),
                                                                                             ("kb", C.c_double)]


class Color(C.Structure):
    """GailHipColor: ColorParams in its own enumerations (gamut 0 P3-D65, 1 DCI-P3, 2 none)."""
    _fields_ = [(n, C.c_int32) for n in ("transfer", "primaries", "toneMap", "toneNorm", "gamut", "lut3dSize")] + [
        (n, C.c_double) for n in ("sdrWhite", "hdrPeak", "cinemaPeak", "hdrRefWhite", "cinemaRefWhite", "knee",
                                  "gamutThreshold")] + [
        ("lut3d", C.c_char_p), ("lut3dData", C.c_void_p), ("lut3dCount", C.c_uint64)]


def set_color(color, transfer=0, primaries=0, tone_map=0, tone_norm=0, gamut=2, sdr_white=48.0, hdr_peak=1000.0,
              cinema_peak=48.0, hdr_ref_white=203.0, cinema_ref_white=16.0, knee=0.25, gamut_threshold=0.8,
              lut3d=None):
    """Fills a Color with ColorParams' defaults for the members not given; lut3d: a float32
    array of n^3 x 3 values (red fastest), kept alive by the caller until the frame is bound."""
    color.transfer, color.primaries, color.toneMap, color.toneNorm, color.gamut = (transfer, primaries, tone_map,
                                                                                  tone_norm, gamut)
    color.sdrWhite, color.hdrPeak, color.cinemaPeak = sdr_white, hdr_peak, cinema_peak
    color.hdrRefWhite, color.cinemaRefWhite, color.knee, color.gamutThreshold = (hdr_ref_white, cinema_ref_white,
                                                                                 knee, gamut_threshold)
    if lut3d is not None:
        n = round((len(lut3d) // 3) ** (1.0 / 3.0))
        color.lut3dSize, color.lut3dData, color.lut3dCount = n, lut3d.ctypes.data, len(lut3d)
    return color


class EncodeSettings(C.Structure):
    _fields_ = [("profile", C.c_int32), ("width", C.c_int32), ("height", C.c_int32), ("baseStep", C.c_double),
                ("slopeFactor", C.c_double), ("bitsPerSymbol", C.c_double), ("input", FrameFormat),
                ("color", Color), ("rcMaxTotal", C.c_uint64), ("rcMaxComp", C.c_uint64)]


class Rate(C.Structure):
    _fields_ = [("lambda_", C.c_double * 3), ("total", C.c_uint64), ("comp", C.c_uint64 * 3),
                ("tilePart", C.c_uint64 * 6), ("distortion", C.c_double), ("iterations", C.c_uint32),
                ("numTileParts", C.c_uint32)]


class Memory(C.Structure):
    _fields_ = [("total", C.c_uint64), ("available", C.c_uint64), ("budgeted", C.c_int32)]


class DeviceDesc(C.Structure):
    _fields_ = [("index", C.c_int32), ("computeUnits", C.c_int32), ("memoryBytes", C.c_uint64),
                ("name", C.c_char * 256), ("arch", C.c_char * 64), ("pciBusId", C.c_char * 32)]


class CblkInfo(C.Structure):
    _fields_ = [(n, C.c_uint32) for n in ("nbps", "numPasses", "dataOffset", "dataLen", "passBase")]


class Tier1(C.Structure):
    _fields_ = [("info", C.POINTER(CblkInfo)), ("passLen", C.POINTER(C.c_uint32)),
                ("passDist", C.POINTER(C.c_int32)), ("data", C.POINTER(C.c_uint8)), ("numInfo", C.c_uint64),
                ("numPasses", C.c_uint64), ("numData", C.c_uint64)]


class FrameTimings(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("upload", "color", "dwt", "quant", "t1", "t1cm", "t1mq", "download",
                                          "gpu", "total", "symbols", "codedBytes")]


class Debug(C.Structure):
    _fields_ = [("full", C.c_int32), ("xyz12", C.POINTER(C.c_uint16)), ("coeffs", C.POINTER(C.c_uint32)),
                ("wavelet", C.POINTER(C.c_float)), ("numXyz12", C.c_uint64), ("numCoeffs", C.c_uint64),
                ("numWavelet", C.c_uint64)]


class BatchOptions(C.Structure):
    _fields_ = [("dwtTile", C.c_int32), ("cmWaves", C.c_int32), ("mqGroup", C.c_int32), ("useGraph", C.c_int32),
                ("mqRead", C.c_int32), ("reserved", C.c_int32), ("poolBytesPerCoef", C.c_double)]


class Slot(C.Structure):
    _fields_ = [("maxWidth", C.c_int32), ("maxHeight", C.c_int32), ("colour", C.c_int32), ("reserved", C.c_int32)]


class DecodeSettings(C.Structure):
    _fields_ = [("reduce", C.c_int32), ("highBits", C.c_int32)]


class CblkCode(C.Structure):
    _fields_ = [("offset", C.c_uint32), ("len", C.c_uint32), ("passes", C.c_uint8), ("nbps", C.c_uint8),
                ("pad", C.c_uint8 * 2)]


# gail::CblkCode as a numpy record (12 bytes, GailHipCblkCode)
CBLK_CODE = np.dtype([("offset", "<u4"), ("len", "<u4"), ("passes", "u1"), ("nbps", "u1"), ("pad", "u1", 2)])


class DecodeTimings(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("t1", "dwt", "color", "total")] + [
        ("wantWavelet", C.c_int32), ("wavelet", C.POINTER(C.c_float) * 3), ("numWavelet", C.c_uint64)]


_vp = C.c_void_p
_F = C.CFUNCTYPE


class Plugin(C.Structure):
    _fields_ = [
        ("abi", C.c_uint32), ("size", C.c_uint32), ("layout", C.c_uint64), ("name", C.c_char_p),
        ("toolkit", C.c_char_p), ("archs", C.c_char_p),
        ("last_error", _F(C.c_char_p)),
        ("device_count", _F(C.c_int)),
        ("device_desc", _F(C.c_int, C.c_int, C.POINTER(DeviceDesc))),
        ("device_usable", _F(C.c_int, C.c_int)),
        ("device_create", _F(C.c_int, C.c_int, _vp, _vp, C.POINTER(_vp))),
        ("device_destroy", _F(None, _vp)),
        ("device_name", _F(C.c_char_p, _vp)),
        ("device_memory", _F(C.c_int, _vp, C.POINTER(Memory))),
        ("encoder_bytes", _F(C.c_uint64, _vp, C.POINTER(Layout), C.POINTER(EncodeSettings))),
        ("encoder_create", _F(C.c_int, _vp, C.POINTER(Layout), C.POINTER(EncodeSettings), C.POINTER(_vp))),
        ("encoder_destroy", _F(None, _vp)),
        ("encoder_input", _F(_vp, _vp)),
        ("encoder_set_slope", _F(None, _vp, C.c_double)),
        ("encoder_encode", _F(C.c_int, _vp, _vp, C.POINTER(Tier1), C.POINTER(FrameTimings), C.POINTER(Debug))),
        ("decoder_bytes", _F(C.c_uint64, _vp, C.POINTER(Layout), C.POINTER(DecodeSettings))),
        ("decoder_create", _F(C.c_int, _vp, C.POINTER(Layout), C.POINTER(DecodeSettings), C.POINTER(_vp))),
        ("decoder_destroy", _F(None, _vp)),
        ("decoder_decode", _F(C.c_int, _vp, C.POINTER(Layout), _vp, C.c_size_t, _vp, C.c_size_t, _vp, C.c_ssize_t,
                              C.POINTER(DecodeTimings))),
        ("batch_bytes", _F(C.c_uint64, _vp, C.POINTER(Layout), C.POINTER(EncodeSettings), C.c_int,
                           C.POINTER(BatchOptions))),
        ("batch_create", _F(C.c_int, _vp, C.POINTER(Layout), C.POINTER(EncodeSettings), C.c_int,
                            C.POINTER(BatchOptions), C.POINTER(_vp))),
        ("batch_destroy", _F(None, _vp)),
        ("batch_input", _F(_vp, _vp, C.c_int)),
        ("batch_upload", _F(C.c_int, _vp, C.c_int, _vp)),
        ("batch_set_slope", _F(None, _vp, C.c_int, C.c_double)),
        ("batch_encode", _F(C.c_int, _vp, C.c_int, C.POINTER(Tier1), C.POINTER(FrameTimings))),
        ("batch_launch", _F(C.c_int, _vp, C.c_int)),
        ("batch_wait", _F(C.c_int, _vp, C.POINTER(Tier1), C.POINTER(FrameTimings))),
        ("batch_debug", _F(C.c_int, _vp, C.c_int, C.POINTER(Debug))),
        ("batch_stage_ms", _F(C.c_int, _vp, C.POINTER(C.c_double), C.c_int)),
        ("slots_bytes", _F(C.c_uint64, _vp, C.POINTER(Slot), C.c_int, C.POINTER(BatchOptions))),
        ("slots_create", _F(C.c_int, _vp, C.POINTER(Slot), C.c_int, C.POINTER(BatchOptions), C.POINTER(_vp))),
        ("batch_bind", _F(C.c_int, _vp, C.POINTER(Layout), C.POINTER(EncodeSettings), C.c_int)),
        ("batch_query", _F(C.c_int, _vp)),
        ("batch_upload_from", _F(C.c_int, _vp, C.c_int, _vp)),
        ("host_alloc", _F(_vp, C.c_uint64)),
        ("host_free", _F(None, _vp)),
        ("batch_autotune", _F(C.c_int, _vp, C.c_int, C.POINTER(BatchOptions), C.POINTER(C.c_double), C.c_int)),
        ("device_alloc", _F(_vp, _vp, C.c_uint64)),
        ("device_free", _F(None, _vp)),
        ("copy", _F(C.c_int, _vp, _vp, C.c_uint64)),
        ("is_device_pointer", _F(C.c_int, _vp)),
        ("batch_rate", _F(C.c_int, _vp, C.c_int, C.POINTER(Rate))),
        ("batch_inputs_ready", _F(C.c_int, _vp)),
        ("encoder_rate", _F(C.c_int, _vp, _vp)),
    ]


# DCI containers: flat 1998x1080 / 3996x2160, scope 2048x858 / 4096x1716, full 2048x1080 / 4096x2160
CONTAINERS = {"2k_flat": (1998, 1080, 0), "2k_scope": (2048, 858, 0), "2k_full": (2048, 1080, 0),
              "4k_flat": (3996, 2160, 1), "4k_scope": (4096, 1716, 1), "4k_full": (4096, 2160, 1)}

_host = None
_plugin = None
_plugin_lib = None


def host():
    """libgail.so (layout, synthetic frames). Raises if it was not built."""
    global _host
    if _host is None:
        if not os.path.exists(HOST_PATH):
            raise FileNotFoundError(f"{HOST_PATH} is missing: run __graft_entry__.build()")
        h = C.CDLL(HOST_PATH)
        h.gail_layout_make.restype = C.c_int
        h.gail_layout_make.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, C.POINTER(Subband), C.c_int,
                                       C.POINTER(C.c_int)]
        h.gail_layout_cblks.restype = C.c_int
        h.gail_layout_cblks.argtypes = [C.c_int, C.c_int, C.c_int, _vp, C.c_int]
        h.gail_synth_xyz12.restype = None
        h.gail_synth_xyz12.argtypes = [C.c_uint64, C.c_int, C.c_int, _vp, C.c_int]
        h.gail_encode_codestream.restype = C.c_int64
        h.gail_encode_codestream.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, _vp, _vp, _vp, _vp, C.c_uint64,
                                             C.c_double, C.c_double, _vp, C.c_uint64, C.POINTER(C.c_double),
                                             C.POINTER(C.c_double), C.c_char_p]
        h.gail_check_codestream.restype = C.c_int
        h.gail_check_codestream.argtypes = [_vp, C.c_uint64, C.c_double, C.c_int, C.c_char_p, C.c_int]
        h.gail_empty_codestream_bytes.restype = C.c_uint64
        h.gail_empty_codestream_bytes.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double]
        _host = h
    return _host


def plugin():
    """The HIP plugin's table, loaded the way libgail loads it (by full path, ABI and
    layout checked). Raises if the plugin is missing or from another build."""
    global _plugin, _plugin_lib
    if _plugin is None:
        if not os.path.exists(PLUGIN_PATH):
            raise FileNotFoundError(f"{PLUGIN_PATH} is missing: run __graft_entry__.build()")
        lib = C.CDLL(PLUGIN_PATH, mode=os.RTLD_NOW | os.RTLD_LOCAL)
        lib.gail_hip_plugin_get.restype = C.POINTER(Plugin)
        lib.gail_hip_plugin_get.argtypes = [C.c_uint32]
        p = lib.gail_hip_plugin_get(ABI)
        if not p:
            raise RuntimeError("libgail_hip.so was built for another plugin ABI")
        t = p.contents
        if t.abi != ABI or t.size != C.sizeof(Plugin):
            raise RuntimeError("libgail_hip.so does not match this binding's plugin table")
        _plugin_lib, _plugin = lib, t
    return _plugin


def synth_frame(seed, width, height, threads=8):
    """Three planar 12-bit X'Y'Z' channels (uint16, shape (3, H, W)) from the seeded generator."""
    a = np.zeros((3, height, width), np.uint16)
    host().gail_synth_xyz12(C.c_uint64(seed), width, height, a.ctypes.data_as(_vp), threads)
    return a


class FrameLayout:
    """makeLayout of a DCI frame in the plain form the plugin takes."""

    def __init__(self, width, height, fourk, base_step=1.0):
        self.width, self.height, self.fourk, self.base_step = width, height, int(fourk), base_step
        self.bands = (Subband * 64)()
        nb = C.c_int(0)
        n = host().gail_layout_make(width, height, self.fourk, base_step, self.bands, 64, C.byref(nb))
        if n < 0:
            raise ValueError(f"invalid frame size {width}x{height}")
        self.num_bands, self.num_cblks = nb.value, n
        self.levels = 6 if fourk else 5
        cb = np.zeros((n, 5), np.int32)
        host().gail_layout_cblks(width, height, self.fourk, cb.ctypes.data_as(_vp), n)
        self.cblks = cb

    def plain(self):
        return Layout(self.fourk, self.width, self.height, self.levels, 2 if self.fourk else 1, self.num_bands,
                      self.num_cblks, self.base_step, C.cast(self.bands, C.POINTER(Subband)))


def xyz_settings(layout):
    s = EncodeSettings()
    s.profile, s.width, s.height, s.baseStep = layout.fourk, layout.width, layout.height, layout.base_step
    s.input.kind, s.input.width, s.input.height = 2, layout.width, layout.height
    set_color(s.color, transfer=5, primaries=4)
    return s


def planar_settings(layout, bits, sx, sy, rgb, full_range, chroma=0, kr=0.2126, kb=0.0722, transfer=0,
                    primaries=0, active=None, offset=(0, 0)):
    s = EncodeSettings()
    s.profile, s.width, s.height, s.baseStep = layout.fourk, layout.width, layout.height, layout.base_step
    f = s.input
    f.kind, f.width, f.height = 1, layout.width, layout.height
    aw, ah = active if active else (layout.width, layout.height)
    f.activeW, f.activeH, f.ox, f.oy = aw, ah, offset[0], offset[1]
    f.sx, f.sy, f.rgb, f.bits, f.fullRange, f.chroma, f.kr, f.kb = sx, sy, rgb, bits, full_range, chroma, kr, kb
    set_color(s.color, transfer=transfer, primaries=primaries)
    return s


def rgb48_settings(layout, **color):
    """Interleaved 16-bit R'G'B' at the container size (the reference's CUDA slot input),
    colour by set_color's keywords."""
    s = EncodeSettings()
    s.profile, s.width, s.height, s.baseStep = layout.fourk, layout.width, layout.height, layout.base_step
    s.input.kind, s.input.width, s.input.height = 0, layout.width, layout.height
    set_color(s.color, **color)
    return s


def _check(status, what):
    if status != 0:
        raise GailError(status, f"{what}: {plugin().last_error().decode(errors='replace')}")


class Tier1Result:
    """One frame's Tier-1 output (gail::Tier1Frame): a copy, or with copy=False views of the
    encoder's own buffers, valid until its next encode or close (the plugin's contract)."""

    def __init__(self, t, copy=True):
        n, npass, nd = t.numInfo, t.numPasses, t.numData
        own = (lambda a: a.copy()) if copy else (lambda a: a)
        self.info = own(np.ctypeslib.as_array(C.cast(t.info, C.POINTER(C.c_uint32)), (n * 5,)).reshape(n, 5)) \
            if n else np.zeros((0, 5), np.uint32)
        self.pass_len = own(np.ctypeslib.as_array(t.passLen, (npass,))) if npass else np.zeros(0, np.uint32)
        self.pass_dist = own(np.ctypeslib.as_array(t.passDist, (npass,))) if npass else np.zeros(0, np.int32)
        self.data = own(np.ctypeslib.as_array(t.data, (nd,))) if nd else np.zeros(0, np.uint8)

    def block(self, i):
        nbps, np_, off, ln, pb = (int(v) for v in self.info[i])
        return dict(nbps=nbps, num_passes=np_, data=bytes(self.data[off:off + ln]),
                    pass_len=[int(v) for v in self.pass_len[pb:pb + np_]],
                    pass_dist=[int(v) for v in self.pass_dist[pb:pb + np_]])


class Device:
    def __init__(self, index=0):
        p = plugin()
        n = p.device_count()
        if n <= 0:
            raise GailError(1, "no HIP device")
        _check(p.device_usable(index), "device_usable")
        self.handle = _vp()
        _check(p.device_create(index, None, None, C.byref(self.handle)), "device_create")

    def name(self):
        return plugin().device_name(self.handle).decode()

    def memory(self):
        m = Memory()
        _check(plugin().device_memory(self.handle, C.byref(m)), "device_memory")
        return m.total, m.available

    def close(self):
        if self.handle:
            plugin().device_destroy(self.handle)
            self.handle = None


class Encoder:
    """One frame at a time (the reference's EncoderSlot)."""

    def __init__(self, device, layout, settings):
        self.layout, self.settings = layout, settings
        self._plain = layout.plain()
        self.handle = _vp()
        _check(plugin().encoder_create(device.handle, C.byref(self._plain), C.byref(settings), C.byref(self.handle)),
               "encoder_create")

    def set_slope(self, t):
        plugin().encoder_set_slope(self.handle, t)

    def encode(self, frame, debug=False, full=True, copy=True):
        """Tier-1 of one frame; copy=False returns views valid until the next encode."""
        frame = np.ascontiguousarray(frame)
        out, tm = Tier1(), FrameTimings()
        dbg = Debug(1 if full else 0) if debug else None
        _check(plugin().encoder_encode(self.handle, frame.ctypes.data_as(_vp), C.byref(out), C.byref(tm),
                                       C.byref(dbg) if dbg else None), "encoder_encode")
        res = Tier1Result(out, copy=copy)
        res.timings = {k: getattr(tm, k) for k, _ in FrameTimings._fields_}
        if dbg:
            res.debug = _debug_arrays(dbg, self.layout)
        return res

    def close(self):
        if self.handle:
            plugin().encoder_destroy(self.handle)
            self.handle = None


def decoded_size(layout, reduce):
    """decodedSize (decode_backend.h:33-40): the picture of a layout with `reduce` levels
    left out."""
    w, h = layout.width, layout.height
    for _ in range(reduce):
        w, h = (w + 1) // 2, (h + 1) // 2
    return w, h


def block_table(t1):
    """The code-block table (CBLK_CODE records) of a frame's Tier-1 output read as its own
    codestream: each block's bytes at its data offset with the passes the frame kept (all of
    them, or those device rate control chose)."""
    info = np.asarray(t1.info, np.uint32).reshape(-1, 5)
    out = np.zeros(len(info), CBLK_CODE)
    out["offset"], out["len"] = info[:, 2], info[:, 3]
    out["passes"], out["nbps"] = np.minimum(info[:, 1], 255), info[:, 0]
    return out


class Decoder:
    """One frame at a time (the reference's DecoderSlot) through the plugin's decoder entries:
    code-block table + codestream in, X'Y'Z' picture (three 16-bit words per pixel) out."""

    def __init__(self, device, layout, reduce=0, high_bits=True):
        self.layout, self.reduce = layout, reduce
        self.settings = DecodeSettings(reduce, 1 if high_bits else 0)
        self._plain = layout.plain()
        self.handle = _vp()
        _check(plugin().decoder_create(device.handle, C.byref(self._plain), C.byref(self.settings),
                                       C.byref(self.handle)), "decoder_create")

    @staticmethod
    def bytes_for(device, layout, reduce=0):
        pl = layout.plain()
        return plugin().decoder_bytes(device.handle, C.byref(pl), C.byref(DecodeSettings(reduce, 1)))

    def decode(self, data, blocks, frame_layout=None, out=None, stride_bytes=None, wavelet=False):
        """Decodes one frame. data: codestream bytes (numpy uint8 or an int device address
        with its size as (address, size)); blocks: CBLK_CODE records, one per code-block.
        Returns (picture as uint16 rows of 3 x width words, timings dict, wavelet planes or
        None). out / stride_bytes: a host array or (device address, pitch) to decode into."""
        fl = (frame_layout or self.layout).plain()
        if isinstance(data, tuple):
            ptr, size = data
        else:
            data = np.ascontiguousarray(data, dtype=np.uint8)
            ptr, size = data.ctypes.data, data.size
        blocks = np.ascontiguousarray(blocks, dtype=CBLK_CODE)
        w, h = decoded_size(self.layout, self.reduce)
        if out is None:
            out = np.zeros((h, w * 3), np.uint16)
            dst, pitch = out.ctypes.data, w * 6
        elif isinstance(out, tuple):
            dst, pitch = out
        else:
            dst, pitch = out.ctypes.data, stride_bytes or out.strides[0]
        tm = DecodeTimings()
        tm.wantWavelet = 1 if wavelet else 0
        _check(plugin().decoder_decode(self.handle, C.byref(fl), _vp(ptr), size, blocks.ctypes.data_as(_vp),
                                       len(blocks), _vp(dst), pitch, C.byref(tm)), "decoder_decode")
        wav = None
        if wavelet:
            n = tm.numWavelet
            wav = np.stack([np.ctypeslib.as_array(tm.wavelet[c], (n,)).copy() for c in range(3)])
            wav = wav.reshape(3, self.layout.height, self.layout.width)
        times = {k: getattr(tm, k) for k in ("t1", "dwt", "color", "total")}
        return (out if not isinstance(out, tuple) else None), times, wav

    def close(self):
        if self.handle:
            plugin().decoder_destroy(self.handle)
            self.handle = None


def _debug_arrays(dbg, layout):
    W, H = layout.width, layout.height
    d = {"codes": np.ctypeslib.as_array(dbg.xyz12, (dbg.numXyz12,)).reshape(3, H, W).copy()}
    if dbg.full:
        d["wavelet"] = np.ctypeslib.as_array(dbg.wavelet, (dbg.numWavelet,)).reshape(3, H, W).copy()
        d["coeffs"] = np.ctypeslib.as_array(dbg.coeffs, (dbg.numCoeffs,)).reshape(-1, 1024).copy()
    return d


class Batch:
    """Many frame slots of any geometry encoded together (the batched form). With `slots`
    (a list of (max_width, max_height, colour)) the batch starts empty and bind() assigns
    frames to its slots for each run (the served form)."""

    def __init__(self, device, layouts=None, settings=None, dwt_tile=10, cm_waves=4, mq_group=256, use_graph=False,
                 pool_bytes_per_coef=0.0, slots=None, mq_read=64):
        self.options = BatchOptions(dwt_tile, cm_waves, mq_group, int(use_graph), mq_read, 0, pool_bytes_per_coef)
        self.handle = _vp()
        if slots is not None:
            self._slots = (Slot * len(slots))(*[Slot(w, h, int(c), 0) for w, h, c in slots])
            self.num_slots = len(slots)
            _check(plugin().slots_create(device.handle, self._slots, len(slots), C.byref(self.options),
                                         C.byref(self.handle)), "slots_create")
            self.layouts, self.settings, self.n = [], [], 0
            return
        self.layouts, self.settings = list(layouts), list(settings)
        n = len(self.layouts)
        self._plain = (Layout * n)(*[l.plain() for l in self.layouts])
        self._settings = (EncodeSettings * n)(*self.settings)
        _check(plugin().batch_create(device.handle, self._plain, self._settings, n, C.byref(self.options),
                                     C.byref(self.handle)), "batch_create")
        self.n = self.num_slots = n

    @staticmethod
    def slots_bytes(device, slots, **options):
        arr = (Slot * len(slots))(*[Slot(w, h, int(c), 0) for w, h, c in slots])
        opt = BatchOptions(options.get("dwt_tile", 3), options.get("cm_waves", 4), options.get("mq_group", 256),
                           int(options.get("use_graph", False)), options.get("mq_read", 64), 0,
                           options.get("pool_bytes_per_coef", 0.0))
        return plugin().slots_bytes(device.handle, arr, len(slots), C.byref(opt))

    def bind(self, layouts, settings):
        """Assigns frames to slots 0..n-1 for the following runs."""
        self.layouts, self.settings = list(layouts), list(settings)
        n = len(self.layouts)
        self._plain = (Layout * n)(*[l.plain() for l in self.layouts])
        self._settings = (EncodeSettings * n)(*self.settings)
        _check(plugin().batch_bind(self.handle, self._plain, self._settings, n), "batch_bind")
        self.n = n

    TUNE = ("dwt_64x64", "dwt_64x32", "dwt_32x64", "dwt_64x16", "cm_1", "cm_2", "cm_4", "cm_8", "mq_64", "mq_128",
            "mq_256", "mq_64_r64", "mq_128_r64", "mq_256_r64", "dwt_strip_120x64", "dwt_strip_120x32",
            "dwt_strip_120x128", "dwt_strip0_120x16", "dwt_strip0_120x8", "dwt_strip0b_120x8", "dwt_stage0_480x8")
    DWT_SHAPES = ("64x64", "64x32", "32x64", "64x16", "strip_120x64", "strip_120x32", "strip_120x128",
                  "strip0_120x16", "strip0_120x8", "strip0b_120x8", "stage0_480x8")

    def autotune(self, reps=2):
        """Times the kernel variants on the bound, uploaded frames and keeps the fastest.
        Returns (chosen options dict, stage times dict in ms)."""
        o = BatchOptions()
        t = (C.c_double * len(self.TUNE))()
        _check(plugin().batch_autotune(self.handle, reps, C.byref(o), t, len(self.TUNE)), "batch_autotune")
        chosen = {"dwt_tile": self.DWT_SHAPES[o.dwtTile], "cm_waves": o.cmWaves,
                  "mq_group": o.mqGroup, "mq_read": o.mqRead}
        return chosen, {k: round(v, 4) for k, v in zip(self.TUNE, list(t))}

    def rate(self, i):
        """Device rate control result of frame i of the last run."""
        r = Rate()
        _check(plugin().batch_rate(self.handle, i, C.byref(r)), "batch_rate")
        return {"lambda": list(r.lambda_), "total": r.total, "comp": list(r.comp),
                "tile_parts": list(r.tilePart)[:r.numTileParts], "distortion": r.distortion,
                "iterations": r.iterations}

    def query(self):
        r = plugin().batch_query(self.handle)
        if r < 0:
            _check(-r, "batch_query")
        return r == 1

    def upload(self, i, frame):
        frame = np.ascontiguousarray(frame)
        _check(plugin().batch_upload(self.handle, i, frame.ctypes.data_as(_vp)), "batch_upload")

    def stage(self, i, frame):
        frame = np.ascontiguousarray(frame)
        dst = plugin().batch_input(self.handle, i)
        C.memmove(dst, frame.ctypes.data, frame.nbytes)

    def set_slope(self, i, t):
        plugin().batch_set_slope(self.handle, i, t)

    def launch(self, from_staging=False):
        _check(plugin().batch_launch(self.handle, int(from_staging)), "batch_launch")

    def wait(self, copy=True):
        outs, tm = (Tier1 * self.n)(), FrameTimings()
        _check(plugin().batch_wait(self.handle, outs if copy else None, C.byref(tm)), "batch_wait")
        timings = {k: getattr(tm, k) for k, _ in FrameTimings._fields_}
        return ([Tier1Result(o) for o in outs] if copy else None), timings

    def encode(self, from_staging=False, copy=True):
        self.launch(from_staging)
        return self.wait(copy)

    def debug(self, i, full=True):
        dbg = Debug(1 if full else 0)
        _check(plugin().batch_debug(self.handle, i, C.byref(dbg)), "batch_debug")
        return _debug_arrays(dbg, self.layouts[i])

    STAGES = ("upload", "color", "dwt", "dwt0", "t1cm", "t1mq", "compact", "download", "gpu", "symbols", "bytes",
              "reruns")

    def stage_ms(self):
        v = (C.c_double * 12)()
        _check(plugin().batch_stage_ms(self.handle, v, 12), "batch_stage_ms")
        return dict(zip(self.STAGES, list(v)))

    def close(self):
        if self.handle:
            plugin().batch_destroy(self.handle)
            self.handle = None


def _t1_arrays(t1):
    """(info5, pass_len, pass_dist, data) of a Tier1Result or an oracle frame dict."""
    if isinstance(t1, dict):
        return (np.ascontiguousarray(t1["info"], np.uint32), np.ascontiguousarray(t1["pass_len"], np.uint32),
                np.ascontiguousarray(t1["pass_dist"], np.int32), np.ascontiguousarray(t1["data"], np.uint8))
    return t1.info, t1.pass_len, t1.pass_dist, t1.data


def encode_codestream(layout, t1, fps=24.0, mbps=250.0, tag="hip7.2", build=True):
    """Rate control + Tier-2 of one frame (libgail.so). Returns (codestream bytes or None,
    lambda, estimated mse)."""
    info, pl, pd, data = _t1_arrays(t1)
    lam, mse = C.c_double(0), C.c_double(0)
    cap = int(fps and 250e6 / (8 * fps)) + 4096 if build else 0
    out = np.zeros(max(cap, 1), np.uint8) if build else None
    n = host().gail_encode_codestream(layout.width, layout.height, layout.fourk, layout.base_step,
                                      info.ctypes.data_as(_vp), pl.ctypes.data_as(_vp), pd.ctypes.data_as(_vp),
                                      data.ctypes.data_as(_vp), len(info), fps, mbps,
                                      out.ctypes.data_as(_vp) if build else None, cap, C.byref(lam), C.byref(mse),
                                      tag.encode())
    if n < 0:
        raise GailError(int(n), "rate control / Tier-2 failed")
    return (bytes(out[:n]) if build else None), lam.value, mse.value


def build_codestream(layout, t1, tag="hip", cap=8 << 20):
    """Tier-2 of a Tier-1 frame whose passes are already chosen (device rate control)."""
    info, pl, pd, data = _t1_arrays(t1)
    h = host()
    h.gail_build_codestream.restype = C.c_int64
    h.gail_build_codestream.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, _vp, _vp, _vp, _vp, C.c_uint64, _vp,
                                        C.c_uint64, C.c_char_p]
    out = np.zeros(cap, np.uint8)
    n = h.gail_build_codestream(layout.width, layout.height, layout.fourk, layout.base_step, info.ctypes.data_as(_vp),
                                pl.ctypes.data_as(_vp), pd.ctypes.data_as(_vp), data.ctypes.data_as(_vp), len(info),
                                out.ctypes.data_as(_vp), cap, tag.encode())
    if n < 0:
        raise GailError(int(n), "Tier-2 failed")
    return bytes(out[:n])


def rate_allocate(layout, t1, max_total, max_comp):
    """Rate control of a Tier-1 frame against explicit byte budgets (libgail.so): a dict of
    the passes kept per block, the per-component lambdas, the codestream's total and
    component sizes, the remaining distortion and the iterations."""
    info, pl, pd, data = _t1_arrays(t1)
    h = host()
    h.gail_rate_allocate.restype = C.c_int
    h.gail_rate_allocate.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, _vp, _vp, _vp, _vp, C.c_uint64,
                                     C.c_uint64, C.c_uint64, _vp, _vp, _vp, _vp]
    passes = np.zeros(len(info), np.uint8)
    lam = np.zeros(3, np.float64)
    sizes = np.zeros(4, np.uint64)
    dist = C.c_double(0)
    it = h.gail_rate_allocate(layout.width, layout.height, layout.fourk, layout.base_step, info.ctypes.data_as(_vp),
                              pl.ctypes.data_as(_vp), pd.ctypes.data_as(_vp), data.ctypes.data_as(_vp), len(info),
                              int(max_total), int(max_comp), passes.ctypes.data_as(_vp), lam.ctypes.data_as(_vp),
                              sizes.ctypes.data_as(_vp), C.byref(dist))
    if it < 0:
        raise GailError(it, "rate control failed")
    return dict(passes=passes, lambdas=lam, total=int(sizes[0]), comp=[int(v) for v in sizes[1:]],
                distortion=dist.value, iterations=it)


def with_passes(t1, passes):
    """A copy of a Tier-1 frame (dict of arrays) whose blocks keep only `passes` passes each,
    for gail_build_codestream."""
    info, pl, pd, data = _t1_arrays(t1)
    info = info.copy()
    info[:, 1] = np.minimum(info[:, 1], np.asarray(passes, np.uint32))
    return dict(info=info, pass_len=pl, pass_dist=pd, data=data)


def dci_limits(fps, mbps=230.0, comp_mbps=200.0):
    """libgail.so's dciLimits: (frame budget, component budget) in bytes."""
    h = host()
    h.gail_dci_limits.restype = None
    h.gail_dci_limits.argtypes = [C.c_double, C.c_double, C.c_double, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]
    t, c = C.c_uint64(0), C.c_uint64(0)
    h.gail_dci_limits(fps, mbps, comp_mbps, C.byref(t), C.byref(c))
    return t.value, c.value


def empty_codestream_bytes(layout):
    h = host()
    h.gail_empty_codestream_bytes.restype = C.c_uint64
    h.gail_empty_codestream_bytes.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double]
    return h.gail_empty_codestream_bytes(layout.width, layout.height, layout.fourk, layout.base_step)


def dci_budget(fps, mbps=250.0, comp_mbps=200.0):
    """(frame budget, component budget) in bytes: floor(Mbit/s * 1e6 / (8 fps)) - 16
    (rate_control.cpp:11-26)."""
    import math
    return math.floor(mbps * 1e6 / (8 * fps)) - 16, math.floor(comp_mbps * 1e6 / (8 * fps)) - 16


def with_rate_control(settings, fps=24.0, mbps=250.0):
    """A copy of encode settings with rate control on the device at fps / mbps."""
    s = EncodeSettings()
    C.pointer(s)[0] = settings
    s.rcMaxTotal, s.rcMaxComp = dci_budget(fps, mbps)
    return s


def rate_lambda(layout, t1, fps=24.0, mbps=250.0):
    return encode_codestream(layout, t1, fps, mbps, build=False)[1]


def check_codestream(cs, fps, fourk):
    """'' when the codestream passes the DCI structural check, else the reason."""
    buf = np.frombuffer(cs, np.uint8)
    msg = C.create_string_buffer(256)
    host().gail_check_codestream(buf.ctypes.data_as(_vp), len(buf), fps, int(fourk), msg, 256)
    return msg.value.decode()


# ---------------------------------------------------------------- library API (gail_amd.h)
class GailInput(C.Structure):
    _fields_ = [("format", C.c_int), ("width", C.c_int), ("height", C.c_int), ("matrix", C.c_int),
                ("full_range", C.c_int), ("chroma_location", C.c_int), ("reserved", C.c_uint64 * 4)]


class GailColor(C.Structure):
    _fields_ = [("mode", C.c_int), ("transfer", C.c_int), ("primaries", C.c_int), ("tone_map", C.c_int),
                ("tone_norm", C.c_int), ("sdr_white", C.c_double), ("hdr_peak", C.c_double),
                ("hdr_ref_white", C.c_double), ("cinema_ref_white", C.c_double), ("cinema_peak", C.c_double),
                ("knee", C.c_double), ("gamut", C.c_int), ("gamut_threshold", C.c_double), ("lut3d", _vp),
                ("lut3d_size", C.c_int), ("reserved", C.c_uint64 * 8)]


class GailConfig(C.Structure):
    _fields_ = [("struct_size", C.c_size_t), ("profile", C.c_int), ("container", C.c_int), ("fps", C.c_double),
                ("max_mbps", C.c_double), ("stereo", C.c_int), ("input", GailInput), ("color", GailColor),
                ("backend", C.c_int), ("devices", _vp), ("num_devices", C.c_int), ("frames_in_flight", C.c_int),
                ("check", C.c_int), ("on_packet", _vp), ("user", _vp), ("log", _vp), ("log_user", _vp),
                ("log_level", C.c_int), ("device_ids", _vp), ("num_device_ids", C.c_int),
                ("max_device_memory", C.c_uint64), ("cpu_threads", C.c_int), ("low_priority", C.c_int)]


class GailFrame(C.Structure):
    _fields_ = [("data", _vp * 3), ("stride", C.c_ssize_t * 3), ("opaque", _vp), ("reserved", C.c_uint64 * 4)]


class GailPacket(C.Structure):
    _fields_ = [("data", _vp), ("size", C.c_size_t), ("index", C.c_int64), ("opaque", _vp), ("device", C.c_int),
                ("compliant", C.c_int), ("psnr_estimate", C.c_double), ("reserved", C.c_uint64 * 4)]


class ServerResult(C.Structure):
    _fields_ = [("id", C.c_uint64), ("stream", C.c_uint32), ("status", C.c_int), ("packet", C.POINTER(GailPacket)),
                ("queued_ms", C.c_double), ("latency_ms", C.c_double), ("batch_frames", C.c_int),
                ("recoded", C.c_int), ("lambda_", C.c_double)]


SERVER_CB = C.CFUNCTYPE(None, _vp, C.POINTER(ServerResult))


class ServerConfig(C.Structure):
    _fields_ = [("struct_size", C.c_size_t), ("device", C.c_int), ("slots_per_job", C.c_int), ("max_jobs", C.c_int),
                ("memory_budget", C.c_uint64), ("cpu_threads", C.c_int), ("check", C.c_int),
                ("batch_window_us", C.c_int), ("host_rate_control", C.c_int), ("on_result", SERVER_CB),
                ("user", _vp)]


class ServerStats(C.Structure):
    _fields_ = [("frames", C.c_uint64), ("failed", C.c_uint64), ("recoded", C.c_uint64), ("seconds", C.c_double),
                ("fps", C.c_double), ("p50_ms", C.c_double), ("p99_ms", C.c_double), ("max_ms", C.c_double),
                ("queued_p50_ms", C.c_double), ("queued_p99_ms", C.c_double), ("mean_batch", C.c_double),
                ("jobs", C.c_int), ("slots_per_job", C.c_int), ("max_frames_in_flight", C.c_int),
                ("job_bytes", C.c_uint64), ("slot_bytes", C.c_uint64), ("jobs_2k", C.c_int), ("jobs_4k", C.c_int),
                ("jobs_made", C.c_int), ("jobs_retired", C.c_int), ("max_jobs", C.c_int),
                ("slots_per_job_2k", C.c_int), ("peak_job_bytes", C.c_uint64), ("slot_bytes_2k", C.c_uint64)]


PIX = {"rgb48": 0, "xyz12": 1, "xyz12_msb": 2, "yuv420p": 3, "yuv420p10": 4, "yuv420p12": 5, "yuv422p": 6,
       "yuv422p10": 7, "yuv422p12": 8, "yuv444p": 9, "yuv444p10": 10, "yuv444p12": 11, "gbrp": 12, "gbrp10": 13,
       "gbrp12": 14, "gbrp16": 15, "xyz12_planar": 16}
CONTAINER_IDS = {"2k_flat": (0, 0), "2k_scope": (0, 1), "2k_full": (0, 2), "4k_flat": (1, 0), "4k_scope": (1, 1),
                 "4k_full": (1, 2)}


def _lib_api():
    h = host()
    if not hasattr(h, "_server_bound"):
        h.gail_config_init.restype = None
        h.gail_config_init.argtypes = [C.POINTER(GailConfig), C.c_size_t]
        h.gail_color_preset.restype = None
        h.gail_color_preset.argtypes = [C.c_int, C.POINTER(GailColor)]
        h.gail_last_error.restype = C.c_char_p
        h.gail_server_config_init.restype = None
        h.gail_server_config_init.argtypes = [C.POINTER(ServerConfig), C.c_size_t]
        h.gail_server_create.restype = C.c_int
        h.gail_server_create.argtypes = [C.POINTER(ServerConfig), C.POINTER(_vp)]
        h.gail_server_submit.restype = C.c_int
        h.gail_server_submit.argtypes = [_vp, C.c_uint64, C.c_uint32, C.POINTER(GailConfig), C.POINTER(GailFrame),
                                         C.c_double]
        h.gail_server_drain.restype = C.c_int
        h.gail_server_drain.argtypes = [_vp]
        h.gail_server_get_stats.restype = C.c_int
        h.gail_server_get_stats.argtypes = [_vp, C.POINTER(ServerStats)]
        h.gail_device_alloc.restype = _vp
        h.gail_device_alloc.argtypes = [C.c_size_t]
        h.gail_device_free.restype = None
        h.gail_device_free.argtypes = [_vp]
        h.gail_copy.restype = C.c_int
        h.gail_copy.argtypes = [_vp, _vp, C.c_size_t]
        h.gail_host_alloc.restype = _vp
        h.gail_host_alloc.argtypes = [C.c_size_t]
        h.gail_host_free.restype = None
        h.gail_host_free.argtypes = [_vp]
        h.gail_server_destroy.restype = None
        h.gail_server_destroy.argtypes = [_vp]
        h._server_bound = True
    return h


def stream_config(container="4k_flat", fmt="xyz12_planar", fps=24.0, mbps=250.0, picture=None, matrix=0,
                  full_range=0, chroma=0, preset=0):
    """A gail_config describing one stream (profile, container, rate, input, colour)."""
    h = _lib_api()
    c = GailConfig()
    h.gail_config_init(C.byref(c), C.sizeof(c))
    c.profile, c.container = CONTAINER_IDS[container]
    c.fps, c.max_mbps = fps, mbps
    c.input.format = PIX[fmt]
    if picture:
        c.input.width, c.input.height = picture
    c.input.matrix, c.input.full_range, c.input.chroma_location = matrix, full_range, chroma
    h.gail_color_preset(preset, C.byref(c.color))
    return c


class ServerError(RuntimeError):
    pass


def _lib_encoder():
    h = _lib_api()
    if not hasattr(h, "_encoder_bound"):
        h.gail_encoder_create.restype = C.c_int
        h.gail_encoder_create.argtypes = [C.POINTER(GailConfig), C.POINTER(_vp)]
        h.gail_encoder_send.restype = C.c_int
        h.gail_encoder_send.argtypes = [_vp, C.POINTER(GailFrame)]
        h.gail_encoder_receive.restype = C.c_int
        h.gail_encoder_receive.argtypes = [_vp, C.POINTER(GailPacket), C.c_int]
        h.gail_encoder_destroy.restype = None
        h.gail_encoder_destroy.argtypes = [_vp]
        h.gail_packet_free.restype = None
        h.gail_packet_free.argtypes = [C.POINTER(GailPacket)]
        h._encoder_bound = True
    return h


class ApiEncoder:
    """The drop-in encoder (gail_encoder_*, include/gail_amd.h): frames sent in order come
    back as codestreams in the same order (pull model, gail.h:329-362)."""

    def __init__(self, cfg):
        h = _lib_encoder()
        self.handle = _vp()
        if h.gail_encoder_create(C.byref(cfg), C.byref(self.handle)) != 0:
            self.handle = _vp()
            raise ServerError(f"gail_encoder_create: {h.gail_last_error().decode()}")

    def send(self, planes):
        """planes: three numpy arrays (planar formats) or one array; copied before the call
        returns, as gail_encoder_send does."""
        f = GailFrame()
        arrs = planes if isinstance(planes, (list, tuple)) else [planes]
        for i, a in enumerate(arrs):
            f.data[i] = a.ctypes.data
            f.stride[i] = a.strides[0] if a.ndim > 1 else 0
        if _lib_encoder().gail_encoder_send(self.handle, C.byref(f)) != 0:
            raise ServerError(f"gail_encoder_send: {_lib_encoder().gail_last_error().decode()}")

    def end(self):
        _lib_encoder().gail_encoder_send(self.handle, None)

    def receive(self):
        """(codestream bytes, frame index, DCI-compliant) of the next frame, or None at the end."""
        h = _lib_encoder()
        pkt = GailPacket()
        st = h.gail_encoder_receive(self.handle, C.byref(pkt), 1)
        if st == 2:  # GAIL_EOF
            return None
        if st != 0:
            raise ServerError(f"gail_encoder_receive: {h.gail_last_error().decode()}")
        data = C.string_at(pkt.data, pkt.size)
        out = (data, pkt.index, pkt.compliant)
        h.gail_packet_free(C.byref(pkt))
        return out

    def close(self):
        if self.handle:
            _lib_encoder().gail_encoder_destroy(self.handle)
            self.handle = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def pinned_empty(nbytes):
    """A uint8 numpy array over pinned host memory (gail_host_alloc), freed with the array."""
    import weakref
    h = _lib_api()
    p = h.gail_host_alloc(nbytes)
    if not p:
        raise ServerError(f"pinned allocation of {nbytes} bytes failed: {h.gail_last_error().decode()}")
    buf = (C.c_uint8 * nbytes).from_address(p)
    weakref.finalize(buf, h.gail_host_free, p)
    return np.frombuffer(buf, np.uint8)


class DeviceFrame:
    """A frame's planes uploaded, back to back, into device memory (gail_device_alloc), for
    submitting frames that are already resident in HBM. The memory is freed by close() or
    when the object is collected."""

    def __init__(self, planes):
        h = _lib_api()
        arrs = [np.ascontiguousarray(p) for p in planes]
        total = sum(a.nbytes for a in arrs)
        self.ptr = h.gail_device_alloc(total)
        if not self.ptr:
            raise ServerError(f"device allocation failed: {h.gail_last_error().decode()}")
        self._free = weakref.finalize(self, h.gail_device_free, self.ptr)
        self.ptrs, self.strides, off = [], [], 0
        for a in arrs:
            if h.gail_copy(self.ptr + off, a.ctypes.data, a.nbytes) != 0:
                raise ServerError(f"upload failed: {h.gail_last_error().decode()}")
            self.ptrs.append(self.ptr + off)
            self.strides.append(a.strides[0] if a.ndim > 1 else 0)
            off += a.nbytes
        self.nbytes = total

    def close(self):
        self._free()
        self.ptr = None


class _ServerSink:
    """The part of a Server its result callback uses. It is kept apart from the Server object:
    the C++ pool threads call the callback thunk until gail_server_destroy returns, so the
    thunk, this sink and the server handle are owned by a finaliser that destroys the server
    first, whether close() runs or the Server is collected without it."""

    def __init__(self, keep, on_done):
        self.keep, self.on_done = keep, on_done
        self.results, self.frames = {}, {}
        self.lock = threading.Lock()

    def on_result(self, _user, rp):
        r = rp.contents
        d = {"status": r.status, "latency_ms": r.latency_ms, "queued_ms": r.queued_ms, "batch": r.batch_frames,
             "recoded": r.recoded, "lambda": r.lambda_, "stream": r.stream}
        if r.packet:
            p = r.packet.contents
            d["size"], d["compliant"], d["psnr"] = p.size, p.compliant, p.psnr_estimate
            if self.keep:
                d["codestream"] = C.string_at(p.data, p.size)
        with self.lock:
            self.results[r.id] = d
            self.frames.pop(r.id, None)
        if self.on_done:
            self.on_done(r.id, d)


def _destroy_server(handle, thunk, sink):
    """Waits for every frame and every pool thread of the server (gail_server_destroy), and
    only then lets the callback thunk, the sink and the frames it keeps alive go."""
    if handle.value:
        _lib_api().gail_server_destroy(handle)
        handle.value = None
    del thunk, sink


class Server:
    """The served form (gail_server_*): frames of many streams admitted into device jobs as
    slots free up. Results land in self.results[id] as dicts; with keep=True each also holds
    the codestream bytes. close() waits for every frame; a Server collected without it is
    closed the same way."""

    def __init__(self, device=0, slots_per_job=8, max_jobs=4, memory_budget=0, cpu_threads=0, check=False,
                 keep=False, batch_window_us=0, on_done=None, host_rate_control=False):
        h = _lib_api()
        self._sink = _ServerSink(keep, on_done)  # on_done: called with (id, result dict)
        self.results = self._sink.results
        cfg = ServerConfig()
        h.gail_server_config_init(C.byref(cfg), C.sizeof(cfg))
        cfg.device, cfg.slots_per_job, cfg.max_jobs = device, slots_per_job, max_jobs
        cfg.memory_budget, cfg.cpu_threads, cfg.check = memory_budget, cpu_threads, int(check)
        cfg.batch_window_us = int(batch_window_us)
        cfg.host_rate_control = int(host_rate_control)
        thunk = SERVER_CB(self._sink.on_result)
        cfg.on_result = thunk
        self.handle = _vp()
        if h.gail_server_create(C.byref(cfg), C.byref(self.handle)) != 0:
            raise ServerError(h.gail_last_error().decode())
        self._close = weakref.finalize(self, _destroy_server, self.handle, thunk, self._sink)

    def submit(self, fid, stream, cfg, planes, slope=-1.0):
        """planes: list of three numpy arrays (planar formats) or one array; kept alive until
        the frame's result arrives."""
        f = GailFrame()
        if isinstance(planes, DeviceFrame):
            arrs = planes
            for i, (ptr, st) in enumerate(zip(planes.ptrs, planes.strides)):
                f.data[i], f.stride[i] = ptr, st
        else:
            arrs = planes if isinstance(planes, (list, tuple)) else [planes]
            for i, a in enumerate(arrs):
                f.data[i] = a.ctypes.data
                f.stride[i] = a.strides[0] if a.ndim > 1 else 0
        if not self.handle:
            raise ServerError("submit: the server is closed")
        with self._sink.lock:  # the planes live until the frame's result arrives
            self._sink.frames[fid] = (arrs, cfg)
        st = _lib_api().gail_server_submit(self.handle, fid, stream, C.byref(cfg), C.byref(f), slope)
        if st != 0:
            with self._sink.lock:
                self._sink.frames.pop(fid, None)
            raise ServerError(f"submit: {_lib_api().gail_last_error().decode()}")

    def drain(self):
        _lib_api().gail_server_drain(self.handle)

    def stats(self):
        s = ServerStats()
        _lib_api().gail_server_get_stats(self.handle, C.byref(s))
        return {k: getattr(s, k) for k, _ in ServerStats._fields_}

    def close(self):
        self._close()
