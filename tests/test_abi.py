"""The C-ABI libraries load and export what include/*.h declares; the plugin table passes
the loader's checks (ABI, size, layout hash of the plain structures). No GPU needed."""
import ctypes as C
import os
import re

import pytest

import libgail_amd as g

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared(header):
    """Functions a header declares as exported (GAIL_API or default visibility)."""
    text = open(os.path.join(ROOT, "include", header)).read()
    decls = re.findall(r"(?:GAIL_API|visibility\(\"default\"\)\)\))[^;{]*?\b(gail_[a-z0-9_]+)\s*\(", text, re.S)
    return set(decls)


def test_plugin_table():
    p = g.plugin()
    assert p.abi == g.ABI and p.size == C.sizeof(g.Plugin)
    assert p.name == b"gail_hip" and p.archs == b"gfx950"
    # every entry is set
    for name, _ in g.Plugin._fields_[6:]:
        assert getattr(p, name), name
    # layout hash over the plain structures, computed the way gail_hip_plugin_layout() does
    sizes = [C.sizeof(t) for t in (g.DeviceDesc, g.Subband, g.Layout, g.FrameFormat, g.Color, g.EncodeSettings,
                                   g.Memory, g.CblkInfo, g.Tier1, g.FrameTimings, g.Debug, g.BatchOptions, g.Slot,
                                   g.Rate, g.DecodeSettings, g.CblkCode, g.DecodeTimings, g.Plugin)]
    h = 1469598103934665603
    for s in sizes:
        h = ((h ^ s) * 1099511628211) & 0xFFFFFFFFFFFFFFFF
    assert p.layout == h


def test_plugin_rejects_other_abi():
    lib = C.CDLL(g.PLUGIN_PATH)
    lib.gail_hip_plugin_get.restype = C.c_void_p
    assert lib.gail_hip_plugin_get(g.ABI + 1) is None


def test_plugin_exports_only_entry():
    lib = C.CDLL(g.PLUGIN_PATH)
    for sym in _declared("gail_hip_plugin.h") - {"gail_hip_plugin_layout"}:
        assert hasattr(lib, sym), sym


def test_host_exports():
    lib = C.CDLL(g.HOST_PATH)
    missing = [s for s in _declared("gail_amd.h") if not hasattr(lib, s)]
    assert not missing, missing


def test_decoder_entries_check_arguments():
    """The decoder entries refuse missing arguments as misuse (GAIL_HIP_MISUSE), and a
    reduction beyond the layout's levels as an error, before touching a device."""
    p = g.plugin()
    out = C.c_void_p()
    assert p.decoder_create(None, None, None, C.byref(out)) == 4
    assert not out.value
    L = g.FrameLayout(1998, 1080, 0)
    pl = L.plain()
    assert p.decoder_bytes(None, C.byref(pl), C.byref(g.DecodeSettings(0, 1))) > 3 * 1998 * 1080 * 12
    assert p.decoder_bytes(None, C.byref(pl), C.byref(g.DecodeSettings(6, 1))) == 0
    assert p.decoder_decode(None, None, None, 0, None, 0, None, 0, None) == 4
    assert C.sizeof(g.CblkCode) == 12 and g.CBLK_CODE.itemsize == 12
