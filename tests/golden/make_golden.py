"""Writes the golden vectors of tests/golden/ from the reference's own CPU sources, compiled
by oracle/Makefile into oracle/_ref/libgailref.so (run in a container that has the
reference checked out; the fixtures travel, the reference does not).

    python tests/golden/make_golden.py
"""
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
from oracle_bind import Reference, random_block  # noqa: E402


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def t1_fixture(ref):
    rng = np.random.default_rng(42)
    coeffs, dims, outs = [], [], []
    # (max_passes, thr, scale): plain, early stop at three thresholds, pass-limited
    variants = [(64, 0.0, 1.0)] * 3 + [(64, 0.5, 0.7), (64, 5.0, 0.7), (64, 50.0, 0.7), (7, 0.0, 1.0)]
    for trial in range(210):
        c, w, h, orient = random_block(rng, trial)
        mp, thr, scale = variants[trial % len(variants)]
        r = ref.t1_encode(c, w, h, orient, max_passes=mp, thr=thr, scale=scale)
        coeffs.append(c)
        dims.append((w, h, orient, mp, r["nbps"], r["num_passes"], len(r["data"])))
        outs.append(r)
    np.savez_compressed(
        os.path.join(HERE, "t1_blocks.npz"),
        coeffs=np.array(coeffs, np.uint32),
        dims=np.array(dims, np.int32),
        params=np.array([variants[t % len(variants)][1:] for t in range(len(dims))], np.float32),
        data=np.frombuffer(b"".join(o["data"] for o in outs), np.uint8),
        pass_len=np.array([v for o in outs for v in o["pass_len"]], np.uint32),
        pass_dist=np.array([v for o in outs for v in o["pass_dist"]], np.int32),
    )


def layout_fixture(ref):
    out = {}
    for (W, H, k) in [(1998, 1080, 0), (2048, 858, 0), (2048, 1080, 0), (3996, 2160, 1), (4096, 1716, 1),
                      (4096, 2160, 1), (301, 157, 0)]:
        bands, cbl = ref.layout(W, H, k)
        rows = [[b.comp, b.res, b.level, b.orient, b.px0, b.py0, b.w, b.h, b.ncbx, b.ncby, b.cblk_base, b.eps, b.mu,
                 b.mb] for b in bands]
        out[f"{W}x{H}"] = dict(fourk=k, bands=rows, delta=[float(b.delta) for b in bands],
                               weight=[float(b.weight) for b in bands], num_cblks=len(cbl),
                               cblks_sha=sha(np.array([[c.band, c.x0, c.y0, c.w, c.h] for c in cbl], np.int32)))
    with open(os.path.join(HERE, "layouts.json"), "w") as f:
        json.dump(out, f, indent=0)


def dwt_fixture(ref):
    rng = np.random.default_rng(7)
    arrays = {}
    for (W, H, NL) in [(67, 45, 3), (131, 97, 4), (64, 2, 2), (3, 70, 2)]:
        x = rng.uniform(-2048, 2047, size=(H, W)).astype(np.float32)
        arrays[f"in_{W}x{H}_{NL}"] = x
        arrays[f"out_{W}x{H}_{NL}"] = ref.fdwt(x, NL)
    np.savez_compressed(os.path.join(HERE, "dwt.npz"), **arrays)


def color_fixture(ref):
    rng = np.random.default_rng(11)
    arrays, meta = {}, {}
    aw, ah, W, H, ox, oy = 61, 33, 67, 41, 3, 4
    for name, (bits, sx, sy, rgb, fr, chroma, kr, kb, tr, pr) in {
            "gbrp12_bt709": (12, 0, 0, 1, 1, 0, 0.2126, 0.0722, 0, 0),
            "yuv420p10_bt709": (10, 1, 1, 0, 0, 0, 0.2126, 0.0722, 0, 0),
            "yuv422p8_center": (8, 1, 0, 0, 1, 1, 0.2627, 0.0593, 1, 1),
    }.items():
        cw, ch = (aw + sx) >> sx, (ah + sy) >> sy
        n = aw * ah + 2 * cw * ch
        src = rng.integers(0, 1 << bits, size=n).astype(np.uint16 if bits > 8 else np.uint8)
        rgb48 = ref.planar_to_rgb48(src, bits, aw, ah, sx, sy, rgb, fr, chroma, kr, kb, W, H, ox, oy)
        planes, xyz = ref.convert_frame(tr, pr, rgb48)
        arrays[f"{name}_src"] = src
        arrays[f"{name}_rgb48"] = rgb48
        arrays[f"{name}_xyz"] = xyz
        arrays[f"{name}_planes"] = planes
        meta[name] = dict(bits=bits, sx=sx, sy=sy, rgb=rgb, full_range=fr, chroma=chroma, kr=kr, kb=kb,
                          transfer=tr, primaries=pr, aw=aw, ah=ah, W=W, H=H, ox=ox, oy=oy)
    for tr, pr in [(0, 0), (1, 1), (4, 2), (5, 4)]:
        lut, cp = ref.color_tables(tr, pr)
        meta[f"tables_{tr}_{pr}"] = dict(lut_sha=sha(lut), params=[float(v) for v in cp])
    np.savez_compressed(os.path.join(HERE, "color.npz"), **arrays)
    with open(os.path.join(HERE, "color.json"), "w") as f:
        json.dump(meta, f, indent=0)


def frame_fixture(ref):
    import libgail_amd as g
    out = {}
    for name, seed, thr in [("2k_flat", 1, 0.0), ("2k_flat", 1, 50.0)]:
        W, H, k = g.CONTAINERS[name]
        x = g.synth_frame(seed, W, H)
        r = ref.encode_frame_xyz(x, k, thr=thr, threads=8)
        out[f"{name}_seed{seed}_thr{thr}"] = dict(
            input=sha(x), coeffs=sha(r["coeffs"]), info=sha(r["info"]), pass_len=sha(r["pass_len"]),
            pass_dist=sha(r["pass_dist"]), data=sha(r["data"]), data_bytes=int(len(r["data"])),
            passes=int(len(r["pass_len"])))
    with open(os.path.join(HERE, "frames.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    ref = Reference()
    t1_fixture(ref)
    layout_fixture(ref)
    dwt_fixture(ref)
    color_fixture(ref)
    frame_fixture(ref)
    for fn in sorted(os.listdir(HERE)):
        print(fn, os.path.getsize(os.path.join(HERE, fn)))
