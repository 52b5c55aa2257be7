"""The oracle against the golden vectors written from the reference's own compiled CPU
sources (tests/golden/make_golden.py): Tier-1 blocks (plain, early-terminated and
pass-limited), layouts of every DCI container, DWT outputs, colour conversions and the
hashes of a whole 2K frame."""
import hashlib
import json
import os

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def test_t1_blocks(oracle):
    z = np.load(os.path.join(GOLDEN, "t1_blocks.npz"))
    dims, params = z["dims"], z["params"]
    data_off = pass_off = 0
    for i in range(len(dims)):
        w, h, orient, mp, nbps, npass, ln = (int(v) for v in dims[i])
        thr, scale = float(params[i][0]), float(params[i][1])
        r = oracle.t1_encode(z["coeffs"][i], w, h, orient, max_passes=mp, thr=thr, scale=scale)
        assert r["nbps"] == nbps and r["num_passes"] == npass, f"block {i}"
        assert r["data"] == bytes(z["data"][data_off:data_off + ln]), f"codeword of block {i}"
        assert r["pass_len"] == [int(v) for v in z["pass_len"][pass_off:pass_off + npass]], f"block {i}"
        assert r["pass_dist"] == [int(v) for v in z["pass_dist"][pass_off:pass_off + npass]], f"block {i}"
        data_off += ln
        pass_off += npass


def test_layouts(oracle):
    with open(os.path.join(GOLDEN, "layouts.json")) as f:
        gold = json.load(f)
    for key, g in gold.items():
        W, H = (int(v) for v in key.split("x"))
        bands, cbl = oracle.layout(W, H, g["fourk"])
        rows = [[b.comp, b.res, b.level, b.orient, b.px0, b.py0, b.w, b.h, b.ncbx, b.ncby, b.cblk_base, b.eps, b.mu,
                 b.mb] for b in bands]
        assert rows == g["bands"], key
        assert [float(np.float32(b.delta)) for b in bands] == g["delta"], key
        assert [b.weight for b in bands] == g["weight"], key
        assert len(cbl) == g["num_cblks"]
        assert sha(np.array([[c.band, c.x0, c.y0, c.w, c.h] for c in cbl], np.int32)) == g["cblks_sha"]


def test_dwt(oracle):
    z = np.load(os.path.join(GOLDEN, "dwt.npz"))
    for k in z.files:
        if not k.startswith("in_"):
            continue
        W, H = (int(v) for v in k.split("_")[1].split("x"))
        NL = int(k.split("_")[2])
        out = oracle.fdwt(z[k], NL)
        assert np.array_equal(out.view(np.uint32), z["out" + k[2:]].view(np.uint32)), k


def test_color(oracle):
    z = np.load(os.path.join(GOLDEN, "color.npz"))
    with open(os.path.join(GOLDEN, "color.json")) as f:
        meta = json.load(f)
    for name, m in meta.items():
        if name.startswith("tables_"):
            tr, pr = (int(v) for v in name.split("_")[1:])
            lut, cp = oracle.color_tables(tr, pr)
            assert sha(lut) == m["lut_sha"], name
            assert [float(v) for v in cp] == m["params"], name
            continue
        rgb48 = oracle.planar_to_rgb48(z[f"{name}_src"], m["bits"], m["aw"], m["ah"], m["sx"], m["sy"], m["rgb"],
                                       m["full_range"], m["chroma"], m["kr"], m["kb"], m["W"], m["H"], m["ox"], m["oy"])
        assert np.array_equal(rgb48, z[f"{name}_rgb48"]), name
        lut, cp = oracle.color_tables(m["transfer"], m["primaries"])
        planes, xyz = oracle.convert_frame(lut, cp, rgb48)
        assert np.array_equal(xyz, z[f"{name}_xyz"]), name
        assert np.array_equal(planes.view(np.uint32), z[f"{name}_planes"].view(np.uint32)), name


@pytest.mark.parametrize("key", ["2k_flat_seed1_thr0.0", "2k_flat_seed1_thr50.0"])
def test_frame_2k(oracle, key):
    """The CPU workload: one 2K flat 1998x1080 frame of the seeded generator, NL 5, whole Tier-1."""
    import libgail_amd as g
    with open(os.path.join(GOLDEN, "frames.json")) as f:
        gold = json.load(f)[key]
    x = g.synth_frame(1, 1998, 1080)
    assert sha(x) == gold["input"], "the synthetic generator changed"
    r = oracle.encode_frame_xyz(x, 0, thr=float(key.split("thr")[1]), threads=8)
    for k in ("coeffs", "info", "pass_len", "pass_dist", "data"):
        assert sha(r[k]) == gold[k], k
