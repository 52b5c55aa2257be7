"""The oracle against the reference compiled from its own sources (oracle/_ref), on fresh
seeded inputs. Skipped where oracle/_ref was not built."""
import numpy as np
import pytest

from oracle_bind import random_block, reference


@pytest.fixture(scope="module")
def ref():
    r = reference()
    if r is None:
        pytest.skip("oracle/_ref/libgailref.so not built")
    return r


def test_t1_fresh_blocks(oracle, ref):
    rng = np.random.default_rng(2024)
    for trial in range(300):
        c, w, h, orient = random_block(rng, trial)
        thr = (0.0, 1.0, 20.0)[trial % 3]
        assert oracle.t1_encode(c, w, h, orient, thr=thr, scale=0.9, trace=True) == \
            ref.t1_encode(c, w, h, orient, thr=thr, scale=0.9, trace=True), trial


def test_dwt_fresh(oracle, ref):
    rng = np.random.default_rng(3)
    for (W, H, NL) in [(157, 301, 5), (8, 8, 2), (5, 300, 3)]:
        x = rng.uniform(-4000, 4000, size=(H, W)).astype(np.float32)
        assert np.array_equal(oracle.fdwt(x, NL).view(np.uint32), ref.fdwt(x, NL).view(np.uint32))


def test_frame_scope_2k(oracle, ref):
    import libgail_amd as g
    x = g.synth_frame(9, 2048, 858)
    a = oracle.encode_frame_xyz(x, 0, thr=10.0, threads=8)
    b = ref.encode_frame_xyz(x, 0, thr=10.0, threads=8)
    for k in ("coeffs", "info", "pass_len", "pass_dist", "data"):
        assert np.array_equal(a[k], b[k]), k
