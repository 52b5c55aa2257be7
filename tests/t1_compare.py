"""Comparison of a Tier-1 frame (info rows, pass tables, codewords) against the oracle's,
block by block, with a readable report of the first differences."""
import numpy as np


def oracle_blocks(ref):
    info = ref["info"]
    out = []
    for i in range(len(info)):
        nbps, np_, off, ln, pb = (int(v) for v in info[i])
        out.append(dict(nbps=nbps, num_passes=np_, data=bytes(ref["data"][off:off + ln]),
                        pass_len=[int(v) for v in ref["pass_len"][pb:pb + np_]],
                        pass_dist=[int(v) for v in ref["pass_dist"][pb:pb + np_]]))
    return out


def compare_tier1(gpu, ref, max_report=5):
    """Returns (number of blocks that differ, report lines)."""
    want = oracle_blocks(ref)
    bad, lines = 0, []
    if len(gpu.info) != len(want):
        return len(want), [f"block count {len(gpu.info)} != {len(want)}"]
    for i, w in enumerate(want):
        g = gpu.block(i)
        if g != w:
            bad += 1
            if len(lines) < max_report:
                diff = [k for k in w if g[k] != w[k]]
                lines.append(f"block {i}: differs in {diff}; gpu nbps/np/len {g['nbps']}/{g['num_passes']}/"
                             f"{len(g['data'])} oracle {w['nbps']}/{w['num_passes']}/{len(w['data'])}")
    return bad, lines


def planar_from_synth(x, bits, sx, sy, rgb):
    """A planar picture derived from a synthetic frame: G'B'R' (rgb) or Y'CbCr with chroma
    subsampled by 2^sx x 2^sy, at the given bit depth, planes back to back (uint16)."""
    shift = 12 - bits
    if rgb:
        planes = [x[1], x[2], x[0]]
        return np.concatenate([(p >> shift if shift >= 0 else p << -shift).astype(np.uint16).ravel() for p in planes])
    Y = (x[1] >> shift).astype(np.uint16)

    def sub(p):
        p = p[::1 << sy, ::1 << sx]
        return (p >> shift).astype(np.uint16)
    return np.concatenate([Y.ravel(), sub(x[0]).ravel(), sub(x[2]).ravel()])


def _gather(base, counts):
    """Indices base[i] + 0..counts[i]-1 for every i, concatenated."""
    counts = np.asarray(counts, np.int64)
    total = int(counts.sum())
    if total == 0:
        return np.zeros(0, np.int64)
    starts = np.repeat(np.asarray(base, np.int64) - np.concatenate([[0], np.cumsum(counts)[:-1]]), counts)
    return starts + np.arange(total, dtype=np.int64)


def tier1_mismatch(gpu, ref):
    """Vectorised form of compare_tier1 for large frames: the number of code-blocks whose
    bit-planes, pass count, pass lengths, pass distortions or codeword bytes differ."""
    gi, ri = np.asarray(gpu.info, np.int64), np.asarray(ref["info"], np.int64)
    if gi.shape != ri.shape:
        return max(len(gi), len(ri))
    bad = (gi[:, 0] != ri[:, 0]) | (gi[:, 1] != ri[:, 1]) | (gi[:, 3] != ri[:, 3])
    ok = np.flatnonzero(~bad)
    npass = gi[ok, 1]
    gp, rp = _gather(gi[ok, 4], npass), _gather(ri[ok, 4], npass)
    owner = np.repeat(np.arange(len(ok)), npass)
    diff = (np.asarray(gpu.pass_len)[gp] != np.asarray(ref["pass_len"])[rp]) | \
           (np.asarray(gpu.pass_dist)[gp] != np.asarray(ref["pass_dist"])[rp])
    bad[ok[np.unique(owner[diff])]] = True
    ok = np.flatnonzero(~bad)
    ln = gi[ok, 3]
    gd, rd = _gather(gi[ok, 2], ln), _gather(ri[ok, 2], ln)
    owner = np.repeat(np.arange(len(ok)), ln)
    diff = np.asarray(gpu.data)[gd] != np.asarray(ref["data"])[rd]
    bad[ok[np.unique(owner[diff])]] = True
    return int(bad.sum())
