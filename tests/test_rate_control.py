"""Rate control (PCRD-opt) and Tier-2 of libgail.so, pinned by what does not come from this
build. The reference's rate_control.cpp and t2.cpp do not compile here (t2.cpp calls
buildVersion(), defined only over the CMake-generated gail_build_version.h), so:

- the reference's own DCI checker (codestream_check.cpp:43-280, compiled from its sources
  into oracle/_ref) passes every codestream and measures its total, component and
  tile-part bytes, which must meet the budgets;
- the reference's own header and packet parser (codestream_parse.cpp, t2_parse.cpp) reads
  every code-block back: its pass count is the one rate control chose, its bytes the first
  bytes of the oracle's Tier-1 codeword, up to the truncation length of that pass;
- the reference's known answers for dciLimits (tests/test_main.cpp:1618-1620) and for
  budgets below the headers (:1622-1647);
- the conditions that define PCRD-opt, checked from their definitions rather than from the
  build's code: every kept truncation point is a vertex of its block's upper concave hull of
  (bytes, weighted distortion reduction) -- the vertex test here is the two-sided slope
  condition, not the incremental scan of rate_control.cpp:38-86 -- each block keeps exactly
  the vertices whose slope is >= its component's lambda, and the remaining distortion the
  rate control reports is the gain of the vertices it dropped;
- the reference's "fills the frame to within a few percent of the cap" check
  (test_main.cpp:1521-1560: total + cap/25 >= cap when the frame has more than the cap to
  code), here on the 48 fps / 651,041-byte stress budget and on noise."""
import numpy as np
import pytest

import libgail_amd as g


def _plane_scale(nbps, k):
    """2^(2 plane) / 4096 of pass k (the first pass codes plane nbps-1, then three per plane)."""
    plane = np.where(k == 0, nbps - 1, nbps - 2 - (k - 1) // 3)
    return np.exp2(2.0 * plane) / 4096.0


def _hull_vertices(rate, dist):
    """Indices j (1-based pass counts) of the upper concave hull vertices of the points
    (0, 0), (rate[0], dist[0]), ..., each point a vertex iff every slope into it from an
    earlier point exceeds every slope out of it to a later one, and its gain is positive.
    Returns (pass counts, slopes from the previous vertex, gains from the previous vertex)."""
    R = np.concatenate([[0.0], rate.astype(np.float64)])
    D = np.concatenate([[0.0], dist])
    n = len(R)
    dR = R[None, :] - R[:, None]          # dR[i, j] = R[j] - R[i]
    dD = D[None, :] - D[:, None]
    with np.errstate(divide="ignore", invalid="ignore"):
        s = np.where(dR > 0, dD / np.where(dR > 0, dR, 1), np.where(dD > 0, np.inf, -np.inf))
    verts = []
    for j in range(1, n):
        left = s[:j, j].min()
        right = s[j, j + 1:].max() if j + 1 < n else -np.inf
        if left > 0 and left > right:
            verts.append(j)
    slopes, gains, prev = [], [], 0
    for j in verts:
        dr = R[j] - R[prev]
        slopes.append(np.inf if dr <= 0 else (D[j] - D[prev]) / dr)
        gains.append(D[j] - D[prev])
        prev = j
    return verts, slopes, gains


def _check_pcrd(layout, t1, res):
    """PCRD-opt conditions of a rate-control result (see the module docstring)."""
    info = np.asarray(t1["info"])
    comps = np.array([layout.bands[int(c[0])].comp for c in layout.cblks])
    bands = [layout.bands[int(c[0])] for c in layout.cblks]
    lam = res["lambdas"]
    remaining = 0.0
    kept_vertex, checked = 0, 0
    for i in range(len(info)):
        nbps, np_, _, _, pb = (int(v) for v in info[i])
        p = int(res["passes"][i])
        if np_ == 0:
            assert p == 0
            continue
        k = np.arange(np_)
        b = bands[i]
        scale = float(b.delta) * float(b.delta) * b.weight
        dist = np.cumsum(np.asarray(t1["pass_dist"][pb:pb + np_], np.float64) * _plane_scale(nbps, k) * scale)
        verts, slopes, gains = _hull_vertices(np.asarray(t1["pass_len"][pb:pb + np_]), dist)
        want = 0
        for v, sl in zip(verts, slopes):
            if sl < lam[comps[i]]:
                break
            want = v
        assert p == want, f"block {i}: keeps {p} passes, the hull at lambda {lam[comps[i]]} gives {want}"
        if p:
            assert p in verts
            kept_vertex += 1
        remaining += sum(gn for v, gn in zip(verts, gains) if v > p)
        checked += 1
    assert checked > 0
    assert res["distortion"] == pytest.approx(remaining, rel=1e-9, abs=1e-6)
    return kept_vertex


def _check_codestream(ref, layout, t1, passes, cs, fps, max_total, max_comp):
    """The reference's checker and parser on a codestream built with `passes` per block."""
    ok, errors, facts = ref.check_codestream(cs, fps)
    assert ok, errors
    assert facts["total"] == len(cs) <= max_total
    for c in range(3):
        assert facts[f"comp{c}"] <= max_comp
    assert (facts["width"], facts["height"]) == (layout.width, layout.height)
    parsed = ref.parse_blocks(cs)
    assert parsed is not None
    _, _, _, codes, _ = parsed
    info = np.asarray(t1["info"])
    assert np.array_equal(codes["passes"].astype(np.int64), np.asarray(passes, np.int64))
    data = np.frombuffer(cs, np.uint8)
    for i in np.flatnonzero(codes["passes"]):
        nbps, np_, off, _, pb = (int(v) for v in info[i])
        p = int(codes["passes"][i])
        ln = int(t1["pass_len"][pb + p - 1])
        assert int(codes["len"][i]) == ln and int(codes["nbps"][i]) == nbps
        o = int(codes["offset"][i])
        assert bytes(data[o:o + ln]) == bytes(t1["data"][off:off + ln])
    return facts


def test_dci_limits_known_answers():
    assert g.dci_limits(24, 230) == (1197900, 1041650)  # test_main.cpp:1618
    assert g.dci_limits(24, 0.0001)[0] == 0 and g.dci_limits(24, 230, 0.0001)[1] == 0
    assert g.dci_limits(0, 230)[0] == 0 and g.dci_limits(24, 0)[0] == 0
    assert g.dci_limits(24, 250, 200) == (1302067, 1041650)
    assert g.dci_limits(48, 250, 200) == (651025, 520817)


@pytest.fixture(scope="module")
def frames(oracle):
    """Oracle Tier-1 of the synthetic 2K and 4K flat frames and of 2K noise."""
    out = {}
    for name, seed in (("2k_flat", 1), ("4k_flat", 1000)):
        W, H, k = g.CONTAINERS[name]
        out[name] = (g.FrameLayout(W, H, k), oracle.encode_frame_xyz(g.synth_frame(seed, W, H), k, threads=8))
    W, H, k = g.CONTAINERS["2k_flat"]
    noise = np.random.default_rng(7).integers(0, 4096, (3, H, W), dtype=np.uint16)
    out["2k_noise"] = (g.FrameLayout(W, H, k), oracle.encode_frame_xyz(noise, k, threads=8))
    return out


@pytest.mark.parametrize("name,fps,mbps", [("2k_flat", 24.0, 250.0), ("2k_flat", 48.0, 250.0),
                                           ("4k_flat", 48.0, 250.0), ("4k_flat", 24.0, 125.0),
                                           ("2k_noise", 24.0, 250.0)])
def test_pcrd_at_dci_budgets(frames, ref, name, fps, mbps):
    L, t1 = frames[name]
    max_total, max_comp = g.dci_limits(fps, mbps, 200.0)
    res = g.rate_allocate(L, t1, max_total, max_comp)
    kept = _check_pcrd(L, t1, res)
    cs, lam, _ = g.encode_codestream(L, t1, fps=fps, mbps=mbps, tag="hip")
    assert lam == pytest.approx(min(res["lambdas"]))
    facts = _check_codestream(ref, L, t1, res["passes"], cs, fps, max_total, max_comp)
    all_bytes = int(np.asarray(t1["data"]).size)
    print(f"[rc] {name} {fps:g} fps {mbps:g} Mbit/s: {len(cs)} of {max_total} bytes, lambda {res['lambdas']}, "
          f"{kept} blocks on a hull vertex, {res['iterations']} measurements, components {facts['comp0']} "
          f"{facts['comp1']} {facts['comp2']}, Tier-1 bytes {all_bytes}")
    assert res["total"] >= len(cs)
    if all_bytes > max_total:  # the cap binds: filled to within 4 % (test_main.cpp:1556)
        assert len(cs) + max_total // 25 >= max_total


def test_pcrd_component_caps(frames, ref):
    """Component caps well below a third of the frame budget: each component's own lambda
    rises until its bytes fit, every component stays under its cap, and the hull conditions
    hold per component."""
    L, t1 = frames["2k_flat"]
    max_total, max_comp = 1302067, 250000
    res = g.rate_allocate(L, t1, max_total, max_comp)
    _check_pcrd(L, t1, res)
    assert all(c <= max_comp for c in res["comp"]) and res["total"] <= max_total
    assert len(set(res["lambdas"].tolist())) > 1
    cs = g.build_codestream(L, g.with_passes(t1, res["passes"]), tag="hip")
    facts = _check_codestream(ref, L, t1, res["passes"], cs, 24.0, max_total, max_comp)
    print(f"[rc] component caps {max_comp}: {facts['comp0']} {facts['comp1']} {facts['comp2']}, "
          f"lambdas {res['lambdas']}")
    # the binding components are filled to within 4 % of their cap
    for c in range(3):
        assert res["comp"][c] + max_comp // 25 >= max_comp


def test_budgets_at_and_below_the_headers(frames, ref):
    """A zero budget drops every pass and the codestream is the headers alone; a budget of
    exactly the header size is met (test_main.cpp:1622-1647)."""
    L, t1 = frames["2k_flat"]
    headers = g.empty_codestream_bytes(L)
    assert 100 < headers < 2000
    res = g.rate_allocate(L, t1, 0, 0)
    assert not res["passes"].any() and res["total"] == headers
    res = g.rate_allocate(L, t1, headers, headers)
    assert res["total"] <= headers
    cs = g.build_codestream(L, g.with_passes(t1, res["passes"]), tag="")
    ok, errors, facts = ref.check_codestream(cs, 24.0)
    assert ok, errors


def test_pcrd_check_catches_planted_errors(frames):
    """The condition check above is not vacuous: a result with one block keeping a pass
    more or a pass fewer than the hull gives, or a remaining distortion off by one block's
    gain, fails it."""
    L, t1 = frames["2k_flat"]
    max_total, max_comp = g.dci_limits(24.0, 250.0, 200.0)
    res = g.rate_allocate(L, t1, max_total, max_comp)
    _check_pcrd(L, t1, res)
    info = np.asarray(t1["info"])
    kept = np.flatnonzero((res["passes"] > 0) & (res["passes"] < info[:, 1]))
    assert len(kept) > 0
    for delta in (+1, -1):
        bad = dict(res)
        bad["passes"] = res["passes"].astype(np.int64)
        bad["passes"][kept[len(kept) // 2]] += delta
        with pytest.raises(AssertionError):
            _check_pcrd(L, t1, bad)
    bad = dict(res)
    bad["distortion"] = res["distortion"] * (1 + 1e-6) + 1.0
    with pytest.raises(AssertionError):
        _check_pcrd(L, t1, bad)
