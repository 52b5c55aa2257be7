"""Benchmark of the GAIL HIP encoder path on one MI355X.

Workload (default, BASELINE configs[1]/[2]): DCI 4K flat 3996x2160 12-bit X'Y'Z' frames of
the seeded synthetic generator, NL 6, 32x32 code-blocks. A step encodes one batch of B frames
(per-frame seeds) through colour/ICT -> 9/7 DWT -> quantisation -> Tier-1 (context modelling,
MQ, truncation lengths) to Tier-1 frames in host memory, with the inputs already resident in
HBM. Early termination uses 0.25 x the rate-control lambda of the preceding frames, as the
reference pipeline does. The plain one-frame-at-a-time path is timed beside it.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--batch B]
"""
import argparse
import json
import os
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def dwt_level0_bytes(W, H):
    """Algorithmic HBM bytes of DWT level 0 per frame: 3 planes of u16 codes read, 3 float
    planes written (the ICT is applied on load)."""
    return 3 * (2 + 4) * W * H


def dwt_all_bytes(W, H, levels):
    b = dwt_level0_bytes(W, H)
    w, h = (W + 1) // 2, (H + 1) // 2
    for _ in range(1, levels):
        b += 3 * 8 * w * h
        w, h = (w + 1) // 2, (h + 1) // 2
    return b


def cpu_baseline(frames, fourk, thr, budget_s=20.0):
    """The reference's CPU flow (compiled from its sources into oracle/_ref when present,
    else the oracle port) on this host's cores, over a bounded sample of the same frames."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_bind import Oracle, reference
    ref = reference()
    kind = "reference" if ref is not None else "port"
    lib = ref if ref is not None else Oracle()
    cores = len(os.sched_getaffinity(0))
    done, t0 = 0, time.perf_counter()
    while done < len(frames) and (time.perf_counter() - t0) < budget_s:
        lib.encode_frame_xyz(frames[done], fourk, thr=thr, threads=cores)
        done += 1
    dt = time.perf_counter() - t0
    return {"value": round(done / dt, 3), "unit": "frames/s", "cores": cores, "kind": kind,
            "sample": f"{done} 4K flat frames, colour/ICT + DWT + quantisation + Tier-1 with the same early "
                      f"termination threshold, {cores} threads, {dt:.1f} s"}


def _planar_picture(g, np, seed, W, H, fmt):
    """A planar picture derived from a synthetic frame, planes back to back (uint16):
    gbrp12 (G, B, R at 12 bits) or yuv420p10 (Y full size, Cb/Cr halved, 10 bits)."""
    x = g.synth_frame(seed, W, H, threads=16)
    if fmt == "gbrp12":
        return [x[1], x[2], x[0]]
    return [(x[1] >> 2).astype(np.uint16), np.ascontiguousarray((x[0] >> 2)[::2, ::2]).astype(np.uint16),
            np.ascontiguousarray((x[2] >> 2)[::2, ::2]).astype(np.uint16)]


def _pin(g, np, planes):
    """Copies planes back to back into pinned host memory (the DMA input path)."""
    buf = g.pinned_empty(sum(p.nbytes for p in planes))
    out, off = [], 0
    for p in planes:
        v = buf[off:off + p.nbytes].view(np.uint16).reshape(p.shape)
        v[...] = p
        out.append(v)
        off += p.nbytes
    return out


def served_run(g, np, offered_fps=300.0, seconds=5.0, seed=7):
    """Served form: Poisson arrivals at offered_fps, profiles uniform over the six DCI sizes,
    inputs alternating 12-bit G'B'R' BT.709 and 10-bit Y'CbCr 4:2:0 (limited range, chroma
    left) at the container size; automatic early termination per stream."""
    import random
    import threading
    rng = random.Random(seed)
    srv = g.Server(slots_per_job=8, max_jobs=4, check=True)
    try:
        names = list(g.CONTAINERS)
        streams, pool = {}, {}
        for si, name in enumerate(names):
            W, H, _ = g.CONTAINERS[name]
            for fi, fmt in enumerate(("gbrp12", "yuv420p10")):
                sid = si * 2 + fi
                streams[sid] = g.stream_config(name, fmt, matrix=0, full_range=1 if fmt == "gbrp12" else 0,
                                               chroma=0, preset=0)
                pool[sid] = [_pin(g, np, _planar_picture(g, np, 5000 + sid * 10 + k, W, H, fmt)) for k in range(2)]
        n = int(offered_fps * seconds)
        t, arrivals = 0.0, []
        for i in range(n):
            t += rng.expovariate(offered_fps)
            arrivals.append((t, rng.randrange(len(streams)), rng.randrange(2)))
        print(f"[bench] served: {n} frames, Poisson {offered_fps} fps offered", file=sys.stderr, flush=True)
        t0 = time.perf_counter()
        for i, (at, sid, k) in enumerate(arrivals):
            now = time.perf_counter() - t0
            if at > now:
                time.sleep(at - now)
            srv.submit(i, sid, streams[sid], pool[sid][k], -1.0)
        srv.drain()
        st = srv.stats()
        bad = sum(1 for r in srv.results.values() if r["status"] != 0 or not r.get("compliant", 0))
        return {"offered_fps": offered_fps, "frames": st["frames"], "fps": round(st["fps"], 2),
                "p50_ms": round(st["p50_ms"], 2), "p99_ms": round(st["p99_ms"], 2),
                "queued_p99_ms": round(st["queued_p99_ms"], 2), "mean_batch": round(st["mean_batch"], 2),
                "jobs": st["jobs"], "slots_per_job": st["slots_per_job"],
                "max_frames_in_flight": st["max_frames_in_flight"],
                "slot_gb": round(st["slot_bytes"] / 1e9, 3), "recoded": st["recoded"], "failed_or_noncompliant": bad}
    finally:
        srv.close()


def stress_run(g, np, frames=512, in_flight=256, distinct=8):
    """4K flat 3996x2160 X'Y'Z' stream at 48 fps rate control (651,025-byte budget): PCRD
    truncation on every frame, up to in_flight frames admitted at once (16 jobs x 16 slots)."""
    slots = 16
    srv = g.Server(slots_per_job=slots, max_jobs=max(1, in_flight // slots), check=True)
    try:
        W, H, _ = g.CONTAINERS["4k_flat"]
        cfg = g.stream_config("4k_flat", "xyz12_planar", fps=48.0)
        src = []
        for k in range(distinct):
            x = g.synth_frame(1000 + k, W, H, threads=16)
            src.append(_pin(g, np, [x[0], x[1], x[2]]))
        # warm-up: one frame per job so every job has run once
        for i in range(srv.stats()["jobs"]):
            srv.submit(10 ** 6 + i, 1, cfg, src[i % distinct], -1.0)
        srv.drain()
        st0 = srv.stats()
        print(f"[bench] stress: {frames} 4K frames at 48 fps rate control, {st0['jobs']} jobs x "
              f"{st0['slots_per_job']} slots", file=sys.stderr, flush=True)
        t0 = time.perf_counter()
        for i in range(frames):
            srv.submit(i, 1, cfg, src[i % distinct], -1.0)
        srv.drain()
        dt = time.perf_counter() - t0
        lat = sorted(srv.results[i]["latency_ms"] for i in range(frames))
        sizes = [srv.results[i]["size"] for i in range(frames)]
        bad = sum(1 for i in range(frames) if srv.results[i]["status"] != 0 or not srv.results[i]["compliant"])
        st = srv.stats()
        return {"frames": frames, "fps": round(frames / dt, 2), "p50_ms": round(float(np.percentile(lat, 50)), 2),
                "p99_ms": round(float(np.percentile(lat, 99)), 2), "max_bytes": max(sizes), "cap_bytes": 651025,
                "jobs": st["jobs"], "slots_per_job": st["slots_per_job"], "max_frames_in_flight": in_flight,
                "recoded": st["recoded"] - st0["recoded"], "failed_or_noncompliant": bad}
    finally:
        srv.close()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--batch", type=int, default=16)
    ap.add_argument("--profile", default="4k_flat")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--plain-frames", type=int, default=8)
    ap.add_argument("--graph", type=int, default=0)
    ap.add_argument("--inflight", type=int, default=2)
    ap.add_argument("--served-seconds", type=float, default=5.0)
    ap.add_argument("--stress-frames", type=int, default=512)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU; multi-GPU figures are not part of it "
              "(frames shard across GPUs with no exchange, see DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    import numpy as np

    import libgail_amd as g
    sys.path.insert(0, os.path.join(ROOT, "tests"))

    W, H, fourk = g.CONTAINERS[args.profile]
    B = args.batch
    layout = g.FrameLayout(W, H, fourk)
    print(f"[bench] generating {B} frames {W}x{H}", file=sys.stderr, flush=True)
    frames = [g.synth_frame(1000 + i, W, H, threads=16) for i in range(B)]
    dev = g.Device(0)

    # Early-termination threshold: 0.25 x lambda of a preceding frame (pipeline.cpp:194),
    # taken here from the rate control of the first frame coded without early termination.
    thr = 0.0
    try:
        import ctypes as C
        h = g.host()
        if hasattr(h, "gail_encode_codestream"):
            enc = g.Encoder(dev, layout, g.xyz_settings(layout))
            r = enc.encode(frames[0])
            enc.close()
            thr = 0.25 * g.rate_lambda(layout, r, fps=24.0, mbps=250.0)
    except Exception as e:  # noqa: BLE001
        print(f"[bench] no rate-control lambda ({e}); early termination off", file=sys.stderr)
    print(f"[bench] early termination threshold {thr:.4g}", file=sys.stderr, flush=True)

    # ---- batched form: B frames resident in HBM per batch, `inflight` batches on their own
    # streams so that one batch's Tier-1 tail overlaps the next batch's DWT and context modelling
    nb = max(1, args.inflight)
    tiles = {"64x64": 0, "64x32": 1, "32x64": 2, "32x32": 3}
    # autotuner: the fastest DWT tile shape, context-modelling and MQ workgroup sizes for
    # this batch, measured on it
    first = g.Batch(dev, [layout] * B, [g.xyz_settings(layout)] * B, use_graph=bool(args.graph))
    for i, f in enumerate(frames):
        first.upload(i, f)
        first.set_slope(i, thr)
    chosen, tune_ms = first.autotune(reps=2)
    print(f"[bench] autotuner: {chosen} {json.dumps(tune_ms)}", file=sys.stderr, flush=True)
    batches = [first] + [g.Batch(dev, [layout] * B, [g.xyz_settings(layout)] * B, use_graph=bool(args.graph),
                                 dwt_tile=tiles[chosen["dwt_tile"]], cm_waves=chosen["cm_waves"],
                                 mq_group=chosen["mq_group"]) for _ in range(nb - 1)]
    for b in batches[1:]:
        for i, f in enumerate(frames):
            b.upload(i, f)
            b.set_slope(i, thr)
    # stage times from sequential runs (no other stream competing)
    stage_acc = []
    for w in range(max(args.warmup, 1) + 2):
        batches[0].encode(copy=False)
        if w >= max(args.warmup, 1):
            stage_acc.append(batches[0].stage_ms())
    med = {k: statistics.median(s[k] for s in stage_acc) for k in stage_acc[0]}
    print(f"[bench] stages per batch (ms): {json.dumps({k: round(v, 3) for k, v in med.items()})}", file=sys.stderr,
          flush=True)
    launched = done = 0
    t0 = time.perf_counter()
    while launched < min(nb, args.steps):
        batches[launched % nb].launch()
        launched += 1
    while done < args.steps:
        batches[done % nb].wait(copy=False)
        done += 1
        if launched < args.steps:
            batches[launched % nb].launch()
            launched += 1
    dt = time.perf_counter() - t0
    ms_step = 1000 * dt / args.steps
    fps = B * args.steps / dt
    print(f"[bench] batched: {fps:.1f} frames/s, {ms_step:.2f} ms per batch of {B}, {nb} in flight",
          file=sys.stderr, flush=True)
    for b in batches:
        b.close()

    # ---- plain path: one frame at a time, upload included as in the reference
    enc = g.Encoder(dev, layout, g.xyz_settings(layout))
    enc.set_slope(thr)
    lat = []
    for i in range(max(args.plain_frames, 1) + 1):
        t1 = time.perf_counter()
        enc.encode(frames[i % B])
        if i:
            lat.append(1000 * (time.perf_counter() - t1))
    enc.close()
    plain_fps = 1000.0 / statistics.mean(lat)

    # roofline: DWT level 0 of the whole batch in one launch (HIP events on the batch stream)
    stage_note = "graph replay: per-stage events off" if args.graph else ""
    roof = None
    if med.get("dwt0", 0) > 0:
        achieved = B * dwt_level0_bytes(W, H) / (med["dwt0"] * 1e-3) / 1e9
        roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None, "kernel": "dwt_level_kernel<level 0>"}

    served = stress = None
    if args.served_seconds > 0:
        served = served_run(g, np, seconds=args.served_seconds)
        print(f"[bench] served: {json.dumps(served)}", file=sys.stderr, flush=True)
    if args.stress_frames > 0:
        stress = stress_run(g, np, frames=args.stress_frames)
        print(f"[bench] stress: {json.dumps(stress)}", file=sys.stderr, flush=True)

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(frames[:4], fourk, thr)

    line = {
        "metric": "4K DCI encode fps, batched/served vs plain, 1 MI355X; p50/p99 frame latency",
        "value": round(fps, 2), "unit": "frames/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(ms_step, 3), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32+int", "data": "synthetic",
        "config": {"workload": f"batched {args.profile} {W}x{H} x{B} frames, Tier-1 to host", "frames_per_step": B,
                   "batches_in_flight": nb,
                   "levels": layout.levels, "code_blocks_per_frame": layout.num_cblks,
                   "slope_threshold": thr, "graph": bool(args.graph), "autotuned": chosen},
        "plain_fps": round(plain_fps, 2),
        "batched_over_plain": round(fps / plain_fps, 3),
        "plain_latency_ms": {"p50": round(float(np.percentile(lat, 50)), 3),
                             "p99": round(float(np.percentile(lat, 99)), 3)},
        "stage_ms_per_batch": {k: round(v, 3) for k, v in med.items()},
        "tier1_symbols_per_s": round(med["symbols"] / (ms_step * 1e-3), 1) if med.get("symbols") else None,
        "autotune_ms": tune_ms,
        "roofline": roof,
        "cpu_baseline": cpu,
        "served": served,
        "stress_4k_48fps": stress,
    }
    if stage_note:
        line["note"] = stage_note
    dev.close()
    print(json.dumps(line))


if __name__ == "__main__":
    main()
