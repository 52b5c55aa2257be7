"""Benchmark of the GAIL HIP encoder path on one MI355X.

Headline (`value`): sustained 4K DCI encode rate of the served form, frames/s of finished
codestreams: DCI 4K flat 3996x2160 12-bit X'Y'Z' frames of the seeded synthetic generator
(NL 6, 32x32 code-blocks), 24 fps rate control at 250 Mbit/s (1,302,067-byte budget), every
codestream built and DCI-checked. The input frames are in pinned host memory and each crosses
PCIe inside the timed region, as the reference's pipeline uploads every frame; up to 256
frames in flight (--stream-inflight: every frame slot of the server's 16 jobs, 170 GB of
HBM). A step is --step-frames (32) frames of the stream: --warmup untimed steps, then
--steps timed ones. The same run from frames resident in HBM (copied within HBM into the job
slots) is reported beside it. Colour/ICT, 9/7 DWT, quantisation, Tier-1 and PCRD rate control
run on the GPU, Tier-2 on the host pool, with early termination at 0.25 x the stream's last
lambda as the reference pipeline does. `plain_fps` is the same encode one frame at a time
(the reference's slot flow); `tier1_batched_fps` the GPU stages alone on batches of 16 frames
resident in HBM. `stress_4k_48fps` is the 48 fps rate-control stream with up to 256 frames in
flight; `served` the mixed-profile request loop at 300 fps of Poisson arrivals.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--batch B]
"""
import argparse
import ctypes
import json
import os
import statistics
import sys
import threading
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# The served form runs its 16 jobs on streams of their own; with 16 hardware queues (HIP's
# default is 4) their kernels and copies interleave on the device instead of queueing behind
# one another: the PCIe stream measured 937 -> 1,058 frames/s, the HBM-resident one 1,410 ->
# 1,524 (32 queues: 701 and 1,057). Read by the HIP runtime when it starts, so set here, before
# anything loads it.
os.environ["GPU_MAX_HW_QUEUES"] = "16"

HBM_PEAK_GBS = 8000.0
# HBM bytes one DWT level-0 launch over 16 4K flat frames moves, per tile shape, from the
# FETCH_SIZE and WRITE_SIZE counter runs in profiles/pmc_traffic_dwt_level0_4k_flat_batch16.md,
# each scaled by what the counter reads for a 1 GiB device copy.
DWT0_TRAFFIC_4K_FLAT_16 = {"64x16": 2491427489, "64x32": 2502515919, "strip0_120x8": 2493219461,
                           "stage0_480x8": 2498583259}


def dwt_level0_kernel(shape):
    """The kernel that runs DWT level 0 for an autotuner tile shape name."""
    if shape.startswith("stage0_"):
        return f"dwt_stage0_kernel<level 0, {shape[len('stage0_'):]} groups of 120x8 strips, line stores through LDS>"
    if shape.startswith("strip0b_"):
        return f"dwt_strip0b_kernel<level 0, {shape[len('strip0b_'):]} strips, input rows loaded ahead, buffer addressing>"
    if shape.startswith("strip0_"):
        return f"dwt_strip0_kernel<level 0, {shape[len('strip0_'):]} strips, input rows loaded ahead>"
    if shape.startswith("strip_"):
        return f"dwt_strip_kernel<level 0, {shape[len('strip_'):]} strips>"
    return f"dwt_level_kernel<level 0, {shape} tiles>"


def dwt_level0_bytes(W, H):
    """Algorithmic HBM bytes of DWT level 0 per frame: 3 planes of u16 codes read, 3 float
    planes written (the ICT is applied on load)."""
    return 3 * (2 + 4) * W * H


def hbm_copy_gbs(g, dev):
    """HBM rate a plain device-to-device copy reaches on this GPU: 1 GiB copied (2 GiB moved)
    with hipMemcpyAsync, timed with HIP events on the null stream, median of 10. It puts a
    kernel's fraction of the published 8 TB/s beside what a copy kernel achieves."""
    hip = ctypes.CDLL("libamdhip64.so")
    vp = ctypes.c_void_p
    hip.hipMemcpyAsync.argtypes = [vp, vp, ctypes.c_size_t, ctypes.c_int, vp]
    hip.hipEventCreate.argtypes = [ctypes.POINTER(vp)]
    hip.hipEventRecord.argtypes = [vp, vp]
    hip.hipEventSynchronize.argtypes = [vp]
    hip.hipEventElapsedTime.argtypes = [ctypes.POINTER(ctypes.c_float), vp, vp]
    hip.hipEventDestroy.argtypes = [vp]
    hip.hipDeviceSynchronize.argtypes = []
    n = 1 << 30
    p = g.plugin()
    a, b = p.device_alloc(dev.handle, n), p.device_alloc(dev.handle, n)
    if not a or not b:
        return None
    e0, e1 = vp(), vp()
    try:
        if hip.hipEventCreate(ctypes.byref(e0)) or hip.hipEventCreate(ctypes.byref(e1)):
            return None
        ts = []
        for i in range(11):
            if hip.hipEventRecord(e0, None) or hip.hipMemcpyAsync(vp(b), vp(a), n, 3, None) \
                    or hip.hipEventRecord(e1, None) or hip.hipEventSynchronize(e1):
                return None
            ms = ctypes.c_float()
            hip.hipEventElapsedTime(ctypes.byref(ms), e0, e1)
            if i:
                ts.append(ms.value)
        return 2 * n / (statistics.median(ts) * 1e-3) / 1e9
    finally:
        hip.hipDeviceSynchronize()
        for e in (e0, e1):
            if e:
                hip.hipEventDestroy(e)
        p.device_free(vp(a))
        p.device_free(vp(b))


def cpu_baseline(frames, fourk, thr, budget_s=12.0):
    """The reference's CPU flow (compiled from its sources into oracle/_ref when present,
    else the oracle port) on this host's cores, over a bounded sample of the same frames:
    the given frames in turn, again and again, until budget_s has passed (at least each once)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_bind import Oracle, reference
    ref = reference()
    kind = "reference" if ref is not None else "port"
    lib = ref if ref is not None else Oracle()
    cores = len(os.sched_getaffinity(0))
    done, t0 = 0, time.perf_counter()
    while done < len(frames) or (time.perf_counter() - t0) < budget_s:
        lib.encode_frame_xyz(frames[done % len(frames)], fourk, thr=thr, threads=cores)
        done += 1
    dt = time.perf_counter() - t0
    return {"value": round(done / dt, 3), "unit": "frames/s", "cores": cores, "kind": kind,
            "sample": f"{done} 4K flat frames ({len(frames)} distinct, in turn), colour/ICT + DWT + quantisation + "
                      f"Tier-1 with the same early "
                      f"termination threshold, {cores} threads, {dt:.1f} s (no rate control or Tier-2: the "
                      f"reference's t2.cpp needs a generated header and does not build here)"}


def plain_parity(np, gpu, frame, fourk, thr, levels):
    """Checks of one plain-path 4K frame against the CPU path, run in the CPU-baseline
    leg with the compiled reference (the oracle where it is not built) as the checker:
    quantised-index mismatches, Tier-1 bit-exactness, and the largest difference between the
    device's float32 DWT and a float64 DWT of the same ICT planes."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_bind import Oracle, reference
    from t1_compare import compare_tier1
    oracle = Oracle()
    ref = reference()
    lib = ref if ref is not None else oracle
    cores = len(os.sched_getaffinity(0))
    want = lib.encode_frame_xyz(frame, fourk, thr=thr, threads=cores)
    gi = (gpu.debug["coeffs"].reshape(-1) & 0x7FFFFFFF) >> 6
    ri = (want["coeffs"].reshape(-1) & 0x7FFFFFFF) >> 6
    d = np.abs(gi.astype(np.int64) - ri.astype(np.int64))
    bad, _ = compare_tier1(gpu, want)
    ict = oracle.xyz_to_ict(frame)
    f64 = max(float(np.max(np.abs(oracle.fdwt_f64(ict[c], levels) - gpu.debug["wavelet"][c].astype(np.float64))))
              for c in range(3))
    return {"checker": "reference" if ref is not None else "port", "quant_mismatch_pm1": int(np.sum(d == 1)),
            "quant_mismatch_gt1": int(np.sum(d > 1)), "sign_mismatches": int(np.sum(
                (gpu.debug["coeffs"].reshape(-1) >> 31) != (want["coeffs"].reshape(-1) >> 31))),
            "tier1_blocks_differing": int(bad), "tier1_bitexact": bad == 0,
            "dwt_fp32_vs_fp64_max_abs": round(f64, 6)}


def decode_bench(g, np, dev, layout, frames, decoders=4, rounds=4, cpu=True):
    """The HIP decoder on the reference's decode workload: 4K frames at 250 Mbit/s (24 fps
    rate control on the device), each frame's code-blocks located (the Tier-1 output read as
    its codestream: the same block bytes the codestream carries), decoded to 12-bit X'Y'Z'
    words (highBits). `decoders` decoders on their own threads and streams; pictures return
    to pinned host memory (the reference's setting), then the same with codestreams and
    pictures resident in HBM. Stage times are medians of single frames in flight."""
    p = g.plugin()
    W, H = layout.width, layout.height
    n = len(frames)
    b = g.Batch(dev, [layout] * n, [g.with_rate_control(g.xyz_settings(layout))] * n)
    for i, f in enumerate(frames):
        b.upload(i, f)
    res, _ = b.encode()
    b.close()
    items = [(np.ascontiguousarray(r.data), g.block_table(r)) for r in res]
    cs_bytes = [len(g.build_codestream(layout, r)) for r in res]
    mbps = statistics.mean(cs_bytes) * 8 * 24 / 1e6
    pic = W * H * 6
    hin = [p.host_alloc(max(d.size, 1)) for d, _ in items]
    dinb = [p.device_alloc(dev.handle, max(d.size, 1)) for d, _ in items]
    hout = [p.host_alloc(pic) for _ in range(decoders)]
    dout = [p.device_alloc(dev.handle, pic) for _ in range(decoders)]
    decs = [g.Decoder(dev, layout, 0, True) for _ in range(decoders)]
    try:
        if not all(hin) or not all(dinb) or not all(hout) or not all(dout):
            raise RuntimeError("decode bench buffers could not be allocated")
        for (d, _), h, dd in zip(items, hin, dinb):
            ctypes.memmove(h, d.ctypes.data, d.size)
            if p.copy(ctypes.c_void_p(dd), ctypes.c_void_p(h), d.size):
                raise RuntimeError("copy failed")
        # single frame in flight: stage times
        stages = []
        for k, (d, blocks) in enumerate(items):
            _, tm, _ = decs[0].decode((hin[k], d.size), blocks, out=(hout[0], W * 6))
            stages.append(tm)
        stage_ms = {key: round(statistics.median(t[key] for t in stages[1:] or stages), 3)
                    for key in ("t1", "dwt", "color", "total")}

        def run(src, dst, total):
            nxt = [0]
            lock = threading.Lock()
            errors = []

            def worker(j):
                try:
                    while True:
                        with lock:
                            k = nxt[0]
                            nxt[0] += 1
                        if k >= total:
                            return
                        d, blocks = items[k % n]
                        decs[j].decode((src[k % n], d.size), blocks, out=(dst[j], W * 6))
                except Exception as e:  # noqa: BLE001
                    errors.append(e)

            ts = [threading.Thread(target=worker, args=(j,)) for j in range(decoders)]
            t0 = time.perf_counter()
            for t in ts:
                t.start()
            for t in ts:
                t.join()
            dt = time.perf_counter() - t0
            if errors:
                raise errors[0]
            return total / dt

        # the floor of one frame's Tier-1 time: its longest code-block decoded alone
        d0, blocks0 = items[0]
        alone = blocks0.copy()
        big = int(np.argmax(alone["len"]))
        keep = alone[big].copy()
        alone["passes"] = 0
        alone[big] = keep
        t_alone = min(decs[0].decode((dinb[0], d0.size), alone, out=(dout[0], W * 6))[1]["t1"] for _ in range(3))
        run(hin, hout, decoders)  # warm-up
        total = n * rounds
        host_fps = run(hin, hout, total)
        hbm_fps = run(dinb, dout, total)
    finally:
        for dcd in decs:
            dcd.close()
        for h in hin + hout:
            if h:
                p.host_free(h)
        for dd in dinb + dout:
            if dd:
                p.device_free(dd)
    out = {"fps_host_memory": round(host_fps, 2), "fps_hbm_resident": round(hbm_fps, 2),
           "decoders": decoders, "frames": total, "distinct_frames": n,
           "stage_ms_one_in_flight": stage_ms, "mbit_per_s_at_24fps": round(mbps, 1),
           "longest_block_alone": {"bytes": int(keep["len"]), "tier1_ms": round(t_alone, 3)},
           "reference_published_fps": {"RTX 3090 CUDA 4K": 121.4, "RTX 5060 Ti CUDA 4K": 57.3},
           "workload": "4K flat 3996x2160, 24 fps / 250 Mbit/s device rate control, decoded to xyz12 words"}
    if cpu:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        from oracle_bind import reference
        ref = reference()
        if ref is not None:
            cores = len(os.sched_getaffinity(0))
            done, t0 = 0, time.perf_counter()
            while done < n and time.perf_counter() - t0 < 15.0:
                d, blocks = items[done]
                ref.decode_frame(W, H, 1, d, blocks, threads=cores)
                done += 1
            dt = time.perf_counter() - t0
            out["cpu_baseline"] = {"value": round(done / dt, 3), "unit": "frames/s", "cores": cores,
                                   "kind": "reference", "sample": f"{done} of the same frames through the "
                                   f"reference's CPU decode slot (cpu_decode_backend.cpp), {cores} threads"}
    return out


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


def _pct(np, v, q):
    return round(float(np.percentile(v, q)), 2) if len(v) else None


def api_encoder_run(g, np, frames, in_flight, count, warm=8):
    """The drop-in encoder API (gail_encoder_send / receive, the reference's gail.h:329-362)
    on a 4K flat X'Y'Z' stream at 24 fps rate control with in_flight frames in flight (the
    reference's benchmark runs 4): one thread sends frames from pinned host memory, the
    calling thread receives the codestreams in order; every frame DCI-checked."""
    cfg = g.stream_config("4k_flat", "xyz12_planar", fps=24.0)
    cfg.frames_in_flight = in_flight
    cfg.check = 1
    enc = g.ApiEncoder(cfg)
    total = warm + count
    sent = [0.0] * total
    err = []

    def sender():
        try:
            for i in range(total):
                sent[i] = time.perf_counter()
                enc.send(frames[i % len(frames)])
            enc.end()
        except Exception as e:  # reported below; the receiver then sees the end or a failure
            err.append(str(e))
            enc.end()

    th = threading.Thread(target=sender)
    th.start()
    lat, bad, t0, n, sizes = [], 0, None, 0, []
    try:
        while True:
            r = enc.receive()
            if r is None:
                break
            data, idx, compliant = r
            now = time.perf_counter()
            if n == warm - 1:
                t0 = now
            if n >= warm:
                lat.append((now - sent[idx]) * 1e3)
                sizes.append(len(data))
                bad += 0 if compliant else 1
            n += 1
        th.join()
    finally:
        enc.close()
    if err or n != total or t0 is None:
        return {"error": err[0] if err else f"{n} of {total} frames came back"}
    dt = time.perf_counter() - t0
    return {"frames": count, "warm_frames": warm, "frames_in_flight": in_flight, "seconds": round(dt, 3),
            "fps": round(count / dt, 2), "p50_ms": _pct(np, lat, 50), "p99_ms": _pct(np, lat, 99),
            "max_bytes": max(sizes), "noncompliant": bad,
            "inputs": "pinned host memory, copied by gail_encoder_send as the reference's API does",
            "reference_published_fps": {"RTX 3090 CUDA 4K": 171.1, "RTX 5060 Ti CUDA 4K": 119.5, "both GPUs": 246.7}}


def served_run(g, np, offered_fps=300.0, seconds=5.0, seed=7):
    """Served form: Poisson arrivals at offered_fps, profiles uniform over the six DCI sizes,
    inputs alternating 12-bit G'B'R' BT.709 and 10-bit Y'CbCr 4:2:0 (limited range, chroma
    left) at the container size; automatic early termination per stream."""
    import random
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
                "jobs_at_end": {"2k": st["jobs_2k"], "4k": st["jobs_4k"]}, "jobs_made": st["jobs_made"],
                "jobs_retired": st["jobs_retired"], "max_jobs": st["max_jobs"],
                "slots_per_job": {"2k": st["slots_per_job_2k"], "4k": st["slots_per_job"]},
                "max_frames_in_flight": st["max_frames_in_flight"],
                "slot_gb": {"2k": round(st["slot_bytes_2k"] / 1e9, 3), "4k": round(st["slot_bytes"] / 1e9, 3)},
                "peak_job_gb": round(st["peak_job_bytes"] / 1e9, 2),
                "admission": "jobs of 2K or 4K slots made as frames of their class wait, against the device's "
                             "free memory less max(512 MiB, 6 %)",
                "recoded": st["recoded"], "failed_or_noncompliant": bad}
    finally:
        srv.close()


def mixed_batch_run(g, np, dev, steps=6, per_profile=16):
    """Batched form over a 64-frame mix: 16 frames each of 2K flat, 2K scope, 4K scope and 4K
    full (per-frame seeds) in one job, one DWT launch per level and one persistent Tier-1
    launch for all their code-blocks, the chain replayed as a HIP graph, tile shapes and
    workgroup sizes autotuned on the batch. Tier-1 output to host memory, two jobs in flight."""
    names = ["2k_flat", "2k_scope", "4k_scope", "4k_full"]
    layouts, frames = [], []
    for n in names:
        W, H, k = g.CONTAINERS[n]
        L = g.FrameLayout(W, H, k)
        for i in range(per_profile):
            layouts.append(L)
            frames.append(g.synth_frame(7000 + 100 * names.index(n) + i, W, H, threads=16))
    settings = [g.xyz_settings(L) for L in layouts]
    first = g.Batch(dev, layouts, settings, use_graph=True)
    for i, f in enumerate(frames):
        first.upload(i, f)
        first.set_slope(i, 0.0)
    chosen, _ = first.autotune(reps=1)
    tiles = {name: i for i, name in enumerate(g.Batch.DWT_SHAPES)}
    second = g.Batch(dev, layouts, settings, use_graph=True, dwt_tile=tiles[chosen["dwt_tile"]],
                     cm_waves=chosen["cm_waves"], mq_group=chosen["mq_group"], mq_read=chosen["mq_read"])
    for i, f in enumerate(frames):
        second.upload(i, f)
        second.set_slope(i, 0.0)
    bs = [first, second]
    for b in bs:  # warm-up (captures the graphs)
        b.encode(copy=False)
    t0 = time.perf_counter()
    for b in bs[:min(2, steps)]:
        b.launch()
    for s in range(steps):  # steps runs in all, two in flight
        bs[s % 2].wait(copy=False)
        if s + 2 < steps:
            bs[s % 2].launch()
    dt = time.perf_counter() - t0
    blocks = sum(L.num_cblks for L in layouts)
    px = sum(L.width * L.height for L in layouts)
    fps = len(frames) * steps / dt
    for b in bs:
        b.close()
    return {"frames_per_job": len(frames), "profiles": names, "jobs_timed": steps, "fps": round(fps, 2),
            "fps_4k_equivalent": round(fps * (px / len(frames)) / (3996 * 2160), 2),
            "code_blocks_per_launch": blocks, "graph": True, "autotuned": chosen,
            "ms_per_job": round(1000 * dt / steps, 2), "early_termination": "off"}


def cpu_config1(g, np):
    """One 2K flat 1998x1080 frame through the CPU restatement of the reference encoder:
    colour/ICT, DWT, quantisation and Tier-1 (oracle, on the host's cores), then this build's
    rate control and Tier-2 on one thread: wall time and the codestream against the cap."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_bind import reference, Oracle
    ref = reference()
    lib = ref if ref is not None else Oracle()
    W, H, k = g.CONTAINERS["2k_flat"]
    x = g.synth_frame(1, W, H, threads=16)
    cores = len(os.sched_getaffinity(0))
    t0 = time.perf_counter()
    t1 = lib.encode_frame_xyz(x, k, threads=cores)
    t_mid = time.perf_counter()
    cs, _, _ = g.encode_codestream(g.FrameLayout(W, H, k), t1, fps=24.0, mbps=250.0, tag="cpu")
    t_end = time.perf_counter()
    return {"frame": "2K flat 1998x1080, seed 1", "wall_ms": round(1000 * (t_end - t0), 1),
            "tier1_ms": round(1000 * (t_mid - t0), 1), "rate_control_tier2_ms": round(1000 * (t_end - t_mid), 1),
            "fps": round(1.0 / (t_end - t0), 2), "cores": cores, "bytes": len(cs), "cap_bytes": 1302067,
            "dci_check": g.check_codestream(cs, 24.0, 0) or "pass",
            "kind": "reference" if ref is not None else "port"}


class StreamBench:
    """One server with jobs of 16 frame slots, as many as frames_in_flight needs (16 jobs for
    256 4K frames, ~170 GB of HBM), for the 4K flat streams: a closed loop with at most a given
    number of frames in flight."""

    def __init__(self, g, np, distinct=8, frames_in_flight=256):
        self.g, self.np = g, np
        self._sem = None
        self.jobs = max(16, -(-frames_in_flight // 16))
        self.srv = g.Server(slots_per_job=16, max_jobs=self.jobs, check=True, on_done=self._done)
        W, H, _ = g.CONTAINERS["4k_flat"]
        self.src = []  # the same frames in pinned host memory (PCIe in the timed loop)
        self.dev = []  # and resident in HBM (copied within HBM into the slots)
        for k in range(distinct):
            x = g.synth_frame(1000 + k, W, H, threads=16)
            self.src.append(_pin(g, np, [x[0], x[1], x[2]]))
            self.dev.append(g.DeviceFrame([x[0], x[1], x[2]]))
        self.next_id = 0
        # the server makes its jobs as frames wait: 16 frames per job at once make them all
        # (and run each once) before anything is timed
        cfg = g.stream_config("4k_flat", "xyz12_planar", fps=24.0)
        for i in self._ids(16 * self.jobs):
            self.srv.submit(i, 100, cfg, self.src[i % distinct], -1.0)
        self.srv.drain()

    def _ids(self, n):
        ids = list(range(self.next_id, self.next_id + n))
        self.next_id += n
        return ids

    def _done(self, _id, _r):
        sem = self._sem
        if sem is not None:
            sem.release()

    def run(self, fps_rc, frames, in_flight, stream, resident=True, warm_frames=4):
        """frames of a 4K flat stream at fps_rc rate control, at most in_flight at once, from
        frames resident in HBM (resident) or from pinned host memory, after warm_frames
        untimed frames of the same stream (they also settle its early-termination threshold)."""
        g, np = self.g, self.np
        pool = self.dev if resident else self.src
        cfg = g.stream_config("4k_flat", "xyz12_planar", fps=fps_rc)
        self._sem = threading.Semaphore(in_flight)
        for i in self._ids(max(warm_frames, 4)):
            self._sem.acquire()
            self.srv.submit(i, stream, cfg, pool[i % len(pool)], -1.0)
        self.srv.drain()
        self._sem = None
        st0 = self.srv.stats()
        ids = self._ids(frames)
        self._sem = threading.Semaphore(in_flight)
        t0 = time.perf_counter()
        for i in ids:
            self._sem.acquire()
            self.srv.submit(i, stream, cfg, pool[i % len(pool)], -1.0)
        self.srv.drain()
        dt = time.perf_counter() - t0
        self._sem = None
        res = [self.srv.results[i] for i in ids]
        lat = [r["latency_ms"] for r in res]
        bad = sum(1 for r in res if r["status"] != 0 or not r.get("compliant", 0))
        cap = int(250e6 / (8 * fps_rc)) - 16
        st = self.srv.stats()
        return {"frames": frames, "warm_frames": max(warm_frames, 4), "seconds": round(dt, 3),
                "fps": round(frames / dt, 2), "p50_ms": _pct(np, lat, 50), "p99_ms": _pct(np, lat, 99),
                "inputs": "resident in HBM" if resident else "pinned host memory (PCIe in the timed loop)",
                "max_frames_in_flight": in_flight, "max_bytes": max(r.get("size", 0) for r in res), "cap_bytes": cap,
                "mean_psnr_db": round(float(np.mean([r.get("psnr", 0) for r in res])), 2),
                "jobs": st["jobs"], "slots_per_job": st["slots_per_job"], "recoded": st["recoded"] - st0["recoded"],
                "failed_or_noncompliant": bad}

    def close(self):
        self.srv.close()
        for d in self.dev:
            d.close()


def pmc_probe(g, np, W, H, fourk, frames, dwt_tile=3):
    """For HBM traffic counters: one plain device-to-device copy of 1 GiB (the reference for
    scaling FETCH_SIZE / WRITE_SIZE), then one batch of the frames through the whole chain."""
    import torch
    a = torch.empty(1 << 30, dtype=torch.uint8, device="cuda:0")
    b = torch.empty_like(a)
    b.copy_(a)
    torch.cuda.synchronize()
    del a, b
    dev = g.Device(0)
    layout = g.FrameLayout(W, H, fourk)
    bt = g.Batch(dev, [layout] * len(frames), [g.xyz_settings(layout)] * len(frames), dwt_tile=dwt_tile, cm_waves=2,
                 mq_group=256, mq_read=64)
    for i, f in enumerate(frames):
        bt.upload(i, f)
        bt.set_slope(i, 1500.0)  # early termination on, as in the stream (0.25 lambda of a 24 fps frame)
    bt.encode(copy=False)
    bt.close()
    dev.close()
    print("[bench] pmc probe done", file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=48,
                    help="timed steps of the headline stream, each --step-frames frames")
    ap.add_argument("--warmup", type=int, default=2, help="untimed steps of the stream before the timed ones")
    ap.add_argument("--step-frames", type=int, default=32)
    ap.add_argument("--t1-batches", type=int, default=10, help="batches of the Tier-1 kernel measurement")
    ap.add_argument("--batch", type=int, default=16)
    ap.add_argument("--profile", default="4k_flat")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--plain-frames", type=int, default=8)
    ap.add_argument("--graph", type=int, default=0)
    ap.add_argument("--inflight", type=int, default=2)
    ap.add_argument("--stress-frames", type=int, default=512)
    ap.add_argument("--served-seconds", type=float, default=5.0)
    ap.add_argument("--mixed-steps", type=int, default=6)
    ap.add_argument("--pmc-probe", action="store_true", help="a 1 GiB copy and one batch, for counter runs")
    ap.add_argument("--pmc-tile", type=int, default=10, help="DWT tile shape of the counter probe's batch")
    ap.add_argument("--decode-frames", type=int, default=8, help="distinct 4K frames of the decoder measurement")
    ap.add_argument("--decoders", type=int, default=8)
    ap.add_argument("--decode-only", action="store_true", help="only the decoder measurement")
    ap.add_argument("--stream-inflight", type=int, default=256,
                    help="frames in flight in the headline stream (256: every slot of the 16 jobs)")
    ap.add_argument("--api-frames", type=int, default=96,
                    help="timed frames of the drop-in encoder API measurement (0: skip it)")
    ap.add_argument("--api-inflight", type=int, default=4, help="frames in flight in that measurement")
    ap.add_argument("--api-only", action="store_true",
                    help="only the drop-in encoder API, at --api-inflight, twice and four times as many in flight")
    ap.add_argument("--tune-only", action="store_true", help="only the autotuner's kernel times on one batch")
    ap.add_argument("--stream-only", action="store_true",
                    help="only the headline stream (pinned host memory) and its HBM-resident twin")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU; multi-GPU figures are not part of it "
              "(frames are independent end to end and shard across GPUs with no exchange)", file=sys.stderr)
        sys.exit(2)

    import numpy as np

    import libgail_amd as g

    W, H, fourk = g.CONTAINERS[args.profile]
    B = args.batch
    layout = g.FrameLayout(W, H, fourk)
    print(f"[bench] generating {B} frames {W}x{H}", file=sys.stderr, flush=True)
    frames = [g.synth_frame(1000 + i, W, H, threads=16) for i in range(B)]
    if args.pmc_probe:
        pmc_probe(g, np, W, H, fourk, frames, dwt_tile=args.pmc_tile)
        return
    if args.api_only:
        W4, H4, _ = g.CONTAINERS["4k_flat"]
        pinned = []
        for k in range(8):
            x = g.synth_frame(1000 + k, W4, H4, threads=16)
            pinned.append(_pin(g, np, [x[0], x[1], x[2]]))
        out = {}
        for n in (args.api_inflight, 2 * args.api_inflight, 4 * args.api_inflight):
            out[n] = api_encoder_run(g, np, pinned, n, args.api_frames)
            print(f"[bench] drop-in encoder API, {n} in flight: {json.dumps(out[n])}", file=sys.stderr, flush=True)
        print(json.dumps({"api_encoder": out}))
        return
    if args.stream_only:
        steps, warm, per = max(args.steps, 1), max(args.warmup, 0), max(args.step_frames, 1)
        sb = StreamBench(g, np, frames_in_flight=args.stream_inflight)
        try:
            out = {"inflight": args.stream_inflight,
                   "stream": sb.run(24.0, steps * per, args.stream_inflight, stream=3, resident=False,
                                    warm_frames=warm * per),
                   "stream_hbm_resident": sb.run(24.0, steps * per, args.stream_inflight, stream=1, resident=True,
                                                 warm_frames=warm * per)}
        finally:
            sb.close()
        print(json.dumps(out))
        return
    dev = g.Device(0)
    if args.decode_only:
        dec = decode_bench(g, np, dev, layout, frames[:max(1, args.decode_frames)], decoders=args.decoders,
                           cpu=not args.no_cpu)
        print(json.dumps({"decode": dec}))
        dev.close()
        return

    # Early-termination threshold of the kernel measurement: 0.25 x lambda of a preceding
    # frame (pipeline.cpp:194), from the rate control of the first frame coded without it.
    enc = g.Encoder(dev, layout, g.xyz_settings(layout))
    thr = 0.25 * g.rate_lambda(layout, enc.encode(frames[0]), fps=24.0, mbps=250.0)
    enc.close()
    print(f"[bench] early termination threshold {thr:.4g}", file=sys.stderr, flush=True)

    # ---- GPU stages alone: batches of B frames resident in HBM, `inflight` batches on their
    # own streams, Tier-1 results copied to host memory
    nb = max(1, args.inflight)
    tiles = {name: i for i, name in enumerate(g.Batch.DWT_SHAPES)}
    first = g.Batch(dev, [layout] * B, [g.xyz_settings(layout)] * B, use_graph=bool(args.graph))
    for i, f in enumerate(frames):
        first.upload(i, f)
        first.set_slope(i, thr)
    chosen, tune_ms = first.autotune(reps=2)
    print(f"[bench] autotuner: {chosen} {json.dumps(tune_ms)}", file=sys.stderr, flush=True)
    if args.tune_only:
        for _ in range(2):  # again, for the run-to-run spread
            print(json.dumps({"chosen": chosen, "ms": first.autotune(reps=2)[1]}))
        first.close()
        dev.close()
        return
    batches = [first] + [g.Batch(dev, [layout] * B, [g.xyz_settings(layout)] * B, use_graph=bool(args.graph),
                                 dwt_tile=tiles[chosen["dwt_tile"]], cm_waves=chosen["cm_waves"],
                                 mq_group=chosen["mq_group"], mq_read=chosen["mq_read"]) for _ in range(nb - 1)]
    for b in batches[1:]:
        for i, f in enumerate(frames):
            b.upload(i, f)
            b.set_slope(i, thr)
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
    while launched < min(nb, args.t1_batches):
        batches[launched % nb].launch()
        launched += 1
    while done < args.t1_batches:
        batches[done % nb].wait(copy=False)
        done += 1
        if launched < args.t1_batches:
            batches[launched % nb].launch()
            launched += 1
    dt = time.perf_counter() - t0
    ms_step = 1000 * dt / args.t1_batches
    t1_fps = B * args.t1_batches / dt
    print(f"[bench] Tier-1 batched: {t1_fps:.1f} frames/s, {ms_step:.2f} ms per batch of {B}", file=sys.stderr,
          flush=True)
    for b in batches:
        b.close()
    roof = None
    if med.get("dwt0", 0) > 0:
        achieved = B * dwt_level0_bytes(W, H) / (med["dwt0"] * 1e-3) / 1e9
        t16 = DWT0_TRAFFIC_4K_FLAT_16.get(chosen["dwt_tile"]) if args.profile == "4k_flat" else None
        traffic = t16 * B // 16 if t16 else None
        roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic, "kernel": dwt_level0_kernel(chosen["dwt_tile"]),
                "launch": f"{B} frames, {dwt_level0_bytes(W, H) * B} bytes algorithmic",
                "traffic_source": "rocprofv3 FETCH_SIZE / WRITE_SIZE, scaled by a 1 GiB copy"}
        copy = hbm_copy_gbs(g, dev)
        if copy:
            roof["copy_gbs"] = round(copy, 1)
            roof["frac_of_copy"] = round(achieved / copy, 4)
            print(f"[bench] 1 GiB device copy: {copy:.1f} GB/s; DWT level 0 {achieved:.1f} GB/s", file=sys.stderr,
                  flush=True)

    # ---- plain path: one frame at a time, upload, Tier-1, rate control and Tier-2 in turn
    enc = g.Encoder(dev, layout, g.xyz_settings(layout))
    lam = 0.0
    lat, lat_t1, plain_stages = [], [], []
    for i in range(max(args.plain_frames, 1) + 1):
        t1 = time.perf_counter()
        enc.set_slope(0.25 * lam)
        r = enc.encode(frames[i % B], copy=False)  # used before the next encode, as the plugin allows
        t2 = time.perf_counter()
        if i:
            plain_stages.append(r.timings)
        _, lam, _ = g.encode_codestream(layout, r, fps=24.0, mbps=250.0, tag="hip")
        if i:
            lat.append(1000 * (time.perf_counter() - t1))
            lat_t1.append(1000 * (t2 - t1))
    # one more plain frame with the debug taps, for the CPU-path checks
    dbg_thr = 0.25 * lam
    enc.set_slope(dbg_thr)
    plain_dbg = enc.encode(frames[0], debug=True)
    enc.close()
    plain_fps = 1000.0 / statistics.mean(lat)
    print(f"[bench] plain: {plain_fps:.1f} frames/s (Tier-1 {statistics.mean(lat_t1):.2f} ms of "
          f"{statistics.mean(lat):.2f} ms)", file=sys.stderr, flush=True)
    mixed = None
    if args.mixed_steps > 0:
        mixed = mixed_batch_run(g, np, dev, steps=args.mixed_steps)
        print(f"[bench] mixed 64-frame batch: {json.dumps(mixed)}", file=sys.stderr, flush=True)
    decode = None
    if args.decode_frames > 0:
        decode = decode_bench(g, np, dev, layout, frames[:args.decode_frames], decoders=args.decoders,
                              cpu=not args.no_cpu)
        print(f"[bench] decoder: {json.dumps(decode)}", file=sys.stderr, flush=True)
    dev.close()

    # ---- served 4K flat streams (the headline: frames handed over from host memory, as the
    # reference's pipeline uploads every frame) and the 48 fps stress stream
    steps, warm, per = max(args.steps, 1), max(args.warmup, 0), max(args.step_frames, 1)
    sb = StreamBench(g, np, frames_in_flight=args.stream_inflight)
    try:
        print(f"[bench] 4K flat stream from pinned host memory, 24 fps rate control, {warm} + {steps} steps of "
              f"{per} frames, {args.stream_inflight} in flight", file=sys.stderr, flush=True)
        main_run = sb.run(24.0, steps * per, args.stream_inflight, stream=3, resident=False, warm_frames=warm * per)
        print(f"[bench] stream: {json.dumps(main_run)}", file=sys.stderr, flush=True)
        hbm_run = sb.run(24.0, steps * per, args.stream_inflight, stream=1, resident=True, warm_frames=warm * per)
        print(f"[bench] stream from HBM-resident frames: {json.dumps(hbm_run)}", file=sys.stderr, flush=True)
        stress = None
        if args.stress_frames > 0:
            print(f"[bench] stress: 4K flat, 48 fps rate control, {args.stress_frames} frames, 256 in flight",
                  file=sys.stderr, flush=True)
            stress = sb.run(48.0, args.stress_frames, 256, stream=2, resident=True)
            print(f"[bench] stress: {json.dumps(stress)}", file=sys.stderr, flush=True)
    finally:
        sb.close()
    api_enc = None
    if args.api_frames > 0:
        api_enc = api_encoder_run(g, np, sb.src, args.api_inflight, args.api_frames)
        print(f"[bench] drop-in encoder API: {json.dumps(api_enc)}", file=sys.stderr, flush=True)
    served = None
    if args.served_seconds > 0:
        served = served_run(g, np, seconds=args.served_seconds)
        print(f"[bench] served: {json.dumps(served)}", file=sys.stderr, flush=True)

    cpu = cfg1 = parity = None
    if not args.no_cpu:
        cpu = cpu_baseline(frames[:4], fourk, thr)
        parity = plain_parity(np, plain_dbg, frames[0], fourk, dbg_thr, layout.levels)
        parity["frame"] = f"{args.profile} seed 1000, early termination threshold {dbg_thr:.4g}"
        print(f"[bench] plain-path checks: {json.dumps(parity)}", file=sys.stderr, flush=True)
        cfg1 = cpu_config1(g, np)
        print(f"[bench] CPU 2K frame: {json.dumps(cfg1)}", file=sys.stderr, flush=True)

    value = main_run["fps"]
    line = {
        "metric": "4K DCI encode fps, batched/served vs plain, 1 MI355X; p50/p99 frame latency",
        "value": value, "unit": "frames/s", "n_gpus": 1, "steps": steps, "warmup": warm,
        "ms_per_step": round(1000.0 * per / value, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32+int", "data": "synthetic",
        "config": {"workload": "served 4K flat 3996x2160 12-bit X'Y'Z' stream, 24 fps / 250 Mbit/s rate control, "
                               "codestreams built and DCI-checked",
                   "frames_per_step": per, "frames_in_flight": args.stream_inflight, "levels": layout.levels,
                   "code_blocks_per_frame": layout.num_cblks,
                   "input": "pinned host memory: every frame crosses PCIe inside the timed region",
                   "autotuned": chosen},
        "latency_ms": {"p50": main_run["p50_ms"], "p99": main_run["p99_ms"]},
        "stream": main_run,
        "stream_hbm_resident": hbm_run,
        "plain_fps": round(plain_fps, 2),
        "batched_over_plain": round(value / plain_fps, 3),
        "plain_latency_ms": {"p50": _pct(np, lat, 50), "p99": _pct(np, lat, 99),
                             "tier1_p50": _pct(np, lat_t1, 50)},
        "plain_stage_ms": {k: round(statistics.median(t[k] for t in plain_stages), 3)
                           for k in ("upload", "color", "dwt", "t1cm", "t1mq", "download", "gpu", "total")},
        "plain_checks": parity,
        "tier1_batched_fps": round(t1_fps, 2),
        "tier1_batch": {"frames": B, "batches_in_flight": nb, "ms_per_batch": round(ms_step, 3),
                        "slope_threshold": round(thr, 3)},
        "stage_ms_per_batch": {k: round(v, 3) for k, v in med.items()},
        "tier1_symbols_per_s": round(med["symbols"] / (ms_step * 1e-3), 1) if med.get("symbols") else None,
        "autotune_ms": tune_ms,
        "roofline": roof,
        "cpu_baseline": cpu,
        "batched_mixed_64": mixed,
        "stress_4k_48fps": stress,
        "served": served,
        "api_encoder": api_enc,
        "cpu_2k_frame": cfg1,
        "decode_4k": decode,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
