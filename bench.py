#!/usr/bin/env python3
"""Throughput of the gfx950 Bayer demosaic path on one MI355X.

Default workload: one 4024x3032 RGGB frame, 12-bit packed (2 px / 3 B),
black level 64, MG ("Malvar-He-Cutler") demosaic -> u16 interleaved RGB,
fused in one kernel launch per frame.  A step is one such frame.  Inputs are
seeded synthetic scenes already resident in HBM; six input/output buffer sets
(549 MB, above the 256 MB Infinity Cache) are rotated so every step reads
and writes HBM.

Prints one JSON line: Mpix/s (whole-job), the HBM roofline of the kernel
(algorithmic bytes = packed input read once + RGB written once, divided by
the mean launch duration measured with HIP events on the launch stream), and
the reference's CPU path timed on this host (oracle/_ref, the reference
compiled from its own sources; the oracle restatement when that is absent).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME]
"""
from __future__ import annotations

import argparse
import hashlib
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "im-operator_amd"))

HBM_PEAK_GBS = 8000.0

WORKLOADS = {
    # name: (w, h, bd, packed, algo, black level, seed, diagonal lines)
    "mg12": (4024, 3032, 12, True, 2, 64.0, 2, 0),
    "prism10": (4024, 3032, 10, True, 8, 64.0, 3, 200),
    "hqli8": (1006, 758, 8, False, 1, None, 1, 0),
    "hqli12": (4024, 3032, 12, True, 1, 64.0, 2, 0),
}

METRIC = "Mpix/s + %HBM roofline, 4024x3032 RAW12 MHC demosaic; batched & fwd+bwd Mpix/s"


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def golden_sha(name):
    """SHA-256 of the reference's output for this workload, if recorded."""
    w, h, bd, packed, algo, bl, seed, lines = WORKLOADS[name]
    try:
        with open(os.path.join(ROOT, "tests", "golden", "golden.json")) as f:
            for c in json.load(f)["cases"]:
                if (c["kind"] == "scene" and c["w"] == w and c["h"] == h and c["algo"] == algo and c["bd"] == bd
                        and c["seed"] == seed):
                    return c["sha256"]
    except OSError:
        pass
    return None


# rocprofv3 summaries holding the measured HBM bytes per launch of each
# workload's kernel (FETCH_SIZE and WRITE_SIZE, each scaled by the factor a
# plain device copy of known size gives in the same run)
TRAFFIC_PROFILES = {
    "mg12": "profiles/k_fast_mg12_hbm_traffic.md",
}


def measured_traffic(name):
    """(bytes per launch, source file) from the committed rocprofv3 summary, or (None, None)."""
    rel = TRAFFIC_PROFILES.get(name)
    if rel is None:
        return None, None
    try:
        with open(os.path.join(ROOT, rel)) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch"):
                    return int(line.split("=")[1]), rel
    except OSError:
        pass
    return None, None


def cpu_baseline(raw, name, budget_s=12.0):
    """The reference CPU path on this host's cores: unpack (pixel::get_raw) ->
    black_level::process_global -> imop::demosaic_cpu, as one frame of work;
    repeated for about budget_s seconds, median frame time."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    w, h, bd, packed, algo, bl, seed, lines = WORKLOADS[name]
    blv = None if bl is None else [bl] * 4
    if O.reference_available():
        fn, kind, threads = O.ref_demosaic, "reference", O.reference().ref_num_threads()
    else:
        fn, kind, threads = O.demosaic, "port", O.oracle().orc_num_threads()
    times = []
    t_end = time.time() + budget_s
    while len(times) < 3 or (time.time() < t_end and len(times) < 200):
        t0 = time.perf_counter()
        e, _ = fn(raw, w, h, 0, algo, bd, packed, blv)
        times.append(time.perf_counter() - t0)
        if e != 0:
            raise RuntimeError(f"CPU baseline returned {e}")
    times.sort()
    med = times[len(times) // 2]
    return {
        "value": round(w * h / med / 1e6, 2),
        "unit": "Mpix/s",
        "cores": int(threads),
        "kind": kind,
        "sample": f"{len(times)} full {w}x{h} frames of the same workload, median frame {med * 1e3:.1f} ms "
                  f"(OpenMP rows, {threads} threads; unpack and black level single-threaded as in the reference)",
    }


def bench_batch(steps, warmup, with_cpu):
    """The mixed-size batch: 48 x 1006x758 + 16 x 4024x3032 MG frames with
    seeded bit depths {8u, 10p, 12p} and black levels {0, 16, 64}, u16 RGB.
    A step is the whole batch.  Times (a) the plain per-frame launches (one
    imop_hip_demosaic per frame, captured into a HIP graph so that host
    launch overhead is not counted; 8-bit frames write u8 there, so this
    baseline moves fewer bytes than the batch), (b) the direct batched call and (c) the graph-captured batch plan with the
    shape the autotuner picks, each over two rotating buffer sets."""
    import numpy as np
    import torch

    import imop_amd
    from imop_amd import scene

    MG = imop_amd.DemosaicAlgorithm.MG
    spec = scene.mixed_batch_spec(seed=4)
    dev = torch.device("cuda:0")
    log(f"bench: batch of {len(spec)} frames, generating scenes")
    large_scene = {}
    raws = []
    for f in spec:
        if f["w"] * f["h"] > 4_000_000:
            key = (f["w"], f["h"], f["bd"])
            if key not in large_scene:
                large_scene[key] = scene.raw_bytes(scene.make_scene(f["h"], f["w"], f["bd"], 0, seed=f["bd"]),
                                                   f["bd"], f["packed"])
            raws.append(large_scene[key])
        else:
            raws.append(scene.raw_bytes(scene.make_scene(f["h"], f["w"], f["bd"], 0, seed=f["seed"]), f["bd"],
                                        f["packed"]))
    nsets = 2
    d_in = [[torch.from_numpy(r.copy()).to(dev) for r in raws] for _ in range(nsets)]
    d_out = [[torch.empty(f["w"] * f["h"] * 6, dtype=torch.uint8, device=dev) for f in spec] for _ in range(nsets)]
    px = sum(f["w"] * f["h"] for f in spec)
    in_bytes = sum(int(r.size) for r in raws)
    alg_bytes = in_bytes + 6 * px
    stream = torch.cuda.current_stream()

    def per_frame(k):
        for f, di, do in zip(spec, d_in[k % nsets], d_out[k % nsets]):
            if imop_amd.demosaic(di, do, f["w"], f["h"], 0, MG, f["bd"], f["packed"], [f["bl"]] * 4) != 0:
                raise RuntimeError("imop_hip_demosaic failed")

    descs = [[imop_amd.frame_desc(di, do, f["w"], f["h"], 0, MG, f["bd"], f["packed"], [f["bl"]] * 4, False, 2)
              for f, di, do in zip(spec, d_in[s], d_out[s])] for s in range(nsets)]
    plans = [imop_amd.BatchPlan(descs[s]) for s in range(nsets)]
    for p in plans:
        if p.error != 0:
            raise RuntimeError(f"batch plan failed: {p.error}")

    # parity: every frame of the batch equals the single-frame path, by value
    per_frame(0)
    torch.cuda.synchronize()
    ref = [do[: f["w"] * f["h"] * (6 if f["bd"] > 8 else 3)].clone() for f, do in zip(spec, d_out[0])]
    plans[0].launch(stream)
    torch.cuda.synchronize()
    ok = True
    for f, do, r in zip(spec, d_out[0], ref):
        got = do.view(torch.int16).to(torch.int32) & 0xFFFF
        want = r.to(torch.int32) if f["bd"] <= 8 else (r.view(torch.int16).to(torch.int32) & 0xFFFF)
        ok = ok and bool(torch.equal(got, want))
    del ref
    parity = "every frame equals the single-frame path" if ok else "MISMATCH"
    log(f"bench: batch parity {parity}")

    # launch shape of the batched kernel: the project's autotuner over graph plans
    def digest():
        acc = 0
        for k, do in enumerate(d_out[0]):
            acc += int((do.view(torch.int32).to(torch.int64) * (2 * k + 1)).sum().item())
        return acc

    from imop_amd import autotune
    log("bench: autotuning the batched kernel")
    best = autotune.tune_batch(descs, digest=digest, verbose=False)
    shape = dict(rows=best[0], waves_per_block=best[1], blocks_per_cu=best[2], order=best[3])
    log(f"bench: batch shape {shape} ({best[4]:.4f} ms)")
    for p in plans:
        p.close()
    plans = [imop_amd.BatchPlan(descs[s]) for s in range(nsets)]

    graphs = []
    for s in range(nsets):
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            per_frame(s)
        graphs.append(g)

    def timed(fn):
        for k in range(warmup):
            fn(k)
        torch.cuda.synchronize()
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record(stream)
        for k in range(steps):
            fn(k)
        b.record(stream)
        torch.cuda.synchronize()
        return a.elapsed_time(b) / steps

    ms_frames_eager = timed(per_frame)
    ms_frames = timed(lambda k: graphs[k % nsets].replay())
    ms_direct = timed(lambda k: imop_amd.demosaic_batch(descs[k % nsets], stream))
    ms_plan = timed(lambda k: plans[k % nsets].launch(stream))
    for p in plans:
        p.close()
    mpx = lambda ms: round(px / (ms / 1e3) / 1e6, 1)
    log(f"bench: batch per-frame {ms_frames:.4f} ms (eager loop {ms_frames_eager:.4f}), direct {ms_direct:.4f} ms, "
        f"plan {ms_plan:.4f} ms")
    out = {
        "workload": "MG demosaic batch 48 x 1006x758 + 16 x 4024x3032, depths {8u,10p,12p}, black level {0,16,64}, "
                    "u16 RGB (seeded)",
        "frames": len(spec),
        "pixels_per_step": px,
        "value": mpx(ms_plan),
        "unit": "Mpix/s",
        "ms_per_step": round(ms_plan, 5),
        "per_frame_launches_mpix_s": mpx(ms_frames),
        "per_frame_launches_eager_mpix_s": mpx(ms_frames_eager),
        "direct_batch_call_mpix_s": mpx(ms_direct),
        "speedup_vs_per_frame": round(ms_frames / ms_plan, 3),
        "batch_shape": shape,
        "roofline": {"bound": "hbm", "achieved": round(alg_bytes / (ms_plan / 1e3) / 1e9, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(alg_bytes / (ms_plan / 1e3) / 1e9 / HBM_PEAK_GBS, 4),
                     "algorithmic_bytes_per_step": alg_bytes},
        "parity": parity,
        "steps": steps,
        "warmup": warmup,
    }
    if with_cpu:
        out["cpu_baseline"] = cpu_baseline_batch(spec, raws)
    return out


def cpu_baseline_batch(spec, raws, budget_s=10.0):
    """The reference CPU path over the batch's frames in order, for about
    budget_s seconds: pixels done / time."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    if O.reference_available():
        fn, kind, threads = O.ref_demosaic, "reference", O.reference().ref_num_threads()
    else:
        fn, kind, threads = O.demosaic, "port", O.oracle().orc_num_threads()
    done_px, n = 0, 0
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < budget_s and n < 4 * len(spec):
        f, raw = spec[n % len(spec)], raws[n % len(spec)]
        e, _ = fn(raw, f["w"], f["h"], 0, 2, f["bd"], f["packed"], [f["bl"]] * 4)
        if e != 0:
            raise RuntimeError(f"CPU baseline returned {e}")
        done_px += f["w"] * f["h"]
        n += 1
    dt = time.perf_counter() - t0
    return {"value": round(done_px / dt / 1e6, 2), "unit": "Mpix/s", "cores": int(threads), "kind": kind,
            "sample": f"the first {n} frames of the batch in order ({done_px} px, {dt:.1f} s), one frame at a time"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="mg12", choices=sorted(WORKLOADS))
    ap.add_argument("--sets", type=int, default=6)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-batch", action="store_true", help="skip the mixed-size batch measurement")
    ap.add_argument("--batch-steps", type=int, default=20)
    ap.add_argument("--copy-mb", type=int, default=0,
                    help="after the timed loop, run 4 plain device copies of this many MiB "
                         "(calibrates rocprofv3 FETCH_SIZE / WRITE_SIZE against known bytes)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU only; --gpus {args.gpus} is not supported "
              "(frames are independent: N GPUs would each run this same per-frame kernel).", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    import imop_amd
    from imop_amd import scene

    w, h, bd, packed, algo, bl, seed, lines = WORKLOADS[args.workload]
    blv = None if bl is None else [bl] * 4
    log(f"bench: workload {args.workload} {w}x{h} bd{bd} packed={packed} algo={algo} bl={bl}")
    vals = scene.make_scene(h, w, bd, 0, seed=seed, n_lines=lines)
    raw = scene.raw_bytes(vals, bd, packed)
    in_b = raw.size
    out_b = imop_amd.rgb_byte_size(w, h, bd)
    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    d_in = [torch.from_numpy(raw.copy()).to(dev) for _ in range(args.sets)]
    d_out = [torch.empty(out_b, dtype=torch.uint8, device=dev) for _ in range(args.sets)]
    stream = torch.cuda.current_stream()

    def launch(k):
        e = imop_amd.demosaic(d_in[k % args.sets], d_out[k % args.sets], w, h, 0, algo, bd, packed, blv,
                              stream=stream)
        if e != 0:
            raise RuntimeError(f"imop_hip_demosaic returned {e}")

    # correctness of the measured configuration against the golden digest
    launch(0)
    torch.cuda.synchronize()
    want = golden_sha(args.workload)
    got = hashlib.sha256(d_out[0].cpu().numpy().tobytes()).hexdigest()
    parity = "unchecked" if want is None else ("bit-exact vs reference digest" if got == want else "MISMATCH")
    log(f"bench: parity {parity}")

    for k in range(args.warmup):
        launch(k)
    torch.cuda.synchronize()
    log(f"bench: warmup {args.warmup} done")

    # Timed region: K back-to-back launches on the launch stream, bracketed by
    # two HIP events on that stream (no events between steps: each step is
    # exactly one kernel launch, so the mean launch duration is total / K).
    start = torch.cuda.Event(enable_timing=True)
    end = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    start.record(stream)
    for k in range(args.steps):
        launch(k)
    end.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    total_ms = start.elapsed_time(end)
    mean_launch_ms = total_ms / args.steps
    ms_per_step = total_ms / args.steps
    px = w * h
    value = px * args.steps / (total_ms / 1e3) / 1e6
    alg_bytes = in_b + out_b
    achieved = alg_bytes / (mean_launch_ms / 1e3) / 1e9
    log(f"bench: {args.steps} steps in {total_ms:.3f} ms (wall {wall * 1e3:.1f} ms); mean launch {mean_launch_ms * 1e3:.2f} us")
    traffic, traffic_src = measured_traffic(args.workload)

    result = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "Mpix/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic",
        "config": {
            "workload": {"mg12": "MG demosaic 4024x3032 RGGB RAW12-packed, black level 64 -> u16 RGB",
                         "prism10": "PRISM demosaic 4024x3032 RGGB RAW10-packed, black level 64 -> u16 RGB",
                         "hqli8": "HQLI (bilinear) 1006x758 RGGB u8 -> u8 RGB",
                         "hqli12": "HQLI (bilinear) 4024x3032 RGGB RAW12-packed, black level 64 -> u16 RGB"}[args.workload],
            "frames_per_step": 1,
            "width": w, "height": h, "bit_depth": bd, "packed": packed, "black_level": bl,
            "buffer_sets": args.sets,
            "parity": parity,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "algorithmic_bytes_per_launch": alg_bytes,
            "mean_launch_us": round(mean_launch_ms * 1e3, 3),
            "traffic_source": traffic_src,
        },
    }
    if args.copy_mb > 0:
        n = args.copy_mb << 20
        src = torch.empty(n, dtype=torch.uint8, device=dev).fill_(1)
        dst = torch.empty(n, dtype=torch.uint8, device=dev)
        for _ in range(4):
            dst.copy_(src)
        torch.cuda.synchronize()
        log(f"bench: {4} calibration copies of {n} bytes done")
    if not args.no_cpu_baseline:
        log("bench: timing the CPU baseline")
        result["cpu_baseline"] = cpu_baseline(raw, args.workload)
    if not args.no_batch and args.workload == "mg12":
        del d_in, d_out
        torch.cuda.empty_cache()
        result["batched"] = bench_batch(args.batch_steps, 3, not args.no_cpu_baseline)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
