"""Benchmark of the raster path on one MI355X.

A step is one whole frame of the 4096x4096, 200k-shape (2.14M quadratic)
scene: every tile's threshold generation, per-column sort and coverage /
compositing sweep, from scene buffers resident in HBM to the BGRA frame in
HBM.  Host serialisation and binning run once, before the timed region, and
are reported separately.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

METRIC = "Gpixels/s + Mcurves/s, 4096² 2M-curve scene; batched & adjoint vs plain"
HBM_PEAK_GBS = 8000.0


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def algorithmic_bytes(scene, geo, jobs):
    """Bytes the path must move once per frame: every tile reads its shape
    records and their strand bytes, the tile records, the substance table,
    and writes the BGRA frame (texel reads are not counted)."""
    shapes = geo.shapes.view(np.uint32).reshape(-1, 4)
    starts = shapes[:, 2].astype(np.int64)
    order = np.argsort(starts)
    ends = np.empty_like(starts)
    ends[order[:-1]] = starts[order[1:]]
    ends[order[-1]] = geo.geo.size // 16
    shape_bytes = {int(s): int(16 * (e - s)) for s, e in zip(starts, ends)}
    refs = jobs.shape_refs.view(np.uint32).reshape(-1, 4)
    g = np.array([shape_bytes[int(s)] for s in refs[:, 2]], np.int64) if len(refs) else np.zeros(0, np.int64)
    per_ref = int((16 + g).sum())
    return per_ref + 32 * jobs.n_tiles + 16 * scene.substances.shape[0] + 4 * scene.width * scene.height


def cpu_baseline(scene, seconds_target=15.0):
    """The oracle (CPU restatement, all host threads) on a bounded sample of
    the same frame: whole raster jobs, taken in order until ~seconds_target."""
    from oracle import oracle as O
    spec = O.default_spec()
    og = O.Geometry(scene)
    oj = O.JobList(spec, og, scene.width, scene.height)
    threads = min(16, os.cpu_count() or 1)
    rnd = O.random_field(spec.random_field_size)
    out = np.zeros(scene.width * scene.height, np.uint32)
    pixels = 0
    jobs_done = 0
    t0 = time.perf_counter()
    for j in range(oj.n_jobs):
        pixels += O.render_jobs(scene, spec, og, oj, [j], out, nthreads=threads, rnd=rnd)
        jobs_done += 1
        if time.perf_counter() - t0 >= seconds_target:
            break
    dt = time.perf_counter() - t0
    return {"value": pixels / dt / 1e9, "unit": "Gpixels/s", "cores": threads, "kind": "port",
            "sample": f"{jobs_done} of {oj.n_jobs} raster jobs ({pixels} pixels) of the same "
                      f"4096x4096 frame, oracle C restatement, {dt:.1f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="only run a plain device copy of this many MiB (calibrates the "
                         "FETCH_SIZE / WRITE_SIZE counters) and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build measures one GPU; the raster path is not sharded across GPUs "
              "(run with --gpus 1)", file=sys.stderr)
        sys.exit(2)

    import torch
    import gudni_amd as G
    from gudni_amd import scenes as S

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    torch.cuda.set_device(0)
    if args.copy_mib:
        n = args.copy_mib * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.int32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(max(1, args.steps)):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": n * 4, "copies": max(1, args.steps)}), flush=True)
        return
    t = time.perf_counter()
    scene = S.config2()
    t_gen = time.perf_counter() - t
    t = time.perf_counter()
    spec, geo, jobs = G.prepare(scene)
    t_host = time.perf_counter() - t
    log(f"scene {scene.width}x{scene.height}: {scene.n_shapes} shapes, {scene.n_curves} curves, "
        f"{jobs.n_tiles} tiles, {jobs.n_refs} tile-shape refs; gen {t_gen:.2f}s host prep {t_host:.2f}s")

    r = G.Rasterizer(spec)
    stream = torch.cuda.current_stream()
    r.set_stream(stream.cuda_stream)
    t = time.perf_counter()
    r.upload_common(geo.geo, scene)
    r.frame_upload(jobs, scene)
    t_upload = time.perf_counter() - t
    npix = scene.width * scene.height
    out = torch.zeros(npix, dtype=torch.int32, device="cuda")

    for _ in range(args.warmup):
        r.frame_render(out.data_ptr())
    r.frame_check()
    torch.cuda.synchronize()

    start = torch.cuda.Event(enable_timing=True)
    stop = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    start.record(stream)
    for _ in range(args.steps):
        r.frame_render(out.data_ptr())
    stop.record(stream)
    torch.cuda.synchronize()
    ms_total = start.elapsed_time(stop)
    r.frame_check()
    ms = ms_total / args.steps

    # per-kernel timing: HIP events around each pass, on the launch stream
    r.enable_timing(True)
    fast, retry = [], []
    stages = np.zeros(4, np.float64)
    for _ in range(max(3, min(args.steps, 10))):
        r.frame_render(out.data_ptr())
        r.frame_check()
        tm = r.last_timings()
        fast.append(float(tm[0]))
        retry.append(float(tm[1]))
        stages += r.last_stage_cycles().astype(np.float64)
    r.enable_timing(False)
    tot = stages[:3].sum()
    stage_share = {k: round(float(v / tot), 4) for k, v in zip(("generate", "sort", "render"), stages[:3])} \
        if tot > 0 else None
    fast_ms = float(np.mean(fast))

    t = time.perf_counter()
    host_frame = out.cpu()
    t_read = time.perf_counter() - t
    del host_frame

    b_alg = algorithmic_bytes(scene, geo, jobs)
    achieved = b_alg / (fast_ms * 1e-3) / 1e9
    gpix = npix / (ms * 1e-3) / 1e9
    mcurves = scene.n_curves / (ms * 1e-3) / 1e6

    cpu = None
    if not args.no_cpu_baseline:
        log("cpu baseline (oracle on a bounded sample)")
        cpu = cpu_baseline(scene, args.cpu_seconds)

    line = {
        "metric": METRIC, "value": gpix, "unit": "Gpixels/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "fp32", "data": "synthetic",
        "mcurves_per_s": mcurves,
        "config": {"workload": "plain raster, 4096x4096 canvas, 200k shapes (180k blobs, 20k glyph "
                               "outlines), 2.14M quadratics, solid + nearest-texture substances",
                   "canvas": [scene.width, scene.height], "shapes": scene.n_shapes,
                   "curves": scene.n_curves, "tiles": jobs.n_tiles, "tile_shape_refs": jobs.n_refs,
                   "coverage_mode": "accurate"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "raster_pass (fast pass)", "kernel_ms": fast_ms,
                     "algorithmic_bytes": b_alg, "retry_pass_ms": float(np.mean(retry)),
                     "stage_cycle_share": stage_share},
        "cpu_baseline": cpu,
        "host_stages_s": {"serialize_and_bin": t_host, "upload": t_upload, "readback": t_read},
    }
    print(json.dumps(line), flush=True)
    r.close()


if __name__ == "__main__":
    main()
