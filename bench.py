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
    the gradient records, the picture bytes once (every texel a picture use
    can reach, read once), and writes the BGRA frame."""
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
    n_grads = np.asarray(scene.gradients, np.float32).reshape(-1, 28).shape[0]
    pict = int(np.asarray(scene.pict_data).size) if len(scene.pict_uses) else 0
    return (per_ref + 32 * jobs.n_tiles + 16 * scene.substances.shape[0] + 112 * n_grads + pict
            + 4 * scene.width * scene.height)


def cpu_info():
    """CPU model, logical CPUs of the machine and the CPUs this process may use
    (its affinity, capped by OMP_NUM_THREADS where the host sets one: the
    share of the host that goes with one GPU)."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as fh:
            for ln in fh:
                if ln.startswith("model name"):
                    model = ln.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    usable = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    share = usable
    if os.environ.get("OMP_NUM_THREADS", "").isdigit():
        share = max(1, min(usable, int(os.environ["OMP_NUM_THREADS"])))
    return {"model": model, "machine_logical_cpus": os.cpu_count(), "affinity_cpus": usable,
            "threads_used": share}


def cpu_baseline(scene, seconds_target=15.0, texture_filter=1):
    """The oracle (CPU restatement) on the host CPUs this process is given, on
    a bounded sample of the same frame: whole raster jobs, taken in order
    until ~seconds_target (the whole 4096^2 frame fits in that on 16 cores)."""
    from oracle import oracle as O
    spec = O.default_spec()
    spec.coverage_mode = 1   # the reference's own arithmetic, as its CPU path would run
    spec.texture_filter = texture_filter
    og = O.Geometry(scene)
    oj = O.JobList(spec, og, scene.width, scene.height)
    info = cpu_info()
    threads = info["threads_used"]
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
                      f"4096x4096 frame in REFERENCE coverage mode, oracle C restatement in the "
                      f"reference's arithmetic, {dt:.1f} s on {threads} threads",
            "whole_frame": jobs_done == oj.n_jobs, "cpu": info}


def measured_traffic(kernel_ms):
    """HBM bytes per fast-pass launch from the committed rocprofv3 counter
    summary (FETCH_SIZE and WRITE_SIZE per dispatch, each scaled by what the
    counter reads for a plain copy of known size), with the source
    fingerprint and commit of the library that was profiled and whether it
    is the library this run uses."""
    from gudni_amd import build as pbuild
    path = os.path.join(ROOT, "profiles", "raster_pass_blobs_4096_traffic.json")
    if not os.path.exists(path):
        return None, None
    with open(path) as fh:
        t = json.load(fh)
    return float(t["traffic_bytes"]), {
        "source": "profiles/raster_pass_blobs_4096_counters.md",
        "fetch_bytes": t["fetch_kib"] * 1024 * t["fetch_factor"],
        "write_bytes": t["write_kib"] * 1024 * t["write_factor"],
        "profiled_kernel_ms": t["fast_pass_avg_ms"], "this_run_kernel_ms": kernel_ms,
        "lds_bank_conflict_rate": t["lds_bank_conflict_rate"],
        "profiled_commit": t.get("commit"), "profiled_fingerprint": t.get("fingerprint"),
        "profiled_library_is_this_run": t.get("fingerprint") == pbuild.fingerprint(),
        "valu_utilization": t.get("valu_utilization")}


def device_binning_leg(r, scene, geo, stream, steps):
    """Stage (a) on the device for the same frame: gudni_scene_bin (tile tree,
    references, jobs, work list) timed with HIP events, then bin + render."""
    import torch
    r.scene_upload(geo, scene)
    r.enable_timing(True)
    bin_ms = []
    for _ in range(max(3, steps)):
        r.scene_bin()
        bin_ms.append(r.last_bin_ms())
    r.enable_timing(False)
    out = torch.zeros(scene.width * scene.height, dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()
    t = time.perf_counter()
    for _ in range(steps):
        r.scene_bin()
        r.frame_render(out.data_ptr())
    r.frame_check()
    torch.cuda.synchronize()
    both = (time.perf_counter() - t) / steps * 1e3
    nt, nr, nw = r.scene_counts()
    return {"bin_ms": float(np.median(bin_ms)), "bin_plus_render_ms": both,
            "bin_plus_render_gpixels_per_s": scene.width * scene.height / (both * 1e-3) / 1e9,
            "tiles": nt, "refs": nr, "work_units": nw}


def copy_bandwidth(mib=2048, reps=10):
    """Achievable HBM bandwidth on this device: a plain device-to-device copy
    of `mib` MiB timed with events; bytes read + written per second."""
    import torch
    n = mib * (1 << 20) // 4
    src = torch.ones(n, dtype=torch.int32, device="cuda")
    dst = torch.empty_like(src)
    dst.copy_(src)
    stream = torch.cuda.current_stream()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(reps):
        dst.copy_(src)
    e1.record(stream)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    del src, dst
    return 2 * n * 4 / (ms * 1e-3) / 1e9


def device_pipeline_leg(r, scene, spec, stream, steps):
    """The whole host side of a frame on the device too: outlines in host
    memory -> upload -> device serialiser -> device binning -> render, against
    the host serialiser + host tile tree + upload + render.  Wall clock from
    outlines in host memory to the frame in HBM."""
    import torch
    import gudni_amd as G
    out = torch.zeros(scene.width * scene.height, dtype=torch.int32, device="cuda")
    r.enable_timing(True)
    ser_ms, bin_ms = [], []
    for _ in range(3):
        r.scene_upload_outlines(scene)
        ser_ms.append(r.last_serialize_ms())
        r.scene_bin()
        bin_ms.append(r.last_bin_ms())
    r.enable_timing(False)
    torch.cuda.synchronize()
    t = time.perf_counter()
    for _ in range(steps):
        r.scene_upload_outlines(scene)
        r.scene_bin()
        r.frame_render(out.data_ptr())
    r.frame_check()
    torch.cuda.synchronize()
    dev = (time.perf_counter() - t) / steps * 1e3
    t = time.perf_counter()
    for _ in range(steps):
        geo = G.Geometry(scene)
        jobs = G.RasterJobs(spec, geo, scene.width, scene.height)
        r.upload_common(geo.geo, scene)
        r.frame_upload(jobs, scene)
        r.frame_render(out.data_ptr())
    r.frame_check()
    torch.cuda.synchronize()
    host = (time.perf_counter() - t) / steps * 1e3
    return {"serialize_ms_device": float(np.median(ser_ms)), "bin_ms_device": float(np.median(bin_ms)),
            "outlines_to_frame_ms": dev, "outlines_to_frame_ms_host_stages": host,
            "speedup": host / dev,
            "outline_bytes": int(scene.pairs.nbytes + scene.tags.nbytes + scene.shape_outline_offsets.nbytes
                                 + scene.outline_pair_offsets.nbytes)}


def render_f32(r, scene, spec):
    """The scene's pre-quantisation RGBA image from the raster path (H, W, 4)."""
    import torch
    import gudni_amd as G
    _, geo, jobs = G.prepare(scene, spec)
    r.upload_common(geo.geo, scene)
    r.frame_upload(jobs, scene)
    out = torch.zeros(scene.width * scene.height, dtype=torch.int32, device="cuda")
    f32 = torch.zeros(scene.width * scene.height * 4, dtype=torch.float32, device="cuda")
    r.frame_render(out.data_ptr(), f32.data_ptr())
    r.frame_check()
    return f32.cpu().numpy().reshape(scene.height, scene.width, 4)


# What the adjoint entry points take (gudni_adjoint_upload): the workloads
# timed below lie inside it.
ADJOINT_INPUTS = ("even-odd shapes with solid, gradient and picture substances, subtract compounds "
                  "and shared substances; nonzero-fill shapes are refused; colour gradients for "
                  "solid substances only")


def adjoint_bytes(scene, spec):
    """Algorithmic bytes of one adjoint step: three passes over the tiles'
    shape records and strands (forward, backward, point pass: recomputation,
    nothing stored), the target read, the RGBA image written, the per-pixel
    dL/d(colour) written and read back, and the gradients written (16 B per
    outline pair, 16 B per substance)."""
    import gudni_amd as G
    _, geo, jobs = G.prepare(scene, spec)
    px = scene.width * scene.height
    return (3 * (algorithmic_bytes(scene, geo, jobs) - 4 * px) + 16 * px + 16 * px + 32 * px
            + 16 * (scene.pairs.size // 4) + 16 * len(scene.substances))


def _adjoint_spec(spec, tile, radix):
    import gudni_amd as G
    return G.default_spec(texture_filter=spec.texture_filter, max_tile_size=tile, sort_radix_bits=radix)


def _time_adjoint(spec, items_fn, runs=2):
    """ms per adjoint step of the items items_fn(rasterizer, spec) builds."""
    import torch
    import gudni_amd as G
    r = G.Rasterizer(spec)
    stream = torch.cuda.current_stream()
    r.set_stream(stream.cuda_stream)
    adj = G.Adjoint(r, items_fn(r, spec))
    try:
        adj.run()
        adj.result(0)
        start = torch.cuda.Event(enable_timing=True)
        stop = torch.cuda.Event(enable_timing=True)
        start.record(stream)
        for _ in range(runs):
            adj.run()
        stop.record(stream)
        adj.result(0)
        return start.elapsed_time(stop) / runs
    finally:
        adj.close()
        r.close()


def adjoint_leg(spec, steps, log_fn, fixed=None):
    """The differentiable workload: a 1024^2, 2,000-shape target and a
    perturbed copy as parameters; one step = forward + backward (loss,
    dL/d points, dL/d RGBA) in one launch.  Tile size and sort radix come
    from the autotuner (timed adjoint steps).  Device time from events around
    the launches; the full step adds re-serialising and re-binning the updated
    parameters on the host, the upload and reading the gradients back."""
    import torch
    import gudni_amd as G
    from gudni_amd import autotune, scenes as S
    base = S.adjoint_1024()
    params = S.perturb(base, S.SEEDS["adjoint_1024"] + 1)

    def items(r, sp):
        return [(params, render_f32(r, base, sp))]

    if fixed:
        (tile, radix), timings = fixed, None
    else:
        tile, radix, timings = autotune.search(lambda ts, rb: _time_adjoint(_adjoint_spec(spec, ts, rb), items),
                                               tiles=autotune.ADJOINT_CANDIDATES)
        log_fn(f"adjoint autotune {json.dumps(timings)} -> tile {tile}, radix {radix}")
    sp = _adjoint_spec(spec, tile, radix)
    r = G.Rasterizer(sp)
    stream = torch.cuda.current_stream()
    r.set_stream(stream.cuda_stream)
    adj = G.Adjoint(r, items(r, sp))
    adj.run()
    loss0, dp, dc, _ = adj.result(0)
    start = torch.cuda.Event(enable_timing=True)
    stop = torch.cuda.Event(enable_timing=True)
    start.record(stream)
    for _ in range(steps):
        adj.run()
    stop.record(stream)
    adj.result(0)
    dev_ms = start.elapsed_time(stop) / steps
    pairs = params.pairs.copy()

    def host_step(pairs, dp):
        pairs = pairs - 0.05 * dp.reshape(pairs.shape)
        adj.update(0, pairs=pairs)
        adj.run()
        loss, dp, dc, _ = adj.result(0)
        return pairs, dp, loss

    pairs, dp, loss = host_step(pairs, dp)   # warm-up: the first update builds the device tables
    t = time.perf_counter()
    for _ in range(steps):
        pairs, dp, loss = host_step(pairs, dp)
    full_ms = (time.perf_counter() - t) / steps * 1e3
    # the same loop with parameters and gradients kept on the device
    pairs_t = torch.tensor(pairs.reshape(-1), dtype=torch.float32, device="cuda")
    dp_t = torch.zeros_like(pairs_t)

    def device_step():
        adj.run()
        adj.result_device(0, d_pairs_ptr=dp_t.data_ptr())
        pairs_t.sub_(0.05 * dp_t)
        adj.update_device(0, pairs_ptr=pairs_t.data_ptr())

    device_step()   # warm-up: first use of the torch kernels
    torch.cuda.synchronize()
    t = time.perf_counter()
    for _ in range(steps):
        device_step()
    adj.check()
    torch.cuda.synchronize()
    dev_loop_ms = (time.perf_counter() - t) / steps * 1e3
    adj.close()
    r.close()
    b_adj = adjoint_bytes(params, sp)
    log_fn(f"adjoint: {dev_ms:.2f} ms per step on the device, {full_ms:.2f} ms with updates from the host, "
           f"{dev_loop_ms:.2f} ms with updates on the device; loss {loss0:.2f} -> {loss:.2f}")
    return {"workload": "1024x1024, 2000 shapes (17k quadratics), solid; params perturbed "
                        "(points + N(0,4px), colours + N(0,0.1))",
            "tile_size": tile, "sort_radix_bits": radix, "autotune_ms": timings,
            "ms_per_step_device": dev_ms, "steps_per_s_device": 1e3 / dev_ms,
            "algorithmic_bytes": b_adj, "achieved_gbs": b_adj / (dev_ms * 1e-3) / 1e9,
            "ms_per_step_with_host_update": full_ms, "steps_per_s_with_host_update": 1e3 / full_ms,
            "ms_per_step_with_device_update": dev_loop_ms,
            "steps_per_s_with_device_update": 1e3 / dev_loop_ms,
            "loss_first": loss0, "loss_after_steps": loss, "n_points": int(params.pairs.size // 2),
            "n_substances": int(len(params.substances)), "inputs_accepted": ADJOINT_INPUTS}


def batched_adjoint_leg(spec, steps, log_fn, fixed=None):
    """64 scenes of 512^2 (500-5,000 shapes, solid + a 1024^2 texture): forward
    and backward of all scenes in one persistent launch, against 64
    per-scene adjoint launches.  Tile size and radix are tuned on the whole
    batch (a sample of a few scenes picked a tile size that was slower for
    all 64)."""
    import torch
    import gudni_amd as G
    from gudni_amd import autotune, scenes as S
    bases = S.adjoint_batch_512()
    params = [S.perturb(b, S.SEEDS["adjoint_batch_512"] + i) for i, b in enumerate(bases)]

    def items_of(k):
        return lambda r, sp: [(params[i], render_f32(r, bases[i], sp)) for i in range(k)]

    if fixed:
        (tile, radix), timings = fixed, None
    else:
        tile, radix, timings = autotune.search(
            lambda ts, rb: _time_adjoint(_adjoint_spec(spec, ts, rb), items_of(len(bases)), runs=1),
            tiles=autotune.ADJOINT_CANDIDATES)
        log_fn(f"batched adjoint autotune {json.dumps(timings)} -> tile {tile}, radix {radix}")
    sp = _adjoint_spec(spec, tile, radix)
    r = G.Rasterizer(sp)
    stream = torch.cuda.current_stream()
    r.set_stream(stream.cuda_stream)
    items = items_of(len(bases))(r, sp)
    adj = G.Adjoint(r, items)
    adj.run()
    adj.result(0)
    start = torch.cuda.Event(enable_timing=True)
    stop = torch.cuda.Event(enable_timing=True)
    start.record(stream)
    for _ in range(steps):
        adj.run()
    stop.record(stream)
    adj.result(0)
    one_ms = start.elapsed_time(stop) / steps
    start.record(stream)
    for _ in range(steps):
        for i in range(len(items)):
            adj.run_scene(i)
    stop.record(stream)
    adj.result(0)
    seq_ms = start.elapsed_time(stop) / steps
    # training steps with every scene's parameters updated on the device
    pairs_t = torch.tensor(np.concatenate([sc.pairs.reshape(-1) for sc, _ in items]), dtype=torch.float32,
                           device="cuda")
    dp_t = torch.zeros_like(pairs_t)

    def device_step():
        adj.run()
        adj.result_device(-1, d_pairs_ptr=dp_t.data_ptr())
        pairs_t.sub_(0.05 * dp_t)
        adj.update_device(-1, pairs_ptr=pairs_t.data_ptr())

    device_step()   # warm-up: the first update rebuilds every scene on the device
    torch.cuda.synchronize()
    t = time.perf_counter()
    for _ in range(steps):
        device_step()
    adj.check()
    torch.cuda.synchronize()
    loop_ms = (time.perf_counter() - t) / steps * 1e3
    adj.close()
    r.close()
    b_adj = sum(adjoint_bytes(sc, sp) for sc, _ in items)
    log_fn(f"batched adjoint: {one_ms:.2f} ms per step in one launch, {seq_ms:.2f} ms as "
           f"{len(items)} launches, {loop_ms:.2f} ms per training step with device updates")
    return {"workload": f"{len(items)} scenes of 512x512, 500-5000 shapes, solid + 1024^2 texture",
            "tile_size": tile, "sort_radix_bits": radix, "autotune_ms": timings,
            "ms_per_step": one_ms, "steps_per_s": 1e3 / one_ms,
            "algorithmic_bytes": b_adj, "achieved_gbs": b_adj / (one_ms * 1e-3) / 1e9,
            "sequential_ms_per_step": seq_ms, "sequential_steps_per_s": 1e3 / seq_ms,
            "ms_per_step_with_device_update": loop_ms, "steps_per_s_with_device_update": 1e3 / loop_ms,
            "speedup_vs_sequential": seq_ms / one_ms,
            "shapes": int(sum(s.n_shapes for s, _ in items)), "inputs_accepted": ADJOINT_INPUTS}


def batched_leg(spec, steps, log_fn, fixed=None):
    """The batched workload: 256 scenes of 256^2..2048^2 with 100..50k shapes,
    tile size and sort radix chosen by the autotuner on the whole batch, all
    scenes in one persistent launch replayed from a captured hipGraph;
    compared with the same scenes rendered by one launch each."""
    import torch
    import gudni_amd as G
    from gudni_amd import autotune, scenes as S
    t = time.perf_counter()
    scenes = S.batch_256()
    t_gen = time.perf_counter() - t
    r = G.Rasterizer(spec)
    stream = torch.cuda.current_stream()
    r.set_stream(stream.cuda_stream)
    t = time.perf_counter()
    if fixed:
        (tile, radix), timings = fixed, None
    else:
        tile, radix, timings = autotune.tune_batch(r, scenes, reps=2, log=log_fn)
    t_tune = time.perf_counter() - t
    items, chosen, radices, t_prep = autotune.prepare_batch(scenes, default_tile=tile, default_radix=radix)
    b_alg = sum(algorithmic_bytes(sc, geo, jobs) for sc, geo, jobs in items)
    t = time.perf_counter()
    b = G.Batch(r, items)
    t_up = time.perf_counter() - t
    pixels = sum(s.width * s.height for s in scenes)
    curves = sum(s.n_curves for s in scenes)
    b.render(use_graph=True)
    b.check()
    start = torch.cuda.Event(enable_timing=True)
    stop = torch.cuda.Event(enable_timing=True)
    start.record(stream)
    for _ in range(steps):
        b.render(use_graph=True)
    stop.record(stream)
    lat = b.check()
    ms = start.elapsed_time(stop) / steps
    # sequential: one launch per scene
    evs = [torch.cuda.Event(enable_timing=True) for _ in range(len(scenes) + 1)]
    evs[0].record(stream)
    for i in range(len(scenes)):
        b.render_scene(i)
        evs[i + 1].record(stream)
    b.check()
    seq = np.array([evs[i].elapsed_time(evs[i + 1]) for i in range(len(scenes))])
    seq_lat = np.cumsum(seq)
    sample = [0, len(scenes) // 2, len(scenes) - 1]
    want = {i: b.frame(i) for i in sample}
    b.close()
    # the same batch serialised and binned on the device from the outlines
    torch.cuda.synchronize()
    t = time.perf_counter()
    bd = G.Batch.from_outlines(r, scenes, chosen, radices)
    t_dev = time.perf_counter() - t
    bd.render(use_graph=True)
    bd.check()
    same = all(np.array_equal(bd.frame(i), want[i]) for i in sample)
    start.record(stream)
    for _ in range(steps):
        bd.render(use_graph=True)
    stop.record(stream)
    bd.check()
    ms_dev = start.elapsed_time(stop) / steps
    bd.close()
    r.close()
    return {
        "scenes": len(scenes), "pixels": pixels, "curves": curves,
        "ms_per_batch": ms, "gpixels_per_s": pixels / (ms * 1e-3) / 1e9,
        "mcurves_per_s": curves / (ms * 1e-3) / 1e6,
        "algorithmic_bytes": b_alg, "achieved_gbs": b_alg / (ms * 1e-3) / 1e9,
        "latency_p50_ms": float(np.percentile(lat, 50)), "latency_p99_ms": float(np.percentile(lat, 99)),
        "sequential_ms": float(seq.sum()),
        "sequential_gpixels_per_s": pixels / (seq.sum() * 1e-3) / 1e9,
        "sequential_latency_p50_ms": float(np.percentile(seq_lat, 50)),
        "sequential_latency_p99_ms": float(np.percentile(seq_lat, 99)),
        "speedup_vs_sequential": float(seq.sum() / ms),
        "tile_size": tile, "sort_radix_bits": radix, "autotune_ms": timings, "graph": True,
        "host_s": {"generate": t_gen, "autotune": t_tune, "serialize_and_bin": t_prep, "upload": t_up},
        "device_built": {"build_s": t_dev, "ms_per_batch": ms_dev, "frames_equal_host_built": same,
                         "sampled_scenes": sample},
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-batch", action="store_true", help="skip the batched workload")
    ap.add_argument("--batch-steps", type=int, default=3)
    ap.add_argument("--no-adjoint", action="store_true", help="skip the differentiable workloads")
    ap.add_argument("--adjoint-only", action="store_true",
                    help="run only the two differentiable workloads and print their results")
    ap.add_argument("--adjoint-tile", type=int, default=0,
                    help="tile size of the differentiable workloads (default: autotuned)")
    ap.add_argument("--adjoint-radix", type=int, default=4)
    ap.add_argument("--sort-radix", type=int, default=-1,
                    help="sort radix bits of the frame (default: chosen by the autotuner)")
    ap.add_argument("--tile", type=int, default=0,
                    help="tile size of the frame (default: chosen by the autotuner)")
    ap.add_argument("--batch-tile", type=int, default=0,
                    help="tile size of every batched scene (with --batch-radix: no batch autotuning)")
    ap.add_argument("--batch-radix", type=int, default=-1)
    ap.add_argument("--texture-filter", type=int, default=1,
                    help="picture sampling of the benchmark scenes: 1 bilinear, 0 nearest")
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
    from gudni_amd import autotune, scenes as S

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
    if args.adjoint_only:
        fixed = (args.adjoint_tile, args.adjoint_radix) if args.adjoint_tile else None
        spec = G.default_spec(texture_filter=args.texture_filter)
        res = {"adjoint": adjoint_leg(spec, max(3, args.batch_steps), log, fixed),
               "batched_adjoint": batched_adjoint_leg(spec, max(2, args.batch_steps), log, fixed)}
        print(json.dumps(res), flush=True)
        return
    t = time.perf_counter()
    scene = S.blobs_4096()
    t_gen = time.perf_counter() - t
    t = time.perf_counter()
    spec, geo, jobs = G.prepare(scene, G.default_spec(texture_filter=args.texture_filter))
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

    # tile size and sort radix for this frame
    def frame_ms(tile, radix):
        j = G.RasterJobs(G.default_spec(texture_filter=args.texture_filter, sort_radix_bits=radix,
                                        max_tile_size=tile), geo, scene.width, scene.height)
        r.frame_upload(j, scene)
        r.frame_render(out.data_ptr())
        r.frame_check()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(3):
            r.frame_render(out.data_ptr())
        e1.record(stream)
        r.frame_check()
        return e0.elapsed_time(e1) / 3

    if args.sort_radix >= 0 and args.tile > 0:
        tile, radix, tune_ms = args.tile, args.sort_radix, None
    else:
        tile, radix, tune_ms = autotune.tune_frame(
            frame_ms, tiles=(args.tile,) if args.tile > 0 else autotune.CANDIDATES,
            radices=(args.sort_radix,) if args.sort_radix >= 0 else (0, 3, 4, 5),
            default_radix=max(args.sort_radix, 0))
        log(f"frame tile size and sort radix: {json.dumps(tune_ms)} -> tile {tile}, radix {radix}")
    spec = G.default_spec(texture_filter=args.texture_filter, sort_radix_bits=radix, max_tile_size=tile)
    jobs = G.RasterJobs(spec, geo, scene.width, scene.height)
    r.close()
    r = G.Rasterizer(spec)
    r.set_stream(stream.cuda_stream)
    r.upload_common(geo.geo, scene)
    r.frame_upload(jobs, scene)

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

    # the same frame in REFERENCE coverage mode, and the pinned spec
    # (256-pixel tiles, Setup.hs:81-83) in both modes, into a frame of their
    # own (`out` keeps the tuned frame for the Level-1 comparison below)
    other_out = torch.zeros_like(out)

    def spec_frame_ms(sp, n=5):
        rr = G.Rasterizer(sp)
        rr.set_stream(stream.cuda_stream)
        rr.upload_common(geo.geo, scene)
        rr.frame_upload(G.RasterJobs(sp, geo, scene.width, scene.height), scene)
        rr.frame_render(other_out.data_ptr())
        rr.frame_check()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(n):
            rr.frame_render(other_out.data_ptr())
        e1.record(stream)
        rr.frame_check()
        rr.close()
        return e0.elapsed_time(e1) / n

    log("reference coverage mode and the pinned spec")
    other_frames = {
        "tuned_spec_reference_mode_ms": spec_frame_ms(G.default_spec(
            texture_filter=args.texture_filter, sort_radix_bits=radix, max_tile_size=tile, coverage_mode=1)),
        "pinned_spec_accurate_ms": spec_frame_ms(G.default_spec(texture_filter=args.texture_filter)),
        "pinned_spec_reference_mode_ms": spec_frame_ms(G.default_spec(texture_filter=args.texture_filter,
                                                                      coverage_mode=1)),
        "pinned_spec": "256-px tiles, 256 threads per tile, maxThresholds 4096, maxShapes 127, radix 0"}
    del other_out
    for k in list(other_frames):
        if k.endswith("_ms"):
            other_frames[k.replace("_ms", "_gpixels_per_s")] = npix / (other_frames[k] * 1e-3) / 1e9
    log(f"other frames: {json.dumps(other_frames)}")

    b_alg = algorithmic_bytes(scene, geo, jobs)
    achieved = b_alg / (fast_ms * 1e-3) / 1e9
    # the same figure with the pinned 256-pixel tiles' (smaller) byte count:
    # smaller tiles read the strands of a shape once per tile it meets
    b_alg_pinned = algorithmic_bytes(scene, geo, G.RasterJobs(G.default_spec(texture_filter=args.texture_filter),
                                                              geo, scene.width, scene.height))
    gpix = npix / (ms * 1e-3) / 1e9
    mcurves = scene.n_curves / (ms * 1e-3) / 1e6

    # Level 1, the reference's generateCall boundary: every raster job with
    # its piles in host memory and the pixels back in a host frame
    # (PCIe-inclusive; not the bench value)
    log("Level 1 (per job, host buffers) of the same frame")
    host_frame = np.zeros(npix, np.uint32)
    t = time.perf_counter()
    r.queue_jobs(jobs, scene, host_frame)
    t_level1 = time.perf_counter() - t
    level1 = {"ms": t_level1 * 1e3, "gpixels_per_s": npix / t_level1 / 1e9, "jobs": jobs.n_jobs,
              "frame_equals_device_frame": bool(np.array_equal(host_frame, out.cpu().numpy().view(np.uint32)))}
    del host_frame

    log("device binning (stage a) of the same frame")
    binning = device_binning_leg(r, scene, geo, stream, max(3, min(args.steps, 10)))
    traffic, traffic_info = measured_traffic(fast_ms)
    copy_gbs = copy_bandwidth()
    log(f"device copy: {copy_gbs:.0f} GB/s achievable")
    log("outlines to frame on the device (serialise + bin + render)")
    pipeline = device_pipeline_leg(r, scene, spec, stream, max(2, min(args.steps, 5)))
    log(f"device pipeline: {json.dumps({k: round(v, 3) for k, v in pipeline.items()})}")

    adjoint = batched_adjoint = None
    if not args.no_adjoint:
        log("adjoint workload (1024^2, 2000 shapes)")
        adjoint = adjoint_leg(spec, max(3, args.batch_steps), log)
        log("batched adjoint workload (64 x 512^2)")
        batched_adjoint = batched_adjoint_leg(spec, max(2, args.batch_steps), log)

    batched = None
    if not args.no_batch:
        log("batched workload (256 scenes)")
        fixed = (args.batch_tile, args.batch_radix) if args.batch_tile > 0 and args.batch_radix >= 0 else None
        batched = batched_leg(spec, args.batch_steps, log, fixed)

    cpu = None
    if not args.no_cpu_baseline:
        log("cpu baseline (oracle on a bounded sample)")
        cpu = cpu_baseline(scene, args.cpu_seconds, args.texture_filter)

    line = {
        "metric": METRIC, "value": gpix, "unit": "Gpixels/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "fp32", "data": "synthetic",
        "mcurves_per_s": mcurves,
        "config": {"workload": "plain raster, 4096x4096 canvas, 200k shapes (180k blobs, 20k glyph "
                               "outlines), 2.14M quadratics; substances 70% solid, 20% linear/radial "
                               "gradients, 10% textures (" + ("bilinear" if args.texture_filter else "nearest")
                               + " sampling)",
                   "canvas": [scene.width, scene.height], "shapes": scene.n_shapes,
                   "curves": scene.n_curves, "tiles": jobs.n_tiles, "tile_shape_refs": jobs.n_refs,
                   "coverage_mode": "accurate", "tile_size": tile, "sort_radix_bits": radix,
                   "autotune_ms": tune_ms},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "traffic_detail": traffic_info,
                     "kernel": "raster_pass (fast pass)", "kernel_ms": fast_ms,
                     "algorithmic_bytes": b_alg, "retry_pass_ms": float(np.mean(retry)),
                     "algorithmic_bytes_256_px_tiles": b_alg_pinned,
                     "frac_with_256_px_tile_bytes": b_alg_pinned / (fast_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                     "peak_measured_copy_gbs": copy_gbs, "frac_of_measured_copy": achieved / copy_gbs,
                     "stage_cycle_share": stage_share},
        "other_frames": other_frames,
        "bin_plus_render_ms": binning["bin_plus_render_ms"],
        "level1_host_buffers": level1,
        "device_binning": binning,
        "device_pipeline": pipeline,
        "batched": batched,
        "adjoint": adjoint,
        "batched_adjoint": batched_adjoint,
        "cpu_baseline": cpu,
        "host_stages_s": {"serialize_and_bin": t_host, "upload": t_upload, "readback": t_read},
    }
    print(json.dumps(line), flush=True)
    r.close()


if __name__ == "__main__":
    main()
