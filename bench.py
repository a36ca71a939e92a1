#!/usr/bin/env python3
"""Training throughput of the gfx950 Gaussian-splatting path on one MI355X.

Workload: 3M Gaussians (seeded 64-blob mixture, scales from the GPU 3-NN initialisation),
1920x1080, SH degree 3, 300 cameras on a radius-5 sphere with targets rendered from the ground
truth and resident in HBM. One step is one full training iteration of the reference schedule
(render, L1+SSIM gradient, backward, Adam, statistics, density control at its 100-iteration
cadence) starting at iteration 3000. The CPU baseline is the oracle timing one full iteration of
the same workload on one host thread. Prints one JSON line.
"""
from __future__ import annotations

import argparse
import importlib.util
import json
import re
import os
import pathlib
import platform
import sys
import time

import numpy as np

REPO = pathlib.Path(__file__).resolve().parent
METRIC = "iters/s + fwd/bwd ms @3M Gaussians 1080p; learned vs rule densify at equal PSNR"
HBM_PEAK_GBS = 8000.0


def load_package():
    if "gs_native_amd" in sys.modules:
        return sys.modules["gs_native_amd"]
    package_dir = REPO / "gs-native_amd"
    spec = importlib.util.spec_from_file_location("gs_native_amd", package_dir / "__init__.py",
                                                  submodule_search_locations=[str(package_dir)])
    module = importlib.util.module_from_spec(spec)
    sys.modules["gs_native_amd"] = module
    spec.loader.exec_module(module)
    return module


def log(message: str) -> None:
    print(f"[bench] {message}", file=sys.stderr, flush=True)


def algorithmic_bytes(n: int, v: float, r: float, p: int, t: int, k: int = 16) -> dict:
    """Compulsory HBM bytes per launch of each stage (one iteration)."""
    return {
        "project": 16 * n + v * (28 + 12 * k + 56),
        "scan": 8 * n,
        "depth_sort": 8 * n,
        "emit": 24 * v + 12 * r,
        "tile_sort": 24 * r,
        "ranges": 8 * r + 8 * (t + 1),
        "raster_forward": 40 * r + 12 * p + 8 * (t + 1),
        "image_gradients": 27 * p,
        "raster_backward": 40 * r + 24 * p + 44 * v,
        "gaussian_gradients": 32 * n + 500 * v + 1624 * n,
    }


def fused_floor_bytes(n: int, r: float) -> dict:
    """Compulsory bytes of the fused gradient + Adam pass itself, which never writes or re-reads a
    58-row gradient buffer: Adam's parameter and moment rows read and written (1392N), the
    statistics (16N), the run offsets (4N) and one flag byte per record (R). The reached records'
    48-byte partials come on top (data-dependent, not counted)."""
    return {"gaussian_gradients": 1412 * n + r}


def profiled_value(key: str, stage: str):
    """A per-kernel figure recorded by a committed rocprofv3 summary of this workload under
    profiles/ ("<key> <stage> = <value>" lines), or None: traffic_bytes_per_launch (FETCH_SIZE +
    WRITE_SIZE scaled by a plain copy of known size) and valu_share (SQ_INSTS_VALU issue share)."""
    for path in sorted((REPO / "profiles").glob("*.md")):
        for line in path.read_text().splitlines():
            match = re.match(r"^(\w+)\s+(\w+)\s*=\s*([0-9.eE+-]+)$", line.strip())
            if match and match.group(1) == key and match.group(2) == stage:
                return float(match.group(3))
    return None


def build_workload(gsn, context, gaussians, width, height, cameras, seed=0, scene="blobs", camera_radius=5.0,
                   hemisphere=False):
    scenes_spec = importlib.util.spec_from_file_location("gs_native_amd.scenes", REPO / "gs-native_amd" / "scenes.py")
    scenes = importlib.util.module_from_spec(scenes_spec)
    scenes_spec.loader.exec_module(scenes)
    if scene == "shell":
        truth = scenes.shell_plane_scene(gaussians, seed=seed, sh_degree=3)
    else:
        truth = scenes.blob_scene(gaussians, seed=seed, sh_degree=3)
    _, mean_dist2 = gsn.knn3(context, truth[0:3].T.copy())
    truth[3:6] = scenes.scales_from_mean_dist2(mean_dist2)[None, :]
    params = scenes.perturb(truth, seed + 1, 0.01)
    focal = scenes.fov_focal(width, 60.0)
    poses = scenes.sphere_cameras(cameras, camera_radius, seed + 2, hemisphere=hemisphere)
    cams = [gsn.Camera.make(width, height, focal, focal, 0.5 * (width - 1), 0.5 * (height - 1), R, t) for R, t in poses]
    centres = np.stack([-R.T @ t for R, t in poses])
    extent = 1.1 * float(np.linalg.norm(centres - centres.mean(0), axis=1).max())
    device_truth = gsn.DeviceBuffer.from_numpy(truth)
    image = gsn.DeviceBuffer((height, width, 3), np.float32)
    targets, visible, records = [], [], []
    projection = None
    for cam in cams:
        projection = gsn.render(context, device_truth, gaussians, cam, (0.0, 0.0, 0.0), image, 3, projection)
        target = np.clip(np.rint(image.numpy() * 255.0), 0, 255).astype(np.uint8)
        targets.append(gsn.DeviceBuffer.from_numpy(target))
    del device_truth
    device_params = gsn.DeviceBuffer.from_numpy(params)
    for cam in cams:
        projection = gsn.render(context, device_params, gaussians, cam, (0.0, 0.0, 0.0), image, 3, projection)
        counts = np.diff(projection.record_offsets.numpy())
        visible.append(int((counts > 0).sum()))
        records.append(int(projection.record_count))
    del device_params
    return params, cams, targets, extent, visible, records


def cpu_baseline(params: np.ndarray, camera, target: np.ndarray) -> dict:
    """The oracle (C, one thread) on one full training iteration of this same workload, first
    camera: projection, render, L1+SSIM gradient, replay backward, chain rule (FP64) and Adam
    (FP32), all N Gaussians at the full image size. One timed run: it takes tens of seconds."""
    sys.path.insert(0, str(REPO))
    import oracle
    first = np.zeros_like(params)
    second = np.zeros_like(params)
    sh_degree = 3
    start = time.perf_counter()
    projected = oracle.project_f64(params, camera, sh_degree)
    image, _ = oracle.render_f64(camera, projected, (0.0, 0.0, 0.0))
    image_grad = oracle.image_gradients_f64(image, target)
    back = oracle.backward_f64(camera, projected, image, image_grad)
    mapped = oracle.map_f64(params, camera, sh_degree, back["screen"], projected["visible"])
    oracle.adam_f32(mapped["gradients"].astype(np.float32), params, first, second, 1.6e-4, 10.0, 1000.0)
    elapsed = time.perf_counter() - start
    count = params.shape[1]
    return {"value": 1.0 / elapsed, "unit": "iters/s", "cores": 1, "kind": "port",
            "sample": f"one full training iteration of this workload ({count} Gaussians, {camera.width}x{camera.height}, "
                      f"SH 3, camera 0): oracle FP64 projection/render/L1+SSIM gradient/replay backward/chain rule + "
                      f"FP32 Adam, one timed run ({elapsed:.1f} s), single thread on "
                      f"{platform.processor() or platform.machine()}"}


def main() -> int:
    parser = argparse.ArgumentParser(description=__doc__)
    parser.add_argument("--gpus", type=int, default=1)
    parser.add_argument("--steps", type=int, default=100)
    parser.add_argument("--warmup", type=int, default=5)
    parser.add_argument("--gaussians", type=int, default=3_000_000)
    parser.add_argument("--width", type=int, default=1920)
    parser.add_argument("--height", type=int, default=1080)
    parser.add_argument("--cameras", type=int, default=300)
    parser.add_argument("--start-iteration", type=int, default=3000)
    parser.add_argument("--scene", choices=["blobs", "shell"], default="blobs",
                        help="blobs: seeded 64-blob mixture; shell: radius-2 shell plus ground plane")
    parser.add_argument("--camera-radius", type=float, default=5.0)
    parser.add_argument("--hemisphere", action="store_true", help="cameras on the upper hemisphere only")
    parser.add_argument("--stage-steps", type=int, default=20,
                        help="iterations of the separate pass that times every stage (before the timed steps)")
    parser.add_argument("--no-cpu-baseline", action="store_true")
    parser.add_argument("--calibration-copy-mib", type=int, default=0,
                        help="after the timed steps, run one device-to-device copy of this many MiB (a plain copy "
                             "kernel of known size for scaling FETCH_SIZE / WRITE_SIZE counters)")
    args = parser.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only: --gpus {args.gpus} is not supported (the training step runs on a "
              f"single MI355X; multi-GPU sharding is not implemented).", file=sys.stderr)
        return 2
    gsn = load_package()
    trainer_spec = importlib.util.spec_from_file_location("gs_native_amd.trainer", REPO / "gs-native_amd" / "trainer.py")
    trainer_module = importlib.util.module_from_spec(trainer_spec)
    trainer_spec.loader.exec_module(trainer_module)
    context = gsn.Context()
    log(f"device {gsn.device_name()}")
    log(f"building workload: {args.gaussians} Gaussians, {args.width}x{args.height}, {args.cameras} cameras")
    params, cameras, targets, extent, visible, records = build_workload(
        gsn, context, args.gaussians, args.width, args.height, args.cameras, scene=args.scene,
        camera_radius=args.camera_radius, hemisphere=args.hemisphere)
    log(f"workload ready: mean visible {np.mean(visible):.0f}, mean records {np.mean(records):.0f}")
    trainer = trainer_module.Trainer(context, params, cameras, targets, extent, completed=args.start_iteration)
    for _ in range(args.warmup):
        trainer.iteration()
    context.synchronize()
    log(f"warm-up done ({args.warmup} iterations), gaussians {trainer.count}")
    p = args.width * args.height
    t = ((args.width + 15) // 16) * ((args.height + 15) // 16)
    # Stage breakdown: a separate pass with an event pair around every stage. Each event record
    # leaves the stream idle for a few microseconds, so the timed steps below record events around
    # the dominant stage only.
    stage_steps = max(args.stage_steps, 0)
    per_stage = {}
    dominant = "gaussian_gradients"
    if stage_steps:
        context.set_timing(True)
        counts = [0, 0, 0]
        for _ in range(stage_steps):
            counts[0] += trainer.count
            counts[1] += trainer.iteration()
            counts[2] += visible[trainer.last_camera]
        stages = context.stage_times()
        context.set_timing(False)
        stage_bytes = algorithmic_bytes(*(c / stage_steps for c in counts), p, t)
        for name, (ms, launches) in stages.items():
            if launches:
                per_stage[name] = {"ms_per_iteration": ms / stage_steps, "launches": launches}
                if name in stage_bytes:
                    per_stage[name]["hbm_gbs_algorithmic"] = stage_bytes[name] / (ms / launches * 1e-3) / 1e9
        candidates = {name: stages[name][0] for name in stage_bytes if stages[name][1]}
        dominant = max(candidates, key=candidates.get)
        log(f"stage pass done ({stage_steps} iterations), dominant stage {dominant}")
    context.set_timing(True, stages=[dominant])
    start_count = trainer.count
    first_timed = trainer.completed + 1
    record_total = 0
    visible_total = 0
    gaussian_total = 0
    begin = time.perf_counter()
    for step in range(args.steps):
        gaussian_total += trainer.count
        record_total += trainer.iteration()
        visible_total += visible[trainer.last_camera]
        if (step + 1) % 25 == 0:
            log(f"step {step + 1}/{args.steps}")
    context.synchronize()
    elapsed = time.perf_counter() - begin
    timed = context.stage_times()
    context.set_timing(False)
    ms_per_step = 1000.0 * elapsed / args.steps
    n_mean = gaussian_total / args.steps
    r_mean = record_total / args.steps
    v_mean = visible_total / args.steps
    bytes_per_launch = algorithmic_bytes(n_mean, v_mean, r_mean, p, t)
    launches = timed[dominant][1]
    avg_ms = timed[dominant][0] / launches
    achieved = bytes_per_launch[dominant] / (avg_ms * 1e-3) / 1e9
    result = {
        "metric": METRIC,
        "value": args.steps / elapsed,
        "unit": "iters/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"{args.gaussians:,}-Gaussian "
                        f"{'shell + ground-plane' if args.scene == 'shell' else '64-blob'} scene, "
                        f"{args.width}x{args.height}, SH degree 3, {args.cameras} cameras "
                        f"({'hemisphere' if args.hemisphere else 'sphere'}, radius {args.camera_radius:g}), "
                        f"training iterations {first_timed}-{first_timed + args.steps - 1} of the reference schedule",
            "gaussians_start": start_count,
            "gaussians_end": trainer.count,
            "mean_records": r_mean,
            "mean_visible": v_mean,
            "density_steps": [h for h in trainer.history],
        },
        "stages": per_stage,
        "stage_pass": {"iterations": stage_steps, "note": "per-stage times from a separate pass before the timed "
                       "steps, with an event pair around every stage; the timed steps record events around "
                       "the dominant stage only"},
        "roofline": {
            "kernel": dominant,
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": None,
            "avg_launch_ms": avg_ms,
            "algorithmic_bytes_per_launch": bytes_per_launch[dominant],
        },
        "cpu_baseline": None,
    }
    floor = fused_floor_bytes(n_mean, r_mean).get(dominant)
    if floor is not None:
        result["roofline"]["fused_floor_bytes_per_launch"] = floor
        result["roofline"]["frac_of_fused_floor"] = floor / (avg_ms * 1e-3) / 1e9 / HBM_PEAK_GBS
    result["roofline"]["traffic"] = profiled_value("traffic_bytes_per_launch", dominant)
    result["roofline"]["valu_share"] = profiled_value("valu_share", dominant)
    if args.calibration_copy_mib > 0:
        size = args.calibration_copy_mib << 20
        source = gsn.DeviceBuffer((size,), np.uint8)
        destination = gsn.DeviceBuffer((size,), np.uint8)
        gsn.call("gsn_memcpy_device", destination.ptr, source.ptr, size)
        context.synchronize()
        log(f"calibration copy of {size} bytes done")
        del source, destination
    if not args.no_cpu_baseline:
        log("timing the CPU baseline: one oracle iteration of this workload (tens of seconds)")
        result["cpu_baseline"] = cpu_baseline(params, cameras[0], targets[0].numpy())
    print(json.dumps(result), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
