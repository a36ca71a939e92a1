"""Peak device memory and time of one training step with the recompute backward (nothing stored
per pixel) against the stored-state backward (final transmittance + last contributor per pixel),
at 3840x2160 with 2M Gaussians from the seeded 64-blob mixture, SH degree 3.

Each form runs in its own process (so that neither inherits the other's allocations); device
memory in use is read from hipMemGetInfo before the scene is uploaded and after each step (every
allocation of the step, scratch included, is still held then, so the reading is the step's peak).
Prints one JSON line.

Usage: python gs-native_amd/memory_comparison.py [--gaussians 2000000] [--size 3840x2160]
"""
from __future__ import annotations

import argparse
import ctypes
import importlib.util
import json
import pathlib
import subprocess
import sys
import time

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent


def _load():
    spec = importlib.util.spec_from_file_location("gs_native_amd", HERE / "__init__.py",
                                                  submodule_search_locations=[str(HERE)])
    module = importlib.util.module_from_spec(spec)
    sys.modules["gs_native_amd"] = module
    spec.loader.exec_module(module)
    return module


def measure(form: str, gaussians: int, width: int, height: int, steps: int) -> dict:
    gsn = _load()
    import gs_native_amd.scenes as scenes
    context = gsn.Context()
    free0, total = gsn.device_memory_info()
    params = scenes.blob_scene(gaussians, seed=5, sh_degree=3)
    _, d2 = gsn.knn3(context, params[0:3].T.copy())
    params[3:6] = scenes.scales_from_mean_dist2(d2)[None, :]
    focal = scenes.fov_focal(width, 60.0)
    R, t = scenes.look_at(np.array([0.5, -1.0, 4.6]))
    cam = gsn.Camera.make(width, height, focal, focal, 0.5 * (width - 1), 0.5 * (height - 1), R, t)
    truth = gsn.DeviceBuffer.from_numpy(params)
    image = gsn.DeviceBuffer((height, width, 3), np.float32)
    projection = gsn.render(context, truth, gaussians, cam, (0.0, 0.0, 0.0), image, 3)
    target = gsn.DeviceBuffer.from_numpy(np.clip(np.rint(image.numpy() * 255.0), 0, 255).astype(np.uint8))
    del truth, projection
    device_params = gsn.DeviceBuffer.from_numpy(scenes.perturb(params, 6, 0.01))
    moments = gsn.Moments(gaussians)
    stats = gsn.DeviceBuffer.zeros((2, gaussians), np.float32)
    abs_stats = gsn.DeviceBuffer.zeros((2, gaussians), np.float32)
    image_grad = gsn.DeviceBuffer((height, width, 3), np.float32)
    state = None
    if form == "stored":
        state = (gsn.DeviceBuffer((width * height,), np.float32), gsn.DeviceBuffer((width * height,), np.int32))
    background = (ctypes.c_float * 3)(0.0, 0.0, 0.0)
    projection = None
    times = []
    for step in range(steps + 1):
        context.synchronize()
        start = time.perf_counter()
        projection = gsn.render(context, device_params, gaussians, cam, (0.0, 0.0, 0.0), image, 3, projection,
                                stored_state=state)
        gsn.image_gradients(context, cam, image, target, image_grad)
        c1, c2 = gsn.adam_corrections(step + 1)
        p = projection
        if state is None:
            gsn.gaussian_gradients_adam(context, p, device_params, 3, image, image_grad, moments, 1e-4, c1, c2,
                                        stats, abs_stats)
        else:
            gsn.call("gsn_gaussian_gradients_adam_stored", context.handle, ctypes.byref(p.camera), p.count, 3,
                     device_params.ptr, p.tile_gaussian_ids.ptr, p.tile_offsets.ptr, p.record_count,
                     p.record_offsets.ptr, p.tile_rects.ptr, p.means.ptr, p.conic_opacities.ptr, p.colors.ptr,
                     state[0].ptr, state[1].ptr, background, image_grad.ptr, moments.ptr, 1e-4, c1, c2,
                     stats.ptr, abs_stats.ptr)
        context.synchronize()
        if step:
            times.append(time.perf_counter() - start)
    free1, _ = gsn.device_memory_info()
    checksum = float(np.abs(device_params.numpy()).astype(np.float64).sum())
    return {"form": form, "peak_bytes": free0 - free1, "total_bytes": total, "records": projection.record_count,
            "ms_per_step": 1000.0 * float(np.median(times)), "params_abs_sum": checksum,
            "per_pixel_state_bytes": 0 if state is None else 8 * width * height}


def main() -> int:
    parser = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    parser.add_argument("--gaussians", type=int, default=2_000_000)
    parser.add_argument("--size", type=str, default="3840x2160")
    parser.add_argument("--steps", type=int, default=5)
    parser.add_argument("--form", choices=["recompute", "stored"], default=None)
    args = parser.parse_args()
    width, height = (int(v) for v in args.size.split("x"))
    if args.form:
        print(json.dumps(measure(args.form, args.gaussians, width, height, args.steps)), flush=True)
        return 0
    results = {}
    for form in ("recompute", "stored"):
        out = subprocess.run([sys.executable, __file__, "--form", form, "--gaussians", str(args.gaussians),
                              "--size", args.size, "--steps", str(args.steps)], capture_output=True, text=True)
        if out.returncode != 0:
            sys.stderr.write(out.stdout + out.stderr)
            return out.returncode
        results[form] = json.loads(out.stdout.strip().splitlines()[-1])
        print(f"[memory] {form}: {results[form]['peak_bytes'] / 2**30:.3f} GiB, "
              f"{results[form]['ms_per_step']:.2f} ms/step", file=sys.stderr, flush=True)
    summary = {
        "gaussians": args.gaussians, "size": args.size, "recompute": results["recompute"],
        "stored": results["stored"],
        "peak_saving_bytes": results["stored"]["peak_bytes"] - results["recompute"]["peak_bytes"],
        "params_relative_difference": abs(results["stored"]["params_abs_sum"] - results["recompute"]["params_abs_sum"])
        / results["recompute"]["params_abs_sum"],
    }
    print(json.dumps(summary), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
