"""Hot-path parity of the HIP training step against the oracle, stage by stage and end to end.

Comparison protocol (so that no Gaussian has to be dropped from a comparison):
* A pixel whose alpha, 0.99-cap or T < 1e-4 decision lies within 1e-4 relative of its threshold
  (flagged by the oracle's FP64 render) may legitimately take the other branch in FP32. Such
  pixels have their dL/dimage set to zero for both the device backward and the FP64 replay: each
  pixel's contribution to every gradient is proportional to its own dL/dimage, so the pixel drops
  out of every gradient on both sides while every Gaussian stays in the comparison. The share of
  such pixels is printed (0.01-0.1% here).
* Rasteriser stages are compared on the device's own FP32 projection (fed to the FP64 replay),
  the chain rule on the device's own screen-space gradients (fed to the FP64 chain rule), so
  each stage's error is its own; the end-to-end comparison runs the whole FP64 chain from the
  parameters.

Bars (FP32 against FP64; relative errors are |gpu - fp64| / |fp64|):
* every element: |gpu - fp64| <= 1e-5 (|fp64| + m), m = the sum of the absolute values of the
  terms FP32 rounding acts on, weighted by the transmittance's condition number (the oracle's
  error-bound magnitude);
* the share of elements within 1e-5 of their own value is printed; elements outside it are the
  ill-conditioned sums of terms of both signs, and they may be at most 12% of a row at the
  10,000-Gaussian 256x256 size (10,000 overlapping splats, ~80 per pixel) and 3% at the
  2,000-Gaussian 512x512 size;
* every row: ||gpu - fp64|| <= 1e-6 ||fp64|| per stage, <= 1e-5 end to end.
"""
import numpy as np
import pytest

from helpers import GpuRun, config0_camera

pytestmark = pytest.mark.gpu

SCREEN_NAMES = ("dmean_x", "dmean_y", "dconic_a", "dconic_b", "dconic_c", "dlogit", "dred", "dgreen", "dblue",
                "|dmean_x|", "|dmean_y|")


def device_projection_f64(oracle, params, cam, sh_degree):
    f32 = oracle.project_f32(params, cam, sh_degree)
    visible = f32["record_counts"][:-1] > 0
    projected = {"means": f32["means"].astype(np.float64), "conics": f32["conic_opacities"].astype(np.float64),
                 "colors": f32["colors"].astype(np.float64), "depths": f32["depths"].astype(np.float64),
                 "tile_rects": np.ascontiguousarray(f32["tile_rects"], np.int32), "visible": visible.astype(np.uint8)}
    return f32, projected


def relative_share(gpu, ref):
    nonzero = ref != 0
    if not nonzero.any():
        return 1.0, 0.0
    error = np.abs(gpu - ref)[nonzero] / np.abs(ref[nonzero])
    return float(np.mean(error <= 1e-5)), float(error.max())


def check_rows(label, gpu, ref, magnitude, min_share, norm_bound, names=None):
    """The three bars above, row by row; returns the printed summary."""
    lines = []
    for row in range(ref.shape[0]):
        name = names[row] if names else f"row {row}"
        g, r = gpu[row].astype(np.float64), ref[row]
        if magnitude is not None:
            bound = 1e-5 * (np.abs(r) + magnitude[row]) + 1e-30
            worst = np.max(np.abs(g - r) / bound)
            assert worst <= 1.0, f"{label} {name}: error above 1e-5 (|fp64| + magnitude), worst ratio {worst:.3g}"
        if not np.any(r != 0):
            assert np.all(g == 0), f"{label} {name}: non-zero gradient where the oracle has none"
            continue
        share, max_rel = relative_share(g, r)
        norm = np.linalg.norm(g - r) / np.linalg.norm(r)
        lines.append(f"{name}: {100 * share:.2f}% within 1e-5 (max {max_rel:.1e}), norm {norm:.1e}")
        assert share >= min_share, f"{label} {name}: only {share:.4f} of the elements within 1e-5 relative"
        assert norm <= norm_bound, f"{label} {name}: normwise error {norm:.3g}"
    print(f"[{label}] " + "; ".join(lines), flush=True)


def raster_and_chain_parity(gsn, context, oracle, params, cam, target, sh_degree, min_share, label):
    """Forward image, backward (screen space) and chain rule, each against the FP64 oracle on the
    device's own inputs. Returns what the end-to-end comparison needs."""
    run = GpuRun(gsn, context, params, cam, (0.0, 0.0, 0.0), sh_degree)
    f32, projected = device_projection_f64(oracle, params, cam, sh_degree)
    # Projection: bit-exact against the FP32 mirror.
    device = run.projected()
    np.testing.assert_array_equal(device["conic_opacities"].view(np.uint32), f32["conic_opacities"].view(np.uint32))
    np.testing.assert_array_equal(device["means"].view(np.uint32), f32["means"].view(np.uint32))
    # Forward: FP64 render of the device's projection.
    reference, ambiguous = oracle.render_f64(cam, projected, (0.0, 0.0, 0.0))
    clear = ambiguous == 0
    image = run.image_host()
    difference = np.abs(image - reference)[clear]
    relative = difference / np.maximum(np.abs(reference[clear]), 1e-30)
    print(f"[{label}] ambiguous pixels {100 * (1 - clear.mean()):.3f}%; image: "
          f"{100 * np.mean(relative <= 1e-5):.3f}% of values within 1e-5 relative, max abs {difference.max():.1e}",
          flush=True)
    assert np.mean(relative <= 1e-5) >= 0.999
    assert difference.max() <= 3e-6
    # Backward: the device's dL/dimage with the ambiguous pixels zeroed, on both sides.
    target_device = gsn.DeviceBuffer.from_numpy(target)
    image_grad = gsn.DeviceBuffer((cam.height, cam.width, 3), np.float32)
    gsn.image_gradients(context, cam, run.image, target_device, image_grad)
    grad = image_grad.numpy()
    grad[~clear] = 0.0
    grad_device = gsn.DeviceBuffer.from_numpy(grad)
    n = params.shape[1]
    gradients = gsn.DeviceBuffer.zeros((58, n), np.float32)
    statistics = gsn.DeviceBuffer.zeros((2, n), np.float32)
    absolute = gsn.DeviceBuffer.zeros((2, n), np.float32)
    screen = gsn.DeviceBuffer.zeros((11, n), np.float32)
    gsn.gaussian_gradients(context, run.projection, run.params, sh_degree, run.image, grad_device, gradients,
                           statistics, absolute, screen)
    context.synchronize()
    back = oracle.backward_f64(cam, projected, reference, grad.astype(np.float64))
    check_rows(f"{label} backward", screen.numpy(), back["screen"], back["magnitude"], min_share, 1e-6,
               SCREEN_NAMES)
    # Chain rule: FP64 on the device's own screen-space gradients.
    device_screen = screen.numpy().astype(np.float64)
    mapped = oracle.map_f64(params, cam, sh_degree, device_screen, projected["visible"])
    check_rows(f"{label} chain rule", gradients.numpy(), mapped["gradients"], None, 0.96, 1e-6)
    np.testing.assert_allclose(statistics.numpy()[0], mapped["statistics"][0], rtol=1e-6, atol=1e-30)
    return {"run": run, "grad": grad, "clear": clear}


def test_config0_full_iteration_matches_oracle(gsn, context, oracle, scenes):
    """10,000 Gaussians (means U[-1,1]^3, log-scales N(-4, 0.5), random axis-angle, opacity logit
    N(0,1), SH degree 0), one 256x256 fov-60 camera at distance 3, target rendered from a copy
    perturbed by N(0, 0.01) and quantised to u8; render -> L1+SSIM gradient -> backward -> one
    fused Adam step with position learning rate 1.6e-4."""
    params = scenes.config0_params(seed=0, count=10_000, sh_degree=0)
    cam = config0_camera(gsn)
    truth = scenes.perturb(params, 1, 0.01)
    target = np.clip(np.rint(oracle.render_reference(truth, cam, (0.0, 0.0, 0.0), 0) * 255.0), 0, 255).astype(np.uint8)
    stages = raster_and_chain_parity(gsn, context, oracle, params, cam, target, 0, 0.88, "config0")
    run, grad = stages["run"], stages["grad"]
    # dL/dimage against FP64 on the device image, away from pixels whose L1 sign is decided
    # within rounding (|I - T| <= 1e-5 |I|).
    image = run.image_host().astype(np.float64)
    fp64_grad = oracle.image_gradients_f64(image, target)
    sign_ambiguous = np.abs(image - target / 255.0) <= 1e-5 * np.abs(image)
    scale = 1.0 / (3 * cam.width * cam.height)
    keep = stages["clear"][..., None] & ~sign_ambiguous
    device_grad = gsn.DeviceBuffer((cam.height, cam.width, 3), np.float32)
    gsn.image_gradients(context, cam, run.image, gsn.DeviceBuffer.from_numpy(target), device_grad)
    # In FP32 the windowed variances E[x^2] - E[x]^2 carry an absolute error of about one ulp of
    # E[x^2] (6e-8), which the SSIM derivative divides by (variance + C2), C2 = 9e-4: up to ~1e-4
    # of the gradient scale 1/(3WH) in flat windows (this scene has a dark background between
    # its splats); elsewhere the error stays below 1e-5 of it.
    error = np.abs(device_grad.numpy() - fp64_grad)[keep] / scale
    print(f"[config0] dL/dimage: {100 * np.mean(error <= 1e-5):.3f}% within 1e-5 of 1/(3WH), "
          f"max {error.max():.1e}", flush=True)
    assert np.mean(error <= 1e-5) >= 0.95
    assert error.max() <= 3e-4
    # End to end: one fused gradient + Adam step against the FP64 chain from the parameters
    # (projection, replay backward and chain rule in FP64 on the same dL/dimage) and FP32 Adam.
    n = params.shape[1]
    moments = gsn.Moments(n)
    statistics = gsn.DeviceBuffer.zeros((2, n), np.float32)
    absolute = gsn.DeviceBuffer.zeros((2, n), np.float32)
    c1, c2 = gsn.adam_corrections(1)
    gsn.gaussian_gradients_adam(context, run.projection, run.params, 0, run.image, gsn.DeviceBuffer.from_numpy(grad),
                                moments, 1.6e-4, c1, c2, statistics, absolute)
    context.synchronize()
    f32 = oracle.project_f32(params, cam, 0)
    projected = oracle.project_f64(params, cam, 0)
    projected["tile_rects"] = np.ascontiguousarray(f32["tile_rects"], np.int32)
    projected["visible"] = ((f32["record_counts"][:-1] > 0) & (projected["visible"] > 0)).astype(np.uint8)
    image64, _ = oracle.render_f64(cam, projected, (0.0, 0.0, 0.0))
    back = oracle.backward_f64(cam, projected, image64, grad.astype(np.float64))
    expected_gradients = oracle.map_f64(params, cam, 0, back["screen"], projected["visible"])["gradients"]
    zeros = np.zeros_like(params)
    expected_params, expected_first, expected_second = oracle.adam_f32(
        expected_gradients.astype(np.float32), params, zeros, zeros, 1.6e-4, c1, c2)
    active = [row for row in range(58) if np.any(expected_gradients[row] != 0)]
    (got_first, got_second), got_params = moments.numpy(), run.params.numpy()
    lines = []
    for row in active:
        norm_m = np.linalg.norm(got_first[row] - expected_first[row]) / np.linalg.norm(expected_first[row])
        norm_v = np.linalg.norm(got_second[row] - expected_second[row]) / np.linalg.norm(expected_second[row])
        assert norm_m <= 1e-5 and norm_v <= 2e-5, (row, norm_m, norm_v)
        # The first step moves each parameter by lr * g / (|g| + 1e-15): lr * sign(g) unless |g|
        # is near 1e-15, where the step follows the gradient's own relative error, and where g
        # changes sign within its error. Those elements are counted and must stay rare.
        step = expected_params[row] - params[row]
        moved = got_params[row] - params[row]
        same = np.abs(moved - step) <= 1e-5 * np.abs(step) + 1e-7 * np.abs(params[row]) + 1e-12
        share = same.mean()
        norm_p = np.linalg.norm(moved - step) / np.linalg.norm(step)
        lines.append(f"row {row}: first {norm_m:.1e}, second {norm_v:.1e}, step {100 * share:.2f}% within 1e-5 "
                     f"(norm {norm_p:.1e})")
        assert share >= 0.99, (row, share)
        assert norm_p <= 0.05, (row, norm_p)
    # Rows without gradient: Adam leaves parameters and moments exactly as they were.
    for row in range(58):
        if row not in active:
            np.testing.assert_array_equal(got_params[row], params[row])
            assert not np.any(got_first[row]) and not np.any(got_second[row])
    print("[config0 end to end] " + "; ".join(lines), flush=True)


@pytest.mark.parametrize("view", range(3))
def test_config2_views_match_oracle(gsn, context, oracle, scenes, view):
    """2,000 Gaussians with the config-0 distributions at SH degree 3, 512x512, fov 60, cameras on a
    radius-3 sphere: forward, backward and chain rule per view."""
    params = scenes.config0_params(seed=21, count=2000, sh_degree=3)
    truth = scenes.perturb(params, 22, 0.01)
    focal = scenes.fov_focal(512, 60.0)
    R, t = scenes.sphere_cameras(3, 3.0, 23)[view]
    cam = gsn.Camera.make(512, 512, focal, focal, 255.5, 255.5, R, t)
    target = np.clip(np.rint(oracle.render_reference(truth, cam, (0.0, 0.0, 0.0), 3) * 255.0), 0, 255).astype(np.uint8)
    raster_and_chain_parity(gsn, context, oracle, params, cam, target, 3, 0.97, f"config2 view {view}")


def test_screen_filling_splats_match_oracle(gsn, context, oracle, scenes):
    """600 large Gaussians (config-0 distributions with log-scales raised by 2.5) at 256x256, SH
    degree 3: most own one to two hundred records, so nearly every 64-record chunk of the
    per-Gaussian pass is owned by few Gaussians and summed by the wave-wide reduction rather than
    lane by lane. Forward, backward and chain rule against FP64. The large overlapping splats make
    the screen-space sums cancel, so fewer elements are within 1e-5 of |fp64| alone (about 75% for
    the position and conic rows, the same share as a lane-by-lane sum of the same records gives;
    the wave-wide sums have the lower normwise error); every element stays within
    1e-5 (|fp64| + magnitude) and the normwise error within 1e-6."""
    params = scenes.config0_params(seed=31, count=600, sh_degree=3)
    params[3:6] += np.float32(2.5)
    truth = scenes.perturb(params, 32, 0.01)
    cam = config0_camera(gsn)
    target = np.clip(np.rint(oracle.render_reference(truth, cam, (0.0, 0.0, 0.0), 3) * 255.0), 0, 255).astype(np.uint8)
    stages = raster_and_chain_parity(gsn, context, oracle, params, cam, target, 3, 0.70, "screen-filling")
    owned = np.diff(stages["run"].projection.record_offsets.numpy())
    print(f"[screen-filling] records per Gaussian: median {np.median(owned):.0f}, max {owned.max()}, "
          f"share with 16 or more {np.mean(owned >= 16):.3f}", flush=True)
    assert np.mean(owned >= 64) >= 0.5


def test_needle_splats_take_the_fp64_conic_and_match_fp64(gsn, context, oracle):
    """Thin splats seen side-on, diagonal on screen: their 2D covariance has det < 1e-3 a c, so
    the device forms det and the conic from exact FP64 products (the one FP64 path on the GPU).
    The conic then matches the FP64 oracle to within the conditioning of its FP32 inputs a, b, c
    (16 ulp x (a c + b^2) / det), and stays bit-exact with the FP32 mirror, which takes the same
    branch."""
    count = 24
    params = np.zeros((58, count), np.float32)
    rng = np.random.default_rng(4)
    angles = np.concatenate([np.linspace(np.pi / 8, 3 * np.pi / 8, count // 2),
                             np.linspace(5 * np.pi / 8, 7 * np.pi / 8, count - count // 2)])
    lengths = np.geomspace(0.5, 3.0, count)[rng.permutation(count)]
    for i in range(count):
        params[0:3, i] = (rng.uniform(-0.3, 0.3), rng.uniform(-0.3, 0.3), 4.0)
        params[3:6, i] = np.log([lengths[i], 1e-3, 1.5e-3])
        params[6:9, i] = (0.0, 0.0, angles[i])  # long axis rotated in the image plane
        params[9, i] = 0.5
        params[10, i] = 0.5 / 0.28209479177387814
    cam = gsn.Camera.make(256, 256, 500.0, 500.0, 127.5, 127.5, None, [0.0, 0.0, 0.0])
    run = GpuRun(gsn, context, params, cam, (0.0, 0.0, 0.0), 0)
    device = run.projected()["conic_opacities"]
    mirror = oracle.project_f32(params, cam, 0)["conic_opacities"]
    np.testing.assert_array_equal(device.view(np.uint32), mirror.view(np.uint32))
    fp64 = oracle.project_f64(params, cam, 0)
    assert np.all(fp64["visible"] == 1)
    worst = 0.0
    for i in range(count):
        q = fp64["conics"][i]
        sigma = np.linalg.inv(np.array([[q[0], q[1]], [q[1], q[2]]]))
        a, b, c = sigma[0, 0], sigma[0, 1], sigma[1, 1]
        det = a * c - b * b
        assert det < 1e-3 * a * c, f"splat {i} is not near-degenerate (det / ac = {det / (a * c):.2e})"
        conditioning = (a * c + b * b) / det
        bound = 16 * np.finfo(np.float32).eps * conditioning
        error = np.max(np.abs(device[i, :3] - q[:3])) / np.max(np.abs(q[:3]))
        worst = max(worst, error / bound)
        assert error <= bound, (i, error, bound)
    print(f"needle conics: worst error / bound = {worst:.3f}", flush=True)
