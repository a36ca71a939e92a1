"""Pins the parts of the CPU oracle that the reference's own tests leave open, each against
something independent of the oracle's own code (no GPU):

* the axis-angle rotation against scipy's Rotation.from_rotvec;
* the 16 SH basis functions against scipy's complex spherical harmonics (real form, graphics
  sign convention) and against orthonormality on the sphere;
* the 2D covariance of one rotated, anisotropic, off-axis Gaussian against the empirical
  covariance of projected samples (the linearisation the EWA projection stands for), and the
  clamped-Jacobian case against a closed form built from scipy rotations;
* front-to-back compositing of stacked Gaussians against hand-computed pixel values: the 0.99
  alpha cap, the skipped alpha < 1/255 splat, depth order independent of id order, and the
  terminating splat (T(1 - alpha) < 1e-4) that is not blended.
Each check is chosen so that a plausible mistake (a transposed rotation, a wrong sign or
constant, a missing clamp, blending the terminating splat) fails it.
"""

import numpy as np
import pytest
from scipy.spatial.transform import Rotation

scipy_special = pytest.importorskip("scipy.special")

SH_C0 = 0.28209479177387814


def make_camera(oracle, width, height, fx, fy, cx, cy, rotation=None, translation=None, near=0.2):
    cam = oracle.Camera()
    cam.fx, cam.fy, cam.cx, cam.cy = fx, fy, cx, cy
    rotation = np.eye(3) if rotation is None else np.asarray(rotation)
    translation = np.zeros(3) if translation is None else np.asarray(translation)
    for k, value in enumerate(rotation.reshape(9)):
        cam.R_c_w[k] = float(value)
    for k, value in enumerate(translation.reshape(3)):
        cam.t_c_w[k] = float(value)
    cam.near_depth, cam.far_depth = near, float(np.finfo(np.float32).max)
    cam.width, cam.height = width, height
    cam.tile_count_x, cam.tile_count_y = (width + 15) // 16, (height + 15) // 16
    return cam


def camera_matrices(cam):
    """The camera's rotation and translation as the FP32 values the oracle reads."""
    return (np.array(cam.R_c_w[:], np.float32).astype(np.float64).reshape(3, 3),
            np.array(cam.t_c_w[:], np.float32).astype(np.float64))


def test_rotation_matches_scipy_rotvec(oracle):
    rng = np.random.default_rng(3)
    axes = rng.normal(size=(300, 3))
    axes /= np.linalg.norm(axes, axis=1, keepdims=True)
    vectors = axes * rng.uniform(0.0, np.pi, size=(300, 1))
    special = np.array([[0.0, 0.0, 0.0], [1e-3, 0.0, 0.0], [0.0, -2e-4, 1e-4], [np.pi - 1e-6, 0.0, 0.0],
                        [0.0, 3.0, 0.0], [-1.0, 2.0, -2.5]])
    for v in np.concatenate([vectors, special]):
        want = Rotation.from_rotvec(v).as_matrix()
        np.testing.assert_allclose(oracle.rotation_f64(v), want, rtol=0, atol=1e-14)
        v32 = v.astype(np.float32)
        want32 = Rotation.from_rotvec(v32.astype(np.float64)).as_matrix()
        np.testing.assert_allclose(oracle.rotation_f32(v32), want32, rtol=0, atol=2e-6)
    # Below the 1e-8 angle threshold the rotation is the identity (renderer.cu:38-49).
    np.testing.assert_array_equal(oracle.rotation_f64([1e-9, 0.0, 0.0]), np.eye(3))


def scipy_real_sh(l, m, direction):
    """Real SH from scipy's complex Y_l^m (Condon-Shortley phase included): m > 0 -> sqrt2 Re,
    m < 0 -> sqrt2 Im of Y_l^|m|, m = 0 -> Y_l^0; z is the polar axis."""
    x, y, z = direction
    theta = np.arccos(np.clip(z, -1.0, 1.0))
    phi = np.arctan2(y, x)
    value = scipy_special.sph_harm_y(l, abs(m), theta, phi)
    if m == 0:
        return value.real
    return np.sqrt(2.0) * (value.real if m > 0 else value.imag)


def test_sh_basis_matches_scipy_real_harmonics(oracle):
    rng = np.random.default_rng(5)
    directions = rng.normal(size=(400, 3))
    directions /= np.linalg.norm(directions, axis=1, keepdims=True)
    directions = np.concatenate([directions, np.eye(3), -np.eye(3)])
    for d in directions:
        basis = oracle.sh_basis_f64(d)
        want = [scipy_real_sh(l, m, d) for l in range(4) for m in range(-l, l + 1)]
        np.testing.assert_allclose(basis, want, rtol=0, atol=1e-13)


def test_sh_basis_is_orthonormal_on_the_sphere(oracle):
    """Gauss-Legendre in cos(theta) x uniform phi integrates products of degree <= 6 exactly."""
    nodes, weights = np.polynomial.legendre.leggauss(12)
    phis = np.arange(24) * (2 * np.pi / 24)
    gram = np.zeros((16, 16))
    for z, w in zip(nodes, weights):
        r = np.sqrt(1.0 - z * z)
        for phi in phis:
            b = oracle.sh_basis_f64([r * np.cos(phi), r * np.sin(phi), z])
            gram += w * (2 * np.pi / 24) * np.outer(b, b)
    np.testing.assert_allclose(gram, np.eye(16), atol=1e-12)


def projected_covariance(conic):
    """Sigma_2 from the stored conic (c/det, -b/det, a/det) = inverse of [[a, b], [b, c]]."""
    q = np.array([[conic[0], conic[1]], [conic[1], conic[2]]])
    return np.linalg.inv(q)


def single_gaussian(position, log_scales, rotation, opacity_logit=0.0, sh=None):
    params = np.zeros((58, 1), np.float32)
    params[0:3, 0] = position
    params[3:6, 0] = log_scales
    params[6:9, 0] = rotation
    params[9, 0] = opacity_logit
    if sh is not None:
        params[10:58, 0] = np.asarray(sh, np.float32).reshape(48)
    return params


def test_projected_covariance_matches_projected_samples(oracle):
    """A small rotated anisotropic Gaussian off the optical axis: Sigma_2 - 0.3 I equals the
    covariance of its samples pushed through the pinhole (the linearisation J W Sigma W^T J^T),
    to sampling accuracy; a transposed rotation is 30-100% off."""
    camera_rotation = Rotation.from_rotvec([0.3, -0.2, 0.1]).as_matrix()
    cam = make_camera(oracle, 200, 180, 300.0, 280.0, 100.0, 90.0, camera_rotation, [0.1, -0.2, 4.0])
    W, t = camera_matrices(cam)
    rotation = np.array([0.5, -1.0, 0.7], np.float32)
    scales = np.array([2e-3, 5e-4, 1e-3], np.float32)
    params = single_gaussian([0.4, 0.3, 0.2], np.log(scales), rotation)
    out = oracle.project_f64(params, cam, 0)
    assert out["visible"][0] == 1
    mean = params[0:3, 0].astype(np.float64)
    p_cam = W @ mean + t
    assert abs(p_cam[0] / p_cam[2]) < 0.65 * cam.width / cam.fx  # Jacobian not clamped
    R = Rotation.from_rotvec(rotation.astype(np.float64)).as_matrix()
    s = np.exp(params[3:6, 0].astype(np.float64))
    sigma3 = sum(s[k] ** 2 * np.outer(R[:, k], R[:, k]) for k in range(3))
    rng = np.random.default_rng(11)
    samples = rng.multivariate_normal(mean, sigma3, size=1_000_000)
    cam_points = samples @ W.T + t
    u = cam.fx * cam_points[:, 0] / cam_points[:, 2] + cam.cx
    v = cam.fy * cam_points[:, 1] / cam_points[:, 2] + cam.cy
    empirical = np.cov(np.stack([u, v]))
    sigma2 = projected_covariance(out["conics"][0]) - 0.3 * np.eye(2)
    scale = np.sqrt(np.outer(np.diag(empirical), np.diag(empirical)))
    assert np.max(np.abs(sigma2 - empirical) / scale) < 0.01
    # The means: the pinhole image of the centre, and its depth.
    np.testing.assert_allclose(out["means"][0], [cam.fx * p_cam[0] / p_cam[2] + cam.cx,
                                                 cam.fy * p_cam[1] / p_cam[2] + cam.cy], rtol=1e-12)
    np.testing.assert_allclose(out["depths"][0], p_cam[2], rtol=1e-12)
    # The test tells a transposed rotation apart.
    sigma3_t = sum(s[k] ** 2 * np.outer(R[k, :], R[k, :]) for k in range(3))
    z = p_cam[2]
    J = np.array([[cam.fx / z, 0.0, -cam.fx * p_cam[0] / z ** 2], [0.0, cam.fy / z, -cam.fy * p_cam[1] / z ** 2]])
    wrong = J @ W @ sigma3_t @ W.T @ J.T
    assert np.max(np.abs(wrong - empirical) / scale) > 0.3


def test_clamped_jacobian_covariance_closed_form(oracle):
    """A large Gaussian whose centre lies outside 1.3x the field of view (x/z = 0.6 > 0.433): the
    Jacobian is evaluated at the clamped x (renderer.cu:164-177); radius, conic and colour against
    closed forms built from scipy rotations and eigenvalues."""
    camera_rotation = Rotation.from_rotvec([0.0, 0.25, -0.1]).as_matrix()
    cam = make_camera(oracle, 200, 180, 300.0, 300.0, 100.0, 90.0, camera_rotation, [0.0, 0.0, 4.0])
    W, t = camera_matrices(cam)
    rotation = np.array([0.2, 0.9, -0.4], np.float32)
    scales = np.array([1.5, 0.4, 0.8], np.float32)
    sh = np.zeros((3, 16), np.float32)
    rng = np.random.default_rng(2)
    sh[:, :] = rng.normal(0.0, 0.3, size=(3, 16))
    # Position chosen so that x_cam / z_cam = 0.6 in the camera frame.
    p_cam_target = np.array([2.4, 0.3, 4.0])
    position = (W.T @ (p_cam_target - t)).astype(np.float32)
    params = single_gaussian(position, np.log(scales), rotation, 0.7, sh)
    out = oracle.project_f64(params, cam, 3)
    assert out["visible"][0] == 1
    p_cam = W @ params[0:3, 0].astype(np.float64) + t
    z = p_cam[2]
    limit_x = 0.65 * cam.width / cam.fx
    limit_y = 0.65 * cam.height / cam.fy
    assert p_cam[0] / z > limit_x
    x = np.clip(p_cam[0] / z, -limit_x, limit_x) * z
    y = np.clip(p_cam[1] / z, -limit_y, limit_y) * z
    J = np.array([[cam.fx / z, 0.0, -cam.fx * x / z ** 2], [0.0, cam.fy / z, -cam.fy * y / z ** 2]])
    R = Rotation.from_rotvec(rotation.astype(np.float64)).as_matrix()
    s = np.exp(params[3:6, 0].astype(np.float64))
    sigma3 = R @ np.diag(s ** 2) @ R.T
    sigma2 = J @ W @ sigma3 @ W.T @ J.T + 0.3 * np.eye(2)
    np.testing.assert_allclose(projected_covariance(out["conics"][0]), sigma2, rtol=1e-10)
    unclamped = np.array([[cam.fx / z, 0.0, -cam.fx * p_cam[0] / z ** 2],
                          [0.0, cam.fy / z, -cam.fy * p_cam[1] / z ** 2]])
    assert np.max(np.abs(unclamped @ W @ sigma3 @ W.T @ unclamped.T + 0.3 * np.eye(2) - sigma2)) > 0.01 * np.max(sigma2)
    # Radius: 3 sigma of the larger eigenvalue, rounded up.
    lam = np.linalg.eigvalsh(sigma2).max()
    radius = np.ceil(3.0 * np.sqrt(lam))
    u = cam.fx * p_cam[0] / z + cam.cx
    v = cam.fy * p_cam[1] / z + cam.cy
    rect = out["tile_rects"][0]
    assert rect[0] == int(np.clip((u - radius) / 16, 0, cam.tile_count_x))
    assert rect[2] == int(np.clip((u + radius + 15) / 16, 0, cam.tile_count_x))
    assert rect[1] == int(np.clip((v - radius) / 16, 0, cam.tile_count_y))
    assert rect[3] == int(np.clip((v + radius + 15) / 16, 0, cam.tile_count_y))
    np.testing.assert_allclose(out["conics"][0][3], 1.0 / (1.0 + np.exp(-0.7)), rtol=1e-7)
    # Colour: view direction from the camera centre (-W^T t) to the mean, SH from scipy.
    direction = params[0:3, 0].astype(np.float64) + W.T @ t
    direction /= np.linalg.norm(direction)
    basis = np.array([scipy_real_sh(l, m, direction) for l in range(4) for m in range(-l, l + 1)])
    colour = np.maximum(0.5 + sh.astype(np.float64) @ basis, 0.0)
    np.testing.assert_allclose(out["colors"][0], colour, rtol=1e-6, atol=1e-7)


def stacked_scene():
    """Four Gaussians on the optical axis of a 16x16 camera (cx = cy = 7, so each projects
    exactly onto pixel (7, 7), where its exponent is 0 and alpha = min(0.99, opacity)). Ids are
    not in depth order. Colours via SH degree 0: colour = 0.5 + C0 * dc."""
    def logit(p):
        return np.log(p / (1.0 - p))
    specs = [  # (depth, opacity, colour); depth order is id 3, 1, 2, 0
        (4.0, 0.99, (0.0, 0.0, 1.0)),     # last: would leave T (1 - 0.99) = 5e-5 < 1e-4, not blended
        (2.0, 0.999, (1.0, 0.0, 0.0)),    # second: alpha capped at 0.99
        (3.0, 0.5, (0.0, 1.0, 0.0)),      # third
        (1.5, 0.003, (0.3, 0.3, 0.3)),    # nearest: alpha 0.003 < 1/255, skipped
    ]
    params = np.zeros((58, len(specs)), np.float32)
    for i, (depth, opacity, colour) in enumerate(specs):
        params[0:3, i] = (0.0, 0.0, depth)
        params[3:6, i] = np.log(0.05)
        params[9, i] = logit(opacity)
        for c in range(3):
            params[10 + 16 * c, i] = (colour[c] - 0.5) / SH_C0
    return params


def test_compositing_of_stacked_gaussians_hand_computed(oracle):
    cam = make_camera(oracle, 16, 16, 10.0, 10.0, 7.0, 7.0)
    background = (0.2, 0.4, 0.6)
    params = stacked_scene()
    image = oracle.render_reference(params, cam, background, 0)
    # Front to back: skip the 0.003 splat; red at alpha 0.99 (T -> 0.01); green at 0.5
    # (T -> 0.005); blue would leave T = 5e-5 < 1e-4, so the pixel is done without it.
    want = np.array([0.99 + 0.005 * 0.2, 0.005 + 0.005 * 0.4, 0.0 + 0.005 * 0.6])
    np.testing.assert_allclose(image[7, 7], want, rtol=1e-6, atol=1e-9)
    # A corner pixel, far outside every splat, shows the background only.
    np.testing.assert_allclose(image[0, 15], background, atol=1e-6)


def test_compositing_order_does_not_follow_ids(oracle):
    cam = make_camera(oracle, 16, 16, 10.0, 10.0, 7.0, 7.0)
    params = stacked_scene()
    permuted = params[:, [2, 0, 3, 1]].copy()
    a = oracle.render_reference(params, cam, (0.2, 0.4, 0.6), 0)
    b = oracle.render_reference(permuted, cam, (0.2, 0.4, 0.6), 0)
    np.testing.assert_array_equal(a, b)
