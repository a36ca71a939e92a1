"""Held-out evaluation (reference demo/render.cpp:107-155) on the GPU:

- gsn_image_quantized_error against NumPy: clamp to [0, 1], x 255 in FP32, round half to even,
  sum of squared 8-bit differences; quantised bytes and sums bit-exact, including exact ties;
- the render program end to end on a COLMAP scene written here (binary sparse model, PNG images
  rendered on the GPU from a ground truth, a PLY model written by an independent writer): its
  psnr.csv lists every held-out view in the reference's format, and each value equals the PSNR
  computed here from the same quantised render.
"""
import csv
import pathlib
import struct
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

REPO = pathlib.Path(__file__).resolve().parents[1]
PROGRAM = REPO / "gs-native_amd" / "demo" / "gs_native_render"


def quantise(image: np.ndarray) -> np.ndarray:
    return np.rint(np.clip(image.astype(np.float32), 0.0, 1.0) * np.float32(255.0)).astype(np.uint8)


def psnr(quantised: np.ndarray, target: np.ndarray) -> tuple[int, float]:
    sse = int(np.sum((quantised.astype(np.int64) - target.astype(np.int64)) ** 2))
    mse = sse / (3.0 * target.shape[0] * target.shape[1] * 255.0 * 255.0)
    return sse, float("inf") if sse == 0 else -10.0 * np.log10(mse)


def test_quantised_error_matches_numpy(gsn, context):
    rng = np.random.default_rng(9)
    height, width = 37, 53
    image = rng.uniform(-0.2, 1.2, size=(height, width, 3)).astype(np.float32)
    # Exact ties: (k + 0.5) / 255 scaled back by 255 lands on k + 0.5 for many k.
    ties = ((np.arange(width * 3) % 255 + 0.5) / 255.0).astype(np.float32).reshape(width, 3)
    image[0] = ties
    image[1] = np.float32(np.nan)
    target = rng.integers(0, 256, size=(height, width, 3), dtype=np.uint8)
    quantised = gsn.DeviceBuffer((height, width, 3), np.uint8)
    sse, value = gsn.quantized_error(context, width, height, gsn.DeviceBuffer.from_numpy(image),
                                     gsn.DeviceBuffer.from_numpy(target), quantised)
    expected = quantise(np.nan_to_num(image, nan=0.0))
    np.testing.assert_array_equal(quantised.numpy(), expected)
    want_sse, want_psnr = psnr(expected, target)
    assert sse == want_sse
    assert value == pytest.approx(want_psnr, rel=1e-12)
    # Identical images: infinite PSNR.
    same, inf = gsn.quantized_error(context, width, height, gsn.DeviceBuffer.from_numpy(expected / np.float32(255.0)),
                                    gsn.DeviceBuffer.from_numpy(expected))
    assert same == 0 and inf == float("inf")


def _quaternion(rotation: np.ndarray) -> tuple[float, float, float, float]:
    from scipy.spatial.transform import Rotation
    x, y, z, w = Rotation.from_matrix(rotation.astype(np.float64)).as_quat()
    return w, x, y, z


def _write_ply(path: pathlib.Path, params: np.ndarray) -> None:
    """Binary little-endian Gaussian PLY: x y z, nx ny nz, f_dc_0..2, f_rest_0..44 (channel-major),
    opacity, scale_0..2, rot_0..3 (w x y z) from the axis-angle rows."""
    from scipy.spatial.transform import Rotation
    n = params.shape[1]
    names = ["x", "y", "z", "nx", "ny", "nz"] + [f"f_dc_{i}" for i in range(3)] + \
            [f"f_rest_{i}" for i in range(45)] + ["opacity"] + [f"scale_{i}" for i in range(3)] + \
            [f"rot_{i}" for i in range(4)]
    header = "ply\nformat binary_little_endian 1.0\n" + f"element vertex {n}\n" + \
             "".join(f"property float {name}\n" for name in names) + "end_header\n"
    columns = np.zeros((n, 62), np.float32)
    columns[:, 0:3] = params[0:3].T
    for c in range(3):
        columns[:, 6 + c] = params[10 + 16 * c]
        columns[:, 9 + 15 * c: 24 + 15 * c] = params[11 + 16 * c: 26 + 16 * c].T
    columns[:, 54] = params[9]
    columns[:, 55:58] = params[3:6].T
    quat = Rotation.from_rotvec(params[6:9].T.astype(np.float64)).as_quat()  # x y z w
    columns[:, 58] = quat[:, 3]
    columns[:, 59:62] = quat[:, 0:3]
    path.write_bytes(header.encode() + columns.tobytes())


def test_render_program_scores_held_out_views(gsn, context, scenes, tmp_path):
    PIL = pytest.importorskip("PIL.Image")
    assert PROGRAM.exists(), f"{PROGRAM} is not built: run __graft_entry__.build()"
    width, height, focal = 160, 120, 140.0
    truth = scenes.config0_params(seed=41, count=3000, sh_degree=3)
    truth[3:6] += np.float32(1.0)
    model = scenes.perturb(truth, 42, 0.02)
    poses = scenes.sphere_cameras(17, 3.0, 43)
    names = [f"view{i:03d}.png" for i in range(len(poses))]
    sparse = tmp_path / "scene" / "sparse" / "0"
    sparse.mkdir(parents=True)
    (tmp_path / "scene" / "images").mkdir()
    # The loader shifts the principal point by -0.5: the file holds cx + 0.5.
    (sparse / "cameras.bin").write_bytes(struct.pack("<QiiQQ4d", 1, 1, 1, width, height, focal, focal,
                                                     0.5 * (width - 1) + 0.5, 0.5 * (height - 1) + 0.5))
    images = bytearray(struct.pack("<Q", len(poses)))
    truth_device = gsn.DeviceBuffer.from_numpy(truth)
    image = gsn.DeviceBuffer((height, width, 3), np.float32)
    cams = []
    for i, ((R, t), name) in enumerate(zip(poses, names)):
        images += struct.pack("<i4d3di", i + 1, *_quaternion(R), *[float(v) for v in t], 1)
        images += name.encode() + b"\0" + struct.pack("<Q", 0)
        cam = gsn.Camera.make(width, height, focal, focal, 0.5 * (width - 1), 0.5 * (height - 1), R, t)
        gsn.render(context, truth_device, truth.shape[1], cam, (0.0, 0.0, 0.0), image, 3)
        PIL.fromarray(quantise(image.numpy()), "RGB").save(tmp_path / "scene" / "images" / name)
        cams.append(cam)
    (sparse / "images.bin").write_bytes(bytes(images))
    points = bytearray(struct.pack("<Q", 4))
    for k in range(4):
        points += struct.pack("<Q3d3BdQ", k + 1, 0.1 * k, 0.0, 0.0, 100, 150, 200, 0.0, 0)
    (sparse / "points3D.bin").write_bytes(bytes(points))
    _write_ply(tmp_path / "model.ply", model)
    result = subprocess.run([str(PROGRAM), f"--scene={tmp_path / 'scene'}", f"--model={tmp_path / 'model.ply'}",
                             f"--output={tmp_path / 'out'}", "--split=test"], capture_output=True, text=True,
                            timeout=120)
    print(result.stdout)
    assert result.returncode == 0, result.stdout + result.stderr
    destination = tmp_path / "out" / "test" / "gs-native"
    rows = list(csv.reader((destination / "psnr.csv").open()))
    assert rows[0] == ["image", "psnr_db"]
    held_out = list(range(0, len(poses), 8))  # every 8th view in name order
    assert [row[0] for row in rows[1:-1]] == [f"{v:05d}.png" for v in range(len(held_out))]
    assert rows[-1][0] == "mean"
    # The model's PLY goes through a quaternion: render the same rows the program read back.
    model_device = gsn.DeviceBuffer.from_numpy(model)
    values = []
    for row, view in zip(rows[1:-1], held_out):
        gsn.render(context, model_device, model.shape[1], cams[view], (0.0, 0.0, 0.0), image, 3)
        target = np.asarray(PIL.open(tmp_path / "scene" / "images" / names[view]).convert("RGB"))
        _, expected = psnr(quantise(image.numpy()), target)
        values.append(float(row[1]))
        assert float(row[1]) == pytest.approx(expected, abs=0.02), (view, row, expected)
        rendered = np.asarray(PIL.open(destination / "renders" / row[0]).convert("RGB"))
        assert rendered.shape == (height, width, 3)
        np.testing.assert_array_equal(np.asarray(PIL.open(destination / "gt" / row[0]).convert("RGB")), target)
    assert float(rows[-1][1]) == pytest.approx(np.mean(values), rel=1e-9)
    assert 15.0 < np.mean(values) < 60.0
