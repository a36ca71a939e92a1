"""File formats on the CPU: the COLMAP loader (reference test/io_test.cpp:20-70 fixture and known
answers), the PNG / PPM decoders and the area downscale, and the Gaussian PLY round trip. The
fixture images are written here with Pillow, an encoder independent of the C++ codec under test;
the expected area downscale is computed from its definition with integer arithmetic. tests/cpp/io_test (built by build()) runs the
checks on the C++ side."""
import pathlib
import struct
import subprocess

import numpy as np
import pytest

PIL = pytest.importorskip("PIL.Image")

REPO = pathlib.Path(__file__).resolve().parents[1]
BINARY = REPO / "tests" / "cpp" / "io_test"


def write_colmap_fixture(root: pathlib.Path, truncate_images: bool = False) -> None:
    """The fixture of the reference's io_test.cpp:20-50: one PINHOLE camera (19x17, f 20/21,
    c 9/8), nine identity-pose images frame08..frame00 (ids 9..1) of the constant colour BGR
    (10, 20, 30), and four points."""
    sparse = root / "sparse" / "0"
    sparse.mkdir(parents=True)
    (root / "images").mkdir()
    (sparse / "cameras.bin").write_bytes(struct.pack("<QiiQQ4d", 1, 1, 1, 19, 17, 20.0, 21.0, 9.0, 8.0))
    images = bytearray(struct.pack("<Q", 9))
    for i in range(8, -1, -1):
        name = f"frame{i:02d}.png"
        images += struct.pack("<i7di", i + 1, 1.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 1)
        images += name.encode() + b"\0"
        images += struct.pack("<Q2dq", 1, 0.0, 0.0, 1)
        pixels = np.zeros((17, 19, 3), np.uint8)
        pixels[:] = (30, 20, 10)  # RGB of the BGR (10, 20, 30) fixture colour
        PIL.fromarray(pixels, "RGB").save(root / "images" / name)
    if truncate_images:
        images = images[:-5]
    (sparse / "images.bin").write_bytes(bytes(images))
    points = bytearray(struct.pack("<Q", 4))
    for i in range(4):
        points += struct.pack("<Q3d3BdQ", i + 1, 0.1 * (i % 2), 0.1 * (i // 2), 3.0, 100, 150, 200, 0.0, 0)
    (sparse / "points3D.bin").write_bytes(bytes(points))


def write_codec_fixture(root: pathlib.Path) -> None:
    codec = root / "codec"
    codec.mkdir()
    rng = np.random.default_rng(3)
    rgb = rng.integers(0, 256, size=(23, 37, 3), dtype=np.uint8)
    PIL.fromarray(rgb, "RGB").save(codec / "rgb.png")
    (codec / "rgb.png.rgb").write_bytes(rgb.tobytes())
    rgba = rng.integers(0, 256, size=(11, 13, 4), dtype=np.uint8)
    PIL.fromarray(rgba, "RGBA").save(codec / "rgba.png")
    (codec / "rgba.png.rgb").write_bytes(np.ascontiguousarray(rgba[..., :3]).tobytes())
    grey = rng.integers(0, 256, size=(9, 14), dtype=np.uint8)
    PIL.fromarray(grey, "L").save(codec / "grey.png")
    (codec / "grey.png.rgb").write_bytes(np.repeat(grey[..., None], 3, axis=2).tobytes())
    # A smooth image: Pillow's encoder picks the Sub / Up / Average / Paeth filters for it.
    y, x = np.mgrid[0:40, 0:50]
    smooth = np.stack([x * 5, y * 6, (x + y) * 2], axis=2).astype(np.uint8)
    PIL.fromarray(smooth, "RGB").save(codec / "filtered.png", optimize=True)
    (codec / "filtered.png.rgb").write_bytes(smooth.tobytes())
    ppm = rng.integers(0, 256, size=(5, 7, 3), dtype=np.uint8)
    (codec / "plain.ppm").write_bytes(b"P6\n# comment\n7 5\n255\n" + ppm.tobytes())
    (codec / "plain.ppm.rgb").write_bytes(ppm.tobytes())
    PIL.fromarray(rgb, "RGB").save(codec / "not_supported.jpg")
    source = rng.integers(0, 256, size=(47, 61, 3), dtype=np.uint8)
    PIL.fromarray(source, "RGB").save(codec / "area_source.png")
    (codec / "area_expected.rgb").write_bytes(area_downscale(source, 20, 16).tobytes())


def area_downscale(source: np.ndarray, width: int, height: int) -> np.ndarray:
    """Area averaging by exact integer arithmetic: each source pixel is replicated into a
    height x width grid of sub-pixels, so that every output pixel covers exactly h x w of them
    (h, w = the source size); their mean is rounded half to even."""
    h, w, _ = source.shape
    fine = np.repeat(np.repeat(source.astype(np.int64), height, axis=0), width, axis=1)
    sums = fine.reshape(height, h, width, w, 3).sum(axis=(1, 3))
    return np.rint(sums / float(h * w)).astype(np.uint8)


def test_file_formats(tmp_path):
    if not BINARY.exists():
        pytest.fail(f"{BINARY} is not built: run __graft_entry__.build()")
    write_colmap_fixture(tmp_path / "colmap")
    write_colmap_fixture(tmp_path / "colmap_truncated", truncate_images=True)
    write_codec_fixture(tmp_path)
    result = subprocess.run([str(BINARY), str(tmp_path)], capture_output=True, text=True, timeout=120)
    print(result.stdout)
    assert result.returncode == 0, result.stdout + result.stderr
    assert "0 failures" in result.stdout
