"""The C-ABI library loads and exports every entry point declared in include/gsn_hip.h (no GPU
needed: nothing here launches a kernel)."""
import ctypes
import pathlib
import re

REPO = pathlib.Path(__file__).resolve().parents[1]


def declared_symbols():
    text = (REPO / "include" / "gsn_hip.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(gsn_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_path():
    names = declared_symbols()
    for required in ("gsn_project_gaussians", "gsn_build_tile_lists", "gsn_render_tiles", "gsn_image_gradients",
                     "gsn_gaussian_gradients", "gsn_gaussian_gradients_adam", "gsn_adam_update", "gsn_remap_history",
                     "gsn_density_count", "gsn_density_write", "gsn_density_accumulate", "gsn_knn3"):
        assert required in names


def test_library_exports_every_declared_symbol(gsn):
    library = ctypes.CDLL(str(gsn.LIBRARY_PATH))
    missing = [name for name in declared_symbols() if not hasattr(library, name)]
    assert not missing, f"missing exports: {missing}"


def test_binding_prototypes_cover_the_header(gsn):
    missing = [name for name in declared_symbols() if name not in gsn.exported_symbols()]
    assert not missing, f"no ctypes prototype for: {missing}"


def test_status_strings_and_argument_errors(gsn):
    assert gsn.lib.gsn_status_string(0) == b"ok"
    assert gsn.lib.gsn_status_string(1) == b"invalid argument"
    # A null context is rejected before any device work.
    status = gsn.lib.gsn_render_tiles(None, None, None, None, None, None, None, None, None)
    assert status == 1
    assert b"gsn_render_tiles" in gsn.lib.gsn_last_error()


def test_camera_layout_matches_reference_struct(gsn):
    # CameraParams: 4 floats, R[9], t[3], near, far, int w, h, int64 pixel_offset, 3 ints.
    assert ctypes.sizeof(gsn.Camera) == 104
    assert gsn.Camera.pixel_offset.offset == 80
    assert gsn.Camera.tile_count_x.offset == 88
