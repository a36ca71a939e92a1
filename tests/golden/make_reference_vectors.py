"""Writes reference_vectors.json: literal known answers from the reference's own tests.

The reference (Naive-q/GS-native) is C++/CUDA and is not run here; each value below is a literal
constant from the cited test or source line, so the fixture can be regenerated from this
script alone.
"""
import json
import pathlib

VECTORS = {
    # solver.cu:14-24, the 11-tap SSIM window as FP32 literals.
    "ssim_window": [1.028380120e-03, 7.598758210e-03, 3.600077330e-02, 1.093606870e-01, 2.130055430e-01,
                    2.660117150e-01, 2.130055430e-01, 1.093606870e-01, 3.600077330e-02, 7.598758210e-03,
                    1.028380120e-03],
    # renderer_test.cu:72-88: lower colour clamp only; expected 0 and 1.35 at the centre pixel.
    "color_clamp": {"red": 0.0, "green": 1.35, "tolerance": 2e-6},
    # renderer_test.cu:26 / solver_test.cu:176, 292, 354.
    "render_linf": 3e-5,
    "statistic_fd_tolerance": {"absolute": 2e-5, "relative": 0.002},
    # solver_test.cu:218-219, 234-235: first Adam step after zeroed history at step 2.
    "adam_reset_step": {"learning_rate": 0.025, "tolerance": 2e-6},
    # density_test.cu, threshold 0.15 unless stated, scene extent 1, percent dense 0.01,
    # min opacity 0.005, seed 42.
    "density": {
        "low_opacity": {"parents": [0, 2], "changes": [0, 1]},
        "clone_then_clear": {"parents": [0, 1, 2, -1], "second_changes": 0, "count": 4},
        "split": {"history": [-1, -1], "parent_ids": [0, 0], "child_scale": 0.125},
        "round_mean": {"thresholds": [0.19, 0.21], "counts": [2, 1], "changes": [1, 0]},
        "oversize": {"parents": [1]},
        "opacity_caps": [0.8, 0.01],
        "all_pruned": {"count": 0},
        "signed_vs_absolute": {"count": 10, "parent_ids": [0, 1, 2, 5, 0, 5, 3, 4, 3, 4],
                               "history": [0, 1, 2, 5, -1, -1, -1, -1, -1, -1]},
        "absolute_mean": {"thresholds": [0.00084, 0.00086], "counts": [2, 1]},
    },
    # knn_test.cu:44-53: order invariance points.
    "knn_order_points": [[100, 0, 0], [0, 0, 0], [1, 0, 0], [3, 0, 0], [7, 0, 0]],
    # knn_test.cu:8-42: exhaustive comparison sizes, seed and tolerance.
    "knn_sizes": [0, 1, 2, 3, 4, 17, 256, 1025],
    "knn_relative": 2e-6,
    # gaussian_splatting_test.cpp:52: 600 iterations on the 4-point scene -> 8 Gaussians.
    "integration_density_count": 8,
}

if __name__ == "__main__":
    path = pathlib.Path(__file__).with_name("reference_vectors.json")
    path.write_text(json.dumps(VECTORS, indent=1) + "\n")
    print(path)
