"""Learned densification: collect outcome-labelled samples from plain-rule training runs, train the
16 -> 64 -> 64 -> 3 proposal network on the GPU (MFMA kernels), and compare it with the plain rule
on unseen scenes (held-out PSNR, wall time to equal PSNR, Gaussian count).

Scenes are seeded mixtures of 64 anisotropic blobs: a ground-truth scene of `truth` Gaussians
renders the targets; training starts from `init` points sampled from the same mixture with the
reference initialisation (gaussian_splatting.cpp:95-128: KNN log-scales, opacity 0.1, SH DC from the
point colour). Training follows the reference schedule (density control every 100 iterations after
500, opacity resets) with a Gaussian-count cap. The plain rule stays the arbiter: the learned policy is
accepted only if its final held-out PSNR is within `tolerance` dB of the rule's.

Usage: python gs-native_amd/densify_experiment.py [--quick] [--out result.json]
Prints one progress line per phase step and a JSON summary.
"""
from __future__ import annotations

import argparse
import importlib.util
import json
import math
import pathlib
import sys
import time

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent


def _load_package():
    if "gs_native_amd" in sys.modules:
        return sys.modules["gs_native_amd"]
    spec = importlib.util.spec_from_file_location("gs_native_amd", HERE / "__init__.py",
                                                  submodule_search_locations=[str(HERE)])
    module = importlib.util.module_from_spec(spec)
    sys.modules["gs_native_amd"] = module
    spec.loader.exec_module(module)
    return module


gsn = _load_package()
import gs_native_amd.scenes as scenes  # noqa: E402
import gs_native_amd.trainer as trainer_module  # noqa: E402

PARAMETERS = 5443


def log(message: str) -> None:
    print(f"[densify] {message}", flush=True)


class Scene:
    def __init__(self, context, seed: int, truth_count: int, init_count: int, width: int, height: int,
                 cameras: int, heldout: int, colour_noise: float = 0.05):
        self.seed = seed
        truth = scenes.blob_scene(truth_count, seed=seed, sh_degree=3, colour_noise=colour_noise)
        _, d2 = gsn.knn3(context, truth[0:3].T.copy())
        truth[3:6] = scenes.scales_from_mean_dist2(d2)[None, :]
        points = scenes.blob_scene(init_count, seed=seed, sh_degree=0, colour_noise=colour_noise)
        init = np.zeros((58, init_count), np.float32)
        init[0:3] = points[0:3]
        _, d2 = gsn.knn3(context, points[0:3].T.copy())
        init[3:6] = scenes.scales_from_mean_dist2(d2)[None, :]
        init[9] = np.float32(math.log(0.1 / 0.9))
        for c in range(3):
            init[10 + 16 * c] = points[10 + 16 * c]
        self.init = init
        focal = scenes.fov_focal(width, 60.0)
        make = lambda poses: [gsn.Camera.make(width, height, focal, focal, 0.5 * (width - 1), 0.5 * (height - 1), R, t)
                              for R, t in poses]
        self.cameras = make(scenes.sphere_cameras(cameras, 5.0, seed + 1000))
        self.heldout_cameras = make(scenes.sphere_cameras(heldout, 5.0, seed + 2000))
        centres = np.stack([np.asarray(c.R_c_w[:], np.float32).reshape(3, 3).T @ -np.asarray(c.t_c_w[:], np.float32)
                            for c in self.cameras])
        self.extent = 1.1 * float(np.linalg.norm(centres - centres.mean(0), axis=1).max())
        device_truth = gsn.DeviceBuffer.from_numpy(truth)
        image = gsn.DeviceBuffer((height, width, 3), np.float32)
        projection = None
        self.targets, self.heldout_targets = [], []
        for cams, out in ((self.cameras, self.targets), (self.heldout_cameras, self.heldout_targets)):
            for cam in cams:
                projection = gsn.render(context, device_truth, truth_count, cam, (0.0, 0.0, 0.0), image, 3, projection)
                target = np.clip(np.rint(image.numpy() * 255.0), 0, 255).astype(np.uint8)
                out.append(gsn.DeviceBuffer.from_numpy(target))
        self.image = image

    def heldout_psnr(self, context, trainer) -> float:
        values = []
        projection = None
        for cam, target in zip(self.heldout_cameras, self.heldout_targets):
            projection = gsn.render(context, trainer.params, trainer.count, cam, (0.0, 0.0, 0.0), self.image,
                                    trainer.sh_degree, projection)
            values.append(gsn.image_metrics(context, cam.width, cam.height, self.image, target)["psnr"])
        return float(np.mean(values))


def run(context, scene: Scene, iterations: int, cap: int, policy=None, outcome_log=None, eval_every: int = 0,
        useful_opacity: float = 0.02):
    """Trains from the scene's initial points; returns the curve [(iteration, seconds, count, psnr)]
    with training wall time only (evaluations excluded)."""
    trainer = trainer_module.Trainer(context, scene.init, scene.cameras, scene.targets, scene.extent,
                                     max_count=cap, seed=scene.seed, policy=policy, outcome_log=outcome_log,
                                     useful_opacity=useful_opacity)
    curve = []
    elapsed = 0.0
    context.synchronize()
    start = time.perf_counter()
    for it in range(1, iterations + 1):
        trainer.iteration()
        if eval_every and (it % eval_every == 0 or it == iterations):
            context.synchronize()
            elapsed += time.perf_counter() - start
            curve.append((it, elapsed, trainer.count, scene.heldout_psnr(context, trainer)))
            context.synchronize()
            start = time.perf_counter()
    context.synchronize()
    elapsed += time.perf_counter() - start
    if not eval_every:
        curve.append((iterations, elapsed, trainer.count, float("nan")))
    return trainer, curve


def train_network(context, features, labels, weights, epochs: int, batch: int, learning_rate: float, seed: int,
                  calibrate: bool = True):
    """Standardises the features, trains on the GPU, folds the standardisation into W1/b1 and
    returns (params, report)."""
    rng = np.random.default_rng(seed)
    order = rng.permutation(features.shape[0])
    features, labels, weights = features[order], labels[order], weights[order]
    validation = max(1, features.shape[0] // 20)
    mean = features[validation:].mean(axis=0)
    std = features[validation:].std(axis=0) + 1e-3
    standardised = ((features - mean) / std).astype(np.float32)
    # The sampling weights keep the loss unbiased; no class rebalancing, so the argmax densifies a
    # Gaussian only where a useful outcome is more likely than not.
    balanced = (weights / weights[validation:].mean()).astype(np.float32)
    params = np.zeros(PARAMETERS, np.float32)
    offset = 0
    for rows, cols in ((64, 16), (64, 0), (64, 64), (64, 0), (3, 64), (3, 0)):
        size = rows * max(cols, 1)
        if cols:
            bound = math.sqrt(6.0 / cols)
            params[offset:offset + size] = rng.uniform(-bound, bound, size)
        offset += size
    network = gsn.ProposalNetwork(context, params)
    x = gsn.DeviceBuffer.from_numpy(standardised[validation:])
    y = gsn.DeviceBuffer.from_numpy(labels[validation:].astype(np.int32))
    w = gsn.DeviceBuffer.from_numpy(balanced[validation:])
    n = x.shape[0]
    batch = min(batch, n)  # small collections: one batch per epoch
    steps = 0
    losses = []
    for epoch in range(epochs):
        for begin in range(0, n - batch + 1, batch):
            want = begin == 0
            loss = network.train_step(x.ptr.value + begin * 64, y.ptr.value + begin * 4, batch, learning_rate,
                                      w.ptr.value + begin * 4, want_loss=want)
            steps += 1
            if want:
                losses.append(loss)
        log(f"epoch {epoch + 1}/{epochs}: loss {losses[-1]:.4f}")
    trained = network.params.numpy()
    # fold x' = (x - mean) / std into the first layer: W1' = W1 / std, b1' = b1 - W1' mean
    w1 = trained[:1024].reshape(64, 16) / std[None, :]
    b1 = trained[1024:1088] - w1 @ mean
    folded = trained.copy()
    folded[:1024] = w1.reshape(-1)
    folded[1024:1088] = b1
    deployed = gsn.ProposalNetwork(context, folded.astype(np.float32))
    vx = gsn.DeviceBuffer.from_numpy(np.ascontiguousarray(features[:validation], np.float32))
    logits = gsn.DeviceBuffer((validation, 3), np.float32)
    deployed.forward(vx, validation, logits)
    validation_logits = logits.numpy().astype(np.float64)
    truth = labels[:validation]
    biases = np.zeros(3)
    if calibrate:
        biases = calibrate_biases(validation_logits, truth, weights[:validation])
        folded[-3:] += biases.astype(np.float32)  # b3 is the last layer's bias
    predicted = np.argmax(validation_logits + biases[None, :], axis=1)
    confusion = [[int(np.sum((truth == a) & (predicted == b))) for b in range(3)] for a in range(3)]
    report = {"samples": int(features.shape[0]), "validation": int(validation), "steps": steps,
              "losses": [float(v) for v in losses], "validation_accuracy": float(np.mean(predicted == truth)),
              "validation_confusion": confusion, "output_biases": [float(v) for v in biases],
              "label_counts": [int(np.sum(labels == c)) for c in range(3)]}
    return folded.astype(np.float32), report


def calibrate_biases(logits: np.ndarray, labels: np.ndarray, weights: np.ndarray) -> np.ndarray:
    """Output biases (keep, clone, split) = (0, d_clone, d_split) chosen on the validation set so
    that the network proposes clones and splits as often as the (sample-weighted) labels hold
    them. Trained on rare densify labels, the plain argmax under-proposes them."""
    total = weights.sum()
    target = {c: float(weights[labels == c].sum() / total) for c in (1, 2)}
    biases = np.zeros(3)
    grid = np.linspace(-8.0, 8.0, 321)
    for _ in range(3):
        for c in (2, 1):
            best, best_error = 0.0, np.inf
            for value in grid:
                trial = biases.copy()
                trial[c] = value
                predicted = np.argmax(logits + trial[None, :], axis=1)
                error = abs(float(weights[predicted == c].sum() / total) - target[c])
                if error < best_error - 1e-12:
                    best, best_error = value, error
            biases[c] = best
    return biases


def time_to_psnr(curve, target):
    for it, seconds, count, psnr in curve:
        if psnr >= target:
            return it, seconds, count
    return None


def main() -> int:
    parser = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    parser.add_argument("--train-seeds", type=int, default=10)
    parser.add_argument("--eval-seeds", type=int, default=3)
    parser.add_argument("--collect-truth", type=int, default=500_000)
    parser.add_argument("--collect-init", type=int, default=50_000)
    parser.add_argument("--collect-size", type=str, default="1248x832")
    parser.add_argument("--collect-cameras", type=int, default=150)
    parser.add_argument("--collect-iterations", type=int, default=3000)
    parser.add_argument("--eval-truth", type=int, default=3_000_000)
    parser.add_argument("--eval-init", type=int, default=200_000)
    parser.add_argument("--eval-size", type=str, default="1920x1080")
    parser.add_argument("--eval-cameras", type=int, default=300)
    parser.add_argument("--eval-heldout", type=int, default=8)
    parser.add_argument("--eval-iterations", type=int, default=7000)
    parser.add_argument("--eval-every", type=int, default=500)
    parser.add_argument("--cap", type=int, default=6_000_000)
    parser.add_argument("--keep-fraction", type=float, default=0.1)
    parser.add_argument("--epochs", type=int, default=4)
    parser.add_argument("--batch", type=int, default=65536)
    parser.add_argument("--learning-rate", type=float, default=2e-3)
    parser.add_argument("--tolerance", type=float, default=0.05)
    parser.add_argument("--useful-opacity", type=float, default=0.02,
                        help="a densified Gaussian's outputs count as useful when observed in the next interval "
                             "and at least this opaque")
    parser.add_argument("--colour-noise", type=float, default=0.05,
                        help="per-Gaussian colour noise of the truth scenes (their fine texture)")
    parser.add_argument("--no-calibration", action="store_true",
                        help="keep the trained output biases (plain argmax) instead of matching the label rates")
    parser.add_argument("--quick", action="store_true", help="small sizes for a functional check")
    parser.add_argument("--out", type=str, default="")
    args = parser.parse_args()
    if args.quick:
        args.train_seeds, args.eval_seeds = 2, 1
        args.collect_truth, args.collect_init, args.collect_size, args.collect_cameras = 100_000, 10_000, "320x240", 32
        args.collect_iterations = 1200
        args.eval_truth, args.eval_init, args.eval_size, args.eval_cameras = 100_000, 10_000, "320x240", 32
        args.eval_iterations, args.eval_every, args.eval_heldout = 1200, 300, 4
        args.epochs, args.batch = 2, 4096
    context = gsn.Context()
    log(f"device {gsn.device_name()}")
    cw, ch = (int(v) for v in args.collect_size.split("x"))
    ew, eh = (int(v) for v in args.eval_size.split("x"))

    outcome_log = trainer_module.OutcomeLog(keep_fraction=args.keep_fraction, seed=1)
    collect_runs = []
    for seed in range(args.train_seeds):
        scene = Scene(context, 100 + seed, args.collect_truth, args.collect_init, cw, ch, args.collect_cameras, 1,
                      args.colour_noise)
        trainer, curve = run(context, scene, args.collect_iterations, args.cap, outcome_log=outcome_log,
                             useful_opacity=args.useful_opacity)
        samples = sum(f.shape[0] for f in outcome_log.features)
        collect_runs.append({"seed": 100 + seed, "final_count": trainer.count, "seconds": curve[-1][1]})
        log(f"collect seed {100 + seed}: {trainer.count} Gaussians, {curve[-1][1]:.1f} s, {samples} samples so far")
        del scene, trainer
    features, labels, weights = outcome_log.arrays()
    represented = float(weights.sum())
    log(f"samples: {features.shape[0]} stored, {represented:.0f} represented, labels {np.bincount(labels, minlength=3)}")
    network_params, training = train_network(context, features, labels, weights, args.epochs, args.batch,
                                              args.learning_rate, seed=7, calibrate=not args.no_calibration)
    training["represented_samples"] = represented
    log(f"trained: validation accuracy {training['validation_accuracy']:.4f}")
    del features, labels, weights, outcome_log
    network = gsn.ProposalNetwork(context, network_params)

    evaluations = []
    for k in range(args.eval_seeds):
        seed = 900 + k
        scene = Scene(context, seed, args.eval_truth, args.eval_init, ew, eh, args.eval_cameras, args.eval_heldout,
                      args.colour_noise)
        results = {}
        for name, policy in (("rule", None), ("learned", network)):
            trainer, curve = run(context, scene, args.eval_iterations, args.cap, policy=policy,
                                 eval_every=args.eval_every)
            results[name] = {"curve": curve, "final_count": trainer.count, "final_psnr": curve[-1][3],
                             "seconds": curve[-1][1], "policy_fallbacks": trainer.policy_fallbacks}
            log(f"eval seed {seed} {name}: PSNR {curve[-1][3]:.3f} dB, {trainer.count} Gaussians, {curve[-1][1]:.1f} s")
            del trainer
        rule, learned = results["rule"], results["learned"]
        reached = time_to_psnr(learned["curve"], rule["final_psnr"] - args.tolerance)
        entry = {
            "seed": seed,
            "rule": rule,
            "learned": learned,
            "accepted": learned["final_psnr"] >= rule["final_psnr"] - args.tolerance,
            "psnr_delta_db": learned["final_psnr"] - rule["final_psnr"],
            "count_reduction": 1.0 - learned["final_count"] / max(rule["final_count"], 1),
            "time_to_rule_psnr": None if reached is None else {
                "iteration": reached[0], "seconds": reached[1], "count": reached[2],
                "time_reduction": 1.0 - reached[1] / max(rule["seconds"], 1e-9),
                "count_reduction": 1.0 - reached[2] / max(rule["final_count"], 1)},
        }
        evaluations.append(entry)
        del scene
    summary = {
        "config": {k: v for k, v in vars(args).items() if k != "out"},
        "collect": collect_runs,
        "training": training,
        "evaluations": evaluations,
        "accepted_all": all(e["accepted"] for e in evaluations),
    }
    text = json.dumps(summary)
    if args.out:
        pathlib.Path(args.out).write_text(json.dumps(summary, indent=1))
        np.save(pathlib.Path(args.out).with_suffix(".params.npy"), network_params)
    print(text, flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
