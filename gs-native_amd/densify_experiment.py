"""Learned densification: collect outcome-labelled samples from plain-rule training runs, train the
16 -> 64 -> 64 -> 3 proposal network on the GPU (MFMA kernels), and compare it with the plain rule
on unseen scenes (held-out PSNR, wall time to equal PSNR, Gaussian count).

Scenes (seeded): a ground-truth scene of `truth` Gaussians renders the targets, by default a
textured radius-2 shell plus ground plane seen from a radius-4 hemisphere of cameras (the
shell+plane workload, with a seeded colour pattern finer than the initial points resolve, so the
rule densifies: 200,000 points grow to about 570,000 Gaussians by iteration 4000 at 1920x1080), or
the 64-blob mixture seen from a sphere. Training starts from `init` points sampled from the same
scene with the reference initialisation (gaussian_splatting.cpp:95-128: KNN log-scales, opacity
0.1, SH DC from the point colour) and follows the reference schedule (density control every 100
iterations after 500, opacity resets) with a Gaussian-count cap.

Labels: at every density step of the plain rule each Gaussian's 16 features are logged with the
outcome of the rule's decision one interval later (gsn_density_outcomes), one label column per
usefulness threshold (a clone or split counts only if all of its outputs are observed and at least
that opaque). One network is trained per (threshold, class-weight power); a validation scene (a
seed used neither for collection nor for evaluation) picks the candidate with the fewest Gaussians
among those within the tolerance of the rule's PSNR. The plain rule stays the arbiter: on each
unseen evaluation scene the chosen policy is accepted only if its final held-out PSNR is within
`tolerance` dB of the rule's.

Usage: python gs-native_amd/densify_experiment.py [--quick] [--out result.json] [--params p.npy]
Prints one progress line per phase step and a JSON summary.
"""
from __future__ import annotations

import argparse
import ctypes
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
import gs_native_amd.workloads as workloads  # noqa: E402

PARAMETERS = 5443


def log(message: str) -> None:
    print(f"[densify] {message}", flush=True)


class Scene:
    def __init__(self, context, seed: int, truth_count: int, init_count: int, width: int, height: int,
                 cameras: int, heldout: int, colour_noise: float = 0.05, texture: float = 0.0,
                 scale: float = 1.0, camera_radius: float = 5.0, kind: str = "blobs"):
        self.seed = seed

        def generate(count, sh_degree):
            if kind == "shell":
                return scenes.shell_plane_scene(count, seed=seed, sh_degree=sh_degree, texture=texture,
                                                colour_noise=colour_noise)
            return scenes.blob_scene(count, seed=seed, sh_degree=sh_degree, colour_noise=colour_noise,
                                     texture=texture, scale=scale)

        truth = generate(truth_count, 3)
        _, d2 = gsn.knn3(context, truth[0:3].T.copy())
        truth[3:6] = scenes.scales_from_mean_dist2(d2)[None, :]
        self.init = workloads.initial_from_points(gsn, context, generate(init_count, 0))
        focal = scenes.fov_focal(width, 60.0)
        make = lambda poses: [gsn.Camera.make(width, height, focal, focal, 0.5 * (width - 1), 0.5 * (height - 1), R, t)
                              for R, t in poses]
        hemisphere = kind == "shell"
        self.cameras = make(scenes.sphere_cameras(cameras, camera_radius, seed + 1000, hemisphere=hemisphere))
        self.heldout_cameras = make(scenes.sphere_cameras(heldout, camera_radius, seed + 2000, hemisphere=hemisphere))
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


def run(scene: Scene, iterations: int, cap: int, policy=None, outcome_log=None, eval_every: int = 0,
        useful_opacity=0.02, threshold_scale: float = 1.0):
    """Trains from the scene's initial points on a context of its own (so every run pays for its
    own scratch growth and no run inherits another's allocations); returns the trainer and the
    curve [(iteration, seconds, count, psnr)] with training wall time only (evaluations excluded).
    threshold_scale multiplies the plain rule's two gradient thresholds (the tuned-rule baseline)."""
    context = gsn.Context()
    trainer = trainer_module.Trainer(context, scene.init, scene.cameras, scene.targets, scene.extent,
                                     max_count=cap, seed=scene.seed, policy=policy, outcome_log=outcome_log,
                                     useful_opacity=useful_opacity)
    if threshold_scale != 1.0:
        plain = trainer.control_density
        trainer.control_density = lambda: plain(gradient_threshold=0.0002 * threshold_scale,
                                                absolute_threshold=0.0008 * threshold_scale)
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


def class_weights(labels: np.ndarray, weights: np.ndarray, power: float) -> np.ndarray:
    """Per-class loss weights (keep, clone, split) = (W / (3 W_c))^power, W_c the sample-weighted
    count of class c: power 1 balances the classes, power 0 leaves the sample weights alone. An
    absent class gets weight 0."""
    total = float(weights.sum())
    result = np.zeros(3)
    for c in range(3):
        mass = float(weights[labels == c].sum())
        result[c] = (total / (3.0 * mass)) ** power if mass > 0 else 0.0
    return result


def train_network(context, features, labels, weights, epochs: int, batch: int, learning_rate: float, seed: int,
                  class_weight_power: float = 1.0):
    """Standardises the features, trains on the GPU with the sample weights times the class
    weights, folds the standardisation into W1/b1 and returns (params, report). The deployed
    decision is the plain argmax of the logits."""
    rng = np.random.default_rng(seed)
    order = rng.permutation(features.shape[0])
    features, labels, weights = features[order], labels[order], weights[order]
    validation = max(1, features.shape[0] // 20)
    mean = features[validation:].mean(axis=0)
    std = features[validation:].std(axis=0) + 1e-3
    per_class = class_weights(labels[validation:], weights[validation:], class_weight_power)
    combined = weights * per_class[labels]
    combined = (combined / combined[validation:].mean()).astype(np.float32)
    params = np.zeros(PARAMETERS, np.float32)
    offset = 0
    for rows, cols in ((64, 16), (64, 0), (64, 64), (64, 0), (3, 64), (3, 0)):
        size = rows * max(cols, 1)
        if cols:
            bound = math.sqrt(6.0 / cols)
            params[offset:offset + size] = rng.uniform(-bound, bound, size)
        offset += size
    network = gsn.ProposalNetwork(context, params)
    # Standardised in chunks: the collection can hold tens of millions of samples.
    n = features.shape[0] - validation
    x = gsn.DeviceBuffer((n, 16), np.float32)
    chunk = 1 << 22
    for begin in range(0, n, chunk):
        part = ((features[validation + begin: validation + begin + chunk] - mean) / std).astype(np.float32)
        call_copy(x, begin * 64, part)
    y = gsn.DeviceBuffer.from_numpy(labels[validation:].astype(np.int32))
    w = gsn.DeviceBuffer.from_numpy(combined[validation:])
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
    del x, y, w
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
    predicted = np.argmax(logits.numpy(), axis=1)
    truth = labels[:validation]
    confusion = [[int(np.sum((truth == a) & (predicted == b))) for b in range(3)] for a in range(3)]
    represented = weights[:validation]
    report = {"samples": int(features.shape[0]), "validation": int(validation), "steps": steps,
              "losses": [float(v) for v in losses], "validation_accuracy": float(np.mean(predicted == truth)),
              "validation_confusion": confusion, "class_weight_power": class_weight_power,
              "class_weights": [float(v) for v in per_class],
              "label_counts": [int(np.sum(labels == c)) for c in range(3)],
              "represented_label_share": [float(weights[labels == c].sum() / weights.sum()) for c in range(3)],
              "validation_proposed_share": [float(represented[predicted == c].sum() / represented.sum())
                                            for c in range(3)]}
    return folded.astype(np.float32), report


def call_copy(buffer, offset_bytes: int, array: np.ndarray) -> None:
    """Host-to-device copy of `array` into `buffer` at a byte offset."""
    array = np.ascontiguousarray(array)
    if offset_bytes + array.nbytes > buffer.nbytes:
        raise ValueError("copy beyond the buffer")
    gsn.call("gsn_memcpy_to_device", ctypes.c_void_p(buffer.ptr.value + offset_bytes),
             array.ctypes.data_as(ctypes.c_void_p), array.nbytes)


def time_to_psnr(curve, target):
    for it, seconds, count, psnr in curve:
        if psnr >= target:
            return it, seconds, count
    return None


def scene_for(context, args, seed: int, size: str, truth: int, init: int, cameras: int, heldout: int) -> Scene:
    width, height = (int(v) for v in size.split("x"))
    return Scene(context, seed, truth, init, width, height, cameras, heldout, args.colour_noise, args.texture,
                 camera_radius=args.camera_radius, kind=args.scene)


def collect(context, args) -> tuple[trainer_module.OutcomeLog, list]:
    """Plain-rule runs on the training seeds, logging (features, outcome labels) at every density
    step, one label column per usefulness threshold."""
    outcome_log = trainer_module.OutcomeLog(keep_fraction=args.keep_fraction, seed=1)
    runs = []
    for seed in range(100, 100 + args.train_seeds):
        scene = scene_for(context, args, seed, args.collect_size, args.collect_truth, args.collect_init,
                          args.collect_cameras, 1)
        trainer, curve = run(scene, args.collect_iterations, args.cap, outcome_log=outcome_log,
                             useful_opacity=args.useful_opacity)
        runs.append({"seed": seed, "final_count": trainer.count, "seconds": curve[-1][1],
                     "density_steps": [list(h[:3]) for h in trainer.history]})
        log(f"collect seed {seed}: {trainer.count} Gaussians, {curve[-1][1]:.1f} s, {outcome_log.samples()} samples "
            f"stored, {outcome_log.represented():.0f} represented so far")
        del scene, trainer
    return outcome_log, runs


def train_candidates(context, args, outcome_log) -> list[dict]:
    """One network per (usefulness threshold, class-weight power)."""
    candidates = []
    for column, threshold in enumerate(args.useful_opacity):
        features, labels, weights = outcome_log.arrays(column)
        log(f"threshold {threshold}: labels {np.bincount(labels, minlength=3)} over {features.shape[0]} samples")
        for power in args.class_weight_power:
            params, report = train_network(context, features, labels, weights, args.epochs, args.batch,
                                           args.learning_rate, seed=7, class_weight_power=power)
            report["useful_opacity"] = threshold
            log(f"trained (threshold {threshold}, power {power}): validation accuracy "
                f"{report['validation_accuracy']:.4f}, proposed share {report['validation_proposed_share']}, "
                f"label share {report['represented_label_share']}")
            candidates.append({"useful_opacity": threshold, "class_weight_power": power, "params": params,
                               "training": report})
        del features, labels, weights
    return candidates


def select(context, args, candidates) -> tuple[int, dict]:
    """Runs the rule and every candidate on a validation scene (a seed used neither for collection
    nor for evaluation) and picks the candidate with the fewest final Gaussians among those within
    the tolerance of the rule's PSNR (the most accurate one if none is)."""
    scene = scene_for(context, args, args.validation_seed, args.eval_size, args.eval_truth, args.eval_init,
                      args.eval_cameras, args.eval_heldout)
    rule, curve = run(scene, args.eval_iterations, args.cap, eval_every=args.eval_iterations)
    rule_psnr, rule_count = curve[-1][3], rule.count
    log(f"validation seed {args.validation_seed} rule: PSNR {rule_psnr:.3f} dB, {rule_count} Gaussians, "
        f"{curve[-1][1]:.1f} s")
    del rule
    trials = []
    for index, candidate in enumerate(candidates):
        network = gsn.ProposalNetwork(context, candidate["params"])
        trainer, curve = run(scene, args.eval_iterations, args.cap, policy=network, eval_every=args.eval_iterations)
        trials.append({"candidate": index, "useful_opacity": candidate["useful_opacity"],
                       "class_weight_power": candidate["class_weight_power"], "psnr": curve[-1][3],
                       "count": trainer.count, "seconds": curve[-1][1], "fallbacks": trainer.policy_fallbacks})
        log(f"validation candidate {index} (threshold {candidate['useful_opacity']}, power "
            f"{candidate['class_weight_power']}): PSNR {curve[-1][3]:.3f} dB, {trainer.count} Gaussians, "
            f"{curve[-1][1]:.1f} s")
        del trainer
    del scene
    within = [t for t in trials if t["psnr"] >= rule_psnr - args.tolerance]
    best = min(within, key=lambda t: t["count"]) if within else max(trials, key=lambda t: t["psnr"])
    return best["candidate"], {"seed": args.validation_seed, "rule_psnr": rule_psnr, "rule_count": rule_count,
                               "trials": trials, "chosen": best["candidate"]}


def with_keep_margin(params: np.ndarray, margin: float) -> np.ndarray:
    """The network with `margin` added to the keep logit's output bias (the parameter vector ends
    with the three output biases keep, clone, split): a Gaussian is then cloned or split only where
    that logit leads keep's by more than the margin."""
    adjusted = np.array(params, np.float32, copy=True)
    adjusted[-3] = np.float32(adjusted[-3] + np.float32(margin))
    return adjusted


def validate(context, args, params) -> tuple[float, float | None, dict]:
    """Chooses the network's keep margin (and, with --rule-scales, the tuned-rule baseline's
    threshold scale) on the validation scenes, none of which is used for collection or evaluation.
    On each scene the plain rule, the rule at each threshold scale and the network at each margin
    train once. A setting qualifies when its held-out PSNR is within the tolerance of the plain
    rule's on every validation scene; the qualifying setting with the fewest Gaussians summed over
    the scenes is chosen (the most accurate one, by its worst scene, if none qualifies)."""
    margins = args.keep_margins or [args.keep_margin]
    scales = [v for v in args.rule_scales if v != 1.0]
    scenes = []
    for seed in args.validation_seeds:
        scene = scene_for(context, args, seed, args.eval_size, args.eval_truth, args.eval_init, args.eval_cameras,
                          args.eval_heldout)
        trainer, curve = run(scene, args.eval_iterations, args.cap, eval_every=args.eval_iterations)
        entry = {"seed": seed, "rule_psnr": curve[-1][3], "rule_count": trainer.count, "rule_seconds": curve[-1][1],
                 "margins": [], "scales": []}
        log(f"validation seed {seed} rule: PSNR {curve[-1][3]:.3f} dB, {trainer.count} Gaussians, {curve[-1][1]:.1f} s")
        del trainer
        for margin in margins:
            network = gsn.ProposalNetwork(context, with_keep_margin(params, margin))
            trainer, curve = run(scene, args.eval_iterations, args.cap, policy=network, eval_every=args.eval_iterations)
            entry["margins"].append({"margin": margin, "psnr": curve[-1][3], "count": trainer.count,
                                     "seconds": curve[-1][1], "fallbacks": trainer.policy_fallbacks})
            log(f"validation seed {seed} keep margin {margin}: PSNR {curve[-1][3]:.3f} dB, {trainer.count} "
                f"Gaussians, {curve[-1][1]:.1f} s")
            del trainer, network
        for scale in scales:
            trainer, curve = run(scene, args.eval_iterations, args.cap, eval_every=args.eval_iterations,
                                 threshold_scale=scale)
            entry["scales"].append({"scale": scale, "psnr": curve[-1][3], "count": trainer.count,
                                    "seconds": curve[-1][1]})
            log(f"validation seed {seed} rule threshold scale {scale}: PSNR {curve[-1][3]:.3f} dB, {trainer.count} "
                f"Gaussians, {curve[-1][1]:.1f} s")
            del trainer
        scenes.append(entry)
        del scene

    def choose(key, values):
        rows = [[next(t for t in scene[key + "s"] if t[key] == value) for scene in scenes] for value in values]
        worst = [min(t["psnr"] - scene["rule_psnr"] for t, scene in zip(row, scenes)) for row in rows]
        totals = [sum(t["count"] for t in row) for row in rows]
        qualifying = [k for k in range(len(values)) if worst[k] >= -args.tolerance]
        best = min(qualifying, key=lambda k: totals[k]) if qualifying else max(range(len(values)), key=lambda k: worst[k])
        return values[best]

    margin = choose("margin", margins)
    scale = choose("scale", scales) if scales else None
    return margin, scale, {"scenes": scenes, "chosen_margin": margin, "chosen_scale": scale}


def evaluate(context, args, network, tuned_scale: float | None = None) -> list[dict]:
    evaluations = []
    for seed in range(args.first_eval_seed, args.first_eval_seed + args.eval_seeds):
        scene = scene_for(context, args, seed, args.eval_size, args.eval_truth, args.eval_init, args.eval_cameras,
                          args.eval_heldout)
        results = {}
        runs = [("rule", None, 1.0), ("learned", network, 1.0)]
        if tuned_scale is not None:
            runs.append(("tuned_rule", None, tuned_scale))
        for name, policy, scale in runs:
            trainer, curve = run(scene, args.eval_iterations, args.cap, policy=policy, eval_every=args.eval_every,
                                 threshold_scale=scale)
            results[name] = {"curve": curve, "final_count": trainer.count, "final_psnr": curve[-1][3],
                             "seconds": curve[-1][1], "policy_fallbacks": trainer.policy_fallbacks}
            log(f"eval seed {seed} {name}: PSNR {curve[-1][3]:.3f} dB, {trainer.count} Gaussians, "
                f"{curve[-1][1]:.1f} s")
            del trainer
        evaluation = judge(seed, results["rule"], results["learned"], args.tolerance)
        if "tuned_rule" in results:
            evaluation["tuned_rule"] = judge(seed, results["rule"], results["tuned_rule"], args.tolerance)
            evaluation["tuned_rule"]["threshold_scale"] = tuned_scale
        evaluations.append(evaluation)
        del scene
    return evaluations


def judge(seed: int, rule: dict, learned: dict, tolerance: float) -> dict:
    """The plain rule is the judge: the learned policy is accepted on a scene when its final
    held-out PSNR is within `tolerance` dB of the rule's."""
    reached = time_to_psnr(learned["curve"], rule["final_psnr"] - tolerance)
    return {
        "seed": seed,
        "rule": rule,
        "learned": learned,
        "accepted": learned["final_psnr"] >= rule["final_psnr"] - tolerance,
        "psnr_delta_db": learned["final_psnr"] - rule["final_psnr"],
        "count_reduction": 1.0 - learned["final_count"] / max(rule["final_count"], 1),
        "time_to_rule_psnr": None if reached is None else {
            "iteration": reached[0], "seconds": reached[1], "count": reached[2],
            "time_reduction": 1.0 - reached[1] / max(rule["seconds"], 1e-9),
            "count_reduction": 1.0 - reached[2] / max(rule["final_count"], 1)},
    }


def floats(text: str) -> list[float]:
    return [float(v) for v in text.split(",") if v]


def main() -> int:
    parser = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    parser.add_argument("--scene", choices=("shell", "blobs"), default="shell",
                        help="shell: textured radius-2 shell + ground plane seen from a hemisphere of cameras; "
                             "blobs: the 64-blob mixture seen from a sphere")
    parser.add_argument("--camera-radius", type=float, default=4.0)
    parser.add_argument("--train-seeds", type=int, default=10)
    parser.add_argument("--validation-seed", type=int, default=500,
                        help="the validation scene of the candidate networks' selection")
    parser.add_argument("--validation-seeds", type=lambda text: [int(v) for v in text.split(",") if v],
                        default=[500, 501], help="validation scenes of the keep-margin and rule-scale selection")
    parser.add_argument("--eval-seeds", type=int, default=3)
    parser.add_argument("--first-eval-seed", type=int, default=900)
    parser.add_argument("--collect-truth", type=int, default=3_000_000)
    parser.add_argument("--collect-init", type=int, default=200_000)
    parser.add_argument("--collect-size", type=str, default="1920x1080")
    parser.add_argument("--collect-cameras", type=int, default=300)
    parser.add_argument("--collect-iterations", type=int, default=7000)
    parser.add_argument("--eval-truth", type=int, default=3_000_000)
    parser.add_argument("--eval-init", type=int, default=200_000)
    parser.add_argument("--eval-size", type=str, default="1920x1080")
    parser.add_argument("--eval-cameras", type=int, default=300)
    parser.add_argument("--eval-heldout", type=int, default=8)
    parser.add_argument("--eval-iterations", type=int, default=7000)
    parser.add_argument("--eval-every", type=int, default=500)
    parser.add_argument("--cap", type=int, default=6_000_000)
    parser.add_argument("--keep-fraction", type=float, default=0.1)
    parser.add_argument("--epochs", type=int, default=2)
    parser.add_argument("--batch", type=int, default=65536)
    parser.add_argument("--learning-rate", type=float, default=2e-3)
    parser.add_argument("--class-weight-power", type=floats, default=[0.5, 1.0],
                        help="comma-separated; loss weight of class c is (W / 3 W_c)^power (1 balances the classes)")
    parser.add_argument("--useful-opacity", type=floats, default=[0.02, 0.1, 0.3],
                        help="comma-separated usefulness thresholds: a densified Gaussian's outputs are useful when "
                             "observed in the next interval and at least this opaque (one label column each)")
    parser.add_argument("--tolerance", type=float, default=0.05)
    parser.add_argument("--colour-noise", type=float, default=0.05,
                        help="per-Gaussian colour noise of the truth scenes")
    parser.add_argument("--texture", type=float, default=0.5,
                        help="amplitude of the truth scenes' seeded colour pattern (structure finer than the "
                             "initial points resolve, which the rule densifies for)")
    parser.add_argument("--params", type=str, default="",
                        help="trained network parameters (.npy): skip collection, training and selection")
    parser.add_argument("--keep-margins", type=floats, default=[],
                        help="keep-logit margins to choose from on the validation scene (comma-separated)")
    parser.add_argument("--keep-margin", type=float, default=0.0,
                        help="keep-logit margin to apply when no --keep-margins sweep is given")
    parser.add_argument("--rule-scales", type=floats, default=[],
                        help="gradient-threshold scales of the plain rule to choose a tuned-rule baseline from on "
                             "the validation scene (comma-separated); the baseline then also runs on every "
                             "evaluation scene")
    parser.add_argument("--skip-eval", action="store_true", help="collect, train and select only")
    parser.add_argument("--quick", action="store_true", help="small sizes for a functional check")
    parser.add_argument("--out", type=str, default="")
    parser.add_argument("--record", type=str, default="",
                        help="write the record the at-scale GPU test checks against: the configuration and the "
                             "first evaluation seed's outcome (needs the evaluation)")
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
    summary = {"config": {k: v for k, v in vars(args).items() if k != "out"}}
    if args.params:
        network_params = np.load(args.params, allow_pickle=False).astype(np.float32)
    else:
        outcome_log, runs = collect(context, args)
        summary["collect"] = {"runs": runs, "stored_samples": outcome_log.samples(),
                              "represented_samples": outcome_log.represented()}
        candidates = train_candidates(context, args, outcome_log)
        del outcome_log
        chosen, summary["selection"] = select(context, args, candidates)
        summary["candidates"] = [{k: v for k, v in c.items() if k != "params"} for c in candidates]
        network_params = candidates[chosen]["params"]
        if args.out:
            np.save(pathlib.Path(args.out).with_suffix(".params.npy"), network_params)
    margin, tuned_scale = args.keep_margin, None
    if args.keep_margins or args.rule_scales:
        margin, tuned_scale, summary["validation"] = validate(context, args, network_params)
    summary["keep_margin"] = margin
    if not args.skip_eval:
        network = gsn.ProposalNetwork(context, with_keep_margin(network_params, margin))
        summary["evaluations"] = evaluate(context, args, network, tuned_scale)
        summary["accepted_all"] = all(e["accepted"] for e in summary["evaluations"])
    if args.out:
        pathlib.Path(args.out).write_text(json.dumps(summary, indent=1))
    if args.record and "evaluations" in summary:
        first = summary["evaluations"][0]
        record = {"config": summary["config"], "keep_margin": margin, "held_out": {
            "seed": first["seed"], "accepted": first["accepted"], "psnr_delta_db": first["psnr_delta_db"],
            "rule_psnr": first["rule"]["final_psnr"], "learned_psnr": first["learned"]["final_psnr"],
            "rule_count": first["rule"]["final_count"], "learned_count": first["learned"]["final_count"]}}
        pathlib.Path(args.record).write_text(json.dumps(record, indent=1))
    print(json.dumps(summary), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
