"""Training loop over the C ABI: the per-iteration schedule of the reference's
GaussianSplatting::Train (src/core/gaussian_splatting.cpp:130-198, 262-287).

Per iteration: camera from an epoch shuffle, SH degree min(3, it/1000), position learning rate
extent * exp-interp(1.6e-4 -> 1.6e-6 over 30000), render, L1+SSIM gradient, fused
gradients + Adam; while it < 15000: statistics accumulation, density control every 100 iterations
after 500 (max screen size 20 after 3000, absolute threshold 8e-4, seed 0) and an opacity cap
every 500 (0.01 every 3000, else 0.8) with the opacity moments cleared. Targets stay resident in
device memory; the only host synchronisation per iteration is the record-count readback.
"""
from __future__ import annotations

import ctypes
import math
import pathlib
import sys

import numpy as np

if __package__:
    from . import CapacityBuffer, DeviceBuffer, Moments, Projection, adam_corrections, call, gaussian_gradients
    from . import gaussian_gradients_adam, gaussian_gradients_adam_project, image_gradients, render, render_projected, _p
else:  # loaded by path
    _pkg = sys.modules["gs_native_amd"]
    DeviceBuffer, Projection, adam_corrections, call = _pkg.DeviceBuffer, _pkg.Projection, _pkg.adam_corrections, _pkg.call
    gaussian_gradients, gaussian_gradients_adam = _pkg.gaussian_gradients, _pkg.gaussian_gradients_adam
    image_gradients, render, _p, Moments = _pkg.image_gradients, _pkg.render, _pkg._p, _pkg.Moments
    CapacityBuffer = _pkg.CapacityBuffer
    gaussian_gradients_adam_project, render_projected = _pkg.gaussian_gradients_adam_project, _pkg.render_projected


FEATURES = 16
FEATURE_STATE_ROWS = 6


class OutcomeLog:
    """Training samples for the densification proposal: the features of every Gaussian at a
    density step paired with the outcome labels of that step (gsn_density_outcomes, one column
    per usefulness threshold of the trainer), collected on the host. Pruned Gaussians (label -1)
    are dropped; samples labelled 0 in every column are kept with probability keep_fraction and
    weighted 1 / keep_fraction, so the weighted loss of every column stays unbiased."""

    def __init__(self, keep_fraction: float = 1.0, seed: int = 0):
        self.keep_fraction = keep_fraction
        self.rng = np.random.default_rng(seed)
        self.features: list[np.ndarray] = []
        self.labels: list[np.ndarray] = []
        self.weights: list[np.ndarray] = []

    def add(self, features: np.ndarray, labels: np.ndarray) -> None:
        labels = labels.reshape(labels.shape[0], -1)
        densified = np.any(labels > 0, axis=1)
        kept = np.all(labels >= 0, axis=1)
        keep = kept & (densified | (self.rng.random(labels.shape[0]) < self.keep_fraction))
        self.features.append(features[keep])
        self.labels.append(labels[keep].astype(np.int8))
        self.weights.append(np.where(densified[keep], 1.0, 1.0 / self.keep_fraction).astype(np.float32))

    def arrays(self, column: int = 0) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
        if not self.features:
            return np.zeros((0, FEATURES), np.float32), np.zeros(0, np.int32), np.zeros(0, np.float32)
        return (np.concatenate(self.features), np.concatenate([l[:, column] for l in self.labels]).astype(np.int32),
                np.concatenate(self.weights))

    def samples(self) -> int:
        return sum(f.shape[0] for f in self.features)

    def represented(self) -> float:
        return float(sum(w.sum() for w in self.weights))


class Trainer:
    """policy: None for the gradient-threshold rule, or a ProposalNetwork whose logits choose the
    densification at every density step. outcome_log: an OutcomeLog to fill with (features,
    outcome label) samples; it labels each density step at the next one. fused (attribute): True
    runs the gradient pass with Adam fused in (no gradient buffer); False writes the [58][N]
    gradients and runs the separate row-major Adam kernel, with bit-identical results (used to
    time Adam on its own). overlap_backward and fuse_projection (attributes, default True) run the
    backward's preparation on the context's second stream and project the next camera's view in
    the gradient pass, which then also adds the density statistics to their accumulators; both
    give bit-identical results. Code that changes `params` outside the
    trainer's own methods must call invalidate_projection() first."""

    def invalidate_projection(self) -> None:
        """Drops the projection the last gradient pass made for the next camera (parameters are
        about to change): the next iteration projects again."""
        self.next_view = None

    def __init__(self, context, params: np.ndarray, cameras: list, targets: list, scene_extent: float,
                 background=(0.0, 0.0, 0.0), max_count: int | None = None, completed: int = 0, seed: int = 0,
                 policy=None, outcome_log: OutcomeLog | None = None, useful_opacity=0.02):
        self.context = context
        self.policy = policy
        self.policy_fallbacks = 0  # density steps where the growth guard handed the decision back to the rule
        self.outcome_log = outcome_log
        # One label column per usefulness threshold (a float or a sequence of them).
        self.useful_opacities = [float(v) for v in np.atleast_1d(useful_opacity)]
        self.features_enabled = policy is not None or outcome_log is not None
        self.pending = None  # the previous density step's decisions, kept for its outcome labels
        self.count = int(params.shape[1])
        # Per-Gaussian state lives in capacity allocations (25% headroom) handed out as views of
        # the current count, so a density step allocates and frees nothing unless the count
        # outgrows them; parameters and moments are double-buffered (density control reads the
        # old and writes the new).
        self._stores = {name: CapacityBuffer(dtype) for name, dtype in (
            ("params_a", np.float32), ("params_b", np.float32), ("moments_a", np.float32),
            ("moments_b", np.float32), ("statistics", np.float32), ("absolute_statistics", np.float32),
            ("sums", np.float32), ("absolute_sums", np.float32), ("observations", np.uint32),
            ("decisions", np.int32), ("offsets", np.int32), ("history", np.int32))}
        self._current = "a"
        self.params = self._stores["params_a"].get((58, self.count))
        self.params.upload(np.ascontiguousarray(params, np.float32))
        self.moments = Moments(self.count, self._stores["moments_a"])
        # The buffers density control writes into (spare parameters and moments, decisions,
        # offsets, history) are allocated now, with the same 25% headroom, instead of at the first
        # density step: a step then allocates nothing unless the count outgrows the headroom.
        self._stores["params_b"].get((58, self.count))
        Moments(self.count, self._stores["moments_b"], zero=False)
        for name in ("decisions", "offsets"):
            self._stores[name].get((self.count + 1, 4))
        self._stores["history"].get((max(self.count, 1),))
        self._allocate_statistics(self.count)
        self.cameras = cameras
        self.targets = [t if isinstance(t, DeviceBuffer) else DeviceBuffer.from_numpy(np.ascontiguousarray(t, np.uint8))
                        for t in targets]
        pixels = max(c.width * c.height for c in cameras)
        self.image = DeviceBuffer((pixels, 3), np.float32)
        self.image_grad = DeviceBuffer((pixels, 3), np.float32)
        self.scene_extent = float(scene_extent)
        self.background = background
        self.max_count = max_count
        self.completed = completed
        self.adam_step = 0
        self.sh_degree = 0
        self.projection = None
        self.overlap_backward = True
        # Fused projection: the gradient pass also projects the updated parameters for the next
        # camera into spare_projection; next_view = (camera, count, SH degree) it was made for, or
        # None when parameters changed since (density control, opacity caps).
        self.fuse_projection = True
        self.spare_projection = None
        self.next_view = None
        self.rng = np.random.default_rng(seed)
        self.order = np.arange(len(cameras))
        self.cursor = len(cameras)
        self.density_counter = 0
        self.last_camera = -1
        self.history = []
        self.feature_state = None
        self.fused = True
        self.gradients = None
        if self.features_enabled:
            self._reset_feature_state()

    def _allocate_statistics(self, count: int) -> None:
        n = max(count, 1)
        # The gradient pass writes both statistics rows of every Gaussian; the accumulators start
        # at zero after every density step.
        self.statistics = self._stores["statistics"].get((2, n))
        self.absolute_statistics = self._stores["absolute_statistics"].get((2, n))
        self.sums = self._stores["sums"].get((n,), zero=True)
        self.absolute_sums = self._stores["absolute_sums"].get((n,), zero=True)
        self.observations = self._stores["observations"].get((n,), zero=True)

    def _reset_feature_state(self) -> None:
        self.feature_state = DeviceBuffer((FEATURE_STATE_ROWS, max(self.count, 1)), np.float32)
        call("gsn_density_features_clear", self.context.handle, self.feature_state.ptr, max(self.count, 1))

    # -- schedule ------------------------------------------------------------------------------
    def position_learning_rate(self, iteration: int) -> float:
        fraction = np.float32(min(1.0, iteration / 30000.0))
        value = np.float32(math.exp(math.log(np.float32(0.00016)) * (1.0 - fraction)
                                    + math.log(np.float32(0.0000016)) * fraction))
        return float(np.float32(self.scene_extent) * value)

    def peek_camera(self) -> int:
        """The camera the next next_camera() returns (when a cycle ends the order is reshuffled now
        instead of then: the generator serves only these shuffles, so the sequence is the same)."""
        if self.cursor >= len(self.order):
            self.rng.shuffle(self.order)
            self.cursor = 0
        return int(self.order[self.cursor])

    def next_camera(self) -> int:
        if self.cursor >= len(self.order):
            self.rng.shuffle(self.order)
            self.cursor = 0
        index = int(self.order[self.cursor])
        self.cursor += 1
        return index

    def iteration(self, camera_index: int | None = None) -> int:
        index = self.next_camera() if camera_index is None else camera_index
        self.last_camera = index
        camera = self.cameras[index]
        iteration = self.completed + 1
        self.sh_degree = min(3, iteration // 1000)
        lr = self.position_learning_rate(iteration)
        if self.overlap_backward:
            # The backward's preparation then runs on a second stream beside the forward raster
            # and the image gradient; results are identical either way.
            self.context.prefetch_backward()
        if self.next_view == (index, self.count, self.sh_degree):
            # The previous gradient pass projected the current parameters for this camera.
            self.projection, self.spare_projection = self.spare_projection, self.projection
            render_projected(self.context, self.projection, self.background, self.image)
        else:
            self.projection = render(self.context, self.params, self.count, camera, self.background, self.image,
                                     self.sh_degree, self.projection)
        self.next_view = None
        image_gradients(self.context, camera, self.image, self.targets[index], self.image_grad)
        self.adam_step += 1
        c1, c2 = adam_corrections(self.adam_step)
        collecting = iteration < 15000  # density statistics are accumulated until iteration 15000
        accumulated = False
        if self.fused and self.fuse_projection and min(3, (iteration + 1) // 1000) == self.sh_degree:
            next_index = self.peek_camera()
            # The pass also adds this view's statistics to the density accumulators.
            accumulate = (self.sums, self.absolute_sums, self.observations) if collecting else None
            self.spare_projection = gaussian_gradients_adam_project(
                self.context, self.projection, self.params, self.sh_degree, self.image, self.image_grad,
                self.moments, lr, c1, c2, self.statistics, self.absolute_statistics, self.cameras[next_index],
                self.spare_projection, accumulate)
            accumulated = accumulate is not None
            self.next_view = (next_index, self.count, self.sh_degree)
        elif self.fused:
            gaussian_gradients_adam(self.context, self.projection, self.params, self.sh_degree, self.image,
                                    self.image_grad, self.moments, lr, c1, c2, self.statistics,
                                    self.absolute_statistics)
        else:
            if self.gradients is None or self.gradients.shape != (58, self.count):
                self.gradients = DeviceBuffer((58, self.count), np.float32)
            gaussian_gradients(self.context, self.projection, self.params, self.sh_degree, self.image,
                               self.image_grad, self.gradients, self.statistics, self.absolute_statistics)
            call("gsn_adam_update", self.context.handle, self.count, self.gradients.ptr, self.params.ptr,
                 self.moments.ptr, lr, c1, c2)
        self.completed = iteration
        records = self.projection.record_count
        if collecting:
            if not accumulated:
                call("gsn_density_accumulate", self.context.handle, self.statistics.ptr, self.absolute_statistics.ptr,
                     self.count, self.sums.ptr, self.absolute_sums.ptr, self.observations.ptr)
            if self.features_enabled:
                call("gsn_density_features_accumulate", self.context.handle, self.statistics.ptr,
                     self.absolute_statistics.ptr, self.projection.depths.ptr, self.projection.conic_opacities.ptr,
                     self.count, self.feature_state.ptr, max(self.count, 1))
            if self.completed > 500 and self.completed % 100 == 0:
                self.control_density()
            if self.completed % 500 == 0:
                self.reset_opacity(0.01 if self.completed % 3000 == 0 else 0.8)
        return records

    # -- density -------------------------------------------------------------------------------
    def reset_opacity(self, cap: float) -> None:
        self.invalidate_projection()
        if cap < 0.5:
            # Opacities collapse: outcomes of the pending density step are not measurable.
            self.pending = None
        logit = float(np.float32(math.log(np.float32(cap) / (np.float32(1.0) - np.float32(cap)))))
        call("gsn_density_reset_opacity", self.context.handle, self.params.ptr, self.count, self.count, logit)
        call("gsn_zero_rows", self.context.handle, self.moments.ptr, self.count, 9, 1)

    def control_density(self, gradient_threshold=0.0002, percent_dense=0.01, min_opacity=0.005,
                        absolute_threshold=0.0008) -> tuple[int, int]:
        self.invalidate_projection()
        count = self.count
        max_screen = 20.0 if self.completed > 3000 else 0.0
        decisions = self._stores["decisions"].get((count + 1, 4))
        offsets = self._stores["offsets"].get((count + 1, 4))
        totals = (ctypes.c_int32 * 4)()
        split_threshold = float(np.float32(percent_dense) * np.float32(self.scene_extent))
        world = float(np.float32(0.1) * np.float32(self.scene_extent))

        features = None
        if self.features_enabled:
            features = DeviceBuffer((max(count, 1), FEATURES), np.float32)
            call("gsn_density_features", self.context.handle, self.params.ptr, count, self.sums.ptr,
                 self.absolute_sums.ptr, self.observations.ptr, self.feature_state.ptr, max(count, 1),
                 gradient_threshold, split_threshold, self.scene_extent, features.ptr)
        if self.outcome_log is not None and self.pending is not None:
            self._log_outcomes(count)
        logits = None
        if self.policy is not None and count > 0:
            logits = DeviceBuffer((count, 3), np.float32)
            self.policy.forward(features, count, logits, context=self.context)

        def count_changes(threshold, absolute, use_policy=True):
            if use_policy and logits is not None and np.isfinite(threshold):
                call("gsn_density_count_proposed", self.context.handle, self.params.ptr, count, self.sums.ptr,
                     self.absolute_sums.ptr, self.observations.ptr, logits.ptr, split_threshold, min_opacity, world,
                     int(max_screen > 0), decisions.ptr, offsets.ptr, totals)
            else:
                call("gsn_density_count", self.context.handle, self.params.ptr, count, self.sums.ptr,
                     self.absolute_sums.ptr, self.observations.ptr, threshold, split_threshold, min_opacity, world,
                     int(max_screen > 0), absolute, decisions.ptr, offsets.ptr, totals)
            t = list(totals)
            return t, t[0] + t[1] + 2 * t[3]

        t, out_count = count_changes(gradient_threshold, absolute_threshold)
        if logits is not None:
            # Growth guard: a proposal that adds more than twice what the plain rule would (and more
            # than 1% of the scene) is not followed; the rule decides this step instead.
            proposed_added = out_count - count
            _, rule_count = count_changes(gradient_threshold, absolute_threshold, use_policy=False)
            if proposed_added > max(2 * max(rule_count - count, 0), count // 100):
                self.policy_fallbacks += 1
                t, out_count = count_changes(gradient_threshold, absolute_threshold, use_policy=False)
            else:
                t, out_count = count_changes(gradient_threshold, absolute_threshold)
        if self.max_count is not None and out_count > self.max_count:
            # Capacity cap: no clones or splits this round, pruning only.
            t, out_count = count_changes(float("inf"), float("inf"))
        spare = "b" if self._current == "a" else "a"
        params = self._stores[f"params_{spare}"].get((58, max(out_count, 1)))
        history = self._stores["history"].get((max(out_count, 1),))
        call("gsn_density_write", self.context.handle, self.params.ptr, count, decisions.ptr, offsets.ptr, totals, 0,
             self.density_counter, params.ptr, out_count, history.ptr, None, None)
        self.density_counter += 1
        moments = Moments(out_count, self._stores[f"moments_{spare}"], zero=False)  # remap writes every pair
        call("gsn_remap_history", self.context.handle, history.ptr, count, out_count, self.moments.ptr, moments.ptr)
        self.params, self.moments, self._current = params, moments, spare
        self.count = out_count
        self._allocate_statistics(out_count)
        if self.features_enabled:
            self._reset_feature_state()
        if self.outcome_log is not None:
            # Rule labels of this step are measured at the next one from what its outputs became.
            self.pending = (decisions, offsets, list(t), count, features.numpy()[:count] if count else None)
        changes = (t[1] + 2 * t[3], count - t[0])
        self.history.append((self.completed, count, out_count, changes))
        return changes

    def _log_outcomes(self, count: int) -> None:
        decisions, offsets, t, previous_count, previous_features = self.pending
        self.pending = None
        if previous_count == 0:
            return
        labels = DeviceBuffer((previous_count,), np.int32)
        totals = (ctypes.c_int32 * 4)(*t)
        columns = []
        for threshold in self.useful_opacities:
            call("gsn_density_outcomes", self.context.handle, decisions.ptr, offsets.ptr, totals, previous_count,
                 self.params.ptr, self.observations.ptr, count, threshold, labels.ptr)
            columns.append(labels.numpy())
        self.outcome_log.add(previous_features, np.stack(columns, axis=1))

    def params_host(self) -> np.ndarray:
        return self.params.numpy()[:, : self.count]
