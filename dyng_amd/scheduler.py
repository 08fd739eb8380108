"""The learned scheduler (no counterpart in the reference): before each update it reads a few
features of the batch and of the tree, predicts with a small network the time each of five
schedules of the update would take, and runs the one predicted fastest. Every schedule yields the
same tree (bit for bit); only the time differs. Without a trained network it is the plain fixed
schedule.

The five schedules (``CHOICES``): repair with the default near-far width, repair with a 4x wider
or a 4x narrower width, repair on the full grid (two workgroups per CU, as a solve from the source
runs: faster on the wide frontiers of large batches, slower on the narrow ones of small batches),
and recomputing the tree from scratch.

The network (12 -> 64 -> 64 -> 5, ReLU) keeps FP32 master weights and is trained on the device by
the library's own forward, backward and Adam kernels; it is evaluated in BF16. It regresses log2
of the engine time of each schedule, trained on traces the plain method generates: for a batch
the graph is applied once and the update of the tree repaired under every schedule on clones of
the tree (``sssp.repair``). A schedule other than the plain one runs only when it is predicted
at least 15 % faster.
"""

from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field

import numpy as np

from . import _lib, sssp
from .batch import EdgeBatch
from .errors import InvalidArgumentError
from .graph import Graph

CHOICES = ("repair", "repair_wide", "repair_narrow", "repair_full_grid", "recompute")
INPUTS, HIDDEN, OUTPUTS = _lib.MLP_INPUTS, _lib.MLP_HIDDEN, _lib.MLP_OUTPUTS
FEATURES = ("log2 ops", "delete fraction", "log2 deletions", "log2 insertions", "log2 vertices",
            "log2 average degree", "tree-edge share of deletions", "log2 tree-edge deletions",
            "mean log2 tail degree", "log2 max tail degree", "log2 last invalidated", "log2 last affected")


class Network:
    """The scheduler's MLP on one device."""

    def __init__(self, seed: int = 1, device: int = 0):
        h = C.c_void_p()
        _lib.check(_lib.lib().dyng_hip_mlp_create(int(device), int(seed) & (2**64 - 1), C.byref(h)))
        self._handle = h

    def __del__(self):
        h = getattr(self, "_handle", None)
        if h is not None and h.value and _lib._lib is not None:
            _lib._lib.dyng_hip_mlp_destroy(h)
            self._handle = None

    @property
    def params(self) -> np.ndarray:
        p = np.empty(_lib.MLP_PARAMS, np.float32)
        _lib.check(_lib.lib().dyng_hip_mlp_get_params(self._handle, p.ctypes.data, len(p)))
        return p

    @params.setter
    def params(self, p) -> None:
        p = np.ascontiguousarray(np.asarray(p, np.float32).reshape(-1))
        if len(p) != _lib.MLP_PARAMS:
            raise InvalidArgumentError(f"scheduler.Network: {_lib.MLP_PARAMS} parameters expected, got {len(p)}")
        _lib.check(_lib.lib().dyng_hip_mlp_set_params(self._handle, p.ctypes.data, len(p)))

    @staticmethod
    def _sets(x, t=None):
        x = np.ascontiguousarray(np.asarray(x, np.float32).reshape(-1, INPUTS))
        if t is None:
            return x, None
        t = np.ascontiguousarray(np.asarray(t, np.float32).reshape(-1, OUTPUTS))
        if len(t) != len(x):
            raise InvalidArgumentError("scheduler.Network: inputs and targets differ in length")
        return x, t

    def gradients(self, x, t) -> tuple[np.ndarray, float]:
        """FP32 gradients of the squared error over the known targets (NaN: not measured), divided
        by the number of outputs, and the loss (trained parameters)."""
        x, t = self._sets(x, t)
        g = np.empty(_lib.MLP_TRAINED, np.float32)
        loss = C.c_float()
        _lib.check(_lib.lib().dyng_hip_mlp_gradients(self._handle, x.ctypes.data, t.ctypes.data, len(x),
                                                     g.ctypes.data, C.byref(loss)))
        return g, float(loss.value)

    def normalize_inputs(self, x) -> None:
        """Set the input normalization to the mean and inverse spread of the sample x."""
        x = np.asarray(x, np.float64).reshape(-1, INPUTS)
        p = self.params
        p[_lib.MLP_TRAINED:_lib.MLP_TRAINED + INPUTS] = x.mean(0)
        sd = x.std(0)
        p[_lib.MLP_TRAINED + INPUTS:] = 1.0 / np.where(sd > 1e-6, sd, 1.0)
        self.params = p

    def train(self, x, t, steps: int = 2000, batch: int = 64, lr: float = 3e-3, seed: int = 1) -> float:
        """Adam on the device; returns the final mean squared error over the set."""
        x, t = self._sets(x, t)
        loss = C.c_float()
        _lib.check(_lib.lib().dyng_hip_mlp_train(self._handle, x.ctypes.data, t.ctypes.data, len(x), int(steps),
                                                 int(batch), float(lr), int(seed), C.byref(loss)))
        return float(loss.value)

    def predict(self, x) -> np.ndarray:
        """BF16 inference."""
        x, _ = self._sets(x)
        y = np.empty((len(x), OUTPUTS), np.float32)
        _lib.check(_lib.lib().dyng_hip_mlp_predict(self._handle, x.ctypes.data, len(x), y.ctypes.data))
        return y


SWITCH_MARGIN = 0.2016  # log2(1.15)
MIN_GAIN_MS = 0.02  # a predicted saving below one run's timing noise is no evidence


def pick(predicted_log2_ms) -> int:
    """The schedule to run: the plain one (0) unless another is predicted at least 15 % faster and
    at least 20 us sooner (the rule of the scheduled update entry point): the plain schedule is the
    fallback, left only on a clear predicted gain."""
    y = np.asarray(predicted_log2_ms, np.float32)
    best = int(np.argmin(y))
    gain = y[best] < y[0] - SWITCH_MARGIN and 2.0 ** float(y[0]) - 2.0 ** float(y[best]) >= MIN_GAIN_MS
    return best if best != 0 and gain else 0


def batch_stats(graph: Graph, batch: EdgeBatch, result: sssp.Result | None) -> _lib.BatchStats:
    """The raw features of a batch on the graph before it is applied."""
    nb, keep = batch._native(graph.num_weights)
    parents, n_tree = None, 0
    if result is not None:
        _, parents, n_tree = result.device_arrays()
    out = _lib.BatchStats()
    _lib.check(_lib.lib().dyng_hip_batch_features(graph._handle, C.byref(nb), parents, n_tree,
                                                  graph.resources.stream_handle, C.byref(out)))
    del keep
    return out


@dataclass
class History:
    invalidated: int = 0
    affected: int = 0


def features(st: _lib.BatchStats, history: History) -> np.ndarray:
    ops = st.num_ins + st.num_del
    n, m = max(st.num_vertices, 1), max(st.num_edges, 1)
    return np.array([
        math.log2(1 + ops), st.num_del / max(1, ops), math.log2(1 + st.num_del), math.log2(1 + st.num_ins),
        math.log2(n), math.log2(m / n), st.tree_deletions / max(1, st.num_del), math.log2(1 + st.tree_deletions),
        st.mean_log_degree, math.log2(1 + st.max_degree), math.log2(1 + history.invalidated),
        math.log2(1 + history.affected)], np.float32)


def default_delta(graph: Graph, objective: int = 0) -> int:
    """sssp_default_delta (sssp.cpp:136-145) of the graph as it is now."""
    _, weight_sum = graph.column(objective)
    m, n = max(graph.num_edges, 1), max(graph.num_vertices, 1)
    return max(1, int(32.0 * (weight_sum / m) / (m / n)))


def settings(choice: int, delta: int) -> tuple[sssp.Tuning, int]:
    """The schedule of a choice: (tuning, near-far width)."""
    if choice == 0:
        return sssp.Tuning(), delta
    if choice == 1:
        return sssp.Tuning(), 4 * delta
    if choice == 2:
        return sssp.Tuning(), max(1, delta // 4)
    if choice == 3:
        return sssp.Tuning(blocks_per_cu=2), delta
    if choice == 4:
        return sssp.Tuning(recompute=True), delta
    raise InvalidArgumentError(f"scheduler: choice {choice} outside [0, {len(CHOICES)})")


def use(result: sssp.Result, choice: int, delta: int) -> None:
    tuning, d = settings(choice, delta)
    result.set_tuning(tuning)
    result.set_options(delta=d)


@dataclass
class Trace:
    """Training traces: features and log2 engine times (ms) of every choice, per batch."""

    x: list = field(default_factory=list)
    t: list = field(default_factory=list)

    def arrays(self) -> tuple[np.ndarray, np.ndarray]:
        return np.asarray(self.x, np.float32), np.asarray(self.t, np.float32)


def collect(graph: Graph, result: sssp.Result, batches, trace: Trace | None = None, repeats: int = 1,
            seed: int = 1) -> Trace:
    """Run the plain method over ``batches`` (an iterable of EdgeBatch), timing every choice on
    clones of the tree for each batch; the result follows the default schedule.

    Every repair of a batch after the first finds the rows and words it touches in the caches, so
    a fixed order would favour the later choices: one untimed repair warms them first, and the
    choices are timed in a shuffled order."""
    trace = trace or Trace()
    history = History()
    rng = np.random.default_rng(seed)
    for batch in batches:
        x = features(batch_stats(graph, batch, result), history)
        delta = default_delta(graph, result.options.objective)
        graph.apply(batch)
        warm = result.clone()
        use(warm, 0, delta)
        sssp.repair(graph, warm)
        del warm
        times = [0.0] * len(CHOICES)
        for choice in rng.permutation(len(CHOICES)):
            best = math.inf
            for _ in range(max(1, repeats)):
                trial = result.clone()
                use(trial, int(choice), delta)
                st = sssp.repair(graph, trial)
                best = min(best, st.engine_ms)
                del trial
            times[int(choice)] = math.log2(max(best, 1e-3))
        use(result, 0, delta)
        st = sssp.repair(graph, result)
        history = History(st.invalidated, st.affected)
        trace.x.append(x)
        trace.t.append(times)
    return trace


def explore(graph: Graph, result: sssp.Result, batches, trace: Trace | None = None, seed: int = 1) -> Trace:
    """Run every batch of ``batches`` for real (``sssp.update``) under a schedule drawn at random,
    recording its engine time as the target of that schedule; the other schedules' targets are
    NaN, which the loss ignores. Unlike ``collect``, these are the times of real updates (the
    caches as a stream leaves them), one run each. The result is left on the plain schedule."""
    trace = trace or Trace()
    history = History()
    rng = np.random.default_rng(seed)
    for batch in batches:
        x = features(batch_stats(graph, batch, result), history)
        choice = int(rng.integers(len(CHOICES)))
        use(result, choice, default_delta(graph, result.options.objective))
        st = sssp.update(graph, batch, result)
        t = [math.nan] * len(CHOICES)
        t[choice] = math.log2(max(st.engine_ms, 1e-3))
        history = History(st.invalidated, st.affected)
        trace.x.append(x)
        trace.t.append(t)
    result.set_tuning(sssp.Tuning())
    result.set_options(delta=0)
    return trace


class Scheduler:
    """Picks each update's schedule; the plain fixed schedule when it has no network."""

    def __init__(self, network: Network | None = None):
        self.network = network
        self.history = History()
        self.choices: list[int] = []

    def fit(self, trace: Trace, steps: int = 3000, batch: int = 64, lr: float = 3e-3, seed: int = 1,
            device: int = 0) -> float:
        x, t = trace.arrays()
        if self.network is None:
            self.network = Network(seed, device)
        self.network.normalize_inputs(x)
        return self.network.train(x, t, steps, batch, lr, seed)

    def choose(self, graph: Graph, batch: EdgeBatch, result: sssp.Result) -> int:
        if self.network is None:
            return 0
        x = features(batch_stats(graph, batch, result), self.history)
        return pick(self.network.predict(x[None, :])[0])

    def update(self, graph: Graph, batch: EdgeBatch, result: sssp.Result) -> sssp.Stats:
        """sssp.update under the chosen schedule: the features and the BF16 inference run on the
        device inside the update call (one extra tiny launch pair, no extra round trip)."""
        nb, keep = batch._native(graph.num_weights)
        out = _lib.Stats()
        choice = C.c_int32(0)
        net = self.network._handle if self.network is not None else None
        _lib.check(_lib.lib().dyng_hip_api_update_scheduled(graph._handle, C.byref(nb), result._handle, net,
                                                            int(self.history.invalidated), int(self.history.affected),
                                                            result._resources.stream_handle, C.byref(out),
                                                            C.byref(choice)))
        del keep
        st = sssp.Stats._from_native(out)
        self.history = History(st.invalidated, st.affected)
        self.choices.append(int(choice.value))
        return st
