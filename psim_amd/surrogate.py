"""Steady-state surrogate for warm starts (include/psim_surrogate.h).

An MLP 32 -> 512 x 4 (ReLU) -> 1 maps one sensor of one batched-geometry structure
(psim_amd.configs.batch_geometry_params) to its steady-state temperature.  It predicts the
linear-response shape u = (T - t_eq) / (t_hot - t_eq), t_eq the mean wall temperature: the
deviational problem is linear in the wall offsets to first order, so u depends on the geometry,
the specularity and the sensor's position, and only weakly on the temperatures.
Its weights and activations are bfloat16; every product runs on the MFMA units with FP32
accumulation, forward and backward, and Adam keeps FP32 master weights
(psim_amd/csrc/surrogate.hip).  It is trained on steady-state fields the GPU Monte Carlo computes
(steady_state_targets) and only proposes each sensor's initial temperature (warm_start_inits);
every reported field still comes from the Monte Carlo, and a cold start (every sensor at t_eq) is
the fallback.
"""
from __future__ import annotations

import ctypes
import math
import os

import numpy as np

from . import PsimError, configs, lib

N_FEATURES = 32
HIDDEN = 512
LAYERS = 4
MAX_BATCH_PHONONS = 400_000_000  # phonon budget of one batched run (~120 GB of HBM)
WEIGHTS_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "surrogate_32x512x4.npz")

_VP = ctypes.c_void_p
_F = ctypes.POINTER(ctypes.c_float)


def _bind():
    L = lib()
    if getattr(L, "_mlp_bound", False):
        return L
    sig = {
        "psim_mlp_create": (ctypes.c_int, [ctypes.POINTER(_VP), ctypes.c_int, ctypes.c_int32, ctypes.c_int32,
                                           ctypes.c_int32]),
        "psim_mlp_destroy": (None, [_VP]),
        "psim_mlp_last_error": (ctypes.c_char_p, [_VP]),
        "psim_mlp_num_params": (ctypes.c_int64, [_VP]),
        "psim_mlp_set_params": (ctypes.c_int, [_VP, _F]),
        "psim_mlp_get_params": (ctypes.c_int, [_VP, _F]),
        "psim_mlp_forward": (ctypes.c_int, [_VP, _F, ctypes.c_int32, _F]),
        "psim_mlp_gradients": (ctypes.c_int, [_VP, _F, _F, ctypes.c_int32, _F, _F]),
        "psim_mlp_train_step": (ctypes.c_int, [_VP, _F, _F, ctypes.c_int32, ctypes.c_float, _F]),
        "psim_mlp_gemm_check": (ctypes.c_int, [ctypes.c_int, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                               ctypes.c_int32, ctypes.c_int32, _F, _F, _F]),
        "psim_mlp_time_gemm": (ctypes.c_int, [ctypes.c_int, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                              ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, _F]),
        "psim_mlp_time_product": (ctypes.c_int, [_VP, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, _F,
                                                 ctypes.POINTER(ctypes.c_int32)]),
        "psim_mlp_set_graph": (ctypes.c_int, [_VP, ctypes.c_int32]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    L._mlp_bound = True
    return L


def _fp(a: np.ndarray):
    return a.ctypes.data_as(_F)


def gemm_check(a: np.ndarray, b: np.ndarray, ta: bool, tb: bool, device: int = 0) -> np.ndarray:
    """op(a) @ op(b) on the network's MFMA GEMM (operands rounded to bf16, FP32 accumulation)."""
    L = _bind()
    a = np.ascontiguousarray(a, np.float32)
    b = np.ascontiguousarray(b, np.float32)
    M = a.shape[1] if ta else a.shape[0]
    K = a.shape[0] if ta else a.shape[1]
    N = b.shape[0] if tb else b.shape[1]
    c = np.zeros((M, N), np.float32)
    rc = L.psim_mlp_gemm_check(int(device), int(ta), int(tb), M, N, K, _fp(a), _fp(b), _fp(c))
    if rc != 0:
        raise PsimError(f"psim_mlp_gemm_check failed ({rc})")
    return c


def time_gemm(M: int, N: int, K: int, ta: bool = False, tb: bool = False, reps: int = 50, device: int = 0) -> float:
    """Average ms of one launch of the network's GEMM kernel on op(A)[M][K] x op(B)[K][N]."""
    L = _bind()
    ms = ctypes.c_float()
    rc = L.psim_mlp_time_gemm(int(device), int(ta), int(tb), int(M), int(N), int(K), int(reps), ctypes.byref(ms))
    if rc != 0:
        raise PsimError(f"psim_mlp_time_gemm failed ({rc})")
    return float(ms.value)


class Surrogate:
    """The MLP on one GPU (psim_mlp)."""

    def __init__(self, n_in: int = N_FEATURES, hidden: int = HIDDEN, layers: int = LAYERS, device: int = 0,
                 seed: int = 0):
        L = _bind()
        h = _VP()
        rc = L.psim_mlp_create(ctypes.byref(h), int(device), int(n_in), int(hidden), int(layers))
        if rc != 0:
            raise PsimError(f"psim_mlp_create failed ({rc})")
        self._h, self._L = h, L
        self.dims = [n_in] + [hidden] * layers + [16]
        self.n_params = int(L.psim_mlp_num_params(h))
        self.set_params(self.initial_params(seed))

    def _check(self, rc: int) -> None:
        if rc != 0:
            raise PsimError(self._L.psim_mlp_last_error(self._h).decode())

    def close(self) -> None:
        if getattr(self, "_h", None):
            self._L.psim_mlp_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def layer_views(self, params: np.ndarray):
        """[(W [in][out], b [out]) per layer] as views into the flat parameter vector."""
        out, off = [], 0
        for k, n in zip(self.dims[:-1], self.dims[1:]):
            w = params[off:off + k * n].reshape(k, n)
            off += k * n
            out.append((w, params[off:off + n]))
            off += n
        return out

    def initial_params(self, seed: int) -> np.ndarray:
        """He-normal hidden layers, a small output layer (column 0 only), zero biases."""
        rng = np.random.default_rng(seed)
        p = np.zeros(self.n_params, np.float32)
        views = self.layer_views(p)
        for i, (w, _) in enumerate(views):
            fan_in = w.shape[0]
            if i + 1 < len(views):
                w[:] = rng.normal(0.0, math.sqrt(2.0 / fan_in), w.shape)
            else:
                w[:, 0] = rng.normal(0.0, math.sqrt(1.0 / fan_in), fan_in)
        return p

    def set_params(self, params: np.ndarray) -> None:
        p = np.ascontiguousarray(params, np.float32)
        assert p.size == self.n_params
        self._check(self._L.psim_mlp_set_params(self._h, _fp(p)))

    def params(self) -> np.ndarray:
        p = np.zeros(self.n_params, np.float32)
        self._check(self._L.psim_mlp_get_params(self._h, _fp(p)))
        return p

    def predict(self, x: np.ndarray) -> np.ndarray:
        x = np.ascontiguousarray(x, np.float32)
        y = np.zeros(x.shape[0], np.float32)
        self._check(self._L.psim_mlp_forward(self._h, _fp(x), x.shape[0], _fp(y)))
        return y

    def gradients(self, x: np.ndarray, t: np.ndarray):
        """(mean squared error, flat gradient) of one batch."""
        x = np.ascontiguousarray(x, np.float32)
        t = np.ascontiguousarray(t, np.float32)
        g = np.zeros(self.n_params, np.float32)
        loss = ctypes.c_float()
        self._check(self._L.psim_mlp_gradients(self._h, _fp(x), _fp(t), x.shape[0], _fp(g), ctypes.byref(loss)))
        return loss.value, g

    def set_graph(self, enable: bool) -> None:
        """Training steps replayed from a captured HIP graph (default) or all eager."""
        self._check(self._L.psim_mlp_set_graph(self._h, int(bool(enable))))

    PRODUCTS = {"forward": 0, "forward_train": 1, "weight_grad": 2, "input_grad": 3}

    def time_product(self, kind: str, rows: int, reps: int = 50) -> tuple:
        """(ms per launch, tile) of one of the network's fused hidden x hidden products at `rows` rows
        (psim_mlp_time_product): forward, forward_train, weight_grad or input_grad."""
        ms = ctypes.c_float()
        tile = ctypes.c_int32()
        self._check(self._L.psim_mlp_time_product(self._h, self.PRODUCTS[kind], int(rows), int(reps), ctypes.byref(ms),
                                                  ctypes.byref(tile)))
        code = tile.value
        if code < 0:  # the kernel that keeps the weight slice in LDS: rows per band x columns
            return float(ms.value), f"{-code // 1000}x{-code % 1000} (weights resident in LDS)"
        return float(ms.value), f"{code // 1000}x{code % 1000}"

    def train_step(self, x: np.ndarray, t: np.ndarray, lr: float) -> float:
        x = np.ascontiguousarray(x, np.float32)
        t = np.ascontiguousarray(t, np.float32)
        loss = ctypes.c_float()
        self._check(self._L.psim_mlp_train_step(self._h, _fp(x), _fp(t), x.shape[0], float(lr), ctypes.byref(loss)))
        return loss.value

    def fit(self, x: np.ndarray, t: np.ndarray, epochs: int = 40, batch: int = 4096, lr: float = 1e-3,
            seed: int = 0, schedule=None) -> list:
        """Adam over shuffled mini-batches with a cosine-decayed learning rate; returns the mean
        training loss of each epoch.  schedule = (first epoch, total epochs) places these epochs on
        a longer schedule (training in several calls)."""
        rng = np.random.default_rng(seed)
        n = x.shape[0]
        steps_per_epoch = max(1, n // batch)
        first, total_epochs = schedule if schedule else (0, epochs)
        total = total_epochs * steps_per_epoch
        history, step = [], first * steps_per_epoch
        for _ in range(epochs):
            order = rng.permutation(n)
            losses = []
            for s in range(steps_per_epoch):
                idx = order[s * batch:(s + 1) * batch]
                rate = lr * 0.5 * (1.0 + math.cos(math.pi * step / total))
                losses.append(self.train_step(x[idx], t[idx], max(rate, lr * 1e-3)))
                step += 1
            history.append(float(np.mean(losses)))
        return history


def sensor_geometry(p: dict):
    """Per sensor of a batched-geometry structure (psim_amd.configs.geometry_from_params): centroid
    x, centroid y, arc-length position of the centroid / structure length, is the wedge, lies in
    the second arm."""
    n, length, width = p["n_sensors"], p["length"], p["width"]
    rows = []
    if not p["kinked"]:
        for i in range(n):
            rows.append(((i + 0.5) * length / n, 0.5 * width, (i + 0.5) / n, 0.0, 0.0))
        return rows
    theta = p["theta"]
    n1 = (n - 1) // 2
    n2 = n - 1 - n1
    l1 = l2 = 0.5 * length
    for i in range(n1):
        rows.append(((i + 0.5) * l1 / n1, 0.5 * width, (i + 0.5) * l1 / n1 / length, 0.0, 0.0))
    q = (l1, width)  # inner corner
    a2 = (q[0] + width * math.sin(theta), q[1] - width * math.cos(theta))
    rows.append(((l1 + a2[0] + q[0]) / 3.0, (a2[1] + q[1]) / 3.0, l1 / length, 1.0, 0.0))
    d = (math.cos(theta), math.sin(theta))
    for j in range(n2):
        u = (j + 0.5) * l2 / n2
        rows.append((0.5 * (a2[0] + q[0]) + u * d[0], 0.5 * (a2[1] + q[1]) + u * d[1], (l1 + u) / length, 0.0, 1.0))
    return rows


def features(p: dict) -> np.ndarray:
    """[n_sensors, 32] inputs for one structure: its shape, wall temperatures, specularity and each
    sensor's centroid and position along it."""
    n = p["n_sensors"]
    length, width, theta, spec = p["length"], p["width"], p["theta"], p["spec"]
    kinked = float(p["kinked"])
    t_eq = 0.5 * (p["t_hot"] + p["t_cold"])
    x = np.zeros((n, N_FEATURES), np.float32)
    for i, (cx, cy, s, wedge, arm2) in enumerate(sensor_geometry(p)):
        x[i] = [kinked, length / 5000.0, width / 500.0, theta / (0.5 * math.pi), n / 200.0, (t_eq - 300.0) / 50.0,
                (p["t_hot"] - p["t_cold"]) / 100.0, spec, s, (i + 0.5) / n, cx / 5000.0, cy / 5000.0, wedge, arm2,
                width / length, math.log10(length) / 4.0, math.log10(width) / 3.0, s * s, s * (1.0 - s),
                min(s, 1.0 - s), spec * s, spec * (1.0 - s), math.cos(theta), math.sin(theta), kinked * s,
                arm2 * (s - 0.5), 10.0 / n, length / width / 100.0, s - 0.5, math.log10(length / width) / 2.0,
                spec * spec, width / length * s]
    return x


def shape_of(p: dict, temps: np.ndarray) -> np.ndarray:
    """u = (T - t_eq) / (t_hot - t_eq) of temperatures of structure p."""
    t_eq = 0.5 * (p["t_hot"] + p["t_cold"])
    return (np.asarray(temps, np.float64) - t_eq) / (p["t_hot"] - t_eq)


def temps_of(p: dict, u) -> list:
    """Temperatures from the shape u (inverse of shape_of)."""
    t_eq = 0.5 * (p["t_hot"] + p["t_cold"])
    return [float(t_eq + (p["t_hot"] - t_eq) * v) for v in np.asarray(u, np.float64)]


def linear_inits(p: dict) -> list:
    """The guess without a network: temperature linear in arc length from the hot to the cold wall."""
    return [p["t_hot"] + s * (p["t_cold"] - p["t_hot"]) for (_, _, s, _, _) in sensor_geometry(p)]


def predict_inits(net: Surrogate, params: list) -> list:
    """Predicted steady-state temperature of every sensor of every structure (one forward pass):
    the warm start's initial temperatures."""
    x = np.concatenate([features(p) for p in params])
    u = net.predict(x)
    out, off = [], 0
    for p in params:
        n = p["n_sensors"]
        out.append(temps_of(p, u[off:off + n]))
        off += n
    return out


def run_relaxations(engine, params: list, inits, phonons: int, steps: int, sim_ns: float, seed: int, reduce,
                    batch: int = 256, threads: int = 16, progress=None) -> list:
    """Runs every structure in periodic mode (every step recorded) from initial sensor temperatures
    inits[i] (None: every sensor at t_eq, the cold start), `phonons` each, in batches of at most
    `batch` structures and MAX_BATCH_PHONONS phonons through psim_batch_run; returns
    reduce(p, temps [S, steps]) of each; progress(message), if given, after every batch."""
    from concurrent.futures import ThreadPoolExecutor

    from . import Model, run_batch

    # every phonon of a batch holds ~300 bytes of HBM (emission pool, two live buffers, lists)
    batch = max(1, min(batch, MAX_BATCH_PHONONS // max(1, phonons)))
    out = []
    with ThreadPoolExecutor(max_workers=threads) as ex:
        for b0 in range(0, len(params), batch):
            chunk = params[b0:b0 + batch]
            cfgs = [configs.geometry_from_params(dict(p, phonons=phonons), steps, sim_ns, sim_type=1,
                                                 step_interval=100, t_init=None if inits is None else inits[b0 + k])
                    for k, p in enumerate(chunk)]
            models = list(ex.map(Model, cfgs))
            run_batch(engine, models, seed + b0)
            out.extend(reduce(p, m.results(0).temps) for p, m in zip(chunk, models))
            del models
            if progress:
                progress(f"  batch done: {len(out)}/{len(params)} structures")
    return out


def gpu_seconds_to_steady_state(engine, params: list, inits, ttss: list, phonons: int, steps: int, sim_ns: float,
                                seed: int, threads: int = 16, progress=None) -> dict:
    """GPU time to bring every structure to its steady state: each run cut at its own time to
    steady state ttss[i] (steps; None: the whole horizon), with its phonon budget scaled so the
    energy per phonon equals that of the full run (the noise of the run whose ttss was measured).
    Runs of equal length share batches.  Returns {"gpu_s", "phonon_steps", "runs"}."""
    from collections import defaultdict
    from concurrent.futures import ThreadPoolExecutor

    from . import Model, run_batch

    dt = sim_ns / steps

    def cut_config(i, n):
        p = params[i]
        init = None if inits is None else inits[i]
        full = configs.geometry_from_params(dict(p, phonons=phonons), steps, sim_ns, sim_type=1, step_interval=100,
                                            t_init=init)
        probe = configs.geometry_from_params(dict(p, phonons=phonons), n, n * dt, sim_type=1, step_interval=100,
                                             t_init=init)
        eff_full = Model(full).prepare(seed)[0]
        eff_cut = Model(probe).prepare(seed)[0]
        n_cut = max(1, int(round(phonons * eff_cut / eff_full)))
        return configs.geometry_from_params(dict(p, phonons=n_cut), n, n * dt, sim_type=1, step_interval=100,
                                            t_init=init), n_cut

    groups = defaultdict(list)
    for i, t in enumerate(ttss):
        groups[steps if t is None else int(t)].append(i)
    gpu_s, phonon_steps, runs = 0.0, 0, 0
    with ThreadPoolExecutor(max_workers=threads) as ex:
        for n, idx in sorted(groups.items()):
            made = list(ex.map(lambda i: cut_config(i, n), idx))
            b0 = 0
            while b0 < len(made):
                # fill a batch up to the phonon cap
                b1, total = b0, 0
                while b1 < len(made) and (b1 == b0 or total + made[b1][1] <= MAX_BATCH_PHONONS):
                    total += made[b1][1]
                    b1 += 1
                models = list(ex.map(Model, [c for c, _ in made[b0:b1]]))
                counters = run_batch(engine, models, seed + runs)
                gpu_s += engine.last_timing()["run_ms"] / 1e3
                phonon_steps += counters["phonon_steps"]
                runs += 1
                del models
                b0 = b1
            if progress:
                progress(f"  cut runs of {n} steps: {len(idx)} structures, {gpu_s:.1f} GPU s so far")
    return {"gpu_s": gpu_s, "phonon_steps": phonon_steps, "runs": runs}


def steady_state_targets(engine, params: list, phonons: int = 100_000, seed: int = 0, batch: int = 256):
    """Training targets: the self-consistent steady-state shape u of every sensor.

    A sensor's heat capacity and scattering table are evaluated at its initial temperature
    (SensorController::updateTables, psim/src/sensorController.cpp:38-50), so the steady state a
    run reaches depends slightly on where it starts; the reference's steady-state iteration moves
    them to the sensor's steady temperature (sensorController.h:63-67, model.cpp:157-176).  Two
    passes per structure do the same: a cold start over 40 ns (4000 steps of 10 ps) averaged over
    its second half, then a run started from that field over 20 ns (2000 steps) averaged over its
    last 1500 steps.  Returns (features [rows, 32], u [rows])."""
    def first_pass(p, temps):
        return temps[:, temps.shape[1] // 2:].mean(axis=1).tolist()

    def second_pass(p, temps):
        return shape_of(p, temps[:, 500:].mean(axis=1))

    us = []
    for b0 in range(0, len(params), batch):
        chunk = params[b0:b0 + batch]
        t1 = run_relaxations(engine, chunk, None, phonons, 4000, 40.0, seed + b0, first_pass, batch)
        us.extend(run_relaxations(engine, chunk, t1, phonons, 2000, 20.0, seed + b0 + 1, second_pass, batch))
    x = np.concatenate([features(p) for p in params])
    return x, np.concatenate(us).astype(np.float32)


def time_to_steady_state(p: dict, temps: np.ndarray, rel_tol: float, window: int = 100, consecutive: int = 3):
    """Steps until the RMS over sensors of the change in window-mean temperature stays below
    rel_tol * |t_hot - t_cold| for `consecutive` windows (psim_amd.analysis.time_to_steady_state)."""
    from . import analysis
    return analysis.time_to_steady_state(temps, window, rel_tol * abs(p["t_hot"] - p["t_cold"]), consecutive)


def bf16_bits(a: np.ndarray) -> np.ndarray:
    """bfloat16 bits of finite float32 values, rounded to nearest even."""
    bits = np.ascontiguousarray(a, np.float32).view(np.uint32)
    return ((bits + np.uint32(0x7FFF) + ((bits >> 16) & np.uint32(1))) >> 16).astype(np.uint16)


def save_weights(net: Surrogate, path: str = WEIGHTS_PATH) -> None:
    """Weights as bfloat16 bits (what the forward pass reads), biases as float32."""
    params = net.params()
    ws, bs = [], []
    for w, b in net.layer_views(params):
        ws.append(bf16_bits(w))
        bs.append(b.astype(np.float32))
    np.savez_compressed(path, dims=np.array(net.dims, np.int32), **{f"w{i}": w for i, w in enumerate(ws)},
                        **{f"b{i}": b for i, b in enumerate(bs)})


def load_weights(path: str = WEIGHTS_PATH, device: int = 0) -> Surrogate:
    """The surrogate with the weights saved by save_weights."""
    with np.load(path, allow_pickle=False) as f:
        dims = [int(d) for d in f["dims"]]
        net = Surrogate(n_in=dims[0], hidden=dims[1], layers=len(dims) - 2, device=device)
        params = np.zeros(net.n_params, np.float32)
        for i, (w, b) in enumerate(net.layer_views(params)):
            w[:] = (f[f"w{i}"].astype(np.uint32) << 16).view(np.float32)
            b[:] = f[f"b{i}"]
    net.set_params(params)
    return net
