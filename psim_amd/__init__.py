"""psim_amd: MI355X-native engine for psim's phonon Monte Carlo loop.

The product is libpsim_amd.so (HIP kernels for gfx950 plus a C++17 host behind the
C ABI of include/psim_gpu.h).  This module binds that ABI with ctypes:

  Engine         a psim_gpu context on one GPU (replaces psim's ModelSimulator)
  Model          a psim model read from the reference's JSON schema (the C++ host:
                 Model/Material/Sensor/Cell setup, SensorInterpreter, OutputManager)
  configs        generators for the workloads in the reference's JSON schema

Nothing here falls back to a CPU implementation: if the library is missing or
cannot load, every call raises.
"""
from __future__ import annotations

import ctypes
import json as _json
import os
from dataclasses import dataclass

import numpy as np

from . import configs  # noqa: F401

_HERE = os.path.dirname(os.path.abspath(__file__))
# PSIM_AMD_LIB names another build of the same sources (the AddressSanitizer build of csrc/asan.mk)
LIB_PATH = os.environ.get("PSIM_AMD_LIB") or os.path.join(_HERE, "libpsim_amd.so")
HEADER_PATH = os.path.join(os.path.dirname(_HERE), "include", "psim_gpu.h")

COUNTER_NAMES = ["phonon_steps", "scatters", "impacts", "relocations", "absorbed", "expired",
                 "not_started", "tallies", "stalled"]

_D = ctypes.POINTER(ctypes.c_double)
_I64 = ctypes.POINTER(ctypes.c_int64)
_I32 = ctypes.POINTER(ctypes.c_int32)
_VP = ctypes.c_void_p
_lib = None


class PsimError(RuntimeError):
    pass


def build() -> None:
    """Compile libpsim_amd.so in-tree (hipcc --offload-arch=gfx950)."""
    import subprocess
    subprocess.run(["make", "-s", "-j8", "-C", os.path.join(_HERE, "csrc")], check=True)


def lib():
    """Load libpsim_amd.so; raises PsimError if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise PsimError(f"{LIB_PATH} is missing: build it with `make -C psim_amd/csrc` "
                        "(or __graft_entry__.build())")
    L = ctypes.CDLL(LIB_PATH)
    sig = {
        "psim_gpu_create": (ctypes.c_int, [ctypes.POINTER(_VP), ctypes.c_int]),
        "psim_gpu_destroy": (None, [_VP]),
        "psim_gpu_last_error": (ctypes.c_char_p, [_VP]),
        "psim_gpu_upload": (ctypes.c_int, [_VP, _VP, _VP, _VP, _VP]),
        "psim_gpu_run": (ctypes.c_int, [_VP, ctypes.c_uint64, _I64, _D, _I64]),
        "psim_gpu_run_resident": (ctypes.c_int, [_VP, ctypes.c_uint64]),
        "psim_gpu_fetch": (ctypes.c_int, [_VP, _I64, _D, _I64]),
        "psim_gpu_set_fate": (ctypes.c_int, [_VP, ctypes.c_int]),
        "psim_gpu_fetch_fate": (ctypes.c_int, [_VP, _I64, _I32, _D]),
        "psim_gpu_last_timing": (ctypes.c_int, [_VP, _D, _I64, _D]),
        "psim_gpu_set_window": (ctypes.c_int, [_VP, ctypes.c_int32]),
        "psim_gpu_set_grid": (ctypes.c_int, [_VP, ctypes.c_int32]),
        "psim_gpu_set_block": (ctypes.c_int, [_VP, ctypes.c_int32]),
        "psim_gpu_get_launch_shape": (ctypes.c_int, [_VP, ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int32),
                                                     ctypes.POINTER(ctypes.c_int32)]),
        "psim_gpu_run_graph": (ctypes.c_int, [_VP, ctypes.c_uint64, ctypes.POINTER(ctypes.c_int64),
                                              ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int64)]),
        "psim_gpu_run_async": (ctypes.c_int, [_VP, ctypes.c_uint64, _VP]),
        "psim_gpu_set_tally_mode": (ctypes.c_int, [_VP, ctypes.c_int32]),
        "psim_gpu_set_tally_precision": (ctypes.c_int, [_VP, ctypes.c_int32]),
        "psim_gpu_autotune": (ctypes.c_int, [_VP, ctypes.c_uint64, ctypes.POINTER(ctypes.c_int32),
                                             ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_double)]),
        "psim_model_from_json": (ctypes.c_int, [ctypes.POINTER(_VP), ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int]),
        "psim_model_destroy": (None, [_VP]),
        "psim_model_last_error": (ctypes.c_char_p, [_VP]),
        "psim_model_sizes": (ctypes.c_int, [_VP, _I32, _I64, _I64]),
        "psim_model_run": (ctypes.c_int, [_VP, _VP, ctypes.c_uint64]),
        "psim_model_results": (ctypes.c_int, [_VP, ctypes.c_int64, _D, _D, _D, _I64, _D, _I64]),
        "psim_model_budget": (ctypes.c_int, [_VP, _D, _I64]),
        "psim_model_prepare": (ctypes.c_int, [_VP, ctypes.c_uint64, _D, _I64, _I64, ctypes.c_int32, _I32]),
        "psim_model_export": (ctypes.c_int, [_VP, ctypes.c_char_p, ctypes.c_double]),
        "psim_model_set_stream": (ctypes.c_int, [_VP, ctypes.c_int32]),
        "psim_batch_run": (ctypes.c_int, [ctypes.POINTER(_VP), ctypes.c_int32, _VP, ctypes.c_uint64,
                                          ctypes.POINTER(ctypes.c_int64), ctypes.c_char_p, ctypes.c_int]),
        "psim_model_upload": (ctypes.c_int, [_VP, _VP, ctypes.c_uint64, _D, _I64]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def exported_symbols_of_header(path: str = HEADER_PATH) -> list[str]:
    """Function names declared in include/psim_gpu.h."""
    import re
    text = open(path).read()
    return sorted(set(re.findall(r"^\s*(?:int|int64_t|void|const char\*)\s+(psim_[a-z_]+)\s*\(", text, re.M)))


def _p(a, t):
    return None if a is None else a.ctypes.data_as(t)


class Engine:
    """A psim_gpu context on one device (the GPU drop-in for psim's ModelSimulator)."""

    def __init__(self, device: int = 0):
        L = lib()
        h = _VP()
        rc = L.psim_gpu_create(ctypes.byref(h), int(device))
        if rc != 0:
            raise PsimError(f"psim_gpu_create failed ({rc}): no usable GPU {device}")
        self._h = h
        self._L = L

    @property
    def handle(self):
        return self._h

    def _check(self, rc: int) -> None:
        if rc != 0:
            raise PsimError(self._L.psim_gpu_last_error(self._h).decode())

    def close(self) -> None:
        if getattr(self, "_h", None):
            self._L.psim_gpu_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def set_fate(self, enable: bool) -> None:
        self._check(self._L.psim_gpu_set_fate(self._h, 1 if enable else 0))

    def set_window(self, steps: int) -> None:
        self._check(self._L.psim_gpu_set_window(self._h, int(steps)))

    def set_grid(self, blocks_per_cu: int) -> None:
        self._check(self._L.psim_gpu_set_grid(self._h, int(blocks_per_cu)))

    def set_block(self, threads: int) -> None:
        """Threads per transport block: 128, 256 (default) or 512."""
        self._check(self._L.psim_gpu_set_block(self._h, int(threads)))

    def launch_shape(self) -> dict:
        w, g, b = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
        self._check(self._L.psim_gpu_get_launch_shape(self._h, ctypes.byref(w), ctypes.byref(g), ctypes.byref(b)))
        return {"window": w.value, "blocks_per_cu": g.value, "block": b.value}

    def set_tally_mode(self, mode: str) -> None:
        """'auto', 'lds' or 'hbm': where recorded steps are tallied (psim_gpu_set_tally_mode)."""
        codes = {"auto": 0, "lds": 1, "hbm": 2}
        self._check(self._L.psim_gpu_set_tally_mode(self._h, codes[mode]))

    def set_tally_precision(self, precision: str) -> None:
        """'fp64' (default) or 'fp32_kahan': the flux tallies' arithmetic (psim_gpu_set_tally_precision)."""
        codes = {"fp64": 0, "fp32_kahan": 1}
        self._check(self._L.psim_gpu_set_tally_precision(self._h, codes[precision]))

    def run_resident(self, seed: int) -> None:
        self._check(self._L.psim_gpu_run_resident(self._h, int(seed)))

    def run_graph(self, n_sensors: int, num_record: int, seed: int):
        """The uploaded model's run replayed from a captured HIP graph: (counts, flux, counters)."""
        counts = np.zeros((n_sensors, num_record), np.int64)
        flux = np.zeros((n_sensors, num_record, 2), np.float64)
        counters = np.zeros(len(COUNTER_NAMES), np.int64)
        self._check(self._L.psim_gpu_run_graph(self._h, int(seed), _p(counts, _I64), _p(flux, _D),
                                               _p(counters, _I64)))
        return counts, flux, dict(zip(COUNTER_NAMES, counters.tolist()))

    def run_async(self, seed: int, stream: int) -> None:
        """Enqueues the uploaded run's graph on a HIP stream (its raw handle, e.g.
        torch.cuda.Stream().cuda_stream) and returns at once; after that stream is synchronised,
        fetch() reads the results."""
        self._check(self._L.psim_gpu_run_async(self._h, int(seed), ctypes.c_void_p(int(stream))))

    def autotune(self, seed: int = 0) -> dict:
        """Tries the launch shapes of psim_gpu_autotune on the uploaded model and keeps the fastest."""
        w = ctypes.c_int32()
        g = ctypes.c_int32()
        ms = ctypes.c_double()
        self._check(self._L.psim_gpu_autotune(self._h, int(seed), ctypes.byref(w), ctypes.byref(g), ctypes.byref(ms)))
        return {"window": w.value, "blocks_per_cu": g.value, "block": self.launch_shape()["block"], "run_ms": ms.value}

    def fetch(self, n_sensors: int, num_record: int):
        counts = np.zeros((n_sensors, num_record), np.int64)
        flux = np.zeros((n_sensors, num_record, 2), np.float64)
        counters = np.zeros(len(COUNTER_NAMES), np.int64)
        self._check(self._L.psim_gpu_fetch(self._h, _p(counts, _I64), _p(flux, _D), _p(counters, _I64)))
        return counts, flux, dict(zip(COUNTER_NAMES, counters.tolist()))

    def fetch_fate(self, n: int):
        step = np.zeros(n, np.int64)
        cause = np.zeros(n, np.int32)
        xy = np.zeros((n, 2), np.float64)
        self._check(self._L.psim_gpu_fetch_fate(self._h, _p(step, _I64), _p(cause, _I32), _p(xy, _D)))
        return step, cause, xy

    def last_timing(self):
        ms = ctypes.c_double()
        n = ctypes.c_int64()
        run = ctypes.c_double()
        self._check(self._L.psim_gpu_last_timing(self._h, ctypes.byref(ms), ctypes.byref(n), ctypes.byref(run)))
        return {"step_kernel_ms": ms.value, "step_launches": n.value, "run_ms": run.value}


@dataclass
class RunResults:
    measurements: np.ndarray  # [S, 6]: t_steady, std, x_flux, std, y_flux, std
    temps: np.ndarray  # [S, R]
    fluxes: np.ndarray  # [S, R, 2]
    counts: np.ndarray  # [S, R] raw tallies
    raw_flux: np.ndarray  # [S, R, 2]
    counters: dict


class Model:
    """A psim model held by the C++ host, read from the reference's JSON schema."""

    def __init__(self, model: dict | str):
        L = lib()
        text = model if isinstance(model, str) else _json.dumps(model)
        h = _VP()
        err = ctypes.create_string_buffer(2048)
        rc = L.psim_model_from_json(ctypes.byref(h), text.encode(), err, 2048)
        if rc != 0:
            raise PsimError(err.value.decode())
        self._h = h
        self._L = L
        ns = ctypes.c_int32()
        nr = ctypes.c_int64()
        runs = ctypes.c_int64()
        L.psim_model_sizes(h, ctypes.byref(ns), ctypes.byref(nr), ctypes.byref(runs))
        self.num_sensors, self.num_record, self.num_runs = ns.value, nr.value, runs.value

    def _check(self, rc: int) -> None:
        if rc != 0:
            raise PsimError(self._L.psim_model_last_error(self._h).decode())

    def __del__(self):
        if getattr(self, "_h", None):
            self._L.psim_model_destroy(self._h)
            self._h = None

    def prepare(self, seed: int):
        """Budget of one run without a GPU: (energy per phonon, total phonons, per-source counts)."""
        eff = ctypes.c_double()
        total = ctypes.c_int64()
        n = ctypes.c_int32()
        self._check(self._L.psim_model_prepare(self._h, int(seed), ctypes.byref(eff), ctypes.byref(total), None, 0,
                                               ctypes.byref(n)))
        counts = np.zeros(n.value, np.int64)
        self._check(self._L.psim_model_prepare(self._h, int(seed), None, None, _p(counts, _I64), n.value, None))
        return eff.value, total.value, counts

    def upload(self, engine: Engine, seed: int):
        """Prepare one run and upload it to the engine: (energy per phonon, total phonons)."""
        eff = ctypes.c_double()
        total = ctypes.c_int64()
        self._check(self._L.psim_model_upload(self._h, engine.handle, int(seed), ctypes.byref(eff),
                                              ctypes.byref(total)))
        return eff.value, total.value

    def run(self, engine: Engine, seed: int) -> None:
        self._check(self._L.psim_model_run(self._h, engine.handle, int(seed)))

    def results(self, run: int = 0) -> RunResults:
        S, R = self.num_sensors, self.num_record
        out = np.zeros((S, 6))
        temps = np.zeros((S, R))
        fluxes = np.zeros((S, R, 2))
        counts = np.zeros((S, R), np.int64)
        raw = np.zeros((S, R, 2))
        counters = np.zeros(len(COUNTER_NAMES), np.int64)
        self._check(self._L.psim_model_results(self._h, int(run), _p(out, _D), _p(temps, _D), _p(fluxes, _D),
                                               _p(counts, _I64), _p(raw, _D), _p(counters, _I64)))
        return RunResults(out, temps, fluxes, counts, raw, dict(zip(COUNTER_NAMES, counters.tolist())))

    def budget(self):
        eff = ctypes.c_double()
        total = ctypes.c_int64()
        self._check(self._L.psim_model_budget(self._h, ctypes.byref(eff), ctypes.byref(total)))
        return eff.value, total.value

    def export(self, json_path: str, time_s: float) -> None:
        self._check(self._L.psim_model_export(self._h, json_path.encode(), float(time_s)))

    def set_stream(self, stream: int) -> None:
        """Random-stream index of this model's phonons; stream m reproduces model m of a batch."""
        self._check(self._L.psim_model_set_stream(self._h, int(stream)))


def run_batch(engine: Engine, models: list, seed: int) -> dict:
    """Runs independent models as one engine upload and one run (psim_batch_run); each model's
    results are then read with model.results(0).  Returns the batch's counters."""
    L = lib()
    handles = (_VP * len(models))(*[m._h for m in models])
    counters = np.zeros(len(COUNTER_NAMES), np.int64)
    err = ctypes.create_string_buffer(2048)
    rc = L.psim_batch_run(handles, len(models), engine.handle, int(seed), _p(counters, _I64), err, 2048)
    if rc != 0:
        raise PsimError(err.value.decode())
    return dict(zip(COUNTER_NAMES, counters.tolist()))
