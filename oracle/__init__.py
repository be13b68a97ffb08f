"""CPU oracle for psim's phonon path -- TEST INFRASTRUCTURE ONLY.

ctypes binding of liboracle_psim.so (a C restatement of the reference, built
by ``make -C oracle``).  Only the tests, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` use this package, as the checker or the
CPU baseline; the product path never does.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle_psim.so")
_lib = None

COUNTER_NAMES = ["phonon_steps", "scatters", "impacts", "relocations", "absorbed", "expired",
                 "not_started", "tallies", "stalled"]

_D = ctypes.POINTER(ctypes.c_double)
_I64 = ctypes.POINTER(ctypes.c_int64)
_I32 = ctypes.POINTER(ctypes.c_int32)


def build() -> None:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def _load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB_PATH):
        build()
    lib = ctypes.CDLL(_LIB_PATH)
    vp = ctypes.c_void_p
    lib.or_model_new.restype = vp
    lib.or_model_new.argtypes = [_D, ctypes.c_char_p, ctypes.c_int]
    lib.or_model_free.argtypes = [vp]
    lib.or_error.restype = ctypes.c_char_p
    lib.or_error.argtypes = [vp]
    lib.or_add_material.argtypes = [vp, ctypes.c_char_p, _D]
    lib.or_add_sensor.argtypes = [vp, ctypes.c_int64, ctypes.c_char_p, ctypes.c_double]
    lib.or_add_cell.argtypes = [vp, _D, ctypes.c_int64, ctypes.c_double]
    lib.or_add_emit.argtypes = [vp, _D]
    lib.or_prepare.argtypes = [vp, ctypes.c_uint64]
    lib.or_set_step_by_step.argtypes = [vp, ctypes.c_int]
    lib.or_set_stream.argtypes = [vp, ctypes.c_int]
    lib.or_plant_flight_fault.argtypes = [vp, ctypes.c_int]
    lib.or_num_sensors.argtypes = [vp]
    lib.or_num_record.restype = ctypes.c_int64
    lib.or_num_record.argtypes = [vp]
    lib.or_total_phonons.restype = ctypes.c_int64
    lib.or_total_phonons.argtypes = [vp]
    lib.or_eff_energy.restype = ctypes.c_double
    lib.or_eff_energy.argtypes = [vp]
    lib.or_num_sources.argtypes = [vp]
    lib.or_sources.argtypes = [vp, _I64, _I32, _D, _I64]
    lib.or_run.argtypes = [vp, ctypes.c_uint64, ctypes.c_int, _I64, _D, _I64, _I64, _I32, _D]
    lib.or_interpret.argtypes = [vp, _I64, _D, _D, _D, _D]
    lib.or_table.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_double, _D, _D, _D, _I64]
    lib.or_material_arrays.argtypes = [vp, ctypes.c_int, _D, _D, _D, _D, _D]
    lib.or_num_temps.restype = ctypes.c_int64
    lib.or_num_temps.argtypes = [vp, ctypes.c_int]
    lib.or_temps.argtypes = [vp, ctypes.c_int, _D, _D]
    lib.or_cell_geometry.argtypes = [vp, ctypes.c_int, _D, _D, _D, _D, _I32, _I32]
    _U32 = ctypes.POINTER(ctypes.c_uint32)
    lib.or_math.argtypes = [ctypes.c_int, _D, _D, ctypes.c_int64]
    lib.or_philox.argtypes = [_U32, _U32, _U32]
    lib.or_philox.restype = None
    lib.or_stream.argtypes = [ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int64,
                              ctypes.c_int, _D]
    lib.or_stream.restype = None
    lib.or_freq_index.argtypes = [_D, _D, ctypes.c_double, ctypes.c_double, _I64, _I32]
    lib.or_freq_index.restype = None
    lib.or_directions.argtypes = [ctypes.c_int, ctypes.c_uint64, ctypes.c_int64, ctypes.c_double, ctypes.c_double, _D]
    lib.or_directions.restype = None
    _lib = lib
    return lib


def _p(a, t):
    return a.ctypes.data_as(t) if a is not None else None


class OracleError(RuntimeError):
    pass


@dataclass
class RunResult:
    counts: np.ndarray  # [S, R] int64
    flux: np.ndarray  # [S, R, 2]
    counters: dict
    fate_step: np.ndarray | None = None
    fate_cause: np.ndarray | None = None
    fate_xy: np.ndarray | None = None


class Model:
    """A psim model built by the oracle from a dict in the psim JSON schema."""

    def __init__(self, model: dict):
        lib = _load()
        self._lib = lib
        s = model["settings"]
        # InputManager passes step_interval only for periodic and transient runs (inputManager.cpp:77-78)
        step_interval = s.get("step_interval", 0) if int(s["sim_type"]) in (1, 2) else 0
        settings = np.array([s.get("num_runs", 1), s["num_measurements"], s["num_phonons"], s["sim_time"],
                             s["t_eq"], s["sim_type"], step_interval,
                             1.0 if s.get("phasor_sim", False) else 0.0], dtype=np.float64)
        err = ctypes.create_string_buffer(512)
        h = lib.or_model_new(_p(settings, _D), err, 512)
        if not h:
            raise OracleError(err.value.decode())
        self._h = h
        self.material_names = []
        for m in model["materials"]:
            d, r = m["d_data"], m["r_data"]
            p = np.array(list(d["la_data"]) + list(d["ta_data"]) + [d["max_freq_la"], d["max_freq_ta"],
                         r["b_l"], r["b_tn"], r["b_tu"], r["b_i"], r["w"]], dtype=np.float64)
            self._check(lib.or_add_material(h, m["name"].encode(), _p(p, _D)))
            self.material_names.append(m["name"])
        for x in model["sensors"]:
            self._check(lib.or_add_sensor(h, int(x["id"]), x["material"].encode(), float(x["t_init"])))
        for c in model["cells"]:
            t = c["triangle"]
            tri = np.array([t["p1"]["x"], t["p1"]["y"], t["p2"]["x"], t["p2"]["y"], t["p3"]["x"], t["p3"]["y"]],
                           dtype=np.float64)
            self._check(lib.or_add_cell(h, _p(tri, _D), int(c["sensorID"]), float(c["specularity"])))
        for e in model["emit_surfaces"]:
            arr = np.array([e["p1"]["x"], e["p1"]["y"], e["p2"]["x"], e["p2"]["y"], e["temp"], e["duration"],
                            e["start_time"]], dtype=np.float64)
            self._check(lib.or_add_emit(h, _p(arr, _D)))
        self.seed = None

    def _check(self, rc):
        if rc != 0:
            raise OracleError(self._lib.or_error(self._h).decode())

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            self._lib.or_model_free(h)
            self._h = None

    def set_stream(self, stream: int) -> None:
        """Random-stream index of every draw: stream m reproduces model m of a batched run."""
        self._check(self._lib.or_set_stream(self._h, int(stream)))
        self.seed = None

    def plant_flight_fault(self, kind: int) -> None:
        """Test hook: a known mistake in the tallies of the boundaries a flight passes (1: one index
        late, 2: into the next sensor, 0: none)."""
        self._check(self._lib.or_plant_flight_fault(self._h, int(kind)))

    def prepare(self, seed: int) -> None:
        self._check(self._lib.or_prepare(self._h, int(seed)))
        self.seed = int(seed)

    @property
    def num_sensors(self) -> int:
        return self._lib.or_num_sensors(self._h)

    @property
    def num_record(self) -> int:
        return self._lib.or_num_record(self._h)

    @property
    def total_phonons(self) -> int:
        return self._lib.or_total_phonons(self._h)

    @property
    def eff_energy(self) -> float:
        return self._lib.or_eff_energy(self._h)

    def sources(self):
        n = self._lib.or_num_sources(self._h)
        cell = np.zeros(n, np.int64)
        kind = np.zeros(n, np.int32)
        energy = np.zeros(n, np.float64)
        count = np.zeros(n, np.int64)
        self._lib.or_sources(self._h, _p(cell, _I64), _p(kind, _I32), _p(energy, _D), _p(count, _I64))
        return {"cell": cell, "kind": kind, "energy": energy, "count": count}

    def run(self, seed: int, nthreads: int = 1, fate: bool = False, step_by_step: bool = False) -> RunResult:
        """step_by_step: the reference's loop as written (one drift per measurement step) in place
        of flights across the boundaries; the two agree up to rounding."""
        if self.seed != int(seed):
            self.prepare(seed)
        self._lib.or_set_step_by_step(self._h, int(bool(step_by_step)))
        S, R = self.num_sensors, self.num_record
        counts = np.zeros((S, R), np.int64)
        flux = np.zeros((S, R, 2), np.float64)
        counters = np.zeros(len(COUNTER_NAMES), np.int64)
        fs = fc = fxy = None
        if fate:
            n = self.total_phonons
            fs = np.zeros(n, np.int64)
            fc = np.zeros(n, np.int32)
            fxy = np.zeros((n, 2), np.float64)
        self._check(self._lib.or_run(self._h, int(seed), int(nthreads), _p(counts, _I64), _p(flux, _D),
                                     _p(counters, _I64), _p(fs, _I64), _p(fc, _I32), _p(fxy, _D)))
        return RunResult(counts, flux, dict(zip(COUNTER_NAMES, counters.tolist())), fs, fc, fxy)

    def interpret(self, counts: np.ndarray, flux: np.ndarray):
        S, R = self.num_sensors, self.num_record
        counts = np.ascontiguousarray(counts, dtype=np.int64)
        flux = np.ascontiguousarray(flux, dtype=np.float64)
        out = np.zeros((S, 6), np.float64)
        temps = np.zeros((S, R), np.float64)
        fluxes = np.zeros((S, R, 2), np.float64)
        self._check(self._lib.or_interpret(self._h, _p(counts, _I64), _p(flux, _D), _p(out, _D), _p(temps, _D),
                                           _p(fluxes, _D)))
        return out, temps, fluxes

    def table(self, kind: int, temp: float, material: int = 0):
        cdf = np.zeros(1000)
        frac = np.zeros(1000)
        s = np.zeros(1)
        idx = np.zeros(1, np.int64)
        self._check(self._lib.or_table(self._h, material, kind, float(temp), _p(cdf, _D), _p(frac, _D), _p(s, _D),
                                       _p(idx, _I64)))
        return cdf, frac, float(s[0]), int(idx[0])

    def material_arrays(self, material: int = 0):
        arrs = [np.zeros(1000) for _ in range(5)]
        self._check(self._lib.or_material_arrays(self._h, material, *[_p(a, _D) for a in arrs]))
        return dict(zip(["freq", "vel_la", "vel_ta", "dens_la", "dens_ta"], arrs))

    def temps(self, material: int = 0):
        n = self._lib.or_num_temps(self._h, material)
        t = np.zeros(n)
        b = np.zeros(n)
        self._lib.or_temps(self._h, material, _p(t, _D), _p(b, _D))
        return t, b

    def cell_geometry(self, cell: int):
        area = np.zeros(1)
        lines = np.zeros((3, 7))
        normals = np.zeros((3, 2))
        spec = np.zeros(3)
        nt = np.zeros(3, np.int32)
        ne = np.zeros(3, np.int32)
        self._lib.or_cell_geometry(self._h, int(cell), _p(area, _D), _p(lines, _D), _p(normals, _D), _p(spec, _D),
                                   _p(nt, _I32), _p(ne, _I32))
        return {"area": float(area[0]), "lines": lines, "normals": normals, "spec": spec, "n_trans": nt, "n_emit": ne}


# ---- test hooks -------------------------------------------------------------------------

MATH_FUNCTIONS = {"log": 0, "exp": 1, "sinh": 2, "cos2pi": 3, "pow4": 4}


def math_fn(name: str, x: np.ndarray) -> np.ndarray:
    lib = _load()
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.zeros_like(x)
    if lib.or_math(MATH_FUNCTIONS[name], _p(x, _D), _p(y, _D), x.size) != 0:
        raise OracleError("bad function")
    return y


def philox(ctr, key) -> np.ndarray:
    lib = _load()
    c = np.ascontiguousarray(ctr, dtype=np.uint32)
    k = np.ascontiguousarray(key, dtype=np.uint32)
    out = np.zeros(4, np.uint32)
    U32 = ctypes.POINTER(ctypes.c_uint32)
    lib.or_philox(c.ctypes.data_as(U32), k.ctypes.data_as(U32), out.ctypes.data_as(U32))
    return out


def stream(seed: int, tag: int, ident: int, step: int, n: int, pairs: bool = False) -> np.ndarray:
    """n single draws of a stream, or n / 2 pair draws (pairs=True)."""
    lib = _load()
    out = np.zeros(n, np.float64)
    lib.or_stream(int(seed), int(tag), int(ident), int(step), int(n), int(bool(pairs)), _p(out, _D))
    return out


def freq_index(cdf, frac, u1: float, u2: float):
    lib = _load()
    c = np.ascontiguousarray(cdf, dtype=np.float64)
    f = np.ascontiguousarray(frac, dtype=np.float64)
    b = np.zeros(1, np.int64)
    p = np.zeros(1, np.int32)
    lib.or_freq_index(_p(c, _D), _p(f, _D), float(u1), float(u2), _p(b, _I64), _p(p, _I32))
    return int(b[0]), int(p[0])


def directions(kind: int, seed: int, n: int, normal=(1.0, 0.0), dirs=None) -> np.ndarray:
    lib = _load()
    out = np.zeros((n, 2), np.float64) if dirs is None else np.ascontiguousarray(dirs, dtype=np.float64).copy()
    lib.or_directions(int(kind), int(seed), int(n), float(normal[0]), float(normal[1]), _p(out, _D))
    return out
