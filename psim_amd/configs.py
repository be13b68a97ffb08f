"""Model descriptions in psim's JSON input schema.

The schema is the one read by psim/src/inputManager.cpp:14-115 and written by
psim_python/psim/builder_tools.py: ``settings``, ``materials``, ``sensors``,
``cells`` (triangles with a sensor id and a specularity) and
``emit_surfaces``.  Units: nm, ns, K.  Rectangles are split into two
triangles exactly as Model::addCell(p1, p2) does
(psim/src/model.cpp:116-122, builder_tools.py:403-413).

The named workloads:
  linear_ss        1000x100 nm Si bar, 10 sensors, walls 310/290 K,
                   deviational steady state, 1e5 phonons, 2000 x 5 ps
  lshape_full      conforming L-shaped Si structure, 40 sensors, walls
                   320/280 K, specularity 0.3, full (non-deviational) steady
                   state, 1e7 phonons, 5000 x 2 ps
  film_transient   1500x150 nm Si film, 30 sensors, 20 K / 100 ps burst on
                   sensors 0-2, transient deviational, 1e8 phonons,
                   20000 x 1 ps, snapshots every 100 steps
"""
from __future__ import annotations

import copy
import json

# Holland silicon: psim_python/psim/pre_builts.py:16-18 (selected at :30).
SILICON_HOLLAND = {
    "name": "Silicon",
    "d_data": {
        "la_data": [-2.01e-7, 9.01e3, 0.0],
        "max_freq_la": 7.63916048e13,
        "ta_data": [-2.26e-7, 5.23e3, 0.0],
        "max_freq_ta": 3.0100793072e13,
    },
    "r_data": {"b_l": 2.0e-24, "b_tn": 9.3e-13, "b_tu": 5.5e-18, "b_i": 1.2e-45, "w": 2.417e13},
}


# Germanium as the reference ships it (psim_python/psim/pre_builts.py:25-28, "not used" there).
GERMANIUM = {
    "name": "Germanium",
    "d_data": {
        "la_data": [-1.50e-7, 5.63e3, 0.0],
        "max_freq_la": 4.45236386e13,
        "ta_data": [-1.13e-7, 2.60e3, 0.0],
        "max_freq_ta": 1.4937724175e13,
    },
    "r_data": {"b_l": 2.3e-24, "b_tn": 30.0e-13, "b_tu": 1.5e-18, "b_i": 0.0, "w": 1.23e13},
}


class ModelBuilder:
    """Accumulates a psim model and serialises it to the JSON schema."""

    def __init__(self, num_measurements: int, sim_time: float, num_phonons: int, t_eq: float,
                 sim_type: int = 0, step_interval: int = 0, num_runs: int = 1, phasor_sim: bool = False):
        self.settings = {
            "num_runs": int(num_runs),
            "num_measurements": int(num_measurements),
            "sim_time": float(sim_time),
            "num_phonons": int(num_phonons),
            "t_eq": float(t_eq),
            "sim_type": int(sim_type),
            "step_interval": int(step_interval),
            "phasor_sim": bool(phasor_sim),
        }
        self.materials = [copy.deepcopy(SILICON_HOLLAND)]
        self.sensors: list[dict] = []
        self.cells: list[dict] = []
        self.surfaces: list[dict] = []

    def add_sensor(self, t_init: float, material: str = "Silicon") -> int:
        sid = len(self.sensors)
        self.sensors.append({"id": sid, "material": material, "t_init": float(t_init)})
        return sid

    def add_triangle(self, p1, p2, p3, sensor_id: int, spec: float) -> None:
        def pt(p):
            return {"x": float(p[0]), "y": float(p[1])}
        self.cells.append({"triangle": {"p1": pt(p1), "p2": pt(p2), "p3": pt(p3)},
                           "sensorID": int(sensor_id), "specularity": float(spec)})

    def add_rectangle(self, lower_left, upper_right, sensor_id: int, spec: float) -> None:
        (x0, y0), (x1, y1) = lower_left, upper_right
        self.add_triangle((x0, y0), (x0, y1), (x1, y0), sensor_id, spec)
        self.add_triangle((x1, y1), (x1, y0), (x0, y1), sensor_id, spec)

    def add_surface(self, p1, p2, temp: float, duration: float | None = None, start_time: float = 0.0) -> None:
        if duration is None:
            duration = self.settings["sim_time"]
        self.surfaces.append({"p1": {"x": float(p1[0]), "y": float(p1[1])},
                              "p2": {"x": float(p2[0]), "y": float(p2[1])},
                              "temp": float(temp), "duration": float(duration),
                              "start_time": float(start_time)})

    def to_dict(self) -> dict:
        return {"settings": dict(self.settings), "materials": copy.deepcopy(self.materials),
                "sensors": copy.deepcopy(self.sensors), "cells": copy.deepcopy(self.cells),
                "emit_surfaces": copy.deepcopy(self.surfaces)}

    def to_json(self) -> str:
        return json.dumps(self.to_dict())


def linear_bar(num_sensors: int = 10, length: float = 1000.0, width: float = 100.0,
               t_hot: float = 310.0, t_cold: float = 290.0, t_init: float = 300.0, t_eq: float = 300.0,
               spec: float = 0.0, num_phonons: int = 100_000, num_measurements: int = 2000,
               sim_time: float = 10.0, sim_type: int = 0, step_interval: int = 0) -> ModelBuilder:
    """A bar split into equal rectangular sensors with emitting end walls."""
    b = ModelBuilder(num_measurements, sim_time, num_phonons, t_eq, sim_type, step_interval)
    dx = length / num_sensors
    for i in range(num_sensors):
        sid = b.add_sensor(t_init)
        b.add_rectangle((i * dx, 0.0), ((i + 1) * dx, width), sid, spec)
    b.add_surface((0.0, 0.0), (0.0, width), t_hot)
    b.add_surface((length, 0.0), (length, width), t_cold)
    return b


def needle_bar(count: int = 6, base: float = 2.0, length: float = 400.0, spec: float = 1.0,
               num_phonons: int = 20000, num_measurements: int = 1000, sim_time: float = 5.0) -> dict:
    """The 1000x100 nm bar (walls 310 / 290 K, deviational about 300 K) with `count` specular
    needles on its top edge, each a triangle of `base` nm on a whole top edge of its own column
    of the bar and `length` nm high.  A phonon running into a needle's tip meets its two walls
    more than MAX_COLLISIONS = 100 times in one drift, and the reference moves it to a random
    point of the cell (ModelSimulator::handleImpacts, psim/src/modelSimulator.cpp:215-218)."""
    b = ModelBuilder(num_measurements, sim_time, num_phonons, 300.0)
    width = 100.0
    xs = [1000.0 * (k + 1) / (count + 1) for k in range(count)]
    cuts = [0.0]
    for x in xs:
        cuts += [x - base / 2, x + base / 2]
    cuts.append(1000.0)
    sensors = []
    for i in range(len(cuts) - 1):
        sid = b.add_sensor(300.0)
        sensors.append(sid)
        b.add_rectangle((cuts[i], 0.0), (cuts[i + 1], width), sid, spec)
    for k, x in enumerate(xs):
        b.add_triangle((x - base / 2, width), (x + base / 2, width), (x, width + length), sensors[2 * k + 1], spec)
    b.add_surface((0.0, 0.0), (0.0, width), 310.0)
    b.add_surface((1000.0, 0.0), (1000.0, width), 290.0)
    return b.to_dict()


def si_ge_bar(num_sensors: int = 10, num_phonons: int = 20000, num_measurements: int = 1000, sim_time: float = 5.0,
              spec: float = 0.3, t_eq: float = 300.0, sim_type: int = 0, step_interval: int = 0,
              phasor_sim: bool = False) -> dict:
    """The linear bar with its second half germanium: every transition between the halves joins
    two materials, where a phonon above the other side's maximum frequency of its branch is
    scattered back diffusely (TransitionSurface::handlePhonon, psim/src/surface.cpp:71-109)."""
    b = ModelBuilder(num_measurements, sim_time, num_phonons, t_eq, sim_type, step_interval, phasor_sim=phasor_sim)
    b.materials.append(copy.deepcopy(GERMANIUM))
    length, width = 100.0 * num_sensors, 100.0
    dx = length / num_sensors
    for i in range(num_sensors):
        sid = b.add_sensor(300.0, "Silicon" if i < num_sensors // 2 else "Germanium")
        b.add_rectangle((i * dx, 0.0), ((i + 1) * dx, width), sid, spec)
    b.add_surface((0.0, 0.0), (0.0, width), 310.0)
    b.add_surface((length, 0.0), (length, width), 290.0)
    return b.to_dict()


def linear_ss(num_phonons: int = 100_000) -> dict:
    """1000x100 nm bar, 10 sensors, 310/290 K, deviational steady state."""
    return linear_bar(num_phonons=num_phonons).to_dict()


def lshape(num_phonons: int = 10_000_000, num_measurements: int = 5000, sim_time: float = 10.0,
           t_hot: float = 320.0, t_cold: float = 280.0, t_init: float = 300.0, t_eq: float = 0.0,
           spec: float = 0.3, sim_type: int = 0, step_interval: int = 0) -> ModelBuilder:
    """Conforming L: a horizontal arm of 19 sensors (100x200 nm), a 200x200 nm
    corner sensor, and a vertical arm of 20 sensors (200x100 nm); every shared
    edge is shared whole.  The hot wall closes the horizontal arm at x = 0, the
    cold wall closes the vertical arm at y = 2200."""
    b = ModelBuilder(num_measurements, sim_time, num_phonons, t_eq, sim_type, step_interval)
    w = 200.0
    n_h, n_v, seg = 19, 20, 100.0
    for i in range(n_h):
        sid = b.add_sensor(t_init)
        b.add_rectangle((i * seg, 0.0), ((i + 1) * seg, w), sid, spec)
    x_c = n_h * seg
    sid = b.add_sensor(t_init)
    b.add_rectangle((x_c, 0.0), (x_c + w, w), sid, spec)
    for j in range(n_v):
        sid = b.add_sensor(t_init)
        b.add_rectangle((x_c, w + j * seg), (x_c + w, w + (j + 1) * seg), sid, spec)
    b.add_surface((0.0, 0.0), (0.0, w), t_hot)
    y_top = w + n_v * seg
    b.add_surface((x_c, y_top), (x_c + w, y_top), t_cold)
    return b


def lshape_full(num_phonons: int = 10_000_000, num_measurements: int = 5000) -> dict:
    return lshape(num_phonons=num_phonons, num_measurements=num_measurements).to_dict()


def film_transient(num_phonons: int = 100_000_000, num_measurements: int = 20000, sim_time: float = 20.0,
                   burst_temp: float = 320.0, burst_duration: float = 0.1, step_interval: int = 100,
                   burst_start: float = 0.0) -> dict:
    """1500x150 nm film in 30 sensors of 50x150 nm at 300 K.  The heat burst is
    a set of transient emitting surfaces at 320 K active for [burst_start, burst_start +
    burst_duration] ns (default [0, 0.1]) on the outer edges of sensors 0-2; outside that window
    they reflect like plain walls.  The far end (x = 1500) is an absorbing 300 K wall.  Transient
    simulations must be deviational (psim/src/model.cpp:58-60)."""
    b = ModelBuilder(num_measurements, sim_time, num_phonons, 300.0, sim_type=2, step_interval=step_interval)
    seg, w = 50.0, 150.0
    for i in range(30):
        sid = b.add_sensor(300.0)
        b.add_rectangle((i * seg, 0.0), ((i + 1) * seg, w), sid, 0.0)
    b.add_surface((0.0, 0.0), (0.0, w), burst_temp, burst_duration, burst_start)
    for i in range(3):
        b.add_surface((i * seg, 0.0), ((i + 1) * seg, 0.0), burst_temp, burst_duration, burst_start)
        b.add_surface((i * seg, w), ((i + 1) * seg, w), burst_temp, burst_duration, burst_start)
    b.add_surface((1500.0, 0.0), (1500.0, w), 300.0)
    return b.to_dict()


def with_settings(model: dict, **settings) -> dict:
    """Copy of a model with some settings replaced (e.g. num_phonons)."""
    m = copy.deepcopy(model)
    m["settings"].update(settings)
    return m


def batch_geometry_params(rng) -> dict:
    """Draws the parameters of one batched-geometry structure from `rng` (random.Random):
      - a straight bar (length 200-5000 nm, width 20-500 nm), or a kinked one: two arms of half
        that length whose second arm turns by 15-90 degrees about the inner corner, joined by a
        wedge-shaped sensor; every shared edge is shared whole;
      - 4-200 sensors along its length (the wedge is one of them), each split into two triangles;
      - emitting walls across both ends at independent temperatures U(250, 350) K;
      - one specularity U(0, 1) for every cell, a phonon budget log-uniform in [1e4, 1e6]."""
    import math

    p = {"kinked": rng.random() < 0.5, "n_sensors": rng.randint(4, 200), "length": rng.uniform(200.0, 5000.0),
         "width": rng.uniform(20.0, 500.0), "t_hot": rng.uniform(250.0, 350.0), "t_cold": rng.uniform(250.0, 350.0),
         "spec": rng.uniform(0.0, 1.0), "phonons": int(round(10.0 ** rng.uniform(4.0, 6.0)))}
    p["theta"] = math.radians(rng.uniform(15.0, 90.0)) if p["kinked"] else 0.0
    return p


def geometry_from_params(p: dict, num_measurements: int = 2000, sim_time: float = 10.0, sim_type: int = 0,
                         step_interval: int = 0, t_init=None) -> dict:
    """The model of one batched-geometry structure: deviational about the mean wall temperature
    (t_eq = mean); every sensor starts at t_init[i] (default t_eq).  Sensors are numbered along the
    structure from the hot wall; a kinked structure has (n - 1) // 2 sensors in its first arm,
    then the wedge, then the second arm."""
    import math

    n_sensors, length, width, spec = p["n_sensors"], p["length"], p["width"], p["spec"]
    t_eq = 0.5 * (p["t_hot"] + p["t_cold"])
    b = ModelBuilder(num_measurements, sim_time, p["phonons"], t_eq, sim_type=sim_type, step_interval=step_interval)
    inits = [t_eq] * n_sensors if t_init is None else [float(v) for v in t_init]

    def quad(a, bb, c, d, sid):  # a-b-c-d counter-clockwise, split along a-c
        b.add_triangle(a, bb, c, sid, spec)
        b.add_triangle(a, c, d, sid, spec)

    if not p["kinked"]:
        xs = [length * i / n_sensors for i in range(n_sensors + 1)]
        for i in range(n_sensors):
            quad((xs[i], 0.0), (xs[i + 1], 0.0), (xs[i + 1], width), (xs[i], width), b.add_sensor(inits[i]))
        b.add_surface((0.0, 0.0), (0.0, width), p["t_hot"])
        b.add_surface((length, 0.0), (length, width), p["t_cold"])
        return b.to_dict()
    theta = p["theta"]
    n1 = (n_sensors - 1) // 2
    n2 = n_sensors - 1 - n1
    l1 = l2 = 0.5 * length
    xs = [l1 * i / n1 for i in range(n1 + 1)]
    for i in range(n1):
        quad((xs[i], 0.0), (xs[i + 1], 0.0), (xs[i + 1], width), (xs[i], width), b.add_sensor(inits[i]))
    q = (l1, width)  # inner corner; the second arm starts on edge a2-q, the first arm's end edge turned by theta
    a = (l1, 0.0)
    a2 = (q[0] + width * math.sin(theta), q[1] - width * math.cos(theta))
    b.add_triangle(a, a2, q, b.add_sensor(inits[n1]), spec)  # the wedge
    d = (math.cos(theta), math.sin(theta))
    lo = [(a2[0] + l2 * d[0] * j / n2, a2[1] + l2 * d[1] * j / n2) for j in range(n2 + 1)]
    hi = [(q[0] + l2 * d[0] * j / n2, q[1] + l2 * d[1] * j / n2) for j in range(n2 + 1)]
    for j in range(n2):
        quad(lo[j], lo[j + 1], hi[j + 1], hi[j], b.add_sensor(inits[n1 + 1 + j]))
    b.add_surface((0.0, 0.0), (0.0, width), p["t_hot"])
    b.add_surface(lo[n2], hi[n2], p["t_cold"])
    return b.to_dict()


def kinked_structure(theta_deg: float, n_sensors: int = 21, length: float = 2000.0, width: float = 200.0,
                     t_hot: float = 320.0, t_cold: float = 280.0, spec: float = 0.5, num_phonons: int = 100_000,
                     num_measurements: int = 2000, sim_time: float = 10.0) -> dict:
    """A kinked structure of the batched-geometry family with fixed parameters: two 1000 x 200 nm
    arms, the second turned by `theta_deg` about the inner corner and joined to the first by a
    wedge sensor, 10 sensors per arm, walls 320 / 280 K (deviational about 300 K), specularity
    0.5, 1e5 phonons, 2000 steps of 5 ps.  Its second arm's edges lie at arbitrary slopes, the
    general branch of Line::getIntersection (psim/src/geometry.cpp:104-138)."""
    import math

    p = {"kinked": True, "n_sensors": n_sensors, "length": length, "width": width, "t_hot": t_hot,
         "t_cold": t_cold, "spec": spec, "phonons": num_phonons, "theta": math.radians(theta_deg)}
    return geometry_from_params(p, num_measurements, sim_time)


def batch_geometry(rng, num_measurements: int = 2000, sim_time: float = 10.0) -> dict:
    """One structure of the batched-geometry workload (batch_geometry_params, geometry_from_params)."""
    return geometry_from_params(batch_geometry_params(rng), num_measurements, sim_time)


def batch_geometries(seed: int = 0, count: int = 256, num_measurements: int = 2000, sim_time: float = 10.0) -> list:
    """The batched-geometry workload: `count` independent structures (batch_geometry) from `seed`."""
    import random

    rng = random.Random(seed)
    return [batch_geometry(rng, num_measurements, sim_time) for _ in range(count)]
