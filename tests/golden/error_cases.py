"""Broken models (psim JSON schema) whose construction the reference rejects.

Shared by make_ref_fixtures.py, which records the reference's message for each case in
ref_errors.json, and by the host tests, which require the same message.
"""
import copy

from psim_amd import configs


def _bar():
    return configs.linear_ss(num_phonons=1000)


def cases():
    out = {}
    c = configs.film_transient(num_phonons=1000, num_measurements=100, sim_time=1.0)
    c["settings"]["t_eq"] = 0.0
    out["transient_full_mode"] = c
    out["periodic_zero_interval"] = configs.linear_bar(num_phonons=1000, sim_type=1, step_interval=0).to_dict()
    c = _bar()
    t = copy.deepcopy(c["cells"][0])
    t["triangle"]["p1"] = {"x": 10.0, "y": 10.0}
    c["cells"].append(t)
    out["cell_inside_cell"] = c
    c = _bar()
    c["cells"].append({"triangle": {"p1": {"x": 50.0, "y": -50.0}, "p2": {"x": 50.0, "y": 150.0},
                                    "p3": {"x": 60.0, "y": 50.0}}, "sensorID": 0, "specularity": 0.0})
    out["crossing_cell"] = c
    c = _bar()
    c["cells"].append(copy.deepcopy(c["cells"][3]))
    out["duplicate_cell"] = c
    c = _bar()
    c["cells"][5]["triangle"]["p3"] = {"x": 200.0, "y": 200.0}
    c["cells"][5]["triangle"]["p2"] = {"x": 150.0, "y": 150.0}
    c["cells"][5]["triangle"]["p1"] = {"x": 100.0, "y": 100.0}
    out["collinear_triangle"] = c
    c = _bar()
    c["emit_surfaces"].append({"p1": {"x": 500.0, "y": 20.0}, "p2": {"x": 500.0, "y": 60.0}, "temp": 300.0,
                               "duration": 10.0, "start_time": 0.0})
    out["emit_off_mesh"] = c
    c = _bar()
    c["emit_surfaces"][0]["duration"] = 5.0
    out["transient_surface_in_steady_state"] = c
    c = _bar()
    c["emit_surfaces"][0]["duration"] = 50.0
    out["emit_duration_too_long"] = c
    c = _bar()
    c["emit_surfaces"].append(copy.deepcopy(c["emit_surfaces"][0]))
    out["emit_overlaps_emit"] = c
    c = _bar()
    c["cells"][0]["sensorID"] = 99
    out["unknown_sensor"] = c
    c = _bar()
    c["sensors"][1]["id"] = 0
    out["duplicate_sensor"] = c
    c = _bar()
    c["sensors"][2]["material"] = "Germanium"
    out["unknown_material"] = c
    c = _bar()
    c["materials"].append(copy.deepcopy(c["materials"][0]))
    out["duplicate_material"] = c
    return out
