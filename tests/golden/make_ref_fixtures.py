"""Write the reference fixtures under tests/golden/ (run in the build container only).

Builds the reference psim from /root/reference with `make -C oracle ref` (its
own sources plus oracle/ref_driver.cpp) and runs it on these workloads:

  linear_ss     the 1000x100 nm bar, deviational steady state, 1e5 phonons, 2000 steps
  lshape_r      the conforming L, full mode, reduced to 1e5 phonons (5000 steps)
  film_r        the transient film, reduced to 1e5 phonons and 2000 steps over 2 ns
  linear_k      the linear_ss bar at 1e6 phonons over 48 runs (statistics only, with every
                run's fluxes): pins the effective conductivity to a few tenths of a percent
  si_ge         the bar with its second half germanium (configs.si_ge_bar), 1e5 phonons,
                2000 steps: transitions between two materials (statistics only)
  phasor        the bar in phasor mode (statistics only)
  needles       the bar with six specular needles (configs.needle_bar): acute corners where
                stuck phonons are relocated (statistics only)
  linear_b      the linear_ss bar at 6e6 phonons, above the reference's PHONON_CUTOFF (5e6), so
                its runs take the builder path (statistics only)
  lshape_full   the L-shape at the benchmark's 1e7 phonons (statistics only; ~25 core-minutes
                per run)
  kinked_15, kinked_90  kinked structures of the batched-geometry family (statistics only)

For each workload it stores
  ref_<name>.json  the model, and over N independent reference runs (the
                   reference RNG cannot be seeded, utils.h:16-21) the per-sensor
                   mean and sd of t_steady / x_flux / y_flux, the per-block
                   (step_interval) temperature series for transient runs, the
                   budget's energy per phonon and the wall times;
  ref_errors.json  the reference's error message for each broken model of error_cases.py;
  ref_<name>_state.npz  the reference's material tables at the workload's
                   temperatures, its temperature grid, one run's raw tallies
                   with the SensorInterpreter results computed from them, and
                   the cell geometry -- bit-exact vectors for the oracle.

Usage: python tests/golden/make_ref_fixtures.py [--runs 16] [--jobs 8]
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import tempfile
from concurrent.futures import ThreadPoolExecutor

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import refmodel  # noqa: E402
from psim_amd import configs  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))
REF_BIN = os.path.join(ROOT, "oracle", "_ref", "psim_ref")


def workloads():
    return {
        "linear_ss": (configs.linear_ss(), [300.0, 310.0, 290.0]),
        "lshape_r": (configs.lshape_full(num_phonons=100_000), [300.0, 320.0, 280.0]),
        "film_r": (configs.film_transient(num_phonons=100_000, num_measurements=2000, sim_time=2.0),
                   [300.0, 320.0]),
        "linear_k": (configs.linear_ss(num_phonons=1_000_000), None),
        "si_ge": (configs.si_ge_bar(num_phonons=100_000, num_measurements=2000, sim_time=10.0), None),
        "phasor": (_phasor_bar(), None),
        "needles": (configs.needle_bar(num_phonons=100_000, num_measurements=2000, sim_time=10.0), None),
        # the linear_ss bar above the reference's PHONON_CUTOFF of 5e6 (modelSimulator.cpp:18), so
        # that its runs take runUsingBuilders (:245-253) instead of runPhononByPhonon
        "linear_b": (configs.linear_ss(num_phonons=6_000_000), None),
        # the benchmark workload at its own size (1e7 phonons, builder path): in full mode the
        # interpreter's bisection is bounded by [min - 10, max + 10] K (sensorInterpreter.cpp:87-95),
        # and the sensors next to the hot wall reach that bound, so their mean temperature depends
        # on the per-step noise and with it on the phonon count; only a fixture of this size
        # pins them
        "lshape_full": (configs.lshape_full(num_phonons=10_000_000), None),
        # kinked structures of the batched-geometry family: arms at 15 and 90 degrees
        "kinked_15": (configs.kinked_structure(15.0), None),
        "kinked_90": (configs.kinked_structure(90.0), None),
    }


def _phasor_bar() -> dict:
    m = configs.linear_bar(num_phonons=100_000, num_measurements=2000, sim_time=10.0, spec=0.3).to_dict()
    m["settings"]["phasor_sim"] = True
    return m


RUNS = {"linear_k": 48}  # runs per workload where the default is not enough


def run_ref(model: dict, workdir: str, tag: str, extra=()) -> dict:
    txt = os.path.join(workdir, f"{tag}.txt")
    out = os.path.join(workdir, f"{tag}.json")
    with open(txt, "w") as f:
        f.write(refmodel.json_to_text(model))
    subprocess.run([REF_BIN, txt, out, *extra], check=True, stdout=subprocess.DEVNULL)
    with open(out) as f:
        return json.load(f)


def block_series(final_temps: list, interval: int) -> list:
    """Per-block means as OutputManager::periodicExport forms them (outputManager.cpp:86-111)."""
    n = len(final_temps)
    out = []
    step = 0
    while step < n - interval + 1:
        out.append(float(np.sum(final_temps[step:step + interval]) / interval))
        step += interval
    return out


def write_error_fixture() -> None:
    """ref_errors.json: the message the reference gives for each broken model (newlines as spaces,
    double quotes as single quotes, as ref_driver.cpp writes them)."""
    sys.path.insert(0, HERE)
    import error_cases

    out = {}
    with tempfile.TemporaryDirectory() as wd:
        for name, model in error_cases.cases().items():
            txt = os.path.join(wd, "m.txt")
            res = os.path.join(wd, "m.json")
            with open(txt, "w") as f:
                f.write(refmodel.json_to_text(model))
            subprocess.run([REF_BIN, txt, res], stdout=subprocess.DEVNULL)
            with open(res) as f:
                r = json.load(f)
            assert not r["ok"], name
            out[name] = r["error"]
    with open(os.path.join(HERE, "ref_errors.json"), "w") as f:
        json.dump(out, f, indent=1)


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--runs", type=int, default=16)
    ap.add_argument("--jobs", type=int, default=8)
    ap.add_argument("--only", default=None)
    args = ap.parse_args()
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "ref"], check=True)
    if args.only in (None, "errors"):
        write_error_fixture()
        if args.only == "errors":
            return
    for name, (model, temps) in workloads().items():
        if args.only and name != args.only:
            continue
        n_runs = RUNS.get(name, args.runs)
        with tempfile.TemporaryDirectory() as wd:
            state = None
            if temps is not None:
                state = run_ref(model, wd, "state", ["--tables", ",".join(str(t) for t in temps), "--geometry"])
            with ThreadPoolExecutor(args.jobs) as ex:
                runs = list(ex.map(lambda i: run_ref(model, wd, f"run{i}"), range(n_runs)))
        s = model["settings"]
        keys = ["t_steady", "x_flux", "y_flux"]
        per = {k: np.array([[m[k] for m in r["measurements"]] for r in runs]) for k in keys}
        fix = {
            "name": name,
            "model": model,
            "runs": n_runs,
            "eff_energy0": [r["eff_energy0"] for r in runs],
            "eff_energy": [r["eff_energy"] for r in runs],
            "total_phonons": [r["total_phonons"] for r in runs],
            "wall_s": [r["wall_s"] for r in runs],
            "mean": {k: per[k].mean(axis=0).tolist() for k in keys},
            "sd": {k: per[k].std(axis=0, ddof=1).tolist() for k in keys},
        }
        if s["sim_type"] != 0:
            interval = int(s["step_interval"])
            series = np.array([[block_series(m["final_temps"], interval) for m in r["measurements"]] for r in runs])
            fix["block_interval"] = interval
            fix["block_mean"] = series.mean(axis=0).tolist()
            fix["block_sd"] = series.std(axis=0, ddof=1).tolist()
        if state is None:  # statistics only: keep every run's values
            fix["per_run"] = {k: per[k].tolist() for k in keys}
        with open(os.path.join(HERE, f"ref_{name}.json"), "w") as f:
            json.dump(fix, f)
        if state is None:
            print(name, "mean wall", np.mean(fix["wall_s"]), "s")
            continue
        arrays = {
            "temps": np.array(state["temps"]),
            "base_sums": np.array(state["base_sums"]),
            "frequencies": np.array(state["frequencies"]),
            "vel_la": np.array(state["vel_la"]),
            "vel_ta": np.array(state["vel_ta"]),
            "table_temps": np.array([t["temp"] for t in state["tables"]]),
            "table_index": np.array([t["index"] for t in state["tables"]]),
            "counts": np.array([x["energies"] for x in state["sensors"]], dtype=np.int64),
            "flux": np.array([x["fluxes"] for x in state["sensors"]], dtype=np.float64),
            "areas": np.array([x["area"] for x in state["sensors"]]),
            "eff_energy0": np.array(state["eff_energy0"]),
            "eff_energy": np.array(state["eff_energy"]),
            "measurements": np.array([[m[k] for k in ["t_steady", "std_t_steady", "x_flux", "std_x_flux",
                                                      "y_flux", "std_y_flux"]] for m in state["measurements"]]),
            "final_temps": np.array([m["final_temps"] for m in state["measurements"]]),
            "cell_area": np.array([c["area"] for c in state["cells"]]),
            "cell_lines": np.array([[e["line"] for e in c["edges"]] for c in state["cells"]]),
            "cell_normals": np.array([[e["normal"] for e in c["edges"]] for c in state["cells"]]),
            "cell_subs": np.array([[[len(e["transitions"]), len(e["emits"])] for e in c["edges"]]
                                   for c in state["cells"]]),
        }
        for kind in ["base", "emit", "scatter"]:
            arrays[f"{kind}_cdf"] = np.array([t[kind]["cdf"] for t in state["tables"]])
            arrays[f"{kind}_frac"] = np.array([[np.nan if v == "nan" else v for v in t[kind]["la_frac"]]
                                              for t in state["tables"]], dtype=np.float64)
            arrays[f"{kind}_sum"] = np.array([t[kind]["sum"] for t in state["tables"]])
        np.savez_compressed(os.path.join(HERE, f"ref_{name}_state.npz"), **arrays)
        print(name, "mean wall", np.mean(fix["wall_s"]), "s")


if __name__ == "__main__":
    main()
