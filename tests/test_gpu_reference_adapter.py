"""The engine as a drop-in inside the reference itself.

oracle/_ref/psim_ref_gpu is the reference's own Model, sensors and SensorInterpreter
(psim/src/model.cpp, sensor.cpp, sensorInterpreter.cpp, ...) built with the engine's ModelSimulator
adapter (oracle/ref_adapter/) in place of psim/src/modelSimulator.cpp, linked with
libpsim_amd.so (`make -C oracle ref-gpu`, in the build container; the binary travels with the tree).
Model::runSimulation calls the adapter's initPhononBuilders and runSimulation exactly where it calls
the reference's (model.cpp:160-161).  With the same seed, the tallies it leaves in the reference's
sensors must equal psim_model_run's (integer counts exactly, flux sums to 1e-9 of the largest entry,
their summation order differs), and the reference's own SensorInterpreter must give the host's
temperatures bit for bit.  For steady-state models the reference's own OutputManager then writes
its ss_ file from those tallies (main.cpp:21, outputManager.cpp:13-38), and psim_model_export's file
must equal it byte for byte but for the wall-clock date in the title line.
"""
import json
import os
import re
import subprocess

import numpy as np
import pytest

import psim_amd
from oracle import refmodel
from psim_amd import configs

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "oracle", "_ref", "psim_ref_gpu")

CASES = {
    "bar_ss_deviational": configs.linear_bar(num_phonons=20000, num_measurements=1000, sim_time=5.0, spec=0.3).to_dict(),
    "lshape_full_mode": configs.lshape_full(num_phonons=20000, num_measurements=1000),
    "film_transient": configs.film_transient(num_phonons=20000, num_measurements=1000, sim_time=1.0, step_interval=50),
    "bar_si_ge": configs.si_ge_bar(),
    "kinked_15": configs.kinked_structure(15.0, num_phonons=20000),
}


@pytest.fixture(scope="module")
def engine():
    e = psim_amd.Engine(0)
    yield e
    e.close()


@pytest.mark.skipif(not os.path.exists(BIN), reason="oracle/_ref/psim_ref_gpu is built only where the reference "
                                                    "sources are (make -C oracle ref-gpu)")
@pytest.mark.parametrize("name", sorted(CASES))
def test_reference_model_runs_the_engine(engine, name, tmp_path):
    cfg = CASES[name]
    seed = 17
    print(f"case {name} seed {seed}", flush=True)
    txt = tmp_path / "model.txt"
    out = tmp_path / "out.json"
    txt.write_text(refmodel.json_to_text(cfg))
    env = dict(os.environ, PSIM_REF_GPU_SEED=str(seed))
    p = subprocess.run([BIN, str(txt), str(out)], env=env, capture_output=True, text=True, timeout=300)
    ref = json.loads(out.read_text())
    assert p.returncode == 0 and ref["ok"], (p.stdout[-500:], p.stderr[-500:], ref.get("error"))
    model = psim_amd.Model(cfg)
    model.run(engine, seed)
    res = model.results(0)
    _, total = model.budget()
    assert ref["total_phonons"] == total
    counts = np.array([s["energies"] for s in ref["sensors"]], dtype=np.int64)
    flux = np.array([s["fluxes"] for s in ref["sensors"]], dtype=np.float64)
    np.testing.assert_array_equal(counts, res.counts)
    scale = max(1.0, float(np.max(np.abs(res.raw_flux))))
    assert np.max(np.abs(flux - res.raw_flux)) <= 1e-9 * scale
    t_ref = np.array([m["t_steady"] for m in ref["measurements"]])
    np.testing.assert_array_equal(t_ref, res.measurements[:, 0])
    temps = np.array([m["final_temps"] for m in ref["measurements"]])
    np.testing.assert_array_equal(temps, res.temps)


# The title's date is the wall clock of the writing process (outputManager.cpp:132-145).
DATE = re.compile(rb" @ \d{4}-\d{2}-\d{2} \d{2}:\d{2}:\d{2} - ")

SS_CASES = {
    "bar_two_runs": configs.with_settings(configs.linear_bar(num_phonons=20000, num_measurements=1000, sim_time=5.0,
                                                             spec=0.3).to_dict(), num_runs=2),
    "lshape_full_mode": CASES["lshape_full_mode"],
    "bar_si_ge": CASES["bar_si_ge"],
    "kinked_15": CASES["kinked_15"],
}


@pytest.mark.skipif(not os.path.exists(BIN), reason="oracle/_ref/psim_ref_gpu is built only where the reference "
                                                    "sources are (make -C oracle ref-gpu)")
@pytest.mark.parametrize("name", sorted(SS_CASES))
def test_ss_file_equals_the_reference_writer(engine, name, tmp_path):
    cfg = SS_CASES[name]
    assert cfg["settings"]["sim_type"] == 0
    seed = 23
    print(f"case {name} seed {seed}", flush=True)
    (tmp_path / "ref").mkdir()
    (tmp_path / "ours").mkdir()
    txt = tmp_path / "model.txt"
    txt.write_text(refmodel.json_to_text(cfg))
    env = dict(os.environ, PSIM_REF_GPU_SEED=str(seed))
    p = subprocess.run([BIN, str(txt), str(tmp_path / "out.json"), "--export", str(tmp_path / "ref" / f"{name} model.json"),
                        "12.5"], env=env, capture_output=True, text=True, timeout=300)
    assert p.returncode == 0, (p.stdout[-500:], p.stderr[-500:])
    model = psim_amd.Model(cfg)
    model.run(engine, seed)
    model.export(str(tmp_path / "ours" / f"{name} model.json"), 12.5)
    ref = (tmp_path / "ref" / f"ss_{name} model.txt").read_bytes()
    ours = (tmp_path / "ours" / f"ss_{name} model.txt").read_bytes()
    assert DATE.search(ref) and DATE.search(ours)
    assert ref.count(b"\n") == 1 + len(cfg["sensors"])
    assert DATE.sub(b" @ - ", ref) == DATE.sub(b" @ - ", ours)
