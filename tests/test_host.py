"""C ABI and C++17 host (CPU only: nothing here launches a kernel).

- libpsim_amd.so loads and exports every function include/psim_gpu.h and
  include/psim_surrogate.h declare;
- the host's budget (bounds, tables, energy per phonon, Bernoulli-rounded counts per
  builder) equals the oracle's for several seeds, so the engine and the oracle start from
  the same phonons;
- model construction rejects what the reference rejects, with the reference's messages
  (psim/src/model.cpp:46-138, cell.cpp:21-25, compositeSurface.cpp:68-87, inputManager.cpp);
- the JSON reader accepts the reference's shipped example input.
"""
import ctypes
import json
import os
import sys

import numpy as np
import pytest

import oracle
import psim_amd
from psim_amd import configs

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
sys.path.insert(0, GOLDEN)
import error_cases  # noqa: E402


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(psim_amd.LIB_PATH)
    names = psim_amd.exported_symbols_of_header()
    assert len(names) >= 20
    surrogate_names = psim_amd.exported_symbols_of_header(os.path.join(os.path.dirname(psim_amd.HEADER_PATH),
                                                                       "psim_surrogate.h"))
    assert len(surrogate_names) == 14  # 11, psim_mlp_forward_noalloc, psim_mlp_time_product, psim_mlp_set_graph
    names += surrogate_names
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing


def test_engine_reports_missing_gpu_cleanly():
    """Without a GPU (or with a bad device index) context creation fails with an error code."""
    L = psim_amd.lib()
    h = ctypes.c_void_p()
    rc = L.psim_gpu_create(ctypes.byref(h), 1 << 20)
    assert rc != 0 and not h.value


@pytest.mark.parametrize("cfg", [
    configs.linear_ss(),
    configs.lshape_full(num_phonons=100_000),
    configs.film_transient(num_phonons=100_000, num_measurements=2000, sim_time=2.0),
    configs.linear_bar(num_phonons=30_000, spec=0.4, sim_type=1, step_interval=20, t_eq=0.0).to_dict(),
], ids=["linear_ss", "lshape", "film", "bar_periodic_full"])
def test_host_budget_equals_oracle(cfg):
    """One host model prepared for several seeds in turn (its material tables built once and kept,
    the temperature bounds being the same) gives each seed's budget as a fresh oracle model does."""
    host = psim_amd.Model(cfg)
    for seed in [0, 1, 12345, 2**40 + 7]:
        eff, total, counts = host.prepare(seed)
        ref = oracle.Model(cfg)
        ref.prepare(seed)
        assert eff == ref.eff_energy
        assert total == ref.total_phonons
        np.testing.assert_array_equal(counts, ref.sources()["count"])


def test_sizes_follow_simulation_type():
    ss = psim_amd.Model(configs.linear_ss())
    assert (ss.num_sensors, ss.num_record) == (10, 200)
    per = psim_amd.Model(configs.linear_bar(sim_type=1, step_interval=10).to_dict())
    assert per.num_record == 2000


def _host_error(cfg):
    with pytest.raises(psim_amd.PsimError) as e:
        psim_amd.Model(cfg)
    return str(e.value).replace("\n", " ").replace('"', "'")


with open(os.path.join(GOLDEN, "ref_errors.json")) as _f:
    REF_ERRORS = json.load(_f)


@pytest.mark.parametrize("name", sorted(error_cases.cases()))
def test_rejects_what_the_reference_rejects(name):
    """Same message as the reference build gives for the same broken model (ref_errors.json)."""
    cfg = error_cases.cases()[name]
    assert _host_error(cfg) == REF_ERRORS[name]
    with pytest.raises(oracle.OracleError):
        oracle.Model(cfg)


def test_steady_state_ignores_step_interval():
    # InputManager reads step_interval only for periodic / transient runs (inputManager.cpp:77-78).
    cfg = configs.linear_ss()
    cfg["settings"]["step_interval"] = 5
    assert psim_amd.Model(cfg).num_record == 200
    oracle.Model(cfg)


def test_rejects_malformed_json():
    with pytest.raises(psim_amd.PsimError):
        psim_amd.Model('{"settings": {"num_measurements": 10,')
    with pytest.raises(psim_amd.PsimError):
        psim_amd.Model('{"settings": {}}')


def test_reads_reference_example_input():
    """psim_python/json/linear_demo.json, the reference's own example input (kept as a fixture)."""
    with open(os.path.join(GOLDEN, "linear_demo.json")) as f:
        text = f.read()
    host = psim_amd.Model(text)
    data = json.loads(text)
    assert host.num_sensors == len(data["sensors"])
    assert host.num_record == int(data["settings"]["num_measurements"] * 0.1)
    eff, total, counts = host.prepare(0)
    ref = oracle.Model(data)
    ref.prepare(0)
    assert eff == ref.eff_energy and total == ref.total_phonons
