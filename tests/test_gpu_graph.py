"""The run captured as a HIP graph (psim_gpu_run_graph, psim_gpu_run_async) and the autotuner
against eager runs.

Every replay, with eager runs and synchronisations between the replays, must give the eager
result of the same seed: integer tallies and counters exactly, flux sums to 1e-9 relative (their
order differs).  Re-capture after a change of launch shape, and the autotuner's choice, must
leave results unchanged too.
"""
import numpy as np
import pytest

import psim_amd
from psim_amd import configs

pytestmark = pytest.mark.gpu

CASES = {
    "bar_periodic_lds_tallies": configs.linear_bar(num_phonons=20000, num_measurements=1000, sim_time=5.0, spec=0.5,
                                                   sim_type=1, step_interval=50).to_dict(),
    "bar_400_sensors_hbm_tallies": configs.linear_bar(num_sensors=400, length=40000.0, num_phonons=30000,
                                                      num_measurements=1000, sim_time=5.0, spec=0.4, sim_type=1,
                                                      step_interval=50).to_dict(),
    "lshape_steady": configs.lshape_full(num_phonons=20000, num_measurements=1000),
}
# the tally path each case is run on (the automatic rule would pick HBM atomics for all three)
TALLY = {"bar_periodic_lds_tallies": "lds", "bar_400_sensors_hbm_tallies": "hbm", "lshape_steady": "lds"}


@pytest.fixture(scope="module")
def engine():
    e = psim_amd.Engine(0)
    yield e
    e.close()


def _eager(engine, model, seed):
    engine.run_resident(seed)
    return engine.fetch(model.num_sensors, model.num_record)


def _same(a, b):
    (ca, fa, na), (cb, fb, nb) = a, b
    np.testing.assert_array_equal(ca, cb)
    assert na == nb
    scale = max(1.0, float(np.max(np.abs(fb))))
    assert np.max(np.abs(fa - fb)) <= 1e-9 * scale


@pytest.mark.parametrize("name", sorted(CASES))
def test_graph_replays_match_eager(engine, name):
    print(f"case {name} tally {TALLY[name]}", flush=True)
    engine.set_window(0)
    engine.set_grid(8)
    engine.set_tally_mode(TALLY[name])
    model = psim_amd.Model(CASES[name])
    model.upload(engine, 0)
    try:
        for seed in (5, 6, 5, 7, 6):
            g = engine.run_graph(model.num_sensors, model.num_record, seed)
            _same(g, _eager(engine, model, seed))
        engine.set_window(7)  # a new launch sequence: captured again
        _same(engine.run_graph(model.num_sensors, model.num_record, 9), _eager(engine, model, 9))
        engine.set_tally_mode("hbm" if TALLY[name] == "lds" else "lds")  # the other path: captured again
        _same(engine.run_graph(model.num_sensors, model.num_record, 9), _eager(engine, model, 9))
    finally:
        engine.set_window(0)
        engine.set_tally_mode("auto")


def test_autotune_keeps_results(engine):
    model = psim_amd.Model(CASES["lshape_steady"])
    model.upload(engine, 0)
    engine.set_window(0)
    engine.set_grid(8)
    engine.set_tally_mode("lds")  # the window choice matters on the LDS path
    reference = _eager(engine, model, 3)
    choice = engine.autotune(seed=1)
    print("autotune choice", choice, flush=True)
    assert choice["window"] in (0, 1024, 256) and choice["blocks_per_cu"] in (8, 3) and choice["run_ms"] > 0
    assert choice["block"] in (128, 256, 512)
    assert engine.launch_shape() == {k: choice[k] for k in ("window", "blocks_per_cu", "block")}
    _same(_eager(engine, model, 3), reference)
    _same(engine.run_graph(model.num_sensors, model.num_record, 3), reference)
    engine.set_window(0)
    engine.set_grid(8)
    engine.set_block(256)
    engine.set_tally_mode("auto")


def _engine_hip_runtime():
    """The HIP runtime libpsim_amd.so is bound to.  It depends on the soname libamdhip64.so.7,
    which torch's bundled runtime carries too: whichever copy was loaded first serves both, and
    opening the soname returns that copy."""
    import ctypes

    psim_amd.lib()
    return ctypes.CDLL("libamdhip64.so.7")


def test_async_runs_on_a_caller_stream_match_eager(engine):
    """psim_gpu_run_async enqueues the graph on the caller's stream and returns; after that
    stream is synchronised the results equal the eager run of the same seed."""
    import ctypes

    print("case async on a caller stream and on the null stream", flush=True)
    hip = _engine_hip_runtime()
    engine.set_window(0)
    engine.set_grid(8)
    engine.set_tally_mode("lds")
    model = psim_amd.Model(CASES["bar_periodic_lds_tallies"])
    model.upload(engine, 0)
    stream = ctypes.c_void_p()
    assert hip.hipStreamCreate(ctypes.byref(stream)) == 0  # default priority
    try:
        for seed in (3, 4, 3):
            engine.run_async(seed, stream.value)
            assert hip.hipStreamSynchronize(stream) == 0
            _same(engine.fetch(model.num_sensors, model.num_record), _eager(engine, model, seed))
    finally:
        assert hip.hipStreamDestroy(stream) == 0
    engine.run_async(8, 0)  # the null stream
    assert hip.hipDeviceSynchronize() == 0
    _same(engine.fetch(model.num_sensors, model.num_record), _eager(engine, model, 8))
    engine.set_tally_mode("auto")
