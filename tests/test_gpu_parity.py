"""GPU engine against the CPU oracle, trajectory by trajectory.

Both sides draw from the same Philox schedule, so every phonon must end at the same
measurement step, for the same reason, at the same position (bit-exact), and the integer
tallies must agree exactly.  Flux tallies are FP64 sums whose order differs (LDS and
reduction order on the GPU, phonon order on the CPU); they must agree to 1e-9 relative of
the largest per-entry magnitude.
"""
import numpy as np
import pytest

import oracle
import psim_amd
from psim_amd import configs

pytestmark = pytest.mark.gpu


def _bar(spec, sim_type=0, step_interval=0, n=20000, steps=1000, t_eq=300.0, sensors=10):
    return configs.linear_bar(num_sensors=sensors, length=100.0 * sensors, num_phonons=n, num_measurements=steps,
                              sim_time=steps * 0.005, spec=spec, sim_type=sim_type, step_interval=step_interval,
                              t_eq=t_eq).to_dict()


CASES = {
    "bar_ss_diffuse": _bar(0.0),
    "bar_ss_specular": _bar(1.0),
    "bar_periodic_mixed": _bar(0.5, sim_type=1, step_interval=50),
    "bar_full_mode": _bar(0.2, t_eq=0.0),
    # 400 sensors: a window's tallies no longer fit LDS, so they go to HBM with atomics
    "bar_400_sensors_periodic": _bar(0.4, sim_type=1, step_interval=50, n=30000, sensors=400),
    "lshape_full": configs.lshape_full(num_phonons=20000, num_measurements=1000),
    "film_transient": configs.film_transient(num_phonons=20000, num_measurements=1000, sim_time=1.0,
                                             step_interval=50),
    # silicon | germanium: transitions between two materials
    "bar_si_ge": configs.si_ge_bar(),
    "bar_si_ge_full_mode": configs.si_ge_bar(t_eq=0.0),
    # phasor mode: fixed speed and branch, no scattering
    "bar_phasor": configs.si_ge_bar(phasor_sim=True),
    # specular needles: phonons stuck in a tip are relocated after 100 collisions in one drift
    "bar_needles_relocation": configs.needle_bar(),
    # a burst that starts at 0.3 ns: before it, its surfaces reflect like plain walls
    "film_late_burst": configs.film_transient(num_phonons=20000, num_measurements=1000, sim_time=1.0,
                                              step_interval=50, burst_start=0.3, burst_duration=0.2),
}


@pytest.fixture(scope="module")
def engine():
    e = psim_amd.Engine(0)
    yield e
    e.close()


def _flux_close(a, b):
    scale = max(1.0, float(np.max(np.abs(b))))
    return np.max(np.abs(a - b)) <= 1e-9 * scale


@pytest.mark.parametrize("tally", ["lds", "hbm"])
@pytest.mark.parametrize("name", sorted(CASES))
def test_trajectories_match_oracle(engine, name, tally):
    """Both tally paths: LDS window partials and FP64 atomics into HBM."""
    cfg = CASES[name]
    seed = 11
    print(f"case {name} tally {tally} seed {seed}", flush=True)
    model = psim_amd.Model(cfg)
    engine.set_fate(True)
    engine.set_tally_mode(tally)
    try:
        model.run(engine, seed)
    finally:
        engine.set_tally_mode("auto")
    res = model.results(0)
    eff, total = model.budget()
    g_step, g_cause, g_xy = engine.fetch_fate(total)
    ref = oracle.Model(cfg)
    o = ref.run(seed, nthreads=8, fate=True)
    assert total == ref.total_phonons and eff == ref.eff_energy
    assert res.counters == o.counters
    if name == "bar_needles_relocation":
        assert o.counters["relocations"] > 0  # the stuck-phonon guard was exercised
    np.testing.assert_array_equal(g_cause, o.fate_cause)
    np.testing.assert_array_equal(g_step, o.fate_step)
    np.testing.assert_array_equal(g_xy, o.fate_xy)
    np.testing.assert_array_equal(res.counts, o.counts)
    assert _flux_close(res.raw_flux, o.flux)
    meas, temps, _ = ref.interpret(o.counts, o.flux)
    np.testing.assert_array_equal(res.temps, temps)  # temperatures depend on the exact counts only
    np.testing.assert_allclose(res.measurements[:, [0, 1]], meas[:, [0, 1]], rtol=0, atol=0)
    np.testing.assert_allclose(res.measurements[:, 2:], meas[:, 2:], rtol=1e-9, atol=1e-9 * np.max(np.abs(meas[:, 2:])))


@pytest.mark.parametrize("sensors,length,launches", [(10, 1000.0, 8), (40, 4000.0, 17)])
def test_recording_windows_fill_the_lds_budget(engine, sensors, length, launches):
    """Default LDS recording windows are sized by the 48 KB budget: W + 1 rows of 20 B per sensor
    (the difference arrays of the window's tallies), at most 128 steps.  10 sensors give 128-step
    windows (1000 steps in 8 launches), 40 sensors 48 KB / 800 B - 1 = 60 steps (17 launches)."""
    cfg = configs.linear_bar(num_sensors=sensors, length=length, num_phonons=20000, num_measurements=1000,
                             sim_time=5.0, spec=0.5, sim_type=1, step_interval=50).to_dict()
    print(f"case {sensors} sensors", flush=True)
    model = psim_amd.Model(cfg)
    engine.set_fate(False)
    engine.set_window(0)
    engine.set_tally_mode("lds")
    try:
        model.run(engine, 3)
        assert engine.last_timing()["step_launches"] == launches
    finally:
        engine.set_tally_mode("auto")


@pytest.mark.parametrize("window", [1, 7, 333])
@pytest.mark.parametrize("name", ["bar_periodic_mixed", "bar_ss_diffuse"])
def test_launch_window_does_not_change_results(engine, name, window):
    """Splitting the run into launches of any length leaves every trajectory unchanged.  In
    steady state the phonons fly across the steps before the recorded ones, so short launches
    make them pause past a launch's end and resume there."""
    cfg = CASES[name]
    print(f"case {name} window {window}", flush=True)
    model = psim_amd.Model(cfg)
    engine.set_fate(True)
    engine.set_window(0)
    engine.set_tally_mode("lds")  # windows of recorded steps in LDS
    model.run(engine, 5)
    base = model.results(0)
    _, total = model.budget()
    base_fate = engine.fetch_fate(total)
    engine.set_window(window)
    try:
        model.run(engine, 5)
        res = model.results(0)
        fate = engine.fetch_fate(total)
    finally:
        engine.set_window(0)
    np.testing.assert_array_equal(res.counts, base.counts)
    assert res.counters == base.counters
    for a, b in zip(fate, base_fate):
        np.testing.assert_array_equal(a, b)
    assert _flux_close(res.raw_flux, base.raw_flux)


@pytest.mark.parametrize("block", [128, 512])
@pytest.mark.parametrize("tally", ["lds", "hbm"])
def test_block_size_does_not_change_results(engine, block, tally):
    """Threads per transport block (a launch shape the autotuner chooses) leave every trajectory
    unchanged: fates, counts and counters equal the 256-thread run's, on both tally paths."""
    cfg = CASES["bar_periodic_mixed"]
    print(f"case bar_periodic_mixed block {block} tally {tally}", flush=True)
    model = psim_amd.Model(cfg)
    engine.set_fate(True)
    engine.set_window(0)
    engine.set_block(256)
    engine.set_tally_mode(tally)
    model.run(engine, 5)
    base = model.results(0)
    _, total = model.budget()
    base_fate = engine.fetch_fate(total)
    engine.set_block(block)
    try:
        model.run(engine, 5)
        res = model.results(0)
        fate = engine.fetch_fate(total)
    finally:
        engine.set_block(256)
        engine.set_tally_mode("auto")
    np.testing.assert_array_equal(res.counts, base.counts)
    assert res.counters == base.counters
    for a, b in zip(fate, base_fate):
        np.testing.assert_array_equal(a, b)
    assert _flux_close(res.raw_flux, base.raw_flux)


def test_each_run_of_a_model_draws_from_its_own_seed(engine):
    """A model with num_runs = 2 (Model::runSimulation's loop, model.cpp:141-182): every run
    starts from reset controllers, and run r draws from seed + r * 0x9E3779B97F4A7C15, so it
    equals the oracle's single run with that seed."""
    cfg = configs.with_settings(_bar(0.3, n=10000), num_runs=2)
    print("case two runs seed 11", flush=True)
    model = psim_amd.Model(cfg)
    engine.set_fate(False)
    model.run(engine, 11)
    for r in range(2):
        res = model.results(r)
        o = oracle.Model(cfg).run((11 + r * 0x9E3779B97F4A7C15) % 2**64, nthreads=8)
        assert res.counters == o.counters
        np.testing.assert_array_equal(res.counts, o.counts)
        assert _flux_close(res.raw_flux, o.flux)
    assert not np.array_equal(model.results(0).counts, model.results(1).counts)


# the FP32 tallies' bounds against FP64: every tally, relative to the run's largest, and the
# per-sensor mean x-fluxes, relative to their largest (measured on an MI355X: 7.0e-7 and 2.0e-6 on
# the LDS path, whose partials are Kahan sums; 7.4e-6 and 8.9e-6 on the HBM path, where every
# recorded segment is an uncompensated FP32 atomic into the run-long difference arrays)
FP32_CASES = {
    "lshape_lds": (configs.lshape_full(num_phonons=200_000, num_measurements=1000), "lds", 2e-6, 1e-5),
    "bar_400_sensors_hbm": (configs.linear_bar(num_sensors=400, length=40000.0, num_phonons=100_000,
                                               num_measurements=1000, sim_time=5.0, spec=0.4, sim_type=1,
                                               step_interval=50).to_dict(), "hbm", 2e-5, 5e-5),
    "film_transient_lds": (configs.film_transient(num_phonons=200_000, num_measurements=2000, sim_time=2.0), "lds",
                           2e-6, 1e-5),
}


@pytest.mark.parametrize("name", sorted(FP32_CASES))
def test_fp32_kahan_tallies_against_fp64(engine, name):
    """The FP32 tally precision (FP32 atomics into the difference arrays, Kahan-compensated FP32
    sums along the steps and across blocks) runs the FP64 run's trajectories on the same seed:
    counts and counters are equal, every flux tally lies within the case's bound of the FP64 run
    (relative to its largest tally), and so do the per-sensor mean x-fluxes (relative to their
    largest): far inside the statistical spread of a run, about 1% per sensor mean."""
    cfg, tally, bound, mean_bound = FP32_CASES[name]
    print(f"case {name} fp32 kahan, tally {tally}", flush=True)
    model = psim_amd.Model(cfg)
    model.upload(engine, 0)
    engine.set_window(0)
    engine.set_tally_mode(tally)
    try:
        engine.run_resident(11)
        c64, f64, n64 = engine.fetch(model.num_sensors, model.num_record)
        engine.set_tally_precision("fp32_kahan")
        engine.run_resident(11)
        c32, f32, n32 = engine.fetch(model.num_sensors, model.num_record)
    finally:
        engine.set_tally_precision("fp64")
        engine.set_tally_mode("auto")
    np.testing.assert_array_equal(c32, c64)
    assert n32 == n64
    scale = float(np.max(np.abs(f64)))
    err = float(np.max(np.abs(f32 - f64)))
    m32, m64 = f32[:, :, 0].mean(axis=1), f64[:, :, 0].mean(axis=1)
    mean_err = float(np.max(np.abs(m32 - m64)) / np.max(np.abs(m64)))
    print(f"max |flux32 - flux64| / max |flux64| = {err / scale:.3e}; sensor mean x-flux {mean_err:.3e}", flush=True)
    assert err <= bound * scale
    assert mean_err <= mean_bound


def test_launch_shape_setters_reject_bad_values(engine):
    """Block sizes other than 128/256/512 threads and unknown tally precisions are refused with an
    error code; the launch shape in use is unchanged by a refused call."""
    L = psim_amd.lib()
    before = engine.launch_shape()
    assert L.psim_gpu_set_block(engine.handle, 100) != 0
    assert L.psim_gpu_set_block(engine.handle, 1024) != 0
    assert L.psim_gpu_set_tally_precision(engine.handle, 2) != 0
    assert engine.launch_shape() == before
    engine.set_block(512)
    assert engine.launch_shape()["block"] == 512
    engine.set_block(256)
