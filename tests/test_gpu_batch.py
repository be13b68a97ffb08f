"""Batched geometries (psim_batch_run) against the oracle, the single runs and the reference.

Model m of a batch draws the random numbers of its single run with stream m, and the batch
changes nothing else:
  - every model's phonons end as the oracle's model at stream m (same step, cause and position,
    bit for bit), its integer tallies are the oracle's and its flux sums agree to 1e-9 relative
    (their order differs), and the batch counters are the sums of the oracle's;
  - every model's raw tallies equal its single run's, and its temperatures bit for bit;
  - kinked structures (arms at 15 and 90 degrees) agree with the reference's own runs within
    3 standard errors per sensor.
"""
import json
import os

import numpy as np
import pytest

import oracle
import psim_amd
from psim_amd import analysis, configs

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module")
def engine():
    e = psim_amd.Engine(0)
    yield e
    e.close()


def _small_batch(n, phonons=60_000):
    ms = configs.batch_geometries(seed=7, count=n)
    for i, m in enumerate(ms):
        m["settings"]["num_phonons"] = min(m["settings"]["num_phonons"], phonons + 1000 * i)
    return ms


def _flux_close(a, b):
    scale = max(1.0, float(np.max(np.abs(b))))
    return np.max(np.abs(a - b)) <= 1e-9 * scale


def _mixed_batch():
    """Ten structures of the generator (four of them kinked) and the two fixed kinked structures at
    15 and 90 degrees, at most 3e4 phonons each."""
    cfgs = configs.batch_geometries(seed=7, count=10)
    cfgs += [configs.kinked_structure(15.0, num_phonons=20_000), configs.kinked_structure(90.0, num_phonons=20_000)]
    for i, m in enumerate(cfgs):
        m["settings"]["num_phonons"] = min(m["settings"]["num_phonons"], 30_000 + 500 * i)
    return cfgs


def test_mixed_batch_is_generator_family():
    cfgs = _mixed_batch()
    kinked = [c for c in cfgs if len(c["cells"]) % 2 == 1]  # the wedge is a single triangle
    assert len(cfgs) >= 8 and len(kinked) >= 3


@pytest.mark.parametrize("tally", ["auto", "lds"])
def test_batch_matches_oracle_at_each_stream(engine, tally):
    """Every model of a batch of twelve (six kinked) against the oracle run at its stream index:
    fates bit for bit, integer tallies exact, counters summed over the models."""
    cfgs = _mixed_batch()
    seed = 21
    print(f"mixed batch of {len(cfgs)}, tally {tally}, seed {seed}", flush=True)
    models = [psim_amd.Model(c) for c in cfgs]
    engine.set_fate(True)
    engine.set_tally_mode(tally)
    try:
        batch_counters = psim_amd.run_batch(engine, models, seed=seed)
    finally:
        engine.set_tally_mode("auto")
    totals = [m.budget()[1] for m in models]
    step, cause, xy = engine.fetch_fate(int(sum(totals)))
    engine.set_fate(False)
    summed = {k: 0 for k in batch_counters}
    first = 0
    for i, (cfg, bm) in enumerate(zip(cfgs, models)):
        ref = oracle.Model(cfg)
        ref.set_stream(i)
        o = ref.run(seed, nthreads=8, fate=True)
        eff, total = bm.budget()
        assert total == ref.total_phonons and eff == ref.eff_energy, i
        sl = slice(first, first + total)
        first += total
        np.testing.assert_array_equal(cause[sl], o.fate_cause, err_msg=f"model {i}")
        np.testing.assert_array_equal(step[sl], o.fate_step, err_msg=f"model {i}")
        np.testing.assert_array_equal(xy[sl], o.fate_xy, err_msg=f"model {i}")
        res = bm.results(0)
        np.testing.assert_array_equal(res.counts, o.counts)
        assert _flux_close(res.raw_flux, o.flux), i
        meas, temps, _ = ref.interpret(o.counts, o.flux)
        np.testing.assert_array_equal(res.temps, temps)
        for k, v in o.counters.items():
            summed[k] += v
    assert batch_counters == summed


def test_batch_equals_single_runs(engine):
    cfgs = _small_batch(10)
    print(f"batch of {len(cfgs)} geometries, sensors {[len(c['sensors']) for c in cfgs]}", flush=True)
    models = [psim_amd.Model(c) for c in cfgs]
    batch_counters = psim_amd.run_batch(engine, models, seed=3)
    total = {k: 0 for k in batch_counters}
    for i, (cfg, bm) in enumerate(zip(cfgs, models)):
        b = bm.results(0)
        single = psim_amd.Model(cfg)
        single.set_stream(i)
        single.run(engine, 3)
        s = single.results(0)
        for k, v in s.counters.items():
            total[k] += v
        np.testing.assert_array_equal(b.counts, s.counts)
        scale = max(1.0, float(np.max(np.abs(s.raw_flux))))
        assert np.max(np.abs(b.raw_flux - s.raw_flux)) <= 1e-9 * scale
        np.testing.assert_array_equal(b.temps, s.temps)
        np.testing.assert_array_equal(b.measurements[:, :2], s.measurements[:, :2])
    assert batch_counters == total


@pytest.mark.parametrize("block", [128, 512])
def test_batch_block_size_does_not_change_results(engine, block):
    """The batched (multi-stream) transport variant at 128 and 512 threads per block gives the
    256-thread run's counters and every model's integer tallies exactly, its flux sums to 1e-9."""
    cfgs = _small_batch(6)
    print(f"batch of {len(cfgs)} geometries, block {block}", flush=True)
    base_models = [psim_amd.Model(c) for c in cfgs]
    engine.set_block(256)
    base = psim_amd.run_batch(engine, base_models, seed=4)
    models = [psim_amd.Model(c) for c in cfgs]
    engine.set_block(block)
    try:
        got = psim_amd.run_batch(engine, models, seed=4)
    finally:
        engine.set_block(256)
    assert got == base
    for bm, m in zip(base_models, models):
        a, b = bm.results(0), m.results(0)
        np.testing.assert_array_equal(a.counts, b.counts)
        assert _flux_close(b.raw_flux, a.raw_flux)


def test_full_batch_of_256_structures(engine):
    """The benchmark's batch at its full size (256 structures of configs.batch_geometries(seed 0),
    1e4-1e6 phonons each, 2000 steps of 5 ps): the batch counters are the sums of the 256 single
    runs' counters, nothing stalls, and every model's tallies equal its single run's."""
    cfgs = configs.batch_geometries(seed=0, count=256)
    print(f"batch of {len(cfgs)} structures, {sum(c['settings']['num_phonons'] for c in cfgs)} phonons", flush=True)
    models = [psim_amd.Model(c) for c in cfgs]
    batch_counters = psim_amd.run_batch(engine, models, seed=5)
    print(f"batch done: {batch_counters}", flush=True)
    assert batch_counters["stalled"] == 0
    summed = {k: 0 for k in batch_counters}
    single = None
    for i, (cfg, bm) in enumerate(zip(cfgs, models)):
        single = psim_amd.Model(cfg)
        single.set_stream(i)
        single.run(engine, 5)
        s = single.results(0)
        b = bm.results(0)
        for k, v in s.counters.items():
            summed[k] += v
        np.testing.assert_array_equal(b.counts, s.counts, err_msg=f"model {i}")
        assert _flux_close(b.raw_flux, s.raw_flux), i
        np.testing.assert_array_equal(b.temps, s.temps)
    assert batch_counters == summed


def _ref_fixture(name):
    with open(os.path.join(GOLDEN, f"ref_{name}.json")) as f:
        return json.load(f)


@pytest.mark.parametrize("name", ["kinked_15", "kinked_90"])
def test_kinked_structures_match_reference(engine, name):
    """16 GPU seeds against the reference's 16 runs of the same kinked structure: every sensor's
    temperature, x-flux and y-flux within 3 standard errors (the per-entry rule)."""
    fix = _ref_fixture(name)
    cfg = fix["model"]
    print(f"case {name}", flush=True)
    model = psim_amd.Model(cfg)
    engine.set_fate(False)
    runs = []
    for seed in range(16):
        model.run(engine, seed)
        runs.append(model.results(0).measurements)
    runs = np.array(runs)
    for col, key in [(0, "t_steady"), (2, "x_flux"), (4, "y_flux")]:
        z = analysis.three_sigma(runs[:, :, col], np.array(fix["mean"][key]), np.array(fix["sd"][key]), fix["runs"])
        assert np.max(z) <= 3.0, (key, np.round(z, 2))


def test_batch_rejects_mixed_settings(engine):
    a, b = _small_batch(2)
    b["settings"]["num_measurements"] = 1000
    with pytest.raises(psim_amd.PsimError, match="must share"):
        psim_amd.run_batch(engine, [psim_amd.Model(a), psim_amd.Model(b)], seed=0)
