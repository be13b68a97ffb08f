"""Batched geometries (psim_batch_run) against single runs.

Model m of a batch draws the random numbers of its single run with stream m, and the batch
changes nothing else: every model's raw tallies must equal its single run's exactly (flux sums to
1e-9 relative, their order differs), its temperatures bit for bit, and the batch counters must be
the sums of the single runs' counters.
"""
import numpy as np
import pytest

import psim_amd
from psim_amd import configs

pytestmark = pytest.mark.gpu


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


def test_batch_rejects_mixed_settings(engine):
    a, b = _small_batch(2)
    b["settings"]["num_measurements"] = 1000
    with pytest.raises(psim_amd.PsimError, match="must share"):
        psim_amd.run_batch(engine, [psim_amd.Model(a), psim_amd.Model(b)], seed=0)
