"""Derived quantities (psim_amd.analysis) on inputs whose answers are known by construction."""
import json
import os
import sys

import numpy as np

from psim_amd import analysis

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
from make_ref_fixtures import block_series  # noqa: E402  (the fixtures' own block definition)


def test_k_eff_linear_profile():
    # Fourier: q = k dT / L  =>  k = q L / dT, whatever the per-sensor noise around q
    q = np.array([9e8, 1.0e9, 1.1e9, 0.9e9, 1.0e9, 7e8])  # interior mean 1e9
    k = analysis.k_eff(q, 1000.0, 310.0, 290.0)
    assert np.isclose(k, 1.0e9 * 1e-6 / 20.0)
    assert np.isclose(analysis.k_eff(q, 1000.0, 310.0, 290.0, interior=False), q.mean() * 1e-6 / 20.0)


def test_block_means_match_fixture_definition():
    rng = np.random.default_rng(3)
    x = rng.normal(300.0, 1.0, 1037)
    np.testing.assert_allclose(analysis.block_means(x, 100), block_series(list(x), 100), rtol=0, atol=1e-12)
    assert analysis.block_means(x[:99], 100).shape == (0,)


def test_time_to_steady_state_known_answers():
    steps = np.arange(3000)
    # a sensor that relaxes linearly for 1000 steps and is flat afterwards
    t = np.where(steps < 1000, 300.0 + 10.0 * (1000 - steps) / 1000.0, 300.0)
    temps = np.vstack([t, t - 5.0])
    # window 9 still moves (0.5 K), the changes 10->11, 11->12, 12->13 are zero: steady at the end
    # of window 13
    assert analysis.time_to_steady_state(temps) == 1400
    # a 40 K one-step spike in window 12 moves its mean by 0.4 K, so the changes 11->12 and
    # 12->13 are loud and the three quiet ones are 13->14 .. 15->16
    spiky = temps.copy()
    spiky[:, 1250] += 40.0
    assert analysis.time_to_steady_state(spiky) == 1700
    # never settles
    assert analysis.time_to_steady_state(np.vstack([300.0 + 0.01 * steps])) is None
    # window-mean noise of 0.01 K is below the 0.05 K tolerance from the start
    rng = np.random.default_rng(0)
    assert analysis.time_to_steady_state(300.0 + rng.normal(0.0, 0.1, (4, 3000))) == 400


def test_three_sigma_scaling():
    rng = np.random.default_rng(1)
    a = rng.normal(1.0, 0.1, (16, 5))
    z = analysis.three_sigma(a, a.mean(axis=0) + 3 * np.sqrt(2 * a.std(axis=0, ddof=1) ** 2 / 16),
                             a.std(axis=0, ddof=1), 16)
    np.testing.assert_allclose(z, 3.0)


def test_three_sigma_without_spread():
    """Entries where neither side varies: equal means score 0, a constant offset fails, and an
    offset within atol passes."""
    a = np.full((16, 3), 300.0)
    a[:, 2] += np.linspace(-1.0, 1.0, 16)  # the third entry varies
    z = analysis.three_sigma(a, [300.0, 300.5, 300.0], [0.0, 0.0, 0.0], 16)
    assert z[0] == 0.0 and np.isinf(z[1]) and np.isfinite(z[2])
    z = analysis.three_sigma(a, [300.0, 300.005, 300.0], [0.0, 0.0, 0.0], 16, atol=0.01)
    assert z[1] == 0.0


def test_k_eff_fixture_is_consistent():
    fix = json.load(open(os.path.join(HERE, "golden", "ref_linear_k.json")))
    per = np.array(fix["per_run"]["x_flux"])
    assert per.shape == (fix["runs"], 10)
    ks = [analysis.k_eff(r, 1000.0, 310.0, 290.0) for r in per]
    # the mean of 48 runs is known to better than 0.5 %, enough to judge a 1 % agreement
    assert np.std(ks, ddof=1) / np.sqrt(len(ks)) / np.mean(ks) < 0.005
