"""The family-wise 3-sigma rule (tests/family.py) that the statistical parity tests apply:
it passes entries that Student-t noise produces and rejects what a real discrepancy does."""
import numpy as np

from family import family_ok

DOF = 30  # 16 runs on each side


def test_noise_passes():
    rng = np.random.default_rng(0)
    passed = sum(family_ok(np.abs(rng.standard_t(DOF, size=30)), DOF)[0] for _ in range(2000))
    assert passed >= 0.995 * 2000  # two bounds at 99.9 % each


def test_one_entry_far_off_fails():
    z = np.full(30, 1.0)
    z[7] = 6.0  # one sensor 6 standard errors off
    assert not family_ok(z, DOF)[0]


def test_several_entries_beyond_three_fail():
    z = np.full(30, 1.0)
    z[:4] = 3.5  # a shift that moves a handful of sensors past 3
    ok, (n, over, allowed, _, _) = family_ok(z, DOF)
    assert not ok and over > allowed


def test_a_single_chance_exceedance_passes():
    z = np.full(30, 1.5)
    z[5] = 3.23  # the size of chance fluctuation seen on 16 seeds
    assert family_ok(z, DOF)[0]
