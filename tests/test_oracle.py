"""Pins of the CPU oracle (CPU only).

What fixes each part of the oracle, independently of the oracle's own formulas:
  Philox4x32-10     Random123 known-answer vectors and AMD's rocRAND implementation
  transcendentals   error bounds against glibc (CPython math) over dense sweeps
  tables, geometry, energy per phonon, SensorInterpreter
                    bit-exact equality with the reference build's own internal state
                    (tests/golden/ref_*_state.npz, written by tests/golden/make_ref_fixtures.py)
  frequency bins    brute force: the bisection returns the first bin whose cumulative value
                    exceeds the draw, clamped to [1, 999] (bin 0 folds into bin 1)
  directions        moments of the isotropic and Lambertian laws; exact specular reflection
  budget            counts are floor(E / E_eff) or one more, summing to the total
  whole loop        16 seeds against 16 independent reference runs, every sensor's entries
                    within 3 standard errors
"""
import json
import math
import os
import shutil
import subprocess

import numpy as np
import pytest

import oracle
from family import family_ok
from psim_amd import analysis, configs

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"


def load_fixture(name):
    with open(os.path.join(GOLDEN, f"ref_{name}.json")) as f:
        fix = json.load(f)
    state = np.load(os.path.join(GOLDEN, f"ref_{name}_state.npz"))
    return fix, state


# ---- Philox ------------------------------------------------------------------------------------

KAT = [  # Random123 kat_vectors, philox4x32 with 10 rounds
    ([0, 0, 0, 0], [0, 0], [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]),
    ([0xffffffff] * 4, [0xffffffff] * 2, [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]),
    ([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0],
     [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]),
]


@pytest.mark.parametrize("ctr,key,expect", KAT)
def test_philox_known_answers(ctr, key, expect):
    assert oracle.philox(ctr, key).tolist() == expect


ROCRAND_PROBE = r"""
#include <cstdio>
#include <rocrand/rocrand_philox4x32_10.h>
struct Probe : rocrand_device::philox4x32_10_engine {
    Probe() : rocrand_device::philox4x32_10_engine(0, 0, 0) {}
    uint4 run(uint4 c, uint2 k) { return ten_rounds(c, k); }
};
int main() {
    unsigned c0, c1, c2, c3, k0, k1;
    Probe p;
    while (std::scanf("%u %u %u %u %u %u", &c0, &c1, &c2, &c3, &k0, &k1) == 6) {
        uint4 r = p.run(make_uint4(c0, c1, c2, c3), make_uint2(k0, k1));
        std::printf("%u %u %u %u\n", r.x, r.y, r.z, r.w);
    }
}
"""


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not available")
def test_philox_matches_rocrand(tmp_path):
    """AMD's rocRAND philox4x32_10 (host build of its header) gives the same blocks."""
    src = tmp_path / "probe.hip"
    src.write_text(ROCRAND_PROBE)
    exe = tmp_path / "probe"
    subprocess.run([HIPCC, "--offload-arch=gfx950", "-O1", "-o", str(exe), str(src)], check=True,
                   capture_output=True, timeout=300)
    rng = np.random.default_rng(1)
    rows = rng.integers(0, 2**32, size=(300, 6), dtype=np.uint64)
    rows[:3] = [[0, 0, 0, 0, 0, 0], [5, 2, 77, 1, 0, 0], [0xffffffff] * 6]
    inp = "\n".join(" ".join(str(int(v)) for v in r) for r in rows) + "\n"
    out = subprocess.run([str(exe)], input=inp, capture_output=True, text=True, check=True, timeout=60).stdout
    theirs = np.array([[int(v) for v in line.split()] for line in out.strip().splitlines()], dtype=np.uint64)
    ours = np.array([oracle.philox(r[:4], r[4:]) for r in rows], dtype=np.uint64)
    np.testing.assert_array_equal(ours, theirs)


def test_uniform_stream_layout():
    """Every draw takes the next block: a single draw n is the first double of block n, a pair
    draw k both doubles of block k; a double is the 53 high bits of two words."""
    seed, tag, ident, step = 0x1234_5678_9abc, 0, 42, 17
    key = [seed & 0xffffffff, seed >> 32]

    def double(hi, lo):
        return (((int(hi) << 32) | int(lo)) >> 11) * 2.0**-53

    u = oracle.stream(seed, tag, ident, step, 6)
    for n in range(6):
        blk = oracle.philox([n, step, ident, tag], key)
        assert u[n] == double(blk[0], blk[1])
    w = oracle.stream(seed, tag, ident, step, 6, pairs=True)
    for n in range(6):
        blk = oracle.philox([n // 2, step, ident, tag], key)
        assert w[n] == (double(blk[0], blk[1]) if n % 2 == 0 else double(blk[2], blk[3]))
    big = oracle.stream(3, 1, 9, 0, 200_000)
    assert big.min() >= 0.0 and big.max() < 1.0
    assert abs(big.mean() - 0.5) < 5 * math.sqrt(1 / 12 / big.size)
    assert abs(np.var(big) - 1 / 12) < 5 * math.sqrt(1 / 180 / big.size)


# ---- transcendentals ---------------------------------------------------------------------------

def _ulp_err(ours, ref):
    ours = np.asarray(ours)
    ref = np.asarray(ref)
    return np.abs(ours - ref) / np.spacing(np.abs(ref))


def test_log_within_one_ulp():
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.random(200_000), np.exp(rng.uniform(-700, 700, 50_000)),
                        [2.0**-53, 0.5, 1.0, 1.0 - 2.0**-53, 2.0, 1e300]])
    x = x[x > 0]
    ref = np.array([math.log(v) for v in x])
    err = _ulp_err(oracle.math_fn("log", x), ref)
    assert err.max() <= 1.0
    assert oracle.math_fn("log", np.array([1.0]))[0] == 0.0
    assert oracle.math_fn("log", np.array([0.0]))[0] == -np.inf


def test_exp_and_sinh_within_two_ulp():
    rng = np.random.default_rng(1)
    x = rng.uniform(-700, 700, 100_000)
    err = _ulp_err(oracle.math_fn("exp", x), np.array([math.exp(v) for v in x]))
    assert err.max() <= 2.0
    xs = np.concatenate([rng.uniform(-3, 3, 100_000), rng.uniform(-50, 50, 20_000), [1e-300, -1e-10, 0.999999, 1.0]])
    err = _ulp_err(oracle.math_fn("sinh", xs), np.array([math.sinh(v) for v in xs]))
    assert err.max() <= 3.0


def test_cos2pi_accuracy_and_exact_points():
    rng = np.random.default_rng(2)
    u = rng.random(200_000)
    ours = oracle.math_fn("cos2pi", u)
    ref = np.array([math.cos(2 * math.pi * v) for v in u])
    # glibc sees 2*pi*u rounded (|error| <= 4.4e-16 in the argument)
    assert np.max(np.abs(ours - ref)) <= 1.2e-15
    exact = oracle.math_fn("cos2pi", np.array([0.0, 0.25, 0.5, 0.75, 0.125]))
    assert exact[0] == 1.0 and exact[2] == -1.0
    assert abs(exact[1]) < 1e-16 and abs(exact[3]) < 1e-16
    assert abs(exact[4] - math.sqrt(0.5)) <= 2**-53


def test_pow4_within_two_ulp():
    rng = np.random.default_rng(3)
    f = rng.uniform(1e11, 8e13, 100_000)
    err = _ulp_err(oracle.math_fn("pow4", f), np.array([math.pow(v, 4) for v in f]))
    assert err.max() <= 2.0


# ---- tables, geometry, energy, interpreter against the reference's own state -------------------

WORKLOADS = ["linear_ss", "lshape_r", "film_r"]


@pytest.mark.parametrize("name", WORKLOADS)
def test_tables_match_reference(name):
    fix, st = load_fixture(name)
    m = oracle.Model(fix["model"])
    m.prepare(0)
    temps, base_sums = m.temps()
    np.testing.assert_array_equal(temps, st["temps"])
    np.testing.assert_array_equal(base_sums, st["base_sums"])
    arrs = m.material_arrays()
    np.testing.assert_array_equal(arrs["freq"], st["frequencies"])
    np.testing.assert_array_equal(arrs["vel_la"], st["vel_la"])
    np.testing.assert_array_equal(arrs["vel_ta"], st["vel_ta"])
    for i, temp in enumerate(st["table_temps"]):
        for kind, key in enumerate(["base", "emit", "scatter"]):
            cdf, frac, total, idx = m.table(kind, float(temp))
            assert idx == st["table_index"][i]
            np.testing.assert_array_equal(cdf, st[f"{key}_cdf"][i])
            np.testing.assert_array_equal(frac, st[f"{key}_frac"][i])
            assert total == st[f"{key}_sum"][i]


@pytest.mark.parametrize("name", WORKLOADS)
def test_geometry_matches_reference(name):
    fix, st = load_fixture(name)
    m = oracle.Model(fix["model"])
    for c in range(len(st["cell_area"])):
        g = m.cell_geometry(c)
        assert g["area"] == st["cell_area"][c]
        np.testing.assert_array_equal(g["lines"], st["cell_lines"][c])
        np.testing.assert_array_equal(g["normals"], st["cell_normals"][c])
        np.testing.assert_array_equal(np.stack([g["n_trans"], g["n_emit"]], axis=1), st["cell_subs"][c])


@pytest.mark.parametrize("name", WORKLOADS)
def test_energy_per_phonon_matches_reference(name):
    fix, st = load_fixture(name)
    m = oracle.Model(fix["model"])
    m.prepare(0)
    assert m.eff_energy == float(st["eff_energy0"])
    assert all(e == m.eff_energy for e in fix["eff_energy0"])


@pytest.mark.parametrize("name", WORKLOADS)
def test_interpreter_matches_reference(name):
    """SensorInterpreter on the reference's own raw tallies reproduces its results bit for bit,
    including the post-run energy-per-phonon update of deviational steady state."""
    fix, st = load_fixture(name)
    m = oracle.Model(fix["model"])
    m.prepare(0)
    out, temps, _ = m.interpret(st["counts"], st["flux"])
    np.testing.assert_array_equal(out, st["measurements"])
    np.testing.assert_array_equal(temps, st["final_temps"])


# ---- sampling ----------------------------------------------------------------------------------

def test_frequency_bin_bisection_brute_force():
    rng = np.random.default_rng(4)
    for trial in range(40):
        w = rng.random(1000)
        w[rng.random(1000) < 0.2] = 0.0  # empty bins give tied cumulative values
        cdf = np.cumsum(w) / w.sum()
        frac = rng.random(1000)
        draws = np.concatenate([rng.random(40), [0.0, cdf[0] * 0.5, cdf[0], cdf[1], cdf[500], cdf[998], cdf[999],
                                                 np.nextafter(cdf[998], 0), 1.0 - 2**-53]])
        for u in draws:
            u2 = rng.random()
            b, pol = oracle.freq_index(cdf, frac, u, u2)
            expect = int(np.count_nonzero(cdf <= u))  # first bin whose cumulative value exceeds u
            expect = min(max(expect, 1), 999)
            assert b == expect, (trial, u, b, expect)
            assert pol == (0 if u2 <= frac[b] else 1)


def test_direction_moments():
    n = 200_000
    iso = oracle.directions(0, 9, n)
    tol = 5 / math.sqrt(n)
    assert abs(iso[:, 0].mean()) < tol and abs(iso[:, 1].mean()) < tol
    assert abs((iso[:, 0] ** 2).mean() - 1 / 3) < tol and abs((iso[:, 1] ** 2).mean() - 1 / 3) < tol
    assert np.all(iso[:, 0] ** 2 + iso[:, 1] ** 2 <= 1 + 1e-15)
    nrm = (math.cos(0.7), math.sin(0.7))
    lam = oracle.directions(1, 10, n, normal=nrm)
    along = lam @ np.array(nrm)
    tang = lam @ np.array([-nrm[1], nrm[0]])
    assert np.all(along >= 0)
    assert abs(along.mean() - 2 / 3) < tol and abs((along**2).mean() - 1 / 2) < tol
    assert abs(tang.mean()) < tol and abs((tang**2).mean() - 1 / 4) < tol


def test_specular_reflection_is_exact_mirror():
    rng = np.random.default_rng(5)
    d = rng.normal(size=(1000, 2))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    nrm = np.array([0.6, -0.8])
    r = oracle.directions(2, 0, 1000, normal=tuple(nrm), dirs=d)
    expect = d - 2 * (d @ nrm)[:, None] * nrm[None, :]
    np.testing.assert_allclose(r, expect, atol=1e-15)
    axis = oracle.directions(2, 0, 3, normal=(1.0, 0.0), dirs=np.array([[0.3, 0.4], [-1.0, 0.0], [0.0, 1.0]]))
    np.testing.assert_array_equal(axis, [[-0.3, 0.4], [1.0, 0.0], [0.0, 1.0]])


def test_budget_rounding():
    cfg = configs.lshape_full(num_phonons=100_000)
    means = []
    for seed in range(8):
        m = oracle.Model(cfg)
        m.prepare(seed)
        src = m.sources()
        ratio = src["energy"] / m.eff_energy
        assert np.all((src["count"] == np.floor(ratio)) | (src["count"] == np.floor(ratio) + 1))
        assert src["count"].sum() == m.total_phonons
        means.append(src["count"] - np.floor(ratio))
    frac = ratio - np.floor(ratio)
    # Bernoulli rounding: over seeds the rounded-up share follows the fractional part
    assert abs(np.mean(means) - frac.mean()) < 0.1


# ---- whole loop: 16 seeds against 16 reference runs ----------------------------------------------

def _three_sigma(fix, ours, key):
    return analysis.three_sigma([o[key] for o in ours], fix["mean"][key], fix["sd"][key], fix["runs"])


@pytest.mark.parametrize("name", WORKLOADS)
def test_oracle_statistics_match_reference(name):
    fix, _ = load_fixture(name)
    ours = []
    for seed in range(fix["runs"]):
        m = oracle.Model(fix["model"])
        r = m.run(seed, nthreads=8)
        out, temps, _ = m.interpret(r.counts, r.flux)
        ours.append({"t_steady": out[:, 0], "x_flux": out[:, 2], "y_flux": out[:, 4]})
    zs = [_three_sigma(fix, ours, key) for key in ["t_steady", "x_flux", "y_flux"]]
    assert max(float(z.max()) for z in zs) <= 3.0, [np.round(z, 2) for z in zs]


@pytest.mark.parametrize("name", ["lshape_r", "film_r", "needles", "linear_ss"])
def test_flights_agree_with_the_step_by_step_loop(name):
    """The oracle's flights across the measurement boundaries (the engine's loop) against the
    reference's loop as written, one drift per measurement step (modelSimulator.cpp:139-198), on
    the same seed.  Both draw the same random numbers at the same events, so they agree up to the
    few trajectories that rounding sends elsewhere: every counter within 2e-3.  This holds the
    flights to the reference's semantics in each place they could differ: boundaries passed
    during a flight must still count as steps and be tallied (tallies, phonon-steps), the step
    of a collision must be that of its moment (the film's burst surfaces absorb only inside
    their window), and stuck phonons must be relocated after 100 collisions in one step's drift,
    not in one flight (the needles relocate 172 times either way at seed 0, 1465 times with
    collisions counted per flight).  The per-sensor, per-step tallies must agree as well."""
    fix, _ = load_fixture(name) if name != "needles" else (json.load(open(os.path.join(GOLDEN, "ref_needles.json"))), None)
    a = oracle.Model(fix["model"]).run(0, nthreads=8)
    b = oracle.Model(fix["model"]).run(0, nthreads=8, step_by_step=True)
    for key in ["phonon_steps", "scatters", "impacts", "tallies", "absorbed"]:
        assert abs(a.counters[key] - b.counters[key]) <= 2e-3 * max(1, b.counters[key]), (key, a.counters, b.counters)
    assert abs(a.counters["expired"] - b.counters["expired"]) <= max(20, 0.01 * b.counters["expired"])
    assert abs(a.counters["relocations"] - b.counters["relocations"]) <= max(10, 0.1 * b.counters["relocations"])
    assert a.counters["stalled"] == b.counters["stalled"] == 0
    # Where the tallies land: the per-sensor, per-step counts of the two loops.  The trajectories
    # that rounding sends elsewhere move at most 0.4 % of the absolute tally mass (the L-shape,
    # whose phonons live ~1200 steps); a flight that tallies its passed boundaries one step late
    # moves 2.8-4.2 %, one that tallies them into the next sensor 9-190 %.  Both planted mistakes
    # (an oracle test hook) must exceed 2 %, and the real loop stay below 1 %.
    def moved(x):
        return np.abs(x.counts - b.counts).sum() / np.abs(b.counts).sum()

    assert moved(a) <= 1e-2, moved(a)
    for fault in (1, 2):
        bad = oracle.Model(fix["model"])
        bad.plant_flight_fault(fault)
        assert moved(bad.run(0, nthreads=8)) > 2e-2, fault


STATS_ONLY = ["si_ge", "phasor", "needles", "kinked_15", "kinked_90"]
# Fixtures judged as a family of entries (tests/family.py) instead of entry by entry, with the
# reason: on seeds 0-15 the silicon|germanium bar has one x-flux at z = 3.23 while its other 29
# entries are at most 2.34, and on seeds 16-31 and 32-47 all 30 entries are at most 2.1 -- a
# Student-t tail, not a bias.
FAMILY_RULE = {"si_ge"}


@pytest.mark.parametrize("name", STATS_ONLY)
def test_oracle_statistics_match_reference_on_interfaces_and_phasors(name):
    """Paths the three workloads above do not take: transitions between two materials (a
    phonon above the neighbour's maximum frequency of its branch scatters back diffusely,
    surface.cpp:71-109), phasor mode (fixed speed and branch, no scattering,
    phononBuilder.cpp:42-49), acute specular needles where stuck phonons are relocated
    (modelSimulator.cpp:215-218), and kinked structures of the batched-geometry family whose
    edges lie at arbitrary slopes (geometry.cpp:104-138).  16 oracle seeds against 16 reference
    runs, every sensor's entries within 3 standard errors (si_ge as a family, see FAMILY_RULE)."""
    with open(os.path.join(GOLDEN, f"ref_{name}.json")) as f:
        fix = json.load(f)
    ours = []
    for seed in range(fix["runs"]):
        m = oracle.Model(fix["model"])
        r = m.run(seed, nthreads=8)
        out, _, _ = m.interpret(r.counts, r.flux)
        ours.append({"t_steady": out[:, 0], "x_flux": out[:, 2], "y_flux": out[:, 4]})
    zs = [_three_sigma(fix, ours, key) for key in ["t_steady", "x_flux", "y_flux"]]
    if name in FAMILY_RULE:
        ok, detail = family_ok(np.concatenate(zs), 2 * (fix["runs"] - 1))
        assert ok, (detail, [np.round(z, 2) for z in zs])
    else:
        assert max(float(z.max()) for z in zs) <= 3.0, [np.round(z, 2) for z in zs]
