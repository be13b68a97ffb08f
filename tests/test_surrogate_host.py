"""Host side of the surrogate (psim_amd/surrogate.py, psim_amd/train_surrogate.py), on the CPU."""
import random

import numpy as np
import torch

from psim_amd import configs, surrogate, train_surrogate


def _cell_centroids(model: dict) -> np.ndarray:
    """Area-weighted centroid of every sensor's triangles, from the model's own cells."""
    acc = np.zeros((len(model["sensors"]), 3))
    for c in model["cells"]:
        t = c["triangle"]
        pts = np.array([[t[k]["x"], t[k]["y"]] for k in ("p1", "p2", "p3")])
        e1, e2 = pts[1] - pts[0], pts[2] - pts[0]
        area = 0.5 * abs(e1[0] * e2[1] - e1[1] * e2[0])
        acc[c["sensorID"], :2] += area * pts.mean(axis=0)
        acc[c["sensorID"], 2] += area
    return acc[:, :2] / acc[:, 2:]


def test_sensor_centroids_match_the_model_cells():
    for i in range(60):
        p = configs.batch_geometry_params(random.Random(i))
        model = configs.geometry_from_params(p, 100, 1.0)
        want = _cell_centroids(model)
        got = np.array([r[:2] for r in surrogate.sensor_geometry(p)])
        assert got.shape == want.shape
        np.testing.assert_allclose(got, want, atol=1e-9 * p["length"])


def test_features_are_finite_and_ordered_along_the_structure():
    for i in range(40):
        p = configs.batch_geometry_params(random.Random(100 + i))
        x = surrogate.features(p)
        assert x.shape == (p["n_sensors"], surrogate.N_FEATURES) and x.dtype == np.float32
        assert np.isfinite(x).all()
        s = x[:, 8]  # arc-length position
        assert np.all(np.diff(s) > 0) and 0 < s[0] and s[-1] < 1


def test_shape_round_trip_and_linear_guess():
    p = configs.batch_geometry_params(random.Random(7))
    temps = np.linspace(p["t_hot"], p["t_cold"], p["n_sensors"])
    u = surrogate.shape_of(p, temps)
    np.testing.assert_allclose(u[[0, -1]], [1.0, -1.0], atol=1e-12)
    np.testing.assert_allclose(surrogate.temps_of(p, u), temps, rtol=0, atol=1e-9)
    lin = np.array(surrogate.linear_inits(p))
    lo, hi = sorted((p["t_hot"], p["t_cold"]))
    assert np.all((lin > lo) & (lin < hi))
    assert np.all(np.diff(lin) * np.sign(p["t_cold"] - p["t_hot"]) > 0)


def test_warm_start_temperatures_reach_the_model():
    p = configs.batch_geometry_params(random.Random(3))
    inits = [300.0 + 0.1 * i for i in range(p["n_sensors"])]
    model = configs.geometry_from_params(p, 100, 1.0, t_init=inits)
    assert [s["t_init"] for s in model["sensors"]] == inits


def test_bf16_rounding_matches_torch():
    rng = np.random.default_rng(0)
    a = np.concatenate([rng.normal(size=10000), [0.0, -0.0, 1.0, 3.0e38, -1e-40, 1.00390625, 1.01171875]])
    a = a.astype(np.float32)
    want = torch.from_numpy(a).to(torch.bfloat16).view(torch.int16).numpy().view(np.uint16)
    np.testing.assert_array_equal(surrogate.bf16_bits(a), want)


def test_training_and_held_out_structures_are_disjoint():
    train = train_surrogate.structures(train_surrogate.TRAIN_SEED, 300)
    held = train_surrogate.structures(train_surrogate.EVAL_SEED, 300)
    key = lambda p: (round(p["length"], 9), round(p["width"], 9), p["n_sensors"])  # noqa: E731
    assert not {key(p) for p in train} & {key(p) for p in held}
    assert all(abs(p["t_hot"] - p["t_cold"]) >= train_surrogate.MIN_WALL_DIFF for p in train + held)


def _fake_relaxation(p, inits, steps, tau_steps):
    """Per-step temperatures relaxing from `inits` (t_eq when None) to the linear profile with
    time constant tau_steps, plus tiny deterministic noise."""
    t_eq = 0.5 * (p["t_hot"] + p["t_cold"])
    target = np.array(surrogate.linear_inits(p))
    start = np.full(p["n_sensors"], t_eq) if inits is None else np.asarray(inits, float)
    t = np.arange(steps)
    noise = 1e-6 * np.sin(np.arange(p["n_sensors"])[:, None] * 0.7 + t[None, :] * 1.3)
    return target[:, None] + (start - target)[:, None] * np.exp(-t / tau_steps)[None, :] + noise


def test_evaluation_bookkeeping(monkeypatch):
    """train_surrogate.evaluate on synthetic relaxations: a start at the settled field reaches
    steady state at the earliest report (4 windows), a start at t_eq later; totals, ratios and
    initial errors follow."""
    params = train_surrogate.structures(train_surrogate.EVAL_SEED, 4)

    def fake_runs(engine, ps, inits, phonons, steps, sim_ns, seed, reduce, batch=256, threads=16, progress=None):
        return [reduce(p, _fake_relaxation(p, None if inits is None else inits[i], steps, 600.0))
                for i, p in enumerate(ps)]

    class PerfectNet:
        """Predicts the linear profile, which the synthetic runs settle to."""

        def predict(self, x):
            return (1.0 - 2.0 * x[:, 8]).astype(np.float32)  # u of the linear profile at position s

    monkeypatch.setattr(surrogate, "run_relaxations", fake_runs)
    res = train_surrogate.evaluate(None, PerfectNet(), params, 1000, 5000, 25.0, 0, linear_count=2,
                                   progress=lambda m: None, cost=False)
    s = res["summary"]
    row = s["criteria"][str(train_surrogate.HEADLINE_TOL)]
    assert row["warm"]["reached"] == 4 and row["warm"]["total_ns"] == 4 * 2.0  # 400 steps of 5 ps each
    assert row["cold"]["total_ns"] > row["warm"]["total_ns"]
    assert row["cold_over_warm_total"] == row["cold"]["total_ns"] / row["warm"]["total_ns"]
    assert s["initial_error_rmse_K"]["warm"] < 1e-3 < s["initial_error_rmse_K"]["cold"]
    assert s["initial_error_rmse_K"]["linear"] < 1e-3
    assert len(res["rows"]) == 4 and set(res["rows"][0]) >= {"cold_ttss", "warm_ttss", "linear_ttss"}


def test_cut_runs_keep_the_energy_per_phonon(monkeypatch):
    """gpu_seconds_to_steady_state groups runs of equal length into one batch and scales each
    budget so that the cut run's energy per phonon equals the full run's."""
    import psim_amd

    params = train_surrogate.structures(train_surrogate.EVAL_SEED, 5)
    batches = []

    def fake_run_batch(engine, models, seed):
        batches.append([(m.num_record, m.prepare(0)) for m in models])
        return {"phonon_steps": 7}

    class Clock:
        def last_timing(self):
            return {"run_ms": 250.0}

    monkeypatch.setattr(psim_amd, "run_batch", fake_run_batch)
    out = surrogate.gpu_seconds_to_steady_state(Clock(), params, None, [400, 1800, None, 400, 2500], 8_000_000,
                                                5000, 25.0, seed=0)
    assert [len(b) for b in batches] == [2, 1, 1, 1]  # 400 x2, 1800, 2500, 5000 (the None)
    assert [b[0][0] for b in batches] == [400, 1800, 2500, 5000]
    assert out == {"gpu_s": 1.0, "phonon_steps": 28, "runs": 4}
    for i, b in zip([0, 1, 4, 2], [batches[0][0], batches[1][0], batches[2][0], batches[3][0]]):
        full = psim_amd.Model(configs.geometry_from_params(dict(params[i], phonons=8_000_000), 5000, 25.0,
                                                          sim_type=1, step_interval=100))
        eff_full = full.prepare(0)[0]
        n_steps, (eff_cut, total_cut, _) = b
        # the cut budget is rounded to whole phonons
        assert abs(eff_cut / eff_full - 1.0) < 1e-5
        assert abs(total_cut - 8_000_000 * n_steps / 5000) <= 2
