"""GPU results against independent runs of the reference build (tests/golden/ref_*.json).

The reference's RNG cannot be seeded, so agreement is statistical: over 16 GPU seeds and 16
reference runs every sensor's mean temperature and x/y flux must agree within 3 standard errors
of the difference, |mean_gpu - mean_ref| <= 3 sqrt(s_gpu^2/16 + s_ref^2/16).  The transient
film is also compared block by block (100-step block means, the periodic export's series):
there the 3-sigma bound is applied to each of the 600 entries as a family, allowing the
expected few exceedances (at most 1 %) and none beyond 4.5 sigma.  The effective
conductivity of the 1000 nm bar (1e6 phonons, 48 runs on each side) must agree within 1 %.
"""
import json
import os

import numpy as np
import pytest

import psim_amd
from psim_amd import analysis

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _fixture(name):
    with open(os.path.join(GOLDEN, f"ref_{name}.json")) as f:
        return json.load(f)


@pytest.fixture(scope="module")
def engine():
    e = psim_amd.Engine(0)
    yield e
    e.close()


def _gpu_runs(engine, model, seeds):
    m = psim_amd.Model(model)
    out = []
    for seed in seeds:
        m.run(engine, seed)
        r = m.results(0)
        out.append(r)
    return out


@pytest.mark.parametrize("name", ["linear_ss", "lshape_r", "film_r"])
def test_sensor_means_within_three_sigma(engine, name):
    fix = _fixture(name)
    print(f"case {name}: {fix['runs']} seeds", flush=True)
    runs = _gpu_runs(engine, fix["model"], range(fix["runs"]))
    for col, key in [(0, "t_steady"), (2, "x_flux"), (4, "y_flux")]:
        z = analysis.three_sigma([r.measurements[:, col] for r in runs], fix["mean"][key], fix["sd"][key],
                                 fix["runs"])
        assert np.all(z <= 3.0), (key, np.round(z, 2))
    if "block_mean" in fix:
        blocks = [analysis.block_means(r.temps, fix["block_interval"]) for r in runs]
        z = analysis.three_sigma(blocks, fix["block_mean"], fix["block_sd"], fix["runs"])
        assert np.mean(z > 3.0) <= 0.01 and z.max() <= 4.5, (np.mean(z > 3.0), z.max())


def test_effective_conductivity_within_one_percent(engine):
    fix = _fixture("linear_k")
    print(f"k_eff: {fix['runs']} seeds of 1e6 phonons", flush=True)
    runs = _gpu_runs(engine, fix["model"], range(fix["runs"]))
    k_gpu = np.array([analysis.k_eff(r.measurements[:, 2], 1000.0, 310.0, 290.0) for r in runs])
    k_ref = np.array([analysis.k_eff(x, 1000.0, 310.0, 290.0) for x in fix["per_run"]["x_flux"]])
    rel = abs(k_gpu.mean() - k_ref.mean()) / k_ref.mean()
    se = np.sqrt(k_gpu.var(ddof=1) / len(k_gpu) + k_ref.var(ddof=1) / len(k_ref)) / k_ref.mean()
    print(f"k_eff gpu {k_gpu.mean():.5g} ref {k_ref.mean():.5g} rel diff {rel:.4%} (se {se:.4%})", flush=True)
    assert rel <= 0.01
    assert rel <= 3 * se
