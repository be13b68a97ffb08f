"""Output side: the ss_ / per_ files (OutputManager::exportResults, psim/src/outputManager.cpp:13-112)
and the psim_gpu command-line driver (psim/src/main.cpp:10-31).

The files are read the way the reference's plotting code reads them
(psim_python/psim/plotting_tools.py:80-92 for ss_, :136-157 for per_) and their numbers must be the
run's results as the default C++ stream prints them (6 significant digits).
"""
import json
import os
import re
import subprocess

import numpy as np
import pytest

import psim_amd
from psim_amd import analysis, configs

pytestmark = pytest.mark.gpu

HEADER = re.compile(r'^(Steady State|Periodic) Results from "(?P<name>[^"]+)" @ \d{4}-\d{2}-\d{2} \d{2}:\d{2}:\d{2} '
                    r'- Time Taken (?P<time>[0-9.e+-]+)\[s\] over (?P<runs>\d+) runs$')


def parse_ss(path):
    """plotting_tools.parse_ss_data: a title line, then six numbers per sensor."""
    lines = open(path).read().splitlines()
    rows = [[float(v) for v in line.split()] for line in lines[1:]]
    return lines[0], np.array(rows)


def parse_per(path):
    """plotting_tools.parse_periodic_data: a title line, then per block its middle step, the
    number of sensors and one line of (T, x-flux, y-flux) per sensor."""
    lines = open(path).read().splitlines()
    body = lines[1:]
    n = int(body[1])
    steps, blocks = [], []
    for i in range(0, len(body), n + 2):
        chunk = body[i:i + n + 2]
        steps.append(int(chunk[0]))
        assert int(chunk[1]) == n
        blocks.append([[float(v) for v in line.split()] for line in chunk[2:]])
    return lines[0], steps, np.array(blocks)


@pytest.fixture(scope="module")
def engine():
    e = psim_amd.Engine(0)
    yield e
    e.close()


def test_steady_state_file(engine, tmp_path):
    cfg = configs.with_settings(configs.linear_bar(num_phonons=20000, num_measurements=1000, sim_time=5.0,
                                                   spec=0.3).to_dict(), num_runs=2)
    src = tmp_path / "bar model.json"
    src.write_text(json.dumps(cfg))
    model = psim_amd.Model(cfg)
    model.run(engine, 3)
    model.export(str(src), 1.25)
    title, rows = parse_ss(tmp_path / "ss_bar model.txt")
    m = HEADER.match(title)
    assert m and m.group(1) == "Steady State" and m.group("name") == "bar model.json"
    assert float(m.group("time")) == 1.25 and int(m.group("runs")) == 2
    # the average over runs of (t_steady, std, x_flux, std, y_flux, std), sensor by sensor
    want = (model.results(0).measurements + model.results(1).measurements) / 2.0
    assert rows.shape == want.shape
    np.testing.assert_allclose(rows, want, rtol=6e-6, atol=0)


def test_periodic_file(engine, tmp_path):
    cfg = configs.linear_bar(num_phonons=20000, num_measurements=1000, sim_time=5.0, spec=0.3, sim_type=1,
                             step_interval=100).to_dict()
    src = tmp_path / "bar.json"
    src.write_text(json.dumps(cfg))
    model = psim_amd.Model(cfg)
    model.run(engine, 4)
    model.export(str(src), 0.5)
    title, steps, blocks = parse_per(tmp_path / "per_bar.txt")
    assert HEADER.match(title).group(1) == "Periodic"
    # every whole 100-step block, labelled by its middle step (outputManager.cpp:86-87)
    assert steps == [50 + 100 * k for k in range(10)]
    res = model.results(0)
    temps = analysis.block_means(res.temps, 100)  # [S, blocks]
    fx = analysis.block_means(res.fluxes[..., 0], 100)
    fy = analysis.block_means(res.fluxes[..., 1], 100)
    want = np.stack([temps.T, fx.T, fy.T], axis=-1)  # [blocks, S, 3]
    assert blocks.shape == want.shape
    np.testing.assert_allclose(blocks, want, rtol=6e-6, atol=1e-300)


def test_command_line_driver(tmp_path):
    """psim_gpu file1.json file2.json: each model read, run and written next to its input; a
    missing file is reported and skipped, as the reference's main does."""
    exe = os.path.join(os.path.dirname(psim_amd.LIB_PATH), "bin", "psim_gpu")
    a = configs.linear_bar(num_phonons=5000, num_measurements=500, sim_time=2.5).to_dict()
    b = configs.linear_bar(num_phonons=5000, num_measurements=500, sim_time=2.5, sim_type=1,
                           step_interval=50).to_dict()
    (tmp_path / "a.json").write_text(json.dumps(a))
    (tmp_path / "b.json").write_text(json.dumps(b))
    out = subprocess.run([exe, str(tmp_path / "a.json"), str(tmp_path / "missing.json"), str(tmp_path / "b.json")],
                         capture_output=True, text=True, timeout=120, env=dict(os.environ, PSIM_SEED="7"))
    assert out.returncode == 0, out.stderr
    assert 'Error opening file at "' in out.stderr
    assert out.stdout.count("Time Taken") == 2 and out.stdout.rstrip().endswith("done")
    title, rows = parse_ss(tmp_path / "ss_a.txt")
    assert HEADER.match(title) and rows.shape == (10, 6)
    _, steps, blocks = parse_per(tmp_path / "per_b.txt")
    assert steps == [25 + 50 * k for k in range(10)] and blocks.shape == (10, 10, 3)
