"""Quantities derived from psim results, computed one way for every backend.

k_eff                 effective thermal conductivity of a bar between two emitting walls
block_means           per-block means of a per-step series, as the reference's periodic export
                      forms them (psim/src/outputManager.cpp:86-111)
time_to_steady_state  the measurement step after which the window-mean sensor temperatures stop
                      moving (RMS over sensors of the change between consecutive windows below a
                      tolerance for several windows in a row)
three_sigma           |mean_a - mean_b| / standard error of the difference, per entry

The reference computes none of these itself; they are applied identically to GPU results, oracle
results and the reference's own tallies.
"""
from __future__ import annotations

import numpy as np


def k_eff(x_flux, length_nm: float, t_hot: float, t_cold: float, interior: bool = True) -> float:
    """Mean x-flux of the sensors (without the two next to the walls when `interior`) times the
    bar length, over the wall temperature difference.  x_flux is per sensor, ordered along the bar."""
    x = np.asarray(x_flux, dtype=np.float64)
    if interior and x.shape[-1] > 2:
        x = x[..., 1:-1]
    return float(np.mean(x) * (length_nm * 1e-9) / (t_hot - t_cold))


def block_means(series, interval: int) -> np.ndarray:
    """Means of consecutive blocks of `interval` steps; a trailing partial block is dropped."""
    s = np.asarray(series, dtype=np.float64)
    n = (s.shape[-1] // interval) * interval
    return s[..., :n].reshape(s.shape[:-1] + (n // interval, interval)).mean(axis=-1)


def time_to_steady_state(temps, window: int = 100, tol: float = 0.05, consecutive: int = 3):
    """temps: [sensors, steps] per-step temperatures from step 0.  Returns the number of steps until
    the RMS over sensors of the change in window-mean temperature between consecutive windows has
    been below `tol` for `consecutive` windows in a row (the end of the last of them), or None."""
    w = block_means(temps, window)  # [S, n_windows]
    if w.shape[-1] < consecutive + 1:
        return None
    rms = np.sqrt(np.mean(np.diff(w, axis=-1) ** 2, axis=0))  # change into window i + 1
    run = 0
    for i, r in enumerate(rms):
        run = run + 1 if r < tol else 0
        if run == consecutive:
            return (i + 2) * window  # rms[i] compares windows i and i + 1
    return None


def three_sigma(a_runs, b_mean, b_sd, b_runs: int, atol: float = 0.0) -> np.ndarray:
    """z = |mean(a) - b_mean| / sqrt(sd(a)^2 / n_a + b_sd^2 / n_b) per entry (a: [runs, ...]).

    Where neither side varies (standard error 0) there is no scale: z is 0 when the means agree
    within `atol` and infinite otherwise, so a constant error cannot pass."""
    a = np.asarray(a_runs, dtype=np.float64)
    se = np.sqrt(a.std(axis=0, ddof=1) ** 2 / a.shape[0] + np.asarray(b_sd, dtype=np.float64) ** 2 / b_runs)
    delta = np.abs(a.mean(axis=0) - np.asarray(b_mean, dtype=np.float64))
    flat = np.where(delta <= atol, 0.0, np.inf)
    return np.where(se > 0, delta / np.where(se > 0, se, 1.0), flat)
