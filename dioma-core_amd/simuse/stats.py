"""Firing statistics of spike rasters and their relative errors.

Rates and ISI CVs follow firing_rates (crates/SiMuse/crates/pillars/src/
kernels/stats/rates.rs:6-35: count / (steps * dt)) and compute_cvs
(stats/isi.rs:30-67: sample standard deviation over mean of each neuron's
ISIs, neurons with at least two ISIs).  Synchrony is chi = sqrt(Var_t(mean_i
s_i(t)) / mean_i Var_t(s_i(t))) on the binary spike trains (Golomb 2007).
"""
from __future__ import annotations

import numpy as np


def rates_hz(raster: np.ndarray, dt_s: float) -> np.ndarray:
    """Per-neuron firing rate of a [steps, n] 0/1 raster."""
    t = raster.shape[0]
    return raster.sum(axis=0, dtype=np.float64) / max(t * dt_s, 1e-300)


def isi_cvs(raster: np.ndarray, dt_s: float) -> np.ndarray:
    """Per-neuron ISI CV (0 for neurons with fewer than two ISIs)."""
    t_idx, n_idx = np.nonzero(raster.T.astype(bool))  # grouped by neuron, ascending time
    out = np.zeros(raster.shape[1], np.float64)
    if t_idx.size == 0:
        return out
    times = n_idx.astype(np.float64) * dt_s
    same = t_idx[1:] == t_idx[:-1]
    isi = np.diff(times)[same]
    who = t_idx[1:][same]
    cnt = np.bincount(who, minlength=raster.shape[1])
    s1 = np.bincount(who, weights=isi, minlength=raster.shape[1])
    s2 = np.bincount(who, weights=isi * isi, minlength=raster.shape[1])
    ok = cnt >= 2
    mean = np.where(ok, s1 / np.maximum(cnt, 1), 0.0)
    var = np.where(ok, (s2 - cnt * mean * mean) / np.maximum(cnt - 1, 1), 0.0)
    out[ok] = np.sqrt(np.maximum(var[ok], 0.0)) / mean[ok]
    return out


def chi(raster: np.ndarray) -> float:
    x = raster.astype(np.float64)
    return float(np.sqrt(x.mean(1).var() / max(float(np.mean(x.var(0))), 1e-30)))


def summary(raster: np.ndarray, dt_s: float) -> dict:
    """Mean rate (Hz), mean CV over neurons with a CV, chi."""
    cv = isi_cvs(raster, dt_s)
    return {"rate_hz": float(rates_hz(raster, dt_s).mean()),
            "isi_cv": float(cv[cv > 0].mean()) if np.any(cv > 0) else 0.0,
            "chi": chi(raster)}


def rel_errors(dev: dict, ref: dict) -> dict:
    """|dev - ref| / |ref| per statistic (absolute error where ref is 0)."""
    out = {}
    for k in ("rate_hz", "isi_cv", "chi"):
        r = ref[k]
        out[k + "_rel_err"] = abs(dev[k] - r) / abs(r) if r != 0 else abs(dev[k])
    return out
