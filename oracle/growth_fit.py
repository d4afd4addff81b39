"""TEST INFRASTRUCTURE, NOT PRODUCT CODE.

numpy restatement of the reference's growth-rate analysis, used by tests/ as
the checker of the library's C++ fitters and device amplitude reduction:

  circular_amplitude   ymgpu2d/analysis/remote_timeseries.py:62-77 (fast path)
  fit_max_r2           ymgpu2d/analysis/batch_analyze.py:44-69
  fit_plateau          ymgpu2d/analysis/batch_analyze.py:72-152

Pinning: tests/test_growth_fit.py checks these against closed forms (a pure
exponential gives its exact rate with R^2 = 1; a planted Fourier mode gives
its amplitude) before comparing the library with them.
"""
import numpy as np


def circular_amplitude(az2, az3, k):
    nz = az2.shape[0]
    sig = az2 + 1j * az3
    ker = np.exp(-2j * np.pi * k * np.arange(nz) / nz)
    F = ker @ sig
    Fc = ker @ np.conj(sig)
    return float(np.abs(F).mean() / nz), float(np.abs(Fc).mean() / nz)


def fit_max_r2(times, amps, t_min=None, t_max=None, min_win_frac=0.15):
    log_a = np.log(np.clip(amps, 1e-30, None))
    mask = np.ones(len(times), dtype=bool)
    if t_min is not None:
        mask &= times >= t_min
    if t_max is not None:
        mask &= times <= t_max
    tw, aw = times[mask], log_a[mask]
    if len(tw) < 4:
        return 0.0, np.nan, np.nan, 0.0
    min_win = max(4, int(len(tw) * min_win_frac))
    best = (0.0, np.nan, np.nan, -1e9)
    for i in range(len(tw)):
        for j in range(i + min_win, len(tw) + 1):
            ts, as_ = tw[i:j], aw[i:j]
            c = np.polyfit(ts, as_, 1)
            if c[0] < 1e-6:
                continue
            res = as_ - np.polyval(c, ts)
            r2 = 1.0 - np.sum(res ** 2) / (np.sum((as_ - as_.mean()) ** 2) + 1e-30)
            if r2 > best[3]:
                best = (c[0], ts[0], ts[-1], r2)
    return best


def fit_plateau(times, amps, win_tu=8.0, min_plateau_tu=10.0, flat_tol=0.06):
    log_a = np.log(np.clip(amps, 1e-30, None))
    empty = dict(gamma=np.nan, t0=np.nan, t1=np.nan, duration=np.nan, has_plateau=False)
    if len(times) < 8:
        return empty
    half = win_tu / 2.0
    tc, slopes = [], []
    for i in range(len(times)):
        m = (times >= times[i] - half) & (times <= times[i] + half)
        if m.sum() < 4:
            continue
        tc.append(times[i])
        slopes.append(np.polyfit(times[m], log_a[m], 1)[0])
    tc, slopes = np.array(tc), np.array(slopes)
    if len(tc) < 4:
        return empty
    pos = slopes > 1e-6
    n, best, i = len(tc), None, 0
    while i < n:
        if not pos[i]:
            i += 1
            continue
        run = [i]
        j = i + 1
        while j < n and pos[j]:
            vals = slopes[run + [j]]
            if (vals.max() - vals.min()) / np.median(vals) <= flat_tol:
                run = run + [j]
                j += 1
            else:
                break
        dur = tc[run[-1]] - tc[run[0]]
        if dur >= min_plateau_tu and (best is None or dur > best[1]):
            best = (run, dur)
        i = run[-1] + 1 if len(run) > 1 else i + 1
    if best is None:
        return empty
    run, dur = best
    t0, t1 = tc[run[0]], tc[run[-1]]
    m = (times >= t0) & (times <= t1)
    return dict(gamma=float(np.polyfit(times[m], log_a[m], 1)[0]), t0=float(t0), t1=float(t1),
                duration=float(dur), has_plateau=True)
