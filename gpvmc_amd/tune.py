"""Autotuning of the mixed-precision path.

The FP32 ratio tables drift from their FP64 values by rounding in every accepted update;
gpvmc_refresh recomputes them from an FP64 LU and reports the drift
max|G_old - G_new| / max|G_new| per walker.  `tune_refresh` finds the longest refresh
interval (in sweeps) whose drift stays under a bound, by doubling the interval until the
bound is crossed.
"""
import numpy as np


def tune_refresh(batch, sweep, bound=1e-6, kmax=64, stat="max"):
    """Runs sweeps on `batch` (callable `sweep()` does one) and returns
    (k, history) where k is the largest interval in {1, 2, 4, ..., kmax} whose drift
    statistic (max or 0.999-quantile over walkers) is <= bound, and history lists
    (k, max, median, quantile_0.999) for every interval tried.  The batch is refreshed
    after every trial, so it ends in a freshly refreshed state."""
    history = []
    k, best = 1, 1
    while True:
        for _ in range(k):
            sweep()
        d = batch.refresh(drift=True)
        row = (k, float(d.max()), float(np.median(d)), float(np.quantile(d, 0.999)))
        history.append(row)
        value = row[1] if stat == "max" else row[3]
        if value > bound:
            break
        best = k
        if k >= kmax:
            break
        k *= 2
    return best, history
