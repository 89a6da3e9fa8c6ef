"""Autotuning of the mixed-precision path.

The FP32 ratio tables drift from their FP64 values by rounding in every accepted update;
gpvmc_refresh recomputes them from an FP64 LU and reports the table drift
max|G_old - G_new| / max|G_new| per walker.  What the drift costs is an error in the
estimator: the local energy measured on drifted tables differs from the one measured on
fresh tables at the same configuration.  `tune_refresh` finds the longest refresh
interval (in sweeps) whose energy error per site, averaged over the batch, stays under a
bound, by doubling the interval until the bound is crossed.
"""
import numpy as np


def tune_refresh(batch, sweep, bound=1e-7, kmax=64):
    """Runs sweeps on `batch` (callable `sweep()` does one) and returns (k, history).

    k is the largest interval in {1, 2, 4, ..., kmax} for which
    mean_w |E_loc(drifted tables) - E_loc(refreshed tables)| <= bound (E_loc per site), the two
    energies taken at the same configurations just before and just after the refresh.
    history lists, for every interval tried, a dict with that statistic ("de_mean"), its
    0.999-quantile over walkers ("de_q999"), and the table drift's max, median and
    0.999-quantile.  The batch is refreshed after every trial, so it ends in a freshly
    refreshed state."""
    history = []
    k, best = 1, 1
    while True:
        for _ in range(k):
            sweep()
        e0 = batch.measure()["eloc"]
        d = batch.refresh(drift=True)
        e1 = batch.measure()["eloc"]
        de = np.abs(e0 - e1)
        row = dict(k=k, de_mean=float(de.mean()), de_q999=float(np.quantile(de, 0.999)),
                   drift_max=float(d.max()), drift_median=float(np.median(d)),
                   drift_q999=float(np.quantile(d, 0.999)))
        history.append(row)
        if not np.isfinite(row["de_mean"]) or row["de_mean"] > bound:
            break
        best = k
        if k >= kmax:
            break
        k *= 2
    return best, history
