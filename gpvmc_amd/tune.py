"""Autotuning of the mixed-precision path.

The FP32 ratio tables drift from their FP64 values by rounding in every accepted update;
gpvmc_refresh recomputes them from an FP64 Gauss-Jordan inverse and reports the table drift
max|G_old - G_new| / max|G_new| per walker.  What the drift costs is an error in the
estimator: the local energy measured on drifted tables differs from the one measured on
fresh tables at the same configuration.  `tune_refresh` finds the longest refresh
interval (in sweeps) whose energy error per site, averaged over the batch, stays under a
bound, by doubling the interval until the bound is crossed.

`pick_launch` picks a context's launch configuration (sweep shape and sweep variant for a
ground-state size, register-resident or streamed kernel for a channel) by timing each
candidate, built by the caller on the same walkers, and keeps the fastest open.
"""
import time

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


def launch_candidates(L, precision, channel=False):
    """Launch configurations to try: dicts of Batch / ChannelBatch keyword arguments plus
    the sweep variant ("variant", ground state only)."""
    from . import binding
    if channel:
        return [dict(kernel="auto"), dict(kernel="streamed")]
    out = []
    for lpr in binding.sweep_shapes(L, precision):
        out.append(dict(lanes_per_row=lpr, variant=binding.SWEEP_ONE_MOVE))
        if precision == binding.MIXED and L == 16 and lpr == 4:
            out.append(dict(lanes_per_row=lpr, variant=binding.SWEEP_WINDOWS))
    return out


def pick_launch(make, candidates, steps, sweeps=2, meas=0, mask=None):
    """Builds the context of every candidate with make(**kwargs) (the "variant" key is set
    on the built context), times `sweeps` walks of `steps` steps (measured every `meas`
    steps) after one warm-up walk, and returns (the fastest context, still open, its
    candidate, history).  A candidate that builds the same kernel as an earlier one (a
    channel without a register-resident shape) or whose variant the shape lacks is
    skipped; the others are closed.  When only one distinct candidate remains it is
    returned untimed."""
    from . import binding
    walk_kw = dict(meas=meas)
    if mask is not None:
        walk_kw["mask"] = mask

    def timed(b):
        b.walk(steps, **walk_kw)
        t0 = time.perf_counter()
        for _ in range(sweeps):
            b.walk(steps, **walk_kw)
        return (time.perf_counter() - t0) / sweeps

    history = []
    best = best_c = best_t = best_key = None
    seen = set()
    for cand in candidates:
        kw = {k: v for k, v in cand.items() if k != "variant"}
        b = make(**kw)
        key = repr(sorted(kw.items()))
        if hasattr(b, "register_resident"):
            key = "register" if b.register_resident else "streamed"
        if "variant" in cand:
            try:
                b.sweep_variant = cand["variant"]
            except binding.GpvmcError:
                b.close()
                continue
            key += " variant %d" % b.sweep_variant
        if key in seen:
            b.close()
            continue
        seen.add(key)
        if best is None:
            best, best_c, best_key = b, cand, key
            continue
        if best_t is None:
            best_t = timed(best)
            history.append(dict(candidate=best_key, ms_per_sweep=round(1e3 * best_t, 3)))
        t = timed(b)
        history.append(dict(candidate=key, ms_per_sweep=round(1e3 * t, 3)))
        if t < best_t:
            best.close()
            best, best_c, best_t, best_key = b, cand, t, key
        else:
            b.close()
    return best, best_c, history
