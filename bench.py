#!/usr/bin/env python3
"""Throughput of the batched Metropolis VMC hot path on one MI355X.

Headline workload (the L=16 configuration): staggered-flux ground state, N=128 particles per
flavour, 16384 walkers on an 8-point (phi, neel) grid (phi in {0.10, 0.1667, 0.2333, 0.30} x
neel in {0, 0.1}, walker w -> point w mod 8), mixed precision (FP32 ratio tables, FP64 LU
refresh, FP64 log-amplitude and sums).  A sweep is L^2 = 256 Metropolis steps on every
walker followed by one ProjHeis + StagMagnZ measurement (meas_interv = L^2, the vmc.cpp:121
default).  One bench step is one refresh period: k sweeps and the FP64 refresh that closes
it, k tuned beforehand (untimed) so that the energy error of drifted tables stays under a
bound; the timed region therefore always holds whole periods, refresh included.

Prints ONE JSON line: value = MC steps/s of the mixed path (whole job, inputs resident in
HBM), the plain-FP64 rate on the same workload, the FP32-vs-FP64 energy error on the most
drifted tables (just before a refresh), energies and staggered magnetisation of both paths
with error bars, the rooflines of the sweep and refresh kernels (HIP events on the
library's stream), further legs (the L=4 reference run, L=8 FP64 with StatSpinStruct, the
L=12 exciton channel, the uneven mixed batch), and a CPU baseline: the reference's own
MetroMC::Step, built from its sources, one chain per host core.
"""
import argparse
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "MC steps/s + ProjHeis E_loc samples/s at L=16; mixed-FP32 vs plain FP64"
PHIS = (0.10, 0.1667, 0.2333, 0.30)
NEELS = (0.00, 0.10)
HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3      # MI355X dense FP32 matrix (and packed-vector) peak
FP64_PEAK_TFLOPS = 78.6


def log(*a):
    print("bench:", *a, file=sys.stderr, flush=True)


def grid_orbitals(L):
    from gpvmc_amd import model
    return [model.ground_occ(model.orbital_table(L, p * np.pi, n)) for p in PHIS for n in NEELS]


def make_ground(precision, L, W, seed, occs, spins=None):
    """A ground-state context on the (phi, neel) grid, started from `spins` or from random
    configurations with the reference's redraw of singular starts (vmc.cpp:222-241)."""
    from gpvmc_amd import binding, model
    b = binding.Batch(L, W, precision, n_groups=len(occs), seed=seed)
    for gi, oc in enumerate(occs):
        b.set_orbitals(gi, oc)
    b.set_groups(np.arange(W, dtype=np.int32) % len(occs))
    if spins is not None:
        ok = b.init(spins)
        if not ok.all():
            raise RuntimeError("singular configuration handed over")
        return b, spins
    spins = model.random_spins(L, W, seed)
    ok = b.init(spins)
    tries = 1
    while not ok.all() and tries < 40:
        spins[~ok] = model.random_spins(L, W, seed + 1000 * tries)[~ok]
        ok = b.init(spins)
        tries += 1
    if not ok.all():
        raise RuntimeError("could not find regular starting configurations")
    return b, spins


def per_walker_stats(acc, groups, G):
    """energy per site and <|m_s|^2> as (mean, standard error) per parameter group and
    over all walkers, from per-walker means (independent chains)"""
    n = np.maximum(acc[:, 3], 1)
    e, m2 = acc[:, 0] / n, acc[:, 6] / n

    def ms(x):
        return [float(x.mean()), float(x.std(ddof=1) / np.sqrt(len(x)))]
    return {"energy_per_site": ms(e), "stag_magn_sq": ms(m2),
            "by_group": [{"phi": PHIS[g // len(NEELS)], "neel": NEELS[g % len(NEELS)],
                          "energy_per_site": ms(e[groups == g]), "stag_magn_sq": ms(m2[groups == g])}
                         for g in range(G)]}


def de_stats(d):
    a = np.abs(d)
    return {"max": float(a.max()), "q999": float(np.quantile(a, 0.999)), "mean_abs": float(a.mean()),
            "mean_signed": float(d.mean())}


# ----------------------------------------------------------------------------- CPU baseline
def cpu_cores():
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    return max(1, min(n, int(os.environ.get("OMP_NUM_THREADS", "16") or 16), 16))


def cpu_baseline(L, budget_s):
    """The reference's own MetroMC::Step (oracle/_ref/ref_harness, built from the reference
    sources, MKL sequential) as one independent chain per host core, rates summed; falls
    back to the oracle's C restatement, one chain per thread, when the reference build is
    absent."""
    P = cpu_cores()
    harness = os.path.join(ROOT, "oracle", "_ref", "ref_harness")
    env = dict(os.environ, MKL_NUM_THREADS="1", OMP_NUM_THREADS="1")
    if os.path.exists(harness):
        try:
            probe = json.loads(subprocess.run([harness, "time", str(L), "0.1667", "0.1", "1", "5000"],
                                              check=True, capture_output=True, text=True, env=env,
                                              timeout=120).stdout)
            n = int(max(5000, min(5e6, probe["steps_per_s"] * budget_s)))
            procs = [subprocess.Popen([harness, "time", str(L), "0.1667", "0.1", str(1 + i), str(n)],
                                      stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, env=env)
                     for i in range(P)]
            rates, secs = [], []
            for p in procs:
                out, err = p.communicate(timeout=600)
                if p.returncode != 0:
                    raise RuntimeError("harness failed: " + err[-300:])
                r = json.loads(out)
                rates.append(r["steps_per_s"])
                secs.append(r["seconds"])
            return {"value": round(sum(rates), 1), "unit": "MC steps/s", "cores": P, "kind": "reference",
                    "sample": "reference MetroMC::Step (LatticeStepper + SlaterDeterminant, MKL sequential), "
                              "%d independent chains on %d host cores, L=%d phi=0.1667 neel=0.1, %d steps each "
                              "in %.1f-%.1f s; per-chain rate %.1f" % (P, P, L, n, min(secs), max(secs),
                                                                       sum(rates) / P)}
        except Exception as e:  # noqa: BLE001
            log("reference harness unavailable: %s" % e)
    from concurrent.futures import ThreadPoolExecutor
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_ref as O
    from gpvmc_amd import model
    occ = O.ground_occ(L, O.orbitals(L, 0.1667 * np.pi, 0.1))
    w = O.Walker(L, occ, model.random_spins(L, 1, 1)[0])
    n = 2000
    t0 = time.perf_counter()
    w.walk(model.philox_uniforms(1, [0], 0, n).reshape(-1))
    n = int(max(n, min(2e6, n / (time.perf_counter() - t0) * budget_s)))

    def chain(i):
        wi = O.Walker(L, occ, model.random_spins(L, 1, 2 + i)[0])
        u = model.philox_uniforms(2 + i, [0], 0, n).reshape(-1)
        t = time.perf_counter()
        wi.walk(u)
        return n / (time.perf_counter() - t)
    with ThreadPoolExecutor(P) as ex:
        rates = list(ex.map(chain, range(P)))
    return {"value": round(sum(rates), 1), "unit": "MC steps/s", "cores": P, "kind": "port",
            "sample": "oracle C restatement, %d chains on %d threads at L=%d, %d steps each" % (P, P, L, n)}


def cpu_eloc_rate(L, budget_s):
    """ProjHeis E_loc samples/s of the oracle's restatement (the reference's ProjHeis needs
    a BLAS zdotu), one configuration per thread, one thread per host core, summed"""
    from concurrent.futures import ThreadPoolExecutor
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_ref as O
    from gpvmc_amd import model
    occ = O.ground_occ(L, O.orbitals(L, 0.1667 * np.pi, 0.1))
    P = cpu_cores()

    def one(i):
        w = O.Walker(L, occ, model.random_spins(L, 1, 2 + i)[0])
        n, t0 = 0, time.perf_counter()
        while time.perf_counter() - t0 < budget_s:
            w.eloc()
            n += 1
        return n / (time.perf_counter() - t0)
    with ThreadPoolExecutor(P) as ex:
        return sum(ex.map(one, range(P))), P


# ----------------------------------------------------------------------------- legs
def c1_leg(seed, n_meas=100000, W_batch=16384):
    """The reference's L=4 run (vmc.cpp defaults, phi=0.25, neel=0.1): one FP64 chain,
    therm 200 L^2 steps, then n_meas measurements every 16 steps (ProjHeis + StagMagnZ);
    the same workload on W_batch independent chains."""
    from gpvmc_amd import binding, model
    L, meas = 4, 16
    occ = model.ground_occ(model.orbital_table(L, 0.25 * np.pi, 0.1))
    out = {"workload": "L=4 phi=0.25 neel=0.1, FP64, therm 3200 steps, %d measurements every 16 steps "
                       "(ProjHeis + StagMagnZ)" % n_meas}
    b, _ = make_ground(binding.FP64, L, 1, seed, [occ])
    b.walk(200 * L * L)
    b.reset_accumulators()
    t0 = time.perf_counter()
    b.walk(n_meas * meas, meas=meas)
    dt = time.perf_counter() - t0
    acc = b.accumulators()
    b.close()
    out["one_chain"] = {"mc_steps_per_s": round(n_meas * meas / dt, 1), "seconds": round(dt, 3),
                        "energy_per_site": round(float(acc[0, 0] / acc[0, 3]), 6),
                        "stag_magn_sq": round(float(acc[0, 6] / acc[0, 3]), 6)}
    n_b = max(1, n_meas // W_batch)
    b, _ = make_ground(binding.FP64, L, W_batch, seed + 1, [occ])
    b.walk(200 * L * L)
    b.reset_accumulators()
    t0 = time.perf_counter()
    b.walk(n_b * 256, meas=meas)
    dt = time.perf_counter() - t0
    acc = b.accumulators()
    b.close()
    st = per_walker_stats(acc, np.zeros(W_batch, int), 1)
    out["batched"] = {"walkers": W_batch, "mc_steps_per_s": round(W_batch * n_b * 256 / dt, 1),
                      "measurements": int(acc[:, 3].sum()),
                      "energy_per_site": [round(x, 7) for x in st["energy_per_site"]],
                      "stag_magn_sq": [round(x, 7) for x in st["stag_magn_sq"]]}
    try:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import exact
        confs = exact.configurations(L)
        psi = exact.spin_basis_psi(confs, occ)
        w = np.abs(psi) ** 2
        e = exact.local_energy_exact(L, confs, psi, model.edges(L)).real
        out["exact_energy_per_site"] = round(float((w * e).sum() / w.sum()), 8)
    except Exception as e:  # noqa: BLE001
        log("exact enumeration unavailable: %s" % e)
    return out


def c2_leg(seed, sweeps=20, W=4096):
    """L=8 FP64 (phi=0.25, neel=0.1, N=32), 4096 walkers, therm 200 L^2 steps, then sweeps
    of L^2 steps with ProjHeis + StagMagnZ + StatSpinStruct measured once per sweep."""
    from gpvmc_amd import binding, model
    L = 8
    occ = model.ground_occ(model.orbital_table(L, 0.25 * np.pi, 0.1))
    b, _ = make_ground(binding.FP64, L, W, seed, [occ])
    mask = binding.MEAS_PROJHEIS | binding.MEAS_STAGMAGNZ | binding.MEAS_STATSPIN
    b.walk(200 * L * L)
    b.walk(L * L, meas=L * L, mask=mask)  # allocates the StatSpinStruct sums
    b.reset_accumulators()
    b.reset_kernel_time()
    t0 = time.perf_counter()
    for _ in range(sweeps):
        b.walk(L * L, meas=L * L, mask=mask)
    dt = time.perf_counter() - t0
    kt = b.kernel_time()
    acc = b.accumulators()
    b.close()
    st = per_walker_stats(acc, np.zeros(W, int), 1)
    return {"workload": "L=8 phi=0.25 neel=0.1, plain FP64, %d walkers, therm 12800 steps, %d sweeps of 64 "
                        "steps + 1 ProjHeis/StagMagnZ/StatSpinStruct measurement" % (W, sweeps),
            "mc_steps_per_s": round(W * L * L * sweeps / dt, 1),
            "eloc_samples_per_s": round(W * sweeps / dt, 1),
            "ms_per_sweep": round(1e3 * dt / sweeps, 3),
            "sweep_kernel_ms": round(kt["sweep_ms"] / max(kt["sweep_launches"], 1), 3),
            "energy_per_site": [round(x, 7) for x in st["energy_per_site"]],
            "stag_magn_sq": [round(x, 7) for x in st["stag_magn_sq"]]}


def exciton_leg(seed, periods, k, W=8192, L=12, n_cmp=1024):
    """Transverse-exciton channel at L=12 (N_up = 73, N_dn = 71, 72 states), every momentum q
    on the 12 x 12 grid (walker w on q = w mod 144), mixed FP32; one period = k sweeps of
    L*L steps and one ProjHeis H/O measurement each, closed by the FP64 refresh.  Also the
    per-amplitude error of the FP32 walk against an FP64 recomputation on the same
    configurations just before the last refresh, and the MFMA contraction's throughput."""
    from gpvmc_amd import binding, channels, model
    qs = [(qx, qy) for qx in range(L) for qy in range(L)]
    chs = [channels.Channel("trans", L, q) for q in qs]
    tab = model.orbital_table(L, 0.25 * np.pi, 0.1)
    Ns = L * L
    spins = model.random_spins(L, W, seed)
    # the trans channel has N + 1 up particles: move one down spin per walker up
    rng = np.random.default_rng(seed)
    for w in range(W):
        dn = np.nonzero(spins[w] == 1)[0]
        spins[w, dn[rng.integers(len(dn))]] = 0
    b = binding.ChannelBatch(L, W, chs, binding.MIXED, seed=seed)
    b.set_orbital_table(0, tab)
    ok = b.init(spins)
    for _ in range(2):
        b.walk(Ns)
    b.refresh()
    b.reset_kernel_time()
    b.reset_accumulators()
    t0 = time.perf_counter()
    for _ in range(periods):
        # one walk of k sweeps, measuring after every L*L steps (MetroMC::Walk(k Ns, Ns))
        b.walk(k * Ns, meas=Ns)
        b.refresh()
    tm = time.perf_counter() - t0
    kt = b.kernel_time()
    acc = b.accumulators()
    # the contraction alone: measurement launches timed on the host around a synchronised call
    t1 = time.perf_counter()
    nm = 3
    for _ in range(nm):
        b.measure()
    t_meas = (time.perf_counter() - t1) / nm
    # drifted state: k more sweeps without a refresh, then FP32 amplitudes vs FP64 ones
    for _ in range(k):
        b.walk(Ns)
    a32 = b.amplitudes()
    sp = b.state()[0]
    n_up, n_dn = chs[0].npt
    b.close()
    Wc = min(W, n_cmp)
    b64 = binding.ChannelBatch(L, Wc, chs, binding.FP64, seed=seed)
    b64.set_orbital_table(0, tab)
    b64.init(sp[:Wc])
    a64 = b64.amplitudes()
    b64.close()
    km = np.argmax(a64.real, axis=1)
    r32 = a32[:Wc] - a32[np.arange(Wc), km][:, None]
    r64 = a64 - a64[np.arange(Wc), km][:, None]
    err = np.abs(np.exp(r32) - np.exp(r64))  # |a_k/a_max|, FP32 vs FP64
    # MFMA work of one measurement per walker: Z_fl = [Ai; Q]_fl x Phi(orbitals, other sites),
    # 8 real flops per complex multiply-add
    nq = (n_dn - 1, 1)
    flops = W * 8 * ((n_up + nq[0]) * n_up * n_dn + (n_dn + nq[1]) * n_dn * n_up)
    sweeps = periods * k
    return {
        "workload": "L=12 transverse exciton, all 144 q (walker w -> q = w mod 144), N_up=73 N_dn=71, "
                    "72 states, %d walkers, mixed FP32; timed: %d refresh period(s) of %d sweeps (144 steps + "
                    "1 ProjHeis H/O measurement each) and the FP64 refresh" % (W, periods, k),
        "mc_steps_per_s": round(W * Ns * sweeps / tm, 1),
        "ms_per_sweep_incl_refresh": round(1e3 * tm / sweeps, 3),
        "sweeps": sweeps, "refreshes": periods,
        "kernel_ms_per_sweep": round(kt["sweep_ms"] / sweeps, 3),
        "walk_launches": kt["sweep_launches"],
        "refresh_ms": round(kt["refresh_ms"] / max(kt["refresh_launches"], 1), 3),
        "measure_ms": round(1e3 * t_meas, 3),
        "p_accept_move": round(float(acc[:, 8].sum() / acc[:, 7].sum()), 4),
        "p_noop": round(float(acc[:, 10].sum() / acc[:, 7].sum()), 4),
        "init_ok": float(np.mean(ok)),
        "amplitude_error_fp32_vs_fp64": {
            "unit": "|a_k/a_max (FP32) - a_k/a_max (FP64)|, same configurations, %d sweeps after the last "
                    "refresh" % k,
            "walkers": Wc, "median": float(np.median(err)), "q99": float(np.quantile(err, 0.99)),
            "max": float(err.max())},
        "contraction_roofline": {"bound": "mfma", "achieved": round(flops / t_meas / 1e12, 3),
                                 "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                                 "frac": round(flops / t_meas / 1e12 / FP32_PEAK_TFLOPS, 4),
                                 "note": "v_mfma_f32_16x16x4_f32 contraction of one measurement "
                                         "(host-timed launch incl. the H/O estimator)"},
    }


def mixed_batch_leg(seed, periods, k, total=20000, Ls=(8, 10, 12, 16, 20)):
    """Walkers of uneven lattice sizes, ground state and longitudinal-exciton channels
    interleaved: lattice side and kind drawn per walker from a seeded categorical
    (uniform over Ls x {groundstate, long}); one context per (L, kind) bucket, all swept
    together through one captured hipGraph per sweep (gpvmc_batch_*).  A sweep is L*L steps
    and one measurement per walker (ProjHeis + StagMagnZ for the ground state, ProjHeis
    H/O for the channels), mixed FP32; one period is k sweeps and the FP64 refresh of every
    bucket.  Each bucket's launch configuration (sweep shape and variant, or register-
    resident vs streamed channel kernel) is picked by timing the candidates on the bucket's
    walkers (gpvmc_amd.tune.pick_launch)."""
    from gpvmc_amd import binding, channels, model, tune
    rng = np.random.default_rng(seed)
    Lw = rng.choice(np.array(Ls), size=total)
    kind = rng.integers(0, 2, size=total)
    sw = binding.Sweeper()
    buckets = []
    members = []
    for L in Ls:
        for kd in (0, 1):
            n = int(np.sum((Lw == L) & (kind == kd)))
            if n == 0:
                continue
            tab = model.orbital_table(L, 0.25 * np.pi, 0.1)
            oks = {}
            if kd == 0:
                def make(lanes_per_row=0):
                    b = binding.Batch(L, n, binding.MIXED, n_groups=1, seed=seed + L,
                                      lanes_per_row=lanes_per_row)
                    b.set_orbitals(0, model.ground_occ(tab))
                    oks[id(b)] = b.init(model.random_spins(L, n, seed + L))
                    return b
                mask = binding.MEAS_PROJHEIS | binding.MEAS_STAGMAGNZ
                cands = tune.launch_candidates(L, binding.MIXED)
            else:
                # four momenta per size, walker w on q = w mod 4
                qs = [(0, 0), (L // 2, L // 2), (1, 0), (L // 4, L // 2)]
                chs = [channels.Channel("long", L, q) for q in qs]
                nk = max(c.nexc for c in chs)
                chs = [c for c in chs if c.nexc == nk]

                def make(kernel="auto"):
                    b = binding.ChannelBatch(L, n, chs, binding.MIXED, seed=seed + L, kernel=kernel)
                    b.set_orbital_table(0, tab)
                    oks[id(b)] = b.init(model.random_spins(L, n, seed + 100 + L))
                    return b
                mask = binding.MEAS_PROJHEIS
                cands = tune.launch_candidates(L, binding.MIXED, channel=True)
            b, chosen, hist = tune.pick_launch(make, cands, L * L, meas=L * L, mask=mask)
            ok = oks[id(b)]
            if kd == 0:
                name = "groundstate (lanes_per_row %d, %s)" % (
                    chosen.get("lanes_per_row", 0),
                    "windows" if b.sweep_variant == binding.SWEEP_WINDOWS else "one move")
            else:
                name = "long (%d states, %s kernel)" % (nk, "register" if b.register_resident else "streamed")
            sw.add(b, L * L, L * L, mask)
            members.append(b)
            buckets.append({"L": L, "kind": name, "walkers": n, "init_ok": float(np.mean(ok)),
                            "tuning": hist, "_kd": kd})
    sw.sweep(1)
    for b in members:
        b.refresh()
        b.reset_accumulators()
        b.reset_kernel_time()
    t0 = time.perf_counter()
    graph_ms = 0.0
    for _ in range(periods):
        graph_ms += sw.sweep(k)
        for b in members:
            b.refresh()
    tm = time.perf_counter() - t0
    sweeps = periods * k
    steps = sum(bk["walkers"] * bk["L"] ** 2 for bk in buckets) * sweeps
    refresh_ms = 0.0
    for bk, b in zip(buckets, members):
        acc = b.accumulators()
        kt = b.kernel_time()
        bk["p_accept_move"] = round(float(acc[:, 8].sum() / max(acc[:, 7].sum(), 1)), 4)
        bk["refresh_ms"] = round(kt["refresh_ms"] / max(kt["refresh_launches"], 1), 3)
        refresh_ms += kt["refresh_ms"]
    # roofline of the largest bucket's kernel (the streamed long channel at the largest L):
    # one sweep on its own, HIP events on its stream; algorithmic bytes per MC step of the
    # streamed channel kernel p_ap 2 N^2 s (y = Ai c, both flavours) + a_ap 4 (N + nq) N s
    # (the Woodbury update reads and writes [Ai; Q] of both flavours), s = 8 (FP32 complex)
    roof = None
    big = [i for i, bk in enumerate(buckets) if bk["_kd"] == 1 and bk["L"] == max(Ls)]
    if big:
        i = big[0]
        b = members[i]
        L = buckets[i]["L"]
        b.reset_kernel_time()
        b.reset_accumulators()
        b.walk(L * L)
        kt = b.kernel_time()
        acc = b.accumulators()
        st = acc[:, 7].sum()
        p_ap = 1.0 - acc[:, 10].sum() / st
        a_ap = acc[:, 8].sum() / st
        N, nq = L * L // 2, L * L // 2
        bytes_launch = st * (p_ap * 2 * N * N * 8 + a_ap * 4 * (N + nq) * N * 8)
        ms = kt["sweep_ms"] / max(kt["sweep_launches"], 1)
        roof = {"kernel": "exc_kernel<float> (streamed long channel, L=%d, %d walkers)" % (L, buckets[i]["walkers"]),
                "bound": "hbm", "achieved": round(bytes_launch / (ms * 1e-3) / 1e9, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(bytes_launch / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                "avg_launch_ms": round(ms, 3), "p_antiparallel": round(float(p_ap), 4),
                "p_accept_move": round(float(a_ap), 4),
                "note": "algorithmic bytes p_ap 2N^2 s + a_ap 4(N+nq)N s per MC step over the launch"}
    sw.close()
    for b in members:
        b.close()
    for bk in buckets:
        bk.pop("_kd")
    return {"workload": "%d walkers, L in %s x {groundstate, long} from a seeded uniform categorical, one "
                        "captured hipGraph per sweep (L*L steps + 1 measurement per walker), mixed FP32; timed: "
                        "%d period(s) of %d sweeps and the FP64 refresh of every bucket" % (total, list(Ls), periods, k),
            "mc_steps_per_s": round(steps / tm, 1), "ms_per_sweep_incl_refresh": round(1e3 * tm / sweeps, 3),
            "graph_ms_per_sweep": round(graph_ms / sweeps, 3),
            "refresh_ms_all_buckets": round(refresh_ms / periods, 3),
            "sweeps": sweeps, "refreshes": periods, "buckets": buckets, "roofline_largest_bucket": roof}


# ----------------------------------------------------------------------------- main
def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2,
                    help="timed refresh periods (k sweeps + the FP64 refresh each)")
    ap.add_argument("--warmup", type=int, default=1, help="untimed refresh periods before the timed ones")
    ap.add_argument("--therm", type=int, default=200,
                    help="untimed thermalisation sweeps first (200 L^2 steps: the reference's therm)")
    ap.add_argument("--L", type=int, default=16)
    ap.add_argument("--walkers", type=int, default=16384)
    ap.add_argument("--refresh-every", type=int, default=0,
                    help="sweeps per FP64 refresh (0: autotune against --eloc-bound)")
    ap.add_argument("--eloc-bound", type=float, default=1e-7,
                    help="autotune bound on the batch mean of |dE_loc| per site between drifted and "
                         "refreshed FP32 tables")
    ap.add_argument("--fp64-steps", type=int, default=2, help="timed sweeps of the FP64 path")
    ap.add_argument("--seed", type=int, default=20261203)
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--exciton-periods", type=int, default=1, help="timed periods of the L=12 exciton leg (0: skip)")
    ap.add_argument("--mixed-periods", type=int, default=1, help="timed periods of the uneven mixed batch (0: skip)")
    ap.add_argument("--c1-measurements", type=int, default=100000, help="measurements of the L=4 leg (0: skip)")
    ap.add_argument("--c2-sweeps", type=int, default=20, help="timed sweeps of the L=8 FP64 leg (0: skip)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build measures one GPU; walkers are independent chains and the "
              "multi-GPU layout (walker shards per GPU, one FP64 reduction per save) is a "
              "design note only.  Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)
    from gpvmc_amd import binding, tune
    L, W = args.L, args.walkers
    N, Ns = L * L // 2, L * L
    G = len(PHIS) * len(NEELS)
    groups = np.arange(W) % G
    mask = binding.MEAS_PROJHEIS | binding.MEAS_STAGMAGNZ
    occs = grid_orbitals(L)

    log("mixed path L=%d W=%d therm=%d sweeps" % (L, W, args.therm))
    bm, _ = make_ground(binding.MIXED, L, W, args.seed, occs)

    def sweep():
        bm.walk(Ns, meas=Ns, mask=mask)
    for i in range(args.therm):
        sweep()
        if (i + 1) % 16 == 0:
            bm.refresh()
    bm.refresh()
    k, tuned = args.refresh_every, None
    if k <= 0:
        # untimed: longest refresh interval whose energy error per site (batch mean of |E_loc
        # on drifted tables - E_loc on refreshed tables|) stays under the bound
        k, hist = tune.tune_refresh(bm, sweep, bound=args.eloc_bound)
        tuned = {"refresh_every_sweeps": k, "criterion": "batch mean |dE_loc| per site <= %g" % args.eloc_bound,
                 "trials": [{kk: (v if kk == "k" else float("%.3g" % v)) for kk, v in h.items()} for h in hist]}
    log("refresh every %d sweeps; warmup %d period(s), timed %d period(s)" % (k, args.warmup, args.steps))

    def period():
        # one walk of k sweeps with a measurement after every L*L steps: MetroMC::Walk(k Ns, Ns),
        # one kernel launch, so the walker's ratio tables are loaded and stored once per period
        bm.walk(k * Ns, meas=Ns, mask=mask)
        bm.refresh()
    for _ in range(args.warmup):
        period()
    bm.reset_accumulators()
    bm.reset_kernel_time()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        period()
    tm = time.perf_counter() - t0
    acc = bm.accumulators()
    kt = bm.kernel_time()
    sweeps = args.steps * k
    value = W * Ns * sweeps / tm
    steps_done = acc[:, 7].sum()
    p_ap = 1.0 - acc[:, 10].sum() / steps_done
    a_ap = acc[:, 8].sum() / steps_done
    stats_mixed = per_walker_stats(acc, groups, G)

    # the E_loc measurement on its own (one ProjHeis + StagMagnZ per walker, no step)
    bm.measure()
    t1 = time.perf_counter()
    nmeas = 5
    for _ in range(nmeas):
        bm.measure()
    t_meas = (time.perf_counter() - t1) / nmeas

    # precision: k sweeps after the last refresh (the most drifted tables of a period), the
    # FP32 estimator before and after the refresh against FP64 on the same configurations
    log("FP32 vs FP64 E_loc on the drifted tables of a full period")
    for _ in range(k):
        sweep()
    m_drift = bm.measure()
    spins = bm.state()[0]
    bm.refresh()
    m_fresh = bm.measure()
    bm.close()
    b64, _ = make_ground(binding.FP64, L, W, args.seed, occs, spins=spins)
    m64 = b64.measure()

    log("plain FP64 path")
    b64.reset_accumulators()
    b64.reset_kernel_time()
    t2 = time.perf_counter()
    for _ in range(args.fp64_steps):
        b64.walk(Ns, meas=Ns, mask=mask)
    t64 = time.perf_counter() - t2
    kt64 = b64.kernel_time()
    acc64 = b64.accumulators()
    stats_fp64 = per_walker_stats(acc64, groups, G)
    b64.close()
    fp64_rate = W * Ns * args.fp64_steps / t64

    # rooflines.  Sweep kernel: the ratio tables stay in registers for a whole launch, in the
    # accumulator layout of the matrix cores at L = 16.  Per MC step the model counts 16 p_ap N
    # flops of ratios and 32 a_ap N^2 of the rank-1 table update, per E_loc sample 16 n_ap N
    # (n_ap ~ p_ap 2 Ns antiparallel bonds); the update dominates and runs as
    # v_mfma_f32_32x32x2_f32 (FP32 matrix peak 157.3 TFLOP/s).  Bytes of the same model
    # (p_ap 4 N s + a_ap 4 N^2 s per step, 2 N^2 s per sample, s = 8) are register/LDS-level.
    s = 8
    spl = W * Ns * k  # one launch = one period of k sweeps
    bytes_launch = spl * (p_ap * 4 * N * s + a_ap * 4 * N * N * s) + k * W * 2 * N * N * s
    flops_update = spl * 32 * a_ap * N * N
    flops_launch = spl * (16 * p_ap * N + 32 * a_ap * N * N) + k * W * 16 * (p_ap * 2 * Ns) * N
    avg_ms = kt["sweep_ms"] / max(kt["sweep_launches"], 1)
    ach_tf = flops_launch / (avg_ms * 1e-3) / 1e12
    traffic = None
    tf = os.path.join(ROOT, "profiles", "sweep_L16_mixed_traffic.md")
    if os.path.exists(tf) and L == 16 and W == 16384:
        for ln in open(tf):
            if ln.startswith("hbm_bytes_per_launch:"):
                traffic = float(ln.split(":")[1])
    mf = L == 16
    roofline = {"bound": "mfma" if mf else "valu", "achieved": round(ach_tf, 3), "peak": FP32_PEAK_TFLOPS,
                "unit": "TFLOP/s", "frac": round(ach_tf / FP32_PEAK_TFLOPS, 5), "traffic": traffic,
                "traffic_source": "rocprofv3 FETCH_SIZE + WRITE_SIZE of this kernel at this workload, scaled by a "
                                  "plain copy's readings (profiles/sweep_L16_mixed_traffic.md); not measured in "
                                  "this run",
                "kernel": "sweep_kernel<float,128,4,RES,MF> (ratio tables in matrix-core tiles)" if mf
                          else "sweep_kernel<float,N,LPR,RES>",
                "avg_launch_ms": round(avg_ms, 3), "sweeps_per_launch": k,
                "update_share_of_flops": round(flops_update / flops_launch, 4),
                "note": "FP32 flops of the per-step model (16 p_ap N + 32 a_ap N^2 per step, 16 n_ap N per E_loc "
                        "sample) per launch over the launch's average duration (HIP events on the context's "
                        "stream).  The rank-1 table update (update_share_of_flops of the model) runs as "
                        "v_mfma_f32_32x32x2_f32, one K = 2 product per 32 x 32 tile (a complex rank-1 update is a "
                        "real rank-2 update of each part); peak = the dense FP32 matrix rate.  Ratios, ProjHeis "
                        "and the bond cache run on the VALU" if mf else
                        "FP32 VALU arithmetic of the per-step model per launch over its average duration",
                "algorithmic_bytes": {"achieved": round(bytes_launch / (avg_ms * 1e-3) / 1e9, 1), "unit": "GB/s",
                                      "note": "register/LDS-level bytes of the per-step model; not HBM traffic "
                                              "(the tables stay in registers)"}}
    ref_ms = kt["refresh_ms"] / max(kt["refresh_launches"], 1)
    ref_flops = 2 * W * 2 * 8.0 * N ** 3  # Gauss-Jordan 8 N^3 and G = Ai Phi 8 N^3 per matrix
    refresh_roof = {"bound": "fp64", "achieved": round(ref_flops / (ref_ms * 1e-3) / 1e12, 3),
                    "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                    "frac": round(ref_flops / (ref_ms * 1e-3) / 1e12 / FP64_PEAK_TFLOPS, 4),
                    "avg_ms": round(ref_ms, 3),
                    "note": "FP64 Gauss-Jordan inverse (8 N^3 flops) + ratio tables G = Ai Phi (8 N^3) per "
                            "matrix, 2 matrices per walker, over the refresh's duration"}

    legs = {}
    if args.c1_measurements > 0:
        log("L=4 reference run")
        legs["c1_L4"] = c1_leg(args.seed, args.c1_measurements)
    if args.c2_sweeps > 0:
        log("L=8 FP64 with StatSpinStruct")
        legs["c2_L8_fp64"] = c2_leg(args.seed, args.c2_sweeps)
    if args.exciton_periods > 0:
        log("transverse-exciton channel, L=12, all q")
        legs["exciton_c4"] = exciton_leg(args.seed, args.exciton_periods, k)
    if args.mixed_periods > 0:
        log("uneven mixed batch, 20000 walkers")
        legs["mixed_batch_c5"] = mixed_batch_leg(args.seed, args.mixed_periods, k)
    cpu = None
    if not args.no_cpu_baseline:
        log("CPU baseline")
        cpu = cpu_baseline(L, args.cpu_seconds)
        rate, P = cpu_eloc_rate(L, min(3.0, args.cpu_seconds))
        cpu["eloc_samples_per_s_port"] = round(rate, 1)
        cpu["eloc_note"] = ("oracle C restatement of ProjHeis::measure, %d threads; the reference's own "
                            "ProjHeis calls the Fortran zdotu shim" % P)

    def z(a, b):
        return round((a[0] - b[0]) / max(np.hypot(a[1], b[1]), 1e-300), 2)
    line = {
        "metric": METRIC, "value": round(value, 1), "unit": "MC steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(1e3 * tm / args.steps, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "fp32", "data": "synthetic",
        "config": {"workload": "L=16 staggered-flux groundstate, N=128, %d walkers, 8-point (phi,neel) grid; "
                               "1 bench step = 1 FP64-refresh period = %d sweeps (256 steps + 1 ProjHeis/"
                               "StagMagnZ measurement each, one walk launch) + the FP64 Gauss-Jordan refresh; therm %d "
                               "sweeps untimed" % (W, k, args.therm),
                   "L": L, "walkers": W, "sweeps_per_step": k, "parallelism": "1 GPU, 1 workgroup per walker"},
        "mc_steps_per_s": round(value, 1),
        "eloc_samples_per_s": round(W * sweeps / tm, 1),
        "eloc_samples_per_s_measurement_kernel": round(W / t_meas, 1),
        "fp64_steps_per_s": round(fp64_rate, 1),
        "mixed_over_fp64": round(value / fp64_rate, 2),
        "sweep_avg_ms": round(avg_ms / k, 3),
        "walk_launch_avg_ms": round(avg_ms, 3), "sweeps_per_walk_launch": k,
        "fp64_sweep_avg_ms": round(kt64["sweep_ms"] / max(kt64["sweep_launches"], 1), 3),
        "refresh_avg_ms": round(ref_ms, 3),
        "refresh_every_sweeps": k, "refresh_tuning": tuned,
        "dE_per_site_fp32_vs_fp64": {
            "drifted": de_stats(m_drift["eloc"].real - m64["eloc"].real),
            "refreshed": de_stats(m_fresh["eloc"].real - m64["eloc"].real),
            "note": "E_loc per site of every walker at the same configuration, FP32 tables %d sweeps after "
                    "the last refresh (drifted) and right after the next one (refreshed) vs FP64" % k},
        "estimates_mixed": stats_mixed, "estimates_fp64": stats_fp64,
        "z_mixed_vs_fp64": {"energy_per_site": z(stats_mixed["energy_per_site"], stats_fp64["energy_per_site"]),
                            "stag_magn_sq": z(stats_mixed["stag_magn_sq"], stats_fp64["stag_magn_sq"])},
        "p_antiparallel": round(float(p_ap), 4), "p_accept_move": round(float(a_ap), 4),
        "roofline": roofline, "refresh_roofline": refresh_roof, "cpu_baseline": cpu,
    }
    line.update(legs)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
