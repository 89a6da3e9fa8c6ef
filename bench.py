#!/usr/bin/env python3
"""Throughput of the batched Metropolis VMC hot path on one MI355X.

Workload: L=16 staggered-flux ground state, N=128 particles per flavour, 16384 walkers
on an 8-point (phi, neel) grid (phi in {0.10, 0.1667, 0.2333, 0.30} x neel in {0, 0.1},
walker w -> point w mod 8).  One bench step = one sweep: L^2 = 256 Metropolis steps on
every walker followed by one ProjHeis + StagMagnZ measurement (meas_interv = L^2, the
vmc.cpp:121 default), with an FP64 Gauss-Jordan refresh of the inverses every
`--refresh-every` sweeps in mixed precision.

Prints ONE JSON line: value = MC steps/s of the mixed-precision path (whole job, inputs
resident in HBM), plus E_loc samples/s, the plain-FP64 rate on the same workload, the
energy difference between the FP32 and FP64 estimators on identical configurations,
the roofline of the sweep kernel (HIP events on the library's stream) and a CPU
baseline (the reference's own MetroMC::Step built from its sources, on host cores).
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
FP32_PEAK_TFLOPS = 157.3


def setup(b, L, W, seed):
    from gpvmc_amd import model
    grid = [(p, n) for p in PHIS for n in NEELS]
    for gi, (p, n) in enumerate(grid):
        b.set_orbitals(gi, model.ground_occ(model.orbital_table(L, p * np.pi, n)))
    b.set_groups(np.arange(W, dtype=np.int32) % len(grid))
    spins = model.random_spins(L, W, seed)
    ok = b.init(spins)
    tries = 1
    while not ok.all() and tries < 10:  # vmc.cpp:230-241: redraw singular starts
        bad = ~ok
        spins[bad] = model.random_spins(L, W, seed + 1000 * tries)[bad]
        ok = b.init(spins)
        tries += 1
    if not ok.all():
        raise RuntimeError("could not find regular starting configurations")
    return spins


def sweep(b, L, mask):
    b.walk(L * L, meas=L * L, mask=mask)


def run_path(precision, L, W, seed, warmup, steps, refresh_every, spins=None, eloc_bound=1e-7):
    from gpvmc_amd import binding, tune
    mask = binding.MEAS_PROJHEIS | binding.MEAS_STAGMAGNZ
    b = binding.Batch(L, W, precision, n_groups=len(PHIS) * len(NEELS), seed=seed)
    if spins is None:
        spins = setup(b, L, W, seed)
    else:
        from gpvmc_amd import model
        grid = [(p, n) for p in PHIS for n in NEELS]
        for gi, (p, n) in enumerate(grid):
            b.set_orbitals(gi, model.ground_occ(model.orbital_table(L, p * np.pi, n)))
        b.set_groups(np.arange(W, dtype=np.int32) % len(grid))
        assert b.init(spins).all()
    for i in range(warmup):
        sweep(b, L, mask)
        if precision == binding.MIXED and (i + 1) % max(refresh_every, 1) == 0:
            b.refresh()
    tuned = None
    if precision == binding.MIXED and refresh_every <= 0:
        # untimed: longest refresh interval whose energy error per site (mean over the
        # batch of |E_loc on drifted tables - E_loc on refreshed tables|) stays under the bound
        refresh_every, hist = tune.tune_refresh(b, lambda: sweep(b, L, mask), bound=eloc_bound)
        tuned = {"refresh_every_sweeps": refresh_every, "eloc_bound_per_site": eloc_bound,
                 "trials": [{k: (v if k == "k" else float("%.3g" % v)) for k, v in h.items()}
                            for h in hist]}
    b.tuned = tuned
    b.refresh_every = refresh_every
    b.reset_accumulators()
    b.reset_kernel_time()
    t0 = time.perf_counter()
    for i in range(steps):
        sweep(b, L, mask)
        # every k sweeps, and always at the end so the timed work leaves FP64-exact tables
        if precision == binding.MIXED and ((i + 1) % refresh_every == 0 or i == steps - 1):
            b.refresh()
    t1 = time.perf_counter()
    return b, t1 - t0


def cpu_baseline(L, budget_s):
    """The reference's own MetroMC::Step (oracle/_ref/ref_harness, built from the
    reference sources) on one host core for a bounded sample; falls back to the oracle's
    C restatement when the reference build is absent."""
    harness = os.path.join(ROOT, "oracle", "_ref", "ref_harness")
    env = dict(os.environ, MKL_NUM_THREADS="1", OMP_NUM_THREADS="1")
    if os.path.exists(harness):
        try:
            probe = json.loads(subprocess.run([harness, "time", str(L), "0.1667", "0.1", "1", "5000"],
                                              check=True, capture_output=True, text=True, env=env,
                                              timeout=120).stdout)
            n = int(max(5000, min(2e6, probe["steps_per_s"] * budget_s)))
            r = json.loads(subprocess.run([harness, "time", str(L), "0.1667", "0.1", "1", str(n)],
                                          check=True, capture_output=True, text=True, env=env,
                                          timeout=600).stdout)
            return {"value": round(r["steps_per_s"], 1), "unit": "MC steps/s", "cores": 1,
                    "kind": "reference",
                    "sample": "reference MetroMC::Step (LatticeStepper + SlaterDeterminant, MKL "
                              "sequential) on one chain, L=%d phi=0.1667 neel=0.1, %d steps in %.1f s"
                              % (L, n, r["seconds"])}
        except Exception as e:  # noqa: BLE001
            err = getattr(e, "stderr", None) or ""
            print("reference harness unavailable: %s %s" % (e, err.strip()[-500:]), file=sys.stderr)
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_ref as O
    from gpvmc_amd import model
    occ = O.ground_occ(L, O.orbitals(L, 0.1667 * np.pi, 0.1))
    w = O.Walker(L, occ, model.random_spins(L, 1, 1)[0])
    n = 2000
    t0 = time.perf_counter()
    w.walk(model.philox_uniforms(1, [0], 0, n).reshape(-1))
    dt = time.perf_counter() - t0
    n = int(max(n, min(2e6, n / dt * budget_s)))
    u = model.philox_uniforms(1, [0], n, n).reshape(-1)
    t0 = time.perf_counter()
    w.walk(u)
    dt = time.perf_counter() - t0
    return {"value": round(n / dt, 1), "unit": "MC steps/s", "cores": 1, "kind": "port",
            "sample": "oracle C restatement, one chain at L=%d, %d steps in %.1f s" % (L, n, dt)}


def cpu_eloc_rate(L, budget_s):
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_ref as O
    from gpvmc_amd import model
    occ = O.ground_occ(L, O.orbitals(L, 0.1667 * np.pi, 0.1))
    w = O.Walker(L, occ, model.random_spins(L, 1, 2)[0])
    n, t0 = 0, time.perf_counter()
    while time.perf_counter() - t0 < budget_s:
        w.eloc()
        n += 1
    return n / (time.perf_counter() - t0)


def exciton_leg(seed, warmup, steps, W=8192, L=12, refresh_every=16, n_cmp=1024):
    """Transverse-exciton channel at L=12 (N_up = 73, N_dn = 71, 72 states), every momentum
    q on the 12 x 12 grid (walker w on q = w mod 144), mixed FP32 with an FP64 refresh every
    `refresh_every` sweeps; a sweep is L*L steps and one ProjHeis H/O measurement.  Also the
    per-amplitude error of the FP32 walk against an FP64 recomputation on the same
    configurations, and the MFMA contraction's throughput."""
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
    for _ in range(warmup):
        b.walk(Ns)
    b.refresh()
    b.reset_kernel_time()
    b.reset_accumulators()
    nref = 0
    t0 = time.time()
    for st in range(steps):
        b.walk(Ns, meas=Ns)
        if (st + 1) % refresh_every == 0 and st + 1 < steps:
            b.refresh()
            nref += 1
    tm = time.time() - t0
    kt = b.kernel_time()
    acc = b.accumulators()
    # the contraction alone: measurement launches timed on the host around a synchronised call
    t1 = time.time()
    nm = 3
    for _ in range(nm):
        b.measure()
    t_meas = (time.time() - t1) / nm
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
    value = W * Ns * steps / tm
    return {
        "workload": "L=12 transverse exciton, all 144 q (walker w -> q = w mod 144), N_up=73 N_dn=71, "
                    "72 states, %d walkers, mixed FP32, 1 sweep = %d steps + 1 ProjHeis H/O "
                    "measurement, FP64 refresh every %d sweeps" % (W, Ns, refresh_every),
        "mc_steps_per_s": round(value, 1),
        "ms_per_sweep": round(1e3 * tm / steps, 3),
        "sweeps": steps, "refreshes": nref,
        "kernel_ms_per_sweep": round(kt["sweep_ms"] / max(steps, 1), 3),  # walk and measurement launches
        "measure_ms": round(1e3 * t_meas, 3),
        "p_accept_move": round(float(acc[:, 8].sum() / acc[:, 7].sum()), 4),
        "p_noop": round(float(acc[:, 10].sum() / acc[:, 7].sum()), 4),
        "init_ok": float(np.mean(ok)),
        "amplitude_error_fp32_vs_fp64": {
            "unit": "|a_k/a_max (FP32) - a_k/a_max (FP64)|, same configurations after the timed walk",
            "walkers": Wc, "median": float(np.median(err)), "q99": float(np.quantile(err, 0.99)),
            "max": float(err.max())},
        "contraction_roofline": {"bound": "mfma", "achieved": round(flops / t_meas / 1e12, 3),
                                 "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                                 "frac": round(flops / t_meas / 1e12 / FP32_PEAK_TFLOPS, 4),
                                 "note": "v_mfma_f32_16x16x4_f32 contraction of one measurement "
                                         "(host-timed launch incl. the H/O estimator)"},
    }


def mixed_batch_leg(seed, sweeps, total=20000, Ls=(8, 10, 12, 16, 20), warmup=1):
    """Walkers of uneven lattice sizes, ground state and longitudinal-exciton channels
    interleaved: lattice side and kind drawn per walker from a seeded categorical
    (uniform over Ls x {groundstate, long}); one context per (L, kind) bucket, all swept
    together through one captured hipGraph per sweep (gpvmc_batch_*).  A sweep is L*L steps
    and one measurement per walker (ProjHeis + StagMagnZ for the ground state, ProjHeis
    H/O for the channels), mixed FP32; an FP64 refresh of every context after warm-up.
    Each bucket's launch configuration (sweep shape and variant, or register-resident vs
    streamed channel kernel) is picked by timing the candidates on the bucket's walkers
    (gpvmc_amd.tune.pick_launch)."""
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
            # the autotuner: every launch configuration of the bucket timed on its own walkers
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
                            "tuning": hist})
    sw.sweep(warmup)
    for b in members:
        b.refresh()
    for b in members:
        b.reset_accumulators()
    ms = sw.sweep(sweeps)
    steps = sum(bk["walkers"] * bk["L"] ** 2 for bk in buckets) * sweeps
    for bk, b in zip(buckets, members):
        acc = b.accumulators()
        bk["p_accept_move"] = round(float(acc[:, 8].sum() / max(acc[:, 7].sum(), 1)), 4)
    sw.close()
    for b in members:
        b.close()
    return {"workload": "%d walkers, L in %s x {groundstate, long} from a seeded uniform categorical, "
                        "one captured hipGraph per sweep (L*L steps + 1 measurement per walker), mixed FP32"
                        % (total, list(Ls)),
            "mc_steps_per_s": round(steps / (ms * 1e-3), 1), "ms_per_sweep": round(ms / sweeps, 3),
            "sweeps": sweeps, "buckets": buckets}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=32,
                    help="timed sweeps (default: two refresh periods at the usual tuned interval)")
    ap.add_argument("--warmup", type=int, default=3, help="untimed thermalisation sweeps")
    ap.add_argument("--L", type=int, default=16)
    ap.add_argument("--walkers", type=int, default=16384)
    ap.add_argument("--refresh-every", type=int, default=0,
                    help="sweeps between FP64 refreshes (0: autotune against --eloc-bound)")
    ap.add_argument("--eloc-bound", type=float, default=1e-7,
                    help="autotune bound on mean |dE_loc| per site between drifted and "
                         "refreshed FP32 tables")
    ap.add_argument("--fp64-steps", type=int, default=2, help="timed sweeps of the FP64 path")
    ap.add_argument("--seed", type=int, default=20261203)
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--exciton-steps", type=int, default=6,
                    help="timed sweeps of the L=12 transverse-exciton leg (0: skip)")
    ap.add_argument("--mixed-sweeps", type=int, default=2,
                    help="timed sweeps of the uneven mixed batch (L in 8..20, ground state and "
                         "long exciton, one hipGraph per sweep; 0: skip)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build measures one GPU; walkers are independent chains and the "
              "multi-GPU layout (walker shards per GPU, one FP64 reduction per save) is a "
              "design note only.  Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)
    from gpvmc_amd import binding
    L, W = args.L, args.walkers
    N = L * L // 2
    print("bench: mixed path L=%d W=%d warmup=%d steps=%d" % (L, W, args.warmup, args.steps),
          file=sys.stderr, flush=True)
    bm, tm = run_path(binding.MIXED, L, W, args.seed, args.warmup, args.steps, args.refresh_every,
                      eloc_bound=args.eloc_bound)
    k_used = bm.refresh_every
    acc = bm.accumulators()
    kt = bm.kernel_time()
    mc_steps = W * L * L * args.steps
    value = mc_steps / tm
    eloc_rate = W * args.steps / tm
    steps_done = acc[:, 7].sum()
    p_ap = 1.0 - acc[:, 10].sum() / steps_done
    a_ap = acc[:, 8].sum() / steps_done
    E = acc[:, 0].sum() / acc[:, 3].sum()
    ms2 = acc[:, 6].sum() / acc[:, 3].sum()

    # identical-configuration energy check: FP32 estimator of the drifted mixed state vs
    # the FP64 estimator after an FP64 LU of the same configurations
    print("bench: identical-configuration FP32 vs FP64 E_loc", file=sys.stderr, flush=True)
    m32 = bm.measure()
    spins = bm.state()[0]
    b64 = binding.Batch(L, W, binding.FP64, n_groups=len(PHIS) * len(NEELS), seed=args.seed)
    from gpvmc_amd import model
    grid = [(p, n) for p in PHIS for n in NEELS]
    for gi, (p, n) in enumerate(grid):
        b64.set_orbitals(gi, model.ground_occ(model.orbital_table(L, p * np.pi, n)))
    b64.set_groups(np.arange(W, dtype=np.int32) % len(grid))
    b64.init(spins)
    m64 = b64.measure()
    dE = np.abs(m32["eloc"].real - m64["eloc"].real)
    b64.close()
    bm.close()

    print("bench: plain FP64 path", file=sys.stderr, flush=True)
    b2, t64 = run_path(binding.FP64, L, W, args.seed, 0, args.fp64_steps, 1, spins=spins)
    kt64 = b2.kernel_time()
    fp64_rate = W * L * L * args.fp64_steps / t64
    b2.close()

    # roofline of the mixed sweep kernel: algorithmic work per launch (per MC step
    # p_ap*4*N*s bytes + a_ap*4*N^2*s bytes, 16*p_ap*N + 32*a_ap*N^2 flops; per E_loc
    # sample 2*N^2*s bytes, 16*n_ap*N flops with n_ap ~ p_ap * 2 Ns antiparallel bonds)
    s = 8
    steps_per_launch = W * L * L
    bytes_launch = steps_per_launch * (p_ap * 4 * N * s + a_ap * 4 * N * N * s) + W * 2 * N * N * s
    flops_launch = steps_per_launch * (16 * p_ap * N + 32 * a_ap * N * N) + W * 16 * (p_ap * 2 * L * L) * N
    avg_ms = kt["sweep_ms"] / max(kt["sweep_launches"], 1)
    ach_gbs = bytes_launch / (avg_ms * 1e-3) / 1e9
    ach_tf = flops_launch / (avg_ms * 1e-3) / 1e12
    traffic = None
    # HBM bytes per launch measured with rocprofv3 FETCH_SIZE / WRITE_SIZE (scaled by a
    # plain copy's readings); recorded in the profile summary at the default workload
    tf = os.path.join(ROOT, "profiles", "sweep_L16_mixed_traffic.md")
    if os.path.exists(tf) and L == 16 and W == 16384:
        for ln in open(tf):
            if ln.startswith("hbm_bytes_per_launch:"):
                traffic = float(ln.split(":")[1])
    # The ratio tables are register-resident for the whole launch, so HBM does not bound
    # the sweep: its roofline is FP32 arithmetic (157.3 TFLOP/s, the same peak for VALU
    # and MFMA FP32 on MI355X).  The byte model is kept beside it, with the measured HBM
    # traffic per launch (rocprofv3 FETCH_SIZE + WRITE_SIZE) when a profile has recorded it.
    roofline = {"bound": "mfma", "achieved": round(ach_tf, 3), "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": round(ach_tf / FP32_PEAK_TFLOPS, 5), "traffic": traffic,
                "kernel": "sweep_kernel<float,128,4,RES>",
                "avg_launch_ms": round(avg_ms, 3),
                "note": "FP32 VALU arithmetic of the survey's per-step model (16 p_ap N + 32 a_ap "
                        "N^2 flops per step, 16 n_ap N per E_loc sample) per launch over the launch's "
                        "average duration (HIP events on the context's stream)",
                "algorithmic_bytes": {"achieved": round(ach_gbs, 1), "unit": "GB/s",
                                      "note": "register/LDS-level bytes of the per-step model; "
                                              "not HBM traffic (the tables stay in VGPRs)"}}
    exc = None
    if args.exciton_steps > 0:
        print("bench: transverse-exciton channel, L=12, all q", file=sys.stderr, flush=True)
        exc = exciton_leg(args.seed, 1, args.exciton_steps)
    mixed = None
    if args.mixed_sweeps > 0:
        print("bench: uneven mixed batch, 20000 walkers", file=sys.stderr, flush=True)
        mixed = mixed_batch_leg(args.seed, args.mixed_sweeps)
    cpu = None
    if not args.no_cpu_baseline:
        print("bench: CPU baseline", file=sys.stderr, flush=True)
        cpu = cpu_baseline(L, args.cpu_seconds)
        cpu["eloc_samples_per_s_port"] = round(cpu_eloc_rate(L, min(3.0, args.cpu_seconds)), 1)
    line = {
        "metric": METRIC, "value": round(value, 1), "unit": "MC steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(1e3 * tm / args.steps, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "fp32", "data": "synthetic",
        "config": {"workload": "L=16 staggered-flux groundstate, N=128, %d walkers, 8-point "
                               "(phi,neel) grid, 1 sweep = %d steps + 1 ProjHeis/StagMagnZ "
                               "measurement, FP64 LU refresh every %d sweep(s); therm shortened "
                               "to %d sweeps" % (W, L * L, k_used, args.warmup),
                   "L": L, "walkers": W, "parallelism": "1 GPU, 1 workgroup per walker"},
        "eloc_samples_per_s": round(eloc_rate, 1),
        "fp64_steps_per_s": round(fp64_rate, 1),
        "mixed_over_fp64": round(value / fp64_rate, 2),
        "fp64_sweep_avg_ms": round(kt64["sweep_ms"] / max(kt64["sweep_launches"], 1), 3),
        "refresh_avg_ms": round(kt["refresh_ms"] / max(kt["refresh_launches"], 1), 3),
        "refresh_every_sweeps": k_used, "refresh_tuning": bm.tuned,
        "dE_per_site_fp32_vs_fp64": {"max": float(dE.max()), "mean": float(dE.mean())},
        "energy_per_site": round(float(E), 6), "stag_magn_sq": round(float(ms2), 6),
        "p_antiparallel": round(float(p_ap), 4), "p_accept_move": round(float(a_ap), 4),
        "roofline": roofline, "cpu_baseline": cpu,
        "exciton_c4": exc,
        "mixed_batch_c5": mixed,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
