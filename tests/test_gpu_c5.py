"""The uneven mixed batch at its own sizes: every bucket kind of the mixed batch (lattice
side L in {8, 10, 12, 16, 20} x {ground state, longitudinal exciton}, mixed FP32) swept
through one captured hipGraph, checked three ways:

  * each bucket ends bit for bit where the same context walked on its own ends;
  * the separate walks' per-step traces against the oracle on the same Philox stream:
    ground-state buckets against O.Walker (LatticeStepper.cpp:36-147, MetroMC.cpp:90-103,
    ProjHeis.cpp:40-158), long L = 8/10/12 against O.ExcWalker over the whole sweep
    (SlaterDeterminant.cpp:96-326; every amplitude recomputed by LU at every step) with
    the ProjHeis H/O sums of the sweep's measurement (ProjHeis.cpp:140-153);
  * long L = 16/20 (streamed kernel, sizes where an oracle chain is too slow): every
    state's amplitude after the sweep against the oracle's LU determinants at the final
    configuration (or_exc_amps).

Also the L = 12 transverse-exciton channel (the C4 size, register-resident kernel with the
MFMA H/O contraction) after one measured step against the oracle's H/O (or_exc_local).

Every bucket first walks one unmeasured warm-up sweep and takes an FP64 refresh, as the
mixed batch does before it is timed; the checks cover the sweep after that.
Mixed-precision bars (FP32 tables from that FP64 LU): decisions equal to the oracle's until
the first near-tie (|min(r, 1) - u| < 1e-3), weight ratios within 1e-4 relative at the
median and 1e-2 at the 99th percentile before it; E_loc of the ground state within 1e-5;
H/O within 1e-4 of their scale; amplitudes a_k / a_max within 1e-2 at the worst state and
2e-3 at the median (FP32, a sweep of Woodbury updates after the refresh) or 1e-8 (FP64),
after aligning the one sign per walker that the slot order of the current state fixes.
The oracle runs on host threads (ctypes releases the GIL), one walker per thread.
"""
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import oracle_ref as O
from gpvmc_amd import binding, channels, model

pytestmark = pytest.mark.gpu

LS = (8, 10, 12, 16, 20)
PHI, NEEL = 0.25, 0.1
POOL = 8


def say(*a):
    print(*a, flush=True)


def rand_spins(L, nup, W, seed):
    rng = np.random.default_rng(seed)
    out = np.ones((W, L * L), np.int32)
    for w in range(W):
        out[w, rng.choice(L * L, nup, replace=False)] = 0
    return out


def pmap(fn, items):
    with ThreadPoolExecutor(max_workers=POOL) as ex:
        return list(ex.map(fn, items))


class Bucket:
    """One (L, kind) bucket: its constructor, walkers and starting configurations."""

    def __init__(self, L, kind, W, seed, prec=binding.MIXED, kernel="auto"):
        self.L, self.kind, self.W, self.seed, self.prec, self.kernel = L, kind, W, seed, prec, kernel
        self.Ns = L * L
        self.tab = model.orbital_table(L, PHI * np.pi, NEEL)
        if kind == "ground":
            self.occ = model.ground_occ(self.tab)
            self.spins = model.random_spins(L, W, seed)
            self.mask = binding.MEAS_PROJHEIS | binding.MEAS_STAGMAGNZ
        else:
            self.ch = channels.Channel("long", L, (1, 0))
            self.spins = rand_spins(L, self.ch.npt[0], W, seed)
            self.mask = binding.MEAS_PROJHEIS

    def make(self, therm=True):
        """The bucket's context after init and, with `therm`, one unmeasured sweep and an
        FP64 refresh (the mixed batch's warm-up: random starts sit next to nodes, where the
        FP32 inverse of the first sweep carries the conditioning of the start)."""
        if self.kind == "ground":
            b = binding.Batch(self.L, self.W, self.prec, seed=self.seed)
            b.set_orbitals(0, self.occ)
        else:
            b = binding.ChannelBatch(self.L, self.W, self.ch, self.prec, seed=self.seed, kernel=self.kernel)
            b.set_orbital_table(0, self.tab)
        ok = b.init(self.spins)
        if therm:
            b.walk(self.Ns)
            b.refresh()
            b.reset_accumulators()
        return b, ok

    def name(self):
        return "%s L=%d %s" % (self.kind, self.L, "fp64" if self.prec == binding.FP64 else "mixed")


def snapshot(b):
    sp, part, la, ph = b.state()
    d = dict(spins=sp, part=part, logabs=la, phase=ph, acc=b.accumulators())
    if isinstance(b, binding.ChannelBatch):
        ho, cs, wt = b.channel_sums()
        d.update(ho=ho, cs=cs, wt=wt)
    return d


def buckets():
    out = []
    for i, L in enumerate(LS):
        out.append(Bucket(L, "ground", 3, 5000 + L))
        out.append(Bucket(L, "long", 2, 6000 + L))
    return out


@pytest.fixture(scope="module")
def c5_runs():
    """Every bucket swept once through the captured graph and once on its own with
    traces; the amplitudes of the long buckets at the end of the separate walk."""
    bks = buckets()
    graph = []
    with binding.Sweeper() as sw:
        made = []
        for bk in bks:
            b, ok = bk.make()
            sw.add(b, bk.Ns, bk.Ns, bk.mask)
            made.append(b)
        say("case c5 graph sweep: %d buckets" % len(made))
        ms = sw.sweep(1)
        assert ms > 0
        graph = [snapshot(b) for b in made]
    for b in made:
        b.close()
    sep = []
    for bk in bks:
        say("case c5 separate walk", bk.name())
        b, ok = bk.make()
        start = snapshot(b)
        out = b.walk(bk.Ns, meas=bk.Ns, mask=bk.mask, trace=True)
        snap = snapshot(b)
        amp = b.amplitudes() if bk.kind == "long" else None
        b.close()
        sep.append(dict(ok=ok, out=out, snap=snap, amp=amp, start=start))
    return bks, graph, sep


def stream(bk):
    """the uniforms of the measured sweep: Philox steps Ns .. 2 Ns - 1 (after the warm-up)"""
    return model.philox_uniforms(bk.seed, range(bk.W), bk.Ns, bk.Ns)


def test_c5_graph_equals_separate_walks(c5_runs):
    bks, graph, sep = c5_runs
    for bk, g, s in zip(bks, graph, sep):
        for key in g:
            assert np.array_equal(g[key], s["snap"][key]), (bk.name(), key)
        assert s["snap"]["acc"][:, 7].sum() == bk.W * bk.Ns
        assert s["ok"].any(), bk.name()


def tracks(o, g_acc, g_ratio, u, nsteps):
    """Steps up to the first decision that differs (which must be a near-tie) and the
    relative weight-ratio errors of the live steps before it."""
    d = np.nonzero(o["accept"] != g_acc)[0]
    upto = d[0] if len(d) else nsteps
    if len(d):
        s = d[0]
        assert abs(min(o["ratio"][s], 1.0) - u[s, 3]) < 1e-3, "decision flipped away from a near-tie"
    live = ~o["noop"][:upto].astype(bool)
    r_o = o["ratio"][:upto][live]
    rel = np.abs(g_ratio[:upto][live] - r_o) / np.maximum(r_o, 1e-300)
    return rel, len(d) == 0


@pytest.mark.parametrize("L", LS)
def test_c5_ground_bucket_tracks_oracle(c5_runs, L):
    bks, _, sep = c5_runs
    i = [j for j, bk in enumerate(bks) if bk.L == L and bk.kind == "ground"][0]
    bk, s = bks[i], sep[i]
    u = stream(bk)
    live_w = [w for w in range(bk.W) if s["ok"][w]]

    def one(w):
        wo = O.Walker(L, bk.occ, s["start"]["spins"][w])
        return wo, wo.walk(u[w].reshape(-1), meas=bk.Ns)

    say("case c5 ground oracle L=%d" % L)
    res = pmap(one, live_w)
    rels, n_same = [], 0
    for w, (wo, o) in zip(live_w, res):
        rel, same = tracks(o, s["out"]["accept"][w], s["out"]["ratio"][w], u[w], bk.Ns)
        rels.append(rel)
        # the measurement after step 0 (MetroMC::Walk's s % meas == 0)
        if o["accept"][0] == s["out"]["accept"][w][0]:
            e_o = o["eloc"][0]
            assert abs(s["out"]["eloc"][w][0] - e_o) < 1e-5 * max(1.0, abs(e_o))
            assert abs(s["out"]["mz"][w][0] - o["mz"][0]) < 1e-12
        if same:
            assert np.array_equal(s["snap"]["spins"][w], wo.spins())
            n_same += 1
    rel = np.concatenate(rels)
    say("  ratio rel err median %.2e q99 %.2e over %d steps" % (np.median(rel), np.quantile(rel, 0.99), len(rel)))
    assert len(rel) > bk.Ns // 2
    assert np.median(rel) <= 1e-4 and np.quantile(rel, 0.99) <= 1e-2
    assert n_same >= 1


@pytest.mark.parametrize("L", [8, 10, 12])
def test_c5_long_bucket_tracks_oracle(c5_runs, L):
    bks, _, sep = c5_runs
    i = [j for j, bk in enumerate(bks) if bk.L == L and bk.kind == "long"][0]
    bk, s = bks[i], sep[i]
    u = stream(bk)
    live_w = [w for w in range(bk.W) if s["ok"][w]]

    def one(w):
        wo = O.ExcWalker(L, bk.tab, bk.ch, s["start"]["spins"][w])
        return wo, wo.walk(u[w].reshape(-1), meas=bk.Ns)

    say("case c5 long oracle L=%d (%d states)" % (L, bk.ch.nexc))
    res = pmap(one, live_w)
    rels, n_ho = [], 0
    for w, (wo, o) in zip(live_w, res):
        rel, same = tracks(o, s["out"]["accept"][w], s["out"]["ratio"][w], u[w], bk.Ns)
        rels.append(rel)
        if o["accept"][0] == s["out"]["accept"][w][0]:
            # H and O of the measurement after step 0
            ho = s["snap"]["ho"][w]
            scale = max(1.0, np.abs(o["HO"]).max())
            np.testing.assert_allclose(ho, o["HO"], rtol=0, atol=1e-4 * scale)
            assert abs(s["out"]["logw"][w][0] - o["logw"][0]) < 1e-3
            n_ho += 1
        if same:
            assert np.array_equal(s["snap"]["spins"][w], wo.spins())
    rel = np.concatenate(rels)
    say("  ratio rel err median %.2e q99 %.2e over %d steps" % (np.median(rel), np.quantile(rel, 0.99), len(rel)))
    assert len(rel) > bk.Ns // 4
    assert np.median(rel) <= 1e-4 and np.quantile(rel, 0.99) <= 1e-2
    assert n_ho >= 1


def amps_match(amp_g, la_o, ph_o, tol):
    """|a_k/a_max| on both sides after aligning the per-walker sign (the slot order of the
    current state relative to increasing orbital order)."""
    g = np.exp(amp_g.real - amp_g.real.max() + 1j * amp_g.imag)
    o = np.exp(la_o - la_o.max() + 1j * ph_o)
    k = int(np.argmax(np.abs(o)))
    sgn = g[k] / o[k]
    assert abs(abs(sgn) - 1.0) < 10 * tol and abs(abs(sgn.real) - 1.0) < 10 * tol, sgn
    err = np.abs(g - np.sign(sgn.real) * o)
    return float(err.max()), float(np.median(err))


@pytest.mark.parametrize("L", [16, 20])
def test_c5_long_bucket_amplitudes_match_oracle(c5_runs, L):
    bks, _, sep = c5_runs
    i = [j for j, bk in enumerate(bks) if bk.L == L and bk.kind == "long"][0]
    bk, s = bks[i], sep[i]
    live_w = [w for w in range(bk.W) if s["ok"][w]]
    part = s["snap"]["part"]
    nup, ndn = bk.ch.npt

    def one(w):
        wo = O.ExcWalker(L, bk.tab, bk.ch, s["snap"]["spins"][w])
        p = np.concatenate([part[w][0][:nup], part[w][1][:ndn]]).astype(np.int32)
        return wo.amps(p)

    say("case c5 long amplitudes L=%d (%d states)" % (L, bk.ch.nexc))
    res = pmap(one, live_w)
    for w, (la_o, ph_o) in zip(live_w, res):
        err, med = amps_match(s["amp"][w], la_o, ph_o, 1e-2)
        say("  walker %d |a_k/a_max| error max %.2e median %.2e" % (w, err, med))
        # FP32 Woodbury updates over a sweep of L^2 steps (N = 128 / 200): measured up to
        # 6e-4 / 1.8e-3 on the worst-conditioned walker (median over its states 1e-4 /
        # 4e-4), 4e-6 / 5e-5 on the other
        assert err < 1e-2 and med < 2e-3
    assert len(live_w) >= 1


@pytest.mark.parametrize("L", [16, 20])
def test_long_streamed_fp64_amplitudes_after_sweep(L):
    """The streamed kernel in FP64 at the large long-channel sizes: after one sweep of
    Woodbury updates every amplitude equals the oracle's LU determinant to 1e-8."""
    bk = Bucket(L, "long", 2, 7000 + L, prec=binding.FP64)
    say("case long fp64 amplitudes L=%d" % L)
    b, ok = bk.make()
    assert not b.register_resident
    b.walk(bk.Ns)
    sp, part, _, _ = b.state()
    amp = b.amplitudes()
    acc = b.accumulators()
    b.close()
    assert acc[:, 8].sum() > 0
    nup, ndn = bk.ch.npt

    def one(w):
        wo = O.ExcWalker(L, bk.tab, bk.ch, sp[w])
        p = np.concatenate([part[w][0][:nup], part[w][1][:ndn]]).astype(np.int32)
        return wo.amps(p)

    live_w = [w for w in range(bk.W) if ok[w]]
    for w, (la_o, ph_o) in zip(live_w, pmap(one, live_w)):
        err, med = amps_match(amp[w], la_o, ph_o, 1e-8)
        say("  walker %d |a_k/a_max| error max %.2e median %.2e" % (w, err, med))
        assert err < 1e-8
    assert len(live_w) >= 1


def test_l12_trans_contraction_matches_oracle():
    """C4's size with its measurement: the L = 12 transverse-exciton channel (N_up = 73,
    N_dn = 71, 72 states) on the register-resident kernel, whose ProjHeis H/O come from the
    MFMA contraction; one step and one measurement after init against the oracle's H/O
    (every swapped amplitude by LU, ProjHeis.cpp:98-153)."""
    L, W, seed = 12, 2, 1313
    chs = [channels.Channel("trans", L, q) for q in ((3, 5), (0, 0))]
    tab = model.orbital_table(L, 0.25 * np.pi, 0.1)
    spins = rand_spins(L, chs[0].npt[0], W, seed)
    say("case l12 trans contraction")
    with binding.ChannelBatch(L, W, chs, binding.MIXED, seed=seed) as b:
        assert b.register_resident
        b.set_orbital_table(0, tab)
        ok = b.init(spins)
        out = b.walk(1, meas=1, trace=True)
        ho, cs, _ = b.channel_sums()
        acc = b.accumulators()
    u = model.philox_uniforms(seed, range(W), 0, 1)
    live_w = [w for w in range(W) if ok[w]]

    def one(w):
        wo = O.ExcWalker(L, tab, chs[w % len(chs)], spins[w])
        return wo.walk(u[w].reshape(-1), meas=1)

    n = 0
    for w, o in zip(live_w, pmap(one, live_w)):
        assert acc[w, 3] == 1
        if o["accept"][0] != out["accept"][w][0]:
            continue
        scale = max(1.0, np.abs(o["HO"]).max())
        err = np.abs(ho[w] - o["HO"]).max() / scale
        say("  walker %d H/O max error / scale %.2e" % (w, err))
        assert err <= 1e-4
        n += 1
    assert n >= 1
