"""GPU parity of channel (exciton) walks against the oracle's exact chain.

The oracle recomputes every state's amplitude by LU at every step (tests/oracle_ref.py,
ExcWalker); the HIP path keeps the inverse of the dominant state and updates it by
Woodbury.  Bars (FP64): identical no-op/accept sequences and configurations, weight
ratios within 1e-8 relative (1e-12 absolute for moves into nodes), ProjHeis H and O sums within 1e-8 of their scale, ln w
within 1e-8.  Mixed FP32: ratios within 1e-3 relative at the median while decisions
agree, H/O on identical configurations within 1e-4 of their scale after a refresh.
"""
import numpy as np
import pytest

import oracle_ref as O
from gpvmc_amd import binding, channels, model

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def rand_spins(L, nup, W, seed):
    rng = np.random.default_rng(seed)
    out = np.ones((W, L * L), np.int32)
    for w in range(W):
        out[w, rng.choice(L * L, nup, replace=False)] = 0
    return out


@pytest.mark.parametrize("name,L,q,W,sweeps", [("trans", 4, (1, 2), 6, 6), ("long", 4, (0, 0), 6, 6),
                                                ("long", 4, (1, 0), 4, 6), ("trans", 6, (2, 3), 3, 3)])
def test_fp64_channel_walk_matches_oracle(name, L, q, W, sweeps):
    say("case channel fp64", name, L, q, W)
    ch = channels.Channel(name, L, q)
    tab = model.orbital_table(L, 0.25 * np.pi, 0.1)
    seed = 500 + L + q[0]
    spins = rand_spins(L, ch.npt[0], W, seed)
    Ns = L * L
    nsteps, meas = sweeps * Ns + 1, Ns
    with binding.ChannelBatch(L, W, ch, binding.FP64, seed=seed) as b:
        b.set_orbital_table(0, tab)
        ok = b.init(spins)
        out = b.walk(nsteps, meas=meas, trace=True)
        ho, cs, wt = b.channel_sums()
        s_f = b.state()[0]
        acc = b.accumulators()
    u = model.philox_uniforms(seed, range(W), 0, nsteps)
    checked = 0
    for w in range(W):
        if not ok[w]:
            continue
        wo = O.ExcWalker(L, tab, ch, spins[w])
        o = wo.walk(u[w].reshape(-1), meas=meas)
        assert np.array_equal(o["noop"], out["noop"][w])
        assert np.array_equal(o["accept"], out["accept"][w])
        live = ~o["noop"].astype(bool)
        r_g, r_o = out["ratio"][w][live], o["ratio"][live]
        # moves into exact nodes (lattice symmetry at L=4) give rounding-level ratios
        # (1e-30 .. 1e-60) on both sides: an absolute floor of 1e-12
        assert (np.abs(r_g - r_o) <= 1e-8 * r_o + 1e-12).all()
        assert np.array_equal(s_f[w], wo.spins())
        nm = o["n_meas"]
        assert acc[w, 3] == nm
        scale = max(1.0, np.abs(o["HO"]).max())
        np.testing.assert_allclose(ho[w], o["HO"], rtol=0, atol=1e-8 * scale)
        np.testing.assert_allclose(out["logw"][w][:nm], o["logw"], rtol=0, atol=1e-8)
        checked += 1
    assert checked > 0


@pytest.mark.parametrize("name,L,q,W", [("trans", 6, (1, 2), 4), ("long", 8, (2, 2), 2)])
def test_mixed_channel_tracks_oracle(name, L, q, W):
    say("case channel mixed", name, L, q, W)
    ch = channels.Channel(name, L, q)
    tab = model.orbital_table(L, 0.1667 * np.pi, 0.1)
    seed = 700 + L
    spins = rand_spins(L, ch.npt[0], W, seed)
    Ns = L * L
    with binding.ChannelBatch(L, W, ch, binding.MIXED, seed=seed) as b:
        b.set_orbital_table(0, tab)
        ok = b.init(spins)
        out = b.walk(2 * Ns, trace=True)
        b.refresh()
        s_f = b.state()[0]
        b.reset_accumulators()
        b.walk(1, meas=1)  # one measurement after one more step
        ho, _, _ = b.channel_sums()
        s_g = b.state()[0]
    u = model.philox_uniforms(seed, range(W), 0, 2 * Ns + 1)
    rels = []
    for w in range(W):
        if not ok[w]:
            continue
        wo = O.ExcWalker(L, tab, ch, spins[w])
        o = wo.walk(u[w, : 2 * Ns].reshape(-1))
        d = np.nonzero(o["accept"] != out["accept"][w])[0]
        upto = d[0] if len(d) else 2 * Ns
        live = ~o["noop"][:upto].astype(bool)
        rels.append(np.abs(out["ratio"][w][:upto][live] - o["ratio"][:upto][live]) / o["ratio"][:upto][live])
        if len(d) == 0:
            assert np.array_equal(s_f[w], wo.spins())
            wo2 = O.ExcWalker(L, tab, ch, s_f[w])
            o2 = wo2.walk(u[w, 2 * Ns: 2 * Ns + 1].reshape(-1), meas=1)
            if np.array_equal(wo2.spins(), s_g[w]):
                scale = max(1.0, np.abs(o2["HO"]).max())
                np.testing.assert_allclose(ho[w], o2["HO"], rtol=0, atol=1e-4 * scale)
    rel = np.concatenate(rels)
    assert np.median(rel) < 1e-3
