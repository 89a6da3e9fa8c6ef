"""Batches of contexts of uneven sizes and channels swept through one captured hipGraph
(gpvmc_batch_*): every member must end where the same context walked on its own ends, bit
for bit (same kernels, same inputs, replays of the captured sweep included)."""
import numpy as np
import pytest

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


def make_members():
    """(constructor, setup) pairs: ground state FP64 L=4 and mixed L=8, a trans channel on
    the register-resident kernel and a long channel on the streamed one"""
    out = []

    def ground(L, W, prec, seed):
        def mk():
            b = binding.Batch(L, W, prec, n_groups=1, seed=seed)
            b.set_orbitals(0, model.ground_occ(model.orbital_table(L, 0.25 * np.pi, 0.1)))
            ok = b.init(model.random_spins(L, W, seed))
            return b, ok
        return mk, (L * L, L * L, binding.MEAS_PROJHEIS | binding.MEAS_STAGMAGNZ)

    def chan(name, L, q, W, prec, seed, kernel):
        def mk():
            ch = channels.Channel(name, L, q)
            b = binding.ChannelBatch(L, W, ch, prec, seed=seed, kernel=kernel)
            b.set_orbital_table(0, model.orbital_table(L, 0.25 * np.pi, 0.1))
            ok = b.init(rand_spins(L, ch.npt[0], W, seed))
            return b, ok
        return mk, (L * L, L * L, binding.MEAS_PROJHEIS)

    out.append(ground(4, 8, binding.FP64, 31))
    out.append(ground(8, 16, binding.MIXED, 32))
    out.append(chan("trans", 4, (1, 2), 6, binding.FP64, 33, "auto"))
    out.append(chan("long", 6, (0, 0), 4, binding.MIXED, 34, "streamed"))
    return out


def snapshot(b):
    d = dict(spins=b.state()[0], acc=b.accumulators())
    if isinstance(b, binding.ChannelBatch):
        ho, cs, wt = b.channel_sums()
        d.update(ho=ho, cs=cs, wt=wt)
    return d


@pytest.mark.parametrize("n_sweeps", [3])
def test_batch_sweeps_equal_separate_walks(n_sweeps):
    members = make_members()
    # separate walks
    ref = []
    for mk, (n, meas, mask) in members:
        b, ok = mk()
        for _ in range(n_sweeps):
            b.walk(n, meas=meas, mask=mask)
        ref.append(snapshot(b))
        b.close()
    # the same contexts through one batch: one captured sweep, replayed
    bs = []
    with binding.Sweeper() as sw:
        for mk, (n, meas, mask) in members:
            b, ok = mk()
            sw.add(b, n, meas, mask)
            bs.append(b)
        say("case batch sweep", len(bs), "members")
        ms1 = sw.sweep(1)
        ms2 = sw.sweep(n_sweeps - 1)
        assert ms1 > 0 and ms2 > 0
        got = [snapshot(b) for b in bs]
    for b in bs:
        b.close()
    for i, (r, g) in enumerate(zip(ref, got)):
        for key in r:
            assert np.array_equal(r[key], g[key]), (i, key)
    # the walks did something
    assert all(r["acc"][:, 7].sum() > 0 for r in ref)


def test_batch_errors():
    with binding.Sweeper() as sw:
        with pytest.raises(binding.GpvmcError):
            sw.sweep(1)  # empty batch
        b = binding.Batch(4, 2, binding.FP64, n_groups=1, seed=1)
        with pytest.raises(binding.GpvmcError):
            sw.add(b, 16)  # not initialised
        b.set_orbitals(0, model.ground_occ(model.orbital_table(4, 0.25 * np.pi, 0.1)))
        b.init(model.random_spins(4, 2, 1))
        sw.add(b, 16)
        with pytest.raises(binding.GpvmcError):
            sw.add(b, 16)  # twice
        assert sw.sweep(2) > 0
        b.close()
