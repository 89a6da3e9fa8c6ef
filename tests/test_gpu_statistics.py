"""Statistical acceptance at the L = 4 configuration of the reference's example run:
4 x 4 periodic lattice, N_up = N_dn = 8, staggered flux phi = 0.25 (x pi), neel = 0.1,
nearest-neighbour swap proposals, ProjHeis + StagMagnZ measured every 16 steps after
200 L^2 thermalisation steps (vmc.cpp:69-94,290-317; MetroMC.cpp:45-103).

The exact variational expectation values come from enumerating all C(16, 8) = 12870 Sz = 0
configurations with amplitudes from the second-quantised definition (tests/exact.py,
independent of the reference's label conventions):
    <E>/site = sum |psi|^2 E_loc / sum |psi|^2        (ProjHeis.cpp:40-158)
    <|m_s|^2> = sum |psi|^2 |m_s|^2 / sum |psi|^2,   m_s = (1/Ns) sum_v e^{i pi (x+y)} (2 up_v - 1)
                                                     (StagMagnZ.cpp:12-43)
The Monte Carlo means must agree within 4 standard errors:
  * one chain with 1e5 measurements, as the reference runs it, error bars from 100 bins of
    1000 consecutive measurements (FP64);
  * 1024 independent walkers with 112 measurements each (114688 in all), error bars from
    the spread of the walkers' means (FP64, and mixed FP32 with an FP64 refresh every
    256 steps).
"""
import numpy as np
import pytest

import exact
from gpvmc_amd import binding, model

pytestmark = pytest.mark.gpu

L, PHI, NEEL = 4, 0.25, 0.1
THERM = 200 * L * L
MEAS = 16


def say(*a):
    print(*a, flush=True)


def exact_values():
    tab = model.orbital_table(L, PHI * np.pi, NEEL)
    occ = model.ground_occ(tab)
    confs = exact.configurations(L)
    psi = exact.spin_basis_psi(confs, occ)
    eloc = exact.local_energy_exact(L, confs, psi, model.edges(L))
    w = np.abs(psi) ** 2
    Z = w.sum()
    x = np.arange(L * L) % L
    y = np.arange(L * L) // L
    stag = np.where((x + y) % 2 == 0, 1.0, -1.0)
    ms = ((2.0 * (confs == 0) - 1.0) * stag[None, :]).sum(axis=1) / (L * L)
    return occ, float((w * eloc.real).sum() / Z), float((w * ms ** 2).sum() / Z), float((w * ms).sum() / Z)


def test_c1_single_chain_matches_exact_enumeration():
    occ, e_ex, m2_ex, m_ex = exact_values()
    say("case c1 single chain: exact E/site %.8f  <m_s^2> %.8f" % (e_ex, m2_ex))
    n_meas, chunk = 100000, 10000
    with binding.Batch(L, 1, binding.FP64, seed=12345) as b:
        b.set_orbitals(0, occ)
        spins = model.random_spins(L, 1, 12345)
        assert b.init(spins).all()
        b.walk(THERM)
        e, m = [], []
        for _ in range(n_meas // chunk):
            out = b.walk(chunk * MEAS, meas=MEAS, trace=True)
            e.append(out["eloc"][0].real)
            m.append(out["mz"][0].real)
        acc = b.accumulators()
    e = np.concatenate(e)
    m = np.concatenate(m)
    assert len(e) == n_meas and acc[0, 3] == n_meas
    eb = e.reshape(100, -1).mean(axis=1)
    mb = (m ** 2).reshape(100, -1).mean(axis=1)
    se_e = eb.std(ddof=1) / np.sqrt(len(eb))
    se_m = mb.std(ddof=1) / np.sqrt(len(mb))
    say("  MC E/site %.6f +- %.6f (exact %.6f), <m_s^2> %.6f +- %.6f (exact %.6f)"
        % (e.mean(), se_e, e_ex, (m ** 2).mean(), se_m, m2_ex))
    assert abs(e.mean() - e_ex) < 4 * se_e
    assert abs((m ** 2).mean() - m2_ex) < 4 * se_m
    assert se_e < 5e-3 and se_m < 5e-3


@pytest.mark.parametrize("prec", [binding.FP64, binding.MIXED])
def test_c1_many_walkers_match_exact_enumeration(prec):
    occ, e_ex, m2_ex, m_ex = exact_values()
    W, seg, nseg = 1024, 256, 7
    say("case c1 walkers", "fp64" if prec == binding.FP64 else "mixed")
    with binding.Batch(L, W, prec, seed=4242) as b:
        b.set_orbitals(0, occ)
        spins = model.random_spins(L, W, 4242)
        ok = b.init(spins)
        tries = 1
        # a third of the L = 4 configurations are exact nodes of this state; the reference
        # redraws singular starts (vmc.cpp:222-241)
        while not ok.all() and tries < 40:
            spins[~ok] = model.random_spins(L, W, 4242 + tries)[~ok]
            ok = b.init(spins)
            tries += 1
        assert ok.all()
        for _ in range(THERM // seg):
            b.walk(seg)
            if prec == binding.MIXED:
                b.refresh()
        b.reset_accumulators()
        for _ in range(nseg):
            b.walk(seg, meas=MEAS)
            if prec == binding.MIXED:
                b.refresh()
        acc = b.accumulators()
    n = acc[:, 3]
    assert (n == nseg * seg // MEAS).all() and n.sum() >= 1e5
    e_w = acc[:, 0] / n
    m2_w = acc[:, 6] / n
    m_w = acc[:, 4] / n
    res = []
    for name, vals, ex in (("E/site", e_w, e_ex), ("<m_s^2>", m2_w, m2_ex), ("<m_s>", m_w, m_ex)):
        se = vals.std(ddof=1) / np.sqrt(W)
        say("  %s MC %.6f +- %.6f exact %.6f" % (name, vals.mean(), se, ex))
        res.append((vals.mean(), se, ex))
    for mean, se, ex in res:
        assert abs(mean - ex) < 4 * se
        assert se < 5e-3
