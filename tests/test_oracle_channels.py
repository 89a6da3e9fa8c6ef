"""Pins of the oracle's channel (exciton) functions (oracle/gpvmc_oracle.c: or_exc_amps,
or_exc_local, or_exc_walk):
  * amplitudes equal numpy determinants of the same matrices;
  * they equal the reference's slot arrangement of each state (state 0's orbitals in
    increasing order, then the stored hop path applied in place) times its HopSign;
  * at L=4 their ratios between states equal those of second-quantised site-basis
    amplitudes (tests/exact.py), which fixes every relative sign;
  * h_k / a_k equals the exact local energy (H psi_k)(alpha) / psi_k(alpha) for every
    state of the trans and long channels on sampled configurations;
  * or_exc_walk replays the reference's own exciton chains (fixtures excwalk_*.npz, from
    the reference compiled in oracle/_ref): same moves, weight ratios and H/O sums.
"""
import os

import numpy as np
import pytest

import exact
import oracle_ref as O
from gpvmc_amd import channels, model

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
CASES = [("trans", 4, (1, 2)), ("long", 4, (0, 0)), ("long", 4, (1, 0)), ("trans", 6, (2, 3))]


def rand_spins(L, nup, rng):
    s = np.ones(L * L, np.int32)
    s[rng.choice(L * L, nup, replace=False)] = 0
    return s


@pytest.mark.parametrize("name,L,q", CASES)
def test_exc_amps_equal_numpy_determinants(name, L, q):
    ch = channels.Channel(name, L, q)
    tab = model.orbital_table(L, 0.25 * np.pi, 0.1)
    rng = np.random.default_rng(L + len(name))
    done = 0
    while done < 4:
        w = O.ExcWalker(L, tab, ch, rand_spins(L, ch.npt[0], rng))
        la, ph = w.amps()
        if la.max() < 0:
            continue  # a node of every state (lattice symmetry): amplitudes are rounding
        done += 1
        pu, pd = w.part[: ch.npt[0]], w.part[ch.npt[0]:]
        for k, (ou, od) in enumerate(ch.states):
            d = np.linalg.det(tab[0][np.ix_(ou, pu)]) * np.linalg.det(tab[1][np.ix_(od, pd)])
            if la[k] < la.max() - 15:
                continue  # an exact zero of this state, evaluated to rounding
            assert abs(np.log(abs(d)) - la[k]) < 1e-9
            assert abs(np.angle(d * np.exp(-1j * ph[k]))) < 1e-9


@pytest.mark.parametrize("name,L,q", CASES[:3])
def test_exc_amps_follow_reference_slot_arrangement_and_hopsign(name, L, q):
    """SlaterDeterminant::VirtUpdate's amplitude of state k: m_amp * GetSign() *
    HopSign(khop) with the rows of state 0 (slots in increasing orbital order, AddState)
    and the path's orbital replacements done in place, slot by slot."""
    ch = channels.Channel(name, L, q)
    tab = model.orbital_table(L, 0.25 * np.pi, 0.1)
    rng = np.random.default_rng(3)
    while True:
        w = O.ExcWalker(L, tab, ch, rand_spins(L, ch.npt[0], rng))
        la, ph = w.amps()
        if la.max() > 0:
            break
    parts = [w.part[: ch.npt[0]], w.part[ch.npt[0]:]]
    for k in range(ch.nexc):
        d = complex(ch.hopsign[0, k])
        for fl in range(2):
            rows = list(ch.states[0][fl])
            for a, b in ch.hops[0][k][fl]:
                rows[rows.index(a)] = b
            d *= np.linalg.det(tab[fl][np.ix_(rows, parts[fl])])
        if la[k] < la.max() - 15:
            continue
        assert abs(np.log(abs(d)) - la[k]) < 1e-9
        assert abs(np.angle(d * np.exp(-1j * ph[k]))) < 1e-9


@pytest.mark.parametrize("name,q", [("trans", (1, 2)), ("long", (1, 0))])
def test_exc_amps_and_local_energy_exact_L4(name, q):
    L = 4
    ch = channels.Channel(name, L, q)
    tab = model.orbital_table(L, 0.25 * np.pi, 0.1)
    confs = exact.configurations_n(L, ch.npt[0])
    psi = np.array([exact.spin_basis_psi_orbitals(confs, tab, u, d) for u, d in ch.states])  # [k][conf]
    edges = O.edges(L)
    eex = np.array([exact.local_energy_exact(L, confs, psi[k], edges) * L * L for k in range(ch.nexc)])
    rng = np.random.default_rng(11)
    scale = np.abs(psi).max(axis=0)
    live = np.nonzero(scale > 1e-6 * scale.max())[0]  # configurations off the common nodes
    sample = rng.choice(live, 200, replace=False)
    checked = 0
    for c in sample:
        w = O.ExcWalker(L, tab, ch, confs[c])
        a, h, wn = w.local()
        big = np.abs(psi[:, c]).max()
        # same state ratios: a_k psi_j == a_j psi_k (a common factor per configuration)
        j = int(np.argmax(np.abs(a)))
        cross = a * psi[j, c] - a[j] * psi[:, c]
        assert np.abs(cross).max() <= 1e-9 * big
        for k in range(ch.nexc):
            if abs(a[k]) < 1e-6:  # a node of state k (scaled: max |a| = 1)
                continue
            assert abs(h[k] / a[k] - eex[k, c]) < 1e-8 * max(1.0, abs(eex[k, c]))
            checked += 1
    assert checked > 100


EXCWALKS = ["excwalk_trans_L4_q1_2_s12345.npz", "excwalk_long_L4_q2_2_s4242.npz",
            "excwalk_trans_L6_q1_2_s31.npz", "excwalk_long_L6_q2_0_s33.npz"]


@pytest.mark.parametrize("name", EXCWALKS)
def test_exc_walk_equals_reference_chain(name):
    """The reference's own exciton chain (its DetUpdate/InvUpdate, whose crossing terms read
    the A of the last Init, linalg.cpp:196,320,457; ProjHeis H/O, ProjHeis.cpp:98-153), run on
    its glibc stream from its start: the oracle's exact chain (every amplitude by LU) on the
    same uniforms makes the same moves, its weight ratios agree within 1e-11 relative and its
    H/O sums within 1e-11 of their scale.  Ratios of moves onto vanishing amplitudes
    (1e-28 .. 1e-32, rounding noise on both sides) have an absolute floor of 1e-20.

    The stale crossing terms cancel: the reference adds them as a block-column operation
    [B11; B21] += [B12; B22] X, X = A_cross - C_cross, on the capacitance B, and InvUpdate
    applies the same X to A^-1 C, so det(B) and the Woodbury product do not depend on X."""
    g = np.load(os.path.join(GOLD, name), allow_pickle=False)
    L, n, meas = int(g["L"]), int(g["nsteps"]), int(g["meas"])
    ch = channels.Channel(str(g["channel"]), L, tuple(int(x) for x in g["q"]))
    tab = model.orbital_table(L, float(g["phi"]) * np.pi, float(g["neel"]))
    pos = int(g["stream_pos"])
    u = O.glibc_uniforms(int(g["seed"]), pos + 4 * n)[pos:]
    w = O.ExcWalker(L, tab, ch, g["init_spins"].astype(np.int32))
    o = w.walk(u, meas=meas)
    moved = np.unpackbits(g["moved"])[:n].astype(bool)
    assert np.array_equal(moved, (o["accept"] & ~o["noop"]).astype(bool))
    r_ref = g["ratio"]
    assert (np.abs(r_ref - o["ratio"]) <= 1e-11 * o["ratio"] + 1e-20).all()
    assert o["n_meas"] == int(g["n_meas"])
    ho = g["ho"].reshape(2 * ch.nexc, ch.nexc)
    scale = np.abs(ho).max()
    assert np.abs(o["HO"] - ho).max() <= 1e-11 * scale
    assert np.array_equal(w.spins(), g["final_spins"].astype(np.int32))
