"""Pins the C oracle (oracle/gpvmc_oracle.c) before anything is checked against it.

Pins, per oracle function:
  * fixtures written by the reference's own compiled classes (tests/golden, made by
    tests/golden/make_ref_fixtures.py from oracle/_ref/ref_harness);
  * brute-force numpy determinants/inverses on random inputs;
  * exact enumeration at L=4 with spin-basis amplitudes built from the second-quantised
    definition (tests/exact.py), independent of the reference's sign conventions;
  * closed forms (orbital unitarity, Neel magnetisation, sum rules).
"""
import os

import numpy as np
import pytest

import exact
import oracle_ref as O

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    return np.load(os.path.join(GOLD, name), allow_pickle=False)


def ratio_close(a, b, rtol=1e-9):
    # The reference's own weight ratio carries ~1e-10 relative rounding: the predicted
    # weight uses q from zgemv (DetUpdate, linalg.cpp:146) while the state weight it is
    # divided by was built from q recomputed by zgemm (InvUpdate, linalg.cpp:434), both
    # through base-10 BigDouble rescaling.  Moves into exact-zero amplitudes give ratios
    # ~1e-60 of pure rounding noise, hence the absolute floor.
    return np.abs(a - b) <= rtol * np.abs(b) + 1e-40


# ---------------------------------------------------------------- lattice (a2)
@pytest.mark.parametrize("L", [4, 6, 8])
def test_lattice_matches_reference(L):
    g = load("lattice_L%d.npz" % L)
    assert np.array_equal(O.neighbours(L), g["neighbours"])
    assert np.array_equal(O.edges(L), g["edges"])


# ---------------------------------------------------------------- orbitals (a1)
@pytest.mark.parametrize("tag", ["L4_a", "L4_b", "L4_c", "L6_a", "L8_a"])
def test_orbitals_match_reference(tag):
    g = load("orbitals_%s.npz" % tag)
    L = int(g["L"])
    bc = g["bc"]
    assert np.array_equal(O.mbz(L, bc[0], bc[1]), g["fock2qn"])
    t = O.orbitals(L, float(g["phi"]) * np.pi, float(g["neel"]), float(g["neel_exp"]), bc[0], bc[1])
    if "table" in g.files:
        np.testing.assert_allclose(t, g["table"], rtol=0, atol=1e-13)
    else:
        np.testing.assert_allclose(t.reshape(-1)[g["index"]], g["values"], rtol=0, atol=1e-13)


@pytest.mark.parametrize("L,phi,neel,ne", [(4, 0.25, 0.1, 0.0), (8, 0.1, 0.3, 0.0),
                                           (6, 0.3, 0.05, 1.0), (12, 0.1667, 0.0, 0.0)])
def test_orbitals_unitary(L, phi, neel, ne):
    # |U|^2+|V|^2 = 1, bands orthogonal, distinct momenta orthogonal per sublattice:
    # Phi_fl Phi_fl^H = Ns I for each flavour.
    t = O.orbitals(L, phi * np.pi, neel, ne)
    Ns = L * L
    for fl in range(2):
        g = t[fl] @ t[fl].conj().T
        np.testing.assert_allclose(g, Ns * np.eye(Ns), atol=1e-9)


def test_orbitals_neel_limit():
    # Large Neel field: band-0 orbitals of flavour 0 live on the even sublattice only.
    L = 4
    t = O.orbitals(L, 0.25 * np.pi, 1e6, 0.0)
    sites = np.arange(L * L)
    odd = ((sites % L) + (sites // L)) % 2 == 1
    occ = t[0, 0::2]
    assert np.abs(occ[:, odd]).max() < 1e-5
    assert np.abs(t[1, 0::2][:, ~odd]).max() < 1e-5


# ---------------------------------------------------------------- linear algebra (a4-a6)
def rand_c(rng, n, m=None):
    m = n if m is None else m
    return rng.standard_normal((n, m)) + 1j * rng.standard_normal((n, m))


@pytest.mark.parametrize("N", [1, 2, 5, 8, 17, 32])
def test_detinv_vs_numpy(N):
    rng = np.random.default_rng(N)
    Acm = rand_c(rng, N)          # A(f, r) stored as Acm[r, f] (reference layout)
    A = Acm.T                     # the mathematical matrix
    Ai, la, ph, sing = O.detinv(Acm)
    assert not sing
    d = np.linalg.det(A)
    assert abs(la - np.log(abs(d))) < 1e-10
    assert abs(np.angle(np.exp(1j * (ph - np.angle(d))))) < 1e-10
    # reference layout of the inverse: Ai[r + f*N] = inv(A)(r, f), i.e. Ai.T in numpy
    np.testing.assert_allclose(Ai.T, np.linalg.inv(A), atol=1e-9)


def test_detinv_singular_threshold():
    # |det|^2 < 1e-9 is singular (linalg.cpp:23)
    A = np.diag([1e-5, 1.0, 1.0]).astype(np.complex128)
    assert O.detinv(A)[3]
    A = np.diag([1e-4, 1.0, 1.0]).astype(np.complex128)
    assert not O.detinv(A)[3]


@pytest.mark.parametrize("N", [3, 8, 16])
def test_ratio_and_sherman_morrison_bruteforce(N):
    rng = np.random.default_rng(100 + N)
    A = rand_c(rng, N)            # mathematical A(f, r)
    Ai_ref = np.ascontiguousarray(np.linalg.inv(A).T)  # reference memory layout
    for ci in (0, N // 2, N - 1):
        col = rng.standard_normal(N) + 1j * rng.standard_normal(N)
        A2 = A.copy()
        A2[:, ci] = col
        q = O.ratio(Ai_ref, ci, col)
        assert abs(q - np.linalg.det(A2) / np.linalg.det(A)) < 1e-9 * abs(q)
        Ai2 = O.sm_update(Ai_ref, ci, col)
        np.testing.assert_allclose(Ai2.T, np.linalg.inv(A2), atol=1e-8)


# ---------------------------------------------------------------- RNG stream and RanInit (a9, a2)
@pytest.mark.parametrize("name", ["walk_L4_s12345.npz", "walk_L6_s4242.npz", "walk_L8_s777.npz"])
def test_glibc_stream_and_raninit(name):
    g = load(name)
    seed = int(g["seed"])
    u = O.glibc_uniforms(seed, 8)
    assert np.array_equal(np.round(u * 2147483647).astype(np.int64), g["glibc_head"])
    L = int(g["L"])
    u = O.glibc_uniforms(seed, 20000)
    spins, used = O.raninit(L, u)
    assert np.array_equal(spins, g["init_attempts"][0])
    assert used > 0


# ---------------------------------------------------------------- the walk (a3, a4, a5, a6, a8)
@pytest.mark.parametrize("name", ["walk_L4_s12345.npz", "walk_L6_s4242.npz", "walk_L8_s777.npz"])
def test_walk_matches_reference(name):
    g = load(name)
    L, N = int(g["L"]), int(g["L"]) ** 2 // 2
    seed, nsteps = int(g["seed"]), int(g["nsteps"])
    u = O.glibc_uniforms(seed, 4 * nsteps + 20000)
    assert len(g["init_attempts"]) == 1  # the fixture seeds start regular
    spins, used = O.raninit(L, u)
    occ = O.ground_occ(L, O.orbitals(L, float(g["phi"]) * np.pi, float(g["neel"]), float(g["neel_exp"])))
    w = O.Walker(L, occ, spins)
    assert w.ok
    assert np.array_equal(w.part.reshape(2, N), g["init_labels"])
    assert abs(w.logabs - float(g["init_logabs"])) < 1e-10
    np.testing.assert_allclose(w.Ai.reshape(2, N * N), g["init_inv"], atol=1e-12)
    out = w.walk(u[used:used + 4 * nsteps])
    moved = np.unpackbits(g["moved"])[:nsteps].astype(bool)
    mine = out["accept"].astype(bool) & ~out["noop"].astype(bool)
    assert np.array_equal(mine, moved)
    k = len(g["ratio"])
    live = ~out["noop"][:k].astype(bool)
    assert ratio_close(out["ratio"][:k][live], g["ratio"][live]).all()
    np.testing.assert_allclose(w.Ai.reshape(2, N * N), g["final_inv"], atol=1e-9)
    assert np.array_equal(w.part.reshape(2, N), g["final_labels"])
    # log10 of the final weight |amp|^2 from BigDouble vs the oracle's running log|amp|
    assert abs(2 * w.logabs / np.log(10) - g["snap_log10w"][-1]) < 1e-8
    assert np.array_equal(w.spins(), g["snaps"][-1])


def test_walk_snapshots_match_reference():
    g = load("walk_L8_s777.npz")
    L, nsteps, every = int(g["L"]), int(g["nsteps"]), int(g["snap_every"])
    u = O.glibc_uniforms(int(g["seed"]), 4 * nsteps + 20000)
    spins, used = O.raninit(L, u)
    occ = O.ground_occ(L, O.orbitals(L, float(g["phi"]) * np.pi, float(g["neel"])))
    w = O.Walker(L, occ, spins)
    for t in range(nsteps // every):
        off = used + 4 * every * t
        w.walk(u[off:off + 4 * every])
        assert np.array_equal(w.spins(), g["snaps"][t])
        assert abs(2 * w.logabs / np.log(10) - g["snap_log10w"][t]) < 1e-8
        # phase of the BigComplex amplitude (SlaterDeterminant m_amp)
        ph = np.angle(complex(g["snap_amp"][t][0], g["snap_amp"][t][1]))
        assert abs(np.angle(np.exp(1j * (w.phase - ph)))) < 1e-8


# ---------------------------------------------------------------- exact enumeration (a10-a12)
@pytest.fixture(scope="module")
def enum4():
    L = 4
    occ = O.ground_occ(L, O.orbitals(L, 0.25 * np.pi, 0.1))
    confs = exact.configurations(L)
    psi = exact.spin_basis_psi(confs, occ)
    return L, occ, confs, psi


def test_eloc_equals_exact_local_energy(enum4):
    L, occ, confs, psi = enum4
    eex = exact.local_energy_exact(L, confs, psi, O.edges(L))
    live = np.abs(psi) > 1e-8 * np.abs(psi).max()
    checked = 0
    for c in np.where(live)[0][::7]:
        w = O.Walker(L, occ, confs[c])
        if not w.ok:
            continue
        assert abs(w.eloc() - eex[c]) < 1e-9
        checked += 1
    assert checked > 500


def test_variational_energy_bounds(enum4):
    L, occ, confs, psi = enum4
    eex = exact.local_energy_exact(L, confs, psi, O.edges(L))
    w = np.abs(psi) ** 2
    E = (w * eex).sum().real / w.sum()
    # exact 4x4 PBC Heisenberg ground state energy is -0.7017802 per site
    assert -0.7017802 <= E < -0.6


def test_stagmagnz_closed_form():
    L = 4
    occ = O.ground_occ(L, O.orbitals(L, 0.25 * np.pi, 0.1))
    sites = np.arange(L * L)
    neel = (((sites % L) + (sites // L)) % 2).astype(np.int32)  # up on even sublattice
    w = O.Walker(L, occ, neel)
    assert abs(w.stagmagnz() - 1.0) < 1e-12
    w = O.Walker(L, occ, 1 - neel)
    assert abs(w.stagmagnz() + 1.0) < 1e-12
    rng = np.random.default_rng(3)
    for _ in range(5):
        s = rng.permutation(neel)
        m = np.sum(np.where(neel == 0, 1, -1) * np.where(s == 0, 1, -1)) / (L * L)
        w = O.Walker(L, occ, s)
        assert abs(w.stagmagnz() - m) < 1e-12


def test_statspinstruct_exact(enum4):
    L, occ, confs, psi = enum4
    Ns = L * L
    szz, spm, smp = exact.flip_flop_expectations(confs, psi)
    wgt = np.abs(psi) ** 2
    acc = np.zeros((5, Ns, Ns), np.complex128)
    Z = 0.0
    for c in np.where(wgt > 1e-14 * wgt.max())[0]:
        w = O.Walker(L, occ, confs[c])
        one = w.statspinstruct()
        acc += wgt[c] * one
        Z += wgt[c]
    acc /= Z
    np.testing.assert_allclose(acc[0], szz, atol=1e-9)
    np.testing.assert_allclose(acc[1], spm, atol=1e-9)
    np.testing.assert_allclose(acc[2], smp, atol=1e-9)
    assert np.abs(acc[3]).max() == 0 and np.abs(acc[4]).max() == 0
    # sum rules at Sz=0: sum_b <Sz_a Sz_b> = 0, diagonal 1/4
    np.testing.assert_allclose(acc[0].sum(axis=1), 0, atol=1e-9)
    np.testing.assert_allclose(np.diag(acc[0]), 0.25, atol=1e-12)


@pytest.mark.parametrize("name", ["eloc_L4_s12345.npz", "eloc_L6_s4242.npz"])
def test_eloc_equals_reference_projheis(name):
    """ProjHeis::measure of the reference (ProjHeis.cpp:40-158, its zdotu ratio path) at the
    configurations of its own chain: E_loc per site equal to the oracle's within 1e-11 relative
    (1e-11 absolute below 1).  The reference reads an inverse carried through its chain's
    Sherman-Morrison updates, the oracle a fresh LU: near a node (|E_loc| ~ 24 per site at
    one L=6 configuration) their rounding differs by a few 1e-12 relative."""
    g = np.load(os.path.join(GOLD, name), allow_pickle=False)
    L = int(g["L"])
    occ = O.ground_occ(L, O.orbitals(L, float(g["phi"]) * np.pi, float(g["neel"])))
    for conf, e in zip(g["confs"], g["eloc"]):
        w = O.Walker(L, occ, conf.astype(np.int32))
        assert w.ok
        assert abs(w.eloc() - e) <= 1e-11 * max(1.0, abs(e)), (conf, e, w.eloc())
