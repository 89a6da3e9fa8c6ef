"""Post-processing numerics (gpvmc_amd/postproc.py; the reference's tools/vmc_postproc), CPU.

Pins, each independent of the code it checks:
* geneigh (proc.py:11-42): the generalised eigen-identity A U = B U diag(E) with U^H B U = 1;
  on a shared null space NaN eigenvalues and, for the rest, the eigenvalues of the problem
  restricted to range(B) computed through an SVD basis; a warning when A leaks into the
  null space and an error when B has no range.
* fermisigns (proc.py:44-72): every channel state (trans, long; L = 4, 6) differs from the
  mean-field reference by one particle-hole pair (h, p), and its sign equals the
  Jordan-Wigner sign of c+_p c_h on the ordered Fock state, computed from the occupations.
* mbz (recspace.py) equals the product's magnetic zone (itself pinned against the
  reference's orbital dumps); evecs (stagflux.py) are eigenvectors of mfham and equal the
  U/V coefficients of the product's orbital table.
* spinops (stagflux.py:62-80): each amplitude equals the real-space matrix element
  (1/N) sum_r e^{i q.r} conj(Phi_c(k, band 1; r)) Phi_a(k - q, band 0; r) of S^+_q (trans)
  or of S^z_q (long; +1/2 for flavour 0, -1/2 for flavour 1) built from the orbital table,
  k - q folded into the magnetic zone.
* get_stat_spin_struct (hl.py:9-66): S(q) from a StatSpinStruct record equals the direct
  double sum (1/N) sum_ij C_ij exp(2 pi i (a dy + b dx) / L), d = r_i - r_j, for each of the
  zz, xx = (+- + -+ + ++ + --)/4 and yy = (+- + -+ - ++ - --)/4 combinations; on the exact
  L = 4 state's <S^z_i S^z_j> (tests/exact.py), S^zz(0) = 0 and sum_q S^zz(q) = N/4.
* get_eig_sys / get_sq_ampl (hl.py:68-145) on a channel's files: H and O conjugated by the
  signs, H v = E O v, the spectral weights summing to s^H O s (O-orthonormal eigenvectors)
  and, with the S(q) rescaling, s^H O s = 2 S(q).
* bunch (load.py:50-82): worked cases and invariants; get_quantity averages its bunches.
"""
import ctypes
import os
import warnings

import numpy as np
import pytest

import exact
import oracle_ref as O
from gpvmc_amd import build, channels, model
from gpvmc_amd import postproc as P


def host_lib():
    if not os.path.exists(build.HOST_LIB):
        pytest.skip("libgpvmc_host.so not built")
    lib = ctypes.CDLL(build.HOST_LIB)
    lib.gpvmc_host_records_write.argtypes = [ctypes.c_char_p] * 4 + [ctypes.c_int32] * 4 + \
        [ctypes.c_void_p, ctypes.c_void_p]
    lib.gpvmc_host_wavefunction_save.argtypes = [ctypes.c_char_p] * 3 + [ctypes.c_int32, ctypes.c_double,
                                                                         ctypes.c_double] + [ctypes.c_int32] * 3
    lib.gpvmc_host_last_error.restype = ctypes.c_char_p
    return lib


def params(L, ch, q, phi=0.25, neel=0.1):
    return dict(L=L, phi=phi, neel=neel, neel_exp=0.0, qx=q[0], qy=q[1], channel=ch, phase_shift_x=1.0,
                phase_shift_y=1.0, stagflux_wav=1.0, sfpnxphz_wav=0.0)


def root_spec(p):
    return ";".join("%s=%s" % kv for kv in p.items()).encode()


def write_records(lib, d, prefix, name, p, mats, stats):
    """mats [write][chain][rows][cols] complex -> records with (re, im) pairs along rows"""
    mats = np.asarray(mats, np.complex128)
    nw, nc, rows, cols = mats.shape
    data = np.zeros((nw, nc, rows, 2 * cols))
    data[..., ::2] = mats.real
    data[..., 1::2] = mats.imag
    st = np.ascontiguousarray(stats, np.float64)
    assert lib.gpvmc_host_records_write(str(d).encode(), prefix.encode(), name.encode(), root_spec(p), nc, nw, rows,
                                        2 * cols, np.ascontiguousarray(data).ctypes.data, st.ctypes.data) == 0, \
        lib.gpvmc_host_last_error()
    return os.path.join(str(d), "%s-%s.h5" % (prefix, name))


def herm(rng, n):
    a = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    return a + a.conj().T


# ----------------------------------------------------------------------------- geneigh
def test_geneigh_full_rank():
    rng = np.random.default_rng(1)
    n = 7
    A = herm(rng, n)
    X = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    B = X @ X.conj().T + np.eye(n)
    E, U = P.geneigh(A, B)
    np.testing.assert_allclose(A @ U, B @ U @ np.diag(E), atol=1e-10)
    np.testing.assert_allclose(U.conj().T @ B @ U, np.eye(n), atol=1e-10)
    assert np.all(np.diff(E) >= 0)


@pytest.mark.parametrize("n,m", [(6, 1), (8, 3)])
def test_geneigh_shared_null_space(n, m):
    rng = np.random.default_rng(n + m)
    V, _ = np.linalg.qr(rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n)))
    b = np.concatenate([np.zeros(m), rng.uniform(0.5, 2.0, n - m)])
    B = V @ np.diag(b) @ V.conj().T
    At = np.zeros((n, n), complex)
    At[m:, m:] = herm(rng, n - m)
    A = V @ At @ V.conj().T
    E, U = P.geneigh(A, B, tol=1e-10)
    assert np.isnan(E[:m]).all() and np.isfinite(E[m:]).all()
    # the problem restricted to range(B), through an SVD basis W of B's range
    u, s, _ = np.linalg.svd(B)
    W = u[:, :n - m]
    Br, Ar = W.conj().T @ B @ W, W.conj().T @ A @ W
    Bh = np.linalg.inv(np.linalg.cholesky(Br))
    ref = np.linalg.eigvalsh(Bh @ Ar @ Bh.conj().T)
    np.testing.assert_allclose(E[m:], ref, atol=1e-9)
    np.testing.assert_allclose(A @ U[:, m:], B @ U[:, m:] @ np.diag(E[m:]), atol=1e-9)


def test_geneigh_warns_and_fails():
    rng = np.random.default_rng(3)
    n = 5
    B = np.diag([0.0, 1.0, 2.0, 3.0, 4.0]).astype(complex)
    A = herm(rng, n)  # couples B's null vector to the rest
    with warnings.catch_warnings(record=True) as w:
        warnings.simplefilter("always")
        P.geneigh(A, B)
    assert any(issubclass(x.category, UserWarning) for x in w)
    with pytest.raises(RuntimeError, match="null-space could not be found"):
        P.geneigh(A, np.zeros((n, n), complex))


# ----------------------------------------------------------------------------- fermisigns
def jordan_wigner_pair_sign(occ, h, p):
    """sign of c+_p c_h |occ> on the ordered Fock state (occ[h] = 1, occ[p] = 0)"""
    n = np.array(occ, int)
    s = (-1) ** n[:h].sum()
    n[h] = 0
    s *= (-1) ** n[:p].sum()
    return s


@pytest.mark.parametrize("name,L,q", [("trans", 4, (1, 2)), ("trans", 6, (2, 3)), ("long", 4, (1, 0)),
                                      ("long", 6, (2, 1)), ("trans", 4, (0, 0))])
def test_fermisigns_equal_jordan_wigner_signs(tmp_path, name, L, q):
    lib = host_lib()
    p = params(L, name, q)
    ch_id = {"trans": 1, "long": 2}[name]
    assert lib.gpvmc_host_wavefunction_save(str(tmp_path).encode(), b"3", root_spec(p), L, 0.25, 0.1, ch_id,
                                            *q) == 0, lib.gpvmc_host_last_error()
    st = P.get_wav(str(tmp_path / "3-WaveFunction.h5"))
    ch = channels.Channel(name, L, q)
    assert st.shape == (ch.nexc, 2 * L * L)
    fs = P.fermisigns(st, p)
    ref = P.refstate(p)
    for s in range(st.shape[0]):
        holes = np.nonzero((ref != 0) & (st[s] == 0))[0]
        parts = np.nonzero((ref == 0) & (st[s] != 0))[0]
        assert len(holes) == len(parts) == 1
        assert fs[s] == jordan_wigner_pair_sign(ref, holes[0], parts[0])


def test_fermisigns_rejects_unbalanced_states():
    p = params(4, "trans", (1, 2))
    st = P.refstate(p)[None, :].copy()
    st[0, 1] = 1
    with pytest.raises(RuntimeError, match="not conserved"):
        P.fermisigns(st, p)


# ----------------------------------------------------------------------------- zone and orbitals
@pytest.mark.parametrize("L,bc", [(4, (1.0, 1.0)), (6, (1.0, 1.0)), (8, (1.0, 0.0)), (10, (0.0, 1.0)),
                                  (12, (1.0, 1.0))])
def test_mbz_equals_product_zone(L, bc):
    p = dict(L=L, phase_shift_x=bc[0], phase_shift_y=bc[1])
    kx, ky = P.mbz(p)
    qn = model.magnetic_zone(L, *bc)
    qn = qn[qn[:, 2] == 0]
    np.testing.assert_allclose(kx, (qn[:, 0] + bc[0] / 2) / L, atol=1e-14)
    np.testing.assert_allclose(ky, (qn[:, 1] + bc[1] / 2) / L, atol=1e-14)


@pytest.mark.parametrize("L,phi,neel", [(4, 0.25, 0.1), (6, 0.1667, 0.05), (8, 0.3, 0.0)])
def test_evecs_diagonalise_mfham_and_equal_orbital_table(L, phi, neel):
    p = params(L, "trans", (0, 0), phi, neel)
    kx, ky = P.mbz(p)
    H, V, E = P.mfham(kx, ky, p), P.evecs(kx, ky, p), P.evals(kx, ky, p)
    for i in range(len(kx)):
        np.testing.assert_allclose(H[:, :, i] @ V[:, :, i], V[:, :, i] @ np.diag(E[:, i]), atol=1e-13)
        np.testing.assert_allclose(V[:, :, i].conj().T @ V[:, :, i], np.eye(4), atol=1e-13)
    tab = model.orbital_table(L, phi * np.pi, neel)
    # orbital 2i + band at momentum i; site 0 = (0, 0) on sublattice A, site 1 = (1, 0) on B
    for i in range(len(kx)):
        ph1 = np.exp(2j * np.pi * kx[i])
        for fl, band, col in ((0, 0, 0), (1, 0, 1), (0, 1, 2), (1, 1, 3)):
            f = 2 * i + band
            np.testing.assert_allclose(tab[fl, f, 0] / np.sqrt(2), V[2 * fl, col, i], atol=1e-13)
            np.testing.assert_allclose(tab[fl, f, 1] / np.sqrt(2) / ph1, V[2 * fl + 1, col, i], atol=1e-13)


def orbital_at(qn, L, kf, band):
    """index of the product's orbital of momentum kf (units of 2 pi, modulo 1) and band, or of
    its magnetic-zone image kf + (1/2, 1/2)"""
    for shift in (0.0, 0.5):
        for f, (a, c, b) in enumerate(qn):
            if b == band and abs(((a + 0.5) / L - kf[0] - shift + 0.25) % 1.0 - 0.25) < 1e-9 \
                    and abs(((c + 0.5) / L - kf[1] - shift + 0.25) % 1.0 - 0.25) < 1e-9:
                return f
    raise AssertionError("no orbital at %s" % (kf,))


@pytest.mark.parametrize("name,L,q", [("trans", 4, (1, 2)), ("trans", 6, (2, 3)), ("trans", 4, (0, 0)),
                                      ("long", 4, (1, 0)), ("long", 6, (2, 1))])
def test_spinops_equal_real_space_matrix_elements(name, L, q):
    phi, neel = 0.25, 0.1
    p = params(L, name, q, phi, neel)
    so = P.spinops(p)
    tab = model.orbital_table(L, phi * np.pi, neel)
    qn = model.magnetic_zone(L)
    N = L * L
    r = np.arange(N)
    eq = np.exp(2j * np.pi * (q[0] * (r % L) + q[1] * (r // L)) / L)
    kx, ky = P.mbz(p)
    checked = 0
    for i in range(len(kx)):
        fc = orbital_at(qn, L, (kx[i], ky[i]), 1)
        fa = orbital_at(qn, L, (kx[i] - q[0] / L, ky[i] - q[1] / L), 0)
        # where delta(k) ~ 1e-17 (L = 2 mod 4) the orbital's phase d/|d| is rounding: only the
        # magnitude of the matrix element is defined there
        degenerate = min(abs(P.delta(kx[i], ky[i], p)),
                         abs(P.delta(kx[i] - q[0] / L, ky[i] - q[1] / L, p))) < 1e-12
        pairs = []
        if name == "trans":
            m = np.sum(eq * np.conj(tab[0, fc]) * tab[1, fa]) / N
            pairs = [(so[0][i], m), (so[1][i], m)]
        else:
            for fl, sgn in ((0, 0.5), (1, -0.5)):
                pairs.append((so[2][2 * i + fl], sgn * np.sum(eq * np.conj(tab[fl, fc]) * tab[fl, fa]) / N))
        for a, b in pairs:
            assert (abs(abs(a) - abs(b)) if degenerate else abs(a - b)) < 1e-13
        checked += not degenerate
    assert checked >= len(kx) - 4


def test_spinops_refuses_long_channel_at_zero():
    with pytest.raises(ValueError):
        P.spinops(params(4, "long", (0, 0)))


# ----------------------------------------------------------------------------- S(q)
def direct_sq(C, L):
    """(1/N) sum_ij C_ij exp(2 pi i (a dy + b dx) / L), d = r_i - r_j, as [a][b]"""
    N = L * L
    r = np.arange(N)
    x, y = r % L, r // L
    dx = x[:, None] - x[None, :]
    dy = y[:, None] - y[None, :]
    out = np.zeros((L, L), complex)
    for a in range(L):
        for b in range(L):
            out[a, b] = np.sum(C * np.exp(2j * np.pi * (a * dy + b * dx) / L)) / N
    return out


@pytest.mark.parametrize("L", [4, 6])
def test_stat_spin_struct_equals_direct_sum(tmp_path, L):
    lib = host_lib()
    rng = np.random.default_rng(L)
    N = L * L
    C = rng.normal(size=(5, N, N)) + 1j * rng.normal(size=(5, N, N))
    f = write_records(lib, tmp_path, "4", "StatSpinStruct", params(L, "groundstate", (0, 0)),
                      C.reshape(1, 1, 5, N * N), [100.0])
    (sxx, syy, szz), (rxx, ryy, rzz) = P.get_stat_spin_struct(f, 1)
    np.testing.assert_allclose(szz[0], direct_sq(C[0], L), atol=1e-11)
    np.testing.assert_allclose(sxx[0], direct_sq(0.25 * (C[1] + C[2] + C[3] + C[4]), L), atol=1e-11)
    np.testing.assert_allclose(syy[0], direct_sq(0.25 * (C[1] + C[2] - C[3] - C[4]), L), atol=1e-11)
    # real space: the zz displacement sums at (dy + L/2, dx + L/2)
    h = L // 2
    r = np.arange(N)
    x, y = r % L, r // L
    for i, j in ((0, 0), (1, 5), (7, 3)):
        dx = (x[i] - x[j] + h) % L - h
        dy = (y[i] - y[j] + h) % L - h
        sel = [(a, b) for a in range(N) for b in range(N)
               if (x[a] - x[b] + h) % L - h == dx and (y[a] - y[b] + h) % L - h == dy]
        want = sum(C[0][a, b] for a, b in sel) / N
        assert abs(rzz[0][dy + h, dx + h] - want) < 1e-12


def test_stat_spin_struct_sum_rules_on_exact_state(tmp_path):
    """<S^z_i S^z_j> of the exact L = 4 variational state: S^zz(0) = 0 (S^z_tot = 0) and
    sum_q S^zz(q) = sum_i <(S^z_i)^2> = N / 4"""
    lib = host_lib()
    L = 4
    N = L * L
    occ = O.ground_occ(L, O.orbitals(L, 0.25 * np.pi, 0.1))
    confs = exact.configurations(L)
    psi = exact.spin_basis_psi(confs, occ)
    w = np.abs(psi) ** 2
    w /= w.sum()
    sz = 0.5 - np.asarray(confs, float)  # flavour 0 (up) -> +1/2
    C = np.zeros((5, N, N), complex)
    C[0] = np.einsum("c,ci,cj->ij", w, sz, sz)
    f = write_records(lib, tmp_path, "4", "StatSpinStruct", params(L, "groundstate", (0, 0)),
                      C.reshape(1, 1, 5, N * N), [1.0])
    (_, _, szz), _ = P.get_stat_spin_struct(f, 1)
    assert abs(szz[0][0, 0]) < 1e-12
    assert abs(szz[0].sum() - N / 4) < 1e-12


def test_stat_spin_struct_old_storage(tmp_path):
    lib = host_lib()
    L, N = 4, 16
    rng = np.random.default_rng(9)
    S = rng.normal(size=(5, N)) + 1j * rng.normal(size=(5, N))
    f = write_records(lib, tmp_path, "4", "StatSpinStruct", params(L, "groundstate", (0, 0)),
                      S.reshape(1, 1, 5, N), [1.0])
    (sxx, syy, szz), (rxx, _, _) = P.get_stat_spin_struct(f, 1)
    np.testing.assert_allclose(szz[0], S[0].reshape(L, L), atol=1e-14)
    np.testing.assert_allclose(sxx[0], 0.25 * (S[1] + S[2] + S[3] + S[4]).reshape(L, L), atol=1e-14)
    np.testing.assert_allclose(np.fft.ifft2(np.fft.ifftshift(rxx[0])) * N, sxx[0], atol=1e-12)


# ----------------------------------------------------------------------------- loader
def test_bunch_worked_cases():
    assert P.bunch([1, 2, 3, 4], 2) == [[1, 4], [3, 2]]
    b, a = P.bunch([5, 1, 4, 2, 3], 2, indices=True)
    # sorted 1 2 3 4 5 (indices 1 3 4 2 0): (1, 5) -> bunch 0; 3 left (not more than 2*2 - 1):
    # 4 -> bunch 1 (total 0), 3 -> bunch 1 (4 < 6), 2 -> bunch 0 (6 < 7)
    assert b == [[1, 5, 2], [4, 3]] and a == [[1, 0, 3], [2, 4]]
    # equal values keep their order (stable sort)
    assert P.bunch([7, 7, 7], 3, indices=True)[1] == [[2], [1], [0]]


@pytest.mark.parametrize("n,nsamp,seed", [(10, 3, 1), (33, 4, 2), (5, 5, 3)])
def test_bunch_invariants(n, nsamp, seed):
    st = list(np.random.default_rng(seed).integers(1, 500, n))
    b, a = P.bunch(st, nsamp, indices=True)
    assert sorted(i for x in a for i in x) == list(range(n))
    for bb, aa in zip(b, a):
        assert bb == [st[i] for i in aa]
    tot = [sum(x) for x in b]
    assert max(tot) - min(tot) <= max(st)


def test_get_quantity_means_of_bunches(tmp_path):
    lib = host_lib()
    rng = np.random.default_rng(4)
    mats = rng.normal(size=(3, 2, 2, 3)) + 1j * rng.normal(size=(3, 2, 2, 3))  # 3 writes, 2 chains
    f = write_records(lib, tmp_path, "6", "ProjHeis", params(4, "trans", (1, 2)), mats, [10.0, 30.0, 20.0])
    Q = P.get_quantity(f, 2)
    # records in name order: rank-1/data-0..2, rank-2/data-0..2; statistics 10 30 20 10 30 20
    recs = [mats[k, c] for c in range(2) for k in range(3)]
    st = [10, 30, 20, 10, 30, 20]
    _, args = P.bunch(st, 2, indices=True)
    for s in range(2):
        np.testing.assert_allclose(Q[s], np.mean([recs[i] for i in args[s]], axis=0), atol=1e-15)


# ----------------------------------------------------------------------------- eigen system
@pytest.mark.parametrize("name,L,q", [("trans", 4, (1, 2)), ("long", 4, (1, 0))])
def test_eig_sys_and_spectral_weights(tmp_path, name, L, q):
    lib = host_lib()
    p = params(L, name, q)
    ch_id = {"trans": 1, "long": 2}[name]
    assert lib.gpvmc_host_wavefunction_save(str(tmp_path).encode(), b"5", root_spec(p), L, 0.25, 0.1, ch_id,
                                            *q) == 0
    nk = channels.Channel(name, L, q).nexc
    rng = np.random.default_rng(nk)
    mats = []
    for w in range(4):
        X = rng.normal(size=(nk, nk)) + 1j * rng.normal(size=(nk, nk))
        Om = X @ X.conj().T + 0.1 * np.eye(nk)
        mats.append([np.vstack([herm(rng, nk), Om])])
    f = write_records(lib, tmp_path, "5", "ProjHeis", p, mats, [100.0] * 4)
    H, Om, E, V = P.get_eig_sys(f, 2)
    Q = P.get_quantity(f, 2)
    st = P.get_wav(str(tmp_path / "5-WaveFunction.h5"))
    fs = np.diag(P.fermisigns(st, p))
    s = P.spinops(p)[1 if name == "trans" else 2]
    W = P.get_sq_ampl(f, 2, O=Om, evecs=V)
    for k in range(2):
        np.testing.assert_allclose(H[k], fs @ Q[k, :nk] @ fs, atol=1e-14)
        np.testing.assert_allclose(Om[k], fs @ Q[k, nk:] @ fs, atol=1e-14)
        np.testing.assert_allclose(H[k] @ V[k], Om[k] @ V[k] @ np.diag(E[k]), atol=1e-9)
        np.testing.assert_allclose(W[k].sum(), (s.conj() @ Om[k] @ s).real, rtol=1e-10)
    # with the S(q) rescaling, s^H O s = 2 S(q) for one sample
    Sq = tuple(np.full((1, L, L), 0.37 + 0.0j) for _ in range(3))
    H1, O1, _, _ = P.get_eig_sys(f, 1, statstruct=Sq)
    np.testing.assert_allclose(s.conj() @ O1[0] @ s, 2 * 0.37, rtol=1e-12)
