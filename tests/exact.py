"""Exact enumeration of the L=4 Sz=0 sector (C(16,8) = 12870 configurations).

The spin-basis amplitude is computed from the second-quantised definition, independently
of the reference's label-keeping conventions: with site-ordered basis states
|alpha> = c+_{0,s0} c+_{1,s1} ... |0> and |psi> = prod_f a+_{f,up} prod_f a+_{f,dn} |0>,
<alpha|psi> = det M, M[i, (f, s)] = delta(s_i, s) phi_s(f, i)  (a 16 x 16 determinant).
In this basis spin operators act with the textbook matrix elements (+1/2 flip-flop).
"""
import itertools

import numpy as np


def configurations(L):
    Ns = L * L
    confs = []
    for ups in itertools.combinations(range(Ns), Ns // 2):
        s = np.ones(Ns, np.int32)
        s[list(ups)] = 0
        confs.append(s)
    return np.array(confs)


def masks(confs):
    w = (1 << np.arange(confs.shape[1], dtype=np.int64))
    return ((confs == 0).astype(np.int64) * w).sum(axis=1)


def spin_basis_psi(confs, occ):
    """occ[fl][site][p] = phi_fl(orbital of p, site); returns psi(alpha) via 16x16 dets."""
    nconf, Ns = confs.shape
    N = Ns // 2
    M = np.zeros((nconf, Ns, Ns), np.complex128)
    for i in range(Ns):
        up = confs[:, i] == 0
        M[up, i, :N] = occ[0, i][None, :]
        M[~up, i, N:] = occ[1, i][None, :]
    return np.linalg.det(M)


def local_energy_exact(L, confs, psi, edges):
    """(H psi)(alpha) / psi(alpha) / Ns for the J=1 Heisenberg model on `edges`."""
    m = masks(confs)
    order = np.argsort(m)
    ms = m[order]
    h = np.zeros(len(confs), np.complex128)
    sz = 0.5 - confs  # +1/2 up, -1/2 down
    for (i, j) in edges:
        h += sz[:, i] * sz[:, j]
        anti = confs[:, i] != confs[:, j]
        mb = m ^ (1 << int(i)) ^ (1 << int(j))
        pos = np.minimum(np.searchsorted(ms, mb), len(ms) - 1)
        idx = order[pos]
        h += np.where(anti, 0.5 * psi[idx] / np.where(psi == 0, 1, psi), 0)
    return h / (L * L)


def flip_flop_expectations(confs, psi):
    """Exact <S^z_a S^z_b>, <S^+_b S^-_a> and <S^-_b S^+_a> for all site pairs."""
    nconf, Ns = confs.shape
    w = np.abs(psi) ** 2
    Z = w.sum()
    sz = 0.5 - confs
    szz = np.einsum("c,ca,cb->ab", w, sz, sz) / Z
    m = masks(confs)
    order = np.argsort(m)
    ms = m[order]
    spm = np.zeros((Ns, Ns), np.complex128)  # <S^+_b S^-_a>
    smp = np.zeros((Ns, Ns), np.complex128)  # <S^-_b S^+_a>
    for a in range(Ns):
        for b in range(Ns):
            if a == b:
                spm[a, a] = (w * (confs[:, a] == 0)).sum() / Z
                smp[a, a] = (w * (confs[:, a] == 1)).sum() / Z
                continue
            # S^+_b S^-_a |beta> = |alpha> needs beta: a up, b down; alpha: a down, b up
            sel = (confs[:, a] == 1) & (confs[:, b] == 0)
            mb = m[sel] ^ (1 << a) ^ (1 << b)
            beta = order[np.searchsorted(ms, mb)]
            spm[a, b] = np.sum(np.conj(psi[sel]) * psi[beta]) / Z
            sel2 = (confs[:, a] == 0) & (confs[:, b] == 1)
            mb2 = m[sel2] ^ (1 << a) ^ (1 << b)
            beta2 = order[np.searchsorted(ms, mb2)]
            smp[a, b] = np.sum(np.conj(psi[sel2]) * psi[beta2]) / Z
    return szz, spm, smp


def configurations_n(L, nup):
    """All configurations of L*L sites with nup particles of flavour 0 (spins[site])."""
    Ns = L * L
    confs = []
    for ups in itertools.combinations(range(Ns), nup):
        s = np.ones(Ns, np.int32)
        s[list(ups)] = 0
        confs.append(s)
    return np.array(confs)


def spin_basis_psi_orbitals(confs, table, orbs_up, orbs_dn):
    """Site-basis amplitude of the Slater state with the given occupied orbitals
    (second-quantised, creation operators in increasing orbital order per flavour, up
    first): psi(alpha) = det M, M[i, (fl, p)] = delta(s_i, fl) table[fl][orb_p][i]."""
    nconf, Ns = confs.shape
    nu = len(orbs_up)
    M = np.zeros((nconf, Ns, Ns), np.complex128)
    for i in range(Ns):
        up = confs[:, i] == 0
        M[up, i, :nu] = table[0][orbs_up, i][None, :]
        M[~up, i, nu:] = table[1][orbs_dn, i][None, :]
    return np.linalg.det(M)
