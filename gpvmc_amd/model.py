"""Host-side model setup for the batched VMC path: lattice tables, the staggered-flux
orbital table of the ground-state channel, and Philox-driven streams/configurations.

These run once per job (O(L^4) orbital table); the per-step work is on the GPU.
"""
import numpy as np

# ------------------------------------------------------------------ lattice
def edges(L):
    """Edge list of SquareLattice (SquareLattice.cpp:22-27): x outer, y inner, first the
    (x,y)->(x,y+1) edge then (x,y)->(x+1,y).  Sites are v = y*L + x."""
    out = []
    for x in range(L):
        for y in range(L):
            out.append((y * L + x, ((y + 1) % L) * L + x))
            out.append((y * L + x, y * L + (x + 1) % L))
    return np.array(out, np.int32)


def neighbours(L):
    """Per-vertex neighbour list in edge-creation order (Edge ctor, Lattice.h:79-84)."""
    Ns = L * L
    nb = [[] for _ in range(Ns)]
    for a, b in edges(L):
        nb[a].append(b)
        nb[b].append(a)
    return np.array(nb, np.int32)


# ------------------------------------------------------------------ orbitals
def magnetic_zone(L, bcx=1.0, bcy=1.0):
    """Orbital quantum numbers (kx, ky, band), kx outer / ky inner over the magnetic
    Brillouin zone, two bands per momentum (StagFluxWaveFunction.cpp:24-41, 89-101)."""
    out = []
    eps = 1e-9
    for kx in range(L):
        for ky in range(L):
            kkx = (kx + bcx / 2.0) / L
            kky = (ky + bcy / 2.0) / L
            inside = (0 <= kkx < 1 and 0 <= kky < 1) and (
                kkx + kky - 0.5 < 0 or kkx + kky - 1.5 >= -eps or kky - kkx - 0.5 > eps
                or kky - kkx + 0.5 <= eps)
            if inside:
                out.append((kx, ky, 0))
                out.append((kx, ky, 1))
    return np.array(out, np.int32)


def orbital_table(L, phi, neel, neel_exp=0.0, bcx=1.0, bcy=1.0):
    """Phi[fl][orbital f][site r] of the staggered-flux + Neel mean-field state
    (StagFluxWaveFunction.cpp:51-163).  `phi` in radians (the command-line value times
    pi, vmc.cpp:120).  Flavour 0 has sigma = +1, flavour 1 sigma = -1."""
    Ns = L * L
    qn = magnetic_zone(L, bcx, bcy)
    if len(qn) != Ns:
        raise ValueError("magnetic zone enumeration found %d orbitals, expected %d" % (len(qn), Ns))
    k0 = (qn[:, 0] + bcx / 2.0) * 2 * np.pi / L
    k1 = (qn[:, 1] + bcy / 2.0) * 2 * np.pi / L
    band = qn[:, 2]
    d = 0.5 * (np.cos(k0) * np.exp(1j * phi) + np.cos(k1) * np.exp(-1j * phi))
    if neel_exp == 0.0:
        nk = np.full(Ns, float(neel))
    else:
        nk = neel * (0.5 * (np.cos(k0) ** 2 + np.cos(k1) ** 2)) ** neel_exp
    om = np.sqrt(nk ** 2 + np.abs(d) ** 2)
    r = np.arange(Ns)
    ry = r // L
    rx = r - ry * L
    odd = (rx + ry) % 2
    phase = np.exp(1j * (np.outer(k0, rx) + np.outer(k1, ry)))
    table = np.zeros((2, Ns, Ns), np.complex128)
    with np.errstate(invalid="raise", divide="raise"):
        for fl, sig in ((0, 1.0), (1, -1.0)):
            up = np.sqrt(0.5 * (1 + sig * nk / om))
            dn = np.sqrt(0.5 * (1 - sig * nk / om))
            U = np.where(band == 0, up, -np.conj(d) / np.abs(d) * dn)
            V = np.where(band == 0, d / np.abs(d) * dn, up)
            table[fl] = np.sqrt(2.0) * phase * np.where(odd[None, :] == 0, U[:, None], V[:, None])
    return table


def ground_occ(table):
    """Occupied-orbital table of the ground state: particle p sits in orbital 2p
    (StagFluxGroundState.cpp:11-14); occ[fl][site][p] = Phi_fl(2p, site)."""
    return np.ascontiguousarray(np.transpose(table[:, 0::2, :], (0, 2, 1)))


# ------------------------------------------------------------------ Philox4x32-10
_M0, _M1 = np.uint64(0xD2511F53), np.uint64(0xCD9E8D57)
_W0, _W1 = np.uint32(0x9E3779B9), np.uint32(0xBB67AE85)
_MASK = np.uint64(0xFFFFFFFF)


def philox4x32_10(ctr, key):
    """Vectorised Philox4x32-10 (Salmon et al., SC'11).  ctr: uint32 [..., 4];
    key: uint32 [..., 2].  Returns uint32 [..., 4]."""
    c = [np.asarray(ctr[..., i], np.uint64) for i in range(4)]
    k0 = np.asarray(key[..., 0], np.uint32)
    k1 = np.asarray(key[..., 1], np.uint32)
    for _ in range(10):
        p0 = _M0 * c[0]
        p1 = _M1 * c[2]
        hi0, lo0 = p0 >> np.uint64(32), p0 & _MASK
        hi1, lo1 = p1 >> np.uint64(32), p1 & _MASK
        c = [hi1 ^ c[1] ^ k0.astype(np.uint64), lo1, hi0 ^ c[3] ^ k1.astype(np.uint64), lo0]
        k0 = (k0 + _W0).astype(np.uint32)
        k1 = (k1 + _W1).astype(np.uint32)
    return np.stack([x.astype(np.uint32) for x in c], axis=-1)


def philox_uniforms(seed, walkers, step0, n_steps):
    """The device stream of gpvmc_walk in Philox mode: for walker w and step s the
    counter is (s_lo, s_hi, w, 0) and the key (seed_lo, seed_hi); u = (x + 1/2) 2^-32.
    Returns float64 [len(walkers), n_steps, 4]."""
    walkers = np.asarray(walkers, np.uint64)
    s = np.arange(step0, step0 + n_steps, dtype=np.uint64)
    W = len(walkers)
    ctr = np.zeros((W, n_steps, 4), np.uint64)
    ctr[..., 0] = (s & _MASK)[None, :]
    ctr[..., 1] = (s >> np.uint64(32))[None, :]
    ctr[..., 2] = walkers[:, None]
    key = np.zeros((W, n_steps, 2), np.uint32)
    key[..., 0] = np.uint32(seed & 0xFFFFFFFF)
    key[..., 1] = np.uint32((seed >> 32) & 0xFFFFFFFF)
    x = philox4x32_10(ctr, key)
    return (x.astype(np.float64) + 0.5) * 2.0 ** -32


def random_spins(L, n_walkers, seed):
    """Half-filled initial configurations: a Fisher-Yates shuffle of N up / N down per
    walker, drawn from Philox4x32-10 with key `seed` and counter (i, 0, w, 1)."""
    Ns = L * L
    N = Ns // 2
    out = np.zeros((n_walkers, Ns), np.int32)
    base = np.array([0] * N + [1] * N, np.int32)
    ctr = np.zeros((n_walkers, Ns, 4), np.uint64)
    ctr[..., 0] = np.arange(Ns, dtype=np.uint64)[None, :]
    ctr[..., 2] = np.arange(n_walkers, dtype=np.uint64)[:, None]
    ctr[..., 3] = 1
    key = np.zeros((n_walkers, Ns, 2), np.uint32)
    key[..., 0] = np.uint32(seed & 0xFFFFFFFF)
    key[..., 1] = np.uint32((seed >> 32) & 0xFFFFFFFF)
    x = philox4x32_10(ctr, key)[..., 0].astype(np.float64) * 2.0 ** -32
    out[:] = base[None, :]
    rows = np.arange(n_walkers)
    for i in range(Ns - 1, 0, -1):
        j = (x[:, i] * (i + 1)).astype(np.int64)
        a = out[rows, i].copy()
        out[rows, i] = out[rows, j]
        out[rows, j] = a
    return out
