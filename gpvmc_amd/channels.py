"""Channels of the staggered-flux wavefunction: the Slater-determinant states a walker's
weight is summed over, and the orbital hops that connect them.

A channel is a list of N_k states.  Each state fixes the occupied orbitals of both
flavours (orbital index f = 2 * (magnetic-zone momentum index) + band).  The walker's
weight is w(x) = sum_k |a_k(x)|^2, where a_k(x) is the product over flavours of the
determinant det[Phi_fl(o_i, site of particle j)] with the state's orbitals o_i in
increasing order (rows) and the lattice particles by label (columns).  This is the
amplitude that SlaterDeterminant::VirtUpdate returns for state k, including the
wavefunction sign GetSign() * HopSign() that relates the reference's slot arrangement to
increasing orbital order (SlaterDeterminant.cpp:133-137, State.cpp:86-115).

The states are those of:
- StagFluxGroundState.cpp:3-15: one state, the even (band-0) orbitals;
- StagFluxTransExciton.cpp:7-44: for each odd orbital 2fe+1, flavour 0 holds the even
  orbitals plus 2fe+1 (N+1 particles) and flavour 1 the even orbitals except the band-0
  orbital at momentum k(2fe+1) - q folded into the magnetic zone (N-1 particles);
- StagFluxLongExciton.cpp:6-66: for each odd orbital, the particle-hole pair in flavour 0
  then in flavour 1, plus the ground state when q is 0 or (L/2, L/2).

`hops[c][k][fl]` is the path WaveFunction::AddState stores in m_exc (WaveFunction.cpp:39-67,
State::get_hop_path, State.cpp:213-231): the orbitals occupied in c but not in k, in
increasing order, paired with those occupied in k but not in c.  `hopsign[c][k]` is
State::HopSign of that path evaluated on state c's occupation.
"""
import numpy as np

from . import model


def _fock_tables(L, bcx, bcy):
    qn = model.magnetic_zone(L, bcx, bcy)
    qn2fock = np.full(2 * L * L, L * L, np.int64)
    for idx, (kx, ky, band) in enumerate(qn):
        qn2fock[(kx * L + ky) * 2 + band] = idx
    return qn, qn2fock


def _inmbz(L, bcx, bcy, kx, ky):
    kkx = (kx + bcx / 2.0) / L
    kky = (ky + bcy / 2.0) / L
    eps = 1e-9
    return (0 <= kkx < 1 and 0 <= kky < 1) and (
        kkx + kky - 0.5 < 0 or kkx + kky - 1.5 >= -eps or kky - kkx - 0.5 > eps
        or kky - kkx + 0.5 <= eps)


def _mbzmod(L, bcx, bcy, k):
    """StagFluxWaveFunction::mbzmod (StagFluxWaveFunction.cpp:103-120)."""
    kx, ky = k
    if _inmbz(L, bcx, bcy, kx, ky):
        return kx, ky
    h = L // 2
    if kx < h and ky < h:
        return kx + h, ky + h
    if kx < h and ky >= h:
        return kx + h, ky - h
    if kx >= h and ky < h:
        return kx - h, ky + h
    return kx - h, ky - h


def _hole_orbital(L, bcx, bcy, qn, qn2fock, fe, q):
    kx, ky = int(qn[2 * fe + 1][0]), int(qn[2 * fe + 1][1])
    Qk = _mbzmod(L, bcx, bcy, ((kx - q[0]) % L, (ky - q[1]) % L))
    return int(qn2fock[(Qk[0] * L + Qk[1]) * 2])


class Channel:
    """States and hop graph of one channel (see the module docstring)."""

    def __init__(self, name, L, q=(0, 0), bcx=1.0, bcy=1.0):
        if name not in ("groundstate", "trans", "long"):
            raise ValueError("channel must be groundstate, trans or long")
        self.name, self.L, self.q = name, L, (int(q[0]) % L, int(q[1]) % L)
        N = L * L // 2
        qn, qn2fock = _fock_tables(L, bcx, bcy)
        even = [2 * f for f in range(N)]
        states = []
        if name == "groundstate":
            states.append((even, even))
        elif name == "trans":
            for fe in range(N):
                hole = _hole_orbital(L, bcx, bcy, qn, qn2fock, fe, self.q)
                states.append((even + [2 * fe + 1], [o for o in even if o != hole]))
        else:
            for fe in range(N):
                hole = _hole_orbital(L, bcx, bcy, qn, qn2fock, fe, self.q)
                ph = [o for o in even if o != hole] + [2 * fe + 1]
                states.append((ph, even))
                states.append((even, ph))
            if self.q in ((0, 0), (L // 2, L // 2)):
                states.append((even, even))
        self.states = [(np.array(sorted(u), np.int32), np.array(sorted(d), np.int32)) for u, d in states]
        self.nexc = len(self.states)
        self.npt = (len(self.states[0][0]), len(self.states[0][1]))
        K = self.nexc
        # every state is E = {0, 2, ..., 2N-2} (the band-0 orbitals) with a few orbitals
        # removed and added per flavour, so each pair's path and sign cost O(1)
        E = set(even)
        rem = [[sorted(E - set(st[fl].tolist())) for fl in range(2)] for st in self.states]
        add = [[sorted(set(st[fl].tolist()) - E) for fl in range(2)] for st in self.states]

        def n_even_between(lo, hi):  # members of E strictly between lo and hi
            a, b = lo + 1, min(hi - 1, 2 * N - 2)
            if b < a:
                return 0
            return b // 2 - (a + 1) // 2 + 1

        self.hops = [[[None, None] for _ in range(K)] for _ in range(K)]
        self.hopsign = np.ones((K, K), np.int8)
        for c in range(K):
            for k in range(K):
                odd = False
                for fl in range(2):
                    rc, ac, rk, ak = rem[c][fl], add[c][fl], rem[k][fl], add[k][fl]
                    # occupied in c but not in k: c's additions k lacks, and E-orbitals k
                    # removed that c kept; the reverse for the targets
                    holes = sorted([o for o in ac if o not in ak] + [o for o in rk if o not in rc])
                    parts = sorted([o for o in ak if o not in ac] + [o for o in rc if o not in rk])
                    path = list(zip(holes, parts))
                    self.hops[c][k][fl] = path
                    # State::HopSign on state c's occupation, hop after hop (paths here
                    # have at most two hops per flavour, where the reference's bookkeeping
                    # is exact)
                    gone, came = list(rc), list(ac)
                    for a, b in path:
                        lo, hi = min(a, b), max(a, b)
                        n_between = (n_even_between(lo, hi) - sum(1 for o in gone if lo < o < hi)
                                     + sum(1 for o in came if lo < o < hi))
                        odd ^= bool(n_between & 1)
                        if a in came:
                            came.remove(a)
                        else:
                            gone.append(a)
                        if b in gone:
                            gone.remove(b)
                        else:
                            came.append(b)
                self.hopsign[c, k] = -1 if odd else 1
        self.max_hops = max(len(self.hops[c][k][fl]) for c in range(K) for k in range(K) for fl in range(2))

    def new_orbitals(self, fl):
        """Orbitals of flavour fl that some hop of the graph fills (the rows a replaced
        slot can take)."""
        out = set()
        for c in range(self.nexc):
            for k in range(self.nexc):
                out.update(b for _, b in self.hops[c][k][fl])
        return np.array(sorted(out), np.int32)
