"""ctypes binding of the C oracle (oracle/liboracle.so) for the test-suite.

The oracle is the checker: tests compare the HIP path against it, and it is itself
pinned against fixtures written by the reference-built harness, brute-force linear
algebra and exact enumeration (tests/test_oracle.py).
"""
import ctypes
import os
import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
_LIB = None

_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_c128p = np.ctypeslib.ndpointer(np.complex128, flags="C_CONTIGUOUS")
_i8p = np.ctypeslib.ndpointer(np.int8, flags="C_CONTIGUOUS")


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(ROOT, "oracle", "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (run __graft_entry__.build())")
        L = ctypes.CDLL(path)
        dp = ctypes.c_void_p
        L.or_neighbours.argtypes = [ctypes.c_int, _i32p]
        L.or_edges.argtypes = [ctypes.c_int, _i32p]
        L.or_mbz.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double, _i32p]
        L.or_orbitals.argtypes = [ctypes.c_int] + [ctypes.c_double] * 5 + [_c128p]
        L.or_ground_occ.argtypes = [ctypes.c_int, _c128p, _c128p]
        L.or_detinv.argtypes = [ctypes.c_int, _c128p, _c128p, ctypes.POINTER(ctypes.c_double),
                                ctypes.POINTER(ctypes.c_double)]
        L.or_ratio.argtypes = [ctypes.c_int, _c128p, ctypes.c_int, _c128p, _f64p]
        L.or_sm_update.argtypes = [ctypes.c_int, _c128p, ctypes.c_int, _c128p]
        L.or_raninit.argtypes = [ctypes.c_int, _f64p, ctypes.c_int64, _i32p]
        L.or_labels.argtypes = [ctypes.c_int, _i32p, _i32p, _i32p]
        L.or_init_walker.argtypes = [ctypes.c_int, _c128p, _i32p, _i32p, _i32p, _c128p,
                                     ctypes.POINTER(ctypes.c_double),
                                     ctypes.POINTER(ctypes.c_double)]
        L.or_walk.argtypes = [ctypes.c_int, _c128p, _i32p, _i32p, _i32p, _c128p,
                              ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double),
                              _f64p, ctypes.c_int64, ctypes.c_int64, dp, dp, dp, dp, dp, dp, dp]
        L.or_eloc.argtypes = [ctypes.c_int, _c128p, _i32p, _i32p, _i32p, _c128p, _f64p]
        L.or_stagmagnz.argtypes = [ctypes.c_int, _i32p, _f64p]
        L.or_statspinstruct.argtypes = [ctypes.c_int, _c128p, _i32p, _i32p, _c128p, _c128p]
        L.or_exc_amps.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _c128p, ctypes.c_int,
                                  _i32p, _i32p, _f64p, _f64p]
        L.or_exc_local.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _c128p, ctypes.c_int,
                                   _i32p, _i32p, _i32p, _c128p, _c128p]
        L.or_exc_local.restype = ctypes.c_double
        L.or_exc_walk.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _c128p, ctypes.c_int,
                                  _i32p, _i32p, _i32p, _i32p, _f64p, ctypes.c_int64, ctypes.c_int64,
                                  dp, dp, dp, dp, dp]
        _LIB = L
    return _LIB


def neighbours(L):
    out = np.zeros((L * L, 4), np.int32)
    lib().or_neighbours(L, out)
    return out


def edges(L):
    out = np.zeros((2 * L * L, 2), np.int32)
    lib().or_edges(L, out)
    return out


def mbz(L, bcx=1.0, bcy=1.0):
    out = np.zeros((L * L, 3), np.int32)
    n = lib().or_mbz(L, bcx, bcy, out)
    return out[:n]


def orbitals(L, phi, neel, neel_exp=0.0, bcx=1.0, bcy=1.0):
    """Full table [2][Ns orbitals][Ns sites]; `phi` in radians."""
    Ns = L * L
    t = np.zeros((2, Ns, Ns), np.complex128)
    rc = lib().or_orbitals(L, phi, neel, neel_exp, bcx, bcy, t)
    if rc != 0:
        raise ValueError("or_orbitals failed (%d)" % rc)
    return t


def ground_occ(L, table):
    Ns = L * L
    occ = np.zeros((2, Ns, Ns // 2), np.complex128)
    lib().or_ground_occ(L, np.ascontiguousarray(table), occ)
    return occ


def detinv(A):
    """A in the reference memory layout (A[r*N+f] = A(f, r)); returns (Ai, logabs, phase, singular)."""
    N = A.shape[0]
    A = np.ascontiguousarray(A, np.complex128)
    Ai = np.zeros_like(A)
    la, ph = ctypes.c_double(), ctypes.c_double()
    s = lib().or_detinv(N, A, Ai, ctypes.byref(la), ctypes.byref(ph))
    return Ai, la.value, ph.value, bool(s)


def ratio(Ai, ci, col):
    q = np.zeros(2)
    lib().or_ratio(Ai.shape[0], np.ascontiguousarray(Ai), int(ci), np.ascontiguousarray(col, np.complex128), q)
    return complex(q[0], q[1])


def sm_update(Ai, ci, col):
    Ai = np.ascontiguousarray(Ai, np.complex128).copy()
    lib().or_sm_update(Ai.shape[0], Ai, int(ci), np.ascontiguousarray(col, np.complex128))
    return Ai


def raninit(L, u):
    spins = np.zeros(L * L, np.int32)
    used = lib().or_raninit(L, np.ascontiguousarray(u, np.float64), len(u), spins)
    return spins, used


class Walker:
    """One chain in the oracle: occupation, labels, FP64 inverses and log-amplitude."""

    def __init__(self, L, occ, spins):
        self.L, self.Ns, self.N = L, L * L, L * L // 2
        self.occ = np.ascontiguousarray(occ, np.complex128)
        self.fock = np.zeros(2 * self.Ns, np.int32)
        self.part = np.zeros(2 * self.N, np.int32)
        self.Ai = np.zeros((2, self.N, self.N), np.complex128)
        la, ph = ctypes.c_double(), ctypes.c_double()
        self.ok = bool(lib().or_init_walker(L, self.occ, np.ascontiguousarray(spins, np.int32),
                                            self.fock, self.part, self.Ai,
                                            ctypes.byref(la), ctypes.byref(ph)))
        self.logabs, self.phase = la.value, ph.value
        self.nbr = neighbours(L)
        self.edges = edges(L)

    def spins(self):
        return (self.fock[: self.Ns] == self.N).astype(np.int32)

    def walk(self, u, meas=0):
        u = np.ascontiguousarray(u, np.float64)
        n = len(u) // 4
        noop = np.zeros(n, np.int8)
        acc = np.zeros(n, np.int8)
        rat = np.zeros(n, np.float64)
        nm = 0 if meas <= 0 else (n + meas - 1) // meas if meas > 1 else n
        eloc = np.zeros(max(nm, 1), np.complex128)
        mz = np.zeros(max(nm, 1), np.complex128)
        logw = np.zeros(max(nm, 1), np.float64)
        cond = np.zeros(n, np.float64)
        la, ph = ctypes.c_double(self.logabs), ctypes.c_double(self.phase)
        got = lib().or_walk(self.L, self.occ, self.nbr, self.fock, self.part, self.Ai,
                            ctypes.byref(la), ctypes.byref(ph), u, n, meas,
                            noop.ctypes.data, acc.ctypes.data, rat.ctypes.data,
                            eloc.ctypes.data, mz.ctypes.data, logw.ctypes.data, cond.ctypes.data)
        self.logabs, self.phase = la.value, ph.value
        return dict(noop=noop, accept=acc, ratio=rat, eloc=eloc[:got], mz=mz[:got], logw=logw[:got],
                    cond=cond)

    def eloc(self):
        out = np.zeros(2)
        lib().or_eloc(self.L, self.occ, self.edges, self.fock, self.part, self.Ai, out)
        return complex(out[0], out[1])

    def stagmagnz(self):
        out = np.zeros(2)
        lib().or_stagmagnz(self.L, self.fock, out)
        return complex(out[0], out[1])

    def statspinstruct(self, acc=None):
        if acc is None:
            acc = np.zeros((5, self.Ns, self.Ns), np.complex128)
        lib().or_statspinstruct(self.L, self.occ, self.fock, self.part, self.Ai, acc)
        return acc


def glibc_uniforms(seed, n):
    """The reference's USE_RNG_STD stream: std::srand(seed); rand()/RAND_MAX (RanGen.cpp:86-116)."""
    libc = ctypes.CDLL("libc.so.6")
    libc.srand(ctypes.c_uint(seed & 0xFFFFFFFF))
    libc.rand.restype = ctypes.c_int
    return np.array([libc.rand() for _ in range(n)], np.float64) / 2147483647.0


class ExcWalker:
    """One chain on a channel in the oracle (exact amplitudes by LU at every step).

    table: full orbital table [2][Ns][Ns]; ch: gpvmc_amd.channels.Channel; spins[site] in
    {0, 1} with ch.npt particles of each flavour.  Labels in increasing site order."""

    def __init__(self, L, table, ch, spins):
        self.L, self.Ns = L, L * L
        self.nup, self.ndn = ch.npt
        self.nk = ch.nexc
        self.tab = np.ascontiguousarray(table, np.complex128)
        self.orbs = np.ascontiguousarray(
            np.array([np.concatenate([u, d]) for u, d in ch.states], np.int32))
        spins = np.asarray(spins)
        self.site = np.zeros(self.Ns, np.int32)
        self.part = np.zeros(self.nup + self.ndn, np.int32)
        cnt = [0, 0]
        for v in range(self.Ns):
            fl = int(spins[v])
            self.site[v] = (fl << 16) | cnt[fl]
            self.part[cnt[fl] + (self.nup if fl else 0)] = v
            cnt[fl] += 1
        assert cnt == [self.nup, self.ndn]
        self.nbr = neighbours(L)

    def spins(self):
        return (self.site >> 16).astype(np.int32)

    def amps(self, part=None):
        la = np.zeros(self.nk)
        ph = np.zeros(self.nk)
        p = self.part if part is None else np.ascontiguousarray(part, np.int32)
        lib().or_exc_amps(self.Ns, self.nup, self.ndn, self.tab, self.nk, self.orbs, p, la, ph)
        return la, ph

    def local(self):
        """(a_k, h_k, w) at the current configuration, scaled so that max |a_k| = 1."""
        a = np.zeros(self.nk, np.complex128)
        h = np.zeros(self.nk, np.complex128)
        w = lib().or_exc_local(self.L, self.nup, self.ndn, self.tab, self.nk, self.orbs,
                               self.site, self.part, a, h)
        return a, h, w

    def walk(self, u, meas=0, HO=None):
        u = np.ascontiguousarray(u, np.float64)
        n = len(u) // 4
        noop = np.zeros(n, np.int8)
        acc = np.zeros(n, np.int8)
        rat = np.zeros(n, np.float64)
        nm = 0 if meas <= 0 else (n + meas - 1) // meas if meas > 1 else n
        if HO is None:
            HO = np.zeros((2 * self.nk, self.nk), np.complex128)
        logw = np.zeros(max(nm, 1), np.float64)
        got = lib().or_exc_walk(self.L, self.nup, self.ndn, self.tab, self.nk, self.orbs,
                                self.nbr, self.site, self.part, u, n, meas, noop.ctypes.data,
                                acc.ctypes.data, rat.ctypes.data, HO.ctypes.data, logw.ctypes.data)
        return dict(noop=noop, accept=acc, ratio=rat, HO=HO, logw=logw[:got], n_meas=got)
