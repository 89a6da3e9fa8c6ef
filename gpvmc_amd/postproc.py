"""Post-processing of the output files: the numerics of the reference's tools/vmc_postproc,
host-side (numpy / scipy), with the same names and argument meanings:

* load.py:     get_attr, get_stat, bunch, argsort, get_quantity, get_wav
* proc.py:     geneigh (11-42), fermisigns (44-72)
* recspace.py: mbz, inmbz, fold, mbzmod
* stagflux.py: delta, neelk, mfham, evals, evecs, spinops, refstate
* hl.py:       get_stat_spin_struct (9-66), get_eig_sys (68-116), get_sq_ampl (118-145)

The files are read through the host library's HDF5 (libgpvmc_host.so), since the image has
no h5py.  Where the reference's Python 2 code cannot run under Python 3 (true division
used as an index, hl.py:73; `array == None` tests, hl.py:120) the intended integer and
identity operations are used.
"""
import copy
import ctypes
import re
import warnings

import numpy as np
import numpy.fft as fft
import numpy.linalg as lg
from numpy import ma
from scipy.linalg import eigh

from . import build

# ----------------------------------------------------------------------------- HDF5 access
_lib = None


def _host():
    global _lib
    if _lib is None:
        lib = ctypes.CDLL(build.HOST_LIB)
        c, v, i64 = ctypes.c_char_p, ctypes.c_void_p, ctypes.c_int64
        lib.gpvmc_host_h5_children.argtypes = [c, c, v, i64]
        lib.gpvmc_host_h5_children.restype = i64
        lib.gpvmc_host_h5_attr_names.argtypes = [c, c, v, i64]
        lib.gpvmc_host_h5_attr_names.restype = i64
        lib.gpvmc_host_h5_attr_value.argtypes = [c, c, c, v, i64, v, i64]
        lib.gpvmc_host_h5_attr_value.restype = i64
        lib.gpvmc_host_h5_shape.argtypes = [c, c, v, ctypes.c_int32]
        lib.gpvmc_host_h5_shape.restype = ctypes.c_int32
        lib.gpvmc_host_h5_read.argtypes = [c, c, v, i64]
        lib.gpvmc_host_h5_read.restype = i64
        _lib = lib
    return _lib


def _names(fn, path, obj):
    need = fn(path.encode(), obj.encode(), None, 0)
    if need < 0:
        raise IOError("cannot list %s in %s" % (obj, path))
    buf = ctypes.create_string_buffer(int(need))
    fn(path.encode(), obj.encode(), buf, need)
    s = buf.value.decode()
    return s.split("\n") if s else []


class H5File:
    """Read-only view of an HDF5 file: group members, datasets as float64 arrays and
    attributes (numbers as arrays, strings as str)."""

    def __init__(self, path):
        self.path = str(path)
        _names(_host().gpvmc_host_h5_children, self.path, "/")

    def keys(self, group="/"):
        return _names(_host().gpvmc_host_h5_children, self.path, group)

    def read(self, dataset):
        lib = _host()
        dims = np.zeros(8, np.int64)
        rank = lib.gpvmc_host_h5_shape(self.path.encode(), dataset.encode(), dims.ctypes.data, 8)
        if rank < 0:
            raise KeyError("%s: no dataset %s" % (self.path, dataset))
        shape = tuple(int(d) for d in dims[:rank])
        out = np.zeros(max(1, int(np.prod(shape))))
        n = lib.gpvmc_host_h5_read(self.path.encode(), dataset.encode(), out.ctypes.data, out.size)
        if n < 0:
            raise IOError("%s: cannot read %s" % (self.path, dataset))
        return out[:n].reshape(shape)

    def attrs(self, obj="/"):
        lib = _host()
        out = {}
        for name in _names(lib.gpvmc_host_h5_attr_names, self.path, obj):
            num = np.zeros(4096)
            s = ctypes.create_string_buffer(4096)
            n = lib.gpvmc_host_h5_attr_value(self.path.encode(), obj.encode(), name.encode(), num.ctypes.data,
                                             num.size, s, 4096)
            if n == -2:
                out[name] = s.value.decode()
            elif n >= 0:
                out[name] = num[:n].copy()
            else:
                raise IOError("%s: cannot read attribute %s of %s" % (self.path, name, obj))
        return out


# ----------------------------------------------------------------------------- load.py
def get_attr(filename):
    """The root attributes of a file as a dict; single numbers as scalars (load.py:87-98 and
    h5py's reading of one-element attributes)"""
    a = H5File(filename).attrs("/")
    return {k: (v[0] if isinstance(v, np.ndarray) and v.size == 1 else v) for k, v in a.items()}


def argsort(seq):
    """indices of seq in increasing order, equal values in their original order (load.py:84-85)"""
    return sorted(range(len(seq)), key=seq.__getitem__)


def bunch(instat, Nsamp, indices=False):
    """The Python loader's grouping (load.py:50-82): the sorted statistics are paired
    smallest with largest into bunch n (n cycling) while more than 2 Nsamp - 1 remain, then
    each one left, largest first, joins the bunch of smallest total.  Returns the bunched
    statistics, and with indices=True also the bunched indices into instat."""
    stats = list(instat)
    args = argsort(stats)
    stats.sort()
    bunched = [[] for _ in range(Nsamp)]
    argbunched = copy.deepcopy(bunched)
    n = 0
    while len(stats) > 2 * Nsamp - 1:
        bunched[n] += [stats[0], stats[-1]]
        argbunched[n] += [args[0], args[-1]]
        stats.pop()
        stats.pop(0)
        args.pop()
        args.pop(0)
        n += 1
        if n == Nsamp:
            n = 0
    bunchedstat = [sum(b) for b in bunched]
    while len(stats) > 0:
        m = int(np.argmin(bunchedstat))
        bunched[m] += [stats[-1]]
        argbunched[m] += [args[-1]]
        bunchedstat[m] += stats[-1]
        stats.pop()
        args.pop()
    return (bunched, argbunched) if indices else bunched


def get_stat(filename, Nsamp=1):
    """(load.py:32-48) every dataset of every group of the files, in name order (an HDF5
    file's own member order), with its "statistics" (1 where absent), bunched into Nsamp
    samples: returns [(file, path)] and the bunched indices"""
    files = filename if isinstance(filename, list) else [filename]
    stats, datapath = [], []
    for f in files:
        h = H5File(f)
        for r in h.keys("/"):
            for d in h.keys("/" + r):
                p = "/{0}/{1}".format(r, d)
                a = h.attrs(p)
                stats.append(int(a["statistics"][0]) if "statistics" in a else 1)
                datapath.append((f, p))
    bunches, args = bunch(stats, Nsamp, indices=True)
    addstat = np.array([sum(bunches[i]) for i in range(Nsamp)])
    print("Average statistics of {0} (min: {1}, max: {2})".format(np.mean(addstat), np.amin(addstat),
                                                                  np.amax(addstat)))
    return datapath, args


def get_quantity(filepath, Nsamp):
    """(load.py:14-30) Nsamp samples of a quantity: each the mean over its bunch of the
    records, read as complex numbers stored as (re, im) pairs along each row"""
    files = filepath if isinstance(filepath, list) else [filepath]
    dpath, args = get_stat(files, Nsamp)
    first = H5File(dpath[args[0][0]][0]).read(dpath[args[0][0]][1])
    dshape = list(first.shape)
    dshape[1] //= 2
    Q = np.zeros([Nsamp] + dshape, dtype=complex)
    for si, bi in enumerate(args):
        for d in bi:
            dat = H5File(dpath[d][0]).read(dpath[d][1])
            Q[si, :] += dat[:, ::2] + 1j * dat[:, 1::2]
        Q[si, :] /= len(bi)
    return Q


def get_wav(filename):
    """(load.py:100-116) the channel's states from a WaveFunction file: the states_<fl>
    occupation datasets side by side, one row per state"""
    attr = get_attr(filename)
    filetype = attr.get("type")
    if filetype is None:
        mo = re.match(r".*/?[0-9]+-(.*)\.h5", filename)
        filetype = mo.groups()[0]
    if filetype != "WaveFunction":
        raise RuntimeError("get_wav: Bad file type '{}'.".format(filetype))
    h = H5File(filename)
    return np.column_stack([h.read("/" + k) for k in h.keys("/")])


# ----------------------------------------------------------------------------- proc.py
def geneigh(A, B, tol=1e-12):
    """(proc.py:11-42) A v = E B v for Hermitian A and positive semi-definite B, also when A
    and B share a null space: the eigenvalues of that null space are NaN.  B's eigenvalues
    up to tol count as null; a warning is issued when A does not vanish on B's null space."""
    if lg.matrix_rank(B, tol) == np.shape(B)[0]:
        return eigh(A, B)
    Be, Bv = eigh(B)
    At = np.dot(np.conj(Bv.T), np.dot(A, Bv))
    Bt = np.diag(Be)
    try:
        idx = next(i for i, v in enumerate(Be) if v > tol)
    except StopIteration:
        raise RuntimeError("geneigh: Rank of B < B.shape[0] but null-space could not be found!")
    m = np.amax(abs(At[0:idx, :].flatten()))
    if m > tol:
        warnings.warn("Maximum non-diagonal element in A written in B null-space is bigger than the tolerance '"
                      + str(tol) + "'.", UserWarning)
    Et, Vt = eigh(At[idx:, idx:], Bt[idx:, idx:])
    Ut = np.zeros(np.shape(A), A.dtype)
    Ut[0:idx, 0:idx] = np.eye(idx)
    Ut[idx:, idx:] = Vt
    U = np.dot(Bv, Ut)
    E = np.concatenate((float("NaN") * np.ones(idx), Et))
    return E, U


def fermisigns(states, params):
    """(proc.py:44-72) the fermionic sign of each state's occupations relative to the
    mean-field reference state: scanning the orbitals in order, each time both a hole (in
    the reference, empty in the state) and a particle are pending, the last pending of each
    are paired, the sign takes (-1)^(occupied orbitals strictly between them) and the
    reference moves the particle from the hole to the particle's orbital"""
    if params.get("sfpnxphz_wav"):
        raise RuntimeError("not implemented: the sfpnxphz wavefunction")
    if not params.get("stagflux_wav"):
        raise RuntimeError("not implemented")
    ref0 = refstate(params)
    fs = np.ones(np.shape(states)[0])
    for s in range(np.shape(states)[0]):
        hole, part = [], []
        ref = copy.deepcopy(ref0)
        for i in range(np.shape(states)[1]):
            if ref[i] == 0 and states[s, i] != 0:
                part.append(i)
            if ref[i] != 0 and states[s, i] == 0:
                hole.append(i)
            if part and hole:
                h = hole.pop()
                p = part.pop()
                fs[s] *= (-1) ** (np.sum(ref[min(h, p) + 1:max(h, p)]))
                ref[h] = 0
                ref[p] = 1
        if hole or part:
            raise RuntimeError("fermisigns: particle number not conserved?")
    return fs


def gaussians(x, x0, A, sig):
    """(proc.py:74-83) sum of Gaussians of weights A, centres x0 and widths sig on the grid x
    (NaN entries masked)"""
    x0m = ma.masked_invalid(np.atleast_1d(x0))
    Am = ma.masked_invalid(np.atleast_1d(A))
    sigm = ma.masked_invalid(np.atleast_1d(sig))
    amp = Am * np.sqrt(0.5 / np.pi) / sigm
    X, X0 = np.meshgrid(x, x0m)
    return np.einsum("...i,...ij->...j", amp,
                     np.exp(-0.5 * (X - X0) ** 2 / np.tile(sigm ** 2, (np.shape(x)[0], 1)).T))


# ----------------------------------------------------------------------------- recspace.py
def fold(k):
    return k - ((k < 0) + (k >= 1)) * np.floor(k)


def inmbz(kx, ky):
    fkx = fold(kx)
    fky = fold(ky)
    tol = 1e-6
    return (fky <= -0.5 + fkx + tol) + (fky < 0.5 - fkx - tol) + (fky >= 1.5 - fkx - tol) + (fky > 0.5 + fkx + tol)


def mbzmod(kx, ky):
    mkx = fold(kx)
    mky = fold(ky)
    mmkx = fold(mkx + 0.5 * (1 - inmbz(mkx, mky)))
    mmky = fold(mky + 0.5 * (1 - inmbz(mkx, mky)))
    return mmkx, mmky


def mbz(params):
    """(recspace.py:3-16) the momenta of the magnetic zone in units of 2 pi, kx outer, ky
    inner, shifted by half the boundary phases"""
    Lx = int(params["L"])
    Ly = int(params["L"])
    shift = [params["phase_shift_x"] / 2.0, params["phase_shift_y"] / 2.0]
    fskx = np.zeros(Lx * Ly // 2)
    fsky = np.zeros(Lx * Ly // 2)
    i = 0
    for kx in range(Lx):
        for ky in range(Ly):
            if inmbz(float(kx + shift[0]) / Lx, float(ky + shift[1]) / Ly):
                fskx[i] = float(kx + shift[0]) / Lx
                fsky[i] = float(ky + shift[1]) / Ly
                i += 1
    return fskx, fsky


# ----------------------------------------------------------------------------- stagflux.py
def delta(kx, ky, params):
    return 0.5 * (np.exp(1j * params["phi"] * np.pi) * np.cos(kx * 2 * np.pi)
                  + np.exp(-1j * params["phi"] * np.pi) * np.cos(ky * 2 * np.pi))


def neelk(kx, ky, params):
    if params.setdefault("neel_exp", 0.0) == 0.0:
        return params["neel"]
    return params["neel"] * np.power(0.5 * (np.cos(kx * 2 * np.pi) ** 2 + np.cos(ky * 2 * np.pi) ** 2),
                                     params["neel_exp"])


def mfham(kx, ky, params):
    """the 4 x 4 mean-field Hamiltonian per momentum, basis (A up, B up, A down, B down)"""
    H = np.zeros([4, 4] + list(np.shape(kx)), dtype=complex)
    nk, dk = neelk(kx, ky, params), delta(kx, ky, params)
    H[0, 0, :] = -nk
    H[0, 1, :] = -np.conj(dk)
    H[1, 0, :] = -dk
    H[1, 1, :] = nk
    H[2, 2, :] = nk
    H[2, 3, :] = -np.conj(dk)
    H[3, 2, :] = -dk
    H[3, 3, :] = -nk
    return H


def evals(kx, ky, params):
    v = np.zeros([4] + list(np.shape(kx)))
    e = np.sqrt(neelk(kx, ky, params) ** 2 + abs(delta(kx, ky, params)) ** 2)
    v[0, :], v[1, :], v[2, :], v[3, :] = -e, -e, e, e
    return v


def evecs(kx, ky, params):
    """columns: spin-up band -1, spin-down band -1, spin-up band +1, spin-down band +1"""
    v = np.zeros([4, 4] + list(np.shape(kx)), dtype=complex)
    ev = evals(kx, ky, params)
    pdk = np.exp(1j * np.angle(delta(kx, ky, params)))
    nk = neelk(kx, ky, params)
    v[0, 0, :] = np.sqrt(0.5 * (1.0 - nk / ev[0, :]))
    v[1, 0, :] = pdk * np.sqrt(0.5 * (1.0 + nk / ev[0, :]))
    v[2, 1, :] = np.sqrt(0.5 * (1.0 + nk / ev[1, :]))
    v[3, 1, :] = pdk * np.sqrt(0.5 * (1.0 - nk / ev[1, :]))
    v[0, 2, :] = -np.conj(pdk) * np.sqrt(0.5 * (1.0 - nk / ev[2, :]))
    v[1, 2, :] = np.sqrt(0.5 * (1.0 + nk / ev[2, :]))
    v[2, 3, :] = -np.conj(pdk) * np.sqrt(0.5 * (1.0 + nk / ev[3, :]))
    v[3, 3, :] = np.sqrt(0.5 * (1.0 - nk / ev[3, :]))
    return v


def spinops(params):
    """(stagflux.py:62-80) the mean-field amplitudes of S^x_q (= S^y_q) and S^z_q on the
    channel's particle-hole states, one entry per magnetic-zone momentum k (S^z: the
    spin-up then the spin-down excitation of k, interleaved).  The long channel at q = 0 or
    (L/2, L/2), whose states include the ground state, is refused (the reference's arrays
    do not fit there either)."""
    kx, ky = mbz(params)
    L = params["L"]
    if (params["qx"], params["qy"]) in [(0, 0), (L / 2, L / 2)] and params["channel"] == "long":
        raise ValueError("spinops: the long channel at q = 0 or (L/2, L/2) is not supported")
    kqx = kx - params["qx"] / params["L"]
    kqy = ky - params["qy"] / params["L"]
    evk = evecs(kx, ky, params)
    evq = evecs(kqx, kqy, params)
    nexc = int(L) ** 2
    so = [None, None, np.zeros(nexc, dtype=complex)]
    so[0] = np.conj(evk[0, 2, :]) * evq[2, 1, :] + np.conj(evk[1, 2, :]) * evq[3, 1, :]
    so[1] = so[0]
    so[2][::2] = 0.5 * (np.conj(evk[0, 2, :]) * evq[0, 0, :] + np.conj(evk[1, 2, :]) * evq[1, 0, :])
    so[2][1::2] = -0.5 * (np.conj(evk[2, 3, :]) * evq[2, 1, :] + np.conj(evk[3, 3, :]) * evq[3, 1, :])
    return so


def refstate(params):
    """the mean-field ground state's orbital occupations: every band -1 orbital"""
    state = np.zeros(2 * int(params["L"]) ** 2)
    state[::2] = 1
    return state


# ----------------------------------------------------------------------------- hl.py
def get_stat_spin_struct(filenames, nsamp):
    """(hl.py:9-66) the static structure factors from StatSpinStruct records: returns
    (Sqxx, Sqyy, Sqzz), (Srxx, Sryy, Srzz), each [nsamp][L][L].  Records of N^2 site pairs
    (rows zz, +-, -+, ++, --; pair i N + j) are summed over the pairs of each displacement
    r_i - r_j (wrapped into [-L/2, L/2), array index (dy + L/2, dx + L/2)) and divided by N;
    S^xx = (+- + -+ + ++ + --) / 4, S^yy = (+- + -+ - ++ - --) / 4; S(q) is N times the
    inverse FFT of the displacement array shifted to the origin, so that
    S(q)[a, b] = (1/N) sum_ij C_ij exp(2 pi i (a dy + b dx) / L).  Records of N entries are
    the older storage of S(q) itself."""
    if not isinstance(filenames, list):
        filenames = [filenames]
    Sq = get_quantity(filenames, nsamp)
    params = get_attr(filenames[0])
    Lx = int(params["L"])
    Ly = int(params["L"])
    hLx, hLy = Lx // 2, Ly // 2
    N = Lx * Ly
    if Sq.shape[2] == N:
        Sqxx = np.reshape(0.25 * (Sq[:, 1, :] + Sq[:, 2, :] + Sq[:, 3, :] + Sq[:, 4, :]), (Sq.shape[0], Lx, Ly))
        Sqyy = np.reshape(0.25 * (Sq[:, 1, :] + Sq[:, 2, :] - Sq[:, 3, :] - Sq[:, 4, :]), (Sq.shape[0], Lx, Ly))
        Sqzz = np.reshape(Sq[:, 0, :], (Sq.shape[0], Lx, Ly))
        Srxx = fft.fftshift(fft.fft2(Sqxx, axes=(1, 2)), axes=(1, 2)) / N
        Sryy = fft.fftshift(fft.fft2(Sqyy, axes=(1, 2)), axes=(1, 2)) / N
        Srzz = fft.fftshift(fft.fft2(Sqzz, axes=(1, 2)), axes=(1, 2)) / N
        return (Sqxx, Sqyy, Sqzz), (Srxx, Sryy, Srzz)
    rx, ry = np.meshgrid(np.arange(Lx, dtype=int), np.arange(Ly, dtype=int))
    rx = rx.ravel()
    ry = ry.ravel()
    rix, rjx = np.meshgrid(rx, rx)
    riy, rjy = np.meshgrid(ry, ry)
    rijx = rjx - rix
    rijy = rjy - riy
    rijx[rijx >= hLx] -= Lx
    rijx[rijx < -hLx] += Lx
    rijy[rijy >= hLy] -= Ly
    rijy[rijy < -hLy] += Ly
    cell = ((rijy + hLy) * Lx + rijx + hLx).ravel()
    Sr = np.zeros((Sq.shape[0], 5, N), dtype=complex)
    for t in range(N):
        Sr[:, :, t] = np.sum(Sq[:, :, cell == t], axis=2) / N
    Srxx = np.reshape(0.25 * np.sum(Sr[:, 1:, :], axis=1), (Sq.shape[0], Lx, Ly))
    Sryy = np.reshape(0.25 * (np.sum(Sr[:, 1:3, :], axis=1) - np.sum(Sr[:, 3:, :], axis=1)), (Sq.shape[0], Lx, Ly))
    Srzz = np.reshape(Sr[:, 0, :], (Sq.shape[0], Lx, Ly))
    Sqxx = fft.ifft2(fft.fftshift(Srxx, axes=(1, 2)), axes=(1, 2)) * N
    Sqyy = fft.ifft2(fft.fftshift(Sryy, axes=(1, 2)), axes=(1, 2)) * N
    Sqzz = fft.ifft2(fft.fftshift(Srzz, axes=(1, 2)), axes=(1, 2)) * N
    return (Sqxx, Sqyy, Sqzz), (Srxx, Sryy, Srzz)


def get_eig_sys(filenames, nsamp, wav=None, statstruct=None, tol=1e-12):
    """(hl.py:68-116) per sample, the channel's H and O matrices from ProjHeis records (H the
    upper, O the lower half), conjugated by the states' fermionic signs, and the generalised
    eigenproblem H v = E O v (geneigh).  With statstruct (the first return value of
    get_stat_spin_struct) O is rescaled so that the channel's S(q) sum rule holds:
    O *= 2 S(q) / (s^dagger O s), s the spinops amplitudes of the channel, S(q) the entry
    (qx, qy) of the first sample's S(q) raveled in its [a][b] order, sample by sample (the
    reference divides each sample by the vector of every sample's s^dagger O s, which
    broadcasts only for one sample).  Returns H, O, the eigenvalues [nsamp][nk] and
    eigenvectors [nsamp][nk][nk]."""
    if not isinstance(filenames, list):
        filenames = [filenames]
    HO = get_quantity(filenames, nsamp)
    half = HO.shape[1] // 2
    H = HO[:, :half, :]
    O = HO[:, half:, :]
    if not wav:
        mo = re.search(r"(.*/?[0-9]+)-ProjHeis\.h5", filenames[0])
        wav = mo.groups()[0] + "-WaveFunction.h5"
    wav_st = get_wav(wav)
    params = get_attr(filenames[0])
    fs = np.diag(fermisigns(wav_st, params))
    for k in range(nsamp):
        H[k, :, :] = np.dot(fs, np.dot(H[k, :, :], fs))
        O[k, :, :] = np.dot(fs, np.dot(O[k, :, :], fs))
    ev = np.zeros(H.shape[:2])
    vecs = np.zeros(H.shape, dtype=complex)
    for s in range(nsamp):
        ev[s, :], vecs[s, :, :] = geneigh(H[s, :, :], O[s, :, :], tol)
    if statstruct:
        qs = [(qx, qy) for qx in range(int(params["L"])) for qy in range(int(params["L"]))]
        idx = qs.index((int(params["qx"]), int(params["qy"])))
        ssqx = statstruct[0]
        ssqz = statstruct[2]
        if len(ssqx.shape) != 2:
            ssqx = ssqx[0, ...].ravel()
            ssqz = ssqz[0, ...].ravel()
        if params["stagflux_wav"]:
            pkq = spinops(params)
            if params["channel"] == "trans":
                sqpm = np.einsum("i,sij,j->s", np.conj(pkq[1]), O, pkq[1])
                for s in range(nsamp):
                    O[s, :, :] *= ssqx[idx] * 2 / sqpm[s]
            elif params["channel"] == "long":
                sqzz = np.einsum("i,sij,j->s", np.conj(pkq[2]), O, pkq[2])
                for s in range(nsamp):
                    O[s, :, :] *= ssqz[idx] * 2 / sqzz[s]
            else:
                raise RuntimeError("Channel not undertood")
        else:
            raise RuntimeError("Not implemented")
    return H, O, ev, vecs


def get_sq_ampl(filenames, nsamp, wav=None, tol=1e-12, O=None, evecs=None):
    """(hl.py:118-145) the spectral weights |s^dagger O v_n|^2 of the channel's eigenstates
    v_n (s: spinops of the channel), [nsamp][nk]"""
    if not isinstance(filenames, list):
        filenames = [filenames]
    if O is None or evecs is None:
        _, O, _, evecs = get_eig_sys(filenames, nsamp, wav, tol=tol)
    params = get_attr(filenames[0])
    if params.get("sfpnxphz_wav"):
        raise RuntimeError("not implemented: the sfpnxphz wavefunction")
    if not params.get("stagflux_wav"):
        raise RuntimeError("sfpnzphx_wav not implemented yet")
    pkq = spinops(params)
    if params["channel"] == "trans":
        return abs(np.einsum("i,sij,sjk->sk", np.conj(pkq[1]), O, evecs)) ** 2
    if params["channel"] == "long":
        return abs(np.einsum("i,sij,sjk->sk", np.conj(pkq[2]), O, evecs)) ** 2
    raise RuntimeError("channel not understood.")
