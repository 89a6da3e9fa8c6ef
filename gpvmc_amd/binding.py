"""ctypes binding of libgpvmc_hip.so (C ABI: include/gpvmc.h)."""
import ctypes
import os

import numpy as np

FP64 = 0
MIXED = 1
MEAS_PROJHEIS = 1
MEAS_STAGMAGNZ = 2
MEAS_OVERLAP = 4
MEAS_STATSPIN = 8
SWEEP_DEFAULT = 0
SWEEP_ONE_MOVE = 1
SWEEP_WINDOWS = 2
ACC_STRIDE = 12

_PKG = os.path.dirname(os.path.abspath(__file__))
_LIB = None

# exported entry points and their ctypes signatures (every one returns int status
# except gpvmc_last_error / gpvmc_version)
_vp = ctypes.c_void_p
SIGNATURES = {
    "gpvmc_last_error": (ctypes.c_char_p, []),
    "gpvmc_version": (ctypes.c_int, []),
    "gpvmc_create": (ctypes.c_int, [_vp, ctypes.POINTER(_vp)]),
    "gpvmc_destroy": (ctypes.c_int, [_vp]),
    "gpvmc_set_orbitals": (ctypes.c_int, [_vp, ctypes.c_int32, _vp]),
    "gpvmc_set_groups": (ctypes.c_int, [_vp, _vp]),
    "gpvmc_init_walkers": (ctypes.c_int, [_vp, _vp, _vp]),
    "gpvmc_init_random": (ctypes.c_int, [_vp, ctypes.c_uint64, ctypes.c_int32, _vp, _vp]),
    "gpvmc_refresh": (ctypes.c_int, [_vp, _vp]),
    "gpvmc_set_sweep_variant": (ctypes.c_int, [_vp, ctypes.c_int32]),
    "gpvmc_sweep_variant": (ctypes.c_int, [_vp]),
    "gpvmc_sweep_shapes": (ctypes.c_int, [ctypes.c_int32, ctypes.c_int32, _vp, ctypes.c_int32]),
    "gpvmc_walk": (ctypes.c_int, [_vp, ctypes.c_int64, ctypes.c_int64, ctypes.c_uint32,
                                  _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "gpvmc_measure": (ctypes.c_int, [_vp, ctypes.c_uint32, _vp, _vp, _vp]),
    "gpvmc_read_accumulators": (ctypes.c_int, [_vp, _vp]),
    "gpvmc_reset_accumulators": (ctypes.c_int, [_vp]),
    "gpvmc_read_statspin": (ctypes.c_int, [_vp, ctypes.c_int32, ctypes.c_int32, _vp, _vp]),
    "gpvmc_read_statspin_groups": (ctypes.c_int, [_vp, _vp, _vp]),
    "gpvmc_read_state": (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp]),
    "gpvmc_read_inverse": (ctypes.c_int, [_vp, ctypes.c_int32, _vp]),
    "gpvmc_read_ratio_table": (ctypes.c_int, [_vp, ctypes.c_int32, _vp]),
    "gpvmc_create_channel": (ctypes.c_int, [_vp, _vp, ctypes.POINTER(_vp)]),
    "gpvmc_create_channels": (ctypes.c_int, [_vp, ctypes.c_int32, _vp, ctypes.POINTER(_vp)]),
    "gpvmc_read_amplitudes": (ctypes.c_int, [_vp, _vp]),
    "gpvmc_channel_kernel": (ctypes.c_int, [_vp, _vp]),
    "gpvmc_batch_create": (ctypes.c_int, [ctypes.c_int32, ctypes.POINTER(_vp)]),
    "gpvmc_batch_add": (ctypes.c_int, [_vp, _vp, ctypes.c_int64, ctypes.c_int64, ctypes.c_uint32]),
    "gpvmc_batch_sweep": (ctypes.c_int, [_vp, ctypes.c_int32, _vp]),
    "gpvmc_batch_destroy": (ctypes.c_int, [_vp]),
    "gpvmc_set_orbital_table": (ctypes.c_int, [_vp, ctypes.c_int32, _vp]),
    "gpvmc_read_channel": (ctypes.c_int, [_vp, ctypes.c_int32, ctypes.c_int32, _vp, _vp, _vp]),
    "gpvmc_kernel_time": (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp]),
    "gpvmc_det_inv": (ctypes.c_int, [ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, _vp, _vp, _vp, _vp, _vp,
                                     _vp]),
    "gpvmc_det_inv_workspace_bytes": (ctypes.c_int, [ctypes.c_int32, ctypes.POINTER(ctypes.c_int64)]),
    "gpvmc_det_inv_noalloc": (ctypes.c_int, [ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, _vp, _vp, _vp, _vp,
                                             _vp, _vp]),
    "gpvmc_reset_kernel_time": (ctypes.c_int, [_vp]),
}


class GpvmcError(RuntimeError):
    pass


class _Config(ctypes.Structure):
    _fields_ = [("L", ctypes.c_int32), ("n_walkers", ctypes.c_int32),
                ("n_groups", ctypes.c_int32), ("precision", ctypes.c_int32),
                ("device", ctypes.c_int32), ("lanes_per_row", ctypes.c_int32),
                ("seed", ctypes.c_uint64)]


class _Channel(ctypes.Structure):
    _fields_ = [("n_up", ctypes.c_int32), ("n_dn", ctypes.c_int32), ("nk", ctypes.c_int32),
                ("state_orb", ctypes.c_void_p), ("hop_n", ctypes.c_void_p),
                ("hop_orb", ctypes.c_void_p), ("hopsign", ctypes.c_void_p)]


def library_path():
    return os.path.join(_PKG, "lib", "libgpvmc_hip.so")


def load_library():
    """Loads the in-tree libgpvmc_hip.so; raises if it is not built."""
    global _LIB
    if _LIB is None:
        path = library_path()
        if not os.path.exists(path):
            raise GpvmcError("libgpvmc_hip.so is not built (run __graft_entry__.build()); "
                             "there is no CPU fallback")
        lib = ctypes.CDLL(path)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = lib
    return _LIB


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def sweep_shapes(L, precision):
    """lanes_per_row of every ground-state sweep shape compiled for (L, precision), the
    default first (no device needed)."""
    lib = load_library()
    n = lib.gpvmc_sweep_shapes(precision, L, None, 0)
    if n < 0:
        raise GpvmcError(lib.gpvmc_last_error().decode())
    out = np.zeros(max(n, 1), np.int32)
    lib.gpvmc_sweep_shapes(precision, L, _ptr(out), n)
    return [int(x) for x in out[:n]]


def _check_rc(lib, rc):
    if rc != 0:
        raise GpvmcError("gpvmc error %d: %s" % (rc, lib.gpvmc_last_error().decode()))


def det_inv_workspace_bytes(n):
    """bytes of device workspace gpvmc_det_inv uses per n x n matrix (0: none)"""
    lib = load_library()
    b = ctypes.c_int64()
    _check_rc(lib, lib.gpvmc_det_inv_workspace_bytes(n, ctypes.byref(b)))
    return b.value


def det_inv(n, batch, a_ptr, ai_ptr, logabs_ptr, phase_ptr, singular_ptr, stream=None, device=0):
    """Batched linalg::DetInv on device memory (gpvmc_det_inv): every argument after
    `batch` is a device address (int); `stream` a hipStream_t handle or None."""
    lib = load_library()
    _check_rc(lib, lib.gpvmc_det_inv(device, n, batch, a_ptr, ai_ptr, logabs_ptr, phase_ptr, singular_ptr,
                                     stream))


def det_inv_noalloc(n, batch, a_ptr, ai_ptr, logabs_ptr, phase_ptr, singular_ptr, stream=None, device=0):
    """gpvmc_det_inv with its workspace in the library's fixed device array (no allocation)"""
    lib = load_library()
    _check_rc(lib, lib.gpvmc_det_inv_noalloc(device, n, batch, a_ptr, ai_ptr, logabs_ptr, phase_ptr,
                                             singular_ptr, stream))


class Batch:
    """A batch of independent walkers on one device (one gpvmc_ctx)."""

    def __init__(self, L, n_walkers, precision=FP64, n_groups=1, seed=0, device=0,
                 lanes_per_row=0):
        self.lib = load_library()
        self.L, self.Ns, self.N = L, L * L, L * L // 2
        self.W, self.G = n_walkers, n_groups
        self.precision = precision
        cfg = _Config(L, n_walkers, n_groups, precision, device, lanes_per_row, seed)
        h = ctypes.c_void_p()
        self._check(self.lib.gpvmc_create(ctypes.byref(cfg), ctypes.byref(h)))
        self.h = h

    def _check(self, rc):
        if rc != 0:
            raise GpvmcError("gpvmc error %d: %s" % (rc, self.lib.gpvmc_last_error().decode()))

    def close(self):
        if getattr(self, "h", None):
            self._check(self.lib.gpvmc_destroy(self.h))
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_orbitals(self, group, occ):
        occ = np.ascontiguousarray(occ, np.complex128)
        assert occ.shape == (2, self.Ns, self.N)
        self._check(self.lib.gpvmc_set_orbitals(self.h, group, _ptr(occ)))

    def set_groups(self, groups):
        g = np.ascontiguousarray(groups, np.int32)
        assert g.shape == (self.W,)
        self._check(self.lib.gpvmc_set_groups(self.h, _ptr(g)))

    def init(self, spins):
        s = np.ascontiguousarray(spins, np.int32).reshape(self.W, self.Ns)
        ok = np.zeros(self.W, np.int32)
        self._check(self.lib.gpvmc_init_walkers(self.h, _ptr(s), _ptr(ok)))
        return ok.astype(bool)

    def init_random(self, seed, max_attempts=10):
        """the random start with redraws on the device (gpvmc_init_random): (ok, attempts)"""
        ok = np.zeros(self.W, np.int32)
        att = np.zeros(self.W, np.int32)
        self._check(self.lib.gpvmc_init_random(self.h, seed, max_attempts, _ptr(ok), _ptr(att)))
        return ok.astype(bool), att

    def refresh(self, drift=False):
        d = np.zeros(self.W) if drift else None
        self._check(self.lib.gpvmc_refresh(self.h, _ptr(d)))
        return d

    @property
    def sweep_variant(self):
        """the kernel that runs: SWEEP_ONE_MOVE or SWEEP_WINDOWS (delayed-update windows)"""
        return self.lib.gpvmc_sweep_variant(self.h)

    @sweep_variant.setter
    def sweep_variant(self, v):
        self._check(self.lib.gpvmc_set_sweep_variant(self.h, int(v)))

    def walk(self, n_steps, meas=0, mask=MEAS_PROJHEIS | MEAS_STAGMAGNZ, uniforms=None,
             trace=False):
        u = None
        if uniforms is not None:
            u = np.ascontiguousarray(uniforms, np.float64).reshape(self.W, n_steps, 4)
        out = {}
        if trace:
            nm = 0 if meas <= 0 else (n_steps if meas == 1 else (n_steps + meas - 1) // meas)
            out = dict(noop=np.zeros((self.W, n_steps), np.int8),
                       accept=np.zeros((self.W, n_steps), np.int8),
                       ratio=np.zeros((self.W, n_steps)),
                       eloc=np.zeros((self.W, max(nm, 1)), np.complex128),
                       mz=np.zeros((self.W, max(nm, 1)), np.complex128),
                       logw=np.zeros((self.W, max(nm, 1))))
        g = out.get
        self._check(self.lib.gpvmc_walk(self.h, n_steps, meas, mask, _ptr(u), _ptr(g("noop")),
                                        _ptr(g("accept")), _ptr(g("ratio")), _ptr(g("eloc")),
                                        _ptr(g("mz")), _ptr(g("logw"))))
        return out

    def measure(self, mask=MEAS_PROJHEIS | MEAS_STAGMAGNZ):
        """Estimators at the current configurations (no step, accumulators untouched)."""
        e = np.zeros(self.W, np.complex128)
        m = np.zeros(self.W, np.complex128)
        lw = np.zeros(self.W)
        self._check(self.lib.gpvmc_measure(self.h, mask, _ptr(e), _ptr(m), _ptr(lw)))
        return dict(eloc=e, mz=m, logw=lw)

    def accumulators(self):
        a = np.zeros((self.W, ACC_STRIDE))
        self._check(self.lib.gpvmc_read_accumulators(self.h, _ptr(a)))
        return a

    def reset_accumulators(self):
        self._check(self.lib.gpvmc_reset_accumulators(self.h))

    def state(self):
        spins = np.zeros((self.W, self.Ns), np.int32)
        part = np.zeros((self.W, 2, self.N), np.int32)
        la = np.zeros(self.W)
        ph = np.zeros(self.W)
        self._check(self.lib.gpvmc_read_state(self.h, _ptr(spins), _ptr(part), _ptr(la), _ptr(ph)))
        return spins, part, la, ph

    def statspin(self, w0=0, nw=None):
        """StatSpinStruct sums of walkers w0 .. w0+nw-1: (acc[nw][5][Ns][Ns] complex, with
        rows szz, spm, smp, spp, smm; n_meas[nw])."""
        nw = self.W - w0 if nw is None else nw
        acc = np.zeros((nw, 5, self.Ns, self.Ns), np.complex128)
        n = np.zeros(nw)
        self._check(self.lib.gpvmc_read_statspin(self.h, w0, nw, _ptr(acc), _ptr(n)))
        return acc, n

    def statspin_groups(self):
        """StatSpinStruct sums added over the walkers of each parameter group."""
        acc = np.zeros((self.G, 5, self.Ns, self.Ns), np.complex128)
        n = np.zeros(self.G)
        self._check(self.lib.gpvmc_read_statspin_groups(self.h, _ptr(acc), _ptr(n)))
        return acc, n

    def inverse(self, w):
        ai = np.zeros((2, self.N, self.N), np.complex128)
        self._check(self.lib.gpvmc_read_inverse(self.h, w, _ptr(ai)))
        return ai

    def ratio_table(self, w):
        g = np.zeros((2, self.N, self.N), np.complex128)
        self._check(self.lib.gpvmc_read_ratio_table(self.h, w, _ptr(g)))
        return g

    def kernel_time(self):
        sm = ctypes.c_double()
        sl = ctypes.c_int64()
        rm = ctypes.c_double()
        rl = ctypes.c_int64()
        self._check(self.lib.gpvmc_kernel_time(self.h, ctypes.byref(sm), ctypes.byref(sl),
                                               ctypes.byref(rm), ctypes.byref(rl)))
        return dict(sweep_ms=sm.value, sweep_launches=sl.value, refresh_ms=rm.value,
                    refresh_launches=rl.value)

    def reset_kernel_time(self):
        self._check(self.lib.gpvmc_reset_kernel_time(self.h))


class ChannelBatch(Batch):
    """A batch of walkers on channels (gpvmc_amd.channels.Channel): the weight is
    sum_k |a_k|^2 over the channel's states, ProjHeis accumulates H and O matrices.
    `channel` is one Channel or a list of Channels of equal sizes (walker w walks on
    channel w mod len(list))."""

    def __init__(self, L, n_walkers, channel, precision=FP64, n_groups=1, seed=0, device=0,
                 kernel="auto"):
        """kernel: "auto" (register-resident where the walker fits) or "streamed"."""
        self.lib = load_library()
        self.L, self.Ns = L, L * L
        if kernel not in ("auto", "streamed"):
            raise ValueError("kernel must be 'auto' or 'streamed'")
        chans = list(channel) if isinstance(channel, (list, tuple)) else [channel]
        self.channels = chans
        self.ch = chans[0]
        self.n_up, self.n_dn = self.ch.npt
        self.N = max(self.ch.npt)
        self.nk = self.ch.nexc
        self.W, self.G = n_walkers, n_groups
        self.precision = precision
        nk = self.nk
        self._keep = []
        structs = (_Channel * len(chans))()
        for q, ch in enumerate(chans):
            if ch.npt != self.ch.npt or ch.nexc != nk:
                raise ValueError("channels of one batch must have equal sizes")
            state_orb = np.ascontiguousarray(
                np.array([np.concatenate([u, d]) for u, d in ch.states], np.int32))
            hop_n = np.zeros((nk, nk, 2), np.int32)
            hop_orb = np.zeros((nk, nk, 2, 2, 2), np.int32)
            for c in range(nk):
                for k in range(nk):
                    for fl in range(2):
                        path = ch.hops[c][k][fl]
                        hop_n[c, k, fl] = len(path)
                        for h, (a, b) in enumerate(path):
                            hop_orb[c, k, fl, h] = (a, b)
            hopsign = np.ascontiguousarray(ch.hopsign.astype(np.int32))
            self._keep += [state_orb, hop_n, hop_orb, hopsign]
            structs[q] = _Channel(self.n_up, self.n_dn, nk, _ptr(state_orb), _ptr(hop_n),
                                  _ptr(hop_orb), _ptr(hopsign))
        cfg = _Config(L, n_walkers, n_groups, precision, device, 0 if kernel == "auto" else -1, seed)
        h = ctypes.c_void_p()
        self._check(self.lib.gpvmc_create_channels(ctypes.byref(cfg), len(chans),
                                                   ctypes.cast(structs, ctypes.c_void_p),
                                                   ctypes.byref(h)))
        self.h = h

    @property
    def register_resident(self):
        v = ctypes.c_int32()
        self._check(self.lib.gpvmc_channel_kernel(self.h, ctypes.byref(v)))
        return bool(v.value)

    def amplitudes(self):
        """ln|a_k| + i arg a_k of every state at each walker's configuration, [W][nk] complex"""
        out = np.zeros((self.W, self.nk, 2))
        self._check(self.lib.gpvmc_read_amplitudes(self.h, _ptr(out)))
        return out[..., 0] + 1j * out[..., 1]

    def set_orbital_table(self, group, table):
        t = np.ascontiguousarray(table, np.complex128)
        assert t.shape == (2, self.Ns, self.Ns)
        self._check(self.lib.gpvmc_set_orbital_table(self.h, group, _ptr(t)))

    def walk(self, n_steps, meas=0, mask=MEAS_PROJHEIS, uniforms=None, trace=False):
        u = None
        if uniforms is not None:
            u = np.ascontiguousarray(uniforms, np.float64).reshape(self.W, n_steps, 4)
        out = {}
        if trace:
            nm = 0 if meas <= 0 else (n_steps if meas == 1 else (n_steps + meas - 1) // meas)
            out = dict(noop=np.zeros((self.W, n_steps), np.int8),
                       accept=np.zeros((self.W, n_steps), np.int8),
                       ratio=np.zeros((self.W, n_steps)),
                       logw=np.zeros((self.W, max(nm, 1))))
        g = out.get
        self._check(self.lib.gpvmc_walk(self.h, n_steps, meas, mask, _ptr(u), _ptr(g("noop")),
                                        _ptr(g("accept")), _ptr(g("ratio")), None, None,
                                        _ptr(g("logw"))))
        return out

    def measure(self, mask=MEAS_PROJHEIS):
        lw = np.zeros(self.W)
        self._check(self.lib.gpvmc_measure(self.h, mask, None, None, _ptr(lw)))
        return dict(logw=lw)

    def channel_sums(self, w0=0, nw=None):
        """(ho[nw][2 nk][nk] complex, current state[nw], weight sum_k |a_k / a_c|^2 [nw])"""
        nw = self.W - w0 if nw is None else nw
        ho = np.zeros((nw, 2 * self.nk, self.nk), np.complex128)
        cs = np.zeros(nw, np.int32)
        wt = np.zeros(nw)
        self._check(self.lib.gpvmc_read_channel(self.h, w0, nw, _ptr(ho), _ptr(cs), _ptr(wt)))
        return ho, cs, wt


class Sweeper:
    """Contexts (Batch or ChannelBatch, any sizes) swept together through one captured
    hipGraph per sweep (gpvmc_batch_*): add(batch, n_steps, meas, mask), then sweep(n)."""

    def __init__(self, device=0):
        self.lib = load_library()
        self.members = []
        h = ctypes.c_void_p()
        rc = self.lib.gpvmc_batch_create(device, ctypes.byref(h))
        if rc != 0:
            raise GpvmcError(self.lib.gpvmc_last_error().decode())
        self.h = h

    def _check(self, rc):
        if rc != 0:
            raise GpvmcError(self.lib.gpvmc_last_error().decode())

    def add(self, batch, n_steps, meas=0, mask=MEAS_PROJHEIS):
        self._check(self.lib.gpvmc_batch_add(self.h, batch.h, n_steps, meas, mask))
        self.members.append(batch)

    def sweep(self, n=1):
        """runs n sweeps of every member; returns their device time in ms"""
        ms = ctypes.c_double()
        self._check(self.lib.gpvmc_batch_sweep(self.h, n, ctypes.byref(ms)))
        return ms.value

    def close(self):
        if self.h is not None:
            self.lib.gpvmc_batch_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
