"""Starting and resuming chains: the device's random start (vmc.cpp:221-242) and resuming
from a FockState file (vmc.cpp:178-220, State.cpp:117-166).

* A chain stopped after n1 steps, its configurations saved to a FockState file by the host
  (SaveFockState), read back (the --spinstate reader) and started again (FockInit labels,
  FP64 LU) continues bit for bit in FP64: on the uniforms the uninterrupted chain used for
  its steps n1 .. n1 + n2 (replayed), the same no-op/accept sequence and the same final
  configurations, and weight ratios within 1e-10 relative (1e-14 absolute: labels differ,
  ratios do not).
* vmc_host --spinstate=<file> starts every chain from the file: with no steps taken, the
  FockState it writes at the end equals the one it read, for one chain (/rank-0) and for a
  batch (/rank-1 .. /rank-W).
"""
import ctypes
import os
import subprocess

import numpy as np
import pytest

from gpvmc_amd import binding, build, channels, model

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def host_lib():
    lib = ctypes.CDLL(build.HOST_LIB)
    lib.gpvmc_host_fockstate_write.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int32, ctypes.c_int32,
                                               ctypes.c_void_p]
    lib.gpvmc_host_fockstate_read.argtypes = [ctypes.c_char_p, ctypes.c_int32, ctypes.c_int32, ctypes.c_void_p]
    lib.gpvmc_host_last_error.restype = ctypes.c_char_p
    return lib


def p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


@pytest.mark.parametrize("L,W", [(4, 6), (8, 4)])
def test_resume_from_fockstate_continues_the_chain(tmp_path, L, W):
    say("case resume", L, W)
    seed, n1, n2 = 606 + L, 5 * L * L, 5 * L * L
    occ = model.ground_occ(model.orbital_table(L, 0.25 * np.pi, 0.1))
    spins0 = model.random_spins(L, W, seed)
    with binding.Batch(L, W, binding.FP64, seed=seed) as b:
        b.set_orbitals(0, occ)
        ok = b.init(spins0)
        b.walk(n1)
        spins1 = b.state()[0]
        ref = b.walk(n2, trace=True)
        spins2 = b.state()[0]
    lib = host_lib()
    assert lib.gpvmc_host_fockstate_write(str(tmp_path).encode(), b"5", L, W, p(spins1)) == 0
    back = np.zeros_like(spins1)
    assert lib.gpvmc_host_fockstate_read(str(tmp_path / "5-FockState.h5").encode(), L, W, p(back)) == 0, \
        lib.gpvmc_host_last_error()
    assert np.array_equal(back, spins1)
    u = model.philox_uniforms(seed, range(W), n1, n2)
    with binding.Batch(L, W, binding.FP64, seed=seed) as b:
        b.set_orbitals(0, occ)
        ok2 = b.init(back)
        out = b.walk(n2, uniforms=u, trace=True)
        spins3 = b.state()[0]
    for w in range(W):
        if not ok[w]:
            continue
        assert ok2[w]
        assert np.array_equal(out["noop"][w], ref["noop"][w])
        assert np.array_equal(out["accept"][w], ref["accept"][w])
        assert np.array_equal(spins3[w], spins2[w])
        live = ~ref["noop"][w].astype(bool)
        r0, r1 = ref["ratio"][w][live], out["ratio"][w][live]
        # a ratio of a move onto a (numerically) vanishing amplitude is the rounding noise of a
        # cancellation, ~1e-60 here: an absolute floor far below any accepted move's ratio
        assert (np.abs(r1 - r0) <= 1e-10 * r0 + 1e-14).all()


@pytest.mark.parametrize("W", [1, 3])
def test_vmc_host_spinstate_starts_from_the_file(tmp_path, W):
    if not os.path.exists(build.HOST_EXE):
        pytest.skip("vmc_host not built")
    say("case vmc_host spinstate", W)
    L = 4
    common = [build.HOST_EXE, f"--L={L}", f"--walkers={W}", "--phi=0.25", "--neel=0.1", "--seed=11",
              f"--dir={tmp_path}", "--verbose=0", "--meas_projheis"]
    r = subprocess.run(common + ["--therm=3", "--samples=1", "--samples_saves_stat=4", "--prefix=1"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    lib = host_lib()
    first = np.zeros((W, L * L), np.int32)
    assert lib.gpvmc_host_fockstate_read(str(tmp_path / "1-FockState.h5").encode(), L, W, p(first)) == 0
    r = subprocess.run(common + ["--therm=0", "--samples=0", "--prefix=2", f"--spinstate={tmp_path / '1-FockState.h5'}"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    second = np.zeros_like(first)
    assert lib.gpvmc_host_fockstate_read(str(tmp_path / "2-FockState.h5").encode(), L, W, p(second)) == 0
    assert np.array_equal(first, second)
    # a file of another batch size is refused
    r = subprocess.run(common[:2] + [f"--walkers={W + 1}"] + common[3:] +
                       ["--therm=0", "--samples=0", "--prefix=3", f"--spinstate={tmp_path / '1-FockState.h5'}"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode != 0 and "rank-" in r.stderr


def raninit_host(L, W, n_up, seed, attempt):
    lib = ctypes.CDLL(build.HOST_LIB)
    lib.gpvmc_host_raninit.argtypes = [ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, ctypes.c_uint64,
                                       ctypes.c_int32, ctypes.c_void_p]
    out = np.zeros((W, L * L), np.int32)
    assert lib.gpvmc_host_raninit(L, W, n_up, seed, attempt, p(out)) == 0
    return out


@pytest.mark.parametrize("kind,L,W", [("ground", 4, 256), ("ground", 8, 64), ("trans", 4, 64), ("long", 4, 48)])
def test_device_random_start_equals_host_draws(kind, L, W):
    """gpvmc_init_random (vmc.cpp:221-242 on the device): each walker's final configuration is
    the host's gpvmc::RanInit draw of its last attempt (LatticeState.cpp:37-82 on the Philox
    stream), every earlier draw of a walker has a zero amplitude and its last one not, and
    the initialised state (amplitude, labels) equals gpvmc_init_walkers on those
    configurations."""
    say("case device random start", kind, L, W)
    seed = 4321 + L
    tab = model.orbital_table(L, 0.25 * np.pi, 0.1)
    if kind == "ground":
        occ = model.ground_occ(tab)
        n_up = L * L // 2

        def make():
            b = binding.Batch(L, W, binding.FP64, seed=3)
            b.set_orbitals(0, occ)
            return b
    else:
        ch = channels.Channel(kind, L, (1, 2) if kind == "trans" else (1, 0))
        n_up = ch.npt[0]

        def make():
            b = binding.ChannelBatch(L, W, ch, binding.FP64, seed=3)
            b.set_orbital_table(0, tab)
            return b
    with make() as b:
        # 40 draws: a random L = 4 start is singular often enough that 10 can run out
        ok, att = b.init_random(seed, 40)
        spins, part, la, ph = b.state()
    assert ok.all() and (att >= 1).all() and (att <= 40).all()
    draws = {t: raninit_host(L, W, n_up, seed, t) for t in range(int(att.max()))}
    for w in range(W):
        assert np.array_equal(spins[w], draws[att[w] - 1][w])
    if kind == "ground" and L == 4:
        assert (att > 1).sum() >= 3  # singular starts are common at L = 4
    # earlier draws are singular, and a host-fed init reproduces the device's start
    with make() as b2:
        ok2 = b2.init(spins)
        s2, part2, la2, ph2 = b2.state()
    assert ok2.all()
    assert np.array_equal(part2, part) and np.array_equal(la2, la) and np.array_equal(ph2, ph)
    redrawn = np.nonzero(att > 1)[0]
    for t in range(int(att.max()) - 1):
        ws = [w for w in redrawn if att[w] - 1 > t]
        if not ws:
            continue
        with make() as b3:
            ok3 = b3.init(draws[t])
        assert not ok3[ws].any()
