"""The reference's output files written by the C++ host (CPU).

Layout pin: the reference's own FileManager (FileManager.cpp:465-575), WaveFunction::Save
(WaveFunction.cpp:85-137) and State::Save (State.cpp:117-166), compiled from its sources
into oracle/_ref/ref_harness (mode h5layout), write a ProjHeis-named record (4 saves, 2 per
sample, with rej/time/statistics), the ground state's and a transverse exciton's
WaveFunction files and a FockState file; the host classes write the same script
(gpvmc_host_h5_script).  The two trees, dumped by the image's h5dump, must be identical:
groups, datasets, datatypes, shapes, attributes and data, except the creation date.

Also: FockState files read back by the --spinstate reader for one chain (/rank-0) and for a
batch (/rank-1 .. /rank-W), its errors, and the reference's RanInit draw (LatticeState.cpp:
37-82) on the reference's own glibc stream against the starts the reference drew.
"""
import ctypes
import json
import os
import re
import subprocess

import numpy as np
import pytest

import oracle_ref as O
from gpvmc_amd import build

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HARNESS = os.path.join(ROOT, "oracle", "_ref", "ref_harness")
H5DUMP = "/opt/conda/bin/h5dump"
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def host_lib():
    if not os.path.exists(build.HOST_LIB):
        pytest.skip("libgpvmc_host.so not built")
    lib = ctypes.CDLL(build.HOST_LIB)
    lib.gpvmc_host_h5_script.argtypes = [ctypes.c_char_p, ctypes.c_int32, ctypes.c_double, ctypes.c_double,
                                         ctypes.c_void_p]
    lib.gpvmc_host_fockstate_write.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int32, ctypes.c_int32,
                                               ctypes.c_void_p]
    lib.gpvmc_host_fockstate_read.argtypes = [ctypes.c_char_p, ctypes.c_int32, ctypes.c_int32, ctypes.c_void_p]
    lib.gpvmc_host_raninit_stream.argtypes = [ctypes.c_int32, ctypes.c_int32, ctypes.c_void_p, ctypes.c_int64,
                                              ctypes.c_void_p]
    lib.gpvmc_host_raninit_stream.restype = ctypes.c_int64
    lib.gpvmc_host_last_error.restype = ctypes.c_char_p
    return lib


def p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def dump(path):
    """h5dump of a file with the file name and the value of the creation date blanked"""
    out = subprocess.run([H5DUMP, str(path)], capture_output=True, text=True, check=True).stdout
    lines = out.splitlines()[1:]  # first line names the file
    text = "\n".join(lines)
    return re.sub(r'(ATTRIBUTE "date" \{.*?DATA \{\s*\(0\): )"[^"]*"', r'\1"<date>"', text, flags=re.S)


@pytest.mark.skipif(not (os.path.exists(HARNESS) and os.path.exists(H5DUMP)),
                    reason="the reference harness (built from /root/reference) or h5dump is absent")
@pytest.mark.parametrize("L,phi,neel,seed", [(4, 0.25, 0.1, 12345), (6, 0.1667, 0.05, 4242)])
def test_hdf5_layout_equals_reference(tmp_path, L, phi, neel, seed):
    ref, ours = tmp_path / "ref", tmp_path / "ours"
    ref.mkdir()
    ours.mkdir()
    r = subprocess.run([HARNESS, "h5layout", str(ref), str(L), str(phi), str(neel), str(seed)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    fock = np.array(json.loads(r.stdout)["fock"], np.int32)
    lib = host_lib()
    assert lib.gpvmc_host_h5_script(str(ours).encode(), L, phi, neel, p(fock)) == 0, lib.gpvmc_host_last_error()
    names = sorted(os.listdir(ref))
    assert names == ["7-FockState.h5", "7-ProjHeis.h5", "7-WaveFunction.h5", "8-WaveFunction.h5"]
    assert sorted(os.listdir(ours)) == names
    for n in names:
        a, b = dump(ref / n), dump(ours / n)
        assert a == b, n


def test_fockstate_roundtrip_and_rank_groups(tmp_path):
    lib = host_lib()
    L = 4
    rng = np.random.default_rng(3)
    for W in (1, 3):
        spins = np.ones((W, L * L), np.int32)
        for w in range(W):
            spins[w, rng.choice(L * L, L * L // 2, replace=False)] = 0
        pre = "r%d" % W
        assert lib.gpvmc_host_fockstate_write(str(tmp_path).encode(), pre.encode(), L, W, p(spins)) == 0
        back = np.zeros_like(spins)
        path = str(tmp_path / (pre + "-FockState.h5")).encode()
        assert lib.gpvmc_host_fockstate_read(path, L, W, p(back)) == 0
        assert np.array_equal(back, spins)
        if os.path.exists(H5DUMP):
            text = subprocess.run([H5DUMP, "-H", path.decode()], capture_output=True, text=True).stdout
            groups = sorted(re.findall(r'GROUP "(rank-\d+)"', text))
            assert groups == (["rank-0"] if W == 1 else ["rank-%d" % (c + 1) for c in range(W)])
    # a file of one chain does not hold a batch of three; a wrong size is refused
    back = np.zeros((3, L * L), np.int32)
    assert lib.gpvmc_host_fockstate_read(str(tmp_path / "r1-FockState.h5").encode(), L, 3, p(back)) == -1
    assert b"rank-1" in lib.gpvmc_host_last_error()
    assert lib.gpvmc_host_fockstate_read(str(tmp_path / "r1-FockState.h5").encode(), 6, 1, p(back)) == -1
    assert lib.gpvmc_host_fockstate_read(str(tmp_path / "none.h5").encode(), L, 1, p(back)) == -1


@pytest.mark.parametrize("name", ["walk_L4_s12345.npz", "walk_L6_s4242.npz", "walk_L8_s777.npz"])
def test_raninit_draw_equals_reference(name):
    """Every RanInit draw the reference made at the start of its walk (fixture init_attempts,
    the glibc rand() stream of its seed consumed in order) equals the host's draw on the same
    stream; the stream position after the draws equals the oracle's."""
    g = np.load(os.path.join(GOLD, name), allow_pickle=False)
    L = int(g["L"])
    u = O.glibc_uniforms(int(g["seed"]), 20000)
    lib = host_lib()
    pos = 0
    for att in g["init_attempts"]:
        spins = np.zeros(L * L, np.int32)
        used = lib.gpvmc_host_raninit_stream(L, L * L // 2, p(u[pos:].copy()), len(u) - pos, p(spins))
        assert used > 0
        assert np.array_equal(spins, att.astype(np.int32))
        pos += used
    _, used_o = O.raninit(L, u)
    if len(g["init_attempts"]) == 1:
        assert pos == used_o
