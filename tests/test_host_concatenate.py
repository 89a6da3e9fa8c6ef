"""The merge of output files into samples (tools/concatenate.cpp:26-233), CPU only.

Pin: the reference's own tool, compiled from its one source file into oracle/_ref/concatenate,
merges the same input files (written by the host FileManager in the reference's layout,
several chains per file so that the records sit in /rank-1 .. /rank-W) into a file whose
h5dump equals ours: root attributes, /rank-1/data-s means (bit for bit) and their
"statistics".  The inputs give every record the same statistics, as the datasets of vmc runs
of one configuration do; with 24 records the reference's std::sort reorders equal keys, which
the grouping must follow.  For records of unequal statistics the reference sums entries of a
copy that it shrinks from both ends (concatenate.cpp:119-133), reading shifted and, past its
end, stale values; the host sums the records' own statistics.  There the grouping is pinned by
its invariants: a partition of the records, totals equal to the members' sums, samples within
one record's statistics of each other, and each output the plain mean of its members.
"""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

from gpvmc_amd import build

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_TOOL = os.path.join(ROOT, "oracle", "_ref", "concatenate")
H5DUMP = "/opt/conda/bin/h5dump"


def host_lib():
    if not os.path.exists(build.HOST_LIB):
        pytest.skip("libgpvmc_host.so not built")
    lib = ctypes.CDLL(build.HOST_LIB)
    lib.gpvmc_host_records_write.argtypes = [ctypes.c_char_p] * 4 + [ctypes.c_int32] * 4 + \
        [ctypes.c_void_p, ctypes.c_void_p]
    lib.gpvmc_host_bunch_records.argtypes = [ctypes.c_int32, ctypes.c_void_p, ctypes.c_int32, ctypes.c_void_p,
                                             ctypes.c_void_p, ctypes.c_void_p]
    lib.gpvmc_host_concatenate.argtypes = [ctypes.c_int32, ctypes.c_void_p, ctypes.c_char_p, ctypes.c_int32]
    lib.gpvmc_host_h5_read.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_void_p, ctypes.c_int64]
    lib.gpvmc_host_h5_read.restype = ctypes.c_int64
    lib.gpvmc_host_h5_attr.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_void_p]
    lib.gpvmc_host_last_error.restype = ctypes.c_char_p
    return lib


def p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def write_inputs(lib, d, files, chains, writes, stats, rows=2, cols=4, seed=0):
    """files x chains x writes records; returns their data [file][chain][write][rows*cols]"""
    rng = np.random.default_rng(seed)
    data = rng.normal(size=(files, writes, chains, rows * cols))
    for f in range(files):
        st = np.ascontiguousarray(stats[f], np.float64)
        assert lib.gpvmc_host_records_write(str(d).encode(), str(f + 1).encode(), b"ProjHeis",
                                            b"L=8;channel=groundstate", chains, writes, rows, cols,
                                            p(np.ascontiguousarray(data[f])), p(st)) == 0, \
            lib.gpvmc_host_last_error()
    return data.transpose(0, 2, 1, 3)


def dump(path):
    out = subprocess.run([H5DUMP, str(path)], capture_output=True, text=True, check=True).stdout
    return "\n".join(out.splitlines()[1:])


def bunch(lib, stats, nsamp):
    n = len(stats)
    st = np.ascontiguousarray(stats, np.int32)
    so = np.zeros(n, np.int32)
    rk = np.zeros(n, np.int32)
    tot = np.zeros(nsamp, np.int32)
    rc = lib.gpvmc_host_bunch_records(n, p(st), nsamp, p(so), p(rk), p(tot))
    return rc, so, rk, tot


@pytest.mark.skipif(not (os.path.exists(REF_TOOL) and os.path.exists(H5DUMP)),
                    reason="the reference tool (built from /root/reference) or h5dump is absent")
@pytest.mark.parametrize("files,chains,writes,nsamp", [(3, 4, 2, 0), (3, 4, 2, 1), (3, 4, 2, 5), (2, 3, 1, 2),
                                                       (2, 2, 3, 3)])
def test_concatenate_equals_reference_tool(tmp_path, files, chains, writes, nsamp):
    lib = host_lib()
    write_inputs(lib, tmp_path, files, chains, writes, [[400.0] * writes] * files, seed=files * 10 + nsamp)
    ins = [str(tmp_path / ("%d-ProjHeis.h5" % (f + 1))) for f in range(files)]
    args = (["-N", str(nsamp)] if nsamp else [])
    r = subprocess.run([REF_TOOL] + args + ins + [str(tmp_path / "ref.h5")], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    r = subprocess.run([build.CONCAT_EXE] + args + ins + [str(tmp_path / "ours.h5")], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    a, b = dump(tmp_path / "ref.h5"), dump(tmp_path / "ours.h5")
    assert a == b
    n_out = nsamp or files * chains * writes
    assert len(re.findall(r'DATASET "data-\d+"', b)) == n_out


def test_concatenate_means_and_statistics(tmp_path):
    """unequal statistics: the output datasets are the plain means of their members, the
    statistics the members' sums, through the C ABI"""
    lib = host_lib()
    files, chains, writes, nsamp = 2, 3, 4, 3
    stats = [[100.0, 250.0, 400.0, 175.0], [300.0, 50.0, 125.0, 800.0]]
    data = write_inputs(lib, tmp_path, files, chains, writes, stats, seed=5)
    ins = [str(tmp_path / ("%d-ProjHeis.h5" % (f + 1))).encode() for f in range(files)]
    arr = (ctypes.c_char_p * files)(*ins)
    out = str(tmp_path / "m.h5").encode()
    assert lib.gpvmc_host_concatenate(files, ctypes.cast(arr, ctypes.c_void_p), out, nsamp) == 0, \
        lib.gpvmc_host_last_error()
    # record order: file, then rank, then data-k
    recs = [(f, c, k) for f in range(files) for c in range(chains) for k in range(writes)]
    rstat = [int(stats[f][k]) for f, c, k in recs]
    rc, so, rk, tot = bunch(lib, rstat, nsamp)
    assert rc == 0
    for s in range(nsamp):
        mem = sorted(np.nonzero(so == s)[0], key=lambda i: rk[i])
        want = np.mean([data[recs[i][0], recs[i][1], recs[i][2]] for i in mem], axis=0)
        got = np.zeros(64)
        n = lib.gpvmc_host_h5_read(out, ("/rank-1/data-%d" % s).encode(), p(got), 64)
        assert n == want.size
        np.testing.assert_allclose(got[:n], want, rtol=1e-15, atol=1e-15)
        st = ctypes.c_double()
        assert lib.gpvmc_host_h5_attr(out, ("/rank-1/data-%d" % s).encode(), b"statistics", ctypes.byref(st)) == 0
        assert st.value == tot[s] == sum(rstat[i] for i in mem)
    L = ctypes.c_double()
    assert lib.gpvmc_host_h5_attr(out, b"/", b"L", ctypes.byref(L)) == 0 and L.value == 8


@pytest.mark.parametrize("n,nsamp,seed", [(7, 2, 1), (40, 5, 2), (41, 20, 3), (9, 9, 4), (100, 7, 5)])
def test_bunching_invariants(n, nsamp, seed):
    lib = host_lib()
    stats = np.random.default_rng(seed).integers(1, 1000, n)
    rc, so, rk, tot = bunch(lib, stats, nsamp)
    assert rc == 0
    # every record in exactly one sample (so[i] is written once per record), no sample empty
    assert set(so.tolist()) == set(range(nsamp))
    for s in range(nsamp):
        mem = np.nonzero(so == s)[0]
        assert sorted(rk[mem].tolist()) == list(range(len(mem)))
        assert tot[s] == stats[mem].sum()
    assert tot.sum() == stats.sum()
    assert tot.max() - tot.min() <= stats.max()


def test_bunching_equal_statistics_and_errors():
    lib = host_lib()
    rc, so, rk, tot = bunch(lib, [5] * 12, 4)
    assert rc == 0 and list(tot) == [15] * 4
    # one sample per record: the largest index first into sample 0, as the reference's
    # second loop hands out the sorted records from the back
    rc, so, rk, tot = bunch(lib, [3, 1, 2], 3)
    assert rc == 0 and list(so) == [0, 2, 1] and list(tot) == [3, 2, 1]
    assert bunch(lib, [1, 2], 3)[0] == -1
    assert b"more samples" in lib.gpvmc_host_last_error()
    assert bunch(lib, [1, 2], 0)[0] == -1


def test_concatenate_errors(tmp_path):
    lib = host_lib()
    write_inputs(lib, tmp_path, 1, 1, 2, [[10.0, 10.0]])  # one chain: /rank-0 only
    single = str(tmp_path / "1-ProjHeis.h5")
    r = subprocess.run([build.CONCAT_EXE, single, str(tmp_path / "o.h5")], capture_output=True, text=True)
    assert r.returncode == 2 and "no /rank-" in r.stderr
    r = subprocess.run([build.CONCAT_EXE, "x"], capture_output=True, text=True)
    assert r.returncode == 1 and "Usage" in r.stdout
    r = subprocess.run([build.CONCAT_EXE, "-N", "abc", single, str(tmp_path / "o.h5")], capture_output=True, text=True)
    assert r.returncode == 1 and "Invalid argument" in r.stderr
    d2 = tmp_path / "two"
    d2.mkdir()
    write_inputs(lib, d2, 1, 2, 1, [[10.0]])
    two = str(d2 / "1-ProjHeis.h5")
    r = subprocess.run([build.CONCAT_EXE, "-N", "3", two, str(d2 / "o.h5")], capture_output=True, text=True)
    assert r.returncode == 2 and "more samples" in r.stderr
    r = subprocess.run([build.CONCAT_EXE, two, str(d2 / "o.h5")], capture_output=True, text=True)
    assert r.returncode == 0
    r = subprocess.run([build.CONCAT_EXE, two, str(d2 / "o.h5")], capture_output=True, text=True)
    assert r.returncode == 2 and "could not create" in r.stderr
    d3 = tmp_path / "three"
    d3.mkdir()
    write_inputs(lib, d3, 1, 2, 1, [[10.0]], rows=2, cols=4)
    os.rename(d3 / "1-ProjHeis.h5", d3 / "a.h5")
    write_inputs(lib, d3, 1, 2, 1, [[10.0]], rows=3, cols=4)
    r = subprocess.run([build.CONCAT_EXE, "-N", "1", str(d3 / "a.h5"), str(d3 / "1-ProjHeis.h5"), str(d3 / "o.h5")],
                       capture_output=True, text=True)
    assert r.returncode == 2 and "differs in size" in r.stderr
