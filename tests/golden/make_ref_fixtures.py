"""Writes the golden fixtures under tests/golden/ from the reference-built harness.

Run in the build container only (it needs /root/reference to build oracle/_ref):
    make -C oracle ref && python tests/golden/make_ref_fixtures.py [base] [exciton] [eloc]

The harness (oracle/ref_harness.cpp) drives the reference's own compiled classes; the
fixtures are plain data (.npz, no pickles): lattice edge order, orbital tables,
RanInit configurations, per-step weight ratios and moves of MetroMC::Step, snapshots,
inverse matrices and excitation graphs; exciton chains with their ProjHeis H/O sums and
ground-state E_loc along a chain.  The uniform stream of each walk is glibc's
rand()/RAND_MAX from the stored seed (the reference's USE_RNG_STD generator); the first
values of that stream are stored too so a libc mismatch is detected.
"""
import ctypes
import json
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
HARNESS = os.path.join(ROOT, "oracle", "_ref", "ref_harness")


def run(*args):
    env = dict(os.environ, MKL_NUM_THREADS="1", OMP_NUM_THREADS="1")
    out = subprocess.run([HARNESS] + [str(a) for a in args], check=True,
                         capture_output=True, text=True, env=env).stdout
    # the reference prints its own warnings (a singular DetInv of a channel state) on stdout
    return json.loads([line for line in out.splitlines() if line.startswith("{")][-1])


def cplx(flat):
    a = np.asarray(flat, np.float64).reshape(-1, 2)
    return a[:, 0] + 1j * a[:, 1]


def glibc_head(seed, n=8):
    libc = ctypes.CDLL("libc.so.6")
    libc.srand(ctypes.c_uint(seed))
    return np.array([libc.rand() for _ in range(n)], np.int64)


def lattice(L):
    d = run("lattice", L)
    np.savez_compressed(os.path.join(HERE, "lattice_L%d.npz" % L),
                        neighbours=np.array(d["neighbours"], np.int32),
                        edges=np.array(d["edges"], np.int32))


def orbitals(L, phi, neel, ne, bcx, bcy, tag, subsample=None):
    d = run("orbitals", L, phi, neel, ne, bcx, bcy)
    re = np.array(d["re"]).reshape(2, L * L, L * L)
    im = np.array(d["im"]).reshape(2, L * L, L * L)
    table = re + 1j * im
    kw = dict(L=L, phi=phi, neel=neel, neel_exp=ne, bc=np.array([bcx, bcy]),
              fock2qn=np.array(d["fock2qn"], np.int32))
    if subsample:
        rng = np.random.default_rng(20260101 + L)
        idx = rng.choice(table.size, subsample, replace=False)
        kw.update(index=idx.astype(np.int64), values=table.reshape(-1)[idx])
    else:
        kw.update(table=table)
    np.savez_compressed(os.path.join(HERE, "orbitals_%s.npz" % tag), **kw)


def walk(L, phi, neel, ne, seed, nsteps, snap_every, keep_ratio):
    d = run("walk", L, phi, neel, ne, seed, nsteps, snap_every)
    N = L * L // 2
    # the harness stream must never hit u = 1.0 (an out-of-range vertex in the reference)
    libc = ctypes.CDLL("libc.so.6")
    libc.srand(ctypes.c_uint(seed))
    need = 4 * nsteps + 10000
    assert all(libc.rand() != 2147483647 for _ in range(need)), "stream hits RAND_MAX"
    amp = d["init_amp"]
    logabs = np.log(abs(complex(amp[0], amp[1]))) + amp[2] * np.log(10.0)
    phase = np.angle(complex(amp[0], amp[1]))
    sw = np.array(d["snap_weight"], np.float64)
    sa = np.array(d["snap_amp"], np.float64)
    np.savez_compressed(
        os.path.join(HERE, "walk_L%d_s%d.npz" % (L, seed)),
        L=L, phi=phi, neel=neel, neel_exp=ne, seed=seed, nsteps=nsteps,
        glibc_head=glibc_head(seed),
        init_attempts=np.array(d["init_attempts"], np.int8),
        init_spins=np.array(d["init_fock"], np.int8),
        init_labels=np.array(d["init_labels"], np.int32),
        init_logabs=logabs, init_phase=phase,
        init_inv=cplx(d["init_inv"]).reshape(2, N * N),
        ratio=np.array(d["ratio"][:keep_ratio], np.float64),
        moved=np.packbits(np.array(d["moved"], np.uint8)),
        snap_every=snap_every,
        snaps=np.array(d["snaps"], np.int8),
        snap_log10w=np.log10(sw[:, 0]) + sw[:, 1],
        snap_amp=sa,
        final_labels=np.array(d["final_labels"], np.int32),
        final_inv=cplx(d["final_inv"]).reshape(2, N * N),
    )


def excgraph(L, channel, qx, qy):
    d = run("excgraph", L, channel, qx, qy)
    hops = d["hops"]
    # flatten: per state k, per flavour, list of (from, to); encode with offsets
    flat, offs = [], [0]
    for k in range(d["nexc"]):
        for fl in range(2):
            for h in hops[k][fl]:
                flat.append(h)
            offs.append(len(flat))
    blk = d["exc_block"]
    bflat, boffs = [], [0]
    for i in range(len(blk)):
        for j in range(len(blk)):
            for fl in range(2):
                for h in blk[i][j][fl]:
                    bflat.append(h)
                boffs.append(len(bflat))
    np.savez_compressed(
        os.path.join(HERE, "excgraph_%s_L%d_q%d_%d.npz" % (channel, L, qx, qy)),
        L=L, q=np.array([qx, qy]), nexc=d["nexc"], npt=np.array(d["npt"]),
        state0_part_up=np.array(d["state0_part"][0], np.int32),
        state0_part_dn=np.array(d["state0_part"][1], np.int32),
        hops=np.array(flat, np.int32).reshape(-1, 2), hop_offsets=np.array(offs, np.int32),
        hopsign=np.array(d["hopsign"], np.int8),
        block=np.array(bflat, np.int32).reshape(-1, 2), block_offsets=np.array(boffs, np.int32),
        block_n=len(blk))


def no_rand_max(seed, need):
    libc = ctypes.CDLL("libc.so.6")
    libc.srand(ctypes.c_uint(seed))
    assert all(libc.rand() != 2147483647 for _ in range(need)), "stream hits RAND_MAX"


def excwalk(channel, L, qx, qy, phi, neel, seed, nsteps, meas):
    """The reference's exciton chain (its DetUpdate/InvUpdate with the crossing terms taken
    from the A of the last Init), ProjHeis H/O sums every `meas` steps."""
    d = run("excwalk", channel, L, qx, qy, phi, neel, seed, nsteps, meas)
    assert d["init_attempts"] == 1, "a retried start leaves ghost labels in the reference's fock"
    no_rand_max(seed, d["stream_pos"] + 4 * nsteps)
    ho = cplx(d["ho"])
    np.savez_compressed(
        os.path.join(HERE, "excwalk_%s_L%d_q%d_%d_s%d.npz" % (channel, L, qx, qy, seed)),
        channel=channel, L=L, q=np.array([qx, qy]), phi=phi, neel=neel, seed=seed, nsteps=nsteps,
        meas=meas, glibc_head=glibc_head(seed), stream_pos=d["stream_pos"], kmax=d["kmax"],
        init_spins=np.array(d["init_fock"], np.int8), ratio=np.array(d["ratio"], np.float64),
        moved=np.packbits(np.array(d["moved"], np.uint8)), n_meas=d["n_meas"],
        ho=ho.reshape(2, -1, int(round(np.sqrt(len(ho) / 2)))),
        final_spins=np.array(d["final_fock"], np.int8))


def eloc(L, phi, neel, seed, nsteps, meas):
    """ProjHeis::measure of the ground state (E_loc per site) at the configurations of a chain"""
    d = run("eloc", L, phi, neel, seed, nsteps, meas)
    e = np.array(d["eloc"], np.float64)
    np.savez_compressed(
        os.path.join(HERE, "eloc_L%d_s%d.npz" % (L, seed)),
        L=L, phi=phi, neel=neel, seed=seed, eloc=e[:, 0] + 1j * e[:, 1],
        confs=np.array(d["confs"], np.int8))


def main(groups):
    if not os.path.exists(HARNESS):
        sys.exit("build the reference harness first: make -C oracle ref")
    if "exciton" in groups:
        excwalk("trans", 4, 1, 2, 0.25, 0.1, 12345, 4000, 16)
        excwalk("long", 4, 2, 2, 0.25, 0.1, 4242, 4000, 16)
        excwalk("trans", 6, 1, 2, 0.25, 0.1, 31, 2000, 36)
        excwalk("long", 6, 2, 0, 0.25, 0.1, 33, 2000, 36)
    if "eloc" in groups:
        eloc(4, 0.25, 0.1, 12345, 2000, 16)
        eloc(6, 0.1667, 0.05, 4242, 3600, 36)
    if "base" not in groups:
        return
    for L in (4, 6, 8):
        lattice(L)
    orbitals(4, 0.25, 0.1, 0.0, 1.0, 1.0, "L4_a")
    orbitals(4, 0.1, 0.0, 0.0, 1.0, 0.0, "L4_b")
    orbitals(4, 0.3, 0.05, 1.5, 0.0, 1.0, "L4_c")
    orbitals(8, 0.25, 0.1, 0.0, 1.0, 1.0, "L8_a", subsample=2000)
    orbitals(6, 0.1667, 0.03, 0.0, 1.0, 1.0, "L6_a", subsample=1000)
    walk(4, 0.25, 0.1, 0.0, 12345, 6000, 500, 6000)
    walk(6, 0.1667, 0.05, 0.0, 4242, 8000, 1000, 3000)
    walk(8, 0.25, 0.1, 0.0, 777, 16000, 2000, 3000)
    for q in ((0, 0), (1, 2), (2, 2)):
        excgraph(4, "trans", *q)
    for q in ((0, 0), (1, 0), (2, 2)):
        excgraph(4, "long", *q)


if __name__ == "__main__":
    # groups: base (lattice, orbitals, walks, excitation graphs), exciton, eloc; default all
    main(sys.argv[1:] or ["base", "exciton", "eloc"])
