"""Channel states and hop graphs against the reference's excitation-graph fixtures
(tests/golden/excgraph_*: StagFluxTransExciton / StagFluxLongExciton built from the
reference sources at L=4, phi=0.1, neel=0.05, bc=(1,1))."""
import glob
import os

import numpy as np
import pytest

from gpvmc_amd import channels

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
FIX = sorted(glob.glob(os.path.join(GOLD, "excgraph_*.npz")))


def _paths(flat, offs, n):
    """offsets -> list over n entries of per-flavour hop lists"""
    out = []
    for i in range(n):
        out.append([[tuple(flat[j]) for j in range(offs[2 * i + fl], offs[2 * i + fl + 1])] for fl in range(2)])
    return out


@pytest.mark.parametrize("path", FIX, ids=[os.path.basename(f) for f in FIX])
def test_channel_graph_matches_reference(path):
    g = np.load(path, allow_pickle=False)
    name = "trans" if "trans" in os.path.basename(path) else "long"
    ch = channels.Channel(name, int(g["L"]), tuple(g["q"]))
    assert ch.nexc == int(g["nexc"])
    assert ch.npt == tuple(int(x) for x in g["npt"])
    # state 0's occupied orbitals by slot (slots follow increasing orbital order)
    np.testing.assert_array_equal(ch.states[0][0], g["state0_part_up"])
    np.testing.assert_array_equal(ch.states[0][1], g["state0_part_dn"])
    hops0 = _paths(g["hops"], g["hop_offsets"], ch.nexc)
    for k in range(ch.nexc):
        assert [list(h) for h in hops0[k]] == [list(map(tuple, ch.hops[0][k][fl])) for fl in range(2)], k
    np.testing.assert_array_equal(ch.hopsign[0], g["hopsign"])
    nb = int(g["block_n"])
    blk = _paths(g["block"], g["block_offsets"], nb * nb)
    for i in range(nb):
        for j in range(nb):
            assert blk[i * nb + j] == [list(map(tuple, ch.hops[i][j][fl])) for fl in range(2)], (i, j)


def test_channel_sizes_and_sign_consistency():
    for name, L, q, nk in (("trans", 4, (1, 2), 8), ("long", 4, (0, 0), 17), ("long", 4, (1, 0), 16),
                           ("groundstate", 4, (0, 0), 1), ("trans", 6, (2, 3), 18)):
        ch = channels.Channel(name, L, q)
        assert ch.nexc == nk
        assert ch.max_hops <= 2
        # hopping c -> k -> c returns the sign of the identity
        for c in range(ch.nexc):
            for k in range(ch.nexc):
                assert ch.hopsign[c, k] * ch.hopsign[k, c] == 1
