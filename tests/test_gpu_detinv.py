"""gpvmc_det_inv (batched linalg::DetInv, linalg.cpp:9-44) on device memory against the
oracle's or_detinv (LU with partial pivoting, pinned against numpy in test_oracle.py) and
numpy.  Sizes cover the register-resident kernels (n <= 128, padded; n in 77 .. 128 is the
panel-blocked one) and the streamed one (n > 128, with its workspace).

Bars: inverse within 1e-10 of the largest oracle entry (times the matrix's condition
estimate |A| |A^-1| / n, which is 1e1 .. 1e3 for these random matrices), ln|det| within
1e-10 per row (n 1e-12 + 1e-10), arg det within 1e-9; singular flags equal to the oracle's.
Device buffers come from hipMalloc of the library's own HIP runtime (tests/hipmem.py)."""
import numpy as np
import pytest

import oracle_ref as O
from gpvmc_amd import binding
from hipmem import DeviceArray

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def run(mats):
    batch, n, _ = mats.shape
    # column-major per matrix: element (i, j) at i + j n  ->  store the transpose row-major
    a = DeviceArray(np.ascontiguousarray(mats.transpose(0, 2, 1)))
    ai = DeviceArray(None, np.complex128, (batch, n, n))
    la = DeviceArray(None, np.float64, (batch,))
    ph = DeviceArray(None, np.float64, (batch,))
    sg = DeviceArray(None, np.int32, (batch,))
    binding.det_inv(n, batch, a.ptr, ai.ptr, la.ptr, ph.ptr, sg.ptr)
    out = ai.get().transpose(0, 2, 1), la.get(), ph.get(), sg.get()
    for d in (a, ai, la, ph, sg):
        d.free()
    return out


def random_mats(n, batch, seed):
    """complex Gaussian entries of unit variance: |det|^2 far above the reference's 1e-9
    singularity threshold at every size here"""
    rng = np.random.default_rng(seed)
    return (rng.normal(size=(batch, n, n)) + 1j * rng.normal(size=(batch, n, n))) / np.sqrt(2)


@pytest.mark.parametrize("n,batch", [(4, 9), (8, 16), (9, 7), (33, 12), (50, 8), (73, 8), (100, 5), (128, 6),
                                     (129, 5), (200, 4), (256, 3)])
def test_det_inv_matches_oracle(n, batch):
    say("case det_inv n=%d batch=%d" % (n, batch))
    mats = random_mats(n, batch, 100 + n)
    ai, la, ph, sg = run(mats)
    for b in range(batch):
        # the oracle takes the same column-major layout, element (i, j) at i + j n
        A_ref = np.ascontiguousarray(mats[b].T)
        Ai_o, la_o, ph_o, s_o = O.detinv(A_ref)
        Ai_o = Ai_o.T
        assert bool(sg[b]) == s_o and not s_o
        cond = np.abs(mats[b]).max() * np.abs(Ai_o).max() * n
        scale = np.abs(Ai_o).max()
        err = np.abs(ai[b] - Ai_o).max()
        assert err <= 1e-13 * cond * scale + 1e-14, (b, err, cond)
        assert abs(la[b] - la_o) <= n * 1e-12 + 1e-10
        assert abs(np.angle(np.exp(1j * (ph[b] - ph_o)))) <= 1e-9
        # an independent pin: numpy
        s_np, l_np = np.linalg.slogdet(mats[b])
        assert abs(la[b] - l_np) <= n * 1e-12 + 1e-10
        assert abs(np.angle(s_np * np.exp(-1j * ph[b]))) <= 1e-9
    say("  max inverse error %.2e" % max(np.abs(ai[b] - np.linalg.inv(mats[b])).max() for b in range(batch)))


@pytest.mark.parametrize("n", [16, 140])
def test_det_inv_flags_singular_matrices(n):
    say("case det_inv singular n=%d" % n)
    mats = random_mats(n, 3, 7 + n)
    # two equal columns (det = 0 up to the rounding of one pivot, which the absolute
    # 1e-9 threshold catches at n = 16 only) and a matrix scaled by 1e-3 (|det|^2 =
    # 1e-6n |det A|^2 < 1e-9 at any n); the third is regular
    mats[0][:, 3] = mats[0][:, 1]
    mats[1] *= 1e-3
    ai, la, ph, sg = run(mats)
    for b in range(3):
        _, la_o, _, s_o = O.detinv(np.ascontiguousarray(mats[b].T))
        if b or n == 16:
            assert bool(sg[b]) == s_o, b
    assert sg[1] == 1 and sg[2] == 0
    if n == 16:
        assert sg[0] == 1


def test_det_inv_rejects_bad_arguments():
    with pytest.raises(binding.GpvmcError, match="outside"):
        binding.det_inv(257, 1, 1, 1, 1, 1, 1)
    with pytest.raises(binding.GpvmcError, match="batch"):
        binding.det_inv(8, 0, 1, 1, 1, 1, 1)
    with pytest.raises(binding.GpvmcError, match="null"):
        binding.det_inv(8, 1, 0, 1, 1, 1, 1)
    assert binding.det_inv_workspace_bytes(128) == 0
    assert binding.det_inv_workspace_bytes(200) == 256 * 200 * 16
