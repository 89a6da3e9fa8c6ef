"""gpvmc_det_inv_noalloc (no device allocation; workspace in the library's fixed device
array) against gpvmc_det_inv on the same inputs: bit-identical inverses, ln|det|, arg det
and singular flags (the same kernel on the same data), on two calls in a row with the
second smaller, both through the streamed kernel that uses the workspace (n > 128) and
through the register-resident one; and a batch beyond the fixed workspace is refused
before anything runs."""
import numpy as np
import pytest

from gpvmc_amd import binding
from hipmem import DeviceArray

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def run(fn, mats):
    batch, n, _ = mats.shape
    a = DeviceArray(np.ascontiguousarray(mats.transpose(0, 2, 1)))
    ai = DeviceArray(None, np.complex128, (batch, n, n))
    la = DeviceArray(None, np.float64, (batch,))
    ph = DeviceArray(None, np.float64, (batch,))
    sg = DeviceArray(None, np.int32, (batch,))
    fn(n, batch, a.ptr, ai.ptr, la.ptr, ph.ptr, sg.ptr)
    out = ai.get(), la.get(), ph.get(), sg.get()
    for d in (a, ai, la, ph, sg):
        d.free()
    return out


def mats_of(n, batch, seed):
    rng = np.random.default_rng(seed)
    return (rng.normal(size=(batch, n, n)) + 1j * rng.normal(size=(batch, n, n))) / np.sqrt(2)


@pytest.mark.parametrize("sizes", [((200, 40), (150, 12)), ((96, 24), (40, 8))])
def test_noalloc_equals_allocating_entry_point(sizes):
    for i, (n, batch) in enumerate(sizes):
        say("case det_inv_noalloc call %d: n=%d batch=%d" % (i, n, batch))
        m = mats_of(n, batch, 31 * n + i)
        ref = run(binding.det_inv, m)
        got = run(binding.det_inv_noalloc, m)
        for r, g in zip(ref, got):
            assert np.array_equal(r, g)
        # and the inverse is an inverse
        inv = ref[0].transpose(0, 2, 1)
        assert np.abs(np.einsum("bij,bjk->bik", m, inv) - np.eye(n)[None]).max() < 1e-9


def test_noalloc_refuses_batches_beyond_its_workspace():
    per = binding.det_inv_workspace_bytes(256)
    too_many = (64 << 20) // per + 1
    with pytest.raises(binding.GpvmcError, match="at most 64 per call"):
        binding.det_inv_noalloc(256, too_many, 1, 1, 1, 1, 1)
