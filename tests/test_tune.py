"""Launch-shape tuning (gpvmc_amd.tune): the shape query of the C ABI and the selection
logic of pick_launch, on the CPU (stand-in contexts with known walk times)."""
import time

import pytest

from gpvmc_amd import binding, tune


def test_sweep_shapes_lists_the_default_first():
    for L in (4, 6, 8, 10, 12, 16, 20):
        for prec in (binding.FP64, binding.MIXED):
            shapes = binding.sweep_shapes(L, prec)
            assert shapes, (L, prec)
            assert all(s in (1, 2, 4, 8, 16) for s in shapes)
    assert binding.sweep_shapes(8, binding.FP64)[0] == 4
    assert binding.sweep_shapes(16, binding.MIXED) == [4]
    with pytest.raises(binding.GpvmcError):
        binding.sweep_shapes(5, binding.FP64)


def test_candidates():
    c16 = tune.launch_candidates(16, binding.MIXED)
    assert dict(lanes_per_row=4, variant=binding.SWEEP_WINDOWS) in c16
    assert c16[0] == dict(lanes_per_row=4, variant=binding.SWEEP_ONE_MOVE)
    assert len(tune.launch_candidates(8, binding.MIXED)) == 2
    assert tune.launch_candidates(20, binding.MIXED, channel=True)[1] == dict(kernel="streamed")


class Fake:
    def __init__(self, t, windows=True):
        self.t, self.windows, self.closed, self.v = t, windows, False, binding.SWEEP_ONE_MOVE

    def walk(self, n, **kw):
        time.sleep(self.t)

    def close(self):
        self.closed = True

    @property
    def sweep_variant(self):
        return self.v

    @sweep_variant.setter
    def sweep_variant(self, v):
        if v == binding.SWEEP_WINDOWS and not self.windows:
            raise binding.GpvmcError("no windows")
        self.v = v


def test_pick_launch_keeps_the_fastest_open_and_closes_the_rest():
    made = []

    def make(lanes_per_row=0):
        f = Fake({4: 0.004, 2: 0.012}[lanes_per_row], windows=lanes_per_row == 4)
        made.append(f)
        return f
    cands = [dict(lanes_per_row=2, variant=binding.SWEEP_ONE_MOVE),
             dict(lanes_per_row=2, variant=binding.SWEEP_WINDOWS),
             dict(lanes_per_row=4, variant=binding.SWEEP_ONE_MOVE)]
    b, c, hist = tune.pick_launch(make, cands, 10)
    assert c == cands[2] and b is made[2] and not b.closed
    assert made[0].closed and made[1].closed  # slower, and a variant the shape lacks
    assert [h["candidate"] for h in hist] == ["[('lanes_per_row', 2)] variant 1", "[('lanes_per_row', 4)] variant 1"]


def test_pick_launch_single_candidate_is_untimed():
    b, c, hist = tune.pick_launch(lambda: Fake(1.0), [dict()], 10)
    assert hist == [] and c == {} and not b.closed


def test_pick_launch_duplicate_kernels_are_untimed():
    made = []

    def make(kernel="auto"):
        f = Fake(1.0)
        f.register_resident = False
        made.append(f)
        return f
    b, c, hist = tune.pick_launch(make, [dict(kernel="auto"), dict(kernel="streamed")], 10)
    assert hist == [] and c == dict(kernel="auto") and b is made[0] and made[1].closed
