"""prism_md_build_pairs_static -- the pair build staged in the library's
static device array instead of memory allocated per call -- against
prism_md_build_pairs on the same state: the same pairs, bit for bit, on two
calls in a row (the second with a smaller system, whose list must not carry
anything over from the first), and a list beyond the static capacity refused
with E_CAPACITY.  Through the C ABI."""
from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _both(pm, top, mode):
    with pm.FusedEngine(top) as eng:
        eng.step(30)
        static = eng.pairs(mode, static=True)
        alloc = eng.pairs(mode)
    return static, alloc


def test_static_pairs_equal_allocating_pairs(pm):
    big = pm.Topology.generate(6000, seed=71)
    small = pm.Topology.generate(900, seed=72)
    for top, mode in ((big, 1), (small, 0)):  # second call: a smaller input
        static, alloc = _both(pm, top, mode)
        print(f"[no-alloc] {top.n_atoms} atoms mode {mode}: {len(static)} pairs", flush=True)
        assert len(static) > 0 and static.shape == alloc.shape and np.array_equal(static, alloc)


def test_static_pairs_capacity_is_an_error(pm):
    top = pm.Topology.generate(20000, seed=73)  # ~7 M pairs within 12 A: above 4,194,304
    with pm.FusedEngine(top) as eng:
        with pytest.raises(pm.PrismMdError) as ex:
            eng.pairs(1, static=True)
        assert ex.value.rc == pm.E_CAPACITY and "exceeds" in str(ex.value)
        assert len(eng.pairs(1)) > (1 << 22)  # the allocating call still serves it
