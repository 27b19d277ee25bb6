"""Binomial checkpoint schedule and HBM-budget plan (host only, no GPU).

The schedule (fwi_hip_revolve_schedule) is checked by simulating it on
symbolic states: every turn must run on the state of its own step, the turns
must visit steps nt-1 .. 0 in that order, snapshots and restores must use
slots 0 .. snaps-1 and restore exactly the state stored.  Its cost (forward
steps of the advances and turns) must equal the optimum of a brute-force
dynamic programme over every split, and the closed form of Griewank's
binomial bound: r nt - C(snaps + r, r - 1) advance steps, r the least
repetition number with C(snaps + r, r) >= nt.
"""
import functools
import math

import numpy as np
import pytest

from fwihip import lib as L
from fwihip.workloads import Geometry

ADVANCE, TAKESHOT, RESTORE, TURN = 0, 1, 2, 3


def simulate(acts, nt, snaps):
    """Run a schedule on symbolic states (the state before step k is k)."""
    slots = {0: 0}                       # the forward pass stores S_0 in slot 0
    live = nt - 1                        # the adjoint pass starts from S_{nt-1}
    # the actions before the first turn ran during the forward pass, which
    # takes their snapshots on its way to step nt
    first = next(i for i, a in enumerate(acts) if a[0] == TURN)
    pos = 0
    for op, a, b, slot in acts[:first]:
        if op == ADVANCE:
            assert a == pos and b > a
            pos = b
        elif op == TAKESHOT:
            assert a == pos and 0 <= slot < snaps
            slots[slot] = a
        else:
            raise AssertionError(f"op {op} in the first sweep")
    assert pos == nt - 1                 # the first sweep reaches the last step's state
    turns = []
    for op, a, b, slot in acts[first:]:
        if op == ADVANCE:
            assert live == a and b > a
            live = b
        elif op == TAKESHOT:
            assert live == a and 0 <= slot < snaps
            slots[slot] = a
        elif op == RESTORE:
            assert slots.get(slot) == a, (slot, a, slots)
            live = a
        else:
            assert live == a, (live, a)
            turns.append(a)
            live = a + 1
    assert turns == list(range(nt - 1, -1, -1))


@functools.lru_cache(maxsize=None)
def best(l, s):
    """Least advance steps to reverse l steps with s slots (one holding the start)."""
    if l <= 1:
        return 0
    if s <= 1:
        return l * (l - 1) // 2
    return min(j + best(l - j, s - 1) + best(j, s) for j in range(1, l))


def beta(s, r):
    return math.comb(s + r, s)


@pytest.mark.parametrize("snaps", [1, 2, 3, 4, 6])
@pytest.mark.parametrize("nt", [1, 2, 3, 7, 20, 37, 61, 100, 151])
def test_schedule_is_valid_and_optimal(nt, snaps):
    acts, fsteps = L.revolve_schedule(nt, snaps)
    acts = [tuple(int(v) for v in a) for a in acts]
    simulate(acts, nt, snaps)
    advances = sum(b - a for op, a, b, _ in acts if op == ADVANCE)
    assert fsteps == advances + nt
    assert advances == best(nt, snaps), (nt, snaps, advances, best(nt, snaps))
    # the binomial bound: r nt - beta(s+1, r-1) advance steps, with r the
    # least repetition number reaching nt
    if nt > 1:
        r = next(r for r in range(nt + 1) if beta(snaps, r) >= nt)
        assert advances == r * nt - beta(snaps + 1, r - 1)


def test_schedule_at_benchmark_sizes():
    for nt, snaps in ((4000, 32), (4000, 128), (10000, 32)):
        acts, fsteps = L.revolve_schedule(nt, snaps)
        simulate([tuple(int(v) for v in a) for a in acts], nt, snaps)
        r = next(r for r in range(nt + 1) if beta(snaps, r) >= nt)
        assert fsteps - nt == r * nt - beta(snaps + 1, r - 1)
        print(f"nt {nt}, {snaps} slots: {len(acts)} actions, {fsteps} forward steps ({fsteps / nt:.2f} x nt)")


def test_budget_selects_frames_or_checkpoints():
    g = Geometry(nx=1024, nz=320, npml=20, dx=10.0, dz=10.0, dt=0.8e-3)
    m, sn, need = L.plan_memory(g, 64, 4000, 1024, 0)            # no budget: the frames
    assert (m, sn) == (0, 0)
    m, sn, b = L.plan_memory(g, 64, 4000, 1024, need)            # exactly enough for the frames
    assert (m, sn, b) == (0, 0, need)
    for budget, fp in ((need - 1, "fp32"), (12 << 30, "fp32"), (12 << 30, "bf16"), (8 << 30, "fp16")):
        m, sn, b = L.plan_memory(g, 64, 4000, 1024, budget, fp)
        assert m == 2 and 3 <= sn <= 128 and b <= budget, (budget, fp, m, sn, b)
    m32 = L.plan_memory(g, 64, 4000, 1024, 12 << 30, "fp32")[1]
    m16 = L.plan_memory(g, 64, 4000, 1024, 12 << 30, "bf16")[1]
    assert m16 > m32                                              # 16-bit slots are about half the size
    with pytest.raises(L.FwiError) as e:                          # not even 3 slots beside the context
        L.plan_memory(g, 64, 4000, 1024, 3 << 30)
    assert e.value.code == L.ENOMEM
