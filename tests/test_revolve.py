"""The binomial checkpoint plan (nim_revolve_plan, a host function): every
step is reversed once, in order nsteps-1 .. 0, with the state before it
current, within the snapshot budget, and in Griewank's optimal number of
forward steps r l - C(c + r, c + 1) for l = C(c + r, c) steps and c
snapshots; for other lengths never more than the next such length needs."""
from math import comb

import pytest

import nim_amd


def simulate(nsteps, nslots):
    plan, fwd = nim_amd.revolve_plan(nsteps, nslots)
    cur, slots, reversed_, steps = 0, {}, [], 0
    for code, a in plan:
        if code == nim_amd.RV_STORE:
            assert 0 <= a < nslots
            slots[a] = cur
        elif code == nim_amd.RV_RESTORE:
            cur = slots[a]
        elif code == nim_amd.RV_ADVANCE:
            assert a > 0
            cur += a
            steps += a
            assert cur < nsteps
        else:
            assert code == nim_amd.RV_REVERSE and a == cur
            reversed_.append(a)
    assert reversed_ == list(range(nsteps - 1, -1, -1))
    assert steps == fwd
    return fwd


@pytest.mark.parametrize("c", [1, 2, 3, 4, 5])
@pytest.mark.parametrize("r", [1, 2, 3, 4])
def test_griewank_optimum(c, r):
    l = comb(c + r, c)
    assert simulate(l, c) == r * l - comb(c + r, c + 1)


def test_lengths_between():
    # every length is reversed correctly, in at most r n forward steps for
    # the smallest r with C(c + r, c) >= n, and never fewer for more steps
    for c in (1, 3, 20):
        prev = -1
        for n in list(range(1, 40)) + [100, 231, 232, 400]:
            r = next(r for r in range(0, 5000) if comb(c + r, c) >= n)
            f = simulate(n, c)
            assert f <= r * n, (n, c, f)
            assert f >= prev
            prev = f
    assert simulate(12, 1) == 12 * 11 // 2


def test_bad_arguments():
    with pytest.raises(RuntimeError):
        nim_amd.revolve_plan(0, 3)
