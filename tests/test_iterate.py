"""Iterate::StopList (include/GooseFEM/Iterate.h) against the reference's own cases
(tests/basic/Iterate.cpp:11-90): host logic, no GPU."""

import goosefem_amd as gf

SORTED = [5e0, 5e1, 5e-1, 5e-2, 5e-3, 5e-4, 4e-4, 3e-4, 2e-4, 1e-4]
UNSORTED = [5e0, 5e1, 5e-1, 5e-2, 5e-3, 5e-4, 4e-4, 5e-4, 4e-4, 4e-4]
UNSORTED2 = [5e0, 5e1, 5e-1, 5e-2, 5e-3, 5e-4, 4e-4, 3e-4, 2e-4, 3e-4, 2e-4, 1e-4, 9e-5, 8e-5]


def run(res, test):
    s = gf.Iterate.StopList(5)
    out = []
    for r in res:
        s.roll_insert(r)
        out.append(test(s))
    return out


def test_all_less():
    assert run(SORTED, lambda s: s.all_less(1e-3)) == [False] * 9 + [True]
    assert run(UNSORTED, lambda s: s.all_less(1e-3)) == [False] * 9 + [True]


def test_descending_and_all_less():
    assert run(SORTED, lambda s: s.descending() and s.all_less(1e-3)) == [False] * 9 + [True]
    assert run(UNSORTED2, lambda s: s.descending() and s.all_less(1e-3)) == [False] * 13 + [True]


def test_reset_and_data():
    s = gf.Iterate.StopList(3)
    assert s.data == [float("inf")] * 3 and not s.all_less(1.0)
    s.roll_insert(1.0)
    assert s.data[-1] == 1.0
    s.reset(2)
    assert s.data == [float("inf")] * 2
