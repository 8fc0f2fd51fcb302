"""Pins the oracle's LBVH topology and build counters independently of its code.

The reference splits a leaf range [b, e) of the Morton-sorted leaves at the
first leaf whose code has the highest bit in which the first and last codes
differ set, or at b + (e - b) / 2 when every code in the range is equal
(morsehgp3d/src/cpu/spatial/lbvh.cpp:831-860); nodes are numbered in strict
postorder with leaves included (lbvh.cpp:862-906); collision groups are runs
of equal codes of length > 1 (lbvh.cpp:514-530).  The checks below restate
those rules with a linear scan (not the oracle's binary search) and rebuild
the whole tree recursively, so a wrong split index, a swapped child or a
wrong counter fails here.
"""

import numpy as np
import pytest

import oracle

INV = (1 << 64) - 1


def _split(codes, b, e):
    first, last = int(codes[b]), int(codes[e - 1])
    if first == last:
        return b + (e - b) // 2
    bit = (first ^ last).bit_length() - 1
    for i in range(b + 1, e):
        if (int(codes[i]) >> bit) & 1:
            return i
    raise AssertionError("no leaf carries the differing bit")


def _expected_tree(codes):
    """(nodes as (left, right, begin, end) in postorder, maximum depth)."""
    nodes = []
    depth = [0]

    def build(b, e, d):
        depth[0] = max(depth[0], d)
        if e - b == 1:
            nodes.append((INV, INV, b, e))
            return len(nodes) - 1
        s = _split(codes, b, e)
        left = build(b, s, d + 1)
        right = build(s, e, d + 1)
        nodes.append((left, right, b, e))
        return len(nodes) - 1

    build(0, len(codes), 0)
    return nodes, depth[0]


def _collisions(codes):
    groups, biggest, i = 0, 0, 0
    while i < len(codes):
        j = i + 1
        while j < len(codes) and codes[j] == codes[i]:
            j += 1
        if j - i > 1:
            groups += 1
            biggest = max(biggest, j - i)
        i = j
    return groups, biggest


def _check_against_closed_form(pts):
    snap = oracle.lbvh_snapshot(pts)
    codes = [int(c) for c in snap.leaves[:, 0]]
    nodes, depth = _expected_tree(codes)
    got = [tuple(int(x) for x in row[6:10]) for row in snap.nodes]
    assert got == nodes
    assert int(snap.counters[2]) == depth
    assert (int(snap.counters[3]), int(snap.counters[4])) == _collisions(codes)
    return snap


def test_hand_computed_six_point_snapshot():
    # Six corners of the unit cube.  The frame is [0, 1] on every axis, so each
    # coordinate quantises to 0 or 2^21 - 1 and code = (4x + 2y + z) * K with
    # K = sum_b 8^b: Morton order = order of 4x + 2y + z = 0, 1, 3, 4, 6, 7.
    pts = oracle.canonicalize([[1, 1, 1], [0, 0, 1], [1, 0, 0], [0, 0, 0], [1, 1, 0], [0, 1, 1]])
    snap = oracle.lbvh_snapshot(pts)
    K = sum(8 ** b for b in range(21))
    assert [int(c) for c in snap.leaves[:, 0]] == [v * K for v in (0, 1, 3, 4, 6, 7)]
    assert [int(i) for i in snap.leaves[:, 1]] == [0, 1, 2, 3, 4, 5]
    # root [0,6) splits on the x bit at leaf 3; [0,3) on y at 2; [0,2) on z at
    # 1; [3,6) on y at 4; [4,6) on z at 5 (worked by hand)
    expected = [
        (INV, INV, 0, 1), (INV, INV, 1, 2), (0, 1, 0, 2), (INV, INV, 2, 3), (2, 3, 0, 3),
        (INV, INV, 3, 4), (INV, INV, 4, 5), (INV, INV, 5, 6), (6, 7, 4, 6), (5, 8, 3, 6), (4, 9, 0, 6),
    ]
    assert [tuple(int(x) for x in row[6:10]) for row in snap.nodes] == expected
    # witnesses (lower xyz, upper xyz): smallest PointId attaining the extreme
    assert [int(x) for x in snap.nodes[4][:6]] == [0, 0, 0, 0, 2, 1]  # points 0, 1, 2
    assert [int(x) for x in snap.nodes[10][:6]] == [0, 0, 0, 3, 2, 1]  # root
    assert [int(x) for x in snap.nodes[9][:6]] == [3, 3, 3, 3, 4, 5]  # points 3, 4, 5
    assert [int(c) for c in snap.counters] == [6, 11, 3, 0, 0]


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_split_rule_on_random_clouds(seed):
    rng = np.random.default_rng(seed)
    pts = oracle.canonicalize(rng.normal(size=(400, 3)) * [1.0, 3.0, 0.2])
    _check_against_closed_form(pts)


def test_split_rule_on_equal_code_runs():
    # clusters far below one Morton bin: long runs of equal codes, split at
    # their midpoints, and collision counters that are not zero
    rng = np.random.default_rng(4)
    centres = rng.random((12, 3))
    pts = np.concatenate([c + rng.random((int(rng.integers(2, 9)), 3)) * 1e-9 for c in centres])
    pts = oracle.canonicalize(pts)
    snap = _check_against_closed_form(pts)
    assert int(snap.counters[3]) == 12 and int(snap.counters[4]) >= 2
