"""Closed-form pins of the oracle's OBB construction and tree split.

Invariants (orthonormal axes, containment, leaf sizes) cannot tell which
eigenvector matrix orientation the box takes, or how a node's triangles are
partitioned.  These cases can:

* OBB axes (OBB.cpp:80-87).  eig3 writes its eigenvectors into the COLUMNS of
  V, but the reference hands glm's column-major storage to eig3 and then takes
  glm's columns, i.e. eig3's ROWS, as the box axes.  A rotated box's corners
  have a known, non-symmetric eigenvector matrix; the box axes must be its
  rows (pinned eig3), and the half-lengths the extents of the points along
  those rows plus FLT_EPSILON (OBB.cpp:105-166), not the tight extents along
  the columns.

* Split rule (OBBtree.cpp:43-108), on meshes whose trees are derived by hand.
  Every triangle lies in z = 0 and is symmetric about the x-axis,
  (a, 0, 0), (b, h, 0), (b, -h, 0) with small dyadic coordinates, so every
  node's covariance is exactly diagonal (the y terms of a triangle cancel in
  sequence, the z terms vanish), eig3 returns a permutation matrix, the box
  axes are exactly the coordinate axes and all projections below are exact:
    - tie:    the middle triangle's projection midpoint equals the centre's
              projection and goes left (`<=`);
    - empty:  the longest axis (y) puts every triangle left, so the next axis
              (x) splits; index halves would pair other triangles;
    - halves: every triangle has the same midpoints on every axis, so the node
              splits by index halves (n / 2 left), and children are swapped so
              the larger surface is on the left, at three levels;
    - swap:   the right child's box is larger, so the children trade places.
  Layout: node 0 holds the root's two children; links are (left, right,
  left count, right count), a count of 0 meaning an internal node's index,
  otherwise a triangle offset into the DFS-ordered triangles (OBBtree.cpp:321-394).

* The three-element sort of the half-lengths (OBBtree.cpp:45-51): libstdc++'s
  std::sort with the reference's comparator, compiled here, against the
  oracle's insertion sort on random, tied, signed-zero, infinite and NaN
  lengths.
"""
import ctypes as C
import os
import shutil
import subprocess

import numpy as np
import pytest

from conftest import ROOT

FLT_EPS = np.float32(np.finfo(np.float32).eps)


def _tri(a, b, h):
    return [a, 0.0, 0.0, b, h, 0.0, b, -h, 0.0]


def _mesh(rows):
    return np.array([_tri(*r) for r in rows], np.float32)


# (triangles, leaf size, expected perm, expected links) — derived by hand, see
# the module docstring and the comments on each case.
CASES = {
    # x in [0, 8], centre 4; midpoints 1, 4, 7: the tie goes left, so
    # {T0, T1} | {T2}; {T0, T1} splits at 2.5; equal leaf boxes, no swaps.
    # With `<` the tie would go right: {T0} | {T1, T2}, swapped -> [1, 2, 0].
    "tie": ([(0, 2, 1), (3, 5, 1), (6, 8, 1)], 1, [0, 1, 2], [[1, 2, 0, 1], [0, 1, 1, 1]]),
    # y half-length 16 > x half-length 4: the y split sends all four left
    # (every midpoint 0 = centre), so x splits at 4: {T0, T2} | {T1, T3};
    # {T0, T2} splits at 1.5, {T1, T3} at 6.5 (T3 left).  Index halves would
    # give {T0, T1} | {T2, T3}.
    "empty": ([(0, 2, 16), (6, 8, 16), (1, 3, 16), (5, 7, 16)], 1, [0, 2, 3, 1],
              [[1, 2, 0, 0], [0, 1, 1, 1], [2, 3, 1, 1]]),
    # every triangle spans x in [0, 16] and is symmetric in y: no axis splits,
    # so index halves: {T0, T1} | {T2, T3, T4} -> {T0} | {T1} (swapped: h 3 > 1),
    # {T2} | {T3, T4} (swapped: the pair's box, h 4, beats h 2), {T3} | {T4}
    # kept (h 4 > 0.5); at the root the h-4 side is larger and is swapped left.
    "halves": ([(0, 16, 1), (0, 16, 3), (0, 16, 2), (0, 16, 4), (0, 16, 0.5)], 1, [3, 4, 2, 1, 0],
               [[1, 3, 0, 0], [2, 2, 0, 1], [0, 1, 1, 1], [3, 4, 1, 1]]),
    # x splits at 4: T0 left, T1 right; T1's box (h 3) has the larger surface,
    # so the children trade places.
    "swap": ([(0, 2, 1), (6, 8, 3)], 1, [1, 0], [[0, 1, 1, 1]]),
    # the same halves mesh with leaf size 2: {T0, T1} is a leaf; {T2, T3, T4}
    # splits by halves into {T2} | {T3, T4}, swapped (the pair is larger);
    # the root swaps as before.
    "halves_L2": ([(0, 16, 1), (0, 16, 3), (0, 16, 2), (0, 16, 4), (0, 16, 0.5)], 2, [3, 4, 2, 0, 1],
                  [[1, 3, 0, 2], [0, 2, 2, 1]]),
}


@pytest.mark.parametrize("name", sorted(CASES))
def test_split_rule_hand_derived(orc, name):
    rows, leaf, perm, links = CASES[name]
    tri = _mesh(rows)
    ex = orc.tree(tri, leaf).export()
    assert ex["perm"].tolist() == perm, name
    assert ex["links"].tolist() == links, name


@pytest.mark.parametrize("name", sorted(CASES))
def test_split_case_boxes_are_axis_aligned(orc, name):
    """The premise of the hand derivations: every box of these meshes has
    the coordinate axes (up to order) as exact side directions."""
    rows, leaf, _, _ = CASES[name]
    ex = orc.tree(_mesh(rows), leaf).export()
    boxes = [ex["root"]] + [o[:12] for o in ex["obbs"]] + [o[12:] for o in ex["obbs"]]
    for b in boxes:
        sides = np.asarray(b[3:12], np.float64).reshape(3, 3)
        assert all(np.count_nonzero(s) == 1 for s in sides)
        assert sorted(int(np.flatnonzero(s)[0]) for s in sides) == [0, 1, 2]


def _rotated_box_corners(half, q):
    x, y, z, w = q / np.linalg.norm(q)
    R = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                  [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                  [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])
    corners = np.array([[sx, sy, sz] for sx in (-1, 1) for sy in (-1, 1) for sz in (-1, 1)], np.float64)
    return (corners * np.asarray(half)) @ R.T


@pytest.mark.parametrize("q", [(0.3, -0.5, 0.7, 0.4), (0.9, 0.1, -0.2, 0.35), (-0.2, 0.6, 0.1, 0.75)])
def test_box_axes_are_rows_of_eigenvectors(orc, q):
    pts = _rotated_box_corners((3.0, 2.0, 1.0), np.array(q)).astype(np.float32)
    p = pts.astype(np.float64)
    A = np.cov(p.T, bias=True)  # the covariance the box is built from, divided by N
    V, d = orc.eig3(A)          # eig3 is pinned bit for bit against the compiled reference
    assert np.allclose(d, [1.0, 4.0, 9.0], atol=1e-5)  # eigenvalues ascending, in the columns of V
    assert np.abs(V - V.T).max() > 0.1  # rows and columns differ: the case decides between them
    box = orc.box_from_points(pts).astype(np.float64)
    c, sides = box[:3], box[3:].reshape(3, 3)
    for k in range(3):
        row = V[k]
        half = np.abs(p @ row).max() + float(FLT_EPS)  # symmetric points: (max - min) / 2 + eps
        assert np.allclose(sides[k], half * row, atol=1e-5), k
        col = V[:, k]
        assert not np.allclose(np.abs(sides[k] / np.linalg.norm(sides[k])), np.abs(col), atol=1e-3)
    assert np.abs(c).max() < 1e-5
    # the rows' box is not the tight one: its volume exceeds the corners' box
    assert np.prod([np.linalg.norm(s) for s in sides]) > 3.0 * 2.0 * 1.0 * 1.05


def test_sort3_oracle_call(orc):
    out = (C.c_int32 * 3)()
    orc.lib.orc_sort3_desc(np.array([1.0, 3.0, 2.0], np.float32), out)
    assert list(out) == [1, 2, 0]
    orc.lib.orc_sort3_desc(np.array([2.0, 2.0, 2.0], np.float32), out)
    assert list(out) == [0, 1, 2]  # ties keep u, v, w order


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++ for libstdc++'s std::sort")
def test_sort3_matches_std_sort(tmp_path, orc):
    vals = [0.0, -0.0, 1.0, 2.0, 2.0, 3.5, float("inf"), float("-inf"), float("nan"), 1e-30, 7.0]
    rng = np.random.default_rng(7)
    triples = [(a, b, c) for a in vals for b in vals for c in vals]
    triples += [tuple(x) for x in rng.choice([0.5, 1.0, 1.5], size=(200, 3))]
    triples += [tuple(x) for x in rng.uniform(0, 4, size=(200, 3)).astype(np.float32)]
    arr = np.array(triples, np.float32)
    inp = tmp_path / "in.bin"
    arr.tofile(inp)
    src = tmp_path / "sort3.cpp"
    src.write_text(r'''
#include <algorithm>
#include <array>
#include <cstdio>
#include <utility>
#include <vector>
int main(int argc, char** argv) {
    std::FILE* f = std::fopen(argv[1], "rb");
    float t[3];
    while (std::fread(t, sizeof(float), 3, f) == 3) {
        // OBBtree.cpp:45-51: pairs of (half-length, axis), the comparator on the length
        std::vector<std::pair<float, std::array<float, 3>>> v;
        for (int k = 0; k < 3; ++k) v.emplace_back(t[k], std::array<float, 3>{float(k), 0.f, 0.f});
        std::sort(v.begin(), v.end(),
                  [](std::pair<float, std::array<float, 3>> a, std::pair<float, std::array<float, 3>> b) {
                      return a.first > b.first;
                  });
        std::printf("%d %d %d\n", int(v[0].second[0]), int(v[1].second[0]), int(v[2].second[0]));
    }
    return 0;
}
''')
    exe = tmp_path / "sort3"
    subprocess.run(["g++", "-O1", "-std=c++17", str(src), "-o", str(exe)], check=True)
    lines = subprocess.run([str(exe), str(inp)], check=True, capture_output=True, text=True).stdout.splitlines()
    assert len(lines) == len(arr)
    out = (C.c_int32 * 3)()
    for t, line in zip(arr, lines):
        orc.lib.orc_sort3_desc(np.ascontiguousarray(t), out)
        assert list(out) == [int(x) for x in line.split()], t
