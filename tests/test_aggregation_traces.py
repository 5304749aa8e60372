"""The greedy aggregation pinned by hand-traced cases.

Each expected result below was traced by hand through the reference's
SpMtx_aggregate (src/datatypes/SpMtx/SpMtx_aggregation.F90:49-641) and
lets_colour / can_add_roundlayer (src/datatypes/Aggregate_mod.F90:287-497),
step by step from the Fortran, and only then compared with the oracle.  The
comments give each trace's deciding steps.  Nodes are numbered
lexicographically, node (r, c) = (r - 1) * m + c on an m x m grid.

Every case runs through the oracle (oracle/doug_oracle.c) and, for the
fine-level cases, through the library's host setup.  Builds of the oracle
with one planted mistake each (-DOR_MUTANT=k) must fail at least one case.
"""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

import doug_amd as D
import oracle as O

ORACLE_DIR = os.path.dirname(O.__file__)


def grid(m, weights=None):
    """5-point m x m matrix in ascending (row, column) order: diagonal d,
    horizontal neighbours h, vertical neighbours v (weights = (d, h, v));
    the default is the Laplacian (4, -1, -1)."""
    d, h, v = weights or (4.0, -1.0, -1.0)
    ii, jj, vv = [], [], []
    for r in range(m):
        for c in range(m):
            k = r * m + c
            for dr, dc, w in ((-1, 0, v), (0, -1, h), (0, 0, d), (0, 1, h), (1, 0, v)):
                if 0 <= r + dr < m and 0 <= c + dc < m:
                    ii.append(k + 1)
                    jj.append((r + dr) * m + c + dc + 1)
                    vv.append(w)
    return m * m, np.array(ii, np.int32), np.array(jj, np.int32), np.array(vv)


# (name, matrix, radius, minasize, maxasize, alpha, expected inner numbers)
CASES = []

# 5x5 Laplacian, radius 1, all connections strong (every scaled entry is
# -1/4), default sizes 4 and 9.
#  seed 1: layer 0 {2,6} (weights 1: allmind = 1); layer 1 sees 3, 7 (twice),
#    11: mind 1 < maxd 2, so 7 is a rounder and joins layer 1; its free
#    neighbours 8, 12 go to the next layer.  Aggregate {1,2,6,7}.  Outer
#    layer 4 (5 nodes, more than layer 3's 4): 4 and 16 at distance 3 are
#    the next good starts.
#  seed 4: layer 1 sees 8 and 10 twice: both rounders -> {3,4,5,8,9,10}; its
#    outer layer 4 adds 19 (distance 3).
#  seed 16: rounders 12, 22 -> {11,12,16,17,21,22}; outer layer 3 adds 18.
#  seed 19 (the first good start left): layer 1 sees 13, 15, 23, 25 twice
#    each: mind = maxd = 2 but mind > allmind = 1 (the minimum kept from the
#    earlier layers), so all four are rounders -> the 9 nodes
#    {13..15,18..20,23..25}.
CASES.append(("5x5 radius 1", grid(5), 1, 4, 9, 0.5,
              [1, 1, 2, 2, 2,
               1, 1, 2, 2, 2,
               3, 3, 4, 4, 4,
               3, 3, 4, 4, 4,
               3, 3, 4, 4, 4]))

# The same grid with maxasize 5 and minasize 4.
#  seed 1: {1,2,6,7} as above (4 <= 5).
#  seed 4: rounder 8 fits (1 + 3 + 1 = 5), rounder 10 does not (6 > 5): the
#    layer fails (ok = 1).  Nodes with stat <= 2 among the first 5 found are
#    still aggregated: 4 (layer 0) and the rounder 8 (stat -2) -> {4,8}; 10
#    and 14 keep their negative marks.
#  seed 16: likewise {12,16} (rounder 12 fits, 22 does not).
#  then singletons {3}, {5}, {9} (their neighbours are aggregated or marked;
#    {3} sets `toosmall`), {10,15} (seed 10 with stat -4, outer 19, 25),
#    {19,20,24,25} (rounder 25), {23}, {11}, {13}, {14}, {17}, {18}, {21},
#    {22}: 16 aggregates.
#  merge (all connections, |a~| = 1/4, alpha 0.5, beta = alpha/4, eater size
#    <= 5): {4,8} touches {3} twice (0.5 >= alpha: eaten by {3});
#    {12,16} is eaten by {11} (0.5); {3,4,8} by {9} (0.5, {1,2,6,7} is too
#    big); {5} (0.25 < alpha) is handed to {9,3,4,8} through 4 (0.25 >= beta);
#    {10,15}: 15 to {14}, then 10 to {9,...}; {23} to {18}; {11,12,16} eaten
#    by {17} (0.5); {13} handed to {9,...}; {14,15} handed to it node by node;
#    {18,23}: 23 to {22}, 18 to {9,...}; {21} to {17,...}; {22,23}: 23 and 22
#    handed on.  Renumbered in node order.
CASES.append(("5x5 maxasize 5", grid(5), 1, 4, 5, 0.5,
              [1, 1, 2, 2, 2,
               1, 1, 2, 2, 2,
               3, 3, 2, 2, 2,
               3, 3, 2, 4, 4,
               3, 3, 2, 4, 4]))

# 7x7 Laplacian, radius 1, sizes 4 and 9.
#  seeds 1 -> {1,2,8,9}, good starts 4, 22; 4 -> {3,4,5,10,11,12}, adds 7, 25;
#  22 -> {15,16,22,23,29,30}, adds 25, 43: the list is now 7, 25, 25, 43 and
#  the repeated 25 is taken before the first entry 7 -> the 9 nodes
#  {17..19,24..26,31..33} (rounders 17, 19, 31, 33); then 7 -> {6,7,13,14},
#  43 -> {36,37,43,44}, 28 -> {20,21,27,28,34,35}, 46 -> {38..40,45..47},
#  49 -> {41,42,48,49} (rounder 41 by mind 2 > allmind 1).
CASES.append(("7x7 radius 1", grid(7), 1, 4, 9, 0.5,
              [1, 1, 2, 2, 2, 5, 5,
               1, 1, 2, 2, 2, 5, 5,
               3, 3, 4, 4, 4, 7, 7,
               3, 3, 4, 4, 4, 7, 7,
               3, 3, 4, 4, 4, 7, 7,
               6, 6, 8, 8, 8, 9, 9,
               6, 6, 8, 8, 8, 9, 9]))

# 7x7 anisotropic grid: diagonal 2.2, horizontal -1, vertical -0.1; radius
# 2, default sizes 12 and 25, strong1 0.2.  Scaled |a~| is 1/2.2 across a
# row and 0.1/2.2 between rows: 0.1/2.2 > 0.2 * 1/2.2 fails, so the strong
# graph is the 7 rows as separate chains.
#  each row: seed (r,1) explores 5 layers along the chain -> aggregate of
#    layers 0..2 {c1,c2,c3}, outer (r,4) -> {c4,c5,c6}, then {c7}; every
#    aggregate has fewer explored nodes than minasize, so `toosmall`.
#  merge (alpha 0.2 over all connections, eater size <= 25): inside a row the
#    strong link (0.45) wins: {c1..c3} eaten by {c4..c6}, that by {c7};
#    the full row then sums 7 weak links (7 x 0.045 = 0.32 >= 0.2) to the
#    row above: rows 2 and 3 join row 1's aggregate, rows 5 and 6 row 4's; a
#    third row would exceed 25 (21 + 7), and the row's hand-over fails on
#    beta = 0.05 > 0.045, so row 4 and row 7 start new aggregates.
CASES.append(("7x7 weak rows radius 2", grid(7, (2.2, -1.0, -0.1)), 2, 12, 25, 0.2,
              [1] * 21 + [2] * 21 + [3] * 7))


def _fine(L, case):
    name, (n, i, j, v), radius, mina, maxa, alpha, _ = case
    O.reset_process(L)
    _, inner, _ = O.aggregate(n, i, j, v, radius, mina, maxa, alpha, L=L)
    return inner


# Coarse-level calls inside the same process as a fine-level one (the case
# "5x5 maxasize 5" above leaves allmind = 1 and beta = 0.5 / 4 = 0.125).
# The coarse matrix: a path 1-2-3-4-5-6, diagonal 2.5, links -1 (scaled
# -0.4) except 5-6 with -0.1 (scaled -0.04: weak in row 5 against 0.6 x 0.4,
# strong in row 6, its only link); the entry (1,2) is stored twice.
def _coarse_path():
    e = [(1, 1, 2.5), (1, 2, -1.0), (1, 2, -1.0), (2, 1, -1.0), (2, 2, 2.5), (2, 3, -1.0),
         (3, 2, -1.0), (3, 3, 2.5), (3, 4, -1.0), (4, 3, -1.0), (4, 4, 2.5), (4, 5, -1.0),
         (5, 4, -1.0), (5, 5, 2.5), (5, 6, -0.1), (6, 5, -0.1), (6, 6, 2.5)]
    i, j, v = zip(*e)
    return 6, np.array(i, np.int32), np.array(j, np.int32), np.array(v)


# (name, minasize, expected inner numbers)
#  saved allmind: seed 1 sees 2 twice (the stored duplicate): mind = maxd =
#    2 > allmind = 1, so 2 is a rounder in layer 0 and 3 joins in layer 1 ->
#    {1,2,3}; 4 -> {4,5}; {6}.  (A fresh process would take allmind = 2 and
#    give {1,2}, {3,4}, {5}, {6}.)  minasize 1: no merge.
#  saved beta: minasize 3 makes {4,5} too small; its best eater {6} has
#    weight 0.04 < alpha and < beta = 0.125, so it is handed to {1,2,3}
#    node by node (0.4 >= beta); {6} stays (0.04 < beta).  (With beta 1e-5
#    of a fresh process {6} would eat {4,5}.)
COARSE_CASES = [("saved allmind", 1, [1, 1, 1, 2, 2, 3]),
                ("saved beta", 3, [1, 1, 1, 1, 1, 2])]


def _coarse(L, minasize):
    _fine(L, CASES[1])
    n, i, j, v = _coarse_path()
    _, inner, _ = O.aggregate(n, i, j, v, 1, minasize, 4, 0.6, coarse=True, L=L)
    return inner


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_oracle_matches_hand_trace(case):
    np.testing.assert_array_equal(_fine(O.lib(), case), case[-1])


@pytest.mark.parametrize("case", COARSE_CASES, ids=[c[0] for c in COARSE_CASES])
def test_oracle_coarse_call_keeps_process_state(case):
    np.testing.assert_array_equal(_coarse(O.lib(), case[1]), case[2])


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_library_matches_hand_trace(case):
    # the same aggregation inside the library's set-up (host-only context)
    name, (n, i, j, v), radius, mina, maxa, alpha, expected = case
    with D.Solver(-1) as s:
        s.set_params(radius1=radius, minasize1=mina, maxasize1=maxa, strong1=alpha,
                     smoothers=1, symmstruct=1, symmnumeric=1)
        s.set_matrix(n, i, j, v, 1)
        s.setup_host()
        num, _, _ = s.fine_aggregates()
    np.testing.assert_array_equal(num, expected)


def test_planted_mistakes_fail_the_traces(tmp_path):
    # one build of the oracle per planted mistake; each must break a case
    src = os.path.join(ORACLE_DIR, "doug_oracle.c")
    procs = {}
    for k in range(1, 9):
        out = str(tmp_path / f"liboracle_m{k}.so")
        procs[k] = (out, subprocess.Popen(
            ["gcc", "-O1", "-std=c11", "-fPIC", "-fopenmp", "-ffp-contract=off",
             f"-DOR_MUTANT={k}", "-shared", "-o", out, src, "-lm"]))
    for k, (out, pr) in procs.items():
        assert pr.wait() == 0, k
        L = O.bind(C.CDLL(out))
        caught = [c[0] for c in CASES if not np.array_equal(_fine(L, c), c[-1])]
        caught += [c[0] for c in COARSE_CASES if not np.array_equal(_coarse(L, c[1]), c[2])]
        print(f"mutant {k}: fails {caught}")
        assert caught, f"planted mistake {k} passes every hand-traced case"
