"""Nested dissection and the binary dissection tree of the parallel solver
(CSMatrix::PartBinaryTreeND, ordmatr.cpp:1204-1396 and 1698-1805).

Pins:
- find_separator equals the reference's own CSMatrix::FindSeparator
  (smatrix.cpp:4355-4488, compiled in oracle/_ref) on every node range of
  dissected grids and on patterns without a split;
- the ordering (this library's dissection; METIS_NodeND, which the
  reference calls, is not part of the image: parity unpinned against METIS)
  is checked by what a nested dissection must satisfy: a permutation, and at
  every tree node no pattern edge between the two children's row ranges;
- the ILU2 / IC2 factors of the dissected matrix equal the reference's
  Ilu2 / Ich2 on the same ordering, bit for bit.
CPU only.
"""

import numpy as np
import pytest

from test_host import _fct_params, need_ref


def sym_pattern(mo):
    """Union of the stored pattern and its transpose (SymmMtr), sorted rows."""
    sprnd, ia, ja, _ = mo.arrays()
    n = len(ia) - 1
    rows = [set() for _ in range(n)]
    for i in range(n):
        for j in ja[ia[i]:ia[i + 1]]:
            rows[i].add(int(j))
            rows[int(j)].add(i)
    sia = np.zeros(n + 1, np.int32)
    sja = []
    for i in range(n):
        r = sorted(rows[i])
        sja.extend(r)
        sia[i + 1] = len(sja)
    return sia, np.array(sja, np.int32)


def sub_pattern(ia, ja, b, e):
    """Rows and columns [b, e) of a pattern, renumbered from 0."""
    sia, sja = [0], []
    for i in range(b, e):
        cols = ja[ia[i]:ia[i + 1]]
        sja.extend(int(c) - b for c in cols if b <= c < e)
        sia.append(len(sja))
    return np.array(sia, np.int32), np.array(sja, np.int32)


def systems(B):
    out = {}
    m, _ = B.gen_heat2d(24, 20, 9, 3)
    out["heat24x20"] = m
    m, _ = B.gen_seepage3d(10, 9, 8, 6, 0.3, 4)
    out["seepage10x9x8"] = m
    m, _ = B.gen_elastic_cell(8, 0.3, 10.0, 0.3, 0.0, 5)
    out["elastic8"] = m
    return out


@pytest.fixture(scope="module")
def dissected(B):
    out = {}
    for name, m in systems(B).items():
        order, blks, b2c = m.part_binary_tree_nd(8)
        out[name] = (m, order, m.reorder(order), blks, b2c)
    return out


def test_nd_order_is_a_permutation_and_tree_covers_rows(B, dissected):
    for name, (m, order, mo, blks, b2c) in dissected.items():
        n = m.info()["nsup"]
        assert sorted(order.tolist()) == list(range(n)), name
        assert blks[0] == 0 and blks[-1] == n and np.all(np.diff(blks) > 0), name
        assert set(b2c.tolist()) <= set(range(8)), name
        if name != "elastic8":
            # every virtual rank owns a leaf; in the periodic cell a node whose
            # separator exceeds a quarter of its rows keeps them all, as the
            # reference's FindSeparator leaves such a node unsplit
            assert set(b2c.tolist()) == set(range(8)), name


def _collect_nodes(ia, ja, b, e, depth, out):
    """The node ranges the reference's PartBinaryTreeND walks: FindSeparator
    of the range, then both children (ordmatr.cpp:1204-1396)."""
    from bcmhip import find_separator

    sia, sja = sub_pattern(ia, ja, b, e)
    s1, s2 = find_separator(sia, sja)
    out.append((b, e, sia, sja, s1, s2))
    if depth == 0 or e - b < 2 or s1 == -1:
        return
    _collect_nodes(ia, ja, b, b + s1, depth - 1, out)
    _collect_nodes(ia, ja, b + s1, b + s2, depth - 1, out)


def _clean_split(ia, ja, b, s):
    """Some k in (b, s) with no pattern edge between rows [b, k) and [k, s)."""
    reach = b  # highest column below s reached by rows [b, k)
    for k in range(b + 1, s):
        cols = ja[ia[k - 1]:ia[k]]
        cols = cols[(cols >= b) & (cols < s)]
        if len(cols):
            reach = max(reach, int(cols.max()))
        if reach < k:
            return k
    return None


def test_dissection_children_are_disconnected(B, dissected):
    """At the top tree levels, the rows before each node's separator fall
    into two parts with no edge between them, and the separator (the rows
    FindSeparator's ind2 marks) is a fraction of the node. (FindSeparator's
    own ind1 may stop a few rows short of the clean split: its bisection is
    the reference's heuristic, matched separately below.)"""
    for name, (m, order, mo, blks, b2c) in dissected.items():
        ia, ja = sym_pattern(mo)
        n = len(ia) - 1
        nodes = []
        _collect_nodes(ia, ja, 0, n, 2, nodes)
        top = nodes[0]
        assert 0 < top[4] < top[5] <= n, name
        assert n - top[5] <= 0.35 * n, (name, n - top[5], n)
        for b, e, _, _, s1, s2 in nodes:
            if s1 == -1 or e - b < 2 * 64:
                continue
            assert _clean_split(ia, ja, b, b + s2) is not None, (name, b, e)


def test_find_separator_equals_reference(B, O, dissected):
    need_ref(O)
    from bcmhip import find_separator

    checked = 0
    for name, (m, order, mo, blks, b2c) in dissected.items():
        ia, ja = sym_pattern(mo)
        nodes = []
        _collect_nodes(ia, ja, 0, len(ia) - 1, 3, nodes)
        for b, e, sia, sja, _, _ in nodes:
            assert find_separator(sia, sja) == O.ref_find_separator(sia, sja), (name, b, e)
            checked += 1
        # patterns with no dissection layout: the natural and the RCM order
        for perm in (np.arange(len(order)), m.rcm_order()):
            p_ia, p_ja = sym_pattern(m.reorder(perm))
            assert find_separator(p_ia, p_ja) == O.ref_find_separator(p_ia, p_ja), name
            checked += 1
    assert checked >= 30


def test_find_separator_small_and_degenerate(B, O):
    need_ref(O)
    from bcmhip import find_separator

    rng = np.random.default_rng(9)
    for n in (0, 1, 2, 3, 4, 5, 8, 13):
        # block-diagonal pair plus a coupling row: [A][B][S]
        rows = [set([i]) for i in range(n)]
        for _ in range(2 * n):
            i, j = rng.integers(0, max(n, 1), 2) if n else (0, 0)
            if n and (i < n // 2) == (j < n // 2):
                rows[i].add(int(j))
                rows[j].add(int(i))
        ia = np.zeros(n + 1, np.int32)
        ja = []
        for i in range(n):
            ja.extend(sorted(rows[i]))
            ia[i + 1] = len(ja)
        ja = np.array(ja, np.int32)
        assert find_separator(ia, ja) == O.ref_find_separator(ia, ja), n


@pytest.mark.parametrize("kind", ["ilu2", "ic2"])
def test_factor_of_dissected_matrix_equals_reference(B, O, kind):
    """The factor the parallel solver builds on one process (the serial ILU2
    / IC2 of the dissected matrix) equals the reference's Ilu2 / Ich2 on the
    same ordering, bit for bit."""
    need_ref(O)
    tau = (1e-2, 1e-4)
    if kind == "ilu2":
        m, _ = B.gen_seepage3d(5, 4, 4, 6, 0.3, 44)
    else:
        m, _ = B.gen_elastic_cell(4, 0.35, 30.0, 0.3, 0.0, 45)
    order, _, _ = m.part_binary_tree_nd(2)
    mo = m.reorder(order)
    A = O.Block(*mo.arrays())
    if kind == "ilu2":
        f = B.Factor.ilu2(mo, _fct_params(B, tau))
        (li, lj, la), (ui, uj, ua) = O.ref_ilu2(A, tau1=tau[0], tau2=tau[1], memory=10.0)
        pairs = ((0, (li, lj, la)), (1, (ui, uj, ua)))
    else:
        f = B.Factor.ic2(mo, _fct_params(B, tau))
        pairs = ((1, O.ref_ich2(A, tau1=tau[0], tau2=tau[1], memory=10.0)),)
    for which, (ri, rj, ra) in pairs:
        ia, ja, a = f.arrays(which, A.nsup)
        assert np.array_equal(ia, ri) and np.array_equal(ja, rj), which
        assert np.array_equal(a, ra), which
