"""Orderings: data_reorder_cuthill_mckee / data_sparse_reorder
(src/cuthill_mckee.cpp:53-265, src/reorder.cpp:167-253) on the CPU.

The reference cannot be compiled here (its headers need MKL), so the
quirk-for-quirk restatement is pinned by hand-traced cases: a 5-node star
traced line by line through the reference loop, and the even-n reverse
(middle pair swapped twice).  The product is then compared with the
oracle restatement on stencil matrices, and the permuted matrix with
A[P][:, P]."""
import ctypes as C

import numpy as np
import pytest


def _star5(orc):
    rows = [[0, 2], [1, 2], [0, 1, 2, 3, 4], [2, 3], [2, 4]]
    row = np.cumsum([0] + [len(r) for r in rows]).astype(np.int32)
    col = np.array([c for r in rows for c in r], np.int32)
    val = np.arange(1, len(col) + 1, dtype=np.float64)
    return orc.CSR(5, 5, row, col, val)


def _cm(lace, A, reverse):
    H = lace.HostCSR(A.row, A.col, A.val)
    m = np.zeros(A.n_rows, np.int32)
    rc = lace.lib().data_reorder_cuthill_mckee(H.ref(), m.ctypes.data_as(C.POINTER(C.c_int)), reverse)
    return rc, m


def test_star_hand_trace(lace, orc):
    """Star centred at 2: the seed is 2 (first row above row 1's degree);
    the window [nnsn[1], nnsn[2]) = nsn[2..5] = (2, 0, 1, 3) drops the
    seed's last neighbour 4; degrees of v+1: node 0 -> 1, node 1 -> 4,
    node 3 -> 1, so R sorts to (1, 0, 3); seeds 1 and 0 add nothing; seed 3
    reaches 4 through nsn[6]."""
    A = _star5(orc)
    want_cm = [2, 1, 0, 3, 4]
    want_rcm = [4, 3, 0, 1, 2]  # odd n: positions 0<->4, 1<->3, 2<->2
    assert list(orc.cuthill_mckee(A)) == want_cm
    assert list(orc.cuthill_mckee(A, reverse=True)) == want_rcm
    for rev, want in ((False, want_cm), (True, want_rcm)):
        rc, m = _cm(lace, A, rev)
        assert rc == 0 and list(m) == want


def test_even_n_reverse_keeps_middle_pair(orc):
    """n = 4: the swap loop runs i = 0, 1, 2, so positions 1 and 2 are
    swapped and swapped back."""
    rows = [[0, 1, 2, 3], [0, 1], [0, 2], [0, 3]]
    row = np.cumsum([0] + [len(r) for r in rows]).astype(np.int32)
    col = np.array([c for r in rows for c in r], np.int32)
    A = orc.CSR(4, 4, row, col, np.ones(len(col)))
    cm = orc.cuthill_mckee(A)
    rcm = orc.cuthill_mckee(A, reverse=True)
    assert cm is not None and rcm is not None
    assert [rcm[0], rcm[1], rcm[2], rcm[3]] == [cm[3], cm[1], cm[2], cm[0]]


def test_path_graph_queue_runs_out(lace, orc):
    """Tridiagonal 4 x 4: the seed is row 0 and the window for row 0 holds
    only nsn[0] = 0 itself, so the reference would read an uninitialised
    queue entry next; the library reports an error instead."""
    row = np.array([0, 2, 5, 8, 10], np.int32)
    col = np.array([0, 1, 0, 1, 2, 1, 2, 3, 2, 3], np.int32)
    A = orc.CSR(4, 4, row, col, np.ones(10))
    assert orc.cuthill_mckee(A) is None
    rc, _ = _cm(lace, A, False)
    assert rc != 0 and b"uninitialised" in lace.lib().lace_last_error()


@pytest.mark.parametrize("gen", ["lap2d", "conv3d", "trefethen"])
def test_product_matches_oracle_and_permutes(lace, orc, gen):
    import os
    from conftest import GOLDEN
    if gen == "lap2d":
        A = orc.laplace2d(9, 7)
    elif gen == "conv3d":
        A = orc.convdiff3d(6, 5, 4)
    else:
        A = orc.read_mtx(os.path.join(GOLDEN, "Trefethen_20.mtx"))
    for rev in (False, True):
        want = orc.cuthill_mckee(A, reverse=rev)
        rc, m = _cm(lace, A, rev)
        if want is None:
            assert rc != 0
            continue
        assert rc == 0 and np.array_equal(m, want)
        assert sorted(m) == list(range(A.n_rows))
        # data_sparse_reorder: row i of the result is row P[i] of A.  It
        # frees and replaces A's arrays (as the reference does), so A is a
        # library-allocated copy.
        H = lace.HostCSR(A.row, A.col, A.val)
        M = lace.DMatrix()
        lace.check(lace.lib().data_zmconvert(H.m, C.byref(M), lace.Magma_CSR, lace.Magma_CSR))
        P = np.zeros(A.n_rows, np.int32); Pinv = np.zeros(A.n_rows, np.int32)
        ip = C.POINTER(C.c_int)
        rc = lace.lib().data_sparse_reorder(C.byref(M), P.ctypes.data_as(ip), Pinv.ctypes.data_as(ip),
                                            2 if rev else 1)
        assert rc == 0 and np.array_equal(P, want)
        assert np.array_equal(Pinv[P], np.arange(A.n_rows))
        n = A.n_rows
        row, col, val = lace.owned_to_numpy(M)
        B = orc.CSR(n, n, row, col, val)
        assert np.array_equal(B.dense(), A.dense()[np.ix_(P, P)])
        assert all(np.all(np.diff(col[row[i]:row[i + 1]]) > 0) for i in range(n))


def test_amd_not_provided(lace, orc):
    A = orc.laplace2d(4, 4)
    H = lace.HostCSR(A.row, A.col, A.val)
    P = np.zeros(16, np.int32); Pinv = np.zeros(16, np.int32)
    ip = C.POINTER(C.c_int)
    # rejected before A's arrays are touched
    assert lace.lib().data_sparse_reorder(H.ref(), P.ctypes.data_as(ip), Pinv.ctypes.data_as(ip), 3) != 0


def _rcm_lib(lace, A):
    H = lace.HostCSR(A.row, A.col, A.val)
    P = np.zeros(A.n_rows, np.int32)
    lace.check(lace.lib().lace_reorder_rcm(H.ref(), P.ctypes.data_as(C.POINTER(C.c_int))))
    return P


def _bandwidth(A, P):
    pos = np.empty_like(P)
    pos[P] = np.arange(len(P))
    ri = A.rowidx()
    return int(np.abs(pos[ri] - pos[A.col]).max())


def _graphs(orc):
    import os
    from conftest import GOLDEN
    rng = np.random.default_rng(3)
    # a scrambled path and a second component (a scrambled star): RCM must
    # lay the path out contiguously
    n = 40
    lab = rng.permutation(n)
    rows, cols = list(range(n)), list(range(n))
    for a, b in zip(lab[:29], lab[1:30]):
        rows += [a, b]; cols += [b, a]
    for leaf in lab[31:]:
        rows += [lab[30], leaf]; cols += [leaf, lab[30]]
    M = orc._coo_to_csr(n, np.array(rows), np.array(cols), np.ones(len(rows)))
    return {"lap2d": orc.laplace2d(9, 7), "conv3d": orc.convdiff3d(6, 5, 4),
            "trefethen": orc.read_mtx(os.path.join(GOLDEN, "Trefethen_20.mtx")), "path_star": M}


def test_textbook_rcm_matches_oracle(lace, orc):
    for name, A in _graphs(orc).items():
        P = _rcm_lib(lace, A)
        assert np.array_equal(P, orc.rcm_textbook(A)), name
        assert sorted(P) == list(range(A.n_rows))


def test_textbook_rcm_properties(orc):
    """The restatement against what RCM fixes: a path comes out contiguous
    (bandwidth 1 within it), the Cuthill-McKee order (before reversal) runs
    through each component by non-decreasing BFS distance from its start,
    and on the grid and path graphs the bandwidth is no worse than scipy's
    RCM (RCM is a heuristic: on the dense Trefethen graph scipy's
    minimum-degree start happens to do better)."""
    import scipy.sparse.csgraph as cg
    for name, A in _graphs(orc).items():
        P = orc.rcm_textbook(A)
        S = A.scipy()
        G = ((S + S.T) != 0).astype(np.int8).tocsr()
        cm = P[::-1]
        # components in CM order: distances from each component's first node
        seen = np.zeros(A.n_rows, bool)
        k = 0
        while k < A.n_rows:
            d = cg.shortest_path(G, unweighted=True, indices=cm[k])
            comp = np.isfinite(d)
            size = int(comp.sum())
            block = cm[k:k + size]
            assert comp[block].all() and not seen[block].any(), name
            assert np.all(np.diff(d[block]) >= 0), name
            seen[block] = True
            k += size
        if name != "trefethen":
            ref = cg.reverse_cuthill_mckee(S.tocsr(), symmetric_mode=False)
            assert _bandwidth(A, P) <= _bandwidth(A, np.asarray(ref, np.int32)), name
    A = _graphs(orc)["path_star"]
    P = orc.rcm_textbook(A)
    pos = np.empty_like(P)
    pos[P] = np.arange(len(P))
    ri = A.rowidx()
    _, comp = cg.connected_components(A.scipy(), directed=False)
    path = np.bincount(comp).argmax()  # the 30-node path outnumbers the star
    path_edges = [(r, c) for r, c in zip(ri, A.col) if r != c and comp[r] == path]
    assert len(path_edges) == 58
    assert max(abs(pos[r] - pos[c]) for r, c in path_edges) == 1


def test_sparse_reorder_textbook_rcm(lace, orc):
    A = orc.convdiff3d(6, 5, 4)
    H = lace.HostCSR(A.row, A.col, A.val)
    M = lace.DMatrix()
    lace.check(lace.lib().data_zmconvert(H.m, C.byref(M), lace.Magma_CSR, lace.Magma_CSR))
    P = np.zeros(A.n_rows, np.int32); Pinv = np.zeros(A.n_rows, np.int32)
    ip = C.POINTER(C.c_int)
    lace.check(lace.lib().data_sparse_reorder(C.byref(M), P.ctypes.data_as(ip), Pinv.ctypes.data_as(ip), 4))
    assert np.array_equal(P, orc.rcm_textbook(A))
    row, col, val = lace.owned_to_numpy(M)
    B = orc.CSR(A.n_rows, A.n_rows, row, col, val)
    assert np.array_equal(B.dense(), A.dense()[np.ix_(P, P)])
    # the 3-D grid: textbook RCM narrows the band, the reference's ordering widens it
    assert _bandwidth(A, P) < _bandwidth(A, np.arange(A.n_rows, dtype=np.int32))


def test_reference_ordering_breaks_synchronous_sweeps(lace, orc):
    """A finding about the RCM pre-step: on the reference's ordering
    (data_sparse_reorder 2, with its quirks) the synchronous ParILU sweeps of
    the 3-D convection-diffusion operator overflow within three sweeps,
    although exact ILU(0) of the same matrix has no zero pivot; on the
    textbook RCM ordering they converge like the natural order."""
    A = orc.convdiff3d(16, 14, 12)
    out = {}
    for name, reorder in (("reference", 2), ("textbook", 4)):
        H = lace.HostCSR(A.row, A.col, A.val)
        M = lace.DMatrix()
        lace.check(lace.lib().data_zmconvert(H.m, C.byref(M), lace.Magma_CSR, lace.Magma_CSR))
        P = np.zeros(A.n_rows, np.int32); Pinv = np.zeros(A.n_rows, np.int32)
        ip = C.POINTER(C.c_int)
        lace.check(lace.lib().data_sparse_reorder(C.byref(M), P.ctypes.data_as(ip), Pinv.ctypes.data_as(ip),
                                                  reorder))
        row, col, val = lace.owned_to_numpy(M)
        B = orc.CSR(A.n_rows, A.n_rows, row, col, val)
        st = orc.ParILUState(B)
        nl = []
        with np.errstate(all="ignore"):
            for _ in range(3):
                st.sweep("jacobi")
                nl.append(orc.ilures(B, *st.factors())[1])
        ex, bad = orc.ilu0(B)
        out[name] = (nl, bad, _bandwidth(A, P))
    assert out["reference"][1] < 0 and not np.isfinite(out["reference"][0][-1])
    assert out["textbook"][1] < 0 and np.all(np.isfinite(out["textbook"][0]))
    assert out["textbook"][0][-1] < out["textbook"][0][0]
    natural = _bandwidth(A, np.arange(A.n_rows, dtype=np.int32))
    assert out["textbook"][2] < natural < out["reference"][2]
