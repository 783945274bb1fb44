"""Value updates on a fixed pattern: the ParILU handle and the solver take new
values of A, and the next setup factors those values, not the ones given at
create.  The autograd binding keeps each node's own values, so one solver
can serve several nodes of a graph."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _perturbed(A, seed=21):
    """A's values scaled by 1 + 0.05 N(0,1) entrywise, diagonal + 2 (the
    7-point operator stays diagonally dominant)."""
    rng = np.random.default_rng(seed)
    ri = np.repeat(np.arange(A.n_rows), np.diff(A.row))
    return A.val * (1.0 + 0.05 * rng.standard_normal(A.nnz)) + 2.0 * (A.col == ri)


@pytest.mark.parametrize("prec", ["f64", "f32", "bf16"])
@pytest.mark.parametrize("name", ["convdiff", "nonsym_pattern"])
def test_parilu_set_values_then_reset(gpu, orc, prec, name):
    """create(A) -> set_values(A') -> reset -> k sweeps equals the oracle's
    sweeps on A' bit for bit (the symmetric-pattern kernels read their own
    position-ordered copies of A, which must follow too)."""
    A = orc.convdiff3d(10, 9, 8)
    if name == "nonsym_pattern":
        # drop one strictly upper entry per row where there is one: the
        # general CSR-L / CSC-U kernel
        keep = np.ones(A.nnz, bool)
        ri = np.repeat(np.arange(A.n_rows), np.diff(A.row))
        ups = np.nonzero(A.col > ri + 1)[0]
        keep[ups[::3]] = False
        rows = ri[keep]
        A = orc._coo_to_csr(A.n_rows, rows, A.col[keep], A.val[keep])
    vals = _perturbed(A)
    if prec == "bf16":
        # new row maxima: the 16-bit row scales follow the new values
        vals = vals * 2.0 ** (np.repeat(np.arange(A.n_rows), np.diff(A.row)) % 5)
    lp, dt = {"f64": (gpu.LACE_F64, np.float64), "f32": (gpu.LACE_F32, np.float32),
              "bf16": (gpu.LACE_BF16, np.float32)}[prec]
    P = gpu.ParILU(gpu.HostCSR.from_csr(A), lp)
    P.sweeps(2)
    P.set_values(vals)
    P.reset()
    A2 = orc.CSR(A.n_rows, A.n_cols, A.row, A.col, vals)
    st = orc.ParILUState(A2, dt, storage="bf16" if prec == "bf16" else None)
    for k in range(4):
        P.sweeps(1)
        st.sweep("jacobi")
    (_, _, lv), (_, _, uv) = P.factors()
    assert np.array_equal(lv.astype(dt), st.Lval) and np.array_equal(uv.astype(dt), st.Uval)
    # ||A'||_F scales the step the drop-in stop rules use
    s = P.sweeps(1, want_step=True)
    st2 = orc.ParILUState(A2, dt, storage="bf16" if prec == "bf16" else None)
    for _ in range(4):
        st2.sweep("jacobi")
    step = st2.sweep("jacobi") / np.linalg.norm(vals)
    assert abs(s - step) <= 1e-6 * step + 1e-300


@pytest.mark.parametrize("inner", [0, 8])
def test_solver_set_values_then_setup(gpu, orc, inner):
    """Solver created with A, given A' and set up, solves exactly as a
    solver created with A' does: same iterations, bit-identical x."""
    A = orc.convdiff3d(12, 11, 10)
    vals = _perturbed(A, 22)
    A2 = orc.CSR(A.n_rows, A.n_cols, A.row, A.col, vals)
    b = gpu.gen_normal(5, A.n_rows)

    def opts():
        o = gpu.default_opts()
        o.search_max, o.restart_max, o.rtol, o.parilu_sweeps = 20, 50, 1e-10, 3
        o.trsv, o.jacobi_sweeps = gpu.TRSV_JACOBI, 5
        o.prec_precond = gpu.LACE_F32 if inner else gpu.LACE_F64
        o.inner_iters = inner
        return o

    S1 = gpu.Solver(gpu.HostCSR.from_csr(A), opts())
    S1.setup()
    S1.solve(b)  # factors and workspace of A in use first
    S1.set_values(vals)
    S1.setup()
    x1, st1 = S1.solve(b)
    S2 = gpu.Solver(gpu.HostCSR.from_csr(A2), opts())
    S2.setup()
    x2, st2 = S2.solve(b)
    info = (st1.converged, st1.iterations, st1.rel_true_residual, st2.converged, st2.iterations,
            st2.rel_true_residual)
    assert st1.converged == 1 and st2.converged == 1, info
    assert st1.iterations == st2.iterations, info
    assert np.array_equal(x1, x2)
    assert np.linalg.norm(b - A2.scipy() @ x1) <= 1.01e-10 * np.linalg.norm(b)
    # the transposed operator follows as well
    l1, _ = S1.solve(b, trans=1)
    l2, _ = S2.solve(b, trans=1)
    assert np.array_equal(l1, l2)


def test_autograd_two_nodes_one_solver(gpu, orc):
    """Two sparse_solve nodes with different theta on one solver, one
    backward: each node's gradient is that of its own theta (central
    differences, h = 1e-6, 1e-6 relative)."""
    torch = pytest.importorskip("torch")
    from lace_amd.autograd import sparse_solve
    A = orc.convdiff3d(9, 8, 7)
    o = gpu.default_opts()
    o.search_max, o.restart_max, o.rtol, o.parilu_sweeps = 50, 200, 1e-13, 3
    o.trsv, o.prec_precond = gpu.TRSV_LEVEL, gpu.LACE_F64
    S = gpu.Solver(gpu.HostCSR.from_csr(A), o)
    S.setup()
    n = A.n_rows
    v2 = _perturbed(A, 23)
    t1 = torch.tensor(A.val, dtype=torch.float64, requires_grad=True)
    t2 = torch.tensor(v2, dtype=torch.float64, requires_grad=True)
    b = torch.tensor(gpu.gen_normal(4, n))
    xt1 = torch.tensor(gpu.gen_normal(3, n))
    xt2 = torch.tensor(gpu.gen_normal(6, n))
    x1 = sparse_solve(t1, b, S)
    x2 = sparse_solve(t2, b, S)
    loss = 0.5 * torch.sum((x1 - xt1) ** 2) + 0.5 * torch.sum((x2 - xt2) ** 2)
    loss.backward()

    def lossv(th1, th2):
        with torch.no_grad():
            y1 = sparse_solve(th1, b, S)
            y2 = sparse_solve(th2, b, S)
            return y1.numpy(), y2.numpy()

    h = 1e-6
    for seed, which in ((400, 1), (401, 2)):
        d = torch.tensor(gpu.gen_normal(seed, A.nnz))
        e1 = d if which == 1 else torch.zeros_like(d)
        e2 = d if which == 2 else torch.zeros_like(d)
        p1, p2 = lossv(t1.detach() + h * e1, t2.detach() + h * e2)
        m1, m2 = lossv(t1.detach() - h * e1, t2.detach() - h * e2)
        fd = (0.5 * float(np.dot(p1 - m1, p1 + m1 - 2 * xt1.numpy()))
              + 0.5 * float(np.dot(p2 - m2, p2 + m2 - 2 * xt2.numpy()))) / (2 * h)
        g = t1.grad if which == 1 else t2.grad
        an = float(torch.dot(g, d))
        assert abs(fd - an) <= 1e-6 * abs(an), (which, fd, an)
