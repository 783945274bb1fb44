"""The command-line drivers with the reference's interface (drivers/
generate_iLU*.cpp, drivers/solver_interface.cpp): flags, MATLAB-syntax
stdout, output file names, and the factor -> solve pipeline through
MatrixMarket files.  The drivers run as child processes on the GPU."""
import glob
import os
import re
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
BIN = os.path.join(os.path.dirname(HERE), "lace_amd", "drivers", "bin")
TREF = os.path.join(HERE, "golden", "Trefethen_20.mtx")

_NUM = r"[-+]?(?:\d+\.?\d*|\.\d+)(?:[eE][-+]?\d+)?"
_STMT = re.compile(r"^\s*[A-Za-z]\w*(?:\(\d+\))?\s*=\s*(?:%s|'[^']*')\s*$" % _NUM)


def matlab_errors(text):
    """Lines of a log that a MATLAB `run` of it would not accept as
    comments or numeric assignments (the post-processing scripts in the
    reference's results/ evaluate the logs this way)."""
    bad = []
    for line in text.splitlines():
        s = line.strip()
        if not s or s.startswith("%"):
            continue
        stmts = [t for t in s.split(";") if t.strip()]
        if not stmts or not all(_STMT.match(t) for t in stmts):
            bad.append(line)
    return bad


def test_matlab_checker_rejects_garbage():
    assert matlab_errors("a = 1;\nb(2) = -3.5e-07;\n% note\nrnorm2 = 1e+00; rtol = 2e-10;\n") == []
    assert matlab_errors("Reading dense matrix: done\n") != []
    assert matlab_errors("x(1) = nan;\n") != []


def _run(args, cwd, env=None):
    e = dict(os.environ)
    e.update(env or {})
    r = subprocess.run(args, cwd=cwd, capture_output=True, text=True, timeout=120, env=e)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "failed" not in r.stderr, r.stdout + r.stderr
    return r.stdout


def _read_dense(path):
    with open(path) as f:
        lines = [l for l in f if not l.startswith("%")]
    m, n = map(int, lines[0].split())
    return np.array([float(v) for v in lines[1:1 + m * n]])


def _history(slog, kind, what="search"):
    pat = r"^FGMRES_%s_%s\((\d+)\) = (%s);$" % (kind, what, _NUM)
    return [(int(k), float(v)) for k, v in re.findall(pat, slog, re.M)]


def test_usage_message():
    r = subprocess.run([os.path.join(BIN, "solver")], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and r.stdout.startswith("Usage")


def test_generate_and_solve_pipeline(gpu, orc, tmp_path):
    out = str(tmp_path)
    log = _run([os.path.join(BIN, "generate_iLU_gpu"), "--matrix", TREF, "--sweeps", "5",
                "--threads", "32", "--outDir", out], out)
    assert "PariLU_v0_3_sweeps = 5" in log and "PariLU_v0_3_omp_num_threads = 32" in log
    assert matlab_errors(log) == []
    fl = os.path.join(out, "Trefethen_20_LpariLUv03_05sweeps_0032threads.mtx")
    fu = os.path.join(out, "Trefethen_20_UpariLUv03_05sweeps_0032threads.mtx")
    assert os.path.exists(fl) and os.path.exists(fu)
    # the factor files hold the device factors: CSRL (unit diagonal last)
    # and CSRU (diagonal first), same pattern as A's triangles
    L = orc.read_mtx(fl)
    U = orc.read_mtx(fu)
    A = orc.read_mtx(TREF)
    assert L.nnz + U.nnz == A.nnz + A.n_rows
    assert np.all(np.abs(L.dense().diagonal() - 1.0) == 0.0)
    # five synchronous sweeps from A's entries: the oracle's factors
    Lo, Uo = orc.parilu(A, 5).factors()
    assert np.allclose(L.dense(), Lo.dense(), rtol=1e-13, atol=1e-13)
    assert np.allclose(U.dense(), Uo.dense(), rtol=1e-13, atol=1e-13)

    for solver, kind in (("FGMRESH", "Householders_mkltrsv"), ("FGMRES", "mkltrsv")):
        slog = _run([os.path.join(BIN, "solver"), "--solver", solver, "--matrix", TREF, "--L", fl,
                     "--U", fu, "--RHS", "ONES", "--outDir", out, "--tolType", "0", "--tol", "1.0e-10",
                     "--searchMax", "1000", "--csrtrsvType", "0", "--monitorOrthog", "1"], out)
        assert matlab_errors(slog) == [], slog
        hist = _history(slog, kind)
        assert len(hist) >= 2, slog
        assert [k for k, _ in hist] == list(range(1, len(hist) + 1))
        assert hist[-1][1] < 1e-10 and all(v >= 1e-10 for _, v in hist[:-1])
        # orthogonality error of the first k basis vectors, ahead of each
        # search line: ||I - Q_k^T Q_k||_inf can only grow with k (the
        # leading block keeps its entries), starts at rounding level, stays
        # there for Householder, and for MGS grows as the residual falls
        # (loss of orthogonality ~ eps * ||r0|| / ||r_k||)
        orth = _history(slog, kind, "ortherr")
        assert [k for k, _ in orth] == [k for k, _ in hist]
        vals = [v for _, v in orth]
        assert all(a <= b for a, b in zip(vals, vals[1:])), (solver, vals)
        assert 0.0 <= vals[0] <= 4.5e-16
        if solver == "FGMRESH":
            assert vals[-1] < 1e-13, vals
            hh_orth = vals[-1]
        else:
            r0 = float(re.search(r"^rnorm2 = (%s);" % _NUM, slog, re.M).group(1))
            for k in range(1, len(vals)):
                assert vals[k] <= 1e3 * 2.2e-16 * r0 / hist[k - 1][1], (k, vals, hist)
            assert vals[-1] > 100 * hh_orth
        lines = slog.splitlines()
        for k in (1, len(hist)):
            i = lines.index(next(l for l in lines if l.startswith("FGMRES_%s_search(%d) =" % (kind, k))))
            assert lines[i - 1].startswith("FGMRES_%s_ortherr(%d) =" % (kind, k))
        chk = float(re.search(r"external check of rnorm2 = (%s);" % _NUM, slog).group(1))
        x = _read_dense(os.path.join(out, f"Trefethen_20_{solver}_05sweeps_0032threads.mtx"))
        b = np.ones(A.n_rows)
        r = np.linalg.norm(b - A.scipy() @ x)
        assert r <= 1e-9 and abs(r - chk) <= 1e-12 + 1e-6 * r


def test_solver_reads_dense_rhs_and_cpu_generator(gpu, orc, tmp_path):
    out = str(tmp_path)
    A = orc.read_mtx(TREF)
    b = np.random.default_rng(2).standard_normal(A.n_rows)
    rhs = os.path.join(out, "b.mtx")
    with open(rhs, "w") as f:
        f.write("%%MatrixMarket matrix array real general\n%d 1\n" % A.n_rows)
        f.writelines("%.17e\n" % v for v in b)
    _run([os.path.join(BIN, "generate_iLU"), "--matrix", TREF, "--sweeps", "3", "--outDir", out], out,
         env={"OMP_NUM_THREADS": "3"})
    fl = os.path.join(out, "Trefethen_20_LpariLUv03_03sweeps_0003threads.mtx")
    fu = os.path.join(out, "Trefethen_20_UpariLUv03_03sweeps_0003threads.mtx")
    slog = _run([os.path.join(BIN, "solver"), "--solver", "FGMRES", "--matrix", TREF, "--L", fl, "--U", fu,
                 "--RHS", rhs, "--outDir", out, "--tolType", "1", "--tol", "1.0e-10", "--searchMax", "60",
                 "--csrtrsvType", "1"], out)
    assert "Reading dense matrix from file" in slog
    assert matlab_errors(slog) == [], slog
    hist = _history(slog, "partrsv")
    assert len(hist) >= 2 and "FGMRES_partrsv_ortherr(1)" not in slog
    # the ParCSRTRSV sweep counts precede each search line
    for k, _ in hist:
        assert "ParCSRTRSV_L(%d) = 2;\nParCSRTRSV_U(%d) = 2;\n" % (k, k) in slog
    # relative tolerance: the third figure of the first line is rtol*rnorm2
    m = re.search(r"^rnorm2 = (%s); tol = (%s); rtol = (%s);$" % (_NUM, _NUM, _NUM), slog, re.M)
    r0, tol, rt = (float(g) for g in m.groups())
    assert tol == 1e-10 and abs(r0 - np.linalg.norm(b)) <= 1e-6 * r0 and abs(rt - tol * r0) <= 1e-6 * rt
    x = _read_dense(os.path.join(out, "Trefethen_20_FGMRES_03sweeps_0003threads.mtx"))
    assert np.linalg.norm(b - A.scipy() @ x) <= 1.01e-10 * np.linalg.norm(b)


def test_reference_script_layout(gpu, orc, tmp_path):
    """The command lines of the reference's drivers/PariLU_testMatrix.sh and
    drivers/Solve_testMatrix.sh, run from a working directory holding the
    executables (here as symbolic links): factors for OMP_NUM_THREADS in
    {2, 4} and 1..2 sweeps, then FGMRES and FGMRESH over every factor pair,
    each log a MATLAB file."""
    work = tmp_path
    for exe in ("generate_iLU", "solver"):
        os.symlink(os.path.join(BIN, exe), str(work / exe))
    pre = work / "PariLU_mi355x"
    sol = work / "Solver_mi355x"
    pre.mkdir()
    sol.mkdir()
    for t in (2, 4):
        for s in (1, 2):
            log = _run(["./generate_iLU", "--matrix", TREF, "--sweeps", str(s), "--outDir", str(pre)], str(work),
                       env={"OMP_NUM_THREADS": str(t)})
            assert matlab_errors(log) == []
    Ls = sorted(glob.glob(str(pre / "Trefethen_20_L*.mtx")))
    Us = sorted(glob.glob(str(pre / "Trefethen_20_U*.mtx")))
    assert [os.path.basename(p) for p in Ls] == [
        "Trefethen_20_LpariLUv03_%02dsweeps_%04dthreads.mtx" % (s, t) for s in (1, 2) for t in (2, 4)]
    A = orc.read_mtx(TREF)
    for s in ("FGMRES", "FGMRESH"):
        for i, (fl, fu) in enumerate(zip(Ls, Us)):
            slog = _run(["./solver", "--solver", s, "--matrix", TREF, "--L", fl, "--U", fu, "--RHS", "ONES",
                         "--outDir", str(sol), "--tolType", "0", "--tol", "1.0e-10", "--searchMax", "1000",
                         "--csrtrsvType", "0", "--monitorOrthog", "1"], str(work))
            (sol / ("log_%s_mi355x_Trefethen_20_%02d.m" % (s, i))).write_text(
                "%% %s\n%% %s\n%s" % (fl, fu, slog))
            assert matlab_errors(slog) == [], slog
            tag = os.path.basename(fl)[len("Trefethen_20_LpariLUv03_"):-len(".mtx")]
            x = _read_dense(str(sol / ("Trefethen_20_%s_%s.mtx" % (s, tag))))
            assert np.linalg.norm(np.ones(A.n_rows) - A.scipy() @ x) <= 1e-9
    assert len(list(sol.glob("log_*.m"))) == 8
