"""The command-line drivers with the reference's interface (drivers/
generate_iLU*.cpp, drivers/solver_interface.cpp): flags, MATLAB-syntax
stdout, output file names, and the factor -> solve pipeline through
MatrixMarket files.  The drivers run as child processes on the GPU."""
import os
import re
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
BIN = os.path.join(os.path.dirname(HERE), "lace_amd", "drivers", "bin")
TREF = os.path.join(HERE, "golden", "Trefethen_20.mtx")


def _run(args, cwd):
    r = subprocess.run(args, cwd=cwd, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "failed" not in r.stderr, r.stdout + r.stderr
    return r.stdout


def _read_dense(path):
    with open(path) as f:
        lines = [l for l in f if not l.startswith("%")]
    m, n = map(int, lines[0].split())
    return np.array([float(v) for v in lines[1:1 + m * n]])


def test_usage_message():
    r = subprocess.run([os.path.join(BIN, "solver")], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and r.stdout.startswith("Usage")


def test_generate_and_solve_pipeline(gpu, orc, tmp_path):
    out = str(tmp_path)
    log = _run([os.path.join(BIN, "generate_iLU_gpu"), "--matrix", TREF, "--sweeps", "5",
                "--threads", "32", "--outDir", out], out)
    assert "PariLU_v0_3_sweeps = 5" in log and "PariLU_v0_3_omp_num_threads = 32" in log
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

    for solver, kind in (("FGMRESH", "Householders_mkltrsv"), ("FGMRES", "mkltrsv")):
        slog = _run([os.path.join(BIN, "solver"), "--solver", solver, "--matrix", TREF, "--L", fl,
                     "--U", fu, "--RHS", "ONES", "--outDir", out, "--tolType", "0", "--tol", "1.0e-10",
                     "--searchMax", "1000", "--csrtrsvType", "0", "--monitorOrthog", "1"], out)
        hist = re.findall(r"^FGMRES_%s_search\((\d+)\) = ([-+0-9.eE]+);$" % kind, slog, re.M)
        assert len(hist) >= 2, slog
        res = [float(v) for _, v in hist]
        assert res[-1] < 1e-10
        chk = float(re.search(r"external check of rnorm2 = ([-+0-9.eE]+);", slog).group(1))
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
    _run([os.path.join(BIN, "generate_iLU"), "--matrix", TREF, "--sweeps", "3", "--outDir", out], out)
    fl = os.path.join(out, "Trefethen_20_LpariLUv03_03sweeps_0001threads.mtx")
    fu = os.path.join(out, "Trefethen_20_UpariLUv03_03sweeps_0001threads.mtx")
    slog = _run([os.path.join(BIN, "solver"), "--solver", "FGMRES", "--matrix", TREF, "--L", fl, "--U", fu,
                 "--RHS", rhs, "--outDir", out, "--tolType", "1", "--tol", "1.0e-10", "--searchMax", "60",
                 "--csrtrsvType", "1"], out)
    assert "Reading dense matrix from file" in slog
    assert re.search(r"^FGMRES_partrsv_search\(1\) = ", slog, re.M)
    x = _read_dense(os.path.join(out, "Trefethen_20_FGMRES_03sweeps_0001threads.mtx"))
    assert np.linalg.norm(b - A.scipy() @ x) <= 1.01e-10 * np.linalg.norm(b)
