// solver -- FGMRES with given ILU factors on MatrixMarket files, with the
// interface of the reference's drivers/solver_interface.cpp:
//   --solver FGMRES|FGMRESH --matrix <A.mtx> --L <L.mtx> --U <U.mtx>
//   --RHS ONES|<b.mtx> --outDir <dir> --tolType 0|1 --tol <t>
//   --searchMax <m> --csrtrsvType <0|1|2|3> --monitorOrthog <0|1>
// The solve runs on the GPU through the drop-in data_fgmres (MGS) or
// data_fgmres_householder entry points, which print the reference's
// MATLAB-syntax progress lines; this driver adds the residual check and the
// summary, and writes the solution to
//   <outDir>/<stem(A)>_<solver>_<SSsweeps_TTTTthreads>.mtx
// (the sweep/thread tag taken from the L file name, "_GE" without one).
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "driver_common.hpp"

int main(int argc, char** argv) {
  if (argc < 7) {
    std::printf("Usage %s --solver <name> --matrix <name> --L <name> --U <name> --RHS <name> --outDir <name> ",
                argv[0]);
    std::printf("--tolType <0/1> --tol <#> --searchMax <#> --csrtrsvType <0/1> --monitorOrthog <0/1>\n");
    return 1;
  }
  const char* solver = flag_value(argc, argv, "--solver", "FGMRES");
  const char* matrix = flag_value(argc, argv, "--matrix", "matrices/Trefethen_20.mtx");
  const char* lname = flag_value(argc, argv, "--L", "matrices/Trefethen_20_L.mtx");
  const char* uname = flag_value(argc, argv, "--U", "matrices/Trefethen_20_U.mtx");
  const char* rhs = flag_value(argc, argv, "--RHS", "ONES");
  const char* outdir = flag_value(argc, argv, "--outDir", "out_");
  data_d_gmres_param par;
  std::memset(&par, 0, sizeof(par));
  par.tol_type = std::atoi(flag_value(argc, argv, "--tolType", "0"));
  par.rtol = std::atof(flag_value(argc, argv, "--tol", "1.0e-3"));
  par.search_max = std::atoi(flag_value(argc, argv, "--searchMax", "50"));
  par.user_csrtrsv_choice = std::atoi(flag_value(argc, argv, "--csrtrsvType", "1"));
  par.monitorOrthog = std::atoi(flag_value(argc, argv, "--monitorOrthog", "0"));
  par.reorth = 0;
  par.restart_max = 1;
  data_d_gmres_log glog;
  std::memset(&glog, 0, sizeof(glog));

  data_d_matrix A, L, U, b, x;
  std::memset(&A, 0, sizeof(A));
  std::memset(&L, 0, sizeof(L));
  std::memset(&U, 0, sizeof(U));
  std::memset(&b, 0, sizeof(b));
  std::memset(&x, 0, sizeof(x));
  A.storage_type = L.storage_type = U.storage_type = Magma_CSR;
  check(data_z_csr_mtx(&A, matrix), "data_z_csr_mtx");
  check(data_z_csr_mtx(&L, lname), "data_z_csr_mtx");
  check(data_z_csr_mtx(&U, uname), "data_z_csr_mtx");
  b.storage_type = Magma_DENSE;
  b.major = MagmaRowMajor;
  if (std::strcmp(rhs, "ONES") == 0) {
    std::printf("%% creating a vector of %d ones for the rhs.\n", A.num_rows);
    check(data_zvinit(&b, A.num_rows, 1, 1.0, MagmaColMajor), "data_zvinit");
  } else {
    check(data_z_dense_mtx(&b, b.major, rhs), "data_z_dense_mtx");
  }
  x.storage_type = Magma_DENSE;
  check(data_zvinit(&x, A.num_rows, 1, 0.0, MagmaColMajor), "data_zvinit");

  int info = 0;
  if (std::strcmp(solver, "FGMRES") == 0) info = data_fgmres(&A, &b, &x, &L, &U, &par, &glog);
  else if (std::strcmp(solver, "FGMRESH") == 0) info = data_fgmres_householder(&A, &b, &x, &L, &U, &par, &glog);
  if (info != 0 && info != DEV_NOTCONVERGED) std::fprintf(stderr, "solve failed (%d): %s\n", info, lace_last_error());

  // r = b - A x checked independently of the solver
  data_d_matrix r;
  std::memset(&r, 0, sizeof(r));
  r.storage_type = Magma_DENSE;
  check(data_zvinit(&r, A.num_rows, 1, 0.0, MagmaColMajor), "data_zvinit");
  check(data_z_spmv(-1.0, &A, &x, 0.0, &r), "data_z_spmv");
  data_zaxpy(A.num_rows, 1.0, b.val, 1, r.val, 1);
  const double rnorm2 = data_dnrm2(A.num_rows, r.val, 1);
  std::printf("%% external check of rnorm2 = %.16e;\n\n", rnorm2);
  std::printf("gmres_search_directions = %d;\n", glog.search_directions);
  std::printf("gmres_solve_time = %e;\n", glog.solve_time);
  std::printf("gmres_initial_residual = %e;\n", glog.initial_residual);
  std::printf("gmres_final_residual = %e;\n", glog.final_residual);
  std::printf("\n\n");
  std::printf("%% ################################################################################\n");
  std::printf("%% Matrix: %s\n%% \t%d -by- %d with %d non-zeros\n", matrix, A.num_rows, A.num_cols, A.nnz);
  std::printf("%% Solver: %s\n", solver);
  std::printf("%% \tsearch directions: %d\n", glog.search_directions);
  std::printf("%% \tsolve time [s]: %e\n", glog.solve_time);
  std::printf("%% \tinitial residual: %e\n", glog.initial_residual);
  std::printf("%% \tfinal residual: %e\n", glog.final_residual);
  std::printf("%% ################################################################################\n");
  std::printf("\n\n");
  std::printf("%% Done.\n");
  std::fflush(stdout);

  // output name: the "SSsweeps_TTTTthreads" tag of the L file, else "GE"
  const std::string lstr(lname);
  const size_t pos = lstr.find("sweeps");
  std::printf("%% pos = %lu\n", (unsigned long)pos);
  std::string tag = "GE";
  if (pos != std::string::npos && pos >= 2) tag = lstr.substr(pos - 2, 20);
  const std::string fx = std::string(outdir) + "/" + stem(matrix) + "_" + solver + "_" + tag + ".mtx";
  check(data_zwrite_dense(x, fx.c_str()), "data_zwrite_dense");

  data_zmfree(&A);
  data_zmfree(&L);
  data_zmfree(&U);
  data_zmfree(&b);
  data_zmfree(&x);
  data_zmfree(&r);
  return 0;
}
