// generate_iLU -- ParILU(0) factors of a MatrixMarket matrix written as
// MatrixMarket files, with the interface of the reference's
// drivers/generate_iLU.cpp (--matrix, --sweeps, --outDir) and
// drivers/generate_iLU_gpu.cpp (LACE_DRIVER_GPU: also --threads).  Prints the
// preconditioner log as MATLAB assignments and writes
//   <outDir>/<stem>_LpariLUv03_<SS>sweeps_<TTTT>threads.mtx   (CSRL)
//   <outDir>/<stem>_UpariLUv03_<SS>sweeps_<TTTT>threads.mtx   (CSRU)
// The sweeps run on the GPU through the drop-in entry points.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "driver_common.hpp"

int main(int argc, char** argv) {
#ifdef LACE_DRIVER_GPU
  if (argc < 9) {
    std::printf("Usage %s --matrix <name> --sweeps <#> --threads <#> --outDir <name>\n", argv[0]);
    return 1;
  }
#else
  if (argc < 7) {
    std::printf("Usage %s --matrix <name> --sweeps <#> --outDir <name>\n", argv[0]);
    return 1;
  }
#endif
  const char* matrix = flag_value(argc, argv, "--matrix", "matrices/Trefethen_20.mtx");
  const char* outdir = flag_value(argc, argv, "--outDir", "out_");
  data_d_preconditioner_log plog;
  std::memset(&plog, 0, sizeof(plog));
  plog.maxSweeps = std::atoi(flag_value(argc, argv, "--sweeps", "1"));
#ifdef LACE_DRIVER_GPU
  plog.omp_num_threads = std::atoi(flag_value(argc, argv, "--threads", "0"));
#endif

  data_d_matrix A;
  std::memset(&A, 0, sizeof(A));
  A.storage_type = Magma_CSR;
  check(data_z_csr_mtx(&A, matrix), "data_z_csr_mtx");

#ifdef LACE_DRIVER_GPU
  std::printf("%% PariLU v0.3 GPU prescribed number of sweeps\n");
#else
  std::printf("%% PariLU v0.3 prescribed number of sweeps\n");
#endif
  data_d_matrix L, U;
  std::memset(&L, 0, sizeof(L));
  std::memset(&U, 0, sizeof(U));
  L.storage_type = Magma_CSRL;
  U.storage_type = Magma_CSCU;
  const dataType reduction = 1.0e-10;
#ifdef LACE_DRIVER_GPU
  data_PariLU_v0_3_gpu_prescribedSweeps(&A, &L, &U, reduction, &plog);
#else
  data_PariLU_v0_3_prescribedSweeps(&A, &L, &U, reduction, &plog);
#endif
  if (plog.sweeps < 0) {
    std::fprintf(stderr, "ParILU failed: %s\n", lace_last_error());
    return 1;
  }
#ifndef LACE_DRIVER_GPU
  plog.omp_num_threads = omp_threads_label();
#endif
  // ||A - LU||_F on LU's pattern and on A's
  dataType ares = 0.0, anl = 0.0;
  data_d_matrix LU;
  std::memset(&LU, 0, sizeof(LU));
  LU.storage_type = Magma_CSR;
  check(data_zmconvert(A, &LU, Magma_CSR, Magma_CSR), "data_zmconvert");
  check(data_zilures(A, L, U, &LU, &ares, &anl), "data_zilures");

  std::printf("PariLU_v0_3_maxSweeps = %d\n", plog.maxSweeps);
  std::printf("PariLU_v0_3_sweeps = %d\n", plog.sweeps);
  std::printf("PariLU_v0_3_omp_num_threads = %d\n", plog.omp_num_threads);
  std::printf("PariLU_v0_3_tol = %e\n", plog.tol);
  std::printf("PariLU_v0_3_finalStep = %e\n", plog.finalStep);
  std::printf("PariLU_v0_3_A_Frobenius = %e\n", plog.A_Frobenius);
  std::printf("PariLU_v0_3_generation_time = %e\n", plog.precond_generation_time);
  std::printf("PariLU_v0_3_initial_residual = %e\n", plog.initial_residual);
  std::printf("PariLU_v0_3_initial_nonlinear_residual = %e\n", plog.initial_nonlinear_residual);
  std::printf("PariLU_v0_3_csrilu0_res = %e\n", ares);
  std::printf("PariLU_v0_3_csrilu0_nonlinres = %e\n", anl);
  std::printf("%% %d\t%d\t%d\t%e\t%e\t%e\t%e\t%e\t%e\t%e\t%e\n", plog.maxSweeps, plog.sweeps,
              plog.omp_num_threads, plog.tol, plog.finalStep, plog.A_Frobenius,
              plog.precond_generation_time, plog.initial_residual, plog.initial_nonlinear_residual, ares,
              anl);

  char suffix[128];
  const std::string base = std::string(outdir) + "/" + stem(matrix);
  std::snprintf(suffix, sizeof(suffix), "_LpariLUv03_%02dsweeps_%04dthreads.mtx", plog.sweeps,
                plog.omp_num_threads);
  const std::string fl = base + suffix;
  std::snprintf(suffix, sizeof(suffix), "_UpariLUv03_%02dsweeps_%04dthreads.mtx", plog.sweeps,
                plog.omp_num_threads);
  const std::string fu = base + suffix;
  check(data_zwrite_csr_mtx(L, L.major, fl.c_str()), "data_zwrite_csr_mtx");
  data_d_matrix Ucsr;
  std::memset(&Ucsr, 0, sizeof(Ucsr));
  Ucsr.storage_type = Magma_CSRU;
  check(data_zmconvert(U, &Ucsr, Magma_CSC, Magma_CSR), "data_zmconvert");
  Ucsr.storage_type = Magma_CSRU;
  Ucsr.fill_mode = MagmaUpper;
  check(data_zwrite_csr_mtx(Ucsr, Ucsr.major, fu.c_str()), "data_zwrite_csr_mtx");

  data_zmfree(&L);
  data_zmfree(&U);
  data_zmfree(&Ucsr);
  data_zmfree(&LU);
  data_zmfree(&A);
  return 0;
}
