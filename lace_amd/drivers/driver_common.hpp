// driver_common.hpp -- shared helpers of the command-line drivers, which keep
// the reference's drivers/*.cpp interface (flags, defaults, MATLAB-syntax
// stdout, output file names) so its shell pipelines (drivers/*.sh) run
// unchanged on top of liblace_mi355x.so.
#pragma once

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "../../include/lace.h"

// Value of the last "--name value" pair on the command line, or dflt.
inline const char* flag_value(int argc, char** argv, const char* name, const char* dflt) {
  const char* v = dflt;
  for (int i = 1; i + 1 < argc; i++)
    if (std::strcmp(argv[i], name) == 0) v = argv[++i];
  return v;
}

// File name without directories and without its last extension
// ("matrices/A.mtx" -> "A").
inline std::string stem(const char* path) {
  std::string s(path);
  const size_t slash = s.find_last_of('/');
  if (slash != std::string::npos) s = s.substr(slash + 1);
  const size_t dot = s.find_last_of('.');
  if (dot != std::string::npos) s = s.substr(0, dot);
  return s;
}

// Exit on a library error, printing its text (the reference's CHECK exits).
inline void check(int rc, const char* what) {
  if (rc != DEV_SUCCESS) {
    std::fprintf(stderr, "%s failed (%d): %s\n", what, rc, lace_last_error());
    std::exit(1);
  }
}

// Thread count named in the CPU driver's output: the first entry of
// OMP_NUM_THREADS, else 1.  The reference's OpenMP sweeps report
// omp_get_num_threads() (src/parilu_v0_3.cpp:114,130) and its scripts
// (drivers/PariLU_testMatrix.sh) vary OMP_NUM_THREADS between runs; the
// sweeps here run on the GPU, and the label keeps those runs' files apart.
inline int omp_threads_label() {
  const char* s = std::getenv("OMP_NUM_THREADS");
  const int t = s ? std::atoi(s) : 0;
  return t > 0 ? t : 1;
}
