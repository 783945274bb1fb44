// dia.hip -- value gather of the diagonal (DIA) form.
#include "dia.hpp"

namespace lace {

template <typename T>
__global__ void dia_gather_kernel(size_t m, const int* __restrict__ map, const T* __restrict__ src,
                                  T* __restrict__ val) {
  const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e < m) val[e] = map[e] >= 0 ? src[map[e]] : T(0);
}

int DiaForm::gather(const void* src, size_t es) {
  if (nd == 0) return DEV_SUCCESS;
  LACE_REQUIRE(es == 4 || es == 8, DEV_ERR_NOT_SUPPORTED, "DIA gather: element size %zu", es);
  const size_t m = (size_t)nd * n;
  if (val.bytes < es * m) LACE_TRY(val.alloc(es * m));
  const unsigned nb = (unsigned)((m + 255) / 256);
  if (es == 8)
    dia_gather_kernel<double><<<nb, 256, 0, stream()>>>(m, map.as<int>(), static_cast<const double*>(src),
                                                        val.as<double>());
  else
    dia_gather_kernel<float><<<nb, 256, 0, stream()>>>(m, map.as<int>(), static_cast<const float*>(src),
                                                       val.as<float>());
  LACE_HIP(hipGetLastError());
  return DEV_SUCCESS;
}

}  // namespace lace
