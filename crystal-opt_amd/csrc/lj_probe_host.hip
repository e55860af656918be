// lj_probe_host.hip -- kernel table and launch of the fused probe kernels
// (the kernels are instantiated in lj_probe_k_*.hip from lj_probe_impl.hpp).
#include <hip/hip_runtime.h>

#include "lj_internal.hpp"

namespace lj {

size_t probe_lds_bytes(const ProbeConfig& c, int lds_filter_words) {
  size_t off = ((size_t)lds_filter_words * 4 + 15) & ~(size_t)15;
  off += (size_t)(c.wg / 64) * 3 * kQueueCap * 4;
  off += 3 * kGroupSlots * 8 + kGroupSlots * 4 + kGroupSlots * kMaxDims * 4;
  off += kStats * 8 + (kMaxDims + 1) * 8;
  off += 16;  // tile counter
  return off;
}

static const void* find_probe(const ProbeConfig& c) {
  int (*tables[])(const ProbeEntry**) = {probe_entries_pht, probe_entries_modp2, probe_entries_mod,
                                         probe_entries_mult, probe_entries_generic};
  for (auto t : tables) {
    const ProbeEntry* e = nullptr;
    const int n = t(&e);
    for (int i = 0; i < n; i++)
      if (e[i].wg == c.wg && e[i].vec == c.vec && e[i].k0 == c.k0) return e[i].fn;
  }
  return nullptr;
}

hipError_t probe_occupancy(const ProbeConfig& c, size_t lds_bytes, int* blocks_per_cu) {
  const void* fn = find_probe(c);
  if (!fn) return hipErrorInvalidValue;
  hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes);
  if (e != hipSuccess) return e;
  return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, fn, c.wg, lds_bytes);
}

hipError_t launch_probe(const ProbeConfig& c, const ProbeArgs& a, int grid, size_t lds_bytes, hipStream_t s) {
  const void* fn = find_probe(c);
  if (!fn) return hipErrorInvalidValue;
  void* kargs[] = {const_cast<ProbeArgs*>(&a)};
  return hipLaunchKernel(fn, dim3(grid), dim3(c.wg), kargs, lds_bytes, s);
}

}  // namespace lj
