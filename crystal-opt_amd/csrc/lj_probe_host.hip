// lj_probe_host.hip -- kernel table and launch of the fused probe kernels
// (the kernels are instantiated in lj_probe_k_*.hip from lj_probe_impl.hpp).
#include <hip/hip_runtime.h>

#include "lj_internal.hpp"

namespace lj {

// The probe's LDS layout (probe_body): wave queues, statistics, sample
// totals, tile counter, the group table, then the staged filters.
size_t probe_lds_bytes(const ProbeConfig& c, int lds_filter_words, int group_slots) {
  size_t off = (size_t)(c.wg / 64) * 3 * kQueueCap * 4;
  off += kStats * 8 + (kMaxDims + 1) * 8 + 16;
  off += (size_t)group_slots * kGroupSlotBytes;
  off = (off + 15) & ~(size_t)15;
  off += (size_t)lds_filter_words * 4;
  return off;
}

static const ProbeEntry* find_entry(const ProbeConfig& c) {
  int (*tables[])(const ProbeEntry**) = {probe_entries_pht, probe_entries_modp2, probe_entries_mod,
                                         probe_entries_mult, probe_entries_generic, probe_entries_mixed};
  for (auto t : tables) {
    const ProbeEntry* e = nullptr;
    const int n = t(&e);
    for (int i = 0; i < n; i++)
      if (e[i].wg == c.wg && e[i].vec == c.vec && e[i].k0 == c.k0) return &e[i];
  }
  return nullptr;
}

bool probe_has_entry(const ProbeConfig& c) { return find_entry(c) != nullptr; }

static const void* find_probe(const ProbeConfig& c) {
  const ProbeEntry* e = find_entry(c);
  return e ? e->fn : nullptr;
}

// Also raises both kernels' dynamic LDS limit to the device maximum (a
// launch with less LDS is unaffected; the limit never has to come down).
hipError_t probe_occupancy(const ProbeConfig& c, size_t lds_bytes, int* blocks_per_cu) {
  const ProbeEntry* en = find_entry(c);
  if (!en) return hipErrorInvalidValue;
  int dev = 0, max_lds = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e == hipSuccess) e = hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev);
  if (e != hipSuccess) return e;
  if ((size_t)max_lds < lds_bytes) return hipErrorInvalidValue;
  e = hipFuncSetAttribute(en->fn, hipFuncAttributeMaxDynamicSharedMemorySize, max_lds);
  if (e != hipSuccess) return e;
  e = hipFuncSetAttribute(en->fn_multi, hipFuncAttributeMaxDynamicSharedMemorySize, max_lds);
  if (e != hipSuccess) return e;
  return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, en->fn, c.wg, lds_bytes);
}

hipError_t launch_probe_multi(const ProbeConfig& c, const ProbeArgs* qs, const uint32_t* wg_begin, int n_q, int grid,
                              size_t lds_bytes, hipStream_t s) {
  const ProbeEntry* en = find_entry(c);
  if (!en) return hipErrorInvalidValue;
  void* kargs[] = {(void*)&qs, (void*)&wg_begin, (void*)&n_q};
  return hipLaunchKernel(en->fn_multi, dim3(grid), dim3(c.wg), kargs, lds_bytes, s);
}

hipError_t launch_probe(const ProbeConfig& c, const ProbeArgs& a, int grid, size_t lds_bytes, hipStream_t s) {
  const void* fn = find_probe(c);
  if (!fn) return hipErrorInvalidValue;
  void* kargs[] = {const_cast<ProbeArgs*>(&a)};
  return hipLaunchKernel(fn, dim3(grid), dim3(c.wg), kargs, lds_bytes, s);
}

}  // namespace lj
