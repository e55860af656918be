// Fused probe kernels for stage-0 kinds: S0_MULT_LDS S0_MULT_GL.
#include "lj_probe_impl.hpp"

namespace lj {
#define LJ_ENTRY(W, V, K) \
  {W, V, K, reinterpret_cast<const void*>(&probe_kernel<W, V, K>), \
   reinterpret_cast<const void*>(&probe_multi_kernel<W, V, K>)},
static const ProbeEntry k_entries[] = {
    LJ_PROBE_SHAPES(LJ_ENTRY, S0_MULT_LDS)
    LJ_PROBE_SHAPES(LJ_ENTRY, S0_MULT_GL)
};
int probe_entries_mult(const ProbeEntry** out) {
  *out = k_entries;
  return (int)(sizeof(k_entries) / sizeof(k_entries[0]));
}
}  // namespace lj
