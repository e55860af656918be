// The fused probe for batches whose queries have different stage-0 kinds, in
// one launch: each workgroup calls the body specialised for its query's kind
// (a uniform switch).  Each body is a function of its own, so its registers
// are allocated as in its own kernel instead of across every kind's code.
#include "lj_probe_impl.hpp"

namespace lj {

// A function's arguments arrive in vector registers: the query's address and
// workgroup numbers are made wave-uniform again, and the arguments are read
// through the constant address space (scalar loads, as a kernel's own
// arguments are; the batch's argument array is not written by the probe).
template <int WG, int VEC, int K0>
__device__ __attribute__((noinline)) void probe_query(const ProbeArgs* a_, uint32_t bid_, uint32_t nb_) {
  const uint64_t p = (uint64_t)(uintptr_t)a_;
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)p);
  const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(p >> 32));
  const auto* ac = (const __attribute__((address_space(4))) ProbeArgs*)(uintptr_t)(((uint64_t)hi << 32) | lo);
  const ProbeArgs& a = *(const ProbeArgs*)ac;
  probe_dispatch<WG, VEC, K0>(a, __builtin_amdgcn_readfirstlane(bid_), __builtin_amdgcn_readfirstlane(nb_));
}

template <int WG, int VEC>
__global__ __launch_bounds__(WG) void probe_mixed_kernel(const ProbeArgs* __restrict__ qs,
                                                         const uint32_t* __restrict__ wg_begin, int32_t n_q) {
  const int q = query_of_block(wg_begin, n_q);
  const uint32_t b0 = wg_begin[q], bid = blockIdx.x - b0, nb = wg_begin[q + 1] - b0;
  const ProbeArgs* a = qs + q;
  switch (a->k0) {
    case S0_PHT: probe_query<WG, VEC, S0_PHT>(a, bid, nb); break;
    case S0_NONE: probe_query<WG, VEC, S0_NONE>(a, bid, nb); break;
    case S0_MODP2_LDS: probe_query<WG, VEC, S0_MODP2_LDS>(a, bid, nb); break;
    case S0_MODP2_GL: probe_query<WG, VEC, S0_MODP2_GL>(a, bid, nb); break;
    case S0_MOD_LDS: probe_query<WG, VEC, S0_MOD_LDS>(a, bid, nb); break;
    case S0_MOD_GL: probe_query<WG, VEC, S0_MOD_GL>(a, bid, nb); break;
    case S0_MULT_LDS: probe_query<WG, VEC, S0_MULT_LDS>(a, bid, nb); break;
    case S0_MULT_GL: probe_query<WG, VEC, S0_MULT_GL>(a, bid, nb); break;
    default: probe_query<WG, VEC, S0_GENERIC>(a, bid, nb); break;
  }
}

static const ProbeEntry k_entries[] = {
    {1024, 1, kS0Mixed, reinterpret_cast<const void*>(&probe_mixed_kernel<1024, 1>),
     reinterpret_cast<const void*>(&probe_mixed_kernel<1024, 1>)},
};
int probe_entries_mixed(const ProbeEntry** out) {
  *out = k_entries;
  return (int)(sizeof(k_entries) / sizeof(k_entries[0]));
}
}  // namespace lj
