// Total-Lagrangian SPH with Randles-Libersky correction (implementation in
// progress; the TL entry points report MFREE_ERR_UNSUPPORTED until then).
#include "tl.h"

namespace mfree {

mfree_status tl_alloc(TlState& t, uint32_t N, const double*, const double*, const std::vector<std::vector<uint32_t>>&,
                      const std::vector<std::vector<uint32_t>>&, const std::vector<uint32_t>&) {
  t.N = N;
  t.msg = "TL method not available in this build";
  return MFREE_ERR_UNSUPPORTED;
}

void tl_free(TlState& t) {
  for (void* p : t.allocs) hipFree(p);
  t.allocs.clear();
}

mfree_status tl_ensure_precomputed(TlState& t, const TlShared&) {
  t.msg = "TL method not available in this build";
  return MFREE_ERR_UNSUPPORTED;
}

void tl_enqueue_step(TlState&, const TlShared&, EventPairs*, const TakeEvent&) {}
void tl_enqueue_eval(TlState&, const TlShared&, double*, double*) {}

mfree_status tl_get_pairs(TlState&, const TlShared&, uint32_t, uint32_t, uint32_t, uint32_t*, uint32_t*, uint64_t,
                          uint64_t*) {
  return MFREE_ERR_UNSUPPORTED;
}

}  // namespace mfree
