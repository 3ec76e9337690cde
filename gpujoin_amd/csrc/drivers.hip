// drivers.hip - the lookup drivers ("approaches") behind gj_index_lookup.
#include "common.hpp"
#include "lookup_dispatch.hpp"

namespace gj {

void lookup_dispatch(const gj_index& ix, const void* probes, uint64_t m, uint32_t* out, const gj_lookup_params& p,
                     hipStream_t s) {
  switch (p.approach) {
    case GJ_PLAIN:
      dispatch_key(ix.key_bytes, [&]<class K>() { lookup_plain<K>(ix, static_cast<const K*>(probes), m, out, p, s); });
      return;
    default:
      throw error(GJ_ERR_UNSUPPORTED, "approach not available");
  }
}

}  // namespace gj
