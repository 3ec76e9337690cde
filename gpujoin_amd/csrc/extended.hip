// extended.hip - index kinds beyond binary search / B-tree / no-op.
#include "common.hpp"
#include "lookup_dispatch.hpp"

namespace gj {

uint64_t extended_memory(const gj_index&) { return 0; }

template <class K>
void build_extended(gj_index& ix, hipStream_t) {
  throw error(GJ_ERR_UNSUPPORTED, std::string("index kind not built in this library: ") + gj_index_type_name(ix.kind));
}
template void build_extended<uint32_t>(gj_index&, hipStream_t);
template void build_extended<uint64_t>(gj_index&, hipStream_t);

template <class K>
void dispatch_extended_plain(const gj_index& ix, const K*, uint64_t, uint32_t*, const gj_lookup_params&, hipStream_t) {
  throw error(GJ_ERR_UNSUPPORTED, std::string("lookup not available for ") + gj_index_type_name(ix.kind));
}
template void dispatch_extended_plain<uint32_t>(const gj_index&, const uint32_t*, uint64_t, uint32_t*,
                                                const gj_lookup_params&, hipStream_t);
template void dispatch_extended_plain<uint64_t>(const gj_index&, const uint64_t*, uint64_t, uint32_t*,
                                                const gj_lookup_params&, hipStream_t);

}  // namespace gj
