// Traversal kernels for the phantom const geometry variant(s) (instantiations of trace_impl.hpp).
#include "trace_impl.hpp"

namespace vk {
using Variant0 = Geo<kPhantom, kRadiusConstant, false>;
VKHRT_INSTANTIATE_TRACE(Variant0)
using Variant1 = Geo<kPhantom, kRadiusConstant, true>;
VKHRT_INSTANTIATE_TRACE(Variant1)
}  // namespace vk
