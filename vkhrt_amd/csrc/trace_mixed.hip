// Traversal kernels for the mixed geometry variant(s) (instantiations of trace_impl.hpp).
#include "trace_impl.hpp"

namespace vk {
using Variant0 = GeoMixed;
VKHRT_INSTANTIATE_TRACE(Variant0)
}  // namespace vk
