// Traversal kernels for the triangle-mesh geometry variant (instantiations of trace_impl.hpp).
#include "trace_impl.hpp"

namespace vk {
using Variant0 = Geo<kTriangles, 0, false>;
VKHRT_INSTANTIATE_TRACE(Variant0)
}  // namespace vk
