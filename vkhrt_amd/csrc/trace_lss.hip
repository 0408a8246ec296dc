// Traversal kernels for the lss geometry variant(s) (instantiations of trace_impl.hpp).
#include "trace_impl.hpp"

namespace vk {
using Variant0 = Geo<kLss, 0, false>;
VKHRT_INSTANTIATE_TRACE(Variant0)
}  // namespace vk
