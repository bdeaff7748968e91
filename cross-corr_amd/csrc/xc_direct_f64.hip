// The double instances of the wave64 direct kernel's tile catalog (xc_direct_kernel.hpp),
// compiled apart from the float ones so that the two build in parallel.
#include "xc_direct_kernel.hpp"

namespace xc {

const void* direct_kernel_f64(const direct_tile& t) {
#define XC_PICK(A, B, C, D) \
    if (t.ty == A && t.rm == B && t.lm == C && t.jb == D) return reinterpret_cast<const void*>(&direct_kernel<double, A, B, C, D>);
    XC_DIRECT_TILES(XC_PICK)
#undef XC_PICK
    return nullptr;
}

}  // namespace xc
