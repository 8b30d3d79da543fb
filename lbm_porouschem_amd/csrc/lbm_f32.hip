// lbm_f32.hip -- FP32 instantiations of the step kernels.
#include "lbm_impl.hpp"

namespace lbmimpl {
Impl* make_impl_f32(lbm_handle* h, int nf, int nfix) {
    return h->d.dim == 3 ? make_for_species<float, float, 3>(h, nf, nfix)
                         : make_for_species<float, float, 2>(h, nf, nfix);
}
}  // namespace lbmimpl
