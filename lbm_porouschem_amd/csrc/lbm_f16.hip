// lbm_f16.hip -- FP16 population storage (deviations from the lattice reference
// values), FP32 arithmetic and macroscopic fields.
#include "lbm_impl.hpp"

namespace lbmimpl {
Impl* make_impl_f16(lbm_handle* h, int nf, int nfix) {
    return h->d.dim == 3 ? make_for_species<float, _Float16, 3>(h, nf, nfix)
                         : make_for_species<float, _Float16, 2>(h, nf, nfix);
}
}  // namespace lbmimpl
