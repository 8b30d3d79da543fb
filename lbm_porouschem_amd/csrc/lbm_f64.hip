// lbm_f64.hip -- FP64 instantiations of the step kernels.  Built with
// -ffp-contract=off so the FP64 path rounds like the CPU oracle (parity <= 1e-10
// also where the reference's eps + 1e-12 denominators make feq ill-conditioned).
#include "lbm_impl.hpp"

namespace lbmimpl {
Impl* make_impl_f64(lbm_handle* h, int nf, int nfix) {
    return h->d.dim == 3 ? make_for_species<double, double, 3>(h, nf, nfix)
                         : make_for_species<double, double, 2>(h, nf, nfix);
}
}  // namespace lbmimpl
