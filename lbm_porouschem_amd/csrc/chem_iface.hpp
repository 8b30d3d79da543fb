// chem_iface.hpp -- what the lattice step needs from the chemistry operator (chem.hip):
// the view of a handle's cell fields the operator reads and the source array it writes.
#pragma once
#include <hip/hip_runtime.h>

#include "../../include/lbm_chem.h"

namespace lbmimpl {

// Cell fields of an attached lattice handle (device pointers, fields in the handle's
// macroscopic layout) and the operator's output.
template <class Real> struct ChemFields {
    long long N;
    const Real* TF;
    const Real* TS;
    const Real* rho;
    const Real* S;       // [nf + nfix][N]: fluid species first, then FIX
    const Real* solid;
    int nf;              // fluid species on the handle (= LBM_CHEM_NSP): biomass at S[nf], char at S[nf + 1]
    double tf_scale, tf_ref, ts_scale, ts_ref;
    double p0, rho_ref, dt;
    double inv_rho_ref, inv_tf_scale, inv_ts_scale;  // reciprocals (the per-cell passes multiply)
    Real* xsrc;          // [16][N]: 12 species dS, biomass dS, char dS, dTF, dTS (lattice units)
    unsigned* list;      // [N] fallback cells of the surrogate pass
    unsigned* count;     // [2]: fallback-list length, work-queue head
    unsigned long long* stats;  // [3]: cells, hull fallbacks, residual fallbacks
};

// Queue the operator over every cell of the handle on `stream`: mode LBM_CHEM_PLAIN or
// LBM_CHEM_SURROGATE.  Returns LBM_OK or an LBM_ERR_* code (message in chem_error(c)).
template <class Real> int chem_run_fields(lbm_chem* c, int mode, const ChemFields<Real>& f, hipStream_t stream);
bool chem_has_surrogate(const lbm_chem* c);
int chem_device(const lbm_chem* c);

}  // namespace lbmimpl
