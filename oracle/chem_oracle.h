/* chem_oracle.h -- internals shared by the CPU oracle's lattice code (lbm_oracle.c)
 * and its chemistry operator (chem_oracle.c).  TEST INFRASTRUCTURE: only the tests
 * and the bench's CPU baseline load the library these files build. */
#ifndef CHEM_ORACLE_H
#define CHEM_ORACLE_H

#include "../include/lbm_chem.h"

struct lbm_chem {
    lbm_chem_mech mech;
    int have_surrogate;
    lbm_surrogate sur;  /* sur.params points at `params` below */
    float params[LBM_MLP_NPARAM];
    char err[256];
};

/* One cell of the plain operator: in[16] -> out[14] over dt; returns accepted steps or -1. */
int chem_oracle_cell(const lbm_chem_mech* m, const double* in, double dt, double* out);

/* One cell of the surrogate with fallback: returns the flag (0 kept, 1 hull, 2 residual);
 * out receives the plain result for a flagged cell when fallback != 0. */
int chem_oracle_surrogate_cell(const lbm_chem* c, const double* in, double dt, int fallback, double* out);

#endif
