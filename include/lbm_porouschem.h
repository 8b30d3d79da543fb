/*
 * lbm_porouschem.h -- C ABI of the coupled lattice-Boltzmann step
 * (flow + thermal + species + gray-LBM porous drag + pyrolysis source).
 *
 * Two shared libraries export exactly this interface:
 *   liblbm_mi355x.so   the product: HIP kernels for gfx950 (lbm_porouschem_amd/csrc)
 *   liblbm_oracle.so   the CPU restatement used only as the test oracle (oracle/)
 *
 * The reference exposes no FFI; its boundary is the Python solver-class API that
 * src/LBM/app.py drives.  Each entry point below replaces one piece of that API
 * (reference file:line given per function).  The Python mirror of the solver
 * classes (lbm_porouschem_amd/solver.py) binds these with ctypes.
 *
 * Conventions
 *   - Plain C, no C++ types across the boundary, no exceptions.
 *   - Every function returns 0 on success or a negative LBM_ERR_* code; the
 *     message is available from lbm_last_error(handle).
 *   - Host arrays are borrowed only for the duration of a call and are laid out
 *     like the reference's Taichi to_numpy(): C-contiguous (nx, ny, nz) with z
 *     fastest; 2D problems use nz = 1.
 *   - Enum values are the integers of the reference's src/LBM/util/flag.py.
 *   - One handle owns one device stream; calls on a handle must be serialised
 *     by the caller; distinct handles may be driven from distinct threads.
 */
#ifndef LBM_POROUSCHEM_H
#define LBM_POROUSCHEM_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LBM_ABI_VERSION 1
#define LBM_MAX_SPECIES 16
#define LBM_MAX_REACTIONS 16
#define LBM_MAX_FACES 6

/* ---- status codes ---- */
#define LBM_OK 0
#define LBM_ERR_INVALID (-1)     /* bad argument / size / field id */
#define LBM_ERR_UNSUPPORTED (-2) /* model combination this backend does not run */
#define LBM_ERR_OOM (-3)         /* device or host allocation failed */
#define LBM_ERR_DEVICE (-4)      /* HIP runtime error */
#define LBM_ERR_DIVERGED (-5)    /* divergence detector tripped (NaN or |u| too large) */
#define LBM_ERR_STATE (-6)       /* call out of order (e.g. step before init) */

/* ---- storage / collision choices (lbm_desc) ---- */
#define LBM_PREC_FP64 0      /* populations and arithmetic in FP64 */
#define LBM_PREC_FP32 1      /* populations and arithmetic in FP32 */
#define LBM_PREC_FP16 2      /* populations stored as FP16 deviations f - w_i*rho0, FP32 arithmetic */
#define LBM_COLL_BGK 0       /* single-relaxation-time, the reference's collision */
#define LBM_COLL_MRT 1       /* D3Q19/D2Q9 multiple-relaxation-time (new; reduces to BGK when all rates = 1/tau) */

/* ---- enums mirrored from src/LBM/util/flag.py:10-109 ---- */
#define LBM_VISC_NONE 0
#define LBM_VISC_CONSTANT 1
#define LBM_VISC_SUTHERLAND 2
#define LBM_VISC_MIXTURE 3
#define LBM_PORO_SPHERICAL 0
#define LBM_PORO_DARCY 1
#define LBM_PORO_DARCYFORCHHEIMER 2
#define LBM_PORO_DARCY_HIGH 3
#define LBM_COND_NONE 0
#define LBM_COND_CONSTANT 1
#define LBM_COND_POLYNOMIAL 2
#define LBM_COND_MIXTURE 3
#define LBM_THERMO_NONE 0
#define LBM_THERMO_CONSTANT 1
#define LBM_THERMO_POLYNOMIAL 2
#define LBM_THERMO_NASA7 3
#define LBM_THERMO_MIXTURE 4
#define LBM_TDIFF_NONE 0
#define LBM_TDIFF_CONSTANT 1
#define LBM_TDIFF_DERIVED 2
#define LBM_TDIFF_PRANDTL 3
#define LBM_DIFF_NONE 0
#define LBM_DIFF_CONSTANT 1
#define LBM_DIFF_SCHMIDT 2
#define LBM_EOS_IDEAL_GAS 0
#define LBM_EOS_INCOMPRESSIBLE 1
#define LBM_RAD_NONE 0
#define LBM_RAD_SURFACE_UNIFORM 1
#define LBM_FORCE_NONE 0
#define LBM_FORCE_GUO 1
#define LBM_BC_PERIODIC 0
#define LBM_BC_ZEROGRADIENT 1
#define LBM_BC_FIXEDVALUE 2
#define LBM_BCFLOW_PERIODIC 0
#define LBM_BCFLOW_INLET 1
#define LBM_BCFLOW_OUTLET 2
#define LBM_BCFLOW_WALL 3
#define LBM_BCFLOW_SYMMETRIC 4
#define LBM_BCFLOW_INLET_FLOW 5
#define LBM_BCMODEL_NONE 0
#define LBM_BCMODEL_NEE 1
#define LBM_BCMODEL_NEBB 2
#define LBM_BCMODEL_ES 3

/* ---- field ids for lbm_set_field / lbm_get_field ----
 * Cell fields hold nx*ny*nz values; `comp` selects a vector component or a
 * species index.  Face profiles hold the face's cell count (face 0/1: ny*nz,
 * 2/3: nx*nz, 4/5: nx*ny) with comp = face*3 + axis (velocity) or face (rho). */
#define LBM_FIELD_RHO 0          /* LBM3D/_core.py:46  rho */
#define LBM_FIELD_VEL 1          /* LBM3D/_core.py:47  v (comp 0..2) */
#define LBM_FIELD_SOLID 2        /* LBM3D/_core.py:48  solid fraction */
#define LBM_FIELD_RHOS 3         /* LBM3D/_core.py:84  rhos (solid density) */
#define LBM_FIELD_COEF_DARCY 4   /* LBM3D/_core.py:93 */
#define LBM_FIELD_COEF_FORCH 5   /* LBM3D/_core.py:94 */
#define LBM_FIELD_TF 6           /* TF.S (lattice units) */
#define LBM_FIELD_TS 7           /* TS.S (lattice units) */
#define LBM_FIELD_TF_DS 8        /* TF.dS */
#define LBM_FIELD_TS_DS 9        /* TS.dS */
#define LBM_FIELD_EXCH_SURF 10   /* TS.exchangeSurface  LBM3D/_thermal.py:116 */
#define LBM_FIELD_EXCH_COEF 11   /* TS.exchangeCoef     LBM3D/_thermal.py:115 */
#define LBM_FIELD_RAD_SURF 12    /* TS.radiation_surface */
#define LBM_FIELD_EMISSIVITY 13  /* TS.emissivity */
#define LBM_FIELD_SPECIES 14     /* species[comp].S */
#define LBM_FIELD_SPECIES_DS 15  /* species[comp].dS */
#define LBM_FIELD_V_BC_PROFILE 16   /* v_bc_profile[face] (comp = face*3+axis) LBM3D/_core.py:59-64 */
#define LBM_FIELD_RHO_BC_PROFILE 17 /* rho_bc_profile[face] (comp = face)      LBM3D/_core.py:65-70 */

#define LBM_DTYPE_F64 0
#define LBM_DTYPE_F32 1

/* ---- reductions for lbm_reduce ---- */
#define LBM_REDUCE_MAX_V 0    /* max |u| over all cells          LBM3D/_io.py:572-576 */
#define LBM_REDUCE_MIN_TF 1   /* min TF.S (lattice units)         LBM3D/_io.py:582-586 */
#define LBM_REDUCE_MAX_TF 2   /* max TF.S (lattice units)         LBM3D/_io.py:577-581 */
#define LBM_REDUCE_MASS 3     /* FP64 sum of rho over cells with solid < 1 */
#define LBM_REDUCE_TF_SUM 4   /* FP64 sum of TF.S over cells with solid < 1 */
#define LBM_REDUCE_TS_SUM 5   /* FP64 sum of TS.S over cells with solid > 0 */
#define LBM_REDUCE_FIX_SUM 6  /* FP64 sum of all FIX species densities (solid-conversion integral) */
#define LBM_REDUCE_SPECIES_SUM 7 /* FP64 sum of species[comp].S over cells with solid < 1 */

typedef struct lbm_desc {
    int abi_version;            /* must equal LBM_ABI_VERSION */
    int dim;                    /* 2 or 3 */
    int nx, ny, nz;             /* nz = 1 in 2D */
    int thermal, chemical, poro, radiation;  /* module switches (LBM3D/_core.py:76-79) */
    int n_species;              /* registered species, fluid and FIX, in registration order */
    int species_fix[LBM_MAX_SPECIES];        /* 1 = FIX (solid) species */
    int precision;              /* LBM_PREC_* */
    int collision;              /* LBM_COLL_* */
    int device;                 /* HIP device ordinal (ignored by the CPU oracle) */
} lbm_desc;

/* Per scalar lattice: TF (fluid temperature) and TS (solid temperature).
 * Mirrors TemperatureFluid / TemperatureSolid (LBM3D/_thermal.py:6-221). */
typedef struct lbm_thermal_params {
    double v_ref, v_scale;      /* physical = S*v_scale + v_ref (LBM3D/_scalarField.py:39-42) */
    int diff_model;             /* LBM_TDIFF_* */
    double diff, Pr;            /* thermal_diff (m^2/s), Prandtl number */
    int cond_model;             /* LBM_COND_* */
    double cond, cond_poly[5];
    int cap_model;              /* LBM_THERMO_* */
    double cm, cm_poly[5];
    double Trange[3];           /* NASA7 temperature ranges */
    double nasa[2][7];          /* NASA7 coefficients (low, high) */
    int bc_type[LBM_MAX_FACES]; /* LBM_BC_* per face */
    double bc_value[LBM_MAX_FACES]; /* s_BC: boundary value in lattice units */
} lbm_thermal_params;

/* Per species: Specie (LBM3D/_chemical.py:7-137). */
typedef struct lbm_species_params {
    double molemass;            /* kg/mol (Specie.molemass = Mmass/1000) */
    int visc_model;             /* LBM_VISC_NONE/CONSTANT/SUTHERLAND */
    double visco, sutherland[2];
    int diff_model;             /* LBM_DIFF_* */
    double diff, Sc;
    int thermo_model;           /* only LBM_THERMO_CONSTANT is reachable in the reference */
    double enthalpy, capa;      /* J/mol, J/(mol K) */
    int cond_model;             /* LBM_COND_CONSTANT / POLYNOMIAL */
    double cond, cond_poly[5];
    int bc_type[LBM_MAX_FACES];
    double bc_value[LBM_MAX_FACES];
} lbm_species_params;

/* One Arrhenius reaction (LBM3D/_chemical.py:141-235). Stoichiometry tables are
 * indexed by species registration order. */
typedef struct lbm_reaction_params {
    double A, Ea, b, Tmin;
    double coef_reactant[LBM_MAX_SPECIES];
    double coef_product[LBM_MAX_SPECIES];
    double coef_rate[LBM_MAX_SPECIES];
} lbm_reaction_params;

typedef struct lbm_params {
    double dx, dt, tol;         /* LBM3D/_core.py:15,21 */
    int eos;                    /* LBM_EOS_* */
    int visc_model;             /* LBM_VISC_* */
    double visco, sutherland[2];
    int force_term;             /* LBM_FORCE_* (app.py:104 forces GUO) */
    int bc_model;               /* LBM_BCMODEL_*: only NEE is run (LBM3D/_core.py:31) */
    int poro_model;             /* LBM_PORO_* */
    int bc_flow[LBM_MAX_FACES]; /* LBM_BCFLOW_* (used by init_boundary and the 2D inflow controller) */
    int bc_v[LBM_MAX_FACES];    /* LBM_BC_* for velocity */
    int bc_rho[LBM_MAX_FACES];  /* LBM_BC_* for density */
    double v_bc[LBM_MAX_FACES][3];
    double rho_bc[LBM_MAX_FACES];
    double flow_bc[LBM_MAX_FACES];
    lbm_thermal_params tf, ts;
    int radiation_model;        /* LBM_RAD_* */
    double t_ambient;
    int n_reactions;
    lbm_species_params species[LBM_MAX_SPECIES];
    lbm_reaction_params reactions[LBM_MAX_REACTIONS];
    double mrt_rates[19];       /* LBM_COLL_MRT only: relaxation rates of the non-conserved moments;
                                   a value <= 0 means "use 1/tau" for that moment */
    int check_divergence;       /* 1: lbm_step returns LBM_ERR_DIVERGED when |u| > max_velocity or NaN */
    double max_velocity;
} lbm_params;

typedef struct lbm_handle lbm_handle;

/* ABI version of the library (compare with LBM_ABI_VERSION). */
int lbm_abi_version(void);

/* Allocate a solver of the given shape and modules.
 * Replaces LBM{2,3}DSolver.__init__ (LBM3D/_core.py:12-107, LBM2D/_core.py:11-101). */
int lbm_create(const lbm_desc* desc, lbm_handle** out);

/* Free all host and device memory of the handle. */
void lbm_destroy(lbm_handle* h);

/* Set the model parameters (setters of LBM3D/_io.py:121-489 collected into one struct).
 * May be called again between steps (boundary values changed by UpdateBCfunc callbacks). */
int lbm_set_params(lbm_handle* h, const lbm_params* p);

/* Copy a host array into a field (Taichi field.from_numpy, LBM3D/_io.py:37-57). */
int lbm_set_field(lbm_handle* h, int field, int comp, const void* host, int dtype, size_t n);

/* Copy a field into a host array (Taichi field.to_numpy).  Values are the
 * reference's fields after the last completed step (post-boundary-condition). */
int lbm_get_field(lbm_handle* h, int field, int comp, void* host, int dtype, size_t n);

/* Equilibrium initialisation of every lattice from the macroscopic fields, solid
 * density from FIX species, inlet profile of face 0.
 * Replaces init_simulation -> init_kernel + init_boundary (LBM3D/_initialize.py:17-88). */
int lbm_init(lbm_handle* h);

/* Advance nsteps time steps (each = LBM3D/_evolution.py:11-25 step(); in 2D also the
 * inflow controller LBM2D/_boundary.py:317-333). */
int lbm_step(lbm_handle* h, int nsteps);

/* Scalar diagnostics (get_max_v/get_min_T/get_max_T, LBM3D/_io.py:542-586) and FP64
 * balances.  `comp` selects the species for LBM_REDUCE_SPECIES_SUM. */
int lbm_reduce(lbm_handle* h, int kind, int comp, double* out);

/* Block until all queued work of the handle has completed. */
int lbm_sync(lbm_handle* h);

/* Number of completed lattice steps since lbm_init (tLattice minus the start offset). */
long long lbm_steps_done(const lbm_handle* h);

/* Kernel timing with device events recorded on the handle's stream around every
 * launch of the given kernel class (diagnostics; off by default).  `enable` = 1
 * starts a fresh record, 0 stops recording.  lbm_profile_read synchronises and
 * returns the summed device time (ms) and the number of launches recorded. */
#define LBM_KERNEL_BULK 0 /* streaming + macro + collision of the interior */
#define LBM_KERNEL_NEE 1  /* boundary extrapolation, all face families */
#define LBM_KERNEL_FACE 2 /* collision of the face cells */
#define LBM_KERNEL_STEP 3 /* everything of one time step */
#define LBM_KERNEL_CHEM 4 /* chemistry operator pass (include/lbm_chem.h), when attached */
int lbm_profile(lbm_handle* h, int enable);
int lbm_profile_read(lbm_handle* h, int kernel, double* total_ms, long long* launches);

/* Snapshot / resume.  The reference's export_snapshot is a stub (LBM3D/_io.py:593-595);
 * these copy the complete evolving state of an initialised handle (populations, FIX
 * densities, lagged fields, macroscopic fields, boundary profiles and the step
 * counter) to and from an opaque host blob.  A blob is valid only for a handle
 * created from the same descriptor and parameters by the same library. */
int lbm_state_size(const lbm_handle* h, size_t* bytes);
int lbm_get_state(lbm_handle* h, void* host, size_t bytes);
int lbm_set_state(lbm_handle* h, const void* host, size_t bytes);

/* sizeof(lbm_desc) for which = 0, sizeof(lbm_params) for 1, and for the chemistry operator
 * (lbm_chem.h) sizeof(lbm_chem_mech) for 2, sizeof(lbm_surrogate) for 3 (binding self-check). */
size_t lbm_struct_size(int which);

/* Message of the last error on this handle ("" if none). */
const char* lbm_last_error(const lbm_handle* h);

#ifdef __cplusplus
}
#endif

#endif /* LBM_POROUSCHEM_H */
