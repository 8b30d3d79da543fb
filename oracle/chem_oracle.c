/*
 * chem_oracle.c -- CPU restatement of the cell-local chemistry operator declared in
 * include/lbm_chem.h.  TEST INFRASTRUCTURE: this file is built into the CPU oracle
 * (liblbm_oracle*.so); only the tests and the bench's CPU baseline load it.
 *
 * The reference has no such operator (its reaction source is the explicit Arrhenius
 * update of src/LBM/LBM3D/_chemical.py:204-270).  What this file restates is the
 * operator as the header defines it:
 *   - gas-phase kinetics of a mechanism table (mass action; three-body collision
 *     efficiencies; Lindemann and Troe fall-off, Gilbert, Luther & Troe, Ber. Bunsenges.
 *     Phys. Chem. 87 (1983) 169; reverse rates through K_c from NASA7 Gibbs energies at
 *     a 1 atm standard state);
 *   - three competing first-order biomass pyrolysis reactions releasing fixed gas
 *     mixtures, with a heat of pyrolysis per kg converted;
 *   - a constant-volume adiabatic energy balance (gas internal energies, solid heat
 *     capacity frozen over the step);
 *   - integrated over dt by the two-stage L-stable Rosenbrock method ROS2 (Verwer, Spee,
 *     Blom & Hundsdorfer, SIAM J. Sci. Comput. 20 (1999) 1456) with its embedded
 *     first-order error estimate, adaptive steps, finite-difference Jacobian and
 *     partially pivoted LU;
 *   - the MLP surrogate (BF16-rounded weights and layer inputs, here accumulated in
 *     double), its training gradient and the Adam update.
 * Compiled with -ffp-contract=off so it rounds like the FP64 device integrator.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "chem_oracle.h"

#define NS LBM_CHEM_NSP
#define NR LBM_CHEM_MAXR
#define NP LBM_CHEM_NPYRO
#define NY (NS + 2) /* implicit state: c[12] (mol/m^3), T (K), rho_B (kg/m^3) */
#define H LBM_MLP_HIDDEN
#define NI LBM_CHEM_NIN
#define NO LBM_CHEM_NOUT

static const double R_GAS = 8.314462618;
static const double P_ATM = 101325.0;

/* ------------------------------------------------------------------ */
/* thermodynamics and rate constants at one temperature                */
/* ------------------------------------------------------------------ */
typedef struct tcache {
    double kf[NR];   /* Arrhenius rate; high-pressure limit for fall-off */
    double k0[NR];   /* fall-off low-pressure limit */
    double lfc[NR];  /* Troe: log10 F_cent */
    double ikc[NR];  /* 1 / K_c */
    double u[NS];    /* molar internal energy, J/mol */
    double cv[NS];   /* molar heat capacity at constant volume, J/(mol K) */
    double kp[NP];   /* pyrolysis rate constants, 1/s */
} tcache;

static void tcache_fill(const lbm_chem_mech* m, double T, tcache* tc) {
    const double lnT = log(T), iT = 1.0 / T;
    double g[NS];
    for (int k = 0; k < NS; ++k) {
        const double* a = T < m->nasa_T[k][1] ? m->nasa[k][0] : m->nasa[k][1];
        const double cpR = a[0] + T * (a[1] + T * (a[2] + T * (a[3] + T * a[4])));
        const double hRT = a[0] + T * (a[1] / 2 + T * (a[2] / 3 + T * (a[3] / 4 + T * a[4] / 5))) + a[5] * iT;
        const double sR = a[0] * lnT + T * (a[1] + T * (a[2] / 2 + T * (a[3] / 3 + T * a[4] / 4))) + a[6];
        g[k] = hRT - sR;
        tc->u[k] = (hRT - 1.0) * R_GAS * T;
        tc->cv[k] = (cpR - 1.0) * R_GAS;
    }
    const double RT_P = R_GAS * T / P_ATM;
    for (int r = 0; r < m->n_reactions; ++r) {
        tc->kf[r] = m->A[r] * exp(m->b[r] * lnT - m->Ea_R[r] * iT);
        if (m->kind[r] >= LBM_CHEM_RXN_LINDEMANN) tc->k0[r] = m->low[r][0] * exp(m->low[r][1] * lnT - m->low[r][2] * iT);
        if (m->kind[r] == LBM_CHEM_RXN_TROE) {
            const double* t = m->troe[r];
            double fc = (1.0 - t[0]) * exp(-T / t[1]) + t[0] * exp(-T / t[2]);
            if (t[3] != 0.0) fc += exp(-t[3] * iT);
            tc->lfc[r] = log10(fc);
        }
        /* K_c = exp(-dG/RT) (P_atm / RT)^dnu, so 1/K_c = exp(dG/RT) (RT / P_atm)^dnu */
        double dG = 0.0, dnu = 0.0;
        for (int i = 0; i < m->n_prod[r]; ++i) {
            dG += m->prod_nu[r][i] * g[m->prod_sp[r][i]];
            dnu += m->prod_nu[r][i];
        }
        for (int i = 0; i < m->n_reac[r]; ++i) {
            dG -= m->reac_nu[r][i] * g[m->reac_sp[r][i]];
            dnu -= m->reac_nu[r][i];
        }
        double ikc = exp(dG);
        const int idn = (int)lround(dnu);
        for (int i = 0; i < idn; ++i) ikc *= RT_P;
        for (int i = 0; i < -idn; ++i) ikc /= RT_P;
        tc->ikc[r] = ikc;
    }
    for (int i = 0; i < NP; ++i) tc->kp[i] = m->pyro_A[i] * exp(-m->pyro_Ea_R[i] * iT);
}

/* right-hand side at state y with the rate constants of y's temperature */
static void rhs(const lbm_chem_mech* m, const tcache* tc, double Csol, const double* y, double* f) {
    const double* c = y;
    double w[NS];
    for (int k = 0; k < NS; ++k) w[k] = 0.0;
    for (int r = 0; r < m->n_reactions; ++r) {
        double k = tc->kf[r];
        if (m->kind[r] != LBM_CHEM_RXN_ELEMENTARY) {
            double cm = 0.0;
            for (int j = 0; j < NS; ++j) cm += m->eff[r][j] * c[j];
            if (m->kind[r] == LBM_CHEM_RXN_THREEBODY) {
                k *= cm;
            } else {
                const double pr = tc->k0[r] * cm / tc->kf[r];
                double F = 1.0;
                if (m->kind[r] == LBM_CHEM_RXN_TROE) {
                    const double lfc = tc->lfc[r];
                    const double lpr = log10(pr > 1e-300 ? pr : 1e-300);
                    const double cc = -0.4 - 0.67 * lfc, nn = 0.75 - 1.27 * lfc;
                    const double f1 = (lpr + cc) / (nn - 0.14 * (lpr + cc));
                    F = pow(10.0, lfc / (1.0 + f1 * f1));
                }
                k *= pr / (1.0 + pr) * F;
            }
        }
        double fw = k, rv = k * tc->ikc[r];
        for (int i = 0; i < m->n_reac[r]; ++i) {
            const double ci = c[m->reac_sp[r][i]];
            for (int t = 0; t < (int)m->reac_nu[r][i]; ++t) fw *= ci;
        }
        for (int i = 0; i < m->n_prod[r]; ++i) {
            const double ci = c[m->prod_sp[r][i]];
            for (int t = 0; t < (int)m->prod_nu[r][i]; ++t) rv *= ci;
        }
        const double q = fw - rv;
        for (int i = 0; i < m->n_reac[r]; ++i) w[m->reac_sp[r][i]] -= m->reac_nu[r][i] * q;
        for (int i = 0; i < m->n_prod[r]; ++i) w[m->prod_sp[r][i]] += m->prod_nu[r][i] * q;
    }
    double heat = 0.0, cv = Csol;
    for (int k = 0; k < NS; ++k) {
        heat += tc->u[k] * w[k];
        cv += c[k] * tc->cv[k];
    }
    double rb = 0.0;
    for (int i = 0; i < NP; ++i) {
        const double ri = tc->kp[i] * y[NS + 1];
        rb += ri;
        heat += m->pyro_dH[i] * ri;
        for (int k = 0; k < NS; ++k) w[k] += m->pyro_gas[i][k] * ri;
    }
    for (int k = 0; k < NS; ++k) f[k] = w[k];
    f[NS] = -heat / cv;
    f[NS + 1] = -rb;
}

/* finite-difference Jacobian J[i][j] = d f_i / d y_j */
static void jacobian(const lbm_chem_mech* m, const tcache* tc, double Csol, double* y, const double* f0,
                     double J[NY][NY]) {
    double f1[NY];
    for (int j = 0; j < NY; ++j) {
        const double yj = y[j];
        const double floor_j = j < NS ? 1e-6 : (j == NS ? 1.0 : 1e-3);
        const double a = fabs(yj) > floor_j ? fabs(yj) : floor_j;
        const double tmp = yj + 1e-7 * a;
        const double d = tmp - yj;
        y[j] = tmp;
        if (j == NS) {
            tcache tt;
            tcache_fill(m, tmp, &tt);
            rhs(m, &tt, Csol, y, f1);
        } else {
            rhs(m, tc, Csol, y, f1);
        }
        y[j] = yj;
        for (int i = 0; i < NY; ++i) J[i][j] = (f1[i] - f0[i]) / d;
    }
}

static void lu_factor(double A[NY][NY], int* piv) {
    for (int k = 0; k < NY; ++k) {
        int p = k;
        double best = fabs(A[k][k]);
        for (int i = k + 1; i < NY; ++i)
            if (fabs(A[i][k]) > best) {
                best = fabs(A[i][k]);
                p = i;
            }
        piv[k] = p;
        if (p != k)
            for (int j = 0; j < NY; ++j) {
                const double t = A[k][j];
                A[k][j] = A[p][j];
                A[p][j] = t;
            }
        if (A[k][k] == 0.0) A[k][k] = 1e-300;
        const double ip = 1.0 / A[k][k];
        for (int i = k + 1; i < NY; ++i) {
            const double l = A[i][k] * ip;
            A[i][k] = l;
            for (int j = k + 1; j < NY; ++j) A[i][j] -= l * A[k][j];
        }
    }
}

static void lu_solve(double A[NY][NY], const int* piv, double* b) {
    for (int k = 0; k < NY; ++k)
        if (piv[k] != k) {
            const double t = b[k];
            b[k] = b[piv[k]];
            b[piv[k]] = t;
        }
    for (int i = 1; i < NY; ++i)
        for (int j = 0; j < i; ++j) b[i] -= A[i][j] * b[j];
    for (int i = NY - 1; i >= 0; --i) {
        for (int j = i + 1; j < NY; ++j) b[i] -= A[i][j] * b[j];
        b[i] /= A[i][i];
    }
}

int chem_oracle_cell(const lbm_chem_mech* m, const double* in, double dt, double* out) {
    const double T0 = in[0], p = in[1], B0 = in[NS + 2], C0 = in[NS + 3];
    double y[NY], c0[NS];
    double sym = 0.0;
    for (int k = 0; k < NS; ++k) sym += in[2 + k] / m->molemass[k];
    const double rho_g = sym > 0.0 ? p / (R_GAS * T0 * sym) : 0.0;
    for (int k = 0; k < NS; ++k) {
        c0[k] = rho_g * in[2 + k] / m->molemass[k];
        y[k] = c0[k];
    }
    y[NS] = T0;
    y[NS + 1] = B0;
    for (int o = 0; o < NO; ++o) out[o] = 0.0;
    if (!(T0 > 0.0) || !(dt > 0.0)) return -1;
    const double Csol = (B0 + C0) * m->cs_solid;
    const double gam = 1.0 + 1.0 / sqrt(2.0);
    double atol[NY];
    for (int k = 0; k < NS; ++k) atol[k] = m->atol_c;
    atol[NS] = m->atol_T;
    atol[NS + 1] = m->atol_B;

    double t = 0.0, h = dt;
    int nacc = 0, natt = 0, rejected = 0, ok = 1, age = 0, have_jac = 0;
    tcache tc;
    double f0[NY], f1[NY], k1[NY], k2[NY], y1[NY], yn[NY], Jac[NY][NY], J[NY][NY];
    int piv[NY];
    for (;;) {
        const int last = h >= dt - t;
        if (last) h = dt - t;
        if (natt >= m->max_steps) {
            ok = 0;
            break;
        }
        ++natt;
        tcache_fill(m, y[NS], &tc);
        rhs(m, &tc, Csol, y, f0);
        if (!have_jac || rejected || age >= m->jac_max_age) {
            jacobian(m, &tc, Csol, y, f0, Jac);
            have_jac = 1;
            age = 0;
        }
        for (int i = 0; i < NY; ++i)
            for (int j = 0; j < NY; ++j) J[i][j] = (i == j ? 1.0 : 0.0) - gam * h * Jac[i][j];
        lu_factor(J, piv);
        for (int i = 0; i < NY; ++i) k1[i] = f0[i];
        lu_solve(J, piv, k1);
        for (int i = 0; i < NY; ++i) y1[i] = y[i] + h * k1[i];
        tcache_fill(m, y1[NS], &tc);
        rhs(m, &tc, Csol, y1, f1);
        for (int i = 0; i < NY; ++i) k2[i] = f1[i] - 2.0 * k1[i];
        lu_solve(J, piv, k2);
        double err = 0.0;
        for (int i = 0; i < NY; ++i) {
            yn[i] = y[i] + 1.5 * h * k1[i] + 0.5 * h * k2[i];
            const double e = 0.5 * h * (k1[i] + k2[i]);
            const double ay = fabs(y[i]) > fabs(yn[i]) ? fabs(y[i]) : fabs(yn[i]);
            const double sc = atol[i] + m->rtol * ay;
            err += (e / sc) * (e / sc);
        }
        err = sqrt(err / NY);
        const double es = err > 1e-10 ? err : 1e-10;
        double fac = 0.9 / sqrt(es);
        if (err <= 1.0) {
            for (int i = 0; i < NY; ++i) y[i] = yn[i];
            ++nacc;
            ++age;
            if (last) break;
            t += h;
            if (fac > 5.0) fac = 5.0;
            if (fac < 0.2) fac = 0.2;
            if (rejected && fac > 1.0) fac = 1.0;
            rejected = 0;
        } else {
            if (fac < 0.2) fac = 0.2;
            if (fac > 1.0) fac = 1.0;
            rejected = 1;
        }
        h *= fac;
    }
    for (int k = 0; k < NS; ++k) out[k] = (y[k] - c0[k]) * m->molemass[k];
    out[NS] = y[NS + 1] - B0;
    out[NS + 1] = y[NS] - T0;
    return ok ? nacc : -1;
}

/* ------------------------------------------------------------------ */
/* element balance                                                      */
/* ------------------------------------------------------------------ */
double lbm_chem_residual(const lbm_chem_mech* m, const double* in, const double* out) {
    const double T = in[0], p = in[1], B = in[NS + 2];
    double sym = 0.0;
    for (int k = 0; k < NS; ++k) sym += in[2 + k] / m->molemass[k];
    const double rho_g = sym > 0.0 && T > 0.0 ? p / (R_GAS * T * sym) : 0.0;
    double dsum = 0.0;
    for (int k = 0; k < NS; ++k) dsum += out[k];
    const double dB = out[NS], dC = -dB - dsum;
    double inv = 0.0, worst = 0.0;
    for (int e = 0; e < 4; ++e) {
        double gas = 0.0, have = 0.0;
        for (int k = 0; k < NS; ++k) {
            gas += m->elements[k][e] * out[k] / m->molemass[k];
            have += m->elements[k][e] * rho_g * in[2 + k] / m->molemass[k];
        }
        have += B / m->biomass_molemass * m->biomass_elements[e];
        double rel = -dB / m->biomass_molemass * m->biomass_elements[e];
        if (e == 0) rel -= dC / m->char_molemass;
        inv += have;
        const double d = fabs(gas - rel);
        if (d > worst) worst = d;
    }
    return inv > 0.0 ? worst / inv : worst;
}

/* ------------------------------------------------------------------ */
/* MLP                                                                  */
/* ------------------------------------------------------------------ */
static float bf16r(float x) {
    /* round to the nearest BF16 value, ties to even (finite inputs) */
    unsigned u;
    memcpy(&u, &x, 4);
    u += 0x7fffu + ((u >> 16) & 1u);
    u &= 0xffff0000u;
    float r;
    memcpy(&r, &u, 4);
    return r;
}

enum { OFF_W1 = 0, OFF_B1 = H * NI, OFF_W2 = OFF_B1 + H, OFF_B2 = OFF_W2 + H * H, OFF_W3 = OFF_B2 + H,
       OFF_B3 = OFF_W3 + NO * H };

typedef struct mlp_trace {
    float xb[NI], h1b[H], h2b[H];
    double z1[H], z2[H], y[NO];
} mlp_trace;

static void mlp_fwd(const float* P, const float* xn, mlp_trace* t) {
    for (int i = 0; i < NI; ++i) t->xb[i] = bf16r(xn[i]);
    for (int j = 0; j < H; ++j) {
        double a = P[OFF_B1 + j];
        for (int i = 0; i < NI; ++i) a += (double)bf16r(P[OFF_W1 + j * NI + i]) * t->xb[i];
        t->z1[j] = a;
        t->h1b[j] = bf16r((float)(a > 0.0 ? a : 0.0));
    }
    for (int j = 0; j < H; ++j) {
        double a = P[OFF_B2 + j];
        for (int i = 0; i < H; ++i) a += (double)bf16r(P[OFF_W2 + j * H + i]) * t->h1b[i];
        t->z2[j] = a;
        t->h2b[j] = bf16r((float)(a > 0.0 ? a : 0.0));
    }
    for (int o = 0; o < NO; ++o) {
        double a = P[OFF_B3 + o];
        for (int i = 0; i < H; ++i) a += (double)bf16r(P[OFF_W3 + o * H + i]) * t->h2b[i];
        t->y[o] = a;
    }
}

int lbm_mlp_forward(lbm_chem* c, const float* params, const float* xn, float* yn, long long n) {
    if (!c || !params || !xn || !yn || n < 0) return LBM_ERR_INVALID;
    mlp_trace t;
    float x[NI];
    for (long long s = 0; s < n; ++s) {
        for (int i = 0; i < NI; ++i) x[i] = xn[(long long)i * n + s];
        mlp_fwd(params, x, &t);
        for (int o = 0; o < NO; ++o) yn[(long long)o * n + s] = (float)t.y[o];
    }
    return LBM_OK;
}

int lbm_mlp_grad(lbm_chem* c, const float* P, const float* xn, const float* yt, long long n, float* grad,
                 double* loss) {
    if (!c || !P || !xn || !yt || !grad || n <= 0) return LBM_ERR_INVALID;
    double* g = (double*)calloc(LBM_MLP_NPARAM, sizeof(double));
    if (!g) return LBM_ERR_OOM;
    double W2b[H * H], W3b[NO * H];
    for (int i = 0; i < H * H; ++i) W2b[i] = bf16r(P[OFF_W2 + i]);
    for (int i = 0; i < NO * H; ++i) W3b[i] = bf16r(P[OFF_W3 + i]);
    const double norm = 1.0 / ((double)n * NO);
    double L = 0.0;
    mlp_trace t;
    float x[NI];
    double dy[NO], dz2[H], dz1[H];
    for (long long s = 0; s < n; ++s) {
        for (int i = 0; i < NI; ++i) x[i] = xn[(long long)i * n + s];
        mlp_fwd(P, x, &t);
        for (int o = 0; o < NO; ++o) {
            const double e = t.y[o] - yt[(long long)o * n + s];
            L += e * e;
            dy[o] = 2.0 * e * norm;
        }
        for (int o = 0; o < NO; ++o) {
            g[OFF_B3 + o] += dy[o];
            for (int i = 0; i < H; ++i) g[OFF_W3 + o * H + i] += dy[o] * t.h2b[i];
        }
        for (int j = 0; j < H; ++j) {
            double a = 0.0;
            for (int o = 0; o < NO; ++o) a += W3b[o * H + j] * dy[o];
            dz2[j] = t.z2[j] > 0.0 ? a : 0.0;
        }
        for (int j = 0; j < H; ++j) {
            g[OFF_B2 + j] += dz2[j];
            for (int i = 0; i < H; ++i) g[OFF_W2 + j * H + i] += dz2[j] * t.h1b[i];
        }
        for (int i = 0; i < H; ++i) {
            double a = 0.0;
            for (int j = 0; j < H; ++j) a += W2b[j * H + i] * dz2[j];
            dz1[i] = t.z1[i] > 0.0 ? a : 0.0;
        }
        for (int j = 0; j < H; ++j) {
            g[OFF_B1 + j] += dz1[j];
            for (int i = 0; i < NI; ++i) g[OFF_W1 + j * NI + i] += dz1[j] * t.xb[i];
        }
    }
    for (int i = 0; i < LBM_MLP_NPARAM; ++i) grad[i] = (float)g[i];
    free(g);
    if (loss) *loss = L * norm;
    return LBM_OK;
}

int lbm_mlp_adam(lbm_chem* c, float* p, const float* g, float* m, float* v, float lr, int t) {
    if (!c || !p || !g || !m || !v || t < 1) return LBM_ERR_INVALID;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const float c1 = (float)(1.0 / (1.0 - pow(0.9, t))), c2 = (float)(1.0 / (1.0 - pow(0.999, t)));
    for (int i = 0; i < LBM_MLP_NPARAM; ++i) {
        m[i] = b1 * m[i] + (1.0f - b1) * g[i];
        v[i] = b2 * v[i] + (1.0f - b2) * g[i] * g[i];
        p[i] -= lr * (m[i] * c1) / (sqrtf(v[i] * c2) + eps);
    }
    return LBM_OK;
}

/* ------------------------------------------------------------------ */
/* context, plain and surrogate operator                                */
/* ------------------------------------------------------------------ */
int lbm_chem_create(const lbm_chem_mech* mech, int device, lbm_chem** out) {
    (void)device;
    if (!mech || !out) return LBM_ERR_INVALID;
    if (mech->n_species != NS || mech->n_reactions < 0 || mech->n_reactions > NR) return LBM_ERR_INVALID;
    lbm_chem* c = (lbm_chem*)calloc(1, sizeof(lbm_chem));
    if (!c) return LBM_ERR_OOM;
    c->mech = *mech;
    *out = c;
    return LBM_OK;
}

void lbm_chem_destroy(lbm_chem* c) { free(c); }

const char* lbm_chem_last_error(const lbm_chem* c) { return c ? c->err : "null chemistry context"; }

int lbm_chem_integrate(lbm_chem* c, const double* in, double* out, int* nsteps, long long n, double dt) {
    if (!c || !in || !out || n < 0) return LBM_ERR_INVALID;
#pragma omp parallel for schedule(dynamic, 64)
    for (long long s = 0; s < n; ++s) {
        double x[NI], y[NO];
        for (int i = 0; i < NI; ++i) x[i] = in[(long long)i * n + s];
        const int st = chem_oracle_cell(&c->mech, x, dt, y);
        for (int o = 0; o < NO; ++o) out[(long long)o * n + s] = y[o];
        if (nsteps) nsteps[s] = st;
    }
    return LBM_OK;
}

int lbm_chem_set_surrogate(lbm_chem* c, const lbm_surrogate* s) {
    if (!c) return LBM_ERR_INVALID;
    if (!s) {
        c->have_surrogate = 0;
        return LBM_OK;
    }
    if (!s->params) return LBM_ERR_INVALID;
    for (int o = 0; o < NO; ++o)
        if (s->out_mul[o] < -1 || s->out_mul[o] >= NI || s->out_kind[o] < 0 || s->out_kind[o] > 3 ||
            (s->out_kind[o] == 2 && !(s->out_lin[o] > 0.0f)) || (s->out_kind[o] == 3 && o >= NS))
            return LBM_ERR_INVALID;
    memcpy(c->params, s->params, sizeof(c->params));
    c->sur = *s;
    c->sur.params = c->params;
    c->have_surrogate = 1;
    return LBM_OK;
}

int chem_oracle_surrogate_cell(const lbm_chem* c, const double* in, double dt, int fallback, double* out) {
    const lbm_surrogate* s = &c->sur;
    int flag = 0;
    float xn[NI];
    for (int i = 0; i < NI; ++i) {
        if (!(in[i] >= (double)s->hull_lo[i] && in[i] <= (double)s->hull_hi[i])) flag = 1; /* NaN: outside */
        xn[i] = ((float)in[i] - s->in_mean[i]) * s->in_scale[i];
    }
    mlp_trace t;
    mlp_fwd(c->params, xn, &t);
    /* decode: output transform by out_kind, per-unit scaling, no gas species below zero */
    double sym = 0.0;
    for (int k = 0; k < NS; ++k) sym += in[2 + k] / c->mech.molemass[k];
    const double rho_g = sym > 0.0 && in[0] > 0.0 ? in[1] / (R_GAS * in[0] * sym) : 0.0;
    for (int o = 0; o < NO; ++o) {
        const double z = (double)((float)t.y[o] * s->out_scale[o] + s->out_mean[o]);
        const double have = o < NS ? rho_g * in[2 + o] : 0.0;
        double v = z;
        if (s->out_kind[o] == 1)
            v = -exp(z);
        else if (s->out_kind[o] == 2)
            v = (double)s->out_lin[o] * sinh(z);
        else if (s->out_kind[o] == 3)
            v = exp(z) - have;
        if (s->out_mul[o] >= 0) v *= in[s->out_mul[o]];
        if (o < NS && v < -have) v = -have;
        out[o] = v;
    }
    if (!flag && !(lbm_chem_residual(&c->mech, in, out) <= s->residual_tol)) flag = 2;
    if (flag && fallback) chem_oracle_cell(&c->mech, in, dt, out);
    return flag;
}

int lbm_chem_surrogate(lbm_chem* c, const float* in, float* out, int* flags, long long n, double dt, int fallback) {
    if (!c || !in || !out || n < 0) return LBM_ERR_INVALID;
    if (!c->have_surrogate) {
        snprintf(c->err, sizeof(c->err), "no surrogate loaded");
        return LBM_ERR_STATE;
    }
#pragma omp parallel for schedule(dynamic, 64)
    for (long long s = 0; s < n; ++s) {
        double x[NI], y[NO];
        for (int i = 0; i < NI; ++i) x[i] = in[(long long)i * n + s];
        const int f = chem_oracle_surrogate_cell(c, x, dt, fallback, y);
        for (int o = 0; o < NO; ++o) out[(long long)o * n + s] = (float)y[o];
        if (flags) flags[s] = f;
    }
    return LBM_OK;
}

/* Oracle-only export: the operator's right-hand side at state y = (c[12] mol/m^3, T K,
 * rho_B kg/m^3) with solid heat capacity Csol (J/(m^3 K)); f receives dy/dt.  The tests
 * integrate it with library ODE solvers and check conservation laws on it. */
int lbm_chem_oracle_rhs(const lbm_chem_mech* m, const double* y, double Csol, double* f) {
    if (!m || !y || !f) return LBM_ERR_INVALID;
    tcache tc;
    tcache_fill(m, y[NS], &tc);
    rhs(m, &tc, Csol, y, f);
    return LBM_OK;
}
