/*
 * spc2 oracle: CPU restatement of the reference's force-summation and stepping
 * path, used ONLY as the checker in tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py.  Nothing in the product links or calls it.
 *
 * Every function follows the reference arithmetic in the same operation order
 * (FP contraction off), with N and L at run time instead of the compile-time
 * kN / kL (Definitions.hpp:20-22).  Citations are into the reference tree.
 */
#ifndef SPC2_ORACLE_H
#define SPC2_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_CHIMERA = 0, ORC_VORTEX = 1 };

typedef struct orc_sys {
  int64_t n;      /* particles */
  double Lx, Ly;  /* periodic box */
  double rho;     /* interaction radius */
  int32_t system; /* ORC_CHIMERA | ORC_VORTEX */
  int32_t threads;
  /* chimera (ChimeraSystemPtr.hpp:16-23); v0 is the vortex "velocity" too */
  double v0, sigma, alpha, D;
  /* vortex arrays (VortexArraysSystemPtr.hpp:16-25) */
  double mu_plus, mu_minus, xi_a, xi_r, kappa;
} orc_sys;

/* ---- inputs ------------------------------------------------------------ */
void orc_splitmix_uniform(uint64_t seed, int64_t n, double lo, double hi, double *out);
void orc_make_state(uint64_t seed, int32_t mode, int64_t n, double L, double rho, double *state_aos);

/* ---- geometry ---------------------------------------------------------- */
void orc_min_image(double L, const double *pts4, int64_t m, double *dxy);
void orc_wrap_points(double L, const double *in2, int64_t m, double *out2);
void orc_wrap_state(const orc_sys *s, double *state_aos);
int32_t orc_num_cells(double L, double rho);

/* ---- cell list --------------------------------------------------------- */
int orc_cell_list(const orc_sys *s, int32_t ncx, int32_t ncy, const double *pos_aos, int32_t *cell_of,
                  int32_t *cell_start, int32_t *sorted, int64_t *clamped);

/* ---- right-hand sides -------------------------------------------------- */
/* counts[0] = candidate pairs tested, counts[1] = pairs with r <= rho,
 * counts[2] = cohesion pairs (vortex r <= xi_r); may be NULL. */
int orc_rhs_k(const orc_sys *s, const double *state, const double *k_prev, double k_coef, double dt, double *k_next,
              int64_t *counts);
/* orc_rhs_k that also writes each particle's term scale (may be NULL): the
 * mean absolute term of its phase derivative, sigma (1/n) sum |sin| (chimera)
 * or (1/(n0+n1)) sum |W sin| + (1/n2) sum |kappa sin| (vortex arrays); 0 for a
 * particle without partners. */
int orc_rhs_k_scaled(const orc_sys *s, const double *state, const double *k_prev, double k_coef, double dt,
                     double *k_next, int64_t *counts, double *tscale);
/* per-particle partner counts of pass 1 at the state: within rho, within xi_r */
int orc_neighbour_counts(const orc_sys *s, const double *state, int32_t *n_rho, int32_t *n_xi_r);
int orc_rhs_ito(const orc_sys *s, const double *state, double *derivative, double *additional, int64_t *counts);

/* ---- steppers (state updated in place, engine wrap applied after) ------ */
int orc_step_se(const orc_sys *s, double *state, double dt, const double *omega, const double *xi, int32_t variant);
int orc_step_srk(const orc_sys *s, double *state, double dt, const double *omega, const double *xi2);
int orc_step_rk4(const orc_sys *s, double *state, double dt, double *kbuf);

/* ---- observer ---------------------------------------------------------- */
void orc_order_param(const double *state, int64_t n, double *R, double *psi);
void orc_order_param_f32(const double *state, int64_t n, float *R, float *psi);

/* ---- noise ------------------------------------------------------------- */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
void orc_philox_normals(uint64_t seed, uint64_t step, uint32_t stream, int64_t n, int32_t per_particle, double *out);
void orc_init_state(uint64_t seed, int64_t n, double Lx, double Ly, double omega_std, double *state_aos, double *omega);
size_t orc_mt_state_size(void);
void orc_mt_seed(void *st, uint32_t seed);
void orc_mt_normals(void *st, int64_t count, double *out);
void orc_mt_uniform(void *st, int64_t count, double lo, double hi, double *out);

#ifdef __cplusplus
}
#endif
#endif
