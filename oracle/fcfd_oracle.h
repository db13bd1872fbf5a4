/*
 * fcfd_oracle.h — CPU oracle for the two-phase step.  TEST INFRASTRUCTURE
 * ONLY: it is the checker that the HIP path is compared against, never part
 * of the product.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it.
 *
 * It is a plain-C restatement of the Fortran reference subroutines, one C
 * function per reference subroutine, keeping the reference's loop order and
 * expression order so that single-threaded runs agree with the compiled
 * reference to round-off.  Every function cites the file:line it restates.
 * State mirrors the reference's p%loc members (now / old / tmp levels).
 */
#ifndef FCFD_ORACLE_H
#define FCFD_ORACLE_H

#include "../include/fcfd_types.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_state orc_state;

/* oracle-only field ids (beyond the shared enum fcfd_field) */
enum orc_extra_field {
    ORC_RHO_OLD = 100, ORC_MU_OLD, ORC_DELTA_OLD, ORC_CURV_OLD, ORC_NX_OLD, ORC_NY_OLD, ORC_NZ_OLD,
    ORC_COE_R, ORC_COE_L, ORC_COE_F, ORC_COE_B, ORC_COE_U, ORC_COE_D, ORC_COE_C
};

orc_state *orc_create(const fcfd_grid *grid, const fcfd_params *prm);
void orc_destroy(orc_state *s);
int orc_set_threads(int n);                   /* OpenMP threads, returns used */

int orc_set(orc_state *s, int field, const double *src);
int orc_get(orc_state *s, int field, double *dst);
int orc_set_time(orc_state *s, double time, int iter);
void orc_get_diag(const orc_state *s, fcfd_diag *d);
void orc_set_imass(orc_state *s, double imass);

/* reference subroutine restatements (names follow the reference) */
int orc_bc(orc_state *s, int field);
int orc_velbc(orc_state *s);
int orc_ls_funs(orc_state *s);
int orc_rho_mu(orc_state *s);
int orc_ls_mv(orc_state *s);
int orc_level_set_rk3_solver(orc_state *s);
int orc_level_set_rk3_redis(orc_state *s, int btn);   /* returns iterations */
int orc_mass_preserving_level_set(orc_state *s);
int orc_uccd(orc_state *s);                  /* UCCD phi derivatives -> normals (test entry) */
int orc_level_set_source(orc_state *s);      /* source of the configured scheme -> grad (test entry) */
int orc_surface_norms(orc_state *s);
int orc_curv(orc_state *s);
int orc_ppe_sor_init(orc_state *s);
int orc_ppe_sor_solver(orc_state *s, double tol);     /* reference relaxation */
int orc_ppe_exact_solver(orc_state *s, double rel_tol);/* converged solve      */
int orc_ppe_solve_rhs(orc_state *s, double rel_tol);  /* converged solve of coe_src alone */
void orc_set_band_max(long ndoubles);  /* direct solve up to this band size, else PCG */
int orc_ns_momentum_correction(orc_state *s);
int orc_ns_linearize(orc_state *s);
int orc_ns_ab_setup(orc_state *s);
int orc_ns_ab_predictor(orc_state *s);
int orc_node_vel(orc_state *s);
int orc_ns_check_convergence_div(orc_state *s);
int orc_switch(orc_state *s);
int orc_init_levels(orc_state *s);                    /* p%loc%init          */
int orc_interface_solver(orc_state *s);               /* mpls()              */
int orc_ns_solver(orc_state *s, int exact_ppe);
int orc_step(orc_state *s, int exact_ppe);            /* one main-loop step  */
int orc_problem_init(orc_state *s, int redis0);       /* post-IC part of problem_init */

/* apply the reference operator of ppe_sor_init: out = A p (interior) */
int orc_ppe_apply(orc_state *s, const double *p, double *out);

#ifdef __cplusplus
}
#endif
#endif
