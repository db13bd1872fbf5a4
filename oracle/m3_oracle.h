/*
 * m3_oracle.h — CPU oracle for the MPLS_RisingBubble (two-bubble, 3D) step.
 * TEST INFRASTRUCTURE ONLY: the checker the HIP path is compared against,
 * never part of the product.  Only tests/ and bench.py's cpu_baseline leg may
 * load it.
 *
 * A plain-C restatement of the reference subroutines of
 * `MPLS_RisingBubble/20190612/`, in their loop and expression order, on the
 * reference's own arrays: every field spans (-2:NX+3, -2:NY+3, -2:NZ+3),
 * i fastest (the same ghosted layout as fcfd_types.h with g = 3).
 */
#ifndef FCFD_M3_ORACLE_H
#define FCFD_M3_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* fields, named after the reference's arrays (DATA_MOD.F90:26-83) */
enum m3_field {
    M3_PHI = 0, M3_PHI1, M3_PHI2, M3_PHI_OLD,
    M3_U, M3_V, M3_W, M3_U_OLD, M3_V_OLD, M3_W_OLD,
    M3_UH, M3_VH, M3_WH,
    M3_US, M3_VS, M3_WS, M3_US_OLD, M3_VS_OLD, M3_WS_OLD,
    M3_U1, M3_V1, M3_W1,
    M3_RHO, M3_RHO_OLD, M3_AMU, M3_AMU_OLD, M3_P, M3_RHS,
    M3_HEAVY, M3_DELTA, M3_GRAD, M3_SGN,
    M3_NORMAL_X, M3_NORMAL_Y, M3_NORMAL_Z, M3_CURV, M3_S0,
    M3_NFIELDS
};

/* problem parameters (SETUP.F90:14-110) and the MPLS / mass bookkeeping of
 * LS_DATA; b1_im, b2_im, imass_ls, ivol_ls, time, iter and rec_mass are
 * updated by the operations that update them in the reference */
typedef struct m3_params {
    int nx, ny, nz;
    int st_force, g_force;
    int press_max_it;
    int iter, rec_mass;
    double dx, dy, dz, dt, rdt, interface_width;
    double ratio_rho, ratio_amu, re, we, fr;
    double tol_p, tol_m, press_omega;
    double b1_im, b2_im, imass_ls, ivol_ls, time;
} m3_params;

typedef struct m3_diag {
    double mass_ls, vol_ls;   /* FIND_MASS / MPLS sums (no cell volume) */
    double div;               /* CHECK_CONVERGENCE: max |div u| */
    double l2sum;             /* L2_U + L2_V + L2_W of the last Picard iteration */
    int ppe_iters;            /* pressure iterations of the last step */
    int picard;               /* Picard iterations of the last SOLVE_NS */
} m3_diag;

typedef struct m3_state m3_state;

m3_state *m3_create(const m3_params *p);
void m3_destroy(m3_state *s);
int m3_set(m3_state *s, int field, const double *src);
int m3_get(const m3_state *s, int field, double *dst);
void m3_set_params(m3_state *s, const m3_params *p);
void m3_get_params(const m3_state *s, m3_params *p);
void m3_get_diag(const m3_state *s, m3_diag *d);

/* reference subroutine restatements (names follow the reference) */
int m3_bc3d_phi(m3_state *s);
int m3_amurho(m3_state *s);
int m3_heavy_f(m3_state *s);                 /* HEAVY_F(PHI) */
int m3_ls_ccd(m3_state *s);                  /* LS_CCD(PHI) */
int m3_cell_face_velocity(m3_state *s);
int m3_ls_source(m3_state *s);               /* CELL_FACE_VELOCITY + LS_sorce(S0) */
int m3_ls_rk3_solver(m3_state *s);
int m3_gradfi(m3_state *s);                  /* sgnf(PHI) + GRADFI(PHI) */
int m3_ls_redistance(m3_state *s);           /* LS_REDISTANCE(PHI, 1) */
int m3_mpls(m3_state *s, double im);         /* MPLS(im) on PHI */
int m3_solve_ls(m3_state *s);
int m3_second_order_upwind(m3_state *s);
int m3_rhsp(m3_state *s);                    /* Adams_Bashforth is separate */
int m3_adams_bashforth(m3_state *s);
int m3_press_poisson(m3_state *s);           /* the reference SOR */
int m3_press_exact(m3_state *s);             /* converged solve of its operator */
int m3_momentum(m3_state *s);                /* SOLVE_MOMENTUM + VELOCITY_BC */
int m3_solve_ns(m3_state *s, int exact_ppe);
int m3_find_mass(m3_state *s);
int m3_store_data(m3_state *s);
int m3_setup_tail(m3_state *s, int exact_ppe);  /* SETUP.F90:210-222 */
int m3_step(m3_state *s, int exact_ppe);        /* MAIN.F90:10-26 without output */

#ifdef __cplusplus
}
#endif
#endif
