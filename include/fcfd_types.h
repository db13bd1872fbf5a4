/*
 * fcfd_types.h — plain-C types shared by the MI355X library (libfcfd) and the
 * CPU oracle: grid descriptor, physical parameters, field identifiers, status
 * codes and diagnostics.
 *
 * Data layout contract (identical to the Fortran reference so host arrays map
 * 1:1, `Two phase flows/newOpenMP/2D/lib/data/branches.f90:64-71`):
 *   every field is a real(8) array over (1-g:nx+g, 1-g:ny+g[, 1-g:nz+g]),
 *   column-major (i fastest), g = ghc = 3.  In 2D nz = 1 and there is no z
 *   ghost layer, so a 2D field holds (nx+2g)*(ny+2g) values.
 *   MAC staggering: u(i) is the east face of cell i, v(j) the north face,
 *   w(k) the top face; phi, p, rho, mu are cell centred
 *   (`Two phase flows/newOpenMP/2D/src/ns_solver/ns_solvers.f90:125`).
 */
#ifndef FCFD_TYPES_H
#define FCFD_TYPES_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes returned by every entry point. */
enum fcfd_status {
    FCFD_OK = 0,
    FCFD_ERR_ARG = 1,            /* bad argument / size / null pointer        */
    FCFD_ERR_HIP = 2,            /* HIP runtime error                          */
    FCFD_ERR_NOT_CONVERGED = 3,  /* iteration cap reached before tolerance     */
    FCFD_ERR_NAN = 4,            /* NaN or Inf met in a reduction              */
    FCFD_ERR_FIELD = 5,          /* field id not held by this context          */
    FCFD_ERR_SINGULAR = 6        /* zero pivot / zero denominator              */
};

/* Grid descriptor (the reference's p%glb node_x.., xstart.., ghc). */
typedef struct fcfd_grid {
    int32_t dim;                 /* 2 or 3                                     */
    int32_t nx, ny, nz;          /* interior cells; nz = 1 in 2D               */
    int32_t ghc;                 /* ghost layers; must be 3                    */
    int32_t pad_;
    double xstart, xend, ystart, yend, zstart, zend;
} fcfd_grid;

/*
 * Physical and numerical parameters after the reference's derivation step
 * (`.../3D+IBM/lib/data/tree.f90:165-201`): dt, rdt and ls_wid are already
 * multiplied by dx; tolerances are absolute values, not exponents.
 */
typedef struct fcfd_params {
    double dt;                   /* time step                                  */
    double rdt;                  /* reinitialisation pseudo time step          */
    double ls_wid;               /* interface half width (1.5 dx)              */
    double rho_12, mu_12;        /* density / viscosity ratio fluid2/fluid1    */
    double re, we, fr;           /* Reynolds, Weber, Froude numbers            */
    double gx, gy, gz;           /* unit gravity direction                     */
    double btn_sf, btn_g;        /* surface tension / gravity switches (0|1)   */
    double imass;                /* reference mass for MPLS (initial mass)     */
    double ppe_tol;              /* PPE relative residual target (1e-10)       */
    double p_w1;                 /* reference relaxation weight (oracle only)  */
    double p_tol;                /* reference max-update tolerance (oracle)    */
    int32_t ubc[2], vbc[2], wbc[2];  /* 1 = no-slip, 2 = slip (bcs.f90)        */
    int32_t redis_interval;      /* reinitialise when iter % interval == 0     */
    int32_t picard_iters;        /* outer Picard iterations per step (4)       */
    int32_t ppe_precision;       /* FCFD_PPE_FP64 or FCFD_PPE_MIXED            */
    int32_t ppe_max_cycles;      /* cap on outer iterations per PPE solve      */
    int32_t ppe_krylov;          /* 1: flexible CG around the V-cycle; 0: plain */
    int32_t ppe_inner;           /* V-cycles per outer iteration (>= 1)        */
    int32_t ls_scheme;           /* FCFD_LS_WENO5 or FCFD_LS_UCCD              */
    int32_t ppe_nu1, ppe_nu2;    /* pre/post RB-GS sweeps per level (0: 2, 2)  */
    int32_t ppe_nu_coarse;       /* pre/post sweeps on levels of <= 1/32 of the
                                    finest cells (0: 6)                         */
    int32_t ppe_fused;           /* FP32 coarse-level sweeps: 0 auto (two per
                                    launch while the tiles fit the device in
                                    one round, else one), 1 one per launch,
                                    2 two per launch, 3 one colour per launch,
                                    4 as 0 on every level below the finest    */
    int32_t ppe_rho_half;        /* 1: the FP32 level-0 smoother and restriction
                                    read the density in FP16, and in 3D those of
                                    level 1 its face coefficients (mixed mode) */
} fcfd_params;

/* Level-set advection source of level_set_rk3_solver: the reference's WENO5
 * flux split, or its upwind-CCD alternative s = -(u phi_x + v phi_y + w phi_z)
 * (`Two phase flows/newOpenMP/3D+IBM/src/interface_solver/
 * level_set_rk3_solver.f90:109,123,137,155-157`). */
enum fcfd_ls_scheme {
    FCFD_LS_WENO5 = 0,
    FCFD_LS_UCCD = 1
};

/* Stencil kernels whose thread-block tile (x threads along i, y along j;
 * one plane per block) can be chosen with fcfd_set_tile.  Every tile gives
 * the same results bit for bit: a cell's arithmetic does not depend on it. */
enum fcfd_tile_kernel {
    FCFD_TILE_PREDICTOR = 0,     /* ns_linearize + upwind advection + AB2      */
    FCFD_TILE_DIFF_SOURCE,       /* sec_part1..3                               */
    FCFD_TILE_PPE_SRC,           /* div u* / dt                                */
    FCFD_TILE_PROJECTION,        /* ns_momentum_correction                     */
    FCFD_TILE_NODE_VEL,          /* node_vel                                   */
    FCFD_NTILE
};

enum fcfd_ppe_precision {
    FCFD_PPE_FP64 = 0,           /* FP64 smoothing/transfer, FP64 everywhere   */
    FCFD_PPE_MIXED = 1           /* FP32 V-cycles inside FP64 defect correction*/
};

/*
 * Field identifiers.  Names follow the reference's p%loc members
 * (`.../3D+IBM/lib/data/branches.f90:40-66`): <member>_<time level>.
 */
enum fcfd_field {
    FCFD_PHI = 0,        /* phi%now                                            */
    FCFD_PHI_OLD,        /* phi%old  (phi at the start of the step)            */
    FCFD_U, FCFD_V, FCFD_W,                 /* vel%{x,y,z}%now (MAC faces)     */
    FCFD_U_OLD, FCFD_V_OLD, FCFD_W_OLD,     /* vel%{x,y,z}%old                 */
    FCFD_NU, FCFD_NV, FCFD_NW,              /* nvel%{x,y,z} (cell centred)     */
    FCFD_SRC_U, FCFD_SRC_V, FCFD_SRC_W,     /* velsrc%{x,y,z}%now              */
    FCFD_SRC_U_OLD, FCFD_SRC_V_OLD, FCFD_SRC_W_OLD, /* velsrc%{x,y,z}%old      */
    FCFD_P,              /* p%now                                              */
    FCFD_HEAVY, FCFD_DELTA, FCFD_SIGN,      /* ls_funs outputs                 */
    FCFD_RHO, FCFD_MU,   /* rho_mu outputs                                     */
    FCFD_GRAD,           /* grad%now                                           */
    FCFD_NX, FCFD_NY, FCFD_NZ,              /* normals%{x,y,z}%now             */
    FCFD_NXX, FCFD_NYY, FCFD_NZZ,           /* normals%{xx,yy,zz}%now          */
    FCFD_NXY, FCFD_NXZ, FCFD_NYZ,           /* normals%{xy,xz,yz}%now          */
    FCFD_CURV,           /* normals%curv%now                                   */
    FCFD_PPE_SRC,        /* coe%src (div u* / dt)                              */
    /* held by the MPLS_RisingBubble step only (allocated on its first call):
     * the two bubbles' level sets and the old density / viscosity
     * (`MPLS_RisingBubble/20190612/DATA_MOD.F90:26-31,52`) */
    FCFD_PHI1, FCFD_PHI2,
    FCFD_RHO_OLD, FCFD_MU_OLD,
    FCFD_NFIELDS
};

/* Diagnostics (the reference's p%glb scalars printed by output.f90). */
typedef struct fcfd_diag {
    double mass, vol;            /* ls_mv sums                                 */
    double imass;                /* MPLS target                                */
    double ppe_rel_resid;        /* last PPE relative residual ||r||/||f||     */
    double vel_div, vel_sdiv;    /* max / mean |div u|                         */
    double mpls_lambda;          /* last MPLS Newton multiplier                */
    int32_t ppe_cycles;          /* V-cycles (or outer corrections) last solve */
    int32_t ppe_cycles_step;     /* summed over the Picard loop of last step   */
    int32_t redis_iters;         /* pseudo steps of the last reinitialisation  */
    int32_t iter;                /* time step counter (p%glb%iter)             */
} fcfd_diag;

/* Bookkeeping of the MPLS_RisingBubble step (`MPLS_RisingBubble/20190612/`
 * DATA_MOD.F90 LS_DATA / PROBLEM_DEF scalars): the per-bubble MPLS targets
 * B1_iM, B2_iM, the initial totals, the last sums and Picard diagnostics.
 * fcfd_m3_setup fills it; it can be read and replaced (restart, tests). */
typedef struct fcfd_m3_state {
    double bubble_imass[2];      /* B1_iM, B2_iM (FIND_MASS.F90:92-118)        */
    double imass, ivol;          /* IMASS_LS, IVOL_LS                          */
    double mass, vol;            /* MASS_LS, VOL_LS of the last FIND_MASS      */
    double div;                  /* DIV: max |div u| of the last iteration     */
    double l2sum;                /* L2_U + L2_V + L2_W of the last iteration   */
    double picard_tol;           /* TOL_M: Picard exit when l2sum < it (1e-5)  */
    int32_t rec_mass;            /* REC_MASS: 0 until the first FIND_MASS      */
    int32_t picard;              /* Picard iterations of the last SOLVE_NS     */
} fcfd_m3_state;

/* Per-case parameters of a batched 2D context (fcfd_create_batch): the
 * physical numbers that differ between the cases of one stack.  Every other
 * parameter (dt, rdt, ls_wid, gravity, switches, tolerances, BCs) is shared
 * and comes from fcfd_params. */
typedef struct fcfd_case {
    double rho_12, mu_12;        /* density / viscosity ratio                 */
    double re, we, fr;           /* Reynolds, Weber, Froude numbers            */
    double imass;                /* MPLS reference mass                        */
} fcfd_case;

#ifdef __cplusplus
}
#endif
#endif /* FCFD_TYPES_H */
