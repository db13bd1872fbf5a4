/*
 * fcfd_oracle.c — CPU restatement of the reference two-phase step.
 *
 * TEST INFRASTRUCTURE: the checker the HIP path is compared against.  Never
 * linked into the product library.  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg load it.
 *
 * Reference files restated (paths under `Two phase flows/newOpenMP/`):
 *   3D: 3D+IBM/src/interface_solver/{level_set_rk3_solver,level_set_rk3_redis,
 *       mass_preserving_level_set,others,others_normals}.f90,
 *       3D+IBM/src/ns_solver/{ns_ppe_solver,ns_solvers,admas_bashforth_solver,
 *       ns_ab_adv,ns_ab_diff}.f90, 3D+IBM/src/others/bcs.f90,
 *       3D+IBM/lib/solver/{ccd_solvers,matrix_solver}.f90
 *   2D: the same files under 2D/.
 * Indices are 0-based: reference cell index r maps to r-1 here.  Loop order
 * (k outer, i inner) and expression order follow the reference so that a
 * single-threaded run agrees with the compiled reference to round-off.
 * Reference x**2.0d0 is written x*x.
 *
 * Each function is pinned against the compiled reference bit for bit, single
 * threaded (tests/test_oracle_vs_reference.py, tests/golden/ref_*.npz); the
 * exact pressure solve against SciPy's sparse LU and the reference relaxation's
 * fixed point (tests/test_oracle_ppe.py, tests/test_oracle_golden.py).
 */
#include "fcfd_oracle.h"
#include "band.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------ state */

enum { L_NOW = 0, L_OLD, L_TMP, L_OLD2, NLEV };

typedef struct { double *lv[NLEV]; } trec;   /* time_recorded: now/old/tmp/old2 */

struct orc_state {
    fcfd_grid gd;
    fcfd_params pr;
    int dim, nx, ny, nz, g, gz;
    long sx, sxy, ntot;
    double dx, dy, dz;
    double time;
    int iter, piter, redis_iters;
    double mass, vol, imass, ppe_err, ppe_rel, vel_div, vel_sdiv, mpls_lam;
    /* p%loc members */
    trec phi, heavy, delta, sign, grad, rho, mu, p;
    trec nx_, ny_, nz_, nxx, nyy, nzz, nxy, nxz, nyz, curv;
    trec u, v, w, nu, nv, nw, su, sv, sw;
    double *coe_r, *coe_l, *coe_f, *coe_b, *coe_u, *coe_d, *coe_c, *coe_src;
    /* tdata%{x,y,z}%{s1,s2,s3,ss1,ss2,l1,l2,l3} */
    double *t_s1[3], *t_s2[3], *t_s3[3], *t_ss1[3], *t_ss2[3], *t_l1, *t_l2, *t_l3;
    /* ccd line work space */
    double *wk;
    /* Cholesky factor of the pinned operator of the last ppe_sor_init (band
     * storage, see ppe_band_factor); NULL until a direct solve needs it */
    double *band;
    long band_n, band_b;
    int band_valid;
};

#define ID(s, i, j, k) ((long)((i) + (s)->g) + (long)((j) + (s)->g) * (s)->sx + (long)((k) + (s)->gz) * (s)->sxy)

static double *falloc(long n) { return (double *)calloc((size_t)n, sizeof(double)); }

static void trec_alloc(orc_state *s, trec *t) {
    for (int l = 0; l < NLEV; ++l) t->lv[l] = falloc(s->ntot);
}
static void trec_free(trec *t) {
    for (int l = 0; l < NLEV; ++l) free(t->lv[l]);
}

static trec *all_trecs(orc_state *s, int idx) {
    trec *list[] = {&s->phi, &s->heavy, &s->delta, &s->sign, &s->grad, &s->rho, &s->mu, &s->p,
                    &s->nx_, &s->ny_, &s->nz_, &s->nxx, &s->nyy, &s->nzz, &s->nxy, &s->nxz, &s->nyz,
                    &s->curv, &s->u, &s->v, &s->w, &s->nu, &s->nv, &s->nw, &s->su, &s->sv, &s->sw};
    int n = (int)(sizeof(list) / sizeof(list[0]));
    return idx < n ? list[idx] : NULL;
}

orc_state *orc_create(const fcfd_grid *grid, const fcfd_params *prm) {
    if (!grid || !prm) return NULL;
    if (grid->dim != 2 && grid->dim != 3) return NULL;
    if (grid->ghc != 3 || grid->nx < 4 || grid->ny < 4) return NULL;
    if (grid->dim == 3 && grid->nz < 4) return NULL;
    if (grid->dim == 2 && grid->nz != 1) return NULL;
    orc_state *s = (orc_state *)calloc(1, sizeof(orc_state));
    s->gd = *grid;
    s->pr = *prm;
    s->dim = grid->dim;
    s->nx = grid->nx; s->ny = grid->ny; s->nz = grid->nz;
    s->g = grid->ghc;
    s->gz = grid->dim == 3 ? grid->ghc : 0;
    s->sx = s->nx + 2 * s->g;
    s->sxy = s->sx * (s->ny + 2 * s->g);
    s->ntot = s->sxy * (s->nz + 2 * s->gz);
    /* dx = (xend - xstart) / node_x  (tree.f90:154-156) */
    s->dx = (grid->xend - grid->xstart) / s->nx;
    s->dy = (grid->yend - grid->ystart) / s->ny;
    s->dz = grid->dim == 3 ? (grid->zend - grid->zstart) / s->nz : 1.0;
    s->imass = prm->imass;
    for (int t = 0; all_trecs(s, t); ++t) trec_alloc(s, all_trecs(s, t));
    double **c[] = {&s->coe_r, &s->coe_l, &s->coe_f, &s->coe_b, &s->coe_u, &s->coe_d, &s->coe_c, &s->coe_src};
    for (int t = 0; t < 8; ++t) *c[t] = falloc(s->ntot);
    for (int d = 0; d < 3; ++d) {
        s->t_s1[d] = falloc(s->ntot); s->t_s2[d] = falloc(s->ntot); s->t_s3[d] = falloc(s->ntot);
        s->t_ss1[d] = falloc(s->ntot); s->t_ss2[d] = falloc(s->ntot);
    }
    s->t_l1 = falloc(s->ntot); s->t_l2 = falloc(s->ntot); s->t_l3 = falloc(s->ntot);
    return s;
}

void orc_destroy(orc_state *s) {
    if (!s) return;
    for (int t = 0; all_trecs(s, t); ++t) trec_free(all_trecs(s, t));
    double *c[] = {s->coe_r, s->coe_l, s->coe_f, s->coe_b, s->coe_u, s->coe_d, s->coe_c, s->coe_src};
    for (int t = 0; t < 8; ++t) free(c[t]);
    for (int d = 0; d < 3; ++d) {
        free(s->t_s1[d]); free(s->t_s2[d]); free(s->t_s3[d]); free(s->t_ss1[d]); free(s->t_ss2[d]);
    }
    free(s->t_l1); free(s->t_l2); free(s->t_l3);
    free(s->band);
    free(s);
}

int orc_set_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
    return omp_get_max_threads();
#else
    (void)n;
    return 1;
#endif
}

static double *field_ptr(orc_state *s, int f) {
    switch (f) {
    case FCFD_PHI: return s->phi.lv[L_NOW];
    case FCFD_PHI_OLD: return s->phi.lv[L_OLD];
    case FCFD_U: return s->u.lv[L_NOW];
    case FCFD_V: return s->v.lv[L_NOW];
    case FCFD_W: return s->w.lv[L_NOW];
    case FCFD_U_OLD: return s->u.lv[L_OLD];
    case FCFD_V_OLD: return s->v.lv[L_OLD];
    case FCFD_W_OLD: return s->w.lv[L_OLD];
    case FCFD_NU: return s->nu.lv[L_NOW];
    case FCFD_NV: return s->nv.lv[L_NOW];
    case FCFD_NW: return s->nw.lv[L_NOW];
    case FCFD_SRC_U: return s->su.lv[L_NOW];
    case FCFD_SRC_V: return s->sv.lv[L_NOW];
    case FCFD_SRC_W: return s->sw.lv[L_NOW];
    case FCFD_SRC_U_OLD: return s->su.lv[L_OLD];
    case FCFD_SRC_V_OLD: return s->sv.lv[L_OLD];
    case FCFD_SRC_W_OLD: return s->sw.lv[L_OLD];
    case FCFD_P: return s->p.lv[L_NOW];
    case FCFD_HEAVY: return s->heavy.lv[L_NOW];
    case FCFD_DELTA: return s->delta.lv[L_NOW];
    case FCFD_SIGN: return s->sign.lv[L_NOW];
    case FCFD_RHO: return s->rho.lv[L_NOW];
    case FCFD_MU: return s->mu.lv[L_NOW];
    case FCFD_GRAD: return s->grad.lv[L_NOW];
    case FCFD_NX: return s->nx_.lv[L_NOW];
    case FCFD_NY: return s->ny_.lv[L_NOW];
    case FCFD_NZ: return s->nz_.lv[L_NOW];
    case FCFD_NXX: return s->nxx.lv[L_NOW];
    case FCFD_NYY: return s->nyy.lv[L_NOW];
    case FCFD_NZZ: return s->nzz.lv[L_NOW];
    case FCFD_NXY: return s->nxy.lv[L_NOW];
    case FCFD_NXZ: return s->nxz.lv[L_NOW];
    case FCFD_NYZ: return s->nyz.lv[L_NOW];
    case FCFD_CURV: return s->curv.lv[L_NOW];
    case FCFD_PPE_SRC: return s->coe_src;
    /* oracle-only ids: the reference's old levels and PPE coefficients */
    case ORC_RHO_OLD: return s->rho.lv[L_OLD];
    case ORC_MU_OLD: return s->mu.lv[L_OLD];
    case ORC_DELTA_OLD: return s->delta.lv[L_OLD];
    case ORC_CURV_OLD: return s->curv.lv[L_OLD];
    case ORC_NX_OLD: return s->nx_.lv[L_OLD];
    case ORC_NY_OLD: return s->ny_.lv[L_OLD];
    case ORC_NZ_OLD: return s->nz_.lv[L_OLD];
    case ORC_COE_R: return s->coe_r;
    case ORC_COE_L: return s->coe_l;
    case ORC_COE_F: return s->coe_f;
    case ORC_COE_B: return s->coe_b;
    case ORC_COE_U: return s->coe_u;
    case ORC_COE_D: return s->coe_d;
    case ORC_COE_C: return s->coe_c;
    default: return NULL;
    }
}

int orc_set(orc_state *s, int f, const double *src) {
    double *p = field_ptr(s, f);
    if (!p || !src) return FCFD_ERR_FIELD;
    memcpy(p, src, sizeof(double) * (size_t)s->ntot);
    s->band_valid = 0;
    return FCFD_OK;
}

int orc_get(orc_state *s, int f, double *dst) {
    double *p = field_ptr(s, f);
    if (!p || !dst) return FCFD_ERR_FIELD;
    memcpy(dst, p, sizeof(double) * (size_t)s->ntot);
    return FCFD_OK;
}

int orc_set_time(orc_state *s, double time, int iter) {
    s->time = time;
    s->iter = iter;
    return FCFD_OK;
}

void orc_set_imass(orc_state *s, double imass) { s->imass = imass; }

void orc_get_diag(const orc_state *s, fcfd_diag *d) {
    memset(d, 0, sizeof(*d));
    d->mass = s->mass; d->vol = s->vol; d->imass = s->imass;
    d->ppe_rel_resid = s->ppe_rel; d->vel_div = s->vel_div; d->vel_sdiv = s->vel_sdiv;
    d->mpls_lambda = s->mpls_lam; d->ppe_cycles = s->piter; d->ppe_cycles_step = s->piter;
    d->redis_iters = s->redis_iters; d->iter = s->iter;
}

/* loop helpers over the interior and the full ghosted box */
#define KLO(s) 0
#define KHI(s) ((s)->nz)
#define FOR_INT(s) for (int k = 0; k < (s)->nz; ++k) for (int j = 0; j < (s)->ny; ++j) for (int i = 0; i < (s)->nx; ++i)
#define ORC_PRAGMA(x) _Pragma(#x)
/* OpenMP versions over the interior (used by the converged pressure solve) */
#define PFOR_INT(s) ORC_PRAGMA(omp parallel for collapse(2)) FOR_INT(s)
/* Deterministic parallel sum over the interior: every thread adds its static
 * share of the (k, j) rows in loop order into a private v, and the thread
 * partials are added in thread order after the loop.  An OpenMP reduction
 * combines them in arrival order, which made the converged pressure differ
 * between runs at the 1e-11 level; this sum is the same in every run with
 * a given thread count, and with one thread it is the sequential sum.
 * Use: PSUM_BEGIN(s, v) { body adding to v } PSUM_END(v) */
#define ORC_MAXT 256
static int orc_nthreads(void) {
    const int n = omp_get_max_threads();
    return n < ORC_MAXT ? n : ORC_MAXT;
}
#define PSUM_BEGIN(s, v)                                                        \
    {                                                                           \
        double v##_p_[ORC_MAXT];                                                \
        int v##_n_ = 1;                                                         \
        ORC_PRAGMA(omp parallel num_threads(orc_nthreads())) {                  \
            double v = 0.0;                                                     \
            ORC_PRAGMA(omp for collapse(2) schedule(static)) FOR_INT(s)
#define PSUM_END(v)                                                             \
            v##_p_[omp_get_thread_num()] = v;                                   \
            if (omp_get_thread_num() == 0) v##_n_ = omp_get_num_threads();       \
        }                                                                       \
        for (int t_ = 0; t_ < v##_n_; ++t_) v += v##_p_[t_];                   \
    }
#define FOR_ALL(s)                                           \
    for (int k = -(s)->gz; k < (s)->nz + (s)->gz; ++k)       \
        for (int j = -(s)->g; j < (s)->ny + (s)->g; ++j)     \
            for (int i = -(s)->g; i < (s)->nx + (s)->g; ++i)

/* ----------------------------------------------------------- bcs.f90 */

/* bc(dat): zero-gradient copy of every ghost layer, x then y then z
 * (3D+IBM/src/others/bcs.f90:1-42; 2D/src/others/bcs.f90:1-26). */
static void bc(orc_state *s, double *d) {
    const int g = s->g, nx = s->nx, ny = s->ny, nz = s->nz, gz = s->gz;
#pragma omp parallel for collapse(2)
    for (int k = -gz; k < nz + gz; ++k)
        for (int j = -g; j < ny + g; ++j)
            for (int ii = 1; ii <= g; ++ii) {
                d[ID(s, -ii, j, k)] = d[ID(s, 0, j, k)];
                d[ID(s, nx - 1 + ii, j, k)] = d[ID(s, nx - 1, j, k)];
            }
#pragma omp parallel for collapse(2)
    for (int k = -gz; k < nz + gz; ++k)
        for (int i = -g; i < nx + g; ++i)
            for (int jj = 1; jj <= g; ++jj) {
                d[ID(s, i, -jj, k)] = d[ID(s, i, 0, k)];
                d[ID(s, i, ny - 1 + jj, k)] = d[ID(s, i, ny - 1, k)];
            }
    if (s->dim == 3) {
#pragma omp parallel for collapse(2)
        for (int j = -g; j < ny + g; ++j)
            for (int i = -g; i < nx + g; ++i)
                for (int kk = 1; kk <= g; ++kk) {
                    d[ID(s, i, j, -kk)] = d[ID(s, i, j, 0)];
                    d[ID(s, i, j, nz - 1 + kk)] = d[ID(s, i, j, nz - 1)];
                }
    }
}

int orc_bc(orc_state *s, int field) {
    double *p = field_ptr(s, field);
    if (!p) return FCFD_ERR_FIELD;
    bc(s, p);
    return FCFD_OK;
}

/* velbc(u,v[,w]): MAC no-slip (1) / slip (2) ghost rules, x then y then z.
 * (3D+IBM/src/others/bcs.f90:44-257; 2D/src/others/bcs.f90:28-148).
 * `nodal` selects nvelbc (cell-centred rules, bcs.f90:259-400 in 3D). */
static void velbc_impl(orc_state *s, double *u, double *v, double *w, int nodal) {
    const int g = s->g, nx = s->nx, ny = s->ny, nz = s->nz, gz = s->gz;
    const int d3 = s->dim == 3;
    /* X direction */
    for (int side = 0; side < 2; ++side) {
        int bt = s->pr.ubc[side];
        if (bt != 1 && bt != 2) continue;
        double sg = bt == 1 ? -1.0 : 1.0;
#pragma omp parallel for collapse(2)
        for (int k = -gz; k < nz + gz; ++k)
            for (int j = -g; j < ny + g; ++j) {
                for (int ii = 1; ii <= g; ++ii) {
                    if (side == 0) {
                        u[ID(s, -ii, j, k)] = sg * u[ID(s, nodal ? ii - 1 : ii - 2, j, k)];
                        v[ID(s, -ii, j, k)] = sg * v[ID(s, ii - 1, j, k)];
                        if (d3) w[ID(s, -ii, j, k)] = sg * w[ID(s, ii - 1, j, k)];
                    } else {
                        u[ID(s, nx - 1 + ii, j, k)] = sg * u[ID(s, nodal ? nx - ii : nx - 1 - ii, j, k)];
                        v[ID(s, nx - 1 + ii, j, k)] = sg * v[ID(s, nx - ii, j, k)];
                        if (d3) w[ID(s, nx - 1 + ii, j, k)] = sg * w[ID(s, nx - ii, j, k)];
                    }
                }
                if (!nodal) u[ID(s, side == 0 ? -1 : nx - 1, j, k)] = 0.0;
            }
    }
    /* Y direction */
    for (int side = 0; side < 2; ++side) {
        int bt = s->pr.vbc[side];
        if (bt != 1 && bt != 2) continue;
        double sg = bt == 1 ? -1.0 : 1.0;
#pragma omp parallel for collapse(2)
        for (int k = -gz; k < nz + gz; ++k)
            for (int i = -g; i < nx + g; ++i) {
                for (int jj = 1; jj <= g; ++jj) {
                    if (side == 0) {
                        u[ID(s, i, -jj, k)] = sg * u[ID(s, i, jj - 1, k)];
                        v[ID(s, i, -jj, k)] = sg * v[ID(s, i, nodal ? jj - 1 : jj - 2, k)];
                        if (d3) w[ID(s, i, -jj, k)] = sg * w[ID(s, i, jj - 1, k)];
                    } else {
                        u[ID(s, i, ny - 1 + jj, k)] = sg * u[ID(s, i, ny - jj, k)];
                        v[ID(s, i, ny - 1 + jj, k)] = sg * v[ID(s, i, nodal ? ny - jj : ny - 1 - jj, k)];
                        if (d3) w[ID(s, i, ny - 1 + jj, k)] = sg * w[ID(s, i, ny - jj, k)];
                    }
                }
                if (!nodal) v[ID(s, i, side == 0 ? -1 : ny - 1, k)] = 0.0;
            }
    }
    if (!d3) return;
    /* Z direction */
    for (int side = 0; side < 2; ++side) {
        int bt = s->pr.wbc[side];
        if (bt != 1 && bt != 2) continue;
        double sg = bt == 1 ? -1.0 : 1.0;
#pragma omp parallel for collapse(2)
        for (int j = -g; j < ny + g; ++j)
            for (int i = -g; i < nx + g; ++i) {
                for (int kk = 1; kk <= g; ++kk) {
                    if (side == 0) {
                        u[ID(s, i, j, -kk)] = sg * u[ID(s, i, j, kk - 1)];
                        v[ID(s, i, j, -kk)] = sg * v[ID(s, i, j, kk - 1)];
                        w[ID(s, i, j, -kk)] = sg * w[ID(s, i, j, nodal ? kk - 1 : kk - 2)];
                    } else {
                        u[ID(s, i, j, nz - 1 + kk)] = sg * u[ID(s, i, j, nz - kk)];
                        v[ID(s, i, j, nz - 1 + kk)] = sg * v[ID(s, i, j, nz - kk)];
                        w[ID(s, i, j, nz - 1 + kk)] = sg * w[ID(s, i, j, nodal ? nz - kk : nz - 1 - kk)];
                    }
                }
                if (!nodal) w[ID(s, i, j, side == 0 ? -1 : nz - 1)] = 0.0;
            }
    }
}

int orc_velbc(orc_state *s) {
    velbc_impl(s, s->u.lv[L_NOW], s->v.lv[L_NOW], s->w.lv[L_NOW], 0);
    return FCFD_OK;
}

/* -------------------------------------------------------- others.f90 */

/* ls_funs: smoothed Heaviside, delta and sign over the full ghosted box
 * (3D+IBM/src/interface_solver/others.f90:56-94; 2D others.f90:49-85). */
int orc_ls_funs(orc_state *s) {
    const double eps = 1.0e-12, pi = acos(-1.0), w = s->pr.ls_wid;
    double *phi = s->phi.lv[L_NOW], *H = s->heavy.lv[L_NOW], *D = s->delta.lv[L_NOW], *S = s->sign.lv[L_NOW];
#pragma omp parallel for collapse(2)
    for (int k = -s->gz; k < s->nz + s->gz; ++k)
        for (int j = -s->g; j < s->ny + s->g; ++j)
            for (int i = -s->g; i < s->nx + s->g; ++i) {
                long id = ID(s, i, j, k);
                double x = phi[id] / w, heavy, hp;
                if (x > 1.0 - eps) { heavy = 1.0; hp = 0.0; }
                else if (x < -1.0 + eps) { heavy = 0.0; hp = 0.0; }
                else {
                    heavy = 0.5 * (1.0 + x + sin(pi * x) / pi);
                    hp = 0.5 * (1.0 + cos(pi * x)) / w;
                }
                heavy = fmax(fmin(heavy, 1.0), 0.0);
                H[id] = heavy;
                D[id] = hp;
                S[id] = 2.0 * heavy - 1.0;
            }
    return FCFD_OK;
}

/* rho_mu (others.f90:96-120 in 3D; :87-109 in 2D) */
int orc_rho_mu(orc_state *s) {
    orc_ls_funs(s);
    double *H = s->heavy.lv[L_NOW], *R = s->rho.lv[L_NOW], *M = s->mu.lv[L_NOW];
    const double r12 = s->pr.rho_12, m12 = s->pr.mu_12;
#pragma omp parallel for
    for (long id = 0; id < s->ntot; ++id) {
        double heavy = H[id];
        R[id] = heavy + r12 * (1.0 - heavy);
        M[id] = heavy + m12 * (1.0 - heavy);
    }
    return FCFD_OK;
}

/* ls_mv: mass = sum rho*H*dV, vol = sum H*dV (others.f90:122-152 in 3D) */
int orc_ls_mv(orc_state *s) {
    const double dv = s->dim == 3 ? s->dx * s->dy * s->dz : s->dx * s->dy;
    orc_rho_mu(s);
    double *H = s->heavy.lv[L_NOW];
    const double r12 = s->pr.rho_12;
    double mass = 0.0, vol = 0.0;
    {   /* deterministic parallel sums, as PSUM_BEGIN / PSUM_END */
        double mp[ORC_MAXT], vp[ORC_MAXT];
        int nt = 1;
#pragma omp parallel num_threads(orc_nthreads())
        {
            double m = 0.0, v = 0.0;
#pragma omp for collapse(2) schedule(static)
            for (int k = 0; k < s->nz; ++k)
                for (int j = 0; j < s->ny; ++j)
                    for (int i = 0; i < s->nx; ++i) {
                        double h = H[ID(s, i, j, k)];
                        double rho = h + r12 * (1.0 - h);
                        m = m + rho * h * dv;
                        v = v + h * dv;
                    }
            mp[omp_get_thread_num()] = m;
            vp[omp_get_thread_num()] = v;
            if (omp_get_thread_num() == 0) nt = omp_get_num_threads();
        }
        for (int t = 0; t < nt; ++t) {
            mass = mass + mp[t];
            vol = vol + vp[t];
        }
    }
    s->mass = mass;
    s->vol = vol;
    return FCFD_OK;
}

/* ------------------------------------------------ level_set_rk3_solver */

/* wenojs_flux_split(f,g,fp,gm): WENO5-JS face values, fp right-biased from f,
 * gm left-biased from g, faces is-ghc+2..ie+ghc-3, EPS = 1e-13, beta without
 * the 1/12 and 1/4 factors (3D+IBM/.../level_set_rk3_solver.f90:162-219). */
static void wenojs_flux_split(const double *f, const double *gg, double *fp, double *gm, int n, int g, long st) {
    const double EPS = 1.0e-13;
#define F(ii) f[(long)(ii) * st]
#define G(ii) gg[(long)(ii) * st]
    for (int i = 2 - g; i <= n + g - 4; ++i) {
        double t1 = G(i - 2) - 2.0 * G(i - 1) + G(i), t2 = G(i - 2) - 4.0 * G(i - 1) + 3.0 * G(i);
        double b1 = 13.0 * (t1 * t1) + 3.0 * (t2 * t2);
        t1 = G(i - 1) - 2.0 * G(i) + G(i + 1); t2 = G(i - 1) - G(i + 1);
        double b2 = 13.0 * (t1 * t1) + 3.0 * (t2 * t2);
        t1 = G(i) - 2.0 * G(i + 1) + G(i + 2); t2 = 3.0 * G(i) - 4.0 * G(i + 1) + G(i + 2);
        double b3 = 13.0 * (t1 * t1) + 3.0 * (t2 * t2);
        double a1 = 1.0 / ((EPS + b1) * (EPS + b1));
        double a2 = 6.0 / ((EPS + b2) * (EPS + b2));
        double a3 = 3.0 / ((EPS + b3) * (EPS + b3));
        double w1 = a1 / (a1 + a2 + a3), w2 = a2 / (a1 + a2 + a3), w3 = a3 / (a1 + a2 + a3);
        gm[(long)i * st] = w1 / 3.0 * G(i - 2) - (7.0 * w1 + w2) / 6.0 * G(i - 1) +
                           (11.0 * w1 + 5.0 * w2 + 2.0 * w3) / 6.0 * G(i) + (2.0 * w2 + 5.0 * w3) / 6.0 * G(i + 1) -
                           w3 / 6.0 * G(i + 2);
    }
    for (int i = 2 - g; i <= n + g - 4; ++i) {
        double t1 = F(i - 1) - 2.0 * F(i) + F(i + 1), t2 = F(i - 1) - 4.0 * F(i) + 3.0 * F(i + 1);
        double b3 = 13.0 * (t1 * t1) + 3.0 * (t2 * t2);
        t1 = F(i) - 2.0 * F(i + 1) + F(i + 2); t2 = F(i) - F(i + 2);
        double b2 = 13.0 * (t1 * t1) + 3.0 * (t2 * t2);
        t1 = F(i + 1) - 2.0 * F(i + 2) + F(i + 3); t2 = 3.0 * F(i + 1) - 4.0 * F(i + 2) + F(i + 3);
        double b1 = 13.0 * (t1 * t1) + 3.0 * (t2 * t2);
        double a1 = 1.0 / ((EPS + b1) * (EPS + b1));
        double a2 = 6.0 / ((EPS + b2) * (EPS + b2));
        double a3 = 3.0 / ((EPS + b3) * (EPS + b3));
        double w1 = a1 / (a1 + a2 + a3), w2 = a2 / (a1 + a2 + a3), w3 = a3 / (a1 + a2 + a3);
        fp[(long)i * st] = w3 * (-F(i - 1) + 5.0 * F(i) + 2.0 * F(i + 1)) / 6.0 +
                           w2 * (2.0 * F(i) + 5.0 * F(i + 1) - F(i + 2)) / 6.0 +
                           w1 * (11.0 * F(i + 1) - 7.0 * F(i + 2) + 2.0 * F(i + 3)) / 6.0;
    }
#undef F
#undef G
}

/* level_set_rk3_source_setup (3D :67-96, 2D :61-84) */
static void ls_source_setup(orc_state *s) {
    double *phi = s->phi.lv[L_NOW];
    double *vel[3] = {s->nu.lv[L_OLD], s->nv.lv[L_OLD], s->nw.lv[L_OLD]};
    for (int d = 0; d < s->dim; ++d) {
        double *u = vel[d], *s1 = s->t_s1[d], *s2 = s->t_s2[d];
#pragma omp parallel for collapse(2)
        for (int k = 0; k < s->nz; ++k)
            for (int j = 0; j < s->ny; ++j)
                for (int i = 0; i < s->nx; ++i) {
                    long id = ID(s, i, j, k);
                    s1[id] = 0.5 * (u[id] + fabs(u[id])) * phi[id];
                    s2[id] = 0.5 * (u[id] - fabs(u[id])) * phi[id];
                }
    }
    for (int d = 0; d < s->dim; ++d) { bc(s, s->t_s1[d]); bc(s, s->t_s2[d]); }
}

/* level_set_rk3_source(s) (3D :98-160, 2D :86-129) */
static void ls_source(orc_state *s, double *out) {
    const int g = s->g;
    /* x lines */
#pragma omp parallel for collapse(2)
    for (int k = 0; k < s->nz; ++k)
        for (int j = 0; j < s->ny; ++j) {
            long o = ID(s, 0, j, k);
            wenojs_flux_split(s->t_s2[0] + o, s->t_s1[0] + o, s->t_ss2[0] + o, s->t_ss1[0] + o, s->nx, g, 1);
        }
#pragma omp parallel for collapse(2)
    for (int k = 0; k < s->nz; ++k)
        for (int i = 0; i < s->nx; ++i) {
            long o = ID(s, i, 0, k);
            wenojs_flux_split(s->t_s2[1] + o, s->t_s1[1] + o, s->t_ss2[1] + o, s->t_ss1[1] + o, s->ny, g, s->sx);
        }
    if (s->dim == 3) {
#pragma omp parallel for collapse(2)
        for (int j = 0; j < s->ny; ++j)
            for (int i = 0; i < s->nx; ++i) {
                long o = ID(s, i, j, 0);
                wenojs_flux_split(s->t_s2[2] + o, s->t_s1[2] + o, s->t_ss2[2] + o, s->t_ss1[2] + o, s->nz, g, s->sxy);
            }
    }
    const double dx = s->dx, dy = s->dy, dz = s->dz;
    double *a1 = s->t_ss1[0], *a2 = s->t_ss2[0], *b1 = s->t_ss1[1], *b2 = s->t_ss2[1];
    double *c1 = s->t_ss1[2], *c2 = s->t_ss2[2];
    const long sx = s->sx, sxy = s->sxy;
#pragma omp parallel for collapse(2)
    for (int k = 0; k < s->nz; ++k)
        for (int j = 0; j < s->ny; ++j)
            for (int i = 0; i < s->nx; ++i) {
                long id = ID(s, i, j, k);
                double v = -(a1[id] + a2[id] - a1[id - 1] - a2[id - 1]) / dx -
                           (b1[id] + b2[id] - b1[id - sx] - b2[id - sx]) / dy;
                if (s->dim == 3) v = v - (c1[id] + c2[id] - c1[id - sxy] - c2[id - sxy]) / dz;
                out[id] = v;
            }
}

static void ls_source_uccd(orc_state *s, double *out);

/* level_set_rk3_solver: SSP-RK3 in increment form (3D :1-65, 2D :1-59); the
 * source is the WENO5 flux split or, with ls_scheme = FCFD_LS_UCCD, the
 * non-conservative UCCD form */
int orc_level_set_rk3_solver(orc_state *s) {
    double *phi = s->phi.lv[L_NOW];
    const double dt = s->pr.dt;
    double *L[3] = {s->t_l1, s->t_l2, s->t_l3};
    for (int st = 0; st < 3; ++st) {
        if (s->pr.ls_scheme == FCFD_LS_UCCD) {
            ls_source_uccd(s, L[st]);
        } else {
            ls_source_setup(s);
            ls_source(s, L[st]);
        }
        double *l1 = s->t_l1, *l2 = s->t_l2, *l3 = s->t_l3;
#pragma omp parallel for collapse(2)
        for (int k = 0; k < s->nz; ++k)
            for (int j = 0; j < s->ny; ++j)
                for (int i = 0; i < s->nx; ++i) {
                    long id = ID(s, i, j, k);
                    double src;
                    if (st == 0) src = l1[id];
                    else if (st == 1) src = (-3.0 * l1[id] + l2[id]) / 4.0;
                    else src = (-l1[id] - l2[id] + 8.0 * l3[id]) / 12.0;
                    phi[id] = phi[id] + src * dt;
                }
        bc(s, phi);
    }
    return FCFD_OK;
}

/* ------------------------------------------------ level_set_rk3_redis */

/* wenojs_flux(f,fp,fm): one-sided WENO5 values for the Godunov gradient,
 * faces is-1..ie, EPS = 1e-10 (3D+IBM/.../level_set_rk3_redis.f90:295-339). */
static void wenojs_flux(const double *f, double *fp, double *fm, int n, long st) {
    const double EPS = 1.0e-10;
#define F(ii) f[(long)(ii) * st]
    for (int i = -1; i <= n - 1; ++i) {
        double t1 = F(i - 2) - 2.0 * F(i - 1) + F(i), t2 = F(i - 2) - 4.0 * F(i - 1) + 3.0 * F(i);
        double b1 = 13.0 * (t1 * t1) + 3.0 * (t2 * t2);
        t1 = F(i - 1) - 2.0 * F(i) + F(i + 1); t2 = F(i - 1) - F(i + 1);
        double b2 = 13.0 * (t1 * t1) + 3.0 * (t2 * t2);
        t1 = F(i) - 2.0 * F(i + 1) + F(i + 2); t2 = 3.0 * F(i) - 4.0 * F(i + 1) + F(i + 2);
        double b3 = 13.0 * (t1 * t1) + 3.0 * (t2 * t2);
        double a1 = 1.0 / ((EPS + b1) * (EPS + b1));
        double a2 = 6.0 / ((EPS + b2) * (EPS + b2));
        double a3 = 3.0 / ((EPS + b3) * (EPS + b3));
        double w1 = a1 / (a1 + a2 + a3), w2 = a2 / (a1 + a2 + a3), w3 = a3 / (a1 + a2 + a3);
        fm[(long)i * st] = w1 / 3.0 * F(i - 2) - (7.0 * w1 + w2) / 6.0 * F(i - 1) +
                           (11.0 * w1 + 5.0 * w2 + 2.0 * w3) / 6.0 * F(i) + (2.0 * w2 + 5.0 * w3) / 6.0 * F(i + 1) -
                           w3 / 6.0 * F(i + 2);
        t1 = F(i - 1) - 2.0 * F(i) + F(i + 1); t2 = F(i - 1) - 4.0 * F(i) + 3.0 * F(i + 1);
        b3 = 13.0 * (t1 * t1) + 3.0 * (t2 * t2);
        t1 = F(i) - 2.0 * F(i + 1) + F(i + 2); t2 = F(i) - F(i + 2);
        b2 = 13.0 * (t1 * t1) + 3.0 * (t2 * t2);
        t1 = F(i + 1) - 2.0 * F(i + 2) + F(i + 3); t2 = 3.0 * F(i + 1) - 4.0 * F(i + 2) + F(i + 3);
        b1 = 13.0 * (t1 * t1) + 3.0 * (t2 * t2);
        a1 = 1.0 / ((EPS + b1) * (EPS + b1));
        a2 = 6.0 / ((EPS + b2) * (EPS + b2));
        a3 = 3.0 / ((EPS + b3) * (EPS + b3));
        w1 = a1 / (a1 + a2 + a3); w2 = a2 / (a1 + a2 + a3); w3 = a3 / (a1 + a2 + a3);
        fp[(long)i * st] = w3 * (-F(i - 1) + 5.0 * F(i) + 2.0 * F(i + 1)) / 6.0 +
                           w2 * (2.0 * F(i) + 5.0 * F(i + 1) - F(i + 2)) / 6.0 +
                           w1 * (11.0 * F(i + 1) - 7.0 * F(i + 2) + 2.0 * F(i + 3)) / 6.0;
    }
#undef F
}

/* level_set_redis_gradient: D-phi, WENO one-sided, Godunov |grad phi|
 * (3D :212-293, 2D :195-251). */
static void redis_gradient(orc_state *s) {
    double *phi = s->phi.lv[L_NOW];
    double *nrm[3] = {s->nx_.lv[L_NOW], s->ny_.lv[L_NOW], s->nz_.lv[L_NOW]};
    const long sx = s->sx, sxy = s->sxy;
    const double dx = s->dx, dy = s->dy, dz = s->dz;
#pragma omp parallel for collapse(2)
    for (int k = 0; k < s->nz; ++k)
        for (int j = 0; j < s->ny; ++j)
            for (int i = 0; i < s->nx; ++i) {
                long id = ID(s, i, j, k);
                nrm[0][id] = (phi[id] - phi[id - 1]) / dx;
                nrm[1][id] = (phi[id] - phi[id - sx]) / dy;
                if (s->dim == 3) nrm[2][id] = (phi[id] - phi[id - sxy]) / dz;
            }
    for (int d = 0; d < s->dim; ++d) bc(s, nrm[d]);
#pragma omp parallel for collapse(2)
    for (int k = 0; k < s->nz; ++k)
        for (int j = 0; j < s->ny; ++j) {
            long o = ID(s, 0, j, k);
            wenojs_flux(nrm[0] + o, s->t_s1[0] + o, s->t_s2[0] + o, s->nx, 1);
        }
#pragma omp parallel for collapse(2)
    for (int k = 0; k < s->nz; ++k)
        for (int i = 0; i < s->nx; ++i) {
            long o = ID(s, i, 0, k);
            wenojs_flux(nrm[1] + o, s->t_s1[1] + o, s->t_s2[1] + o, s->ny, sx);
        }
    if (s->dim == 3) {
#pragma omp parallel for collapse(2)
        for (int j = 0; j < s->ny; ++j)
            for (int i = 0; i < s->nx; ++i) {
                long o = ID(s, i, j, 0);
                wenojs_flux(nrm[2] + o, s->t_s1[2] + o, s->t_s2[2] + o, s->nz, sxy);
            }
    }
    double *G = s->grad.lv[L_NOW], *S0 = s->sign.lv[L_TMP];
#pragma omp parallel for collapse(2)
    for (int k = 0; k < s->nz; ++k)
        for (int j = 0; j < s->ny; ++j)
            for (int i = 0; i < s->nx; ++i) {
                long id = ID(s, i, j, k);
                double upm = -fmin(s->t_s1[0][id], 0.0), upp = fmax(s->t_s1[0][id], 0.0);
                double umm = -fmin(s->t_s2[0][id], 0.0), ump = fmax(s->t_s2[0][id], 0.0);
                double vpm = -fmin(s->t_s1[1][id], 0.0), vpp = fmax(s->t_s1[1][id], 0.0);
                double vmm = -fmin(s->t_s2[1][id], 0.0), vmp = fmax(s->t_s2[1][id], 0.0);
                double a, b, c = 0.0;
                if (S0[id] >= 0.0) {
                    a = fmax(upm, ump); b = fmax(vpm, vmp);
                    if (s->dim == 3) {
                        double wpm = -fmin(s->t_s1[2][id], 0.0), wmp = fmax(s->t_s2[2][id], 0.0);
                        c = fmax(wpm, wmp);
                    }
                } else {
                    a = fmax(upp, umm); b = fmax(vpp, vmm);
                    if (s->dim == 3) {
                        double wpp = fmax(s->t_s1[2][id], 0.0), wmm = -fmin(s->t_s2[2][id], 0.0);
                        c = fmax(wpp, wmm);
                    }
                }
                G[id] = s->dim == 3 ? sqrt(a * a + b * b + c * c) : sqrt(a * a + b * b);
            }
}

/* level_set_redis_lambda(btn): mass-constrained correction, 3^d stencil with
 * centre weight 16 (2D) / 51 (3D); lambda = delta (2D) or density weighted
 * (3D) (3D :341-410, 2D :300-361). */
static void redis_lambda(orc_state *s, int btn) {
    double *s1 = s->t_s1[0], *s2 = s->t_s2[0], *s3 = s->t_s3[0];
    if (btn == 0) {
        FOR_INT(s) s1[ID(s, i, j, k)] = 0.0;
        return;
    }
    orc_ls_funs(s);
    double *H = s->heavy.lv[L_NOW], *D = s->delta.lv[L_NOW], *G = s->grad.lv[L_NOW], *S0 = s->sign.lv[L_TMP];
    const double r12 = s->pr.rho_12;
#pragma omp parallel for collapse(2)
    for (int k = 0; k < s->nz; ++k)
        for (int j = 0; j < s->ny; ++j)
            for (int i = 0; i < s->nx; ++i) {
                long id = ID(s, i, j, k);
                double lam = s->dim == 3 ? (2.0 * (1.0 - r12) * H[id] + r12) * D[id] : D[id];
                s2[id] = lam * S0[id] * (G[id] - 1.0);
                s3[id] = G[id] * D[id] * lam;
            }
    bc(s, s2);
    bc(s, s3);
    const long sx = s->sx, sxy = s->sxy;
    const double wc = s->dim == 3 ? 51.0 : 16.0;
#pragma omp parallel for collapse(2)
    for (int k = 0; k < s->nz; ++k)
        for (int j = 0; j < s->ny; ++j)
            for (int i = 0; i < s->nx; ++i) {
                long id = ID(s, i, j, k);
                double a = wc * s2[id], b = wc * s3[id];
                int kl = s->dim == 3 ? -1 : 0, kh = s->dim == 3 ? 1 : 0;
                for (int kk = kl; kk <= kh; ++kk)
                    for (int jj = -1; jj <= 1; ++jj)
                        for (int ii = -1; ii <= 1; ++ii) {
                            long o = id + ii + jj * sx + kk * sxy;
                            a = a + s2[o];
                            b = b + s3[o];
                        }
                s1[id] = 0.0;
                if (fabs(b) > 1.0e-12) s1[id] = a / b * D[id];
            }
}

/* level_set_rk3_redis_solver(btn) (3D :138-210, 2D :127-193) */
static void redis_solver(orc_state *s, int btn) {
    double *phi = s->phi.lv[L_NOW], *G = s->grad.lv[L_NOW], *S0 = s->sign.lv[L_TMP], *s1 = s->t_s1[0];
    const double rdt = s->pr.rdt;
    for (int st = 0; st < 3; ++st) {
        redis_gradient(s);
        redis_lambda(s, btn);
#pragma omp parallel for collapse(2)
        for (int k = 0; k < s->nz; ++k)
            for (int j = 0; j < s->ny; ++j)
                for (int i = 0; i < s->nx; ++i) {
                    long id = ID(s, i, j, k);
                    double l = (S0[id] - s1[id]) * G[id] - S0[id], src;
                    if (st == 0) { s->t_l1[id] = l; src = l; }
                    else if (st == 1) { s->t_l2[id] = l; src = (-3.0 * s->t_l1[id] + l) / 4.0; }
                    else { s->t_l3[id] = l; src = (-s->t_l1[id] - s->t_l2[id] + 8.0 * l) / 12.0; }
                    phi[id] = phi[id] - rdt * src;
                }
        bc(s, phi);
    }
}

/* level_set_rk3_redis(btn) + level_set_redis_init/stable
 * (3D :1-136, 2D :1-125).  Returns the pseudo-step count. */
int orc_level_set_rk3_redis(orc_state *s, int btn) {
    double *phi = s->phi.lv[L_NOW], *ptmp = s->phi.lv[L_TMP];
    double *S0 = s->sign.lv[L_TMP];
    orc_ls_funs(s);
    FOR_INT(s) S0[ID(s, i, j, k)] = s->sign.lv[L_NOW][ID(s, i, j, k)];
    if (btn == 0) {
        FOR_INT(s) S0[ID(s, i, j, k)] = phi[ID(s, i, j, k)];
        redis_gradient(s);
        double grad = 0.0;
        FOR_INT(s) grad = fmax(grad, s->grad.lv[L_NOW][ID(s, i, j, k)]);
        FOR_INT(s) phi[ID(s, i, j, k)] = phi[ID(s, i, j, k)] / grad;
        bc(s, phi);
    }
    double timestop;
    if (btn == 0) {
        double lx = s->gd.xend - s->gd.xstart, ly = s->gd.yend - s->gd.ystart;
        double m = lx > ly ? lx : ly;
        if (s->dim == 3) { double lz = s->gd.zend - s->gd.zstart; m = m > lz ? m : lz; }
        timestop = 1.5 * m;
    } else {
        timestop = 3.0 * s->dx;
    }
    const double tol = s->dim == 3 ? 1.0e-5 : 1.0e-8;
    int iter = 0;
    double time = 0.0;
    for (;;) {
        iter = iter + 1;
        time = time + s->pr.rdt;
        FOR_INT(s) ptmp[ID(s, i, j, k)] = phi[ID(s, i, j, k)];
        redis_solver(s, btn);
        double error = 0.0;
        FOR_INT(s) error = fmax(error, fabs(ptmp[ID(s, i, j, k)] - phi[ID(s, i, j, k)]));
        if (time > timestop || error <= tol) break;
    }
    s->redis_iters = iter;
    return iter;
}

/* ----------------------------------------------------------- CCD lines */

/* 2x2 block-tridiagonal line system of the CCD family: rows (1,2,3) =
 * (sub, diag, super) per point, as working_matrix's A, B, AA, BB
 * (3D+IBM/lib/solver/ccd_solvers.f90:5-11). */
typedef struct { double *A, *B, *AA, *BB; } twin_mat;
#define M(X, r, i) X[(i) * 3 + (r) - 1]

static twin_mat twin_alloc(int N) {
    twin_mat m;
    m.A = (double *)calloc(3 * (size_t)N, sizeof(double)); m.B = (double *)calloc(3 * (size_t)N, sizeof(double));
    m.AA = (double *)calloc(3 * (size_t)N, sizeof(double)); m.BB = (double *)calloc(3 * (size_t)N, sizeof(double));
    return m;
}
static void twin_free(twin_mat *m) { free(m->A); free(m->B); free(m->AA); free(m->BB); }

/* working_matrix_bc: one-sided closure rows of both ends (ccd_solvers.f90:76-130) */
static void twin_bc(twin_mat *m, int N, double dx) {
    const int is = 0, ie = N - 1;
    M(m->A, 2, is) = 1.0; M(m->B, 2, is) = 0.0; M(m->AA, 2, is) = 0.0; M(m->BB, 2, is) = 1.0;
    M(m->A, 3, is) = 2.0; M(m->B, 3, is) = -dx; M(m->AA, 3, is) = -2.5 / dx; M(m->BB, 3, is) = 8.5;
    M(m->A, 2, ie) = 1.0; M(m->B, 2, ie) = 0.0; M(m->AA, 2, ie) = 0.0; M(m->BB, 2, ie) = 1.0;
    M(m->A, 1, ie) = 2.0; M(m->B, 1, ie) = dx; M(m->AA, 1, ie) = 2.5 / dx; M(m->BB, 1, ie) = 8.5;
}

/* TWIN_DEC (3D+IBM/lib/solver/matrix_solver.f90:6-48) */
static void twin_dec(twin_mat *m, int N) {
    double *A = m->A, *B = m->B, *AA = m->AA, *BB = m->BB;
    const int is = 0, ie = N - 1;
    M(A, 1, is) = 0.0; M(AA, 1, is) = 0.0; M(A, 3, ie) = 0.0; M(AA, 3, ie) = 0.0;
    M(B, 1, is) = 0.0; M(BB, 1, is) = 0.0; M(B, 3, ie) = 0.0; M(BB, 3, ie) = 0.0;
    for (int i = is + 1; i <= ie; ++i) {
        double Den = M(A, 2, i - 1) * M(BB, 2, i - 1) - M(AA, 2, i - 1) * M(B, 2, i - 1);
        double SC1 = -(M(AA, 2, i - 1) * M(B, 1, i) - M(A, 1, i) * M(BB, 2, i - 1)) / Den;
        double SC2 = (M(B, 1, i) * M(A, 2, i - 1) - M(B, 2, i - 1) * M(A, 1, i)) / Den;
        M(A, 1, i) = SC1;
        M(A, 2, i) = M(A, 2, i) - (SC1 * M(A, 3, i - 1) + SC2 * M(AA, 3, i - 1));
        M(B, 1, i) = SC2;
        M(B, 2, i) = M(B, 2, i) - (SC1 * M(B, 3, i - 1) + SC2 * M(BB, 3, i - 1));
        SC1 = -(M(AA, 2, i - 1) * M(BB, 1, i) - M(AA, 1, i) * M(BB, 2, i - 1)) / Den;
        SC2 = (M(BB, 1, i) * M(A, 2, i - 1) - M(B, 2, i - 1) * M(AA, 1, i)) / Den;
        M(AA, 1, i) = SC1;
        M(AA, 2, i) = M(AA, 2, i) - (SC1 * M(A, 3, i - 1) + SC2 * M(AA, 3, i - 1));
        M(BB, 1, i) = SC2;
        M(BB, 2, i) = M(BB, 2, i) - (SC1 * M(B, 3, i - 1) + SC2 * M(BB, 3, i - 1));
    }
}

/* TWIN_BKS (matrix_solver.f90:50-78): S, SS rhs in, solution out */
static void twin_bks(const twin_mat *m, double *S, double *SS, int N) {
    const double *A = m->A, *B = m->B, *AA = m->AA, *BB = m->BB;
    const int is = 0, ie = N - 1;
    for (int i = is + 1; i <= ie; ++i) {
        S[i] = S[i] - (M(A, 1, i) * S[i - 1] + M(B, 1, i) * SS[i - 1]);
        SS[i] = SS[i] - (M(AA, 1, i) * S[i - 1] + M(BB, 1, i) * SS[i - 1]);
    }
    double Den = M(A, 2, ie) * M(BB, 2, ie) - M(AA, 2, ie) * M(B, 2, ie);
    double SolS = -(M(B, 2, ie) * SS[ie] - M(BB, 2, ie) * S[ie]) / Den;
    double SolSS = (M(A, 2, ie) * SS[ie] - M(AA, 2, ie) * S[ie]) / Den;
    S[ie] = SolS; SS[ie] = SolSS;
    for (int i = ie - 1; i >= is; --i) {
        S[i] = S[i] - (M(A, 3, i) * SolS + M(B, 3, i) * SolSS);
        SS[i] = SS[i] - (M(AA, 3, i) * SolS + M(BB, 3, i) * SolSS);
        Den = M(A, 2, i) * M(BB, 2, i) - M(AA, 2, i) * M(B, 2, i);
        SolS = -(M(B, 2, i) * SS[i] - M(BB, 2, i) * S[i]) / Den;
        SolSS = (M(A, 2, i) * SS[i] - M(AA, 2, i) * S[i]) / Den;
        S[i] = SolS; SS[i] = SolSS;
    }
}

/* one-sided closure sources of both ends (asgn_src_ccd / _uccd,
 * ccd_solvers.f90:287-306,308-341) */
#define FF(i) f[(long)(i) * st]
static void src_ends(const double *f, long st, int N, double dx, double *S, double *SS) {
    const int is = 0, ie = N - 1;
    const double dx2 = dx * dx;
    S[is] = (-3.5 * FF(is) + 4.0 * FF(is + 1) - 0.5 * FF(is + 2)) / dx;
    SS[is] = (34.0 / 3.0 * FF(is) - 83.0 / 4.0 * FF(is + 1) + 10.0 * FF(is + 2) - 7.0 / 12.0 * FF(is + 3)) / dx2;
    S[ie] = -(-3.5 * FF(ie) + 4.0 * FF(ie - 1) - 0.5 * FF(ie - 2)) / dx;
    SS[ie] = (34.0 / 3.0 * FF(ie) - 83.0 / 4.0 * FF(ie - 1) + 10.0 * FF(ie - 2) - 7.0 / 12.0 * FF(ie - 3)) / dx2;
}

/* working_matrix_init_ccd + bc + TWIN_DEC, then asgn_src_ccd + TWIN_BKS
 * (3D+IBM/lib/solver/ccd_solvers.f90:76-159,287-306,379-399;
 *  3D+IBM/lib/solver/matrix_solver.f90:6-78).  Line of N points, stride st. */
static void ccd_line(const double *f, double *fx, double *fxx, int N, long st, double dx) {
    twin_mat m = twin_alloc(N);
    double *S = (double *)malloc(sizeof(double) * N), *SS = (double *)malloc(sizeof(double) * N);
    const int is = 0, ie = N - 1;
    for (int i = is + 1; i <= ie - 1; ++i) {
        M(m.A, 1, i) = 7.0 / 16.0; M(m.A, 2, i) = 1.0; M(m.A, 3, i) = 7.0 / 16.0;
        M(m.B, 1, i) = dx / 16.0; M(m.B, 2, i) = 0.0; M(m.B, 3, i) = -dx / 16.0;
        M(m.AA, 1, i) = -9.0 / 8.0 / dx; M(m.AA, 2, i) = 0.0; M(m.AA, 3, i) = 9.0 / 8.0 / dx;
        M(m.BB, 1, i) = -1.0 / 8.0; M(m.BB, 2, i) = 1.0; M(m.BB, 3, i) = -1.0 / 8.0;
    }
    twin_bc(&m, N, dx);
    twin_dec(&m, N);
    const double dx2 = dx * dx;
    src_ends(f, st, N, dx, S, SS);
    for (int i = is + 1; i <= ie - 1; ++i) {
        S[i] = 15.0 / 16.0 * (-FF(i - 1) + FF(i + 1)) / dx;
        SS[i] = (3.0 * FF(i - 1) - 6.0 * FF(i) + 3.0 * FF(i + 1)) / dx2;
    }
    twin_bks(&m, S, SS, N);
    for (int i = is; i <= ie; ++i) { fx[(long)i * st] = S[i]; fxx[(long)i * st] = SS[i]; }
    twin_free(&m);
    free(S); free(SS);
}

/* root_solve("uccd"): working_matrix_init_uccd + bc, TWIN_DEC of the up- and
 * down-wind systems, asgn_src_uccd, TWIN_BKS of both, then fx = SU where
 * u > 0 else SD, fxx = (SSU + SSD)/2 (3D+IBM/lib/solver/ccd_solvers.f90:
 * 76-130,161-210,308-341,400-412).  Line of N points, stride st. */
static void uccd_line(const double *f, const double *u, double *fx, double *fxx, int N, long st, double dx) {
    twin_mat U = twin_alloc(N), D = twin_alloc(N);
    double *SU = (double *)malloc(sizeof(double) * N), *SSU = (double *)malloc(sizeof(double) * N);
    double *SD = (double *)malloc(sizeof(double) * N), *SSD = (double *)malloc(sizeof(double) * N);
    const int is = 0, ie = N - 1;
    const double A1 = 0.875, B1 = 0.1251282341599089, B2 = -0.2487176584009104, B3 = 0.0001282341599089;
    twin_bc(&U, N, dx);
    twin_bc(&D, N, dx);
    for (int i = is + 1; i <= ie - 1; ++i) {
        M(U.A, 1, i) = A1; M(U.A, 2, i) = 1.0; M(U.A, 3, i) = 0.0;
        M(U.B, 1, i) = B1 * dx; M(U.B, 2, i) = B2 * dx; M(U.B, 3, i) = B3 * dx;
        M(U.AA, 1, i) = -9.0 / 8.0 / dx; M(U.AA, 2, i) = 0.0; M(U.AA, 3, i) = 9.0 / 8.0 / dx;
        M(U.BB, 1, i) = -1.0 / 8.0; M(U.BB, 2, i) = 1.0; M(U.BB, 3, i) = -1.0 / 8.0;
        M(D.A, 1, i) = 0.0; M(D.A, 2, i) = 1.0; M(D.A, 3, i) = A1;
        M(D.B, 1, i) = -B3 * dx; M(D.B, 2, i) = -B2 * dx; M(D.B, 3, i) = -B1 * dx;
        M(D.AA, 1, i) = -9.0 / 8.0 / dx; M(D.AA, 2, i) = 0.0; M(D.AA, 3, i) = 9.0 / 8.0 / dx;
        M(D.BB, 1, i) = -1.0 / 8.0; M(D.BB, 2, i) = 1.0; M(D.BB, 3, i) = -1.0 / 8.0;
    }
    twin_dec(&U, N);
    twin_dec(&D, N);
    const double dx2 = dx * dx;
    src_ends(f, st, N, dx, SU, SSU);
    src_ends(f, st, N, dx, SD, SSD);
    const double c3 = -0.06096119008109, c2 = 1.99692238016218, c1 = -1.93596119008109;
    for (int i = is + 1; i <= ie - 1; ++i) {
        SU[i] = (c1 * FF(i - 1) + c2 * FF(i) + c3 * FF(i + 1)) / dx;
        SSU[i] = (3.0 * FF(i - 1) - 6.0 * FF(i) + 3.0 * FF(i + 1)) / dx2;
        SD[i] = -(c3 * FF(i - 1) + c2 * FF(i) + c1 * FF(i + 1)) / dx;
        SSD[i] = (3.0 * FF(i - 1) - 6.0 * FF(i) + 3.0 * FF(i + 1)) / dx2;
    }
    twin_bks(&U, SU, SSU, N);
    twin_bks(&D, SD, SSD, N);
    for (int i = is; i <= ie; ++i) {
        fx[(long)i * st] = u[(long)i * st] > 0.0 ? SU[i] : SD[i];
        fxx[(long)i * st] = 0.5 * (SSU[i] + SSD[i]);
    }
    twin_free(&U); twin_free(&D);
    free(SU); free(SSU); free(SD); free(SSD);
}
#undef FF
#undef M

/* ccd along direction d of every line of the full ghosted box */
static void ccd_dir(orc_state *s, int d, const double *f, double *fx, double *fxx) {
    const int g = s->g, gz = s->gz;
    if (d == 0) {
#pragma omp parallel for collapse(2)
        for (int k = -gz; k < s->nz + gz; ++k)
            for (int j = -g; j < s->ny + g; ++j) {
                long o = ID(s, -g, j, k);
                ccd_line(f + o, fx + o, fxx + o, s->nx + 2 * g, 1, s->dx);
            }
    } else if (d == 1) {
#pragma omp parallel for collapse(2)
        for (int k = -gz; k < s->nz + gz; ++k)
            for (int i = -g; i < s->nx + g; ++i) {
                long o = ID(s, i, -g, k);
                ccd_line(f + o, fx + o, fxx + o, s->ny + 2 * g, s->sx, s->dy);
            }
    } else {
#pragma omp parallel for collapse(2)
        for (int j = -g; j < s->ny + g; ++j)
            for (int i = -g; i < s->nx + g; ++i) {
                long o = ID(s, i, j, -g);
                ccd_line(f + o, fx + o, fxx + o, s->nz + 2 * g, s->sxy, s->dz);
            }
    }
}

/* root_solve("uccd") along direction d on every interior line (the lines
 * the reference's UCCD level-set source walks, with their ghost ends:
 * DomainDecomposition/3D/src/interface_solver/level_set_rk3_solver.f90:206-226;
 * the commented alternative of 3D+IBM/.../level_set_rk3_solver.f90:109,123,137) */
static void uccd_dir(orc_state *s, int d, const double *f, const double *u, double *fx, double *fxx) {
    const int g = s->g, gz = s->gz;
    (void)gz;
    if (d == 0) {
#pragma omp parallel for collapse(2)
        for (int k = 0; k < s->nz; ++k)
            for (int j = 0; j < s->ny; ++j) {
                long o = ID(s, -g, j, k);
                uccd_line(f + o, u + o, fx + o, fxx + o, s->nx + 2 * g, 1, s->dx);
            }
    } else if (d == 1) {
#pragma omp parallel for collapse(2)
        for (int k = 0; k < s->nz; ++k)
            for (int i = 0; i < s->nx; ++i) {
                long o = ID(s, i, -g, k);
                uccd_line(f + o, u + o, fx + o, fxx + o, s->ny + 2 * g, s->sx, s->dy);
            }
    } else {
#pragma omp parallel for collapse(2)
        for (int j = 0; j < s->ny; ++j)
            for (int i = 0; i < s->nx; ++i) {
                long o = ID(s, i, j, -g);
                uccd_line(f + o, u + o, fx + o, fxx + o, s->nz + 2 * g, s->sxy, s->dz);
            }
    }
}

/* UCCD derivatives of phi with the cell velocity nvel%old as upwind
 * selector: fx -> normals x/y/z, fxx -> normals xx/yy/zz (test entry) */
int orc_uccd(orc_state *s) {
    double *phi = s->phi.lv[L_NOW];
    double *vel[3] = {s->nu.lv[L_OLD], s->nv.lv[L_OLD], s->nw.lv[L_OLD]};
    double *fx[3] = {s->nx_.lv[L_NOW], s->ny_.lv[L_NOW], s->nz_.lv[L_NOW]};
    double *fxx[3] = {s->nxx.lv[L_NOW], s->nyy.lv[L_NOW], s->nzz.lv[L_NOW]};
    for (int d = 0; d < s->dim; ++d) uccd_dir(s, d, phi, vel[d], fx[d], fxx[d]);
    return FCFD_OK;
}

/* UCCD level-set source s = -(u phi_x + v phi_y + w phi_z), phi_x from
 * upwind CCD (DomainDecomposition/3D/.../level_set_rk3_solver.f90:228-236;
 * the commented form of 3D+IBM/.../level_set_rk3_solver.f90:155-157) */
static void ls_source_uccd(orc_state *s, double *out) {
    double *phi = s->phi.lv[L_NOW];
    double *vel[3] = {s->nu.lv[L_OLD], s->nv.lv[L_OLD], s->nw.lv[L_OLD]};
    for (int d = 0; d < s->dim; ++d) uccd_dir(s, d, phi, vel[d], s->t_s1[d], s->t_s2[d]);
    double *ax = s->t_s1[0], *ay = s->t_s1[1], *az = s->t_s1[2];
    double *u = vel[0], *v = vel[1], *w = vel[2];
#pragma omp parallel for collapse(2)
    for (int k = 0; k < s->nz; ++k)
        for (int j = 0; j < s->ny; ++j)
            for (int i = 0; i < s->nx; ++i) {
                long id = ID(s, i, j, k);
                double val = -u[id] * ax[id] - v[id] * ay[id];
                if (s->dim == 3) val = val - w[id] * az[id];
                out[id] = val;
            }
}

/* the level-set source of the configured scheme for the current phi and
 * nvel%old, written to the grad field (test entry: no RK stage, no bc) */
int orc_level_set_source(orc_state *s) {
    double *out = s->grad.lv[L_NOW];
    if (s->pr.ls_scheme == FCFD_LS_UCCD) {
        ls_source_uccd(s, out);
    } else {
        ls_source_setup(s);
        ls_source(s, out);
    }
    return FCFD_OK;
}

/* surface_norms (3D+IBM/.../others_normals.f90:1-83; 2D :1-40) */
int orc_surface_norms(orc_state *s) {
    double *phi = s->phi.lv[L_NOW], *cv = s->curv.lv[L_NOW];
    ccd_dir(s, 0, phi, s->nx_.lv[L_NOW], s->nxx.lv[L_NOW]);
    ccd_dir(s, 1, phi, s->ny_.lv[L_NOW], s->nyy.lv[L_NOW]);
    if (s->dim == 3) ccd_dir(s, 2, phi, s->nz_.lv[L_NOW], s->nzz.lv[L_NOW]);
    ccd_dir(s, 1, s->nx_.lv[L_NOW], s->nxy.lv[L_NOW], cv);
    if (s->dim == 3) {
        ccd_dir(s, 2, s->nx_.lv[L_NOW], s->nxz.lv[L_NOW], cv);
        ccd_dir(s, 2, s->ny_.lv[L_NOW], s->nyz.lv[L_NOW], cv);
    }
    bc(s, s->nx_.lv[L_NOW]); bc(s, s->nxx.lv[L_NOW]);
    bc(s, s->ny_.lv[L_NOW]); bc(s, s->nyy.lv[L_NOW]);
    if (s->dim == 3) { bc(s, s->nz_.lv[L_NOW]); bc(s, s->nzz.lv[L_NOW]); }
    bc(s, s->nxy.lv[L_NOW]);
    if (s->dim == 3) { bc(s, s->nxz.lv[L_NOW]); bc(s, s->nyz.lv[L_NOW]); }
    return FCFD_OK;
}

/* curv (3D others_normals.f90:133-167; 2D :76-104) */
int orc_curv(orc_state *s) {
    orc_surface_norms(s);
    double *c = s->curv.lv[L_NOW];
    double *X = s->nx_.lv[L_NOW], *Y = s->ny_.lv[L_NOW], *Z = s->nz_.lv[L_NOW];
    double *XX = s->nxx.lv[L_NOW], *YY = s->nyy.lv[L_NOW], *ZZ = s->nzz.lv[L_NOW];
    double *XY = s->nxy.lv[L_NOW], *XZ = s->nxz.lv[L_NOW], *YZ = s->nyz.lv[L_NOW];
#pragma omp parallel for collapse(2)
    for (int k = 0; k < s->nz; ++k)
        for (int j = 0; j < s->ny; ++j)
            for (int i = 0; i < s->nx; ++i) {
                long id = ID(s, i, j, k);
                double fx = X[id], fy = Y[id], fxx = XX[id], fyy = YY[id], fxy = XY[id];
                if (s->dim == 3) {
                    double fz = Z[id], fzz = ZZ[id], fxz = XZ[id], fyz = YZ[id];
                    c[id] = ((fyy + fzz) * (fx * fx) + (fxx + fzz) * (fy * fy) + (fxx + fyy) * (fz * fz) -
                             2.0 * (fxy * fx * fy + fxz * fx * fz + fyz * fy * fz)) /
                            pow(fx * fx + fy * fy + fz * fz + 1.0e-12, 1.5);
                } else {
                    c[id] = (fyy * (fx * fx) + fxx * (fy * fy) - 2.0 * fxy * fx * fy) / pow(fx * fx + fy * fy + 1.0e-12, 1.5);
                }
            }
    bc(s, c);
    return FCFD_OK;
}

/* ------------------------------------------ mass_preserving_level_set */

/* mass_preserving_level_set: 2 Newton steps on the mass
 * (3D+IBM/.../mass_preserving_level_set.f90:1-54; 2D :1-47). */
int orc_mass_preserving_level_set(orc_state *s) {
    double *phi = s->phi.lv[L_NOW], *G = s->grad.lv[L_NOW], *H = s->heavy.lv[L_NOW], *D = s->delta.lv[L_NOW];
    double *X = s->nx_.lv[L_NOW], *Y = s->ny_.lv[L_NOW], *Z = s->nz_.lv[L_NOW];
    const double r12 = s->pr.rho_12;
    for (int it = 1; it <= 2; ++it) {
        orc_ls_mv(s);
        orc_surface_norms(s);
        double lam = 0.0;
        PSUM_BEGIN(s, lam) {
                    long id = ID(s, i, j, k);
                    double plam;
                    if (s->dim == 3) {
                        G[id] = sqrt(X[id] * X[id] + Y[id] * Y[id] + Z[id] * Z[id]);
                        plam = D[id] * D[id] * G[id];
                        plam = plam * (2.0 * (1.0 - r12) * H[id] + r12) * s->dx * s->dy * s->dz;
                    } else {
                        G[id] = sqrt(X[id] * X[id] + Y[id] * Y[id]);
                        plam = D[id] * D[id] * G[id];
                        plam = plam * (2.0 * (1.0 - r12) * H[id] + r12) * s->dx * s->dy;
                    }
                    lam = lam + plam;
        } PSUM_END(lam)
        lam = (s->imass - s->mass) / lam;
        s->mpls_lam = lam;
        FOR_INT(s) {
            long id = ID(s, i, j, k);
            phi[id] = phi[id] + lam * D[id] * G[id];
        }
        bc(s, phi);
    }
    return FCFD_OK;
}

/* ---------------------------------------------------------- ns solver */

/* ppe_sor_init: 1/rho face coefficients with Neumann folding
 * (3D+IBM/src/ns_solver/ns_ppe_solver.f90:1-65; 2D :1-48). */
int orc_ppe_sor_init(orc_state *s) {
    double *R = s->rho.lv[L_NOW];
    const long sx = s->sx, sxy = s->sxy;
    const double dx2 = s->dx * s->dx, dy2 = s->dy * s->dy, dz2 = s->dz * s->dz;
#pragma omp parallel for collapse(2)
    for (int k = 0; k < s->nz; ++k)
        for (int j = 0; j < s->ny; ++j)
            for (int i = 0; i < s->nx; ++i) {
                long id = ID(s, i, j, k);
                double r = 2.0 / (R[id] + R[id + 1]), l = 2.0 / (R[id] + R[id - 1]);
                double f = 2.0 / (R[id] + R[id + sx]), b = 2.0 / (R[id] + R[id - sx]);
                double u = 0.0, d = 0.0;
                r = r / dx2; l = l / dx2; f = f / dy2; b = b / dy2;
                double c;
                if (s->dim == 3) {
                    u = 2.0 / (R[id] + R[id + sxy]); d = 2.0 / (R[id] + R[id - sxy]);
                    u = u / dz2; d = d / dz2;
                    c = -(r + l + f + b + u + d);
                } else {
                    c = -(r + l + f + b);
                }
                if (i == 0) { c = c + l; l = 0.0; }
                if (i == s->nx - 1) { c = c + r; r = 0.0; }
                if (j == 0) { c = c + b; b = 0.0; }
                if (j == s->ny - 1) { c = c + f; f = 0.0; }
                if (s->dim == 3) {
                    if (k == 0) { c = c + d; d = 0.0; }
                    if (k == s->nz - 1) { c = c + u; u = 0.0; }
                }
                s->coe_r[id] = r; s->coe_l[id] = l; s->coe_f[id] = f; s->coe_b[id] = b;
                s->coe_u[id] = u; s->coe_d[id] = d; s->coe_c[id] = c;
            }
    s->band_valid = 0;
    return FCFD_OK;
}

/* src = div(u*)/dt (ns_ppe_solver.f90:275-287 in 3D, :60-69 in 2D) */
static void ppe_src(orc_state *s) {
    double *U = s->u.lv[L_NOW], *V = s->v.lv[L_NOW], *W = s->w.lv[L_NOW];
    const long sx = s->sx, sxy = s->sxy;
#pragma omp parallel for collapse(2)
    for (int k = 0; k < s->nz; ++k)
        for (int j = 0; j < s->ny; ++j)
            for (int i = 0; i < s->nx; ++i) {
                long id = ID(s, i, j, k);
                if (s->dim == 3)
                    s->coe_src[id] = ((U[id] - U[id - 1]) / s->dx + (V[id] - V[id - sx]) / s->dy +
                                      (W[id] - W[id - sxy]) / s->dz) / s->pr.dt;
                else
                    s->coe_src[id] = ((U[id] - U[id - 1]) / s->dx + (V[id] - V[id - sx]) / s->dy) / s->pr.dt;
            }
}

/* ns_momentum_correction (3D ns_ppe_solver.f90:174-205; 2D :146-171) */
int orc_ns_momentum_correction(orc_state *s) {
    double *U = s->u.lv[L_NOW], *V = s->v.lv[L_NOW], *W = s->w.lv[L_NOW], *P = s->p.lv[L_NOW], *R = s->rho.lv[L_NOW];
    const long sx = s->sx, sxy = s->sxy;
    const double dt = s->pr.dt;
#pragma omp parallel for collapse(2)
    for (int k = 0; k < s->nz; ++k)
        for (int j = 0; j < s->ny; ++j)
            for (int i = 0; i < s->nx; ++i) {
                long id = ID(s, i, j, k);
                double rho = (R[id] + R[id + 1]) / 2.0, px = (P[id + 1] - P[id]) / s->dx;
                U[id] = U[id] - dt * px / rho;
                rho = (R[id] + R[id + sx]) / 2.0;
                double py = (P[id + sx] - P[id]) / s->dy;
                V[id] = V[id] - dt * py / rho;
                if (s->dim == 3) {
                    rho = (R[id] + R[id + sxy]) / 2.0;
                    double pz = (P[id + sxy] - P[id]) / s->dz;
                    W[id] = W[id] - dt * pz / rho;
                }
            }
    orc_velbc(s);
    return FCFD_OK;
}

/* ppe_sor_solver(tol): the reference's in-place relaxation, run in the
 * sequential (single-thread) order; ends with ns_momentum_correction
 * (3D ns_ppe_solver.f90:67-172; 2D :50-144). */
int orc_ppe_sor_solver(orc_state *s, double tol) {
    ppe_src(s);
    double *P = s->p.lv[L_NOW], *T = s->p.lv[L_TMP];
    const long sx = s->sx, sxy = s->sxy;
    const double w = s->pr.p_w1;
    const double ncell = (double)s->nx * (double)s->ny * (double)s->nz;
    double err = 0.0;
    int rc = FCFD_OK;
    for (;;) {
        s->piter = s->piter + 1;
        FOR_INT(s) {
            long id = ID(s, i, j, k);
            T[id] = P[id];
            double pv = s->coe_src[id] - s->coe_r[id] * P[id + 1] - s->coe_l[id] * P[id - 1] - s->coe_f[id] * P[id + sx] -
                        s->coe_b[id] * P[id - sx];
            if (s->dim == 3) pv = pv - s->coe_u[id] * P[id + sxy] - s->coe_d[id] * P[id - sxy];
            P[id] = pv / s->coe_c[id];
        }
        double sump = 0.0;
        FOR_INT(s) sump = sump + P[ID(s, i, j, k)];
        sump = sump / ncell;
        err = 0.0;
        FOR_INT(s) {
            long id = ID(s, i, j, k);
            P[id] = P[id] - sump;
            P[id] = w * P[id] + (1.0 - w) * T[id];
            err = fmax(err, fabs(P[id] - T[id]));
        }
        bc(s, P);
        if (err < tol) break;
        if (err > 10.0 && s->piter > 100000) { rc = FCFD_ERR_NOT_CONVERGED; break; }
        if (s->piter > 5000000) { rc = FCFD_ERR_NOT_CONVERGED; break; }
    }
    s->ppe_err = err;
    orc_ns_momentum_correction(s);
    return rc;
}

/* A p with the ppe_sor_init coefficients (interior), out = c p + sum c_nb p_nb */
int orc_ppe_apply(orc_state *s, const double *P, double *out) {
    const long sx = s->sx, sxy = s->sxy;
#pragma omp parallel for collapse(2)
    for (int k = 0; k < s->nz; ++k)
        for (int j = 0; j < s->ny; ++j)
            for (int i = 0; i < s->nx; ++i) {
                long id = ID(s, i, j, k);
                double v = s->coe_c[id] * P[id] + s->coe_r[id] * P[id + 1] + s->coe_l[id] * P[id - 1] +
                           s->coe_f[id] * P[id + sx] + s->coe_b[id] * P[id - sx];
                if (s->dim == 3) v += s->coe_u[id] * P[id + sxy] + s->coe_d[id] * P[id - sxy];
                out[id] = v;
            }
    return FCFD_OK;
}

/* ------------------------------------------------ converged PPE solve
 * Not in the reference: the fixed point its relaxation converges to, solved
 * directly.  The ppe_sor_init operator A (c p + sum c_nb p_nb over the
 * interior, zero-flux walls) is symmetric with the constants as null space,
 * and the relaxation's iterate satisfies A p = src - mean(src), mean(p) = 0.
 *
 * M = -A with the first cell pinned (its row and column replaced by the
 * identity) is symmetric positive definite and banded: cells are numbered
 * i + nx (j + ny k), so the couplings of a row lie within b = nx (2D) or
 * nx ny (3D) of the diagonal.  ppe_band_factor forms its Cholesky factor in
 * band storage (row r keeps columns r-b..r at r(b+1) + (c - r + b)),
 * ppe_band_solve the two triangular sweeps.  Backward stable, so the
 * solution is exact to the operator's conditioning; two steps of iterative
 * refinement with the reference-form residual remove the rest.  Dropping
 * the pinned cell's equation loses nothing: the rhs sums to zero and the
 * columns of A do, so that equation follows from the others. */

static long ppe_cell(const orc_state *s, int i, int j, int k) {
    return (long)i + (long)s->nx * ((long)j + (long)s->ny * (long)k);
}

/* ---- band Cholesky (shared with the MPLS_RisingBubble oracle, band.h) ---- */
int orc_band_alloc(orc_band *B, long n, long b) {
    const long w = b + 1;
    if (!B->L || B->n != n || B->b != b) {
        free(B->L);
        B->L = (double *)malloc(sizeof(double) * (size_t)(n * w));
        if (!B->L) return FCFD_ERR_ARG;
        B->n = n;
        B->b = b;
    }
    memset(B->L, 0, sizeof(double) * (size_t)(n * w));
    return FCFD_OK;
}

/* right-looking band Cholesky in place: column j scales, then updates rows
 * j+1..j+b */
int orc_band_cholesky(orc_band *B) {
    const long n = B->n, b = B->b, w = b + 1;
    double *L = B->L;
    double *col = (double *)malloc(sizeof(double) * (size_t)w);
    for (long j = 0; j < n; ++j) {
        double *rj = L + j * w + b;
        if (!(rj[0] > 0.0)) { free(col); return FCFD_ERR_SINGULAR; }
        const double d = sqrt(rj[0]);
        rj[0] = d;
        const long iend = j + b < n - 1 ? j + b : n - 1;
        const long m = iend - j;
        for (long t = 1; t <= m; ++t) {
            double *e = L + (j + t) * w + b - t;   /* M[j+t][j] */
            *e /= d;
            col[t] = *e;
        }
#pragma omp parallel for schedule(static) if (m >= 256)
        for (long t = 1; t <= m; ++t) {
            const long i = j + t;
            const double lij = col[t];
            if (lij == 0.0) continue;
            double *ri = L + i * w + b - i;   /* ri[c] = M[i][c] */
            for (long q = 1; q <= t; ++q) ri[j + q] -= lij * col[q];
        }
    }
    free(col);
    return FCFD_OK;
}

/* x <- M^-1 x, x[0] pinned to 0 (row 0 of M is the identity) */
void orc_band_solve(const orc_band *B, double *x) {
    const long n = B->n, b = B->b, w = b + 1;
    const double *L = B->L;
    x[0] = 0.0;
    for (long i = 0; i < n; ++i) {   /* L y = x */
        const long c0 = i - b > 0 ? i - b : 0;
        const double *ri = L + i * w + b - i;
        double acc = x[i];
        for (long c = c0; c < i; ++c) acc -= ri[c] * x[c];
        x[i] = acc / ri[i];
    }
    for (long i = n - 1; i >= 0; --i) {   /* L^T x = y */
        const double *ri = L + i * w + b - i;
        x[i] /= ri[i];
        const long c0 = i - b > 0 ? i - b : 0;
        const double xi = x[i];
        for (long c = c0; c < i; ++c) x[c] -= ri[c] * xi;
    }
}

static int ppe_band_factor(orc_state *s) {
    const long n = (long)s->nx * s->ny * s->nz;
    const long b = s->dim == 3 ? (long)s->nx * s->ny : s->nx;
    const long w = b + 1;
    orc_band B = {s->band, s->band_n, s->band_b};
    int rc = orc_band_alloc(&B, n, b);
    s->band = B.L; s->band_n = B.n; s->band_b = B.b;
    if (rc != FCFD_OK) return rc;
    double *L = s->band;
    /* lower band of M = -A: left (-1), back (-nx), down (-nx ny) and the diagonal */
    FOR_INT(s) {
        const long id = ID(s, i, j, k), r = ppe_cell(s, i, j, k);
        double *row = L + r * w + b;   /* row[c - r] for c <= r */
        row[0] = -s->coe_c[id];
        if (i > 0) row[-1] = -s->coe_l[id];
        if (j > 0) row[-(long)s->nx] = -s->coe_b[id];
        if (s->dim == 3 && k > 0) row[-b] = -s->coe_d[id];
    }
    /* pin cell 0: identity row and column */
    L[b] = 1.0;
    if (n > 1) L[1 * w + b - 1] = 0.0;
    if (n > s->nx) L[(long)s->nx * w + b - s->nx] = 0.0;
    if (s->dim == 3 && n > b) L[b * w] = 0.0;
    rc = orc_band_cholesky(&B);
    if (rc != FCFD_OK) return rc;
    s->band_valid = 1;
    return FCFD_OK;
}

static void ppe_band_solve(const orc_state *s, double *x) {
    const orc_band B = {s->band, s->band_n, s->band_b};
    orc_band_solve(&B, x);
}

/* A P = coe_src - mean(coe_src), mean(P) = 0 (interior), ghosts by bc. */
static int ppe_direct(orc_state *s) {
    const long n = (long)s->nx * s->ny * s->nz;
    double *P = s->p.lv[L_NOW];
    if (!s->band_valid) {
        int rc = ppe_band_factor(s);
        if (rc != FCFD_OK) return rc;
    }
    double *f = falloc(s->ntot), *Ap = falloc(s->ntot), *x = falloc(n);
    const double ncell = (double)n;
    double mean = 0.0;
    FOR_INT(s) mean += s->coe_src[ID(s, i, j, k)];
    mean /= ncell;
    FOR_INT(s) { long id = ID(s, i, j, k); f[id] = s->coe_src[id] - mean; }
    FOR_INT(s) { long id = ID(s, i, j, k); P[id] = 0.0; }
    for (int pass = 0; pass < 3; ++pass) {
        /* x = M^-1 (A P - f) = -A^-1 r with r = f - A P */
        orc_ppe_apply(s, P, Ap);
        FOR_INT(s) { long id = ID(s, i, j, k); x[ppe_cell(s, i, j, k)] = Ap[id] - f[id]; }
        ppe_band_solve(s, x);
        FOR_INT(s) { long id = ID(s, i, j, k); P[id] += x[ppe_cell(s, i, j, k)]; }
    }
    double pm = 0.0;
    FOR_INT(s) pm += P[ID(s, i, j, k)];
    pm /= ncell;
    FOR_INT(s) P[ID(s, i, j, k)] -= pm;
    bc(s, P);
    orc_ppe_apply(s, P, Ap);
    double rn = 0.0, fn = 0.0;
    FOR_INT(s) {
        long id = ID(s, i, j, k);
        double r = f[id] - Ap[id];
        rn += r * r;
        fn += f[id] * f[id];
    }
    s->ppe_rel = fn > 0.0 ? sqrt(rn / fn) : 0.0;
    s->piter += 3;
    free(f); free(Ap); free(x);
    return FCFD_OK;
}

/* Largest band factor the converged solve forms (doubles); bigger systems
 * use the projected Jacobi-PCG below.  Settable so tests can force either. */
static long orc_band_max = 400L * 1000L * 1000L;
void orc_set_band_max(long ndoubles) { orc_band_max = ndoubles; }

/* Jacobi-preconditioned CG on -A p = -(src - mean src) to rel_tol, the fallback
 * of the converged solve for systems whose band factor is too large.  The
 * preconditioned residual is projected onto zero mean (the range of A), so
 * the search directions carry no null-space component that rounding could
 * let grow. */
static int ppe_pcg(orc_state *s, double rel_tol) {
    const long n = s->ntot;
    double *P = s->p.lv[L_NOW];
    double *r = falloc(n), *z = falloc(n), *q = falloc(n), *Aq = falloc(n), *f = falloc(n);
    const double ncell = (double)s->nx * (double)s->ny * (double)s->nz;
    double mean = 0.0;
    PSUM_BEGIN(s, mean) mean += s->coe_src[ID(s, i, j, k)]; PSUM_END(mean)
    mean /= ncell;
    double fn = 0.0;
    PSUM_BEGIN(s, fn) { long id = ID(s, i, j, k); f[id] = -(s->coe_src[id] - mean); fn += f[id] * f[id]; } PSUM_END(fn)
    fn = sqrt(fn);
    orc_ppe_apply(s, P, Aq);
    double rn = 0.0, rz = 0.0, zm = 0.0;
    PSUM_BEGIN(s, rn) {
        long id = ID(s, i, j, k);
        r[id] = f[id] + Aq[id];
        rn += r[id] * r[id];
    } PSUM_END(rn)
    int it = 0, rc = FCFD_OK;
    if (fn == 0.0) fn = 1.0;
    if (sqrt(rn) / fn > rel_tol) {
        PSUM_BEGIN(s, zm) { long id = ID(s, i, j, k); z[id] = r[id] / (-s->coe_c[id]); zm += z[id]; } PSUM_END(zm)
        zm /= ncell;
        PSUM_BEGIN(s, rz) { long id = ID(s, i, j, k); z[id] -= zm; q[id] = z[id]; rz += r[id] * z[id]; } PSUM_END(rz)
        for (it = 1; it <= 200000; ++it) {
            orc_ppe_apply(s, q, Aq);
            double qAq = 0.0;
            PSUM_BEGIN(s, qAq) { long id = ID(s, i, j, k); Aq[id] = -Aq[id]; qAq += q[id] * Aq[id]; } PSUM_END(qAq)
            double alpha = rz / qAq;
            rn = 0.0;
            PSUM_BEGIN(s, rn) {
                long id = ID(s, i, j, k);
                P[id] += alpha * q[id];
                r[id] -= alpha * Aq[id];
                rn += r[id] * r[id];
            } PSUM_END(rn)
            if (sqrt(rn) / fn <= rel_tol) break;
            double rz2 = 0.0;
            zm = 0.0;
            PSUM_BEGIN(s, zm) { long id = ID(s, i, j, k); z[id] = r[id] / (-s->coe_c[id]); zm += z[id]; } PSUM_END(zm)
            zm /= ncell;
            PSUM_BEGIN(s, rz2) { long id = ID(s, i, j, k); z[id] -= zm; rz2 += r[id] * z[id]; } PSUM_END(rz2)
            double beta = rz2 / rz;
            rz = rz2;
            PFOR_INT(s) { long id = ID(s, i, j, k); q[id] = z[id] + beta * q[id]; }
        }
        if (it > 200000) rc = FCFD_ERR_NOT_CONVERGED;
    }
    double pm = 0.0;
    PSUM_BEGIN(s, pm) pm += P[ID(s, i, j, k)]; PSUM_END(pm)
    pm /= ncell;
    PFOR_INT(s) P[ID(s, i, j, k)] -= pm;
    bc(s, P);
    s->ppe_rel = sqrt(rn) / fn;
    s->piter += it;
    free(r); free(z); free(q); free(Aq); free(f);
    return rc;
}

/* The converged solve of coe_src, then mean(p) = 0: direct while the band
 * factor fits orc_band_max doubles (rel_tol unused: the solution is exact to
 * rounding), otherwise projected Jacobi-PCG to rel_tol. */
int orc_ppe_solve_rhs(orc_state *s, double rel_tol) {
    const long n = (long)s->nx * s->ny * s->nz;
    const long b = s->dim == 3 ? (long)s->nx * s->ny : s->nx;
    if (n * (b + 1) <= orc_band_max) return ppe_direct(s);
    return ppe_pcg(s, rel_tol);
}

/* Converged replacement of ppe_sor_solver: src = div(u*)/dt, the converged
 * solve, then ns_momentum_correction, as the relaxation ends with it. */
int orc_ppe_exact_solver(orc_state *s, double rel_tol) {
    ppe_src(s);
    int rc = orc_ppe_solve_rhs(s, rel_tol);
    orc_ns_momentum_correction(s);
    return rc;
}

/* find_stag_vel (3D bcs.f90:259-774; 2D bcs.f90:265-400).  Of the ghost
 * rules that follow the interior averages only the wall-face zeroing reaches
 * an interior index that the advection source reads: V(ie)=W(ie)=0 for the
 * east wall, U(je)=WW(je)=0 for the north wall, UU(ke)=VV(ke)=0 for the top
 * wall (3D bcs.f90:561-562,653-654,745-746).  The remaining ghost writes
 * land on scratch cells no later loop reads and are not restated. */
int orc_ns_linearize(orc_state *s) {
    double *Ut = s->u.lv[L_TMP], *Vt = s->v.lv[L_TMP], *Wt = s->w.lv[L_TMP];
    memcpy(Ut, s->u.lv[L_NOW], sizeof(double) * (size_t)s->ntot);
    memcpy(Vt, s->v.lv[L_NOW], sizeof(double) * (size_t)s->ntot);
    if (s->dim == 3) memcpy(Wt, s->w.lv[L_NOW], sizeof(double) * (size_t)s->ntot);
    /* U = tdata%x%s1, V = tdata%y%s1, W = tdata%z%s1, UU = x%s2, VV = y%s2, WW = z%s2 */
    double *U = s->t_s1[0], *V = s->t_s1[1], *W = s->t_s1[2], *UU = s->t_s2[0], *VV = s->t_s2[1], *WW = s->t_s2[2];
    const long sx = s->sx, sxy = s->sxy;
#pragma omp parallel for collapse(2)
    for (int k = 0; k < s->nz; ++k)
        for (int j = 0; j < s->ny; ++j)
            for (int i = 0; i < s->nx; ++i) {
                long id = ID(s, i, j, k);
                V[id] = 0.25 * (Vt[id] + Vt[id - sx] + Vt[id + 1] + Vt[id + 1 - sx]);
                U[id] = 0.25 * (Ut[id] + Ut[id - 1] + Ut[id + sx] + Ut[id - 1 + sx]);
                if (s->dim == 3) {
                    W[id] = 0.25 * (Wt[id] + Wt[id - sxy] + Wt[id + 1] + Wt[id + 1 - sxy]);
                    WW[id] = 0.25 * (Wt[id] + Wt[id - sxy] + Wt[id + sx] + Wt[id + sx - sxy]);
                    UU[id] = 0.25 * (Ut[id] + Ut[id - 1] + Ut[id + sxy] + Ut[id - 1 + sxy]);
                    VV[id] = 0.25 * (Vt[id] + Vt[id - sx] + Vt[id + sxy] + Vt[id - sx + sxy]);
                }
            }
    int ub = s->pr.ubc[1], vb = s->pr.vbc[1], wb = s->pr.wbc[1];
    if (ub == 1 || ub == 2)
        for (int k = 0; k < s->nz; ++k)
            for (int j = 0; j < s->ny; ++j) {
                V[ID(s, s->nx - 1, j, k)] = 0.0;
                if (s->dim == 3) W[ID(s, s->nx - 1, j, k)] = 0.0;
            }
    if (vb == 1 || vb == 2)
        for (int k = 0; k < s->nz; ++k)
            for (int i = 0; i < s->nx; ++i) {
                U[ID(s, i, s->ny - 1, k)] = 0.0;
                if (s->dim == 3) WW[ID(s, i, s->ny - 1, k)] = 0.0;
            }
    if (s->dim == 3 && (wb == 1 || wb == 2))
        for (int j = 0; j < s->ny; ++j)
            for (int i = 0; i < s->nx; ++i) {
                UU[ID(s, i, j, s->nz - 1)] = 0.0;
                VV[ID(s, i, j, s->nz - 1)] = 0.0;
            }
    return FCFD_OK;
}

/* sec_part1..3 (3D ns_ab_diff.f90:10-198; 2D :10-135) into velsrc%tmp */
static void ns_ab_diff_sec(orc_state *s) {
    const long sx = s->sx, sxy = s->sxy;
    const double dx = s->dx, dy = s->dy, dz = s->dz, re = s->pr.re, m12 = s->pr.mu_12;
    const double dx2 = dx * dx, dy2 = dy * dy, dz2 = dz * dz;
    double *Uo = s->u.lv[L_OLD], *Vo = s->v.lv[L_OLD], *Wo = s->w.lv[L_OLD];
    double *Ro = s->rho.lv[L_OLD], *Mo = s->mu.lv[L_OLD], *Do = s->delta.lv[L_OLD], *Po = s->phi.lv[L_OLD];
    double *Co = s->curv.lv[L_OLD], *NXo = s->nx_.lv[L_OLD], *NYo = s->ny_.lv[L_OLD], *NZo = s->nz_.lv[L_OLD];
    double *Tx = s->su.lv[L_TMP], *Ty = s->sv.lv[L_TMP], *Tz = s->sw.lv[L_TMP];
    const int d3 = s->dim == 3;
    /* sec_part1 */
#pragma omp parallel for collapse(2)
    for (int k = 0; k < s->nz; ++k)
        for (int j = 0; j < s->ny; ++j)
            for (int i = 0; i < s->nx; ++i) {
                long id = ID(s, i, j, k);
                double *Fs[3] = {Uo, Vo, Wo};
                long nb[3] = {1, sx, sxy};
                double *T[3] = {Tx, Ty, Tz};
                for (int c = 0; c < s->dim; ++c) {
                    double *F = Fs[c];
                    double rho = 0.5 * (Ro[id] + Ro[id + nb[c]]);
                    double mu = 0.5 * (Mo[id] + Mo[id + nb[c]]);
                    double xx = (F[id + 1] - 2.0 * F[id] + F[id - 1]) / dx2;
                    double yy = (F[id + sx] - 2.0 * F[id] + F[id - sx]) / dy2;
                    double dif_x = mu / rho * xx / re, dif_y = mu / rho * yy / re;
                    if (d3) {
                        double zz = (F[id + sxy] - 2.0 * F[id] + F[id - sxy]) / dz2;
                        double dif_z = mu / rho * zz / re;
                        T[c][id] = dif_x + dif_y + dif_z;
                    } else {
                        T[c][id] = dif_x + dif_y;
                    }
                }
            }
    /* sec_part2 */
#pragma omp parallel for collapse(2)
    for (int k = 0; k < s->nz; ++k)
        for (int j = 0; j < s->ny; ++j)
            for (int i = 0; i < s->nx; ++i) {
                long id = ID(s, i, j, k);
                double rho, delta, ux, uy, uz = 0, vx, vy, vz = 0, wx = 0, wy = 0, wz = 0, phix, phiy, phiz = 0;
                double dif_x, dif_y, dif_z = 0;
                /* x momentum */
                rho = 0.5 * (Ro[id] + Ro[id + 1]);
                delta = 0.5 * (Do[id] + Do[id + 1]);
                ux = 0.5 * (Uo[id + 1] - Uo[id - 1]) / dx;
                uy = 0.5 * (Uo[id + sx] - Uo[id - sx]) / dy;
                if (d3) uz = 0.5 * (Uo[id + sxy] - Uo[id - sxy]) / dz;
                vx = 0.5 * (Vo[id + 1] - Vo[id] + Vo[id + 1 - sx] - Vo[id - sx]) / dx;
                if (d3) wx = 0.5 * (Wo[id + 1] - Wo[id] + Wo[id + 1 - sxy] - Wo[id - sxy]) / dx;
                phix = (Po[id + 1] - Po[id]) / dx;
                phiy = 0.25 * (Po[id + 1 + sx] - Po[id + 1 - sx] + Po[id + sx] - Po[id - sx]) / dy;
                if (d3) phiz = 0.25 * (Po[id + 1 + sxy] - Po[id + 1 - sxy] + Po[id + sxy] - Po[id - sxy]) / dz;
                dif_x = (1.0 - m12) * delta * phix * 2.0 * ux / (rho * re);
                dif_y = (1.0 - m12) * delta * phiy * (uy + vx) / (rho * re);
                if (d3) {
                    dif_z = (1.0 - m12) * delta * phiz * (uz + wx) / (rho * re);
                    Tx[id] = Tx[id] + dif_x + dif_y + dif_z;
                } else {
                    Tx[id] = Tx[id] + dif_x + dif_y;
                }
                /* y momentum */
                rho = 0.5 * (Ro[id] + Ro[id + sx]);
                delta = 0.5 * (Do[id] + Do[id + sx]);
                vx = 0.5 * (Vo[id + 1] - Vo[id - 1]) / dx;
                vy = 0.5 * (Vo[id + sx] - Vo[id - sx]) / dy;
                if (d3) vz = 0.5 * (Vo[id + sxy] - Vo[id - sxy]) / dz;
                uy = 0.5 * (Uo[id + sx] - Uo[id] + Uo[id - 1 + sx] - Uo[id - 1]) / dy;
                if (d3) wy = 0.5 * (Wo[id + sx] - Wo[id] + Wo[id + sx - sxy] - Wo[id - sxy]) / dy;
                phix = 0.25 * (Po[id + 1] - Po[id - 1] + Po[id + 1 + sx] - Po[id - 1 + sx]) / dx;
                phiy = (Po[id + sx] - Po[id]) / dy;
                if (d3) phiz = 0.25 * (Po[id + sx + sxy] - Po[id + sx - sxy] + Po[id + sxy] - Po[id - sxy]) / dz;
                dif_x = (1.0 - m12) * delta * phix * (uy + vx) / (rho * re);
                dif_y = (1.0 - m12) * delta * phiy * 2.0 * vy / (rho * re);
                if (d3) {
                    dif_z = (1.0 - m12) * delta * phiz * (wy + vz) / (rho * re);
                    Ty[id] = Ty[id] + dif_x + dif_y + dif_z;
                } else {
                    Ty[id] = Ty[id] + dif_x + dif_y;
                }
                if (!d3) continue;
                /* z momentum */
                rho = 0.5 * (Ro[id] + Ro[id + sxy]);
                delta = 0.5 * (Do[id] + Do[id + sxy]);
                wx = 0.5 * (Wo[id + 1] - Wo[id - 1]) / dx;
                wy = 0.5 * (Wo[id + sx] - Wo[id - sx]) / dy;
                wz = 0.5 * (Wo[id + sxy] - Wo[id - sxy]) / dz;
                uz = 0.5 * (Uo[id + sxy] - Uo[id] + Uo[id - 1 + sxy] - Uo[id - 1]) / dz;
                vz = 0.5 * (Vo[id + sxy] - Vo[id] + Vo[id - sx + sxy] - Vo[id - sx]) / dz;
                phix = 0.25 * (Po[id + 1 + sxy] - Po[id - 1 + sxy] + Po[id + 1] - Po[id - 1]) / dx;
                phiy = 0.25 * (Po[id + sx + sxy] - Po[id - sx + sxy] + Po[id + sx] - Po[id - sx]) / dy;
                phiz = (Po[id + sxy] - Po[id]) / dz;
                dif_x = (1.0 - m12) * delta * phix * (uz + wx) / (rho * re);
                dif_y = (1.0 - m12) * delta * phiy * (vz + wy) / (rho * re);
                dif_z = (1.0 - m12) * delta * phiz * 2.0 * wz / (rho * re);
                Tz[id] = Tz[id] + dif_x + dif_y + dif_z;
            }
    /* sec_part3 */
    const double gvec[3] = {s->pr.gx, s->pr.gy, s->pr.gz};
    const double bg = s->pr.btn_g, fr = s->pr.fr, bsf = s->pr.btn_sf, we = s->pr.we;
#pragma omp parallel for collapse(2)
    for (int k = 0; k < s->nz; ++k)
        for (int j = 0; j < s->ny; ++j)
            for (int i = 0; i < s->nx; ++i) {
                long id = ID(s, i, j, k);
                long nb[3] = {1, sx, sxy};
                double *N[3] = {NXo, NYo, NZo};
                double *T[3] = {Tx, Ty, Tz};
                for (int c = 0; c < s->dim; ++c) {
                    double rho = 0.5 * (Ro[id] + Ro[id + nb[c]]);
                    double delta = 0.5 * (Do[id] + Do[id + nb[c]]);
                    double curv = (Co[id] + Co[id + nb[c]]) / 2.0;
                    double ph = 0.5 * (N[c][id] + N[c][id + nb[c]]);
                    T[c][id] = T[c][id] + gvec[c] * bg / fr - bsf * curv * delta * ph / (we * rho);
                }
            }
}

/* ns_ab_adv_source_sec: 2nd-order upwind (3D ns_ab_adv.f90:1-73; 2D :1-42) */
static void ns_ab_adv_source_sec(orc_state *s) {
    const long sx = s->sx, sxy = s->sxy;
    const double dx = s->dx, dy = s->dy, dz = s->dz;
    double *Uo = s->u.lv[L_OLD], *Vo = s->v.lv[L_OLD], *Wo = s->w.lv[L_OLD];
    double *Ut = s->u.lv[L_TMP], *Vt = s->v.lv[L_TMP], *Wt = s->w.lv[L_TMP];
    double *Sx = s->su.lv[L_NOW], *Sy = s->sv.lv[L_NOW], *Sz = s->sw.lv[L_NOW];
    const int d3 = s->dim == 3;
#pragma omp parallel for collapse(2)
    for (int k = 0; k < s->nz; ++k)
        for (int j = 0; j < s->ny; ++j)
            for (int i = 0; i < s->nx; ++i) {
                long id = ID(s, i, j, k);
                double *F[3] = {Uo, Vo, Wo};
                /* advecting velocities per component: (u, v, w) */
                double adv[3][3];
                adv[0][0] = Ut[id]; adv[0][1] = s->t_s1[1][id]; adv[0][2] = d3 ? s->t_s1[2][id] : 0.0;
                adv[1][0] = s->t_s1[0][id]; adv[1][1] = Vt[id]; adv[1][2] = d3 ? s->t_s2[2][id] : 0.0;
                adv[2][0] = d3 ? s->t_s2[0][id] : 0.0; adv[2][1] = d3 ? s->t_s2[1][id] : 0.0;
                adv[2][2] = d3 ? Wt[id] : 0.0;
                double *Out[3] = {Sx, Sy, Sz};
                for (int c = 0; c < s->dim; ++c) {
                    double *G = F[c];
                    double u = adv[c][0], v = adv[c][1], w = adv[c][2];
                    double xp = 0.5 * (-G[id + 2] + 4.0 * G[id + 1] - 3.0 * G[id]) / dx;
                    double xm = 0.5 * (G[id - 2] - 4.0 * G[id - 1] + 3.0 * G[id]) / dx;
                    double yp = 0.5 * (-G[id + 2 * sx] + 4.0 * G[id + sx] - 3.0 * G[id]) / dy;
                    double ym = 0.5 * (G[id - 2 * sx] - 4.0 * G[id - sx] + 3.0 * G[id]) / dy;
                    if (d3) {
                        double zp = 0.5 * (-G[id + 2 * sxy] + 4.0 * G[id + sxy] - 3.0 * G[id]) / dz;
                        double zm = 0.5 * (G[id - 2 * sxy] - 4.0 * G[id - sxy] + 3.0 * G[id]) / dz;
                        Out[c][id] = -((u + fabs(u)) * xm + (u - fabs(u)) * xp) / 2.0 -
                                     ((v + fabs(v)) * ym + (v - fabs(v)) * yp) / 2.0 -
                                     ((w + fabs(w)) * zm + (w - fabs(w)) * zp) / 2.0;
                    } else {
                        Out[c][id] = -((u + fabs(u)) * xm + (u - fabs(u)) * xp) / 2.0 -
                                     ((v + fabs(v)) * ym + (v - fabs(v)) * yp) / 2.0;
                    }
                }
            }
    (void)Wt;
}

/* ns_ab_setup (3D admas_bashforth_solver.f90:230-252; 2D :152-171) */
int orc_ns_ab_setup(orc_state *s) {
    ns_ab_diff_sec(s);
    ns_ab_adv_source_sec(s);
    double *N[3] = {s->su.lv[L_NOW], s->sv.lv[L_NOW], s->sw.lv[L_NOW]};
    double *T[3] = {s->su.lv[L_TMP], s->sv.lv[L_TMP], s->sw.lv[L_TMP]};
    for (int c = 0; c < s->dim; ++c) {
        double *n = N[c], *t = T[c];
        FOR_INT(s) { long id = ID(s, i, j, k); n[id] = n[id] + t[id]; }
    }
    return FCFD_OK;
}

/* ns_ab_predictor: AB2 (3D admas_bashforth_solver.f90:283-311; 2D :202-225) */
int orc_ns_ab_predictor(orc_state *s) {
    trec *V[3] = {&s->u, &s->v, &s->w};
    trec *S[3] = {&s->su, &s->sv, &s->sw};
    const double dt = s->pr.dt;
    for (int c = 0; c < s->dim; ++c) {
        double *vn = V[c]->lv[L_NOW], *vo = V[c]->lv[L_OLD], *sn = S[c]->lv[L_NOW], *so = S[c]->lv[L_OLD];
        FOR_INT(s) {
            long id = ID(s, i, j, k);
            double src = 1.5 * sn[id] - 0.5 * so[id];
            vn[id] = vo[id] + dt * src;
        }
    }
    orc_velbc(s);
    return FCFD_OK;
}

/* node_vel (3D ns_solvers.f90:128-147; 2D :117-133) */
int orc_node_vel(orc_state *s) {
    double *U = s->u.lv[L_NOW], *V = s->v.lv[L_NOW], *W = s->w.lv[L_NOW];
    double *NU = s->nu.lv[L_NOW], *NV = s->nv.lv[L_NOW], *NW = s->nw.lv[L_NOW];
    const long sx = s->sx, sxy = s->sxy;
    FOR_INT(s) {
        long id = ID(s, i, j, k);
        NU[id] = 0.5 * (U[id - 1] + U[id]);
        NV[id] = 0.5 * (V[id - sx] + V[id]);
        if (s->dim == 3) NW[id] = 0.5 * (W[id - sxy] + W[id]);
    }
    velbc_impl(s, NU, NV, NW, 1);
    return FCFD_OK;
}

/* ns_check_convergence_div (3D ns_solvers.f90:59-90; 2D :55-83) */
int orc_ns_check_convergence_div(orc_state *s) {
    double *U = s->u.lv[L_NOW], *V = s->v.lv[L_NOW], *W = s->w.lv[L_NOW];
    const long sx = s->sx, sxy = s->sxy;
    double div = 0.0, sumdiv = 0.0;
    FOR_INT(s) {
        long id = ID(s, i, j, k);
        double d = (U[id] - U[id - 1]) / s->dx + (V[id] - V[id - sx]) / s->dy;
        if (s->dim == 3) d = d + (W[id] - W[id - sxy]) / s->dz;
        div = fmax(div, fabs(d));
        sumdiv = sumdiv + fabs(d);
    }
    s->vel_sdiv = sumdiv / ((double)s->nx * (double)s->ny * (double)s->nz);
    s->vel_div = div;
    return FCFD_OK;
}

static void copy_level(orc_state *s, trec *t, int dst, int src) {
    memcpy(t->lv[dst], t->lv[src], sizeof(double) * (size_t)s->ntot);
}

/* manager_switch: old2 <- old, old <- now (3D+IBM/lib/data/tree.f90:229-252) */
int orc_switch(orc_state *s) {
    trec *list[] = {&s->phi, &s->rho, &s->mu, &s->delta, &s->heavy, &s->sign, &s->nx_, &s->ny_, &s->nz_,
                    &s->nxx, &s->nyy, &s->nzz, &s->nxy, &s->nxz, &s->nyz, &s->curv, &s->u, &s->v, &s->w,
                    &s->nu, &s->nv, &s->nw, &s->su, &s->sv, &s->sw, &s->p};
    for (size_t t = 0; t < sizeof(list) / sizeof(list[0]); ++t) {
        copy_level(s, list[t], L_OLD2, L_OLD);
        copy_level(s, list[t], L_OLD, L_NOW);
    }
    return FCFD_OK;
}

/* job_loc_init: old = old2 = tmp = now (3D+IBM/lib/data/branches.f90:145-166) */
int orc_init_levels(orc_state *s) {
    trec *list[] = {&s->phi, &s->heavy, &s->delta, &s->grad, &s->sign, &s->p, &s->rho, &s->mu,
                    &s->nx_, &s->ny_, &s->nz_, &s->nxx, &s->nyy, &s->nzz, &s->nxy, &s->nxz, &s->nyz, &s->curv,
                    &s->u, &s->v, &s->w, &s->nu, &s->nv, &s->nw, &s->su, &s->sv, &s->sw};
    for (size_t t = 0; t < sizeof(list) / sizeof(list[0]); ++t) {
        copy_level(s, list[t], L_OLD, L_NOW);
        copy_level(s, list[t], L_OLD2, L_NOW);
        copy_level(s, list[t], L_TMP, L_NOW);
    }
    return FCFD_OK;
}

/* mpls(): advect, reinitialise every redis_interval steps, mass correction
 * (3D+IBM/src/interface_solver/interface_solver.f90:25-34; the reference
 * hard-codes interval 5). */
int orc_interface_solver(orc_state *s) {
    orc_level_set_rk3_solver(s);
    int iv = s->pr.redis_interval > 0 ? s->pr.redis_interval : 5;
    if (s->iter % iv == 0) orc_level_set_rk3_redis(s, 1);
    orc_mass_preserving_level_set(s);
    return FCFD_OK;
}

/* ns_solver = ns_init + ns_ab_solver_SOR + diagnostics + node_vel
 * (3D ns_solvers.f90:1-31, admas_bashforth_solver.f90:254-281).  With
 * exact_ppe the relaxation is replaced by the converged solve to ppe_tol. */
int orc_ns_solver(orc_state *s, int exact_ppe) {
    int rc = FCFD_OK;
    orc_rho_mu(s);
    if (s->pr.btn_sf > 0.0) orc_curv(s);
    orc_ppe_sor_init(s);
    s->piter = 0;
    int np = s->pr.picard_iters > 0 ? s->pr.picard_iters : 4;
    for (int it = 1; it <= np; ++it) {
        orc_ns_linearize(s);
        orc_ns_ab_setup(s);
        orc_ns_ab_predictor(s);
        int r = exact_ppe ? orc_ppe_exact_solver(s, s->pr.ppe_tol) : orc_ppe_sor_solver(s, s->pr.p_tol);
        if (r != FCFD_OK) rc = r;
    }
    orc_ns_check_convergence_div(s);
    orc_node_vel(s);
    return rc;
}

/* one pass of the main loop body (3D+IBM/src/main.f90:15-24) */
int orc_step(orc_state *s, int exact_ppe) {
    s->time = s->time + s->pr.dt;
    s->iter = s->iter + 1;
    orc_switch(s);
    orc_interface_solver(s);
    int rc = orc_ns_solver(s, exact_ppe);
    orc_ls_mv(s);
    return rc;
}

/* problem_init after the initial condition is set
 * (3D+IBM/src/setup/problem_init.f90:109-133, without IBM and plotting):
 * bc, redis(0), node_vel, ns_init, p%loc%init, ns_ab_setup, ls_mv, imass. */
int orc_problem_init(orc_state *s, int redis0) {
    bc(s, s->phi.lv[L_NOW]);
    orc_velbc(s);
    if (redis0) orc_level_set_rk3_redis(s, 0);
    orc_node_vel(s);
    orc_rho_mu(s);
    if (s->pr.btn_sf > 0.0) orc_curv(s);
    orc_init_levels(s);
    orc_ns_ab_setup(s);
    orc_ls_mv(s);
    s->imass = s->mass;
    s->time = 0.0;
    s->iter = 0;
    return FCFD_OK;
}
