/*
 * m3_oracle.c — CPU restatement of the MPLS_RisingBubble reference step.
 *
 * TEST INFRASTRUCTURE: the checker the HIP path is compared against.  Never
 * linked into the product library.  Only tests/ and bench.py's cpu_baseline
 * leg load it.
 *
 * Reference files restated (paths under `MPLS_RisingBubble/20190612/`):
 *   LS_SOLVER.F90 (SOLVE_LS, LS_RK3_SOLVER, CELL_FACE_VELOCITY, LS_sorce,
 *   MPLS), LS_MAINTAIN.F90, LS_REDISTANCE.F90 (LS_REDISTANCE, LS_REDIS,
 *   LS_REDIS_SORCE, RIS_MASS, GRADFI, FV_DERI, gradf, HEAVY_F, HEAV, SGNN,
 *   sgnf), SCHEME_COMPACT.F90 (UCCD, CCD, ASSIGN_*, TWIN_DEC, TWIN_BKS),
 *   SCHEME_WENO.F90 (weno_js), FIND_CURVATURE.F90 (LS_CCD), AMURHO.F90,
 *   BC.F90 (bc3d, bcu, bcv, bcw with VEL_BC = 0), SOLVE_NS.F90,
 *   SECOND_ORDER_UPWIND.F90, SOLVE_PRESSURE_POISSON.F90, FIND_MASS.F90 (the
 *   level-set sums), STORE_DATA.F90, FIND_ERROR.F90.
 * Fortran indices are kept: arrays span (-2:NX+3, -2:NY+3, -2:NZ+3), i
 * fastest.  Loop order and expression order follow the reference, so a
 * single-threaded run agrees with the compiled reference (oracle/_ref/refm3)
 * bit for bit, with two deliberate departures:
 *   - LS_CCD's z-direction loops run J = 1..NODE_Z over the y index
 *     (FIND_CURVATURE.F90:128,157), which is wrong when NODE_Y != NODE_Z;
 *     here they run J = 1..NODE_Y.  Parity runs use NODE_Y = NODE_Z.
 *   - the pressure solve can be the reference's SOR (PRESS_POISSON) or the
 *     converged solve of the same operator (band Cholesky, band.h).
 * Each function is pinned against the compiled reference bit for bit
 * (tests/test_m3_oracle.py, tests/golden/m3_ref.npz); the converged pressure
 * solve as the fixed point of the reference's SOR.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "band.h"
#include "m3_oracle.h"

#define M3PI 3.14159265358979323846264338327950288419
#define M3EPS 1.0e-12

struct m3_state {
    m3_params pr;
    int nx, ny, nz;
    long sx, sxy, ntot;
    double *f[M3_NFIELDS];
    /* internal work arrays of the reference's DUMMY_DATA / FLUID_PROPERTIES */
    double *u_tmp, *v_tmp, *w_tmp, *phi_tmp, *s1, *s2, *a, *b, *c, *ff, *gg, *hh;
    double *up, *um, *vp, *vm, *wp, *wm, *p_old;
    /* scalars */
    double mass_ls, vol_ls, div, l2sum;
    int ppe_iters, picard;
    orc_band band;
    int band_valid;
};

#define IX(s, i, j, k) ((long)((i) + 2) + (long)((j) + 2) * (s)->sx + (long)((k) + 2) * (s)->sxy)
#define F_(s, n) ((s)->f[n])

static double *zalloc(long n) { return (double *)calloc((size_t)n, sizeof(double)); }
static inline double sq(double x) { return x * x; }

m3_state *m3_create(const m3_params *p) {
    if (!p || p->nx < 4 || p->ny < 4 || p->nz < 4) return NULL;
    m3_state *s = (m3_state *)calloc(1, sizeof(m3_state));
    s->pr = *p;
    s->nx = p->nx; s->ny = p->ny; s->nz = p->nz;
    s->sx = s->nx + 6;
    s->sxy = s->sx * (s->ny + 6);
    s->ntot = s->sxy * (s->nz + 6);
    for (int n = 0; n < M3_NFIELDS; ++n) s->f[n] = zalloc(s->ntot);
    double **w[] = {&s->u_tmp, &s->v_tmp, &s->w_tmp, &s->phi_tmp, &s->s1, &s->s2, &s->a, &s->b, &s->c, &s->ff,
                    &s->gg, &s->hh, &s->up, &s->um, &s->vp, &s->vm, &s->wp, &s->wm, &s->p_old};
    for (size_t t = 0; t < sizeof(w) / sizeof(w[0]); ++t) *w[t] = zalloc(s->ntot);
    return s;
}

void m3_destroy(m3_state *s) {
    if (!s) return;
    for (int n = 0; n < M3_NFIELDS; ++n) free(s->f[n]);
    double *w[] = {s->u_tmp, s->v_tmp, s->w_tmp, s->phi_tmp, s->s1, s->s2, s->a, s->b, s->c, s->ff,
                   s->gg, s->hh, s->up, s->um, s->vp, s->vm, s->wp, s->wm, s->p_old};
    for (size_t t = 0; t < sizeof(w) / sizeof(w[0]); ++t) free(w[t]);
    free(s->band.L);
    free(s);
}

int m3_set(m3_state *s, int n, const double *src) {
    if (n < 0 || n >= M3_NFIELDS || !src) return 1;
    memcpy(s->f[n], src, sizeof(double) * (size_t)s->ntot);
    s->band_valid = 0;
    return 0;
}

int m3_get(const m3_state *s, int n, double *dst) {
    if (n < 0 || n >= M3_NFIELDS || !dst) return 1;
    memcpy(dst, s->f[n], sizeof(double) * (size_t)s->ntot);
    return 0;
}

void m3_set_params(m3_state *s, const m3_params *p) {
    if (p->nx == s->nx && p->ny == s->ny && p->nz == s->nz) s->pr = *p;
}

void m3_get_params(const m3_state *s, m3_params *p) { *p = s->pr; }

void m3_get_diag(const m3_state *s, m3_diag *d) {
    d->mass_ls = s->mass_ls;
    d->vol_ls = s->vol_ls;
    d->div = s->div;
    d->l2sum = s->l2sum;
    d->ppe_iters = s->ppe_iters;
    d->picard = s->picard;
}

/* ------------------------------------------------------------------ BC.F90 */

/* bc3d (BC.F90:10-60): zero-gradient ghosts, x then y then z */
static void bc3d(m3_state *s, double *fi) {
    const int NX = s->nx, NY = s->ny, NZ = s->nz;
    for (int k = 1; k <= NZ; ++k)
        for (int j = 1; j <= NY; ++j) {
            const double a = fi[IX(s, 1, j, k)], b = fi[IX(s, NX, j, k)];
            fi[IX(s, 0, j, k)] = a; fi[IX(s, -1, j, k)] = a; fi[IX(s, -2, j, k)] = a;
            fi[IX(s, NX + 1, j, k)] = b; fi[IX(s, NX + 2, j, k)] = b; fi[IX(s, NX + 3, j, k)] = b;
        }
    for (int k = 1; k <= NZ; ++k)
        for (int i = -2; i <= NX + 3; ++i) {
            const double a = fi[IX(s, i, 1, k)], b = fi[IX(s, i, NY, k)];
            fi[IX(s, i, 0, k)] = a; fi[IX(s, i, -1, k)] = a; fi[IX(s, i, -2, k)] = a;
            fi[IX(s, i, NY + 1, k)] = b; fi[IX(s, i, NY + 2, k)] = b; fi[IX(s, i, NY + 3, k)] = b;
        }
    for (int j = -2; j <= NY + 3; ++j)
        for (int i = -2; i <= NX + 3; ++i) {
            const double a = fi[IX(s, i, j, 1)], b = fi[IX(s, i, j, NZ)];
            fi[IX(s, i, j, 0)] = a; fi[IX(s, i, j, -1)] = a; fi[IX(s, i, j, -2)] = a;
            fi[IX(s, i, j, NZ + 1)] = b; fi[IX(s, i, j, NZ + 2)] = b; fi[IX(s, i, j, NZ + 3)] = b;
        }
}

/* bcu / bcv / bcw with VEL_BC = 0 (BC.F90:62-118, 250-..., no-slip): the
 * normal component is zero on its walls and odd across them, the tangential
 * components are odd about the wall faces.  dir = 0, 1, 2 for u, v, w. */
static void bc_vel(m3_state *s, double *fi, int dir) {
    const int NX = s->nx, NY = s->ny, NZ = s->nz;
#define A3(i, j, k) fi[IX(s, i, j, k)]
    if (dir == 0) {
        for (int k = 1; k <= NZ; ++k)
            for (int j = 1; j <= NY; ++j) {
                A3(0, j, k) = 0.0; A3(-1, j, k) = -A3(1, j, k); A3(-2, j, k) = -A3(2, j, k);
                A3(NX, j, k) = 0.0; A3(NX + 1, j, k) = -A3(NX - 1, j, k); A3(NX + 2, j, k) = -A3(NX - 2, j, k);
                A3(NX + 3, j, k) = -A3(NX - 3, j, k);
            }
        for (int k = 1; k <= NZ; ++k)
            for (int i = -2; i <= NX + 3; ++i) {
                A3(i, 0, k) = -A3(i, 1, k); A3(i, -1, k) = -A3(i, 2, k); A3(i, -2, k) = -A3(i, 3, k);
                A3(i, NY + 1, k) = -A3(i, NY, k); A3(i, NY + 2, k) = -A3(i, NY - 1, k);
                A3(i, NY + 3, k) = -A3(i, NY - 2, k);
            }
        for (int j = -2; j <= NY + 3; ++j)
            for (int i = -2; i <= NX + 3; ++i) {
                A3(i, j, 0) = -A3(i, j, 1); A3(i, j, -1) = -A3(i, j, 2); A3(i, j, -2) = -A3(i, j, 3);
                A3(i, j, NZ + 1) = -A3(i, j, NZ); A3(i, j, NZ + 2) = -A3(i, j, NZ - 1);
                A3(i, j, NZ + 3) = -A3(i, j, NZ - 2);
            }
    } else if (dir == 1) {
        for (int k = 1; k <= NZ; ++k)
            for (int i = 1; i <= NX; ++i) {
                A3(i, 0, k) = 0.0; A3(i, -1, k) = -A3(i, 1, k); A3(i, -2, k) = -A3(i, 2, k);
                A3(i, NY, k) = 0.0; A3(i, NY + 1, k) = -A3(i, NY - 1, k); A3(i, NY + 2, k) = -A3(i, NY - 2, k);
                A3(i, NY + 3, k) = -A3(i, NY - 3, k);
            }
        for (int k = 1; k <= NZ; ++k)
            for (int j = -2; j <= NY + 3; ++j) {
                A3(0, j, k) = -A3(1, j, k); A3(-1, j, k) = -A3(2, j, k); A3(-2, j, k) = -A3(3, j, k);
                A3(NX + 1, j, k) = -A3(NX, j, k); A3(NX + 2, j, k) = -A3(NX - 1, j, k);
                A3(NX + 3, j, k) = -A3(NX - 2, j, k);
            }
        for (int j = -2; j <= NY + 3; ++j)
            for (int i = -2; i <= NX + 3; ++i) {
                A3(i, j, 0) = -A3(i, j, 1); A3(i, j, -1) = -A3(i, j, 2); A3(i, j, -2) = -A3(i, j, 3);
                A3(i, j, NZ + 1) = -A3(i, j, NZ); A3(i, j, NZ + 2) = -A3(i, j, NZ - 1);
                A3(i, j, NZ + 3) = -A3(i, j, NZ - 2);
            }
    } else {
        for (int j = 1; j <= NY; ++j)
            for (int i = 1; i <= NX; ++i) {
                A3(i, j, 0) = 0.0; A3(i, j, -1) = -A3(i, j, 1); A3(i, j, -2) = -A3(i, j, 2);
                A3(i, j, NZ) = 0.0; A3(i, j, NZ + 1) = -A3(i, j, NZ - 1); A3(i, j, NZ + 2) = -A3(i, j, NZ - 2);
                A3(i, j, NZ + 3) = -A3(i, j, NZ - 3);
            }
        for (int k = -2; k <= NZ + 3; ++k)
            for (int j = 1; j <= NY; ++j) {
                A3(0, j, k) = -A3(1, j, k); A3(-1, j, k) = -A3(2, j, k); A3(-2, j, k) = -A3(3, j, k);
                A3(NX + 1, j, k) = -A3(NX, j, k); A3(NX + 2, j, k) = -A3(NX - 1, j, k);
                A3(NX + 3, j, k) = -A3(NX - 2, j, k);
            }
        for (int k = -2; k <= NZ + 3; ++k)
            for (int i = -2; i <= NX + 3; ++i) {
                A3(i, 0, k) = -A3(i, 1, k); A3(i, -1, k) = -A3(i, 2, k); A3(i, -2, k) = -A3(i, 3, k);
                A3(i, NY + 1, k) = -A3(i, NY, k); A3(i, NY + 2, k) = -A3(i, NY - 1, k);
                A3(i, NY + 3, k) = -A3(i, NY - 2, k);
            }
    }
#undef A3
}

/* ------------------------------------------------------- LS_REDISTANCE.F90 */

/* HEAV (LS_REDISTANCE.F90:330-348): smoothed Heaviside and its derivative */
static double heav(double x, double w, double *hp) {
    double h;
    if (x > w) {
        h = 1.0; *hp = 0.0;
    } else if (x < -w) {
        h = 0.0; *hp = 0.0;
    } else {
        h = 0.5 * (1.0 + x / w + 1.0 / M3PI * sin(M3PI * x / w));
        *hp = 0.5 * (1.0 / w + 1.0 / w * cos(M3PI * x / w));
    }
    return h;
}

/* HEAVY_F (LS_REDISTANCE.F90:310-328): HEAVY, DELTA over the interior, then
 * bc3d(HEAVY) */
static void heavy_f(m3_state *s, const double *T) {
    double *H = F_(s, M3_HEAVY), *D = F_(s, M3_DELTA);
    const double w = s->pr.interface_width;
    for (int k = 1; k <= s->nz; ++k)
        for (int j = 1; j <= s->ny; ++j)
            for (int i = 1; i <= s->nx; ++i) {
                const long id = IX(s, i, j, k);
                double hp;
                H[id] = heav(T[id], w, &hp);
                D[id] = hp;
                if (H[id] < M3EPS) H[id] = 0.0;
            }
    bc3d(s, H);
}

/* sgnf (LS_REDISTANCE.F90:360-376): SGN = 2 (H - 0.5) */
static void sgnf(m3_state *s, const double *T) {
    double *S = F_(s, M3_SGN);
    const double w = s->pr.interface_width;
    for (int k = 1; k <= s->nz; ++k)
        for (int j = 1; j <= s->ny; ++j)
            for (int i = 1; i <= s->nx; ++i) {
                double hp;
                const double h = heav(T[IX(s, i, j, k)], w, &hp);
                S[IX(s, i, j, k)] = 2.0 * (h - 0.5);
            }
}

/* weno_js (SCHEME_WENO.F90:47-101) along one line of stride st whose point
 * i (Fortran index -2..n+3) is f0[(i + 2) st]: fm(i), fp(i) for i = 0..n */
static void weno_js(const double *f0, double *fp0, double *fm0, long st, int n) {
#define FV(i) f0[(long)((i) + 2) * st]
    for (int i = 0; i <= n; ++i) {
        const double b1 = 13.0 * sq(FV(i - 2) - 2.0 * FV(i - 1) + FV(i)) + 3.0 * sq(FV(i - 2) - 4.0 * FV(i - 1) + 3.0 * FV(i));
        const double b2 = 13.0 * sq(FV(i - 1) - 2.0 * FV(i) + FV(i + 1)) + 3.0 * sq(FV(i - 1) - FV(i + 1));
        const double b3 = 13.0 * sq(FV(i) - 2.0 * FV(i + 1) + FV(i + 2)) + 3.0 * sq(3.0 * FV(i) - 4.0 * FV(i + 1) + FV(i + 2));
        const double a1 = 1.0 / sq(M3EPS + b1), a2 = 6.0 / sq(M3EPS + b2), a3 = 3.0 / sq(M3EPS + b3);
        const double w1 = a1 / (a1 + a2 + a3), w2 = a2 / (a1 + a2 + a3), w3 = a3 / (a1 + a2 + a3);
        fm0[(long)(i + 2) * st] = w1 / 3.0 * FV(i - 2) - (7.0 * w1 + w2) / 6.0 * FV(i - 1) +
                                  (11.0 * w1 + 5.0 * w2 + 2.0 * w3) / 6.0 * FV(i) + (2.0 * w2 + 5.0 * w3) / 6.0 * FV(i + 1) -
                                  w3 / 6.0 * FV(i + 2);
    }
    for (int i = 0; i <= n; ++i) {
        const double b3 = 13.0 * sq(FV(i - 1) - 2.0 * FV(i) + FV(i + 1)) + 3.0 * sq(FV(i - 1) - 4.0 * FV(i) + 3.0 * FV(i + 1));
        const double b2 = 13.0 * sq(FV(i) - 2.0 * FV(i + 1) + FV(i + 2)) + 3.0 * sq(FV(i) - FV(i + 2));
        const double b1 = 13.0 * sq(FV(i + 1) - 2.0 * FV(i + 2) + FV(i + 3)) + 3.0 * sq(3.0 * FV(i + 1) - 4.0 * FV(i + 2) + FV(i + 3));
        const double a1 = 1.0 / sq(M3EPS + b1), a2 = 6.0 / sq(M3EPS + b2), a3 = 3.0 / sq(M3EPS + b3);
        const double w1 = a1 / (a1 + a2 + a3), w2 = a2 / (a1 + a2 + a3), w3 = a3 / (a1 + a2 + a3);
        fp0[(long)(i + 2) * st] = w3 * (-FV(i - 1) + 5.0 * FV(i) + 2.0 * FV(i + 1)) / 6.0 +
                                  w2 * (2.0 * FV(i) + 5.0 * FV(i + 1) - FV(i + 2)) / 6.0 +
                                  w1 * (11.0 * FV(i + 1) - 7.0 * FV(i + 2) + 2.0 * FV(i + 3)) / 6.0;
    }
#undef FV
}

/* gradf (LS_REDISTANCE.F90:282-308): Godunov |grad| by the sign of fi0 */
static double gradf(double up, double um, double vp, double vm, double wp, double wm, double fi0) {
    const double upm = -fmin(up, 0.0), upp = fmax(up, 0.0), umm = -fmin(um, 0.0), ump = fmax(um, 0.0);
    const double vpm = -fmin(vp, 0.0), vpp = fmax(vp, 0.0), vmm = -fmin(vm, 0.0), vmp = fmax(vm, 0.0);
    const double wpm = -fmin(wp, 0.0), wpp = fmax(wp, 0.0), wmm = -fmin(wm, 0.0), wmp = fmax(wm, 0.0);
    if (fi0 >= 0.0) return sqrt(sq(fmax(upm, ump)) + sq(fmax(vpm, vmp)) + sq(fmax(wpm, wmp)));
    return sqrt(sq(fmax(upp, umm)) + sq(fmax(vpp, vmm)) + sq(fmax(wpp, wmm)));
}

/* FV_DERI (LS_REDISTANCE.F90:240-280): backward differences, bc3d, weno_js
 * along every interior line of each direction */
static void fv_deri(m3_state *s, const double *T) {
    const int NX = s->nx, NY = s->ny, NZ = s->nz;
    for (int k = 1; k <= NZ; ++k)
        for (int j = 1; j <= NY; ++j)
            for (int i = 1; i <= NX; ++i) {
                const long id = IX(s, i, j, k);
                s->ff[id] = (T[id] - T[id - 1]) / s->pr.dx;
                s->gg[id] = (T[id] - T[id - s->sx]) / s->pr.dy;
                s->hh[id] = (T[id] - T[id - s->sxy]) / s->pr.dz;
            }
    bc3d(s, s->ff);
    bc3d(s, s->gg);
    bc3d(s, s->hh);
    for (int k = 1; k <= NZ; ++k)
        for (int j = 1; j <= NY; ++j) {
            const long o = IX(s, -2, j, k);
            weno_js(s->ff + o, s->up + o, s->um + o, 1, NX);
        }
    for (int k = 1; k <= NZ; ++k)
        for (int i = 1; i <= NX; ++i) {
            const long o = IX(s, i, -2, k);
            weno_js(s->gg + o, s->vp + o, s->vm + o, s->sx, NY);
        }
    for (int j = 1; j <= NY; ++j)
        for (int i = 1; i <= NX; ++i) {
            const long o = IX(s, i, j, -2);
            weno_js(s->hh + o, s->wp + o, s->wm + o, s->sxy, NZ);
        }
}

/* GRADFI (LS_REDISTANCE.F90:222-238) */
static void gradfi(m3_state *s, const double *T) {
    fv_deri(s, T);
    double *G = F_(s, M3_GRAD), *S = F_(s, M3_SGN);
    for (int k = 1; k <= s->nz; ++k)
        for (int j = 1; j <= s->ny; ++j)
            for (int i = 1; i <= s->nx; ++i) {
                const long id = IX(s, i, j, k);
                G[id] = gradf(s->up[id], s->um[id], s->vp[id], s->vm[id], s->wp[id], s->wm[id], S[id]);
            }
}

/* RIS_MASS with BTN = 1 (LS_REDISTANCE.F90:148-211): the mass-constrained
 * correction C = delta (51 a + sum_27 a) / (51 b + sum_27 b) with
 * a = SGN delta (|grad| - 1), b = |grad| delta^2; the 27-point sums in the
 * column-major element order of the section a(i-1:i+1, j-1:j+1, k-1:k+1) */
static void ris_mass(m3_state *s, const double *T) {
    heavy_f(s, T);
    const double *S = F_(s, M3_SGN), *D = F_(s, M3_DELTA), *G = F_(s, M3_GRAD);
    for (int k = 1; k <= s->nz; ++k)
        for (int j = 1; j <= s->ny; ++j)
            for (int i = 1; i <= s->nx; ++i) {
                const long id = IX(s, i, j, k);
                s->a[id] = (S[id] * D[id]) * (G[id] - 1.0);
                s->b[id] = G[id] * sq(D[id]);
            }
    bc3d(s, s->a);
    bc3d(s, s->b);
    for (int k = 1; k <= s->nz; ++k)
        for (int j = 1; j <= s->ny; ++j)
            for (int i = 1; i <= s->nx; ++i) {
                const long id = IX(s, i, j, k);
                double sa = 0.0, sb = 0.0;
                for (int kk = -1; kk <= 1; ++kk)
                    for (int jj = -1; jj <= 1; ++jj)
                        for (int ii = -1; ii <= 1; ++ii) {
                            const long q = id + ii + jj * s->sx + kk * s->sxy;
                            sa += s->a[q];
                            sb += s->b[q];
                        }
                const double suma = 51.0 * s->a[id] + sa, sumb = 51.0 * s->b[id] + sb;
                double c = sumb != 0.0 ? suma / sumb : 0.0;
                s->c[id] = c * D[id];
            }
}

/* LS_REDIS_SORCE (LS_REDISTANCE.F90:128-146) */
static void redis_source(m3_state *s, double *out) {
    const double *S = F_(s, M3_SGN), *G = F_(s, M3_GRAD);
    for (int k = 1; k <= s->nz; ++k)
        for (int j = 1; j <= s->ny; ++j)
            for (int i = 1; i <= s->nx; ++i) {
                const long id = IX(s, i, j, k);
                out[id] = (S[id] - s->c[id]) * G[id] - S[id];
            }
}

/* LS_REDIS with BTN = 1 (LS_REDISTANCE.F90:68-126): three RK stages */
static void ls_redis(m3_state *s, double *T) {
    double *S0 = F_(s, M3_S0);
    const double rdt = s->pr.rdt;
    for (int stage = 0; stage < 3; ++stage) {
        double *out = stage == 0 ? S0 : (stage == 1 ? s->s1 : s->s2);
        gradfi(s, T);
        ris_mass(s, T);
        redis_source(s, out);
        for (int k = 1; k <= s->nz; ++k)
            for (int j = 1; j <= s->ny; ++j)
                for (int i = 1; i <= s->nx; ++i) {
                    const long id = IX(s, i, j, k);
                    if (stage == 0) T[id] = T[id] - rdt * S0[id];
                    else if (stage == 1) T[id] = T[id] - rdt / 4.0 * (-3.0 * S0[id] + s->s1[id]);
                    else T[id] = T[id] - rdt / 12.0 * (-S0[id] - s->s1[id] + 8.0 * s->s2[id]);
                }
        bc3d(s, T);
    }
}

/* LS_REDISTANCE(TARGET, 1) (LS_REDISTANCE.F90:1-66): sign, then pseudo steps
 * until RED_TIME > 2 INTERFACE_WIDTH */
static void ls_redistance(m3_state *s, double *T) {
    sgnf(s, T);
    double red_time = 0.0;
    for (;;) {
        red_time = red_time + s->pr.rdt;
        for (int k = 1; k <= s->nz; ++k)
            for (int j = 1; j <= s->ny; ++j)
                for (int i = 1; i <= s->nx; ++i) s->phi_tmp[IX(s, i, j, k)] = T[IX(s, i, j, k)];
        ls_redis(s, T);
        if (red_time > 2.0 * s->pr.interface_width) break;
    }
}

/* ------------------------------------------------------ SCHEME_COMPACT.F90 */

/* the line work space of the compact schemes, 1-based: X[(r - 1) * n + i - 1] */
typedef struct {
    int n;
    double *A, *B, *AA, *BB, *AD, *BD, *AAD, *BBD, *S, *SS, *SD, *SSD, *U, *F;
} m3_line;

static m3_line line_alloc(int n) {
    m3_line L;
    L.n = n;
    double **p[] = {&L.A, &L.B, &L.AA, &L.BB, &L.AD, &L.BD, &L.AAD, &L.BBD};
    for (int t = 0; t < 8; ++t) *p[t] = zalloc(3L * n);
    double **q[] = {&L.S, &L.SS, &L.SD, &L.SSD, &L.U, &L.F};
    for (int t = 0; t < 6; ++t) *q[t] = zalloc(n);
    return L;
}

static void line_free(m3_line *L) {
    double *p[] = {L->A, L->B, L->AA, L->BB, L->AD, L->BD, L->AAD, L->BBD, L->S, L->SS, L->SD, L->SSD, L->U, L->F};
    for (int t = 0; t < 14; ++t) free(p[t]);
}

#define M2(X, r, i) (X)[((r) - 1) * n + (i) - 1]
#define V1(X, i) (X)[(i) - 1]

/* ASSIGN_CCD_TWIN (SCHEME_COMPACT.F90:54-141) */
static void assign_ccd(int n, double Dx, double *A, double *B, double *AA, double *BB) {
    const double A1 = 7.0 / 16.0, A3 = 7.0 / 16.0, B1 = 1.0 / 16.0, B3 = -1.0 / 16.0;
    const double AA1 = -9.0 / 8.0, AA3 = 9.0 / 8.0, BB1 = -1.0 / 8.0, BB3 = -1.0 / 8.0;
    M2(A, 2, 1) = 1.0; M2(B, 2, 1) = 0.0; M2(AA, 2, 1) = 0.0; M2(BB, 2, 1) = 1.0;
    M2(A, 3, 1) = 2.0; M2(B, 3, 1) = -Dx; M2(AA, 3, 1) = -2.5 / Dx; M2(BB, 3, 1) = 8.5;
    for (int i = 2; i <= n - 1; ++i) {
        M2(A, 1, i) = A1; M2(B, 1, i) = B1 * Dx; M2(AA, 1, i) = AA1 / Dx; M2(BB, 1, i) = BB1;
        M2(A, 2, i) = 1.0; M2(B, 2, i) = 0.0; M2(AA, 2, i) = 0.0; M2(BB, 2, i) = 1.0;
        M2(A, 3, i) = A3; M2(B, 3, i) = B3 * Dx; M2(AA, 3, i) = AA3 / Dx; M2(BB, 3, i) = BB3;
    }
    M2(A, 1, n) = 2.0; M2(B, 1, n) = Dx; M2(AA, 1, n) = 2.5 / Dx; M2(BB, 1, n) = 8.5;
    M2(A, 2, n) = 1.0; M2(B, 2, n) = 0.0; M2(AA, 2, n) = 0.0; M2(BB, 2, n) = 1.0;
}

/* ASSIGN_CCD_UD_TWIN (SCHEME_COMPACT.F90:143-326), M-DRPCCD coefficients */
static void assign_ccd_ud(int n, double Dx, double *A, double *B, double *AA, double *BB, double *AD, double *BD,
                          double *AAD, double *BBD) {
    const double A1 = 0.875, B1 = 0.125100442885824, B2 = -0.248995571141759, B3 = 0.0001004428858241;
    const double A3D = 0.875, B1D = -0.0001004428858241, B2D = 0.248995571141759, B3D = -0.125100442885824;
    const double AA1 = -9.0 / 8.0, AA3 = 9.0 / 8.0, BB1 = -1.0 / 8.0, BB3 = -1.0 / 8.0;
    M2(A, 2, 1) = 1.0; M2(B, 2, 1) = 0.0; M2(AA, 2, 1) = 0.0; M2(BB, 2, 1) = 1.0;
    M2(A, 3, 1) = 2.0; M2(B, 3, 1) = -Dx; M2(AA, 3, 1) = -2.5 / Dx; M2(BB, 3, 1) = 8.5;
    for (int i = 2; i <= n - 1; ++i) {
        M2(A, 1, i) = A1; M2(B, 1, i) = B1 * Dx; M2(AA, 1, i) = AA1 / Dx; M2(BB, 1, i) = BB1;
        M2(A, 2, i) = 1.0; M2(B, 2, i) = B2 * Dx; M2(AA, 2, i) = 0.0; M2(BB, 2, i) = 1.0;
        M2(A, 3, i) = 0.0; M2(B, 3, i) = B3 * Dx; M2(AA, 3, i) = AA3 / Dx; M2(BB, 3, i) = BB3;
    }
    M2(A, 1, n) = 2.0; M2(B, 1, n) = Dx; M2(AA, 1, n) = 2.5 / Dx; M2(BB, 1, n) = 8.5;
    M2(A, 2, n) = 1.0; M2(B, 2, n) = 0.0; M2(AA, 2, n) = 0.0; M2(BB, 2, n) = 1.0;
    M2(AD, 2, 1) = 1.0; M2(BD, 2, 1) = 0.0; M2(AAD, 2, 1) = 0.0; M2(BBD, 2, 1) = 1.0;
    M2(AD, 3, 1) = 2.0; M2(BD, 3, 1) = -Dx; M2(AAD, 3, 1) = -2.5 / Dx; M2(BBD, 3, 1) = 8.5;
    for (int i = 2; i <= n - 1; ++i) {
        M2(AD, 1, i) = 0.0; M2(BD, 1, i) = B1D * Dx; M2(AAD, 1, i) = AA1 / Dx; M2(BBD, 1, i) = BB1;
        M2(AD, 2, i) = 1.0; M2(BD, 2, i) = B2D * Dx; M2(AAD, 2, i) = 0.0; M2(BBD, 2, i) = 1.0;
        M2(AD, 3, i) = A3D; M2(BD, 3, i) = B3D * Dx; M2(AAD, 3, i) = AA3 / Dx; M2(BBD, 3, i) = BB3;
    }
    M2(AD, 1, n) = 2.0; M2(BD, 1, n) = Dx; M2(AAD, 1, n) = 2.5 / Dx; M2(BBD, 1, n) = 8.5;
    M2(AD, 2, n) = 1.0; M2(BD, 2, n) = 0.0; M2(AAD, 2, n) = 0.0; M2(BBD, 2, n) = 1.0;
}

/* TWIN_DEC (SCHEME_COMPACT.F90:328-377) */
static void twin_dec(int n, double *A, double *B, double *AA, double *BB) {
    M2(A, 1, 1) = 0.0; M2(AA, 1, 1) = 0.0; M2(A, 3, n) = 0.0; M2(AA, 3, n) = 0.0;
    M2(B, 1, 1) = 0.0; M2(BB, 1, 1) = 0.0; M2(B, 3, n) = 0.0; M2(BB, 3, n) = 0.0;
    for (int i = 2; i <= n; ++i) {
        const double Den = M2(A, 2, i - 1) * M2(BB, 2, i - 1) - M2(AA, 2, i - 1) * M2(B, 2, i - 1);
        double SC1 = -(M2(AA, 2, i - 1) * M2(B, 1, i) - M2(A, 1, i) * M2(BB, 2, i - 1)) / Den;
        double SC2 = (M2(B, 1, i) * M2(A, 2, i - 1) - M2(B, 2, i - 1) * M2(A, 1, i)) / Den;
        M2(A, 1, i) = SC1;
        M2(A, 2, i) = M2(A, 2, i) - (SC1 * M2(A, 3, i - 1) + SC2 * M2(AA, 3, i - 1));
        M2(B, 1, i) = SC2;
        M2(B, 2, i) = M2(B, 2, i) - (SC1 * M2(B, 3, i - 1) + SC2 * M2(BB, 3, i - 1));
        SC1 = -(M2(AA, 2, i - 1) * M2(BB, 1, i) - M2(AA, 1, i) * M2(BB, 2, i - 1)) / Den;
        SC2 = (M2(BB, 1, i) * M2(A, 2, i - 1) - M2(B, 2, i - 1) * M2(AA, 1, i)) / Den;
        M2(AA, 1, i) = SC1;
        M2(AA, 2, i) = M2(AA, 2, i) - (SC1 * M2(A, 3, i - 1) + SC2 * M2(AA, 3, i - 1));
        M2(BB, 1, i) = SC2;
        M2(BB, 2, i) = M2(BB, 2, i) - (SC1 * M2(B, 3, i - 1) + SC2 * M2(BB, 3, i - 1));
    }
}

/* ASSIGN_CCD_TWIN_SRC (SCHEME_COMPACT.F90:379-437) */
static void assign_ccd_src(int n, double Dx, const double *U, double *S, double *SS) {
    const double C1 = -15.0 / 16.0, C3 = 15.0 / 16.0, CC1 = 3.0, CC2 = -6.0, CC3 = 3.0;
    V1(S, 1) = (-3.5 * V1(U, 1) + 4.0 * V1(U, 2) - 0.5 * V1(U, 3)) / Dx;
    V1(SS, 1) = (34.0 / 3.0 * V1(U, 1) - 83.0 / 4.0 * V1(U, 2) + 10.0 * V1(U, 3) - 7.0 / 12.0 * V1(U, 4)) / (Dx * Dx);
    for (int i = 2; i <= n - 1; ++i) {
        V1(S, i) = (C1 * V1(U, i - 1) + C3 * V1(U, i + 1)) / Dx;
        V1(SS, i) = (CC1 * V1(U, i - 1) + CC2 * V1(U, i) + CC3 * V1(U, i + 1)) / (Dx * Dx);
    }
    V1(S, n) = -(-3.5 * V1(U, n) + 4.0 * V1(U, n - 1) - 0.5 * V1(U, n - 2)) / Dx;
    V1(SS, n) = (34.0 / 3.0 * V1(U, n) - 83.0 / 4.0 * V1(U, n - 1) + 10.0 * V1(U, n - 2) - 7.0 / 12.0 * V1(U, n - 3)) /
                (Dx * Dx);
}

/* ASSIGN_CCD_UD_TWIN_SRC (SCHEME_COMPACT.F90:439-546), M-DRPCCD sources */
static void assign_ccd_ud_src(int n, double Dx, const double *U, double *S, double *SS, double *SD, double *SSD) {
    const double C1 = -1.93629468537011, C2 = 1.99758937102742, C3 = -0.061294685370110;
    const double C1D = 0.061294685370110, C2D = -1.99758937102742, C3D = 1.93629468537011;
    const double CC1 = 3.0, CC2 = -6.0, CC3 = 3.0;
    const double s1 = (-3.5 * V1(U, 1) + 4.0 * V1(U, 2) - 0.5 * V1(U, 3)) / Dx;
    const double ss1 = (34.0 / 3.0 * V1(U, 1) - 83.0 / 4.0 * V1(U, 2) + 10.0 * V1(U, 3) - 7.0 / 12.0 * V1(U, 4)) / (Dx * Dx);
    const double sn = -(-3.5 * V1(U, n) + 4.0 * V1(U, n - 1) - 0.5 * V1(U, n - 2)) / Dx;
    const double ssn = (34.0 / 3.0 * V1(U, n) - 83.0 / 4.0 * V1(U, n - 1) + 10.0 * V1(U, n - 2) - 7.0 / 12.0 * V1(U, n - 3)) /
                       (Dx * Dx);
    V1(S, 1) = s1; V1(SS, 1) = ss1;
    for (int i = 2; i <= n - 1; ++i) {
        V1(S, i) = (C1 * V1(U, i - 1) + C2 * V1(U, i) + C3 * V1(U, i + 1)) / Dx;
        V1(SS, i) = (CC1 * V1(U, i - 1) + CC2 * V1(U, i) + CC3 * V1(U, i + 1)) / (Dx * Dx);
    }
    V1(S, n) = sn; V1(SS, n) = ssn;
    V1(SD, 1) = s1; V1(SSD, 1) = ss1;
    for (int i = 2; i <= n - 1; ++i) {
        V1(SD, i) = (C1D * V1(U, i - 1) + C2D * V1(U, i) + C3D * V1(U, i + 1)) / Dx;
        V1(SSD, i) = (CC1 * V1(U, i - 1) + CC2 * V1(U, i) + CC3 * V1(U, i + 1)) / (Dx * Dx);
    }
    V1(SD, n) = sn; V1(SSD, n) = ssn;
}

/* TWIN_BKS (SCHEME_COMPACT.F90:548-593) */
static void twin_bks(int n, const double *A, const double *B, const double *AA, const double *BB, double *S, double *SS) {
    for (int i = 2; i <= n; ++i) {
        V1(S, i) = V1(S, i) - (M2(A, 1, i) * V1(S, i - 1) + M2(B, 1, i) * V1(SS, i - 1));
        V1(SS, i) = V1(SS, i) - (M2(AA, 1, i) * V1(S, i - 1) + M2(BB, 1, i) * V1(SS, i - 1));
    }
    double Den = M2(A, 2, n) * M2(BB, 2, n) - M2(AA, 2, n) * M2(B, 2, n);
    double SolS = -(M2(B, 2, n) * V1(SS, n) - M2(BB, 2, n) * V1(S, n)) / Den;
    double SolSS = (M2(A, 2, n) * V1(SS, n) - M2(AA, 2, n) * V1(S, n)) / Den;
    V1(S, n) = SolS;
    V1(SS, n) = SolSS;
    for (int i = n - 1; i >= 1; --i) {
        V1(S, i) = V1(S, i) - (M2(A, 3, i) * SolS + M2(B, 3, i) * SolSS);
        V1(SS, i) = V1(SS, i) - (M2(AA, 3, i) * SolS + M2(BB, 3, i) * SolSS);
        Den = M2(A, 2, i) * M2(BB, 2, i) - M2(AA, 2, i) * M2(B, 2, i);
        SolS = -(M2(B, 2, i) * V1(SS, i) - M2(BB, 2, i) * V1(S, i)) / Den;
        SolSS = (M2(A, 2, i) * V1(SS, i) - M2(AA, 2, i) * V1(S, i)) / Den;
        V1(S, i) = SolS;
        V1(SS, i) = SolSS;
    }
}

/* CCD (SCHEME_COMPACT.F90:32-52) of the n line values f (stride st) into
 * fx, fxx (same stride) */
static void ccd_line(m3_line *L, double Dx, const double *f, double *fx, double *fxx, long st) {
    const int n = L->n;
    assign_ccd(n, Dx, L->A, L->B, L->AA, L->BB);
    twin_dec(n, L->A, L->B, L->AA, L->BB);
    for (int i = 1; i <= n; ++i) V1(L->U, i) = f[(long)(i - 1) * st];
    assign_ccd_src(n, Dx, L->U, L->S, L->SS);
    twin_bks(n, L->A, L->B, L->AA, L->BB, L->S, L->SS);
    for (int i = 1; i <= n; ++i) {
        fx[(long)(i - 1) * st] = V1(L->S, i);
        fxx[(long)(i - 1) * st] = V1(L->SS, i);
    }
}

/* UCCD (SCHEME_COMPACT.F90:1-30): upwind / downwind systems, fx by the sign
 * of u */
static void uccd_line(m3_line *L, double Dx, const double *u, const double *f, double *fx, long st) {
    const int n = L->n;
    assign_ccd_ud(n, Dx, L->A, L->B, L->AA, L->BB, L->AD, L->BD, L->AAD, L->BBD);
    twin_dec(n, L->A, L->B, L->AA, L->BB);
    twin_dec(n, L->AD, L->BD, L->AAD, L->BBD);
    for (int i = 1; i <= n; ++i) V1(L->F, i) = f[(long)(i - 1) * st];
    assign_ccd_ud_src(n, Dx, L->F, L->S, L->SS, L->SD, L->SSD);
    twin_bks(n, L->A, L->B, L->AA, L->BB, L->S, L->SS);
    twin_bks(n, L->AD, L->BD, L->AAD, L->BBD, L->SD, L->SSD);
    for (int i = 1; i <= n; ++i) fx[(long)(i - 1) * st] = u[(long)(i - 1) * st] >= 0.0 ? V1(L->S, i) : V1(L->SD, i);
}
#undef M2
#undef V1

/* CCD of every interior line of one direction: out1 = d/dx, out2 = d2/dx2 */
static void ccd_dir(m3_state *s, int dir, const double *T, double *o1, double *o2) {
    const int n[3] = {s->nx, s->ny, s->nz};
    const double h[3] = {s->pr.dx, s->pr.dy, s->pr.dz};
    const long st[3] = {1, s->sx, s->sxy};
    m3_line L = line_alloc(n[dir]);
    if (dir == 0) {
        for (int k = 1; k <= s->nz; ++k)
            for (int j = 1; j <= s->ny; ++j) {
                const long o = IX(s, 1, j, k);
                ccd_line(&L, h[0], T + o, o1 + o, o2 + o, st[0]);
            }
    } else if (dir == 1) {
        for (int k = 1; k <= s->nz; ++k)
            for (int i = 1; i <= s->nx; ++i) {
                const long o = IX(s, i, 1, k);
                ccd_line(&L, h[1], T + o, o1 + o, o2 + o, st[1]);
            }
    } else {
        for (int j = 1; j <= s->ny; ++j)   /* the reference runs j to NODE_Z here */
            for (int i = 1; i <= s->nx; ++i) {
                const long o = IX(s, i, j, 1);
                ccd_line(&L, h[2], T + o, o1 + o, o2 + o, st[2]);
            }
    }
    line_free(&L);
}

/* ------------------------------------------------------ FIND_CURVATURE.F90 */

/* LS_CCD (FIND_CURVATURE.F90:99-199): CCD first / second derivatives, the
 * cross derivatives, curvature; normals = first derivatives */
static void ls_ccd(m3_state *s, const double *T) {
    heavy_f(s, T);
    ccd_dir(s, 0, T, s->a, s->ff);
    ccd_dir(s, 1, T, s->b, s->gg);
    ccd_dir(s, 2, T, s->c, s->hh);
    ccd_dir(s, 0, s->c, s->up, s->um);   /* fz_x */
    ccd_dir(s, 1, s->a, s->vp, s->vm);   /* fx_y */
    ccd_dir(s, 2, s->b, s->wp, s->wm);   /* fy_z */
    double *K = F_(s, M3_CURV), *NX_ = F_(s, M3_NORMAL_X), *NY_ = F_(s, M3_NORMAL_Y), *NZ_ = F_(s, M3_NORMAL_Z);
    for (int k = 1; k <= s->nz; ++k)
        for (int j = 1; j <= s->ny; ++j)
            for (int i = 1; i <= s->nx; ++i) {
                const long id = IX(s, i, j, k);
                const double fx = s->a[id], fy = s->b[id], fz = s->c[id];
                const double fxx = s->ff[id], fyy = s->gg[id], fzz = s->hh[id];
                const double fxy = s->vp[id], fxz = s->up[id], fyz = s->wp[id];
                K[id] = (fxx * (sq(fy) + sq(fz)) + fyy * (sq(fx) + sq(fz)) + fzz * (sq(fx) + sq(fy)) -
                         2.0 * (fxy * fx * fy + fxz * fx * fz + fyz * fy * fz)) /
                        pow(sq(fx) + sq(fy) + sq(fz) + M3EPS, 1.5);
                NX_[id] = fx;
                NY_[id] = fy;
                NZ_[id] = fz;
            }
    bc3d(s, K);
    bc3d(s, NX_);
    bc3d(s, NY_);
    bc3d(s, NZ_);
}

/* ---------------------------------------------------------------- AMURHO.F90 */

/* AMURHO (AMURHO.F90:1-46) with INTERFACE_METHOD = 0: the second pair of
 * assignments is the one that stays, rho = rho_12 H + (1 - H) (phi > 0 is
 * the bubble) */
static void amurho(m3_state *s) {
    heavy_f(s, F_(s, M3_PHI));
    double *H = F_(s, M3_HEAVY), *R = F_(s, M3_RHO), *M = F_(s, M3_AMU);
    for (int k = 1; k <= s->nz; ++k)
        for (int j = 1; j <= s->ny; ++j)
            for (int i = 1; i <= s->nx; ++i) {
                const long id = IX(s, i, j, k);
                R[id] = H[id] + s->pr.ratio_rho * (1.0 - H[id]);
                M[id] = H[id] + s->pr.ratio_amu * (1.0 - H[id]);
                R[id] = s->pr.ratio_rho * H[id] + (1.0 - H[id]);
                M[id] = s->pr.ratio_amu * H[id] + (1.0 - H[id]);
            }
    bc3d(s, R);
    bc3d(s, M);
}

/* ------------------------------------------------------------- LS_SOLVER.F90 */

/* CELL_FACE_VELOCITY (LS_SOLVER.F90:93-116): cell velocities UH, VH, WH */
static void cell_face_velocity(m3_state *s) {
    const double *U = F_(s, M3_U), *V = F_(s, M3_V), *W = F_(s, M3_W);
    double *UH = F_(s, M3_UH), *VH = F_(s, M3_VH), *WH = F_(s, M3_WH);
    for (int k = 1; k <= s->nz; ++k)
        for (int j = 1; j <= s->ny; ++j)
            for (int i = 1; i <= s->nx; ++i) {
                const long id = IX(s, i, j, k);
                UH[id] = 0.5 * (U[id] + U[id - 1]);
                VH[id] = 0.5 * (V[id] + V[id - s->sx]);
                WH[id] = 0.5 * (W[id] + W[id - s->sxy]);
            }
    bc3d(s, UH);
    bc3d(s, VH);
    bc3d(s, WH);
}

/* LS_sorce with LS_SOLVER = 10 (LS_SOLVER.F90:118-228): UCCD derivatives
 * along each direction, s = -(F UH + G VH + H WH) */
static void ls_source(m3_state *s, double *S) {
    const double *P = F_(s, M3_PHI), *UH = F_(s, M3_UH), *VH = F_(s, M3_VH), *WH = F_(s, M3_WH);
    {
        m3_line L = line_alloc(s->nx);
        for (int k = 1; k <= s->nz; ++k)
            for (int j = 1; j <= s->ny; ++j) {
                const long o = IX(s, 1, j, k);
                uccd_line(&L, s->pr.dx, UH + o, P + o, s->ff + o, 1);
            }
        line_free(&L);
    }
    {
        m3_line L = line_alloc(s->ny);
        for (int k = 1; k <= s->nz; ++k)
            for (int i = 1; i <= s->nx; ++i) {
                const long o = IX(s, i, 1, k);
                uccd_line(&L, s->pr.dy, VH + o, P + o, s->gg + o, s->sx);
            }
        line_free(&L);
    }
    {
        m3_line L = line_alloc(s->nz);
        for (int j = 1; j <= s->ny; ++j)
            for (int i = 1; i <= s->nx; ++i) {
                const long o = IX(s, i, j, 1);
                uccd_line(&L, s->pr.dz, WH + o, P + o, s->hh + o, s->sxy);
            }
        line_free(&L);
    }
    for (int k = 1; k <= s->nz; ++k)
        for (int j = 1; j <= s->ny; ++j)
            for (int i = 1; i <= s->nx; ++i) {
                const long id = IX(s, i, j, k);
                S[id] = -(s->ff[id] * UH[id] + s->gg[id] * VH[id] + s->hh[id] * WH[id]);
            }
}

/* LS_RK3_SOLVER (LS_SOLVER.F90:39-91) */
static void ls_rk3_solver(m3_state *s) {
    double *P = F_(s, M3_PHI), *S0 = F_(s, M3_S0);
    const double dt = s->pr.dt;
    cell_face_velocity(s);
    for (int stage = 0; stage < 3; ++stage) {
        double *out = stage == 0 ? S0 : (stage == 1 ? s->s1 : s->s2);
        ls_source(s, out);
        for (int k = 1; k <= s->nz; ++k)
            for (int j = 1; j <= s->ny; ++j)
                for (int i = 1; i <= s->nx; ++i) {
                    const long id = IX(s, i, j, k);
                    if (stage == 0) P[id] = P[id] + dt * S0[id];
                    else if (stage == 1) P[id] = P[id] + dt / 4.0 * (-3.0 * S0[id] + s->s1[id]);
                    else P[id] = P[id] + dt / 12.0 * (-S0[id] - s->s1[id] + 8.0 * s->s2[id]);
                }
        bc3d(s, P);
    }
}

/* MPLS(im) (LS_SOLVER.F90:257-321): five corrections of PHI along
 * delta |grad phi| with the dt-scaled multiplier */
static void mpls(m3_state *s, double im) {
    double *P = F_(s, M3_PHI), *H = F_(s, M3_HEAVY), *D = F_(s, M3_DELTA), *R = F_(s, M3_RHO);
    double *G = F_(s, M3_GRAD);
    const double *NX_ = F_(s, M3_NORMAL_X), *NY_ = F_(s, M3_NORMAL_Y), *NZ_ = F_(s, M3_NORMAL_Z);
    for (int it = 1; it <= 5; ++it) {
        amurho(s);
        ls_ccd(s, P);
        double vol = 0.0, mass = 0.0, fs = 0.0;
        for (int k = 1; k <= s->nz; ++k)
            for (int j = 1; j <= s->ny; ++j)
                for (int i = 1; i <= s->nx; ++i) {
                    const long id = IX(s, i, j, k);
                    s->a[id] = 1.0;
                    G[id] = sqrt(sq(NX_[id]) + sq(NY_[id]) + sq(NZ_[id]));
                }
        for (int k = 1; k <= s->nz; ++k)
            for (int j = 1; j <= s->ny; ++j)
                for (int i = 1; i <= s->nx; ++i) {
                    const long id = IX(s, i, j, k);
                    vol = vol + H[id];
                    mass = mass + H[id] * R[id];
                    fs = fs + (2.0 * (s->pr.ratio_rho - 1.0) * H[id] + 1.0) * sq(D[id]) * G[id] * s->a[id];
                }
        s->vol_ls = vol;
        s->mass_ls = mass;
        fs = (im - mass) / (s->pr.dt * fs);
        for (int k = 1; k <= s->nz; ++k)
            for (int j = 1; j <= s->ny; ++j)
                for (int i = 1; i <= s->nx; ++i) {
                    const long id = IX(s, i, j, k);
                    P[id] = P[id] + fs * s->a[id] * D[id] * G[id] * s->pr.dt;
                }
    }
}

static void copy_all(m3_state *s, double *dst, const double *src) {
    memcpy(dst, src, sizeof(double) * (size_t)s->ntot);
}

/* LS_MAINTAIN with INTERFACE_METHOD = 0 (LS_MAINTAIN.F90:1-20) */
static void ls_maintain(m3_state *s) {
    if (s->pr.iter % 5 == 0) ls_redistance(s, F_(s, M3_PHI));
}

/* SOLVE_LS (LS_SOLVER.F90:1-37): each bubble's level set advected,
 * maintained and corrected to its own mass, then merged by max and
 * redistanced */
static void solve_ls(m3_state *s) {
    double *P = F_(s, M3_PHI), *P1 = F_(s, M3_PHI1), *P2 = F_(s, M3_PHI2);
    copy_all(s, P, P1);
    ls_rk3_solver(s);
    ls_maintain(s);
    mpls(s, s->pr.b1_im);
    copy_all(s, P1, P);
    copy_all(s, P, P2);
    ls_rk3_solver(s);
    ls_maintain(s);
    mpls(s, s->pr.b2_im);
    copy_all(s, P2, P);
    for (int k = 1; k <= s->nz; ++k)
        for (int j = 1; j <= s->ny; ++j)
            for (int i = 1; i <= s->nx; ++i) {
                const long id = IX(s, i, j, k);
                P[id] = fmax(P1[id], P2[id]);
            }
    bc3d(s, P);
    ls_redistance(s, P);
}

/* --------------------------------------------------- SECOND_ORDER_UPWIND.F90 */

/* VELA / VELB / VELC (SECOND_ORDER_UPWIND.F90:10-256) with ST_FORCE = 0 or 1:
 * second-order upwind advection of the old velocity by the current one,
 * viscous term, CSF surface tension, gravity on w */
static void second_order_upwind(m3_state *s) {
    const double dx = s->pr.dx, dy = s->pr.dy, dz = s->pr.dz;
    const double DX2 = dx * dx, DY2 = dy * dy, DZ2 = dz * dz;
    const double RE2X = s->pr.re * DX2, RE2Y = s->pr.re * DY2, RE2Z = s->pr.re * DZ2;
    const double FR2 = s->pr.fr * s->pr.fr, we = s->pr.we, w = s->pr.interface_width;
    const double *U = F_(s, M3_U), *V = F_(s, M3_V), *W = F_(s, M3_W);
    const double *UO = F_(s, M3_U_OLD), *VO = F_(s, M3_V_OLD), *WO = F_(s, M3_W_OLD);
    const double *MO = F_(s, M3_AMU_OLD), *RO = F_(s, M3_RHO_OLD), *P = F_(s, M3_PHI), *K = F_(s, M3_CURV);
    const double *NX_ = F_(s, M3_NORMAL_X), *NY_ = F_(s, M3_NORMAL_Y), *NZ_ = F_(s, M3_NORMAL_Z);
    double *US = F_(s, M3_US), *VS = F_(s, M3_VS), *WS = F_(s, M3_WS);
    const long sx = s->sx, sxy = s->sxy;
    for (int k = 1; k <= s->nz; ++k)
        for (int j = 1; j <= s->ny; ++j)
            for (int i = 1; i <= s->nx - 1; ++i) {
                const long id = IX(s, i, j, k);
                const double uu = U[id];
                const double vv = 0.25 * (V[id + 1] + V[id] + V[id + 1 - sx] + V[id - sx]);
                const double ww = 0.25 * (W[id + 1] + W[id] + W[id + 1 - sxy] + W[id - sxy]);
                const double up = 0.5 * (uu + fabs(uu)), um = 0.5 * (uu - fabs(uu));
                const double vp = 0.5 * (vv + fabs(vv)), vm = 0.5 * (vv - fabs(vv));
                const double wp = 0.5 * (ww + fabs(ww)), wm = 0.5 * (ww - fabs(ww));
                const double cm = MO[id] + MO[id + 1], cr = RO[id] + RO[id + 1], cc = cm / cr;
                double detf;
                heav(0.5 * (P[id] + P[id + 1]), w, &detf);
                const double curvk = 0.5 * (K[id] + K[id + 1]), fin = 0.5 * (NX_[id] + NX_[id + 1]);
                const double a1 = 0.5 / dx * (up * (3.0 * UO[id] - 4.0 * UO[id - 1] + UO[id - 2]) +
                                              um * (-UO[id + 2] + 4.0 * UO[id + 1] - 3.0 * UO[id]));
                const double a2 = 0.5 / dy * (vp * (3.0 * UO[id] - 4.0 * UO[id - sx] + UO[id - 2 * sx]) +
                                              vm * (-UO[id + 2 * sx] + 4.0 * UO[id + sx] - 3.0 * UO[id]));
                const double a3 = 0.5 / dz * (wp * (3.0 * UO[id] - 4.0 * UO[id - sxy] + UO[id - 2 * sxy]) +
                                              wm * (-UO[id + 2 * sxy] + 4.0 * UO[id + sxy] - 3.0 * UO[id]));
                const double a4 = 1.0 / RE2X * cc * (UO[id + 1] - 2.0 * UO[id] + UO[id - 1]);
                const double a5 = 1.0 / RE2Y * cc * (UO[id + sx] - 2.0 * UO[id] + UO[id - sx]);
                const double a6 = 1.0 / RE2Z * cc * (UO[id + sxy] - 2.0 * UO[id] + UO[id - sxy]);
                const double a7 = s->pr.st_force == 0 ? 0.0 : 2.0 / (cr * we) * curvk * detf * fin;
                US[id] = a1 + a2 + a3 - a4 - a5 - a6 + a7;
            }
    for (int k = 1; k <= s->nz; ++k)
        for (int j = 1; j <= s->ny - 1; ++j)
            for (int i = 1; i <= s->nx; ++i) {
                const long id = IX(s, i, j, k);
                const double uu = 0.25 * (U[id + sx] + U[id] + U[id - 1 + sx] + U[id - 1]);
                const double vv = V[id];
                const double ww = 0.25 * (W[id + sx] + W[id] + W[id + sx - sxy] + W[id - sxy]);
                const double up = 0.5 * (uu + fabs(uu)), um = 0.5 * (uu - fabs(uu));
                const double vp = 0.5 * (vv + fabs(vv)), vm = 0.5 * (vv - fabs(vv));
                const double wp = 0.5 * (ww + fabs(ww)), wm = 0.5 * (ww - fabs(ww));
                const double cm = MO[id] + MO[id + sx], cr = RO[id] + RO[id + sx], cc = cm / cr;
                double detf;
                heav(0.5 * (P[id] + P[id + sx]), w, &detf);
                const double curvk = 0.5 * (K[id] + K[id + sx]), fin = 0.5 * (NY_[id] + NY_[id + sx]);
                const double b1 = 0.5 / dx * (up * (3.0 * VO[id] - 4.0 * VO[id - 1] + VO[id - 2]) +
                                              um * (-VO[id + 2] + 4.0 * VO[id + 1] - 3.0 * VO[id]));
                const double b2 = 0.5 / dy * (vp * (3.0 * VO[id] - 4.0 * VO[id - sx] + VO[id - 2 * sx]) +
                                              vm * (-VO[id + 2 * sx] + 4.0 * VO[id + sx] - 3.0 * VO[id]));
                const double b3 = 0.5 / dz * (wp * (3.0 * VO[id] - 4.0 * VO[id - sxy] + VO[id - 2 * sxy]) +
                                              wm * (-VO[id + 2 * sxy] + 4.0 * VO[id + sxy] - 3.0 * VO[id]));
                const double b4 = 1.0 / RE2X * cc * (VO[id + 1] - 2.0 * VO[id] + VO[id - 1]);
                const double b5 = 1.0 / RE2Y * cc * (VO[id + sx] - 2.0 * VO[id] + VO[id - sx]);
                const double b6 = 1.0 / RE2Z * cc * (VO[id + sxy] - 2.0 * VO[id] + VO[id - sxy]);
                const double b7 = s->pr.st_force == 0 ? 0.0 : 2.0 / (cr * we) * curvk * detf * fin;
                VS[id] = b1 + b2 + b3 - b4 - b5 - b6 + b7;
            }
    for (int k = 1; k <= s->nz - 1; ++k)
        for (int j = 1; j <= s->ny; ++j)
            for (int i = 1; i <= s->nx; ++i) {
                const long id = IX(s, i, j, k);
                const double uu = 0.25 * (U[id + sxy] + U[id] + U[id - 1 + sxy] + U[id - 1]);
                const double vv = 0.25 * (V[id + sxy] + V[id] + V[id - sx + sxy] + V[id - sx]);
                const double ww = W[id];
                const double up = 0.5 * (uu + fabs(uu)), um = 0.5 * (uu - fabs(uu));
                const double vp = 0.5 * (vv + fabs(vv)), vm = 0.5 * (vv - fabs(vv));
                const double wp = 0.5 * (ww + fabs(ww)), wm = 0.5 * (ww - fabs(ww));
                const double cm = MO[id] + MO[id + sxy], cr = RO[id] + RO[id + sxy], cc = cm / cr;
                double detf;
                heav(0.5 * (P[id] + P[id + sxy]), w, &detf);
                const double curvk = 0.5 * (K[id] + K[id + sxy]), fin = 0.5 * (NZ_[id] + NZ_[id + sxy]);
                const double c1 = 0.5 / dx * (up * (3.0 * WO[id] - 4.0 * WO[id - 1] + WO[id - 2]) +
                                              um * (-WO[id + 2] + 4.0 * WO[id + 1] - 3.0 * WO[id]));
                const double c2 = 0.5 / dy * (vp * (3.0 * WO[id] - 4.0 * WO[id - sx] + WO[id - 2 * sx]) +
                                              vm * (-WO[id + 2 * sx] + 4.0 * WO[id + sx] - 3.0 * WO[id]));
                const double c3 = 0.5 / dz * (wp * (3.0 * WO[id] - 4.0 * WO[id - sxy] + WO[id - 2 * sxy]) +
                                              wm * (-WO[id + 2 * sxy] + 4.0 * WO[id + sxy] - 3.0 * WO[id]));
                const double c4 = 1.0 / RE2X * cc * (WO[id + 1] - 2.0 * WO[id] + WO[id - 1]);
                const double c5 = 1.0 / RE2Y * cc * (WO[id + sx] - 2.0 * WO[id] + WO[id - sx]);
                const double c6 = 1.0 / RE2Z * cc * (WO[id + sxy] - 2.0 * WO[id] + WO[id - sxy]);
                const double c7 = s->pr.st_force == 0 ? 0.0 : 2.0 / (cr * we) * curvk * detf * fin;
                WS[id] = c1 + c2 + c3 - c4 - c5 - c6 + c7;
                if (s->pr.g_force == 1) WS[id] = WS[id] + 1.0 / FR2;
            }
}

/* Adams_Bashforth (SOLVE_NS.F90:147-166) */
static void adams_bashforth(m3_state *s) {
    const double *US = F_(s, M3_US), *VS = F_(s, M3_VS), *WS = F_(s, M3_WS);
    const double *USO = F_(s, M3_US_OLD), *VSO = F_(s, M3_VS_OLD), *WSO = F_(s, M3_WS_OLD);
    double *U1 = F_(s, M3_U1), *V1_ = F_(s, M3_V1), *W1 = F_(s, M3_W1);
    for (int k = 1; k <= s->nz; ++k)
        for (int j = 1; j <= s->ny; ++j)
            for (int i = 1; i <= s->nx; ++i) {
                const long id = IX(s, i, j, k);
                U1[id] = 1.5 * US[id] - 0.5 * USO[id];
                V1_[id] = 1.5 * VS[id] - 0.5 * VSO[id];
                W1[id] = 1.5 * WS[id] - 0.5 * WSO[id];
            }
}

/* RHSP (SOLVE_PRESSURE_POISSON.F90:12-58): the wall-face sources zeroed,
 * RHS = div(u_old / dt - U1).  The reference zeroes U1 on the x walls for
 * K = 0..NODE_X+1 only (its loop bound names NODE_X); restated as written
 * (kept inside the array).  In a run nothing else writes U1 there (VELA
 * leaves US(NODE_X) at zero), so the planes beyond hold zero anyway. */
static void rhsp(m3_state *s) {
    double *U1 = F_(s, M3_U1), *V1_ = F_(s, M3_V1), *W1 = F_(s, M3_W1), *RHS = F_(s, M3_RHS);
    const double *UO = F_(s, M3_U_OLD), *VO = F_(s, M3_V_OLD), *WO = F_(s, M3_W_OLD);
    const int NX = s->nx, NY = s->ny, NZ = s->nz;
    const int kx = NX + 1 < NZ + 3 ? NX + 1 : NZ + 3;
    for (int k = 0; k <= kx; ++k)
        for (int j = 0; j <= NY + 1; ++j) {
            U1[IX(s, 0, j, k)] = 0.0;
            U1[IX(s, NX, j, k)] = 0.0;
        }
    for (int k = 0; k <= NZ + 1; ++k)
        for (int i = 0; i <= NX + 1; ++i) {
            V1_[IX(s, i, 0, k)] = 0.0;
            V1_[IX(s, i, NY, k)] = 0.0;
        }
    for (int j = 0; j <= NY + 1; ++j)
        for (int i = 0; i <= NX + 1; ++i) {
            W1[IX(s, i, j, 0)] = 0.0;
            W1[IX(s, i, j, NZ)] = 0.0;
        }
    const double dt = s->pr.dt;
    for (int k = 1; k <= NZ; ++k)
        for (int j = 1; j <= NY; ++j)
            for (int i = 1; i <= NX; ++i) {
                const long id = IX(s, i, j, k);
                RHS[id] = (UO[id] / dt - UO[id - 1] / dt - U1[id] + U1[id - 1]) / s->pr.dx +
                          (VO[id] / dt - VO[id - s->sx] / dt - V1_[id] + V1_[id - s->sx]) / s->pr.dy +
                          (WO[id] / dt - WO[id - s->sxy] / dt - W1[id] + W1[id - s->sxy]) / s->pr.dz;
            }
}

/* the coefficients of PRESS_POISSON for cell (i,j,k), walls folded */
static void press_coef(const m3_state *s, int i, int j, int k, double *cl, double *cr, double *cf, double *cb,
                       double *cu, double *cd, double *cc) {
    const double *R = s->f[M3_RHO];
    const long id = IX(s, i, j, k);
    const double DX2 = s->pr.dx * s->pr.dx, DY2 = s->pr.dy * s->pr.dy, DZ2 = s->pr.dz * s->pr.dz;
    const double RR = 0.5 * (R[id] + R[id + 1]), RL = 0.5 * (R[id] + R[id - 1]);
    const double RF = 0.5 * (R[id] + R[id + s->sx]), RB = 0.5 * (R[id] + R[id - s->sx]);
    const double RU = 0.5 * (R[id] + R[id + s->sxy]), RD = 0.5 * (R[id] + R[id - s->sxy]);
    double CL = 1.0 / (DX2 * RL), CR = 1.0 / (DX2 * RR), CF = 1.0 / (DY2 * RF), CB = 1.0 / (DY2 * RB);
    double CU = 1.0 / (DZ2 * RU), CD = 1.0 / (DZ2 * RD);
    double CC = -(CL + CR + CF + CB + CU + CD);
    if (i == 1) { CC = CC + CL; CL = 0.0; }
    if (i == s->nx) { CC = CC + CR; CR = 0.0; }
    if (j == 1) { CC = CC + CB; CB = 0.0; }
    if (j == s->ny) { CC = CC + CF; CF = 0.0; }
    if (k == 1) { CC = CC + CD; CD = 0.0; }
    if (k == s->nz) { CC = CC + CU; CU = 0.0; }
    *cl = CL; *cr = CR; *cf = CF; *cb = CB; *cu = CU; *cd = CD; *cc = CC;
}

/* PRESS_POISSON (SOLVE_PRESSURE_POISSON.F90:60-180): in-place SOR (run in
 * its sequential order), mean removal, until max |dP| < TOL_P */
int m3_press_poisson(m3_state *s) {
    double *P = F_(s, M3_P), *RHS = F_(s, M3_RHS), *PO = s->p_old;
    const double om = s->pr.press_omega;
    const double ncell = (double)(s->nx * s->ny * s->nz);
    int it = 0;
    for (;;) {
        double linf = 0.0;
        ++it;
        for (int k = 1; k <= s->nz; ++k)
            for (int j = 1; j <= s->ny; ++j)
                for (int i = 1; i <= s->nx; ++i) {
                    const long id = IX(s, i, j, k);
                    double CL, CR, CF, CB, CU, CD, CC;
                    PO[id] = P[id];
                    press_coef(s, i, j, k, &CL, &CR, &CF, &CB, &CU, &CD, &CC);
                    const double pold = P[id];
                    double pcal = RHS[id] - CL * P[id - 1] - CR * P[id + 1] - CB * P[id - s->sx] - CF * P[id + s->sx] -
                                  CD * P[id - s->sxy] - CU * P[id + s->sxy];
                    pcal = pcal / CC;
                    P[id] = om * pcal + (1.0 - om) * pold;
                }
        double sump = 0.0;
        for (int k = 1; k <= s->nz; ++k)
            for (int j = 1; j <= s->ny; ++j)
                for (int i = 1; i <= s->nx; ++i) sump = sump + P[IX(s, i, j, k)];
        sump = sump / ncell;
        for (int k = 1; k <= s->nz; ++k)
            for (int j = 1; j <= s->ny; ++j)
                for (int i = 1; i <= s->nx; ++i) {
                    const long id = IX(s, i, j, k);
                    P[id] = P[id] - sump;
                    linf = fmax(linf, fabs(P[id] - PO[id]));
                }
        if (linf < s->pr.tol_p) break;
        if (it > s->pr.press_max_it) { s->ppe_iters += it; return 3; }
    }
    s->ppe_iters += it;
    return 0;
}

/* the operator of PRESS_POISSON applied to P over the interior */
static void press_apply(const m3_state *s, const double *P, double *out) {
    for (int k = 1; k <= s->nz; ++k)
        for (int j = 1; j <= s->ny; ++j)
            for (int i = 1; i <= s->nx; ++i) {
                const long id = IX(s, i, j, k);
                double CL, CR, CF, CB, CU, CD, CC;
                press_coef(s, i, j, k, &CL, &CR, &CF, &CB, &CU, &CD, &CC);
                out[id] = CC * P[id] + CL * P[id - 1] + CR * P[id + 1] + CB * P[id - s->sx] + CF * P[id + s->sx] +
                          CD * P[id - s->sxy] + CU * P[id + s->sxy];
            }
}

/* Not in the reference: the fixed point of PRESS_POISSON solved directly,
 * A P = RHS - mean(RHS), mean(P) = 0, by the band Cholesky of the pinned
 * operator (cell (1,1,1)) and two steps of refinement.  The ghosts of P stay
 * as they are (the reference never sets them; its operator multiplies them by
 * zero). */
int m3_press_exact(m3_state *s) {
    const int NX = s->nx, NY = s->ny, NZ = s->nz;
    const long n = (long)NX * NY * NZ, b = (long)NX * NY, w = b + 1;
    double *P = F_(s, M3_P), *RHS = F_(s, M3_RHS);
    if (!s->band_valid) {
        if (orc_band_alloc(&s->band, n, b) != 0) return 1;
        double *L = s->band.L;
        for (int k = 1; k <= NZ; ++k)
            for (int j = 1; j <= NY; ++j)
                for (int i = 1; i <= NX; ++i) {
                    const long r = (i - 1) + (long)NX * ((j - 1) + (long)NY * (k - 1));
                    double CL, CR, CF, CB, CU, CD, CC;
                    press_coef(s, i, j, k, &CL, &CR, &CF, &CB, &CU, &CD, &CC);
                    double *row = L + r * w + b;
                    row[0] = -CC;
                    if (i > 1) row[-1] = -CL;
                    if (j > 1) row[-(long)NX] = -CB;
                    if (k > 1) row[-b] = -CD;
                }
        L[b] = 1.0;
        L[1 * w + b - 1] = 0.0;
        L[(long)NX * w + b - NX] = 0.0;
        L[b * w] = 0.0;
        if (orc_band_cholesky(&s->band) != 0) return 6;
        s->band_valid = 1;
    }
    double *f = zalloc(s->ntot), *Ap = zalloc(s->ntot), *x = zalloc(n);
    double mean = 0.0;
    for (int k = 1; k <= NZ; ++k)
        for (int j = 1; j <= NY; ++j)
            for (int i = 1; i <= NX; ++i) mean += RHS[IX(s, i, j, k)];
    mean /= (double)n;
    for (int k = 1; k <= NZ; ++k)
        for (int j = 1; j <= NY; ++j)
            for (int i = 1; i <= NX; ++i) {
                const long id = IX(s, i, j, k);
                f[id] = RHS[id] - mean;
                P[id] = 0.0;
            }
    for (int pass = 0; pass < 3; ++pass) {
        press_apply(s, P, Ap);
        for (int k = 1; k <= NZ; ++k)
            for (int j = 1; j <= NY; ++j)
                for (int i = 1; i <= NX; ++i) {
                    const long id = IX(s, i, j, k);
                    x[(i - 1) + (long)NX * ((j - 1) + (long)NY * (k - 1))] = Ap[id] - f[id];
                }
        orc_band_solve(&s->band, x);
        for (int k = 1; k <= NZ; ++k)
            for (int j = 1; j <= NY; ++j)
                for (int i = 1; i <= NX; ++i)
                    P[IX(s, i, j, k)] += x[(i - 1) + (long)NX * ((j - 1) + (long)NY * (k - 1))];
    }
    double pm = 0.0;
    for (int k = 1; k <= NZ; ++k)
        for (int j = 1; j <= NY; ++j)
            for (int i = 1; i <= NX; ++i) pm += P[IX(s, i, j, k)];
    pm /= (double)n;
    for (int k = 1; k <= NZ; ++k)
        for (int j = 1; j <= NY; ++j)
            for (int i = 1; i <= NX; ++i) P[IX(s, i, j, k)] -= pm;
    free(f); free(Ap); free(x);
    s->ppe_iters += 3;
    return 0;
}

/* SOLVE_MOMENTUM (SOLVE_NS.F90:102-145) */
static void solve_momentum(m3_state *s) {
    double *U = F_(s, M3_U), *V = F_(s, M3_V), *W = F_(s, M3_W);
    double *U1 = F_(s, M3_U1), *V1_ = F_(s, M3_V1), *W1 = F_(s, M3_W1);
    const double *UO = F_(s, M3_U_OLD), *VO = F_(s, M3_V_OLD), *WO = F_(s, M3_W_OLD);
    const double *P = F_(s, M3_P), *R = F_(s, M3_RHO);
    const double dt = s->pr.dt;
    for (int k = 1; k <= s->nz; ++k)
        for (int j = 1; j <= s->ny; ++j)
            for (int i = 1; i <= s->nx - 1; ++i) {
                const long id = IX(s, i, j, k);
                U1[id] = -U1[id] - 2.0 * (P[id + 1] - P[id]) / (s->pr.dx * (R[id] + R[id + 1]));
                U1[id] = UO[id] + U1[id] * dt;
                U[id] = U1[id];
            }
    for (int k = 1; k <= s->nz; ++k)
        for (int j = 1; j <= s->ny - 1; ++j)
            for (int i = 1; i <= s->nx; ++i) {
                const long id = IX(s, i, j, k);
                V1_[id] = -V1_[id] - 2.0 * (P[id + s->sx] - P[id]) / (s->pr.dy * (R[id] + R[id + s->sx]));
                V1_[id] = VO[id] + V1_[id] * dt;
                V[id] = V1_[id];
            }
    for (int k = 1; k <= s->nz - 1; ++k)
        for (int j = 1; j <= s->ny; ++j)
            for (int i = 1; i <= s->nx; ++i) {
                const long id = IX(s, i, j, k);
                W1[id] = -W1[id] - 2.0 * (P[id + s->sxy] - P[id]) / (s->pr.dz * (R[id] + R[id + s->sxy]));
                W1[id] = WO[id] + W1[id] * dt;
                W[id] = W1[id];
            }
}

/* FIND_ERROR with BTN = 2 (FIND_ERROR.F90:1-56): sqrt(mean((a - b)^2)) */
static double l2_err(const m3_state *s, const double *a, const double *b) {
    double r = 0.0;
    for (int k = 1; k <= s->nz; ++k)
        for (int j = 1; j <= s->ny; ++j)
            for (int i = 1; i <= s->nx; ++i) {
                const long id = IX(s, i, j, k);
                r = r + sq(a[id] - b[id]);
            }
    return sqrt(r / (double)(s->nz * s->ny * s->nx));
}

/* CHECK_CONVERGENCE (SOLVE_NS.F90:51-86): returns the added INITER */
static int check_convergence(m3_state *s) {
    const double *U = F_(s, M3_U), *V = F_(s, M3_V), *W = F_(s, M3_W);
    const double l2u = l2_err(s, U, s->u_tmp), l2v = l2_err(s, V, s->v_tmp), l2w = l2_err(s, W, s->w_tmp);
    s->l2sum = l2u + l2v + l2w;
    double div = 0.0;
    for (int k = 1; k <= s->nz; ++k)
        for (int j = 1; j <= s->ny; ++j)
            for (int i = 1; i <= s->nx; ++i) {
                const long id = IX(s, i, j, k);
                const double res = (U[id] - U[id - 1]) / s->pr.dx + (V[id] - V[id - s->sx]) / s->pr.dy +
                                   (W[id] - W[id - s->sxy]) / s->pr.dz;
                div = fmax(div, fabs(res));
            }
    s->div = div;
    return s->l2sum < s->pr.tol_m ? 5 : 0;
}

/* SOLVE_NS (SOLVE_NS.F90:1-49) */
static int solve_ns(m3_state *s, int exact_ppe) {
    int rc = 0;
    amurho(s);
    if (s->pr.st_force != 0) ls_ccd(s, F_(s, M3_PHI));
    s->band_valid = 0;   /* the operator of this step's RHO */
    s->ppe_iters = 0;
    int initer = 0;
    s->picard = 0;
    for (;;) {
        initer = initer + 1;
        copy_all(s, s->u_tmp, F_(s, M3_U));
        copy_all(s, s->v_tmp, F_(s, M3_V));
        copy_all(s, s->w_tmp, F_(s, M3_W));
        second_order_upwind(s);
        if (s->pr.rec_mass == 0) return 0;
        adams_bashforth(s);
        rhsp(s);
        int r = exact_ppe ? m3_press_exact(s) : m3_press_poisson(s);
        if (r) rc = r;
        solve_momentum(s);
        initer += check_convergence(s);
        bc_vel(s, F_(s, M3_U), 0);
        bc_vel(s, F_(s, M3_V), 1);
        bc_vel(s, F_(s, M3_W), 2);
        s->picard++;
        if (initer > 3) break;
    }
    return rc;
}

/* FIND_MASS (FIND_MASS.F90:1-162), the level-set part: volume and mass sums;
 * the first call (REC_MASS = 0) sets the bubble targets B1_iM, B2_iM from
 * PHI1 and PHI2 and the initial totals.  The per-bubble history files and
 * the VOF sums are output only. */
static void find_mass(m3_state *s) {
    amurho(s);
    const double *H = F_(s, M3_HEAVY), *R = F_(s, M3_RHO);
    double vol = 0.0, mass = 0.0;
    for (int k = 1; k <= s->nz; ++k)
        for (int j = 1; j <= s->ny; ++j)
            for (int i = 1; i <= s->nx; ++i) {
                const long id = IX(s, i, j, k);
                vol = vol + H[id];
                mass = mass + H[id] * R[id];
            }
    s->vol_ls = vol;
    s->mass_ls = mass;
    if (s->pr.rec_mass == 0) {
        double *P = F_(s, M3_PHI);
        copy_all(s, s->phi_tmp, P);
        double im[2];
        for (int q = 0; q < 2; ++q) {
            copy_all(s, P, F_(s, q == 0 ? M3_PHI1 : M3_PHI2));
            amurho(s);
            double m = 0.0;
            for (int k = 1; k <= s->nz; ++k)
                for (int j = 1; j <= s->ny; ++j)
                    for (int i = 1; i <= s->nx; ++i) {
                        const long id = IX(s, i, j, k);
                        m = m + H[id] * R[id];
                    }
            im[q] = m;
        }
        s->pr.b1_im = im[0];
        s->pr.b2_im = im[1];
        copy_all(s, P, s->phi_tmp);
        s->pr.imass_ls = mass;
        s->pr.ivol_ls = vol;
        s->pr.rec_mass = 1;
    }
}

/* STORE_DATA (STORE_DATA.F90:1-33) */
static void store_data(m3_state *s) {
    copy_all(s, F_(s, M3_PHI_OLD), F_(s, M3_PHI));
    copy_all(s, F_(s, M3_U_OLD), F_(s, M3_U));
    copy_all(s, F_(s, M3_V_OLD), F_(s, M3_V));
    copy_all(s, F_(s, M3_W_OLD), F_(s, M3_W));
    copy_all(s, F_(s, M3_US_OLD), F_(s, M3_US));
    copy_all(s, F_(s, M3_VS_OLD), F_(s, M3_VS));
    copy_all(s, F_(s, M3_WS_OLD), F_(s, M3_WS));
    copy_all(s, F_(s, M3_RHO_OLD), F_(s, M3_RHO));
    copy_all(s, F_(s, M3_AMU_OLD), F_(s, M3_AMU));
}

/* ---------------------------------------------------------- entry points */
int m3_bc3d_phi(m3_state *s) { bc3d(s, F_(s, M3_PHI)); return 0; }
int m3_amurho(m3_state *s) { amurho(s); return 0; }
int m3_heavy_f(m3_state *s) { heavy_f(s, F_(s, M3_PHI)); return 0; }
int m3_ls_ccd(m3_state *s) { ls_ccd(s, F_(s, M3_PHI)); return 0; }
int m3_cell_face_velocity(m3_state *s) { cell_face_velocity(s); return 0; }
int m3_ls_source(m3_state *s) { cell_face_velocity(s); ls_source(s, F_(s, M3_S0)); return 0; }
int m3_ls_rk3_solver(m3_state *s) { ls_rk3_solver(s); return 0; }
int m3_gradfi(m3_state *s) { sgnf(s, F_(s, M3_PHI)); gradfi(s, F_(s, M3_PHI)); return 0; }
int m3_ls_redistance(m3_state *s) { ls_redistance(s, F_(s, M3_PHI)); return 0; }
int m3_mpls(m3_state *s, double im) { mpls(s, im); return 0; }
int m3_solve_ls(m3_state *s) { solve_ls(s); return 0; }
int m3_second_order_upwind(m3_state *s) { second_order_upwind(s); return 0; }
int m3_rhsp(m3_state *s) { rhsp(s); return 0; }
int m3_adams_bashforth(m3_state *s) { adams_bashforth(s); return 0; }
int m3_momentum(m3_state *s) {
    solve_momentum(s);
    bc_vel(s, F_(s, M3_U), 0);
    bc_vel(s, F_(s, M3_V), 1);
    bc_vel(s, F_(s, M3_W), 2);
    return 0;
}
int m3_solve_ns(m3_state *s, int exact_ppe) { return solve_ns(s, exact_ppe); }
int m3_find_mass(m3_state *s) { find_mass(s); return 0; }
int m3_store_data(m3_state *s) { store_data(s); return 0; }

int m3_setup_tail(m3_state *s, int exact_ppe) {
    bc3d(s, F_(s, M3_PHI));
    bc3d(s, F_(s, M3_U));
    bc3d(s, F_(s, M3_V));
    bc3d(s, F_(s, M3_W));
    ls_redistance(s, F_(s, M3_PHI));
    s->pr.time = s->pr.iter * s->pr.dt;
    amurho(s);
    store_data(s);
    int rc = solve_ns(s, exact_ppe);
    store_data(s);
    find_mass(s);
    return rc;
}

int m3_step(m3_state *s, int exact_ppe) {
    s->pr.iter = s->pr.iter + 1;
    s->pr.time = s->pr.time + s->pr.dt;
    solve_ls(s);
    int rc = solve_ns(s, exact_ppe);
    find_mass(s);
    store_data(s);
    return rc;
}
