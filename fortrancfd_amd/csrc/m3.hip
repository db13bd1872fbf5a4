// m3.hip — the two-bubble step of MPLS_RisingBubble on MI355X
// (`MPLS_RisingBubble/20190612/`, 3D):
//   * SOLVE_LS: each bubble's level set advected by SSP-RK3 with the upwind
//     combined compact (M-DRPCCD) source, reinitialised every fifth step,
//     corrected to its own mass by five MPLS iterations, then the two merged
//     by max and reinitialised            (LS_SOLVER.F90:1-321, LS_MAINTAIN.F90)
//   * LS_REDISTANCE with the RIS_MASS constraint     (LS_REDISTANCE.F90:1-384)
//   * AMURHO (phi > 0 is the bubble: rho = rho_12 H + 1 - H), LS_CCD
//     normals and curvature                (AMURHO.F90, FIND_CURVATURE.F90:99-199)
//   * SOLVE_NS: second-order upwind momentum source, Adams-Bashforth, the
//     pressure solve (the multigrid of mg.hip on PRESS_POISSON's operator),
//     the projection, up to four Picard iterations with the reference's
//     early exit            (SOLVE_NS.F90, SECOND_ORDER_UPWIND.F90,
//                            SOLVE_PRESSURE_POISSON.F90)
//   * FIND_MASS's level-set sums and the per-bubble targets, STORE_DATA.
//
// Fields (the reference's arrays -> context fields): PHI, PHI_OLD, PHI1,
// PHI2; U, V, W (MAC) and their OLD level; UH, VH, WH -> NU, NV, NW; US, VS,
// WS -> SRC_U.., their OLD level -> SRC_U_OLD..; RHO, AMU -> RHO, MU; RHO_OLD,
// AMU_OLD -> RHO_OLD, MU_OLD; P; RHS -> PPE_SRC; HEAVY, DELTA, SGN, GRAD;
// NORMAL_X.. -> NX..; CURV.  U1, V1, W1 and the pseudo-time sources live in
// context scratch.  Everything is FP64; one thread per interior cell, 64
// threads along i.
#include <algorithm>
#include <cmath>
#include <vector>

#include "fcfd_internal.h"
#include "weno.h"

using namespace fcfd;

namespace fcfd {

namespace {

constexpr double kPi = 3.14159265358979323846;
constexpr double kEps = 1.0e-12;   // EPS of the reference's PRECISION module

// HEAV (LS_REDISTANCE.F90:330-348): smoothed Heaviside, hp its derivative
__device__ __forceinline__ double m3_heav(double x, double w, double &hp) {
    if (x > w) { hp = 0.0; return 1.0; }
    if (x < -w) { hp = 0.0; return 0.0; }
    double s, c;
    sincos(kPi * x / w, &s, &c);
    hp = 0.5 * (1.0 / w + 1.0 / w * c);
    return 0.5 * (1.0 + x / w + 1.0 / kPi * s);
}

// HEAVY_F / AMURHO over the interior: H (zero below EPS), delta and, when R
// is given, rho = rho_12 H + (1 - H), mu = mu_12 H + (1 - H)
__global__ void k_m3_props(const double *__restrict__ phi, double *__restrict__ H, double *__restrict__ D,
                           double *__restrict__ R, double *__restrict__ M, Geom G, double w, double r12, double m12) {
    int i, j, k;
    if (!cell_of_flat(G, i, j, k)) return;
    const long id = G.id(i, j, k);
    double hp;
    double h = m3_heav(phi[id], w, hp);
    if (h < kEps) h = 0.0;
    H[id] = h;
    D[id] = hp;
    if (R) {
        R[id] = r12 * h + (1.0 - h);
        M[id] = m12 * h + (1.0 - h);
    }
}

// CELL_FACE_VELOCITY (LS_SOLVER.F90:93-116)
__global__ void k_m3_cell_vel(const double *__restrict__ U, const double *__restrict__ V, const double *__restrict__ W,
                              double *__restrict__ UH, double *__restrict__ VH, double *__restrict__ WH, Geom G) {
    int i, j, k;
    if (!cell_of_flat(G, i, j, k)) return;
    const long id = G.id(i, j, k);
    UH[id] = 0.5 * (U[id] + U[id - 1]);
    VH[id] = 0.5 * (V[id] + V[id - G.sx]);
    WH[id] = 0.5 * (W[id] + W[id - G.sxy]);
}

// one SSP-RK3 stage of LS_RK3_SOLVER (LS_SOLVER.F90:39-91): the source
// s = -(F UH + G VH + H WH) from the UCCD derivatives, the increment added to
// phi in place (each cell reads only itself)
__global__ void k_m3_adv_stage(double *__restrict__ phi, const double *__restrict__ uh, const double *__restrict__ vh,
                               const double *__restrict__ wh, const double *__restrict__ fx,
                               const double *__restrict__ fy, const double *__restrict__ fz, double *__restrict__ s0,
                               double *__restrict__ s1, Geom G, double dt, int stage) {
    int i, j, k;
    if (!cell_of_flat(G, i, j, k)) return;
    const long id = G.id(i, j, k);
    const double s = -(fx[id] * uh[id] + fy[id] * vh[id] + fz[id] * wh[id]);
    if (stage == 0) {
        s0[id] = s;
        phi[id] = phi[id] + dt * s;
    } else if (stage == 1) {
        s1[id] = s;
        phi[id] = phi[id] + dt / 4.0 * (-3.0 * s0[id] + s);
    } else {
        phi[id] = phi[id] + dt / 12.0 * (-s0[id] - s1[id] + 8.0 * s);
    }
}

// sgnf (LS_REDISTANCE.F90:360-376)
__global__ void k_m3_sign(const double *__restrict__ phi, double *__restrict__ S, Geom G, double w) {
    int i, j, k;
    if (!cell_of_flat(G, i, j, k)) return;
    const long id = G.id(i, j, k);
    double hp;
    const double h = m3_heav(phi[id], w, hp);
    S[id] = 2.0 * (h - 0.5);
}

__device__ __forceinline__ double sq(double x) { return x * x; }

// weno_js (SCHEME_WENO.F90:47-101) at one point is the WENO5-JS pair of
// weno.h with EPS = 1e-12: the left-biased value from f(i-2..i+2), the
// right-biased one from f(i-1..i+3), the same smoothness indicators and
// weights d = (1, 6, 3), normalised there with one division instead of six.

// the one-sided WENO values of the backward differences along a line:
// F(m) = (T(m) - T(m-1)) / h for interior m, F of a ghost index = F of the
// nearest interior index (FV_DERI's bc3d); n the line's interior length.
// The seven values T(i-3 .. i+3) lie inside the 3 ghost layers for every
// i, and every clamped difference is one of their six neighbour
// differences: F(clamp(i-2+q)) = d[clamp(q, 2-i, n+1-i)].  So the loads
// take no branch, the caller issues the three lines of a cell together
// (m3_line_load), and only the cells within two of a wall remap their
// differences.
__device__ __forceinline__ void m3_line_load(const double *__restrict__ T, long id, long st, double (&g)[7]) {
#pragma unroll
    for (int r = 0; r < 7; ++r) g[r] = T[id + (long)(r - 3) * st];
}

__device__ __forceinline__ void m3_line_weno(const double (&g)[7], int i, int n, double ih, double &up, double &um) {
    double f[6];
#pragma unroll
    for (int q = 0; q < 6; ++q) f[q] = (g[q + 1] - g[q]) * ih;
    if (i < 2 || i > n - 4) clamp_window(f, 2 - i, n + 1 - i);
    um = weno_left(f[0], f[1], f[2], f[3], f[4], kEps);
    up = weno_right(f[1], f[2], f[3], f[4], f[5], kEps);
}

// GRADFI + the pointwise part of RIS_MASS (LS_REDISTANCE.F90:148-238):
// |grad T| by Godunov on the WENO values, delta of T, and the constraint
// terms a = SGN delta (|grad| - 1), b = |grad| delta^2
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(1, 4)))
k_m3_redis_a(const double *__restrict__ T, const double *__restrict__ S, double *__restrict__ Gr,
                             double *__restrict__ D, double *__restrict__ A, double *__restrict__ B, Geom G, double w) {
    int i, j, k;
    if (!cell_of_flat(G, i, j, k)) return;
    const long id = G.id(i, j, k);
    double gx[7], gy[7], gz[7];
    m3_line_load(T, id, 1, gx);
    m3_line_load(T, id, G.sx, gy);
    m3_line_load(T, id, G.sxy, gz);
    const double sg = S[id];
    double up, um, vp, vm, wp, wm;
    m3_line_weno(gx, i, G.nx, 1.0 / G.dx, up, um);
    m3_line_weno(gy, j, G.ny, 1.0 / G.dy, vp, vm);
    m3_line_weno(gz, k, G.nz, 1.0 / G.dz, wp, wm);
    double g;
    if (sg >= 0.0)
        g = sqrt(sq(fmax(-fmin(up, 0.0), fmax(um, 0.0))) + sq(fmax(-fmin(vp, 0.0), fmax(vm, 0.0))) +
                 sq(fmax(-fmin(wp, 0.0), fmax(wm, 0.0))));
    else
        g = sqrt(sq(fmax(fmax(up, 0.0), -fmin(um, 0.0))) + sq(fmax(fmax(vp, 0.0), -fmin(vm, 0.0))) +
                 sq(fmax(fmax(wp, 0.0), -fmin(wm, 0.0))));
    double hp;
    m3_heav(gx[3], w, hp);
    Gr[id] = g;
    D[id] = hp;
    A[id] = (sg * hp) * (g - 1.0);
    B[id] = g * sq(hp);
}

// RIS_MASS's 27-point constraint (a, b zero-gradient beyond the walls, as
// bc3d leaves them) and one RK stage of LS_REDIS (LS_REDISTANCE.F90:68-146)
// on T in place: s = (SGN - C) |grad| - SGN, C = delta (51 a + sum a) /
// (51 b + sum b).  C enters only where delta != 0, the interface band of
// |T| <= INTERFACE_WIDTH (well under 1 % of the cells of a bubble problem);
// elsewhere C = 0 and the cell needs none of the 54 values of a and b, so a
// thread per cell reads them only in the band (3 x 3 plane sums of planes
// k-1, k, k+1 in the order of the reference's loop nest).
constexpr int M3_BT = 256;
__global__ void __launch_bounds__(M3_BT) k_m3_redis_b(double *__restrict__ T, const double *__restrict__ S,
                                                     const double *__restrict__ Gr, const double *__restrict__ D,
                                                     const double *__restrict__ A, const double *__restrict__ B,
                                                     double *__restrict__ s0, double *__restrict__ s1, Geom G,
                                                     double rdt, int stage) {
    int i, j, k;
    if (!cell_of_flat(G, i, j, k)) return;
    const long id = G.id(i, j, k);
    const double d = D[id], sg = S[id], g = Gr[id];
    double c = 0.0;
    if (d != 0.0) {
        const int im = clampi(i - 1, 0, G.nx - 1), ip = clampi(i + 1, 0, G.nx - 1);
        const int jm = clampi(j - 1, 0, G.ny - 1), jp = clampi(j + 1, 0, G.ny - 1);
        // 3 x 3 sums of a and b over plane kk (clamped), rows j-1, j, j+1
        auto plane = [&](int kk, double &pa, double &pb) {
            const int kc = clampi(kk, 0, G.nz - 1);
            pa = 0.0;
            pb = 0.0;
            const int js[3] = {jm, j, jp};
#pragma unroll
            for (int r = 0; r < 3; ++r) {
                const long row = G.id(0, js[r], kc);
                pa += A[row + im];
                pa += A[row + i];
                pa += A[row + ip];
                pb += B[row + im];
                pb += B[row + i];
                pb += B[row + ip];
            }
        };
        double am, bm, ac, bc, ap, bp;
        plane(k - 1, am, bm);
        plane(k, ac, bc);
        plane(k + 1, ap, bp);
        const double suma = 51.0 * A[id] + (am + ac + ap), sumb = 51.0 * B[id] + (bm + bc + bp);
        c = (sumb != 0.0 ? suma / sumb : 0.0) * d;
    }
    const double s = (sg - c) * g - sg;
    if (stage == 0) {
        s0[id] = s;
        T[id] = T[id] - rdt * s;
    } else if (stage == 1) {
        s1[id] = s;
        T[id] = T[id] - rdt / 4.0 * (-3.0 * s0[id] + s);
    } else {
        T[id] = T[id] - rdt / 12.0 * (-s0[id] - s1[id] + 8.0 * s);
    }
}

// deterministic sums over the interior: RB blocks, each thread a fixed
// strided subset, then fixed-shape tree reductions (k_m3_final)
constexpr int RB = 1024, RT = 256;

template <int NS>
__device__ __forceinline__ void m3_block_store(double (&acc)[NS], double *partial) {
    __shared__ double sh[NS][RT];
    for (int s = 0; s < NS; ++s) sh[s][threadIdx.x] = acc[s];
    __syncthreads();
    for (int w = RT / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int s = 0; s < NS; ++s) sh[s][threadIdx.x] += sh[s][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int s = 0; s < NS; ++s) partial[s * RB + blockIdx.x] = sh[s][0];
}

// out[s] = sum over blocks of partial[s * RB + b], one block
__global__ void k_m3_final(const double *__restrict__ partial, double *__restrict__ out, int ns) {
    __shared__ double sh[RT];
    for (int s = 0; s < ns; ++s) {
        double a = 0.0;
        for (int b = threadIdx.x; b < RB; b += RT) a += partial[s * RB + b];
        sh[threadIdx.x] = a;
        __syncthreads();
        for (int w = RT / 2; w > 0; w >>= 1) {
            if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
            __syncthreads();
        }
        if (threadIdx.x == 0) out[s] = sh[0];
        __syncthreads();
    }
}

// level-set volume and mass (FIND_MASS.F90:20-37; no cell volume) and, with
// the normals, |grad phi| -> Gr and the MPLS denominator
// sum (2 (rho_12 - 1) H + 1) delta^2 |grad phi| (LS_SOLVER.F90:283-296)
__global__ void k_m3_sums(const double *__restrict__ H, const double *__restrict__ R, const double *__restrict__ D,
                          const double *__restrict__ nx_, const double *__restrict__ ny_, const double *__restrict__ nz_,
                          double *__restrict__ Gr, double *__restrict__ partial, Geom G, double r12) {
    double acc[3] = {0.0, 0.0, 0.0};
    const long n = (long)G.nx * G.ny * G.nz;
    for (long t = (long)blockIdx.x * RT + threadIdx.x; t < n; t += (long)RB * RT) {
        int i, j, k;
        cell_of_index(t, G, i, j, k);
        const long id = G.id(i, j, k);
        const double h = H[id];
        acc[0] += h;
        acc[1] += h * R[id];
        if (nx_) {
            const double g = sqrt(sq(nx_[id]) + sq(ny_[id]) + sq(nz_[id]));
            Gr[id] = g;
            acc[2] += (2.0 * (r12 - 1.0) * h + 1.0) * sq(D[id]) * g;
        }
    }
    m3_block_store<3>(acc, partial);
}

// MPLS update (LS_SOLVER.F90:298-318): lambda = (im - mass) / (dt fs) from
// the reduced sums, phi += lambda delta |grad phi| dt
__global__ void k_m3_mpls_update(double *__restrict__ phi, const double *__restrict__ D, const double *__restrict__ Gr,
                                 const double *__restrict__ sums, double *__restrict__ lam_out, Geom G, double im,
                                 double dt) {
    int i = -1, j = -1, k = -1;
    const bool in = cell_of_flat(G, i, j, k);
    const double lam = (im - sums[1]) / (dt * sums[2]);
    if (in && i == 0 && j == 0 && k == 0) *lam_out = lam;
    if (!in) return;
    const long id = G.id(i, j, k);
    phi[id] = phi[id] + lam * D[id] * Gr[id] * dt;
}

// phi = max(phi1, phi2) over the interior (LS_SOLVER.F90:24-32)
__global__ void k_m3_merge(const double *__restrict__ p1, const double *__restrict__ p2, double *__restrict__ phi,
                           Geom G) {
    int i, j, k;
    if (!cell_of_flat(G, i, j, k)) return;
    const long id = G.id(i, j, k);
    phi[id] = fmax(p1[id], p2[id]);
}

// LS_CCD's curvature (FIND_CURVATURE.F90:169-190)
__global__ void k_m3_curv(const double *__restrict__ X, const double *__restrict__ Y, const double *__restrict__ Z,
                          const double *__restrict__ XX, const double *__restrict__ YY, const double *__restrict__ ZZ,
                          const double *__restrict__ XY, const double *__restrict__ XZ, const double *__restrict__ YZ,
                          double *__restrict__ C, Geom G) {
    int i, j, k;
    if (!cell_of_flat(G, i, j, k)) return;
    const long id = G.id(i, j, k);
    const double fx = X[id], fy = Y[id], fz = Z[id];
    C[id] = (XX[id] * (sq(fy) + sq(fz)) + YY[id] * (sq(fx) + sq(fz)) + ZZ[id] * (sq(fx) + sq(fy)) -
             2.0 * (XY[id] * fx * fy + XZ[id] * fx * fz + YZ[id] * fy * fz)) /
            pow(sq(fx) + sq(fy) + sq(fz) + kEps, 1.5);
}

// VELA / VELB / VELC + Adams_Bashforth (SECOND_ORDER_UPWIND.F90:10-256,
// SOLVE_NS.F90:147-166): the momentum source of each face (u faces i <
// NX, v faces j < NY, w faces k < NZ: 2nd-order upwind advection of the old
// velocity by the current one, viscous term with (mu_a + mu_b) / (rho_a +
// rho_b) of the old level, CSF with the smoothed delta of the face, gravity
// 1/Fr^2 on w) into US, VS, WS; then U1 = 1.5 US - 0.5 US_old on every
// interior face
struct M3Vel {
    const double *u, *v, *w, *uo, *vo, *wo;
};
struct M3Src {
    double *us, *vs, *ws;
    const double *uso, *vso, *wso;
    double *u1, *v1, *w1;
};
struct M3Phys {
    double re, we, fr2, wid, st_force, g_force;
};

__device__ __forceinline__ double upwind2(double a, const double *__restrict__ o, long id, long d, double h) {
    const double ap = 0.5 * (a + fabs(a)), am = 0.5 * (a - fabs(a));
    return 0.5 / h * (ap * (3.0 * o[id] - 4.0 * o[id - d] + o[id - 2 * d]) +
                      am * (-o[id + 2 * d] + 4.0 * o[id + d] - 3.0 * o[id]));
}

// The smoothed delta of the three faces is formed first (its band test is a
// branch, and loads after a branch cannot be issued before it); the
// component loop then has no branches, so its loads issue together.  The
// outermost face of each direction (not updated: its source keeps the old
// value) is formed too and not selected; its stencil stays inside the
// ghost layers.
__global__ void __launch_bounds__(M3_BT) __attribute__((amdgpu_waves_per_eu(1, 4)))
k_m3_predict(M3Vel V, M3Src Sr, const double *__restrict__ mo, const double *__restrict__ ro,
             const double *__restrict__ phi, const double *__restrict__ K, const double *__restrict__ nx_,
             const double *__restrict__ ny_, const double *__restrict__ nz_, Geom G, M3Phys ph) {
    int i, j, k;
    if (!cell_of_flat(G, i, j, k)) return;
    const long id = G.id(i, j, k), sx = G.sx, sxy = G.sxy;
    const double dx = G.dx, dy = G.dy, dz = G.dz;
    const double re2x = ph.re * (dx * dx), re2y = ph.re * (dy * dy), re2z = ph.re * (dz * dz);
    const long off[3] = {1, sx, sxy};
    const double *old[3] = {V.uo, V.vo, V.wo};
    const double *nrm[3] = {nx_, ny_, nz_};
    double *src[3] = {Sr.us, Sr.vs, Sr.ws};
    const double *srco[3] = {Sr.uso, Sr.vso, Sr.wso};
    double *u1[3] = {Sr.u1, Sr.v1, Sr.w1};
    const bool inner[3] = {i < G.nx - 1, j < G.ny - 1, k < G.nz - 1};
    double detf[3], pf[3];
    const double pc = phi[id];
#pragma unroll
    for (int c = 0; c < 3; ++c) pf[c] = phi[id + off[c]];
#pragma unroll
    for (int c = 0; c < 3; ++c) m3_heav(0.5 * (pc + pf[c]), ph.wid, detf[c]);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const double s_old = src[c][id];
        const long d = off[c];
        double a[3];   // advecting velocity at this face, per direction
        if (c == 0) {
            a[0] = V.u[id];
            a[1] = 0.25 * (V.v[id + 1] + V.v[id] + V.v[id + 1 - sx] + V.v[id - sx]);
            a[2] = 0.25 * (V.w[id + 1] + V.w[id] + V.w[id + 1 - sxy] + V.w[id - sxy]);
        } else if (c == 1) {
            a[0] = 0.25 * (V.u[id + sx] + V.u[id] + V.u[id - 1 + sx] + V.u[id - 1]);
            a[1] = V.v[id];
            a[2] = 0.25 * (V.w[id + sx] + V.w[id] + V.w[id + sx - sxy] + V.w[id - sxy]);
        } else {
            a[0] = 0.25 * (V.u[id + sxy] + V.u[id] + V.u[id - 1 + sxy] + V.u[id - 1]);
            a[1] = 0.25 * (V.v[id + sxy] + V.v[id] + V.v[id - sx + sxy] + V.v[id - sx]);
            a[2] = V.w[id];
        }
        const double *q = old[c];
        const double cm = mo[id] + mo[id + d], cr = ro[id] + ro[id + d], cc = cm / cr;
        const double curvk = 0.5 * (K[id] + K[id + d]), fin = 0.5 * (nrm[c][id] + nrm[c][id + d]);
        const double a1 = upwind2(a[0], q, id, 1, dx), a2 = upwind2(a[1], q, id, sx, dy),
                     a3 = upwind2(a[2], q, id, sxy, dz);
        const double a4 = 1.0 / re2x * cc * (q[id + 1] - 2.0 * q[id] + q[id - 1]);
        const double a5 = 1.0 / re2y * cc * (q[id + sx] - 2.0 * q[id] + q[id - sx]);
        const double a6 = 1.0 / re2z * cc * (q[id + sxy] - 2.0 * q[id] + q[id - sxy]);
        const double a7 = ph.st_force != 0.0 ? 2.0 / (cr * ph.we) * curvk * detf[c] * fin : 0.0;
        double s = a1 + a2 + a3 - a4 - a5 - a6 + a7;
        if (c == 2 && ph.g_force != 0.0) s = s + 1.0 / ph.fr2;
        s = inner[c] ? s : s_old;
        if (inner[c]) src[c][id] = s;
        u1[c][id] = 1.5 * s - 0.5 * srco[c][id];
    }
}

// RHSP (SOLVE_PRESSURE_POISSON.F90:12-58): RHS = div(u_old / dt - U1) with
// U1 zero on the wall faces
__global__ void k_m3_rhs(const double *__restrict__ uo, const double *__restrict__ vo, const double *__restrict__ wo,
                         const double *__restrict__ u1, const double *__restrict__ v1, const double *__restrict__ w1,
                         double *__restrict__ rhs, Geom G, double dt) {
    int i, j, k;
    if (!cell_of_flat(G, i, j, k)) return;
    const long id = G.id(i, j, k), sx = G.sx, sxy = G.sxy;
    const double ue = i < G.nx - 1 ? u1[id] : 0.0, uw = i > 0 ? u1[id - 1] : 0.0;
    const double vn = j < G.ny - 1 ? v1[id] : 0.0, vs = j > 0 ? v1[id - sx] : 0.0;
    const double wt = k < G.nz - 1 ? w1[id] : 0.0, wb = k > 0 ? w1[id - sxy] : 0.0;
    rhs[id] = (uo[id] / dt - uo[id - 1] / dt - ue + uw) / G.dx + (vo[id] / dt - vo[id - sx] / dt - vn + vs) / G.dy +
              (wo[id] / dt - wo[id - sxy] / dt - wt + wb) / G.dz;
}

// SOLVE_MOMENTUM (SOLVE_NS.F90:102-145): u = u_old + dt (-U1 - 2 dp / (h
// (rho_a + rho_b))) on the inner faces, with the squared change of each
// component summed for CHECK_CONVERGENCE's L2 norms (the faces not updated
// do not change)
__global__ void k_m3_momentum(double *__restrict__ u, double *__restrict__ v, double *__restrict__ w,
                              const double *__restrict__ uo, const double *__restrict__ vo,
                              const double *__restrict__ wo, const double *__restrict__ u1,
                              const double *__restrict__ v1, const double *__restrict__ w1,
                              const double *__restrict__ p, const double *__restrict__ rho, double *__restrict__ partial,
                              Geom G, double dt) {
    // every load of a cell before the face tests (the stencils of the
    // outermost faces, which are not updated, stay inside the ghost layers)
    double acc[3] = {0.0, 0.0, 0.0};
    const long n = (long)G.nx * G.ny * G.nz;
    for (long t = (long)blockIdx.x * RT + threadIdx.x; t < n; t += (long)RB * RT) {
        int i, j, k;
        cell_of_index(t, G, i, j, k);
        const long id = G.id(i, j, k);
        const double pc = p[id], rc = rho[id];
        const double pe = p[id + 1], pn = p[id + G.sx], pt = p[id + G.sxy];
        const double re = rho[id + 1], rn = rho[id + G.sx], rt = rho[id + G.sxy];
        const double uc = u[id], vc = v[id], wc = w[id];
        double au = -u1[id] - 2.0 * (pe - pc) / (G.dx * (rc + re));
        au = uo[id] + au * dt;
        double av = -v1[id] - 2.0 * (pn - pc) / (G.dy * (rc + rn));
        av = vo[id] + av * dt;
        double aw = -w1[id] - 2.0 * (pt - pc) / (G.dz * (rc + rt));
        aw = wo[id] + aw * dt;
        if (i < G.nx - 1) {
            acc[0] += sq(au - uc);
            u[id] = au;
        }
        if (j < G.ny - 1) {
            acc[1] += sq(av - vc);
            v[id] = av;
        }
        if (k < G.nz - 1) {
            acc[2] += sq(aw - wc);
            w[id] = aw;
        }
    }
    m3_block_store<3>(acc, partial);
}

}  // namespace

// ------------------------------------------------------------ host side


// the fields only the MPLS_RisingBubble step holds, allocated on first use
// (by an M3 entry point or a push / pull of one of them)
int m3_fields(fcfd_ctx *c) {
    // the kernels number the interior cells in 32 bits (grid_flat)
    if (c->G.dim != 3 || !flat_fits(c->G)) return FCFD_ERR_ARG;
    for (int f : {FCFD_PHI1, FCFD_PHI2, FCFD_RHO_OLD, FCFD_MU_OLD})
        if (!c->f[f]) {
            FCFD_CHECK(hipMalloc((void **)&c->f[f], sizeof(double) * c->G.ntot));
            FCFD_CHECK(hipMemsetAsync(c->f[f], 0, sizeof(double) * c->G.ntot, c->stream));
        }
    return FCFD_OK;
}

// HEAVY_F (props without rho) or AMURHO (with): interior, then the ghost
// copies the reference makes (HEAVY; RHO and AMU)
static int m3_props(fcfd_ctx *c, const double *phi, int with_rho) {
    const Geom &G = c->G;
    double **F = c->f;
    {
        ProfScope ps_(c, R_MISC);
        hipLaunchKernelGGL(k_m3_props, grid_flat(G, M3_BT), dim3(M3_BT), 0, c->stream, phi, F[FCFD_HEAVY],
                           F[FCFD_DELTA], with_rho ? F[FCFD_RHO] : nullptr, with_rho ? F[FCFD_MU] : nullptr, G,
                           c->pr.ls_wid, c->pr.rho_12, c->pr.mu_12);
        FCFD_CHECK(hipGetLastError());
    }
    FCFD_TRY(bc_scalar(c, F[FCFD_HEAVY]));
    if (with_rho) {
        FCFD_TRY(bc_scalar(c, F[FCFD_RHO]));
        FCFD_TRY(bc_scalar(c, F[FCFD_MU]));
    }
    return FCFD_OK;
}

// LS_RK3_SOLVER on PHI: cell velocities, then three stages with the UCCD
// derivatives of the interior lines (M-DRPCCD) upwind by UH, VH, WH
static int m3_ls_rk3(fcfd_ctx *c) {
    const Geom &G = c->G;
    double **F = c->f;
    {
        ProfScope ps_(c, R_MISC);
        hipLaunchKernelGGL(k_m3_cell_vel, grid_flat(G, M3_BT), dim3(M3_BT), 0, c->stream, F[FCFD_U], F[FCFD_V],
                           F[FCFD_W], F[FCFD_NU], F[FCFD_NV], F[FCFD_NW], G);
        FCFD_CHECK(hipGetLastError());
    }
    FCFD_TRY(bc_scalar(c, F[FCFD_NU]));
    FCFD_TRY(bc_scalar(c, F[FCFD_NV]));
    FCFD_TRY(bc_scalar(c, F[FCFD_NW]));
    double *d[3] = {c->un, c->vn, c->wn};
    const double *vel[3] = {F[FCFD_NU], F[FCFD_NV], F[FCFD_NW]};
    for (int s = 0; s < 3; ++s) {
        for (int dir = 0; dir < 3; ++dir) FCFD_TRY(uccd_lines_set(c, 1, dir, F[FCFD_PHI], vel[dir], d[dir], nullptr));
        {
            ProfScope ps_(c, R_LS_ADV);
            hipLaunchKernelGGL(k_m3_adv_stage, grid_flat(G, M3_BT), dim3(M3_BT), 0, c->stream, F[FCFD_PHI], vel[0],
                               vel[1], vel[2], d[0], d[1], d[2], c->l1, c->l2, G, c->pr.dt, s);
            FCFD_CHECK(hipGetLastError());
        }
        FCFD_TRY(bc_scalar(c, F[FCFD_PHI]));
    }
    return FCFD_OK;
}

// the pseudo-step count of LS_REDISTANCE with BTN = 1: RED_TIME accumulates
// RDT until it exceeds 2 INTERFACE_WIDTH (LS_REDISTANCE.F90:44-64)
static int m3_redis_steps(const fcfd_ctx *c) {
    double t = 0.0;
    int n = 0;
    if (!(c->pr.rdt > 0.0)) return 0;
    do {
        t = t + c->pr.rdt;
        ++n;
    } while (!(t > 2.0 * c->pr.ls_wid) && n < 100000);
    return n;
}

// LS_REDISTANCE(T, 1) on T (a ghosted field whose ghosts hold bc3d copies)
static int m3_redistance(fcfd_ctx *c, double *T) {
    const Geom &G = c->G;
    double **F = c->f;
    if (!flat_fits(G)) return FCFD_ERR_ARG;
    hipLaunchKernelGGL(k_m3_sign, grid_flat(G, M3_BT), dim3(M3_BT), 0, c->stream, T, F[FCFD_SIGN], G, c->pr.ls_wid);
    FCFD_CHECK(hipGetLastError());
    const dim3 ga = grid_flat(G, M3_BT);
    const int n = m3_redis_steps(c);
    for (int it = 0; it < n; ++it)
        for (int s = 0; s < 3; ++s) {
            {
                ProfScope ps_(c, R_REDIS_A);
                hipLaunchKernelGGL(k_m3_redis_a, ga, dim3(M3_BT), 0, c->stream, T, F[FCFD_SIGN], F[FCFD_GRAD],
                                   F[FCFD_DELTA], c->s2, c->s3, G, c->pr.ls_wid);
                FCFD_CHECK(hipGetLastError());
            }
            {
                ProfScope ps_(c, R_REDIS_B);
                hipLaunchKernelGGL(k_m3_redis_b, ga, dim3(M3_BT), 0, c->stream, T, F[FCFD_SIGN], F[FCFD_GRAD],
                                   F[FCFD_DELTA], c->s2, c->s3, c->l1, c->l2, G, c->pr.rdt, s);
                FCFD_CHECK(hipGetLastError());
            }
            FCFD_TRY(bc_scalar(c, T));
        }
    c->diag.redis_iters = n;
    return FCFD_OK;
}

// the interface sums into result[0..2] (volume, mass, MPLS denominator),
// |grad phi| from NX, NY, NZ into GRAD when with_grad
static int m3_sums(fcfd_ctx *c, int with_grad) {
    const Geom &G = c->G;
    double **F = c->f;
    ProfScope ps_(c, R_MASS);
    hipLaunchKernelGGL(k_m3_sums, dim3(RB), dim3(RT), 0, c->stream, F[FCFD_HEAVY], F[FCFD_RHO], F[FCFD_DELTA],
                       with_grad ? F[FCFD_NX] : nullptr, F[FCFD_NY], F[FCFD_NZ], F[FCFD_GRAD], c->red.partial, G,
                       c->pr.rho_12);
    FCFD_CHECK(hipGetLastError());
    hipLaunchKernelGGL(k_m3_final, dim3(1), dim3(RT), 0, c->stream, c->red.partial, c->red.result, 3);
    FCFD_CHECK(hipGetLastError());
    return FCFD_OK;
}

// MPLS(im) on PHI: five corrections, all on the device
static int m3_mpls(fcfd_ctx *c, double im) {
    const Geom &G = c->G;
    double **F = c->f;
    for (int it = 0; it < 5; ++it) {
        FCFD_TRY(m3_props(c, F[FCFD_PHI], 1));
        FCFD_TRY(ccd_lines_set(c, 1, 0, F[FCFD_PHI], F[FCFD_NX], c->cx));
        FCFD_TRY(ccd_lines_set(c, 1, 1, F[FCFD_PHI], F[FCFD_NY], c->cy));
        FCFD_TRY(ccd_lines_set(c, 1, 2, F[FCFD_PHI], F[FCFD_NZ], c->cz));
        FCFD_TRY(m3_sums(c, 1));
        ProfScope ps_(c, R_MPLS_UPD);
        hipLaunchKernelGGL(k_m3_mpls_update, grid_flat(G, M3_BT), dim3(M3_BT), 0, c->stream, F[FCFD_PHI],
                           F[FCFD_DELTA], F[FCFD_GRAD], c->red.result, c->red.result + RES_LAMBDA, G, im, c->pr.dt);
        FCFD_CHECK(hipGetLastError());
    }
    return FCFD_OK;
}

// LS_CCD(PHI): first / second and cross CCD derivatives of the interior
// lines, the curvature, the ghost copies of the curvature and normals
static int m3_ls_ccd(fcfd_ctx *c) {
    const Geom &G = c->G;
    double **F = c->f;
    FCFD_TRY(m3_props(c, F[FCFD_PHI], 0));
    FCFD_TRY(ccd_lines_set(c, 1, 0, F[FCFD_PHI], F[FCFD_NX], F[FCFD_NXX]));
    FCFD_TRY(ccd_lines_set(c, 1, 1, F[FCFD_PHI], F[FCFD_NY], F[FCFD_NYY]));
    FCFD_TRY(ccd_lines_set(c, 1, 2, F[FCFD_PHI], F[FCFD_NZ], F[FCFD_NZZ]));
    FCFD_TRY(ccd_lines_set(c, 1, 0, F[FCFD_NZ], F[FCFD_NXZ], c->cx));   // FZ_X
    FCFD_TRY(ccd_lines_set(c, 1, 1, F[FCFD_NX], F[FCFD_NXY], c->cx));   // FX_Y
    FCFD_TRY(ccd_lines_set(c, 1, 2, F[FCFD_NY], F[FCFD_NYZ], c->cx));   // FY_Z
    {
        ProfScope ps_(c, R_CURV);
        hipLaunchKernelGGL(k_m3_curv, grid_flat(G, M3_BT), dim3(M3_BT), 0, c->stream, F[FCFD_NX], F[FCFD_NY],
                           F[FCFD_NZ], F[FCFD_NXX], F[FCFD_NYY], F[FCFD_NZZ], F[FCFD_NXY], F[FCFD_NXZ], F[FCFD_NYZ],
                           F[FCFD_CURV], G);
        FCFD_CHECK(hipGetLastError());
    }
    for (int f : {FCFD_CURV, FCFD_NX, FCFD_NY, FCFD_NZ}) FCFD_TRY(bc_scalar(c, F[f]));
    return FCFD_OK;
}

static int m3_solve_ls(fcfd_ctx *c) {
    const Geom &G = c->G;
    double **F = c->f;
    const double im[2] = {c->m3.bubble_imass[0], c->m3.bubble_imass[1]};
    const int fb[2] = {FCFD_PHI1, FCFD_PHI2};
    for (int b = 0; b < 2; ++b) {
        FCFD_TRY(copy_field(c, F[FCFD_PHI], F[fb[b]]));
        FCFD_TRY(m3_ls_rk3(c));
        if (c->iter % 5 == 0) FCFD_TRY(m3_redistance(c, F[FCFD_PHI]));   // LS_MAINTAIN
        FCFD_TRY(m3_mpls(c, im[b]));
        FCFD_TRY(copy_field(c, F[fb[b]], F[FCFD_PHI]));
    }
    hipLaunchKernelGGL(k_m3_merge, grid_flat(G, M3_BT), dim3(M3_BT), 0, c->stream, F[FCFD_PHI1], F[FCFD_PHI2],
                       F[FCFD_PHI], G);
    FCFD_CHECK(hipGetLastError());
    FCFD_TRY(bc_scalar(c, F[FCFD_PHI]));
    return m3_redistance(c, F[FCFD_PHI]);
}

static int m3_store_data(fcfd_ctx *c) {
    double **F = c->f;
    const int pairs[][2] = {{FCFD_PHI_OLD, FCFD_PHI},     {FCFD_U_OLD, FCFD_U},         {FCFD_V_OLD, FCFD_V},
                            {FCFD_W_OLD, FCFD_W},         {FCFD_SRC_U_OLD, FCFD_SRC_U}, {FCFD_SRC_V_OLD, FCFD_SRC_V},
                            {FCFD_SRC_W_OLD, FCFD_SRC_W}, {FCFD_RHO_OLD, FCFD_RHO},     {FCFD_MU_OLD, FCFD_MU}};
    for (auto &pr : pairs) FCFD_TRY(copy_field(c, F[pr[0]], F[pr[1]]));
    return FCFD_OK;
}

static int m3_predict(fcfd_ctx *c) {
    const Geom &G = c->G;
    double **F = c->f;
    M3Vel V{F[FCFD_U], F[FCFD_V], F[FCFD_W], F[FCFD_U_OLD], F[FCFD_V_OLD], F[FCFD_W_OLD]};
    M3Src S{F[FCFD_SRC_U], F[FCFD_SRC_V], F[FCFD_SRC_W], F[FCFD_SRC_U_OLD], F[FCFD_SRC_V_OLD], F[FCFD_SRC_W_OLD],
            c->un, c->vn, c->wn};
    M3Phys ph{c->pr.re, c->pr.we, c->pr.fr * c->pr.fr, c->pr.ls_wid, c->pr.btn_sf, c->pr.btn_g};
    ProfScope ps_(c, R_PRED);
    hipLaunchKernelGGL(k_m3_predict, grid_flat(G, M3_BT), dim3(M3_BT), 0, c->stream, V, S, F[FCFD_MU_OLD],
                       F[FCFD_RHO_OLD], F[FCFD_PHI], F[FCFD_CURV], F[FCFD_NX], F[FCFD_NY], F[FCFD_NZ], G, ph);
    FCFD_CHECK(hipGetLastError());
    return FCFD_OK;
}

// SOLVE_NS: AMURHO, FIND_CURVATURE, the pressure operator of this RHO, then
// up to four Picard iterations; the iteration count and the early exit
// (L2 change below TOL_M) follow CHECK_CONVERGENCE
static int m3_solve_ns(fcfd_ctx *c) {
    const Geom &G = c->G;
    double **F = c->f;
    FCFD_TRY(m3_props(c, F[FCFD_PHI], 1));
    if (c->pr.btn_sf != 0.0) FCFD_TRY(m3_ls_ccd(c));
    c->m3.picard = 0;
    int initer = 0, rc = FCFD_OK, total = 0;
    bool ready = false;
    const double tol = c->pr.ppe_tol > 0.0 ? c->pr.ppe_tol : 1e-10;
    for (;;) {
        initer = initer + 1;
        FCFD_TRY(m3_predict(c));
        if (!c->m3.rec_mass) return FCFD_OK;
        {
            ProfScope ps_(c, R_PPE_SRC);
            hipLaunchKernelGGL(k_m3_rhs, grid_flat(G, M3_BT), dim3(M3_BT), 0, c->stream, F[FCFD_U_OLD],
                               F[FCFD_V_OLD], F[FCFD_W_OLD], c->un, c->vn, c->wn, F[FCFD_PPE_SRC], G, c->pr.dt);
            FCFD_CHECK(hipGetLastError());
        }
        if (!ready) {
            FCFD_TRY(mg_setup(c));
            ready = true;
        }
        int cyc = 0;
        const int r = ppe_solve(c, tol, c->pr.ppe_precision, &cyc, nullptr);
        if (r != FCFD_OK && r != FCFD_ERR_NOT_CONVERGED) return r;
        if (r != FCFD_OK) rc = r;
        total += cyc;
        {
            ProfScope ps_(c, R_PROJ);
            hipLaunchKernelGGL(k_m3_momentum, dim3(RB), dim3(RT), 0, c->stream, F[FCFD_U], F[FCFD_V], F[FCFD_W],
                               F[FCFD_U_OLD], F[FCFD_V_OLD], F[FCFD_W_OLD], c->un, c->vn, c->wn, F[FCFD_P],
                               F[FCFD_RHO], c->red.partial, G, c->pr.dt);
            FCFD_CHECK(hipGetLastError());
            hipLaunchKernelGGL(k_m3_final, dim3(1), dim3(RT), 0, c->stream, c->red.partial, c->red.result + 3, 3);
            FCFD_CHECK(hipGetLastError());
        }
        FCFD_TRY(div_check(c, 0));
        double l2[3];
        FCFD_CHECK(hipMemcpyAsync(l2, c->red.result + 3, sizeof(l2), hipMemcpyDeviceToHost, c->stream));
        FCFD_CHECK(hipStreamSynchronize(c->stream));
        const double ncell = (double)G.nx * (double)G.ny * (double)G.nz;
        c->m3.l2sum = sqrt(l2[0] / ncell) + sqrt(l2[1] / ncell) + sqrt(l2[2] / ncell);
        if (c->m3.l2sum < c->m3.picard_tol) initer = initer + 5;
        FCFD_TRY(velbc(c, F[FCFD_U], F[FCFD_V], F[FCFD_W], 0));
        c->m3.picard++;
        if (initer > 3) break;
    }
    c->diag.ppe_cycles_step = total;
    return rc;
}

// FIND_MASS's level-set part: AMURHO, the volume and mass sums; on the
// first call the two bubbles' targets from PHI1, PHI2 and the initial totals
static int m3_find_mass(fcfd_ctx *c) {
    double **F = c->f;
    FCFD_TRY(m3_props(c, F[FCFD_PHI], 1));
    FCFD_TRY(m3_sums(c, 0));
    double h[2];
    FCFD_CHECK(hipMemcpyAsync(h, c->red.result, sizeof(h), hipMemcpyDeviceToHost, c->stream));
    FCFD_CHECK(hipStreamSynchronize(c->stream));
    c->m3.vol = h[0];
    c->m3.mass = h[1];
    if (!c->m3.rec_mass) {
        const int fb[2] = {FCFD_PHI1, FCFD_PHI2};
        for (int b = 0; b < 2; ++b) {
            FCFD_TRY(m3_props(c, F[fb[b]], 1));
            FCFD_TRY(m3_sums(c, 0));
            double g[2];
            FCFD_CHECK(hipMemcpyAsync(g, c->red.result, sizeof(g), hipMemcpyDeviceToHost, c->stream));
            FCFD_CHECK(hipStreamSynchronize(c->stream));
            c->m3.bubble_imass[b] = g[1];
        }
        c->m3.imass = c->m3.mass;
        c->m3.ivol = c->m3.vol;
        c->m3.rec_mass = 1;
    }
    c->diag.mass = c->m3.mass;
    c->diag.vol = c->m3.vol;
    c->diag_pending = 0;
    return FCFD_OK;
}

}  // namespace fcfd

// ==================================================================== C ABI

#define M3_CTX(c)                             \
    if (!(c)) return FCFD_ERR_ARG;            \
    FCFD_TRY(fcfd::m3_fields(c))

extern "C" int fcfd_m3_get_state(fcfd_ctx *c, fcfd_m3_state *s) {
    if (!c || !s) return FCFD_ERR_ARG;
    FCFD_CHECK(hipStreamSynchronize(c->stream));
    *s = c->m3;
    s->div = c->diag.vel_div;
    return FCFD_OK;
}

extern "C" int fcfd_m3_set_state(fcfd_ctx *c, const fcfd_m3_state *s) {
    if (!c || !s) return FCFD_ERR_ARG;
    c->m3 = *s;
    return FCFD_OK;
}

extern "C" int fcfd_m3_amurho(fcfd_ctx *c) {
    M3_CTX(c);
    return fcfd::m3_props(c, c->f[FCFD_PHI], 1);
}

extern "C" int fcfd_m3_ls_ccd(fcfd_ctx *c) {
    M3_CTX(c);
    return fcfd::m3_ls_ccd(c);
}

extern "C" int fcfd_m3_ls_rk3_solver(fcfd_ctx *c) {
    M3_CTX(c);
    return fcfd::m3_ls_rk3(c);
}

extern "C" int fcfd_m3_ls_redistance(fcfd_ctx *c) {
    M3_CTX(c);
    return fcfd::m3_redistance(c, c->f[FCFD_PHI]);
}

extern "C" int fcfd_m3_mpls(fcfd_ctx *c, double im) {
    M3_CTX(c);
    return fcfd::m3_mpls(c, im);
}

extern "C" int fcfd_m3_solve_ls(fcfd_ctx *c) {
    M3_CTX(c);
    return fcfd::m3_solve_ls(c);
}

extern "C" int fcfd_m3_solve_ns(fcfd_ctx *c) {
    M3_CTX(c);
    return fcfd::m3_solve_ns(c);
}

extern "C" int fcfd_m3_find_mass(fcfd_ctx *c) {
    M3_CTX(c);
    return fcfd::m3_find_mass(c);
}

extern "C" int fcfd_m3_store_data(fcfd_ctx *c) {
    M3_CTX(c);
    return fcfd::m3_store_data(c);
}

extern "C" int fcfd_m3_setup(fcfd_ctx *c) {
    M3_CTX(c);
    double **F = c->f;
    FCFD_TRY(fcfd::bc_scalar(c, F[FCFD_PHI]));
    FCFD_TRY(fcfd::bc_scalar(c, F[FCFD_U]));
    FCFD_TRY(fcfd::bc_scalar(c, F[FCFD_V]));
    FCFD_TRY(fcfd::bc_scalar(c, F[FCFD_W]));
    FCFD_TRY(fcfd::m3_redistance(c, F[FCFD_PHI]));
    c->time = c->iter * c->pr.dt;
    FCFD_TRY(fcfd::m3_props(c, F[FCFD_PHI], 1));
    FCFD_TRY(fcfd::m3_store_data(c));
    c->m3.rec_mass = 0;
    FCFD_TRY(fcfd::m3_solve_ns(c));
    FCFD_TRY(fcfd::m3_store_data(c));
    return fcfd::m3_find_mass(c);
}

extern "C" int fcfd_m3_step(fcfd_ctx *c) {
    M3_CTX(c);
    c->iter = c->iter + 1;
    c->time = c->time + c->pr.dt;
    FCFD_TRY(fcfd::m3_solve_ls(c));
    int rc = fcfd::m3_solve_ns(c);
    if (rc != FCFD_OK && rc != FCFD_ERR_NOT_CONVERGED) return rc;
    FCFD_TRY(fcfd::m3_find_mass(c));
    FCFD_TRY(fcfd::m3_store_data(c));
    return rc;
}
