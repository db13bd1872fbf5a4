// ns.hip — momentum predictor, pressure source, projection and cell
// velocity of the two-phase step.
//
// Reference (paths under `Two phase flows/newOpenMP/3D+IBM/src/`):
//   sec_part1..3 (viscous, variable-viscosity, gravity + CSF)  ns_solver/ns_ab_diff.f90:10-198
//   ns_linearize + find_stag_vel (cross-face averages)           ns_solver/ns_solvers.f90:33-57,
//                                                                 others/bcs.f90:457-774
//   ns_ab_adv_source_sec (2nd-order upwind advection)            ns_solver/ns_ab_adv.f90:1-73
//   ns_ab_setup + ns_ab_predictor (AB2)                           ns_solver/admas_bashforth_solver.f90:230-311
//   ppe source, ns_momentum_correction                           ns_solver/ns_ppe_solver.f90:275-287,372-403
//   node_vel, ns_check_convergence_div                            ns_solver/ns_solvers.f90:59-147
// 2D twins under `2D/src/`.
#include "fcfd_internal.h"

using namespace fcfd;

namespace fcfd {

constexpr double kPi = 3.141592653589793;

__device__ __forceinline__ void lsfun2(double phi, double w, double &H, double &D) {
    const double eps = 1.0e-12;
    double x = phi / w;
    if (x > 1.0 - eps) { H = 1.0; D = 0.0; }
    else if (x < -1.0 + eps) { H = 0.0; D = 0.0; }
    else {
        double s, c;
        sincos(kPi * x, &s, &c);
        H = 0.5 * (1.0 + x + s / kPi);
        D = 0.5 * (1.0 + c) / w;
    }
    H = fmax(fmin(H, 1.0), 0.0);
}

struct Phys {
    double re, m12, r12, w;      // Reynolds, mu ratio, rho ratio, interface width
    double g[3], btn_g, fr, btn_sf, we, dt;
};

// sec_part1..3 evaluated once per step: they read only old-level fields
// (rho, mu, delta of phi_old; vel_old; curvature and CCD normals of the
// previous ns_init), so the reference's four Picard re-evaluations all give
// this same value.  rho/mu/delta are recomputed from phi_old with the ls_funs
// formula instead of being stored.
template <int DIM>
__global__ void __launch_bounds__(256) k_diff_source(const double *__restrict__ Po, const double *__restrict__ Uo,
                                                     const double *__restrict__ Vo, const double *__restrict__ Wo,
                                                     const double *__restrict__ Co, const double *__restrict__ NXo,
                                                     const double *__restrict__ NYo, const double *__restrict__ NZo,
                                                     double *__restrict__ Tx, double *__restrict__ Ty,
                                                     double *__restrict__ Tz, Geom G, Phys P,
                                                     const fcfd_case *__restrict__ cp, int slab) {
    int bx_, by_, bz_;
    xcd_block_slab(bx_, by_, bz_, slab);
    const int i = bx_ * blockDim.x + threadIdx.x, j = by_ * blockDim.y + threadIdx.y, k = bz_;
    if (i >= G.nx || j >= G.ny) return;
    if (cp) {   // a batched stack: this slice's case
        const fcfd_case &q = cp[case_of<DIM>(k)];
        P.re = q.re; P.m12 = q.mu_12; P.r12 = q.rho_12; P.fr = q.fr; P.we = q.we;
    }
    const long id = G.id(i, j, k);
    const long sx = G.sx, sxy = G.sxy;
    const double dx = G.dx, dy = G.dy, dz = G.dz;
    const double dx2 = dx * dx, dy2 = dy * dy, dz2 = dz * dz;
    const double re = P.re, m12 = P.m12, r12 = P.r12;
    const long nb[3] = {1, sx, sxy};
    const double *Fs[3] = {Uo, Vo, Wo};
    const double *N[3] = {NXo, NYo, NZo};
    double *T[3] = {Tx, Ty, Tz};
    double hc, dc;
    lsfun2(Po[id], P.w, hc, dc);
    const double rc = hc + r12 * (1.0 - hc), mc = hc + m12 * (1.0 - hc);
#pragma unroll
    for (int c = 0; c < DIM; ++c) {
        double hn, dn;
        lsfun2(Po[id + nb[c]], P.w, hn, dn);
        const double rn = hn + r12 * (1.0 - hn), mn = hn + m12 * (1.0 - hn);
        const double rho = 0.5 * (rc + rn), mu = 0.5 * (mc + mn), delta = 0.5 * (dc + dn);
        const double *F = Fs[c];
        // sec_part1
        double xx = (F[id + 1] - 2.0 * F[id] + F[id - 1]) / dx2;
        double yy = (F[id + sx] - 2.0 * F[id] + F[id - sx]) / dy2;
        double v = mu / rho * xx / re + mu / rho * yy / re;
        if (DIM == 3) {
            double zz = (F[id + sxy] - 2.0 * F[id] + F[id - sxy]) / dz2;
            v = v + mu / rho * zz / re;
        }
        // sec_part2
        double dif_x, dif_y, dif_z = 0.0;
        if (c == 0) {
            double ux = 0.5 * (Uo[id + 1] - Uo[id - 1]) / dx;
            double uy = 0.5 * (Uo[id + sx] - Uo[id - sx]) / dy;
            double vx = 0.5 * (Vo[id + 1] - Vo[id] + Vo[id + 1 - sx] - Vo[id - sx]) / dx;
            double phix = (Po[id + 1] - Po[id]) / dx;
            double phiy = 0.25 * (Po[id + 1 + sx] - Po[id + 1 - sx] + Po[id + sx] - Po[id - sx]) / dy;
            dif_x = (1.0 - m12) * delta * phix * 2.0 * ux / (rho * re);
            dif_y = (1.0 - m12) * delta * phiy * (uy + vx) / (rho * re);
            if (DIM == 3) {
                double uz = 0.5 * (Uo[id + sxy] - Uo[id - sxy]) / dz;
                double wx = 0.5 * (Wo[id + 1] - Wo[id] + Wo[id + 1 - sxy] - Wo[id - sxy]) / dx;
                double phiz = 0.25 * (Po[id + 1 + sxy] - Po[id + 1 - sxy] + Po[id + sxy] - Po[id - sxy]) / dz;
                dif_z = (1.0 - m12) * delta * phiz * (uz + wx) / (rho * re);
            }
        } else if (c == 1) {
            double vx = 0.5 * (Vo[id + 1] - Vo[id - 1]) / dx;
            double vy = 0.5 * (Vo[id + sx] - Vo[id - sx]) / dy;
            double uy = 0.5 * (Uo[id + sx] - Uo[id] + Uo[id - 1 + sx] - Uo[id - 1]) / dy;
            double phix = 0.25 * (Po[id + 1] - Po[id - 1] + Po[id + 1 + sx] - Po[id - 1 + sx]) / dx;
            double phiy = (Po[id + sx] - Po[id]) / dy;
            dif_x = (1.0 - m12) * delta * phix * (uy + vx) / (rho * re);
            dif_y = (1.0 - m12) * delta * phiy * 2.0 * vy / (rho * re);
            if (DIM == 3) {
                double vz = 0.5 * (Vo[id + sxy] - Vo[id - sxy]) / dz;
                double wy = 0.5 * (Wo[id + sx] - Wo[id] + Wo[id + sx - sxy] - Wo[id - sxy]) / dy;
                double phiz = 0.25 * (Po[id + sx + sxy] - Po[id + sx - sxy] + Po[id + sxy] - Po[id - sxy]) / dz;
                dif_z = (1.0 - m12) * delta * phiz * (wy + vz) / (rho * re);
            }
        } else {
            double wx = 0.5 * (Wo[id + 1] - Wo[id - 1]) / dx;
            double wy = 0.5 * (Wo[id + sx] - Wo[id - sx]) / dy;
            double wz = 0.5 * (Wo[id + sxy] - Wo[id - sxy]) / dz;
            double uz = 0.5 * (Uo[id + sxy] - Uo[id] + Uo[id - 1 + sxy] - Uo[id - 1]) / dz;
            double vz = 0.5 * (Vo[id + sxy] - Vo[id] + Vo[id - sx + sxy] - Vo[id - sx]) / dz;
            double phix = 0.25 * (Po[id + 1 + sxy] - Po[id - 1 + sxy] + Po[id + 1] - Po[id - 1]) / dx;
            double phiy = 0.25 * (Po[id + sx + sxy] - Po[id - sx + sxy] + Po[id + sx] - Po[id - sx]) / dy;
            double phiz = (Po[id + sxy] - Po[id]) / dz;
            dif_x = (1.0 - m12) * delta * phix * (uz + wx) / (rho * re);
            dif_y = (1.0 - m12) * delta * phiy * (vz + wy) / (rho * re);
            dif_z = (1.0 - m12) * delta * phiz * 2.0 * wz / (rho * re);
        }
        v = DIM == 3 ? v + dif_x + dif_y + dif_z : v + dif_x + dif_y;
        // sec_part3
        double curv = (Co[id] + Co[id + nb[c]]) / 2.0;
        double ph = 0.5 * (N[c][id] + N[c][id + nb[c]]);
        v = v + P.g[c] * P.btn_g / P.fr - P.btn_sf * curv * delta * ph / (P.we * rho);
        T[c][id] = v;
    }
}

// ns_linearize + ns_ab_adv_source_sec + ns_ab_setup + ns_ab_predictor fused:
// cross-face averages of the Picard iterate (find_stag_vel, including the
// east/north/top wall-face zeroing of its ghost rules), second-order upwind
// advection of vel_old, S = adv + viscous/body-force source, and the AB2
// predictor u* = u_old + dt (1.5 S - 0.5 S_old).  write_src = 0 leaves S
// unstored: the Picard iterations of a step before the last, whose S the
// next iteration overwrites (only the last one's S becomes the next step's
// S_old), then move 120 instead of 144 bytes per cell.
template <int DIM, bool SETUP_ONLY>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(1, 4))) k_predictor(
    const double *__restrict__ Ut, const double *__restrict__ Vt, const double *__restrict__ Wt,
    const double *__restrict__ Uo, const double *__restrict__ Vo, const double *__restrict__ Wo,
    const double *__restrict__ Dx, const double *__restrict__ Dy, const double *__restrict__ Dz,
    const double *__restrict__ Sox, const double *__restrict__ Soy, const double *__restrict__ Soz,
    double *__restrict__ Sx, double *__restrict__ Sy, double *__restrict__ Sz, double *__restrict__ Un,
    double *__restrict__ Vn, double *__restrict__ Wn, Geom G, double dt, int zero_e, int zero_n, int zero_t,
    int write_src, int slab) {
    int bx_, by_, bz_;
    xcd_block_slab(bx_, by_, bz_, slab);
    const int i = bx_ * blockDim.x + threadIdx.x, j = by_ * blockDim.y + threadIdx.y, k = bz_;
    if (i >= G.nx || j >= G.ny) return;
    const long id = G.id(i, j, k);
    const long sx = G.sx, sxy = G.sxy;
    // the one-sided differences scaled by 1/h: three divisions per thread
    // instead of 6 per component (an FP64 division is ~10 instructions); the
    // products differ from the quotients by at most an ulp
    const double idx = 1.0 / G.dx, idy = 1.0 / G.dy, idz = 1.0 / G.dz;
    // advecting velocities (u, v, w) at each MAC face
    double adv[3][3];
    adv[0][0] = Ut[id];
    adv[0][1] = 0.25 * (Vt[id] + Vt[id - sx] + Vt[id + 1] + Vt[id + 1 - sx]);
    adv[1][0] = 0.25 * (Ut[id] + Ut[id - 1] + Ut[id + sx] + Ut[id - 1 + sx]);
    adv[1][1] = Vt[id];
    if (zero_e && i == G.nx - 1) adv[0][1] = 0.0;
    if (zero_n && j == G.ny - 1) adv[1][0] = 0.0;
    if (DIM == 3) {
        adv[0][2] = 0.25 * (Wt[id] + Wt[id - sxy] + Wt[id + 1] + Wt[id + 1 - sxy]);
        adv[1][2] = 0.25 * (Wt[id] + Wt[id - sxy] + Wt[id + sx] + Wt[id + sx - sxy]);
        adv[2][0] = 0.25 * (Ut[id] + Ut[id - 1] + Ut[id + sxy] + Ut[id - 1 + sxy]);
        adv[2][1] = 0.25 * (Vt[id] + Vt[id - sx] + Vt[id + sxy] + Vt[id - sx + sxy]);
        adv[2][2] = Wt[id];
        if (zero_e && i == G.nx - 1) adv[0][2] = 0.0;
        if (zero_n && j == G.ny - 1) adv[1][2] = 0.0;
        if (zero_t && k == G.nz - 1) { adv[2][0] = 0.0; adv[2][1] = 0.0; }
    }
    const double *F[3] = {Uo, Vo, Wo};
    const double *D[3] = {Dx, Dy, Dz};
    const double *So[3] = {Sox, Soy, Soz};
    double *S[3] = {Sx, Sy, Sz};
    double *Vn_[3] = {Un, Vn, Wn};
#pragma unroll
    for (int c = 0; c < DIM; ++c) {
        const double *g = F[c];
        const double u = adv[c][0], v = adv[c][1];
        double xp = 0.5 * (-g[id + 2] + 4.0 * g[id + 1] - 3.0 * g[id]) * idx;
        double xm = 0.5 * (g[id - 2] - 4.0 * g[id - 1] + 3.0 * g[id]) * idx;
        double yp = 0.5 * (-g[id + 2 * sx] + 4.0 * g[id + sx] - 3.0 * g[id]) * idy;
        double ym = 0.5 * (g[id - 2 * sx] - 4.0 * g[id - sx] + 3.0 * g[id]) * idy;
        double a = -((u + fabs(u)) * xm + (u - fabs(u)) * xp) / 2.0 - ((v + fabs(v)) * ym + (v - fabs(v)) * yp) / 2.0;
        if (DIM == 3) {
            const double w = adv[c][2];
            double zp = 0.5 * (-g[id + 2 * sxy] + 4.0 * g[id + sxy] - 3.0 * g[id]) * idz;
            double zm = 0.5 * (g[id - 2 * sxy] - 4.0 * g[id - sxy] + 3.0 * g[id]) * idz;
            a = a - ((w + fabs(w)) * zm + (w - fabs(w)) * zp) / 2.0;
        }
        double s = a + D[c][id];
        if (SETUP_ONLY || write_src) S[c][id] = s;
        if (!SETUP_ONLY) {
            double src = 1.5 * s - 0.5 * So[c][id];
            Vn_[c][id] = g[id] + dt * src;
        }
    }
}

// coe%src = div(u*) / dt (ns_ppe_solver.f90:275-287)
template <int DIM>
__global__ void k_ppe_src(const double *__restrict__ U, const double *__restrict__ V, const double *__restrict__ W,
                          double *__restrict__ f, Geom G, double dt, int slab) {
    int bx_, by_, bz_;
    xcd_block_slab(bx_, by_, bz_, slab);
    const int i = bx_ * blockDim.x + threadIdx.x, j = by_ * blockDim.y + threadIdx.y, k = bz_;
    if (i >= G.nx || j >= G.ny) return;
    const long id = G.id(i, j, k);
    double d = (U[id] - U[id - 1]) / G.dx + (V[id] - V[id - G.sx]) / G.dy;
    if (DIM == 3) d = d + (W[id] - W[id - G.sxy]) / G.dz;
    f[id] = d / dt;
}

// ns_momentum_correction (ns_ppe_solver.f90:379-399)
template <int DIM>
__global__ void k_projection(double *__restrict__ U, double *__restrict__ V, double *__restrict__ W,
                             const double *__restrict__ P, const double *__restrict__ R, Geom G, double dt,
                             int slab) {
    int bx_, by_, bz_;
    xcd_block_slab(bx_, by_, bz_, slab);
    const int i = bx_ * blockDim.x + threadIdx.x, j = by_ * blockDim.y + threadIdx.y, k = bz_;
    if (i >= G.nx || j >= G.ny) return;
    const long id = G.id(i, j, k);
    double rho = (R[id] + R[id + 1]) / 2.0, px = (P[id + 1] - P[id]) / G.dx;
    U[id] = U[id] - dt * px / rho;
    rho = (R[id] + R[id + G.sx]) / 2.0;
    double py = (P[id + G.sx] - P[id]) / G.dy;
    V[id] = V[id] - dt * py / rho;
    if (DIM == 3) {
        rho = (R[id] + R[id + G.sxy]) / 2.0;
        double pz = (P[id + G.sxy] - P[id]) / G.dz;
        W[id] = W[id] - dt * pz / rho;
    }
}

// node_vel (ns_solvers.f90:133-143)
template <int DIM>
__global__ void k_node_vel(const double *__restrict__ U, const double *__restrict__ V, const double *__restrict__ W,
                           double *__restrict__ NU, double *__restrict__ NV, double *__restrict__ NW, Geom G,
                           int slab) {
    int bx_, by_, bz_;
    xcd_block_slab(bx_, by_, bz_, slab);
    const int i = bx_ * blockDim.x + threadIdx.x, j = by_ * blockDim.y + threadIdx.y, k = bz_;
    if (i >= G.nx || j >= G.ny) return;
    const long id = G.id(i, j, k);
    NU[id] = 0.5 * (U[id - 1] + U[id]);
    NV[id] = 0.5 * (V[id - G.sx] + V[id]);
    if (DIM == 3) NW[id] = 0.5 * (W[id - G.sxy] + W[id]);
}

// max and sum of |div u| (ns_check_convergence_div :70-88): block partials.
// A grid-stride loop over the flattened interior with 32-bit index
// arithmetic (the host keeps the cell count plus one stride below 2^32).
template <int DIM>
__global__ void __launch_bounds__(256) k_div(const double *__restrict__ U, const double *__restrict__ V,
                                             const double *__restrict__ W, double *__restrict__ partial, Geom G,
                                             int nblocks) {
    __shared__ double smax[256], ssum[256];
    double mx = 0.0, sm = 0.0;
    const unsigned nx = (unsigned)G.nx, nxy = nx * (unsigned)G.ny, n = nxy * (unsigned)G.nz;
    for (unsigned t = blockIdx.x * 256u + threadIdx.x; t < n; t += (unsigned)nblocks * 256u) {
        const unsigned k = t / nxy, r = t - k * nxy, j = r / nx, i = r - j * nx;
        const long id = G.id((int)i, (int)j, (int)k);
        double d = (U[id] - U[id - 1]) / G.dx + (V[id] - V[id - G.sx]) / G.dy;
        if (DIM == 3) d = d + (W[id] - W[id - G.sxy]) / G.dz;
        mx = fmax(mx, fabs(d));
        sm += fabs(d);
    }
    smax[threadIdx.x] = mx;
    ssum[threadIdx.x] = sm;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            smax[threadIdx.x] = fmax(smax[threadIdx.x], smax[threadIdx.x + w]);
            ssum[threadIdx.x] += ssum[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        partial[blockIdx.x] = smax[0];
        partial[nblocks + blockIdx.x] = ssum[0];
    }
}

static Phys phys_of(fcfd_ctx *c) {
    Phys P;
    P.re = c->pr.re; P.m12 = c->pr.mu_12; P.r12 = c->pr.rho_12; P.w = c->pr.ls_wid;
    P.g[0] = c->pr.gx; P.g[1] = c->pr.gy; P.g[2] = c->pr.gz;
    P.btn_g = c->pr.btn_g; P.fr = c->pr.fr; P.btn_sf = c->pr.btn_sf; P.we = c->pr.we; P.dt = c->pr.dt;
    return P;
}

// viscous / body-force source from the start-of-step fields
int ns_diff_source(fcfd_ctx *c) {
    const Geom &G = c->G;
    double **F = c->f;
    dim3 blk = tile_blk(c, FCFD_TILE_DIFF_SOURCE), grd = grid_interior(G, blk);
    Phys P = phys_of(c);
    ProfScope ps_(c, R_DIFF);
    if (G.dim == 3)
        hipLaunchKernelGGL(k_diff_source<3>, grd, blk, 0, c->stream, F[FCFD_PHI_OLD], F[FCFD_U_OLD], F[FCFD_V_OLD],
                           F[FCFD_W_OLD], F[FCFD_CURV], F[FCFD_NX], F[FCFD_NY], F[FCFD_NZ], c->dsu, c->dsv, c->dsw,
                           G, P, (const fcfd_case *)nullptr, c->tile[FCFD_TILE_DIFF_SOURCE][2]);
    else
        hipLaunchKernelGGL(k_diff_source<2>, grd, blk, 0, c->stream, F[FCFD_PHI_OLD], F[FCFD_U_OLD], F[FCFD_V_OLD],
                           nullptr, F[FCFD_CURV], F[FCFD_NX], F[FCFD_NY], nullptr, c->dsu, c->dsv, nullptr, G, P,
                           c->cases_dev, c->tile[FCFD_TILE_DIFF_SOURCE][2]);
    FCFD_CHECK(hipGetLastError());
    return FCFD_OK;
}

// predictor: reads vel (Picard iterate), writes SRC and the predicted
// velocity into the spare buffers, then swaps them in as vel now
int ns_predictor(fcfd_ctx *c, const double *ut, const double *vt, const double *wt, int write_src) {
    const Geom &G = c->G;
    double **F = c->f;
    dim3 blk = tile_blk(c, FCFD_TILE_PREDICTOR), grd = grid_interior(G, blk);
    int ze = c->pr.ubc[1] == 1 || c->pr.ubc[1] == 2;
    int zn = c->pr.vbc[1] == 1 || c->pr.vbc[1] == 2;
    int zt = c->pr.wbc[1] == 1 || c->pr.wbc[1] == 2;
    {
        ProfScope ps_(c, R_PRED);
        if (G.dim == 3)
            hipLaunchKernelGGL((k_predictor<3, false>), grd, blk, 0, c->stream, ut, vt, wt, F[FCFD_U_OLD], F[FCFD_V_OLD],
                               F[FCFD_W_OLD], c->dsu, c->dsv, c->dsw, F[FCFD_SRC_U_OLD], F[FCFD_SRC_V_OLD],
                               F[FCFD_SRC_W_OLD], F[FCFD_SRC_U], F[FCFD_SRC_V], F[FCFD_SRC_W], c->un, c->vn, c->wn, G,
                               c->pr.dt, ze, zn, zt, write_src, c->tile[FCFD_TILE_PREDICTOR][2]);
        else
            hipLaunchKernelGGL((k_predictor<2, false>), grd, blk, 0, c->stream, ut, vt, nullptr, F[FCFD_U_OLD], F[FCFD_V_OLD],
                               nullptr, c->dsu, c->dsv, nullptr, F[FCFD_SRC_U_OLD], F[FCFD_SRC_V_OLD], nullptr,
                               F[FCFD_SRC_U], F[FCFD_SRC_V], nullptr, c->un, c->vn, nullptr, G, c->pr.dt, ze, zn, zt,
                               write_src, c->tile[FCFD_TILE_PREDICTOR][2]);
        FCFD_CHECK(hipGetLastError());
    }
    // the predicted field becomes vel now; its ghost values follow velbc
    std::swap(c->f[FCFD_U], c->un);
    std::swap(c->f[FCFD_V], c->vn);
    if (G.dim == 3) std::swap(c->f[FCFD_W], c->wn);
    return velbc(c, F[FCFD_U], F[FCFD_V], F[FCFD_W], 0);
}

// ns_ab_setup alone (problem_init): SRC = adv + diff, no predictor
int ns_setup_only(fcfd_ctx *c) {
    const Geom &G = c->G;
    double **F = c->f;
    dim3 blk = tile_blk(c, FCFD_TILE_PREDICTOR), grd = grid_interior(G, blk);
    int ze = c->pr.ubc[1] == 1 || c->pr.ubc[1] == 2;
    int zn = c->pr.vbc[1] == 1 || c->pr.vbc[1] == 2;
    int zt = c->pr.wbc[1] == 1 || c->pr.wbc[1] == 2;
    if (G.dim == 3)
        hipLaunchKernelGGL((k_predictor<3, true>), grd, blk, 0, c->stream, F[FCFD_U], F[FCFD_V], F[FCFD_W], F[FCFD_U_OLD],
                           F[FCFD_V_OLD], F[FCFD_W_OLD], c->dsu, c->dsv, c->dsw, F[FCFD_SRC_U_OLD],
                           F[FCFD_SRC_V_OLD], F[FCFD_SRC_W_OLD], F[FCFD_SRC_U], F[FCFD_SRC_V], F[FCFD_SRC_W],
                           c->un, c->vn, c->wn, G, c->pr.dt, ze, zn, zt, 1, c->tile[FCFD_TILE_PREDICTOR][2]);
    else
        hipLaunchKernelGGL((k_predictor<2, true>), grd, blk, 0, c->stream, F[FCFD_U], F[FCFD_V], nullptr, F[FCFD_U_OLD],
                           F[FCFD_V_OLD], nullptr, c->dsu, c->dsv, nullptr, F[FCFD_SRC_U_OLD], F[FCFD_SRC_V_OLD],
                           nullptr, F[FCFD_SRC_U], F[FCFD_SRC_V], nullptr, c->un, c->vn, nullptr, G, c->pr.dt, ze,
                           zn, zt, 1, c->tile[FCFD_TILE_PREDICTOR][2]);
    FCFD_CHECK(hipGetLastError());
    return FCFD_OK;
}

int ppe_src(fcfd_ctx *c) {
    const Geom &G = c->G;
    double **F = c->f;
    dim3 blk = tile_blk(c, FCFD_TILE_PPE_SRC), grd = grid_interior(G, blk);
    ProfScope ps_(c, R_PPE_SRC);
    if (G.dim == 3)
        hipLaunchKernelGGL(k_ppe_src<3>, grd, blk, 0, c->stream, F[FCFD_U], F[FCFD_V], F[FCFD_W], F[FCFD_PPE_SRC], G,
                           c->pr.dt, c->tile[FCFD_TILE_PPE_SRC][2]);
    else
        hipLaunchKernelGGL(k_ppe_src<2>, grd, blk, 0, c->stream, F[FCFD_U], F[FCFD_V], nullptr, F[FCFD_PPE_SRC], G,
                           c->pr.dt, c->tile[FCFD_TILE_PPE_SRC][2]);
    FCFD_CHECK(hipGetLastError());
    return FCFD_OK;
}

int projection(fcfd_ctx *c) {
    const Geom &G = c->G;
    double **F = c->f;
    dim3 blk = tile_blk(c, FCFD_TILE_PROJECTION), grd = grid_interior(G, blk);
    {
        ProfScope ps_(c, R_PROJ);
        if (G.dim == 3)
            hipLaunchKernelGGL(k_projection<3>, grd, blk, 0, c->stream, F[FCFD_U], F[FCFD_V], F[FCFD_W], F[FCFD_P],
                               F[FCFD_RHO], G, c->pr.dt, c->tile[FCFD_TILE_PROJECTION][2]);
        else
            hipLaunchKernelGGL(k_projection<2>, grd, blk, 0, c->stream, F[FCFD_U], F[FCFD_V], nullptr, F[FCFD_P],
                               F[FCFD_RHO], G, c->pr.dt, c->tile[FCFD_TILE_PROJECTION][2]);
        FCFD_CHECK(hipGetLastError());
    }
    return velbc(c, F[FCFD_U], F[FCFD_V], F[FCFD_W], 0);
}

int node_vel(fcfd_ctx *c) {
    const Geom &G = c->G;
    double **F = c->f;
    dim3 blk = tile_blk(c, FCFD_TILE_NODE_VEL), grd = grid_interior(G, blk);
    {
        ProfScope ps_(c, R_NODE_VEL);
        if (G.dim == 3)
            hipLaunchKernelGGL(k_node_vel<3>, grd, blk, 0, c->stream, F[FCFD_U], F[FCFD_V], F[FCFD_W], F[FCFD_NU],
                               F[FCFD_NV], F[FCFD_NW], G, c->tile[FCFD_TILE_NODE_VEL][2]);
        else
            hipLaunchKernelGGL(k_node_vel<2>, grd, blk, 0, c->stream, F[FCFD_U], F[FCFD_V], nullptr, F[FCFD_NU],
                               F[FCFD_NV], nullptr, G, c->tile[FCFD_TILE_NODE_VEL][2]);
        FCFD_CHECK(hipGetLastError());
    }
    return velbc(c, F[FCFD_NU], F[FCFD_NV], F[FCFD_NW], 1);
}

// the block partials of k_div reduced in a fixed order into
// out[0] = max |div u|, out[1] = mean |div u|
__global__ void k_div_final(const double *__restrict__ partial, int nblocks, double inv_ncell,
                            double *__restrict__ out) {
    __shared__ double smax[256], ssum[256];
    double mx = 0.0, sm = 0.0;
    for (int b = threadIdx.x; b < nblocks; b += 256) {
        mx = fmax(mx, partial[b]);
        sm += partial[nblocks + b];
    }
    smax[threadIdx.x] = mx;
    ssum[threadIdx.x] = sm;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            smax[threadIdx.x] = fmax(smax[threadIdx.x], smax[threadIdx.x + w]);
            ssum[threadIdx.x] += ssum[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[0] = smax[0];
        out[1] = ssum[0] * inv_ncell;
    }
}

// ns_check_convergence_div: max and mean |div u| over the interior (of every
// case of a batched stack together) into result slots RES_VDIV, RES_VDIV + 1
// of case 0; with to_host the diagnostics are read back at once, otherwise
// they travel with the step's end-of-step gather.
int div_check(fcfd_ctx *c, int to_host) {
    const Geom &G = c->G;
    double **F = c->f;
    // 8 blocks per CU: with one block per CU the loop's dependent loads left
    // the launch latency-bound (222 us at 256^3)
    const int nb = 2048;
    if ((double)G.nx * G.ny * G.nz + (double)nb * 256 >= 4294967296.0) return FCFD_ERR_ARG;
    if (G.dim == 3)
        hipLaunchKernelGGL(k_div<3>, dim3(nb), dim3(256), 0, c->stream, F[FCFD_U], F[FCFD_V], F[FCFD_W],
                           c->red.partial, G, nb);
    else
        hipLaunchKernelGGL(k_div<2>, dim3(nb), dim3(256), 0, c->stream, F[FCFD_U], F[FCFD_V], nullptr,
                           c->red.partial, G, nb);
    FCFD_CHECK(hipGetLastError());
    hipLaunchKernelGGL(k_div_final, dim3(1), dim3(256), 0, c->stream, c->red.partial, nb,
                       1.0 / ((double)G.nx * (double)G.ny * (double)G.nz), c->red.result + RES_VDIV);
    FCFD_CHECK(hipGetLastError());
    if (!to_host) return FCFD_OK;
    double h[2];
    FCFD_CHECK(hipMemcpyAsync(h, c->red.result + RES_VDIV, sizeof(h), hipMemcpyDeviceToHost, c->stream));
    FCFD_CHECK(hipStreamSynchronize(c->stream));
    c->diag.vel_div = h[0];
    c->diag.vel_sdiv = h[1];
    return FCFD_OK;
}

}  // namespace fcfd
