// levelset.hip — level-set kernels of the two-phase step on MI355X:
//   * ls_funs / rho_mu / ls_mv        (`.../interface_solver/others.f90:56-176`)
//   * WENO5 flux-split advection with SSP-RK3
//                                     (`.../interface_solver/level_set_rk3_solver.f90:1-219`)
//   * PDE reinitialisation with the mass-constrained lambda
//                                     (`.../interface_solver/level_set_rk3_redis.f90:1-410`)
//   * mass-preserving correction      (`.../interface_solver/mass_preserving_level_set.f90:1-54`)
// Paths are under `Two phase flows/newOpenMP/3D+IBM/` (2D twins under `2D/`).
//
// Layout: reference column-major ghosted arrays.  One thread per interior cell,
// 64 threads along i (the contiguous axis) per wavefront, so every stencil
// load of a wavefront is a contiguous 512-byte run.  Intermediate arrays that
// the reference fills and then ghost-copies with bc() (the flux split s1/s2,
// the backward differences, the lambda numerator/denominator) are recomputed
// or read with clamped indices instead of being stored with ghosts.
#include <algorithm>
#include <mutex>

#include "fcfd_internal.h"
#include "weno.h"

using namespace fcfd;

namespace fcfd {

constexpr double kPi = 3.141592653589793;  // dacos(-1.0_8)

// ls_funs for one value (others.f90:62-87)
__device__ __forceinline__ void lsfun(double phi, double w, double &H, double &D) {
    const double eps = 1.0e-12;
    double x = phi / w;
    if (x > 1.0 - eps) {
        H = 1.0; D = 0.0;
    } else if (x < -1.0 + eps) {
        H = 0.0; D = 0.0;
    } else {
        double s, c;
        sincos(kPi * x, &s, &c);
        H = 0.5 * (1.0 + x + s / kPi);
        D = 0.5 * (1.0 + c) / w;
    }
    H = fmax(fmin(H, 1.0), 0.0);
}

// ------------------------------------------------------------- properties
// ls_funs + rho_mu over the full ghosted box (others.f90:56-120)
// (cp != NULL: a batched 2D stack, case t / slab)
__global__ void k_ls_props(const double *__restrict__ phi, double *__restrict__ H, double *__restrict__ D,
                           double *__restrict__ S, double *__restrict__ R, double *__restrict__ M, long n,
                           double w, double r12, double m12, const fcfd_case *__restrict__ cp, long slab) {
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    if (cp) {
        const fcfd_case &q = cp[t / slab];
        r12 = q.rho_12;
        m12 = q.mu_12;
    }
    double h, d;
    lsfun(phi[t], w, h, d);
    if (H) H[t] = h;
    if (D) D[t] = d;
    if (S) S[t] = 2.0 * h - 1.0;
    if (R) R[t] = h + r12 * (1.0 - h);
    if (M) M[t] = h + m12 * (1.0 - h);
}

int ls_props(fcfd_ctx *c, const double *phi, double *H, double *D, double *S, double *rho, double *mu) {
    long n = c->G.ntot;
    ProfScope ps_(c, R_MISC);
    hipLaunchKernelGGL(k_ls_props, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, phi, H, D, S, rho, mu,
                       n, c->pr.ls_wid, c->pr.rho_12, c->pr.mu_12, c->cases_dev, c->G.sxy);
    FCFD_CHECK(hipGetLastError());
    return FCFD_OK;
}

// ------------------------------------------------------------- reductions
// Deterministic sums over the interior: a fixed grid of RBLK blocks, each
// thread accumulates a fixed strided subset in order, then fixed-shape tree
// reductions in LDS; a second single-block pass sums the block partials.
constexpr int RBLK = 1024;
constexpr int RTHR = 256;


template <int NS>
__device__ __forceinline__ void block_sum_store(double (&acc)[NS], double *partial, int nblocks) {
    __shared__ double sh[NS][RTHR];
    for (int s = 0; s < NS; ++s) sh[s][threadIdx.x] = acc[s];
    __syncthreads();
    for (int w = RTHR / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int s = 0; s < NS; ++s) sh[s][threadIdx.x] += sh[s][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int s = 0; s < NS; ++s) partial[s * nblocks + blockIdx.x] = sh[s][0];
}

// second pass, one block of RTHR threads per case q (blockIdx.x):
// result[q*RES_STRIDE + s] = sum_b partial[s*nb*ncase + q*nb + b]
__global__ void k_reduce_final(const double *__restrict__ partial, double *__restrict__ result, int ns, int nb) {
    __shared__ double sh[RTHR];
    const long ntot = (long)nb * gridDim.x;
    partial += (long)blockIdx.x * nb;
    result += (long)blockIdx.x * RES_STRIDE;
    for (int s = 0; s < ns; ++s) {
        double a = 0.0;
        for (int b = threadIdx.x; b < nb; b += RTHR) a += partial[s * ntot + b];
        sh[threadIdx.x] = a;
        __syncthreads();
        for (int w = RTHR / 2; w > 0; w >>= 1) {
            if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
            __syncthreads();
        }
        if (threadIdx.x == 0) result[s] = sh[0];
        __syncthreads();
    }
}

// mass / volume (ls_mv, others.f90:129-152) and, with nrm != null, the MPLS
// lambda denominator sum delta^2 |grad phi| (2(1-rho12)H + rho12) dV with
// |grad phi| from the CCD normals, written to grad
// (mass_preserving_level_set.f90:17-34).  blockIdx.y is the case of a
// batched 2D stack (its slice; cp holds its rho_12), 0 otherwise; each case
// reduces over its own RBLK blocks in the same order as a single case.
template <int DIM>
__global__ void k_mass_sums(const double *__restrict__ phi, const double *__restrict__ nx_,
                            const double *__restrict__ ny_, const double *__restrict__ nz_,
                            double *__restrict__ grad, double *__restrict__ partial, Geom G, double w,
                            double r12, int with_lam, const fcfd_case *__restrict__ cp) {
    double acc[3] = {0.0, 0.0, 0.0};
    const int q = blockIdx.y;
    if (cp) r12 = cp[q].rho_12;
    const long n = DIM == 3 ? (long)G.nx * G.ny * G.nz : (long)G.nx * G.ny;
    const double dv = DIM == 3 ? G.dx * G.dy * G.dz : G.dx * G.dy;
    for (long t = (long)blockIdx.x * RTHR + threadIdx.x; t < n; t += (long)RBLK * RTHR) {
        int i, j, k;
        cell_of_index(t, G, i, j, k);
        if (DIM == 2) k = q;
        long id = G.id(i, j, k);
        double h, d;
        lsfun(phi[id], w, h, d);
        double rho = h + r12 * (1.0 - h);
        acc[0] += rho * h * dv;
        acc[1] += h * dv;
        if (with_lam) {
            double fx = nx_[id], fy = ny_[id], g;
            if (DIM == 3) {
                double fz = nz_[id];
                g = sqrt(fx * fx + fy * fy + fz * fz);
            } else {
                g = sqrt(fx * fx + fy * fy);
            }
            grad[id] = g;
            double plam = d * d * g;
            plam = DIM == 3 ? plam * (2.0 * (1.0 - r12) * h + r12) * G.dx * G.dy * G.dz
                            : plam * (2.0 * (1.0 - r12) * h + r12) * G.dx * G.dy;
            acc[2] += plam;
        }
    }
    block_sum_store<3>(acc, partial + (long)q * RBLK, RBLK * gridDim.y);
}

// phi += lambda delta |grad phi|, lambda = (imass - mass) / denominator
// computed here from the reduced sums (mass_preserving_level_set.f90:36-48)
// (per case of a batched stack: its sums, its imass from cp, its lambda slot)
template <int DIM>
__global__ void k_mpls_update(double *__restrict__ phi, const double *__restrict__ grad,
                              const double *__restrict__ sums, double *__restrict__ lam_out, Geom G, double w,
                              double imass, const fcfd_case *__restrict__ cp) {
    int bx_, by_, bz_;
    xcd_block(bx_, by_, bz_);
    int i = bx_ * blockDim.x + threadIdx.x, j = by_ * blockDim.y + threadIdx.y, k = bz_;
    const int q = case_of<DIM>(k);
    sums += (long)q * RES_STRIDE;
    if (cp) imass = cp[q].imass;
    double lam = (imass - sums[0]) / sums[2];
    if (i == 0 && j == 0 && (DIM == 2 || k == 0)) lam_out[(long)q * RES_STRIDE] = lam;
    if (i >= G.nx || j >= G.ny) return;
    long id = G.id(i, j, k);
    double h, d;
    lsfun(phi[id], w, h, d);
    phi[id] = phi[id] + lam * d * grad[id];
}

int mass_sums(fcfd_ctx *c, const double *phi, int with_lam) {
    const Geom &G = c->G;
    ProfScope ps_(c, R_MASS);
    if (G.dim == 3)
        hipLaunchKernelGGL(k_mass_sums<3>, dim3(RBLK), dim3(RTHR), 0, c->stream, phi, c->f[FCFD_NX], c->f[FCFD_NY],
                           c->f[FCFD_NZ], c->f[FCFD_GRAD], c->red.partial, G, c->pr.ls_wid, c->pr.rho_12, with_lam,
                           (const fcfd_case *)nullptr);
    else
        hipLaunchKernelGGL(k_mass_sums<2>, dim3(RBLK, c->nb), dim3(RTHR), 0, c->stream, phi, c->f[FCFD_NX],
                           c->f[FCFD_NY], nullptr, c->f[FCFD_GRAD], c->red.partial, G, c->pr.ls_wid, c->pr.rho_12,
                           with_lam, c->cases_dev);
    FCFD_CHECK(hipGetLastError());
    hipLaunchKernelGGL(k_reduce_final, dim3(c->nb), dim3(RTHR), 0, c->stream, c->red.partial, c->red.result, 3,
                       RBLK);
    FCFD_CHECK(hipGetLastError());
    return FCFD_OK;
}

int fetch_result(fcfd_ctx *c, int n) {
    FCFD_CHECK(hipMemcpyAsync(c->red.host, c->red.result, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
    FCFD_CHECK(hipStreamSynchronize(c->stream));
    return FCFD_OK;
}

// ls_mv; mass and vol of case 0 (a batched stack leaves every case's pair at
// red.host[q*RES_STRIDE], [q*RES_STRIDE + 1])
int ls_mv(fcfd_ctx *c, double *mass, double *vol) {
    FCFD_TRY(ls_props(c, c->f[FCFD_PHI], c->f[FCFD_HEAVY], c->f[FCFD_DELTA], c->f[FCFD_SIGN], c->f[FCFD_RHO],
                      c->f[FCFD_MU]));
    FCFD_TRY(mass_sums(c, c->f[FCFD_PHI], 0));
    FCFD_TRY(fetch_result(c, RES_STRIDE * c->nb));
    c->diag.mass = c->red.host[0];
    c->diag.vol = c->red.host[1];
    if (mass) *mass = c->red.host[0];
    if (vol) *vol = c->red.host[1];
    return FCFD_OK;
}

// ------------------------------------------------ ls_mv without allocation
// fcfd_ls_mv_noalloc: the kernels of ls_mv (k_ls_props, k_mass_sums,
// k_reduce_final) on caller-owned fields, with their reduction partials,
// reduced sums and per-case ratios in one file-scope device array instead of
// context memory.  Every call on a device shares that device's copy of the
// array, so calls that overlap on one device (two streams, two threads) race;
// the array keeps the last call's contents, and nothing in it is read before
// this call has written it.
constexpr int NOALLOC_MAX_CASES = FCFD_LS_MV_NOALLOC_MAX_CASES;
constexpr long NA_PARTIAL = 3L * RBLK * NOALLOC_MAX_CASES;            // k_mass_sums block partials
constexpr long NA_RESULT = (long)RES_STRIDE * NOALLOC_MAX_CASES;      // k_reduce_final output
constexpr long NA_CASES = (long)(sizeof(fcfd_case) / sizeof(double)) * NOALLOC_MAX_CASES;
constexpr long NA_WORDS = NA_PARTIAL + NA_RESULT + NA_CASES;          // 1.58 MB
static __device__ __align__(256) double scratch[NA_WORDS];

struct CaseRatios {
    double r12[NOALLOC_MAX_CASES], m12[NOALLOC_MAX_CASES];
};

__global__ void k_put_cases(CaseRatios cr, int n, fcfd_case *__restrict__ dst) {
    const int q = threadIdx.x;
    if (q >= n) return;
    fcfd_case c{};
    c.rho_12 = cr.r12[q];
    c.mu_12 = cr.m12[q];
    dst[q] = c;
}

// sums[2q], sums[2q + 1] = mass, volume of case q
__global__ void k_put_sums(const double *__restrict__ result, int n, double *__restrict__ sums) {
    const int t = threadIdx.x;
    if (t >= 2 * n) return;
    sums[t] = result[(long)(t >> 1) * RES_STRIDE + (t & 1)];
}

}  // namespace fcfd

extern "C" int fcfd_ls_mv_noalloc(const fcfd_grid *grid, const fcfd_params *params, int ncases,
                                  const fcfd_case *cases, const double *phi, double *heavy, double *delta,
                                  double *sign, double *rho, double *mu, double *sums, void *stream) {
    using namespace fcfd;
    if (!grid || !params || !phi || !sums) return FCFD_ERR_ARG;
    if (grid->dim != 2 && grid->dim != 3) return FCFD_ERR_ARG;
    if (grid->ghc != 3 || grid->nx < 4 || grid->ny < 4) return FCFD_ERR_ARG;
    if (grid->dim == 3 && (grid->nz < 4 || ncases != 1 || cases)) return FCFD_ERR_ARG;
    if (grid->dim == 2 && grid->nz != 1) return FCFD_ERR_ARG;
    if (ncases < 1) return FCFD_ERR_ARG;
    if (ncases > NOALLOC_MAX_CASES) {
        fprintf(stderr, "libfcfd: fcfd_ls_mv_noalloc: %d cases exceed its fixed scratch of %d cases\n", ncases,
                NOALLOC_MAX_CASES);
        return FCFD_ERR_ARG;
    }
    if (ncases > 1 && !cases) return FCFD_ERR_ARG;
    // the device address of this device's copy of the scratch array, looked
    // up once per device (the array's name is not a device address on the host)
    static std::mutex mu_;
    static double *dev_scratch[64] = {};
    int dev = 0;
    FCFD_CHECK(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64) return FCFD_ERR_ARG;
    double *base;
    {
        std::lock_guard<std::mutex> lk(mu_);
        if (!dev_scratch[dev]) {
            void *ptr = nullptr;
            FCFD_CHECK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
            dev_scratch[dev] = (double *)ptr;
        }
        base = dev_scratch[dev];
    }
    double *partial = base, *result = base + NA_PARTIAL;
    fcfd_case *cp = nullptr;
    hipStream_t st = (hipStream_t)stream;
    Geom G;
    G.dim = grid->dim;
    G.nx = grid->nx; G.ny = grid->ny; G.nz = grid->dim == 3 ? grid->nz : ncases;
    G.g = grid->ghc; G.gz = grid->dim == 3 ? grid->ghc : 0;
    G.sx = G.nx + 2 * G.g;
    G.sxy = G.sx * (G.ny + 2 * G.g);
    G.ntot = G.sxy * (G.nz + 2 * G.gz);
    G.dx = (grid->xend - grid->xstart) / G.nx;
    G.dy = (grid->yend - grid->ystart) / G.ny;
    G.dz = grid->dim == 3 ? (grid->zend - grid->zstart) / G.nz : 1.0;
    if (ncases > 1) {
        CaseRatios cr{};
        for (int q = 0; q < ncases; ++q) {
            cr.r12[q] = cases[q].rho_12;
            cr.m12[q] = cases[q].mu_12;
        }
        cp = (fcfd_case *)(base + NA_PARTIAL + NA_RESULT);
        hipLaunchKernelGGL(k_put_cases, dim3(1), dim3(NOALLOC_MAX_CASES), 0, st, cr, ncases, cp);
        FCFD_CHECK(hipGetLastError());
    }
    const long n = G.ntot;
    hipLaunchKernelGGL(k_ls_props, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, phi, heavy, delta, sign,
                       rho, mu, n, params->ls_wid, params->rho_12, params->mu_12, (const fcfd_case *)cp, G.sxy);
    FCFD_CHECK(hipGetLastError());
    if (G.dim == 3)
        hipLaunchKernelGGL(k_mass_sums<3>, dim3(RBLK), dim3(RTHR), 0, st, phi, (const double *)nullptr,
                           (const double *)nullptr, (const double *)nullptr, (double *)nullptr, partial, G,
                           params->ls_wid, params->rho_12, 0, (const fcfd_case *)nullptr);
    else
        hipLaunchKernelGGL(k_mass_sums<2>, dim3(RBLK, ncases), dim3(RTHR), 0, st, phi, (const double *)nullptr,
                           (const double *)nullptr, (const double *)nullptr, (double *)nullptr, partial, G,
                           params->ls_wid, params->rho_12, 0, (const fcfd_case *)cp);
    FCFD_CHECK(hipGetLastError());
    hipLaunchKernelGGL(k_reduce_final, dim3(ncases), dim3(RTHR), 0, st, (const double *)partial, result, 3, RBLK);
    FCFD_CHECK(hipGetLastError());
    hipLaunchKernelGGL(k_put_sums, dim3(1), dim3(2 * NOALLOC_MAX_CASES), 0, st, (const double *)result, ncases,
                       sums);
    FCFD_CHECK(hipGetLastError());
    return FCFD_OK;
}

namespace fcfd {

int mpls(fcfd_ctx *c) {
    const Geom &G = c->G;
    double *phi = c->f[FCFD_PHI];
    dim3 blk(64, 4), grd = grid_interior(G, blk);
    double *lam_dev = c->red.result + RES_LAMBDA;
    for (int it = 0; it < 2; ++it) {
        // surface_norms: only the first derivatives enter the correction
        FCFD_TRY(ccd_dir(c, 0, phi, c->f[FCFD_NX], c->cx));
        FCFD_TRY(ccd_dir(c, 1, phi, c->f[FCFD_NY], c->cy));
        if (G.dim == 3) FCFD_TRY(ccd_dir(c, 2, phi, c->f[FCFD_NZ], c->cz));
        FCFD_TRY(mass_sums(c, phi, 1));
        {
            ProfScope ps_(c, R_MPLS_UPD);
            if (G.dim == 3)
                hipLaunchKernelGGL(k_mpls_update<3>, grd, blk, 0, c->stream, phi, c->f[FCFD_GRAD], c->red.result,
                                   lam_dev, G, c->pr.ls_wid, c->pr.imass, (const fcfd_case *)nullptr);
            else
                hipLaunchKernelGGL(k_mpls_update<2>, grd, blk, 0, c->stream, phi, c->f[FCFD_GRAD], c->red.result,
                                   lam_dev, G, c->pr.ls_wid, c->pr.imass, c->cases_dev);
            FCFD_CHECK(hipGetLastError());
        }
        FCFD_TRY(bc_scalar(c, phi));
    }
    return FCFD_OK;
}

// ------------------------------------------------------------- advection
constexpr int LS_KC = 16;   // planes per block of the 3D stage

// One SSP-RK3 stage of level_set_rk3_solver, fused: velocity flux split
// (source_setup :67-96, with the bc() copy of s1/s2 realised by clamped
// indices), WENO5-JS faces (:162-219), flux divergence (:145-158) and the
// increment-form stage update (:13-60), with every face computed once.
// T = float is the single-precision variant (same operation order, FP32
// weights of weno.h); the velocity is vs (nu, nv, nw).
// A block covers a 64 x 4 tile and, in 3D, walks LS_KC planes of it (2D:
// the block's slice, a case of a batched stack).  Per plane its 256 threads
// compute the 4 x 65 x-faces and 5 x 64 y-faces of the tile into LDS (580
// faces where two faces per cell and direction take 1024), then each cell
// differences them.  The z flux of face k+1/2 is the k-1/2 flux of the next
// plane, from a 7-plane window of (w, phi) splits carried in registers.  The
// LDS plane is double buffered, one barrier per plane.  Face value F =
// WENO_left(s1 at f-2..f+2) + WENO_right(s2 at f-1..f+3) at clamped
// positions; L sums F(+1/2) - F(-1/2) per direction.
template <int DIM, typename T>
__global__ void __launch_bounds__(256) k_ls_adv_stage_kf(const T *__restrict__ phi, T *__restrict__ phin,
                                                         const T *__restrict__ nu, const T *__restrict__ nv,
                                                         const T *__restrict__ nw, T *__restrict__ l1,
                                                         T *__restrict__ l2, Geom G, T dt, T vs, int stage) {
    // Products are fused into adds only within one statement here: the flux
    // splits s1, s2 of the z window are formed in the plane that loads them,
    // and with fusion across statements a face whose splits were formed in
    // the same plane (the first face of a walk) rounded differently from the
    // same face formed from carried splits.  A face now has one value
    // wherever a walk starts, which keeps block-by-block runs bit for bit
    // equal to whole-field ones.
#pragma clang fp contract(on)
    constexpr int TX = 64, TY = 4, NFX = TY * (TX + 1), NFY = (TY + 1) * TX;
    __shared__ T fxs[2][TY][TX + 1], fys[2][TY + 1][TX];
    int bx_, by_, bz_;
    xcd_block(bx_, by_, bz_);
    const int tx = threadIdx.x, ty = threadIdx.y, t = ty * TX + tx;
    const int i0 = bx_ * TX, j0 = by_ * TY, i = i0 + tx, j = j0 + ty;
    // 3D: LS_KC planes per block; 2D: the block's slice (a case of a batched stack)
    const int k0 = DIM == 3 ? bz_ * LS_KC : bz_, k1 = DIM == 3 ? min(k0 + LS_KC, G.nz) : bz_ + 1;
    const bool in = i < G.nx && j < G.ny;
    const int ic = min(i, G.nx - 1), jc = min(j, G.ny - 1);
    const T half = (T)0.5, EPS = (T)1.0e-13;
    const T hx = (T)G.dx, hy = (T)G.dy, hz = (T)G.dz;
    // face f + 1/2 of a line: positions clamp(f - 2 + m), m = 0..5, into
    // [-lo, n - 1 + hi] (lo, hi = 0 on physical faces), at stride st from the
    // line start a0
    auto face = [&](const T *__restrict__ vel, long a0, long st, int f, int lo, int hi) -> T {
        T s1[6], s2[6];
#pragma unroll
        for (int m = 0; m < 6; ++m) {
            const long a = a0 + (long)clampi(f - 2 + m, lo, hi) * st;
            const T u = vs * vel[a], p = phi[a];
            s1[m] = half * (u + fabs(u)) * p;
            s2[m] = half * (u - fabs(u)) * p;
        }
        return weno_left(s1[0], s1[1], s1[2], s1[3], s1[4], EPS) + weno_right(s2[1], s2[2], s2[3], s2[4], s2[5], EPS);
    };
    // z window of this column: positions clamp(k - 3 + m), m = 0..6
    T s1z[7], s2z[7], Fl = 0;
    if (DIM == 3) {
#pragma unroll
        for (int m = 0; m < 7; ++m) {
            const long a = G.id(ic, jc, clampi(k0 - 3 + m, -G.oz0, G.nz - 1 + G.oz1));
            const T u = vs * nw[a], p = phi[a];
            s1z[m] = half * (u + fabs(u)) * p;
            s2z[m] = half * (u - fabs(u)) * p;
        }
        Fl = weno_left(s1z[0], s1z[1], s1z[2], s1z[3], s1z[4], EPS) +
             weno_right(s2z[1], s2z[2], s2z[3], s2z[4], s2z[5], EPS);
    }
    for (int k = k0; k < k1; ++k) {
        const int b = k & 1;
        for (int e = t; e < NFX + NFY; e += TX * TY) {
            if (e < NFX) {   // x-face i0 - 1 + c + 1/2 of row j0 + r
                const int r = e / (TX + 1), c = e - r * (TX + 1);
                fxs[b][r][c] = face(nu, G.id(0, min(j0 + r, G.ny - 1), k), 1, i0 - 1 + c, -G.ox0, G.nx - 1 + G.ox1);
            } else {         // y-face j0 - 1 + r + 1/2 of column i0 + c
                const int q = e - NFX, r = q / TX, c = q - r * TX;
                fys[b][r][c] = face(nv, G.id(min(i0 + c, G.nx - 1), 0, k), G.sx, j0 - 1 + r, -G.oy0,
                                    G.ny - 1 + G.oy1);
            }
        }
        __syncthreads();
        // face k + 1/2: gm from s1(k-2..k+2), fp from s2(k-1..k+3)
        T Fr = 0;
        if (DIM == 3)
            Fr = weno_left(s1z[1], s1z[2], s1z[3], s1z[4], s1z[5], EPS) +
                 weno_right(s2z[2], s2z[3], s2z[4], s2z[5], s2z[6], EPS);
        if (in) {
            const long id = G.id(i, j, k);
            T L = -(fxs[b][ty][tx + 1] - fxs[b][ty][tx]) / hx;
            L = L + (-(fys[b][ty + 1][tx] - fys[b][ty][tx]) / hy);
            if (DIM == 3) L = L + (-(Fr - Fl) / hz);
            T src;
            if (stage == 0) {
                l1[id] = L;
                src = L;
            } else if (stage == 1) {
                l2[id] = L;
                src = ((T)-3.0 * l1[id] + L) / (T)4.0;
            } else {
                src = (-l1[id] - l2[id] + (T)8.0 * L) / (T)12.0;
            }
            phin[id] = phi[id] + src * dt;
        }
        if (DIM == 3) {
            Fl = Fr;
            // slide the z window to plane k + 1
#pragma unroll
            for (int m = 0; m < 6; ++m) {
                s1z[m] = s1z[m + 1];
                s2z[m] = s2z[m + 1];
            }
            const long a = G.id(ic, jc, clampi(k + 4, -G.oz0, G.nz - 1 + G.oz1));
            const T u = vs * nw[a], p = phi[a];
            s1z[6] = half * (u + fabs(u)) * p;
            s2z[6] = half * (u - fabs(u)) * p;
        }
    }
}

// One SSP-RK3 stage with the UCCD source L = -(u phi_x + v phi_y + w phi_z)
// (the reference's commented alternative source,
// `.../interface_solver/level_set_rk3_solver.f90:155-157`), from the upwind
// CCD derivatives fx, fy, fz of the stage input.
template <int DIM>
__global__ void __launch_bounds__(256) k_ls_uccd_stage(const double *__restrict__ phi, double *__restrict__ phin,
                                                       const double *__restrict__ nu, const double *__restrict__ nv,
                                                       const double *__restrict__ nw, const double *__restrict__ fx,
                                                       const double *__restrict__ fy, const double *__restrict__ fz,
                                                       double *__restrict__ l1, double *__restrict__ l2, Geom G,
                                                       double dt, int stage) {
    int bx_, by_, bz_;
    xcd_block(bx_, by_, bz_);
    const int i = bx_ * blockDim.x + threadIdx.x, j = by_ * blockDim.y + threadIdx.y, k = bz_;
    if (i >= G.nx || j >= G.ny) return;
    const long id = G.id(i, j, k);
    double L = -nu[id] * fx[id] - nv[id] * fy[id];
    if (DIM == 3) L = L - nw[id] * fz[id];
    double src;
    if (stage == 0) {
        l1[id] = L;
        src = L;
    } else if (stage == 1) {
        l2[id] = L;
        src = (-3.0 * l1[id] + L) / 4.0;
    } else {
        src = (-l1[id] - l2[id] + 8.0 * L) / 12.0;
    }
    phin[id] = phi[id] + src * dt;
}

// launch grid of k_ls_adv_stage_kf: LS_KC planes per block
static dim3 grid_km(const Geom &G, dim3 blk) {
    return dim3((G.nx + blk.x - 1) / blk.x, (G.ny + blk.y - 1) / blk.y, (G.nz + LS_KC - 1) / LS_KC);
}

// one WENO5 SSP-RK3 stage from phi into phin with the context's velocity
// (the stage sources stay in l1/l2 as in ls_advect); no ghost update
int ls_adv_stage(fcfd_ctx *c, const double *phi, double *phin, int stage) {
    const Geom &G = c->G;
    dim3 blk(64, 4), grd = grid_interior(G, blk);
    if (G.dim == 3)
        hipLaunchKernelGGL((k_ls_adv_stage_kf<3, double>), grid_km(G, blk), blk, 0, c->stream, phi, phin,
                           c->f[FCFD_NU], c->f[FCFD_NV], c->f[FCFD_NW], c->l1, c->l2, G, c->pr.dt, c->vscale, stage);
    else
        hipLaunchKernelGGL((k_ls_adv_stage_kf<2, double>), grd, blk, 0, c->stream, phi, phin, c->f[FCFD_NU],
                           c->f[FCFD_NV], nullptr, c->l1, c->l2, G, c->pr.dt, c->vscale, stage);
    FCFD_CHECK(hipGetLastError());
    return FCFD_OK;
}

int ls_adv_stage_geom(hipStream_t st, const Geom &G, const double *phi, double *phin, const double *nu,
                      const double *nv, const double *nw, double *l1, double *l2, double dt, int stage) {
    dim3 blk(64, 4);
    if (G.dim == 3)
        hipLaunchKernelGGL((k_ls_adv_stage_kf<3, double>), grid_km(G, blk), blk, 0, st, phi, phin, nu, nv, nw, l1, l2,
                           G, dt, 1.0, stage);
    else
        hipLaunchKernelGGL((k_ls_adv_stage_kf<2, double>), grid_interior(G, blk), blk, 0, st, phi, phin, nu, nv,
                           nullptr, l1, l2, G, dt, 1.0, stage);
    FCFD_CHECK(hipGetLastError());
    return FCFD_OK;
}

int ls_advect(fcfd_ctx *c) {
    const Geom &G = c->G;
    dim3 blk(64, 4), grd = grid_interior(G, blk);
    double *a = c->f[FCFD_PHI], *b = c->phiB;
    if (c->pr.ls_scheme == FCFD_LS_UCCD) {
        if (c->vscale != 1.0) {
            fprintf(stderr, "libfcfd: the velocity scale applies to the WENO5 scheme only\n");
            return FCFD_ERR_ARG;
        }
        // upwind-CCD derivatives park in the predictor's spare velocity buffers
        double *d[3] = {c->un, c->vn, c->wn};
        const double *vel[3] = {c->f[FCFD_NU], c->f[FCFD_NV], c->f[FCFD_NW]};
        for (int s = 0; s < 3; ++s) {
            for (int dir = 0; dir < G.dim; ++dir) FCFD_TRY(uccd_dir(c, dir, a, vel[dir], d[dir], nullptr));
            {
                ProfScope ps_(c, R_LS_ADV);
                if (G.dim == 3)
                    hipLaunchKernelGGL(k_ls_uccd_stage<3>, grd, blk, 0, c->stream, a, b, vel[0], vel[1], vel[2], d[0],
                                       d[1], d[2], c->l1, c->l2, G, c->pr.dt, s);
                else
                    hipLaunchKernelGGL(k_ls_uccd_stage<2>, grd, blk, 0, c->stream, a, b, vel[0], vel[1], nullptr, d[0],
                                       d[1], nullptr, c->l1, c->l2, G, c->pr.dt, s);
                FCFD_CHECK(hipGetLastError());
            }
            FCFD_TRY(bc_scalar(c, b));
            double *t = a; a = b; b = t;
        }
        c->phiB = c->f[FCFD_PHI];
        c->f[FCFD_PHI] = a;
        return FCFD_OK;
    }
    for (int s = 0; s < 3; ++s) {
        {
            ProfScope ps_(c, R_LS_ADV);
            if (G.dim == 3)
                hipLaunchKernelGGL((k_ls_adv_stage_kf<3, double>), grid_km(G, blk), blk, 0, c->stream, a, b,
                                   c->f[FCFD_NU], c->f[FCFD_NV], c->f[FCFD_NW], c->l1, c->l2, G, c->pr.dt, c->vscale,
                                   s);
            else
                hipLaunchKernelGGL((k_ls_adv_stage_kf<2, double>), grd, blk, 0, c->stream, a, b, c->f[FCFD_NU],
                                   c->f[FCFD_NV], nullptr, c->l1, c->l2, G, c->pr.dt, c->vscale, s);
            FCFD_CHECK(hipGetLastError());
        }
        FCFD_TRY(bc_scalar(c, b));
        double *t = a; a = b; b = t;
    }
    // three stages end in the buffer that was phiB: swap roles
    c->phiB = c->f[FCFD_PHI];
    c->f[FCFD_PHI] = a;
    return FCFD_OK;
}

// one FP32 WENO5 SSP-RK3 stage from phi into phin with caller arrays (the
// velocity times the context's scale); no ghost update
int ls_adv_stage_f32(fcfd_ctx *c, const float *phi, float *phin, const float *nu, const float *nv, const float *nw,
                     float *l1, float *l2, int stage) {
    const Geom &G = c->G;
    dim3 blk(64, 4);
    const float dt = (float)c->pr.dt, vs = (float)c->vscale;
    if (G.dim == 3)
        hipLaunchKernelGGL((k_ls_adv_stage_kf<3, float>), grid_km(G, blk), blk, 0, c->stream, phi, phin, nu, nv, nw,
                           l1, l2, G, dt, vs, stage);
    else
        hipLaunchKernelGGL((k_ls_adv_stage_kf<2, float>), grid_interior(G, blk), blk, 0, c->stream, phi, phin, nu,
                           nv, (const float *)nullptr, l1, l2, G, dt, vs, stage);
    FCFD_CHECK(hipGetLastError());
    return FCFD_OK;
}

// One SSP-RK3 step in FP32 of a caller's level set phi with the caller's
// FP32 cell velocity times the context's scale: stage 0 phi -> B, stage 1
// B -> C, stage 2 C -> phi (each stage reads its input and writes another
// array), the zero-gradient ghost copy after each stage as in ls_advect.
int ls_advect_f32(fcfd_ctx *c, float *phi, const float *nu, const float *nv, const float *nw) {
    const Geom &G = c->G;
    if (!c->adv32[0])
        for (int q = 0; q < 4; ++q) FCFD_CHECK(hipMalloc((void **)&c->adv32[q], sizeof(float) * G.ntot));
    float *B = c->adv32[0], *Cb = c->adv32[1], *l1 = c->adv32[2], *l2 = c->adv32[3];
    dim3 blk(64, 4), grd = grid_interior(G, blk);
    const float dt = (float)c->pr.dt, vs = (float)c->vscale;
    float *in[3] = {phi, B, Cb}, *out[3] = {B, Cb, phi};
    for (int s = 0; s < 3; ++s) {
        {
            ProfScope ps_(c, R_LS_ADV);
            if (G.dim == 3)
                hipLaunchKernelGGL((k_ls_adv_stage_kf<3, float>), grid_km(G, blk), blk, 0, c->stream, in[s], out[s], nu,
                                   nv, nw, l1,
                                   l2, G, dt, vs, s);
            else
                hipLaunchKernelGGL((k_ls_adv_stage_kf<2, float>), grd, blk, 0, c->stream, in[s], out[s], nu, nv,
                                   (const float *)nullptr, l1, l2, G, dt, vs, s);
            FCFD_CHECK(hipGetLastError());
        }
        FCFD_TRY(bc_scalar_f32(c, out[s]));
    }
    return FCFD_OK;
}

// ------------------------------------------------------------- reinit
// Pass A of one reinitialisation stage: backward differences of phi with
// the bc() ghost copy as clamped indices (redis_gradient :219-233), one-sided
// WENO5 values, EPS = 1e-10 (wenojs_flux :295-339), Godunov |grad phi| by the
// frozen sign (:262-291), then the lambda numerator/denominator terms
// (redis_lambda :366-380; 2D: lambda = delta, 3D: density weighted).
template <int DIM>
__global__ void __launch_bounds__(256) k_redis_a(const double *__restrict__ phi, const double *__restrict__ S0,
                                                 double *__restrict__ grad, double *__restrict__ s2,
                                                 double *__restrict__ s3, const unsigned long long *__restrict__ done,
                                                 Geom G, double w, double r12, int btn) {
    int bx_, by_, bz_;
    xcd_block(bx_, by_, bz_);
    const int i = bx_ * blockDim.x + threadIdx.x, j = by_ * blockDim.y + threadIdx.y, k = bz_;
    if (done && done[(long)case_of<DIM>(k) * FLAG_STRIDE + 1]) return;
    if (i >= G.nx || j >= G.ny) return;
    const long id = G.id(i, j, k);
    const int n[3] = {G.nx, G.ny, G.nz};
    const int pos[3] = {i, j, k};
    const long st[3] = {1, G.sx, G.sxy};
    const double ih[3] = {1.0 / G.dx, 1.0 / G.dy, 1.0 / G.dz};
    double sel[3];
    // phi at pos-4 .. pos+3 of the three lines, all loads before any branch
    // (pos-4 is clamped to the outermost ghost layer; that value is never
    // selected below)
    double gw[3][8];
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
        gw[d][0] = phi[id + (long)(max(pos[d] - 4, -3) - pos[d]) * st[d]];
#pragma unroll
        for (int r = 1; r < 8; ++r) gw[d][r] = phi[id + (long)(r - 4) * st[d]];
    }
    const bool pos_sign = S0[id] >= 0.0;
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
        // D-phi at pos-3 .. pos+3, ghost entries copying the boundary one:
        // nr[m] = D(clamp(pos - 3 + m, 0, n - 1)) = dw[clamp(m, 3 - pos, n + 2 - pos)]
        // with dw[m] = D(pos - 3 + m) the window's neighbour differences
        double nr[7];
#pragma unroll
        for (int m = 0; m < 7; ++m) nr[m] = (gw[d][m + 1] - gw[d][m]) * ih[d];
        if (pos[d] < 3 || pos[d] > n[d] - 4) clamp_window(nr, 3 - pos[d], n[d] + 2 - pos[d]);
        const double EPS = 1.0e-10;
        double fm = weno_left(nr[1], nr[2], nr[3], nr[4], nr[5], EPS);   // tdata%s2
        double fp = weno_right(nr[2], nr[3], nr[4], nr[5], nr[6], EPS);  // tdata%s1
        if (pos_sign) sel[d] = fmax(-fmin(fp, 0.0), fmax(fm, 0.0));
        else sel[d] = fmax(fmax(fp, 0.0), -fmin(fm, 0.0));
    }
    double g = DIM == 3 ? sqrt(sel[0] * sel[0] + sel[1] * sel[1] + sel[2] * sel[2])
                        : sqrt(sel[0] * sel[0] + sel[1] * sel[1]);
    grad[id] = g;
    if (btn != 0) {
        double hh, dd;
        lsfun(phi[id], w, hh, dd);
        double lam = DIM == 3 ? (2.0 * (1.0 - r12) * hh + r12) * dd : dd;
        s2[id] = lam * S0[id] * (g - 1.0);
        s3[id] = g * dd * lam;
    }
}

// Pass B: 3^d weighted sums (centre weight 16 in 2D, 51 in 3D, bc() copy as
// clamped indices), correction s1 (redis_lambda :385-408), stage source and
// in-place increment-form SSP-RK3 update (redis_solver :150-208).  The last
// stage also folds max|phi - phi_start| into the flag shards (reinit stop
// test :42-54).
template <int DIM>
__global__ void __launch_bounds__(256) k_redis_b(double *__restrict__ phi, const double *__restrict__ S0,
                                                 const double *__restrict__ grad, const double *__restrict__ s2,
                                                 const double *__restrict__ s3, double *__restrict__ l1,
                                                 double *__restrict__ l2, const double *__restrict__ phis,
                                                 unsigned long long *__restrict__ flags, Geom G, double w,
                                                 double rdt, int btn, int stage) {
    int bx_, by_, bz_;
    xcd_block(bx_, by_, bz_);
    const int i = bx_ * blockDim.x + threadIdx.x, j = by_ * blockDim.y + threadIdx.y, k = bz_;
    flags += (long)case_of<DIM>(k) * FLAG_STRIDE;   // this case's flag block (a block covers one slice)
    if (flags[1]) return;
    bool in = i < G.nx && j < G.ny;
    double err = 0.0;
    if (in) {
        const long id = G.id(i, j, k);
        double hh, dd;
        double p0 = phi[id];
        lsfun(p0, w, hh, dd);
        double s1 = 0.0;
        if (btn != 0) {
            const double wc = DIM == 3 ? 51.0 : 16.0;
            double a = wc * s2[id], b = wc * s3[id];
            const int kl = DIM == 3 ? -1 : 0, kh = DIM == 3 ? 1 : 0;
            for (int kk = kl; kk <= kh; ++kk) {
                int kq = DIM == 3 ? clampi(k + kk, 0, G.nz - 1) : k;
                for (int jj = -1; jj <= 1; ++jj) {
                    int jq = clampi(j + jj, 0, G.ny - 1);
                    for (int ii = -1; ii <= 1; ++ii) {
                        int iq = clampi(i + ii, 0, G.nx - 1);
                        long o = G.id(iq, jq, kq);
                        a = a + s2[o];
                        b = b + s3[o];
                    }
                }
            }
            if (fabs(b) > 1.0e-12) s1 = a / b * dd;
        }
        double S = S0[id];
        double l = (S - s1) * grad[id] - S, src;
        if (stage == 0) { l1[id] = l; src = l; }
        else if (stage == 1) { l2[id] = l; src = (-3.0 * l1[id] + l) / 4.0; }
        else { src = (-l1[id] - l2[id] + 8.0 * l) / 12.0; }
        double pn = p0 - rdt * src;
        phi[id] = pn;
        if (stage == 2) err = fabs(phis[id] - pn);
    }
    if (stage == 2) {
        // block max, then one atomic per block into one of NSHARD slots (a
        // single slot serialises tens of thousands of atomics); max of
        // non-negative doubles is order independent, compared as bit patterns
        __shared__ double sh[4];
        const double m = block_max256(err, sh);
        if (threadIdx.x + threadIdx.y == 0) {
            const unsigned b = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
            atomicMax(&flags[FLAG_SHARD0 + b % NSHARD], (unsigned long long)__double_as_longlong(m));
        }
    }
}

// Pass B in 3D, RB_KC planes per thread.  The constraint term is
// (a / b) delta, and delta = 0 off the interface band, so only band cells
// (a thin shell, a few % of the cells) read the 3x3x3 sums of s2 and s3,
// as 3x3 plane sums of planes k-1, k, k+1 (clamped).  Same update as
// k_redis_b; the sums associate differently (plane by plane), a change of
// rounding only.  The stage and btn are template arguments so that every
// load of a plane is issued at the top of its iteration.
constexpr int RB_KC = 16;
template <int STAGE, bool BTN>
__global__ void __launch_bounds__(256) k_redis_b_km(double *__restrict__ phi, const double *__restrict__ S0,
                                                    const double *__restrict__ grad, const double *__restrict__ s2,
                                                    const double *__restrict__ s3, double *__restrict__ l1,
                                                    double *__restrict__ l2, const double *__restrict__ phis,
                                                    unsigned long long *__restrict__ flags, Geom G, double w,
                                                    double rdt) {
    int bx_, by_, bz_;
    xcd_block(bx_, by_, bz_);
    const int i = bx_ * blockDim.x + threadIdx.x, j = by_ * blockDim.y + threadIdx.y;
    const int k0 = bz_ * RB_KC, k1 = min(k0 + RB_KC, G.nz);
    if (flags[1]) return;
    const bool in = i < G.nx && j < G.ny;
    double err = 0.0;
    if (in) {
        const int im = max(i - 1, 0), ip = min(i + 1, G.nx - 1), jm = max(j - 1, 0), jp = min(j + 1, G.ny - 1);
        const long c0 = G.id(i, j, 0);
        const long o[9] = {G.id(im, jm, 0) - c0, G.id(i, jm, 0) - c0, G.id(ip, jm, 0) - c0,
                           G.id(im, j, 0) - c0,  0,                    G.id(ip, j, 0) - c0,
                           G.id(im, jp, 0) - c0, G.id(i, jp, 0) - c0,  G.id(ip, jp, 0) - c0};
        auto plane = [&](long c, double &a, double &b) {
            double t2[9], t3[9];
#pragma unroll
            for (int q = 0; q < 9; ++q) {
                t2[q] = s2[c + o[q]];
                t3[q] = s3[c + o[q]];
            }
            a = 0.0;
            b = 0.0;
#pragma unroll
            for (int q = 0; q < 9; ++q) {
                a += t2[q];
                b += t3[q];
            }
        };
        long id = G.id(i, j, k0);
        for (int k = k0; k < k1; ++k, id += G.sxy) {
            // every load of the plane first
            const double p0 = phi[id], S = S0[id], gr = grad[id];
            const double l1o = STAGE >= 1 ? l1[id] : 0.0, l2o = STAGE == 2 ? l2[id] : 0.0;
            const double pst = STAGE == 2 ? phis[id] : 0.0;
            double hh, dd;
            lsfun(p0, w, hh, dd);
            // the constraint enters as (a / b) delta: only cells of the
            // interface band (delta != 0) read the 27-point sums
            double s1 = 0.0;
            if (BTN && dd != 0.0) {
                double am, bm, ac, bc, ap = 0.0, bp = 0.0;
                plane(id - (k > 0 ? G.sxy : 0), am, bm);
                plane(id, ac, bc);
                if (k + 1 <= G.nz - 1) plane(id + G.sxy, ap, bp);
                else { ap = ac; bp = bc; }
                const double a = 51.0 * s2[id] + am + ac + ap, b = 51.0 * s3[id] + bm + bc + bp;
                if (fabs(b) > 1.0e-12) s1 = a / b * dd;
            }
            const double l = (S - s1) * gr - S;
            double src;
            if (STAGE == 0) { l1[id] = l; src = l; }
            else if (STAGE == 1) { l2[id] = l; src = (-3.0 * l1o + l) / 4.0; }
            else { src = (-l1o - l2o + 8.0 * l) / 12.0; }
            const double pn = p0 - rdt * src;
            phi[id] = pn;
            if (STAGE == 2) err = fmax(err, fabs(pst - pn));
        }
    }
    if (STAGE == 2) {
        __shared__ double sh[4];
        const double m = block_max256(err, sh);
        if (threadIdx.x + threadIdx.y == 0) {
            const unsigned b = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
            atomicMax(&flags[FLAG_SHARD0 + b % NSHARD], (unsigned long long)__double_as_longlong(m));
        }
    }
}

static void launch_redis_b_km(fcfd_ctx *c, dim3 g, dim3 blk, double *phi, const double *grad, int btn, int stage) {
    const Geom &G = c->G;
    const double w = c->pr.ls_wid, rdt = c->pr.rdt;
#define FCFD_RB(S_, B_)                                                                                              \
    hipLaunchKernelGGL((k_redis_b_km<S_, B_>), g, blk, 0, c->stream, phi, c->s0, grad, c->s2, c->s3, c->l1, c->l2, \
                       c->phiS, c->flag, G, w, rdt)
    if (btn) {
        if (stage == 0) FCFD_RB(0, true);
        else if (stage == 1) FCFD_RB(1, true);
        else FCFD_RB(2, true);
    } else {
        if (stage == 0) FCFD_RB(0, false);
        else if (stage == 1) FCFD_RB(1, false);
        else FCFD_RB(2, false);
    }
#undef FCFD_RB
}

// end of one pseudo step: count it, stop when the max over the shards is
// <= tol, clear the shards (one block of NSHARD threads per case)
__global__ void __launch_bounds__(NSHARD) k_redis_check(unsigned long long *flags, double tol) {
    __shared__ double sh[4];
    flags += (long)blockIdx.x * FLAG_STRIDE;
    if (flags[1]) return;
    const int t = threadIdx.x;
    const double e = __longlong_as_double((long long)flags[FLAG_SHARD0 + t]);
    const double err = block_max256(e, sh);
    flags[FLAG_SHARD0 + t] = 0ull;
    if (t == 0) {
        flags[2] += 1;
        if (err <= tol) flags[1] = 1;
    }
}

__global__ void k_copy_interior(double *__restrict__ dst, const double *__restrict__ src,
                                const unsigned long long *__restrict__ done, Geom G) {
    int bx_, by_, bz_;
    xcd_block(bx_, by_, bz_);
    const int i = bx_ * blockDim.x + threadIdx.x, j = by_ * blockDim.y + threadIdx.y, k = bz_;
    if (done && done[(G.dim == 2 ? (long)k : 0L) * FLAG_STRIDE + 1]) return;
    if (i >= G.nx || j >= G.ny) return;
    long id = G.id(i, j, k);
    dst[id] = src[id];
}

__global__ void k_sign_init(const double *__restrict__ phi, double *__restrict__ S0, Geom G, double w, int btn) {
    int bx_, by_, bz_;
    xcd_block(bx_, by_, bz_);
    const int i = bx_ * blockDim.x + threadIdx.x, j = by_ * blockDim.y + threadIdx.y, k = bz_;
    if (i >= G.nx || j >= G.ny) return;
    long id = G.id(i, j, k);
    if (btn == 0) {
        S0[id] = phi[id];
    } else {
        double h, d;
        lsfun(phi[id], w, h, d);
        S0[id] = 2.0 * h - 1.0;
    }
}

// max of grad over the interior into flags[3] (level_set_redis_stable :113-122)
__global__ void k_max_interior(const double *__restrict__ a, unsigned long long *flags, Geom G) {
    int bx_, by_, bz_;
    xcd_block(bx_, by_, bz_);
    const int i = bx_ * blockDim.x + threadIdx.x, j = by_ * blockDim.y + threadIdx.y, k = bz_;
    __shared__ double sh[4];
    double v = 0.0;
    if (i < G.nx && j < G.ny) v = a[G.id(i, j, k)];
    v = block_max256(v, sh);
    const long q = G.dim == 2 ? k : 0;   // the case of a batched stack
    if (threadIdx.x + threadIdx.y == 0)
        atomicMax(&flags[q * FLAG_STRIDE + 3], (unsigned long long)__double_as_longlong(v));
}

__global__ void k_scale_by_flag(double *__restrict__ phi, const unsigned long long *flags, Geom G) {
    int bx_, by_, bz_;
    xcd_block(bx_, by_, bz_);
    const int i = bx_ * blockDim.x + threadIdx.x, j = by_ * blockDim.y + threadIdx.y, k = bz_;
    if (i >= G.nx || j >= G.ny) return;
    double gmax = __longlong_as_double((long long)flags[(G.dim == 2 ? (long)k : 0L) * FLAG_STRIDE + 3]);
    long id = G.id(i, j, k);
    phi[id] = phi[id] / gmax;
}

int redis(fcfd_ctx *c, int btn, int *iters) {
    const Geom &G = c->G;
    dim3 blk(64, 4), grd = grid_interior(G, blk);
    double *phi = c->f[FCFD_PHI], *grad = c->f[FCFD_GRAD];
    const double w = c->pr.ls_wid, r12 = c->pr.rho_12, rdt = c->pr.rdt;
    FCFD_CHECK(hipMemsetAsync(c->flag, 0, (size_t)FLAG_STRIDE * c->nb * sizeof(unsigned long long), c->stream));
    // level_set_redis_init / stable (:67-136)
    hipLaunchKernelGGL(k_sign_init, grd, blk, 0, c->stream, phi, c->s0, G, w, btn);
    if (btn == 0) {
        if (G.dim == 3)
            hipLaunchKernelGGL(k_redis_a<3>, grd, blk, 0, c->stream, phi, c->s0, grad, c->s2, c->s3, nullptr, G, w,
                               r12, 0);
        else
            hipLaunchKernelGGL(k_redis_a<2>, grd, blk, 0, c->stream, phi, c->s0, grad, c->s2, c->s3, nullptr, G, w,
                               r12, 0);
        hipLaunchKernelGGL(k_max_interior, grd, blk, 0, c->stream, grad, c->flag, G);
        hipLaunchKernelGGL(k_scale_by_flag, grd, blk, 0, c->stream, phi, c->flag, G);
        FCFD_TRY(bc_scalar(c, phi));
    }
    // pseudo-time budget: the reference stops once the accumulated time exceeds
    // timestop (:18-28,54); replay that accumulation to fix the launch count
    double timestop;
    if (btn == 0) {
        double m = c->gd.xend - c->gd.xstart;
        m = fmax(m, c->gd.yend - c->gd.ystart);
        if (G.dim == 3) m = fmax(m, c->gd.zend - c->gd.zstart);
        timestop = 1.5 * m;
    } else {
        timestop = 3.0 * G.dx;
    }
    int nmax = 0;
    for (double t = 0.0;;) {
        nmax += 1;
        t = t + rdt;
        if (t > timestop) break;
        if (nmax > 10000000) return FCFD_ERR_ARG;
    }
    const double tol = G.dim == 3 ? 1.0e-5 : 1.0e-8;
    const dim3 gkm(grd.x, grd.y, (G.nz + RB_KC - 1) / RB_KC);
    // The stop test runs on the device (k_redis_check); a stopped case's
    // kernels return at once.  The done words of pseudo-step `it` are copied
    // to the host behind it, and before step it+2 is queued the host reads
    // them, so it stops queueing once every case has stopped while one step
    // stays queued ahead.  Slots of flag_host past the 4 nb words read below.
    for (int q = 0; q < 2; ++q)
        if (!c->redis_ev[q]) FCFD_CHECK(hipEventCreateWithFlags(&c->redis_ev[q], hipEventDisableTiming));
    unsigned long long *poll[2] = {c->flag_host + 4 * c->nb, c->flag_host + 5 * c->nb};
    for (int it = 0; it < nmax; ++it) {
        if (it >= 2) {
            FCFD_CHECK(hipEventSynchronize(c->redis_ev[it & 1]));
            bool all = true;
            for (int q = 0; q < c->nb; ++q) all = all && poll[it & 1][q] != 0ull;
            if (all) break;
        }
        hipLaunchKernelGGL(k_copy_interior, grd, blk, 0, c->stream, c->phiS, phi, c->flag, G);
        for (int s = 0; s < 3; ++s) {
            {
                ProfScope ps_(c, R_REDIS_A);
                if (G.dim == 3)
                    hipLaunchKernelGGL(k_redis_a<3>, grd, blk, 0, c->stream, phi, c->s0, grad, c->s2, c->s3, c->flag,
                                       G, w, r12, btn);
                else
                    hipLaunchKernelGGL(k_redis_a<2>, grd, blk, 0, c->stream, phi, c->s0, grad, c->s2, c->s3, c->flag,
                                       G, w, r12, btn);
            }
            {
                ProfScope ps_(c, R_REDIS_B);
                if (G.dim == 3)
                    launch_redis_b_km(c, gkm, blk, phi, grad, btn, s);
                else
                    hipLaunchKernelGGL(k_redis_b<2>, grd, blk, 0, c->stream, phi, c->s0, grad, c->s2, c->s3, c->l1,
                                       c->l2, c->phiS, c->flag, G, w, rdt, btn, s);
            }
            FCFD_CHECK(hipGetLastError());
            FCFD_TRY(bc_scalar(c, phi));
        }
        hipLaunchKernelGGL(k_redis_check, dim3(c->nb), dim3(NSHARD), 0, c->stream, c->flag, tol);
        FCFD_TRY(gather_to_host(c, poll[it & 1], c->flag + 1, FLAG_STRIDE, 1, c->nb));
        FCFD_CHECK(hipEventRecord(c->redis_ev[it & 1], c->stream));
    }
    // words 0..3 of every case's flag block
    FCFD_TRY(gather_to_host(c, c->flag_host, c->flag, FLAG_STRIDE, 4, c->nb));
    FCFD_CHECK(hipStreamSynchronize(c->stream));
    int most = 0;   // a batched stack reports the largest pseudo-step count of its cases
    for (int q = 0; q < c->nb; ++q) most = std::max(most, (int)c->flag_host[4 * q + 2]);
    c->diag.redis_iters = most;
    if (iters) *iters = most;
    return FCFD_OK;
}

}  // namespace fcfd
