// ctx.hip — context lifetime, field transfer, ghost-cell boundary conditions
// and deterministic reductions.
//
// Boundary rules follow `Two phase flows/newOpenMP/3D+IBM/src/others/bcs.f90`
// (2D: `.../2D/src/others/bcs.f90`).
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>

#include "fcfd_internal.h"

using namespace fcfd;

namespace fcfd {

// ---------------------------------------------------------------- bc(dat)
// Zero-gradient copy of all ghost layers, x then y then z (bcs.f90:113-154).
// Because every pass copies along one axis over the full extent of the other
// axes, the final value of a ghost cell is the value of the cell whose
// indices are clamped to the interior on every axis.  One thread per ghost
// cell; the ghost shell is enumerated as z slabs, then y slabs, then x slabs.
template <int DIM, typename T>
__global__ void k_bc_scalar(T *__restrict__ d, Geom G, long nzslab, long nyslab, long total) {
    // 32-bit index arithmetic: the host keeps the ghost count below 2^32
    const unsigned t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (unsigned)total) return;
    const unsigned g = G.g, X = G.nx + 2 * g, Y = G.ny + 2 * g, nzs = (unsigned)nzslab, nys = (unsigned)nyslab;
    int i, j, k;
    if (DIM == 3 && t < nzs) {                       // k in ghost layers, all i, j
        const unsigned per = X * Y, layer = t / per, r = t - layer * per, jr = r / X;
        j = (int)jr - (int)g;
        i = (int)(r - jr * X) - (int)g;
        k = (int)layer < G.gz ? (int)layer - G.gz : G.nz + ((int)layer - G.gz);
    } else if (t < nzs + nys) {                      // j in ghost layers, interior k, all i
        const unsigned u = t - nzs, per = X * 2 * g, kk = u / per, r = u - kk * per, layer = r / X;
        k = (int)kk;
        i = (int)(r - layer * X) - (int)g;
        j = layer < g ? (int)layer - (int)g : G.ny + ((int)layer - (int)g);
    } else {                                         // i in ghost layers, interior j, k
        const unsigned u = t - nzs - nys, per = (unsigned)G.ny * 2 * g, kk = u / per, r = u - kk * per;
        const unsigned jr = r / (2 * g), layer = r - jr * (2 * g);
        k = (int)kk;
        j = (int)jr;
        i = layer < g ? (int)layer - (int)g : G.nx + ((int)layer - (int)g);
    }
    // 2D: k is the case of a batched stack (0 otherwise), which has no z ghosts
    int ci = clampi(i, 0, G.nx - 1), cj = clampi(j, 0, G.ny - 1), ck = DIM == 3 ? clampi(k, 0, G.nz - 1) : k;
    d[G.id(i, j, k)] = d[G.id(ci, cj, ck)];
}

template <typename T>
static int bc_scalar_t(fcfd_ctx *c, T *d) {
    const Geom &G = c->G;
    const int g = G.g;
    long X = G.nx + 2 * g, Y = G.ny + 2 * g;
    long nzslab = G.dim == 3 ? X * Y * 2 * G.gz : 0;
    long nyslab = X * 2 * g * G.nz;
    long nxslab = (long)2 * g * G.ny * G.nz;
    long total = nzslab + nyslab + nxslab;
    int blk = 256;
    long nb = (total + blk - 1) / blk;
    if (nb * blk >= 4294967296L) return FCFD_ERR_ARG;   // the kernel numbers ghosts in 32 bits
    ProfScope ps_(c, R_BC);
    if (G.dim == 3)
        hipLaunchKernelGGL((k_bc_scalar<3, T>), dim3((unsigned)nb), dim3(blk), 0, c->stream, d, G, nzslab, nyslab, total);
    else
        hipLaunchKernelGGL((k_bc_scalar<2, T>), dim3((unsigned)nb), dim3(blk), 0, c->stream, d, G, nzslab, nyslab, total);
    FCFD_CHECK(hipGetLastError());
    return FCFD_OK;
}

int bc_scalar(fcfd_ctx *c, double *d) { return bc_scalar_t(c, d); }
int bc_scalar_f32(fcfd_ctx *c, float *d) { return bc_scalar_t(c, d); }

// ------------------------------------------------------------ velbc/nvelbc
// MAC no-slip (1) / slip (2) rules of velbc (bcs.f90:156-369) and the
// cell-centred rules of nvelbc (bcs.f90:371-455).  Three launches, x then y
// then z, each over the full extent of the two other axes like the
// reference loops, so corner ghosts end with the same values.
struct VelBC {
    int lo, hi;      // bc type on the low / high wall of this axis (1|2, else none)
    int nodal;
};

// For AX = 0 the ghost cells of a row are g consecutive words, so lane x
// covers layer x % g + 1 of row x / g: the g writes of a row leave as one
// access instead of g accesses each touching 64 rows.  Within a row only
// layer 1 of the face-normal component reads a ghost cell (itself), and
// the same lane then zeroes the wall face, so the layers are independent.
template <int AX, int DIM>
__global__ void k_velbc(double *__restrict__ u, double *__restrict__ v, double *__restrict__ w, Geom G, VelBC b) {
    // thread over the 2D set of the two other axes (full ghosted extents)
    const int g = G.g;
    int a = blockIdx.x * blockDim.x + threadIdx.x;
    int lay = 0;   // AX = 0: the one ghost layer of this lane (1..g)
    if (AX == 0) {
        lay = a % g + 1;
        a /= g;
    }
    int bb = blockIdx.y * blockDim.y + threadIdx.y;
    int na, nb;
    if (AX == 0) { na = G.ny + 2 * g; nb = G.nz + 2 * G.gz; }
    else if (AX == 1) { na = G.nx + 2 * g; nb = G.nz + 2 * G.gz; }
    else { na = G.nx + 2 * g; nb = G.ny + 2 * g; }
    if (a >= na || bb >= nb) return;
    double *comp[3] = {u, v, w};
    const int n = AX == 0 ? G.nx : (AX == 1 ? G.ny : G.nz);
    // per side: every source value first, then the ghost stores, then the
    // wall zero of the normal component (the ghosts alias the arrays they
    // are copied from, so in load-store order each load would wait for the
    // previous store); the layers ii of a lane are at most 3 (g = 3)
    constexpr int NL = AX == 0 ? 1 : 3;
    for (int side = 0; side < 2; ++side) {
        const int bt = side == 0 ? b.lo : b.hi;
        if (bt != 1 && bt != 2) continue;
        const double sg = bt == 1 ? -1.0 : 1.0;
        double val[DIM][NL];
        long dsti[DIM][NL];
#pragma unroll
        for (int cpt = 0; cpt < DIM; ++cpt) {
            const double *f = comp[cpt];
            const bool normal = (cpt == AX) && !b.nodal;   // face-normal component
#pragma unroll
            for (int q = 0; q < NL; ++q) {
                const int ii = AX == 0 ? lay : min(1 + q, g);   // (layers past g repeat the last one)
                const int dst = side == 0 ? -ii : n - 1 + ii;
                int src;
                if (side == 0) src = normal ? ii - 2 : ii - 1;
                else src = normal ? n - 1 - ii : n - ii;
                long ld, ls;
                if (AX == 0) { ld = G.id(dst, a - g, bb - G.gz); ls = G.id(src, a - g, bb - G.gz); }
                else if (AX == 1) { ld = G.id(a - g, dst, bb - G.gz); ls = G.id(a - g, src, bb - G.gz); }
                else { ld = G.id(a - g, bb - g, dst); ls = G.id(a - g, bb - g, src); }
                val[cpt][q] = f[ls];
                dsti[cpt][q] = ld;
            }
        }
#pragma unroll
        for (int cpt = 0; cpt < DIM; ++cpt)
#pragma unroll
            for (int q = 0; q < NL; ++q) comp[cpt][dsti[cpt][q]] = sg * val[cpt][q];
#pragma unroll
        for (int cpt = 0; cpt < DIM; ++cpt) {
            const bool normal = (cpt == AX) && !b.nodal;
            if (normal && (AX != 0 || lay == 1)) {
                const int wall = side == 0 ? -1 : n - 1;
                const long lw = AX == 0 ? G.id(wall, a - g, bb - G.gz)
                                        : (AX == 1 ? G.id(a - g, wall, bb - G.gz) : G.id(a - g, bb - g, wall));
                comp[cpt][lw] = 0.0;
            }
        }
    }
}

int velbc(fcfd_ctx *c, double *u, double *v, double *w, int nodal) {
    const Geom &G = c->G;
    const int g = G.g;
    dim3 blk(64, 4);
    ProfScope ps_(c, R_BC);
    for (int ax = 0; ax < G.dim; ++ax) {
        VelBC b;
        const int32_t *t = ax == 0 ? c->pr.ubc : (ax == 1 ? c->pr.vbc : c->pr.wbc);
        b.lo = t[0]; b.hi = t[1]; b.nodal = nodal;
        int na = ax == 0 ? (G.ny + 2 * g) * g : G.nx + 2 * g;   // x: one lane per ghost layer of a row
        int nb = ax == 2 ? G.ny + 2 * g : G.nz + 2 * G.gz;
        dim3 grd((na + 63) / 64, (nb + 3) / 4);
        if (G.dim == 3) {
            if (ax == 0) hipLaunchKernelGGL((k_velbc<0, 3>), grd, blk, 0, c->stream, u, v, w, G, b);
            if (ax == 1) hipLaunchKernelGGL((k_velbc<1, 3>), grd, blk, 0, c->stream, u, v, w, G, b);
            if (ax == 2) hipLaunchKernelGGL((k_velbc<2, 3>), grd, blk, 0, c->stream, u, v, w, G, b);
        } else {
            if (ax == 0) hipLaunchKernelGGL((k_velbc<0, 2>), grd, blk, 0, c->stream, u, v, w, G, b);
            if (ax == 1) hipLaunchKernelGGL((k_velbc<1, 2>), grd, blk, 0, c->stream, u, v, w, G, b);
        }
        FCFD_CHECK(hipGetLastError());
    }
    return FCFD_OK;
}

int copy_field(fcfd_ctx *c, double *dst, const double *src) {
    FCFD_CHECK(hipMemcpyAsync(dst, src, sizeof(double) * c->G.ntot, hipMemcpyDeviceToDevice, c->stream));
    return FCFD_OK;
}

}  // namespace fcfd

// ==================================================================== C ABI

static int alloc_field(double **p, long n) {
    if (hipMalloc((void **)p, sizeof(double) * n) != hipSuccess) return FCFD_ERR_HIP;
    if (hipMemset(*p, 0, sizeof(double) * n) != hipSuccess) return FCFD_ERR_HIP;
    return FCFD_OK;
}

namespace fcfd {
__global__ void k_gather8(const unsigned long long *__restrict__ src, long pitch, int width, int rows,
                          unsigned long long *__restrict__ dst) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= width * rows) return;
    dst[t] = src[(long)(t / width) * pitch + t % width];
}

int gather_to_host(fcfd_ctx *c, void *host, const void *dev, long pitch, int width, int rows) {
    if (rows == 1) {
        FCFD_CHECK(hipMemcpyAsync(host, dev, sizeof(unsigned long long) * width, hipMemcpyDeviceToHost, c->stream));
        return FCFD_OK;
    }
    if (width * rows > 16 * c->nb) return FCFD_ERR_ARG;
    hipLaunchKernelGGL(k_gather8, dim3((width * rows + 255) / 256), dim3(256), 0, c->stream,
                       (const unsigned long long *)dev, pitch, width, rows, c->pack);
    FCFD_CHECK(hipGetLastError());
    FCFD_CHECK(hipMemcpyAsync(host, c->pack, sizeof(unsigned long long) * width * rows, hipMemcpyDeviceToHost,
                              c->stream));
    return FCFD_OK;
}

// per-case parameter table of a batched stack (nothing for nb == 1, whose
// kernels take the scalars of the params)
int upload_cases(fcfd_ctx *c) {
    if (c->nb == 1) return FCFD_OK;
    if (!c->cases_dev) FCFD_CHECK(hipMalloc((void **)&c->cases_dev, sizeof(fcfd_case) * c->nb));
    FCFD_CHECK(hipMemcpyAsync(c->cases_dev, c->cases.data(), sizeof(fcfd_case) * c->nb, hipMemcpyHostToDevice,
                              c->stream));
    return FCFD_OK;
}
}  // namespace fcfd

// an ordinary context (ncases = 0) or a batched 2D stack of ncases cases
extern "C" int fcfd_ppe_fp16_density_ok(double rho_min, double rho_max) {
    // FP16's normal range is [6.104e-5, 65504].  Level 0 stores rho itself;
    // level 1 (3D) stores averaged face coefficients 2/(rho_a + rho_b), which
    // lie in [1/rho_max, 1/rho_min].  A small margin keeps the rounding of
    // those averages inside the range too.
    const double lo = 6.2e-5, hi = 6.5e4;
    if (!(rho_min > 0.0) || !(rho_max >= rho_min)) return 0;
    return rho_min >= lo && rho_max <= hi && 1.0 / rho_min <= hi && 1.0 / rho_max >= lo;
}

// params.ppe_rho_half is a request: the FP16 copies are only made when the
// density of every case, between rho_12 and 1 (rho = H + rho_12 (1 - H)),
// fits FP16 as fcfd_ppe_fp16_density_ok checks; otherwise the FP32 copies
// are read and a note is printed once per context.
static void guard_fp16(fcfd_ctx *c) {
    if (!c->pr.ppe_rho_half) return;
    double lo = std::min(1.0, c->pr.rho_12), hi = std::max(1.0, c->pr.rho_12);
    for (const fcfd_case &q : c->cases) {
        lo = std::min(lo, std::min(1.0, q.rho_12));
        hi = std::max(hi, std::max(1.0, q.rho_12));
    }
    if (!fcfd_ppe_fp16_density_ok(lo, hi)) {
        fprintf(stderr, "libfcfd: density range [%g, %g] exceeds FP16; ppe_rho_half ignored\n", lo, hi);
        c->pr.ppe_rho_half = 0;
    }
}

static int create_impl(const fcfd_grid *grid, const fcfd_params *params, int ncases, const fcfd_case *cases,
                       fcfd_ctx **out) {
    if (!grid || !params || !out) return FCFD_ERR_ARG;
    if (grid->dim != 2 && grid->dim != 3) return FCFD_ERR_ARG;
    if (grid->ghc != 3) return FCFD_ERR_ARG;
    if (grid->nx < 4 || grid->ny < 4) return FCFD_ERR_ARG;
    if (grid->dim == 3 && grid->nz < 4) return FCFD_ERR_ARG;
    if (grid->dim == 2 && grid->nz != 1) return FCFD_ERR_ARG;
    if (ncases > 0 && (grid->dim != 2 || !cases)) return FCFD_ERR_ARG;
    fcfd_ctx *c = new fcfd_ctx();
    memset(c->f, 0, sizeof(c->f));
    c->gd = *grid;
    c->pr = *params;
    c->nb = ncases > 0 ? ncases : 1;
    // 64 x 4 tiles; in 3D each XCD walks the planes of 8-row slabs, so the
    // z-neighbour planes the stencils re-read are still in its L2
    for (auto &t : c->tile) { t[0] = 64; t[1] = 4; t[2] = grid->dim == 3 ? 8 : 0; }
    if (ncases > 0) c->cases.assign(cases, cases + ncases);
    guard_fp16(c);
    Geom &G = c->G;
    G.dim = grid->dim;
    G.nx = grid->nx; G.ny = grid->ny; G.nz = ncases > 0 ? ncases : grid->nz;
    G.g = grid->ghc; G.gz = grid->dim == 3 ? grid->ghc : 0;
    G.sx = G.nx + 2 * G.g;
    G.sxy = G.sx * (G.ny + 2 * G.g);
    G.ntot = G.sxy * (G.nz + 2 * G.gz);
    G.dx = (grid->xend - grid->xstart) / G.nx;
    G.dy = (grid->yend - grid->ystart) / G.ny;
    G.dz = grid->dim == 3 ? (grid->zend - grid->zstart) / G.nz : 1.0;
    if (hipGetDevice(&c->device) != hipSuccess) { delete c; return FCFD_ERR_HIP; }
    if (hipDeviceGetAttribute(&c->ncu, hipDeviceAttributeMultiprocessorCount, c->device) != hipSuccess) {
        delete c;
        return FCFD_ERR_HIP;
    }
    if (hipStreamCreate(&c->stream) != hipSuccess) { delete c; return FCFD_ERR_HIP; }
    const long n = G.ntot;
    int rc = FCFD_OK;
    for (int f = 0; f < FCFD_NFIELDS && rc == FCFD_OK; ++f) {
        bool zf = f == FCFD_W || f == FCFD_W_OLD || f == FCFD_NW || f == FCFD_SRC_W || f == FCFD_SRC_W_OLD ||
                  f == FCFD_NZ || f == FCFD_NZZ || f == FCFD_NXZ || f == FCFD_NYZ;
        if (zf && G.dim == 2) continue;
        // the MPLS_RisingBubble fields are allocated by its first call
        if (f == FCFD_PHI1 || f == FCFD_PHI2 || f == FCFD_RHO_OLD || f == FCFD_MU_OLD) continue;
        rc = alloc_field(&c->f[f], n);
    }
    double **scr[] = {&c->phiB, &c->phiS, &c->l1, &c->l2, &c->s2, &c->s3, &c->s0, &c->cx, &c->cy,
                      &c->cz, &c->dsu, &c->dsv, &c->dsw, &c->un, &c->vn, &c->wn};
    for (auto p : scr) {
        if (rc != FCFD_OK) break;
        rc = alloc_field(p, n);
    }
    // reduction partials: 8 slots x 1024 blocks per case, or two slots per
    // block of an interior launch with 64x4 blocks, whichever is larger
    const long nbc = c->nb;
    {
        long nbi = (long)((G.nx + 63) / 64) * ((G.ny + 3) / 4) * G.nz;
        long need = std::max(8L * 1024 * nbc, 2 * nbi + 16);
        c->red.nblocks = (int)((need + 7) / 8);
    }
    // + 4 slots x 256 second-stage partials per case of the level-0 reductions
    if (rc == FCFD_OK) rc = alloc_field(&c->red.partial, 8L * c->red.nblocks + 4 * 256 * nbc);
    if (rc == FCFD_OK) c->red.partial2 = c->red.partial + 8L * c->red.nblocks;
    if (rc == FCFD_OK) rc = alloc_field(&c->red.result, RES_STRIDE * nbc);
    if (rc == FCFD_OK && hipHostMalloc((void **)&c->red.host, 64 * nbc * sizeof(double)) != hipSuccess)
        rc = FCFD_ERR_HIP;
    if (rc == FCFD_OK && hipHostMalloc((void **)&c->diag_host, DIAG_W * nbc * sizeof(double)) != hipSuccess)
        rc = FCFD_ERR_HIP;
    if (rc == FCFD_OK) memset(c->diag_host, 0, DIAG_W * nbc * sizeof(double));
    if (rc == FCFD_OK && hipMalloc((void **)&c->pack, 16 * nbc * sizeof(unsigned long long)) != hipSuccess)
        rc = FCFD_ERR_HIP;
    if (rc == FCFD_OK && hipMalloc((void **)&c->flag, FLAG_STRIDE * nbc * sizeof(unsigned long long)) != hipSuccess)
        rc = FCFD_ERR_HIP;
    if (rc == FCFD_OK &&
        hipHostMalloc((void **)&c->flag_host, FLAG_STRIDE * nbc * sizeof(unsigned long long)) != hipSuccess)
        rc = FCFD_ERR_HIP;
    if (rc == FCFD_OK) rc = upload_cases(c);
    if (rc == FCFD_OK) rc = ccd_setup(c);
    c->rho32 = nullptr;
    c->rho16 = nullptr;
    c->mg_scalars = nullptr;
    c->mg_ready = 0;
    c->time = 0.0;
    c->iter = 0;
    c->diag_pending = 0;
    memset(&c->diag, 0, sizeof(c->diag));
    c->diag.imass = params->imass;
    if (rc == FCFD_OK && hipStreamSynchronize(c->stream) != hipSuccess) rc = FCFD_ERR_HIP;
    if (rc != FCFD_OK) {
        fcfd_destroy(c);
        return rc;
    }
    *out = c;
    return FCFD_OK;
}

extern "C" int fcfd_create(const fcfd_grid *grid, const fcfd_params *params, fcfd_ctx **out) {
    return create_impl(grid, params, 0, nullptr, out);
}

extern "C" int fcfd_create_batch(const fcfd_grid *grid, const fcfd_params *params, int ncases,
                                 const fcfd_case *cases, fcfd_ctx **out) {
    if (ncases < 1) return FCFD_ERR_ARG;
    return create_impl(grid, params, ncases, cases, out);
}

extern "C" int fcfd_batch_size(const fcfd_ctx *c) { return c ? c->nb : 0; }

extern "C" int fcfd_set_cases(fcfd_ctx *c, const fcfd_case *cases) {
    if (!c || !cases || c->nb == 1) return FCFD_ERR_ARG;
    c->cases.assign(cases, cases + c->nb);
    guard_fp16(c);
    FCFD_TRY(mg_drop_graphs(c));
    return upload_cases(c);
}

extern "C" int fcfd_get_cases(const fcfd_ctx *c, fcfd_case *cases) {
    if (!c || !cases || c->nb == 1) return FCFD_ERR_ARG;
    std::copy(c->cases.begin(), c->cases.end(), cases);
    return FCFD_OK;
}

extern "C" int fcfd_destroy(fcfd_ctx *c) {
    if (!c) return FCFD_OK;
    hipStreamSynchronize(c->stream);
    for (int f = 0; f < FCFD_NFIELDS; ++f)
        if (c->f[f]) hipFree(c->f[f]);
    double *scr[] = {c->phiB, c->phiS, c->l1, c->l2, c->s2, c->s3, c->s0, c->cx, c->cy,
                     c->cz, c->dsu, c->dsv, c->dsw, c->un, c->vn, c->wn,
                     c->red.partial, c->red.result};
    for (auto p : scr)
        if (p) hipFree(p);
    for (auto p : c->ccd_coef)
        if (p) hipFree(p);
    for (auto p : c->ccd_seg)
        if (p) hipFree(p);
    if (c->ccd_z) hipFree(c->ccd_z);
    for (auto p : c->adj)
        if (p) hipFree(p);
    for (auto p : c->adv32)
        if (p) hipFree(p);
    for (auto p : c->adj32)
        if (p) hipFree(p);
    for (auto e : c->cg_ev)
        if (e) hipEventDestroy(e);
    if (c->red.host) hipHostFree(c->red.host);
    if (c->diag_host) hipHostFree(c->diag_host);
    if (c->cases_dev) hipFree(c->cases_dev);
    if (c->pack) hipFree(c->pack);
    if (c->flag) hipFree(c->flag);
    if (c->flag_host) hipHostFree(c->flag_host);
    for (hipEvent_t e : c->redis_ev)
        if (e) hipEventDestroy(e);
    for (hipEvent_t e : c->prof.pool) hipEventDestroy(e);
    mg_free(c);
    hipStreamDestroy(c->stream);
    delete c;
    return FCFD_OK;
}

extern "C" int fcfd_set_params(fcfd_ctx *c, const fcfd_params *p) {
    if (!c || !p) return FCFD_ERR_ARG;
    c->pr = *p;
    guard_fp16(c);
    c->diag.imass = p->imass;
    // captured V-cycles bake in the sweep counts: capture again on next use
    return mg_drop_graphs(c);
}

extern "C" int64_t fcfd_field_size(const fcfd_ctx *c) { return c ? (int64_t)c->G.ntot : 0; }

static double *field_or_null(fcfd_ctx *c, int f) {
    if (!c || f < 0 || f >= FCFD_NFIELDS) return nullptr;
    const bool m3f = f == FCFD_PHI1 || f == FCFD_PHI2 || f == FCFD_RHO_OLD || f == FCFD_MU_OLD;
    if (m3f && !c->f[f] && fcfd::m3_fields(c) != FCFD_OK) return nullptr;
    return c->f[f];
}

extern "C" int fcfd_push(fcfd_ctx *c, int f, const double *host) {
    double *d = field_or_null(c, f);
    if (!d) return FCFD_ERR_FIELD;
    if (!host) return FCFD_ERR_ARG;
    FCFD_CHECK(hipMemcpyAsync(d, host, sizeof(double) * c->G.ntot, hipMemcpyHostToDevice, c->stream));
    FCFD_CHECK(hipStreamSynchronize(c->stream));
    return FCFD_OK;
}

extern "C" int fcfd_pull(fcfd_ctx *c, int f, double *host) {
    double *d = field_or_null(c, f);
    if (!d) return FCFD_ERR_FIELD;
    if (!host) return FCFD_ERR_ARG;
    FCFD_CHECK(hipMemcpyAsync(host, d, sizeof(double) * c->G.ntot, hipMemcpyDeviceToHost, c->stream));
    FCFD_CHECK(hipStreamSynchronize(c->stream));
    return FCFD_OK;
}

extern "C" int fcfd_push_device(fcfd_ctx *c, int f, const double *dev) {
    double *d = field_or_null(c, f);
    if (!d) return FCFD_ERR_FIELD;
    if (!dev) return FCFD_ERR_ARG;
    FCFD_CHECK(hipMemcpyAsync(d, dev, sizeof(double) * c->G.ntot, hipMemcpyDeviceToDevice, c->stream));
    return FCFD_OK;
}

extern "C" int fcfd_pull_device(fcfd_ctx *c, int f, double *dev) {
    double *d = field_or_null(c, f);
    if (!d) return FCFD_ERR_FIELD;
    if (!dev) return FCFD_ERR_ARG;
    FCFD_CHECK(hipMemcpyAsync(dev, d, sizeof(double) * c->G.ntot, hipMemcpyDeviceToDevice, c->stream));
    return FCFD_OK;
}

extern "C" int fcfd_sync(fcfd_ctx *c) {
    if (!c) return FCFD_ERR_ARG;
    FCFD_CHECK(hipStreamSynchronize(c->stream));
    return FCFD_OK;
}

extern "C" int fcfd_get_diag(fcfd_ctx *c, fcfd_diag *out) {
    if (!c || !out) return FCFD_ERR_ARG;
    FCFD_CHECK(hipStreamSynchronize(c->stream));
    if (c->diag_pending) {
        c->diag.mass = c->diag_host[0];
        c->diag.vol = c->diag_host[1];
        c->diag.mpls_lambda = c->diag_host[RES_LAMBDA];
        c->diag.vel_div = c->diag_host[RES_VDIV];
        c->diag.vel_sdiv = c->diag_host[RES_VDIV + 1];
        c->diag_pending = 0;
    }
    *out = c->diag;
    out->iter = c->iter;
    out->imass = c->pr.imass;
    return FCFD_OK;
}

extern "C" int fcfd_get_case_diag(fcfd_ctx *c, double *mass, double *vol, double *imass) {
    if (!c) return FCFD_ERR_ARG;
    FCFD_CHECK(hipStreamSynchronize(c->stream));
    for (int q = 0; q < c->nb; ++q) {
        if (mass) mass[q] = c->diag_host[(long)q * DIAG_W];
        if (vol) vol[q] = c->diag_host[(long)q * DIAG_W + 1];
        if (imass) imass[q] = c->nb > 1 ? c->cases[q].imass : c->pr.imass;
    }
    return FCFD_OK;
}

extern "C" int fcfd_set_time(fcfd_ctx *c, double time, int iter) {
    if (!c) return FCFD_ERR_ARG;
    c->time = time;
    c->iter = iter;
    return FCFD_OK;
}

extern "C" int fcfd_set_tile(fcfd_ctx *c, int kernel, int tx, int ty, int slab) {
    if (!c || kernel < 0 || kernel >= FCFD_NTILE) return FCFD_ERR_ARG;
    if ((tx != 64 && tx != 128 && tx != 256) || ty < 1 || tx * ty > 256) return FCFD_ERR_ARG;
    if (slab < 0 || slab > 4096) return FCFD_ERR_ARG;
    c->tile[kernel][0] = tx;
    c->tile[kernel][1] = ty;
    c->tile[kernel][2] = slab;
    return FCFD_OK;
}

extern "C" int fcfd_get_tile(const fcfd_ctx *c, int kernel, int *tx, int *ty, int *slab) {
    if (!c || kernel < 0 || kernel >= FCFD_NTILE || !tx || !ty) return FCFD_ERR_ARG;
    *tx = c->tile[kernel][0];
    *ty = c->tile[kernel][1];
    if (slab) *slab = c->tile[kernel][2];
    return FCFD_OK;
}

extern "C" int fcfd_set_velocity_scale(fcfd_ctx *c, double s) {
    if (!c || !std::isfinite(s)) return FCFD_ERR_ARG;
    c->vscale = s;
    return FCFD_OK;
}

extern "C" int fcfd_set_imass(fcfd_ctx *c, double imass) {
    if (!c) return FCFD_ERR_ARG;
    c->pr.imass = imass;
    c->diag.imass = imass;
    return FCFD_OK;
}

extern "C" int fcfd_bc(fcfd_ctx *c, int f) {
    double *d = field_or_null(c, f);
    if (!d) return FCFD_ERR_FIELD;
    return bc_scalar(c, d);
}

extern "C" int fcfd_velbc(fcfd_ctx *c) {
    if (!c) return FCFD_ERR_ARG;
    return velbc(c, c->f[FCFD_U], c->f[FCFD_V], c->f[FCFD_W], 0);
}
