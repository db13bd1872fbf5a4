// adjoint.hip — discrete adjoint of the WENO5 flux-split level-set advection
// with SSP-RK3 (the forward is k_ls_adv_stage in levelset.hip, following
// `Two phase flows/newOpenMP/3D+IBM/src/interface_solver/
// level_set_rk3_solver.f90:1-219`).  There is no counterpart in the
// reference: this is the differentiation entry point of the path.
//
// One forward step with the velocity frozen over the step is
//   phi1 = phi  + dt L(phi)
//   phi2 = phi1 + dt/4  (-3 L(phi) + L(phi1))
//   phi3 = phi2 + dt/12 (-L(phi) - L(phi1) + 8 L(phi2)),
//   L(phi)_i = -sum_d (F_{i+1/2} - F_{i-1/2}) / h_d,
//   F_{f} = WENO_left(s1 at f-2..f+2) + WENO_right(s2 at f-1..f+3),
//   s1 = (u + |u|)/2 phi, s2 = (u - |u|)/2 phi (u: the direction's velocity
//   component at the cell), every stencil position clamped to the interior.
// Its adjoint, given lambda3 = dJ/dphi3:
//   lambda2 = lambda3 + J(phi2)^T (8 dt/12 lambda3)
//   lambda1 = lambda2 + J(phi1)^T (dt/4 lambda2 - dt/12 lambda3)
//   lambda0 = lambda1 + J(phi)^T  (dt lambda1 - 3 dt/4 lambda2 - dt/12 lambda3),
// with J = dL/dphi, and the same stage seeds applied to dL/du give the
// velocity adjoint.  The stage states phi1, phi2 are recomputed from phi.
#include "fcfd_internal.h"
#include "weno.h"

using namespace fcfd;

namespace fcfd {

// w = c1 l1 + c2 l2 + c3 l3 over the interior (the seed of one stage)
__global__ void k_adj_seed(double *__restrict__ w, const double *__restrict__ l1, const double *__restrict__ l2,
                           const double *__restrict__ l3, double c1, double c2, double c3, Geom G) {
    int bx_, by_, bz_;
    xcd_block(bx_, by_, bz_);
    const int i = bx_ * blockDim.x + threadIdx.x, j = by_ * blockDim.y + threadIdx.y, k = bz_;
    if (i >= G.nx || j >= G.ny) return;
    const long id = G.id(i, j, k);
    double v = c1 * l1[id];
    if (l2) v += c2 * l2[id];
    if (l3) v += c3 * l3[id];
    w[id] = v;
}

// out = base + J(psi)^T w over the interior, and the velocity adjoint of
// the stage added into ub/vb/wb (when given).  Gather form: cell j visits
// every face f of each direction whose clamped WENO stencil can contain j
// (f = j-3 .. j+2, the faces -1/2 .. n-1/2 existing), forms
// Fbar_f = (w_{f+1} - w_f)/h (the two cells the face enters with -1/h and
// +1/h), runs both reconstructions in reverse and keeps the parts of the
// stencil slots that are j.  Deterministic: no atomics.
template <int DIM>
__global__ void __launch_bounds__(256) k_ls_adv_adj(const double *__restrict__ psi, const double *__restrict__ nu,
                                                    const double *__restrict__ nv, const double *__restrict__ nw,
                                                    const double *__restrict__ w, const double *__restrict__ base,
                                                    double *__restrict__ out, double *__restrict__ ub,
                                                    double *__restrict__ vb, double *__restrict__ wb, Geom G,
                                                    double vs) {
    int bx_, by_, bz_;
    xcd_block(bx_, by_, bz_);
    const int i = bx_ * blockDim.x + threadIdx.x, j = by_ * blockDim.y + threadIdx.y, k = bz_;
    if (i >= G.nx || j >= G.ny) return;
    const long id = G.id(i, j, k);
    const double *vel[3] = {nu, nv, nw};
    double *velb[3] = {ub, vb, wb};
    const int n[3] = {G.nx, G.ny, G.nz};
    const int pos[3] = {i, j, k};
    const long st[3] = {1, G.sx, G.sxy};
    const double h[3] = {G.dx, G.dy, G.dz};
    const double EPS = 1.0e-13;
    double pb = 0.0;
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
        const int p0 = pos[d], nd = n[d];
        // s1, s2 at the clamped positions p0-5 .. p0+5, the seed at p0-3 .. p0+3
        double s1[11], s2[11], wv[7];
#pragma unroll
        for (int o = -5; o <= 5; ++o) {
            const int q = clampi(p0 + o, 0, nd - 1);
            const long a = id + (long)(q - p0) * st[d];
            const double u = vs * vel[d][a], ph = psi[a];
            s1[o + 5] = 0.5 * (u + fabs(u)) * ph;
            s2[o + 5] = 0.5 * (u - fabs(u)) * ph;
        }
#pragma unroll
        for (int o = -3; o <= 3; ++o) {
            const int q = p0 + o;
            wv[o + 3] = (q >= 0 && q < nd) ? w[id + (long)o * st[d]] : 0.0;
        }
        double s1b = 0.0, s2b = 0.0;
#pragma unroll
        for (int fo = -3; fo <= 2; ++fo) {   // face f = p0 + fo between cells f and f+1
            const int f = p0 + fo;
            if (f < -1 || f > nd - 1) continue;
            const double fbar = (wv[fo + 4] - wv[fo + 3]) / h[d];
            double gb[5], fb[5];
            weno_left_adj(&s1[fo + 3], EPS, fbar, gb);    // s1 at f-2 .. f+2
            weno_right_adj(&s2[fo + 4], EPS, fbar, fb);   // s2 at f-1 .. f+3
#pragma unroll
            for (int m = 0; m < 5; ++m) {
                if (clampi(f - 2 + m, 0, nd - 1) == p0) s1b += gb[m];
                if (clampi(f - 1 + m, 0, nd - 1) == p0) s2b += fb[m];
            }
        }
        const double u = vs * vel[d][id], ph = psi[id];
        pb += s1b * 0.5 * (u + fabs(u)) + s2b * 0.5 * (u - fabs(u));
        if (velb[d]) {   // d/d(stored velocity) = vs d/du
            const double sg = u > 0.0 ? 1.0 : (u < 0.0 ? -1.0 : 0.0);
            velb[d][id] += vs * ((s1b * 0.5 * (1.0 + sg) + s2b * 0.5 * (1.0 - sg)) * ph);
        }
    }
    out[id] = base[id] + pb;
}

static int adj_scratch(fcfd_ctx *c) {
    if (c->adj[0]) return FCFD_OK;
    for (int q = 0; q < 5; ++q) {
        if (hipMalloc((void **)&c->adj[q], sizeof(double) * c->G.ntot) != hipSuccess) return FCFD_ERR_HIP;
        if (hipMemsetAsync(c->adj[q], 0, sizeof(double) * c->G.ntot, c->stream) != hipSuccess) return FCFD_ERR_HIP;
    }
    return FCFD_OK;
}

int ls_adv_adjoint_step(fcfd_ctx *c, const double *phi, double *lambda, double *ub, double *vb, double *wb) {
    const Geom &G = c->G;
    if (c->pr.ls_scheme != FCFD_LS_WENO5) {
        fprintf(stderr, "libfcfd: the advection adjoint covers the WENO5 scheme only\n");
        return FCFD_ERR_ARG;
    }
    FCFD_TRY(adj_scratch(c));
    double *phi1 = c->adj[0], *phi2 = c->adj[1], *lam1 = c->adj[2], *lam2 = c->adj[3], *w = c->adj[4];
    dim3 blk(64, 4), grd = grid_interior(G, blk);
    const double dt = c->pr.dt;
    double **F = c->f;
    ProfScope ps_(c, R_LS_ADV);
    // stage states phi1, phi2 of the forward step (interior; L reads no ghost)
    FCFD_TRY(ls_adv_stage(c, phi, phi1, 0));
    FCFD_TRY(ls_adv_stage(c, phi1, phi2, 1));
    auto adj = [&](const double *psi, const double *base, double *out) {
        if (G.dim == 3)
            hipLaunchKernelGGL(k_ls_adv_adj<3>, grd, blk, 0, c->stream, psi, F[FCFD_NU], F[FCFD_NV], F[FCFD_NW], w,
                               base, out, ub, vb, wb, G, c->vscale);
        else
            hipLaunchKernelGGL(k_ls_adv_adj<2>, grd, blk, 0, c->stream, psi, F[FCFD_NU], F[FCFD_NV], nullptr, w,
                               base, out, ub, vb, nullptr, G, c->vscale);
    };
    // lambda2 = lambda3 + J(phi2)^T (8 dt/12 lambda3)
    hipLaunchKernelGGL(k_adj_seed, grd, blk, 0, c->stream, w, lambda, nullptr, nullptr, 8.0 * dt / 12.0, 0.0, 0.0,
                       G);
    adj(phi2, lambda, lam2);
    // lambda1 = lambda2 + J(phi1)^T (dt/4 lambda2 - dt/12 lambda3)
    hipLaunchKernelGGL(k_adj_seed, grd, blk, 0, c->stream, w, lam2, lambda, nullptr, dt / 4.0, -dt / 12.0, 0.0, G);
    adj(phi1, lam2, lam1);
    // lambda0 = lambda1 + J(phi)^T (dt lambda1 - 3 dt/4 lambda2 - dt/12 lambda3)
    hipLaunchKernelGGL(k_adj_seed, grd, blk, 0, c->stream, w, lam1, lam2, lambda, dt, -3.0 * dt / 4.0, -dt / 12.0,
                       G);
    adj(phi, lam1, lambda);
    FCFD_CHECK(hipGetLastError());
    return FCFD_OK;
}

}  // namespace fcfd

extern "C" int fcfd_level_set_rk3_solver_f32(fcfd_ctx *ctx, float *phi, const float *nu, const float *nv,
                                             const float *nw) {
    if (!ctx || !phi || !nu || !nv || (ctx->G.dim == 3 && !nw)) return FCFD_ERR_ARG;
    if (ctx->pr.ls_scheme != FCFD_LS_WENO5) return FCFD_ERR_ARG;
    if (hipSetDevice(ctx->device) != hipSuccess) return FCFD_ERR_HIP;
    return fcfd::ls_advect_f32(ctx, phi, nu, nv, nw);
}

extern "C" int fcfd_ls_adv_adjoint_step(fcfd_ctx *ctx, const double *phi, double *lambda, double *ubar,
                                        double *vbar, double *wbar) {
    if (!ctx || !phi || !lambda) return FCFD_ERR_ARG;
    if (hipSetDevice(ctx->device) != hipSuccess) return FCFD_ERR_HIP;
    return fcfd::ls_adv_adjoint_step(ctx, phi, lambda, ubar, vbar, ctx->G.dim == 3 ? wbar : nullptr);
}
