// adjoint.hip — discrete adjoint of the WENO5 flux-split level-set advection
// with SSP-RK3 (the forward is k_ls_adv_stage_kf in levelset.hip, following
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
template <typename T>
__global__ void k_adj_seed(T *__restrict__ w, const T *__restrict__ l1, const T *__restrict__ l2,
                           const T *__restrict__ l3, T c1, T c2, T c3, Geom G) {
    int bx_, by_, bz_;
    xcd_block(bx_, by_, bz_);
    const int i = bx_ * blockDim.x + threadIdx.x, j = by_ * blockDim.y + threadIdx.y, k = bz_;
    if (i >= G.nx || j >= G.ny) return;
    const long id = G.id(i, j, k);
    T v = c1 * l1[id];
    if (l2) v += c2 * l2[id];
    if (l3) v += c3 * l3[id];
    w[id] = v;
}

// out = base + J(psi)^T w over the interior, and the velocity adjoint of the
// stage added into ub/vb/wb (when given), one direction per pass: the x
// pass writes out = base + (x part), the y and z passes add theirs.  Each
// face's reverse reconstructions are computed once (the value of face f,
// between cells f and f+1, is WENO_left(s1 at f-2..f+2) +
// WENO_right(s2 at f-1..f+3); its seed is Fbar_f = (w_{f+1} - w_f)/h, w = 0
// outside the line) and cell c collects the slots of faces c-3 .. c+2 whose
// clamped stencil position is c, so boundary cells receive every slot that
// the zero-gradient clamp folds onto them.  Deterministic: no atomics.
// T = float is the adjoint of the FP32 forward (its weights, weno.h).
constexpr double ADJ_EPS = 1.0e-13;

// the velocity and level-set adjoints of one cell from its s1, s2 adjoints
template <typename T>
__device__ __forceinline__ void adj_cell(T s1b, T s2b, T u, T ph, T vs, T &pb, T &velb) {
    const T half = (T)0.5, one = (T)1;
    pb = s1b * half * (u + fabs(u)) + s2b * half * (u - fabs(u));
    const T sg = u > (T)0 ? one : (u < (T)0 ? -one : (T)0);
    velb = vs * ((s1b * half * (one + sg) + s2b * half * (one - sg)) * ph);
}

// reverse reconstructions of one face from its line samples: s1, s2 at the
// clamped positions f-2 .. f+3
template <typename T>
__device__ __forceinline__ void face_adj(const T *s1, const T *s2, T fbar, T *gb, T *fb) {
    weno_left_adj<T>(s1, (T)ADJ_EPS, fbar, gb);       // s1 at f-2 .. f+2
    weno_right_adj<T>(s2 + 1, (T)ADJ_EPS, fbar, fb);  // s2 at f-1 .. f+3
}

constexpr int ADJ_TX = 64, ADJ_TY = 4;
constexpr int ADJ_NFX = ADJ_TX + 5;   // x faces i0-3 .. i0+65 of a 64-cell row
constexpr int ADJ_KC = 16;            // z cells per thread

// x pass: block 64 x 4 cells of one plane; the 69 faces of each row are
// computed by the row's wave (two passes) into LDS, then each cell gathers
template <int DIM, typename T>
__global__ void __launch_bounds__(256) k_adj_x(const T *__restrict__ psi, const T *__restrict__ nu,
                                               const T *__restrict__ w, const T *__restrict__ base,
                                               T *__restrict__ out, T *__restrict__ ub, Geom G, T vs) {
    __shared__ T sg[ADJ_TY][ADJ_NFX][5], sf[ADJ_TY][ADJ_NFX][5];
    int bx_, by_, bz_;
    xcd_block(bx_, by_, bz_);
    const int lane = threadIdx.x, r = threadIdx.y;
    const int i0 = bx_ * ADJ_TX, j = by_ * ADJ_TY + r, k = bz_;
    const int n = G.nx;
    const T half = (T)0.5, zero = (T)0;
    // the 4 x 69 faces over the 256 threads (one wave takes a second turn)
    for (int e = r * ADJ_TX + lane; e < ADJ_TY * ADJ_NFX; e += ADJ_TX * ADJ_TY) {
        const int rr = e / ADJ_NFX, t = e - rr * ADJ_NFX, jj = by_ * ADJ_TY + rr;
        const int f = i0 - 3 + t;
        T gb[5] = {zero, zero, zero, zero, zero}, fb[5] = {zero, zero, zero, zero, zero};
        if (jj < G.ny && f >= -1 && f <= n - 1) {
            const long row = G.id(0, jj, k);
            const T wl = f >= 0 ? w[row + f] : zero, wr = f + 1 <= n - 1 ? w[row + f + 1] : zero;
            T s1[6], s2[6];
#pragma unroll
            for (int m = 0; m < 6; ++m) {
                const long a = row + clampi(f - 2 + m, 0, n - 1);
                const T u = vs * nu[a], ph = psi[a];
                s1[m] = half * (u + fabs(u)) * ph;
                s2[m] = half * (u - fabs(u)) * ph;
            }
            face_adj<T>(s1, s2, (wr - wl) / (T)G.dx, gb, fb);
        }
#pragma unroll
        for (int m = 0; m < 5; ++m) {
            sg[rr][t][m] = gb[m];
            sf[rr][t][m] = fb[m];
        }
    }
    __syncthreads();
    const bool rowok = j < G.ny;
    const long row = rowok ? G.id(0, j, k) : 0;
    const int i = i0 + lane;
    if (!rowok || i >= n) return;
    T s1b = zero, s2b = zero;
#pragma unroll
    for (int fo = -3; fo <= 2; ++fo) {
        const int f = i + fo, t = lane + fo + 3;
        if (f < -1 || f > n - 1) continue;
#pragma unroll
        for (int m = 0; m < 5; ++m) {
            if (clampi(f - 2 + m, 0, n - 1) == i) s1b += sg[r][t][m];
            if (clampi(f - 1 + m, 0, n - 1) == i) s2b += sf[r][t][m];
        }
    }
    const long id = row + i;
    T pb, vb;
    adj_cell<T>(s1b, s2b, vs * nu[id], psi[id], vs, pb, vb);
    out[id] = base[id] + pb;
    if (ub) ub[id] += vb;
}

// y and z passes: each thread walks the faces of one line along axis AX
// (1: y, 2: z) that reach its ADJ_KC cells, keeping the samples of positions
// f-2 .. f+3 and the s1 / s2 adjoints of cells f-2 .. f+3 in registers;
// after face f no later face reaches cell f-2, which is then complete.
// Threads run along x (coalesced); blockIdx.y picks the other index.
template <int AX, typename T>
__global__ void __launch_bounds__(256) k_adj_walk(const T *__restrict__ psi, const T *__restrict__ vel,
                                                  const T *__restrict__ w, T *__restrict__ out,
                                                  T *__restrict__ velbar, Geom G, T vs) {
    int bx_, by_, bz_;
    xcd_block(bx_, by_, bz_);
    const int i = bx_ * blockDim.x + threadIdx.x, o = by_ * blockDim.y + threadIdx.y;
    const int n = AX == 1 ? G.ny : G.nz;
    const int k0 = bz_ * ADJ_KC, k1 = min(k0 + ADJ_KC, n);
    if (i >= G.nx || o >= (AX == 1 ? G.nz : G.ny)) return;
    const long c0 = AX == 1 ? G.id(i, 0, o) : G.id(i, o, 0), st = AX == 1 ? G.sx : G.sxy;
    const T h = (T)(AX == 1 ? G.dy : G.dz), half = (T)0.5, zero = (T)0;
    const int fa = max(-1, k0 - 3), fz = min(n - 1, k1 + 1);
    T s1[6], s2[6], acc1[6], acc2[6];
#pragma unroll
    for (int m = 0; m < 6; ++m) {
        const long a = c0 + clampi(fa - 2 + m, 0, n - 1) * st;
        const T u = vs * vel[a], ph = psi[a];
        s1[m] = half * (u + fabs(u)) * ph;
        s2[m] = half * (u - fabs(u)) * ph;
        acc1[m] = zero;
        acc2[m] = zero;
    }
    auto finish = [&](int c, T s1b, T s2b) {
        const long id = c0 + c * st;
        T pb, vb;
        adj_cell<T>(s1b, s2b, vs * vel[id], psi[id], vs, pb, vb);
        out[id] = out[id] + pb;
        if (velbar) velbar[id] += vb;
    };
    T wl = fa >= 0 ? w[c0 + fa * st] : zero;
    for (int f = fa; f <= fz; ++f) {
        const T wr = f + 1 <= n - 1 ? w[c0 + (f + 1) * st] : zero;
        T gb[5], fb[5];
        face_adj<T>(s1, s2, (wr - wl) / h, gb, fb);
        const int b = f - 2;   // cell of window slot 0
        if (f - 2 >= 0 && f + 3 <= n - 1) {
#pragma unroll
            for (int m = 0; m < 5; ++m) {
                acc1[m] += gb[m];
                acc2[m + 1] += fb[m];
            }
        } else {   // near a wall: slots the clamp folds onto the boundary cell
#pragma unroll
            for (int m = 0; m < 5; ++m) {
                const int c1 = clampi(f - 2 + m, 0, n - 1) - b, c2 = clampi(f - 1 + m, 0, n - 1) - b;
#pragma unroll
                for (int q = 0; q < 6; ++q) {
                    acc1[q] += c1 == q ? gb[m] : zero;
                    acc2[q] += c2 == q ? fb[m] : zero;
                }
            }
        }
        if (b >= k0 && b < k1) finish(b, acc1[0], acc2[0]);
#pragma unroll
        for (int m = 0; m < 5; ++m) {
            acc1[m] = acc1[m + 1];
            acc2[m] = acc2[m + 1];
            s1[m] = s1[m + 1];
            s2[m] = s2[m + 1];
        }
        acc1[5] = zero;
        acc2[5] = zero;
        {
            const long a = c0 + clampi(f + 4, 0, n - 1) * st;   // position (f+1) + 3
            const T u = vs * vel[a], ph = psi[a];
            s1[5] = half * (u + fabs(u)) * ph;
            s2[5] = half * (u - fabs(u)) * ph;
        }
        wl = wr;
    }
    // at the far wall the last cells complete after the last face
#pragma unroll
    for (int m = 0; m < 5; ++m) {
        const int c = fz - 1 + m;
        if (c >= k0 && c < k1 && c <= n - 1) finish(c, acc1[m], acc2[m]);
    }
}

static int adj_scratch(fcfd_ctx *c) {
    if (c->adj[0]) return FCFD_OK;
    for (int q = 0; q < 5; ++q) {
        if (hipMalloc((void **)&c->adj[q], sizeof(double) * c->G.ntot) != hipSuccess) return FCFD_ERR_HIP;
        if (hipMemsetAsync(c->adj[q], 0, sizeof(double) * c->G.ntot, c->stream) != hipSuccess) return FCFD_ERR_HIP;
    }
    return FCFD_OK;
}

// the three reverse stages of one step (see the file head), for T = double
// (the context's NU, NV, NW) or float (caller velocity); scratch: stage
// states phi1, phi2, adjoints lam1, lam2, the seed w
template <typename T>
static int adjoint_stages(fcfd_ctx *c, const T *phi, const T *phi1, const T *phi2, T *lambda, T *lam1, T *lam2,
                          T *w, const T *nu, const T *nv, const T *nw, T *ub, T *vb, T *wb) {
    const Geom &G = c->G;
    dim3 blk(64, 4), grd = grid_interior(G, blk);
    const T dt = (T)c->pr.dt, vs = (T)c->vscale;
    // walkers: 64 x 4 threads over (x, z) or (x, y) in 3D, 256 along x in 2D
    const dim3 bw = G.dim == 3 ? dim3(64, 4) : dim3(256, 1);
    const dim3 gy((G.nx + bw.x - 1) / bw.x, (G.nz + bw.y - 1) / bw.y, (G.ny + ADJ_KC - 1) / ADJ_KC);
    const dim3 gz((G.nx + bw.x - 1) / bw.x, (G.ny + bw.y - 1) / bw.y, (G.nz + ADJ_KC - 1) / ADJ_KC);
    auto adj = [&](const T *psi, const T *base, T *out) {
        if (G.dim == 3)
            hipLaunchKernelGGL((k_adj_x<3, T>), grd, blk, 0, c->stream, psi, nu, w, base, out, ub, G, vs);
        else
            hipLaunchKernelGGL((k_adj_x<2, T>), grd, blk, 0, c->stream, psi, nu, w, base, out, ub, G, vs);
        hipLaunchKernelGGL((k_adj_walk<1, T>), gy, bw, 0, c->stream, psi, nv, w, out, vb, G, vs);
        if (G.dim == 3) hipLaunchKernelGGL((k_adj_walk<2, T>), gz, bw, 0, c->stream, psi, nw, w, out, wb, G, vs);
    };
    const T one = (T)1, three = (T)3, four = (T)4, eight = (T)8, twelve = (T)12;
    // lambda2 = lambda3 + J(phi2)^T (8 dt/12 lambda3)
    hipLaunchKernelGGL((k_adj_seed<T>), grd, blk, 0, c->stream, w, (const T *)lambda, (const T *)nullptr,
                       (const T *)nullptr, eight * dt / twelve, (T)0, (T)0, G);
    adj(phi2, lambda, lam2);
    // lambda1 = lambda2 + J(phi1)^T (dt/4 lambda2 - dt/12 lambda3)
    hipLaunchKernelGGL((k_adj_seed<T>), grd, blk, 0, c->stream, w, (const T *)lam2, (const T *)lambda,
                       (const T *)nullptr, dt / four, -dt / twelve, (T)0, G);
    adj(phi1, lam2, lam1);
    // lambda0 = lambda1 + J(phi)^T (dt lambda1 - 3 dt/4 lambda2 - dt/12 lambda3)
    hipLaunchKernelGGL((k_adj_seed<T>), grd, blk, 0, c->stream, w, (const T *)lam1, (const T *)lam2,
                       (const T *)lambda, one * dt, -three * dt / four, -dt / twelve, G);
    adj(phi, lam1, lambda);
    FCFD_CHECK(hipGetLastError());
    return FCFD_OK;
}

int ls_adv_adjoint_step(fcfd_ctx *c, const double *phi, double *lambda, double *ub, double *vb, double *wb) {
    if (c->pr.ls_scheme != FCFD_LS_WENO5) {
        fprintf(stderr, "libfcfd: the advection adjoint covers the WENO5 scheme only\n");
        return FCFD_ERR_ARG;
    }
    FCFD_TRY(adj_scratch(c));
    double *phi1 = c->adj[0], *phi2 = c->adj[1], *lam1 = c->adj[2], *lam2 = c->adj[3], *w = c->adj[4];
    double **F = c->f;
    ProfScope ps_(c, R_LS_ADV);
    // stage states phi1, phi2 of the forward step (interior; L reads no ghost)
    FCFD_TRY(ls_adv_stage(c, phi, phi1, 0));
    FCFD_TRY(ls_adv_stage(c, phi1, phi2, 1));
    return adjoint_stages<double>(c, phi, phi1, phi2, lambda, lam1, lam2, w, F[FCFD_NU], F[FCFD_NV], F[FCFD_NW], ub,
                                  vb, wb);
}

// the FP32 adjoint of one fcfd_level_set_rk3_solver_f32 step on caller arrays
int ls_adv_adjoint_step_f32(fcfd_ctx *c, const float *phi, float *lambda, float *ub, float *vb, float *wb,
                            const float *nu, const float *nv, const float *nw) {
    const Geom &G = c->G;
    if (!c->adj32[0])
        for (int q = 0; q < 7; ++q) {
            FCFD_CHECK(hipMalloc((void **)&c->adj32[q], sizeof(float) * G.ntot));
            FCFD_CHECK(hipMemsetAsync(c->adj32[q], 0, sizeof(float) * G.ntot, c->stream));
        }
    float *phi1 = c->adj32[0], *phi2 = c->adj32[1], *lam1 = c->adj32[2], *lam2 = c->adj32[3], *w = c->adj32[4];
    float *l1 = c->adj32[5], *l2 = c->adj32[6];
    ProfScope ps_(c, R_LS_ADV);
    FCFD_TRY(ls_adv_stage_f32(c, phi, phi1, nu, nv, nw, l1, l2, 0));
    FCFD_TRY(ls_adv_stage_f32(c, phi1, phi2, nu, nv, nw, l1, l2, 1));
    return adjoint_stages<float>(c, phi, phi1, phi2, lambda, lam1, lam2, w, nu, nv, nw, ub, vb, wb);
}

}  // namespace fcfd

extern "C" int fcfd_level_set_rk3_solver_f32(fcfd_ctx *ctx, float *phi, const float *nu, const float *nv,
                                             const float *nw) {
    if (!ctx || !phi || !nu || !nv || (ctx->G.dim == 3 && !nw)) return FCFD_ERR_ARG;
    if (ctx->pr.ls_scheme != FCFD_LS_WENO5) return FCFD_ERR_ARG;
    if (hipSetDevice(ctx->device) != hipSuccess) return FCFD_ERR_HIP;
    return fcfd::ls_advect_f32(ctx, phi, nu, nv, nw);
}

extern "C" int fcfd_ls_adv_adjoint_step_f32(fcfd_ctx *ctx, const float *phi, float *lambda, float *ubar,
                                            float *vbar, float *wbar, const float *nu, const float *nv,
                                            const float *nw) {
    if (!ctx || !phi || !lambda || !nu || !nv || (ctx->G.dim == 3 && !nw)) return FCFD_ERR_ARG;
    if (ctx->pr.ls_scheme != FCFD_LS_WENO5) return FCFD_ERR_ARG;
    if (hipSetDevice(ctx->device) != hipSuccess) return FCFD_ERR_HIP;
    return fcfd::ls_adv_adjoint_step_f32(ctx, phi, lambda, ubar, vbar, ctx->G.dim == 3 ? wbar : nullptr, nu, nv,
                                         ctx->G.dim == 3 ? nw : nullptr);
}

extern "C" int fcfd_ls_adv_adjoint_step(fcfd_ctx *ctx, const double *phi, double *lambda, double *ubar,
                                        double *vbar, double *wbar) {
    if (!ctx || !phi || !lambda) return FCFD_ERR_ARG;
    if (hipSetDevice(ctx->device) != hipSuccess) return FCFD_ERR_HIP;
    return fcfd::ls_adv_adjoint_step(ctx, phi, lambda, ubar, vbar, ctx->G.dim == 3 ? wbar : nullptr);
}
