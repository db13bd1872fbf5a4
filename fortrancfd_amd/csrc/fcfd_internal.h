// fcfd_internal.h — device context and helpers shared by the HIP sources of
// libfcfd.  Everything here is MI355X (gfx950, wave64) code.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <vector>

#include "../../include/fcfd.h"

#define FCFD_CHECK(call)                                                                \
    do {                                                                                \
        hipError_t e_ = (call);                                                         \
        if (e_ != hipSuccess) {                                                         \
            fprintf(stderr, "libfcfd: HIP error %s at %s:%d\n", hipGetErrorString(e_), \
                    __FILE__, __LINE__);                                                \
            return FCFD_ERR_HIP;                                                        \
        }                                                                               \
    } while (0)

#define FCFD_TRY(expr)              \
    do {                            \
        int rc_ = (expr);           \
        if (rc_ != FCFD_OK) return rc_; \
    } while (0)

namespace fcfd {

constexpr int WAVE = 64;

// Geometry of one ghosted field: interior (nx, ny, nz), ghost width g (gz = 0
// in 2D), strides of the column-major reference layout.
struct Geom {
    int dim, nx, ny, nz, g, gz;
    long sx, sxy, ntot;
    double dx, dy, dz;
    // Ghost depth that holds a neighbour's data on each face (x low / high,
    // y, z): 0 on a physical face, where kernels read clamped indices as the
    // reference's zero-gradient copies; up to g on a face between blocks of
    // a multi-block layout, whose ghosts a halo exchange fills.
    int ox0 = 0, ox1 = 0, oy0 = 0, oy1 = 0, oz0 = 0, oz1 = 0;
    __host__ __device__ long id(int i, int j, int k) const {
        return (long)(i + g) + (long)(j + g) * sx + (long)(k + gz) * sxy;
    }
};

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// d[m] <- d[clamp(m, lo, hi)] for m = 0 .. W-1 (0 <= lo <= hi < W) by
// selects, without indexing registers at run time: the clamped differences
// of a line near a wall taken from the unclamped ones of its load window
template <int W>
__device__ __forceinline__ void clamp_window(double (&d)[W], int lo, int hi) {
    double vlo = d[0], vhi = d[0];
#pragma unroll
    for (int r = 1; r < W; ++r) {
        vlo = lo == r ? d[r] : vlo;
        vhi = hi == r ? d[r] : vhi;
    }
#pragma unroll
    for (int m = 0; m < W; ++m) d[m] = m < lo ? vlo : (m > hi ? vhi : d[m]);
}

// One level of the pressure multigrid hierarchy (see mg.hip).
struct MgLevel {
    int nx, ny, nz;          // cells
    int cx, cy, cz;          // coarsening factor (1|2) to the next level
    double hx, hy, hz;       // spacing
    long sx, sxy, ntot;      // layout with one ghost layer (finest uses Geom)
    // FP64 arrays (level >= 1): solution/error, rhs, residual, face coeffs
    double *e64, *f64, *r64, *bx64, *by64, *bz64;
    // FP32 mirrors for the mixed-precision inner cycle; e32b is the second
    // error buffer of the fused multi-sweep smoother (read one, write the other)
    float *e32, *f32, *r32, *bx32, *by32, *bz32, *e32b;
    // FP16 face coefficients of level 1 in 3D (params.ppe_rho_half), read by
    // its FP32 smoothing and restriction; NULL elsewhere
    _Float16 *bx16, *by16, *bz16;
};

// Kernel regions timed by the built-in HIP-event profiler (fcfd_prof_*).
enum Region {
    R_LS_ADV, R_REDIS_A, R_REDIS_B, R_MASS, R_MPLS_UPD, R_CCD, R_CURV, R_DIFF, R_PRED, R_PPE_SRC,
    R_PROJ, R_NODE_VEL, R_BC, R_SMOOTH0, R_SMOOTHC, R_RESTRICT0, R_RESTRICTC, R_PROLONG0, R_PROLONGC,
    R_RESID0, R_COARSE, R_FCG, R_MISC, NREG
};

struct ProfRec {
    int reg;
    int ev;   // index of the start event in the pool (end = ev + 1)
};

struct Prof {
    int on = 0;
    std::vector<hipEvent_t> pool;
    std::vector<ProfRec> recs;
    size_t used = 0;
};

struct Reduce {
    double *partial;         // per-block partial results (nslots * nblocks)
    double *partial2;        // second-stage partials (4 slots x 256), inside `partial`
    double *result;          // reduced values (device)
    double *host;            // pinned host mirror
    int nblocks;
};

}  // namespace fcfd

struct fcfd_ctx {
    fcfd_grid gd;
    fcfd_params pr;
    fcfd::Geom G;
    int tile[FCFD_NTILE][3];   // thread-block tile (x, y) and slab rows per enum fcfd_tile_kernel
    hipStream_t stream;
    int device;
    int ncu;                   // compute units of the device
    // ---- state (device, full ghosted arrays) ----
    double *f[FCFD_NFIELDS];   // indexed by enum fcfd_field; NULL when not held
    // ---- scratch ----
    double *phiB, *phiS;       // ping-pong phi buffers / stage start copy
    double *l1, *l2;           // stored RK stage sources
    double *s2, *s3, *s0;      // reinit work: lambda numerator/denominator, frozen sign
    double *cx, *cy, *cz;      // CCD second-derivative scratch / cross-derivative temp
    double *dsu, *dsv, *dsw;   // viscous + body force source (sec_part1..3), once per step
    double *un, *vn, *wn;      // predictor output buffer (ping-pong with u, v, w)
    double *ccd_coef[18];      // factorised line matrices [kind * 3 + direction]: CCD, UCCD up / down on
                               // ghosted lines, the same on interior lines (MPLS_RisingBubble set)
    double *ccd_seg[18];       // the same per segment for the partitioned y / z solve, + reduced block LU
    double *ccd_z;             // partitioned solve: segment end values / interface values per line
    double *adj[5];            // advection adjoint: stage states phi1, phi2, adjoints lambda1, lambda2, seed
    float *adv32[4];           // FP32 advection: two stage buffers and the stage sources l1, l2
    float *adj32[7];           // FP32 adjoint: phi1, phi2, lambda1, lambda2, seed, stage sources l1, l2
    double vscale = 1.0;       // factor on NU/NV/NW in the WENO5 advection and its adjoint
    // ---- batched 2D stack (fcfd_create_batch): case q is the slice k = q ----
    int nb = 1;                        // cases in the stack; 1 for an ordinary context
    fcfd_case *cases_dev = nullptr;    // per-case parameters on the device; NULL when nb == 1
    std::vector<fcfd_case> cases;      // host copy
    // ---- reductions ----
    fcfd::Reduce red;
    unsigned long long *flag;  // device flags (reinit error max bits, done, count)
    unsigned long long *flag_host;
    hipEvent_t redis_ev[2] = {nullptr, nullptr};   // reinit stop polls in flight
    // ---- multigrid ----
    std::vector<fcfd::MgLevel> mg;
    float *rho32;              // FP32 copy of rho for the mixed inner cycle
    _Float16 *rho16;           // FP16 copy (params.ppe_rho_half)
    double *mg_scalars;        // device scalars of the solve (norms)
    int mg_ready;
    // level-0 Krylov vectors of the flexible CG (FP64) and FP32 previous z
    double *kr, *kp, *kq, *kz, *kzo;
    float *kzo32;
    // captured V-cycles: [kind][inner cycles 1..4]; kind 0 fp64 on p, 1 fp64 on
    // z with the CG dots, 2 fp32 on e32, 3 fp32 on e32 with the CG dots
    hipGraphExec_t vgraph[2][4][5];                      // [host status slot][kind][inner]
    int *mg_skip;              // device flag: every case's flexible-CG solve has ended
    hipEvent_t cg_ev[2];       // end of the flexible-CG iteration of each host status slot
    std::vector<std::pair<int, int>> vgraph_prof[4][5];  // (region, event) pairs inside each graph
    int capture_kind = -1, capture_inner = 0;            // graph being captured (or -1)
    double prof_gms[fcfd::NREG] = {};                    // timings folded in from graph replays
    int64_t prof_gcnt[fcfd::NREG] = {};
    // ---- bookkeeping ----
    double time;
    int iter;
    int diag_pending;          // step diagnostics of every case waiting in diag_host
    double *diag_host;         // pinned: result words 0..DIAG_W-1 of case q at [q*DIAG_W]
    unsigned long long *pack;  // device staging for per-case values gathered to the host (16 words per case)
    fcfd_diag diag;
    fcfd_m3_state m3 = {{0.0, 0.0}, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 1.0e-5, 0, 0};   // MPLS_RisingBubble step
    fcfd::Prof prof;
};

namespace fcfd {
// Records a start/end HIP event pair on the context stream around the launches
// of its scope when profiling is on.
struct ProfScope {
    fcfd_ctx *c;
    int reg, ev;
    ProfScope(fcfd_ctx *c_, int reg_);
    ~ProfScope();
};
}  // namespace fcfd

namespace fcfd {

// Sum over the threads of a 256-thread block in a fixed order: a butterfly
// inside each wave (lane 0's value is used), then the four wave sums in
// order.  Every thread gets the result; every thread of the block must call
// it; sh4 is LDS scratch of 4 doubles that may be reused after the call.
__device__ __forceinline__ double block_sum256(double v, double *sh4) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, WAVE);
    const int t = threadIdx.x + threadIdx.y * blockDim.x;
    if ((t & (WAVE - 1)) == 0) sh4[t / WAVE] = v;
    __syncthreads();
    const double s = (sh4[0] + sh4[1]) + (sh4[2] + sh4[3]);
    __syncthreads();
    return s;
}

// Workgroups of a launch are dealt to the 8 XCDs round-robin by linear id
// (x fastest, then y, then z).  Renumbering them so that XCD x owns the
// contiguous range of logical blocks [x*nb/8, (x+1)*nb/8) keeps the k-1 and
// k+1 planes a stencil re-reads in the L2 of the XCD that read them first.
// Returns the logical (bx, by, bz) of the calling workgroup.
constexpr unsigned NXCD = 8;
__device__ __forceinline__ void xcd_block(int &bx, int &by, int &bz) {
    const unsigned nbx = gridDim.x, nby = gridDim.y, nb = nbx * nby * gridDim.z;
    const unsigned w = blockIdx.x + nbx * (blockIdx.y + nby * blockIdx.z);
    const unsigned q = nb / NXCD, rem = nb % NXCD, x = w % NXCD;
    const unsigned l = x * q + (x < rem ? x : rem) + w / NXCD;
    bx = (int)(l % nbx);
    const unsigned r = l / nbx;
    by = (int)(r % nby);
    bz = (int)(r / nby);
}

// xcd_block with another order of the logical blocks: with slab = S > 0 the
// tile rows are grouped in slabs of S rows, and an XCD's contiguous range
// walks the planes of one slab (x fastest, then the slab's rows, then the
// planes) before the next slab.  A plane's footprint per XCD is then S tile
// rows instead of the whole plane, so the k-2 .. k+2 planes a z-stencil
// re-reads are still in the XCD's L2.  slab = 0 is xcd_block's order.
__device__ __forceinline__ void xcd_block_slab(int &bx, int &by, int &bz, int slab) {
    const unsigned nbx = gridDim.x, nby = gridDim.y, nbz = gridDim.z, nb = nbx * nby * nbz;
    const unsigned w = blockIdx.x + nbx * (blockIdx.y + nby * blockIdx.z);
    const unsigned q = nb / NXCD, rem = nb % NXCD, x = w % NXCD;
    const unsigned l = x * q + (x < rem ? x : rem) + w / NXCD;
    if (slab <= 0) {
        bx = (int)(l % nbx);
        const unsigned r = l / nbx;
        by = (int)(r % nby);
        bz = (int)(r / nby);
        return;
    }
    const unsigned S = (unsigned)slab, per = nbx * S * nbz;
    const unsigned sl = l / per, rows = min(S, nby - sl * S);
    const unsigned within = l - sl * per;
    bx = (int)(within % nbx);
    const unsigned r = within / nbx;
    by = (int)(sl * S + r % rows);
    bz = (int)(r / rows);
}

// Max over the threads of a 256-thread block (values >= 0 or any order-free
// max); every thread gets the result; every thread of the block must call it.
__device__ __forceinline__ double block_max256(double v, double *sh4) {
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, WAVE));
    const int t = threadIdx.x + threadIdx.y * blockDim.x;
    if ((t & (WAVE - 1)) == 0) sh4[t / WAVE] = v;
    __syncthreads();
    const double m = fmax(fmax(sh4[0], sh4[1]), fmax(sh4[2], sh4[3]));
    __syncthreads();
    return m;
}

// device flag words: [0] unused, [1] reinit done, [2] reinit pseudo-step
// count, [3] max |grad phi| bits; [FLAG_SHARD0, +NSHARD) per-block max slots.
// A batched stack holds one such block of FLAG_STRIDE words per case.
constexpr int FLAG_SHARD0 = 16, NSHARD = 256, FLAG_STRIDE = FLAG_SHARD0 + NSHARD;
// reduced values per case in Reduce::result: [0..2] mass sums, [RES_LAMBDA]
// the last MPLS multiplier, [RES_VDIV], [RES_VDIV + 1] max / mean |div u|
// (case 0 only); the first DIAG_W words of every case are gathered to the
// host after each step
constexpr int RES_STRIDE = 16, RES_LAMBDA = 8, RES_VDIV = 9, DIAG_W = 11;

// case index of slice k: a batched 2D stack keeps case q in slice k = q;
// 3D contexts hold one case
template <int DIM>
__device__ __forceinline__ int case_of(int k) { return DIM == 2 ? k : 0; }

// launch helpers --------------------------------------------------------
inline dim3 tile_blk(const fcfd_ctx *c, int kernel) { return dim3(c->tile[kernel][0], c->tile[kernel][1]); }
// interior cell (i, j, k) of a one-dimensional launch over the nx * ny * nz
// interior cells in memory order (x fastest; no idle lanes at the row ends
// as a 2D tile has when nx is not a multiple of its width), the blocks in
// xcd_block's XCD-contiguous order; false past the last cell.  The launch
// (grid_flat) needs nx * ny * nz < 2^32.
__device__ __forceinline__ bool cell_of_flat(const Geom &G, int &i, int &j, int &k) {
    int bx, by, bz;
    xcd_block(bx, by, bz);
    const unsigned t = (unsigned)bx * blockDim.x + threadIdx.x;
    const unsigned nxy = (unsigned)G.nx * (unsigned)G.ny;
    if (t >= nxy * (unsigned)G.nz) return false;
    const unsigned kk = t / nxy, r = t - kk * nxy, jj = r / (unsigned)G.nx;
    i = (int)(r - jj * (unsigned)G.nx);
    j = (int)jj;
    k = (int)kk;
    return true;
}

// cell (i, j, k) of flattened interior index t (i fastest), by 32-bit
// division where t fits (a 64-bit one is several times longer)
__device__ __forceinline__ void cell_of_index(long t, const Geom &G, int &i, int &j, int &k) {
    if (t < 4294967296L) {
        const unsigned tt = (unsigned)t, nx = (unsigned)G.nx, ny = (unsigned)G.ny;
        const unsigned r = tt / nx, kk = r / ny;
        i = (int)(tt - r * nx);
        j = (int)(r - kk * ny);
        k = (int)kk;
    } else {
        i = (int)(t % G.nx);
        const long r = t / G.nx;
        j = (int)(r % G.ny);
        k = (int)(r / G.ny);
    }
}

inline bool flat_fits(const Geom &G) { return (double)G.nx * G.ny * G.nz < 4294967296.0 - 1024.0; }

inline dim3 grid_flat(const Geom &G, int blk) {
    return dim3((unsigned)(((long)G.nx * G.ny * G.nz + blk - 1) / blk));
}

inline dim3 grid_interior(const Geom &G, dim3 blk) {
    return dim3((G.nx + blk.x - 1) / blk.x, (G.ny + blk.y - 1) / blk.y, G.nz);
}

// kernels implemented in the .hip files (host wrappers)
int upload_cases(fcfd_ctx *c);
// rows of `width` 8-byte words, `pitch` words apart on the device, into
// consecutive host words (one contiguous copy; a gather kernel first when
// there is more than one row); stream ordered
int gather_to_host(fcfd_ctx *c, void *host, const void *dev, long pitch, int width, int rows);
int bc_scalar(fcfd_ctx *c, double *d);
int bc_scalar_f32(fcfd_ctx *c, float *d);
int velbc(fcfd_ctx *c, double *u, double *v, double *w, int nodal);
int ls_props(fcfd_ctx *c, const double *phi, double *H, double *D, double *S, double *rho, double *mu);
int reduce_sums(fcfd_ctx *c, int kind, const double *a, const double *b, const double *cc, const double *d);
int ccd_dir(fcfd_ctx *c, int dir, const double *f, double *fx, double *fxx);
int ccd_lines_set(fcfd_ctx *c, int set, int dir, const double *f, double *fx, double *fxx);
int m3_fields(fcfd_ctx *c);
int uccd_lines_set(fcfd_ctx *c, int set, int dir, const double *f, const double *u, double *fx, double *fxx);
int uccd_dir(fcfd_ctx *c, int dir, const double *f, const double *u, double *fx, double *fxx);
// one WENO5 SSP-RK3 stage of the FP64 advection on the field geometry G
// (which may carry open faces) with caller pointers, queued on st
int ls_adv_stage_f32(fcfd_ctx *c, const float *phi, float *phin, const float *nu, const float *nv, const float *nw,
                     float *l1, float *l2, int stage);
int ls_adv_stage_geom(hipStream_t st, const Geom &G, const double *phi, double *phin, const double *nu,
                      const double *nv, const double *nw, double *l1, double *l2, double dt, int stage);
int ccd_setup(fcfd_ctx *c);
int surface_norms(fcfd_ctx *c, const double *phi, int full);
int curv_kernel(fcfd_ctx *c);
int ls_advect(fcfd_ctx *c);
int ls_adv_stage(fcfd_ctx *c, const double *phi, double *phin, int stage);
int ls_advect_f32(fcfd_ctx *c, float *phi, const float *nu, const float *nv, const float *nw);
int ls_adv_adjoint_step(fcfd_ctx *c, const double *phi, double *lambda, double *ub, double *vb, double *wb);
int redis(fcfd_ctx *c, int btn, int *iters);
int mpls(fcfd_ctx *c);
int ls_mv(fcfd_ctx *c, double *mass, double *vol);
int ns_diff_source(fcfd_ctx *c);
int ns_predictor(fcfd_ctx *c, const double *ut, const double *vt, const double *wt, int write_src = 1);
int ppe_src(fcfd_ctx *c);
int projection(fcfd_ctx *c);
int node_vel(fcfd_ctx *c);
int div_check(fcfd_ctx *c, int to_host);
int mg_setup(fcfd_ctx *c);
int mg_free(fcfd_ctx *c);
int mg_drop_graphs(fcfd_ctx *c);
int prof_graph_accum(fcfd_ctx *c, int kind, int inner);
int ppe_solve(fcfd_ctx *c, double rel_tol, int precision, int *cycles, double *rel, int normalize = 1);
int copy_field(fcfd_ctx *c, double *dst, const double *src);

}  // namespace fcfd
