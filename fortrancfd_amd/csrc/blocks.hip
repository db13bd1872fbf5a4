// blocks.hip — the multi-block layout of the reference's domain
// decomposition on one MI355X: a grid split into gx x gy x gz blocks, each a
// ghosted array of its own, with the halo exchange that fills the ghosts of
// faces between blocks (`DomainDecomposition/3D/lib/pointer/ptr_roots.f90:
// 157-320`, the block ranges of `.../lib/data/branches.f90:84-121`).
//
// Layout.  Block (bx, by, bz) owns the reference's index ranges
// is = bx N / gx + 1 .. ie = (bx + 1) N / gx (integer division, 1-based) along
// each axis, and stores them as a column-major array over is-g..ie+g (z: ks-gz
// .. ke+gz, gz = 0 in 2D), i fastest.  The blocks of a field live back to back
// in one device allocation, the "stack": z-layers of blocks, in each layer the
// rows of blocks along y, in each row the blocks along x, so that block
// (bx, by, bz) starts at PZ[bz] SX SY + EZ (PY[by] SX + EY PX[bx]) with E the
// extents (length + 2 ghost) and P, S their prefix and total sums per axis.
//
// Halo exchange.  The reference copies, block by block in loop order, the
// ghost slabs of each direction from the neighbour's array, ghosts of the
// other directions included, then wraps the periodic directions.  Its face
// ghosts are right after one pass; its edge and corner ghosts take values
// from neighbours that may not have been refreshed yet in that pass, and
// settle after at most three.  One launch here writes every ghost cell of
// every block from the cell's owner: along each axis the block whose range
// holds the coordinate (wrapped on a periodic axis), or, outside the domain
// on a non-periodic axis, the boundary block itself, whose physical ghost is
// then the source.  That is the reference's settled state, and the source of
// a copy is never a cell the same launch writes, so the launch is race free
// and independent of the order of its threads.
#include <algorithm>
#include <cstring>

#include "fcfd_internal.h"

using namespace fcfd;

namespace {

constexpr int MAXB = FCFD_BLOCKS_MAX_PER_AXIS;

// everything a kernel needs to address any block of a stack
struct BlkDesc {
    int dim, g, gz;
    int n[3], nb[3], per[3];
    int start[3][MAXB + 1];   // 0-based first cell of each block; start[a][nb] = n[a]
    long pre[3][MAXB + 1];    // prefix sums of the extents (len + 2 ghost)
    long sx, sy;              // total extent sums along x and y
};

__device__ __forceinline__ int ghost_of(const BlkDesc &D, int a) { return a == 2 ? D.gz : D.g; }

__host__ __device__ __forceinline__ long blk_off(const BlkDesc &D, int bx, int by, int bz) {
    const long ey = D.pre[1][by + 1] - D.pre[1][by], ez = D.pre[2][bz + 1] - D.pre[2][bz];
    return D.pre[2][bz] * D.sx * D.sy + ez * (D.pre[1][by] * D.sx + ey * D.pre[0][bx]);
}

// owner block and its local 0-based index along axis a of global coordinate X
__device__ __forceinline__ void owner(const BlkDesc &D, int a, int X, int &ob, int &lx) {
    const int n = D.n[a], nb = D.nb[a];
    if (X < 0 || X >= n) {
        if (D.per[a]) {
            X = X < 0 ? X + n : X - n;
        } else {
            ob = X < 0 ? 0 : nb - 1;
            lx = X - D.start[a][ob];
            return;
        }
    }
    int b = 0;
    while (b + 1 < nb && X >= D.start[a][b + 1]) ++b;
    ob = b;
    lx = X - D.start[a][b];
}

// decode thread t of block b's array: local (i, j, k), 0-based interior
// indices (ghosts negative or >= len); returns false past the array
__device__ __forceinline__ bool local_cell(const BlkDesc &D, int bx, int by, int bz, long t, int &i, int &j, int &k) {
    const int lx = D.start[0][bx + 1] - D.start[0][bx], ly = D.start[1][by + 1] - D.start[1][by];
    const int lz = D.start[2][bz + 1] - D.start[2][bz];
    const long ex = lx + 2 * D.g, ey = ly + 2 * D.g, ez = lz + 2 * D.gz;
    if (t >= ex * ey * ez) return false;
    i = (int)(t % ex) - D.g;
    j = (int)((t / ex) % ey) - D.g;
    k = (int)(t / (ex * ey)) - D.gz;
    return true;
}

__device__ __forceinline__ long local_id(const BlkDesc &D, int bx, int by, int i, int j, int k) {
    const long ex = D.start[0][bx + 1] - D.start[0][bx] + 2 * D.g;
    const long ey = D.start[1][by + 1] - D.start[1][by] + 2 * D.g;
    return (long)(i + D.g) + (long)(j + D.g) * ex + (long)(k + D.gz) * ex * ey;
}

__device__ __forceinline__ void block_coords(const BlkDesc &D, int &bx, int &by, int &bz) {
    const int b = blockIdx.y;
    bx = b % D.nb[0];
    by = (b / D.nb[0]) % D.nb[1];
    bz = b / (D.nb[0] * D.nb[1]);
}

// ptrpart_sync: every ghost cell of every block from its owner (see above)
__global__ void __launch_bounds__(256) k_blocks_sync(double *__restrict__ stack, BlkDesc D) {
    int bx, by, bz, i, j, k;
    block_coords(D, bx, by, bz);
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (!local_cell(D, bx, by, bz, t, i, j, k)) return;
    const int X = D.start[0][bx] + i, Y = D.start[1][by] + j, Z = D.start[2][bz] + k;
    int ox, oy, oz, li, lj, lk;
    owner(D, 0, X, ox, li);
    owner(D, 1, Y, oy, lj);
    owner(D, 2, Z, oz, lk);
    if (ox == bx && oy == by && oz == bz && li == i && lj == j && lk == k) return;   // interior or own physical ghost
    stack[blk_off(D, bx, by, bz) + t] = stack[blk_off(D, ox, oy, oz) + local_id(D, ox, oy, li, lj, lk)];
}

// job_bc(0) on every block: each ghost cell takes the block's interior cell
// at the clamped index (the x, then y, then z zero-gradient copies over the
// full extents, `.../lib/data/branches.f90:240-290`)
__global__ void __launch_bounds__(256) k_blocks_bc(double *__restrict__ stack, BlkDesc D) {
    int bx, by, bz, i, j, k;
    block_coords(D, bx, by, bz);
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (!local_cell(D, bx, by, bz, t, i, j, k)) return;
    const int lx = D.start[0][bx + 1] - D.start[0][bx], ly = D.start[1][by + 1] - D.start[1][by];
    const int lz = D.start[2][bz + 1] - D.start[2][bz];
    const int ci = clampi(i, 0, lx - 1), cj = clampi(j, 0, ly - 1), ck = clampi(k, 0, lz - 1);
    if (ci == i && cj == j && ck == k) return;
    const long o = blk_off(D, bx, by, bz);
    stack[o + t] = stack[o + local_id(D, bx, by, ci, cj, ck)];
}

// global ghosted field <-> stack: scatter copies every cell of every block
// (ghosts included) from the field at the same global index; gather copies
// the interiors back
__global__ void __launch_bounds__(256) k_blocks_scatter(const double *__restrict__ field, double *__restrict__ stack,
                                                        BlkDesc D, Geom G) {
    int bx, by, bz, i, j, k;
    block_coords(D, bx, by, bz);
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (!local_cell(D, bx, by, bz, t, i, j, k)) return;
    stack[blk_off(D, bx, by, bz) + t] = field[G.id(D.start[0][bx] + i, D.start[1][by] + j, D.start[2][bz] + k)];
}

__global__ void __launch_bounds__(256) k_blocks_gather(const double *__restrict__ stack, double *__restrict__ field,
                                                       BlkDesc D, Geom G) {
    int bx, by, bz, i, j, k;
    block_coords(D, bx, by, bz);
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (!local_cell(D, bx, by, bz, t, i, j, k)) return;
    const int lx = D.start[0][bx + 1] - D.start[0][bx], ly = D.start[1][by + 1] - D.start[1][by];
    const int lz = D.start[2][bz + 1] - D.start[2][bz];
    if (i < 0 || i >= lx || j < 0 || j >= ly || k < 0 || k >= lz) return;
    field[G.id(D.start[0][bx] + i, D.start[1][by] + j, D.start[2][bz] + k)] = stack[blk_off(D, bx, by, bz) + t];
}

// ptrpart_check_periodic: per block of threads, the sum of squared
// differences between the periodic ghosts of the first / last blocks along
// each periodic axis and the interior cells they mirror (the reference's two
// terms per ghost cell); partial sums per workgroup, reduced on the host in
// a fixed order
__global__ void __launch_bounds__(256) k_blocks_check_per(const double *__restrict__ stack, double *__restrict__ part,
                                                          BlkDesc D) {
    __shared__ double sh[256];
    int bx, by, bz, i, j, k;
    block_coords(D, bx, by, bz);
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    double e = 0.0;
    if (local_cell(D, bx, by, bz, t, i, j, k)) {
        const int b3[3] = {bx, by, bz}, l3[3] = {i, j, k};
        int len[3];
        for (int a = 0; a < 3; ++a) len[a] = D.start[a][b3[a] + 1] - D.start[a][b3[a]];
        // the reference sums over each periodic axis a the ghost layers of
        // that axis only, over the full extents of the other two
        for (int a = 0; a < 3; ++a) {
            if (!D.per[a]) continue;
            const int ga = ghost_of(D, a);
            const bool lowg = b3[a] == 0 && l3[a] < 0 && l3[a] >= -ga;
            const bool highg = b3[a] == D.nb[a] - 1 && l3[a] >= len[a] && l3[a] < len[a] + ga;
            if (!lowg && !highg) continue;
            int o[3] = {bx, by, bz}, q[3] = {i, j, k};
            if (lowg) {   // dat(1 - i) of block 0 against dat(ie + 1 - i) of the last block
                o[a] = D.nb[a] - 1;
                q[a] = (D.start[a][D.nb[a]] - D.start[a][D.nb[a] - 1]) + l3[a];
            } else {      // dat(ie + i) of the last block against dat(i) of block 0
                o[a] = 0;
                q[a] = l3[a] - len[a];
            }
            const double d = stack[blk_off(D, bx, by, bz) + t] -
                             stack[blk_off(D, o[0], o[1], o[2]) + local_id(D, o[0], o[1], q[0], q[1], q[2])];
            e += d * d;
        }
    }
    sh[threadIdx.x] = e;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[(long)blockIdx.y * gridDim.x + blockIdx.x] = sh[0];
}

}  // namespace

struct fcfd_blocks {
    BlkDesc D;
    Geom G;             // the global field's geometry
    int nblk;
    long maxcells;      // largest block array
    long total;         // doubles in a stack
    double dt;          // advection time step (params.dt at creation)
    int device;
    double *a = nullptr, *b = nullptr, *l1 = nullptr, *l2 = nullptr;   // advection scratch stacks
    double *part = nullptr;
    std::vector<double> part_host;
};

namespace {

dim3 blk_grid(const fcfd_blocks *B) { return dim3((unsigned)((B->maxcells + 255) / 256), (unsigned)B->nblk); }

Geom block_geom(const fcfd_blocks *B, int bx, int by, int bz) {
    const BlkDesc &D = B->D;
    Geom g = B->G;
    g.nx = D.start[0][bx + 1] - D.start[0][bx];
    g.ny = D.start[1][by + 1] - D.start[1][by];
    g.nz = D.start[2][bz + 1] - D.start[2][bz];
    g.sx = g.nx + 2 * g.g;
    g.sxy = g.sx * (g.ny + 2 * g.g);
    g.ntot = g.sxy * (g.nz + 2 * g.gz);
    // faces shared with another block (or wrapped) read their ghosts
    g.ox0 = (bx > 0 || D.per[0]) ? g.g : 0;
    g.ox1 = (bx < D.nb[0] - 1 || D.per[0]) ? g.g : 0;
    g.oy0 = (by > 0 || D.per[1]) ? g.g : 0;
    g.oy1 = (by < D.nb[1] - 1 || D.per[1]) ? g.g : 0;
    g.oz0 = (bz > 0 || D.per[2]) ? g.gz : 0;
    g.oz1 = (bz < D.nb[2] - 1 || D.per[2]) ? g.gz : 0;
    return g;
}

int stack_scratch(fcfd_blocks *B) {
    if (B->a) return FCFD_OK;
    for (double **p : {&B->a, &B->b, &B->l1, &B->l2})
        FCFD_CHECK(hipMalloc((void **)p, sizeof(double) * B->total));
    return FCFD_OK;
}

}  // namespace

// ==================================================================== C ABI

extern "C" int fcfd_blocks_create(const fcfd_grid *grid, const fcfd_params *params, int gx, int gy, int gz,
                                  int xper, int yper, int zper, fcfd_blocks **out) {
    if (!grid || !out) return FCFD_ERR_ARG;
    *out = nullptr;
    if (grid->dim != 2 && grid->dim != 3) return FCFD_ERR_ARG;
    if (grid->ghc != 3 || grid->nx < 1 || grid->ny < 1 || grid->nz < 1) return FCFD_ERR_ARG;
    if (grid->dim == 2 && (grid->nz != 1 || gz != 1 || zper)) return FCFD_ERR_ARG;
    const int nb[3] = {gx, gy, gz}, n[3] = {grid->nx, grid->ny, grid->nz};
    for (int a = 0; a < 3; ++a)
        if (nb[a] < 1 || nb[a] > MAXB) return FCFD_ERR_ARG;
    fcfd_blocks *B = new fcfd_blocks();
    BlkDesc &D = B->D;
    memset(&D, 0, sizeof(D));
    D.dim = grid->dim;
    D.g = grid->ghc;
    D.gz = grid->dim == 3 ? grid->ghc : 0;
    const int per[3] = {xper != 0, yper != 0, zper != 0};
    for (int a = 0; a < 3; ++a) {
        D.n[a] = n[a];
        D.nb[a] = nb[a];
        D.per[a] = per[a];
        const int ga = a == 2 ? D.gz : D.g;
        // IS = ID N / G + 1, IE = (ID + 1) N / G (branches.f90:106-113), 0-based here
        for (int b = 0; b <= nb[a]; ++b) D.start[a][b] = (int)((long)b * n[a] / nb[a]);
        D.pre[a][0] = 0;
        for (int b = 0; b < nb[a]; ++b) {
            const int len = D.start[a][b + 1] - D.start[a][b];
            // a block thinner than the ghost width cannot hand its neighbour
            // a full halo (the reference stops there too, branches.f90:115-118)
            if (len < std::max(ga, 1)) { delete B; return FCFD_ERR_ARG; }
            D.pre[a][b + 1] = D.pre[a][b] + len + 2 * ga;
        }
    }
    D.sx = D.pre[0][nb[0]];
    D.sy = D.pre[1][nb[1]];
    B->nblk = nb[0] * nb[1] * nb[2];
    B->total = D.pre[2][nb[2]] * D.sx * D.sy;
    long mx = 0;
    for (int bz = 0; bz < nb[2]; ++bz)
        for (int by = 0; by < nb[1]; ++by)
            for (int bx = 0; bx < nb[0]; ++bx)
                mx = std::max(mx, (D.pre[0][bx + 1] - D.pre[0][bx]) * (D.pre[1][by + 1] - D.pre[1][by]) *
                                      (D.pre[2][bz + 1] - D.pre[2][bz]));
    B->maxcells = mx;
    Geom &G = B->G;
    G.dim = grid->dim;
    G.nx = n[0]; G.ny = n[1]; G.nz = n[2];
    G.g = D.g; G.gz = D.gz;
    G.sx = G.nx + 2 * G.g;
    G.sxy = G.sx * (G.ny + 2 * G.g);
    G.ntot = G.sxy * (G.nz + 2 * G.gz);
    G.dx = (grid->xend - grid->xstart) / G.nx;
    G.dy = (grid->yend - grid->ystart) / G.ny;
    G.dz = grid->dim == 3 ? (grid->zend - grid->zstart) / G.nz : 1.0;
    B->dt = params ? params->dt : 0.0;
    if (hipGetDevice(&B->device) != hipSuccess) { delete B; return FCFD_ERR_HIP; }
    *out = B;
    return FCFD_OK;
}

extern "C" int fcfd_blocks_destroy(fcfd_blocks *B) {
    if (!B) return FCFD_ERR_ARG;
    for (double *p : {B->a, B->b, B->l1, B->l2, B->part})
        if (p) hipFree(p);
    delete B;
    return FCFD_OK;
}

extern "C" int64_t fcfd_blocks_stack_size(const fcfd_blocks *B) { return B ? (int64_t)B->total : -1; }

extern "C" int fcfd_blocks_range(const fcfd_blocks *B, int bx, int by, int bz, int32_t range[6], int64_t *offset) {
    if (!B || !range) return FCFD_ERR_ARG;
    const BlkDesc &D = B->D;
    if (bx < 0 || bx >= D.nb[0] || by < 0 || by >= D.nb[1] || bz < 0 || bz >= D.nb[2]) return FCFD_ERR_ARG;
    const int b3[3] = {bx, by, bz};
    for (int a = 0; a < 3; ++a) {
        range[2 * a] = D.start[a][b3[a]] + 1;
        range[2 * a + 1] = D.start[a][b3[a] + 1];
    }
    if (offset) *offset = blk_off(D, bx, by, bz);
    return FCFD_OK;
}

extern "C" int fcfd_blocks_scatter(fcfd_blocks *B, const double *field, double *stack, void *stream) {
    if (!B || !field || !stack) return FCFD_ERR_ARG;
    hipLaunchKernelGGL(k_blocks_scatter, blk_grid(B), dim3(256), 0, (hipStream_t)stream, field, stack, B->D, B->G);
    FCFD_CHECK(hipGetLastError());
    return FCFD_OK;
}

extern "C" int fcfd_blocks_gather(fcfd_blocks *B, const double *stack, double *field, void *stream) {
    if (!B || !field || !stack) return FCFD_ERR_ARG;
    hipLaunchKernelGGL(k_blocks_gather, blk_grid(B), dim3(256), 0, (hipStream_t)stream, stack, field, B->D, B->G);
    FCFD_CHECK(hipGetLastError());
    return FCFD_OK;
}

extern "C" int fcfd_blocks_bc(fcfd_blocks *B, double *stack, void *stream) {
    if (!B || !stack) return FCFD_ERR_ARG;
    hipLaunchKernelGGL(k_blocks_bc, blk_grid(B), dim3(256), 0, (hipStream_t)stream, stack, B->D);
    FCFD_CHECK(hipGetLastError());
    return FCFD_OK;
}

extern "C" int fcfd_blocks_sync(fcfd_blocks *B, double *stack, void *stream) {
    if (!B || !stack) return FCFD_ERR_ARG;
    hipLaunchKernelGGL(k_blocks_sync, blk_grid(B), dim3(256), 0, (hipStream_t)stream, stack, B->D);
    FCFD_CHECK(hipGetLastError());
    return FCFD_OK;
}

extern "C" int fcfd_blocks_check_per(fcfd_blocks *B, const double *stack, double *err, void *stream) {
    if (!B || !stack || !err) return FCFD_ERR_ARG;
    const dim3 grd = blk_grid(B);
    const long np = (long)grd.x * grd.y;
    if (!B->part) {
        FCFD_CHECK(hipMalloc((void **)&B->part, sizeof(double) * np));
        B->part_host.resize(np);
    }
    hipStream_t st = (hipStream_t)stream;
    hipLaunchKernelGGL(k_blocks_check_per, grd, dim3(256), 0, st, stack, B->part, B->D);
    FCFD_CHECK(hipGetLastError());
    FCFD_CHECK(hipMemcpyAsync(B->part_host.data(), B->part, sizeof(double) * np, hipMemcpyDeviceToHost, st));
    FCFD_CHECK(hipStreamSynchronize(st));
    double e = 0.0;
    for (long q = 0; q < np; ++q) e += B->part_host[q];
    *err = e;
    return FCFD_OK;
}

extern "C" int fcfd_blocks_level_set_rk3_solver(fcfd_blocks *B, double *phi, const double *nu, const double *nv,
                                                const double *nw, void *stream) {
    if (!B || !phi || !nu || !nv || (B->G.dim == 3 && !nw)) return FCFD_ERR_ARG;
    FCFD_TRY(stack_scratch(B));
    hipStream_t st = (hipStream_t)stream;
    const BlkDesc &D = B->D;
    // stage inputs and outputs: phi -> a -> b -> phi
    const double *in[3] = {phi, B->a, B->b};
    double *out[3] = {B->a, B->b, phi};
    for (int s = 0; s < 3; ++s) {
        for (int bz = 0; bz < D.nb[2]; ++bz)
            for (int by = 0; by < D.nb[1]; ++by)
                for (int bx = 0; bx < D.nb[0]; ++bx) {
                    const Geom g = block_geom(B, bx, by, bz);
                    const long o = blk_off(D, bx, by, bz);
                    FCFD_TRY(ls_adv_stage_geom(st, g, in[s] + o, out[s] + o, nu + o, nv + o, nw ? nw + o : nullptr,
                                               B->l1 + o, B->l2 + o, B->dt, s));
                }
        FCFD_TRY(fcfd_blocks_bc(B, out[s], stream));
        FCFD_TRY(fcfd_blocks_sync(B, out[s], stream));
    }
    return FCFD_OK;
}
