// mg.hip — multigrid solve of the variable-coefficient pressure Poisson
// equation on MI355X.
//
// Operator: the reference's ppe_sor_init
// (`Two phase flows/newOpenMP/3D+IBM/src/ns_solver/ns_ppe_solver.f90:1-65`):
//   sum_nb c_nb (p_nb - p_P) = f,  c_nb = 2/(rho_P + rho_nb)/h^2 on interior
//   faces and 0 on wall faces (the reference folds the wall coefficient into
//   c_P, which is the same homogeneous Neumann operator),
// with f = div(u*)/dt and the reference's normalisation mean(p) = 0
// (ns_ppe_solver.f90:316-339).  The reference relaxes this system with an
// in-place point iteration; here it is solved by V-cycles to a relative
// residual ||f - A p|| / ||f||.
//
// Multigrid: cell-centred, red-black Gauss-Seidel V(nu1,nu2) cycles.  Level 0
// is the ghosted field layout and takes its face coefficients straight from
// rho (one read of rho instead of six coefficient arrays).  Coarse levels
// store face coefficients beta = 1/rho_face averaged over the fine faces that
// make up a coarse face, restriction averages the children (the reference's
// 8-cell mean, `DomainDecomposition/3D/src/others/multigrid_solver.f90:68-72`),
// prolongation is piecewise constant.  An axis is coarsened only while its
// spacing is within 1.5x of the finest spacing, so anisotropic cells
// (dz = 2 dx) are first semi-coarsened to isotropy.  The coarsest level is
// solved in FP64 by Jacobi-preconditioned CG inside one workgroup.
//
// Precision modes: FCFD_PPE_FP64 runs the V-cycles on p in FP64.
// FCFD_PPE_MIXED keeps p, f and the residual in FP64 and runs each V-cycle in
// FP32 on the error equation A e = r (iterative refinement / defect
// correction); only the outer residual, the update, the coarsest solve and
// the norms are FP64.
#include <algorithm>
#include <cmath>
#include <cstdlib>

#include "fcfd_internal.h"

using namespace fcfd;

namespace fcfd {

// layout of one level: interior (nx,ny,nz), ghost width g (gz = 0 in 2D).
// skip: while a flexible-CG iteration is queued ahead of the host's reading
// of the previous one's status, its kernels read this device flag first and
// return at once when the solve has already ended (see ppe_solve_t)
struct Lay {
    int nx, ny, nz, g, gz;
    long sx, sxy;
    const int *skip = nullptr;
    __host__ __device__ long id(int i, int j, int k) const {
        return (long)(i + g) + (long)(j + g) * sx + (long)(k + gz) * sxy;
    }
    __device__ bool skipped() const { return skip && *skip != 0; }
};

// the skip flag given to the layouts made on this host thread (set while a
// pipelined flexible-CG iteration is launched or captured)
static thread_local const int *t_skip = nullptr;

static Lay lay_of(const Geom &G) {
    Lay L;
    L.nx = G.nx; L.ny = G.ny; L.nz = G.nz; L.g = G.g; L.gz = G.gz; L.sx = G.sx; L.sxy = G.sxy;
    L.skip = t_skip;
    return L;
}

static Lay lay_of(const MgLevel &M, int dim) {
    Lay L;
    L.nx = M.nx; L.ny = M.ny; L.nz = M.nz; L.g = 1; L.gz = dim == 3 ? 1 : 0; L.sx = M.sx; L.sxy = M.sxy;
    L.skip = t_skip;
    return L;
}

// one case's layout: a 2D level of a batched stack holds one case per slice
static Lay lay_case(const MgLevel &M, int dim) {
    Lay L = lay_of(M, dim);
    if (dim == 2) L.nz = 1;
    return L;
}

// cells of one case on level M
static long case_cells(const MgLevel &M, int dim) { return (long)M.nx * M.ny * (dim == 3 ? M.nz : 1); }

// Colour-split layout of the FP32 level-0 arrays of the mixed cycle.  Cell
// (i,j,k) has colour c = (i+j+k) & 1 (2D: (i+j) & 1, k being the case of a
// batched stack) and lives in colour array c (the first
// or second `half` elements) at slot m = i >> 1 of row (j,k); each colour
// array has mh = (nx+1)/2 slots per row plus one ghost slot on every side.
// All six neighbours of a cell have the other colour, so a red-black
// half-sweep streams whole cache lines of exactly the arrays it needs.
// Slots that map to no cell (ghosts, and i = nx for odd nx) stay zero.
struct Split {
    int nx, ny, nz, kz;   // cells; kz = 1 in 3D (ghost plane below k = 0), 0 in 2D
    long sx, sxy, half;   // row and plane pitch of one colour array; its size
    __host__ __device__ long row(int j, int k) const { return (long)(j + 1) * sx + (long)(k + kz) * sxy; }
    __host__ __device__ long at(int i, int j, int k) const {
        return ((i + j + (kz ? k : 0)) & 1) * half + (long)((i >> 1) + 1) + row(j, k);
    }
};

static Split split_of(const Geom &G) {
    Split S;
    S.nx = G.nx; S.ny = G.ny; S.nz = G.nz; S.kz = G.dim == 3 ? 1 : 0;
    S.sx = (G.nx + 1) / 2 + 2;
    S.sxy = S.sx * (G.ny + 2);
    S.half = S.sxy * (G.dim == 3 ? G.nz + 2 : G.nz);   // 2D: one plane per case of a batched stack
    return S;
}

struct Inv2 {
    double x, y, z;  // 1/h^2 per axis
};

// face coefficient of level 0 between cells with densities a and b.  The
// FP32 operator only preconditions (the FP64 outer iteration owns the
// accuracy), so it uses the hardware reciprocal (about 1 ulp) instead of an
// IEEE division sequence; FP64 stays correctly rounded.
template <typename T>
__device__ __forceinline__ T beta_rho(T a, T b) {
    return (T)2 / (a + b);
}
template <>
__device__ __forceinline__ float beta_rho<float>(float a, float b) {
    return 2.0f * __builtin_amdgcn_rcpf(a + b);
}

// Gauss-Seidel correction (sum c (e_nb - e_P) - f) / sum c, same precision
// rule as beta_rho
template <typename T>
__device__ __forceinline__ T gs_step(T num, T sc) {
    return num / sc;
}
template <>
__device__ __forceinline__ float gs_step<float>(float num, float sc) {
    return num * __builtin_amdgcn_rcpf(sc);
}

// Wall-masked face coefficient c = beta_rho(r, rn) * h, or 0 across a wall.
// The neighbour density is always loaded (ghost cells exist on every level)
// and the wall test selects without a branch.
template <typename T>
__device__ __forceinline__ T face(bool open, T r, T rn, T h) {
    const T c = beta_rho<T>(r, rn) * h;
    return open ? c : (T)0;
}

// Coarse-grid correction folded into the first post-smoothing sweep: the
// value of a fine cell is e + ec(parent) until that cell has been updated.
template <typename T>
struct Prol {
    const T *ec;
    Lay C;
    int lx, ly, lz;   // log2 of the coarsening factor (1 or 2) per axis
    // parent of fine (i, j, k) by arithmetic shifts; i = -1 or nx maps to a
    // coarse ghost cell (zero, and its face coefficient is 0)
    __device__ __forceinline__ T at(int i, int j, int k) const { return ec[C.id(i >> lx, j >> ly, k >> lz)]; }
};

// Face coefficients of cell (i,j,k): sum_c = sum c_nb and
// sum_ce = sum c_nb (e_nb - e_P), and the centre value ep.  FINE: from rho;
// otherwise from stored beta arrays whose wall and ghost entries are zero.
// PROL 0: plain values; 1: centre and neighbours get the coarse correction
// (first colour after prolongation); 2: only the centre gets it (second
// colour: its neighbours are first-colour cells, already corrected).
// 3 and 4: the same two half-sweeps for a zero initial guess: the values are
// taken as 0 instead of read (3: centre and neighbours, 4: centre only), so
// the error array needs no clearing between cycles.
template <typename T, int DIM, bool FINE, int PROL, typename CT = T>
__device__ __forceinline__ void stencil(const T *__restrict__ e, const T *__restrict__ rho,
                                        const CT *__restrict__ bx, const CT *__restrict__ by,
                                        const CT *__restrict__ bz, const Lay &L, Inv2 ih, int i, int j, int k,
                                        long id, const Prol<T> &P, T &sum_c, T &sum_ce, T &ep) {
    // the values first, so that every load of the cell issues before the
    // coefficient arithmetic (one memory round trip instead of two)
    auto nb = [&](long o, int ii, int jj, int kk) -> T {
        if (PROL == 3) return (T)0;
        T v = e[id + o];
        if (PROL == 1) v += P.at(ii, jj, kk);
        return v;
    };
    const T vE = nb(1, i + 1, j, k), vW = nb(-1, i - 1, j, k), vN = nb(L.sx, i, j + 1, k),
            vS = nb(-L.sx, i, j - 1, k);
    const T vT = DIM == 3 ? nb(L.sxy, i, j, k + 1) : (T)0, vB = DIM == 3 ? nb(-L.sxy, i, j, k - 1) : (T)0;
    if (PROL >= 3) ep = (T)0;
    else {
        ep = e[id];
        if (PROL != 0) ep += P.at(i, j, k);
    }
    T cw, ce_, cs, cn, cb = 0, ct = 0;
    if (FINE) {
        T r = rho[id];
        ce_ = face<T>(i < L.nx - 1, r, rho[id + 1], (T)ih.x);
        cw = face<T>(i > 0, r, rho[id - 1], (T)ih.x);
        cn = face<T>(j < L.ny - 1, r, rho[id + L.sx], (T)ih.y);
        cs = face<T>(j > 0, r, rho[id - L.sx], (T)ih.y);
        if (DIM == 3) {
            ct = face<T>(k < L.nz - 1, r, rho[id + L.sxy], (T)ih.z);
            cb = face<T>(k > 0, r, rho[id - L.sxy], (T)ih.z);
        }
    } else {
        ce_ = (T)bx[id] * (T)ih.x;
        cw = (T)bx[id - 1] * (T)ih.x;
        cn = (T)by[id] * (T)ih.y;
        cs = (T)by[id - L.sx] * (T)ih.y;
        if (DIM == 3) {
            ct = (T)bz[id] * (T)ih.z;
            cb = (T)bz[id - L.sxy] * (T)ih.z;
        }
    }
    sum_c = ce_ + cw + cn + cs;
    sum_ce = ce_ * (vE - ep) + cw * (vW - ep) + cn * (vN - ep) + cs * (vS - ep);
    if (DIM == 3) {
        sum_c += ct + cb;
        sum_ce += ct * (vT - ep) + cb * (vB - ep);
    }
}

// (A e)_P = sum_nb c_nb (e_nb - e_P), formed from the differences: in the
// gas (rho ~ 1e-3) c_nb is ~1000x larger than in the liquid, and the form
// sum c e_nb - (sum c) e_P would cancel terms of size c |e| and put the
// round-off floor of the residual above 1e-10 relative on fine grids.
template <typename T, int DIM, bool FINE, typename CT = T>
__device__ __forceinline__ T apply_A(const T *__restrict__ e, const T *__restrict__ rho, const CT *__restrict__ bx,
                                     const CT *__restrict__ by, const CT *__restrict__ bz, const Lay &L, Inv2 ih,
                                     int i, int j, int k, long id) {
    const T ep = e[id];
    T acc;
    if (FINE) {
        T r = rho[id];
        acc = face<T>(i < L.nx - 1, r, rho[id + 1], (T)ih.x) * (e[id + 1] - ep);
        acc += face<T>(i > 0, r, rho[id - 1], (T)ih.x) * (e[id - 1] - ep);
        acc += face<T>(j < L.ny - 1, r, rho[id + L.sx], (T)ih.y) * (e[id + L.sx] - ep);
        acc += face<T>(j > 0, r, rho[id - L.sx], (T)ih.y) * (e[id - L.sx] - ep);
        if (DIM == 3) {
            acc += face<T>(k < L.nz - 1, r, rho[id + L.sxy], (T)ih.z) * (e[id + L.sxy] - ep);
            acc += face<T>(k > 0, r, rho[id - L.sxy], (T)ih.z) * (e[id - L.sxy] - ep);
        }
    } else {
        acc = (T)bx[id] * (T)ih.x * (e[id + 1] - ep) + (T)bx[id - 1] * (T)ih.x * (e[id - 1] - ep) +
              (T)by[id] * (T)ih.y * (e[id + L.sx] - ep) + (T)by[id - L.sx] * (T)ih.y * (e[id - L.sx] - ep);
        if (DIM == 3)
            acc += (T)bz[id] * (T)ih.z * (e[id + L.sxy] - ep) + (T)bz[id - L.sxy] * (T)ih.z * (e[id - L.sxy] - ep);
    }
    return acc;
}

// Dot products of the flexible CG folded into the last half-sweep of a
// V-cycle: r.z and r.zold over both cells of every thread's pair (z = the
// cycle's output), block sums into partial[b] and partial[nb + b].
template <typename T>
struct Dots {
    const double *r;
    const T *zo;
    double *partial;
};

// one colour of a red-black Gauss-Seidel sweep: e_P = (sum c e_nb - f) / sum c.
// Thread x covers two cells of a row; it updates the one whose colour matches.
template <typename T, int DIM, bool FINE, int PROL, bool DOTS, typename CT = T>
__global__ void __launch_bounds__(256) k_smooth(T *__restrict__ e, const T *__restrict__ f, const T *__restrict__ rho,
                                                const CT *__restrict__ bx, const CT *__restrict__ by,
                                                const CT *__restrict__ bz, Lay L, Inv2 ih, int color, Prol<T> P,
                                                Dots<T> D) {
    if (L.skipped()) return;
    int bx_, by_, bz_;
    xcd_block(bx_, by_, bz_);
    const int j = by_ * blockDim.y + threadIdx.y, k = bz_;
    const int i0 = 2 * (bx_ * blockDim.x + threadIdx.x), par_ = (j + (DIM == 3 ? k : 0) + color) & 1;
    const int i = i0 + par_;
    const bool upd = i < L.nx && j < L.ny;
    T enew = (T)0;
    if (upd) {
        const long id = L.id(i, j, k);
        T sc, sce, ep;
        stencil<T, DIM, FINE, PROL, CT>(e, rho, bx, by, bz, L, ih, i, j, k, id, P, sc, sce, ep);
        // correction form e_P += (sum c (e_nb - e_P) - f) / sum c: the Gauss-Seidel
        // update without cancelling terms of size c |e|.  Formed for every
        // cell and selected (sc = 0 only off the grid), so that no branch
        // holds back the value loads until the coefficients have arrived.
        const T d = gs_step<T>(sce - f[id], sc);
        enew = sc > (T)0 ? ep + d : ep;
        e[id] = enew;
    }
    if (DOTS) {
        __shared__ double sh[4];
        double rz = 0.0, rzo = 0.0;
        if (upd) {
            const long id = L.id(i, j, k);
            rz = D.r[id] * (double)enew;
            rzo = D.r[id] * (double)D.zo[id];
        }
        const int ip = i0 + 1 - par_;   // the pair's other cell: final from the previous half-sweep
        if (ip < L.nx && j < L.ny) {
            const long idp = L.id(ip, j, k);
            rz += D.r[idp] * (double)e[idp];
            rzo += D.r[idp] * (double)D.zo[idp];
        }
        rz = block_sum256(rz, sh);
        rzo = block_sum256(rzo, sh);
        if (threadIdx.x + threadIdx.y == 0) {
            const long b = bx_ + (long)gridDim.x * (by_ + (long)gridDim.y * bz_);
            const long nb = (long)gridDim.x * gridDim.y * gridDim.z;
            D.partial[b] = rz;
            D.partial[nb + b] = rzo;
        }
    }
}

// Gauss-Seidel value of one cell of the colour-split FP32 level 0 from its
// density r, the neighbour densities (E, W, N, S, T, B) and the neighbour
// values e*: (sum c e_nb - f) / sum c, the walls closing faces.  ZERO: the
// neighbour values are 0 (zero initial guess).  The sums use explicit fmaf,
// so the rounding does not depend on how the caller is compiled.
template <int DIM, bool ZERO>
__device__ __forceinline__ float split_gs(float r, float rE, float rW, float rN, float rS, float rT, float rB,
                                          float eE, float eW, float eN, float eS, float eT, float eB, float fv,
                                          int i, int j, int k, const Split &S, Inv2 ih) {
    const float cE = face<float>(i < S.nx - 1, r, rE, (float)ih.x);
    const float cW = face<float>(i > 0, r, rW, (float)ih.x);
    const float cN = face<float>(j < S.ny - 1, r, rN, (float)ih.y);
    const float cS = face<float>(j > 0, r, rS, (float)ih.y);
    float cT = 0.f, cB = 0.f;
    if (DIM == 3) {
        cT = face<float>(k < S.nz - 1, r, rT, (float)ih.z);
        cB = face<float>(k > 0, r, rB, (float)ih.z);
    }
    float sc = cE + cW + cN + cS, sce = 0.f;
    if (!ZERO) sce = fmaf(cS, eS, fmaf(cN, eN, fmaf(cW, eW, cE * eE)));
    if (DIM == 3) {
        sc += cT + cB;
        if (!ZERO) sce += fmaf(cB, eB, cT * eT);
    }
    // formed unconditionally and selected (sc = 0 only off the grid): a
    // branch here would hold back the value loads behind the density loads
    const float d = (sce - fv) * __builtin_amdgcn_rcpf(sc);
    return sc > 0.f ? d : 0.f;
}

// One colour of the red-black Gauss-Seidel sweep on the colour-split FP32
// level 0 (the mixed cycle): e_P = (sum c e_nb - f) / sum c.  A half-sweep
// reads the other colour's e, this colour's f and both colours' rho and
// writes this colour's e: 2.5 values per cell, where an interleaved array
// moves whole lines of e, f, rho and e again.  The centre's previous value
// does not enter the update, so MODE 0: plain; 1: the neighbours carry the
// coarse-grid correction e + ec(parent) (first half-sweep after
// prolongation; the second colour reads first-colour cells already
// updated); 2: zero initial guess, the neighbours are 0 (first pre-sweep
// half, so the error array needs no clearing).  DOTS: the CG dots of the
// thread's pair (2m, 2m+1) as in k_smooth.
template <int DIM, int MODE, bool DOTS, typename RT>
__global__ void __launch_bounds__(256) k_smooth_split(float *__restrict__ e, const float *__restrict__ f,
                                                      const RT *__restrict__ rho, Split S, Inv2 ih, int color,
                                                      Prol<float> P, Dots<float> D, Lay L) {
    if (L.skipped()) return;
    int bx_, by_, bz_;
    xcd_block(bx_, by_, bz_);
    const int m = bx_ * blockDim.x + threadIdx.x, j = by_ * blockDim.y + threadIdx.y, k = bz_;
    const int s = (j + (DIM == 3 ? k : 0) + color) & 1, i = 2 * m + s;
    const long id = (long)(m + 1) + S.row(j, k);
    const long cc = color ? S.half : 0, oc = S.half - cc;   // this colour, the other colour
    const bool upd = i < S.nx && j < S.ny;
    // the dots' operands of this cell and of the pair's other cell, loaded
    // before the update's store: the other colour's e is not written by this
    // half-sweep, but the compiler cannot tell the two halves of e apart and
    // would issue these loads one round trip each after the store
    const int ip = 2 * m + 1 - s;   // the pair's other cell: final from the previous half-sweep
    const bool oth = ip < S.nx && j < S.ny;
    float zo_c = 0.f, f_o = 0.f, e_o = 0.f, zo_o = 0.f;
    if (DOTS) {   // unconditional, at a valid slot for threads past the grid
        const long idd = (upd || oth) ? id : S.row(0, 0) + 1;
        zo_c = D.zo[cc + idd];
        f_o = f[oc + idd];
        e_o = e[oc + idd];
        zo_o = D.zo[oc + idd];
    }
    float enew = 0.f;
    if (upd) {
        const RT *ro = rho + oc;
        const float *eo = e + oc;
        // the x neighbours i-1 and i+1 are slots m-1+s and m+s of the other colour
        const long xw = id - 1 + s, xe = id + s;
        auto nb = [&](long o, int ii, int jj, int kk) -> float {
            if (MODE == 2) return 0.f;
            float v = eo[o];
            if (MODE == 1) v += P.at(ii, jj, kk);
            return v;
        };
        const long ot = DIM == 3 ? id + S.sxy : id, ob = DIM == 3 ? id - S.sxy : id;
        enew = split_gs<DIM, MODE == 2>((float)rho[cc + id], (float)ro[xe], (float)ro[xw], (float)ro[id + S.sx],
                                        (float)ro[id - S.sx], DIM == 3 ? (float)ro[ot] : 0.f,
                                        DIM == 3 ? (float)ro[ob] : 0.f, nb(xe, i + 1, j, k), nb(xw, i - 1, j, k),
                                        nb(id + S.sx, i, j + 1, k), nb(id - S.sx, i, j - 1, k),
                                        DIM == 3 ? nb(ot, i, j, k + 1) : 0.f, DIM == 3 ? nb(ob, i, j, k - 1) : 0.f,
                                        f[cc + id], i, j, k, S, ih);
        e[cc + id] = enew;
    }
    if (DOTS) {
        // r is taken as f, the FP32 right-hand side of this cycle (r rounded
        // to FP32), which the update has just read for this cell: the dots only
        // set the flexible-CG step lengths, while the FP64 residual update
        // keeps the outer iteration exact, and the pair costs 4 B per cell
        // instead of the 8 B of the FP64 residual
        __shared__ double sh[4];
        double rz = 0.0, rzo = 0.0;
        if (upd) {
            const double rv = (double)f[cc + id];
            rz = rv * (double)enew;
            rzo = rv * (double)zo_c;
        }
        if (oth) {
            const double rv = (double)f_o;
            rz += rv * (double)e_o;
            rzo += rv * (double)zo_o;
        }
        rz = block_sum256(rz, sh);
        rzo = block_sum256(rzo, sh);
        if (threadIdx.x + threadIdx.y == 0) {
            const long b = bx_ + (long)gridDim.x * (by_ + (long)gridDim.y * bz_);
            const long nb = (long)gridDim.x * gridDim.y * gridDim.z;
            D.partial[b] = rz;
            D.partial[nb + b] = rzo;
        }
    }
}

// (A e)_P of the colour-split FP32 level 0, difference form as apply_A
template <int DIM, typename RT>
__device__ __forceinline__ float apply_A_split(const float *__restrict__ e, const RT *__restrict__ rho,
                                               const Split &S, Inv2 ih, int i, int j, int k) {
    const long a = S.at(i, j, k);
    const float ep = e[a], r = (float)rho[a];
    long n = S.at(i + 1, j, k);
    float acc = face<float>(i < S.nx - 1, r, (float)rho[n], (float)ih.x) * (e[n] - ep);
    n = S.at(i - 1, j, k);
    acc += face<float>(i > 0, r, (float)rho[n], (float)ih.x) * (e[n] - ep);
    n = S.at(i, j + 1, k);
    acc += face<float>(j < S.ny - 1, r, (float)rho[n], (float)ih.y) * (e[n] - ep);
    n = S.at(i, j - 1, k);
    acc += face<float>(j > 0, r, (float)rho[n], (float)ih.y) * (e[n] - ep);
    if (DIM == 3) {
        n = S.at(i, j, k + 1);
        acc += face<float>(k < S.nz - 1, r, (float)rho[n], (float)ih.z) * (e[n] - ep);
        n = S.at(i, j, k - 1);
        acc += face<float>(k > 0, r, (float)rho[n], (float)ih.z) * (e[n] - ep);
    }
    return acc;
}

// k_resid_restrict for the colour-split FP32 level 0 (C2 as there)
template <int DIM, typename RT, int C2 = 0>
__global__ void __launch_bounds__(256) k_resid_restrict_split(const float *__restrict__ e, const float *__restrict__ f,
                                                              const RT *__restrict__ rho, Split S, Inv2 ih,
                                                              float *__restrict__ fc, float *__restrict__ ec, Lay C,
                                                              int cx, int cy, int cz) {
    if (C.skipped()) return;
    int bx_, by_, bz_;
    xcd_block(bx_, by_, bz_);
    const int I = bx_ * blockDim.x + threadIdx.x, J = by_ * blockDim.y + threadIdx.y, K = bz_;
    if (I >= C.nx || J >= C.ny) return;
    if (C2) { cx = 2; cy = 2; cz = C2; }
    float acc = 0.f;
    int nch = 0;
#pragma unroll
    for (int kk = 0; kk < (C2 ? C2 : cz); ++kk)
#pragma unroll
        for (int jj = 0; jj < (C2 ? 2 : cy); ++jj)
#pragma unroll
            for (int ii = 0; ii < (C2 ? 2 : cx); ++ii) {
                const int i = I * cx + ii, j = J * cy + jj, k = K * cz + kk;
                // the last aggregate of an odd axis has one child
                if (i >= S.nx || j >= S.ny || (DIM == 3 && k >= S.nz)) continue;
                acc += f[S.at(i, j, k)] - apply_A_split<DIM, RT>(e, rho, S, ih, i, j, k);
                ++nch;
            }
    const long cid = C.id(I, J, K);
    fc[cid] = acc / (float)nch;
    ec[cid] = 0.f;
}

// FP32 (and, for the FP16-density smoother, FP16) density of level 0 in the
// colour-split layout
__global__ void k_rho_split(const double *__restrict__ rho, Lay L, float *__restrict__ out,
                            _Float16 *__restrict__ out16, Split S) {
    int bx_, by_, bz_;
    xcd_block(bx_, by_, bz_);
    const int i = bx_ * blockDim.x + threadIdx.x, j = by_ * blockDim.y + threadIdx.y, k = bz_;
    if (i >= L.nx || j >= L.ny) return;
    const double r = rho[L.id(i, j, k)];
    out[S.at(i, j, k)] = (float)r;
    if (out16) out16[S.at(i, j, k)] = (_Float16)r;
}

// NS full red-black sweeps of a coarse level (stored face coefficients,
// one-ghost layout) in one launch.  A dependent launch costs ~4 us however
// small its grid, and the small levels take 6 sweeps each.  Each block owns
// an output tile and stages the tile plus a halo of H = 2 NS cells of e, f
// and the face coefficients in LDS; half-sweep h updates its colour's cells
// within distance 2 NS - 1 - h of the tile, so after the last half-sweep
// the tile holds exactly what 2 NS single-colour launches produce (same
// update, same operation order as stencil() + gs_step).  Positions outside
// the ghosted array are staged as 0 and only interior cells are updated.
// The result goes to a second array eo: blocks stage their halo from ei
// while other blocks finish, so writing back into ei would race.
// prol: the coarse-grid correction ec(parent) is added to every staged
// value, which is what the folded first post-sweep (PROL 1 then 2) reads.
template <int DIM>
struct FusedTile {
    static constexpr int TX = DIM == 3 ? 16 : 32, TY = DIM == 3 ? 8 : 16, TZ = DIM == 3 ? 8 : 1;
};
constexpr int FUSED_THREADS = 1024;
// PROL: the first launch after a prolongation adds the coarse correction as
// it stages e (a template argument: a run-time flag put a branch between the
// staging rounds' loads, one memory round trip each)
template <int DIM, int NS, bool PROL>
__global__ void __launch_bounds__(FUSED_THREADS) k_smooth_fused(const float *__restrict__ e, float *__restrict__ eo,
                                                                const float *__restrict__ f,
                                                                const float *__restrict__ bx,
                                                                const float *__restrict__ by,
                                                                const float *__restrict__ bz, Lay L, Inv2 ih,
                                                                int color0, Prol<float> P) {
    if (L.skipped()) return;
    constexpr int H = 2 * NS, TX = FusedTile<DIM>::TX, TY = FusedTile<DIM>::TY, TZ = FusedTile<DIM>::TZ;
    constexpr int EX = TX + 2 * H, EY = TY + 2 * H, EZ = DIM == 3 ? TZ + 2 * H : 1, NE = EX * EY * EZ;
    constexpr int SY = EX, SZ = EX * EY;
    __shared__ float se[NE], sf[NE], sbx[NE], sby[NE], sbz[DIM == 3 ? NE : 1];
    int bx_, by_, bz_;
    xcd_block(bx_, by_, bz_);
    const int t = threadIdx.x;
    const int i0 = bx_ * TX - H, j0 = by_ * TY - H, k0 = DIM == 3 ? bz_ * TZ - H : bz_;   // 2D: the slice
    const float hx = (float)ih.x, hy = (float)ih.y, hz = (float)ih.z;
    // staging: every load of the block's NIT rounds is issued before the
    // first is waited for (a loop with a branch around its loads costs one
    // memory latency per round); positions outside the ghosted array load a
    // clamped in-array address and stage 0
    constexpr int NIT = (NE + FUSED_THREADS - 1) / FUSED_THREADS;
    float ev[NIT], fv[NIT], b1[NIT], b2[NIT], b3[NIT];
#pragma unroll
    for (int it = 0; it < NIT; ++it) {
        const int q = min(t + it * FUSED_THREADS, NE - 1);
        const int qx = q % EX, qy = (q / EX) % EY, qz = q / SZ;
        const int i = i0 + qx, j = j0 + qy, k = k0 + qz;
        const bool inarr = i >= -1 && i <= L.nx && j >= -1 && j <= L.ny && (DIM == 2 || (k >= -1 && k <= L.nz));
        const int ic = clampi(i, -1, L.nx), jc = clampi(j, -1, L.ny), kc = DIM == 3 ? clampi(k, -1, L.nz) : k;
        const long id = L.id(ic, jc, kc);
        float e_ = e[id];
        if (PROL) e_ += P.at(ic, jc, kc);
        const float f_ = f[id], x_ = bx[id], y_ = by[id], z_ = DIM == 3 ? bz[id] : 0.f;
        ev[it] = inarr ? e_ : 0.f;
        fv[it] = inarr ? f_ : 0.f;
        b1[it] = inarr ? x_ : 0.f;
        b2[it] = inarr ? y_ : 0.f;
        b3[it] = inarr ? z_ : 0.f;
    }
#pragma unroll
    for (int it = 0; it < NIT; ++it) {
        const int q = t + it * FUSED_THREADS;
        if (q < NE) {
            se[q] = ev[it];
            sf[q] = fv[it];
            sbx[q] = b1[it];
            sby[q] = b2[it];
            if (DIM == 3) sbz[q] = b3[it];
        }
    }
    __syncthreads();
    // half-sweep h: cells of colour (color0 + h) & 1 within distance
    // D = 2 NS - 1 - h of the tile; the loop is unrolled, so the region
    // extents are constants and only the colour's cells are visited (x
    // steps by 2 from the row's first cell of that colour)
#pragma unroll
    for (int h = 0; h < 2 * NS; ++h) {
        const int color = (color0 + h) & 1, D = 2 * NS - 1 - h;
        const int rx = TX + 2 * D, ry = TY + 2 * D, rz = DIM == 3 ? TZ + 2 * D : 1;
        const int mx = (rx + 1) / 2, nq = mx * ry * rz;
        const int ox = H - D, oy = H - D, oz = DIM == 3 ? H - D : 0;
        for (int q = t; q < nq; q += FUSED_THREADS) {
            const int m = q % mx, r = q / mx, ly = oy + r % ry, lz = oz + r / ry;
            const int par = (color - i0 - ox - j0 - ly - (DIM == 3 ? k0 + lz : 0)) & 1;
            const int lx = ox + par + 2 * m;
            const int i = i0 + lx, j = j0 + ly, k = k0 + lz;
            if (lx >= ox + rx || i < 0 || i >= L.nx || j < 0 || j >= L.ny || (DIM == 3 && (k < 0 || k >= L.nz)))
                continue;
            const int c = lx + ly * SY + lz * SZ;
            const float ep = se[c];
            const float ce_ = sbx[c] * hx, cw = sbx[c - 1] * hx, cn = sby[c] * hy, cs = sby[c - SY] * hy;
            float sc = ce_ + cw + cn + cs;
            float sce = ce_ * (se[c + 1] - ep) + cw * (se[c - 1] - ep) + cn * (se[c + SY] - ep) + cs * (se[c - SY] - ep);
            if (DIM == 3) {
                const float ct = sbz[c] * hz, cb = sbz[c - SZ] * hz;
                sc += ct + cb;
                sce += ct * (se[c + SZ] - ep) + cb * (se[c - SZ] - ep);
            }
            se[c] = sc > 0.f ? ep + gs_step<float>(sce - sf[c], sc) : ep;
        }
        __syncthreads();
    }
    for (int q = t; q < TX * TY * TZ; q += FUSED_THREADS) {
        const int lx = H + q % TX, ly = H + (q / TX) % TY, lz = (DIM == 3 ? H : 0) + q / (TX * TY);
        const int i = i0 + lx, j = j0 + ly, k = k0 + lz;
        if (i < L.nx && j < L.ny && (DIM == 2 || k < L.nz)) eo[L.id(i, j, k)] = se[lx + ly * SY + lz * SZ];
    }
}

// residual of the level, restricted to the coarse right-hand side by the
// mean over the children; the coarse error is zeroed in the same pass.
// C2 (1 or 2): the factors are 2 in x and y and C2 in z (1 along the cases
// of a 2D stack), known at compile time, so the children's loads are issued
// together rather than one child per memory round trip; same summation order
// as the runtime form (C2 = 0).
template <typename T, int DIM, bool FINE, int C2 = 0, typename CT = T>
__global__ void __launch_bounds__(256) k_resid_restrict(const T *__restrict__ e, const T *__restrict__ f,
                                                        const T *__restrict__ rho, const CT *__restrict__ bx,
                                                        const CT *__restrict__ by, const CT *__restrict__ bz, Lay L,
                                                        Inv2 ih, T *__restrict__ fc, T *__restrict__ ec, Lay C,
                                                        int cx, int cy, int cz) {
    if (L.skipped()) return;
    int bx_, by_, bz_;
    xcd_block(bx_, by_, bz_);
    const int I = bx_ * blockDim.x + threadIdx.x, J = by_ * blockDim.y + threadIdx.y, K = bz_;
    if (I >= C.nx || J >= C.ny) return;
    if (C2) { cx = 2; cy = 2; cz = C2; }
    T acc = 0;
    int nch = 0;
#pragma unroll
    for (int kk = 0; kk < (C2 ? C2 : cz); ++kk)
#pragma unroll
        for (int jj = 0; jj < (C2 ? 2 : cy); ++jj)
#pragma unroll
            for (int ii = 0; ii < (C2 ? 2 : cx); ++ii) {
                int i = I * cx + ii, j = J * cy + jj, k = K * cz + kk;
                // the last aggregate of an odd axis has one child
                if (i >= L.nx || j >= L.ny || (DIM == 3 && k >= L.nz)) continue;
                long id = L.id(i, j, k);
                acc += f[id] - apply_A<T, DIM, FINE, CT>(e, rho, bx, by, bz, L, ih, i, j, k, id);
                ++nch;
            }
    const long cid = C.id(I, J, K);
    fc[cid] = acc / (T)nch;
    ec[cid] = (T)0;
}

// e_fine += e_coarse(parent): piecewise-constant prolongation.  (Trilinear
// interpolation between coarse centres converges worse here: at a 1000:1
// density jump the error is not smooth across the interface, and the
// flexible-CG iteration count from p = 0 to 1e-10 rose from 13 to 21 with it
// at 256^3.)
template <typename T>
__global__ void __launch_bounds__(256) k_prolong(T *__restrict__ e, const T *__restrict__ ec, Lay L, Lay C, int cx,
                                                 int cy, int cz) {
    if (L.skipped()) return;
    int bx_, by_, bz_;
    xcd_block(bx_, by_, bz_);
    const int i = bx_ * blockDim.x + threadIdx.x, j = by_ * blockDim.y + threadIdx.y, k = bz_;
    if (i >= L.nx || j >= L.ny) return;
    e[L.id(i, j, k)] += ec[C.id(i / cx, j / cy, k / cz)];
}

// coarse face coefficients: mean of the fine faces forming each coarse face
template <int DIM, bool FINE>
__global__ void k_coef_restrict(const double *__restrict__ rho, const double *__restrict__ bxf,
                                const double *__restrict__ byf, const double *__restrict__ bzf, Lay L,
                                double *__restrict__ bxc, double *__restrict__ byc, double *__restrict__ bzc, Lay C,
                                int cx, int cy, int cz) {
    const int I = blockIdx.x * blockDim.x + threadIdx.x, J = blockIdx.y * blockDim.y + threadIdx.y, K = blockIdx.z;
    if (I >= C.nx || J >= C.ny) return;
    const long cid = C.id(I, J, K);
    // the last aggregate of an odd axis has one child: fine indices past the
    // level are not children, and each mean runs over the faces that exist
    // (the +x face of the last coarse cell is the wall either way)
    const int nz_f = DIM == 3 ? L.nz : 1 << 30;
    // x faces at fine i = I*cx + cx - 1
    {
        double s = 0.0;
        int m = 0;
        int i = I * cx + cx - 1;
        for (int kk = 0; kk < cz; ++kk)
            for (int jj = 0; jj < cy; ++jj) {
                int j = J * cy + jj, k = K * cz + kk;
                if (j >= L.ny || k >= nz_f) continue;
                long id = L.id(i, j, k);
                double b;
                if (FINE) b = i < L.nx - 1 ? 2.0 / (rho[id] + rho[id + 1]) : 0.0;
                else b = bxf[id];
                s += b;
                ++m;
            }
        bxc[cid] = I < C.nx - 1 ? s / m : 0.0;
    }
    {
        double s = 0.0;
        int m = 0;
        int j = J * cy + cy - 1;
        for (int kk = 0; kk < cz; ++kk)
            for (int ii = 0; ii < cx; ++ii) {
                int i = I * cx + ii, k = K * cz + kk;
                if (i >= L.nx || k >= nz_f) continue;
                long id = L.id(i, j, k);
                double b;
                if (FINE) b = j < L.ny - 1 ? 2.0 / (rho[id] + rho[id + L.sx]) : 0.0;
                else b = byf[id];
                s += b;
                ++m;
            }
        byc[cid] = J < C.ny - 1 ? s / m : 0.0;
    }
    if (DIM == 3) {
        double s = 0.0;
        int m = 0;
        int k = K * cz + cz - 1;
        for (int jj = 0; jj < cy; ++jj)
            for (int ii = 0; ii < cx; ++ii) {
                int i = I * cx + ii, j = J * cy + jj;
                if (i >= L.nx || j >= L.ny) continue;
                long id = L.id(i, j, k);
                double b;
                if (FINE) b = k < L.nz - 1 ? 2.0 / (rho[id] + rho[id + L.sxy]) : 0.0;
                else b = bzf[id];
                s += b;
                ++m;
            }
        bzc[cid] = K < C.nz - 1 ? s / m : 0.0;
    }
}

__global__ void k_to_f32(const double *__restrict__ a, float *__restrict__ b, long n) {
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) b[t] = (float)a[t];
}

__global__ void k_to_f16(const double *__restrict__ a, _Float16 *__restrict__ b, long n) {
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) b[t] = (_Float16)a[t];
}

// Fixed-order two-stage sum of the block partials of a level-0 launch, per
// case q (blockIdx.z; a batched stack's case owns the contiguous logical
// blocks [q nb, (q+1) nb) of every slot): stage 1, block (b, s, q) sums the
// contiguous chunk b of slot s into partial2[(q ns + s) 256 + b]; stage 2
// (one block per case) sums the 256 chunk sums per slot.
__global__ void __launch_bounds__(256) k_sum_stage1(const double *__restrict__ partial, long nb, long chunk,
                                                    double *__restrict__ partial2, const int *__restrict__ skip) {
    if (skip && *skip) return;
    __shared__ double sh[4];
    const int s = blockIdx.y, q = blockIdx.z;
    const long ntot = nb * gridDim.z;
    const long b0 = (long)blockIdx.x * chunk, b1 = b0 + chunk < nb ? b0 + chunk : nb;
    double a = 0.0;
    for (long b = b0 + threadIdx.x; b < b1; b += 256) a += partial[s * ntot + q * nb + b];
    a = block_sum256(a, sh);
    if (threadIdx.x == 0) partial2[((long)q * gridDim.y + s) * 256 + blockIdx.x] = a;
}

// Device scalars of a solve (mg_scalars[0..63]).  The flexible-CG
// coefficients are formed on the device so an iteration needs one host
// read (the residual norm and the breakdown flag) and replays as one graph.
// A batched stack holds one block of SCS scalars per case; the dense coarse
// inverses follow the nb blocks.  SC_DONE freezes a case whose solve has
// ended while others of the stack go on: its alpha and beta are 0 from then
// on, so the x += alpha p still pending from its last iteration is the only
// update it receives, as at the end of a single solve.
// The stopping state of each case's solve follows (k_cg_start, k_cg_check):
// ||f||, the relative residual, its best value, iterations, iterations
// since the best, a non-finite flag, the starting relative residual (the
// host reads SC_DONE .. SC_R0), and the solve's tolerance and iteration
// limit.
enum {
    SC_RZ = 0, SC_RZO = 1, SC_PQ = 2, SC_RR = 3, SC_MEAN = 4, SC_RZ_OLD = 5, SC_BETA = 6, SC_ALPHA = 7,
    SC_FIRST = 8, SC_BREAK = 9, SC_DONE = 10, SC_FN = 11, SC_R = 12, SC_BEST = 13, SC_K = 14, SC_SINCE = 15,
    SC_NAN = 16, SC_R0 = 17, SC_NST = SC_R0 - SC_DONE + 1, SC_TOL = 18, SC_MAXC = 19
};
constexpr int SCS = 64;
enum { SUM_STORE = 0, SUM_CG_BETA = 1, SUM_CG_ALPHA = 2 };

// op SUM_CG_BETA: out = {r.z, r.zold} -> beta = (r.z - r.zold) / rz_old
// (0 on the first iteration), rz_old = r.z (Polak-Ribiere);
// op SUM_CG_ALPHA: out = {p.Ap} -> alpha = rz_old / p.Ap.  A is negative
// definite on the zero-mean space and the cycle approximates its inverse, so
// r.z < 0 and p.Ap < 0 until round-off; otherwise alpha = 0 and the
// breakdown flag stops the iteration.
__global__ void __launch_bounds__(256) k_sum_stage2(const double *__restrict__ partial2, int ns,
                                                    double *__restrict__ out, int op, double *__restrict__ sc,
                                                    const int *__restrict__ skip) {
    if (skip && *skip) return;
    __shared__ double sh[4];
    const int q = blockIdx.x;   // case
    partial2 += (long)q * ns * 256;
    out += (long)q * SCS;
    sc += (long)q * SCS;
    for (int s = 0; s < ns; ++s) {
        double a = block_sum256(partial2[s * 256 + threadIdx.x], sh);
        if (threadIdx.x == 0) out[s] = a;
    }
    const bool done = sc[SC_DONE] != 0.0;
    if (threadIdx.x == 0 && op == SUM_CG_BETA) {
        const double rz = out[0], rzo = out[1];
        sc[SC_BETA] = (sc[SC_FIRST] != 0.0 || done) ? 0.0 : (rz - rzo) / sc[SC_RZ_OLD];
        sc[SC_RZ_OLD] = rz;
        sc[SC_FIRST] = 0.0;
    } else if (threadIdx.x == 0 && op == SUM_CG_ALPHA) {
        const double pq = out[0], rz = sc[SC_RZ_OLD];
        const bool ok = pq < 0.0 && rz < 0.0 && !done;
        sc[SC_ALPHA] = ok ? rz / pq : 0.0;
        if (!ok && !done) sc[SC_BREAK] = 1.0;
    }
}

// The same sums in one launch: a block of SUM1_T threads per case q sums
// each slot's nbq partials (four independent chains per thread, so their
// loads overlap; then the waves in order), then the scalar step op as in
// k_sum_stage2.  Two dependent launches of ~4.4 us become one.
constexpr int SUM1_T = 1024;
__global__ void __launch_bounds__(SUM1_T) k_sum_one(const double *__restrict__ partial, long nbq, int ns,
                                                   double *__restrict__ out, int op, double *__restrict__ sc,
                                                   const int *__restrict__ skip) {
    if (skip && *skip) return;
    __shared__ double sh[SUM1_T / WAVE];
    const int q = blockIdx.x;   // case
    const long ntot = nbq * gridDim.x;
    out += (long)q * SCS;
    sc += (long)q * SCS;
    for (int s = 0; s < ns; ++s) {
        const double *pp = partial + s * ntot + q * nbq;
        double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
        long b = threadIdx.x;
        for (; b + 3 * SUM1_T < nbq; b += 4 * SUM1_T) {
            a0 += pp[b];
            a1 += pp[b + SUM1_T];
            a2 += pp[b + 2 * SUM1_T];
            a3 += pp[b + 3 * SUM1_T];
        }
        for (; b < nbq; b += SUM1_T) a0 += pp[b];
        double a = (a0 + a1) + (a2 + a3);
        for (int o = WAVE / 2; o > 0; o >>= 1) a += __shfl_xor(a, o, WAVE);
        if ((threadIdx.x & (WAVE - 1)) == 0) sh[threadIdx.x / WAVE] = a;
        __syncthreads();
        if (threadIdx.x == 0) {
            double t = 0.0;
            for (int w = 0; w < SUM1_T / WAVE; ++w) t += sh[w];
            out[s] = t;
        }
        __syncthreads();
    }
    const bool done = sc[SC_DONE] != 0.0;
    if (threadIdx.x == 0 && op == SUM_CG_BETA) {
        const double rz = out[0], rzo = out[1];
        sc[SC_BETA] = (sc[SC_FIRST] != 0.0 || done) ? 0.0 : (rz - rzo) / sc[SC_RZ_OLD];
        sc[SC_RZ_OLD] = rz;
        sc[SC_FIRST] = 0.0;
    } else if (threadIdx.x == 0 && op == SUM_CG_ALPHA) {
        const double pq = out[0], rz = sc[SC_RZ_OLD];
        const bool ok = pq < 0.0 && rz < 0.0 && !done;
        sc[SC_ALPHA] = ok ? rz / pq : 0.0;
        if (!ok && !done) sc[SC_BREAK] = 1.0;
    }
}

// start of a flexible-CG solve: first iteration, no breakdown (case blockIdx.x)
__global__ void k_cg_init(double *sc) {
    sc += (long)blockIdx.x * SCS;
    sc[SC_FIRST] = 1.0;
    sc[SC_BREAK] = 0.0;
    sc[SC_RZ_OLD] = 0.0;
    sc[SC_BETA] = 0.0;
    sc[SC_ALPHA] = 0.0;
    sc[SC_DONE] = 0.0;
}

// freeze case q of a batched solve (see SC_DONE)
__global__ void k_cg_done(double *sc, int q) { sc[(long)q * SCS + SC_DONE] = 1.0; }

// *skip = every case's solve has ended (one block of CG_THREADS threads)
constexpr int CG_THREADS = 256;
__device__ __forceinline__ void cg_all_done(const double *sc, int nq, int *skip) {
    int done = 1;
    for (int q = threadIdx.x; q < nq; q += CG_THREADS) done &= sc[(long)q * SCS + SC_DONE] != 0.0;
    const int all = __syncthreads_and(done);
    if (threadIdx.x == 0) *skip = all ? 1 : 0;
}

// Stopping state of the flexible-CG solve of every case from the initial
// sums ||r0||^2 (slot SC_RZ) and ||f||^2 (slot SC_RZO): the relative
// residual r = ||r0|| / ||f|| (||f|| = 1 for f = 0); a case ends at once
// when r <= tol.  Thread t: cases t, t + CG_THREADS, ...; the skip flag is
// set when all have ended.
__global__ void __launch_bounds__(CG_THREADS) k_cg_start(double *sc, int nq, double tol, int maxc, int *skip) {
    for (int q = threadIdx.x; q < nq; q += CG_THREADS) {
        double *s = sc + (long)q * SCS;
        const double fn = s[SC_RZO] > 0.0 ? sqrt(s[SC_RZO]) : 1.0, r = sqrt(s[SC_RZ]) / fn;
        s[SC_TOL] = tol;
        s[SC_MAXC] = (double)maxc;
        s[SC_FN] = fn;
        s[SC_R] = r;
        s[SC_R0] = r;
        s[SC_BEST] = r;
        s[SC_K] = 0.0;
        s[SC_SINCE] = 0.0;
        s[SC_NAN] = isfinite(r) ? 0.0 : 1.0;
        if (!(r > tol && 0 < maxc) || !isfinite(r)) s[SC_DONE] = 1.0;
    }
    cg_all_done(sc, nq, skip);
}

// After an iteration: a case that has not ended counts it, forms its
// relative residual sqrt(||r||^2) / ||f||, and ends on a breakdown, on
// r <= tol, after maxc iterations (both as k_cg_start stored them), or when r has not fallen below 0.9 of
// its best for 4 iterations (the FP64 round-off floor); an ended case is
// frozen (SC_DONE) for the iterations still queued.  A non-finite r ends
// the case and is flagged.
__global__ void __launch_bounds__(CG_THREADS) k_cg_check(double *sc, int nq, int *skip) {
    if (*skip) return;
    for (int q = threadIdx.x; q < nq; q += CG_THREADS) {
        double *s = sc + (long)q * SCS;
        const double tol = s[SC_TOL], maxc = s[SC_MAXC];
        if (s[SC_DONE] == 0.0) {
            if (s[SC_BREAK] != 0.0) {
                s[SC_DONE] = 1.0;
            } else {
                const double k = s[SC_K] + 1.0, r = sqrt(s[SC_RR]) / s[SC_FN];
                s[SC_K] = k;
                s[SC_R] = r;
                bool done = false;
                if (!isfinite(r)) {
                    s[SC_NAN] = 1.0;
                    done = true;
                } else {
                    if (r < 0.9 * s[SC_BEST]) {
                        s[SC_BEST] = r;
                        s[SC_SINCE] = 0.0;
                    } else {
                        s[SC_SINCE] += 1.0;
                        if (s[SC_SINCE] >= 4.0) done = true;
                    }
                    if (!(r > tol && k < maxc)) done = true;
                }
                if (done) s[SC_DONE] = 1.0;
            }
        }
    }
    cg_all_done(sc, nq, skip);
}

// sum of an interior field into block partials; in 3D a block sums SUM_KC
// planes (each thread its column's, loads issued together, the planes past
// the top clamped and not added), in 2D its slice (a case of a batched stack)
constexpr int SUM_KC = 8;
template <int DIM>
__global__ void __launch_bounds__(256) k_sum(const double *__restrict__ a, Lay L, double *__restrict__ partial) {
    __shared__ double sh[4];
    int bx_, by_, bz_;
    xcd_block(bx_, by_, bz_);
    const int i = bx_ * blockDim.x + threadIdx.x, j = by_ * blockDim.y + threadIdx.y;
    double v = 0.0;
    if (i < L.nx && j < L.ny) {
        if (DIM == 3) {
            const int k0 = bz_ * SUM_KC;
            double x[SUM_KC];
#pragma unroll
            for (int q = 0; q < SUM_KC; ++q) x[q] = a[L.id(i, j, min(k0 + q, L.nz - 1))];
#pragma unroll
            for (int q = 0; q < SUM_KC; ++q) v += k0 + q < L.nz ? x[q] : 0.0;
        } else {
            v = a[L.id(i, j, bz_)];
        }
    }
    v = block_sum256(v, sh);
    if (threadIdx.x + threadIdx.y == 0) partial[bx_ + (long)gridDim.x * (by_ + (long)gridDim.y * bz_)] = v;
}

// a -= scale * s[0]  (interior; s[q SCS] for case q of a batched stack)
__global__ void k_sub_scalar(double *__restrict__ a, Lay L, const double *__restrict__ s, double scale, int dim) {
    int bx_, by_, bz_;
    xcd_block(bx_, by_, bz_);
    const int i = bx_ * blockDim.x + threadIdx.x, j = by_ * blockDim.y + threadIdx.y, k = bz_;
    if (i >= L.nx || j >= L.ny) return;
    a[L.id(i, j, k)] -= s[(dim == 2 ? (long)k : 0L) * SCS] * scale;
}

// copy f into the working rhs minus its mean (compatibility of the Neumann
// problem): out = f - s[0]*scale
__global__ void k_center_copy(const double *__restrict__ f, double *__restrict__ out, Lay L,
                              const double *__restrict__ s, double scale, int dim) {
    int bx_, by_, bz_;
    xcd_block(bx_, by_, bz_);
    const int i = bx_ * blockDim.x + threadIdx.x, j = by_ * blockDim.y + threadIdx.y, k = bz_;
    if (i >= L.nx || j >= L.ny) return;
    long id = L.id(i, j, k);
    out[id] = f[id] - s[(dim == 2 ? (long)k : 0L) * SCS] * scale;
}

// p += e (FP32 colour-split correction into the FP64 solution), except for
// the frozen cases of a batched stack (SC_DONE of case k in sc)
__global__ void k_add_f32(double *__restrict__ p, const float *__restrict__ e, Lay L, Split S,
                          const double *__restrict__ sc, int dim) {
    int bx_, by_, bz_;
    xcd_block(bx_, by_, bz_);
    const int i = bx_ * blockDim.x + threadIdx.x, j = by_ * blockDim.y + threadIdx.y, k = bz_;
    if (i >= L.nx || j >= L.ny) return;
    if (sc[(dim == 2 ? (long)k : 0L) * SCS + SC_DONE] != 0.0) return;
    p[L.id(i, j, k)] += (double)e[S.at(i, j, k)];
}

// Coarsest level: Jacobi-preconditioned CG in FP64 inside one workgroup.
// The rhs is projected to zero mean; the solution is returned with zero mean.
// A batched 2D stack runs one workgroup per case on its slice (L describes
// one slice; workspace 6 n per case).
template <typename T, int DIM>
__global__ void __launch_bounds__(1024) k_coarse_cg(T *__restrict__ eout, const T *__restrict__ fin,
                                                    const T *__restrict__ bxT, const T *__restrict__ byT,
                                                    const T *__restrict__ bzT, const double *__restrict__ bx,
                                                    const double *__restrict__ by, const double *__restrict__ bz,
                                                    double *__restrict__ ws, Lay L, Inv2 ih, int maxit, double rtol) {
    (void)bxT; (void)byT; (void)bzT;
    if (L.skipped()) return;
    __shared__ double red[1024];
    const int nt = blockDim.x, t = threadIdx.x;
    const long n = (long)L.nx * L.ny * L.nz;
    const int kq = DIM == 2 ? (int)blockIdx.x : 0;
    ws += (long)kq * 6 * n;
    double *x = ws, *r = ws + n, *z = ws + 2 * n, *q = ws + 3 * n, *Aq = ws + 4 * n, *dg = ws + 5 * n;
    auto block_sum = [&](double v) -> double {
        red[t] = v;
        __syncthreads();
        for (int w = nt / 2; w > 0; w >>= 1) {
            if (t < w) red[t] += red[t + w];
            __syncthreads();
        }
        double s = red[0];
        __syncthreads();
        return s;
    };
    auto coords = [&](long c, int &i, int &j, int &k) {
        i = (int)(c % L.nx);
        long rr = c / L.nx;
        j = (int)(rr % L.ny);
        k = (int)(rr / L.ny) + kq;
    };
    // Apply -A (positive semi-definite) to v at cell c
    auto negA = [&](const double *v, long c) -> double {
        int i, j, k;
        coords(c, i, j, k);
        long id = L.id(i, j, k);
        double ce_ = bx[id] * ih.x, cw = bx[id - 1] * ih.x, cn = by[id] * ih.y, cs = by[id - L.sx] * ih.y;
        double sc = ce_ + cw + cn + cs;
        double acc = sc * v[c];
        if (i < L.nx - 1) acc -= ce_ * v[c + 1];
        if (i > 0) acc -= cw * v[c - 1];
        if (j < L.ny - 1) acc -= cn * v[c + L.nx];
        if (j > 0) acc -= cs * v[c - L.nx];
        if (DIM == 3) {
            double ct = bz[id] * ih.z, cb = bz[id - L.sxy] * ih.z;
            long pz = (long)L.nx * L.ny;
            if (k < L.nz - 1) acc -= ct * v[c + pz];
            if (k > 0) acc -= cb * v[c - pz];
            sc += ct + cb;
            acc += (ct + cb) * v[c];
        }
        return acc;
    };
    // rhs mean
    double s = 0.0;
    for (long c = t; c < n; c += nt) {
        int i, j, k;
        coords(c, i, j, k);
        s += (double)fin[L.id(i, j, k)];
    }
    const double mean = block_sum(s) / (double)n;
    double bb = 0.0;
    for (long c = t; c < n; c += nt) {
        int i, j, k;
        coords(c, i, j, k);
        long id = L.id(i, j, k);
        double b = -((double)fin[id] - mean);   // solve -A x = -f
        x[c] = 0.0;
        r[c] = b;
        double d = bx[id] * ih.x + bx[id - 1] * ih.x + by[id] * ih.y + by[id - L.sx] * ih.y;
        if (DIM == 3) d += bz[id] * ih.z + bz[id - L.sxy] * ih.z;
        dg[c] = d > 0.0 ? 1.0 / d : 0.0;
        bb += b * b;
    }
    bb = block_sum(bb);
    double rz = 0.0;
    for (long c = t; c < n; c += nt) {
        z[c] = r[c] * dg[c];
        q[c] = z[c];
        rz += r[c] * z[c];
    }
    rz = block_sum(rz);
    const double stop = rtol * rtol * bb;
    double rr = bb;
    for (int it = 0; it < maxit && rr > stop && bb > 0.0; ++it) {
        double qAq = 0.0;
        for (long c = t; c < n; c += nt) {
            Aq[c] = negA(q, c);
            qAq += q[c] * Aq[c];
        }
        qAq = block_sum(qAq);
        if (!(qAq > 0.0)) break;
        const double alpha = rz / qAq;
        double rn = 0.0, rz2 = 0.0;
        for (long c = t; c < n; c += nt) {
            x[c] += alpha * q[c];
            r[c] -= alpha * Aq[c];
            z[c] = r[c] * dg[c];
            rn += r[c] * r[c];
            rz2 += r[c] * z[c];
        }
        rn = block_sum(rn);
        rz2 = block_sum(rz2);
        const double beta = rz2 / rz;
        rz = rz2;
        rr = rn;
        for (long c = t; c < n; c += nt) q[c] = z[c] + beta * q[c];
        __syncthreads();
    }
    double xm = 0.0;
    for (long c = t; c < n; c += nt) xm += x[c];
    xm = block_sum(xm) / (double)n;
    for (long c = t; c < n; c += nt) {
        int i, j, k;
        coords(c, i, j, k);
        eout[L.id(i, j, k)] = (T)(x[c] - xm);
    }
}

// Coarsest levels of at most CDENSE cells are solved exactly with a dense
// inverse built once per coefficient setup.  The matrix is B = -A (symmetric,
// positive semi-definite, null space = constants) shifted to
// M = B + sigma 1 1^T, which is positive definite; for a zero-mean rhs b the
// solution of M x = b has zero mean and solves B x = b.  sigma = mean(diag B)/n
// keeps the shift at the scale of the operator.
constexpr int CDENSE = 128;

__device__ __forceinline__ void cell_coords(const Lay &L, int c, int &i, int &j, int &k) {
    i = c % L.nx;
    const int r = c / L.nx;
    j = r % L.ny;
    k = r / L.ny;
}

// interior (i, j, k) of a ghosted index of a small one-ghost level layout
// (32-bit arithmetic: a 64-bit division is a long software sequence)
__device__ __forceinline__ void cell_coords_g(const Lay &L, int id, int &i, int &j, int &k) {
    const int sxy = (int)L.sxy, sx = (int)L.sx;
    const int kq = id / sxy, r = id - kq * sxy, jq = r / sx;
    k = kq - L.gz;
    j = jq - 1;
    i = r - jq * sx - 1;
}

// Gauss-Jordan inversion without pivoting (M is SPD) by one block of
// CINV_THREADS threads, the matrix held in registers: thread t owns column
// t % 128 of the rows t / 128 + NRG q.  Per pivot only the pivot row and
// column go through LDS; the update of the owned elements is register
// arithmetic.  Writes M^-1 (symmetric) to minv[n*n].
constexpr int CINV_THREADS = 1024;
template <int DIM>
__global__ void __launch_bounds__(CINV_THREADS) k_coarse_inv(const double *__restrict__ bx,
                                                             const double *__restrict__ by,
                                                             const double *__restrict__ bz, Lay L, Inv2 ih,
                                                             double *__restrict__ minv) {
    constexpr int NRG = CINV_THREADS / CDENSE;           // row groups
    constexpr int NQ = CDENSE * CDENSE / CINV_THREADS;   // owned rows per thread
    constexpr int NW = CINV_THREADS / WAVE;
    __shared__ double colg[NRG][NQ], row[CDENSE];
    __shared__ double sh[NW];
    const int t = threadIdx.x, n = L.nx * L.ny * L.nz;
    const int cc = t % CDENSE, rg = t / CDENSE;
    // a batched 2D stack inverts case blockIdx.x's block (L: one slice)
    const int kq = DIM == 2 ? (int)blockIdx.x : 0;
    minv += (long)kq * n * n;
    double dsum = 0.0;
    if (t < n) {
        int i, j, k;
        cell_coords(L, t, i, j, k);
        k += kq;
        const long id = L.id(i, j, k);
        double d = (bx[id] + bx[id - 1]) * ih.x + (by[id] + by[id - L.sx]) * ih.y;
        if (DIM == 3) d += (bz[id] + bz[id - L.sxy]) * ih.z;
        dsum = d;
    }
    // fixed-order block sum: wave butterflies, then the waves in order
    for (int o = 32; o > 0; o >>= 1) dsum += __shfl_xor(dsum, o, WAVE);
    if ((t & (WAVE - 1)) == 0) sh[t / WAVE] = dsum;
    __syncthreads();
    double tot = 0.0;
    for (int w = 0; w < NW; ++w) tot += sh[w];
    const double sigma = tot / ((double)n * (double)n);
    // M = B + sigma 1 1^T.  B is symmetric (the face coefficient between two
    // cells is one stored value), so column cc is formed from cell cc's own
    // faces: the diagonal d(cc) and -c(face) for the neighbour rows.
    double m[NQ];
    {
        int i = 0, j = 0, k = 0;
        double ce_ = 0.0, cw = 0.0, cn = 0.0, cs = 0.0, ct = 0.0, cb = 0.0, d = 0.0;
        if (cc < n) {
            cell_coords(L, cc, i, j, k);
            k += kq;
            const long id = L.id(i, j, k);
            ce_ = bx[id] * ih.x; cw = bx[id - 1] * ih.x; cn = by[id] * ih.y; cs = by[id - L.sx] * ih.y;
            d = ce_ + cw + cn + cs;
            if (DIM == 3) {
                ct = bz[id] * ih.z;
                cb = bz[id - L.sxy] * ih.z;
                d += ct + cb;
            }
        }
        const int pz = L.nx * L.ny;
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            const int r = rg + NRG * q;
            double v = sigma;
            if (r == cc) v += d;
            else if (i < L.nx - 1 && r == cc + 1) v -= ce_;
            else if (i > 0 && r == cc - 1) v -= cw;
            else if (j < L.ny - 1 && r == cc + L.nx) v -= cn;
            else if (j > 0 && r == cc - L.nx) v -= cs;
            else if (DIM == 3 && k < L.nz - 1 && r == cc + pz) v -= ct;
            else if (DIM == 3 && k > 0 && r == cc - pz) v -= cb;
            m[q] = v;
        }
    }
    for (int kk = 0; kk < n; ++kk) {
        // pivot column (owned by the threads of column kk), stored by row
        // group so that each thread reads its rows as consecutive words
        if (cc == kk) {
#pragma unroll
            for (int q = 0; q < NQ; ++q) colg[rg][q] = m[q];
        }
        if (rg == kk % NRG && cc < n) {
            double v = 0.0;
#pragma unroll
            for (int q = 0; q < NQ; ++q) v = q == kk / NRG ? m[q] : v;
            row[cc] = v;
        }
        __syncthreads();
        const double piv = colg[kk % NRG][kk / NRG];
        const double nrow = (cc == kk ? 1.0 : row[cc]) / piv;   // new pivot row, this column
        double cv[NQ];
#pragma unroll
        for (int q = 0; q < NQ; ++q) cv[q] = colg[rg][q];
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            const int r = rg + NRG * q;
            m[q] = r == kk ? nrow : (cc == kk ? 0.0 : m[q]) - cv[q] * nrow;
        }
        __syncthreads();
    }
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        const int r = rg + NRG * q;
        if (r < n && cc < n) minv[r * n + cc] = m[q];
    }
}

// e = M^-1 b with b = -(f - mean f), then the zero-mean projection of e;
// one block of CDENSE threads, one cell per thread
template <typename T>
__global__ void __launch_bounds__(CDENSE) k_coarse_dense(T *__restrict__ eout, const T *__restrict__ fin,
                                                         const double *__restrict__ minv, Lay L, int dim) {
    if (L.skipped()) return;
    constexpr int NW = CDENSE / WAVE;
    __shared__ double b[CDENSE];
    __shared__ double ws[NW];
    const int t = threadIdx.x, n = L.nx * L.ny * L.nz;
    const int kq = dim == 2 ? (int)blockIdx.x : 0;   // case of a batched 2D stack (L: one slice)
    minv += (long)kq * n * n;
    int i = 0, j = 0, k = kq;
    double f = 0.0;
    if (t < n) {
        cell_coords(L, t, i, j, k);
        k += kq;
        f = (double)fin[L.id(i, j, k)];
    }
    // fixed-order block sum: butterfly per wave (lane 0's order), then waves in order
    auto block_sum = [&](double v) {
        for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, WAVE);
        if ((t & (WAVE - 1)) == 0) ws[t / WAVE] = v;
        __syncthreads();
        double sum = 0.0;
        for (int w = 0; w < NW; ++w) sum += ws[w];
        __syncthreads();
        return sum;
    };
    const double fm = block_sum(f) / (double)n;
    if (t < n) b[t] = fm - f;
    __syncthreads();
    double x = 0.0;
    if (t < n)
        for (int c = 0; c < n; ++c) x += minv[c * n + t] * b[c];
    const double xm = block_sum(x) / (double)n;
    if (t < n) eout[L.id(i, j, k)] = (T)(x - xm);
}

// Small levels in one workgroup: every level of at most TAIL_MAX cells, down
// to the dense coarsest solve and back up, runs inside a single launch with
// barriers between the phases.  On MI355X a dependent launch costs ~4 us
// however small the grid, and these levels take up to 12 sweeps each; a phase
// with one cell per thread costs ~1-2 us.  (With 8 cells per thread read from
// global memory, an 8192-cell level inside the workgroup was slower than
// separate launches: the sweeps wait on dependent L2 loads.)
//
// FP32 cycles may in addition take one fully coarsened "top" level of up to
// TAIL_TOP cells (8192 = 8 per thread) whose error (TAIL_LDS floats with
// padding) and rhs live in LDS and whose six face coefficients per cell live
// in the registers of the thread that owns the cell; a thread owns whole
// coarse cells with all their children, so the restriction needs no
// exchange.  FP64 would need twice the registers (spills) and keeps TAIL_MAX.
//
// The levels of at most TAIL_MAX cells (TAIL_CPT cells per thread) keep their
// error in LDS for the whole cycle (ghosted boxes of TAIL_ELDS elements in
// all, ghosts 0): the owner thread holds the cell's six face coefficients
// and rhs in registers, loaded from global memory once per down and once per
// up phase, so a half-sweep waits on LDS instead of on dependent L2 loads.
// Each level's error is written back at the end of its up phase.
constexpr int TAIL_MAX = 2048, TAIL_THREADS = 1024, TAIL_LEVELS = 8, TAIL_CPT = TAIL_MAX / TAIL_THREADS;
constexpr int TAIL_TOP = 8192, TAIL_LDS = 12288, TAIL_SLOTS = TAIL_TOP / TAIL_THREADS;
constexpr int TAIL_ELDS = 4096;

template <typename T>
struct TailLevel {
    T *e;
    T *f;
    const T *bx, *by, *bz;
    Lay L;
    Inv2 ih;
    int cx, cy, cz;   // coarsening factors to the next level
    int nu1, nu2;
};

template <typename T>
struct TailArgs {
    TailLevel<T> lv[TAIL_LEVELS];
    int n;                 // levels in the tail; the last is solved densely
    int top;               // lv[0] is an LDS/register top level (FP32 only)
    const double *minv;    // dense inverse of the last level (k_coarse_inv)
    int lds;               // levels from lv[top] on keep their error in LDS
    int eoff[TAIL_LEVELS]; // offset of each such level's ghosted box
};

template <typename T, int DIM>
__global__ void __launch_bounds__(TAIL_THREADS) k_mg_tail(TailArgs<T> A) {
    if (A.lv[0].L.skipped()) return;
    __shared__ double red[TAIL_THREADS / WAVE];
    __shared__ double bvec[CDENSE];
    const int t = threadIdx.x;
    // a batched 2D stack runs one workgroup per case on its slice kq; the
    // level layouts describe one slice
    const int kq = DIM == 2 ? (int)blockIdx.x : 0;
    auto cell_coords = [kq](const Lay &L, int c, int &i, int &j, int &k) {
        fcfd::cell_coords(L, c, i, j, k);
        k += kq;
    };
    auto ncells = [](const Lay &L) { return L.nx * L.ny * L.nz; };
    auto gs = [&](const TailLevel<T> &V, int color) {
        const Lay &L = V.L;
        const int n = ncells(L);
        const Prol<T> P0{nullptr, Lay{}, 0, 0, 0};
        for (int c = t; c < n; c += TAIL_THREADS) {
            int i, j, k;
            cell_coords(L, c, i, j, k);
            if (((i + j + (DIM == 3 ? k : 0)) & 1) != color) continue;   // 2D: k is the case
            const long id = L.id(i, j, k);
            T sc, sce, ep;
            stencil<T, DIM, false, 0, T>(V.e, nullptr, V.bx, V.by, V.bz, L, V.ih, i, j, k, id, P0, sc, sce, ep);
            V.e[id] = sc > (T)0 ? ep + gs_step<T>(sce - V.f[id], sc) : ep;
        }
        __syncthreads();
    };
    // fixed-order sum over the block: wave butterflies, then the waves in order
    auto block_sum = [&](double v) {
        for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, WAVE);
        if ((t & (WAVE - 1)) == 0) red[t / WAVE] = v;
        __syncthreads();
        double s = 0.0;
        for (int w = 0; w < TAIL_THREADS / WAVE; ++w) s += red[w];
        __syncthreads();
        return s;
    };
    const int nl = A.n;
    // ---- LDS-resident levels (A.lds): error in a ghosted box per level
    __shared__ T el_[TAIL_ELDS];
    __shared__ T rt_[TAIL_MAX];
    auto lbox = [](const Lay &L, int i, int j, int k) {   // k: 0 in 2D
        return (i + 1) + (L.nx + 2) * ((j + 1) + (DIM == 3 ? (L.ny + 2) * (k + 1) : 0));
    };
    auto boxn = [](const Lay &L) { return (L.nx + 2) * (L.ny + 2) * (DIM == 3 ? L.nz + 2 : 1); };
    // the owner's face coefficients (E W N S T B, as stencil() forms them) and
    // rhs of its cells c = t + u TAIL_THREADS
    auto load_regs = [&](const TailLevel<T> &V, T (&cf)[TAIL_CPT][6], T (&fv)[TAIL_CPT]) {
        const Lay &L = V.L;
#pragma unroll
        for (int u = 0; u < TAIL_CPT; ++u) {
            const int c = t + u * TAIL_THREADS;
            if (c < ncells(L)) {
                int i, j, k;
                cell_coords(L, c, i, j, k);
                const long id = L.id(i, j, k);
                cf[u][0] = (T)V.bx[id] * (T)V.ih.x;
                cf[u][1] = (T)V.bx[id - 1] * (T)V.ih.x;
                cf[u][2] = (T)V.by[id] * (T)V.ih.y;
                cf[u][3] = (T)V.by[id - L.sx] * (T)V.ih.y;
                cf[u][4] = DIM == 3 ? (T)V.bz[id] * (T)V.ih.z : (T)0;
                cf[u][5] = DIM == 3 ? (T)V.bz[id - L.sxy] * (T)V.ih.z : (T)0;
                fv[u] = V.f[id];
            }
        }
    };
    // one colour of Gauss-Seidel on an LDS-resident level, stencil()'s order
    auto gs_lds = [&](const TailLevel<T> &V, T *el, const T (&cf)[TAIL_CPT][6], const T (&fv)[TAIL_CPT],
                      int color) {
        const Lay &L = V.L;
#pragma unroll
        for (int u = 0; u < TAIL_CPT; ++u) {
            const int c = t + u * TAIL_THREADS;
            if (c >= ncells(L)) continue;
            int i, j, k;
            cell_coords(L, c, i, j, k);
            const int kl = DIM == 3 ? k : 0;
            if (((i + j + kl) & 1) != color) continue;
            const int li = lbox(L, i, j, kl), bsx = L.nx + 2, bsxy = bsx * (L.ny + 2);
            const T ep = el[li];
            const T vE = el[li + 1], vW = el[li - 1], vN = el[li + bsx], vS = el[li - bsx];
            T sc = cf[u][0] + cf[u][1] + cf[u][2] + cf[u][3];
            T sce = cf[u][0] * (vE - ep) + cf[u][1] * (vW - ep) + cf[u][2] * (vN - ep) + cf[u][3] * (vS - ep);
            if (DIM == 3) {
                sc += cf[u][4] + cf[u][5];
                sce += cf[u][4] * (el[li + bsxy] - ep) + cf[u][5] * (el[li - bsxy] - ep);
            }
            el[li] = sc > (T)0 ? ep + gs_step<T>(sce - fv[u], sc) : ep;
        }
        __syncthreads();
    };
    // ---- top level in LDS + registers (FP32 only; A.top is never set for FP64).
    // Fully coarsened (2x2x2, or 2x2 in 2D): thread t owns coarse cell(s)
    // cc = t + TAIL_THREADS * g and slot sl = g * NCH + ch holds child ch =
    // ii + 2 jj + 4 kk.  The error is stored child-major, es[ch * stride +
    // pad + cc], so a wave's accesses to one child are consecutive words
    // (no bank conflicts) and every neighbour sits at a uniform offset;
    // `pad` zero words (one coarse plane) before and after each child block
    // stand in for the missing neighbours below the first and above the last
    // coarse plane (their face coefficients are 0).
    constexpr int NCH = DIM == 3 ? 8 : 4, NG = TAIL_SLOTS / NCH;
    float *es = nullptr, *fs = nullptr;   // LDS error (child-major) and rhs (slot-major)
    float tc[TAIL_SLOTS][6];              // face coefficients E W N S T B of each slot
    int a0 = 0, cnx = 0, plane = 0, stride = 0, nct = 0;
    // index of slot sl's cell in es
    auto top_li = [&](int sl) { return (sl % NCH) * stride + plane + t + TAIL_THREADS * (sl / NCH); };
    // one colour of Gauss-Seidel on the top level; same operation order as stencil()
    auto top_gs = [&](int color) {
        // opaque copies of the offsets: keeps the compiler from holding six
        // neighbour addresses per slot in registers across the sweeps
        int st = stride, cx_ = cnx, pl = plane;
        asm volatile("" : "+s"(st), "+s"(cx_), "+s"(pl));
#pragma unroll
        for (int sl = 0; sl < TAIL_SLOTS; ++sl) {
            const int ch = sl % NCH, ii = ch & 1, jj = (ch >> 1) & 1, kk = ch >> 2;
            if (((ii + jj + kk) & 1) != color || t + TAIL_THREADS * (sl / NCH) >= nct) continue;
            const int li = (sl % NCH) * st + pl + t + TAIL_THREADS * (sl / NCH);
            const int oxp = ii ? 1 - st : st, oxm = ii ? -st : st - 1;
            const int oyp = jj ? cx_ - 2 * st : 2 * st, oym = jj ? -2 * st : 2 * st - cx_;
            const float ep = es[li];
            float sc = tc[sl][0] + tc[sl][1] + tc[sl][2] + tc[sl][3];
            float sce = tc[sl][0] * (es[li + oxp] - ep) + tc[sl][1] * (es[li + oxm] - ep) +
                        tc[sl][2] * (es[li + oyp] - ep) + tc[sl][3] * (es[li + oym] - ep);
            if (DIM == 3) {
                const int ozp = kk ? pl - 4 * st : 4 * st, ozm = kk ? -4 * st : 4 * st - pl;
                sc += tc[sl][4] + tc[sl][5];
                sce += tc[sl][4] * (es[li + ozp] - ep) + tc[sl][5] * (es[li + ozm] - ep);
            }
            es[li] = sc > 0.f ? ep + gs_step<float>(sce - fs[sl * TAIL_THREADS + t], sc) : ep;
        }
        __syncthreads();
    };
    // global (ghosted) index of slot sl's cell
    auto top_gid = [&](const TailLevel<T> &F, const TailLevel<T> &C, int sl) -> long {
        const int ch = sl % NCH;
        int I, J, K;
        cell_coords(C.L, t + TAIL_THREADS * (sl / NCH), I, J, K);
        return F.L.id(2 * I + (ch & 1), 2 * J + ((ch >> 1) & 1), DIM == 3 ? 2 * K + (ch >> 2) : K);
    };
    if constexpr (sizeof(T) == 4) {
        __shared__ float es_[TAIL_LDS], fs_[TAIL_TOP];
        es = es_;
        fs = fs_;
        if (A.top) {
            a0 = 1;
            const TailLevel<T> &F = A.lv[0];
            const TailLevel<T> &C = A.lv[1];
            const Lay &L = F.L;
            cnx = C.L.nx;
            plane = DIM == 3 ? C.L.nx * C.L.ny : C.L.nx;
            nct = ncells(C.L);
            stride = nct + 2 * plane;
            for (int q = t; q < NCH * stride; q += TAIL_THREADS) es[q] = 0.f;   // zero error and pads
#pragma unroll
            for (int sl = 0; sl < TAIL_SLOTS; ++sl) {
                if (t + TAIL_THREADS * (sl / NCH) < nct) {
                    const long id = top_gid(F, C, sl);
                    tc[sl][0] = F.bx[id] * (float)F.ih.x;
                    tc[sl][1] = F.bx[id - 1] * (float)F.ih.x;
                    tc[sl][2] = F.by[id] * (float)F.ih.y;
                    tc[sl][3] = F.by[id - L.sx] * (float)F.ih.y;
                    tc[sl][4] = DIM == 3 ? F.bz[id] * (float)F.ih.z : 0.f;
                    tc[sl][5] = DIM == 3 ? F.bz[id - L.sxy] * (float)F.ih.z : 0.f;
                    fs[sl * TAIL_THREADS + t] = F.f[id];
                } else {
                    for (int q = 0; q < 6; ++q) tc[sl][q] = 0.f;
                }
            }
            __syncthreads();
            for (int sw = 0; sw < F.nu1; ++sw) { top_gs(0); top_gs(1); }
            // restriction: a coarse cell's children are its NCH slots, summed in
            // k_resid_restrict's order (ii fastest); A e in apply_A's order
#pragma unroll
            for (int g = 0; g < NG; ++g) {
                const int cc = t + TAIL_THREADS * g;
                if (cc >= nct) continue;
                float acc = 0.f;
#pragma unroll
                for (int ch = 0; ch < NCH; ++ch) {
                    const int sl = g * NCH + ch, ii = ch & 1, jj = (ch >> 1) & 1, kk = ch >> 2;
                    const int li = top_li(sl);
                    const int oxp = ii ? 1 - stride : stride, oxm = ii ? -stride : stride - 1;
                    const int oyp = jj ? cnx - 2 * stride : 2 * stride, oym = jj ? -2 * stride : 2 * stride - cnx;
                    const float ep = es[li];
                    float ae = tc[sl][0] * (es[li + oxp] - ep) + tc[sl][1] * (es[li + oxm] - ep) +
                               tc[sl][2] * (es[li + oyp] - ep) + tc[sl][3] * (es[li + oym] - ep);
                    if (DIM == 3) {
                        const int ozp = kk ? plane - 4 * stride : 4 * stride, ozm = kk ? -4 * stride : 4 * stride - plane;
                        ae += tc[sl][4] * (es[li + ozp] - ep) + tc[sl][5] * (es[li + ozm] - ep);
                    }
                    acc += fs[sl * TAIL_THREADS + t] - ae;
                }
                int I, J, K;
                cell_coords(C.L, cc, I, J, K);
                const long cid = C.L.id(I, J, K);
                C.f[cid] = acc / (float)NCH;
                C.e[cid] = 0.f;
            }
            __syncthreads();
        }
    }
    // down: smooth, restrict the residual, clear the coarse error
    for (int a = a0; a < nl - 1 && A.lds; ++a) {
        const TailLevel<T> &F = A.lv[a];
        const TailLevel<T> &C = A.lv[a + 1];
        const Lay &L = F.L;
        T *el = el_ + A.eoff[a];
        T cf[TAIL_CPT][6], fv[TAIL_CPT];
        for (int q = t; q < boxn(L); q += TAIL_THREADS) el[q] = (T)0;
        __syncthreads();
        load_regs(F, cf, fv);
        const int n = ncells(L);
        if (a == a0)   // the first level's error as it was left (later ones start at 0)
            for (int c = t; c < n; c += TAIL_THREADS) {
                int i, j, k;
                cell_coords(L, c, i, j, k);
                el[lbox(L, i, j, DIM == 3 ? k : 0)] = F.e[L.id(i, j, k)];
            }
        __syncthreads();
        for (int s = 0; s < F.nu1; ++s) { gs_lds(F, el, cf, fv, 0); gs_lds(F, el, cf, fv, 1); }
        // residual of each cell by its owner in apply_A's order, then the
        // coarse sums over the children in k_resid_restrict's order
#pragma unroll
        for (int u = 0; u < TAIL_CPT; ++u) {
            const int c = t + u * TAIL_THREADS;
            if (c >= n) continue;
            int i, j, k;
            cell_coords(L, c, i, j, k);
            const int li = lbox(L, i, j, DIM == 3 ? k : 0), bsx = L.nx + 2, bsxy = bsx * (L.ny + 2);
            const T ep = el[li];
            T acc = cf[u][0] * (el[li + 1] - ep) + cf[u][1] * (el[li - 1] - ep) + cf[u][2] * (el[li + bsx] - ep) +
                    cf[u][3] * (el[li - bsx] - ep);
            if (DIM == 3) acc += cf[u][4] * (el[li + bsxy] - ep) + cf[u][5] * (el[li - bsxy] - ep);
            rt_[c] = fv[u] - acc;
        }
        __syncthreads();
        const int nc = ncells(C.L);
        for (int cc = t; cc < nc; cc += TAIL_THREADS) {
            int I, J, K;
            cell_coords(C.L, cc, I, J, K);
            const int KL = DIM == 3 ? K : 0;
            T acc = 0;
            int nch = 0;
            for (int kk = 0; kk < F.cz; ++kk)
                for (int jj = 0; jj < F.cy; ++jj)
                    for (int ii = 0; ii < F.cx; ++ii) {
                        const int i = I * F.cx + ii, j = J * F.cy + jj, k = KL * F.cz + kk;
                        if (i >= L.nx || j >= L.ny || (DIM == 3 && k >= L.nz)) continue;
                        acc += rt_[i + L.nx * (j + L.ny * k)];
                        ++nch;
                    }
            const long cid = C.L.id(I, J, K);
            C.f[cid] = acc / (T)nch;
            C.e[cid] = (T)0;
        }
        __syncthreads();
    }
    for (int a = a0; a < nl - 1 && !A.lds; ++a) {
        const TailLevel<T> &F = A.lv[a];
        const TailLevel<T> &C = A.lv[a + 1];
        for (int s = 0; s < F.nu1; ++s) { gs(F, 0); gs(F, 1); }
        const int nc = ncells(C.L);
        for (int cc = t; cc < nc; cc += TAIL_THREADS) {
            int I, J, K;
            cell_coords(C.L, cc, I, J, K);
            T acc = 0;
            int nch = 0;
            for (int kk = 0; kk < F.cz; ++kk)
                for (int jj = 0; jj < F.cy; ++jj)
                    for (int ii = 0; ii < F.cx; ++ii) {
                        const int i = I * F.cx + ii, j = J * F.cy + jj, k = K * F.cz + kk;
                        // the last aggregate of an odd axis has one child (2D: k is the case)
                        if (i >= F.L.nx || j >= F.L.ny || (DIM == 3 && k >= F.L.nz)) continue;
                        const long id = F.L.id(i, j, k);
                        acc += F.f[id] - apply_A<T, DIM, false, T>(F.e, nullptr, F.bx, F.by, F.bz, F.L, F.ih, i, j, k, id);
                        ++nch;
                    }
            const long cid = C.L.id(I, J, K);
            C.f[cid] = acc / (T)nch;
            C.e[cid] = (T)0;
        }
        __syncthreads();
    }
    // coarsest: e = M^-1 b, b = -(f - mean f), zero-mean result
    {
        const TailLevel<T> &V = A.lv[nl - 1];
        const int n = ncells(V.L);
        int i = 0, j = 0, k = kq;
        double fv = 0.0;
        if (t < n) {
            cell_coords(V.L, t, i, j, k);
            fv = (double)V.f[V.L.id(i, j, k)];
        }
        const double fm = block_sum(fv) / (double)n;
        if (t < n) bvec[t] = fm - fv;
        __syncthreads();
        double x = 0.0;
        const double *minv = A.minv + (long)kq * n * n;
        if (t < n)
            for (int cc = 0; cc < n; ++cc) x += minv[cc * n + t] * bvec[cc];
        const double xm = block_sum(t < n ? x : 0.0) / (double)n;
        if (t < n) V.e[V.L.id(i, j, k)] = (T)(x - xm);
        if (A.lds) {
            T *el = el_ + A.eoff[nl - 1];
            for (int q = t; q < boxn(V.L); q += TAIL_THREADS) el[q] = (T)0;
            __syncthreads();
            if (t < n) el[lbox(V.L, i, j, DIM == 3 ? k : 0)] = (T)(x - xm);
        }
        __syncthreads();
    }
    // up: prolong (piecewise constant), smooth; LDS levels write their error back
    for (int a = nl - 2; a >= a0 && A.lds; --a) {
        const TailLevel<T> &F = A.lv[a];
        const TailLevel<T> &C = A.lv[a + 1];
        const Lay &L = F.L;
        T *el = el_ + A.eoff[a];
        const T *elc = el_ + A.eoff[a + 1];
        T cf[TAIL_CPT][6], fv[TAIL_CPT];
        load_regs(F, cf, fv);
        const int n = ncells(L);
        for (int c = t; c < n; c += TAIL_THREADS) {
            int i, j, k;
            cell_coords(L, c, i, j, k);
            const int kl = DIM == 3 ? k : 0;
            el[lbox(L, i, j, kl)] += elc[lbox(C.L, i / F.cx, j / F.cy, kl / F.cz)];
        }
        __syncthreads();
        for (int s = 0; s < F.nu2; ++s) { gs_lds(F, el, cf, fv, 1); gs_lds(F, el, cf, fv, 0); }
        for (int c = t; c < n; c += TAIL_THREADS) {
            int i, j, k;
            cell_coords(L, c, i, j, k);
            F.e[L.id(i, j, k)] = el[lbox(L, i, j, DIM == 3 ? k : 0)];
        }
        __syncthreads();
    }
    for (int a = nl - 2; a >= a0 && !A.lds; --a) {
        const TailLevel<T> &F = A.lv[a];
        const TailLevel<T> &C = A.lv[a + 1];
        const int n = ncells(F.L);
        for (int c = t; c < n; c += TAIL_THREADS) {
            int i, j, k;
            cell_coords(F.L, c, i, j, k);
            F.e[F.L.id(i, j, k)] += C.e[C.L.id(i / F.cx, j / F.cy, k / F.cz)];
        }
        __syncthreads();
        for (int s = 0; s < F.nu2; ++s) { gs(F, 1); gs(F, 0); }
    }
    if constexpr (sizeof(T) == 4) {
        if (A.top) {
            const TailLevel<T> &F = A.lv[0];
            const TailLevel<T> &C = A.lv[1];
#pragma unroll
            for (int g = 0; g < NG; ++g) {
                const int cc = t + TAIL_THREADS * g;
                if (cc >= nct) continue;
                int I, J, K;
                cell_coords(C.L, cc, I, J, K);
                const float ec = C.e[C.L.id(I, J, K)];
#pragma unroll
                for (int ch = 0; ch < NCH; ++ch) es[top_li(g * NCH + ch)] += ec;
            }
            __syncthreads();
            for (int sw = 0; sw < F.nu2; ++sw) { top_gs(1); top_gs(0); }
#pragma unroll
            for (int sl = 0; sl < TAIL_SLOTS; ++sl)
                if (t + TAIL_THREADS * (sl / NCH) < nct) F.e[top_gid(F, C, sl)] = es[top_li(sl)];
        }
    }
}

// ------------------------------------------------------------ host side

// 256-thread block for rows of w threads: 64 x 4, or 32 x 8 where that
// leaves fewer lanes idle past the row end (140 cells: 160 instead of 192
// threads per row; the 70-pair colour rows: 96 instead of 128)
static dim3 mg_blk(int w) {
    const long p64 = (w + 63) / 64 * 64L, p32 = (w + 31) / 32 * 32L;
    return p32 < p64 ? dim3(32, 8) : dim3(64, 4);
}

static dim3 grd3(int nx, int ny, int nz, dim3 blk) {
    return dim3((nx + blk.x - 1) / blk.x, (ny + blk.y - 1) / blk.y, nz);
}

static Inv2 inv2(double hx, double hy, double hz) {
    Inv2 r;
    r.x = 1.0 / (hx * hx);
    r.y = 1.0 / (hy * hy);
    r.z = 1.0 / (hz * hz);
    return r;
}

int mg_free(fcfd_ctx *c) {
    for (auto &M : c->mg) {
        void *ptrs[] = {M.e64, M.f64, M.r64, M.bx64, M.by64, M.bz64, M.e32, M.f32, M.r32, M.bx32, M.by32, M.bz32,
                        M.e32b, M.bx16, M.by16, M.bz16};
        for (void *p : ptrs)
            if (p) hipFree(p);
    }
    c->mg.clear();
    if (c->rho32) hipFree(c->rho32);
    if (c->rho16) hipFree(c->rho16);
    if (c->mg_scalars) hipFree(c->mg_scalars);
    if (c->mg_skip) hipFree(c->mg_skip);
    c->mg_skip = nullptr;
    void *kv[] = {c->kr, c->kp, c->kq, c->kz, c->kzo, c->kzo32};
    for (void *p : kv)
        if (p) hipFree(p);
    c->kr = c->kp = c->kq = c->kz = c->kzo = nullptr;
    c->kzo32 = nullptr;
    c->rho32 = nullptr;
    c->rho16 = nullptr;
    c->mg_scalars = nullptr;
    c->mg_ready = 0;
    return mg_drop_graphs(c);
}

int mg_drop_graphs(fcfd_ctx *c) {
    for (int a = 0; a < 4; ++a)
        for (int b = 0; b < 5; ++b) {
            for (int s = 0; s < 2; ++s) {
                if (c->vgraph[s][a][b]) hipGraphExecDestroy(c->vgraph[s][a][b]);
                c->vgraph[s][a][b] = nullptr;
            }
            c->vgraph_prof[a][b].clear();
        }
    return FCFD_OK;
}

template <typename T>
static int alloc0(T **p, long n) {
    if (hipMalloc((void **)p, sizeof(T) * n) != hipSuccess) return FCFD_ERR_HIP;
    if (hipMemset(*p, 0, sizeof(T) * n) != hipSuccess) return FCFD_ERR_HIP;
    return FCFD_OK;
}

// Build the level hierarchy once per grid (sizes only).  Level 0 describes
// the field layout; its arrays are the context's fields (FP64) and, for the
// mixed mode, e32/f32/rho32 (and the FP32 z_old) in the colour-split layout.
static int mg_build(fcfd_ctx *c) {
    const Geom &G = c->G;
    mg_free(c);
    MgLevel L0 = {};
    L0.nx = G.nx; L0.ny = G.ny; L0.nz = G.nz;
    L0.hx = G.dx; L0.hy = G.dy; L0.hz = G.dim == 3 ? G.dz : 1.0;
    L0.sx = G.sx; L0.sxy = G.sxy; L0.ntot = G.ntot;
    c->mg.push_back(L0);
    for (;;) {
        MgLevel &F = c->mg.back();
        long ncell = case_cells(F, G.dim);
        if (ncell <= CDENSE && c->mg.size() > 1) break;
        double hmin = G.dim == 3 ? std::min({F.hx, F.hy, F.hz}) : std::min(F.hx, F.hy);
        int cf[3] = {1, 1, 1};
        const int n[3] = {F.nx, F.ny, F.nz};
        const double h[3] = {F.hx, F.hy, F.hz};
        bool any = false;
        // an axis of n >= 3 cells pairs its cells into (n + 1) / 2 aggregates;
        // on an odd axis the last aggregate is a single cell
        for (int d = 0; d < G.dim; ++d)
            if (n[d] >= 3 && h[d] <= 1.5 * hmin) { cf[d] = 2; any = true; }
        if (!any) {
            if (c->mg.size() == 1) {
                fprintf(stderr, "libfcfd: PPE multigrid needs at least 3 cells on some axis\n");
                return FCFD_ERR_ARG;
            }
            break;
        }
        F.cx = cf[0]; F.cy = cf[1]; F.cz = cf[2];
        MgLevel C = {};
        C.nx = (F.nx + cf[0] - 1) / cf[0]; C.ny = (F.ny + cf[1] - 1) / cf[1];
        C.nz = G.dim == 3 ? (F.nz + cf[2] - 1) / cf[2] : F.nz;   // 2D: the cases
        C.hx = F.hx * cf[0]; C.hy = F.hy * cf[1]; C.hz = F.hz * cf[2];
        C.sx = C.nx + 2;
        C.sxy = C.sx * (C.ny + 2);
        C.ntot = C.sxy * (C.nz + (G.dim == 3 ? 2 : 0));
        c->mg.push_back(C);
    }
    c->mg.back().cx = c->mg.back().cy = c->mg.back().cz = 1;
    for (size_t l = 1; l < c->mg.size(); ++l) {
        MgLevel &M = c->mg[l];
        long n = M.ntot;
        FCFD_TRY(alloc0(&M.e64, n)); FCFD_TRY(alloc0(&M.f64, n));
        FCFD_TRY(alloc0(&M.bx64, n)); FCFD_TRY(alloc0(&M.by64, n));
        if (G.dim == 3) FCFD_TRY(alloc0(&M.bz64, n));
        FCFD_TRY(alloc0(&M.e32, n)); FCFD_TRY(alloc0(&M.f32, n)); FCFD_TRY(alloc0(&M.e32b, n));
        FCFD_TRY(alloc0(&M.bx32, n)); FCFD_TRY(alloc0(&M.by32, n));
        if (G.dim == 3) FCFD_TRY(alloc0(&M.bz32, n));
        if (G.dim == 3 && l == 1) {   // FP16 coefficients of level 1 (ppe_rho_half)
            FCFD_TRY(alloc0(&M.bx16, n)); FCFD_TRY(alloc0(&M.by16, n)); FCFD_TRY(alloc0(&M.bz16, n));
        }
    }
    MgLevel &M0 = c->mg[0];
    const long nsplit = 2 * split_of(G).half;
    FCFD_TRY(alloc0(&M0.e32, nsplit));
    FCFD_TRY(alloc0(&M0.f32, nsplit));
    FCFD_TRY(alloc0(&c->rho32, nsplit));
    FCFD_TRY(alloc0(&c->rho16, nsplit));
    FCFD_TRY(alloc0(&M0.f64, G.ntot));   // zero-mean working rhs
    FCFD_TRY(alloc0(&c->kr, G.ntot));
    FCFD_TRY(alloc0(&c->kp, G.ntot));
    FCFD_TRY(alloc0(&c->kq, G.ntot));
    FCFD_TRY(alloc0(&c->kz, G.ntot));
    FCFD_TRY(alloc0(&c->kzo, G.ntot));
    FCFD_TRY(alloc0(&c->kzo32, nsplit));
    // per-case scalars + coarse CG workspace (6 vectors) or the dense coarse
    // inverse of every case
    long ncoarse = case_cells(c->mg.back(), G.dim);
    if (c->mg.size() < 2) {
        fprintf(stderr, "libfcfd: the PPE multigrid needs at least two levels\n");
        return FCFD_ERR_ARG;
    }
    FCFD_TRY(alloc0(&c->mg_scalars, SCS * c->nb + c->nb * std::max(6 * ncoarse, ncoarse * ncoarse)));
    FCFD_TRY(alloc0(&c->mg_skip, 1));
    for (auto &e : c->cg_ev)
        if (!e) FCFD_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    return FCFD_OK;
}

// dense coarse inverses (case q at + q n^2) after the per-case scalars
static double *coarse_inv(fcfd_ctx *c) { return c->mg_scalars + SCS * c->nb; }

// ppe_sor_init: coefficients of every level from rho%now
int mg_setup(fcfd_ctx *c) {
    const Geom &G = c->G;
    if (c->mg.empty()) FCFD_TRY(mg_build(c));
    dim3 blk(64, 4);
    for (size_t l = 1; l < c->mg.size(); ++l) {
        MgLevel &F = c->mg[l - 1], &C = c->mg[l];
        Lay LF = l == 1 ? lay_of(G) : lay_of(F, G.dim), LC = lay_of(C, G.dim);
        dim3 g = grd3(C.nx, C.ny, C.nz, blk);
        if (l == 1) {
            if (G.dim == 3)
                hipLaunchKernelGGL((k_coef_restrict<3, true>), g, blk, 0, c->stream, c->f[FCFD_RHO], nullptr, nullptr,
                                   nullptr, LF, C.bx64, C.by64, C.bz64, LC, F.cx, F.cy, F.cz);
            else
                hipLaunchKernelGGL((k_coef_restrict<2, true>), g, blk, 0, c->stream, c->f[FCFD_RHO], nullptr, nullptr,
                                   nullptr, LF, C.bx64, C.by64, nullptr, LC, F.cx, F.cy, 1);
        } else {
            if (G.dim == 3)
                hipLaunchKernelGGL((k_coef_restrict<3, false>), g, blk, 0, c->stream, nullptr, F.bx64, F.by64, F.bz64,
                                   LF, C.bx64, C.by64, C.bz64, LC, F.cx, F.cy, F.cz);
            else
                hipLaunchKernelGGL((k_coef_restrict<2, false>), g, blk, 0, c->stream, nullptr, F.bx64, F.by64,
                                   nullptr, LF, C.bx64, C.by64, nullptr, LC, F.cx, F.cy, 1);
        }
        FCFD_CHECK(hipGetLastError());
        long n = C.ntot;
        unsigned nb = (unsigned)((n + 255) / 256);
        hipLaunchKernelGGL(k_to_f32, dim3(nb), dim3(256), 0, c->stream, C.bx64, C.bx32, n);
        hipLaunchKernelGGL(k_to_f32, dim3(nb), dim3(256), 0, c->stream, C.by64, C.by32, n);
        if (G.dim == 3) hipLaunchKernelGGL(k_to_f32, dim3(nb), dim3(256), 0, c->stream, C.bz64, C.bz32, n);
        if (C.bx16) {
            hipLaunchKernelGGL(k_to_f16, dim3(nb), dim3(256), 0, c->stream, C.bx64, C.bx16, n);
            hipLaunchKernelGGL(k_to_f16, dim3(nb), dim3(256), 0, c->stream, C.by64, C.by16, n);
            hipLaunchKernelGGL(k_to_f16, dim3(nb), dim3(256), 0, c->stream, C.bz64, C.bz16, n);
        }
    }
    hipLaunchKernelGGL(k_rho_split, grd3(G.nx, G.ny, G.nz, blk), blk, 0, c->stream, c->f[FCFD_RHO], lay_of(G),
                       c->rho32, c->rho16, split_of(G));
    FCFD_CHECK(hipGetLastError());
    {
        MgLevel &M = c->mg.back();
        const long n = case_cells(M, G.dim);
        if (c->mg.size() > 1 && n <= CDENSE) {
            Lay L = lay_case(M, G.dim);
            Inv2 ih = inv2(M.hx, M.hy, M.hz);
            if (G.dim == 3)
                hipLaunchKernelGGL(k_coarse_inv<3>, dim3(1), dim3(CINV_THREADS), 0, c->stream, M.bx64, M.by64,
                                   M.bz64, L, ih, coarse_inv(c));
            else
                hipLaunchKernelGGL(k_coarse_inv<2>, dim3(c->nb), dim3(CINV_THREADS), 0, c->stream, M.bx64, M.by64,
                                   nullptr, L, ih, coarse_inv(c));
            FCFD_CHECK(hipGetLastError());
        }
    }
    c->mg_ready = 1;
    return FCFD_OK;
}

// One V-cycle from level l with error e_l (level 0: the solution itself)
// and rhs f_l of precision T.
template <typename T, int DIM>
static int vcycle(fcfd_ctx *c, int l, T *e0, const T *f0, const T *rho0, int nu1, int nu2,
                  const Dots<T> *dots = nullptr, bool zero_init = false) {
    const Geom &G = c->G;
    const int nl = (int)c->mg.size();
    if (dots && nu2 < 1) return FCFD_ERR_ARG;        // the dots come from the last post-smoothing sweep
    // FP32 cycle: level 0 may read the density in FP16 (the split copy rho16)
    const bool half_rho = sizeof(T) == 4 && c->pr.ppe_rho_half != 0 && c->rho16;
    if (zero_init && nu1 < 1) return FCFD_ERR_ARG;   // the first pre-sweep stands in for clearing e
    auto arrays = [&](int lv, T *&e, const T *&f, const T *&bx, const T *&by, const T *&bz) {
        MgLevel &M = c->mg[lv];
        if (lv == 0) { e = e0; f = f0; bx = by = bz = nullptr; return; }
        if constexpr (sizeof(T) == 8) {
            e = (T *)M.e64; f = (const T *)M.f64; bx = (const T *)M.bx64; by = (const T *)M.by64; bz = (const T *)M.bz64;
        } else {
            e = (T *)M.e32; f = (const T *)M.f32; bx = (const T *)M.bx32; by = (const T *)M.by32; bz = (const T *)M.bz32;
        }
    };
    // sweeps of RB-GS; with prol the coarse correction is added by the first
    // sweep, with dots the last half-sweep of level 0 forms the CG dots
    auto smooth = [&](int lv, int sweeps, bool reverse, const Prol<T> *prol, const Dots<T> *dt,
                      bool zinit) -> int {
        MgLevel &M = c->mg[lv];
        Lay L = lv == 0 ? lay_of(G) : lay_of(M, G.dim);
        Inv2 ih = inv2(M.hx, M.hy, M.hz);
        T *e; const T *f, *bx, *by, *bz;
        arrays(lv, e, f, bx, by, bz);
        const dim3 blk = mg_blk((M.nx + 1) / 2);   // a thread per cell pair of a row
        dim3 g((unsigned)(((M.nx + 1) / 2 + blk.x - 1) / blk.x), (M.ny + blk.y - 1) / blk.y, M.nz);
        const Prol<T> P0 = prol ? *prol : Prol<T>{nullptr, Lay{}, 0, 0, 0};
        const Dots<T> D0 = dt ? *dt : Dots<T>{nullptr, nullptr, nullptr};
        if constexpr (sizeof(T) == 4) {
            // small FP32 levels: fused sweeps ping-ponging between e32 and
            // e32b; an even number of launches ends back in e32.  Two sweeps
            // per launch (one 123 KB block per CU) while the tiles fit in one
            // round on the device, else one sweep per launch (two blocks per CU)
            const long n0c = (long)c->mg[0].nx * c->mg[0].ny * c->mg[0].nz;
            // 0 auto, 1 / 2 sweeps per launch, 3 off, 4 auto on every level >= 1
            const int fused = c->pr.ppe_fused;
            const bool small = fused == 4 || 32L * M.nx * M.ny * M.nz <= n0c;
            if (fused != 3 && lv >= 1 && M.e32b && sweeps >= 2 && small && !dt && !zinit) {
                using FT = FusedTile<DIM>;
                const dim3 gf((M.nx + FT::TX - 1) / FT::TX, (M.ny + FT::TY - 1) / FT::TY,
                              DIM == 3 ? (M.nz + FT::TZ - 1) / FT::TZ : M.nz);
                const int c0 = reverse ? 1 : 0;
                int n2, n1;
                if (fused == 2 || ((fused == 0 || fused == 4) && (long)gf.x * gf.y * gf.z <= c->ncu)) {
                    n2 = sweeps / 2;
                    n1 = sweeps % 2;
                    if ((n2 + n1) % 2) { n2 -= 1; n1 += 2; }
                } else {
                    n2 = sweeps % 2;   // one two-sweep launch keeps an odd sweep count even in launches
                    n1 = sweeps - 2 * n2;
                }
                float *bufs[2] = {M.e32, M.e32b};
                for (int q = 0; q < n2 + n1; ++q) {
                    ProfScope ps_(c, R_SMOOTHC);
                    const int pm = prol && q == 0;
                    const float *ei = bufs[q & 1];
                    float *eo = bufs[(q + 1) & 1];
#define FCFD_FUSED(NSW, PR)                                                                                    \
    hipLaunchKernelGGL((k_smooth_fused<DIM, NSW, PR>), gf, dim3(FUSED_THREADS), 0, c->stream, ei, eo, f, bx, by, bz, \
                       L, ih, c0, P0)
                    if (q < n2) {
                        if (pm) FCFD_FUSED(2, true);
                        else FCFD_FUSED(2, false);
                    } else {
                        if (pm) FCFD_FUSED(1, true);
                        else FCFD_FUSED(1, false);
                    }
#undef FCFD_FUSED
                }
                FCFD_CHECK(hipGetLastError());
                return FCFD_OK;
            }
        }
        for (int s = 0; s < sweeps; ++s)
            for (int h = 0; h < 2; ++h) {
                ProfScope ps_(c, lv == 0 ? R_SMOOTH0 : R_SMOOTHC);
                const int color = reverse ? 1 - h : h;
                const int pm = (zinit && s == 0) ? 3 + h : ((prol && s == 0) ? 1 + h : 0);
                const bool dm = dt && lv == 0 && s == sweeps - 1 && h == 1;
#define FCFD_SMOOTH(FINE, PM, DM)                                                                              \
    hipLaunchKernelGGL((k_smooth<T, DIM, FINE, PM, DM>), g, blk, 0, c->stream, e, f, FINE ? rho0 : (const T *)nullptr, \
                       bx, by, bz, L, ih, color, P0, D0)
                if (lv == 0) {
                    if constexpr (sizeof(T) == 4) {
                        // colour-split FP32 level 0: the second half of a
                        // prolongation or zero-guess sweep is a plain one
                        const Split S = split_of(G);
#define FCFD_SPLIT(MD, DM)                                                                                     \
    do {                                                                                                       \
        if (half_rho)                                                                                          \
            hipLaunchKernelGGL((k_smooth_split<DIM, MD, DM, _Float16>), g, blk, 0, c->stream, e, f,            \
                               (const _Float16 *)c->rho16, S, ih, color, P0, D0, L);                          \
        else                                                                                                   \
            hipLaunchKernelGGL((k_smooth_split<DIM, MD, DM, float>), g, blk, 0, c->stream, e, f,               \
                               (const float *)rho0, S, ih, color, P0, D0, L);                                 \
    } while (0)
                        if (dm) FCFD_SPLIT(0, true);
                        else if (pm == 1) FCFD_SPLIT(1, false);
                        else if (pm == 3) FCFD_SPLIT(2, false);
                        else FCFD_SPLIT(0, false);
#undef FCFD_SPLIT
                    } else {
                        if (dm) {
                            if (pm == 2) FCFD_SMOOTH(true, 2, true);
                            else FCFD_SMOOTH(true, 0, true);
                        } else if (pm == 1) FCFD_SMOOTH(true, 1, false);
                        else if (pm == 2) FCFD_SMOOTH(true, 2, false);
                        else if (pm == 3) FCFD_SMOOTH(true, 3, false);
                        else if (pm == 4) FCFD_SMOOTH(true, 4, false);
                        else FCFD_SMOOTH(true, 0, false);
                    }
                } else {
                    bool done16 = false;
                    if constexpr (sizeof(T) == 4) {
                        if (half_rho && M.bx16) {   // FP16 face coefficients (level 1 in 3D)
                            const _Float16 *hx = M.bx16, *hy = M.by16, *hz = M.bz16;
#define FCFD_SMOOTH16(PM)                                                                                      \
    hipLaunchKernelGGL((k_smooth<T, DIM, false, PM, false, _Float16>), g, blk, 0, c->stream, e, f,            \
                       (const T *)nullptr, hx, hy, hz, L, ih, color, P0, D0)
                            if (pm == 1) FCFD_SMOOTH16(1);
                            else if (pm == 2) FCFD_SMOOTH16(2);
                            else FCFD_SMOOTH16(0);
#undef FCFD_SMOOTH16
                            done16 = true;
                        }
                    }
                    if (!done16) {
                        if (pm == 1) FCFD_SMOOTH(false, 1, false);
                        else if (pm == 2) FCFD_SMOOTH(false, 2, false);
                        else FCFD_SMOOTH(false, 0, false);
                    }
                }
#undef FCFD_SMOOTH
            }
        FCFD_CHECK(hipGetLastError());
        return FCFD_OK;
    };
    // The coarse-grid correction of this aggregation hierarchy is its weak
    // part: levels of <= 1/32 of the finest cells get more sweeps (measured at
    // 256^3: 6 sweeps there cut flexible-CG iterations from 12 to 10 and
    // plain V-cycles from 30 to 19; the launch-bound small levels cost little).
    const long n0 = (long)c->mg[0].nx * c->mg[0].ny * c->mg[0].nz;
    const int nuc = c->pr.ppe_nu_coarse > 0 ? c->pr.ppe_nu_coarse : 6;
    auto sweeps = [&](int lv, int nu) {
        const MgLevel &M = c->mg[lv];
        return 32L * M.nx * M.ny * M.nz <= n0 ? nuc : nu;
    };
    // levels [lt, nl) run inside one single-workgroup launch (k_mg_tail) when
    // they are all small and the coarsest has a dense inverse; else lt = nl
    int lt = nl, top = 0;
    auto ncell_of = [&](int lv) { return case_cells(c->mg[lv], G.dim); };   // per case
    if (nl > 1 && ncell_of(nl - 1) <= CDENSE) {
        for (int lv = std::max(l, 1); lv < nl; ++lv)
            if (ncell_of(lv) <= TAIL_MAX) { lt = lv; break; }
        // FP32: one more level above, resident in LDS and registers
        if constexpr (sizeof(T) == 4) {
            const int lv = lt - 1;
            const MgLevel &F = c->mg[lv];
            const int nch = DIM == 3 ? 8 : 4;
            const bool full = F.cx == 2 && F.cy == 2 && (DIM == 2 || F.cz == 2) && F.nx % 2 == 0 &&
                              F.ny % 2 == 0 && (DIM == 2 || F.nz % 2 == 0);
            if (lt < nl && lv >= std::max(l, 1) && full && ncell_of(lv) <= TAIL_TOP) {
                const MgLevel &C = c->mg[lv + 1];
                const long plane = DIM == 3 ? (long)C.nx * C.ny : C.nx;
                if (nch * (ncell_of(lv + 1) + 2 * plane) <= TAIL_LDS) { lt = lv; top = 1; }
            }
        }
        if (nl - lt > TAIL_LEVELS) { lt = nl; top = 0; }
    }
    const int ldown = lt < nl ? lt : nl - 1;   // levels smoothed by separate launches: [l, ldown)
    // down
    for (int lv = l; lv < ldown; ++lv) {
        FCFD_TRY(smooth(lv, sweeps(lv, nu1), false, nullptr, nullptr, zero_init && lv == 0));
        MgLevel &M = c->mg[lv], &C = c->mg[lv + 1];
        Lay L = lv == 0 ? lay_of(G) : lay_of(M, G.dim), LC = lay_of(C, G.dim);
        Inv2 ih = inv2(M.hx, M.hy, M.hz);
        T *e; const T *f, *bx, *by, *bz;
        arrays(lv, e, f, bx, by, bz);
        T *ec; const T *fc_, *a, *b, *cc;
        arrays(lv + 1, ec, fc_, a, b, cc);
        const dim3 blk = mg_blk(C.nx);
        dim3 g = grd3(C.nx, C.ny, C.nz, blk);
        ProfScope ps_(c, lv == 0 ? R_RESTRICT0 : R_RESTRICTC);
        const int mz = DIM == 3 ? M.cz : 1;
        const int c2 = M.cx == 2 && M.cy == 2 && (mz == 1 || mz == 2) ? mz : 0;   // compile-time children
        // the kernel instances by compile-time children (0: runtime factors)
#define FCFD_RR3(KERNEL_0, KERNEL_1, KERNEL_2, ...)                                                          \
    do {                                                                                                       \
        if (c2 == 2) hipLaunchKernelGGL(KERNEL_2, g, blk, 0, c->stream, __VA_ARGS__);                           \
        else if (c2 == 1) hipLaunchKernelGGL(KERNEL_1, g, blk, 0, c->stream, __VA_ARGS__);                      \
        else hipLaunchKernelGGL(KERNEL_0, g, blk, 0, c->stream, __VA_ARGS__);                                   \
    } while (0)
        if (lv == 0) {
            if constexpr (sizeof(T) == 4) {
                const Split S0 = split_of(G);
                if (half_rho) {
                    const _Float16 *r16 = (const _Float16 *)c->rho16;
                    FCFD_RR3((k_resid_restrict_split<DIM, _Float16, 0>), (k_resid_restrict_split<DIM, _Float16, 1>),
                             (k_resid_restrict_split<DIM, _Float16, 2>), e, f, r16, S0, ih, (T *)fc_, ec, LC, M.cx,
                             M.cy, mz);
                } else {
                    const float *r32 = (const float *)rho0;
                    FCFD_RR3((k_resid_restrict_split<DIM, float, 0>), (k_resid_restrict_split<DIM, float, 1>),
                             (k_resid_restrict_split<DIM, float, 2>), e, f, r32, S0, ih, (T *)fc_, ec, LC, M.cx,
                             M.cy, mz);
                }
            } else {
                FCFD_RR3((k_resid_restrict<T, DIM, true, 0>), (k_resid_restrict<T, DIM, true, 1>),
                         (k_resid_restrict<T, DIM, true, 2>), e, f, rho0, bx, by, bz, L, ih, (T *)fc_, ec, LC, M.cx,
                         M.cy, mz);
            }
        } else if (sizeof(T) == 4 && half_rho && M.bx16) {
            // the FP16 face coefficients the level's smoothing reads
            if constexpr (sizeof(T) == 4) {
                const _Float16 *hx = M.bx16, *hy = M.by16, *hz = M.bz16;
                FCFD_RR3((k_resid_restrict<T, DIM, false, 0, _Float16>), (k_resid_restrict<T, DIM, false, 1, _Float16>),
                         (k_resid_restrict<T, DIM, false, 2, _Float16>), e, f, (const T *)nullptr, hx, hy, hz, L, ih,
                         (T *)fc_, ec, LC, M.cx, M.cy, mz);
            }
        } else {
            FCFD_RR3((k_resid_restrict<T, DIM, false, 0>), (k_resid_restrict<T, DIM, false, 1>),
                     (k_resid_restrict<T, DIM, false, 2>), e, f, (const T *)nullptr, bx, by, bz, L, ih, (T *)fc_, ec,
                     LC, M.cx, M.cy, mz);
        }
#undef FCFD_RR3
        FCFD_CHECK(hipGetLastError());
    }
    if (lt < nl) {
        TailArgs<T> A{};
        A.n = nl - lt;
        A.top = top;
        A.minv = coarse_inv(c);
        for (int q = 0; q < A.n; ++q) {
            const int lv = lt + q;
            MgLevel &M = c->mg[lv];
            TailLevel<T> &V = A.lv[q];
            T *e; const T *f, *bx, *by, *bz;
            arrays(lv, e, f, bx, by, bz);
            V.e = e; V.f = (T *)f; V.bx = bx; V.by = by; V.bz = bz;
            V.L = lay_case(M, G.dim);
            V.ih = inv2(M.hx, M.hy, M.hz);
            V.cx = M.cx; V.cy = M.cy; V.cz = DIM == 3 ? M.cz : 1;
            V.nu1 = sweeps(lv, nu1);
            V.nu2 = sweeps(lv, nu2);
        }
        // the levels below the top keep their error in LDS when their
        // ghosted boxes fit
        {
            int off = 0;
            bool fit = true;
            for (int q = A.top; q < A.n; ++q) {
                const Lay &L = A.lv[q].L;
                A.eoff[q] = off;
                off += (L.nx + 2) * (L.ny + 2) * (DIM == 3 ? L.nz + 2 : 1);
                fit = fit && (long)L.nx * L.ny * (DIM == 3 ? L.nz : 1) <= TAIL_MAX;
            }
            A.lds = fit && off <= TAIL_ELDS ? 1 : 0;
        }
        ProfScope ps_(c, R_COARSE);
        hipLaunchKernelGGL((k_mg_tail<T, DIM>), dim3(DIM == 2 ? c->nb : 1), dim3(TAIL_THREADS), 0, c->stream, A);
        FCFD_CHECK(hipGetLastError());
    } else {
        // coarsest: dense inverse or FP64 CG in one workgroup
        MgLevel &M = c->mg[nl - 1];
        Lay L = lay_of(M, G.dim);
        Inv2 ih = inv2(M.hx, M.hy, M.hz);
        T *e; const T *f, *bx, *by, *bz;
        arrays(nl - 1, e, f, bx, by, bz);
        long n = case_cells(M, G.dim);
        ProfScope ps_(c, R_COARSE);
        if (nl > 1 && n <= CDENSE)
            hipLaunchKernelGGL((k_coarse_dense<T>), dim3(DIM == 2 ? c->nb : 1), dim3(CDENSE), 0, c->stream, e, f,
                               (const double *)coarse_inv(c), lay_case(M, G.dim), DIM);
        else   // one workgroup per case
            hipLaunchKernelGGL((k_coarse_cg<T, DIM>), dim3(DIM == 2 ? c->nb : 1), dim3(n >= 1024 ? 1024 : 256), 0,
                               c->stream, e, f, bx, by, bz, M.bx64, M.by64, M.bz64, coarse_inv(c),
                               lay_case(M, G.dim), ih, (int)std::min<long>(n, 96), 0.0);
        FCFD_CHECK(hipGetLastError());
    }
    // up
    for (int lv = ldown - 1; lv >= l; --lv) {
        MgLevel &M = c->mg[lv], &C = c->mg[lv + 1];
        Lay L = lv == 0 ? lay_of(G) : lay_of(M, G.dim), LC = lay_of(C, G.dim);
        T *e; const T *f, *bx, *by, *bz;
        arrays(lv, e, f, bx, by, bz);
        T *ec; const T *fc_, *a, *b, *cc;
        arrays(lv + 1, ec, fc_, a, b, cc);
        if (nu2 > 0) {
            // prolongation folded into the first post-smoothing sweep
            const Prol<T> P{(const T *)ec, LC, M.cx == 2, M.cy == 2, DIM == 3 && M.cz == 2};
            FCFD_TRY(smooth(lv, sweeps(lv, nu2), true, &P, lv == 0 ? dots : nullptr, false));
        } else {
            if (lv == 0 && sizeof(T) == 4) return FCFD_ERR_ARG;   // the split level 0 prolongs inside a sweep
            ProfScope ps_(c, lv == 0 ? R_PROLONG0 : R_PROLONGC);
            const dim3 blk = mg_blk(M.nx);
            hipLaunchKernelGGL((k_prolong<T>), grd3(M.nx, M.ny, M.nz, blk), blk, 0, c->stream, e, (const T *)ec, L,
                               LC, M.cx, M.cy, DIM == 3 ? M.cz : 1);
            FCFD_CHECK(hipGetLastError());
        }
    }
    return FCFD_OK;
}

// ------------------------------------------------------ outer iteration

// FP64 residual r = f - A p of level 0 with ||r||^2 and (with_f) ||f||^2
// folded into block partials; optional outputs: r64 (Krylov residual), r32
// (rhs of the FP32 cycle).  The cycles' error arrays need no clearing: the
// first pre-smoothing half-sweeps take a zero initial guess.  With mean
// (per case, times mscale) the rhs is f - mean, formed here as k_center_copy
// forms it, so the flexible CG needs no centred copy of the source.
template <int DIM>
__global__ void __launch_bounds__(256) k_resid_out(const double *__restrict__ p, const double *__restrict__ f,
                                                   const double *__restrict__ rho, Lay L, Inv2 ih,
                                                   double *__restrict__ r64, float *__restrict__ r32, Split S,
                                                   double *__restrict__ partial, int with_f,
                                                   const double *__restrict__ mean, double mscale) {
    __shared__ double sh[4];
    int bx_, by_, bz_;
    // 3D: each XCD walks the planes of 8-row slabs, so the k-1 and k+1 planes
    // of p and rho are still in its L2 (whole planes per XCD re-fetched them:
    // 1.67x the algorithmic bytes at 256^3)
    if (DIM == 3) xcd_block_slab(bx_, by_, bz_, 8);
    else xcd_block(bx_, by_, bz_);
    const int i = bx_ * blockDim.x + threadIdx.x, j = by_ * blockDim.y + threadIdx.y, k = bz_;
    double rr = 0.0, ff = 0.0;
    if (i < L.nx && j < L.ny) {
        long id = L.id(i, j, k);
        const double fv = mean ? f[id] - mean[(DIM == 2 ? (long)k : 0L) * SCS] * mscale : f[id];
        double r = fv - apply_A<double, DIM, true, double>(p, rho, nullptr, nullptr, nullptr, L, ih, i, j, k, id);
        rr = r * r;
        if (with_f) ff = fv * fv;
        if (r64) r64[id] = r;
        if (r32) r32[S.at(i, j, k)] = (float)r;
    }
    rr = block_sum256(rr, sh);
    ff = block_sum256(ff, sh);
    if (threadIdx.x + threadIdx.y == 0) {
        long b = bx_ + (long)gridDim.x * (by_ + (long)gridDim.y * bz_);
        long nb = (long)gridDim.x * gridDim.y * gridDim.z;
        partial[b] = rr;
        partial[nb + b] = ff;
    }
}

// new direction: p = z + beta p, q = A z + beta q (so q = A p), z_old = z,
// partial sums of p.q.  The previous iteration's x += alpha p is applied
// here, where the old p is read anyway (alpha = 0 on the first iteration;
// the last one is applied by k_fcg_finish).  z and z_old are FP64 in the
// field layout, or FP32 in the colour-split layout S.  Each thread walks
// FCG_KC planes of its (i, j) column and carries z and rho of the planes
// k-1, k, k+1 in registers: a plane-per-block launch re-fetched both from
// HBM, because the ten streams of one plane overflow an XCD's 4 MB L2
// before the next plane needs them again.  At the default occupancy target
// (8 waves per SIMD, <= 64 VGPRs) the compiler issues the neighbour loads of
// a plane one at a time, each followed by a wait; allowing 5 waves per SIMD
// lets it issue them together (about 7 % faster at 256^3).  Four planes
// per thread: twice the blocks of eight, which hides more latency than the
// two boundary planes per column re-read cost (256^3 step 0.25 %, M3 0.6 %
// faster than eight).
constexpr int FCG_KC = 4;
template <typename T, int DIM>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(1, 5))) k_fcg_dir(const T *__restrict__ z, T *__restrict__ zo,
                                                 double *__restrict__ p, double *__restrict__ q,
                                                 double *__restrict__ x, const double *__restrict__ rho, Lay L,
                                                 Split S, Inv2 ih, const double *__restrict__ sc,
                                                 double *__restrict__ partial) {
    if (L.skipped()) return;
    __shared__ double sh[4];
    int bx_, by_, bz_;
    xcd_block(bx_, by_, bz_);
    const int i = bx_ * blockDim.x + threadIdx.x, j = by_ * blockDim.y + threadIdx.y;
    // 3D: FCG_KC planes per block; 2D: the block's slice (the case of a batched stack)
    const int k0 = DIM == 3 ? bz_ * FCG_KC : bz_, k1 = DIM == 3 ? min(k0 + FCG_KC, L.nz) : bz_ + 1;
    sc += (long)case_of<DIM>(bz_) * SCS;
    const double beta = sc[SC_BETA], alpha_prev = sc[SC_ALPHA];
    auto zid = [&](int ii, int jj, int kk) -> long {
        if constexpr (sizeof(T) == 4) return S.at(ii, jj, kk);
        else return L.id(ii, jj, kk);
    };
    double pq = 0.0;
    if (i < L.nx && j < L.ny) {
        long id = L.id(i, j, k0);
        // A z with the FP64 operator; z promoted.  (b, c, t): planes k-1, k, k+1
        T zc = z[zid(i, j, k0)];
        double zb = 0.0, rb = 0.0, rc = rho[id];
        if (DIM == 3) {
            zb = (double)z[zid(i, j, k0 - 1)];
            rb = rho[id - L.sxy];
        }
        for (int k = k0; k < k1; ++k, id += L.sxy) {
            const double zp = (double)zc;
            T zt = (T)0;
            double rt = 0.0, acc = 0.0;
            acc += face<double>(i < L.nx - 1, rc, rho[id + 1], ih.x) * ((double)z[zid(i + 1, j, k)] - zp);
            acc += face<double>(i > 0, rc, rho[id - 1], ih.x) * ((double)z[zid(i - 1, j, k)] - zp);
            acc += face<double>(j < L.ny - 1, rc, rho[id + L.sx], ih.y) * ((double)z[zid(i, j + 1, k)] - zp);
            acc += face<double>(j > 0, rc, rho[id - L.sx], ih.y) * ((double)z[zid(i, j - 1, k)] - zp);
            if (DIM == 3) {
                zt = z[zid(i, j, k + 1)];
                rt = rho[id + L.sxy];
                acc += face<double>(k < L.nz - 1, rc, rt, ih.z) * ((double)zt - zp);
                acc += face<double>(k > 0, rc, rb, ih.z) * (zb - zp);
            }
            const double po = p[id];
            x[id] += alpha_prev * po;
            double pn = zp + beta * po;
            double qn = acc + beta * q[id];
            p[id] = pn;
            q[id] = qn;
            zo[zid(i, j, k)] = zc;
            pq += pn * qn;
            zb = zp; zc = zt; rb = rc; rc = rt;
        }
    }
    pq = block_sum256(pq, sh);
    if (threadIdx.x + threadIdx.y == 0) partial[bx_ + (long)gridDim.x * (by_ + (long)gridDim.y * bz_)] = pq;
}

// r -= alpha q, ||r||^2; the new residual also becomes the rhs of the next
// preconditioning cycle (r32 for the FP32 cycle; the FP64 cycle reads r)
__global__ void __launch_bounds__(256) k_fcg_update(const double *__restrict__ q, double *__restrict__ r, Lay L,
                                                    const double *__restrict__ sc, float *__restrict__ r32, Split S,
                                                    double *__restrict__ partial, int dim) {
    if (L.skipped()) return;
    __shared__ double sh[4];
    int bx_, by_, bz_;
    xcd_block(bx_, by_, bz_);
    const int i = bx_ * blockDim.x + threadIdx.x, j = by_ * blockDim.y + threadIdx.y, k = bz_;
    const double alpha = sc[(dim == 2 ? (long)k : 0L) * SCS + SC_ALPHA];
    double rr = 0.0;
    if (i < L.nx && j < L.ny) {
        long id = L.id(i, j, k);
        double rn = r[id] - alpha * q[id];
        r[id] = rn;
        rr = rn * rn;
        if (r32) r32[S.at(i, j, k)] = (float)rn;
    }
    rr = block_sum256(rr, sh);
    if (threadIdx.x + threadIdx.y == 0) partial[bx_ + (long)gridDim.x * (by_ + (long)gridDim.y * bz_)] = rr;
}

// the last iteration's x += alpha p
__global__ void __launch_bounds__(256) k_fcg_finish(double *__restrict__ x, const double *__restrict__ p, Lay L,
                                                    const double *__restrict__ sc, int dim) {
    int bx_, by_, bz_;
    xcd_block(bx_, by_, bz_);
    const int i = bx_ * blockDim.x + threadIdx.x, j = by_ * blockDim.y + threadIdx.y, k = bz_;
    if (i >= L.nx || j >= L.ny) return;
    const long id = L.id(i, j, k);
    x[id] += sc[(dim == 2 ? (long)k : 0L) * SCS + SC_ALPHA] * p[id];
}

// out[q SCS + s] = sums of the ns slots of the block partials of a level-0
// launch of nb blocks, per case q of the context, then the scalar step `op`
// (SUM_*) on each case's scalars
// Up to SUM1_MAX partials per case one block reduces them in one launch;
// beyond, the 256-block first stage is faster than one block's reads (at
// 256^3 a single block over the 65536 partials of a plane-per-block launch
// made the step 1 % slower, while the batched 2D stacks, 128-2048 partials
// per case, ran 7 % faster in one launch).
constexpr long SUM1_MAX = 16384;
static int sum_partials(fcfd_ctx *c, long nb, int ns, double *out, int op = SUM_STORE) {
    const long nbq = nb / c->nb;
    if (nbq <= SUM1_MAX) {
        hipLaunchKernelGGL(k_sum_one, dim3(c->nb), dim3(SUM1_T), 0, c->stream, c->red.partial, nbq, ns, out, op,
                           c->mg_scalars, t_skip);
    } else {
        const long chunk = (nbq + 255) / 256;
        hipLaunchKernelGGL(k_sum_stage1, dim3(256, ns, c->nb), dim3(256), 0, c->stream, c->red.partial, nbq, chunk,
                           c->red.partial2, t_skip);
        hipLaunchKernelGGL(k_sum_stage2, dim3(c->nb), dim3(256), 0, c->stream, c->red.partial2, ns, out, op,
                           c->mg_scalars, t_skip);
    }
    FCFD_CHECK(hipGetLastError());
    return FCFD_OK;
}

// sum ns slots of the block partials of a level-0 launch and bring them to
// the host: out[q ns + s] for case q
static int sums_to_host(fcfd_ctx *c, long nb, int ns, double *out) {
    FCFD_TRY(sum_partials(c, nb, ns, c->mg_scalars));
    FCFD_TRY(gather_to_host(c, c->red.host, c->mg_scalars, SCS, ns, c->nb));
    FCFD_CHECK(hipStreamSynchronize(c->stream));
    for (int q = 0; q < c->nb * ns; ++q) out[q] = c->red.host[q];
    return FCFD_OK;
}

// mean of an interior FP64 field into mg_scalars[q SCS + SC_MEAN] per case (device)
static int field_mean(fcfd_ctx *c, const double *a) {
    const Geom &G = c->G;
    dim3 blk = mg_blk(G.nx), g = grd3(G.nx, G.ny, G.nz, blk);
    if (G.dim == 3) g.z = (G.nz + SUM_KC - 1) / SUM_KC;
    long nb = (long)g.x * g.y * g.z;
    if (G.dim == 3) hipLaunchKernelGGL(k_sum<3>, g, blk, 0, c->stream, a, lay_of(G), c->red.partial);
    else hipLaunchKernelGGL(k_sum<2>, g, blk, 0, c->stream, a, lay_of(G), c->red.partial);
    FCFD_TRY(sum_partials(c, nb, 1, c->mg_scalars + SC_MEAN));
    return FCFD_OK;
}

// Replay `body` from a graph captured on its first use for (kind, inner):
// the level arrays, rho and the level-0 vectors are fixed for the context and
// every scalar the body needs lives in device memory.  With the profiler on,
// the body is launched directly (events inside a graph cannot be timed).
template <typename Body>
static int run_graph(fcfd_ctx *c, int kind, int inner, Body body, int slot = 0) {
    static const bool no_graph = getenv("FCFD_NO_GRAPH") != nullptr;
    if (no_graph || c->prof.on) return body();
    hipGraphExec_t &ex = c->vgraph[slot][kind][inner];
    if (!ex) {
        hipGraph_t graph;
        c->capture_kind = kind;
        c->capture_inner = inner;
        c->vgraph_prof[kind][inner].clear();
        FCFD_CHECK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        int rc = body();
        hipError_t e2 = hipStreamEndCapture(c->stream, &graph);
        c->capture_kind = -1;
        if (rc != FCFD_OK) return rc;
        FCFD_CHECK(e2);
        FCFD_CHECK(hipGraphInstantiate(&ex, graph, nullptr, nullptr, 0));
        FCFD_CHECK(hipGraphDestroy(graph));
    }
    FCFD_CHECK(hipGraphLaunch(ex, c->stream));
    return FCFD_OK;
}

// `inner` V-cycles of one kind: 0 FP64 on (P, f64); 2 FP32 on (e32, f32)
template <int DIM>
static int vcycles(fcfd_ctx *c, int kind, int inner) {
    MgLevel &M0 = c->mg[0];
    const int nu1 = c->pr.ppe_nu1 > 0 ? c->pr.ppe_nu1 : 2, nu2 = c->pr.ppe_nu2 > 0 ? c->pr.ppe_nu2 : 2;
    for (int t = 0; t < inner; ++t) {
        if (kind == 0) FCFD_TRY((vcycle<double, DIM>(c, 0, c->f[FCFD_P], M0.f64, c->f[FCFD_RHO], nu1, nu2)));
        else FCFD_TRY((vcycle<float, DIM>(c, 0, M0.e32, M0.f32, c->rho32, nu1, nu2, nullptr, t == 0)));
    }
    return FCFD_OK;
}

// One flexible-CG iteration, all on the device: z = V-cycle(s)(r) (kind 1:
// FP64 z, kind 3: FP32 z) with r.z and r.zold folded into the last sweep,
// beta; p = z + beta p, q = A z + beta q, z_old = z, p.q; alpha; x += alpha p,
// r -= alpha q, the next cycle's rhs; ||r||^2; the stopping test of every
// case (k_cg_check); then the stopping state to host buffer `slot`.
template <int DIM>
static int fcg_iteration(fcfd_ctx *c, int kind, int inner, int slot) {
    const Geom &G = c->G;
    dim3 blk = mg_blk(G.nx), g = grd3(G.nx, G.ny, G.nz, blk);
    const long nb = (long)g.x * g.y * g.z;
    // blocks of the level-0 half-sweep that folds in the CG dots (vcycle's smooth)
    const dim3 bs = mg_blk((G.nx + 1) / 2);
    const long nb_s = (long)(((G.nx + 1) / 2 + bs.x - 1) / bs.x) * ((G.ny + bs.y - 1) / bs.y) * G.nz;
    const dim3 gdir(g.x, g.y, DIM == 3 ? (G.nz + FCG_KC - 1) / FCG_KC : G.nz);
    const Lay L0 = lay_of(G);
    const Split S0 = split_of(G);
    const Inv2 ih = inv2(G.dx, G.dy, G.dz);
    MgLevel &M0 = c->mg[0];
    double *sc = c->mg_scalars, *rho = c->f[FCFD_RHO];
    const bool mixed = kind == 3;
    const int nu1 = c->pr.ppe_nu1 > 0 ? c->pr.ppe_nu1 : 2, nu2 = c->pr.ppe_nu2 > 0 ? c->pr.ppe_nu2 : 2;
    const Dots<double> d64{c->kr, c->kzo, c->red.partial};
    const Dots<float> d32{c->kr, c->kzo32, c->red.partial};
    for (int t = 0; t < inner; ++t) {
        const bool last = t == inner - 1;
        if (mixed)
            FCFD_TRY((vcycle<float, DIM>(c, 0, M0.e32, M0.f32, c->rho32, nu1, nu2, last ? &d32 : nullptr, t == 0)));
        else
            FCFD_TRY((vcycle<double, DIM>(c, 0, c->kz, c->kr, rho, nu1, nu2, last ? &d64 : nullptr, t == 0)));
    }
    FCFD_TRY(sum_partials(c, nb_s, 2, sc + SC_RZ, SUM_CG_BETA));
    {
        ProfScope ps_(c, R_FCG);
        if (mixed)
            hipLaunchKernelGGL((k_fcg_dir<float, DIM>), gdir, blk, 0, c->stream, M0.e32, c->kzo32, c->kp, c->kq,
                               c->f[FCFD_P], rho, L0, S0, ih, sc, c->red.partial);
        else
            hipLaunchKernelGGL((k_fcg_dir<double, DIM>), gdir, blk, 0, c->stream, c->kz, c->kzo, c->kp, c->kq,
                               c->f[FCFD_P], rho, L0, S0, ih, sc, c->red.partial);
        FCFD_CHECK(hipGetLastError());
    }
    FCFD_TRY(sum_partials(c, (long)gdir.x * gdir.y * gdir.z, 1, sc + SC_PQ, SUM_CG_ALPHA));
    {
        ProfScope ps_(c, R_FCG);
        hipLaunchKernelGGL(k_fcg_update, g, blk, 0, c->stream, c->kq, c->kr, L0, sc, mixed ? M0.f32 : nullptr, S0,
                           c->red.partial, DIM);
        FCFD_CHECK(hipGetLastError());
    }
    FCFD_TRY(sum_partials(c, nb, 1, sc + SC_RR, SUM_STORE));
    hipLaunchKernelGGL(k_cg_check, dim3(1), dim3(CG_THREADS), 0, c->stream, sc, c->nb, c->mg_skip);
    FCFD_CHECK(hipGetLastError());
    // SC_DONE .. SC_NAN of every case, packed
    FCFD_TRY(gather_to_host(c, c->red.host + (long)slot * SC_NST * c->nb, sc + SC_DONE, SCS, SC_NST, c->nb));
    return FCFD_OK;
}

// The solve runs the loop of a single system for every case of the context
// at once (one case unless it is a batched 2D stack).  Each case keeps its
// own stopping state (on the device for flexible CG: k_cg_start,
// k_cg_check); once a case's loop would have ended it is frozen on the
// device (SC_DONE; the mixed defect correction skips its update), so its
// result is that of its own solve.  FP64 plain V-cycles cannot freeze a case
// and go on improving finished ones until the last case stops.
// the host queues the next flexible-CG iteration ahead while the predicted
// residual after the iteration in flight exceeds AHEAD_MARGIN x tolerance
constexpr double AHEAD_MARGIN = 3.0;
template <int DIM>
static int ppe_solve_t(fcfd_ctx *c, double rel_tol, int precision, int *cycles, double *rel, int normalize) {
    const Geom &G = c->G;
    dim3 blk = mg_blk(G.nx), g = grd3(G.nx, G.ny, G.nz, blk);
    const long nb = (long)g.x * g.y * g.z;
    const Lay L0 = lay_of(G);
    const Split S0 = split_of(G);
    const Inv2 ih = inv2(G.dx, G.dy, G.dz);
    MgLevel &M0 = c->mg[0];
    const int nq = c->nb;
    const double ncell = DIM == 3 ? (double)G.nx * (double)G.ny * (double)G.nz : (double)G.nx * (double)G.ny;
    double *p = c->f[FCFD_P], *f = M0.f64, *rho = c->f[FCFD_RHO];
    const bool mixed = precision == FCFD_PPE_MIXED;
    const bool krylov = c->pr.ppe_krylov != 0;
    const int inner = std::max(1, std::min(c->pr.ppe_inner > 0 ? c->pr.ppe_inner : 1, 4));
    const int kind = mixed ? (krylov ? 3 : 2) : (krylov ? 1 : 0);
    const int maxc = c->pr.ppe_max_cycles > 0 ? c->pr.ppe_max_cycles : 200;
    static const bool verbose = getenv("FCFD_MG_VERBOSE") != nullptr;
    // compatibility of the Neumann problem: f = src - mean(src), per case;
    // the flexible CG reads f once (r0), so its residual kernel centres the
    // source itself, while plain V-cycles read the centred copy every cycle
    FCFD_TRY(field_mean(c, c->f[FCFD_PPE_SRC]));
    if (!krylov) {
        hipLaunchKernelGGL(k_center_copy, g, blk, 0, c->stream, c->f[FCFD_PPE_SRC], f, L0, c->mg_scalars + SC_MEAN,
                           1.0 / ncell, DIM);
        FCFD_CHECK(hipGetLastError());
    }
    // r0 = f - A p
    {
        ProfScope ps_(c, R_RESID0);
        if (krylov)
            hipLaunchKernelGGL(k_resid_out<DIM>, g, blk, 0, c->stream, p, (const double *)c->f[FCFD_PPE_SRC], rho,
                               L0, ih, c->kr, mixed ? M0.f32 : nullptr, S0, c->red.partial, 1,
                               (const double *)c->mg_scalars + SC_MEAN, 1.0 / ncell);
        else
            hipLaunchKernelGGL(k_resid_out<DIM>, g, blk, 0, c->stream, p, (const double *)f, rho, L0, ih,
                               (double *)nullptr, mixed ? M0.f32 : nullptr, S0, c->red.partial, 1,
                               (const double *)nullptr, 0.0);
        FCFD_CHECK(hipGetLastError());
    }
    std::vector<double> s2(2 * nq);
    struct Case_ {
        double fn, r, best;
        int k, since;
        bool done;
    };
    std::vector<Case_> cs(nq);
    int active = 0;
    if (krylov) {
        // ||r0||^2 and ||f||^2 stay on the device: k_cg_start forms each
        // case's stopping state there, and the host learns it with the first
        // iteration's status (no host round trip before the first cycle)
        FCFD_TRY(sum_partials(c, nb, 2, c->mg_scalars));
        for (Case_ &a : cs) a = Case_{1.0, 0.0, 0.0, 0, 0, false};
        active = nq;
    } else {
        FCFD_TRY(sums_to_host(c, nb, 2, s2.data()));
        for (int q = 0; q < nq; ++q) {
            Case_ &a = cs[q];
            a.fn = s2[2 * q + 1] > 0.0 ? sqrt(s2[2 * q + 1]) : 1.0;
            a.r = sqrt(s2[2 * q]) / a.fn;
            a.best = a.r;
            a.k = a.since = 0;
            if (!std::isfinite(a.r)) return FCFD_ERR_NAN;
            a.done = !(a.r > rel_tol && a.k < maxc);
            active += !a.done;
        }
    }
    auto rmax = [&]() {
        double m = 0.0;
        for (const Case_ &a : cs) m = std::max(m, a.r);
        return m;
    };
    if (verbose && !krylov) fprintf(stderr, "fcfd mg[%s]: it 0 rel %.3e\n", mixed ? "mixed" : "fp64", rmax());
    hipLaunchKernelGGL(k_cg_init, dim3(nq), dim3(1), 0, c->stream, c->mg_scalars);
    if (krylov)
        hipLaunchKernelGGL(k_cg_start, dim3(1), dim3(CG_THREADS), 0, c->stream, c->mg_scalars, nq, rel_tol, maxc,
                           c->mg_skip);
    else
        for (int q = 0; q < nq; ++q)
            if (cs[q].done) hipLaunchKernelGGL(k_cg_done, dim3(1), dim3(1), 0, c->stream, c->mg_scalars, q);
    FCFD_CHECK(hipGetLastError());
    int it = 0;
    if (krylov && active > 0) {
        // Flexible CG, one graph per iteration, pipelined: while the host
        // reads iteration it's stopping state (k_cg_check), iteration it + 1
        // is already queued, so the device does not wait on the host between
        // iterations.  An iteration queued after the solve has ended finds
        // the skip flag set and all its kernels return at once, but its
        // launches still cost time: the host queues ahead only while the
        // residual's last rate of decrease predicts that the iteration in
        // flight will not end every case.  Two graphs alternate, each
        // writing its own host status buffer.
        auto launch = [&](int slot) -> int {
            t_skip = c->mg_skip;
            const int rc = run_graph(
                c, kind, inner, [&]() { return fcg_iteration<DIM>(c, kind, inner, slot); }, slot);
            t_skip = nullptr;
            FCFD_TRY(rc);
            FCFD_CHECK(hipEventRecord(c->cg_ev[slot], c->stream));
            return FCFD_OK;
        };
        // the starting residual arrives with the first status (SC_R0)
        std::vector<double> rprev(nq, 0.0);
        // With the region profiler on, every launch is timed: an iteration
        // queued ahead and then skipped would enter the region averages as a
        // near-empty launch, so the profiled solve does not queue ahead.
        const bool pipelined = !c->prof.on;
        FCFD_TRY(launch(0));
        int queued = 1;
        bool ahead = pipelined;
        for (;;) {
            if (ahead && queued == it + 1 && queued < maxc) {
                FCFD_TRY(launch(queued & 1));
                ++queued;
            }
            const int slot = it & 1;
            FCFD_CHECK(hipEventSynchronize(c->cg_ev[slot]));
            FCFD_TRY(prof_graph_accum(c, kind, inner));
            ++it;
            const double *h = c->red.host + (long)slot * SC_NST * nq;
            bool nan = false;
            active = 0;
            ahead = false;
            for (int q = 0; q < nq; ++q) {
                const double *hq = h + (long)q * SC_NST;
                Case_ &a = cs[q];
                a.done = hq[0] != 0.0;
                a.k = (int)hq[SC_K - SC_DONE];
                a.r = hq[SC_R - SC_DONE];
                a.since = (int)hq[SC_SINCE - SC_DONE];
                nan = nan || hq[SC_NAN - SC_DONE] != 0.0;
                if (it == 1) rprev[q] = hq[SC_R0 - SC_DONE];
                active += !a.done;
                // residual after the next iteration at the last rate, with a
                // margin: an iteration queued in vain costs ~0.4 ms of
                // launches at 256^3, one not queued in time a host round trip
                if (!a.done && !(a.r * (a.r / rprev[q]) <= AHEAD_MARGIN * rel_tol)) ahead = pipelined;
                rprev[q] = a.r;
            }
            if (verbose) fprintf(stderr, "fcfd mg[%s+fcg]: it %d rel %.3e\n", mixed ? "mixed" : "fp64", it, rmax());
            if (nan) {
                FCFD_CHECK(hipStreamSynchronize(c->stream));
                return FCFD_ERR_NAN;
            }
            if (active == 0) break;
            if (queued == it) {   // nothing in flight: the next iteration is due now
                FCFD_TRY(launch(queued & 1));
                ++queued;
            }
        }
    }
    while (!krylov && active > 0) {
        // defect correction (mixed) or plain V-cycles on p (FP64)
        FCFD_TRY(run_graph(c, kind, inner, [&]() { return vcycles<DIM>(c, kind, inner); }));
        if (mixed) {
            hipLaunchKernelGGL(k_add_f32, g, blk, 0, c->stream, p, M0.e32, L0, S0, c->mg_scalars, DIM);
            FCFD_CHECK(hipGetLastError());
        }
        {
            ProfScope ps_(c, R_RESID0);
            hipLaunchKernelGGL(k_resid_out<DIM>, g, blk, 0, c->stream, p, (const double *)f, rho, L0, ih,
                               (double *)nullptr, mixed ? M0.f32 : nullptr, S0, c->red.partial, 0,
                               (const double *)nullptr, 0.0);
            FCFD_CHECK(hipGetLastError());
        }
        FCFD_TRY(sums_to_host(c, nb, 1, s2.data()));
        FCFD_TRY(prof_graph_accum(c, kind, inner));
        ++it;
        for (int q = 0; q < nq; ++q) {
            Case_ &a = cs[q];
            if (a.done) continue;
            ++a.k;
            a.r = sqrt(s2[q]) / a.fn;
            if (!std::isfinite(a.r)) return FCFD_ERR_NAN;
            // stagnation: the FP64 residual has reached its round-off floor
            if (a.r < 0.9 * a.best) { a.best = a.r; a.since = 0; }
            else if (++a.since >= 4) a.done = true;
            if (!(a.r > rel_tol && a.k < maxc)) a.done = true;
        }
        active = 0;
        for (const Case_ &a : cs) active += !a.done;
        if (verbose) fprintf(stderr, "fcfd mg[%s]: it %d rel %.3e\n", mixed ? "mixed" : "fp64", it, rmax());
        if (active > 0)
            for (int q = 0; q < nq; ++q)
                if (cs[q].done) hipLaunchKernelGGL(k_cg_done, dim3(1), dim3(1), 0, c->stream, c->mg_scalars, q);
    }
    int rc = FCFD_OK, kmax = 0;
    for (const Case_ &a : cs) {
        kmax = std::max(kmax, a.k);
        if (a.r > rel_tol && !(a.r <= 10.0 * rel_tol && a.since >= 4)) rc = FCFD_ERR_NOT_CONVERGED;
    }
    // (no iteration ran when every case had converged at the start)
    if (krylov && kmax > 0) {
        hipLaunchKernelGGL(k_fcg_finish, g, blk, 0, c->stream, p, c->kp, L0, c->mg_scalars, DIM);
        FCFD_CHECK(hipGetLastError());
    }
    const double r = rmax();
    // mean(p) = 0 per case as the reference's iteration enforces (skipped on
    // request: a Picard solve whose p only enters gradients and the next
    // solve's start, A p being blind to a constant), then ghost copy
    if (normalize) {
        FCFD_TRY(field_mean(c, p));
        hipLaunchKernelGGL(k_sub_scalar, g, blk, 0, c->stream, p, L0, c->mg_scalars + SC_MEAN, 1.0 / ncell, DIM);
        FCFD_CHECK(hipGetLastError());
    }
    FCFD_TRY(bc_scalar(c, p));
    if (cycles) *cycles = kmax;
    if (rel) *rel = r;
    c->diag.ppe_cycles = kmax;
    c->diag.ppe_rel_resid = r;
    return rc;
}

int ppe_solve(fcfd_ctx *c, double rel_tol, int precision, int *cycles, double *rel, int normalize) {
    if (!c->mg_ready) FCFD_TRY(mg_setup(c));
    if (c->G.dim == 3) return ppe_solve_t<3>(c, rel_tol, precision, cycles, rel, normalize);
    return ppe_solve_t<2>(c, rel_tol, precision, cycles, rel, normalize);
}

}  // namespace fcfd
