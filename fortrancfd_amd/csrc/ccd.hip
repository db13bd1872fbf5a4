// ccd.hip — combined compact difference (CCD) derivatives along grid lines
// and the interface curvature.
//
// Reference: `Two phase flows/newOpenMP/3D+IBM/lib/solver/ccd_solvers.f90`
// (init_ccd :132-159, boundary closures :76-130, sources :287-306, driver
// :379-399) and `.../lib/solver/matrix_solver.f90` (TWIN_DEC :6-48,
// TWIN_BKS :50-78); surface_norms / curv in
// `.../src/interface_solver/others_normals.f90:1-83,133-167`.
//
// The 2x2 block-tridiagonal CCD matrix depends only on the line length and
// spacing, so it is factorised once per direction at context creation
// (the reference refactorises it for every line of every call).  A solve is
// then one forward and one backward sweep per line; one GPU thread owns one
// line.  For y and z lines consecutive threads hold consecutive i, so every
// sweep step is a coalesced row access and the kernel streams at HBM speed.
// x lines walk the contiguous axis, each lane in its own row: k_ccd_lines_x
// stages them through LDS tiles so that the wave's loads and stores are row
// segments (measured at 256^3: ~0.35 ms per x sweep, from ~0.88 ms with
// per-lane strided access; y or z ~0.19 ms).
#include <cmath>
#include <type_traits>
#include <vector>

#include "fcfd_internal.h"

using namespace fcfd;

namespace fcfd {

// Per-point coefficients of the line solve, packed for the two sweeps:
// forward [A1 B1 AA1 BB1], backward [A3 B3 AA3 BB3 A2/D AA2/D B2/D BB2/D]
// with D the determinant of the diagonal block (the reference divides by D;
// the reciprocal products are folded in here once).
enum { PF_A1, PF_B1, PF_AA1, PF_BB1, PB_A3, PB_B3, PB_AA3, PB_BB3, PB_A2D, PB_AA2D, PB_B2D, PB_BB2D, NPACK };

// line matrices of the CCD family (working_matrix rows 1..3 = sub, diag, super):
// the newOpenMP set on ghosted lines (n + 6 points) and the MPLS_RisingBubble
// set on interior lines (n points; its UCCD has the M-DRPCCD coefficients,
// `MPLS_RisingBubble/20190612/SCHEME_COMPACT.F90:189-200,482-490`)
enum { MAT_CCD = 0, MAT_UCCD_UP = 1, MAT_UCCD_DOWN = 2, MAT_M3_CCD = 3, MAT_M3_UP = 4, MAT_M3_DOWN = 5, NMAT = 6 };

// first-derivative source coefficients of the upwind system (the downwind
// one mirrors them) and the choice at u = 0: newOpenMP takes the downwind
// value there (ccd_solvers.f90:400-412), MPLS_RisingBubble the upwind one
// (SCHEME_COMPACT.F90:17-25)
struct UccdSrc {
    double c1, c2, c3;
    int ge;
};
constexpr UccdSrc UCCD_N3{-1.93596119008109, 1.99692238016218, -0.06096119008109, 0};
constexpr UccdSrc UCCD_M3{-1.93629468537011, 1.99758937102742, -0.061294685370110, 1};

// Matrix of one kind on a line of N points, spacing dx: interior rows of
// init_ccd / init_uccd and the closure rows of working_matrix_bc
// (`.../lib/solver/ccd_solvers.f90:76-210`).  Row i: eq. 1 (first
// derivative) A(r, i) f'(i-2+r) + B(r, i) f''(i-2+r), eq. 2 (second
// derivative) AA(r, i) f'(.) + BB(r, i) f''(.), r = 1 sub, 2 diagonal, 3 super.
struct LineMat {
    int N;
    std::vector<double> A, B, AA, BB;
    double &M(std::vector<double> &X, int r, int i) { return X[i * 3 + r - 1]; }
};

static LineMat ccd_matrix(int kind, int N, double dx) {
    LineMat L;
    L.N = N;
    L.A.assign(3 * N, 0.0); L.B.assign(3 * N, 0.0); L.AA.assign(3 * N, 0.0); L.BB.assign(3 * N, 0.0);
    auto M = [&](std::vector<double> &X, int r, int i) -> double & { return X[i * 3 + r - 1]; };
    std::vector<double> &A = L.A, &B = L.B, &AA = L.AA, &BB = L.BB;
    const int is = 0, ie = N - 1;
    const bool m3 = kind >= MAT_M3_CCD;
    if (m3) kind -= MAT_M3_CCD;
    const double A1 = 0.875;
    const double B1 = m3 ? 0.125100442885824 : 0.1251282341599089;
    const double B2 = m3 ? -0.248995571141759 : -0.2487176584009104;
    const double B3 = m3 ? 0.0001004428858241 : 0.0001282341599089;
    for (int i = is + 1; i <= ie - 1; ++i) {
        if (kind == MAT_CCD) {
            M(A, 1, i) = 7.0 / 16.0; M(A, 2, i) = 1.0; M(A, 3, i) = 7.0 / 16.0;
            M(B, 1, i) = dx / 16.0; M(B, 2, i) = 0.0; M(B, 3, i) = -dx / 16.0;
        } else if (kind == MAT_UCCD_UP) {
            M(A, 1, i) = A1; M(A, 2, i) = 1.0; M(A, 3, i) = 0.0;
            M(B, 1, i) = B1 * dx; M(B, 2, i) = B2 * dx; M(B, 3, i) = B3 * dx;
        } else {
            M(A, 1, i) = 0.0; M(A, 2, i) = 1.0; M(A, 3, i) = A1;
            M(B, 1, i) = -B3 * dx; M(B, 2, i) = -B2 * dx; M(B, 3, i) = -B1 * dx;
        }
        M(AA, 1, i) = -9.0 / 8.0 / dx; M(AA, 2, i) = 0.0; M(AA, 3, i) = 9.0 / 8.0 / dx;
        M(BB, 1, i) = -1.0 / 8.0; M(BB, 2, i) = 1.0; M(BB, 3, i) = -1.0 / 8.0;
    }
    M(A, 2, is) = 1.0; M(B, 2, is) = 0.0; M(AA, 2, is) = 0.0; M(BB, 2, is) = 1.0;
    M(A, 3, is) = 2.0; M(B, 3, is) = -dx; M(AA, 3, is) = -2.5 / dx; M(BB, 3, is) = 8.5;
    M(A, 2, ie) = 1.0; M(B, 2, ie) = 0.0; M(AA, 2, ie) = 0.0; M(BB, 2, ie) = 1.0;
    M(A, 1, ie) = 2.0; M(B, 1, ie) = dx; M(AA, 1, ie) = 2.5 / dx; M(BB, 1, ie) = 8.5;
    M(A, 1, is) = 0.0; M(AA, 1, is) = 0.0; M(A, 3, ie) = 0.0; M(AA, 3, ie) = 0.0;
    M(B, 1, is) = 0.0; M(BB, 1, is) = 0.0; M(B, 3, ie) = 0.0; M(BB, 3, ie) = 0.0;
    return L;
}

// first point of segment s of a line of N points cut into P segments whose
// lengths differ by at most one
__host__ __device__ inline int seg_start(int s, int N, int P) { return (int)((unsigned)s * (unsigned)N / (unsigned)P); }

// TWIN_DEC (`.../lib/solver/matrix_solver.f90:6-48`) in place, packed per
// point.  With P > 1 the elimination restarts at the first point of every
// segment: the factors are those of each segment's own system, and the
// forward coefficients of a segment's first row keep the row's coupling to
// the previous point (unused by a sweep that starts from zero, applied to the
// known neighbour value by one that does not).
static void ccd_pack(LineMat L, int P, std::vector<double> &pk) {
    const int N = L.N;
    auto M = [&](std::vector<double> &X, int r, int i) -> double & { return X[i * 3 + r - 1]; };
    std::vector<double> &A = L.A, &B = L.B, &AA = L.AA, &BB = L.BB;
    std::vector<char> start(N, 0);
    for (int s = 0; s < P; ++s) start[seg_start(s, N, P)] = 1;
    for (int i = 1; i < N; ++i) {
        if (start[i]) continue;
        double Den = M(A, 2, i - 1) * M(BB, 2, i - 1) - M(AA, 2, i - 1) * M(B, 2, i - 1);
        double SC1 = -(M(AA, 2, i - 1) * M(B, 1, i) - M(A, 1, i) * M(BB, 2, i - 1)) / Den;
        double SC2 = (M(B, 1, i) * M(A, 2, i - 1) - M(B, 2, i - 1) * M(A, 1, i)) / Den;
        M(A, 1, i) = SC1;
        M(A, 2, i) = M(A, 2, i) - (SC1 * M(A, 3, i - 1) + SC2 * M(AA, 3, i - 1));
        M(B, 1, i) = SC2;
        M(B, 2, i) = M(B, 2, i) - (SC1 * M(B, 3, i - 1) + SC2 * M(BB, 3, i - 1));
        SC1 = -(M(AA, 2, i - 1) * M(BB, 1, i) - M(AA, 1, i) * M(BB, 2, i - 1)) / Den;
        SC2 = (M(BB, 1, i) * M(A, 2, i - 1) - M(B, 2, i - 1) * M(AA, 1, i)) / Den;
        M(AA, 1, i) = SC1;
        M(AA, 2, i) = M(AA, 2, i) - (SC1 * M(A, 3, i - 1) + SC2 * M(AA, 3, i - 1));
        M(BB, 1, i) = SC2;
        M(BB, 2, i) = M(BB, 2, i) - (SC1 * M(B, 3, i - 1) + SC2 * M(BB, 3, i - 1));
    }
    pk.assign((size_t)NPACK * N, 0.0);
    for (int i = 0; i < N; ++i) {
        double *q = &pk[(size_t)i * NPACK];
        const double D = M(A, 2, i) * M(BB, 2, i) - M(AA, 2, i) * M(B, 2, i);
        q[PF_A1] = M(A, 1, i); q[PF_B1] = M(B, 1, i); q[PF_AA1] = M(AA, 1, i); q[PF_BB1] = M(BB, 1, i);
        q[PB_A3] = M(A, 3, i); q[PB_B3] = M(B, 3, i); q[PB_AA3] = M(AA, 3, i); q[PB_BB3] = M(BB, 3, i);
        q[PB_A2D] = M(A, 2, i) / D; q[PB_AA2D] = M(AA, 2, i) / D;
        q[PB_B2D] = M(B, 2, i) / D; q[PB_BB2D] = M(BB, 2, i) / D;
    }
}

static void ccd_factor(int kind, int N, double dx, std::vector<double> &pk) {
    ccd_pack(ccd_matrix(kind, N, dx), 1, pk);
}

// ---- partitioned line solve ---------------------------------------------
// One thread per line leaves a sweep with only as many threads as lines,
// too few to cover the memory latency of the chip (140 x 140 z lines are 77
// workgroups of 256).  The line of N points is cut into P segments
// of at most SEG points.  With x0 the solution of a segment's own system
// (its rows without their couplings to the neighbouring segments), the
// line's solution on segment [a, b] is
//   x(i) = x0(i) - V(i) x(a-1) - W(i) x(b+1),
// V, W the 2 x 2 "spikes" (the segment's inverse times its coupling blocks),
// fixed by the grid.  The values at the P - 1 interfaces, u_s = (x(b_s),
// x(a_{s+1})), solve a block-tridiagonal system of 4 x 4 blocks, also fixed
// by the grid, whose block LU is formed here once.  The solve is then
//   k_seg_lines<.., false>: per segment, x0 at its two ends;
//   k_seg_reduce:           per line, the reduced forward / backward sweeps;
//   k_seg_lines<.., true>:  per segment, the segment's own system once more
//                           with x(a-1) and x(b+1) known, written out.
// Every line has P threads in each sweep and f is read twice.
constexpr int SEG = 16;
// reduced block LU per interface: multiplier m (4 x 4), inverse of the
// modified diagonal Dinv (4 x 4), super block U (4 x 4), row-major
constexpr int NRED = 48;

// the local solve of segment [a, b] (starting from zero neighbours) for a
// right side given per point (r1, r2), with the packed segment factors
static void seg_solve_host(const std::vector<double> &lt, int a, int b, const double *r, double *x) {
    const int n = b - a + 1;
    std::vector<double> S(n), SS(n);
    double s = 0.0, ss = 0.0;
    for (int q = 0; q < n; ++q) {
        const double *c = &lt[(size_t)(a + q) * NPACK];
        const double Sn = r[2 * q] - (c[PF_A1] * s + c[PF_B1] * ss);
        const double SSn = r[2 * q + 1] - (c[PF_AA1] * s + c[PF_BB1] * ss);
        s = Sn; ss = SSn;
        S[q] = s; SS[q] = ss;
    }
    double xs = 0.0, xss = 0.0;
    for (int q = n - 1; q >= 0; --q) {
        const double *c = &lt[(size_t)(a + q) * NPACK];
        const double Si = S[q] - (c[PB_A3] * xs + c[PB_B3] * xss);
        const double SSi = SS[q] - (c[PB_AA3] * xs + c[PB_BB3] * xss);
        xs = c[PB_BB2D] * Si - c[PB_B2D] * SSi;
        xss = c[PB_A2D] * SSi - c[PB_AA2D] * Si;
        x[2 * q] = xs; x[2 * q + 1] = xss;
    }
}

// 4 x 4 inverse by Gauss-Jordan with partial pivoting
static void inv4(const double *m, double *out) {
    double a[4][8];
    for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 8; ++c) a[r][c] = c < 4 ? m[r * 4 + c] : (c - 4 == r ? 1.0 : 0.0);
    for (int c = 0; c < 4; ++c) {
        int p = c;
        for (int r = c + 1; r < 4; ++r)
            if (fabs(a[r][c]) > fabs(a[p][c])) p = r;
        for (int k = 0; k < 8; ++k) std::swap(a[c][k], a[p][k]);
        const double d = a[c][c];
        for (int k = 0; k < 8; ++k) a[c][k] /= d;
        for (int r = 0; r < 4; ++r)
            if (r != c) {
                const double f = a[r][c];
                for (int k = 0; k < 8; ++k) a[r][k] -= f * a[c][k];
            }
    }
    for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 4; ++c) out[r * 4 + c] = a[r][c + 4];
}

static void mul4(const double *a, const double *b, double *c) {
    for (int r = 0; r < 4; ++r)
        for (int k = 0; k < 4; ++k) {
            double s = 0.0;
            for (int j = 0; j < 4; ++j) s += a[r * 4 + j] * b[j * 4 + k];
            c[r * 4 + k] = s;
        }
}

static int seg_count(int N) { return (N + SEG - 1) / SEG; }

// segment factors, reduced block LU and segment end functionals of one line
// matrix, in one buffer: [N x NPACK segment factors][(P - 1) x NRED reduced
// LU][P x 4 x (SEG + 2) end functionals]
static void ccd_seg_tables(int kind, int N, double dx, std::vector<double> &out) {
    const LineMat M0 = ccd_matrix(kind, N, dx);
    const int P = seg_count(N);
    std::vector<double> lt;
    ccd_pack(M0, P, lt);
    // spikes per point: V (2 x 2) then W (2 x 2), rows (f', f''), columns
    // the (f', f'') of x(a-1) / x(b+1)
    std::vector<double> sp((size_t)N * 8, 0.0);
    auto Mv = [&](const std::vector<double> &X, int r, int i) { return X[i * 3 + r - 1]; };
    std::vector<double> r(2 * SEG), x(2 * SEG);
    for (int s = 0; s < P; ++s) {
        const int a = seg_start(s, N, P), b = seg_start(s + 1, N, P) - 1, n = b - a + 1;
        for (int col = 0; col < 2; ++col) {
            if (a > 0) {   // V column: coupling of row a to x(a-1)
                std::fill(r.begin(), r.end(), 0.0);
                r[0] = col == 0 ? Mv(M0.A, 1, a) : Mv(M0.B, 1, a);
                r[1] = col == 0 ? Mv(M0.AA, 1, a) : Mv(M0.BB, 1, a);
                seg_solve_host(lt, a, b, r.data(), x.data());
                for (int q = 0; q < n; ++q) {
                    sp[(size_t)(a + q) * 8 + 0 + col] = x[2 * q];
                    sp[(size_t)(a + q) * 8 + 2 + col] = x[2 * q + 1];
                }
            }
            if (b < N - 1) {   // W column: coupling of row b to x(b+1)
                std::fill(r.begin(), r.end(), 0.0);
                r[2 * (n - 1)] = col == 0 ? Mv(M0.A, 3, b) : Mv(M0.B, 3, b);
                r[2 * (n - 1) + 1] = col == 0 ? Mv(M0.AA, 3, b) : Mv(M0.BB, 3, b);
                seg_solve_host(lt, a, b, r.data(), x.data());
                for (int q = 0; q < n; ++q) {
                    sp[(size_t)(a + q) * 8 + 4 + col] = x[2 * q];
                    sp[(size_t)(a + q) * 8 + 6 + col] = x[2 * q + 1];
                }
            }
        }
    }
    // reduced rows of interface s (unknowns u_s = (x(b_s), x(a_{s+1}))):
    //   x(b_s)     + V_s(b_s) x(b_{s-1})     + W_s(b_s) x(a_{s+1})         = x0(b_s)
    //   x(a_{s+1}) + V_{s+1}(a_{s+1}) x(b_s) + W_{s+1}(a_{s+1}) x(a_{s+2}) = x0(a_{s+1})
    const int NI = P - 1;
    std::vector<double> red((size_t)NI * NRED, 0.0);
    double Dinv_prev[16] = {0}, Uprev[16] = {0};
    for (int s = 0; s < NI; ++s) {
        const int an = seg_start(s + 1, N, P), bs = an - 1;
        const double *Vb = &sp[(size_t)bs * 8], *Wb = Vb + 4;
        const double *Va = &sp[(size_t)an * 8], *Wa = Va + 4;
        double D[16] = {0}, L[16] = {0}, U[16] = {0};
        for (int k = 0; k < 4; ++k) D[k * 4 + k] = 1.0;
        D[0 * 4 + 2] = Wb[0]; D[0 * 4 + 3] = Wb[1]; D[1 * 4 + 2] = Wb[2]; D[1 * 4 + 3] = Wb[3];
        D[2 * 4 + 0] = Va[0]; D[2 * 4 + 1] = Va[1]; D[3 * 4 + 0] = Va[2]; D[3 * 4 + 1] = Va[3];
        if (s > 0) { L[0] = Vb[0]; L[1] = Vb[1]; L[4] = Vb[2]; L[5] = Vb[3]; }
        if (s < NI - 1) { U[2 * 4 + 2] = Wa[0]; U[2 * 4 + 3] = Wa[1]; U[3 * 4 + 2] = Wa[2]; U[3 * 4 + 3] = Wa[3]; }
        double m[16] = {0}, Dm[16];
        if (s > 0) {
            mul4(L, Dinv_prev, m);
            double mU[16];
            mul4(m, Uprev, mU);
            for (int k = 0; k < 16; ++k) D[k] -= mU[k];
        }
        inv4(D, Dm);
        double *o = &red[(size_t)s * NRED];
        for (int k = 0; k < 16; ++k) { o[k] = m[k]; o[16 + k] = Dm[k]; o[32 + k] = U[k]; }
        std::copy(Dm, Dm + 16, Dinv_prev);
        std::copy(U, U + 16, Uprev);
    }
    // end functionals: x0(a) and x0(b) of segment s as weights on f(a - 1 ..
    // a + n) (the sources are linear in f), E[(s * 4 + c) * (SEG + 2) + k],
    // c = 0, 1 the (f', f'') at a, 2, 3 at b; found by solving for unit f
    std::vector<double> E((size_t)P * 4 * (SEG + 2), 0.0);
    {
        const double idx = 1.0 / dx, idx2 = idx * idx;
        const int form = kind % 3;   // 0 CCD, 1 up-wind, 2 down-wind
        const UccdSrc cs = kind >= MAT_M3_CCD ? UCCD_M3 : UCCD_N3;
        for (int s = 0; s < P; ++s) {
            const int a = seg_start(s, N, P), b = seg_start(s + 1, N, P) - 1, n = b - a + 1;
            for (int k = 0; k <= n + 1; ++k) {
                const int fi = a - 1 + k;
                if (fi < 0 || fi >= N) continue;
                auto F = [&](int j) { return j == fi ? 1.0 : 0.0; };
                for (int q = 0; q < n; ++q) {
                    const int i = a + q;
                    double r1, r2;
                    if (i == 0) {
                        r1 = (-3.5 * F(0) + 4.0 * F(1) - 0.5 * F(2)) * idx;
                        r2 = (34.0 / 3.0 * F(0) - 83.0 / 4.0 * F(1) + 10.0 * F(2) - 7.0 / 12.0 * F(3)) * idx2;
                    } else if (i == N - 1) {
                        const double e0 = F(N - 1), e1 = F(N - 2), e2 = F(N - 3), e3 = F(N - 4);
                        r1 = -(-3.5 * e0 + 4.0 * e1 - 0.5 * e2) * idx;
                        r2 = (34.0 / 3.0 * e0 - 83.0 / 4.0 * e1 + 10.0 * e2 - 7.0 / 12.0 * e3) * idx2;
                    } else {
                        const double fm = F(i - 1), fc = F(i), fp = F(i + 1);
                        r2 = (3.0 * fm - 6.0 * fc + 3.0 * fp) * idx2;
                        r1 = form == 0 ? 15.0 / 16.0 * (fp - fm) * idx
                                       : (form == 1 ? (cs.c1 * fm + cs.c2 * fc + cs.c3 * fp) * idx
                                                    : -(cs.c3 * fm + cs.c2 * fc + cs.c1 * fp) * idx);
                    }
                    r[2 * q] = r1;
                    r[2 * q + 1] = r2;
                }
                seg_solve_host(lt, a, b, r.data(), x.data());
                double *e = &E[(size_t)s * 4 * (SEG + 2) + k];
                e[0] = x[0];
                e[SEG + 2] = x[1];
                e[2 * (SEG + 2)] = x[2 * (n - 1)];
                e[3 * (SEG + 2)] = x[2 * (n - 1) + 1];
            }
        }
    }
    out = lt;
    out.insert(out.end(), red.begin(), red.end());
    out.insert(out.end(), E.begin(), E.end());
}

int ccd_setup(fcfd_ctx *c) {
    const Geom &G = c->G;
    const int n[3] = {G.nx, G.ny, G.nz};
    const double h[3] = {G.dx, G.dy, G.dz};
    const long nl[3] = {(long)G.ny * G.nz, (long)G.nx * G.nz, (long)G.nx * G.ny};
    long zmax = 0;
    for (int kind = 0; kind < NMAT; ++kind)
        for (int d = 0; d < 3; ++d) {
            double *&dst = c->ccd_coef[kind * 3 + d];
            double *&seg = c->ccd_seg[kind * 3 + d];
            dst = seg = nullptr;
            if (d >= G.dim) continue;
            const int N = n[d] + (kind >= MAT_M3_CCD ? 0 : 2 * G.g);
            std::vector<double> pk;
            ccd_factor(kind, N, h[d], pk);
            FCFD_CHECK(hipMalloc((void **)&dst, sizeof(double) * pk.size()));
            FCFD_CHECK(hipMemcpy(dst, pk.data(), sizeof(double) * pk.size(), hipMemcpyHostToDevice));
            // the partitioned solve numbers lines and segment points in 32 bits
            // (one launch dimension counts the groups of 64 x lines or 256 y / z lines)
            if (nl[d] >= (1L << 31) || seg_count(N) > 65535 || (nl[d] + (d == 0 ? 63 : 255)) / (d == 0 ? 64 : 256) > 65535)
                continue;
            ccd_seg_tables(kind, N, h[d], pk);
            FCFD_CHECK(hipMalloc((void **)&seg, sizeof(double) * pk.size()));
            FCFD_CHECK(hipMemcpy(seg, pk.data(), sizeof(double) * pk.size(), hipMemcpyHostToDevice));
            zmax = std::max(zmax, 2L * seg_count(N) * 4 * nl[d]);
        }
    c->ccd_z = nullptr;
    if (zmax) FCFD_CHECK(hipMalloc((void **)&c->ccd_z, sizeof(double) * zmax));
    return FCFD_OK;
}

// asgn_src_ccd + TWIN_BKS along one line; ghost line ends included.  One
// thread per line.  The packed coefficient table of the direction is staged
// in LDS once per workgroup (every line reads the same row at the same step,
// so the reads are LDS broadcasts); a line too long for the LDS table
// (in_lds = 0) reads the same rows from global memory, where they stay in
// the caches for the same reason.  Each sweep moves through the line in
// chunks of CH points: the loads of chunk c+1 are issued before chunk c's
// dependent recurrence runs, so memory latency overlaps the arithmetic.
constexpr int CCD_THREADS = 256;
constexpr long CCD_LDS_MAX = 64 * 1024;

template <int DIR, int CH>
__global__ void __launch_bounds__(CCD_THREADS) k_ccd_lines(const double *__restrict__ f, double *__restrict__ fx,
                                                           double *__restrict__ fxx, const double *__restrict__ cog,
                                                           Geom G, int in_lds, int ext) {
    extern __shared__ double lds_co[];
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    int a, b, N;
    long base, st, nlines;
    double dx;
    if (DIR == 0) {
        nlines = (long)G.ny * G.nz; N = G.nx + 2 * ext; st = 1; dx = G.dx;
    } else if (DIR == 1) {
        nlines = (long)G.nx * G.nz; N = G.ny + 2 * ext; st = G.sx; dx = G.dy;
    } else {
        nlines = (long)G.nx * G.ny; N = G.nz + 2 * ext; st = G.sxy; dx = G.dz;
    }
    const double *co = cog;
    if (in_lds) {
        for (int e = threadIdx.x; e < N * NPACK; e += CCD_THREADS) lds_co[e] = cog[e];
        __syncthreads();
        co = lds_co;
    }
    if (t >= nlines) return;
    if (DIR == 0) {
        a = (int)(t % G.ny); b = (int)(t / G.ny);
        base = G.id(-ext, a, b);
    } else if (DIR == 1) {
        a = (int)(t % G.nx); b = (int)(t / G.nx);
        base = G.id(a, -ext, b);
    } else {
        a = (int)(t % G.nx); b = (int)(t / G.nx);
        base = G.id(a, b, -ext);
    }
    const double *F = f + base;
    double *X = fx + base, *XX = fxx + base;
    const double idx = 1.0 / dx, idx2 = idx * idx;
    // forward elimination with the source of asgn_src_ccd
    const double f0 = F[0], f1 = F[st], f2 = F[2 * st], f3 = F[3 * st];
    double S = (-3.5 * f0 + 4.0 * f1 - 0.5 * f2) * idx;
    double SS = (34.0 / 3.0 * f0 - 83.0 / 4.0 * f1 + 10.0 * f2 - 7.0 / 12.0 * f3) * idx2;
    X[0] = S;
    XX[0] = SS;
    double fm = f0, fc = f1;  // f(i-1), f(i)
    // interior points i = 1 .. N-2 need f(i+1): chunk c covers i0 = 1 + c*CH
    double fp[CH], fq[CH];
#pragma unroll
    for (int u = 0; u < CH; ++u)
        if (1 + u < N - 1) fp[u] = F[(long)(2 + u) * st];
    for (int i0 = 1; i0 < N - 1; i0 += CH) {
        const int n0 = i0 + CH;
#pragma unroll
        for (int u = 0; u < CH; ++u)
            if (n0 + u < N - 1) fq[u] = F[(long)(n0 + u + 1) * st];
#pragma unroll
        for (int u = 0; u < CH; ++u) {
            const int i = i0 + u;
            if (i < N - 1) {
                const double *q = co + i * NPACK;
                const double s = 15.0 / 16.0 * (fp[u] - fm) * idx;
                const double ss = (3.0 * fm - 6.0 * fc + 3.0 * fp[u]) * idx2;
                const double Sn = s - (q[PF_A1] * S + q[PF_B1] * SS);
                const double SSn = ss - (q[PF_AA1] * S + q[PF_BB1] * SS);
                S = Sn; SS = SSn;
                X[(long)i * st] = S;
                XX[(long)i * st] = SS;
                fm = fc; fc = fp[u];
            }
        }
#pragma unroll
        for (int u = 0; u < CH; ++u) fp[u] = fq[u];
    }
    const int ie = N - 1;
    const double e0 = F[(long)ie * st], e1 = F[(long)(ie - 1) * st], e2 = F[(long)(ie - 2) * st],
                 e3 = F[(long)(ie - 3) * st];
    {
        const double *q = co + ie * NPACK;
        const double s = -(-3.5 * e0 + 4.0 * e1 - 0.5 * e2) * idx;
        const double ss = (34.0 / 3.0 * e0 - 83.0 / 4.0 * e1 + 10.0 * e2 - 7.0 / 12.0 * e3) * idx2;
        const double Sn = s - (q[PF_A1] * S + q[PF_B1] * SS);
        const double SSn = ss - (q[PF_AA1] * S + q[PF_BB1] * SS);
        S = q[PB_BB2D] * Sn - q[PB_B2D] * SSn;
        SS = q[PB_A2D] * SSn - q[PB_AA2D] * Sn;
        X[(long)ie * st] = S;
        XX[(long)ie * st] = SS;
    }
    // back substitution, i = ie-1 .. 0 in chunks; the forward values of the
    // chunk are read before any of them is overwritten
    double xv[CH], xxv[CH], yv[CH], yyv[CH];
#pragma unroll
    for (int u = 0; u < CH; ++u)
        if (ie - 1 - u >= 0) {
            xv[u] = X[(long)(ie - 1 - u) * st];
            xxv[u] = XX[(long)(ie - 1 - u) * st];
        }
    for (int i1 = ie - 1; i1 >= 0; i1 -= CH) {
        const int n1 = i1 - CH;
#pragma unroll
        for (int u = 0; u < CH; ++u)
            if (n1 - u >= 0) {
                yv[u] = X[(long)(n1 - u) * st];
                yyv[u] = XX[(long)(n1 - u) * st];
            }
#pragma unroll
        for (int u = 0; u < CH; ++u) {
            const int i = i1 - u;
            if (i >= 0) {
                const double *q = co + i * NPACK;
                const double Si = xv[u] - (q[PB_A3] * S + q[PB_B3] * SS);
                const double SSi = xxv[u] - (q[PB_AA3] * S + q[PB_BB3] * SS);
                S = q[PB_BB2D] * Si - q[PB_B2D] * SSi;
                SS = q[PB_A2D] * SSi - q[PB_AA2D] * Si;
                X[(long)i * st] = S;
                XX[(long)i * st] = SS;
            }
        }
#pragma unroll
        for (int u = 0; u < CH; ++u) {
            xv[u] = yv[u];
            xxv[u] = yyv[u];
        }
    }
}

// x lines: the same two sweeps, with every global access of a wave made
// coalesced.  A block is one wave owning 64 consecutive lines; the sweeps
// move along the lines in CX-point chunks staged through LDS tiles (one
// row per line, padded to CX+1 words against bank conflicts): the wave
// loads and stores a tile as 128-byte row segments, then each lane runs its
// line's recurrence over its tile row.  The next chunk is prefetched into
// registers while the current one is swept.  The coefficient row of step i
// is the same for every lane and is read through the scalar cache.
constexpr int CCDX_CH = 16, CCDX_TS = CCDX_CH + 1, CCDX_LINES = 64;
static_assert(CCDX_LINES == WAVE && CCDX_LINES % CCDX_CH == 0, "one wave per block, whole tile rows per access");

__global__ void __launch_bounds__(CCDX_LINES) k_ccd_lines_x(const double *__restrict__ f, double *__restrict__ fx,
                                                            double *__restrict__ fxx, const double *__restrict__ co,
                                                            Geom G, int ext) {
    __shared__ double tF[CCDX_LINES * CCDX_TS], tX[CCDX_LINES * CCDX_TS], tXX[CCDX_LINES * CCDX_TS];
    const int lane = threadIdx.x;
    const long nlines = (long)G.ny * G.nz, L0 = (long)blockIdx.x * CCDX_LINES, t = L0 + lane;
    const int N = G.nx + 2 * ext;
    const bool valid = t < nlines;
    long mybase = 0;
    if (valid) mybase = G.id(-ext, (int)(t % G.ny), (int)(t / G.ny));
    // tile element e = r*64 + lane of a cooperative access: line l = e / CX, point u = e % CX
    constexpr int RPI = CCDX_LINES / CCDX_CH;   // lines per wave instruction
    const int lsub = lane / CCDX_CH, usub = lane % CCDX_CH;
    long rbase[CCDX_CH];
    bool rval[CCDX_CH];
#pragma unroll
    for (int r = 0; r < CCDX_CH; ++r) {
        const int l = r * RPI + lsub;
        rbase[r] = __shfl(mybase, l, CCDX_LINES);
        rval[r] = L0 + l < nlines;
    }
    const double idx = 1.0 / G.dx, idx2 = idx * idx;
    const double *F = f + mybase;
    double *X = fx + mybase, *XX = fxx + mybase;
    // ---- forward sweep (asgn_src_ccd sources, elimination)
    double S = 0.0, SS = 0.0, fm = 0.0, fc = 0.0;
    if (valid) {
        const double f0 = F[0], f1 = F[1], f2 = F[2], f3 = F[3];
        S = (-3.5 * f0 + 4.0 * f1 - 0.5 * f2) * idx;
        SS = (34.0 / 3.0 * f0 - 83.0 / 4.0 * f1 + 10.0 * f2 - 7.0 / 12.0 * f3) * idx2;
        X[0] = S;
        XX[0] = SS;
        fm = f0;
        fc = f1;
    }
    // chunk c holds f(j0 .. j0+CX-1), j0 = 2 + c*CX: the f(i+1) of steps i = j0-1+u
    double pf[CCDX_CH];
    auto load_f = [&](int j0) {
#pragma unroll
        for (int r = 0; r < CCDX_CH; ++r)
            if (rval[r] && j0 + usub < N) pf[r] = f[rbase[r] + j0 + usub];
    };
    auto put = [&](double *tile, const double *v) {
#pragma unroll
        for (int r = 0; r < CCDX_CH; ++r) tile[(r * RPI + lsub) * CCDX_TS + usub] = v[r];
    };
    load_f(2);
    for (int j0 = 2; j0 < N; j0 += CCDX_CH) {
        put(tF, pf);
        __syncthreads();
        if (j0 + CCDX_CH < N) load_f(j0 + CCDX_CH);
        if (valid) {
            const double *row = tF + lane * CCDX_TS;
#pragma unroll
            for (int u = 0; u < CCDX_CH; ++u) {
                const int i = j0 - 1 + u;
                if (i <= N - 2) {
                    const double *q = co + i * NPACK;
                    const double fp = row[u];
                    const double s = 15.0 / 16.0 * (fp - fm) * idx;
                    const double ss = (3.0 * fm - 6.0 * fc + 3.0 * fp) * idx2;
                    const double Sn = s - (q[PF_A1] * S + q[PF_B1] * SS);
                    const double SSn = ss - (q[PF_AA1] * S + q[PF_BB1] * SS);
                    S = Sn; SS = SSn;
                    tX[lane * CCDX_TS + u] = S;
                    tXX[lane * CCDX_TS + u] = SS;
                    fm = fc; fc = fp;
                }
            }
        }
        __syncthreads();
#pragma unroll
        for (int r = 0; r < CCDX_CH; ++r) {
            const int l = r * RPI + lsub, i = j0 - 1 + usub;
            if (rval[r] && i <= N - 2) {
                fx[rbase[r] + i] = tX[l * CCDX_TS + usub];
                fxx[rbase[r] + i] = tXX[l * CCDX_TS + usub];
            }
        }
        __syncthreads();
    }
    const int ie = N - 1;
    if (valid) {
        const double e0 = F[ie], e1 = F[ie - 1], e2 = F[ie - 2], e3 = F[ie - 3];
        const double *q = co + ie * NPACK;
        const double s = -(-3.5 * e0 + 4.0 * e1 - 0.5 * e2) * idx;
        const double ss = (34.0 / 3.0 * e0 - 83.0 / 4.0 * e1 + 10.0 * e2 - 7.0 / 12.0 * e3) * idx2;
        const double Sn = s - (q[PF_A1] * S + q[PF_B1] * SS);
        const double SSn = ss - (q[PF_AA1] * S + q[PF_BB1] * SS);
        S = q[PB_BB2D] * Sn - q[PB_B2D] * SSn;
        SS = q[PB_A2D] * SSn - q[PB_AA2D] * Sn;
        X[ie] = S;
        XX[ie] = SS;
    }
    // ---- backward sweep over i = N-2 .. 1 by the same chunks in reverse,
    // then i = 0; the forward values of a chunk are read before it is rewritten
    const int nch = (N - 2 + CCDX_CH - 1) / CCDX_CH;   // chunks with j0 = 2 + c*CX < N
    double px[CCDX_CH], pxx[CCDX_CH];
    auto load_x = [&](int j0) {
#pragma unroll
        for (int r = 0; r < CCDX_CH; ++r) {
            const int i = j0 - 1 + usub;
            if (rval[r] && i <= N - 2) {
                px[r] = fx[rbase[r] + i];
                pxx[r] = fxx[rbase[r] + i];
            }
        }
    };
    __syncthreads();   // the forward stores of this wave are complete before they are read back
    load_x(2 + (nch - 1) * CCDX_CH);
    for (int c = nch - 1; c >= 0; --c) {
        const int j0 = 2 + c * CCDX_CH;
        put(tX, px);
        put(tXX, pxx);
        __syncthreads();
        if (c > 0) load_x(j0 - CCDX_CH);
        if (valid) {
            double *rx = tX + lane * CCDX_TS, *rxx = tXX + lane * CCDX_TS;
#pragma unroll
            for (int u = CCDX_CH - 1; u >= 0; --u) {
                const int i = j0 - 1 + u;
                if (i <= N - 2) {
                    const double *q = co + i * NPACK;
                    const double Si = rx[u] - (q[PB_A3] * S + q[PB_B3] * SS);
                    const double SSi = rxx[u] - (q[PB_AA3] * S + q[PB_BB3] * SS);
                    S = q[PB_BB2D] * Si - q[PB_B2D] * SSi;
                    SS = q[PB_A2D] * SSi - q[PB_AA2D] * Si;
                    rx[u] = S;
                    rxx[u] = SS;
                }
            }
        }
        __syncthreads();
#pragma unroll
        for (int r = 0; r < CCDX_CH; ++r) {
            const int l = r * RPI + lsub, i = j0 - 1 + usub;
            if (rval[r] && i <= N - 2) {
                fx[rbase[r] + i] = tX[l * CCDX_TS + usub];
                fxx[rbase[r] + i] = tXX[l * CCDX_TS + usub];
            }
        }
        __syncthreads();
    }
    if (valid) {
        const double *q = co;
        const double Si = X[0] - (q[PB_A3] * S + q[PB_B3] * SS);
        const double SSi = XX[0] - (q[PB_AA3] * S + q[PB_BB3] * SS);
        X[0] = q[PB_BB2D] * Si - q[PB_B2D] * SSi;
        XX[0] = q[PB_A2D] * SSi - q[PB_AA2D] * Si;
    }
}

// Partitioned line solve kernels.  zb holds per system y, segment s and line
// t four values at zb[((y * P + s) * 4 + c) * nl + t]: c = 0, 1 the
// (f', f'') of the segment's first point, c = 2, 3 of its last: the ends of
// the segment's own solution x0 (k_seg_ends*), then, after k_seg_reduce*,
// the line's solution at the interfaces.  NS = 1: CCD (asgn_src_ccd); NS =
// 2: the up- and down-wind UCCD systems side by side (root_solve "uccd").
// For y and z lines a thread is one (line, segment) and the wave's 64 lanes
// are 64 consecutive lines (consecutive i: coalesced row accesses) of one
// segment, so every coefficient read is uniform over the wave.
constexpr int SEGT = 256;

// (segment, first line) of a y / z block: as for x lines (segx_block) the
// launch is (P, line groups), so the P segments of a group, whose f reads
// overlap by one point at each end, run consecutively on one XCD
__device__ __forceinline__ void segyz_block(int &s, long &t) {
    int bx, by, bz;
    xcd_block(bx, by, bz);
    s = bx;
    t = (long)by * SEGT + threadIdx.x;
}

// line geometry of the y (DIR 1) or z (DIR 2) lines
template <int DIR>
__device__ __forceinline__ void seg_line_geom(const Geom &G, int ext, long &nl, int &N, long &st, double &dx) {
    if constexpr (DIR == 1) { nl = (long)G.nx * G.nz; N = G.ny + 2 * ext; st = G.sx; dx = G.dy; }
    else { nl = (long)G.nx * G.ny; N = G.nz + 2 * ext; st = G.sxy; dx = G.dz; }
}

// x0(a), x0(b) of every system from the end functionals E (uniform over the
// wave: one segment), fv[k] = f(a - 1 + k)
template <int NS>
__device__ __forceinline__ void seg_ends_store(const double (&fv)[SEG + 2], const double *end_u,
                                               const double *end_d, double *__restrict__ zb, long t, long nl, int s,
                                               int P) {
#pragma unroll
    for (int y = 0; y < NS; ++y) {
        const double *E = (y ? end_d : end_u) + (long)s * 4 * (SEG + 2);
        double *Z = zb + (long)y * P * 4 * nl + t;
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            double acc = 0.0;
#pragma unroll
            for (int k = 0; k < SEG + 2; ++k) acc = fma(E[c * (SEG + 2) + k], fv[k], acc);
            Z[((long)s * 4 + c) * nl] = acc;
        }
    }
}

// the segment ends of the y / z lines: SEG + 2 loads and 4 NS dot products
// per thread, no recurrence
template <int DIR, int NS>
__global__ void __launch_bounds__(SEGT) k_seg_ends(const double *__restrict__ f, double *__restrict__ zb,
                                                   const double *__restrict__ end_u,
                                                   const double *__restrict__ end_d, Geom G, int ext, int P) {
    long nl, st;
    int N;
    double dx;
    seg_line_geom<DIR>(G, ext, nl, N, st, dx);
    int s;
    long t;
    segyz_block(s, t);
    if (t >= nl) return;
    const int a = seg_start(s, N, P);
    const int lb = (int)((unsigned)t / (unsigned)G.nx), la = (int)((unsigned)t - (unsigned)lb * (unsigned)G.nx);
    const long line = DIR == 1 ? G.id(la, -ext, lb) : G.id(la, lb, -ext);
    // all loads unconditional (indices clamped to the line), issued together;
    // the functionals weigh the positions off the line with 0
    double fv[SEG + 2];
#pragma unroll
    for (int k = 0; k < SEG + 2; ++k) fv[k] = f[line + (long)min(max(a - 1 + k, 0), N - 1) * st];
    seg_ends_store<NS>(fv, end_u, end_d, zb, t, nl, s, P);
}

// The sweeps of one segment (n <= SEG points from line index a, N points on
// the line) for NS systems with the segment's neighbours known: fv[k] =
// f(a - 1 + k) where the line has it, p1, p2 = x(a-1) seed the forward
// sweep, x1, x2 = x(b+1) the backward one (zero past the line's ends).  On
// return S[y][q], SS[y][q] hold the (f', f'') of system y at point a + q.
template <int NS>
__device__ __forceinline__ void seg_sweep(const double (&fv)[SEG + 2], int a, int n, int N, double idx, double idx2,
                                          const double *tab_u, const double *tab_d, const UccdSrc &cs,
                                          double (&p1)[NS], double (&p2)[NS], double (&x1)[NS], double (&x2)[NS],
                                          double (&S)[NS][SEG], double (&SS)[NS][SEG]) {
    const double *tab[2] = {tab_u, tab_d};
#pragma unroll
    for (int q = 0; q < SEG; ++q) {
        if (q < n) {
            const int i = a + q;
            double r1[NS], r2;
            if (i == 0) {
                const double f0 = fv[q + 1], f1 = fv[q + 2], f2 = fv[q + 3], f3 = fv[q + 4 < SEG + 2 ? q + 4 : SEG + 1];
                r1[0] = (-3.5 * f0 + 4.0 * f1 - 0.5 * f2) * idx;
                r2 = (34.0 / 3.0 * f0 - 83.0 / 4.0 * f1 + 10.0 * f2 - 7.0 / 12.0 * f3) * idx2;
                if (NS == 2) r1[NS - 1] = r1[0];
            } else if (i == N - 1) {
                const double e0 = fv[q + 1], e1 = fv[q], e2 = fv[q >= 1 ? q - 1 : 0], e3 = fv[q >= 2 ? q - 2 : 0];
                r1[0] = -(-3.5 * e0 + 4.0 * e1 - 0.5 * e2) * idx;
                r2 = (34.0 / 3.0 * e0 - 83.0 / 4.0 * e1 + 10.0 * e2 - 7.0 / 12.0 * e3) * idx2;
                if (NS == 2) r1[NS - 1] = r1[0];
            } else {
                const double fm = fv[q], fc = fv[q + 1], fp = fv[q + 2];
                r2 = (3.0 * fm - 6.0 * fc + 3.0 * fp) * idx2;
                if (NS == 1) {
                    r1[0] = 15.0 / 16.0 * (fp - fm) * idx;
                } else {
                    r1[0] = (cs.c1 * fm + cs.c2 * fc + cs.c3 * fp) * idx;
                    r1[NS - 1] = -(cs.c3 * fm + cs.c2 * fc + cs.c1 * fp) * idx;
                }
            }
#pragma unroll
            for (int y = 0; y < NS; ++y) {
                const double *co = tab[y] + (long)i * NPACK;
                const double Sn = r1[y] - (co[PF_A1] * p1[y] + co[PF_B1] * p2[y]);
                const double SSn = r2 - (co[PF_AA1] * p1[y] + co[PF_BB1] * p2[y]);
                p1[y] = S[y][q] = Sn;
                p2[y] = SS[y][q] = SSn;
            }
        }
    }
#pragma unroll
    for (int q = SEG - 1; q >= 0; --q) {
        if (q < n) {
            const int i = a + q;
#pragma unroll
            for (int y = 0; y < NS; ++y) {
                const double *co = tab[y] + (long)i * NPACK;
                const double Si = S[y][q] - (co[PB_A3] * x1[y] + co[PB_B3] * x2[y]);
                const double SSi = SS[y][q] - (co[PB_AA3] * x1[y] + co[PB_BB3] * x2[y]);
                x1[y] = co[PB_BB2D] * Si - co[PB_B2D] * SSi;
                x2[y] = co[PB_A2D] * SSi - co[PB_AA2D] * Si;
                S[y][q] = x1[y];
                SS[y][q] = x2[y];
            }
        }
    }
}

// x(a-1) and x(b+1) of line t's segment s from the reduced solve
template <int NS>
__device__ __forceinline__ void seg_neighbours(const double *__restrict__ zb, long t, long nl, int s, int P,
                                               double (&p1)[NS], double (&p2)[NS], double (&x1)[NS],
                                               double (&x2)[NS]) {
#pragma unroll
    for (int y = 0; y < NS; ++y) {
        p1[y] = p2[y] = x1[y] = x2[y] = 0.0;
        const double *Z = zb + (long)y * P * 4 * nl + t;
        if (s > 0) { p1[y] = Z[((long)(s - 1) * 4 + 2) * nl]; p2[y] = Z[((long)(s - 1) * 4 + 3) * nl]; }
        if (s < P - 1) { x1[y] = Z[((long)(s + 1) * 4 + 0) * nl]; x2[y] = Z[((long)(s + 1) * 4 + 1) * nl]; }
    }
}

// the segment's system with its neighbours known, written to fx, fxx (UCCD:
// fx the up-wind value where u > 0, else the down-wind one, fxx their mean)
template <int DIR, int NS>
__global__ void __launch_bounds__(SEGT) k_seg_lines(const double *__restrict__ f, const double *__restrict__ u,
                                                    double *__restrict__ fx, double *__restrict__ fxx,
                                                    const double *__restrict__ zb, const double *__restrict__ tab_u,
                                                    const double *__restrict__ tab_d, Geom G, int ext, int P,
                                                    UccdSrc cs) {
    long nl, st;
    int N;
    double dx;
    seg_line_geom<DIR>(G, ext, nl, N, st, dx);
    int s;
    long t;
    segyz_block(s, t);
    if (t >= nl) return;
    const int a = seg_start(s, N, P), n = seg_start(s + 1, N, P) - a;
    const int lb = (int)((unsigned)t / (unsigned)G.nx), la = (int)((unsigned)t - (unsigned)lb * (unsigned)G.nx);
    const long line = DIR == 1 ? G.id(la, -ext, lb) : G.id(la, lb, -ext);
    const long base = line + (long)a * st;
    const double idx = 1.0 / dx, idx2 = idx * idx;
    // f(a - 1 .. a + n): all SEG + 2 loads unconditional, at indices clamped
    // to the line, so they are issued together; the sweeps read none of the
    // values past the segment's range or the line's ends
    double fv[SEG + 2];
#pragma unroll
    for (int k = 0; k < SEG + 2; ++k) fv[k] = f[line + (long)min(max(a - 1 + k, 0), N - 1) * st];
    // upwind choice per point, from u loaded up front
    unsigned up = 0;
    if (NS == 2) {
        double uv[SEG];   // unconditional loads (clamped to the line), issued together
#pragma unroll
        for (int q = 0; q < SEG; ++q) uv[q] = u[line + (long)min(a + q, N - 1) * st];
#pragma unroll
        for (int q = 0; q < SEG; ++q)
            up |= (q < n && (uv[q] > 0.0 || (cs.ge && uv[q] == 0.0))) ? (1u << q) : 0u;
    }
    double p1[NS], p2[NS], x1[NS], x2[NS];
    seg_neighbours<NS>(zb, t, nl, s, P, p1, p2, x1, x2);
    double S[NS][SEG], SS[NS][SEG];
    seg_sweep<NS>(fv, a, n, N, idx, idx2, tab_u, tab_d, cs, p1, p2, x1, x2, S, SS);
#pragma unroll
    for (int q = 0; q < SEG; ++q)
        if (q < n) {
            const long o = base + (long)q * st;
            if (NS == 1) {
                fx[o] = S[0][q];
                fxx[o] = SS[0][q];
            } else {
                fx[o] = ((up >> q) & 1u) ? S[0][q] : S[NS - 1][q];
                if (fxx) fxx[o] = 0.5 * (SS[0][q] + SS[NS - 1][q]);
            }
        }
}

// The partitioned solve along x, the contiguous axis.  A block is one wave:
// 64 consecutive x lines (rows) and one segment.  A row's segment values are
// contiguous, so the wave moves them through an LDS tile of 64 rows x SEGX_TS
// (odd, against bank conflicts on the per-row reads): the loads and stores
// number the tile's elements row by row across the wave's lanes, a few rows
// per instruction, each row a contiguous run; each lane then works on its
// own row from registers.
constexpr int SEGX_TS = SEG + 3;

// (line group, segment) of an x-line block.  The launch is (P, groups): the
// segments of a line group are consecutive logical blocks, and xcd_block
// deals consecutive logical blocks to one XCD, so the cache lines that two
// neighbouring segments of a row share (a row's 18-value span is not 128-B
// aligned) are fetched into that XCD's L2 once.
__device__ __forceinline__ void segx_block(long &t0, int &s) {
    int bx, by, bz;
    xcd_block(bx, by, bz);
    s = bx;
    t0 = (long)by * WAVE;
}

// the block's rows (x lines t0 ..) and f(a - 1 .. a + n) of each, clipped to
// the line, into the tile; fv[k] = f(a - 1 + k) of this lane's row
__device__ __forceinline__ void segx_stage(const double *__restrict__ f, double *tile, long *rowb, const Geom &G,
                                           int ext, long nl, long t0, int rows, int a, int n, int N,
                                           double (&fv)[SEG + 2]) {
    const int lane = threadIdx.x;
    const long t = t0 + lane;
    if (t < nl) {
        const unsigned k = (unsigned)t / (unsigned)G.ny;
        rowb[lane] = G.id(-ext, (int)((unsigned)t - k * (unsigned)G.ny), (int)k);
    } else {
        rowb[lane] = 0;
    }
    __syncthreads();
    const int k0 = a > 0 ? 0 : 1, k1 = a + n < N ? n + 1 : n;   // tile columns k0 .. k1
    const int W = k1 - k0 + 1, tot = rows * W;
    {
        // at most SEG + 2 elements per lane, all loads issued before the
        // first is waited for (surplus lanes repeat the last element)
        double v[SEG + 2];
#pragma unroll
        for (int m = 0; m < SEG + 2; ++m) {
            const int e = min(lane + m * WAVE, tot - 1), r = e / W;
            v[m] = f[rowb[r] + a - 1 + k0 + (e - r * W)];
        }
#pragma unroll
        for (int m = 0; m < SEG + 2; ++m) {
            const int e = lane + m * WAVE, r = e / W;
            if (e < tot) tile[r * SEGX_TS + k0 + (e - r * W)] = v[m];
        }
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < SEG + 2; ++k) fv[k] = (k >= k0 && k <= k1) ? tile[lane * SEGX_TS + k] : 0.0;
}

template <int NS>
__global__ void __launch_bounds__(WAVE) k_seg_ends_x(const double *__restrict__ f, double *__restrict__ zb,
                                                     const double *__restrict__ end_u,
                                                     const double *__restrict__ end_d, Geom G, int ext, int P) {
    __shared__ double tile[WAVE * SEGX_TS];
    __shared__ long rowb[WAVE];
    const long nl = (long)G.ny * G.nz;
    const int N = G.nx + 2 * ext;
    long t0;
    int s;
    segx_block(t0, s);
    const long t = t0 + threadIdx.x;
    const int rows = (int)min((long)WAVE, nl - t0);
    const int a = seg_start(s, N, P), n = seg_start(s + 1, N, P) - a;
    double fv[SEG + 2];
    segx_stage(f, tile, rowb, G, ext, nl, t0, rows, a, n, N, fv);
    if (t < nl) seg_ends_store<NS>(fv, end_u, end_d, zb, t, nl, s, P);
}

template <int NS>
__global__ void __launch_bounds__(WAVE) k_seg_lines_x(const double *__restrict__ f, const double *__restrict__ u,
                                                      double *__restrict__ fx, double *__restrict__ fxx,
                                                      const double *__restrict__ zb, const double *__restrict__ tab_u,
                                                      const double *__restrict__ tab_d, Geom G, int ext, int P,
                                                      UccdSrc cs) {
    __shared__ double tile[WAVE * SEGX_TS];
    __shared__ long rowb[WAVE];
    const long nl = (long)G.ny * G.nz;
    const int N = G.nx + 2 * ext;
    const int lane = threadIdx.x;
    long t0;
    int s;
    segx_block(t0, s);
    const long t = t0 + lane;
    const int rows = (int)min((long)WAVE, nl - t0);
    const int a = seg_start(s, N, P), n = seg_start(s + 1, N, P) - a;
    double fv[SEG + 2];
    segx_stage(f, tile, rowb, G, ext, nl, t0, rows, a, n, N, fv);
    // upwind choice per point: u through the same tile
    unsigned up = 0;
    if (NS == 2) {
        __syncthreads();
        const int totu = rows * n;
        double v[SEG];
#pragma unroll
        for (int m = 0; m < SEG; ++m) {
            const int e = min(lane + m * WAVE, totu - 1), r = e / n;
            v[m] = u[rowb[r] + a + (e - r * n)];
        }
#pragma unroll
        for (int m = 0; m < SEG; ++m) {
            const int e = lane + m * WAVE, r = e / n;
            if (e < totu) tile[r * SEGX_TS + (e - r * n)] = v[m];
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < SEG; ++q)
            if (q < n) {
                const double uv = tile[lane * SEGX_TS + q];
                up |= (uv > 0.0 || (cs.ge && uv == 0.0)) ? (1u << q) : 0u;
            }
    }
    const double idx = 1.0 / G.dx, idx2 = idx * idx;
    double p1[NS], p2[NS], x1[NS], x2[NS];
    double S[NS][SEG], SS[NS][SEG];
    if (t < nl) {
        seg_neighbours<NS>(zb, t, nl, s, P, p1, p2, x1, x2);
        seg_sweep<NS>(fv, a, n, N, idx, idx2, tab_u, tab_d, cs, p1, p2, x1, x2, S, SS);
    }
    // fx, then fxx, out through the tile
#pragma unroll
    for (int o2 = 0; o2 < 2; ++o2) {
        if (o2 == 1 && !fxx) break;
        __syncthreads();   // the tile's previous contents are consumed
        if (t < nl) {
#pragma unroll
            for (int q = 0; q < SEG; ++q)
                if (q < n) {
                    double v;
                    if (NS == 1) v = o2 == 0 ? S[0][q] : SS[0][q];
                    else v = o2 == 0 ? (((up >> q) & 1u) ? S[0][q] : S[NS - 1][q]) : 0.5 * (SS[0][q] + SS[NS - 1][q]);
                    tile[lane * SEGX_TS + q] = v;
                }
        }
        __syncthreads();
        double *out = o2 == 0 ? fx : fxx;
        for (int e = lane; e < rows * n; e += WAVE) {
            const int r = e / n, q = e - r * n;
            out[rowb[r] + a + q] = tile[r * SEGX_TS + q];
        }
    }
}

// The reduced block-tridiagonal solve of one line and system (blockIdx.y:
// 0 the CCD or up-wind system, 1 the down-wind one): forward d'_s = d_s -
// m_s d'_{s-1}, backward u_s = Dinv_s (d'_s - U_s u_{s+1}), with d_s =
// (x0(b_s), x0(a_{s+1})) read from zb and u_s = (x(b_s), x(a_{s+1})) written
// back over them.
__global__ void __launch_bounds__(SEGT) k_seg_reduce(double *__restrict__ zb, const double *__restrict__ red_u,
                                                     const double *__restrict__ red_d, long nl, int P) {
    const long t = (long)blockIdx.x * SEGT + threadIdx.x;
    if (t >= nl) return;
    const int NI = P - 1, y = blockIdx.y;
    const double *R = y ? red_d : red_u;
    double *Z = zb + (long)y * P * 4 * nl + t;
    // slots of interface s: (s, 2), (s, 3), (s + 1, 0), (s + 1, 1)
    auto slot = [&](int s, int k) -> double & {
        return Z[(k < 2 ? (long)s * 4 + 2 + k : (long)(s + 1) * 4 + k - 2) * nl];
    };
    double dp[4] = {0.0, 0.0, 0.0, 0.0};
    for (int s = 0; s < NI; ++s) {
        const double *m = R + (long)s * NRED;
        double d[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) d[k] = slot(s, k);
#pragma unroll
        for (int r = 0; r < 4; ++r)
            d[r] -= m[r * 4 + 0] * dp[0] + m[r * 4 + 1] * dp[1] + m[r * 4 + 2] * dp[2] + m[r * 4 + 3] * dp[3];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            if (s < NI - 1) slot(s, k) = d[k];
            dp[k] = d[k];
        }
    }
    double un[4] = {0.0, 0.0, 0.0, 0.0};
    for (int s = NI - 1; s >= 0; --s) {
        const double *Di = R + (long)s * NRED + 16, *U = Di + 16;
        double w[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) w[k] = s == NI - 1 ? dp[k] : slot(s, k);
#pragma unroll
        for (int r = 0; r < 4; ++r)
            w[r] -= U[r * 4 + 0] * un[0] + U[r * 4 + 1] * un[1] + U[r * 4 + 2] * un[2] + U[r * 4 + 3] * un[3];
#pragma unroll
        for (int r = 0; r < 4; ++r)
            un[r] = Di[r * 4 + 0] * w[0] + Di[r * 4 + 1] * w[1] + Di[r * 4 + 2] * w[2] + Di[r * 4 + 3] * w[3];
#pragma unroll
        for (int k = 0; k < 4; ++k) slot(s, k) = un[k];
    }
}

// k_seg_reduce for lines of at most MAXI interfaces: all d_s are loaded up
// front into registers, so the loads of a line overlap instead of costing one
// memory latency per interface and sweep (MAXI = 20: 160 of the registers),
// and the coefficients come from LDS.
constexpr int SEG_MAXI = 20;
template <int MAXI>
__global__ void __launch_bounds__(SEGT) k_seg_reduce_r(double *__restrict__ zb, const double *__restrict__ red_u,
                                                       const double *__restrict__ red_d, long nl, int P) {
    // the system's reduced LU staged in LDS: read through the scalar cache it
    // costs one cache-miss latency per interface and sweep on the critical path
    __shared__ double R[MAXI * NRED];
    const int NI = P - 1, y = blockIdx.y;
    const double *Rg = y ? red_d : red_u;
    for (int e = threadIdx.x; e < NI * NRED; e += SEGT) R[e] = Rg[e];
    __syncthreads();
    const long t = (long)blockIdx.x * SEGT + threadIdx.x;
    if (t >= nl) return;
    double *Z = zb + (long)y * P * 4 * nl + t;
    auto slot = [&](int s, int k) -> double & {
        return Z[(k < 2 ? (long)s * 4 + 2 + k : (long)(s + 1) * 4 + k - 2) * nl];
    };
    double d[MAXI][4];
#pragma unroll
    for (int s = 0; s < MAXI; ++s)
        if (s < NI) {
#pragma unroll
            for (int k = 0; k < 4; ++k) d[s][k] = slot(s, k);
        }
#pragma unroll
    for (int s = 1; s < MAXI; ++s)
        if (s < NI) {
            const double *m = R + (long)s * NRED;
#pragma unroll
            for (int r = 0; r < 4; ++r)
                d[s][r] -= m[r * 4 + 0] * d[s - 1][0] + m[r * 4 + 1] * d[s - 1][1] + m[r * 4 + 2] * d[s - 1][2] +
                           m[r * 4 + 3] * d[s - 1][3];
        }
    double un[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int s = MAXI - 1; s >= 0; --s)
        if (s < NI) {
            const double *Di = R + (long)s * NRED + 16, *U = Di + 16;
            double w[4];
#pragma unroll
            for (int r = 0; r < 4; ++r)
                w[r] = d[s][r] -
                       (U[r * 4 + 0] * un[0] + U[r * 4 + 1] * un[1] + U[r * 4 + 2] * un[2] + U[r * 4 + 3] * un[3]);
#pragma unroll
            for (int r = 0; r < 4; ++r)
                un[r] = Di[r * 4 + 0] * w[0] + Di[r * 4 + 1] * w[1] + Di[r * 4 + 2] * w[2] + Di[r * 4 + 3] * w[3];
#pragma unroll
            for (int k = 0; k < 4; ++k) slot(s, k) = un[k];
        }
}

// the partitioned solve of the y (dir 1) or z (dir 2) lines: NS = 1 CCD
// with table k0, NS = 2 UCCD with tables k0 (up) and k1 (down)
// the partitioned solve of the lines along DIR: NS = 1 CCD with table k0,
// NS = 2 UCCD with tables k0 (up) and k1 (down)
template <int DIR, int NS>
static int seg_solve(fcfd_ctx *c, int ext, int k0, int k1, const double *f, const double *u, double *fx, double *fxx,
                     UccdSrc cs) {
    const Geom &G = c->G;
    const long nl = DIR == 0 ? (long)G.ny * G.nz : (DIR == 1 ? (long)G.nx * G.nz : (long)G.nx * G.ny);
    const int N = (DIR == 0 ? G.nx : (DIR == 1 ? G.ny : G.nz)) + 2 * ext;
    const int P = seg_count(N);
    const double *tu = c->ccd_seg[k0 * 3 + DIR], *td = c->ccd_seg[k1 * 3 + DIR];
    const long red_off = (long)N * NPACK, end_off = red_off + (long)(P - 1) * NRED;
    const unsigned gx = (unsigned)((nl + SEGT - 1) / SEGT), gw = (unsigned)((nl + WAVE - 1) / WAVE);
    if (P > 1) {
        if constexpr (DIR == 0)
            hipLaunchKernelGGL(k_seg_ends_x<NS>, dim3((unsigned)P, gw), dim3(WAVE), 0, c->stream, f, c->ccd_z,
                               tu + end_off, td + end_off, G, ext, P);
        else
            hipLaunchKernelGGL((k_seg_ends<DIR, NS>), dim3((unsigned)P, gx), dim3(SEGT), 0, c->stream, f, c->ccd_z,
                               tu + end_off, td + end_off, G, ext, P);
        if (P - 1 <= SEG_MAXI)
            hipLaunchKernelGGL(k_seg_reduce_r<SEG_MAXI>, dim3(gx, NS), dim3(SEGT), 0, c->stream, c->ccd_z,
                               tu + red_off, td + red_off, nl, P);
        else
            hipLaunchKernelGGL(k_seg_reduce, dim3(gx, NS), dim3(SEGT), 0, c->stream, c->ccd_z, tu + red_off,
                               td + red_off, nl, P);
    }
    if constexpr (DIR == 0)
        hipLaunchKernelGGL(k_seg_lines_x<NS>, dim3((unsigned)P, gw), dim3(WAVE), 0, c->stream, f, u, fx, fxx,
                           c->ccd_z, tu, td, G, ext, P, cs);
    else
        hipLaunchKernelGGL((k_seg_lines<DIR, NS>), dim3((unsigned)P, gx), dim3(SEGT), 0, c->stream, f, u, fx, fxx,
                           c->ccd_z, tu, td, G, ext, P, cs);
    FCFD_CHECK(hipGetLastError());
    return FCFD_OK;
}

// CCD of f along direction dir: set 0 the newOpenMP lines through the ghosts
// (n + 6 points), set 1 the MPLS_RisingBubble interior lines (n points)
int ccd_lines_set(fcfd_ctx *c, int set, int dir, const double *f, double *fx, double *fxx) {
    const Geom &G = c->G;
    const int ext = set ? 0 : G.g;
    long nl = dir == 0 ? (long)G.ny * G.nz : (dir == 1 ? (long)G.nx * G.nz : (long)G.nx * G.ny);
    const int N = (dir == 0 ? G.nx : (dir == 1 ? G.ny : G.nz)) + 2 * ext;
    const size_t need = sizeof(double) * NPACK * N;
    const int in_lds = (long)need <= CCD_LDS_MAX;   // else the table is read from global memory
    const size_t lds = in_lds ? need : 0;
    constexpr int CH = 8;
    dim3 grd((unsigned)((nl + CCD_THREADS - 1) / CCD_THREADS));
    ProfScope ps_(c, R_CCD);
    const int kc = set ? MAT_M3_CCD : MAT_CCD;
    const double *co = c->ccd_coef[kc * 3 + dir];
    if (dir == 0 && c->ccd_seg[kc * 3 + 0]) return seg_solve<0, 1>(c, ext, kc, kc, f, nullptr, fx, fxx, UccdSrc{});
    if (dir == 1 && c->ccd_seg[kc * 3 + 1]) return seg_solve<1, 1>(c, ext, kc, kc, f, nullptr, fx, fxx, UccdSrc{});
    if (dir == 2 && c->ccd_seg[kc * 3 + 2]) return seg_solve<2, 1>(c, ext, kc, kc, f, nullptr, fx, fxx, UccdSrc{});
    if (dir == 0)
        hipLaunchKernelGGL(k_ccd_lines_x, dim3((unsigned)((nl + CCDX_LINES - 1) / CCDX_LINES)), dim3(CCDX_LINES), 0,
                           c->stream, f, fx, fxx, co, G, ext);
    else if (dir == 1)
        hipLaunchKernelGGL((k_ccd_lines<1, CH>), grd, dim3(CCD_THREADS), lds, c->stream, f, fx, fxx, co, G, in_lds,
                           ext);
    else
        hipLaunchKernelGGL((k_ccd_lines<2, CH>), grd, dim3(CCD_THREADS), lds, c->stream, f, fx, fxx, co, G, in_lds,
                           ext);
    FCFD_CHECK(hipGetLastError());
    return FCFD_OK;
}

int ccd_dir(fcfd_ctx *c, int dir, const double *f, double *fx, double *fxx) {
    return ccd_lines_set(c, 0, dir, f, fx, fxx);
}

// root_solve("uccd") along one line (`.../lib/solver/ccd_solvers.f90:
// 308-341,400-412`): the up- and down-wind systems share the second-
// derivative source and differ in the first-derivative source and matrix;
// both are swept together, and the result keeps fx = SU where u > 0, else SD,
// fxx = (SSU + SSD)/2 (fxx optional).  Forward results park in fx (SU) and
// three scratch arrays (SSU, SD, SSD).  One thread per line, CH-point chunks
// with the next chunk's loads issued before the current chunk's recurrence.
template <int DIR, int CH>
__global__ void __launch_bounds__(CCD_THREADS) k_uccd_lines(const double *__restrict__ f, const double *__restrict__ u,
                                                            double *__restrict__ fx, double *__restrict__ fxx,
                                                            double *__restrict__ t1, double *__restrict__ t2,
                                                            double *__restrict__ t3, const double *__restrict__ cog_u,
                                                            const double *__restrict__ cog_d, Geom G, int use_lds,
                                                            int ext, UccdSrc cs) {
    extern __shared__ double lds[];
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    long nlines, st;
    int N;
    double dx;
    if (DIR == 0) { nlines = (long)G.ny * G.nz; N = G.nx + 2 * ext; st = 1; dx = G.dx; }
    else if (DIR == 1) { nlines = (long)G.nx * G.nz; N = G.ny + 2 * ext; st = G.sx; dx = G.dy; }
    else { nlines = (long)G.nx * G.ny; N = G.nz + 2 * ext; st = G.sxy; dx = G.dz; }
    const double *cu = cog_u, *cd = cog_d;
    if (use_lds) {
        for (int e = threadIdx.x; e < N * NPACK; e += CCD_THREADS) {
            lds[e] = cog_u[e];
            lds[N * NPACK + e] = cog_d[e];
        }
        __syncthreads();
        cu = lds;
        cd = lds + N * NPACK;
    }
    if (t >= nlines) return;
    long base;
    if (DIR == 0) base = G.id(-ext, (int)(t % G.ny), (int)(t / G.ny));
    else if (DIR == 1) base = G.id((int)(t % G.nx), -ext, (int)(t / G.nx));
    else base = G.id((int)(t % G.nx), (int)(t / G.nx), -ext);
    const double *F = f + base, *U = u + base;
    double *X = fx + base, *XX = fxx ? fxx + base : nullptr, *T1 = t1 + base, *T2 = t2 + base, *T3 = t3 + base;
    const double idx = 1.0 / dx, idx2 = idx * idx;
    const double c1 = cs.c1, c2 = cs.c2, c3 = cs.c3;
    // boundary rows (shared by both systems)
    const double f0 = F[0], f1 = F[st], f2 = F[2 * st], f3 = F[3 * st];
    double Su = (-3.5 * f0 + 4.0 * f1 - 0.5 * f2) * idx;
    double SSu = (34.0 / 3.0 * f0 - 83.0 / 4.0 * f1 + 10.0 * f2 - 7.0 / 12.0 * f3) * idx2;
    double Sd = Su, SSd = SSu;
    X[0] = Su; T1[0] = SSu; T2[0] = Sd; T3[0] = SSd;
    double fm = f0, fc = f1;
    double fp[CH], fq[CH];
#pragma unroll
    for (int q = 0; q < CH; ++q)
        if (1 + q < N - 1) fp[q] = F[(long)(2 + q) * st];
    for (int i0 = 1; i0 < N - 1; i0 += CH) {
        const int n0 = i0 + CH;
#pragma unroll
        for (int q = 0; q < CH; ++q)
            if (n0 + q < N - 1) fq[q] = F[(long)(n0 + q + 1) * st];
#pragma unroll
        for (int q = 0; q < CH; ++q) {
            const int i = i0 + q;
            if (i < N - 1) {
                const double *a = cu + i * NPACK, *b = cd + i * NPACK;
                const double ss = (3.0 * fm - 6.0 * fc + 3.0 * fp[q]) * idx2;
                const double su = (c1 * fm + c2 * fc + c3 * fp[q]) * idx;
                const double sd = -(c3 * fm + c2 * fc + c1 * fp[q]) * idx;
                const double Sun = su - (a[PF_A1] * Su + a[PF_B1] * SSu);
                const double SSun = ss - (a[PF_AA1] * Su + a[PF_BB1] * SSu);
                const double Sdn = sd - (b[PF_A1] * Sd + b[PF_B1] * SSd);
                const double SSdn = ss - (b[PF_AA1] * Sd + b[PF_BB1] * SSd);
                Su = Sun; SSu = SSun; Sd = Sdn; SSd = SSdn;
                const long o = (long)i * st;
                X[o] = Su; T1[o] = SSu; T2[o] = Sd; T3[o] = SSd;
                fm = fc; fc = fp[q];
            }
        }
#pragma unroll
        for (int q = 0; q < CH; ++q) fp[q] = fq[q];
    }
    const int ie = N - 1;
    {
        const double e0 = F[(long)ie * st], e1 = F[(long)(ie - 1) * st], e2 = F[(long)(ie - 2) * st],
                     e3 = F[(long)(ie - 3) * st];
        const double s = -(-3.5 * e0 + 4.0 * e1 - 0.5 * e2) * idx;
        const double ss = (34.0 / 3.0 * e0 - 83.0 / 4.0 * e1 + 10.0 * e2 - 7.0 / 12.0 * e3) * idx2;
        const double *a = cu + ie * NPACK, *b = cd + ie * NPACK;
        const double Sun = s - (a[PF_A1] * Su + a[PF_B1] * SSu);
        const double SSun = ss - (a[PF_AA1] * Su + a[PF_BB1] * SSu);
        const double Sdn = s - (b[PF_A1] * Sd + b[PF_B1] * SSd);
        const double SSdn = ss - (b[PF_AA1] * Sd + b[PF_BB1] * SSd);
        Su = a[PB_BB2D] * Sun - a[PB_B2D] * SSun;
        SSu = a[PB_A2D] * SSun - a[PB_AA2D] * Sun;
        Sd = b[PB_BB2D] * Sdn - b[PB_B2D] * SSdn;
        SSd = b[PB_A2D] * SSdn - b[PB_AA2D] * Sdn;
        const long o = (long)ie * st;
        X[o] = (U[o] > 0.0 || (cs.ge && U[o] == 0.0)) ? Su : Sd;
        if (XX) XX[o] = 0.5 * (SSu + SSd);
    }
    // back substitution of both systems
    for (int i1 = ie - 1; i1 >= 0; i1 -= CH) {
        double xu[CH], xxu[CH], xd[CH], xxd[CH], uv[CH];
#pragma unroll
        for (int q = 0; q < CH; ++q)
            if (i1 - q >= 0) {
                const long o = (long)(i1 - q) * st;
                xu[q] = X[o]; xxu[q] = T1[o]; xd[q] = T2[o]; xxd[q] = T3[o]; uv[q] = U[o];
            }
#pragma unroll
        for (int q = 0; q < CH; ++q) {
            const int i = i1 - q;
            if (i >= 0) {
                const double *a = cu + i * NPACK, *b = cd + i * NPACK;
                const double Siu = xu[q] - (a[PB_A3] * Su + a[PB_B3] * SSu);
                const double SSiu = xxu[q] - (a[PB_AA3] * Su + a[PB_BB3] * SSu);
                const double Sid = xd[q] - (b[PB_A3] * Sd + b[PB_B3] * SSd);
                const double SSid = xxd[q] - (b[PB_AA3] * Sd + b[PB_BB3] * SSd);
                Su = a[PB_BB2D] * Siu - a[PB_B2D] * SSiu;
                SSu = a[PB_A2D] * SSiu - a[PB_AA2D] * Siu;
                Sd = b[PB_BB2D] * Sid - b[PB_B2D] * SSid;
                SSd = b[PB_A2D] * SSid - b[PB_AA2D] * Sid;
                const long o = (long)i * st;
                X[o] = (uv[q] > 0.0 || (cs.ge && uv[q] == 0.0)) ? Su : Sd;
                if (XX) XX[o] = 0.5 * (SSu + SSd);
            }
        }
    }
}

// UCCD of f along direction dir, upwind by u; set 0: the newOpenMP lines
// through the ghosts, set 1: the MPLS_RisingBubble interior lines with its
// M-DRPCCD coefficients; scratch: the context's cx, cy, cz
// k_uccd_lines along x, the contiguous axis: one wave owns 64 lines and
// moves them through CX-point LDS tiles, as k_ccd_lines_x does, so global
// accesses are row segments of CX values (CX = 8: four 9 x 64 tiles, 18 KB,
// keep two waves per SIMD resident).  Forward: a chunk of f comes in,
// each lane runs both systems over it and leaves SU, SSU, SD, SSD in four
// tiles (SSD over the f tile, read just before), which go out together.
// Backward: the four forward tiles of a chunk come in, each lane back-
// substitutes both systems and leaves SU, SD and (SSU + SSD) / 2; the choice
// between SU and SD by the sign of u is made as the tile goes out, where the
// loads of u are coalesced.  Same arithmetic per point as k_uccd_lines.
template <int CX>
__global__ void __launch_bounds__(CCDX_LINES) k_uccd_lines_x(const double *__restrict__ f,
                                                             const double *__restrict__ u,
                                                             double *__restrict__ fx, double *__restrict__ fxx,
                                                             double *__restrict__ t1, double *__restrict__ t2,
                                                             double *__restrict__ t3, const double *__restrict__ cu,
                                                             const double *__restrict__ cd, Geom G, int ext,
                                                             UccdSrc cs) {
    constexpr int TS = CX + 1;
    __shared__ double tA[CCDX_LINES * TS], tB[CCDX_LINES * TS], tC[CCDX_LINES * TS], tD[CCDX_LINES * TS];
    const int lane = threadIdx.x;
    const long nlines = (long)G.ny * G.nz, L0 = (long)blockIdx.x * CCDX_LINES, t = L0 + lane;
    const int N = G.nx + 2 * ext;
    const bool valid = t < nlines;
    long mybase = 0;
    if (valid) mybase = G.id(-ext, (int)(t % G.ny), (int)(t / G.ny));
    constexpr int RPI = CCDX_LINES / CX;
    const int lsub = lane / CX, usub = lane % CX;
    int rbase[CX];   // row starts (the caller keeps fields below 2^31 values)
    bool rval[CX];
#pragma unroll
    for (int r = 0; r < CX; ++r) {
        const int l = r * RPI + lsub;
        rbase[r] = __shfl((int)mybase, l, CCDX_LINES);
        rval[r] = L0 + l < nlines;
    }
    const double idx = 1.0 / G.dx, idx2 = idx * idx;
    const double c1 = cs.c1, c2 = cs.c2, c3 = cs.c3;
    const double *F = f + mybase;
    auto put = [&](double *tile, const double *v) {
#pragma unroll
        for (int r = 0; r < CX; ++r) tile[(r * RPI + lsub) * TS + usub] = v[r];
    };
    // ---- forward sweep: boundary row, then the chunks
    double Su = 0.0, SSu = 0.0, Sd = 0.0, SSd = 0.0, fm = 0.0, fc = 0.0;
    if (valid) {
        const double f0 = F[0], f1 = F[1], f2 = F[2], f3 = F[3];
        Su = (-3.5 * f0 + 4.0 * f1 - 0.5 * f2) * idx;
        SSu = (34.0 / 3.0 * f0 - 83.0 / 4.0 * f1 + 10.0 * f2 - 7.0 / 12.0 * f3) * idx2;
        Sd = Su;
        SSd = SSu;
        fx[mybase] = Su; t1[mybase] = SSu; t2[mybase] = Sd; t3[mybase] = SSd;
        fm = f0;
        fc = f1;
    }
    double pf[CX];
    auto load_f = [&](int j0) {
#pragma unroll
        for (int r = 0; r < CX; ++r)
            if (rval[r] && j0 + usub < N) pf[r] = f[rbase[r] + j0 + usub];
    };
    load_f(2);
    for (int j0 = 2; j0 < N; j0 += CX) {
        put(tD, pf);
        __syncthreads();
        if (j0 + CX < N) load_f(j0 + CX);
        if (valid) {
            double *ra = tA + lane * TS, *rb = tB + lane * TS, *rc = tC + lane * TS,
                   *rd = tD + lane * TS;
#pragma unroll 2
            for (int q = 0; q < CX; ++q) {
                const int i = j0 - 1 + q;
                if (i <= N - 2) {
                    const double *a = cu + i * NPACK, *b = cd + i * NPACK;
                    const double fp = rd[q];
                    const double ss = (3.0 * fm - 6.0 * fc + 3.0 * fp) * idx2;
                    const double su = (c1 * fm + c2 * fc + c3 * fp) * idx;
                    const double sd = -(c3 * fm + c2 * fc + c1 * fp) * idx;
                    const double Sun = su - (a[PF_A1] * Su + a[PF_B1] * SSu);
                    const double SSun = ss - (a[PF_AA1] * Su + a[PF_BB1] * SSu);
                    const double Sdn = sd - (b[PF_A1] * Sd + b[PF_B1] * SSd);
                    const double SSdn = ss - (b[PF_AA1] * Sd + b[PF_BB1] * SSd);
                    Su = Sun; SSu = SSun; Sd = Sdn; SSd = SSdn;
                    ra[q] = Su; rb[q] = SSu; rc[q] = Sd; rd[q] = SSd;
                    fm = fc; fc = fp;
                }
            }
        }
        __syncthreads();
#pragma unroll
        for (int r = 0; r < CX; ++r) {
            const int l = r * RPI + lsub, i = j0 - 1 + usub;
            if (rval[r] && i <= N - 2) {
                const long o = rbase[r] + i;
                fx[o] = tA[l * TS + usub];
                t1[o] = tB[l * TS + usub];
                t2[o] = tC[l * TS + usub];
                t3[o] = tD[l * TS + usub];
            }
        }
        __syncthreads();
    }
    const int ie = N - 1;
    if (valid) {
        const double e0 = F[ie], e1 = F[ie - 1], e2 = F[ie - 2], e3 = F[ie - 3];
        const double s = -(-3.5 * e0 + 4.0 * e1 - 0.5 * e2) * idx;
        const double ss = (34.0 / 3.0 * e0 - 83.0 / 4.0 * e1 + 10.0 * e2 - 7.0 / 12.0 * e3) * idx2;
        const double *a = cu + ie * NPACK, *b = cd + ie * NPACK;
        const double Sun = s - (a[PF_A1] * Su + a[PF_B1] * SSu);
        const double SSun = ss - (a[PF_AA1] * Su + a[PF_BB1] * SSu);
        const double Sdn = s - (b[PF_A1] * Sd + b[PF_B1] * SSd);
        const double SSdn = ss - (b[PF_AA1] * Sd + b[PF_BB1] * SSd);
        Su = a[PB_BB2D] * Sun - a[PB_B2D] * SSun;
        SSu = a[PB_A2D] * SSun - a[PB_AA2D] * Sun;
        Sd = b[PB_BB2D] * Sdn - b[PB_B2D] * SSdn;
        SSd = b[PB_A2D] * SSdn - b[PB_AA2D] * Sdn;
        const double uv = u[mybase + ie];
        fx[mybase + ie] = (uv > 0.0 || (cs.ge && uv == 0.0)) ? Su : Sd;
        if (fxx) fxx[mybase + ie] = 0.5 * (SSu + SSd);
    }
    // ---- back substitution over i = N-2 .. 1 by the chunks in reverse, then i = 0
    // (the four forward arrays of a chunk go straight into the tiles: held
    // in registers for the next chunk they would take 128 VGPRs)
    const int nch = (N - 2 + CX - 1) / CX;
    __syncthreads();
    for (int cix = nch - 1; cix >= 0; --cix) {
        const int j0 = 2 + cix * CX;
#pragma unroll
        for (int r = 0; r < CX; ++r) {
            const int i = j0 - 1 + usub, e = (r * RPI + lsub) * TS + usub;
            if (rval[r] && i <= N - 2) {
                const long o = rbase[r] + i;
                tA[e] = fx[o]; tB[e] = t1[o]; tC[e] = t2[o]; tD[e] = t3[o];
            }
        }
        __syncthreads();
        if (valid) {
            double *ra = tA + lane * TS, *rb = tB + lane * TS, *rc = tC + lane * TS,
                   *rd = tD + lane * TS;
#pragma unroll 2
            for (int q = CX - 1; q >= 0; --q) {
                const int i = j0 - 1 + q;
                if (i <= N - 2) {
                    const double *a = cu + i * NPACK, *b = cd + i * NPACK;
                    const double Siu = ra[q] - (a[PB_A3] * Su + a[PB_B3] * SSu);
                    const double SSiu = rb[q] - (a[PB_AA3] * Su + a[PB_BB3] * SSu);
                    const double Sid = rc[q] - (b[PB_A3] * Sd + b[PB_B3] * SSd);
                    const double SSid = rd[q] - (b[PB_AA3] * Sd + b[PB_BB3] * SSd);
                    Su = a[PB_BB2D] * Siu - a[PB_B2D] * SSiu;
                    SSu = a[PB_A2D] * SSiu - a[PB_AA2D] * Siu;
                    Sd = b[PB_BB2D] * Sid - b[PB_B2D] * SSid;
                    SSd = b[PB_A2D] * SSid - b[PB_AA2D] * Sid;
                    ra[q] = Su; rc[q] = Sd; rb[q] = 0.5 * (SSu + SSd);
                }
            }
        }
        __syncthreads();
#pragma unroll
        for (int r = 0; r < CX; ++r) {
            const int l = r * RPI + lsub, i = j0 - 1 + usub;
            if (rval[r] && i <= N - 2) {
                const long o = rbase[r] + i;
                const double uv = u[o];
                fx[o] = (uv > 0.0 || (cs.ge && uv == 0.0)) ? tA[l * TS + usub] : tC[l * TS + usub];
                if (fxx) fxx[o] = tB[l * TS + usub];
            }
        }
        __syncthreads();
    }
    if (valid) {
        const double *a = cu, *b = cd;
        const double Siu = fx[mybase] - (a[PB_A3] * Su + a[PB_B3] * SSu);
        const double SSiu = t1[mybase] - (a[PB_AA3] * Su + a[PB_BB3] * SSu);
        const double Sid = t2[mybase] - (b[PB_A3] * Sd + b[PB_B3] * SSd);
        const double SSid = t3[mybase] - (b[PB_AA3] * Sd + b[PB_BB3] * SSd);
        Su = a[PB_BB2D] * Siu - a[PB_B2D] * SSiu;
        SSu = a[PB_A2D] * SSiu - a[PB_AA2D] * Siu;
        Sd = b[PB_BB2D] * Sid - b[PB_B2D] * SSid;
        SSd = b[PB_A2D] * SSid - b[PB_AA2D] * Sid;
        const double uv = u[mybase];
        fx[mybase] = (uv > 0.0 || (cs.ge && uv == 0.0)) ? Su : Sd;
        if (fxx) fxx[mybase] = 0.5 * (SSu + SSd);
    }
}

int uccd_lines_set(fcfd_ctx *c, int set, int dir, const double *f, const double *u, double *fx, double *fxx) {
    const Geom &G = c->G;
    const int ext = set ? 0 : G.g;
    const long nl = dir == 0 ? (long)G.ny * G.nz : (dir == 1 ? (long)G.nx * G.nz : (long)G.nx * G.ny);
    const int N = (dir == 0 ? G.nx : (dir == 1 ? G.ny : G.nz)) + 2 * ext;
    const size_t need = 2 * sizeof(double) * NPACK * N;
    const int use_lds = (long)need <= CCD_LDS_MAX;
    const size_t lds = use_lds ? need : 0;
    constexpr int CH = 8;
    dim3 grd((unsigned)((nl + CCD_THREADS - 1) / CCD_THREADS));
    const int ku = set ? MAT_M3_UP : MAT_UCCD_UP, kd = set ? MAT_M3_DOWN : MAT_UCCD_DOWN;
    const double *cu = c->ccd_coef[ku * 3 + dir], *cd = c->ccd_coef[kd * 3 + dir];
    const UccdSrc cs = set ? UCCD_M3 : UCCD_N3;
    ProfScope ps_(c, R_CCD);
    if (dir == 0 && c->ccd_seg[ku * 3 + 0]) return seg_solve<0, 2>(c, ext, ku, kd, f, u, fx, fxx, cs);
    if (dir == 1 && c->ccd_seg[ku * 3 + 1]) return seg_solve<1, 2>(c, ext, ku, kd, f, u, fx, fxx, cs);
    if (dir == 2 && c->ccd_seg[ku * 3 + 2]) return seg_solve<2, 2>(c, ext, ku, kd, f, u, fx, fxx, cs);
    if (dir == 0 && G.ntot < (1L << 31))
        hipLaunchKernelGGL(k_uccd_lines_x<8>, dim3((unsigned)((nl + CCDX_LINES - 1) / CCDX_LINES)), dim3(CCDX_LINES), 0,
                           c->stream, f, u, fx, fxx, c->cx, c->cy, c->cz, cu, cd, G, ext, cs);
    else if (dir == 0)
        hipLaunchKernelGGL((k_uccd_lines<0, CH>), grd, dim3(CCD_THREADS), lds, c->stream, f, u, fx, fxx, c->cx, c->cy,
                           c->cz, cu, cd, G, use_lds, ext, cs);
    else if (dir == 1)
        hipLaunchKernelGGL((k_uccd_lines<1, CH>), grd, dim3(CCD_THREADS), lds, c->stream, f, u, fx, fxx, c->cx, c->cy,
                           c->cz, cu, cd, G, use_lds, ext, cs);
    else
        hipLaunchKernelGGL((k_uccd_lines<2, CH>), grd, dim3(CCD_THREADS), lds, c->stream, f, u, fx, fxx, c->cx, c->cy,
                           c->cz, cu, cd, G, use_lds, ext, cs);
    FCFD_CHECK(hipGetLastError());
    return FCFD_OK;
}

int uccd_dir(fcfd_ctx *c, int dir, const double *f, const double *u, double *fx, double *fxx) {
    return uccd_lines_set(c, 0, dir, f, u, fx, fxx);
}

// surface_norms: the reference solves every line of the ghosted box and then
// ghost-copies the results.  Lines through ghost cells of phi repeat interior
// lines (phi's ghosts are copies), so only interior lines are solved here;
// the first derivatives are ghost-copied before the cross passes read them.
int surface_norms(fcfd_ctx *c, const double *phi, int full) {
    const Geom &G = c->G;
    double **F = c->f;
    FCFD_TRY(ccd_dir(c, 0, phi, F[FCFD_NX], F[FCFD_NXX]));
    FCFD_TRY(ccd_dir(c, 1, phi, F[FCFD_NY], F[FCFD_NYY]));
    if (G.dim == 3) FCFD_TRY(ccd_dir(c, 2, phi, F[FCFD_NZ], F[FCFD_NZZ]));
    FCFD_TRY(bc_scalar(c, F[FCFD_NX]));
    FCFD_TRY(bc_scalar(c, F[FCFD_NY]));
    if (G.dim == 3) FCFD_TRY(bc_scalar(c, F[FCFD_NZ]));
    if (!full) return FCFD_OK;
    FCFD_TRY(ccd_dir(c, 1, F[FCFD_NX], F[FCFD_NXY], c->cx));
    if (G.dim == 3) {
        FCFD_TRY(ccd_dir(c, 2, F[FCFD_NX], F[FCFD_NXZ], c->cx));
        FCFD_TRY(ccd_dir(c, 2, F[FCFD_NY], F[FCFD_NYZ], c->cx));
    }
    FCFD_TRY(bc_scalar(c, F[FCFD_NXX]));
    FCFD_TRY(bc_scalar(c, F[FCFD_NYY]));
    FCFD_TRY(bc_scalar(c, F[FCFD_NXY]));
    if (G.dim == 3) {
        FCFD_TRY(bc_scalar(c, F[FCFD_NZZ]));
        FCFD_TRY(bc_scalar(c, F[FCFD_NXZ]));
        FCFD_TRY(bc_scalar(c, F[FCFD_NYZ]));
    }
    return FCFD_OK;
}

// curvature from the CCD derivatives (others_normals.f90:145-165 in 3D,
// 2D :244-260)
template <int DIM>
__global__ void k_curv(const double *__restrict__ X, const double *__restrict__ Y, const double *__restrict__ Z,
                       const double *__restrict__ XX, const double *__restrict__ YY, const double *__restrict__ ZZ,
                       const double *__restrict__ XY, const double *__restrict__ XZ, const double *__restrict__ YZ,
                       double *__restrict__ C, Geom G) {
    int bx_, by_, bz_;
    xcd_block(bx_, by_, bz_);
    const int i = bx_ * blockDim.x + threadIdx.x, j = by_ * blockDim.y + threadIdx.y, k = bz_;
    if (i >= G.nx || j >= G.ny) return;
    long id = G.id(i, j, k);
    double fx = X[id], fy = Y[id], fxx = XX[id], fyy = YY[id], fxy = XY[id];
    if (DIM == 3) {
        double fz = Z[id], fzz = ZZ[id], fxz = XZ[id], fyz = YZ[id];
        C[id] = ((fyy + fzz) * (fx * fx) + (fxx + fzz) * (fy * fy) + (fxx + fyy) * (fz * fz) -
                 2.0 * (fxy * fx * fy + fxz * fx * fz + fyz * fy * fz)) /
                pow(fx * fx + fy * fy + fz * fz + 1.0e-12, 1.5);
    } else {
        C[id] = (fyy * (fx * fx) + fxx * (fy * fy) - 2.0 * fxy * fx * fy) / pow(fx * fx + fy * fy + 1.0e-12, 1.5);
    }
}

int curv_kernel(fcfd_ctx *c) {
    const Geom &G = c->G;
    double **F = c->f;
    dim3 blk(64, 4), grd = grid_interior(G, blk);
    {
        ProfScope ps_(c, R_CURV);
        if (G.dim == 3)
            hipLaunchKernelGGL(k_curv<3>, grd, blk, 0, c->stream, F[FCFD_NX], F[FCFD_NY], F[FCFD_NZ], F[FCFD_NXX],
                               F[FCFD_NYY], F[FCFD_NZZ], F[FCFD_NXY], F[FCFD_NXZ], F[FCFD_NYZ], F[FCFD_CURV], G);
        else
            hipLaunchKernelGGL(k_curv<2>, grd, blk, 0, c->stream, F[FCFD_NX], F[FCFD_NY], nullptr, F[FCFD_NXX],
                               F[FCFD_NYY], nullptr, F[FCFD_NXY], nullptr, nullptr, F[FCFD_CURV], G);
        FCFD_CHECK(hipGetLastError());
    }
    return bc_scalar(c, F[FCFD_CURV]);
}

}  // namespace fcfd
