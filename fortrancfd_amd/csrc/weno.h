// weno.h — WENO5-JS face reconstructions of the level-set solvers and their
// reverse-mode derivatives (the adjoint advection).  Device code for gfx950.
#pragma once

#include <hip/hip_runtime.h>

namespace fcfd {

// WENO5-JS nonlinear weights w_k = a_k / sum a, a_k = d_k / (EPS + beta_k)^2,
// d = (1, 6, 3) (level_set_rk3_solver.f90:170-216).  Multiplying every a_k
// by q1 q2 q3 (q_k = (EPS + beta_k)^2) leaves the weights unchanged and needs
// one division instead of six: w_k = (d_k prod_{m != k} q_m) / sum_l (...).
// With |phi| and |u| of O(1) the products stay far inside the FP64 range
// (q >= EPS^2 = 1e-26, so a product of two q is >= 1e-52).
__device__ __forceinline__ void weno_weights(double b1, double b2, double b3, double EPS, double &w1, double &w2,
                                             double &w3) {
    const double q1 = (EPS + b1) * (EPS + b1), q2 = (EPS + b2) * (EPS + b2), q3 = (EPS + b3) * (EPS + b3);
    const double a1 = q2 * q3, a2 = 6.0 * (q1 * q3), a3 = 3.0 * (q1 * q2);
    const double r = 1.0 / (a1 + a2 + a3);
    w1 = a1 * r;
    w2 = a2 * r;
    w3 = a3 * r;
}

// WENO5-JS left-biased face value from g(i-2..i+2) (level_set_rk3_solver.f90:170-191)
__device__ __forceinline__ double weno_left(double g0, double g1, double g2, double g3, double g4, double EPS) {
    double t1 = g0 - 2.0 * g1 + g2, t2 = g0 - 4.0 * g1 + 3.0 * g2;
    const double b1 = 13.0 * (t1 * t1) + 3.0 * (t2 * t2);
    t1 = g1 - 2.0 * g2 + g3; t2 = g1 - g3;
    const double b2 = 13.0 * (t1 * t1) + 3.0 * (t2 * t2);
    t1 = g2 - 2.0 * g3 + g4; t2 = 3.0 * g2 - 4.0 * g3 + g4;
    const double b3 = 13.0 * (t1 * t1) + 3.0 * (t2 * t2);
    double w1, w2, w3;
    weno_weights(b1, b2, b3, EPS, w1, w2, w3);
    constexpr double r6 = 1.0 / 6.0;
    return w1 * (1.0 / 3.0) * g0 - (7.0 * w1 + w2) * r6 * g1 + (11.0 * w1 + 5.0 * w2 + 2.0 * w3) * r6 * g2 +
           (2.0 * w2 + 5.0 * w3) * r6 * g3 - w3 * r6 * g4;
}

// WENO5-JS right-biased face value from f(i-1..i+3) (level_set_rk3_solver.f90:193-216)
__device__ __forceinline__ double weno_right(double f0, double f1, double f2, double f3, double f4, double EPS) {
    double t1 = f0 - 2.0 * f1 + f2, t2 = f0 - 4.0 * f1 + 3.0 * f2;
    const double b3 = 13.0 * (t1 * t1) + 3.0 * (t2 * t2);
    t1 = f1 - 2.0 * f2 + f3; t2 = f1 - f3;
    const double b2 = 13.0 * (t1 * t1) + 3.0 * (t2 * t2);
    t1 = f2 - 2.0 * f3 + f4; t2 = 3.0 * f2 - 4.0 * f3 + f4;
    const double b1 = 13.0 * (t1 * t1) + 3.0 * (t2 * t2);
    double w1, w2, w3;
    weno_weights(b1, b2, b3, EPS, w1, w2, w3);
    constexpr double r6 = 1.0 / 6.0;
    return w3 * (-f0 + 5.0 * f1 + 2.0 * f2) * r6 + w2 * (2.0 * f1 + 5.0 * f2 - f3) * r6 +
           w1 * (11.0 * f2 - 7.0 * f3 + 2.0 * f4) * r6;
}

// FP32 forms for the single-precision advection.  The weights use
// a_k = d_k / q_k: q_k >= EPS^2 = 1e-26 is a normal float, while the product
// of two q of the FP64 form (>= 1e-52) would underflow to zero where the
// flux is locally flat (beta = 0 on all three stencils).
__device__ __forceinline__ void weno_weights(float b1, float b2, float b3, float EPS, float &w1, float &w2,
                                             float &w3) {
    const float a1 = 1.0f / ((EPS + b1) * (EPS + b1)), a2 = 6.0f / ((EPS + b2) * (EPS + b2)),
                a3 = 3.0f / ((EPS + b3) * (EPS + b3));
    const float r = 1.0f / (a1 + a2 + a3);
    w1 = a1 * r;
    w2 = a2 * r;
    w3 = a3 * r;
}

__device__ __forceinline__ float weno_left(float g0, float g1, float g2, float g3, float g4, float EPS) {
    float t1 = g0 - 2.0f * g1 + g2, t2 = g0 - 4.0f * g1 + 3.0f * g2;
    const float b1 = 13.0f * (t1 * t1) + 3.0f * (t2 * t2);
    t1 = g1 - 2.0f * g2 + g3; t2 = g1 - g3;
    const float b2 = 13.0f * (t1 * t1) + 3.0f * (t2 * t2);
    t1 = g2 - 2.0f * g3 + g4; t2 = 3.0f * g2 - 4.0f * g3 + g4;
    const float b3 = 13.0f * (t1 * t1) + 3.0f * (t2 * t2);
    float w1, w2, w3;
    weno_weights(b1, b2, b3, EPS, w1, w2, w3);
    constexpr float r6 = 1.0f / 6.0f;
    return w1 * (1.0f / 3.0f) * g0 - (7.0f * w1 + w2) * r6 * g1 + (11.0f * w1 + 5.0f * w2 + 2.0f * w3) * r6 * g2 +
           (2.0f * w2 + 5.0f * w3) * r6 * g3 - w3 * r6 * g4;
}

__device__ __forceinline__ float weno_right(float f0, float f1, float f2, float f3, float f4, float EPS) {
    float t1 = f0 - 2.0f * f1 + f2, t2 = f0 - 4.0f * f1 + 3.0f * f2;
    const float b3 = 13.0f * (t1 * t1) + 3.0f * (t2 * t2);
    t1 = f1 - 2.0f * f2 + f3; t2 = f1 - f3;
    const float b2 = 13.0f * (t1 * t1) + 3.0f * (t2 * t2);
    t1 = f2 - 2.0f * f3 + f4; t2 = 3.0f * f2 - 4.0f * f3 + f4;
    const float b1 = 13.0f * (t1 * t1) + 3.0f * (t2 * t2);
    float w1, w2, w3;
    weno_weights(b1, b2, b3, EPS, w1, w2, w3);
    constexpr float r6 = 1.0f / 6.0f;
    return w3 * (-f0 + 5.0f * f1 + 2.0f * f2) * r6 + w2 * (2.0f * f1 + 5.0f * f2 - f3) * r6 +
           w1 * (11.0f * f2 - 7.0f * f3 + 2.0f * f4) * r6;
}

// ---------------------------------------------------------------- adjoints
// Reverse mode of weno_weights: given wbar_k = dJ/dw_k, the dJ/dbeta_k.
// w_k = a_k r with r = 1 / sum a, a1 = q2 q3, a2 = 6 q1 q3, a3 = 3 q1 q2,
// q_k = (EPS + b_k)^2.
__device__ __forceinline__ void weno_weights_adj(double b1, double b2, double b3, double EPS, double wb1,
                                                 double wb2, double wb3, double &bb1, double &bb2, double &bb3) {
    const double e1 = EPS + b1, e2 = EPS + b2, e3 = EPS + b3;
    const double q1 = e1 * e1, q2 = e2 * e2, q3 = e3 * e3;
    const double a1 = q2 * q3, a2 = 6.0 * (q1 * q3), a3 = 3.0 * (q1 * q2);
    const double r = 1.0 / (a1 + a2 + a3);
    const double rb = wb1 * a1 + wb2 * a2 + wb3 * a3;
    const double sb = -rb * r * r;
    const double ab1 = wb1 * r + sb, ab2 = wb2 * r + sb, ab3 = wb3 * r + sb;
    const double qb1 = 6.0 * ab2 * q3 + 3.0 * ab3 * q2;
    const double qb2 = ab1 * q3 + 3.0 * ab3 * q1;
    const double qb3 = ab1 * q2 + 6.0 * ab2 * q1;
    bb1 = 2.0 * e1 * qb1;
    bb2 = 2.0 * e2 * qb2;
    bb3 = 2.0 * e3 * qb3;
}

// FP32 form of weno_weights_adj for the weights of the FP32 forward,
// a_k = d_k / q_k (same function, other arithmetic): da_k/db_k = -2 a_k / e_k.
__device__ __forceinline__ void weno_weights_adj(float b1, float b2, float b3, float EPS, float wb1, float wb2,
                                                 float wb3, float &bb1, float &bb2, float &bb3) {
    const float e1 = EPS + b1, e2 = EPS + b2, e3 = EPS + b3;
    const float a1 = 1.0f / (e1 * e1), a2 = 6.0f / (e2 * e2), a3 = 3.0f / (e3 * e3);
    const float r = 1.0f / (a1 + a2 + a3);
    const float rb = wb1 * a1 + wb2 * a2 + wb3 * a3;
    const float sb = -rb * r * r;
    const float ab1 = wb1 * r + sb, ab2 = wb2 * r + sb, ab3 = wb3 * r + sb;
    bb1 = -2.0f * ab1 * a1 / e1;
    bb2 = -2.0f * ab2 * a2 / e2;
    bb3 = -2.0f * ab3 * a3 / e3;
}

// d(weno_left)/dg_m times vb, into gb[0..4]: the value is
// w1 P1 + w2 P2 + w3 P3 with P1 = (2g0 - 7g1 + 11g2)/6, P2 = (-g1 + 5g2 + 2g3)/6,
// P3 = (2g2 + 5g3 - g4)/6, the weights from the smoothness indicators b_k.
// T = double or float (the weights of that precision's forward).
template <typename T>
__device__ __forceinline__ void weno_left_adj(const T *g, T EPS, T vb, T *gb) {
    const T c2 = 2, c3 = 3, c4 = 4, c5 = 5, c6 = 6, c7 = 7, c11 = 11, c13 = 13, c26 = 26;
    const T t11 = g[0] - c2 * g[1] + g[2], t12 = g[0] - c4 * g[1] + c3 * g[2];
    const T t21 = g[1] - c2 * g[2] + g[3], t22 = g[1] - g[3];
    const T t31 = g[2] - c2 * g[3] + g[4], t32 = c3 * g[2] - c4 * g[3] + g[4];
    const T b1 = c13 * (t11 * t11) + c3 * (t12 * t12);
    const T b2 = c13 * (t21 * t21) + c3 * (t22 * t22);
    const T b3 = c13 * (t31 * t31) + c3 * (t32 * t32);
    T w1, w2, w3;
    weno_weights(b1, b2, b3, EPS, w1, w2, w3);
    const T r6 = (T)1 / c6;
    const T P1 = (c2 * g[0] - c7 * g[1] + c11 * g[2]) * r6;
    const T P2 = (-g[1] + c5 * g[2] + c2 * g[3]) * r6;
    const T P3 = (c2 * g[2] + c5 * g[3] - g[4]) * r6;
    T bb1, bb2, bb3;
    weno_weights_adj(b1, b2, b3, EPS, vb * P1, vb * P2, vb * P3, bb1, bb2, bb3);
    const T p1 = vb * w1 * r6, p2 = vb * w2 * r6, p3 = vb * w3 * r6;
    // through the candidate stencils
    gb[0] = c2 * p1;
    gb[1] = -c7 * p1 - p2;
    gb[2] = c11 * p1 + c5 * p2 + c2 * p3;
    gb[3] = c2 * p2 + c5 * p3;
    gb[4] = -p3;
    // through the smoothness indicators
    const T u11 = c26 * t11 * bb1, u12 = c6 * t12 * bb1;
    const T u21 = c26 * t21 * bb2, u22 = c6 * t22 * bb2;
    const T u31 = c26 * t31 * bb3, u32 = c6 * t32 * bb3;
    gb[0] += u11 + u12;
    gb[1] += -c2 * u11 - c4 * u12 + u21 + u22;
    gb[2] += u11 + c3 * u12 - c2 * u21 + u31 + c3 * u32;
    gb[3] += u21 - u22 - c2 * u31 - c4 * u32;
    gb[4] += u31 + u32;
}

// d(weno_right)/df_m times vb, into fb[0..4]: the value is
// w3 R3 + w2 R2 + w1 R1 with R3 = (-f0 + 5f1 + 2f2)/6, R2 = (2f1 + 5f2 - f3)/6,
// R1 = (11f2 - 7f3 + 2f4)/6; b3 from f0..f2, b2 from f1..f3, b1 from f2..f4.
template <typename T>
__device__ __forceinline__ void weno_right_adj(const T *f, T EPS, T vb, T *fb) {
    const T c2 = 2, c3 = 3, c4 = 4, c5 = 5, c6 = 6, c7 = 7, c11 = 11, c13 = 13, c26 = 26;
    const T t31 = f[0] - c2 * f[1] + f[2], t32 = f[0] - c4 * f[1] + c3 * f[2];
    const T t21 = f[1] - c2 * f[2] + f[3], t22 = f[1] - f[3];
    const T t11 = f[2] - c2 * f[3] + f[4], t12 = c3 * f[2] - c4 * f[3] + f[4];
    const T b3 = c13 * (t31 * t31) + c3 * (t32 * t32);
    const T b2 = c13 * (t21 * t21) + c3 * (t22 * t22);
    const T b1 = c13 * (t11 * t11) + c3 * (t12 * t12);
    T w1, w2, w3;
    weno_weights(b1, b2, b3, EPS, w1, w2, w3);
    const T r6 = (T)1 / c6;
    const T R3 = (-f[0] + c5 * f[1] + c2 * f[2]) * r6;
    const T R2 = (c2 * f[1] + c5 * f[2] - f[3]) * r6;
    const T R1 = (c11 * f[2] - c7 * f[3] + c2 * f[4]) * r6;
    T bb1, bb2, bb3;
    weno_weights_adj(b1, b2, b3, EPS, vb * R1, vb * R2, vb * R3, bb1, bb2, bb3);
    const T p1 = vb * w1 * r6, p2 = vb * w2 * r6, p3 = vb * w3 * r6;
    fb[0] = -p3;
    fb[1] = c5 * p3 + c2 * p2;
    fb[2] = c2 * p3 + c5 * p2 + c11 * p1;
    fb[3] = -p2 - c7 * p1;
    fb[4] = c2 * p1;
    const T u31 = c26 * t31 * bb3, u32 = c6 * t32 * bb3;
    const T u21 = c26 * t21 * bb2, u22 = c6 * t22 * bb2;
    const T u11 = c26 * t11 * bb1, u12 = c6 * t12 * bb1;
    fb[0] += u31 + u32;
    fb[1] += -c2 * u31 - c4 * u32 + u21 + u22;
    fb[2] += u31 + c3 * u32 - c2 * u21 + u11 + c3 * u12;
    fb[3] += u21 - u22 - c2 * u11 - c4 * u12;
    fb[4] += u11 + u12;
}

}  // namespace fcfd
