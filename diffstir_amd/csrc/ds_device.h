// ds_device.h -- device-side building blocks of the MLS-MPM stirring substep (gfx950).
//
// Shared by every kernel of libdiffstir_hip.so: quadratic B-spline weights, the three
// rheology models and their partial derivatives, the FP64 3x3 signed SVD and its backward,
// and the per-node grid update with its adjoint.  FP32 everywhere except the SVD, its
// backward and the pressure term built from J (FP64).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ds {

enum Species : int { LIQUID = 0, STICK = 1, CONTAINER = 2 };

// State layout: component-major SoA over all particles of all instances, 24 floats each.
enum StateSlot : int { SX = 0, SV = 3, SC = 6, SF = 15, NSLOT = 24 };

// One contiguous run of particles of one (instance, species); runs start on a multiple of
// 64 so a wavefront never mixes instances or species.
struct Group {
    int inst;
    int species;
    int offset;
    int count;
};

// Scene constants in the form the kernels use.
struct DevScene {
    int n_grid, bound, model, k_from_param;
    float dx, inv_dx, dt;
    float gdt;          // dt * gravity: added to the liquid's y velocity (simulation_functions.py:152)
    float wall_r2, wall_cx, wall_cz;
    float m_liq, m_stick, m_cont;
    float alpha;        // -dt * vol_liq * 4 * inv_dx^2 (simulation_functions.py:150)
    float c4;           // 4 * inv_dx (simulation_functions.py:244)
    double k_bulk;
    // tiled kernels: a stencil uses its chunk's LDS tile when its base lies within tile_win cells
    // of the tile origin on every axis (largest 5 = tile nodes - 3); a smaller window sends more
    // stencils to the global fall-back grid (a diagnostic; the results are the same)
    int tile_win;
    int count_fallback;   // 1: P2G counts its fall-back stencils (flag words 2-3, a 64-bit counter)
};

// ---------------------------------------------------------------------------- B-spline

// Quadratic B-spline weights of one particle (simulation_functions.py:108-110): base is the
// truncating cast of x/dx - 0.5.  Returns false when the 3x3x3 stencil leaves the grid or the
// position is not finite; base is then clamped so every access stays in bounds.
struct Stencil {
    int b[3];
    float fx[3];
    float w[3][3];
    float dw[3][3];   // d w / d fx
};

__device__ __forceinline__ bool stencil(const DevScene& sc, float x0, float x1, float x2, Stencil& s)
{
    const float xs[3] = {x0, x1, x2};
    bool ok = true;
#pragma unroll
    for (int a = 0; a < 3; a++) {
        float X = xs[a] * sc.inv_dx;
        bool fin = isfinite(X);
        int b = fin ? (int)(X - 0.5f) : 0;
        if (!fin || b < 0 || b > sc.n_grid - 3) {
            ok = false;
            b = b < 0 ? 0 : (b > sc.n_grid - 3 ? sc.n_grid - 3 : b);
            if (!fin) X = (float)b + 1.0f;
        }
        s.b[a] = b;
        float f = X - (float)b;
        s.fx[a] = f;
        s.w[a][0] = 0.5f * (1.5f - f) * (1.5f - f);
        s.w[a][1] = 0.75f - (f - 1.0f) * (f - 1.0f);
        s.w[a][2] = 0.5f * (f - 0.5f) * (f - 0.5f);
        s.dw[a][0] = f - 1.5f;
        s.dw[a][1] = -2.0f * (f - 1.0f);
        s.dw[a][2] = f - 0.5f;
    }
    return ok;
}

// ---------------------------------------------------------------------------- rheology

// eta(gamma; params) with its partial derivatives (simulation_functions.py:119-146).
struct Visc {
    float eta;
    float d_dg;
    float d_dth[4];
};

__device__ __forceinline__ Visc viscosity(int model, const float th[4], float g)
{
    Visc r;
    r.d_dth[0] = r.d_dth[1] = r.d_dth[2] = r.d_dth[3] = 0.0f;
    if (model == 0) {           // Carreau
        float e0 = th[0], ei = th[1], n = th[2], lam = th[3];
        float lg = lam * g;
        float q = 1.0f + lg * lg;
        float e = (n - 1.0f) * 0.5f;
        float lq = __logf(q);
        float qe = __expf(e * lq);
        float dq = e * qe / q;              // d q^e / dq
        r.eta = ei + (e0 - ei) * qe;
        r.d_dth[0] = qe;
        r.d_dth[1] = 1.0f - qe;
        r.d_dth[2] = (e0 - ei) * qe * lq * 0.5f;
        r.d_dth[3] = (e0 - ei) * dq * 2.0f * lg * g;
        r.d_dg = (e0 - ei) * dq * 2.0f * lg * lam;
    } else if (model == 1) {    // Cross (exponent n - 1, simulation_functions.py:134)
        float e0 = th[0], ei = th[1], n = th[2], lam = th[3];
        float lg = lam * g;
        float llg = __logf(lg);
        float rr = __expf((n - 1.0f) * llg);
        float D = 1.0f + rr;
        float iD = 1.0f / D;
        float c = -(e0 - ei) * iD * iD;
        float dr = (n - 1.0f) * rr / lg;    // d rr / d lg
        r.eta = ei + (e0 - ei) * iD;
        r.d_dth[0] = iD;
        r.d_dth[1] = 1.0f - iD;
        r.d_dth[2] = c * rr * llg;
        r.d_dth[3] = c * dr * g;
        r.d_dg = c * dr * lam;
    } else {                    // Herschel-Bulkley
        float K = th[0], ty = th[1], n = th[2];
        float lgm = __logf(g);
        float gn = __expf((n - 1.0f) * lgm);
        float de = K * gn;
        float ig = 1.0f / g;
        r.eta = ty * ig;
        r.d_dth[1] = ig;
        r.d_dg = -ty * ig * ig;
        if (de > 0.0f) {
            r.eta += de;
            r.d_dth[0] = gn;
            r.d_dth[2] = de * lgm;
            r.d_dg += (n - 1.0f) * de * ig;
        }
    }
    return r;
}

// ---------------------------------------------------------------------------- FP64 SVD

// FP64 reciprocal and reciprocal square root: the hardware estimates (v_rcp_f64, v_rsq_f64)
// refined by two Newton-Raphson steps each.  Every step doubles the number of correct bits,
// so any estimate good to 2^-14 ends within an ulp or two of the exact value.  For positive,
// normal, finite arguments only (no IEEE special-case handling): the SVD below keeps its
// arguments there or discards the result.
__device__ __forceinline__ double rcp64(double x)
{
    double r = __builtin_amdgcn_rcp(x);
    double e = fma(-x, r, 1.0);
    r = fma(r, e, r);
    e = fma(-x, r, 1.0);
    return fma(r, e, r);
}

__device__ __forceinline__ double rsqrt64(double a)
{
    double y = __builtin_amdgcn_rsq(a);
    const double h = 0.5 * a;
    double t = fma(-h * y, y, 0.5);     // y <- y (3 - a y^2) / 2
    y = fma(y, t, y);
    t = fma(-h * y, y, 0.5);
    return fma(y, t, y);
}

// One cyclic-Jacobi rotation on the symmetric 3x3 A, annihilating A[p][q] and (kV)
// accumulating the rotation into the columns of V (Golub & Van Loan, Matrix Computations,
// 4th ed., sec. 8.5.2).  With d = a_qq - a_pp and e = 2 a_pq, the tangent of the rotation
// angle is the smaller root of t^2 + 2 (d/e) t - 1 = 0, t = sgn(d) e / (|d| + sqrt(d^2 + e^2)),
// written without the quotient d/e; c = 1/sqrt(1 + t^2), s = t c.
template <int P, int Q, bool kV>
__device__ __forceinline__ void jacobi_rot(double A[3][3], double V[3][3])
{
    constexpr int R = 3 - P - Q;
    const double apq = A[P][Q];
    const double d = A[Q][Q] - A[P][P];
    const double e = 2.0 * apq;
    const double r2 = fma(d, d, e * e);
    const double den = fabs(d) + r2 * rsqrt64(r2);
    const double t = r2 > 0.0 ? (d >= 0.0 ? e : -e) * rcp64(den) : 0.0;
    const double c = rsqrt64(fma(t, t, 1.0));
    const double s = t * c;
    A[P][P] -= t * apq;
    A[Q][Q] += t * apq;
    A[P][Q] = A[Q][P] = 0.0;
    const double arp = A[R][P], arq = A[R][Q];
    A[R][P] = A[P][R] = c * arp - s * arq;
    A[R][Q] = A[Q][R] = s * arp + c * arq;
    if (kV) {
#pragma unroll
        for (int i = 0; i < 3; i++) {
            const double vp = V[i][P], vq = V[i][Q];
            V[i][P] = c * vp - s * vq;
            V[i][Q] = s * vp + c * vq;
        }
    }
}

// A = F^T F diagonalised by cyclic Jacobi sweeps (at most 8; stop once the off-diagonal mass is
// below 1e-32 of the diagonal's, i.e. relative 1e-16).
template <bool kV>
__device__ __forceinline__ void jacobi_eig_FtF(const double F[9], double A[3][3], double V[3][3])
{
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++)
            A[i][j] = F[0 * 3 + i] * F[0 * 3 + j] + F[1 * 3 + i] * F[1 * 3 + j] + F[2 * 3 + i] * F[2 * 3 + j];
    for (int sweep = 0; sweep < 8; sweep++) {
        const double off = A[0][1] * A[0][1] + A[0][2] * A[0][2] + A[1][2] * A[1][2];
        const double dia = A[0][0] * A[0][0] + A[1][1] * A[1][1] + A[2][2] * A[2][2];
        if (off <= 1e-32 * dia) break;
        jacobi_rot<0, 1, kV>(A, V);
        jacobi_rot<0, 2, kV>(A, V);
        jacobi_rot<1, 2, kV>(A, V);
    }
}

__device__ __forceinline__ void swap_cols(double V[3][3], double S[3], int a, int b)
{
    double t = S[a]; S[a] = S[b]; S[b] = t;
#pragma unroll
    for (int i = 0; i < 3; i++) {
        double u = V[i][a];
        V[i][a] = V[i][b];
        V[i][b] = -u;   // a swap with one sign flip keeps det V = +1
    }
}

// Signed SVD in FP64: F = U diag(S) V^T with U, V rotations, S[0] >= S[1] >= |S[2]| and
// S[2] < 0 when det F < 0.  Cyclic Jacobi on F^T F gives V and S^2; U comes from a
// Gram-Schmidt QR of F V, the construction of McAdams et al., "Computing the Singular Value
// Decomposition of 3x3 matrices with minimal branching and elementary floating point
// operations" (2011), here with exact FP64 rotations.  S[0]*S[1]*S[2] = det F for any V
// because F V = U R with R upper triangular and diag(R) = S.
__device__ __forceinline__ void svd3(const double F[9], double U[9], double S[3], double Vo[9])
{
    double A[3][3];
    double V[3][3] = {{1.0, 0.0, 0.0}, {0.0, 1.0, 0.0}, {0.0, 0.0, 1.0}};
    jacobi_eig_FtF<true>(F, A, V);
    double L[3] = {A[0][0], A[1][1], A[2][2]};
    if (L[0] < L[1]) swap_cols(V, L, 0, 1);
    if (L[0] < L[2]) swap_cols(V, L, 0, 2);
    if (L[1] < L[2]) swap_cols(V, L, 1, 2);
    // B = F V, columns b0, b1, b2
    double b[3][3];
#pragma unroll
    for (int c = 0; c < 3; c++)
#pragma unroll
        for (int i = 0; i < 3; i++)
            b[c][i] = F[i * 3 + 0] * V[0][c] + F[i * 3 + 1] * V[1][c] + F[i * 3 + 2] * V[2][c];
    double u0[3], u1[3], u2[3];
    const double nn0 = b[0][0] * b[0][0] + b[0][1] * b[0][1] + b[0][2] * b[0][2];
    double n0 = 0.0;
    if (nn0 > 1e-280) {
        const double in = rsqrt64(nn0);
        n0 = nn0 * in;
        u0[0] = b[0][0] * in; u0[1] = b[0][1] * in; u0[2] = b[0][2] * in;
    } else {
        u0[0] = 1.0; u0[1] = 0.0; u0[2] = 0.0;
    }
    const double d01 = u0[0] * b[1][0] + u0[1] * b[1][1] + u0[2] * b[1][2];
    const double r1[3] = {b[1][0] - d01 * u0[0], b[1][1] - d01 * u0[1], b[1][2] - d01 * u0[2]};
    const double nn1 = r1[0] * r1[0] + r1[1] * r1[1] + r1[2] * r1[2];
    double n1 = 0.0;
    if (nn1 > 1e-280) {
        const double in = rsqrt64(nn1);
        n1 = nn1 * in;
        u1[0] = r1[0] * in; u1[1] = r1[1] * in; u1[2] = r1[2] * in;
    } else {
        // any unit vector orthogonal to u0
        const double ax = fabs(u0[0]) < 0.9 ? 1.0 : 0.0, ay = 1.0 - ax;
        const double t[3] = {-u0[2] * ay, u0[2] * ax, u0[0] * ay - u0[1] * ax};
        const double nt = rsqrt64(t[0] * t[0] + t[1] * t[1] + t[2] * t[2]);
        u1[0] = t[0] * nt; u1[1] = t[1] * nt; u1[2] = t[2] * nt;
    }
    u2[0] = u0[1] * u1[2] - u0[2] * u1[1];
    u2[1] = u0[2] * u1[0] - u0[0] * u1[2];
    u2[2] = u0[0] * u1[1] - u0[1] * u1[0];
    S[0] = n0;
    S[1] = n1;
    S[2] = u2[0] * b[2][0] + u2[1] * b[2][1] + u2[2] * b[2][2];
#pragma unroll
    for (int i = 0; i < 3; i++) {
        U[i * 3 + 0] = u0[i];
        U[i * 3 + 1] = u1[i];
        U[i * 3 + 2] = u2[i];
        Vo[i * 3 + 0] = V[i][0];
        Vo[i * 3 + 1] = V[i][1];
        Vo[i * 3 + 2] = V[i][2];
    }
}

// The singular values alone, in the same signed convention as svd3: S[0], S[1] = sqrt of the two
// largest eigenvalues of A = F^T F, S[2] = det F / (S[0] S[1]) (so the smallest carries the sign
// of det F, and S[0]*S[1]*S[2] = det F up to rounding).  The eigenvalues come in closed form (O. K. Smith,
// "Eigenvalues of a symmetric 3 x 3 matrix", Comm. ACM 4(4), 1961): with q = tr(A)/3,
// B = A - q I, p = sqrt(tr(B^2)/6) and r = det(B)/(2 p^3) in [-1, 1], they are
// q + 2 p cos(phi + 2 pi k / 3), k = 0, 1, 2, phi = acos(r)/3 (cos(phi) from the triple-angle
// cubic below, without trigonometric calls).  B is formed exactly enough
// for F near a rotation (a_ii - q loses nothing), and the sum of the three is 3q, so the
// middle one is taken from the other two.  Errors are of order 1e-16 of the largest
// eigenvalue; where two eigenvalues nearly coincide their individual errors grow (to about
// 1e-12 of the largest for a coincident pair at the top, more for a coincident pair below a
// much larger one), which leaves the product with S[2] below unaffected.
__device__ __forceinline__ void svd3_values(const double F[9], double S[3])
{
    double A[3][3];
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = i; j < 3; j++)
            A[i][j] = F[0 * 3 + i] * F[0 * 3 + j] + F[1 * 3 + i] * F[1 * 3 + j] + F[2 * 3 + i] * F[2 * 3 + j];
    const double a01 = A[0][1], a02 = A[0][2], a12 = A[1][2];
    const double q = (A[0][0] + A[1][1] + A[2][2]) * (1.0 / 3.0);
    const double b00 = A[0][0] - q, b11 = A[1][1] - q, b22 = A[2][2] - q;
    const double off = a01 * a01 + a02 * a02 + a12 * a12;
    const double p2 = (b00 * b00 + b11 * b11 + b22 * b22 + 2.0 * off) * (1.0 / 6.0);
    double l0 = q, l1 = q, l2 = q;
    if (p2 > 0.0) {
        const double p = sqrt(p2);
        const double detB = b00 * (b11 * b22 - a12 * a12) - a01 * (a01 * b22 - a12 * a02) + a02 * (a01 * a12 - b11 * a02);
        const double r = fmin(1.0, fmax(-1.0, detB * rcp64(2.0 * p2 * p)));
        // c = cos(acos(r) / 3), the root in [1/2, 1] of 4 c^3 - 3 c = r, by Newton's method from
        // above (the cubic is convex and increasing there, so the iterates fall monotonically);
        // the start 1/2 + sqrt((1 + r) / 6) is the root's expansion at r = -1 and lies above it
        double c = 0.5 + sqrt((1.0 + r) * (1.0 / 6.0));
#pragma unroll
        for (int it = 0; it < 5; it++) {
            const double g = (4.0 * c * c - 3.0) * c - r;
            const double gp = 12.0 * c * c - 3.0;
            c -= gp > 0.0 ? g * rcp64(gp) : 0.0;
        }
        // cos(phi + 2 pi / 3) = -c / 2 - (sqrt(3) / 2) sin(phi), sin(phi) = sqrt(1 - c^2)
        const double sn = sqrt(fmax(0.0, 1.0 - c * c));
        l0 = q + 2.0 * p * c;
        l2 = q - p * (c + 1.7320508075688772 * sn);
        l1 = 3.0 * q - l0 - l2;
    }
    l0 = fmax(l0, 0.0);
    l1 = fmax(l1, 0.0);
    l2 = fmax(l2, 0.0);
    const double hi = fmax(fmax(l0, l1), l2);
    const double mid = fmax(fmin(l0, l1), fmin(fmax(l0, l1), l2));
    const double detF = F[0] * (F[4] * F[8] - F[5] * F[7]) - F[1] * (F[3] * F[8] - F[5] * F[6])
                      + F[2] * (F[3] * F[7] - F[4] * F[6]);
    S[0] = hi > 0.0 ? hi * rsqrt64(hi) : 0.0;
    S[1] = mid > 0.0 ? mid * rsqrt64(mid) : 0.0;
    // the smallest from |det F| = S0 S1 S2: the closed form's eigenvalues carry errors of order
    // 1e-16 of the largest, which the smallest cannot afford when F is far from a rotation (a
    // liquid sheared by a long rollout: S0 / S2 of 1e4 and more); this keeps S2, and the product
    // J that the stress uses, accurate to rounding; capped at S1, which it can only pass when F
    // is numerically rank-deficient (S1 then holds rounding noise and |det F| / S0 S1 is noise
    // divided by noise)
    const double s01 = S[0] * S[1];
    const double s2 = s01 > 0.0 ? fmin(fabs(detF) / s01, S[1]) : 0.0;
    S[2] = detF < 0.0 ? -s2 : s2;
}

// clamp of simulation_functions.py:61-67
__device__ __forceinline__ double svd_clamp(double a)
{
    return a >= 0.0 ? fmax(a, 1e-6) : fmin(a, -1e-6);
}

// SVD backward of simulation_functions.py:70-87 (PyTorch's formula):
//   gF = U (f o (U^T gU - gU^T U)) S V^T + U S (f o (V^T gV - gV^T V)) V^T + U diag(gS) V^T
// with f_ij = 1 / clamp(s_j^2 - s_i^2), f_ii = 0.
__device__ __forceinline__ void svd3_backward(const double U[9], const double S[3], const double V[9],
                                              const double gU[9], const double gS[3], const double gV[9],
                                              double gF[9])
{
    double s2[3] = {S[0] * S[0], S[1] * S[1], S[2] * S[2]};
    double f[3][3];
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) f[i][j] = (i == j) ? 0.0 : 1.0 / svd_clamp(s2[j] - s2[i]);
    double Ku[3][3], Kv[3][3];
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) {
            double a = 0.0, bb = 0.0, c = 0.0, d = 0.0;
#pragma unroll
            for (int k = 0; k < 3; k++) {
                a += U[k * 3 + i] * gU[k * 3 + j];     // (U^T gU)_ij
                bb += gU[k * 3 + i] * U[k * 3 + j];    // (gU^T U)_ij
                c += V[k * 3 + i] * gV[k * 3 + j];     // (V^T gV)_ij
                d += gV[k * 3 + i] * V[k * 3 + j];     // (gV^T V)_ij
            }
            Ku[i][j] = f[i][j] * (a - bb);
            Kv[i][j] = f[i][j] * (c - d);
        }
    // inner = Ku S + S Kv + diag(gS)
    double M[3][3];
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) M[i][j] = Ku[i][j] * S[j] + S[i] * Kv[i][j] + (i == j ? gS[i] : 0.0);
    // gF = U M V^T
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) {
            double acc = 0.0;
#pragma unroll
            for (int k = 0; k < 3; k++) {
                double um = U[i * 3 + 0] * M[0][k] + U[i * 3 + 1] * M[1][k] + U[i * 3 + 2] * M[2][k];
                acc += um * V[j * 3 + k];
            }
            gF[i * 3 + j] = acc;
        }
}

// The same backward when only the singular values carry an adjoint (gU = gV = 0): the
// u- and v-terms of svd3_backward vanish identically, leaving U diag(gS) V^T.
__device__ __forceinline__ void svd3_backward_sigma(const double U[9], const double V[9], const double gS[3],
                                                    double gF[9])
{
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++)
            gF[i * 3 + j] = U[i * 3 + 0] * gS[0] * V[j * 3 + 0] + U[i * 3 + 1] * gS[1] * V[j * 3 + 1]
                          + U[i * 3 + 2] * gS[2] * V[j * 3 + 2];
}

// The SVD backward of J = S0 S1 S2 up to the factor dL/dJ: with gS = dL/dJ (S1 S2, S0 S2, S0 S1),
// U diag(gS) V^T = dL/dJ * M, M = U diag(S1 S2, S0 S2, S0 S1) V^T.
__device__ __forceinline__ void svd3_j_factor(const double U[9], const double S[3], const double V[9], double M[9])
{
    const double g[3] = {S[1] * S[2], S[0] * S[2], S[0] * S[1]};
    svd3_backward_sigma(U, V, g, M);
}

// ---------------------------------------------------------------------------- liquid stress

// Everything P2G needs from one liquid particle (simulation_functions.py:111-153):
// F_tmp = (I + dt C) F, J from the FP64 SVD of F_tmp, gamma = |C|_F + 1e-6, eta(gamma),
// S = eta sym(C) + I (k/2)(J - 1/J - 1) J, affine A = alpha S + m C.
struct LiquidP2G {
    float Fn[9];
    float A[9];
    float eta;
    float gamma;
    double J;
    double U[9], S[3], V[9];
};

// F_tmp = (I + dt C) F in FP32 (simulation_functions.py:39-42), the F of the next step.
__device__ __forceinline__ void f_tmp(const DevScene& sc, const float C[9], const float F[9], float Fn[9])
{
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) {
            float s = F[i * 3 + j];
#pragma unroll
            for (int k = 0; k < 3; k++) s = fmaf(sc.dt * C[i * 3 + k], F[k * 3 + j], s);
            Fn[i * 3 + j] = s;
        }
}

// J = S0 S1 S2 of the FP64 SVD of F_tmp (the values-only form).
__device__ __forceinline__ double liquid_j(const float Fn[9])
{
    double Fd[9], S[3];
#pragma unroll
    for (int i = 0; i < 9; i++) Fd[i] = (double)Fn[i];
    svd3_values(Fd, S);
    return S[0] * S[1] * S[2];
}

// J and the SVD backward factor M of J (svd3_j_factor) from the full FP64 SVD of F_tmp.
__device__ __forceinline__ double liquid_j_factor(const float Fn[9], double M[9])
{
    double Fd[9], U[9], S[3], V[9];
#pragma unroll
    for (int i = 0; i < 9; i++) Fd[i] = (double)Fn[i];
    svd3(Fd, U, S, V);
    svd3_j_factor(U, S, V, M);
    return S[0] * S[1] * S[2];
}

// The stress half of liquid_p2g for a given J: gamma, eta(gamma) and A.
__device__ __forceinline__ void liquid_stress(const DevScene& sc, const float th[4], double kb, const float C[9],
                                              double J, LiquidP2G& q)
{
    q.J = J;
    float n2 = 0.0f;
#pragma unroll
    for (int i = 0; i < 9; i++) n2 = fmaf(C[i], C[i], n2);
    q.gamma = sqrtf(n2) + 1e-6f;
    Visc vv = viscosity(sc.model, th, q.gamma);
    q.eta = vv.eta;
    float pr = (float)((kb * 0.5) * (J - 1.0 / J - 1.0) * J);
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) {
            float s = q.eta * 0.5f * (C[i * 3 + j] + C[j * 3 + i]) + (i == j ? pr : 0.0f);
            q.A[i * 3 + j] = sc.alpha * s + sc.m_liq * C[i * 3 + j];
        }
}

template <bool kFull>
__device__ __forceinline__ void liquid_p2g(const DevScene& sc, const float th[4], double kb, const float C[9],
                                           const float F[9], LiquidP2G& q)
{
    f_tmp(sc, C, F, q.Fn);
    double J;
    if (kFull) {
        double Fd[9];
#pragma unroll
        for (int i = 0; i < 9; i++) Fd[i] = (double)q.Fn[i];
        svd3(Fd, q.U, q.S, q.V);
        J = q.S[0] * q.S[1] * q.S[2];
    } else {
        J = liquid_j(q.Fn);
    }
    liquid_stress(sc, th, kb, C, J, q);
}

// ---------------------------------------------------------------------------- grid update

// Per-node grid update (simulation_functions.py:199-225) with the branch record the adjoint
// replays: v = p / (m + 1e-10); outside the cylinder radius the normal component is removed;
// box bounds zero outward components; the top face zeroes all three.
struct NodeRec {
    float inv_m;
    float v0[3];
    float n[3];
    bool wall;
    bool mask[3];
};

__device__ __forceinline__ void gridop_node(const DevScene& sc, float4 P, int i, int j, int k, float vo[3], NodeRec& r)
{
    r.inv_m = 1.0f / (P.w + 1e-10f);
    float v[3] = {r.inv_m * P.x, r.inv_m * P.y, r.inv_m * P.z};
    r.v0[0] = v[0]; r.v0[1] = v[1]; r.v0[2] = v[2];
    float d0 = (float)i * sc.dx - sc.wall_cx;
    float d2 = (float)k * sc.dx - sc.wall_cz;
    float dd = d0 * d0 + d2 * d2;
    r.wall = dd > sc.wall_r2;
    r.n[0] = r.n[1] = r.n[2] = 0.0f;
    if (r.wall) {
        float in = 1.0f / sqrtf(dd);
        r.n[0] = d0 * in;
        r.n[2] = d2 * in;
        float vn = v[0] * r.n[0] + v[2] * r.n[2];
        v[0] -= r.n[0] * vn;
        v[2] -= r.n[2] * vn;
    }
    int n = sc.n_grid, bd = sc.bound;
    r.mask[0] = r.mask[1] = r.mask[2] = false;
    if (i < bd && v[0] < 0.0f) { v[0] = 0.0f; r.mask[0] = true; }
    if (i > n - bd && v[0] > 0.0f) { v[0] = 0.0f; r.mask[0] = true; }
    if (k < bd && v[2] < 0.0f) { v[2] = 0.0f; r.mask[2] = true; }
    if (k > n - bd && v[2] > 0.0f) { v[2] = 0.0f; r.mask[2] = true; }
    if (j < bd && v[1] < 0.0f) { v[1] = 0.0f; r.mask[1] = true; }
    if (j > n - bd && v[1] > 0.0f) {
        v[0] = v[1] = v[2] = 0.0f;
        r.mask[0] = r.mask[1] = r.mask[2] = true;
    }
    vo[0] = v[0]; vo[1] = v[1]; vo[2] = v[2];
}

// Adjoint of gridop_node: given dL/dv_out, returns dL/d(momentum) in xyz and dL/d(mass) in w.
__device__ __forceinline__ float4 gridop_node_adj(const NodeRec& r, float g0, float g1, float g2)
{
    float g[3] = {r.mask[0] ? 0.0f : g0, r.mask[1] ? 0.0f : g1, r.mask[2] ? 0.0f : g2};
    if (r.wall) {
        float gn = g[0] * r.n[0] + g[2] * r.n[2];
        g[0] -= r.n[0] * gn;
        g[2] -= r.n[2] * gn;
    }
    float4 o;
    o.x = r.inv_m * g[0];
    o.y = r.inv_m * g[1];
    o.z = r.inv_m * g[2];
    o.w = -r.inv_m * (g[0] * r.v0[0] + g[1] * r.v0[1] + g[2] * r.v0[2]);
    return o;
}

// ---------------------------------------------------------------------------- wave reductions

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// A workgroup's FP64 partial of a sum that workgroups add with atomics, rounded to a multiple of
// 2^-32.  Sums of such values stay multiples of 2^-32, and below 2^21 in magnitude every one of
// them is exact in FP64 (53-bit significands), so the atomic adds give the same bits in any order.
// The rounding moves a partial by at most 2^-33.
__device__ __forceinline__ double exact_grain(double x)
{
    return rint(x * 4294967296.0) * (1.0 / 4294967296.0);
}

}  // namespace ds
