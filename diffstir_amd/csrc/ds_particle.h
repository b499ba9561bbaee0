// ds_particle.h -- per-particle work of the substep and its adjoint, written once and used by
// both kernel families.  Each routine takes grid accessors: global-memory accessors for the
// per-particle kernels (kernels_ref.hip), LDS-tile accessors with a global fall-back for the
// tiled kernels (kernels_tiled.hip).
//
// Accessor interface (node coordinates are global grid indices):
//   scatter:  void add(int i, int j, int k, float px, float py, float pz, float m)
//   gather3:  float3 get(int i, int j, int k)              (grid velocity or its adjoint)
//   gather4:  float4 get(int i, int j, int k)              (momentum/mass adjoints)
//   scatter3: void add(int i, int j, int k, float x, float y, float z)
#pragma once

#include "ds_device.h"

namespace ds {

struct ParticleIn {
    float x[3], v[3], C[9], F[9];
};

// Liquid P2G constants of one instance.
struct InstParams {
    float th[4];
    double kb;
};

__device__ __forceinline__ InstParams load_params(const DevScene& sc, const double* params, int inst)
{
    InstParams ip;
    const double* pp = params + 4 * inst;
#pragma unroll
    for (int a = 0; a < 4; a++) ip.th[a] = (float)pp[a];
    ip.kb = sc.k_from_param ? pp[3] : sc.k_bulk;
    return ip;
}

__device__ __forceinline__ float species_mass(const DevScene& sc, int sp)
{
    return sp == LIQUID ? sc.m_liq : (sp == STICK ? sc.m_stick : sc.m_cont);
}

// ---------------------------------------------------------------------------- stencil sums
//
// The 27 weights factor, W(i,j,k) = wx_i wy_j wz_k, and the node offset o = (i,j,k) enters every
// scatter and gather linearly (affine momenta, the C of G2P and their adjoints).  So the
// routines below form the affine terms from their value at offset 0 by adding columns, and
// reduce the gathers row by row (a row: the nodes k = 0, 1, 2 at fixed i, j) before one
// weighted update per row, instead of forming every product per node.  The offsets are the
// unrolled loop indices, so add_k costs nothing for k = 0 and one add for k = 1.

// y + k x for a node offset k in {0, 1, 2}.  Both arms are formed before the choice so that
// the choice is a select, which folds away for a constant k (fma(1, x, y) is x + y), rather
// than a branch: a chain of such branches lets the optimiser sink the arithmetic of the whole
// stencil past all 27 gathers, with every gathered value live at once.
__device__ __forceinline__ float add_k(int k, float x, float y)
{
    const float r = fmaf((float)k, x, y);
    return k == 0 ? y : r;
}

// Accumulators start at -0, the identity of IEEE addition, so the first add is exact and free.
constexpr float kZero = -0.0f;

// One row of a gathered vector g: t0 = sum_k wz_k g, t1 = sum_k k wz_k g.
struct RowSums {
    float t0[3] = {kZero, kZero, kZero};
    float t1[3] = {kZero, kZero, kZero};
    __device__ __forceinline__ void add(int k, float wz, float gx, float gy, float gz)
    {
        const float h[3] = {wz * gx, wz * gy, wz * gz};
#pragma unroll
        for (int a = 0; a < 3; a++) {
            t0[a] += h[a];
            t1[a] = add_k(k, h[a], t1[a]);
        }
    }
};

// Zeroth and first moments of a gathered vector over the stencil: m0 = sum W g,
// m1[c] = sum W g o_c.
struct Moments {
    float m0[3] = {kZero, kZero, kZero};
    float m1[3][3] = {{kZero, kZero, kZero}, {kZero, kZero, kZero}, {kZero, kZero, kZero}};
    // row (i, j) with weight wij = wx_i wy_j
    __device__ __forceinline__ void add_row(int i, int j, float wij, const RowSums& r)
    {
#pragma unroll
        for (int a = 0; a < 3; a++) {
            const float p = wij * r.t0[a];
            m0[a] += p;
            m1[0][a] = add_k(i, p, m1[0][a]);
            m1[1][a] = add_k(j, p, m1[1][a]);
            m1[2][a] = fmaf(wij, r.t1[a], m1[2][a]);
        }
    }
    // sum W g_a (o_c - fx_c): the first moment about the particle
    __device__ __forceinline__ float centred(int a, int c, const float fx[3]) const
    {
        return fmaf(-fx[c], m0[a], m1[c][a]);
    }
};

// Columns and offset-0 value of an affine field y(o) = y0 + K (o - fx) s over the stencil:
// col[c][a] = K[a][c] s, base = y0 - K fx s, so y(o) = base + sum_c o_c col[c].
struct Affine {
    float col[3][3];
    float base[3];
    __device__ __forceinline__ Affine(const float K[9], const float y0[3], float s, const float fx[3])
    {
#pragma unroll
        for (int c = 0; c < 3; c++)
#pragma unroll
            for (int a = 0; a < 3; a++) col[c][a] = K[a * 3 + c] * s;
#pragma unroll
        for (int a = 0; a < 3; a++)
            base[a] = fmaf(-col[2][a], fx[2], fmaf(-col[1][a], fx[1], fmaf(-col[0][a], fx[0], y0[a])));
    }
    // value moved by offset k along axis c
    __device__ __forceinline__ void step(int c, int k, const float in[3], float out[3]) const
    {
#pragma unroll
        for (int a = 0; a < 3; a++) out[a] = add_k(k, col[c][a], in[a]);
    }
};

// Row sums of a per-node scalar Wb for sum_nodes Wb dW, dW = (dwx wy wz, wx dwy wz, wx wy dwz):
// a row holds P = sum_k wz_k Wb, Q = sum_k dwz_k Wb and adds (dwx wy P, wx dwy P, wx wy Q).
struct WeightGrad {
    float g[3] = {kZero, kZero, kZero};
    __device__ __forceinline__ void add_row(float wx, float dwx, float wy, float dwy, float P, float Q)
    {
        g[0] = fmaf(dwx * wy, P, g[0]);
        g[1] = fmaf(wx * dwy, P, g[1]);
        g[2] = fmaf(wx * wy, Q, g[2]);
    }
};

// ---------------------------------------------------------------------------- forward

// What one particle scatters in P2G (simulation_functions.py:148-163): affine A, momentum m u
// and mass m.  For liquid, Fn receives F_tmp.
struct P2GPayload {
    float A[9];
    float mu[3];
    float m;
};

// Jin: J of the particle when already known (computed from the same F_tmp by the previous
// step's G2P), else nullptr to compute it here.
__device__ __forceinline__ void p2g_prepare(const DevScene& sc, int sp, const InstParams& ip, const ParticleIn& q,
                                            float Fn[9], P2GPayload& pl, const double* Jin = nullptr)
{
    float u[3];
    pl.m = species_mass(sc, sp);
    if (sp == LIQUID) {
        LiquidP2G L;
        if (Jin) {
            f_tmp(sc, q.C, q.F, L.Fn);
            liquid_stress(sc, ip.th, ip.kb, q.C, *Jin, L);
        } else {
            liquid_p2g<false>(sc, ip.th, ip.kb, q.C, q.F, L);
        }
#pragma unroll
        for (int a = 0; a < 9; a++) {
            Fn[a] = L.Fn[a];
            pl.A[a] = L.A[a];
        }
        u[0] = q.v[0];
        u[1] = q.v[1] + sc.gdt;
        u[2] = q.v[2];
    } else {
#pragma unroll
        for (int a = 0; a < 9; a++) pl.A[a] = pl.m * q.C[a];
        u[0] = q.v[0]; u[1] = q.v[1]; u[2] = q.v[2];
    }
#pragma unroll
    for (int a = 0; a < 3; a++) pl.mu[a] = pl.m * u[a];
}

// The 27-node scatter of P2G: w (m u + A dpos), w m with dpos = (offset - fx) dx.
template <class Scatter>
__device__ __forceinline__ void p2g_scatter(const DevScene& sc, const Stencil& s, const P2GPayload& pl, Scatter& acc)
{
    const float* A = pl.A;
#pragma unroll
    for (int i = 0; i < 3; i++) {
        const float d0 = ((float)i - s.fx[0]) * sc.dx;
#pragma unroll
        for (int j = 0; j < 3; j++) {
            const float d1 = ((float)j - s.fx[1]) * sc.dx;
            const float wij = s.w[0][i] * s.w[1][j];
#pragma unroll
            for (int k = 0; k < 3; k++) {
                const float d2 = ((float)k - s.fx[2]) * sc.dx;
                const float W = wij * s.w[2][k];
                acc.add(s.b[0] + i, s.b[1] + j, s.b[2] + k,
                        W * (pl.mu[0] + A[0] * d0 + A[1] * d1 + A[2] * d2),
                        W * (pl.mu[1] + A[3] * d0 + A[4] * d1 + A[5] * d2),
                        W * (pl.mu[2] + A[6] * d0 + A[7] * d1 + A[8] * d2), W * pl.m);
            }
        }
    }
}

// P2G of one particle (simulation_functions.py:104-196).  For liquid, Fn receives F_tmp.
template <class Scatter>
__device__ __forceinline__ void p2g_particle(const DevScene& sc, int sp, const InstParams& ip, const ParticleIn& q,
                                             const Stencil& s, float Fn[9], Scatter& acc)
{
    P2GPayload pl;
    p2g_prepare(sc, sp, ip, q, Fn, pl);
    p2g_scatter(sc, s, pl, acc);
}

// G2P gather of one particle (simulation_functions.py:237-244): nv = sum w v, nC = 4/dx sum w v (x)(o - fx).
template <class Gather>
__device__ __forceinline__ void g2p_gather(const DevScene& sc, const Stencil& s, Gather& acc, float nv[3], float nC[9])
{
    Moments M;
#pragma unroll
    for (int i = 0; i < 3; i++) {
        asm volatile("" ::: "memory");
#pragma unroll
        for (int j = 0; j < 3; j++) {
            RowSums r;
#pragma unroll
            for (int k = 0; k < 3; k++) {
                const float3 g = acc.get(s.b[0] + i, s.b[1] + j, s.b[2] + k);
                r.add(k, s.w[2][k], g.x, g.y, g.z);
            }
            M.add_row(i, j, s.w[0][i] * s.w[1][j], r);
        }
    }
#pragma unroll
    for (int a = 0; a < 3; a++) {
        nv[a] = M.m0[a];
#pragma unroll
        for (int c = 0; c < 3; c++) nC[a * 3 + c] = sc.c4 * M.centred(a, c, s.fx);
    }
}

// ---------------------------------------------------------------------------- adjoint

// Adjoint of G2P for one particle, scatter half.  vt is the adjoint carried by the gathered
// velocity (liquid: vbar' + dt xbar'; stick: -dvbar; container: 0), Cb the adjoint of the
// gathered C: dL/dv_grid(node) += w (vt + 4/dx Cb (offset - fx)).
template <class Scatter>
__device__ __forceinline__ void g2p_adj_scatter(const DevScene& sc, const Stencil& s, const float vt[3],
                                                const float Cb[9], Scatter& acc)
{
    const Affine e(Cb, vt, sc.c4, s.fx);
#pragma unroll
    for (int i = 0; i < 3; i++) {
        float bi[3];
        e.step(0, i, e.base, bi);
#pragma unroll
        for (int j = 0; j < 3; j++) {
            float bj[3];
            e.step(1, j, bi, bj);
            const float wij = s.w[0][i] * s.w[1][j];
#pragma unroll
            for (int k = 0; k < 3; k++) {
                float q[3];
                e.step(2, k, bj, q);
                const float W = wij * s.w[2][k];
                acc.add(s.b[0] + i, s.b[1] + j, s.b[2] + k, W * q[0], W * q[1], W * q[2]);
            }
        }
    }
}

// Adjoint of G2P for one liquid particle, position half: the weight-gradient part of dL/dx
// (the weights and offsets of the gather depend on x).  kRolled keeps the loop over x-slices
// rolled: the form for gathers from global memory, whose 27 inlined loads would otherwise
// all be in flight at once.
template <bool kRolled = false, class Gather>
__device__ __forceinline__ void g2p_adj_xgrad(const DevScene& sc, const Stencil& s, const float vt[3],
                                              const float Cb[9], Gather& gv, float xb[3])
{
    // dL/dfx = sum Wb dW - c4 Cb^T sum W g, Wb = g . e with e(o) = vt + c4 Cb (o - fx)
    const Affine e(Cb, vt, sc.c4, s.fx);
    WeightGrad wg;
    float m0[3] = {kZero, kZero, kZero};
    auto slice = [&](int i) {
        // one x-slice of nodes at a time (see p2g_adj_gather)
        asm volatile("" ::: "memory");
        const float w0 = i == 0 ? s.w[0][0] : (i == 1 ? s.w[0][1] : s.w[0][2]);
        const float dw0 = i == 0 ? s.dw[0][0] : (i == 1 ? s.dw[0][1] : s.dw[0][2]);
        float bi[3];
        e.step(0, i, e.base, bi);
#pragma unroll
        for (int j = 0; j < 3; j++) {
            float bj[3];
            e.step(1, j, bi, bj);
            RowSums r;
            float P = kZero, Q = kZero;
#pragma unroll
            for (int k = 0; k < 3; k++) {
                float q[3];
                e.step(2, k, bj, q);
                const float3 g = gv.get(s.b[0] + i, s.b[1] + j, s.b[2] + k);
                const float Wb = g.x * q[0] + g.y * q[1] + g.z * q[2];
                P = fmaf(s.w[2][k], Wb, P);
                Q = fmaf(s.dw[2][k], Wb, Q);
                r.add(k, s.w[2][k], g.x, g.y, g.z);
            }
            wg.add_row(w0, dw0, s.w[1][j], s.dw[1][j], P, Q);
            const float wij = w0 * s.w[1][j];
#pragma unroll
            for (int a = 0; a < 3; a++) m0[a] = fmaf(wij, r.t0[a], m0[a]);
        }
    };
    if (kRolled) {
#pragma unroll 1
        for (int i = 0; i < 3; i++) slice(i);
    } else {
#pragma unroll
        for (int i = 0; i < 3; i++) slice(i);
    }
    // c4 Cb^T sum W g with the columns of c4 Cb
#pragma unroll
    for (int c = 0; c < 3; c++)
        xb[c] = (wg.g[c] - (e.col[c][0] * m0[0] + e.col[c][1] * m0[1] + e.col[c][2] * m0[2])) * sc.inv_dx;
}

// Adjoint of G2P for one particle: scatter of dL/dv_grid and, for liquid (with_x), the
// weight-gradient part of dL/dx.
template <bool with_x, class Gather, class Scatter>
__device__ __forceinline__ void g2p_adj_particle(const DevScene& sc, const Stencil& s, const float vt[3],
                                                 const float Cb[9], Gather& gv, Scatter& acc, float xb[3])
{
    g2p_adj_scatter(sc, s, vt, Cb, acc);
    if (with_x) g2p_adj_xgrad(sc, s, vt, Cb, gv, xb);
    else xb[0] = xb[1] = xb[2] = 0.0f;
}

// Adjoint of P2G for one liquid particle, gather half: with the scattered affine A and
// u = v + dt g, gathers (dL/dmomentum, dL/dmass) and accumulates ub = dL/du, Ab = dL/dA and the
// weight/offset part of dL/dfx.
template <class Gather>
__device__ __forceinline__ void p2g_adj_gather(const DevScene& sc, const Stencil& s, const float A[9], const float u[3],
                                               Gather& acc, float ub[3], float Ab[9], float fxb[3])
{
    // with g the momentum adjoint and mom(o) = m u + A (o - fx) dx: ub = m sum W g,
    // Ab = dx sum W g (x) (o - fx), dL/dfx = sum Wb dW - dx A^T sum W g, Wb = g . mom + g_m m
    const float m = sc.m_liq;
    const float mu[3] = {m * u[0], m * u[1], m * u[2]};
    const Affine mom(A, mu, sc.dx, s.fx);
    Moments M;
    WeightGrad wg;
#pragma unroll
    for (int i = 0; i < 3; i++) {
        // one x-slice of nodes at a time: keeps the compiler from hoisting all 27 gathers
        // (and their registers) ahead of the arithmetic
        asm volatile("" ::: "memory");
        float bi[3];
        mom.step(0, i, mom.base, bi);
#pragma unroll
        for (int j = 0; j < 3; j++) {
            float bj[3];
            mom.step(1, j, bi, bj);
            RowSums r;
            float P = kZero, Q = kZero;
#pragma unroll
            for (int k = 0; k < 3; k++) {
                float q[3];
                mom.step(2, k, bj, q);
                const float4 gq = acc.get(s.b[0] + i, s.b[1] + j, s.b[2] + k);
                const float Wb = gq.x * q[0] + gq.y * q[1] + gq.z * q[2] + gq.w * m;
                P = fmaf(s.w[2][k], Wb, P);
                Q = fmaf(s.dw[2][k], Wb, Q);
                r.add(k, s.w[2][k], gq.x, gq.y, gq.z);
            }
            wg.add_row(s.w[0][i], s.dw[0][i], s.w[1][j], s.dw[1][j], P, Q);
            M.add_row(i, j, s.w[0][i] * s.w[1][j], r);
        }
    }
#pragma unroll
    for (int a = 0; a < 3; a++) {
        ub[a] = m * M.m0[a];
#pragma unroll
        for (int c = 0; c < 3; c++) Ab[a * 3 + c] = sc.dx * M.centred(a, c, s.fx);
    }
    // dx A^T sum W g with the columns of A dx (A itself is then dead after the set-up)
#pragma unroll
    for (int c = 0; c < 3; c++)
        fxb[c] = wg.g[c] - (mom.col[c][0] * M.m0[0] + mom.col[c][1] * M.m0[1] + mom.col[c][2] * M.m0[2]);
}

// Adjoint of P2G for one liquid particle, stress half: from Ab = dL/dA, back through
// A = alpha S + m C, S = eta(|C|) sym(C) + I (k/2)(J - 1/J - 1) J, J = S0 S1 S2 of the FP64 SVD
// of F_tmp (simulation_functions.py:111-153) and F_tmp = (I + dt C) F.  L holds the particle's
// P2G quantities (eta, gamma, J); M is the SVD backward factor of J (svd3_j_factor); Fnb_in is
// the adjoint of F_tmp from the next step's F and xb_partial the G2P part of dL/dx.  Writes the
// adjoints of x, v, C, F and adds dL/dparams into tb.
__device__ __forceinline__ void p2g_adj_stress(const DevScene& sc, const InstParams& ip, const ParticleIn& q,
                                               const LiquidP2G& L, const double M[9], const float ub[3],
                                               const float Ab[9], const float fxb[3], const float Fnb_in[9],
                                               const float xb_partial[3], float xo[3], float vo[3], float Co[9],
                                               float Fo[9], double tb[4])
{
    const float m = sc.m_liq;
    const float* C = q.C;
    float Cb[9], Sb[9];
#pragma unroll
    for (int a = 0; a < 9; a++) {
        Cb[a] = m * Ab[a];
        Sb[a] = sc.alpha * Ab[a];
    }
    float etab = 0.0f;
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
        for (int c = 0; c < 3; c++) etab += Sb[a * 3 + c] * 0.5f * (C[a * 3 + c] + C[c * 3 + a]);
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
        for (int c = 0; c < 3; c++) Cb[a * 3 + c] += L.eta * 0.5f * (Sb[a * 3 + c] + Sb[c * 3 + a]);
    const double trS = (double)Sb[0] + (double)Sb[4] + (double)Sb[8];
    const double J = L.J;
    const double Jb = trS * (ip.kb * 0.5) * (2.0 * J - 1.0);
    if (sc.k_from_param) tb[3] += trS * 0.5 * (J - 1.0 / J - 1.0) * J;
    const Visc vv = viscosity(sc.model, ip.th, L.gamma);
#pragma unroll
    for (int a = 0; a < 4; a++) tb[a] += (double)etab * (double)vv.d_dth[a];
    const float gb = etab * vv.d_dg;
    float n2 = 0.0f;
#pragma unroll
    for (int a = 0; a < 9; a++) n2 = fmaf(C[a], C[a], n2);
    if (n2 > 0.0f) {
        const float inr = 1.0f / sqrtf(n2);
#pragma unroll
        for (int a = 0; a < 9; a++) Cb[a] += gb * C[a] * inr;
    }
    // J = S0 S1 S2 from the FP64 SVD of F_tmp: dL/dF_tmp = U diag(dL/dS) V^T = Jb M
    float Fnb[9];
#pragma unroll
    for (int a = 0; a < 9; a++) Fnb[a] = Fnb_in[a] + (float)(Jb * M[a]);
    // F_tmp = (I + dt C) F
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
        for (int c = 0; c < 3; c++) {
            float sF = 0.0f, sC2 = 0.0f;
#pragma unroll
            for (int k = 0; k < 3; k++) {
                const float Mka = (k == a ? 1.0f : 0.0f) + sc.dt * C[k * 3 + a];
                sF += Mka * Fnb[k * 3 + c];
                sC2 += Fnb[a * 3 + k] * q.F[c * 3 + k];
            }
            Fo[a * 3 + c] = sF;
            Cb[a * 3 + c] += sc.dt * sC2;
        }
#pragma unroll
    for (int a = 0; a < 3; a++) xo[a] = xb_partial[a] + fxb[a] * sc.inv_dx;
#pragma unroll
    for (int a = 0; a < 3; a++) vo[a] = ub[a];
#pragma unroll
    for (int a = 0; a < 9; a++) Co[a] = Cb[a];
}

// Adjoint of P2G for one liquid particle (gather and stress halves, P2G quantities computed here).
template <class Gather>
__device__ __forceinline__ void p2g_adj_liquid(const DevScene& sc, const InstParams& ip, const ParticleIn& q,
                                               const Stencil& s, const float Fnb_in[9], const float xb_partial[3],
                                               Gather& acc, float xo[3], float vo[3], float Co[9], float Fo[9],
                                               double tb[4])
{
    LiquidP2G L;
    liquid_p2g<true>(sc, ip.th, ip.kb, q.C, q.F, L);
    double M[9];
    svd3_j_factor(L.U, L.S, L.V, M);
    const float u[3] = {q.v[0], q.v[1] + sc.gdt, q.v[2]};
    float ub[3], Ab[9], fxb[3];
    p2g_adj_gather(sc, s, L.A, u, acc, ub, Ab, fxb);
    p2g_adj_stress(sc, ip, q, L, M, ub, Ab, fxb, Fnb_in, xb_partial, xo, vo, Co, Fo, tb);
}

// Adjoint of P2G for a stick or container particle: only C carries an adjoint
// (positions and velocities of these species do not depend on the parameters).
template <class Gather>
__device__ __forceinline__ void p2g_adj_rigid(const DevScene& sc, float m, const Stencil& s, Gather& acc, float Co[9])
{
    // Cb = dx sum W g (x) (o - fx), g the momentum adjoint
    Moments M;
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) {
            RowSums r;
#pragma unroll
            for (int k = 0; k < 3; k++) {
                const float4 gq = acc.get(s.b[0] + i, s.b[1] + j, s.b[2] + k);
                r.add(k, s.w[2][k], gq.x, gq.y, gq.z);
            }
            M.add_row(i, j, s.w[0][i] * s.w[1][j], r);
        }
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
        for (int c = 0; c < 3; c++) Co[a * 3 + c] = m * (sc.dx * M.centred(a, c, s.fx));
}

// ---------------------------------------------------------------------------- global accessors

__device__ __forceinline__ long node_index(int n, int i, int j, int k)
{
    return ((long)i * n + j) * n + k;
}

struct ScatterGlobal4 {
    float4* g;
    int n;
    __device__ __forceinline__ void add(int i, int j, int k, float a, float b, float c, float d)
    {
        float4* q = g + node_index(n, i, j, k);
        atomicAdd(&q->x, a);
        atomicAdd(&q->y, b);
        atomicAdd(&q->z, c);
        atomicAdd(&q->w, d);
    }
};

struct ScatterGlobal3 {
    float4* g;
    int n;
    __device__ __forceinline__ void add(int i, int j, int k, float a, float b, float c)
    {
        float4* q = g + node_index(n, i, j, k);
        atomicAdd(&q->x, a);
        atomicAdd(&q->y, b);
        atomicAdd(&q->z, c);
    }
};

// The fall-back scatter of the tiled kernels, with sums that do not depend on the order of the
// atomic adds.  A particle whose contributions are below `bound` < 2^ex in magnitude adds each one
// rounded to a multiple of its grain 2^(ex - 26) (an error of at most half the grain, under 2^-26
// of the bound and so under FP32's own rounding) into an FP64 grid (4 doubles per node).  The rounding is (x + M) - M in FP64 with
// M = 1.5 * 2^(ex + 26): x + M stays in M's binade, whose spacing is the grain.  Every partial sum
// at a node is then a multiple of the smallest grain among the particles that reached it and
// stays exact in FP64 (53-bit significands) while those particles' bounds lie within
// 2^(26 - log2(count)) of each other, so the atomic adds give the same bits in any order.
// Outside that range the sum is still exact to FP64 rounding.
__device__ __forceinline__ double fallback_magic(float bound)
{
    int ex = 0;
    (void)frexpf(bound, &ex);   // bound < 2^ex
    return ldexp(1.5, ex + 26);
}

__device__ __forceinline__ double grain_round(float x, double M)
{
    return ((double)x + M) - M;
}

struct ScatterExact4 {
    double* g;
    int n;
    double M;
    __device__ __forceinline__ void add(int i, int j, int k, float a, float b, float c, float d)
    {
        double* q = g + 4 * node_index(n, i, j, k);
        atomicAdd(q + 0, grain_round(a, M));
        atomicAdd(q + 1, grain_round(b, M));
        atomicAdd(q + 2, grain_round(c, M));
        atomicAdd(q + 3, grain_round(d, M));
    }
};

struct ScatterExact3 {
    double* g;
    int n;
    double M;
    __device__ __forceinline__ void add(int i, int j, int k, float a, float b, float c)
    {
        double* q = g + 4 * node_index(n, i, j, k);
        atomicAdd(q + 0, grain_round(a, M));
        atomicAdd(q + 1, grain_round(b, M));
        atomicAdd(q + 2, grain_round(c, M));
    }
};

// Bound on the P2G contributions of a particle with payload pl: W (mu + A d) per component with
// W <= 1 and |d| <= 1.5 dx per axis, and W m.
__device__ __forceinline__ float p2g_bound(const DevScene& sc, const P2GPayload& pl)
{
    float b = fabsf(pl.m);
#pragma unroll
    for (int c = 0; c < 3; c++)
        b = fmaxf(b, fabsf(pl.mu[c]) + 1.5f * sc.dx * (fabsf(pl.A[c * 3]) + fabsf(pl.A[c * 3 + 1]) + fabsf(pl.A[c * 3 + 2])));
    return b * 1.001f;
}

// Bound on the adjoint G2P contributions W (vt + c4 Cb (o - fx)) with W <= 1 and |o - fx| <= 1.5.
__device__ __forceinline__ float g2p_adj_bound(const DevScene& sc, const float vt[3], const float Cb[9])
{
    float b = 0.0f;
#pragma unroll
    for (int a = 0; a < 3; a++)
        b = fmaxf(b, fabsf(vt[a]) + 1.5f * fabsf(sc.c4) * (fabsf(Cb[a * 3]) + fabsf(Cb[a * 3 + 1]) + fabsf(Cb[a * 3 + 2])));
    return b * 1.001f;
}

// velocity from a precomputed v_out grid
struct GatherVout {
    const float4* g;
    int n;
    __device__ __forceinline__ float3 get(int i, int j, int k) const
    {
        const float4 v = g[node_index(n, i, j, k)];
        return make_float3(v.x, v.y, v.z);
    }
};

// velocity computed on the fly from the momentum/mass grid (grid update per node)
struct GatherVfromP {
    const float4* g;
    DevScene sc;
    __device__ __forceinline__ float3 get(int i, int j, int k) const
    {
        float v[3];
        NodeRec r;
        gridop_node(sc, g[node_index(sc.n_grid, i, j, k)], i, j, k, v, r);
        return make_float3(v[0], v[1], v[2]);
    }
};

// (dL/dmomentum, dL/dmass) from a precomputed grid
struct GatherPbar {
    const float4* g;
    int n;
    __device__ __forceinline__ float4 get(int i, int j, int k) const { return g[node_index(n, i, j, k)]; }
};

// (dL/dmomentum, dL/dmass) computed on the fly from the momentum/mass grid and dL/dv_out
struct GatherPbarFromP {
    const float4* g;
    const float4* vb;
    DevScene sc;
    __device__ __forceinline__ float4 get(int i, int j, int k) const
    {
        const long id = node_index(sc.n_grid, i, j, k);
        float v[3];
        NodeRec r;
        gridop_node(sc, g[id], i, j, k, v, r);
        const float4 b = vb[id];
        return gridop_node_adj(r, b.x, b.y, b.z);
    }
};

}  // namespace ds
