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
#pragma unroll
    for (int a = 0; a < 3; a++) nv[a] = 0.0f;
#pragma unroll
    for (int a = 0; a < 9; a++) nC[a] = 0.0f;
#pragma unroll
    for (int i = 0; i < 3; i++) {
        const float d0 = (float)i - s.fx[0];
#pragma unroll
        for (int j = 0; j < 3; j++) {
            const float d1 = (float)j - s.fx[1];
            const float wij = s.w[0][i] * s.w[1][j];
#pragma unroll
            for (int k = 0; k < 3; k++) {
                const float d2 = (float)k - s.fx[2];
                const float W = wij * s.w[2][k];
                const float3 g = acc.get(s.b[0] + i, s.b[1] + j, s.b[2] + k);
                const float wv[3] = {W * g.x, W * g.y, W * g.z};
#pragma unroll
                for (int a = 0; a < 3; a++) {
                    nv[a] += wv[a];
                    nC[a * 3 + 0] += sc.c4 * wv[a] * d0;
                    nC[a * 3 + 1] += sc.c4 * wv[a] * d1;
                    nC[a * 3 + 2] += sc.c4 * wv[a] * d2;
                }
            }
        }
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
#pragma unroll
    for (int i = 0; i < 3; i++) {
#pragma unroll
        for (int j = 0; j < 3; j++) {
#pragma unroll
            for (int k = 0; k < 3; k++) {
                const float d[3] = {(float)i - s.fx[0], (float)j - s.fx[1], (float)k - s.fx[2]};
                const float W = s.w[0][i] * s.w[1][j] * s.w[2][k];
                float e[3];
#pragma unroll
                for (int a = 0; a < 3; a++)
                    e[a] = vt[a] + sc.c4 * (Cb[a * 3 + 0] * d[0] + Cb[a * 3 + 1] * d[1] + Cb[a * 3 + 2] * d[2]);
                acc.add(s.b[0] + i, s.b[1] + j, s.b[2] + k, W * e[0], W * e[1], W * e[2]);
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
    float fxb[3] = {0.0f, 0.0f, 0.0f};
    auto slice = [&](int i) {
        // one x-slice of nodes at a time (see p2g_adj_gather)
        asm volatile("" ::: "memory");
        const float w0 = i == 0 ? s.w[0][0] : (i == 1 ? s.w[0][1] : s.w[0][2]);
        const float dw0 = i == 0 ? s.dw[0][0] : (i == 1 ? s.dw[0][1] : s.dw[0][2]);
#pragma unroll
        for (int j = 0; j < 3; j++) {
#pragma unroll
            for (int k = 0; k < 3; k++) {
                const float d[3] = {(float)i - s.fx[0], (float)j - s.fx[1], (float)k - s.fx[2]};
                const float W = w0 * s.w[1][j] * s.w[2][k];
                float e[3];
#pragma unroll
                for (int a = 0; a < 3; a++)
                    e[a] = vt[a] + sc.c4 * (Cb[a * 3 + 0] * d[0] + Cb[a * 3 + 1] * d[1] + Cb[a * 3 + 2] * d[2]);
                const float3 g = gv.get(s.b[0] + i, s.b[1] + j, s.b[2] + k);
                const float Wb = g.x * e[0] + g.y * e[1] + g.z * e[2];
                const float dW[3] = {dw0 * s.w[1][j] * s.w[2][k], w0 * s.dw[1][j] * s.w[2][k],
                                     w0 * s.w[1][j] * s.dw[2][k]};
#pragma unroll
                for (int c = 0; c < 3; c++) {
                    const float ctg = Cb[0 * 3 + c] * g.x + Cb[1 * 3 + c] * g.y + Cb[2 * 3 + c] * g.z;
                    fxb[c] += Wb * dW[c] - sc.c4 * W * ctg;
                }
            }
        }
    };
    if (kRolled) {
#pragma unroll 1
        for (int i = 0; i < 3; i++) slice(i);
    } else {
#pragma unroll
        for (int i = 0; i < 3; i++) slice(i);
    }
#pragma unroll
    for (int a = 0; a < 3; a++) xb[a] = fxb[a] * sc.inv_dx;
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
    const float m = sc.m_liq;
#pragma unroll
    for (int a = 0; a < 3; a++) ub[a] = fxb[a] = 0.0f;
#pragma unroll
    for (int a = 0; a < 9; a++) Ab[a] = 0.0f;
#pragma unroll
    for (int i = 0; i < 3; i++) {
        // one x-slice of nodes at a time: keeps the compiler from hoisting all 27 gathers
        // (and their registers) ahead of the arithmetic
        asm volatile("" ::: "memory");
#pragma unroll
        for (int j = 0; j < 3; j++) {
#pragma unroll
            for (int k = 0; k < 3; k++) {
                const float dp[3] = {((float)i - s.fx[0]) * sc.dx, ((float)j - s.fx[1]) * sc.dx,
                                     ((float)k - s.fx[2]) * sc.dx};
                const float W = s.w[0][i] * s.w[1][j] * s.w[2][k];
                const float dW[3] = {s.dw[0][i] * s.w[1][j] * s.w[2][k], s.w[0][i] * s.dw[1][j] * s.w[2][k],
                                     s.w[0][i] * s.w[1][j] * s.dw[2][k]};
                const float4 gq = acc.get(s.b[0] + i, s.b[1] + j, s.b[2] + k);
                const float gg[3] = {gq.x, gq.y, gq.z};
                float mom[3];
#pragma unroll
                for (int a = 0; a < 3; a++) mom[a] = m * u[a] + A[a * 3] * dp[0] + A[a * 3 + 1] * dp[1] + A[a * 3 + 2] * dp[2];
#pragma unroll
                for (int a = 0; a < 3; a++) {
                    ub[a] += W * m * gg[a];
#pragma unroll
                    for (int c = 0; c < 3; c++) Ab[a * 3 + c] += W * gg[a] * dp[c];
                }
                const float Wb = gg[0] * mom[0] + gg[1] * mom[1] + gg[2] * mom[2] + gq.w * m;
#pragma unroll
                for (int c = 0; c < 3; c++) {
                    const float atg = A[0 * 3 + c] * gg[0] + A[1 * 3 + c] * gg[1] + A[2 * 3 + c] * gg[2];
                    fxb[c] += Wb * dW[c] - sc.dx * W * atg;
                }
            }
        }
    }
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
    float Cb[9] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++)
#pragma unroll
            for (int k = 0; k < 3; k++) {
                const float dp[3] = {((float)i - s.fx[0]) * sc.dx, ((float)j - s.fx[1]) * sc.dx,
                                     ((float)k - s.fx[2]) * sc.dx};
                const float W = s.w[0][i] * s.w[1][j] * s.w[2][k];
                const float4 gq = acc.get(s.b[0] + i, s.b[1] + j, s.b[2] + k);
                const float gg[3] = {gq.x, gq.y, gq.z};
#pragma unroll
                for (int a = 0; a < 3; a++)
#pragma unroll
                    for (int c = 0; c < 3; c++) Cb[a * 3 + c] += W * gg[a] * dp[c];
            }
#pragma unroll
    for (int a = 0; a < 9; a++) Co[a] = m * Cb[a];
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
