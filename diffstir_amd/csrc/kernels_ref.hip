// kernels_ref.hip -- per-particle kernels of the substep and its adjoint (one thread per
// particle or node, grid scatters through global float atomics).
//
// These are the straightforward forms of the substep: one launch per phase, dense grid,
// every P2G contribution an atomic add in L2.  The tiled fused kernels (kernels_tiled.hip)
// are checked against them; they also serve the recompute of the first and last step of
// each adjoint segment.
#include "ds_device.h"

namespace ds {

__global__ void __launch_bounds__(256) k_fill_f4(float4* __restrict__ p, long n)
{
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    long stride = (long)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}

__global__ void __launch_bounds__(256) k_fill_f32(float* __restrict__ p, long n, float v)
{
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    long stride = (long)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = v;
}

__global__ void __launch_bounds__(256) k_fill_f64(double* __restrict__ p, long n)
{
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    long stride = (long)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = 0.0;
}

// State at step 0 (simulation_functions.py:306-319): x from the initial positions, v = 0,
// C = 0, F = I.  Padding lanes get a harmless in-grid position.
__global__ void __launch_bounds__(256) k_init_state(float* __restrict__ st, const float* __restrict__ x0,
                                                    const int* __restrict__ meta, int P, float pad_pos)
{
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P) return;
    bool real = meta[p] >= 0;
#pragma unroll
    for (int a = 0; a < 3; a++) st[(SX + a) * (long)P + p] = real ? x0[3 * (long)p + a] : pad_pos;
#pragma unroll
    for (int a = 0; a < 3; a++) st[(SV + a) * (long)P + p] = 0.0f;
#pragma unroll
    for (int a = 0; a < 9; a++) st[(SC + a) * (long)P + p] = 0.0f;
#pragma unroll
    for (int a = 0; a < 9; a++) st[(SF + a) * (long)P + p] = (a == 0 || a == 4 || a == 8) ? 1.0f : 0.0f;
}

// ---------------------------------------------------------------------------- forward

// P2G of simulation_functions.py:104-196 for all species of all instances.  Reads x, v, C, F
// from `in`, writes F_tmp into `fout` (NULL: not written), scatters into grid (momentum xyz,
// mass w) with float atomics.
__global__ void __launch_bounds__(256) k_p2g_ref(const float* __restrict__ in, float* __restrict__ fout,
                                                 const int* __restrict__ meta, const Group* __restrict__ groups,
                                                 const double* __restrict__ params, float4* __restrict__ grid,
                                                 DevScene sc, int P, unsigned* __restrict__ flag)
{
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P) return;
    int g = meta[p];
    if (g < 0) return;
    Group G = groups[g];
    const long LP = P;
    float x[3], v[3], C[9];
#pragma unroll
    for (int a = 0; a < 3; a++) x[a] = in[(SX + a) * LP + p];
#pragma unroll
    for (int a = 0; a < 3; a++) v[a] = in[(SV + a) * LP + p];
#pragma unroll
    for (int a = 0; a < 9; a++) C[a] = in[(SC + a) * LP + p];
    Stencil s;
    if (!stencil(sc, x[0], x[1], x[2], s)) atomicOr(flag, 1u);
    float A[9], u[3], m;
    if (G.species == LIQUID) {
        float F[9];
#pragma unroll
        for (int a = 0; a < 9; a++) F[a] = in[(SF + a) * LP + p];
        float th[4];
        const double* pp = params + 4 * G.inst;
#pragma unroll
        for (int a = 0; a < 4; a++) th[a] = (float)pp[a];
        double kb = sc.k_from_param ? pp[3] : sc.k_bulk;
        LiquidP2G q;
        liquid_p2g(sc, th, kb, C, F, q);
        if (fout) {
#pragma unroll
            for (int a = 0; a < 9; a++) fout[(SF + a) * LP + p] = q.Fn[a];
        }
#pragma unroll
        for (int a = 0; a < 9; a++) A[a] = q.A[a];
        m = sc.m_liq;
        u[0] = v[0];
        u[1] = v[1] + sc.gdt;
        u[2] = v[2];
    } else {
        m = G.species == STICK ? sc.m_stick : sc.m_cont;
#pragma unroll
        for (int a = 0; a < 9; a++) A[a] = m * C[a];
        u[0] = v[0]; u[1] = v[1]; u[2] = v[2];
    }
    const int n = sc.n_grid;
    float4* gi = grid + (long)G.inst * n * n * n;
    float mu[3] = {m * u[0], m * u[1], m * u[2]};
#pragma unroll
    for (int i = 0; i < 3; i++) {
        float d0 = ((float)i - s.fx[0]) * sc.dx;
#pragma unroll
        for (int j = 0; j < 3; j++) {
            float d1 = ((float)j - s.fx[1]) * sc.dx;
            float wij = s.w[0][i] * s.w[1][j];
#pragma unroll
            for (int k = 0; k < 3; k++) {
                float d2 = ((float)k - s.fx[2]) * sc.dx;
                float W = wij * s.w[2][k];
                float4* node = gi + ((long)(s.b[0] + i) * n + (s.b[1] + j)) * n + (s.b[2] + k);
                atomicAdd(&node->x, W * (mu[0] + A[0] * d0 + A[1] * d1 + A[2] * d2));
                atomicAdd(&node->y, W * (mu[1] + A[3] * d0 + A[4] * d1 + A[5] * d2));
                atomicAdd(&node->z, W * (mu[2] + A[6] * d0 + A[7] * d1 + A[8] * d2));
                atomicAdd(&node->w, W * m);
            }
        }
    }
}

// Grid update of simulation_functions.py:199-225 over every node of every instance.
__global__ void __launch_bounds__(256) k_gridop_ref(const float4* __restrict__ grid, float4* __restrict__ vout,
                                                    DevScene sc, int n_inst)
{
    const int n = sc.n_grid;
    const long nn = (long)n * n * n;
    long id = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= nn * n_inst) return;
    long node = id % nn;
    int i = (int)(node / ((long)n * n)), j = (int)((node / n) % n), k = (int)(node % n);
    float v[3];
    NodeRec r;
    gridop_node(sc, grid[id], i, j, k, v, r);
    vout[id] = make_float4(v[0], v[1], v[2], 0.0f);
}

// G2P of simulation_functions.py:228-290.  Reads x from `in`, writes x, v, C to `out`.  The
// stick's dv = v_kin - v_grid is summed per instance into dvsum[inst][f+1] in FP64
// (simulation_functions.py:266-267, 330-334); dvsum NULL skips it.
__global__ void __launch_bounds__(256) k_g2p_ref(const float* __restrict__ in, float* __restrict__ out,
                                                 const int* __restrict__ meta, const Group* __restrict__ groups,
                                                 const float4* __restrict__ vout, const float* __restrict__ traj,
                                                 int T, int f, double* __restrict__ dvsum, DevScene sc, int P,
                                                 unsigned* __restrict__ flag)
{
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    int g = p < P ? meta[p] : -1;
    // lanes of one wavefront share a group (groups are 64-aligned); padding lanes take part
    // in the stick reduction with zeros.
    int gw = __shfl(g, 0, 64);
    double dv[3] = {0.0, 0.0, 0.0};
    const long LP = P;
    if (g >= 0) {
        Group G = groups[g];
        float x[3];
#pragma unroll
        for (int a = 0; a < 3; a++) x[a] = in[(SX + a) * LP + p];
        Stencil s;
        if (!stencil(sc, x[0], x[1], x[2], s)) atomicOr(flag, 1u);
        const int n = sc.n_grid;
        const float4* vi = vout + (long)G.inst * n * n * n;
        float nv[3] = {0.0f, 0.0f, 0.0f};
        float nC[9] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int i = 0; i < 3; i++) {
            float d0 = (float)i - s.fx[0];
#pragma unroll
            for (int j = 0; j < 3; j++) {
                float d1 = (float)j - s.fx[1];
                float wij = s.w[0][i] * s.w[1][j];
#pragma unroll
                for (int k = 0; k < 3; k++) {
                    float d2 = (float)k - s.fx[2];
                    float W = wij * s.w[2][k];
                    float4 gv = vi[((long)(s.b[0] + i) * n + (s.b[1] + j)) * n + (s.b[2] + k)];
                    float wv[3] = {W * gv.x, W * gv.y, W * gv.z};
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
        if (G.species == LIQUID) {
#pragma unroll
            for (int a = 0; a < 3; a++) {
                out[(SV + a) * LP + p] = nv[a];
                out[(SX + a) * LP + p] = x[a] + sc.dt * nv[a];
            }
        } else if (G.species == STICK) {
            const float* tr = traj + ((long)G.inst * T + f) * 3;
#pragma unroll
            for (int a = 0; a < 3; a++) {
                float vk = (tr[3 + a] - tr[a]) / sc.dt;
                dv[a] = (double)(vk - nv[a]);
                out[(SV + a) * LP + p] = vk;
                out[(SX + a) * LP + p] = x[a] + vk * sc.dt;
            }
        } else {
#pragma unroll
            for (int a = 0; a < 3; a++) {
                out[(SV + a) * LP + p] = 0.0f;
                out[(SX + a) * LP + p] = x[a];
            }
        }
#pragma unroll
        for (int a = 0; a < 9; a++) out[(SC + a) * LP + p] = nC[a];
    }
    if (dvsum && gw >= 0 && groups[gw].species == STICK) {
        double s0 = wave_sum(dv[0]), s1 = wave_sum(dv[1]), s2 = wave_sum(dv[2]);
        if ((threadIdx.x & 63) == 0) {
            double* d = dvsum + ((long)groups[gw].inst * T + f + 1) * 3;
            atomicAdd(d + 0, s0);
            atomicAdd(d + 1, s1);
            atomicAdd(d + 2, s2);
        }
    }
}

// ---------------------------------------------------------------------------- adjoint

// Adjoint of G2P.  `st` is the state at step f, `adj` holds the adjoints of the state at f+1
// (x, v, C for liquid; C for stick and container), dvbar[inst][f+1] the adjoint of the stick's
// dv.  Scatters dL/dv_out into vbar and writes the liquid's partial dL/dx (the direct x' = x
// path plus the weight-gradient terms of the gather) into xbar.
__global__ void __launch_bounds__(256) k_g2p_adj_ref(const float* __restrict__ st, const float* __restrict__ adj,
                                                     const int* __restrict__ meta, const Group* __restrict__ groups,
                                                     const float4* __restrict__ vout, const double* __restrict__ dvbar,
                                                     int T, int f, float4* __restrict__ vbar, float* __restrict__ xbar,
                                                     DevScene sc, int P)
{
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P) return;
    int g = meta[p];
    if (g < 0) return;
    Group G = groups[g];
    const long LP = P;
    float x[3];
#pragma unroll
    for (int a = 0; a < 3; a++) x[a] = st[(SX + a) * LP + p];
    Stencil s;
    stencil(sc, x[0], x[1], x[2], s);
    float Cb[9];
#pragma unroll
    for (int a = 0; a < 9; a++) Cb[a] = adj[(SC + a) * LP + p];
    float vt[3] = {0.0f, 0.0f, 0.0f};
    bool liq = G.species == LIQUID;
    if (liq) {
#pragma unroll
        for (int a = 0; a < 3; a++) vt[a] = adj[(SV + a) * LP + p] + sc.dt * adj[(SX + a) * LP + p];
    } else if (G.species == STICK) {
        const double* d = dvbar + ((long)G.inst * T + f + 1) * 3;
#pragma unroll
        for (int a = 0; a < 3; a++) vt[a] = -(float)d[a];
    }
    const int n = sc.n_grid;
    const float4* vi = vout + (long)G.inst * n * n * n;
    float4* bi = vbar + (long)G.inst * n * n * n;
    float fxb[3] = {0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int i = 0; i < 3; i++) {
#pragma unroll
        for (int j = 0; j < 3; j++) {
#pragma unroll
            for (int k = 0; k < 3; k++) {
                float d[3] = {(float)i - s.fx[0], (float)j - s.fx[1], (float)k - s.fx[2]};
                float W = s.w[0][i] * s.w[1][j] * s.w[2][k];
                long id = ((long)(s.b[0] + i) * n + (s.b[1] + j)) * n + (s.b[2] + k);
                float cd[3];
#pragma unroll
                for (int a = 0; a < 3; a++) cd[a] = Cb[a * 3 + 0] * d[0] + Cb[a * 3 + 1] * d[1] + Cb[a * 3 + 2] * d[2];
                atomicAdd(&bi[id].x, W * (vt[0] + sc.c4 * cd[0]));
                atomicAdd(&bi[id].y, W * (vt[1] + sc.c4 * cd[1]));
                atomicAdd(&bi[id].z, W * (vt[2] + sc.c4 * cd[2]));
                if (liq) {
                    float4 gv = vi[id];
                    float gg[3] = {gv.x, gv.y, gv.z};
                    float Wb = gg[0] * (vt[0] + sc.c4 * cd[0]) + gg[1] * (vt[1] + sc.c4 * cd[1])
                             + gg[2] * (vt[2] + sc.c4 * cd[2]);
                    float dW[3] = {s.dw[0][i] * s.w[1][j] * s.w[2][k], s.w[0][i] * s.dw[1][j] * s.w[2][k],
                                   s.w[0][i] * s.w[1][j] * s.dw[2][k]};
#pragma unroll
                    for (int c = 0; c < 3; c++) {
                        float ctg = Cb[0 * 3 + c] * gg[0] + Cb[1 * 3 + c] * gg[1] + Cb[2 * 3 + c] * gg[2];
                        fxb[c] += Wb * dW[c] - sc.c4 * W * ctg;
                    }
                }
            }
        }
    }
    if (liq) {
#pragma unroll
        for (int a = 0; a < 3; a++) xbar[a * LP + p] = adj[(SX + a) * LP + p] + fxb[a] * sc.inv_dx;
    }
}

// Adjoint of the grid update, per node: vbar -> (dL/d momentum, dL/d mass) into pbar.
__global__ void __launch_bounds__(256) k_gridop_adj_ref(const float4* __restrict__ grid, const float4* __restrict__ vbar,
                                                        float4* __restrict__ pbar, DevScene sc, int n_inst)
{
    const int n = sc.n_grid;
    const long nn = (long)n * n * n;
    long id = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= nn * n_inst) return;
    long node = id % nn;
    int i = (int)(node / ((long)n * n)), j = (int)((node / n) % n), k = (int)(node % n);
    float v[3];
    NodeRec r;
    gridop_node(sc, grid[id], i, j, k, v, r);
    float4 gb = vbar[id];
    pbar[id] = gridop_node_adj(r, gb.x, gb.y, gb.z);
}

// Adjoint of P2G.  `st` is the state at step f; adj holds dL/dF of step f+1 in its F slots on
// entry and receives dL/d(x, v, C, F) of step f (liquid; C only for stick and container).
// xbar is the partial dL/dx from k_g2p_adj_ref.  dL/dparams is accumulated in FP64 per
// instance into thbar[inst][4].
__global__ void __launch_bounds__(256) k_p2g_adj_ref(const float* __restrict__ st, float* __restrict__ adj,
                                                     const float* __restrict__ xbar, const int* __restrict__ meta,
                                                     const Group* __restrict__ groups, const double* __restrict__ params,
                                                     const float4* __restrict__ pbar, double* __restrict__ thbar,
                                                     DevScene sc, int P)
{
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    int g = p < P ? meta[p] : -1;
    int gw = __shfl(g, 0, 64);
    double tb[4] = {0.0, 0.0, 0.0, 0.0};
    const long LP = P;
    if (g >= 0) {
        Group G = groups[g];
        float x[3], C[9];
#pragma unroll
        for (int a = 0; a < 3; a++) x[a] = st[(SX + a) * LP + p];
#pragma unroll
        for (int a = 0; a < 9; a++) C[a] = st[(SC + a) * LP + p];
        Stencil s;
        stencil(sc, x[0], x[1], x[2], s);
        const int n = sc.n_grid;
        const float4* pb = pbar + (long)G.inst * n * n * n;
        if (G.species == LIQUID) {
            float v[3], F[9];
#pragma unroll
            for (int a = 0; a < 3; a++) v[a] = st[(SV + a) * LP + p];
#pragma unroll
            for (int a = 0; a < 9; a++) F[a] = st[(SF + a) * LP + p];
            float th[4];
            const double* pp = params + 4 * G.inst;
#pragma unroll
            for (int a = 0; a < 4; a++) th[a] = (float)pp[a];
            double kb = sc.k_from_param ? pp[3] : sc.k_bulk;
            LiquidP2G q;
            liquid_p2g(sc, th, kb, C, F, q);
            const float m = sc.m_liq;
            float u[3] = {v[0], v[1] + sc.gdt, v[2]};
            float ub[3] = {0.0f, 0.0f, 0.0f}, Ab[9] = {0.0f}, fxb[3] = {0.0f, 0.0f, 0.0f};
#pragma unroll
            for (int i = 0; i < 3; i++) {
#pragma unroll
                for (int j = 0; j < 3; j++) {
#pragma unroll
                    for (int k = 0; k < 3; k++) {
                        float dp[3] = {((float)i - s.fx[0]) * sc.dx, ((float)j - s.fx[1]) * sc.dx,
                                       ((float)k - s.fx[2]) * sc.dx};
                        float W = s.w[0][i] * s.w[1][j] * s.w[2][k];
                        float dW[3] = {s.dw[0][i] * s.w[1][j] * s.w[2][k], s.w[0][i] * s.dw[1][j] * s.w[2][k],
                                       s.w[0][i] * s.w[1][j] * s.dw[2][k]};
                        float4 gq = pb[((long)(s.b[0] + i) * n + (s.b[1] + j)) * n + (s.b[2] + k)];
                        float gg[3] = {gq.x, gq.y, gq.z};
                        float mom[3];
#pragma unroll
                        for (int a = 0; a < 3; a++)
                            mom[a] = m * u[a] + q.A[a * 3] * dp[0] + q.A[a * 3 + 1] * dp[1] + q.A[a * 3 + 2] * dp[2];
#pragma unroll
                        for (int a = 0; a < 3; a++) {
                            ub[a] += W * m * gg[a];
#pragma unroll
                            for (int c = 0; c < 3; c++) Ab[a * 3 + c] += W * gg[a] * dp[c];
                        }
                        float Wb = gg[0] * mom[0] + gg[1] * mom[1] + gg[2] * mom[2] + gq.w * m;
#pragma unroll
                        for (int c = 0; c < 3; c++) {
                            float atg = q.A[0 * 3 + c] * gg[0] + q.A[1 * 3 + c] * gg[1] + q.A[2 * 3 + c] * gg[2];
                            fxb[c] += Wb * dW[c] - sc.dx * W * atg;
                        }
                    }
                }
            }
            float Cb[9], Sb[9];
#pragma unroll
            for (int a = 0; a < 9; a++) { Cb[a] = m * Ab[a]; Sb[a] = sc.alpha * Ab[a]; }
            float etab = 0.0f;
#pragma unroll
            for (int a = 0; a < 3; a++)
#pragma unroll
                for (int c = 0; c < 3; c++) etab += Sb[a * 3 + c] * 0.5f * (C[a * 3 + c] + C[c * 3 + a]);
#pragma unroll
            for (int a = 0; a < 3; a++)
#pragma unroll
                for (int c = 0; c < 3; c++) Cb[a * 3 + c] += q.eta * 0.5f * (Sb[a * 3 + c] + Sb[c * 3 + a]);
            double trS = (double)Sb[0] + (double)Sb[4] + (double)Sb[8];
            double J = q.J;
            double Jb = trS * (kb * 0.5) * (2.0 * J - 1.0);
            if (sc.k_from_param) tb[3] += trS * 0.5 * (J - 1.0 / J - 1.0) * J;
            Visc vv = viscosity(sc.model, th, q.gamma);
#pragma unroll
            for (int a = 0; a < 4; a++) tb[a] += (double)etab * (double)vv.d_dth[a];
            float gb = etab * vv.d_dg;
            float nrm = q.gamma - 1e-6f;
            float n2 = 0.0f;
#pragma unroll
            for (int a = 0; a < 9; a++) n2 = fmaf(C[a], C[a], n2);
            if (n2 > 0.0f) {
                float inr = 1.0f / sqrtf(n2);
                (void)nrm;
#pragma unroll
                for (int a = 0; a < 9; a++) Cb[a] += gb * C[a] * inr;
            }
            // J = S0 S1 S2: dJ/dS_i = prod of the other two; back through the SVD.
            double gS[3] = {Jb * q.S[1] * q.S[2], Jb * q.S[0] * q.S[2], Jb * q.S[0] * q.S[1]};
            double gFn[9];
            svd3_backward_sigma(q.U, q.V, gS, gFn);
            float Fnb[9];
#pragma unroll
            for (int a = 0; a < 9; a++) Fnb[a] = adj[(SF + a) * LP + p] + (float)gFn[a];
            // F_tmp = (I + dt C) F
            float Fb[9];
#pragma unroll
            for (int a = 0; a < 3; a++)
#pragma unroll
                for (int c = 0; c < 3; c++) {
                    float sF = 0.0f, sC2 = 0.0f;
#pragma unroll
                    for (int k = 0; k < 3; k++) {
                        float Mka = (k == a ? 1.0f : 0.0f) + sc.dt * C[k * 3 + a];
                        sF += Mka * Fnb[k * 3 + c];
                        sC2 += Fnb[a * 3 + k] * F[c * 3 + k];
                    }
                    Fb[a * 3 + c] = sF;
                    Cb[a * 3 + c] += sc.dt * sC2;
                }
#pragma unroll
            for (int a = 0; a < 3; a++) adj[(SX + a) * LP + p] = xbar[a * LP + p] + fxb[a] * sc.inv_dx;
#pragma unroll
            for (int a = 0; a < 3; a++) adj[(SV + a) * LP + p] = ub[a];
#pragma unroll
            for (int a = 0; a < 9; a++) adj[(SC + a) * LP + p] = Cb[a];
#pragma unroll
            for (int a = 0; a < 9; a++) adj[(SF + a) * LP + p] = Fb[a];
        } else {
            const float m = G.species == STICK ? sc.m_stick : sc.m_cont;
            float Cb[9] = {0.0f};
#pragma unroll
            for (int i = 0; i < 3; i++)
#pragma unroll
                for (int j = 0; j < 3; j++)
#pragma unroll
                    for (int k = 0; k < 3; k++) {
                        float dp[3] = {((float)i - s.fx[0]) * sc.dx, ((float)j - s.fx[1]) * sc.dx,
                                       ((float)k - s.fx[2]) * sc.dx};
                        float W = s.w[0][i] * s.w[1][j] * s.w[2][k];
                        float4 gq = pb[((long)(s.b[0] + i) * n + (s.b[1] + j)) * n + (s.b[2] + k)];
                        float gg[3] = {gq.x, gq.y, gq.z};
#pragma unroll
                        for (int a = 0; a < 3; a++)
#pragma unroll
                            for (int c = 0; c < 3; c++) Cb[a * 3 + c] += W * gg[a] * dp[c];
                    }
#pragma unroll
            for (int a = 0; a < 9; a++) adj[(SC + a) * LP + p] = m * Cb[a];
        }
    }
    if (gw >= 0 && groups[gw].species == LIQUID) {
        double r0 = wave_sum(tb[0]), r1 = wave_sum(tb[1]), r2 = wave_sum(tb[2]), r3 = wave_sum(tb[3]);
        if ((threadIdx.x & 63) == 0) {
            double* d = thbar + 4 * groups[gw].inst;
            atomicAdd(d + 0, r0);
            atomicAdd(d + 1, r1);
            atomicAdd(d + 2, r2);
            atomicAdd(d + 3, r3);
        }
    }
}

// ---------------------------------------------------------------------------- test kernels

__global__ void k_svd3_test(const double* __restrict__ F, int n, double* __restrict__ U, double* __restrict__ S,
                            double* __restrict__ V)
{
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    double f[9], u[9], s[3], v[9];
    for (int a = 0; a < 9; a++) f[a] = F[9 * (long)p + a];
    svd3(f, u, s, v);
    for (int a = 0; a < 9; a++) { U[9 * (long)p + a] = u[a]; V[9 * (long)p + a] = v[a]; }
    for (int a = 0; a < 3; a++) S[3 * (long)p + a] = s[a];
}

__global__ void k_svd3_bwd_test(const double* __restrict__ U, const double* __restrict__ S,
                                const double* __restrict__ V, const double* __restrict__ gU,
                                const double* __restrict__ gS, const double* __restrict__ gV, int n,
                                double* __restrict__ gF)
{
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    double u[9], s[3], v[9], gu[9], gs[3], gv[9], gf[9];
    for (int a = 0; a < 9; a++) {
        u[a] = U[9 * (long)p + a]; v[a] = V[9 * (long)p + a];
        gu[a] = gU[9 * (long)p + a]; gv[a] = gV[9 * (long)p + a];
    }
    for (int a = 0; a < 3; a++) { s[a] = S[3 * (long)p + a]; gs[a] = gS[3 * (long)p + a]; }
    svd3_backward(u, s, v, gu, gs, gv, gf);
    for (int a = 0; a < 9; a++) gF[9 * (long)p + a] = gf[a];
}

}  // namespace ds
