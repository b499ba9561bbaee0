// kernels_ref.hip -- per-particle kernels of the substep and its adjoint (one thread per
// particle or node, grid scatters through global float atomics).
//
// These are the straightforward forms of the substep: one launch per phase, dense grid,
// every P2G contribution an atomic add in L2.  The tiled fused kernels (kernels_tiled.hip)
// are checked against them; they also serve the recompute of the first and last step of
// each adjoint segment.
#include "ds_particle.h"

namespace ds {

// Waits `ticks` of the constant-rate wall clock (100 MHz on gfx9), then ends: put in front of a
// sampled substep's launches so that the host has queued them all before the GPU reaches them
// (timed launches are slower to submit than to run, and a GPU waiting for the host would count
// that wait into the samples).
__global__ void k_hold(long ticks)
{
    const long t0 = (long)wall_clock64();
    while ((long)wall_clock64() - t0 < ticks) __builtin_amdgcn_s_sleep(2);
}

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

__device__ __forceinline__ void load_particle(const float* __restrict__ st, long LP, int p, bool with_F, ParticleIn& q)
{
#pragma unroll
    for (int a = 0; a < 3; a++) q.x[a] = st[(SX + a) * LP + p];
#pragma unroll
    for (int a = 0; a < 3; a++) q.v[a] = st[(SV + a) * LP + p];
#pragma unroll
    for (int a = 0; a < 9; a++) q.C[a] = st[(SC + a) * LP + p];
    if (with_F) {
#pragma unroll
        for (int a = 0; a < 9; a++) q.F[a] = st[(SF + a) * LP + p];
    } else {
#pragma unroll
        for (int a = 0; a < 9; a++) q.F[a] = (a == 0 || a == 4 || a == 8) ? 1.0f : 0.0f;
    }
}

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
    const Group G = groups[g];
    const long LP = P;
    ParticleIn q;
    load_particle(in, LP, p, G.species == LIQUID, q);
    Stencil s;
    if (!stencil(sc, q.x[0], q.x[1], q.x[2], s)) atomicOr(flag, 1u);
    const InstParams ip = load_params(sc, params, G.inst);
    const int n = sc.n_grid;
    ScatterGlobal4 acc{grid + (long)G.inst * n * n * n, n};
    float Fn[9];
    p2g_particle(sc, G.species, ip, q, s, Fn, acc);
    if (fout && G.species == LIQUID) {
#pragma unroll
        for (int a = 0; a < 9; a++) fout[(SF + a) * LP + p] = Fn[a];
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

// Particle update after the gather (simulation_functions.py:246-290): liquid advects, the stick
// moves kinematically and reports dv = v_kin - v_grid, the container keeps x and v = 0.
__device__ __forceinline__ void g2p_write(float* __restrict__ out, long LP, int p, int sp, const float x[3],
                                          const float nv[3], const float nC[9], const float* tr, float dt,
                                          double dv[3])
{
    if (sp == LIQUID) {
#pragma unroll
        for (int a = 0; a < 3; a++) {
            out[(SV + a) * LP + p] = nv[a];
            out[(SX + a) * LP + p] = x[a] + dt * nv[a];
        }
    } else if (sp == STICK) {
#pragma unroll
        for (int a = 0; a < 3; a++) {
            const float vk = (tr[3 + a] - tr[a]) / dt;
            dv[a] = (double)(vk - nv[a]);
            out[(SV + a) * LP + p] = vk;
            out[(SX + a) * LP + p] = x[a] + vk * dt;
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
        const Group G = groups[g];
        float x[3];
#pragma unroll
        for (int a = 0; a < 3; a++) x[a] = in[(SX + a) * LP + p];
        Stencil s;
        if (!stencil(sc, x[0], x[1], x[2], s)) atomicOr(flag, 1u);
        const int n = sc.n_grid;
        GatherVout acc{vout + (long)G.inst * n * n * n, n};
        float nv[3], nC[9];
        g2p_gather(sc, s, acc, nv, nC);
        g2p_write(out, LP, p, G.species, x, nv, nC, traj + ((long)G.inst * T + f) * 3, sc.dt, dv);
    }
    if (dvsum && gw >= 0 && groups[gw].species == STICK) {
        double s0 = wave_sum(dv[0]), s1 = wave_sum(dv[1]), s2 = wave_sum(dv[2]);
        if ((threadIdx.x & 63) == 0) {
            double* d = dvsum + ((long)groups[gw].inst * T + f + 1) * 3;
            atomicAdd(d + 0, exact_grain(s0));
            atomicAdd(d + 1, exact_grain(s1));
            atomicAdd(d + 2, exact_grain(s2));
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
    const Group G = groups[g];
    const long LP = P;
    float x[3];
#pragma unroll
    for (int a = 0; a < 3; a++) x[a] = st[(SX + a) * LP + p];
    Stencil s;
    stencil(sc, x[0], x[1], x[2], s);
    float Cb[9];
#pragma unroll
    for (int a = 0; a < 9; a++) Cb[a] = adj[(SC + a) * LP + p];
    const int n = sc.n_grid;
    GatherVout gv{vout + (long)G.inst * n * n * n, n};
    ScatterGlobal3 acc{vbar + (long)G.inst * n * n * n, n};
    float vt[3] = {0.0f, 0.0f, 0.0f}, xb[3];
    if (G.species == LIQUID) {
        float xa[3];
#pragma unroll
        for (int a = 0; a < 3; a++) {
            xa[a] = adj[(SX + a) * LP + p];
            vt[a] = adj[(SV + a) * LP + p] + sc.dt * xa[a];
        }
        g2p_adj_particle<true>(sc, s, vt, Cb, gv, acc, xb);
#pragma unroll
        for (int a = 0; a < 3; a++) xbar[a * LP + p] = xa[a] + xb[a];
    } else {
        if (G.species == STICK) {
            const double* d = dvbar + ((long)G.inst * T + f + 1) * 3;
#pragma unroll
            for (int a = 0; a < 3; a++) vt[a] = -(float)d[a];
        }
        g2p_adj_particle<false>(sc, s, vt, Cb, gv, acc, xb);
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
        const Group G = groups[g];
        const int n = sc.n_grid;
        GatherPbar acc{pbar + (long)G.inst * n * n * n, n};
        if (G.species == LIQUID) {
            ParticleIn q;
            load_particle(st, LP, p, true, q);
            Stencil s;
            stencil(sc, q.x[0], q.x[1], q.x[2], s);
            const InstParams ip = load_params(sc, params, G.inst);
            float Fnb[9], xp[3], xo[3], vo[3], Co[9], Fo[9];
#pragma unroll
            for (int a = 0; a < 9; a++) Fnb[a] = adj[(SF + a) * LP + p];
#pragma unroll
            for (int a = 0; a < 3; a++) xp[a] = xbar[a * LP + p];
            p2g_adj_liquid(sc, ip, q, s, Fnb, xp, acc, xo, vo, Co, Fo, tb);
#pragma unroll
            for (int a = 0; a < 3; a++) adj[(SX + a) * LP + p] = xo[a];
#pragma unroll
            for (int a = 0; a < 3; a++) adj[(SV + a) * LP + p] = vo[a];
#pragma unroll
            for (int a = 0; a < 9; a++) adj[(SC + a) * LP + p] = Co[a];
#pragma unroll
            for (int a = 0; a < 9; a++) adj[(SF + a) * LP + p] = Fo[a];
        } else {
            float x[3];
#pragma unroll
            for (int a = 0; a < 3; a++) x[a] = st[(SX + a) * LP + p];
            Stencil s;
            stencil(sc, x[0], x[1], x[2], s);
            float Co[9];
            p2g_adj_rigid(sc, species_mass(sc, G.species), s, acc, Co);
#pragma unroll
            for (int a = 0; a < 9; a++) adj[(SC + a) * LP + p] = Co[a];
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

__global__ void k_svd3_values_test(const double* __restrict__ F, int n, double* __restrict__ S)
{
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    double f[9], s[3];
    for (int a = 0; a < 9; a++) f[a] = F[9 * (long)p + a];
    svd3_values(f, s);
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
