// diffstir.hip -- libdiffstir_hip.so: the C ABI of include/diffstir.h over the gfx950 kernels.
//
// One translation unit: the kernel files are included here so that every kernel and the
// host code that launches it are compiled together (no relocatable device code).
//
// Device memory of a context (P = particles of all instances and species, each
// (instance, species) run padded to a multiple of 64; G = n_grid^3 nodes per instance):
//   state      24 x P float   current x, v, C, F (component-major SoA)
//   ckpt       nck x 24 x P   checkpoint every ckpt_every substeps
//   seg        K x 24 x P     states of one segment during the adjoint recompute
//   adj, xbar  24 x P, 3 x P  adjoint state, partial dL/dx
//   grid/vout/vbar/pbar       4 x B x G float4 (momentum+mass, velocity, their adjoints)
//   dvsum/dvbar               B x T x 3 double (stick sum of dv and its adjoint)
//   params/thbar              B x 4 double
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/diffstir.h"
#include "ds_device.h"
#include "kernels_ref.hip"

using namespace ds;

struct ds_ctx {
    ds_scene_desc desc;
    DevScene dsc;
    int B = 0, T = 0, K = 0, ncp = 0;
    int P = 0;
    std::vector<int> nl, ns, nc;
    std::vector<Group> groups;
    std::vector<int> meta_h;
    std::vector<float> x0_h;      // [P][3] initial positions in particle order
    std::vector<float> traj_h;    // [B][T][3]
    std::vector<double> params_h; // [B][4]
    int step = 0;
    bool reset_done = false;

    int* d_meta = nullptr;
    Group* d_groups = nullptr;
    float* d_x0 = nullptr;
    float* d_state = nullptr;
    float* d_ckpt = nullptr;
    float* d_seg = nullptr;
    float* d_adj = nullptr;
    float* d_xbar = nullptr;
    float4* d_grid = nullptr;
    float4* d_vout = nullptr;
    float4* d_vbar = nullptr;
    float4* d_pbar = nullptr;
    double* d_dvsum = nullptr;
    double* d_dvbar = nullptr;
    double* d_params = nullptr;
    double* d_thbar = nullptr;
    float* d_traj = nullptr;
    unsigned* d_flag = nullptr;

    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    double fwd_ms = 0.0, bwd_ms = 0.0;
    std::string err;

    long gridNodes() const { return (long)desc.n_grid * desc.n_grid * desc.n_grid; }
    long stateFloats() const { return (long)NSLOT * P; }
};

namespace {

thread_local std::string g_create_err;

int fail(ds_ctx* c, int code, const std::string& msg)
{
    if (c) c->err = msg;
    else g_create_err = msg;
    return code;
}

#define HIPCHK(ctx, expr)                                                                       \
    do {                                                                                        \
        hipError_t e_ = (expr);                                                                 \
        if (e_ != hipSuccess)                                                                   \
            return fail(ctx, DS_EHIP, std::string(#expr) + ": " + hipGetErrorString(e_));       \
    } while (0)

template <typename T>
int dalloc(ds_ctx* c, T** p, size_t count)
{
    if (count == 0) count = 1;
    hipError_t e = hipMalloc((void**)p, count * sizeof(T));
    if (e != hipSuccess) {
        *p = nullptr;
        return fail(c, DS_ENOMEM, std::string("hipMalloc of ") + std::to_string(count * sizeof(T)) +
                                      " bytes failed: " + hipGetErrorString(e));
    }
    return DS_OK;
}

inline unsigned nblocks(long n, int bs = 256) { return (unsigned)((n + bs - 1) / bs); }

int fill_zero_f4(ds_ctx* c, float4* p, long n)
{
    unsigned nb = std::min<long>(nblocks(n), 8192);
    hipLaunchKernelGGL(k_fill_f4, dim3(nb), dim3(256), 0, c->stream, p, n);
    return DS_OK;
}

int fill_zero_f32(ds_ctx* c, float* p, long n)
{
    unsigned nb = std::min<long>(nblocks(n), 8192);
    hipLaunchKernelGGL(k_fill_f32, dim3(nb), dim3(256), 0, c->stream, p, n, 0.0f);
    return DS_OK;
}

int fill_zero_f64(ds_ctx* c, double* p, long n)
{
    unsigned nb = std::min<long>(nblocks(n), 8192);
    hipLaunchKernelGGL(k_fill_f64, dim3(nb), dim3(256), 0, c->stream, p, n);
    return DS_OK;
}

DevScene make_dev_scene(const ds_scene_desc& d)
{
    DevScene s;
    s.n_grid = d.n_grid;
    s.bound = d.bound;
    s.model = d.model;
    s.k_from_param = d.k_from_param;
    double dx = 1.0 / d.n_grid, inv_dx = (double)d.n_grid;
    s.dx = (float)dx;
    s.inv_dx = (float)inv_dx;
    s.dt = (float)d.dt;
    s.gdt = (float)(d.dt * d.gravity);
    s.wall_r2 = (float)(d.wall_r * d.wall_r);
    s.wall_cx = (float)d.wall_cx;
    s.wall_cz = (float)d.wall_cz;
    s.m_liq = (float)d.m_liq;
    s.m_stick = (float)d.m_stick;
    s.m_cont = (float)d.m_cont;
    s.alpha = (float)(-d.dt * d.vol_liq * 4.0 * inv_dx * inv_dx);
    s.c4 = (float)(4.0 * inv_dx);
    s.k_bulk = d.k_bulk;
    return s;
}

int check_flag(ds_ctx* c, const char* where)
{
    unsigned f = 0;
    HIPCHK(c, hipMemcpyAsync(&f, c->d_flag, sizeof(unsigned), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (f) {
        unsigned z = 0;
        HIPCHK(c, hipMemcpy(c->d_flag, &z, sizeof(unsigned), hipMemcpyHostToDevice));
        return fail(c, DS_ENONFINITE, std::string(where) +
                                          ": a particle left the grid's interior or became non-finite "
                                          "(the simulation is unstable for these parameters)");
    }
    return DS_OK;
}

// One forward substep with the per-particle kernels: state `in` at step f -> `out` at f+1.
void step_ref(ds_ctx* c, const float* in, float* out, int f, bool record)
{
    const long G = c->gridNodes();
    fill_zero_f4(c, c->d_grid, G * c->B);
    hipLaunchKernelGGL(k_p2g_ref, dim3(nblocks(c->P)), dim3(256), 0, c->stream, in, out, c->d_meta, c->d_groups,
                       c->d_params, c->d_grid, c->dsc, c->P, c->d_flag);
    hipLaunchKernelGGL(k_gridop_ref, dim3(nblocks(G * c->B)), dim3(256), 0, c->stream, c->d_grid, c->d_vout,
                       c->dsc, c->B);
    hipLaunchKernelGGL(k_g2p_ref, dim3(nblocks(c->P)), dim3(256), 0, c->stream, in, out, c->d_meta, c->d_groups,
                       c->d_vout, c->d_traj, c->T, f, record ? c->d_dvsum : nullptr, c->dsc, c->P, c->d_flag);
}

// Adjoint of one substep with the per-particle kernels: `st` is the state at step f.
void step_adj_ref(ds_ctx* c, const float* st, int f)
{
    const long G = c->gridNodes();
    fill_zero_f4(c, c->d_grid, G * c->B);
    hipLaunchKernelGGL(k_p2g_ref, dim3(nblocks(c->P)), dim3(256), 0, c->stream, st, (float*)nullptr, c->d_meta,
                       c->d_groups, c->d_params, c->d_grid, c->dsc, c->P, c->d_flag);
    hipLaunchKernelGGL(k_gridop_ref, dim3(nblocks(G * c->B)), dim3(256), 0, c->stream, c->d_grid, c->d_vout,
                       c->dsc, c->B);
    fill_zero_f4(c, c->d_vbar, G * c->B);
    hipLaunchKernelGGL(k_g2p_adj_ref, dim3(nblocks(c->P)), dim3(256), 0, c->stream, st, c->d_adj, c->d_meta,
                       c->d_groups, c->d_vout, c->d_dvbar, c->T, f, c->d_vbar, c->d_xbar, c->dsc, c->P);
    hipLaunchKernelGGL(k_gridop_adj_ref, dim3(nblocks(G * c->B)), dim3(256), 0, c->stream, c->d_grid, c->d_vbar,
                       c->d_pbar, c->dsc, c->B);
    hipLaunchKernelGGL(k_p2g_adj_ref, dim3(nblocks(c->P)), dim3(256), 0, c->stream, st, c->d_adj, c->d_xbar,
                       c->d_meta, c->d_groups, c->d_params, c->d_pbar, c->d_thbar, c->dsc, c->P);
}

void free_all(ds_ctx* c)
{
    void* ptrs[] = {c->d_meta, c->d_groups, c->d_x0, c->d_state, c->d_ckpt, c->d_seg, c->d_adj, c->d_xbar,
                    c->d_grid, c->d_vout, c->d_vbar, c->d_pbar, c->d_dvsum, c->d_dvbar, c->d_params,
                    c->d_thbar, c->d_traj, c->d_flag};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    if (c->stream) (void)hipStreamDestroy(c->stream);
}

}  // namespace

extern "C" {

int ds_create(const ds_scene_desc* desc, int32_t n_inst, const int32_t* n_liquid, const int32_t* n_stick,
              const int32_t* n_container, int32_t max_steps, int32_t ckpt_every, ds_ctx** out)
{
    if (!out) return fail(nullptr, DS_EINVAL, "out is NULL");
    *out = nullptr;
    if (!desc || !n_liquid || !n_stick || !n_container) return fail(nullptr, DS_EINVAL, "NULL argument");
    if (n_inst < 1) return fail(nullptr, DS_EINVAL, "n_inst must be >= 1");
    if (desc->n_grid < 8 || desc->n_grid > 1024) return fail(nullptr, DS_EINVAL, "n_grid out of range [8, 1024]");
    if (desc->model < 0 || desc->model > 2) return fail(nullptr, DS_EINVAL, "model must be 0, 1 or 2");
    if (desc->k_from_param && desc->model != DS_MODEL_HERSCHEL_BULKLEY)
        return fail(nullptr, DS_EINVAL, "k_from_param needs the Herschel-Bulkley model (params[3] is free only there)");
    if (!(desc->dt > 0.0)) return fail(nullptr, DS_EINVAL, "dt must be > 0");
    if (max_steps < 2) return fail(nullptr, DS_EINVAL, "max_steps must be >= 2");
    if (ckpt_every < 1) return fail(nullptr, DS_EINVAL, "ckpt_every must be >= 1");

    ds_ctx* c = new ds_ctx();
    c->desc = *desc;
    c->dsc = make_dev_scene(*desc);
    c->B = n_inst;
    c->T = max_steps;
    c->K = ckpt_every;
    c->ncp = (max_steps - 2) / ckpt_every + 1;
    long P = 0;
    for (int b = 0; b < n_inst; b++) {
        int cnt[3] = {n_liquid[b], n_stick[b], n_container[b]};
        c->nl.push_back(cnt[0]);
        c->ns.push_back(cnt[1]);
        c->nc.push_back(cnt[2]);
        for (int sp = 0; sp < 3; sp++) {
            if (cnt[sp] < 0) {
                delete c;
                return fail(nullptr, DS_EINVAL, "negative particle count");
            }
            Group g{b, sp, (int)P, cnt[sp]};
            c->groups.push_back(g);
            P += ((long)cnt[sp] + 63) / 64 * 64;
        }
    }
    if (P < 64) P = 64;
    if (P > (1l << 30)) {
        delete c;
        return fail(nullptr, DS_EINVAL, "too many particles");
    }
    c->P = (int)P;
    c->meta_h.assign(P, -1);
    for (size_t g = 0; g < c->groups.size(); g++)
        for (int i = 0; i < c->groups[g].count; i++) c->meta_h[c->groups[g].offset + i] = (int)g;
    c->x0_h.assign(3 * P, 0.5f);
    c->traj_h.assign((size_t)c->B * c->T * 3, 0.0f);
    c->params_h.assign((size_t)c->B * 4, 0.0);

    int rc;
    const long G = c->gridNodes();
#define ALLOC(ptr, n)                      \
    if ((rc = dalloc(c, &ptr, (n))) != 0) { \
        std::string m = c->err;            \
        free_all(c);                       \
        delete c;                          \
        return fail(nullptr, rc, m);       \
    }
    ALLOC(c->d_meta, (size_t)P);
    ALLOC(c->d_groups, c->groups.size());
    ALLOC(c->d_x0, (size_t)3 * P);
    ALLOC(c->d_state, (size_t)c->stateFloats());
    ALLOC(c->d_ckpt, (size_t)c->stateFloats() * c->ncp);
    ALLOC(c->d_seg, (size_t)c->stateFloats() * c->K);
    ALLOC(c->d_adj, (size_t)c->stateFloats());
    ALLOC(c->d_xbar, (size_t)3 * P);
    ALLOC(c->d_grid, (size_t)G * c->B);
    ALLOC(c->d_vout, (size_t)G * c->B);
    ALLOC(c->d_vbar, (size_t)G * c->B);
    ALLOC(c->d_pbar, (size_t)G * c->B);
    ALLOC(c->d_dvsum, (size_t)c->B * c->T * 3);
    ALLOC(c->d_dvbar, (size_t)c->B * c->T * 3);
    ALLOC(c->d_params, (size_t)c->B * 4);
    ALLOC(c->d_thbar, (size_t)c->B * 4);
    ALLOC(c->d_traj, (size_t)c->B * c->T * 3);
    ALLOC(c->d_flag, 1);
#undef ALLOC
    hipError_t e = hipStreamCreate(&c->stream);
    if (e == hipSuccess) e = hipEventCreate(&c->ev0);
    if (e == hipSuccess) e = hipEventCreate(&c->ev1);
    if (e == hipSuccess) e = hipMemcpy(c->d_meta, c->meta_h.data(), sizeof(int) * P, hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = hipMemcpy(c->d_groups, c->groups.data(), sizeof(Group) * c->groups.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemset(c->d_flag, 0, sizeof(unsigned));
    if (e != hipSuccess) {
        std::string m = hipGetErrorString(e);
        free_all(c);
        delete c;
        return fail(nullptr, DS_EHIP, "context set-up: " + m);
    }
    *out = c;
    return DS_OK;
}

void ds_destroy(ds_ctx* ctx)
{
    if (!ctx) return;
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    free_all(ctx);
    delete ctx;
}

const char* ds_last_error(const ds_ctx* ctx)
{
    return ctx ? ctx->err.c_str() : g_create_err.c_str();
}

int ds_set_particles(ds_ctx* c, int32_t inst, const float* xl, const float* xs, const float* xc)
{
    if (!c) return DS_EINVAL;
    if (inst < 0 || inst >= c->B) return fail(c, DS_EINVAL, "instance index out of range");
    const float* src[3] = {xl, xs, xc};
    for (int sp = 0; sp < 3; sp++) {
        const Group& g = c->groups[inst * 3 + sp];
        if (g.count > 0 && !src[sp]) return fail(c, DS_EINVAL, "NULL positions for a non-empty species");
        for (int i = 0; i < g.count; i++)
            for (int a = 0; a < 3; a++) c->x0_h[3 * ((size_t)g.offset + i) + a] = src[sp][3 * (size_t)i + a];
    }
    c->reset_done = false;
    return DS_OK;
}

int ds_set_trajectory(ds_ctx* c, int32_t inst, const float* traj)
{
    if (!c) return DS_EINVAL;
    if (inst < 0 || inst >= c->B || !traj) return fail(c, DS_EINVAL, "bad instance or NULL trajectory");
    std::memcpy(c->traj_h.data() + (size_t)inst * c->T * 3, traj, sizeof(float) * 3 * c->T);
    HIPCHK(c, hipMemcpyAsync(c->d_traj + (size_t)inst * c->T * 3, c->traj_h.data() + (size_t)inst * c->T * 3,
                             sizeof(float) * 3 * c->T, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return DS_OK;
}

int ds_set_params(ds_ctx* c, const double* params)
{
    if (!c) return DS_EINVAL;
    if (!params) return fail(c, DS_EINVAL, "NULL params");
    for (int i = 0; i < 4 * c->B; i++)
        if (!std::isfinite(params[i])) return fail(c, DS_EINVAL, "non-finite parameter");
    std::memcpy(c->params_h.data(), params, sizeof(double) * 4 * c->B);
    HIPCHK(c, hipMemcpyAsync(c->d_params, c->params_h.data(), sizeof(double) * 4 * c->B, hipMemcpyHostToDevice,
                             c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return DS_OK;
}

int ds_reset(ds_ctx* c)
{
    if (!c) return DS_EINVAL;
    HIPCHK(c, hipMemcpyAsync(c->d_x0, c->x0_h.data(), sizeof(float) * 3 * c->P, hipMemcpyHostToDevice, c->stream));
    hipLaunchKernelGGL(k_init_state, dim3(nblocks(c->P)), dim3(256), 0, c->stream, c->d_state, c->d_x0, c->d_meta,
                       c->P, 0.5f);
    fill_zero_f64(c, c->d_dvsum, (long)c->B * c->T * 3);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->step = 0;
    c->reset_done = true;
    return DS_OK;
}

int ds_forward(ds_ctx* c, int32_t nsteps)
{
    if (!c) return DS_EINVAL;
    if (!c->reset_done) return fail(c, DS_ESTATE, "ds_reset must be called before ds_forward");
    if (nsteps < 0 || c->step + nsteps > c->T - 1)
        return fail(c, DS_EINVAL, "nsteps exceeds the trajectory (max_steps - 1 substeps in total)");
    HIPCHK(c, hipEventRecord(c->ev0, c->stream));
    const long SF = c->stateFloats();
    for (int i = 0; i < nsteps; i++) {
        int s = c->step;
        if (s % c->K == 0)
            HIPCHK(c, hipMemcpyAsync(c->d_ckpt + (size_t)(s / c->K) * SF, c->d_state, sizeof(float) * SF,
                                     hipMemcpyDeviceToDevice, c->stream));
        step_ref(c, c->d_state, c->d_state, s, true);
        c->step++;
    }
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipEventRecord(c->ev1, c->stream));
    HIPCHK(c, hipEventSynchronize(c->ev1));
    float ms = 0.0f;
    HIPCHK(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
    c->fwd_ms = ms;
    return check_flag(c, "ds_forward");
}

int ds_current_step(const ds_ctx* c, int32_t* step)
{
    if (!c || !step) return DS_EINVAL;
    *step = c->step;
    return DS_OK;
}

int ds_get_forces(ds_ctx* c, double* forces)
{
    if (!c) return DS_EINVAL;
    if (!forces) return fail(c, DS_EINVAL, "NULL forces");
    size_t n = (size_t)c->B * c->T * 3;
    HIPCHK(c, hipMemcpyAsync(forces, c->d_dvsum, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    double scale = c->desc.m_stick / c->desc.dt;
    for (size_t i = 0; i < n; i++) forces[i] *= scale;
    return DS_OK;
}

int ds_backward(ds_ctx* c, const double* dl_dforces, double* dl_dparams)
{
    if (!c) return DS_EINVAL;
    if (!dl_dforces || !dl_dparams) return fail(c, DS_EINVAL, "NULL argument");
    if (!c->reset_done || c->step < 1) return fail(c, DS_ESTATE, "ds_backward needs a recorded forward rollout");
    const int S = c->step;
    // dL/d(dv sum at step s) = m_stick/dt * dL/dF(s)
    std::vector<double> dvb((size_t)c->B * c->T * 3);
    double scale = c->desc.m_stick / c->desc.dt;
    for (size_t i = 0; i < dvb.size(); i++) dvb[i] = scale * dl_dforces[i];
    HIPCHK(c, hipMemcpyAsync(c->d_dvbar, dvb.data(), sizeof(double) * dvb.size(), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipEventRecord(c->ev0, c->stream));
    const long SF = c->stateFloats();
    fill_zero_f32(c, c->d_adj, SF);
    fill_zero_f64(c, c->d_thbar, (long)c->B * 4);
    for (int cp = (S - 1) / c->K; cp >= 0; cp--) {
        int s0 = cp * c->K;
        int s1 = std::min(s0 + c->K, S);
        HIPCHK(c, hipMemcpyAsync(c->d_seg, c->d_ckpt + (size_t)cp * SF, sizeof(float) * SF, hipMemcpyDeviceToDevice,
                                 c->stream));
        for (int s = s0; s < s1 - 1; s++) {
            int k = s - s0;
            step_ref(c, c->d_seg + (size_t)k * SF, c->d_seg + (size_t)(k + 1) * SF, s, false);
        }
        for (int s = s1 - 1; s >= s0; s--) step_adj_ref(c, c->d_seg + (size_t)(s - s0) * SF, s);
    }
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipEventRecord(c->ev1, c->stream));
    HIPCHK(c, hipMemcpyAsync(dl_dparams, c->d_thbar, sizeof(double) * 4 * c->B, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipEventSynchronize(c->ev1));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    float ms = 0.0f;
    HIPCHK(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
    c->bwd_ms = ms;
    return check_flag(c, "ds_backward");
}

static int read_group(ds_ctx* c, int gidx, int slot0, int ncomp, float* out)
{
    if (!out) return DS_OK;
    const Group& g = c->groups[gidx];
    if (g.count == 0) return DS_OK;
    std::vector<float> tmp((size_t)ncomp * g.count);
    for (int a = 0; a < ncomp; a++)
        HIPCHK(c, hipMemcpyAsync(tmp.data() + (size_t)a * g.count, c->d_state + (size_t)(slot0 + a) * c->P + g.offset,
                                 sizeof(float) * g.count, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (int i = 0; i < g.count; i++)
        for (int a = 0; a < ncomp; a++) out[(size_t)i * ncomp + a] = tmp[(size_t)a * g.count + i];
    return DS_OK;
}

int ds_read_liquid(ds_ctx* c, int32_t inst, float* x, float* v, float* C, float* F)
{
    if (!c) return DS_EINVAL;
    if (inst < 0 || inst >= c->B) return fail(c, DS_EINVAL, "instance index out of range");
    int g = inst * 3 + LIQUID, rc;
    if ((rc = read_group(c, g, SX, 3, x))) return rc;
    if ((rc = read_group(c, g, SV, 3, v))) return rc;
    if ((rc = read_group(c, g, SC, 9, C))) return rc;
    return read_group(c, g, SF, 9, F);
}

int ds_read_stick(ds_ctx* c, int32_t inst, float* x, float* v, float* C)
{
    if (!c) return DS_EINVAL;
    if (inst < 0 || inst >= c->B) return fail(c, DS_EINVAL, "instance index out of range");
    int g = inst * 3 + STICK, rc;
    if ((rc = read_group(c, g, SX, 3, x))) return rc;
    if ((rc = read_group(c, g, SV, 3, v))) return rc;
    return read_group(c, g, SC, 9, C);
}

int ds_synchronize(ds_ctx* c)
{
    if (!c) return DS_EINVAL;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return DS_OK;
}

int ds_last_timing(const ds_ctx* c, double* fwd_ms, double* bwd_ms)
{
    if (!c) return DS_EINVAL;
    if (fwd_ms) *fwd_ms = c->fwd_ms;
    if (bwd_ms) *bwd_ms = c->bwd_ms;
    return DS_OK;
}

int ds_svd3(const double* F, int32_t n, double* U, double* S, double* V)
{
    if (!F || !U || !S || !V || n < 1) return fail(nullptr, DS_EINVAL, "bad argument");
    double *dF, *dU, *dS, *dV;
    if (hipMalloc((void**)&dF, 9 * sizeof(double) * n) != hipSuccess) return fail(nullptr, DS_ENOMEM, "hipMalloc");
    if (hipMalloc((void**)&dU, 9 * sizeof(double) * n) != hipSuccess) { (void)hipFree(dF); return fail(nullptr, DS_ENOMEM, "hipMalloc"); }
    if (hipMalloc((void**)&dS, 3 * sizeof(double) * n) != hipSuccess) { (void)hipFree(dF); (void)hipFree(dU); return fail(nullptr, DS_ENOMEM, "hipMalloc"); }
    if (hipMalloc((void**)&dV, 9 * sizeof(double) * n) != hipSuccess) { (void)hipFree(dF); (void)hipFree(dU); (void)hipFree(dS); return fail(nullptr, DS_ENOMEM, "hipMalloc"); }
    hipError_t e = hipMemcpy(dF, F, 9 * sizeof(double) * n, hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_svd3_test, dim3(nblocks(n)), dim3(256), 0, 0, dF, n, dU, dS, dV);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpy(U, dU, 9 * sizeof(double) * n, hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(S, dS, 3 * sizeof(double) * n, hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(V, dV, 9 * sizeof(double) * n, hipMemcpyDeviceToHost);
    (void)hipFree(dF); (void)hipFree(dU); (void)hipFree(dS); (void)hipFree(dV);
    if (e != hipSuccess) return fail(nullptr, DS_EHIP, hipGetErrorString(e));
    return DS_OK;
}

int ds_svd3_backward(const double* U, const double* S, const double* V, const double* gU, const double* gS,
                     const double* gV, int32_t n, double* gF)
{
    if (!U || !S || !V || !gU || !gS || !gV || !gF || n < 1) return fail(nullptr, DS_EINVAL, "bad argument");
    const size_t nb9 = 9 * sizeof(double) * n, nb3 = 3 * sizeof(double) * n;
    double* buf;
    if (hipMalloc((void**)&buf, 5 * nb9 + 2 * nb3) != hipSuccess) return fail(nullptr, DS_ENOMEM, "hipMalloc");
    double *dU = buf, *dV = dU + 9 * (size_t)n, *dgU = dV + 9 * (size_t)n, *dgV = dgU + 9 * (size_t)n,
           *dgF = dgV + 9 * (size_t)n, *dS = dgF + 9 * (size_t)n, *dgS = dS + 3 * (size_t)n;
    hipError_t e = hipMemcpy(dU, U, nb9, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dV, V, nb9, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dgU, gU, nb9, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dgV, gV, nb9, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dS, S, nb3, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dgS, gS, nb3, hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_svd3_bwd_test, dim3(nblocks(n)), dim3(256), 0, 0, dU, dS, dV, dgU, dgS, dgV, n, dgF);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpy(gF, dgF, nb9, hipMemcpyDeviceToHost);
    (void)hipFree(buf);
    if (e != hipSuccess) return fail(nullptr, DS_EHIP, hipGetErrorString(e));
    return DS_OK;
}

}  // extern "C"
