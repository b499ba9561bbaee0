// diffstir.hip -- libdiffstir_hip.so: the C ABI of include/diffstir.h over the gfx950 kernels.
//
// One translation unit: the kernel files are included here so that every kernel and the
// host code that launches it are compiled together (no relocatable device code).
//
// Device memory of a context (P = particle slots of all instances and species: each
// (instance, species) run padded to a multiple of 64 in the initial layout; G = n_grid^3
// nodes per instance, B instances, K = checkpoint interval):
//   state           24 x P float   current x, v, C, F (component-major SoA)
//   ckpt            nck x 24 x P   state at every K-th substep (sorted order in tiled mode)
//   seg             K x 24 x P     states of one segment during the adjoint recompute
//   adj, adj2, xbar 24 x P, 24 x P, 3 x P   adjoint state (+ re-sort buffer), partial dL/dx
//   dvsum/dvbar     B x T x 3 double (stick sum of dv per step and its adjoint)
//   params/thbar    B x 4 double
// per-particle kernels:  grid, vout, vbar, pbar       4 x B x G float4
// tiled kernels:         grid[2], vbar[2]             4 x B x G float4 (double buffers)
//                        seg_grid                     K x B x G float4 (momentum/mass per step)
//                        bucket sort scratch, chunk tables and re-sort maps per checkpoint
#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/diffstir.h"
#include "ds_particle.h"
#include "kernels_ref.hip"
#include "kernels_tiled.hip"

using namespace ds;

struct ds_ctx {
    ds_scene_desc desc;
    DevScene dsc;
    int B = 0, T = 0, K = 0, ncp = 0;
    int sort_every = 0;                // substeps between bucket sorts (0: the checkpoint interval)
    int E = 1, nslot = 0;              // sorts per checkpoint segment, sort table slots (ncp x E)
    int P = 0, nreal = 0;
    int mode = DS_KERNELS_TILED;       // requested
    int ppt = 2;                       // rows per wavefront of the tiled scatter kernels (chunk = 4 * ppt rows)
    int ppt_sorted = 2;                // ppt the current chunk tables were built with
    int gppt = 1;                      // particles per thread of the tiled gather kernels (256 * gppt)
    int gppt_sorted = 1;
    bool adj_split = false;            // adjoint P2G as two kernels (gather, stress) instead of one
    int rparts = 1;                    // wavefronts per node block of the grid reductions (1, 2, 4)
    int g2p_split = 1;                 // workgroups per gather chunk in G2P (DS_OPT_G2P_SPLIT)
    bool fuse_step = false;            // G2P fused with the next step's P2G (DS_OPT_FUSED_STEP)
    bool packed = false;               // scatter kernels on the packed (gather) chunks with claim passes
    bool use_graph = false;            // replay whole-rollout forward / adjoint launch sequences as hipGraphs
    bool overlap = true;               // adjoint: recompute segment k-1 on a second stream during segment k's adjoint
    long chunk_grid = 16384;           // workgroups of a chunk kernel launch (chunks beyond it loop)
    bool capturing = false;            // a launch sequence is being captured (no kernel timing)
    hipGraphExec_t fwd_exec = nullptr, bwd_exec = nullptr;
    int alloc_mode = -1;               // mode whose buffers are allocated
    std::vector<int> nl, ns, nc;
    std::vector<Group> groups;
    std::vector<int> meta_h, pid_h;
    std::vector<float> x0_h;      // [P][3] initial positions in the initial layout
    std::vector<float> traj_h;    // [B][T][3]
    std::vector<double> params_h; // [B][4]
    int step = 0;
    bool reset_done = false;
    // the parameters or a trajectory changed after the rollout was recorded: the adjoint's
    // recompute would replay a different rollout than the recorded forces (cleared by ds_reset)
    bool tape_stale = false;

    // common
    int* d_meta = nullptr;
    Group* d_groups = nullptr;
    float* d_x0 = nullptr;
    float* d_state = nullptr;
    float* d_ckpt = nullptr;
    float* d_seg = nullptr;
    float* d_adj = nullptr;
    float* d_xbar = nullptr;
    double* d_dvsum = nullptr;
    double* d_dvbar = nullptr;
    double* d_params = nullptr;
    double* d_thbar = nullptr;
    float* d_traj = nullptr;
    unsigned* d_flag = nullptr;
    // per-particle kernels
    float4* d_grid = nullptr;
    float4* d_vout = nullptr;
    float4* d_vbar = nullptr;
    float4* d_pbar = nullptr;
    // tiled kernels
    float4* d_vgrid = nullptr;     // grid velocity of the current step
    float4* d_pbar2 = nullptr;     // dL/d(momentum, mass) of the current adjoint step
    double* d_pfb = nullptr;       // fall-back momentum/mass grid, FP64 x 4 per node (stencils outside their tile)
    double* d_vbfb = nullptr;      // fall-back dL/dv grid, FP64 x 4 per node
    float4* d_staging = nullptr;   // per-chunk staging slots
    long staging_cap = 0;          // slots allocated
    float4* d_seg_grid = nullptr;
    float* d_seg_aff = nullptr;    // K x 9 x P liquid affine A per step of a segment
    float* d_adj2 = nullptr;
    int* d_meta2 = nullptr;
    int* d_pid = nullptr;
    int* d_pid2 = nullptr;
    int* d_cnt = nullptr;
    int* d_celloff = nullptr;
    int* d_bcount = nullptr;
    int* d_bnch = nullptr;         // ncp x NBK chunks per bucket
    int* d_bstart = nullptr;
    int* d_cstart = nullptr;       // ncp x NBK first chunk of each bucket
    int* d_tot = nullptr;
    int* d_tsum = nullptr;        // 4 x tiles: tile sums of the bucket scans
    int* d_key = nullptr;
    int* d_rank = nullptr;
    int* d_rrpos = nullptr;
    int* d_clist = nullptr;        // P: particles in cell-major order within buckets (sort scratch)
    Chunk* d_chunks = nullptr;    // ncp x maxch
    int* d_rows = nullptr;        // ncp x (P + 1) row table: first slot of every row
    GChunk* d_gchunks = nullptr;  // ncp x maxgch chunks of the gather-only kernels
    int* d_bgch = nullptr;        // ncp x NBK gather chunks per bucket
    int* d_gstart = nullptr;      // ncp x NBK first gather chunk of each bucket
    int* d_cov = nullptr;         // ncp x B x NR^3 reduction blocks reached by a chunk tile
    int* d_fbflag = nullptr;      // B x NR^3 reduction blocks marked by a fall-back scatter this step
    double* d_j = nullptr;        // P: J of each liquid particle for the next P2G
    double* d_seg_mj = nullptr;   // K x 10 x P: J and its SVD backward factor per step of a segment
    float* d_gadj = nullptr;      // 18 x P: gathered quantities of the adjoint P2G
    double* d_thpart = nullptr;   // maxgch x 4: per-chunk parameter-gradient partials of a segment
    float* d_seg_tmp = nullptr;   // a step's state before the permutation of a mid-segment sort (E > 1)
    float* d_seg_tmp_b = nullptr; // the same for the second segment buffer set
    // second segment buffer set and the recompute's own scratch, for the overlapped adjoint
    float* d_seg_b = nullptr;
    float4* d_seg_grid_b = nullptr;
    float* d_seg_aff_b = nullptr;
    double* d_seg_mj_b = nullptr;
    float4* d_staging_r = nullptr;
    int* d_fbflag_r = nullptr;
    bool overlap_alloc = false;
    hipStream_t stream2 = nullptr;     // the recompute stream (default priority, like the main stream)
    hipEvent_t ev_fork = nullptr, ev_rec[2] = {nullptr, nullptr}, ev_adj[2] = {nullptr, nullptr};
    bool xgrad_stream = false;         // adjoint P2G's position gradient on its own stream (DS_OPT_XGRAD_STREAM)
    bool xgrad_alloc = false;
    hipStream_t stream3 = nullptr;     // the position-gradient stream (default priority)
    hipEvent_t ev_step = nullptr, ev_xg = nullptr;
    float* d_xp = nullptr;             // [3][P] position gradient of the current adjoint step
    int* d_brows = nullptr;       // rows per bucket
    int* d_rstart = nullptr;      // first row of each bucket
    int* d_src = nullptr;         // ncp x P
    int NB = 0, NBK = 0;
    int NR = 0;                   // reduction blocks per axis (n / 4 + 1: block r owns nodes [4r - 1, 4r + 3))
    long maxch = 0, maxgch = 0;
    int* d_counts = nullptr;      // ncp x 4: sort totals of each checkpoint (particles, chunks, rows,
                                  // gather chunks), read by the kernels; the host never reads them back

    // sampled per-kernel timing
    int kt_every = 0;
    bool kt_on = false;                                    // sampling the current substep
    std::vector<hipEvent_t> kt_pool;
    std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> kt_pending;
    double kt_ms[DS_NKERNELS] = {0};
    int64_t kt_n[DS_NKERNELS] = {0};

    hipStream_t stream = nullptr;       // where the context's work is queued: own_stream or the caller's
    hipStream_t own_stream = nullptr;   // created with the context
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

#define RCHK(expr)                \
    do {                          \
        int r_ = (expr);          \
        if (r_ != DS_OK) return r_; \
    } while (0)

template <typename T>
int dalloc(ds_ctx* c, T** p, size_t count)
{
    if (count == 0) count = 1;
    hipError_t e = hipMalloc((void**)p, count * sizeof(T));
    if (e != hipSuccess) {
        *p = nullptr;
        (void)hipGetLastError();
        return fail(c, DS_ENOMEM, std::string("hipMalloc of ") + std::to_string(count * sizeof(T)) +
                                      " bytes failed: " + hipGetErrorString(e));
    }
    return DS_OK;
}

template <typename T>
void dfree(T*& p)
{
    if (p) (void)hipFree(p);
    p = nullptr;
}

inline unsigned nblocks(long n, int bs = 256) { return (unsigned)((n + bs - 1) / bs); }

void fill_zero_f4(ds_ctx* c, float4* p, long n)
{
    unsigned nb = (unsigned)std::min<long>(nblocks(n), 8192);
    hipLaunchKernelGGL(k_fill_f4, dim3(nb), dim3(256), 0, c->stream, p, n);
}

void fill_zero_f32(ds_ctx* c, float* p, long n)
{
    unsigned nb = (unsigned)std::min<long>(nblocks(n), 8192);
    hipLaunchKernelGGL(k_fill_f32, dim3(nb), dim3(256), 0, c->stream, p, n, 0.0f);
}

void fill_zero_f64(ds_ctx* c, double* p, long n)
{
    unsigned nb = (unsigned)std::min<long>(nblocks(n), 8192);
    hipLaunchKernelGGL(k_fill_f64, dim3(nb), dim3(256), 0, c->stream, p, n);
}

void fill_i32(ds_ctx* c, int* p, long n, int v)
{
    unsigned nb = (unsigned)std::min<long>(nblocks(n), 8192);
    hipLaunchKernelGGL(k_fill_i32, dim3(nb), dim3(256), 0, c->stream, p, n, v);
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
    s.tile_win = TS - 3;
    s.count_fallback = 0;
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
        if (f & 2u)
            return fail(c, DS_EHIP, std::string(where) + ": the bucket sort's totals do not match its tables");
        return fail(c, DS_ENONFINITE, std::string(where) +
                                          ": a particle left the grid's interior or became non-finite "
                                          "(the simulation is unstable for these parameters)");
    }
    return DS_OK;
}

// ---------------------------------------------------------------------------- kernel timing

hipEvent_t kt_event(ds_ctx* c)
{
    if (!c->kt_pool.empty()) {
        hipEvent_t e = c->kt_pool.back();
        c->kt_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

// Times the launches of its scope when the current substep is sampled.  Bracketing form (the
// default): events recorded on the stream before and after the scope's launches, so the time
// includes the dispatch gaps between them.  Dispatch form (`dispatch` true): the scope makes one
// launch through DS_LAUNCH_T / DS_PPT_DISPATCH*_T, which hands both events to
// hipExtLaunchKernel.  Measured on MI355X against the kernel trace: the dispatch after a packet
// that records an event starts about 4.6 us after it (untimed launches follow each other with no
// gap), so a sample is the kernel's execution plus that gap; kt_hold keeps the GPU from also
// waiting for the host, which submits timed launches more slowly than they run.
struct KernelTimer {
    ds_ctx* c;
    int id;
    bool dispatch;
    hipEvent_t a = nullptr, b = nullptr;
    KernelTimer(ds_ctx* c_, int id_, bool dispatch_ = false) : c(c_), id(id_), dispatch(dispatch_)
    {
        if (!c->kt_on) return;
        a = kt_event(c);
        b = kt_event(c);
        if (!a || !b) {
            if (a) c->kt_pool.push_back(a);
            if (b) c->kt_pool.push_back(b);
            a = b = nullptr;
            return;
        }
        if (!dispatch) (void)hipEventRecord(a, c->stream);
    }
    ~KernelTimer()
    {
        if (!a || !b) return;
        if (!dispatch) (void)hipEventRecord(b, c->stream);
        c->kt_pending.push_back({id, {a, b}});
    }
};

// One launch, timed by a dispatch-form KernelTimer when it samples (never during capture: the
// timer samples only outside it).
#define DS_LAUNCH_T(kt, kern, grid, block, shm, stream, ...)                                     \
    do {                                                                                         \
        if ((kt).a) hipExtLaunchKernelGGL(kern, grid, block, shm, stream, (kt).a, (kt).b, 0, __VA_ARGS__); \
        else hipLaunchKernelGGL(kern, grid, block, shm, stream, __VA_ARGS__);                    \
    } while (0)

// At the start of a sampled substep: hold the stream for 0.5 ms so that the host queues the
// substep's timed launches before the GPU reaches them.
void kt_hold(ds_ctx* c, hipStream_t s)
{
    if (c->kt_on) hipLaunchKernelGGL(k_hold, dim3(1), dim3(64), 0, s, 50000l);
}

// After a synchronisation: add the elapsed time of every recorded pair.
void kt_collect(ds_ctx* c)
{
    for (auto& pr : c->kt_pending) {
        float ms = 0.0f;
        if (hipEventElapsedTime(&ms, pr.second.first, pr.second.second) == hipSuccess) {
            c->kt_ms[pr.first] += ms;
            c->kt_n[pr.first] += 1;
        }
        c->kt_pool.push_back(pr.second.first);
        c->kt_pool.push_back(pr.second.second);
    }
    c->kt_pending.clear();
}

// Drop captured launch sequences (after an option or buffer change).
void drop_graphs(ds_ctx* c)
{
    if (c->fwd_exec) (void)hipGraphExecDestroy(c->fwd_exec);
    if (c->bwd_exec) (void)hipGraphExecDestroy(c->bwd_exec);
    c->fwd_exec = c->bwd_exec = nullptr;
}

void free_overlap(ds_ctx* c)
{
    dfree(c->d_seg_b);
    dfree(c->d_seg_grid_b);
    dfree(c->d_seg_aff_b);
    dfree(c->d_seg_mj_b);
    dfree(c->d_staging_r);
    dfree(c->d_fbflag_r);
    dfree(c->d_seg_tmp_b);
    c->overlap_alloc = false;
    dfree(c->d_xp);
    c->xgrad_alloc = false;
}

void free_mode(ds_ctx* c)
{
    drop_graphs(c);
    dfree(c->d_seg);
    dfree(c->d_grid);
    dfree(c->d_vout);
    dfree(c->d_vbar);
    dfree(c->d_pbar);
    dfree(c->d_vgrid);
    dfree(c->d_pbar2);
    dfree(c->d_pfb);
    dfree(c->d_vbfb);
    dfree(c->d_staging);
    c->staging_cap = 0;
    dfree(c->d_seg_grid);
    dfree(c->d_seg_aff);
    dfree(c->d_adj2);
    dfree(c->d_meta2);
    dfree(c->d_pid);
    dfree(c->d_pid2);
    dfree(c->d_cnt);
    dfree(c->d_celloff);
    dfree(c->d_bcount);
    dfree(c->d_bnch);
    dfree(c->d_bstart);
    dfree(c->d_cstart);
    dfree(c->d_tot);
    dfree(c->d_tsum);
    dfree(c->d_key);
    dfree(c->d_rank);
    dfree(c->d_rrpos);
    dfree(c->d_clist);
    dfree(c->d_chunks);
    dfree(c->d_rows);
    dfree(c->d_counts);
    dfree(c->d_gchunks);
    dfree(c->d_bgch);
    dfree(c->d_gstart);
    dfree(c->d_cov);
    dfree(c->d_fbflag);
    dfree(c->d_j);
    dfree(c->d_seg_mj);
    dfree(c->d_gadj);
    dfree(c->d_thpart);
    dfree(c->d_seg_tmp);
    dfree(c->d_brows);
    dfree(c->d_rstart);
    dfree(c->d_src);
    free_overlap(c);
    c->alloc_mode = -1;
}

void free_all(ds_ctx* c)
{
    for (auto& pr : c->kt_pending) {
        (void)hipEventDestroy(pr.second.first);
        (void)hipEventDestroy(pr.second.second);
    }
    c->kt_pending.clear();
    for (hipEvent_t e : c->kt_pool) (void)hipEventDestroy(e);
    c->kt_pool.clear();
    free_mode(c);
    dfree(c->d_meta);
    dfree(c->d_groups);
    dfree(c->d_x0);
    dfree(c->d_state);
    dfree(c->d_ckpt);
    dfree(c->d_adj);
    dfree(c->d_xbar);
    dfree(c->d_dvsum);
    dfree(c->d_dvbar);
    dfree(c->d_params);
    dfree(c->d_thbar);
    dfree(c->d_traj);
    dfree(c->d_flag);
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    for (hipEvent_t* e : {&c->ev_fork, &c->ev_rec[0], &c->ev_rec[1], &c->ev_adj[0], &c->ev_adj[1], &c->ev_step,
                          &c->ev_xg}) {
        if (*e) (void)hipEventDestroy(*e);
        *e = nullptr;
    }
    if (c->stream3) (void)hipStreamDestroy(c->stream3);
    if (c->stream2) (void)hipStreamDestroy(c->stream2);
    if (c->own_stream) (void)hipStreamDestroy(c->own_stream);
    c->ev0 = c->ev1 = nullptr;
    c->stream = c->own_stream = c->stream2 = c->stream3 = nullptr;
}

// Buffers that depend on the kernel family, allocated at the first ds_reset after a change.
int alloc_mode(ds_ctx* c)
{
    if (c->alloc_mode == c->mode) return DS_OK;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    free_mode(c);
    const long G = c->gridNodes() * c->B;
    const size_t SF = (size_t)c->stateFloats();
    int rc;
#define MALLOC(ptr, n)                                  \
    if ((rc = dalloc(c, &ptr, (size_t)(n))) != DS_OK) { \
        free_mode(c);                                   \
        return rc;                                      \
    }
    MALLOC(c->d_seg, SF * c->K);
    if (c->mode == DS_KERNELS_PER_PARTICLE) {
        MALLOC(c->d_grid, G);
        MALLOC(c->d_vout, G);
        MALLOC(c->d_vbar, G);
        MALLOC(c->d_pbar, G);
    } else {
        c->NB = (c->desc.n_grid + TB - 1) / TB;
        c->NR = c->desc.n_grid / TB + 1;
        c->NBK = c->B * 3 * c->NB * c->NB * c->NB;
        c->E = c->K / (c->sort_every > 0 ? c->sort_every : c->K);
        c->nslot = c->ncp * c->E;
        // chunks of a bucket: ceil(rows / (WAVES * PPT)) <= rows / WAVES + 1, and rows <= particles
        c->maxch = (long)c->P / WAVES + std::min<long>(c->P, c->NBK) + 1;
        c->maxgch = (long)c->P / WG + std::min<long>(c->P, c->NBK) + 1;
        // staging entries are addressed with 32-bit indices (slot * TNODES + node)
        if (std::max(c->maxch, c->maxgch) * (long)TNODES > 2147483647L)
            return fail(c, DS_EINVAL, "too many particles for the tiled kernels in one context");
        MALLOC(c->d_vgrid, G);
        MALLOC(c->d_pbar2, G);
        MALLOC(c->d_pfb, 4 * G);
        MALLOC(c->d_vbfb, 4 * G);
        MALLOC(c->d_seg_grid, (size_t)G * c->K);
        MALLOC(c->d_seg_aff, (size_t)9 * c->P * c->K);
        MALLOC(c->d_adj2, SF);
        MALLOC(c->d_meta2, c->P);
        MALLOC(c->d_pid, c->P);
        MALLOC(c->d_pid2, c->P);
        MALLOC(c->d_cnt, (size_t)c->NBK * 64);
        MALLOC(c->d_celloff, (size_t)c->NBK * 64);
        MALLOC(c->d_bcount, c->NBK);
        MALLOC(c->d_bnch, (size_t)c->NBK * c->nslot);
        MALLOC(c->d_bstart, c->NBK);
        MALLOC(c->d_cstart, (size_t)c->NBK * c->nslot);
        MALLOC(c->d_tot, 4);
        MALLOC(c->d_tsum, (size_t)4 * ((c->NBK + SCAN_TILE - 1) / SCAN_TILE));
        MALLOC(c->d_bgch, (size_t)c->NBK * c->nslot);
        MALLOC(c->d_gstart, (size_t)c->NBK * c->nslot);
        MALLOC(c->d_gchunks, (size_t)c->maxgch * c->nslot);
        MALLOC(c->d_cov, (size_t)c->B * c->NR * c->NR * c->NR * c->nslot);
        MALLOC(c->d_fbflag, (size_t)c->B * c->NR * c->NR * c->NR);
        MALLOC(c->d_j, c->P);
        MALLOC(c->d_seg_mj, (size_t)10 * c->P * c->K);
        MALLOC(c->d_gadj, (size_t)18 * c->P);
        MALLOC(c->d_thpart, (size_t)4 * c->maxgch);
        MALLOC(c->d_brows, c->NBK);
        MALLOC(c->d_rstart, c->NBK);
        MALLOC(c->d_rows, (size_t)(c->P + 1) * c->nslot);
        MALLOC(c->d_key, c->P);
        MALLOC(c->d_rank, c->P);
        MALLOC(c->d_rrpos, c->P);
        MALLOC(c->d_clist, c->P);
        MALLOC(c->d_chunks, (size_t)c->maxch * c->nslot);
        MALLOC(c->d_src, (size_t)c->P * c->nslot);
        MALLOC(c->d_counts, (size_t)4 * c->nslot);
        fill_i32(c, c->d_counts, 4L * c->nslot, 0);
        if (c->E > 1) MALLOC(c->d_seg_tmp, SF);
        // staging slots for either chunk set at its largest
        c->staging_cap = std::max(c->maxch, c->maxgch);
        MALLOC(c->d_staging, (size_t)c->staging_cap * TNODES);
        fill_i32(c, c->d_fbflag, (long)c->B * c->NR * c->NR * c->NR, 0);
        fill_zero_f64(c, c->d_pfb, 4 * G);
        fill_zero_f64(c, c->d_vbfb, 4 * G);
    }
#undef MALLOC
    c->alloc_mode = c->mode;
    return DS_OK;
}

// The position-gradient buffer, stream and events of the adjoint (tiled family), allocated at
// the first ds_backward that uses them.
int alloc_xgrad(ds_ctx* c)
{
    if (c->xgrad_alloc || c->mode != DS_KERNELS_TILED) return DS_OK;
    RCHK(dalloc(c, &c->d_xp, (size_t)3 * c->P));
    if (!c->stream3) HIPCHK(c, hipStreamCreate(&c->stream3));
    for (hipEvent_t* e : {&c->ev_step, &c->ev_xg})
        if (!*e) HIPCHK(c, hipEventCreateWithFlags(e, hipEventDisableTiming));
    c->xgrad_alloc = true;
    return DS_OK;
}

// The overlapped adjoint's second buffer set, staging, flags, stream and events (tiled family),
// allocated at the first ds_backward that uses them.
int alloc_overlap(ds_ctx* c)
{
    if (c->overlap_alloc || c->mode != DS_KERNELS_TILED) return DS_OK;
    const long G = c->gridNodes() * c->B;
    const size_t SF = (size_t)c->stateFloats();
    int rc;
#define MALLOC(ptr, n)                                  \
    if ((rc = dalloc(c, &ptr, (size_t)(n))) != DS_OK) { \
        free_overlap(c);                                \
        return rc;                                      \
    }
    MALLOC(c->d_seg_b, SF * c->K);
    MALLOC(c->d_seg_grid_b, (size_t)G * c->K);
    MALLOC(c->d_seg_aff_b, (size_t)9 * c->P * c->K);
    MALLOC(c->d_seg_mj_b, (size_t)10 * c->P * c->K);
    MALLOC(c->d_staging_r, (size_t)c->staging_cap * TNODES);
    MALLOC(c->d_fbflag_r, (size_t)c->B * c->NR * c->NR * c->NR);
    if (c->E > 1) MALLOC(c->d_seg_tmp_b, SF);
#undef MALLOC
    fill_i32(c, c->d_fbflag_r, (long)c->B * c->NR * c->NR * c->NR, 0);
    if (!c->stream2) HIPCHK(c, hipStreamCreate(&c->stream2));
    for (hipEvent_t* e : {&c->ev_fork, &c->ev_rec[0], &c->ev_rec[1], &c->ev_adj[0], &c->ev_adj[1]})
        if (!*e) HIPCHK(c, hipEventCreateWithFlags(e, hipEventDisableTiming));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->overlap_alloc = true;
    return DS_OK;
}

// ---------------------------------------------------------------------------- per-particle path

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

int forward_ref(ds_ctx* c, int nsteps)
{
    const long SF = c->stateFloats();
    for (int i = 0; i < nsteps; i++) {
        const int s = c->step;
        if (s % c->K == 0)
            HIPCHK(c, hipMemcpyAsync(c->d_ckpt + (size_t)(s / c->K) * SF, c->d_state, sizeof(float) * SF,
                                     hipMemcpyDeviceToDevice, c->stream));
        step_ref(c, c->d_state, c->d_state, s, true);
        c->step++;
    }
    return DS_OK;
}

int backward_ref(ds_ctx* c)
{
    const int S = c->step;
    const long SF = c->stateFloats();
    for (int cp = (S - 1) / c->K; cp >= 0; cp--) {
        const int s0 = cp * c->K;
        const int s1 = std::min(s0 + c->K, S);
        HIPCHK(c, hipMemcpyAsync(c->d_seg, c->d_ckpt + (size_t)cp * SF, sizeof(float) * SF, hipMemcpyDeviceToDevice,
                                 c->stream));
        for (int s = s0; s < s1 - 1; s++) {
            const int k = s - s0;
            step_ref(c, c->d_seg + (size_t)k * SF, c->d_seg + (size_t)(k + 1) * SF, s, false);
        }
        for (int s = s1 - 1; s >= s0; s--) step_adj_ref(c, c->d_seg + (size_t)(s - s0) * SF, s);
    }
    return DS_OK;
}

// ---------------------------------------------------------------------------- tiled path

// Sort the current state into checkpoint slot cp (bucket order, round-robin within buckets),
// build that checkpoint's chunk table and record where every slot came from.
// Bucket sort of the current state (d_state) into dst, with the chunk tables, row table,
// reached-block flags and new-to-old slot map in table slot `slot`.
int sort_into(ds_ctx* c, int slot, float* dst)
{
    const int cp = slot;   // table slot (checkpoint x sorts per segment)
    Chunk* chunks = c->d_chunks + (size_t)cp * c->maxch;
    int* src = c->d_src + (size_t)cp * c->P;
    int* bnch = c->d_bnch + (size_t)cp * c->NBK;
    int* cstart = c->d_cstart + (size_t)cp * c->NBK;
    int* bgch = c->d_bgch + (size_t)cp * c->NBK;
    int* gstart = c->d_gstart + (size_t)cp * c->NBK;
    fill_i32(c, c->d_cnt, (long)c->NBK * 64, 0);
    hipLaunchKernelGGL(k_sort_count, dim3(nblocks(c->P)), dim3(256), 0, c->stream, c->d_state, c->d_meta,
                       c->d_groups, c->d_key, c->d_rank, c->d_cnt, c->dsc, c->NB, c->P);
    const int rpc = WAVES * c->ppt;   // rows per chunk
    const int gch = WG * c->gppt;     // particles per gather chunk
    c->ppt_sorted = c->ppt;
    c->gppt_sorted = c->gppt;
    int* rows = c->d_rows + (size_t)cp * (c->P + 1);
    GChunk* gchunks = c->d_gchunks + (size_t)cp * c->maxgch;
    hipLaunchKernelGGL(k_bucket_totals, dim3(nblocks((long)c->NBK * 64)), dim3(256), 0, c->stream, c->d_cnt,
                       c->d_celloff, c->d_bcount, c->d_brows, bnch, bgch, c->NBK, rpc, gch);
    Scan4 q{{c->d_bcount, bnch, c->d_brows, bgch}, {c->d_bstart, cstart, c->d_rstart, gstart}};
    int* tot = c->d_counts + 4 * cp;
    const int ntiles = (c->NBK + SCAN_TILE - 1) / SCAN_TILE;
    hipLaunchKernelGGL(k_scan4_sums, dim3(ntiles), dim3(SCAN_TILE), 0, c->stream, q, c->NBK, c->d_tsum);
    hipLaunchKernelGGL(k_scan4_tiles, dim3(1), dim3(SCAN_TILE), 0, c->stream, c->d_tsum, ntiles, tot);
    hipLaunchKernelGGL(k_scan4_apply, dim3(ntiles), dim3(SCAN_TILE), 0, c->stream, q, c->NBK, c->d_tsum);
    // deterministic ranks within each cell (by original particle index)
    hipLaunchKernelGGL(k_cell_list, dim3(nblocks(c->P)), dim3(256), 0, c->stream, c->d_key, c->d_rank, c->d_celloff,
                       c->d_bstart, c->d_clist, c->P);
    hipLaunchKernelGGL(k_rank_fix, dim3(nblocks(c->P)), dim3(256), 0, c->stream, c->d_key, c->d_cnt, c->d_celloff,
                       c->d_bstart, c->d_clist, c->d_pid, c->d_rank, c->P);
    const long nbl = (long)c->B * c->NR * c->NR * c->NR;
    int* cov = c->d_cov + (size_t)cp * nbl;
    fill_i32(c, cov, nbl, 0);
    hipLaunchKernelGGL(k_bucket_layout, dim3(nblocks((long)c->NBK * 64)), dim3(256), 0, c->stream, c->d_cnt,
                       c->d_celloff, c->d_bstart, c->d_rstart, cstart, gstart, c->d_rrpos, rows, chunks, gchunks,
                       cov, c->NB, c->NR, c->NBK, rpc, gch);
    fill_i32(c, c->d_meta2, c->P, -1);
    hipLaunchKernelGGL(k_sort_scatter, dim3(nblocks(c->P)), dim3(256), 0, c->stream, c->d_state, dst, c->d_key,
                       c->d_rank, c->d_celloff, c->d_bstart, c->d_rrpos, c->d_meta, c->d_meta2, c->d_pid, c->d_pid2,
                       src, c->P);
    // the sorted particle tables back into place (fixed buffers: a captured launch sequence
    // replays with the same pointers)
    hipLaunchKernelGGL(k_copy_i32, dim3(std::min<long>(nblocks(c->P), 4096)), dim3(256), 0, c->stream, c->d_meta,
                       c->d_meta2, (long)c->P);
    hipLaunchKernelGGL(k_copy_i32, dim3(std::min<long>(nblocks(c->P), 4096)), dim3(256), 0, c->stream, c->d_pid,
                       c->d_pid2, (long)c->P);
    hipLaunchKernelGGL(k_check_sort, dim3(1), dim3(64), 0, c->stream, tot, c->nreal, c->maxch, c->maxgch, c->d_flag);
    return DS_OK;
}

// Workgroups of a chunk kernel launch: the chunk count lives on the device, so the grid is fixed
// (workgroups beyond the count exit at once; a count above the grid is covered by the loop).
// (A larger grid lets the hardware balance chunks of uneven cost over the chip; workgroups past
// the count exit after one load.  Measured on the 32-instance workload: 4096 -> 16384 workgroups
// took 4-8 % off the scatter and gather kernels; one instance is unaffected.)
constexpr long CHUNK_GRID = 16384;
inline dim3 chunk_grid(const ds_ctx* c, long maxchunks)
{
    return dim3((unsigned)std::max<long>(1, std::min<long>(maxchunks, c->chunk_grid)));
}

#define DS_PPT_DISPATCH(ppt, kern, ...)                                                          \
    do {                                                                                         \
        if ((ppt) == 1) hipLaunchKernelGGL(kern<1>, __VA_ARGS__);                                \
        else if ((ppt) == 2) hipLaunchKernelGGL(kern<2>, __VA_ARGS__);                           \
        else hipLaunchKernelGGL(kern<4>, __VA_ARGS__);                                           \
    } while (0)
#define DS_PPT_DISPATCH2(ppt, kern, b, ...)                                                      \
    do {                                                                                         \
        if ((ppt) == 1) hipLaunchKernelGGL((kern<1, b>), __VA_ARGS__);                           \
        else if ((ppt) == 2) hipLaunchKernelGGL((kern<2, b>), __VA_ARGS__);                      \
        else hipLaunchKernelGGL((kern<4, b>), __VA_ARGS__);                                      \
    } while (0)
#define DS_PPT_DISPATCH_T(kt, ppt, kern, ...)                                                    \
    do {                                                                                         \
        if ((ppt) == 1) DS_LAUNCH_T(kt, kern<1>, __VA_ARGS__);                                   \
        else if ((ppt) == 2) DS_LAUNCH_T(kt, kern<2>, __VA_ARGS__);                              \
        else DS_LAUNCH_T(kt, kern<4>, __VA_ARGS__);                                              \
    } while (0)
#define DS_PPT_DISPATCH3_T(kt, ppt, kern, b, b2, ...)                                            \
    do {                                                                                         \
        if ((ppt) == 1) DS_LAUNCH_T(kt, (kern<1, b, b2>), __VA_ARGS__);                          \
        else if ((ppt) == 2) DS_LAUNCH_T(kt, (kern<2, b, b2>), __VA_ARGS__);                     \
        else DS_LAUNCH_T(kt, (kern<4, b, b2>), __VA_ARGS__);                                     \
    } while (0)
#define DS_PPT_DISPATCH2_T(kt, ppt, kern, b, ...)                                                \
    do {                                                                                         \
        if ((ppt) == 1) DS_LAUNCH_T(kt, (kern<1, b>), __VA_ARGS__);                              \
        else if ((ppt) == 2) DS_LAUNCH_T(kt, (kern<2, b>), __VA_ARGS__);                         \
        else DS_LAUNCH_T(kt, (kern<4, b>), __VA_ARGS__);                                         \
    } while (0)
// A grid reduction with `parts` wavefronts per node block (64 * parts threads).
#define DS_PARTS_DISPATCH_T(kt, parts, kern, grid, stream, ...)                                  \
    do {                                                                                         \
        if ((parts) == 1) DS_LAUNCH_T(kt, kern<1>, grid, dim3(64), 0, stream, __VA_ARGS__);      \
        else if ((parts) == 2) DS_LAUNCH_T(kt, kern<2>, grid, dim3(128), 0, stream, __VA_ARGS__); \
        else DS_LAUNCH_T(kt, kern<4>, grid, dim3(256), 0, stream, __VA_ARGS__);                  \
    } while (0)

// Where the forward-type kernels (P2G, grid reduction, G2P, J) run: the stream and the scratch
// they write besides their outputs.  The forward and a sequential recompute use the main stream's;
// a recompute overlapped with the adjoint uses the second stream and its own staging and flags.
struct Exec {
    hipStream_t stream;
    float4* staging;
    int* fbflag;
};
inline Exec main_exec(const ds_ctx* c) { return Exec{c->stream, c->d_staging, c->d_fbflag}; }

// P2G of a checkpoint segment's chunks into the staging slots.
void p2g_tiled(ds_ctx* c, const Exec& x, const float* in, float* fout, int cp, float* aout)
{
    const int* counts = c->d_counts + 4 * cp;
    {
        KernelTimer kt(c, DS_K_P2G, true);
        if (c->packed)
            DS_PPT_DISPATCH2_T(kt, c->gppt_sorted, k_p2g_tiled, GChunk, chunk_grid(c, c->maxgch), dim3(WG), 0, x.stream,
                               in, fout, aout, c->d_gchunks + (size_t)cp * c->maxgch, nullptr, counts + 3,
                               (int)c->maxgch, c->d_params, x.staging, c->d_pfb, x.fbflag, c->NR, c->d_j, c->dsc,
                               c->P, c->d_flag);
        else
            DS_PPT_DISPATCH2_T(kt, c->ppt_sorted, k_p2g_tiled, Chunk, chunk_grid(c, c->maxch), dim3(WG), 0, x.stream, in,
                               fout, aout, c->d_chunks + (size_t)cp * c->maxch, c->d_rows + (size_t)cp * (c->P + 1),
                               counts + 1, (int)c->maxch, c->d_params, x.staging, c->d_pfb, x.fbflag, c->NR,
                               c->d_j, c->dsc, c->P, c->d_flag);
    }
}

// The grid reduction of the staging slots that P2G (or the fused G2P + P2G) wrote for table slot
// cp: velocity into d_vgrid and, when pout is given, momentum/mass into pout.
void grid_reduce_tiled(ds_ctx* c, const Exec& x, int cp, float4* pout)
{
    // the scatter's chunk set: rows of distinct stencil bases, or packed gather chunks
    const int* bnch = (c->packed ? c->d_bgch : c->d_bnch) + (size_t)cp * c->NBK;
    const int* cstart = (c->packed ? c->d_gstart : c->d_cstart) + (size_t)cp * c->NBK;
    const int nr3 = c->NR * c->NR * c->NR;
    KernelTimer kt(c, DS_K_GRID_REDUCE, true);
    const int* cov = c->d_cov + (size_t)cp * c->B * nr3;
    DS_PARTS_DISPATCH_T(kt, c->rparts, k_grid_reduce, dim3(c->B * nr3), x.stream, x.staging, bnch, cstart, cov,
                        c->d_pfb, x.fbflag, pout, c->d_vgrid, c->dsc, c->NB, c->NR);
}

void g2p_tiled(ds_ctx* c, const Exec& x, const float* in, float* out, int cp, int f, bool record,
               double* mj = nullptr)
{
    KernelTimer kt(c, mj ? DS_K_G2P_FACTOR : DS_K_G2P, true);
    const GChunk* gch = c->d_gchunks + (size_t)cp * c->maxgch;
    const int* ngc = c->d_counts + 4 * cp + 3;
    double* dv = record ? c->d_dvsum : nullptr;
    // the forward's chunks of gppt particles per thread as `split` workgroups of gppt / split
    // each (the recompute's G2P keeps whole chunks: there a faster G2P only competes harder with
    // the adjoint on the other stream)
    const int split = mj ? 1 : std::max(1, std::min(c->g2p_split, c->gppt_sorted));
    const int ppt = c->gppt_sorted / split;
    const dim3 grid = chunk_grid(c, c->maxgch * split);
    if (mj)
        DS_PPT_DISPATCH2_T(kt, ppt, k_g2p_tiled, true, grid, dim3(WG), 0, x.stream, in, out, gch, ngc, (int)c->maxgch,
                           c->d_vgrid, c->d_traj, c->T, f, dv, c->d_j, mj, c->dsc, c->P, c->d_flag, split);
    else
        DS_PPT_DISPATCH2_T(kt, ppt, k_g2p_tiled, false, grid, dim3(WG), 0, x.stream, in, out, gch, ngc, (int)c->maxgch,
                           c->d_vgrid, c->d_traj, c->T, f, dv, c->d_j, mj, c->dsc, c->P, c->d_flag, split);
}

// The forward's G2P of step f fused with the P2G of step f + 1 (row chunks of table slot cp, which
// both steps use): x, v, C of step f + 1 into out, F of step f + 2 into fout, the stick force of
// step f + 1, the staging slots of step f + 1.
void g2p_p2g_tiled(ds_ctx* c, const Exec& x, const float* in, const float* fin, float* out, float* fout, int cp,
                   int f)
{
    KernelTimer kt(c, DS_K_G2P_P2G, true);
    const int* counts = c->d_counts + 4 * cp;
    DS_PPT_DISPATCH_T(kt, c->ppt_sorted, k_g2p_p2g_tiled, chunk_grid(c, c->maxch), dim3(WG), 0, x.stream, in, fin, out,
                      fout, c->d_chunks + (size_t)cp * c->maxch, c->d_rows + (size_t)cp * (c->P + 1), counts + 1,
                      (int)c->maxch, c->d_vgrid, c->d_traj, c->T, f, c->d_dvsum, c->d_params, x.staging, c->d_pfb,
                      x.fbflag, c->NR, c->dsc, c->P, c->d_flag);
}

// The fused G2P + P2G applies: the option is on and the scatter runs on row chunks.
inline bool fused_step(const ds_ctx* c) { return c->fuse_step && !c->packed; }

// J of every liquid particle of the checkpoint state `st` into d_j, for the segment's first P2G
// (and, given mj, J and its SVD backward factor into mj).
void liquid_j_tiled(ds_ctx* c, const Exec& x, const float* st, int cp, double* mj = nullptr)
{
    const GChunk* gch = c->d_gchunks + (size_t)cp * c->maxgch;
    const int* ngc = c->d_counts + 4 * cp + 3;
    if (mj)
        DS_PPT_DISPATCH2(c->gppt_sorted, k_liquid_j, true, chunk_grid(c, c->maxgch), dim3(WG), 0, x.stream, st, gch,
                         ngc, (int)c->maxgch, c->d_j, mj, c->dsc, c->P);
    else
        DS_PPT_DISPATCH2(c->gppt_sorted, k_liquid_j, false, chunk_grid(c, c->maxgch), dim3(WG), 0, x.stream, st, gch,
                         ngc, (int)c->maxgch, c->d_j, mj, c->dsc, c->P);
}

// Substeps between sorts and the table slot of step s: the sorts of checkpoint segment cp sit
// at its start and every R substeps after it, in slots cp * E .. cp * E + E - 1.
inline int sort_interval(const ds_ctx* c) { return c->K / c->E; }
inline int table_slot(const ds_ctx* c, int s) { return (s / c->K) * c->E + (s % c->K) / sort_interval(c); }

int forward_tiled(ds_ctx* c, int nsteps)
{
    const long SF = c->stateFloats();
    const int R = sort_interval(c);
    // the P2G of the current step was done by the previous step's fused G2P + P2G
    bool p2g_done = false;
    for (int i = 0; i < nsteps; i++) {
        const int s = c->step;
        const int cp = s / c->K;
        const int slot = table_slot(c, s);
        const float* in = c->d_state;
        c->kt_on = !c->capturing && c->kt_every > 0 && s % c->kt_every == 0;
        kt_hold(c, c->stream);
        if (s % R == 0) {
            // a checkpoint sort keeps its sorted state; a sort inside a segment goes through the
            // segment buffer, which nothing else uses during a forward rollout
            float* dst = s % c->K == 0 ? c->d_ckpt + (size_t)cp * SF : c->d_seg;
            {
                KernelTimer kt(c, DS_K_SORT);
                RCHK(sort_into(c, slot, dst));
            }
            in = dst;
            liquid_j_tiled(c, main_exec(c), in, slot);
        }
        if (!p2g_done) p2g_tiled(c, main_exec(c), in, c->d_state, slot, nullptr);
        grid_reduce_tiled(c, main_exec(c), slot, nullptr);
        // the next step's P2G rides on this step's G2P when it is part of this call and runs under
        // the same sort (a call ends on a complete state: the last G2P stands alone)
        p2g_done = fused_step(c) && i + 1 < nsteps && (s + 1) % R != 0;
        if (p2g_done)
            g2p_p2g_tiled(c, main_exec(c), in, c->d_state, c->d_state, c->d_state, slot, s);
        else
            g2p_tiled(c, main_exec(c), in, c->d_state, slot, s, true);
        c->step++;
    }
    return DS_OK;
}

// One buffer set of a checkpoint segment: each step's state, momentum/mass grid, liquid affine,
// and J with its SVD backward factor.
struct SegSet {
    float* seg;
    float4* grid;
    float* aff;
    double* mj;
    float* tmp;   // a step's state before a mid-segment sort's permutation (sorts per segment > 1)
};
inline SegSet seg_set(const ds_ctx* c, int i)
{
    return i == 0 ? SegSet{c->d_seg, c->d_seg_grid, c->d_seg_aff, c->d_seg_mj, c->d_seg_tmp}
                  : SegSet{c->d_seg_b, c->d_seg_grid_b, c->d_seg_aff_b, c->d_seg_mj_b, c->d_seg_tmp_b};
}

// Substeps [s0, s1) of checkpoint segment cp (s1 = min(s0 + K, recorded steps)).
inline void segment_range(const ds_ctx* c, int cp, int S, int& s0, int& L)
{
    s0 = cp * c->K;
    L = std::min(s0 + c->K, S) - s0;
}

// Recompute checkpoint segment cp from its checkpoint into buffer set z, keeping each step's
// state, momentum/mass grid, affine and J factor.
void recompute_segment(ds_ctx* c, const Exec& x, int cp, int S, const SegSet& z)
{
    const long SF = c->stateFloats();
    const long G = c->gridNodes() * c->B;
    int s0, L;
    segment_range(c, cp, S, s0, L);
    const float* ck = c->d_ckpt + (size_t)cp * SF;
    const int R = sort_interval(c);
    auto st_of = [&](int k) -> float* { return k == 0 ? (float*)ck : z.seg + (size_t)k * SF; };
    auto mj_of = [&](int k) -> double* { return z.mj + (size_t)k * 10 * c->P; };
    liquid_j_tiled(c, x, ck, table_slot(c, s0), z.mj);
    for (int k = 0; k < L; k++) {
        const int s = s0 + k;
        const int slot = table_slot(c, s);
        c->kt_on = !c->capturing && c->kt_every > 0 && s % c->kt_every == 0;
        kt_hold(c, x.stream);
        // the forward sorted again before step k + 1: this step writes the old order into tmp,
        // which the forward's permutation of that sort takes to the step's buffer
        const bool resort = k + 1 < L && (k + 1) % R == 0;
        float* next = k + 1 < L ? (resort ? z.tmp : st_of(k + 1)) : nullptr;
        // (the recompute keeps G2P and P2G apart: beside the adjoint on the other stream, G2P on
        // the scatter's sparse rows costs the adjoint more than the saved launch gains)
        p2g_tiled(c, x, st_of(k), next, slot, z.aff + (size_t)k * 9 * c->P);
        grid_reduce_tiled(c, x, slot, z.grid + (size_t)k * G);
        if (k + 1 >= L) continue;
        if (!resort) {
            g2p_tiled(c, x, st_of(k), next, slot, s, false, mj_of(k + 1));
            continue;
        }
        g2p_tiled(c, x, st_of(k), next, slot, s, false);
        const int slot2 = table_slot(c, s + 1);
        hipLaunchKernelGGL(k_permute_state, dim3(nblocks(c->nreal)), dim3(256), 0, x.stream, z.tmp, st_of(k + 1),
                           c->d_src + (size_t)slot2 * c->P, c->nreal, c->P);
        liquid_j_tiled(c, x, st_of(k + 1), slot2, mj_of(k + 1));
    }
}

// Adjoint of checkpoint segment cp from its recomputed buffer set z, last step first, on the
// main stream; then the segment's parameter-gradient partials into d_thbar and the particle
// adjoint back into the previous checkpoint's order.
int adjoint_segment(ds_ctx* c, int cp, int S, const SegSet& z)
{
    const long SF = c->stateFloats();
    const long G = c->gridNodes() * c->B;
    const int nr3 = c->NR * c->NR * c->NR;
    const dim3 rgrid = chunk_grid(c, c->maxch), ggrid = chunk_grid(c, c->maxgch);
    const int maxch = (int)c->maxch, maxgch = (int)c->maxgch;
    const int R = sort_interval(c);
    int s0, L;
    segment_range(c, cp, S, s0, L);
    const float* ck = c->d_ckpt + (size_t)cp * SF;
    auto st_of = [&](int k) -> float* { return k == 0 ? (float*)ck : z.seg + (size_t)k * SF; };
    auto grid_of = [&](int k) -> float4* { return z.grid + (size_t)k * G; };
    auto aff_of = [&](int k) -> float* { return z.aff + (size_t)k * 9 * c->P; };
    auto mj_of = [&](int k) -> double* { return z.mj + (size_t)k * 10 * c->P; };
    fill_zero_f64(c, c->d_thpart, 4L * c->maxgch);
    // the position gradient of step k runs on stream3 beside the adjoint G2P and grid reduction
    // of step k: it waits for everything the main stream did before (ev_step: the previous step's
    // adjoint P2G wrote the particle adjoint it reads and read d_xp), and the adjoint P2G of step
    // k waits for it (ev_xg)
    const bool xs = c->xgrad_stream && c->xgrad_alloc;
    if (xs) HIPCHK(c, hipEventRecord(c->ev_step, c->stream));
    for (int k = L - 1; k >= 0; k--) {
        const int s = s0 + k;
        // the tables of the sort this step ran under
        const int slot = table_slot(c, s);
        const int* counts = c->d_counts + 4 * slot;   // particles, row chunks, rows, gather chunks
        const Chunk* chunks = c->d_chunks + (size_t)slot * c->maxch;
        const int* rows = c->d_rows + (size_t)slot * (c->P + 1);
        const GChunk* gchunks = c->d_gchunks + (size_t)slot * c->maxgch;
        const int* bnch = (c->packed ? c->d_bgch : c->d_bnch) + (size_t)slot * c->NBK;
        const int* cstart = (c->packed ? c->d_gstart : c->d_cstart) + (size_t)slot * c->NBK;
        c->kt_on = !c->capturing && c->kt_every > 0 && s % c->kt_every == 0;
        kt_hold(c, c->stream);
        if (xs) {
            HIPCHK(c, hipStreamWaitEvent(c->stream3, c->ev_step, 0));
            KernelTimer kt(c, DS_K_P2G_ADJ_XGRAD, true);
            DS_PPT_DISPATCH_T(kt, c->gppt_sorted, k_p2g_adj_xgrad, ggrid, dim3(WG), 0, c->stream3, st_of(k), c->d_adj,
                              gchunks, counts + 3, maxgch, grid_of(k), c->d_xp, c->dsc, c->P);
            HIPCHK(c, hipEventRecord(c->ev_xg, c->stream3));
        }
        {
            KernelTimer kt(c, DS_K_G2P_ADJ, true);
            if (c->packed)
                DS_PPT_DISPATCH2_T(kt, c->gppt_sorted, k_g2p_adj_tiled, GChunk, ggrid, dim3(WG), 0, c->stream,
                                   st_of(k), c->d_adj, gchunks, nullptr, counts + 3, maxgch, c->d_dvbar, c->T, s,
                                   c->d_staging, c->d_vbfb, c->d_fbflag, c->NR, c->dsc, c->P);
            else
                DS_PPT_DISPATCH2_T(kt, c->ppt_sorted, k_g2p_adj_tiled, Chunk, rgrid, dim3(WG), 0, c->stream, st_of(k),
                                   c->d_adj, chunks, rows, counts + 1, maxch, c->d_dvbar, c->T, s, c->d_staging,
                                   c->d_vbfb, c->d_fbflag, c->NR, c->dsc, c->P);
        }
        {
            KernelTimer kt(c, DS_K_VBAR_REDUCE, true);
            DS_PARTS_DISPATCH_T(kt, c->rparts, k_vbar_reduce, dim3(c->B * nr3), c->stream, c->d_staging, bnch,
                                cstart, c->d_cov + (size_t)slot * c->B * nr3, c->d_vbfb, c->d_fbflag, grid_of(k),
                                c->d_pbar2, c->dsc, c->NB, c->NR);
        }
        if (xs) HIPCHK(c, hipStreamWaitEvent(c->stream, c->ev_xg, 0));
        if (c->adj_split) {
            {
                KernelTimer kt(c, DS_K_P2G_ADJ, true);
                if (xs)
                    DS_PPT_DISPATCH3_T(kt, c->gppt_sorted, k_p2g_adj_tiled, false, true, ggrid, dim3(WG), 0, c->stream,
                                       st_of(k), aff_of(k), c->d_adj, gchunks, counts + 3, maxgch, grid_of(k),
                                       c->d_pbar2, c->d_gadj, c->d_params, mj_of(k), c->d_thpart, c->d_xp, c->dsc,
                                       c->P);
                else
                    DS_PPT_DISPATCH3_T(kt, c->gppt_sorted, k_p2g_adj_tiled, false, false, ggrid, dim3(WG), 0, c->stream,
                                       st_of(k), aff_of(k), c->d_adj, gchunks, counts + 3, maxgch, grid_of(k),
                                       c->d_pbar2, c->d_gadj, c->d_params, mj_of(k), c->d_thpart, c->d_xp, c->dsc,
                                       c->P);
            }
            KernelTimer kt(c, DS_K_P2G_ADJ_STRESS, true);
            DS_PPT_DISPATCH_T(kt, c->gppt_sorted, k_p2g_adj_stress, ggrid, dim3(WG), 0, c->stream, st_of(k), c->d_adj,
                              gchunks, counts + 3, maxgch, c->d_params, c->d_gadj, mj_of(k), c->d_thpart, c->dsc,
                              c->P);
        } else {
            KernelTimer kt(c, DS_K_P2G_ADJ, true);
            if (xs)
                DS_PPT_DISPATCH3_T(kt, c->gppt_sorted, k_p2g_adj_tiled, true, true, ggrid, dim3(WG), 0, c->stream,
                                   st_of(k), aff_of(k), c->d_adj, gchunks, counts + 3, maxgch, grid_of(k), c->d_pbar2,
                                   c->d_gadj, c->d_params, mj_of(k), c->d_thpart, c->d_xp, c->dsc, c->P);
            else
                DS_PPT_DISPATCH3_T(kt, c->gppt_sorted, k_p2g_adj_tiled, true, false, ggrid, dim3(WG), 0, c->stream,
                                   st_of(k), aff_of(k), c->d_adj, gchunks, counts + 3, maxgch, grid_of(k), c->d_pbar2,
                                   c->d_gadj, c->d_params, mj_of(k), c->d_thpart, c->d_xp, c->dsc, c->P);
        }
        if (k % R != 0) {
            if (xs) HIPCHK(c, hipEventRecord(c->ev_step, c->stream));
            continue;
        }
        // first step after a sort: the sort's parameter-gradient partials, and the adjoint back
        // into the particle order before that sort (none before the very first one)
        hipLaunchKernelGGL(k_thpart_reduce, dim3(c->B), dim3(256), 0, c->stream, gchunks, counts + 3, maxgch,
                           c->d_thpart, c->d_thbar);
        if (k > 0) fill_zero_f64(c, c->d_thpart, 4L * c->maxgch);
        if (k > 0 || cp > 0) {
            hipLaunchKernelGGL(k_unsort_adj, dim3(nblocks(c->nreal)), dim3(256), 0, c->stream, c->d_adj, c->d_adj2,
                               c->d_src + (size_t)slot * c->P, c->nreal, c->P);
            hipLaunchKernelGGL(k_copy_f32, dim3((unsigned)std::min<long>(nblocks(SF), 8192)), dim3(256), 0,
                               c->stream, c->d_adj, c->d_adj2, SF);
        }
        if (xs) HIPCHK(c, hipEventRecord(c->ev_step, c->stream));
    }
    return DS_OK;
}

// The checkpointed adjoint of the recorded rollout.  Sequential: recompute segment k, then its
// adjoint, for k from the last segment down.  Overlapped: segment k-1 is recomputed on the
// second stream (into the other buffer set, with its own staging and flags) while segment k's
// adjoint runs on the main stream; events order each recompute after the adjoint that last read
// its buffer set, and each adjoint after its recompute.  Both orders run the same kernels on
// the same data, so the results are identical.
int backward_tiled(ds_ctx* c)
{
    const int S = c->step;
    const int last = (S - 1) / c->K;
    if (!c->overlap || !c->overlap_alloc) {
        for (int cp = last; cp >= 0; cp--) {
            recompute_segment(c, main_exec(c), cp, S, seg_set(c, 0));
            RCHK(adjoint_segment(c, cp, S, seg_set(c, 0)));
        }
        return DS_OK;
    }
    const Exec xr{c->stream2, c->d_staging_r, c->d_fbflag_r};
    HIPCHK(c, hipEventRecord(c->ev_fork, c->stream));
    HIPCHK(c, hipStreamWaitEvent(c->stream2, c->ev_fork, 0));
    recompute_segment(c, xr, last, S, seg_set(c, last & 1));
    HIPCHK(c, hipEventRecord(c->ev_rec[last & 1], c->stream2));
    for (int cp = last; cp >= 0; cp--) {
        if (cp > 0) {
            // buffer set (cp - 1) & 1 was last read by the adjoint of segment cp + 1
            if (cp + 1 <= last) HIPCHK(c, hipStreamWaitEvent(c->stream2, c->ev_adj[(cp + 1) & 1], 0));
            recompute_segment(c, xr, cp - 1, S, seg_set(c, (cp - 1) & 1));
            HIPCHK(c, hipEventRecord(c->ev_rec[(cp - 1) & 1], c->stream2));
        }
        HIPCHK(c, hipStreamWaitEvent(c->stream, c->ev_rec[cp & 1], 0));
        RCHK(adjoint_segment(c, cp, S, seg_set(c, cp & 1)));
        HIPCHK(c, hipEventRecord(c->ev_adj[cp & 1], c->stream));
    }
    return DS_OK;
}

int backward_tiled_n(ds_ctx* c, int) { return backward_tiled(c); }

// Copy one species of one instance out of the current state, in the order it was given.
int read_species(ds_ctx* c, int inst, int sp, int slot0, int ncomp, float* out)
{
    if (!out) return DS_OK;
    const Group& g = c->groups[inst * 3 + sp];
    if (g.count == 0) return DS_OK;
    if (c->alloc_mode != DS_KERNELS_TILED) {
        std::vector<float> tmp((size_t)ncomp * g.count);
        for (int a = 0; a < ncomp; a++)
            HIPCHK(c, hipMemcpyAsync(tmp.data() + (size_t)a * g.count, c->d_state + (size_t)(slot0 + a) * c->P + g.offset,
                                     sizeof(float) * g.count, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        for (int i = 0; i < g.count; i++)
            for (int a = 0; a < ncomp; a++) out[(size_t)i * ncomp + a] = tmp[(size_t)a * g.count + i];
        return DS_OK;
    }
    std::vector<int> meta(c->P), pid(c->P);
    std::vector<float> tmp((size_t)ncomp * c->P);
    HIPCHK(c, hipMemcpyAsync(meta.data(), c->d_meta, sizeof(int) * c->P, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(pid.data(), c->d_pid, sizeof(int) * c->P, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(tmp.data(), c->d_state + (size_t)slot0 * c->P, sizeof(float) * ncomp * c->P,
                             hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    const int gid = inst * 3 + sp;
    for (int p = 0; p < c->P; p++) {
        if (meta[p] != gid) continue;
        const int i = pid[p];
        if (i < 0 || i >= g.count) return fail(c, DS_EHIP, "particle index table is corrupt");
        for (int a = 0; a < ncomp; a++) out[(size_t)i * ncomp + a] = tmp[(size_t)a * c->P + p];
    }
    return DS_OK;
}

// ---------------------------------------------------------------------------- no-allocation substep

// Scratch of ds_substep_noalloc: one instance's dense grid (momentum/mass, then velocity), the
// slot-to-group table, the group table, the stick's dv sums and the failure flag.  One copy per
// GPU, held until the process ends; every call on a GPU shares it.
constexpr int NA_MAX_GRID = 64;                                   // nodes per axis
constexpr long NA_MAX_NODES = (long)NA_MAX_GRID * NA_MAX_GRID * NA_MAX_GRID;
constexpr long NA_MAX_SLOTS = 1l << 18;                           // particle slots (runs padded to 64)
constexpr size_t NA_GRID_BYTES = sizeof(float4) * NA_MAX_NODES;  // 4 MiB each
constexpr size_t NA_META_BYTES = sizeof(int) * NA_MAX_SLOTS;     // 1 MiB
constexpr size_t NA_TAB_BYTES = 256;                              // groups 48 B, dv sums at 64, flag at 128
constexpr size_t NA_SCRATCH_BYTES = 2 * NA_GRID_BYTES + NA_META_BYTES + NA_TAB_BYTES;

static __device__ __align__(256) unsigned char scratch[NA_SCRATCH_BYTES];

constexpr int NA_MAX_DEVICES = 64;
std::mutex g_na_mu;
void* g_na_ptr[NA_MAX_DEVICES] = {};   // device address of `scratch` per device

// The device address of the scratch array on the current device, looked up once per device.
int na_scratch(void** out)
{
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return fail(nullptr, DS_EHIP, std::string("hipGetDevice: ") + hipGetErrorString(e));
    if (dev < 0 || dev >= NA_MAX_DEVICES) return fail(nullptr, DS_EINVAL, "device index beyond the scratch table");
    std::lock_guard<std::mutex> lock(g_na_mu);
    if (!g_na_ptr[dev]) {
        void* ptr = nullptr;
        e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            return fail(nullptr, DS_EHIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_na_ptr[dev] = ptr;
    }
    *out = g_na_ptr[dev];
    return DS_OK;
}

// Slot-to-group table of one instance (liquid, stick, container runs, each padded to 64), the
// group table, cleared dv sums and flag.
__global__ void __launch_bounds__(256) k_na_setup(int* __restrict__ meta, Group* __restrict__ groups,
                                                  double* __restrict__ dvsum, unsigned* __restrict__ flag, int nl,
                                                  int ns, int nc, int P)
{
    const int o1 = (nl + 63) / 64 * 64, o2 = o1 + (ns + 63) / 64 * 64;
    const long stride = (long)gridDim.x * blockDim.x;
    for (long p = (long)blockIdx.x * blockDim.x + threadIdx.x; p < P; p += stride) {
        int g = -1;
        if (p < nl) g = 0;
        else if (p >= o1 && p < o1 + ns) g = 1;
        else if (p >= o2 && p < o2 + nc) g = 2;
        meta[p] = g;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        groups[0] = Group{0, LIQUID, 0, nl};
        groups[1] = Group{0, STICK, o1, ns};
        groups[2] = Group{0, CONTAINER, o2, nc};
        for (int i = 0; i < 6; i++) dvsum[i] = 0.0;
        flag[0] = 0u;
    }
}

// Force of the substep (m_stick / dt * sum dv, the row ds_get_forces reports) and the flag.
__global__ void k_na_finish(const double* __restrict__ dvsum, const unsigned* __restrict__ flag, double scale,
                            double* __restrict__ force_out, uint32_t* __restrict__ status_out)
{
    if (threadIdx.x < 3) force_out[threadIdx.x] = scale * dvsum[3 + threadIdx.x];
    if (threadIdx.x == 0) status_out[0] = flag[0] ? 1u : 0u;
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
    long P = 0, nreal = 0;
    for (int b = 0; b < n_inst; b++) {
        const int cnt[3] = {n_liquid[b], n_stick[b], n_container[b]};
        c->nl.push_back(cnt[0]);
        c->ns.push_back(cnt[1]);
        c->nc.push_back(cnt[2]);
        for (int sp = 0; sp < 3; sp++) {
            if (cnt[sp] < 0) {
                delete c;
                return fail(nullptr, DS_EINVAL, "negative particle count");
            }
            c->groups.push_back(Group{b, sp, (int)P, cnt[sp]});
            P += ((long)cnt[sp] + 63) / 64 * 64;
            nreal += cnt[sp];
        }
    }
    if (P < 64) P = 64;
    if (P > (1l << 30)) {
        delete c;
        return fail(nullptr, DS_EINVAL, "too many particles");
    }
    c->P = (int)P;
    c->nreal = (int)nreal;
    c->meta_h.assign(P, -1);
    c->pid_h.assign(P, -1);
    for (size_t g = 0; g < c->groups.size(); g++)
        for (int i = 0; i < c->groups[g].count; i++) {
            c->meta_h[c->groups[g].offset + i] = (int)g;
            c->pid_h[c->groups[g].offset + i] = i;
        }
    c->x0_h.assign(3 * P, 0.5f);
    c->traj_h.assign((size_t)c->B * c->T * 3, 0.0f);
    c->params_h.assign((size_t)c->B * 4, 0.0);

    int rc;
#define ALLOC(ptr, n)                          \
    if ((rc = dalloc(c, &ptr, (n))) != 0) {     \
        std::string m = c->err;                \
        free_all(c);                           \
        delete c;                              \
        return fail(nullptr, rc, m);           \
    }
    ALLOC(c->d_meta, (size_t)P);
    ALLOC(c->d_groups, c->groups.size());
    ALLOC(c->d_x0, (size_t)3 * P);
    ALLOC(c->d_state, (size_t)c->stateFloats());
    ALLOC(c->d_ckpt, (size_t)c->stateFloats() * c->ncp);
    ALLOC(c->d_adj, (size_t)c->stateFloats());
    ALLOC(c->d_xbar, (size_t)3 * P);
    ALLOC(c->d_dvsum, (size_t)c->B * c->T * 3);
    ALLOC(c->d_dvbar, (size_t)c->B * c->T * 3);
    ALLOC(c->d_params, (size_t)c->B * 4);
    ALLOC(c->d_thbar, (size_t)c->B * 4);
    ALLOC(c->d_traj, (size_t)c->B * c->T * 3);
    ALLOC(c->d_flag, 4);   // [0] failure bits, [2..3] the 64-bit fall-back counter
#undef ALLOC
    hipError_t e = hipStreamCreate(&c->own_stream);
    c->stream = c->own_stream;
    if (e == hipSuccess) e = hipEventCreate(&c->ev0);
    if (e == hipSuccess) e = hipEventCreate(&c->ev1);
    if (e == hipSuccess)
        e = hipMemcpy(c->d_groups, c->groups.data(), sizeof(Group) * c->groups.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemset(c->d_flag, 0, 4 * sizeof(unsigned));
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
    if (ctx->own_stream && ctx->own_stream != ctx->stream) (void)hipStreamSynchronize(ctx->own_stream);
    free_all(ctx);
    delete ctx;
}

int ds_set_stream(ds_ctx* c, void* stream)
{
    if (!c) return DS_EINVAL;
    hipStream_t next = stream ? static_cast<hipStream_t>(stream) : c->own_stream;
    if (next == c->stream) return DS_OK;
    // the work queued so far comes before the caller's next work on the new stream
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->stream = next;
    return DS_OK;
}

int ds_set_option(ds_ctx* c, int32_t option, int32_t value)
{
    if (!c) return DS_EINVAL;
    if (option == DS_OPT_KERNELS) {
        if (value != DS_KERNELS_TILED && value != DS_KERNELS_PER_PARTICLE)
            return fail(c, DS_EINVAL, "DS_OPT_KERNELS takes DS_KERNELS_TILED or DS_KERNELS_PER_PARTICLE");
        if (value != c->mode) c->reset_done = false;
        c->mode = value;
        return DS_OK;
    }
    if (option == DS_OPT_KERNEL_TIMING) {
        if (value < 0) return fail(c, DS_EINVAL, "DS_OPT_KERNEL_TIMING takes N >= 0");
        c->kt_every = value;
        return DS_OK;
    }
    if (option == DS_OPT_PARTICLES_PER_THREAD || option == DS_OPT_GATHER_PARTICLES_PER_THREAD) {
        if (value != 1 && value != 2 && value != 4)
            return fail(c, DS_EINVAL, "particles per thread must be 1, 2 or 4");
        int& dst = option == DS_OPT_PARTICLES_PER_THREAD ? c->ppt : c->gppt;
        if (value != dst) {
            c->reset_done = false;
            drop_graphs(c);
        }
        dst = value;
        return DS_OK;
    }
    if (option == DS_OPT_GRAPH) {
        if (value != 0 && value != 1) return fail(c, DS_EINVAL, "DS_OPT_GRAPH takes 0 or 1");
        c->use_graph = value == 1;
        if (!c->use_graph) drop_graphs(c);
        return DS_OK;
    }
    if (option == DS_OPT_SCATTER_CHUNKS) {
        if (value != DS_SCATTER_ROWS && value != DS_SCATTER_PACKED)
            return fail(c, DS_EINVAL, "DS_OPT_SCATTER_CHUNKS takes DS_SCATTER_ROWS or DS_SCATTER_PACKED");
        if (c->packed != (value == DS_SCATTER_PACKED)) drop_graphs(c);
        c->packed = value == DS_SCATTER_PACKED;
        return DS_OK;
    }
    if (option == DS_OPT_CHUNK_GRID) {
        if (value != 0 && (value < 64 || value > 1048576))
            return fail(c, DS_EINVAL, "DS_OPT_CHUNK_GRID takes 0 (default) or 64..1048576 workgroups");
        const long g = value == 0 ? CHUNK_GRID : value;
        if (g != c->chunk_grid) drop_graphs(c);
        c->chunk_grid = g;
        return DS_OK;
    }
    if (option == DS_OPT_SORT_EVERY) {
        if (value < 0 || value > c->K || (value > 0 && c->K % value != 0))
            return fail(c, DS_EINVAL, "DS_OPT_SORT_EVERY takes 0 (the checkpoint interval) or a divisor of it");
        const int v = value == c->K ? 0 : value;
        if (v != c->sort_every) {
            // the sort tables are sized by the sorts per segment: reallocated at the next reset
            if (c->stream) (void)hipStreamSynchronize(c->stream);
            free_mode(c);
            c->reset_done = false;
        }
        c->sort_every = v;
        return DS_OK;
    }
    if (option == DS_OPT_G2P_SPLIT) {
        if (value != 1 && value != 2 && value != 4) return fail(c, DS_EINVAL, "DS_OPT_G2P_SPLIT takes 1, 2 or 4");
        if (c->g2p_split != value) drop_graphs(c);
        c->g2p_split = (int)value;
        return DS_OK;
    }
    if (option == DS_OPT_XGRAD_STREAM) {
        if (value != 0 && value != 1) return fail(c, DS_EINVAL, "DS_OPT_XGRAD_STREAM takes 0 or 1");
        if (c->xgrad_stream != (value == 1)) drop_graphs(c);
        c->xgrad_stream = value == 1;
        return DS_OK;
    }
    if (option == DS_OPT_REDUCE_PARTS) {
        if (value != 1 && value != 2 && value != 4) return fail(c, DS_EINVAL, "DS_OPT_REDUCE_PARTS takes 1, 2 or 4");
        if (c->rparts != value) drop_graphs(c);
        c->rparts = (int)value;
        return DS_OK;
    }
    if (option == DS_OPT_OVERLAP) {
        if (value != 0 && value != 1) return fail(c, DS_EINVAL, "DS_OPT_OVERLAP takes 0 or 1");
        if (c->overlap != (value == 1)) drop_graphs(c);
        c->overlap = value == 1;
        return DS_OK;
    }
    if (option == DS_OPT_TILE_WINDOW) {
        if (value < 0 || value > TS - 3) return fail(c, DS_EINVAL, "DS_OPT_TILE_WINDOW takes 0..5");
        if (c->dsc.tile_win != value) drop_graphs(c);   // the scene constants are kernel arguments
        c->dsc.tile_win = (int)value;
        return DS_OK;
    }
    if (option == DS_OPT_COUNT_FALLBACK) {
        if (value != 0 && value != 1) return fail(c, DS_EINVAL, "DS_OPT_COUNT_FALLBACK takes 0 or 1");
        if (c->dsc.count_fallback != (int)value) drop_graphs(c);
        c->dsc.count_fallback = (int)value;
        return DS_OK;
    }
    if (option == DS_OPT_FUSED_STEP) {
        if (value != 0 && value != 1) return fail(c, DS_EINVAL, "DS_OPT_FUSED_STEP takes 0 or 1");
        if (c->fuse_step != (value == 1)) drop_graphs(c);
        c->fuse_step = value == 1;
        return DS_OK;
    }
    if (option == DS_OPT_ADJOINT_P2G) {
        if (value != DS_ADJ_P2G_FUSED && value != DS_ADJ_P2G_SPLIT)
            return fail(c, DS_EINVAL, "DS_OPT_ADJOINT_P2G takes DS_ADJ_P2G_FUSED or DS_ADJ_P2G_SPLIT");
        if (c->adj_split != (value == DS_ADJ_P2G_SPLIT)) drop_graphs(c);
        c->adj_split = value == DS_ADJ_P2G_SPLIT;
        return DS_OK;
    }
    return fail(c, DS_EINVAL, "unknown option");
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
    static const char* names[3] = {"liquid", "stick", "container"};
    // every stencil inside the grid, as the kernels require (their device flag would otherwise
    // fail the first substep): the base cell int(x/dx - 0.5) in [0, n_grid - 3] per axis
    const float inv_dx = c->dsc.inv_dx;
    const int n = c->desc.n_grid;
    for (int sp = 0; sp < 3; sp++) {
        const Group& g = c->groups[inst * 3 + sp];
        if (g.count > 0 && !src[sp]) return fail(c, DS_EINVAL, "NULL positions for a non-empty species");
        for (int i = 0; i < g.count; i++)
            for (int a = 0; a < 3; a++) {
                const float X = src[sp][3 * (size_t)i + a] * inv_dx;
                const int b = std::isfinite(X) ? (int)(X - 0.5f) : -1;
                if (b < 0 || b > n - 3)
                    return fail(c, DS_EINVAL, std::string(names[sp]) + " particle " + std::to_string(i) +
                                                  " is outside the grid's interior or not finite");
            }
    }
    for (int sp = 0; sp < 3; sp++) {
        const Group& g = c->groups[inst * 3 + sp];
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
    if (c->step > 0) c->tape_stale = true;
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
    if (c->step > 0 && std::memcmp(c->params_h.data(), params, sizeof(double) * 4 * c->B) != 0) c->tape_stale = true;
    std::memcpy(c->params_h.data(), params, sizeof(double) * 4 * c->B);
    HIPCHK(c, hipMemcpyAsync(c->d_params, c->params_h.data(), sizeof(double) * 4 * c->B, hipMemcpyHostToDevice,
                             c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return DS_OK;
}

int ds_reset(ds_ctx* c)
{
    if (!c) return DS_EINVAL;
    RCHK(alloc_mode(c));
    HIPCHK(c, hipMemcpyAsync(c->d_x0, c->x0_h.data(), sizeof(float) * 3 * c->P, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->d_meta, c->meta_h.data(), sizeof(int) * c->P, hipMemcpyHostToDevice, c->stream));
    if (c->mode == DS_KERNELS_TILED) {
        HIPCHK(c, hipMemcpyAsync(c->d_pid, c->pid_h.data(), sizeof(int) * c->P, hipMemcpyHostToDevice, c->stream));
    }
    hipLaunchKernelGGL(k_init_state, dim3(nblocks(c->P)), dim3(256), 0, c->stream, c->d_state, c->d_x0, c->d_meta,
                       c->P, 0.5f);
    fill_zero_f64(c, c->d_dvsum, (long)c->B * c->T * 3);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->step = 0;
    c->reset_done = true;
    c->tape_stale = false;
    return DS_OK;
}

// Capture the launch sequence of body() on the context's stream into an executable graph
// (nothing runs while capturing; the host-side step counter is restored).
int capture_graph(ds_ctx* c, int (*body)(ds_ctx*, int), int arg, hipGraphExec_t* out)
{
    const int step0 = c->step;
    HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    c->capturing = true;
    const int rc = body(c, arg);
    c->capturing = false;
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(c->stream, &g);
    c->step = step0;
    if (rc != DS_OK) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess || !g) return fail(c, DS_EHIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    const hipError_t ei = hipGraphInstantiate(out, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (ei != hipSuccess) {
        *out = nullptr;
        return fail(c, DS_EHIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(ei));
    }
    return DS_OK;
}

int ds_forward(ds_ctx* c, int32_t nsteps)
{
    if (!c) return DS_EINVAL;
    if (!c->reset_done) return fail(c, DS_ESTATE, "ds_reset must be called before ds_forward");
    if (nsteps < 0 || c->step + nsteps > c->T - 1)
        return fail(c, DS_EINVAL, "nsteps exceeds the trajectory (max_steps - 1 substeps in total)");
    const bool graph = c->use_graph && c->mode == DS_KERNELS_TILED && c->step == 0 && nsteps == c->T - 1;
    if (graph && !c->fwd_exec) RCHK(capture_graph(c, forward_tiled, nsteps, &c->fwd_exec));
    HIPCHK(c, hipEventRecord(c->ev0, c->stream));
    if (graph) {
        // the whole rollout (sorts included) as one graph launch
        HIPCHK(c, hipGraphLaunch(c->fwd_exec, c->stream));
        c->step += nsteps;
    } else if (c->mode == DS_KERNELS_TILED) {
        RCHK(forward_tiled(c, nsteps));
    } else {
        RCHK(forward_ref(c, nsteps));
    }
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipEventRecord(c->ev1, c->stream));
    HIPCHK(c, hipEventSynchronize(c->ev1));
    float ms = 0.0f;
    HIPCHK(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
    c->fwd_ms = ms;
    c->kt_on = false;
    kt_collect(c);
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
    const size_t n = (size_t)c->B * c->T * 3;
    HIPCHK(c, hipMemcpyAsync(forces, c->d_dvsum, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    const double scale = c->desc.m_stick / c->desc.dt;
    for (size_t i = 0; i < n; i++) forces[i] *= scale;
    return DS_OK;
}

int ds_backward(ds_ctx* c, const double* dl_dforces, double* dl_dparams)
{
    if (!c) return DS_EINVAL;
    if (!dl_dforces || !dl_dparams) return fail(c, DS_EINVAL, "NULL argument");
    if (!c->reset_done || c->step < 1) return fail(c, DS_ESTATE, "ds_backward needs a recorded forward rollout");
    if (c->tape_stale)
        return fail(c, DS_ESTATE, "the parameters or a trajectory changed after the rollout was recorded: "
                                  "ds_reset and ds_forward again before ds_backward");
    // dL/d(dv sum at step s) = m_stick/dt * dL/dF(s)
    std::vector<double> dvb((size_t)c->B * c->T * 3);
    const double scale = c->desc.m_stick / c->desc.dt;
    for (size_t i = 0; i < dvb.size(); i++) dvb[i] = scale * dl_dforces[i];
    HIPCHK(c, hipMemcpyAsync(c->d_dvbar, dvb.data(), sizeof(double) * dvb.size(), hipMemcpyHostToDevice, c->stream));
    if (c->mode == DS_KERNELS_TILED && c->overlap) RCHK(alloc_overlap(c));
    if (c->mode == DS_KERNELS_TILED && c->xgrad_stream) RCHK(alloc_xgrad(c));
    const bool graph = c->use_graph && c->mode == DS_KERNELS_TILED && c->step == c->T - 1;
    if (graph && !c->bwd_exec) RCHK(capture_graph(c, backward_tiled_n, 0, &c->bwd_exec));
    HIPCHK(c, hipEventRecord(c->ev0, c->stream));
    fill_zero_f32(c, c->d_adj, c->stateFloats());
    fill_zero_f64(c, c->d_thbar, (long)c->B * 4);
    if (graph) HIPCHK(c, hipGraphLaunch(c->bwd_exec, c->stream));
    else if (c->mode == DS_KERNELS_TILED) RCHK(backward_tiled(c));
    else RCHK(backward_ref(c));
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipEventRecord(c->ev1, c->stream));
    HIPCHK(c, hipMemcpyAsync(dl_dparams, c->d_thbar, sizeof(double) * 4 * c->B, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipEventSynchronize(c->ev1));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    float ms = 0.0f;
    HIPCHK(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
    c->bwd_ms = ms;
    c->kt_on = false;
    kt_collect(c);
    return check_flag(c, "ds_backward");
}

int ds_read_liquid(ds_ctx* c, int32_t inst, float* x, float* v, float* C, float* F)
{
    if (!c) return DS_EINVAL;
    if (inst < 0 || inst >= c->B) return fail(c, DS_EINVAL, "instance index out of range");
    if (!c->reset_done) return fail(c, DS_ESTATE, "no state: call ds_reset first");
    RCHK(read_species(c, inst, LIQUID, SX, 3, x));
    RCHK(read_species(c, inst, LIQUID, SV, 3, v));
    RCHK(read_species(c, inst, LIQUID, SC, 9, C));
    return read_species(c, inst, LIQUID, SF, 9, F);
}

int ds_read_stick(ds_ctx* c, int32_t inst, float* x, float* v, float* C)
{
    if (!c) return DS_EINVAL;
    if (inst < 0 || inst >= c->B) return fail(c, DS_EINVAL, "instance index out of range");
    if (!c->reset_done) return fail(c, DS_ESTATE, "no state: call ds_reset first");
    RCHK(read_species(c, inst, STICK, SX, 3, x));
    RCHK(read_species(c, inst, STICK, SV, 3, v));
    return read_species(c, inst, STICK, SC, 9, C);
}

int ds_substep_noalloc(const ds_scene_desc* desc, int32_t n_liquid, int32_t n_stick, int32_t n_container,
                       const double* params, const float* state_in, float* state_out, const float* traj2,
                       double* force_out, uint32_t* status_out, void* stream)
{
    if (!desc || !params || !state_in || !state_out || !traj2 || !force_out || !status_out)
        return fail(nullptr, DS_EINVAL, "NULL argument");
    if (desc->n_grid < 8 || desc->n_grid > NA_MAX_GRID)
        return fail(nullptr, DS_EINVAL, "n_grid out of the no-allocation range [8, 64]");
    if (desc->model < 0 || desc->model > 2) return fail(nullptr, DS_EINVAL, "model must be 0, 1 or 2");
    if (desc->k_from_param && desc->model != DS_MODEL_HERSCHEL_BULKLEY)
        return fail(nullptr, DS_EINVAL, "k_from_param needs the Herschel-Bulkley model (params[3] is free only there)");
    if (!(desc->dt > 0.0)) return fail(nullptr, DS_EINVAL, "dt must be > 0");
    if (n_liquid < 0 || n_stick < 0 || n_container < 0) return fail(nullptr, DS_EINVAL, "negative particle count");
    const long P = ((long)n_liquid + 63) / 64 * 64 + ((long)n_stick + 63) / 64 * 64 + ((long)n_container + 63) / 64 * 64;
    if (P > NA_MAX_SLOTS)
        return fail(nullptr, DS_EINVAL, "more particle slots than the no-allocation limit 262144 (runs padded to 64)");
    if (P == 0) return fail(nullptr, DS_EINVAL, "no particles");
    void* base = nullptr;
    RCHK(na_scratch(&base));
    unsigned char* b = static_cast<unsigned char*>(base);
    float4* grid = reinterpret_cast<float4*>(b);
    float4* vout = reinterpret_cast<float4*>(b + NA_GRID_BYTES);
    int* meta = reinterpret_cast<int*>(b + 2 * NA_GRID_BYTES);
    unsigned char* tab = b + 2 * NA_GRID_BYTES + NA_META_BYTES;
    Group* groups = reinterpret_cast<Group*>(tab);
    double* dvsum = reinterpret_cast<double*>(tab + 64);
    unsigned* flag = reinterpret_cast<unsigned*>(tab + 128);
    const DevScene dsc = make_dev_scene(*desc);
    const hipStream_t st = static_cast<hipStream_t>(stream);
    const long nodes = (long)desc->n_grid * desc->n_grid * desc->n_grid;
    const int Pi = (int)P;
    hipLaunchKernelGGL(k_na_setup, dim3(nblocks(P)), dim3(256), 0, st, meta, groups, dvsum, flag, (int)n_liquid,
                       (int)n_stick, (int)n_container, Pi);
    hipLaunchKernelGGL(k_fill_f4, dim3(nblocks(nodes)), dim3(256), 0, st, grid, nodes);
    hipLaunchKernelGGL(k_p2g_ref, dim3(nblocks(P)), dim3(256), 0, st, state_in, state_out, meta,
                       (const Group*)groups, params, grid, dsc, Pi, flag);
    hipLaunchKernelGGL(k_gridop_ref, dim3(nblocks(nodes)), dim3(256), 0, st, (const float4*)grid, vout, dsc, 1);
    hipLaunchKernelGGL(k_g2p_ref, dim3(nblocks(P)), dim3(256), 0, st, state_in, state_out, meta,
                       (const Group*)groups, (const float4*)vout, traj2, 2, 0, dvsum, dsc, Pi, flag);
    hipLaunchKernelGGL(k_na_finish, dim3(1), dim3(64), 0, st, (const double*)dvsum, (const unsigned*)flag,
                       desc->m_stick / desc->dt, force_out, status_out);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(nullptr, DS_EHIP, std::string("ds_substep_noalloc launch: ") + hipGetErrorString(e));
    return DS_OK;
}

int ds_fallback_count(ds_ctx* c, int64_t* count, int32_t reset)
{
    if (!c || !count) return DS_EINVAL;
    unsigned long long n = 0;
    HIPCHK(c, hipMemcpyAsync(&n, c->d_flag + 2, sizeof(n), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    *count = (int64_t)n;
    if (reset) {
        const unsigned long long z = 0;
        HIPCHK(c, hipMemcpyAsync(c->d_flag + 2, &z, sizeof(z), hipMemcpyHostToDevice, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
    }
    return DS_OK;
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

int ds_kernel_timing(ds_ctx* c, double* total_ms, int64_t* launches, int32_t reset)
{
    if (!c) return DS_EINVAL;
    for (int k = 0; k < DS_NKERNELS; k++) {
        if (total_ms) total_ms[k] = c->kt_ms[k];
        if (launches) launches[k] = c->kt_n[k];
        if (reset) {
            c->kt_ms[k] = 0.0;
            c->kt_n[k] = 0;
        }
    }
    return DS_OK;
}

int ds_svd3(const double* F, int32_t n, double* U, double* S, double* V)
{
    if (!F || !U || !S || !V || n < 1) return fail(nullptr, DS_EINVAL, "bad argument");
    const size_t nb9 = 9 * sizeof(double) * n, nb3 = 3 * sizeof(double) * n;
    double* buf;
    if (hipMalloc((void**)&buf, 3 * nb9 + nb3) != hipSuccess) {
        (void)hipGetLastError();
        return fail(nullptr, DS_ENOMEM, "hipMalloc");
    }
    double *dF = buf, *dU = dF + 9 * (size_t)n, *dV = dU + 9 * (size_t)n, *dS = dV + 9 * (size_t)n;
    hipError_t e = hipMemcpy(dF, F, nb9, hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_svd3_test, dim3(nblocks(n)), dim3(256), 0, 0, dF, n, dU, dS, dV);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpy(U, dU, nb9, hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(S, dS, nb3, hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(V, dV, nb9, hipMemcpyDeviceToHost);
    (void)hipFree(buf);
    if (e != hipSuccess) return fail(nullptr, DS_EHIP, hipGetErrorString(e));
    return DS_OK;
}

int ds_svd3_values(const double* F, int32_t n, double* S)
{
    if (!F || !S || n < 1) return fail(nullptr, DS_EINVAL, "bad argument");
    const size_t nb9 = 9 * sizeof(double) * n, nb3 = 3 * sizeof(double) * n;
    double* buf;
    if (hipMalloc((void**)&buf, nb9 + nb3) != hipSuccess) {
        (void)hipGetLastError();
        return fail(nullptr, DS_ENOMEM, "hipMalloc");
    }
    double *dF = buf, *dS = dF + 9 * (size_t)n;
    hipError_t e = hipMemcpy(dF, F, nb9, hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_svd3_values_test, dim3(nblocks(n)), dim3(256), 0, 0, dF, n, dS);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpy(S, dS, nb3, hipMemcpyDeviceToHost);
    (void)hipFree(buf);
    if (e != hipSuccess) return fail(nullptr, DS_EHIP, hipGetErrorString(e));
    return DS_OK;
}

int ds_svd3_backward(const double* U, const double* S, const double* V, const double* gU, const double* gS,
                     const double* gV, int32_t n, double* gF)
{
    if (!U || !S || !V || !gU || !gS || !gV || !gF || n < 1) return fail(nullptr, DS_EINVAL, "bad argument");
    const size_t nb9 = 9 * sizeof(double) * n, nb3 = 3 * sizeof(double) * n;
    double* buf;
    if (hipMalloc((void**)&buf, 5 * nb9 + 2 * nb3) != hipSuccess) {
        (void)hipGetLastError();
        return fail(nullptr, DS_ENOMEM, "hipMalloc");
    }
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
