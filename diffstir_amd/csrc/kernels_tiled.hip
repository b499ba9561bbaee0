// kernels_tiled.hip -- the substep and its adjoint on sorted particles with LDS grid tiles,
// without floating-point atomics on the hot path.
//
// Particles are re-sorted at every checkpoint (every K substeps, or every R dividing K) into
// buckets of (instance, species, 4x4x4-cell block).  Inside a bucket the order is round-robin over the
// block's cells (the first particle of every occupied cell, then the second, ...), so the 64
// lanes of a wavefront sit in different cells.  A bucket is cut into chunks of up to
// 256 * PPT particles; one workgroup (4 wavefronts, PPT particles per thread) takes a chunk.
//
// Each chunk works on an 8x8x8-node tile covering nodes [o-1, o+7) per axis (o = the block's
// first cell): every stencil whose base cell lies in [o-1, o+4], so particles may drift one
// cell out of their block within a segment.  A stencil that drifts further goes to a global
// fall-back grid with FP64 atomics of grain-rounded contributions (rare; exact sums, so their
// order does not matter).
//
// Scatter without atomics.  LDS float atomics (ds_add_f32) run at about 1/24 of the rate of a
// plain LDS read-modify-write on gfx950, and global float atomics are slow as well.  So:
//   * every wavefront owns a private LDS tile;
//   * before scattering, the lanes of a wavefront claim their stencil base node in a per-wave
//     table; lanes that share a base are served in successive passes, so within one pass the
//     27 read-modify-writes of different lanes never hit the same node in the same
//     instruction, and a wavefront's LDS operations complete in program order (compiler
//     memory barriers keep the compiler from reordering them);
//   * the workgroup sums its four wave tiles and stores the sum in the chunk's staging slot
//     (8 KB of plain coalesced stores);
//   * a reduction kernel, one thread per grid node (one to four wavefronts per 4x4x4 node
//     block), adds the staging slots of the chunks whose tiles cover the node plus the
//     fall-back grid, and applies the grid update there.  The sum order is fixed, so the grid
//     is deterministic.
//
//   k_p2g_tiled       P2G of a chunk into its staging slot
//   k_grid_reduce     staged momentum/mass -> grid update -> velocity (and momentum/mass)
//   k_g2p_tiled       gather from a velocity tile, particle update, stick force
//   k_g2p_adj_tiled   adjoint G2P: dL/dv scatter to staging
//   k_vbar_reduce     staged dL/dv -> adjoint grid update -> dL/d(momentum, mass)
//   k_p2g_adj_xgrad   adjoint P2G's position gradient through G2P's gather (optional, own stream)
//   k_p2g_adj_tiled   adjoint P2G gathers and stress adjoint, FP64 parameter-gradient reduction
// The P2G wave tiles hold one float4 (momentum, mass) per node in the staging slot's node order
// (x 64, y 8, z 1), so a node's read-modify-write is one 16-byte LDS read and one write.  The
// other LDS tiles are component planes.  The adjoint G2P's wave tiles and the forward G2P's
// velocity tile use strides (x 68, y 8, z 1): for the 64 cells of a block the node offsets of one
// stencil point fall in 32 distinct banks per half-wavefront.  The adjoint P2G's velocity tiles
// use (x 112, y 12, z 1), with which its tile reads take fewer LDS instructions (more of them
// pair); its dL/d(momentum, mass) tile holds one float4 per node in staging order.
#include "ds_particle.h"

namespace ds {

constexpr int TB = 4;                 // cells per block axis
constexpr int TS = 8;                 // tile nodes per axis
constexpr int TSY = 8;
constexpr int TSX = 68;
constexpr int TILE = TS * TSX;        // floats per component plane
constexpr int TNODES = TS * TS * TS;  // nodes of a staging slot (dense, index lx*64 + ly*8 + lz)
constexpr int WG = 256;               // threads per workgroup
constexpr int WAVES = WG / 64;

// A chunk: up to WAVES * PPT consecutive rows of one bucket (row r of a bucket holds the r-th
// particle of every cell with more than r particles, so the lanes of a row have distinct
// stencil bases).  Rows are global indices into the checkpoint's row table: row k covers
// slots [rows[k], rows[k + 1]).
struct Chunk {
    int inst, species, ox, oy, oz, start, row0, nrows;
};

// A chunk for the gather-only kernels (G2P, adjoint P2G), which have no scatter conflicts: a
// contiguous run of up to WG * PPT particles of one bucket.
struct GChunk {
    int inst, species, ox, oy, oz, start, count, pad;
};

// The slots of lane `lane` in row `lr` (local to the chunk) of chunk ch: sets p and returns
// whether the lane has a particle there.  lr is uniform across the wavefront.
__device__ __forceinline__ bool chunk_slot(const Chunk& ch, const int* __restrict__ rows, int lr, int lane, int& p)
{
    if (lr >= ch.nrows) return false;
    const int rs = rows[ch.row0 + lr], re = rows[ch.row0 + lr + 1];
    p = rs + lane;
    return p < re;
}

// The particle of (iteration r, wavefront, lane) in a scatter kernel: row r * WAVES + wave of a
// row chunk, or slot r * WG + wave * 64 + lane of a packed (gather) chunk, whose lanes may share
// stencil bases (the claim protocol then serves them in successive passes).
__device__ __forceinline__ bool scatter_slot(const Chunk& ch, const int* __restrict__ rows, int r, int wave, int lane,
                                             int& p)
{
    return chunk_slot(ch, rows, r * WAVES + wave, lane, p);
}

__device__ __forceinline__ bool scatter_slot(const GChunk& ch, const int* __restrict__, int r, int wave, int lane,
                                             int& p)
{
    const int idx = r * WG + wave * 64 + lane;
    p = ch.start + idx;
    return idx < ch.count;
}

__device__ __forceinline__ int tile_id(int lx, int ly, int lz) { return lx * TSX + ly * TSY + lz; }

// A component-plane layout: node (lx, ly, lz) at lx * XSTR + ly * YSTR + lz of each plane, the
// planes `size` floats apart.
template <int XSTR, int YSTR>
struct Planes {
    static constexpr int size = TS * XSTR;
    __device__ static __forceinline__ int id(int lx, int ly, int lz) { return lx * XSTR + ly * YSTR + lz; }
};
using PlanesS = Planes<TSX, TSY>;   // scatter wave tiles, forward gathers
using PlanesG = Planes<112, 12>;    // adjoint P2G velocity tiles (1.4 KB more LDS per plane)

// win: DevScene::tile_win, at most TS - 3 (the stencil's 3 nodes per axis stay inside the tile)
__device__ __forceinline__ bool stencil_in_tile(const Stencil& s, int x0, int y0, int z0, int win)
{
    const unsigned a = (unsigned)(s.b[0] - x0), b = (unsigned)(s.b[1] - y0), c = (unsigned)(s.b[2] - z0);
    return a <= (unsigned)win && b <= (unsigned)win && c <= (unsigned)win;
}

// Compiler-only memory barrier: keeps LDS accesses in program order without waiting for
// them (a wavefront's LDS operations execute in order in hardware).
__device__ __forceinline__ void lds_order() { asm volatile("" ::: "memory"); }

// Makes a stencil's offsets and weights opaque to the compiler at this point.  The claim loop
// below runs the scatter in whichever pass the lane wins; without this, loop-invariant code
// motion computes all 27 node contributions before the loop and keeps them live across it
// (over 100 extra registers, halving the wavefronts a SIMD can hold).
__device__ __forceinline__ Stencil opaque(const Stencil& s)
{
    Stencil t = s;
#pragma unroll
    for (int a = 0; a < 3; a++) asm volatile("" : "+v"(t.fx[a]));
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
        for (int b = 0; b < 3; b++) asm volatile("" : "+v"(t.w[a][b]));
    return t;
}

// Reduction blocks: the per-node reductions run on 4x4x4 node blocks shifted by one node, block
// r owning nodes [4r - 1, 4r + 3) per axis (NR = n / 4 + 1 blocks per axis).  Every node of block
// r then lies in the tiles of the same buckets, r - 1 and r per axis, so the block's nodes share
// one list of covering chunk slots.
__device__ __forceinline__ int rblock(int node) { return (node + 1) / TB; }

// Marks the reduction blocks a stencil touches (its corner nodes' blocks) in the per-step
// fall-back flags, so the reduction visits them even where no chunk tile covers them.
__device__ __forceinline__ void mark_fallback_blocks(int* __restrict__ fbflag, int inst, const Stencil& s, int NR)
{
    const int nr3 = NR * NR * NR;
#pragma unroll
    for (int q = 0; q < 8; q++) {
        const int rx = rblock(s.b[0] + ((q >> 2) & 1) * 2), ry = rblock(s.b[1] + ((q >> 1) & 1) * 2),
                  rz = rblock(s.b[2] + (q & 1) * 2);
        atomicOr(fbflag + (long)inst * nr3 + (rx * NR + ry) * NR + rz, 1);
    }
}

// Read-modify-write scatter into a wavefront's private tile (see the claim protocol above).
// The barrier after each node keeps the next node's reads behind this node's writes: another
// lane's stencil may reach the same node at a different offset.  WaveScatter4: one float4 per
// node (TNODES per wave, staging order); WaveScatter21 (below): a float2 and a float plane.
struct WaveScatter4 {
    float4* t;
    int x0, y0, z0;
    __device__ __forceinline__ void add(int i, int j, int k, float a, float b, float c, float d)
    {
        float4* q = t + ((i - x0) * TS + (j - y0)) * TS + (k - z0);
        const float4 o = *q;
        *q = make_float4(o.x + a, o.y + b, o.z + c, o.w + d);
        lds_order();
    }
};

// WaveScatter21: a float2 plane (first two components) followed by a float plane (the third), so
// a node's read-modify-write is one 8-byte and one 4-byte LDS read and write.
struct WaveScatter21 {
    float* t;
    int x0, y0, z0;
    __device__ __forceinline__ void add(int i, int j, int k, float a, float b, float c)
    {
        const int id = tile_id(i - x0, j - y0, k - z0);
        float2* q2 = reinterpret_cast<float2*>(t) + id;
        float* q1 = t + 2 * TILE + id;
        const float2 o = *q2;
        const float oc = *q1;
        *q2 = make_float2(o.x + a, o.y + b);
        *q1 = oc + c;
        lds_order();
    }
};

// Gather accessors for a workgroup tile.  A particle whose stencil left the tile gathers from
// global memory instead; the kernels branch once per particle between the two accessors, so
// each inlined 27-node gather is a single straight-line form.
template <class L>
struct TileVT {
    const float* t;
    int x0, y0, z0;
    __device__ __forceinline__ float3 get(int i, int j, int k) const
    {
        const int id = L::id(i - x0, j - y0, k - z0);
        return make_float3(t[id], t[L::size + id], t[2 * L::size + id]);
    }
};
using TileV = TileVT<PlanesS>;

// A float4 tile in staging node order (x 64, y 8, z 1): one 16-byte read per node.
struct TileQ4 {
    const float4* t;
    int x0, y0, z0;
    __device__ __forceinline__ float4 get(int i, int j, int k) const
    {
        return t[((i - x0) * TS + (j - y0)) * TS + (k - z0)];
    }
};

__device__ __forceinline__ void load_particle_t(const float* __restrict__ st, long LP, int p, bool with_F, ParticleIn& q)
{
#pragma unroll
    for (int a = 0; a < 3; a++) q.x[a] = st[(SX + a) * LP + p];
#pragma unroll
    for (int a = 0; a < 3; a++) q.v[a] = st[(SV + a) * LP + p];
#pragma unroll
    for (int a = 0; a < 9; a++) q.C[a] = st[(SC + a) * LP + p];
#pragma unroll
    for (int a = 0; a < 9; a++) q.F[a] = with_F ? st[(SF + a) * LP + p] : ((a == 0 || a == 4 || a == 8) ? 1.0f : 0.0f);
}

// Zeroes nfloats (a multiple of 4) of a 16-byte aligned LDS array with 16-byte stores.
__device__ __forceinline__ void zero_lds(float* t, int nfloats)
{
    float4* t4 = reinterpret_cast<float4*>(t);
    for (int e = threadIdx.x; e < nfloats / 4; e += WG) t4[e] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}

__device__ __forceinline__ int node_key(const Stencil& s, int x0, int y0, int z0)
{
    return ((s.b[0] - x0) * TS + (s.b[1] - y0)) * TS + (s.b[2] - z0);
}

// Sum the float4 wave tiles (TNODES each, staging order) of the workgroup into the chunk's
// staging slot.
__device__ __forceinline__ void store_staging4(const float4* wt, float4* __restrict__ slot)
{
    for (int e = threadIdx.x; e < TNODES; e += WG) {
        float4 v = wt[e];
#pragma unroll
        for (int w = 1; w < WAVES; w++) {
            const float4 u = wt[w * TNODES + e];
            v.x += u.x;
            v.y += u.y;
            v.z += u.z;
            v.w += u.w;
        }
        slot[e] = v;
    }
}

// Sum the WaveScatter21 tiles (3 * TILE floats each) of the workgroup into the chunk's staging slot:
// TNODES float4 in (lx, ly, lz) order, the fourth component 0.
__device__ __forceinline__ void store_staging21(const float* wt, float4* __restrict__ slot)
{
    for (int e = threadIdx.x; e < TNODES; e += WG) {
        const int lx = e / (TS * TS), ly = (e / TS) % TS, lz = e % TS;
        const int id = tile_id(lx, ly, lz);
        float v[3] = {0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int w = 0; w < WAVES; w++) {
            const float2 u = reinterpret_cast<const float2*>(wt + w * 3 * TILE)[id];
            v[0] += u.x;
            v[1] += u.y;
            v[2] += wt[w * 3 * TILE + 2 * TILE + id];
        }
        slot[e] = make_float4(v[0], v[1], v[2], 0.0f);
    }
}

// ---------------------------------------------------------------------------- forward

// Row chunks: at most 128 registers (four wavefronts per SIMD, as the LDS allows); the packed
// form's claim loop needs more.
template <class CH> struct P2GMinWaves { static constexpr int value = 1; };
template <> struct P2GMinWaves<Chunk> { static constexpr int value = 4; };

template <int PPT, class CH>
__global__ void __launch_bounds__(WG, P2GMinWaves<CH>::value) k_p2g_tiled(const float* __restrict__ in, float* __restrict__ fout,
                                                  float* __restrict__ aout, const CH* __restrict__ chunks,
                                                  const int* __restrict__ rows,
                                                  const int* __restrict__ nchunks_p, int maxchunks,
                                                  const double* __restrict__ params, float4* __restrict__ staging,
                                                  double* __restrict__ fallback, int* __restrict__ fbflag, int NR,
                                                  const double* __restrict__ jin, DevScene sc, int P,
                                                  unsigned* __restrict__ flag)
{
    // chunk count of the checkpoint, written by the bucket sort (bounded by the table size)
    const int nchunks = min(*nchunks_p, maxchunks);
    __shared__ float4 wt[WAVES * TNODES];
    __shared__ unsigned char own[WAVES * TNODES];
    const int n = sc.n_grid;
    const long nn = (long)n * n * n;
    const long LP = P;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    unsigned char* ow = own + wave * TNODES;
    for (int c = blockIdx.x; c < nchunks; c += gridDim.x) {
        const CH ch = chunks[c];
        const int x0 = ch.ox - 1, y0 = ch.oy - 1, z0 = ch.oz - 1;
        const bool liq = ch.species == LIQUID;
        // the first row's particles are loaded before the tile is cleared, and each next row's
        // while the current one scatters
        int p = 0;
        bool valid = scatter_slot(ch, rows, 0, wave, lane, p);
        ParticleIn q;
        double J = 1.0;
        if (valid) {
            load_particle_t(in, LP, p, liq, q);
            if (liq) J = jin[p];
        }
        zero_lds(reinterpret_cast<float*>(wt), WAVES * 4 * TNODES);
        const InstParams ip = load_params(sc, params, ch.inst);
        WaveScatter4 acc{wt + wave * TNODES, x0, y0, z0};
#pragma unroll 1
        for (int r = 0; r < PPT; r++) {
            Stencil s;
            P2GPayload pl;
            bool it = false;
            if (valid) {
                if (!stencil(sc, q.x[0], q.x[1], q.x[2], s)) atomicOr(flag, 1u);
                float Fn[9];
                p2g_prepare(sc, ch.species, ip, q, Fn, pl, &J);
                if (fout && ch.species == LIQUID) {
#pragma unroll
                    for (int a = 0; a < 9; a++) fout[(SF + a) * LP + p] = Fn[a];
                }
                if (aout && ch.species == LIQUID) {
#pragma unroll
                    for (int a = 0; a < 9; a++) aout[a * LP + p] = pl.A[a];
                }
                it = stencil_in_tile(s, x0, y0, z0, sc.tile_win);
                if (!it) {
                    mark_fallback_blocks(fbflag, ch.inst, s, NR);
                    if (sc.count_fallback) atomicAdd(reinterpret_cast<unsigned long long*>(flag + 2), 1ull);
                    ScatterExact4 g{fallback + 4 * (long)ch.inst * nn, n, fallback_magic(p2g_bound(sc, pl))};
                    p2g_scatter(sc, s, pl, g);
                }
            }
            int pn = 0;
            const bool vn = r + 1 < PPT && scatter_slot(ch, rows, r + 1, wave, lane, pn);
            if (vn) {
                load_particle_t(in, LP, pn, liq, q);
                if (liq) J = jin[pn];
            }
            if (r == 0) __syncthreads();   // every wavefront's tile is cleared
            bool pending = valid && it;
            const int key = pending ? node_key(s, x0, y0, z0) : 0;
            while (__ballot(pending)) {
                lds_order();
                if (pending) ow[key] = (unsigned char)lane;
                lds_order();
                if (pending && ow[key] == lane) {
                    p2g_scatter(sc, opaque(s), pl, acc);
                    pending = false;
                }
            }
            valid = vn;
            p = pn;
        }
        __syncthreads();
        store_staging4(wt, staging + (long)c * TNODES);
        __syncthreads();
    }
}

// The chunk counts and first chunks of the 8 buckets covering reduction block (rx, ry, rz) (buckets
// rx - 1 + dx, dx in {0, 1}, per axis), all three species, into LDS (index sp * 8 + dx * 4 + dy * 2
// + dz; 0 chunks outside the grid).
__device__ __forceinline__ void load_covering_buckets(int* s_nc, int* s_cs, const int* __restrict__ bnch,
                                                      const int* __restrict__ cstart, int inst, int rx, int ry,
                                                      int rz, int NB)
{
    for (int e = threadIdx.x; e < 24; e += blockDim.x) {
        const int sp = e >> 3;
        const int cx = rx - 1 + ((e >> 2) & 1), cy = ry - 1 + ((e >> 1) & 1), cz = rz - 1 + (e & 1);
        int nc = 0, cs = 0;
        if (cx >= 0 && cy >= 0 && cz >= 0 && cx < NB && cy < NB && cz < NB) {
            const int bucket = ((inst * 3 + sp) * NB + cx) * NB * NB + cy * NB + cz;
            nc = bnch[bucket];
            cs = cstart[bucket];
        }
        s_nc[e] = nc;
        s_cs[e] = cs;
    }
}

// Sum, in a fixed order, the staging entries of node (lx, ly, lz) of a reduction block from every
// chunk whose tile covers it.  The node lies in the tiles of buckets r - 1 and r per axis, at index
// l + 4 - 4d in the tile of bucket r - 1 + d.  The covering entries form one list, the same for
// every node of the block: the 24 (species, covering bucket) pairs in order, and within a pair
// its chunks in order.  The walk over the list is uniform across the wavefront (scalar); it goes
// in batches of KB entries whose addresses are formed first and whose loads are then all in
// flight at once, so a batch costs one memory latency, and the 64 lanes of a load read 64 nodes
// of one chunk slot.  This call takes list entries part, part + nparts, ... (the parts of a split
// reduction).
template <int KB>
__device__ __forceinline__ void sum_covering_list(float4& acc, const float4* __restrict__ staging, const int* s_nc,
                                                  const int* s_cs, bool inb, int lx, int ly, int lz, int ncomp,
                                                  int part, int nparts)
{
    // the node's index in the slot of pair q's bucket
    auto node_in_tile = [&](int q) {
        const int dx = (q >> 2) & 1, dy = (q >> 1) & 1, dz = q & 1;
        return ((lx + 4 - 4 * dx) * TS + (ly + 4 - 4 * dy)) * TS + (lz + 4 - 4 * dz);
    };
    // walk position: chunk k of pair q (q == 24: past the end); n, cs: pair q's chunk count and
    // first chunk
    int q = 0, k = part;
    int n = __builtin_amdgcn_readfirstlane(s_nc[0]), cs = __builtin_amdgcn_readfirstlane(s_cs[0]);
    auto settle = [&]() {
        while (q < 24 && k >= n) {
            k -= n;
            if (++q < 24) {
                n = __builtin_amdgcn_readfirstlane(s_nc[q]);
                cs = __builtin_amdgcn_readfirstlane(s_cs[q]);
            }
        }
    };
    settle();
    while (q < 24) {
        int a[KB];
#pragma unroll
        for (int j = 0; j < KB; j++) {
            a[j] = q < 24 && inb ? (cs + k) * TNODES + node_in_tile(q) : -1;
            k += nparts;
            settle();
        }
        float4 v[KB];
#pragma unroll
        for (int j = 0; j < KB; j++) v[j] = a[j] >= 0 ? staging[a[j]] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
#pragma unroll
        for (int j = 0; j < KB; j++) {
            acc.x += v[j].x;
            acc.y += v[j].y;
            acc.z += v[j].z;
            if (ncomp == 4) acc.w += v[j].w;
        }
    }
}

// Entries of a node's covering list loaded together (16 float4: 64 registers).
constexpr int REDUCE_BATCH = 16;

// The staged sum of one node of a reduction block in NPART parts: wavefront w of the workgroup
// takes entries w, w + NPART, ... of the block's covering list (so a block covered by many chunks
// waits for fewer memory latencies) and part 0 adds the others' partial sums in wavefront order.
// Part 0 also takes the fall-back grid (read and cleared) first.  Every thread of the workgroup
// must call this; the result is meaningful in part 0 only.
template <int NPART>
__device__ __forceinline__ float4 staged_node_sum(const float4* __restrict__ staging, const int* s_nc, const int* s_cs,
                                                  double* __restrict__ fallback, bool fl, long gid, bool inb, int lx,
                                                  int ly, int lz, int ncomp)
{
    __shared__ float4 s_part[NPART > 1 ? (NPART - 1) * 64 : 1];
    const int node = threadIdx.x & 63, part = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    float4 acc = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (inb && fl && part == 0) {
        // the fall-back grid's exact FP64 sums, rounded once
        double2* f = reinterpret_cast<double2*>(fallback + 4 * gid);
        const double2 a = f[0], b = f[1];
        acc = make_float4((float)a.x, (float)a.y, (float)b.x, (float)b.y);
        f[0] = make_double2(0.0, 0.0);
        f[1] = make_double2(0.0, 0.0);
    }
    sum_covering_list<REDUCE_BATCH>(acc, staging, s_nc, s_cs, inb, lx, ly, lz, ncomp, part, NPART);
    if (NPART > 1) {
        if (part > 0) s_part[(part - 1) * 64 + node] = acc;
        __syncthreads();
        if (part == 0) {
#pragma unroll
            for (int w = 0; w < NPART - 1; w++) {
                const float4 t = s_part[w * 64 + node];
                acc.x += t.x;
                acc.y += t.y;
                acc.z += t.z;
                acc.w += t.w;
            }
        }
    }
    return acc;
}

// One thread per grid node (NPART wavefronts per reduction block, see staged_node_sum): the
// sum of the staging slots of the chunks whose tiles cover the node, in a fixed order, plus the
// fall-back grid (read and cleared for the next step); then the grid update.  Writes the
// velocity (vout) and, when pout is given, the momentum/mass.
// Only blocks that a chunk tile reaches (cov, fixed for a checkpoint segment) or that a
// fall-back scatter marked this step (fbflag, cleared here) are visited.  The others keep stale
// values that no kernel reads: gathers read nodes inside a chunk's tile or inside the stencil of
// a particle that scattered into them in the same step.
template <int NPART>
__global__ void __launch_bounds__(64 * NPART) k_grid_reduce(const float4* __restrict__ staging,
                                                            const int* __restrict__ bnch, const int* __restrict__ cstart,
                                                            const int* __restrict__ cov, double* __restrict__ fallback,
                                                            int* __restrict__ fbflag, float4* __restrict__ pout,
                                                            float4* __restrict__ vout, DevScene sc, int NB,
                                                            int NR)
{
    const int fl = fbflag[blockIdx.x];
    if (!cov[blockIdx.x] && !fl) return;
    const int n = sc.n_grid;
    const long nn = (long)n * n * n;
    const int nr3 = NR * NR * NR;
    const int inst = blockIdx.x / nr3, blk = blockIdx.x % nr3;
    const int rx = blk / (NR * NR), ry = (blk / NR) % NR, rz = blk % NR;
    const int node = threadIdx.x & 63;
    const int lx = node >> 4, ly = (node >> 2) & 3, lz = node & 3;
    const int gx = rx * TB - 1 + lx, gy = ry * TB - 1 + ly, gz = rz * TB - 1 + lz;
    __shared__ int s_nc[24], s_cs[24];
    load_covering_buckets(s_nc, s_cs, bnch, cstart, inst, rx, ry, rz, NB);
    __syncthreads();
    const bool inb = gx >= 0 && gy >= 0 && gz >= 0 && gx < n && gy < n && gz < n;
    const long gid = (long)inst * nn + (inb ? node_index(n, gx, gy, gz) : 0);
    if (fl && threadIdx.x == 0) fbflag[blockIdx.x] = 0;
    const float4 acc = staged_node_sum<NPART>(staging, s_nc, s_cs, fallback, fl, gid, inb, lx, ly, lz, 4);
    if (!inb || threadIdx.x >= 64) return;
    if (pout) pout[gid] = acc;
    float v[3];
    NodeRec r;
    gridop_node(sc, acc, gx, gy, gz, v, r);
    vout[gid] = make_float4(v[0], v[1], v[2], 0.0f);
}

// kFactor (the recompute of the adjoint): the full FP64 SVD, storing J and its SVD backward
// factor M for the adjoint P2G of the next step in mj ([10][P]: M row-major, then J).
// split: workgroups per chunk; part h of a chunk takes its particles [h, h + 1) * WG * PPT, so a
// table of 512-particle chunks runs as 256-particle workgroups (each loads the chunk's tile).
template <int PPT, bool kFactor>
__global__ void __launch_bounds__(WG) k_g2p_tiled(const float* __restrict__ in, float* __restrict__ out,
                                                  const GChunk* __restrict__ chunks, const int* __restrict__ nchunks_p, int maxchunks,
                                                  const float4* __restrict__ vgrid, const float* __restrict__ traj,
                                                  int T, int f, double* __restrict__ dvsum, double* __restrict__ jout,
                                                  double* __restrict__ mj, DevScene sc, int P,
                                                  unsigned* __restrict__ flag, int split)
{
    // chunk count of the checkpoint, written by the bucket sort (bounded by the table size)
    const int nchunks = min(*nchunks_p, maxchunks);
    __shared__ float tv[3 * TILE];
    __shared__ double red[WAVES][3];
    const int n = sc.n_grid;
    const long nn = (long)n * n * n;
    const long LP = P;
    for (int w = blockIdx.x; w < nchunks * split; w += gridDim.x) {
        const int c = w / split, base = (w % split) * (WG * PPT);
        const GChunk ch = chunks[c];
        if (base >= ch.count) continue;   // uniform across the workgroup
        const int x0 = ch.ox - 1, y0 = ch.oy - 1, z0 = ch.oz - 1;
        const float4* vi = vgrid + (long)ch.inst * nn;
        for (int e = threadIdx.x; e < TNODES; e += WG) {
            const int lx = e / (TS * TS), ly = (e / TS) % TS, lz = e % TS;
            const int gx = x0 + lx, gy = y0 + ly, gz = z0 + lz;
            float4 v = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            if (gx >= 0 && gy >= 0 && gz >= 0 && gx < n && gy < n && gz < n) v = vi[node_index(n, gx, gy, gz)];
            const int id = tile_id(lx, ly, lz);
            tv[id] = v.x;
            tv[TILE + id] = v.y;
            tv[2 * TILE + id] = v.z;
        }
        __syncthreads();
        double dv[3] = {0.0, 0.0, 0.0};
        const float* tr = traj + ((long)ch.inst * T + f) * 3;
#pragma unroll 1
        for (int r = 0; r < PPT; r++) {
            const int idx = base + r * WG + threadIdx.x;
            if (idx >= ch.count) break;
            const int p = ch.start + idx;
            float x[3], Fo[9];
#pragma unroll
            for (int a = 0; a < 3; a++) x[a] = in[(SX + a) * LP + p];
            // F of the next step, written to out by this step's P2G (read before out is written)
            const bool want_j = jout && ch.species == LIQUID;
            if (want_j) {
#pragma unroll
                for (int a = 0; a < 9; a++) Fo[a] = out[(SF + a) * LP + p];
            }
            Stencil s;
            if (!stencil(sc, x[0], x[1], x[2], s)) atomicOr(flag, 1u);
            float nv[3], nC[9];
            if (stencil_in_tile(s, x0, y0, z0, sc.tile_win)) {
                TileV g{tv, x0, y0, z0};
                g2p_gather(sc, s, g, nv, nC);
            } else {
                GatherVout g{vi, n};
                g2p_gather(sc, s, g, nv, nC);
            }
            if (ch.species == LIQUID) {
#pragma unroll
                for (int a = 0; a < 3; a++) {
                    out[(SV + a) * LP + p] = nv[a];
                    out[(SX + a) * LP + p] = x[a] + sc.dt * nv[a];
                }
                if (want_j) {
                    // J of the next step's P2G: F_tmp = (I + dt C') F'
                    float Fn[9];
                    f_tmp(sc, nC, Fo, Fn);
                    if (kFactor) {
                        double M[9];
                        const double J = liquid_j_factor(Fn, M);
#pragma unroll
                        for (int a = 0; a < 9; a++) mj[a * LP + p] = M[a];
                        mj[9 * LP + p] = J;
                        jout[p] = J;
                    } else {
                        jout[p] = liquid_j(Fn);
                    }
                }
            } else if (ch.species == STICK) {
#pragma unroll
                for (int a = 0; a < 3; a++) {
                    const float vk = (tr[3 + a] - tr[a]) / sc.dt;
                    dv[a] += (double)(vk - nv[a]);
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
        if (ch.species == STICK && dvsum) {
            const double s0 = wave_sum(dv[0]), s1 = wave_sum(dv[1]), s2 = wave_sum(dv[2]);
            const int w = threadIdx.x >> 6;
            if ((threadIdx.x & 63) == 0) {
                red[w][0] = s0;
                red[w][1] = s1;
                red[w][2] = s2;
            }
            __syncthreads();
            if (threadIdx.x < 3) {
                double acc2 = 0.0;
                for (int k = 0; k < WAVES; k++) acc2 += red[k][threadIdx.x];
                atomicAdd(dvsum + ((long)ch.inst * T + f + 1) * 3 + threadIdx.x, exact_grain(acc2));
            }
        }
        __syncthreads();
    }
}

// J of every liquid particle of a state (F_tmp = (I + dt C) F, values-only FP64 SVD), for the
// first P2G after a checkpoint; later steps get J from the previous step's G2P.  kFactor: the
// full SVD, also storing J and M in mj as k_g2p_tiled does.
template <int PPT, bool kFactor>
__global__ void __launch_bounds__(WG) k_liquid_j(const float* __restrict__ st, const GChunk* __restrict__ chunks,
                                                 const int* __restrict__ nchunks_p, int maxchunks, double* __restrict__ jout, double* __restrict__ mj,
                                                 DevScene sc, int P)
{
    // chunk count of the checkpoint, written by the bucket sort (bounded by the table size)
    const int nchunks = min(*nchunks_p, maxchunks);
    const long LP = P;
    for (int c = blockIdx.x; c < nchunks; c += gridDim.x) {
        const GChunk ch = chunks[c];
        if (ch.species != LIQUID) continue;
#pragma unroll 1
        for (int r = 0; r < PPT; r++) {
            const int idx = r * WG + threadIdx.x;
            if (idx >= ch.count) break;
            const int p = ch.start + idx;
            float C[9], F[9], Fn[9];
#pragma unroll
            for (int a = 0; a < 9; a++) {
                C[a] = st[(SC + a) * LP + p];
                F[a] = st[(SF + a) * LP + p];
            }
            f_tmp(sc, C, F, Fn);
            if (kFactor) {
                double M[9];
                const double J = liquid_j_factor(Fn, M);
#pragma unroll
                for (int a = 0; a < 9; a++) mj[a * LP + p] = M[a];
                mj[9 * LP + p] = J;
                jout[p] = J;
            } else {
                jout[p] = liquid_j(Fn);
            }
        }
    }
}

// The forward's G2P of step f fused with the P2G of step f + 1, on the row chunks of one sort (no
// re-sort between the two steps).  Each lane takes one particle of a row: the gather from the
// chunk's velocity tile and the particle update of step f (as k_g2p_tiled), then, with the updated
// particle still in registers, F_tmp, J (FP64 singular values), the stress and the scatter of step
// f + 1 into the wavefront's tile (as k_p2g_tiled).  One launch instead of two per substep, and the
// particle state is not written by one kernel and read back by the next; G2P then runs on the
// scatter's rows (one particle per occupied cell) instead of full gather chunks.
//   in:   state of step f (x is read)          fin:  F of step f + 1 (written by the previous P2G)
//   out:  x, v, C of step f + 1                fout: F of step f + 2
// In the forward in, fin, out and fout are one buffer (each thread reads its particle before it
// writes it), so they are not restrict-qualified.  The claim table holds only the bases a tile
// takes (node keys up to 365) and the velocity tile has dense planes, so four workgroups fit the
// LDS of a CU.
constexpr int OWN_KEYS = 368;   // node_key of a base in [0, 5]^3 is at most (5 * 8 + 5) * 8 + 5 = 365
using PlanesD = Planes<TS * TS, TS>;

template <int PPT>
__global__ void __launch_bounds__(WG, 4) k_g2p_p2g_tiled(const float* in, const float* fin, float* out, float* fout,
                                                         const Chunk* __restrict__ chunks,
                                                         const int* __restrict__ rows, const int* __restrict__ nchunks_p,
                                                         int maxchunks, const float4* __restrict__ vgrid,
                                                         const float* __restrict__ traj, int T, int f,
                                                         double* __restrict__ dvsum, const double* __restrict__ params,
                                                         float4* __restrict__ staging, double* __restrict__ fallback,
                                                         int* __restrict__ fbflag, int NR, DevScene sc, int P,
                                                         unsigned* __restrict__ flag)
{
    // chunk count of the checkpoint, written by the bucket sort (bounded by the table size)
    const int nchunks = min(*nchunks_p, maxchunks);
    __shared__ float4 wt[WAVES * TNODES];
    __shared__ unsigned char own[WAVES * OWN_KEYS];
    __shared__ float tv[3 * PlanesD::size];
    __shared__ double red[WAVES][3];
    const int n = sc.n_grid;
    const long nn = (long)n * n * n;
    const long LP = P;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    unsigned char* ow = own + wave * OWN_KEYS;
    for (int c = blockIdx.x; c < nchunks; c += gridDim.x) {
        const Chunk ch = chunks[c];
        const int x0 = ch.ox - 1, y0 = ch.oy - 1, z0 = ch.oz - 1;
        const bool liq = ch.species == LIQUID;
        const float4* vi = vgrid + (long)ch.inst * nn;
        // the first row's particles are loaded before the tiles are set up, and each next row's
        // while the current one scatters
        int p = 0;
        bool valid = scatter_slot(ch, rows, 0, wave, lane, p);
        float xs[3], Fs[9];
        auto load_xf = [&](int pp) {
#pragma unroll
            for (int a = 0; a < 3; a++) xs[a] = in[(SX + a) * LP + pp];
#pragma unroll
            for (int a = 0; a < 9; a++)
                Fs[a] = liq ? fin[(SF + a) * LP + pp] : ((a == 0 || a == 4 || a == 8) ? 1.0f : 0.0f);
        };
        if (valid) load_xf(p);
        for (int e = threadIdx.x; e < TNODES; e += WG) {
            const int lx = e / (TS * TS), ly = (e / TS) % TS, lz = e % TS;
            const int gx = x0 + lx, gy = y0 + ly, gz = z0 + lz;
            float4 v = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            if (gx >= 0 && gy >= 0 && gz >= 0 && gx < n && gy < n && gz < n) v = vi[node_index(n, gx, gy, gz)];
            const int id = PlanesD::id(lx, ly, lz);
            tv[id] = v.x;
            tv[PlanesD::size + id] = v.y;
            tv[2 * PlanesD::size + id] = v.z;
        }
        zero_lds(reinterpret_cast<float*>(wt), WAVES * 4 * TNODES);
        const InstParams ip = load_params(sc, params, ch.inst);
        WaveScatter4 acc{wt + wave * TNODES, x0, y0, z0};
        const float* tr = traj + ((long)ch.inst * T + f) * 3;
        double dv[3] = {0.0, 0.0, 0.0};
        __syncthreads();   // velocity tile loaded, wave tiles cleared
#pragma unroll 1
        for (int r = 0; r < PPT; r++) {
            Stencil s;
            P2GPayload pl;
            bool it = false;
            if (valid) {
                // G2P of step f
                ParticleIn q;
                {
                    Stencil sg;
                    if (!stencil(sc, xs[0], xs[1], xs[2], sg)) atomicOr(flag, 1u);
                    float nv[3];
                    if (stencil_in_tile(sg, x0, y0, z0, sc.tile_win)) {
                        TileVT<PlanesD> g{tv, x0, y0, z0};
                        g2p_gather(sc, sg, g, nv, q.C);
                    } else {
                        GatherVout g{vi, n};
                        g2p_gather(sc, sg, g, nv, q.C);
                    }
                    if (liq) {
#pragma unroll
                        for (int a = 0; a < 3; a++) {
                            q.v[a] = nv[a];
                            q.x[a] = xs[a] + sc.dt * nv[a];
                        }
                    } else if (ch.species == STICK) {
#pragma unroll
                        for (int a = 0; a < 3; a++) {
                            const float vk = (tr[3 + a] - tr[a]) / sc.dt;
                            dv[a] += (double)(vk - nv[a]);
                            q.v[a] = vk;
                            q.x[a] = xs[a] + vk * sc.dt;
                        }
                    } else {
#pragma unroll
                        for (int a = 0; a < 3; a++) {
                            q.v[a] = 0.0f;
                            q.x[a] = xs[a];
                        }
                    }
#pragma unroll
                    for (int a = 0; a < 3; a++) {
                        out[(SX + a) * LP + p] = q.x[a];
                        out[(SV + a) * LP + p] = q.v[a];
                    }
#pragma unroll
                    for (int a = 0; a < 9; a++) out[(SC + a) * LP + p] = q.C[a];
#pragma unroll
                    for (int a = 0; a < 9; a++) q.F[a] = Fs[a];
                }
                // P2G of step f + 1
                if (!stencil(sc, q.x[0], q.x[1], q.x[2], s)) atomicOr(flag, 1u);
                double J = 1.0;
                if (liq) {
                    float Ft[9];
                    f_tmp(sc, q.C, q.F, Ft);
                    J = liquid_j(Ft);
                }
                float Fn[9];
                p2g_prepare(sc, ch.species, ip, q, Fn, pl, &J);
                if (liq) {
#pragma unroll
                    for (int a = 0; a < 9; a++) fout[(SF + a) * LP + p] = Fn[a];
                }
                it = stencil_in_tile(s, x0, y0, z0, sc.tile_win);
                if (!it) {
                    mark_fallback_blocks(fbflag, ch.inst, s, NR);
                    if (sc.count_fallback) atomicAdd(reinterpret_cast<unsigned long long*>(flag + 2), 1ull);
                    ScatterExact4 g{fallback + 4 * (long)ch.inst * nn, n, fallback_magic(p2g_bound(sc, pl))};
                    p2g_scatter(sc, s, pl, g);
                }
            }
            int pn = 0;
            const bool vn = r + 1 < PPT && scatter_slot(ch, rows, r + 1, wave, lane, pn);
            if (vn) load_xf(pn);
            bool pending = valid && it;
            const int key = pending ? node_key(s, x0, y0, z0) : 0;
            while (__ballot(pending)) {
                lds_order();
                if (pending) ow[key] = (unsigned char)lane;
                lds_order();
                if (pending && ow[key] == lane) {
                    p2g_scatter(sc, opaque(s), pl, acc);
                    pending = false;
                }
            }
            valid = vn;
            p = pn;
        }
        if (ch.species == STICK && dvsum) {
            const double s0 = wave_sum(dv[0]), s1 = wave_sum(dv[1]), s2 = wave_sum(dv[2]);
            if (lane == 0) {
                red[wave][0] = s0;
                red[wave][1] = s1;
                red[wave][2] = s2;
            }
            __syncthreads();
            if (threadIdx.x < 3) {
                double acc2 = 0.0;
                for (int k = 0; k < WAVES; k++) acc2 += red[k][threadIdx.x];
                atomicAdd(dvsum + ((long)ch.inst * T + f + 1) * 3 + threadIdx.x, exact_grain(acc2));
            }
        }
        __syncthreads();
        store_staging4(wt, staging + (long)c * TNODES);
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------- adjoint

// Velocity tile (layout L) from the momentum/mass grid: the grid update at every node of the tile.
template <class L>
__device__ __forceinline__ void load_velocity_tile(float* tv, const float4* __restrict__ gi, const DevScene& sc,
                                                   int x0, int y0, int z0)
{
    const int n = sc.n_grid;
    for (int e = threadIdx.x; e < TNODES; e += WG) {
        const int lx = e / (TS * TS), ly = (e / TS) % TS, lz = e % TS;
        const int gx = x0 + lx, gy = y0 + ly, gz = z0 + lz;
        float v[3] = {0.0f, 0.0f, 0.0f};
        if (gx >= 0 && gy >= 0 && gz >= 0 && gx < n && gy < n && gz < n) {
            NodeRec r;
            gridop_node(sc, gi[node_index(n, gx, gy, gz)], gx, gy, gz, v, r);
        }
        const int id = L::id(lx, ly, lz);
        tv[id] = v[0];
        tv[L::size + id] = v[1];
        tv[2 * L::size + id] = v[2];
    }
}

template <int PPT, class CH>
__global__ void __launch_bounds__(WG) k_g2p_adj_tiled(const float* __restrict__ st, const float* __restrict__ adj,
                                                      const CH* __restrict__ chunks, const int* __restrict__ rows,
                                                      const int* __restrict__ nchunks_p, int maxchunks, const double* __restrict__ dvbar, int T, int f,
                                                      float4* __restrict__ staging, double* __restrict__ fallback,
                                                      int* __restrict__ fbflag, int NR, DevScene sc, int P)
{
    // chunk count of the checkpoint, written by the bucket sort (bounded by the table size)
    const int nchunks = min(*nchunks_p, maxchunks);
    __shared__ __align__(16) float wt[WAVES * 3 * TILE];
    __shared__ unsigned char own[WAVES * TNODES];
    const int n = sc.n_grid;
    const long nn = (long)n * n * n;
    const long LP = P;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    unsigned char* ow = own + wave * TNODES;
    for (int c = blockIdx.x; c < nchunks; c += gridDim.x) {
        const CH ch = chunks[c];
        const int x0 = ch.ox - 1, y0 = ch.oy - 1, z0 = ch.oz - 1;
        const bool liq = ch.species == LIQUID;
        float vstick[3] = {0.0f, 0.0f, 0.0f};
        if (ch.species == STICK) {
            const double* d = dvbar + ((long)ch.inst * T + f + 1) * 3;
#pragma unroll
            for (int a = 0; a < 3; a++) vstick[a] = -(float)d[a];
        }
        // the first row's particles are loaded before the tile is cleared, and each next row's
        // while the current one scatters
        auto load = [&](int q, float x[3], float Cb[9], float vt[3]) {
#pragma unroll
            for (int a = 0; a < 3; a++) x[a] = st[(SX + a) * LP + q];
#pragma unroll
            for (int a = 0; a < 9; a++) Cb[a] = adj[(SC + a) * LP + q];
#pragma unroll
            for (int a = 0; a < 3; a++) vt[a] = liq ? adj[(SV + a) * LP + q] + sc.dt * adj[(SX + a) * LP + q] : vstick[a];
        };
        int p = 0;
        bool valid = scatter_slot(ch, rows, 0, wave, lane, p);
        float x[3], Cb[9], vt[3];
        if (valid) load(p, x, Cb, vt);
        zero_lds(wt, WAVES * 3 * TILE);
        WaveScatter21 acc{wt + wave * 3 * TILE, x0, y0, z0};
#pragma unroll 1
        for (int r = 0; r < PPT; r++) {
            Stencil s;
            bool it = false;
            float Cc[9], vc[3];
            if (valid) {
                stencil(sc, x[0], x[1], x[2], s);
                it = stencil_in_tile(s, x0, y0, z0, sc.tile_win);
#pragma unroll
                for (int a = 0; a < 9; a++) Cc[a] = Cb[a];
#pragma unroll
                for (int a = 0; a < 3; a++) vc[a] = vt[a];
                if (!it) {
                    mark_fallback_blocks(fbflag, ch.inst, s, NR);
                    ScatterExact3 g{fallback + 4 * (long)ch.inst * nn, n, fallback_magic(g2p_adj_bound(sc, vc, Cc))};
                    g2p_adj_scatter(sc, s, vc, Cc, g);
                }
            }
            int pn = 0;
            const bool vn = r + 1 < PPT && scatter_slot(ch, rows, r + 1, wave, lane, pn);
            if (vn) load(pn, x, Cb, vt);
            if (r == 0) __syncthreads();   // every wavefront's tile is cleared
            bool pending = valid && it;
            const int key = pending ? node_key(s, x0, y0, z0) : 0;
            while (__ballot(pending)) {
                lds_order();
                if (pending) ow[key] = (unsigned char)lane;
                lds_order();
                if (pending && ow[key] == lane) {
                    g2p_adj_scatter(sc, opaque(s), vc, Cc, acc);
                    pending = false;
                }
            }
            valid = vn;
        }
        __syncthreads();
        store_staging21(wt, staging + (long)c * TNODES);
        __syncthreads();
    }
}

// One thread per grid node: dL/dv at the node (staging slots of the covering chunks plus the
// fall-back grid, cleared for the next step), then the adjoint of the grid update with the
// node's momentum/mass -> dL/d(momentum, mass) into pbar.
template <int NPART>
__global__ void __launch_bounds__(64 * NPART, 3) k_vbar_reduce(const float4* __restrict__ staging,
                                                            const int* __restrict__ bnch, const int* __restrict__ cstart,
                                                            const int* __restrict__ cov, double* __restrict__ fallback,
                                                            int* __restrict__ fbflag, const float4* __restrict__ grid,
                                                            float4* __restrict__ pbar, DevScene sc, int NB,
                                                            int NR)
{
    const int fl = fbflag[blockIdx.x];
    if (!cov[blockIdx.x] && !fl) return;
    const int n = sc.n_grid;
    const long nn = (long)n * n * n;
    const int nr3 = NR * NR * NR;
    const int inst = blockIdx.x / nr3, blk = blockIdx.x % nr3;
    const int rx = blk / (NR * NR), ry = (blk / NR) % NR, rz = blk % NR;
    const int node = threadIdx.x & 63;
    const int lx = node >> 4, ly = (node >> 2) & 3, lz = node & 3;
    const int gx = rx * TB - 1 + lx, gy = ry * TB - 1 + ly, gz = rz * TB - 1 + lz;
    __shared__ int s_nc[24], s_cs[24];
    load_covering_buckets(s_nc, s_cs, bnch, cstart, inst, rx, ry, rz, NB);
    __syncthreads();
    const bool inb = gx >= 0 && gy >= 0 && gz >= 0 && gx < n && gy < n && gz < n;
    const long gid = (long)inst * nn + (inb ? node_index(n, gx, gy, gz) : 0);
    if (fl && threadIdx.x == 0) fbflag[blockIdx.x] = 0;
    const float4 acc = staged_node_sum<NPART>(staging, s_nc, s_cs, fallback, fl, gid, inb, lx, ly, lz, 3);
    if (!inb || threadIdx.x >= 64) return;
    float v[3];
    NodeRec r;
    gridop_node(sc, grid[gid], gx, gy, gz, v, r);
    pbar[gid] = gridop_node_adj(r, acc.x, acc.y, acc.z);
}

// The stress half of the adjoint P2G of liquid particle p (xp, ub, Ab, fxb gathered): with the
// J and SVD backward factor stored by the recompute pass, the adjoint of the particle's state
// written in place of the next step's, and the parameter gradient added to tb.
__device__ __forceinline__ void p2g_adj_stress_at(const DevScene& sc, const InstParams& ip,
                                                  const float* __restrict__ st, float* __restrict__ adj,
                                                  const double M[9], double J, long LP, int p, const float xp[3],
                                                  const float ub[3], const float Ab[9], const float fxb[3],
                                                  double tb[4])
{
    ParticleIn q;
#pragma unroll
    for (int a = 0; a < 9; a++) {
        q.C[a] = st[(SC + a) * LP + p];
        q.F[a] = st[(SF + a) * LP + p];
    }
    float Fnb[9], xo[3], vo[3], Co[9], Fo[9];
#pragma unroll
    for (int a = 0; a < 9; a++) Fnb[a] = adj[(SF + a) * LP + p];
    LiquidP2G L;
    liquid_stress(sc, ip.th, ip.kb, q.C, J, L);
    p2g_adj_stress(sc, ip, q, L, M, ub, Ab, fxb, Fnb, xp, xo, vo, Co, Fo, tb);
#pragma unroll
    for (int a = 0; a < 3; a++) adj[(SX + a) * LP + p] = xo[a];
#pragma unroll
    for (int a = 0; a < 3; a++) adj[(SV + a) * LP + p] = vo[a];
#pragma unroll
    for (int a = 0; a < 9; a++) adj[(SC + a) * LP + p] = Co[a];
#pragma unroll
    for (int a = 0; a < 9; a++) adj[(SF + a) * LP + p] = Fo[a];
}

// Workgroup sum of the four FP64 parameter-gradient partials into the chunk's slot of thpart.
__device__ __forceinline__ void add_chunk_partials(double (*red)[4], const double tb[4], double* __restrict__ thpart,
                                                   int c)
{
    const double r0 = wave_sum(tb[0]), r1 = wave_sum(tb[1]), r2 = wave_sum(tb[2]), r3 = wave_sum(tb[3]);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        red[w][0] = r0;
        red[w][1] = r1;
        red[w][2] = r2;
        red[w][3] = r3;
    }
    __syncthreads();
    if (threadIdx.x < 4) {
        double a2 = 0.0;
        for (int k = 0; k < WAVES; k++) a2 += red[k][threadIdx.x];
        thpart[4 * (long)c + threadIdx.x] += a2;
    }
}

// Adjoint P2G, gather half: for liquid particles the G2P part of dL/dx (xp, through the gather
// of this step's velocity) and the gather of dL/d(momentum, mass) with the stored affine (ub, Ab,
// fxb); for stick and container particles their whole adjoint (dL/dC).  kFused: the stress half
// follows in the same thread (p2g_adj_stress_at, partials into thpart); otherwise the gathered
// quantities go to gout ([18][P]: xp, ub, Ab, fxb) for k_p2g_adj_stress.  kXp: the position
// gradient through G2P's gather was computed by k_p2g_adj_xgrad into xpin ([3][P]), so neither it
// nor the velocity tile is computed here.
template <int PPT, bool kFused, bool kXp>
__global__ void __launch_bounds__(WG) k_p2g_adj_tiled(const float* __restrict__ st, const float* __restrict__ ain,
                                                      float* __restrict__ adj, const GChunk* __restrict__ chunks,
                                                      const int* __restrict__ nchunks_p, int maxchunks, const float4* __restrict__ grid,
                                                      const float4* __restrict__ pbar, float* __restrict__ gout,
                                                      const double* __restrict__ params, const double* __restrict__ mj,
                                                      double* __restrict__ thpart, const float* __restrict__ xpin,
                                                      DevScene sc, int P)
{
    // chunk count of the checkpoint, written by the bucket sort (bounded by the table size)
    const int nchunks = min(*nchunks_p, maxchunks);
    __shared__ float4 tq[TNODES];
    __shared__ float tv[3 * PlanesG::size];
    __shared__ double red[WAVES][4];
    const int n = sc.n_grid;
    const long nn = (long)n * n * n;
    const long LP = P;
    for (int c = blockIdx.x; c < nchunks; c += gridDim.x) {
        const GChunk ch = chunks[c];
        const int x0 = ch.ox - 1, y0 = ch.oy - 1, z0 = ch.oz - 1;
        const float4* qi = pbar + (long)ch.inst * nn;
        for (int e = threadIdx.x; e < TNODES; e += WG) {
            const int lx = e / (TS * TS), ly = (e / TS) % TS, lz = e % TS;
            const int gx = x0 + lx, gy = y0 + ly, gz = z0 + lz;
            float4 v = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            if (gx >= 0 && gy >= 0 && gz >= 0 && gx < n && gy < n && gz < n) v = qi[node_index(n, gx, gy, gz)];
            tq[e] = v;
        }
        const float4* gi = grid + (long)ch.inst * nn;
        if (!kXp && ch.species == LIQUID) load_velocity_tile<PlanesG>(tv, gi, sc, x0, y0, z0);
        __syncthreads();
        double tb[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll 1
        for (int r = 0; r < PPT; r++) {
            const int idx = r * WG + threadIdx.x;
            if (idx >= ch.count) break;
            const int p = ch.start + idx;
            float x[3];
#pragma unroll
            for (int a = 0; a < 3; a++) x[a] = st[(SX + a) * LP + p];
            Stencil s;
            stencil(sc, x[0], x[1], x[2], s);
            const bool it = stencil_in_tile(s, x0, y0, z0, sc.tile_win);
            if (ch.species == LIQUID) {
                // the affine stored by the recompute pass and the velocity, loaded before the
                // position-gradient gather so that their latency overlaps it
                float A[9];
#pragma unroll
                for (int a = 0; a < 9; a++) A[a] = ain[a * LP + p];
                const float u[3] = {st[SV * LP + p], st[(SV + 1) * LP + p] + sc.gdt, st[(SV + 2) * LP + p]};
                // with the position gradient computed here as well and two particles per thread
                // or more (two wavefronts per SIMD either way) the SVD backward factor and J are
                // loaded this early too; otherwise they are loaded where they are used, which keeps
                // the other forms at their occupancy (three wavefronts per SIMD with the gradient on
                // its own stream, two with one particle per thread)
                constexpr bool kEarlyMJ = kFused && !kXp && PPT > 1;
                double M[9], J = 1.0;
                auto load_mj = [&]() {
#pragma unroll
                    for (int a = 0; a < 9; a++) M[a] = mj[a * LP + p];
                    J = mj[9 * LP + p];
                };
                if (kEarlyMJ) load_mj();
                float xp[3];
                if (kXp) {
#pragma unroll
                    for (int a = 0; a < 3; a++) xp[a] = xpin[a * LP + p];
                    if (!kFused) {
#pragma unroll
                        for (int a = 0; a < 3; a++) gout[a * LP + p] = xp[a];
                    }
                } else {
                    // the G2P part of dL/dx (its weights and offsets depend on x): dL/dx of the
                    // next step plus the gradient through the gather of this step's velocity
                    float xa[3], vt[3], Cb[9];
#pragma unroll
                    for (int a = 0; a < 3; a++) {
                        xa[a] = adj[(SX + a) * LP + p];
                        vt[a] = adj[(SV + a) * LP + p] + sc.dt * xa[a];
                    }
#pragma unroll
                    for (int a = 0; a < 9; a++) Cb[a] = adj[(SC + a) * LP + p];
                    if (it) {
                        TileVT<PlanesG> gv{tv, x0, y0, z0};
                        g2p_adj_xgrad(sc, s, vt, Cb, gv, xp);
                    } else {
                        GatherVfromP gv{gi, sc};
                        g2p_adj_xgrad<true>(sc, s, vt, Cb, gv, xp);
                    }
#pragma unroll
                    for (int a = 0; a < 3; a++) xp[a] += xa[a];
                    if (!kFused) {
#pragma unroll
                        for (int a = 0; a < 3; a++) gout[a * LP + p] = xp[a];
                    }
                }
                asm volatile("" ::: "memory");
                // the gather with the affine
                float ub[3], Ab[9], fxb[3];
                if (it) {
                    TileQ4 g{tq, x0, y0, z0};
                    p2g_adj_gather(sc, s, A, u, g, ub, Ab, fxb);
                } else {
                    GatherPbar g{qi, n};
                    p2g_adj_gather(sc, s, A, u, g, ub, Ab, fxb);
                }
                if (kFused) {
                    asm volatile("" ::: "memory");
                    if (!kEarlyMJ) load_mj();
                    p2g_adj_stress_at(sc, load_params(sc, params, ch.inst), st, adj, M, J, LP, p, xp, ub, Ab, fxb,
                                      tb);
                } else {
#pragma unroll
                    for (int a = 0; a < 3; a++) gout[(3 + a) * LP + p] = ub[a];
#pragma unroll
                    for (int a = 0; a < 9; a++) gout[(6 + a) * LP + p] = Ab[a];
#pragma unroll
                    for (int a = 0; a < 3; a++) gout[(15 + a) * LP + p] = fxb[a];
                }
            } else {
                float Co[9];
                if (it) {
                    TileQ4 g{tq, x0, y0, z0};
                    p2g_adj_rigid(sc, species_mass(sc, ch.species), s, g, Co);
                } else {
                    GatherPbar g{qi, n};
                    p2g_adj_rigid(sc, species_mass(sc, ch.species), s, g, Co);
                }
#pragma unroll
                for (int a = 0; a < 9; a++) adj[(SC + a) * LP + p] = Co[a];
            }
        }
        if (kFused && ch.species == LIQUID) add_chunk_partials(red, tb, thpart, c);
        __syncthreads();
    }
}

// The position gradient of the adjoint P2G on its own, for liquid particles: dL/dx of the next
// step plus the gradient through G2P's gather of this step's velocity (weights and offsets
// depend on x), into xp ([3][P]).  It needs neither dL/d(momentum, mass) nor the stress, so it
// runs beside the adjoint G2P and grid reduction of the same step, on another stream.
template <int PPT>
__global__ void __launch_bounds__(WG) k_p2g_adj_xgrad(const float* __restrict__ st, const float* __restrict__ adj,
                                                      const GChunk* __restrict__ chunks, const int* __restrict__ nchunks_p,
                                                      int maxchunks, const float4* __restrict__ grid,
                                                      float* __restrict__ xp, DevScene sc, int P)
{
    // chunk count of the checkpoint, written by the bucket sort (bounded by the table size)
    const int nchunks = min(*nchunks_p, maxchunks);
    __shared__ float tv[3 * PlanesG::size];
    const long nn = (long)sc.n_grid * sc.n_grid * sc.n_grid;
    const long LP = P;
    for (int c = blockIdx.x; c < nchunks; c += gridDim.x) {
        const GChunk ch = chunks[c];
        if (ch.species != LIQUID) continue;   // uniform across the workgroup
        const int x0 = ch.ox - 1, y0 = ch.oy - 1, z0 = ch.oz - 1;
        const float4* gi = grid + (long)ch.inst * nn;
        load_velocity_tile<PlanesG>(tv, gi, sc, x0, y0, z0);
        __syncthreads();
#pragma unroll 1
        for (int r = 0; r < PPT; r++) {
            const int idx = r * WG + threadIdx.x;
            if (idx >= ch.count) break;
            const int p = ch.start + idx;
            float x[3], xa[3], vt[3], Cb[9], g[3];
#pragma unroll
            for (int a = 0; a < 3; a++) x[a] = st[(SX + a) * LP + p];
            Stencil s;
            stencil(sc, x[0], x[1], x[2], s);
#pragma unroll
            for (int a = 0; a < 3; a++) {
                xa[a] = adj[(SX + a) * LP + p];
                vt[a] = adj[(SV + a) * LP + p] + sc.dt * xa[a];
            }
#pragma unroll
            for (int a = 0; a < 9; a++) Cb[a] = adj[(SC + a) * LP + p];
            if (stencil_in_tile(s, x0, y0, z0, sc.tile_win)) {
                TileVT<PlanesG> gv{tv, x0, y0, z0};
                g2p_adj_xgrad(sc, s, vt, Cb, gv, g);
            } else {
                GatherVfromP gv{gi, sc};
                g2p_adj_xgrad<true>(sc, s, vt, Cb, gv, g);
            }
#pragma unroll
            for (int a = 0; a < 3; a++) xp[a * LP + p] = g[a] + xa[a];
        }
        __syncthreads();
    }
}

// Adjoint P2G, stress half, one thread per liquid particle: from the gathered quantities and
// the J and SVD backward factor stored by the recompute pass, the adjoint of the particle's
// state (written in place of the next step's) and the FP64 parameter gradient (wave, then
// workgroup sums into the chunk's slot of thpart).
template <int PPT>
__global__ void __launch_bounds__(WG) k_p2g_adj_stress(const float* __restrict__ st, float* __restrict__ adj,
                                                       const GChunk* __restrict__ chunks, const int* __restrict__ nchunks_p, int maxchunks,
                                                       const double* __restrict__ params,
                                                       const float* __restrict__ gin, const double* __restrict__ mj,
                                                       double* __restrict__ thpart, DevScene sc, int P)
{
    // chunk count of the checkpoint, written by the bucket sort (bounded by the table size)
    const int nchunks = min(*nchunks_p, maxchunks);
    __shared__ double red[WAVES][4];
    const long LP = P;
    for (int c = blockIdx.x; c < nchunks; c += gridDim.x) {
        const GChunk ch = chunks[c];
        if (ch.species != LIQUID) continue;
        double tb[4] = {0.0, 0.0, 0.0, 0.0};
        const InstParams ip = load_params(sc, params, ch.inst);
#pragma unroll 1
        for (int r = 0; r < PPT; r++) {
            const int idx = r * WG + threadIdx.x;
            if (idx >= ch.count) break;
            const int p = ch.start + idx;
            float xp[3], ub[3], Ab[9], fxb[3];
#pragma unroll
            for (int a = 0; a < 3; a++) {
                xp[a] = gin[a * LP + p];
                ub[a] = gin[(3 + a) * LP + p];
                fxb[a] = gin[(15 + a) * LP + p];
            }
#pragma unroll
            for (int a = 0; a < 9; a++) Ab[a] = gin[(6 + a) * LP + p];
            double M[9];
#pragma unroll
            for (int a = 0; a < 9; a++) M[a] = mj[a * LP + p];
            p2g_adj_stress_at(sc, ip, st, adj, M, mj[9 * LP + p], LP, p, xp, ub, Ab, fxb, tb);
        }
        add_chunk_partials(red, tb, thpart, c);
        __syncthreads();
    }
}

// The per-chunk FP64 parameter-gradient partials of one sort's gather chunks, added into thbar:
// one workgroup per instance, in a fixed order (thread t takes the instance's chunks t, t + 256,
// ..., then the wave sums and the wavefronts in order), so the sum has the same bits in every run.
// An instance's chunks are contiguous in the table (bucket order, instance first); the workgroup
// finds them by binary search over the chunks' instance numbers.
__global__ void __launch_bounds__(256) k_thpart_reduce(const GChunk* __restrict__ chunks,
                                                       const int* __restrict__ nchunks_p, int maxchunks,
                                                       const double* __restrict__ thpart, double* __restrict__ thbar)
{
    __shared__ double red[WAVES][4];
    const int nchunks = min(*nchunks_p, maxchunks);
    const int inst = blockIdx.x;
    auto first_at_least = [&](int key) {
        int lo = 0, hi = nchunks;
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (chunks[mid].inst < key) lo = mid + 1;
            else hi = mid;
        }
        return lo;
    };
    const int c0 = first_at_least(inst), c1 = first_at_least(inst + 1);
    double v[4] = {0.0, 0.0, 0.0, 0.0};
    for (int c = c0 + threadIdx.x; c < c1; c += blockDim.x) {
#pragma unroll
        for (int k = 0; k < 4; k++) v[k] += thpart[4 * (long)c + k];
    }
    const int w = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const double r = wave_sum(v[k]);
        if ((threadIdx.x & 63) == 0) red[w][k] = r;
    }
    __syncthreads();
    if (threadIdx.x < 4) {
        double a = 0.0;
        for (int k = 0; k < WAVES; k++) a += red[k][threadIdx.x];
        thbar[4 * inst + threadIdx.x] += a;
    }
}

// ---------------------------------------------------------------------------- bucket sort

// Bucket of a particle: (instance, species, block) with the block from its stencil base cell.
__global__ void __launch_bounds__(256) k_sort_count(const float* __restrict__ st, const int* __restrict__ meta,
                                                    const Group* __restrict__ groups, int* __restrict__ key,
                                                    int* __restrict__ rank, int* __restrict__ cnt, DevScene sc, int NB,
                                                    int P)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P) return;
    const int g = meta[p];
    if (g < 0) {
        key[p] = -1;
        return;
    }
    const Group G = groups[g];
    const long LP = P;
    Stencil s;
    stencil(sc, st[(SX + 0) * LP + p], st[(SX + 1) * LP + p], st[(SX + 2) * LP + p], s);
    const int bx = s.b[0] / TB, by = s.b[1] / TB, bz = s.b[2] / TB;
    const int cib = (s.b[0] % TB) * TB * TB + (s.b[1] % TB) * TB + (s.b[2] % TB);
    const int bucket = ((G.inst * 3 + G.species) * NB + bx) * NB * NB + by * NB + bz;
    const int cell = bucket * 64 + cib;
    key[p] = cell;
    rank[p] = atomicAdd(&cnt[cell], 1);
}

__device__ __forceinline__ int wave_incl_scan(int v)
{
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int t = __shfl_up(v, o, 64);
        if (lane >= o) v += t;
    }
    return v;
}

// One wavefront per bucket: particle count, row count (the largest cell count), chunk count
// (rows_per_chunk rows each), gather-chunk count (gchunk particles each) and each cell's
// offset in the bucket.
__global__ void __launch_bounds__(256) k_bucket_totals(const int* __restrict__ cnt, int* __restrict__ cell_off,
                                                       int* __restrict__ bcount, int* __restrict__ brows,
                                                       int* __restrict__ bnch, int* __restrict__ bgch, int NBK,
                                                       int rows_per_chunk, int gchunk)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int bucket = (int)(t >> 6);
    if (bucket >= NBK) return;
    const int lane = threadIdx.x & 63;
    const int nc = cnt[(long)bucket * 64 + lane];
    const int incl = wave_incl_scan(nc);
    cell_off[(long)bucket * 64 + lane] = incl - nc;
    const int total = __shfl(incl, 63, 64);
    int mx = nc;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) mx = max(mx, __shfl_xor(mx, o, 64));
    if (lane == 0) {
        bcount[bucket] = total;
        brows[bucket] = mx;
        bnch[bucket] = (mx + rows_per_chunk - 1) / rows_per_chunk;
        bgch[bucket] = (total + gchunk - 1) / gchunk;
    }
}

// Exclusive scans of four int arrays a[k] into o[k] (tot[k] gets the sums), in three passes
// over tiles of 1024 elements: the tile sums (k_scan4_sums), a scan of the tile sums in place by
// one workgroup (k_scan4_tiles), and the scan within each tile plus its offset (k_scan4_apply).
struct Scan4 {
    const int* a[4];
    int* o[4];
};
constexpr int SCAN_TILE = 1024;

// Inclusive scan of one value per thread across a 1024-thread workgroup (ws: 17 ints of LDS);
// `total` gets the workgroup's sum.
__device__ __forceinline__ int block_incl_scan(int v, int* ws, int& total)
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int sc = wave_incl_scan(v);
    if (lane == 63) ws[w] = sc;
    __syncthreads();
    if (threadIdx.x < 64) {
        const int x = threadIdx.x < 16 ? ws[threadIdx.x] : 0;
        const int y = wave_incl_scan(x);
        if (threadIdx.x < 16) ws[threadIdx.x] = y - x;   // exclusive wave offsets
        if (threadIdx.x == 15) ws[16] = y;
    }
    __syncthreads();
    const int r = ws[w] + sc;
    total = ws[16];
    __syncthreads();
    return r;
}

__global__ void __launch_bounds__(SCAN_TILE) k_scan4_sums(Scan4 q, int n, int* __restrict__ tsum)
{
    __shared__ int ws[17];
    const long i = (long)blockIdx.x * SCAN_TILE + threadIdx.x;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const int v = i < n ? q.a[k][i] : 0;
        int total;
        block_incl_scan(v, ws, total);
        if (threadIdx.x == 0) tsum[(long)k * gridDim.x + blockIdx.x] = total;
    }
}

__global__ void __launch_bounds__(SCAN_TILE) k_scan4_tiles(int* __restrict__ tsum, int ntiles, int* __restrict__ tot)
{
    __shared__ int ws[17];
    for (int k = 0; k < 4; k++) {
        int carry = 0;
        for (int base = 0; base < ntiles; base += SCAN_TILE) {
            const int i = base + threadIdx.x;
            const int v = i < ntiles ? tsum[(long)k * ntiles + i] : 0;
            int total;
            const int r = block_incl_scan(v, ws, total);
            if (i < ntiles) tsum[(long)k * ntiles + i] = carry + r - v;
            carry += total;
        }
        if (threadIdx.x == 0) tot[k] = carry;
    }
}

__global__ void __launch_bounds__(SCAN_TILE) k_scan4_apply(Scan4 q, int n, const int* __restrict__ tsum)
{
    __shared__ int ws[17];
    const long i = (long)blockIdx.x * SCAN_TILE + threadIdx.x;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const int v = i < n ? q.a[k][i] : 0;
        int total;
        const int r = block_incl_scan(v, ws, total);
        if (i < n) q.o[k][i] = tsum[(long)k * gridDim.x + blockIdx.x] + r - v;
    }
}

// One wavefront per bucket: the round-robin slot of every (cell, rank) pair (slot = the
// particles of earlier rows + the occupied cells before this one in this row), the bucket's
// row table, its chunk entries (rows_per_chunk rows each) and its gather chunks (gchunk
// particles each).  rows[k] is a global slot; the entry after a bucket's last row is the next
// bucket's first slot, written by both.
__global__ void __launch_bounds__(256) k_bucket_layout(const int* __restrict__ cnt, const int* __restrict__ cell_off,
                                                       const int* __restrict__ bstart, const int* __restrict__ rstart,
                                                       const int* __restrict__ cstart, const int* __restrict__ gstart,
                                                       int* __restrict__ rrpos, int* __restrict__ rows,
                                                       Chunk* __restrict__ chunks, GChunk* __restrict__ gchunks,
                                                       int* __restrict__ cov, int NB, int NR, int NBK, int rows_per_chunk,
                                                       int gchunk)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int bucket = (int)(t >> 6);
    if (bucket >= NBK) return;
    const int lane = threadIdx.x & 63;
    const int nc = cnt[(long)bucket * 64 + lane];
    if (__ballot(nc > 0) == 0ull) return;
    int mx = nc;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) mx = max(mx, __shfl_xor(mx, o, 64));
    const int bs = bstart[bucket], rs = rstart[bucket], cs = cstart[bucket];
    const int nb3 = NB * NB * NB;
    const int inst = bucket / (3 * nb3), sp = (bucket / nb3) % 3, blk = bucket % nb3;
    const int bx = blk / (NB * NB), by = (blk / NB) % NB, bz = blk % NB;
    if (lane < 8) {
        // the reduction blocks this bucket's tiles reach: b and b + 1 per axis (the tile covers
        // nodes [4b - 1, 4b + 7))
        const int cx = bx + ((lane >> 2) & 1), cy = by + ((lane >> 1) & 1), cz = bz + (lane & 1);
        if (cx < NR && cy < NR && cz < NR) cov[(long)inst * NR * NR * NR + (cx * NR + cy) * NR + cz] = 1;
    }
    int total = nc;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) total += __shfl_xor(total, o, 64);
    const int ngc = (total + gchunk - 1) / gchunk, gs = gstart[bucket];
    for (int i = lane; i < ngc; i += 64) {
        GChunk g;
        g.inst = inst;
        g.species = sp;
        g.ox = bx * TB;
        g.oy = by * TB;
        g.oz = bz * TB;
        g.start = bs + i * gchunk;
        g.count = min(gchunk, total - i * gchunk);
        g.pad = 0;
        gchunks[gs + i] = g;
    }
    const int off = cell_off[(long)bucket * 64 + lane];
    const unsigned long long below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    int S = 0;
    for (int r = 0; r < mx; r++) {
        const bool act = nc > r;
        const unsigned long long m = __ballot(act);
        if (act) rrpos[bs + off + r] = S + __popcll(m & below);
        if (lane == 0) {
            rows[rs + r] = bs + S;
            if (r % rows_per_chunk == 0) {
                Chunk c;
                c.inst = inst;
                c.species = sp;
                c.ox = bx * TB;
                c.oy = by * TB;
                c.oz = bz * TB;
                c.start = bs + S;
                c.row0 = rs + r;
                c.nrows = min(rows_per_chunk, mx - r);
                chunks[cs + r / rows_per_chunk] = c;
            }
        }
        S += __popcll(m);
    }
    if (lane == 0) rows[rs + mx] = bs + S;
}

// Move every particle to its sorted slot; src[q] records where slot q came from.
// Particle p into the cell-major list of its bucket at its provisional rank (the order in which
// the counting atomics happened to run).
__global__ void __launch_bounds__(256) k_cell_list(const int* __restrict__ key, const int* __restrict__ rank,
                                                   const int* __restrict__ cell_off, const int* __restrict__ bstart,
                                                   int* __restrict__ clist, int P)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P) return;
    const int cell = key[p];
    if (cell < 0) return;
    clist[bstart[cell / 64] + cell_off[cell] + rank[p]] = p;
}

// The rank of particle p in its cell by original particle index, so that the sorted order (and
// with it every summation order downstream) does not depend on the order of the atomics.
__global__ void __launch_bounds__(256) k_rank_fix(const int* __restrict__ key, const int* __restrict__ cnt,
                                                  const int* __restrict__ cell_off, const int* __restrict__ bstart,
                                                  const int* __restrict__ clist, const int* __restrict__ pid,
                                                  int* __restrict__ rank, int P)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P) return;
    const int cell = key[p];
    if (cell < 0) return;
    const int* l = clist + bstart[cell / 64] + cell_off[cell];
    const int n = cnt[cell], me = pid[p];
    int r = 0;
    for (int j = 0; j < n; j++) r += pid[l[j]] < me ? 1 : 0;
    rank[p] = r;
}

__global__ void __launch_bounds__(256) k_sort_scatter(const float* __restrict__ in, float* __restrict__ out,
                                                      const int* __restrict__ key, const int* __restrict__ rank,
                                                      const int* __restrict__ cell_off, const int* __restrict__ bstart,
                                                      const int* __restrict__ rrpos, const int* __restrict__ meta_in,
                                                      int* __restrict__ meta_out, const int* __restrict__ pid_in,
                                                      int* __restrict__ pid_out, int* __restrict__ src, int P)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P) return;
    const int cell = key[p];
    if (cell < 0) return;
    const int bs = bstart[cell / 64];
    const int q = bs + rrpos[bs + cell_off[cell] + rank[p]];
    const long LP = P;
#pragma unroll
    for (int a = 0; a < NSLOT; a++) out[a * LP + q] = in[a * LP + p];
    meta_out[q] = meta_in[p];
    pid_out[q] = pid_in[p];
    src[q] = p;
}

// The bucket sort's totals against what the tables hold: every particle placed, chunk and row
// counts within their tables; otherwise flag bit 2 (reported by the host after the call).
__global__ void k_check_sort(const int* __restrict__ tot, int nreal, long maxch, long maxgch,
                             unsigned* __restrict__ flag)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        if (tot[0] != nreal || tot[1] < 0 || tot[1] > maxch || tot[2] > nreal || tot[3] < 0 || tot[3] > maxgch)
            atomicOr(flag, 2u);
    }
}

__global__ void __launch_bounds__(256) k_copy_i32(int* __restrict__ dst, const int* __restrict__ src, long n)
{
    const long stride = (long)gridDim.x * blockDim.x;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) dst[i] = src[i];
}

__global__ void __launch_bounds__(256) k_copy_f32(float* __restrict__ dst, const float* __restrict__ src, long n)
{
    const long stride = (long)gridDim.x * blockDim.x;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) dst[i] = src[i];
}

__global__ void __launch_bounds__(256) k_fill_i32(int* __restrict__ p, long n, int v)
{
    const long stride = (long)gridDim.x * blockDim.x;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = v;
}

// Adjoint carried across a re-sort: slot q of the later order came from slot src[q] of the
// earlier order.
// The state of a recomputed step into the order a sort of the forward gave it at that step:
// slot q of out takes slot src[q] of in (real particles only; padding slots keep stale values
// that no chunk reads).
__global__ void __launch_bounds__(256) k_permute_state(const float* __restrict__ in, float* __restrict__ out,
                                                       const int* __restrict__ src, int nreal, int P)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nreal) return;
    const int p = src[q];
    const long LP = P;
#pragma unroll
    for (int a = 0; a < NSLOT; a++) out[a * LP + q] = in[a * LP + p];
}

__global__ void __launch_bounds__(256) k_unsort_adj(const float* __restrict__ in, float* __restrict__ out,
                                                    const int* __restrict__ src, int nreal, int P)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nreal) return;
    const int p = src[q];
    const long LP = P;
#pragma unroll
    for (int a = 0; a < NSLOT; a++) out[a * LP + p] = in[a * LP + q];
}

}  // namespace ds
