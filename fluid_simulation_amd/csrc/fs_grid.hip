// fs_grid.hip -- Warp-compatible uniform hash grid for gfx950.
//
// Build = four kernels (five with a pre-op) on one stream, no host sync, no
// memset:
//   0. k_pre_op       (DFSPH/PBF) the solver's position update before the
//                     build, in original order, when the hash runs in the
//                     previous sorted order
//   1. k_hash_count   per particle, in the previous build's sorted order:
//                     bucket id, rank = atomicAdd(count), both stored at the
//                     processing position; alongside, in original order, the
//                     mass, rest density, flags and mask of each point packed
//                     into one record for the fix-up
//   2. k_bucket_scan  single-pass decoupled look-back exclusive scan of the
//                     dense bucket counts -> cell_start[dims+1]
//   3. k_scatter      slot = cell_start[b] + rank for the point at each
//                     processing position; re-zero the counter
//   4. k_fix_gather   order each bucket's members by original index (the
//                     stable (bucket, index) permutation of Warp's build),
//                     and gather particle data into that sorted order.
// The scan's look-back descriptors are re-zeroed by k_scatter, so every build
// (also one replayed from a captured graph) starts from a clean state.
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <map>
#include <string>
#include <vector>

#include "fs_internal.h"

namespace fs {

static thread_local char g_err[512];

void set_error(const char* fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

int32_t hip_check(hipError_t e, const char* what)
{
    if (e == hipSuccess) return FS_OK;
    set_error("%s: %s", what, hipGetErrorString(e));
    return FS_ERR_HIP;
}

bool g_profiling = false;

struct ProfRecord {
    const char* name;
    hipEvent_t ev[2];
};
static std::vector<ProfRecord> g_prof_live;
static std::vector<hipEvent_t> g_prof_pool;
static std::map<std::string, std::pair<double, long long>> g_prof_totals;

static hipEvent_t prof_event()
{
    if (!g_prof_pool.empty()) {
        hipEvent_t e = g_prof_pool.back();
        g_prof_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    (void)hipEventCreate(&e);
    return e;
}

void prof_mark(const char* name, hipStream_t s, bool begin)
{
    if (begin) {
        ProfRecord r{name, {prof_event(), prof_event()}};
        (void)hipEventRecord(r.ev[0], s);
        g_prof_live.push_back(r);
    } else if (!g_prof_live.empty()) {
        (void)hipEventRecord(g_prof_live.back().ev[1], s);
    }
}

static void prof_collect()
{
    for (auto& r : g_prof_live) {
        float ms = 0.0f;
        (void)hipEventSynchronize(r.ev[1]);
        (void)hipEventElapsedTime(&ms, r.ev[0], r.ev[1]);
        auto& t = g_prof_totals[r.name];
        t.first += ms;
        t.second += 1;
        g_prof_pool.push_back(r.ev[0]);
        g_prof_pool.push_back(r.ev[1]);
    }
    g_prof_live.clear();
}

constexpr int kScanBlock = 256;
constexpr int kScanItems = 32;
constexpr int kScanTile = kScanBlock * kScanItems;

static inline size_t align_up(size_t v) { return (v + 255) & ~(size_t)255; }

GridView grid_view(const fs_grid* g)
{
    GridView v;
    v.dx = g->dim_x;
    v.dy = g->dim_y;
    v.dz = g->dim_z;
    v.dxy = g->dim_x * g->dim_y;
    v.round_mode = g->round_mode;
    v.npts = g->num_points;
    v.base = 0;
    v.inv = g->cell_width_inv;
    v.start = g->cell_start;
    return v;
}

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------

struct HashArgs {
    GridView g;
    PreOp pre;
    int n;
    const int* order;   // processing order (the previous build's sorted order) or null
    int* cell_of;
    int* rank;
    int* count;
    GatherOp pack;      // with pack.packed set, also packs every point's fix-up words
};

// Position of point i after the solver's pre-op; writes the pre-op outputs.
template <int MODE>
__device__ __forceinline__ v3 pre_point(const PreOp& p, int i)
{
    v3 x = ld3(p.x_in, i);
    const float dt = p.dt_device ? *p.dt_device : p.dt;
    if (MODE == kPreIntegrate) {
        // integrate_positions  solver_dfsph.py:340-353
        bool dyn = (p.flags[i] & 1) != 0 && p.mask[i] == 0;
        if (dyn) x = vadd(x, muls(ld3(p.v_in, i), dt));
        st3(p.x_out, i, x);
    } else if (MODE == kPrePredict) {
        bool dyn = (p.flags[i] & 1) != 0 && p.mask[i] == 0;
        v3 v = ld3(p.v_in, i);
        if (dyn) {
            // predict_velocity_with_gravity  solver_pbf.py:192-208
            int w = p.world[i];
            w = w > 0 ? w : 0;
            v = vadd(v, smul(dt, ld3(p.gravity, w)));
            // predict_positions  solver_pbf.py:211-224
            x = vadd(x, smul(dt, v));
            // apply_position_bounds  solver_pbf.py:366-399
            float r = p.particle_radius, e = p.boundary_epsilon;
            float x_min = -p.bound[0] + r + e, x_max = p.bound[0] - r - e;
            float y_min = -p.bound[1] + r + e, y_max = p.bound[1] - r - e;
            float z_min = r + e, z_max = p.bound[2] * 2.0f - r - e;
            if (x.x < x_min) x.x = x_min;
            if (x.x > x_max) x.x = x_max;
            if (x.y < y_min) x.y = y_min;
            if (x.y > y_max) x.y = y_max;
            if (x.z < z_min) x.z = z_min;
            if (x.z > z_max) x.z = z_max;
        }
        st3(p.v_out, i, v);
        st3(p.x_out, i, x);
    }
    return x;
}

// The pre-op alone, in original order (coalesced): used when the hash kernel
// visits points in the previous build's sorted order, where the pre-op's
// loads and stores of a wave would each touch a separate cache line.
template <int MODE>
__global__ __launch_bounds__(kElemBlock) void k_pre_op(PreOp p, int n)
{
    const int i = blockIdx.x * kElemBlock + threadIdx.x;
    if (i < n) pre_point<MODE>(p, i);
}

// Pre-op (when fused), bucket id of point i; writes the bucket id and
// returns it.
template <int MODE>
__device__ __forceinline__ int hash_point(const HashArgs& a, int i)
{
    const PreOp& p = a.pre;
    const v3 x = pre_point<MODE>(p, i);
    int c;
    if (p.scenes) {
        // segmented grid: the point's scene selects cell width, dims and bucket base
        const fs_scene& sc = p.scenes[p.scene_of[i]];
        GridView g = a.g;
        g.dx = sc.dims[0];
        g.dy = sc.dims[1];
        g.dz = sc.dims[2];
        g.dxy = sc.dims[0] * sc.dims[1];
        g.base = sc.bucket_base;
        g.inv = sc.cell_width_inv;
        c = bucket_of(g, x.x, x.y, x.z);
    } else {
        c = bucket_of(a.g, x.x, x.y, x.z);
    }
    return c;
}

// The per-point words the fix-up gathers besides the positions and
// velocities, packed into one 16-byte record of point i: (mass, rest
// density, flags, mask).  An absent array's word is loaded from the flags
// array instead (always present, at least n words), so the four loads are
// straight-line code; the fix-up reads such a word as 0.
__device__ __forceinline__ int4 pack_load(const GatherOp& ga, int i)
{
    const int* fl = ga.flags;
    const int m = (ga.mass ? reinterpret_cast<const int*>(ga.mass) : fl)[i];
    const int rd = (ga.wvol ? reinterpret_cast<const int*>(ga.rest_density) : fl)[i];
    const int mk = (ga.mask ? ga.mask : fl)[i];
    return make_int4(m, rd, fl[i], mk);
}

// Bucket every point and take a slot rank from its bucket counter.  Lanes are
// fed in the previous build's sorted order when there is one, so equal buckets
// sit in runs of adjacent lanes: the first lane of each run adds the run's
// length to the counter (one atomic instruction for the wave, whatever the
// number of distinct buckets) and the run's lanes take consecutive ranks from
// it.  A bucket split over several runs just takes several atomics.  Ranks only
// need to be distinct: k_fix_gather orders each bucket by original index
// afterwards.
// ORDERED: a.order is set; PACK: a.pack.packed is set.  Both as template
// flags, the processing-order load and the packed record's loads are
// straight-line code issued together (lanes past n load point n - 1 and
// write nothing).
template <int MODE, bool ORDERED, bool PACK>
__global__ __launch_bounds__(kElemBlock) void k_hash_count(HashArgs a)
{
    const int t = blockIdx.x * kElemBlock + threadIdx.x;
    if (a.pre.control_reset && t < 16) a.pre.control_reset[t] = 0;
    const bool valid = t < a.n;   // the valid lanes of a wave are a prefix
    const int tl = min(t, a.n - 1);
    int i = 0;
    int4 pk = make_int4(0, 0, 0, 0);
    if (a.n > 0) {
        i = ORDERED ? a.order[tl] : tl;
        // the packed record of point t (original order: coalesced), stored
        // after the bucket so that its loads overlap the bucket's
        if constexpr (PACK) pk = pack_load(a.pack, tl);
    }
    int c = -1;
    if (valid) c = hash_point<MODE>(a, i);
    if (PACK && valid) a.pack.packed[t] = pk;
    const int lane = threadIdx.x & 63;
    const int prev = __shfl_up(c, 1, 64);
    const bool head = valid && (lane == 0 || prev != c);
    const unsigned long long heads = __ballot(head);
    const unsigned long long valids = __ballot(valid);
    const unsigned long long below = (lane == 63) ? ~0ull : ((2ull << lane) - 1ull);   // lanes 0..lane
    int base = 0;
    if (head) {
        const unsigned long long later = heads & ~below;
        const int next = later ? __ffsll((long long)later) - 1 : 64;
        const unsigned long long run = (next == 64 ? ~0ull : ((1ull << next) - 1ull)) & ~((1ull << lane) - 1ull);
        base = atomicAdd(&a.count[c], __popcll(run & valids));
    }
    const unsigned long long mine = heads & below;
    const int lead = mine ? 63 - __clzll((long long)mine) : 0;
    const int lead_base = __shfl(base, lead, 64);
    // bucket and rank in processing order: coalesced stores (k_scatter maps
    // position t back to its point)
    if (valid) {
        a.cell_of[t] = c;
        a.rank[t] = lead_base + (lane - lead);
    }
}


// Descriptor of one scan tile: [63:62] status (0 not ready, 1 aggregate,
// 2 inclusive prefix), [31:0] value.  Flag and value travel in one 64-bit
// word, so relaxed agent-scope atomics suffice.  k_scatter re-zeroes the
// descriptors after every scan, so the next build starts clean.
__device__ __forceinline__ uint64_t scan_desc(uint32_t status, int value)
{
    return ((uint64_t)status << 62) | (uint64_t)(uint32_t)value;
}

// Single-pass exclusive scan with decoupled look-back.  Tile = blockIdx.x:
// workgroups are dispatched in index order, so every predecessor of a
// resident tile has started and will publish.  Each wave scans 2048 counters
// of the tile in 8 rounds of 64 lanes x int4 (every load and store
// instruction covers 1 KB of contiguous memory), carrying the running total
// from round to round.
__global__ __launch_bounds__(kScanBlock) void k_bucket_scan(const int* __restrict__ in, int* __restrict__ out,
                                                            int m, int ntiles, uint64_t* desc)
{
    constexpr int kWaves = kScanBlock / 64;
    constexpr int kRounds = kScanItems / 4;
    constexpr int kWaveItems = kScanTile / kWaves;
    static_assert(kRounds * 256 == kWaveItems, "scan tiling");
    __shared__ int s_wsum[kWaves];
    __shared__ int s_excl;
    const int t = threadIdx.x, lane = t & 63, wid = t >> 6;
    const int tile = blockIdx.x;
    const int wbase = tile * kScanTile + wid * kWaveItems;

    int4 q[kRounds];
#pragma unroll
    for (int j = 0; j < kRounds; ++j) {
        const int b = wbase + j * 256 + lane * 4;
        if (b + 4 <= m) {
            q[j] = *reinterpret_cast<const int4*>(in + b);
        } else {
            q[j].x = (b < m) ? in[b] : 0;
            q[j].y = (b + 1 < m) ? in[b + 1] : 0;
            q[j].z = (b + 2 < m) ? in[b + 2] : 0;
            q[j].w = 0;
        }
    }
    int v[kScanItems];   // per round: the lane's 4 counters, inclusive-scanned
    int lex[kRounds];    // exclusive offset of the lane's group within the wave
    int wsum = 0;
#pragma unroll
    for (int j = 0; j < kRounds; ++j) {
        v[4 * j] = q[j].x;
        v[4 * j + 1] = v[4 * j] + q[j].y;
        v[4 * j + 2] = v[4 * j + 1] + q[j].z;
        v[4 * j + 3] = v[4 * j + 2] + q[j].w;
        const int tsum = v[4 * j + 3];
        int x = tsum;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            int y = __shfl_up(x, off, 64);
            if (lane >= off) x += y;
        }
        lex[j] = wsum + x - tsum;
        wsum += __shfl(x, 63, 64);
    }
    if (lane == 0) s_wsum[wid] = wsum;
    __syncthreads();
    int wprefix = 0, agg = 0;
#pragma unroll
    for (int w = 0; w < kWaves; ++w) {
        int ws = s_wsum[w];
        if (w < wid) wprefix += ws;
        agg += ws;
    }

    if (wid == 0) {
        int excl = 0;
        if (tile == 0) {
            if (lane == 0) __hip_atomic_store(&desc[0], scan_desc(2, agg), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        } else {
            if (lane == 0)
                __hip_atomic_store(&desc[tile], scan_desc(1, agg), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            int look = tile - 1;
            while (true) {
                const int idx = look - lane;
                int status = 2, val = 0;
                if (idx >= 0) {
                    uint64_t d = __hip_atomic_load(&desc[idx], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    status = (int)(d >> 62);
                    val = (int)(uint32_t)d;
                }
                if (__any(status == 0)) continue;
                unsigned long long pm = __ballot(status == 2);
                const int first = pm ? (__ffsll(pm) - 1) : 64;
                int contrib = (lane <= first) ? val : 0;
#pragma unroll
                for (int off = 32; off > 0; off >>= 1) contrib += __shfl_xor(contrib, off, 64);
                excl += contrib;
                if (pm) break;
                look -= 64;
            }
            if (lane == 0)
                __hip_atomic_store(&desc[tile], scan_desc(2, excl + agg), __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
        }
        if (lane == 0) s_excl = excl;
    }
    __syncthreads();
    const int wexcl = s_excl + wprefix;
#pragma unroll
    for (int j = 0; j < kRounds; ++j) {
        const int b = wbase + j * 256 + lane * 4;
        const int e = wexcl + lex[j];
        int4 o;
        o.x = e;
        o.y = e + v[4 * j];
        o.z = e + v[4 * j + 1];
        o.w = e + v[4 * j + 2];
        if (b + 4 <= m) {
            *reinterpret_cast<int4*>(out + b) = o;
        } else {
            if (b < m) out[b] = o.x;
            if (b + 1 < m) out[b + 1] = o.y;
            if (b + 2 < m) out[b + 2] = o.z;
        }
    }
    if (tile == ntiles - 1 && t == 0) out[m] = s_excl + agg;
}

__global__ __launch_bounds__(kElemBlock) void k_scatter(int n, const int* __restrict__ order,
                                                        const int* __restrict__ cell_t, const int* __restrict__ rank_t,
                                                        const int* __restrict__ start, int2* __restrict__ scatter,
                                                        int* __restrict__ count, uint64_t* scan_desc_clear, int ntiles)
{
    const int t = blockIdx.x * kElemBlock + threadIdx.x;
    if (t < ntiles) scan_desc_clear[t] = 0;
    if (t >= n) return;
    // position t of the hash's processing order holds point order[t] (or t)
    const int i = order ? order[t] : t;
    const int c = cell_t[t];
    const int slot = start[c] + rank_t[t];
    // (index, bucket): the fix-up reads the bucket with the index instead of
    // a dependent bucket load
    if ((unsigned)slot < (unsigned)n) scatter[slot] = make_int2(i, c);
    count[c] = 0;
}

constexpr int kFixBatch = 8;   // bucket members whose indices the fix-up loads at once

// The gathered streams are template flags (POS: particle data at all;
// PACKED: the build's packed records; VEL: velocities; SCENE: scene ids), so
// every load of a point's data is straight-line code issued together with
// the bucket-range loads, ahead of the rank loop.
template <bool POS, bool PACKED, bool VEL, bool SCENE>
__global__ __launch_bounds__(kElemBlock) void k_fix_gather(int n, const int2* __restrict__ scatter,
                                                           const int* __restrict__ start, int* __restrict__ point_ids,
                                                           int* __restrict__ point_cells, GatherOp ga)
{
    const int k = blockIdx.x * kElemBlock + threadIdx.x;
    if (k >= n) return;
    const int2 e = scatter[k];
    const int i = e.x;
    const int c = e.y;
    const int sa = start[c], sb = start[c + 1];
    v3 x = V(0.0f, 0.0f, 0.0f), u = V(0.0f, 0.0f, 0.0f);
    float m = 0.0f, rd = 0.0f;
    int fl = 0, mk = 0, sc = 0;
    if constexpr (POS) {
        x = ld3(ga.x, i);
        if constexpr (PACKED) {
            // words of absent arrays (pack_load) read as 0
            const int4 pk = ga.packed[i];
            m = ga.mass ? __int_as_float(pk.x) : 0.0f;
            rd = ga.wvol ? __int_as_float(pk.y) : 0.0f;
            fl = pk.z;
            mk = ga.mask ? pk.w : 0;
        } else {
            if (ga.mass) m = ga.mass[i];
            if (ga.wvol) rd = ga.rest_density[i];
            fl = ga.flags[i];
            if (ga.mask) mk = ga.mask[i];
        }
        if constexpr (SCENE) sc = ga.scene_of[i];
        if constexpr (VEL) u = ld3(ga.v, i);
    }
    // rank = members of the bucket with a smaller index; each trip loads
    // kFixBatch members' indices at once (slots clamped into the bucket,
    // those past its end not counted)
    const int a = min(max(sa, 0), n), b = min(max(sb, a), n);
    int r = 0;
    for (int j = a; j < b; j += kFixBatch) {
        int t[kFixBatch];
#pragma unroll
        for (int q = 0; q < kFixBatch; ++q) t[q] = scatter[min(j + q, b - 1)].x;
#pragma unroll
        for (int q = 0; q < kFixBatch; ++q) r += (j + q < b && t[q] < i) ? 1 : 0;
    }
    const int f = a + r;
    point_ids[f] = i;
    point_cells[f] = c;
    if constexpr (POS) {
        ga.pos[f] = make_float4(x.x, x.y, x.z, m);
        // compute_particle_volume (solver_wcsph.py:109-116) times rest density,
        // as compute_density forms mass_nei (solver_wcsph.py:148)
        if (ga.wvol) ga.wvol[f] = rd * (m / fmax_(rd, 1.0e-6f));
        const int act = fl & 1;
        ga.sflags[f] = act | ((act && mk == 0) ? kDynamic : 0);
        if (ga.any_inactive && __ballot(!act) && (threadIdx.x & 63) == __ffsll((long long)__ballot(1)) - 1)
            atomicOr(ga.any_inactive, 1);
        if constexpr (SCENE) ga.sscene[f] = sc;
        if constexpr (VEL) ga.vel[f] = make_float4(u.x, u.y, u.z, 0.0f);
    }
}

__global__ __launch_bounds__(kElemBlock) void k_export_tables(int m, const int* __restrict__ start,
                                                              int* __restrict__ starts, int* __restrict__ ends)
{
    const int c = blockIdx.x * kElemBlock + threadIdx.x;
    if (c >= m) return;
    int s = start[c], e = start[c + 1];
    starts[c] = (s == e) ? 0 : s;
    ends[c] = (s == e) ? 0 : e;
}

// Diagnostics: candidates of hash_grid_query(points[i], radius), counted or
// listed in visit order, optionally filtered by the kernels' support test.
template <bool LIST>
__global__ __launch_bounds__(kElemBlock) void k_candidates(GridView g, int n, const float* __restrict__ pts,
                                                           const int* __restrict__ point_ids, float radius, int mode,
                                                           int* __restrict__ counts,
                                                           const int64_t* __restrict__ offsets,
                                                           int* __restrict__ indices)
{
    const int i = blockIdx.x * kElemBlock + threadIdx.x;
    if (i >= n) return;
    v3 xi = ld3(pts, i);
    QueryRows it;
    it.begin(g, xi.x, xi.y, xi.z, radius);
    int cnt = 0;
    int64_t o = LIST ? offsets[i] : 0;
    if (it.valid()) {
        while (true) {
            for (int k = it.k; k < it.kend; ++k) {
                int j = point_ids[k];
                if (mode >= 0) {
                    float r = vlen(vsub(xi, ld3(pts, j)));
                    if (mode == 0 && !(r / radius <= 1.0f)) continue;
                    if (mode == 1 && !(r < radius)) continue;
                }
                if (LIST) indices[o + cnt] = j;
                ++cnt;
            }
            if (!it.advance(g)) break;
        }
    }
    if (!LIST) counts[i] = cnt;
}

// ---------------------------------------------------------------------------
// host
// ---------------------------------------------------------------------------

static int scan_tiles(const fs_grid* g)
{
    long long m = (long long)g->dim_x * g->dim_y * g->dim_z;
    return (int)((m + kScanTile - 1) / kScanTile);
}

int32_t build_grid(fs_grid* g, int n, float radius, const PreOp& pre, const GatherOp& ga, hipStream_t s)
{
    if (n < 0 || n > g->capacity) {
        set_error("grid build: %d points exceed the reserved capacity %d", n, g->capacity);
        return FS_ERR_CAPACITY;
    }
    if (!(radius > 0.0f)) {
        set_error("grid build: radius must be positive");
        return FS_ERR_ARG;
    }
    if (ga.pos && ga.scene_of && !ga.packed) {
        set_error("grid build: scene ids need packed records");
        return FS_ERR_ARG;
    }
    const bool reorder = g->num_points == n && g->cell_width > 0.0f;
    g->num_points = n;
    g->cell_width = radius;
    g->cell_width_inv = 1.0f / radius;
    const GridView gv = grid_view(g);
    const int m = g->dim_x * g->dim_y * g->dim_z;
    // the previous build's sorted order, when it covered the same points
    const int* order = reorder ? g->point_ids : nullptr;
    HashArgs ha{gv, pre, n, order, g->cell_of, g->rank, g->cell_count};
    if (ga.pos && ga.packed) ha.pack = ga;
    const int nb = blocks_for(n, kElemBlock);
    if (order && nb > 0 && pre.mode != kPrePlain) {
        // pre-op in original order first; the hash then buckets its output
        if (pre.mode == kPreIntegrate) FS_LAUNCH(k_pre_op<kPreIntegrate>, dim3(nb), dim3(kElemBlock), 0, s, pre, n);
        else FS_LAUNCH(k_pre_op<kPrePredict>, dim3(nb), dim3(kElemBlock), 0, s, pre, n);
        ha.pre.mode = kPrePlain;
        ha.pre.x_in = pre.x_out;
    }
    if (nb > 0 || pre.control_reset) {
        const int nbh = nb > 0 ? nb : 1;
        const bool pack = ha.pack.packed != nullptr;
        switch (ha.pre.mode) {
#define HASH_LAUNCH(M)                                                                                       \
    do {                                                                                                     \
        if (ha.order && pack) FS_LAUNCH((k_hash_count<M, true, true>), dim3(nbh), dim3(kElemBlock), 0, s, ha); \
        else if (ha.order) FS_LAUNCH((k_hash_count<M, true, false>), dim3(nbh), dim3(kElemBlock), 0, s, ha);  \
        else if (pack) FS_LAUNCH((k_hash_count<M, false, true>), dim3(nbh), dim3(kElemBlock), 0, s, ha);      \
        else FS_LAUNCH((k_hash_count<M, false, false>), dim3(nbh), dim3(kElemBlock), 0, s, ha);              \
    } while (0)
        case kPreIntegrate: HASH_LAUNCH(kPreIntegrate); break;
        case kPrePredict: HASH_LAUNCH(kPrePredict); break;
        default: HASH_LAUNCH(kPrePlain); break;
#undef HASH_LAUNCH
        }
    }
    const int ntiles = scan_tiles(g);
    FS_LAUNCH(k_bucket_scan, dim3(ntiles), dim3(kScanBlock), 0, s, g->cell_count, g->cell_start, m, ntiles,
              g->scan_state);
    const int nbs = nb > blocks_for(ntiles, kElemBlock) ? nb : blocks_for(ntiles, kElemBlock);
    FS_LAUNCH(k_scatter, dim3(nbs), dim3(kElemBlock), 0, s, n, order, g->cell_of, g->rank, g->cell_start,
              reinterpret_cast<int2*>(g->scatter_ids),
              g->cell_count, g->scan_state, ntiles);
    if (nb == 0) return FS_OK;
    const int2* sids = reinterpret_cast<const int2*>(g->scatter_ids);
    const bool pos = ga.pos != nullptr, packed = pos && ga.packed, vel = pos && ga.v, scene = pos && ga.scene_of;
#define FIX_LAUNCH(P, K, V, S)                                                                                   \
    FS_LAUNCH((k_fix_gather<P, K, V, S>), dim3(nb), dim3(kElemBlock), 0, s, n, sids, g->cell_start, g->point_ids, \
              g->point_cells, ga)
    if (!pos) FIX_LAUNCH(false, false, false, false);
    else if (!packed) {
        if (vel) FIX_LAUNCH(true, false, true, false);
        else FIX_LAUNCH(true, false, false, false);
    } else if (scene) {
        if (vel) FIX_LAUNCH(true, true, true, true);
        else FIX_LAUNCH(true, true, false, true);
    } else {
        if (vel) FIX_LAUNCH(true, true, true, false);
        else FIX_LAUNCH(true, true, false, false);
    }
#undef FIX_LAUNCH
    return FS_OK;
}

}  // namespace fs

using namespace fs;

extern "C" {

const char* fs_last_error(void) { return g_err; }

void fs_profile_enable(int32_t on)
{
    if (!on) prof_collect();
    g_profiling = on != 0;
}

void fs_profile_reset(void)
{
    prof_collect();
    g_prof_totals.clear();
}

int32_t fs_profile_report(char* buf, size_t len)
{
    prof_collect();
    std::string out;
    for (auto& kv : g_prof_totals) {
        char line[256];
        snprintf(line, sizeof(line), "%s %lld %.6f\n", kv.first.c_str(), kv.second.second, kv.second.first);
        out += line;
    }
    if (buf && len) {
        size_t n = out.size() < len - 1 ? out.size() : len - 1;
        memcpy(buf, out.data(), n);
        buf[n] = 0;
    }
    return (int32_t)out.size();
}

int32_t fs_version(void) { return 100; }

size_t fs_grid_workspace_bytes(int32_t capacity, int32_t dim_x, int32_t dim_y, int32_t dim_z)
{
    if (capacity < 0 || dim_x <= 0 || dim_y <= 0 || dim_z <= 0) return 0;
    size_t m = (size_t)dim_x * dim_y * dim_z;
    size_t cap = (size_t)(capacity > 0 ? capacity : 1);
    size_t ntiles = (m + kScanTile - 1) / kScanTile;
    size_t b = 0;
    b += align_up(cap * 4) * 4;                 // point_ids, point_cells, cell_of, rank
    b += align_up(cap * 8);                     // scatter_ids: (index, bucket) pairs
    b += align_up((m + 1) * 4);                 // cell_start
    b += align_up(m * 4);                       // cell_count
    b += align_up(ntiles * 8);                  // scan_state
    b += align_up(8 * 4);                       // sync
    return b;
}

int32_t fs_grid_init(fs_grid* g, void* workspace, size_t bytes, int32_t capacity, int32_t dim_x, int32_t dim_y,
                     int32_t dim_z, int32_t round_mode, fs_stream_t stream)
{
    if (!g || !workspace || capacity < 0 || dim_x <= 0 || dim_y <= 0 || dim_z <= 0) {
        set_error("fs_grid_init: invalid arguments");
        return FS_ERR_ARG;
    }
    if (capacity > kMaxParticles) {
        set_error("fs_grid_init: capacity %d exceeds the maximum of %d points", capacity, kMaxParticles);
        return FS_ERR_ARG;
    }
    if ((long long)dim_x * dim_y * dim_z >= (1LL << 30)) {
        set_error("fs_grid_init: grid dims too large");
        return FS_ERR_ARG;
    }
    size_t need = fs_grid_workspace_bytes(capacity, dim_x, dim_y, dim_z);
    if (bytes < need) {
        set_error("fs_grid_init: workspace %zu bytes < required %zu", bytes, need);
        return FS_ERR_WORKSPACE;
    }
    hipStream_t s = (hipStream_t)stream;
    size_t m = (size_t)dim_x * dim_y * dim_z;
    size_t cap = (size_t)(capacity > 0 ? capacity : 1);
    size_t ntiles = (m + kScanTile - 1) / kScanTile;
    char* p = (char*)workspace;
    auto take = [&](size_t nbytes) {
        char* r = p;
        p += align_up(nbytes);
        return r;
    };
    memset(g, 0, sizeof(*g));
    g->dim_x = dim_x;
    g->dim_y = dim_y;
    g->dim_z = dim_z;
    g->round_mode = round_mode ? 1 : 0;
    g->capacity = capacity;
    g->point_ids = (int32_t*)take(cap * 4);
    g->point_cells = (int32_t*)take(cap * 4);
    g->cell_of = (int32_t*)take(cap * 4);
    g->rank = (int32_t*)take(cap * 4);
    g->scatter_ids = (int32_t*)take(cap * 8);
    g->cell_start = (int32_t*)take((m + 1) * 4);
    g->cell_count = (int32_t*)take(m * 4);
    g->scan_state = (uint64_t*)take(ntiles * 8);
    g->sync = (uint32_t*)take(8 * 4);
    int32_t rc = hip_check(hipMemsetAsync(g->cell_count, 0, m * 4, s), "fs_grid_init memset");
    if (rc) return rc;
    rc = hip_check(hipMemsetAsync(g->cell_start, 0, (m + 1) * 4, s), "fs_grid_init memset");
    if (rc) return rc;
    rc = hip_check(hipMemsetAsync(g->scan_state, 0, ntiles * 8, s), "fs_grid_init memset");
    if (rc) return rc;
    return hip_check(hipMemsetAsync(g->sync, 0, 8 * 4, s), "fs_grid_init memset");
}

int32_t fs_grid_build(fs_grid* g, const float* points, int32_t n, float radius, fs_stream_t stream)
{
    if (!g || (n > 0 && !points)) {
        set_error("fs_grid_build: invalid arguments");
        return FS_ERR_ARG;
    }
    PreOp pre;
    pre.mode = kPrePlain;
    pre.x_in = points;
    GatherOp ga;
    return build_grid(g, n, radius, pre, ga, (hipStream_t)stream);
}

int32_t fs_grid_export_tables(const fs_grid* g, int32_t* cell_starts, int32_t* cell_ends, fs_stream_t stream)
{
    if (!g || !cell_starts || !cell_ends) {
        set_error("fs_grid_export_tables: invalid arguments");
        return FS_ERR_ARG;
    }
    int m = g->dim_x * g->dim_y * g->dim_z;
    hipLaunchKernelGGL(k_export_tables, dim3(blocks_for(m, kElemBlock)), dim3(kElemBlock), 0, (hipStream_t)stream, m,
                       g->cell_start, cell_starts, cell_ends);
    FS_LAUNCH_CHECK("k_export_tables");
    return FS_OK;
}

int32_t fs_grid_count_candidates(const fs_grid* g, const float* points, float radius, int32_t within_mode,
                                 int32_t* counts, fs_stream_t stream)
{
    if (!g || !counts || (g->num_points > 0 && !points) || within_mode < -1 || within_mode > 1) {
        set_error("fs_grid_count_candidates: invalid arguments");
        return FS_ERR_ARG;
    }
    int n = g->num_points;
    if (n == 0) return FS_OK;
    hipLaunchKernelGGL(k_candidates<false>, dim3(blocks_for(n, kElemBlock)), dim3(kElemBlock), 0,
                       (hipStream_t)stream, grid_view(g), n, points, g->point_ids, radius, within_mode, counts,
                       (const int64_t*)nullptr, (int*)nullptr);
    FS_LAUNCH_CHECK("k_candidates<count>");
    return FS_OK;
}

int32_t fs_grid_list_candidates(const fs_grid* g, const float* points, float radius, int32_t within_mode,
                                const int64_t* offsets, int32_t* indices, fs_stream_t stream)
{
    if (!g || !offsets || !indices || (g->num_points > 0 && !points) || within_mode < -1 || within_mode > 1) {
        set_error("fs_grid_list_candidates: invalid arguments");
        return FS_ERR_ARG;
    }
    int n = g->num_points;
    if (n == 0) return FS_OK;
    hipLaunchKernelGGL(k_candidates<true>, dim3(blocks_for(n, kElemBlock)), dim3(kElemBlock), 0, (hipStream_t)stream,
                       grid_view(g), n, points, g->point_ids, radius, within_mode, (int*)nullptr, offsets, indices);
    FS_LAUNCH_CHECK("k_candidates<list>");
    return FS_OK;
}

}  // extern "C"
