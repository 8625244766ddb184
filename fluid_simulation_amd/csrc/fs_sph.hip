// fs_sph.hip -- neighbour-loop passes and substep orchestration of the WCSPH,
// DFSPH and PBF solvers on gfx950.
//
// Each pass is one launch of one lane per particle, lanes in the grid's sorted
// order (slot s -> original index i = point_ids[s]) and workgroups mapped so
// each XCD walks one contiguous run of slots.  Neighbour data is read from
// copies gathered into sorted order (contiguous bucket rows); per-particle
// results are written at the original index, and neighbour-visible results
// also into sorted scratch for the next pass.  Sums run over candidates in
// hash_grid_query visit order with the reference's float32 expressions, so a
// pass reproduces the reference kernel's arithmetic operation for operation.
//
// A pass that scans the grid records each particle's within-cut candidates
// (visit order) in a neighbour list; the next pass over the same positions
// and the same grid replays it instead of scanning the 27 buckets again.  A
// particle's candidates depend on its current position (the query range is
// recomputed from it), so every pass after the positions change scans again.
#include "fs_coop.h"
#include "fs_internal.h"

#include <atomic>
#include <mutex>

namespace fs {

constexpr int kPassBlock = 256;
constexpr int kMinPassBlock = 64;   // smallest workgroup of a tunable pass
constexpr int kScanBlock = 64;  // workgroup of the separation pass
constexpr int kSepCap = 24;     // separation: few survivors, a shorter list keeps occupancy
constexpr int kReplayCap = 16;  // ... of a list-replaying pass (only its overflow fallback scans)
constexpr float kCutSlack = 1.0001f;   // candidates with d2 > h^2*slack have q > 1

// loop-control words in fs_work::control (int32[64])
enum Control : int {
    kStopRho = 0,
    kStopVel = 1,
    kItRho = 2,
    kCorrRho = 3,
    kItVel = 4,
    kCorrVel = 5,
    kAnyInactive = 6,   // set by the builds when some particle is inactive
    kListNeed = FS_CONTROL_LIST_NEED,   // the longest list an overflowing particle needed
    kControlWords = 64,
};

// The substep's dt: the device scalar when one is given (a device dt of 0
// marks an idle substep, which the caller's driver uses once a frame's time
// is spent), else the launch parameter.  Returns false for an idle substep.
__device__ __forceinline__ bool substep_dt(const float* dt_device, float host_dt, float& dt)
{
    dt = dt_device ? *dt_device : host_dt;
    return !(dt_device && dt == 0.0f);
}

struct PassCommon {
    GridView g;
    int n;
    const int* __restrict__ perm;
    const float4* __restrict__ pos;
    const int* __restrict__ sflags;
    float h;
    float cut2;
    NbrList L;
};

// What a two-phase neighbour body fetches of neighbour k (sorted slot):
// its flags and position (+ mass in w), plus the pass's own neighbour data.
struct NbrX {
    int flags;
    float4 x;
};
struct NbrXS {   // + one scalar
    int flags;
    float4 x;
    float s;
};
struct NbrXV {   // + one float4
    int flags;
    float4 x;
    float4 v;
};
struct NbrXVS {  // + a float4 and a scalar
    int flags;
    float4 x;
    float4 v;
    float s;
};

struct AdjNbr {   // reverse force pass: flags, x + mass, abar + c, v + m/rho, p/max(rho^2)
    int flags;
    float4 x;
    float4 abar;
    float4 v;
    float P;
};

__device__ __forceinline__ int pass_slot() { return xcd_tile() * kPassBlock + threadIdx.x; }
template <int BLOCK>
__device__ __forceinline__ int pass_slot_b() { return xcd_tile() * BLOCK + threadIdx.x; }
// slot of a lane's particle with LPP adjacent lanes per particle
template <int BLOCK, int LPP>
__device__ __forceinline__ int pass_slot_g() { return xcd_tile() * (BLOCK / LPP) + (int)(threadIdx.x / LPP); }

// Per-lane view of a batch scene: its bucket table, cell width and radius.
__device__ __forceinline__ void scene_view(PassCommon& c, const fs_scene& sc)
{
    c.g.dx = sc.dims[0];
    c.g.dy = sc.dims[1];
    c.g.dz = sc.dims[2];
    c.g.dxy = sc.dims[0] * sc.dims[1];
    c.g.base = sc.bucket_base;
    c.g.inv = sc.cell_width_inv;
    c.h = sc.smoothing_length;
    c.cut2 = c.h * c.h * kCutSlack;
    c.L.cut2 = c.cut2;
}

// neighbour loop of a scanning pass (SCAN_*) or of a list-replaying pass
// (REPLAY_*); the _B forms are for kernels templated on their block size BLOCK
#define NBR_LOOP(blk, rep, unr, c, s, px, py, pz, count, body)                                                 \
    for_each_neighbor<blk, CAP, rep, kBatch, unr>((c).g, (c).pos, (s), (px), (py), (pz), (c).h, (c).cut2, lds,  \
                                                  (c).L, count, body)
#define SCAN_NEIGHBORS(...) NBR_LOOP(kPassBlock, false, kReplayUnroll, __VA_ARGS__)
#define REPLAY_NEIGHBORS(...) NBR_LOOP(kPassBlock, true, kReplayUnroll, __VA_ARGS__)
#define SCAN_NEIGHBORS_B(...) NBR_LOOP(BLOCK, false, kReplayUnroll, __VA_ARGS__)
#define REPLAY_NEIGHBORS_B(...) NBR_LOOP(BLOCK, true, kReplayUnroll, __VA_ARGS__)
// a replaying pass templated on its block size and replay unroll (UNR)
#define REPLAY_NEIGHBORS_BU(...) NBR_LOOP(BLOCK, true, UNR, __VA_ARGS__)

// Scanning neighbour loop of a pass whose body is fold(term(k, data)): the
// one-lane compacting scan (LPP = 1) or the group scan of LPP lanes per
// particle (scan_split, U survivors per lane fetched together).  A term of a
// skipped neighbour is zero, which leaves every sum unchanged (a
// round-to-nearest sum from +0 never becomes -0).  Returns the number of
// candidates.
template <int BLOCK, int CAP, int LPP, int U, class Count, class Load, class Term, class Fold>
__device__ __forceinline__ int scan_pass(const PassCommon& c, int s, float px, float py, float pz, int* lds,
                                         Count&& count, Load load, Term term, Fold fold)
{
    if constexpr (LPP == 1) {
        return for_each_neighbor<BLOCK, CAP, false>(
            c.g, c.pos, s, px, py, pz, c.h, c.cut2, lds, c.L, count,
            phased(load, [&](int k, const decltype(load(0))& d) { fold(term(k, d)); }));
    } else {
        NbrList scan = c.L;
        if (scan.mode == kListUse) scan.mode = kListNone;
        return scan_split<BLOCK, CAP, LPP, kBatch / LPP, U>(c.g, c.pos, s, px, py, pz, c.h, c.cut2, lds, scan, count,
                                                              load, term, fold);
    }
}

// ---------------------------------------------------------------------------
// DFSPH residual: each block of the rate pass stores its partial sums; a
// one-block kernel then reduces them in a fixed order and decides the early exit
// ---------------------------------------------------------------------------
template <int BLOCK>
__device__ void residual_partials(double v, double c, double* partials)
{
    __shared__ double s_v[BLOCK / 64];
    __shared__ double s_c[BLOCK / 64];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        v += __shfl_xor(v, off, 64);
        c += __shfl_xor(c, off, 64);
    }
    if (lane == 0) {
        s_v[wid] = v;
        s_c[wid] = c;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double bv = 0.0, bc = 0.0;
        for (int w = 0; w < BLOCK / 64; ++w) {
            bv += s_v[w];
            bc += s_c[w];
        }
        partials[2 * blockIdx.x] = bv;
        partials[2 * blockIdx.x + 1] = bc;
    }
}

// The reference's residual is numpy's float32 mean over the dynamic
// particles, in original order (solver_dfsph.py:579-596): np.add.reduce walks
// the compressed array in buffers of 8192 elements, adding each buffer's
// pairwise sum to a float32 running total, and np.mean divides that total by
// the count in float64 and rounds the quotient to float32.  The pairwise sum
// of one buffer splits a range of n > 128 elements at n/2 rounded down to a
// multiple of 8 and sums a range of 8..128 elements with eight interleaved
// accumulators, fewer than 8 one after another.
constexpr int kNpBuffer = 8192;
constexpr int kNpLeaf = 128;

__device__ __forceinline__ int np_split(int n)
{
    int n2 = n / 2;
    return n2 - n2 % 8;
}

// pairwise sum of a range of at most kNpLeaf elements
__device__ float np_leaf_sum(const float* __restrict__ a, int n)
{
    if (n < 8) {
        float r = -0.0f;
        for (int i = 0; i < n; ++i) r += a[i];
        return r;
    }
    float r0 = a[0], r1 = a[1], r2 = a[2], r3 = a[3], r4 = a[4], r5 = a[5], r6 = a[6], r7 = a[7];
    int i = 8;
    for (; i < n - (n % 8); i += 8) {
        r0 += a[i];
        r1 += a[i + 1];
        r2 += a[i + 2];
        r3 += a[i + 3];
        r4 += a[i + 4];
        r5 += a[i + 5];
        r6 += a[i + 6];
        r7 += a[i + 7];
    }
    float res = ((r0 + r1) + (r2 + r3)) + ((r4 + r5) + (r6 + r7));
    for (; i < n; ++i) res += a[i];
    return res;
}

// The residual value of one particle as the reference forms it:
// np.maximum(x, 0.0) keeps a NaN.
__device__ __forceinline__ float np_max0(float x) { return x != x ? x : (x >= 0.0f ? x : 0.0f); }

struct ResidualArgs {
    const double* __restrict__ partials;   // per-block (sum, count) of the rate pass
    int nb;
    int* control;
    int stop_idx, it_idx, iter;
    double eta;
    // the exact form: values in original order (vals - sub, sub optional) of
    // the dynamic particles, and a scratch of n floats for their compression
    int n;
    const int* __restrict__ flags;
    const int* __restrict__ mask;
    const float* __restrict__ vals;
    const float* __restrict__ sub;
    float* scratch;
    int force_exact;   // decide every iteration with the exact float32 mean
    const float* dt_device;   // idle substep: no iteration
};

// One block.  The rate pass's float64 partial sums give the mean to about
// 1e-15; numpy's float32 mean differs from the exact mean by less than
// (40 + buffers) * 2^-24 relative (each value passes through at most 25
// roundings in its leaf, 7 pairwise levels and one per later buffer; all
// values are >= 0).  A residual that far from eta decides the break as the
// float32 mean would; a closer one (or any NaN) is decided by forming the
// float32 mean exactly as numpy does.
__global__ __launch_bounds__(kPassBlock) void k_dfsph_residual(ResidualArgs a)
{
    __shared__ double s_v[kPassBlock / 64];
    __shared__ double s_c[kPassBlock / 64];
    __shared__ int s_exact;
    __shared__ int s_scan[kPassBlock / 64 + 1];
    __shared__ int s_leaf_lo[2 * kNpBuffer / kNpLeaf + 8];
    __shared__ int s_leaf_n[2 * kNpBuffer / kNpLeaf + 8];
    __shared__ float s_leaf_sum[2 * kNpBuffer / kNpLeaf + 8];
    __shared__ int s_nleaf;
    if (a.control[a.stop_idx]) return;
    float dt_unused;
    if (!substep_dt(a.dt_device, 1.0f, dt_unused)) return;
    // the partials eight at a time per thread, so the loads of a batch are
    // in flight together (the float64 order only moves the mean within the
    // margin below; the exact path does not use it)
    constexpr int kBatchP = 8;
    const double2* __restrict__ pp = reinterpret_cast<const double2*>(a.partials);
    double bv[kBatchP], bc[kBatchP];
#pragma unroll
    for (int u = 0; u < kBatchP; ++u) bv[u] = bc[u] = 0.0;
    int b = threadIdx.x;
    for (; b + (kBatchP - 1) * kPassBlock < a.nb; b += kBatchP * kPassBlock) {
        double2 q[kBatchP];
#pragma unroll
        for (int u = 0; u < kBatchP; ++u) q[u] = pp[b + u * kPassBlock];
#pragma unroll
        for (int u = 0; u < kBatchP; ++u) {
            bv[u] += q[u].x;
            bc[u] += q[u].y;
        }
    }
    for (; b < a.nb; b += kPassBlock) {
        const double2 q = pp[b];
        bv[0] += q.x;
        bc[0] += q.y;
    }
    double tv = ((bv[0] + bv[1]) + (bv[2] + bv[3])) + ((bv[4] + bv[5]) + (bv[6] + bv[7]));
    double tc = ((bc[0] + bc[1]) + (bc[2] + bc[3])) + ((bc[4] + bc[5]) + (bc[6] + bc[7]));
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        tv += __shfl_xor(tv, off, 64);
        tc += __shfl_xor(tc, off, 64);
    }
    if (lane == 0) {
        s_v[wid] = tv;
        s_c[wid] = tc;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double sv = 0.0, sc = 0.0;
        for (int w = 0; w < kPassBlock / 64; ++w) {
            sv += s_v[w];
            sc += s_c[w];
        }
        a.control[a.it_idx] += 1;
        int exact = 0;
        if (a.iter > 0) {
            if (!(sc > 0.0)) {
                // no dynamic particle: the reference's residual is 0.0
                if (0.0 <= a.eta) a.control[a.stop_idx] = 1;
            } else {
                const double residual = sv / sc;
                const double buffers = ceil(sc / (double)kNpBuffer);
                const double margin = residual * ((40.0 + buffers) * 0x1.0p-24 + sc * 0x1.0p-52);
                if (a.force_exact || !(residual == residual) || fabs(residual - a.eta) <= margin) exact = 1;
                else if (residual <= a.eta) a.control[a.stop_idx] = 1;
            }
        }
        s_exact = exact;
    }
    __syncthreads();
    if (!s_exact) return;

    // compress the dynamic particles' values into scratch, in original order
    constexpr int kItems = 8;
    int base_out = 0;
    for (int base = 0; base < a.n; base += kPassBlock * kItems) {
        const int i0 = base + threadIdx.x * kItems;
        float v[kItems];
        int cnt = 0;
#pragma unroll
        for (int u = 0; u < kItems; ++u) {
            const int i = i0 + u;
            const bool dyn = i < a.n && (a.flags[i] & 1) != 0 && a.mask[i] == 0;
            v[u] = 0.0f;
            if (dyn) {
                const float x = a.sub ? a.vals[i] - a.sub[i] : a.vals[i];
                v[u] = np_max0(x);
                ++cnt;
            }
        }
        // exclusive scan of the per-thread counts over the block
        int incl = cnt;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const int t = __shfl_up(incl, off, 64);
            if (lane >= off) incl += t;
        }
        if (lane == 63) s_scan[wid] = incl;
        __syncthreads();
        if (threadIdx.x == 0) {
            int run = 0;
            for (int w = 0; w < kPassBlock / 64; ++w) {
                const int t = s_scan[w];
                s_scan[w] = run;
                run += t;
            }
            s_scan[kPassBlock / 64] = run;
        }
        __syncthreads();
        int out = base_out + s_scan[wid] + incl - cnt;
#pragma unroll
        for (int u = 0; u < kItems; ++u) {
            const int i = i0 + u;
            if (i < a.n && (a.flags[i] & 1) != 0 && a.mask[i] == 0) a.scratch[out++] = v[u];
        }
        base_out += s_scan[kPassBlock / 64];
        __syncthreads();
    }
    const int nd = base_out;

    // buffer by buffer: pairwise sums of the leaves in parallel, the tree and
    // the running total on one lane
    float total = 0.0f;
    for (int b0 = 0; b0 < nd; b0 += kNpBuffer) {
        const int m = min(kNpBuffer, nd - b0);
        if (threadIdx.x == 0) {
            int st_lo[32], st_n[32], sp = 0, nl = 0;
            st_lo[sp] = 0;
            st_n[sp++] = m;
            while (sp > 0) {
                --sp;
                const int lo = st_lo[sp], len = st_n[sp];
                if (len <= kNpLeaf) {
                    s_leaf_lo[nl] = lo;
                    s_leaf_n[nl++] = len;
                } else {
                    const int n2 = np_split(len);
                    st_lo[sp] = lo + n2;
                    st_n[sp++] = len - n2;
                    st_lo[sp] = lo;
                    st_n[sp++] = n2;
                }
            }
            s_nleaf = nl;
        }
        __syncthreads();
        for (int l = threadIdx.x; l < s_nleaf; l += kPassBlock)
            s_leaf_sum[l] = np_leaf_sum(a.scratch + b0 + s_leaf_lo[l], s_leaf_n[l]);
        __syncthreads();
        if (threadIdx.x == 0) {
            // post-order walk of the same tree: a frame is (length, stage)
            int f_n[32], f_stage[32], fp = 0, next = 0, vp = 0;
            float vals[32];
            f_n[fp] = m;
            f_stage[fp++] = 0;
            while (fp > 0) {
                const int len = f_n[fp - 1];
                if (len <= kNpLeaf) {
                    --fp;
                    vals[vp++] = s_leaf_sum[next++];
                    continue;
                }
                const int n2 = np_split(len);
                if (f_stage[fp - 1] == 0) {
                    f_stage[fp - 1] = 1;
                    f_n[fp] = n2;
                    f_stage[fp++] = 0;
                } else if (f_stage[fp - 1] == 1) {
                    f_stage[fp - 1] = 2;
                    f_n[fp] = len - n2;
                    f_stage[fp++] = 0;
                } else {
                    --fp;
                    const float rgt = vals[--vp];
                    const float lft = vals[--vp];
                    vals[vp++] = lft + rgt;
                }
            }
            total = total + vals[0];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        // np.mean: float32(float64(total) / count); the break test in float64
        const float mean = nd > 0 ? (float)((double)total / (double)nd) : 0.0f;
        if ((double)mean <= a.eta) a.control[a.stop_idx] = 1;
    }
}

// ---------------------------------------------------------------------------
// WCSPH
// ---------------------------------------------------------------------------

struct WcsphDensityArgs {
    PassCommon c;
    const fs_scene* __restrict__ scenes;   // batch: scene table and sorted scene ids
    const int* __restrict__ sscene;
    const float* __restrict__ wvol;
    const float* __restrict__ rest_density;
    const float* __restrict__ stiffness;
    const float* __restrict__ exponent;
    int clamp;
    const float* dt_device;   // plain pass: a device dt of 0 makes the pass idle
    const float* rho_in;      // idle pass: carried over to rho / pressure (optional)
    const float* pressure_in;
    float* rho;
    float* pressure;
    float4* sca;    // rho, p, p/max(rho^2,1e-6), -
    float4* vel;    // .w <- m/max(rho,1e-6)
};

// compute_density_pressure  solver_wcsph.py:119-158, for slot s with the
// pass parameters in `c` (LPP lanes per particle; the group's first lane
// writes); loop(xi, load, term, fold) runs the neighbour loop
template <int LPP, class Loop>
__device__ __forceinline__ void wcsph_density_core(const WcsphDensityArgs& a, const PassCommon& c, int clamp, int s,
                                                   Loop&& loop)
{
    const bool lead = (threadIdx.x & (LPP - 1)) == 0;
    const int i = c.perm[s];
    const float4 xi = c.pos[s];
    if (!(c.sflags[s] & kActive)) {
        if (lead) {
            a.rho[i] = 0.0f;
            a.pressure[i] = 0.0f;
            a.sca[s] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            a.vel[s].w = xi.w / fmax_(0.0f, 1.0e-6f);
            if (c.L.mode == kListBuild) c.L.cnt[s] = 0;
        }
        return;
    }
    const Smooth sm = smooth_of(c.h);
    float rho_temp = 0.0f;
    auto load = [&](int k) { return NbrXS{c.sflags[k], c.pos[k], a.wvol[k]}; };
    auto term = [&](int k, const NbrXS& nb) -> float {
        if (!(nb.flags & kActive)) return 0.0f;
        v3 dist = vsub(xyz(xi), xyz(nb.x));
        return nb.s * cubic_w(vlen_nr(dist), sm);
    };
    auto fold = [&](float t) { rho_temp += t; };
    loop(xi, load, term, fold);
    if (!lead) return;
    const float rho_i = rho_temp;
    float p = a.stiffness[i] * (fpow_(rho_i / fmax_(a.rest_density[i], 1.0e-6f), a.exponent[i]) - 1.0f);
    if (clamp) p = fmax_(p, 0.0f);
    a.rho[i] = rho_i;
    a.pressure[i] = p;
    a.sca[s] = make_float4(rho_i, p, p / fmax_(rho_i * rho_i, 1.0e-6f), 0.0f);
    a.vel[s].w = xi.w / fmax_(rho_i, 1.0e-6f);
}

template <int BLOCK, int CAP, int LPP, int U>
__device__ __forceinline__ void wcsph_density_lane(const WcsphDensityArgs& a, const PassCommon& c, int clamp, int s,
                                                   int* lds)
{
    wcsph_density_core<LPP>(a, c, clamp, s, [&](const float4& xi, auto& load, auto& term, auto& fold) {
        scan_pass<BLOCK, CAP, LPP, U>(c, s, xi.x, xi.y, xi.z, lds, NoCount{}, load, term, fold);
    });
}

// The keep radius of a scanning pass: the list's record radius when it
// records, else the pass cut (as in scan_neighbors / scan_split).
__device__ __forceinline__ float scan_keep2(const PassCommon& c)
{
    return c.L.mode == kListBuild ? fmax_(c.L.cut2, c.cut2) : c.cut2;
}

// Batched passes: slots are sorted by bucket and every scene owns a
// consecutive bucket range, so a wave almost always holds a single scene.
// The lanes take the wave's scenes one at a time (readfirstlane), each with
// its scene parameters uniform, so they stay in scalar registers as the
// plain pass's launch parameters do.  A scene with dt 0 is idle: its state,
// density, pressure and acceleration stay as they are (neighbours never
// cross scenes).
#define FOR_MY_SCENE(a, s, sc, ...)                                   \
    do {                                                              \
        const int mine_ = (a).sscene[s];                              \
        while (true) {                                                \
            const int sid_ = __builtin_amdgcn_readfirstlane(mine_);   \
            if (mine_ == sid_) {                                      \
                const fs_scene& sc = (a).scenes[sid_];                \
                __VA_ARGS__                                           \
                break;                                                \
            }                                                         \
        }                                                             \
    } while (0)

template <int BLOCK, int CAP, bool BATCH, int LPP, int U>
__global__ __launch_bounds__(BLOCK) void k_wcsph_density(WcsphDensityArgs a)
{
    __shared__ int lds[BLOCK / LPP * CAP];
    const int s = pass_slot_g<BLOCK, LPP>();
    if (s >= a.c.n) return;
    if constexpr (BATCH) {
        FOR_MY_SCENE(a, s, sc, {
            if (sc.dt != 0.0f) {
                PassCommon c = a.c;
                scene_view(c, sc);
                wcsph_density_lane<BLOCK, CAP, LPP, U>(a, c, sc.clamp_negative_pressure, s, lds);
            }
        });
    } else {
        if (a.dt_device && *a.dt_device == 0.0f) {   // idle substep: carry rho and p over
            if ((threadIdx.x & (LPP - 1)) == 0) {
                if (a.rho_in) a.rho[s] = a.rho_in[s];
                if (a.pressure_in) a.pressure[s] = a.pressure_in[s];
            }
            return;
        }
        wcsph_density_lane<BLOCK, CAP, LPP, U>(a, a.c, a.clamp, s, lds);
    }
}

// The density/EOS pass with the cooperative cell scan (fs_coop.h): a
// workgroup of 64 LPP threads per tile of 64 slots, NCH chunk slots of
// staged candidates and up to MAXC chunks per group.
template <int NCH, int MAXC, bool BATCH, int LPP, int U>
__global__ __launch_bounds__(kCoopTile * LPP) void k_wcsph_density_coop(WcsphDensityArgs a)
{
    __shared__ CoopShared<NCH, MAXC> sh;
    const int s0 = xcd_tile() * kCoopTile;
    if constexpr (!BATCH) {
        if (a.dt_device && *a.dt_device == 0.0f) {   // idle substep: carry rho and p over
            const int s = s0 + (int)threadIdx.x;
            if (threadIdx.x < kCoopTile && s < a.c.n) {
                if (a.rho_in) a.rho[s] = a.rho_in[s];
                if (a.pressure_in) a.pressure[s] = a.pressure_in[s];
            }
            return;
        }
    }
    const int tl = s0 + (int)(threadIdx.x & 63);
    const int sid = (BATCH && tl < a.c.n) ? a.sscene[tl] : 0;
    coop_test<NCH, MAXC, LPP>(a.c.pos, a.c.sflags, a.c.n, s0, sid,
                              [&](int id, GridView& g, float& r, float& keep2) -> bool {
                                  if constexpr (BATCH) {
                                      const fs_scene& sc = a.scenes[id];
                                      if (sc.dt == 0.0f) return false;
                                      PassCommon c = a.c;
                                      scene_view(c, sc);
                                      g = c.g;
                                      r = c.h;
                                      keep2 = scan_keep2(c);
                                  } else {
                                      g = a.c.g;
                                      r = a.c.h;
                                      keep2 = scan_keep2(a.c);
                                  }
                                  return true;
                              },
                              sh);
    const int p = (int)(threadIdx.x / LPP);
    const int s = s0 + p;
    if (s >= a.c.n) return;
    auto run = [&](const PassCommon& c, int clamp) {
        const float keep2 = scan_keep2(c);
        wcsph_density_core<LPP>(a, c, clamp, s, [&](const float4& xi, auto& load, auto& term, auto& fold) {
            coop_neighbors<NCH, MAXC, LPP, U>(sh, c.g, c.pos, p, s, xi.x, xi.y, xi.z, c.h, keep2, c.L, load, term,
                                              fold);
        });
    };
    if constexpr (BATCH) {
        FOR_MY_SCENE(a, s, sc, {
            if (sc.dt != 0.0f) {
                PassCommon c = a.c;
                scene_view(c, sc);
                run(c, sc.clamp_negative_pressure);
            }
        });
    } else {
        run(a.c, a.clamp);
    }
}

struct WcsphForceArgs {
    PassCommon c;
    const fs_scene* __restrict__ scenes;
    const int* __restrict__ sscene;
    const float4* __restrict__ vel;
    const float4* __restrict__ sca;
    const float* __restrict__ surface_tension;
    const float* __restrict__ viscosity;
    const float* __restrict__ gravity;
    const int* __restrict__ world;
    const float* __restrict__ x_in;
    const float* __restrict__ v_in;
    float particle_radius;
    float dt;
    float bound[3];
    float damping;
    const float* dt_device;   // plain pass: replaces dt when set
    float* accel;
    float* x_out;
    float* v_out;
};

// acceleration_wcsph + integrate_velocity + integrate_positions + apply_bounds
// solver_wcsph.py:161-305.  Per neighbour the operations of
// cal_acc_with_non_pressure and cal_acc_with_pressure (:63-106) are evaluated
// with the kernel gradient computed once and the per-particle quotients
// p/max(rho^2,1e-6), m/max(rho,1e-6) precomputed -- the same float operations
// on the same operands.  A zero cohesion (gamma) or viscosity (mu) coefficient
// contributes an exact zero, so its term is skipped.
// A neighbour's three updates of the particle's acceleration, applied in
// this order: cohesion (COH only: surface tension != 0), viscosity, pressure.
// With mu = 0 the viscosity term is an exact zero, and adding it leaves the
// sum unchanged (a round-to-nearest sum from +0 never becomes -0).
template <bool COH>
struct ForceTerms {
    v3 visc, pres;
};
template <>
struct ForceTerms<true> {
    v3 coh, visc, pres;
};

template <int BLOCK, int CAP, int UNR, int LPP, bool COH>
__device__ __forceinline__ v3 wcsph_force_sum(const WcsphForceArgs& a, const PassCommon& c, float pr, int s,
                                              v3 xi, v3 vi, float dp_i, float m_i, float mu, float gamma, int* lds)
{
    const float h = c.h;
    const Smooth sm = smooth_of(h);
    const float visc_den0 = 0.01f * (h * h);
    const float mu10 = 10.0f * mu;
    v3 acc = V(0.0f, 0.0f, 0.0f);
    const __amdgpu_buffer_rsrc_t rp = float4_rsrc(c.pos, c.n), rv = float4_rsrc(a.vel, c.n),
                                 rsc = float4_rsrc(a.sca, c.n);
    for_each_neighbor_split<BLOCK, CAP, UNR, LPP>(
        c.g, c.pos, s, xi.x, xi.y, xi.z, c.h, c.cut2, lds, c.L,
        [&](int k) {
            const int o = slot_off(k);
            return NbrXVS{c.sflags[k], ld_f4(rp, o, 0), ld_f4(rv, o, 0), ld_f32(rsc, o, 8)};
        },
        [&](int k, const NbrXVS& nb) -> ForceTerms<COH> {
            ForceTerms<COH> t;
            if constexpr (COH) t.coh = V(0.0f, 0.0f, 0.0f);
            t.visc = V(0.0f, 0.0f, 0.0f);
            t.pres = V(0.0f, 0.0f, 0.0f);
            if (k == s || !(nb.flags & kActive)) return t;
            const float4 xj = nb.x;
            const v3 d = vsub(xi, xyz(xj));
            const LenGrad lg = cubic_len_grad(d, sm);
            const float distance = lg.len;
            const float mass_nei = xj.w;
            const v3 grad = lg.grad;
            if constexpr (COH) t.coh = muls(smul(gamma / m_i * mass_nei, d), cubic_w(fmax_(distance, pr), sm));
            if (mu != 0.0f) {
                const float4 vj = nb.v;
                float v_current_nei = vdot(vsub(vi, xyz(vj)), d);
                float sv = mu10 * vj.w * v_current_nei / (distance * distance + visc_den0);
                t.visc = smul(sv, grad);
            }
            t.pres = smul(-mass_nei * (dp_i + nb.s), grad);
            return t;
        },
        [&](const ForceTerms<COH>& t) {
            if constexpr (COH) acc = vsub(acc, t.coh);
            acc = vadd(acc, t.visc);
            acc = vadd(acc, t.pres);
        });
    return acc;
}

// acceleration_wcsph + integrate_velocity + integrate_positions + apply_bounds
// solver_wcsph.py:161-305.  Per neighbour the operations of
// cal_acc_with_non_pressure and cal_acc_with_pressure (:63-106) are evaluated
// with the kernel gradient computed once and the per-particle quotients
// p/max(rho^2,1e-6), m/max(rho,1e-6) precomputed -- the same float operations
// on the same operands.  A zero cohesion (gamma) coefficient contributes an
// exact zero, so its term is skipped.  LPP lanes per particle share the
// neighbour terms; the group's first lane writes.
template <int BLOCK, int CAP, int UNR, int LPP>
__device__ __forceinline__ void wcsph_force_lane(const WcsphForceArgs& a, const PassCommon& c, float pr, float dt,
                                                 float dmp, float sx, float sy, float sz, int s, int* lds)
{
    const bool lead = (threadIdx.x & (LPP - 1)) == 0;
    const int i = c.perm[s];
    if (!(c.sflags[s] & kDynamic)) {
        if (lead) {
            st3(a.accel, i, V(0.0f, 0.0f, 0.0f));
            st3(a.v_out, i, ld3(a.v_in, i));
            st3(a.x_out, i, ld3(a.x_in, i));
        }
        return;
    }
    const float4 xi4 = c.pos[s];
    const v3 xi = xyz(xi4);
    const v3 vi = xyz(a.vel[s]);
    const float4 si = a.sca[s];
    const float dp_i = si.z, m_i = xi4.w;
    const float mu = a.viscosity[i], gamma = a.surface_tension[i];
    const v3 acc = gamma != 0.0f
                       ? wcsph_force_sum<BLOCK, CAP, UNR, LPP, true>(a, c, pr, s, xi, vi, dp_i, m_i, mu, gamma, lds)
                       : wcsph_force_sum<BLOCK, CAP, UNR, LPP, false>(a, c, pr, s, xi, vi, dp_i, m_i, mu, gamma, lds);
    if (!lead) return;
    int w = a.world[i];
    w = w > 0 ? w : 0;
    const v3 acc_g = vadd(acc, ld3(a.gravity, w));
    st3(a.accel, i, acc_g);
    v3 v = vadd(vi, smul(dt, acc_g));
    v3 x = vadd(xi, muls(v, dt));
    if (x.x < -sx) { x.x = -sx; if (v.x < 0.0f) v.x = v.x * dmp; }
    if (x.x > sx) { x.x = sx; if (v.x > 0.0f) v.x = v.x * dmp; }
    if (x.y < -sy) { x.y = -sy; if (v.y < 0.0f) v.y = v.y * dmp; }
    if (x.y > sy) { x.y = sy; if (v.y > 0.0f) v.y = v.y * dmp; }
    if (x.z < 0.0f) { x.z = 0.0f; if (v.z < 0.0f) v.z = v.z * dmp; }
    if (x.z > sz * 2.0f) { x.z = sz * 2.0f; if (v.z > 0.0f) v.z = v.z * dmp; }
    st3(a.v_out, i, v);
    st3(a.x_out, i, x);
}

// PPT particles per lane: the workgroup covers PPT consecutive tiles of BLOCK
// slots and each lane takes one slot of each, one after another.  LPP lanes
// per particle (PPT = 1): the workgroup covers BLOCK / LPP slots.
template <int BLOCK, int CAP, bool BATCH, int UNR, int PPT, int LPP>
__global__ __launch_bounds__(BLOCK) void k_wcsph_force(WcsphForceArgs a)
{
    static_assert(PPT == 1 || LPP == 1, "particles per lane or lanes per particle");
    __shared__ int lds[BLOCK * CAP];
    const int tile = xcd_tile();
    const bool lead = (threadIdx.x & (LPP - 1)) == 0;
    for (int q = 0; q < PPT; ++q) {
        const int s = LPP > 1 ? pass_slot_g<BLOCK, LPP>() : (tile * PPT + q) * BLOCK + (int)threadIdx.x;
        if (s >= a.c.n) return;
        if constexpr (BATCH) {
            FOR_MY_SCENE(a, s, sc, {
                if (sc.dt != 0.0f) {
                    PassCommon c = a.c;
                    scene_view(c, sc);
                    wcsph_force_lane<BLOCK, CAP, UNR, LPP>(a, c, sc.particle_radius, sc.dt, sc.boundary_damping,
                                                           sc.bound[0], sc.bound[1], sc.bound[2], s, lds);
                } else if (lead) {   // idle scene: carry the state over, keep its acceleration
                    const int i = a.c.perm[s];
                    st3(a.v_out, i, ld3(a.v_in, i));
                    st3(a.x_out, i, ld3(a.x_in, i));
                }
            });
        } else {
            const float dt = a.dt_device ? *a.dt_device : a.dt;
            if (dt == 0.0f) {   // idle substep (device dt 0): carry the state over
                if (lead) {
                    const int i = a.c.perm[s];
                    st3(a.v_out, i, ld3(a.v_in, i));
                    st3(a.x_out, i, ld3(a.x_in, i));
                }
                continue;
            }
            wcsph_force_lane<BLOCK, CAP, UNR, LPP>(a, a.c, a.particle_radius, dt, a.damping, a.bound[0], a.bound[1],
                                                   a.bound[2], s, lds);
        }
    }
}

// ---------------------------------------------------------------------------
// WCSPH adjoint: reverse-mode derivative of one substep (surface tension 0),
// in gather form so that every particle's adjoint is one lane's sum over its
// recorded neighbour list, in visit order, with no atomics.
// ---------------------------------------------------------------------------

struct WadjIntegrateArgs {
    int n;
    const int* __restrict__ perm;
    const int* __restrict__ sflags;
    const float* __restrict__ x_in;
    const float* __restrict__ v_in;
    const float* __restrict__ accel;
    const float* __restrict__ xbar_out;
    const float* __restrict__ vbar_out;
    const float* __restrict__ viscosity;
    float dt;
    float bound[3];
    float damping;
    float4* abar;   // sorted: dL/da, 10*viscosity
    float4* xbar;   // sorted partial dL/dx_in
    float4* vbar;   // sorted partial dL/dv_in
};

// reverse of integrate_velocity, integrate_positions, apply_bounds
// (solver_wcsph.py:231-305)
__global__ __launch_bounds__(kPassBlock) void k_wadj_integrate(WadjIntegrateArgs a)
{
    const int s = pass_slot();
    if (s >= a.n) return;
    const int i = a.perm[s];
    const v3 xb = ld3(a.xbar_out, i), vb = ld3(a.vbar_out, i);
    const float c = 10.0f * a.viscosity[i];
    if (!(a.sflags[s] & kDynamic)) {
        a.xbar[s] = make_float4(xb.x, xb.y, xb.z, 0.0f);
        a.vbar[s] = make_float4(vb.x, vb.y, vb.z, 0.0f);
        a.abar[s] = make_float4(0.0f, 0.0f, 0.0f, c);
        return;
    }
    const v3 u = vadd(ld3(a.v_in, i), smul(a.dt, ld3(a.accel, i)));
    const v3 p = vadd(ld3(a.x_in, i), muls(u, a.dt));
    const float lo[3] = {-a.bound[0], -a.bound[1], 0.0f};
    const float hi[3] = {a.bound[0], a.bound[1], a.bound[2] * 2.0f};
    const float pc[3] = {p.x, p.y, p.z}, uc[3] = {u.x, u.y, u.z};
    float xbc[3] = {xb.x, xb.y, xb.z}, vbc[3] = {vb.x, vb.y, vb.z};
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        if (pc[d] < lo[d]) {
            xbc[d] = 0.0f;
            if (uc[d] < 0.0f) vbc[d] = vbc[d] * a.damping;
        } else if (pc[d] > hi[d]) {
            xbc[d] = 0.0f;
            if (uc[d] > 0.0f) vbc[d] = vbc[d] * a.damping;
        }
    }
    const v3 xbu = V(xbc[0], xbc[1], xbc[2]);
    const v3 vbu = vadd(V(vbc[0], vbc[1], vbc[2]), smul(a.dt, xbu));
    const v3 ab = smul(a.dt, vbu);
    a.xbar[s] = make_float4(xbu.x, xbu.y, xbu.z, 0.0f);
    a.vbar[s] = make_float4(vbu.x, vbu.y, vbu.z, 0.0f);
    a.abar[s] = make_float4(ab.x, ab.y, ab.z, c);
}

struct WadjForceArgs {
    PassCommon c;
    const float4* __restrict__ vel;    // v, m/max(rho,1e-6)
    const float4* __restrict__ sca;    // rho, p, p/max(rho^2,1e-6), -
    const float4* __restrict__ abar;
    const float* __restrict__ rest_density;
    const float* __restrict__ stiffness;
    const float* __restrict__ exponent;
    int clamp;
    float4* xbar;
    float4* vbar;
    float* rhobar;                     // sorted dL/drho
    float* grad_stiffness;             // original order, +=
    float* grad_viscosity;
};

// a neighbour's updates of the reverse force sums, applied in this order:
// xb += x1, xb += x2, Pb += P, xb += x3, xb -= x4, vb += v, mub += mu, wb += w
struct AdjForceTerms {
    v3 x1, x2, x3, x4, v;
    float P, mu, w;
};

// lanes per particle of the adjoint passes (few particles: more waves)
constexpr int kAdjLpp = 4;

// reverse of acceleration_wcsph (pressure + viscosity) and of the EOS
// (solver_wcsph.py:92-106,119-228): per particle k, the pairs (k,j) of its
// own acceleration and (j,k) where it is the neighbour.  LPP lanes per
// particle share the neighbour terms; the group's first lane writes.
template <int CAP, int LPP>
__global__ __launch_bounds__(kPassBlock) void k_wadj_force(WadjForceArgs a)
{
    __shared__ int lds[kPassBlock * CAP];
    const int s = pass_slot_g<kPassBlock, LPP>();
    const bool lead = (threadIdx.x & (LPP - 1)) == 0;
    if (s >= a.c.n) return;
    const PassCommon& c = a.c;
    if (!(c.sflags[s] & kActive)) {
        if (lead) a.rhobar[s] = 0.0f;
        return;
    }
    const int i = c.perm[s];
    const float4 xk4 = c.pos[s];
    const float4 vk4 = a.vel[s];
    const float4 sk = a.sca[s];
    const float4 ak4 = a.abar[s];
    const v3 xk = xyz(xk4), vk = xyz(vk4), ak = xyz(ak4);
    const float rho_k = sk.x, p_k = sk.y, P_k = sk.z, m_k = xk4.w, w_k = vk4.w, c_k = ak4.w;
    const float h = c.h;
    const float eps2 = 0.01f * (h * h);
    v3 xb = V(0.0f, 0.0f, 0.0f), vb = V(0.0f, 0.0f, 0.0f);
    float Pb = 0.0f, wb = 0.0f, mub = 0.0f;
    // skipped neighbours give zero terms, which leave every sum unchanged
    for_each_neighbor_split<kPassBlock, CAP, 2, LPP>(
        c.g, c.pos, s, xk.x, xk.y, xk.z, c.h, c.cut2, lds, c.L,
        [&](int j) { return AdjNbr{c.sflags[j], c.pos[j], a.abar[j], a.vel[j], a.sca[j].z}; },
        [&](int j, const AdjNbr& nb) -> AdjForceTerms {
            const v3 z = V(0.0f, 0.0f, 0.0f);
            AdjForceTerms t{z, z, z, z, z, 0.0f, 0.0f, 0.0f};
            if (j == s || !(nb.flags & kActive)) return t;
            const float4 xj4 = nb.x;
            const v3 r = vsub(xk, xyz(xj4));
            const float d = vlen_nr(r);
            if (!(d / h <= 1.0f)) return t;
            const v3 G = cubic_grad_r(r, d, h);
            const float4 aj4 = nb.abar;
            const float4 vj4 = nb.v;
            const v3 aj = xyz(aj4);
            const float m_j = xj4.w;
            const float P_j = nb.P;
            const float w_j = vj4.w;
            const float c_j = aj4.w;
            const float den = d * d + eps2;
            const v3 dv = vsub(vk, xyz(vj4));
            const float sv = vdot(dv, r) / den;
            const float Gak = vdot(G, ak);
            const float Gaj = vdot(G, aj);
            const v3 Hak = cubic_hess_mul(r, d, h, ak);
            const v3 Haj = cubic_hess_mul(r, d, h, aj);
            const v3 ds = vdivs(vsub(dv, smul(2.0f * sv, r)), den);
            t.x1 = smul(-m_j * (P_k + P_j), Hak);
            t.x2 = smul(m_k * (P_j + P_k), Haj);
            t.P = -m_j * Gak + m_k * Gaj;
            const float ck_wj = c_k * w_j;
            const float cj_wk = c_j * w_k;
            t.x3 = smul(ck_wj, vadd(smul(Gak, ds), smul(sv, Hak)));
            t.x4 = smul(cj_wk, vadd(smul(Gaj, ds), smul(sv, Haj)));
            t.v = smul((ck_wj * Gak - cj_wk * Gaj) / den, r);
            t.mu = 10.0f * w_j * sv * Gak;
            t.w = -c_j * sv * Gaj;
            return t;
        },
        [&](const AdjForceTerms& t) {
            xb = vadd(xb, t.x1);
            xb = vadd(xb, t.x2);
            Pb += t.P;
            xb = vadd(xb, t.x3);
            xb = vsub(xb, t.x4);
            vb = vadd(vb, t.v);
            mub += t.mu;
            wb += t.w;
        });
    if (!lead) return;
    float rb = 0.0f;
    const float pb = Pb / fmax_(rho_k * rho_k, 1.0e-6f);
    if (rho_k * rho_k > 1.0e-6f) rb += Pb * (-2.0f * p_k / (rho_k * rho_k * rho_k));
    if (rho_k > 1.0e-6f) rb += wb * (-m_k / (rho_k * rho_k));
    const float rd = fmax_(a.rest_density[i], 1.0e-6f);
    const float ratio = rho_k / rd;
    const float gam = a.exponent[i];
    const float B = a.stiffness[i];
    const float p_raw = B * (fpow_(ratio, gam) - 1.0f);
    float sb = 0.0f;
    if (!a.clamp || p_raw > 0.0f) {
        rb += pb * B * gam * fpow_(ratio, gam - 1.0f) / rd;
        sb = pb * (fpow_(ratio, gam) - 1.0f);
    }
    a.rhobar[s] = rb;
    const float4 x0 = a.xbar[s], v0 = a.vbar[s];
    a.xbar[s] = make_float4(x0.x + xb.x, x0.y + xb.y, x0.z + xb.z, 0.0f);
    a.vbar[s] = make_float4(v0.x + vb.x, v0.y + vb.y, v0.z + vb.z, 0.0f);
    a.grad_stiffness[i] += sb;
    a.grad_viscosity[i] += mub;
}

struct WadjDensityArgs {
    PassCommon c;
    const float* __restrict__ wvol;
    const float* __restrict__ rhobar;
    const float4* __restrict__ xbar_s;
    const float4* __restrict__ vbar_s;
    float* xbar_in;
    float* vbar_in;
};

// reverse of the density sum of compute_density_pressure (solver_wcsph.py:119-158)
// and the final write of dL/dx_in, dL/dv_in
template <int CAP, int LPP>
__global__ __launch_bounds__(kPassBlock) void k_wadj_density(WadjDensityArgs a)
{
    __shared__ int lds[kPassBlock * CAP];
    const int s = pass_slot_g<kPassBlock, LPP>();
    if (s >= a.c.n) return;
    const PassCommon& c = a.c;
    const int i = c.perm[s];
    const float4 x0 = a.xbar_s[s];
    v3 xb = V(0.0f, 0.0f, 0.0f);
    if (c.sflags[s] & kActive) {
        const float4 xk4 = c.pos[s];
        const v3 xk = xyz(xk4);
        const float wv_k = a.wvol[s];
        const float rb_k = a.rhobar[s];
        const float h = c.h;
        for_each_neighbor_split<kPassBlock, CAP, kReplayUnroll, LPP>(
            c.g, c.pos, s, xk.x, xk.y, xk.z, c.h, c.cut2, lds, c.L,
            [&](int j) { return NbrXV{c.sflags[j], c.pos[j], make_float4(a.wvol[j], a.rhobar[j], 0.0f, 0.0f)}; },
            [&](int j, const NbrXV& nb) -> v3 {
                if (j == s || !(nb.flags & kActive)) return V(0.0f, 0.0f, 0.0f);
                const v3 r = vsub(xk, xyz(nb.x));
                const float d = vlen_nr(r);
                if (!(d / h <= 1.0f)) return V(0.0f, 0.0f, 0.0f);
                return smul(rb_k * nb.v.x + nb.v.y * wv_k, cubic_grad_r(r, d, h));
            },
            [&](const v3& t) { xb = vadd(xb, t); });
    }
    if ((threadIdx.x & (LPP - 1)) != 0) return;
    st3(a.xbar_in, i, V(x0.x + xb.x, x0.y + xb.y, x0.z + xb.z));
    const float4 v0 = a.vbar_s[s];
    st3(a.vbar_in, i, V(v0.x, v0.y, v0.z));
}

// ---------------------------------------------------------------------------
// DFSPH
// ---------------------------------------------------------------------------

struct DfDensityArgs {
    PassCommon c;
    const float* __restrict__ wvol;
    int min_neighbors;
    const float* __restrict__ kappa_warm;   // warm start source (state_in.sph.kappa / kappa_v) or null
    float* kappa_commit;                    // state_out.sph.kappa / kappa_v when warm starting
    float* rho;
    float* factor;
    int* neighbor_count;
    float4* sca;   // rho, factor, kappa, kappa/max(rho,1e-6)
    float4* vel;   // .w <- m/max(rho,1e-6)
    const int* any_inactive;   // set by the grid build when some particle is inactive
    const float* dt_device;    // idle substep (device dt 0): nothing to do
};

// compute_density + compute_factor_dfsph  solver_dfsph.py:106-174 (one pass)
// a neighbour's contributions to the density, the gradient sum and the sum
// of squared gradients
struct FactorTerms {
    float rho;
    v3 grad;
    float grad_sq;
};

template <int BLOCK, int CAP, int LPP, int U>
__global__ __launch_bounds__(BLOCK) void k_dfsph_density_factor(DfDensityArgs a)
{
    __shared__ int lds[BLOCK / LPP * CAP];
    const int s = pass_slot_g<BLOCK, LPP>();
    const bool lead = (threadIdx.x & (LPP - 1)) == 0;
    if (s >= a.c.n) return;
    float dt_unused;
    if (!substep_dt(a.dt_device, 1.0f, dt_unused)) return;
    const int i = a.c.perm[s];
    const float4 xi4 = a.c.pos[s];
    if (!(a.c.sflags[s] & kActive)) {
        if (lead) {
            a.rho[i] = 0.0f;
            a.factor[i] = 0.0f;
            a.neighbor_count[i] = 0;
            float kw = a.kappa_warm ? a.kappa_warm[i] : 0.0f;
            if (a.kappa_warm) a.kappa_commit[i] = kw;
            a.sca[s] = make_float4(0.0f, 0.0f, kw, kw / fmax_(0.0f, 1.0e-6f));
            a.vel[s].w = xi4.w / fmax_(0.0f, 1.0e-6f);
            if (a.c.L.mode == kListBuild) a.c.L.cnt[s] = 0;
        }
        return;
    }
    const v3 xi = xyz(xi4);
    const float h = a.c.h;
    const Smooth sm = smooth_of(h);
    float rho_temp = 0.0f;
    v3 sum_grad = V(0.0f, 0.0f, 0.0f);
    float sum_grad_sq = 0.0f;
    auto load = [&](int k) { return NbrXS{a.c.sflags[k], a.c.pos[k], a.wvol[k]}; };
    auto term = [&](int k, const NbrXS& nb) -> FactorTerms {
        FactorTerms t{0.0f, V(0.0f, 0.0f, 0.0f), 0.0f};
        if (!(nb.flags & kActive)) return t;
        const float4 xj = nb.x;
        v3 r = vsub(xi, xyz(xj));
        const LenGrad lg = cubic_len_grad(r, sm);
        t.rho = nb.s * cubic_w(lg.len, sm);
        t.grad = smul(xj.w, lg.grad);
        t.grad_sq = vdot(t.grad, t.grad);
        return t;
    };
    auto fold = [&](const FactorTerms& t) {
        rho_temp += t.rho;
        sum_grad = vadd(sum_grad, t.grad);
        sum_grad_sq += t.grad_sq;
    };
    // count = active candidates other than the particle itself, no distance
    // test (solver_dfsph.py:150-160)
    int count = 0;
    // A query range repeats a bucket only where an axis has two or more cells
    // at or below -kOrigin (they all hash to bucket 0 of that axis), i.e.
    // where its first cell lies below -kOrigin; the particle's own bucket may
    // then be visited more than once, and the reference skips it every time.
    const GridView& gv = a.c.g;
    const bool repeats = cell_coord((xi.x - h) * gv.inv, gv.round_mode) < -kOrigin ||
                         cell_coord((xi.y - h) * gv.inv, gv.round_mode) < -kOrigin ||
                         cell_coord((xi.z - h) * gv.inv, gv.round_mode) < -kOrigin;
    if (*a.any_inactive || repeats) {
        auto counter = [&](int k) { count += (k != s && (a.c.sflags[k] & kActive)) ? 1 : 0; };
        scan_pass<BLOCK, CAP, LPP, U>(a.c, s, xi.x, xi.y, xi.z, lds, counter, load, term, fold);
        count = group_sum<LPP>(count);   // each lane counted the candidates it tested
    } else {
        // every particle active and no bucket repeated: all candidates but
        // the particle itself, which is among its own candidates exactly once
        // (its bucket lies in the query range)
        const int total = scan_pass<BLOCK, CAP, LPP, U>(a.c, s, xi.x, xi.y, xi.z, lds, NoCount{}, load, term, fold);
        count = total > 0 ? total - 1 : 0;
    }
    if (!lead) return;
    const float denom = vdot(sum_grad, sum_grad) + sum_grad_sq;
    const float f = (count < a.min_neighbors || denom < 1.0e-6f) ? 0.0f : 1.0f / denom;
    a.rho[i] = rho_temp;
    a.factor[i] = f;
    a.neighbor_count[i] = count;
    float kw = 0.0f;
    if (a.kappa_warm) {
        kw = a.kappa_warm[i];
        a.kappa_commit[i] = kw;
    }
    a.sca[s] = make_float4(rho_temp, f, kw, kw / fmax_(rho_temp, 1.0e-6f));
    a.vel[s].w = xi4.w / fmax_(rho_temp, 1.0e-6f);
}

struct DfNonPressureArgs {
    PassCommon c;
    const float4* __restrict__ vel;
    const float* __restrict__ surface_tension;
    const float* __restrict__ viscosity;
    const float* __restrict__ gravity;
    const int* __restrict__ world;
    float particle_radius;
    float dt;
    const float* dt_device;
    float* accel;
    float* v_star;
    float4* vstar_s;
};

// a neighbour's updates of the non-pressure acceleration, in this order:
// cohesion (COH only: surface tension != 0), then viscosity
template <bool COH>
struct NonPressureTerms {
    v3 visc;
};
template <>
struct NonPressureTerms<true> {
    v3 coh, visc;
};

template <int BLOCK, int CAP, int UNR, int LPP, bool COH>
__device__ __forceinline__ v3 df_nonpressure_sum(const DfNonPressureArgs& a, int s, v3 xi, v3 vi, float m_i,
                                                 float mu, float gamma, int* lds)
{
    const float h = a.c.h, pr = a.particle_radius;
    const Smooth sm = smooth_of(h);
    const float visc_den0 = 0.01f * (h * h);
    const float mu10 = 10.0f * mu;
    v3 acc = V(0.0f, 0.0f, 0.0f);
    // zero terms (skipped neighbours, mu = 0) leave the sums unchanged
    for_each_neighbor_split<BLOCK, CAP, UNR, LPP>(
        a.c.g, a.c.pos, s, xi.x, xi.y, xi.z, a.c.h, a.c.cut2, lds, a.c.L,
        [&](int k) { return NbrXV{a.c.sflags[k], a.c.pos[k], a.vel[k]}; },
        [&](int k, const NbrXV& nb) -> NonPressureTerms<COH> {
            NonPressureTerms<COH> t;
            if constexpr (COH) t.coh = V(0.0f, 0.0f, 0.0f);
            t.visc = V(0.0f, 0.0f, 0.0f);
            if (k == s || !(nb.flags & kActive)) return t;
            const float4 xj = nb.x;
            const v3 d = vsub(xi, xyz(xj));
            const LenGrad lg = cubic_len_grad(d, sm);
            const float dist = lg.len;
            if constexpr (COH) t.coh = muls(smul(gamma / m_i * xj.w, d), cubic_w(fmax_(dist, pr), sm));
            if (mu != 0.0f) {
                const float4 vj = nb.v;
                float v_current_nei = vdot(vsub(vi, xyz(vj)), d);
                float sv = mu10 * vj.w * v_current_nei / (dist * dist + visc_den0);
                t.visc = smul(sv, lg.grad);
            }
            return t;
        },
        [&](const NonPressureTerms<COH>& t) {
            if constexpr (COH) acc = vsub(acc, t.coh);
            acc = vadd(acc, t.visc);
        });
    return acc;
}

// acceleration_non_pressure + predict_velocity  solver_dfsph.py:177-242
template <int BLOCK, int CAP, int UNR, int LPP>
__global__ __launch_bounds__(BLOCK) void k_dfsph_nonpressure(DfNonPressureArgs a)
{
    __shared__ int lds[BLOCK * CAP];
    const int s = pass_slot_g<BLOCK, LPP>();
    const bool lead = (threadIdx.x & (LPP - 1)) == 0;
    if (s >= a.c.n) return;
    float dt;
    if (!substep_dt(a.dt_device, a.dt, dt)) return;
    const int i = a.c.perm[s];
    if (!(a.c.sflags[s] & kDynamic)) {
        if (lead) {
            st3(a.accel, i, V(0.0f, 0.0f, 0.0f));
            st3(a.v_star, i, V(0.0f, 0.0f, 0.0f));
            a.vstar_s[s] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        }
        return;
    }
    const float4 xi4 = a.c.pos[s];
    const v3 xi = xyz(xi4);
    const v3 vi = xyz(a.vel[s]);
    const float m_i = xi4.w, mu = a.viscosity[i], gamma = a.surface_tension[i];
    v3 acc = V(0.0f, 0.0f, 0.0f);
    if (gamma != 0.0f) acc = df_nonpressure_sum<BLOCK, CAP, UNR, LPP, true>(a, s, xi, vi, m_i, mu, gamma, lds);
    else if (mu != 0.0f) acc = df_nonpressure_sum<BLOCK, CAP, UNR, LPP, false>(a, s, xi, vi, m_i, mu, gamma, lds);
    if (!lead) return;
    int w = a.world[i];
    w = w > 0 ? w : 0;
    const v3 acc_g = vadd(acc, ld3(a.gravity, w));
    st3(a.accel, i, acc_g);
    const v3 vs = vadd(vi, smul(dt, acc_g));
    st3(a.v_star, i, vs);
    a.vstar_s[s] = make_float4(vs.x, vs.y, vs.z, 0.0f);
}

struct DfRateArgs {
    PassCommon c;
    const float4* __restrict__ vstar_s;
    const float* __restrict__ rest_density;
    float dt;
    const float* dt_device;
    float max_kappa;
    int density_loop;   // 1: density loop (rho_star, kappa), 0: divergence loop (kappa_v)
    float* d_rho_dt;
    float* rho_star;
    float4* sca;
    double* partials;
    int* control;
};

// compute_density_change_rate (+ predict_density_star, compute_kappa_density
// or compute_kappa_divergence, and the host residual)  solver_dfsph.py:245-306,
// 356-372, 579-596
template <int BLOCK, int CAP, int UNR, int LPP>
__global__ __launch_bounds__(BLOCK) void k_dfsph_rate(DfRateArgs a)
{
    __shared__ int lds[BLOCK * CAP];
    const int stop_idx = a.density_loop ? kStopRho : kStopVel;
    if (a.control[stop_idx]) return;
    float dt;
    if (!substep_dt(a.dt_device, a.dt, dt)) return;
    const int s = pass_slot_g<BLOCK, LPP>();
    const bool lead = (threadIdx.x & (LPP - 1)) == 0;   // the lane that writes the particle's results
    double res_v = 0.0, res_c = 0.0;
    if (s < a.c.n) {
        const int i = a.c.perm[s];
        const int fl = a.c.sflags[s];
        float4 si = a.sca[s];
        if (!(fl & kActive)) {
            if (lead) {
                a.d_rho_dt[i] = 0.0f;
                if (a.density_loop) a.rho_star[i] = 0.0f;
                si.z = 0.0f;
                si.w = 0.0f / fmax_(si.x, 1.0e-6f);
                a.sca[s] = si;
            }
        } else {
            const v3 xi = xyz(a.c.pos[s]);
            const v3 vi = xyz(a.vstar_s[s]);
            const Smooth sm = smooth_of(a.c.h);
            float div_val = 0.0f;
            const __amdgpu_buffer_rsrc_t rp = float4_rsrc(a.c.pos, a.c.n), rv = float4_rsrc(a.vstar_s, a.c.n);
            // a skipped (inactive) neighbour contributes +0: div_val starts at
            // +0 and a round-to-nearest sum from +0 never becomes -0, so adding
            // +0 leaves it unchanged
            for_each_neighbor_split<BLOCK, CAP, UNR, LPP>(
                a.c.g, a.c.pos, s, xi.x, xi.y, xi.z, a.c.h, a.c.cut2, lds, a.c.L,
                [&](int k) {
                    const int o = slot_off(k);
                    return NbrXV{a.c.sflags[k], ld_f4(rp, o, 0), ld_f4(rv, o, 0)};
                },
                [&](int k, const NbrXV& nb) -> float {
                    if (!(nb.flags & kActive)) return 0.0f;
                    const float4 xj = nb.x;
                    v3 grad_w = cubic_len_grad(vsub(xi, xyz(xj)), sm).grad;
                    return xj.w * vdot(vsub(vi, xyz(nb.v)), grad_w);
                },
                [&](float t) { div_val += t; });
            if (lead) {
                a.d_rho_dt[i] = div_val;
                const bool dyn = (fl & kDynamic) != 0;
                float kap = 0.0f;
                if (a.density_loop) {
                    const float rs = si.x + dt * div_val;
                    a.rho_star[i] = rs;
                    const float diff = rs - a.rest_density[i];
                    const float err = fmax_(diff, 0.0f);
                    if (dyn) {
                        res_v = (double)np_max0(diff);
                        res_c = 1.0;
                        kap = fmin_((err / (dt * dt)) * si.y, a.max_kappa);
                    }
                } else {
                    const float dc = fmax_(div_val, 0.0f);
                    if (dyn) {
                        res_v = (double)np_max0(div_val);
                        res_c = 1.0;
                        kap = fmin_((1.0f / dt) * dc * si.x * si.y, a.max_kappa);
                    }
                }
                si.z = kap;
                si.w = kap / fmax_(si.x, 1.0e-6f);
                a.sca[s] = si;
            }
        }
    }
    residual_partials<BLOCK>(res_v, res_c, a.partials);
}

struct DfCorrArgs {
    PassCommon c;
    const float4* __restrict__ sca;
    float dt;
    const float* dt_device;
    int stop_idx;   // -1: unconditional (warm start)
    int corr_idx;
    float* v_star;
    float4* vstar_s;
    float* kappa_commit;
    int* control;
};

// apply_constant_density_correction / apply_divergence_free_correction
// solver_dfsph.py:309-337, 375-403 (and the kappa write of :290-306, 356-372)
template <int BLOCK, int CAP, int UNR, int LPP>
__global__ __launch_bounds__(BLOCK) void k_dfsph_correction(DfCorrArgs a)
{
    __shared__ int lds[BLOCK * CAP];
    if (a.stop_idx >= 0 && a.control[a.stop_idx]) return;
    float dt;
    if (!substep_dt(a.dt_device, a.dt, dt)) return;
    const int s = pass_slot_g<BLOCK, LPP>();
    const bool lead = (threadIdx.x & (LPP - 1)) == 0;
    if (blockIdx.x == 0 && threadIdx.x == 0 && a.corr_idx >= 0) a.control[a.corr_idx] += 1;
    if (s >= a.c.n) return;
    const int i = a.c.perm[s];
    const float4 si = a.sca[s];
    if (a.kappa_commit && lead) a.kappa_commit[i] = si.z;
    if (!(a.c.sflags[s] & kDynamic)) return;
    const v3 xi = xyz(a.c.pos[s]);
    const Smooth sm = smooth_of(a.c.h);
    const float kri = si.w;
    v3 dv = V(0.0f, 0.0f, 0.0f);
    const __amdgpu_buffer_rsrc_t rp = float4_rsrc(a.c.pos, a.c.n), rsc = float4_rsrc(a.sca, a.c.n);
    // skipped neighbours contribute +0 (see k_dfsph_rate)
    for_each_neighbor_split<BLOCK, CAP, UNR, LPP>(
        a.c.g, a.c.pos, s, xi.x, xi.y, xi.z, a.c.h, a.c.cut2, lds, a.c.L,
        [&](int k) {
            const int o = slot_off(k);
            return NbrXS{a.c.sflags[k], ld_f4(rp, o, 0), ld_f32(rsc, o, 12)};
        },
        [&](int k, const NbrXS& nb) -> v3 {
            if (!(nb.flags & kActive)) return V(0.0f, 0.0f, 0.0f);
            const float4 xj = nb.x;
            float coeff = xj.w * (kri + nb.s);
            return smul(coeff, cubic_len_grad(vsub(xi, xyz(xj)), sm).grad);
        },
        [&](const v3& t) { dv = vadd(dv, t); });
    if (!lead) return;
    const float4 vs4 = a.vstar_s[s];
    const v3 vs = vsub(xyz(vs4), smul(dt, dv));
    st3(a.v_star, i, vs);
    a.vstar_s[s] = make_float4(vs.x, vs.y, vs.z, 0.0f);
}

struct DfFinalArgs {
    int n;
    const int* __restrict__ flags;
    const int* __restrict__ mask;
    const float* __restrict__ v_star;
    float* x_out;
    float* v_out;
    float bound[3];
    float damping;
    const int* control;
    int* iterations;
    // idle substep (device dt 0): carry the input state over
    const float* dt_device;
};

// state_in -> state_out of an idle substep (pointers may be null)
struct DfCarry {
    const float* x_in;
    const float* v_in;
    const float* rho_in;
    const float* factor_in;
    const float* rho_star_in;
    const float* d_rho_dt_in;
    const float* kappa_in;
    const float* kappa_v_in;
    const int* neighbor_count_in;
    const float* v_star_in;
    float* rho;
    float* factor;
    float* rho_star;
    float* d_rho_dt;
    float* kappa;
    float* kappa_v;
    int* neighbor_count;
    float* v_star;
};

// copy v_star -> particle_qd, apply_bounds (DFSPH variant)  solver_dfsph.py:406-442,885-898
__global__ __launch_bounds__(kElemBlock) void k_dfsph_finalize(DfFinalArgs a, DfCarry c)
{
    const int t = blockIdx.x * kElemBlock + threadIdx.x;
    if (a.iterations && t < 4) a.iterations[t] = a.control[kItRho + t];
    if (t >= a.n) return;
    float dt_unused;
    if (!substep_dt(a.dt_device, 1.0f, dt_unused)) {
        st3(a.x_out, t, ld3(c.x_in, t));
        st3(a.v_out, t, ld3(c.v_in, t));
        if (c.rho_in) c.rho[t] = c.rho_in[t];
        if (c.factor_in) c.factor[t] = c.factor_in[t];
        if (c.rho_star_in) c.rho_star[t] = c.rho_star_in[t];
        if (c.d_rho_dt_in) c.d_rho_dt[t] = c.d_rho_dt_in[t];
        if (c.kappa_in) c.kappa[t] = c.kappa_in[t];
        if (c.kappa_v_in) c.kappa_v[t] = c.kappa_v_in[t];
        if (c.neighbor_count_in) c.neighbor_count[t] = c.neighbor_count_in[t];
        if (c.v_star_in) st3(c.v_star, t, ld3(c.v_star_in, t));
        return;
    }
    v3 v = ld3(a.v_star, t);
    if (!((a.flags[t] & 1) != 0 && a.mask[t] == 0)) {
        st3(a.v_out, t, v);
        return;
    }
    v3 x = ld3(a.x_out, t);
    const float sx = a.bound[0], sy = a.bound[1], sz = a.bound[2], d = a.damping;
    if (x.x < -sx) { x.x = -sx; v.x = v.x * d; }
    if (x.x > sx) { x.x = sx; v.x = v.x * d; }
    if (x.y < -sy) { x.y = -sy; v.y = v.y * d; }
    if (x.y > sy) { x.y = sy; v.y = v.y * d; }
    if (x.z < 0.0f) { x.z = 0.0f; v.z = v.z * d; }
    if (x.z > sz * 2.0f) { x.z = sz * 2.0f; v.z = v.z * d; }
    st3(a.x_out, t, x);
    st3(a.v_out, t, v);
}

// ---------------------------------------------------------------------------
// PBF
// ---------------------------------------------------------------------------

struct PbfBounds {
    float bound[3];
    float particle_radius;
    float boundary_epsilon;
    __device__ __forceinline__ v3 clamp(v3 x) const
    {
        // apply_position_bounds  solver_pbf.py:366-399
        float r = particle_radius, e = boundary_epsilon;
        float x_min = -bound[0] + r + e, x_max = bound[0] - r - e;
        float y_min = -bound[1] + r + e, y_max = bound[1] - r - e;
        float z_min = r + e, z_max = bound[2] * 2.0f - r - e;
        if (x.x < x_min) x.x = x_min;
        if (x.x > x_max) x.x = x_max;
        if (x.y < y_min) x.y = y_min;
        if (x.y > y_max) x.y = y_max;
        if (x.z < z_min) x.z = z_min;
        if (x.z > z_max) x.z = z_max;
        return x;
    }
    // apply_bounds  solver_pbf.py:466-517
    __device__ __forceinline__ void clamp_damp(v3& x, v3& v, float d) const
    {
        float r = particle_radius, e = boundary_epsilon;
        float x_min = -bound[0] + r + e, x_max = bound[0] - r - e;
        float y_min = -bound[1] + r + e, y_max = bound[1] - r - e;
        float z_min = r + e, z_max = bound[2] * 2.0f - r - e;
        if (x.x < x_min) { x.x = x_min; if (v.x < 0.0f) v.x = v.x * d; }
        if (x.x > x_max) { x.x = x_max; if (v.x > 0.0f) v.x = v.x * d; }
        if (x.y < y_min) { x.y = y_min; if (v.y < 0.0f) v.y = v.y * d; }
        if (x.y > y_max) { x.y = y_max; if (v.y > 0.0f) v.y = v.y * d; }
        if (x.z < z_min) { x.z = z_min; if (v.z < 0.0f) v.z = v.z * d; }
        if (x.z > z_max) { x.z = z_max; if (v.z > 0.0f) v.z = v.z * d; }
    }
};

struct PbfDensityArgs {
    PassCommon c;
    const float* __restrict__ wvol;
    float* rho;
    float4* sca;
    // final-density mode: update_velocity_from_positions (+ finalize when no XSPH)
    int final_pass;
    int finalize;
    const float* __restrict__ x_in;
    const float* __restrict__ v_in;
    float dt;
    float damping;
    PbfBounds bounds;
    float4* vel;
    float* x_out;
    float* v_out;
};

// compute_density (poly6)  solver_pbf.py:114-141; in the final pass also
// update_velocity_from_positions :402-416 and, without XSPH, :939-954.
// vel[s].w <- m/max(rho,1e-6) for the non-pressure and XSPH passes.
template <int BLOCK, int CAP, int LPP, int U>
__global__ __launch_bounds__(BLOCK) void k_pbf_density(PbfDensityArgs a)
{
    __shared__ int lds[BLOCK / LPP * CAP];
    PassCommon c = a.c;
    const int s = pass_slot_g<BLOCK, LPP>();
    const bool lead = (threadIdx.x & (LPP - 1)) == 0;
    if (s >= c.n) return;
    const int i = c.perm[s];
    const int fl = c.sflags[s];
    const float4 xi4 = c.pos[s];
    const v3 xi = xyz(xi4);
    float rho_temp = 0.0f;
    if (fl & kActive) {
        const float h = c.h;
        scan_pass<BLOCK, CAP, LPP, U>(
            c, s, xi.x, xi.y, xi.z, lds, NoCount{}, [&](int k) { return NbrXS{c.sflags[k], c.pos[k], a.wvol[k]}; },
            [&](int k, const NbrXS& nb) -> float {
                if (!(nb.flags & kActive)) return 0.0f;
                v3 r = vsub(xi, xyz(nb.x));
                return nb.s * poly6_w(vlen_nr(r), h);
            },
            [&](float t) { rho_temp += t; });
    } else if (c.L.mode == kListBuild && lead) {
        c.L.cnt[s] = 0;
    }
    if (!lead) return;
    a.rho[i] = rho_temp;
    a.sca[s] = make_float4(rho_temp, 0.0f, 0.0f, 0.0f);
    const float mr = xi4.w / fmax_(rho_temp, 1.0e-6f);
    if (!a.final_pass) {
        a.vel[s].w = mr;
        return;
    }
    v3 v;
    const bool dyn = (fl & kDynamic) != 0;
    if (dyn) v = vdivs(vsub(xi, ld3(a.x_in, i)), fmax_(a.dt, 1.0e-8f));
    else v = ld3(a.v_in, i);
    a.vel[s] = make_float4(v.x, v.y, v.z, mr);
    if (a.finalize) {
        v3 x = xi;
        if (dyn) a.bounds.clamp_damp(x, v, a.damping);
        st3(a.x_out, i, x);
        st3(a.v_out, i, v);
    }
}

struct PbfNonPressureArgs {
    PassCommon c;
    const float4* __restrict__ vel;
    const float* __restrict__ surface_tension;
    const float* __restrict__ viscosity;
    float particle_radius;
    float dt;
    PbfBounds bounds;
    float* accel;
    float* x_pred;
    float* v_pred;
    float4* pos_next;
};

// acceleration_non_pressure (PBF) + apply_non_pressure_prediction +
// apply_position_bounds  solver_pbf.py:144-189, 227-242, 784-796
template <int BLOCK, int CAP, int UNR, int LPP, bool COH>
__device__ __forceinline__ v3 pbf_nonpressure_sum(const PbfNonPressureArgs& a, int s, v3 xi, v3 vi, float m_i,
                                                  float mu, float gamma, int* lds)
{
    const float h = a.c.h, pr = a.particle_radius;
    const float visc_den0 = 0.01f * (h * h);
    const float mu10 = 10.0f * mu;
    v3 acc = V(0.0f, 0.0f, 0.0f);
    // zero terms (skipped neighbours, mu = 0, no relative motion) leave the
    // sums unchanged
    for_each_neighbor_split<BLOCK, CAP, UNR, LPP>(
        a.c.g, a.c.pos, s, xi.x, xi.y, xi.z, a.c.h, a.c.cut2, lds, a.c.L,
        [&](int k) { return NbrXV{a.c.sflags[k], a.c.pos[k], a.vel[k]}; },
        [&](int k, const NbrXV& nb) -> NonPressureTerms<COH> {
            NonPressureTerms<COH> t;
            if constexpr (COH) t.coh = V(0.0f, 0.0f, 0.0f);
            t.visc = V(0.0f, 0.0f, 0.0f);
            if (k == s || !(nb.flags & kActive)) return t;
            const float4 xj = nb.x;
            const v3 d = vsub(xi, xyz(xj));
            const float dist = vlen_nr(d);
            if constexpr (COH) t.coh = muls(smul(gamma / m_i * xj.w, d), poly6_w(fmax_(dist, pr), h));
            if (mu != 0.0f) {
                const float4 vj = nb.v;
                float v_current_nei = vdot(vsub(vi, xyz(vj)), d);
                if (v_current_nei != 0.0f) {
                    float sv = mu10 * vj.w * v_current_nei / (dist * dist + visc_den0);
                    t.visc = smul(sv, spiky_grad_r(d, dist, h));
                }
            }
            return t;
        },
        [&](const NonPressureTerms<COH>& t) {
            if constexpr (COH) acc = vsub(acc, t.coh);
            acc = vadd(acc, t.visc);
        });
    return acc;
}

template <int BLOCK, int CAP, int UNR, int LPP>
__global__ __launch_bounds__(BLOCK) void k_pbf_nonpressure(PbfNonPressureArgs a)
{
    __shared__ int lds[BLOCK * CAP];
    const int s = pass_slot_g<BLOCK, LPP>();
    const bool lead = (threadIdx.x & (LPP - 1)) == 0;
    if (s < a.c.n) {
        const int i = a.c.perm[s];
        const float4 xi4 = a.c.pos[s];
        if (!(a.c.sflags[s] & kDynamic)) {
            if (lead) {
                st3(a.accel, i, V(0.0f, 0.0f, 0.0f));
                a.pos_next[s] = xi4;
            }
        } else {
            const v3 xi = xyz(xi4);
            const float4 vi4 = a.vel[s];
            const v3 vi = xyz(vi4);
            const float m_i = xi4.w, mu = a.viscosity[i], gamma = a.surface_tension[i];
            v3 acc = V(0.0f, 0.0f, 0.0f);
            if (gamma != 0.0f) acc = pbf_nonpressure_sum<BLOCK, CAP, UNR, LPP, true>(a, s, xi, vi, m_i, mu, gamma, lds);
            else if (mu != 0.0f)
                acc = pbf_nonpressure_sum<BLOCK, CAP, UNR, LPP, false>(a, s, xi, vi, m_i, mu, gamma, lds);
            if (lead) {
                st3(a.accel, i, acc);
                const v3 delta_v = smul(a.dt, acc);
                const v3 v = vadd(vi, delta_v);
                v3 x = vadd(xi, smul(a.dt, delta_v));
                x = a.bounds.clamp(x);
                st3(a.v_pred, i, v);
                st3(a.x_pred, i, x);
                a.pos_next[s] = make_float4(x.x, x.y, x.z, xi4.w);
            }
        }
    }
}

struct PbfLambdaArgs {
    PassCommon c;
    const float* __restrict__ rest_density;
    float lambda_epsilon;
    int min_neighbors;
    float* pbf_lambda;
    float* rho;
    int* neighbor_count;
    float4* sca;
};

// compute_pbf_lambda  solver_pbf.py:245-302
// a neighbour's contributions to the PBF constraint: the neighbour count,
// density, gradient and squared gradient
struct LambdaTerms {
    int count;
    float rho;
    v3 grad;
    float grad_sq;
};

template <int BLOCK, int CAP, int LPP, int U>
__global__ __launch_bounds__(BLOCK) void k_pbf_lambda(PbfLambdaArgs a)
{
    __shared__ int lds[BLOCK / LPP * CAP];
    PassCommon c = a.c;
    const int s = pass_slot_g<BLOCK, LPP>();
    const bool lead = (threadIdx.x & (LPP - 1)) == 0;
    if (s >= c.n) return;
    const int i = c.perm[s];
    if (!(c.sflags[s] & kDynamic)) {
        if (lead) {
            a.pbf_lambda[i] = 0.0f;
            a.rho[i] = 0.0f;
            a.neighbor_count[i] = 0;
            a.sca[s] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        }
        return;
    }
    const float rho0_i = fmax_(a.rest_density[i], 1.0e-6f);
    const v3 xi = xyz(c.pos[s]);
    const float h = c.h;
    float rho_i = 0.0f;
    v3 grad_i = V(0.0f, 0.0f, 0.0f);
    float sum_grad_sq = 0.0f;
    int count = 0;
    scan_pass<BLOCK, CAP, LPP, U>(
        c, s, xi.x, xi.y, xi.z, lds, NoCount{}, [&](int k) { return NbrX{c.sflags[k], c.pos[k]}; },
        [&](int k, const NbrX& nb) -> LambdaTerms {
            LambdaTerms t{0, 0.0f, V(0.0f, 0.0f, 0.0f), 0.0f};
            if (!(nb.flags & kActive)) return t;
            const float4 xj = nb.x;
            v3 xij = vsub(xi, xyz(xj));
            float r = vlen_nr(xij);
            t.count = (k != s && r < h) ? 1 : 0;
            float mass_j = xj.w;
            t.rho = mass_j * poly6_w(r, h);
            t.grad = smul(-(mass_j / rho0_i), spiky_grad_r(xij, r, h));
            t.grad_sq = vdot(t.grad, t.grad);
            return t;
        },
        [&](const LambdaTerms& t) {
            count += t.count;
            rho_i += t.rho;
            sum_grad_sq += t.grad_sq;
            grad_i = vsub(grad_i, t.grad);
        });
    if (!lead) return;
    sum_grad_sq += vdot(grad_i, grad_i);
    const float c_i = fmax_(rho_i / rho0_i - 1.0f, 0.0f);
    a.rho[i] = rho_i;
    a.neighbor_count[i] = count;
    float lam;
    if (a.min_neighbors > 0 && count < a.min_neighbors) lam = 0.0f;
    else lam = -c_i / (sum_grad_sq + a.lambda_epsilon);
    a.pbf_lambda[i] = lam;
    a.sca[s] = make_float4(lam, 0.0f, 0.0f, 0.0f);
}

struct PbfDeltaArgs {
    PassCommon c;
    const float4* __restrict__ sca;
    const float* __restrict__ rest_density;
    float scorr_k, scorr_n, scorr_w_delta_q, max_delta_position;
    int apply_bounds;
    PbfBounds bounds;
    float* delta_p;
    float* x_pred;
    float4* pos_next;
};

// compute_delta_position + apply_delta_position (+ apply_position_bounds after
// the last iteration)  solver_pbf.py:305-363, 860-872
template <int BLOCK, int CAP, int UNR, bool SCORR, int LPP>
__global__ __launch_bounds__(BLOCK) void k_pbf_delta(PbfDeltaArgs a)
{
    __shared__ int lds[BLOCK * CAP];
    PassCommon c = a.c;
    const int s = pass_slot_g<BLOCK, LPP>();
    const bool lead = (threadIdx.x & (LPP - 1)) == 0;
    if (s < c.n) {
        const int i = c.perm[s];
        const float4 xi4 = c.pos[s];
        if (!(c.sflags[s] & kDynamic)) {
            if (lead) {
                st3(a.delta_p, i, V(0.0f, 0.0f, 0.0f));
                a.pos_next[s] = xi4;
                st3(a.x_pred, i, xyz(xi4));
            }
        } else {
            const v3 xi = xyz(xi4);
            const float rho0_i = fmax_(a.rest_density[i], 1.0e-6f);
            const float lam_i = a.sca[s].x;
            const float h = c.h;
            v3 dp = V(0.0f, 0.0f, 0.0f);
            // skipped neighbours contribute +0, which leaves the sums unchanged
            for_each_neighbor_split<BLOCK, CAP, UNR, LPP>(
                c.g, c.pos, s, xi.x, xi.y, xi.z, c.h, c.cut2, lds, c.L,
                [&](int k) { return NbrXS{c.sflags[k], c.pos[k], a.sca[k].x}; },
                [&](int k, const NbrXS& nb) -> v3 {
                    if (k == s || !(nb.flags & kActive)) return V(0.0f, 0.0f, 0.0f);
                    const float4 xj = nb.x;
                    v3 xij = vsub(xi, xyz(xj));
                    const float rn = vlen_nr(xij);
                    v3 grad = spiky_grad_r(xij, rn, h);
                    float scorr = 0.0f;
                    if (SCORR) scorr = compute_scorr(rn, h, a.scorr_k, a.scorr_n, a.scorr_w_delta_q);
                    return smul((lam_i + nb.s + scorr) * xj.w, grad);
                },
                [&](const v3& t) { dp = vadd(dp, t); });
            if (!lead) return;
            dp = vdivs(dp, rho0_i);
            if (a.max_delta_position > 0.0f) {
                float dp_len = vlen(dp);
                if (dp_len > a.max_delta_position) dp = muls(dp, a.max_delta_position / fmax_(dp_len, 1.0e-12f));
            }
            st3(a.delta_p, i, dp);
            v3 x = vadd(xi, dp);
            if (a.apply_bounds) x = a.bounds.clamp(x);
            st3(a.x_pred, i, x);
            a.pos_next[s] = make_float4(x.x, x.y, x.z, xi4.w);
        }
    }
}

struct PbfXsphArgs {
    PassCommon c;
    const float4* __restrict__ vel;
    float xsph_c;
    float damping;
    PbfBounds bounds;
    float* x_out;
    float* v_out;
};

// apply_xsph_viscosity + copies + apply_bounds  solver_pbf.py:434-463, 919-954
template <int BLOCK, int CAP, int UNR, int LPP>
__global__ __launch_bounds__(BLOCK) void k_pbf_xsph(PbfXsphArgs a)
{
    __shared__ int lds[BLOCK * CAP];
    PassCommon c = a.c;
    const int s = pass_slot_g<BLOCK, LPP>();
    const bool lead = (threadIdx.x & (LPP - 1)) == 0;
    if (s >= c.n) return;
    const int i = c.perm[s];
    v3 x = xyz(c.pos[s]);
    const v3 vi = xyz(a.vel[s]);
    if (!(c.sflags[s] & kDynamic)) {
        if (lead) {
            st3(a.x_out, i, x);
            st3(a.v_out, i, vi);
        }
        return;
    }
    const float h = c.h;
    v3 corr = V(0.0f, 0.0f, 0.0f);
    // skipped neighbours contribute +0, which leaves the sums unchanged
    for_each_neighbor_split<BLOCK, CAP, UNR, LPP>(
        c.g, c.pos, s, x.x, x.y, x.z, c.h, c.cut2, lds, c.L,
        [&](int k) { return NbrXV{c.sflags[k], c.pos[k], a.vel[k]}; },
        [&](int k, const NbrXV& nb) -> v3 {
            if (k == s || !(nb.flags & kActive)) return V(0.0f, 0.0f, 0.0f);
            const float4 xj = nb.x;
            const float4 vj = nb.v;
            float w = poly6_w(vlen_nr(vsub(x, xyz(xj))), h);
            return muls(smul(vj.w, vsub(xyz(vj), vi)), w);
        },
        [&](const v3& t) { corr = vadd(corr, t); });
    if (!lead) return;
    v3 v = vadd(vi, smul(a.xsph_c, corr));
    a.bounds.clamp_damp(x, v, a.damping);
    st3(a.x_out, i, x);
    st3(a.v_out, i, v);
}

// ---------------------------------------------------------------------------
// APIC particle separation
// ---------------------------------------------------------------------------

struct SepArgs {
    PassCommon c;        // c.h: query radius (= cell width); c.cut2: rest_dist^2 with slack
    float rest_dist;
    float strength;
    float wall_band;
    float lo[3], hi[3];  // box corners
    float clo[3], chi[3];  // clamp bounds: lo + wall_margin, hi - wall_margin
    float* x;            // positions [n,3], original order, updated in place
    float* delta;        // optional [n,3]: the displacement applied
};

// accumulate_particle_separation + apply_particle_separation_delta
// (solver_apic.py:193-268), fused.  Positions are read only from the sorted
// copy the grid build gathered, so particle i's new position is written in
// place while other lanes still read its old one: the reference's two
// launches (all displacements, then all updates) in one.
template <int BLOCK, int CAP>
__global__ __launch_bounds__(BLOCK) void k_separation(SepArgs a)
{
    __shared__ int lds[BLOCK * CAP];
    PassCommon c = a.c;
    const int s = pass_slot_b<BLOCK>();
    if (s >= c.n) return;
    const int i = c.perm[s];
    v3 disp = V(0.0f, 0.0f, 0.0f);
    if (c.sflags[s] & kActive) {
        const v3 xi = xyz(c.pos[s]);
        const float rest = a.rest_dist, strength = a.strength;
        float count = 0.0f;
        // cells of the rest distance hold about one particle, so a row has
        // about three candidates: batches of four waste less than eight
        for_each_neighbor<BLOCK, CAP, false, 4>(c.g, c.pos, s, xi.x, xi.y, xi.z, c.h, c.cut2, lds, c.L, NoCount{}, phased(
            [&](int k) { return NbrX{c.sflags[k], c.pos[k]}; },
            [&](int k, const NbrX& nb) {
                if (k == s || !(nb.flags & kActive)) return;
                const v3 rij = vsub(xi, xyz(nb.x));
                const float d = vlen_nr(rij);
                if (d > 1.0e-8f && d < rest) {
                    disp = vadd(disp, muls(vdivs(rij, d), strength * (rest - d)));
                    count += 1.0f;
                }
            }));
        if (count > 0.0f) disp = vdivs(disp, count);
        else disp = V(0.0f, 0.0f, 0.0f);
        // into-wall components removed near the walls (:233-245)
        if (xi.x <= a.lo[0] + a.wall_band && disp.x < 0.0f) disp.x = 0.0f;
        if (xi.x >= a.hi[0] - a.wall_band && disp.x > 0.0f) disp.x = 0.0f;
        if (xi.y <= a.lo[1] + a.wall_band && disp.y < 0.0f) disp.y = 0.0f;
        if (xi.y >= a.hi[1] - a.wall_band && disp.y > 0.0f) disp.y = 0.0f;
        if (xi.z <= a.lo[2] + a.wall_band && disp.z < 0.0f) disp.z = 0.0f;
        if (xi.z >= a.hi[2] - a.wall_band && disp.z > 0.0f) disp.z = 0.0f;
        // wp.clamp(v, lo, hi) = min(max(lo, v), hi)  (:262-267)
        const v3 p = vadd(xi, disp);
        st3(a.x, i, V(fmin_(fmax_(a.clo[0], p.x), a.chi[0]), fmin_(fmax_(a.clo[1], p.y), a.chi[1]),
                      fmin_(fmax_(a.clo[2], p.z), a.chi[2])));
    }
    if (a.delta) st3(a.delta, i, disp);
}

// ---------------------------------------------------------------------------
// launch variants of the tunable passes: workgroup size, the depth of the
// per-lane LDS survivor list (which bounds occupancy through LDS), the number
// of neighbours a list replay fetches together and the particles per lane.
// Every variant visits the same candidates in the same order, so results are
// bit-identical; only speed differs.  The selection is process-wide.
// ---------------------------------------------------------------------------

struct Variant {
    int block;
    int cap;       // LDS survivor-list entries per lane
    int unroll;    // neighbours whose data a list replay fetches together (per lane)
    int ppt;       // particles per lane
    int lpp = 1;   // lanes per particle (block / lpp >= 64 particles per workgroup)
    int coop = 0;  // cooperative cell scan (fs_coop.h): chunks per group (cap = chunk slots per tile); 0: per-lane
};
// variant 0 of each table is the default: the fastest at the developed
// 401,856-particle dam break
constexpr Variant kScanVariants[] = {{128, 48, 4, 1, 2}, {64, 24, 4, 1},     {64, 32, 4, 1},     {64, 64, 4, 1},
                                     {128, 24, 4, 1},    {128, 32, 4, 1},    {128, 16, 4, 1},    {256, 32, 4, 1},
                                     {256, 24, 4, 1},    {256, 16, 4, 1},    {64, 40, 4, 1},     {64, 56, 4, 1},
                                     {64, 48, 4, 1},     {128, 32, 4, 1, 2}, {128, 64, 2, 1, 2}, {256, 48, 2, 1, 4},
                                     {256, 32, 2, 1, 4}, {128, 40, 4, 1, 2}, {128, 56, 4, 1, 2},
                                     {256, 40, 4, 1, 2},
                                     // cooperative cell scan (fs_coop.h): 64 lpp threads, cap chunk
                                     // slots per tile, coop chunks per group (the second holds the
                                     // 2,048-candidate queries of dense inputs).  Measured slower than
                                     // the per-lane scan at the dam break and on clustered points
                                     {128, 48, 4, 1, 2, 8},   {128, 64, 4, 1, 2, 32}};
constexpr Variant kReplayVariants[] = {{128, 16, 4, 1, 2}, {128, 16, 4, 1},    {256, 16, 4, 1},    {512, 16, 4, 1},
                                       {64, 16, 8, 1},     {256, 16, 8, 1},    {64, 16, 4, 2},     {256, 16, 4, 2},
                                       {256, 16, 4, 1, 4}, {256, 16, 2, 1, 4}, {64, 16, 4, 1},     {256, 16, 4, 1, 2}};
// the scanning (density, lambda) and list-replaying (non-pressure, delta,
// XSPH) passes of PBF
constexpr Variant kSolverScanVariants[] = {{128, 48, 4, 1, 2}, {64, 40, 4, 1},     {64, 56, 4, 1},
                                           {64, 32, 4, 1},     {128, 32, 4, 1},    {64, 48, 4, 1},
                                           {128, 32, 4, 1, 2}, {256, 48, 2, 1, 4}, {256, 32, 2, 1, 4},
                                           {128, 40, 4, 1, 2}, {128, 56, 4, 1, 2}, {256, 40, 4, 1, 2}};
// DFSPH's scanning pass (density/factor): the same shapes, the deeper
// survivor list first (its 35-neighbour lists at the developed dam break)
constexpr Variant kDfsphScanVariants[] = {{128, 56, 4, 1, 2}, {64, 40, 4, 1},     {64, 56, 4, 1},
                                          {64, 32, 4, 1},     {128, 32, 4, 1},    {64, 48, 4, 1},
                                          {128, 32, 4, 1, 2}, {256, 48, 2, 1, 4}, {256, 32, 2, 1, 4},
                                          {128, 40, 4, 1, 2}, {128, 48, 4, 1, 2}, {256, 40, 4, 1, 2}};
constexpr Variant kSolverReplayVariants[] = {{256, 16, 4, 1, 4}, {64, 16, 4, 1},     {128, 16, 4, 1},
                                             {512, 16, 4, 1},    {256, 16, 8, 1},    {64, 16, 8, 1},
                                             {256, 16, 4, 1},    {256, 16, 2, 1, 4}, {256, 16, 4, 1, 2},
                                             {128, 16, 4, 1, 2}};
// DFSPH's list-replaying passes (non-pressure, rate, correction)
constexpr Variant kDfsphReplayVariants[] = {
    {256, 16, 4, 1, 4}, {64, 16, 4, 1, 1},  {128, 16, 4, 1, 1}, {512, 16, 4, 1, 1},
    {256, 16, 8, 1, 1}, {64, 16, 8, 1, 1},  {256, 16, 4, 1, 1}, {256, 16, 2, 1, 4},
    {512, 16, 4, 1, 4}, {512, 16, 2, 1, 4}, {256, 16, 4, 1, 2}, {128, 16, 4, 1, 2}};

enum TunePass : int {
    kTuneWcsphDensity = FS_TUNE_WCSPH_DENSITY,
    kTuneWcsphForce = FS_TUNE_WCSPH_FORCE,
    kTuneBatchDensity = FS_TUNE_BATCH_DENSITY,
    kTuneBatchForce = FS_TUNE_BATCH_FORCE,
    kTuneDfsphScan = FS_TUNE_DFSPH_SCAN,
    kTuneDfsphReplay = FS_TUNE_DFSPH_REPLAY,
    kTunePbfScan = FS_TUNE_PBF_SCAN,
    kTunePbfReplay = FS_TUNE_PBF_REPLAY,
};
enum VariantTable : int {
    kTabScan = 0,
    kTabReplay = 1,
    kTabSolverScan = 2,
    kTabSolverReplay = 3,
    kTabDfsphReplay = 4,
    kTabDfsphScan = 5
};
constexpr int kTunePasses = 8;   // FS_TUNE_WCSPH_DENSITY .. FS_TUNE_PBF_REPLAY
constexpr int kTuneTable[kTunePasses] = {kTabScan,      kTabReplay,      kTabScan,       kTabReplay,
                                         kTabDfsphScan, kTabDfsphReplay, kTabSolverScan, kTabSolverReplay};

constexpr int table_size(int t)
{
    return t == kTabScan     ? (int)(sizeof(kScanVariants) / sizeof(Variant))
           : t == kTabReplay ? (int)(sizeof(kReplayVariants) / sizeof(Variant))
           : t == kTabSolverScan ? (int)(sizeof(kSolverScanVariants) / sizeof(Variant))
           : t == kTabSolverReplay ? (int)(sizeof(kSolverReplayVariants) / sizeof(Variant))
           : t == kTabDfsphScan ? (int)(sizeof(kDfsphScanVariants) / sizeof(Variant))
                                : (int)(sizeof(kDfsphReplayVariants) / sizeof(Variant));
}
constexpr Variant variant_at(int t, int i)
{
    return t == kTabScan     ? kScanVariants[i]
           : t == kTabReplay ? kReplayVariants[i]
           : t == kTabSolverScan ? kSolverScanVariants[i]
           : t == kTabSolverReplay ? kSolverReplayVariants[i]
           : t == kTabDfsphScan ? kDfsphScanVariants[i]
                                : kDfsphReplayVariants[i];
}

// Every launch shape covers at least kMinPassBlock particles per workgroup:
// fs_work_bytes sizes the rate pass's per-workgroup residual partials (and
// the tile mappings assume whole 64-slot tiles) on that bound.
constexpr bool table_fits(int t)
{
    for (int i = 0; i < table_size(t); ++i) {
        const Variant v = variant_at(t, i);
        if (v.lpp < 1 || v.block % v.lpp != 0 || v.block / v.lpp < kMinPassBlock || (v.block / v.lpp) % 64 != 0)
            return false;
    }
    return true;
}
static_assert(table_fits(kTabScan) && table_fits(kTabReplay) && table_fits(kTabSolverScan) &&
                  table_fits(kTabSolverReplay) && table_fits(kTabDfsphScan) && table_fits(kTabDfsphReplay),
              "a launch variant covers fewer than kMinPassBlock particles per workgroup");

static std::atomic<int> g_tune[kTunePasses];
static std::mutex g_tune_mutex;

static int tune_count(int pass) { return table_size(kTuneTable[pass]); }
static int tuned(int pass) { return g_tune[pass].load(std::memory_order_relaxed); }
static Variant tuned_variant(int pass)
{
    const int t = kTuneTable[pass], v = tuned(pass);
    switch (t) {
    case kTabScan: return kScanVariants[v];
    case kTabReplay: return kReplayVariants[v];
    case kTabSolverScan: return kSolverScanVariants[v];
    case kTabSolverReplay: return kSolverReplayVariants[v];
    case kTabDfsphScan: return kDfsphScanVariants[v];
    default: return kDfsphReplayVariants[v];
    }
}

// f.template operator()<V>() with V = variant v of table TABLE
template <int TABLE, int I = 0, class F>
static int32_t with_variant(int v, F&& f)
{
    if constexpr (I < table_size(TABLE)) {
        if (v == I) return f.template operator()<variant_at(TABLE, I)>();
        return with_variant<TABLE, I + 1>(v, f);
    } else {
        set_error("unknown launch variant %d", v);
        return FS_ERR_ARG;
    }
}

template <bool BATCH>
static int32_t launch_density(int v, int n, hipStream_t s, const WcsphDensityArgs& a)
{
    return with_variant<kTabScan>(v, [&]<Variant V>() -> int32_t {
        if constexpr (V.coop > 0) {
            FS_LAUNCH((k_wcsph_density_coop<V.cap, V.coop, BATCH, V.lpp, V.unroll>),
                      dim3(blocks_for(n, kCoopTile)), dim3(kCoopTile * V.lpp), 0, s, a);
        } else {
            FS_LAUNCH((k_wcsph_density<V.block, V.cap, BATCH, V.lpp, V.unroll>), dim3(blocks_for(n, V.block / V.lpp)),
                      dim3(V.block), 0, s, a);
        }
        return FS_OK;
    });
}

template <bool BATCH>
static int32_t launch_force(int v, int n, hipStream_t s, const WcsphForceArgs& a)
{
    return with_variant<kTabReplay>(v, [&]<Variant V>() -> int32_t {
        FS_LAUNCH((k_wcsph_force<V.block, V.cap, BATCH, V.unroll, V.ppt, V.lpp>),
                  dim3(blocks_for(n, V.block * V.ppt / V.lpp)), dim3(V.block), 0, s, a);
        return FS_OK;
    });
}

// ---------------------------------------------------------------------------
// host orchestration
// ---------------------------------------------------------------------------

static NbrList list_of(const fs_work* w, int n, int mode, float cut2)
{
    NbrList L;
    L.idx = w->nbr;
    L.cnt = w->nbr_count;
    L.cap = w->list_capacity;
    L.n = n;
    L.mode = (w->list_capacity > 0 && w->nbr && w->nbr_count) ? mode : kListNone;
    L.cut2 = cut2;
    L.need = w->control ? w->control + kListNeed : nullptr;
    return L;
}

static PassCommon make_common(const fs_grid* g, const fs_work* w, const float4* pos, float h, int list_mode)
{
    PassCommon c;
    c.g = grid_view(g);
    c.n = g->num_points;
    c.perm = g->point_ids;
    c.pos = pos;
    c.sflags = w->flags;
    c.h = h;
    c.cut2 = h * h * kCutSlack;
    c.L = list_of(w, g->num_points, list_mode, c.cut2);
    return c;
}

static int32_t check_common(const fs_grid* g, const fs_work* w, const fs_model* m)
{
    if (!g || !w || !m) {
        set_error("null grid, workspace or model");
        return FS_ERR_ARG;
    }
    if (m->n < 0 || m->n > g->capacity || m->n > w->capacity) {
        set_error("particle count %d exceeds grid capacity %d / workspace capacity %d", m->n, g->capacity,
                  w->capacity);
        return FS_ERR_CAPACITY;
    }
    if (m->n > 0 && (!m->flags || !m->mask || !m->world || !m->mass || !m->gravity || !m->rest_density ||
                     !m->surface_tension || !m->viscosity || m->num_worlds < 1)) {
        set_error("model arrays missing");
        return FS_ERR_ARG;
    }
    return FS_OK;
}

#define LAUNCH_PASS(kernel, nb, s, args) FS_LAUNCH((kernel), dim3(nb), dim3(kPassBlock), 0, s, args)

static GatherOp gather_for(const fs_model* m, fs_work* w, const float* x, const float* v, float4* vel_dst)
{
    GatherOp ga;
    ga.x = x;
    ga.v = v;
    ga.mass = m->mass;
    ga.rest_density = m->rest_density;
    ga.flags = m->flags;
    ga.mask = m->mask;
    ga.pos = reinterpret_cast<float4*>(w->pos);
    ga.vel = vel_dst;
    ga.wvol = w->wvol;
    ga.sflags = w->flags;
    ga.packed = reinterpret_cast<int4*>(w->packed);
    return ga;
}

}  // namespace fs

using namespace fs;

extern "C" {

size_t fs_work_bytes(int32_t capacity, int32_t list_capacity)
{
    if (capacity < 0 || list_capacity < 0) return 0;
    size_t cap = (size_t)(capacity > 0 ? capacity : 1);
    auto al = [](size_t v) { return (v + 255) & ~(size_t)255; };
    size_t nblocks = (cap + kMinPassBlock - 1) / kMinPassBlock;   // residual partials of the rate pass
    return al(cap * 16) * 4 + al(cap * 4) * 2 + al(cap * 12) * 2 + al(cap * 4 * (size_t)list_capacity) +
           al(cap * 4) + al(nblocks * 16 + 16) + al(kControlWords * 4) + al(8 * 4) + al(cap * 16);
}

int32_t fs_work_init(fs_work* w, void* workspace, size_t bytes, int32_t capacity, int32_t list_capacity,
                     fs_stream_t stream)
{
    if (!w || !workspace || capacity < 0 || list_capacity < 0) {
        set_error("fs_work_init: invalid arguments");
        return FS_ERR_ARG;
    }
    if (capacity > kMaxParticles) {
        set_error("fs_work_init: capacity %d exceeds the maximum of %d particles", capacity, kMaxParticles);
        return FS_ERR_ARG;
    }
    size_t need = fs_work_bytes(capacity, list_capacity);
    if (bytes < need) {
        set_error("fs_work_init: workspace %zu bytes < required %zu", bytes, need);
        return FS_ERR_WORKSPACE;
    }
    size_t cap = (size_t)(capacity > 0 ? capacity : 1);
    size_t nblocks = (cap + kMinPassBlock - 1) / kMinPassBlock;
    auto al = [](size_t v) { return (v + 255) & ~(size_t)255; };
    char* p = (char*)workspace;
    auto take = [&](size_t n) {
        char* r = p;
        p += al(n);
        return r;
    };
    w->capacity = capacity;
    w->list_capacity = list_capacity;
    w->pos = (float*)take(cap * 16);
    w->pos_alt = (float*)take(cap * 16);
    w->vel = (float*)take(cap * 16);
    w->sca = (float*)take(cap * 16);
    w->wvol = (float*)take(cap * 4);
    w->flags = (int32_t*)take(cap * 4);
    w->x_pred = (float*)take(cap * 12);
    w->v_pred = (float*)take(cap * 12);
    w->nbr = list_capacity > 0 ? (int32_t*)take(cap * 4 * (size_t)list_capacity) : nullptr;
    w->nbr_count = (int32_t*)take(cap * 4);
    w->partials = (double*)take(nblocks * 16 + 16);
    w->control = (int32_t*)take(kControlWords * 4);
    w->sync = (uint32_t*)take(8 * 4);
    w->packed = (int32_t*)take(cap * 16);
    hipStream_t s = (hipStream_t)stream;
    int32_t rc = hip_check(hipMemsetAsync(w->control, 0, kControlWords * 4, s), "fs_work_init memset");
    if (rc) return rc;
    return hip_check(hipMemsetAsync(w->sync, 0, 8 * 4, s), "fs_work_init memset");
}

}  // extern "C"

namespace fs {

// grid build + density/EOS pass (records neighbour lists) + force/integrate
// pass (replays them): SolverWCSPH.step  solver_wcsph.py:431-535
// build (solver_wcsph.py:456) + compute_density_pressure (:458-474); the
// pass records the neighbour lists
static int32_t wcsph_density_eos(fs_grid* g, fs_work* w, const fs_model* m, const fs_wcsph_params* prm,
                                 const fs_wcsph_io* io, hipStream_t s)
{
    const int n = m->n;
    const float h = prm->smoothing_length;
    PreOp pre;
    pre.x_in = io->x_in;
    int32_t rc = build_grid(g, n, h, pre, gather_for(m, w, io->x_in, io->v_in, reinterpret_cast<float4*>(w->vel)), s);
    if (rc) return rc;
    const float4* pos = reinterpret_cast<const float4*>(w->pos);
    WcsphDensityArgs da;
    da.c = make_common(g, w, pos, h, kListBuild);
    da.wvol = w->wvol;
    da.rest_density = m->rest_density;
    da.stiffness = m->stiffness;
    da.exponent = m->exponent;
    da.clamp = prm->clamp_negative_pressure;
    da.dt_device = prm->dt_device;
    da.rho_in = io->rho_in;
    da.pressure_in = io->pressure_in;
    da.rho = io->rho;
    da.pressure = io->pressure;
    da.sca = reinterpret_cast<float4*>(w->sca);
    da.vel = reinterpret_cast<float4*>(w->vel);
    da.scenes = nullptr;
    da.sscene = nullptr;
    return launch_density<false>(tuned(0), n, s, da);
}

// acceleration_wcsph + integrate_velocity + integrate_positions + apply_bounds
// (solver_wcsph.py:475-535), replaying the density pass's lists
static int32_t wcsph_force_integrate(fs_grid* g, fs_work* w, const fs_model* m, const fs_wcsph_params* prm,
                                     const fs_wcsph_io* io, hipStream_t s)
{
    const int n = m->n;
    const float h = prm->smoothing_length;
    const float4* pos = reinterpret_cast<const float4*>(w->pos);
    WcsphForceArgs fa;
    fa.c = make_common(g, w, pos, h, kListUse);
    fa.vel = reinterpret_cast<const float4*>(w->vel);
    fa.sca = reinterpret_cast<const float4*>(w->sca);
    fa.surface_tension = m->surface_tension;
    fa.viscosity = m->viscosity;
    fa.gravity = m->gravity;
    fa.world = m->world;
    fa.x_in = io->x_in;
    fa.v_in = io->v_in;
    fa.particle_radius = prm->particle_radius;
    fa.dt = prm->dt;
    for (int d = 0; d < 3; ++d) fa.bound[d] = prm->bound[d];
    fa.damping = prm->boundary_damping;
    fa.dt_device = prm->dt_device;
    fa.accel = io->accel;
    fa.x_out = io->x_out;
    fa.v_out = io->v_out;
    fa.scenes = nullptr;
    fa.sscene = nullptr;
    return launch_force<false>(tuned(1), n, s, fa);
}

static int32_t wcsph_forward(fs_grid* g, fs_work* w, const fs_model* m, const fs_wcsph_params* prm,
                             const fs_wcsph_io* io, hipStream_t s)
{
    int32_t rc = wcsph_density_eos(g, w, m, prm, io, s);
    if (rc) return rc;
    return wcsph_force_integrate(g, w, m, prm, io, s);
}

static int32_t check_wcsph(const fs_grid* g, const fs_work* w, const fs_model* m, const fs_wcsph_params* prm,
                           const fs_wcsph_io* io, const char* who)
{
    int32_t rc = check_common(g, w, m);
    if (rc) return rc;
    if (!prm || !io) {
        set_error("%s: null params or io", who);
        return FS_ERR_ARG;
    }
    if (m->n > 0 && (!io->x_in || !io->v_in || !io->x_out || !io->v_out || !io->rho || !io->pressure ||
                     !io->accel || !m->stiffness || !m->exponent)) {
        set_error("%s: missing arrays", who);
        return FS_ERR_ARG;
    }
    return FS_OK;
}

static size_t adjoint_scratch_bytes(size_t cap)
{
    auto al = [](size_t v) { return (v + 255) & ~(size_t)255; };
    return al(cap * 16) * 3 + al(cap * 4);
}

}  // namespace fs

extern "C" {

int32_t fs_tune_variant_count(int32_t pass)
{
    if (pass < 0 || pass >= kTunePasses) {
        set_error("fs_tune_variant_count: unknown pass %d", pass);
        return FS_ERR_ARG;
    }
    return tune_count(pass);
}

int32_t fs_tune_describe(int32_t pass, int32_t variant, int32_t* block, int32_t* list_depth)
{
    if (pass < 0 || pass >= kTunePasses || variant < 0 || variant >= tune_count(pass)) {
        set_error("fs_tune_describe: unknown pass %d / variant %d", pass, variant);
        return FS_ERR_ARG;
    }
    const Variant v = variant_at(kTuneTable[pass], variant);
    if (block) *block = v.block;
    if (list_depth) *list_depth = v.cap;
    return FS_OK;
}

int32_t fs_tune_describe_ex(int32_t pass, int32_t variant, int32_t* block, int32_t* list_depth, int32_t* unroll,
                            int32_t* particles_per_thread, int32_t* lanes_per_particle)
{
    int32_t rc = fs_tune_describe(pass, variant, block, list_depth);
    if (rc) return rc;
    const Variant v = variant_at(kTuneTable[pass], variant);
    if (unroll) *unroll = v.unroll;
    if (particles_per_thread) *particles_per_thread = v.ppt;
    if (lanes_per_particle) *lanes_per_particle = v.lpp;
    return FS_OK;
}

int32_t fs_tune_variant_coop(int32_t pass, int32_t variant)
{
    if (pass < 0 || pass >= kTunePasses || variant < 0 || variant >= tune_count(pass)) {
        set_error("fs_tune_variant_coop: unknown pass %d / variant %d", pass, variant);
        return FS_ERR_ARG;
    }
    return variant_at(kTuneTable[pass], variant).coop;
}

int32_t fs_tune_select(int32_t pass, int32_t variant)
{
    std::lock_guard<std::mutex> lock(g_tune_mutex);
    if (pass < 0 || pass >= kTunePasses || variant < 0 || variant >= tune_count(pass)) {
        set_error("fs_tune_select: unknown pass %d / variant %d", pass, variant);
        return FS_ERR_ARG;
    }
    g_tune[pass].store(variant, std::memory_order_relaxed);
    return FS_OK;
}

int32_t fs_tune_selected(int32_t pass)
{
    std::lock_guard<std::mutex> lock(g_tune_mutex);
    if (pass < 0 || pass >= kTunePasses) {
        set_error("fs_tune_selected: unknown pass %d", pass);
        return FS_ERR_ARG;
    }
    return tuned(pass);
}

int32_t fs_wcsph_step(fs_grid* g, fs_work* w, const fs_model* m, const fs_wcsph_params* prm, const fs_wcsph_io* io,
                      fs_stream_t stream)
{
    int32_t rc = check_wcsph(g, w, m, prm, io, "fs_wcsph_step");
    if (rc) return rc;
    if (m->n == 0) return FS_OK;
    return wcsph_forward(g, w, m, prm, io, (hipStream_t)stream);
}

int32_t fs_wcsph_density_eos(fs_grid* g, fs_work* w, const fs_model* m, const fs_wcsph_params* prm,
                             const fs_wcsph_io* io, fs_stream_t stream)
{
    int32_t rc = check_wcsph(g, w, m, prm, io, "fs_wcsph_density_eos");
    if (rc) return rc;
    if (m->n == 0) return FS_OK;
    return wcsph_density_eos(g, w, m, prm, io, (hipStream_t)stream);
}

int32_t fs_wcsph_force_integrate(fs_grid* g, fs_work* w, const fs_model* m, const fs_wcsph_params* prm,
                                 const fs_wcsph_io* io, fs_stream_t stream)
{
    int32_t rc = check_wcsph(g, w, m, prm, io, "fs_wcsph_force_integrate");
    if (rc) return rc;
    if (m->n == 0) return FS_OK;
    if (g->num_points != m->n) {
        set_error("fs_wcsph_force_integrate: the grid does not hold this model's particles (run fs_wcsph_density_eos "
                  "first)");
        return FS_ERR_ARG;
    }
    return wcsph_force_integrate(g, w, m, prm, io, (hipStream_t)stream);
}

size_t fs_wcsph_adjoint_bytes(int32_t capacity)
{
    if (capacity < 0) return 0;
    return adjoint_scratch_bytes((size_t)(capacity > 0 ? capacity : 1));
}

int32_t fs_wcsph_adjoint_step(fs_grid* g, fs_work* w, const fs_model* m, const fs_wcsph_params* prm,
                              const fs_wcsph_adjoint_io* io, size_t scratch_bytes, fs_stream_t stream)
{
    if (!io) {
        set_error("fs_wcsph_adjoint_step: null io");
        return FS_ERR_ARG;
    }
    int32_t rc = check_wcsph(g, w, m, prm, &io->forward, "fs_wcsph_adjoint_step");
    if (rc) return rc;
    const int n = m->n;
    if (n == 0) return FS_OK;
    if (!io->xbar_out || !io->vbar_out || !io->xbar_in || !io->vbar_in || !io->grad_stiffness ||
        !io->grad_viscosity || !io->scratch) {
        set_error("fs_wcsph_adjoint_step: missing adjoint arrays");
        return FS_ERR_ARG;
    }
    if (scratch_bytes < adjoint_scratch_bytes((size_t)n)) {
        set_error("fs_wcsph_adjoint_step: scratch %zu bytes < required %zu", scratch_bytes,
                  adjoint_scratch_bytes((size_t)n));
        return FS_ERR_WORKSPACE;
    }
    if (w->list_capacity <= 0) {
        set_error("fs_wcsph_adjoint_step: the workspace needs neighbour lists (list_capacity > 0)");
        return FS_ERR_ARG;
    }
    hipStream_t s = (hipStream_t)stream;
    rc = wcsph_forward(g, w, m, prm, &io->forward, s);
    if (rc) return rc;
    auto al = [](size_t v) { return (v + 255) & ~(size_t)255; };
    char* p = (char*)io->scratch;
    float4* abar = (float4*)p;
    float4* xbar = (float4*)(p + al((size_t)n * 16));
    float4* vbar = (float4*)(p + 2 * al((size_t)n * 16));
    float* rhobar = (float*)(p + 3 * al((size_t)n * 16));
    const int nb = blocks_for(n, kPassBlock);
    const float h = prm->smoothing_length;
    const float4* pos = reinterpret_cast<const float4*>(w->pos);

    WadjIntegrateArgs ia;
    ia.n = n;
    ia.perm = g->point_ids;
    ia.sflags = w->flags;
    ia.x_in = io->forward.x_in;
    ia.v_in = io->forward.v_in;
    ia.accel = io->forward.accel;
    ia.xbar_out = io->xbar_out;
    ia.vbar_out = io->vbar_out;
    ia.viscosity = m->viscosity;
    ia.dt = prm->dt;
    for (int d = 0; d < 3; ++d) ia.bound[d] = prm->bound[d];
    ia.damping = prm->boundary_damping;
    ia.abar = abar;
    ia.xbar = xbar;
    ia.vbar = vbar;
    LAUNCH_PASS(k_wadj_integrate, nb, s, ia);

    WadjForceArgs fa;
    fa.c = make_common(g, w, pos, h, kListUse);
    fa.vel = reinterpret_cast<const float4*>(w->vel);
    fa.sca = reinterpret_cast<const float4*>(w->sca);
    fa.abar = abar;
    fa.rest_density = m->rest_density;
    fa.stiffness = m->stiffness;
    fa.exponent = m->exponent;
    fa.clamp = prm->clamp_negative_pressure;
    fa.xbar = xbar;
    fa.vbar = vbar;
    fa.rhobar = rhobar;
    fa.grad_stiffness = io->grad_stiffness;
    fa.grad_viscosity = io->grad_viscosity;
    const int nb_group = blocks_for(n, kPassBlock / kAdjLpp);
    LAUNCH_PASS((k_wadj_force<kReplayCap, kAdjLpp>), nb_group, s, fa);

    WadjDensityArgs da;
    da.c = make_common(g, w, pos, h, kListUse);
    da.wvol = w->wvol;
    da.rhobar = rhobar;
    da.xbar_s = xbar;
    da.vbar_s = vbar;
    da.xbar_in = io->xbar_in;
    da.vbar_in = io->vbar_in;
    LAUNCH_PASS((k_wadj_density<kReplayCap, kAdjLpp>), nb_group, s, da);
    return FS_OK;
}

int32_t fs_batch_wcsph_step(fs_grid* g, fs_work* w, const fs_model* m, const fs_scene* scenes, int32_t num_scenes,
                            const fs_batch_io* io, fs_stream_t stream)
{
    int32_t rc = check_common(g, w, m);
    if (rc) return rc;
    if (!scenes || num_scenes <= 0 || !io) {
        set_error("fs_batch_wcsph_step: null scene table or io");
        return FS_ERR_ARG;
    }
    const int n = m->n;
    if (n == 0) return FS_OK;
    if (!io->x_in || !io->v_in || !io->x_out || !io->v_out || !io->rho || !io->pressure || !io->accel ||
        !io->scene_of || !io->scene_sorted || !m->stiffness || !m->exponent) {
        set_error("fs_batch_wcsph_step: missing arrays");
        return FS_ERR_ARG;
    }
    hipStream_t s = (hipStream_t)stream;
    PreOp pre;
    pre.x_in = io->x_in;
    pre.scenes = scenes;
    pre.scene_of = io->scene_of;
    GatherOp ga = gather_for(m, w, io->x_in, io->v_in, reinterpret_cast<float4*>(w->vel));
    ga.scene_of = io->scene_of;
    ga.sscene = io->scene_sorted;
    // the radius only sets the (unused) uniform cell width of the segmented table
    rc = build_grid(g, n, 1.0f, pre, ga, s);
    if (rc) return rc;
    const int nb = blocks_for(n, kPassBlock);
    const float4* pos = reinterpret_cast<const float4*>(w->pos);

    WcsphDensityArgs da;
    da.c = make_common(g, w, pos, 1.0f, kListBuild);
    da.scenes = scenes;
    da.sscene = io->scene_sorted;
    da.wvol = w->wvol;
    da.rest_density = m->rest_density;
    da.stiffness = m->stiffness;
    da.exponent = m->exponent;
    da.clamp = 1;
    da.dt_device = nullptr;
    da.rho_in = nullptr;
    da.pressure_in = nullptr;
    da.rho = io->rho;
    da.pressure = io->pressure;
    da.sca = reinterpret_cast<float4*>(w->sca);
    da.vel = reinterpret_cast<float4*>(w->vel);
    rc = launch_density<true>(tuned(2), n, s, da);
    if (rc) return rc;

    WcsphForceArgs fa;
    fa.c = make_common(g, w, pos, 1.0f, kListUse);
    fa.scenes = scenes;
    fa.sscene = io->scene_sorted;
    fa.vel = reinterpret_cast<const float4*>(w->vel);
    fa.sca = reinterpret_cast<const float4*>(w->sca);
    fa.surface_tension = m->surface_tension;
    fa.viscosity = m->viscosity;
    fa.gravity = m->gravity;
    fa.world = m->world;
    fa.x_in = io->x_in;
    fa.v_in = io->v_in;
    fa.particle_radius = 0.0f;
    fa.dt = 0.0f;
    for (int d = 0; d < 3; ++d) fa.bound[d] = 0.0f;
    fa.damping = 0.0f;
    fa.dt_device = nullptr;
    fa.accel = io->accel;
    fa.x_out = io->x_out;
    fa.v_out = io->v_out;
    return launch_force<true>(tuned(3), n, s, fa);
}

// One DFSPH substep as a sequence of passes sharing the grid and workspace.
struct DfPlan {
    fs_grid* g;
    fs_work* w;
    const fs_model* m;
    const fs_dfsph_params* prm;
    const fs_dfsph_io* io;
    hipStream_t s;
    int n, nb;
    float h, dt;
    float4* vstar_s;
    float4* sca;
    float4* vel;
    const float4* pos;
};

static int32_t df_plan(fs_grid* g, fs_work* w, const fs_model* m, const fs_dfsph_params* prm, const fs_dfsph_io* io,
                       fs_stream_t stream, const char* what, DfPlan& p)
{
    int32_t rc = check_common(g, w, m);
    if (rc) return rc;
    if (!prm || !io) {
        set_error("%s: null params or io", what);
        return FS_ERR_ARG;
    }
    if (m->n > 0 && (!io->x_in || !io->v_in || !io->x_out || !io->v_out || !io->rho || !io->factor ||
                     !io->rho_star || !io->d_rho_dt || !io->kappa || !io->kappa_v || !io->neighbor_count ||
                     !io->v_star || !io->accel || (prm->warm_start && (!io->kappa_in || !io->kappa_v_in)))) {
        set_error("%s: missing arrays", what);
        return FS_ERR_ARG;
    }
    p.g = g;
    p.w = w;
    p.m = m;
    p.prm = prm;
    p.io = io;
    p.s = (hipStream_t)stream;
    p.n = m->n;
    p.nb = blocks_for(m->n, kPassBlock);
    p.h = prm->smoothing_length;
    p.dt = prm->dt;
    p.vstar_s = reinterpret_cast<float4*>(w->pos_alt);
    p.sca = reinterpret_cast<float4*>(w->sca);
    p.vel = reinterpret_cast<float4*>(w->vel);
    p.pos = reinterpret_cast<const float4*>(w->pos);
    return FS_OK;
}

// stage 0: build on x_in (solver_dfsph.py:624) + compute_density +
// compute_factor_dfsph (:626-653); stage 1: the same passes on the grid the
// integrate step rebuilt at x_out (:785-812)
static int32_t df_density_factor(const DfPlan& p, bool after_integrate)
{
    const fs_model* m = p.m;
    const fs_dfsph_io* io = p.io;
    fs_work* w = p.w;
    if (!after_integrate) {
        PreOp pre;
        pre.x_in = io->x_in;
        pre.control_reset = w->control;
        GatherOp ga = gather_for(m, w, io->x_in, io->v_in, p.vel);
        ga.any_inactive = w->control + kAnyInactive;
        int32_t rc = build_grid(p.g, p.n, p.h, pre, ga, p.s);
        if (rc) return rc;
    }
    DfDensityArgs dfa;
    dfa.c = make_common(p.g, w, p.pos, p.h, kListBuild);
    dfa.any_inactive = w->control + kAnyInactive;
    dfa.wvol = w->wvol;
    dfa.min_neighbors = p.prm->min_neighbors_for_factor;
    dfa.kappa_warm = p.prm->warm_start ? (after_integrate ? io->kappa_v_in : io->kappa_in) : nullptr;
    dfa.kappa_commit = after_integrate ? io->kappa_v : io->kappa;
    dfa.rho = io->rho;
    dfa.factor = io->factor;
    dfa.neighbor_count = io->neighbor_count;
    dfa.sca = p.sca;
    dfa.vel = p.vel;
    dfa.dt_device = p.prm->dt_device;
    return with_variant<kTabDfsphScan>(tuned(kTuneDfsphScan), [&]<Variant V>() -> int32_t {
        FS_LAUNCH((k_dfsph_density_factor<V.block, V.cap, V.lpp, V.unroll>), dim3(blocks_for(p.n, V.block / V.lpp)),
                  dim3(V.block), 0, p.s, dfa);
        return FS_OK;
    });
}

// acceleration_non_pressure + predict_velocity  (solver_dfsph.py:654-688)
static int32_t df_nonpressure_predict(const DfPlan& p)
{
    DfNonPressureArgs npa;
    npa.c = make_common(p.g, p.w, p.pos, p.h, kListUse);
    npa.vel = p.vel;
    npa.surface_tension = p.m->surface_tension;
    npa.viscosity = p.m->viscosity;
    npa.gravity = p.m->gravity;
    npa.world = p.m->world;
    npa.particle_radius = p.prm->particle_radius;
    npa.dt = p.dt;
    npa.dt_device = p.prm->dt_device;
    npa.accel = p.io->accel;
    npa.v_star = p.io->v_star;
    npa.vstar_s = p.vstar_s;
    return with_variant<kTabDfsphReplay>(tuned(kTuneDfsphReplay), [&]<Variant V>() -> int32_t {
        FS_LAUNCH((k_dfsph_nonpressure<V.block, V.cap, V.unroll, V.lpp>), dim3(blocks_for(p.n, V.block / V.lpp)),
                  dim3(V.block), 0, p.s, npa);
        return FS_OK;
    });
}

// apply_constant_density_correction (loop 0) or apply_divergence_free_correction
// (loop 1)  (solver_dfsph.py:309-337, 375-403).  The warm-start form
// (:689-707, :814-832) applies the warm-started kappa and neither stops nor
// counts; the loop form returns at once once the loop has stopped.
static int32_t df_correction(const DfPlan& p, int loop, bool warm)
{
    DfCorrArgs ca;
    ca.c = make_common(p.g, p.w, p.pos, p.h, kListUse);
    ca.sca = p.sca;
    ca.dt = p.dt;
    ca.dt_device = p.prm->dt_device;
    ca.v_star = p.io->v_star;
    ca.vstar_s = p.vstar_s;
    ca.control = p.w->control;
    if (warm) {
        ca.stop_idx = -1;
        ca.corr_idx = -1;
        ca.kappa_commit = nullptr;
    } else {
        ca.stop_idx = loop == 0 ? kStopRho : kStopVel;
        ca.corr_idx = loop == 0 ? kCorrRho : kCorrVel;
        ca.kappa_commit = loop == 0 ? p.io->kappa : p.io->kappa_v;
    }
    return with_variant<kTabDfsphReplay>(tuned(kTuneDfsphReplay), [&]<Variant V>() -> int32_t {
        FS_LAUNCH((k_dfsph_correction<V.block, V.cap, V.unroll, V.lpp>), dim3(blocks_for(p.n, V.block / V.lpp)),
                  dim3(V.block), 0, p.s, ca);
        return FS_OK;
    });
}

// compute_density_change_rate + predict_density_star + compute_kappa_density
// (loop 0) or compute_kappa_divergence (loop 1), with the residual's
// per-block partial sums  (solver_dfsph.py:712-750, 835-866)
static int32_t df_rate_kappa(const DfPlan& p, int loop)
{
    DfRateArgs ra;
    ra.c = make_common(p.g, p.w, p.pos, p.h, kListUse);
    ra.vstar_s = p.vstar_s;
    ra.rest_density = p.m->rest_density;
    ra.dt = p.dt;
    ra.dt_device = p.prm->dt_device;
    ra.d_rho_dt = p.io->d_rho_dt;
    ra.rho_star = p.io->rho_star;
    ra.sca = p.sca;
    ra.partials = p.w->partials;
    ra.control = p.w->control;
    ra.density_loop = loop == 0 ? 1 : 0;
    ra.max_kappa = loop == 0 ? p.prm->max_kappa : p.prm->max_kappa_v;
    return with_variant<kTabDfsphReplay>(tuned(kTuneDfsphReplay), [&]<Variant V>() -> int32_t {
        FS_LAUNCH((k_dfsph_rate<V.block, V.cap, V.unroll, V.lpp>), dim3(blocks_for(p.n, V.block / V.lpp)),
                  dim3(V.block), 0, p.s, ra);
        return FS_OK;
    });
}

// _density_residual / _divergence_residual and the loop's break rule
// (solver_dfsph.py:574-596, 733-735, 849-851), decided on the device
static int32_t df_residual(const DfPlan& p, int loop, int iteration)
{
    ResidualArgs ra;
    ra.partials = p.w->partials;
    const Variant rv = tuned_variant(kTuneDfsphReplay);
    ra.nb = blocks_for(p.n, rv.block / rv.lpp);   // blocks of the rate pass
    ra.control = p.w->control;
    ra.stop_idx = loop == 0 ? kStopRho : kStopVel;
    ra.it_idx = loop == 0 ? kItRho : kItVel;
    ra.iter = iteration;
    ra.eta = loop == 0 ? p.prm->eta_density : p.prm->eta_divergence;
    ra.n = p.n;
    ra.flags = p.m->flags;
    ra.mask = p.m->mask;
    ra.vals = loop == 0 ? p.io->rho_star : p.io->d_rho_dt;
    ra.sub = loop == 0 ? p.m->rest_density : nullptr;
    ra.scratch = p.w->x_pred;   // unused by DFSPH otherwise; holds >= n floats
    ra.force_exact = p.prm->exact_residual;
    ra.dt_device = p.prm->dt_device;
    FS_LAUNCH(k_dfsph_residual, dim3(1), dim3(kPassBlock), 0, p.s, ra);
    return FS_OK;
}

// integrate_positions fused into the second build's hash (solver_dfsph.py:769-783)
static int32_t df_integrate(const DfPlan& p)
{
    const fs_model* m = p.m;
    const fs_dfsph_io* io = p.io;
    PreOp pre2;
    pre2.mode = kPreIntegrate;
    pre2.x_in = io->x_in;
    pre2.v_in = io->v_star;
    pre2.x_out = io->x_out;
    pre2.flags = m->flags;
    pre2.mask = m->mask;
    pre2.dt = p.dt;
    pre2.dt_device = p.prm->dt_device;
    GatherOp ga2 = gather_for(m, p.w, io->x_out, io->v_star, p.vstar_s);
    ga2.any_inactive = p.w->control + kAnyInactive;   // same particles: sets the same flag again
    return build_grid(p.g, p.n, p.h, pre2, ga2, p.s);
}

// apply_bounds (DFSPH variant) and the outputs  (solver_dfsph.py:886-898)
static int32_t df_finalize(const DfPlan& p)
{
    DfFinalArgs fa;
    fa.n = p.n;
    fa.flags = p.m->flags;
    fa.mask = p.m->mask;
    fa.v_star = p.io->v_star;
    fa.x_out = p.io->x_out;
    fa.v_out = p.io->v_out;
    for (int d = 0; d < 3; ++d) fa.bound[d] = p.prm->bound[d];
    fa.damping = p.prm->boundary_damping;
    fa.control = p.w->control;
    fa.iterations = p.io->iterations;
    fa.dt_device = p.prm->dt_device;
    const fs_dfsph_io* io = p.io;
    DfCarry c;
    c.x_in = io->x_in;
    c.v_in = io->v_in;
    c.rho_in = io->rho_in;
    c.factor_in = io->factor_in;
    c.rho_star_in = io->rho_star_in;
    c.d_rho_dt_in = io->d_rho_dt_in;
    c.kappa_in = io->kappa_in;
    c.kappa_v_in = io->kappa_v_in;
    c.neighbor_count_in = io->neighbor_count_in;
    c.v_star_in = io->v_star_in;
    c.rho = io->rho;
    c.factor = io->factor;
    c.rho_star = io->rho_star;
    c.d_rho_dt = io->d_rho_dt;
    c.kappa = io->kappa;
    c.kappa_v = io->kappa_v;
    c.neighbor_count = io->neighbor_count;
    c.v_star = io->v_star;
    FS_LAUNCH(k_dfsph_finalize, dim3(blocks_for(p.n, kElemBlock)), dim3(kElemBlock), 0, p.s, fa, c);
    return FS_OK;
}

int32_t fs_dfsph_step(fs_grid* g, fs_work* w, const fs_model* m, const fs_dfsph_params* prm, const fs_dfsph_io* io,
                      fs_stream_t stream)
{
    DfPlan p;
    int32_t rc = df_plan(g, w, m, prm, io, stream, "fs_dfsph_step", p);
    if (rc) return rc;
    if (p.n == 0) return FS_OK;
#define DF_TRY(call)    \
    do {                \
        rc = (call);    \
        if (rc) return rc; \
    } while (0)
    DF_TRY(df_density_factor(p, false));
    DF_TRY(df_nonpressure_predict(p));
    if (prm->warm_start) DF_TRY(df_correction(p, 0, true));
    for (int it = 0; it < prm->max_rho_iterations; ++it) {   // solver_dfsph.py:711-767
        DF_TRY(df_rate_kappa(p, 0));
        DF_TRY(df_residual(p, 0, it));
        DF_TRY(df_correction(p, 0, false));
    }
    DF_TRY(df_integrate(p));
    DF_TRY(df_density_factor(p, true));
    if (prm->warm_start) DF_TRY(df_correction(p, 1, true));
    for (int it = 0; it < prm->max_vel_iterations; ++it) {   // solver_dfsph.py:834-883
        DF_TRY(df_rate_kappa(p, 1));
        DF_TRY(df_residual(p, 1, it));
        DF_TRY(df_correction(p, 1, false));
    }
    return df_finalize(p);
#undef DF_TRY
}

// the passes of fs_dfsph_step, one call each
static int32_t df_pass(fs_grid* g, fs_work* w, const fs_model* m, const fs_dfsph_params* prm, const fs_dfsph_io* io,
                       fs_stream_t stream, const char* what, int which, int loop, int iteration)
{
    DfPlan p;
    int32_t rc = df_plan(g, w, m, prm, io, stream, what, p);
    if (rc) return rc;
    if (p.n == 0) return FS_OK;
    if ((which != 0 && which != 6) && g->num_points != m->n) {
        set_error("%s: the grid does not hold this model's particles (run fs_dfsph_density_factor first)", what);
        return FS_ERR_ARG;
    }
    if ((which == 2 || which == 3 || which == 4) && loop != 0 && loop != 1) {
        set_error("%s: loop must be 0 (density) or 1 (divergence)", what);
        return FS_ERR_ARG;
    }
    switch (which) {
    case 0: return df_density_factor(p, loop != 0);
    case 1: return df_nonpressure_predict(p);
    case 2: return df_rate_kappa(p, loop);
    case 3: return df_residual(p, loop, iteration);
    case 4: return df_correction(p, loop, iteration < 0);
    case 5: return df_integrate(p);
    default: return df_finalize(p);
    }
}

int32_t fs_particle_separation(fs_grid* g, fs_work* w, const fs_separation_params* prm, float* positions,
                               const int32_t* flags, float* delta, int32_t n, fs_stream_t stream)
{
    if (!g || !w || !prm) {
        set_error("fs_particle_separation: null grid, workspace or params");
        return FS_ERR_ARG;
    }
    if (n < 0 || n > g->capacity || n > w->capacity) {
        set_error("particle count %d exceeds grid capacity %d / workspace capacity %d", n, g->capacity, w->capacity);
        return FS_ERR_CAPACITY;
    }
    if (n == 0) return FS_OK;
    if (!positions || !flags) {
        set_error("fs_particle_separation: missing positions or flags");
        return FS_ERR_ARG;
    }
    if (!(prm->query_radius > 0.0f) || prm->iterations < 1) {
        set_error("fs_particle_separation: query_radius must be positive and iterations >= 1");
        return FS_ERR_ARG;
    }
    hipStream_t s = (hipStream_t)stream;
    GatherOp ga;
    ga.x = positions;
    ga.flags = flags;
    ga.pos = reinterpret_cast<float4*>(w->pos);
    ga.sflags = w->flags;
    SepArgs a;
    a.rest_dist = prm->rest_dist;
    a.strength = prm->strength;
    a.wall_band = prm->wall_band;
    for (int d = 0; d < 3; ++d) {
        a.lo[d] = prm->bound_lo[d];
        a.hi[d] = prm->bound_hi[d];
        a.clo[d] = prm->bound_lo[d] + prm->wall_margin;
        a.chi[d] = prm->bound_hi[d] - prm->wall_margin;
    }
    a.x = positions;
    a.delta = delta;
    for (int it = 0; it < prm->iterations; ++it) {
        PreOp pre;
        pre.x_in = positions;
        int32_t rc = build_grid(g, n, prm->query_radius, pre, ga, s);   // :1036
        if (rc) return rc;
        a.c = make_common(g, w, reinterpret_cast<const float4*>(w->pos), prm->query_radius, kListNone);
        a.c.cut2 = prm->rest_dist * prm->rest_dist * kCutSlack;
        FS_LAUNCH((k_separation<kScanBlock, kSepCap>), dim3(blocks_for(n, kScanBlock)), dim3(kScanBlock), 0, s,
                  a);   // :1037-1066
    }
    return FS_OK;
}

// One PBF substep as a sequence of passes sharing the grid and workspace.
// Predicted positions ping-pong between the workspace's two sorted position
// streams; which one holds them follows from the iteration alone.
struct PbfPlan {
    fs_grid* g;
    fs_work* w;
    const fs_model* m;
    const fs_pbf_params* prm;
    const fs_pbf_io* io;
    hipStream_t s;
    int n, nb;
    float h, dt;
    float4* sca;
    float4* vel;
    float4* bufs[2];
    PbfBounds bounds;
};

static int32_t pbf_plan(fs_grid* g, fs_work* w, const fs_model* m, const fs_pbf_params* prm, const fs_pbf_io* io,
                        fs_stream_t stream, const char* what, PbfPlan& p)
{
    int32_t rc = check_common(g, w, m);
    if (rc) return rc;
    if (!prm || !io) {
        set_error("%s: null params or io", what);
        return FS_ERR_ARG;
    }
    if (m->n > 0 && (!io->x_in || !io->v_in || !io->x_out || !io->v_out || !io->rho || !io->pbf_lambda ||
                     !io->neighbor_count || !io->delta_p || !io->accel)) {
        set_error("%s: missing arrays", what);
        return FS_ERR_ARG;
    }
    p.g = g;
    p.w = w;
    p.m = m;
    p.prm = prm;
    p.io = io;
    p.s = (hipStream_t)stream;
    p.n = m->n;
    p.nb = blocks_for(m->n, kPassBlock);
    p.h = prm->smoothing_length;
    p.dt = prm->dt;
    p.sca = reinterpret_cast<float4*>(w->sca);
    p.vel = reinterpret_cast<float4*>(w->vel);
    p.bufs[0] = reinterpret_cast<float4*>(w->pos);
    p.bufs[1] = reinterpret_cast<float4*>(w->pos_alt);
    for (int d = 0; d < 3; ++d) p.bounds.bound[d] = prm->bound[d];
    p.bounds.particle_radius = prm->particle_radius;
    p.bounds.boundary_epsilon = prm->boundary_epsilon;
    return FS_OK;
}

// The stream holding the predicted positions before iteration `it` (it =
// max_iterations: after the loop).  The non-pressure pass writes stream 1;
// each delta pass writes the other one; a per-iteration rebuild gathers into
// stream 0.
static int pbf_cur(const fs_pbf_params* prm, int it)
{
    if (prm->rebuild_grid_each_iteration) return it == 0 ? 1 : 0;
    return (1 + it) & 1;
}

// predict_velocity_with_gravity + predict_positions + apply_position_bounds
// fused into the build's hash, then the build (solver_pbf.py:702-737)
static int32_t pbf_predict(const PbfPlan& p)
{
    const fs_model* m = p.m;
    PreOp pre;
    pre.mode = kPrePredict;
    pre.x_in = p.io->x_in;
    pre.v_in = p.io->v_in;
    pre.x_out = p.w->x_pred;
    pre.v_out = p.w->v_pred;
    pre.flags = m->flags;
    pre.mask = m->mask;
    pre.world = m->world;
    pre.gravity = m->gravity;
    pre.dt = p.dt;
    for (int d = 0; d < 3; ++d) pre.bound[d] = p.prm->bound[d];
    pre.particle_radius = p.prm->particle_radius;
    pre.boundary_epsilon = p.prm->boundary_epsilon;
    return build_grid(p.g, p.n, p.h, pre, gather_for(m, p.w, p.w->x_pred, p.w->v_pred, p.vel), p.s);
}

// Positions change after the non-pressure pass and after every delta pass:
// the pass that follows a change scans (and records), the one after it, at
// the same positions, replays.
//
// compute_density (solver_pbf.py:738-751), or with `final_pass` the density
// at the final positions + update_velocity_from_positions (+ apply_bounds
// when there is no XSPH pass)  (:878-917, :939-954)
static int32_t pbf_density(const PbfPlan& p, bool final_pass)
{
    const int cur = final_pass ? pbf_cur(p.prm, p.prm->max_iterations) : 0;
    PbfDensityArgs da;
    da.c = make_common(p.g, p.w, p.bufs[cur], p.h, kListBuild);
    da.wvol = p.w->wvol;
    da.rho = p.io->rho;
    da.sca = p.sca;
    da.final_pass = final_pass ? 1 : 0;
    da.finalize = (final_pass && !(p.prm->xsph_c > 0.0f)) ? 1 : 0;
    da.x_in = p.io->x_in;
    da.v_in = p.io->v_in;
    da.dt = p.dt;
    da.damping = p.prm->boundary_damping;
    da.bounds = p.bounds;
    da.vel = p.vel;
    da.x_out = p.io->x_out;
    da.v_out = p.io->v_out;
    return with_variant<kTabSolverScan>(tuned(kTunePbfScan), [&]<Variant V>() -> int32_t {
        FS_LAUNCH((k_pbf_density<V.block, V.cap, V.lpp, V.unroll>), dim3(blocks_for(p.n, V.block / V.lpp)),
                  dim3(V.block), 0, p.s, da);
        return FS_OK;
    });
}

// acceleration_non_pressure + apply_non_pressure_prediction +
// apply_position_bounds  (solver_pbf.py:752-796)
static int32_t pbf_nonpressure(const PbfPlan& p)
{
    PbfNonPressureArgs na;
    na.c = make_common(p.g, p.w, p.bufs[0], p.h, kListUse);
    na.vel = p.vel;
    na.surface_tension = p.m->surface_tension;
    na.viscosity = p.m->viscosity;
    na.particle_radius = p.prm->particle_radius;
    na.dt = p.dt;
    na.bounds = p.bounds;
    na.accel = p.io->accel;
    na.x_pred = p.w->x_pred;
    na.v_pred = p.w->v_pred;
    na.pos_next = p.bufs[1];
    return with_variant<kTabSolverReplay>(tuned(kTunePbfReplay), [&]<Variant V>() -> int32_t {
        FS_LAUNCH((k_pbf_nonpressure<V.block, V.cap, V.unroll, V.lpp>), dim3(blocks_for(p.n, V.block / V.lpp)),
                  dim3(V.block), 0, p.s, na);
        return FS_OK;
    });
}

// compute_pbf_lambda of iteration `it`  (solver_pbf.py:801-817)
static int32_t pbf_lambda(const PbfPlan& p, int it)
{
    PbfLambdaArgs la;
    la.c = make_common(p.g, p.w, p.bufs[pbf_cur(p.prm, it)], p.h, kListBuild);
    la.rest_density = p.m->rest_density;
    la.lambda_epsilon = p.prm->lambda_epsilon;
    la.min_neighbors = p.prm->min_neighbors_for_lambda;
    la.pbf_lambda = p.io->pbf_lambda;
    la.rho = p.io->rho;
    la.neighbor_count = p.io->neighbor_count;
    la.sca = p.sca;
    return with_variant<kTabSolverScan>(tuned(kTunePbfScan), [&]<Variant V>() -> int32_t {
        FS_LAUNCH((k_pbf_lambda<V.block, V.cap, V.lpp, V.unroll>), dim3(blocks_for(p.n, V.block / V.lpp)),
                  dim3(V.block), 0, p.s, la);
        return FS_OK;
    });
}

// compute_delta_position + apply_delta_position (+ apply_position_bounds and
// the rebuild when rebuilding each iteration, or after the last iteration)
// of iteration `it`  (solver_pbf.py:818-875)
static int32_t pbf_delta(const PbfPlan& p, int it)
{
    const fs_pbf_params* prm = p.prm;
    const int cur = pbf_cur(prm, it);
    PbfDeltaArgs pa;
    pa.c = make_common(p.g, p.w, p.bufs[cur], p.h, kListUse);
    pa.sca = p.sca;
    pa.rest_density = p.m->rest_density;
    pa.scorr_k = prm->scorr_k;
    pa.scorr_n = prm->scorr_n;
    pa.scorr_w_delta_q = prm->scorr_w_delta_q;
    pa.max_delta_position = prm->max_delta_position;
    pa.apply_bounds = (it == prm->max_iterations - 1 || prm->rebuild_grid_each_iteration) ? 1 : 0;
    pa.bounds = p.bounds;
    pa.delta_p = p.io->delta_p;
    pa.x_pred = p.w->x_pred;
    pa.pos_next = p.bufs[cur ^ 1];
    int32_t rc = with_variant<kTabSolverReplay>(tuned(kTunePbfReplay), [&]<Variant V>() -> int32_t {
        if (prm->scorr_k != 0.0f)
            FS_LAUNCH((k_pbf_delta<V.block, V.cap, V.unroll, true, V.lpp>), dim3(blocks_for(p.n, V.block / V.lpp)),
                      dim3(V.block), 0, p.s, pa);
        else
            FS_LAUNCH((k_pbf_delta<V.block, V.cap, V.unroll, false, V.lpp>), dim3(blocks_for(p.n, V.block / V.lpp)),
                      dim3(V.block), 0, p.s, pa);
        return FS_OK;
    });
    if (rc) return rc;
    if (prm->rebuild_grid_each_iteration) {
        PreOp rp;
        rp.x_in = p.w->x_pred;
        return build_grid(p.g, p.n, p.h, rp, gather_for(p.m, p.w, p.w->x_pred, nullptr, nullptr), p.s);
    }
    return FS_OK;
}

// apply_xsph_viscosity + apply_bounds  (solver_pbf.py:919-954); without XSPH
// (xsph_c <= 0) the final density pass has finished the substep and this
// pass does nothing
static int32_t pbf_xsph_finalize(const PbfPlan& p)
{
    if (!(p.prm->xsph_c > 0.0f)) return FS_OK;
    PbfXsphArgs xa;
    xa.c = make_common(p.g, p.w, p.bufs[pbf_cur(p.prm, p.prm->max_iterations)], p.h, kListUse);
    xa.vel = p.vel;
    xa.xsph_c = p.prm->xsph_c;
    xa.damping = p.prm->boundary_damping;
    xa.bounds = p.bounds;
    xa.x_out = p.io->x_out;
    xa.v_out = p.io->v_out;
    return with_variant<kTabSolverReplay>(tuned(kTunePbfReplay), [&]<Variant V>() -> int32_t {
        FS_LAUNCH((k_pbf_xsph<V.block, V.cap, V.unroll, V.lpp>), dim3(blocks_for(p.n, V.block / V.lpp)), dim3(V.block),
                  0, p.s, xa);
        return FS_OK;
    });
}

int32_t fs_pbf_step(fs_grid* g, fs_work* w, const fs_model* m, const fs_pbf_params* prm, const fs_pbf_io* io,
                    fs_stream_t stream)
{
    PbfPlan p;
    int32_t rc = pbf_plan(g, w, m, prm, io, stream, "fs_pbf_step", p);
    if (rc) return rc;
    if (p.n == 0) return FS_OK;
#define PBF_TRY(call)      \
    do {                   \
        rc = (call);       \
        if (rc) return rc; \
    } while (0)
    PBF_TRY(pbf_predict(p));
    PBF_TRY(pbf_density(p, false));
    PBF_TRY(pbf_nonpressure(p));
    for (int it = 0; it < prm->max_iterations; ++it) {   // solver_pbf.py:800-858
        PBF_TRY(pbf_lambda(p, it));
        PBF_TRY(pbf_delta(p, it));
    }
    PBF_TRY(pbf_density(p, true));
    return pbf_xsph_finalize(p);
#undef PBF_TRY
}

// the passes of fs_pbf_step, one call each
static int32_t pbf_pass(fs_grid* g, fs_work* w, const fs_model* m, const fs_pbf_params* prm, const fs_pbf_io* io,
                        fs_stream_t stream, const char* what, int which, int iteration)
{
    PbfPlan p;
    int32_t rc = pbf_plan(g, w, m, prm, io, stream, what, p);
    if (rc) return rc;
    if (p.n == 0) return FS_OK;
    if (which != 0 && g->num_points != m->n) {
        set_error("%s: the grid does not hold this model's particles (run fs_pbf_predict first)", what);
        return FS_ERR_ARG;
    }
    if ((which == 3 || which == 4) && (iteration < 0 || iteration >= prm->max_iterations)) {
        set_error("%s: iteration %d outside [0, max_iterations = %d)", what, iteration, prm->max_iterations);
        return FS_ERR_ARG;
    }
    switch (which) {
    case 0: return pbf_predict(p);
    case 1: return pbf_density(p, false);
    case 2: return pbf_nonpressure(p);
    case 3: return pbf_lambda(p, iteration);
    case 4: return pbf_delta(p, iteration);
    case 5: return pbf_density(p, true);
    default: return pbf_xsph_finalize(p);
    }
}

int32_t fs_dfsph_density_factor(fs_grid* g, fs_work* w, const fs_model* m, const fs_dfsph_params* prm,
                                const fs_dfsph_io* io, int32_t after_integrate, fs_stream_t stream)
{
    return df_pass(g, w, m, prm, io, stream, "fs_dfsph_density_factor", 0, after_integrate ? 1 : 0, 0);
}
int32_t fs_dfsph_nonpressure_predict(fs_grid* g, fs_work* w, const fs_model* m, const fs_dfsph_params* prm,
                                     const fs_dfsph_io* io, fs_stream_t stream)
{
    return df_pass(g, w, m, prm, io, stream, "fs_dfsph_nonpressure_predict", 1, 0, 0);
}
int32_t fs_dfsph_rate_kappa(fs_grid* g, fs_work* w, const fs_model* m, const fs_dfsph_params* prm,
                            const fs_dfsph_io* io, int32_t loop, fs_stream_t stream)
{
    return df_pass(g, w, m, prm, io, stream, "fs_dfsph_rate_kappa", 2, loop, 0);
}
int32_t fs_dfsph_residual(fs_grid* g, fs_work* w, const fs_model* m, const fs_dfsph_params* prm,
                          const fs_dfsph_io* io, int32_t loop, int32_t iteration, fs_stream_t stream)
{
    return df_pass(g, w, m, prm, io, stream, "fs_dfsph_residual", 3, loop, iteration);
}
int32_t fs_dfsph_correction(fs_grid* g, fs_work* w, const fs_model* m, const fs_dfsph_params* prm,
                            const fs_dfsph_io* io, int32_t loop, int32_t warm_start, fs_stream_t stream)
{
    return df_pass(g, w, m, prm, io, stream, "fs_dfsph_correction", 4, loop, warm_start ? -1 : 0);
}
int32_t fs_dfsph_integrate(fs_grid* g, fs_work* w, const fs_model* m, const fs_dfsph_params* prm,
                           const fs_dfsph_io* io, fs_stream_t stream)
{
    return df_pass(g, w, m, prm, io, stream, "fs_dfsph_integrate", 5, 0, 0);
}
int32_t fs_dfsph_finalize(fs_grid* g, fs_work* w, const fs_model* m, const fs_dfsph_params* prm,
                          const fs_dfsph_io* io, fs_stream_t stream)
{
    return df_pass(g, w, m, prm, io, stream, "fs_dfsph_finalize", 6, 0, 0);
}

int32_t fs_pbf_predict(fs_grid* g, fs_work* w, const fs_model* m, const fs_pbf_params* prm, const fs_pbf_io* io,
                       fs_stream_t stream)
{
    return pbf_pass(g, w, m, prm, io, stream, "fs_pbf_predict", 0, 0);
}
int32_t fs_pbf_density(fs_grid* g, fs_work* w, const fs_model* m, const fs_pbf_params* prm, const fs_pbf_io* io,
                       fs_stream_t stream)
{
    return pbf_pass(g, w, m, prm, io, stream, "fs_pbf_density", 1, 0);
}
int32_t fs_pbf_nonpressure(fs_grid* g, fs_work* w, const fs_model* m, const fs_pbf_params* prm,
                           const fs_pbf_io* io, fs_stream_t stream)
{
    return pbf_pass(g, w, m, prm, io, stream, "fs_pbf_nonpressure", 2, 0);
}
int32_t fs_pbf_lambda(fs_grid* g, fs_work* w, const fs_model* m, const fs_pbf_params* prm, const fs_pbf_io* io,
                      int32_t iteration, fs_stream_t stream)
{
    return pbf_pass(g, w, m, prm, io, stream, "fs_pbf_lambda", 3, iteration);
}
int32_t fs_pbf_delta(fs_grid* g, fs_work* w, const fs_model* m, const fs_pbf_params* prm, const fs_pbf_io* io,
                     int32_t iteration, fs_stream_t stream)
{
    return pbf_pass(g, w, m, prm, io, stream, "fs_pbf_delta", 4, iteration);
}
int32_t fs_pbf_final_density(fs_grid* g, fs_work* w, const fs_model* m, const fs_pbf_params* prm,
                             const fs_pbf_io* io, fs_stream_t stream)
{
    return pbf_pass(g, w, m, prm, io, stream, "fs_pbf_final_density", 5, 0);
}
int32_t fs_pbf_xsph_finalize(fs_grid* g, fs_work* w, const fs_model* m, const fs_pbf_params* prm,
                             const fs_pbf_io* io, fs_stream_t stream)
{
    return pbf_pass(g, w, m, prm, io, stream, "fs_pbf_xsph_finalize", 6, 0);
}

}  // extern "C"
