// fs_coop.h -- the cooperative cell scan of the scanning passes.
//
// Particles whose hash_grid_query(p, r) visits the same cells (the same cell
// range on every axis; in practice the particles of one cell) share every
// candidate, in the same visit order.  The per-lane scan tests each
// particle's ~200-300 candidates on its own lane: lanes of one wave run
// different row lengths and finish at different times, and every lane repeats
// the row bookkeeping.  Here a workgroup takes a tile of 64 consecutive
// sorted slots and, for each group of its particles with one query:
//
//  1. stages the query's candidate slots, in visit order, 64 per chunk, in
//     LDS (the 27-cell neighbourhood, once for the whole group);
//  2. loads each chunk's candidate positions once (one candidate per lane,
//     kCoopBatch chunks at a time) and tests them against every particle of
//     the group, whose relative position and threshold are read with
//     readlane: per particle and chunk three FMAs, one compare and a
//     two-word select;
//  3. keeps, per particle and chunk, the 64-bit mask of candidates that pass
//     (bit order = visit order).
//
// Measured on gfx950 it is slower than the per-lane scan, at the developed
// dam break (groups of about 6 particles: the per-group set-up costs as much
// as the tests it shares) and on clustered points alike; the autotuner keeps
// two variants of the WCSPH density pass and measures them on every input it
// tunes.
//
// The test is the identity |p-c|^2 = |c'|^2 - 2 p'.c' + |p'|^2 on
// coordinates relative to the group's first particle o (p' = p - o,
// c' = c - o), compared with keep2 (1 + 1e-3) - |p'|^2.  It is a superset
// test: group members lie within 2r of o on every axis, so every candidate
// the exact prefilter keeps (d2 <= keep2, |c'| <= 4.5 r) rounds within about
// 300 ulp(r^2) of its exact value, far inside the 1e-3 widening; a NaN
// candidate passes both tests.  The walk below re-tests each passing
// candidate with the per-lane scan's own prefilter (prefilter_d2 <= keep2),
// so the survivors, their order, the recorded lists and every sum are those
// of the per-lane scan.
//
// The walk then runs LPP lanes per particle (64 LPP threads per workgroup,
// one wave per 64 lanes; the test phase gives the tile's groups to the waves
// in turn): the group's lanes take survivors in rank order (lane q the ranks
// q, q + LPP, ...), evaluate their terms and fold them in visit order
// through DPP broadcasts, as the list replay does.
//
// Particles the cooperative path does not take -- a non-finite position, a
// smoothing length outside [2^-20, 2^20], a group member more than 2r from o,
// or a group whose chunks overflow the workgroup's LDS (a clump with more
// than 64 MAXC candidates, or more than NCH chunks in the tile) -- scan on
// their own (scan_direct, the group's first lane), with the same result.
#pragma once

#include "fs_common.h"

namespace fs {

constexpr int kCoopTile = 64;            // particles per workgroup tile
constexpr int kCoopBatch = 8;            // chunks whose candidates a wave holds in registers at once
constexpr float kCoopMargin = 1.0e-3f;   // relative widening of the superset test

template <int NCH, int MAXC>
struct CoopShared {
    int slots[NCH * kCoopTile];              // chunk c's candidate slots at [c*64, c*64+64)
    uint32_t masks[2 * MAXC * kCoopTile];    // word w (= 2j + half) of chunk j of particle p at [w*64 + p]
                                             // (MAXC: chunks per group)
    int2 ginfo[kCoopTile];                   // per particle: first chunk, mask words; words -1: scan alone
};

// The cell range of hash_grid_query(p, r) (QueryRows::begin's first steps):
// equal keys give equal visit sequences.
struct QueryKey {
    int xs, ys, zs, xe, ye, ze;
};
__device__ __forceinline__ QueryKey query_key(const GridView& g, float qx, float qy, float qz, float r)
{
    QueryKey k;
    k.xs = cell_coord((qx - r) * g.inv, g.round_mode);
    k.ys = cell_coord((qy - r) * g.inv, g.round_mode);
    k.zs = cell_coord((qz - r) * g.inv, g.round_mode);
    k.xe = min(cell_coord((qx + r) * g.inv, g.round_mode), k.xs + g.dx - 1);
    k.ye = min(cell_coord((qy + r) * g.inv, g.round_mode), k.ys + g.dy - 1);
    k.ze = min(cell_coord((qz + r) * g.inv, g.round_mode), k.zs + g.dz - 1);
    return k;
}

__device__ __forceinline__ float readlane_f(float v, int lane)
{
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}

__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Whether the cooperative test is exact for a pass's radius and keep
// radius (uniform).
__device__ __forceinline__ bool coop_usable(float r, float keep2)
{
    return r >= 0x1p-20f && r <= 0x1p20f && keep2 >= 0.0f && keep2 <= 1.01f * (r * r);
}

// Test phase for the tile of particles s0 .. s0+63: lane l of every wave
// holds particle s0 + l; W waves (the workgroup) share the groups.  sid is
// the lane's scene (0 for a plain pass); view(sid, g, r, keep2) gives the
// scene's grid view, query radius and keep radius, and returns false for a
// scene the pass skips.  Fills sh.slots, sh.masks and sh.ginfo; ends with a
// workgroup barrier.  Every thread of the workgroup must call it.
//
// Per group the memory round trips are batched: the query's row pieces
// (QueryRows order: z, then y, then the x run's parts) are enumerated one per
// lane and their slot ranges loaded together; the pieces' slots are written
// to the group's LDS chunks in visit order; all chunks' candidate positions
// are loaded together; the members' relative positions are read with
// readlane.
template <int NCH, int MAXC, int W, class SceneView>
__device__ __forceinline__ void coop_test(const float4* __restrict__ pos, const int* __restrict__ sflags, int n,
                                          int s0, int sid, SceneView&& view, CoopShared<NCH, MAXC>& sh)
{
    static_assert(NCH % W == 0, "chunk slots per wave");
    const int lane = (int)(threadIdx.x & 63);
    const int w = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int s = s0 + lane;
    const bool in = s < n;
    const float4 xi = in ? pos[s] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    const int fl = in ? sflags[s] : 0;
    GridView gv;
    float rv = 0.0f, kv = 0.0f;
    const bool live = in && view(sid, gv, rv, kv);
    const bool fin = __builtin_isfinite(xi.x) && __builtin_isfinite(xi.y) && __builtin_isfinite(xi.z);
    const bool elig = live && coop_usable(rv, kv) && (fl & kActive) && fin;
    if (w == 0) sh.ginfo[lane] = make_int2(0, (fl & kActive) ? -1 : 0);
    __syncthreads();
    QueryKey kq{0, 0, 0, 0, 0, 0};
    if (elig) kq = query_key(gv, xi.x, xi.y, xi.z, rv);
    constexpr int kPerWave = NCH / W;
    int cnext = w * kPerWave;
    const int cend = cnext + kPerWave;
    uint64_t remaining = __ballot(elig);
    int gidx = 0;
    while (remaining) {
        const int L = (int)__builtin_ctzll(remaining);
        const int lsid = __builtin_amdgcn_readlane(sid, L);
        const int kxs = __builtin_amdgcn_readlane(kq.xs, L), kys = __builtin_amdgcn_readlane(kq.ys, L);
        const int kzs = __builtin_amdgcn_readlane(kq.zs, L), kxe = __builtin_amdgcn_readlane(kq.xe, L);
        const int kye = __builtin_amdgcn_readlane(kq.ye, L), kze = __builtin_amdgcn_readlane(kq.ze, L);
        const bool same = sid == lsid && kq.xs == kxs && kq.ys == kys && kq.zs == kzs && kq.xe == kxe &&
                          kq.ye == kye && kq.ze == kze;
        const uint64_t gm0 = __ballot(same) & remaining;
        remaining &= ~gm0;
        if ((gidx++ % W) != w) continue;
        GridView g;
        float r, keep2;
        view(lsid, g, r, keep2);
        const float keep2c = keep2 * (1.0f + kCoopMargin);
        const float spread = 2.0f * r;
        const float ox = readlane_f(xi.x, L), oy = readlane_f(xi.y, L), oz = readlane_f(xi.z, L);
        const float ax = xi.x - ox, ay = xi.y - oy, az = xi.z - oz;
        const bool near_o = fabsf(ax) <= spread && fabsf(ay) <= spread && fabsf(az) <= spread;
        const uint64_t gm = gm0 & __ballot(near_o);
        if (gm == 0) continue;
        const bool member = (gm >> lane) & 1ull;
        // the query's pieces (QueryRows::begin / load_row)
        int xs = kxs, lenx = kxe - kxs + 1;
        const int ny = kye - kys + 1, nz = kze - kzs + 1;
        int xrep = 0;
        if (lenx > 0 && xs < -kOrigin) {
            xrep = min(kxe, -kOrigin) - xs;
            xs += xrep;
            lenx -= xrep;
        }
        const int hx0 = hash_axis(xs, g.dx);
        const bool empty = lenx <= 0 || ny <= 0 || nz <= 0;
        const int seg1 = min(lenx, g.dx - hx0);
        const int seg2 = lenx - seg1;
        const int ppr = xrep + 1 + (seg2 > 0 ? 1 : 0);   // pieces per row
        const int np = empty ? 0 : ny * nz * ppr;
        bool over = np > 64;
        int pk = 0, plen = 0;   // piece `lane`: first slot and length
        if (!over && lane < np) {
            const int row = lane / ppr;
            const int part = lane - row * ppr;
            const int zi = row / ny;
            const int yi = row - zi * ny;
            const int a = g.base + hash_axis(kzs + zi, g.dz) * g.dxy + hash_axis(kys + yi, g.dy) * g.dx +
                          (part == xrep ? hx0 : 0);
            const int b = a + (part < xrep ? 1 : (part == xrep ? seg1 : seg2));
            pk = min(max(g.start[a], 0), g.npts);
            plen = min(max(g.start[b], pk), g.npts) - pk;
        }
        // the pieces' slots, in visit order, into the group's chunks
        const int cfirst = cnext;
        int total = 0;
        if (!over) {
            for (int j = 0; j < np; ++j) {
                const int len = __builtin_amdgcn_readlane(plen, j);
                if (len <= 0) continue;
                if (total + len > 64 * MAXC || cfirst * 64 + total + len > cend * 64) {
                    over = true;
                    break;
                }
                const int k0 = __builtin_amdgcn_readlane(pk, j);
                for (int o = 0; o < len; o += 64)
                    if (o + lane < len) sh.slots[cfirst * 64 + total + o + lane] = k0 + o + lane;
                total += len;
            }
        }
        if (over) continue;   // its members scan alone (ginfo stays -1)
        const int nch = (total + 63) >> 6;
        cnext = cfirst + nch;
        wave_lds_sync();
        // the chunks' candidates kCoopBatch chunks at a time: c' = c - o and
        // |c'|^2, tested against every member
        const __amdgpu_buffer_rsrc_t rs = float4_rsrc(pos, g.npts);
        // members: -2p' and the threshold keep2' - |p'|^2
        const float nax = -2.0f * ax, nay = -2.0f * ay, naz = -2.0f * az;
        const float thr = __builtin_fmaf(-az, az, __builtin_fmaf(-ay, ay, __builtin_fmaf(-ax, ax, keep2c)));
        for (int cb = 0; cb < nch; cb += kCoopBatch) {
            const int nb = min(kCoopBatch, nch - cb);
            float cx[kCoopBatch], cy[kCoopBatch], cz[kCoopBatch], cc[kCoopBatch];
            int slot[kCoopBatch];
#pragma unroll
            for (int c = 0; c < kCoopBatch; ++c)
                slot[c] = (c < nb && 64 * (cb + c) + lane < total) ? sh.slots[(cfirst + cb + c) * 64 + lane] : 0;
#pragma unroll
            for (int c = 0; c < kCoopBatch; ++c) {
                if (c < nb) {
                    const float4 q = ld_f4(rs, (int)((unsigned)slot[c] * 16u), 0);
                    cx[c] = q.x - ox;
                    cy[c] = q.y - oy;
                    cz[c] = q.z - oz;
                    cc[c] = __builtin_fmaf(cz[c], cz[c], __builtin_fmaf(cy[c], cy[c], cx[c] * cx[c]));
                }
            }
            int mlo[kCoopBatch], mhi[kCoopBatch];
#pragma unroll
            for (int c = 0; c < kCoopBatch; ++c) mlo[c] = mhi[c] = 0;
            uint64_t pm = gm;
            while (pm) {
                const int p = (int)__builtin_ctzll(pm);
                pm &= pm - 1ull;
                const float px = readlane_f(nax, p), py = readlane_f(nay, p), pz = readlane_f(naz, p);
                const float pt = readlane_f(thr, p);
                const bool me = lane == p;
#pragma unroll
                for (int c = 0; c < kCoopBatch; ++c) {
                    if (c < nb) {
                        float t = __builtin_fmaf(px, cx[c], cc[c]);
                        t = __builtin_fmaf(py, cy[c], t);
                        t = __builtin_fmaf(pz, cz[c], t);
                        const int left = total - 64 * (cb + c);
                        const uint64_t vm = left >= 64 ? ~0ull : ((1ull << left) - 1ull);
                        const uint64_t hm = __ballot(!(t > pt)) & vm;
                        if (me) {
                            mlo[c] = (int)(uint32_t)hm;
                            mhi[c] = (int)(uint32_t)(hm >> 32);
                        }
                    }
                }
            }
            if (member) {
#pragma unroll
                for (int c = 0; c < kCoopBatch; ++c) {
                    if (c < nb) {
                        sh.masks[(2 * (cb + c)) * kCoopTile + lane] = (uint32_t)mlo[c];
                        sh.masks[(2 * (cb + c) + 1) * kCoopTile + lane] = (uint32_t)mhi[c];
                    }
                }
            }
        }
        if (member) {
            sh.ginfo[lane] = make_int2(cfirst, 2 * nch);
        }
    }
    __syncthreads();
}

// Sum of an integer over the group's lower lanes and over the whole group
// (LPP = 1 included).
template <int LPP>
__device__ __forceinline__ int coop_prefix(int x, int& total)
{
    if constexpr (LPP == 1) {
        total = x;
        return 0;
    } else {
        return group_prefix<LPP>(x, total);
    }
}

// Survivors of particle p (slot s) in visit order from the test phase's
// masks, LPP lanes per particle (every lane of the group active and walking
// in step), U survivors per lane fetched together.  Each candidate the
// superset test passed is re-tested with the per-lane scan's prefilter; the
// survivors are recorded (active ones, kListBuild) and folded exactly as
// scan_split does.
template <int NCH, int MAXC, int LPP, int U, class Load, class Term, class Fold>
__device__ __forceinline__ void coop_walk(const CoopShared<NCH, MAXC>& sh, int p, int s, int2 gi, float px, float py,
                                          float pz, float keep2, const NbrList& L, Load& load, Term& term, Fold& fold)
{
    const int q = (int)(threadIdx.x & (LPP - 1));
    const bool record = L.mode == kListBuild;
    const ptrdiff_t n = L.n;
    const int nw = gi.y;
    const int cb = gi.x * kCoopTile;
    int wi = 0;
    uint32_t cur = nw > 0 ? sh.masks[p] : 0u;
    int written = 0;
    auto next = [&]() -> int {
        while (cur == 0u) {
            if (++wi >= nw) return -1;
            cur = sh.masks[wi * kCoopTile + p];
        }
        const int b = __builtin_ctz(cur);
        cur &= cur - 1u;
        return sh.slots[cb + wi * 32 + b];
    };
    bool done = false;
    while (!done) {
        int k[U];
        int last = 0;
#pragma unroll
        for (int u = 0; u < U; ++u) {
            int mine = -1;
#pragma unroll
            for (int ql = 0; ql < LPP; ++ql) {
                last = next();
                if (ql == q) mine = last;
            }
            k[u] = mine;
        }
        decltype(load(0)) d[U];
        int v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            d[u] = load(k[u] < 0 ? s : k[u]);
            const float ddx = px - d[u].x.x, ddy = py - d[u].x.y, ddz = pz - d[u].x.z;
            v[u] = (k[u] >= 0 && !(prefilter_d2(ddx, ddy, ddz) > keep2)) ? 1 : 0;
        }
        if (record) {
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int act = (v[u] && (d[u].flags & kActive)) ? 1 : 0;
                int nact;
                const int at = written + coop_prefix<LPP>(act, nact);
                if (act && at < L.cap) L.idx[s + (ptrdiff_t)at * n] = k[u];
                written += nact;
            }
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const auto t = term(k[u] < 0 ? s : k[u], d[u]);
            if (group_bcast<LPP, 0>(v[u])) fold(group_bcast<LPP, 0>(t));
            if constexpr (LPP >= 2)
                if (group_bcast<LPP, 1>(v[u])) fold(group_bcast<LPP, 1>(t));
            if constexpr (LPP == 4) {
                if (group_bcast<LPP, 2>(v[u])) fold(group_bcast<LPP, 2>(t));
                if (group_bcast<LPP, 3>(v[u])) fold(group_bcast<LPP, 3>(t));
            }
        }
        done = last < 0;
    }
    if (record && q == 0) store_list_count(L, s, written);
}

// One lane's scan of hash_grid_query(p, r) with no LDS: every candidate in
// visit order, the prefilter, then load/record/term/fold for the survivors
// (the cooperative scan's fallback).  count(k) runs for every candidate.
template <class Count, class Load, class Term, class Fold>
__device__ __forceinline__ void scan_direct(const GridView& g, const float4* __restrict__ pos, int s, float px,
                                            float py, float pz, float r, float keep2, const NbrList& L, Count&& count,
                                            Load& load, Term& term, Fold& fold)
{
    const bool record = L.mode == kListBuild;
    int written = 0;
    int* __restrict__ out = L.idx + s;
    QueryRows it;
    it.begin(g, px, py, pz, r);
    if (it.valid()) {
        do {
            for (int k = it.k; k < it.kend; ++k) {
                count(k);
                const float4 c = pos[k];
                if (!(prefilter_d2(px - c.x, py - c.y, pz - c.z) > keep2)) {
                    const auto d = load(k);
                    if (record && (d.flags & kActive)) {
                        if (written < L.cap) {
                            *out = k;
                            out += L.n;
                        }
                        ++written;
                    }
                    fold(term(k, d));
                }
            }
        } while (it.advance(g));
    }
    if (record) store_list_count(L, s, written);
}

// Neighbour loop of particle p of the tile (slot s) after coop_test: the
// masks' walk, or the group's first lane scanning alone.  After a scan alone
// only that lane's sums are the particle's.
template <int NCH, int MAXC, int LPP, int U, class Load, class Term, class Fold>
__device__ __forceinline__ void coop_neighbors(const CoopShared<NCH, MAXC>& sh, const GridView& g,
                                               const float4* __restrict__ pos, int p, int s, float px, float py,
                                               float pz, float r, float keep2, const NbrList& L, Load& load,
                                               Term& term, Fold& fold)
{
    const int2 gi = sh.ginfo[p];
    if (gi.y >= 0) {
        coop_walk<NCH, MAXC, LPP, U>(sh, p, s, gi, px, py, pz, keep2, L, load, term, fold);
    } else if ((threadIdx.x & (LPP - 1)) == 0) {
        scan_direct(g, pos, s, px, py, pz, r, keep2, L, NoCount{}, load, term, fold);
    }
}

}  // namespace fs
