// fs_common.h -- device-side building blocks shared by the gfx950 kernels:
// hash-grid coordinates and hashing (Warp HashGrid semantics), the SPH kernel
// functions in the reference's float32 evaluation order, and the compacting
// neighbour loop.
//
// Floating point: every file is compiled with -ffp-contract=off and correctly
// rounded float32 division/sqrt, so each expression below is evaluated exactly
// as written (no FMA contraction), in the reference's left-to-right order.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

#include "fluidsim.h"

namespace fs {

constexpr int kOrigin = 1 << 20;
// Largest particle / point count: positions are float4s read through buffer
// resources with 32-bit byte offsets (< 4 GiB).
constexpr int kMaxParticles = FS_MAX_PARTICLES;
constexpr float kPi = 3.14159265358979323846f;

// ---------------------------------------------------------------------------
// error reporting (host)
// ---------------------------------------------------------------------------
void set_error(const char* fmt, ...);
int32_t hip_check(hipError_t e, const char* what);
#define FS_LAUNCH_CHECK(what)                                             \
    do {                                                                  \
        hipError_t e_ = hipGetLastError();                                \
        if (e_ != hipSuccess) return ::fs::hip_check(e_, what);           \
    } while (0)

// ---------------------------------------------------------------------------
// small float3 algebra with explicit evaluation order
// ---------------------------------------------------------------------------
struct v3 {
    float x, y, z;
};
__device__ __forceinline__ v3 V(float x, float y, float z) { return v3{x, y, z}; }
__device__ __forceinline__ v3 vadd(v3 a, v3 b) { return V(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ v3 vsub(v3 a, v3 b) { return V(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ v3 smul(float s, v3 a) { return V(s * a.x, s * a.y, s * a.z); }
__device__ __forceinline__ v3 muls(v3 a, float s) { return V(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ v3 vdivs(v3 a, float s) { return V(a.x / s, a.y / s, a.z / s); }
__device__ __forceinline__ float vdot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ float vlen(v3 a) { return sqrtf(vdot(a, a)); }
__device__ __forceinline__ float fmax_(float a, float b) { return a > b ? a : b; }
__device__ __forceinline__ float fmin_(float a, float b) { return a < b ? a : b; }
__device__ __forceinline__ float fpow_(float a, float b) { return (float)pow((double)a, (double)b); }
__device__ __forceinline__ v3 ld3(const float* __restrict__ p, int i) { return V(p[3 * i], p[3 * i + 1], p[3 * i + 2]); }
__device__ __forceinline__ void st3(float* __restrict__ p, int i, v3 v)
{
    p[3 * i] = v.x;
    p[3 * i + 1] = v.y;
    p[3 * i + 2] = v.z;
}
__device__ __forceinline__ v3 xyz(float4 q) { return V(q.x, q.y, q.z); }

// ParticleFlags.ACTIVE is bit 0 (solver_wcsph.py:53-60).  Sorted copies keep
// bit 0 = active and bit 1 = dynamic (active and sph:particle_mask == 0).
constexpr int kActive = 1;
constexpr int kDynamic = 2;

// ---------------------------------------------------------------------------
// kernel functions (reference float32 order)
// ---------------------------------------------------------------------------

// get_cubic  solver_wcsph.py:20-33
__device__ __forceinline__ float cubic_w(float r_norm, float h)
{
    float res = 0.0f;
    float k = 8.0f / (kPi * (h * h * h));
    float q = r_norm / h;
    if (q <= 1.0f) {
        if (q <= 0.5f) {
            float q2 = q * q;
            float q3 = q2 * q;
            res = k * (6.0f * q3 - 6.0f * q2 + 1.0f);
        } else {
            float t = 1.0f - q;
            res = 2.0f * k * (t * t * t);
        }
    }
    return res;
}

// get_cubic_derivative  solver_wcsph.py:36-50, with r_norm = wp.length(r)
// supplied by the caller (it is the same float value every time)
__device__ __forceinline__ v3 cubic_grad_r(v3 r, float r_norm, float h)
{
    v3 res = V(0.0f, 0.0f, 0.0f);
    float k = 8.0f / (kPi * (h * h * h));
    float q = r_norm / h;
    if (r_norm > 1.0e-5f && q <= 1.0f) {
        v3 grad_q = vdivs(r, r_norm * h);
        if (q <= 0.5f) {
            res = smul(6.0f * k * q * (3.0f * q - 2.0f), grad_q);
        } else {
            float q_term = 1.0f - q;
            res = smul(-6.0f * k * q_term * q_term, grad_q);
        }
    }
    return res;
}
__device__ __forceinline__ v3 cubic_grad(v3 r, float h) { return cubic_grad_r(r, vlen(r), h); }

// ---------------------------------------------------------------------------
// Correctly rounded quotients and square roots without the range wrapping
// ---------------------------------------------------------------------------
// With correctly rounded float32 division, hipcc expands x / d into
//   v_div_scale of d and of x, r = v_rcp(d) refined by one Newton step
//   r += r (1 - d r), q = x r, two corrections q += r (x - d q) (the second
//   one in v_div_fmas), and v_div_fixup;
// and sqrtf(x) into v_sqrt_f32 followed by a one-ulp correction from two FMA
// residuals, wrapped in a 2^32 scale for x < 2^-96 and a class test that
// passes 0 and +inf through.  The wrapping only acts outside the ranges below:
// for d normal with a normal reciprocal, x = 0 or |x| >= 2^-103, a normal
// quotient and exponents of x and d less than 96 apart, v_div_scale returns
// its operand and clears VCC (v_div_fmas is then a plain FMA) and v_div_fixup
// only gives a zero quotient the sign of x; sqrt keeps x = 0 and x >= 2^-96
// unscaled.  div_nr and sqrt_nr run the same core on operands their callers
// keep inside those ranges, and div_nr takes the refined reciprocal as an
// argument, so quotients with one divisor (a gradient's three components, a
// pass's r / h) share it.  fs_check_fast_math compares them with the IEEE
// forms on the GPU.
__device__ __forceinline__ float rcp_nr(float d)
{
    const float r = __builtin_amdgcn_rcpf(d);
    return __builtin_fmaf(__builtin_fmaf(-d, r, 1.0f), r, r);
}
// x / d for d > 0, r = rcp_nr(d)
__device__ __forceinline__ float div_nr(float x, float d, float r)
{
    float q = x * r;
    q = __builtin_fmaf(__builtin_fmaf(-d, q, x), r, q);
    q = __builtin_fmaf(__builtin_fmaf(-d, q, x), r, q);
    return __builtin_copysignf(q, x);
}
// x / d for x >= 0 (+0 stays +0 without the sign copy)
__device__ __forceinline__ float div_nr_pos(float x, float d, float r)
{
    float q = x * r;
    q = __builtin_fmaf(__builtin_fmaf(-d, q, x), r, q);
    return __builtin_fmaf(__builtin_fmaf(-d, q, x), r, q);
}
// sqrtf(x) for x = 0 or 2^-96 <= x < inf
__device__ __forceinline__ float sqrt_nr(float x)
{
    const float s = __builtin_amdgcn_sqrtf(x);
    const float lo = __uint_as_float(__float_as_uint(s) - 1u);
    const float hi = __uint_as_float(__float_as_uint(s) + 1u);
    const float t = __builtin_fmaf(-lo, s, x) <= 0.0f ? lo : s;
    return __builtin_fmaf(-hi, s, x) > 0.0f ? hi : t;
}

// vlen(a) through sqrt_nr where its range holds (a.a = 0 or 2^-96 <= a.a <
// inf), the correctly rounded sqrtf elsewhere (a tiny, infinite or NaN a.a):
// equal to vlen(a) for every a, in about half the instructions.
__device__ __forceinline__ float vlen_nr(v3 a)
{
    const float d2 = vdot(a, a);
    float len = sqrt_nr(d2);
    // outside the range (bits of 2^-96: 0x0f800000, of inf: 0x7f800000; a
    // NaN is above either): a branch that a wave skips unless a lane needs
    // it (the empty asm keeps the compiler from evaluating sqrtf on every
    // lane ahead of the branch)
    const uint32_t b = __float_as_uint(d2);
    const bool in_range = (b == 0u) | (b - 0x0f800000u < 0x70000000u);
    if (!in_range) {
        float x = d2;
        asm volatile("" : "+v"(x));
        len = sqrtf(x);
    }
    return len;
}

// A pass's smoothing length with its refined reciprocal.  fast: h lies in
// [2^-20, 2^20], where every quotient the kernel functions form from an
// in-range separation (sep_in_range) is inside div_nr's ranges.
struct Smooth {
    float h;
    float rh;
    bool fast;
};
__device__ __forceinline__ Smooth smooth_of(float h)
{
    const bool fast = h >= 0x1p-20f && h <= 0x1p20f;
    return Smooth{h, fast ? rcp_nr(h) : 1.0f, fast};
}
// Every component of r is 0 or of magnitude >= 2^-48: then r.r is 0 or
// >= 2^-96 and every nonzero numerator of the gradient is >= 2^-48.
// (bits << 1) - 1 drops the sign and maps +-0 to 0xffffffff.
__device__ __forceinline__ bool sep_in_range(v3 r)
{
    const uint32_t kx = (__float_as_uint(r.x) << 1) - 1u;
    const uint32_t ky = (__float_as_uint(r.y) << 1) - 1u;
    const uint32_t kz = (__float_as_uint(r.z) << 1) - 1u;
    return min(kx, min(ky, kz)) >= (0x27800000u << 1) - 1u;   // 0x27800000 = 2^-48
}

// get_cubic with q = r_norm / h from the shared reciprocal.  Equal to
// cubic_w(r_norm, h) for every r_norm in [0, 2^100 h]: div_nr's ranges hold
// for r_norm = 0 and r_norm >= 2^-103, and below that both quotients are
// under 2^-83, where 6q^3 - 6q^2 + 1 rounds to 1 whatever q's low bits.
__device__ __forceinline__ float cubic_w(float r_norm, const Smooth& s)
{
    if (!s.fast) return cubic_w(r_norm, s.h);
    const float h = s.h;
    float res = 0.0f;
    float k = 8.0f / (kPi * (h * h * h));
    float q = div_nr_pos(r_norm, h, s.rh);
    if (q <= 1.0f) {
        if (q <= 0.5f) {
            float q2 = q * q;
            float q3 = q2 * q;
            res = k * (6.0f * q3 - 6.0f * q2 + 1.0f);
        } else {
            float t = 1.0f - q;
            res = 2.0f * k * (t * t * t);
        }
    }
    return res;
}

// |r| and cubic_grad_r(r, |r|, h): the shared-reciprocal forms when r and h
// are in range, the IEEE forms otherwise.  Bit-identical to
// (vlen(r), cubic_grad(r, h)).
struct LenGrad {
    float len;
    v3 grad;
};
__device__ __forceinline__ LenGrad cubic_len_grad(v3 r, const Smooth& s)
{
    LenGrad o;
    if (s.fast && sep_in_range(r)) {
        const float h = s.h;
        o.len = sqrt_nr(vdot(r, r));
        o.grad = V(0.0f, 0.0f, 0.0f);
        float k = 8.0f / (kPi * (h * h * h));
        float q = div_nr_pos(o.len, h, s.rh);
        if (o.len > 1.0e-5f && q <= 1.0f) {
            const float D = o.len * h;
            const float rD = rcp_nr(D);
            v3 grad_q = V(div_nr(r.x, D, rD), div_nr(r.y, D, rD), div_nr(r.z, D, rD));
            if (q <= 0.5f) {
                o.grad = smul(6.0f * k * q * (3.0f * q - 2.0f), grad_q);
            } else {
                float q_term = 1.0f - q;
                o.grad = smul(-6.0f * k * q_term * q_term, grad_q);
            }
        }
    } else {
        o.len = vlen(r);
        o.grad = cubic_grad_r(r, o.len, s.h);
    }
    return o;
}

// H(r) u with H = d(cubic_grad)/dr on cubic_grad's branches (zero where it is
// zero): cubic_grad = g(d) r, H u = g u + (g'(d)/d)(r.u) r, with
//   q <= 0.5: g = 6k q(3q-2)/(d h),  g' = 18k/h^3
//   q >  0.5: g = -6k(1-q)^2/(d h),  g' = 6k/h (2(1-q)/(h d) + (1-q)^2/d^2)
__device__ __forceinline__ v3 cubic_hess_mul(v3 r, float r_norm, float h, v3 u)
{
    float k = 8.0f / (kPi * (h * h * h));
    float q = r_norm / h;
    if (!(r_norm > 1.0e-5f && q <= 1.0f)) return V(0.0f, 0.0f, 0.0f);
    float g, gpd;
    if (q <= 0.5f) {
        g = 6.0f * k * q * (3.0f * q - 2.0f) / (r_norm * h);
        gpd = 18.0f * k / (h * h * h) / r_norm;
    } else {
        float t = 1.0f - q;
        g = -6.0f * k * t * t / (r_norm * h);
        gpd = 6.0f * k / h * (2.0f * t / (h * r_norm) + t * t / (r_norm * r_norm)) / r_norm;
    }
    return vadd(smul(g, u), smul(gpd * vdot(r, u), r));
}

// PBF get_cubic = poly6  solver_pbf.py:23-32
__device__ __forceinline__ float poly6_w(float r_norm, float h)
{
    float res = 0.0f;
    if (h > 0.0f && r_norm < h) {
        float h2 = h * h;
        float rhs = h2 - r_norm * r_norm;
        float h3 = h * h * h;
        float k = 315.0f / (64.0f * kPi * (h3 * h3 * h3));
        res = k * (rhs * rhs * rhs);
    }
    return res;
}

// PBF get_cubic_derivative = spiky gradient  solver_pbf.py:35-45
__device__ __forceinline__ v3 spiky_grad_r(v3 r, float r_norm, float h)
{
    v3 res = V(0.0f, 0.0f, 0.0f);
    if (h > 0.0f && r_norm > 1.0e-6f && r_norm < h) {
        float h3 = h * h * h;
        float h6 = h3 * h3;
        float coeff = -45.0f / (kPi * h6);
        float term = (h - r_norm) * (h - r_norm) / r_norm;
        res = smul(coeff * term, r);
    }
    return res;
}
__device__ __forceinline__ v3 spiky_grad(v3 r, float h) { return spiky_grad_r(r, vlen(r), h); }

// cal_acc_with_non_pressure  solver_wcsph.py:63-89 (= solver_dfsph.py:67-93)
__device__ __forceinline__ v3 acc_non_pressure(v3 a, float rho_nei, v3 vel, v3 vel_nei, float mass, float mass_nei,
                                               float mu, float gamma, v3 d, float e_dist, float distance, float h)
{
    a = vsub(a, muls(smul(gamma / mass * mass_nei, d), cubic_w(e_dist, h)));
    float v_current_nei = vdot(vsub(vel, vel_nei), d);
    float s = 10.0f * mu * (mass_nei / fmax_(rho_nei, 1.0e-6f)) * v_current_nei /
              (distance * distance + 0.01f * (h * h));
    return vadd(a, smul(s, cubic_grad(d, h)));
}

// PBF cal_acc_with_non_pressure  solver_pbf.py:58-87
__device__ __forceinline__ v3 pbf_acc_non_pressure(v3 a, float rho_nei, v3 vel, v3 vel_nei, float mass,
                                                   float mass_nei, float mu, float gamma, v3 d, float e_dist,
                                                   float distance, float h)
{
    if (gamma != 0.0f) a = vsub(a, muls(smul(gamma / mass * mass_nei, d), poly6_w(e_dist, h)));
    if (mu != 0.0f) {
        float v_current_nei = vdot(vsub(vel, vel_nei), d);
        if (v_current_nei != 0.0f) {
            float s = 10.0f * mu * (mass_nei / fmax_(rho_nei, 1.0e-6f)) * v_current_nei /
                      (distance * distance + 0.01f * (h * h));
            a = vadd(a, smul(s, spiky_grad(d, h)));
        }
    }
    return a;
}

// cal_acc_with_pressure  solver_wcsph.py:92-106
__device__ __forceinline__ v3 acc_pressure(v3 a, v3 d, float p, float p_nei, float rho, float rho_nei,
                                           float mass_nei, float h)
{
    float dp_i = p / fmax_(rho * rho, 1.0e-6f);
    float dp_nei = p_nei / fmax_(rho_nei * rho_nei, 1.0e-6f);
    return vadd(a, smul(-mass_nei * (dp_i + dp_nei), cubic_grad(d, h)));
}

// compute_scorr  solver_pbf.py:90-101
__device__ __forceinline__ float compute_scorr(float r_norm, float h, float k, float n, float w_delta_q)
{
    if (k == 0.0f || w_delta_q <= 1.0e-6f) return 0.0f;
    float ratio = poly6_w(r_norm, h) / w_delta_q;
    return -k * fpow_(ratio, n);
}

// ---------------------------------------------------------------------------
// hash grid
// ---------------------------------------------------------------------------

__host__ __device__ __forceinline__ int cell_coord(float v, int round_mode)
{
    float f = round_mode ? truncf(v) : floorf(v);
    if (f > 1.0e9f) f = 1.0e9f;
    if (f < -1.0e9f) f = -1.0e9f;
    if (!(f == f)) f = 0.0f;
    return (int)f;
}

__host__ __device__ __forceinline__ int hash_axis(int c, int dim)
{
    c += kOrigin;
    if (c < 0) c = 0;
    return c % dim;
}

struct GridView {
    int dx, dy, dz, dxy;
    int round_mode;
    int npts;                       // points in the build (slot ranges are clamped to it)
    int base;                       // first bucket of this (scene's) table
    float inv;
    const int* __restrict__ start;  // [buckets + 1]
};

__device__ __forceinline__ int bucket_of(const GridView& g, float x, float y, float z)
{
    int cx = cell_coord(x * g.inv, g.round_mode);
    int cy = cell_coord(y * g.inv, g.round_mode);
    int cz = cell_coord(z * g.inv, g.round_mode);
    return g.base + hash_axis(cz, g.dz) * g.dxy + hash_axis(cy, g.dy) * g.dx + hash_axis(cx, g.dx);
}

// Iterator over the sorted-slot ranges visited by hash_grid_query(p, r):
// rows (y,z) in Warp order; within a row the x run is one contiguous range
// of sorted slots, or two when the run wraps past the last bucket, and x
// cells below -kOrigin (which hash to bucket 0 like -kOrigin itself) add one
// range of bucket 0 each ahead of it.
// The hashed y/z of successive rows are stepped, not recomputed: moving from
// virtual coordinate c-1 to c, hash_axis adds one and wraps at the dimension,
// except that coordinates at or below -kOrigin all hash to 0.  The row bases
// are kept premultiplied (y by dx, z by dx*dy), so a row costs no integer
// division or multiplication.
__device__ __forceinline__ int next_axis_base(int base, int c, int stride, int limit)
{
    return c + kOrigin <= 0 ? 0 : (base + stride == limit ? 0 : base + stride);
}

struct QueryRows {
    int hx0, seg1, seg2;     // x run: [hx0, hx0+seg1) then [0, seg2)
    int xrep;                // one-bucket parts of bucket 0 ahead of the run (x cells below -kOrigin)
    int ys, ny, zs, nz;      // virtual y/z start and extent
    int yi, zi, part;        // current row
    int yb0, yb, zb;         // hash_axis(ys, dy) * dx, and the current row's y and z bases
    int k, kend;             // current slot range

    __device__ __forceinline__ void begin(const GridView& g, float qx, float qy, float qz, float r)
    {
        int xs = cell_coord((qx - r) * g.inv, g.round_mode);
        int ysv = cell_coord((qy - r) * g.inv, g.round_mode);
        int zsv = cell_coord((qz - r) * g.inv, g.round_mode);
        int xe = cell_coord((qx + r) * g.inv, g.round_mode);
        int ye = cell_coord((qy + r) * g.inv, g.round_mode);
        int ze = cell_coord((qz + r) * g.inv, g.round_mode);
        xe = min(xe, xs + g.dx - 1);
        ye = min(ye, ysv + g.dy - 1);
        ze = min(ze, zsv + g.dz - 1);
        int lenx = xe - xs + 1;
        ny = ye - ysv + 1;
        nz = ze - zsv + 1;
        ys = ysv;
        zs = zsv;
        // x cells at or below -kOrigin all hash to bucket 0, which the query
        // then visits once per such cell; all but the last of them become
        // one-bucket parts ahead of the run, which starts at bucket 0
        xrep = 0;
        if (lenx > 0 && xs < -kOrigin) {
            xrep = min(xe, -kOrigin) - xs;
            xs += xrep;
            lenx -= xrep;
        }
        hx0 = hash_axis(xs, g.dx);
        if (lenx <= 0 || ny <= 0 || nz <= 0) {
            nz = 0;
            ny = 1;
            lenx = 0;
        }
        seg1 = min(lenx, g.dx - hx0);
        seg2 = lenx - seg1;
        yi = 0;
        zi = 0;
        part = 0;
        k = 0;
        kend = 0;
        yb0 = hash_axis(ys, g.dy) * g.dx;
        yb = yb0;
        zb = hash_axis(zs, g.dz) * g.dxy;
        if (nz > 0) load_row(g);
    }
    __device__ __forceinline__ void load_row(const GridView& g)
    {
        int a = g.base + zb + yb + (part == xrep ? hx0 : 0);
        int b = a + (part < xrep ? 1 : (part == xrep ? seg1 : seg2));
        k = min(max(g.start[a], 0), g.npts);
        kend = min(max(g.start[b], k), g.npts);
    }
    // Move to the next row; false when the query is exhausted.
    __device__ __forceinline__ bool advance(const GridView& g)
    {
        if (part < xrep || (part == xrep && seg2 > 0)) {
            ++part;
        } else {
            part = 0;
            if (++yi == ny) {
                yi = 0;
                yb = yb0;
                if (++zi == nz) return false;
                zb = next_axis_base(zb, zs + zi, g.dxy, g.dxy * g.dz);
            } else {
                yb = next_axis_base(yb, ys + yi, g.dx, g.dxy);
            }
        }
        load_row(g);
        return true;
    }
    __device__ __forceinline__ bool valid() const { return nz > 0; }
    // Advance past empty rows; false when the query is exhausted.
    __device__ __forceinline__ bool skip_empty(const GridView& g)
    {
        while (k >= kend)
            if (!advance(g)) return false;
        return true;
    }
};

// Per-particle neighbour list: the active within-cut candidates of a query in
// visit order, entry e of sorted slot s at idx[e * n + s] (coalesced across
// lanes).
// cnt[s] > cap marks an overflowed list; its users fall back to scanning.
struct NbrList {
    int* idx;
    int* cnt;
    int cap;
    int n;
    int mode;      // kListNone: scan only; kListBuild: scan and record; kListUse: replay
    float cut2;    // squared record radius (>= the pass cut; larger for a skin)
    int* need;     // optional: atomic max of the lengths of lists that overflowed
};

// Length of slot s's recorded list (cap + 1 marks an overflowed list; the
// length it needed then goes to L.need, which the host reads to grow the
// lists).
__device__ __forceinline__ void store_list_count(const NbrList& L, int s, int written)
{
    if (written <= L.cap) {
        L.cnt[s] = written;
    } else {
        L.cnt[s] = L.cap + 1;
        if (L.need) atomicMax(L.need, written);
    }
}
constexpr int kListNone = 0;
constexpr int kListBuild = 1;
constexpr int kListUse = 2;
constexpr int kBatch = 8;

struct NoCount {
    __device__ __forceinline__ void operator()(int) const {}
};

// gfx9 buffer resource over `count` float4s: 32-bit offsets, loads past the
// end return zeros.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t float4_rsrc(const float4* p, int count)
{
    // num_records in bytes, unsigned 32-bit: the hosts cap particle counts
    // at kMaxParticles so that it and every byte offset fit
    return __builtin_amdgcn_make_buffer_rsrc((void*)p, (short)0, (int)((unsigned)count * 16u), 0x00020000);
}
__device__ __forceinline__ float4 ld_f4(__amdgpu_buffer_rsrc_t r, int base_bytes, int imm)
{
    return __builtin_bit_cast(float4, __builtin_amdgcn_raw_buffer_load_b128(r, base_bytes, imm, 0));
}
__device__ __forceinline__ float ld_f32(__amdgpu_buffer_rsrc_t r, int base_bytes, int imm)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, base_bytes, imm, 0));
}
// Byte offset of slot k in a sorted float4 stream (one 32-bit offset serves
// every stream a neighbour body reads; slot counts are capped so it fits).
__device__ __forceinline__ int slot_off(int k) { return (int)((unsigned)k * 16u); }

// Squared distance for the survivor prefilter only: fused multiply-adds (3
// instructions instead of 5).  It differs from the reference's written-order
// value by a few ulps, far inside the relative slack of the cuts it is
// compared with (1e-4), so it never drops a candidate the exact test inside
// the pass bodies would keep; the bodies still decide with the exact value.
__device__ __forceinline__ float prefilter_d2(float dx, float dy, float dz)
{
    return __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx));
}

// A two-phase neighbour body: load(k) fetches what the body needs of
// neighbour k (unconditionally, no control flow), body(k, data) uses it.  The
// loops fetch the data of several neighbours before running their bodies, so
// those loads are in flight together instead of one round trip per
// neighbour; the bodies still run one after another in visit order, so every
// sum is formed in the same order as with a plain body(k).
template <class Load, class Body>
struct Phased {
    Load load;
    Body body;
};
template <class Load, class Body>
__device__ __forceinline__ Phased<Load, Body> phased(Load load, Body body)
{
    return Phased<Load, Body>{load, body};
}
template <class T>
struct IsPhased {
    static constexpr bool value = false;
};
template <class Load, class Body>
struct IsPhased<Phased<Load, Body>> {
    static constexpr bool value = true;
};
template <class Body>
__device__ __forceinline__ void run_body(Body& body, int k)
{
    if constexpr (IsPhased<typename std::decay<Body>::type>::value) body.body(k, body.load(k));
    else body(k);
}
// Data of a neighbour taken from a recorded list.  Lists hold active
// neighbours only (the scanning loop records a survivor after its flags are
// loaded), so the flags are known and their load is dropped; every body
// tests no flag of a neighbour but kActive.
template <class Body>
__device__ __forceinline__ auto load_listed(Body& body, int k)
{
    auto d = body.load(k);
    d.flags = kActive;
    return d;
}
template <class Body>
__device__ __forceinline__ void run_listed(Body& body, int k)
{
    if constexpr (IsPhased<typename std::decay<Body>::type>::value) body.body(k, load_listed(body, k));
    else body(k);
}
constexpr int kPhasedUnroll = 4;   // neighbours whose data is fetched together
constexpr int kReplayUnroll = 4;   // ... in the list-replaying loop

// Visit the candidates of hash_grid_query(p, r) in Warp order (returning how
// many there were), but call
// `body(k)` (k = sorted slot) only for those with squared distance not above
// `cut2` -- every skipped candidate has q > 1 and contributes exactly zero to
// every kernel sum here.  Candidate positions are loaded BATCH at a time
// (kBatch unless the pass sets it: short rows waste less with small batches);
// survivors are compacted through a per-lane LDS list of CAP entries so that
// the expensive body runs on full waves.  `count(k)` is called for every
// candidate (DFSPH's unfiltered neighbour count).  With a kListBuild list the
// active survivors (radius^2 <= L.cut2) are also recorded for later passes:
// every pass skips inactive neighbours, so leaving them out of the lists
// changes no result.
//
// The primary form loads through a buffer resource (32-bit offsets, an
// immediate offset per candidate; slots past a row's end are read and
// ignored, slots past the array read as zeros) and compacts branch-free.
// LEAN is the fallback of the list-replaying passes: global loads clamped to
// the row and branchy compaction, which keep that rarely taken path from
// raising the register count (and lowering the occupancy) of its kernel.
template <int BLOCK, int CAP, bool LEAN = false, int BATCH = kBatch, class Count, class Body>
__device__ __forceinline__ int scan_neighbors(const GridView& g, const float4* __restrict__ pos, int s, float px,
                                               float py, float pz, float r, float cut2, int* __restrict__ lds,
                                               const NbrList& L, Count&& count, Body&& body)
{
    static_assert(CAP >= 2 * BATCH, "list capacity");
    // lists hold active neighbours only, which needs the flags of the phased
    // payload at record time (the lean scan never records)
    static_assert(LEAN || IsPhased<typename std::decay<Body>::type>::value, "scanning bodies must be phased");
    constexpr bool kCounts = !__is_same(typename std::decay<Count>::type, NoCount);
    const bool record = L.mode == kListBuild;
    const float keep2 = record ? fmax_(L.cut2, cut2) : cut2;
    QueryRows it;
    it.begin(g, px, py, pz, r);
    int written = 0;
    int candidates = 0;
    // entry `written` of slot s; stepped by n so that cap * n may pass 2^31
    int* __restrict__ out = L.idx + s;
    // Every trip of the batch loop below tests one batch: the current range is
    // kept non-empty, and a lane whose batch ends its row moves to the next
    // row in the same trip (the bucket-range loads of that row are issued
    // before the batch's tests, so both round trips overlap).
    if (it.valid() && it.skip_empty(g)) {
        int* list = lds + threadIdx.x;
        bool done = false;
        while (true) {
            int cnt = 0;
            while (!done && cnt <= CAP - BATCH) {
                const int k0 = it.k;
                const int rem = it.kend - k0;
                const int take = rem < BATCH ? rem : BATCH;
                float4 q[BATCH];
                const bool row_end = take == rem;
                if constexpr (LEAN) {
#pragma unroll
                    for (int u = 0; u < BATCH; ++u) q[u] = pos[k0 + (u < take ? u : take - 1)];
                    it.k = k0 + take;
                    if (row_end && !it.advance(g)) done = true;
#pragma unroll
                    for (int u = 0; u < BATCH; ++u) {
                        if (u < take) {
                            count(k0 + u);
                            float ddx = px - q[u].x, ddy = py - q[u].y, ddz = pz - q[u].z;
                            float d2 = prefilter_d2(ddx, ddy, ddz);
                            if (!(d2 > keep2)) list[(cnt++) * BLOCK] = k0 + u;
                        }
                    }
                } else {
                    const __amdgpu_buffer_rsrc_t rs = float4_rsrc(pos, g.npts);
                    const int off = (int)((unsigned)k0 * 16u);
#pragma unroll
                    for (int u = 0; u < BATCH; ++u) q[u] = ld_f4(rs, off, u * 16);
                    it.k = k0 + take;
                    if (row_end && !it.advance(g)) done = true;
#pragma unroll
                    for (int u = 0; u < BATCH; ++u) {
                        if (kCounts && u < take) count(k0 + u);
                        float ddx = px - q[u].x, ddy = py - q[u].y, ddz = pz - q[u].z;
                        float d2 = prefilter_d2(ddx, ddy, ddz);
                        list[cnt * BLOCK] = k0 + u;
                        cnt += (u < take && !(d2 > keep2)) ? 1 : 0;
                    }
                }
                candidates += take;
                if (row_end && !done && !it.skip_empty(g)) done = true;
            }
            auto rec = [&](int k) {
                if (written < L.cap) {
                    *out = k;
                    out += L.n;
                }
                ++written;
            };
            int e = 0;
            if constexpr (IsPhased<typename std::decay<Body>::type>::value) {
                for (; e + kPhasedUnroll <= cnt; e += kPhasedUnroll) {
                    int k[kPhasedUnroll];
#pragma unroll
                    for (int u = 0; u < kPhasedUnroll; ++u) k[u] = list[(e + u) * BLOCK];
                    decltype(body.load(0)) d[kPhasedUnroll];
#pragma unroll
                    for (int u = 0; u < kPhasedUnroll; ++u) d[u] = body.load(k[u]);
                    if (record) {
#pragma unroll
                        for (int u = 0; u < kPhasedUnroll; ++u)
                            if (d[u].flags & kActive) rec(k[u]);
                    }
#pragma unroll
                    for (int u = 0; u < kPhasedUnroll; ++u) body.body(k[u], d[u]);
                }
            }
            for (; e < cnt; ++e) {
                const int k = list[e * BLOCK];
                if constexpr (IsPhased<typename std::decay<Body>::type>::value) {
                    const auto d = body.load(k);
                    if (record && (d.flags & kActive)) rec(k);
                    body.body(k, d);
                } else {
                    run_body(body, k);
                }
            }
            if (done) break;
        }
    }
    if (record) store_list_count(L, s, written);
    return candidates;
}

// Neighbour loop of one pass.  A replaying pass (REPLAY) walks the list the
// scanning pass recorded at the same positions, and scans only when there is
// no list or it overflowed (the lean scan); a scanning pass always scans with
// the primary loop and records when its list is in build mode.  The role is
// a template parameter so that each kernel carries only the loop it runs.
template <int BLOCK, int CAP, bool REPLAY, int BATCH = kBatch, int UNR = kReplayUnroll, class Count, class Body>
__device__ __forceinline__ int for_each_neighbor(const GridView& g, const float4* __restrict__ pos, int s, float px,
                                                  float py, float pz, float r, float cut2, int* __restrict__ lds,
                                                  const NbrList& L, Count&& count, Body&& body)
{
    if constexpr (REPLAY) {
        if (L.mode == kListUse) {
            const int m = L.cnt[s];
            if (m <= L.cap) {
                // pointer stepping keeps the addressing 64-bit (cap * n may
                // pass 2^31)
                const int* __restrict__ idx = L.idx + s;
                const ptrdiff_t n = L.n;
                int e = 0;
                if constexpr (IsPhased<typename std::decay<Body>::type>::value) {
                    // software-pipelined by one group: the next group's
                    // indices are issued before this group's neighbour data,
                    // so a group costs one round trip instead of two (long
                    // lists are latency chains)
                    constexpr int U = UNR;   // neighbours whose data is fetched together
                    if (m >= U) {
                        int k[U];
#pragma unroll
                        for (int u = 0; u < U; ++u) k[u] = idx[u * n];
                        idx += U * n;
                        for (e = U;; e += U) {
                            const bool more = e + U <= m;
                            int nk[U];
#pragma unroll
                            for (int u = 0; u < U; ++u) nk[u] = more ? idx[u * n] : 0;
                            if (more) idx += U * n;
                            decltype(load_listed(body, 0)) d[U];
#pragma unroll
                            for (int u = 0; u < U; ++u) d[u] = load_listed(body, k[u]);
#pragma unroll
                            for (int u = 0; u < U; ++u) body.body(k[u], d[u]);
                            if (!more) break;
#pragma unroll
                            for (int u = 0; u < U; ++u) k[u] = nk[u];
                        }
                    }
                    // the last partial group in one round trip: slots past the
                    // list's end fetch the particle's own data and are skipped
                    const int r = m - e;
                    if (r > 0) {
                        int k[U];
#pragma unroll
                        for (int u = 0; u < U; ++u) k[u] = u < r ? idx[u * n] : s;
                        decltype(load_listed(body, 0)) d[U];
#pragma unroll
                        for (int u = 0; u < U; ++u) d[u] = load_listed(body, k[u]);
#pragma unroll
                        for (int u = 0; u < U; ++u)
                            if (u < r) body.body(k[u], d[u]);
                        e = m;
                    }
                } else {
                    for (; e + 4 <= m; e += 4) {
                        const int k0 = idx[0], k1 = idx[n], k2 = idx[2 * n], k3 = idx[3 * n];
                        idx += 4 * n;
                        body(k0);
                        body(k1);
                        body(k2);
                        body(k3);
                    }
                }
                for (; e < m; ++e, idx += n) run_listed(body, *idx);
                return 0;
            }
        }
        NbrList none = L;
        none.mode = kListNone;
        return scan_neighbors<BLOCK, CAP, true, BATCH>(g, pos, s, px, py, pz, r, cut2, lds, none, count, body);
    } else {
        NbrList scan = L;
        if (scan.mode == kListUse) scan.mode = kListNone;
        return scan_neighbors<BLOCK, CAP, false, BATCH>(g, pos, s, px, py, pz, r, cut2, lds, scan, count, body);
    }
}

// Word W of v as held by lane W of this lane's group of LPP adjacent lanes
// (DPP quad permutation: no LDS, usually folded into the consuming
// instruction).  Every lane of the group must be active.
template <int LPP, int W, class T>
__device__ __forceinline__ T group_bcast(const T& v)
{
    if constexpr (LPP == 1) {
        return v;
    } else {
        static_assert(LPP == 2 || LPP == 4, "lanes per particle");
        static_assert(sizeof(T) % 4 == 0, "word-sized fields");
        constexpr int ctrl = LPP == 4 ? W * 0x55 : (W == 0 ? 0xA0 : 0xF5);   // [W,W,W,W] / [0,0,2,2] / [1,1,3,3]
        struct Words {
            int w[sizeof(T) / 4];
        };
        Words in = __builtin_bit_cast(Words, v), out;
#pragma unroll
        for (int i = 0; i < (int)(sizeof(T) / 4); ++i) out.w[i] = __builtin_amdgcn_mov_dpp(in.w[i], ctrl, 0xF, 0xF, true);
        return __builtin_bit_cast(T, out);
    }
}

// List replay with LPP lanes per particle.  Lane q of the particle's group
// evaluates term(k, data) for the entries e = q, q + LPP, ... (U of them per
// trip, their indices fetched one trip ahead and their data together); the
// group then hands the terms to fold() in entry order, lane by lane, so the
// particle's sums are formed in the same order as by one lane walking the
// list.  Every lane of the group folds every term and ends with the same sums.
// The data loads are load_listed's: lists hold active neighbours only.
template <int LPP, int U, class Load, class Term, class Fold>
__device__ __forceinline__ void replay_split(const NbrList& L, int s, int m, Load& load, Term& term, Fold& fold)
{
    const int q = (int)(threadIdx.x & (LPP - 1));
    const int* __restrict__ base = L.idx + s;
    const ptrdiff_t n = L.n;
    constexpr int kStep = LPP * U;   // entries per trip of the group
    int k[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const int e = u * LPP + q;
        k[u] = e < m ? base[(ptrdiff_t)e * n] : s;
    }
    for (int e0 = 0; e0 < m; e0 += kStep) {
        int nk[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int e = e0 + kStep + u * LPP + q;
            nk[u] = e < m ? base[(ptrdiff_t)e * n] : s;
        }
        decltype(load(0)) d[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            d[u] = load(k[u]);
            d[u].flags = kActive;
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const auto t = term(k[u], d[u]);
            if (e0 + u * LPP < m) fold(group_bcast<LPP, 0>(t));
            if constexpr (LPP >= 2)
                if (e0 + u * LPP + 1 < m) fold(group_bcast<LPP, 1>(t));
            if constexpr (LPP == 4) {
                if (e0 + u * LPP + 2 < m) fold(group_bcast<LPP, 2>(t));
                if (e0 + u * LPP + 3 < m) fold(group_bcast<LPP, 3>(t));
            }
        }
#pragma unroll
        for (int u = 0; u < U; ++u) k[u] = nk[u];
    }
}

// Neighbour loop of a list-replaying pass split into term (per neighbour,
// order-free) and fold (accumulation in visit order), run by LPP lanes per
// particle.  Without a usable list (no list, or an overflowed one) the
// group's first lane scans alone with body fold(term(k, data)); then only
// that lane's sums are the particle's.
template <int BLOCK, int CAP, int UNR, int LPP, class Load, class Term, class Fold>
__device__ __forceinline__ void for_each_neighbor_split(const GridView& g, const float4* __restrict__ pos, int s,
                                                        float px, float py, float pz, float r, float cut2,
                                                        int* __restrict__ lds, const NbrList& L, Load load, Term term,
                                                        Fold fold)
{
    if (L.mode == kListUse) {
        const int m = L.cnt[s];
        if (m <= L.cap) {
            replay_split<LPP, UNR>(L, s, m, load, term, fold);
            return;
        }
    }
    if ((threadIdx.x & (LPP - 1)) == 0) {
        NbrList none = L;
        none.mode = kListNone;
        scan_neighbors<BLOCK, CAP, true, kBatch>(
            g, pos, s, px, py, pz, r, cut2, lds, none, NoCount{},
            phased(load, [&](int k, const decltype(load(0))& d) { fold(term(k, d)); }));
    }
}

// Sum of an integer over the lane's group of LPP lanes (every lane of the
// group active).
template <int LPP>
__device__ __forceinline__ int group_sum(int x)
{
    if constexpr (LPP == 1) {
        return x;
    } else if constexpr (LPP == 2) {
        return group_bcast<2, 0>(x) + group_bcast<2, 1>(x);
    } else {
        return (group_bcast<4, 0>(x) + group_bcast<4, 1>(x)) + (group_bcast<4, 2>(x) + group_bcast<4, 3>(x));
    }
}

// Exclusive prefix (over the group's lower lanes) and total of an integer
// over the lane's group of LPP lanes (every lane of the group active).
template <int LPP>
__device__ __forceinline__ int group_prefix(int x, int& total)
{
    const int q = (int)(threadIdx.x & (LPP - 1));
    if constexpr (LPP == 2) {
        const int x0 = group_bcast<2, 0>(x), x1 = group_bcast<2, 1>(x);
        total = x0 + x1;
        return q == 1 ? x0 : 0;
    } else {
        const int x0 = group_bcast<4, 0>(x), x1 = group_bcast<4, 1>(x);
        const int x2 = group_bcast<4, 2>(x), x3 = group_bcast<4, 3>(x);
        total = (x0 + x1) + (x2 + x3);
        return (q > 0 ? x0 : 0) + (q > 1 ? x1 : 0) + (q > 2 ? x2 : 0);
    }
}

// Scanning loop with LPP lanes per particle (LPP = 2 or 4).  The group walks
// the query's rows together (every lane keeps the same row state); each trip
// takes the next LPP * BL candidates of the current row, lane q testing
// candidates k0 + q BL .. k0 + q BL + BL - 1, so lane order is visit order.
// Survivors are compacted, in that order, into the group's LDS list (entry e
// at lds[e * BLOCK / LPP + group]) at offsets from a prefix over the group.  The list is then
// flushed as replay_split does a recorded list: lane q evaluates term() for
// entries q, q + LPP, ... (U at a time) and the group folds the terms in
// entry order; active survivors are recorded in order when L is in build
// mode.  count(k) runs on the lane that tests candidate k (callers sum it
// over the group).  Returns the number of candidates (the same on every lane
// of the group).
template <int BLOCK, int CAP, int LPP, int BL, int U, class Count, class Load, class Term, class Fold>
__device__ __forceinline__ int scan_split(const GridView& g, const float4* __restrict__ pos, int s, float px,
                                          float py, float pz, float r, float cut2, int* __restrict__ lds,
                                          const NbrList& L, Count&& count, Load& load, Term& term, Fold& fold)
{
    static_assert(LPP == 2 || LPP == 4, "lanes per particle");
    constexpr int kGroups = BLOCK / LPP;
    constexpr int kTrip = LPP * BL;   // candidates per trip of the group
    static_assert(CAP >= 2 * kTrip, "list capacity");
    constexpr bool kCounts = !__is_same(typename std::decay<Count>::type, NoCount);
    const int q = (int)(threadIdx.x & (LPP - 1));
    int* __restrict__ list = lds + threadIdx.x / LPP;
    const bool record = L.mode == kListBuild;
    const float keep2 = record ? fmax_(L.cut2, cut2) : cut2;
    const ptrdiff_t n = L.n;
    QueryRows it;
    it.begin(g, px, py, pz, r);
    int written = 0;
    int candidates = 0;
    if (it.valid() && it.skip_empty(g)) {
        const __amdgpu_buffer_rsrc_t rs = float4_rsrc(pos, g.npts);
        bool done = false;
        while (true) {
            int cnt = 0;
            while (!done && cnt <= CAP - kTrip) {
                const int k0 = it.k;
                const int rem = it.kend - k0;
                const int take = rem < kTrip ? rem : kTrip;
                const bool row_end = take == rem;
                const int kq = k0 + q * BL;
                float4 c[BL];
                const int off = (int)((unsigned)kq * 16u);
#pragma unroll
                for (int u = 0; u < BL; ++u) c[u] = ld_f4(rs, off, u * 16);
                it.k = k0 + take;
                if (row_end && !it.advance(g)) done = true;
                bool hit[BL];
                int mine = 0;
#pragma unroll
                for (int u = 0; u < BL; ++u) {
                    const bool in = q * BL + u < take;
                    if (kCounts && in) count(kq + u);
                    float ddx = px - c[u].x, ddy = py - c[u].y, ddz = pz - c[u].z;
                    hit[u] = in && !(prefilter_d2(ddx, ddy, ddz) > keep2);
                    mine += hit[u] ? 1 : 0;
                }
                // survivors of the group's lower lanes come first
                int total;
                int at = cnt + group_prefix<LPP>(mine, total);
#pragma unroll
                for (int u = 0; u < BL; ++u) {
                    if (hit[u]) list[at * kGroups] = kq + u;
                    at += hit[u] ? 1 : 0;
                }
                cnt += total;
                candidates += take;
                if (row_end && !done && !it.skip_empty(g)) done = true;
            }
            // the group's list entries are written by its own lanes (one
            // wave); keep the reads below after those writes
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            for (int e0 = 0; e0 < cnt; e0 += LPP * U) {
                int k[U];
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    const int e = e0 + u * LPP + q;
                    k[u] = e < cnt ? list[e * kGroups] : s;
                }
                decltype(load(0)) d[U];
#pragma unroll
                for (int u = 0; u < U; ++u) d[u] = load(k[u]);
                if (record) {
#pragma unroll
                    for (int u = 0; u < U; ++u) {
                        const int act = (e0 + u * LPP + q < cnt && (d[u].flags & kActive)) ? 1 : 0;
                        int nact;
                        const int p = written + group_prefix<LPP>(act, nact);
                        if (act && p < L.cap) L.idx[s + (ptrdiff_t)p * n] = k[u];
                        written += nact;
                    }
                }
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    const auto t = term(k[u], d[u]);
                    if (e0 + u * LPP < cnt) fold(group_bcast<LPP, 0>(t));
                    if (e0 + u * LPP + 1 < cnt) fold(group_bcast<LPP, 1>(t));
                    if constexpr (LPP == 4) {
                        if (e0 + u * LPP + 2 < cnt) fold(group_bcast<LPP, 2>(t));
                        if (e0 + u * LPP + 3 < cnt) fold(group_bcast<LPP, 3>(t));
                    }
                }
            }
            // the next trip's writes may reuse entries read above
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            if (done) break;
        }
    }
    if (record && q == 0) store_list_count(L, s, written);
    return candidates;
}

// Block index -> tile index so that each of the 8 XCDs (which receive
// workgroups round-robin) processes one contiguous run of tiles.
__device__ __forceinline__ int xcd_tile()
{
    int b = blockIdx.x, nb = gridDim.x;
    int x = b & 7, q = b >> 3;
    int per = nb >> 3, rem = nb & 7;
    return x * per + min(x, rem) + q;
}

}  // namespace fs
