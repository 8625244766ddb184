// fs_fastmath.hip -- on-device comparison of the shared-reciprocal quotient
// and square-root forms of fs_common.h (div_nr, sqrt_nr, vlen_nr,
// cubic_len_grad, cubic_w on a Smooth) with the correctly rounded forms hipcc emits for '/'
// and sqrtf.  A diagnostic: the neighbour passes use the fast forms only on
// the operand ranges this kernel covers.
#include "fs_internal.h"

namespace fs {
namespace {

// splitmix64 finaliser: 64 well-mixed bits from a counter
__device__ __forceinline__ uint64_t mix64(uint64_t z)
{
    z += 0x9e3779b97f4a7c15ull;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

__device__ __forceinline__ float make_float(uint32_t sign, int exp2, uint32_t mant)
{
    return __uint_as_float((sign << 31) | ((uint32_t)(exp2 + 127) << 23) | (mant & 0x7fffffu));
}

__device__ __forceinline__ bool differ(float a, float b) { return __float_as_uint(a) != __float_as_uint(b); }

// one separation component for mode 3: +-0, a tiny value that takes the
// IEEE fallback, or a uniform value in [-1.2 h, 1.2 h]
__device__ __forceinline__ float component(uint64_t u, float h)
{
    const uint32_t kind = (uint32_t)(u & 15u);
    const uint32_t sign = (uint32_t)(u >> 4) & 1u;
    if (kind == 0) return sign ? -0.0f : 0.0f;
    if (kind == 1) return make_float(sign, -80 + (int)((u >> 5) % 40u), (uint32_t)(u >> 16));
    const float f = (float)((u >> 40) & 0xffffffu) * 0x1p-24f;   // [0, 1)
    return (2.0f * f - 1.0f) * (1.2f * h);
}

// mode 0: x = bits(first + t): sqrt_nr(x) vs sqrtf(x)
// mode 1: (x, d) from a counter hash, d in [2^-90, 2^91), |x| = 0 or in
//         [2^-100, 2^101) with |x / d| in about [2^-91, 2^91]: div_nr vs '/'
// mode 2: x = bits(first + t), d = bits(seed): div_nr vs '/'
// mode 3: a random separation r and smoothing length h (mostly in
//         [2^-20, 2^20), some outside): cubic_len_grad(r, smooth_of(h)) vs
//         (vlen(r), cubic_grad(r, h)), cubic_w on a Smooth vs on h, and
//         vlen_nr(r) vs vlen(r)
__global__ __launch_bounds__(kElemBlock) void k_check_fast_math(int mode, uint64_t first, uint64_t count,
                                                                uint32_t seed, unsigned long long* out)
{
    const uint64_t stride = (uint64_t)gridDim.x * kElemBlock;
    for (uint64_t t = (uint64_t)blockIdx.x * kElemBlock + threadIdx.x; t < count; t += stride) {
        bool bad = false;
        if (mode == 0) {
            const float x = __uint_as_float((uint32_t)(first + t));
            bad = differ(sqrt_nr(x), sqrtf(x));
        } else if (mode == 1) {
            const uint64_t u = mix64(((uint64_t)seed << 40) ^ (first + t));
            const uint64_t w = mix64(u);
            const int ed = -90 + (int)(u % 181u);
            const int eq = -90 + (int)((u >> 8) % 181u);
            int ex = ed + eq;
            ex = ex < -100 ? -100 : (ex > 100 ? 100 : ex);
            const float d = make_float(0u, ed, (uint32_t)(u >> 20));
            float x = make_float((uint32_t)(w >> 63), ex, (uint32_t)w);
            if (((w >> 40) & 15u) == 0) x = ((w >> 44) & 1u) ? -0.0f : 0.0f;
            bad = differ(div_nr(x, d, rcp_nr(d)), x / d);
        } else if (mode == 2) {
            const float x = __uint_as_float((uint32_t)(first + t));
            const float d = __uint_as_float(seed);
            bad = differ(div_nr(x, d, rcp_nr(d)), x / d);
        } else {
            const uint64_t u = mix64(((uint64_t)seed << 40) ^ (first + t));
            const bool wide = (u & 15u) == 0;
            const float e = wide ? -30.0f + 60.0f * (float)((u >> 4) & 0xffffu) * 0x1p-16f
                                 : -20.0f + 40.0f * (float)((u >> 4) & 0xffffu) * 0x1p-16f;
            const float h = exp2f(e);
            const v3 r = V(component(mix64(u ^ 1u), h), component(mix64(u ^ 2u), h), component(mix64(u ^ 3u), h));
            const Smooth sm = smooth_of(h);
            const LenGrad lg = cubic_len_grad(r, sm);
            const float len = vlen(r);
            const v3 g = cubic_grad_r(r, len, h);
            bad = differ(lg.len, len) || differ(lg.grad.x, g.x) || differ(lg.grad.y, g.y) || differ(lg.grad.z, g.z) ||
                  differ(cubic_w(len, sm), cubic_w(len, h)) || differ(vlen_nr(r), len);
        }
        if (bad) {
            atomicAdd(&out[0], 1ull);
            atomicMin(&out[1], (unsigned long long)t);
        }
    }
}

__global__ void k_init_check(unsigned long long* out)
{
    out[0] = 0ull;
    out[1] = ~0ull;
}

}  // namespace
}  // namespace fs

using namespace fs;

extern "C" int32_t fs_check_fast_math(int32_t mode, uint64_t first, uint64_t count, uint32_t seed, uint64_t* result,
                                      fs_stream_t stream)
{
    if (mode < 0 || mode > 3 || !result) {
        set_error("fs_check_fast_math: invalid arguments");
        return FS_ERR_ARG;
    }
    hipStream_t s = (hipStream_t)stream;
    unsigned long long* d_out = nullptr;
    int32_t rc = hip_check(hipMalloc((void**)&d_out, 2 * sizeof(unsigned long long)), "fs_check_fast_math: hipMalloc");
    if (rc) return rc;
    unsigned long long h_out[2] = {0ull, 0ull};
    auto run = [&]() -> int32_t {
        FS_LAUNCH(k_init_check, dim3(1), dim3(1), 0, s, d_out);
        if (count > 0) {
            const uint64_t want = (count + kElemBlock - 1) / kElemBlock;
            const int blocks = (int)(want < 16384 ? want : 16384);
            FS_LAUNCH(k_check_fast_math, dim3(blocks), dim3(kElemBlock), 0, s, (int)mode, first, count, seed, d_out);
        }
        int32_t r = hip_check(hipMemcpyAsync(h_out, d_out, sizeof(h_out), hipMemcpyDeviceToHost, s),
                              "fs_check_fast_math: copy");
        return r ? r : hip_check(hipStreamSynchronize(s), "fs_check_fast_math: synchronize");
    };
    rc = run();
    hipFree(d_out);
    if (rc) return rc;
    result[0] = h_out[0];
    result[1] = h_out[1];
    return FS_OK;
}
