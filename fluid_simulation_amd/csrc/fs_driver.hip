// fs_driver.hip -- device reductions of the dam-break driver: CFL maxima and
// the non-finite state check (examples/dam_break/simulation.py:19-53).
#include "fs_internal.h"

#include <algorithm>
#include <mutex>

// Scratch of fs_batch_cfl_dt_noalloc: the per-scene maxima, (|v|^2, |a|^2) as
// float bits, for at most kNoAllocMaxScenes scenes.  Each device holds its own
// copy of this array until the process ends.  Host code never passes the
// name itself to a kernel: it takes the device address from
// hipGetSymbolAddress, once per device (noalloc_scratch below).
constexpr int kNoAllocMaxScenes = 4096;
static __device__ __align__(256) uint32_t scratch[2 * kNoAllocMaxScenes];

namespace fs {

constexpr int kRedBlock = 256;
constexpr int kRedUnroll = 4;   // points per thread per trip of the maxima reductions

__device__ __forceinline__ float norm2_or_inf(v3 u)
{
    float u2 = vdot(u, u);
    if (!(isfinite(u.x) && isfinite(u.y) && isfinite(u.z) && isfinite(u2))) u2 = INFINITY;
    return u2;
}

__global__ void k_fill_zero(uint32_t* p, int n)
{
    int i = threadIdx.x;
    if (i < n) p[i] = 0u;
}

// zero n words (any n), grid-stride; a kernel rather than hipMemsetAsync so
// that the clear also records correctly into a captured graph
__global__ __launch_bounds__(256) void k_fill_zero_n(uint32_t* p, int n)
{
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) p[i] = 0u;
}

// particle_norm_maxima  simulation.py:19-36.  Values are >= 0 (or +inf), so
// their float bits order like unsigned integers.
__global__ __launch_bounds__(kRedBlock) void k_norm_maxima(const float* __restrict__ v, int nv,
                                                           const float* __restrict__ a, int na, uint32_t* out)
{
    __shared__ float s_m[2][kRedBlock / 64];
    const int n = nv > na ? nv : na;
    float m0 = 0.0f, m1 = 0.0f;
    // kRedUnroll points per trip, their loads issued together (a zero vector
    // leaves a maximum unchanged; maxima do not depend on the order)
    const int stride = gridDim.x * kRedBlock;
    for (int i0 = blockIdx.x * kRedBlock + threadIdx.x; i0 < n; i0 += kRedUnroll * stride) {
        v3 vv[kRedUnroll], aa[kRedUnroll];
#pragma unroll
        for (int u = 0; u < kRedUnroll; ++u) {
            const int i = i0 + u * stride;
            vv[u] = i < nv ? ld3(v, i) : V(0.0f, 0.0f, 0.0f);
            aa[u] = i < na ? ld3(a, i) : V(0.0f, 0.0f, 0.0f);
        }
#pragma unroll
        for (int u = 0; u < kRedUnroll; ++u) {
            m0 = fmax_(m0, norm2_or_inf(vv[u]));
            m1 = fmax_(m1, norm2_or_inf(aa[u]));
        }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        m0 = fmax_(m0, __shfl_xor(m0, off, 64));
        m1 = fmax_(m1, __shfl_xor(m1, off, 64));
    }
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) {
        s_m[0][wid] = m0;
        s_m[1][wid] = m1;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < kRedBlock / 64; ++w) {
            m0 = fmax_(m0, s_m[0][w]);
            m1 = fmax_(m1, s_m[1][w]);
        }
        atomicMax(&out[0], __float_as_uint(m0));
        atomicMax(&out[1], __float_as_uint(m1));
    }
}

// check_particle_finite  simulation.py:39-53
__global__ __launch_bounds__(kRedBlock) void k_check_finite(const float* __restrict__ x, const float* __restrict__ v,
                                                            int n, const float* __restrict__ rho, int nd,
                                                            uint32_t* out)
{
    int bad = 0;
    for (int i = blockIdx.x * kRedBlock + threadIdx.x; i < n; i += gridDim.x * kRedBlock) {
        v3 p = ld3(x, i), u = ld3(v, i);
        bool b = !(isfinite(p.x) && isfinite(p.y) && isfinite(p.z) && isfinite(u.x) && isfinite(u.y) &&
                   isfinite(u.z));
        if (i < nd) b = b || !isfinite(rho[i]);
        bad |= b ? 1 : 0;
    }
    if (__any(bad) && (threadIdx.x & 63) == 0) atomicMax(&out[0], 1u);
}

// few blocks: each ends with one same-address atomic, which serialise
// Per-scene maxima of |v|^2 and |a|^2: blockIdx.y picks the scene, blocks along
// x stride over its particles; one atomicMax per block and quantity.
__global__ __launch_bounds__(kRedBlock) void k_scene_maxima(const fs_scene* __restrict__ scenes,
                                                            const float* __restrict__ v, const float* __restrict__ a,
                                                            int use_accel, uint32_t* out)
{
    __shared__ float s_m[2][kRedBlock / 64];
    const fs_scene sc = scenes[blockIdx.y];
    if (blockIdx.x * kRedBlock >= sc.particle_count) return;   // past a smaller scene: nothing to add
    float m0 = 0.0f, m1 = 0.0f;
    const int stride = gridDim.x * kRedBlock;
    for (int k0 = blockIdx.x * kRedBlock + threadIdx.x; k0 < sc.particle_count; k0 += kRedUnroll * stride) {
        v3 vv[kRedUnroll], aa[kRedUnroll];
#pragma unroll
        for (int u = 0; u < kRedUnroll; ++u) {
            const int k = k0 + u * stride;
            const bool in = k < sc.particle_count;
            vv[u] = in ? ld3(v, sc.particle_offset + k) : V(0.0f, 0.0f, 0.0f);
            aa[u] = (in && use_accel) ? ld3(a, sc.particle_offset + k) : V(0.0f, 0.0f, 0.0f);
        }
#pragma unroll
        for (int u = 0; u < kRedUnroll; ++u) {
            m0 = fmax_(m0, norm2_or_inf(vv[u]));
            m1 = fmax_(m1, norm2_or_inf(aa[u]));
        }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        m0 = fmax_(m0, __shfl_xor(m0, off, 64));
        m1 = fmax_(m1, __shfl_xor(m1, off, 64));
    }
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) {
        s_m[0][wid] = m0;
        s_m[1][wid] = m1;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < kRedBlock / 64; ++w) {
            m0 = fmax_(m0, s_m[0][w]);
            m1 = fmax_(m1, s_m[1][w]);
        }
        atomicMax(&out[2 * blockIdx.y], __float_as_uint(m0));
        atomicMax(&out[2 * blockIdx.y + 1], __float_as_uint(m1));
    }
}

// DamBreakSimulation.step's dt policy per scene  (simulation.py:166-187)
__global__ void k_scene_dt(fs_scene* scenes, int num_scenes, uint32_t* maxima, double stiffness, double exponent,
                           double rest_density, double viscosity, double frame_dt, int substeps, double* remaining,
                           int* status)
{
    const int sidx = blockIdx.x * blockDim.x + threadIdx.x;
    if (sidx >= num_scenes) return;
    const float v2 = __uint_as_float(maxima[2 * sidx]);
    const float a2 = __uint_as_float(maxima[2 * sidx + 1]);
    maxima[2 * sidx] = 0u;
    maxima[2 * sidx + 1] = 0u;
    double rem = remaining[sidx];
    double dt = 0.0;
    if (rem > 1.0e-12) {
        if (!isfinite(v2) || !isfinite(a2)) atomicMax(status, 1);
        const double speed = sqrt((double)v2);
        const double accel_max = sqrt((double)a2);
        const double h = scenes[sidx].cfl_smoothing_length > 0.0 ? scenes[sidx].cfl_smoothing_length
                                                                 : (double)scenes[sidx].smoothing_length;
        double limit = frame_dt / substeps;
        const double sound_speed = sqrt(stiffness * exponent / rest_density);
        limit = fmin(limit, 0.25 * h / (sound_speed + speed));
        if (accel_max > 0.0) limit = fmin(limit, 0.25 * sqrt(h / accel_max));
        if (viscosity > 0.0) limit = fmin(limit, 0.025 * (h * h) / viscosity);   // 0.025 * h**2 / viscosity
        dt = fmin(rem, limit);
        if (dt < 1.0e-10) atomicMax(status, 2);
        remaining[sidx] = rem - dt;
    }
    scenes[sidx].dt = (float)dt;
}

// Device address of `scratch` on the current device, looked up once per device.
constexpr int kNoAllocMaxDevices = 64;
static std::mutex g_noalloc_mutex;
static uint32_t* g_noalloc_ptr[kNoAllocMaxDevices] = {};

static int32_t noalloc_scratch(hipStream_t stream, uint32_t** out)
{
    int dev = 0;
    int32_t rc = hip_check(hipGetDevice(&dev), "fs_batch_cfl_dt_noalloc: hipGetDevice");
    if (rc) return rc;
    if (stream) {
        int sdev = 0;
        rc = hip_check(hipStreamGetDevice(stream, &sdev), "fs_batch_cfl_dt_noalloc: hipStreamGetDevice");
        if (rc) return rc;
        if (sdev != dev) {
            set_error("fs_batch_cfl_dt_noalloc: the stream belongs to device %d, the current device is %d", sdev,
                      dev);
            return FS_ERR_ARG;
        }
    }
    if (dev < 0 || dev >= kNoAllocMaxDevices) {
        set_error("fs_batch_cfl_dt_noalloc: device %d outside the supported 0..%d", dev, kNoAllocMaxDevices - 1);
        return FS_ERR_ARG;
    }
    std::lock_guard<std::mutex> lock(g_noalloc_mutex);
    if (!g_noalloc_ptr[dev]) {
        void* p = nullptr;
        rc = hip_check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "fs_batch_cfl_dt_noalloc: hipGetSymbolAddress");
        if (rc) return rc;
        g_noalloc_ptr[dev] = static_cast<uint32_t*>(p);
    }
    *out = g_noalloc_ptr[dev];
    return FS_OK;
}

static int32_t check_cfl_args(const char* who, const fs_scene* scenes, int32_t num_scenes, const float* v,
                              const float* a, int32_t use_accel, const double* remaining, const int32_t* status,
                              int32_t substeps)
{
    if (!scenes || num_scenes <= 0 || num_scenes > 65535 || !v || (use_accel && !a) || !remaining || !status ||
        substeps <= 0) {
        set_error("%s: invalid arguments", who);
        return FS_ERR_ARG;
    }
    return FS_OK;
}

static int32_t launch_cfl(fs_scene* scenes, int32_t num_scenes, int32_t max_scene_particles, const float* v,
                          const float* a, int32_t use_accel, double stiffness, double exponent, double rest_density,
                          double viscosity, double frame_dt, int32_t substeps, double* remaining, uint32_t* maxima,
                          int32_t* status, hipStream_t s)
{
    // blocks per scene: enough for the largest scene at about 32 particles
    // per lane (8 trips of kRedUnroll), at most 128; a block past a smaller
    // scene's particles returns at once
    const int per_scene = max_scene_particles > 0 ? std::min(128, blocks_for(max_scene_particles, kRedBlock * 32)) : 8;
    FS_LAUNCH(k_scene_maxima, dim3(per_scene, num_scenes), dim3(kRedBlock), 0, s, scenes, v, a, use_accel, maxima);
    FS_LAUNCH(k_scene_dt, dim3(blocks_for(num_scenes, 64)), dim3(64), 0, s, scenes, num_scenes, maxima, stiffness,
              exponent, rest_density, viscosity, frame_dt, substeps, remaining, status);
    return FS_OK;
}

// DamBreakSimulation.step's DFSPH dt policy per scene (simulation.py:166-187):
// limit = frame_dt/substeps, and 0.4 spacing / vmax when vmax > 0
__global__ void k_scene_dt_dfsph(fs_scene* scenes, int num_scenes, uint32_t* maxima, double spacing, double frame_dt,
                                 int substeps, double* remaining, int* status)
{
    const int sidx = blockIdx.x * blockDim.x + threadIdx.x;
    if (sidx >= num_scenes) return;
    const float v2 = __uint_as_float(maxima[2 * sidx]);
    maxima[2 * sidx] = 0u;
    maxima[2 * sidx + 1] = 0u;
    double rem = remaining[sidx];
    double dt = 0.0;
    if (rem > 1.0e-12) {
        if (!isfinite(v2)) atomicMax(status, 1);
        const double speed = sqrt((double)v2);
        double limit = frame_dt / substeps;
        if (speed > 0.0) limit = fmin(limit, 0.4 * spacing / speed);
        dt = fmin(rem, limit);
        if (dt < 1.0e-10) atomicMax(status, 2);
        remaining[sidx] = rem - dt;
    }
    scenes[sidx].dt = (float)dt;
}

static int red_blocks(int n)
{
    int b = blocks_for(n, kRedBlock * 8);
    if (b > 128) b = 128;
    return b > 0 ? b : 1;
}

}  // namespace fs

using namespace fs;

extern "C" {

int32_t fs_norm_maxima(const float* v, int32_t nv, const float* a, int32_t na, float* maxima, fs_stream_t stream)
{
    if (!maxima || nv < 0 || na < 0 || (nv > 0 && !v) || (na > 0 && !a)) {
        set_error("fs_norm_maxima: invalid arguments");
        return FS_ERR_ARG;
    }
    hipStream_t s = (hipStream_t)stream;
    uint32_t* out = reinterpret_cast<uint32_t*>(maxima);
    FS_LAUNCH(k_fill_zero, dim3(1), dim3(64), 0, s, out, 2);
    int n = nv > na ? nv : na;
    if (n == 0) return FS_OK;
    FS_LAUNCH(k_norm_maxima, dim3(red_blocks(n)), dim3(kRedBlock), 0, s, v, nv, a, na, out);
    return FS_OK;
}

int32_t fs_batch_cfl_dt(fs_scene* scenes, int32_t num_scenes, int32_t max_scene_particles, const float* v,
                        const float* a, int32_t use_accel, double stiffness, double exponent, double rest_density,
                        double viscosity, double frame_dt, int32_t substeps, double* remaining, uint32_t* scratch,
                        int32_t* status, fs_stream_t stream)
{
    int32_t rc = check_cfl_args("fs_batch_cfl_dt", scenes, num_scenes, v, a, use_accel, remaining, status, substeps);
    if (rc) return rc;
    if (!scratch) {
        set_error("fs_batch_cfl_dt: invalid arguments (null scratch)");
        return FS_ERR_ARG;
    }
    return launch_cfl(scenes, num_scenes, max_scene_particles, v, a, use_accel, stiffness, exponent, rest_density,
                      viscosity, frame_dt, substeps, remaining, scratch, status, (hipStream_t)stream);
}

int32_t fs_dfsph_cfl_dt(fs_scene* scenes, int32_t num_scenes, int32_t max_scene_particles, const float* v,
                        double spacing, double frame_dt, int32_t substeps, double* remaining, uint32_t* scratch,
                        int32_t* status, fs_stream_t stream)
{
    int32_t rc = check_cfl_args("fs_dfsph_cfl_dt", scenes, num_scenes, v, nullptr, 0, remaining, status, substeps);
    if (rc) return rc;
    if (!scratch) {
        set_error("fs_dfsph_cfl_dt: invalid arguments (null scratch)");
        return FS_ERR_ARG;
    }
    hipStream_t s = (hipStream_t)stream;
    const int per_scene = max_scene_particles > 0 ? std::min(128, blocks_for(max_scene_particles, kRedBlock * 8)) : 8;
    FS_LAUNCH(k_scene_maxima, dim3(per_scene, num_scenes), dim3(kRedBlock), 0, s, scenes, v, nullptr, 0, scratch);
    FS_LAUNCH(k_scene_dt_dfsph, dim3(blocks_for(num_scenes, 64)), dim3(64), 0, s, scenes, num_scenes, scratch, spacing,
              frame_dt, substeps, remaining, status);
    return FS_OK;
}

int32_t fs_batch_cfl_dt_noalloc(fs_scene* scenes, int32_t num_scenes, int32_t max_scene_particles, const float* v,
                                const float* a, int32_t use_accel, double stiffness, double exponent,
                                double rest_density, double viscosity, double frame_dt, int32_t substeps,
                                double* remaining, int32_t* status, fs_stream_t stream)
{
    int32_t rc = check_cfl_args("fs_batch_cfl_dt_noalloc", scenes, num_scenes, v, a, use_accel, remaining, status,
                                substeps);
    if (rc) return rc;
    if (num_scenes > kNoAllocMaxScenes) {
        set_error("fs_batch_cfl_dt_noalloc: %d scenes exceed the fixed scratch of %d scenes (use fs_batch_cfl_dt)",
                  num_scenes, kNoAllocMaxScenes);
        return FS_ERR_CAPACITY;
    }
    hipStream_t s = (hipStream_t)stream;
    uint32_t* maxima = nullptr;
    rc = noalloc_scratch(s, &maxima);
    if (rc) return rc;
    // the array keeps the previous call's contents: the maxima must start at 0
    const int words = 2 * num_scenes;
    FS_LAUNCH(k_fill_zero_n, dim3(std::min(32, blocks_for(words, 256))), dim3(256), 0, s, maxima, words);
    return launch_cfl(scenes, num_scenes, max_scene_particles, v, a, use_accel, stiffness, exponent, rest_density,
                      viscosity, frame_dt, substeps, remaining, maxima, status, s);
}

int32_t fs_check_finite(const float* x, const float* v, int32_t n, const float* density, int32_t nd,
                        int32_t* invalid, fs_stream_t stream)
{
    if (!invalid || n < 0 || nd < 0 || nd > n || (n > 0 && (!x || !v)) || (nd > 0 && !density)) {
        set_error("fs_check_finite: invalid arguments");
        return FS_ERR_ARG;
    }
    hipStream_t s = (hipStream_t)stream;
    uint32_t* out = reinterpret_cast<uint32_t*>(invalid);
    FS_LAUNCH(k_fill_zero, dim3(1), dim3(64), 0, s, out, 1);
    if (n == 0) return FS_OK;
    FS_LAUNCH(k_check_finite, dim3(red_blocks(n)), dim3(kRedBlock), 0, s, x, v, n, density, nd, out);
    return FS_OK;
}

}  // extern "C"
