// fs_internal.h -- host-side interfaces shared between the library's
// translation units (grid build stages used by the solver steps).
#pragma once

#include "fs_common.h"

namespace fs {

constexpr int kElemBlock = 256;   // elementwise / per-slot kernels

// What the hash kernel does to each original particle before bucketing it.
enum PreMode : int {
    kPrePlain = 0,        // bucket points[i]
    kPreIntegrate = 1,    // x_out = x_in + v*dt for dynamic particles (solver_dfsph.py:340-353)
    kPrePredict = 2,      // v_pred, x_pred, position bounds (solver_pbf.py:192-224,366-399)
};

struct PreOp {
    int mode = kPrePlain;
    const float* x_in = nullptr;   // positions to bucket (mode 0) or x_in
    const float* v_in = nullptr;   // velocity (mode 1: v_star, mode 2: v_in)
    float* x_out = nullptr;        // mode 1/2 output positions
    float* v_out = nullptr;        // mode 2: v_pred
    const int* flags = nullptr;
    const int* mask = nullptr;
    const int* world = nullptr;
    const float* gravity = nullptr;
    float dt = 0.0f;
    const float* dt_device = nullptr;   // replaces dt when set
    float bound[3] = {0, 0, 0};
    float particle_radius = 0.0f;
    float boundary_epsilon = 0.0f;
    int* control_reset = nullptr;  // zeroed by the hash kernel when non-null
    const fs_scene* scenes = nullptr;   // segmented grid: per-scene cell width, dims, bucket base
    const int* scene_of = nullptr;      // [n] scene of each point
};

// Particle data copied into sorted order by the final build stage.
struct GatherOp {
    const float* x = nullptr;      // positions in original order (what was bucketed)
    const float* v = nullptr;      // optional velocity source
    const float* mass = nullptr;           // optional (pos.w = 0 without it)
    const float* rest_density = nullptr;   // read only when wvol is set
    const int* flags = nullptr;
    const int* mask = nullptr;             // optional (no mask: dynamic = active)
    float4* pos = nullptr;         // x,y,z,mass
    float4* vel = nullptr;         // vx,vy,vz,0   (when v != nullptr)
    float* wvol = nullptr;         // optional: rest_density * volume
    int* sflags = nullptr;         // bit0 active, bit1 dynamic
    const int* scene_of = nullptr; // optional: scene ids, gathered to sscene
    int* sscene = nullptr;
    int* any_inactive = nullptr;   // optional: set to 1 when some point is inactive
    // optional scratch [n]: (mass, rest_density, flags, mask) of each point in
    // original order, packed by the build before the fix-up so that the
    // fix-up fetches one 16-byte record per point instead of four words
    int4* packed = nullptr;
};

// Full build: hash (+pre-op), bucket scan, scatter, stable fix-up (+gather).
int32_t build_grid(fs_grid* g, int n, float radius, const PreOp& pre, const GatherOp& ga, hipStream_t s);

GridView grid_view(const fs_grid* g);

inline int blocks_for(int n, int block) { return (n + block - 1) / block; }

// Optional per-launch timing (fs_profile_enable): events recorded on the
// launch stream around each kernel, summed per kernel name on collection.
extern bool g_profiling;
void prof_mark(const char* name, hipStream_t s, bool begin);

#define FS_LAUNCH(kernel, grid, block, shmem, stream, ...)                         \
    do {                                                                           \
        if (::fs::g_profiling) ::fs::prof_mark(#kernel, stream, true);             \
        hipLaunchKernelGGL(kernel, grid, block, shmem, stream, __VA_ARGS__);       \
        if (::fs::g_profiling) ::fs::prof_mark(#kernel, stream, false);            \
        FS_LAUNCH_CHECK(#kernel);                                                  \
    } while (0)

}  // namespace fs
