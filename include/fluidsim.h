/*
 * fluidsim.h -- C-ABI of libfluidsim_hip.so, the MI355X (gfx950) particle-fluid
 * hot path: Warp-compatible hash-grid build and query, and the neighbour-loop
 * passes of the WCSPH, DFSPH and PBF solvers.
 *
 * Conventions
 *   - Every pointer argument named in an *_io / fs_model struct is a DEVICE
 *     pointer owned by the caller (e.g. torch.Tensor.data_ptr()).  Vectors are
 *     AoS float32 xyz with stride 12 bytes (Warp vec3 / torch (N,3)).
 *   - Workspaces are caller-allocated device memory sized by *_bytes().  No
 *     entry point allocates device memory; none synchronises the host except
 *     the ones whose name ends in _readback.
 *   - fs_stream_t is a hipStream_t (NULL = default stream).  Work is enqueued
 *     in order on that stream.
 *   - Return value: FS_OK (0) or a negative FS_ERR_* code; fs_last_error()
 *     then describes the failure (thread-local).
 *
 * Reference interfaces replaced (SParksLz/fluid_simulation):
 *   - warp HashGrid(128,128,128).reserve/build and hash_grid_query, as used by
 *     solver/sph/solver_wcsph.py:398-402,456; solver/sph/solver_dfsph.py:541-545,
 *     624,783; solver/pbf/solver_pbf.py:634-638,737,858.
 *   - SolverWCSPH.step   solver/sph/solver_wcsph.py:431-535
 *   - SolverDFSPH.step   solver/sph/solver_dfsph.py:598-898
 *   - SolverPBF.step     solver/pbf/solver_pbf.py:676-954
 *   - particle_norm_maxima / check_particle_finite
 *                        examples/dam_break/simulation.py:19-53
 */
#ifndef FLUIDSIM_H
#define FLUIDSIM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef void* fs_stream_t;

enum {
    FS_OK = 0,
    FS_ERR_ARG = -1,        /* bad argument (null pointer, negative size, ...) */
    FS_ERR_CAPACITY = -2,   /* more points than the grid / workspace reserved  */
    FS_ERR_HIP = -3,        /* a HIP runtime call failed                       */
    FS_ERR_WORKSPACE = -4   /* workspace too small                             */
};

/* Description of the last error on the calling thread ("" if none). */
const char* fs_last_error(void);
/* Library ABI version (major*100 + minor). */
int32_t fs_version(void);

/* Per-kernel timing for benchmarks (not thread-safe; host-synchronising when
 * collected).  While enabled, every kernel the library launches is bracketed
 * by HIP events on its stream.  fs_profile_report writes one line per kernel,
 * "name launches total_ms", into buf (NUL-terminated, truncated to len) and
 * returns the full text length; it synchronises the recorded events. */
void fs_profile_enable(int32_t on);
void fs_profile_reset(void);
int32_t fs_profile_report(char* buf, size_t len);

/* ------------------------------------------------------------------------ */
/* Launch-variant table of the tunable passes (no counterpart in the       */
/* reference).  Each tunable pass has a few launch variants (workgroup size, */
/* depth of the per-lane LDS survivor list, neighbours a list replay fetches */
/* together, particles per lane, lanes per particle) with bit-identical     */
/* results; the selection is process-wide, guarded by a mutex, variant 0 by */
/* default.                                                                 */
/* ------------------------------------------------------------------------ */

#define FS_TUNE_WCSPH_DENSITY 0   /* fs_wcsph_step density/EOS pass          */
#define FS_TUNE_WCSPH_FORCE 1     /* fs_wcsph_step force/integrate pass      */
#define FS_TUNE_BATCH_DENSITY 2   /* fs_batch_wcsph_step density/EOS pass    */
#define FS_TUNE_BATCH_FORCE 3     /* fs_batch_wcsph_step force/integrate pass */
#define FS_TUNE_DFSPH_SCAN 4      /* DFSPH density/factor passes             */
#define FS_TUNE_DFSPH_REPLAY 5    /* DFSPH non-pressure, rate, correction    */
#define FS_TUNE_PBF_SCAN 6        /* PBF density and lambda passes           */
#define FS_TUNE_PBF_REPLAY 7      /* PBF non-pressure, delta, XSPH passes     */

int32_t fs_tune_variant_count(int32_t pass);
int32_t fs_tune_describe(int32_t pass, int32_t variant, int32_t* block, int32_t* list_depth);
/* ... and the replay unroll (neighbours a lane fetches together), particles
 * per lane and lanes per particle of a variant (any pointer may be NULL) */
int32_t fs_tune_describe_ex(int32_t pass, int32_t variant, int32_t* block, int32_t* list_depth, int32_t* unroll,
                            int32_t* particles_per_thread, int32_t* lanes_per_particle);
/* ... and its scan kind: 0 for the per-lane scan, else the cooperative cell
 * scan (candidates staged per query group in LDS and tested against every
 * particle of the group at once) with this many 64-candidate chunks per
 * group; list_depth then counts the tile's chunk slots.  Negative on error. */
int32_t fs_tune_variant_coop(int32_t pass, int32_t variant);
int32_t fs_tune_select(int32_t pass, int32_t variant);
int32_t fs_tune_selected(int32_t pass);

/* Diagnostic (no counterpart in the reference): compares, on the device,
 * the shared-reciprocal float32 quotient and square-root forms the
 * neighbour passes use on in-range operands with the correctly rounded '/'
 * and sqrtf.  mode 0: sqrt of x = bits(first + t); 1: random in-range (x, d)
 * pairs; 2: x = bits(first + t) over d = bits(seed); 3: random separations
 * and smoothing lengths through the cubic kernel's |r|, gradient and value;
 * t runs over [0, count).  result (host uint64[2]) = number of bitwise
 * mismatches and the first mismatching t (UINT64_MAX if none).
 * Synchronises the stream. */
int32_t fs_check_fast_math(int32_t mode, uint64_t first, uint64_t count, uint32_t seed, uint64_t* result,
                           fs_stream_t stream);

/* ------------------------------------------------------------------------ */
/* Hash grid                                                                */
/* ------------------------------------------------------------------------ */

/* Device-resident uniform hash grid with Warp HashGrid semantics:
 *   bucket of a point  = hash(round(x/w), round(y/w), round(z/w)) with
 *                        hash(c) = max(c + 2^20, 0) mod dim per axis and
 *                        id = bz*dim_x*dim_y + by*dim_x + bx;
 *   sorted order       = stable by (bucket, original index);
 *   query(p, r)        = buckets of cells round((p-r)/w)..round((p+r)/w) per
 *                        axis (at most dim per axis), x fastest, then y, z.
 * Fields are filled by fs_grid_init / fs_grid_build; read them, do not write. */
typedef struct fs_grid {
    int32_t dim_x, dim_y, dim_z;
    int32_t round_mode;     /* 0 = floor (default), 1 = truncate toward zero  */
    int32_t capacity;       /* points reserved                                 */
    int32_t num_points;     /* points in the last build                        */
    float cell_width;       /* radius of the last build                        */
    float cell_width_inv;   /* 1/cell_width in float32                         */
    int32_t* point_ids;     /* [capacity] sorted slot -> original index        */
    int32_t* point_cells;   /* [capacity] sorted bucket keys                   */
    int32_t* cell_of;       /* [capacity] scratch: bucket of the point at each
                               position of the build's processing order (the
                               previous build's sorted order when it covered
                               the same points, else original order); the
                               bucket of original point point_ids[s] is
                               point_cells[s]                                   */
    int32_t* cell_start;    /* [dims+1] exclusive prefix over buckets           */
    int32_t* cell_count;    /* [dims] per-bucket counters, zero between builds  */
    int32_t* rank;          /* [capacity] scratch, in the same order as cell_of */
    int32_t* scatter_ids;   /* [2*capacity] scratch: (index, bucket) pairs     */
    uint64_t* scan_state;   /* look-back descriptors of the bucket scan         */
    uint32_t* sync;         /* [8] self-resetting counters                      */
} fs_grid;

/* Largest grid / workspace capacity (points, particles). */
#define FS_MAX_PARTICLES ((1 << 28) - 1)

/* Bytes of device workspace for a grid reserving `capacity` points. */
size_t fs_grid_workspace_bytes(int32_t capacity, int32_t dim_x, int32_t dim_y, int32_t dim_z);

/* Carve `workspace` (device, >= fs_grid_workspace_bytes) into the grid arrays
 * and zero the counters on `stream`.  Warp HashGrid(dim_x,dim_y,dim_z) +
 * reserve(capacity)  (solver_wcsph.py:398-402).  capacity is at most
 * FS_MAX_PARTICLES (positions are addressed with 32-bit byte offsets). */
int32_t fs_grid_init(fs_grid* grid, void* workspace, size_t workspace_bytes, int32_t capacity, int32_t dim_x,
                     int32_t dim_y, int32_t dim_z, int32_t round_mode, fs_stream_t stream);

/* Build from `n` AoS float3 points with cell width `radius`.
 * Warp HashGrid.build(points, radius)  (solver_wcsph.py:456). */
int32_t fs_grid_build(fs_grid* grid, const float* points, int32_t n, float radius, fs_stream_t stream);

/* Warp-layout bucket tables of the last build: cell_starts/cell_ends [dims],
 * 0/0 for empty buckets (as Warp's build leaves them). */
int32_t fs_grid_export_tables(const fs_grid* grid, int32_t* cell_starts, int32_t* cell_ends, fs_stream_t stream);

/* Per-point candidate counts (counts[i] for original point i) of
 * hash_grid_query(points[i], radius) on the last build.
 * within_mode: -1 all candidates (self and aliased buckets included),
 *               0 cubic-spline support  fl(sqrt(d2))/radius <= 1,
 *               1 PBF support           fl(sqrt(d2)) < radius. */
int32_t fs_grid_count_candidates(const fs_grid* grid, const float* points, float radius, int32_t within_mode,
                                 int32_t* counts, fs_stream_t stream);

/* Candidate indices (original numbering) in query visit order, CSR layout:
 * indices[offsets[i] .. offsets[i+1]) for point i.  offsets: int64 [n+1]
 * (e.g. an exclusive scan of fs_grid_count_candidates). */
int32_t fs_grid_list_candidates(const fs_grid* grid, const float* points, float radius, int32_t within_mode,
                                const int64_t* offsets, int32_t* indices, fs_stream_t stream);

/* ------------------------------------------------------------------------ */
/* Solver workspace and model                                               */
/* ------------------------------------------------------------------------ */

/* Per-particle model arrays in original order (Newton Model + sph: MODEL
 * attributes, solver_wcsph.py:320-392).  stiffness/exponent are read by
 * WCSPH only and may be NULL otherwise. */
typedef struct fs_model {
    int32_t n;
    int32_t num_worlds;
    const int32_t* flags;          /* ParticleFlags, bit 0 = ACTIVE            */
    const int32_t* mask;           /* sph:particle_mask                       */
    const int32_t* world;          /* particle_world                          */
    const float* mass;
    const float* gravity;          /* [num_worlds*3]                          */
    const float* rest_density;
    const float* surface_tension;
    const float* viscosity;
    const float* stiffness;
    const float* exponent;
} fs_model;

/* Solver scratch carved from one caller-allocated workspace: particle data
 * gathered into the grid's sorted order, plus reduction scratch. */
typedef struct fs_work {
    int32_t capacity;
    int32_t list_capacity; /* neighbour-list entries per particle (0: no lists) */
    float* pos;         /* [cap*4] x,y,z,mass  sorted                          */
    float* pos_alt;     /* [cap*4] ping-pong positions (PBF)                    */
    float* vel;         /* [cap*4] vx,vy,vz,-  sorted                          */
    float* sca;         /* [cap*4] neighbour-visible per-pass scalars, sorted  */
    float* wvol;        /* [cap]   rest_density*volume, sorted                 */
    int32_t* flags;     /* [cap]   flags (bit0 active, bit1 dynamic), sorted   */
    float* x_pred;      /* [cap*3] original order (PBF x_pred / DFSPH scratch)  */
    float* v_pred;      /* [cap*3] original order (PBF v_pred)                 */
    int32_t* nbr;       /* [list_capacity*cap] neighbour lists (entry-major)    */
    int32_t* nbr_count; /* [cap] list lengths (list_capacity+1 = overflowed)    */
    double* partials;   /* residual partial sums                               */
    int32_t* control;   /* [64] loop control: stop flags, iteration counters;
                           word FS_CONTROL_LIST_NEED: see below            */
    uint32_t* sync;     /* [8]  self-resetting counters                        */
    int32_t* packed;    /* [cap*4] grid-build scratch, original order: mass,
                           rest density, flags, mask of each particle packed
                           into one 16-byte record (repacked every build)    */
} fs_work;

/* list_capacity: entries per particle of the neighbour lists recorded by the
 * first pass after each grid build and replayed by later passes over the
 * same candidates (0 disables them; every pass then scans the grid).  The
 * lists take capacity * list_capacity * 4 bytes; capacity <= FS_MAX_PARTICLES. */
size_t fs_work_bytes(int32_t capacity, int32_t list_capacity);

/* control[FS_CONTROL_LIST_NEED]: the longest list a particle needed whose list
 * overflowed list_capacity (0 while none has).  fs_work_init clears it, and
 * so does every fs_dfsph_step (its first grid build clears the loop-control
 * words); the host reads it to re-initialise a workspace with longer lists. */
#define FS_CONTROL_LIST_NEED 7
int32_t fs_work_init(fs_work* work, void* workspace, size_t workspace_bytes, int32_t capacity,
                     int32_t list_capacity, fs_stream_t stream);

/* ------------------------------------------------------------------------ */
/* WCSPH                                                                    */
/* ------------------------------------------------------------------------ */

typedef struct fs_wcsph_params {
    float smoothing_length;        /* h = particle_length * coff             */
    float particle_radius;         /* e_dist floor in the cohesion term      */
    float bound[3];                /* bound_width, bound_height, bound_length */
    float boundary_damping;
    float dt;
    int32_t clamp_negative_pressure;
    /* Optional device scalar that replaces `dt` (NULL: use `dt`).  With it a
     * substep needs no host value, so a device-side CFL (fs_batch_cfl_dt on
     * a one-scene table) and graph capture drive whole frames.  A device dt
     * of 0 copies the state (x, v, and rho/pressure when io->rho_in and
     * io->pressure_in are set) to the outputs and leaves the acceleration
     * unchanged. */
    const float* dt_device;
} fs_wcsph_params;

typedef struct fs_wcsph_io {
    const float* x_in;   /* state_in.particle_q   */
    const float* v_in;   /* state_in.particle_qd  */
    float* x_out;        /* state_out.particle_q  */
    float* v_out;        /* state_out.particle_qd */
    float* rho;          /* state_out.sph.rho     */
    float* pressure;     /* state_out.sph.pressure */
    float* accel;        /* solver._particle_accel [n*3] */
    /* Optional (may be NULL): state_in.sph.rho / .pressure.  A substep whose
     * device dt is 0 copies them to rho / pressure, so that an idle substep
     * carries the whole state over, density and pressure included. */
    const float* rho_in;
    const float* pressure_in;
} fs_wcsph_io;

/* One WCSPH substep: build, compute_density_pressure, acceleration_wcsph,
 * integrate_velocity, integrate_positions, apply_bounds
 * (SolverWCSPH.step, solver_wcsph.py:431-535). */
int32_t fs_wcsph_step(fs_grid* grid, fs_work* work, const fs_model* model, const fs_wcsph_params* params,
                      const fs_wcsph_io* io, fs_stream_t stream);
/* fs_wcsph_step is exactly these two passes, in order, on one stream:
 * build + compute_density_pressure (solver_wcsph.py:456-474; records the
 * neighbour lists), then acceleration_wcsph + integrate_velocity +
 * integrate_positions + apply_bounds (:475-535; replays them). */
int32_t fs_wcsph_density_eos(fs_grid* grid, fs_work* work, const fs_model* model, const fs_wcsph_params* params,
                             const fs_wcsph_io* io, fs_stream_t stream);
int32_t fs_wcsph_force_integrate(fs_grid* grid, fs_work* work, const fs_model* model,
                                 const fs_wcsph_params* params, const fs_wcsph_io* io, fs_stream_t stream);

/* WCSPH adjoint.  No counterpart in the reference, whose gradients would come
 * from Warp's autodiff tape; this is the reverse-mode derivative of
 * fs_wcsph_step (SolverWCSPH.step, solver_wcsph.py:431-535): density, EOS
 * pressure, pressure and viscosity forces, symplectic Euler and the
 * outward-damping bounds.  Surface tension is taken as 0 (its cohesion term
 * is not differentiated); callers with non-zero surface tension must not use
 * it. */
typedef struct fs_wcsph_adjoint_io {
    fs_wcsph_io forward;       /* the substep's inputs; the forward is recomputed
                                  into its outputs (x_out, v_out, rho, pressure,
                                  accel), as fs_wcsph_step would write them     */
    const float* xbar_out;     /* [n*3] dL/dx_out                               */
    const float* vbar_out;     /* [n*3] dL/dv_out                               */
    float* xbar_in;            /* [n*3] dL/dx_in  (may alias xbar_out)          */
    float* vbar_in;            /* [n*3] dL/dv_in  (may alias vbar_out)          */
    float* grad_stiffness;     /* [n] += dL/d(stiffness_i) of this substep      */
    float* grad_viscosity;     /* [n] += dL/d(viscosity_i) of this substep      */
    void* scratch;             /* device, >= fs_wcsph_adjoint_bytes(n)          */
} fs_wcsph_adjoint_io;

size_t fs_wcsph_adjoint_bytes(int32_t capacity);

/* One substep's adjoint: recompute the forward substep (grid build, density
 * pass recording the neighbour lists, force pass), then three passes over
 * the same lists -- reverse integration/bounds, reverse force + EOS, reverse
 * density -- each a gather per particle with no atomics.  `work` must have
 * list_capacity > 0. */
int32_t fs_wcsph_adjoint_step(fs_grid* grid, fs_work* work, const fs_model* model, const fs_wcsph_params* params,
                              const fs_wcsph_adjoint_io* io, size_t scratch_bytes, fs_stream_t stream);

/* ------------------------------------------------------------------------ */
/* DFSPH                                                                    */
/* ------------------------------------------------------------------------ */

typedef struct fs_dfsph_params {
    double eta_density;        /* compared in float64, as the reference's   */
    double eta_divergence;     /* Python-side residual test (:733, :850)    */
    float smoothing_length;
    float particle_radius;
    float bound[3];
    float boundary_damping;
    float dt;
    float max_kappa;
    float max_kappa_v;
    int32_t max_rho_iterations;
    int32_t max_vel_iterations;
    int32_t warm_start;
    int32_t min_neighbors_for_factor;
    int32_t exact_residual;    /* 1: form numpy's float32 mean for every break
                                  test (slow; for checks).  0: only when the
                                  float64 mean lies within float32 rounding
                                  of eta                                      */
    /* Optional device scalar that replaces `dt` (NULL: use `dt`), as for
     * WCSPH: a device-side CFL (fs_dfsph_cfl_dt) and graph capture then
     * drive whole frames.  A device dt of 0 makes the substep idle: it
     * copies x_in, v_in and the *_in state arrays that are set to the
     * outputs and changes nothing else. */
    const float* dt_device;
} fs_dfsph_params;

typedef struct fs_dfsph_io {
    const float* x_in;
    const float* v_in;
    const float* kappa_in;     /* state_in.sph.kappa   (warm start)  */
    const float* kappa_v_in;   /* state_in.sph.kappa_v (warm start)  */
    float* x_out;
    float* v_out;
    float* rho;
    float* factor;
    float* rho_star;
    float* d_rho_dt;
    float* kappa;
    float* kappa_v;
    int32_t* neighbor_count;
    float* v_star;             /* [n*3] */
    float* accel;              /* [n*3] solver._particle_accel */
    int32_t* iterations;       /* device int32[4]: density iterations,
                                  density corrections, divergence iterations,
                                  divergence corrections (may be NULL) */
    /* Optional (may be NULL): the state_in.sph arrays an idle substep (device
     * dt 0) carries over to the outputs, with kappa_in / kappa_v_in. */
    const float* rho_in;
    const float* factor_in;
    const float* rho_star_in;
    const float* d_rho_dt_in;
    const int32_t* neighbor_count_in;
    const float* v_star_in;    /* [n*3] */
} fs_dfsph_io;

/* One DFSPH substep (SolverDFSPH.step, solver_dfsph.py:598-898).  The
 * residual early exits (:733-735, :849-851) are decided on the device with
 * the reference's arithmetic: residual = numpy's float32 mean of max(x, 0)
 * over the dynamic particles in original order (:579-596).  A float64 mean
 * decides when it lies farther from eta than float32 rounding can move the
 * float32 mean; otherwise the float32 mean is formed in numpy's order
 * (8192-element buffers, pairwise within each).  The remaining passes of a
 * finished loop return at once. */
int32_t fs_dfsph_step(fs_grid* grid, fs_work* work, const fs_model* model, const fs_dfsph_params* params,
                      const fs_dfsph_io* io, fs_stream_t stream);
/* The passes of fs_dfsph_step, one call each; the substep is exactly:
 *   fs_dfsph_density_factor(after_integrate = 0)   build on x_in + compute_density + compute_factor_dfsph (:624-653)
 *   fs_dfsph_nonpressure_predict                   acceleration_non_pressure + predict_velocity (:654-688)
 *   [fs_dfsph_correction(loop 0, warm_start = 1)]  when warm starting (:689-707)
 *   max_rho_iterations x:                          (:711-767)
 *     fs_dfsph_rate_kappa(loop 0)                  compute_density_change_rate + predict_density_star + compute_kappa_density
 *     fs_dfsph_residual(loop 0, iteration)         _density_residual and the break rule, decided on the device
 *     fs_dfsph_correction(loop 0, warm_start = 0)  apply_constant_density_correction
 *   fs_dfsph_integrate                             integrate_positions + build on x_out (:769-783)
 *   fs_dfsph_density_factor(after_integrate = 1)   compute_density + compute_factor_dfsph (:785-812)
 *   [fs_dfsph_correction(loop 1, warm_start = 1)]  when warm starting (:814-832)
 *   max_vel_iterations x: rate_kappa, residual, correction with loop 1 (:834-883)
 *   fs_dfsph_finalize                              apply_bounds and the outputs (:886-898)
 * Once a loop's residual has stopped it, that loop's later rate and
 * correction calls return at once (on the device). */
int32_t fs_dfsph_density_factor(fs_grid* grid, fs_work* work, const fs_model* model, const fs_dfsph_params* params,
                                const fs_dfsph_io* io, int32_t after_integrate, fs_stream_t stream);
int32_t fs_dfsph_nonpressure_predict(fs_grid* grid, fs_work* work, const fs_model* model,
                                     const fs_dfsph_params* params, const fs_dfsph_io* io, fs_stream_t stream);
int32_t fs_dfsph_rate_kappa(fs_grid* grid, fs_work* work, const fs_model* model, const fs_dfsph_params* params,
                            const fs_dfsph_io* io, int32_t loop, fs_stream_t stream);
int32_t fs_dfsph_residual(fs_grid* grid, fs_work* work, const fs_model* model, const fs_dfsph_params* params,
                          const fs_dfsph_io* io, int32_t loop, int32_t iteration, fs_stream_t stream);
int32_t fs_dfsph_correction(fs_grid* grid, fs_work* work, const fs_model* model, const fs_dfsph_params* params,
                            const fs_dfsph_io* io, int32_t loop, int32_t warm_start, fs_stream_t stream);
int32_t fs_dfsph_integrate(fs_grid* grid, fs_work* work, const fs_model* model, const fs_dfsph_params* params,
                           const fs_dfsph_io* io, fs_stream_t stream);
int32_t fs_dfsph_finalize(fs_grid* grid, fs_work* work, const fs_model* model, const fs_dfsph_params* params,
                          const fs_dfsph_io* io, fs_stream_t stream);

/* ------------------------------------------------------------------------ */
/* PBF                                                                      */
/* ------------------------------------------------------------------------ */

typedef struct fs_pbf_params {
    float smoothing_length;
    float particle_radius;
    float bound[3];
    float boundary_damping;
    float boundary_epsilon;
    float dt;
    float lambda_epsilon;
    float scorr_k;
    float scorr_n;
    float scorr_w_delta_q;     /* poly6(scorr_q*h)  (solver_pbf.py:671-674) */
    float max_delta_position;
    float xsph_c;
    int32_t max_iterations;
    int32_t min_neighbors_for_lambda;
    int32_t rebuild_grid_each_iteration;
} fs_pbf_params;

typedef struct fs_pbf_io {
    const float* x_in;
    const float* v_in;
    float* x_out;
    float* v_out;
    float* rho;
    float* pbf_lambda;
    int32_t* neighbor_count;
    float* delta_p;            /* [n*3] */
    float* accel;              /* [n*3] solver._particle_accel */
} fs_pbf_io;

/* One PBF substep (SolverPBF.step, solver_pbf.py:676-954). */
int32_t fs_pbf_step(fs_grid* grid, fs_work* work, const fs_model* model, const fs_pbf_params* params,
                    const fs_pbf_io* io, fs_stream_t stream);
/* The passes of fs_pbf_step, one call each; the substep is exactly:
 *   fs_pbf_predict          predict_velocity_with_gravity + predict_positions + apply_position_bounds + build (:702-737)
 *   fs_pbf_density          compute_density (:738-751)
 *   fs_pbf_nonpressure      acceleration_non_pressure + apply_non_pressure_prediction + bounds (:752-796)
 *   max_iterations x:       (:800-858)
 *     fs_pbf_lambda(i)      compute_pbf_lambda
 *     fs_pbf_delta(i)       compute_delta_position + apply_delta_position (+ bounds and rebuild as configured)
 *   fs_pbf_final_density    compute_density + update_velocity_from_positions (+ apply_bounds without XSPH) (:878-917)
 *   fs_pbf_xsph_finalize    apply_xsph_viscosity + apply_bounds; nothing when xsph_c <= 0 (:919-954)
 * Predicted positions alternate between the workspace's two sorted position
 * streams; which one a pass reads follows from the iteration. */
int32_t fs_pbf_predict(fs_grid* grid, fs_work* work, const fs_model* model, const fs_pbf_params* params,
                       const fs_pbf_io* io, fs_stream_t stream);
int32_t fs_pbf_density(fs_grid* grid, fs_work* work, const fs_model* model, const fs_pbf_params* params,
                       const fs_pbf_io* io, fs_stream_t stream);
int32_t fs_pbf_nonpressure(fs_grid* grid, fs_work* work, const fs_model* model, const fs_pbf_params* params,
                           const fs_pbf_io* io, fs_stream_t stream);
int32_t fs_pbf_lambda(fs_grid* grid, fs_work* work, const fs_model* model, const fs_pbf_params* params,
                      const fs_pbf_io* io, int32_t iteration, fs_stream_t stream);
int32_t fs_pbf_delta(fs_grid* grid, fs_work* work, const fs_model* model, const fs_pbf_params* params,
                     const fs_pbf_io* io, int32_t iteration, fs_stream_t stream);
int32_t fs_pbf_final_density(fs_grid* grid, fs_work* work, const fs_model* model, const fs_pbf_params* params,
                             const fs_pbf_io* io, fs_stream_t stream);
int32_t fs_pbf_xsph_finalize(fs_grid* grid, fs_work* work, const fs_model* model, const fs_pbf_params* params,
                             const fs_pbf_io* io, fs_stream_t stream);

/* ------------------------------------------------------------------------ */
/* APIC particle separation                                                 */
/* ------------------------------------------------------------------------ */

/* Host-side values of SolverAPIC._apply_particle_separation
 * (solver/apic/solver_apic.py:1003-1066) and SolverAPIC.__init__ (:640-650). */
typedef struct fs_separation_params {
    float rest_dist;     /* push apart pairs closer than this                    */
    float query_radius;  /* max(rest_dist, 1e-5): grid cell width, query radius  */
    float strength;      /* separation_strength                                  */
    float wall_band;     /* max(wall_push_band, 2 wall_margin): into-wall kill   */
    float wall_margin;   /* positions clamped to [lo + margin, hi - margin]      */
    float bound_lo[3];
    float bound_hi[3];
    int32_t iterations;  /* max(1, separation_iterations)                        */
} fs_separation_params;

/* `iterations` x (grid build at the current positions with cell width
 * query_radius, accumulate_particle_separation, apply_particle_separation_delta;
 * solver_apic.py:193-268, 1036-1066).  The grid is the caller's HashGrid of
 * the box dims; the workspace may have list_capacity 0.  positions [n*3] are
 * updated in place; delta [n*3] (optional) receives the last iteration's
 * displacement (zero for inactive particles, which are not moved). */
int32_t fs_particle_separation(fs_grid* grid, fs_work* work, const fs_separation_params* params,
                               float* positions, const int32_t* flags, float* delta, int32_t n,
                               fs_stream_t stream);

/* ------------------------------------------------------------------------ */
/* Batched WCSPH: many independent scenes in one launch per pass            */
/* ------------------------------------------------------------------------ */

/* One scene of a batch (device array of these; no counterpart in the
 * reference, which solves one scene per solver).  Particles of scene s are the
 * contiguous range [particle_offset, particle_offset + particle_count) of the
 * batch arrays; its buckets are [bucket_base, bucket_base + dims[0]*dims[1]*
 * dims[2]) of one segmented table.  With dims[a] = 128 or dims[a] >= the
 * scene's cell span + 4, a scene's candidate sets and visit order equal those
 * of a plain 128^3 grid, so each scene evolves exactly as a plain solve. */
typedef struct fs_scene {
    int32_t particle_offset;
    int32_t particle_count;
    int32_t bucket_base;
    int32_t dims[3];
    float smoothing_length;       /* h: query radius and cell width        */
    float cell_width_inv;         /* float32 1/h, as the plain build forms */
    float particle_radius;
    float bound[3];
    float boundary_damping;
    float dt;                     /* this scene's substep; 0 leaves the
                                     scene's state, rho, pressure and
                                     acceleration as they are               */
    int32_t clamp_negative_pressure;
    int32_t pad_;
    double cfl_smoothing_length;  /* h as the host holds it (float64), for the
                                     CFL limits of fs_batch_cfl_dt; 0: use
                                     smoothing_length                        */
} fs_scene;

typedef struct fs_batch_io {
    const float* x_in;
    const float* v_in;
    float* x_out;
    float* v_out;
    float* rho;
    float* pressure;
    float* accel;
    const int32_t* scene_of;      /* [n] scene id of each particle          */
    int32_t* scene_sorted;        /* [n] scratch: scene ids in sorted order */
} fs_batch_io;

/* One WCSPH substep of every scene of the batch: one segmented grid build
 * and the two neighbour passes of fs_wcsph_step, each a single launch over
 * all particles of all scenes.  `grid` must be initialised with dim_x = total
 * buckets, dim_y = dim_z = 1. */
int32_t fs_batch_wcsph_step(fs_grid* grid, fs_work* work, const fs_model* model, const fs_scene* scenes,
                            int32_t num_scenes, const fs_batch_io* io, fs_stream_t stream);

/* Per-scene CFL substep of the dam-break driver (simulation.py:166-187), on
 * the device, in float64 as the reference's host code evaluates it:
 *   limit = min(frame_dt/substeps, 0.25h/(c0+vmax), 0.25 sqrt(h/amax) if
 *           amax > 0, 0.025 h^2/viscosity if viscosity > 0),
 *   c0 = sqrt(stiffness*exponent/rest_density),
 *   scenes[s].dt = min(remaining[s], limit) while remaining[s] > 1e-12
 *   (else 0), remaining[s] -= dt,
 * with h the scene's cfl_smoothing_length (float64, as the host's
 * SolverWCSPH.smoothing_length) and vmax/amax the scene's maxima of |v| / |a|
 * (a ignored when use_accel is
 * 0).  status[0] = 1 if a maximum is non-finite, 2 if a scene's dt falls
 * below 1e-10 (the reference's RuntimeError cases).  scratch: device uint32
 * [2*num_scenes], zero on first use and left zero.  max_scene_particles (the
 * largest particle_count in the table; the table itself is device memory)
 * sizes the reduction. */
int32_t fs_batch_cfl_dt(fs_scene* scenes, int32_t num_scenes, int32_t max_scene_particles, const float* v,
                        const float* a, int32_t use_accel, double stiffness, double exponent, double rest_density,
                        double viscosity, double frame_dt, int32_t substeps, double* remaining, uint32_t* scratch,
                        int32_t* status, fs_stream_t stream);

/* The DFSPH dt policy of the dam-break driver (simulation.py:166-187) per
 * scene, on the device: limit = frame_dt/substeps, lowered to
 * 0.4 spacing/vmax when vmax > 0; scenes[s].dt = min(remaining[s], limit)
 * while remaining[s] > 1e-12 (else 0), remaining[s] -= dt.  status and
 * scratch as for fs_batch_cfl_dt. */
int32_t fs_dfsph_cfl_dt(fs_scene* scenes, int32_t num_scenes, int32_t max_scene_particles, const float* v,
                        double spacing, double frame_dt, int32_t substeps, double* remaining, uint32_t* scratch,
                        int32_t* status, fs_stream_t stream);

/* fs_batch_cfl_dt without the scratch argument; it allocates nothing.  No
 * counterpart in the reference: it is for callers that cannot allocate device
 * memory (code under HIP graph capture, serving loops, harnesses that forbid
 * hipMalloc).  Its scratch is a fixed device array inside the library, one
 * per device, for at most 4096 scenes (32 KiB); more scenes return
 * FS_ERR_CAPACITY.  The array is cleared by a kernel at the start of each
 * call.  Calls that overlap on one device (concurrent streams, or several
 * host threads) share that array and race: serialise them.  The stream must
 * belong to the current device. */
int32_t fs_batch_cfl_dt_noalloc(fs_scene* scenes, int32_t num_scenes, int32_t max_scene_particles, const float* v,
                                const float* a, int32_t use_accel, double stiffness, double exponent,
                                double rest_density, double viscosity, double frame_dt, int32_t substeps,
                                double* remaining, int32_t* status, fs_stream_t stream);

/* ------------------------------------------------------------------------ */
/* Driver diagnostics  (examples/dam_break/simulation.py:19-53)           */
/* ------------------------------------------------------------------------ */

/* maxima[0] = max |v|^2 over nv points, maxima[1] = max |a|^2 over na points
 * (non-finite -> +inf); maxima is device float[2], overwritten. */
int32_t fs_norm_maxima(const float* v, int32_t nv, const float* a, int32_t na, float* maxima, fs_stream_t stream);

/* invalid[0] = 1 if any x or v component (or density, first nd) is
 * non-finite, else 0; invalid is device int32[1], overwritten. */
int32_t fs_check_finite(const float* x, const float* v, int32_t n, const float* density, int32_t nd,
                        int32_t* invalid, fs_stream_t stream);

#ifdef __cplusplus
}
#endif

#endif /* FLUIDSIM_H */
