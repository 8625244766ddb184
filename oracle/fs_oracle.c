/*
 * fs_oracle.c -- CPU oracle for the particle-fluid hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (fluid_simulation_amd/)
 * links, loads or calls this file; only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg use it, as the checker.
 *
 * It is a plain-C restatement of the reference's Warp kernels and of the Warp
 * HashGrid semantics those kernels rely on.  Every function cites the
 * reference file:line it follows.  Reference files (SParksLz/fluid_simulation):
 *   solver/sph/solver_wcsph.py, solver/sph/solver_dfsph.py,
 *   solver/pbf/solver_pbf.py, examples/dam_break/simulation.py.
 *
 * Build: twice, with ORACLE_REAL=float (the FP32 restatement the GPU path is
 * compared with) and ORACLE_REAL=double (the FP64 error reference).  Always
 * with -ffp-contract=off: every expression is evaluated as written, in the
 * reference's left-to-right order, with IEEE single or double rounding.
 *
 * Arithmetic conventions (the reference runs Warp kernels in float32):
 *   - literals are rounded to R (Warp float literals are float32);
 *   - wp.pi is the R nearest pi;
 *   - wp.length(v) = sqrt((x*x + y*y) + z*z);
 *   - wp.max(a,b) = a > b ? a : b, wp.min(a,b) = a < b ? a : b;
 *   - vec3 / scalar divides each component;
 *   - wp.pow(a,b) is evaluated as (R)pow((double)a,(double)b).
 *
 * Hash grid (Warp HashGrid, third-party, restated from its documented
 * behaviour and its call sites, e.g. solver_wcsph.py:398-402,456,134,143):
 *   cell coordinate c = round(x * inv)  (round = floor by default; the
 *     truncating variant is selectable with round_mode=1);
 *   bucket axis  b = max(c + 2^20, 0) mod dim;
 *   bucket id    = bz*(dx*dy) + by*dx + bx;
 *   build        = stable sort of (bucket, original index); cell_starts /
 *     cell_ends tables zeroed then filled from sorted-key transitions;
 *   query(p, r)  = per axis cells round((p-r)*inv) .. min(round((p+r)*inv),
 *     start+dim-1); visit x fastest, then y, then z; yield every point of each
 *     visited bucket in sorted order -- self included, aliased buckets
 *     included, no distance test.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifndef ORACLE_REAL
#define ORACLE_REAL float
#endif
typedef ORACLE_REAL R;

#define RC(v) ((R)(v))

#ifdef _OPENMP
#include <omp.h>
#endif

static inline R rsqrt_(R v) { return (sizeof(R) == 4) ? (R)sqrtf((float)v) : (R)sqrt((double)v); }
static inline R rfloor_(R v) { return (sizeof(R) == 4) ? (R)floorf((float)v) : (R)floor((double)v); }
static inline R rtrunc_(R v) { return (sizeof(R) == 4) ? (R)truncf((float)v) : (R)trunc((double)v); }
static inline R rmax(R a, R b) { return a > b ? a : b; }
static inline R rmin(R a, R b) { return a < b ? a : b; }
static inline R rpow(R a, R b) { return (R)pow((double)a, (double)b); }
static inline R rpi(void) { return (R)3.14159265358979323846; }

typedef struct { R x, y, z; } v3;
static inline v3 V(R x, R y, R z) { v3 r = {x, y, z}; return r; }
static inline v3 vload(const R* a, int i) { return V(a[3 * i], a[3 * i + 1], a[3 * i + 2]); }
static inline void vstore(R* a, int i, v3 v) { a[3 * i] = v.x; a[3 * i + 1] = v.y; a[3 * i + 2] = v.z; }
static inline v3 vadd(v3 a, v3 b) { return V(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline v3 vsub(v3 a, v3 b) { return V(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline v3 smul(R s, v3 a) { return V(s * a.x, s * a.y, s * a.z); }
static inline v3 muls(v3 a, R s) { return V(a.x * s, a.y * s, a.z * s); }
static inline v3 vdivs(v3 a, R s) { return V(a.x / s, a.y / s, a.z / s); }
static inline R vdot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline R vlen(v3 a) { return rsqrt_(vdot(a, a)); }
static const v3 VZERO = {0, 0, 0};

/* ParticleFlags.ACTIVE is bit 0 in Newton.  solver_wcsph.py:53-60 */
static inline int is_dynamic(int flags, int mask) { return (flags & 1) != 0 && mask == 0; }
static inline int is_active(int flags) { return (flags & 1) != 0; }

/* ------------------------------------------------------------------------ */
/* Hash grid                                                                */
/* ------------------------------------------------------------------------ */

typedef struct {
    int dim_x, dim_y, dim_z;
    int round_mode;        /* 0 floor, 1 truncate toward zero */
    int num_points;
    R cell_width;
    R cell_width_inv;
    const int* point_ids;  /* [num_points]  sorted slot -> original index */
    const int* cell_starts;/* [dim^3] */
    const int* cell_ends;  /* [dim^3] */
} ora_grid;

static inline int cell_coord(R v, int round_mode)
{
    R f = round_mode ? rtrunc_(v) : rfloor_(v);
    /* keep the integer conversion defined for coordinates far outside */
    if (f > RC(1.0e9)) f = RC(1.0e9);
    if (f < RC(-1.0e9)) f = RC(-1.0e9);
    if (!(f == f)) f = RC(0.0); /* NaN coordinate -> cell 0 */
    return (int)f;
}

static inline int hash_axis(int c, int dim)
{
    c += (1 << 20);
    if (c < 0) c = 0;
    return c % dim;
}

static inline int hash_index(int x, int y, int z, int dx, int dy, int dz)
{
    return hash_axis(z, dz) * (dx * dy) + hash_axis(y, dy) * dx + hash_axis(x, dx);
}

/* Bucket of every point, in original order.  Warp HashGrid.build, called at
 * solver_wcsph.py:456, solver_dfsph.py:624,783, solver_pbf.py:737. */
void ora_grid_cells(int n, const R* x, R radius, int dx, int dy, int dz, int round_mode, int* cell_of)
{
    R inv = RC(1.0) / radius;
    for (int i = 0; i < n; ++i) {
        int cx = cell_coord(x[3 * i] * inv, round_mode);
        int cy = cell_coord(x[3 * i + 1] * inv, round_mode);
        int cz = cell_coord(x[3 * i + 2] * inv, round_mode);
        cell_of[i] = hash_index(cx, cy, cz, dx, dy, dz);
    }
}

/* Full build: cell ids, stable (bucket, index) sort, start/end tables.
 * point_cells receives the sorted keys (Warp keeps keys sorted in place). */
int ora_grid_build(int n, const R* x, R radius, int dx, int dy, int dz, int round_mode,
                   int* cell_of, int* point_ids, int* point_cells, int* cell_starts, int* cell_ends)
{
    const int m = dx * dy * dz;
    ora_grid_cells(n, x, radius, dx, dy, dz, round_mode, cell_of);
    int* cnt = (int*)calloc((size_t)m + 1, sizeof(int));
    if (!cnt) return -1;
    for (int i = 0; i < n; ++i) cnt[cell_of[i] + 1]++;
    for (int c = 0; c < m; ++c) cnt[c + 1] += cnt[c];
    for (int i = 0; i < n; ++i) point_ids[cnt[cell_of[i]]++] = i; /* ascending i: stable */
    free(cnt);
    memset(cell_starts, 0, sizeof(int) * (size_t)m);
    memset(cell_ends, 0, sizeof(int) * (size_t)m);
    for (int k = 0; k < n; ++k) point_cells[k] = cell_of[point_ids[k]];
    for (int k = 0; k < n; ++k) {
        int c = point_cells[k];
        if (k == 0) {
            cell_starts[c] = 0;
        } else {
            int p = point_cells[k - 1];
            if (c != p) {
                cell_starts[c] = k;
                cell_ends[p] = k;
            }
        }
        if (k == n - 1) cell_ends[c] = k + 1;
    }
    return 0;
}

typedef struct { int xs, ys, zs, xe, ye, ze; } ora_range;

static inline ora_range query_range(const ora_grid* g, v3 p, R radius)
{
    ora_range q;
    R inv = g->cell_width_inv;
    q.xs = cell_coord((p.x - radius) * inv, g->round_mode);
    q.ys = cell_coord((p.y - radius) * inv, g->round_mode);
    q.zs = cell_coord((p.z - radius) * inv, g->round_mode);
    q.xe = cell_coord((p.x + radius) * inv, g->round_mode);
    q.ye = cell_coord((p.y + radius) * inv, g->round_mode);
    q.ze = cell_coord((p.z + radius) * inv, g->round_mode);
    if (q.xe > q.xs + g->dim_x - 1) q.xe = q.xs + g->dim_x - 1;
    if (q.ye > q.ys + g->dim_y - 1) q.ye = q.ys + g->dim_y - 1;
    if (q.ze > q.zs + g->dim_z - 1) q.ze = q.zs + g->dim_z - 1;
    return q;
}

/* for (j in hash_grid_query(grid, P, RADIUS)) BODY -- a `continue` inside
 * BODY skips to the next candidate, as in the reference's loops. */
#define FOR_CANDIDATES(G, P, RADIUS, J, ...)                                               \
    do {                                                                                   \
        ora_range q_ = query_range((G), (P), (RADIUS));                                    \
        for (int z_ = q_.zs; z_ <= q_.ze; ++z_)                                            \
            for (int y_ = q_.ys; y_ <= q_.ye; ++y_)                                        \
                for (int x_ = q_.xs; x_ <= q_.xe; ++x_) {                                  \
                    int c_ = hash_index(x_, y_, z_, (G)->dim_x, (G)->dim_y, (G)->dim_z);   \
                    for (int k_ = (G)->cell_starts[c_]; k_ < (G)->cell_ends[c_]; ++k_) {   \
                        int J = (G)->point_ids[k_];                                        \
                        __VA_ARGS__                                                        \
                    }                                                                      \
                }                                                                          \
    } while (0)

/* Candidate list of each listed query point, in visit order (CSR).  With
 * radius_test < 0 every candidate is listed; otherwise only j whose distance
 * passes the cubic-spline support test fl(sqrt(d2))/h <= 1 (mode 0) or the
 * strict PBF test fl(sqrt(d2)) < h (mode 1).  offsets has n_query+1 entries;
 * call with indices=NULL to size. */
long long ora_grid_candidates(const ora_grid* g, const R* x, int n_query, const int* query_ids,
                              R radius, int within_mode, long long* offsets, int* indices)
{
    long long total = 0;
    for (int t = 0; t < n_query; ++t) {
        int i = query_ids ? query_ids[t] : t;
        v3 xi = vload(x, i);
        if (offsets) offsets[t] = total;
        FOR_CANDIDATES(g, xi, radius, j, {
            if (within_mode >= 0) {
                R r = vlen(vsub(xi, vload(x, j)));
                if (within_mode == 0 && !(r / radius <= RC(1.0))) continue;
                if (within_mode == 1 && !(r < radius)) continue;
            }
            if (indices) indices[total] = j;
            total++;
        });
    }
    if (offsets) offsets[n_query] = total;
    return total;
}

/* ------------------------------------------------------------------------ */
/* Kernel functions                                                         */
/* ------------------------------------------------------------------------ */

/* get_cubic  solver_wcsph.py:20-33 (= solver_dfsph.py:24-37) */
static inline R cubic_w(R r_norm, R radius)
{
    R res = RC(0.0);
    R h = radius;
    R k = RC(8.0) / (rpi() * (h * h * h));
    R q = r_norm / h;
    if (q <= RC(1.0)) {
        if (q <= RC(0.5)) {
            R q2 = q * q;
            R q3 = q2 * q;
            res = k * (RC(6.0) * q3 - RC(6.0) * q2 + RC(1.0));
        } else {
            R t = RC(1.0) - q;
            res = RC(2.0) * k * (t * t * t);
        }
    }
    return res;
}

/* get_cubic_derivative  solver_wcsph.py:36-50 (= solver_dfsph.py:40-54) */
static inline v3 cubic_grad(v3 r, R smoothing_length)
{
    v3 res = VZERO;
    R r_norm = vlen(r);
    R h = smoothing_length;
    R k = RC(8.0) / (rpi() * (h * h * h));
    R q = r_norm / h;
    if (r_norm > RC(1.0e-5) && q <= RC(1.0)) {
        v3 grad_q = vdivs(r, r_norm * h);
        if (q <= RC(0.5)) {
            res = smul(RC(6.0) * k * q * (RC(3.0) * q - RC(2.0)), grad_q);
        } else {
            R q_term = RC(1.0) - q;
            res = smul(RC(-6.0) * k * q_term * q_term, grad_q);
        }
    }
    return res;
}

/* PBF get_cubic = poly6  solver_pbf.py:23-32 */
static inline R poly6_w(R r_norm, R radius)
{
    R res = RC(0.0);
    R h = radius;
    if (h > RC(0.0) && r_norm < h) {
        R h2 = h * h;
        R rhs = h2 - r_norm * r_norm;
        R h3 = h * h * h;
        R k = RC(315.0) / (RC(64.0) * rpi() * (h3 * h3 * h3));
        res = k * (rhs * rhs * rhs);
    }
    return res;
}

/* PBF get_cubic_derivative = spiky gradient  solver_pbf.py:35-45 */
static inline v3 spiky_grad(v3 r, R smoothing_length)
{
    v3 res = VZERO;
    R r_norm = vlen(r);
    R h = smoothing_length;
    if (h > RC(0.0) && r_norm > RC(1.0e-6) && r_norm < h) {
        R h3 = h * h * h;
        R h6 = h3 * h3;
        R coeff = RC(-45.0) / (rpi() * h6);
        R term = (h - r_norm) * (h - r_norm) / r_norm;
        res = smul(coeff * term, r);
    }
    return res;
}

/* cal_acc_with_non_pressure  solver_wcsph.py:63-89 (= solver_dfsph.py:67-93) */
static inline v3 acc_non_pressure(v3 a, R rho_nei, v3 vel, v3 vel_nei, R mass, R mass_nei, R mu, R gamma,
                                  v3 d, R e_dist, R distance, R h)
{
    a = vsub(a, muls(smul(gamma / mass * mass_nei, d), cubic_w(e_dist, h)));
    R v_current_nei = vdot(vsub(vel, vel_nei), d);
    R s = RC(10.0) * mu * (mass_nei / rmax(rho_nei, RC(1.0e-6))) * v_current_nei
          / (distance * distance + RC(0.01) * (h * h));
    a = vadd(a, smul(s, cubic_grad(d, h)));
    return a;
}

/* PBF cal_acc_with_non_pressure  solver_pbf.py:58-87 (skip branches) */
static inline v3 pbf_acc_non_pressure(v3 a, R rho_nei, v3 vel, v3 vel_nei, R mass, R mass_nei, R mu, R gamma,
                                      v3 d, R e_dist, R distance, R h)
{
    if (gamma != RC(0.0)) a = vsub(a, muls(smul(gamma / mass * mass_nei, d), poly6_w(e_dist, h)));
    if (mu != RC(0.0)) {
        R v_current_nei = vdot(vsub(vel, vel_nei), d);
        if (v_current_nei != RC(0.0)) {
            R s = RC(10.0) * mu * (mass_nei / rmax(rho_nei, RC(1.0e-6))) * v_current_nei
                  / (distance * distance + RC(0.01) * (h * h));
            a = vadd(a, smul(s, spiky_grad(d, h)));
        }
    }
    return a;
}

/* cal_acc_with_pressure  solver_wcsph.py:92-106 */
static inline v3 acc_pressure(v3 a, v3 d, R p, R p_nei, R rho, R rho_nei, R mass_nei, R h)
{
    R dp_i = p / rmax(rho * rho, RC(1.0e-6));
    R dp_nei = p_nei / rmax(rho_nei * rho_nei, RC(1.0e-6));
    return vadd(a, smul(-mass_nei * (dp_i + dp_nei), cubic_grad(d, h)));
}

/* compute_scorr  solver_pbf.py:90-101 */
static inline R compute_scorr(R r_norm, R h, R scorr_k, R scorr_n, R w_delta_q)
{
    if (scorr_k == RC(0.0) || w_delta_q <= RC(1.0e-6)) return RC(0.0);
    R ratio = poly6_w(r_norm, h) / w_delta_q;
    return -scorr_k * rpow(ratio, scorr_n);
}

/* exported scalar kernels, for closed-form checks */
R ora_cubic_w(R r, R h) { return cubic_w(r, h); }
void ora_cubic_grad(const R* r, R h, R* out) { vstore(out, 0, cubic_grad(vload(r, 0), h)); }
R ora_poly6_w(R r, R h) { return poly6_w(r, h); }
void ora_spiky_grad(const R* r, R h, R* out) { vstore(out, 0, spiky_grad(vload(r, 0), h)); }

#define PARFOR _Pragma("omp parallel for schedule(static)")

/* ------------------------------------------------------------------------ */
/* WCSPH  (solver/sph/solver_wcsph.py)                                      */
/* ------------------------------------------------------------------------ */

/* compute_particle_volume  solver_wcsph.py:109-116 (= dfsph :96-103, pbf :104-111) */
void ora_particle_volume(int n, const R* mass, const R* rest_density, R* volume)
{
    for (int t = 0; t < n; ++t) volume[t] = mass[t] / rmax(rest_density[t], RC(1.0e-6));
}

/* compute_density_pressure  solver_wcsph.py:119-158, for particle i */
static inline void wcsph_density_one(const ora_grid* g, int i, const int* flags, const R* x, const R* rest_density,
                                     const R* stiffness, const R* exponent, const R* volume, R h,
                                     int clamp_negative_pressure, R* rho, R* pressure)
{
    if (!is_active(flags[i])) { rho[i] = RC(0.0); pressure[i] = RC(0.0); return; }
    v3 xi = vload(x, i);
    R rho_temp = RC(0.0);
    FOR_CANDIDATES(g, xi, h, j, {
        if (!is_active(flags[j])) continue;
        v3 dist = vsub(xi, vload(x, j));
        R mass_nei = rest_density[j] * volume[j];
        rho_temp += mass_nei * cubic_w(vlen(dist), h);
    });
    R rho_i = rho_temp;
    rho[i] = rho_i;
    R p = stiffness[i] * (rpow(rho_i / rmax(rest_density[i], RC(1.0e-6)), exponent[i]) - RC(1.0));
    if (clamp_negative_pressure) p = rmax(p, RC(0.0));
    pressure[i] = p;
}

void ora_wcsph_density_pressure(const ora_grid* g, const int* flags, const R* x, const R* rest_density,
                                const R* stiffness, const R* exponent, const R* volume, R h,
                                int clamp_negative_pressure, R* rho, R* pressure)
{
    PARFOR
    for (int t = 0; t < g->num_points; ++t)
        wcsph_density_one(g, g->point_ids[t], flags, x, rest_density, stiffness, exponent, volume, h,
                          clamp_negative_pressure, rho, pressure);
}

/* The same for the listed particles only (rho/pressure written at those
 * indices): per-particle parity checks of large inputs. */
void ora_wcsph_density_pressure_ids(const ora_grid* g, int n_query, const int* ids, const int* flags, const R* x,
                                    const R* rest_density, const R* stiffness, const R* exponent, const R* volume,
                                    R h, int clamp_negative_pressure, R* rho, R* pressure)
{
    PARFOR
    for (int t = 0; t < n_query; ++t)
        wcsph_density_one(g, ids[t], flags, x, rest_density, stiffness, exponent, volume, h, clamp_negative_pressure,
                          rho, pressure);
}

/* acceleration_wcsph  solver_wcsph.py:161-228, for particle i */
static inline void wcsph_accel_one(const ora_grid* g, int i, const int* flags, const int* mask, const R* x,
                                   const R* v, const R* rho, const R* pressure, R particle_radius, const R* mass,
                                   const R* surface_tension, const R* viscosity, const R* gravity, const int* world,
                                   R h, R* accel)
{
    if (!is_dynamic(flags[i], mask[i])) { vstore(accel, i, VZERO); return; }
    v3 xi = vload(x, i);
    R rho_i = rho[i];
    R p_i = pressure[i];
    v3 acc = VZERO;
    FOR_CANDIDATES(g, xi, h, j, {
        if (j == i || !is_active(flags[j])) continue;
        v3 direction = vsub(xi, vload(x, j));
        R distance = vlen(direction);
        R mass_nei = mass[j];
        R rho_nei = rho[j];
        R p_nei = pressure[j];
        acc = acc_non_pressure(acc, rho_nei, vload(v, i), vload(v, j), mass[i], mass_nei, viscosity[i],
                               surface_tension[i], direction, rmax(distance, particle_radius), distance, h);
        acc = acc_pressure(acc, direction, p_i, p_nei, rho_i, rho_nei, mass_nei, h);
    });
    int w = world[i] > 0 ? world[i] : 0;
    vstore(accel, i, vadd(acc, vload(gravity, w)));
}

void ora_wcsph_acceleration(const ora_grid* g, const int* flags, const int* mask, const R* x, const R* v,
                            const R* rho, const R* pressure, R particle_radius, const R* mass,
                            const R* surface_tension, const R* viscosity, const R* gravity,
                            const int* world, R h, R* accel)
{
    PARFOR
    for (int t = 0; t < g->num_points; ++t)
        wcsph_accel_one(g, g->point_ids[t], flags, mask, x, v, rho, pressure, particle_radius, mass, surface_tension,
                        viscosity, gravity, world, h, accel);
}

/* The same for the listed particles only. */
void ora_wcsph_acceleration_ids(const ora_grid* g, int n_query, const int* ids, const int* flags, const int* mask,
                                const R* x, const R* v, const R* rho, const R* pressure, R particle_radius,
                                const R* mass, const R* surface_tension, const R* viscosity, const R* gravity,
                                const int* world, R h, R* accel)
{
    PARFOR
    for (int t = 0; t < n_query; ++t)
        wcsph_accel_one(g, ids[t], flags, mask, x, v, rho, pressure, particle_radius, mass, surface_tension,
                        viscosity, gravity, world, h, accel);
}

/* integrate_velocity  solver_wcsph.py:231-244 */
void ora_wcsph_integrate_velocity(int n, const int* flags, const int* mask, const R* v_in, const R* a, R dt, R* v_out)
{
    for (int t = 0; t < n; ++t) {
        if (!is_dynamic(flags[t], mask[t])) { vstore(v_out, t, vload(v_in, t)); continue; }
        vstore(v_out, t, vadd(vload(v_in, t), smul(dt, vload(a, t))));
    }
}

/* integrate_positions  solver_wcsph.py:247-260 (= dfsph :340-353 with v_star) */
void ora_integrate_positions(int n, const int* flags, const int* mask, const R* x_in, const R* v, R dt, R* x_out)
{
    for (int t = 0; t < n; ++t) {
        if (!is_dynamic(flags[t], mask[t])) { vstore(x_out, t, vload(x_in, t)); continue; }
        vstore(x_out, t, vadd(vload(x_in, t), muls(vload(v, t), dt)));
    }
}

/* apply_bounds  solver_wcsph.py:263-305 (outward_only=1);
 *               solver_dfsph.py:406-442 (outward_only=0) */
void ora_sph_apply_bounds(int n, const int* flags, const int* mask, R* x, R* v, R sx, R sy, R sz,
                          R damping, int outward_only)
{
    for (int t = 0; t < n; ++t) {
        if (!is_dynamic(flags[t], mask[t])) continue;
        v3 p = vload(x, t), u = vload(v, t);
        if (p.x < -sx) { p.x = -sx; if (!outward_only || u.x < RC(0.0)) u.x = u.x * damping; }
        if (p.x > sx) { p.x = sx; if (!outward_only || u.x > RC(0.0)) u.x = u.x * damping; }
        if (p.y < -sy) { p.y = -sy; if (!outward_only || u.y < RC(0.0)) u.y = u.y * damping; }
        if (p.y > sy) { p.y = sy; if (!outward_only || u.y > RC(0.0)) u.y = u.y * damping; }
        if (p.z < RC(0.0)) { p.z = RC(0.0); if (!outward_only || u.z < RC(0.0)) u.z = u.z * damping; }
        if (p.z > sz * RC(2.0)) { p.z = sz * RC(2.0); if (!outward_only || u.z > RC(0.0)) u.z = u.z * damping; }
        vstore(x, t, p);
        vstore(v, t, u);
    }
}

/* ------------------------------------------------------------------------ */
/* WCSPH adjoint (no counterpart in the reference, whose gradients would come */
/* from Warp's tape): reverse-mode derivative of one SolverWCSPH.step        */
/* (solver_wcsph.py:431-535) with surface_tension = 0, in gather form.       */
/* ------------------------------------------------------------------------ */

/* H(r) u, H = d(cubic_grad)/dr, on the branches of cubic_grad (zero where
 * cubic_grad is identically zero).  cubic_grad = g(d) r with
 *   q <= 0.5: g = 6k q(3q-2)/(d h),  g' = 18k/h^3
 *   q >  0.5: g = -6k (1-q)^2/(d h), g' = 6k/h (2(1-q)/(h d) + (1-q)^2/d^2)
 * and H u = g u + (g'/d)(r.u) r. */
static inline v3 cubic_hess_mul(v3 r, R r_norm, R h, v3 u)
{
    R k = RC(8.0) / (rpi() * (h * h * h));
    R q = r_norm / h;
    if (!(r_norm > RC(1.0e-5) && q <= RC(1.0))) return VZERO;
    R g, gpd;
    if (q <= RC(0.5)) {
        g = RC(6.0) * k * q * (RC(3.0) * q - RC(2.0)) / (r_norm * h);
        gpd = RC(18.0) * k / (h * h * h) / r_norm;
    } else {
        R t = RC(1.0) - q;
        g = RC(-6.0) * k * t * t / (r_norm * h);
        gpd = RC(6.0) * k / h * (RC(2.0) * t / (h * r_norm) + t * t / (r_norm * r_norm)) / r_norm;
    }
    return vadd(smul(g, u), smul(gpd * vdot(r, u), r));
}

/* Pass A: reverse of integrate_velocity, integrate_positions and the
 * outward-damping apply_bounds (solver_wcsph.py:231-305).  Writes the
 * direct parts of xbar_in / vbar_in and abar = dL/da. */
void ora_wcsph_adj_integrate(int n, const int* flags, const int* mask, const R* x_in, const R* v_in, const R* accel,
                             R dt, R sx, R sy, R sz, R damping, const R* xbar_out, const R* vbar_out,
                             R* xbar_in, R* vbar_in, R* abar)
{
    for (int t = 0; t < n; ++t) {
        v3 xb = vload(xbar_out, t), vb = vload(vbar_out, t);
        if (!is_dynamic(flags[t], mask[t])) {
            vstore(xbar_in, t, xb);
            vstore(vbar_in, t, vb);
            vstore(abar, t, VZERO);
            continue;
        }
        v3 u = vadd(vload(v_in, t), smul(dt, vload(accel, t)));
        v3 p = vadd(vload(x_in, t), muls(u, dt));
        R lo[3] = {-sx, -sy, RC(0.0)}, hi[3] = {sx, sy, sz * RC(2.0)};
        R pc[3] = {p.x, p.y, p.z}, uc[3] = {u.x, u.y, u.z};
        R xbc[3] = {xb.x, xb.y, xb.z}, vbc[3] = {vb.x, vb.y, vb.z};
        for (int a = 0; a < 3; ++a) {
            if (pc[a] < lo[a]) {
                xbc[a] = RC(0.0);
                if (uc[a] < RC(0.0)) vbc[a] = vbc[a] * damping;
            } else if (pc[a] > hi[a]) {
                xbc[a] = RC(0.0);
                if (uc[a] > RC(0.0)) vbc[a] = vbc[a] * damping;
            }
        }
        v3 xbu = V(xbc[0], xbc[1], xbc[2]);
        v3 vbu = vadd(V(vbc[0], vbc[1], vbc[2]), smul(dt, xbu));
        vstore(xbar_in, t, xbu);
        vstore(vbar_in, t, vbu);
        vstore(abar, t, smul(dt, vbu));
    }
}

/* Pass B: reverse of acceleration_wcsph (pressure + viscosity, gamma = 0) and
 * of the EOS (solver_wcsph.py:92-106,119-228), gathered per particle k over
 * the pairs (k,j) [k's own acceleration] and (j,k) [k as a neighbour].
 * Accumulates into xbar_in / vbar_in and grad_stiffness / grad_viscosity,
 * writes rhobar (dL/drho_k). */
void ora_wcsph_adj_force(const ora_grid* g, const int* flags, const R* x, const R* v,
                         const R* rho, const R* pressure, const R* mass, const R* rest_density, const R* stiffness,
                         const R* exponent, const R* viscosity, R h, int clamp_negative_pressure, const R* abar,
                         R* xbar_in, R* vbar_in, R* rhobar, R* grad_stiffness, R* grad_viscosity)
{
    PARFOR
    for (int t = 0; t < g->num_points; ++t) {
        int k = g->point_ids[t];
        if (!is_active(flags[k])) { rhobar[k] = RC(0.0); continue; }
        v3 xk = vload(x, k), vk = vload(v, k), ak = vload(abar, k);
        R rho_k = rho[k], p_k = pressure[k], m_k = mass[k];
        R P_k = p_k / rmax(rho_k * rho_k, RC(1.0e-6));
        R w_k = m_k / rmax(rho_k, RC(1.0e-6));
        R c_k = RC(10.0) * viscosity[k];
        R eps2 = RC(0.01) * (h * h);
        v3 xb = VZERO, vb = VZERO;
        R Pb = RC(0.0), wb = RC(0.0), mub = RC(0.0);
        FOR_CANDIDATES(g, xk, h, j, {
            if (j == k || !is_active(flags[j])) continue;
            v3 r = vsub(xk, vload(x, j));
            R d = vlen(r);
            if (!(d / h <= RC(1.0))) continue;
            v3 G = cubic_grad(r, h);
            v3 aj = vload(abar, j);
            R m_j = mass[j];
            R rho_j = rho[j];
            R P_j = pressure[j] / rmax(rho_j * rho_j, RC(1.0e-6));
            R w_j = m_j / rmax(rho_j, RC(1.0e-6));
            R c_j = RC(10.0) * viscosity[j];
            R den = d * d + eps2;
            v3 dv = vsub(vk, vload(v, j));
            R s = vdot(dv, r) / den;
            R Gak = vdot(G, ak), Gaj = vdot(G, aj);
            v3 Hak = cubic_hess_mul(r, d, h, ak), Haj = cubic_hess_mul(r, d, h, aj);
            v3 ds = vdivs(vsub(dv, smul(RC(2.0) * s, r)), den);
            /* pressure: f(k,j) = -m_j (P_k+P_j) G,  f(j,k) = +m_k (P_j+P_k) G */
            xb = vadd(xb, smul(-m_j * (P_k + P_j), Hak));
            xb = vadd(xb, smul(m_k * (P_j + P_k), Haj));
            Pb += -m_j * Gak + m_k * Gaj;
            /* viscosity: f(k,j) = c_k w_j s G,  f(j,k) = -c_j w_k s G */
            R ck_wj = c_k * w_j, cj_wk = c_j * w_k;
            xb = vadd(xb, smul(ck_wj, vadd(smul(Gak, ds), smul(s, Hak))));
            xb = vsub(xb, smul(cj_wk, vadd(smul(Gaj, ds), smul(s, Haj))));
            vb = vadd(vb, smul((ck_wj * Gak - cj_wk * Gaj) / den, r));
            mub += RC(10.0) * w_j * s * Gak;
            wb += -c_j * s * Gaj;
        });
        R rb = RC(0.0);
        R pb = Pb / rmax(rho_k * rho_k, RC(1.0e-6));
        if (rho_k * rho_k > RC(1.0e-6)) rb += Pb * (RC(-2.0) * p_k / (rho_k * rho_k * rho_k));
        if (rho_k > RC(1.0e-6)) rb += wb * (-m_k / (rho_k * rho_k));
        R rd = rmax(rest_density[k], RC(1.0e-6));
        R ratio = rho_k / rd;
        R gam = exponent[k];
        R p_raw = stiffness[k] * (rpow(ratio, gam) - RC(1.0));
        R sb = RC(0.0);
        if (!clamp_negative_pressure || p_raw > RC(0.0)) {
            rb += pb * stiffness[k] * gam * rpow(ratio, gam - RC(1.0)) / rd;
            sb = pb * (rpow(ratio, gam) - RC(1.0));
        }
        rhobar[k] = rb;
        vstore(xbar_in, k, vadd(vload(xbar_in, k), xb));
        vstore(vbar_in, k, vadd(vload(vbar_in, k), vb));
        grad_stiffness[k] += sb;
        grad_viscosity[k] += mub;
    }
}

/* Pass C: reverse of compute_density_pressure's density sum
 * (solver_wcsph.py:119-158): xbar_k += sum_j (rhobar_k wv_j + rhobar_j wv_k) G. */
void ora_wcsph_adj_density(const ora_grid* g, const int* flags, const R* x, const R* rest_density,
                           const R* volume, R h, const R* rhobar, R* xbar_in)
{
    PARFOR
    for (int t = 0; t < g->num_points; ++t) {
        int k = g->point_ids[t];
        if (!is_active(flags[k])) continue;
        v3 xk = vload(x, k);
        R wv_k = rest_density[k] * volume[k];
        R rb_k = rhobar[k];
        v3 xb = VZERO;
        FOR_CANDIDATES(g, xk, h, j, {
            if (j == k || !is_active(flags[j])) continue;
            v3 r = vsub(xk, vload(x, j));
            R d = vlen(r);
            if (!(d / h <= RC(1.0))) continue;
            R wv_j = rest_density[j] * volume[j];
            xb = vadd(xb, smul(rb_k * wv_j + rhobar[j] * wv_k, cubic_grad(r, h)));
        });
        vstore(xbar_in, k, vadd(vload(xbar_in, k), xb));
    }
}

/* ------------------------------------------------------------------------ */
/* DFSPH  (solver/sph/solver_dfsph.py)                                      */
/* ------------------------------------------------------------------------ */

/* compute_density  solver_dfsph.py:106-133 (poly6=0); solver_pbf.py:114-141 (poly6=1) */
void ora_density(const ora_grid* g, const int* flags, const R* x, const R* rest_density, const R* volume,
                 R h, int poly6, R* rho)
{
    PARFOR
    for (int t = 0; t < g->num_points; ++t) {
        int i = g->point_ids[t];
        if (!is_active(flags[i])) { rho[i] = RC(0.0); continue; }
        v3 xi = vload(x, i);
        R rho_temp = RC(0.0);
        FOR_CANDIDATES(g, xi, h, j, {
            if (!is_active(flags[j])) continue;
            v3 r = vsub(xi, vload(x, j));
            R mass_j = rest_density[j] * volume[j];
            R w = poly6 ? poly6_w(vlen(r), h) : cubic_w(vlen(r), h);
            rho_temp += mass_j * w;
        });
        rho[i] = rho_temp;
    }
}

/* compute_factor_dfsph  solver_dfsph.py:136-174 */
void ora_dfsph_factor(const ora_grid* g, const int* flags, const R* x, const R* mass, R h,
                      int min_neighbors_for_factor, R* factor, int* neighbor_count)
{
    PARFOR
    for (int t = 0; t < g->num_points; ++t) {
        int i = g->point_ids[t];
        if (!is_active(flags[i])) { factor[i] = RC(0.0); neighbor_count[i] = 0; continue; }
        v3 xi = vload(x, i);
        v3 sum_grad = VZERO;
        R sum_grad_sq = RC(0.0);
        int count = 0;
        FOR_CANDIDATES(g, xi, h, j, {
            if (!is_active(flags[j])) continue;
            if (j != i) count += 1;
            v3 grad = smul(mass[j], cubic_grad(vsub(xi, vload(x, j)), h));
            sum_grad = vadd(sum_grad, grad);
            sum_grad_sq += vdot(grad, grad);
        });
        neighbor_count[i] = count;
        R denom = vdot(sum_grad, sum_grad) + sum_grad_sq;
        if (count < min_neighbors_for_factor || denom < RC(1.0e-6)) factor[i] = RC(0.0);
        else factor[i] = RC(1.0) / denom;
    }
}

/* acceleration_non_pressure  solver_dfsph.py:177-226 (cubic, gravity);
 *                            solver_pbf.py:144-189 (poly6/spiky, skip branches, no gravity) */
void ora_accel_non_pressure(const ora_grid* g, const int* flags, const int* mask, const R* x, const R* v,
                            const R* rho, R particle_radius, const R* mass, const R* surface_tension,
                            const R* viscosity, const R* gravity, const int* world, R h, int pbf, R* accel)
{
    PARFOR
    for (int t = 0; t < g->num_points; ++t) {
        int i = g->point_ids[t];
        if (!is_dynamic(flags[i], mask[i])) { vstore(accel, i, VZERO); continue; }
        v3 xi = vload(x, i);
        v3 acc = VZERO;
        FOR_CANDIDATES(g, xi, h, j, {
            if (j == i || !is_active(flags[j])) continue;
            v3 d = vsub(xi, vload(x, j));
            R dist = vlen(d);
            if (pbf)
                acc = pbf_acc_non_pressure(acc, rho[j], vload(v, i), vload(v, j), mass[i], mass[j], viscosity[i],
                                           surface_tension[i], d, rmax(dist, particle_radius), dist, h);
            else
                acc = acc_non_pressure(acc, rho[j], vload(v, i), vload(v, j), mass[i], mass[j], viscosity[i],
                                       surface_tension[i], d, rmax(dist, particle_radius), dist, h);
        });
        if (pbf) {
            vstore(accel, i, acc);
        } else {
            int w = world[i] > 0 ? world[i] : 0;
            vstore(accel, i, vadd(acc, vload(gravity, w)));
        }
    }
}

/* predict_velocity  solver_dfsph.py:229-242 */
void ora_dfsph_predict_velocity(int n, const int* flags, const int* mask, const R* v, const R* a, R dt, R* v_star)
{
    for (int t = 0; t < n; ++t) {
        if (!is_dynamic(flags[t], mask[t])) { vstore(v_star, t, VZERO); continue; }
        vstore(v_star, t, vadd(vload(v, t), smul(dt, vload(a, t))));
    }
}

/* compute_density_change_rate  solver_dfsph.py:245-272 */
void ora_dfsph_density_change_rate(const ora_grid* g, const int* flags, const R* x, const R* v, const R* mass,
                                   R h, R* d_rho_dt)
{
    PARFOR
    for (int t = 0; t < g->num_points; ++t) {
        int i = g->point_ids[t];
        if (!is_active(flags[i])) { d_rho_dt[i] = RC(0.0); continue; }
        v3 xi = vload(x, i);
        v3 vi = vload(v, i);
        R div_val = RC(0.0);
        FOR_CANDIDATES(g, xi, h, j, {
            if (!is_active(flags[j])) continue;
            v3 grad_w = cubic_grad(vsub(xi, vload(x, j)), h);
            div_val += mass[j] * vdot(vsub(vi, vload(v, j)), grad_w);
        });
        d_rho_dt[i] = div_val;
    }
}

/* predict_density_star  solver_dfsph.py:275-287 */
void ora_dfsph_predict_density_star(int n, const int* flags, const R* rho, const R* d_rho_dt, R dt, R* rho_star)
{
    for (int t = 0; t < n; ++t) {
        if (!is_active(flags[t])) { rho_star[t] = RC(0.0); continue; }
        rho_star[t] = rho[t] + dt * d_rho_dt[t];
    }
}

/* compute_kappa_density  solver_dfsph.py:290-306 */
void ora_dfsph_kappa_density(int n, const int* flags, const int* mask, const R* rho_star, const R* rest_density,
                             const R* factor, R dt, R max_kappa, R* kappa)
{
    for (int t = 0; t < n; ++t) {
        if (!is_dynamic(flags[t], mask[t])) { kappa[t] = RC(0.0); continue; }
        R rho_err = rmax(rho_star[t] - rest_density[t], RC(0.0));
        kappa[t] = rmin((rho_err / (dt * dt)) * factor[t], max_kappa);
    }
}

/* apply_constant_density_correction  solver_dfsph.py:309-337;
 * apply_divergence_free_correction   solver_dfsph.py:375-403 (same body) */
void ora_dfsph_correction(const ora_grid* g, const int* flags, const int* mask, const R* x, R* v_star,
                          const R* mass, const R* rho, const R* kappa, R h, R dt)
{
    R* dv_all = (R*)malloc(sizeof(R) * 3 * (size_t)(g->num_points > 0 ? g->num_points : 1));
    int* upd = (int*)calloc((size_t)(g->num_points > 0 ? g->num_points : 1), sizeof(int));
    /* v_star is only written at i and never read across particles, so the
     * update is applied after the gather (identical to the in-place kernel). */
    PARFOR
    for (int t = 0; t < g->num_points; ++t) {
        int i = g->point_ids[t];
        if (!is_dynamic(flags[i], mask[i])) continue;
        v3 xi = vload(x, i);
        v3 dv = VZERO;
        FOR_CANDIDATES(g, xi, h, j, {
            if (!is_active(flags[j])) continue;
            R coeff = mass[j] * (kappa[i] / rmax(rho[i], RC(1.0e-6)) + kappa[j] / rmax(rho[j], RC(1.0e-6)));
            dv = vadd(dv, smul(coeff, cubic_grad(vsub(xi, vload(x, j)), h)));
        });
        vstore(dv_all, i, dv);
        upd[i] = 1;
    }
    for (int i = 0; i < g->num_points; ++i)
        if (upd[i]) vstore(v_star, i, vsub(vload(v_star, i), smul(dt, vload(dv_all, i))));
    free(dv_all);
    free(upd);
}

/* compute_kappa_divergence  solver_dfsph.py:356-372 */
void ora_dfsph_kappa_divergence(int n, const int* flags, const int* mask, const R* d_rho_dt, const R* rho,
                                const R* factor, R dt, R max_kappa_v, R* kappa_v)
{
    for (int t = 0; t < n; ++t) {
        if (!is_dynamic(flags[t], mask[t])) { kappa_v[t] = RC(0.0); continue; }
        R density_change = rmax(d_rho_dt[t], RC(0.0));
        kappa_v[t] = rmin((RC(1.0) / dt) * density_change * rho[t] * factor[t], max_kappa_v);
    }
}

/* ------------------------------------------------------------------------ */
/* PBF  (solver/pbf/solver_pbf.py)                                          */
/* ------------------------------------------------------------------------ */

/* predict_velocity_with_gravity  solver_pbf.py:192-208 */
void ora_pbf_predict_velocity(int n, const int* flags, const int* mask, const R* v_in, const R* gravity,
                              const int* world, R dt, R* v_pred)
{
    for (int t = 0; t < n; ++t) {
        if (!is_dynamic(flags[t], mask[t])) { vstore(v_pred, t, vload(v_in, t)); continue; }
        int w = world[t] > 0 ? world[t] : 0;
        vstore(v_pred, t, vadd(vload(v_in, t), smul(dt, vload(gravity, w))));
    }
}

/* predict_positions  solver_pbf.py:211-224 */
void ora_pbf_predict_positions(int n, const int* flags, const int* mask, const R* x_in, const R* v_pred, R dt,
                               R* x_pred)
{
    for (int t = 0; t < n; ++t) {
        if (!is_dynamic(flags[t], mask[t])) { vstore(x_pred, t, vload(x_in, t)); continue; }
        vstore(x_pred, t, vadd(vload(x_in, t), smul(dt, vload(v_pred, t))));
    }
}

/* apply_non_pressure_prediction  solver_pbf.py:227-242 */
void ora_pbf_apply_non_pressure(int n, const int* flags, const int* mask, const R* a, R dt, R* v_pred, R* x_pred)
{
    for (int t = 0; t < n; ++t) {
        if (!is_dynamic(flags[t], mask[t])) continue;
        v3 delta_v = smul(dt, vload(a, t));
        vstore(v_pred, t, vadd(vload(v_pred, t), delta_v));
        vstore(x_pred, t, vadd(vload(x_pred, t), smul(dt, delta_v)));
    }
}

/* compute_pbf_lambda  solver_pbf.py:245-302 */
void ora_pbf_lambda(const ora_grid* g, const int* flags, const int* mask, const R* x_pred, const R* mass,
                    const R* rest_density, R h, R lambda_epsilon, int min_neighbors_for_lambda,
                    R* pbf_lambda, R* rho, int* neighbor_count)
{
    PARFOR
    for (int t = 0; t < g->num_points; ++t) {
        int i = g->point_ids[t];
        if (!is_dynamic(flags[i], mask[i])) {
            pbf_lambda[i] = RC(0.0); rho[i] = RC(0.0); neighbor_count[i] = 0; continue;
        }
        R rho0_i = rmax(rest_density[i], RC(1.0e-6));
        v3 xi = vload(x_pred, i);
        R rho_i = RC(0.0);
        v3 grad_i = VZERO;
        R sum_grad_sq = RC(0.0);
        int count = 0;
        FOR_CANDIDATES(g, xi, h, j, {
            if (!is_active(flags[j])) continue;
            v3 xij = vsub(xi, vload(x_pred, j));
            R r = vlen(xij);
            if (j != i && r < h) count += 1;
            R mass_j = mass[j];
            rho_i += mass_j * poly6_w(r, h);
            v3 grad_j = smul(-(mass_j / rho0_i), spiky_grad(xij, h));
            sum_grad_sq += vdot(grad_j, grad_j);
            grad_i = vsub(grad_i, grad_j);
        });
        sum_grad_sq += vdot(grad_i, grad_i);
        R c_i = rmax(rho_i / rho0_i - RC(1.0), RC(0.0));
        rho[i] = rho_i;
        neighbor_count[i] = count;
        if (min_neighbors_for_lambda > 0 && count < min_neighbors_for_lambda) pbf_lambda[i] = RC(0.0);
        else pbf_lambda[i] = -c_i / (sum_grad_sq + lambda_epsilon);
    }
}

/* compute_delta_position  solver_pbf.py:305-350 */
void ora_pbf_delta_position(const ora_grid* g, const int* flags, const int* mask, const R* x_pred, const R* mass,
                            const R* rest_density, const R* pbf_lambda, R h, R scorr_k, R scorr_n,
                            R scorr_w_delta_q, R max_delta_position, R* delta_p)
{
    PARFOR
    for (int t = 0; t < g->num_points; ++t) {
        int i = g->point_ids[t];
        if (!is_dynamic(flags[i], mask[i])) { vstore(delta_p, i, VZERO); continue; }
        v3 xi = vload(x_pred, i);
        R rho0_i = rmax(rest_density[i], RC(1.0e-6));
        v3 dp = VZERO;
        FOR_CANDIDATES(g, xi, h, j, {
            if (j == i || !is_active(flags[j])) continue;
            v3 xij = vsub(xi, vload(x_pred, j));
            v3 grad = spiky_grad(xij, h);
            R scorr = RC(0.0);
            if (scorr_k != RC(0.0)) scorr = compute_scorr(vlen(xij), h, scorr_k, scorr_n, scorr_w_delta_q);
            dp = vadd(dp, smul((pbf_lambda[i] + pbf_lambda[j] + scorr) * mass[j], grad));
        });
        dp = vdivs(dp, rho0_i);
        if (max_delta_position > RC(0.0)) {
            R dp_len = vlen(dp);
            if (dp_len > max_delta_position) dp = muls(dp, max_delta_position / rmax(dp_len, RC(1.0e-12)));
        }
        vstore(delta_p, i, dp);
    }
}

/* apply_delta_position  solver_pbf.py:353-363 */
void ora_pbf_apply_delta(int n, const int* flags, const int* mask, R* x_pred, const R* delta_p)
{
    for (int t = 0; t < n; ++t) {
        if (!is_dynamic(flags[t], mask[t])) continue;
        vstore(x_pred, t, vadd(vload(x_pred, t), vload(delta_p, t)));
    }
}

/* apply_position_bounds  solver_pbf.py:366-399 */
void ora_pbf_position_bounds(int n, const int* flags, const int* mask, R* x_pred, R sx, R sy, R sz,
                             R particle_radius, R boundary_epsilon)
{
    for (int t = 0; t < n; ++t) {
        if (!is_dynamic(flags[t], mask[t])) continue;
        v3 p = vload(x_pred, t);
        R x_min = -sx + particle_radius + boundary_epsilon;
        R x_max = sx - particle_radius - boundary_epsilon;
        R y_min = -sy + particle_radius + boundary_epsilon;
        R y_max = sy - particle_radius - boundary_epsilon;
        R z_min = particle_radius + boundary_epsilon;
        R z_max = sz * RC(2.0) - particle_radius - boundary_epsilon;
        if (p.x < x_min) p.x = x_min;
        if (p.x > x_max) p.x = x_max;
        if (p.y < y_min) p.y = y_min;
        if (p.y > y_max) p.y = y_max;
        if (p.z < z_min) p.z = z_min;
        if (p.z > z_max) p.z = z_max;
        vstore(x_pred, t, p);
    }
}

/* update_velocity_from_positions  solver_pbf.py:402-416 */
void ora_pbf_update_velocity(int n, const int* flags, const int* mask, const R* x_in, const R* x_pred,
                             const R* v_in, R dt, R* v_out)
{
    for (int t = 0; t < n; ++t) {
        if (!is_dynamic(flags[t], mask[t])) { vstore(v_out, t, vload(v_in, t)); continue; }
        vstore(v_out, t, vdivs(vsub(vload(x_pred, t), vload(x_in, t)), rmax(dt, RC(1.0e-8))));
    }
}

/* apply_xsph_viscosity  solver_pbf.py:434-463 */
void ora_pbf_xsph(const ora_grid* g, const int* flags, const int* mask, const R* x, const R* v_in, const R* mass,
                  const R* rho, R h, R xsph_c, R* v_out)
{
    PARFOR
    for (int t = 0; t < g->num_points; ++t) {
        int i = g->point_ids[t];
        if (!is_dynamic(flags[i], mask[i])) { vstore(v_out, i, vload(v_in, i)); continue; }
        v3 vi = vload(v_in, i);
        v3 xi = vload(x, i);
        v3 corr = VZERO;
        FOR_CANDIDATES(g, xi, h, j, {
            if (j == i || !is_active(flags[j])) continue;
            R w = poly6_w(vlen(vsub(xi, vload(x, j))), h);
            corr = vadd(corr, muls(smul(mass[j] / rmax(rho[j], RC(1.0e-6)), vsub(vload(v_in, j), vi)), w));
        });
        vstore(v_out, i, vadd(vi, smul(xsph_c, corr)));
    }
}

/* apply_bounds  solver_pbf.py:466-517 */
void ora_pbf_apply_bounds(int n, const int* flags, const int* mask, R* x, R* v, R sx, R sy, R sz,
                          R particle_radius, R boundary_epsilon, R damping)
{
    for (int t = 0; t < n; ++t) {
        if (!is_dynamic(flags[t], mask[t])) continue;
        v3 p = vload(x, t), u = vload(v, t);
        R x_min = -sx + particle_radius + boundary_epsilon;
        R x_max = sx - particle_radius - boundary_epsilon;
        R y_min = -sy + particle_radius + boundary_epsilon;
        R y_max = sy - particle_radius - boundary_epsilon;
        R z_min = particle_radius + boundary_epsilon;
        R z_max = sz * RC(2.0) - particle_radius - boundary_epsilon;
        if (p.x < x_min) { p.x = x_min; if (u.x < RC(0.0)) u.x = u.x * damping; }
        if (p.x > x_max) { p.x = x_max; if (u.x > RC(0.0)) u.x = u.x * damping; }
        if (p.y < y_min) { p.y = y_min; if (u.y < RC(0.0)) u.y = u.y * damping; }
        if (p.y > y_max) { p.y = y_max; if (u.y > RC(0.0)) u.y = u.y * damping; }
        if (p.z < z_min) { p.z = z_min; if (u.z < RC(0.0)) u.z = u.z * damping; }
        if (p.z > z_max) { p.z = z_max; if (u.z > RC(0.0)) u.z = u.z * damping; }
        vstore(x, t, p);
        vstore(v, t, u);
    }
}

/* ------------------------------------------------------------------------ */
/* Driver diagnostics  (examples/dam_break/simulation.py)                   */
/* ------------------------------------------------------------------------ */

/* particle_norm_maxima  simulation.py:19-36: max |v|^2, max |a|^2 with
 * non-finite values mapped to +inf */
void ora_norm_maxima(int nv, const R* v, int na, const R* a, R* maxima)
{
    R m0 = RC(0.0), m1 = RC(0.0);
    for (int i = 0; i < nv; ++i) {
        v3 u = vload(v, i);
        R u2 = vdot(u, u);
        if (!(isfinite(u.x) && isfinite(u.y) && isfinite(u.z) && isfinite(u2))) u2 = (R)INFINITY;
        if (u2 > m0) m0 = u2;
        if (i < na) {
            v3 b = vload(a, i);
            R b2 = vdot(b, b);
            if (!(isfinite(b.x) && isfinite(b.y) && isfinite(b.z) && isfinite(b2))) b2 = (R)INFINITY;
            if (b2 > m1) m1 = b2;
        }
    }
    maxima[0] = m0;
    maxima[1] = m1;
}

/* ------------------------------------------------------------------------ */
/* APIC particle separation                                                 */
/* ------------------------------------------------------------------------ */

/* accumulate_particle_separation  solver_apic.py:193-248.  One thread per
 * original index i; candidates of hash_grid_query(grid, x_i, query_radius)
 * in visit order; lo/hi are the box corners (float32 vec3 in the reference),
 * wall_band the into-wall kill band. */
void ora_separation_accumulate(const ora_grid* g, const int* flags, const R* x, R rest_dist, R query_radius,
                               R strength, const R* lo, const R* hi, R wall_band, R* delta)
{
    PARFOR
    for (int i = 0; i < g->num_points; ++i) {
        if (!is_active(flags[i])) {
            vstore(delta, i, VZERO);
            continue;
        }
        v3 xi = vload(x, i);
        v3 disp = VZERO;
        R count = RC(0.0);
        FOR_CANDIDATES(g, xi, query_radius, j, {
            if (j != i && is_active(flags[j])) {
                v3 rij = vsub(xi, vload(x, j));
                R d = vlen(rij);
                if (d > RC(1.0e-8) && d < rest_dist) {
                    v3 nrm = vdivs(rij, d);
                    disp = vadd(disp, muls(nrm, strength * (rest_dist - d)));
                    count += RC(1.0);
                }
            }
        });
        if (count > RC(0.0)) disp = vdivs(disp, count);
        else disp = VZERO;
        if (xi.x <= lo[0] + wall_band && disp.x < RC(0.0)) disp.x = RC(0.0);
        if (xi.x >= hi[0] - wall_band && disp.x > RC(0.0)) disp.x = RC(0.0);
        if (xi.y <= lo[1] + wall_band && disp.y < RC(0.0)) disp.y = RC(0.0);
        if (xi.y >= hi[1] - wall_band && disp.y > RC(0.0)) disp.y = RC(0.0);
        if (xi.z <= lo[2] + wall_band && disp.z < RC(0.0)) disp.z = RC(0.0);
        if (xi.z >= hi[2] - wall_band && disp.z > RC(0.0)) disp.z = RC(0.0);
        vstore(delta, i, disp);
    }
}

/* wp.clamp(x, a, b) = min(max(a, x), b) */
static inline R clamp_(R v, R a, R b) { return rmin(rmax(a, v), b); }

/* apply_particle_separation_delta  solver_apic.py:251-268 (in place). */
void ora_separation_apply(int n, const int* flags, R* x, const R* delta, const R* lo, const R* hi, R wall_margin)
{
    const v3 l = V(lo[0] + wall_margin, lo[1] + wall_margin, lo[2] + wall_margin);
    const v3 h = V(hi[0] - wall_margin, hi[1] - wall_margin, hi[2] - wall_margin);
    PARFOR
    for (int i = 0; i < n; ++i) {
        if (!is_active(flags[i])) continue;
        v3 p = vadd(vload(x, i), vload(delta, i));
        vstore(x, i, V(clamp_(p.x, l.x, h.x), clamp_(p.y, l.y, h.y), clamp_(p.z, l.z, h.z)));
    }
}

void ora_set_threads(int nthreads)
{
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
}

int ora_real_bytes(void) { return (int)sizeof(R); }
