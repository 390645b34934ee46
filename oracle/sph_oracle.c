/*
 * CPU oracle for the SPH particle step -- TEST INFRASTRUCTURE ONLY.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load this library,
 * and only as the checker / CPU baseline; the product path never calls it.
 *
 * Exports (see oracle/oracle.py for the Python wrapper):
 *   orc_step_f64   FP64 oracle of one step (neighbour sets from the FP32 predicate)
 *   orc_viscosity_f64 / _f32   density and viscous acceleration of one state (test hook)
 *   orc_step_f32   same-order FP32 restatement of the gfx950 kernels (bit-exact target;
 *                  also the CPU baseline, OpenMP over particles)
 *   orc_neighbors_brute   O(N^2) neighbour sets under the FP32 predicate
 *   orc_sorted_order      fluid ids in cell-sorted order (stable by key, then id)
 *
 * Pinned by tests/test_oracle.py (closed forms, invariants, brute force, linear algebra);
 * each function's header comment names the reference lines it follows.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

typedef struct {
    int n, n_worlds;
    const float* q0;       /* model positions (boundary positions are taken from here) */
    const float* qd0;
    const float* mass;
    const float* inv_mass;
    const int* flags;
    const int* world;      /* NULL = world 0 */
    const float* gravity;  /* n_worlds x 3 */
    float v_max;
    const double* support;          /* n_worlds */
    const double* boundary_volume;  /* n_worlds */
    const double* domain_lo;        /* n_worlds x 3 */
    const double* domain_hi;
    double rest_density, viscosity, omega, tolerance;
    int min_iters, max_iters, max_neighbors, warm_start;
} orc_problem;

typedef struct {
    float lo[3];
    float inv_w;
    int dim[3];
    int cell_off;
} orc_gw;

typedef struct {
    int nf, nb, cap, overflow, max_count;
    int* fid;       /* fluid ids (ascending) */
    int* fworld;    /* world per fluid-list entry */
    int* fperm;     /* sorted fluid s -> fluid-list index */
    int* fworld_s;  /* world per sorted fluid */
    int* is_fluid;  /* per particle */
    int* nbr;       /* s * cap + k -> sorted fluid index */
    int* ncnt;
    float* bxs;     /* sorted boundary positions */
    int* bstart;    /* CSR of boundary neighbours per sorted fluid */
    int* blist;
    int* w_fstart;
    int* w_nf;
} orc_layout;

/* ---- cell geometry: identical FP32 operations to the gfx950 code --------------------- */

static int cell_raw(float x, float lo, float inv_w) {
    float f = floorf((x - lo) * inv_w);
    f = fminf(fmaxf(f, -1.0e9f), 1.0e9f);
    return (int)f;
}

static int clampi(int c, int dim) { return c < 0 ? 0 : (c >= dim ? dim - 1 : c); }

/* Cells are split into ORC_XSUB slabs along x for sorting: the solver's order is
 * (world, cz, cy, cx, slab, id) with slab coordinate clamp(floor(XSUB (x - lo) inv_w)). */
#define ORC_XSUB_BITS 2
#define ORC_XSUB (1 << ORC_XSUB_BITS)

static uint32_t key_of(const orc_gw* g, float x, float y, float z) {
    float f = floorf(((x - g->lo[0]) * g->inv_w) * (float)ORC_XSUB);
    f = fminf(fmaxf(f, -1.0e9f), 1.0e9f);
    int gx = (int)f;
    const int top = g->dim[0] * ORC_XSUB - 1;
    gx = gx < 0 ? 0 : (gx > top ? top : gx);
    const int cy = clampi(cell_raw(y, g->lo[1], g->inv_w), g->dim[1]);
    const int cz = clampi(cell_raw(z, g->lo[2], g->inv_w), g->dim[2]);
    return (uint32_t)(g->cell_off + (cz * g->dim[1] + cy) * g->dim[0]) * ORC_XSUB + (uint32_t)gx;
}

static void axis_range(float x, float r, float lo, float inv_w, int dim, int* c0, int* n) {
    int a = clampi(cell_raw(x - r, lo, inv_w), dim);
    int b = clampi(cell_raw(x + r, lo, inv_w), dim);
    *c0 = a;
    *n = b - a + 1;
}

typedef struct {
    uint32_t key;
    int idx;
} kv_t;

static int kv_cmp(const void* a, const void* b) {
    const kv_t* x = (const kv_t*)a;
    const kv_t* y = (const kv_t*)b;
    if (x->key != y->key) return x->key < y->key ? -1 : 1;
    return x->idx < y->idx ? -1 : (x->idx > y->idx);
}

/* Per-world cell boxes: same double formulas and float roundings as the solver set-up. */
static size_t make_grids(const orc_problem* pb, orc_gw* gws) {
    size_t off = 0;
    for (int w = 0; w < pb->n_worlds; ++w) {
        const double H = pb->support[w];
        const double cw = H * (1.0 + 1.0 / 128.0);
        for (int a = 0; a < 3; ++a) {
            const double lo = pb->domain_lo[3 * w + a], hi = pb->domain_hi[3 * w + a];
            double d = ceil((hi - lo) / cw) + 1.0;
            if (d < 3.0) d = 3.0;
            gws[w].lo[a] = (float)lo;
            gws[w].dim[a] = (int)d;
        }
        gws[w].inv_w = (float)(1.0 / cw);
        gws[w].cell_off = (int)off;
        off += (size_t)gws[w].dim[0] * gws[w].dim[1] * gws[w].dim[2];
    }
    return off;
}

/* Hash, stably sort and tabulate; visits candidates in the same order as grid_visit. */
static int orc_layout_build(const orc_problem* pb, const float* q_in, orc_layout* L) {
    memset(L, 0, sizeof(*L));
    const int N = pb->n, W = pb->n_worlds;
    L->cap = pb->max_neighbors;
    L->is_fluid = (int*)calloc((size_t)N + 1, sizeof(int));
    L->fid = (int*)malloc(sizeof(int) * ((size_t)N + 1));
    L->fworld = (int*)malloc(sizeof(int) * ((size_t)N + 1));
    int* bid = (int*)malloc(sizeof(int) * ((size_t)N + 1));
    int* bworld = (int*)malloc(sizeof(int) * ((size_t)N + 1));
    L->w_nf = (int*)calloc((size_t)W, sizeof(int));
    L->w_fstart = (int*)calloc((size_t)W, sizeof(int));
    int nf = 0, nb = 0;
    for (int i = 0; i < N; ++i) {
        int w = pb->world ? pb->world[i] : 0;
        if (w < 0) w = 0;
        if (w >= W) return 1;
        const int active = (pb->flags[i] & 1) != 0;
        if (active && pb->inv_mass[i] > 0.f) {
            L->fid[nf] = i;
            L->fworld[nf] = w;
            L->is_fluid[i] = 1;
            L->w_nf[w]++;
            nf++;
        } else if (active) {
            bid[nb] = i;
            bworld[nb] = w;
            nb++;
        }
    }
    L->nf = nf;
    L->nb = nb;
    for (int w = 1; w < W; ++w) L->w_fstart[w] = L->w_fstart[w - 1] + L->w_nf[w - 1];

    orc_gw* gws = (orc_gw*)malloc(sizeof(orc_gw) * (size_t)W);
    const size_t cells = make_grids(pb, gws);
    int* ft = (int*)calloc(2 * (cells + 1), sizeof(int));
    int* bt = (int*)calloc(2 * (cells + 1), sizeof(int));

    /* fluid */
    kv_t* kv = (kv_t*)malloc(sizeof(kv_t) * ((size_t)(nf > nb ? nf : nb) + 1));
    for (int k = 0; k < nf; ++k) {
        const int id = L->fid[k];
        kv[k].key = key_of(&gws[L->fworld[k]], q_in[3 * id], q_in[3 * id + 1], q_in[3 * id + 2]);
        kv[k].idx = k;
    }
    qsort(kv, (size_t)nf, sizeof(kv_t), kv_cmp);
    L->fperm = (int*)malloc(sizeof(int) * ((size_t)nf + 1));
    L->fworld_s = (int*)malloc(sizeof(int) * ((size_t)nf + 1));
    float* fx = (float*)malloc(sizeof(float) * 3 * ((size_t)nf + 1));
    for (int s = 0; s < nf; ++s) {
        L->fperm[s] = kv[s].idx;
        L->fworld_s[s] = L->fworld[kv[s].idx];
        const int id = L->fid[kv[s].idx];
        fx[3 * s] = q_in[3 * id];
        fx[3 * s + 1] = q_in[3 * id + 1];
        fx[3 * s + 2] = q_in[3 * id + 2];
        const uint32_t cell = kv[s].key >> ORC_XSUB_BITS;
        if (s == 0 || (kv[s - 1].key >> ORC_XSUB_BITS) != cell) ft[2 * cell] = s;
        if (s == nf - 1 || (kv[s + 1].key >> ORC_XSUB_BITS) != cell) ft[2 * cell + 1] = s + 1;
    }
    /* boundary (positions of the model at creation) */
    for (int k = 0; k < nb; ++k) {
        const int id = bid[k];
        kv[k].key = key_of(&gws[bworld[k]], pb->q0[3 * id], pb->q0[3 * id + 1], pb->q0[3 * id + 2]);
        kv[k].idx = k;
    }
    qsort(kv, (size_t)nb, sizeof(kv_t), kv_cmp);
    L->bxs = (float*)malloc(sizeof(float) * 3 * ((size_t)nb + 1));
    for (int s = 0; s < nb; ++s) {
        const int id = bid[kv[s].idx];
        L->bxs[3 * s] = pb->q0[3 * id];
        L->bxs[3 * s + 1] = pb->q0[3 * id + 1];
        L->bxs[3 * s + 2] = pb->q0[3 * id + 2];
        const uint32_t cell = kv[s].key >> ORC_XSUB_BITS;
        if (s == 0 || (kv[s - 1].key >> ORC_XSUB_BITS) != cell) bt[2 * cell] = s;
        if (s == nb - 1 || (kv[s + 1].key >> ORC_XSUB_BITS) != cell) bt[2 * cell + 1] = s + 1;
    }
    free(kv);

    /* neighbour visits */
    const int cap = L->cap;
    L->nbr = (int*)malloc(sizeof(int) * ((size_t)nf * cap + 1));
    L->ncnt = (int*)calloc((size_t)nf + 1, sizeof(int));
    L->bstart = (int*)calloc((size_t)nf + 2, sizeof(int));
    size_t bcap = (size_t)nf * 16 + 16, bn = 0;
    L->blist = (int*)malloc(sizeof(int) * bcap);
    for (int pass = 0; pass < 2; ++pass) {
        const int* tab = pass == 0 ? ft : bt;
        const float* px = pass == 0 ? fx : L->bxs;
        for (int s = 0; s < nf; ++s) {
            const orc_gw* g = &gws[L->fworld_s[s]];
            const double H = pb->support[L->fworld_s[s]];
            const float Hf = (float)H, H2 = (float)(H * H);
            const float qx = fx[3 * s], qy = fx[3 * s + 1], qz = fx[3 * s + 2];
            const float rp = Hf + Hf * (1.0f / 256.0f);
            int x0, nx, y0, ny, z0, nz;
            axis_range(qx, rp, g->lo[0], g->inv_w, g->dim[0], &x0, &nx);
            axis_range(qy, rp, g->lo[1], g->inv_w, g->dim[1], &y0, &ny);
            axis_range(qz, rp, g->lo[2], g->inv_w, g->dim[2], &z0, &nz);
            int c = 0;
            if (pass == 1) L->bstart[s] = (int)bn;
            for (int iz = 0; iz < nz; ++iz)
                for (int iy = 0; iy < ny; ++iy) {
                    const int row = g->cell_off + ((z0 + iz) * g->dim[1] + (y0 + iy)) * g->dim[0];
                    for (int ix = 0; ix < nx; ++ix) {
                        const int cell = row + x0 + ix;
                        for (int j = tab[2 * cell]; j < tab[2 * cell + 1]; ++j) {
                            const float dx = qx - px[3 * j], dy = qy - px[3 * j + 1], dz = qz - px[3 * j + 2];
                            const float d2 = (dx * dx + dy * dy) + dz * dz;
                            if (!(d2 < H2)) continue;
                            if (pass == 0) {
                                if (j == s) continue;
                                if (c < cap) L->nbr[(size_t)s * cap + c] = j;
                                ++c;
                            } else {
                                if (bn == bcap) {
                                    bcap *= 2;
                                    L->blist = (int*)realloc(L->blist, sizeof(int) * bcap);
                                }
                                L->blist[bn++] = j;
                            }
                        }
                    }
                }
            if (pass == 0) {
                if (c > cap) L->overflow = 1;
                if (c > L->max_count) L->max_count = c;
                L->ncnt[s] = c < cap ? c : cap;
            }
        }
        if (pass == 1) L->bstart[nf] = (int)bn;
    }
    free(fx);
    free(ft);
    free(bt);
    free(gws);
    free(bid);
    free(bworld);
    return 0;
}

static void orc_layout_free(orc_layout* L) {
    free(L->fid); free(L->fworld); free(L->fperm); free(L->fworld_s); free(L->is_fluid);
    free(L->nbr); free(L->ncnt); free(L->bxs); free(L->bstart); free(L->blist);
    free(L->w_fstart); free(L->w_nf);
    memset(L, 0, sizeof(*L));
}

/* Sum of e over one world's sorted run in the fixed order of the gfx950 reduction:
 * chunks of 256 reduced by the LDS tree (offsets 128..1), then chunk partials summed by
 * 64 lanes (lane l takes chunks l, l+64, ...) and the lanes by the tree 32..1. */
static double orc_world_reduce(const double* e, int start, int n) {
    const int nch = (n + 255) / 256;
    double lanes[64];
    for (int l = 0; l < 64; ++l) lanes[l] = 0.0;
    double buf[256];
    for (int c = 0; c < nch; ++c) {
        for (int t = 0; t < 256; ++t) {
            const int r = c * 256 + t;
            buf[t] = r < n ? e[start + r] : 0.0;
        }
        for (int off = 128; off > 0; off >>= 1)
            for (int t = 0; t < off; ++t) buf[t] += buf[t + off];
        lanes[c % 64] += buf[0];
    }
    for (int off = 32; off > 0; off >>= 1)
        for (int l = 0; l < off; ++l) lanes[l] += lanes[l + off];
    return lanes[0];
}

/* ---- precision instances ------------------------------------------------------------- */

#define REAL float
#define SUFFIX _f32
#define RSQRT sqrtf
#define RMAX fmaxf
#define RABS fabsf
#define ORC_IS_F32 1
#include "sph_oracle_impl.h"
#undef REAL
#undef SUFFIX
#undef RSQRT
#undef RMAX
#undef RABS
#undef ORC_IS_F32

#define REAL double
#define SUFFIX _f64
#define RSQRT sqrt
#define RMAX fmax
#define RABS fabs
#define ORC_IS_F32 0
#include "sph_oracle_impl.h"
#undef REAL
#undef SUFFIX
#undef RSQRT
#undef RMAX
#undef RABS
#undef ORC_IS_F32

/* ---- exported -------------------------------------------------------------------------- */

/* One SPH step (same contract as nsph_solver_step plus dumps).
 *   f_in: NULL, or the caller's force per particle (state_in.particle_f, N x 3), added to the
 *         SPH force before integration (integrate_particles, solver.py:24-46, integrates the
 *         total force of particle_f)
 *   pprev: previous-step pressure in fluid-list order (ascending fluid id), in/out
 *   forced_iters: NULL, or per world the exact number of Jacobi updates (-1 = stop test)
 *   returns 0, 3 (neighbour overflow) or 6 (a world hit max_iters). */
int orc_step_f32(const orc_problem* pb, const float* q_in, const float* qd_in, const float* f_in,
                 float* pprev, float dt, const int* forced_iters, float* q_out, float* qd_out, float* f_out,
                 float* rho_out, float* p_out, int* nbr_count_out, int* iters_out,
                 double* err_out, int nthreads) {
    return orc_step_impl_f32(pb, q_in, qd_in, f_in, pprev, dt, forced_iters, q_out, qd_out, f_out,
                             rho_out, p_out, nbr_count_out, iters_out, err_out, nthreads);
}

int orc_step_f64(const orc_problem* pb, const float* q_in, const float* qd_in, const float* f_in,
                 float* pprev, float dt, const int* forced_iters, float* q_out, float* qd_out, float* f_out,
                 float* rho_out, float* p_out, int* nbr_count_out, int* iters_out,
                 double* err_out, int nthreads) {
    return orc_step_impl_f64(pb, q_in, qd_in, f_in, pprev, dt, forced_iters, q_out, qd_out, f_out,
                             rho_out, p_out, nbr_count_out, iters_out, err_out, nthreads);
}

/* Cubic spline of support H at r^2 (FP64): W and (dW/dr)/r (test hook). */
void orc_kernel_f64(double H, double r2, double* W, double* gf) {
    phys_t_f64 P;
    const float g[3] = {0.f, 0.f, 0.f};
    make_phys_f64(&P, H, 0.0, 1000.0, 0.0, g);
    kern_wg_f64(&P, r2, W, gf);
}

/* Density and viscous acceleration of one state, particle order (see orc_viscosity_impl). */
int orc_viscosity_f64(const orc_problem* pb, const float* q_in, const float* qd_in, double* av, double* rho) {
    return orc_viscosity_impl_f64(pb, q_in, qd_in, av, rho);
}

int orc_viscosity_f32(const orc_problem* pb, const float* q_in, const float* qd_in, double* av, double* rho) {
    return orc_viscosity_impl_f32(pb, q_in, qd_in, av, rho);
}

/* IISPH operator pieces in FP64 (see orc_operator_impl). */
int orc_operator_f64(const orc_problem* pb, const float* q_in, const float* qd_in, float dt,
                     const double* p_in, double* rho, double* aii, double* s, double* Ap,
                     double* ap) {
    return orc_operator_impl_f64(pb, q_in, qd_in, dt, p_in, rho, aii, s, Ap, ap);
}

/* O(N^2) neighbour sets: for active i, all active j != i with FP32
 * (dx*dx + dy*dy) + dz*dz < H2, d = x_i - x_j, in ascending j.  Lists row-major i*cap + k.
 * Returns 3 when a list exceeds cap (counts hold the true sizes). */
int orc_neighbors_brute(const float* x, const int* flags, int n, float H2, int cap, int* counts,
                        int* idx, int nthreads) {
    int overflow = 0;
#pragma omp parallel for num_threads(nthreads) schedule(dynamic, 64) reduction(| : overflow)
    for (int i = 0; i < n; ++i) {
        int c = 0;
        if (!flags || (flags[i] & 1)) {
            const float xi = x[3 * i], yi = x[3 * i + 1], zi = x[3 * i + 2];
            for (int j = 0; j < n; ++j) {
                if (j == i || (flags && !(flags[j] & 1))) continue;
                const float dx = xi - x[3 * j], dy = yi - x[3 * j + 1], dz = zi - x[3 * j + 2];
                const float d2 = (dx * dx + dy * dy) + dz * dz;
                if (d2 < H2) {
                    if (c < cap) idx[(size_t)i * cap + c] = j;
                    ++c;
                }
            }
        }
        counts[i] = c;
        if (c > cap) overflow = 1;
    }
    return overflow ? 3 : 0;
}

/* Fluid particle ids in the solver's cell-sorted order, and the fluid/boundary counts. */
int orc_sorted_order(const orc_problem* pb, const float* q_in, int* ids_out, int* nf_out, int* nb_out) {
    orc_layout L;
    int rc = orc_layout_build(pb, q_in, &L);
    if (rc) return rc;
    for (int s = 0; s < L.nf; ++s) ids_out[s] = L.fid[L.fperm[s]];
    *nf_out = L.nf;
    *nb_out = L.nb;
    orc_layout_free(&L);
    return 0;
}

/* Fluid-fluid neighbour lists as the solver builds them (particle ids, slot-major
 * idx[k * n + id] like nsph_solver_neighbors). */
int orc_solver_neighbors(const orc_problem* pb, const float* q_in, int* counts, int* idx) {
    orc_layout L;
    int rc = orc_layout_build(pb, q_in, &L);
    if (rc) return rc;
    for (int i = 0; i < pb->n; ++i) counts[i] = 0;
    for (int s = 0; s < L.nf; ++s) {
        const int id = L.fid[L.fperm[s]];
        counts[id] = L.ncnt[s];
        for (int k = 0; k < L.ncnt[s]; ++k)
            idx[(size_t)k * pb->n + id] = L.fid[L.fperm[L.nbr[(size_t)s * L.cap + k]]];
    }
    rc = L.overflow ? 3 : 0;
    orc_layout_free(&L);
    return rc;
}

int orc_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
