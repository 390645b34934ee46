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
    double skin;             /* 0: lists searched every step; > 0: kept candidates          */
    int max_skin_neighbors;  /* capacity of the kept candidates                             */
} orc_problem;

typedef struct {
    float lo[3];
    float inv_w;  /* inverse cell width along x */
    float inv_t;  /* inverse cell height along z: 2 inv_w (half-height cells) */
    int dim[3];
    int cell_off;
} orc_gw;

typedef struct {
    int nf, nb, nbs, cap, overflow, max_count;  /* nb boundary entries: nbs static, then moving */
    int* fid;       /* fluid ids (ascending) */
    int* fworld;    /* world per fluid-list entry */
    int* fperm;     /* sorted fluid s -> fluid-list index */
    int* fworld_s;  /* world per sorted fluid */
    int* is_fluid;  /* per particle */
    int* nbr;       /* s * cap + k -> sorted fluid index */
    int* ncnt;
    float* bxs;     /* sorted boundary positions: static (model), then moving (state_in) */
    float* bvs;     /* their velocities (0 for static) */
    int* bstart;    /* CSR of boundary neighbours per sorted fluid */
    int* blist;
    int* w_fstart;
    int* w_nf;
    float* xbk;     /* search position of each fluid-list entry (x, y, z) */
} orc_layout;

/* ---- cell geometry: identical FP32 operations to the gfx950 code --------------------- */

static int cell_raw(float x, float lo, float inv_w) {
    float f = floorf((x - lo) * inv_w);
    f = fminf(fmaxf(f, -1.0e9f), 1.0e9f);
    return (int)f;
}

static int clampi(int c, int dim) { return c < 0 ? 0 : (c >= dim ? dim - 1 : c); }

/* Cells are w wide along x and y and w / 2 high along z, and split into ORC_XSUB slabs along
 * x for sorting: the solver's order is (world, cz, cy, cx, slab, id) with slab coordinate
 * clamp(floor(XSUB (x - lo) inv_w)), cy = clamp(floor((y - lo) inv_w)) and
 * cz = clamp(floor((z - lo) inv_t)). */
#define ORC_XSUB_BITS 2
#define ORC_XSUB (1 << ORC_XSUB_BITS)

static uint32_t key_of(const orc_gw* g, float x, float y, float z) {
    float f = floorf(((x - g->lo[0]) * g->inv_w) * (float)ORC_XSUB);
    f = fminf(fmaxf(f, -1.0e9f), 1.0e9f);
    int gx = (int)f;
    const int top = g->dim[0] * ORC_XSUB - 1;
    gx = gx < 0 ? 0 : (gx > top ? top : gx);
    const int cy = clampi(cell_raw(y, g->lo[1], g->inv_w), g->dim[1]);
    const int cz = clampi(cell_raw(z, g->lo[2], g->inv_t), g->dim[2]);
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
        const double cw = H * (1.0 + pb->skin) * (1.0 + 1.0 / 128.0);  /* the search radius */
        for (int a = 0; a < 3; ++a) {
            const double lo = pb->domain_lo[3 * w + a], hi = pb->domain_hi[3 * w + a];
            double d = ceil((hi - lo) / (a < 2 ? cw : 0.5 * cw)) + 1.0;
            if (d < 3.0) d = 3.0;
            gws[w].lo[a] = (float)lo;
            gws[w].dim[a] = (int)d;
        }
        gws[w].inv_w = (float)(1.0 / cw);
        gws[w].inv_t = 2.0f * gws[w].inv_w;
        gws[w].cell_off = (int)off;
        off += (size_t)gws[w].dim[0] * gws[w].dim[1] * gws[w].dim[2];
    }
    return off;
}

/* Squared rebuild distance and search radius of a world with support H (the same double
 * formulas and float roundings as the solver set-up): a pair within H now was within
 * Rs = H (1 + skin) of each other's search positions when each particle lies within
 * (Rs - H) / 2 of its own; thr keeps a 1e-4 relative margin for the FP32 roundings. */
static void skin_radii(double H, double skin, float* thr2, float* Rs, float* Rs2) {
    const double rs = H * (1.0 + skin);
    const double thr = 0.5 * (rs - H) * (1.0 - 1.0e-4);
    *thr2 = (float)(thr * thr);
    *Rs = (float)rs;
    *Rs2 = (float)(rs * rs);
}

/* Candidates of the sorted points pts (x, y, z) within r of point s, in the cell visiting
 * order of grid_visit (z, y, x over the clamped axis ranges, then sorted order inside a cell),
 * FP32 predicate d2 < r2 with d = pts[s] - pts[j]; j == s is skipped.  Writes at most cap ids
 * to out and returns the number found. */
static int scan_cells(const orc_gw* g, const int* tab, const float* pts, int s, float r, float r2, int cap,
                      int* out) {
    const float qx = pts[3 * s], qy = pts[3 * s + 1], qz = pts[3 * s + 2];
    const float rp = r + r * (1.0f / 256.0f);
    int x0, nx, y0, ny, z0, nz;
    axis_range(qx, rp, g->lo[0], g->inv_w, g->dim[0], &x0, &nx);
    axis_range(qy, rp, g->lo[1], g->inv_w, g->dim[1], &y0, &ny);
    axis_range(qz, rp, g->lo[2], g->inv_t, g->dim[2], &z0, &nz);
    int c = 0;
    for (int iz = 0; iz < nz; ++iz)
        for (int iy = 0; iy < ny; ++iy) {
            const int row = g->cell_off + ((z0 + iz) * g->dim[1] + (y0 + iy)) * g->dim[0];
            for (int ix = 0; ix < nx; ++ix) {
                const int cell = row + x0 + ix;
                for (int j = tab[2 * cell]; j < tab[2 * cell + 1]; ++j) {
                    if (j == s) continue;
                    const float dx = qx - pts[3 * j], dy = qy - pts[3 * j + 1], dz = qz - pts[3 * j + 2];
                    const float d2 = (dx * dx + dy * dy) + dz * dz;
                    if (!(d2 < r2)) continue;
                    if (c < cap) out[c] = j;
                    ++c;
                }
            }
        }
    return c;
}

/* Hash, stably sort and tabulate; visits candidates in the same order as grid_visit.
 *
 * Search positions (skin > 0): a world whose fluid particles all lie within the rebuild
 * distance of their xb_in entries (FP32 (dx dx + dy dy) + dz dz <= thr2, d = x - xb; NaN
 * fails) is sorted and searched around xb_in, any other world (and every world when xb_in is
 * NULL or skin = 0) around the current positions.  Keys, sort and cell table are those of the
 * search positions; the candidates of a particle are the points within Rs of its search
 * position (search positions of both), and its list the candidates within H of it now (current
 * positions, FP32 d2 < H^2), in candidate order, i.e. ascending sorted index. */
static int orc_layout_build(const orc_problem* pb, const float* q_in, const float* qd_in, const float* xb_in,
                            orc_layout* L) {
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
    /* zero-mass particles: static boundary when the model gives them no velocity (hashed from
     * the model's positions), moving boundary otherwise (hashed from state_in every step);
     * statics first, each group in ascending id */
    int nbs = 0;
    {
        int* mid = (int*)malloc(sizeof(int) * ((size_t)nb + 1));
        int* mworld = (int*)malloc(sizeof(int) * ((size_t)nb + 1));
        int nm = 0;
        for (int k = 0; k < nb; ++k) {
            const int id = bid[k];
            const int rest = pb->qd0[3 * id] == 0.f && pb->qd0[3 * id + 1] == 0.f && pb->qd0[3 * id + 2] == 0.f;
            if (rest) {
                bid[nbs] = id;
                bworld[nbs] = bworld[k];
                nbs++;
            } else {
                mid[nm] = id;
                mworld[nm] = bworld[k];
                nm++;
            }
        }
        for (int k = 0; k < nm; ++k) {
            bid[nbs + k] = mid[k];
            bworld[nbs + k] = mworld[k];
        }
        free(mid);
        free(mworld);
    }
    L->nf = nf;
    L->nb = nb;
    L->nbs = nbs;
    for (int w = 1; w < W; ++w) L->w_fstart[w] = L->w_fstart[w - 1] + L->w_nf[w - 1];

    orc_gw* gws = (orc_gw*)malloc(sizeof(orc_gw) * (size_t)W);
    const size_t cells = make_grids(pb, gws);
    int* ft = (int*)calloc(2 * (cells + 1), sizeof(int));
    int* bt = (int*)calloc(2 * (cells + 1), sizeof(int));
    int* kt = (int*)calloc(2 * (cells + 1), sizeof(int));

    /* search positions: per world, state_in's when every particle lies within the rebuild
     * distance of them, else the current ones */
    const int skin_on = pb->skin > 0.0;
    int* from_xb = (int*)calloc((size_t)W + 1, sizeof(int));
    if (skin_on && xb_in) {
        for (int w = 0; w < W; ++w) from_xb[w] = 1;
        for (int k = 0; k < nf; ++k) {
            const int id = L->fid[k], w = L->fworld[k];
            float thr2, Rs, Rs2;
            skin_radii(pb->support[w], pb->skin, &thr2, &Rs, &Rs2);
            const float dx = q_in[3 * id] - xb_in[3 * id], dy = q_in[3 * id + 1] - xb_in[3 * id + 1],
                        dz = q_in[3 * id + 2] - xb_in[3 * id + 2];
            const float d2 = (dx * dx + dy * dy) + dz * dz;
            if (!(d2 <= thr2)) from_xb[w] = 0;
        }
    }
    L->xbk = (float*)malloc(sizeof(float) * 3 * ((size_t)nf + 1));
    for (int k = 0; k < nf; ++k) {
        const int id = L->fid[k];
        const float* b = from_xb[L->fworld[k]] ? xb_in : q_in;
        L->xbk[3 * k] = b[3 * id];
        L->xbk[3 * k + 1] = b[3 * id + 1];
        L->xbk[3 * k + 2] = b[3 * id + 2];
    }
    free(from_xb);

    /* fluid */
    kv_t* kv = (kv_t*)malloc(sizeof(kv_t) * ((size_t)(nf > nb ? nf : nb) + 1));
    for (int k = 0; k < nf; ++k) {
        kv[k].key = key_of(&gws[L->fworld[k]], L->xbk[3 * k], L->xbk[3 * k + 1], L->xbk[3 * k + 2]);
        kv[k].idx = k;
    }
    qsort(kv, (size_t)nf, sizeof(kv_t), kv_cmp);
    L->fperm = (int*)malloc(sizeof(int) * ((size_t)nf + 1));
    L->fworld_s = (int*)malloc(sizeof(int) * ((size_t)nf + 1));
    float* fx = (float*)malloc(sizeof(float) * 3 * ((size_t)nf + 1));
    float* fb = (float*)malloc(sizeof(float) * 3 * ((size_t)nf + 1));
    for (int s = 0; s < nf; ++s) {
        L->fperm[s] = kv[s].idx;
        L->fworld_s[s] = L->fworld[kv[s].idx];
        const int id = L->fid[kv[s].idx];
        fx[3 * s] = q_in[3 * id];
        fx[3 * s + 1] = q_in[3 * id + 1];
        fx[3 * s + 2] = q_in[3 * id + 2];
        fb[3 * s] = L->xbk[3 * kv[s].idx];
        fb[3 * s + 1] = L->xbk[3 * kv[s].idx + 1];
        fb[3 * s + 2] = L->xbk[3 * kv[s].idx + 2];
        const uint32_t cell = kv[s].key >> ORC_XSUB_BITS;
        if (s == 0 || (kv[s - 1].key >> ORC_XSUB_BITS) != cell) ft[2 * cell] = s;
        if (s == nf - 1 || (kv[s + 1].key >> ORC_XSUB_BITS) != cell) ft[2 * cell + 1] = s + 1;
    }
    /* boundary: static from the model's positions, moving from state_in's, each group
     * stably sorted by key with its own cell table (entries relative to the group's start) */
    L->bxs = (float*)malloc(sizeof(float) * 3 * ((size_t)nb + 1));
    L->bvs = (float*)calloc(3 * ((size_t)nb + 1), sizeof(float));
    for (int grp = 0; grp < 2; ++grp) {
        const int first = grp == 0 ? 0 : nbs, cnt = grp == 0 ? nbs : nb - nbs;
        const float* src = grp == 0 ? pb->q0 : q_in;
        int* tab = grp == 0 ? bt : kt;
        for (int k = 0; k < cnt; ++k) {
            const int id = bid[first + k];
            kv[k].key = key_of(&gws[bworld[first + k]], src[3 * id], src[3 * id + 1], src[3 * id + 2]);
            kv[k].idx = k;
        }
        qsort(kv, (size_t)cnt, sizeof(kv_t), kv_cmp);
        for (int s = 0; s < cnt; ++s) {
            const int id = bid[first + kv[s].idx];
            for (int a = 0; a < 3; ++a) {
                L->bxs[3 * (first + s) + a] = src[3 * id + a];
                L->bvs[3 * (first + s) + a] = (grp == 1 && qd_in) ? qd_in[3 * id + a] : 0.f;
            }
            const uint32_t cell = kv[s].key >> ORC_XSUB_BITS;
            if (s == 0 || (kv[s - 1].key >> ORC_XSUB_BITS) != cell) tab[2 * cell] = s;
            if (s == cnt - 1 || (kv[s + 1].key >> ORC_XSUB_BITS) != cell) tab[2 * cell + 1] = s + 1;
        }
    }
    free(kv);

    /* fluid lists: candidates around the search positions (radius Rs; H without a skin), then
     * those within H of the current positions */
    const int cap = L->cap;
    L->nbr = (int*)malloc(sizeof(int) * ((size_t)nf * cap + 1));
    L->ncnt = (int*)calloc((size_t)nf + 1, sizeof(int));
    {
        const int capS = skin_on ? pb->max_skin_neighbors : cap;
        int* cand = (int*)malloc(sizeof(int) * ((size_t)capS + 1));
        for (int s = 0; s < nf; ++s) {
            const int w = L->fworld_s[s];
            const double H = pb->support[w];
            const float Hf = (float)H, H2 = (float)(H * H);
            float thr2, Rs, Rs2;
            skin_radii(H, pb->skin, &thr2, &Rs, &Rs2);
            int ns = skin_on ? scan_cells(&gws[w], ft, fb, s, Rs, Rs2, capS, cand)
                             : scan_cells(&gws[w], ft, fx, s, Hf, H2, capS, cand);
            if (ns > capS) {
                L->overflow = 1;
                ns = capS;
            }
            int c = 0;
            for (int t = 0; t < ns; ++t) {
                const int j = cand[t];
                const float dx = fx[3 * s] - fx[3 * j], dy = fx[3 * s + 1] - fx[3 * j + 1],
                            dz = fx[3 * s + 2] - fx[3 * j + 2];
                const float d2 = (dx * dx + dy * dy) + dz * dz;
                if (!(d2 < H2)) continue;
                if (c < cap) L->nbr[(size_t)s * cap + c] = j;
                ++c;
            }
            if (c > cap) L->overflow = 1;
            if (c > L->max_count) L->max_count = c;
            L->ncnt[s] = c < cap ? c : cap;
        }
        free(cand);
    }
    /* boundary visits around the current positions (radius H, same cell order) */
    L->bstart = (int*)calloc((size_t)nf + 2, sizeof(int));
    size_t bcap = (size_t)nf * 16 + 16, bn = 0;
    L->blist = (int*)malloc(sizeof(int) * bcap);
    for (int s = 0; s < nf; ++s) {
        const orc_gw* g = &gws[L->fworld_s[s]];
        const double H = pb->support[L->fworld_s[s]];
        const float Hf = (float)H, H2 = (float)(H * H);
        const float qx = fx[3 * s], qy = fx[3 * s + 1], qz = fx[3 * s + 2];
        const float rp = Hf + Hf * (1.0f / 256.0f);
        int x0, nx, y0, ny, z0, nz;
        axis_range(qx, rp, g->lo[0], g->inv_w, g->dim[0], &x0, &nx);
        axis_range(qy, rp, g->lo[1], g->inv_w, g->dim[1], &y0, &ny);
        axis_range(qz, rp, g->lo[2], g->inv_t, g->dim[2], &z0, &nz);
        L->bstart[s] = (int)bn;
        for (int grp = 0; grp < 2; ++grp) {  /* static boundary, then moving */
            const int* tab = grp == 0 ? bt : kt;
            const int first = grp == 0 ? 0 : nbs;
            for (int iz = 0; iz < nz; ++iz)
                for (int iy = 0; iy < ny; ++iy) {
                    const int row = g->cell_off + ((z0 + iz) * g->dim[1] + (y0 + iy)) * g->dim[0];
                    for (int ix = 0; ix < nx; ++ix) {
                        const int cell = row + x0 + ix;
                        for (int j = first + tab[2 * cell]; j < first + tab[2 * cell + 1]; ++j) {
                            const float dx = qx - L->bxs[3 * j], dy = qy - L->bxs[3 * j + 1],
                                        dz = qz - L->bxs[3 * j + 2];
                            const float d2 = (dx * dx + dy * dy) + dz * dz;
                            if (!(d2 < H2)) continue;
                            if (bn == bcap) {
                                bcap *= 2;
                                L->blist = (int*)realloc(L->blist, sizeof(int) * bcap);
                            }
                            L->blist[bn++] = j;
                        }
                    }
                }
        }
    }
    L->bstart[nf] = (int)bn;
    free(fx);
    free(fb);
    free(ft);
    free(bt);
    free(kt);
    free(gws);
    free(bid);
    free(bworld);
    return 0;
}

static void orc_layout_free(orc_layout* L) {
    free(L->fid); free(L->fworld); free(L->fperm); free(L->fworld_s); free(L->is_fluid);
    free(L->nbr); free(L->ncnt); free(L->bxs); free(L->bvs); free(L->bstart); free(L->blist);
    free(L->w_fstart); free(L->w_nf); free(L->xbk);
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
 *   xb_in: NULL, or the search positions of state_in (N x 3, see orc_layout_build)
 *   pprev: previous-step pressure in fluid-list order (ascending fluid id), in/out
 *   forced_iters: NULL, or per world the exact number of Jacobi updates (-1 = stop test)
 *   xb_out: NULL, or receives the step's search positions (a non-fluid particle's position)
 *   returns 0, 3 (neighbour overflow) or 6 (a world hit max_iters). */
int orc_step_f32(const orc_problem* pb, const float* q_in, const float* qd_in, const float* f_in,
                 const float* xb_in, float* pprev, float dt, const int* forced_iters, float* q_out,
                 float* qd_out, float* f_out, float* rho_out, float* p_out, float* xb_out,
                 int* nbr_count_out, int* iters_out, double* err_out, int nthreads) {
    return orc_step_impl_f32(pb, q_in, qd_in, f_in, xb_in, pprev, dt, forced_iters, q_out, qd_out, f_out,
                             rho_out, p_out, xb_out, nbr_count_out, iters_out, err_out, nthreads);
}

int orc_step_f64(const orc_problem* pb, const float* q_in, const float* qd_in, const float* f_in,
                 const float* xb_in, float* pprev, float dt, const int* forced_iters, float* q_out,
                 float* qd_out, float* f_out, float* rho_out, float* p_out, float* xb_out,
                 int* nbr_count_out, int* iters_out, double* err_out, int nthreads) {
    return orc_step_impl_f64(pb, q_in, qd_in, f_in, xb_in, pprev, dt, forced_iters, q_out, qd_out, f_out,
                             rho_out, p_out, xb_out, nbr_count_out, iters_out, err_out, nthreads);
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

/* Fluid particle ids in the solver's cell-sorted order, and the fluid/boundary counts
 * (xb_in: NULL or the search positions, as in orc_step_f32). */
int orc_sorted_order(const orc_problem* pb, const float* q_in, const float* xb_in, int* ids_out, int* nf_out,
                     int* nb_out) {
    orc_layout L;
    int rc = orc_layout_build(pb, q_in, NULL, xb_in, &L);
    if (rc) return rc;
    for (int s = 0; s < L.nf; ++s) ids_out[s] = L.fid[L.fperm[s]];
    *nf_out = L.nf;
    *nb_out = L.nb;
    orc_layout_free(&L);
    return 0;
}

/* Fluid-fluid neighbour lists as the solver builds them (particle ids, slot-major
 * idx[k * n + id] like nsph_solver_neighbors; xb_in as in orc_step_f32). */
int orc_solver_neighbors(const orc_problem* pb, const float* q_in, const float* xb_in, int* counts, int* idx) {
    orc_layout L;
    int rc = orc_layout_build(pb, q_in, NULL, xb_in, &L);
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
