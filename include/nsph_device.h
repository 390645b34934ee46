/*
 * nsph_device.h -- in-kernel queries of a built nsph_grid, for HIP consumer kernels.
 *
 * Replaces Warp's kernel-side HashGrid interface as the reference consumers use it:
 *   hash_grid_point_id(grid, tid)        kernels_contact.py:54, xpbd/kernels.py:244
 *   hash_grid_query(grid, x, r)          kernels_contact.py:68, xpbd/kernels.py:257
 *   hash_grid_query_next(query, index)   kernels_contact.py:71, xpbd/kernels.py:262
 *
 * Host side: nsph_grid_device_view() (declared in nsph.h) fills an nsph_grid_device with
 * the device pointers and geometry of the grid's last build; pass it to a kernel by value.
 * The view stays valid until the grid is built again or destroyed.
 *
 * Device side (HIP translation units): nsph_grid_query_begin / nsph_grid_query_next visit
 * every point of the cells overlapping [x - r, x + r] per axis (cells taken modulo the grid
 * dims, as the periodic HashGrid hashes them), z outer, then y, then x, and inside a cell in
 * ascending point id: a superset of the points within r that the consumer filters by distance,
 * as the reference's loops do.  nsph_grid_point_id maps a thread index to a point in
 * cell-sorted order, so that neighbouring threads handle neighbouring points.
 */
#ifndef NSPH_DEVICE_H
#define NSPH_DEVICE_H

#include "nsph.h"

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>

/* Query state of one thread. */
typedef struct nsph_grid_query_t {
    int x0, y0, z0;  /* first cell per axis (unwrapped) */
    int nx, ny, nz;  /* cells per axis */
    int k;           /* next cell, linear over nz * ny * nx with x fastest */
    int pos, end;    /* next sorted position in the current cell and its end */
} nsph_grid_query_t;

__device__ inline int nsph_grid_cell_floor(float v, float inv_cell) {
    float f = floorf(v * inv_cell);
    f = fminf(fmaxf(f, -1.0e9f), 1.0e9f);
    return (int)f;
}

__device__ inline int nsph_grid_wrap(int c, int dim) {
    c %= dim;
    return c < 0 ? c + dim : c;
}

/* Point id of thread tid in cell-sorted order, or -1 past the last point
 * (hash_grid_point_id). */
__device__ inline int nsph_grid_point_id(const nsph_grid_device* g, int tid) {
    return (tid >= 0 && tid < g->n) ? (int)g->ids[tid] : -1;
}

/* Start a query around (x, y, z) with radius r (hash_grid_query).  The range is padded by
 * r / 256 so that FP32 rounding of the cell coordinates never drops a point within r. */
__device__ inline nsph_grid_query_t nsph_grid_query_begin(const nsph_grid_device* g, float x, float y, float z,
                                                          float r) {
    nsph_grid_query_t q;
    const float rp = r + r * (1.0f / 256.0f);
    const float c[3] = {x, y, z};
    int lo[3], n[3];
    for (int a = 0; a < 3; ++a) {
        lo[a] = nsph_grid_cell_floor(c[a] - rp, g->inv_cell);
        n[a] = nsph_grid_cell_floor(c[a] + rp, g->inv_cell) - lo[a] + 1;
        if (n[a] > g->dim[a]) {  /* the range covers the whole periodic axis: each cell once */
            lo[a] = 0;
            n[a] = g->dim[a];
        }
    }
    q.x0 = lo[0];
    q.y0 = lo[1];
    q.z0 = lo[2];
    q.nx = n[0];
    q.ny = n[1];
    q.nz = n[2];
    q.k = 0;
    q.pos = 0;
    q.end = 0;
    return q;
}

/* Next candidate (hash_grid_query_next): returns 0 when the query is exhausted, else 1 with
 * *id the point id and *pos its sorted position (g->points[4 * pos .. 4 * pos + 2] is its
 * position at the build). */
__device__ inline int nsph_grid_query_next(const nsph_grid_device* g, nsph_grid_query_t* q, int* id, int* pos) {
    while (q->pos >= q->end) {
        if (q->k >= q->nx * q->ny * q->nz) return 0;
        const int ix = q->k % q->nx, iy = (q->k / q->nx) % q->ny, iz = q->k / (q->nx * q->ny);
        const int cx = nsph_grid_wrap(q->x0 + ix, g->dim[0]);
        const int cy = nsph_grid_wrap(q->y0 + iy, g->dim[1]);
        const int cz = nsph_grid_wrap(q->z0 + iz, g->dim[2]);
        const int cell = (cz * g->dim[1] + cy) * g->dim[0] + cx;
        q->pos = g->cells[2 * cell];
        q->end = g->cells[2 * cell + 1];
        ++q->k;
    }
    *pos = q->pos;
    *id = (int)g->ids[q->pos];
    ++q->pos;
    return 1;
}
#endif /* __HIPCC__ */

#endif /* NSPH_DEVICE_H */
