// Internal definitions shared by the libnsph translation units (gfx950 only).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include <string>

#include "../../include/nsph.h"

namespace nsph {

// ---- error handling -------------------------------------------------------------------

void set_error(const char* fmt, ...);

// Resolve the static soft-contact workspace of `device` (current); called by nsph_ctx_create.
int soft_static_init(int device);

#define NSPH_HIP(call)                                                                     \
    do {                                                                                   \
        hipError_t e_ = (call);                                                            \
        if (e_ != hipSuccess) {                                                            \
            ::nsph::set_error("%s failed: %s (%s:%d)", #call, hipGetErrorString(e_),       \
                              __FILE__, __LINE__);                                         \
            return NSPH_ERR_HIP;                                                           \
        }                                                                                  \
    } while (0)

#define NSPH_CHECK(cond, code, ...)                                                        \
    do {                                                                                   \
        if (!(cond)) {                                                                     \
            ::nsph::set_error(__VA_ARGS__);                                                \
            return (code);                                                                 \
        }                                                                                  \
    } while (0)

#define NSPH_TRY(expr)                                                                     \
    do {                                                                                   \
        int r_ = (expr);                                                                   \
        if (r_ != NSPH_OK) return r_;                                                      \
    } while (0)

// ---- device buffers ---------------------------------------------------------------------

struct DevBuf {
    void* ptr = nullptr;
    size_t bytes = 0;
};

int buf_ensure(DevBuf& b, size_t bytes);  // grows (hipMalloc) only when too small
void buf_free(DevBuf& b);
template <class T> inline T* bp(const DevBuf& b) { return reinterpret_cast<T*>(b.ptr); }

// ---- cell geometry --------------------------------------------------------------------
//
// One grid "world": a box of dim[0]*dim[1]*dim[2] cells of width 1/inv_w starting at lo.
// wrap = 1: cell index modulo dim (periodic hash, as the reference HashGrid);
// wrap = 0: cell index clamped into the box (particles outside land in the border cells).
// Key of a cell = cell_off + (cz*dim_y + cy)*dim_x + cx, so keys are world-major and,
// inside a world, x-fastest.

struct GridWorld {
    float lo[3];
    float inv_w;
    int dim[3];
    int wrap;
    int cell_off;
    float inv_t;  // SPH grids: inverse height of the z cells, 2 inv_w (half-height cells)
    int pad[2];
};

__host__ __device__ inline int cell_coord(float x, float lo, float inv_w, int dim, int wrap) {
    float f = floorf((x - lo) * inv_w);
    f = fminf(fmaxf(f, -1.0e9f), 1.0e9f);
    int c = (int)f;
    if (wrap) {
        c %= dim;
        if (c < 0) c += dim;
    } else {
        c = c < 0 ? 0 : (c >= dim ? dim - 1 : c);
    }
    return c;
}

// Unclamped, unwrapped integer cell coordinate (used for query ranges).
__host__ __device__ inline int cell_raw(float x, float lo, float inv_w) {
    float f = floorf((x - lo) * inv_w);
    f = fminf(fmaxf(f, -1.0e9f), 1.0e9f);
    return (int)f;
}

__host__ __device__ inline uint32_t cell_key(const GridWorld& g, float x, float y, float z) {
    int cx = cell_coord(x, g.lo[0], g.inv_w, g.dim[0], g.wrap);
    int cy = cell_coord(y, g.lo[1], g.inv_w, g.dim[1], g.wrap);
    int cz = cell_coord(z, g.lo[2], g.inv_w, g.dim[2], g.wrap);
    return (uint32_t)(g.cell_off + (cz * g.dim[1] + cy) * g.dim[0] + cx);
}

// Per-axis query range [c0, c0 + n) (n >= 1) for the interval [x - r, x + r].
// Wrapped axes: n is capped at dim and cells are taken modulo dim by the caller.
// Clamped axes: both ends are clamped into [0, dim-1] (monotone, so every point within r
// of x whose own clamped cell is c satisfies c0 <= c < c0 + n).
__host__ __device__ inline void axis_range(float x, float r, float lo, float inv_w, int dim,
                                           int wrap, int* c0, int* n) {
    int a = cell_raw(x - r, lo, inv_w);
    int b = cell_raw(x + r, lo, inv_w);
    if (wrap) {
        int m = b - a + 1;
        if (m > dim) {
            a = 0;
            m = dim;
        }
        *c0 = a;
        *n = m;
    } else {
        a = a < 0 ? 0 : (a >= dim ? dim - 1 : a);
        b = b < 0 ? 0 : (b >= dim ? dim - 1 : b);
        *c0 = a;
        *n = b - a + 1;
    }
}

__host__ __device__ inline int wrap_idx(int c, int dim, int wrap) {
    if (!wrap) return c;
    c %= dim;
    return c < 0 ? c + dim : c;
}

// A query radius is padded by this relative amount when cell ranges are computed, so that
// FP32 rounding of (x - r - lo) * inv_w can never exclude a point that is strictly within r.
// The guarantee holds while |x - lo| < 2^14 cell widths.
#define NSPH_RANGE_PAD (1.0f / 256.0f)

// Visit every hashed point within `r` of (qx,qy,qz).  Candidates are taken cell by cell:
// z outer, y, then x over the axis ranges, and inside a cell in sorted order (= ascending
// input index, the sort being stable).  A candidate j is accepted when d2 < r2max with
//   d = q - x_j,   d2 = (dx*dx + dy*dy) + dz*dz   (FP32, no contraction: -ffp-contract=off)
// and then passed to visit(j, dx, dy, dz, d2).
template <class Visit>
__device__ inline void grid_visit(const GridWorld& g, const int2* __restrict__ table,
                                  const float4* __restrict__ xs, float qx, float qy, float qz,
                                  float r, float r2max, Visit visit) {
    const float rp = r + r * NSPH_RANGE_PAD;
    int x0, nx, y0, ny, z0, nz;
    axis_range(qx, rp, g.lo[0], g.inv_w, g.dim[0], g.wrap, &x0, &nx);
    axis_range(qy, rp, g.lo[1], g.inv_w, g.dim[1], g.wrap, &y0, &ny);
    axis_range(qz, rp, g.lo[2], g.inv_w, g.dim[2], g.wrap, &z0, &nz);
    for (int iz = 0; iz < nz; ++iz) {
        const int cz = wrap_idx(z0 + iz, g.dim[2], g.wrap);
        for (int iy = 0; iy < ny; ++iy) {
            const int cy = wrap_idx(y0 + iy, g.dim[1], g.wrap);
            const int row = g.cell_off + (cz * g.dim[1] + cy) * g.dim[0];
            for (int ix = 0; ix < nx; ++ix) {
                const int cx = wrap_idx(x0 + ix, g.dim[0], g.wrap);
                const int2 se = table[row + cx];
                for (int j = se.x; j < se.y; ++j) {
                    const float4 p = xs[j];
                    const float dx = qx - p.x, dy = qy - p.y, dz = qz - p.z;
                    const float d2 = (dx * dx + dy * dy) + dz * dz;
                    if (d2 < r2max) visit(j, dx, dy, dz, d2);
                }
            }
        }
    }
}

// ---- radix sort ---------------------------------------------------------------------------

struct RadixSort {
    DevBuf keys_alt, vals_alt, hist, digit_tot;
    int capacity = 0;
};
int radix_reserve(RadixSort& rs, int n);
// Stable LSD radix sort of (keys, vals) over the low `bits` bits. Ping-pongs between the
// given buffers and rs's alternates; *out_keys/*out_vals point to whichever holds the result
// (which buffer depends on n and bits only).  guard: NULL, or a device flag; when it reads 0
// every kernel of the sort returns at once and the buffers keep their contents.
int radix_sort(RadixSort& rs, uint32_t* keys, uint32_t* vals, int n, int bits,
               hipStream_t stream, uint32_t** out_keys, uint32_t** out_vals, const int* guard = nullptr);
// The buffers radix_sort(rs, keys, vals, n, bits, ...) leaves its result in, without sorting.
void radix_result(const RadixSort& rs, uint32_t* keys, uint32_t* vals, int n, int bits, uint32_t** out_keys,
                  uint32_t** out_vals);
void radix_free(RadixSort& rs);

// ---- cell table (start,end) per key -------------------------------------------------------

// Cell tables are int2 (start, end) per key, end exclusive, (0,0) when empty. A build
// clears only the cells that the previous build occupied.
__global__ void k_cells_clear_prev(const uint32_t* __restrict__ prev_keys, int n_prev,
                                   int2* __restrict__ table);
__global__ void k_cells_mark(const uint32_t* __restrict__ keys, int n, int2* __restrict__ table);
__global__ void k_fill_int2(int2* p, size_t n, int2 v);
__global__ void k_fill_i32(int* p, size_t n, int v);

int bits_for(uint32_t max_key);

// Read-only view of a built HashGrid for consumer kernels (cell table, sorted point ids).
struct GridView {
    GridWorld g;
    const int2* table;
    const uint32_t* perm;
    int n;
    float4* px;  // scratch of n entries for the consumers' packed data in sorted order
    float4* pv;
    const int* inv;  // sorted position of each point
};

}  // namespace nsph

struct nsph_grid;
int nsph_grid_view(const nsph_grid* g, nsph::GridView* out);
int nsph_grid_stream(const nsph_grid* g, hipStream_t* out);

struct nsph_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
};

// Before an object's buffers are freed: make its device current and wait for the work
// queued on its context's stream, which may still read them.
static inline hipError_t nsph_ctx_drain(const nsph_ctx* c) {
    if (!c) return hipSuccess;
    const hipError_t e = hipSetDevice(c->device);
    return e != hipSuccess ? e : hipStreamSynchronize(c->stream);
}
