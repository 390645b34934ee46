// Context, buffers, cell tables and the public HashGrid (periodic hash grid) of libnsph.
//
// The HashGrid mirrors warp.HashGrid as Newton uses it (builder.py:9851-9855 creates
// HashGrid(128,128,128); solver_xpbd.py:136-139 reserves; solver_xpbd.py:315-319 builds with
// a radius; kernels query with hash_grid_point_id / hash_grid_query / hash_grid_query_next,
// kernels_contact.py:54-73).  Build = cell key per point, stable radix sort of (key, id),
// reordered float4 copy of the points, (start,end) per cell.

#include <stdarg.h>
#include <string.h>

#include <cmath>

#include "common.h"
#include "../../include/nsph_device.h"

namespace nsph {

static thread_local char g_err[1024] = {0};

void set_error(const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

int buf_ensure(DevBuf& b, size_t bytes) {
    if (bytes <= b.bytes && b.ptr) return NSPH_OK;
    if (b.ptr) {
        NSPH_HIP(hipFree(b.ptr));
        b.ptr = nullptr;
        b.bytes = 0;
    }
    size_t sz = bytes < 256 ? 256 : bytes;
    NSPH_HIP(hipMalloc(&b.ptr, sz));
    b.bytes = sz;
    return NSPH_OK;
}

void buf_free(DevBuf& b) {
    if (b.ptr) (void)hipFree(b.ptr);
    b.ptr = nullptr;
    b.bytes = 0;
}

__global__ void k_fill_int2(int2* p, size_t n, int2 v) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = v;
}

__global__ void k_fill_i32(int* p, size_t n, int v) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = v;
}

__global__ void k_cells_clear_prev(const uint32_t* __restrict__ prev_keys, int n_prev,
                                   int2* __restrict__ table) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_prev) return;
    uint32_t k = prev_keys[i];
    if (i == 0 || prev_keys[i - 1] != k) table[k] = make_int2(0, 0);
}

__global__ void k_cells_mark(const uint32_t* __restrict__ keys, int n, int2* __restrict__ table) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t k = keys[i];
    if (i == 0 || keys[i - 1] != k) table[k].x = i;
    if (i == n - 1 || keys[i + 1] != k) table[k].y = i + 1;
}

// Key of each point (HashGrid: single periodic world at origin 0).
__global__ void k_grid_keys(const float* __restrict__ pts, int n, GridWorld g,
                            uint32_t* __restrict__ keys, uint32_t* __restrict__ vals) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    keys[i] = cell_key(g, pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]);
    vals[i] = (uint32_t)i;
}

// Build positions in sorted order, and the inverse permutation (sorted position of each point).
__global__ void k_grid_reorder(const float* __restrict__ pts, const uint32_t* __restrict__ perm,
                               int n, float4* __restrict__ xs, int* __restrict__ inv) {
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    uint32_t id = perm[s];
    xs[s] = make_float4(pts[3 * id], pts[3 * id + 1], pts[3 * id + 2], 0.0f);
    inv[id] = s;
}

__global__ void k_grid_neighbors(GridWorld g, const int2* __restrict__ table,
                                 const float4* __restrict__ xs, const uint32_t* __restrict__ perm,
                                 const int* __restrict__ flags, int n, float r, float r2max,
                                 int max_nbrs, int* __restrict__ counts, int* __restrict__ idx,
                                 int* __restrict__ overflow) {
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const int i = (int)perm[s];
    if (flags && (flags[i] & NSPH_PARTICLE_ACTIVE) == 0) {
        counts[i] = 0;
        return;
    }
    const float4 q = xs[s];
    int c = 0;
    grid_visit(g, table, xs, q.x, q.y, q.z, r, r2max,
               [&](int j, float, float, float, float) {
                   const int id = (int)perm[j];
                   if (id == i) return;
                   if (flags && (flags[id] & NSPH_PARTICLE_ACTIVE) == 0) return;
                   if (c < max_nbrs) idx[(size_t)c * n + i] = id;
                   ++c;
               });
    counts[i] = c;
    if (c > max_nbrs) atomicOr(overflow, 1);
}

__global__ void k_grid_query(GridWorld g, const int2* __restrict__ table,
                             const float4* __restrict__ xs, const uint32_t* __restrict__ perm,
                             const float* __restrict__ q, int nq, float r, float r2max,
                             int max_nbrs, int* __restrict__ counts, int* __restrict__ idx,
                             int* __restrict__ overflow) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nq) return;
    int c = 0;
    grid_visit(g, table, xs, q[3 * k], q[3 * k + 1], q[3 * k + 2], r, r2max,
               [&](int j, float, float, float, float) {
                   if (c < max_nbrs) idx[(size_t)c * nq + k] = (int)perm[j];
                   ++c;
               });
    counts[k] = c;
    if (c > max_nbrs) atomicOr(overflow, 1);
}

// Points within r of each query point, through the public in-kernel query (nsph_device.h).
__global__ void k_grid_count(nsph_grid_device g, const float* __restrict__ q, int nq, float r, float r2max,
                             int* __restrict__ counts) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nq) return;
    const float x = q[3 * k], y = q[3 * k + 1], z = q[3 * k + 2];
    nsph_grid_query_t it = nsph_grid_query_begin(&g, x, y, z, r);
    int id, pos, c = 0;
    while (nsph_grid_query_next(&g, &it, &id, &pos)) {
        const float dx = x - g.points[4 * pos], dy = y - g.points[4 * pos + 1], dz = z - g.points[4 * pos + 2];
        if ((dx * dx + dy * dy) + dz * dz < r2max) ++c;
    }
    counts[k] = c;
}

}  // namespace nsph

using namespace nsph;

struct nsph_grid {
    nsph_ctx* ctx = nullptr;
    int dim[3] = {0, 0, 0};
    int reserved = 0;
    int n = 0;          // points of the last build
    int built = 0;
    float radius = 0.f;
    GridWorld g{};
    DevBuf keys, vals, xs, table, flag;
    DevBuf px, pv;  // consumers' packed per-particle data in sorted order (scratch)
    DevBuf inv;     // sorted position of each point of the last build
    uint32_t* sorted_keys = nullptr;  // result buffers of the last sort (keys / perm)
    uint32_t* perm = nullptr;
    int n_prev = 0;  // points of the previous build (their cells are cleared next build)
    RadixSort rs;
};

int nsph_grid_view(const nsph_grid* g, GridView* out) {
    NSPH_CHECK(g != nullptr && out != nullptr, NSPH_ERR_INVALID, "grid view: NULL grid");
    NSPH_CHECK(g->built, NSPH_ERR_NOT_BUILT, "grid view: the grid has not been built");
    out->g = g->g;
    out->table = bp<int2>(g->table);
    out->perm = g->perm;
    out->n = g->n;
    out->px = bp<float4>(g->px);
    out->inv = bp<int>(g->inv);
    out->pv = bp<float4>(g->pv);
    return NSPH_OK;
}

int nsph_grid_stream(const nsph_grid* g, hipStream_t* out) {
    NSPH_CHECK(g != nullptr && out != nullptr, NSPH_ERR_INVALID, "grid stream: NULL grid");
    *out = g->ctx->stream;
    return NSPH_OK;
}

extern "C" {

const char* nsph_last_error(void) { return g_err; }
int nsph_abi_version(void) { return NSPH_ABI_VERSION; }

int nsph_ctx_create(int device, void* stream, nsph_ctx** out) {
    NSPH_CHECK(out != nullptr, NSPH_ERR_INVALID, "nsph_ctx_create: out is NULL");
    int count = 0;
    NSPH_HIP(hipGetDeviceCount(&count));
    NSPH_CHECK(device >= 0 && device < count, NSPH_ERR_INVALID,
               "nsph_ctx_create: device %d out of range (%d devices)", device, count);
    NSPH_HIP(hipSetDevice(device));
    NSPH_TRY(soft_static_init(device));
    nsph_ctx* c = new nsph_ctx();
    c->device = device;
    c->stream = (hipStream_t)stream;
    *out = c;
    return NSPH_OK;
}

int nsph_ctx_set_stream(nsph_ctx* ctx, void* stream) {
    NSPH_CHECK(ctx != nullptr, NSPH_ERR_INVALID, "nsph_ctx_set_stream: NULL context");
    ctx->stream = (hipStream_t)stream;
    return NSPH_OK;
}

int nsph_ctx_synchronize(nsph_ctx* ctx) {
    NSPH_CHECK(ctx != nullptr, NSPH_ERR_INVALID, "nsph_ctx_synchronize: NULL context");
    NSPH_HIP(hipStreamSynchronize(ctx->stream));
    return NSPH_OK;
}

int nsph_ctx_destroy(nsph_ctx* ctx) {
    delete ctx;
    return NSPH_OK;
}

int nsph_grid_create(nsph_ctx* ctx, int dim_x, int dim_y, int dim_z, nsph_grid** out) {
    NSPH_CHECK(ctx != nullptr && out != nullptr, NSPH_ERR_INVALID, "nsph_grid_create: NULL argument");
    NSPH_CHECK(dim_x >= 3 && dim_y >= 3 && dim_z >= 3, NSPH_ERR_INVALID,
               "nsph_grid_create: every dim must be >= 3 (got %d, %d, %d)", dim_x, dim_y, dim_z);
    NSPH_CHECK((double)dim_x * dim_y * dim_z < 2147483647.0, NSPH_ERR_INVALID,
               "nsph_grid_create: dim_x*dim_y*dim_z must fit in int32");
    nsph_grid* g = new nsph_grid();
    g->ctx = ctx;
    g->dim[0] = dim_x;
    g->dim[1] = dim_y;
    g->dim[2] = dim_z;
    const size_t cells = (size_t)dim_x * dim_y * dim_z;
    int rc = buf_ensure(g->table, cells * sizeof(int2));
    if (rc == NSPH_OK) rc = buf_ensure(g->flag, 256);
    if (rc != NSPH_OK) {
        delete g;
        return rc;
    }
    hipLaunchKernelGGL(k_fill_int2, dim3(1024), dim3(256), 0, ctx->stream, bp<int2>(g->table),
                       cells, make_int2(0, 0));
    NSPH_HIP(hipGetLastError());
    *out = g;
    return NSPH_OK;
}

int nsph_grid_reserve(nsph_grid* g, int n) {
    NSPH_CHECK(g != nullptr && n >= 0, NSPH_ERR_INVALID, "nsph_grid_reserve: bad argument");
    if (n <= g->reserved) return NSPH_OK;
    // The previous build's keys live in these buffers; a regrow drops them, so clear the
    // whole table instead of the previous cells at the next build.
    NSPH_TRY(buf_ensure(g->keys, (size_t)n * 4));
    NSPH_TRY(buf_ensure(g->vals, (size_t)n * 4));
    NSPH_TRY(buf_ensure(g->xs, (size_t)n * 16));
    NSPH_TRY(buf_ensure(g->px, (size_t)n * 16));
    NSPH_TRY(buf_ensure(g->pv, (size_t)n * 16));
    NSPH_TRY(buf_ensure(g->inv, (size_t)n * 4));
    NSPH_TRY(radix_reserve(g->rs, n));
    g->reserved = n;
    if (g->n_prev > 0) {
        const size_t cells = (size_t)g->dim[0] * g->dim[1] * g->dim[2];
        hipLaunchKernelGGL(k_fill_int2, dim3(1024), dim3(256), 0, g->ctx->stream, bp<int2>(g->table),
                           cells, make_int2(0, 0));
        NSPH_HIP(hipGetLastError());
        g->n_prev = 0;
    }
    g->built = 0;
    return NSPH_OK;
}

int nsph_grid_build(nsph_grid* g, const float* d_points, int n, float radius) {
    NSPH_CHECK(g != nullptr, NSPH_ERR_INVALID, "nsph_grid_build: NULL grid");
    NSPH_CHECK(n >= 0 && (n == 0 || d_points != nullptr), NSPH_ERR_INVALID,
               "nsph_grid_build: bad points");
    NSPH_CHECK(radius > 0.f && std::isfinite(radius), NSPH_ERR_INVALID,
               "nsph_grid_build: radius must be positive and finite (got %g)", (double)radius);
    if (n > g->reserved) NSPH_TRY(nsph_grid_reserve(g, n));
    hipStream_t st = g->ctx->stream;
    // clear the cells occupied by the previous build
    if (g->n_prev > 0) {
        hipLaunchKernelGGL(k_cells_clear_prev, dim3((g->n_prev + 255) / 256), dim3(256), 0, st,
                           g->sorted_keys, g->n_prev, bp<int2>(g->table));
    }
    GridWorld gw{};
    gw.lo[0] = gw.lo[1] = gw.lo[2] = 0.f;
    gw.inv_w = (float)(1.0 / (double)radius);
    gw.dim[0] = g->dim[0];
    gw.dim[1] = g->dim[1];
    gw.dim[2] = g->dim[2];
    gw.wrap = 1;
    gw.cell_off = 0;
    g->g = gw;
    g->radius = radius;
    if (n > 0) {
        hipLaunchKernelGGL(k_grid_keys, dim3((n + 255) / 256), dim3(256), 0, st, d_points, n, gw,
                           bp<uint32_t>(g->keys), bp<uint32_t>(g->vals));
        const uint32_t max_key = (uint32_t)((size_t)g->dim[0] * g->dim[1] * g->dim[2] - 1);
        NSPH_TRY(radix_sort(g->rs, bp<uint32_t>(g->keys), bp<uint32_t>(g->vals), n,
                            bits_for(max_key), st, &g->sorted_keys, &g->perm));
        hipLaunchKernelGGL(k_grid_reorder, dim3((n + 255) / 256), dim3(256), 0, st, d_points,
                           g->perm, n, bp<float4>(g->xs), bp<int>(g->inv));
        hipLaunchKernelGGL(k_cells_mark, dim3((n + 255) / 256), dim3(256), 0, st, g->sorted_keys,
                           n, bp<int2>(g->table));
    }
    NSPH_HIP(hipGetLastError());
    g->n = n;
    g->n_prev = n;
    g->built = 1;
    return NSPH_OK;
}

int nsph_grid_point_ids(const nsph_grid* g, int* d_ids) {
    NSPH_CHECK(g != nullptr && d_ids != nullptr, NSPH_ERR_INVALID, "nsph_grid_point_ids: NULL argument");
    NSPH_CHECK(g->built, NSPH_ERR_NOT_BUILT, "nsph_grid_point_ids: grid not built");
    if (g->n > 0)
        NSPH_HIP(hipMemcpyAsync(d_ids, g->perm, (size_t)g->n * 4, hipMemcpyDeviceToDevice,
                                g->ctx->stream));
    return NSPH_OK;
}

int nsph_grid_num_points(const nsph_grid* g, int* n) {
    NSPH_CHECK(g != nullptr && n != nullptr, NSPH_ERR_INVALID, "nsph_grid_num_points: NULL argument");
    *n = g->built ? g->n : 0;
    return NSPH_OK;
}

static int read_overflow(const nsph_grid* g, int* ov) {
    NSPH_HIP(hipMemcpyAsync(ov, g->flag.ptr, 4, hipMemcpyDeviceToHost, g->ctx->stream));
    NSPH_HIP(hipStreamSynchronize(g->ctx->stream));
    return NSPH_OK;
}

int nsph_grid_neighbors(const nsph_grid* g, const int* d_flags, float radius, int max_neighbors,
                        int* d_counts, int* d_idx) {
    NSPH_CHECK(g != nullptr && d_counts != nullptr && (d_idx != nullptr || max_neighbors == 0),
               NSPH_ERR_INVALID, "nsph_grid_neighbors: NULL argument");
    NSPH_CHECK(g->built, NSPH_ERR_NOT_BUILT, "nsph_grid_neighbors: grid not built");
    NSPH_CHECK(radius > 0.f && max_neighbors >= 0, NSPH_ERR_INVALID, "nsph_grid_neighbors: bad radius/capacity");
    hipStream_t st = g->ctx->stream;
    const int n = g->n;
    if (n == 0) return NSPH_OK;
    hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, st, bp<int>(g->flag), (size_t)1, 0);
    hipLaunchKernelGGL(k_grid_neighbors, dim3((n + 127) / 128), dim3(128), 0, st, g->g,
                       bp<int2>(g->table), bp<float4>(g->xs), g->perm, d_flags, n, radius,
                       radius * radius, max_neighbors, d_counts, d_idx, bp<int>(g->flag));
    NSPH_HIP(hipGetLastError());
    int ov = 0;
    NSPH_TRY(read_overflow(g, &ov));
    NSPH_CHECK(!ov, NSPH_ERR_CAPACITY, "nsph_grid_neighbors: a list exceeded max_neighbors=%d",
               max_neighbors);
    return NSPH_OK;
}

int nsph_grid_query(const nsph_grid* g, const float* d_q, int nq, float radius, int max_neighbors,
                    int* d_counts, int* d_idx) {
    NSPH_CHECK(g != nullptr && d_counts != nullptr && (d_idx != nullptr || max_neighbors == 0),
               NSPH_ERR_INVALID, "nsph_grid_query: NULL argument");
    NSPH_CHECK(g->built, NSPH_ERR_NOT_BUILT, "nsph_grid_query: grid not built");
    NSPH_CHECK(radius > 0.f && max_neighbors >= 0 && nq >= 0, NSPH_ERR_INVALID, "nsph_grid_query: bad argument");
    if (nq == 0) return NSPH_OK;
    NSPH_CHECK(d_q != nullptr, NSPH_ERR_INVALID, "nsph_grid_query: NULL query points");
    hipStream_t st = g->ctx->stream;
    hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, st, bp<int>(g->flag), (size_t)1, 0);
    hipLaunchKernelGGL(k_grid_query, dim3((nq + 127) / 128), dim3(128), 0, st, g->g,
                       bp<int2>(g->table), bp<float4>(g->xs), g->perm, d_q, nq, radius,
                       radius * radius, max_neighbors, d_counts, d_idx, bp<int>(g->flag));
    NSPH_HIP(hipGetLastError());
    int ov = 0;
    NSPH_TRY(read_overflow(g, &ov));
    NSPH_CHECK(!ov, NSPH_ERR_CAPACITY, "nsph_grid_query: a list exceeded max_neighbors=%d", max_neighbors);
    return NSPH_OK;
}

int nsph_grid_device_view(const nsph_grid* g, nsph_grid_device* out) {
    NSPH_CHECK(g != nullptr && out != nullptr, NSPH_ERR_INVALID, "nsph_grid_device_view: NULL argument");
    NSPH_CHECK(g->built, NSPH_ERR_NOT_BUILT, "nsph_grid_device_view: grid not built");
    out->cells = bp<int>(g->table);
    out->ids = g->perm;
    out->points = bp<float>(g->xs);
    for (int a = 0; a < 3; ++a) out->dim[a] = g->dim[a];
    out->inv_cell = g->g.inv_w;
    out->n = g->n;
    return NSPH_OK;
}

int nsph_grid_count(const nsph_grid* g, const float* d_q, int nq, float radius, int* d_counts) {
    NSPH_CHECK(g != nullptr && d_counts != nullptr, NSPH_ERR_INVALID, "nsph_grid_count: NULL argument");
    NSPH_CHECK(radius > 0.f && nq >= 0, NSPH_ERR_INVALID, "nsph_grid_count: bad argument");
    nsph_grid_device view;
    NSPH_TRY(nsph_grid_device_view(g, &view));
    if (nq == 0) return NSPH_OK;
    NSPH_CHECK(d_q != nullptr, NSPH_ERR_INVALID, "nsph_grid_count: NULL query points");
    hipLaunchKernelGGL(k_grid_count, dim3((nq + 255) / 256), dim3(256), 0, g->ctx->stream, view, d_q, nq, radius,
                       radius * radius, d_counts);
    NSPH_HIP(hipGetLastError());
    return NSPH_OK;
}

int nsph_grid_destroy(nsph_grid* g) {
    if (!g) return NSPH_OK;
    const hipError_t e = nsph_ctx_drain(g->ctx);
    buf_free(g->keys);
    buf_free(g->vals);
    buf_free(g->xs);
    buf_free(g->table);
    buf_free(g->flag);
    buf_free(g->px);
    buf_free(g->pv);
    buf_free(g->inv);
    radix_free(g->rs);
    delete g;
    NSPH_HIP(e);
    return NSPH_OK;
}

}  // extern "C"
