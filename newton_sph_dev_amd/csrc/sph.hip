// SPH particle step for gfx950: IISPH (Ihmsen et al., IEEE TVCG 20(3), 2014) with a cubic
// spline kernel (Monaghan 1992) of support H, static boundary particles (Akinci et al. 2012
// style pseudo-mass psi = rho0 * V_b), relaxed Jacobi pressure iterations and a
// deterministic FP64 mean-compression stop test.
//
// It sits behind SolverBase.step(state_in, state_out, control, contacts, dt)
// (newton/_src/solvers/solver.py:301-316) and ends with the reference particle integrator
// (solver.py:10-47, integrate_particles): v1 = v0 + (f*w + g*step(-w))*dt, |v1| clamp,
// x1 = x0 + v1*dt, inactive particles keep x and leave v untouched.
//
// Layout (HBM, fluid particles in cell-sorted order s; one world = one contiguous run):
//   xs[s]   float4 (x, y, z, mass)         vs[s]   float4 (v, mass / rho)
//   rho[s], aii[s], src[s], p[s]   float
//   xpr[s]  float4 (x, y, z, p/rho^2)   (positions again, so that the pressure-acceleration
//                                     gather of a world of one fluid mass loads 16 bytes)
//   gb[s]   float4 (sum_b gradW_ib, boundary count)
//   avisc[s] float4 (a_visc, velocity divergence term)   ap[s] float4 (a^p, 0)
//   vadv[s]  float4 (v_adv, pr0), pr0 = p0 / rho^2 of the warm start (0 under the zero one)
//   nbr4[q*Nf + s] int4  fluid-fluid neighbour list, slots 4q..4q+3 of particle s (lanes
//                        of a wave read 64 consecutive quads); ascending in s
//   fp[s]   float4 (SPH force, p) at the step's end;  inv[k] sorted position of fluid index k
// Boundary particles are hashed once into their own table (bxs, btable); only the density
// pass visits them, and it folds them into gb, so the other passes never touch them.
//
// Each neighbour gather pass is bound by the L1's cache-line accesses (a lane's load of a
// neighbour's data is one access), so the layouts above keep the loads per neighbour low:
// one float4 for density and pressure acceleration, two for the other passes.
//
// Work partition: a "chunk" is 256 consecutive sorted fluid particles of one world.
// Kernels run one workgroup per chunk, so the world of a workgroup is uniform and every
// per-world quantity (cell box, kernel constants, convergence flag) is a scalar load.
// The per-iteration FP64 reduction sums each chunk with a fixed LDS tree, then each world's
// chunks with one wave in a fixed lane order, so the result is bitwise repeatable and equal
// between a world run alone and the same world inside a batch.

#include <math.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <cstring>
#include <cmath>

#include <vector>

#include "common.h"
#include "mlp_dev.h"

// device view of an MLP handle (defined in mlp.hip)
nsph::MlpDev nsph_mlp_device_view(const nsph_mlp* m);

namespace nsph {

constexpr int CHUNK = 256;

struct PhysWorld {
    float H, H2, invH, sigma;        // support, H^2, 1/H, 8/(pi H^3)
    float sigma2, gsig2, gsig6, W0;  // 2 sigma, sigma/H^2, 6 sigma/H, W(0) = sigma
    float psi, eps2, visc10, rq;     // rho0*V_b, 0.01 H^2, 10 nu, padded query radius
    float g[3], rho0;
    float Rs, Rs2, thr2;  // skin radius H (1 + skin), its square, squared rebuild displacement
    int skin_on;          // lists are kept between steps (skin > 0)
};

// sqrtf(x), correctly rounded, in fewer instructions for 2^-96 <= x < inf: the hardware root
// s (within 1 ulp) is replaced by its lower neighbour when that one's square is not below x,
// and by its upper neighbour when that one's square is below x (residuals x - s' s exact by
// FMA), which leaves the root nearest sqrt(x).  sqrtf's own lowering adds to the same
// correction a scaling for arguments below 2^-96 and a pass-through of 0 and inf; such
// arguments (and NaN) go to sqrtf here.
__device__ inline float sqrt_rn(float x) {
    if (__builtin_expect((__float_as_uint(x) - 0x0F800000u) >= (0x7F800000u - 0x0F800000u), 0))
        return sqrtf(x);
    const float s = __builtin_amdgcn_sqrtf(x);
    const float dn = __uint_as_float(__float_as_uint(s) - 1u);
    const float up = __uint_as_float(__float_as_uint(s) + 1u);
    float t = fmaf(-dn, s, x) <= 0.0f ? dn : s;
    t = fmaf(-up, s, x) > 0.0f ? up : t;
    return t;
}

// W(r) and (dW/dr)/r of the cubic spline for r^2 = r2 < H^2.
__device__ inline void kern_wg(const PhysWorld& P, float r2, float* W, float* gf) {
    const float r = sqrtf(r2);
    const float q = r * P.invH;
    if (q <= 0.5f) {
        float a = 6.0f * q;
        a = a * q;
        a = a * (q - 1.0f);
        *W = P.sigma * (a + 1.0f);
        *gf = P.gsig2 * (18.0f * q - 12.0f);
    } else {
        const float u = fmaxf(1.0f - q, 0.0f);
        *W = P.sigma2 * ((u * u) * u);
        *gf = -(P.gsig6 * (u * u)) / r;
    }
}

__device__ inline float kern_gf(const PhysWorld& P, float r2) {
    const float r = sqrtf(r2);
    const float q = r * P.invH;
    if (q <= 0.5f) return P.gsig2 * (18.0f * q - 12.0f);
    const float u = fmaxf(1.0f - q, 0.0f);
    return -(P.gsig6 * (u * u)) / r;
}

// kern_gf with sqrt_rn (the same bits): the pressure passes' form
__device__ inline float kern_gf_p(const PhysWorld& P, float r2) {
    const float r = sqrt_rn(r2);
    const float q = r * P.invH;
    if (q <= 0.5f) return P.gsig2 * (18.0f * q - 12.0f);
    const float u = fmaxf(1.0f - q, 0.0f);
    return -(P.gsig6 * (u * u)) / r;
}

// Neumaier compensated add.
__device__ inline void neumaier(float& s, float& c, float t) {
    const float u = s + t;
    if (fabsf(s) >= fabsf(t))
        c += (s - u) + t;
    else
        c += (t - u) + s;
    s = u;
}

__device__ inline float dot3(float ax, float ay, float az, float bx, float by, float bz) {
    return (ax * bx + ay * by) + az * bz;
}

// Three consecutive floats of record k (12-byte stride), one dwordx3 access.
struct F3 {
    float x, y, z;
};
__device__ inline float3 load3(const float* p, int k) {
    const F3 v = *reinterpret_cast<const F3*>(p + 3 * (size_t)k);
    return make_float3(v.x, v.y, v.z);
}
__device__ inline void store3(float* p, int k, float x, float y, float z) {
    *reinterpret_cast<F3*>(p + 3 * (size_t)k) = F3{x, y, z};
}

struct SphDev {
    int Nf, cap;
    int nsub;  // workgroups per chunk for the gather kernels (CHUNK / gather block size)
    // per-world tables
    const GridWorld* gw;
    const PhysWorld* phys;
    const int* w_fstart;
    const int* w_nf;
    const int* w_cfirst;
    const int* w_nchunks;
    const int* chunk_world;
    // boundary (static)
    const int2* btable;
    const int* bcend;
    const float4* bxs;
    const unsigned char* breach;  // per cell: 1 if a boundary particle lies in its 3x3x5 block
    // moving boundary (zero-mass particles with a velocity), hashed every step from state_in;
    // kxs == NULL when there is none
    const int2* ktable;
    const int* kcend;
    const float4* kxs;
    const float4* kvs;
    const unsigned char* kbreach;  // per cell: 1 if a moving boundary particle lies in its 3x3x5 block
    float* kvb;  // per sorted fluid particle: sum_b v_b . gradW_ib over moving boundary particles
    // fluid grid
    int2* ftable;
    int* fcend;
    // sorted fluid arrays
    float4* xs;
    float4* vs;
    float* rho;
    float* aii;
    float4* gb;
    float4* avisc;
    float4* vadv;
    float* src;
    float* p;
    float4* xpr;
    float4* fp;  // (force, p) of the step's end, sorted order
    int* mixed;  // per world: 1 when its fluid particles do not all have one mass
    // pressure acceleration a^p of the current iteration, three floats per particle: a 12-byte
    // record puts 10 2/3 particles in a cache line where a 16-byte one puts 8, so a wave-wide
    // load of it touches fewer lines
    float* a3;
    float* pws;  // warm-start input in sorted order
    float* p0;   // warm-start pressure the iterations started from, sorted order
    int4* nbr4;
    int* ncnt;
    float2* aux;  // (boundary neighbours, distance to the nearest boundary particle / H, 1 if none)
    // iteration control
    double* partial;
    int* iters;
    double* err;
    int* conv;
    int* act;
    int* flags;  // [0] overflow, [2] non-finite
    int* wmax;   // per wave of the neighbour kernel: its largest uncapped count
    int* bq_count;   // [0]: particles queued for the boundary visit this step
    int2* bq_id;     // queued (sorted index, world)
    float4* bq_a;    // their fluid density state: (sum, compensation, Gf.x, Gf.y)
    float2* bq_b;    //                            (Gf.z, S)
    // lists kept between steps (skin radius Rs): per sorted particle the candidates within Rs
    // of its search position at the world's last search (capS of them at most, 0 = no skin)
    int capS;
    int4* snb4;
    int* scnt;
    float4* xbs;     // search positions, sorted order
    float4* xbk;     // search positions, fluid index order
    uint32_t* okey;  // sort key of fluid index k at its world's last search
    int* wflag;      // per world: bit 0 a particle moved too far, bit 1 new search positions
    int* rb;         // per world: 0 kept, 1 search at the positions, 2 at state_in's search positions
    int* anyrb;      // [0]: some world searches this step (guards the sort and reorder)
    int* built;      // [0]: a sorted layout exists (the first step searches every world)
    int nwave;   // 4 per chunk
    // whole-world solve: worlds in the order the workgroups take them (largest first) and
    // the queue position (cleared by k_sph_begin)
    const int* w_order;
    int* wq;
};

// Workgroups are dispatched round-robin over the 8 XCDs, each with its own L2. Renumber them
// so that XCD x runs one contiguous range of logical workgroups: neighbouring chunks (which
// gather each other's particles) then share an L2 instead of being spread over all eight.
constexpr int NUM_XCD = 8;
__device__ inline int xcd_swizzle(int i, int n) {
    const int x = i % NUM_XCD, k = i / NUM_XCD;
    const int q = n / NUM_XCD, r = n % NUM_XCD;
    return x * q + min(x, r) + k;
}

// Map workgroup -> (world, rank inside world, sorted index). Returns false past the world's end.
__device__ inline bool chunk_slot(const SphDev& D, int* w, int* s) {
    const int c = xcd_swizzle(blockIdx.x, gridDim.x);
    const int ww = __builtin_amdgcn_readfirstlane(D.chunk_world[c]);
    const int r = (c - D.w_cfirst[ww]) * CHUNK + (int)threadIdx.x;
    *w = ww;
    *s = D.w_fstart[ww] + r;
    return r < D.w_nf[ww];
}

// Same with a chunk split over nsub workgroups of CHUNK/nsub threads (gather kernels without
// a chunk reduction; the split changes scheduling only, never the arithmetic).
__device__ inline bool chunk_slot_sub(const SphDev& D, int nsub, int* w, int* s) {
    const int b = xcd_swizzle(blockIdx.x, gridDim.x);
    const int c = b / nsub;
    const int sub = b - c * nsub;
    const int ww = __builtin_amdgcn_readfirstlane(D.chunk_world[c]);
    const int r = (c - D.w_cfirst[ww]) * CHUNK + sub * (int)blockDim.x + (int)threadIdx.x;
    *w = ww;
    *s = D.w_fstart[ww] + r;
    return r < D.w_nf[ww];
}

// ---- build ------------------------------------------------------------------------------
//
// Cells are w wide along x and y and w / 2 along z (inverse width inv_t = 2 inv_w, so the z
// coordinate in half cells is exactly twice the full-cell one): a row of cells (cy, cz) is a
// slab along x one particle spacing high at the support H = 2 spacings, so the 64
// consecutive sorted particles of a wave lie in one or two such rows and their neighbour
// lists run through the same rows in step, which keeps the number of distinct cache lines of
// a wave-wide neighbour load low (about a third fewer than with cubic cells on a lattice).
// A query of radius r <= w spans at most 3 cells along x and y and 5 along z.
// Sort keys split each cell into NSPH_XSUB slabs along x:
//   key = (cell_off + (cz*dim_y + cy)*dim_x) * XSUB + gx,   gx = clamp(floor(XSUB f_x), 0, XSUB dim_x - 1)
// with f_x = (x - lo_x) inv_w, the same FP32 value the cell coordinate is taken from, so
// gx >> XSUB_BITS is the cell's cx, cy = clamp(floor((y - lo_y) inv_w)) and
// cz = clamp(floor((z - lo_z) inv_t)).
// Sorted order is therefore (world, cz, cy, cx, slab, id).
// The cell table keeps one int2 per cell: x = first sorted index of the cell, y = the counts
// of its XSUB slabs packed as bytes (slab k in bits 8k..8k+7), 0 for an empty cell.  A query
// can then clip each row of cells to the x-extent of its sphere at slab resolution.  A cell
// of NSPH_SLAB_SAT or more points (points outside the box pile up in its clamped border
// cells) stores NSPH_SLAB_SAT as its count; such a cell is visited whole, up to its end
// index kept in a separate per-cell array.

#define NSPH_XSUB_BITS 2
#define NSPH_XSUB (1 << NSPH_XSUB_BITS)
#define NSPH_SLAB_SAT 255

__host__ __device__ inline int slab_coord(float x, float lo, float inv_w, int dim) {
    float f = floorf(((x - lo) * inv_w) * (float)NSPH_XSUB);
    f = fminf(fmaxf(f, -1.0e9f), 1.0e9f);
    const int c = (int)f;
    const int top = dim * NSPH_XSUB - 1;
    return c < 0 ? 0 : (c > top ? top : c);
}

__host__ __device__ inline uint32_t slab_key(const GridWorld& g, float x, float y, float z) {
    const int gx = slab_coord(x, g.lo[0], g.inv_w, g.dim[0]);
    const int cy = cell_coord(y, g.lo[1], g.inv_w, g.dim[1], 0);
    const int cz = cell_coord(z, g.lo[2], g.inv_t, g.dim[2], 0);
    return (uint32_t)(g.cell_off + (cz * g.dim[1] + cy) * g.dim[0]) * NSPH_XSUB + (uint32_t)gx;
}

// Table entries of sorted position s (keys sorted): the cell's first index, its end, and
// its slab prefix word: byte k - 1 (k = 1 .. 4) is the number of the cell's points in slabs
// 0 .. k - 1, so byte 3 is the cell's count.  The cell's last position counts the slabs by
// looking back at most NSPH_SLAB_SAT - 1 keys; a cell that long stores NSPH_SLAB_SAT in
// byte 3 (saturated) and nothing else.
__device__ inline void mark_cell(const uint32_t* __restrict__ skeys, int s, int n,
                                 int2* __restrict__ table, int* __restrict__ cend) {
    const uint32_t key = skeys[s];
    const uint32_t cell = key >> NSPH_XSUB_BITS;
    if (s == 0 || (skeys[s - 1] >> NSPH_XSUB_BITS) != cell) table[cell].x = s;
    if (s == n - 1 || (skeys[s + 1] >> NSPH_XSUB_BITS) != cell) {
        cend[cell] = s + 1;
        const int stop = s - (NSPH_SLAB_SAT - 1) > 0 ? s - (NSPH_SLAB_SAT - 1) : 0;
        uint32_t acc = 1u << (8u * (key & (NSPH_XSUB - 1)));  // per-slab counts, one byte each
        int j = s;
        while (j > stop && (skeys[j - 1] >> NSPH_XSUB_BITS) == cell) {
            --j;
            acc += 1u << (8u * (skeys[j] & (NSPH_XSUB - 1)));
        }
        const int cnt = s + 1 - j;  // at most NSPH_SLAB_SAT, which also stands for more
        uint32_t w = (uint32_t)NSPH_SLAB_SAT << 24;
        if (cnt < NSPH_SLAB_SAT) {
            const uint32_t p1 = acc & 0xffu, p2 = p1 + ((acc >> 8) & 0xffu), p3 = p2 + ((acc >> 16) & 0xffu);
            w = p1 | (p2 << 8) | (p3 << 16) | ((uint32_t)cnt << 24);
        }
        table[cell].y = (int)w;
    }
}

// mark_cell for a whole wave of consecutive sorted indices s (every lane calls it; lanes with
// s >= n only pass values along).  The neighbours' keys come from the adjacent lanes, and the
// per-slab counts from the cell's first lane in the wave come from a segmented inclusive scan;
// only a cell that began before the wave walks back through the keys in memory, from the
// wave's first index, with the same stop as mark_cell, so the table words are the same.
__device__ inline void mark_cell_wave(const uint32_t* __restrict__ skeys, int s, int n,
                                      int2* __restrict__ table, int* __restrict__ cend) {
    const int lane = threadIdx.x & 63;
    const bool valid = s < n;
    const uint32_t key = valid ? skeys[s] : 0u;
    const uint32_t cell = key >> NSPH_XSUB_BITS;
    uint32_t kprev = (uint32_t)__shfl_up((int)key, 1, 64);
    uint32_t knext = (uint32_t)__shfl_down((int)key, 1, 64);
    if (lane == 0 && valid && s > 0) kprev = skeys[s - 1];
    if (lane == 63 && valid && s + 1 < n) knext = skeys[s + 1];
    const bool start = s == 0 || lane == 0 || (kprev >> NSPH_XSUB_BITS) != cell;
    const bool before = lane == 0 && s > 0 && (kprev >> NSPH_XSUB_BITS) == cell;  // cell began earlier
    const bool end = s == n - 1 || (knext >> NSPH_XSUB_BITS) != cell;
    int seg = start ? lane : 0;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const int o = __shfl_up(seg, d, 64);
        seg = o > seg ? o : seg;
    }
    uint32_t acc = 1u << (8u * (key & (NSPH_XSUB - 1)));
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t a = (uint32_t)__shfl_up((int)acc, d, 64);
        if (lane - d >= seg) acc += a;
    }
    const bool walk = __shfl((int)before, 0, 64) != 0 && seg == 0;
    if (!valid) return;
    if (start && !before) table[cell].x = s;
    if (end) {
        cend[cell] = s + 1;
        const int stop = s - (NSPH_SLAB_SAT - 1) > 0 ? s - (NSPH_SLAB_SAT - 1) : 0;
        int j = s - (lane - seg);
        if (walk) {
            while (j > stop && (skeys[j - 1] >> NSPH_XSUB_BITS) == cell) {
                --j;
                acc += 1u << (8u * (skeys[j] & (NSPH_XSUB - 1)));
            }
        }
        const int cnt = s + 1 - j;
        uint32_t w = (uint32_t)NSPH_SLAB_SAT << 24;
        if (cnt < NSPH_SLAB_SAT) {
            const uint32_t p1 = acc & 0xffu, p2 = p1 + ((acc >> 8) & 0xffu), p3 = p2 + ((acc >> 16) & 0xffu);
            w = p1 | (p2 << 8) | (p3 << 16) | ((uint32_t)cnt << 24);
        }
        table[cell].y = (int)w;
    }
}

// Particles in the new sorted order, cell table, inverse permutation; with kept lists also
// the search positions in sorted order.  Runs only when some world searches anew.
__global__ void k_sph_reorder(SphDev D, const float* __restrict__ q, const float* __restrict__ qd,
                              const float* __restrict__ mass, const int* __restrict__ fid,
                              const uint32_t* __restrict__ perm, const float* __restrict__ pprev,
                              const float* __restrict__ p_in,
                              const uint32_t* __restrict__ skeys, int n, float4* __restrict__ xs,
                              float4* __restrict__ vs, float* __restrict__ pws,
                              int2* __restrict__ ftable, int* __restrict__ fcend,
                              int* __restrict__ inv) {
    if (*D.anyrb == 0) return;
    const int s = blockIdx.x * blockDim.x + threadIdx.x;  // blockDim.x: a multiple of 64
    if (s == 0) *D.built = 1;
    if (s < n) {
        const int k = (int)perm[s];
        const int id = fid[k];
        inv[k] = s;
        xs[s] = make_float4(q[3 * id], q[3 * id + 1], q[3 * id + 2], mass[id]);
        vs[s] = make_float4(qd[3 * id], qd[3 * id + 1], qd[3 * id + 2], 0.0f);
        pws[s] = p_in ? p_in[id] : pprev[k];
        if (D.capS > 0) D.xbs[s] = D.xbk[k];
    }
    mark_cell_wave(skeys, s, n, ftable, fcend);
}

// Particles in the sorted order of the last search (one thread per sorted fluid particle);
// nothing to do before the first search.
__global__ void k_sph_reorder_kept(SphDev D, const float* __restrict__ q, const float* __restrict__ qd,
                                   const float* __restrict__ mass, const int* __restrict__ fid,
                                   const uint32_t* __restrict__ perm, const float* __restrict__ pprev,
                                   const float* __restrict__ p_in, int n) {
    if (*D.built == 0) return;
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const int k = (int)perm[s];
    const int id = fid[k];
    D.xs[s] = make_float4(q[3 * id], q[3 * id + 1], q[3 * id + 2], mass[id]);
    D.vs[s] = make_float4(qd[3 * id], qd[3 * id + 1], qd[3 * id + 2], 0.0f);
    D.pws[s] = p_in ? p_in[id] : pprev[k];
}

// Search positions b of the step's input (one thread per fluid index k): state_in's (xb_in,
// particle order) or, without them, those of the solver's last search (none before the first:
// NaN).  Per world, bit 0 of wflag: some particle lies farther than the rebuild distance from
// its b (FP32 (dx dx + dy dy) + dz dz > thr2 with d = x - b, or non-finite);  bit 1: some b
// differs from the solver's copy, or there is no copy yet (the kept candidates are not those of b).
__global__ void k_sph_displace(SphDev D, int n, const float* __restrict__ q, const int* __restrict__ fid,
                               const int* __restrict__ fworld, const float* __restrict__ xb_in) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int id = fid[k];
    const int w = fworld[k];
    const bool built = *D.built != 0;
    const float4 c = built ? D.xbk[k] : make_float4(NAN, NAN, NAN, 0.0f);
    float bx = c.x, by = c.y, bz = c.z;
    bool mis = false;
    if (xb_in) {
        bx = xb_in[3 * id];
        by = xb_in[3 * id + 1];
        bz = xb_in[3 * id + 2];
        mis = !built || __float_as_uint(bx) != __float_as_uint(c.x) || __float_as_uint(by) != __float_as_uint(c.y) ||
              __float_as_uint(bz) != __float_as_uint(c.z);
    }
    const float dx = q[3 * id] - bx, dy = q[3 * id + 1] - by, dz = q[3 * id + 2] - bz;
    const float d2 = (dx * dx + dy * dy) + dz * dz;
    const bool far = !(d2 <= D.phys[w].thr2);  // NaN (a non-finite position or b) counts as far
    const int bits = (far ? 1 : 0) | (mis ? 2 : 0);
    // one atomic per wave when the wave's particles share a world (the usual case)
    const uint64_t act = __ballot(1);
    const int leader = __ffsll((unsigned long long)act) - 1;
    const int w0 = __shfl(w, leader, 64);
    if (__ballot(w != w0) == 0ull) {
        const int any = (__ballot(bits & 1) ? 1 : 0) | (__ballot(bits & 2) ? 2 : 0);
        if ((int)__lane_id() == leader && any) atomicOr(&D.wflag[w0], any);
    } else if (bits) {
        atomicOr(&D.wflag[w], bits);
    }
}

// One thread per world: how the world's lists are made this step (rb):
//   0  kept: the candidates of the last search are filtered again
//   1  searched anew around the current positions (any particle farther than the rebuild
//      distance from its search position, no skin, or forced)
//   2  searched anew around state_in's search positions (they are not those of the kept
//      candidates, and every particle is within the rebuild distance of them)
// Cases 0 and 2 give the same lists: a pure function of the positions and the search positions.
__global__ void k_sph_decide(SphDev D, int W, int force) {
    const int w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= W) return;
    const int f = D.wflag[w];
    int r = 0;
    if (D.w_nf[w] > 0) {
        if (!D.phys[w].skin_on || force || (f & 1)) r = 1;
        else if (f & 2) r = 2;
    }
    D.rb[w] = r;
    if (r) atomicOr(D.anyrb, 1);
}

// Start of a step (threads over worlds and waves): iteration counters and flags cleared.
__device__ inline void begin_one(const SphDev& D, int W, int k) {
    if (k < W) {
        D.iters[k] = 0;
        D.err[k] = 0.0;
        D.mixed[k] = 0;
        D.conv[k] = D.w_nf[k] == 0 ? 1 : 0;
        D.wflag[k] = 0;
    }
    if (k == 0) {
        D.act[0] = 0;
        D.act[1] = W;
        D.flags[0] = 0;
        D.flags[2] = 0;
        D.bq_count[0] = 0;
        D.wq[0] = 0;
    }
    if (k < D.nwave) D.wmax[k] = 0;
}

__global__ void k_sph_begin(SphDev D, int W) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    begin_one(D, W, k);
    if (k == 0) D.anyrb[0] = 0;
}

// Sort keys when some world searches anew (one thread per fluid index k): the cell of k's
// previous sorted key (prev_keys, zero-filled before the first build) is cleared; a searching
// world's particle gets the key of its search position (the current one for rb = 1, state_in's
// for rb = 2), which becomes its xbk; a kept world's keeps the key of its last search, so the
// stable sort leaves that world's order as it was.  prev_keys may be the keys array itself;
// thread k reads entry k before writing it.
__global__ void k_sph_keys(SphDev D, int n, const uint32_t* prev_keys, int2* __restrict__ table,
                           const float* __restrict__ q, const float* __restrict__ xb_in,
                           const int* __restrict__ fid, const int* __restrict__ fworld,
                           const GridWorld* __restrict__ gw, uint32_t* keys, uint32_t* __restrict__ vals) {
    if (*D.anyrb == 0) return;
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) {
        table[prev_keys[k] >> NSPH_XSUB_BITS] = make_int2(0, 0);
        const int id = fid[k];
        const int w = fworld[k];
        const int r = D.rb[w];
        uint32_t key;
        if (r) {
            const float* b = r == 2 ? xb_in : q;
            const float bx = b[3 * id], by = b[3 * id + 1], bz = b[3 * id + 2];
            key = slab_key(gw[w], bx, by, bz);
            if (D.capS > 0) D.xbk[k] = make_float4(bx, by, bz, 0.0f);
        } else {
            key = D.okey[k];
        }
        D.okey[k] = key;
        keys[k] = key;
        vals[k] = (uint32_t)k;
    }
}

// Start of a step without kept candidates, one launch (threads over worlds, waves and fluid
// indices): the clearing of k_sph_begin, what k_sph_decide decides when no world keeps its
// candidates (every world with fluid searches: rb = 1, anyrb = 1), and k_sph_keys for rb = 1
// (the previous build's cell cleared, the key of the current position).
__global__ void k_sph_begin_keys(SphDev D, int W, int n, const uint32_t* prev_keys, int2* __restrict__ table,
                                 const float* __restrict__ q, const int* __restrict__ fid,
                                 const int* __restrict__ fworld, const GridWorld* __restrict__ gw, uint32_t* keys,
                                 uint32_t* __restrict__ vals) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    begin_one(D, W, k);
    if (k < W) D.rb[k] = D.w_nf[k] > 0 ? 1 : 0;
    if (k == 0) D.anyrb[0] = 1;
    if (k < n) {
        table[prev_keys[k] >> NSPH_XSUB_BITS] = make_int2(0, 0);
        const int id = fid[k];
        const uint32_t key = slab_key(gw[fworld[k]], q[3 * id], q[3 * id + 1], q[3 * id + 2]);
        D.okey[k] = key;
        keys[k] = key;
        vals[k] = (uint32_t)k;
    }
}

// ---- neighbour lists ---------------------------------------------------------------------
//
// Lists are int4 quads: nbr4[q * Nf + s] holds slots 4q .. 4q+3 of sorted particle s, so a
// wave reads 64 consecutive 16-byte quads per load.  Gather passes fetch one quad, issue the
// four neighbour loads together, then accumulate in slot order (the order is fixed; only the
// loads are batched).

// Number of points in slabs 0 .. k-1 (k = 0 .. 4) of a cell whose slab prefix word is w.
__device__ inline int slab_prefix(uint32_t w, int k) {
    return k == 0 ? 0 : (int)((w >> (8 * k - 8)) & 0xffu);
}

__device__ inline bool slab_saturated(uint32_t w) { return (w >> 24) == (uint32_t)NSPH_SLAB_SAT; }

// Sorted range of the points of row cells x0 .. x0+nx-1 (consecutive keys, so one range)
// whose slab lies in [fa, fb] (global slab coordinates); empty when lo >= hi.  c[ix] is the
// table entry of cell `cell0 + ix`; a saturated cell contributes its whole range.
__device__ inline int2 row_slab_range(const int2* __restrict__ c, int nx, int x0, int fa, int fb,
                                      const int* __restrict__ cend, int cell0) {
    int lo = 0x7fffffff, hi = 0;
#pragma unroll
    for (int ix = 0; ix < 3; ++ix) {
        if (ix >= nx) break;
        const int base = (x0 + ix) * NSPH_XSUB;
        const int a = fa - base > 0 ? fa - base : 0;
        const int b = fb - base < NSPH_XSUB - 1 ? fb - base : NSPH_XSUB - 1;
        const uint32_t w = (uint32_t)c[ix].y;
        if (a <= b && w != 0u) {
            const bool sat = slab_saturated(w);
            const int u = sat ? c[ix].x : c[ix].x + slab_prefix(w, a);
            const int v = sat ? cend[cell0 + ix] : c[ix].x + slab_prefix(w, b + 1);
            if (u < v) {
                lo = min(lo, u);
                hi = max(hi, v);
            }
        }
    }
    return make_int2(lo, hi);
}

// Rows of the query sphere, in the order z, y: scan(lo, hi) is called with the sorted range
// [lo, hi) of each row (y0 + iy, z0 + iz) of the (clamped) query box that the sphere meets,
// k = 3 iz + iy ascending.  r <= the cell width w, so the box has at most 3 x 5 rows (z cells
// are w / 2 high) of at most 3 cells.  Each row is clipped to the sphere: with gy, gz the
// distances (in cell widths) from the query to the row's y and z extent (border rows extend
// to infinity, as clamping puts outside points there), the row is skipped when
// gy^2 + gz^2 >= rc^2 and otherwise only the slabs meeting [f_x - hx, f_x + hx],
// hx = sqrt(rc^2 - gy^2 - gz^2), are kept.  Clipping drops only points farther than r, so the
// accepted points and their order are those of the unclipped rows.  rc = r/w (1 + 1/64) + 1/64
// cells covers the FP32 rounding of the cell coordinates (at most 2^-8 cells per axis while
// |x - lo| < 2^14 cell widths; the z coordinate in half cells is exactly twice the full-cell
// one, so its distances halve exactly).
// The row loop stays rolled (one copy of the set-up and of the caller's scan keeps the
// kernels small enough for the instruction cache); the table entries of row k + 1 are
// loaded before row k is scanned, so their latency hides behind the scan.
#define NSPH_ZROWS 5  // z cells a query spans at most (3 along x and y)
template <class Scan>
__device__ inline void rows_for_each(const GridWorld& g, const int2* __restrict__ table,
                                     const int* __restrict__ cend, float qx, float qy, float qz,
                                     float r, Scan scan) {
    const float rp = r + r * NSPH_RANGE_PAD;
    int x0, nx, y0, ny, z0, nz;
    axis_range(qx, rp, g.lo[0], g.inv_w, g.dim[0], 0, &x0, &nx);
    axis_range(qy, rp, g.lo[1], g.inv_w, g.dim[1], 0, &y0, &ny);
    axis_range(qz, rp, g.lo[2], g.inv_t, g.dim[2], 0, &z0, &nz);
    const float fx = (qx - g.lo[0]) * g.inv_w, fy = (qy - g.lo[1]) * g.inv_w, fz = (qz - g.lo[2]) * g.inv_t;
    const float rcell = r * g.inv_w;
    const float rc = rcell + rcell * (1.0f / 64.0f) + (1.0f / 64.0f);
    const float rc2 = rc * rc;
    const int xa = x0 * NSPH_XSUB, xb = (x0 + nx) * NSPH_XSUB - 1;
    // distance from v to cell c's extent, in cells of that axis
    auto gap = [](float v, int c, int dim) {
        const float e0 = c == 0 ? -INFINITY : (float)c, e1 = c == dim - 1 ? INFINITY : (float)(c + 1);
        return fmaxf(fmaxf(e0 - v, v - e1), 0.0f);
    };
    // squared distances (in cell widths) to the three y rows and the five z rows, once per
    // query, held in registers that rotate as the rows advance (y fastest): (gy0, gz0) are
    // always those of the next row to be loaded
    float gy0, gy1, gy2, gz0, gz1, gz2, gz3, gz4;
    {
        float t;
        t = gap(fy, y0, g.dim[1]), gy0 = t * t;
        t = gap(fy, y0 + 1, g.dim[1]), gy1 = t * t;
        t = gap(fy, y0 + 2, g.dim[1]), gy2 = t * t;
        t = 0.5f * gap(fz, z0, g.dim[2]), gz0 = t * t;
        t = 0.5f * gap(fz, z0 + 1, g.dim[2]), gz1 = t * t;
        t = 0.5f * gap(fz, z0 + 2, g.dim[2]), gz2 = t * t;
        t = 0.5f * gap(fz, z0 + 3, g.dim[2]), gz3 = t * t;
        t = 0.5f * gap(fz, z0 + 4, g.dim[2]), gz4 = t * t;
    }
    static_assert(NSPH_ZROWS == 5, "five z rows");
    auto advance = [&](int k) {  // row k was loaded: move to row k + 1
        const float t = gy0;
        gy0 = gy1;
        gy1 = gy2;
        gy2 = t;
        if (k % 3 == 2) {
            gz0 = gz1;
            gz1 = gz2;
            gz2 = gz3;
            gz3 = gz4;
        }
    };
    const int row0 = g.cell_off + (z0 * g.dim[1] + y0) * g.dim[0] + x0;
    // clip row k to the sphere and load its table entries (the row's three cells: cells past a
    // row of fewer than three are loaded too, from the next row or the table's two padding
    // entries, and ignored); false when the row is skipped
    auto load_row = [&](int k, int2 c[3], int* fa, int* fb, int* row) {
        const int iz = k / 3, iy = k - 3 * (k / 3);
        if (!(iz < nz && iy < ny)) return false;
        const float h2 = (rc2 - gy0) - gz0;
        if (!(h2 > 0.0f)) return false;
        // the 1-ulp hardware root: the 1/64-cell pad of rc covers its error
        const float hx = __builtin_amdgcn_sqrtf(h2);
        float a = floorf((fx - hx) * (float)NSPH_XSUB), b = floorf((fx + hx) * (float)NSPH_XSUB);
        a = fminf(fmaxf(a, -1.0e9f), 1.0e9f);
        b = fminf(fmaxf(b, -1.0e9f), 1.0e9f);
        // both ends clamped into the loaded cells' slabs, as points are clamped into slabs
        *fa = min(max((int)a, xa), xb);
        *fb = max(min((int)b, xb), xa);
        if (*fa > *fb) return false;
        *row = row0 + (iz * g.dim[1] + iy) * g.dim[0];
        const int2* __restrict__ tp = table + *row;
#pragma unroll
        for (int ix = 0; ix < 3; ++ix) c[ix] = tp[ix];
        return true;
    };
    constexpr int K = 3 * NSPH_ZROWS;
    int2 c[3];
    int fa = 0, fb = 0, row = 0;
    bool ok = load_row(0, c, &fa, &fb, &row);
    advance(0);
#pragma unroll 1
    for (int k = 0; k < K; ++k) {
        int2 cn[3];
        int fan = 0, fbn = 0, rown = 0;
        const bool okn = k < K - 1 && load_row(k + 1, cn, &fan, &fbn, &rown);
        advance(k + 1);
        // every lane calls scan (an empty range for a skipped row), so a scan may exchange
        // data between lanes
        int2 rr = make_int2(0, 0);
        if (ok) rr = row_slab_range(c, nx, x0, fa, fb, cend, row);
        scan(rr.x, rr.y);
        ok = okn;
        fa = fan;
        fb = fbn;
        row = rown;
#pragma unroll
        for (int ix = 0; ix < 3; ++ix) c[ix] = cn[ix];
    }
}

// Candidates j within r of (qx,qy,qz) in the order z, y, then ascending sorted index;
// calls f(j, dx, dy, dz, d2) for d2 < r2max (d = q - x_j).  Positions are fetched B at a time
// (RV_BATCH by default).
#define RV_BATCH 4
#define XS_PAD 8  // positions readable past the last fluid particle
template <int B = RV_BATCH, class F>
__device__ inline void rows_visit(const GridWorld& g, const int2* __restrict__ table, const int* __restrict__ cend,
                                  const float4* __restrict__ xs, float qx, float qy, float qz,
                                  float r, float r2max, F f) {
    rows_for_each(g, table, cend, qx, qy, qz, r, [&](int lo, int hi) {
        for (int j = lo; j < hi; j += B) {
            float4 p[B];
#pragma unroll
            for (int u = 0; u < B; ++u) p[u] = xs[min(j + u, hi - 1)];
#pragma unroll
            for (int u = 0; u < B; ++u) {
                if (j + u < hi) {
                    const float dx = qx - p[u].x, dy = qy - p[u].y, dz = qz - p[u].z;
                    const float d2 = (dx * dx + dy * dy) + dz * dz;
                    if (d2 < r2max) f(j + u, dx, dy, dz, d2);
                }
            }
        }
    });
}

// Lane exchange within groups of four lanes (DPP quad_perm): lane l reads lane
// (l & ~3) + sel[l & 3] of v.
#define QPERM(a, b, c, d) ((a) | ((b) << 2) | ((c) << 4) | ((d) << 6))
template <int CTRL>
__device__ inline int quad_perm(int v) {
    return __builtin_amdgcn_mov_dpp(v, CTRL, 0xF, 0xF, true);
}
template <int CTRL>
__device__ inline float quad_perm(float v) {
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), CTRL, 0xF, 0xF, true));
}

// Maximum of v over the 16 lanes of each DPP row, in every lane of the row (all lanes of the
// wave active): two quad exchanges, then rotations by 4 and 8 lanes within the row.
__device__ inline int row16_max(int v) {
    v = max(v, quad_perm<QPERM(1, 0, 3, 2)>(v));
    v = max(v, quad_perm<QPERM(2, 3, 0, 1)>(v));
    v = max(v, __builtin_amdgcn_mov_dpp(v, 0x124, 0xF, 0xF, true));  // row_ror:4
    v = max(v, __builtin_amdgcn_mov_dpp(v, 0x128, 0xF, 0xF, true));  // row_ror:8
    return v;
}

// rows_visit for callers that need only the accepted indices, in the same order: each row
// range is tested in windows of 32 candidates into an accept bitmask, and f(j) is then
// called for the accepted candidates in ascending order.  The per-candidate work of a wave
// no longer includes the caller's (divergent) accept code.  Candidate t of a window of n
// enters the mask at bit n4 - 1 - t (m = 2m + accept; n4 = n rounded up to the batch), so
// each test costs a compare and an add-with-carry; a window reads up to XS_PAD - 1 positions
// past its range, which xs must hold (padding) and which are masked off.  The lanes of a wave
// are consecutive particles of one or two cell rows, so the candidates of neighbouring lanes
// are neighbouring positions and the per-lane loads share cache lines.
template <class F>
__device__ inline void rows_visit_ids(const GridWorld& g, const int2* __restrict__ table,
                                      const int* __restrict__ cend,
                                      const float4* __restrict__ xs, float qx, float qy, float qz,
                                      float r, float r2max, F f) {
    rows_for_each(g, table, cend, qx, qy, qz, r, [&](int lo, int hi) {
        for (int base = lo; base < hi; base += 32) {
            const int n = min(hi - base, 32);
            const float4* __restrict__ xb = xs + base;
            uint32_t m = 0u;
            int n4 = 0;
            for (; n4 < n; n4 += RV_BATCH) {
                float4 p[RV_BATCH];
#pragma unroll
                for (int v = 0; v < RV_BATCH; ++v) p[v] = xb[n4 + v];
#pragma unroll
                for (int v = 0; v < RV_BATCH; ++v) {
                    const float dx = qx - p[v].x, dy = qy - p[v].y, dz = qz - p[v].z;
                    const float d2 = (dx * dx + dy * dy) + dz * dz;
                    m = m + m + (d2 < r2max ? 1u : 0u);
                }
            }
            m &= ~((1u << (n4 - n)) - 1u);  // the tail past the row
            while (m != 0u) {
                const int bit = 31 - __builtin_clz(m);
                m ^= 1u << bit;
                f(base + (n4 - 1 - bit));
            }
        }
    });
}

// Gather helper: for each list entry of s in slot order, load(j) fetches the neighbour data
// (B = 4 or 8 entries, one or two quads, at a time) and use(data, j) accumulates in slot
// order.  Missing entries of the last batch repeat the first one and are not used.
// quad(q) returns list quad q of s.
template <class T, int B, class Quad, class Load, class Use>
__device__ inline void for_each_nbr_q(int n, Quad quad, Load load, Use use) {
    static_assert(B == 4 || B == 8, "batch of 4 or 8");
    for (int k = 0; k < n; k += B) {
        const int m = n - k;
        const int4 q0 = quad(k >> 2);
        const int j0 = q0.x, j1 = m > 1 ? q0.y : j0, j2 = m > 2 ? q0.z : j0, j3 = m > 3 ? q0.w : j0;
        if (B == 4) {
            const T a0 = load(j0), a1 = load(j1), a2 = load(j2), a3 = load(j3);
            use(a0, j0);
            if (m > 1) use(a1, j1);
            if (m > 2) use(a2, j2);
            if (m > 3) use(a3, j3);
        } else {
            const int4 q1 = quad((k >> 2) + (m > 4 ? 1 : 0));
            const int j4 = m > 4 ? q1.x : j0, j5 = m > 5 ? q1.y : j0, j6 = m > 6 ? q1.z : j0,
                      j7 = m > 7 ? q1.w : j0;
            const T a0 = load(j0), a1 = load(j1), a2 = load(j2), a3 = load(j3);
            const T a4 = load(j4), a5 = load(j5), a6 = load(j6), a7 = load(j7);
            use(a0, j0);
            if (m > 1) use(a1, j1);
            if (m > 2) use(a2, j2);
            if (m > 3) use(a3, j3);
            if (m > 4) use(a4, j4);
            if (m > 5) use(a5, j5);
            if (m > 6) use(a6, j6);
            if (m > 7) use(a7, j7);
        }
    }
}

template <class T, int B, class Load, class Use>
__device__ inline void for_each_nbr(const SphDev& D, int s, Load load, Use use) {
    const int4* __restrict__ L = D.nbr4 + s;
    const size_t stride = (size_t)D.Nf;
    for_each_nbr_q<T, B>(D.ncnt[s], [&](int q) { return L[(size_t)q * stride]; }, load, use);
}

// Neighbour lists fused with the density pass, in two phases so that the expensive
// per-neighbour work is not paid for every candidate of a divergent wave:
//   phase 1  list:  j != i, |x_i - x_j|^2 < H^2 (FP32 predicate), candidate order
//   phase 2  rho_i = m_i W(0) + sum_j m_j W_ij + psi sum_b W_ib   (Neumaier, list then
//                    boundary order)
//            Gf = sum_j m_j gradW_ij,  S = sum_j m_j |gradW_ij|^2,  Gb = sum_b gradW_ib
//            G = Gf + psi Gb,          a_ii = -dt^2 ((G.G + m_i S) / rho_i^2)
// a_ii is the diagonal of the operator (A p)_i of k_sph_pupdate.  rho is also stored in
// vs[s].w as m_i / rho_i, the factor the viscosity pass takes of each neighbour, so that it
// fetches it with the velocity and divides once per particle instead of once per pair.
// End of the density pass for sorted particle s: rho = (fluid + boundary compensated sum),
// G = Gf + psi Gb, a_ii = -dt^2 ((G.G + m_i S) / rho^2), and the per-particle outputs.
__device__ inline void density_finish(const SphDev& D, const PhysWorld& P, int s, float mi, float sum, float cmp,
                                      float gfx, float gfy, float gfz, float S, float gbx, float gby, float gbz,
                                      float mr2, int nb, float dt2, float kvb = 0.f) {
    const float rho = sum + cmp;
    const float Gx = gfx + P.psi * gbx, Gy = gfy + P.psi * gby, Gz = gfz + P.psi * gbz;
    const float num = dot3(Gx, Gy, Gz, Gx, Gy, Gz) + mi * S;
    D.rho[s] = rho;
    reinterpret_cast<float*>(D.vs + s)[3] = mi / rho;
    D.aux[s] = make_float2((float)nb, nb > 0 ? sqrtf(mr2) * P.invH : 1.0f);
    D.aii[s] = -(dt2 * (num / (rho * rho)));
    D.gb[s] = make_float4(gbx, gby, gbz, (float)nb);
    if (D.kvb) D.kvb[s] = kvb;
}

// Six waves per SIMD (77-81 VGPRs, no scratch) hide more of the gathers' latency than five.
// Variants (the launch configuration picks one; every variant gives the same bits): LQ list
// quads per particle staged in LDS, LQ * blockDim * 16 bytes per workgroup, which also sets
// how many workgroups of 256 share a CU (six up to LQ = 6, five at LQ = 8).  A register
// target below six waves does not change the allocation, so occupancy is tuned through LQ.
// SKIN = false: the candidate scan of every step gives the list (radius H).  SKIN = true:
// candidates within the skin radius Rs are kept between steps; a world that rebuilds scans
// them anew (sorted order of this build), and every world then filters its kept candidates to
// the list (d^2 < H^2, same FP32 predicate, same ascending order) while adding the density
// terms, so the list equals the one a scan with radius H in this sorted order would give.
template <bool SKIN, int LQ>
__global__ __launch_bounds__(CHUNK) __attribute__((amdgpu_waves_per_eu(6, 8))) void k_sph_nbr_density(SphDev D, float dt2) {
    int w, s;
    if (!chunk_slot_sub(D, D.nsub, &w, &s)) return;
    const GridWorld g = D.gw[w];
    const PhysWorld P = D.phys[w];
    const float4 xi = D.xs[s];
    if (xi.w != D.xs[D.w_fstart[w]].w) atomicOr(&D.mixed[w], 1);
    const int cx = cell_coord(xi.x, g.lo[0], g.inv_w, g.dim[0], 0);
    const int cy = cell_coord(xi.y, g.lo[1], g.inv_w, g.dim[1], 0);
    const int cz = cell_coord(xi.z, g.lo[2], g.inv_t, g.dim[2], 0);
    const int Nf = D.Nf, cap = D.cap;
    int4* __restrict__ out = D.nbr4 + s;
    int c = 0;
    int4 quad = make_int4(0, 0, 0, 0);
    // The first LQ quads of a particle wait in LDS (lq[q * blockDim + t]): each entry is
    // written to its slot there as it is accepted, and the quads are stored after the search,
    // when the whole wave writes quad q of its 64 particles at once (storing each quad when it
    // completes would leave lines partly written and evicted several times).  Entries past
    // the LQ quads shift into a register quad from .w, which is stored when its fourth entry
    // arrives.  Slots past the count hold whatever was there; readers stop at the count.
    extern __shared__ int4 lq[];
    int* __restrict__ lqi = reinterpret_cast<int*>(lq);
    const int t = (int)threadIdx.x, bd = (int)blockDim.x;
    auto append = [&](int j) {
        if (j == s) return;
        if (c < 4 * LQ) {
            lqi[((c >> 2) * bd + t) * 4 + (c & 3)] = j;
        } else if (c < cap) {
            quad = make_int4(quad.y, quad.z, quad.w, j);
            if ((c & 3) == 3) out[(size_t)(c >> 2) * Nf] = quad;
        }
        ++c;
    };
    // the last, partial quad past the LDS, the LDS-staged quads, the count and the per-wave
    // maximum
    auto finish_list = [&]() {
        const int cl = c < cap ? c : cap;
        const int u = cl & 3;  // entries of the last, partial quad, at its top
        if (cl > 4 * LQ) {
            if (u == 1) out[(size_t)(cl >> 2) * Nf] = make_int4(quad.w, 0, 0, 0);
            if (u == 2) out[(size_t)(cl >> 2) * Nf] = make_int4(quad.z, quad.w, 0, 0);
            if (u == 3) out[(size_t)(cl >> 2) * Nf] = make_int4(quad.y, quad.z, quad.w, 0);
        }
        const int nq = min((cl + 3) >> 2, LQ);
        for (int q = 0; q < nq; ++q) out[(size_t)q * Nf] = lq[q * bd + t];
        D.ncnt[s] = cl;
        if (c > cap) atomicOr(&D.flags[0], 1);
        // Largest count of the step: each wave stores its own and k_sph_report takes the
        // maximum.  (A per-lane atomicMax on one address compiles to a scalar loop over the
        // wave's lanes, about a tenth of this kernel's time; one same-address atomic per wave
        // still costs several percent across the eight XCDs.)
        const int wv = (int)(blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6));
        if (__ballot(1) == ~0ull) {
            const int r = row16_max(c);
            const int mx = max(max(__builtin_amdgcn_readlane(r, 0), __builtin_amdgcn_readlane(r, 16)),
                               max(__builtin_amdgcn_readlane(r, 32), __builtin_amdgcn_readlane(r, 48)));
            if (__lane_id() == 0u) D.wmax[wv] = mx;
        } else {
            atomicMax(&D.wmax[wv], c);  // a wave with inactive lanes (the end of a world)
        }
    };

    float sum = xi.w * P.W0, cmp = 0.0f;
    float gfx = 0.f, gfy = 0.f, gfz = 0.f, S = 0.f;
    auto dens = [&](float mj, float dx, float dy, float dz, float r2) {
        float W, gf;
        kern_wg(P, r2, &W, &gf);
        neumaier(sum, cmp, mj * W);
        const float gx = gf * dx, gy = gf * dy, gz = gf * dz;
        gfx += mj * gx;
        gfy += mj * gy;
        gfz += mj * gz;
        S += mj * dot3(gx, gy, gz, gx, gy, gz);
    };
    if constexpr (!SKIN) {
        rows_visit_ids(g, D.ftable, D.fcend, D.xs, xi.x, xi.y, xi.z, P.H, P.H2, append);
        finish_list();
        // the first LQ quads are read back from LDS
        for_each_nbr_q<float4, 8>(
            c < cap ? c : cap,
            [&](int q) { return q < LQ ? lq[q * bd + t] : out[(size_t)q * Nf]; },
            [&](int j) { return D.xs[j]; },
            [&](const float4& xj, int) {
                const float dx = xi.x - xj.x, dy = xi.y - xj.y, dz = xi.z - xj.z;
                dens(xj.w, dx, dy, dz, (dx * dx + dy * dy) + dz * dz);
            });
    } else {
        int4* __restrict__ sk = D.snb4 + s;
        int ns;
        if (__builtin_amdgcn_readfirstlane(D.rb[w]) != 0) {
            const int capS = D.capS;
            int cs = 0;
            int4 sq = make_int4(0, 0, 0, 0);
            auto keep = [&](int j) {
                if (j == s) return;
                if (cs < capS) {
                    sq = make_int4(sq.y, sq.z, sq.w, j);
                    if ((cs & 3) == 3) sk[(size_t)(cs >> 2) * Nf] = sq;
                }
                ++cs;
            };
            const float4 xq = D.xbs[s];
            rows_visit_ids(g, D.ftable, D.fcend, D.xbs, xq.x, xq.y, xq.z, P.Rs, P.Rs2, keep);
            ns = cs < capS ? cs : capS;
            const int u = ns & 3;
            if (u == 1) sk[(size_t)(ns >> 2) * Nf] = make_int4(sq.w, 0, 0, 0);
            if (u == 2) sk[(size_t)(ns >> 2) * Nf] = make_int4(sq.z, sq.w, 0, 0);
            if (u == 3) sk[(size_t)(ns >> 2) * Nf] = make_int4(sq.y, sq.z, sq.w, 0);
            D.scnt[s] = ns;
            if (cs > capS) atomicOr(&D.flags[0], 1);
        } else {
            ns = D.scnt[s];
        }
        // the thread reads back the quads it stored above (same address, program order)
        for_each_nbr_q<float4, 8>(
            ns, [&](int q) { return sk[(size_t)q * Nf]; }, [&](int j) { return D.xs[j]; },
            [&](const float4& xj, int j) {
                const float dx = xi.x - xj.x, dy = xi.y - xj.y, dz = xi.z - xj.z;
                const float r2 = (dx * dx + dy * dy) + dz * dz;
                if (r2 < P.H2) {
                    const bool within = c < cap;  // as the list: only its first cap entries count
                    append(j);
                    if (within) dens(xj.w, dx, dy, dz, r2);
                }
            });
        finish_list();
    }

    // Particles whose cell has boundary particles in its 3x3x5 block are queued with their
    // fluid sums; k_sph_bnd_density adds the boundary terms in full waves (the visit is
    // per-lane and a wave holds few such particles).  The others finish here.
    const int cell = g.cell_off + (cz * g.dim[1] + cy) * g.dim[0] + cx;
    const bool breach = D.breach[cell] != 0 || (D.kbreach != nullptr && D.kbreach[cell] != 0);
    const uint64_t bm = __ballot(breach);
    if (bm) {
        const int lane = (int)__lane_id();
        const int leader = __ffsll((unsigned long long)bm) - 1;
        int base = 0;
        if (lane == leader) base = atomicAdd(D.bq_count, __popcll(bm));
        base = __shfl(base, leader, 64);
        if (breach) {
            const int slot = base + __popcll(bm & ((1ull << lane) - 1ull));
            D.bq_id[slot] = make_int2(s, w);
            D.bq_a[slot] = make_float4(sum, cmp, gfx, gfy);
            D.bq_b[slot] = make_float2(gfz, S);
            return;
        }
    }
    density_finish(D, P, s, xi.w, sum, cmp, gfx, gfy, gfz, S, 0.f, 0.f, 0.f, P.H2, 0, dt2);
}

#define BND_BATCH 8
// Boundary terms of queued particle i (sorted index id.x in world w), continuing its
// compensated density sum in the boundary candidate order, then the same finish as the
// neighbour kernel.  MOVING: moving boundary particles exist (visited after the static ones).
template <bool MOVING>
__device__ inline void bnd_visit(const SphDev& D, int i, int s, int w, float dt2) {
    const GridWorld g = D.gw[w];
    const PhysWorld P = D.phys[w];
    const float4 xi = D.xs[s];
    const float4 a = D.bq_a[i];
    const float2 b = D.bq_b[i];
    float sum = a.x, cmp = a.y;
    float gbx = 0.f, gby = 0.f, gbz = 0.f, mr2 = P.H2;
    int nb = 0;
    rows_visit<BND_BATCH>(g, D.btable, D.bcend, D.bxs, xi.x, xi.y, xi.z, P.H, P.H2,
               [&](int, float dx, float dy, float dz, float r2) {
                   float W, gf;
                   kern_wg(P, r2, &W, &gf);
                   neumaier(sum, cmp, P.psi * W);
                   gbx += gf * dx;
                   gby += gf * dy;
                   gbz += gf * dz;
                   mr2 = fminf(mr2, r2);
                   ++nb;
               });
    // moving boundary particles after the static ones, same terms, plus v_b . gradW_ib
    float kvb = 0.f;
    if constexpr (MOVING)
        rows_visit<BND_BATCH>(g, D.ktable, D.kcend, D.kxs, xi.x, xi.y, xi.z, P.H, P.H2,
                   [&](int j, float dx, float dy, float dz, float r2) {
                       float W, gf;
                       kern_wg(P, r2, &W, &gf);
                       neumaier(sum, cmp, P.psi * W);
                       const float gx = gf * dx, gy = gf * dy, gz = gf * dz;
                       gbx += gx;
                       gby += gy;
                       gbz += gz;
                       const float4 vb = D.kvs[j];
                       kvb += dot3(vb.x, vb.y, vb.z, gx, gy, gz);
                       mr2 = fminf(mr2, r2);
                       ++nb;
                   });
    density_finish(D, P, s, xi.w, sum, cmp, a.z, a.w, b.x, b.y, gbx, gby, gbz, mr2, nb, dt2, kvb);
}

// Boundary terms of the queued particles in full waves.  Grid-stride over the queue (its
// length is known on the device only).
template <bool MOVING>
__global__ __launch_bounds__(256) void k_sph_bnd_density(SphDev D, float dt2) {
    const int n = *D.bq_count;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const int2 id = D.bq_id[i];
        bnd_visit<MOVING>(D, i, id.x, id.y, dt2);
    }
}

struct XV {
    float4 x, v;  // v.w = rho
};
struct XS {
    float4 x;
    float s;
};
struct XA {
    float4 x, a;
};

// Viscosity and advected velocity:
//   a_visc = 10 nu sum_j (m_j/rho_j) ((v_ij . x_ij) / (r^2 + 0.01 H^2)) gradW_ij
//   v_adv  = v + dt (g + a_visc)
// vadv.w carries the particle's p0 / rho^2 of the previous-pressure warm start, so that
// the source pass fetches it with v_adv (one load per neighbour instead of three).
__device__ inline void advect_one(const SphDev& D, const PhysWorld& P, int s, float dt, int warm) {
    const float4 xi = D.xs[s];
    const float4 vi = D.vs[s];
    float ax = 0.f, ay = 0.f, az = 0.f, dv = 0.f;
    for_each_nbr<XV, 8>(
        D, s, [&](int j) { return XV{D.xs[j], D.vs[j]}; },
        [&](const XV& o, int) {
            const float dx = xi.x - o.x.x, dy = xi.y - o.x.y, dz = xi.z - o.x.z;
            const float r2 = (dx * dx + dy * dy) + dz * dz;
            const float gf = kern_gf(P, r2);
            const float vx = dot3(vi.x - o.v.x, vi.y - o.v.y, vi.z - o.v.z, dx, dy, dz);
            const float c = o.v.w * (vx / (r2 + P.eps2));  // (m_j / rho_j) (v_ij . x_ij) / (r^2 + eps^2)
            ax += c * (gf * dx);
            ay += c * (gf * dy);
            az += c * (gf * dz);
            dv += o.x.w * (gf * vx);
        });
    const float avx = P.visc10 * ax, avy = P.visc10 * ay, avz = P.visc10 * az;
    // avisc.w: sum_j m_j (v_i - v_j) . gradW_ij + psi (v_i . Gb_i - sum_b v_b . gradW_ib), the
    // velocity divergence term of the source with the velocities at the step's start (an input
    // of the learned warm start, which runs before the source pass)
    const float4 gb = D.gb[s];
    dv += P.psi * (dot3(vi.x, vi.y, vi.z, gb.x, gb.y, gb.z) - (D.kvb ? D.kvb[s] : 0.f));
    D.avisc[s] = make_float4(avx, avy, avz, dv);
    const float rho = D.rho[s];
    const float pr0 = warm == NSPH_WARM_PREV ? (0.5f * D.pws[s]) / (rho * rho) : 0.0f;
    D.vadv[s] = make_float4(vi.x + dt * (P.g[0] + avx), vi.y + dt * (P.g[1] + avy),
                            vi.z + dt * (P.g[2] + avz), pr0);
}

__global__ __launch_bounds__(CHUNK) void k_sph_advect(SphDev D, float dt, int warm) {
    int w, s;
    if (!chunk_slot_sub(D, D.nsub, &w, &s)) return;
    advect_one(D, D.phys[w], s, dt, warm);
}

// Source term and warm start:
//   s_i = (rho0 - rho_i) - dt (sum_j m_j (v_adv_i - v_adv_j) . gradW_ij + psi v_adv_i . Gb_i)
//   p_i = 0 | 0.5 p_prev | learned;  pr_i = p_i / rho_i^2
// Source term fused with the first pressure acceleration: one pass over the list gives s_i
// and a^p_i of iteration 0 from pr_j = p0_j / rho_j^2 (vadv.w, formed with the same operations
// as pri here by k_sph_advect for the zero and previous-pressure warm starts, or by
// k_sph_mlp_warm, which runs between the two passes, for the learned one), the same
// operations, operands and order as a source pass followed by k_sph_pacc.
__device__ inline void source_pacc_one(const SphDev& D, const PhysWorld& P, int s, float dt, int warm) {
    const float4 xi = D.xs[s];
    const float4 vi = D.vadv[s];
    const float rho = D.rho[s];
    const float p0 = warm == NSPH_WARM_PREV ? 0.5f * D.pws[s] : (warm == NSPH_WARM_MLP ? D.p0[s] : 0.f);
    const float pri = p0 / (rho * rho);
    float sum = 0.f, ax = 0.f, ay = 0.f, az = 0.f;
    for_each_nbr<XA, 8>(
        D, s, [&](int j) { return XA{D.xs[j], D.vadv[j]}; },
        [&](const XA& o, int) {
            const float dx = xi.x - o.x.x, dy = xi.y - o.x.y, dz = xi.z - o.x.z;
            const float r2 = (dx * dx + dy * dy) + dz * dz;
            const float gf = kern_gf(P, r2);
            sum += o.x.w * dot3(vi.x - o.a.x, vi.y - o.a.y, vi.z - o.a.z, gf * dx, gf * dy, gf * dz);
            const float prj = o.a.w;
            const float c = o.x.w * (pri + prj);
            ax += c * (gf * dx);
            ay += c * (gf * dy);
            az += c * (gf * dz);
        });
    const float4 gb = D.gb[s];
    // boundary: psi sum_b (v_i - v_b) . gradW_ib (v_b = 0 for static boundary particles)
    sum += P.psi * (dot3(vi.x, vi.y, vi.z, gb.x, gb.y, gb.z) - (D.kvb ? D.kvb[s] : 0.f));
    D.src[s] = (P.rho0 - rho) - dt * sum;
    D.p[s] = p0;
    D.p0[s] = p0;
    D.xpr[s] = make_float4(xi.x, xi.y, xi.z, pri);
    const float cb = P.psi * pri;
    store3(D.a3, s, -(ax + cb * gb.x), -(ay + cb * gb.y), -(az + cb * gb.z));
}

__global__ __launch_bounds__(CHUNK) void k_sph_source_pacc(SphDev D, float dt, int warm) {
    int w, s;
    if (!chunk_slot_sub(D, D.nsub, &w, &s)) return;
    source_pacc_one(D, D.phys[w], s, dt, warm);
}

// Pressure acceleration  a^p_i = -( sum_j m_j (pr_i + pr_j) gradW_ij + (psi pr_i) Gb_i ).
// In a world whose fluid particles all have one mass, m_j = m_i and the gather loads only
// xpr[j]; otherwise it also loads xs[j] for m_j.  Both give the same operations and values.
__device__ inline void press_acc(const SphDev& D, const PhysWorld& P, int w, int s, float4 xi,
                                 float* ox, float* oy, float* oz) {
    const float pri = D.xpr[s].w;
    float ax = 0.f, ay = 0.f, az = 0.f;
    auto term = [&](float ox_, float oy_, float oz_, float mj, float prj) {
        const float dx = xi.x - ox_, dy = xi.y - oy_, dz = xi.z - oz_;
        const float r2 = (dx * dx + dy * dy) + dz * dz;
        const float gf = kern_gf_p(P, r2);
        const float c = mj * (pri + prj);
        ax += c * (gf * dx);
        ay += c * (gf * dy);
        az += c * (gf * dz);
    };
    if (__builtin_amdgcn_readfirstlane(D.mixed[w]) == 0) {
        const float m = xi.w;
        for_each_nbr<float4, 8>(
            D, s, [&](int j) { return D.xpr[j]; },
            [&](const float4& o, int) { term(o.x, o.y, o.z, m, o.w); });
    } else {
        for_each_nbr<XS, 8>(
            D, s, [&](int j) { return XS{D.xs[j], D.xpr[j].w}; },
            [&](const XS& o, int) { term(o.x.x, o.x.y, o.x.z, o.x.w, o.s); });
    }
    const float4 gb = D.gb[s];
    const float cb = P.psi * pri;
    *ox = -(ax + cb * gb.x);
    *oy = -(ay + cb * gb.y);
    *oz = -(az + cb * gb.z);
}

__device__ inline void pacc_one(const SphDev& D, const PhysWorld& P, int w, int s) {
    const float4 xi = D.xs[s];
    float ax, ay, az;
    press_acc(D, P, w, s, xi, &ax, &ay, &az);
    store3(D.a3, s, ax, ay, az);
}

__global__ __launch_bounds__(CHUNK) void k_sph_pacc(SphDev D) {
    int w, s;
    const bool ok = chunk_slot_sub(D, D.nsub, &w, &s);
    if (D.conv[w]) return;
    if (!ok) return;
    pacc_one(D, D.phys[w], w, s);
}

// (A p)_i = dt^2 ( sum_j m_j (a^p_i - a^p_j) . gradW_ij + psi a^p_i . Gb_i )
// e_i     = max((A p)_i - s_i, 0)   (compression of the predicted density, kg/m^3)
// p_i    <- a_ii != 0 ? max(p_i + omega (s_i - (A p)_i) / a_ii, 0) : 0
// partial[chunk] = sum of e_i over the chunk (FP64, fixed LDS tree).
struct XF3 {
    float4 x;  // x, y, z, m
    float3 a;  // a^p
};

__device__ inline double pupdate_one(const SphDev& D, const PhysWorld& P, int w, int s, float dt2, float omega) {
    const float4 xi = D.xs[s];
    const float3 ai = load3(D.a3, s);
    float sum = 0.f;
    for_each_nbr<XF3, 8>(
        D, s, [&](int j) { return XF3{D.xs[j], load3(D.a3, j)}; },
        [&](const XF3& o, int) {
            const float dx = xi.x - o.x.x, dy = xi.y - o.x.y, dz = xi.z - o.x.z;
            const float r2 = (dx * dx + dy * dy) + dz * dz;
            const float gf = kern_gf_p(P, r2);
            sum += o.x.w * dot3(ai.x - o.a.x, ai.y - o.a.y, ai.z - o.a.z, gf * dx, gf * dy, gf * dz);
        });
    const float4 gb = D.gb[s];
    sum += P.psi * dot3(ai.x, ai.y, ai.z, gb.x, gb.y, gb.z);
    const float Ap = dt2 * sum;
    const float si = D.src[s];
    const double e = (double)fmaxf(Ap - si, 0.0f);
    const float a = D.aii[s];
    float pn = 0.f;
    if (a != 0.0f) pn = fmaxf(D.p[s] + omega * (si - Ap) / a, 0.0f);
    D.p[s] = pn;
    const float rho = D.rho[s];
    D.xpr[s] = make_float4(xi.x, xi.y, xi.z, pn / (rho * rho));
    return e;
}

// Sum of red[t] over each run of CHUNK threads (t = 0 .. CHUNK - 1 of the run), left in the
// run's first entry: the fixed LDS tree of the chunk partials.  Every thread of the workgroup
// calls it.
__device__ inline void chunk_tree(double* red, int t) {
    const int u = t & (CHUNK - 1);
    __syncthreads();
    for (int off = CHUNK / 2; off > 0; off >>= 1) {
        if (u < off) red[t] += red[t + off];
        __syncthreads();
    }
}

__global__ __launch_bounds__(CHUNK) void k_sph_pupdate(SphDev D, float dt2, float omega) {
    __shared__ double red[CHUNK];
    int w, s;
    const bool ok = chunk_slot(D, &w, &s);
    if (D.conv[w]) return;
    const int t = threadIdx.x;
    red[t] = ok ? pupdate_one(D, D.phys[w], w, s, dt2, omega) : 0.0;
    chunk_tree(red, t);
    if (t == 0) D.partial[xcd_swizzle(blockIdx.x, gridDim.x)] = red[0];
}

// One wave per world: err_w = (sum of its chunk partials) / (n_w rho0); count the update;
// decide convergence. act[it & 1] counts the worlds still iterating after iteration `it`.
// world_reduce: one wave (lane = 0 .. 63) reduces world w's chunk partials and, in lane 0,
// counts the update, decides convergence, stores it in conv[w] and returns it (0 = iterate on,
// 1 = converged, 2 = stopped at max_iters).
__device__ inline int world_reduce(const SphDev& D, int w, int lane, double tol, int min_iters, int max_iters) {
    const int first = D.w_cfirst[w], nch = D.w_nchunks[w];
    double acc = 0.0;
    int k = lane;
    // lane sums partials lane, lane+64, ... in order; sixteen (then eight) loads are issued
    // before the adds
    for (; k + 15 * 64 < nch; k += 16 * 64) {
        double v[16];
#pragma unroll
        for (int u = 0; u < 16; ++u) v[u] = D.partial[first + k + u * 64];
#pragma unroll
        for (int u = 0; u < 16; ++u) acc += v[u];
    }
    for (; k + 7 * 64 < nch; k += 8 * 64) {
        double v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) v[u] = D.partial[first + k + u * 64];
#pragma unroll
        for (int u = 0; u < 8; ++u) acc += v[u];
    }
    for (; k < nch; k += 64) acc += D.partial[first + k];
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
    if (lane == 0) {
        const double err = acc / ((double)D.w_nf[w] * (double)D.phys[w].rho0);
        const int n_it = D.iters[w] + 1;
        D.iters[w] = n_it;
        D.err[w] = err;
        int c = 0;
        if (n_it >= min_iters && err <= tol) c = 1;
        else if (n_it >= max_iters) c = 2;
        D.conv[w] = c;
        return c;
    }
    return 0;
}

__global__ void k_sph_reduce(SphDev D, int W, int it, double tol, int min_iters, int max_iters) {
    const int lane = threadIdx.x & 63;
    const int w = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    if (blockIdx.x == 0 && threadIdx.x == 0) D.act[(it + 1) & 1] = 0;
    if (w >= W) return;
    if (D.conv[w]) return;
    const int c = world_reduce(D, w, lane, tol, min_iters, max_iters);
    if (lane == 0 && c == 0) atomicAdd(&D.act[it & 1], 1);
}

// Step summary over the worlds with fluid (one workgroup): max and sum of the iteration
// counts, worlds not converged, max density error, the neighbour and finiteness flags, and
// (act_slot >= 0) the still-iterating count act[act_slot].
struct RepDev {
    int act, max_it, sum_it, unconv, overflow, maxnbr, nonfinite, seq;  // seq: written last
    int searched;
    double max_err;
};
__global__ __launch_bounds__(1024) void k_sph_report(SphDev D, int W, int act_slot, int seq, RepDev* __restrict__ out) {
    constexpr int NT = 1024;
    __shared__ int smax[NT], ssum[NT], sun[NT], snb[NT], ssr[NT];
    __shared__ double serr[NT];
    const int t = threadIdx.x;
    int mx = 0, sm = 0, un = 0, nb = 0, sr = 0;
    double er = 0.0;
    // per-wave list maxima, four per load (nwave = 4 per chunk), several loads in flight
    const int4* __restrict__ wm = reinterpret_cast<const int4*>(D.wmax);
    const int n4 = D.nwave >> 2;
#pragma unroll 4
    for (int i = t; i < n4; i += NT) {
        const int4 v = wm[i];
        nb = max(nb, max(max(v.x, v.y), max(v.z, v.w)));
    }
    for (int w = t; w < W; w += NT) {
        if (D.w_nf[w] == 0) continue;
        const int it = D.iters[w];
        mx = max(mx, it);
        sm += it;
        un += D.conv[w] != 1 ? 1 : 0;
        er = fmax(er, D.err[w]);
        sr += D.rb[w] != 0 ? 1 : 0;
    }
    smax[t] = mx;
    ssum[t] = sm;
    sun[t] = un;
    snb[t] = nb;
    ssr[t] = sr;
    serr[t] = er;
    __syncthreads();
    for (int off = NT / 2; off > 0; off >>= 1) {
        if (t < off) {
            smax[t] = max(smax[t], smax[t + off]);
            ssum[t] += ssum[t + off];
            sun[t] += sun[t + off];
            snb[t] = max(snb[t], snb[t + off]);
            ssr[t] += ssr[t + off];
            serr[t] = fmax(serr[t], serr[t + off]);
        }
        __syncthreads();
    }
    if (t == 0) {
        RepDev r;
        r.act = act_slot >= 0 ? D.act[act_slot] : 0;
        r.max_it = smax[0];
        r.sum_it = ssum[0];
        r.unconv = sun[0];
        r.overflow = D.flags[0];
        r.maxnbr = snb[0];
        r.nonfinite = D.flags[2];
        r.seq = seq - 1;
        r.searched = ssr[0];
        r.max_err = serr[0];
        *out = r;  // pinned host memory
        __threadfence_system();
        // the host waits for this sequence number; the fields above are visible before it
        *reinterpret_cast<volatile int*>(&out->seq) = seq;
        __threadfence_system();
    }
}

// Pressure acceleration from the final p and the SPH force f = m (a_visc + a^p), sorted
// order: fp[s] = (f, p).  With a guard (the still-iterating count of the last iteration),
// the step's end is enqueued before the host knows whether every world has converged and
// does nothing unless it has (the host then iterates on and enqueues it again).
__device__ inline void final_one(const SphDev& D, const PhysWorld& P, int w, int s) {
    const float4 xi = D.xs[s];
    float ax, ay, az;
    press_acc(D, P, w, s, xi, &ax, &ay, &az);
    const float4 av = D.avisc[s];
    const float m = xi.w;
    D.fp[s] = make_float4(m * (av.x + ax), m * (av.y + ay), m * (av.z + az), D.p[s]);
}

__global__ __launch_bounds__(CHUNK) void k_sph_final(SphDev D, const int* __restrict__ guard) {
    if (guard && *guard != 0) return;
    int w, s;
    if (!chunk_slot_sub(D, D.nsub, &w, &s)) return;
    final_one(D, D.phys[w], w, s);
}

// The reference integrator (solver.py:24-46) on fluid particles, in particle order (fluid
// index k, sorted position inv[k]), so that the particle-order outputs are written
// coalesced:  v1 = v0 + (f w + g) dt;  |v1| <= v_max;  x1 = x0 + v1 dt.
struct StepEndArgs {
    const int* inv;     // sorted position of fluid index k
    const int* fid;     // particle id of fluid index k
    const int* fworld;  // world of fluid index k
    const int* oid;     // particle ids of the non-fluid particles
    const float* inv_mass;
    const int* flags;
    const int* world;  // model.particle_world (NULL: all world 0)
    int W, nf, no;
    float dt, v_max;
    const float* q_in;
    const float* qd_in;
    const float* f_in;  // caller's forces (NULL: none)
    float* q_out;
    float* qd_out;
    float* f_out;  // SPH force out (NULL: not written)
    float* p_out;  // state_out.sph.pressure (NULL: none)
    float* xb_out;  // state_out.sph.search_q (NULL: none)
    float* pprev;   // solver's copy of the pressure, fluid index order
};

__device__ inline void integrate_one(const SphDev& D, const StepEndArgs& A, int k) {
    const int* __restrict__ fid = A.fid;
    const float* __restrict__ inv_mass = A.inv_mass;
    const float* __restrict__ f_in = A.f_in;
    float* __restrict__ q_out = A.q_out;
    float* __restrict__ qd_out = A.qd_out;
    float* __restrict__ f_out = A.f_out;
    float* __restrict__ p_out = A.p_out;
    float* __restrict__ xb_out = A.xb_out;
    const float dt = A.dt, v_max = A.v_max;
    const int s = A.inv[k];
    const int id = fid[k];
    const PhysWorld P = D.phys[A.fworld[k]];
    const float4 f = D.fp[s];
    // x0 and v0 from state_in in particle order (coalesced): the sorted xs[s] and vs[s] hold
    // copies of the same values, which would be gathered through the permutation
    const float4 xi = make_float4(A.q_in[3 * id], A.q_in[3 * id + 1], A.q_in[3 * id + 2], 0.f);
    const float4 v0 = make_float4(A.qd_in[3 * id], A.qd_in[3 * id + 1], A.qd_in[3 * id + 2], 0.f);
    const float wi = inv_mass[id];
    // total force: the caller's (state_in.particle_f) plus the SPH force
    float tx = f.x, ty = f.y, tz = f.z;
    if (f_in) {
        tx = f_in[3 * id] + f.x;
        ty = f_in[3 * id + 1] + f.y;
        tz = f_in[3 * id + 2] + f.z;
    }
    float v1x = v0.x + (tx * wi + P.g[0]) * dt;
    float v1y = v0.y + (ty * wi + P.g[1]) * dt;
    float v1z = v0.z + (tz * wi + P.g[2]) * dt;
    const float mag = sqrtf(dot3(v1x, v1y, v1z, v1x, v1y, v1z));
    if (mag > v_max) {
        const float sc = v_max / mag;
        v1x *= sc;
        v1y *= sc;
        v1z *= sc;
    }
    const float x1 = xi.x + v1x * dt, y1 = xi.y + v1y * dt, z1 = xi.z + v1z * dt;
    q_out[3 * id] = x1;
    q_out[3 * id + 1] = y1;
    q_out[3 * id + 2] = z1;
    qd_out[3 * id] = v1x;
    qd_out[3 * id + 1] = v1y;
    qd_out[3 * id + 2] = v1z;
    if (f_out) {
        f_out[3 * id] = f.x;
        f_out[3 * id + 1] = f.y;
        f_out[3 * id + 2] = f.z;
    }
    A.pprev[k] = f.w;
    if (p_out) p_out[id] = f.w;
    if (xb_out) {
        const float4 b = D.xbk[k];
        xb_out[3 * id] = b.x;
        xb_out[3 * id + 1] = b.y;
        xb_out[3 * id + 2] = b.z;
    }
    if (!(isfinite(x1) && isfinite(y1) && isfinite(z1) && isfinite(v1x) && isfinite(v1y) &&
          isfinite(v1z)))
        atomicOr(&D.flags[2], 1);
}

// ---- learned warm start -----------------------------------------------------------------
//
// Features of sorted fluid particle s (all from the prologue's buffers):
//   f0 = 100 (rho - rho0) / rho0                  density error, %
//   f1 = 100 dt (sum_j m_j v_ij . gradW_ij + psi (v_i . Gb_i - sum_b v_b . gradW_ib)) / rho0
//        with the velocities at the step's start (k_sph_advect, avisc.w): the density change
//        from the velocity divergence, %
//   f2 = (fluid + boundary neighbours) / 32
//   f3 = p_prev / p_unit                          previous pressure
//   f4 = 100 |v_adv| dt / H                       displacement per step, % of H
//   f5 = distance to the nearest boundary particle / H (1 if none within H)
// p_unit = 0.01 rho0 / (-a_ii) is the pressure that one Jacobi step assigns to a 1 % density
// error.  The network predicts the correction to the IISPH previous-pressure warm start:
// target (p - 0.5 p_prev) / p_unit, warm start p0 = max(0.5 p_prev + y p_unit, 0), so an
// all-zero network reproduces the 0.5 p_prev warm start exactly.
__device__ inline float mlp_features(const SphDev& D, const PhysWorld& P, int s, float dt,
                                     float* f) {
    const float rho = D.rho[s], a = D.aii[s];
    const float2 ax = D.aux[s];
    const float4 v = D.vadv[s];
    const float punit = a < 0.0f ? (P.rho0 * 0.01f) / (-a) : 0.0f;
    f[0] = (rho - P.rho0) / P.rho0 * 100.0f;
    f[1] = (dt * D.avisc[s].w) / P.rho0 * 100.0f;  // dt x velocity divergence term, in % of rho0
    f[2] = ((float)D.ncnt[s] + ax.x) * (1.0f / 32.0f);
    f[3] = punit > 0.0f ? D.pws[s] / punit : 0.0f;
    f[4] = sqrtf(dot3(v.x, v.y, v.z, v.x, v.y, v.z)) * dt * P.invH * 100.0f;
    f[5] = ax.y;
    f[6] = 0.0f;
    f[7] = 0.0f;
    return punit;
}

// Features and the training target ((converged pressure - 0.5 p_prev) / p_unit) of the last
// step, sorted order.
__global__ __launch_bounds__(CHUNK) void k_sph_mlp_features(SphDev D, float dt, float* __restrict__ X,
                                                            float* __restrict__ T) {
    int w, s;
    if (!chunk_slot(D, &w, &s)) return;
    const PhysWorld P = D.phys[w];
    float f[MLP_XPAD];
    const float punit = mlp_features(D, P, s, dt, f);
#pragma unroll
    for (int j = 0; j < MLP_XPAD; ++j) X[(size_t)s * MLP_XPAD + j] = f[j];
    T[s] = punit > 0.0f ? (D.p[s] - 0.5f * D.pws[s]) / punit : 0.0f;
}

// Warm start: p0 = max(0.5 p_prev + clamp(MLP(features), +-MLP_Y_MAX) p_unit, 0), written to p
// and pr before the iterations.  The learned correction applies only to particles with at
// least MLP_GATE_NEIGHBOURS fluid + boundary neighbours: at the free surface and in splashes
// -a_ii is small, p_unit large, and any network output there becomes a pressure spike that a
// mean-error stop test does not see (such particles keep 0.5 p_prev).
#define MLP_GATE_NEIGHBOURS 24.0f
#define MLP_Y_MAX 8.0f
// LDS writes of a wave made visible to its other lanes.
__device__ inline void wave_lds_sync() {
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Warm start of the 64 sorted particles of one wave (lane l holds particle s, ok when it
// exists; every lane of the wave calls it, the MFMA tiles need them all).  xw ([64][MLP_XPAD]
// floats), h1w ([16][MLP_H] bf16) and yw ([64] floats) are the wave's own LDS.
__device__ inline void mlp_warm_wave(const SphDev& D, const MlpRegs& R, const PhysWorld& P, int s, bool ok,
                                     float dt, float* xw, __bf16* h1w, float* yw) {
    const int l = threadIdx.x & 63, g = l >> 4, c = l & 15;
    float f[MLP_XPAD];
    float punit = 0.0f;
    if (ok) {
        punit = mlp_features(D, P, s, dt, f);
    } else {
#pragma unroll
        for (int j = 0; j < MLP_XPAD; ++j) f[j] = 0.0f;
    }
#pragma unroll
    for (int j = 0; j < MLP_XPAD; ++j) xw[l * MLP_XPAD + j] = f[j];
    wave_lds_sync();
    for (int h = 0; h < 4; ++h) {
        float y[4];
        f32x4 c2[4];
        mlp_tile_forward(R, xw + 16 * h * MLP_XPAD, h1w, y, c2);
        if (c < 4) yw[16 * h + 4 * g + c] = c == 0 ? y[0] : (c == 1 ? y[1] : (c == 2 ? y[2] : y[3]));
        __builtin_amdgcn_wave_barrier();
    }
    wave_lds_sync();
    if (ok) {
        const bool bulk = f[2] * 32.0f >= MLP_GATE_NEIGHBOURS;
        const float y = bulk ? fminf(fmaxf(yw[l], -MLP_Y_MAX), MLP_Y_MAX) : 0.0f;
        const float p0 = fmaxf(0.5f * D.pws[s] + y * punit, 0.0f);
        const float rho = D.rho[s];
        D.p0[s] = p0;
        reinterpret_cast<float*>(D.vadv + s)[3] = p0 / (rho * rho);  // pr_j of the source pass
    }
}

__global__ __launch_bounds__(CHUNK) void k_sph_mlp_warm(SphDev D, MlpDev M, float dt) {
    __shared__ __align__(16) float xs[CHUNK * MLP_XPAD];
    __shared__ __align__(16) __bf16 h1s[CHUNK / 64][16 * MLP_H];
    __shared__ float ys[CHUNK];
    int w, s;
    const bool ok = chunk_slot(D, &w, &s);
    const int wave = threadIdx.x >> 6;
    MlpRegs R;
    mlp_load_regs(M, R);
    mlp_warm_wave(D, R, D.phys[w], s, ok, dt, xs + wave * 64 * MLP_XPAD, h1s[wave], ys + wave * 64);
}

// ---- whole-world solve -------------------------------------------------------------------
//
// The pressure solve and forces of whole worlds in one persistent launch: each workgroup takes
// the next world from a queue (worlds in descending fluid count, so the large ones start first
// and the small ones fill in behind them) and runs, for that world alone, the passes the
// per-pass schedule launches over all worlds -- advection, source + first pressure
// acceleration, then Jacobi iterations (pressure acceleration, pressure update with the chunk
// partials, the world's reduction and stop test) until the world has converged or reached
// max_iters, and the forces -- with a workgroup barrier between passes.  START_SOURCE: the
// advection (and the learned warm start between it and the source pass) ran as their own
// launches over all worlds before this one (the network's operands would otherwise hold
// registers through every pass).  START_RESUME: the first iterations ran as per-pass launches
// over all worlds; this launch finishes the worlds still iterating, one at a time, and the
// forces follow as a per-pass launch.  Worlds
// never exchange data, so no workgroup waits for another, and iteration control needs neither
// the host nor a grid-wide barrier.  The passes call the per-pass kernels' particle functions
// in the same order, the chunk partials come from the same CHUNK-thread tree and the world's
// sum from the same one-wave reduction, so every result bit equals the per-pass schedule's.
// NB threads per workgroup (a multiple of CHUNK): a pass covers NB particles of the world at
// a time.  Within a workgroup, global stores are visible to its other waves after a barrier.
enum SolveStart { START_ADVECT = 0, START_SOURCE = 1, START_RESUME = 2 };
template <int NB, int START>
__global__ __launch_bounds__(NB) void k_sph_solve_worlds(SphDev D, float dt, float dt2, float omega, int warm,
                                                         double tol, int min_iters, int max_iters, int W) {
    static_assert(NB % CHUNK == 0, "whole chunks per pass");
    __shared__ double ered[2][NB];
    __shared__ int s_w, s_conv;
    const int t = threadIdx.x;
    for (;;) {
        if (t == 0) s_w = atomicAdd(D.wq, 1);
        __syncthreads();
        // uniform across the workgroup: the world's constants become scalar loads
        const int k = __builtin_amdgcn_readfirstlane(s_w);
        if (k >= W) break;
        const int w = D.w_order[k];
        const int nf = D.w_nf[w];
        // resuming: worlds that the per-pass iterations already stopped are done
        if (nf > 0 && (START != START_RESUME || D.conv[w] == 0)) {
            const PhysWorld P = D.phys[w];
            const int f0 = D.w_fstart[w], c0 = D.w_cfirst[w], nr = D.w_nchunks[w] * CHUNK;
            if constexpr (START == START_ADVECT) {
                for (int r = t; r < nf; r += NB) advect_one(D, P, f0 + r, dt, warm);
                __syncthreads();
            }
            if constexpr (START != START_RESUME) {
                for (int r = t; r < nf; r += NB) source_pacc_one(D, P, f0 + r, dt, warm);
                __syncthreads();
            }
            for (int it = START == START_RESUME ? 1 : 0;; ++it) {
                if (it > 0) {
                    for (int r = t; r < nf; r += NB) pacc_one(D, P, w, f0 + r);
                    __syncthreads();
                }
                // Each thread updates one particle per batch of NB; the chunk tree of each
                // CHUNK-thread run then finishes in the run's first wave: lane l adds the
                // entries l, l + 64, l + 128, l + 192 as the tree's first two levels do (pairs
                // 128 apart, then 64 apart) and the other six levels (32 .. 1 apart) by shifting
                // down across lanes -- the same additions of the same operands as the LDS tree
                // of k_sph_pupdate.  The entries alternate between two LDS buffers, so one
                // barrier per batch suffices: a buffer is written again only after the next
                // batch's barrier, which its readers have passed.
                const int lane = t & 63;
                int nb = 0;
                for (int b = 0; b < nr; b += NB, ++nb) {
                    const int r = b + t;
                    double* red = ered[nb & 1];
                    red[t] = r < nf ? pupdate_one(D, P, w, f0 + r, dt2, omega) : 0.0;
                    __syncthreads();
                    if ((t & (CHUNK - 1)) < 64 && b + (t & ~(CHUNK - 1)) < nr) {
                        const double* e = red + (t & ~(CHUNK - 1)) + lane;
                        double v = (e[0] + e[128]) + (e[64] + e[192]);
#pragma unroll
                        for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
                        if (lane == 0) D.partial[c0 + (b + t) / CHUNK] = v;
                    }
                }
                __syncthreads();
                if (t < 64) {
                    const int c = world_reduce(D, w, t, tol, min_iters, max_iters);
                    if (t == 0) s_conv = c;
                }
                __syncthreads();
                if (__builtin_amdgcn_readfirstlane(s_conv) != 0) break;
            }
            if constexpr (START != START_RESUME)
                for (int r = t; r < nf; r += NB) final_one(D, P, w, f0 + r);
        }
        __syncthreads();  // every thread has read s_w before it is replaced
    }
}

// Non-fluid particles, reference integrator with the caller's force (zero if none)
// (solver.py:24-46).
__device__ inline void passthrough_one(const SphDev& D, const StepEndArgs& A, int k) {
    const float* __restrict__ q = A.q_in;
    const float* __restrict__ qd = A.qd_in;
    const float* __restrict__ inv_mass = A.inv_mass;
    const float* __restrict__ f_in = A.f_in;
    float* __restrict__ q_out = A.q_out;
    float* __restrict__ qd_out = A.qd_out;
    float* __restrict__ f_out = A.f_out;
    float* __restrict__ p_out = A.p_out;
    float* __restrict__ xb_out = A.xb_out;
    const float dt = A.dt, v_max = A.v_max;
    const int id = A.oid[k];
    const float x0 = q[3 * id], y0 = q[3 * id + 1], z0 = q[3 * id + 2];
    if (xb_out) {  // not used by the solver: the particle's position
        xb_out[3 * id] = x0;
        xb_out[3 * id + 1] = y0;
        xb_out[3 * id + 2] = z0;
    }
    if (f_out) {
        f_out[3 * id] = 0.f;
        f_out[3 * id + 1] = 0.f;
        f_out[3 * id + 2] = 0.f;
    }
    if (p_out) p_out[id] = 0.f;
    if ((A.flags[id] & NSPH_PARTICLE_ACTIVE) == 0) {
        q_out[3 * id] = x0;
        q_out[3 * id + 1] = y0;
        q_out[3 * id + 2] = z0;
        return;
    }
    int wd = A.world ? A.world[id] : 0;
    wd = wd < 0 ? 0 : (wd >= A.W ? A.W - 1 : wd);
    const PhysWorld& P = D.phys[wd];
    const float wi = inv_mass[id];
    const float stp = (-wi < 0.0f) ? 1.0f : 0.0f;
    const float fx = f_in ? f_in[3 * id] : 0.0f, fy = f_in ? f_in[3 * id + 1] : 0.0f,
                fz = f_in ? f_in[3 * id + 2] : 0.0f;
    float v1x = qd[3 * id] + (fx * wi + P.g[0] * stp) * dt;
    float v1y = qd[3 * id + 1] + (fy * wi + P.g[1] * stp) * dt;
    float v1z = qd[3 * id + 2] + (fz * wi + P.g[2] * stp) * dt;
    const float mag = sqrtf(dot3(v1x, v1y, v1z, v1x, v1y, v1z));
    if (mag > v_max) {
        const float sc = v_max / mag;
        v1x *= sc;
        v1y *= sc;
        v1z *= sc;
    }
    q_out[3 * id] = x0 + v1x * dt;
    q_out[3 * id + 1] = y0 + v1y * dt;
    q_out[3 * id + 2] = z0 + v1z * dt;
    qd_out[3 * id] = v1x;
    qd_out[3 * id + 1] = v1y;
    qd_out[3 * id + 2] = v1z;
}

// The step's end: threads 0 .. nf - 1 integrate the fluid particles (reference integrator in
// particle order, with the caller's force and the SPH force), threads nf .. nf + no - 1 the
// non-fluid ones (reference integrator with the caller's force), one launch.  guard: nullptr,
// or the still-iterating count that must be 0 for the fluid part to run; other = 0 leaves the
// non-fluid particles out (a step whose end is enqueued a second time).
__global__ void k_sph_step_end(SphDev D, StepEndArgs A, const int* __restrict__ guard, int other) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < A.nf) {
        if (guard && *guard != 0) return;
        integrate_one(D, A, k);
    } else if (other && k < A.nf + A.no) {
        passthrough_one(D, A, k - A.nf);
    }
}

// Boundary build (once): keys, then reorder into bxs and mark btable.
__global__ void k_bnd_keys(const float* __restrict__ q, const int* __restrict__ bid,
                           const int* __restrict__ bworld, const GridWorld* __restrict__ gw, int n,
                           uint32_t* __restrict__ keys, uint32_t* __restrict__ vals) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int id = bid[k];
    const GridWorld g = gw[bworld[k]];
    keys[k] = slab_key(g, q[3 * id], q[3 * id + 1], q[3 * id + 2]);
    vals[k] = (uint32_t)k;
}

__global__ void k_bnd_reorder(const float* __restrict__ q, const int* __restrict__ bid,
                              const uint32_t* __restrict__ perm, const uint32_t* __restrict__ skeys,
                              int n, float4* __restrict__ bxs, int2* __restrict__ btable,
                              int* __restrict__ bcend) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const int id = bid[perm[s]];
    bxs[s] = make_float4(q[3 * id], q[3 * id + 1], q[3 * id + 2], 0.f);
    mark_cell(skeys, s, n, btable, bcend);
}

// breach[c] = 1 when any cell of the 3x3x5 block around c (clamped to the world's box) holds
// a boundary particle.  Every boundary candidate of a fluid particle in cell c lies in that
// block (query ranges span at most cx-1 .. cx+1, cy-1 .. cy+1 and cz-2 .. cz+2: the query
// radius is below the cell width, twice the z cell height), so a zero lets the density pass
// skip the boundary visit.
__global__ void k_bnd_reach(const GridWorld* __restrict__ gw, int W, int n_cells,
                            const int2* __restrict__ btable, unsigned char* __restrict__ breach) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n_cells) return;
    int lo = 0, hi = W - 1;  // last world whose cells start at or before c
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (gw[mid].cell_off <= c) lo = mid;
        else hi = mid - 1;
    }
    const GridWorld g = gw[lo];
    const int local = c - g.cell_off;
    const int cx = local % g.dim[0], cy = (local / g.dim[0]) % g.dim[1], cz = local / (g.dim[0] * g.dim[1]);
    unsigned char any = 0;
    for (int dz = -NSPH_ZROWS / 2; dz <= NSPH_ZROWS / 2; ++dz) {
        const int z = cz + dz;
        if (z < 0 || z >= g.dim[2]) continue;
        for (int dy = -1; dy <= 1; ++dy) {
            const int y = cy + dy;
            if (y < 0 || y >= g.dim[1]) continue;
            for (int dx = -1; dx <= 1; ++dx) {
                const int x = cx + dx;
                if (x < 0 || x >= g.dim[0]) continue;
                if (btable[g.cell_off + (z * g.dim[1] + y) * g.dim[0] + x].y != 0) any = 1;
            }
        }
    }
    breach[c] = any;
}

// Moving boundary (zero-mass particles with a velocity), every step, one thread per moving
// particle k: the cell of the previous step's key of k is cleared (prev_keys, zero-filled before
// the first step), then k's key from its position in state_in.  prev_keys may be the keys array
// itself; thread k reads entry k before writing it.
__global__ void k_kin_keys(int n, const uint32_t* prev_keys, int2* __restrict__ table, const float* __restrict__ q,
                           const int* __restrict__ kid, const int* __restrict__ kworld,
                           const GridWorld* __restrict__ gw, uint32_t* keys, uint32_t* __restrict__ vals) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    table[prev_keys[k] >> NSPH_XSUB_BITS] = make_int2(0, 0);
    const int id = kid[k];
    keys[k] = slab_key(gw[kworld[k]], q[3 * id], q[3 * id + 1], q[3 * id + 2]);
    vals[k] = (uint32_t)k;
}

// Moving boundary in sorted order (positions and velocities of state_in), its cell table, and
// kbreach (zeroed before) over the 3x3x5 block around each occupied cell, marked by the cell's
// first particle.
__global__ void k_kin_reorder(const float* __restrict__ q, const float* __restrict__ qd, const int* __restrict__ kid,
                              const int* __restrict__ kworld, const GridWorld* __restrict__ gw,
                              const uint32_t* __restrict__ perm, const uint32_t* __restrict__ skeys, int n,
                              float4* __restrict__ kxs, float4* __restrict__ kvs, int2* __restrict__ table,
                              int* __restrict__ cend, unsigned char* __restrict__ kbreach) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const int k = (int)perm[s];
    const int id = kid[k];
    kxs[s] = make_float4(q[3 * id], q[3 * id + 1], q[3 * id + 2], 0.f);
    kvs[s] = make_float4(qd[3 * id], qd[3 * id + 1], qd[3 * id + 2], 0.f);
    mark_cell(skeys, s, n, table, cend);
    const int cell = (int)(skeys[s] >> NSPH_XSUB_BITS);
    if (s > 0 && (int)(skeys[s - 1] >> NSPH_XSUB_BITS) == cell) return;
    const GridWorld g = gw[kworld[k]];
    const int local = cell - g.cell_off;
    const int cx = local % g.dim[0], cy = (local / g.dim[0]) % g.dim[1], cz = local / (g.dim[0] * g.dim[1]);
    for (int dz = -NSPH_ZROWS / 2; dz <= NSPH_ZROWS / 2; ++dz) {
        const int z = cz + dz;
        if (z < 0 || z >= g.dim[2]) continue;
        for (int dy = -1; dy <= 1; ++dy) {
            const int y = cy + dy;
            if (y < 0 || y >= g.dim[1]) continue;
            for (int dx = -1; dx <= 1; ++dx) {
                const int x = cx + dx;
                if (x < 0 || x >= g.dim[0]) continue;
                kbreach[g.cell_off + (z * g.dim[1] + y) * g.dim[0] + x] = 1;
            }
        }
    }
}

// Sorted -> particle order for the debugging / dump entry points.
__global__ void k_sph_fields(const uint32_t* __restrict__ perm, const int* __restrict__ fid, int n,
                             const float* __restrict__ rho, const float* __restrict__ p,
                             const int* __restrict__ ncnt, float* __restrict__ o_rho,
                             float* __restrict__ o_p, int* __restrict__ o_cnt) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const int id = fid[perm[s]];
    if (o_rho) o_rho[id] = rho[s];
    if (o_p) o_p[id] = p[s];
    if (o_cnt) o_cnt[id] = ncnt[s];
}

__global__ void k_sph_export_nbrs(const uint32_t* __restrict__ perm, const int* __restrict__ fid,
                                  int Nf, int N, int cap, const int4* __restrict__ nbr4,
                                  const int* __restrict__ ncnt, int* __restrict__ counts,
                                  int* __restrict__ idx) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= Nf) return;
    const int id = fid[perm[s]];
    const int n = ncnt[s];
    counts[id] = n;
    for (int k = 0; k < n && k < cap; ++k) {
        const int4 q = nbr4[(size_t)(k >> 2) * Nf + s];
        const int u = k & 3;
        const int j = u == 0 ? q.x : (u == 1 ? q.y : (u == 2 ? q.z : q.w));
        idx[(size_t)k * N + id] = fid[perm[j]];
    }
}

__global__ void k_f32_zero(float* p, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.f;
}

static PhysWorld make_phys(double H, double Vb, double rho0, double nu, const float* g, double skin) {
    PhysWorld P{};
    const double sigma = 8.0 / (M_PI * H * H * H);
    P.H = (float)H;
    P.H2 = (float)(H * H);
    P.invH = (float)(1.0 / H);
    P.sigma = (float)sigma;
    P.sigma2 = (float)(2.0 * sigma);
    P.gsig2 = (float)(sigma / (H * H));
    P.gsig6 = (float)(6.0 * sigma / H);
    P.W0 = (float)sigma;
    P.psi = (float)(rho0 * Vb);
    P.eps2 = (float)(0.01 * H * H);
    P.visc10 = (float)(10.0 * nu);
    P.rq = (float)H;
    P.g[0] = g[0];
    P.g[1] = g[1];
    P.g[2] = g[2];
    P.rho0 = (float)rho0;
    // Kept lists: a pair within H now was within Rs at the build when each particle moved less
    // than (Rs - H) / 2; the rebuild threshold keeps a 1e-4 relative margin against the FP32
    // rounding of the distances and displacements.
    const double Rs = H * (1.0 + skin);
    const double thr = 0.5 * (Rs - H) * (1.0 - 1.0e-4);
    P.Rs = (float)Rs;
    P.Rs2 = (float)(Rs * Rs);
    P.thr2 = (float)(thr * thr);
    P.skin_on = skin > 0.0 ? 1 : 0;
    return P;
}

}  // namespace nsph

using namespace nsph;

// Kernel classes timed by nsph_solver_set_timing (names in nsph_kernel_name).
enum KernelId {
    KT_BUILD = 0,  // cell clear + keys + radix sort + reorder/mark
    KT_NEIGHBORS,  // neighbour lists fused with the density pass
    KT_ADVECT,
    KT_SOURCE,
    KT_MLP,
    KT_PACC,
    KT_PUPDATE,
    KT_REDUCE,
    KT_FINAL,
    KT_INTEGRATE,
    KT_SOLVE,  // whole-world solve (advection to forces of every world, one launch)
    KT_COUNT
};
static const char* kKernelNames[KT_COUNT] = {"build", "neighbors_density", "advect", "source", "mlp_warm_start",
                                            "pressure_acc", "pressure_update", "reduce", "final",
                                            "integrate", "solve_worlds"};

struct StepGraph {
    nsph_state_in in;
    nsph_state_out out;
    float dt;
    int slots;
    int persist;          // whole-world solve captured (1 + its per-pass iterations; 0: none)
    int warm;             // warm-start mode captured
    const nsph_mlp* mlp;  // MLP whose weights the captured warm start reads
    hipGraphExec_t pro, epi;
};

struct EventPair {
    int kind;
    hipEvent_t a, b;
};

struct nsph_solver {
    nsph_ctx* ctx = nullptr;
    // per-kernel timing (off unless enabled)
    int timing = 0;
    std::vector<EventPair> ev_pool;
    size_t ev_used = 0;
    double t_ms[KT_COUNT] = {0};
    long t_count[KT_COUNT] = {0};
    int N = 0, W = 0, Nf = 0, Nb = 0, Nk = 0, No = 0, C = 0, cap = 0, cap4 = 0, capS = 0;
    double skin = 0.0;
    int force_rebuild = 0;
    int gblock = CHUNK;  // gather-kernel workgroup size (64, 128 or 256)
    int persist = 0;      // 1: whole-world solve (k_sph_solve_worlds) instead of a host-checked one
    int pblock = CHUNK;   // its workgroup size (256, 512 or 1024)
    int ppasses = 0;      // Jacobi iterations run per pass over all worlds before it (0: none)
    int pgrid[3][3] = {};  // its grid per (start, block size): resident workgroups, 0 = not yet queried
    int nbr_variant = 0;  // neighbour-kernel variant (index into kNbrQuads)
    int key_bits = 0;
    size_t total_cells = 0;
    double rho0 = 1000, omega = 0.5, tol = 1e-3;
    int min_iters = 2, max_iters = 100, warm = 0;
    float v_max = 1e5f;
    const float* d_mass = nullptr;  // model arrays (caller-owned, must outlive the solver)
    const float* d_inv_mass = nullptr;
    const int* d_flags = nullptr;
    const int* d_world = nullptr;
    std::vector<int> h_nf, h_iters_last;
    // device tables
    DevBuf gw, phys, w_fstart, w_nf, w_cfirst, w_nchunks, chunk_world, w_order, wq;
    DevBuf fid, fworld, oid;
    DevBuf btable, bxs, ftable, breach, bcend, fcend;
    DevBuf keys, vals;
    uint32_t* skeys = nullptr;
    uint32_t* perm = nullptr;
    RadixSort rs;
    // moving boundary particles: ids, worlds, keys (kskeys/kperm: sorted output), sorted
    // positions and velocities, cell table, reach bytes, per-fluid v_b . gradW sums
    DevBuf kid, kworld, kkeys, kvals, kxs, kvs, ktable, kcend, kbreach, kvb;
    uint32_t* kskeys = nullptr;
    uint32_t* kperm = nullptr;
    RadixSort rsk;
    DevBuf xs, vs, rho, aii, gb, avisc, vadv, src, p, xpr, mixed, fp, inv, a3, pws, p0, pprev, nbr, ncnt, aux;
    DevBuf partial, iters, err, conv, act, flags, wmax, bq_count, bq_id, bq_a, bq_b;
    DevBuf snb, scnt, xbs, xbk, okey, wflag, rb, anyrb, built;
    int* h_pinned = nullptr;  // [0] act value, [16..] RepDev
    RepDev* d_rep_host = nullptr;  // device address of the pinned RepDev
    int rep_seq = 0;               // sequence number of the last step summary
    int stepped = 0;
    int last_iters_launched = 0;
    int mlp_ready = 0;
    nsph_mlp* mlp = nullptr;
    float last_dt = 0.f;
    hipStream_t cap_stream = nullptr;
    std::vector<StepGraph> graphs;
};

static SphDev make_dev(nsph_solver* S) {
    SphDev D{};
    D.Nf = S->Nf;
    D.cap = S->cap;
    D.nsub = CHUNK / S->gblock;
    D.gw = bp<GridWorld>(S->gw);
    D.phys = bp<PhysWorld>(S->phys);
    D.w_fstart = bp<int>(S->w_fstart);
    D.w_nf = bp<int>(S->w_nf);
    D.w_cfirst = bp<int>(S->w_cfirst);
    D.w_nchunks = bp<int>(S->w_nchunks);
    D.chunk_world = bp<int>(S->chunk_world);
    D.btable = bp<int2>(S->btable);
    D.bcend = bp<int>(S->bcend);
    D.fcend = bp<int>(S->fcend);
    D.bxs = bp<float4>(S->bxs);
    D.breach = bp<unsigned char>(S->breach);
    const bool moving = S->Nk > 0;
    D.ktable = moving ? bp<int2>(S->ktable) : nullptr;
    D.kcend = moving ? bp<int>(S->kcend) : nullptr;
    D.kxs = moving ? bp<float4>(S->kxs) : nullptr;
    D.kvs = moving ? bp<float4>(S->kvs) : nullptr;
    D.kbreach = moving ? bp<unsigned char>(S->kbreach) : nullptr;
    D.kvb = moving ? bp<float>(S->kvb) : nullptr;
    D.ftable = bp<int2>(S->ftable);
    D.xs = bp<float4>(S->xs);
    D.vs = bp<float4>(S->vs);
    D.rho = bp<float>(S->rho);
    D.aii = bp<float>(S->aii);
    D.gb = bp<float4>(S->gb);
    D.avisc = bp<float4>(S->avisc);
    D.vadv = bp<float4>(S->vadv);
    D.src = bp<float>(S->src);
    D.p = bp<float>(S->p);
    D.xpr = bp<float4>(S->xpr);
    D.fp = bp<float4>(S->fp);
    D.mixed = bp<int>(S->mixed);
    D.a3 = bp<float>(S->a3);
    D.pws = bp<float>(S->pws);
    D.p0 = bp<float>(S->p0);
    D.nbr4 = bp<int4>(S->nbr);
    D.ncnt = bp<int>(S->ncnt);
    D.aux = bp<float2>(S->aux);
    D.partial = bp<double>(S->partial);
    D.iters = bp<int>(S->iters);
    D.err = bp<double>(S->err);
    D.conv = bp<int>(S->conv);
    D.act = bp<int>(S->act);
    D.flags = bp<int>(S->flags);
    D.wmax = bp<int>(S->wmax);
    D.bq_count = bp<int>(S->bq_count);
    D.bq_id = bp<int2>(S->bq_id);
    D.bq_a = bp<float4>(S->bq_a);
    D.bq_b = bp<float2>(S->bq_b);
    D.capS = S->capS;
    D.snb4 = bp<int4>(S->snb);
    D.scnt = bp<int>(S->scnt);
    D.xbs = bp<float4>(S->xbs);
    D.xbk = bp<float4>(S->xbk);
    D.okey = bp<uint32_t>(S->okey);
    D.wflag = bp<int>(S->wflag);
    D.rb = bp<int>(S->rb);
    D.anyrb = bp<int>(S->anyrb);
    D.built = bp<int>(S->built);
    D.nwave = 4 * S->C;
    D.w_order = bp<int>(S->w_order);
    D.wq = bp<int>(S->wq);
    return D;
}

template <class T>
static int upload(DevBuf& b, const std::vector<T>& v, hipStream_t st) {
    NSPH_TRY(buf_ensure(b, v.size() * sizeof(T) + 16));
    if (!v.empty())
        NSPH_HIP(hipMemcpyAsync(b.ptr, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, st));
    return NSPH_OK;
}

static void solver_free(nsph_solver* S) {
    DevBuf* bufs[] = {&S->gw, &S->phys, &S->w_fstart, &S->w_nf, &S->w_cfirst, &S->w_nchunks,
                      &S->chunk_world, &S->fid, &S->fworld, &S->oid, &S->btable, &S->bxs, &S->breach, &S->bcend, &S->fcend,
                      &S->ftable, &S->keys, &S->vals, &S->xs, &S->vs, &S->rho, &S->aii, &S->gb,
                      &S->avisc, &S->vadv, &S->src, &S->p, &S->xpr, &S->mixed, &S->fp, &S->inv, &S->a3, &S->pws, &S->p0, &S->pprev,
                      &S->nbr, &S->ncnt, &S->aux, &S->partial, &S->iters, &S->err, &S->conv, &S->act,
                      &S->flags, &S->wmax, &S->bq_count, &S->bq_id, &S->bq_a, &S->bq_b,
                      &S->snb, &S->scnt, &S->xbs, &S->xbk, &S->okey, &S->wflag, &S->rb, &S->anyrb,
                      &S->built, &S->kid, &S->kworld, &S->kkeys, &S->kvals, &S->kxs, &S->kvs, &S->ktable,
                      &S->kcend, &S->kbreach, &S->kvb, &S->w_order, &S->wq};
    for (DevBuf* b : bufs) buf_free(*b);
    radix_free(S->rs);
    radix_free(S->rsk);
    if (S->h_pinned) (void)hipHostFree(S->h_pinned);
    S->h_pinned = nullptr;
    for (StepGraph& g : S->graphs) {
        if (g.pro) (void)hipGraphExecDestroy(g.pro);
        if (g.epi) (void)hipGraphExecDestroy(g.epi);
    }
    S->graphs.clear();
    if (S->cap_stream) (void)hipStreamDestroy(S->cap_stream);
    S->cap_stream = nullptr;
    for (EventPair& e : S->ev_pool) {
        (void)hipEventDestroy(e.a);
        (void)hipEventDestroy(e.b);
    }
    S->ev_pool.clear();
}

// Timing brackets: record an event pair around a group of launches (timing mode only).
static int tic(nsph_solver* S, int kind, EventPair** out) {
    *out = nullptr;
    if (!((S->timing >> kind) & 1)) return NSPH_OK;
    if (S->ev_used == S->ev_pool.size()) {
        EventPair e{};
        NSPH_HIP(hipEventCreate(&e.a));
        NSPH_HIP(hipEventCreate(&e.b));
        S->ev_pool.push_back(e);
    }
    EventPair* e = &S->ev_pool[S->ev_used++];
    e->kind = kind;
    NSPH_HIP(hipEventRecord(e->a, S->ctx->stream));
    *out = e;
    return NSPH_OK;
}

static int toc(nsph_solver* S, EventPair* e) {
    if (e) NSPH_HIP(hipEventRecord(e->b, S->ctx->stream));
    return NSPH_OK;
}

static int flush_timing(nsph_solver* S) {
    if (S->ev_used == 0) return NSPH_OK;
    NSPH_HIP(hipStreamSynchronize(S->ctx->stream));
    for (size_t i = 0; i < S->ev_used; ++i) {
        float ms = 0.f;
        NSPH_HIP(hipEventElapsedTime(&ms, S->ev_pool[i].a, S->ev_pool[i].b));
        S->t_ms[S->ev_pool[i].kind] += ms;
        S->t_count[S->ev_pool[i].kind] += 1;
    }
    S->ev_used = 0;
    return NSPH_OK;
}

#define TIMED(kind, ...)                                                                 \
    do {                                                                                 \
        EventPair* ep_;                                                                  \
        NSPH_TRY(tic(S, (kind), &ep_));                                                  \
        __VA_ARGS__;                                                                     \
        NSPH_TRY(toc(S, ep_));                                                           \
    } while (0)

static int solver_init(nsph_solver* S, const nsph_model_desc* m, const nsph_sph_params* prm) {
    hipStream_t st = S->ctx->stream;
    const int N = m->n_particles, W = m->n_worlds;
    S->N = N;
    S->W = W;
    S->rho0 = prm->rest_density;
    S->omega = prm->omega;
    S->tol = prm->tolerance;
    S->min_iters = prm->min_iters;
    S->max_iters = prm->max_iters;
    S->warm = prm->warm_start;
    S->cap = prm->max_neighbors;
    S->cap4 = (prm->max_neighbors + 3) / 4;
    S->skin = prm->skin;
    S->capS = prm->skin > 0.0 ? prm->max_skin_neighbors : 0;
    S->v_max = m->v_max;
    S->d_mass = m->d_mass;
    S->d_inv_mass = m->d_inv_mass;
    S->d_flags = m->d_flags;
    S->d_world = m->d_world;

    // classify on the host (creation time only)
    std::vector<float> q(3 * (size_t)N), qd(3 * (size_t)N), inv_mass(N);
    std::vector<int> flags(N), world(N, 0);
    if (N > 0) {
        NSPH_HIP(hipMemcpyAsync(q.data(), m->d_q, q.size() * 4, hipMemcpyDeviceToHost, st));
        NSPH_HIP(hipMemcpyAsync(qd.data(), m->d_qd, qd.size() * 4, hipMemcpyDeviceToHost, st));
        NSPH_HIP(hipMemcpyAsync(inv_mass.data(), m->d_inv_mass, (size_t)N * 4, hipMemcpyDeviceToHost, st));
        NSPH_HIP(hipMemcpyAsync(flags.data(), m->d_flags, (size_t)N * 4, hipMemcpyDeviceToHost, st));
        if (m->d_world)
            NSPH_HIP(hipMemcpyAsync(world.data(), m->d_world, (size_t)N * 4, hipMemcpyDeviceToHost, st));
        NSPH_HIP(hipStreamSynchronize(st));
    }
    std::vector<int> fid, fworld, bid, bworld, oid, kid, kworld;
    std::vector<int> nf(W, 0), nb(W, 0);
    for (int i = 0; i < N; ++i) {
        int w = world[i] < 0 ? 0 : world[i];
        NSPH_CHECK(w < W, NSPH_ERR_INVALID, "nsph_solver_create: particle %d has world %d >= n_worlds %d",
                   i, world[i], W);
        const bool active = (flags[i] & NSPH_PARTICLE_ACTIVE) != 0;
        if (active && inv_mass[i] > 0.f) {
            fid.push_back(i);
            fworld.push_back(w);
            nf[w]++;
        } else {
            oid.push_back(i);
            if (active) {
                // zero-mass: static boundary at rest, moving boundary with a velocity (the
                // integrator moves it with that velocity, solver.py:24-46)
                const bool rest = qd[3 * (size_t)i] == 0.f && qd[3 * (size_t)i + 1] == 0.f &&
                                  qd[3 * (size_t)i + 2] == 0.f;
                if (rest) {
                    bid.push_back(i);
                    bworld.push_back(w);
                    nb[w]++;
                } else {
                    kid.push_back(i);
                    kworld.push_back(w);
                }
            }
        }
    }
    // fluid of each world must be contiguous in sorted order: keys are world-major, so sorting
    // by key groups worlds.  Fluid ids are ascending; the stable sort keeps id order per cell.
    S->Nf = (int)fid.size();
    S->Nb = (int)bid.size();
    S->Nk = (int)kid.size();
    S->No = (int)oid.size();
    S->h_nf = nf;

    // per-world grid boxes and kernel constants
    std::vector<GridWorld> gws(W);
    std::vector<PhysWorld> phys(W);
    size_t off = 0;
    for (int w = 0; w < W; ++w) {
        const double H = prm->h_support[w];
        NSPH_CHECK(H > 0.0 && std::isfinite(H), NSPH_ERR_INVALID, "nsph_solver_create: bad support radius in world %d", w);
        const double cw = H * (1.0 + prm->skin) * (1.0 + 1.0 / 128.0);  // cell width: the search radius
        GridWorld g{};
        for (int a = 0; a < 3; ++a) {
            const double lo = prm->h_domain_lo[3 * w + a], hi = prm->h_domain_hi[3 * w + a];
            NSPH_CHECK(hi > lo, NSPH_ERR_INVALID, "nsph_solver_create: empty domain in world %d", w);
            const double wa = a < 2 ? cw : 0.5 * cw;  // z cells are half as high
            double d = std::ceil((hi - lo) / wa) + 1.0;
            if (d < 3.0) d = 3.0;
            NSPH_CHECK(d < 1.0e6, NSPH_ERR_INVALID, "nsph_solver_create: domain too large for the cell size in world %d", w);
            g.lo[a] = (float)lo;
            g.dim[a] = (int)d;
        }
        g.inv_w = (float)(1.0 / cw);
        g.inv_t = 2.0f * g.inv_w;  // z
        g.wrap = 0;
        g.cell_off = (int)off;
        off += (size_t)g.dim[0] * g.dim[1] * g.dim[2];
        NSPH_CHECK(off < (size_t)1 << 31, NSPH_ERR_INVALID, "nsph_solver_create: more than 2^31 cells in total");
        gws[w] = g;
        phys[w] = make_phys(H, prm->h_boundary_volume[w], prm->rest_density, prm->viscosity,
                            m->h_gravity + 3 * w, prm->skin);
    }
    S->total_cells = off;
    NSPH_CHECK(off * NSPH_XSUB < (size_t)1 << 32, NSPH_ERR_INVALID,
               "nsph_solver_create: more than 2^30 cells in total");
    S->key_bits = bits_for((uint32_t)(off * NSPH_XSUB - 1));

    // chunk map
    std::vector<int> fstart(W), cfirst(W), nch(W), chunk_world;
    int acc = 0, cacc = 0;
    for (int w = 0; w < W; ++w) {
        fstart[w] = acc;
        cfirst[w] = cacc;
        nch[w] = (nf[w] + CHUNK - 1) / CHUNK;
        for (int c = 0; c < nch[w]; ++c) chunk_world.push_back(w);
        acc += nf[w];
        cacc += nch[w];
    }
    S->C = cacc;

    NSPH_TRY(upload(S->gw, gws, st));
    NSPH_TRY(upload(S->phys, phys, st));
    NSPH_TRY(upload(S->w_fstart, fstart, st));
    NSPH_TRY(upload(S->w_nf, nf, st));
    NSPH_TRY(upload(S->w_cfirst, cfirst, st));
    NSPH_TRY(upload(S->w_nchunks, nch, st));
    NSPH_TRY(upload(S->chunk_world, chunk_world, st));
    // the whole-world solve's queue: descending fluid count, ties by world index (the host
    // vector lives until the stream is synchronised at the end of this function)
    std::vector<int> order(W);
    for (int w = 0; w < W; ++w) order[w] = w;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return nf[a] > nf[b]; });
    NSPH_TRY(upload(S->w_order, order, st));
    NSPH_TRY(buf_ensure(S->wq, 16));
    NSPH_TRY(upload(S->fid, fid, st));
    NSPH_TRY(upload(S->fworld, fworld, st));
    NSPH_TRY(upload(S->oid, oid, st));

    const int Nf = S->Nf, Nb = S->Nb;
    const int nmax = Nf > Nb ? Nf : Nb;
    NSPH_TRY(radix_reserve(S->rs, nmax > 0 ? nmax : 1));
    NSPH_TRY(buf_ensure(S->keys, (size_t)(nmax + 1) * 4));
    NSPH_TRY(buf_ensure(S->vals, (size_t)(nmax + 1) * 4));
    // two entries past the last cell: a row's three table entries are loaded together
    NSPH_TRY(buf_ensure(S->btable, (off + 2) * sizeof(int2)));
    NSPH_TRY(buf_ensure(S->ftable, (off + 2) * sizeof(int2)));
    NSPH_TRY(buf_ensure(S->bcend, (off + 1) * sizeof(int)));
    NSPH_TRY(buf_ensure(S->fcend, (off + 1) * sizeof(int)));
    hipLaunchKernelGGL(k_fill_int2, dim3(2048), dim3(256), 0, st, bp<int2>(S->btable), off + 2, make_int2(0, 0));
    hipLaunchKernelGGL(k_fill_int2, dim3(2048), dim3(256), 0, st, bp<int2>(S->ftable), off + 2, make_int2(0, 0));
    NSPH_TRY(buf_ensure(S->bxs, (size_t)(Nb + 1) * 16));
    NSPH_TRY(buf_ensure(S->flags, 16));

    // static boundary: hash once
    if (Nb > 0) {
        DevBuf dbid, dbworld;
        int rc = upload(dbid, bid, st);
        if (rc == NSPH_OK) rc = upload(dbworld, bworld, st);
        if (rc == NSPH_OK) {
            hipLaunchKernelGGL(k_bnd_keys, dim3((Nb + 255) / 256), dim3(256), 0, st, m->d_q,
                               bp<int>(dbid), bp<int>(dbworld), bp<GridWorld>(S->gw), Nb,
                               bp<uint32_t>(S->keys), bp<uint32_t>(S->vals));
            uint32_t *sk, *sp;
            rc = radix_sort(S->rs, bp<uint32_t>(S->keys), bp<uint32_t>(S->vals), Nb, S->key_bits, st, &sk, &sp);
            if (rc == NSPH_OK) {
                hipLaunchKernelGGL(k_bnd_reorder, dim3((Nb + 255) / 256), dim3(256), 0, st, m->d_q,
                                   bp<int>(dbid), sp, sk, Nb, bp<float4>(S->bxs), bp<int2>(S->btable),
                                   bp<int>(S->bcend));
                if (hipStreamSynchronize(st) != hipSuccess) {
                    rc = NSPH_ERR_HIP;
                    set_error("nsph_solver_create: boundary build failed");
                }
            }
        }
        buf_free(dbid);
        buf_free(dbworld);
        NSPH_TRY(rc);
        NSPH_HIP(hipGetLastError());
    }

    // moving boundary: tables zeroed, keys zero-filled for the first step's clearing
    if (S->Nk > 0) {
        const int Nk = S->Nk;
        NSPH_TRY(upload(S->kid, kid, st));
        NSPH_TRY(upload(S->kworld, kworld, st));
        NSPH_TRY(radix_reserve(S->rsk, Nk));
        NSPH_TRY(buf_ensure(S->kkeys, (size_t)(Nk + 1) * 4));
        NSPH_TRY(buf_ensure(S->kvals, (size_t)(Nk + 1) * 4));
        NSPH_TRY(buf_ensure(S->kxs, (size_t)(Nk + 1) * 16));
        NSPH_TRY(buf_ensure(S->kvs, (size_t)(Nk + 1) * 16));
        NSPH_TRY(buf_ensure(S->ktable, (off + 2) * sizeof(int2)));
        NSPH_TRY(buf_ensure(S->kcend, (off + 1) * sizeof(int)));
        NSPH_TRY(buf_ensure(S->kbreach, (off + 4) & ~(size_t)3));
        NSPH_TRY(buf_ensure(S->kvb, ((size_t)Nf + 1) * 4));
        hipLaunchKernelGGL(k_fill_int2, dim3(2048), dim3(256), 0, st, bp<int2>(S->ktable), off + 2, make_int2(0, 0));
        hipLaunchKernelGGL(k_fill_i32, dim3(256), dim3(256), 0, st, bp<int>(S->kkeys), (size_t)(Nk + 1), 0);
        S->kskeys = bp<uint32_t>(S->kkeys);
        uint32_t* unused = nullptr;
        radix_result(S->rsk, bp<uint32_t>(S->kkeys), bp<uint32_t>(S->kvals), Nk, S->key_bits, &unused, &S->kperm);
        NSPH_HIP(hipGetLastError());
    }

    NSPH_TRY(buf_ensure(S->breach, off + 1));
    hipLaunchKernelGGL(k_bnd_reach, dim3((unsigned)((off + 255) / 256)), dim3(256), 0, st, bp<GridWorld>(S->gw), W,
                       (int)off, bp<int2>(S->btable), bp<unsigned char>(S->breach));
    NSPH_HIP(hipGetLastError());

    // the first step clears the cells named by S->skeys: point it at zero-filled keys (cell 0,
    // already empty), after the boundary build has finished with the key buffer
    hipLaunchKernelGGL(k_fill_i32, dim3(256), dim3(256), 0, st, bp<int>(S->keys), (size_t)(nmax + 1), 0);
    S->skeys = bp<uint32_t>(S->keys);
    // The permutation is read before the sort by the kept-order copy (guarded on the device
    // until a first sort exists), and a step captured before any eager step holds this pointer:
    // point it at the buffer every sort of Nf keys leaves its result in.
    {
        uint32_t* unused = nullptr;
        radix_result(S->rs, bp<uint32_t>(S->keys), bp<uint32_t>(S->vals), Nf, S->key_bits, &unused, &S->perm);
    }
    const size_t nf1 = (size_t)Nf + 1;
    NSPH_TRY(buf_ensure(S->xs, (nf1 + XS_PAD) * 16));  // windows read past a row's end
    NSPH_TRY(buf_ensure(S->vs, nf1 * 16));
    NSPH_TRY(buf_ensure(S->rho, nf1 * 4));
    NSPH_TRY(buf_ensure(S->aii, nf1 * 4));
    NSPH_TRY(buf_ensure(S->gb, nf1 * 16));
    NSPH_TRY(buf_ensure(S->avisc, nf1 * 16));
    NSPH_TRY(buf_ensure(S->vadv, nf1 * 16));
    NSPH_TRY(buf_ensure(S->src, nf1 * 4));
    NSPH_TRY(buf_ensure(S->p, nf1 * 4));
    NSPH_TRY(buf_ensure(S->xpr, nf1 * 16));
    NSPH_TRY(buf_ensure(S->fp, nf1 * 16));
    NSPH_TRY(buf_ensure(S->inv, nf1 * 4));
    NSPH_TRY(buf_ensure(S->mixed, ((size_t)W + 1) * 4));
    NSPH_TRY(buf_ensure(S->a3, nf1 * 12));
    NSPH_TRY(buf_ensure(S->pws, nf1 * 4));
    NSPH_TRY(buf_ensure(S->p0, nf1 * 4));
    NSPH_TRY(buf_ensure(S->pprev, nf1 * 4));
    NSPH_TRY(buf_ensure(S->nbr, nf1 * (size_t)S->cap4 * 16));
    NSPH_TRY(buf_ensure(S->ncnt, nf1 * 4));
    NSPH_TRY(buf_ensure(S->aux, nf1 * 8));
    NSPH_TRY(buf_ensure(S->partial, ((size_t)S->C + 1) * 8));
    NSPH_TRY(buf_ensure(S->wmax, ((size_t)S->C + 1) * 16));
    NSPH_TRY(buf_ensure(S->bq_count, 16));
    NSPH_TRY(buf_ensure(S->bq_id, nf1 * 8));
    NSPH_TRY(buf_ensure(S->bq_a, nf1 * 16));
    NSPH_TRY(buf_ensure(S->bq_b, nf1 * 8));
    NSPH_TRY(buf_ensure(S->snb, nf1 * (size_t)((S->capS + 3) / 4) * 16 + 16));
    NSPH_TRY(buf_ensure(S->scnt, nf1 * 4));
    NSPH_TRY(buf_ensure(S->xbs, (nf1 + XS_PAD) * 16));  // searched like xs: windows past a row's end
    NSPH_TRY(buf_ensure(S->xbk, nf1 * 16));
    NSPH_TRY(buf_ensure(S->okey, nf1 * 4));
    NSPH_TRY(buf_ensure(S->wflag, ((size_t)W + 1) * 4));
    NSPH_TRY(buf_ensure(S->rb, ((size_t)W + 1) * 4));
    NSPH_TRY(buf_ensure(S->anyrb, 16));
    NSPH_TRY(buf_ensure(S->built, 16));
    hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, st, bp<int>(S->built), (size_t)4, 0);
    NSPH_TRY(buf_ensure(S->iters, ((size_t)W + 1) * 4));
    NSPH_TRY(buf_ensure(S->err, ((size_t)W + 1) * 8));
    NSPH_TRY(buf_ensure(S->conv, ((size_t)W + 1) * 4));
    NSPH_TRY(buf_ensure(S->act, 16));
    NSPH_TRY(buf_ensure(S->flags, 16));
    hipLaunchKernelGGL(k_f32_zero, dim3((unsigned)((nf1 + 255) / 256)), dim3(256), 0, st, bp<float>(S->pprev), nf1);
    hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, st, bp<int>(S->flags), (size_t)4, 0);
    NSPH_HIP(hipGetLastError());
    // [0] act value, [16..] RepDev.  Coherent (fine-grained) so that the report kernel's
    // system-scope stores reach the host while the stream is still running.
    NSPH_HIP(hipHostMalloc((void**)&S->h_pinned, 128, hipHostMallocMapped | hipHostMallocCoherent));
    // The block may be one a destroyed solver used: clear its last sequence number, or the
    // first summary wait could match it and return that solver's report.
    std::memset(S->h_pinned, 0, 128);
    NSPH_HIP(hipHostGetDevicePointer((void**)&S->d_rep_host, S->h_pinned + 16, 0));
    NSPH_HIP(hipStreamSynchronize(st));
    return NSPH_OK;
}

// Enqueue the step summary (k_sph_report writes it straight into pinned host memory, so no
// copy follows), wait for the stream once and read it.
static int summarise_step(nsph_solver* S, hipStream_t st, int act_slot, RepDev* out) {
    const int want = S->rep_seq = (S->rep_seq % 1000000000) + 1;
    hipLaunchKernelGGL(k_sph_report, dim3(1), dim3(1024), 0, st, make_dev(S), S->W, act_slot, want,
                       S->d_rep_host);
    NSPH_HIP(hipGetLastError());
    // Wait for the summary's sequence number without letting the thread sleep (a blocking
    // stream wait can add an interrupt wake-up to every step).  The stream is asked for
    // errors every 1024 polls; if it went idle without the number, synchronise to report why.
    volatile RepDev* h = reinterpret_cast<volatile RepDev*>(S->h_pinned + 16);
    for (long k = 1; h->seq != want; ++k) {
        if ((k & 1023) == 0) {
            const hipError_t e = hipStreamQuery(st);
            if (e == hipErrorNotReady) continue;
            NSPH_HIP(e);
            if (h->seq == want) break;
            NSPH_HIP(hipStreamSynchronize(st));
            NSPH_CHECK(h->seq == want, NSPH_ERR_HIP, "step summary: the stream finished without it");
        }
    }
    std::atomic_thread_fence(std::memory_order_acquire);
    out->act = h->act;
    out->max_it = h->max_it;
    out->sum_it = h->sum_it;
    out->unconv = h->unconv;
    out->overflow = h->overflow;
    out->maxnbr = h->maxnbr;
    out->nonfinite = h->nonfinite;
    out->seq = h->seq;
    out->searched = h->searched;
    out->max_err = h->max_err;
    return NSPH_OK;
}

static void report_from(const RepDev& d, nsph_step_report* rep) {
    nsph_step_report r{};
    r.max_iterations = d.max_it;
    r.sum_iterations = d.sum_it;
    r.unconverged_worlds = d.unconv;
    r.max_density_error = d.max_err;
    r.neighbor_overflow = d.overflow;
    r.max_neighbor_count = d.maxnbr;
    r.searched_worlds = d.searched;
    r.status = NSPH_OK;
    if (r.unconverged_worlds) r.status = NSPH_ERR_NOT_CONVERGED;
    if (d.overflow) r.status = NSPH_ERR_CAPACITY;
    if (d.nonfinite) r.status = NSPH_ERR_NONFINITE;
    *rep = r;
}

static int collect_report(nsph_solver* S, nsph_step_report* rep) {
    RepDev d;
    NSPH_TRY(summarise_step(S, S->ctx->stream, -1, &d));
    report_from(d, rep);
    return NSPH_OK;
}

extern "C" {

int nsph_solver_create(nsph_ctx* ctx, const nsph_model_desc* m, const nsph_sph_params* prm,
                       nsph_solver** out) {
    NSPH_CHECK(ctx && m && prm && out, NSPH_ERR_INVALID, "nsph_solver_create: NULL argument");
    NSPH_CHECK(m->n_particles >= 0 && m->n_worlds >= 1, NSPH_ERR_INVALID,
               "nsph_solver_create: need n_particles >= 0 and n_worlds >= 1");
    NSPH_CHECK(m->n_particles == 0 || (m->d_q && m->d_qd && m->d_mass && m->d_inv_mass && m->d_flags),
               NSPH_ERR_INVALID, "nsph_solver_create: missing model array");
    NSPH_CHECK(m->h_gravity && prm->h_support && prm->h_boundary_volume && prm->h_domain_lo &&
                   prm->h_domain_hi, NSPH_ERR_INVALID, "nsph_solver_create: missing per-world parameter");
    NSPH_CHECK(prm->max_neighbors >= 1 && prm->max_neighbors <= 4096, NSPH_ERR_INVALID,
               "nsph_solver_create: max_neighbors must be in [1, 4096]");
    NSPH_CHECK(prm->max_iters >= 1 && prm->min_iters >= 0 && prm->min_iters <= prm->max_iters,
               NSPH_ERR_INVALID, "nsph_solver_create: need 0 <= min_iters <= max_iters, max_iters >= 1");
    NSPH_CHECK(prm->rest_density > 0 && prm->omega > 0 && prm->tolerance >= 0, NSPH_ERR_INVALID,
               "nsph_solver_create: rest_density, omega must be > 0 and tolerance >= 0");
    NSPH_CHECK(prm->warm_start >= NSPH_WARM_ZERO && prm->warm_start <= NSPH_WARM_MLP, NSPH_ERR_INVALID,
               "nsph_solver_create: unknown warm_start mode %d", prm->warm_start);
    NSPH_CHECK(prm->skin == 0.0 || (prm->skin >= 0.01 && prm->skin <= 1.0), NSPH_ERR_INVALID,
               "nsph_solver_create: skin must be 0 or in [0.01, 1] (got %g)", prm->skin);
    NSPH_CHECK(prm->skin == 0.0 || (prm->max_skin_neighbors >= prm->max_neighbors &&
                                    prm->max_skin_neighbors <= 8192),
               NSPH_ERR_INVALID, "nsph_solver_create: max_skin_neighbors must be in [max_neighbors, 8192]");
    NSPH_HIP(hipSetDevice(ctx->device));
    nsph_solver* S = new nsph_solver();
    S->ctx = ctx;
    int rc = solver_init(S, m, prm);
    if (rc != NSPH_OK) {
        solver_free(S);
        delete S;
        return rc;
    }
    *out = S;
    return NSPH_OK;
}

}  // extern "C"

// ---- step pieces: prologue (build, lists, density, advection, source), Jacobi iterations,
// epilogue (final pressure force, integration, non-fluid particles).  The eager step runs
// them back to back; a captured step holds prologue + a fixed number of iteration slots in
// one graph and the epilogue in another.

// Neighbour-kernel variants: LDS list quads per particle (nsph_solver_set_search).
static const int kNbrQuads[] = {6, 4, 8, 2};
#define NBR_VARIANTS ((int)(sizeof(kNbrQuads) / sizeof(kNbrQuads[0])))

template <bool SKIN>
static void launch_nbr_variant(int v, dim3 gg, dim3 gb, hipStream_t st, const SphDev& D, float dt2) {
    const size_t lds = (size_t)kNbrQuads[v] * gb.x * 16;
    switch (v) {
        case 1: hipLaunchKernelGGL((k_sph_nbr_density<SKIN, 4>), gg, gb, lds, st, D, dt2); break;
        case 2: hipLaunchKernelGGL((k_sph_nbr_density<SKIN, 8>), gg, gb, lds, st, D, dt2); break;
        case 3: hipLaunchKernelGGL((k_sph_nbr_density<SKIN, 2>), gg, gb, lds, st, D, dt2); break;
        default: hipLaunchKernelGGL((k_sph_nbr_density<SKIN, 6>), gg, gb, lds, st, D, dt2); break;
    }
}

static void launch_nbr_density(nsph_solver* S, bool keep, dim3 gg, dim3 gb, hipStream_t st, const SphDev& D,
                               float dt2) {
    if (keep) launch_nbr_variant<true>(S->nbr_variant, gg, gb, st, D, dt2);
    else launch_nbr_variant<false>(S->nbr_variant, gg, gb, st, D, dt2);
}

// Whole-world solve: as many workgroups as the GPU holds at once (at most one per world); a
// workgroup that finds the queue empty exits at once, so the grid never waits on itself.
template <int NB, int START>
static int launch_solve_t(nsph_solver* S, hipStream_t st, const SphDev& D, float dt) {
    int& grid = S->pgrid[START][NB == 256 ? 0 : (NB == 512 ? 1 : 2)];
    if (grid == 0) {
        int per_cu = 0, ncu = 0;
        NSPH_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_sph_solve_worlds<NB, START>, NB, 0));
        NSPH_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, S->ctx->device));
        NSPH_CHECK(per_cu > 0 && ncu > 0, NSPH_ERR_HIP, "whole-world solve: no resident workgroup of %d threads", NB);
        grid = per_cu * ncu;
    }
    const int g = std::max(1, std::min(grid, S->W));
    const float dt2 = (float)((double)dt * (double)dt);
    hipLaunchKernelGGL((k_sph_solve_worlds<NB, START>), dim3(g), dim3(NB), 0, st, D, dt, dt2, (float)S->omega,
                       S->warm, S->tol, S->min_iters, S->max_iters, S->W);
    NSPH_HIP(hipGetLastError());
    return NSPH_OK;
}

template <int NB>
static int launch_solve_b(nsph_solver* S, hipStream_t st, const SphDev& D, float dt, int start) {
    switch (start) {
        case START_SOURCE: return launch_solve_t<NB, START_SOURCE>(S, st, D, dt);
        case START_RESUME: return launch_solve_t<NB, START_RESUME>(S, st, D, dt);
        default: return launch_solve_t<NB, START_ADVECT>(S, st, D, dt);
    }
}

static int launch_solve(nsph_solver* S, hipStream_t st, const SphDev& D, float dt, int start) {
    switch (S->pblock) {
        case 512: return launch_solve_b<512>(S, st, D, dt, start);
        case 1024: return launch_solve_b<1024>(S, st, D, dt, start);
        default: return launch_solve_b<256>(S, st, D, dt, start);
    }
}

static int enqueue_prologue(nsph_solver* S, hipStream_t st, const nsph_state_in& in, float dt) {
    const float* q_in = in.d_q;
    const float* qd_in = in.d_qd;
    const int Nf = S->Nf, W = S->W, C = S->C;
    const float dt2 = (float)((double)dt * (double)dt);
    SphDev D = make_dev(S);
    EventPair* ep_build = nullptr;
    if (Nf > 0) NSPH_TRY(tic(S, KT_BUILD, &ep_build));
    // With a skin, the particles are first put in the sorted order of the last search and each
    // world's search positions are checked; the sort and reorder below then run (guarded on the
    // device) only when some world searches anew, and leave the other worlds' order unchanged.
    // Without a skin every world searches every step, and the step starts in one launch.
    const bool keep = S->skin > 0.0;
    if (Nf > 0 && !keep) {
        const int nt = std::max(std::max(W, D.nwave), Nf);
        hipLaunchKernelGGL(k_sph_begin_keys, dim3((nt + 255) / 256), dim3(256), 0, st, D, W, Nf, S->skeys,
                           bp<int2>(S->ftable), q_in, bp<int>(S->fid), bp<int>(S->fworld), bp<GridWorld>(S->gw),
                           bp<uint32_t>(S->keys), bp<uint32_t>(S->vals));
        S->force_rebuild = 0;
    } else {
        hipLaunchKernelGGL(k_sph_begin, dim3((std::max(W, D.nwave) + 255) / 256), dim3(256), 0, st, D, W);
        if (Nf == 0) return NSPH_OK;
        hipLaunchKernelGGL(k_sph_reorder_kept, dim3((Nf + 255) / 256), dim3(256), 0, st, D, q_in, qd_in,
                           S->d_mass, bp<int>(S->fid), S->perm, bp<float>(S->pprev), in.d_p, Nf);
        hipLaunchKernelGGL(k_sph_displace, dim3((Nf + 255) / 256), dim3(256), 0, st, D, Nf, q_in, bp<int>(S->fid),
                           bp<int>(S->fworld), in.d_xb);
        hipLaunchKernelGGL(k_sph_decide, dim3((W + 255) / 256), dim3(256), 0, st, D, W, S->force_rebuild);
        S->force_rebuild = 0;
        hipLaunchKernelGGL(k_sph_keys, dim3((Nf + 255) / 256), dim3(256), 0, st, D, Nf, S->skeys,
                           bp<int2>(S->ftable), q_in, in.d_xb, bp<int>(S->fid), bp<int>(S->fworld),
                           bp<GridWorld>(S->gw), bp<uint32_t>(S->keys), bp<uint32_t>(S->vals));
    }
    NSPH_TRY(radix_sort(S->rs, bp<uint32_t>(S->keys), bp<uint32_t>(S->vals), Nf, S->key_bits,
                        st, &S->skeys, &S->perm, D.anyrb));
    hipLaunchKernelGGL(k_sph_reorder, dim3((Nf + 255) / 256), dim3(256), 0, st, D, q_in, qd_in,
                       S->d_mass, bp<int>(S->fid), S->perm, bp<float>(S->pprev), in.d_p,
                       S->skeys, Nf, D.xs, D.vs, D.pws, D.ftable, D.fcend, bp<int>(S->inv));
    if (S->Nk > 0) {
        // moving boundary: hashed from state_in every step
        const int Nk = S->Nk;
        const size_t words = (S->total_cells + 4) / 4;
        hipLaunchKernelGGL(k_fill_i32, dim3((unsigned)std::min<size_t>((words + 255) / 256, 4096)), dim3(256), 0, st,
                           bp<int>(S->kbreach), words, 0);
        hipLaunchKernelGGL(k_kin_keys, dim3((Nk + 255) / 256), dim3(256), 0, st, Nk, S->kskeys, bp<int2>(S->ktable),
                           q_in, bp<int>(S->kid), bp<int>(S->kworld), bp<GridWorld>(S->gw), bp<uint32_t>(S->kkeys),
                           bp<uint32_t>(S->kvals));
        NSPH_TRY(radix_sort(S->rsk, bp<uint32_t>(S->kkeys), bp<uint32_t>(S->kvals), Nk, S->key_bits, st, &S->kskeys,
                            &S->kperm));
        hipLaunchKernelGGL(k_kin_reorder, dim3((Nk + 255) / 256), dim3(256), 0, st, q_in, qd_in, bp<int>(S->kid),
                           bp<int>(S->kworld), bp<GridWorld>(S->gw), S->kperm, S->kskeys, Nk, bp<float4>(S->kxs),
                           bp<float4>(S->kvs), bp<int2>(S->ktable), bp<int>(S->kcend), bp<unsigned char>(S->kbreach));
    }
    NSPH_TRY(toc(S, ep_build));
    const dim3 gg(C * D.nsub), gb(S->gblock);
    TIMED(KT_NEIGHBORS, {
        launch_nbr_density(S, keep, gg, gb, st, D, dt2);
        // particles near static or moving boundary particles were queued for this pass
        if (S->Nb > 0 || S->Nk > 0) {
            const dim3 bg((unsigned)std::min(S->C, 2048));
            if (S->Nk > 0) hipLaunchKernelGGL(k_sph_bnd_density<true>, bg, dim3(256), 0, st, D, dt2);
            else hipLaunchKernelGGL(k_sph_bnd_density<false>, bg, dim3(256), 0, st, D, dt2);
        }
    });
    if (S->persist && S->ppasses == 0 && S->warm != NSPH_WARM_MLP) {
        // advection to forces of every world in one launch
        TIMED(KT_SOLVE, NSPH_TRY(launch_solve(S, st, D, dt, START_ADVECT)));
        S->last_dt = dt;
        return NSPH_OK;
    }
    TIMED(KT_ADVECT, hipLaunchKernelGGL(k_sph_advect, gg, gb, 0, st, D, dt, S->warm));
    if (S->warm == NSPH_WARM_MLP)
        TIMED(KT_MLP, hipLaunchKernelGGL(k_sph_mlp_warm, dim3(C), dim3(CHUNK), 0, st, D,
                                         nsph_mlp_device_view(S->mlp), dt));
    if (S->persist && S->ppasses == 0) {
        // source to forces of every world in one launch
        TIMED(KT_SOLVE, NSPH_TRY(launch_solve(S, st, D, dt, START_SOURCE)));
        S->last_dt = dt;
        return NSPH_OK;
    }
    TIMED(KT_SOURCE, hipLaunchKernelGGL(k_sph_source_pacc, gg, gb, 0, st, D, dt, S->warm));
    NSPH_HIP(hipGetLastError());
    S->last_dt = dt;
    return NSPH_OK;
}

// Iterations it0 .. it1-1; with check_every > 0 the host reads the number of still-iterating
// worlds every check_every iterations (a stream sync) and stops launching when it is zero.
static int enqueue_iterations(nsph_solver* S, hipStream_t st, float dt, int it0, int it1,
                              int check_every, int* launched_to) {
    const int W = S->W, C = S->C;
    const float dt2 = (float)((double)dt * (double)dt);
    SphDev D = make_dev(S);
    const int red_blocks = (W * 64 + 255) / 256;
    int launched = it0;
    if (S->Nf > 0) {
        for (int it = it0; it < it1; ++it) {
            // iteration 0's pressure acceleration comes from the fused source pass
            if (it > 0)
                TIMED(KT_PACC, hipLaunchKernelGGL(k_sph_pacc, dim3(C * D.nsub), dim3(S->gblock), 0, st, D));
            TIMED(KT_PUPDATE, hipLaunchKernelGGL(k_sph_pupdate, dim3(C), dim3(CHUNK), 0, st, D, dt2,
                                                 (float)S->omega));
            TIMED(KT_REDUCE, hipLaunchKernelGGL(k_sph_reduce, dim3(red_blocks), dim3(256), 0, st, D,
                                                W, it, S->tol, S->min_iters, S->max_iters));
            launched = it + 1;
            if (check_every > 0 && launched >= S->min_iters && (launched % check_every) == 0 &&
                launched < S->max_iters) {
                NSPH_HIP(hipMemcpyAsync(S->h_pinned, bp<int>(S->act) + (it & 1), 4,
                                        hipMemcpyDeviceToHost, st));
                NSPH_HIP(hipStreamSynchronize(st));
                if (S->h_pinned[0] == 0) break;
            }
        }
    }
    NSPH_HIP(hipGetLastError());
    *launched_to = launched;
    return NSPH_OK;
}

// guard: nullptr, or the still-iterating count that must be 0 for the fluid part to run.
// boundary: also integrate the non-fluid particles (unguarded: they do not depend on the
// pressure solve, so a step that enqueues its end again skips them the second time).
static int enqueue_epilogue(nsph_solver* S, hipStream_t st, const nsph_state_in& in, const nsph_state_out& out,
                            float dt, const int* guard = nullptr, bool boundary = true) {
    const int W = S->W, C = S->C;
    SphDev D = make_dev(S);
    if (S->Nf > 0 && !(S->persist && S->ppasses == 0))  // a whole-world solve from the start ends with the forces
        TIMED(KT_FINAL, hipLaunchKernelGGL(k_sph_final, dim3(C * D.nsub), dim3(S->gblock), 0, st, D, guard));
    const int other = S->No > 0 && boundary ? 1 : 0;
    const int n = S->Nf + (other ? S->No : 0);
    if (n > 0) {
        StepEndArgs A{};
        A.inv = bp<int>(S->inv);
        A.fid = bp<int>(S->fid);
        A.fworld = bp<int>(S->fworld);
        A.oid = bp<int>(S->oid);
        A.inv_mass = S->d_inv_mass;
        A.flags = S->d_flags;
        A.world = S->d_world;
        A.W = W;
        A.nf = S->Nf;
        A.no = S->No;
        A.dt = dt;
        A.v_max = S->v_max;
        A.q_in = in.d_q;
        A.qd_in = in.d_qd;
        A.f_in = in.d_f;
        A.q_out = out.d_q;
        A.qd_out = out.d_qd;
        A.f_out = out.d_f;
        A.p_out = out.d_p;
        A.xb_out = S->capS > 0 ? out.d_xb : nullptr;
        A.pprev = bp<float>(S->pprev);
        TIMED(KT_INTEGRATE, hipLaunchKernelGGL(k_sph_step_end, dim3((n + 255) / 256), dim3(256), 0, st, D, A, guard,
                                               other));
    }
    NSPH_HIP(hipGetLastError());
    return NSPH_OK;
}

// After the prologue of a whole-world solve that starts with per-pass iterations: those
// iterations, then the launch that finishes the worlds still iterating (nothing to do when the
// prologue's persistent launch solved every world).
static int enqueue_world_tail(nsph_solver* S, hipStream_t st, float dt) {
    if (!S->persist || S->ppasses == 0 || S->Nf == 0) return NSPH_OK;
    int launched = 0;
    NSPH_TRY(enqueue_iterations(S, st, dt, 0, std::min(S->ppasses, S->max_iters), 0, &launched));
    const SphDev D = make_dev(S);
    TIMED(KT_SOLVE, NSPH_TRY(launch_solve(S, st, D, dt, START_RESUME)));
    return NSPH_OK;
}

static int check_step_args(nsph_solver* S, const nsph_state_in* in, const nsph_state_out* out, float dt) {
    NSPH_CHECK(S != nullptr && in != nullptr && out != nullptr, NSPH_ERR_INVALID, "nsph_solver_step: NULL argument");
    NSPH_CHECK(S->N == 0 || (in->d_q && in->d_qd && out->d_q && out->d_qd), NSPH_ERR_INVALID,
               "nsph_solver_step: NULL state array");
    NSPH_CHECK(dt > 0.f && std::isfinite(dt), NSPH_ERR_INVALID, "nsph_solver_step: dt must be positive");
    NSPH_CHECK(S->N == 0 || (in->d_q != out->d_q && in->d_qd != out->d_qd &&
                             (in->d_p == nullptr || in->d_p != out->d_p) &&
                             (in->d_xb == nullptr || in->d_xb != out->d_xb)),
               NSPH_ERR_INVALID, "nsph_solver_step: state_in and state_out must not alias");
    NSPH_CHECK(S->warm != NSPH_WARM_MLP || S->mlp_ready, NSPH_ERR_INVALID,
               "nsph_solver_step: the MLP warm start needs weights (nsph_solver_set_mlp)");
    return NSPH_OK;
}

extern "C" {

int nsph_solver_step(nsph_solver* S, const nsph_state_in* in_, const nsph_state_out* out_, float dt,
                     int check_every, nsph_step_report* report) {
    NSPH_TRY(check_step_args(S, in_, out_, dt));
    const nsph_state_in in = *in_;
    const nsph_state_out out = *out_;
    NSPH_CHECK(check_every >= 0, NSPH_ERR_INVALID, "nsph_solver_step: check_every must be >= 0");
    hipStream_t st = S->ctx->stream;
    if (S->timing && S->ev_used > 4096) NSPH_TRY(flush_timing(S));
    NSPH_TRY(enqueue_prologue(S, st, in, dt));
    int launched = 0;
    if (S->persist) {
        // iteration control is on the device: one summary read, when a report is asked for
        NSPH_TRY(enqueue_world_tail(S, st, dt));
        NSPH_TRY(enqueue_epilogue(S, st, in, out, dt));
        S->last_iters_launched = -1;
        S->stepped = 1;
        if (report) NSPH_TRY(collect_report(S, report));
        return NSPH_OK;
    }
    if (check_every == 0 || S->Nf == 0) {
        // no host checks: every iteration slot is enqueued (converged worlds skip them)
        NSPH_TRY(enqueue_iterations(S, st, dt, 0, S->max_iters, 0, &launched));
        NSPH_TRY(enqueue_epilogue(S, st, in, out, dt));
        S->last_iters_launched = launched;
        S->stepped = 1;
        if (report) NSPH_TRY(collect_report(S, report));
        return NSPH_OK;
    }
    // min_iters iterations, then the step's end guarded by "every world converged", so the
    // common step needs one synchronisation (the summary).  Otherwise iterate on, checking
    // every check_every iterations, and enqueue the end again.  Iteration counts are decided
    // on the device per world, so the results do not depend on this schedule.
    const int first = std::min(std::max(S->min_iters, 1), S->max_iters);
    NSPH_TRY(enqueue_iterations(S, st, dt, 0, first, 0, &launched));
    for (bool boundary = true;; boundary = false) {
        const bool last = launched >= S->max_iters;
        const int slot = (launched - 1) & 1;
        NSPH_TRY(enqueue_epilogue(S, st, in, out, dt, last ? nullptr : bp<int>(S->act) + slot, boundary));
        RepDev d;
        NSPH_TRY(summarise_step(S, st, last ? -1 : slot, &d));
        if (last || d.act == 0) {
            if (report) report_from(d, report);
            break;
        }
        NSPH_TRY(enqueue_iterations(S, st, dt, launched, S->max_iters, check_every, &launched));
    }
    S->last_iters_launched = launched;
    S->stepped = 1;
    return NSPH_OK;
}

int nsph_solver_graph_capture(nsph_solver* S, const nsph_state_in* in, const nsph_state_out* out, float dt,
                              int iter_slots, int* handle) {
    NSPH_TRY(check_step_args(S, in, out, dt));
    NSPH_CHECK(handle != nullptr, NSPH_ERR_INVALID, "nsph_solver_graph_capture: NULL handle");
    NSPH_CHECK(iter_slots >= 1 && iter_slots <= S->max_iters, NSPH_ERR_INVALID,
               "nsph_solver_graph_capture: iter_slots must be in [1, max_iters]");
    if (!S->cap_stream) NSPH_HIP(hipStreamCreateWithFlags(&S->cap_stream, hipStreamNonBlocking));
    NSPH_HIP(hipStreamSynchronize(S->ctx->stream));
    const int timing = S->timing;
    S->timing = 0;
    const int forced = S->force_rebuild;
    S->force_rebuild = 0;
    StepGraph G{};
    G.in = *in;
    G.out = *out;
    G.dt = dt;
    G.slots = S->persist ? S->max_iters : iter_slots;
    G.persist = S->persist ? 1 + S->ppasses : 0;
    G.warm = S->warm;
    G.mlp = S->warm == NSPH_WARM_MLP ? S->mlp : nullptr;
    hipStream_t cs = S->cap_stream;
    int rc = NSPH_OK;
    for (int part = 0; part < 2 && rc == NSPH_OK; ++part) {
        hipGraph_t g = nullptr;
        if (hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal) != hipSuccess) {
            rc = NSPH_ERR_HIP;
            set_error("nsph_solver_graph_capture: hipStreamBeginCapture failed");
            break;
        }
        int launched = 0;
        if (part == 0) {
            rc = enqueue_prologue(S, cs, G.in, dt);
            if (rc == NSPH_OK && !S->persist) rc = enqueue_iterations(S, cs, dt, 0, iter_slots, 0, &launched);
            if (rc == NSPH_OK && S->persist) rc = enqueue_world_tail(S, cs, dt);
        } else {
            rc = enqueue_epilogue(S, cs, G.in, G.out, dt);
        }
        hipError_t e = hipStreamEndCapture(cs, &g);
        if (rc == NSPH_OK && e != hipSuccess) {
            rc = NSPH_ERR_HIP;
            set_error("nsph_solver_graph_capture: capture failed: %s", hipGetErrorString(e));
        }
        if (rc == NSPH_OK) {
            hipGraphExec_t ex = nullptr;
            e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
            if (e != hipSuccess) {
                rc = NSPH_ERR_HIP;
                set_error("nsph_solver_graph_capture: instantiate failed: %s", hipGetErrorString(e));
            } else if (part == 0) {
                G.pro = ex;
            } else {
                G.epi = ex;
            }
        }
        if (g) (void)hipGraphDestroy(g);
    }
    S->timing = timing;
    S->force_rebuild = forced;
    if (rc != NSPH_OK) {
        if (G.pro) (void)hipGraphExecDestroy(G.pro);
        if (G.epi) (void)hipGraphExecDestroy(G.epi);
        return rc;
    }
    S->graphs.push_back(G);
    *handle = (int)S->graphs.size() - 1;
    return NSPH_OK;
}

int nsph_solver_graph_step(nsph_solver* S, int handle, nsph_step_report* report) {
    NSPH_CHECK(S != nullptr, NSPH_ERR_INVALID, "nsph_solver_graph_step: NULL solver");
    NSPH_CHECK(handle >= 0 && handle < (int)S->graphs.size() && S->graphs[handle].pro, NSPH_ERR_INVALID,
               "nsph_solver_graph_step: unknown graph %d", handle);
    const StepGraph& G = S->graphs[handle];
    NSPH_CHECK(G.warm == S->warm && G.mlp == (S->warm == NSPH_WARM_MLP ? S->mlp : nullptr), NSPH_ERR_INVALID,
               "nsph_solver_graph_step: graph %d was captured with another warm start or MLP; capture again",
               handle);
    NSPH_CHECK(G.persist == (S->persist ? 1 + S->ppasses : 0), NSPH_ERR_INVALID,
               "nsph_solver_graph_step: graph %d was captured with another solve schedule; capture again", handle);
    hipStream_t st = S->ctx->stream;
    NSPH_HIP(hipGraphLaunch(G.pro, st));
    int launched = G.persist ? -1 : G.slots;
    if (!G.persist && G.slots < S->max_iters && S->Nf > 0) {
        // the slots may not have been enough: read the still-iterating count once and finish
        // eagerly (same kernels, same order, so results equal the eager step's)
        NSPH_HIP(hipMemcpyAsync(S->h_pinned, bp<int>(S->act) + ((G.slots - 1) & 1), 4,
                                hipMemcpyDeviceToHost, st));
        NSPH_HIP(hipStreamSynchronize(st));
        if (S->h_pinned[0] != 0)
            NSPH_TRY(enqueue_iterations(S, st, G.dt, G.slots, S->max_iters, 1, &launched));
    }
    S->last_iters_launched = launched;
    NSPH_HIP(hipGraphLaunch(G.epi, st));
    S->stepped = 1;
    if (report) NSPH_TRY(collect_report(S, report));
    return NSPH_OK;
}

int nsph_solver_graph_destroy(nsph_solver* S, int handle) {
    NSPH_CHECK(S != nullptr, NSPH_ERR_INVALID, "nsph_solver_graph_destroy: NULL solver");
    NSPH_CHECK(handle >= 0 && handle < (int)S->graphs.size(), NSPH_ERR_INVALID,
               "nsph_solver_graph_destroy: unknown graph %d", handle);
    StepGraph& G = S->graphs[handle];
    if (G.pro || G.epi) NSPH_HIP(hipStreamSynchronize(S->ctx->stream));
    if (G.pro) NSPH_HIP(hipGraphExecDestroy(G.pro));
    if (G.epi) NSPH_HIP(hipGraphExecDestroy(G.epi));
    G.pro = nullptr;
    G.epi = nullptr;
    return NSPH_OK;
}

int nsph_solver_report(nsph_solver* S, nsph_step_report* report) {
    NSPH_CHECK(S && report, NSPH_ERR_INVALID, "nsph_solver_report: NULL argument");
    NSPH_CHECK(S->stepped, NSPH_ERR_NOT_BUILT, "nsph_solver_report: no step has run");
    return collect_report(S, report);
}

int nsph_solver_fields(nsph_solver* S, float* d_rho, float* d_p, int* d_cnt) {
    NSPH_CHECK(S != nullptr, NSPH_ERR_INVALID, "nsph_solver_fields: NULL solver");
    NSPH_CHECK(S->stepped, NSPH_ERR_NOT_BUILT, "nsph_solver_fields: no step has run");
    hipStream_t st = S->ctx->stream;
    const size_t N = (size_t)S->N;
    if (N == 0) return NSPH_OK;
    if (d_rho) hipLaunchKernelGGL(k_f32_zero, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, st, d_rho, N);
    if (d_p) hipLaunchKernelGGL(k_f32_zero, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, st, d_p, N);
    if (d_cnt) hipLaunchKernelGGL(k_fill_i32, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, st, d_cnt, N, 0);
    if (S->Nf > 0)
        hipLaunchKernelGGL(k_sph_fields, dim3((S->Nf + 255) / 256), dim3(256), 0, st, S->perm,
                           bp<int>(S->fid), S->Nf, bp<float>(S->rho), bp<float>(S->p),
                           bp<int>(S->ncnt), d_rho, d_p, d_cnt);
    NSPH_HIP(hipGetLastError());
    return NSPH_OK;
}

int nsph_solver_world_stats(nsph_solver* S, int* h_iters, double* h_err, int* h_state) {
    NSPH_CHECK(S != nullptr, NSPH_ERR_INVALID, "nsph_solver_world_stats: NULL solver");
    NSPH_CHECK(S->stepped, NSPH_ERR_NOT_BUILT, "nsph_solver_world_stats: no step has run");
    hipStream_t st = S->ctx->stream;
    const size_t W = (size_t)S->W;
    if (h_iters) NSPH_HIP(hipMemcpyAsync(h_iters, S->iters.ptr, W * 4, hipMemcpyDeviceToHost, st));
    if (h_err) NSPH_HIP(hipMemcpyAsync(h_err, S->err.ptr, W * 8, hipMemcpyDeviceToHost, st));
    if (h_state) NSPH_HIP(hipMemcpyAsync(h_state, S->conv.ptr, W * 4, hipMemcpyDeviceToHost, st));
    NSPH_HIP(hipStreamSynchronize(st));
    return NSPH_OK;
}

int nsph_solver_neighbors(nsph_solver* S, int* d_counts, int* d_idx) {
    NSPH_CHECK(S && d_counts && d_idx, NSPH_ERR_INVALID, "nsph_solver_neighbors: NULL argument");
    NSPH_CHECK(S->stepped, NSPH_ERR_NOT_BUILT, "nsph_solver_neighbors: no step has run");
    hipStream_t st = S->ctx->stream;
    const size_t N = (size_t)S->N;
    if (N == 0) return NSPH_OK;
    hipLaunchKernelGGL(k_fill_i32, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, st, d_counts, N, 0);
    if (S->Nf > 0)
        hipLaunchKernelGGL(k_sph_export_nbrs, dim3((S->Nf + 255) / 256), dim3(256), 0, st, S->perm,
                           bp<int>(S->fid), S->Nf, S->N, S->cap, bp<int4>(S->nbr), bp<int>(S->ncnt),
                           d_counts, d_idx);
    NSPH_HIP(hipGetLastError());
    return NSPH_OK;
}

int nsph_solver_set_warm_start(nsph_solver* S, int mode) {
    NSPH_CHECK(S != nullptr, NSPH_ERR_INVALID, "nsph_solver_set_warm_start: NULL solver");
    NSPH_CHECK(mode >= NSPH_WARM_ZERO && mode <= NSPH_WARM_MLP, NSPH_ERR_INVALID,
               "nsph_solver_set_warm_start: unknown mode %d", mode);
    S->warm = mode;
    return NSPH_OK;
}

int nsph_solver_rebuild_lists(nsph_solver* S) {
    NSPH_CHECK(S != nullptr, NSPH_ERR_INVALID, "nsph_solver_rebuild_lists: NULL solver");
    S->force_rebuild = 1;
    return NSPH_OK;
}

int nsph_solver_reset_pressure(nsph_solver* S) {
    NSPH_CHECK(S != nullptr, NSPH_ERR_INVALID, "nsph_solver_reset_pressure: NULL solver");
    const size_t n = (size_t)S->Nf + 1;
    hipLaunchKernelGGL(k_f32_zero, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, S->ctx->stream,
                       bp<float>(S->pprev), n);
    NSPH_HIP(hipGetLastError());
    return NSPH_OK;
}

int nsph_solver_warm_pressure(nsph_solver* S, float* d_p0) {
    NSPH_CHECK(S && d_p0, NSPH_ERR_INVALID, "nsph_solver_warm_pressure: NULL argument");
    NSPH_CHECK(S->stepped, NSPH_ERR_NOT_BUILT, "nsph_solver_warm_pressure: no step has run");
    hipStream_t st = S->ctx->stream;
    const size_t N = (size_t)S->N;
    if (N == 0) return NSPH_OK;
    hipLaunchKernelGGL(k_f32_zero, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, st, d_p0, N);
    if (S->Nf > 0)
        hipLaunchKernelGGL(k_sph_fields, dim3((S->Nf + 255) / 256), dim3(256), 0, st, S->perm,
                           bp<int>(S->fid), S->Nf, (const float*)nullptr, bp<float>(S->p0), (const int*)nullptr,
                           (float*)nullptr, d_p0, (int*)nullptr);
    NSPH_HIP(hipGetLastError());
    return NSPH_OK;
}

int nsph_solver_num_fluid(const nsph_solver* S, int* n_fluid, int* n_boundary) {
    NSPH_CHECK(S != nullptr, NSPH_ERR_INVALID, "nsph_solver_num_fluid: NULL solver");
    if (n_fluid) *n_fluid = S->Nf;
    if (n_boundary) *n_boundary = S->Nb + S->Nk;
    return NSPH_OK;
}

int nsph_solver_set_mlp(nsph_solver* S, nsph_mlp* mlp) {
    NSPH_CHECK(S != nullptr, NSPH_ERR_INVALID, "nsph_solver_set_mlp: NULL solver");
    S->mlp = mlp;
    S->mlp_ready = mlp != nullptr;
    if (!mlp && S->warm == NSPH_WARM_MLP) S->warm = NSPH_WARM_ZERO;
    return NSPH_OK;
}

int nsph_solver_mlp_features(nsph_solver* S, float* d_x, float* d_target) {
    NSPH_CHECK(S && d_x && d_target, NSPH_ERR_INVALID, "nsph_solver_mlp_features: NULL argument");
    NSPH_CHECK(S->stepped, NSPH_ERR_NOT_BUILT, "nsph_solver_mlp_features: no step has run");
    if (S->Nf > 0) {
        hipLaunchKernelGGL(k_sph_mlp_features, dim3(S->C), dim3(CHUNK), 0, S->ctx->stream, make_dev(S),
                           S->last_dt, d_x, d_target);
        NSPH_HIP(hipGetLastError());
    }
    return NSPH_OK;
}

int nsph_solver_set_launch(nsph_solver* S, int gather_block) {
    NSPH_CHECK(S != nullptr, NSPH_ERR_INVALID, "nsph_solver_set_launch: NULL solver");
    NSPH_CHECK(gather_block == 64 || gather_block == 128 || gather_block == 256, NSPH_ERR_INVALID,
               "nsph_solver_set_launch: gather_block must be 64, 128 or 256 (got %d)", gather_block);
    S->gblock = gather_block;
    return NSPH_OK;
}

int nsph_solver_set_search(nsph_solver* S, int lds_quads) {
    NSPH_CHECK(S != nullptr, NSPH_ERR_INVALID, "nsph_solver_set_search: NULL solver");
    for (int v = 0; v < NBR_VARIANTS; ++v) {
        if (kNbrQuads[v] == lds_quads) {
            S->nbr_variant = v;
            return NSPH_OK;
        }
    }
    set_error("nsph_solver_set_search: no neighbour-kernel variant with %d LDS list quads", lds_quads);
    return NSPH_ERR_INVALID;
}

int nsph_solver_set_solve(nsph_solver* S, int mode, int block, int pass_iters) {
    NSPH_CHECK(S != nullptr, NSPH_ERR_INVALID, "nsph_solver_set_solve: NULL solver");
    NSPH_CHECK(mode == NSPH_SOLVE_PASSES || mode == NSPH_SOLVE_WORLDS, NSPH_ERR_INVALID,
               "nsph_solver_set_solve: mode must be NSPH_SOLVE_PASSES or NSPH_SOLVE_WORLDS");
    NSPH_CHECK(block == 256 || block == 512 || block == 1024, NSPH_ERR_INVALID,
               "nsph_solver_set_solve: block must be 256, 512 or 1024");
    NSPH_CHECK(pass_iters >= 0, NSPH_ERR_INVALID, "nsph_solver_set_solve: pass_iters must be >= 0");
    S->persist = mode == NSPH_SOLVE_WORLDS ? 1 : 0;
    S->pblock = block;
    S->ppasses = pass_iters;
    return NSPH_OK;
}

int nsph_solver_set_timing(nsph_solver* S, int enable) {
    NSPH_CHECK(S != nullptr, NSPH_ERR_INVALID, "nsph_solver_set_timing: NULL solver");
    NSPH_TRY(flush_timing(S));
    S->timing = enable;
    for (int k = 0; k < KT_COUNT; ++k) {
        S->t_ms[k] = 0.0;
        S->t_count[k] = 0;
    }
    return NSPH_OK;
}

int nsph_solver_timing(nsph_solver* S, double* h_ms, int* h_launches) {
    NSPH_CHECK(S && h_ms && h_launches, NSPH_ERR_INVALID, "nsph_solver_timing: NULL argument");
    NSPH_TRY(flush_timing(S));
    for (int k = 0; k < KT_COUNT; ++k) {
        h_ms[k] = S->t_ms[k];
        h_launches[k] = (int)S->t_count[k];
    }
    return NSPH_OK;
}

int nsph_kernel_count(void) { return KT_COUNT; }

const char* nsph_kernel_name(int k) { return (k >= 0 && k < KT_COUNT) ? kKernelNames[k] : ""; }

int nsph_solver_iterations_launched(const nsph_solver* S, int* n) {
    NSPH_CHECK(S && n, NSPH_ERR_INVALID, "nsph_solver_iterations_launched: NULL argument");
    *n = S->last_iters_launched;
    return NSPH_OK;
}

int nsph_solver_destroy(nsph_solver* S) {
    if (!S) return NSPH_OK;
    const hipError_t e = nsph_ctx_drain(S->ctx);
    solver_free(S);
    delete S;
    NSPH_HIP(e);
    return NSPH_OK;
}

}  // extern "C"
