// Particle-shape soft contacts against static shapes (body = -1), the reference's
//   create_soft_contacts              newton/_src/geometry/kernels.py:1000-1141 (launch collide.py:1194-1230)
//   solve_particle_shape_contacts     newton/_src/solvers/xpbd/kernels.py:114-222
//   eval_particle_body_contact        newton/_src/solvers/semi_implicit/kernels_contact.py:144-262
// with the signed-distance functions of geometry/kernels.py:132-600 for planes, spheres, boxes,
// capsules and cylinders (capsule and cylinder axes along local Z, as create_soft_contacts uses).
//
// Contacts are generated without an atomic counter: one thread per particle counts its
// contacts over all shapes, a scan gives each particle its first slot, and a second pass
// writes them, so the list is ordered by (particle, shape) and identical run to run.
// soft_contact_count holds the total (which may exceed the capacity, as the reference's
// counter does); slots at or beyond the capacity are not written.  The consumers loop over
// each particle's own contacts (start[i] .. start[i+1]) instead of adding atomically, so their
// sums are deterministic too.

#include <cmath>

#include <mutex>

#include "common.h"

namespace nsph {

enum GeoType { GEO_PLANE = 1, GEO_SPHERE = 3, GEO_CAPSULE = 4, GEO_CYLINDER = 6, GEO_BOX = 7 };
constexpr int SHAPE_COLLIDE_PARTICLES = 1 << 2;

struct V3 {
    float x, y, z;
};
__device__ inline V3 v3(float x, float y, float z) { return V3{x, y, z}; }
__device__ inline V3 operator+(V3 a, V3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ inline V3 operator-(V3 a, V3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ inline V3 operator*(V3 a, float s) { return v3(a.x * s, a.y * s, a.z * s); }
__device__ inline V3 operator/(V3 a, float s) { return v3(a.x / s, a.y / s, a.z / s); }
__device__ inline float dot(V3 a, V3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
__device__ inline float len(V3 a) { return sqrtf(dot(a, a)); }
__device__ inline V3 cross(V3 a, V3 b) {
    return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
__device__ inline V3 unit_or_zero(V3 a) {
    const float l = len(a);
    return l > 0.0f ? a * (1.0f / l) : v3(0.f, 0.f, 0.f);
}
__device__ inline V3 ld3(const float* p, int i) { return v3(p[3 * i], p[3 * i + 1], p[3 * i + 2]); }
__device__ inline void st3(float* p, int i, V3 a) {
    p[3 * i] = a.x;
    p[3 * i + 1] = a.y;
    p[3 * i + 2] = a.z;
}

// Rotation by the unit quaternion (x, y, z, w): v' = v + 2w (u x v) + 2 u x (u x v).
__device__ inline V3 qrot(float4 q, V3 v) {
    const V3 u = v3(q.x, q.y, q.z);
    const V3 t = cross(u, v) * 2.0f;
    return v + t * q.w + cross(u, t);
}
__device__ inline V3 qrot_inv(float4 q, V3 v) { return qrot(make_float4(-q.x, -q.y, -q.z, q.w), v); }

// ---- signed distances (negative inside) and outward gradients, shape-local frame ---------

__device__ inline float sdf_sphere(V3 p, float r) { return len(p) - r; }
__device__ inline V3 sdf_sphere_grad(V3 p) {
    const float l = len(p);
    return l > 1.0e-8f ? p * (1.0f / l) : v3(0.f, 0.f, 1.f);
}

__device__ inline float sdf_box(V3 p, V3 h) {
    const float qx = fabsf(p.x) - h.x, qy = fabsf(p.y) - h.y, qz = fabsf(p.z) - h.z;
    const V3 e = v3(fmaxf(qx, 0.f), fmaxf(qy, 0.f), fmaxf(qz, 0.f));
    return len(e) + fminf(fmaxf(qx, fmaxf(qy, qz)), 0.0f);
}
__device__ inline float sgn(float a) { return a > 0.f ? 1.f : (a < 0.f ? -1.f : 0.f); }
__device__ inline V3 sdf_box_grad(V3 p, V3 h) {
    const float qx = fabsf(p.x) - h.x, qy = fabsf(p.y) - h.y, qz = fabsf(p.z) - h.z;
    if (qx > 0.f || qy > 0.f || qz > 0.f) {
        const V3 c = v3(fminf(fmaxf(p.x, -h.x), h.x), fminf(fmaxf(p.y, -h.y), h.y), fminf(fmaxf(p.z, -h.z), h.z));
        return unit_or_zero(p - c);
    }
    if ((qx > qy && qx > qz) || (qy == 0.f && qz == 0.f)) return v3(sgn(p.x), 0.f, 0.f);
    if ((qy > qx && qy > qz) || (qx == 0.f && qz == 0.f)) return v3(0.f, sgn(p.y), 0.f);
    return v3(0.f, 0.f, sgn(p.z));
}

__device__ inline float sdf_capsule(V3 p, float r, float hh) {
    if (p.z > hh) return len(v3(p.x, p.y, p.z - hh)) - r;
    if (p.z < -hh) return len(v3(p.x, p.y, p.z + hh)) - r;
    return len(v3(p.x, p.y, 0.f)) - r;
}
__device__ inline V3 sdf_capsule_grad(V3 p, float hh) {
    V3 v = p.z > hh ? v3(p.x, p.y, p.z - hh) : (p.z < -hh ? v3(p.x, p.y, p.z + hh) : v3(p.x, p.y, 0.f));
    const float l = len(v);
    if (l > 1.0e-8f) return v * (1.0f / l);
    return v3(0.f, 0.f, p.z < -hh ? -1.f : 1.f);
}

__device__ inline float sdf_cylinder(V3 p, float r, float hh) {
    const float dx = len(v3(p.x, p.y, 0.f)) - r, dy = fabsf(p.z) - hh;
    const float ex = fmaxf(dx, 0.f), ey = fmaxf(dy, 0.f);
    return fminf(fmaxf(dx, dy), 0.f) + sqrtf(ex * ex + ey * ey);
}
__device__ inline V3 sdf_cylinder_grad(V3 p, float r, float hh) {
    const V3 v = v3(p.x, p.y, 0.f);
    const float l = len(v);
    const V3 radial = l > 1.0e-8f ? v * (1.0f / l) : v3(0.f, 0.f, 1.f);
    const V3 axial = v3(0.f, 0.f, sgn(p.z));
    const float dx = l - r, dy = fabsf(p.z) - hh;
    if (dx > 0.f && dy > 0.f) {
        const V3 g = radial * dx + axial * dy;
        const float gl = len(g);
        return gl > 1.0e-8f ? g * (1.0f / gl) : radial;
    }
    return dx > dy ? radial : axial;
}

// Plane: local XY quad of half extents (w, l); with w <= 0 or l <= 0 the infinite plane z = 0.
// The finite quad uses the reference's Chebyshev distance max(|x| - w, |y| - l, |z|).
__device__ inline float sdf_plane(V3 p, float w, float l) {
    if (w > 0.f && l > 0.f) return fmaxf(fmaxf(fabsf(p.x) - w, fabsf(p.y) - l), fabsf(p.z));
    return p.z;
}

struct ShapeView {
    const float* xform;  // (S, 7): px py pz qx qy qz qw
    const int* type;
    const float* scale;  // (S, 3)
    const int* world;
    const int* flags;
    int count;
};

// Distance and normal of local point p to shape s; false for a type without particle contacts.
__device__ inline bool shape_eval(const ShapeView& S, int s, V3 p, float* d, V3* n) {
    const V3 sc = ld3(S.scale, s);
    switch (S.type[s]) {
        case GEO_SPHERE: *d = sdf_sphere(p, sc.x); *n = sdf_sphere_grad(p); return true;
        case GEO_BOX: *d = sdf_box(p, sc); *n = sdf_box_grad(p, sc); return true;
        case GEO_CAPSULE: *d = sdf_capsule(p, sc.x, sc.y); *n = sdf_capsule_grad(p, sc.y); return true;
        case GEO_CYLINDER: *d = sdf_cylinder(p, sc.x, sc.y); *n = sdf_cylinder_grad(p, sc.x, sc.y); return true;
        case GEO_PLANE: *d = sdf_plane(p, sc.x * 0.5f, sc.y * 0.5f); *n = v3(0.f, 0.f, 1.f); return true;
        default: return false;
    }
}

__device__ inline bool pair_active(const ShapeView& S, int s, int pworld) {
    if ((S.flags[s] & SHAPE_COLLIDE_PARTICLES) == 0) return false;
    const int sw = S.world[s];
    return !(pworld != -1 && sw != -1 && pworld != sw);
}

// Visits the contacts of particle i in shape order: f(s, local point, d, n, pose).
template <class F>
__device__ inline void particle_shape_contacts(const ShapeView& S, V3 x, float radius, int pworld, float margin,
                                               F f) {
    for (int s = 0; s < S.count; ++s) {
        if (!pair_active(S, s, pworld)) continue;
        const V3 pos = v3(S.xform[7 * s], S.xform[7 * s + 1], S.xform[7 * s + 2]);
        const float4 q = make_float4(S.xform[7 * s + 3], S.xform[7 * s + 4], S.xform[7 * s + 5], S.xform[7 * s + 6]);
        const V3 local = qrot_inv(q, x - pos);
        float d = 1.0e6f;
        V3 n = v3(0.f, 0.f, 0.f);
        if (!shape_eval(S, s, local, &d, &n)) continue;
        if (d < margin + radius) f(s, local, d, n, pos, q);
    }
}

// Contacts per particle (work[i]) and per workgroup of 256 particles (bsum[block]).
__global__ __launch_bounds__(256) void k_soft_count(int n, const float* __restrict__ x,
                                                    const float* __restrict__ radius, const int* __restrict__ flags,
                                                    const int* __restrict__ world, ShapeView S, float margin,
                                                    int* __restrict__ counts, int* __restrict__ bsum) {
    __shared__ int wsum[4];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    int c = 0;
    if (i < n && (flags[i] & NSPH_PARTICLE_ACTIVE) != 0)
        particle_shape_contacts(S, ld3(x, i), radius[i], world ? world[i] : -1, margin,
                                [&](int, V3, float, V3, V3, float4) { ++c; });
    if (i < n) counts[i] = c;
    int v = c;
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) bsum[blockIdx.x] = (wsum[0] + wsum[1]) + (wsum[2] + wsum[3]);
}

// Exclusive scan of the nb workgroup sums in place (one workgroup of 1024 threads walking the
// array in tiles of 4096); start[n] and *total_out receive the total.
__global__ __launch_bounds__(1024) void k_soft_scan_blocks(int nb, int* __restrict__ bsum, int n,
                                                           int* __restrict__ start, int* __restrict__ total_out) {
    __shared__ int sh[1024];
    __shared__ int carry;
    const int t = threadIdx.x;
    if (t == 0) carry = 0;
    __syncthreads();
    for (int base = 0; base < nb; base += 1024 * 4) {
        int v[4], sum = 0;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int i = base + t * 4 + k;
            v[k] = i < nb ? bsum[i] : 0;
            sum += v[k];
        }
        sh[t] = sum;
        __syncthreads();
        for (int off = 1; off < 1024; off <<= 1) {
            const int y = t >= off ? sh[t - off] : 0;
            __syncthreads();
            sh[t] += y;
            __syncthreads();
        }
        int run = carry + sh[t] - sum;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int i = base + t * 4 + k;
            if (i < nb) bsum[i] = run;
            run += v[k];
        }
        __syncthreads();
        if (t == 1023) carry += sh[1023];
        __syncthreads();
    }
    if (t == 0) {
        start[n] = carry;
        total_out[0] = carry;
    }
}

// First contact slot of each particle (its workgroup's offset plus the exclusive scan of the
// workgroup's counts), then the contacts themselves.
__global__ __launch_bounds__(256) void k_soft_write(int n, const float* __restrict__ x,
                                                    const float* __restrict__ radius, const int* __restrict__ flags,
                                                    const int* __restrict__ world, ShapeView S, float margin,
                                                    const int* __restrict__ counts, const int* __restrict__ boff,
                                                    int* __restrict__ start, int cap, int* __restrict__ c_particle,
                                                    int* __restrict__ c_shape, float* __restrict__ c_body_pos,
                                                    float* __restrict__ c_body_vel, float* __restrict__ c_normal) {
    __shared__ int sh[256];
    const int t = threadIdx.x;
    const int i = blockIdx.x * blockDim.x + t;
    const int c = i < n ? counts[i] : 0;
    sh[t] = c;
    __syncthreads();
    for (int off = 1; off < 256; off <<= 1) {
        const int y = t >= off ? sh[t - off] : 0;
        __syncthreads();
        sh[t] += y;
        __syncthreads();
    }
    if (i >= n) return;
    int k = boff[blockIdx.x] + sh[t] - c;
    start[i] = k;
    if ((flags[i] & NSPH_PARTICLE_ACTIVE) == 0) return;
    particle_shape_contacts(S, ld3(x, i), radius[i], world ? world[i] : -1, margin,
                            [&](int s, V3 local, float d, V3 nrm, V3 pos, float4 q) {
                                if (k < cap) {
                                    // contact point and normal of a static shape in world space
                                    st3(c_body_pos, k, qrot(q, local - nrm * d) + pos);
                                    st3(c_body_vel, k, v3(0.f, 0.f, 0.f));
                                    st3(c_normal, k, qrot(q, nrm));
                                    c_particle[k] = i;
                                    c_shape[k] = s;
                                }
                                ++k;
                            });
}

// xpbd/kernels.py:114-222 for static shapes (body velocity zero, w2 = 0): per contact
//   c = n . (x - p_b) - r;  skip if c > ka;  mu = (particle_mu + shape_mu) / 2
//   delta = (normalize(v_t) max(mu c, -|v_t| dt) - n c) / w1 * relaxation, added as w1 delta.
__global__ void k_xpbd_shape_contacts(int n, const float* __restrict__ x, const float* __restrict__ v,
                                      const float* __restrict__ inv_mass, const float* __restrict__ radius,
                                      const int* __restrict__ flags, const float* __restrict__ shape_mu,
                                      float particle_mu, float particle_ka, const int* __restrict__ start,
                                      int cap, const int* __restrict__ c_shape, const float* __restrict__ c_body_pos,
                                      const float* __restrict__ c_body_vel, const float* __restrict__ c_normal,
                                      float dt, float relaxation, float* __restrict__ deltas) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if ((flags[i] & NSPH_PARTICLE_ACTIVE) == 0) return;
    const int k0 = min(start[i], cap), k1 = min(start[i + 1], cap);
    if (k0 >= k1) return;
    const float w1 = inv_mass[i];
    if (w1 == 0.0f) return;
    const V3 px = ld3(x, i), pv = ld3(v, i);
    V3 acc = v3(0.f, 0.f, 0.f);
    for (int k = k0; k < k1; ++k) {
        const V3 bx = ld3(c_body_pos, k), nrm = ld3(c_normal, k);
        const float c = dot(nrm, px - bx) - radius[i];
        if (c > particle_ka) continue;
        const float mu = 0.5f * (particle_mu + shape_mu[c_shape[k]]);
        const V3 rv = pv - ld3(c_body_vel, k);
        const float vn = dot(nrm, rv);
        const V3 vt = rv - nrm * vn;
        const float lambda_f = fmaxf(mu * c, -len(vt) * dt);
        const V3 dtot = ((unit_or_zero(vt) * lambda_f - nrm * c) / w1) * relaxation;
        acc = acc + dtot * w1;
    }
    st3(deltas, i, ld3(deltas, i) + acc);
}

// kernels_contact.py:144-262 for static shapes: per contact with c <= ka
//   f = n c ke + n min(v_n, 0) kd + normalize(v_t) min(kf |v_t|, |mu c ke|);  particle_f -= f
// with material constants averaged between particle and shape.
__global__ void k_eval_shape_contacts(int n, const float* __restrict__ x, const float* __restrict__ v,
                                      const float* __restrict__ radius, const int* __restrict__ flags,
                                      const float* __restrict__ s_ke, const float* __restrict__ s_kd,
                                      const float* __restrict__ s_kf, const float* __restrict__ s_mu,
                                      float p_ke, float p_kd, float p_kf, float p_mu, float p_ka,
                                      const int* __restrict__ start, int cap, const int* __restrict__ c_shape,
                                      const float* __restrict__ c_body_pos, const float* __restrict__ c_body_vel,
                                      const float* __restrict__ c_normal, float* __restrict__ f_out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if ((flags[i] & NSPH_PARTICLE_ACTIVE) == 0) return;
    const int k0 = min(start[i], cap), k1 = min(start[i + 1], cap);
    if (k0 >= k1) return;
    const V3 px = ld3(x, i), pv = ld3(v, i);
    V3 f = ld3(f_out, i);
    for (int k = k0; k < k1; ++k) {
        const int s = c_shape[k];
        const V3 bx = ld3(c_body_pos, k), nrm = ld3(c_normal, k);
        const float c = dot(nrm, px - bx) - radius[i];
        if (c > p_ka) continue;
        const float ke = 0.5f * (p_ke + s_ke[s]), kd = 0.5f * (p_kd + s_kd[s]);
        const float kf = 0.5f * (p_kf + s_kf[s]), mu = 0.5f * (p_mu + s_mu[s]);
        const V3 rv = pv - ld3(c_body_vel, k);
        const float vn = dot(nrm, rv);
        const V3 vt = rv - nrm * vn;
        const V3 fn = (nrm * c) * ke;
        const V3 fd = (nrm * fminf(vn, 0.0f)) * kd;
        const V3 ft = unit_or_zero(vt) * fminf(kf * len(vt), fabsf(mu * c * ke));
        f = f - (fn + (fd + ft));
    }
    st3(f_out, i, f);
}

// Shape corrections after an SPH step: a particle with a nonzero correction d moves by d and
// gains d/dt of velocity (|v| clamped to v_max); every other particle is left bit for bit.
__global__ void k_apply_shape_deltas(int n, const int* __restrict__ flags, const float* __restrict__ deltas,
                                     float dt, float v_max, float* __restrict__ x, float* __restrict__ v) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if ((flags[i] & NSPH_PARTICLE_ACTIVE) == 0) return;
    const V3 d = ld3(deltas, i);
    if (d.x == 0.0f && d.y == 0.0f && d.z == 0.0f) return;
    st3(x, i, ld3(x, i) + d);
    V3 vn = ld3(v, i) + d / dt;
    const float m = len(vn);
    if (m > v_max) vn = vn * (v_max / m);
    st3(v, i, vn);
}

}  // namespace nsph

using namespace nsph;

extern "C" {

static int soft_contacts(const char* who, nsph_ctx* ctx, int n, const float* d_x, const float* d_radius,
                         const int* d_flags, const int* d_world, int n_shapes, const float* d_shape_xform,
                         const int* d_shape_type, const float* d_shape_scale, const int* d_shape_world,
                         const int* d_shape_flags, float margin, int capacity, int* d_work, int* d_start, int* d_count,
                         int* d_particle, int* d_shape, float* d_body_pos, float* d_body_vel, float* d_normal) {
    hipStream_t st = ctx->stream;
    if (n == 0 || n_shapes == 0) {
        hipLaunchKernelGGL(k_fill_i32, dim3((n + 256) / 256 + 1), dim3(256), 0, st, d_start, (size_t)n + 1, 0);
        hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, st, d_count, (size_t)1, 0);
        NSPH_HIP(hipGetLastError());
        return NSPH_OK;
    }
    NSPH_CHECK(d_x && d_radius && d_flags && d_shape_xform && d_shape_type && d_shape_scale && d_shape_world &&
                   d_shape_flags && d_particle && d_shape && d_body_pos && d_body_vel && d_normal,
               NSPH_ERR_INVALID, "%s: NULL array", who);
    ShapeView S{d_shape_xform, d_shape_type, d_shape_scale, d_shape_world, d_shape_flags, n_shapes};
    const int nb = (n + 255) / 256;
    int* counts = d_work;
    int* bsum = d_work + n;
    hipLaunchKernelGGL(k_soft_count, dim3(nb), dim3(256), 0, st, n, d_x, d_radius, d_flags, d_world, S, margin,
                       counts, bsum);
    hipLaunchKernelGGL(k_soft_scan_blocks, dim3(1), dim3(1024), 0, st, nb, bsum, n, d_start, d_count);
    hipLaunchKernelGGL(k_soft_write, dim3(nb), dim3(256), 0, st, n, d_x, d_radius, d_flags, d_world, S, margin,
                       counts, bsum, d_start, capacity, d_particle, d_shape, d_body_pos, d_body_vel, d_normal);
    NSPH_HIP(hipGetLastError());
    return NSPH_OK;
}

}  // extern "C"

// Workspace of nsph_soft_contacts_static: a file-scope device array, one copy per GPU, held
// until the process ends.  Host code never passes the array's name to a kernel (on the host it
// is only a stand-in for the symbol): nsph_soft_static_init resolves its device address once
// per device, with that device current, and every use goes through that address.
static __device__ __align__(256) int g_soft_scratch[NSPH_SOFT_WORK_INTS(NSPH_SOFT_STATIC_MAX_PARTICLES)];
static constexpr int kMaxDevices = 64;
static int* g_soft_scratch_ptr[kMaxDevices] = {};
static std::mutex g_soft_scratch_mu;

namespace nsph {
int soft_static_init(int device) {
    NSPH_CHECK(device >= 0 && device < kMaxDevices, NSPH_ERR_INVALID, "soft-contact scratch: device %d out of range",
               device);
    std::lock_guard<std::mutex> lock(g_soft_scratch_mu);
    if (g_soft_scratch_ptr[device]) return NSPH_OK;
    int* p = nullptr;
    NSPH_HIP(hipGetSymbolAddress((void**)&p, HIP_SYMBOL(g_soft_scratch)));
    NSPH_CHECK(p != nullptr, NSPH_ERR_HIP, "soft-contact scratch: no device address");
    g_soft_scratch_ptr[device] = p;
    return NSPH_OK;
}
}  // namespace nsph

extern "C" {

int nsph_soft_contacts(nsph_ctx* ctx, int n, const float* d_x, const float* d_radius, const int* d_flags,
                       const int* d_world, int n_shapes, const float* d_shape_xform, const int* d_shape_type,
                       const float* d_shape_scale, const int* d_shape_world, const int* d_shape_flags,
                       float margin, int capacity, int* d_work, int* d_start, int* d_count, int* d_particle,
                       int* d_shape, float* d_body_pos, float* d_body_vel, float* d_normal) {
    NSPH_CHECK(ctx && n >= 0 && n_shapes >= 0 && capacity >= 0, NSPH_ERR_INVALID,
               "nsph_soft_contacts: bad argument");
    NSPH_CHECK(d_start && d_count && d_work, NSPH_ERR_INVALID, "nsph_soft_contacts: NULL output");
    return soft_contacts("nsph_soft_contacts", ctx, n, d_x, d_radius, d_flags, d_world, n_shapes, d_shape_xform,
                         d_shape_type, d_shape_scale, d_shape_world, d_shape_flags, margin, capacity, d_work, d_start,
                         d_count, d_particle, d_shape, d_body_pos, d_body_vel, d_normal);
}

int nsph_soft_contacts_static(nsph_ctx* ctx, int n, const float* d_x, const float* d_radius, const int* d_flags,
                              const int* d_world, int n_shapes, const float* d_shape_xform, const int* d_shape_type,
                              const float* d_shape_scale, const int* d_shape_world, const int* d_shape_flags,
                              float margin, int capacity, int* d_start, int* d_count, int* d_particle, int* d_shape,
                              float* d_body_pos, float* d_body_vel, float* d_normal) {
    NSPH_CHECK(ctx && n >= 0 && n_shapes >= 0 && capacity >= 0, NSPH_ERR_INVALID,
               "nsph_soft_contacts_static: bad argument");
    NSPH_CHECK(d_start && d_count, NSPH_ERR_INVALID, "nsph_soft_contacts_static: NULL output");
    NSPH_CHECK(n <= NSPH_SOFT_STATIC_MAX_PARTICLES, NSPH_ERR_CAPACITY,
               "nsph_soft_contacts_static: %d particles exceed the static workspace (at most %d)", n,
               NSPH_SOFT_STATIC_MAX_PARTICLES);
    NSPH_CHECK(ctx->device >= 0 && ctx->device < kMaxDevices && g_soft_scratch_ptr[ctx->device], NSPH_ERR_INVALID,
               "nsph_soft_contacts_static: the static workspace of device %d is not set up (nsph_ctx_create)",
               ctx->device);
    return soft_contacts("nsph_soft_contacts_static", ctx, n, d_x, d_radius, d_flags, d_world, n_shapes,
                         d_shape_xform, d_shape_type, d_shape_scale, d_shape_world, d_shape_flags, margin, capacity,
                         g_soft_scratch_ptr[ctx->device], d_start, d_count, d_particle, d_shape, d_body_pos,
                         d_body_vel, d_normal);
}

int nsph_xpbd_shape_contacts(nsph_ctx* ctx, int n, const float* d_x, const float* d_v, const float* d_inv_mass,
                             const float* d_radius, const int* d_flags, const float* d_shape_mu, float particle_mu,
                             float particle_ka, const int* d_start, int capacity, const int* d_shape,
                             const float* d_body_pos, const float* d_body_vel, const float* d_normal, float dt,
                             float relaxation, float* d_deltas) {
    NSPH_CHECK(ctx && n >= 0 && capacity >= 0, NSPH_ERR_INVALID, "nsph_xpbd_shape_contacts: bad argument");
    if (n == 0 || capacity == 0) return NSPH_OK;
    NSPH_CHECK(d_x && d_v && d_inv_mass && d_radius && d_flags && d_shape_mu && d_start && d_shape && d_body_pos &&
                   d_body_vel && d_normal && d_deltas,
               NSPH_ERR_INVALID, "nsph_xpbd_shape_contacts: NULL array");
    hipLaunchKernelGGL(k_xpbd_shape_contacts, dim3((n + 255) / 256), dim3(256), 0, ctx->stream, n, d_x, d_v,
                       d_inv_mass, d_radius, d_flags, d_shape_mu, particle_mu, particle_ka, d_start, capacity, d_shape,
                       d_body_pos, d_body_vel, d_normal, dt, relaxation, d_deltas);
    NSPH_HIP(hipGetLastError());
    return NSPH_OK;
}

int nsph_eval_shape_contacts(nsph_ctx* ctx, int n, const float* d_x, const float* d_v, const float* d_radius,
                             const int* d_flags, const float* d_shape_ke, const float* d_shape_kd,
                             const float* d_shape_kf, const float* d_shape_mu, float ke, float kd, float kf, float mu,
                             float ka, const int* d_start, int capacity, const int* d_shape, const float* d_body_pos,
                             const float* d_body_vel, const float* d_normal, float* d_f) {
    NSPH_CHECK(ctx && n >= 0 && capacity >= 0, NSPH_ERR_INVALID, "nsph_eval_shape_contacts: bad argument");
    if (n == 0 || capacity == 0) return NSPH_OK;
    NSPH_CHECK(d_x && d_v && d_radius && d_flags && d_shape_ke && d_shape_kd && d_shape_kf && d_shape_mu && d_start &&
                   d_shape && d_body_pos && d_body_vel && d_normal && d_f,
               NSPH_ERR_INVALID, "nsph_eval_shape_contacts: NULL array");
    hipLaunchKernelGGL(k_eval_shape_contacts, dim3((n + 255) / 256), dim3(256), 0, ctx->stream, n, d_x, d_v,
                       d_radius, d_flags, d_shape_ke, d_shape_kd, d_shape_kf, d_shape_mu, ke, kd, kf, mu, ka, d_start,
                       capacity, d_shape, d_body_pos, d_body_vel, d_normal, d_f);
    NSPH_HIP(hipGetLastError());
    return NSPH_OK;
}

int nsph_apply_shape_deltas(nsph_ctx* ctx, int n, const int* d_flags, const float* d_deltas, float dt, float v_max,
                            float* d_x, float* d_v) {
    NSPH_CHECK(ctx && n >= 0 && dt > 0.f, NSPH_ERR_INVALID, "nsph_apply_shape_deltas: bad argument");
    if (n == 0) return NSPH_OK;
    NSPH_CHECK(d_flags && d_deltas && d_x && d_v, NSPH_ERR_INVALID, "nsph_apply_shape_deltas: NULL array");
    hipLaunchKernelGGL(k_apply_shape_deltas, dim3((n + 255) / 256), dim3(256), 0, ctx->stream, n, d_flags, d_deltas,
                       dt, v_max, d_x, d_v);
    NSPH_HIP(hipGetLastError());
    return NSPH_OK;
}

}  // extern "C"
