// The reference's own HashGrid consumers and particle integrator, re-hosted on the HIP grid:
//   integrate_particles              newton/_src/solvers/solver.py:10-47
//   solve_particle_particle_contacts newton/_src/solvers/xpbd/kernels.py:225-289
//   apply_particle_deltas            newton/_src/solvers/xpbd/kernels.py:776-806
//   eval_particle_contact            newton/_src/solvers/semi_implicit/kernels_contact.py:11-84
// One thread per particle in the grid's cell-sorted order (hash_grid_point_id), candidates
// from the grid's cells, distances from the current positions as the reference does (the
// grid may have been built from older positions).  Each thread writes only its own particle,
// so the reference's atomic_add into deltas becomes a plain store.
//
// A candidate's data is fetched in sorted order from arrays packed at the start of each call
// (px = (x, r) or four NaNs for an inactive particle, pv = (v, inverse mass)): one 16-byte
// load per candidate, from the contiguous run of its cell, instead of five loads through the
// permutation.  A NaN distance fails the contact test exactly as the reference's
// inactive-particle check skips the candidate, and the visit order is the reference's, so the
// results are the same.

#include <cmath>

#include "common.h"

namespace nsph {

// Candidates (sorted positions j; the point id is perm[j]) whose cell overlaps [x - r, x + r],
// in cell-visit order.
template <class Visit>
__device__ inline void grid_candidates(const GridView& G, float qx, float qy, float qz, float r,
                                       Visit visit) {
    const float rp = r + r * NSPH_RANGE_PAD;
    int x0, nx, y0, ny, z0, nz;
    axis_range(qx, rp, G.g.lo[0], G.g.inv_w, G.g.dim[0], G.g.wrap, &x0, &nx);
    axis_range(qy, rp, G.g.lo[1], G.g.inv_w, G.g.dim[1], G.g.wrap, &y0, &ny);
    axis_range(qz, rp, G.g.lo[2], G.g.inv_w, G.g.dim[2], G.g.wrap, &z0, &nz);
    for (int iz = 0; iz < nz; ++iz) {
        const int cz = wrap_idx(z0 + iz, G.g.dim[2], G.g.wrap);
        for (int iy = 0; iy < ny; ++iy) {
            const int cy = wrap_idx(y0 + iy, G.g.dim[1], G.g.wrap);
            const int row = G.g.cell_off + (cz * G.g.dim[1] + cy) * G.g.dim[0];
            for (int ix = 0; ix < nx; ++ix) {
                const int cx = wrap_idx(x0 + ix, G.g.dim[0], G.g.wrap);
                const int2 se = G.table[row + cx];
                for (int j = se.x; j < se.y; ++j) visit(j);
            }
        }
    }
}

// True when a pair is certainly farther apart than touching distance plus cohesion c, so that
// err = |n| - r_i - r_j > c without taking the square root: |n|^2 > 1.001 (r_i + r_j + c)^2
// means |n| > 1.0005 (r_i + r_j + c).  With c >= 0 that margin is far above the FP32 rounding
// of |n|^2, of d = |n| and of d - r_i - r_j (a few units in the last place of r_i + r_j + c),
// so a pair it rejects is one the exact test rejects; every other pair takes the exact test.
__device__ inline bool surely_apart(float3 n, float ri, float rj, float c) {
    const float lim = ri + rj + c;
    return c >= 0.0f && (n.x * n.x + n.y * n.y + n.z * n.z) > (lim * lim) * 1.001f;
}

__device__ inline float3 ld3(const float* p, int i) { return make_float3(p[3 * i], p[3 * i + 1], p[3 * i + 2]); }

// px[s] = (x, r) of point i (NaN for an inactive one), pv[s] = (v, inverse mass or 0), s = inv[i]:
// read in point order (coalesced), written to the sorted positions.
__global__ void k_pack_particles(GridView G, const float* __restrict__ x, const float* __restrict__ v,
                                 const float* __restrict__ radius, const float* __restrict__ invmass,
                                 const int* __restrict__ flags) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= G.n) return;
    const int s = G.inv[i];
    const float nan = __int_as_float(0x7fc00000);
    G.px[s] = (flags[i] & NSPH_PARTICLE_ACTIVE) ? make_float4(x[3 * i], x[3 * i + 1], x[3 * i + 2], radius[i])
                                                 : make_float4(nan, nan, nan, nan);
    G.pv[s] = make_float4(v[3 * i], v[3 * i + 1], v[3 * i + 2], invmass ? invmass[i] : 0.0f);
}
__device__ inline void st3(float* p, int i, float3 v) {
    p[3 * i] = v.x;
    p[3 * i + 1] = v.y;
    p[3 * i + 2] = v.z;
}
__device__ inline float dotf(float3 a, float3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
__device__ inline float lenf(float3 a) { return sqrtf(dotf(a, a)); }

// solver.py:10-47
__global__ void k_integrate_particles(int n, const float* __restrict__ x, const float* __restrict__ v,
                                      const float* __restrict__ f, const float* __restrict__ w,
                                      const int* __restrict__ flags, const int* __restrict__ world,
                                      const float* __restrict__ gravity, int n_gravity, float dt,
                                      float v_max, float* __restrict__ x_new, float* __restrict__ v_new) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float3 x0 = ld3(x, i);
    if ((flags[i] & NSPH_PARTICLE_ACTIVE) == 0) {
        st3(x_new, i, x0);
        return;
    }
    const float3 v0 = ld3(v, i);
    const float3 f0 = f ? ld3(f, i) : make_float3(0.f, 0.f, 0.f);
    const float inv_mass = w[i];
    int wd = world ? world[i] : 0;
    wd = wd < 0 ? 0 : (wd >= n_gravity ? n_gravity - 1 : wd);
    const float3 g = ld3(gravity, wd);
    const float stp = (-inv_mass < 0.0f) ? 1.0f : 0.0f;
    float3 v1 = make_float3(v0.x + (f0.x * inv_mass + g.x * stp) * dt, v0.y + (f0.y * inv_mass + g.y * stp) * dt,
                            v0.z + (f0.z * inv_mass + g.z * stp) * dt);
    const float mag = lenf(v1);
    if (mag > v_max) {
        const float sc = v_max / mag;
        v1 = make_float3(v1.x * sc, v1.y * sc, v1.z * sc);
    }
    st3(x_new, i, make_float3(x0.x + v1.x * dt, x0.y + v1.y * dt, x0.z + v1.z * dt));
    st3(v_new, i, v1);
}

// xpbd/kernels.py:225-289.  The reference atomic_adds into deltas[i]; here thread tid owns
// particle i = perm[tid] alone (perm is a permutation), so a plain read-add-store is the same
// accumulation without atomics.  Inactive particles leave deltas untouched.
__global__ void k_xpbd_particle_contacts(GridView G, const int* __restrict__ flags, float k_mu, float k_cohesion,
                                         float max_radius, float dt, float relaxation,
                                         float* __restrict__ deltas) {
    const int tid = blockIdx.x * blockDim.x + threadIdx.x;
    if (tid >= G.n) return;
    const int i = (int)G.perm[tid];
    if ((flags[i] & NSPH_PARTICLE_ACTIVE) == 0) return;
    float3 delta = make_float3(0.f, 0.f, 0.f);
    {
        const float4 pi = G.px[tid], qi = G.pv[tid];
        const float3 xi = make_float3(pi.x, pi.y, pi.z), vi = make_float3(qi.x, qi.y, qi.z);
        const float ri = pi.w, w1 = qi.w;
        grid_candidates(G, xi.x, xi.y, xi.z, ri + max_radius + k_cohesion, [&](int j) {
            if (j == tid) return;
            const float4 pj = G.px[j];
            float3 nrm = make_float3(xi.x - pj.x, xi.y - pj.y, xi.z - pj.z);
            if (surely_apart(nrm, ri, pj.w, k_cohesion)) return;
            const float d = lenf(nrm);
            const float err = d - ri - pj.w;
            if (err <= k_cohesion) {
                const float4 qj = G.pv[j];
                const float w2 = qj.w;
                const float denom = w1 + w2;
                if (!(denom > 0.0f)) return;
                nrm = make_float3(nrm.x / d, nrm.y / d, nrm.z / d);
                const float3 vrel = make_float3(vi.x - qj.x, vi.y - qj.y, vi.z - qj.z);
                const float lambda_n = err;
                const float3 delta_n = make_float3(nrm.x * lambda_n, nrm.y * lambda_n, nrm.z * lambda_n);
                const float vn = dotf(nrm, vrel);
                const float3 vt = make_float3(vrel.x - nrm.x * vn, vrel.y - nrm.y * vn, vrel.z - nrm.z * vn);
                const float vtl = lenf(vt);
                const float lambda_f = fmaxf(k_mu * lambda_n, -vtl * dt);
                // normalize(vt) is zero for a zero vector
                const float3 tdir = vtl > 0.0f ? make_float3(vt.x / vtl, vt.y / vtl, vt.z / vtl) : make_float3(0.f, 0.f, 0.f);
                const float3 delta_f = make_float3(tdir.x * lambda_f, tdir.y * lambda_f, tdir.z * lambda_f);
                delta.x += (delta_f.x - delta_n.x) / denom;
                delta.y += (delta_f.y - delta_n.y) / denom;
                delta.z += (delta_f.z - delta_n.z) / denom;
            }
        });
        delta = make_float3(delta.x * w1, delta.y * w1, delta.z * w1);
        delta = make_float3(delta.x * relaxation, delta.y * relaxation, delta.z * relaxation);
    }
    const float3 d0 = ld3(deltas, i);
    st3(deltas, i, make_float3(d0.x + delta.x, d0.y + delta.y, d0.z + delta.z));
}

// xpbd/kernels.py:776-806
__global__ void k_apply_particle_deltas(int n, const float* __restrict__ x_orig, const float* __restrict__ x_pred,
                                        const int* __restrict__ flags, const float* __restrict__ delta,
                                        float dt, float v_max, float* __restrict__ x_out,
                                        float* __restrict__ v_out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if ((flags[i] & NSPH_PARTICLE_ACTIVE) == 0) return;
    const float3 x0 = ld3(x_orig, i), xp = ld3(x_pred, i), d = ld3(delta, i);
    const float3 xn = make_float3(xp.x + d.x, xp.y + d.y, xp.z + d.z);
    float3 vn = make_float3((xn.x - x0.x) / dt, (xn.y - x0.y) / dt, (xn.z - x0.z) / dt);
    const float mag = lenf(vn);
    if (mag > v_max) {
        const float sc = v_max / mag;
        vn = make_float3(vn.x * sc, vn.y * sc, vn.z * sc);
    }
    st3(x_out, i, xn);
    st3(v_out, i, vn);
}

// kernels_contact.py:11-32
__device__ inline float3 particle_force(float3 n, float3 v, float c, float k_n, float k_d, float k_f, float k_mu) {
    const float vn = dotf(n, v);
    const float jn = c * k_n;
    const float jd = fminf(vn, 0.0f) * k_d;
    const float fn = jn + jd;
    float3 vt = make_float3(v.x - n.x * vn, v.y - n.y * vn, v.z - n.z * vn);
    const float vs = lenf(vt);
    if (vs > 0.0f) vt = make_float3(vt.x / vs, vt.y / vs, vt.z / vs);
    const float ft = fminf(vs * k_f, k_mu * fabsf(fn));
    return make_float3(-n.x * fn - vt.x * ft, -n.y * fn - vt.y * ft, -n.z * fn - vt.z * ft);
}

// kernels_contact.py:35-84 (overwrites f[i] for active particles)
__global__ void k_eval_particle_contact(GridView G, const int* __restrict__ flags, float k_contact, float k_damp,
                                        float k_friction, float k_mu, float k_cohesion, float max_radius,
                                        float* __restrict__ f_out) {
    const int tid = blockIdx.x * blockDim.x + threadIdx.x;
    if (tid >= G.n) return;
    const int i = (int)G.perm[tid];
    if ((flags[i] & NSPH_PARTICLE_ACTIVE) == 0) return;
    const float4 pi = G.px[tid], qi = G.pv[tid];
    const float3 xi = make_float3(pi.x, pi.y, pi.z), vi = make_float3(qi.x, qi.y, qi.z);
    const float ri = pi.w;
    float3 f = make_float3(0.f, 0.f, 0.f);
    grid_candidates(G, xi.x, xi.y, xi.z, ri + max_radius + k_cohesion, [&](int j) {
        if (j == tid) return;
        const float4 pj = G.px[j];
        float3 nrm = make_float3(xi.x - pj.x, xi.y - pj.y, xi.z - pj.z);
        if (surely_apart(nrm, ri, pj.w, k_cohesion)) return;
        const float d = lenf(nrm);
        const float err = d - ri - pj.w;
        if (err <= k_cohesion) {
            const float4 qj = G.pv[j];
            nrm = make_float3(nrm.x / d, nrm.y / d, nrm.z / d);
            const float3 vrel = make_float3(vi.x - qj.x, vi.y - qj.y, vi.z - qj.z);
            const float3 fc = particle_force(nrm, vrel, err, k_contact, k_damp, k_friction, k_mu);
            f.x += fc.x;
            f.y += fc.y;
            f.z += fc.z;
        }
    });
    st3(f_out, i, f);
}

}  // namespace nsph

using namespace nsph;


extern "C" {

int nsph_particles_integrate(nsph_ctx* ctx, int n, const float* d_x, const float* d_v, const float* d_f,
                             const float* d_inv_mass, const int* d_flags, const int* d_world,
                             const float* d_gravity, int n_gravity, float dt, float v_max, float* d_x_out,
                             float* d_v_out) {
    NSPH_CHECK(ctx && n >= 0 && n_gravity >= 1, NSPH_ERR_INVALID, "nsph_particles_integrate: bad argument");
    if (n == 0) return NSPH_OK;
    NSPH_CHECK(d_x && d_v && d_inv_mass && d_flags && d_gravity && d_x_out && d_v_out, NSPH_ERR_INVALID,
               "nsph_particles_integrate: NULL array");
    hipLaunchKernelGGL(k_integrate_particles, dim3((n + 255) / 256), dim3(256), 0, ctx->stream, n, d_x, d_v, d_f,
                       d_inv_mass, d_flags, d_world, d_gravity, n_gravity, dt, v_max, d_x_out, d_v_out);
    NSPH_HIP(hipGetLastError());
    return NSPH_OK;
}

int nsph_xpbd_particle_contacts(nsph_grid* grid, int n, const float* d_x, const float* d_v, const float* d_inv_mass,
                                const float* d_radius, const int* d_flags, float mu, float cohesion,
                                float max_radius, float dt, float relaxation, float* d_deltas) {
    GridView G;
    NSPH_TRY(nsph_grid_view(grid, &G));
    NSPH_CHECK(d_x && d_v && d_inv_mass && d_radius && d_flags && d_deltas, NSPH_ERR_INVALID,
               "nsph_xpbd_particle_contacts: NULL array");
    NSPH_CHECK(n == G.n, NSPH_ERR_INVALID,
               "nsph_xpbd_particle_contacts: the grid holds %d points but %d particles were passed", G.n, n);
    if (G.n == 0) return NSPH_OK;
    hipStream_t st;
    NSPH_TRY(nsph_grid_stream(grid, &st));
    hipLaunchKernelGGL(k_pack_particles, dim3((G.n + 255) / 256), dim3(256), 0, st, G, d_x, d_v, d_radius,
                       d_inv_mass, d_flags);
    hipLaunchKernelGGL(k_xpbd_particle_contacts, dim3((G.n + 127) / 128), dim3(128), 0, st, G, d_flags, mu,
                       cohesion, max_radius, dt, relaxation, d_deltas);
    NSPH_HIP(hipGetLastError());
    return NSPH_OK;
}

int nsph_apply_particle_deltas(nsph_ctx* ctx, int n, const float* d_x_orig, const float* d_x_pred,
                               const int* d_flags, const float* d_deltas, float dt, float v_max,
                               float* d_x_out, float* d_v_out) {
    NSPH_CHECK(ctx && n >= 0 && dt > 0.f, NSPH_ERR_INVALID, "nsph_apply_particle_deltas: bad argument");
    if (n == 0) return NSPH_OK;
    NSPH_CHECK(d_x_orig && d_x_pred && d_flags && d_deltas && d_x_out && d_v_out, NSPH_ERR_INVALID,
               "nsph_apply_particle_deltas: NULL array");
    hipLaunchKernelGGL(k_apply_particle_deltas, dim3((n + 255) / 256), dim3(256), 0, ctx->stream, n, d_x_orig,
                       d_x_pred, d_flags, d_deltas, dt, v_max, d_x_out, d_v_out);
    NSPH_HIP(hipGetLastError());
    return NSPH_OK;
}

int nsph_eval_particle_contact(nsph_grid* grid, int n, const float* d_x, const float* d_v, const float* d_radius,
                               const int* d_flags, float ke, float kd, float kf, float mu, float cohesion,
                               float max_radius, float* d_f) {
    GridView G;
    NSPH_TRY(nsph_grid_view(grid, &G));
    NSPH_CHECK(d_x && d_v && d_radius && d_flags && d_f, NSPH_ERR_INVALID, "nsph_eval_particle_contact: NULL array");
    NSPH_CHECK(n == G.n, NSPH_ERR_INVALID,
               "nsph_eval_particle_contact: the grid holds %d points but %d particles were passed", G.n, n);
    if (G.n == 0) return NSPH_OK;
    hipStream_t st;
    NSPH_TRY(nsph_grid_stream(grid, &st));
    hipLaunchKernelGGL(k_pack_particles, dim3((G.n + 255) / 256), dim3(256), 0, st, G, d_x, d_v, d_radius,
                       (const float*)nullptr, d_flags);
    hipLaunchKernelGGL(k_eval_particle_contact, dim3((G.n + 127) / 128), dim3(128), 0, st, G, d_flags, ke, kd,
                       kf, mu, cohesion, max_radius, d_f);
    NSPH_HIP(hipGetLastError());
    return NSPH_OK;
}

}  // extern "C"
