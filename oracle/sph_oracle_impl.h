/*
 * Precision-generic body of the CPU SPH oracle; included by sph_oracle.c once with
 * REAL=float (suffix _f32, the same-order FP32 restatement) and once with REAL=double
 * (suffix _f64, the FP64 oracle).  TEST INFRASTRUCTURE ONLY.
 *
 * The FP32 instance performs the same IEEE operations in the same order as the gfx950
 * kernels (no contraction, correctly rounded / and sqrt), so it must agree bit for bit.
 * The FP64 instance uses the same neighbour sets (defined by the FP32 predicate) and does
 * all arithmetic in double.
 *
 * Method (no SPH exists in the Newton reference; the conventions it does define are kept):
 *   - input generator: newton/_src/sim/builder.py:8083-8103, 9815-9843 (oracle/oracle.py)
 *   - fluid = ACTIVE and inv_mass > 0; zero-mass ACTIVE particles are static boundary
 *     (builder.py:7205-7206, solver.py:27-39); flags per newton/_src/geometry/flags.py:13
 *   - neighbour predicate style of kernels_contact.py:68-73 / xpbd/kernels.py:257-264:
 *     ACTIVE, j != i, here |x_i - x_j|^2 < H^2 in FP32
 *   - IISPH, Ihmsen et al., "Implicit Incompressible SPH", IEEE TVCG 20(3), 2014;
 *     cubic spline, Monaghan, "Smoothed particle hydrodynamics", ARAA 30, 1992;
 *     boundary pseudo-mass psi = rho0 V_b, Akinci et al., ACM TOG 31(4), 2012
 *   - integrator: solver.py:24-46 (integrate_particles)
 */

#define CAT2(a, b) a##b
#define CAT(a, b) CAT2(a, b)
#define FN(name) CAT(name, SUFFIX)

typedef struct {
    REAL H, H2, invH, sigma, sigma2, gsig2, gsig6, W0, psi, eps2, visc10, rho0;
    REAL g[3];
} FN(phys_t);

static void FN(make_phys)(FN(phys_t) * P, double H, double Vb, double rho0, double nu,
                          const float* g) {
    const double sigma = 8.0 / (M_PI * H * H * H);
    P->H = (REAL)H;
    P->H2 = (REAL)(H * H);
    P->invH = (REAL)(1.0 / H);
    P->sigma = (REAL)sigma;
    P->sigma2 = (REAL)(2.0 * sigma);
    P->gsig2 = (REAL)(sigma / (H * H));
    P->gsig6 = (REAL)(6.0 * sigma / H);
    P->W0 = (REAL)sigma;
    P->psi = (REAL)(rho0 * Vb);
    P->eps2 = (REAL)(0.01 * H * H);
    P->visc10 = (REAL)(10.0 * nu);
    P->rho0 = (REAL)rho0;
    P->g[0] = (REAL)g[0];
    P->g[1] = (REAL)g[1];
    P->g[2] = (REAL)g[2];
}

static void FN(kern_wg)(const FN(phys_t) * P, REAL r2, REAL* W, REAL* gf) {
    const REAL r = RSQRT(r2);
    const REAL q = r * P->invH;
    if (q <= (REAL)0.5) {
        REAL a = (REAL)6.0 * q;
        a = a * q;
        a = a * (q - (REAL)1.0);
        *W = P->sigma * (a + (REAL)1.0);
        *gf = P->gsig2 * ((REAL)18.0 * q - (REAL)12.0);
    } else {
        const REAL u = RMAX((REAL)1.0 - q, (REAL)0.0);
        *W = P->sigma2 * ((u * u) * u);
        *gf = -(P->gsig6 * (u * u)) / r;
    }
}

static REAL FN(kern_gf)(const FN(phys_t) * P, REAL r2) {
    REAL W, gf;
    FN(kern_wg)(P, r2, &W, &gf);
    return gf;
}

static void FN(neumaier)(REAL* s, REAL* c, REAL t) {
    const REAL u = *s + t;
    if (RABS(*s) >= RABS(t))
        *c += (*s - u) + t;
    else
        *c += (t - u) + *s;
    *s = u;
}

static REAL FN(dot3)(REAL ax, REAL ay, REAL az, REAL bx, REAL by, REAL bz) {
    return (ax * bx + ay * by) + az * bz;
}

/* Density (Neumaier-compensated, list then boundary order), a_ii and the boundary gradient
 * sum of every sorted fluid particle (the density half of k_sph_nbr_density). */
static void FN(pass_density)(const FN(phys_t) * P, const orc_layout* Lp, const REAL* xs, REAL* rho, REAL* aii,
                             REAL* gb, REAL* kvb, REAL dt2, int nthreads) {
    const orc_layout L = *Lp;
    const int Nf = L.nf;
    const int* nbr = L.nbr;
    const int* ncnt = L.ncnt;
    const int cap = L.cap;
#pragma omp parallel for num_threads(nthreads) schedule(static)
    for (int s = 0; s < Nf; ++s) {
        const FN(phys_t)* Pw = &P[L.fworld_s[s]];
        const REAL xi = xs[4 * s], yi = xs[4 * s + 1], zi = xs[4 * s + 2], mi = xs[4 * s + 3];
        REAL sum = mi * Pw->W0, cmp = 0, gfx = 0, gfy = 0, gfz = 0, S = 0;
        for (int k = 0; k < ncnt[s]; ++k) {
            const int j = nbr[(size_t)s * cap + k];
            const REAL dx = xi - xs[4 * j], dy = yi - xs[4 * j + 1], dz = zi - xs[4 * j + 2];
            const REAL r2 = (dx * dx + dy * dy) + dz * dz;
            REAL Wv, gf;
            FN(kern_wg)(Pw, r2, &Wv, &gf);
            const REAL mj = xs[4 * j + 3];
            FN(neumaier)(&sum, &cmp, mj * Wv);
            const REAL gx = gf * dx, gy = gf * dy, gz = gf * dz;
            gfx += mj * gx;
            gfy += mj * gy;
            gfz += mj * gz;
            S += mj * FN(dot3)(gx, gy, gz, gx, gy, gz);
        }
        /* boundary: static entries, then moving ones, which also add v_b . gradW_ib to kvb */
        REAL gbx = 0, gby = 0, gbz = 0, kv = 0;
        for (int k = L.bstart[s]; k < L.bstart[s + 1]; ++k) {
            const int b = L.blist[k];
            const REAL dx = xi - (REAL)L.bxs[3 * b], dy = yi - (REAL)L.bxs[3 * b + 1],
                       dz = zi - (REAL)L.bxs[3 * b + 2];
            const REAL r2 = (dx * dx + dy * dy) + dz * dz;
            REAL Wv, gf;
            FN(kern_wg)(Pw, r2, &Wv, &gf);
            FN(neumaier)(&sum, &cmp, Pw->psi * Wv);
            const REAL gx = gf * dx, gy = gf * dy, gz = gf * dz;
            gbx += gx;
            gby += gy;
            gbz += gz;
            if (b >= L.nbs)
                kv += FN(dot3)((REAL)L.bvs[3 * b], (REAL)L.bvs[3 * b + 1], (REAL)L.bvs[3 * b + 2], gx, gy, gz);
        }
        kvb[s] = kv;
        const REAL r = sum + cmp;
        const REAL Gx = gfx + Pw->psi * gbx, Gy = gfy + Pw->psi * gby, Gz = gfz + Pw->psi * gbz;
        const REAL num = FN(dot3)(Gx, Gy, Gz, Gx, Gy, Gz) + mi * S;
        rho[s] = r;
        aii[s] = -(dt2 * (num / (r * r)));
        gb[3 * s] = gbx;
        gb[3 * s + 1] = gby;
        gb[3 * s + 2] = gbz;
    }
}

/* Viscosity a_visc = 10 nu sum_j (m_j/rho_j) (v_ij . x_ij) / (r^2 + 0.01 H^2) gradW_ij and the
 * advected velocity v + dt (g + a_visc) (k_sph_advect). */
static void FN(pass_viscosity)(const FN(phys_t) * P, const orc_layout* Lp, const REAL* xs, const REAL* vs,
                               const REAL* rho, REAL* av, REAL* va, REAL rdt, int nthreads) {
    const orc_layout L = *Lp;
    const int Nf = L.nf;
    const int* nbr = L.nbr;
    const int* ncnt = L.ncnt;
    const int cap = L.cap;
#pragma omp parallel for num_threads(nthreads) schedule(static)
    for (int s = 0; s < Nf; ++s) {
        const FN(phys_t)* Pw = &P[L.fworld_s[s]];
        const REAL xi = xs[4 * s], yi = xs[4 * s + 1], zi = xs[4 * s + 2];
        const REAL vxi = vs[3 * s], vyi = vs[3 * s + 1], vzi = vs[3 * s + 2];
        REAL ax = 0, ay = 0, az = 0;
        for (int k = 0; k < ncnt[s]; ++k) {
            const int j = nbr[(size_t)s * cap + k];
            const REAL dx = xi - xs[4 * j], dy = yi - xs[4 * j + 1], dz = zi - xs[4 * j + 2];
            const REAL r2 = (dx * dx + dy * dy) + dz * dz;
            const REAL gf = FN(kern_gf)(Pw, r2);
            const REAL vx = FN(dot3)(vxi - vs[3 * j], vyi - vs[3 * j + 1], vzi - vs[3 * j + 2], dx, dy, dz);
            const REAL c = (xs[4 * j + 3] / rho[j]) * (vx / (r2 + Pw->eps2));
            ax += c * (gf * dx);
            ay += c * (gf * dy);
            az += c * (gf * dz);
        }
        const REAL avx = Pw->visc10 * ax, avy = Pw->visc10 * ay, avz = Pw->visc10 * az;
        av[3 * s] = avx;
        av[3 * s + 1] = avy;
        av[3 * s + 2] = avz;
        va[3 * s] = vxi + rdt * (Pw->g[0] + avx);
        va[3 * s + 1] = vyi + rdt * (Pw->g[1] + avy);
        va[3 * s + 2] = vzi + rdt * (Pw->g[2] + avz);
    }
}

/* One step. See orc_step_* in sph_oracle.c for the argument contract. */
static int FN(orc_step_impl)(const orc_problem* pb, const float* q_in, const float* qd_in,
                             const float* f_in, const float* xb_in, float* pprev, float dt,
                             const int* forced_iters, float* q_out, float* qd_out, float* f_out,
                             float* rho_out, float* p_out, float* xb_out, int* nbr_count_out,
                             int* iters_out, double* err_out, int nthreads) {
    const int N = pb->n, W = pb->n_worlds;
    orc_layout L;
    int rc = orc_layout_build(pb, q_in, qd_in, xb_in, &L);
    if (rc) return rc;
    const int Nf = L.nf;
    FN(phys_t)* P = (FN(phys_t)*)calloc((size_t)W, sizeof(FN(phys_t)));
    for (int w = 0; w < W; ++w)
        FN(make_phys)(&P[w], pb->support[w], pb->boundary_volume[w], pb->rest_density,
                      pb->viscosity, pb->gravity + 3 * w);
    const double dtd = (double)dt;
    const REAL rdt = (REAL)dt;
#if ORC_IS_F32
    const REAL dt2 = (float)(dtd * dtd);
#else
    const REAL dt2 = dtd * dtd;
#endif
    REAL* xs = (REAL*)malloc(sizeof(REAL) * 4 * (size_t)(Nf + 1));
    REAL* vs = (REAL*)malloc(sizeof(REAL) * 3 * (size_t)(Nf + 1));
    REAL* rho = (REAL*)malloc(sizeof(REAL) * (size_t)(Nf + 1));
    REAL* aii = (REAL*)malloc(sizeof(REAL) * (size_t)(Nf + 1));
    REAL* gb = (REAL*)malloc(sizeof(REAL) * 3 * (size_t)(Nf + 1));
    REAL* kvb = (REAL*)malloc(sizeof(REAL) * (size_t)(Nf + 1));
    REAL* av = (REAL*)malloc(sizeof(REAL) * 3 * (size_t)(Nf + 1));
    REAL* va = (REAL*)malloc(sizeof(REAL) * 3 * (size_t)(Nf + 1));
    REAL* src = (REAL*)malloc(sizeof(REAL) * (size_t)(Nf + 1));
    REAL* p = (REAL*)malloc(sizeof(REAL) * (size_t)(Nf + 1));
    REAL* pr = (REAL*)malloc(sizeof(REAL) * (size_t)(Nf + 1));
    REAL* ap = (REAL*)malloc(sizeof(REAL) * 3 * (size_t)(Nf + 1));
    double* e = (double*)malloc(sizeof(double) * (size_t)(Nf + 1));
    int* conv = (int*)calloc((size_t)W, sizeof(int));
    int* iters = (int*)calloc((size_t)W, sizeof(int));
    double* errw = (double*)calloc((size_t)W, sizeof(double));

    for (int s = 0; s < Nf; ++s) {
        const int id = L.fid[L.fperm[s]];
        xs[4 * s] = (REAL)q_in[3 * id];
        xs[4 * s + 1] = (REAL)q_in[3 * id + 1];
        xs[4 * s + 2] = (REAL)q_in[3 * id + 2];
        xs[4 * s + 3] = (REAL)pb->mass[id];
        vs[3 * s] = (REAL)qd_in[3 * id];
        vs[3 * s + 1] = (REAL)qd_in[3 * id + 1];
        vs[3 * s + 2] = (REAL)qd_in[3 * id + 2];
    }
    const int* nbr = L.nbr;
    const int* ncnt = L.ncnt;
    const int cap = L.cap;

    FN(pass_density)(P, &L, xs, rho, aii, gb, kvb, dt2, nthreads);
    FN(pass_viscosity)(P, &L, xs, vs, rho, av, va, rdt, nthreads);
    /* source term, warm start */
#pragma omp parallel for num_threads(nthreads) schedule(static)
    for (int s = 0; s < Nf; ++s) {
        const FN(phys_t)* Pw = &P[L.fworld_s[s]];
        const REAL xi = xs[4 * s], yi = xs[4 * s + 1], zi = xs[4 * s + 2];
        const REAL vxi = va[3 * s], vyi = va[3 * s + 1], vzi = va[3 * s + 2];
        REAL sum = 0;
        for (int k = 0; k < ncnt[s]; ++k) {
            const int j = nbr[(size_t)s * cap + k];
            const REAL dx = xi - xs[4 * j], dy = yi - xs[4 * j + 1], dz = zi - xs[4 * j + 2];
            const REAL r2 = (dx * dx + dy * dy) + dz * dz;
            const REAL gf = FN(kern_gf)(Pw, r2);
            sum += xs[4 * j + 3] * FN(dot3)(vxi - va[3 * j], vyi - va[3 * j + 1], vzi - va[3 * j + 2],
                                            gf * dx, gf * dy, gf * dz);
        }
        /* boundary: psi sum_b (v_i - v_b) . gradW_ib */
        sum += Pw->psi * (FN(dot3)(vxi, vyi, vzi, gb[3 * s], gb[3 * s + 1], gb[3 * s + 2]) - kvb[s]);
        src[s] = (Pw->rho0 - rho[s]) - rdt * sum;
        REAL p0 = 0;
        const REAL pp = (REAL)pprev[L.fperm[s]];
        if (pb->warm_start == 1) p0 = (REAL)0.5 * pp;
        else if (pb->warm_start == 2) p0 = pp;
        p[s] = p0;
        pr[s] = p0 / (rho[s] * rho[s]);
    }
    for (int w = 0; w < W; ++w) conv[w] = (L.w_nf[w] == 0);
    /* relaxed Jacobi */
    for (int it = 0; it < pb->max_iters; ++it) {
        int active = 0;
        for (int w = 0; w < W; ++w) active += (conv[w] == 0);
        if (!active) break;
#pragma omp parallel for num_threads(nthreads) schedule(static)
        for (int s = 0; s < Nf; ++s) {
            const int w = L.fworld_s[s];
            if (conv[w]) continue;
            const FN(phys_t)* Pw = &P[w];
            const REAL xi = xs[4 * s], yi = xs[4 * s + 1], zi = xs[4 * s + 2];
            const REAL pri = pr[s];
            REAL ax = 0, ay = 0, az = 0;
            for (int k = 0; k < ncnt[s]; ++k) {
                const int j = nbr[(size_t)s * cap + k];
                const REAL dx = xi - xs[4 * j], dy = yi - xs[4 * j + 1], dz = zi - xs[4 * j + 2];
                const REAL r2 = (dx * dx + dy * dy) + dz * dz;
                const REAL gf = FN(kern_gf)(Pw, r2);
                const REAL c = xs[4 * j + 3] * (pri + pr[j]);
                ax += c * (gf * dx);
                ay += c * (gf * dy);
                az += c * (gf * dz);
            }
            const REAL cb = Pw->psi * pri;
            ap[3 * s] = -(ax + cb * gb[3 * s]);
            ap[3 * s + 1] = -(ay + cb * gb[3 * s + 1]);
            ap[3 * s + 2] = -(az + cb * gb[3 * s + 2]);
        }
#pragma omp parallel for num_threads(nthreads) schedule(static)
        for (int s = 0; s < Nf; ++s) {
            const int w = L.fworld_s[s];
            if (conv[w]) continue;
            const FN(phys_t)* Pw = &P[w];
            const REAL xi = xs[4 * s], yi = xs[4 * s + 1], zi = xs[4 * s + 2];
            const REAL axi = ap[3 * s], ayi = ap[3 * s + 1], azi = ap[3 * s + 2];
            REAL sum = 0;
            for (int k = 0; k < ncnt[s]; ++k) {
                const int j = nbr[(size_t)s * cap + k];
                const REAL dx = xi - xs[4 * j], dy = yi - xs[4 * j + 1], dz = zi - xs[4 * j + 2];
                const REAL r2 = (dx * dx + dy * dy) + dz * dz;
                const REAL gf = FN(kern_gf)(Pw, r2);
                sum += xs[4 * j + 3] * FN(dot3)(axi - ap[3 * j], ayi - ap[3 * j + 1], azi - ap[3 * j + 2],
                                                gf * dx, gf * dy, gf * dz);
            }
            sum += Pw->psi * FN(dot3)(axi, ayi, azi, gb[3 * s], gb[3 * s + 1], gb[3 * s + 2]);
            const REAL Ap = dt2 * sum;
            const REAL si = src[s];
            e[s] = (double)RMAX(Ap - si, (REAL)0);
            REAL pn = 0;
            if (aii[s] != (REAL)0) pn = RMAX(p[s] + (REAL)pb->omega * (si - Ap) / aii[s], (REAL)0);
            p[s] = pn;
            pr[s] = pn / (rho[s] * rho[s]);
        }
        for (int w = 0; w < W; ++w) {
            if (conv[w]) continue;
            const double err = orc_world_reduce(e, L.w_fstart[w], L.w_nf[w]) /
                               ((double)L.w_nf[w] * (double)P[w].rho0);
            const int n_it = ++iters[w];
            errw[w] = err;
            int c = 0;
            if (forced_iters && forced_iters[w] >= 0) {
                if (n_it >= forced_iters[w]) c = 1;
            } else if (n_it >= pb->min_iters && err <= pb->tolerance) {
                c = 1;
            } else if (n_it >= pb->max_iters) {
                c = 2;
            }
            conv[w] = c;
        }
    }
    /* final pressure acceleration, force, reference integrator, scatter */
#pragma omp parallel for num_threads(nthreads) schedule(static)
    for (int s = 0; s < Nf; ++s) {
        const FN(phys_t)* Pw = &P[L.fworld_s[s]];
        const REAL xi = xs[4 * s], yi = xs[4 * s + 1], zi = xs[4 * s + 2], m = xs[4 * s + 3];
        const REAL pri = pr[s];
        REAL ax = 0, ay = 0, az = 0;
        for (int k = 0; k < ncnt[s]; ++k) {
            const int j = nbr[(size_t)s * cap + k];
            const REAL dx = xi - xs[4 * j], dy = yi - xs[4 * j + 1], dz = zi - xs[4 * j + 2];
            const REAL r2 = (dx * dx + dy * dy) + dz * dz;
            const REAL gf = FN(kern_gf)(Pw, r2);
            const REAL c = xs[4 * j + 3] * (pri + pr[j]);
            ax += c * (gf * dx);
            ay += c * (gf * dy);
            az += c * (gf * dz);
        }
        const REAL cb = Pw->psi * pri;
        ax = -(ax + cb * gb[3 * s]);
        ay = -(ay + cb * gb[3 * s + 1]);
        az = -(az + cb * gb[3 * s + 2]);
        const REAL fx = m * (av[3 * s] + ax), fy = m * (av[3 * s + 1] + ay), fz = m * (av[3 * s + 2] + az);
        const int k = L.fperm[s];
        const int id = L.fid[k];
        const REAL wi = (REAL)pb->inv_mass[id];
        /* total force: the caller's force on the particle (state_in.particle_f) plus SPH */
        REAL tx = fx, ty = fy, tz = fz;
        if (f_in) {
            tx = (REAL)f_in[3 * id] + fx;
            ty = (REAL)f_in[3 * id + 1] + fy;
            tz = (REAL)f_in[3 * id + 2] + fz;
        }
        REAL v1x = vs[3 * s] + (tx * wi + Pw->g[0]) * rdt;
        REAL v1y = vs[3 * s + 1] + (ty * wi + Pw->g[1]) * rdt;
        REAL v1z = vs[3 * s + 2] + (tz * wi + Pw->g[2]) * rdt;
        const REAL mag = RSQRT(FN(dot3)(v1x, v1y, v1z, v1x, v1y, v1z));
        const REAL vmax = (REAL)pb->v_max;
        if (mag > vmax) {
            const REAL sc = vmax / mag;
            v1x *= sc;
            v1y *= sc;
            v1z *= sc;
        }
        q_out[3 * id] = (float)(xi + v1x * rdt);
        q_out[3 * id + 1] = (float)(yi + v1y * rdt);
        q_out[3 * id + 2] = (float)(zi + v1z * rdt);
        qd_out[3 * id] = (float)v1x;
        qd_out[3 * id + 1] = (float)v1y;
        qd_out[3 * id + 2] = (float)v1z;
        if (f_out) {
            f_out[3 * id] = (float)fx;
            f_out[3 * id + 1] = (float)fy;
            f_out[3 * id + 2] = (float)fz;
        }
        if (rho_out) rho_out[id] = (float)rho[s];
        if (p_out) p_out[id] = (float)p[s];
        if (nbr_count_out) nbr_count_out[id] = ncnt[s];
        if (xb_out) {
            xb_out[3 * id] = L.xbk[3 * k];
            xb_out[3 * id + 1] = L.xbk[3 * k + 1];
            xb_out[3 * id + 2] = L.xbk[3 * k + 2];
        }
    }
    for (int s = 0; s < Nf; ++s) pprev[L.fperm[s]] = (float)p[s];
    /* non-fluid particles: reference integrator with zero force */
    for (int i = 0; i < N; ++i) {
        if (L.is_fluid[i]) continue;
        if (f_out) f_out[3 * i] = f_out[3 * i + 1] = f_out[3 * i + 2] = 0.f;
        if (rho_out) rho_out[i] = 0.f;
        if (p_out) p_out[i] = 0.f;
        if (nbr_count_out) nbr_count_out[i] = 0;
        if (xb_out) {  /* not used by the step: the particle's position */
            xb_out[3 * i] = q_in[3 * i];
            xb_out[3 * i + 1] = q_in[3 * i + 1];
            xb_out[3 * i + 2] = q_in[3 * i + 2];
        }
        if ((pb->flags[i] & 1) == 0) {
            q_out[3 * i] = q_in[3 * i];
            q_out[3 * i + 1] = q_in[3 * i + 1];
            q_out[3 * i + 2] = q_in[3 * i + 2];
            continue;
        }
        int wd = pb->world ? pb->world[i] : 0;
        wd = wd < 0 ? 0 : (wd >= W ? W - 1 : wd);
        const float wi = pb->inv_mass[i];
        const float stp = (-wi < 0.0f) ? 1.0f : 0.0f;
        float v1[3];
        for (int a = 0; a < 3; ++a) {
            const float fa = f_in ? f_in[3 * i + a] : 0.0f;
            v1[a] = qd_in[3 * i + a] + (fa * wi + pb->gravity[3 * wd + a] * stp) * dt;
        }
        const float mag = sqrtf((v1[0] * v1[0] + v1[1] * v1[1]) + v1[2] * v1[2]);
        if (mag > pb->v_max) {
            const float sc = pb->v_max / mag;
            for (int a = 0; a < 3; ++a) v1[a] *= sc;
        }
        for (int a = 0; a < 3; ++a) {
            q_out[3 * i + a] = q_in[3 * i + a] + v1[a] * dt;
            qd_out[3 * i + a] = v1[a];
        }
    }
    for (int w = 0; w < W; ++w) {
        if (iters_out) iters_out[w] = iters[w];
        if (err_out) err_out[w] = errw[w];
    }
    int status = 0;
    for (int w = 0; w < W; ++w)
        if (L.w_nf[w] > 0 && conv[w] != 1) status = 6;
    if (L.overflow) status = 3;
    free(P); free(xs); free(vs); free(rho); free(aii); free(gb); free(kvb); free(av); free(va); free(src);
    free(p); free(pr); free(ap); free(e); free(conv); free(iters); free(errw);
    orc_layout_free(&L);
    return status;
}

/* The pieces of one step before the Jacobi loop, and the IISPH operator applied to a given
 * pressure field p_in (fluid-list order). Outputs in fluid-list order:
 *   rho, a_ii, s (source), Ap = (A p_in), ap = a^p(p_in) (3 per particle).
 * Test hook: lets tests check A against its diagonal, its linearity and the density sum. */
static int FN(orc_operator_impl)(const orc_problem* pb, const float* q_in, const float* qd_in,
                                 float dt, const double* p_in, double* rho_o, double* aii_o,
                                 double* s_o, double* Ap_o, double* ap_o) {
    const int W = pb->n_worlds;
    orc_layout L;
    int rc = orc_layout_build(pb, q_in, qd_in, NULL, &L);
    if (rc) return rc;
    const int Nf = L.nf;
    FN(phys_t)* P = (FN(phys_t)*)calloc((size_t)W, sizeof(FN(phys_t)));
    for (int w = 0; w < W; ++w)
        FN(make_phys)(&P[w], pb->support[w], pb->boundary_volume[w], pb->rest_density,
                      pb->viscosity, pb->gravity + 3 * w);
    const REAL rdt = (REAL)dt;
    const REAL dt2 = (REAL)((double)dt * (double)dt);
    REAL* xs = (REAL*)malloc(sizeof(REAL) * 4 * (size_t)(Nf + 1));
    REAL* rho = (REAL*)malloc(sizeof(REAL) * (size_t)(Nf + 1));
    REAL* gb = (REAL*)malloc(sizeof(REAL) * 3 * (size_t)(Nf + 1));
    REAL* kvb = (REAL*)malloc(sizeof(REAL) * (size_t)(Nf + 1));
    REAL* pr = (REAL*)malloc(sizeof(REAL) * (size_t)(Nf + 1));
    REAL* ap = (REAL*)malloc(sizeof(REAL) * 3 * (size_t)(Nf + 1));
    REAL* va = (REAL*)malloc(sizeof(REAL) * 3 * (size_t)(Nf + 1));
    for (int s = 0; s < Nf; ++s) {
        const int id = L.fid[L.fperm[s]];
        for (int a = 0; a < 3; ++a) xs[4 * s + a] = (REAL)q_in[3 * id + a];
        xs[4 * s + 3] = (REAL)pb->mass[id];
    }
    const int cap = L.cap;
    for (int s = 0; s < Nf; ++s) {
        const FN(phys_t)* Pw = &P[L.fworld_s[s]];
        REAL sum = xs[4 * s + 3] * Pw->W0, cmp = 0, gfx = 0, gfy = 0, gfz = 0, S = 0;
        for (int k = 0; k < L.ncnt[s]; ++k) {
            const int j = L.nbr[(size_t)s * cap + k];
            const REAL dx = xs[4 * s] - xs[4 * j], dy = xs[4 * s + 1] - xs[4 * j + 1], dz = xs[4 * s + 2] - xs[4 * j + 2];
            REAL Wv, gf;
            FN(kern_wg)(Pw, (dx * dx + dy * dy) + dz * dz, &Wv, &gf);
            FN(neumaier)(&sum, &cmp, xs[4 * j + 3] * Wv);
            gfx += xs[4 * j + 3] * gf * dx;
            gfy += xs[4 * j + 3] * gf * dy;
            gfz += xs[4 * j + 3] * gf * dz;
            S += xs[4 * j + 3] * (gf * gf) * ((dx * dx + dy * dy) + dz * dz);
        }
        REAL gbx = 0, gby = 0, gbz = 0, kv = 0;
        for (int k = L.bstart[s]; k < L.bstart[s + 1]; ++k) {
            const int b = L.blist[k];
            const REAL dx = xs[4 * s] - (REAL)L.bxs[3 * b], dy = xs[4 * s + 1] - (REAL)L.bxs[3 * b + 1],
                       dz = xs[4 * s + 2] - (REAL)L.bxs[3 * b + 2];
            REAL Wv, gf;
            FN(kern_wg)(Pw, (dx * dx + dy * dy) + dz * dz, &Wv, &gf);
            FN(neumaier)(&sum, &cmp, Pw->psi * Wv);
            gbx += gf * dx;
            gby += gf * dy;
            gbz += gf * dz;
            if (b >= L.nbs)
                kv += FN(dot3)((REAL)L.bvs[3 * b], (REAL)L.bvs[3 * b + 1], (REAL)L.bvs[3 * b + 2], gf * dx, gf * dy,
                               gf * dz);
        }
        kvb[s] = kv;
        rho[s] = sum + cmp;
        gb[3 * s] = gbx;
        gb[3 * s + 1] = gby;
        gb[3 * s + 2] = gbz;
        const REAL Gx = gfx + Pw->psi * gbx, Gy = gfy + Pw->psi * gby, Gz = gfz + Pw->psi * gbz;
        const int k = L.fperm[s];
        rho_o[k] = (double)rho[s];
        aii_o[k] = (double)(-(dt2 * ((FN(dot3)(Gx, Gy, Gz, Gx, Gy, Gz) + xs[4 * s + 3] * S) / (rho[s] * rho[s]))));
        /* advected velocity without viscosity (the tests use nu = 0 here) */
        const int id = L.fid[k];
        for (int a = 0; a < 3; ++a) va[3 * s + a] = (REAL)qd_in[3 * id + a] + rdt * Pw->g[a];
        pr[s] = (REAL)p_in[k] / (rho[s] * rho[s]);
    }
    for (int s = 0; s < Nf; ++s) {
        const FN(phys_t)* Pw = &P[L.fworld_s[s]];
        REAL sum = 0, ax = 0, ay = 0, az = 0;
        for (int k = 0; k < L.ncnt[s]; ++k) {
            const int j = L.nbr[(size_t)s * cap + k];
            const REAL dx = xs[4 * s] - xs[4 * j], dy = xs[4 * s + 1] - xs[4 * j + 1], dz = xs[4 * s + 2] - xs[4 * j + 2];
            const REAL gf = FN(kern_gf)(Pw, (dx * dx + dy * dy) + dz * dz);
            sum += xs[4 * j + 3] * FN(dot3)(va[3 * s] - va[3 * j], va[3 * s + 1] - va[3 * j + 1],
                                            va[3 * s + 2] - va[3 * j + 2], gf * dx, gf * dy, gf * dz);
            const REAL c = xs[4 * j + 3] * (pr[s] + pr[j]);
            ax += c * (gf * dx);
            ay += c * (gf * dy);
            az += c * (gf * dz);
        }
        sum += Pw->psi * (FN(dot3)(va[3 * s], va[3 * s + 1], va[3 * s + 2], gb[3 * s], gb[3 * s + 1], gb[3 * s + 2]) -
                          kvb[s]);
        const REAL cb = Pw->psi * pr[s];
        ap[3 * s] = -(ax + cb * gb[3 * s]);
        ap[3 * s + 1] = -(ay + cb * gb[3 * s + 1]);
        ap[3 * s + 2] = -(az + cb * gb[3 * s + 2]);
        s_o[L.fperm[s]] = (double)((Pw->rho0 - rho[s]) - rdt * sum);
    }
    for (int s = 0; s < Nf; ++s) {
        const FN(phys_t)* Pw = &P[L.fworld_s[s]];
        REAL sum = 0;
        for (int k = 0; k < L.ncnt[s]; ++k) {
            const int j = L.nbr[(size_t)s * cap + k];
            const REAL dx = xs[4 * s] - xs[4 * j], dy = xs[4 * s + 1] - xs[4 * j + 1], dz = xs[4 * s + 2] - xs[4 * j + 2];
            const REAL gf = FN(kern_gf)(Pw, (dx * dx + dy * dy) + dz * dz);
            sum += xs[4 * j + 3] * FN(dot3)(ap[3 * s] - ap[3 * j], ap[3 * s + 1] - ap[3 * j + 1],
                                            ap[3 * s + 2] - ap[3 * j + 2], gf * dx, gf * dy, gf * dz);
        }
        sum += Pw->psi * FN(dot3)(ap[3 * s], ap[3 * s + 1], ap[3 * s + 2], gb[3 * s], gb[3 * s + 1], gb[3 * s + 2]);
        const int k = L.fperm[s];
        Ap_o[k] = (double)(dt2 * sum);
        for (int a = 0; a < 3; ++a) ap_o[3 * k + a] = (double)ap[3 * s + a];
    }
    free(P); free(xs); free(rho); free(gb); free(kvb); free(pr); free(ap); free(va);
    orc_layout_free(&L);
    return 0;
}

/* Density and viscosity of one state (test hook): a_visc per particle (particle order, 3 per
 * particle, 0 for non-fluid), computed by the same passes as the step. */
static int FN(orc_viscosity_impl)(const orc_problem* pb, const float* q_in, const float* qd_in, double* av_o,
                                  double* rho_o) {
    const int W = pb->n_worlds;
    orc_layout L;
    int rc = orc_layout_build(pb, q_in, qd_in, NULL, &L);
    if (rc) return rc;
    const int Nf = L.nf;
    FN(phys_t)* P = (FN(phys_t)*)calloc((size_t)W, sizeof(FN(phys_t)));
    for (int w = 0; w < W; ++w)
        FN(make_phys)(&P[w], pb->support[w], pb->boundary_volume[w], pb->rest_density,
                      pb->viscosity, pb->gravity + 3 * w);
    REAL* xs = (REAL*)malloc(sizeof(REAL) * 4 * (size_t)(Nf + 1));
    REAL* vs = (REAL*)malloc(sizeof(REAL) * 3 * (size_t)(Nf + 1));
    REAL* rho = (REAL*)malloc(sizeof(REAL) * (size_t)(Nf + 1));
    REAL* aii = (REAL*)malloc(sizeof(REAL) * (size_t)(Nf + 1));
    REAL* gb = (REAL*)malloc(sizeof(REAL) * 3 * (size_t)(Nf + 1));
    REAL* av = (REAL*)malloc(sizeof(REAL) * 3 * (size_t)(Nf + 1));
    REAL* va = (REAL*)malloc(sizeof(REAL) * 3 * (size_t)(Nf + 1));
    for (int s = 0; s < Nf; ++s) {
        const int id = L.fid[L.fperm[s]];
        for (int a = 0; a < 3; ++a) {
            xs[4 * s + a] = (REAL)q_in[3 * id + a];
            vs[3 * s + a] = (REAL)qd_in[3 * id + a];
        }
        xs[4 * s + 3] = (REAL)pb->mass[id];
    }
    REAL* kvb = (REAL*)malloc(sizeof(REAL) * (size_t)(Nf + 1));
    FN(pass_density)(P, &L, xs, rho, aii, gb, kvb, (REAL)0, 1);
    FN(pass_viscosity)(P, &L, xs, vs, rho, av, va, (REAL)0, 1);
    for (int i = 0; i < 3 * pb->n; ++i) av_o[i] = 0.0;
    for (int i = 0; i < pb->n; ++i) rho_o[i] = 0.0;
    for (int s = 0; s < Nf; ++s) {
        const int id = L.fid[L.fperm[s]];
        for (int a = 0; a < 3; ++a) av_o[3 * id + a] = (double)av[3 * s + a];
        rho_o[id] = (double)rho[s];
    }
    free(P); free(xs); free(vs); free(rho); free(aii); free(gb); free(kvb); free(av); free(va);
    orc_layout_free(&L);
    return 0;
}

#undef FN
#undef CAT
#undef CAT2
