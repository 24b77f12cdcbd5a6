/* Scalar C restatement of the jax_tinyfluids step -- TEST INFRASTRUCTURE ONLY.
 *
 * Included twice by tf_oracle.c, once with REAL=double (the oracle) and once
 * with REAL=float (the reference's own precision). Every expression keeps the
 * reference's evaluation order so that, compiled with -ffp-contract=off, the
 * float64 build is bit-equal to the NumPy restatement (oracle/np_tinyfluids.py).
 *
 * Reference (relative to the reference repository root):
 *   fluid.py            = tinyfluids/jax_tinyfluids/fluid.py
 *   time_integration.py = tinyfluids/jax_tinyfluids/time_integration.py
 *
 * Layout: state[5][nx][ny][nz], C order, primitive variables rho, vx, vy, vz, p
 * (fluid.py:10-22). Periodic neighbours as jnp.roll (fluid.py:528, 548).
 */

#define CAT_(a, b) a##b
#define CAT(a, b) CAT_(a, b)
#define FN(name) CAT(CAT(tfo_, SUFFIX), CAT(_, name))

/* np.maximum / np.minimum: NaN in either operand gives NaN */
static inline REAL FN(nmax)(REAL a, REAL b) { return (a != a || b != b) ? (a + b) : (a > b ? a : b); }
static inline REAL FN(nmin)(REAL a, REAL b) { return (a != a || b != b) ? (a + b) : (a < b ? a : b); }

typedef struct { REAL v[5]; } FN(vec5);

/* fluid.py:28-67 conserved_state_from_primitive, one cell */
static inline void FN(p2c)(const REAL *w, REAL g, REAL *u) {
    REAL rho = w[0], p = w[4];
    REAL ut = SQRT(w[1] * w[1] + w[2] * w[2] + w[3] * w[3]);
    REAL E = p / (g - (REAL)1) + (REAL)0.5 * rho * (ut * ut);
    u[0] = rho;
    u[1] = rho * w[1];
    u[2] = rho * w[2];
    u[3] = rho * w[3];
    u[4] = E;
}

/* fluid.py:69-103 primitive_state_from_conserved, one cell */
static inline void FN(c2p)(const REAL *u, REAL g, REAL *w) {
    REAL rho = u[0], E = u[4];
    REAL ux = u[1] / rho, uy = u[2] / rho, uz = u[3] / rho;
    REAL uu = SQRT(ux * ux + uy * uy + uz * uz);
    REAL p = (g - (REAL)1) * rho * (E / rho - (REAL)0.5 * (uu * uu));
    w[0] = rho;
    w[1] = ux;
    w[2] = uy;
    w[3] = uz;
    w[4] = p;
}

/* fluid.py:105-117 */
static inline REAL FN(sound)(REAL rho, REAL p, REAL g) { return SQRT(g * p / rho); }

/* fluid.py:119-160 get_wave_speeds (d = velocity index = axis) */
static inline void FN(wave_speeds)(const REAL *wl, const REAL *wr, REAL g, int d, REAL *sl, REAL *sr) {
    REAL cl = FN(sound)(wl[0], wl[4], g), cr = FN(sound)(wr[0], wr[4], g);
    *sr = FN(nmax)(FN(nmax)(wl[d] + cl, wr[d] + cr), (REAL)0);
    *sl = FN(nmin)(FN(nmin)(wl[d] - cl, wr[d] - cr), (REAL)0);
}

/* fluid.py:206-259 _euler_flux, one cell */
static inline void FN(euler_flux)(const REAL *w, REAL g, int d, REAL *f) {
    REAL rho = w[0], p = w[4];
    REAL ut = SQRT(w[1] * w[1] + w[2] * w[2] + w[3] * w[3]);
    REAL E = p / (g - (REAL)1) + (REAL)0.5 * rho * (ut * ut);
    REAL fv[5] = {rho, w[1] * rho, w[2] * rho, w[3] * rho, p + E};
    for (int k = 0; k < 5; ++k) f[k] = w[d] * fv[k];
    f[d] = f[d] + p;
}

/* fluid.py:261-305 _hll_solver, one interface */
static inline void FN(hll)(const REAL *wl, const REAL *wr, REAL g, int d, REAL *f) {
    REAL fl[5], fr[5], ul[5], ur[5], sl, sr;
    FN(euler_flux)(wl, g, d, fl);
    FN(euler_flux)(wr, g, d, fr);
    FN(wave_speeds)(wl, wr, g, d, &sl, &sr);
    FN(p2c)(wl, g, ul);
    FN(p2c)(wr, g, ur);
    for (int k = 0; k < 5; ++k)
        f[k] = (sr * fl[k] - sl * fr[k] + sl * sr * (ur[k] - ul[k])) / (sr - sl);
}

static inline void FN(load)(const REAL *s, size_t nc, size_t idx, REAL *w) {
    for (int k = 0; k < 5; ++k) w[k] = s[k * nc + idx];
}

/* fluid.py:506-557 _evolve_state_along_axis: in -> out, axis in {1,2,3}.
 * Each interface flux is a pure function of its two cells, so computing it in
 * any order gives the reference's values exactly. */
void FN(sweep)(const REAL *in, REAL *out, int nx, int ny, int nz, REAL dx, REAL dt, REAL g, int axis) {
    const size_t nc = (size_t)nx * ny * nz;
    const int n[3] = {nx, ny, nz};
    const size_t stride[3] = {(size_t)ny * nz, (size_t)nz, 1};
    const int a = axis - 1;
    const int o1 = a == 0 ? 1 : 0, o2 = a == 2 ? 1 : 2; /* the two other axes */
    const REAL inv_dx = (REAL)1 / dx;
    const long nlines = (long)n[o1] * n[o2];
#pragma omp parallel
    {
        REAL fprev[5], fnext[5], fwrap[5], wl[5], wc[5], u[5], w[5];
#pragma omp for schedule(static)
        for (long line = 0; line < nlines; ++line) {
            const int i1 = (int)(line / n[o2]), i2 = (int)(line % n[o2]);
            const size_t base = (size_t)i1 * stride[o1] + (size_t)i2 * stride[o2];
            const int N = n[a];
            const size_t s = stride[a];
            /* interface 0 sits between cell N-1 and cell 0 */
            FN(load)(in, nc, base + (size_t)(N - 1) * s, wl);
            FN(load)(in, nc, base, wc);
            FN(hll)(wl, wc, g, axis, fprev);
            for (int k = 0; k < 5; ++k) fwrap[k] = fprev[k];
            for (int i = 0; i < N; ++i) {
                /* interface i+1 sits between cell i and cell i+1 (periodic) */
                if (i + 1 < N) {
                    REAL wn[5];
                    FN(load)(in, nc, base + (size_t)(i + 1) * s, wn);
                    FN(hll)(wc, wn, g, axis, fnext);
                    FN(p2c)(wc, g, u);
                    for (int k = 0; k < 5; ++k) u[k] = u[k] + inv_dx * (fprev[k] - fnext[k]) * dt;
                    FN(c2p)(u, g, w);
                    for (int k = 0; k < 5; ++k) out[k * nc + base + (size_t)i * s] = w[k];
                    for (int k = 0; k < 5; ++k) { fprev[k] = fnext[k]; wc[k] = wn[k]; }
                } else {
                    FN(p2c)(wc, g, u);
                    for (int k = 0; k < 5; ++k) u[k] = u[k] + inv_dx * (fprev[k] - fwrap[k]) * dt;
                    FN(c2p)(u, g, w);
                    for (int k = 0; k < 5; ++k) out[k * nc + base + (size_t)i * s] = w[k];
                }
            }
        }
    }
}

/* fluid.py:382-413 _open_boundaries with one ghost cell: the sequential
 * left/right copies along x, then y, then z. */
void FN(open_boundaries)(REAL *s, int nx, int ny, int nz) {
    const size_t nc = (size_t)nx * ny * nz;
    for (int k = 0; k < 5; ++k) {
        REAL *f = s + k * nc;
        for (int y = 0; y < ny; ++y)
            for (int z = 0; z < nz; ++z) {
                f[((size_t)0 * ny + y) * nz + z] = f[((size_t)1 * ny + y) * nz + z];
                f[((size_t)(nx - 1) * ny + y) * nz + z] = f[((size_t)(nx - 2) * ny + y) * nz + z];
            }
        for (int x = 0; x < nx; ++x)
            for (int z = 0; z < nz; ++z) {
                f[((size_t)x * ny + 0) * nz + z] = f[((size_t)x * ny + 1) * nz + z];
                f[((size_t)x * ny + (ny - 1)) * nz + z] = f[((size_t)x * ny + (ny - 2)) * nz + z];
            }
        for (int x = 0; x < nx; ++x)
            for (int y = 0; y < ny; ++y) {
                f[((size_t)x * ny + y) * nz + 0] = f[((size_t)x * ny + y) * nz + 1];
                f[((size_t)x * ny + y) * nz + (nz - 1)] = f[((size_t)x * ny + y) * nz + (nz - 2)];
            }
    }
}

/* fluid.py:425-504 _cfl_time_step: for each axis, the max over periodic
 * interfaces of max(|S_R|, |S_L|); dt = C_CFL * dx / max. */
REAL FN(cfl_time_step)(const REAL *s, int nx, int ny, int nz, REAL dx, REAL g, REAL ccfl) {
    const size_t nc = (size_t)nx * ny * nz;
    const int n[3] = {nx, ny, nz};
    const size_t stride[3] = {(size_t)ny * nz, (size_t)nz, 1};
    REAL axis_max[3];
    for (int a = 0; a < 3; ++a) {
        REAL mr = -(REAL)INFINITY, ml = -(REAL)INFINITY;
        int nan_seen = 0;
#pragma omp parallel for reduction(max : mr, ml) reduction(| : nan_seen) schedule(static)
        for (long x = 0; x < nx; ++x) {
            REAL wl[5], wr[5], sl, sr;
            for (int y = 0; y < ny; ++y)
                for (int z = 0; z < nz; ++z) {
                    int idx[3] = {(int)x, y, z};
                    size_t here = (size_t)x * stride[0] + (size_t)y * stride[1] + (size_t)z;
                    int il = idx[a] == 0 ? n[a] - 1 : idx[a] - 1;
                    size_t left = here - (size_t)idx[a] * stride[a] + (size_t)il * stride[a];
                    FN(load)(s, nc, left, wl);
                    FN(load)(s, nc, here, wr);
                    FN(wave_speeds)(wl, wr, g, a + 1, &sl, &sr);
                    REAL ar = sr < 0 ? -sr : sr, al = sl < 0 ? -sl : sl;
                    if (ar != ar || al != al) nan_seen = 1;
                    if (ar > mr) mr = ar;
                    if (al > ml) ml = al;
                }
        }
        axis_max[a] = nan_seen ? (REAL)NAN : FN(nmax)(mr, ml);
    }
    REAL m = FN(nmax)(FN(nmax)(axis_max[0], axis_max[1]), axis_max[2]);
    return ccfl * dx / m;
}

/* fluid.py:560-578 _evolve_state: [open boundaries] then X, Y, Z sweeps.
 * s is updated in place; tmp holds one state. */
void FN(evolve_state)(REAL *s, REAL *tmp, int nx, int ny, int nz, REAL dx, REAL dt, REAL g, int handle_boundaries) {
    const size_t n = (size_t)5 * nx * ny * nz;
    if (handle_boundaries) FN(open_boundaries)(s, nx, ny, nz);
    FN(sweep)(s, tmp, nx, ny, nz, dx, dt, g, 1);
    FN(sweep)(tmp, s, nx, ny, nz, dx, dt, g, 2);
    FN(sweep)(s, tmp, nx, ny, nz, dx, dt, g, 3);
    memcpy(s, tmp, n * sizeof(REAL));
}

/* time_integration.py:63-92 with a fixed step count when t_final is +inf;
 * returns the number of iterations, *t_out the accumulated time (REAL). */
long FN(time_integration)(REAL *s, REAL *tmp, int nx, int ny, int nz, REAL dx, REAL t_final, REAL g,
                          int handle_boundaries, long max_steps, REAL ccfl, REAL *t_out, REAL *dts) {
    REAL t = (REAL)0;
    long it = 0;
    while (t < t_final && (max_steps < 0 || it < max_steps)) {
        REAL dt = FN(cfl_time_step)(s, nx, ny, nz, dx, g, ccfl);
        FN(evolve_state)(s, tmp, nx, ny, nz, dx, dt, g, handle_boundaries);
        if (dts) dts[it] = dt;
        t = t + dt;
        ++it;
    }
    if (t_out) *t_out = t;
    return it;
}

void FN(prim_to_cons)(const REAL *in, REAL *out, size_t ncells, REAL g) {
#pragma omp parallel for schedule(static)
    for (long i = 0; i < (long)ncells; ++i) {
        REAL w[5], u[5];
        FN(load)(in, ncells, (size_t)i, w);
        FN(p2c)(w, g, u);
        for (int k = 0; k < 5; ++k) out[k * ncells + (size_t)i] = u[k];
    }
}

void FN(cons_to_prim)(const REAL *in, REAL *out, size_t ncells, REAL g) {
#pragma omp parallel for schedule(static)
    for (long i = 0; i < (long)ncells; ++i) {
        REAL u[5], w[5];
        FN(load)(in, ncells, (size_t)i, u);
        FN(c2p)(u, g, w);
        for (int k = 0; k < 5; ++k) out[k * ncells + (size_t)i] = w[k];
    }
}

/* Totals of the conserved variables, accumulated in double per x-plane and
 * then over planes (fixed order, so the result does not depend on threads). */
void FN(conserved_totals)(const REAL *s, int nx, int ny, int nz, double g, double *out5) {
    const size_t nc = (size_t)nx * ny * nz;
    double *part = (double *)calloc((size_t)nx * 5, sizeof(double));
#pragma omp parallel for schedule(static)
    for (long x = 0; x < nx; ++x) {
        double acc[5] = {0, 0, 0, 0, 0};
        for (size_t i = (size_t)x * ny * nz; i < (size_t)(x + 1) * ny * nz; ++i) {
            double w[5], u[5];
            for (int k = 0; k < 5; ++k) w[k] = (double)s[k * nc + i];
            tfo_f64_p2c(w, g, u);
            for (int k = 0; k < 5; ++k) acc[k] += u[k];
        }
        for (int k = 0; k < 5; ++k) part[x * 5 + k] = acc[k];
    }
    for (int k = 0; k < 5; ++k) {
        double a = 0;
        for (long x = 0; x < nx; ++x) a += part[x * 5 + k];
        out5[k] = a;
    }
    free(part);
}

#undef FN
