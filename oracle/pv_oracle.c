/*
 * pv_oracle.c -- CPU FP64 restatement of PointVortex's velocity backend.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP backend: only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.  The
 * product library (libpvhip.so) never links or calls it, and has no CPU
 * fallback through it.
 *
 * Each function restates the reference algorithm with the same formulas, the
 * same loop order and the same OpenMP pragmas, so that the results agree with
 * the reference CPU build to the last bit or within a few ulps.  Citations are
 * to the reference tree (src/...).  Pinned by tests/test_oracle.py against the
 * reference's closed-form known-answer tests, the golden vectors under
 * tests/golden/ (written by the reference's own CPU code) and invariants.
 *
 * Error handling mirrors the reference's exceptions: every entry point returns
 * PVO_OK or the kind of exception the reference would throw, and the message
 * text is available from pvo_last_error().
 */
#include <math.h>
#include <stddef.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define PVO_OK 0
#define PVO_INVALID_ARGUMENT 1
#define PVO_OUT_OF_RANGE 2
#define PVO_RUNTIME 3
#define PVO_LOGIC 4

#define PVO_INFINITE 0
#define PVO_PERIODIC 1
#define PVO_DISK 2

static const double kPi = 3.14159265358979323846264338327950288;

static _Thread_local char g_error[256];

static int fail(int kind, const char *message) {
    snprintf(g_error, sizeof g_error, "%s", message);
    return kind;
}

const char *pvo_last_error(void) { return g_error; }

typedef struct pvo_kernel {
    int geometry;
    int image_layers;
    double core_radius;
    double length_x;
    double length_y;
    double radius;
} pvo_kernel;

/* src/vortex.h:7-14 (validateVortexArrays): finite positions and circulations. */
static int validate_arrays(const double *x, const double *y, const double *g, size_t n) {
    for (size_t i = 0; i < n; ++i)
        if (!isfinite(x[i]) || !isfinite(y[i]) || !isfinite(g[i]))
            return fail(PVO_INVALID_ARGUMENT, "vortex data must be finite");
    return PVO_OK;
}

/* src/compute.cpp:31-39 (validatePeriodicCirculation). */
int pvo_validate_periodic_circulation(const double *g, size_t n) {
    double total = 0.0, absolute_total = 0.0;
    for (size_t i = 0; i < n; ++i) {
        total += g[i];
        absolute_total += fabs(g[i]);
    }
    if (fabs(total) > 1e-12 * fmax(1.0, absolute_total))
        return fail(PVO_INVALID_ARGUMENT, "periodic box requires zero total circulation");
    return PVO_OK;
}

/* src/compute.cpp:41-48 (validateDiskPositions). */
int pvo_validate_disk_positions(const double *x, const double *y, size_t n, double r2) {
    for (size_t i = 0; i < n; ++i)
        if (x[i] * x[i] + y[i] * y[i] >= r2)
            return fail(PVO_INVALID_ARGUMENT, "vortex lies on or outside the disk");
    return PVO_OK;
}

/* Constructor checks: src/compute.cpp:50-56 (infinite), 165-173 (periodic),
 * 258-261 (disk), and src/backend_common.cpp:4-13 (geometry selection). */
int pvo_check_kernel(const pvo_kernel *k) {
    if (k->geometry == PVO_INFINITE) {
        const double c = k->core_radius, c2 = c * c;
        if (!(c >= 0.0) || !isfinite(c2) || (c > 0.0 && c2 == 0.0))
            return fail(PVO_INVALID_ARGUMENT,
                        "core radius is negative or outside the supported numeric range");
        return PVO_OK;
    }
    if (k->geometry == PVO_PERIODIC) {
        const double lx = k->length_x, ly = k->length_y;
        if (!(lx > 0.0) || !(ly > 0.0) || !isfinite(lx) || !isfinite(ly) ||
            !isfinite(2.0 * kPi / lx) || k->image_layers < 0 || k->image_layers > 64)
            return fail(PVO_INVALID_ARGUMENT, "invalid periodic-box parameters");
        if (fabs(lx - ly) > 1e-13 * fmax(lx, ly))
            return fail(PVO_INVALID_ARGUMENT,
                        "Weiss-McWilliams kernel requires a square periodic box");
        return PVO_OK;
    }
    if (k->geometry == PVO_DISK) {
        const double r = k->radius, r2 = r * r;
        if (!(r > 0.0) || !isfinite(r2) || r2 == 0.0)
            return fail(PVO_INVALID_ARGUMENT, "invalid disk parameters");
        return PVO_OK;
    }
    return fail(PVO_INVALID_ARGUMENT, "unsupported boundary condition");
}

/* src/compute.cpp:88-133 (InfinitePlaneKernel::evaluateRange). */
static int evaluate_infinite(const pvo_kernel *k, const double *px, const double *py,
                             const double *pg, size_t count, double *pu, double *pv,
                             size_t begin, size_t end) {
    const double eps2 = k->core_radius * k->core_radius;
    const double inverse_two_pi = 0.5 / kPi;
    int singular = 0;
#pragma omp parallel for schedule(static) if (count >= 256) reduction(| : singular)
    for (ptrdiff_t target = (ptrdiff_t)begin; target < (ptrdiff_t)end; ++target) {
        double u = 0.0, v = 0.0;
#pragma omp simd reduction(+ : u, v) reduction(| : singular)
        for (ptrdiff_t source = 0; source < (ptrdiff_t)count; ++source) {
            if (source == target)
                continue;
            const double dx = px[target] - px[source];
            const double dy = py[target] - py[source];
            const double den = dx * dx + dy * dy + eps2;
            if (den == 0.0) {
                singular = 1;
                continue;
            }
            const double c = inverse_two_pi * pg[source] / den;
            u -= c * dy;
            v += c * dx;
        }
        pu[target] = u;
        pv[target] = v;
    }
    if (singular)
        return fail(PVO_RUNTIME, "coincident vortices in singular point-vortex kernel");
    return PVO_OK;
}

/* src/compute.cpp:235-240: 2(sinh^2(a/2)+sin^2(b/2)), infinite beyond |a|>40. */
static inline double periodic_denominator(double a, double sin_half_b) {
    if (fabs(a) > 40.0)
        return INFINITY;
    const double s = sinh(0.5 * a);
    return 2.0 * (s * s + sin_half_b * sin_half_b);
}

/* src/compute.cpp:203-256 (PeriodicBoxKernel::evaluateRange): Weiss--McWilliams sum
 * over 2L+1 image rows, complementary orientations for u and v. */
static int evaluate_periodic(const pvo_kernel *k, const double *x, const double *y,
                             const double *g, size_t count, double *pu, double *pv,
                             size_t begin, size_t end) {
    int status = pvo_validate_periodic_circulation(g, count);
    if (status)
        return status;
    const double lx = k->length_x, ly = k->length_y;
    const int layers = k->image_layers;
    const double wave = 2.0 * kPi / lx;
    const double scale = 0.5 / lx;
    int singular = 0;
#pragma omp parallel for schedule(static) if (count >= 256) reduction(| : singular)
    for (ptrdiff_t target = (ptrdiff_t)begin; target < (ptrdiff_t)end; ++target) {
        double u = 0.0, v = 0.0;
        for (ptrdiff_t source = 0; source < (ptrdiff_t)count; ++source) {
            const double sdx = wave * remainder(x[target] - x[source], lx);
            const double sdy = wave * remainder(y[target] - y[source], ly);
            const double sin_x = sin(sdx), sin_y = sin(sdy);
            const double sin_half_x = sin(0.5 * sdx);
            const double sin_half_y = sin(0.5 * sdy);
            for (int m = -layers; m <= layers; ++m) {
                if (source == target && m == 0)
                    continue;
                const double shifted_x = sdx - 2.0 * kPi * m;
                const double shifted_y = sdy - 2.0 * kPi * m;
                const double den_u = periodic_denominator(shifted_x, sin_half_y);
                const double den_v = periodic_denominator(shifted_y, sin_half_x);
                if (den_u == 0.0 || den_v == 0.0) {
                    singular = 1;
                    continue;
                }
                u -= scale * g[source] * sin_y / den_u;
                v += scale * g[source] * sin_x / den_v;
            }
        }
        pu[target] = u;
        pv[target] = v;
    }
    if (singular)
        return fail(PVO_RUNTIME, "coincident periodic vortices");
    return PVO_OK;
}

/* src/compute.cpp:286-336 (DiskKernel::evaluateRange): direct term for j != i plus
 * the circle-theorem image of every source, the target's own image included. */
static int evaluate_disk(const pvo_kernel *k, const double *x, const double *y, const double *g,
                         size_t count, double *pu, double *pv, size_t begin, size_t end) {
    const double radius = k->radius;
    int status = pvo_validate_disk_positions(x, y, count, radius * radius);
    if (status)
        return status;
    const double inverse_two_pi = 0.5 / kPi;
    const double inverse_radius = 1.0 / radius;
    int singular = 0;
#pragma omp parallel for schedule(static) if (count >= 256) reduction(| : singular)
    for (ptrdiff_t target = (ptrdiff_t)begin; target < (ptrdiff_t)end; ++target) {
        double u = 0.0, v = 0.0;
        const double tx = x[target] * inverse_radius, ty = y[target] * inverse_radius;
        for (ptrdiff_t source = 0; source < (ptrdiff_t)count; ++source) {
            if (source != target) {
                const double dx = x[target] - x[source];
                const double dy = y[target] - y[source];
                const double den = dx * dx + dy * dy;
                if (den == 0.0) {
                    singular = 1;
                    continue;
                }
                const double c = inverse_two_pi * g[source] / den;
                u -= c * dy;
                v += c * dx;
            }
            const double sx = x[source] * inverse_radius, sy = y[source] * inverse_radius;
            const double a = 1.0 - (tx * sx + ty * sy);
            const double b = ty * sx - tx * sy;
            const double den = a * a + b * b;
            const double image_x = -a * sx - b * sy;
            const double image_y = -a * sy + b * sx;
            const double c = -inverse_two_pi * g[source] * inverse_radius / den;
            u -= c * image_y;
            v += c * image_x;
        }
        pu[target] = u;
        pv[target] = v;
    }
    if (singular)
        return fail(PVO_RUNTIME, "coincident vortices in disk");
    return PVO_OK;
}

/* VelocityKernel::evaluateRange for the configured geometry (src/compute.h:24-26). */
int pvo_evaluate_range(const pvo_kernel *k, const double *x, const double *y, const double *g,
                       size_t n, double *u, double *v, size_t begin, size_t end) {
    int status = pvo_check_kernel(k);
    if (status)
        return status;
    status = validate_arrays(x, y, g, n);
    if (status)
        return status;
    if (begin > end || end > n)
        return fail(PVO_OUT_OF_RANGE, "invalid target-vortex range");
    switch (k->geometry) {
    case PVO_INFINITE:
        return evaluate_infinite(k, x, y, g, n, u, v, begin, end);
    case PVO_PERIODIC:
        return evaluate_periodic(k, x, y, g, n, u, v, begin, end);
    default:
        return evaluate_disk(k, x, y, g, n, u, v, begin, end);
    }
}

/* src/compute.cpp:60-69 (VelocityKernel::evaluate): full range plus a finiteness check. */
int pvo_evaluate(const pvo_kernel *k, const double *x, const double *y, const double *g,
                 size_t n, double *u, double *v) {
    int status = pvo_evaluate_range(k, x, y, g, n, u, v, 0, n);
    if (status)
        return status;
    for (size_t i = 0; i < n; ++i)
        if (!isfinite(u[i]) || !isfinite(v[i]))
            return fail(PVO_RUNTIME, "non-finite velocity; check scales and close encounters");
    return PVO_OK;
}

/* src/compute.cpp:11-27 (computeMoments): circulation, linear and angular impulse. */
int pvo_moments(const double *x, const double *y, const double *g, size_t n, double out[4]) {
    int status = validate_arrays(x, y, g, n);
    if (status)
        return status;
    double c = 0.0, ix = 0.0, iy = 0.0, a = 0.0;
    for (size_t i = 0; i < n; ++i) {
        const double gamma = g[i];
        c += gamma;
        ix += gamma * x[i];
        iy += gamma * y[i];
        a += gamma * (x[i] * x[i] + y[i] * y[i]);
    }
    if (!isfinite(c) || !isfinite(ix) || !isfinite(iy) || !isfinite(a))
        return fail(PVO_RUNTIME, "non-finite vortex moment; check input scales");
    out[0] = c;
    out[1] = ix;
    out[2] = iy;
    out[3] = a;
    return PVO_OK;
}

/* src/compute.cpp:134-156 (computeInvariants with a core radius): the infinite-plane
 * Hamiltonian -sum_{i<j} G_i G_j log(r^2+eps^2) / 4pi. */
static int hamiltonian_infinite(const pvo_kernel *k, const double *x, const double *y,
                                const double *g, size_t n, double *out) {
    const double eps2 = k->core_radius * k->core_radius;
    double h = 0.0;
    for (size_t i = 0; i < n; ++i) {
        for (size_t j = i + 1; j < n; ++j) {
            const double dx = x[i] - x[j];
            const double dy = y[i] - y[j];
            const double r2 = dx * dx + dy * dy + eps2;
            if (r2 == 0.0)
                return fail(PVO_RUNTIME, "coincident vortices");
            h -= g[i] * g[j] * log(r2) / (4.0 * kPi);
        }
    }
    *out = h;
    return PVO_OK;
}

/* src/compute.cpp:174-202 (PeriodicBoxKernel::hamiltonian): all 2L+1 image rows,
 * no |a|>40 cutoff, with the quadratic lattice-limit term. */
static int hamiltonian_periodic(const pvo_kernel *k, const double *x, const double *y,
                                const double *g, size_t n, double *out) {
    const double lx = k->length_x, ly = k->length_y;
    const int layers = k->image_layers;
    const double wave = 2.0 * kPi / lx;
    double h = 0.0;
    for (size_t i = 0; i < n; ++i) {
        for (size_t j = i + 1; j < n; ++j) {
            const double dx = wave * remainder(x[i] - x[j], lx);
            const double dy = wave * remainder(y[i] - y[j], ly);
            double e = -dx * dx / (2.0 * kPi);
            for (int m = -layers; m <= layers; ++m) {
                const double shifted = dx - 2.0 * kPi * m;
                const double sh = sinh(0.5 * shifted);
                const double sn = sin(0.5 * dy);
                const double den = 2.0 * (sh * sh + sn * sn);
                if (den == 0.0)
                    return fail(PVO_RUNTIME, "coincident periodic vortices");
                const double arg = 2.0 * kPi * m;
                const double mag = fabs(arg);
                const double log_cosh = mag + log1p(exp(-2.0 * mag)) - log(2.0);
                e += log(den) - log_cosh;
            }
            h -= g[i] * g[j] * e / (4.0 * kPi);
        }
    }
    *out = h;
    return PVO_OK;
}

/* src/compute.cpp:262-285 (DiskKernel::hamiltonian): self-image terms plus pair terms. */
static int hamiltonian_disk(const pvo_kernel *k, const double *x, const double *y,
                            const double *g, size_t n, double *out) {
    const double radius = k->radius, r2 = radius * radius;
    double h = 0.0;
    for (size_t i = 0; i < n; ++i) {
        const double ri2 = x[i] * x[i] + y[i] * y[i];
        if (ri2 >= r2)
            return fail(PVO_INVALID_ARGUMENT, "vortex outside disk");
        h += g[i] * g[i] * log((r2 - ri2) / radius) / (4.0 * kPi);
        for (size_t j = i + 1; j < n; ++j) {
            const double dx = x[i] - x[j], dy = y[i] - y[j];
            const double num = (dx * dx + dy * dy) * r2;
            const double dot = x[i] * x[j] + y[i] * y[j];
            const double cross = y[i] * x[j] - x[i] * y[j];
            const double den = (r2 - dot) * (r2 - dot) + cross * cross;
            if (num == 0.0)
                return fail(PVO_RUNTIME, "coincident vortices in disk");
            h -= g[i] * g[j] * log(num / den) / (4.0 * kPi);
        }
    }
    *out = h;
    return PVO_OK;
}

/* VelocityKernel::hamiltonian (src/compute.h:27) for the configured geometry. */
int pvo_hamiltonian(const pvo_kernel *k, const double *x, const double *y, const double *g,
                    size_t n, double *out) {
    int status = pvo_check_kernel(k);
    if (status)
        return status;
    status = validate_arrays(x, y, g, n);
    if (status)
        return status;
    switch (k->geometry) {
    case PVO_INFINITE:
        return hamiltonian_infinite(k, x, y, g, n, out);
    case PVO_PERIODIC:
        return hamiltonian_periodic(k, x, y, g, n, out);
    default:
        return hamiltonian_disk(k, x, y, g, n, out);
    }
}

/* src/compute.cpp:157-163 (computeInvariants(vortices, kernel)): moments + Hamiltonian.
 * out = {circulation, linear_impulse_x, linear_impulse_y, angular_impulse, hamiltonian}. */
int pvo_invariants(const pvo_kernel *k, const double *x, const double *y, const double *g,
                   size_t n, double out[5]) {
    int status = pvo_moments(x, y, g, n, out);
    if (status)
        return status;
    status = pvo_hamiltonian(k, x, y, g, n, &out[4]);
    if (status)
        return status;
    if (!isfinite(out[4]))
        return fail(PVO_RUNTIME, "non-finite Hamiltonian; check input scales");
    return PVO_OK;
}

/* ---------------------------------------------------------------------------------
 * Integrators: src/timestep.h:8-34 and src/timestep.cpp:1-120.
 * ------------------------------------------------------------------------------- */

/* Dormand--Prince tableau rows a_{s,j} (src/timestep.h:8-15). */
static const double kDopriA[7][7] = {
    {0},
    {1.0 / 5.0},
    {3.0 / 40.0, 9.0 / 40.0},
    {44.0 / 45.0, -56.0 / 15.0, 32.0 / 9.0},
    {19372.0 / 6561.0, -25360.0 / 2187.0, 64448.0 / 6561.0, -212.0 / 729.0},
    {9017.0 / 3168.0, -355.0 / 33.0, 46732.0 / 5247.0, 49.0 / 176.0, -5103.0 / 18656.0},
    {35.0 / 384.0, 0.0, 500.0 / 1113.0, 125.0 / 192.0, -2187.0 / 6784.0, 11.0 / 84.0}};
/* Fifth- and fourth-order weights (src/timestep.cpp:69-73). */
static const double kB5[7] = {35.0 / 384.0,     0.0, 500.0 / 1113.0, 125.0 / 192.0,
                              -2187.0 / 6784.0, 11.0 / 84.0, 0.0};
static const double kB4[7] = {5179.0 / 57600.0, 0.0,           7571.0 / 16695.0, 393.0 / 640.0,
                              -92097.0 / 339200.0, 187.0 / 2100.0, 1.0 / 40.0};

typedef struct pvo_integrator {
    size_t n;
    double *ix, *iy; /* initial_ */
    double *tx, *ty; /* temporary_ */
    double *kx[7], *ky[7];
    int fsal_valid;
} pvo_integrator;

/* src/timestep.cpp:4-8 (constructor) and 9-18 (ensureSize). */
static int ensure_size(pvo_integrator *it, size_t n) {
    if (it->n == n && it->ix)
        return PVO_OK;
    free(it->ix);
    free(it->iy);
    free(it->tx);
    free(it->ty);
    for (int s = 0; s < 7; ++s) {
        free(it->kx[s]);
        free(it->ky[s]);
    }
    const size_t bytes = (n ? n : 1) * sizeof(double);
    it->ix = calloc(1, bytes);
    it->iy = calloc(1, bytes);
    it->tx = calloc(1, bytes);
    it->ty = calloc(1, bytes);
    for (int s = 0; s < 7; ++s) {
        it->kx[s] = calloc(1, bytes);
        it->ky[s] = calloc(1, bytes);
    }
    it->n = n;
    it->fsal_valid = 0;
    return PVO_OK;
}

pvo_integrator *pvo_integrator_create(size_t n) {
    pvo_integrator *it = calloc(1, sizeof *it);
    if (it)
        ensure_size(it, n);
    return it;
}

void pvo_integrator_destroy(pvo_integrator *it) {
    if (!it)
        return;
    free(it->ix);
    free(it->iy);
    free(it->tx);
    free(it->ty);
    for (int s = 0; s < 7; ++s) {
        free(it->kx[s]);
        free(it->ky[s]);
    }
    free(it);
}

/* src/timestep.h:25 (invalidateCachedDerivative). */
void pvo_integrator_invalidate(pvo_integrator *it) { it->fsal_valid = 0; }

/* src/timestep.cpp:19-30 (makeStage): temporary = initial + dt * sum_{j<stage} c_j k_j. */
static void make_stage(pvo_integrator *it, int stage, const double *c, double dt) {
    for (size_t i = 0; i < it->n; ++i) {
        double dx = 0.0, dy = 0.0;
        for (int j = 0; j < stage; ++j) {
            dx += c[j] * it->kx[j][i];
            dy += c[j] * it->ky[j][i];
        }
        it->tx[i] = it->ix[i] + dt * dx;
        it->ty[i] = it->iy[i] + dt * dy;
    }
}

/* src/timestep.cpp:31-61 (rk4Step). */
int pvo_rk4_step(pvo_integrator *it, const pvo_kernel *k, double *x, double *y, const double *g,
                 size_t n, double dt) {
    if (!isfinite(dt) || !(dt > 0.0))
        return fail(PVO_INVALID_ARGUMENT, "timestep must be finite and positive");
    int status = validate_arrays(x, y, g, n);
    if (status)
        return status;
    ensure_size(it, n);
    memcpy(it->ix, x, n * sizeof(double));
    memcpy(it->iy, y, n * sizeof(double));
    it->fsal_valid = 0;
    static const double a2[7] = {0.5}, a3[7] = {0.0, 0.5}, a4[7] = {0.0, 0.0, 1.0};
    if ((status = pvo_evaluate(k, it->ix, it->iy, g, n, it->kx[0], it->ky[0])))
        return status;
    make_stage(it, 1, a2, dt);
    if ((status = pvo_evaluate(k, it->tx, it->ty, g, n, it->kx[1], it->ky[1])))
        return status;
    make_stage(it, 2, a3, dt);
    if ((status = pvo_evaluate(k, it->tx, it->ty, g, n, it->kx[2], it->ky[2])))
        return status;
    make_stage(it, 3, a4, dt);
    if ((status = pvo_evaluate(k, it->tx, it->ty, g, n, it->kx[3], it->ky[3])))
        return status;
    for (size_t i = 0; i < n; ++i) {
        it->tx[i] = it->ix[i] + dt *
                                    (it->kx[0][i] + 2.0 * it->kx[1][i] + 2.0 * it->kx[2][i] +
                                     it->kx[3][i]) /
                                    6.0;
        it->ty[i] = it->iy[i] + dt *
                                    (it->ky[0][i] + 2.0 * it->ky[1][i] + 2.0 * it->ky[2][i] +
                                     it->ky[3][i]) /
                                    6.0;
    }
    if ((status = validate_arrays(it->tx, it->ty, g, n)))
        return status;
    memcpy(x, it->tx, n * sizeof(double));
    memcpy(y, it->ty, n * sizeof(double));
    return PVO_OK;
}

static inline double clampd(double v, double lo, double hi) {
    return v < lo ? lo : (hi < v ? hi : v);
}

/* src/timestep.cpp:62-119 (dopri5Step): FSAL, max-norm error, controller, <=32 rejects.
 * out = {accepted dt, suggested dt, normalized error, rejected steps}. */
int pvo_dopri5_step(pvo_integrator *it, const pvo_kernel *k, double *x, double *y,
                    const double *g, size_t n, double dt, double atol, double rtol,
                    double min_dt, double max_dt, double out[4]) {
    if (!isfinite(dt) || !(dt > 0.0))
        return fail(PVO_INVALID_ARGUMENT, "timestep must be finite and positive");
    int status = validate_arrays(x, y, g, n);
    if (status)
        return status;
    if (it->n != n)
        ensure_size(it, n);
    memcpy(it->ix, x, n * sizeof(double));
    memcpy(it->iy, y, n * sizeof(double));
    unsigned rejected = 0;
    if (!it->fsal_valid)
        if ((status = pvo_evaluate(k, it->ix, it->iy, g, n, it->kx[0], it->ky[0])))
            return status;
    for (;;) {
        for (int s = 1; s < 7; ++s) {
            make_stage(it, s, kDopriA[s], dt);
            if ((status = pvo_evaluate(k, it->tx, it->ty, g, n, it->kx[s], it->ky[s])))
                return status;
        }
        double error = 0.0;
        for (size_t i = 0; i < n; ++i) {
            const double x5 = it->tx[i], y5 = it->ty[i];
            double ex = 0.0, ey = 0.0;
            for (int s = 0; s < 7; ++s) {
                ex += dt * (kB5[s] - kB4[s]) * it->kx[s][i];
                ey += dt * (kB5[s] - kB4[s]) * it->ky[s][i];
            }
            const double scx = atol + rtol * fmax(fabs(it->ix[i]), fabs(x5));
            const double scy = atol + rtol * fmax(fabs(it->iy[i]), fabs(y5));
            if (!isfinite(ex) || !isfinite(ey) || !isfinite(scx) || !isfinite(scy) ||
                !(scx > 0.0) || !(scy > 0.0))
                return fail(PVO_RUNTIME, "non-finite adaptive error estimate or invalid tolerance");
            error = fmax(error, fabs(ex) / scx);
            error = fmax(error, fabs(ey) / scy);
        }
        const double factor = error == 0.0 ? 5.0 : clampd(0.9 * pow(error, -0.2), 0.2, 5.0);
        const double suggested = clampd(dt * factor, min_dt, max_dt);
        if (error <= 1.0) {
            memcpy(x, it->tx, n * sizeof(double));
            memcpy(y, it->ty, n * sizeof(double));
            double *t;
            t = it->kx[0], it->kx[0] = it->kx[6], it->kx[6] = t;
            t = it->ky[0], it->ky[0] = it->ky[6], it->ky[6] = t;
            it->fsal_valid = 1;
            out[0] = dt;
            out[1] = suggested;
            out[2] = error;
            out[3] = (double)rejected;
            return PVO_OK;
        }
        if (dt <= min_dt || ++rejected > 32)
            return fail(PVO_RUNTIME, "adaptive integrator could not satisfy tolerance");
        dt = fmax(min_dt, fmin(suggested, dt * 0.9));
    }
}

/* src/dipole.cpp:87-119 (the search part of DipoleManager::process): every i<j pair with
 * nonzero circulations of opposite sign closer than the removal distance (periodic: minimum
 * image via remainder), plus, in a disk, the real/image wall candidate (i, n).
 * Candidates are written in the reference's discovery order; *count is the total found
 * even when it exceeds cap. */
int pvo_dipole_candidates(int geometry, double length_x, double length_y, double radius,
                          double distance, const double *x, const double *y, const double *g,
                          size_t n, uint64_t *first, uint64_t *second, double *d2, size_t cap,
                          size_t *count) {
    const int periodic = geometry == PVO_PERIODIC;
    const double thr2 = distance * distance;
    size_t found = 0;
    for (size_t i = 0; i < n; ++i) {
        if (g[i] == 0.0)
            continue;
        for (size_t j = i + 1; j < n; ++j) {
            if (g[j] == 0.0 || signbit(g[i]) == signbit(g[j]))
                continue;
            const double dx = periodic ? remainder(x[i] - x[j], length_x) : x[i] - x[j];
            const double dy = periodic ? remainder(y[i] - y[j], length_y) : y[i] - y[j];
            const double dd = dx * dx + dy * dy;
            if (dd < thr2) {
                if (found < cap) {
                    first[found] = i;
                    second[found] = j;
                    d2[found] = dd;
                }
                ++found;
            }
        }
        if (geometry == PVO_DISK) {
            const double rr = radius * radius;
            const double ri2 = x[i] * x[i] + y[i] * y[i];
            if (ri2 > 0.0) {
                const double image_distance = (rr - ri2) / sqrt(ri2);
                if (image_distance * image_distance < thr2) {
                    if (found < cap) {
                        first[found] = i;
                        second[found] = n;
                        d2[found] = image_distance * image_distance;
                    }
                    ++found;
                }
            }
        }
    }
    *count = found;
    return PVO_OK;
}
