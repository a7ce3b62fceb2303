/*
 * smoc_oracle.c -- CPU restatement of SimpleMOC-kernel's segment attenuation.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, never the product): loaded by
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg.
 *
 * Compiled with strict IEEE settings (-O2 -ffp-contract=off, no fast-math) so
 * the FP32 functions evaluate the reference's expressions in C's literal
 * order; see oracle/Makefile.  Citations are to the reference tree
 * (ANL-CESAR/SimpleMOC-kernel, Version 4).
 */
#include "smoc_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* Placeholder geometry constants of src/cpu/kernel.c:99-104 (ds and zin are
 * per-segment inputs here; the reference fixes them at 0.7f and 0.3f). */
static const float K_DZ = 0.1f;
static const float K_WEIGHT = 0.5f;
static const float K_MU = 0.9f;
static const float K_MU2 = 0.3f;

/* src/cpu/init.c:81-117.  N = (int)(maxVal*sqrt(1/(8*precision*0.01)));
 * dx = maxVal/N; pair n = (-exp(-n dx), 1 + (n dx - 1) exp(-n dx)) -- the
 * slope's sign is the reference's (the derivative of 1-e^-x is +e^-x; the
 * reference stores -e^-x, and this restatement keeps it). */
int oracle_build_table(float precision, float maxVal, oracle_table *t)
{
    int N = (int)(maxVal * sqrt(1.0 / (8.0 * precision * 0.01)));
    if (N <= 0 || 2 * N > (int)(sizeof(t->values) / sizeof(t->values[0])))
        return -1;
    float dx = maxVal / (float)N;
    for (int n = 0; n < N; n++) {
        float exponential = (float)exp(-n * dx);
        t->values[2 * n] = -exponential;
        t->values[2 * n + 1] = 1 + (n * dx - 1) * exponential;
    }
    t->dx = dx;
    t->maxVal = maxVal - dx;
    t->N = N;
    return 0;
}

/* src/cpu/kernel.c:340-355: interval = (int)(x/dx + 0.5f*dx) (0.5f*dx, not
 * 0.5f, as in the reference). */
int oracle_table_interval(const oracle_table *t, float x)
{
    if (x > t->maxVal)
        return -1;
    return (int)(x / t->dx + 0.5f * t->dx);
}

/* src/cpu/kernel.c:337-361 */
float oracle_table_lookup_f32(const oracle_table *t, float x)
{
    int interval = oracle_table_interval(t, x);
    if (interval < 0)
        return 1.0f;
    interval = interval * 2;
    float slope = t->values[interval];
    float intercept = t->values[interval + 1];
    float val = slope * x + intercept;
    return val;
}

void oracle_decode_z(float z, int F, int *f, float *zin)
{
    float zf = z * (float)F;
    int fi = (int)zf;
    if (fi > F - 1)
        fi = F - 1;
    if (fi < 0)
        fi = 0;
    *f = fi;
    *zin = ((zf - (float)fi) - 0.5f) * K_DZ;
}

/* src/cpu/kernel.c:75-333, one (segment, group) at a time in the literal
 * operand order of each reference loop. */
void oracle_segment_f32(int G, int F, const float *sig, const float *src, int fai,
                        float ds, float zin, int exp_mode, const oracle_table *tab,
                        float *psi, float *flux)
{
    const float dz = K_DZ, weight = K_WEIGHT, mu = K_MU, mu2 = K_MU2;
    float *FSR_flux = flux + (long)fai * G; /* kernel.c:109 */
    for (int g = 0; g < G; g++) {
        float q0, q1, q2;
        if (fai == 0) { /* kernel.c:111-136 */
            const float y2 = src[(long)fai * G + g];
            const float y3 = src[(long)(fai + 1) * G + g];
            const float c0 = y2;
            const float c1 = (y3 - y2) / dz;
            q0 = c0 + c1 * zin;
            q1 = c1;
            q2 = 0;
        } else if (fai == F - 1) { /* kernel.c:137-162 */
            const float y1 = src[(long)(fai - 1) * G + g];
            const float y2 = src[(long)fai * G + g];
            const float c0 = y2;
            const float c1 = (y2 - y1) / dz;
            q0 = c0 + c1 * zin;
            q1 = c1;
            q2 = 0;
        } else { /* kernel.c:163-191 */
            const float y1 = src[(long)(fai - 1) * G + g];
            const float y2 = src[(long)fai * G + g];
            const float y3 = src[(long)(fai + 1) * G + g];
            const float c0 = y2;
            const float c1 = (y1 - y3) / (2.f * dz);
            const float c2 = (y1 - 2.f * y2 + y3) / (2.f * dz * dz);
            q0 = c0 + c1 * zin + c2 * zin * zin;
            q1 = c1 + 2.f * c2 * zin;
            q2 = c2;
        }
        /* kernel.c:200-208 */
        const float sigT = sig[g];
        const float tau = sigT * ds;
        const float sigT2 = sigT * sigT;
        /* kernel.c:216-223 */
        float expVal;
        if (exp_mode == ORACLE_EXP_TABLE)
            expVal = oracle_table_lookup_f32(tab, tau);
        else
            expVal = 1.f - expf(-tau);
        /* kernel.c:233-237 */
        const float reuse = tau * (tau - 2.f) + 2.f * expVal / (sigT * sigT2);
        /* kernel.c:245-252 */
        const float flux_integral =
            (q0 * tau + (sigT * psi[g] - q0) * expVal) / sigT2 + q1 * mu * reuse +
            q2 * mu2 * (tau * (tau * (tau - 3.f) + 6.f) - 6.f * expVal) /
                (3.f * sigT2 * sigT2);
        /* kernel.c:259-263, 274-277 */
        const float tally = weight * flux_integral;
        FSR_flux[g] += tally;
        /* kernel.c:289-332 */
        const float t1 = q0 * expVal / sigT;
        const float t2 = q1 * mu * (tau - expVal) / sigT2;
        const float t3 = q2 * mu2 * reuse;
        const float t4 = psi[g] * (1.f - expVal);
        psi[g] = t1 + t2 + t3 + t4;
    }
}

/* The FP64 flux integral's q2 term (kernel.c:250-252) is multiplied by this
 * constant, 1 in every build of the checker.  tests/test_oracle.py compiles a
 * mutant with -DORACLE_Q2_TERM_SCALE=2 to show that its pin of this term
 * against the reference's golden tallies fails on such a slip. */
#ifndef ORACLE_Q2_TERM_SCALE
#define ORACLE_Q2_TERM_SCALE 1.0
#endif

void oracle_segment_f64(int G, int F, const float *sig, const float *src, int fai,
                        float ds_f, float zin_f, int exp_mode, const oracle_table *tab,
                        double *psi, double *flux)
{
    const double dz = K_DZ, weight = K_WEIGHT, mu = K_MU, mu2 = K_MU2;
    const double ds = ds_f, zin = zin_f;
    double *FSR_flux = flux + (long)fai * G;
    for (int g = 0; g < G; g++) {
        double q0, q1, q2;
        if (fai == 0) {
            const double y2 = src[(long)fai * G + g];
            const double y3 = src[(long)(fai + 1) * G + g];
            const double c1 = (y3 - y2) / dz;
            q0 = y2 + c1 * zin;
            q1 = c1;
            q2 = 0;
        } else if (fai == F - 1) {
            const double y1 = src[(long)(fai - 1) * G + g];
            const double y2 = src[(long)fai * G + g];
            const double c1 = (y2 - y1) / dz;
            q0 = y2 + c1 * zin;
            q1 = c1;
            q2 = 0;
        } else {
            const double y1 = src[(long)(fai - 1) * G + g];
            const double y2 = src[(long)fai * G + g];
            const double y3 = src[(long)(fai + 1) * G + g];
            const double c1 = (y1 - y3) / (2.0 * dz);
            const double c2 = (y1 - 2.0 * y2 + y3) / (2.0 * dz * dz);
            q0 = y2 + c1 * zin + c2 * zin * zin;
            q1 = c1 + 2.0 * c2 * zin;
            q2 = c2;
        }
        const double sigT = sig[g];
        const double tau = sigT * ds;
        const double sigT2 = sigT * sigT;
        double expVal;
        if (exp_mode == ORACLE_EXP_TABLE) {
            /* interval from the FP32 tau, exactly as the FP32 path picks it */
            const float tau_f = sig[g] * ds_f;
            const int iv = oracle_table_interval(tab, tau_f);
            if (iv < 0)
                expVal = 1.0;
            else
                expVal = (double)tab->values[2 * iv] * tau + (double)tab->values[2 * iv + 1];
        } else {
            expVal = 1.0 - exp(-tau);
        }
        const double reuse = tau * (tau - 2.0) + 2.0 * expVal / (sigT * sigT2);
        const double flux_integral =
            (q0 * tau + (sigT * psi[g] - q0) * expVal) / sigT2 + q1 * mu * reuse +
            ORACLE_Q2_TERM_SCALE * q2 * mu2 * (tau * (tau * (tau - 3.0) + 6.0) - 6.0 * expVal) /
                (3.0 * sigT2 * sigT2);
        FSR_flux[g] += weight * flux_integral;
        const double t1 = q0 * expVal / sigT;
        const double t2 = q1 * mu * (tau - expVal) / sigT2;
        const double t3 = q2 * mu2 * reuse;
        const double t4 = psi[g] * (1.0 - expVal);
        psi[g] = t1 + t2 + t3 + t4;
    }
}

static int oracle_table_for(const oracle_problem *p, oracle_table *tab)
{
    if (p->exp_mode == ORACLE_EXP_TABLE)
        return oracle_build_table(0.01f, 10.0f, tab); /* src/cpu/main.c:33 */
    memset(tab, 0, sizeof(*tab));
    return 0;
}

static void sweep_track_f64(const oracle_problem *p, const oracle_table *tab, int64_t t,
                            double *flux, double *psi)
{
    const int G = p->G, F = p->F;
    for (int g = 0; g < G; g++)
        psi[g] = p->psi_in[t * G + g];
    for (int64_t k = p->track_off[t]; k < p->track_off[t + 1]; k++) {
        int f;
        float zin;
        oracle_decode_z(p->z[k], F, &f, &zin);
        const int32_t r = p->region[k];
        oracle_segment_f64(G, F, p->sigT + (long)r * G, p->fine_source + (long)r * F * G, f,
                           p->len[k], zin, p->exp_mode, tab, psi, flux + (long)r * F * G);
    }
}

void oracle_sweep_f64(const oracle_problem *p, double *fine_flux, double *psi_out,
                      int nthreads)
{
    static oracle_table tab;
    oracle_table_for(p, &tab);
    const int64_t nflux = (int64_t)p->R * p->F * p->G;
    if (nthreads <= 1) {
        for (int64_t t = 0; t < p->T; t++)
            sweep_track_f64(p, &tab, t, fine_flux, psi_out + t * p->G);
        return;
    }
#ifdef _OPENMP
    double **priv = (double **)calloc((size_t)nthreads, sizeof(double *));
    #pragma omp parallel num_threads(nthreads)
    {
        const int th = omp_get_thread_num();
        double *mine = (double *)calloc((size_t)nflux, sizeof(double));
        priv[th] = mine;
        #pragma omp for schedule(static)
        for (int64_t t = 0; t < p->T; t++)
            sweep_track_f64(p, &tab, t, mine, psi_out + t * p->G);
    }
    for (int th = 0; th < nthreads; th++) {
        if (!priv[th])
            continue;
        #pragma omp parallel for num_threads(nthreads) schedule(static)
        for (int64_t i = 0; i < nflux; i++)
            fine_flux[i] += priv[th][i];
        free(priv[th]);
    }
    free(priv);
#else
    for (int64_t t = 0; t < p->T; t++)
        sweep_track_f64(p, &tab, t, fine_flux, psi_out + t * p->G);
#endif
}

void oracle_sweep_f32(const oracle_problem *p, float *fine_flux, float *psi_out)
{
    static oracle_table tab;
    oracle_table_for(p, &tab);
    const int G = p->G, F = p->F;
    for (int64_t t = 0; t < p->T; t++) {
        float *psi = psi_out + t * G;
        for (int g = 0; g < G; g++)
            psi[g] = p->psi_in[t * G + g];
        for (int64_t k = p->track_off[t]; k < p->track_off[t + 1]; k++) {
            int f;
            float zin;
            oracle_decode_z(p->z[k], F, &f, &zin);
            const int32_t r = p->region[k];
            oracle_segment_f32(G, F, p->sigT + (long)r * G, p->fine_source + (long)r * F * G,
                               f, p->len[k], zin, p->exp_mode, &tab, psi,
                               fine_flux + (long)r * F * G);
        }
    }
}

double oracle_baseline_f32(const oracle_problem *p, float *fine_flux, float *psi_out,
                           int nthreads)
{
    static oracle_table tab;
    oracle_table_for(p, &tab);
    const int G = p->G, F = p->F;
#ifdef _OPENMP
    const int64_t nlocks = (int64_t)p->R * F;
    omp_lock_t *locks = (omp_lock_t *)malloc((size_t)nlocks * sizeof(omp_lock_t));
    for (int64_t i = 0; i < nlocks; i++)
        omp_init_lock(&locks[i]);
    const double t0 = omp_get_wtime();
    #pragma omp parallel num_threads(nthreads)
    {
        float *tmp = (float *)malloc((size_t)F * G * sizeof(float));
        #pragma omp for schedule(dynamic, 1)
        for (int64_t t = 0; t < p->T; t++) {
            float *psi = psi_out + t * G;
            for (int g = 0; g < G; g++)
                psi[g] = p->psi_in[t * G + g];
            for (int64_t k = p->track_off[t]; k < p->track_off[t + 1]; k++) {
                int f;
                float zin;
                oracle_decode_z(p->z[k], F, &f, &zin);
                const int32_t r = p->region[k];
                /* compute this segment's tally privately, then add it under
                 * the (region, interval) lock as kernel.c:265-281 does */
                memset(tmp + (long)f * G, 0, (size_t)G * sizeof(float));
                oracle_segment_f32(G, F, p->sigT + (long)r * G,
                                   p->fine_source + (long)r * F * G, f, p->len[k], zin,
                                   p->exp_mode, &tab, psi, tmp);
                omp_set_lock(&locks[(int64_t)r * F + f]);
                float *dst = fine_flux + ((long)r * F + f) * G;
                for (int g = 0; g < G; g++)
                    dst[g] += tmp[(long)f * G + g];
                omp_unset_lock(&locks[(int64_t)r * F + f]);
            }
        }
        free(tmp);
    }
    const double t1 = omp_get_wtime();
    for (int64_t i = 0; i < nlocks; i++)
        omp_destroy_lock(&locks[i]);
    free(locks);
    return t1 - t0;
#else
    (void)nthreads;
    oracle_sweep_f32(p, fine_flux, psi_out);
    return -1.0;
#endif
}

double oracle_sum_f64(const double *x, int64_t n)
{
    double s = 0.0;
    for (int64_t i = 0; i < n; i++)
        s += x[i];
    return s;
}

double oracle_sum_f32(const float *x, int64_t n)
{
    double s = 0.0;
    for (int64_t i = 0; i < n; i++)
        s += (double)x[i];
    return s;
}


/* ---- track producer restatement ------------------------------------------- */
/* counter-based generator of the library: draw(key, i) = mix(key + (i+1) *
 * golden), key = mix(seed * golden + stream); the tracer's stream id is 9 and
 * track t uses draw indices t * 64 .. t * 64 + 63. */
static uint64_t tr_mix(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
#define TR_GOLDEN 0x9E3779B97F4A7C15ULL
static double tr_u53(uint64_t key, uint64_t i)
{
    return (double)(tr_mix(key + (i + 1) * TR_GOLDEN) >> 11) * 0x1p-53;
}
static double tr_face(double p, double d, double hi)
{
    if (d > 0.0)
        return (hi - p) / d;
    if (d < 0.0)
        return (0.0 - p) / d;
    return 1e300;
}

void oracle_trace_ray(const oracle_geometry *g, uint64_t seed, int64_t t, oracle_ray *r)
{
    const uint64_t key = tr_mix(seed * TR_GOLDEN + 9);
    const double X = g->nx * g->pitch, Y = g->ny * g->pitch, Z = g->height;
    const uint64_t b = (uint64_t)t * 64;
    const double px = X * tr_u53(key, b), py = Y * tr_u53(key, b + 1), pz = Z * tr_u53(key, b + 2);
    double dx = 0.0, dy = 0.0, dz = 1.0;
    /* isotropic direction, Marsaglia's disc method (up to 30 attempts) */
    for (int k = 0; k < 30; k++) {
        const double a = 2.0 * tr_u53(key, b + 3 + 2 * k) - 1.0;
        const double c = 2.0 * tr_u53(key, b + 4 + 2 * k) - 1.0;
        const double s = a * a + c * c;
        if (s > 0.0 && s < 1.0) {
            const double w = 2.0 * sqrt(1.0 - s);
            dx = a * w;
            dy = c * w;
            dz = 1.0 - 2.0 * s;
            break;
        }
    }
    const double back = fmin(fmin(tr_face(px, -dx, X), tr_face(py, -dy, Y)), tr_face(pz, -dz, Z));
    const double fwd = fmin(fmin(tr_face(px, dx, X), tr_face(py, dy, Y)), tr_face(pz, dz, Z));
    r->ex = px - back * dx;
    r->ey = py - back * dy;
    r->ez = pz - back * dz;
    r->dx = dx;
    r->dy = dy;
    r->dz = dz;
    r->L = back + fwd;
}

static int tr_clampi(int v, int hi) { return v < 0 ? 0 : (v > hi ? hi : v); }

int64_t oracle_trace_track(const oracle_geometry *g, uint64_t seed, int64_t t, int64_t cap,
                           float *len, int32_t *region, float *z)
{
    oracle_ray r;
    oracle_trace_ray(g, seed, t, &r);
    const int nx = g->nx, ny = g->ny, nz = g->nz;
    const double p = g->pitch, rr = g->pin_radius * g->pin_radius, dzl = g->height / nz;
    int ix = tr_clampi((int)floor((r.ex + 1e-9 * r.dx) / p), nx - 1);
    int iy = tr_clampi((int)floor((r.ey + 1e-9 * r.dy) / p), ny - 1);
    const int sx = r.dx > 0.0 ? 1 : -1, sy = r.dy > 0.0 ? 1 : -1;
    int64_t n = 0;
    double tc = 0.0;
    while (tc < r.L && ix >= 0 && ix < nx && iy >= 0 && iy < ny) {
        /* next cell faces, as parameters from the entry point */
        const double tx = r.dx != 0.0 ? ((ix + (sx > 0 ? 1 : 0)) * p - r.ex) / r.dx : 1e300;
        const double ty = r.dy != 0.0 ? ((iy + (sy > 0 ? 1 : 0)) * p - r.ey) / r.dy : 1e300;
        double tend = fmin(fmin(tx, ty), r.L);
        if (tend < tc)
            tend = tc;
        /* the cell's pin circle */
        const double cx = (ix + 0.5) * p, cy = (iy + 0.5) * p;
        const double ox = r.ex - cx, oy = r.ey - cy;
        const double A = r.dx * r.dx + r.dy * r.dy;
        const double B = ox * r.dx + oy * r.dy;
        const double C = ox * ox + oy * oy - rr;
        double cut[4];
        int nc = 0;
        cut[nc++] = tc;
        if (A > 0.0) {
            const double disc = B * B - A * C;
            if (disc > 0.0) {
                const double sq = sqrt(disc);
                const double s1 = (-B - sq) / A, s2 = (-B + sq) / A;
                if (s1 > tc && s1 < tend)
                    cut[nc++] = s1;
                if (s2 > tc && s2 < tend)
                    cut[nc++] = s2;
            }
        }
        cut[nc++] = tend;
        for (int k = 0; k + 1 < nc; k++) {
            const double a = cut[k], b = cut[k + 1];
            if (!(b > a))
                continue;
            const double sm = 0.5 * (a + b);
            const double qx = r.ex + sm * r.dx - cx, qy = r.ey + sm * r.dy - cy;
            const int outside = qx * qx + qy * qy < rr ? 0 : 1;
            const int64_t r2 = ((int64_t)iy * nx + ix) * 2 + outside;
            double s0 = a;
            while (s0 < b) {
                const double zs = r.ez + s0 * r.dz;
                int layer;
                double s1;
                if (r.dz > 0.0) {
                    layer = tr_clampi((int)floor(zs / dzl), nz - 1);
                    s1 = ((layer + 1) * dzl - r.ez) / r.dz;
                } else if (r.dz < 0.0) {
                    layer = tr_clampi((int)ceil(zs / dzl) - 1, nz - 1);
                    s1 = (layer * dzl - r.ez) / r.dz;
                } else {
                    layer = tr_clampi((int)floor(zs / dzl), nz - 1);
                    s1 = b;
                }
                if (!(s1 > s0))
                    s1 = s0 + 1e-12;
                if (!(s1 > s0) || s1 > b)
                    s1 = b;
                const double zm = r.ez + 0.5 * (s0 + s1) * r.dz;
                double zf = (zm - layer * dzl) / dzl;
                zf = zf < 0.0 ? 0.0 : (zf >= 1.0 ? 0x1.fffffep-1 : zf);
                if (n < cap) {
                    len[n] = (float)(s1 - s0);
                    region[n] = (int32_t)(r2 * nz + layer);
                    z[n] = (float)zf;
                }
                n++;
                s0 = s1;
            }
        }
        tc = tend;
        if (tend >= r.L)
            break;
        if (tx <= ty)
            ix += sx;
        if (ty <= tx)
            iy += sy;
    }
    return n;
}
