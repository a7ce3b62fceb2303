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
            q2 * mu2 * (tau * (tau * (tau - 3.0) + 6.0) - 6.0 * expVal) /
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
