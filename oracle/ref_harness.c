/*
 * ref_harness.c -- drives the reference's own CPU attenuation code.
 *
 * TEST INFRASTRUCTURE ONLY.  oracle/Makefile compiles this file together with
 * the reference sources src/cpu/kernel.c and src/cpu/init.c, where they lie
 * in the reference tree, into oracle/_ref/libsmocref_{exp,table}[_strict].so.
 * Nothing of the reference is copied into this repository; this file only
 * calls its entry points:
 *   set_default_input, initialize_sources      src/cpu/init.c:4-78
 *   buildExponentialTable                       src/cpu/init.c:81-117
 *   allocate_simd_vectors                       src/cpu/init.c:141-174
 *   attenuate_segment, interpolateTable         src/cpu/kernel.c:75-361
 *   get_time                                    src/cpu/init.c:194-204
 * The reference fixes ds = 0.7 and zin = 0.3 inside attenuate_segment
 * (kernel.c:100,104), so results from here are reference-constant results.
 */
#include "SimpleMOC-kernel_header.h"

#include <stdint.h>

static Input *g_I;
static Source *g_S;
static Table *g_T;
static SIMD_Vectors g_vec;

/* Set up a reference problem of G groups, R regions, F fine intervals.
 * initialize_sources fills the arrays with rand(); ref_set_regions then
 * overwrites them with caller data. */
int ref_init(int G, int R, int F)
{
    g_I = set_default_input();
    g_I->egroups = G;
    g_I->source_3D_regions = R;
    g_I->fine_axial_intervals = F;
    g_S = initialize_sources(g_I);
    g_T = buildExponentialTable(0.01, 10.0, g_I);
    g_vec = allocate_simd_vectors(g_I);
    return 0;
}

void ref_set_regions(const float *sigT, const float *fine_source, const float *fine_flux)
{
    const long G = g_I->egroups, R = g_I->source_3D_regions, F = g_I->fine_axial_intervals;
    /* initialize_sources allocates each field as one flat region-major array
     * (init.c:35-54); region 0's pointer is its base. */
    memcpy(g_S[0].sigT, sigT, (size_t)(R * G) * sizeof(float));
    memcpy(g_S[0].fine_source, fine_source, (size_t)(R * F * G) * sizeof(float));
    memcpy(g_S[0].fine_flux, fine_flux, (size_t)(R * F * G) * sizeof(float));
}

void ref_get_flux(float *fine_flux)
{
    const long G = g_I->egroups, R = g_I->source_3D_regions, F = g_I->fine_axial_intervals;
    memcpy(fine_flux, g_S[0].fine_flux, (size_t)(R * F * G) * sizeof(float));
}

/* Attenuate n segments in order with one carried state_flux vector psi[G]. */
void ref_run_segments(long n, const int32_t *region, const int32_t *fai, float *psi)
{
    for (long i = 0; i < n; i++)
        attenuate_segment(g_I, g_S, region[i], fai[i], psi, &g_vec, g_T);
}

void ref_table(float *values, int *N, float *dx, float *maxVal)
{
    *N = g_T->N;
    *dx = g_T->dx;
    *maxVal = g_T->maxVal;
    memcpy(values, g_T->values, (size_t)(2 * g_T->N) * sizeof(float));
}

float ref_interp(float x) { return interpolateTable(g_T, x); }

/* CPU baseline: the reference's attenuate_segment over tracks, threads over
 * tracks with the reference's dynamic schedule and per-(region, interval)
 * locks (kernel.c:6,43,265-281).  Region and interval ids come from the
 * caller's segment stream instead of rand_r (kernel.c:47-50).  Returns the
 * seconds between two get_time() calls around the parallel loop, the
 * reference's timed region (main.c:45-47). */
double ref_sweep(long T, const int64_t *track_off, const int32_t *region, const int32_t *fai,
                 const float *psi_in, float *psi_out, int nthreads)
{
    const int G = g_I->egroups;
#ifdef OPENMP
    omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
    const double start = get_time();
    #pragma omp parallel
    {
        SIMD_Vectors vec = allocate_simd_vectors(g_I);
        #pragma omp for schedule(dynamic, 1)
        for (long t = 0; t < T; t++) {
            float *psi = psi_out + t * G;
            memcpy(psi, psi_in + t * G, (size_t)G * sizeof(float));
            for (int64_t k = track_off[t]; k < track_off[t + 1]; k++)
                attenuate_segment(g_I, g_S, region[k], fai[k], psi, &vec, g_T);
        }
        free(vec.q0);
    }
    const double stop = get_time();
    return stop - start;
}
