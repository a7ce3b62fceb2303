/*
 * smoc_host.c -- host-side parts of libsmoc: reference defaults, the
 * exponential table, segment geometry decode, the seeded input generator,
 * the (region, interval) sort and the result checksums.  Pure C, no GPU.
 *
 * Compiled without fast-math and with -ffp-contract=off: the generator and
 * smoc_decode_z must produce the same floats as the device code and as any
 * restatement of them (see tests/test_host.py).
 */
#include "smoc.h"
#include "smoc_internal.h"

#include <math.h>
#include <omp.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* ---- errors ------------------------------------------------------------- */
static __thread char g_err[512];

void smoc_set_error(const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

const char *smoc_last_error(void) { return g_err; }

int smoc_abi_version(void) { return SMOC_ABI_VERSION; }

const char *smoc_strerror(int code)
{
    switch (code) {
    case SMOC_OK: return "ok";
    case SMOC_ERR_ARG: return "invalid argument";
    case SMOC_ERR_NODEV: return "no usable HIP device";
    case SMOC_ERR_HIP: return "HIP runtime error";
    case SMOC_ERR_STATE: return "call out of order";
    case SMOC_ERR_NOMEM: return "out of memory";
    case SMOC_ERR_RANGE: return "fixed-point tally out of range";
    default: return "unknown error";
    }
}

/* ---- reference defaults (src/cpu/init.c:4-24, src/cuda/init.cu:30-44) ---- */
int smoc_derive_regions(smoc_input *in)
{
    if (!in || in->decomp_assemblies_ax <= 0) {
        smoc_set_error("smoc_derive_regions: decomp_assemblies_ax must be > 0");
        return SMOC_ERR_ARG;
    }
    /* src/cpu/main.c:18-19 */
    in->source_3D_regions = (int)ceil((double)in->source_2D_regions *
                                      in->coarse_axial_intervals / in->decomp_assemblies_ax);
    return SMOC_OK;
}

int smoc_default_input(smoc_input *in)
{
    if (!in) {
        smoc_set_error("smoc_default_input: NULL");
        return SMOC_ERR_ARG;
    }
    in->source_2D_regions = 5000;
    in->coarse_axial_intervals = 27;
    in->fine_axial_intervals = 5;
    in->decomp_assemblies_ax = 20;
    in->segments = 50000000;
    in->egroups = 128;
    in->seg_per_block = 100;
    return smoc_derive_regions(in);
}

/* ---- exponential table (src/cpu/init.c:81-117) ---------------------------- */
int smoc_build_exp_table(float precision, float maxVal, float *values, int capacity, int *N_out,
                         float *dx_out, float *maxVal_out)
{
    if (!(precision > 0.0f) || !(maxVal > 0.0f) || !values || !N_out || !dx_out || !maxVal_out) {
        smoc_set_error("smoc_build_exp_table: bad argument");
        return SMOC_ERR_ARG;
    }
    const int N = (int)(maxVal * sqrt(1.0 / (8.0 * precision * 0.01)));
    if (N <= 0 || 2 * N > capacity) {
        smoc_set_error("smoc_build_exp_table: table of %d intervals does not fit %d floats", N,
                       capacity);
        return SMOC_ERR_ARG;
    }
    const float dx = maxVal / (float)N;
    for (int n = 0; n < N; n++) {
        /* exp() in double of the float n*dx, stored as float; the slope keeps
         * the reference's sign (-exp(-n dx)). */
        const float e = (float)exp(-n * dx);
        values[2 * n] = -e;
        values[2 * n + 1] = 1 + (n * dx - 1) * e;
    }
    *N_out = N;
    *dx_out = dx;
    *maxVal_out = maxVal - dx;
    return SMOC_OK;
}

float smoc_exp_table_lookup(const float *values, float dx, float maxVal, float x)
{
    if (x > maxVal)
        return 1.0f;
    const int i = 2 * (int)(x / dx + 0.5f * dx); /* src/cpu/kernel.c:344,355 */
    return values[i] * x + values[i + 1];
}

/* ---- segment geometry --------------------------------------------------- */
void smoc_decode_z(float z, int F, int *f, float *zin)
{
    const float zf = z * (float)F;
    int fi = (int)zf;
    if (fi > F - 1)
        fi = F - 1;
    if (fi < 0)
        fi = 0;
    *f = fi;
    *zin = ((zf - (float)fi) - 0.5f) * SMOC_DZ;
}

/* ---- counter-based generator --------------------------------------------- */
/* splitmix64's output function applied to a Weyl sequence: draw(key, i) =
 * mix(key + (i+1) * golden).  Each array has its own key = mix(seed * golden
 * + stream id). */
#define SMOC_GOLDEN 0x9E3779B97F4A7C15ULL

static inline uint64_t mix64(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

static inline uint64_t stream_key(uint64_t seed, uint64_t stream)
{
    return mix64(seed * SMOC_GOLDEN + stream);
}

static inline uint64_t draw(uint64_t key, uint64_t i) { return mix64(key + (i + 1) * SMOC_GOLDEN); }

/* top 24 bits -> [0, 1), exact in FP32 */
static inline float u01(uint64_t x) { return (float)(x >> 40) * 0x1p-24f; }

static inline float uab(uint64_t x, float a, float b) { return a + (b - a) * u01(x); }

/* Lemire multiply-shift on the top 32 bits -> [0, n) */
static inline int32_t below(uint64_t x, uint32_t n)
{
    return (int32_t)(((x >> 32) * (uint64_t)n) >> 32);
}

enum {
    ST_SIGT = 1,
    ST_SOURCE = 2,
    ST_LEN = 3,
    ST_REGION = 4,
    ST_Z = 5,
    ST_PSI = 6,
    ST_TRACK = 7,
    ST_SCATTER = 8
};

int smoc_generate_regions(uint64_t seed, int G, int R, int F, float *sigT, float *fine_source)
{
    if (G <= 0 || R <= 0 || F <= 0 || !sigT || !fine_source) {
        smoc_set_error("smoc_generate_regions: bad argument");
        return SMOC_ERR_ARG;
    }
    const uint64_t ks = stream_key(seed, ST_SIGT), kq = stream_key(seed, ST_SOURCE);
    const int64_t nsig = (int64_t)R * G, nsrc = (int64_t)R * F * G;
    #pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < nsig; i++)
        sigT[i] = uab(draw(ks, (uint64_t)i), 0.1f, 10.0f);
    #pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < nsrc; i++)
        fine_source[i] = u01(draw(kq, (uint64_t)i));
    return SMOC_OK;
}

int smoc_generate_segments(uint64_t seed, int64_t S, int R, float *len, int32_t *region, float *z)
{
    if (S < 0 || R <= 0 || (S > 0 && (!len || !region || !z))) {
        smoc_set_error("smoc_generate_segments: bad argument");
        return SMOC_ERR_ARG;
    }
    const uint64_t kl = stream_key(seed, ST_LEN), kr = stream_key(seed, ST_REGION),
                   kz = stream_key(seed, ST_Z);
    #pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < S; i++) {
        len[i] = uab(draw(kl, (uint64_t)i), 0.0f, 0.1f);
        region[i] = below(draw(kr, (uint64_t)i), (uint32_t)R);
        z[i] = u01(draw(kz, (uint64_t)i));
    }
    return SMOC_OK;
}

int smoc_generate_psi(uint64_t seed, int64_t T, int G, float *psi_in)
{
    if (T < 0 || G <= 0 || (T > 0 && !psi_in)) {
        smoc_set_error("smoc_generate_psi: bad argument");
        return SMOC_ERR_ARG;
    }
    const uint64_t kp = stream_key(seed, ST_PSI);
    const int64_t n = T * G;
    #pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++)
        psi_in[i] = u01(draw(kp, (uint64_t)i));
    return SMOC_OK;
}

int smoc_generate_scatter(uint64_t seed, int R, int G, float cmax, float *scatter)
{
    if (R <= 0 || G <= 0 || !scatter || !(cmax >= 0.f && cmax < 1.f)) {
        smoc_set_error("smoc_generate_scatter: bad argument (cmax must be in [0, 1))");
        return SMOC_ERR_ARG;
    }
    const uint64_t ks = stream_key(seed, ST_SCATTER);
    const int64_t n = (int64_t)R * G;
    #pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++)
        scatter[i] = cmax * u01(draw(ks, (uint64_t)i));
    return SMOC_OK;
}

int smoc_tracks_fixed(int64_t S, int L, int64_t *track_off, int64_t *T_out)
{
    if (S < 0 || L <= 0 || !track_off || !T_out) {
        smoc_set_error("smoc_tracks_fixed: bad argument");
        return SMOC_ERR_ARG;
    }
    const int64_t T = (S + L - 1) / L;
    for (int64_t t = 0; t < T; t++)
        track_off[t] = t * L;
    track_off[T] = S;
    *T_out = T;
    return SMOC_OK;
}

int smoc_tracks_lognormal(uint64_t seed, int64_t T, double median, double sigma, int lo, int hi,
                          int64_t *track_off, int64_t *S_out)
{
    if (T < 0 || !(median > 0.0) || !(sigma >= 0.0) || lo < 1 || hi < lo || !track_off ||
        !S_out) {
        smoc_set_error("smoc_tracks_lognormal: bad argument");
        return SMOC_ERR_ARG;
    }
    const uint64_t kt = stream_key(seed, ST_TRACK);
    const double mu = log(median);
    int64_t acc = 0;
    track_off[0] = 0;
    for (int64_t t = 0; t < T; t++) {
        /* Box-Muller on two 53-bit uniforms, u1 in (0, 1] */
        const uint64_t x1 = draw(kt, (uint64_t)(2 * t)), x2 = draw(kt, (uint64_t)(2 * t + 1));
        const double u1 = (double)((x1 >> 11) + 1) * 0x1p-53;
        const double u2 = (double)(x2 >> 11) * 0x1p-53;
        const double n = sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
        double L = floor(exp(mu + sigma * n) + 0.5);
        if (L < lo)
            L = lo;
        if (L > hi)
            L = hi;
        acc += (int64_t)L;
        track_off[t + 1] = acc;
    }
    *S_out = acc;
    return SMOC_OK;
}

int smoc_sort_segments(int64_t S, int R, int F, float *len, int32_t *region, float *z)
{
    if (S < 0 || R <= 0 || F < 2 || (S > 0 && (!len || !region || !z))) {
        smoc_set_error("smoc_sort_segments: bad argument");
        return SMOC_ERR_ARG;
    }
    const int64_t K = (int64_t)R * F;
    int64_t *cnt = (int64_t *)calloc((size_t)K + 1, sizeof(int64_t));
    int32_t *key = (int32_t *)malloc((size_t)(S > 0 ? S : 1) * sizeof(int32_t));
    float *tl = (float *)malloc((size_t)(S > 0 ? S : 1) * sizeof(float));
    float *tz = (float *)malloc((size_t)(S > 0 ? S : 1) * sizeof(float));
    int32_t *tr = (int32_t *)malloc((size_t)(S > 0 ? S : 1) * sizeof(int32_t));
    if (!cnt || !key || !tl || !tz || !tr) {
        free(cnt), free(key), free(tl), free(tz), free(tr);
        smoc_set_error("smoc_sort_segments: host allocation of %lld segments failed",
                       (long long)S);
        return SMOC_ERR_NOMEM;
    }
    for (int64_t i = 0; i < S; i++) {
        if (region[i] < 0 || region[i] >= R) {
            free(cnt), free(key), free(tl), free(tz), free(tr);
            smoc_set_error("smoc_sort_segments: region[%lld] = %d outside [0, %d)",
                           (long long)i, region[i], R);
            return SMOC_ERR_ARG;
        }
        int f;
        float zin;
        smoc_decode_z(z[i], F, &f, &zin);
        key[i] = region[i] * F + f;
        cnt[key[i] + 1]++;
    }
    for (int64_t k = 0; k < K; k++)
        cnt[k + 1] += cnt[k];
    for (int64_t i = 0; i < S; i++) {
        const int64_t d = cnt[key[i]]++;
        tl[d] = len[i];
        tr[d] = region[i];
        tz[d] = z[i];
    }
    memcpy(len, tl, (size_t)S * sizeof(float));
    memcpy(region, tr, (size_t)S * sizeof(int32_t));
    memcpy(z, tz, (size_t)S * sizeof(float));
    free(cnt), free(key), free(tl), free(tz), free(tr);
    return SMOC_OK;
}

int smoc_set_host_threads(int n)
{
    if (n > 0)
        omp_set_num_threads(n);
    return SMOC_OK;
}

/* ---- checksums ------------------------------------------------------------ */
double smoc_checksum_f64(const double *x, int64_t n)
{
    double s = 0.0;
    for (int64_t i = 0; i < n; i++)
        s += x[i];
    return s;
}

double smoc_checksum_f32(const float *x, int64_t n)
{
    double s = 0.0;
    for (int64_t i = 0; i < n; i++)
        s += (double)x[i];
    return s;
}
