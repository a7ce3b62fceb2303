/*
 * smoc_fixture.c -- binary fixture files for problems and results (host C).
 *
 * Layout (little-endian, no padding between fields):
 *   problem: "SMOCPRB1" | u32 version=1 | i32 G R F | i64 S T | u64 seed |
 *            u32 flags | u64 hash | f32 sigT[R*G] | f32 fine_source[R*F*G] |
 *            f32 len[S] | i32 region[S] | f32 z[S] | i64 track_off[T+1] |
 *            f32 psi_in[T*G]
 *   results: "SMOCRES1" | u32 version=1 | i32 G R F | i64 T |
 *            i32 storage exp_mode tally | f64 fine_flux_checksum
 *            state_flux_checksum | u64 hash | f64 fine_flux[R*F*G] |
 *            f32 psi_out[T*G]
 * hash: 64-bit FNV-1a over the header (with its hash field zero) followed by
 * the array bytes in file order, so a changed shape field is caught too.
 */
#include "smoc.h"
#include "smoc_internal.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define FIX_VERSION 1u
static const char PRB_MAGIC[8] = {'S', 'M', 'O', 'C', 'P', 'R', 'B', '1'};
static const char RES_MAGIC[8] = {'S', 'M', 'O', 'C', 'R', 'E', 'S', '1'};

typedef struct {
    char magic[8];
    uint32_t version;
    int32_t G, R, F;
    int64_t S, T;
    uint64_t seed;
    uint32_t flags;
    uint64_t hash;
} __attribute__((packed)) prb_header;

typedef struct {
    char magic[8];
    uint32_t version;
    int32_t G, R, F;
    int64_t T;
    int32_t storage, exp_mode, tally;
    double ck_fine, ck_state;
    uint64_t hash;
} __attribute__((packed)) res_header;

static uint64_t fnv1a(uint64_t h, const void *p, size_t n)
{
    const unsigned char *b = (const unsigned char *)p;
    for (size_t i = 0; i < n; i++) {
        h ^= b[i];
        h *= 1099511628211ULL;
    }
    return h;
}
#define FNV_INIT 14695981039346656037ULL

typedef struct {
    const void *p;
    size_t n;  /* bytes */
} span;

static int write_file(const char *path, const void *hdr, size_t hdr_bytes, const span *a, int na)
{
    FILE *f = fopen(path, "wb");
    if (!f) {
        smoc_set_error("cannot open %s for writing", path);
        return SMOC_ERR_ARG;
    }
    int ok = fwrite(hdr, 1, hdr_bytes, f) == hdr_bytes;
    for (int i = 0; ok && i < na; i++)
        ok = a[i].n == 0 || fwrite(a[i].p, 1, a[i].n, f) == a[i].n;
    ok = (fclose(f) == 0) && ok;
    if (!ok) {
        smoc_set_error("write to %s failed", path);
        return SMOC_ERR_ARG;
    }
    return SMOC_OK;
}

/* read the arrays after a header, verifying size and hash; h0 is the hash of
 * the header with its hash field zero */
static int read_arrays(FILE *f, const char *path, uint64_t h0, uint64_t want_hash, span *a,
                       int na)
{
    uint64_t h = h0;
    for (int i = 0; i < na; i++) {
        if (a[i].n && fread((void *)a[i].p, 1, a[i].n, f) != a[i].n) {
            smoc_set_error("%s is truncated", path);
            return SMOC_ERR_ARG;
        }
        h = fnv1a(h, a[i].p, a[i].n);
    }
    if (fgetc(f) != EOF) {
        smoc_set_error("%s has trailing bytes", path);
        return SMOC_ERR_ARG;
    }
    if (h != want_hash) {
        smoc_set_error("%s: array hash mismatch (file corrupt)", path);
        return SMOC_ERR_ARG;
    }
    return SMOC_OK;
}

#define FIX_MAX_COUNT (INT64_C(1) << 40) /* bound on S and T; keeps every byte count in range */

/* the file's size must be exactly what its header implies; checked before
 * any array is read, so a damaged header cannot make a reader size its
 * buffers from it */
static int size_ok(FILE *f, const char *path, uint64_t want)
{
    const long here = ftell(f);
    if (here < 0 || fseek(f, 0, SEEK_END) != 0) {
        smoc_set_error("%s: cannot determine the file size", path);
        return SMOC_ERR_ARG;
    }
    const long end = ftell(f);
    if (end < 0 || fseek(f, here, SEEK_SET) != 0) {
        smoc_set_error("%s: cannot determine the file size", path);
        return SMOC_ERR_ARG;
    }
    if ((uint64_t)end < want) {
        smoc_set_error("%s is truncated (%ld bytes, header implies %llu)", path, end,
                       (unsigned long long)want);
        return SMOC_ERR_ARG;
    }
    if ((uint64_t)end > want) {
        smoc_set_error("%s has trailing bytes (%ld bytes, header implies %llu)", path, end,
                       (unsigned long long)want);
        return SMOC_ERR_ARG;
    }
    return SMOC_OK;
}

static uint64_t prb_bytes(const prb_header *h)
{
    const uint64_t G = (uint64_t)h->G, R = (uint64_t)h->R, F = (uint64_t)h->F;
    const uint64_t S = (uint64_t)h->S, T = (uint64_t)h->T;
    return sizeof(*h) + 4 * (R * G + R * F * G + 3 * S + T * G) + 8 * (T + 1);
}

static uint64_t res_bytes(const res_header *h)
{
    const uint64_t G = (uint64_t)h->G, R = (uint64_t)h->R, F = (uint64_t)h->F;
    const uint64_t T = (uint64_t)h->T;
    return sizeof(*h) + 8 * R * F * G + 4 * T * G;
}

static int shape_ok(int64_t G, int64_t R, int64_t F, int64_t T)
{
    return G >= 1 && G <= 65536 && R >= 1 && F >= 1 && T >= 0 && T <= FIX_MAX_COUNT &&
           R * F <= INT32_MAX;
}

/* hash of a header whose hash field is zero */
static uint64_t header_hash(const void *hdr, size_t bytes)
{
    return fnv1a(FNV_INIT, hdr, bytes);
}

int smoc_problem_write(const char *path, const smoc_problem_info *in, const float *sigT,
                       const float *fine_source, const float *len, const int32_t *region,
                       const float *z, const int64_t *track_off, const float *psi_in)
{
    if (!path || !in || !sigT || !fine_source || !track_off || !shape_ok(in->G, in->R, in->F, in->T) ||
        in->S < 0 || in->S > FIX_MAX_COUNT || (in->S > 0 && (!len || !region || !z)) || (in->T > 0 && !psi_in) ||
        track_off[0] != 0 || track_off[in->T] != in->S) {
        smoc_set_error("smoc_problem_write: bad argument");
        return SMOC_ERR_ARG;
    }
    const int64_t G = in->G, R = in->R, F = in->F, S = in->S, T = in->T;
    span a[7] = {{sigT, (size_t)(R * G) * 4},  {fine_source, (size_t)(R * F * G) * 4},
                 {len, (size_t)S * 4},         {region, (size_t)S * 4},
                 {z, (size_t)S * 4},           {track_off, (size_t)(T + 1) * 8},
                 {psi_in, (size_t)(T * G) * 4}};
    prb_header h;
    memset(&h, 0, sizeof(h));
    memcpy(h.magic, PRB_MAGIC, 8);
    h.version = FIX_VERSION;
    h.G = (int32_t)G, h.R = (int32_t)R, h.F = (int32_t)F, h.S = S, h.T = T;
    h.seed = in->seed;
    h.flags = in->flags;
    h.hash = header_hash(&h, sizeof(h));
    for (int i = 0; i < 7; i++)
        h.hash = fnv1a(h.hash, a[i].p, a[i].n);
    return write_file(path, &h, sizeof(h), a, 7);
}

static int read_prb_header(FILE *f, const char *path, prb_header *h)
{
    if (fread(h, 1, sizeof(*h), f) != sizeof(*h) || memcmp(h->magic, PRB_MAGIC, 8) != 0) {
        smoc_set_error("%s is not a smoc problem file", path);
        return SMOC_ERR_ARG;
    }
    if (h->version != FIX_VERSION || !shape_ok(h->G, h->R, h->F, h->T) || h->S < 0 ||
        h->S > FIX_MAX_COUNT) {
        smoc_set_error("%s: unsupported version %u or bad shape", path, h->version);
        return SMOC_ERR_ARG;
    }
    return size_ok(f, path, prb_bytes(h));
}

int smoc_problem_read_info(const char *path, smoc_problem_info *info)
{
    if (!path || !info) {
        smoc_set_error("smoc_problem_read_info: NULL argument");
        return SMOC_ERR_ARG;
    }
    FILE *f = fopen(path, "rb");
    if (!f) {
        smoc_set_error("cannot open %s", path);
        return SMOC_ERR_ARG;
    }
    prb_header h;
    int rc = read_prb_header(f, path, &h);
    fclose(f);
    if (rc)
        return rc;
    info->G = h.G, info->R = h.R, info->F = h.F, info->S = h.S, info->T = h.T;
    info->seed = h.seed;
    info->flags = h.flags;
    return SMOC_OK;
}

int smoc_problem_read(const char *path, float *sigT, float *fine_source, float *len,
                      int32_t *region, float *z, int64_t *track_off, float *psi_in)
{
    if (!path || !sigT || !fine_source || !track_off) {
        smoc_set_error("smoc_problem_read: NULL argument");
        return SMOC_ERR_ARG;
    }
    FILE *f = fopen(path, "rb");
    if (!f) {
        smoc_set_error("cannot open %s", path);
        return SMOC_ERR_ARG;
    }
    prb_header h;
    int rc = read_prb_header(f, path, &h);
    if (rc == SMOC_OK) {
        const int64_t G = h.G, R = h.R, F = h.F, S = h.S, T = h.T;
        if ((S > 0 && (!len || !region || !z)) || (T > 0 && !psi_in)) {
            smoc_set_error("smoc_problem_read: NULL array for a non-empty problem");
            rc = SMOC_ERR_ARG;
        } else {
            span a[7] = {{sigT, (size_t)(R * G) * 4},  {fine_source, (size_t)(R * F * G) * 4},
                         {len, (size_t)S * 4},         {region, (size_t)S * 4},
                         {z, (size_t)S * 4},           {track_off, (size_t)(T + 1) * 8},
                         {psi_in, (size_t)(T * G) * 4}};
            prb_header z0 = h;
            z0.hash = 0;
            rc = read_arrays(f, path, header_hash(&z0, sizeof(z0)), h.hash, a, 7);
        }
    }
    fclose(f);
    return rc;
}

int smoc_results_write(const char *path, const smoc_results_info *in, const double *fine_flux,
                       const float *psi_out)
{
    if (!path || !in || !fine_flux || !shape_ok(in->G, in->R, in->F, in->T) ||
        (in->T > 0 && !psi_out)) {
        smoc_set_error("smoc_results_write: bad argument");
        return SMOC_ERR_ARG;
    }
    const int64_t G = in->G, R = in->R, F = in->F, T = in->T;
    span a[2] = {{fine_flux, (size_t)(R * F * G) * 8}, {psi_out, (size_t)(T * G) * 4}};
    res_header h;
    memset(&h, 0, sizeof(h));
    memcpy(h.magic, RES_MAGIC, 8);
    h.version = FIX_VERSION;
    h.G = (int32_t)G, h.R = (int32_t)R, h.F = (int32_t)F, h.T = T;
    h.storage = in->storage, h.exp_mode = in->exp_mode, h.tally = in->tally;
    h.ck_fine = smoc_checksum_f64(fine_flux, R * F * G);
    h.ck_state = T > 0 ? smoc_checksum_f32(psi_out, T * G) : 0.0;
    h.hash = header_hash(&h, sizeof(h));
    for (int i = 0; i < 2; i++)
        h.hash = fnv1a(h.hash, a[i].p, a[i].n);
    return write_file(path, &h, sizeof(h), a, 2);
}

static int read_res_header(FILE *f, const char *path, res_header *h)
{
    if (fread(h, 1, sizeof(*h), f) != sizeof(*h) || memcmp(h->magic, RES_MAGIC, 8) != 0) {
        smoc_set_error("%s is not a smoc results file", path);
        return SMOC_ERR_ARG;
    }
    if (h->version != FIX_VERSION || !shape_ok(h->G, h->R, h->F, h->T)) {
        smoc_set_error("%s: unsupported version %u or bad shape", path, h->version);
        return SMOC_ERR_ARG;
    }
    return size_ok(f, path, res_bytes(h));
}

int smoc_results_read_info(const char *path, smoc_results_info *info)
{
    if (!path || !info) {
        smoc_set_error("smoc_results_read_info: NULL argument");
        return SMOC_ERR_ARG;
    }
    FILE *f = fopen(path, "rb");
    if (!f) {
        smoc_set_error("cannot open %s", path);
        return SMOC_ERR_ARG;
    }
    res_header h;
    int rc = read_res_header(f, path, &h);
    fclose(f);
    if (rc)
        return rc;
    info->G = h.G, info->R = h.R, info->F = h.F, info->T = h.T;
    info->storage = h.storage, info->exp_mode = h.exp_mode, info->tally = h.tally;
    info->fine_flux_checksum = h.ck_fine;
    info->state_flux_checksum = h.ck_state;
    return SMOC_OK;
}

int smoc_results_read(const char *path, double *fine_flux, float *psi_out)
{
    if (!path || !fine_flux) {
        smoc_set_error("smoc_results_read: NULL argument");
        return SMOC_ERR_ARG;
    }
    FILE *f = fopen(path, "rb");
    if (!f) {
        smoc_set_error("cannot open %s", path);
        return SMOC_ERR_ARG;
    }
    res_header h;
    int rc = read_res_header(f, path, &h);
    if (rc == SMOC_OK) {
        if (h.T > 0 && !psi_out) {
            smoc_set_error("smoc_results_read: NULL psi_out for T = %lld", (long long)h.T);
            rc = SMOC_ERR_ARG;
        } else {
            span a[2] = {{fine_flux, (size_t)((int64_t)h.R * h.F * h.G) * 8},
                         {psi_out, (size_t)(h.T * h.G) * 4}};
            res_header z0 = h;
            z0.hash = 0;
            rc = read_arrays(f, path, header_hash(&z0, sizeof(z0)), h.hash, a, 2);
        }
    }
    fclose(f);
    return rc;
}

int smoc_results_compare(const char *path, const double *fine_flux, const float *psi_out,
                         double *fine_checksum_rel, double *state_checksum_rel,
                         double *fine_normwise, double *state_normwise)
{
    smoc_results_info info;
    int rc = smoc_results_read_info(path, &info);
    if (rc)
        return rc;
    if (!fine_flux || (info.T > 0 && !psi_out)) {
        smoc_set_error("smoc_results_compare: NULL results");
        return SMOC_ERR_ARG;
    }
    const int64_t nf = (int64_t)info.R * info.F * info.G, np = info.T * info.G;
    double *rf = (double *)malloc((size_t)(nf > 0 ? nf : 1) * sizeof(double));
    float *rp = (float *)malloc((size_t)(np > 0 ? np : 1) * sizeof(float));
    if (!rf || !rp) {
        free(rf), free(rp);
        smoc_set_error("smoc_results_compare: host allocation failed");
        return SMOC_ERR_NOMEM;
    }
    rc = smoc_results_read(path, rf, rp);
    if (rc == SMOC_OK) {
        double sf = 0, sp = 0, df = 0, dp = 0, nfr = 0, npr = 0;
        for (int64_t i = 0; i < nf; i++) {
            sf += fine_flux[i];
            df += (fine_flux[i] - rf[i]) * (fine_flux[i] - rf[i]);
            nfr += rf[i] * rf[i];
        }
        for (int64_t i = 0; i < np; i++) {
            sp += psi_out[i];
            dp += ((double)psi_out[i] - rp[i]) * ((double)psi_out[i] - rp[i]);
            npr += (double)rp[i] * rp[i];
        }
        if (fine_checksum_rel)
            *fine_checksum_rel = fabs(sf - info.fine_flux_checksum) /
                                 fmax(fabs(info.fine_flux_checksum), 1e-300);
        if (state_checksum_rel)
            *state_checksum_rel = fabs(sp - info.state_flux_checksum) /
                                  fmax(fabs(info.state_flux_checksum), 1e-300);
        if (fine_normwise)
            *fine_normwise = nfr > 0 ? sqrt(df / nfr) : sqrt(df);
        if (state_normwise)
            *state_normwise = npr > 0 ? sqrt(dp / npr) : sqrt(dp);
    }
    free(rf), free(rp);
    return rc;
}
