/*
 * smoc_internal.h -- declarations shared by the host C code, the C-ABI
 * implementation and the HIP kernels of libsmoc.  Not installed.
 */
#ifndef SMOC_INTERNAL_H
#define SMOC_INTERNAL_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Placeholder geometry constants of src/cpu/kernel.c:99-103 (ds and zin are
 * per-segment inputs in this library). */
#define SMOC_DZ 0.1f
#define SMOC_WEIGHT 0.5f
#define SMOC_MU 0.9f
#define SMOC_MU2 0.3f

/* exp table capacity in floats (the reference table has N = 353 intervals) */
#define SMOC_TABLE_CAP 1024

/* On a region-sorted stream that can be chunked (a track is longer than
 * SMOC_RUN_SPLIT segments), a run (the stretch of one track in one region
 * interval whose tally is summed in registers) also ends before every stream
 * index that is a multiple of SMOC_RUN_SPLIT (a split point).  For a chunk
 * length C that is a multiple of SMOC_RUN_SPLIT, chunked sweeps cut tracks at
 * split points only (smoc_plan.h), so every chunk begins a run, and the run
 * partition, with it every deterministic and FIXED64 sum, is the same whether
 * the tracks are chunked or not.  A stream whose tracks are all shorter cannot
 * be chunked (a chunk at least as long as every track turns chunking off), and
 * its runs are not split.  A power of two. */
#define SMOC_RUN_SPLIT 1024

/* flags of a chunked sweep's work-list entries (entry = item id << 2 | flags) */
#define SMOC_ITEM_XFER 1 /* transfer pass of a middle chunk                     */
#define SMOC_ITEM_EMIT 2 /* sweep of a first chunk that emits its psi_out       */
#define SMOC_ITEM_LIN 3  /* linear chunk: swept from psi = 0, emits its transfer */

void smoc_set_error(const char *fmt, ...);

/* Arguments of the attenuation sweep.  One work item is (track or chunk,
 * group slice); a slice covers 64 * groups_per_lane consecutive groups. */
typedef struct {
    const void *tuples;          /* packed region data [R*F][Gp] (see pack)   */
    const float *len;            /* [S] */
    const int32_t *code;         /* [S] (region * F + f) << 2 | fit kind (0 interior,
                                    1 first, 2 last fine interval) */
    const float *zin;            /* [S] axial offset inside the fine interval */
    const int64_t *track_off;    /* [T+1] */
    const float *psi_in;         /* [T][Gp] */
    float *psi_out;              /* [T][Gp] */
    void *flux;                  /* ATOMIC_*: fine_flux [R*F][Gp]; DET_*: run partials
                                    [runs][Gp]; FIXED64: int64 fine_flux whose
                                    preceding 8 bytes count overflows */
    const int64_t *track_run_off;/* DET_*: first run id of each track [T+1]  */
    const float *table;          /* exp table, 2N floats (EXP_TABLE)          */
    const float *row_thr;        /* [R*F] EXP_ACCURATE: segments with len below it
                                    have sigT len < 1 on every group of the row */
    /* chunked sweeps (tracks split into chunks) */
    const int64_t *items;        /* NULL: item i = track i; else [n][4]: track, begin,
                                    end, first run id */
    const int64_t *item_list;    /* NULL: work entry w = item w; else entries
                                    item id << 2 | flags (1: transfer pass,
                                    2: sweep and emit psi_out as a transfer) */
    const float *chunk_psi;      /* [items][Gp] incoming psi of each chunk, or NULL */
    float *xfer;                 /* transfer pass output [items][2][Gp]: B, M */
    float *wrun;                 /* linear chunks: per run, the tally per unit of
                                    the chunk's incoming psi [runs][Gp]       */
    int64_t n_items;             /* work entries: tracks, chunks or list length */
    int xfer_mode;               /* 1: phase-1 launch, mode chosen per entry */
    int64_t *queue;              /* persistent mode: work counter, or NULL    */
    int queue_batch;             /* persistent mode: consecutive work items a
                                    wave takes per counter update (>= 1)     */
    float tab_dx, tab_max;
    int tab_n;
    int G, Gp, F, slices;        /* Gp: padded row stride (multiple of 128) */
    int split;                   /* 1: runs also end at split points (a stream
                                    that can be chunked, see SMOC_RUN_SPLIT) */
    int runs;                    /* 1: sum runs of equal (region, interval) in
                                    registers; 0: every segment is a run */
    float fix_limit;             /* FIXED64: bound on one segment's tally */
} smoc_sweep_args;

typedef struct {
    const int64_t *track_item_off; /* [T+1] first chunk of each track */
    const float *xfer;             /* [items][2][Gp] */
    const float *psi_in;           /* [T][Gp] */
    float *chunk_psi;              /* [items][Gp] */
    float *psi_out;                /* linear chunks: [T][Gp] psi_out of the tracks of
                                      more than one chunk, from the last chunk's
                                      transfer; NULL otherwise */
    int64_t T;
    int Gp, slices;
} smoc_scan_args;

/* Correction of linear chunks (k_lin_fix): for every chunk c in list (chunks
 * swept from psi = 0) and every run r of c, w chunk_psi[c] wrun[r] (w the
 * track weight) joins the run's tally: added to its partial (deterministic
 * tallies), or atomically to its (region, interval) row of fine_flux. */
typedef struct {
    const int64_t *items;          /* [n_items][4]: track, begin, end, first run id */
    const int64_t *list;           /* chunk ids << 2 */
    const float *chunk_psi;        /* [items][Gp] */
    const float *wrun;             /* [runs][Gp] */
    void *out;                     /* DET_F64: partials, double [runs][Gp]; DET_KAHAN:
                                      float2 [runs][Gp]; ATOMIC_F32 / ATOMIC_F64:
                                      fine_flux [R*F][Gp] */
    const int32_t *run_row;        /* atomic tallies: the row of each run [runs] */
    int64_t n_list, n_items, n_runs;
    int G, Gp, slices;
} smoc_fix_args;

typedef struct {
    const float *sigT;        /* [R][G] FP32 master */
    const float *fine_source; /* [R][F][G] FP32 master */
    void *tuples;             /* out: [R*F][Gp] */
    int G, Gp, R, F;
} smoc_pack_args;

typedef struct {
    const void *part;            /* run partials */
    const int64_t *fsr_run_off;  /* [R*F+1] */
    const int64_t *fsr_run_idx;  /* [n_runs] run ids grouped by (region, interval) */
    const float *init;           /* fine_flux init [R*F][Gp] or NULL */
    void *flux;                  /* out fine_flux [R*F][Gp]: f64 (DET_F64) / f32 (DET_KAHAN) */
    int64_t n_fsr;
    int G, Gp, slices;           /* lanes of groups >= G (row padding) are skipped */
    int single;                  /* partials are single FP32 tallies */
} smoc_reduce_args;

/* Source update of an outer iteration (see k_source_update). */
typedef struct {
    const void *flux;       /* fine_flux [R*F][Gp], f32 or f64 */
    const double *vol;      /* [R*F] sum of w * len per (region, interval) */
    const float *sigT;      /* [R][G] FP32 master */
    const float *scatter;   /* [R][G] c = sigma_s / sigT */
    const float *q_ext;     /* [R][F][G] external source */
    float *source;          /* [R][F][G] in: current master, out: updated */
    double *partial;        /* [blocks][2] per-block sums */
    double *sums;           /* [2] sum (q_new - q_old)^2, sum q_old^2 */
    int G, Gp, R, F;
} smoc_update_args;

/* kernel launchers (smoc_kernels.hip); return 0 or a hipError_t value */
int smoc_launch_pack(const smoc_pack_args *a, int storage, void *stream);
int smoc_launch_row_thr(const void *tuples, float *thr, int64_t nrow, int G, int Gp,
                        int storage, void *stream);
int smoc_launch_reset(void *flux, const float *init, int64_t n, int f64, void *stream);
int smoc_launch_sweep(const smoc_sweep_args *a, int storage, int gpl, int exp_mode, int tally,
                      int waves_per_block, int grid_blocks, void *stream);
int smoc_launch_reduce(const smoc_reduce_args *a, int tally, int gpl, int waves_per_block,
                       void *stream);
int smoc_launch_scan(const smoc_scan_args *a, int gpl, void *stream);
int smoc_launch_lin_fix(const smoc_fix_args *a, int tally, int gpl, void *stream);
int smoc_launch_zero_i64(int64_t *p, int64_t n, void *stream);
int64_t smoc_update_blocks(int64_t n);
int smoc_launch_source_update(const smoc_update_args *a, int flux_f64, void *stream);
int smoc_launch_encode_stream(float *zin_io, int32_t *code_io, int F, int64_t S, void *stream);
/* chunk plan of a chunked sweep on the device (k_plan): items [n][4], the
 * phase-1 / phase-2 work lists and track_item_off [T+1], for chunk length
 * chunk >= 1; the sizes follow from track_off as for the host plan.  Each
 * item's first run id is its first segment, or, with track_run_off and
 * split_run given (deterministic tally of a region-sorted stream; chunk a
 * multiple of SMOC_RUN_SPLIT), the id of the run that begins there */
int smoc_launch_plan(const int64_t *track_off, int64_t T, int chunk, int64_t *items,
                     int64_t *list1, int64_t *list2, int64_t *track_item_off,
                     const int64_t *track_run_off, const int64_t *split_run, void *stream);
int smoc_launch_fixed_to_f64(int64_t *flux, const float *init, int64_t n, void *stream);
/* plain copy of n16 16-byte elements, kernel variant 0..SMOC_COPY_VARIANTS-1 */
int smoc_launch_copy(const void *src, void *dst, int64_t n16, int variant, int cus,
                     void *stream);
/* k_tally_probe over S segments into [rows][Gp] FP32; grid = ceil(ceil(S / 100) / 4) */
int smoc_launch_tally_probe(void *flux, int64_t S, int rows, int G, int Gp, int grid,
                            void *stream);
/* bytes per packed tuple for a storage type */
int smoc_tuple_bytes(int storage);

#ifdef __cplusplus
}
#endif
#endif
