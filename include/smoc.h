/*
 * smoc.h -- C ABI of libsmoc.so, the MI355X (gfx950) segment-attenuation
 * library for SimpleMOC-kernel's hot path.
 *
 * Plain C: no HIP, torch or CUDA types cross this boundary.  Host buffers are
 * caller-owned and copied in; device buffers belong to the context.  Every
 * function returns SMOC_OK (0) or a negative SMOC_ERR_* code and never
 * exits the process; smoc_last_error() gives the message of the most recent
 * failure on the calling thread.  One context per host thread at a time; a
 * context owns one HIP stream.
 *
 * Reference interfaces replaced (ANL-CESAR/SimpleMOC-kernel, Version 4):
 *   run_kernel(Input*, Source*, Table*)            src/cpu/SimpleMOC-kernel_header.h:81
 *       and __global__ run_kernel(Input, Source*, Source_Arrays, Table*,
 *       curandState*, float*, int)                 src/cuda/SimpleMOC-kernel_header.h:66-68
 *       launched at src/cuda/main.cu:90          -> smoc_run
 *   attenuate_segment(...)                        src/cpu/SimpleMOC-kernel_header.h:82-84
 *                                                 -> the sweep kernels behind smoc_run
 *   initialize_sources / initialize_device_sources
 *                                                 src/cpu/init.c:26-78, src/cuda/init.cu:69-127
 *                                                 -> smoc_set_regions (+ smoc_generate_regions)
 *   the segment draws QSR_id/FAI_id = rand_r()    src/cpu/kernel.c:47-50, src/cuda/kernel.cu:52-60
 *                                                 -> smoc_set_segments (+ smoc_generate_segments)
 *   buildExponentialTable / interpolateTable      src/cpu/init.c:81-117, src/cpu/kernel.c:337-361
 *                                                 -> smoc_build_exp_table, SMOC_EXP_TABLE
 *   set_default_input + derived region count     src/cpu/init.c:4-24, src/cpu/main.c:18-19
 *                                                 -> smoc_default_input
 */
#ifndef SMOC_H
#define SMOC_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SMOC_ABI_VERSION 2

/* ---- error codes ------------------------------------------------------ */
#define SMOC_OK 0
#define SMOC_ERR_ARG (-1)         /* invalid argument or unsupported combination */
#define SMOC_ERR_NODEV (-2)       /* no usable HIP device */
#define SMOC_ERR_HIP (-3)         /* a HIP runtime call failed */
#define SMOC_ERR_STATE (-4)       /* call out of order (e.g. run before set_*) */
#define SMOC_ERR_NOMEM (-5)       /* device or host allocation failed */
#define SMOC_ERR_RANGE (-6)       /* a FIXED64 tally left its fixed-point range */

/* ---- variants ---------------------------------------------------------- */
/* Storage of sigT and fine_source in HBM; arithmetic is FP32 in every case. */
#define SMOC_STORE_F32 0
#define SMOC_STORE_BF16 1
#define SMOC_STORE_F16 2

/* ev = 1 - exp(-tau), as in src/cpu/kernel.c:216-223:
 *   NATIVE    1 - exp(-tau) with the hardware exponential (literal form);
 *   TABLE     the reference interpolation table (smoc_build_exp_table(0.01,
 *             10), src/cpu/main.c:33, lookup src/cpu/kernel.c:337-361);
 *   ACCURATE  native exponential, but for tau < 1 ev, tau - ev and the
 *             quadratic-source polynomial tau^3 - 3tau^2 + 6tau - 6ev come
 *             from their Taylor series in tau, which removes the
 *             cancellation the literal FP32 form suffers for small tau. */
#define SMOC_EXP_NATIVE 0
#define SMOC_EXP_TABLE 1
#define SMOC_EXP_ACCURATE 2

/* fine_flux tally.
 *   ATOMIC_F32  FP32 device atomics into an FP32 fine_flux (reference CUDA
 *               src/cuda/kernel.cu:181); order-dependent rounding.
 *   ATOMIC_F64  FP64 device atomics into an FP64 fine_flux.
 *   DET_F64     deterministic: each run of consecutive segments of a track
 *               in one (region, interval) is summed in FP64 registers (on a
 *               region-sorted stream, FP32 sums of blocks of up to 8
 *               segments join the FP64 run sum) and stored; a segmented
 *               reduction then sums the runs of every (region, interval) in
 *               a fixed order.  Bit-exact run to run and across launch
 *               shapes.  On a region-sorted stream with a track longer than
 *               1024 segments a run also ends before every stream index that
 *               is a multiple of 1024, and chunked sweeps need a chunk length
 *               that is a multiple of 1024 (others would split runs:
 *               SMOC_ERR_ARG); tracks are then cut at such indices only.
 *   DET_KAHAN   as DET_F64 with compensated (Kahan) FP32 sums (of the
 *               FP32 block sums, on a region-sorted stream).
 *   FIXED64     deterministic: each run is rounded to a 64-bit fixed-point
 *               integer (2^-32 resolution, |fine_flux| < 2^31) and added with
 *               integer atomics, which are associative, so the result is
 *               bit-exact run to run in any order; converted to FP64 after
 *               the sweep.  Every contribution (on a region-sorted stream,
 *               every FP32 sum of a block of up to 8 segments, against 8x the
 *               bound) is checked against a per-contribution bound that keeps
 *               the sums in range; smoc_run fails with SMOC_ERR_RANGE if one
 *               exceeded it.  Chunk lengths on a region-sorted stream as for
 *               DET_F64. */
#define SMOC_TALLY_ATOMIC_F32 0
#define SMOC_TALLY_ATOMIC_F64 1
#define SMOC_TALLY_DET_F64 2
#define SMOC_TALLY_DET_KAHAN 3
#define SMOC_TALLY_FIXED64 4

typedef struct smoc_ctx smoc_ctx;

typedef struct {
    int device;      /* HIP device ordinal (reference CUDA -d, src/cuda/io.cu:149-158) */
    int egroups;     /* G, energy groups (-e)                                  */
    int regions;     /* R, 3D source regions                                    */
    int fine_axial;  /* F, fine axial intervals per region (>= 2)               */
    int storage;     /* SMOC_STORE_*                                            */
    int exp_mode;    /* SMOC_EXP_*                                              */
    int tally;       /* SMOC_TALLY_*                                            */
} smoc_params;

/* Launch configuration.  Zero fields mean "choose": the library's default
 * or, with autotune = 1, the fastest candidate measured on this problem.
 * None of these fields changes results except chunk (splitting a track into
 * chunks re-associates its psi recurrence through the affine scan) and
 * segment_flush (tallies summed per segment instead of per run), both
 * rounding-level changes. */
typedef struct {
    int seg_per_block;   /* -p: segments per work item; used as chunk when
                            chunk is 0 (0 = whole tracks)                     */
    int groups_per_lane; /* 1 or 2 energy groups per lane                     */
    int waves_per_block; /* wavefronts per workgroup (1..16; default 4)     */
    int chunk;           /* split tracks longer than this many segments into
                            chunks of about this length joined by an affine
                            psi scan (0 = never); the first chunk of a track
                            may be up to 1023 segments longer when chunk is
                            a multiple of 1024                                */
    int persistent;      /* 1: persistent workgroups pulling work from a queue */
    int use_graph;       /* 1: capture the run into a hipGraph, replay it      */
    int autotune;        /* 1: measure candidates on first run, keep the best  */
    int reduce_waves;    /* deterministic tallies: wavefronts per workgroup of
                            the segmented reduction (1..16; default 4); the
                            summation order does not depend on it           */
    int segment_flush;   /* 1: tally every segment on its own even when the
                            stream is region-sorted (the loop a stream in
                            generation order takes), instead of summing runs
                            in registers; a measurement switch that shows
                            what the run sums save.  Atomic tallies and
                            FIXED64 only: on a region-sorted stream the
                            deterministic tallies refuse it (SMOC_ERR_ARG),
                            since their run bookkeeping follows the stream */
} smoc_launch;

typedef struct {
    double total_ms;   /* whole run, HIP events on the context stream        */
    double reset_ms;   /* fine_flux reset                                    */
    double xfer_ms;    /* chunked runs: transfer pass + psi scan (chunked
                          deterministic runs of a region-sorted stream, which
                          sweep every chunk once: the chunk plan only)       */
    double sweep_ms;   /* attenuation sweep kernel (the whole graph with
                          use_graph, whose phases are not timed apart; with
                          the chunked deterministic runs above, also the psi
                          scan and the per-run correction)                   */
    double reduce_ms;  /* deterministic segmented reduction (DET_* only)     */
    int64_t segments;  /* segments swept                                     */
    int64_t work_items;/* (track or chunk) x group-slice items launched      */
    smoc_launch used;  /* launch configuration actually used                 */
} smoc_timing;

/* Reference defaults (src/cpu/init.c:4-24, src/cuda/init.cu:30-44). */
typedef struct {
    int source_2D_regions;      /* 5000 */
    int coarse_axial_intervals; /* 27 */
    int fine_axial_intervals;   /* 5 */
    int decomp_assemblies_ax;   /* 20 */
    int64_t segments;           /* 50,000,000 */
    int egroups;                /* 128 */
    int seg_per_block;          /* 100 (CUDA -p) */
    int source_3D_regions;      /* derived: ceil(2D * coarse / decomp) = 6750 */
} smoc_input;

/* ---- host-side utilities (no GPU needed) ------------------------------- */
int smoc_abi_version(void);
const char *smoc_strerror(int code);
const char *smoc_last_error(void);

/* set_default_input + source_3D_regions (src/cpu/init.c:4-24, main.c:18-19) */
int smoc_default_input(smoc_input *in);
/* recompute in->source_3D_regions from the other fields */
int smoc_derive_regions(smoc_input *in);

/* buildExponentialTable(precision, maxVal) (src/cpu/init.c:81-117).
 * values receives 2N floats (slope, intercept per interval); capacity is in
 * floats.  The slope is stored as -exp(-n dx), as the reference does. */
int smoc_build_exp_table(float precision, float maxVal, float *values, int capacity,
                         int *N, float *dx, float *maxVal_out);
/* interpolateTable (src/cpu/kernel.c:337-361) on the host, for checks */
float smoc_exp_table_lookup(const float *values, float dx, float maxVal, float x);

/* Axial entry point decode used by every kernel: f = min((int)(z*F), F-1),
 * zin = ((z*F - f) - 0.5) * dz, all in FP32 (dz = 0.1, src/cpu/kernel.c:99). */
void smoc_decode_z(float z, int F, int *f, float *zin);

/* Seeded, counter-based input generator: element i of each array depends
 * only on (seed, array, i), so results do not depend on thread count.
 *   sigT ~ U(0.1, 10) [R][G]; fine_source ~ U(0,1) [R][F][G]. */
int smoc_generate_regions(uint64_t seed, int G, int R, int F, float *sigT, float *fine_source);
/*   len ~ U(0, 0.1); region ~ U[0, R); z ~ U(0, 1), for segments [0, S). */
int smoc_generate_segments(uint64_t seed, int64_t S, int R, float *len, int32_t *region,
                           float *z);
/*   psi_in ~ U(0,1) [T][G]. */
int smoc_generate_psi(uint64_t seed, int64_t T, int G, float *psi_in);
/*   scattering ratio ~ cmax * U(0,1) [R][G], 0 <= cmax < 1 (outer iteration;
 *   no counterpart in the reference). */
int smoc_generate_scatter(uint64_t seed, int R, int G, float cmax, float *scatter);
/* Tracks of L consecutive segments (the last one shorter when S % L != 0).
 * Returns the track count via *T; track_off needs ceil(S/L)+1 entries. */
int smoc_tracks_fixed(int64_t S, int L, int64_t *track_off, int64_t *T);
/* T tracks with lognormal lengths (median, sigma) rounded and clipped to
 * [lo, hi]; track_off [T+1]; *S receives the total segment count. */
int smoc_tracks_lognormal(uint64_t seed, int64_t T, double median, double sigma, int lo,
                          int hi, int64_t *track_off, int64_t *S);
/* Stable counting sort of the segment stream by (region, fine interval)
 * key, in place.  Tracks stay defined as contiguous ranges of the stream. */
int smoc_sort_segments(int64_t S, int R, int F, float *len, int32_t *region, float *z);

/* Host threads for the generator and sort (reference -t, src/cpu/io.c:114-120);
 * n <= 0 leaves the OpenMP default. */
int smoc_set_host_threads(int n);

/* Fixed-index-order FP64 sums, the result checksums. */
double smoc_checksum_f64(const double *x, int64_t n);
double smoc_checksum_f32(const float *x, int64_t n);

/* ---- fixture files -------------------------------------------------------
 * Binary files that carry a problem (inputs) or a sweep's results between
 * hosts, so parity can be checked where the oracle does not run.  Both are
 * little-endian: a fixed header (magic, version, shape) followed by the raw
 * arrays, with a 64-bit FNV-1a hash of the arrays in the header.  Readers
 * check magic, version, sizes and hash and fail with SMOC_ERR_ARG (bad or
 * corrupt file) or SMOC_ERR_NOMEM/SMOC_ERR_ARG on I/O errors; the message
 * names the problem. */
typedef struct {
    int G, R, F;
    int64_t S, T;
    uint64_t seed;   /* generator seed, informational (0 if unknown) */
    uint32_t flags;  /* bit 0: stream sorted by (region, fine interval) */
} smoc_problem_info;

typedef struct {
    int G, R, F;
    int64_t T;
    int storage, exp_mode, tally; /* how the results were produced */
    double fine_flux_checksum;    /* fixed-order FP64 sums */
    double state_flux_checksum;
} smoc_results_info;

/* problem file: sigT [R][G], fine_source [R][F][G], len/region/z [S],
 * track_off [T+1], psi_in [T][G] */
int smoc_problem_write(const char *path, const smoc_problem_info *info, const float *sigT,
                       const float *fine_source, const float *len, const int32_t *region,
                       const float *z, const int64_t *track_off, const float *psi_in);
int smoc_problem_read_info(const char *path, smoc_problem_info *info);
int smoc_problem_read(const char *path, float *sigT, float *fine_source, float *len,
                      int32_t *region, float *z, int64_t *track_off, float *psi_in);
/* results file: fine_flux [R][F][G] (FP64), psi_out [T][G]; the checksums
 * are computed by the writer */
int smoc_results_write(const char *path, const smoc_results_info *info, const double *fine_flux,
                       const float *psi_out);
int smoc_results_read_info(const char *path, smoc_results_info *info);
int smoc_results_read(const char *path, double *fine_flux, float *psi_out);
/* Compare results against a results file: checksum relative errors and
 * normwise (l2) relative errors of fine_flux and psi_out (any pointer may be
 * NULL).  Shapes must match. */
int smoc_results_compare(const char *path, const double *fine_flux, const float *psi_out,
                         double *fine_checksum_rel, double *state_checksum_rel,
                         double *fine_normwise, double *state_normwise);

/* ---- device path -------------------------------------------------------- */
int smoc_device_count(int *n);
/* Device name and compute-unit count for reports. */
int smoc_device_info(int device, char *name, int name_cap, int *compute_units);

/* Measurement utility (no reference counterpart): copies `bytes` bytes
 * between two device buffers with a plain vectorised copy kernel `reps`
 * times and returns the mean kernel time.  Used to calibrate HBM traffic
 * counters and to measure the attainable copy bandwidth next to the roofline.
 * variant (SMOC_COPY_VARIANTS of them) picks the copy kernel:
 *   0  4 x 16 B per lane, one pass, one 256-thread workgroup per 16 KiB
 *   1  8 x 16 B per lane in flight, grid-stride over 8 workgroups per CU
 *   2  as 1 with nontemporal loads and stores
 *   3  16 x 16 B per lane, one pass, one workgroup per 64 KiB */
#define SMOC_COPY_VARIANTS 4
int smoc_probe_copy(int device, int64_t bytes, int reps, int variant, double *mean_ms);

/* Measurement utility (no reference counterpart): the unsorted sweep's
 * per-segment FP32 atomic tally alone -- S segments of 100 per wavefront, each
 * adding 1.0 to the G groups of a pseudo-random row of a [rows][G padded to
 * 128] array, with no gathers and no arithmetic.  Returns the mean kernel time
 * of `reps` launches (after one warm-up): the floor that the atomic tally sets
 * for k_sweep on a stream in generation order.  Needs 1 <= G <= 65536,
 * 1 <= rows, 0 <= S, reps >= 1. */
int smoc_probe_tally(int device, int64_t S, int rows, int G, int reps, double *mean_ms);

int smoc_create(smoc_ctx **ctx, const smoc_params *p);
int smoc_destroy(smoc_ctx *ctx);

/* Region data: sigT [R][G], fine_source [R][F][G] (FP32 masters, quantised on
 * the device to the context's storage type), fine_flux_init [R][F][G] or
 * NULL for zeros.  May be called again to replace the data. */
int smoc_set_regions(smoc_ctx *ctx, const float *sigT, const float *fine_source,
                     const float *fine_flux_init);

/* Segment stream (structure of arrays) and tracks.  Track t sweeps segments
 * [track_off[t], track_off[t+1]) in order, carrying its angular flux
 * starting from psi_in[t][0..G).  track_off[0] must be 0 and
 * track_off[T] == S; region ids must lie in [0, R). */
int smoc_set_segments(smoc_ctx *ctx, int64_t S, const float *len, const int32_t *region,
                      const float *z, int64_t T, const int64_t *track_off,
                      const float *psi_in);

/* One sweep: fine_flux := fine_flux_init + all tallies; psi_out := outgoing
 * flux of every track.  Inputs are not modified, so repeated runs start from
 * the same state.  timing may be NULL.  launch may be NULL (defaults). */
int smoc_run(smoc_ctx *ctx, const smoc_launch *launch, smoc_timing *timing);

/* ---- allocation-free sweep ------------------------------------------------
 * No counterpart in the reference: an entry point for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  The same sweep as smoc_run, but the
 * scratch that smoc_run allocates in the context on first use -- the
 * persistent work counters and the chunk plan of a chunked sweep (work items
 * and lists, each chunk's incoming psi and affine transfer) -- comes from a
 * fixed device array of SMOC_NOALLOC_SCRATCH_BYTES per GPU, which the library
 * holds from its first smoc_create on that GPU until the process ends.  The
 * plan is rebuilt in that array by a device kernel on every call, so the call
 * allocates neither device nor host memory.  Inputs and outputs are the
 * context's, as for smoc_run (the smoc_set_* calls allocate, this one does
 * not).  launch->autotune and launch->use_graph must be 0; a chunk plan that
 * does not fit the array is refused with SMOC_ERR_ARG, naming the bytes it
 * needs.  All calls on one GPU share the array and leave their contents in
 * it: calls that overlap on one GPU (several host threads or contexts at
 * once) race; calls made one after another do not. */
#define SMOC_NOALLOC_SCRATCH_BYTES (64u << 20)
int smoc_run_noalloc(smoc_ctx *ctx, const smoc_launch *launch, smoc_timing *timing);

/* Copy results to the host: fine_flux [R][F][G] as double, psi_out [T][G].
 * Either pointer may be NULL. */
int smoc_get_results(smoc_ctx *ctx, double *fine_flux, float *psi_out);

/* ---- track producer: ray tracing through a pin-cell lattice ----------------
 * Replaces the reference's random region and interval draws
 * (src/cpu/kernel.c:47-50) by the step a transport code performs before the
 * sweep.  Geometry: nx x ny square cells of side pitch, each with a centred
 * pin of radius pin_radius, extruded over height and cut into nz axial
 * layers.  Region ids ((iy * nx + ix) * 2 + outside) * nz + layer, so the
 * context needs R = 2 nx ny nz regions.  Track t starts from a point drawn
 * uniformly in the box with an isotropic direction and runs between the two
 * box faces it meets; its segments are the pieces between cell faces, pin
 * surfaces and layer planes; z is a segment midpoint's position in its
 * layer, in [0, 1).  Traced on the GPU in FP64. */
typedef struct {
    int nx, ny, nz;
    double pitch, pin_radius, height;
} smoc_geometry;

/* Segments per track: track_off [T+1] (exclusive prefix sums) and S. */
int smoc_trace_count(int device, const smoc_geometry *geometry, uint64_t seed, int64_t T,
                     int64_t *track_off, int64_t *S);
/* The segments themselves, len / region / z [S], for the track_off that
 * smoc_trace_count returned with the same geometry, seed and T. */
int smoc_trace_segments(int device, const smoc_geometry *geometry, uint64_t seed, int64_t T,
                        const int64_t *track_off, float *len, int32_t *region, float *z);

/* ---- outer (source) iteration ----------------------------------------------
 * No counterpart in the reference kernel: it is the step after the sweep in
 * the full application the kernel was cut from (README.txt:40-50 describes
 * the kernel as that application's inner loop).  A fixed-source problem with
 * isotropic within-group scattering: after each sweep, the angle-averaged
 * flux of (region r, interval f, group g) is estimated by track length,
 *   phi = fine_flux / vol,  vol = sum of 0.5 * len over the (r, f) segments,
 * and the source becomes
 *   fine_source[r][f][g] = q_ext[r][f][g] + scatter[r][g] * sigT[r][g] * phi
 * (FP64 arithmetic, stored as FP32 and repacked for the next sweep; rows
 * without segments keep q_ext).  With scatter < 1 the iteration converges to
 * the fixed point of source and flux. */

/* scatter [R][G], the scattering ratio sigma_s / sigT in [0, 1); q_ext
 * [R][F][G], or NULL to take the current fine_source as the external source.
 * Call after smoc_set_regions (which clears it). */
int smoc_set_scattering(smoc_ctx *ctx, const float *scatter, const float *q_ext);

/* One source update from the last smoc_run's fine_flux.  change (may be
 * NULL) receives ||q_new - q_old||_2 / ||q_old||_2. */
int smoc_source_update(smoc_ctx *ctx, double *change);

/* n_outer iterations of (smoc_run, source update).  timing sums the runs'
 * timings (a graph launch replays its graph every iteration); update_ms is
 * the total time of the updates, change the last update's relative change.
 * Any output pointer may be NULL.  Afterwards fine_flux and psi_out are those
 * of the last sweep and fine_source the last update's. */
int smoc_iterate(smoc_ctx *ctx, const smoc_launch *launch, int n_outer, smoc_timing *timing,
                 double *update_ms, double *change);

/* Copy the current fine_source [R][F][G] (FP32 master) to the host. */
int smoc_get_source(smoc_ctx *ctx, float *fine_source);

#ifdef __cplusplus
}
#endif
#endif
