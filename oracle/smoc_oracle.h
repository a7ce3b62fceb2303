/*
 * smoc_oracle.h -- CPU restatement of SimpleMOC-kernel's segment attenuation.
 *
 * TEST INFRASTRUCTURE ONLY.  This code is the checker for the HIP product
 * path: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * may load it.  The product library never links or calls it.
 *
 * Every function restates a piece of the reference (ANL-CESAR/SimpleMOC-kernel,
 * Version 4) and cites the file:line it follows.  Parity pins (see the tests):
 *   oracle_build_table / oracle_table_lookup_f32
 *       -- bit-exact against the reference's own buildExponentialTable /
 *          interpolateTable compiled from src/cpu (oracle/_ref) and against the
 *          known answers committed in tests/golden/.
 *   oracle_segment_f32
 *       -- bit-exact against the reference attenuate_segment (strict IEEE build,
 *          reference constants ds=0.7, zin=0.3) on golden vectors; closed forms
 *          (Beer-Lambert, flat source) and brute force in FP64.
 *   oracle_segment_f64 / oracle_sweep_f64
 *       -- against oracle_segment_f32 at FP32 tolerance, and against the
 *          closed forms in FP64 (exact up to FP64 rounding).
 *   oracle_decode_z
 *       -- against a numpy restatement of the same float operations.
 */
#ifndef SMOC_ORACLE_H
#define SMOC_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* exp modes, same numbering as the product's SMOC_EXP_* */
#define ORACLE_EXP_NATIVE 0
#define ORACLE_EXP_TABLE 1

typedef struct {
    float values[1024]; /* slope, intercept pairs; N <= 512 */
    float dx;
    float maxVal;
    int N;
} oracle_table;

/* src/cpu/init.c:81-117 (buildExponentialTable) */
int oracle_build_table(float precision, float maxVal, oracle_table *t);
/* src/cpu/kernel.c:337-361 (interpolateTable) */
float oracle_table_lookup_f32(const oracle_table *t, float x);
/* Interval index the lookup uses, -1 for x > maxVal (kernel.c:340-355). */
int oracle_table_interval(const oracle_table *t, float x);

/* Axial entry point decode: z in [0,1) -> fine interval f and offset zin,
 * evaluated in FP32 (the build's segment geometry; the reference uses the
 * constant zin = 0.3f, src/cpu/kernel.c:100). */
void oracle_decode_z(float z, int F, int *f, float *zin);

/* One segment over G groups, literal FP32 order of src/cpu/kernel.c:108-332.
 * sig: sigT row [G]; src: the region's fine_source block [F][G];
 * flux: the region's fine_flux block [F][G] (+= tally); psi: [G] in/out. */
void oracle_segment_f32(int G, int F, const float *sig, const float *src, int fai,
                        float ds, float zin, int exp_mode, const oracle_table *tab,
                        float *psi, float *flux);

/* Same formulas in FP64 (float-valued constants promoted to double).
 * Table mode: interval chosen from the FP32 tau exactly as the FP32 path
 * does, slope/intercept evaluated in FP64. */
void oracle_segment_f64(int G, int F, const float *sig, const float *src, int fai,
                        float ds, float zin, int exp_mode, const oracle_table *tab,
                        double *psi, double *flux);

typedef struct {
    int G, R, F;
    int exp_mode;
    const float *sigT;        /* [R][G]    */
    const float *fine_source; /* [R][F][G] */
    int64_t S;
    const float *len;         /* [S] segment length (cm)           */
    const int32_t *region;    /* [S] source region id in [0, R)    */
    const float *z;           /* [S] axial entry point in [0, 1)   */
    int64_t T;
    const int64_t *track_off; /* [T+1] segments of track t: [off[t], off[t+1]) */
    const float *psi_in;      /* [T][G] */
} oracle_problem;

/* Sweep every track in order, carrying psi along the track
 * (src/cpu/kernel.c:29-30,43-55).  fine_flux [R][F][G] is accumulated
 * (+=); psi_out [T][G] receives each track's outgoing flux.
 * nthreads > 1 splits tracks over OpenMP threads with private tallies that
 * are summed in thread order at the end (FP64 only). */
void oracle_sweep_f64(const oracle_problem *p, double *fine_flux, double *psi_out,
                      int nthreads);
/* Sequential literal FP32 sweep (tally in float, segment order). */
void oracle_sweep_f32(const oracle_problem *p, float *fine_flux, float *psi_out);

/* OpenMP CPU baseline in the reference's structure: threads over tracks,
 * dynamic schedule, one omp lock per (region, fine interval) around the
 * tally (src/cpu/kernel.c:6,43,265-281).  Returns seconds of the sweep. */
double oracle_baseline_f32(const oracle_problem *p, float *fine_flux, float *psi_out,
                           int nthreads);

/* Fixed-order FP64 sums (checksum definitions). */
double oracle_sum_f64(const double *x, int64_t n);
double oracle_sum_f32(const float *x, int64_t n);

/* ---- track producer (ray tracing through a pin-cell lattice) --------------
 * The reference draws regions at random (src/cpu/kernel.c:47-50); the
 * library's producer ray-traces tracks through an nx x ny lattice of pin
 * cells with nz axial layers (see the product's smoc_geometry).  This is a
 * restatement of that tracer in C, FP64 without contraction: the library's
 * GPU tracer must give the same segments bit for bit.  Pins (tests): the
 * segment lengths of a track add up to its chord, every segment's region is
 * that of its midpoint by brute force, z lies in [0, 1) and matches the
 * midpoint's height, and the share of track length inside pins approaches
 * the pin volume fraction pi r^2 / pitch^2. */
typedef struct {
    int nx, ny, nz;
    double pitch, pin_radius, height;
} oracle_geometry;

typedef struct {
    double ex, ey, ez, dx, dy, dz, L;
} oracle_ray;

/* Track t's chord (entry point, unit direction, length). */
void oracle_trace_ray(const oracle_geometry *g, uint64_t seed, int64_t t, oracle_ray *r);
/* Segments of track t: writes up to cap of them (len, region, z may be NULL
 * when cap is 0) and returns how many the track has. */
int64_t oracle_trace_track(const oracle_geometry *g, uint64_t seed, int64_t t, int64_t cap,
                           float *len, int32_t *region, float *z);

#ifdef __cplusplus
}
#endif
#endif
