/*
 * smoc_main.c -- host driver: the reference's command line and report around
 * the libsmoc C ABI.
 *
 * Flow of the reference driver (src/cuda/main.cu:3-115, src/cpu/main.c:3-62):
 * defaults -> read_CLI -> derived region count -> input summary -> build
 * inputs -> timed attenuation run -> report "Runtime" and "Time per
 * Intersection" (= runtime / (segments * groups), main.c:55-58).  Inputs come
 * from the seeded generator of libsmoc instead of rand(), and the report adds
 * the fine_flux and state_flux checksums the reference never computes.
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "smoc.h"

typedef struct {
    smoc_input in;
    int device;
    int regions_override;
    int storage, exp_mode, tally;
    int track_len;          /* fixed-length tracks */
    int64_t lognormal_tracks;/* > 0: lognormal batches instead of fixed tracks */
    int sorted;
    uint64_t seed;
    int steps;
    int nthreads;
    int p_given;
    const char *save_problem, *load_problem, *save_results, *check_results;
    double tol;
    int outer;
    double scatter;
    int64_t trace_tracks;  /* > 0: ray-traced tracks instead of generated segments */
    smoc_geometry geo;
    smoc_launch launch;
} options;

static void border_print(void)
{
    printf("================================================================================\n");
}

static void center_print(const char *s, int width)
{
    int pad = (width - (int)strlen(s)) / 2;
    printf("%*s%s\n", pad > 0 ? pad : 0, "", s);
}

/* comma-grouped integer (src/cpu/io.c:49-68 prints the same grouping) */
static void fancy_int(int64_t a)
{
    char buf[32], out[48];
    snprintf(buf, sizeof(buf), "%lld", (long long)a);
    int n = (int)strlen(buf), o = 0;
    for (int i = 0; i < n; i++) {
        out[o++] = buf[i];
        if (buf[i] != '-' && (n - 1 - i) % 3 == 0 && i != n - 1)
            out[o++] = ',';
    }
    out[o] = 0;
    printf("%s\n", out);
}

static void print_CLI_error(void)
{
    printf("Usage: ./smoc <options>\n");
    printf("Options include:\n");
    printf("  -t <threads>          Host threads for input generation\n");
    printf("  -s <segments>         Number of segments to process\n");
    printf("  -e <energy groups>    Number of energy groups\n");
    printf("  -p <segs per block>   Segments per block (chunk length of long tracks)\n");
    printf("  -d <device ID>        HIP device ID number\n");
    printf("  --storage f32|bf16|f16             sigT / fine_source storage\n");
    printf("  --exp native|table|accurate        1-exp(-tau) evaluation\n");
    printf("  --tally atomic-f32|atomic-f64|det-f64|det-kahan|fixed64\n");
    printf("  --track-len L          fixed track length in segments (default 100)\n");
    printf("  --lognormal T          T tracks of lognormal(4000, 1.0) length in [16, 200000]\n");
    printf("  --sorted               sort the stream by (region, fine interval)\n");
    printf("  --regions R            3D source regions (default: derived, 6750)\n");
    printf("  --seed N               generator seed (default 42)\n");
    printf("  --steps K              timed runs (default 1)\n");
    printf("  --gpl 1|2 --wpb W --persistent --graph --autotune   launch tuning\n");
    printf("  --save-problem F / --load-problem F   write / read the inputs (fixture file)\n");
    printf("  --save-results F / --check-results F  write results / compare with a results\n");
    printf("                         file (exit 3 when an error exceeds --tol, default 1e-5)\n");
    printf("  --trace T              T ray-traced tracks through a pin-cell lattice (regions\n");
    printf("                         and segments from the geometry; -s and --regions unused)\n");
    printf("  --geometry nx,ny,nz[,pitch,radius,height]  lattice for --trace (default\n");
    printf("                         50,50,27,1.26,0.54,366: 5000 2D regions, 27 layers)\n");
    printf("  --outer N --scatter C  then N outer (source) iterations with scattering ratios\n");
    printf("                         C * U(0,1) per region and group (default C 0.5)\n");
    printf("See README for the default run values\n");
    exit(1);
}

static int parse_choice(const char *v, const char *const *names, int n)
{
    for (int i = 0; i < n; i++)
        if (strcmp(v, names[i]) == 0)
            return i;
    print_CLI_error();
    return -1;
}

static double wall_s(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

/* read_CLI (src/cuda/io.cu:107-162): unknown options print usage and exit(1) */
static void read_CLI(int argc, char *argv[], options *o)
{
    static const char *const stores[] = {"f32", "bf16", "f16"};
    static const char *const exps[] = {"native", "table", "accurate"};
    static const char *const tallies[] = {"atomic-f32", "atomic-f64", "det-f64", "det-kahan",
                                          "fixed64"};
    for (int i = 1; i < argc; i++) {
        const char *arg = argv[i];
#define NEXT() (++i < argc ? argv[i] : (print_CLI_error(), ""))
        if (strcmp(arg, "-t") == 0)
            o->nthreads = atoi(NEXT());
        else if (strcmp(arg, "-s") == 0)
            o->in.segments = atoll(NEXT());
        else if (strcmp(arg, "-e") == 0)
            o->in.egroups = atoi(NEXT());
        else if (strcmp(arg, "-p") == 0) {
            o->in.seg_per_block = atoi(NEXT());
            o->p_given = 1;
        }
        else if (strcmp(arg, "-d") == 0)
            o->device = atoi(NEXT());
        else if (strcmp(arg, "--storage") == 0)
            o->storage = parse_choice(NEXT(), stores, 3);
        else if (strcmp(arg, "--exp") == 0)
            o->exp_mode = parse_choice(NEXT(), exps, 3);
        else if (strcmp(arg, "--tally") == 0)
            o->tally = parse_choice(NEXT(), tallies, 5);
        else if (strcmp(arg, "--track-len") == 0)
            o->track_len = atoi(NEXT());
        else if (strcmp(arg, "--lognormal") == 0)
            o->lognormal_tracks = atoll(NEXT());
        else if (strcmp(arg, "--sorted") == 0)
            o->sorted = 1;
        else if (strcmp(arg, "--regions") == 0)
            o->regions_override = atoi(NEXT());
        else if (strcmp(arg, "--seed") == 0)
            o->seed = strtoull(NEXT(), NULL, 10);
        else if (strcmp(arg, "--steps") == 0)
            o->steps = atoi(NEXT());
        else if (strcmp(arg, "--gpl") == 0)
            o->launch.groups_per_lane = atoi(NEXT());
        else if (strcmp(arg, "--wpb") == 0)
            o->launch.waves_per_block = atoi(NEXT());
        else if (strcmp(arg, "--persistent") == 0)
            o->launch.persistent = 1;
        else if (strcmp(arg, "--graph") == 0)
            o->launch.use_graph = 1;
        else if (strcmp(arg, "--autotune") == 0)
            o->launch.autotune = 1;
        else if (strcmp(arg, "--save-problem") == 0)
            o->save_problem = NEXT();
        else if (strcmp(arg, "--load-problem") == 0)
            o->load_problem = NEXT();
        else if (strcmp(arg, "--save-results") == 0)
            o->save_results = NEXT();
        else if (strcmp(arg, "--check-results") == 0)
            o->check_results = NEXT();
        else if (strcmp(arg, "--tol") == 0)
            o->tol = atof(NEXT());
        else if (strcmp(arg, "--trace") == 0)
            o->trace_tracks = atoll(NEXT());
        else if (strcmp(arg, "--geometry") == 0) {
            const char *gs = NEXT();
            if (sscanf(gs, "%d,%d,%d,%lf,%lf,%lf", &o->geo.nx, &o->geo.ny, &o->geo.nz,
                       &o->geo.pitch, &o->geo.pin_radius, &o->geo.height) < 3)
                print_CLI_error();
        }
        else if (strcmp(arg, "--outer") == 0)
            o->outer = atoi(NEXT());
        else if (strcmp(arg, "--scatter") == 0)
            o->scatter = atof(NEXT());
        else
            print_CLI_error();
#undef NEXT
    }
    if (o->in.segments < 0 || o->in.egroups < 1 || o->in.seg_per_block < 1 || o->track_len < 1 ||
        o->steps < 1 || o->nthreads < 0)
        print_CLI_error();
}

static void die(const char *what, int rc)
{
    fprintf(stderr, "smoc: %s failed: %s (%s)\n", what, smoc_strerror(rc), smoc_last_error());
    exit(2);
}

int main(int argc, char *argv[])
{
    options o;
    memset(&o, 0, sizeof(o));
    smoc_default_input(&o.in); /* 50,000,000 segments x 128 groups, as the reference */
    o.track_len = 100;
    o.seed = 42;
    o.steps = 1;
    o.exp_mode = SMOC_EXP_ACCURATE; /* --exp native: the reference's literal form */
    o.storage = SMOC_STORE_F32;
    o.tally = SMOC_TALLY_ATOMIC_F32;
    o.tol = 1e-5;
    o.scatter = 0.5;
    o.geo.nx = 50, o.geo.ny = 50, o.geo.nz = 27;
    o.geo.pitch = 1.26, o.geo.pin_radius = 0.54, o.geo.height = 366.0;
    read_CLI(argc, argv, &o);
    if (o.outer < 0 || !(o.scatter >= 0.0 && o.scatter < 1.0)) {
        fprintf(stderr, "--outer must be >= 0 and --scatter in [0, 1)\n");
        return 1;
    }
    smoc_set_host_threads(o.nthreads);
    smoc_derive_regions(&o.in);
    int R = o.regions_override > 0 ? o.regions_override : o.in.source_3D_regions;
    int G = o.in.egroups, F = o.in.fine_axial_intervals;
    smoc_problem_info pinfo;
    memset(&pinfo, 0, sizeof(pinfo));
    if (o.load_problem) {
        /* the fixture file fixes the problem: its shape overrides -s -e --regions */
        int rc = smoc_problem_read_info(o.load_problem, &pinfo);
        if (rc)
            die("smoc_problem_read_info", rc);
        G = pinfo.G, R = pinfo.R, F = pinfo.F;
        o.in.egroups = G, o.in.fine_axial_intervals = F, o.in.segments = pinfo.S;
        o.seed = pinfo.seed, o.sorted = (int)(pinfo.flags & 1u), o.lognormal_tracks = 0;
    }
    /* -p: segments per work item (tracks longer than p are split into chunks
     * joined by the psi scan); without -p tracks are swept whole */
    o.launch.seg_per_block = o.p_given ? o.in.seg_per_block : 0;

    int64_t S = o.in.segments, T = 0;
    int64_t *track_off = NULL;
    double trace_s = 0.0;
    if (o.trace_tracks > 0 && !o.load_problem) {
        R = 2 * o.geo.nx * o.geo.ny * o.geo.nz;
        T = o.trace_tracks;
        track_off = (int64_t *)malloc(sizeof(int64_t) * (size_t)(T + 1));
        if (!track_off)
            die("track allocation", SMOC_ERR_NOMEM);
        const double t0 = wall_s();
        int rc = smoc_trace_count(o.device, &o.geo, o.seed, T, track_off, &S);
        if (rc)
            die("smoc_trace_count", rc);
        trace_s = wall_s() - t0;
        o.in.segments = S;
    } else if (o.load_problem) {
        T = pinfo.T;
        track_off = (int64_t *)malloc(sizeof(int64_t) * (size_t)(T + 1));
        if (!track_off)
            die("track allocation", SMOC_ERR_NOMEM);
    } else if (o.lognormal_tracks > 0) {
        T = o.lognormal_tracks;
        track_off = (int64_t *)malloc(sizeof(int64_t) * (size_t)(T + 1));
        if (!track_off)
            die("track allocation", SMOC_ERR_NOMEM);
        int rc = smoc_tracks_lognormal(o.seed, T, 4000.0, 1.0, 16, 200000, track_off, &S);
        if (rc)
            die("smoc_tracks_lognormal", rc);
    } else {
        const int64_t Tm = (S + o.track_len - 1) / o.track_len;
        track_off = (int64_t *)malloc(sizeof(int64_t) * (size_t)(Tm + 1));
        if (!track_off)
            die("track allocation", SMOC_ERR_NOMEM);
        int rc = smoc_tracks_fixed(S, o.track_len, track_off, &T);
        if (rc)
            die("smoc_tracks_fixed", rc);
    }

    char devname[256] = "unknown";
    int cus = 0;
    smoc_device_info(o.device, devname, sizeof(devname), &cus);
    const double tuple_b = o.storage == SMOC_STORE_F32 ? 16.0 : 8.0;
    const double mem_mb =
        ((double)R * F * G * (tuple_b + 8.0) + (double)S * 12.0 + (double)T * G * 8.0) / 1048576.0;

    border_print();
    center_print("SimpleMOC-kernel segment attenuation", 80);
    center_print("MI355X / HIP build", 80);
    border_print();
    center_print("INPUT SUMMARY", 80);
    border_print();
    printf("%-25s%s (%d CUs)\n", "Device:", devname, cus);
    printf("%-25s%d\n", "Energy Groups:", G);
    printf("%-25s%d\n", "2D Source Regions:", o.in.source_2D_regions);
    printf("%-25s%d\n", "Coarse Axial Intervals:", o.in.coarse_axial_intervals);
    printf("%-25s%d\n", "Fine Axial Intervals:", F);
    printf("%-25s%d\n", "Axial Decomposition:", o.in.decomp_assemblies_ax);
    printf("%-25s%d\n", "3D Source Regions:", R);
    printf("%-25s", "Segments:");
    fancy_int(S);
    printf("%-25s", "Tracks:");
    fancy_int(T);
    printf("%-25s%.2f\n", "Memory Estimate (MB):", mem_mb);
    if (o.p_given)
        printf("%-25s%d\n", "Segments per block:", o.in.seg_per_block);
    else
        printf("%-25s%s\n", "Segments per block:", "whole tracks");
    printf("%-25s%s\n", "Exponential Table:", o.exp_mode == SMOC_EXP_TABLE ? "ON" : "OFF");
    printf("%-25s%s\n", "Exponential:",
           (const char *[]){"native", "table", "accurate (series for tau < 1)"}[o.exp_mode]);
    printf("%-25s%s\n", "Storage:",
           o.storage == SMOC_STORE_F32 ? "f32" : (o.storage == SMOC_STORE_BF16 ? "bf16" : "f16"));
    printf("%-25s%s\n", "Tally:",
           (const char *[]){"atomic-f32", "atomic-f64", "det-f64", "det-kahan", "fixed64"}[o.tally]);
    printf("%-25s%s\n", "Stream order:", o.sorted ? "sorted by (region, interval)" : "track");
    if (o.trace_tracks > 0)
        printf("%-25s%d x %d pin cells (pitch %.3g, pin radius %.3g), %d layers over %.4g\n",
               "Traced geometry:", o.geo.nx, o.geo.ny, o.geo.pitch, o.geo.pin_radius, o.geo.nz,
               o.geo.height);
    border_print();

    center_print("INITIALIZATION", 80);
    border_print();
    if (o.load_problem)
        printf("Reading inputs from %s...\n", o.load_problem);
    else
        printf("Generating seeded inputs (seed %llu)...\n", (unsigned long long)o.seed);
    float *sigT = (float *)malloc(sizeof(float) * (size_t)R * G);
    float *src = (float *)malloc(sizeof(float) * (size_t)R * F * G);
    float *len = (float *)malloc(sizeof(float) * (size_t)(S > 0 ? S : 1));
    int32_t *region = (int32_t *)malloc(sizeof(int32_t) * (size_t)(S > 0 ? S : 1));
    float *z = (float *)malloc(sizeof(float) * (size_t)(S > 0 ? S : 1));
    float *psi_in = (float *)malloc(sizeof(float) * (size_t)(T > 0 ? T : 1) * G);
    if (!sigT || !src || !len || !region || !z || !psi_in)
        die("host allocation", SMOC_ERR_NOMEM);
    int rc;
    if (o.load_problem) {
        if ((rc = smoc_problem_read(o.load_problem, sigT, src, len, region, z, track_off, psi_in)))
            die("smoc_problem_read", rc);
    } else {
        if ((rc = smoc_generate_regions(o.seed, G, R, F, sigT, src)))
            die("smoc_generate_regions", rc);
        if (o.trace_tracks > 0) {
            const double t0 = wall_s();
            if ((rc = smoc_trace_segments(o.device, &o.geo, o.seed, T, track_off, len, region,
                                          z)))
                die("smoc_trace_segments", rc);
            trace_s += wall_s() - t0;
            printf("Ray tracing: %.3f s for %lld tracks (count and fill passes, with copies)\n",
                   trace_s, (long long)T);
        } else if ((rc = smoc_generate_segments(o.seed, S, R, len, region, z))) {
            die("smoc_generate_segments", rc);
        }
        if ((rc = smoc_generate_psi(o.seed, T, G, psi_in)))
            die("smoc_generate_psi", rc);
        if (o.sorted && (rc = smoc_sort_segments(S, R, F, len, region, z)))
            die("smoc_sort_segments", rc);
    }
    if (o.save_problem) {
        smoc_problem_info wi = {G, R, F, S, T, o.seed, o.sorted ? 1u : 0u};
        if ((rc = smoc_problem_write(o.save_problem, &wi, sigT, src, len, region, z, track_off,
                                     psi_in)))
            die("smoc_problem_write", rc);
        printf("Wrote inputs to %s\n", o.save_problem);
    }

    smoc_params p = {o.device, G, R, F, o.storage, o.exp_mode, o.tally};
    smoc_ctx *ctx = NULL;
    if ((rc = smoc_create(&ctx, &p)))
        die("smoc_create", rc);
    printf("Uploading and packing region data...\n");
    if ((rc = smoc_set_regions(ctx, sigT, src, NULL)))
        die("smoc_set_regions", rc);
    printf("Uploading segment stream...\n");
    if ((rc = smoc_set_segments(ctx, S, len, region, z, T, track_off, psi_in)))
        die("smoc_set_segments", rc);
    printf("Initialization Complete.\n");
    border_print();
    center_print("SIMULATION", 80);
    border_print();
    printf("Attenuating fluxes across segments...\n");

    smoc_timing tm, best;
    memset(&best, 0, sizeof(best));
    double sum_ms = 0.0;
    for (int k = 0; k < o.steps; k++) {
        if ((rc = smoc_run(ctx, &o.launch, &tm)))
            die("smoc_run", rc);
        sum_ms += tm.total_ms;
        if (k == 0 || tm.total_ms < best.total_ms)
            best = tm;
    }
    printf("Simulation Complete.\n");

    double *flux = (double *)malloc(sizeof(double) * (size_t)R * F * G);
    float *psi_out = (float *)malloc(sizeof(float) * (size_t)(T > 0 ? T : 1) * G);
    if (!flux || !psi_out)
        die("host allocation", SMOC_ERR_NOMEM);
    if ((rc = smoc_get_results(ctx, flux, psi_out)))
        die("smoc_get_results", rc);
    const double ck_flux = smoc_checksum_f64(flux, (int64_t)R * F * G);
    const double ck_psi = smoc_checksum_f32(psi_out, T * G);

    border_print();
    center_print("RESULTS SUMMARY", 80);
    border_print();
    const double mean_s = sum_ms / o.steps / 1000.0;
    const double tpi = mean_s / (double)S / (double)G * 1.0e9; /* src/cpu/main.c:55-58 */
    printf("%-25s%.6f seconds\n", "Runtime:", mean_s);
    printf("%-25s%.6f ns\n", "Time per Intersection:", tpi);
    printf("%-25s%.4e\n", "Attenuations/s:", (double)S * G / mean_s);
    printf("%-25s%.3f ms (chunk plan/transfer/scan %.3f, sweep %.3f, reduce %.3f)\n", "Best run:",
           best.total_ms, best.xfer_ms, best.sweep_ms, best.reduce_ms);
    printf("%-25sgpl %d, wpb %d, chunk %d, persistent %d, graph %d, reduce wpb %d\n", "Launch:",
           best.used.groups_per_lane, best.used.waves_per_block, best.used.chunk,
           best.used.persistent, best.used.use_graph, best.used.reduce_waves);
    printf("%-25s%.17g\n", "fine_flux checksum:", ck_flux);
    printf("%-25s%.17g\n", "state_flux checksum:", ck_psi);
    border_print();
    if (o.outer > 0) {
        float *sc = (float *)malloc(sizeof(float) * (size_t)R * G);
        if (!sc)
            die("host allocation", SMOC_ERR_NOMEM);
        if ((rc = smoc_generate_scatter(o.seed, R, G, (float)o.scatter, sc)))
            die("smoc_generate_scatter", rc);
        if ((rc = smoc_set_scattering(ctx, sc, NULL)))
            die("smoc_set_scattering", rc);
        free(sc);
        smoc_timing ot;
        double upd_ms = 0.0, chg = 0.0;
        if ((rc = smoc_iterate(ctx, &o.launch, o.outer, &ot, &upd_ms, &chg)))
            die("smoc_iterate", rc);
        if ((rc = smoc_get_results(ctx, flux, psi_out)))
            die("smoc_get_results", rc);
        center_print("OUTER ITERATION", 80);
        border_print();
        printf("%-25s%d (scattering ratio up to %.2f)\n", "Outer iterations:", o.outer, o.scatter);
        printf("%-25s%.3f ms sweep, %.3f ms source update\n", "Per iteration:",
               ot.total_ms / o.outer, upd_ms / o.outer);
        printf("%-25s%.3e\n", "Last source change:", chg);
        printf("%-25s%.17g\n", "fine_flux checksum:", smoc_checksum_f64(flux, (int64_t)R * F * G));
        printf("%-25s%.17g\n", "state_flux checksum:", smoc_checksum_f32(psi_out, T * G));
        border_print();
    }
    int status = 0;
    if (o.save_results) {
        smoc_results_info ri = {G, R, F, T, o.storage, o.exp_mode, o.tally, 0.0, 0.0};
        if ((rc = smoc_results_write(o.save_results, &ri, flux, psi_out)))
            die("smoc_results_write", rc);
        printf("Wrote results to %s\n", o.save_results);
    }
    if (o.check_results) {
        double e[4];
        if ((rc = smoc_results_compare(o.check_results, flux, psi_out, &e[0], &e[1], &e[2],
                                       &e[3])))
            die("smoc_results_compare", rc);
        double worst = 0.0;
        for (int i = 0; i < 4; i++)
            worst = e[i] > worst ? e[i] : worst;
        printf("%-25sfine_flux %.3e, state_flux %.3e (checksums); %.3e, %.3e (normwise)\n",
               "Check vs results file:", e[0], e[1], e[2], e[3]);
        status = worst <= o.tol ? 0 : 3;
        printf("%-25s%s (tolerance %.1e)\n", "Check:", status ? "FAILED" : "passed", o.tol);
    }

    smoc_destroy(ctx);
    free(flux), free(psi_out), free(sigT), free(src), free(len), free(region), free(z),
        free(psi_in), free(track_off);
    return status;
}
