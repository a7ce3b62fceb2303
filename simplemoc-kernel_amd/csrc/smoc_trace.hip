// smoc_trace.hip -- the segment stream's producer: straight tracks ray-traced
// through a lattice of pin cells (gfx950).
//
// The reference draws each segment's region and axial interval at random
// (src/cpu/kernel.c:47-50); a transport code obtains them by ray tracing
// tracks through its geometry.  This is that step for a simple geometry: an
// nx x ny lattice of square cells of side `pitch`, each with a centred
// cylindrical pin of radius `pin_radius`, extruded over `height` and cut into
// nz axial layers.  Region ids: ((iy * nx + ix) * 2 + outside) * nz + layer,
// so R = 2 nx ny nz.  A segment's z is its midpoint's position inside its
// layer, in [0, 1), which the sweep decodes into a fine interval and offset.
//
// Tracks: track t starts from a point drawn uniformly in the box with an
// isotropic direction (Marsaglia's disc method: square roots only), and is
// extended both ways to the box surface.  One thread traces one track: a
// counting pass sizes the stream, a second pass writes it.  All arithmetic is
// FP64 without contraction, with correctly rounded division and square root,
// so the same operations in C on the host give the same segments bit for bit.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

#include "smoc.h"
#include "smoc_internal.h"

namespace {

// the library's counter-based generator: draw(key, i) = mix(key + (i + 1) *
// golden), key = mix(seed * golden + stream)
constexpr uint64_t kGolden = 0x9E3779B97F4A7C15ULL;
constexpr uint64_t kStreamTrace = 9;
constexpr int kDrawsPerTrack = 64;   // draw indices t * 64 .. t * 64 + 63
constexpr int kDiscAttempts = 30;    // rejection attempts for the direction

__host__ __device__ __forceinline__ uint64_t mix64(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
__device__ __forceinline__ double u53(uint64_t key, uint64_t i)
{
    return (double)(mix64(key + (i + 1) * kGolden) >> 11) * 0x1p-53;
}

struct Ray {
    double ex, ey, ez;  // entry point on the box surface
    double dx, dy, dz;  // unit direction
    double L;           // length inside the box
};

// distance from p along d to the slab [0, hi] boundary ahead (d != 0), or a
// huge value for d == 0
__device__ __forceinline__ double to_face(double p, double d, double hi)
{
#pragma clang fp contract(off)
    if (d > 0.0)
        return (hi - p) / d;
    if (d < 0.0)
        return (0.0 - p) / d;
    return 1e300;
}

__device__ Ray make_ray(const smoc_geometry &g, uint64_t key, int64_t t)
{
#pragma clang fp contract(off)
    const double X = g.nx * g.pitch, Y = g.ny * g.pitch, Z = g.height;
    const uint64_t b = (uint64_t)t * kDrawsPerTrack;
    const double px = X * u53(key, b), py = Y * u53(key, b + 1), pz = Z * u53(key, b + 2);
    double dx = 0.0, dy = 0.0, dz = 1.0;
    for (int k = 0; k < kDiscAttempts; k++) {
        const double a = 2.0 * u53(key, b + 3 + 2 * k) - 1.0;
        const double c = 2.0 * u53(key, b + 4 + 2 * k) - 1.0;
        const double s = a * a + c * c;
        if (s > 0.0 && s < 1.0) {
            const double w = 2.0 * sqrt(1.0 - s);
            dx = a * w;
            dy = c * w;
            dz = 1.0 - 2.0 * s;
            break;
        }
    }
    const double back = fmin(fmin(to_face(px, -dx, X), to_face(py, -dy, Y)), to_face(pz, -dz, Z));
    const double fwd = fmin(fmin(to_face(px, dx, X), to_face(py, dy, Y)), to_face(pz, dz, Z));
    Ray r;
    r.ex = px - back * dx;
    r.ey = py - back * dy;
    r.ez = pz - back * dz;
    r.dx = dx;
    r.dy = dy;
    r.dz = dz;
    r.L = back + fwd;
    return r;
}

// Trace one track; emit(len, region, z) per segment.  Cells are walked with a
// DDA in x and y; inside a cell the pin circle splits the chord, and every
// piece is cut at the axial layer planes.
template <typename Emit>
__device__ void trace(const smoc_geometry &g, const Ray &r, Emit &&emit)
{
#pragma clang fp contract(off)
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    const double p = g.pitch, rr = g.pin_radius * g.pin_radius, dzl = g.height / nz;
    // starting cell: the cell of a point just inside the box along the ray
    int ix = (int)floor((r.ex + 1e-9 * r.dx) / p), iy = (int)floor((r.ey + 1e-9 * r.dy) / p);
    ix = ix < 0 ? 0 : (ix > nx - 1 ? nx - 1 : ix);
    iy = iy < 0 ? 0 : (iy > ny - 1 ? ny - 1 : iy);
    const int sx = r.dx > 0.0 ? 1 : -1, sy = r.dy > 0.0 ? 1 : -1;
    double t = 0.0;
    while (t < r.L && ix >= 0 && ix < nx && iy >= 0 && iy < ny) {
        const double tx = r.dx != 0.0 ? ((ix + (sx > 0 ? 1 : 0)) * p - r.ex) / r.dx : 1e300;
        const double ty = r.dy != 0.0 ? ((iy + (sy > 0 ? 1 : 0)) * p - r.ey) / r.dy : 1e300;
        double tend = fmin(fmin(tx, ty), r.L);
        if (tend < t)
            tend = t;
        // pin circle of this cell: |e + s d - c|^2 = r^2 in the plane
        const double cx = (ix + 0.5) * p, cy = (iy + 0.5) * p;
        const double ox = r.ex - cx, oy = r.ey - cy;
        const double A = r.dx * r.dx + r.dy * r.dy;
        const double B = ox * r.dx + oy * r.dy;
        const double C = ox * ox + oy * oy - rr;
        double cut[4];
        int nc = 0;
        cut[nc++] = t;
        if (A > 0.0) {
            const double disc = B * B - A * C;
            if (disc > 0.0) {
                const double sq = sqrt(disc);
                const double s1 = (-B - sq) / A, s2 = (-B + sq) / A;
                if (s1 > t && s1 < tend)
                    cut[nc++] = s1;
                if (s2 > t && s2 < tend)
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
            // axial layers crossed by [a, b]
            double s0 = a;
            while (s0 < b) {
                const double zs = r.ez + s0 * r.dz;
                int layer;
                double s1;
                if (r.dz > 0.0) {
                    layer = (int)floor(zs / dzl);
                    layer = layer < 0 ? 0 : (layer > nz - 1 ? nz - 1 : layer);
                    s1 = ((layer + 1) * dzl - r.ez) / r.dz;
                } else if (r.dz < 0.0) {
                    layer = (int)ceil(zs / dzl) - 1;
                    layer = layer < 0 ? 0 : (layer > nz - 1 ? nz - 1 : layer);
                    s1 = (layer * dzl - r.ez) / r.dz;
                } else {
                    layer = (int)floor(zs / dzl);
                    layer = layer < 0 ? 0 : (layer > nz - 1 ? nz - 1 : layer);
                    s1 = b;
                }
                if (!(s1 > s0))  // a plane exactly at s0: step past it
                    s1 = s0 + 1e-12;
                if (!(s1 > s0) || s1 > b)
                    s1 = b;
                const double zm = r.ez + 0.5 * (s0 + s1) * r.dz;
                double zf = (zm - layer * dzl) / dzl;
                zf = zf < 0.0 ? 0.0 : (zf >= 1.0 ? 0x1.fffffep-1 : zf);
                emit(s1 - s0, r2 * nz + layer, zf);
                s0 = s1;
            }
        }
        t = tend;
        if (tend >= r.L)
            break;
        if (tx <= ty)
            ix += sx;
        if (ty <= tx)
            iy += sy;
    }
}

__global__ void __launch_bounds__(64) k_trace_count(smoc_geometry g, uint64_t key, int64_t T,
                                                    int64_t *count)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= T)
        return;
    const Ray r = make_ray(g, key, t);
    int64_t n = 0;
    trace(g, r, [&](double, int64_t, double) { ++n; });
    count[t] = n;
}

__global__ void __launch_bounds__(64) k_trace_fill(smoc_geometry g, uint64_t key, int64_t T,
                                                   const int64_t *off, float *len,
                                                   int32_t *region, float *z)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= T)
        return;
    const Ray r = make_ray(g, key, t);
    int64_t k = off[t];
    const int64_t ke = off[t + 1];
    trace(g, r, [&](double l, int64_t reg, double zf) {
        if (k < ke) {
            len[k] = (float)l;
            region[k] = (int32_t)reg;
            z[k] = (float)zf;
        }
        ++k;
    });
}

bool geometry_ok(const smoc_geometry *g)
{
    return g && g->nx >= 1 && g->ny >= 1 && g->nz >= 1 && g->pitch > 0.0 &&
           g->pin_radius >= 0.0 && g->pin_radius < 0.5 * g->pitch && g->height > 0.0 &&
           (int64_t)2 * g->nx * g->ny * g->nz <= INT32_MAX / 4;
}

#define TRACE_HIPCHK(call)                                                                     \
    do {                                                                                       \
        hipError_t e_ = (call);                                                                \
        if (e_ != hipSuccess) {                                                                \
            smoc_set_error("%s failed: %s", #call, hipGetErrorString(e_));                     \
            return SMOC_ERR_HIP;                                                               \
        }                                                                                      \
    } while (0)

// device buffer released on scope exit
template <typename T> struct Buf {
    T *p = nullptr;
    ~Buf() { (void)hipFree(p); }
    hipError_t alloc(size_t n) { return hipMalloc(reinterpret_cast<void **>(&p), (n ? n : 1) * sizeof(T)); }
};

int check_args(int device, const smoc_geometry *g, int64_t T, const char *what)
{
    if (!geometry_ok(g) || T < 0 || T > (int64_t)INT32_MAX) {
        smoc_set_error("%s: need nx, ny, nz >= 1, pitch > 0, 0 <= pin_radius < pitch / 2, "
                       "height > 0, 2 nx ny nz < 2^29 and 0 <= T < 2^31", what);
        return SMOC_ERR_ARG;
    }
    int n = 0;
    int rc = smoc_device_count(&n);
    if (rc != SMOC_OK)
        return rc;
    if (device < 0 || device >= n) {
        smoc_set_error("%s: device %d not available (%d HIP devices)", what, device, n);
        return SMOC_ERR_NODEV;
    }
    TRACE_HIPCHK(hipSetDevice(device));
    return SMOC_OK;
}

}  // namespace

extern "C" int smoc_trace_count(int device, const smoc_geometry *g, uint64_t seed, int64_t T,
                                int64_t *track_off, int64_t *S)
{
    int rc = check_args(device, g, T, "smoc_trace_count");
    if (rc != SMOC_OK)
        return rc;
    if (!track_off || !S) {
        smoc_set_error("smoc_trace_count: NULL output");
        return SMOC_ERR_ARG;
    }
    const uint64_t key = mix64(seed * kGolden + kStreamTrace);
    std::vector<int64_t> cnt((size_t)T);
    if (T > 0) {
        Buf<int64_t> d;
        TRACE_HIPCHK(d.alloc((size_t)T));
        hipLaunchKernelGGL(k_trace_count, dim3((unsigned)((T + 63) / 64)), dim3(64), 0, 0, *g, key,
                           T, d.p);
        TRACE_HIPCHK(hipGetLastError());
        TRACE_HIPCHK(hipMemcpy(cnt.data(), d.p, sizeof(int64_t) * T, hipMemcpyDeviceToHost));
    }
    track_off[0] = 0;
    for (int64_t t = 0; t < T; t++)
        track_off[t + 1] = track_off[t] + cnt[(size_t)t];
    *S = track_off[T];
    return SMOC_OK;
}

extern "C" int smoc_trace_segments(int device, const smoc_geometry *g, uint64_t seed, int64_t T,
                                   const int64_t *track_off, float *len, int32_t *region,
                                   float *z)
{
    int rc = check_args(device, g, T, "smoc_trace_segments");
    if (rc != SMOC_OK)
        return rc;
    if (!track_off || track_off[0] != 0) {
        smoc_set_error("smoc_trace_segments: track_off must come from smoc_trace_count");
        return SMOC_ERR_ARG;
    }
    // k_trace_fill writes track t's segments at [track_off[t], track_off[t+1]),
    // bounded by track_off[t+1] only: offsets that decrease would let a track
    // write past the S-sized buffers, so they are refused here
    for (int64_t t = 0; t < T; t++)
        if (track_off[t + 1] < track_off[t]) {
            smoc_set_error("smoc_trace_segments: track_off not monotone at %lld", (long long)t);
            return SMOC_ERR_ARG;
        }
    const int64_t S = track_off[T];
    if (S > 0 && (!len || !region || !z)) {
        smoc_set_error("smoc_trace_segments: NULL output");
        return SMOC_ERR_ARG;
    }
    if (S == 0)
        return SMOC_OK;
    const uint64_t key = mix64(seed * kGolden + kStreamTrace);
    Buf<int64_t> doff;
    Buf<float> dlen, dz;
    Buf<int32_t> dreg;
    TRACE_HIPCHK(doff.alloc((size_t)T + 1));
    TRACE_HIPCHK(dlen.alloc((size_t)S));
    TRACE_HIPCHK(dz.alloc((size_t)S));
    TRACE_HIPCHK(dreg.alloc((size_t)S));
    TRACE_HIPCHK(hipMemcpy(doff.p, track_off, sizeof(int64_t) * (T + 1), hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_trace_fill, dim3((unsigned)((T + 63) / 64)), dim3(64), 0, 0, *g, key, T,
                       doff.p, dlen.p, dreg.p, dz.p);
    TRACE_HIPCHK(hipGetLastError());
    TRACE_HIPCHK(hipMemcpy(len, dlen.p, sizeof(float) * S, hipMemcpyDeviceToHost));
    TRACE_HIPCHK(hipMemcpy(region, dreg.p, sizeof(int32_t) * S, hipMemcpyDeviceToHost));
    TRACE_HIPCHK(hipMemcpy(z, dz.p, sizeof(float) * S, hipMemcpyDeviceToHost));
    return SMOC_OK;
}
