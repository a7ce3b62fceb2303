// smoc_kernels.hip -- gfx950 kernels of the segment-attenuation sweep.
//
// The sweep restates the arithmetic of the reference attenuate_segment
// (src/cpu/kernel.c:108-332; CUDA twin src/cuda/kernel.cu:90-192) for one
// (segment, energy group) pair per lane and step, with the segment stream
// walked in track order so the angular flux psi is carried in registers.
//
// Mapping (CDNA4, wave64):
//   * one wavefront sweeps one work item = (track or chunk of a track, group
//     slice); a slice is 64 * GPL consecutive groups, lane l owning groups
//     slice*64*GPL + j*64 + l for j < GPL;
//   * the segment stream is structure of arrays, decoded once at upload
//     (k_encode_stream) into len, code = (region * F + f) << 2 | fit kind and
//     the axial offset zin, and read with wave-uniform addresses, i.e. scalar
//     loads, ahead of the arithmetic together with the vector gather of the
//     next segment's packed region tuple;
//   * region data is packed once per smoc_set_regions into one tuple per
//     (region, fine interval, group): FP32 storage keeps (sigT, c0, c1, c2)
//     of the source fit, 16-bit storage keeps (sigT, y[f-1], y[f], y[f+1])
//     and fits in registers;
//   * consecutive segments of a track in the same (region, interval) form a
//     run whose tally is summed in registers and flushed once (an atomic,
//     or a plain store of the run partial in the deterministic modes); on a
//     region-sorted stream the tuple, the source fit and every quantity that
//     depends only on sigT are computed once per run, not per segment;
//   * block -> work mapping is XCD-aware: the 8 XCDs each get a contiguous
//     range of work items (hardware dispatches blocks round-robin over XCDs).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <type_traits>

#include "smoc.h"
#include "smoc_internal.h"
#include "smoc_plan.h"

namespace {

constexpr float kLog2e = 1.4426950408889634f;
constexpr float kDz = SMOC_DZ;
// divisions by the fit's constant denominators (kernel.c:129,155,183-184)
constexpr float kInvDz = 1.0f / kDz;
constexpr float kInv2Dz = 1.0f / (2.f * kDz);
constexpr float kInv2Dz2 = 1.0f / (2.f * kDz * kDz);

__device__ __forceinline__ float rcp(float x) { return __builtin_amdgcn_rcpf(x); }

// segments of a sorted run evaluated together by the series path
constexpr int kSeries = 8;

// s_waitcnt immediate (gfx9 encoding) that waits for lgkmcnt = 0 only:
// vmcnt 63 (bits 3:0 and 15:14), expcnt 7 (bits 6:4), lgkmcnt 0 (bits 11:8)
constexpr int kWaitLgkm0 = 0xC07F;

// ---- storage formats -------------------------------------------------------
template <int STORE> struct Tup;
template <> struct Tup<SMOC_STORE_F32> {
    using raw = float4;
    static __device__ __forceinline__ raw one() { return make_float4(1.f, 0.f, 0.f, 0.f); }
};
template <> struct Tup<SMOC_STORE_BF16> {
    using raw = uint2;
    static __device__ __forceinline__ raw one() { return make_uint2(0x3F80u, 0u); }
};
template <> struct Tup<SMOC_STORE_F16> {
    using raw = uint2;
    static __device__ __forceinline__ raw one() { return make_uint2(0x3C00u, 0u); }
};

__device__ __forceinline__ float bf_lo(uint32_t w) { return __uint_as_float(w << 16); }
__device__ __forceinline__ float bf_hi(uint32_t w) { return __uint_as_float(w & 0xFFFF0000u); }
__device__ __forceinline__ float h_to_f(uint32_t h16)
{
    return (float)__builtin_bit_cast(_Float16, (uint16_t)h16);
}
// round to nearest even (finite inputs)
__device__ __forceinline__ uint32_t f_to_bf(float x)
{
    uint32_t u = __float_as_uint(x);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return u >> 16;
}
__device__ __forceinline__ uint32_t f_to_h(float x)
{
    return (uint32_t)__builtin_bit_cast(uint16_t, (_Float16)x);
}

template <int STORE>
__device__ __forceinline__ float4 unpack(typename Tup<STORE>::raw r)
{
    if constexpr (STORE == SMOC_STORE_F32) {
        return r;
    } else if constexpr (STORE == SMOC_STORE_BF16) {
        return make_float4(bf_lo(r.x), bf_hi(r.x), bf_lo(r.y), bf_hi(r.y));
    } else {
        return make_float4(h_to_f(r.x & 0xFFFFu), h_to_f(r.x >> 16), h_to_f(r.y & 0xFFFFu),
                           h_to_f(r.y >> 16));
    }
}

// ---- segment geometry -----------------------------------------------------
struct RawSeg {
    float len, zin;
    int code;  // (region * F + f) << 2 | kind; -1 past the end
};

// fit kinds (kernel.c:111-191): interior, first and last fine interval
enum : int { KIND_INTERIOR = 0, KIND_FIRST = 1, KIND_LAST = 2 };

struct Seg {
    float ds;   // segment length
    float zin;  // axial offset inside the fine interval
    int kind;   // KIND_*
    int row;    // region * F + f, -1 past the end of the item
};

__device__ __forceinline__ RawSeg load_raw(const float *__restrict__ len,
                                           const int32_t *__restrict__ code,
                                           const float *__restrict__ zin, int64_t k)
{
    return RawSeg{len[k], zin[k], code[k]};
}

__device__ __forceinline__ Seg decode(const RawSeg r)
{
    return Seg{r.len, r.zin, r.code & 3, r.code >> 2};
}

// z -> (f, zin) with the same FP32 operations as smoc_decode_z on the host
// (no contraction), so the fine interval and offset are bit-identical.
__device__ __forceinline__ void decode_z(float z, int F, int &f, float &zin)
{
#pragma clang fp contract(off)
    const float zf = z * (float)F;
    int fi = (int)zf;
    fi = fi > F - 1 ? F - 1 : fi;
    fi = fi < 0 ? 0 : fi;
    f = fi;
    zin = ((zf - (float)fi) - 0.5f) * kDz;
}

// Source fit coefficients c0, c1, c2 of one tuple (kernel.c:111-191); the
// source at zin is q0 = c0 + c1 zin + c2 zin^2, its slope q1 = c1 + 2 c2 zin.
template <int STORE>
__device__ __forceinline__ void fit_coeffs(const float4 t, int kind, float &c0, float &c1,
                                           float &c2)
{
    if constexpr (STORE == SMOC_STORE_F32) {
        c0 = t.y;  // coefficients were fitted by the pack kernel
        c1 = t.z;
        c2 = t.w;
    } else {
        c0 = t.z;
        if (kind == KIND_FIRST) {
            c1 = (t.w - t.z) * kInvDz;
            c2 = 0.f;
        } else if (kind == KIND_LAST) {
            c1 = (t.z - t.y) * kInvDz;
            c2 = 0.f;
        } else {
            c1 = (t.y - t.w) * kInv2Dz;
            c2 = (__builtin_fmaf(-2.f, t.z, t.y) + t.w) * kInv2Dz2;
        }
    }
}

template <int STORE>
__device__ __forceinline__ void fit(const float4 t, int kind, float zin, float zin2, float &q0,
                                    float &q1, float &q2)
{
    float c0, c1, c2;
    fit_coeffs<STORE>(t, kind, c0, c1, c2);
    q0 = c0 + c1 * zin + c2 * zin2;
    q1 = c1 + 2.f * c2 * zin;
    q2 = c2;
}

struct ExpTab {
    const float *lds;  // 2N floats in LDS
    float dx, hdx, tmax;
    int n;
};

// ev = 1 - e^-tau and the two cancellation-prone combinations of the flux
// integral, tme = tau - ev and poly = tau(tau(tau-3)+6) - 6ev (kernel.c:235-
// 252, 301).
struct EvTerms {
    float ev, one_m_ev, tme, poly;
};

// Series form for 0 <= tau < 1: r4 = sum_{k>=0} (-tau)^k / (k+4)!,
// r = 1/2 - tau/6 + tau^2 r4; then ev = tau - tau^2 r, tme = tau^2 r and
// poly = 6 tau^4 r4 without subtracting nearly equal numbers.  r4 is the
// degree-4 polynomial closest to it in relative error on [0, 1] (Lawson's
// reweighted least squares, FP32 coefficients; csrc/fit_r4.py recomputes
// them): 2.1e-7 relative evaluated in FP32, and ev within 8e-8 absolute of
// 1 - e^-tau over [0, 1) (the degree-6 Taylor polynomial: 1.0e-7), with two
// fused multiply-adds fewer.
__device__ __forceinline__ float r4_series(float tau)
{
    float r4 = 0x1.3cd4c4p-16f;
    r4 = __builtin_fmaf(r4, tau, -0x1.95bd10p-13f);
    r4 = __builtin_fmaf(r4, tau, 0x1.6ba232p-10f);
    r4 = __builtin_fmaf(r4, tau, -0x1.110f3ep-7f);
    return __builtin_fmaf(r4, tau, 0x1.555552p-5f);
}

__device__ __forceinline__ EvTerms ev_series(float tau)
{
    const float r4 = r4_series(tau);
    const float t2 = tau * tau;
    const float r = (0.5f - tau * (1.f / 6.f)) + t2 * r4;
    const float tme = t2 * r;
    const float ev = tau - tme;
    return EvTerms{ev, 1.f - ev, tme, 6.f * (t2 * t2) * r4};
}

__device__ __forceinline__ EvTerms ev_direct(float tau, float ev, float one_m_ev)
{
    return EvTerms{ev, one_m_ev, tau - ev, tau * (tau * (tau - 3.f) + 6.f) - 6.f * ev};
}

template <int EXPM>
__device__ __forceinline__ EvTerms ev_terms(float tau, const ExpTab &tb)
{
    if constexpr (EXPM == SMOC_EXP_TABLE) {
        // interpolateTable (kernel.c:337-361), interval from the IEEE quotient
        float ev;
        if (tau > tb.tmax) {
            ev = 1.f;
        } else {
            int iv = (int)(tau / tb.dx + tb.hdx);
            iv = iv < 0 ? 0 : (iv > tb.n - 1 ? tb.n - 1 : iv);
            const float2 sl = reinterpret_cast<const float2 *>(tb.lds)[iv];
            ev = sl.x * tau + sl.y;
        }
        return ev_direct(tau, ev, 1.f - ev);
    } else if constexpr (EXPM == SMOC_EXP_NATIVE) {
        const float e = __builtin_amdgcn_exp2f(-tau * kLog2e);  // e^-tau, tau >= 0
        return ev_direct(tau, 1.f - e, e);
    } else {
        // series for tau < 1; the direct forms only where a lane needs them
        // (wave-uniform test, so waves with tau < 1 everywhere skip them)
        EvTerms o = ev_series(tau);
        if (__any(tau >= 1.f)) {
            const float e = __builtin_amdgcn_exp2f(-tau * kLog2e);
            const EvTerms d = ev_direct(tau, 1.f - e, e);
            if (tau >= 1.f)
                o = d;
        }
        return o;
    }
}

// Attenuation of one (segment, group): kernel.c:200-332.  Updates psi and
// returns the flux integral (the caller applies the 0.5 weight, an exact
// scaling, once per flushed run).  The first term of the integral,
// (q0 tau + (sigT psi - q0) ev) / sigT^2, is evaluated as
// q0 (tau - ev) / sigT^2 + psi ev / sigT, which shares ev / sigT and
// (tau - ev) / sigT^2 with the psi update and uses the cancellation-free
// tau - ev of ev_terms.
template <int EXPM>
__device__ __forceinline__ float attenuate(float sig, float q0, float q1, float q2, float ds,
                                           float &psi, float &one_m_ev, const ExpTab &tb)
{
    const float tau = sig * ds;
    const EvTerms x = ev_terms<EXPM>(tau, tb);
    const float ev = x.ev;
    const float i1 = rcp(sig);
    const float i2 = i1 * i1;
    const float e1 = ev * i1;     // ev / sigT
    const float m2 = x.tme * i2;  // (tau - ev) / sigT^2
    const float reuse = tau * (tau - 2.f) + 2.f * e1 * i2;
    const float q1m = q1 * SMOC_MU, q2m = q2 * SMOC_MU2;
    const float flux_integral =
        q0 * m2 + psi * e1 + q1m * reuse + q2m * x.poly * (i2 * i2 * (1.f / 3.f));
    psi = q0 * e1 + q1m * m2 + q2m * reuse + psi * x.one_m_ev;
    one_m_ev = x.one_m_ev;
    return flux_integral;
}

// ---- tally accumulators ------------------------------------------------------
// SINGLE: every segment is its own run (unsorted stream), so a deterministic
// partial is one FP32 tally and is stored as FP32 without loss.
template <int TALLY, bool SINGLE> struct Acc;
template <bool SINGLE> struct Acc<SMOC_TALLY_ATOMIC_F32, SINGLE> {
    float s = 0.f;
    __device__ __forceinline__ void add(float x) { s += x; }
    __device__ __forceinline__ void flush(void *flux, int64_t row, int64_t, int G, int g)
    {
        unsafeAtomicAdd(static_cast<float *>(flux) + row * G + g, SMOC_WEIGHT * s);
        s = 0.f;
    }
};
template <bool SINGLE> struct Acc<SMOC_TALLY_ATOMIC_F64, SINGLE> {
    double s = 0.0;
    __device__ __forceinline__ void add(float x) { s += (double)x; }
    __device__ __forceinline__ void flush(void *flux, int64_t row, int64_t, int G, int g)
    {
        unsafeAtomicAdd(static_cast<double *>(flux) + row * G + g, (double)SMOC_WEIGHT * s);
        s = 0.0;
    }
};
template <> struct Acc<SMOC_TALLY_DET_F64, false> {
    double s = 0.0;
    __device__ __forceinline__ void add(float x) { s += (double)x; }
    __device__ __forceinline__ void flush(void *part, int64_t, int64_t run, int G, int g)
    {
        static_cast<double *>(part)[run * G + g] = (double)SMOC_WEIGHT * s;
        s = 0.0;
    }
};
template <> struct Acc<SMOC_TALLY_DET_KAHAN, false> {
    float s = 0.f, c = 0.f;
    __device__ __forceinline__ void add(float x)
    {
#pragma clang fp contract(off)
        const float y = x - c;
        const float t = s + y;
        c = (t - s) - y;
        s = t;
    }
    __device__ __forceinline__ void flush(void *part, int64_t, int64_t run, int G, int g)
    {
        static_cast<float2 *>(part)[run * G + g] = make_float2(SMOC_WEIGHT * s, SMOC_WEIGHT * c);
        s = 0.f;
        c = 0.f;
    }
};
// 64-bit fixed point with 32 fractional bits: integer atomics are
// associative, so the sum is bit-exact whatever the order.  The host bounds
// every contribution by lim = (2^31 - 1) / (most segments in one region
// interval), so no sum can wrap; a contribution beyond lim (or NaN) raises
// the overflow counter and the run reports SMOC_ERR_RANGE.
constexpr double kFixScale = 4294967296.0;  // 2^32
template <bool SINGLE> struct Acc<SMOC_TALLY_FIXED64, SINGLE> {
    double s = 0.0;
    bool bad = false;
    float lim = 0.f;
    __device__ __forceinline__ void add(float x)
    {
        bad |= !(fabsf(x) <= lim);  // x is unweighted: lim is twice the bound
        s += (double)x;
    }
    __device__ __forceinline__ void flush(void *flux, int64_t row, int64_t, int G, int g)
    {
        if (bad) {
            atomicAdd(reinterpret_cast<unsigned int *>(static_cast<int64_t *>(flux) - 1), 1u);
            bad = false;
            s = 0.0;
        }
        const long long q = __double2ll_rn((double)SMOC_WEIGHT * s * kFixScale);
        atomicAdd(reinterpret_cast<unsigned long long *>(static_cast<int64_t *>(flux) + row * G + g),
                  (unsigned long long)q);
        s = 0.0;
    }
};
template <typename A> __device__ __forceinline__ void acc_init(A &, float) {}
template <bool SINGLE>
__device__ __forceinline__ void acc_init(Acc<SMOC_TALLY_FIXED64, SINGLE> &acc, float lim)
{
    acc.lim = lim / SMOC_WEIGHT;
}

// Block sums (region-sorted stream): the flux integrals of a block of up to 8
// segments of a run are summed in FP32, and the block sum joins the run's
// accumulator with one add (compensated for DET_KAHAN).  FIXED64 checks the
// block sum against n times its per-contribution bound: a block of n in-bound
// contributions passes, so no run of in-bound contributions can wrap.
template <typename A> __device__ __forceinline__ void acc_add_block(A &acc, float sum, int)
{
    acc.add(sum);
}
template <bool SINGLE>
__device__ __forceinline__ void acc_add_block(Acc<SMOC_TALLY_FIXED64, SINGLE> &acc, float sum,
                                              int n)
{
    acc.bad |= !(fabsf(sum) <= (float)n * acc.lim);
    acc.s += (double)sum;
}

template <int TALLY> struct AccSingle {
    float s = 0.f;
    __device__ __forceinline__ void add(float x) { s = x; }
    __device__ __forceinline__ void flush(void *part, int64_t, int64_t run, int G, int g)
    {
        static_cast<float *>(part)[run * G + g] = SMOC_WEIGHT * s;
    }
};
template <> struct Acc<SMOC_TALLY_DET_F64, true> : AccSingle<SMOC_TALLY_DET_F64> {};
template <> struct Acc<SMOC_TALLY_DET_KAHAN, true> : AccSingle<SMOC_TALLY_DET_KAHAN> {};

__device__ __forceinline__ uint32_t xcd_remap(uint32_t b, uint32_t nb)
{
    const uint32_t x = b & 7u, i = b >> 3, per = nb >> 3, rem = nb & 7u;
    return x < rem ? x * (per + 1) + i : rem * (per + 1) + (x - rem) * per + i;
}

// ---- the sweep ------------------------------------------------------------------
__device__ __forceinline__ int64_t uniform64(int64_t x)
{
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)x);
    const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)((uint64_t)x >> 32));
    return (int64_t)(((uint64_t)hi << 32) | lo);
}

// The segment-stream pointers reach sweep_item as separate restrict kernel
// parameters so that their wave-uniform loads compile to scalar loads.
//
// MODE_SWEEP: attenuate and tally.  MODE_XFER: the affine transfer of the
// item's segments, psi_out = M psi_in + B per group (kernel.c:289-332 is
// affine in psi), computed from psi_in = 0 and a running product of 1 - ev;
// no tally.  A scan over a track's chunks turns the transfers into every
// chunk's incoming psi.
//
// Chunked runs go in two launches.  Phase 1 sweeps every track's first chunk
// (its incoming psi is psi_in) and, in the same launch, runs the transfer pass
// of the middle chunks; phase 2, after the scan, sweeps the remaining chunks.
// MODE_MIXED kernels pick the mode per work entry from its ITEM_* flags.
//
// MODE_LIN (linear chunks; a region-sorted stream whose chunks begin runs,
// with any tally but FIXED64): every chunk after a track's first is swept once,
// in phase 1, from psi = 0.  A segment's tally is affine in the chunk's
// incoming psi: its psi term is psi_k ev_k / sigT with psi_k = psi0_k +
// psi_in P_k, P_k the product of 1 - ev over the chunk's earlier segments.
// The run's flush takes the tally from psi0 and wrun[run] the run's sum of
// P_k ev_k / sigT; the chunk emits its transfer (psi0_out, P_end).  After the
// scan, k_lin_fix adds psi_in times wrun to each run's tally and k_scan gives
// the tracks' psi_out.  No segment is swept twice.

enum { MODE_SWEEP = 0, MODE_XFER = 1, MODE_MIXED = 2, MODE_LIN = 3 };
enum { ITEM_XFER = SMOC_ITEM_XFER, ITEM_EMIT = SMOC_ITEM_EMIT, ITEM_LIN = SMOC_ITEM_LIN };

// Per-lane constants of one run (a track's consecutive segments in one
// (region, interval)): everything of kernel.c:111-332 that depends on the
// region tuple only.  With them a segment costs tau, ev and the two affine
// updates.
struct RunK {
    float sig, i1, i3x2;      // sigT, 1/sigT, 2/sigT^3
    float c0s, c1s, c2s;      // source fit / sigT: q0 / sigT = c0s + zin (c1s + c2s zin)
    float na1, m2s;           // -mu c1, 2 mu sigT:
                              // mu q1 = mu (c1 + 2 c2 zin) = 2 mu sigT (c1s + c2s zin) - mu c1
    float q2m;                // mu2 q2: the integral's q2 term is
                              // (2 mu2 q2 / sigT^4) poly / 6 = mu2 q2 (2 ds^4) r4
};

// The run loop's arithmetic (run_consts, attenuate_run, series_a, series_b and
// the per-segment powers) is written with explicit fused multiply-adds and no
// contraction, so every instantiation -- 1 or 2 groups per lane, any exp or
// tally variant -- performs the same operations in the same order and the
// deterministic tallies stay bit-exact across launch shapes.
template <int STORE>
__device__ __forceinline__ RunK run_consts(const float4 t, int kind)
{
#pragma clang fp contract(off)
    RunK k;
    float c0, c1, c2;
    k.sig = t.x;
    k.i1 = rcp(t.x);
    k.i3x2 = (2.f * (k.i1 * k.i1)) * k.i1;
    fit_coeffs<STORE>(t, kind, c0, c1, c2);
    k.c0s = c0 * k.i1;
    k.c1s = c1 * k.i1;
    k.c2s = c2 * k.i1;
    k.na1 = -SMOC_MU * c1;
    k.m2s = (2.f * SMOC_MU) * k.sig;
    k.q2m = SMOC_MU2 * c2;
    return k;
}

// attenuate() with the run constants hoisted: same formulas, the products
// of constants regrouped (rounding-level differences only).
template <int EXPM>
__device__ __forceinline__ float attenuate_run(const RunK &k, float ds, float zin,
                                               float &psi, float &one_m_ev, float &ev,
                                               const ExpTab &tb)
{
#pragma clang fp contract(off)
    const float tau = k.sig * ds;
    const EvTerms x = ev_terms<EXPM>(tau, tb);
    const float i2 = k.i1 * k.i1;  // 1/sigT^2 (this path is rare on the benchmark streams)
    const float m2 = x.tme * i2;
    const float reuse = __builtin_fmaf(x.ev, k.i3x2, tau * (tau - 2.f));
    const float inner = __builtin_fmaf(k.c2s, zin, k.c1s);
    const float q0s = __builtin_fmaf(inner, zin, k.c0s);  // q0 / sigT
    const float q1m = __builtin_fmaf(inner, k.m2s, k.na1);
    const float kfi6 = (2.f * k.q2m) * (i2 * i2);  // 2 mu2 q2 / sigT^4
    const float flux_integral = __builtin_fmaf(
        q0s * k.sig, m2,
        __builtin_fmaf(psi, x.ev * k.i1, __builtin_fmaf(q1m, reuse, kfi6 * (x.poly * (1.f / 6.f)))));
    psi = __builtin_fmaf(q0s, x.ev, __builtin_fmaf(q1m, m2, __builtin_fmaf(k.q2m, reuse,
                                                                          psi * x.one_m_ev)));
    one_m_ev = x.one_m_ev;
    ev = x.ev;
    return flux_integral;
}

// attenuate_run() for a segment on which every lane of the wave has tau < 1
// (ACCURATE exp, checked once per segment on the scalar unit): the series
// alone, with tau^2 shared by the series and reuse, and poly / 6 = tau^4 r4.
// Split in two: series_a holds everything that does not depend on psi, so a
// wave can evaluate it for a block of segments before the psi updates run one
// after the other (series_b).  HAS_C2 = false: a first or last fine interval, whose
// fit is linear (c2 = 0, so q2 = 0 and q1 = c1): the terms in c2 are exactly
// zero and skipped, with the same results.
struct SegA {
    float ev;   // 1 - e^-tau
    float src;  // the source part S of psi' = S + psi (1 - ev)
};

// ds2 = ds^2 and ds4x2 = 2 ds^4 are the segment's (shared by every group):
// tau - ev = tau^2 r, so (tau - ev) / sigT^2 = ds^2 r, and the q2 term of
// the integral, (2 mu2 q2 / sigT^4) tau^4 r4, is mu2 q2 (2 ds^4) r4.
// With q0 / sigT from the fit (c0s, c1s, c2s), the source's q0 term is
// (q0 / sigT) ev and needs no ev / sigT; reuse takes ev times 2 / sigT^3.
// The psi-free part of the flux integral goes into three block sums, each a
// single fused multiply-add per segment, whose run-constant factors are applied
// once per block: b0 = sum (q0 / sigT) m2 (times sigT), bsum = sum mu q1 reuse
// (times 1) and s2 = sum (2 ds^4) r4 (times mu2 q2).
template <bool HAS_C2>
__device__ __forceinline__ SegA series_a(const RunK &k, float ds, float ds2, float ds4x2, float zin,
                                         float &b0, float &bsum, float &s2)
{
#pragma clang fp contract(off)
    const float tau = k.sig * ds;
    const float r4 = r4_series(tau);
    const float t2 = tau * tau;
    // r = 1/2 - tau/6 + tau^2 r4 in Horner form: the -1/6 is the literal of a
    // two-operand FMA and 1/2 an inline constant, so no constant is held in an
    // SGPR (VALU instructions that read SGPRs pair less often)
    const float r = __builtin_fmaf(tau, __builtin_fmaf(tau, r4, -1.f / 6.f), 0.5f);
    SegA a;
    a.ev = __builtin_fmaf(-t2, r, tau);  // one fused rounding
    const float m2 = ds2 * r;  // (tau - ev) / sigT^2
    const float reuse = __builtin_fmaf(a.ev, k.i3x2, __builtin_fmaf(-2.f, tau, t2));
    if constexpr (HAS_C2) {
        // Horner in zin: no zin^2 per segment; mu q1 from the same inner term
        const float inner = __builtin_fmaf(k.c2s, zin, k.c1s);
        const float q0s = __builtin_fmaf(inner, zin, k.c0s);  // q0 / sigT
        const float q1m = __builtin_fmaf(inner, k.m2s, k.na1);
        b0 = __builtin_fmaf(q0s, m2, b0);
        bsum = __builtin_fmaf(q1m, reuse, bsum);
        s2 = __builtin_fmaf(ds4x2, r4, s2);
        a.src = __builtin_fmaf(q0s, a.ev, __builtin_fmaf(q1m, m2, k.q2m * reuse));
    } else {
        const float q0s = __builtin_fmaf(k.c1s, zin, k.c0s);
        b0 = __builtin_fmaf(q0s, m2, b0);
        bsum = __builtin_fmaf(-k.na1, reuse, bsum);
        a.src = __builtin_fmaf(q0s, a.ev, -k.na1 * m2);
    }
    return a;
}

// psi' = S + psi (1 - ev) as psi + (S - psi ev): the small change is formed
// first (one FMA), then added to psi, so only one rounding happens at the
// magnitude of psi (a run near its steady state psi = S / ev would otherwise
// accumulate the roundings of every term).  The psi term of the flux
// integral, psi ev / sigT, is summed as psi ev into psum; the block multiplies
// that sum by 1 / sigT once, so a segment keeps two values (ev, S) instead of
// three between series_a and series_b.
__device__ __forceinline__ void series_b(const SegA &a, float &psi, float &psum)
{
#pragma clang fp contract(off)
    psum = __builtin_fmaf(psi, a.ev, psum);
    psi = psi + __builtin_fmaf(-psi, a.ev, a.src);
}

template <int STORE, int GPL, int EXPM, int TALLY, bool RUNS, int MODE>
__device__ __forceinline__ void sweep_item(const smoc_sweep_args &a,
                                           const float *__restrict__ len,
                                           const int32_t *__restrict__ code,
                                           const float *__restrict__ zin,
                                           const int64_t *__restrict__ track_off,
                                           const int64_t *__restrict__ items, int64_t item,
                                           int lane, const ExpTab &tb)
{
    using raw_t = typename Tup<STORE>::raw;
    const int64_t w = a.slices == 1 ? item : item / a.slices;
    const int sl = (int)(item - w * a.slices);
    // list entries: item id << 2 | ITEM_* flags
    const int64_t entry = a.item_list ? uniform64(a.item_list[w]) : (w << 2);
    const int64_t it = entry >> 2;
    const bool emit_xfer = (entry & ITEM_EMIT) != 0;

    int64_t t, kb, ke, run = 0;
    if (items) {
        t = uniform64(items[4 * it]);
        kb = uniform64(items[4 * it + 1]);
        ke = uniform64(items[4 * it + 2]);
        run = uniform64(items[4 * it + 3]);
    } else {
        t = it;
        kb = uniform64(track_off[t]);
        ke = uniform64(track_off[t + 1]);
        if constexpr (TALLY == SMOC_TALLY_DET_F64 || TALLY == SMOC_TALLY_DET_KAHAN)
            run = uniform64(a.track_run_off[t]);
    }

    // Device rows are padded to Gp >= slices * 64 * GPL groups.  Lanes past G
    // load the last valid group's data (same cache lines, no extra traffic),
    // compute on it and never flush; their psi stores land in padding.
    const int Gp = a.Gp, G = a.G;
    int g[GPL], gi[GPL];
    // mprod: the running product of 1 - ev (MODE_XFER, MODE_LIN); wacc: the
    // run's sum of mprod ev / sigT (MODE_LIN)
    float psi[GPL], mprod[GPL];
    double wacc[GPL];
    Acc<TALLY, !RUNS> acc[GPL];
    const float *pin = a.chunk_psi ? a.chunk_psi + it * Gp : a.psi_in + t * Gp;
#pragma unroll
    for (int j = 0; j < GPL; j++) {
        g[j] = sl * 64 * GPL + j * 64 + lane;
        gi[j] = g[j] < G ? g[j] : G - 1;
        psi[j] = MODE == MODE_XFER || MODE == MODE_LIN ? 0.f : pin[gi[j]];
        mprod[j] = 1.f;
        wacc[j] = 0.0;
        acc_init(acc[j], a.fix_limit);
    }

    const raw_t *__restrict__ tup = static_cast<const raw_t *>(a.tuples);
    // segment indices fit 32 bits (S <= 2^31 - 1 is checked on upload), so
    // the loop bounds are 32-bit scalar compares; k + kSeries cannot wrap
    // unsigned
    const uint32_t kb32 = (uint32_t)kb, ke32 = (uint32_t)ke;
    // the stream entry of segment k, marked past-the-end (code -1) beyond the
    // item: loaded unconditionally (the arrays are padded and k <= S + 1), so
    // only the mark is a select
    auto raw_at = [&](uint32_t k) {
        RawSeg r = load_raw(len, code, zin, (int64_t)k);
        r.code = k < ke32 ? r.code : -1;
        return r;
    };
    auto flush_all = [&](int row) {
#pragma unroll
        for (int j = 0; j < GPL; j++)
            if (g[j] < G)
                acc[j].flush(a.flux, row, run, Gp, g[j]);
        ++run;
    };

    if constexpr (!RUNS) {
        // Unsorted stream: every segment is its own run.  Software pipeline
        // over the item's segments: while segment k is computed from tuple
        // buffer X (or Y), the tuples of k+1 are in flight into the other
        // buffer and the stream entries of k+2 into scalar registers.  Two
        // explicit buffers avoid register copies that would force a wait on
        // the freshly issued gathers.  Gathers and flushes are unconditional
        // so that the compiler's memory-counter waits stay exact: a load
        // issued on only one side of a branch makes every later wait assume
        // the shorter path and stall on the loads just issued.  Past the end,
        // row 0 is read and unused.
        auto gather = [&](raw_t (&buf)[GPL], const Seg &sg) {
            const int64_t row = sg.row >= 0 ? sg.row : 0;
            const raw_t *__restrict__ base = tup + row * Gp;  // uniform base + lane offset
#pragma unroll
            for (int j = 0; j < GPL; j++)
                buf[j] = base[(uint32_t)gi[j]];
        };
        auto step = [&](const raw_t (&buf)[GPL], const Seg &cur) {
            const float zin2 = cur.zin * cur.zin;
#pragma unroll
            for (int j = 0; j < GPL; j++) {
                const float4 r = unpack<STORE>(buf[j]);
                float q0, q1, q2, m;
                fit<STORE>(r, cur.kind, cur.zin, zin2, q0, q1, q2);
                const float tally = attenuate<EXPM>(r.x, q0, q1, q2, cur.ds, psi[j], m, tb);
                if constexpr (MODE == MODE_XFER)
                    mprod[j] *= m;
                else
                    acc[j].add(tally);
            }
            if constexpr (MODE == MODE_SWEEP)
                flush_all(cur.row);
        };
        if (kb32 < ke32) {
            raw_t X[GPL], Y[GPL];
#pragma unroll
            for (int j = 0; j < GPL; j++)
                X[j] = Y[j] = Tup<STORE>::one();
            Seg s0 = decode(raw_at(kb32));
            gather(X, s0);
            Seg s1 = decode(raw_at(kb32 + 1));
            for (uint32_t k = kb32;;) {
                gather(Y, s1);
                const RawSeg r2 = raw_at(k + 2);
                step(X, s0);
                const Seg s2 = decode(r2);
                if (++k >= ke32)
                    break;
                gather(X, s2);
                const RawSeg r3 = raw_at(k + 2);
                step(Y, s1);
                const Seg s3 = decode(r3);
                if (++k >= ke32)
                    break;
                s0 = s2;
                s1 = s3;
            }
        }
    } else {
        // Region-sorted stream: a track stays in one (region, interval) for
        // many segments.  The outer loop starts a run: it gathers the tuple,
        // derives the run constants and, at the end, flushes the run's tally
        // summed in registers.  The inner loop walks the run in blocks: the
        // stream entries of the next kSeries segments are loaded ahead (the
        // device arrays are padded, so loads past the item's end are in
        // bounds; blocks never reach past it), and up to kSeries consecutive
        // segments that all take the series branch have their psi-free parts
        // evaluated together before their psi updates run in order.  On a
        // stream that can be chunked (a.split) a run also ends before every
        // stream index that is a multiple of SMOC_RUN_SPLIT (a split point;
        // the run table splits there too).  Chunks of a multiple of that
        // length are cut at split points, so each begins a run, and the blocks
        // and the sums do not depend on the chunking.
        if (kb32 < ke32) {
            RunK rk[GPL];
            uint32_t k = kb32;
            RawSeg cur = load_raw(len, code, zin, k);
            for (;;) {
                const int ccode = cur.code;
                // the run ends by rend at the latest: the item's end or, on a
                // stream that can be chunked, the next split point (one min per
                // run: capping each block by both would compile to a vector
                // min3 of scalars)
                const uint32_t rend =
                    a.split ? min(ke32, (k | (SMOC_RUN_SPLIT - 1)) + 1) : ke32;
                {
                    const raw_t *__restrict__ base = tup + (int64_t)(ccode >> 2) * Gp;
                    raw_t b[GPL];
#pragma unroll
                    for (int j = 0; j < GPL; j++)
                        b[j] = base[(uint32_t)gi[j]];
#pragma unroll
                    for (int j = 0; j < GPL; j++)
                        rk[j] = run_consts<STORE>(unpack<STORE>(b[j]), ccode & 3);
                }
                // per-row threshold: the same for every launch shape, so the
                // branch taken (and the rounding) does not depend on which
                // groups share the wave
                const uint32_t thr =
                    EXPM == SMOC_EXP_ACCURATE ? __float_as_uint(a.row_thr[ccode >> 2]) : 0u;
                // Scalar loads may return out of order, so a wait for any of
                // them is lgkmcnt(0).  Complete the threshold's load here:
                // pending on entry to the segment loop, it would make every
                // iteration wait for the loads just issued for the next
                // segment, defeating the prefetch.
                __builtin_amdgcn_s_waitcnt(kWaitLgkm0);
                // the run's segments; the fit kind picks the specialization
                // once per run (a linear fit drops its zero c2 terms)
                auto walk = [&](auto has_c2) {
                    constexpr bool HAS_C2 = decltype(has_c2)::value;
                    auto series_ok = [&](const RawSeg &sg) {
                        return EXPM == SMOC_EXP_ACCURATE && __float_as_uint(sg.len) < thr;
                    };
                    // blocks of up to NB = 8 segments (the same length for
                    // every tally and launch shape, so the block sums and the
                    // results do not depend on the shape); 1 where the exp
                    // mode has no series branch (one segment a step)
                    constexpr int NB = EXPM != SMOC_EXP_ACCURATE ? 1 : kSeries;
                    for (;;) {
                        // the next NB segments, from one 64-bit base per array
                        // so that the scalar loads take constant offsets
                        // (the arrays are padded, so the loads past the
                        // item's end are in bounds; such entries are never
                        // used: blocks are capped by the segments left)
                        const float *__restrict__ lp = len + (size_t)k + 1;
                        const int32_t *__restrict__ cp = code + (size_t)k + 1;
                        const float *__restrict__ zp = zin + (size_t)k + 1;
                        RawSeg nx[NB];
#pragma unroll
                        for (int u = 0; u < NB; u++)
                            nx[u] = RawSeg{lp[u], zp[u], cp[u]};
                        if (series_ok(cur)) {
                            // up to NB segments of the run at once when
                            // all take the series: their psi-free parts are
                            // independent of each other.  Common case first,
                            // in a few scalar operations: every look-ahead
                            // entry continues the run below the threshold.
                            // len bits < thr <=> thr - 1 - bits has its sign
                            // bit clear: both are below 2^31 (len >= 0 is
                            // checked on upload, thr > 0), and -0.0, whose
                            // bits are 2^31, sets it as series_ok() rejects it
                            uint32_t diff = 0, below = 0;
#pragma unroll
                            for (int u = 0; u < NB - 1; u++) {
                                diff |= (uint32_t)(nx[u].code ^ ccode);
                                below |= thr - 1u - __float_as_uint(nx[u].len);
                            }
                            uint32_t n = NB;
                            if ((diff | (below & 0x80000000u)) != 0) {
                                n = 1;
                                while (n < NB && nx[n - 1].code == ccode &&
                                       series_ok(nx[n - 1]))
                                    ++n;
                            }
                            n = min(n, rend - k);  // segments k .. k + n - 1 are the run's
                            auto block = [&](auto cnt) {
                                constexpr int N = decltype(cnt)::value;
                                SegA a[N][GPL];
                                // block sums: series_a's three psi-free parts,
                                // series_b's sum of psi ev and (MODE_LIN) the
                                // sum of mprod ev
                                float b0[GPL], bs[GPL], s2[GPL], ps[GPL], ws[GPL];
#pragma unroll
                                for (int j = 0; j < GPL; j++)
                                    b0[j] = bs[j] = s2[j] = ps[j] = ws[j] = 0.f;
#pragma unroll
                                for (int u = 0; u < N; u++) {
                                    const RawSeg &sg = u == 0 ? cur : nx[u - 1];
                                    const float d2 = sg.len * sg.len;
                                    const float d4x2 = HAS_C2 ? (2.f * d2) * d2 : 0.f;
#pragma unroll
                                    for (int j = 0; j < GPL; j++)
                                        a[u][j] = series_a<HAS_C2>(rk[j], sg.len, d2, d4x2, sg.zin,
                                                                   b0[j], bs[j], s2[j]);
                                }
#pragma unroll
                                for (int j = 0; j < GPL; j++) {
#pragma unroll
                                    for (int u = 0; u < N; u++) {
                                        if constexpr (MODE == MODE_XFER)
                                            mprod[j] *= 1.f - a[u][j].ev;
                                        if constexpr (MODE == MODE_LIN) {
                                            ws[j] = __builtin_fmaf(mprod[j], a[u][j].ev, ws[j]);
                                            mprod[j] = __builtin_fmaf(-mprod[j], a[u][j].ev, mprod[j]);
                                        }
                                        series_b(a[u][j], psi[j], ps[j]);
                                    }
                                    // the block's flux integrals, summed in FP32, join
                                    // the run's accumulator once:
                                    // bsum + mu2 q2 s2 + sigT b0 + (sum psi ev) / sigT
                                    if constexpr (MODE != MODE_XFER) {
                                        float t = bs[j];
                                        if constexpr (HAS_C2)
                                            t = __builtin_fmaf(s2[j], rk[j].q2m, t);
                                        t = __builtin_fmaf(b0[j], rk[j].sig, t);
                                        acc_add_block(acc[j], __builtin_fmaf(ps[j], rk[j].i1, t), N);
                                        if constexpr (MODE == MODE_LIN)
                                            wacc[j] += (double)(ws[j] * rk[j].i1);
                                    }
                                }
                                k += N;
                                cur = nx[N - 1];
                            };
                            // the largest power of two <= n (and <= NB)
                            if (NB >= 8 && n >= 8)
                                block(std::integral_constant<int, (NB >= 8 ? 8 : 1)>{});
                            else if (NB >= 4 && n >= 4)
                                block(std::integral_constant<int, (NB >= 4 ? 4 : 1)>{});
                            else if (NB >= 2 && n >= 2)
                                block(std::integral_constant<int, (NB >= 2 ? 2 : 1)>{});
                            else
                                block(std::integral_constant<int, 1>{});
                        } else {
#pragma unroll
                            for (int j = 0; j < GPL; j++) {
                                float m, ev;
                                const float fi = attenuate_run<EXPM>(rk[j], cur.len, cur.zin,
                                                                     psi[j], m, ev, tb);
                                if constexpr (MODE == MODE_XFER) {
                                    mprod[j] *= m;
                                } else {
                                    acc[j].add(fi);
                                    if constexpr (MODE == MODE_LIN) {
                                        wacc[j] += (double)((mprod[j] * ev) * rk[j].i1);
                                        mprod[j] *= m;
                                    }
                                }
                            }
                            ++k;
                            cur = nx[0];
                        }
                        if (k >= rend || cur.code != ccode)
                            break;
                    }
                };
                if ((ccode & 3) == KIND_INTERIOR)
                    walk(std::true_type{});
                else
                    walk(std::false_type{});
                if constexpr (MODE == MODE_LIN) {
#pragma unroll
                    for (int j = 0; j < GPL; j++) {
                        if (g[j] < G)
                            a.wrun[run * Gp + g[j]] = (float)wacc[j];
                        wacc[j] = 0.0;
                    }
                }
                if constexpr (MODE == MODE_SWEEP || MODE == MODE_LIN)
                    flush_all(ccode >> 2);
                if (k >= ke32)
                    break;
            }
        }
    }

    if (MODE == MODE_XFER || MODE == MODE_LIN || emit_xfer) {
        // transfer of this chunk; a swept first chunk knows its psi_out
        // exactly and passes it on as B with M = 0
        float *xb = a.xfer + (2 * it) * Gp, *xm = a.xfer + (2 * it + 1) * Gp;
#pragma unroll
        for (int j = 0; j < GPL; j++) {
            xb[g[j]] = psi[j];
            xm[g[j]] = MODE == MODE_XFER || MODE == MODE_LIN ? mprod[j] : 0.f;
        }
    } else if (!items || ke == uniform64(track_off[t + 1])) {
        // chunks other than a track's last carry no psi_out
        float *pout = a.psi_out + t * Gp;
#pragma unroll
        for (int j = 0; j < GPL; j++)
            pout[g[j]] = psi[j];
    }
}

template <int STORE, int GPL, int EXPM, int TALLY, bool RUNS, int MODE>
__device__ __forceinline__ void dispatch_item(const smoc_sweep_args &a,
                                              const float *__restrict__ len,
                                              const int32_t *__restrict__ code,
                                              const float *__restrict__ zin,
                                              const int64_t *__restrict__ track_off,
                                              const int64_t *__restrict__ items, int64_t item,
                                              int lane, const ExpTab &tb)
{
    if constexpr (MODE == MODE_MIXED) {
        const int64_t w = a.slices == 1 ? item : item / a.slices;
        const int flags = (int)(uniform64(a.item_list[w]) & 3);
        constexpr bool LIN = RUNS && TALLY != SMOC_TALLY_FIXED64;
        if (LIN && flags == ITEM_LIN) {
            if constexpr (LIN)
                sweep_item<STORE, GPL, EXPM, TALLY, RUNS, MODE_LIN>(a, len, code, zin, track_off,
                                                                    items, item, lane, tb);
        } else if (flags == ITEM_XFER) {
            sweep_item<STORE, GPL, EXPM, TALLY, RUNS, MODE_XFER>(a, len, code, zin, track_off,
                                                                 items, item, lane, tb);
        } else {
            sweep_item<STORE, GPL, EXPM, TALLY, RUNS, MODE_SWEEP>(a, len, code, zin, track_off,
                                                                  items, item, lane, tb);
        }
    } else {
        sweep_item<STORE, GPL, EXPM, TALLY, RUNS, MODE>(a, len, code, zin, track_off, items,
                                                        item, lane, tb);
    }
}

template <int EXPM>
__device__ __forceinline__ ExpTab stage_table(const smoc_sweep_args &a, float *s_tab)
{
    ExpTab tb{nullptr, 0.f, 0.f, 0.f, 0};
    if constexpr (EXPM == SMOC_EXP_TABLE) {
        for (int i = threadIdx.x; i < 2 * a.tab_n; i += blockDim.x)
            s_tab[i] = a.table[i];
        __syncthreads();
        tb.lds = s_tab;
        tb.dx = a.tab_dx;
        tb.hdx = 0.5f * a.tab_dx;
        tb.tmax = a.tab_max;
        tb.n = a.tab_n;
    }
    return tb;
}

// One wave per work item, XCD-aware block order.
template <int STORE, int GPL, int EXPM, int TALLY, bool RUNS, int MODE>
__global__ void __launch_bounds__(1024)
    k_sweep(smoc_sweep_args a, const float *__restrict__ len, const int32_t *__restrict__ code,
            const float *__restrict__ zin, const int64_t *__restrict__ track_off,
            const int64_t *__restrict__ items)
{
    __shared__ float s_tab[EXPM == SMOC_EXP_TABLE ? SMOC_TABLE_CAP : 1];
    const ExpTab tb = stage_table<EXPM>(a, s_tab);
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int wpb = blockDim.x >> 6;
    const int64_t item = (int64_t)xcd_remap(blockIdx.x, gridDim.x) * wpb + wave;
    if (item >= a.n_items * a.slices)
        return;
    dispatch_item<STORE, GPL, EXPM, TALLY, RUNS, MODE>(a, len, code, zin, track_off, items, item,
                                                       lane, tb);
}

// Persistent variant: each wave pulls work items from a device counter until
// none remain; every wave exits after at most n_work + 1 pulls.
template <int STORE, int GPL, int EXPM, int TALLY, bool RUNS, int MODE>
__global__ void __launch_bounds__(1024)
    k_sweep_persistent(smoc_sweep_args a, const float *__restrict__ len,
                       const int32_t *__restrict__ code, const float *__restrict__ zin,
                       const int64_t *__restrict__ track_off, const int64_t *__restrict__ items)
{
    __shared__ float s_tab[EXPM == SMOC_EXP_TABLE ? SMOC_TABLE_CAP : 1];
    const ExpTab tb = stage_table<EXPM>(a, s_tab);
    const int lane = threadIdx.x & 63;
    const int64_t n_work = a.n_items * a.slices;
    // a wave takes queue_batch consecutive items per counter update: with
    // millions of short items one counter updated per item serialises the
    // grid on that address; the counter only grows, so every wave reaches
    // first >= n_work and exits
    const int64_t batch = a.queue_batch > 0 ? a.queue_batch : 1;
    for (;;) {
        unsigned long long got = 0;
        if (lane == 0)
            got = atomicAdd(reinterpret_cast<unsigned long long *>(a.queue),
                            (unsigned long long)batch);
        const int64_t first = uniform64((int64_t)__shfl(got, 0));
        if (first >= n_work)
            break;
        const int64_t last = first + batch < n_work ? first + batch : n_work;
        for (int64_t item = first; item < last; ++item)
            dispatch_item<STORE, GPL, EXPM, TALLY, RUNS, MODE>(a, len, code, zin, track_off,
                                                               items, item, lane, tb);
    }
}

// Scan of chunk transfers along each track: chunk_psi[first chunk] = psi_in,
// chunk_psi[c+1] = B_c + M_c chunk_psi[c]; with linear chunks, the psi_out
// of a track of several chunks is B + M psi_in of its last chunk.  One wave
// per (track, slice).
template <int GPL>
__global__ void __launch_bounds__(256) k_scan(smoc_scan_args a)
{
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int64_t item = (int64_t)blockIdx.x * (blockDim.x >> 6) + wave;
    if (item >= a.T * a.slices)
        return;
    const int64_t t = item / a.slices;
    const int sl = (int)(item - t * a.slices);
    const int Gp = a.Gp;
    const int64_t cb = a.track_item_off[t], ce = a.track_item_off[t + 1];
    // the outputs do not overlap the transfers, so the loads of a group of
    // chunks go out together ahead of the chain of FMAs (a long track has
    // about a hundred chunks)
    const float *__restrict__ xf = a.xfer;
    float *__restrict__ cpsi = a.chunk_psi;
#pragma unroll
    for (int j = 0; j < GPL; j++) {
        const int g = sl * 64 * GPL + j * 64 + lane;
        float psi = a.psi_in[t * Gp + g];
        int64_t c = cb;
        for (; c + 8 <= ce; c += 8) {
            float bv[8], mv[8];
#pragma unroll
            for (int u = 0; u < 8; u++) {
                bv[u] = xf[(2 * (c + u)) * Gp + g];
                mv[u] = xf[(2 * (c + u) + 1) * Gp + g];
            }
#pragma unroll
            for (int u = 0; u < 8; u++) {
                cpsi[(c + u) * Gp + g] = psi;
                if (c + u + 1 < ce)
                    psi = bv[u] + mv[u] * psi;
            }
        }
        for (; c < ce; ++c) {
            cpsi[c * Gp + g] = psi;
            if (c + 1 < ce)
                psi = xf[(2 * c) * Gp + g] + xf[(2 * c + 1) * Gp + g] * psi;
        }
        if (a.psi_out && ce - cb > 1)
            a.psi_out[t * Gp + g] = xf[(2 * (ce - 1)) * Gp + g] +
                                    xf[(2 * (ce - 1) + 1) * Gp + g] * psi;
    }
}

// Linear chunks: each run r of a chunk c swept from psi = 0 gets w psi_in(c)
// wrun[r] added (w the track weight): to its partial in the partial's own
// arithmetic (FP64, or a compensated FP32 step), or with an atomic add to its
// row of fine_flux.  One wave per (list entry, slice); every run belongs to
// one chunk, so each partial is written once.
template <int TALLY, int GPL>
__global__ void __launch_bounds__(256) k_lin_fix(smoc_fix_args a)
{
#pragma clang fp contract(off)
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int64_t item = (int64_t)blockIdx.x * (blockDim.x >> 6) + wave;
    if (item >= a.n_list * a.slices)
        return;
    const int64_t w = item / a.slices;
    const int sl = (int)(item - w * a.slices);
    const int Gp = a.Gp;
    const int64_t c = uniform64(a.list[w]) >> 2;
    // items are in track order and a track's chunks in stream order, and runs
    // are numbered the same way, so chunk c's runs end where the next item's
    // begin (the last item's at n_runs)
    const int64_t r0 = uniform64(a.items[4 * c + 3]);
    const int64_t r1 = c + 1 < a.n_items ? uniform64(a.items[4 * (c + 1) + 3]) : a.n_runs;
#pragma unroll
    for (int j = 0; j < GPL; j++) {
        const int g = sl * 64 * GPL + j * 64 + lane;
        if (g >= a.G)
            continue;
        const float pin = a.chunk_psi[c * Gp + g];
        for (int64_t r = r0; r < r1; ++r) {
            const float wr = a.wrun[r * Gp + g];
            if constexpr (TALLY == SMOC_TALLY_ATOMIC_F32) {
                const int64_t row = a.run_row[r];
                unsafeAtomicAdd(static_cast<float *>(a.out) + row * Gp + g,
                                SMOC_WEIGHT * (pin * wr));
            } else if constexpr (TALLY == SMOC_TALLY_ATOMIC_F64) {
                const int64_t row = a.run_row[r];
                unsafeAtomicAdd(static_cast<double *>(a.out) + row * Gp + g,
                                (double)SMOC_WEIGHT * ((double)pin * (double)wr));
            } else if constexpr (TALLY == SMOC_TALLY_DET_F64) {
                double *p = static_cast<double *>(a.out) + r * Gp + g;
                *p = *p + (double)SMOC_WEIGHT * ((double)pin * (double)wr);
            } else {
                float2 *p = static_cast<float2 *>(a.out) + r * Gp + g;
                float2 v = *p;  // (s, c): s - c, as the run accumulator keeps it
                const float y = SMOC_WEIGHT * (pin * wr) - v.y;
                const float s = v.x + y;
                v.y = (s - v.x) - y;
                v.x = s;
                *p = v;
            }
        }
    }
}

// ---- chunk plan built on the device ---------------------------------------------
// The work items of a chunked sweep, written without host memory (the
// allocation-free entry point builds them in its scratch on every call).  Track
// t is cut into smoc_chunk_count chunks (smoc_plan.h), items [i][4] = (track,
// begin, end, first run id), list1 holds each track's first chunk (ITEM_EMIT
// when it has more) and its middle chunks (ITEM_XFER), list2 every chunk after
// the first, tio[t] the first chunk of track t and tio[T] the chunk count.  The
// first run id is begin, or, given the run tables of a region-sorted stream (C
// a multiple of SMOC_RUN_SPLIT, so every chunk after a track's first begins at
// a split point), track_run_off[t] for the first chunk and the split_run entry
// of the chunk's split point for the others; split_run holds, track by track,
// the run id at each of the track's smoc_split_count split points.  With the
// run tables the chunks are linear (MODE_LIN): list1 then holds every chunk,
// those after a track's first with ITEM_LIN.  These are the items of the host
// plan; the lists are in track order.  One workgroup of 1024 threads: thread i
// takes a contiguous range of tracks, counts its chunks, list entries and
// split points, a scan over the threads gives its offsets, then it writes.
__global__ void __launch_bounds__(1024)
    k_plan(const int64_t *__restrict__ track_off, int64_t T, int C, int64_t *__restrict__ items,
           int64_t *__restrict__ list1, int64_t *__restrict__ list2, int64_t *__restrict__ tio,
           const int64_t *__restrict__ track_run_off, const int64_t *__restrict__ split_run)
{
    __shared__ int64_t s_n[1024], s_1[1024], s_2[1024], s_s[1024];
    const int tid = threadIdx.x;
    const int64_t per = (T + 1023) / 1024;
    const int64_t t0 = min(T, (int64_t)tid * per), t1 = min(T, t0 + per);
    auto chunks = [&](int64_t t) { return smoc_chunk_count(track_off[t], track_off[t + 1], C); };
    auto splits = [&](int64_t t) { return smoc_split_count(track_off[t], track_off[t + 1]); };
    // linear chunks (given the run tables): every chunk in list1
    const bool lin = split_run != nullptr;
    int64_t n = 0, n1 = 0, n2 = 0, ns = 0;
    for (int64_t t = t0; t < t1; ++t) {
        const int64_t nc = chunks(t);
        n += nc;
        n1 += lin ? nc : nc > 1 ? nc - 1 : 1;
        n2 += nc - 1;
        ns += splits(t);
    }
    s_n[tid] = n;
    s_1[tid] = n1;
    s_2[tid] = n2;
    s_s[tid] = ns;
    __syncthreads();
    for (int d = 1; d < 1024; d <<= 1) {  // inclusive scan over the threads
        int64_t a = 0, b = 0, c = 0, e = 0;
        if (tid >= d) {
            a = s_n[tid - d];
            b = s_1[tid - d];
            c = s_2[tid - d];
            e = s_s[tid - d];
        }
        __syncthreads();
        s_n[tid] += a;
        s_1[tid] += b;
        s_2[tid] += c;
        s_s[tid] += e;
        __syncthreads();
    }
    int64_t i = s_n[tid] - n, p1 = s_1[tid] - n1, p2 = s_2[tid] - n2, sp = s_s[tid] - ns;
    for (int64_t t = t0; t < t1; ++t) {
        const int64_t b = track_off[t], e = track_off[t + 1], nc = chunks(t);
        tio[t] = i;
        for (int64_t k = 0; k < nc; ++k) {
            const int64_t cb = smoc_chunk_begin(b, k, C), ce = smoc_chunk_end(b, e, k, C);
            int64_t *it = items + 4 * (i + k);
            it[0] = t;
            it[1] = cb;
            it[2] = ce;
            it[3] = !split_run ? cb
                    : k == 0   ? track_run_off[t]
                               : split_run[sp + smoc_split_index(b, cb) - 1];
        }
        list1[p1++] = (i << 2) | (nc > 1 ? ITEM_EMIT : 0);
        for (int64_t k = 1; k < (lin ? nc : nc - 1); ++k)
            list1[p1++] = ((i + k) << 2) | (lin ? ITEM_LIN : ITEM_XFER);
        for (int64_t k = 1; k < nc; ++k)
            list2[p2++] = (i + k) << 2;
        i += nc;
        sp += splits(t);
    }
    if (tid == 1023)  // the last thread's range ends at T: i is the chunk count
        tio[T] = i;
}

// ---- packing of region data (once per smoc_set_regions) ------------------------
template <int STORE>
__global__ void __launch_bounds__(256) k_pack(smoc_pack_args a)
{
#pragma clang fp contract(off)
    const int64_t n = (int64_t)a.R * a.F * a.Gp;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const int G = a.G, Gp = a.Gp, F = a.F;
    const int g = (int)(i % Gp);
    const int64_t rf = i / Gp;
    using raw_t = typename Tup<STORE>::raw;
    if (g >= G) {  // padding lanes compute on sigT = 1, zero source
        reinterpret_cast<raw_t *>(a.tuples)[i] = Tup<STORE>::one();
        return;
    }
    const int f = (int)(rf % F);
    const int64_t r = rf / F;
    const float sig = a.sigT[r * G + g];
    const float *src = a.fine_source + r * F * G;
    const float ylo = f > 0 ? src[(int64_t)(f - 1) * G + g] : 0.f;
    const float ymid = src[(int64_t)f * G + g];
    const float yhi = f < F - 1 ? src[(int64_t)(f + 1) * G + g] : 0.f;
    if constexpr (STORE == SMOC_STORE_F32) {
        // literal source fit coefficients, IEEE divisions (kernel.c:128-189)
        float c1, c2;
        if (f == 0) {
            c1 = (yhi - ymid) / kDz;
            c2 = 0.f;
        } else if (f == F - 1) {
            c1 = (ymid - ylo) / kDz;
            c2 = 0.f;
        } else {
            c1 = (ylo - yhi) / (2.f * kDz);
            c2 = (ylo - 2.f * ymid + yhi) / (2.f * kDz * kDz);
        }
        reinterpret_cast<float4 *>(a.tuples)[i] = make_float4(sig, ymid, c1, c2);
    } else if constexpr (STORE == SMOC_STORE_BF16) {
        reinterpret_cast<uint2 *>(a.tuples)[i] =
            make_uint2(f_to_bf(sig) | (f_to_bf(ylo) << 16), f_to_bf(ymid) | (f_to_bf(yhi) << 16));
    } else {
        reinterpret_cast<uint2 *>(a.tuples)[i] =
            make_uint2(f_to_h(sig) | (f_to_h(ylo) << 16), f_to_h(ymid) | (f_to_h(yhi) << 16));
    }
}

// Series threshold of each row (region, interval), for the series branch: with
// ds below it, sigT ds < 1 for every group of the row.  thr = (1 - 2^-20) /
// max_g sigT (IEEE division, sigT as the sweep unpacks it) gives sigT ds <
// (1 - 2^-20)(1 + 2^-24) < 1 - 2^-21 before rounding, so the FP32 product
// stays below 1.  The sweep compares float bits as unsigned integers, which
// orders non-negative floats numerically.  One wave per row.
template <int STORE>
__global__ void __launch_bounds__(256) k_row_thr(const void *tuples, float *thr, int64_t nrow,
                                                 int G, int Gp)
{
    const int lane = threadIdx.x & 63;
    const int64_t row = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    if (row >= nrow)
        return;
    using raw_t = typename Tup<STORE>::raw;
    const raw_t *t = static_cast<const raw_t *>(tuples) + row * Gp;
    float m = 0.f;
    for (int g = lane; g < G; g += 64)
        m = fmaxf(m, unpack<STORE>(t[g]).x);
#pragma unroll
    for (int o = 32; o > 0; o >>= 1)
        m = fmaxf(m, __shfl_xor(m, o));
    if (lane == 0)
        thr[row] = (1.f - 0x1p-20f) / m;
}

// ---- source update of an outer (source) iteration -------------------------------
// After a sweep, the track-length estimate of the angle-averaged flux of
// (region r, interval f, group g) is phi = fine_flux / vol, vol = sum of
// w * len over the segments in (r, f); isotropic within-group scattering with
// ratio c = sigma_s / sigT gives the new source
//   q[r][f][g] = q_ext[r][f][g] + c[r][g] * sigT[r][g] * phi[r][f][g]
// (FP64 arithmetic, stored as the FP32 master).  Rows without segments keep
// q_ext.  Each block also writes its partial sums of (q_new - q_old)^2 and
// q_old^2 in a fixed order, for a deterministic change norm.
template <typename FT>
__global__ void __launch_bounds__(256) k_source_update(smoc_update_args a)
{
    __shared__ double s_d[256], s_o[256];
    const int64_t n = (int64_t)a.R * a.F * a.G;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    double d2 = 0.0, o2 = 0.0;
    if (i < n) {
        const int G = a.G;
        const int64_t row = i / G;
        const int g = (int)(i - row * G);
        const int64_t r = row / a.F;
        const double v = a.vol[row];
        const double phi = v > 0.0 ? (double)static_cast<const FT *>(a.flux)[row * a.Gp + g] / v
                                   : 0.0;
        const double q = (double)a.q_ext[i] +
                         (double)a.scatter[r * G + g] * (double)a.sigT[r * G + g] * phi;
        const float qf = (float)q;
        const float old = a.source[i];
        a.source[i] = qf;
        d2 = ((double)qf - old) * ((double)qf - old);
        o2 = (double)old * old;
    }
    s_d[threadIdx.x] = d2;
    s_o[threadIdx.x] = o2;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) {
            s_d[threadIdx.x] += s_d[threadIdx.x + o];
            s_o[threadIdx.x] += s_o[threadIdx.x + o];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        a.partial[2 * blockIdx.x] = s_d[0];
        a.partial[2 * blockIdx.x + 1] = s_o[0];
    }
}

// Sum of the per-block partials in a fixed order (one block of 256: thread t
// sums blocks t, t + 256, ... in turn, then a tree over the threads).
__global__ void __launch_bounds__(256) k_sum_partials(const double *partial, int64_t nblk,
                                                      double *out)
{
    __shared__ double s_d[256], s_o[256];
    double d = 0.0, o = 0.0;
    for (int64_t b = threadIdx.x; b < nblk; b += 256) {
        d += partial[2 * b];
        o += partial[2 * b + 1];
    }
    s_d[threadIdx.x] = d;
    s_o[threadIdx.x] = o;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) {
            s_d[threadIdx.x] += s_d[threadIdx.x + h];
            s_o[threadIdx.x] += s_o[threadIdx.x + h];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[0] = s_d[0];
        out[1] = s_o[0];
    }
}

// ---- segment stream encoding (once per smoc_set_segments) -----------------------
// In place: zin_io holds z on entry and zin on exit; code_io holds the region
// id on entry and (region * F + f) << 2 | fit kind on exit.
__global__ void __launch_bounds__(256) k_encode_stream(float *zin_io, int32_t *code_io, int F,
                                                       int64_t S)
{
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= S)
        return;
    int f;
    float zi;
    decode_z(zin_io[k], F, f, zi);
    const int kind = f == 0 ? KIND_FIRST : (f == F - 1 ? KIND_LAST : KIND_INTERIOR);
    zin_io[k] = zi;
    code_io[k] = ((code_io[k] * F + f) << 2) | kind;
}

// ---- resets --------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(256) k_reset(T *flux, const float *init, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        flux[i] = init ? (T)init[i] : (T)0;
}

// Plain copy: U independent 16-byte loads in flight per lane, then the U
// stores; STRIDE: the grid strides over the array (a fixed grid sized to the
// CUs), else one pass with one workgroup per 256 U elements; NT: nontemporal
// loads and stores (the copy's data is touched once).
typedef float vec4f __attribute__((ext_vector_type(4)));
template <int U, bool STRIDE, bool NT>
__global__ void __launch_bounds__(256) k_copy(const vec4f *__restrict__ src,
                                               vec4f *__restrict__ dst, int64_t n)
{
    const int64_t step = (int64_t)gridDim.x * 256 * U;
    for (int64_t base = (int64_t)blockIdx.x * 256 * U + threadIdx.x; base < n; base += step) {
        vec4f v[U];
#pragma unroll
        for (int u = 0; u < U; u++)
            if (base + u * 256 < n)
                v[u] = NT ? __builtin_nontemporal_load(src + base + u * 256) : src[base + u * 256];
#pragma unroll
        for (int u = 0; u < U; u++)
            if (base + u * 256 < n) {
                if (NT)
                    __builtin_nontemporal_store(v[u], dst + base + u * 256);
                else
                    dst[base + u * 256] = v[u];
            }
        if (!STRIDE)
            break;
    }
}

// Tally-only probe: the per-segment FP32 atomic flush of the sweep on an
// unsorted stream, without the gathers and the arithmetic.  Wave w flushes
// segments [w*100, w*100 + 100) of S, segment s into row mix(s) % rows of a
// [rows][Gp] array, one atomic per group with the lanes past G masked, as the
// sweep does; what it measures is the memory-side atomic rate for that
// pattern, the floor of the unsorted sweep with an atomic tally.
__global__ void __launch_bounds__(256) k_tally_probe(float *flux, int64_t S, int rows, int G,
                                                      int Gp)
{
    const int lane = threadIdx.x & 63;
    const int64_t s0 = ((int64_t)blockIdx.x * 4 + (threadIdx.x >> 6)) * 100;
    const int64_t s1 = s0 + 100 < S ? s0 + 100 : S;
    for (int64_t s = s0; s < s1; s++) {
        uint32_t h = (uint32_t)s * 0x9E3779B1u;  // Fibonacci hashing, then a xorshift-multiply
        h ^= h >> 15;
        h *= 0x2C1B3C6Du;
        h ^= h >> 12;
        float *row = flux + (int64_t)(h % (uint32_t)rows) * Gp;
        for (int g = lane; g < G; g += 64)
            unsafeAtomicAdd(row + g, 1.0f);
    }
}

// FIXED64 result, in place: int64 fixed point -> init + value as FP64.
__global__ void __launch_bounds__(256) k_fixed_to_f64(int64_t *flux, const float *init, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        const double v = (double)flux[i] * (1.0 / kFixScale);
        reinterpret_cast<double *>(flux)[i] = init ? (double)init[i] + v : v;
    }
}

__global__ void __launch_bounds__(64) k_zero_i64(int64_t *p, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        p[i] = 0;
}

// ---- deterministic segmented reduction over run partials -------------------------
// One wave per ((region, interval), group slice); the runs of a (region,
// interval) are summed in ascending run-id order, whatever the launch shape.
// Lanes of the row padding (groups >= G) neither read partials nor write: no
// sweep writes their partials and no reader uses their fine_flux.
template <int TALLY, int GPL, bool SINGLE>
__global__ void __launch_bounds__(1024) k_reduce(smoc_reduce_args a)
{
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int wpb = blockDim.x >> 6;
    const int64_t item = (int64_t)xcd_remap(blockIdx.x, gridDim.x) * wpb + wave;
    if (item >= a.n_fsr * a.slices)
        return;
    const int64_t i = item / a.slices;
    const int sl = (int)(item - i * a.slices);
    const int Gp = a.Gp;
    const int64_t rb = a.fsr_run_off[i], re = a.fsr_run_off[i + 1];
    const int64_t *__restrict__ idx = a.fsr_run_idx;
    if constexpr (TALLY == SMOC_TALLY_DET_F64) {
        // both group slots in one loop, 8 runs of loads in flight per slot
        // before their adds; each slot's sum still takes the runs in
        // ascending order.  Padding lanes load group G-1's entries (the same
        // cache lines as the last real group) and store nothing.
        using part_t = typename std::conditional<SINGLE, float, double>::type;
        const part_t *__restrict__ part = static_cast<const part_t *>(a.part);
        int gi[GPL];
        bool on[GPL];
        double s[GPL];
#pragma unroll
        for (int j = 0; j < GPL; j++) {
            const int g = sl * 64 * GPL + j * 64 + lane;
            on[j] = g < a.G;
            gi[j] = on[j] ? g : a.G - 1;
            s[j] = a.init && on[j] ? (double)a.init[i * Gp + g] : 0.0;
        }
        int64_t r = rb;
        for (; r + 8 <= re; r += 8) {
            part_t v[8][GPL];
#pragma unroll
            for (int u = 0; u < 8; u++) {
                const part_t *row = part + idx[r + u] * Gp;
#pragma unroll
                for (int j = 0; j < GPL; j++)
                    v[u][j] = row[gi[j]];
            }
#pragma unroll
            for (int j = 0; j < GPL; j++)
#pragma unroll
                for (int u = 0; u < 8; u++)
                    s[j] += (double)v[u][j];
        }
        for (; r < re; ++r) {
            const part_t *row = part + idx[r] * Gp;
#pragma unroll
            for (int j = 0; j < GPL; j++)
                s[j] += (double)row[gi[j]];
        }
#pragma unroll
        for (int j = 0; j < GPL; j++)
            if (on[j])
                static_cast<double *>(a.flux)[i * Gp + sl * 64 * GPL + j * 64 + lane] = s[j];
        return;
    }
#pragma unroll
    for (int j = 0; j < GPL; j++) {
        const int g = sl * 64 * GPL + j * 64 + lane;
        if (g >= a.G)
            continue;
        const int64_t o = i * Gp + g;
        {
            // DET_KAHAN: compensated FP32, one slot at a time
#pragma clang fp contract(off)
            float s = a.init ? a.init[o] : 0.f, c = 0.f;
            for (int64_t r = rb; r < re; ++r) {
                float px, pc;
                if constexpr (SINGLE) {
                    px = static_cast<const float *>(a.part)[idx[r] * Gp + g];
                    pc = 0.f;
                } else {
                    const float2 p = static_cast<const float2 *>(a.part)[idx[r] * Gp + g];
                    px = p.x;
                    pc = p.y;
                }
                float y = px - c;
                float t = s + y;
                c = (t - s) - y;
                s = t;
                if constexpr (!SINGLE) {
                    y = -pc - c;
                    t = s + y;
                    c = (t - s) - y;
                    s = t;
                }
            }
            static_cast<float *>(a.flux)[o] = s - c;
        }
    }
}

template <int STORE, int GPL, int EXPM, int TALLY, bool RUNS, int MODE>
int launch_sweep_m(const smoc_sweep_args *a, int wpb, int grid, hipStream_t s)
{
    if (a->queue)
        hipLaunchKernelGGL((k_sweep_persistent<STORE, GPL, EXPM, TALLY, RUNS, MODE>), dim3(grid),
                           dim3(64 * wpb), 0, s, *a, a->len, a->code, a->zin, a->track_off,
                           a->items);
    else
        hipLaunchKernelGGL((k_sweep<STORE, GPL, EXPM, TALLY, RUNS, MODE>), dim3(grid),
                           dim3(64 * wpb), 0, s, *a, a->len, a->code, a->zin, a->track_off,
                           a->items);
    return (int)hipGetLastError();
}

template <int STORE, int GPL, int EXPM, int TALLY, bool RUNS>
int launch_sweep_r(const smoc_sweep_args *a, int wpb, int grid, hipStream_t s)
{
    if (a->xfer_mode)
        return launch_sweep_m<STORE, GPL, EXPM, TALLY, RUNS, MODE_MIXED>(a, wpb, grid, s);
    return launch_sweep_m<STORE, GPL, EXPM, TALLY, RUNS, MODE_SWEEP>(a, wpb, grid, s);
}

template <int STORE, int GPL, int EXPM, int TALLY>
int launch_sweep_t(const smoc_sweep_args *a, int wpb, int grid, hipStream_t s)
{
    if (a->runs)
        return launch_sweep_r<STORE, GPL, EXPM, TALLY, true>(a, wpb, grid, s);
    return launch_sweep_r<STORE, GPL, EXPM, TALLY, false>(a, wpb, grid, s);
}

template <int STORE, int GPL, int EXPM>
int launch_sweep_tally(const smoc_sweep_args *a, int tally, int wpb, int grid, hipStream_t s)
{
    switch (tally) {
    case SMOC_TALLY_ATOMIC_F32:
        return launch_sweep_t<STORE, GPL, EXPM, SMOC_TALLY_ATOMIC_F32>(a, wpb, grid, s);
    case SMOC_TALLY_ATOMIC_F64:
        return launch_sweep_t<STORE, GPL, EXPM, SMOC_TALLY_ATOMIC_F64>(a, wpb, grid, s);
    case SMOC_TALLY_DET_F64:
        return launch_sweep_t<STORE, GPL, EXPM, SMOC_TALLY_DET_F64>(a, wpb, grid, s);
    case SMOC_TALLY_DET_KAHAN:
        return launch_sweep_t<STORE, GPL, EXPM, SMOC_TALLY_DET_KAHAN>(a, wpb, grid, s);
    case SMOC_TALLY_FIXED64:
        return launch_sweep_t<STORE, GPL, EXPM, SMOC_TALLY_FIXED64>(a, wpb, grid, s);
    }
    return (int)hipErrorInvalidValue;
}

template <int STORE, int GPL>
int launch_sweep_exp(const smoc_sweep_args *a, int expm, int tally, int wpb, int grid,
                     hipStream_t s)
{
    switch (expm) {
    case SMOC_EXP_TABLE:
        return launch_sweep_tally<STORE, GPL, SMOC_EXP_TABLE>(a, tally, wpb, grid, s);
    case SMOC_EXP_ACCURATE:
        return launch_sweep_tally<STORE, GPL, SMOC_EXP_ACCURATE>(a, tally, wpb, grid, s);
    case SMOC_EXP_NATIVE:
        return launch_sweep_tally<STORE, GPL, SMOC_EXP_NATIVE>(a, tally, wpb, grid, s);
    }
    return (int)hipErrorInvalidValue;
}

template <int STORE>
int launch_sweep_gpl(const smoc_sweep_args *a, int gpl, int expm, int tally, int wpb, int grid,
                     hipStream_t s)
{
    if (gpl == 2)
        return launch_sweep_exp<STORE, 2>(a, expm, tally, wpb, grid, s);
    return launch_sweep_exp<STORE, 1>(a, expm, tally, wpb, grid, s);
}

inline unsigned blocks_for(int64_t n, int per) { return (unsigned)((n + per - 1) / per); }

}  // namespace

extern "C" int smoc_tuple_bytes(int storage)
{
    return storage == SMOC_STORE_F32 ? 16 : 8;
}

extern "C" int smoc_launch_pack(const smoc_pack_args *a, int storage, void *stream)
{
    const int64_t n = (int64_t)a->R * a->F * a->Gp;
    if (n == 0)
        return 0;
    hipStream_t s = static_cast<hipStream_t>(stream);
    const unsigned nb = blocks_for(n, 256);
    switch (storage) {
    case SMOC_STORE_F32: hipLaunchKernelGGL(k_pack<SMOC_STORE_F32>, dim3(nb), dim3(256), 0, s, *a); break;
    case SMOC_STORE_BF16: hipLaunchKernelGGL(k_pack<SMOC_STORE_BF16>, dim3(nb), dim3(256), 0, s, *a); break;
    case SMOC_STORE_F16: hipLaunchKernelGGL(k_pack<SMOC_STORE_F16>, dim3(nb), dim3(256), 0, s, *a); break;
    default: return (int)hipErrorInvalidValue;
    }
    return (int)hipGetLastError();
}

extern "C" int smoc_launch_row_thr(const void *tuples, float *thr, int64_t nrow, int G, int Gp,
                                   int storage, void *stream)
{
    if (nrow == 0)
        return 0;
    hipStream_t s = static_cast<hipStream_t>(stream);
    const unsigned nb = blocks_for(nrow, 4);
    switch (storage) {
    case SMOC_STORE_F32: hipLaunchKernelGGL(k_row_thr<SMOC_STORE_F32>, dim3(nb), dim3(256), 0, s, tuples, thr, nrow, G, Gp); break;
    case SMOC_STORE_BF16: hipLaunchKernelGGL(k_row_thr<SMOC_STORE_BF16>, dim3(nb), dim3(256), 0, s, tuples, thr, nrow, G, Gp); break;
    case SMOC_STORE_F16: hipLaunchKernelGGL(k_row_thr<SMOC_STORE_F16>, dim3(nb), dim3(256), 0, s, tuples, thr, nrow, G, Gp); break;
    default: return (int)hipErrorInvalidValue;
    }
    return (int)hipGetLastError();
}

extern "C" int smoc_launch_reset(void *flux, const float *init, int64_t n, int f64, void *stream)
{
    if (n == 0)
        return 0;
    hipStream_t s = static_cast<hipStream_t>(stream);
    const unsigned nb = blocks_for(n, 256);
    if (f64)
        hipLaunchKernelGGL(k_reset<double>, dim3(nb), dim3(256), 0, s, static_cast<double *>(flux), init, n);
    else
        hipLaunchKernelGGL(k_reset<float>, dim3(nb), dim3(256), 0, s, static_cast<float *>(flux), init, n);
    return (int)hipGetLastError();
}

extern "C" int smoc_launch_copy(const void *src, void *dst, int64_t n16, int variant, int cus,
                                void *stream)
{
    if (n16 == 0)
        return 0;
    hipStream_t s = static_cast<hipStream_t>(stream);
    const vec4f *a = static_cast<const vec4f *>(src);
    vec4f *b = static_cast<vec4f *>(dst);
    const unsigned stride_grid = (unsigned)std::max<int64_t>(
        1, std::min<int64_t>((int64_t)cus * 8, (n16 + 2047) / 2048));
    switch (variant) {
    case 0: hipLaunchKernelGGL((k_copy<4, false, false>), dim3(blocks_for(n16, 1024)), dim3(256), 0, s, a, b, n16); break;
    case 1: hipLaunchKernelGGL((k_copy<8, true, false>), dim3(stride_grid), dim3(256), 0, s, a, b, n16); break;
    case 2: hipLaunchKernelGGL((k_copy<8, true, true>), dim3(stride_grid), dim3(256), 0, s, a, b, n16); break;
    case 3: hipLaunchKernelGGL((k_copy<16, false, false>), dim3(blocks_for(n16, 4096)), dim3(256), 0, s, a, b, n16); break;
    default: return (int)hipErrorInvalidValue;
    }
    return (int)hipGetLastError();
}

extern "C" int smoc_launch_tally_probe(void *flux, int64_t S, int rows, int G, int Gp, int grid,
                                       void *stream)
{
    if (S == 0)
        return 0;
    hipLaunchKernelGGL(k_tally_probe, dim3(grid), dim3(256), 0, static_cast<hipStream_t>(stream),
                       static_cast<float *>(flux), S, rows, G, Gp);
    return (int)hipGetLastError();
}

extern "C" int64_t smoc_update_blocks(int64_t n) { return n > 0 ? blocks_for(n, 256) : 0; }

extern "C" int smoc_launch_source_update(const smoc_update_args *a, int flux_f64, void *stream)
{
    const int64_t n = (int64_t)a->R * a->F * a->G;
    const int64_t nb = smoc_update_blocks(n);
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (nb > 0) {
        if (flux_f64)
            hipLaunchKernelGGL(k_source_update<double>, dim3((unsigned)nb), dim3(256), 0, s, *a);
        else
            hipLaunchKernelGGL(k_source_update<float>, dim3((unsigned)nb), dim3(256), 0, s, *a);
    }
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(256), 0, s, a->partial, nb, a->sums);
    return (int)hipGetLastError();
}

extern "C" int smoc_launch_encode_stream(float *zin_io, int32_t *code_io, int F, int64_t S,
                                         void *stream)
{
    if (S <= 0)
        return 0;
    hipLaunchKernelGGL(k_encode_stream, dim3(blocks_for(S, 256)), dim3(256), 0,
                       (hipStream_t)stream, zin_io, code_io, F, S);
    return (int)hipGetLastError();
}

extern "C" int smoc_launch_plan(const int64_t *track_off, int64_t T, int chunk, int64_t *items,
                                int64_t *list1, int64_t *list2, int64_t *track_item_off,
                                const int64_t *track_run_off, const int64_t *split_run,
                                void *stream)
{
    if (T <= 0 || chunk <= 0)
        return T < 0 || chunk <= 0 ? (int)hipErrorInvalidValue : 0;
    if (split_run && (!track_run_off || chunk % SMOC_RUN_SPLIT != 0))
        return (int)hipErrorInvalidValue;
    hipLaunchKernelGGL(k_plan, dim3(1), dim3(1024), 0, static_cast<hipStream_t>(stream),
                       track_off, T, chunk, items, list1, list2, track_item_off, track_run_off,
                       split_run);
    return (int)hipGetLastError();
}

extern "C" int smoc_launch_fixed_to_f64(int64_t *flux, const float *init, int64_t n,
                                        void *stream)
{
    if (n == 0)
        return 0;
    hipLaunchKernelGGL(k_fixed_to_f64, dim3(blocks_for(n, 256)), dim3(256), 0,
                       static_cast<hipStream_t>(stream), flux, init, n);
    return (int)hipGetLastError();
}

extern "C" int smoc_launch_zero_i64(int64_t *p, int64_t n, void *stream)
{
    if (n == 0)
        return 0;
    hipLaunchKernelGGL(k_zero_i64, dim3(blocks_for(n, 64)), dim3(64), 0,
                       static_cast<hipStream_t>(stream), p, n);
    return (int)hipGetLastError();
}

extern "C" int smoc_launch_sweep(const smoc_sweep_args *a, int storage, int gpl, int exp_mode,
                                 int tally, int wpb, int grid_blocks, void *stream)
{
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (grid_blocks <= 0)
        return 0;
    switch (storage) {
    case SMOC_STORE_F32: return launch_sweep_gpl<SMOC_STORE_F32>(a, gpl, exp_mode, tally, wpb, grid_blocks, s);
    case SMOC_STORE_BF16: return launch_sweep_gpl<SMOC_STORE_BF16>(a, gpl, exp_mode, tally, wpb, grid_blocks, s);
    case SMOC_STORE_F16: return launch_sweep_gpl<SMOC_STORE_F16>(a, gpl, exp_mode, tally, wpb, grid_blocks, s);
    }
    return (int)hipErrorInvalidValue;
}

template <int TALLY, int GPL>
int launch_reduce_t(const smoc_reduce_args *a, unsigned nb, dim3 blk, hipStream_t s)
{
    if (a->single)
        hipLaunchKernelGGL((k_reduce<TALLY, GPL, true>), dim3(nb), blk, 0, s, *a);
    else
        hipLaunchKernelGGL((k_reduce<TALLY, GPL, false>), dim3(nb), blk, 0, s, *a);
    return (int)hipGetLastError();
}

extern "C" int smoc_launch_scan(const smoc_scan_args *a, int gpl, void *stream)
{
    const int64_t n = a->T * a->slices;
    if (n == 0)
        return 0;
    hipStream_t s = static_cast<hipStream_t>(stream);
    const unsigned nb = blocks_for(n, 4);
    if (gpl == 2)
        hipLaunchKernelGGL(k_scan<2>, dim3(nb), dim3(256), 0, s, *a);
    else
        hipLaunchKernelGGL(k_scan<1>, dim3(nb), dim3(256), 0, s, *a);
    return (int)hipGetLastError();
}

extern "C" int smoc_launch_lin_fix(const smoc_fix_args *a, int tally, int gpl, void *stream)
{
    const int64_t n = a->n_list * a->slices;
    if (n == 0)
        return 0;
    if (tally == SMOC_TALLY_FIXED64 || tally < 0 || tally > SMOC_TALLY_FIXED64 ||
        ((tally == SMOC_TALLY_ATOMIC_F32 || tally == SMOC_TALLY_ATOMIC_F64) && !a->run_row))
        return (int)hipErrorInvalidValue;
    hipStream_t s = static_cast<hipStream_t>(stream);
    const unsigned nb = blocks_for(n, 4);
    auto go = [&](auto tc) {
        constexpr int T = decltype(tc)::value;
        if (gpl == 2)
            hipLaunchKernelGGL((k_lin_fix<T, 2>), dim3(nb), dim3(256), 0, s, *a);
        else
            hipLaunchKernelGGL((k_lin_fix<T, 1>), dim3(nb), dim3(256), 0, s, *a);
    };
    switch (tally) {
    case SMOC_TALLY_ATOMIC_F32: go(std::integral_constant<int, SMOC_TALLY_ATOMIC_F32>{}); break;
    case SMOC_TALLY_ATOMIC_F64: go(std::integral_constant<int, SMOC_TALLY_ATOMIC_F64>{}); break;
    case SMOC_TALLY_DET_F64: go(std::integral_constant<int, SMOC_TALLY_DET_F64>{}); break;
    default: go(std::integral_constant<int, SMOC_TALLY_DET_KAHAN>{}); break;
    }
    return (int)hipGetLastError();
}

extern "C" int smoc_launch_reduce(const smoc_reduce_args *a, int tally, int gpl, int wpb,
                                  void *stream)
{
    hipStream_t s = static_cast<hipStream_t>(stream);
    const int64_t n = a->n_fsr * a->slices;
    if (n == 0)
        return 0;
    const unsigned nb = blocks_for(n, wpb);
    const dim3 blk(64 * wpb);
    if (tally == SMOC_TALLY_DET_F64)
        return gpl == 2 ? launch_reduce_t<SMOC_TALLY_DET_F64, 2>(a, nb, blk, s)
                        : launch_reduce_t<SMOC_TALLY_DET_F64, 1>(a, nb, blk, s);
    if (tally == SMOC_TALLY_DET_KAHAN)
        return gpl == 2 ? launch_reduce_t<SMOC_TALLY_DET_KAHAN, 2>(a, nb, blk, s)
                        : launch_reduce_t<SMOC_TALLY_DET_KAHAN, 1>(a, nb, blk, s);
    return (int)hipErrorInvalidValue;
}
