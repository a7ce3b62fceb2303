// smoc_api.cpp -- the C ABI of libsmoc (include/smoc.h) on the HIP runtime.
//
// A context owns its device buffers and one HIP stream.  Host inputs are
// copied in once (smoc_set_regions / smoc_set_segments) into padded device
// layouts; smoc_run only launches kernels on resident data, so its timing is
// whole-sweep throughput with inputs already in HBM.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <vector>

#include "smoc.h"
#include "smoc_internal.h"
#include "smoc_plan.h"

namespace {

// zeroed entries after the last segment of each device stream array
constexpr size_t kStreamPad = 64;

#define HIPCHK(call)                                                                           \
    do {                                                                                       \
        hipError_t e_ = (call);                                                                \
        if (e_ != hipSuccess) {                                                                \
            smoc_set_error("%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), __FILE__,    \
                           __LINE__);                                                          \
            return SMOC_ERR_HIP;                                                               \
        }                                                                                      \
    } while (0)

#define LAUNCHCHK(expr)                                                                        \
    do {                                                                                       \
        int e_ = (expr);                                                                       \
        if (e_ != 0) {                                                                         \
            smoc_set_error("kernel launch %s failed: %s", #expr,                               \
                           hipGetErrorString((hipError_t)e_));                                 \
            return SMOC_ERR_HIP;                                                               \
        }                                                                                      \
    } while (0)

template <typename T> struct DevBuf {
    T *p = nullptr;
    size_t n = 0;  // elements
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    ~DevBuf() { release(); }
    int alloc(size_t count)
    {
        release();
        if (count == 0)
            return SMOC_OK;
        hipError_t e = hipMalloc(reinterpret_cast<void **>(&p), count * sizeof(T));
        if (e != hipSuccess) {
            p = nullptr;
            smoc_set_error("hipMalloc of %zu bytes failed: %s", count * sizeof(T),
                           hipGetErrorString(e));
            return SMOC_ERR_NOMEM;
        }
        n = count;
        return SMOC_OK;
    }
    void release()
    {
        if (p)
            (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
};

inline int round_up(int x, int m) { return (x + m - 1) / m * m; }

// Chunking of long tracks for one chunk length C.
struct Plan {
    int chunk = 0;
    bool linear = false;  // linear chunks (MODE_LIN entries, run ids from the tables)
    int64_t n_items = 0, n1 = 0, n2 = 0;
    DevBuf<int64_t> items;           // [n_items][4]
    DevBuf<int64_t> list1, list2;    // phase-1 / phase-2 work entries
    DevBuf<int64_t> track_item_off;  // [T+1]
    DevBuf<float> chunk_psi;         // [n_items][Gp]
    DevBuf<float> xfer;              // [n_items][2][Gp]
    void release()
    {
        items.release();
        list1.release();
        list2.release();
        track_item_off.release();
        chunk_psi.release();
        xfer.release();
        chunk = 0;
        linear = false;
        n_items = n1 = n2 = 0;
    }
};

}  // namespace

struct smoc_ctx {
    smoc_params p{};
    int Gp = 0;  // padded group stride
    hipStream_t stream = nullptr;
    hipEvent_t ev[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
    int cu_count = 0;

    // region data
    DevBuf<unsigned char> tuples;
    DevBuf<float> row_thr;  // [R*F] series threshold per row (EXP_ACCURATE)
    DevBuf<float> flux_init;  // [R*F][Gp] or empty (zeros)
    DevBuf<unsigned char> flux;  // f32 or f64 [R*F][Gp]
    DevBuf<float> sig_m, src_m;  // FP32 masters on the device: sigT [R][G], source [R][F][G]
    DevBuf<float> table;
    float tab_dx = 0.f, tab_max = 0.f;
    int tab_n = 0;
    bool regions_set = false;

    // segments
    int64_t S = 0, T = 0;
    DevBuf<float> len, zin;  // zin: axial offset decoded from z at upload
    DevBuf<int32_t> code;    // (region * F + f) << 2 | fit kind
    DevBuf<int64_t> track_off;
    DevBuf<float> psi_in, psi_out;  // [T][Gp]
    bool segments_set = false;
    bool runs = false;  // sorted-enough stream: sum runs in registers
    bool split = false;  // runs also end at split points (runs, a track > SMOC_RUN_SPLIT)
    float fix_limit = 0.f;

    // deterministic tally bookkeeping
    int64_t n_runs = 0;
    DevBuf<int64_t> track_run_off, fsr_run_off, fsr_run_idx;
    // region-sorted stream: the id of the run that begins at each split point
    // inside a track (a stream index that is a multiple of SMOC_RUN_SPLIT),
    // track by track; host copies of it and of track_run_off for the chunk plan
    DevBuf<int64_t> split_run;
    std::vector<int64_t> h_track_run_off, h_split_run;
    DevBuf<unsigned char> part;
    // region-sorted stream: per run, its tally per unit of incoming psi when
    // its chunk is swept from psi = 0 (linear chunks) [n_runs][Gp]; allocated
    // when a track is long enough to be chunked
    DevBuf<float> wrun;
    // atomic tallies on such a stream: the (region, interval) row of each run
    DevBuf<int32_t> run_row;

    // outer (source) iteration
    DevBuf<float> scatter, q_ext;  // [R][G] scattering ratio, [R][F][G] external source
    DevBuf<double> vol;            // [R*F] sum of w * len per (region, interval)
    DevBuf<double> upd;            // per-block partials of the change norm, then 2 sums
    bool scatter_set = false;

    DevBuf<int64_t> queue;  // persistent work counters
    std::vector<int64_t> h_track_off;
    int64_t max_track = 0;
    Plan plan;
    hipGraphExec_t graph_exec = nullptr;
    smoc_launch graph_launch{};
    smoc_launch tuned{};
    bool tuned_valid = false;
    bool ran = false;

    int64_t n_fsr() const { return (int64_t)p.regions * p.fine_axial; }
    bool flux_f64() const
    {
        return p.tally == SMOC_TALLY_ATOMIC_F64 || p.tally == SMOC_TALLY_DET_F64 ||
               p.tally == SMOC_TALLY_FIXED64;
    }
    bool fixed() const { return p.tally == SMOC_TALLY_FIXED64; }
    // fine_flux starts 256 bytes into its buffer; with FIXED64 the 8 bytes
    // just before it count FIXED64 contributions beyond the bound (dropped)
    int64_t *fixed_flux() const { return reinterpret_cast<int64_t *>(flux.p + 256); }
    bool det() const { return p.tally == SMOC_TALLY_DET_F64 || p.tally == SMOC_TALLY_DET_KAHAN; }
};

// ---- scratch of the allocation-free entry point ---------------------------------
// One array per GPU in the library's code object: the runtime allocates it
// when the module loads on a device and it lives until the process ends.  Host
// code must not use the array's name as an address (that is a host-side
// stand-in); every use goes through the per-device address that
// hipGetSymbolAddress returns, looked up once per device by smoc_create.
static __device__ __align__(256) unsigned char scratch[SMOC_NOALLOC_SCRATCH_BYTES];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
unsigned char *g_scratch[kMaxDevices] = {};
}  // namespace

// With `device` current: its scratch address, looked up on first use.
static int scratch_init(int device)
{
    if (device < 0 || device >= kMaxDevices) {
        smoc_set_error("device %d beyond the %d devices the scratch table holds", device,
                       kMaxDevices);
        return SMOC_ERR_NODEV;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch[device])
        return SMOC_OK;
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) {
        smoc_set_error("hipGetSymbolAddress(scratch) on device %d failed: %s", device,
                       hipGetErrorString(e));
        return SMOC_ERR_HIP;
    }
    g_scratch[device] = static_cast<unsigned char *>(ptr);
    return SMOC_OK;
}

static unsigned char *scratch_of(int device)
{
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    return device >= 0 && device < kMaxDevices ? g_scratch[device] : nullptr;
}

// A captured graph holds the device addresses of every buffer its kernels
// touch; whatever frees or reallocates one of them drops the graph first.
static void drop_graph(smoc_ctx *c)
{
    if (c->graph_exec) {
        (void)hipGraphExecDestroy(c->graph_exec);
        c->graph_exec = nullptr;
    }
}

// Stream and events of a measurement probe, released on every return path.
struct ProbeTimer {
    hipStream_t s = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    ProbeTimer() = default;
    ProbeTimer(const ProbeTimer &) = delete;
    ProbeTimer &operator=(const ProbeTimer &) = delete;
    ~ProbeTimer()
    {
        if (e0)
            (void)hipEventDestroy(e0);
        if (e1)
            (void)hipEventDestroy(e1);
        if (s)
            (void)hipStreamDestroy(s);
    }
};

// One warm-up launch, then `reps` launches between two events on a stream of
// their own: the mean time per launch.
template <typename Launch> static int time_probe(Launch launch, int reps, double *mean_ms)
{
    ProbeTimer t;
    HIPCHK(hipStreamCreateWithFlags(&t.s, hipStreamNonBlocking));
    HIPCHK(hipEventCreate(&t.e0));
    HIPCHK(hipEventCreate(&t.e1));
    LAUNCHCHK(launch(t.s));
    HIPCHK(hipEventRecord(t.e0, t.s));
    for (int r = 0; r < reps; r++)
        LAUNCHCHK(launch(t.s));
    HIPCHK(hipEventRecord(t.e1, t.s));
    HIPCHK(hipEventSynchronize(t.e1));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, t.e0, t.e1));
    *mean_ms = ms / reps;
    return SMOC_OK;
}

extern "C" {

int smoc_device_count(int *n)
{
    if (!n) {
        smoc_set_error("smoc_device_count: NULL");
        return SMOC_ERR_ARG;
    }
    *n = 0;
    int c = 0;
    hipError_t e = hipGetDeviceCount(&c);
    if (e != hipSuccess) {
        smoc_set_error("hipGetDeviceCount: %s", hipGetErrorString(e));
        return SMOC_ERR_NODEV;
    }
    *n = c;
    return SMOC_OK;
}

int smoc_device_info(int device, char *name, int name_cap, int *compute_units)
{
    hipDeviceProp_t prop;
    hipError_t e = hipGetDeviceProperties(&prop, device);
    if (e != hipSuccess) {
        smoc_set_error("hipGetDeviceProperties(%d): %s", device, hipGetErrorString(e));
        return SMOC_ERR_NODEV;
    }
    if (name && name_cap > 0)
        snprintf(name, (size_t)name_cap, "%s (%s)", prop.name[0] ? prop.name : "AMD GPU",
                 prop.gcnArchName);
    if (compute_units)
        *compute_units = prop.multiProcessorCount;
    return SMOC_OK;
}

int smoc_probe_copy(int device, int64_t bytes, int reps, int variant, double *mean_ms)
{
    if (bytes <= 0 || bytes % 16 != 0 || reps < 1 || !mean_ms || variant < 0 ||
        variant >= SMOC_COPY_VARIANTS) {
        smoc_set_error("smoc_probe_copy: bytes must be a positive multiple of 16, reps >= 1, "
                       "0 <= variant < %d", SMOC_COPY_VARIANTS);
        return SMOC_ERR_ARG;
    }
    const int64_t nblk = (bytes / 16 + 1023) / 1024;
    if (nblk > INT32_MAX) {
        smoc_set_error("smoc_probe_copy: %lld bytes is too large", (long long)bytes);
        return SMOC_ERR_ARG;
    }
    int ndev = 0;
    int rc = smoc_device_count(&ndev);
    if (rc != SMOC_OK)
        return rc;
    if (device < 0 || device >= ndev) {
        smoc_set_error("smoc_probe_copy: device %d not available", device);
        return SMOC_ERR_NODEV;
    }
    HIPCHK(hipSetDevice(device));
    DevBuf<unsigned char> a, b;
    rc = a.alloc((size_t)bytes);
    if (rc == SMOC_OK)
        rc = b.alloc((size_t)bytes);
    if (rc != SMOC_OK)
        return rc;
    HIPCHK(hipMemset(a.p, 1, (size_t)bytes));
    HIPCHK(hipDeviceSynchronize());
    int cus = 0;
    HIPCHK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device));
    return time_probe(
        [&](hipStream_t s) { return smoc_launch_copy(a.p, b.p, bytes / 16, variant, cus, s); },
        reps, mean_ms);
}

int smoc_probe_tally(int device, int64_t S, int rows, int G, int reps, double *mean_ms)
{
    if (S < 0 || rows < 1 || G < 1 || G > 65536 || reps < 1 || !mean_ms) {
        smoc_set_error("smoc_probe_tally: need S >= 0, rows >= 1, 1 <= G <= 65536, reps >= 1");
        return SMOC_ERR_ARG;
    }
    const int64_t nblk = ((S + 99) / 100 + 3) / 4;
    if (nblk > INT32_MAX) {
        smoc_set_error("smoc_probe_tally: %lld segments is too many", (long long)S);
        return SMOC_ERR_ARG;
    }
    int ndev = 0;
    int rc = smoc_device_count(&ndev);
    if (rc != SMOC_OK)
        return rc;
    if (device < 0 || device >= ndev) {
        smoc_set_error("smoc_probe_tally: device %d not available", device);
        return SMOC_ERR_NODEV;
    }
    HIPCHK(hipSetDevice(device));
    const int Gp = round_up(G, 128);
    DevBuf<float> flux;
    rc = flux.alloc((size_t)rows * Gp);
    if (rc != SMOC_OK)
        return rc;
    HIPCHK(hipMemset(flux.p, 0, sizeof(float) * (size_t)rows * Gp));
    HIPCHK(hipDeviceSynchronize());
    const int grid = (int)(nblk > 0 ? nblk : 1);
    return time_probe(
        [&](hipStream_t s) {
            return smoc_launch_tally_probe(flux.p, S, rows, G, Gp, grid, s);
        },
        reps, mean_ms);
}

int smoc_create(smoc_ctx **out, const smoc_params *p)
{
    if (!out || !p) {
        smoc_set_error("smoc_create: NULL argument");
        return SMOC_ERR_ARG;
    }
    *out = nullptr;
    if (p->egroups < 1 || p->egroups > 65536 || p->regions < 1 || p->fine_axial < 2 ||
        (int64_t)p->regions * p->fine_axial > INT32_MAX / 4) {
        smoc_set_error("smoc_create: need 1 <= egroups <= 65536, regions >= 1, fine_axial >= 2, "
                       "regions * fine_axial < 2^29 (got G=%d R=%d F=%d)",
                       p->egroups, p->regions, p->fine_axial);
        return SMOC_ERR_ARG;
    }
    if (p->storage < SMOC_STORE_F32 || p->storage > SMOC_STORE_F16 ||
        p->exp_mode < SMOC_EXP_NATIVE || p->exp_mode > SMOC_EXP_ACCURATE ||
        p->tally < SMOC_TALLY_ATOMIC_F32 || p->tally > SMOC_TALLY_FIXED64) {
        smoc_set_error("smoc_create: unknown storage %d / exp_mode %d / tally %d", p->storage,
                       p->exp_mode, p->tally);
        return SMOC_ERR_ARG;
    }
    int ndev = 0;
    int rc = smoc_device_count(&ndev);
    if (rc != SMOC_OK)
        return rc;
    if (ndev <= 0 || p->device < 0 || p->device >= ndev) {
        smoc_set_error("smoc_create: device %d not available (%d HIP devices)", p->device, ndev);
        return SMOC_ERR_NODEV;
    }
    HIPCHK(hipSetDevice(p->device));
    rc = scratch_init(p->device);
    if (rc != SMOC_OK)
        return rc;
    smoc_ctx *c = new (std::nothrow) smoc_ctx();
    if (!c) {
        smoc_set_error("smoc_create: host allocation failed");
        return SMOC_ERR_NOMEM;
    }
    c->p = *p;
    c->Gp = round_up(p->egroups, 128);
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    for (int i = 0; e == hipSuccess && i < 5; i++)
        e = hipEventCreate(&c->ev[i]);
    if (e == hipSuccess) {
        hipDeviceProp_t prop;
        e = hipGetDeviceProperties(&prop, p->device);
        c->cu_count = prop.multiProcessorCount;
    }
    if (e != hipSuccess) {
        smoc_set_error("smoc_create: stream/event setup failed: %s", hipGetErrorString(e));
        smoc_destroy(c);
        return SMOC_ERR_HIP;
    }
    if (p->exp_mode == SMOC_EXP_TABLE) {
        float vals[SMOC_TABLE_CAP];
        rc = smoc_build_exp_table(0.01f, 10.0f, vals, SMOC_TABLE_CAP, &c->tab_n, &c->tab_dx,
                                  &c->tab_max);
        if (rc == SMOC_OK)
            rc = c->table.alloc((size_t)2 * c->tab_n);
        if (rc == SMOC_OK) {
            e = hipMemcpy(c->table.p, vals, sizeof(float) * 2 * c->tab_n, hipMemcpyHostToDevice);
            if (e != hipSuccess) {
                smoc_set_error("smoc_create: table upload failed: %s", hipGetErrorString(e));
                rc = SMOC_ERR_HIP;
            }
        }
        if (rc != SMOC_OK) {
            smoc_destroy(c);
            return rc;
        }
    }
    *out = c;
    return SMOC_OK;
}

int smoc_destroy(smoc_ctx *c)
{
    if (!c)
        return SMOC_OK;
    (void)hipSetDevice(c->p.device);
    if (c->stream)
        (void)hipStreamSynchronize(c->stream);
    c->tuples.release();
    c->row_thr.release();
    c->sig_m.release();
    c->src_m.release();
    c->scatter.release();
    c->q_ext.release();
    c->vol.release();
    c->upd.release();
    c->flux_init.release();
    c->flux.release();
    c->table.release();
    c->len.release();
    c->zin.release();
    c->code.release();
    c->track_off.release();
    c->psi_in.release();
    c->psi_out.release();
    c->track_run_off.release();
    c->fsr_run_off.release();
    c->fsr_run_idx.release();
    c->split_run.release();
    c->part.release();
    c->wrun.release();
    c->run_row.release();
    c->queue.release();
    c->plan.release();
    if (c->graph_exec)
        (void)hipGraphExecDestroy(c->graph_exec);
    for (auto &ev : c->ev)
        if (ev)
            (void)hipEventDestroy(ev);
    if (c->stream)
        (void)hipStreamDestroy(c->stream);
    delete c;
    return SMOC_OK;
}

// Copy a [rows][G] host array into a [rows][Gp] device array (padding = 0).
// Rows of `cols` elements between contiguous host memory and device rows
// `pitch` elements apart.  A 2D copy straight to or from pageable host memory
// moves one row at a time (seconds for 5e5 rows), so the host side is one
// contiguous copy through a device staging buffer and the strided part is a
// device-to-device 2D copy.  Both return with the copies complete.
extern "C++" {
template <typename T>
static int rows_to_device(T *dst, int64_t pitch, const T *src, int64_t rows, int cols,
                          hipStream_t s)
{
    if (rows == 0)
        return SMOC_OK;
    const size_t n = (size_t)rows * cols;
    if (pitch == cols) {
        HIPCHK(hipMemcpyAsync(dst, src, sizeof(T) * n, hipMemcpyHostToDevice, s));
        HIPCHK(hipStreamSynchronize(s));
        return SMOC_OK;
    }
    DevBuf<T> stage;
    int rc = stage.alloc(n);
    if (rc != SMOC_OK)
        return rc;
    HIPCHK(hipMemcpyAsync(stage.p, src, sizeof(T) * n, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpy2DAsync(dst, sizeof(T) * pitch, stage.p, sizeof(T) * cols, sizeof(T) * cols,
                            (size_t)rows, hipMemcpyDeviceToDevice, s));
    HIPCHK(hipStreamSynchronize(s));
    return SMOC_OK;
}

template <typename T>
static int rows_to_host(T *dst, const T *src, int64_t pitch, int64_t rows, int cols,
                        hipStream_t s)
{
    if (rows == 0)
        return SMOC_OK;
    const size_t n = (size_t)rows * cols;
    if (pitch == cols) {
        HIPCHK(hipMemcpyAsync(dst, src, sizeof(T) * n, hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        return SMOC_OK;
    }
    DevBuf<T> stage;
    int rc = stage.alloc(n);
    if (rc != SMOC_OK)
        return rc;
    HIPCHK(hipMemcpy2DAsync(stage.p, sizeof(T) * cols, src, sizeof(T) * pitch, sizeof(T) * cols,
                            (size_t)rows, hipMemcpyDeviceToDevice, s));
    HIPCHK(hipMemcpyAsync(dst, stage.p, sizeof(T) * n, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    return SMOC_OK;
}
}  // extern "C++"

// host [rows][G] -> device [rows][Gp], the padding zeroed
static int upload_padded(float *dst, const float *src, int64_t rows, int G, int Gp,
                         hipStream_t s)
{
    if (rows == 0)
        return SMOC_OK;
    HIPCHK(hipMemsetAsync(dst, 0, sizeof(float) * (size_t)rows * Gp, s));
    return rows_to_device(dst, Gp, src, rows, G, s);
}

int smoc_set_regions(smoc_ctx *c, const float *sigT, const float *fine_source,
                     const float *fine_flux_init)
{
    if (!c || !sigT || !fine_source) {
        smoc_set_error("smoc_set_regions: NULL argument");
        return SMOC_ERR_ARG;
    }
    HIPCHK(hipSetDevice(c->p.device));
    const int G = c->p.egroups, R = c->p.regions, F = c->p.fine_axial;
    const int64_t nsig = (int64_t)R * G, nsrc = (int64_t)R * F * G;
    for (int64_t i = 0; i < nsig; i++)
        if (!std::isfinite(sigT[i]) || !(sigT[i] > 0.f)) {
            smoc_set_error("smoc_set_regions: sigT[%lld] = %g must be finite and > 0",
                           (long long)i, (double)sigT[i]);
            return SMOC_ERR_ARG;
        }
    for (int64_t i = 0; i < nsrc; i++)
        if (!std::isfinite(fine_source[i])) {
            smoc_set_error("smoc_set_regions: fine_source[%lld] is not finite", (long long)i);
            return SMOC_ERR_ARG;
        }
    const int64_t nrow = c->n_fsr();
    // the buffers below are reallocated: until this call succeeds the context
    // holds no region data, and no captured graph may replay on freed memory
    HIPCHK(hipStreamSynchronize(c->stream));
    drop_graph(c);
    c->regions_set = false;
    c->ran = false;
    c->tuned_valid = false;
    c->scatter_set = false;
    int rc = c->tuples.alloc((size_t)nrow * c->Gp * smoc_tuple_bytes(c->p.storage));
    if (rc == SMOC_OK)
        rc = c->row_thr.alloc((size_t)nrow);
    if (rc == SMOC_OK)
        rc = c->flux.alloc((size_t)nrow * c->Gp * (c->flux_f64() ? 8 : 4) + 256);
    if (rc != SMOC_OK)
        return rc;
    // FP32 masters are kept on the device and packed (and quantised) there;
    // a source update rewrites the source master and repacks
    rc = c->sig_m.alloc((size_t)nsig);
    if (rc == SMOC_OK)
        rc = c->src_m.alloc((size_t)nsrc);
    if (rc != SMOC_OK)
        return rc;
    HIPCHK(hipMemcpyAsync(c->sig_m.p, sigT, sizeof(float) * nsig, hipMemcpyHostToDevice,
                          c->stream));
    HIPCHK(hipMemcpyAsync(c->src_m.p, fine_source, sizeof(float) * nsrc, hipMemcpyHostToDevice,
                          c->stream));
    smoc_pack_args pa{c->sig_m.p, c->src_m.p, c->tuples.p, G, c->Gp, R, F};
    LAUNCHCHK(smoc_launch_pack(&pa, c->p.storage, c->stream));
    LAUNCHCHK(smoc_launch_row_thr(c->tuples.p, c->row_thr.p, nrow, G, c->Gp, c->p.storage,
                                  c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (fine_flux_init) {
        rc = c->flux_init.alloc((size_t)nrow * c->Gp);
        if (rc == SMOC_OK)
            rc = upload_padded(c->flux_init.p, fine_flux_init, nrow, G, c->Gp, c->stream);
        if (rc != SMOC_OK)
            return rc;
    } else {
        c->flux_init.release();
    }
    c->regions_set = true;
    return SMOC_OK;
}

// Run bookkeeping: runs are numbered track by track in stream order; when
// c->runs a run is a maximal stretch of one track's segments with the same
// (region, interval) that does not cross a split point (a stream index that
// is a multiple of SMOC_RUN_SPLIT: the sweep ends its runs there too), else a
// single segment.  split_run holds the id of the run that begins at each
// split point inside a track, track by track.  The deterministic tallies get
// all of it: fsr_run_idx lists, for every (region, interval), its runs in
// ascending id, and part holds the partials.  An atomic FP32/FP64 tally on a
// sorted stream that can be chunked (a track longer than SMOC_RUN_SPLIT) gets
// the track and split-point tables, run_row and wrun, for linear chunks.
static_assert((SMOC_RUN_SPLIT & (SMOC_RUN_SPLIT - 1)) == 0, "split length: a power of two");
static int build_runs(smoc_ctx *c, int64_t S, const int32_t *region, const float *z, int64_t T,
                      const int64_t *track_off)
{
    const int F = c->p.fine_axial;
    const int64_t K = c->n_fsr();
    const bool det = c->det();
    // linear chunks need wrun; a stream can be chunked only when a track is
    // longer than the chunk length, which is then at least SMOC_RUN_SPLIT
    const bool chunkable = c->split;
    if (!det && !chunkable) {
        c->track_run_off.release();
        c->split_run.release();
        c->wrun.release();
        c->run_row.release();
        c->h_track_run_off.clear();
        c->h_split_run.clear();
        c->n_runs = 0;
        return SMOC_OK;
    }
    std::vector<int32_t> run_key;
    std::vector<int64_t> tro((size_t)T + 1), split;
    try {
        run_key.reserve(c->runs ? (size_t)T + 1024 : (size_t)S);
        if (chunkable)
            split.reserve((size_t)(S / SMOC_RUN_SPLIT) + 1);
        for (int64_t t = 0; t < T; t++) {
            tro[t] = (int64_t)run_key.size();
            int32_t prev = -1;
            for (int64_t k = track_off[t]; k < track_off[t + 1]; k++) {
                int f;
                float zin;
                smoc_decode_z(z[k], F, &f, &zin);
                const int32_t key = region[k] * F + f;
                const bool at_split =
                    chunkable && k > track_off[t] && (k & (SMOC_RUN_SPLIT - 1)) == 0;
                if (at_split)
                    split.push_back((int64_t)run_key.size());
                if (!c->runs || key != prev || at_split)
                    run_key.push_back(key);
                prev = key;
            }
        }
        c->h_track_run_off.clear();
        c->h_split_run.clear();
        if (c->runs) {
            c->h_track_run_off = tro;
            c->h_split_run = split;
        }
    } catch (...) {
        smoc_set_error("smoc_set_segments: host allocation for run bookkeeping failed");
        return SMOC_ERR_NOMEM;
    }
    tro[T] = (int64_t)run_key.size();
    if (c->runs)
        c->h_track_run_off[T] = tro[T];
    const int64_t n_runs = tro[T];
    int rc = c->track_run_off.alloc((size_t)T + 1);
    if (rc == SMOC_OK)
        rc = c->split_run.alloc(std::max<size_t>(split.size(), 1));
    if (rc == SMOC_OK && chunkable)
        rc = c->wrun.alloc((size_t)std::max<int64_t>(n_runs, 1) * c->Gp);
    else if (rc == SMOC_OK)
        c->wrun.release();
    if (rc != SMOC_OK)
        return rc;
    if (!split.empty())
        HIPCHK(hipMemcpy(c->split_run.p, split.data(), sizeof(int64_t) * split.size(),
                         hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(c->track_run_off.p, tro.data(), sizeof(int64_t) * (T + 1),
                     hipMemcpyHostToDevice));
    if (det) {
        std::vector<int64_t> fro, idx;
        try {
            fro.assign((size_t)K + 1, 0);
            idx.resize((size_t)std::max<int64_t>(n_runs, 1));
            for (int64_t r = 0; r < n_runs; r++)
                fro[run_key[r] + 1]++;
            for (int64_t k = 0; k < K; k++)
                fro[k + 1] += fro[k];
            std::vector<int64_t> pos(fro.begin(), fro.end() - 1);
            for (int64_t r = 0; r < n_runs; r++)
                idx[pos[run_key[r]]++] = r;
        } catch (...) {
            smoc_set_error("smoc_set_segments: host allocation for run bookkeeping failed");
            return SMOC_ERR_NOMEM;
        }
        rc = c->fsr_run_off.alloc((size_t)K + 1);
        if (rc == SMOC_OK)
            rc = c->fsr_run_idx.alloc((size_t)std::max<int64_t>(n_runs, 1));
        if (rc == SMOC_OK)
            rc = c->part.alloc((size_t)std::max<int64_t>(n_runs, 1) * c->Gp * (c->runs ? 8 : 4));
        if (rc != SMOC_OK)
            return rc;
        HIPCHK(hipMemcpy(c->fsr_run_off.p, fro.data(), sizeof(int64_t) * (K + 1),
                         hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(c->fsr_run_idx.p, idx.data(),
                         sizeof(int64_t) * std::max<int64_t>(n_runs, 1), hipMemcpyHostToDevice));
        c->run_row.release();
    } else {
        // run keys are the (region, interval) rows
        rc = c->run_row.alloc((size_t)std::max<int64_t>(n_runs, 1));
        if (rc != SMOC_OK)
            return rc;
        if (n_runs > 0)
            HIPCHK(hipMemcpy(c->run_row.p, run_key.data(), sizeof(int32_t) * n_runs,
                             hipMemcpyHostToDevice));
    }
    c->n_runs = n_runs;
    return SMOC_OK;
}

int smoc_set_segments(smoc_ctx *c, int64_t S, const float *len, const int32_t *region,
                      const float *z, int64_t T, const int64_t *track_off, const float *psi_in)
{
    if (!c || S < 0 || T < 0 || !track_off || (S > 0 && (!len || !region || !z)) ||
        (T > 0 && !psi_in)) {
        smoc_set_error("smoc_set_segments: bad argument");
        return SMOC_ERR_ARG;
    }
    if (S > INT32_MAX) {
        smoc_set_error("smoc_set_segments: %lld segments exceed 2^31-1", (long long)S);
        return SMOC_ERR_ARG;
    }
    HIPCHK(hipSetDevice(c->p.device));
    const int R = c->p.regions, G = c->p.egroups;
    // every index a kernel derives from the stream is checked here
    if (track_off[0] != 0 || track_off[T] != S) {
        smoc_set_error("smoc_set_segments: track_off[0] = %lld and track_off[T] = %lld must be "
                       "0 and S = %lld",
                       (long long)track_off[0], (long long)track_off[T], (long long)S);
        return SMOC_ERR_ARG;
    }
    for (int64_t t = 0; t < T; t++)
        if (track_off[t + 1] < track_off[t]) {
            smoc_set_error("smoc_set_segments: track_off not monotone at %lld", (long long)t);
            return SMOC_ERR_ARG;
        }
    int64_t same = 0;
    for (int64_t k = 0; k < S; k++) {
        if (region[k] < 0 || region[k] >= R) {
            smoc_set_error("smoc_set_segments: region[%lld] = %d outside [0, %d)", (long long)k,
                           region[k], R);
            return SMOC_ERR_ARG;
        }
        if (!std::isfinite(len[k]) || !std::isfinite(z[k]) || len[k] < 0.f) {
            smoc_set_error("smoc_set_segments: segment %lld has len %g, z %g", (long long)k,
                           (double)len[k], (double)z[k]);
            return SMOC_ERR_ARG;
        }
        if (k > 0 && region[k] == region[k - 1])
            same++;
    }
    for (int64_t i = 0; i < T * G; i++)
        if (!std::isfinite(psi_in[i])) {
            smoc_set_error("smoc_set_segments: psi_in[%lld] is not finite", (long long)i);
            return SMOC_ERR_ARG;
        }
    // A stream sorted by region keeps most neighbours in the same region;
    // then runs are summed in registers.
    c->runs = S > 1 && 2 * same >= S;
    {
        int64_t longest = 0;
        for (int64_t t = 0; t < T; t++)
            longest = std::max<int64_t>(longest, track_off[t + 1] - track_off[t]);
        c->split = c->runs && longest > SMOC_RUN_SPLIT;
    }
    // FIXED64: no fine_flux sum may reach 2^31, so bound each contribution by
    // 2^31 / (most segments landing in one region interval)
    c->fix_limit = 0.f;
    if (c->fixed()) {
        std::vector<int32_t> cnt((size_t)c->n_fsr(), 0);
        int32_t most = 1;
        for (int64_t k = 0; k < S; k++) {
            int f;
            float zin;
            smoc_decode_z(z[k], c->p.fine_axial, &f, &zin);
            most = std::max(most, ++cnt[(size_t)region[k] * c->p.fine_axial + f]);
        }
        c->fix_limit = (float)(2147483647.0 / most * 0.999);
    }

    // the buffers below are reallocated: until this call succeeds the context
    // holds no segment data, no chunk plan and no captured graph
    HIPCHK(hipStreamSynchronize(c->stream));
    drop_graph(c);
    c->segments_set = false;
    c->ran = false;
    c->tuned_valid = false;
    c->plan.release();
    c->S = c->T = 0;
    c->n_runs = 0;
    c->max_track = 0;
    c->h_track_off.clear();
    c->h_track_run_off.clear();
    c->h_split_run.clear();
    // the stream arrays carry kStreamPad zeroed entries past S: the sweep
    // loads one segment ahead without a bounds branch
    int rc = c->len.alloc((size_t)S + kStreamPad);
    if (rc == SMOC_OK)
        rc = c->zin.alloc((size_t)S + kStreamPad);
    if (rc == SMOC_OK)
        rc = c->code.alloc((size_t)S + kStreamPad);
    if (rc == SMOC_OK)
        rc = c->track_off.alloc((size_t)T + 1);
    if (rc == SMOC_OK)
        rc = c->psi_in.alloc((size_t)T * c->Gp);
    if (rc == SMOC_OK)
        rc = c->psi_out.alloc((size_t)T * c->Gp);
    if (rc != SMOC_OK)
        return rc;
    HIPCHK(hipMemset(c->len.p + S, 0, sizeof(float) * kStreamPad));
    HIPCHK(hipMemset(c->zin.p + S, 0, sizeof(float) * kStreamPad));
    HIPCHK(hipMemset(c->code.p + S, 0, sizeof(int32_t) * kStreamPad));
    if (S > 0) {
        HIPCHK(hipMemcpy(c->len.p, len, sizeof(float) * S, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(c->zin.p, z, sizeof(float) * S, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(c->code.p, region, sizeof(int32_t) * S, hipMemcpyHostToDevice));
        // z -> zin and region -> row code, in place, once per upload
        HIPCHK((hipError_t)smoc_launch_encode_stream(c->zin.p, c->code.p, c->p.fine_axial, S,
                                                     c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
    }
    HIPCHK(hipMemcpy(c->track_off.p, track_off, sizeof(int64_t) * (T + 1),
                     hipMemcpyHostToDevice));
    rc = upload_padded(c->psi_in.p, psi_in, T, G, c->Gp, c->stream);
    if (rc != SMOC_OK)
        return rc;
    {
        // track-length volume of each (region, interval): sum of w * len, in
        // segment order (FP64), for the source update of an outer iteration
        std::vector<double> v((size_t)c->n_fsr(), 0.0);
        for (int64_t k = 0; k < S; k++) {
            int f;
            float zi;
            smoc_decode_z(z[k], c->p.fine_axial, &f, &zi);
            v[(size_t)region[k] * c->p.fine_axial + f] += (double)SMOC_WEIGHT * len[k];
        }
        rc = c->vol.alloc(v.size());
        if (rc != SMOC_OK)
            return rc;
        HIPCHK(hipMemcpy(c->vol.p, v.data(), sizeof(double) * v.size(), hipMemcpyHostToDevice));
    }
    if (!c->fixed()) {  // (releases what an earlier stream needed when nothing is)
        rc = build_runs(c, S, region, z, T, track_off);
        if (rc != SMOC_OK)
            return rc;
    }
    try {
        c->h_track_off.assign(track_off, track_off + T + 1);
    } catch (...) {
        smoc_set_error("smoc_set_segments: host allocation for track offsets failed");
        return SMOC_ERR_NOMEM;
    }
    int64_t longest = 0;
    for (int64_t t = 0; t < T; t++)
        longest = std::max<int64_t>(longest, track_off[t + 1] - track_off[t]);
    c->max_track = longest;
    c->S = S;
    c->T = T;
    c->segments_set = true;
    return SMOC_OK;
}

static void fill_launch(const smoc_ctx *c, const smoc_launch *in, smoc_launch *l)
{
    smoc_launch d{};
    if (in)
        d = *in;
    if (d.groups_per_lane != 1 && d.groups_per_lane != 2)
        d.groups_per_lane = c->p.egroups > 64 ? 2 : 1;
    // default shape: 4 waves per workgroup (measured best, or within 2 % of
    // it, on unsorted and region-sorted streams alike)
    if (d.waves_per_block < 1 || d.waves_per_block > 16)
        d.waves_per_block = 4;
    if (d.chunk < 0)
        d.chunk = 0;
    if (d.chunk == 0 && d.seg_per_block > 0)
        d.chunk = d.seg_per_block;  // -p: segments per work item
    if (d.chunk > 0 && d.chunk >= c->max_track)
        d.chunk = 0;  // nothing to split
    d.persistent = d.persistent ? 1 : 0;
    d.use_graph = d.use_graph ? 1 : 0;
    if (d.reduce_waves < 1 || d.reduce_waves > 16)
        d.reduce_waves = 4;
    d.segment_flush = d.segment_flush ? 1 : 0;
    *l = d;
}

// Linear chunks (MODE_LIN): a chunked sweep that sums runs in registers (a
// region-sorted stream, no segment_flush) with a deterministic tally, or an
// atomic FP32/FP64 one at a chunk length that is a multiple of SMOC_RUN_SPLIT
// (at other lengths chunks need not begin runs, and the transfer pass runs).
static bool linear_chunks(const smoc_ctx *c, int chunk, bool segment_flush)
{
    return chunk > 0 && c->runs && !segment_flush && !c->fixed() &&
           (c->det() || chunk % SMOC_RUN_SPLIT == 0);
}

// Work items of a chunked sweep: every track is cut into chunks of about C
// segments (smoc_plan.h; an empty track is one empty chunk).  Phase 1 sweeps
// each track's first chunk (a multi-chunk track's first chunk also emits its
// psi_out as a transfer) and runs the transfer pass of the middle chunks;
// phase 2 sweeps the remaining chunks.  Both lists are ordered longest first.
// An item's first run id is its first segment, where every segment is a run,
// or with linear chunks the run that begins at its first segment: C is then a
// multiple of SMOC_RUN_SPLIT, so each chunk after a track's first begins at
// one of the track's split points.  With linear chunks phase 1 sweeps every
// chunk, those after a track's first from psi = 0, and list2 names the chunks
// whose runs k_lin_fix corrects.
static int build_plan(smoc_ctx *c, int C, bool linear)
{
    Plan &pl = c->plan;
    // the plan buffers are reallocated below; a graph captured with the old
    // ones (any chunk length) must not replay, and until this plan is complete
    // no chunk length is planned
    HIPCHK(hipStreamSynchronize(c->stream));
    drop_graph(c);
    pl.chunk = 0;
    pl.linear = false;
    pl.n_items = pl.n1 = pl.n2 = 0;
    if (linear && (C % SMOC_RUN_SPLIT != 0 || !c->wrun.p)) {  // refused by check_launch
        smoc_set_error("smoc_run: no linear chunks of %d segments for this stream", C);
        return SMOC_ERR_ARG;
    }
    std::vector<int64_t> items, tio, l1, l2;
    int64_t n = 0;
    try {
        tio.resize((size_t)c->T + 1);
        std::vector<std::pair<int64_t, int64_t>> p1, p2;  // (length, entry)
        const std::vector<int64_t> &off = c->h_track_off;
        int64_t sp = 0;  // the track's first entry in the split-point table
        for (int64_t t = 0; t < c->T; t++) {
            tio[t] = n;
            const int64_t b = off[t], e = off[t + 1];
            const int64_t nc = smoc_chunk_count(b, e, C);
            for (int64_t k = 0; k < nc; k++, n++) {
                const int64_t cb = smoc_chunk_begin(b, k, C);
                const int64_t ce = smoc_chunk_end(b, e, k, C);
                items.push_back(t);
                items.push_back(cb);
                items.push_back(ce);
                items.push_back(!linear ? cb
                                : k == 0 ? c->h_track_run_off[t]
                                         : c->h_split_run[sp + smoc_split_index(b, cb) - 1]);
                if (k == 0)
                    p1.emplace_back(ce - cb, (n << 2) | (nc > 1 ? SMOC_ITEM_EMIT : 0));
                else
                    p2.emplace_back(ce - cb, n << 2);
                if (k > 0 && (linear || k + 1 < nc))
                    p1.emplace_back(ce - cb, (n << 2) | (linear ? SMOC_ITEM_LIN : SMOC_ITEM_XFER));
            }
            sp += smoc_split_count(b, e);
        }
        tio[c->T] = n;
        auto longest_first = [](const std::pair<int64_t, int64_t> &x,
                                const std::pair<int64_t, int64_t> &y) { return x.first > y.first; };
        std::stable_sort(p1.begin(), p1.end(), longest_first);
        std::stable_sort(p2.begin(), p2.end(), longest_first);
        for (auto &x : p1)
            l1.push_back(x.second);
        for (auto &x : p2)
            l2.push_back(x.second);
    } catch (...) {
        smoc_set_error("smoc_run: host allocation for the chunk plan failed");
        return SMOC_ERR_NOMEM;
    }
    int rc = pl.items.alloc(std::max<size_t>(items.size(), 4));
    if (rc == SMOC_OK)
        rc = pl.list1.alloc(std::max<size_t>(l1.size(), 1));
    if (rc == SMOC_OK)
        rc = pl.list2.alloc(std::max<size_t>(l2.size(), 1));
    if (rc == SMOC_OK)
        rc = pl.track_item_off.alloc(tio.size());
    if (rc == SMOC_OK)
        rc = pl.chunk_psi.alloc((size_t)std::max<int64_t>(n, 1) * c->Gp);
    if (rc == SMOC_OK)
        rc = pl.xfer.alloc((size_t)std::max<int64_t>(n, 1) * 2 * c->Gp);
    if (rc != SMOC_OK)
        return rc;
    if (!items.empty())
        HIPCHK(hipMemcpy(pl.items.p, items.data(), sizeof(int64_t) * items.size(),
                         hipMemcpyHostToDevice));
    if (!l1.empty())
        HIPCHK(hipMemcpy(pl.list1.p, l1.data(), sizeof(int64_t) * l1.size(),
                         hipMemcpyHostToDevice));
    if (!l2.empty())
        HIPCHK(hipMemcpy(pl.list2.p, l2.data(), sizeof(int64_t) * l2.size(),
                         hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(pl.track_item_off.p, tio.data(), sizeof(int64_t) * tio.size(),
                     hipMemcpyHostToDevice));
    pl.chunk = C;
    pl.linear = linear;
    pl.n_items = n;
    pl.n1 = (int64_t)l1.size();
    pl.n2 = (int64_t)l2.size();
    return SMOC_OK;
}

// Device buffers a run uses beyond the context's inputs and outputs: the
// persistent work counters and, for a chunked sweep, the chunk plan.
// smoc_run takes them from context allocations (ctx_bufs); smoc_run_noalloc
// from its scratch array, where the plan is rebuilt on the device at the start
// of every run (plan_chunk > 0).
struct RunBufs {
    int64_t *queue = nullptr;  // [2] work counters of the two sweep phases
    int64_t *items = nullptr, *list1 = nullptr, *list2 = nullptr, *track_item_off = nullptr;
    float *chunk_psi = nullptr, *xfer = nullptr;
    int64_t n_items = 0, n1 = 0, n2 = 0;
    int plan_chunk = 0;
};

// Launch combinations the loaded problem cannot take.
static int check_launch(const smoc_ctx *c, const smoc_launch &l, const char *what)
{
    const bool runs = c->runs && !l.segment_flush;
    // FIXED64 too: the run sums are what is rounded to fixed point.  Runs also
    // end at split points (stream indices that are multiples of
    // SMOC_RUN_SPLIT), where chunks of a multiple of that length are cut, so
    // such chunks split no run; other lengths would make results depend on
    // the chunk length.
    if (l.chunk > 0 && (c->det() || c->fixed()) && runs && l.chunk % SMOC_RUN_SPLIT != 0) {
        smoc_set_error("%s: chunked sweeps of a region-sorted stream with a deterministic "
                       "or FIXED64 tally need a chunk length that is a multiple of %d "
                       "(got %d)",
                       what, SMOC_RUN_SPLIT, l.chunk);
        return SMOC_ERR_ARG;
    }
    // linear chunks need wrun, allocated for every stream that can be chunked
    // (a track longer than the chunk length, so longer than SMOC_RUN_SPLIT)
    if (linear_chunks(c, l.chunk, l.segment_flush != 0) && !c->wrun.p) {
        smoc_set_error("%s: no per-run buffer for chunks of this stream", what);
        return SMOC_ERR_STATE;
    }
    if (l.segment_flush && c->det() && c->runs) {
        smoc_set_error("%s: segment_flush on a region-sorted stream needs an atomic or "
                       "FIXED64 tally (the deterministic run bookkeeping follows the stream)",
                       what);
        return SMOC_ERR_ARG;
    }
    return SMOC_OK;
}

// The context's plan for l.chunk (built if it is another chunk length) and
// work counters (allocated on first use).
static int ctx_bufs(smoc_ctx *c, const smoc_launch &l, RunBufs *b)
{
    // a plan holds linear or transfer entries: rebuilt when the launch needs
    // the other kind (a segment_flush launch at the same chunk length)
    const bool linear = linear_chunks(c, l.chunk, l.segment_flush != 0);
    if (l.chunk > 0 && (c->plan.chunk != l.chunk || c->plan.linear != linear)) {
        int rc = build_plan(c, l.chunk, linear);
        if (rc != SMOC_OK)
            return rc;
    }
    if (l.persistent && !c->queue.p) {
        int rc = c->queue.alloc(2);
        if (rc != SMOC_OK)
            return rc;
    }
    *b = RunBufs{};
    b->queue = c->queue.p;
    if (l.chunk > 0) {
        Plan &pl = c->plan;
        b->items = pl.items.p;
        b->list1 = pl.list1.p;
        b->list2 = pl.list2.p;
        b->track_item_off = pl.track_item_off.p;
        b->chunk_psi = pl.chunk_psi.p;
        b->xfer = pl.xfer.p;
        b->n_items = pl.n_items;
        b->n1 = pl.n1;
        b->n2 = pl.n2;
    }
    return SMOC_OK;
}

// Enqueue one whole sweep on the context stream.  ev (4 events) may be NULL
// (graph capture): reset | plan + transfer + scan | sweep | reduce.
static int enqueue_run(smoc_ctx *c, const smoc_launch &l, const RunBufs &b, hipEvent_t *ev)
{
    const int G = c->p.egroups;
    const int gpl = l.groups_per_lane, wpb = l.waves_per_block;
    const int slices = (G + 64 * gpl - 1) / (64 * gpl);
    const int64_t nrow = c->n_fsr();
    const bool chunked = l.chunk > 0;
    // per-run sums unless the caller asks for a tally per segment
    const bool runs = c->runs && !l.segment_flush;

    smoc_sweep_args a{};
    a.tuples = c->tuples.p;
    a.len = c->len.p;
    a.row_thr = c->row_thr.p;
    a.code = c->code.p;
    a.zin = c->zin.p;
    a.track_off = c->track_off.p;
    a.psi_in = c->psi_in.p;
    a.psi_out = c->psi_out.p;
    a.flux = c->det() ? (void *)c->part.p : (void *)(c->flux.p + 256);
    a.track_run_off = c->track_run_off.p;
    a.table = c->table.p;
    a.tab_dx = c->tab_dx;
    a.tab_max = c->tab_max;
    a.tab_n = c->tab_n;
    a.G = G;
    a.Gp = c->Gp;
    a.F = c->p.fine_axial;
    a.slices = slices;
    a.runs = runs ? 1 : 0;
    a.split = c->split ? 1 : 0;
    a.fix_limit = c->fix_limit;
    // persistent grids: as many workgroups as can be resident (8 waves per
    // SIMD, 4 SIMDs per CU)
    const int64_t resident_blocks = (int64_t)c->cu_count * std::max(1, 32 / wpb);
    auto grid_for = [&](int64_t n_entries) {
        int64_t g = (n_entries * slices + wpb - 1) / wpb;
        if (l.persistent)
            g = std::min<int64_t>(g, resident_blocks);
        return g;
    };
    // persistent work batches: about 1/16 of a wave's fair share per counter
    // update (1 for the few long chunks of a batched run), at most 64
    auto batch_for = [&](int64_t n_entries) {
        const int64_t fair = n_entries * slices / std::max<int64_t>(1, resident_blocks * wpb);
        return (int)std::min<int64_t>(64, std::max<int64_t>(1, fair / 16));
    };

    if (ev)
        HIPCHK(hipEventRecord(ev[0], c->stream));
    if (c->fixed())  // counter (first 8 bytes) and the fixed-point sums
        LAUNCHCHK(smoc_launch_zero_i64(reinterpret_cast<int64_t *>(c->flux.p),
                                       nrow * c->Gp + 32, c->stream));
    else if (!c->det())
        LAUNCHCHK(smoc_launch_reset(c->flux.p + 256, c->flux_init.p, nrow * c->Gp,
                                    c->flux_f64() ? 1 : 0, c->stream));
    if (ev)
        HIPCHK(hipEventRecord(ev[1], c->stream));
    // linear chunks: every chunk begins a run (MODE_LIN)
    const bool linear = linear_chunks(c, l.chunk, l.segment_flush != 0);
    if (chunked) {
        if (b.plan_chunk > 0)
            LAUNCHCHK(smoc_launch_plan(
                c->track_off.p, c->T, b.plan_chunk, b.items, b.list1, b.list2,
                b.track_item_off, linear ? c->track_run_off.p : nullptr,
                linear ? c->split_run.p : nullptr, c->stream));
        if (linear && ev)  // the sweep time then covers phase 1, scan and correction
            HIPCHK(hipEventRecord(ev[2], c->stream));
        // phase 1: first chunks (swept from psi_in) + transfers of middle chunks,
        // or with linear chunks every chunk
        a.items = b.items;
        a.xfer = b.xfer;
        a.wrun = c->wrun.p;
        smoc_sweep_args p1 = a;
        p1.xfer_mode = 1;
        p1.item_list = b.list1;
        p1.chunk_psi = nullptr;
        p1.n_items = b.n1;
        if (l.persistent) {
            p1.queue = b.queue;
            p1.queue_batch = batch_for(b.n1);
            LAUNCHCHK(smoc_launch_zero_i64(b.queue, 1, c->stream));
        }
        LAUNCHCHK(smoc_launch_sweep(&p1, c->p.storage, gpl, c->p.exp_mode, c->p.tally, wpb,
                                    (int)grid_for(b.n1), c->stream));
        if (b.n2 > 0) {
            smoc_scan_args sa{b.track_item_off, b.xfer, c->psi_in.p, b.chunk_psi,
                              linear ? c->psi_out.p : nullptr, c->T, c->Gp, slices};
            LAUNCHCHK(smoc_launch_scan(&sa, gpl, c->stream));
        }
        if (linear) {
            // each run of a chunk after a track's first gets its incoming psi's share
            smoc_fix_args fa{b.items, b.list2, b.chunk_psi, c->wrun.p,
                             c->det() ? (void *)c->part.p : (void *)(c->flux.p + 256),
                             c->run_row.p, b.n2, b.n_items, c->n_runs, G, c->Gp, slices};
            LAUNCHCHK(smoc_launch_lin_fix(&fa, c->p.tally, gpl, c->stream));
            a.n_items = 0;  // no phase 2
        } else {
            if (ev)
                HIPCHK(hipEventRecord(ev[2], c->stream));
            // phase 2: the remaining chunks from the scanned psi
            a.item_list = b.list2;
            a.chunk_psi = b.chunk_psi;
            a.n_items = b.n2;
        }
    } else {
        a.n_items = c->T;
        if (ev)
            HIPCHK(hipEventRecord(ev[2], c->stream));
    }
    if (a.n_items > 0) {
        if (l.persistent) {
            a.queue = b.queue + 1;
            a.queue_batch = batch_for(a.n_items);
            LAUNCHCHK(smoc_launch_zero_i64(a.queue, 1, c->stream));
        }
        LAUNCHCHK(smoc_launch_sweep(&a, c->p.storage, gpl, c->p.exp_mode, c->p.tally, wpb,
                                    (int)grid_for(a.n_items), c->stream));
    }
    if (ev)
        HIPCHK(hipEventRecord(ev[3], c->stream));
    if (c->det()) {
        smoc_reduce_args r{};
        r.part = c->part.p;
        r.fsr_run_off = c->fsr_run_off.p;
        r.fsr_run_idx = c->fsr_run_idx.p;
        r.init = c->flux_init.p;
        r.flux = c->flux.p + 256;
        r.n_fsr = nrow;
        r.G = G;
        r.Gp = c->Gp;
        r.slices = slices;
        r.single = c->runs ? 0 : 1;
        LAUNCHCHK(smoc_launch_reduce(&r, c->p.tally, gpl, l.reduce_waves, c->stream));
    }
    if (c->fixed())
        LAUNCHCHK(smoc_launch_fixed_to_f64(c->fixed_flux(), c->flux_init.p, nrow * c->Gp,
                                           c->stream));
    if (ev)
        HIPCHK(hipEventRecord(ev[4], c->stream));
    return SMOC_OK;
}

static bool same_launch(const smoc_launch &x, const smoc_launch &y)
{
    return x.groups_per_lane == y.groups_per_lane && x.waves_per_block == y.waves_per_block &&
           x.chunk == y.chunk && x.persistent == y.persistent &&
           x.reduce_waves == y.reduce_waves && x.segment_flush == y.segment_flush;
}

// One eager run on the given buffers, timed with the context's events.
static int run_eager(smoc_ctx *c, const smoc_launch &l, const RunBufs &b, smoc_timing *tm)
{
    int rc = enqueue_run(c, l, b, c->ev);
    if (rc != SMOC_OK)
        return rc;
    HIPCHK(hipEventSynchronize(c->ev[4]));
    float ms[4] = {0.f, 0.f, 0.f, 0.f};
    for (int i = 0; i < 4; i++)
        HIPCHK(hipEventElapsedTime(&ms[i], c->ev[i], c->ev[i + 1]));
    *tm = smoc_timing{};
    tm->reset_ms = ms[0];
    tm->xfer_ms = ms[1];
    tm->sweep_ms = ms[2];
    tm->reduce_ms = ms[3];
    tm->total_ms = (double)ms[0] + ms[1] + ms[2] + ms[3];
    return SMOC_OK;
}

// One run, timed with events on the context stream; with use_graph the
// sequence is captured once into a hipGraph and replayed.
static int run_once(smoc_ctx *c, const smoc_launch &l, smoc_timing *tm)
{
    int rc = check_launch(c, l, "smoc_run");
    if (rc != SMOC_OK)
        return rc;
    RunBufs b;
    if (l.use_graph) {
        if (!c->graph_exec || !same_launch(c->graph_launch, l)) {
            drop_graph(c);
            // plan buffers must exist before capture (no allocation inside)
            rc = ctx_bufs(c, l, &b);
            if (rc != SMOC_OK)
                return rc;
            hipGraph_t g = nullptr;
            HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
            rc = enqueue_run(c, l, b, nullptr);
            hipError_t e = hipStreamEndCapture(c->stream, &g);
            if (rc != SMOC_OK) {
                if (g)
                    (void)hipGraphDestroy(g);
                return rc;
            }
            if (e != hipSuccess) {
                smoc_set_error("hipStreamEndCapture: %s", hipGetErrorString(e));
                return SMOC_ERR_HIP;
            }
            e = hipGraphInstantiate(&c->graph_exec, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (e != hipSuccess) {
                c->graph_exec = nullptr;
                smoc_set_error("hipGraphInstantiate: %s", hipGetErrorString(e));
                return SMOC_ERR_HIP;
            }
            c->graph_launch = l;
        }
        HIPCHK(hipEventRecord(c->ev[0], c->stream));
        HIPCHK(hipGraphLaunch(c->graph_exec, c->stream));
        HIPCHK(hipEventRecord(c->ev[4], c->stream));
        HIPCHK(hipEventSynchronize(c->ev[4]));
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, c->ev[0], c->ev[4]));
        *tm = smoc_timing{};
        tm->total_ms = ms;
        tm->sweep_ms = ms;  // one graph: phases are not timed separately
        return SMOC_OK;
    }
    rc = ctx_bufs(c, l, &b);
    if (rc != SMOC_OK)
        return rc;
    return run_eager(c, l, b, tm);
}

// Measures candidate launch shapes on the loaded problem (each run is a
// complete, correct sweep) and keeps the fastest: first the track split and
// persistence, then groups per lane, then waves per workgroup.  Candidates run
// eagerly (the graph of the chosen shape replays the same kernels).  Each one
// is timed in alternation with the current best, so that a drift of the GPU's
// clocks during the search favours neither, and must beat it by more than the
// timing noise.
static int autotune(smoc_ctx *c, const smoc_launch &base, smoc_launch *best_out)
{
    auto time_once = [&](const smoc_launch &l, double *ms) {
        smoc_timing t{};
        int rc = run_once(c, l, &t);
        *ms = t.total_ms;
        return rc;
    };
    smoc_launch best = base;
    best.autotune = 0;
    best.use_graph = 0;
    best.chunk = 0;
    best.persistent = 0;
    // runs of the starting shape first, for at least 200 ms of GPU time, so
    // that no candidate is timed on a GPU still raising its clocks
    {
        double spent = 0.0, ms = 0.0;
        for (int r = 0; r < 1000 && (r < 3 || spent < 200.0); r++) {
            int rc = time_once(best, &ms);
            if (rc != SMOC_OK)
                return rc;
            spent += ms;
        }
    }
    auto consider = [&](smoc_launch cand) {
        fill_launch(c, &cand, &cand);
        if (same_launch(cand, best))
            return SMOC_OK;
        double ms = 0.0, t_best = 1e30, t_cand = 1e30;
        int r = time_once(cand, &ms);  // warm-up (plans, caches)
        if (r == SMOC_ERR_ARG)         // combination not supported for this problem
            return SMOC_OK;
        if (r != SMOC_OK)
            return r;
        for (int k = 0; k < 5; k++) {
            if ((r = time_once(best, &ms)) != SMOC_OK)
                return r;
            t_best = std::min(t_best, ms);
            if ((r = time_once(cand, &ms)) != SMOC_OK)
                return r;
            t_cand = std::min(t_cand, ms);
        }
        if (t_cand < 0.99 * t_best)  // a change must gain more than timing noise
            best = cand;
        return SMOC_OK;
    };
    int rc = SMOC_OK;
    const int64_t mean_len = c->T > 0 ? c->S / c->T : 0;
    {
        smoc_launch cand = best;
        cand.persistent = 1;
        if ((rc = consider(cand)) != SMOC_OK)
            return rc;
    }
    // coordinate search over chunk length (with persistence), groups per lane
    // and waves per workgroup; a second pass when the first changed anything,
    // since the best chunk length can depend on the workgroup shape
    for (int pass = 0; pass < 2; pass++) {
        const smoc_launch before = best;
        if (c->max_track > 4 * std::max<int64_t>(mean_len, 1)) {
            const smoc_launch base_split = best;
            for (int C : {1024, 2048, 4096, 8192, 16384}) {
                for (int pers = 0; pers < 2; pers++) {
                    smoc_launch cand = base_split;
                    cand.chunk = C;
                    cand.seg_per_block = 0;
                    cand.persistent = pers;
                    if ((rc = consider(cand)) != SMOC_OK)
                        return rc;
                }
            }
        }
        if (c->p.egroups > 64) {
            smoc_launch cand = best;
            cand.groups_per_lane = 3 - best.groups_per_lane;
            if ((rc = consider(cand)) != SMOC_OK)
                return rc;
        }
        for (int w : {1, 2, 4, 8, 16}) {
            smoc_launch cand = best;
            cand.waves_per_block = w;
            if ((rc = consider(cand)) != SMOC_OK)
                return rc;
        }
        if (same_launch(best, before))
            break;
    }
    if (c->det()) {  // the reduction's tile: waves per workgroup
        for (int w : {1, 2, 8, 16}) {
            smoc_launch cand = best;
            cand.reduce_waves = w;
            if ((rc = consider(cand)) != SMOC_OK)
                return rc;
        }
    }
    best.use_graph = base.use_graph;
    *best_out = best;
    return SMOC_OK;
}

// After a completed run: the FIXED64 range report, the results flag and the
// caller's timing record (n_items: work entries before the group slices).
static int finish_run(smoc_ctx *c, const smoc_launch &l, const smoc_timing &tm, int64_t n_items,
                      smoc_timing *timing, const char *what)
{
    if (c->fixed()) {
        uint32_t overflows = 0;
        HIPCHK(hipMemcpy(&overflows, c->flux.p + 248, sizeof(overflows), hipMemcpyDeviceToHost));
        if (overflows) {
            smoc_set_error("%s: %u FIXED64 run sums exceeded the per-contribution bound "
                           "and were dropped; fine_flux is incomplete (use an FP64 tally)",
                           what, overflows);
            return SMOC_ERR_RANGE;
        }
    }
    c->ran = true;
    if (timing) {
        const int G = c->p.egroups;
        const int slices = (G + 64 * l.groups_per_lane - 1) / (64 * l.groups_per_lane);
        *timing = tm;
        timing->segments = c->S;
        timing->work_items = n_items * slices;
        timing->used = l;
    }
    return SMOC_OK;
}

int smoc_run(smoc_ctx *c, const smoc_launch *launch, smoc_timing *timing)
{
    if (!c) {
        smoc_set_error("smoc_run: NULL context");
        return SMOC_ERR_ARG;
    }
    if (!c->regions_set || !c->segments_set) {
        smoc_set_error("smoc_run: call smoc_set_regions and smoc_set_segments first");
        return SMOC_ERR_STATE;
    }
    HIPCHK(hipSetDevice(c->p.device));
    smoc_launch l;
    fill_launch(c, launch, &l);
    if (l.autotune) {
        if (!c->tuned_valid) {
            int rc = autotune(c, l, &c->tuned);
            if (rc != SMOC_OK)
                return rc;
            c->tuned_valid = true;
        }
        const int graph = l.use_graph;
        l = c->tuned;
        l.use_graph = graph;
        l.autotune = 1;
    }
    smoc_timing tm{};
    int rc = run_once(c, l, &tm);
    if (rc != SMOC_OK)
        return rc;
    return finish_run(c, l, tm, l.chunk > 0 ? c->plan.n_items : c->T, timing, "smoc_run");
}

int smoc_run_noalloc(smoc_ctx *c, const smoc_launch *launch, smoc_timing *timing)
{
    const char *what = "smoc_run_noalloc";
    if (!c) {
        smoc_set_error("%s: NULL context", what);
        return SMOC_ERR_ARG;
    }
    if (!c->regions_set || !c->segments_set) {
        smoc_set_error("%s: call smoc_set_regions and smoc_set_segments first", what);
        return SMOC_ERR_STATE;
    }
    smoc_launch l;
    fill_launch(c, launch, &l);
    if (l.autotune || l.use_graph) {
        smoc_set_error("%s: autotune and use_graph allocate; pass them as 0", what);
        return SMOC_ERR_ARG;
    }
    int rc = check_launch(c, l, what);
    if (rc != SMOC_OK)
        return rc;
    HIPCHK(hipSetDevice(c->p.device));
    unsigned char *base = scratch_of(c->p.device);
    if (!base) {
        smoc_set_error("%s: no scratch address for device %d", what, c->p.device);
        return SMOC_ERR_STATE;
    }
    // layout: work counters, then the chunk plan (items, the two work lists,
    // track_item_off, chunk_psi, xfer), each 256-byte aligned; sizes from the
    // host copy of the track offsets, as the device plan will make them
    RunBufs b;
    size_t off = 0;
    auto take = [&](size_t bytes) {
        unsigned char *p = base + off;
        off += (bytes + 255) / 256 * 256;
        return p;
    };
    b.queue = reinterpret_cast<int64_t *>(take(2 * sizeof(int64_t)));
    if (l.chunk > 0) {
        const int64_t C = l.chunk;
        const bool lin = linear_chunks(c, l.chunk, l.segment_flush != 0);
        int64_t n = 0, n1 = 0, n2 = 0;
        for (int64_t t = 0; t < c->T; t++) {
            const int64_t nc = smoc_chunk_count(c->h_track_off[t], c->h_track_off[t + 1], C);
            n += nc;
            n1 += lin ? nc : nc > 1 ? nc - 1 : 1;  // linear chunks: all
            n2 += nc - 1;
        }
        const size_t gp = (size_t)c->Gp;
        const size_t need = 256 + ((size_t)n * 32 + 255) / 256 * 256 +
                            ((size_t)std::max<int64_t>(n1, 1) * 8 + 255) / 256 * 256 +
                            ((size_t)std::max<int64_t>(n2, 1) * 8 + 255) / 256 * 256 +
                            ((size_t)(c->T + 1) * 8 + 255) / 256 * 256 +
                            ((size_t)n * gp * 4 + 255) / 256 * 256 +
                            ((size_t)n * gp * 8 + 255) / 256 * 256;
        if (need > (size_t)SMOC_NOALLOC_SCRATCH_BYTES) {
            smoc_set_error("%s: the plan of %lld chunks of %d segments needs %zu bytes of "
                           "scratch, more than its %zu (use longer chunks or smoc_run)",
                           what, (long long)n, l.chunk, need,
                           (size_t)SMOC_NOALLOC_SCRATCH_BYTES);
            return SMOC_ERR_ARG;
        }
        b.items = reinterpret_cast<int64_t *>(take((size_t)n * 32));
        b.list1 = reinterpret_cast<int64_t *>(take((size_t)std::max<int64_t>(n1, 1) * 8));
        b.list2 = reinterpret_cast<int64_t *>(take((size_t)std::max<int64_t>(n2, 1) * 8));
        b.track_item_off = reinterpret_cast<int64_t *>(take((size_t)(c->T + 1) * 8));
        b.chunk_psi = reinterpret_cast<float *>(take((size_t)n * gp * 4));
        b.xfer = reinterpret_cast<float *>(take((size_t)n * gp * 8));
        b.n_items = n;
        b.n1 = n1;
        b.n2 = n2;
        b.plan_chunk = l.chunk;  // rebuilt in the scratch by this run
    }
    smoc_timing tm{};
    rc = run_eager(c, l, b, &tm);
    if (rc != SMOC_OK)
        return rc;
    return finish_run(c, l, tm, l.chunk > 0 ? b.n_items : c->T, timing, what);
}

int smoc_get_results(smoc_ctx *c, double *fine_flux, float *psi_out)
{
    if (!c) {
        smoc_set_error("smoc_get_results: NULL context");
        return SMOC_ERR_ARG;
    }
    if (!c->ran) {
        smoc_set_error("smoc_get_results: no completed smoc_run since the last set_*");
        return SMOC_ERR_STATE;
    }
    HIPCHK(hipSetDevice(c->p.device));
    const int G = c->p.egroups, Gp = c->Gp;
    const int64_t nrow = c->n_fsr();
    int rc = SMOC_OK;
    if (fine_flux) {
        if (c->flux_f64()) {
            rc = rows_to_host(fine_flux, reinterpret_cast<const double *>(c->flux.p + 256), Gp,
                              nrow, G, c->stream);
        } else {
            std::vector<float> tmp((size_t)nrow * G);
            rc = rows_to_host(tmp.data(), reinterpret_cast<const float *>(c->flux.p + 256), Gp,
                              nrow, G, c->stream);
            for (size_t i = 0; rc == SMOC_OK && i < tmp.size(); i++)
                fine_flux[i] = tmp[i];
        }
    }
    if (rc == SMOC_OK && psi_out && c->T > 0)
        rc = rows_to_host(psi_out, c->psi_out.p, Gp, c->T, G, c->stream);
    return rc;
}

int smoc_set_scattering(smoc_ctx *c, const float *scatter, const float *q_ext)
{
    if (!c || !scatter) {
        smoc_set_error("smoc_set_scattering: NULL argument");
        return SMOC_ERR_ARG;
    }
    if (!c->regions_set) {
        smoc_set_error("smoc_set_scattering: call smoc_set_regions first");
        return SMOC_ERR_STATE;
    }
    HIPCHK(hipSetDevice(c->p.device));
    const int G = c->p.egroups, R = c->p.regions, F = c->p.fine_axial;
    const int64_t nsc = (int64_t)R * G, nsrc = (int64_t)R * F * G;
    for (int64_t i = 0; i < nsc; i++)
        if (!(scatter[i] >= 0.f && scatter[i] < 1.f)) {
            smoc_set_error("smoc_set_scattering: scatter[%lld] = %g must be in [0, 1)",
                           (long long)i, (double)scatter[i]);
            return SMOC_ERR_ARG;
        }
    if (q_ext)
        for (int64_t i = 0; i < nsrc; i++)
            if (!std::isfinite(q_ext[i])) {
                smoc_set_error("smoc_set_scattering: q_ext[%lld] is not finite", (long long)i);
                return SMOC_ERR_ARG;
            }
    int rc = c->scatter.alloc((size_t)nsc);
    if (rc == SMOC_OK)
        rc = c->q_ext.alloc((size_t)nsrc);
    if (rc == SMOC_OK)
        rc = c->upd.alloc((size_t)(2 * smoc_update_blocks(nsrc) + 2));
    if (rc != SMOC_OK)
        return rc;
    HIPCHK(hipMemcpy(c->scatter.p, scatter, sizeof(float) * nsc, hipMemcpyHostToDevice));
    if (q_ext)
        HIPCHK(hipMemcpy(c->q_ext.p, q_ext, sizeof(float) * nsrc, hipMemcpyHostToDevice));
    else  // the current source is the external one
        HIPCHK(hipMemcpy(c->q_ext.p, c->src_m.p, sizeof(float) * nsrc, hipMemcpyDeviceToDevice));
    c->scatter_set = true;
    return SMOC_OK;
}

// New source from the last sweep's fine_flux, repacked into the tuples the
// next sweep reads (same buffers, so a captured sweep graph stays valid).
static int source_update(smoc_ctx *c, double *change, double *ms)
{
    const int G = c->p.egroups, R = c->p.regions, F = c->p.fine_axial;
    const int64_t nsrc = (int64_t)R * F * G, nblk = smoc_update_blocks(nsrc);
    smoc_update_args u{};
    u.flux = c->flux.p + 256;
    u.vol = c->vol.p;
    u.sigT = c->sig_m.p;
    u.scatter = c->scatter.p;
    u.q_ext = c->q_ext.p;
    u.source = c->src_m.p;
    u.partial = c->upd.p;
    u.sums = c->upd.p + 2 * nblk;
    u.G = G, u.Gp = c->Gp, u.R = R, u.F = F;
    HIPCHK(hipEventRecord(c->ev[0], c->stream));
    LAUNCHCHK(smoc_launch_source_update(&u, c->flux_f64() ? 1 : 0, c->stream));
    smoc_pack_args pa{c->sig_m.p, c->src_m.p, c->tuples.p, G, c->Gp, R, F};
    LAUNCHCHK(smoc_launch_pack(&pa, c->p.storage, c->stream));
    HIPCHK(hipEventRecord(c->ev[1], c->stream));
    double sums[2] = {0.0, 0.0};
    HIPCHK(hipMemcpyAsync(sums, u.sums, sizeof(sums), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (ms) {
        float t = 0.f;
        HIPCHK(hipEventElapsedTime(&t, c->ev[0], c->ev[1]));
        *ms = t;
    }
    if (change)
        *change = sums[1] > 0.0 ? std::sqrt(sums[0] / sums[1]) : std::sqrt(sums[0]);
    return SMOC_OK;
}

int smoc_source_update(smoc_ctx *c, double *change)
{
    if (!c) {
        smoc_set_error("smoc_source_update: NULL context");
        return SMOC_ERR_ARG;
    }
    if (!c->scatter_set || !c->ran) {
        smoc_set_error("smoc_source_update: needs smoc_set_scattering and a completed smoc_run");
        return SMOC_ERR_STATE;
    }
    HIPCHK(hipSetDevice(c->p.device));
    return source_update(c, change, nullptr);
}

int smoc_iterate(smoc_ctx *c, const smoc_launch *launch, int n_outer, smoc_timing *timing,
                 double *update_ms, double *change)
{
    if (!c || n_outer < 1) {
        smoc_set_error("smoc_iterate: NULL context or n_outer < 1");
        return SMOC_ERR_ARG;
    }
    if (!c->scatter_set) {
        smoc_set_error("smoc_iterate: call smoc_set_scattering first");
        return SMOC_ERR_STATE;
    }
    smoc_timing sum{}, tm{};
    double upd_ms = 0.0, chg = 0.0;
    for (int it = 0; it < n_outer; it++) {
        int rc = smoc_run(c, launch, &tm);
        if (rc != SMOC_OK)
            return rc;
        double ms = 0.0;
        rc = source_update(c, &chg, &ms);
        if (rc != SMOC_OK)
            return rc;
        upd_ms += ms;
        sum.total_ms += tm.total_ms;
        sum.reset_ms += tm.reset_ms;
        sum.xfer_ms += tm.xfer_ms;
        sum.sweep_ms += tm.sweep_ms;
        sum.reduce_ms += tm.reduce_ms;
        sum.segments += tm.segments;
        sum.work_items = tm.work_items;
        sum.used = tm.used;
    }
    if (timing)
        *timing = sum;
    if (update_ms)
        *update_ms = upd_ms;
    if (change)
        *change = chg;
    return SMOC_OK;
}

int smoc_get_source(smoc_ctx *c, float *fine_source)
{
    if (!c || !fine_source) {
        smoc_set_error("smoc_get_source: NULL argument");
        return SMOC_ERR_ARG;
    }
    if (!c->regions_set) {
        smoc_set_error("smoc_get_source: call smoc_set_regions first");
        return SMOC_ERR_STATE;
    }
    HIPCHK(hipSetDevice(c->p.device));
    const int64_t nsrc = (int64_t)c->p.regions * c->p.fine_axial * c->p.egroups;
    HIPCHK(hipMemcpy(fine_source, c->src_m.p, sizeof(float) * nsrc, hipMemcpyDeviceToHost));
    return SMOC_OK;
}

}  // extern "C"
