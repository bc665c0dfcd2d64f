// cf_engine.hip -- host side of the MI355X theory pair-counting engine and its C ABI.
//
// Implements the five reference entry points (include/corrfunc_theory.h):
// option handling, bin setup, grid selection and the finalisation (autocorr
// doubling, self pairs, averages, analytic RR) run on the host exactly as the
// reference drivers do; gridding, the cell-pair work list and the pair counting
// run on the GPU (cf_kernels.hip).
#include <hip/hip_runtime.h>

#include <ctype.h>
#include <math.h>
#include <signal.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/time.h>

#include <algorithm>
#include <array>
#include <atomic>
#include <memory>
#include <mutex>
#include <new>
#include <stdexcept>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "cf_kernels.hip"
#include "cf_sched.hip"
#include "cf_model.hip"
#include "cf_bgraph.hip"
#include "../../include/corrfunc_theory.h"
#include "../../include/corrfunc_amd.h"
#include "../../include/corrfunc_mocks.h"

namespace cfamd {

#define CF_CHECK(call)                                                                          \
    do {                                                                                        \
        hipError_t e_ = (call);                                                                 \
        if (e_ != hipSuccess) {                                                                 \
            fprintf(stderr, "corrfunc_amd: HIP error %s at %s:%d (%s)\n", hipGetErrorString(e_), \
                    __FILE__, __LINE__, #call);                                                 \
            throw HipFailure();                                                                 \
        }                                                                                       \
    } while (0)

struct HipFailure {};
struct InputError {};

// --------------------------------------------------------------- state
// Grow-only device buffers, reused across calls (the entry points are not
// re-entrant, like the reference's). In fixed mode (cf_amd_countpairs_noalloc)
// every buffer of the call is carved out of one caller-independent device
// region instead, 256-byte aligned, and nothing is allocated or freed; a call
// that needs more than the region throws InputError after printing why.
struct Arena {
    std::unordered_map<std::string, std::pair<void*, size_t>> slots;
    char* fixed = nullptr;   // base of the region in fixed mode, else nullptr
    size_t fixed_cap = 0, fixed_off = 0;
    std::unordered_map<std::string, std::pair<void*, size_t>> fixed_slots;
    void begin_fixed(void* base, size_t cap)
    {
        fixed = (char*)base;
        fixed_cap = cap;
        fixed_off = 0;
        fixed_slots.clear();
    }
    void end_fixed()
    {
        fixed = nullptr;
        fixed_slots.clear();
    }
    void* get(const std::string& name, size_t bytes)
    {
        if (bytes == 0) bytes = 256;
        if (fixed) {
            auto f = fixed_slots.find(name);
            if (f != fixed_slots.end() && f->second.second >= bytes) return f->second.first;
            const size_t off = (fixed_off + 255) & ~(size_t)255;
            if (off > fixed_cap || bytes > fixed_cap - off) {
                fprintf(stderr, "Error: the call needs more than the %zu bytes of the engine's fixed device "
                                "scratch region (at %s: %zu bytes more); use the allocating entry point for "
                                "inputs this large\n", fixed_cap, name.c_str(), bytes);
                throw InputError();
            }
            fixed_off = off + bytes;
            fixed_slots[name] = {fixed + off, bytes};
            return fixed + off;
        }
        auto it = slots.find(name);
        if (it != slots.end() && it->second.second >= bytes) return it->second.first;
        if (it != slots.end()) {
            // queued work may still read the old buffer
            CF_CHECK(hipDeviceSynchronize());
            CF_CHECK(hipFree(it->second.first));
            slots.erase(it);
        }
        void* p = nullptr;
        size_t cap = bytes + bytes / 8;
        CF_CHECK(hipMalloc(&p, cap));
        slots[name] = {p, cap};
        return p;
    }
};

// The work-list cost model: network parameters on the device, and the store
// of recorded (features, log2 ticks) training samples.
struct CostModel {
    bool trained = false;
    float* d_params = nullptr;  // kNumParams
    float* d_norm = nullptr;    // kNumNorm
    bool recording = false;
    int cap = 1 << 21;          // samples kept in the store
    float* d_sx = nullptr;      // [cap][kNumFeatures]
    float* d_sy = nullptr;      // [cap]
    int* d_count = nullptr;     // samples appended so far (may exceed cap)
};

// Device scratch of cf_amd_countpairs_noalloc: every buffer that entry point's
// kernels use besides its inputs and outputs, carved from this array (the
// engine's fixed arena mode). Each GPU the library runs on holds its own copy
// for the life of the process. Host code never passes the array's name to a
// kernel or a copy: it is a host-side stand-in; the device address comes from
// hipGetSymbolAddress (Engine::scratch_ptr).
constexpr size_t kScratchBytes = CF_AMD_NOALLOC_SCRATCH_BYTES;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

struct Engine {
    bool init = false;
    int device = 0;
    void* scratch_ptr = nullptr;  // device address of `scratch` on `device` (nullptr: lookup failed)
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr, ev_list = nullptr;
    Arena arena;
    int engine_kind = CF_AMD_ENGINE_TUNED;
    int sched_kind = CF_AMD_SCHED_LEARNED;
    cf_amd_stats last;
    int ncu = 256;
    size_t lds_max = 64 * 1024; // LDS one workgroup may allocate (160 KiB on gfx950)
    double ns_per_tick = 10.0;  // device wall clock period
    CostModel model;
    DevLattice hlats[2];
    // autotuner
    bool autotune = false;
    bool tuning = false;            // inside a candidate run
    int force_R = 0;                // histogram replication override (0: automatic)
    int force_fetch = 0;            // queue batch override (0: CORRFUNC_AMD_FETCH or the work-list rule)
    int last_autotuned = 0;         // 0 off, 1 cached choice, 2 tuned by this call
    bool tune_loaded = false;       // CORRFUNC_AMD_TUNE_CACHE read
    std::unordered_map<std::string, std::array<int, 5>> tune_cache;  // refine x, y, z, R, fetch
};

static Engine& eng()
{
    static Engine E;
    if (!E.init) {
        CF_CHECK(hipGetDevice(&E.device));
        CF_CHECK(hipStreamCreateWithFlags(&E.stream, hipStreamNonBlocking));
        CF_CHECK(hipEventCreate(&E.ev0));
        CF_CHECK(hipEventCreate(&E.ev1));
        CF_CHECK(hipEventCreate(&E.ev_list));
        const char* k = getenv("CORRFUNC_AMD_ENGINE");
        if (k && strcmp(k, "plain") == 0) E.engine_kind = CF_AMD_ENGINE_PLAIN;
        if (const char* sk = getenv("CORRFUNC_AMD_SCHED")) {
            if (strcmp(sk, "none") == 0) E.sched_kind = CF_AMD_SCHED_NONE;
            else if (strcmp(sk, "heuristic") == 0) E.sched_kind = CF_AMD_SCHED_HEURISTIC;
        }
        hipDeviceProp_t prop;
        CF_CHECK(hipGetDeviceProperties(&prop, E.device));
        E.ncu = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
        if (prop.sharedMemPerBlock >= 64 * 1024) E.lds_max = prop.sharedMemPerBlock;
        if (const char* lm = getenv("CORRFUNC_AMD_LDS_MAX")) {
            const long v = atol(lm);
            if (v >= 32 * 1024 && (size_t)v < E.lds_max) E.lds_max = (size_t)v;
        }
        int khz = 0;
        if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, E.device) == hipSuccess && khz > 0)
            E.ns_per_tick = 1e6 / khz;
        if (const char* at = getenv("CORRFUNC_AMD_AUTOTUNE")) E.autotune = atoi(at) != 0;
        memset(&E.last, 0, sizeof E.last);
        // the scratch array's address on this device, once, with the device current
        void* sp = nullptr;
        const hipError_t se = hipGetSymbolAddress(&sp, HIP_SYMBOL(scratch));
        if (se == hipSuccess) {
            E.scratch_ptr = sp;
        } else {
            (void)hipGetLastError();
            fprintf(stderr, "corrfunc_amd: hipGetSymbolAddress(scratch) failed: %s; cf_amd_countpairs_noalloc is "
                            "unavailable\n", hipGetErrorString(se));
        }
        E.init = true;
    }
    return E;
}

static double now_s()
{
    struct timeval t;
    gettimeofday(&t, NULL);
    return t.tv_sec + 1e-6 * t.tv_usec;
}

// --------------------------------------------------------- interrupts
// As the reference (utils/macros.h:145-167, theory/DD/countpairs_impl.c.src:
// 31-37, 191-193, 554-566): while a counting call runs, SIGTERM, SIGINT and
// SIGHUP are caught; the call then stops at its next phase boundary (after
// gridding, after the cell pairs, after the pair kernel) and returns
// EXIT_FAILURE. Unlike the reference, the previous handlers are restored on
// every return and the next call starts uninterrupted. The outermost scope of
// nested engine calls (the autotuner's candidate runs) installs the handlers.
static volatile sig_atomic_t g_interrupt_signo = 0;
static int g_interrupt_depth = 0;

static void on_interrupt(int signo) { g_interrupt_signo = signo; }

struct InterruptScope {
    static constexpr int kSignals[3] = {SIGTERM, SIGINT, SIGHUP};
    struct sigaction prev[3];  // the caller's dispositions, flags and masks, restored exactly
    bool saved[3] = {false, false, false};
    bool outer;
    InterruptScope() : outer(g_interrupt_depth++ == 0)
    {
        if (!outer) return;
        g_interrupt_signo = 0;
        struct sigaction sa;
        memset(&sa, 0, sizeof sa);
        sa.sa_handler = on_interrupt;
        sigemptyset(&sa.sa_mask);
        sa.sa_flags = SA_RESTART;
        for (int i = 0; i < 3; i++) {
            saved[i] = sigaction(kSignals[i], &sa, &prev[i]) == 0;
            if (!saved[i]) fprintf(stderr, "Can not handle signal = %d\n", kSignals[i]);
        }
    }
    ~InterruptScope()
    {
        g_interrupt_depth--;
        if (!outer) return;
        for (int i = 0; i < 3; i++)
            if (saved[i] && sigaction(kSignals[i], &prev[i], nullptr) != 0)
                fprintf(stderr, "Could not reset signal handler for signal = %d\n", kSignals[i]);
    }
};

// Throws InputError (the entry point returns EXIT_FAILURE) after a caught signal.
// CORRFUNC_AMD_TEST_SIGNAL=<signo> raises that signal at the first check, to
// exercise the path from a test.
static void check_interrupt()
{
    if (const char* ts = getenv("CORRFUNC_AMD_TEST_SIGNAL"))
        if (!g_interrupt_signo && atoi(ts) > 0) raise(atoi(ts));
    if (g_interrupt_signo) {
        // nothing queued may still write into buffers the unwinding frees
        (void)hipStreamSynchronize(eng().stream);
        fprintf(stderr, "Received signal = `%s' (signo = %d). Aborting \n", strsignal(g_interrupt_signo),
                (int)g_interrupt_signo);
        throw InputError();
    }
}

// Clears device words with a fill kernel on the engine's stream (no memset
// nodes, so the sequence can be captured into a graph).
static void fill_i32(int* p, int64_t n, int v)
{
    Engine& E = eng();
    if (n <= 0) return;
    k_fill_i32<<<(unsigned)std::min<int64_t>(1024, (n + 255) / 256), 256, 0, E.stream>>>(n, p, v);
    CF_CHECK(hipGetLastError());
}
static void fill_u64(void* p, int64_t n)
{
    Engine& E = eng();
    if (n <= 0) return;
    k_fill_u64<<<(unsigned)std::min<int64_t>(1024, (n + 255) / 256), 256, 0, E.stream>>>(
        n, (unsigned long long*)p, 0ull);
    CF_CHECK(hipGetLastError());
}

// -------------------------------------------------- bins (utils.c:62-106)
static int setup_bins(const char* fname, double* rmin, double* rmax, int* nbin, std::vector<double>& rupp)
{
    FILE* fp = fopen(fname, "rt");
    if (!fp) {
        fprintf(stderr, "Error: could not open bin file `%s'\n", fname);
        return EXIT_FAILURE;
    }
    char line[10000];
    int64_t nl = 0;
    while (fgets(line, sizeof line, fp)) {
        char* c = line;
        while (*c != '\0' && isspace((unsigned char)*c)) c++;
        if (*c != '\0' && *c != '#') nl++;
    }
    fclose(fp);
    *nbin = (int)nl + 1;
    rupp.assign((size_t)*nbin + 1, 0.0);
    fp = fopen(fname, "r");
    if (!fp) return EXIT_FAILURE;
    char buf[1000];
    int index = 1;
    double lo, hi;
    while (fgets(buf, sizeof buf, fp)) {
        if (sscanf(buf, "%lf %lf", &lo, &hi) == 2) {
            if (index == 1) {
                *rmin = lo;
                rupp[0] = lo;
            }
            if (index < *nbin + 1) rupp[index] = hi;
            index++;
        }
    }
    fclose(fp);
    *rmax = rupp[index - 1 < *nbin ? index - 1 : *nbin];
    rupp[*nbin] = *rmax;
    rupp[*nbin - 1] = *rmax;
    return EXIT_SUCCESS;
}

// ------------------------------------------------ input residency
static bool is_device_ptr(const void* p)
{
    if (!p) return false;
    hipPointerAttribute_t a;
    hipError_t e = hipPointerGetAttributes(&a, p);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    return a.type == hipMemoryTypeDevice;
}

// Returns a device pointer holding n elements of `esize` bytes from `p`
// (copied when p is host memory).
static const void* to_device(const std::string& slot, const void* p, int64_t n, size_t esize)
{
    if (!p) return nullptr;
    if (is_device_ptr(p)) return p;
    Engine& E = eng();
    void* d = E.arena.get(slot, (size_t)n * esize);
    CF_CHECK(hipMemcpyAsync(d, p, (size_t)n * esize, hipMemcpyHostToDevice, E.stream));
    return d;
}

// ----------------------------------------------------------- helpers
// Data extent (x, y, z min then max) into res[0..5] on the device (no sync).
static void device_minmax_async(const void* X, const void* Y, const void* Z, int64_t n, size_t esize, double* res)
{
    Engine& E = eng();
    const int nb = 512;
    double* part = (double*)E.arena.get("mm_part", sizeof(double) * 6 * nb);
    if (esize == 4)
        k_minmax<float><<<nb, 256, 0, E.stream>>>(n, (const float*)X, (const float*)Y, (const float*)Z, part);
    else
        k_minmax<double><<<nb, 256, 0, E.stream>>>(n, (const double*)X, (const double*)Y, (const double*)Z, part);
    k_minmax_final<<<1, 64, 0, E.stream>>>(nb, part, res);
    CF_CHECK(hipGetLastError());
}

static void device_minmax(const void* X, const void* Y, const void* Z, int64_t n, size_t esize, double out[6])
{
    Engine& E = eng();
    double* res = (double*)E.arena.get("mm_res", sizeof(double) * 6);
    device_minmax_async(X, Y, Z, n, esize, res);
    CF_CHECK(hipMemcpyAsync(out, res, sizeof(double) * 6, hipMemcpyDeviceToHost, E.stream));
    CF_CHECK(hipStreamSynchronize(E.stream));
}

// exclusive scan of n int32 values; out has n+1 entries (out[n] = total)
static void device_scan(const int* in, int* out, int64_t n, const std::string& tag, int depth = 0)
{
    Engine& E = eng();
    if (n <= 0) {
        fill_i32(out, 1, 0);
        return;
    }
    const int64_t per = (int64_t)kScanBlock * kScanItems;
    const int64_t nblocks = (n + per - 1) / per;
    int* bsum = (int*)E.arena.get(tag + "_bs" + std::to_string(depth), sizeof(int) * (nblocks + 1));
    int* boff = (int*)E.arena.get(tag + "_bo" + std::to_string(depth), sizeof(int) * (nblocks + 1));
    k_scan_blocks<<<(unsigned)nblocks, kScanBlock, 0, E.stream>>>(n, in, out, bsum);
    CF_CHECK(hipGetLastError());
    if (nblocks > 1) {
        device_scan(bsum, boff, nblocks, tag, depth + 1);
        k_scan_add<<<(unsigned)nblocks, kScanBlock, 0, E.stream>>>(n, out, boff);
        CF_CHECK(hipGetLastError());
        // total = boff[nblocks]
        CF_CHECK(hipMemcpyAsync(out + n, boff + nblocks, sizeof(int), hipMemcpyDeviceToDevice, E.stream));
    } else {
        CF_CHECK(hipMemcpyAsync(out + n, bsum, sizeof(int), hipMemcpyDeviceToDevice, E.stream));
    }
}

struct GridSpec {
    int nx, ny, nz;
    double xmin, ymin, zmin, xmax, ymax, zmax;
    double xbs, ybs, zbs, xinv, yinv, zinv;
    double qreach;  // largest separation a pair may have along one axis (quantization range)
};

// gridlink_utils.c.src:31-49
static int get_binsize(double xdiff, double xwrap, double rmax, int refine, int max_ncells, double* binsize, int* nl)
{
    int nmesh = (int)(refine * xdiff / rmax);
    nmesh = nmesh < 1 ? 1 : nmesh;
    if (xwrap > 0 && rmax >= xwrap / 2) {
        fprintf(stderr, "Error: rmax=%f must be less than half of periodic boxize=%f to avoid double-counting particles\n",
                rmax, xwrap);
        return EXIT_FAILURE;
    }
    if (nmesh > max_ncells) nmesh = max_ncells;
    if (nmesh < 2) nmesh = 2;
    *binsize = xdiff / nmesh;
    *nl = nmesh;
    return EXIT_SUCCESS;
}

static int make_grid(GridSpec& g, const double ext[6], const double gsz[3], const double wrap[3], const int refine[3],
                     int max_cells)
{
    g.xmin = ext[0]; g.ymin = ext[1]; g.zmin = ext[2];
    g.xmax = ext[3]; g.ymax = ext[4]; g.zmax = ext[5];
    if (get_binsize(g.xmax - g.xmin, wrap[0], gsz[0], refine[0], max_cells, &g.xbs, &g.nx) ||
        get_binsize(g.ymax - g.ymin, wrap[1], gsz[1], refine[1], max_cells, &g.ybs, &g.ny) ||
        get_binsize(g.zmax - g.zmin, wrap[2], gsz[2], refine[2], max_cells, &g.zbs, &g.nz))
        return EXIT_FAILURE;
    g.xinv = g.xbs > 0 ? 1.0 / g.xbs : 0.;
    g.yinv = g.ybs > 0 ? 1.0 / g.ybs : 0.;
    g.zinv = g.zbs > 0 ? 1.0 / g.zbs : 0.;
    return EXIT_SUCCESS;
}

// Device counting sort of one particle set into the grid (gridlink_impl.c.src:65-436).
// Counting-sort gridding launched on the stream without waiting: particles
// outside the grid are counted into *oob (zeroed by the caller) and skipped.
static int gridlink_launch(const std::string& tag, const GridSpec& g, int64_t n, const void* X, const void* Y,
                           const void* Z, const void* W, size_t esize, DevLattice& L, unsigned long long* oob)
{
    Engine& E = eng();
    L.nx = g.nx; L.ny = g.ny; L.nz = g.nz;
    L.ncells = (int64_t)g.nx * g.ny * g.nz;
    L.n = n;
    L.xmin = g.xmin; L.ymin = g.ymin; L.zmin = g.zmin;
    L.xbs = g.xbs; L.ybs = g.ybs; L.zbs = g.zbs;
    L.xinv = g.xinv; L.yinv = g.yinv; L.zinv = g.zinv;
    const int64_t nb = L.ncells * kSlabs;
    if (nb >= (int64_t)INT32_MAX || n >= (int64_t)INT32_MAX) {
        fprintf(stderr, "Error: grid of %lld cells or %lld particles exceeds the engine's 32-bit indexing\n",
                (long long)L.ncells, (long long)n);
        return EXIT_FAILURE;
    }
    int* counts = (int*)E.arena.get(tag + "_cnt", sizeof(int) * (nb + 1));
    int* fill = (int*)E.arena.get(tag + "_fill", sizeof(int) * (nb + 1));
    int* bucket = (int*)E.arena.get(tag + "_bucket", sizeof(int) * (n + 1));
    L.bstart = (int*)E.arena.get(tag + "_bstart", sizeof(int) * (nb + 1));
    L.x = (double*)E.arena.get(tag + "_x", sizeof(double) * n);
    L.y = (double*)E.arena.get(tag + "_y", sizeof(double) * n);
    L.z = (double*)E.arena.get(tag + "_z", sizeof(double) * n);
    L.w = W ? (double*)E.arena.get(tag + "_w", sizeof(double) * n) : nullptr;
    L.d4 = (double4*)E.arena.get(tag + "_d4", sizeof(double4) * (n + kQPad));
    // quantum: kQMax quanta cover the largest cell side plus the pair reach, so
    // every position within reach of a cell has an unclamped quantized offset
    L.qinv = (double)kQMax / (fmax(g.xbs, fmax(g.ybs, g.zbs)) + g.qreach);
    L.q = (uint2*)E.arena.get(tag + "_q", sizeof(uint2) * (n + kQPad));
    L.bounds = (double*)E.arena.get(tag + "_bounds", sizeof(double) * 6 * L.ncells);
    fill_i32(counts, nb + 1, 0);
    fill_i32(fill, nb + 1, 0);
    GridArgs ga{g.xmin, g.ymin, g.zmin, g.xinv, g.yinv, g.zinv, g.xbs, g.ybs, g.zbs, g.nx, g.ny, g.nz, L.qinv};
    const unsigned blocks = (unsigned)((n + 255) / 256);
    if (n > 0) {
        if (esize == 4)
            k_bucket_count<float><<<blocks, 256, 0, E.stream>>>(n, (const float*)X, (const float*)Y, (const float*)Z,
                                                                 ga, bucket, counts, oob);
        else
            k_bucket_count<double><<<blocks, 256, 0, E.stream>>>(n, (const double*)X, (const double*)Y,
                                                                  (const double*)Z, ga, bucket, counts, oob);
        CF_CHECK(hipGetLastError());
    }
    device_scan(counts, L.bstart, nb, tag + "_scan");
    if (n > 0) {
        if (esize == 4)
            k_scatter<float><<<blocks, 256, 0, E.stream>>>(n, (const float*)X, (const float*)Y, (const float*)Z,
                                                            (const float*)W, ga, bucket, L.bstart, fill, L.x, L.y,
                                                            L.z, L.w, L.d4, L.q);
        else
            k_scatter<double><<<blocks, 256, 0, E.stream>>>(n, (const double*)X, (const double*)Y, (const double*)Z,
                                                             (const double*)W, ga, bucket, L.bstart, fill, L.x,
                                                             L.y, L.z, L.w, L.d4, L.q);
        CF_CHECK(hipGetLastError());
    }
    k_cell_bounds<<<(unsigned)((L.ncells + 3) / 4), 256, 0, E.stream>>>(L.ncells, L.bstart, L.x, L.y, L.z,
                                                                           L.bounds);
    CF_CHECK(hipGetLastError());
    return EXIT_SUCCESS;
}

static int check_oob(unsigned long long n_oob)
{
    if (n_oob) {
        fprintf(stderr, "Error: %llu particles are out of bounds. Check periodic wrapping?\n", n_oob);
        return EXIT_FAILURE;
    }
    return EXIT_SUCCESS;
}

static int gridlink_device(const std::string& tag, const GridSpec& g, int64_t n, const void* X, const void* Y,
                           const void* Z, const void* W, size_t esize, DevLattice& L)
{
    Engine& E = eng();
    unsigned long long* oob = (unsigned long long*)E.arena.get(tag + "_oob", 8);
    fill_u64(oob, 1);
    if (gridlink_launch(tag, g, n, X, Y, Z, W, esize, L, oob) != EXIT_SUCCESS) return EXIT_FAILURE;
    unsigned long long h_oob = 0;
    CF_CHECK(hipMemcpyAsync(&h_oob, oob, 8, hipMemcpyDeviceToHost, E.stream));
    CF_CHECK(hipStreamSynchronize(E.stream));
    return check_oob(h_oob);
}

// Device work list of neighbour-cell pairs (gridlink_impl.c.src:439-625), in
// two steps: cellpairs_flag marks the kept (cell, neighbour) candidates and
// scans them (the count lands at plan.pos[plan.total] on the device);
// cellpairs_emit writes the records in the reference's order and adds the pairs
// tested to *dtested.
struct CpPlan {
    NgbArgs a;
    int nngb = 0;
    int64_t total = 0;
    int* flags = nullptr;
    int* pos = nullptr;
    DevLattice L1, L2;
};

static int cellpairs_flag(const DevLattice& L1, const DevLattice& L2, const int refine[3], const int per[3],
                          const double wrap[3], double max3, double max2, double max1, int min_sep_opt, int autocorr,
                          CpPlan& cp)
{
    Engine& E = eng();
    NgbArgs& a = cp.a;
    a.nx = L1.nx; a.ny = L1.ny; a.nz = L1.nz;
    a.rx = refine[0]; a.ry = refine[1]; a.rz = refine[2];
    a.px = per[0]; a.py = per[1]; a.pz = per[2];
    a.autocorr = autocorr;
    a.min_sep_opt = min_sep_opt;
    a.dupcheck = (per[0] || per[1] || per[2]) &&
                 (L1.nx < 2 * refine[0] + 1 || L1.ny < 2 * refine[1] + 1 || L1.nz < 2 * refine[2] + 1);
    a.xwrap = wrap[0]; a.ywrap = wrap[1]; a.zwrap = wrap[2];
    a.max3 = max3; a.max2 = max2; a.max1 = max1;
    cp.nngb = (2 * refine[0] + 1) * (2 * refine[1] + 1) * (2 * refine[2] + 1);
    cp.total = L1.ncells * cp.nngb;
    cp.L1 = L1;
    cp.L2 = L2;
    if (cp.total >= (int64_t)INT32_MAX) {
        fprintf(stderr, "Error: %lld candidate cell pairs exceed the engine's 32-bit indexing\n", (long long)cp.total);
        return EXIT_FAILURE;
    }
    cp.flags = (int*)E.arena.get("cp_flags", sizeof(int) * (cp.total + 1));
    cp.pos = (int*)E.arena.get("cp_pos", sizeof(int) * (cp.total + 1));
    const unsigned blocks = (unsigned)((cp.total + 255) / 256);
    k_cellpair_flag<<<blocks, 256, 0, E.stream>>>(a, L1.ncells, cp.nngb, L1.bstart, L2.bstart, L1.bounds, L2.bounds,
                                                  cp.flags);
    CF_CHECK(hipGetLastError());
    device_scan(cp.flags, cp.pos, cp.total, "cp_scan");
    return EXIT_SUCCESS;
}

static void cellpairs_emit(const CpPlan& cp, CellPairRec* out, double* dtested, int cat)
{
    Engine& E = eng();
    const unsigned blocks = (unsigned)std::min<int64_t>((cp.total + 255) / 256, 1024);
    k_cellpair_emit<<<blocks, 256, 0, E.stream>>>(cp.a, cp.L1.ncells, cp.nngb, cp.L1.bstart, cp.L2.bstart,
                                                  cp.L1.bounds, cp.L2.bounds, cp.pos, out, dtested, cat);
    CF_CHECK(hipGetLastError());
}

static int cellpairs_device(const DevLattice& L1, const DevLattice& L2, const int refine[3], const int per[3],
                            const double wrap[3], double max3, double max2, double max1, int min_sep_opt,
                            int autocorr, CellPairRec** out, int* ncp_out, double* tested, int cat = 0)
{
    Engine& E = eng();
    CpPlan cp;
    if (cellpairs_flag(L1, L2, refine, per, wrap, max3, max2, max1, min_sep_opt, autocorr, cp) != EXIT_SUCCESS)
        return EXIT_FAILURE;
    int ncp = 0;
    CF_CHECK(hipMemcpyAsync(&ncp, cp.pos + cp.total, sizeof(int), hipMemcpyDeviceToHost, E.stream));
    CF_CHECK(hipStreamSynchronize(E.stream));
    CellPairRec* cps = (CellPairRec*)E.arena.get("cp_list", sizeof(CellPairRec) * (ncp + 1));
    double* dtested = (double*)E.arena.get("cp_tested", sizeof(double));
    fill_u64(dtested, 1);
    cellpairs_emit(cp, cps, dtested, cat);
    CF_CHECK(hipMemcpyAsync(tested, dtested, sizeof(double), hipMemcpyDeviceToHost, E.stream));
    CF_CHECK(hipStreamSynchronize(E.stream));
    *out = cps;
    *ncp_out = ncp;
    return EXIT_SUCCESS;
}

// ------------------------------------------------------------- driver
struct CallSpec {
    int mode;
    int64_t N1, N2;
    void *X1, *Y1, *Z1, *X2, *Y2, *Z2;
    int autocorr;
    const char* binfile;
    double pimax;       // rp_pi / wp
    double mu_max;      // s_mu
    int nmu_bins;
    double boxsize;     // xi / wp
    int numthreads;
    int cell_timer = 0; // record per-cell-pair timings (wp's c_cell_timer)
};

static_assert(sizeof(TimingRow) == sizeof(struct api_cell_timings), "timing rows must match api_cell_timings");

struct CallOut {
    std::vector<TimingRow> timings;  // per cell pair, when the cell timer is on
    std::vector<double> rupp;
    int nbin = 0, n2 = 0;
    std::vector<uint64_t> np;
    std::vector<double> avg, wavg;
    double wsum = 0, wsq = 0;  // sum of weights / of squared weights over D1
    int need_w = 0;
};

static int resolve_isa(int isa_in)
{
    if (isa_in >= 0) return isa_in;
#if defined(__x86_64__) && !defined(__HIP_DEVICE_COMPILE__)
    __builtin_cpu_init();
    if (__builtin_cpu_supports("avx512f")) return AVX512F;
    if (__builtin_cpu_supports("avx2")) return AVX2;
    if (__builtin_cpu_supports("avx")) return AVX;
    if (__builtin_cpu_supports("sse4.2")) return SSE42;
#endif
    return FALLBACK;
}

static int r2form_for(int mode, int isa_id)
{
    switch (mode) {
    case MODE_DD: return isa_id == AVX512F ? R2_FMA_DZ_DY_DX : R2_SUM_XY_Z;
    case MODE_XI: return isa_id == AVX512F ? R2_FMA_DZ_DY_DX : ((isa_id == AVX || isa_id == AVX2) ? R2_X_SUM_YZ : R2_SUM_XY_Z);
    case MODE_RPPI: return isa_id == AVX512F ? RP_FMA : RP_SUM;
    case MODE_WP: return (isa_id == AVX512F || isa_id == AVX2) ? RP_FMA : RP_SUM;
    case MODE_SMU: return isa_id == AVX512F ? R2_FMA_DX_DY_DZ : R2_SUM_XY_Z;
    }
    return R2_SUM_XY_Z;
}


// Refine heuristics, grid, boost (countpairs_impl.c.src:272-332 and the per-mode
// variants: rp_pi_impl:277-329, s_mu_impl:276-340, xi_impl:212-270, wp_impl:240-290).
// Updates o->bin_refine_factors as the reference does.
static int choose_grid(int mode, int64_t N1, double rmax, double pimax, double boxsize, const double ext[6],
                       const double wrap[3], struct config_options* o, GridSpec& g, int refine[3], int64_t N2 = 0)
{
    const int is2d = mode == MODE_RPPI || mode == MODE_WP;
    if (mode == MODE_RPPI_MOCKS || mode == MODE_SMU_MOCKS) {
        // on-sky engines: rmax is the 3-D limit; refine 1 on an axis whose extent
        // exceeds 20 rmax; boost with the larger set's density
        // (countpairs_rp_pi_mocks_impl.c.src:404-477)
        if (get_bin_refine_scheme(o) == BINNING_DFL)
            for (int d = 0; d < 3; d++)
                if (rmax < 0.05 * (ext[3 + d] - ext[d])) o->bin_refine_factors[d] = 1;
        const double gsz[3] = {rmax, rmax, rmax};
        for (int d = 0; d < 3; d++) refine[d] = o->bin_refine_factors[d];
        if (make_grid(g, ext, gsz, wrap, refine, o->max_cells_per_dim) != EXIT_SUCCESS) return EXIT_FAILURE;
        const double ncell = (double)g.nx * g.ny * g.nz;
        const double avg_np = (double)(N2 > N1 ? N2 : N1) / ncell;
        const int max_nmesh = std::max(g.nx, std::max(g.ny, g.nz));
        if ((max_nmesh <= 10 || avg_np >= 250) && max_nmesh < o->max_cells_per_dim &&
            get_bin_refine_scheme(o) == BINNING_DFL) {
            o->bin_refine_factors[0] += 1;
            o->bin_refine_factors[1] += 1;
            for (int d = 0; d < 3; d++) refine[d] = o->bin_refine_factors[d];
            if (make_grid(g, ext, gsz, wrap, refine, o->max_cells_per_dim) != EXIT_SUCCESS) return EXIT_FAILURE;
        }
        g.qreach = rmax * (1.0 + 1e-6);
        return EXIT_SUCCESS;
    }
    if (get_bin_refine_scheme(o) == BINNING_DFL) {
        if (mode == MODE_DD) {
            for (int d = 0; d < 3; d++)
                if (rmax < 0.05 * wrap[d]) o->bin_refine_factors[d] = 1;
        } else if (mode == MODE_RPPI) {
            if (rmax < 0.05 * wrap[0]) o->bin_refine_factors[0] = 1;
            if (rmax < 0.05 * wrap[1]) o->bin_refine_factors[1] = 1;
            if (pimax < 0.05 * wrap[2]) o->bin_refine_factors[2] = 1;
        } else if (mode == MODE_XI) {
            if (rmax < 0.05 * boxsize)
                for (int d = 0; d < 3; d++) o->bin_refine_factors[d] = 1;
        } else if (mode == MODE_WP) {
            if (rmax < 0.05 * boxsize) o->bin_refine_factors[0] = o->bin_refine_factors[1] = 1;
            if (pimax < 0.05 * boxsize) o->bin_refine_factors[2] = 1;
        }
    }
    double gsz[3] = {rmax, rmax, rmax};
    if (is2d || mode == MODE_SMU) gsz[2] = pimax;
    refine[0] = o->bin_refine_factors[0];
    refine[1] = o->bin_refine_factors[1];
    refine[2] = o->bin_refine_factors[2];
    if (make_grid(g, ext, gsz, wrap, refine, o->max_cells_per_dim) != EXIT_SUCCESS) return EXIT_FAILURE;
    // boost (countpairs_impl.c.src:298-332; DDsmu variant s_mu_impl:313-340)
    if (mode == MODE_SMU) {
        if (g.nx <= 10 && g.ny <= 10 && g.nz <= 10 && o->max_cells_per_dim >= 10 &&
            get_bin_refine_scheme(o) == BINNING_DFL) {
            for (int d = 0; d < 3; d++) o->bin_refine_factors[d] *= 1;
        }
    } else {
        const double avg_np = ((double)N1) / ((double)g.nx * g.ny * g.nz);
        const int max_nmesh = g.nx > g.ny ? (g.nx > g.nz ? g.nx : g.nz) : (g.ny > g.nz ? g.ny : g.nz);
        if ((max_nmesh <= 10 || avg_np >= 250) && max_nmesh < o->max_cells_per_dim &&
            get_bin_refine_scheme(o) == BINNING_DFL) {
            o->bin_refine_factors[0] += 1;
            o->bin_refine_factors[1] += 1;
        }
    }
    refine[0] = o->bin_refine_factors[0];
    refine[1] = o->bin_refine_factors[1];
    refine[2] = o->bin_refine_factors[2];
    if (make_grid(g, ext, gsz, wrap, refine, o->max_cells_per_dim) != EXIT_SUCCESS) return EXIT_FAILURE;
    g.qreach = fmax(rmax, pimax) * (1.0 + 1e-6);

    return EXIT_SUCCESS;
}

// Launch parameters of the pair kernels: rounding order, cut values, edge
// windows, slab window and the bin lookup table (the integer filter's
// thresholds depend on the secondary set's quantum: set_quant_filter).
static void build_params(int mode, int isa_id, int nbin, int n2, int nslots, const std::vector<double>& rupp_sqr,
                         int sorted, double rmax, double pimax, double inv_d2, double sqr_mu_max, int need_avg,
                         int need_w, int autocorr, const double wrap[3], double coord_scale, double cs, int rmaxr,
                         PairParams& P, std::vector<unsigned short>& lut)
{
    memset(&P, 0, sizeof P);
    const int is2d = mode == MODE_RPPI || mode == MODE_WP;
    P.mode = mode;
    P.r2form = r2form_for(mode, isa_id);
    P.float_index = (isa_id == AVX512F) && (mode == MODE_RPPI || mode == MODE_SMU);
    P.nbin = nbin;
    P.n2 = n2;
    P.nslots = nslots;
    P.need_avg = need_avg;
    P.need_w = need_w;
    P.autocorr = autocorr;
    P.sorted_edges = sorted;
    P.sqr_rmax = rupp_sqr[nbin - 1];
    P.sqr_rmin = rupp_sqr[0];
    P.pimax = pimax;
    P.inv_d2 = inv_d2;
    P.sqr_mu_max = sqr_mu_max;
    if (mode == MODE_SMU) {
        const double a = sqrt(P.sqr_rmax - 0.0 - 0.0);
        P.max_all_dz_same = pimax < a ? pimax : a;
    } else {
        P.max_all_dz_same = sqrt(P.sqr_rmax - 0.0 - 0.0);
    }
    P.xwrap = wrap[0];
    P.ywrap = wrap[1];
    P.zwrap = wrap[2];
    // Edge windows: the reference's pruning tests can only reject pairs whose
    // separation lies within a few ulps of the coordinate scale below rmax / pimax.
    P.near_r2 = 1e-9 * P.sqr_rmax + 1e-9 * rmax * coord_scale;
    P.near_dz = 1e-9 * pimax + 1e-9 * coord_scale;
    P.zreach = (is2d || mode == MODE_SMU) ? pimax : sqrt(P.sqr_rmax);
    P.slab_margin = 1e-7 * (coord_scale + cs);

    // ---- bin lookup table: for each 1/16-octave bucket of (float)r^2, the number
    // of edges strictly below the bucket (a lower bound of the bin index)
    lut.clear();
    {
        auto bucket = [](double v) {
            const float f = (float)v;
            uint32_t b;
            memcpy(&b, &f, 4);
            return (int)(b >> 19);
        };
        double minpos = INFINITY;
        for (int m = 0; m + 1 < nbin; m++)
            if (rupp_sqr[m] > 0 && rupp_sqr[m] < minpos) minpos = rupp_sqr[m];
        if (!std::isfinite(minpos)) minpos = P.sqr_rmax > 0 ? P.sqr_rmax : 1.0;
        int base = bucket(minpos) - 1;
        base = base < 0 ? 0 : base;
        const int top = bucket(P.sqr_rmax) + 1;
        const int n = top - base + 1;
        lut.assign(n > 0 ? n : 1, 0);
        for (int idx = 0; idx < n; idx++) {
            const uint32_t bits = (uint32_t)(base + idx) << 19;
            float lo_f;
            memcpy(&lo_f, &bits, 4);
            const double thr = (double)lo_f * (1.0 - 0x1p-20);
            int cnt = 0;
            for (int m = 0; m + 1 < nbin; m++)
                if (rupp_sqr[m] < thr) cnt++;
            lut[idx] = (unsigned short)cnt;
        }
        P.lut_base = base;
        P.lut_n = (int)lut.size();
        if (nbin > 65535) P.sorted_edges = 0;
        // The kernels step up from lut[idx] while e[k] <= r2. A pair in bucket idx
        // has thr(idx) <= r2 < lo(idx + 1) * (1 + 2^-24) ((float)r2 rounds by at
        // most 2^-24), so at most one step is needed when no such interval holds
        // two edges e[0..nbin-2].
        P.lut_single = 1;
        for (int idx = 0; idx < n && P.lut_single; idx++) {
            auto lo_of = [&](int b) {
                const uint32_t bits = (uint32_t)(base + b) << 19;
                float f;
                memcpy(&f, &bits, 4);
                return (double)f;
            };
            const double thr = lo_of(idx) * (1.0 - 0x1p-20);
            const double top = idx + 1 < n ? lo_of(idx + 1) * (1.0 + 0x1p-23) : INFINITY;
            int cnt = 0;
            for (int m = 0; m + 1 < nbin; m++)
                if (rupp_sqr[m] >= thr && rupp_sqr[m] < top) cnt++;
            if (cnt > 1) P.lut_single = 0;
        }
    }
    {
        const char* dm = getenv("CORRFUNC_AMD_DENSE");
        P.dense_mode = dm ? atoi(dm) : 0;
        const char* ds = getenv("CORRFUNC_AMD_STRIDE");
        // dense cross-cell items take contiguous primaries by default (measured
        // against strided ones: on-sky DD(rp, pi) -8.5 %, DDrppi -1.8 %, xi on
        // NFW halos -0.6 %, DD and DDsmu equal)
        P.dense_stride = ds ? atoi(ds) : 0;
        const char* df = getenv("CORRFUNC_AMD_DENSE_FRAC");
        // measured (bench.py --sweep CORRFUNC_AMD_DENSE_FRAC): DDsmu cross
        // (workload 5) 2.71 s at 0.35, 2.60 s at 0.2-0.25, 3.18 s at 0; DDrppi
        // and DD within 2 % from 0.2 to 0.5
        P.dense_frac = df ? (float)atof(df) : 0.25f;
        // sub-window length of the sparse scan (tests shorten it to exercise the
        // window changes; a multiple of 4 below 2^26)
        const char* mw = getenv("CORRFUNC_AMD_MAX_WINDOW");
        P.max_window = mw ? std::max(4, std::min(1 << 25, atoi(mw))) & ~3 : 1 << 25;
        const char* qb = getenv("CORRFUNC_AMD_QBLOCK_LOG2");
        // blocks of 32 items per sub-queue (measured, bench.py --sweep
        // CORRFUNC_AMD_QBLOCK_LOG2=0..8): DDsmu cross -1.8 %, DDrppi -1.5 %, DD and xi
        // within 1 % (the kernels are bound by instruction issue, not L2 misses)
        P.qblock_log2 = qb ? std::max(0, std::min(12, atoi(qb))) : 5;
        const char* fe = getenv("CORRFUNC_AMD_FETCH");
        P.fetch = fe ? std::max(1, atoi(fe)) : 0;  // 0: chosen at launch from the work list
        if (eng().force_fetch > 0) P.fetch = eng().force_fetch;
    }

}

// Device copy of the (primary, secondary) lattice pair of a single call.
// Thresholds of the integer pair filter of a secondary set (quantum 1/L.qinv).
// Rounding both positions of a pair to the nearest quantum moves their
// difference by at most one quantum per axis, so in quanta a pair at distance d
// has |dq| <= d*qinv + sqrt(3) (3-D) and |dqz| <= |dz|*qinv + 1; the bounds
// below add slack to these, so every pair the FP64 test can accept passes.
static void set_quant_filter(DevLattice& L, const PairParams& P)
{
    const bool is2d = P.mode == MODE_RPPI || P.mode == MODE_WP;
    const double rhi = sqrt(P.sqr_rmax) * (1.0 + 1e-9) * L.qinv + 2.0;
    const double rlo = sqrt(P.sqr_rmin) * (1.0 - 1e-9) * L.qinv - 2.0;
    const double hi = floor(rhi * rhi) + 1.0;
    const double lo = rlo > 0 ? floor(rlo * rlo) : 0.0;
    L.qrange = (unsigned)(hi - lo);
    L.qlo_neg = -(int)lo;
    L.qz = is2d ? (int)ceil(P.pimax * (1.0 + 1e-9) * L.qinv + 1.0) + 1 : 0;
}

static const DevLattice* lats_for(const DevLattice& L1, const DevLattice& L2)
{
    Engine& E = eng();
    E.hlats[0] = L1;  // the engine keeps the host copy alive until the copy has run
    E.hlats[1] = L2;
    DevLattice* lats = (DevLattice*)E.arena.get("lats", sizeof(E.hlats));
    CF_CHECK(hipMemcpyAsync(lats, E.hlats, sizeof(E.hlats), hipMemcpyHostToDevice, E.stream));
    return lats;
}

// ---------------------------------------------------------- tuned launch
// the queue hands out 4 items per atomic from 100 items per resident wave, when
// there are at most 8 items per cell pair (few split cell pairs)
constexpr int kBatchItemsPerWave = 100;
constexpr double kBatchItemsPerCellPair = 8.0;
constexpr int kSplitPerWave = 4;  // a work item is split when it exceeds 1/(4 * resident waves) of the total cost

// launch_tuned's answer when even a one-copy LDS layout (edges, lookup,
// queues) does not fit a workgroup: the call is counted by the plain kernel
constexpr int kNoTunedFit = 2;

struct TunedRun {
    int nitems = 0;
    int fetch = 0;                        // work items per queue atomic
    int global_hist = 0;                  // 1: histogram in global memory (too large for LDS)
    int drain_pairs = 0;                  // survivors per lane per drain of the tuned kernel
    int threads = 0;                      // workgroup size of the tuned kernel
    unsigned long long* ticks = nullptr;  // per cell pair wall-clock ticks (device), when timed
    int* cu = nullptr;                    // compute unit of the last item of each cell pair (device)
};

// Samples store of the cost model, allocated on first use.
static void ensure_sample_store()
{
    Engine& E = eng();
    CostModel& M = E.model;
    if (M.d_sx) return;
    CF_CHECK(hipMalloc(&M.d_sx, sizeof(float) * (size_t)M.cap * kNumFeatures));
    CF_CHECK(hipMalloc(&M.d_sy, sizeof(float) * (size_t)M.cap));
    CF_CHECK(hipMalloc(&M.d_count, sizeof(int)));
    k_fill_i32<<<1, 64, 0, E.stream>>>(1, M.d_count, 0);
    CF_CHECK(hipGetLastError());
}

// The tuned pair kernel for P's mode, rounding order and `batch`, and its
// launch shape (see the comment in the body); sets P.hist_copies and
// P.prim_cache. Returns kNoTunedFit when even a one-copy layout does not fit a
// workgroup's LDS.
struct TunedShape {
    const void* kfn = nullptr;
    bool gh = false;
    int threads = 0, nwaves = 0, per_cu = 0, dp = 1;
    size_t lds = 0;
};

// The tuned kernel of mode M, rounding order F and batch form B for a launch:
// the global-histogram form (gh; runtime flags, not batched) or the
// LDS-histogram form of flag mask fl, with dp survivors per lane per drain.
template <int M, int F, bool B>
static const void* tuned_kernel(bool gh, int dp, int fl)
{
    if (gh) {
        if constexpr (B) {
            return nullptr;
        } else {
            return dp == 2 ? (const void*)&k_pairs_tuned<M, F, false, true, 2, kFlRuntime>
                           : (const void*)&k_pairs_tuned<M, F, false, true, 1, kFlRuntime>;
        }
    }
#define CF_FL(X)                                                                                               \
    case X:                                                                                                    \
        return dp == 2 ? (const void*)&k_pairs_tuned<M, F, B, false, 2, X>                                     \
                       : (const void*)&k_pairs_tuned<M, F, B, false, 1, X>;
    switch (fl) {
        CF_FL(0) CF_FL(1) CF_FL(2) CF_FL(3) CF_FL(4) CF_FL(5) CF_FL(6) CF_FL(7)
    }
#undef CF_FL
    return nullptr;
}

static int tuned_shape(int batch, PairParams& P, int nslots, int nbin, TunedShape& out)
{
    Engine& E = eng();
    const size_t per_slot = 8 + (P.need_avg ? 8 : 0) + (P.need_w ? 8 : 0);
    // LDS of everything but the histogram for a workgroup of `nw` waves:
    // squared edges, bin lookup, survivor queues
    auto lds_fixed_for = [&](int nw) {
        return (size_t)nbin * 8 + (size_t)((P.lut_n + 1) & ~1) * 2 + (size_t)nw * kQueueAlloc * 4;
    };
    // a histogram whose single copy does not fit next to the rest of a 4-wave
    // workgroup is kept in global memory (no replication; FP64 atomics)
    const bool gh = !batch && (size_t)nslots * per_slot + lds_fixed_for(4) > E.lds_max;
    if (lds_fixed_for(4) > E.lds_max || (batch && (size_t)nslots * per_slot * 4 + lds_fixed_for(4) > E.lds_max))
        return kNoTunedFit;
    // the kernel of this mode and rounding order with 1 or 2 survivors per lane per drain
    // the launch's flags as the LDS-histogram kernels take them (compile-time
    // sums and bin lookup; the AVX512F index rule follows the rounding order)
    const int fl = (P.need_avg ? kFlAvg : 0) | (P.need_w ? kFlW : 0) | (P.sorted_edges && P.lut_single ? kFlLut1 : 0);
    const bool findex = (P.mode == MODE_RPPI && P.r2form == RP_FMA) || (P.mode == MODE_SMU && P.r2form == R2_FMA_DX_DY_DZ);
    if ((P.float_index != 0) != findex) {
        fprintf(stderr, "Error: index rule %d does not match rounding order %d\n", P.float_index, P.r2form);
        return EXIT_FAILURE;
    }
    auto kernel_for = [&](int dp) -> const void* {
        const void* k = nullptr;
#define CF_K(M, F, B)                                                                                          \
    if (P.mode == M && P.r2form == F && batch == B) k = tuned_kernel<M, F, B>(gh, dp, fl);
        CF_K(MODE_DD, R2_SUM_XY_Z, 0) CF_K(MODE_DD, R2_FMA_DZ_DY_DX, 0)
        CF_K(MODE_XI, R2_SUM_XY_Z, 0) CF_K(MODE_XI, R2_X_SUM_YZ, 0) CF_K(MODE_XI, R2_FMA_DZ_DY_DX, 0)
        CF_K(MODE_RPPI, RP_SUM, 0) CF_K(MODE_RPPI, RP_FMA, 0) CF_K(MODE_WP, RP_SUM, 0) CF_K(MODE_WP, RP_FMA, 0)
        CF_K(MODE_SMU, R2_SUM_XY_Z, 0) CF_K(MODE_SMU, R2_FMA_DX_DY_DZ, 0)
        CF_K(MODE_RPPI_MOCKS, R2_SUM_XY_Z, 0) CF_K(MODE_SMU_MOCKS, R2_SUM_XY_Z, 0)
        CF_K(MODE_DD, R2_SUM_XY_Z, 1) CF_K(MODE_DD, R2_FMA_DZ_DY_DX, 1)
#undef CF_K
        return k;
    };
    if (!kernel_for(1)) {
        fprintf(stderr, "Error: no pair kernel for mode %d / rounding order %d%s\n", P.mode, P.r2form,
                batch ? " (batched)" : "");
        return EXIT_FAILURE;
    }
    // Launch shape for a workgroup of `thr` threads: LDS histogram replication
    // R (power of 2, <= 32, histograms <= the LDS limit less 24 KiB) = the
    // largest that keeps min(16, best achievable) waves per CU resident (on-sky
    // modes: the best achievable, since their survivors' FP64 decision needs
    // the latency hiding more than the histogram needs copies; measured on
    // DDrppi_mocks, R 4 at 16 waves is 15 % slower than R 2 at 24, while
    // DDrppi gains 4 %); the primaries cache where it costs no resident
    // workgroup; `per_cu` = resident workgroups per CU (registers and LDS),
    // which is also how many persistent workgroups per CU are launched.
    struct Shape {
        int threads, nwaves, R, prim, per_cu;
        size_t lds;
    };
    auto shape_for = [&](const void* kfn, int thr) {
        Shape sh;
        sh.threads = thr;
        sh.nwaves = thr / 64;
        const int regions = gh ? 0 : (batch ? sh.nwaves : 1);
        const size_t lds_fixed = lds_fixed_for(sh.nwaves);
        const size_t lds_prim = 16 + (size_t)sh.nwaves * 64 * sizeof(double4);
        auto lds_for = [&](int r) { return (size_t)regions * nslots * per_slot * r + lds_fixed; };
        auto resident = [&](size_t l) {
            int nb = 0;
            if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kfn, thr, l) != hipSuccess) {
                (void)hipGetLastError();
                nb = (int)((160 * 1024) / (l + 1024));
            }
            return nb < 1 ? 1 : (nb > 8 ? 8 : nb);
        };
        const bool onsky = P.mode == MODE_RPPI_MOCKS || P.mode == MODE_SMU_MOCKS;
        const int best = resident(lds_for(1)) * sh.nwaves;
        const int target = onsky ? best : std::min(16, best);
        int R = 1;
        for (int r = 2; r <= 32 && !gh; r *= 2) {
            if ((size_t)regions * nslots * per_slot * r > E.lds_max - 24 * 1024) break;
            if (lds_for(r) > E.lds_max) break;  // (the workgroup's queues included)
            if (resident(lds_for(r)) * sh.nwaves < target) break;
            R = r;
        }
        if (const char* rc = getenv("CORRFUNC_AMD_HIST_COPIES")) {
            // a power of two that fits the workgroup's LDS (others are ignored)
            const int v = atoi(rc);
            if (!gh && v >= 1 && v <= 64 && (v & (v - 1)) == 0 && lds_for(v) <= E.lds_max) R = v;
        }
        if (E.force_R > 0 && E.force_R < R) R = E.force_R;
        sh.R = R;
        const size_t lds_base = lds_for(R);
        sh.prim = lds_base + lds_prim <= E.lds_max && resident(lds_base + lds_prim) >= resident(lds_base) ? 1 : 0;
        if (const char* pc = getenv("CORRFUNC_AMD_PRIM_CACHE")) sh.prim = atoi(pc) != 0;
        sh.lds = lds_base + (sh.prim ? lds_prim : 0);
        sh.per_cu = resident(sh.lds);
        return sh;
    };
    // workgroups of 8 waves (a histogram shared by more waves; measured as fast
    // or faster), or of 4 where 8 would keep fewer waves resident or exceed LDS;
    // of 16 (one per CU) where that keeps as many waves resident with more
    // histogram copies: the same LDS then holds one histogram with twice the
    // copies instead of two (measured: DDsmu cross 2.52 -> 2.02 s, DDrppi 18.4 ->
    // 17.4 ms; clustered lanes that land in one slot spread over more words)
    auto best_shape = [&](const void* kfn) {
        Shape sh = shape_for(kfn, 256);
        const Shape s8 = shape_for(kfn, 512);
        if (s8.lds <= E.lds_max && s8.per_cu * s8.nwaves >= sh.per_cu * sh.nwaves) sh = s8;
        const Shape s16 = shape_for(kfn, 1024);
        if (s16.lds <= E.lds_max && s16.per_cu * s16.nwaves >= sh.per_cu * sh.nwaves && s16.R > sh.R) sh = s16;
        if (const char* wg = getenv("CORRFUNC_AMD_THREADS")) {
            const int t = atoi(wg);
            sh = t == 1024 && s16.lds <= E.lds_max ? s16 : (t == 512 && s8.lds <= E.lds_max ? s8 : shape_for(kfn, 256));
        }
        return sh;
    };
    // two survivors per lane per drain where that kernel keeps as many waves
    // resident and as many histogram copies (launches whose LDS already limits
    // the waves per SIMD to five or fewer; measured: DDsmu 2.5 % and DDrppi
    // 4.5 % faster there, DD and xi 5-6 % slower where six waves fit)
    int dp = 1;
    const void* kfn = kernel_for(1);
    Shape sh = best_shape(kfn);
    {
        const void* k2 = kernel_for(2);
        const Shape s2 = best_shape(k2);
        bool two = s2.per_cu * s2.nwaves >= sh.per_cu * sh.nwaves && s2.R >= sh.R;
        if (const char* d = getenv("CORRFUNC_AMD_DRAIN_PAIRS")) two = atoi(d) == 2;
        if (two) {
            dp = 2;
            kfn = k2;
            sh = s2;
        }
    }
    P.hist_copies = sh.R;
    P.prim_cache = sh.prim;
    if (sh.lds > E.lds_max) {
        fprintf(stderr, "Error: LDS requirement %zu B of the pair kernel exceeds the %zu B a workgroup can have\n",
                sh.lds, E.lds_max);
        return EXIT_FAILURE;
    }
    if (sh.lds > 64 * 1024)
        CF_CHECK(hipFuncSetAttribute(kfn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sh.lds));
    out.kfn = kfn;
    out.gh = gh;
    out.dp = dp;
    out.threads = sh.threads;
    out.nwaves = sh.nwaves;
    out.per_cu = sh.per_cu;
    out.lds = sh.lds;
    return EXIT_SUCCESS;
}


// Builds the work list of `ncp` cell pairs (ordered and split by the selected
// scheduler) and launches the tuned pair kernel on it. `batch`: cell pair k
// belongs to catalog cps[k].pad, whose lattice is lats[pad] and whose histogram
// is g_*[pad * nslots]. Records E.ev_list before the work list, E.ev0 / E.ev1
// around the pair kernel. With `timed`, per-cell-pair ticks are recorded (and
// appended to the cost model's samples when recording is on).
static int launch_tuned(int ncp, const CellPairRec* cps, const DevLattice* lats, int batch, PairParams& P, int nslots,
                        int nbin, const double* d_edges, const unsigned short* d_lut, unsigned long long* g_np,
                        double* g_avg, double* g_w, bool timed, TunedRun& T)
{
    Engine& E = eng();
    TunedShape ts;
    const int src = tuned_shape(batch, P, nslots, nbin, ts);
    if (src != EXIT_SUCCESS) return src;
    const void* kfn = ts.kfn;
    const bool gh = ts.gh;
    const int threads = ts.threads, nwaves = ts.nwaves;
    const size_t lds = ts.lds;
    const int per_cu = ts.per_cu;
    const long max_blocks = (long)E.ncu * per_cu;

    // ---- cost prediction and order
    CF_CHECK(hipEventRecord(E.ev_list, E.stream));
    const int nb = (ncp + 255) / 256;
    const int sched = E.sched_kind;
    const bool learned = sched == CF_AMD_SCHED_LEARNED && E.model.trained;
    const bool record = timed && E.model.recording;
    float* feats = nullptr;
    if (learned || record) {
        feats = (float*)E.arena.get("s_feats", sizeof(float) * kNumFeatures * (size_t)ncp);
        k_cp_features<<<nb, 256, 0, E.stream>>>(ncp, cps, lats, batch, P, feats);
        CF_CHECK(hipGetLastError());
    }
    float* cost = nullptr;
    int *order = nullptr, *nparts = nullptr, *nheavy = nullptr;
    double* total = nullptr;
    if (sched != CF_AMD_SCHED_NONE) {
        cost = (float*)E.arena.get("s_cost", sizeof(float) * ncp);
        if (learned)
            k_mlp_predict<<<nb, 256, 0, E.stream>>>(ncp, feats, E.model.d_norm, E.model.d_params, cost);
        else
            k_cost_heuristic<<<nb, 256, 0, E.stream>>>(ncp, cps, lats, batch, cost);
        CF_CHECK(hipGetLastError());
        total = (double*)E.arena.get("s_total", sizeof(double));
        k_fill_u64<<<1, 64, 0, E.stream>>>(1, (unsigned long long*)total, 0ull);  // +0.0
        k_cost_total<<<std::min(nb, 1024), 256, 0, E.stream>>>(ncp, cost, total);
        const int nsb = (ncp + kSortBlock - 1) / kSortBlock;
        const int64_t nmat = (int64_t)kSortBuckets * nsb;
        int* bcount = (int*)E.arena.get("s_bcount", sizeof(int) * (size_t)(nmat + 1));
        int* boff = (int*)E.arena.get("s_boff", sizeof(int) * (size_t)(nmat + 1));
        const char* hv = getenv("CORRFUNC_AMD_HEAVY");
        const float heavy = hv ? (float)atof(hv) : kHeavyFactor;
        k_cost_hist<<<nsb, 256, 0, E.stream>>>(ncp, cost, total, heavy, nsb, bcount);
        CF_CHECK(hipGetLastError());
        device_scan(bcount, boff, nmat, "s_scan");
        order = (int*)E.arena.get("s_order", sizeof(int) * ncp);
        k_cost_place<<<nsb, 256, 0, E.stream>>>(ncp, cost, total, heavy, nsb, boff, order);
        CF_CHECK(hipGetLastError());
        nheavy = boff + (int64_t)(kSortBuckets - 1) * nsb;  // cell pairs ahead of the light bucket
        nparts = (int*)E.arena.get("s_nparts", sizeof(int) * ncp);
    }

    // ---- work items
    int* nch = (int*)E.arena.get("wi_n", sizeof(int) * (ncp + 1));
    int* off = (int*)E.arena.get("wi_off", sizeof(int) * (ncp + 1));
    const char* sp = getenv("CORRFUNC_AMD_SPLIT");
    const double split = sp ? atof(sp) : (double)kSplitPerWave;
    // split <= 0: no splitting (cap = total)
    const double cap_div = split > 0 ? (double)max_blocks * nwaves * split : 1.0;
    k_count_chunks<<<nb, 256, 0, E.stream>>>(ncp, order, cps, lats, batch, cost, total, cap_div, nparts, nch);
    CF_CHECK(hipGetLastError());
    device_scan(nch, off, ncp, "wi_scan");
    int nitems = 0;
    CF_CHECK(hipMemcpyAsync(&nitems, off + ncp, sizeof(int), hipMemcpyDeviceToHost, E.stream));
    CF_CHECK(hipStreamSynchronize(E.stream));
    WorkItem* items = (WorkItem*)E.arena.get("wi_items", sizeof(WorkItem) * ((size_t)nitems + 1));
    k_emit_chunks<<<nb, 256, 0, E.stream>>>(ncp, order, cps, lats, batch, P, off, nparts, items);
    CF_CHECK(hipGetLastError());
    int* counter = (int*)E.arena.get("wi_counter", sizeof(int) * kQueues * kCounterStride);
    k_fill_i32<<<1, 256, 0, E.stream>>>(kQueues * kCounterStride, counter, 0);
    const int* d_nitems = off + ncp;
    // items taken one at a time: the heavy cell pairs' (all of them without an order)
    int* single = (int*)E.arena.get("wi_single", sizeof(int));
    if (nheavy)
        k_single_boundary<<<1, 64, 0, E.stream>>>(nheavy, off, single);
    else
        k_fill_i32<<<1, 64, 0, E.stream>>>(1, single, INT32_MAX);
    CF_CHECK(hipGetLastError());
    long blocks = max_blocks;
    const long needed = ((long)nitems + nwaves - 1) / nwaves;
    if (blocks > needed) blocks = needed > 0 ? needed : 1;
    // Queue batch: with many small items the serialised sub-queue atomics
    // (~14 ns each) become a large share of the launch, and 4 items per atomic
    // is faster (uniform DD: 138 items per wave, 1.3 per cell pair, -11 %; xi
    // 1.8k and 4.6, -3 %; batched DD -6 %). Where heavy cell pairs were split
    // into many large items (clustered DDrppi 14.6 items per cell pair, DDsmu
    // cross 18) or items are few per wave, batching is slower (+2 to +9 %).
    if (P.fetch <= 0)
        P.fetch = (double)nitems >= (double)kBatchItemsPerWave * blocks * nwaves &&
                          (double)nitems <= kBatchItemsPerCellPair * (double)ncp
                      ? 4
                      : 1;

    unsigned long long* ticks = nullptr;
    int* cu = nullptr;
    if (timed) {
        ticks = (unsigned long long*)E.arena.get("t_ticks", sizeof(unsigned long long) * ncp);
        cu = (int*)E.arena.get("t_cu", sizeof(int) * ncp);
        k_fill_u64<<<nb, 256, 0, E.stream>>>(ncp, ticks, 0ull);
        k_fill_i32<<<nb, 256, 0, E.stream>>>(ncp, cu, -1);
    }

    void* kargs[] = {&P, &lats, &cps, &items, &d_nitems, &counter, &d_edges, &d_lut, &g_np, &g_avg, &g_w, &ticks, &cu, &single};
    CF_CHECK(hipEventRecord(E.ev0, E.stream));
    CF_CHECK(hipLaunchKernel(kfn, dim3((unsigned)blocks), dim3(threads), kargs, lds, E.stream));
    CF_CHECK(hipGetLastError());
    CF_CHECK(hipEventRecord(E.ev1, E.stream));
    if (record) {
        ensure_sample_store();
        k_append_samples<<<nb, 256, 0, E.stream>>>(ncp, feats, ticks, E.model.cap, E.model.d_count, E.model.d_sx,
                                                   E.model.d_sy);
        CF_CHECK(hipGetLastError());
    }
    T.nitems = nitems;
    T.fetch = P.fetch;
    T.global_hist = gh ? 1 : 0;
    T.drain_pairs = ts.dp;
    T.threads = ts.threads;
    T.ticks = ticks;
    T.cu = cu;
    return EXIT_SUCCESS;
}

// ---------------------------------------------------------------- autotuner
// Signature of a workload for the autotuner's cache: mode, the octave of each
// set's size, rmax and pimax relative to the box (1/16 octave), the box's
// periodic axes, weights, averages and input precision.
static std::string tune_key(int mode, int64_t N1, int64_t N2, int autocorr, double rmax, double pimax,
                            const double ext[6], const double wrap[3], int need_w, int need_avg, size_t esize,
                            int nbin)
{
    double span = 0;
    for (int d = 0; d < 3; d++) span = fmax(span, wrap[d] > 0 ? wrap[d] : ext[3 + d] - ext[d]);
    auto q = [&](double v) { return v > 0 && span > 0 ? (int)lround(16.0 * log2(v / span)) : -999; };
    char buf[256];
    snprintf(buf, sizeof buf, "m%d n%d n%d a%d r%d p%d w%d%d%d W%d A%d e%d b%d", mode,
             N1 > 0 ? (int)floor(log2((double)N1)) : -1, N2 > 0 ? (int)floor(log2((double)N2)) : -1, autocorr,
             q(rmax), q(pimax), wrap[0] > 0, wrap[1] > 0, wrap[2] > 0, need_w, need_avg, (int)esize, nbin);
    return buf;
}

static void tune_cache_load()
{
    Engine& E = eng();
    const char* path = getenv("CORRFUNC_AMD_TUNE_CACHE");
    if (!path) return;
    FILE* fp = fopen(path, "r");
    if (!fp) return;
    char line[512];
    while (fgets(line, sizeof line, fp)) {
        char* bar = strchr(line, '|');
        if (!bar) continue;
        *bar = '\0';
        std::array<int, 5> v;
        if (sscanf(bar + 1, "%d %d %d %d %d", &v[0], &v[1], &v[2], &v[3], &v[4]) == 5) E.tune_cache[line] = v;
    }
    fclose(fp);
}

static void tune_cache_store(const std::string& key, const std::array<int, 5>& v)
{
    const char* path = getenv("CORRFUNC_AMD_TUNE_CACHE");
    if (!path) return;
    FILE* fp = fopen(path, "a");
    if (!fp) return;
    fprintf(fp, "%s|%d %d %d %d %d\n", key.c_str(), v[0], v[1], v[2], v[3], v[4]);
    fclose(fp);
}

static int run_core(const CallSpec& c, struct config_options* o, struct extra_options* extra, CallOut& R);

// Candidate bin refine factors (x, y, z) and histogram replications the
// autotuner times; the fastest complete call wins.
static const int kTuneRefine[][3] = {{1, 1, 1}, {2, 2, 1}, {2, 2, 2}, {3, 3, 1}, {3, 3, 2}, {4, 4, 2}};
static const int kTuneR[] = {0, 8, 2};
static const int kTuneFetch[] = {1, 4};

// Looks up (or measures) the refine factors and histogram replication of a
// workload. Measuring runs the call once per candidate.
static int autotune_choice(const CallSpec& c, const struct config_options* o, struct extra_options* extra,
                           const std::string& key, std::array<int, 5>& out)
{
    Engine& E = eng();
    if (!E.tune_loaded) {
        tune_cache_load();
        E.tune_loaded = true;
    }
    auto it = E.tune_cache.find(key);
    if (it != E.tune_cache.end()) {
        out = it->second;
        E.last_autotuned = std::max(E.last_autotuned, 1);
        return EXIT_SUCCESS;
    }
    // one handler scope over all candidate runs; the forced settings are
    // cleared however the measuring ends (an error or an interrupt throws)
    InterruptScope interrupts;
    struct TuneReset {
        Engine& E;
        ~TuneReset() { E.force_fetch = 0; E.force_R = 0; E.tuning = false; }
    } reset{E};
    E.tuning = true;
    double best = 1e300;
    std::array<int, 5> pick = {2, 2, 1, 0, 1};
    for (const auto& f : kTuneRefine) {
        struct config_options oc = *o;
        set_custom_bin_refine_factors(&oc, f);
        E.force_R = 0;
        E.force_fetch = 1;
        CallOut tmp;
        const double t0 = now_s();
        const int st = run_core(c, &oc, extra, tmp);
        const double dt = now_s() - t0;
        if (st == EXIT_SUCCESS && dt < best) {
            best = dt;
            pick = {f[0], f[1], f[2], 0, 1};
        }
    }
    if (E.engine_kind != CF_AMD_ENGINE_PLAIN) {
        for (int r : kTuneR) {
            if (r == 0) continue;
            struct config_options oc = *o;
            set_custom_bin_refine_factors(&oc, pick.data());
            E.force_R = r;
            CallOut tmp;
            const double t0 = now_s();
            const int st = run_core(c, &oc, extra, tmp);
            const double dt = now_s() - t0;
            if (st == EXIT_SUCCESS && dt < best) {
                best = dt;
                pick[3] = r;
            }
        }
        for (int fch : kTuneFetch) {
            if (fch == 1) continue;
            struct config_options oc = *o;
            set_custom_bin_refine_factors(&oc, pick.data());
            E.force_R = pick[3];
            E.force_fetch = fch;
            CallOut tmp;
            const double t0 = now_s();
            const int st = run_core(c, &oc, extra, tmp);
            const double dt = now_s() - t0;
            if (st == EXIT_SUCCESS && dt < best) {
                best = dt;
                pick[4] = fch;
            }
        }
    }
    E.tune_cache[key] = pick;
    tune_cache_store(key, pick);
    out = pick;
    E.last_autotuned = 2;
    return EXIT_SUCCESS;
}

static int run_core(const CallSpec& c, struct config_options* o, struct extra_options* extra, CallOut& R)
{
    Engine& E = eng();
    InterruptScope interrupts;
    cf_amd_stats st;
    memset(&st, 0, sizeof st);
    const double t_start = now_s();
    const int mode = c.mode;
    const int is2d = mode == MODE_RPPI || mode == MODE_WP;
    const bool xiwp = mode == MODE_XI || mode == MODE_WP;
    const bool onsky = mode == MODE_RPPI_MOCKS || mode == MODE_SMU_MOCKS;
    const size_t esize = o->float_type;
    const int need_w = extra->weight_method != NONE;
    R.need_w = need_w;
    if (need_w && extra->weight_method != PAIR_PRODUCT) {
        fprintf(stderr, "Error: unknown weight method %d\n", (int)extra->weight_method);
        return EXIT_FAILURE;
    }
    if (need_w && (extra->weights0.weights[0] == nullptr || (!c.autocorr && extra->weights1.weights[0] == nullptr))) {
        fprintf(stderr, "Error: weight_method requires weights arrays\n");
        return EXIT_FAILURE;
    }
    // ---- bins
    double rmin = 0, rmax = 0;
    int nbin = 0;
    if (setup_bins(c.binfile, &rmin, &rmax, &nbin, R.rupp) != EXIT_SUCCESS) return EXIT_FAILURE;
    if (!((onsky ? rmin > 0.0 : rmin >= 0.0) && rmax > 0.0 && rmin < rmax && nbin > 0)) {
        fprintf(stderr, "Error: Could not setup with R bins correctly. (rmin = %lf, rmax = %lf, with nbins = %d). "
                        "Expected non-zero rmin/rmax with rmax > rmin and nbins >=1 \n", rmin, rmax, nbin);
        return EXIT_FAILURE;
    }
    R.nbin = nbin;
    if (mode == MODE_SMU || mode == MODE_SMU_MOCKS) {
        if (c.mu_max <= 0.0 || c.mu_max > 1.0) {
            fprintf(stderr, "Error: max_mu = %lf. Please change it to be > 0 and <= 1.0\n", c.mu_max);
            return EXIT_FAILURE;
        }
        if (c.nmu_bins < 1) {
            fprintf(stderr, "Error: Number of mu bins = %d must be at least 1\n", c.nmu_bins);
            return EXIT_FAILURE;
        }
    }
    int n2 = 0;
    double pimax = 0, inv_d2 = 0, sqr_mu_max = 0;
    if (mode == MODE_RPPI) {
        n2 = (int)c.pimax;
        pimax = c.pimax;
        const double dpi = pimax / n2;
        inv_d2 = 1.0 / dpi;
    } else if (mode == MODE_WP) {
        pimax = c.pimax;
    } else if (mode == MODE_SMU) {
        n2 = c.nmu_bins;
        pimax = rmax * c.mu_max;
        const double dmu = c.mu_max / n2;
        inv_d2 = 1.0 / dmu;
        sqr_mu_max = c.mu_max * c.mu_max;
    } else if (mode == MODE_RPPI_MOCKS) {
        // pi column width pimax / npibin; P.pimax and P.sqr_mu_max carry pimax, pimax^2
        n2 = (int)c.pimax;
        pimax = c.pimax;
        const double dpi = pimax / n2;
        inv_d2 = 1.0 / dpi;
        sqr_mu_max = pimax * pimax;
    } else if (mode == MODE_SMU_MOCKS) {
        n2 = c.nmu_bins;
        pimax = c.mu_max;
        const double dmu = c.mu_max / (double)n2;
        inv_d2 = 1.0 / dmu;
        sqr_mu_max = c.mu_max * c.mu_max;
    } else {
        pimax = rmax;
    }
    // on-sky engines: the largest 3-D separation a counted pair can have
    const double max_sep = mode == MODE_RPPI_MOCKS ? sqrt(rmax * rmax + pimax * pimax) : rmax;
    R.n2 = n2;
    const int64_t nslots =
        (mode == MODE_RPPI || mode == MODE_SMU || onsky) ? (int64_t)(n2 + 1) * (nbin + 1) : nbin;
    std::vector<double> rupp_sqr(nbin);
    for (int i = 0; i < nbin; i++) rupp_sqr[i] = R.rupp[i] * R.rupp[i];
    int sorted = 1;
    for (int i = 1; i + 1 < nbin; i++)
        if (!(rupp_sqr[i] >= rupp_sqr[i - 1])) sorted = 0;

    // ---- inputs in HBM
    const void *X1 = to_device("in_x1", c.X1, c.N1, esize), *Y1 = to_device("in_y1", c.Y1, c.N1, esize),
               *Z1 = to_device("in_z1", c.Z1, c.N1, esize);
    const void *W1 = need_w ? to_device("in_w1", extra->weights0.weights[0], c.N1, esize) : nullptr;
    const void *X2 = nullptr, *Y2 = nullptr, *Z2 = nullptr, *W2 = nullptr;
    if (!c.autocorr) {
        X2 = to_device("in_x2", c.X2, c.N2, esize);
        Y2 = to_device("in_y2", c.Y2, c.N2, esize);
        Z2 = to_device("in_z2", c.Z2, c.N2, esize);
        W2 = need_w ? to_device("in_w2", extra->weights1.weights[0], c.N2, esize) : nullptr;
    }

    // ---- extent and wraps
    double ext[6], wrap[3];
    int per[3];
    if (xiwp) {
        ext[0] = ext[1] = ext[2] = 0.0;
        ext[3] = ext[4] = ext[5] = c.boxsize;
        wrap[0] = wrap[1] = wrap[2] = c.boxsize;
        per[0] = per[1] = per[2] = 1;
    } else {
        double e1[6];
        device_minmax(X1, Y1, Z1, c.N1, esize, e1);
        memcpy(ext, e1, sizeof ext);
        if (!c.autocorr) {
            double e2[6];
            device_minmax(X2, Y2, Z2, c.N2, esize, e2);
            for (int d = 0; d < 3; d++) {
                ext[d] = e2[d] < ext[d] ? e2[d] : ext[d];
                ext[3 + d] = e2[3 + d] > ext[3 + d] ? e2[3 + d] : ext[3 + d];
            }
        }
        if (!onsky && o->periodic && o->boxsize == BOXSIZE_NOTGIVEN) {
            fprintf(stderr, "boxsize = %g must be specified with periodic wrap. Please specify a non-zero boxsize, or "
                            "zero to detect the particle extent, or -1 to make a dimension non-periodic.\n", o->boxsize);
            return EXIT_FAILURE;
        }
        const double by = o->boxsize_y == BOXSIZE_NOTGIVEN ? o->boxsize : o->boxsize_y;
        const double bz = o->boxsize_z == BOXSIZE_NOTGIVEN ? o->boxsize : o->boxsize_z;
        per[0] = o->periodic && o->boxsize_x >= 0;
        per[1] = o->periodic && by >= 0;
        per[2] = o->periodic && bz >= 0;
        wrap[0] = per[0] ? (o->boxsize_x > 0 ? o->boxsize_x : (ext[3] - ext[0])) : 0.;
        wrap[1] = per[1] ? (by > 0 ? by : (ext[4] - ext[1])) : 0.;
        wrap[2] = per[2] ? (bz > 0 ? bz : (ext[5] - ext[2])) : 0.;
        if (onsky) {
            per[0] = per[1] = per[2] = 0;
            wrap[0] = wrap[1] = wrap[2] = 0.;
        }
    }

    // ---- autotuned refine factors / histogram replication
    if (!E.tuning) {
        E.force_R = 0;
        E.force_fetch = 0;
        E.last_autotuned = 0;
        if (E.autotune && get_bin_refine_scheme(o) == BINNING_DFL && !onsky) {
            std::array<int, 5> ch;
            const std::string key = tune_key(mode, c.N1, c.autocorr ? 0 : c.N2, c.autocorr, rmax, pimax, ext, wrap,
                                             need_w, o->need_avg_sep, esize, nbin);
            if (autotune_choice(c, o, extra, key, ch) != EXIT_SUCCESS) return EXIT_FAILURE;
            set_custom_bin_refine_factors(o, ch.data());
            E.force_R = ch[3];
            E.force_fetch = ch[4];
        }
    }

    // ---- grid
    int refine[3];
    GridSpec g;
    if (choose_grid(mode, c.N1, onsky ? max_sep : rmax, pimax, c.boxsize, ext, wrap, o, g, refine, c.N2) != EXIT_SUCCESS)
        return EXIT_FAILURE;

    // ---- gridding on the GPU
    DevLattice L1, L2;
    if (gridlink_device("l1", g, c.N1, X1, Y1, Z1, W1, esize, L1) != EXIT_SUCCESS) return EXIT_FAILURE;
    if (!c.autocorr) {
        if (gridlink_device("l2", g, c.N2, X2, Y2, Z2, W2, esize, L2) != EXIT_SUCCESS) return EXIT_FAILURE;
    } else {
        L2 = L1;
    }
    const double t_grid = now_s();
    check_interrupt();

    // ---- cell pairs
    double max3 = -1, max2 = -1, max1 = -1;
    if (onsky) max3 = max_sep;
    else if (mode == MODE_DD || mode == MODE_XI) max3 = rmax;
    else if (is2d) { max2 = rmax; max1 = pimax; }
    else { max3 = rmax; max1 = pimax; }
    CellPairRec* cps = nullptr;
    int ncp = 0;
    double tested = 0;
    if (cellpairs_device(L1, L2, refine, per, wrap, max3, max2, max1, o->enable_min_sep_opt, c.autocorr, &cps, &ncp,
                         &tested) != EXIT_SUCCESS)
        return EXIT_FAILURE;
    check_interrupt();

    // ---- launch parameters
    PairParams P;
    std::vector<unsigned short> lut;
    {
        const double coord_scale = fabs(ext[0]) + fabs(ext[1]) + fabs(ext[2]) + fabs(ext[3]) + fabs(ext[4]) +
                                   fabs(ext[5]) + wrap[0] + wrap[1] + wrap[2] + rmax;
        const double cs = fmax(g.xbs, fmax(g.ybs, g.zbs));
        const int rmaxr = refine[0] > refine[1] ? (refine[0] > refine[2] ? refine[0] : refine[2])
                                                : (refine[1] > refine[2] ? refine[1] : refine[2]);
        build_params(mode, resolve_isa(o->instruction_set), nbin, n2, (int)nslots, rupp_sqr, sorted, rmax, pimax,
                     inv_d2, sqr_mu_max, o->need_avg_sep, need_w, c.autocorr, wrap, coord_scale, cs, rmaxr, P, lut);
        if (onsky) {
            // 3-D limit for the pruning, the filter and the slab window; rp^2 limit
            // for the (rp, pi) bins (countpairs_rp_pi_mocks_impl.c.src:303-306)
            P.sqr_rpmax = rupp_sqr[nbin - 1];
            P.sqr_rmax = mode == MODE_RPPI_MOCKS ? rmax * rmax + pimax * pimax : rupp_sqr[nbin - 1];
            P.max_all_dz_same = sqrt(P.sqr_rmax - 0.0 - 0.0);
            P.zreach = sqrt(P.sqr_rmax);
            P.near_r2 = 1e-9 * P.sqr_rmax + 1e-9 * max_sep * coord_scale;
        }
        set_quant_filter(L1, P);
        set_quant_filter(L2, P);
    }

    // ---- histograms
    unsigned long long* g_np = (unsigned long long*)E.arena.get("h_np", sizeof(uint64_t) * nslots);
    double* g_avg = (double*)E.arena.get("h_avg", sizeof(double) * nslots);
    double* g_w = (double*)E.arena.get("h_w", sizeof(double) * nslots);
    double* d_edges = (double*)E.arena.get("edges", sizeof(double) * nbin);
    k_fill_u64<<<(unsigned)((nslots + 255) / 256), 256, 0, E.stream>>>(nslots, g_np, 0ull);
    k_fill_u64<<<(unsigned)((nslots + 255) / 256), 256, 0, E.stream>>>(nslots, (unsigned long long*)g_avg, 0ull);
    k_fill_u64<<<(unsigned)((nslots + 255) / 256), 256, 0, E.stream>>>(nslots, (unsigned long long*)g_w, 0ull);
    CF_CHECK(hipMemcpyAsync(d_edges, rupp_sqr.data(), sizeof(double) * nbin, hipMemcpyHostToDevice, E.stream));
    unsigned short* d_lut = (unsigned short*)E.arena.get("lut", sizeof(unsigned short) * lut.size());
    CF_CHECK(hipMemcpyAsync(d_lut, lut.data(), sizeof(unsigned short) * lut.size(), hipMemcpyHostToDevice, E.stream));

    int nitems = 0;
    TunedRun T;
    bool tuned = ncp > 0 && E.engine_kind != CF_AMD_ENGINE_PLAIN;
    if (tuned) {
        const int rc = launch_tuned(ncp, cps, lats_for(L1, L2), 0, P, (int)nslots, nbin, d_edges, d_lut, g_np,
                                    g_avg, g_w, c.cell_timer || E.model.recording, T);
        if (rc == kNoTunedFit) {
            // more edges than a workgroup's LDS holds next to the queues: plain kernel
            tuned = false;
        } else if (rc != EXIT_SUCCESS) {
            return EXIT_FAILURE;
        }
        nitems = T.nitems;
    }
    if (!tuned) {
        // the plain kernel keeps an FP64 LDS histogram when it fits 64 KiB,
        // else adds to the global histogram
        const size_t lds_hist = (size_t)nslots * 24 + (size_t)nbin * 8;
        const bool plain_gh = lds_hist + 4096 > 64 * 1024;
        T.global_hist = plain_gh ? 1 : 0;
        CF_CHECK(hipEventRecord(E.ev_list, E.stream));
        CF_CHECK(hipEventRecord(E.ev0, E.stream));
        if (ncp > 0) {
            if (plain_gh)
                k_pairs_plain<true><<<ncp, 128, 0, E.stream>>>(P, L1, L2, cps, d_edges, g_np, g_avg, g_w);
            else
                k_pairs_plain<false><<<ncp, 128, lds_hist, E.stream>>>(P, L1, L2, cps, d_edges, g_np, g_avg, g_w);
            CF_CHECK(hipGetLastError());
        }
        CF_CHECK(hipEventRecord(E.ev1, E.stream));
    }
    R.np.assign(nslots, 0);
    R.avg.assign(nslots, 0.0);
    R.wavg.assign(nslots, 0.0);
    CF_CHECK(hipMemcpyAsync(R.np.data(), g_np, sizeof(uint64_t) * nslots, hipMemcpyDeviceToHost, E.stream));
    CF_CHECK(hipMemcpyAsync(R.avg.data(), g_avg, sizeof(double) * nslots, hipMemcpyDeviceToHost, E.stream));
    CF_CHECK(hipMemcpyAsync(R.wavg.data(), g_w, sizeof(double) * nslots, hipMemcpyDeviceToHost, E.stream));
    if (need_w) {
        double* ws = (double*)E.arena.get("wsums", sizeof(double) * 2);
        fill_u64(ws, 2);
        if (c.N1 > 0) k_weight_sums<<<256, 256, 0, E.stream>>>(c.N1, L1.w, ws);
        CF_CHECK(hipGetLastError());
        double h[2];
        CF_CHECK(hipMemcpyAsync(h, ws, sizeof h, hipMemcpyDeviceToHost, E.stream));
        CF_CHECK(hipStreamSynchronize(E.stream));
        R.wsum = h[0];
        R.wsq = h[1];
    }
    if (c.cell_timer && tuned) {
        // per-cell-pair timings in the order of the cell-pair list (wp_impl:433-463)
        TimingRow* rows = (TimingRow*)E.arena.get("t_rows", sizeof(TimingRow) * ncp);
        k_timing_rows<<<(ncp + 255) / 256, 256, 0, E.stream>>>(ncp, cps, lats_for(L1, L2), 0, T.ticks, T.cu,
                                                               E.ns_per_tick, rows);
        CF_CHECK(hipGetLastError());
        R.timings.resize(ncp);
        CF_CHECK(hipMemcpyAsync(R.timings.data(), rows, sizeof(TimingRow) * ncp, hipMemcpyDeviceToHost, E.stream));
    }
    CF_CHECK(hipStreamSynchronize(E.stream));
    check_interrupt();
    float kms = 0, lms = 0;
    CF_CHECK(hipEventElapsedTime(&kms, E.ev0, E.ev1));
    CF_CHECK(hipEventElapsedTime(&lms, E.ev_list, E.ev0));

    st.list_ms = lms;
    st.hist_copies = tuned && !T.global_hist ? P.hist_copies : 0;
    st.queue_fetch = tuned ? T.fetch : 0;
    st.global_hist = T.global_hist;
    st.drain_pairs = tuned ? T.drain_pairs : 0;
    st.workgroup_threads = tuned ? T.threads : 0;
    st.autotuned = E.last_autotuned;
    st.sched = tuned ? (E.sched_kind == CF_AMD_SCHED_LEARNED && !E.model.trained ? CF_AMD_SCHED_HEURISTIC
                                                                                  : E.sched_kind)
                     : CF_AMD_SCHED_NONE;
    st.nmesh[0] = g.nx; st.nmesh[1] = g.ny; st.nmesh[2] = g.nz;
    st.refine[0] = refine[0]; st.refine[1] = refine[1]; st.refine[2] = refine[2];
    st.num_cell_pairs = ncp;
    st.num_work_items = nitems;
    st.tested_pairs = tested;
    st.kernel_ms = kms;
    st.grid_ms = (t_grid - t_start) * 1e3;
    st.total_ms = (now_s() - t_start) * 1e3;
    // the plain kernel also runs where the tuned one cannot (kNoTunedFit)
    st.engine = ncp > 0 && !tuned ? CF_AMD_ENGINE_PLAIN : E.engine_kind;
    E.last = st;
    return EXIT_SUCCESS;
}


// ----------------------------------------------------------------- batches
// DD(r) auto-correlations of many independent catalogs in one device work list
// and one persistent pair-kernel launch. Catalog k gets exactly the results that
// countpairs() gives for it with a copy of `options`.
static int run_batch(int ncat, const int64_t* ND, void** X, void** Y, void** Z, void** W, const char* binfile,
                     struct config_options* options, struct extra_options* extra, std::vector<CallOut>& outs)
{
    Engine& E = eng();
    InterruptScope interrupts;
    const double t_start = now_s();
    const size_t esize = options->float_type;
    const int need_w = extra->weight_method != NONE;
    E.last_autotuned = 0;
    E.force_R = 0;
    E.force_fetch = 0;
    int batch_fetch = 0;          // the autotuner's queue batch for the largest catalog
    int64_t batch_fetch_n = -1;
    if (need_w && extra->weight_method != PAIR_PRODUCT) {
        fprintf(stderr, "Error: unknown weight method %d\n", (int)extra->weight_method);
        return EXIT_FAILURE;
    }
    std::vector<double> rupp;
    double rmin = 0, rmax = 0;
    int nbin = 0;
    if (setup_bins(binfile, &rmin, &rmax, &nbin, rupp) != EXIT_SUCCESS) return EXIT_FAILURE;
    if (!(rmin >= 0.0 && rmax > 0.0 && rmin < rmax && nbin > 0)) {
        fprintf(stderr, "Error: Could not setup with R bins correctly. (rmin = %lf, rmax = %lf, with nbins = %d).\n",
                rmin, rmax, nbin);
        return EXIT_FAILURE;
    }
    std::vector<double> rupp_sqr(nbin);
    for (int i = 0; i < nbin; i++) rupp_sqr[i] = rupp[i] * rupp[i];
    int sorted = 1;
    for (int i = 1; i + 1 < nbin; i++)
        if (!(rupp_sqr[i] >= rupp_sqr[i - 1])) sorted = 0;
    const int nslots = nbin;

    // Phased over all catalogs, one host synchronisation per phase:
    //   1. inputs and extents            -> sync (extents)
    //   2. grids (host), gridding        (no sync)
    //   3. cell-pair flags + scans, counts gathered on the device -> sync
    //   4. flags + scans again, records emitted into the concatenated list
    std::vector<DevLattice> hl(ncat);
    std::vector<const void*> dx(ncat), dy(ncat), dz(ncat), dw(ncat);
    for (int k = 0; k < ncat; k++) {
        const std::string tag = "b" + std::to_string(k);
        if (need_w && W[k] == nullptr) {
            fprintf(stderr, "Error: catalog %d has no weights\n", k);
            return EXIT_FAILURE;
        }
        dx[k] = to_device(tag + "x", X[k], ND[k], esize);
        dy[k] = to_device(tag + "y", Y[k], ND[k], esize);
        dz[k] = to_device(tag + "z", Z[k], ND[k], esize);
        dw[k] = need_w ? to_device(tag + "w", W[k], ND[k], esize) : nullptr;
    }
    if (options->periodic && options->boxsize == BOXSIZE_NOTGIVEN) {
        fprintf(stderr, "boxsize must be specified with periodic wrap.\n");
        return EXIT_FAILURE;
    }
    std::vector<double> exts(6 * (size_t)ncat);
    {
        double* d_ext = (double*)E.arena.get("b_ext", sizeof(double) * 6 * ncat);
        for (int k = 0; k < ncat; k++) device_minmax_async(dx[k], dy[k], dz[k], ND[k], esize, d_ext + 6 * k);
        CF_CHECK(hipMemcpyAsync(exts.data(), d_ext, sizeof(double) * 6 * ncat, hipMemcpyDeviceToHost, E.stream));
        CF_CHECK(hipStreamSynchronize(E.stream));
    }
    std::vector<std::array<int, 3>> refines(ncat), pers(ncat);
    double wrap0[3] = {0, 0, 0}, coord_scale = 0, cs_max = 0, tested_all = 0;
    int rmaxr = 1;
    unsigned long long* d_oob = (unsigned long long*)E.arena.get("b_oob", sizeof(unsigned long long));
    fill_u64(d_oob, 1);
    for (int k = 0; k < ncat; k++) {
        struct config_options oc = *options;
        const double* ext = &exts[6 * (size_t)k];
        double wrap[3];
        int per[3];
        const double by = oc.boxsize_y == BOXSIZE_NOTGIVEN ? oc.boxsize : oc.boxsize_y;
        const double bz = oc.boxsize_z == BOXSIZE_NOTGIVEN ? oc.boxsize : oc.boxsize_z;
        per[0] = oc.periodic && oc.boxsize_x >= 0;
        per[1] = oc.periodic && by >= 0;
        per[2] = oc.periodic && bz >= 0;
        wrap[0] = per[0] ? (oc.boxsize_x > 0 ? oc.boxsize_x : (ext[3] - ext[0])) : 0.;
        wrap[1] = per[1] ? (by > 0 ? by : (ext[4] - ext[1])) : 0.;
        wrap[2] = per[2] ? (bz > 0 ? bz : (ext[5] - ext[2])) : 0.;
        if (k == 0) {
            memcpy(wrap0, wrap, sizeof wrap0);
        } else if (wrap[0] != wrap0[0] || wrap[1] != wrap0[1] || wrap[2] != wrap0[2]) {
            fprintf(stderr, "Error: the catalogs of a batch must share their periodic wraps\n");
            return EXIT_FAILURE;
        }
        if (E.autotune && get_bin_refine_scheme(&oc) == BINNING_DFL && ND[k] > 0) {
            // per size class: tuned on a single-catalog call of this catalog
            std::array<int, 5> ch;
            const std::string key = tune_key(MODE_DD, ND[k], 0, 1, rmax, rmax, ext, wrap, need_w, oc.need_avg_sep,
                                             esize, nbin);
            CallSpec cs{MODE_DD, ND[k], 0, X[k], Y[k], Z[k], nullptr, nullptr, nullptr, 1, binfile, 0, 0, 0, 0, 1};
            struct extra_options ek = *extra;
            ek.weights0.weights[0] = need_w ? W[k] : nullptr;
            if (autotune_choice(cs, &oc, &ek, key, ch) != EXIT_SUCCESS) return EXIT_FAILURE;
            set_custom_bin_refine_factors(&oc, ch.data());
            if (ND[k] > batch_fetch_n) {
                batch_fetch_n = ND[k];
                batch_fetch = ch[4];
            }
        }
        int refine[3];
        GridSpec g;
        if (choose_grid(MODE_DD, ND[k], rmax, rmax, 0.0, ext, wrap, &oc, g, refine) != EXIT_SUCCESS)
            return EXIT_FAILURE;
        if (gridlink_launch("b" + std::to_string(k), g, ND[k], dx[k], dy[k], dz[k], dw[k], esize, hl[k], d_oob) !=
            EXIT_SUCCESS)
            return EXIT_FAILURE;
        refines[k] = {refine[0], refine[1], refine[2]};
        pers[k] = {per[0], per[1], per[2]};
        for (int d = 0; d < 6; d++) coord_scale = fmax(coord_scale, fabs(ext[d]));
        cs_max = fmax(cs_max, fmax(g.xbs, fmax(g.ybs, g.zbs)));
        for (int d = 0; d < 3; d++) rmaxr = refine[d] > rmaxr ? refine[d] : rmaxr;
    }
    // cell-pair counts per catalog
    std::vector<int> ncps(ncat);
    {
        int* d_ncp = (int*)E.arena.get("b_ncp", sizeof(int) * ncat);
        for (int k = 0; k < ncat; k++) {
            CpPlan cp;
            if (cellpairs_flag(hl[k], hl[k], refines[k].data(), pers[k].data(), wrap0, rmax, -1, -1,
                               options->enable_min_sep_opt, 1, cp) != EXIT_SUCCESS)
                return EXIT_FAILURE;
            k_copy_int<<<1, 64, 0, E.stream>>>(cp.pos + cp.total, d_ncp + k);
        }
        CF_CHECK(hipGetLastError());
        unsigned long long h_oob = 0;
        CF_CHECK(hipMemcpyAsync(ncps.data(), d_ncp, sizeof(int) * ncat, hipMemcpyDeviceToHost, E.stream));
        CF_CHECK(hipMemcpyAsync(&h_oob, d_oob, sizeof(h_oob), hipMemcpyDeviceToHost, E.stream));
        CF_CHECK(hipStreamSynchronize(E.stream));
        if (check_oob(h_oob) != EXIT_SUCCESS) return EXIT_FAILURE;
    }
    int ncp_all = 0;
    std::vector<int> cp_off(ncat);
    for (int k = 0; k < ncat; k++) {
        cp_off[k] = ncp_all;
        if ((int64_t)ncp_all + ncps[k] >= (int64_t)INT32_MAX) {
            fprintf(stderr, "Error: the batch's cell pairs exceed the engine's 32-bit indexing\n");
            return EXIT_FAILURE;
        }
        ncp_all += ncps[k];
    }
    check_interrupt();
    coord_scale = 6 * coord_scale + wrap0[0] + wrap0[1] + wrap0[2] + rmax;
    // concatenated work list, emitted in place
    CellPairRec* cps = (CellPairRec*)E.arena.get("bcp_all", sizeof(CellPairRec) * ((size_t)ncp_all + 1));
    double* d_tested = (double*)E.arena.get("b_tested", sizeof(double) * ncat);
    fill_u64(d_tested, ncat);
    for (int k = 0; k < ncat; k++) {
        if (ncps[k] == 0) continue;
        CpPlan cp;
        if (cellpairs_flag(hl[k], hl[k], refines[k].data(), pers[k].data(), wrap0, rmax, -1, -1,
                           options->enable_min_sep_opt, 1, cp) != EXIT_SUCCESS)
            return EXIT_FAILURE;
        cellpairs_emit(cp, cps + cp_off[k], d_tested + k, k);
    }
    {
        std::vector<double> tk(ncat);
        CF_CHECK(hipMemcpyAsync(tk.data(), d_tested, sizeof(double) * ncat, hipMemcpyDeviceToHost, E.stream));
        CF_CHECK(hipStreamSynchronize(E.stream));
        for (int k = 0; k < ncat; k++) tested_all += tk[k];
    }
    const double t_grid = now_s();

    PairParams P;
    std::vector<unsigned short> lut;
    E.force_fetch = batch_fetch;
    build_params(MODE_DD, resolve_isa(options->instruction_set), nbin, 0, nslots, rupp_sqr, sorted, rmax, rmax, 0,
                 0, options->need_avg_sep, need_w, 1, wrap0, coord_scale, cs_max, rmaxr, P, lut);
    E.force_fetch = 0;
    for (int k = 0; k < ncat; k++) set_quant_filter(hl[k], P);
    DevLattice* lats = (DevLattice*)E.arena.get("blats", sizeof(DevLattice) * ncat);
    CF_CHECK(hipMemcpyAsync(lats, hl.data(), sizeof(DevLattice) * ncat, hipMemcpyHostToDevice, E.stream));

    const size_t tot = (size_t)ncat * nslots;
    unsigned long long* g_np = (unsigned long long*)E.arena.get("bh_np", sizeof(uint64_t) * tot);
    double* g_avg = (double*)E.arena.get("bh_avg", sizeof(double) * tot);
    double* g_w = (double*)E.arena.get("bh_w", sizeof(double) * tot);
    double* d_edges = (double*)E.arena.get("edges", sizeof(double) * nbin);
    unsigned short* d_lut = (unsigned short*)E.arena.get("lut", sizeof(unsigned short) * lut.size());
    const unsigned fb = (unsigned)std::min<size_t>(4096, (tot + 255) / 256);
    k_fill_u64<<<fb, 256, 0, E.stream>>>((int64_t)tot, g_np, 0ull);
    k_fill_u64<<<fb, 256, 0, E.stream>>>((int64_t)tot, (unsigned long long*)g_avg, 0ull);
    k_fill_u64<<<fb, 256, 0, E.stream>>>((int64_t)tot, (unsigned long long*)g_w, 0ull);
    CF_CHECK(hipMemcpyAsync(d_edges, rupp_sqr.data(), sizeof(double) * nbin, hipMemcpyHostToDevice, E.stream));
    CF_CHECK(hipMemcpyAsync(d_lut, lut.data(), sizeof(unsigned short) * lut.size(), hipMemcpyHostToDevice, E.stream));

    int nitems = 0;
    TunedRun T;
    if (ncp_all > 0) {
        const int rc = launch_tuned(ncp_all, cps, lats, 1, P, nslots, nbin, d_edges, d_lut, g_np, g_avg, g_w,
                                    E.model.recording, T);
        if (rc == kNoTunedFit) {
            fprintf(stderr, "Error: %d bins are more than the batched kernel's LDS holds; count the catalogs with "
                            "countpairs one at a time\n", nbin - 1);
            return EXIT_FAILURE;
        }
        if (rc != EXIT_SUCCESS) return EXIT_FAILURE;
        nitems = T.nitems;
    } else {
        CF_CHECK(hipEventRecord(E.ev_list, E.stream));
        CF_CHECK(hipEventRecord(E.ev0, E.stream));
        CF_CHECK(hipEventRecord(E.ev1, E.stream));
    }
    std::vector<uint64_t> np(tot);
    std::vector<double> avg(tot), wv(tot);
    CF_CHECK(hipMemcpyAsync(np.data(), g_np, sizeof(uint64_t) * tot, hipMemcpyDeviceToHost, E.stream));
    CF_CHECK(hipMemcpyAsync(avg.data(), g_avg, sizeof(double) * tot, hipMemcpyDeviceToHost, E.stream));
    CF_CHECK(hipMemcpyAsync(wv.data(), g_w, sizeof(double) * tot, hipMemcpyDeviceToHost, E.stream));
    std::vector<double> wsums(2 * (size_t)ncat, 0.0);
    if (need_w) {
        double* ws = (double*)E.arena.get("bwsums", sizeof(double) * 2 * ncat);
        fill_u64(ws, 2 * (int64_t)ncat);
        for (int k = 0; k < ncat; k++)
            if (ND[k] > 0) k_weight_sums<<<64, 256, 0, E.stream>>>(ND[k], hl[k].w, ws + 2 * k);
        CF_CHECK(hipGetLastError());
        CF_CHECK(hipMemcpyAsync(wsums.data(), ws, sizeof(double) * 2 * ncat, hipMemcpyDeviceToHost, E.stream));
    }
    CF_CHECK(hipStreamSynchronize(E.stream));
    check_interrupt();
    float kms = 0, lms = 0;
    CF_CHECK(hipEventElapsedTime(&kms, E.ev0, E.ev1));
    CF_CHECK(hipEventElapsedTime(&lms, E.ev_list, E.ev0));
    outs.assign(ncat, CallOut());
    for (int k = 0; k < ncat; k++) {
        CallOut& R = outs[k];
        R.rupp = rupp;
        R.nbin = nbin;
        R.need_w = need_w;
        R.np.assign(np.begin() + (size_t)k * nslots, np.begin() + (size_t)(k + 1) * nslots);
        R.avg.assign(avg.begin() + (size_t)k * nslots, avg.begin() + (size_t)(k + 1) * nslots);
        R.wavg.assign(wv.begin() + (size_t)k * nslots, wv.begin() + (size_t)(k + 1) * nslots);
        R.wsum = wsums[2 * k];
        R.wsq = wsums[2 * k + 1];
    }
    cf_amd_stats st;
    memset(&st, 0, sizeof st);
    st.num_cell_pairs = ncp_all;
    st.num_work_items = nitems;
    st.tested_pairs = tested_all;
    st.kernel_ms = kms;
    st.grid_ms = (t_grid - t_start) * 1e3;
    st.total_ms = (now_s() - t_start) * 1e3;
    st.engine = CF_AMD_ENGINE_TUNED;
    st.list_ms = lms;
    st.hist_copies = P.hist_copies;
    st.queue_fetch = T.fetch;
    st.drain_pairs = T.drain_pairs;
    st.workgroup_threads = T.threads;
    st.autotuned = E.last_autotuned;
    st.sched = E.sched_kind == CF_AMD_SCHED_LEARNED && !E.model.trained ? CF_AMD_SCHED_HEURISTIC : E.sched_kind;
    E.last = st;
    return EXIT_SUCCESS;
}

// ---------------------------------------------------------- graph batches
// A batch of DD(r) autocorrelations captured once into a hipGraph and replayed
// on the current contents of the catalogs' device arrays (cf_bgraph.hip for
// the device side). The capture runs the engine's regular batch once on the
// capture-time positions to size the cell-pair list and the work list (with a
// margin); every other buffer is sized from the box.
struct BatchGraph {
    int ncat = 0;
    std::vector<int64_t> nd;
    size_t esize = 8;
    int need_w = 0, need_avg = 0;
    int nbin = 0;
    std::vector<double> rupp;
    std::vector<void*> owned;            // device buffers of the graph
    std::vector<int64_t> ncells_cap;
    int pairs_cap = 0, items_cap = 0;
    unsigned long long* g_np = nullptr;
    double *g_avg = nullptr, *g_w = nullptr, *ws = nullptr, *tested = nullptr;
    int *fail = nullptr, *cp_off = nullptr, *nitems_eff = nullptr;
    BgCat* cats = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int hist_copies = 0, queue_fetch = 0, drain_pairs = 0, threads = 0, sched = CF_AMD_SCHED_NONE;
    void* alloc(size_t bytes)
    {
        void* p = nullptr;
        CF_CHECK(hipMalloc(&p, bytes ? bytes : 1));
        owned.push_back(p);
        return p;
    }
    ~BatchGraph()
    {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        for (void* p : owned) (void)hipFree(p);
    }
};

// Exclusive scan with block-sum buffers allocated at capture (no arena:
// buffers a captured graph reads must outlive any later arena regrowth).
struct GraphScan {
    std::vector<std::pair<int*, int*>> lv;  // per level: block sums, block offsets
    void init(int64_t nmax, BatchGraph& G)
    {
        const int64_t per = (int64_t)kScanBlock * kScanItems;
        int64_t n = nmax > 0 ? nmax : 1;
        for (;;) {
            const int64_t nb = (n + per - 1) / per;
            lv.push_back({(int*)G.alloc(sizeof(int) * (nb + 1)), (int*)G.alloc(sizeof(int) * (nb + 1))});
            if (nb <= 1) break;
            n = nb;
        }
    }
    void run(const int* in, int* out, int64_t n, hipStream_t st, size_t level = 0)
    {
        if (n <= 0) {
            k_fill_i32<<<1, 64, 0, st>>>(1, out, 0);
            return;
        }
        const int64_t per = (int64_t)kScanBlock * kScanItems;
        const int64_t nblocks = (n + per - 1) / per;
        k_scan_blocks<<<(unsigned)nblocks, kScanBlock, 0, st>>>(n, in, out, lv[level].first);
        if (nblocks > 1) {
            run(lv[level].first, lv[level].second, nblocks, st, level + 1);
            k_scan_add<<<(unsigned)nblocks, kScanBlock, 0, st>>>(n, out, lv[level].second);
            k_copy_int<<<1, 64, 0, st>>>(lv[level].second + nblocks, out + n);
        } else {
            k_copy_int<<<1, 64, 0, st>>>(lv[level].first, out + n);
        }
    }
};

static void fix_refine(struct config_options* o);
static void fix_maxcells(struct config_options* o);

static int bg_create(int ncat, const int64_t* ND, void** X, void** Y, void** Z, void** W, const char* binfile,
                     struct config_options* options, struct extra_options* extra, BatchGraph& G)
{
    Engine& E = eng();
    const size_t esize = options->float_type;
    const int need_w = extra->weight_method != NONE;
    if (need_w && extra->weight_method != PAIR_PRODUCT) {
        fprintf(stderr, "Error: unknown weight method %d\n", (int)extra->weight_method);
        return EXIT_FAILURE;
    }
    // the grid capacities come from the box: a periodic box with its size given
    const double by = options->boxsize_y == BOXSIZE_NOTGIVEN ? options->boxsize : options->boxsize_y;
    const double bz = options->boxsize_z == BOXSIZE_NOTGIVEN ? options->boxsize : options->boxsize_z;
    const double wrap[3] = {options->boxsize_x, by, bz};
    if (!options->periodic || !(wrap[0] > 0 && wrap[1] > 0 && wrap[2] > 0)) {
        fprintf(stderr, "Error: a graph batch needs a periodic box with its size given on every axis\n");
        return EXIT_FAILURE;
    }
    for (int k = 0; k < ncat; k++) {
        const void* ins[4] = {X[k], Y[k], Z[k], need_w ? W[k] : nullptr};
        for (int d = 0; d < 4; d++)
            if (ND[k] > 0 && (d < 3 || need_w) && (ins[d] == nullptr || !is_device_ptr(ins[d]))) {
                fprintf(stderr, "Error: a graph batch reads positions%s from device memory (catalog %d)\n",
                        need_w ? " and weights" : "", k);
                return EXIT_FAILURE;
            }
    }
    // the regular batch on the capture-time positions sizes the two lists
    {
        struct config_options oc = *options;
        std::vector<CallOut> outs;
        const int sched = E.sched_kind;
        if (run_batch(ncat, ND, X, Y, Z, W, binfile, &oc, extra, outs) != EXIT_SUCCESS) return EXIT_FAILURE;
        (void)sched;
    }
    const double ncp0 = (double)E.last.num_cell_pairs, items0 = (double)E.last.num_work_items;
    const int fetch0 = E.last.queue_fetch;
    std::vector<double> rupp;
    double rmin = 0, rmax = 0;
    int nbin = 0;
    if (setup_bins(binfile, &rmin, &rmax, &nbin, rupp) != EXIT_SUCCESS) return EXIT_FAILURE;
    std::vector<double> rupp_sqr(nbin);
    for (int i = 0; i < nbin; i++) rupp_sqr[i] = rupp[i] * rupp[i];
    int sorted = 1;
    for (int i = 1; i + 1 < nbin; i++)
        if (!(rupp_sqr[i] >= rupp_sqr[i - 1])) sorted = 0;
    G.ncat = ncat;
    G.nd.assign(ND, ND + ncat);
    G.esize = esize;
    G.need_w = need_w;
    G.need_avg = options->need_avg_sep;
    G.nbin = nbin;
    G.rupp = rupp;
    const int nslots = nbin;

    BgOpts bo;
    memset(&bo, 0, sizeof bo);
    bo.rmax = rmax;
    bo.sqr_rmax = rupp_sqr[nbin - 1];
    bo.sqr_rmin = rupp_sqr[0];
    int refine_cap[3];
    {
        struct config_options oc = *options;
        fix_maxcells(&oc);
        fix_refine(&oc);
        bo.custom = get_bin_refine_scheme(&oc) != BINNING_DFL;
        bo.max_cells = oc.max_cells_per_dim;
        for (int d = 0; d < 3; d++) {
            bo.wrap[d] = wrap[d];
            bo.per[d] = 1;
            bo.refine[d] = oc.bin_refine_factors[d];
            int r = bo.refine[d];
            if (!bo.custom && rmax < 0.05 * wrap[d]) r = 1;
            if (!bo.custom && d < 2) r += 1;  // the boost
            refine_cap[d] = std::max(r, bo.refine[d] + (d < 2 && !bo.custom ? 1 : 0));
        }
    }
    bo.min_sep_opt = options->enable_min_sep_opt;
    // cells: nmesh = (int)(refine * extent / rmax) with the extent within the
    // box, plus one for positions on its faces
    int64_t ncap = 1;
    for (int d = 0; d < 3; d++) {
        int64_t m = (int64_t)(refine_cap[d] * wrap[d] / rmax) + 1;
        m = std::max<int64_t>(2, std::min<int64_t>(m, bo.max_cells));
        ncap *= m;
    }
    const int64_t nngb_cap = (int64_t)(2 * refine_cap[0] + 1) * (2 * refine_cap[1] + 1) * (2 * refine_cap[2] + 1);
    const int64_t cand_cap = ncap * nngb_cap;
    if (ncap * kSlabs >= (int64_t)INT32_MAX || cand_cap >= (int64_t)INT32_MAX) {
        fprintf(stderr, "Error: the graph batch's grid capacity exceeds the engine's 32-bit indexing\n");
        return EXIT_FAILURE;
    }
    bo.ncells_cap = ncap;
    G.pairs_cap = (int)std::min<double>((double)INT32_MAX - 1, ncp0 * 1.25 + 4096);
    G.items_cap = (int)std::min<double>((double)INT32_MAX - 1, items0 * 1.25 + 4096);

    // launch parameters, with the coordinate scale of a box whose positions lie
    // within [0, wrap] (a wider edge window than the data's only sends more
    // pairs through the pruning replay)
    double coord_scale = 0, cs = 0;
    for (int d = 0; d < 3; d++) {
        coord_scale += 2 * wrap[d] + wrap[d];
        cs = fmax(cs, wrap[d]);
    }
    coord_scale += rmax;
    PairParams P;
    std::vector<unsigned short> lut;
    build_params(MODE_DD, resolve_isa(options->instruction_set), nbin, 0, nslots, rupp_sqr, sorted, rmax, rmax, 0,
                 0, options->need_avg_sep, need_w, 1, wrap, coord_scale, cs, refine_cap[0], P, lut);
    TunedShape ts;
    const int src = tuned_shape(1, P, nslots, nbin, ts);
    if (src == kNoTunedFit) {
        fprintf(stderr, "Error: %d bins are more than the batched kernel's LDS holds\n", nbin - 1);
        return EXIT_FAILURE;
    }
    if (src != EXIT_SUCCESS) return EXIT_FAILURE;
    // the queue batch the regular batch chose for these catalogs (launch_tuned's rule)
    if (P.fetch <= 0) P.fetch = fetch0 > 0 ? fetch0 : 1;
    G.hist_copies = P.hist_copies;
    G.queue_fetch = P.fetch;
    G.drain_pairs = ts.dp;
    G.threads = ts.threads;
    const long blocks = (long)E.ncu * ts.per_cu;

    // ---- buffers
    const size_t tot = (size_t)ncat * nslots;
    G.g_np = (unsigned long long*)G.alloc(sizeof(uint64_t) * tot);
    G.g_avg = (double*)G.alloc(sizeof(double) * tot);
    G.g_w = (double*)G.alloc(sizeof(double) * tot);
    G.ws = (double*)G.alloc(sizeof(double) * 2 * (size_t)ncat);
    G.tested = (double*)G.alloc(sizeof(double) * (size_t)ncat);
    G.fail = (int*)G.alloc(sizeof(int));
    G.cp_off = (int*)G.alloc(sizeof(int) * ((size_t)ncat + 1));
    G.nitems_eff = (int*)G.alloc(sizeof(int));
    G.cats = (BgCat*)G.alloc(sizeof(BgCat) * (size_t)ncat);
    DevLattice* lats = (DevLattice*)G.alloc(sizeof(DevLattice) * (size_t)ncat);
    double* ext = (double*)G.alloc(sizeof(double) * 6 * (size_t)ncat);
    double* mm_part = (double*)G.alloc(sizeof(double) * 6 * 512);
    unsigned long long* oob = (unsigned long long*)G.alloc(sizeof(unsigned long long));
    int* flags = (int*)G.alloc(sizeof(int) * (size_t)(cand_cap + 1));
    int* pos = (int*)G.alloc(sizeof(int) * (size_t)(cand_cap + 1));
    CellPairRec* cps = (CellPairRec*)G.alloc(sizeof(CellPairRec) * ((size_t)G.pairs_cap + 1));
    int* nch = (int*)G.alloc(sizeof(int) * ((size_t)G.pairs_cap + 1));
    int* ioff = (int*)G.alloc(sizeof(int) * ((size_t)G.pairs_cap + 1));
    WorkItem* items = (WorkItem*)G.alloc(sizeof(WorkItem) * ((size_t)G.items_cap + 1));
    int* counter = (int*)G.alloc(sizeof(int) * kQueues * kCounterStride);
    int* single = (int*)G.alloc(sizeof(int));
    double* d_edges = (double*)G.alloc(sizeof(double) * nbin);
    unsigned short* d_lut = (unsigned short*)G.alloc(sizeof(unsigned short) * lut.size());
    CF_CHECK(hipMemcpy(d_edges, rupp_sqr.data(), sizeof(double) * nbin, hipMemcpyHostToDevice));
    CF_CHECK(hipMemcpy(d_lut, lut.data(), sizeof(unsigned short) * lut.size(), hipMemcpyHostToDevice));
    const int64_t nb_cap = ncap * kSlabs;
    std::vector<DevLattice> tmpl(ncat);
    std::vector<int*> bucket(ncat), counts(ncat), fillb(ncat);
    for (int k = 0; k < ncat; k++) {
        const int64_t n = ND[k];
        DevLattice& L = tmpl[k];
        memset(&L, 0, sizeof L);
        L.x = (double*)G.alloc(sizeof(double) * (size_t)n);
        L.y = (double*)G.alloc(sizeof(double) * (size_t)n);
        L.z = (double*)G.alloc(sizeof(double) * (size_t)n);
        L.w = need_w ? (double*)G.alloc(sizeof(double) * (size_t)n) : nullptr;
        L.d4 = (double4*)G.alloc(sizeof(double4) * (size_t)(n + kQPad));
        L.q = (uint2*)G.alloc(sizeof(uint2) * (size_t)(n + kQPad));
        L.bstart = (int*)G.alloc(sizeof(int) * (size_t)(nb_cap + 1));
        L.bounds = (double*)G.alloc(sizeof(double) * 6 * (size_t)ncap);
        bucket[k] = (int*)G.alloc(sizeof(int) * (size_t)(n + 1));
        counts[k] = (int*)G.alloc(sizeof(int) * (size_t)(nb_cap + 1));
        fillb[k] = (int*)G.alloc(sizeof(int) * (size_t)(nb_cap + 1));
    }
    // scheduler buffers (as launch_tuned's)
    const int sched = E.sched_kind;
    const bool learned = sched == CF_AMD_SCHED_LEARNED && E.model.trained;
    const int64_t nsb = ((int64_t)G.pairs_cap + kSortBlock - 1) / kSortBlock;
    const int64_t nmat = (int64_t)kSortBuckets * nsb;
    float* cost = (float*)G.alloc(sizeof(float) * (size_t)G.pairs_cap);
    float* feats = learned ? (float*)G.alloc(sizeof(float) * kNumFeatures * (size_t)G.pairs_cap) : nullptr;
    double* total = (double*)G.alloc(sizeof(double));
    int* bcount = (int*)G.alloc(sizeof(int) * (size_t)(nmat + 1));
    int* boff = (int*)G.alloc(sizeof(int) * (size_t)(nmat + 1));
    int* order = (int*)G.alloc(sizeof(int) * (size_t)G.pairs_cap);
    int* nparts = (int*)G.alloc(sizeof(int) * (size_t)G.pairs_cap);
    const char* hv = getenv("CORRFUNC_AMD_HEAVY");
    const float heavy = hv ? (float)atof(hv) : kHeavyFactor;
    const char* sp = getenv("CORRFUNC_AMD_SPLIT");
    const double split = sp ? atof(sp) : (double)kSplitPerWave;
    const double cap_div = split > 0 ? (double)blocks * ts.nwaves * split : 1.0;
    G.sched = learned ? CF_AMD_SCHED_LEARNED : (sched == CF_AMD_SCHED_NONE ? CF_AMD_SCHED_NONE
                                                                           : CF_AMD_SCHED_HEURISTIC);
    GraphScan scan;
    scan.init(std::max<int64_t>(std::max<int64_t>(nb_cap, cand_cap), std::max<int64_t>(G.pairs_cap, nmat)), G);

    // ---- the captured sequence (fill kernels only: no memset nodes)
    hipStream_t st = E.stream;
    auto fill_i = [&](int* p, int64_t n, int v) {
        k_fill_i32<<<(unsigned)std::min<int64_t>(1024, (n + 255) / 256), 256, 0, st>>>(n, p, v);
    };
    auto fill_z = [&](void* p, int64_t n) {
        k_fill_u64<<<(unsigned)std::min<int64_t>(1024, (n + 255) / 256), 256, 0, st>>>(
            n, (unsigned long long*)p, 0ull);
    };
    CF_CHECK(hipStreamSynchronize(st));
    CF_CHECK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    fill_i(G.fail, 1, 0);
    fill_i(G.cp_off, ncat + 1, 0);
    fill_z(G.tested, ncat);
    fill_z(oob, 1);
    fill_z(G.g_np, (int64_t)tot);
    fill_z(G.g_avg, (int64_t)tot);
    fill_z(G.g_w, (int64_t)tot);
    fill_z(G.ws, 2 * (int64_t)ncat);
    for (int k = 0; k < ncat; k++) {
        // extent
        if (esize == 4)
            k_minmax<float><<<512, 256, 0, st>>>(ND[k], (const float*)X[k], (const float*)Y[k], (const float*)Z[k],
                                                 mm_part);
        else
            k_minmax<double><<<512, 256, 0, st>>>(ND[k], (const double*)X[k], (const double*)Y[k],
                                                  (const double*)Z[k], mm_part);
        k_minmax_final<<<1, 64, 0, st>>>(512, mm_part, ext + 6 * k);
        // grid
        k_bg_grid<<<1, 64, 0, st>>>(k, bo, ext, ND[k], tmpl[k], G.cats, lats, G.fail);
        // counting sort into the grid
        const int64_t n = ND[k];
        const BgCat* ck = G.cats + k;
        fill_i(counts[k], nb_cap + 1, 0);
        fill_i(fillb[k], nb_cap + 1, 0);
        const unsigned nblk = (unsigned)((n + 255) / 256);
        if (n > 0) {
            if (esize == 4)
                k_bg_bucket_count<float><<<nblk, 256, 0, st>>>(n, (const float*)X[k], (const float*)Y[k],
                                                               (const float*)Z[k], ck, bucket[k], counts[k], oob);
            else
                k_bg_bucket_count<double><<<nblk, 256, 0, st>>>(n, (const double*)X[k], (const double*)Y[k],
                                                                (const double*)Z[k], ck, bucket[k], counts[k], oob);
        }
        scan.run(counts[k], tmpl[k].bstart, nb_cap, st);
        if (n > 0) {
            if (esize == 4)
                k_bg_scatter<float><<<nblk, 256, 0, st>>>(n, (const float*)X[k], (const float*)Y[k],
                                                          (const float*)Z[k], (const float*)(need_w ? W[k] : nullptr),
                                                          ck, bucket[k], tmpl[k].bstart, fillb[k], tmpl[k].x,
                                                          tmpl[k].y, tmpl[k].z, tmpl[k].w, tmpl[k].d4, tmpl[k].q);
            else
                k_bg_scatter<double><<<nblk, 256, 0, st>>>(n, (const double*)X[k], (const double*)Y[k],
                                                           (const double*)Z[k],
                                                           (const double*)(need_w ? W[k] : nullptr), ck, bucket[k],
                                                           tmpl[k].bstart, fillb[k], tmpl[k].x, tmpl[k].y,
                                                           tmpl[k].z, tmpl[k].w, tmpl[k].d4, tmpl[k].q);
        }
        k_bg_cell_bounds<<<(unsigned)((ncap + 3) / 4), 256, 0, st>>>(ck, tmpl[k].bstart, tmpl[k].x, tmpl[k].y,
                                                                      tmpl[k].z, tmpl[k].bounds);
        // neighbour-cell pairs, appended to the list at cp_off[k]
        k_bg_cellpair_flag<<<(unsigned)((cand_cap + 255) / 256), 256, 0, st>>>(ck, cand_cap, tmpl[k].bstart,
                                                                                tmpl[k].bounds, flags);
        scan.run(flags, pos, cand_cap, st);
        k_bg_cp_offset<<<1, 64, 0, st>>>(k, pos, cand_cap, G.cp_off, G.pairs_cap, G.fail);
        k_bg_cellpair_emit<<<(unsigned)std::min<int64_t>((cand_cap + 255) / 256, 1024), 256, 0, st>>>(
            ck, cand_cap, tmpl[k].bstart, tmpl[k].bounds, pos, G.cp_off, k, G.pairs_cap, cps, G.tested);
        if (need_w && n > 0) k_weight_sums<<<64, 256, 0, st>>>(n, tmpl[k].w, G.ws + 2 * k);
    }
    k_oob_flag<<<1, 64, 0, st>>>(oob, G.fail, kBgFailOob);
    // work list as launch_tuned builds it (cost order, heavy cell pairs first
    // and split), over the capacity with the device's cell-pair count
    const int* ncp_dev = G.cp_off + ncat;
    const unsigned pblk = (unsigned)(((int64_t)G.pairs_cap + 255) / 256);
    int* nheavy = nullptr;
    if (sched != CF_AMD_SCHED_NONE) {
        if (learned) {
            k_cp_features<<<pblk, 256, 0, st>>>(G.pairs_cap, cps, lats, 1, P, feats, ncp_dev);
            k_mlp_predict<<<pblk, 256, 0, st>>>(G.pairs_cap, feats, E.model.d_norm, E.model.d_params, cost, ncp_dev);
        } else {
            k_cost_heuristic<<<pblk, 256, 0, st>>>(G.pairs_cap, cps, lats, 1, cost, ncp_dev);
        }
        fill_z(total, 1);
        k_cost_total<<<std::min<unsigned>(pblk, 1024), 256, 0, st>>>(G.pairs_cap, cost, total, ncp_dev);
        k_cost_hist<<<(unsigned)nsb, 256, 0, st>>>(G.pairs_cap, cost, total, heavy, (int)nsb, bcount, ncp_dev);
        scan.run(bcount, boff, nmat, st);
        k_cost_place<<<(unsigned)nsb, 256, 0, st>>>(G.pairs_cap, cost, total, heavy, (int)nsb, boff, order,
                                                    ncp_dev);
        nheavy = boff + (int64_t)(kSortBuckets - 1) * nsb;
    }
    k_count_chunks<<<pblk, 256, 0, st>>>(G.pairs_cap, sched != CF_AMD_SCHED_NONE ? order : nullptr, cps, lats, 1,
                                         sched != CF_AMD_SCHED_NONE ? cost : nullptr, total, cap_div,
                                         sched != CF_AMD_SCHED_NONE ? nparts : nullptr, nch, ncp_dev);
    scan.run(nch, ioff, G.pairs_cap, st);
    k_bg_items_clamp<<<1, 64, 0, st>>>(ioff + G.pairs_cap, G.items_cap, G.nitems_eff, G.fail);
    k_emit_chunks<<<pblk, 256, 0, st>>>(G.pairs_cap, sched != CF_AMD_SCHED_NONE ? order : nullptr, cps, lats, 1, P,
                                        ioff, sched != CF_AMD_SCHED_NONE ? nparts : nullptr, items, ncp_dev,
                                        G.items_cap);
    fill_i(counter, kQueues * kCounterStride, 0);
    if (nheavy)
        k_single_boundary<<<1, 64, 0, st>>>(nheavy, ioff, single);
    else
        fill_i(single, 1, INT32_MAX);
    {
        const int* d_nitems = G.nitems_eff;
        unsigned long long* ticks = nullptr;
        int* cu = nullptr;
        const DevLattice* clats = lats;
        const CellPairRec* ccps = cps;
        const WorkItem* citems = items;
        const double* cedges = d_edges;
        const unsigned short* clut = d_lut;
        const int* csingle = single;
        void* kargs[] = {&P, &clats, &ccps, &citems, &d_nitems, &counter, &cedges, &clut, &G.g_np, &G.g_avg, &G.g_w,
                         &ticks, &cu, &csingle};
        (void)hipLaunchKernel(ts.kfn, dim3((unsigned)blocks), dim3(ts.threads), kargs, ts.lds, st);
    }
    const hipError_t le = hipGetLastError();
    hipGraph_t graph = nullptr;
    const hipError_t ce = hipStreamEndCapture(st, &graph);
    if (le != hipSuccess || ce != hipSuccess) {
        fprintf(stderr, "corrfunc_amd: graph capture failed: %s\n", hipGetErrorString(le != hipSuccess ? le : ce));
        if (graph) (void)hipGraphDestroy(graph);
        throw HipFailure();
    }
    G.graph = graph;
    CF_CHECK(hipGraphInstantiate(&G.exec, graph, nullptr, nullptr, 0));
    G.ncells_cap.assign(ncat, ncap);
    return EXIT_SUCCESS;
}

// Replays the graph and packs the results as cf_amd_countpairs_batch does.
static int bg_launch(BatchGraph& G, std::vector<CallOut>& outs)
{
    Engine& E = eng();
    InterruptScope interrupts;
    const double t_start = now_s();
    const int ncat = G.ncat, nslots = G.nbin;
    const size_t tot = (size_t)ncat * nslots;
    CF_CHECK(hipEventRecord(E.ev0, E.stream));
    CF_CHECK(hipGraphLaunch(G.exec, E.stream));
    CF_CHECK(hipEventRecord(E.ev1, E.stream));
    std::vector<uint64_t> np(tot);
    std::vector<double> avg(tot), wv(tot), ws(2 * (size_t)ncat), tk(ncat);
    std::vector<BgCat> cats(ncat);
    int fail = 0, ncp = 0, nitems = 0;
    CF_CHECK(hipMemcpyAsync(np.data(), G.g_np, sizeof(uint64_t) * tot, hipMemcpyDeviceToHost, E.stream));
    CF_CHECK(hipMemcpyAsync(avg.data(), G.g_avg, sizeof(double) * tot, hipMemcpyDeviceToHost, E.stream));
    CF_CHECK(hipMemcpyAsync(wv.data(), G.g_w, sizeof(double) * tot, hipMemcpyDeviceToHost, E.stream));
    CF_CHECK(hipMemcpyAsync(ws.data(), G.ws, sizeof(double) * 2 * ncat, hipMemcpyDeviceToHost, E.stream));
    CF_CHECK(hipMemcpyAsync(tk.data(), G.tested, sizeof(double) * ncat, hipMemcpyDeviceToHost, E.stream));
    CF_CHECK(hipMemcpyAsync(cats.data(), G.cats, sizeof(BgCat) * ncat, hipMemcpyDeviceToHost, E.stream));
    CF_CHECK(hipMemcpyAsync(&fail, G.fail, sizeof(int), hipMemcpyDeviceToHost, E.stream));
    CF_CHECK(hipMemcpyAsync(&ncp, G.cp_off + ncat, sizeof(int), hipMemcpyDeviceToHost, E.stream));
    CF_CHECK(hipMemcpyAsync(&nitems, G.nitems_eff, sizeof(int), hipMemcpyDeviceToHost, E.stream));
    CF_CHECK(hipStreamSynchronize(E.stream));
    check_interrupt();
    if (fail) {
        fprintf(stderr, "Error: the graph batch cannot count these positions (%s%s%s%s%s); create the graph again "
                        "for them\n",
                fail & kBgFailGrid ? "rmax >= half the box " : "", fail & kBgFailCells ? "more cells " : "",
                fail & kBgFailOob ? "positions outside the grid " : "",
                fail & kBgFailPairs ? "more cell pairs " : "", fail & kBgFailItems ? "more work items " : "");
        return EXIT_FAILURE;
    }
    float gms = 0;
    CF_CHECK(hipEventElapsedTime(&gms, E.ev0, E.ev1));
    outs.assign(ncat, CallOut());
    double tested_all = 0;
    for (int k = 0; k < ncat; k++) {
        CallOut& R = outs[k];
        R.rupp = G.rupp;
        R.nbin = G.nbin;
        R.need_w = G.need_w;
        R.np.assign(np.begin() + (size_t)k * nslots, np.begin() + (size_t)(k + 1) * nslots);
        R.avg.assign(avg.begin() + (size_t)k * nslots, avg.begin() + (size_t)(k + 1) * nslots);
        R.wavg.assign(wv.begin() + (size_t)k * nslots, wv.begin() + (size_t)(k + 1) * nslots);
        R.wsum = ws[2 * k];
        R.wsq = ws[2 * k + 1];
        tested_all += tk[k];
    }
    cf_amd_stats st;
    memset(&st, 0, sizeof st);
    st.num_cell_pairs = ncp;
    st.num_work_items = nitems;
    st.tested_pairs = tested_all;
    st.kernel_ms = gms;  // the whole replayed graph
    st.total_ms = (now_s() - t_start) * 1e3;
    st.engine = CF_AMD_ENGINE_TUNED;
    st.hist_copies = G.hist_copies;
    st.queue_fetch = G.queue_fetch;
    st.drain_pairs = G.drain_pairs;
    st.workgroup_threads = G.threads;
    st.sched = G.sched;
    if (ncat > 0) {
        st.nmesh[0] = cats[ncat - 1].g.nx; st.nmesh[1] = cats[ncat - 1].g.ny; st.nmesh[2] = cats[ncat - 1].g.nz;
        for (int d = 0; d < 3; d++) st.refine[d] = cats[ncat - 1].refine[d];
    }
    E.last = st;
    return EXIT_SUCCESS;
}

// ------------------------------------------------------------ cost model
static_assert(CF_AMD_MODEL_FEATURES == kNumFeatures && CF_AMD_MODEL_PARAMS == kNumParams &&
                  CF_AMD_MODEL_NORM == kNumNorm, "cost model sizes of corrfunc_amd.h");

static void model_alloc()
{
    CostModel& M = eng().model;
    if (M.d_params) return;
    CF_CHECK(hipMalloc(&M.d_params, sizeof(float) * kNumParams));
    CF_CHECK(hipMalloc(&M.d_norm, sizeof(float) * kNumNorm));
}

// splitmix64 (Steele, Lea & Flood 2014) and Box-Muller normals, for seeded
// weight initialisation and minibatch permutations.
struct Rng {
    uint64_t s;
    uint64_t next()
    {
        uint64_t z = (s += 0x9e3779b97f4a7c15ull);
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
        return z ^ (z >> 31);
    }
    double uniform() { return ((next() >> 11) + 0.5) * (1.0 / 9007199254740992.0); }
    double normal() { return sqrt(-2.0 * log(uniform())) * cos(6.283185307179586 * uniform()); }
};

// Standardisation of features and targets, from the samples (FP64 on the host).
static std::vector<float> model_norm(const std::vector<float>& hx, const std::vector<float>& hy, int64_t n)
{
    std::vector<float> norm(kNumNorm);
    for (int f = 0; f < kNumFeatures; f++) {
        double m = 0, v = 0;
        for (int64_t k = 0; k < n; k++) m += hx[(size_t)k * kNumFeatures + f];
        m /= (double)n;
        for (int64_t k = 0; k < n; k++) {
            const double d = hx[(size_t)k * kNumFeatures + f] - m;
            v += d * d;
        }
        const double sd = sqrt(v / (double)n);
        norm[f] = (float)m;
        norm[kNumFeatures + f] = sd > 1e-6 ? (float)(1.0 / sd) : 0.f;
    }
    double m = 0, v = 0;
    for (int64_t k = 0; k < n; k++) m += hy[k];
    m /= (double)n;
    for (int64_t k = 0; k < n; k++) v += (hy[k] - m) * (hy[k] - m);
    const double sd = sqrt(v / (double)n);
    norm[2 * kNumFeatures] = (float)m;
    norm[2 * kNumFeatures + 1] = sd > 1e-6 ? (float)sd : 1.f;
    return norm;
}

// Loss and gradient of `nb` samples (rows idx[0..nb), or 0..nb) into d_grad /
// d_loss, with the loss scaled by 1/inv_n.
static void model_grad_dev(int nb, const int* idx, const float* dX, const float* dY, const float* d_norm,
                           const float* d_params, float inv_n, float* d_grad, float* d_loss)
{
    Engine& E = eng();
    const int nblk = (nb + kGradBlock - 1) / kGradBlock;
    float* partial = (float*)E.arena.get("m_partial", sizeof(float) * (size_t)nblk * kNumParams);
    float* ploss = (float*)E.arena.get("m_ploss", sizeof(float) * (size_t)nblk);
    k_mlp_grad<<<nblk, kGradBlock, 0, E.stream>>>(nb, idx, dX, dY, d_norm, d_params, inv_n, partial, ploss);
    CF_CHECK(hipGetLastError());
    k_mlp_reduce<<<(kNumParams + 1 + 255) / 256, 256, 0, E.stream>>>(nblk, partial, ploss, d_grad, d_loss);
    CF_CHECK(hipGetLastError());
}

static int model_train_dev(const float* dX, const float* dY, int64_t n, int epochs, int batch, float lr,
                           unsigned seed, float* loss_out)
{
    Engine& E = eng();
    CostModel& M = E.model;
    if (n < 2 || epochs < 1 || batch < 1 || !(lr > 0)) {
        fprintf(stderr, "Error: cost model training needs n >= 2 samples, epochs >= 1, batch >= 1, lr > 0\n");
        return EXIT_FAILURE;
    }
    if (n > (1 << 30)) {
        fprintf(stderr, "Error: too many training samples (%lld)\n", (long long)n);
        return EXIT_FAILURE;
    }
    model_alloc();
    std::vector<float> hx((size_t)n * kNumFeatures), hy((size_t)n);
    CF_CHECK(hipMemcpyAsync(hx.data(), dX, sizeof(float) * hx.size(), hipMemcpyDeviceToHost, E.stream));
    CF_CHECK(hipMemcpyAsync(hy.data(), dY, sizeof(float) * hy.size(), hipMemcpyDeviceToHost, E.stream));
    CF_CHECK(hipStreamSynchronize(E.stream));
    const std::vector<float> norm = model_norm(hx, hy, n);
    // He-normal initialisation of the ReLU layers, zero biases
    Rng rng{0x5eedull ^ ((uint64_t)seed << 17)};
    std::vector<float> w(kNumParams, 0.f);
    for (int k = 0; k < kHidden * kNumFeatures; k++) w[kOffW1 + k] = (float)(rng.normal() * sqrt(2.0 / kNumFeatures));
    for (int k = 0; k < kHidden * kHidden; k++) w[kOffW2 + k] = (float)(rng.normal() * sqrt(2.0 / kHidden));
    for (int k = 0; k < kHidden; k++) w[kOffW3 + k] = (float)(rng.normal() * sqrt(1.0 / kHidden));
    float* d_m = (float*)E.arena.get("m_adam_m", sizeof(float) * kNumParams);
    float* d_v = (float*)E.arena.get("m_adam_v", sizeof(float) * kNumParams);
    float* d_grad = (float*)E.arena.get("m_grad", sizeof(float) * kNumParams);
    float* d_loss = (float*)E.arena.get("m_loss", sizeof(float));
    int* d_idx = (int*)E.arena.get("m_idx", sizeof(int) * (size_t)n);
    CF_CHECK(hipMemcpyAsync(M.d_params, w.data(), sizeof(float) * kNumParams, hipMemcpyHostToDevice, E.stream));
    CF_CHECK(hipMemcpyAsync(M.d_norm, norm.data(), sizeof(float) * kNumNorm, hipMemcpyHostToDevice, E.stream));
    k_fill_i32<<<8, 256, 0, E.stream>>>(kNumParams, (int*)d_m, 0);  // +0.0f
    k_fill_i32<<<8, 256, 0, E.stream>>>(kNumParams, (int*)d_v, 0);
    std::vector<int> perm((size_t)n);
    for (int64_t k = 0; k < n; k++) perm[k] = (int)k;
    int step = 0;
    for (int e = 0; e < epochs; e++) {
        for (int64_t k = n - 1; k > 0; k--) std::swap(perm[k], perm[rng.next() % (uint64_t)(k + 1)]);
        CF_CHECK(hipMemcpyAsync(d_idx, perm.data(), sizeof(int) * (size_t)n, hipMemcpyHostToDevice, E.stream));
        for (int64_t b0 = 0; b0 < n; b0 += batch) {
            const int nb = (int)std::min<int64_t>(batch, n - b0);
            model_grad_dev(nb, d_idx + b0, dX, dY, M.d_norm, M.d_params, 1.f / nb, d_grad, d_loss);
            k_adam<<<(kNumParams + 255) / 256, 256, 0, E.stream>>>(M.d_params, d_grad, d_m, d_v,
                                                                  adam_consts(++step, lr));
            CF_CHECK(hipGetLastError());
        }
        CF_CHECK(hipStreamSynchronize(E.stream));  // perm is rewritten next epoch
    }
    // loss over all samples, in chunks
    double total = 0;
    const int64_t chunk = 1 << 16;
    for (int64_t b0 = 0; b0 < n; b0 += chunk) {
        const int nb = (int)std::min<int64_t>(chunk, n - b0);
        model_grad_dev(nb, nullptr, dX + b0 * kNumFeatures, dY + b0, M.d_norm, M.d_params, (float)(1.0 / n), d_grad,
                       d_loss);
        float l = 0;
        CF_CHECK(hipMemcpyAsync(&l, d_loss, sizeof(float), hipMemcpyDeviceToHost, E.stream));
        CF_CHECK(hipStreamSynchronize(E.stream));
        total += l;
    }
    if (loss_out) *loss_out = (float)total;
    M.trained = true;
    return EXIT_SUCCESS;
}

static int check_common(const struct config_options* o)
{
    if (!(o->float_type == sizeof(float) || o->float_type == sizeof(double))) {
        fprintf(stderr, "ERROR: In %s> Can only handle doubles or floats. Got an array of size = %zu\n", "corrfunc_amd",
                o->float_type);
        return EXIT_FAILURE;
    }
    if (strncmp(o->version, CORRFUNC_API_VERSION, sizeof(o->version) - 1) != 0 &&
        strncmp(o->version, "2.5.3", sizeof(o->version) - 1) != 0) {
        fprintf(stderr, "Error: Do not know this API version = `%s'. Expected version = `%s'\n", o->version,
                CORRFUNC_API_VERSION);
        return EXIT_FAILURE;
    }
    return EXIT_SUCCESS;
}

static void fix_refine(struct config_options* o)
{
    for (int i = 0; i < 3; i++) {
        if (o->bin_refine_factors[i] < 1) {
            fprintf(stderr, "Warning: bin refine factor along axis = %d *must* be >=1. Instead found bin refine factor =%d\n",
                    i, o->bin_refine_factors[i]);
            reset_bin_refine_factors(o);
            break;
        }
    }
}

static void fix_maxcells(struct config_options* o)
{
    if (o->max_cells_per_dim == 0) {
        fprintf(stderr, "Warning: Max. cells per dimension is set to 0 - resetting to `NLATMAX' = %d\n", NLATMAX);
        o->max_cells_per_dim = NLATMAX;
    }
}

// autocorr doubling, self pairs, averages (countpairs_impl.c.src:608-664)
static void finalize(CallOut& R, int autocorr, int64_t ND1, int need_avg, int self_slot)
{
    const size_t nb = R.np.size();
    if (autocorr == 1) {
        for (size_t i = 0; i < nb; i++) {
            R.np[i] *= 2;
            if (need_avg) R.avg[i] *= 2.0;
            if (R.need_w) R.wavg[i] *= 2.0;
        }
        if (R.rupp[0] <= 0.0) {
            R.np[self_slot] += (uint64_t)ND1;
            if (R.need_w) R.wavg[1] += R.wsq;
        }
    }
    for (size_t i = 0; i < nb; i++) {
        if (R.np[i] > 0) {
            if (need_avg) R.avg[i] /= (double)R.np[i];
            if (R.need_w) R.wavg[i] /= (double)R.np[i];
        }
    }
}

// Every entry point runs under one process-wide lock: the engine's device
// buffers, stream and host copies are shared by all calls, so calls from
// several threads are serialised (the reference's functions are not
// re-entrant either; its Python extension holds the GIL's place with the
// caller's). No exception leaves an extern "C" function.
static std::recursive_mutex& engine_lock()
{
    static std::recursive_mutex m;
    return m;
}

template <typename F>
static int guarded(F&& f)
{
    std::lock_guard<std::recursive_mutex> hold(engine_lock());
    try {
        return f();
    } catch (const HipFailure&) {
        return EXIT_FAILURE;
    } catch (const InputError&) {
        return EXIT_FAILURE;
    } catch (const std::bad_alloc&) {
        fprintf(stderr, "corrfunc_amd: out of host memory\n");
        return EXIT_FAILURE;
    } catch (const std::exception& ex) {
        fprintf(stderr, "corrfunc_amd: %s\n", ex.what());
        return EXIT_FAILURE;
    }
}


// ------------------------------------------------------------ on-sky inputs
// (RA, DEC, cz or comoving distance) -> Cartesian FP64 on the host, with the
// reference's steps and FP64 operation order (countpairs_rp_pi_mocks_impl.c.src:
// 43-101 input fixes, 309-390 distances and positions; utils/set_cosmo_dist.c:
// 25-71 the distance table; utils/cosmology_params.c:26-56 the cosmologies).
constexpr double kSpeedOfLight = 299800.0;  // km/s (utils/set_cosmo_dist.h:19)
constexpr double kPiOver180 = 0.017453292519943295769236907684886127134428718885417254560971;

// The caller's arrays are fixed in place, in their own precision, as the reference does.
template <typename T>
static int sky_fix_inputs(int64_t n, T* ra, T* dec, T* cz, int is_comoving)
{
    if (n == 0) return EXIT_SUCCESS;
    if (!ra || !dec || !cz) {
        fprintf(stderr, "Input arrays can not be NULL. Have RA = %p DEC = %p cz = %p\n", (void*)ra, (void*)dec,
                (void*)cz);
        return EXIT_FAILURE;
    }
    bool fix_ra = false, fix_dec = false;
    T max_cz = 0;
    for (int64_t i = 0; i < n; i++) {
        if (cz[i] > max_cz) max_cz = cz[i];
        if (ra[i] < 0) fix_ra = true;
        if (dec[i] > 90) fix_dec = true;
        if (dec[i] > 180) {
            fprintf(stderr, "theta[%lld] = %g should be less than 180 deg\n", (long long)i, (double)dec[i]);
            return EXIT_FAILURE;
        }
    }
    const bool fix_cz = max_cz < 10 && !is_comoving;
    if (fix_ra) fprintf(stderr, "Out of range values found for ra. Expected ra in [0.0,360.0], found [-180,180] -- fixing that\n");
    if (fix_dec) fprintf(stderr, "Out of range values found for dec. Expected dec in [-90.0,90.0], found [0,180] -- fixing that\n");
    if (fix_cz) fprintf(stderr, "Out of range values found for cz. Expected `cz' but found `z' (max = %g) -- multiplying by c\n", (double)max_cz);
    for (int64_t i = 0; i < n && (fix_ra || fix_dec || fix_cz); i++) {
        if (fix_ra) ra[i] += (T)180.0;
        if (fix_dec) dec[i] -= (T)90.0;
        if (fix_cz) cz[i] *= (T)kSpeedOfLight;
    }
    return EXIT_SUCCESS;
}

static int sky_fix(int64_t n, void* ra, void* dec, void* cz, size_t esize, int is_comoving)
{
    return esize == 4 ? sky_fix_inputs(n, (float*)ra, (float*)dec, (float*)cz, is_comoving)
                      : sky_fix_inputs(n, (double*)ra, (double*)dec, (double*)cz, is_comoving);
}

static int sky_cosmology(int which, double* om, double* ol)
{
    if (which == 1) *om = 0.25;        // LasDamas
    else if (which == 2) *om = 0.302;  // Planck
    else {
        fprintf(stderr, "ERROR: Cosmology=%d not implemented\n", which);
        return EXIT_FAILURE;
    }
    *ol = 1.0 - *om;
    return EXIT_SUCCESS;
}

// Comoving distance (Mpc/h) at z = k/max_size, k = 1, 2, ..., up to zmax: Simpson's
// rule over steps of 1e-2/max_size of 1/E(z), flat FRW.
static int sky_distance_table(double zmax, int max_size, double om, double ol, std::vector<double>& zc,
                              std::vector<double>& dc)
{
    zc.clear();
    dc.clear();
    const double ok = 1.0 - om - ol;
    const double dh = kSpeedOfLight * 0.01 / 1.0;
    const double step = 1.0 / max_size;
    const double dz = 1e-2 * step;
    double acc = 0.0, e2 = 1.0, znext = step;
    for (double z = 2.0 * dz; z < zmax; z += 2.0 * dz) {
        const double e0 = e2;
        const double zm = 1 + z - dz, zh = 1 + z;
        const double e1 = 1.0 / sqrt(om * (zm * zm * zm) + ok * zm + ol);
        e2 = 1.0 / sqrt(om * (zh * zh * zh) + ok * zh + ol);
        acc += dz * (e0 + 4. * e1 + e2) / 3.;
        if (z > (znext - 1e-10) && z < (znext + 1e-10)) {
            if ((int)zc.size() >= max_size) break;
            zc.push_back(z);
            dc.push_back(acc * dh);
            znext += step;
        }
    }
    return (int)zc.size();
}

// Piecewise-linear interpolation in the table (strictly increasing xa); false
// outside it. The interval is the last lo <= n-2 with xa[lo] <= x (what a
// bisection finds), reached from the guess of an evenly spaced table by a walk:
// the redshift table is spaced evenly to ~1e-10, so the walk is 0-1 steps.
static bool sky_interp(const std::vector<double>& xa, const std::vector<double>& ya, double x, double* y)
{
    const size_t n = xa.size();
    if (n < 2 || !(x >= xa[0] && x <= xa[n - 1])) return false;
    const double span = xa[n - 1] - xa[0];
    size_t lo = span > 0.0 ? (size_t)((x - xa[0]) / span * (double)(n - 1)) : 0;
    if (lo > n - 2) lo = n - 2;
    while (lo > 0 && xa[lo] > x) lo--;
    while (lo + 1 <= n - 2 && xa[lo + 1] <= x) lo++;
    const double w = xa[lo + 1] - xa[lo];
    *y = w > 0.0 ? ya[lo] + (x - xa[lo]) / w * (ya[lo + 1] - ya[lo]) : 0.0;
    return true;
}

static double sky_value(const void* p, size_t esize, int64_t i)
{
    return esize == 4 ? (double)((const float*)p)[i] : ((const double*)p)[i];
}

struct SkySet {
    std::vector<double> x, y, z, w;
};

// Both sets' Cartesian positions (and FP64 weights).
static int sky_positions(int64_t n1, const void* ra1, const void* dec1, const void* cz1, const void* w1, int64_t n2,
                         const void* ra2, const void* dec2, const void* cz2, const void* w2, int autocorr,
                         size_t esize, int is_comoving, int cosmology, int nthreads, SkySet& s1, SkySet& s2)
{
    double om, ol;
    if (sky_cosmology(cosmology, &om, &ol) != EXIT_SUCCESS) return EXIT_FAILURE;
    std::vector<double> zc, dc;
    const double inv_c = 1.0 / kSpeedOfLight;
    if (!is_comoving) {
        double czmax = 0.0;
        for (int64_t i = 0; i < n1; i++)
            if (sky_value(cz1, esize, i) > czmax) czmax = sky_value(cz1, esize, i);
        if (!autocorr)
            for (int64_t i = 0; i < n2; i++)
                if (sky_value(cz2, esize, i) > czmax) czmax = sky_value(cz2, esize, i);
        sky_distance_table(czmax * inv_c + 0.01, 10000, om, ol, zc, dc);
    }
    for (int set = 0; set < (autocorr ? 1 : 2); set++) {
        const int64_t n = set ? n2 : n1;
        const void *ra = set ? ra2 : ra1, *dec = set ? dec2 : dec1, *cz = set ? cz2 : cz1, *w = set ? w2 : w1;
        SkySet& S = set ? s2 : s1;
        S.x.resize(n);
        S.y.resize(n);
        S.z.resize(n);
        if (w) S.w.resize(n);
        else S.w.clear();  // no weights: the caller's size check must not see an earlier call's
        // element-wise, so the result does not depend on the split across host threads
        std::atomic<int64_t> bad{-1};
        auto convert = [&](int64_t lo, int64_t hi) {
            for (int64_t i = lo; i < hi; i++) {
                double d = sky_value(cz, esize, i);
                if (!is_comoving && !sky_interp(zc, dc, d * inv_c, &d)) {
                    bad = i;
                    return;
                }
                const double de = sky_value(dec, esize, i), ra_i = sky_value(ra, esize, i);
                const double cd = cos(de * kPiOver180);
                S.x[i] = d * cd * cos(ra_i * kPiOver180);
                S.y[i] = d * cd * sin(ra_i * kPiOver180);
                S.z[i] = d * sin(de * kPiOver180);
                if (w) S.w[i] = sky_value(w, esize, i);
            }
        };
        const int nt = (int)std::max<int64_t>(1, std::min<int64_t>(nthreads, n / 65536));
        std::vector<std::thread> pool;
        for (int t = 1; t < nt; t++) pool.emplace_back(convert, n * t / nt, n * (t + 1) / nt);
        convert(0, n / nt);
        for (auto& th : pool) th.join();
        if (bad >= 0) {
            fprintf(stderr, "Error: cz = %g is outside the tabulated redshift range [%g, %g] of the distance "
                            "conversion\n", sky_value(cz, esize, bad), zc.empty() ? 0.0 : zc.front() * kSpeedOfLight,
                    zc.empty() ? 0.0 : zc.back() * kSpeedOfLight);
            return EXIT_FAILURE;
        }
    }
    return EXIT_SUCCESS;
}

// Shared driver of the two on-sky entry points: fixes, converts, counts, doubles
// and averages (countpairs_rp_pi_mocks_impl.c.src:200-847, s_mu_mocks_impl:200-850).
static int sky_run(int mode, int64_t ND1, void* RA1, void* DEC1, void* CZ1, int64_t ND2, void* RA2, void* DEC2,
                   void* CZ2, int numthreads, int autocorr, const char* binfile, double pimax, double mu_max,
                   int nmu_bins, int cosmology, struct config_options* o, struct extra_options* extra, CallOut& R)
{
    if (autocorr != 0 && autocorr != 1) {
        fprintf(stderr, "Error: autocorr = %d must be 0 or 1\n", autocorr);
        return EXIT_FAILURE;
    }
    const size_t esize = o->float_type;
    if (sky_fix(ND1, RA1, DEC1, CZ1, esize, o->is_comoving_dist) != EXIT_SUCCESS) return EXIT_FAILURE;
    if (!autocorr && sky_fix(ND2, RA2, DEC2, CZ2, esize, o->is_comoving_dist) != EXIT_SUCCESS) return EXIT_FAILURE;
    double om, ol;
    if (sky_cosmology(cosmology, &om, &ol) != EXIT_SUCCESS) return EXIT_FAILURE;
    const int need_w = extra->weight_method != NONE;
    // kept across calls (resized, not re-allocated and re-zeroed, when a later
    // call has the same or fewer points); like the device arena, one per process
    static SkySet keep[2];
    SkySet &s1 = keep[0], &s2 = keep[1];
    if (sky_positions(ND1, RA1, DEC1, CZ1, need_w ? extra->weights0.weights[0] : nullptr, ND2, RA2, DEC2, CZ2,
                      need_w && !autocorr ? extra->weights1.weights[0] : nullptr, autocorr, esize,
                      o->is_comoving_dist, cosmology, numthreads, s1, s2) != EXIT_SUCCESS)
        return EXIT_FAILURE;
    if (need_w && (s1.w.size() != (size_t)ND1 || (!autocorr && s2.w.size() != (size_t)ND2))) {
        fprintf(stderr, "Error: weight_method requires weights arrays\n");
        return EXIT_FAILURE;
    }
    // the counting runs in FP64 on the converted positions
    struct extra_options ex = *extra;
    if (need_w) {
        ex.weights0.weights[0] = s1.w.data();
        ex.weights1.weights[0] = autocorr ? s1.w.data() : s2.w.data();
    }
    const size_t saved = o->float_type;
    o->float_type = sizeof(double);
    // N2 as given: the boost rule takes the larger of the two sets' densities
    CallSpec c{mode, ND1, ND2, s1.x.data(), s1.y.data(), s1.z.data(),
               autocorr ? nullptr : s2.x.data(), autocorr ? nullptr : s2.y.data(), autocorr ? nullptr : s2.z.data(),
               autocorr, binfile, pimax, mu_max, nmu_bins, 0.0, numthreads};
    const int st = run_core(c, o, &ex, R);
    o->float_type = saved;
    return st;
}

}  // namespace cfamd

using namespace cfamd;

// ===================================================================== C ABI
extern "C" {

void free_results(results_countpairs* r)
{
    if (!r) return;
    free(r->rupp); r->rupp = NULL;
    free(r->npairs); r->npairs = NULL;
    free(r->rpavg); r->rpavg = NULL;
    free(r->weightavg); r->weightavg = NULL;
    r->nbin = 0;
}

void free_results_rp_pi(results_countpairs_rp_pi* r)
{
    if (!r) return;
    free(r->rupp); r->rupp = NULL;
    free(r->npairs); r->npairs = NULL;
    free(r->rpavg); r->rpavg = NULL;
    free(r->weightavg); r->weightavg = NULL;
    r->nbin = 0;
    r->npibin = 0;
}

void free_results_s_mu(results_countpairs_s_mu* r)
{
    if (!r) return;
    free(r->supp); r->supp = NULL;
    free(r->npairs); r->npairs = NULL;
    free(r->savg); r->savg = NULL;
    free(r->weightavg); r->weightavg = NULL;
    r->nsbin = 0;
    r->nmu_bins = 0;
}

void free_results_xi(results_countpairs_xi* r)
{
    if (!r) return;
    free(r->rupp); r->rupp = NULL;
    free(r->xi); r->xi = NULL;
    free(r->npairs); r->npairs = NULL;
    free(r->ravg); r->ravg = NULL;
    free(r->weightavg); r->weightavg = NULL;
    r->nbin = 0;
}

void free_results_wp(results_countpairs_wp* r)
{
    if (!r) return;
    free(r->rupp); r->rupp = NULL;
    free(r->wp); r->wp = NULL;
    free(r->npairs); r->npairs = NULL;
    free(r->rpavg); r->rpavg = NULL;
    free(r->weightavg); r->weightavg = NULL;
    r->nbin = 0;
}

int countpairs(const int64_t ND1, void* X1, void* Y1, void* Z1, const int64_t ND2, void* X2, void* Y2, void* Z2,
               const int numthreads, const int autocorr, const char* binfile, results_countpairs* results,
               struct config_options* options, struct extra_options* extra)
{
    if (check_common(options) != EXIT_SUCCESS) return EXIT_FAILURE;
    return guarded([&]() -> int {
        struct extra_options dummy;
        if (extra == NULL) { dummy = get_extra_options(NONE); extra = &dummy; }
        const double t0 = now_s();
        fix_maxcells(options);
        fix_refine(options);
        options->sort_on_z = 1;
        CallSpec c{MODE_DD, ND1, ND2, X1, Y1, Z1, X2, Y2, Z2, autocorr, binfile, 0, 0, 0, 0, numthreads};
        CallOut R;
        if (run_core(c, options, extra, R) != EXIT_SUCCESS) return EXIT_FAILURE;
        finalize(R, autocorr, ND1, options->need_avg_sep, 1);
        const int nbin = R.nbin;
        results->nbin = nbin;
        results->npairs = (uint64_t*)malloc(sizeof(uint64_t) * nbin);
        results->rupp = (double*)malloc(sizeof(double) * nbin);
        results->rpavg = (double*)calloc(nbin, sizeof(double));
        results->weightavg = (double*)calloc(nbin, sizeof(double));
        if (!results->npairs || !results->rupp || !results->rpavg || !results->weightavg) {
            free_results(results);
            return EXIT_FAILURE;
        }
        for (int i = 0; i < nbin; i++) {
            results->npairs[i] = R.np[i];
            results->rupp[i] = R.rupp[i];
            results->rpavg[i] = options->need_avg_sep ? R.avg[i] : 0.0;
            results->weightavg[i] = R.need_w ? R.wavg[i] : 0.0;
        }
        reset_bin_refine_factors(options);
        if (options->c_api_timer) options->c_api_time = now_s() - t0;
        return EXIT_SUCCESS;
    });
}

// countpairs() with caller-owned results and no device allocation: every
// buffer the call's kernels use lives in the engine's fixed scratch array
// (kScratchBytes per GPU). Calls that overlap on one GPU race on it; the
// engine's lock serialises calls from threads of one process.
int cf_amd_countpairs_noalloc(const int64_t ND1, void* X1, void* Y1, void* Z1, const int64_t ND2, void* X2, void* Y2,
                              void* Z2, const int numthreads, const int autocorr, const char* binfile,
                              results_countpairs* results, struct config_options* options,
                              struct extra_options* extra)
{
    if (check_common(options) != EXIT_SUCCESS) return EXIT_FAILURE;
    return guarded([&]() -> int {
        Engine& E = eng();
        int dev = -1;
        CF_CHECK(hipGetDevice(&dev));
        if (dev != E.device || E.scratch_ptr == nullptr) {
            fprintf(stderr, "Error: cf_amd_countpairs_noalloc runs on the engine's device %d (current %d) with its "
                            "scratch array\n", E.device, dev);
            return EXIT_FAILURE;
        }
        if (results == NULL || results->npairs == NULL || results->rupp == NULL || results->rpavg == NULL ||
            results->weightavg == NULL) {
            fprintf(stderr, "Error: cf_amd_countpairs_noalloc writes into the caller's results arrays "
                            "(npairs, rupp, rpavg, weightavg of results->nbin entries)\n");
            return EXIT_FAILURE;
        }
        struct extra_options dummy;
        if (extra == NULL) { dummy = get_extra_options(NONE); extra = &dummy; }
        const int need_w = extra->weight_method != NONE;
        const void* ins[8] = {X1, Y1, Z1, need_w ? extra->weights0.weights[0] : nullptr, nullptr, nullptr, nullptr,
                              nullptr};
        if (!autocorr) {
            ins[4] = X2; ins[5] = Y2; ins[6] = Z2;
            ins[7] = need_w ? extra->weights1.weights[0] : nullptr;
        }
        for (int k = 0; k < 8; k++) {
            const int64_t n = k < 4 ? ND1 : ND2;
            if (ins[k] && n > 0 && !is_device_ptr(ins[k])) {
                fprintf(stderr, "Error: cf_amd_countpairs_noalloc takes positions and weights in device memory\n");
                return EXIT_FAILURE;
            }
        }
        if (E.model.recording) {
            fprintf(stderr, "Error: cost recording allocates its sample store; turn it off for "
                            "cf_amd_countpairs_noalloc\n");
            return EXIT_FAILURE;
        }
        const int cap = results->nbin;
        const double t0 = now_s();
        fix_maxcells(options);
        fix_refine(options);
        options->sort_on_z = 1;
        CallSpec c{MODE_DD, ND1, ND2, X1, Y1, Z1, X2, Y2, Z2, autocorr, binfile, 0, 0, 0, 0, numthreads};
        CallOut R;
        // the autotuner's candidate runs are left out: they would grow the scratch use
        const bool autotune = E.autotune;
        E.autotune = false;
        E.arena.begin_fixed(E.scratch_ptr, kScratchBytes);
        int rc = EXIT_FAILURE;
        try {
            rc = run_core(c, options, extra, R);
        } catch (...) {
            E.arena.end_fixed();
            E.autotune = autotune;
            throw;
        }
        E.arena.end_fixed();
        E.autotune = autotune;
        if (rc != EXIT_SUCCESS) return EXIT_FAILURE;
        finalize(R, autocorr, ND1, options->need_avg_sep, 1);
        const int nbin = R.nbin;
        if (nbin > cap) {
            fprintf(stderr, "Error: %d bin edges do not fit the caller's results arrays of %d entries\n", nbin, cap);
            return EXIT_FAILURE;
        }
        results->nbin = nbin;
        for (int i = 0; i < nbin; i++) {
            results->npairs[i] = R.np[i];
            results->rupp[i] = R.rupp[i];
            results->rpavg[i] = options->need_avg_sep ? R.avg[i] : 0.0;
            results->weightavg[i] = R.need_w ? R.wavg[i] : 0.0;
        }
        reset_bin_refine_factors(options);
        if (options->c_api_timer) options->c_api_time = now_s() - t0;
        return EXIT_SUCCESS;
    });
}

int countpairs_rp_pi(const int64_t ND1, void* X1, void* Y1, void* Z1, const int64_t ND2, void* X2, void* Y2, void* Z2,
                     const int numthreads, const int autocorr, const char* binfile, const double pimax,
                     results_countpairs_rp_pi* results, struct config_options* options, struct extra_options* extra)
{
    if (check_common(options) != EXIT_SUCCESS) return EXIT_FAILURE;
    return guarded([&]() -> int {
        struct extra_options dummy;
        if (extra == NULL) { dummy = get_extra_options(NONE); extra = &dummy; }
        const double t0 = now_s();
        options->sort_on_z = 1;
        fix_refine(options);
        fix_maxcells(options);
        CallSpec c{MODE_RPPI, ND1, ND2, X1, Y1, Z1, X2, Y2, Z2, autocorr, binfile, pimax, 0, 0, 0, numthreads};
        CallOut R;
        if (run_core(c, options, extra, R) != EXIT_SUCCESS) return EXIT_FAILURE;
        const int npibin = R.n2;
        finalize(R, autocorr, ND1, options->need_avg_sep, npibin + 1);
        const int nbin = R.nbin;
        const int64_t tot = (int64_t)(npibin + 1) * (nbin + 1);
        results->nbin = nbin;
        results->npibin = npibin;
        results->pimax = pimax;
        results->npairs = (uint64_t*)calloc(tot, sizeof(uint64_t));
        results->rupp = (double*)malloc(sizeof(double) * nbin);
        results->rpavg = (double*)calloc(tot, sizeof(double));
        results->weightavg = (double*)calloc(tot, sizeof(double));
        if (!results->npairs || !results->rupp || !results->rpavg || !results->weightavg) {
            free_results_rp_pi(results);
            return EXIT_FAILURE;
        }
        for (int i = 0; i < nbin; i++) {
            results->rupp[i] = R.rupp[i];
            for (int j = 0; j < npibin; j++) {
                const int64_t idx = (int64_t)i * (npibin + 1) + j;
                results->npairs[idx] = R.np[idx];
                results->rpavg[idx] = options->need_avg_sep ? R.avg[idx] : 0.0;
                results->weightavg[idx] = R.need_w ? R.wavg[idx] : 0.0;
            }
        }
        reset_bin_refine_factors(options);
        if (options->c_api_timer) options->c_api_time = now_s() - t0;
        return EXIT_SUCCESS;
    });
}

int countpairs_s_mu(const int64_t ND1, void* X1, void* Y1, void* Z1, const int64_t ND2, void* X2, void* Y2, void* Z2,
                    const int numthreads, const int autocorr, const char* sbinfile, const double mu_max,
                    const int nmu_bins, results_countpairs_s_mu* results, struct config_options* options,
                    struct extra_options* extra)
{
    if (check_common(options) != EXIT_SUCCESS) return EXIT_FAILURE;
    return guarded([&]() -> int {
        struct extra_options dummy;
        if (extra == NULL) { dummy = get_extra_options(NONE); extra = &dummy; }
        const double t0 = now_s();
        options->sort_on_z = 1;
        fix_refine(options);
        fix_maxcells(options);
        if (options->fast_divide_and_NR_steps >= MAX_FAST_DIVIDE_NR_STEPS) options->fast_divide_and_NR_steps = 0;
        CallSpec c{MODE_SMU, ND1, ND2, X1, Y1, Z1, X2, Y2, Z2, autocorr, sbinfile, 0, mu_max, nmu_bins, 0, numthreads};
        CallOut R;
        if (run_core(c, options, extra, R) != EXIT_SUCCESS) return EXIT_FAILURE;
        finalize(R, autocorr, ND1, options->need_avg_sep, nmu_bins + 1);
        const int nsbin = R.nbin;
        const int64_t tot = (int64_t)(nmu_bins + 1) * (nsbin + 1);
        results->nsbin = nsbin;
        results->nmu_bins = nmu_bins;
        results->mu_max = mu_max;
        results->mu_min = 0.0;
        results->npairs = (uint64_t*)calloc(tot, sizeof(uint64_t));
        results->supp = (double*)malloc(sizeof(double) * nsbin);
        results->savg = (double*)calloc(tot, sizeof(double));
        results->weightavg = (double*)calloc(tot, sizeof(double));
        if (!results->npairs || !results->supp || !results->savg || !results->weightavg) {
            free_results_s_mu(results);
            return EXIT_FAILURE;
        }
        for (int i = 0; i < nsbin; i++) {
            results->supp[i] = R.rupp[i];
            for (int j = 0; j < nmu_bins; j++) {
                const int64_t idx = (int64_t)i * (nmu_bins + 1) + j;
                results->npairs[idx] = R.np[idx];
                results->savg[idx] = options->need_avg_sep ? R.avg[idx] : 0.0;
                results->weightavg[idx] = R.need_w ? R.wavg[idx] : 0.0;
            }
        }
        reset_bin_refine_factors(options);
        if (options->c_api_timer) options->c_api_time = now_s() - t0;
        return EXIT_SUCCESS;
    });
}

static int xi_wp_run(int mode, const int64_t ND, void* X, void* Y, void* Z, double boxsize, int numthreads,
                     const char* binfile, double pimax, struct config_options* options, struct extra_options*& extra,
                     struct extra_options* dummy, CallOut& R)
{
    if (extra == NULL) { *dummy = get_extra_options(NONE); extra = dummy; }
    options->periodic = 1;
    options->autocorr = 1;
    options->sort_on_z = 1;
    fix_maxcells(options);
    fix_refine(options);
    CallSpec c{mode, ND, 0, X, Y, Z, nullptr, nullptr, nullptr, 1, binfile, pimax, 0, 0, boxsize, numthreads};
    c.cell_timer = mode == MODE_WP && options->c_cell_timer;
    if (run_core(c, options, extra, R) != EXIT_SUCCESS) return EXIT_FAILURE;
    finalize(R, 1, ND, options->need_avg_sep, 1);
    if (c.cell_timer) {
        // the reference's allocate_cell_timer (utils/defs.h:448-462): an existing
        // array of at least num_cell_pairs rows is reused, else replaced
        const int64_t ncp = (int64_t)R.timings.size();
        if (options->totncells_timings < ncp) {
            free_cell_timings(options);
            options->cell_timings = (struct api_cell_timings*)calloc(ncp, sizeof(struct api_cell_timings));
            if (options->cell_timings == NULL)
                fprintf(stderr, "Warning: In %s> Could not allocate memory to store the API timings per cell. \n",
                        "countpairs_wp");
            else
                options->totncells_timings = ncp;
        }
        if (options->cell_timings != NULL && options->totncells_timings >= ncp && ncp > 0)
            memcpy(options->cell_timings, R.timings.data(), sizeof(struct api_cell_timings) * ncp);
    }
    return EXIT_SUCCESS;
}

int countpairs_xi(const int64_t ND1, void* X1, void* Y1, void* Z1, const double boxsize, const int numthreads,
                  const char* binfile, results_countpairs_xi* results, struct config_options* options,
                  struct extra_options* extra)
{
    if (check_common(options) != EXIT_SUCCESS) return EXIT_FAILURE;
    return guarded([&]() -> int {
        const double t0 = now_s();
        struct extra_options dummy;
        CallOut R;
        if (xi_wp_run(MODE_XI, ND1, X1, Y1, Z1, boxsize, numthreads, binfile, 0, options, extra, &dummy, R))
            return EXIT_FAILURE;
        const int nbin = R.nbin;
        results->nbin = nbin;
        results->npairs = (uint64_t*)malloc(sizeof(uint64_t) * nbin);
        results->xi = (double*)malloc(sizeof(double) * nbin);
        results->rupp = (double*)malloc(sizeof(double) * nbin);
        results->ravg = (double*)malloc(sizeof(double) * nbin);
        results->weightavg = (double*)calloc(nbin, sizeof(double));
        if (!results->npairs || !results->xi || !results->rupp || !results->ravg || !results->weightavg) {
            free_results_xi(results);
            return EXIT_FAILURE;
        }
        // analytic RR (countpairs_xi_impl.c.src:581-628)
        double wsum = (double)ND1, wsq = (double)ND1;
        if (R.need_w) {
            wsum = R.wsum;
            wsq = (double)ND1 + R.wsq;
        }
        const double prefac = wsum * (wsum - wsum / ND1) / (boxsize * boxsize * boxsize);
        double rlow = 0.0;
        for (int i = 0; i < nbin; i++) {
            results->npairs[i] = R.np[i];
            results->rupp[i] = R.rupp[i];
            results->ravg[i] = options->need_avg_sep ? R.avg[i] : 0.0;
            results->weightavg[i] = R.need_w ? R.wavg[i] : 0.0;
            double weight0 = (double)results->npairs[i];
            if (R.need_w) weight0 *= results->weightavg[i];
            const double vol = 4.0 / 3.0 * M_PI * (R.rupp[i] * R.rupp[i] * R.rupp[i] - rlow * rlow * rlow);
            if (vol > 0.0) {
                double wr = prefac * vol;
                if (rlow <= 0.) wr += wsq;
                results->xi[i] = weight0 / wr - 1.0;
            } else {
                results->xi[i] = -2.0;
            }
            rlow = results->rupp[i];
        }
        reset_bin_refine_factors(options);
        if (options->c_api_timer) options->c_api_time = now_s() - t0;
        return EXIT_SUCCESS;
    });
}

int countpairs_wp(const int64_t ND1, void* X1, void* Y1, void* Z1, const double boxsize, const int numthreads,
                  const char* binfile, const double pimax, results_countpairs_wp* results,
                  struct config_options* options, struct extra_options* extra)
{
    if (check_common(options) != EXIT_SUCCESS) return EXIT_FAILURE;
    return guarded([&]() -> int {
        const double t0 = now_s();
        struct extra_options dummy;
        CallOut R;
        if (xi_wp_run(MODE_WP, ND1, X1, Y1, Z1, boxsize, numthreads, binfile, pimax, options, extra, &dummy, R))
            return EXIT_FAILURE;
        const int nbin = R.nbin;
        results->nbin = nbin;
        results->pimax = pimax;
        results->npairs = (uint64_t*)malloc(sizeof(uint64_t) * nbin);
        results->wp = (double*)malloc(sizeof(double) * nbin);
        results->rupp = (double*)malloc(sizeof(double) * nbin);
        results->rpavg = (double*)malloc(sizeof(double) * nbin);
        results->weightavg = (double*)calloc(nbin, sizeof(double));
        if (!results->npairs || !results->wp || !results->rupp || !results->rpavg || !results->weightavg) {
            free_results_wp(results);
            return EXIT_FAILURE;
        }
        // analytic RR (countpairs_wp_impl.c.src:619-659)
        double wsum = (double)ND1, wsq = (double)ND1;
        if (R.need_w) {
            wsum = R.wsum;
            wsq = (double)ND1 + R.wsq;
        }
        const double prefac = wsum * (wsum - wsum / ND1) / (boxsize * boxsize * boxsize);
        double rlow = 0.0;
        const double twice_pimax = 2.0 * pimax;
        for (int i = 0; i < nbin; i++) {
            results->npairs[i] = R.np[i];
            results->rupp[i] = R.rupp[i];
            results->rpavg[i] = options->need_avg_sep ? R.avg[i] : 0.0;
            results->weightavg[i] = R.need_w ? R.wavg[i] : 0.0;
            double weight0 = (double)results->npairs[i];
            if (R.need_w) weight0 *= results->weightavg[i];
            const double vol = M_PI * (results->rupp[i] * results->rupp[i] - rlow * rlow) * twice_pimax;
            if (vol > 0.0) {
                double wr = prefac * vol;
                if (rlow <= 0.) wr += wsq;
                results->wp[i] = (weight0 / wr - 1) * twice_pimax;
            } else {
                results->wp[i] = -2.0 * twice_pimax;
            }
            rlow = results->rupp[i];
        }
        reset_bin_refine_factors(options);
        // wp reports nanoseconds (countpairs_wp_impl.c.src:669-673); the other modes seconds
        if (options->c_api_timer) options->c_api_time = (now_s() - t0) * 1e9;
        return EXIT_SUCCESS;
    });
}

// ------------------------------------------------------------ extensions
int cf_amd_set_engine(int kind)
{
    return guarded([&]() -> int {
        if (kind != CF_AMD_ENGINE_TUNED && kind != CF_AMD_ENGINE_PLAIN) return EXIT_FAILURE;
        eng().engine_kind = kind;
        return EXIT_SUCCESS;
    });
}

int cf_amd_get_engine(void) { return guarded([&]() -> int { return eng().engine_kind; }); }

int cf_amd_last_stats(cf_amd_stats* out)
{
    if (!out) return EXIT_FAILURE;
    return guarded([&]() -> int {
        *out = eng().last;
        return EXIT_SUCCESS;
    });
}


int cf_amd_countpairs_batch(int ncat, const int64_t* ND, void** X, void** Y, void** Z, void** W, const int numthreads,
                            const char* binfile, results_countpairs* results, struct config_options* options,
                            struct extra_options* extra)
{
    (void)numthreads;
    if (ncat <= 0 || !ND || !X || !Y || !Z || !results) return EXIT_FAILURE;
    if (check_common(options) != EXIT_SUCCESS) return EXIT_FAILURE;
    return guarded([&]() -> int {
        struct extra_options dummy;
        if (extra == NULL) { dummy = get_extra_options(NONE); extra = &dummy; }
        const double t0 = now_s();
        fix_maxcells(options);
        fix_refine(options);
        options->sort_on_z = 1;
        std::vector<CallOut> outs;
        if (run_batch(ncat, ND, X, Y, Z, W, binfile, options, extra, outs) != EXIT_SUCCESS) return EXIT_FAILURE;
        for (int k = 0; k < ncat; k++) {
            CallOut& R = outs[k];
            finalize(R, 1, ND[k], options->need_avg_sep, 1);
            const int nbin = R.nbin;
            results_countpairs* r = &results[k];
            r->nbin = nbin;
            r->npairs = (uint64_t*)malloc(sizeof(uint64_t) * nbin);
            r->rupp = (double*)malloc(sizeof(double) * nbin);
            r->rpavg = (double*)calloc(nbin, sizeof(double));
            r->weightavg = (double*)calloc(nbin, sizeof(double));
            if (!r->npairs || !r->rupp || !r->rpavg || !r->weightavg) {
                for (int q = 0; q <= k; q++) free_results(&results[q]);
                return EXIT_FAILURE;
            }
            for (int i = 0; i < nbin; i++) {
                r->npairs[i] = R.np[i];
                r->rupp[i] = R.rupp[i];
                r->rpavg[i] = options->need_avg_sep ? R.avg[i] : 0.0;
                r->weightavg[i] = R.need_w ? R.wavg[i] : 0.0;
            }
        }
        reset_bin_refine_factors(options);
        if (options->c_api_timer) options->c_api_time = now_s() - t0;
        return EXIT_SUCCESS;
    });
}

struct cf_amd_batch_graph {
    cfamd::BatchGraph g;
    struct config_options options;
};

int cf_amd_batch_graph_create(int ncat, const int64_t* ND, void** X, void** Y, void** Z, void** W,
                              const int numthreads, const char* binfile, struct config_options* options,
                              struct extra_options* extra, cf_amd_batch_graph** graph)
{
    (void)numthreads;
    if (ncat <= 0 || !ND || !X || !Y || !Z || !graph) return EXIT_FAILURE;
    if (check_common(options) != EXIT_SUCCESS) return EXIT_FAILURE;
    *graph = nullptr;
    return guarded([&]() -> int {
        struct extra_options dummy;
        if (extra == NULL) { dummy = get_extra_options(NONE); extra = &dummy; }
        fix_maxcells(options);
        fix_refine(options);
        options->sort_on_z = 1;
        std::unique_ptr<cf_amd_batch_graph> h(new cf_amd_batch_graph());
        h->options = *options;
        if (bg_create(ncat, ND, X, Y, Z, W, binfile, options, extra, h->g) != EXIT_SUCCESS) return EXIT_FAILURE;
        reset_bin_refine_factors(options);
        *graph = h.release();
        return EXIT_SUCCESS;
    });
}

int cf_amd_batch_graph_launch(cf_amd_batch_graph* graph, results_countpairs* results)
{
    if (!graph || !results) return EXIT_FAILURE;
    return guarded([&]() -> int {
        const double t0 = now_s();
        std::vector<CallOut> outs;
        if (bg_launch(graph->g, outs) != EXIT_SUCCESS) return EXIT_FAILURE;
        const int need_avg = graph->options.need_avg_sep;
        for (int k = 0; k < graph->g.ncat; k++) {
            CallOut& R = outs[k];
            finalize(R, 1, graph->g.nd[k], need_avg, 1);
            const int nbin = R.nbin;
            results_countpairs* r = &results[k];
            r->nbin = nbin;
            r->npairs = (uint64_t*)malloc(sizeof(uint64_t) * nbin);
            r->rupp = (double*)malloc(sizeof(double) * nbin);
            r->rpavg = (double*)calloc(nbin, sizeof(double));
            r->weightavg = (double*)calloc(nbin, sizeof(double));
            if (!r->npairs || !r->rupp || !r->rpavg || !r->weightavg) {
                for (int q = 0; q <= k; q++) free_results(&results[q]);
                return EXIT_FAILURE;
            }
            for (int i = 0; i < nbin; i++) {
                r->npairs[i] = R.np[i];
                r->rupp[i] = R.rupp[i];
                r->rpavg[i] = need_avg ? R.avg[i] : 0.0;
                r->weightavg[i] = R.need_w ? R.wavg[i] : 0.0;
            }
        }
        if (graph->options.c_api_timer) graph->options.c_api_time = now_s() - t0;
        return EXIT_SUCCESS;
    });
}

int cf_amd_batch_graph_destroy(cf_amd_batch_graph* graph)
{
    if (!graph) return EXIT_SUCCESS;
    return guarded([&]() -> int {
        CF_CHECK(hipStreamSynchronize(eng().stream));
        delete graph;
        return EXIT_SUCCESS;
    });
}

int cf_amd_set_autotune(int on)
{
    return guarded([&]() -> int {
        eng().autotune = on != 0;
        return EXIT_SUCCESS;
    });
}

int cf_amd_get_autotune(void) { return guarded([&]() -> int { return eng().autotune ? 1 : 0; }); }

int cf_amd_autotune_clear(void)
{
    return guarded([&]() -> int {
        eng().tune_cache.clear();
        eng().tune_loaded = false;  // re-read CORRFUNC_AMD_TUNE_CACHE (if set) at the next tuned call
        return EXIT_SUCCESS;
    });
}

int cf_amd_set_scheduler(int kind)
{
    return guarded([&]() -> int {
        if (kind != CF_AMD_SCHED_NONE && kind != CF_AMD_SCHED_HEURISTIC && kind != CF_AMD_SCHED_LEARNED)
            return EXIT_FAILURE;
        eng().sched_kind = kind;
        return EXIT_SUCCESS;
    });
}

int cf_amd_get_scheduler(void) { return guarded([&]() -> int { return eng().sched_kind; }); }

int cf_amd_record_costs(int on)
{
    return guarded([&]() -> int {
        if (on) ensure_sample_store();
        eng().model.recording = on != 0;
        return EXIT_SUCCESS;
    });
}

int64_t cf_amd_cost_samples(float* X, float* y, int64_t cap)
{
    int64_t out = -1;
    if (guarded([&]() -> int {
            Engine& E = eng();
            ensure_sample_store();
            int count = 0;
            CF_CHECK(hipMemcpyAsync(&count, E.model.d_count, sizeof(int), hipMemcpyDeviceToHost, E.stream));
            CF_CHECK(hipStreamSynchronize(E.stream));
            const int64_t have = std::min<int64_t>(count, E.model.cap);
            const int64_t k = std::min<int64_t>(have, cap > 0 ? cap : 0);
            if (X && k > 0)
                CF_CHECK(hipMemcpyAsync(X, E.model.d_sx, sizeof(float) * (size_t)k * kNumFeatures,
                                        hipMemcpyDeviceToHost, E.stream));
            if (y && k > 0)
                CF_CHECK(hipMemcpyAsync(y, E.model.d_sy, sizeof(float) * (size_t)k, hipMemcpyDeviceToHost, E.stream));
            CF_CHECK(hipStreamSynchronize(E.stream));
            out = have;
            return EXIT_SUCCESS;
        }) != EXIT_SUCCESS)
        return -1;
    return out;
}

int cf_amd_clear_cost_samples(void)
{
    return guarded([&]() -> int {
        Engine& E = eng();
        ensure_sample_store();
        k_fill_i32<<<1, 64, 0, E.stream>>>(1, E.model.d_count, 0);
        CF_CHECK(hipGetLastError());
        CF_CHECK(hipStreamSynchronize(E.stream));
        return EXIT_SUCCESS;
    });
}

int cf_amd_model_train(const float* X, const float* y, int64_t n, int epochs, int batch, float lr, unsigned seed,
                       float* loss)
{
    return guarded([&]() -> int {
        Engine& E = eng();
        if (!X) {
            ensure_sample_store();
            int count = 0;
            CF_CHECK(hipMemcpyAsync(&count, E.model.d_count, sizeof(int), hipMemcpyDeviceToHost, E.stream));
            CF_CHECK(hipStreamSynchronize(E.stream));
            const int64_t have = std::min<int64_t>(count, E.model.cap);
            return model_train_dev(E.model.d_sx, E.model.d_sy, have, epochs, batch, lr, seed, loss);
        }
        if (!y || n < 1) return EXIT_FAILURE;
        float* dX = (float*)E.arena.get("m_tx", sizeof(float) * (size_t)n * kNumFeatures);
        float* dY = (float*)E.arena.get("m_ty", sizeof(float) * (size_t)n);
        CF_CHECK(hipMemcpyAsync(dX, X, sizeof(float) * (size_t)n * kNumFeatures, hipMemcpyHostToDevice, E.stream));
        CF_CHECK(hipMemcpyAsync(dY, y, sizeof(float) * (size_t)n, hipMemcpyHostToDevice, E.stream));
        return model_train_dev(dX, dY, n, epochs, batch, lr, seed, loss);
    });
}

int cf_amd_model_get(float* params, float* norm)
{
    return guarded([&]() -> int {
        Engine& E = eng();
        if (!E.model.trained) return EXIT_FAILURE;
        if (params)
            CF_CHECK(hipMemcpyAsync(params, E.model.d_params, sizeof(float) * kNumParams, hipMemcpyDeviceToHost,
                                    E.stream));
        if (norm)
            CF_CHECK(hipMemcpyAsync(norm, E.model.d_norm, sizeof(float) * kNumNorm, hipMemcpyDeviceToHost, E.stream));
        CF_CHECK(hipStreamSynchronize(E.stream));
        return EXIT_SUCCESS;
    });
}

int cf_amd_model_set(const float* params, const float* norm)
{
    return guarded([&]() -> int {
        Engine& E = eng();
        if (!params || !norm) return EXIT_FAILURE;
        model_alloc();
        CF_CHECK(hipMemcpyAsync(E.model.d_params, params, sizeof(float) * kNumParams, hipMemcpyHostToDevice,
                                E.stream));
        CF_CHECK(hipMemcpyAsync(E.model.d_norm, norm, sizeof(float) * kNumNorm, hipMemcpyHostToDevice, E.stream));
        CF_CHECK(hipStreamSynchronize(E.stream));
        E.model.trained = true;
        return EXIT_SUCCESS;
    });
}

int cf_amd_model_reset(void)
{
    return guarded([&]() -> int {
        eng().model.trained = false;
        return EXIT_SUCCESS;
    });
}

int cf_amd_model_trained(void) { return guarded([&]() -> int { return eng().model.trained ? 1 : 0; }); }

int cf_amd_model_predict(const float* X, int64_t n, float* out)
{
    return guarded([&]() -> int {
        Engine& E = eng();
        if (!E.model.trained || !X || !out || n < 0 || n > (1 << 30)) return EXIT_FAILURE;
        if (n == 0) return EXIT_SUCCESS;
        float* dX = (float*)E.arena.get("m_px", sizeof(float) * (size_t)n * kNumFeatures);
        float* dO = (float*)E.arena.get("m_po", sizeof(float) * (size_t)n);
        CF_CHECK(hipMemcpyAsync(dX, X, sizeof(float) * (size_t)n * kNumFeatures, hipMemcpyHostToDevice, E.stream));
        k_mlp_predict<<<(unsigned)((n + 255) / 256), 256, 0, E.stream>>>((int)n, dX, E.model.d_norm,
                                                                         E.model.d_params, dO);
        CF_CHECK(hipGetLastError());
        CF_CHECK(hipMemcpyAsync(out, dO, sizeof(float) * (size_t)n, hipMemcpyDeviceToHost, E.stream));
        CF_CHECK(hipStreamSynchronize(E.stream));
        return EXIT_SUCCESS;
    });
}

int cf_amd_model_grad(const float* X, const float* y, int64_t n, const float* params, const float* norm,
                      float* grad, float* loss)
{
    return guarded([&]() -> int {
        Engine& E = eng();
        if (!X || !y || !params || !norm || !grad || n < 1 || n > (1 << 20)) return EXIT_FAILURE;
        float* dX = (float*)E.arena.get("m_gx", sizeof(float) * (size_t)n * kNumFeatures);
        float* dY = (float*)E.arena.get("m_gy", sizeof(float) * (size_t)n);
        float* dP = (float*)E.arena.get("m_gp", sizeof(float) * kNumParams);
        float* dN = (float*)E.arena.get("m_gn", sizeof(float) * kNumNorm);
        float* dG = (float*)E.arena.get("m_gg", sizeof(float) * kNumParams);
        float* dL = (float*)E.arena.get("m_gl", sizeof(float));
        CF_CHECK(hipMemcpyAsync(dX, X, sizeof(float) * (size_t)n * kNumFeatures, hipMemcpyHostToDevice, E.stream));
        CF_CHECK(hipMemcpyAsync(dY, y, sizeof(float) * (size_t)n, hipMemcpyHostToDevice, E.stream));
        CF_CHECK(hipMemcpyAsync(dP, params, sizeof(float) * kNumParams, hipMemcpyHostToDevice, E.stream));
        CF_CHECK(hipMemcpyAsync(dN, norm, sizeof(float) * kNumNorm, hipMemcpyHostToDevice, E.stream));
        model_grad_dev((int)n, nullptr, dX, dY, dN, dP, (float)(1.0 / n), dG, dL);
        CF_CHECK(hipMemcpyAsync(grad, dG, sizeof(float) * kNumParams, hipMemcpyDeviceToHost, E.stream));
        float l = 0;
        CF_CHECK(hipMemcpyAsync(&l, dL, sizeof(float), hipMemcpyDeviceToHost, E.stream));
        CF_CHECK(hipStreamSynchronize(E.stream));
        if (loss) *loss = l;
        return EXIT_SUCCESS;
    });
}

int cf_amd_model_adam_step(float* params, const float* grad, float* m, float* v, int step, float lr)
{
    if (!params || !grad || !m || !v || step < 1) return EXIT_FAILURE;
    return guarded([&]() -> int {
        Engine& E = eng();
        float* d = (float*)E.arena.get("adam_check", sizeof(float) * 4 * kNumParams);
        const size_t b = sizeof(float) * kNumParams;
        CF_CHECK(hipMemcpyAsync(d, params, b, hipMemcpyHostToDevice, E.stream));
        CF_CHECK(hipMemcpyAsync(d + kNumParams, grad, b, hipMemcpyHostToDevice, E.stream));
        CF_CHECK(hipMemcpyAsync(d + 2 * kNumParams, m, b, hipMemcpyHostToDevice, E.stream));
        CF_CHECK(hipMemcpyAsync(d + 3 * kNumParams, v, b, hipMemcpyHostToDevice, E.stream));
        k_adam<<<(kNumParams + 255) / 256, 256, 0, E.stream>>>(d, d + kNumParams, d + 2 * kNumParams,
                                                                d + 3 * kNumParams, adam_consts(step, lr));
        CF_CHECK(hipGetLastError());
        CF_CHECK(hipMemcpyAsync(params, d, b, hipMemcpyDeviceToHost, E.stream));
        CF_CHECK(hipMemcpyAsync(m, d + 2 * kNumParams, b, hipMemcpyDeviceToHost, E.stream));
        CF_CHECK(hipMemcpyAsync(v, d + 3 * kNumParams, b, hipMemcpyDeviceToHost, E.stream));
        CF_CHECK(hipStreamSynchronize(E.stream));
        return EXIT_SUCCESS;
    });
}

const char* cf_amd_version(void) { return "corrfunc_amd 0.1 (reference API 2.5.3, gfx950)"; }

void* cf_amd_device_alloc(size_t bytes)
{
    void* p = nullptr;
    if (guarded([&]() -> int {
            eng();
            CF_CHECK(hipMalloc(&p, bytes ? bytes : 1));
            return EXIT_SUCCESS;
        }) != EXIT_SUCCESS)
        return nullptr;
    return p;
}

int cf_amd_device_free(void* p)
{
    return guarded([&]() -> int {
        CF_CHECK(hipFree(p));
        return EXIT_SUCCESS;
    });
}

int cf_amd_memcpy(void* dst, const void* src, size_t bytes, int kind)
{
    return guarded([&]() -> int {
        Engine& E = eng();
        if (kind != CF_AMD_H2D && kind != CF_AMD_D2H && kind != CF_AMD_D2D) {
            fprintf(stderr, "cf_amd_memcpy: unknown copy kind %d\n", kind);
            return EXIT_FAILURE;
        }
        if (kind == CF_AMD_D2D && ((uintptr_t)dst % 16 == 0) && ((uintptr_t)src % 16 == 0)) {
            // device copies run through the engine's own kernel (16-byte words + tail)
            const int64_t n16 = (int64_t)(bytes / 16);
            const int tail = (int)(bytes % 16);
            const int64_t want = (n16 + 255) / 256;
            const unsigned blocks = (unsigned)std::max<int64_t>(1, std::min<int64_t>(want, 256 * 32));
            k_copy16<<<blocks, 256, 0, E.stream>>>(n16, (const uint4*)src, (uint4*)dst, tail,
                                                   (const unsigned char*)src + n16 * 16,
                                                   (unsigned char*)dst + n16 * 16);
            CF_CHECK(hipGetLastError());
        } else {
            const hipMemcpyKind k = kind == CF_AMD_H2D ? hipMemcpyHostToDevice
                                  : kind == CF_AMD_D2H ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice;
            CF_CHECK(hipMemcpyAsync(dst, src, bytes, k, E.stream));
        }
        CF_CHECK(hipStreamSynchronize(E.stream));
        return EXIT_SUCCESS;
    });
}

void free_results_mocks(results_countpairs_mocks* r)
{
    if (!r) return;
    free(r->npairs);
    free(r->rupp);
    free(r->rpavg);
    free(r->weightavg);
    r->npairs = nullptr;
    r->rupp = r->rpavg = r->weightavg = nullptr;
}

void free_results_mocks_s_mu(results_countpairs_mocks_s_mu* r)
{
    if (!r) return;
    free(r->npairs);
    free(r->supp);
    free(r->savg);
    free(r->weightavg);
    r->npairs = nullptr;
    r->supp = r->savg = r->weightavg = nullptr;
}

int countpairs_mocks(const int64_t ND1, void* RA1, void* DEC1, void* CZ1, const int64_t ND2, void* RA2, void* DEC2,
                     void* CZ2, const int numthreads, const int autocorr, const char* binfile, const double pimax,
                     const int cosmology, results_countpairs_mocks* results, struct config_options* options,
                     struct extra_options* extra)
{
    if (check_common(options) != EXIT_SUCCESS) return EXIT_FAILURE;
    return guarded([&]() -> int {
        struct extra_options dummy;
        if (extra == NULL) { dummy = get_extra_options(NONE); extra = &dummy; }
        const double t0 = now_s();
        options->sort_on_z = 1;
        fix_refine(options);
        fix_maxcells(options);
        if (options->fast_divide_and_NR_steps >= MAX_FAST_DIVIDE_NR_STEPS) options->fast_divide_and_NR_steps = 0;
        CallOut R;
        if (ND1 == 0 || (autocorr == 0 && ND2 == 0)) {
            // nothing to count: bins and zero counts
            double rmin, rmax;
            int nbin;
            if (setup_bins(binfile, &rmin, &rmax, &nbin, R.rupp) != EXIT_SUCCESS) return EXIT_FAILURE;
            R.nbin = nbin;
            R.n2 = (int)pimax;
            const size_t tot = (size_t)(R.n2 + 1) * (nbin + 1);
            R.np.assign(tot, 0);
            R.avg.assign(tot, 0.0);
            R.wavg.assign(tot, 0.0);
            R.need_w = extra->weight_method != NONE;
        } else {
            if (sky_run(MODE_RPPI_MOCKS, ND1, RA1, DEC1, CZ1, ND2, RA2, DEC2, CZ2, numthreads, autocorr, binfile,
                        pimax, 0.0, 0, cosmology, options, extra, R) != EXIT_SUCCESS) {
                reset_bin_refine_factors(options);
                return EXIT_FAILURE;
            }
            finalize(R, autocorr, ND1, options->need_avg_sep, R.n2 + 1);
        }
        const int npibin = R.n2, nbin = R.nbin;
        const int64_t tot = (int64_t)(npibin + 1) * (nbin + 1);
        results->nbin = nbin;
        results->npibin = npibin;
        results->pimax = pimax;
        results->npairs = (uint64_t*)calloc(tot, sizeof(uint64_t));
        results->rupp = (double*)malloc(sizeof(double) * nbin);
        results->rpavg = (double*)calloc(tot, sizeof(double));
        results->weightavg = (double*)calloc(tot, sizeof(double));
        if (!results->npairs || !results->rupp || !results->rpavg || !results->weightavg) {
            free_results_mocks(results);
            return EXIT_FAILURE;
        }
        for (int i = 0; i < nbin; i++) {
            results->rupp[i] = R.rupp[i];
            for (int j = 0; j < npibin; j++) {
                const int64_t idx = (int64_t)i * (npibin + 1) + j;
                results->npairs[idx] = R.np[idx];
                results->rpavg[idx] = options->need_avg_sep ? R.avg[idx] : 0.0;
                results->weightavg[idx] = R.need_w ? R.wavg[idx] : 0.0;
            }
        }
        reset_bin_refine_factors(options);
        if (options->c_api_timer) options->c_api_time = now_s() - t0;
        return EXIT_SUCCESS;
    });
}

int countpairs_mocks_s_mu(const int64_t ND1, void* RA1, void* DEC1, void* CZ1, const int64_t ND2, void* RA2,
                          void* DEC2, void* CZ2, const int numthreads, const int autocorr, const char* sbinfile,
                          const double mu_max, const int nmu_bins, const int cosmology,
                          results_countpairs_mocks_s_mu* results, struct config_options* options,
                          struct extra_options* extra)
{
    if (check_common(options) != EXIT_SUCCESS) return EXIT_FAILURE;
    return guarded([&]() -> int {
        struct extra_options dummy;
        if (extra == NULL) { dummy = get_extra_options(NONE); extra = &dummy; }
        const double t0 = now_s();
        options->sort_on_z = 1;
        fix_refine(options);
        fix_maxcells(options);
        if (options->fast_divide_and_NR_steps >= MAX_FAST_DIVIDE_NR_STEPS) options->fast_divide_and_NR_steps = 0;
        CallOut R;
        if (ND1 == 0 || (autocorr == 0 && ND2 == 0)) {
            double rmin, rmax;
            int nbin;
            if (setup_bins(sbinfile, &rmin, &rmax, &nbin, R.rupp) != EXIT_SUCCESS) return EXIT_FAILURE;
            if (mu_max <= 0.0 || mu_max > 1.0 || nmu_bins < 1) {
                fprintf(stderr, "Error: need 0 < mu_max <= 1 and nmu_bins >= 1 (got %lf, %d)\n", mu_max, nmu_bins);
                return EXIT_FAILURE;
            }
            R.nbin = nbin;
            R.n2 = nmu_bins;
            const size_t tot = (size_t)(R.n2 + 1) * (nbin + 1);
            R.np.assign(tot, 0);
            R.avg.assign(tot, 0.0);
            R.wavg.assign(tot, 0.0);
            R.need_w = extra->weight_method != NONE;
        } else {
            if (sky_run(MODE_SMU_MOCKS, ND1, RA1, DEC1, CZ1, ND2, RA2, DEC2, CZ2, numthreads, autocorr, sbinfile, 0.0,
                        mu_max, nmu_bins, cosmology, options, extra, R) != EXIT_SUCCESS) {
                reset_bin_refine_factors(options);
                return EXIT_FAILURE;
            }
            finalize(R, autocorr, ND1, options->need_avg_sep, R.n2 + 1);
        }
        const int n2 = R.n2, nbin = R.nbin;
        const int64_t tot = (int64_t)(n2 + 1) * (nbin + 1);
        results->nsbin = nbin;
        results->nmu_bins = n2;
        results->mu_max = mu_max;
        results->mu_min = 0.0;
        results->npairs = (uint64_t*)calloc(tot, sizeof(uint64_t));
        results->supp = (double*)malloc(sizeof(double) * nbin);
        results->savg = (double*)calloc(tot, sizeof(double));
        results->weightavg = (double*)calloc(tot, sizeof(double));
        if (!results->npairs || !results->supp || !results->savg || !results->weightavg) {
            free_results_mocks_s_mu(results);
            return EXIT_FAILURE;
        }
        for (int i = 0; i < nbin; i++) {
            results->supp[i] = R.rupp[i];
            for (int j = 0; j < n2; j++) {
                const int64_t idx = (int64_t)i * (n2 + 1) + j;
                results->npairs[idx] = R.np[idx];
                results->savg[idx] = options->need_avg_sep ? R.avg[idx] : 0.0;
                results->weightavg[idx] = R.need_w ? R.wavg[idx] : 0.0;
            }
        }
        reset_bin_refine_factors(options);
        if (options->c_api_timer) options->c_api_time = now_s() - t0;
        return EXIT_SUCCESS;
    });
}

}  // extern "C"
