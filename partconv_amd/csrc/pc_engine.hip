// pc_engine.hip -- host engine and C ABI of the MI355X partitioned-convolution
// library (libpartconv_amd.so).
//
// What the reference does with FFTW plans, pthread level workers and condvars
// (src/partconv.cpp:70-648, src/partconvMulti.cpp:23-1134) is done here with:
//   * the kernels of pc_kernels.hpp (level jobs batched across instances);
//   * one HIP stream per partition level (levels >= 1; at most PC_LEVEL_STREAMS
//     streams, levels beyond that share round-robin) plus the tick stream;
//   * events for the deadlines: a level's block started at tick f_s must be
//     complete before the tick kernel that first reads it (strict: tick
//     f_s + k - 1; relaxed: tick f_s + k), enforced with hipStreamWaitEvent;
//   * a host-visible completion flag that the tick's last workgroup raises, so
//     a tick costs one launch plus a spin on pinned memory.
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>
#include <hsa/hsa.h>
#include <hsa/hsa_ext_amd.h>
#include <immintrin.h>

#include <algorithm>
#include <array>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/partconv_amd.h"
#include "pc_io.hpp"
#include "pc_kernels.hpp"
#include "pc_mimo_rt.hpp"

#define PC_API extern "C" __attribute__((visibility("default")))
#define PC_LEVEL_STREAMS 3
#define PC_BIG_MIN 4096       // block sizes >= this use the four-step kernels
#define PC_MAX_BLOCK 65536
#define PC_MIN_BLOCK 16
#define PC_TW_MAX (2 * PC_MAX_BLOCK)
#define PC_GO_WORD 1          // release counter: the word after the completion flag (same line)
#define PC_SPIN_LIMIT 200000000ull  // persistent-kernel waits: 2 s at 100 MHz (s_memrealtime)
#define PC_SESSION_IDLE_S 1.0       // host sessions restart after this long without a release
#define PC_TICK_TIMEOUT_S 10.0
#define PC_RELEASE_WINDOW_MAX 16
#define PC_LL_MAX_WORDS 16384       // host sessions use tagged input words up to nIn * N0 of these
#define PC_WG_FLAGS_MAX 1024        // direct mix: per-workgroup host flags up to this many instances
                                    // (the host polls them; beyond, device tickets raise one flag)
#define PC_DH_PT 8                  // dL/dh: partitions per thread (register window)
#define PC_TUNE_TIE 0.05            // tuners: candidates within 5% of the fastest are equal

static thread_local std::string g_err;

static int fail(int code, const std::string &msg)
{
    g_err = msg;
    return code;
}

#define HIPCHK(x)                                                                        \
    do {                                                                                 \
        hipError_t e_ = (x);                                                             \
        if (e_ != hipSuccess)                                                            \
            return fail(PC_ERR_HIP, std::string(#x) + ": " + hipGetErrorString(e_));    \
    } while (0)

static double now_s()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

static int ilog2(long long x)
{
    int l = 0;
    while ((1ll << l) < x) l++;
    return l;
}
static bool is_pow2(long long x) { return x > 0 && (x & (x - 1)) == 0; }

// ----------------------------------------------------------------------------
// Twiddle table, one per device: tw[n + t] = exp(-2 pi i t / n), double -> float.
// ----------------------------------------------------------------------------
static cf *g_tw[64];

static int get_twiddles(int dev, cf **out)
{
    if (dev < 0 || dev >= 64) return fail(PC_ERR_ARG, "device index out of range");
    if (!g_tw[dev]) {
        std::vector<cf> h((size_t)2 * PC_TW_MAX);
        h[0] = make_float2(1.f, 0.f);
        for (long long n = 1; n <= PC_TW_MAX; n <<= 1)
            for (long long t = 0; t < n; t++) {
                const double a = -2.0 * M_PI * (double)t / (double)n;
                h[n + t] = make_float2((float)cos(a), (float)sin(a));
            }
        cf *d = nullptr;
        HIPCHK(hipMalloc(&d, h.size() * sizeof(cf)));
        HIPCHK(hipMemcpy(d, h.data(), h.size() * sizeof(cf), hipMemcpyHostToDevice));
        g_tw[dev] = d;
    }
    *out = g_tw[dev];
    return PC_OK;
}

// ----------------------------------------------------------------------------
// Device memory that the host writes through the PCIe BAR.
//
// On large-BAR systems the HSA runtime maps device memory into the host's
// address space.  Blocks come from the device's fine-grained global pool with
// CPU access granted (each checked by a host write and read-back).  They are
// never returned to the runtime -- device memory that kernels have polled is
// kept until the process ends -- but freed blocks are reused by later engines.
// Uses:
//   * release words: the host releases each tick of the persistent kernel by
//     writing a counter that the kernel polls.  In pinned host memory every
//     poll is a read across the bus (about 1.2 us on MI355X); in device memory
//     the kernel polls locally and only the host's write crosses the bus
//     (posted), measured 0.6-0.8 us less per host round trip at 16 workgroups.
//     Engines take 64-byte slots of one page per device.
//   * the input frames of host sessions (tagged words, see ll_tag).
// Without a large BAR, or if any step fails, callers fall back to pinned host
// memory.
// ----------------------------------------------------------------------------
#define PC_REL_SLOTS 256
#define PC_REL_STRIDE 8  // 64-byte slots, in 8-byte words

struct BarDevice {
    bool tried = false, ok = false;
    hsa_amd_memory_pool_t pool{};
    hsa_agent_t cpu{};
    std::vector<std::pair<size_t, void *>> free_blocks;
    unsigned long long *rel_base = nullptr;
    std::vector<int> free_slots;
};
static std::mutex g_bar_mu;
static BarDevice g_bar[64];
static int g_hsa_inits = 0;  // successful hsa_init calls of bar_init (g_bar_mu held)

struct HsaFind {
    uint32_t bdf = 0, domain = 0;
    hsa_agent_t gpu{}, cpu{};
    bool has_gpu = false, has_cpu = false;
    hsa_amd_memory_pool_t pool{};
    bool has_pool = false;
};

static hsa_status_t rel_agent_cb(hsa_agent_t a, void *data)
{
    HsaFind *f = static_cast<HsaFind *>(data);
    hsa_device_type_t t;
    if (hsa_agent_get_info(a, HSA_AGENT_INFO_DEVICE, &t) != HSA_STATUS_SUCCESS) return HSA_STATUS_SUCCESS;
    if (t == HSA_DEVICE_TYPE_CPU && !f->has_cpu) {
        f->cpu = a;
        f->has_cpu = true;
    } else if (t == HSA_DEVICE_TYPE_GPU && !f->has_gpu) {
        uint32_t bdf = 0, dom = 0;
        hsa_agent_get_info(a, (hsa_agent_info_t)HSA_AMD_AGENT_INFO_BDFID, &bdf);
        hsa_agent_get_info(a, (hsa_agent_info_t)HSA_AMD_AGENT_INFO_DOMAIN, &dom);
        if (bdf == f->bdf && dom == f->domain) {
            f->gpu = a;
            f->has_gpu = true;
        }
    }
    return HSA_STATUS_SUCCESS;
}

static hsa_status_t rel_pool_cb(hsa_amd_memory_pool_t pool, void *data)
{
    HsaFind *f = static_cast<HsaFind *>(data);
    hsa_amd_segment_t seg;
    uint32_t flags = 0;
    if (hsa_amd_memory_pool_get_info(pool, HSA_AMD_MEMORY_POOL_INFO_SEGMENT, &seg) != HSA_STATUS_SUCCESS)
        return HSA_STATUS_SUCCESS;
    if (seg != HSA_AMD_SEGMENT_GLOBAL) return HSA_STATUS_SUCCESS;
    hsa_amd_memory_pool_get_info(pool, HSA_AMD_MEMORY_POOL_INFO_GLOBAL_FLAGS, &flags);
    if ((flags & HSA_AMD_MEMORY_POOL_GLOBAL_FLAG_FINE_GRAINED) && !f->has_pool) {
        f->pool = pool;
        f->has_pool = true;
    }
    return HSA_STATUS_SUCCESS;
}

// Find the device's fine-grained pool and the CPU agent (once); g_bar_mu held.
static bool bar_init(int dev)
{
    BarDevice &B = g_bar[dev];
    if (B.tried) return B.ok;
    B.tried = true;
    int large_bar = 0, bus = 0, devno = 0, domain = 0;
    if (hipDeviceGetAttribute(&large_bar, hipDeviceAttributeIsLargeBar, dev) != hipSuccess || !large_bar) return false;
    if (hipDeviceGetAttribute(&bus, hipDeviceAttributePciBusId, dev) != hipSuccess ||
        hipDeviceGetAttribute(&devno, hipDeviceAttributePciDeviceId, dev) != hipSuccess ||
        hipDeviceGetAttribute(&domain, hipDeviceAttributePciDomainID, dev) != hipSuccess)
        return false;
    if (hsa_init() != HSA_STATUS_SUCCESS) return false;
    g_hsa_inits++;  // paired with hsa_shut_down in pc_shutdown
    HsaFind f;
    f.bdf = ((uint32_t)bus << 8) | ((uint32_t)devno << 3);
    f.domain = (uint32_t)domain;
    hsa_iterate_agents(rel_agent_cb, &f);
    if (!f.has_gpu || !f.has_cpu) return false;
    hsa_amd_agent_iterate_memory_pools(f.gpu, rel_pool_cb, &f);
    if (!f.has_pool) return false;
    B.pool = f.pool;
    B.cpu = f.cpu;
    B.ok = true;
    return true;
}

// A zeroed block of `bytes` (rounded up to 4 KiB) of host-writable device
// memory on `dev`, or nullptr; g_bar_mu held.
static void *bar_alloc_locked(int dev, size_t bytes)
{
    if (!bar_init(dev)) return nullptr;
    BarDevice &B = g_bar[dev];
    bytes = (bytes + 4095) & ~(size_t)4095;
    void *p = nullptr;
    for (size_t i = 0; i < B.free_blocks.size(); i++)
        if (B.free_blocks[i].first == bytes) {
            p = B.free_blocks[i].second;
            B.free_blocks.erase(B.free_blocks.begin() + i);
            break;
        }
    if (!p) {
        if (hsa_amd_memory_pool_allocate(B.pool, bytes, 0, &p) != HSA_STATUS_SUCCESS || !p) return nullptr;
        if (hsa_amd_agents_allow_access(1, &B.cpu, nullptr, p) != HSA_STATUS_SUCCESS) return nullptr;  // kept
        volatile unsigned long long *h = static_cast<volatile unsigned long long *>(p);
        h[0] = 0x5ca1ab1eull;  // the host must read back what it wrote
        _mm_mfence();
        if (h[0] != 0x5ca1ab1eull) return nullptr;
    }
    volatile unsigned long long *h = static_cast<volatile unsigned long long *>(p);
    for (size_t i = 0; i < bytes / sizeof(unsigned long long); i++) h[i] = 0;
    _mm_mfence();
    return p;
}

static void *bar_alloc(int dev, size_t bytes)
{
    if (dev < 0 || dev >= 64) return nullptr;
    std::lock_guard<std::mutex> lk(g_bar_mu);
    return bar_alloc_locked(dev, bytes);
}

static void bar_free(int dev, void *p, size_t bytes)
{
    if (!p || dev < 0 || dev >= 64) return;
    std::lock_guard<std::mutex> lk(g_bar_mu);
    g_bar[dev].free_blocks.push_back({(bytes + 4095) & ~(size_t)4095, p});
}

// A release word for an engine on `dev`: a device-memory slot (returns the
// slot index) or -1 when none is available.
static int release_slot_get(int dev, unsigned long long **word)
{
    if (dev < 0 || dev >= 64) return -1;
    std::lock_guard<std::mutex> lk(g_bar_mu);
    BarDevice &B = g_bar[dev];
    if (!B.rel_base) {
        if (B.tried && !B.ok) return -1;
        B.rel_base = static_cast<unsigned long long *>(
            bar_alloc_locked(dev, (size_t)PC_REL_SLOTS * PC_REL_STRIDE * sizeof(unsigned long long)));
        if (!B.rel_base) return -1;
        for (int s = PC_REL_SLOTS - 1; s >= 0; s--) B.free_slots.push_back(s);
    }
    if (B.free_slots.empty()) return -1;
    const int s = B.free_slots.back();
    B.free_slots.pop_back();
    *word = B.rel_base + (size_t)s * PC_REL_STRIDE;
    *(volatile unsigned long long *)*word = 0;
    _mm_mfence();
    return s;
}

static void release_slot_put(int dev, int slot)
{
    if (slot < 0 || dev < 0 || dev >= 64) return;
    std::lock_guard<std::mutex> lk(g_bar_mu);
    g_bar[dev].free_slots.push_back(slot);
}

// ----------------------------------------------------------------------------
// Live handles and library shutdown.
//
// Every handle a *_create returns is recorded until its *_destroy takes it.
// pc_shutdown() destroys the handles still live -- stopping resident
// persistent kernels and joining the launcher threads -- then returns the BAR
// blocks and the twiddle tables and pairs bar_init's hsa_init with
// hsa_shut_down.  It runs from an atexit handler registered right after the
// library's first HIP call (so it runs before the HIP runtime's own exit-time
// teardown, which was registered earlier) and from the Python loader's atexit
// hook.  After it, *_destroy of a handle it destroyed is a no-op and *_create
// fails: a destructor that runs late at exit never touches a torn-down runtime.
// ----------------------------------------------------------------------------
enum { HK_MULTI = 1, HK_FILTER, HK_CONV, HK_MIMO, HK_MIMO_RT };
static std::mutex g_live_mu;
static std::vector<std::pair<void *, int>> g_live;
static bool g_shut = false;

static bool live_add(void *p, int kind)
{
    std::lock_guard<std::mutex> lk(g_live_mu);
    if (g_shut) return false;
    g_live.push_back({p, kind});
    return true;
}

// True if p was live (it is no longer); false for a handle already destroyed
// (by pc_shutdown or an earlier *_destroy) or never returned by *_create.
static bool live_take(void *p)
{
    std::lock_guard<std::mutex> lk(g_live_mu);
    for (size_t i = 0; i < g_live.size(); i++)
        if (g_live[i].first == p) {
            g_live.erase(g_live.begin() + i);
            return true;
        }
    return false;
}

static bool lib_shut()
{
    std::lock_guard<std::mutex> lk(g_live_mu);
    return g_shut;
}

PC_API void pc_shutdown(void);
static void shutdown_at_exit() { pc_shutdown(); }

// Called after a HIP call has succeeded: the HIP runtime is initialised, so an
// atexit handler registered now runs before its teardown.
static void hip_live()
{
    static std::once_flag once;
    std::call_once(once, [] { atexit(shutdown_at_exit); });
}

// ----------------------------------------------------------------------------
// Persistent-kernel launch.
//
// Cooperative launch first: the runtime then guarantees that every workgroup
// is resident at once.  The device grants cooperative launches to one process
// at a time (while a process that made one is alive, another process's
// cooperative launch fails), and these kernels use no grid barrier, so on
// refusal the kernel is launched plainly if it fits the device's occupancy.
// Its workgroups then stamp a pinned word each on entry (stamp_present), and
// the host releases the first tick only after seeing every stamp within
// PC_RESIDENT_S; otherwise it stops the kernel (every wait in it ends at the
// stop code) and falls back like any other failed launch.
// ----------------------------------------------------------------------------
#define PC_RESIDENT_S 0.1

struct Residency {
    unsigned *h = nullptr;  // pinned words, host view
    unsigned *d = nullptr;  // device view
    int cap = 0;
    unsigned tag = 0;
};

static void resident_free(Residency &R)
{
    if (R.h) hipHostFree(R.h);
    R = Residency();
}

static int resident_alloc(Residency &R, int n)
{
    if (R.cap >= n) return PC_OK;
    resident_free(R);
    HIPCHK(hipHostMalloc(&R.h, (size_t)n * sizeof(unsigned), hipHostMallocCoherent | hipHostMallocMapped));
    memset(R.h, 0, (size_t)n * sizeof(unsigned));
    HIPCHK(hipHostGetDevicePointer((void **)&R.d, R.h, 0));
    R.cap = n;
    return PC_OK;
}

// Launch fn over `grid` workgroups of nt threads; present / tag are the
// kernel argument's stamp fields (args points at that argument).  *plain: the
// launch was not cooperative and the caller must resident_wait before the
// first release.
static hipError_t launch_persistent(const void *fn, int grid, int nt, void **args, size_t lds, hipStream_t s,
                                    Residency &R, int ncu, unsigned **present, unsigned *tag, bool *plain)
{
    *present = nullptr;
    *plain = false;
    hipError_t e = hipLaunchCooperativeKernel(fn, dim3(grid), dim3(nt), args, (unsigned)lds, s);
    if (e == hipSuccess) return e;
    (void)hipGetLastError();
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, fn, nt, lds) != hipSuccess || (long long)nb * ncu < grid ||
        resident_alloc(R, grid) != PC_OK) {
        (void)hipGetLastError();
        return e;
    }
    if (++R.tag == 0) R.tag = 1;
    *present = R.d;
    *tag = R.tag;
    const hipError_t e2 = hipLaunchKernel(fn, dim3(grid), dim3(nt), args, lds, s);
    if (e2 != hipSuccess) {
        (void)hipGetLastError();
        *present = nullptr;
        return e;
    }
    *plain = true;
    return hipSuccess;
}

// Every workgroup of a plain launch stamped the current tag in time.
static bool resident_wait(const Residency &R, int grid)
{
    const double t0 = now_s();
    const volatile unsigned *h = R.h;
    for (int w = 0; w < grid; w++)
        while (h[w] != R.tag)
            if (now_s() - t0 > PC_RESIDENT_S) return false;
    return true;
}

// ----------------------------------------------------------------------------
// Partition rules (PartConvFilter::setup validation, src/partconv.cpp:86-182)
// ----------------------------------------------------------------------------
static int check_partition(std::vector<int> &N, std::vector<int> &P, long long ir_len,
                           std::vector<int> &delay, std::vector<int> &T)
{
    const int M = (int)N.size();
    if (M < 1 || (int)P.size() != M) return fail(PC_ERR_PARTITION, "need M >= 1 levels");
    if (M > PC_MAX_LEVELS) return fail(PC_ERR_PARTITION, "too many levels");
    for (int i = 0; i < M; i++)
        if (!is_pow2(N[i])) {
            char m[160];
            snprintf(m, sizeof m, "PartConvFilter ERROR: block size of level %d must be a power of 2 (N[%d] = %d)", i, i, N[i]);
            return fail(PC_ERR_PARTITION, m);
        }
    for (int i = 1; i < M; i++)
        if (N[i] <= N[i - 1]) {
            char m[160];
            snprintf(m, sizeof m, "PartConvFilter ERROR: N[%d]=%d <= N[%d]=%d", i, N[i], i - 1, N[i - 1]);
            return fail(PC_ERR_PARTITION, m);
        }
    long long position = 2ll * N[0];
    delay.assign(M, 0);
    for (int i = 0; i < M - 1; i++) {
        delay[i] = (int)(position / N[0]);
        position += (long long)N[i] * P[i];
        if (position < 2ll * N[i + 1]) {
            char m[160];
            snprintf(m, sizeof m, "PartConvFilter ERROR: sufficient delay not allowed for block size of level %d", i + 2);
            return fail(PC_ERR_PARTITION, m);
        }
    }
    delay[M - 1] = (int)(position / N[0]);
    const long long rest = ir_len - (position - 2ll * N[0]);
    // float division then ceil, as src/partconv.cpp:169
    const int plast = (int)ceilf((float)rest / (float)N[M - 1]);
    if (rest < 0 || plast <= 0) {
        char m[160];
        snprintf(m, sizeof m, "PartConvFilter ERROR: last partitioning level (%d) not required", M);
        return fail(PC_ERR_PARTITION, m);
    }
    P[M - 1] = plast;
    T.assign(M, 0);
    long long t = 0;
    for (int i = 0; i < M; i++) {
        T[i] = (int)t;
        t += (long long)N[i] * P[i];
    }
    if (t > (1ll << 30)) return fail(PC_ERR_PARTITION, "impulse response too long");
    return PC_OK;
}

// ----------------------------------------------------------------------------
// Engine state
// ----------------------------------------------------------------------------
struct LevelSpec {
    int N, P, T, delay, kshift;
    bool big;
    BigShape sh;
    size_t off_H, off_fdl, off_ybuf, off_scr, off_S;
};

struct InstDesc {
    int mute = 0, in_ch = 0, out_ch = 0;
    float scale = 1.f;
    std::vector<int> N, P;
    std::vector<float> ir;
};

struct Inst {
    InstDesc d;
    std::vector<LevelSpec> lv;
    std::vector<int> gmap;  // global level -> local level or -1
    int R;
    size_t off_ring;
};

struct Batch {
    int g = 0, N = 0;
    bool big = false;
    BigShape sh{0, 0};
    LevelJob *d_jobs = nullptr;
    int njobs = 0;
};

struct Group {
    std::vector<int> members;
    TickJob *d_tick = nullptr;
    LevelJob *d_l0 = nullptr;     // level-0 jobs (filter API)
    std::vector<Batch> batches;   // levels >= 1
};

static BigShape big_shape(int N)
{
    const int lg = ilog2(N);
    BigShape s;
    s.logB = std::min(7, lg / 2);
    s.logA = lg - s.logB;
    return s;
}

// Launches of host-launched level blocks during persistent runs, issued by a
// helper thread so that the tick thread is not held by hipLaunchKernel calls
// (a few microseconds each) between two releases.  The tick thread queues
// (level, start tick); the helper launches the level's batches on its stream,
// records the level's event and then publishes done[g] = start tick + 1.  The
// tick that first reads the block waits for done[g] before it queries the
// event.  The reference starts its level workers from the audio thread with
// condition variables (src/partconvMulti.cpp:860-903); here the helper only
// issues launches, the GPU does the work.
struct LevelLauncher {
    std::thread th;
    std::mutex mu;
    std::condition_variable cv;
    std::vector<std::pair<int, uint64_t>> q;
    bool stop = false;
    int busy = 0;  // queued or in progress
    std::unique_ptr<std::atomic<uint64_t>[]> done;
    int err = 0;
    std::string msg;
};

// A level's block launches as one executable HIP graph (pcm_set_level_graphs):
// the kernel nodes of every batch of the level, chained in launch order; per
// block start only the tick-dependent arguments (block index, FDL slot,
// output half, next block's delay-line sum) change, through
// hipGraphExecKernelNodeSetParams, and one hipGraphLaunch on the level's stream
// replaces the 2-4 launches per batch.
struct GArgs {  // argument storage of one kernel node (kernelParams point into it)
    const LevelJob *jobs = nullptr;
    BigShape sh{0, 0};
    int N0 = 0, N = 0, wpj = 0, ypar = 0, mode = 0;
    uint64_t m = 0, bcur = 0, b = 0;
    const cf *tw = nullptr;
    void *ptrs[8] = {};
};
struct GNode {
    hipGraphNode_t node = nullptr;
    hipKernelNodeParams p{};
    std::unique_ptr<GArgs> a;
};
struct LevelGraph {
    hipGraph_t g = nullptr;
    hipGraphExec_t x = nullptr;
    std::vector<GNode> nodes;
};

struct Engine {
    int dev = 0, verbosity = 0, relaxed = 0, mode = PC_BATCHED;
    int FS = 0, N0 = 0, nIn = 0, nOut = 0;
    bool ready = false;
    std::vector<Inst> inst;
    std::vector<int> gsize, gshift;  // per global level: k_g (frames), log2
    char *arena = nullptr;
    size_t arena_bytes = 0;
    cf *tw = nullptr;
    hipStream_t s_main = nullptr;
    std::vector<hipStream_t> s_lvl;   // indexed by stream slot
    std::vector<int> g_stream;        // global level -> stream slot
    std::vector<hipEvent_t> ev_lvl;   // per global level
    std::vector<bool> lvl_pending;
    std::unique_ptr<LevelLauncher> lau;   // persistent runs: level launches off the tick thread
    std::vector<bool> lvl_async;          // level g's pending block was queued to it
    std::vector<uint64_t> lvl_start;      // start tick of level g's pending block
    hipEvent_t ev_tick = nullptr;
    std::vector<Group> groups;
    int mix = 1;
    float *y_inst = nullptr;
    int *d_src_off = nullptr, *d_src_list = nullptr;
    unsigned *d_ticket = nullptr;
    unsigned *d_ticket_ch = nullptr;     // per output channel (reduce mix)
    unsigned *d_ticket_grp = nullptr;  // per group of PC_TICKET_GROUP workgroups + global (direct mix)
    size_t ticket_grp_bytes = 0;
    int *d_empty = nullptr;              // output channels without a source
    int nactive = 0, nempty = 0;
    float *h_in = nullptr, *h_out = nullptr;
    unsigned long long *h_flag = nullptr;
    float *d_in_map = nullptr, *d_out_map = nullptr;   // device views of the pinned buffers
    unsigned long long *d_flag_map = nullptr;
    // release counter of the persistent kernel: a device-memory slot written
    // through the BAR (go_slot >= 0), else word PC_GO_WORD of h_flag
    unsigned long long *go_h = nullptr, *go_d = nullptr;
    int go_slot = -1;
    // per-workgroup flags (direct mix): persistent runs use nwg flags
    // PC_WG_FLAG_STRIDE words apart from h_wg, per-tick launches nwg packed
    // flags after them; wg_persist says which set the current tick uses
    unsigned long long *h_wg = nullptr, *d_wg_map = nullptr;
    bool wg_persist = false;
    // tagged input words of host sessions (see ll_tag): written by the host
    // into device memory through the BAR; ll_active while a session uses them
    unsigned long long *ll_in = nullptr;
    bool ll_active = false;
    int nwg = 0;
    bool use_wgflags = false;   // direct mix: one host flag per workgroup (up to PC_WG_FLAGS_MAX)
    int wg_scan = 0;            // tick_done: first workgroup not yet seen done for wg_scan_key
    unsigned long long wg_scan_key = ~0ull;
    uint64_t frame = 0;
    // kernel probe: HIP events around one kernel of one global level
    int probe_g = -1;
    unsigned long long *d_trace = nullptr;   // persistent-kernel timestamps (optional)
    long long trace_cap = 0;
    std::vector<hipEvent_t> pa, pb;
    std::vector<bool> pused;
    int pnext = 0;
    double probe_ms = 0, probe_bytes = 0;
    long long probe_n = 0;
    // level workers of persistent runs (short-period small levels, batched engine)
    std::vector<char> g_worker;       // per global level: served by workers in persistent runs
    std::vector<WorkJob> wjobs_h;
    WorkJob *d_wjobs = nullptr;
    int nwjobs = 0, lds_w = 0;        // worker-level jobs, largest worker block size
    int tw_lds = 0;                   // persistent runs with workers: twiddle-table prefix staged in LDS
    int nworkers = -1;                // worker workgroups (-1: not planned yet)
    bool self_levels = false;         // no room for workers: tick workgroups run those levels
    long long persist_budget = 0;     // workgroups a persistent kernel may take (half the device)
    int2 *d_wrange = nullptr;
    unsigned long long *d_lvl_done = nullptr, *d_dev_tick = nullptr;
    bool workers_active = false;      // a persistent run with workers owns those levels
    bool sess_active = false;         // a persistent kernel is serving ticks
    uint64_t sess_end = 0;            // first tick the session's kernel does not serve
    double sess_last = 0;             // host time of the session's last release
    bool sess_served_late = false;    // persist_end accounted for a tick served while stopping
    int host_session = 1;             // pcm_run through a persistent session (batched engine);
                                      // 2: without tagged words
    int sess_host = 0;                // the active session serves host buffers
    int persist_workers = 1;          // short levels in persistent runs: 1 in-kernel, 2 self mode, 0 host
    double pend_t0 = -1;              // submit time of a tick a stalled session left unserved (-1: none)
    int release_window = 1;           // device-resident persistent runs: ticks in flight (pcm_set_release_window)
    uint64_t win_rel = 0;             // release-window run: first tick not released (0: none released)
    std::vector<double> win_trel;     // release-window run: release time of tick f at f % window
    int ncu = 0;                      // compute units of the device
    int part_q = 1;                   // k_partial float4 groups per thread (pcm_set_partial_width)
    bool level_graphs = false;        // level blocks launched as graphs (pcm_set_level_graphs)
    std::vector<LevelGraph> lgraph;   // per global level, built at first use
    Residency resid;                  // residency stamps of plain persistent launches
    // held level (pcm_stall_level): the next block of global level stall_g
    // waits behind a host function that sleeps stall_us on the level's stream
    std::atomic<int> stall_g{-1};
    std::atomic<long long> stall_us{0};
    // stats
    std::vector<double> lat;
    size_t lat_pos = 0;
    bool lat_wrap = false;
    unsigned long long ticks = 0, misses = 0, stalls = 0;
    double lat_sum = 0, lat_max = 0;
};

static hipStream_t level_stream(Engine &E, int g) { return E.s_lvl[E.g_stream[g]]; }
static int persist_end(Engine &E, bool stop);
static void launcher_stop(Engine &E);

// Frame layouts of a tick: the caller's interleaved device buffers, or the
// engine's pinned host staging, which is planar (channel c's N0 samples
// contiguous) so each workgroup reads and writes whole 64-byte lines across the
// bus; run() converts to and from the caller's interleaved buffers on the host.
static void set_io_layout(const Engine &E, TickArgs &a, bool in_staging, bool out_staging)
{
    a.in_cs = in_staging ? E.N0 : 1;
    a.in_ss = in_staging ? 1 : E.nIn;
    a.out_cs = out_staging ? E.N0 : 1;
    a.out_ss = out_staging ? 1 : E.nOut;
}

// Words of the per-workgroup flag block: the persistent set, then the packed set.
static size_t wg_flag_words(const Engine &E)
{
    return (size_t)std::max(E.nwg, 1) * (PC_WG_FLAG_STRIDE + 1);
}

static void level_graphs_free(Engine &E)
{
    for (LevelGraph &G : E.lgraph) {
        if (G.x) hipGraphExecDestroy(G.x);
        if (G.g) hipGraphDestroy(G.g);
    }
    E.lgraph.clear();
}

static void engine_free(Engine &E)
{
    if (E.sess_active) (void)persist_end(E, true);
    launcher_stop(E);
    if (E.s_main) hipStreamSynchronize(E.s_main);
    for (hipStream_t s : E.s_lvl) hipStreamSynchronize(s);
    hipFree(E.d_wjobs);
    hipFree(E.d_lvl_done);
    hipFree(E.d_dev_tick);
    E.d_wjobs = nullptr;
    E.d_lvl_done = E.d_dev_tick = nullptr;
    hipFree(E.d_wrange);
    E.d_wrange = nullptr;
    E.nworkers = -1;
    E.nwjobs = 0;
    E.wjobs_h.clear();
    E.g_worker.clear();
    for (Group &G : E.groups) {
        hipFree(G.d_tick);
        hipFree(G.d_l0);
        for (Batch &b : G.batches) hipFree(b.d_jobs);
    }
    E.groups.clear();
    hipFree(E.arena);
    E.arena = nullptr;
    hipFree(E.y_inst);
    hipFree(E.d_src_off);
    hipFree(E.d_src_list);
    hipFree(E.d_ticket);
    hipFree(E.d_ticket_ch);
    hipFree(E.d_ticket_grp);
    E.d_ticket_grp = nullptr;
    hipFree(E.d_empty);
    E.d_ticket_ch = nullptr;
    E.d_empty = nullptr;
    E.y_inst = nullptr;
    E.d_src_off = E.d_src_list = nullptr;
    E.d_ticket = nullptr;
    if (E.h_in) hipHostFree(E.h_in);
    if (E.h_out) hipHostFree(E.h_out);
    if (E.h_flag) hipHostFree(E.h_flag);
    release_slot_put(E.dev, E.go_slot);
    E.go_slot = -1;
    E.go_h = E.go_d = nullptr;
    if (E.h_wg) hipHostFree(E.h_wg);
    bar_free(E.dev, E.ll_in, (size_t)E.nIn * E.N0 * sizeof(unsigned long long));
    E.ll_in = nullptr;
    E.ll_active = false;
    E.h_wg = E.d_wg_map = nullptr;
    E.h_in = E.h_out = nullptr;
    E.h_flag = nullptr;
    for (hipEvent_t e : E.ev_lvl) hipEventDestroy(e);
    E.ev_lvl.clear();
    hipFree(E.d_trace);
    E.d_trace = nullptr;
    E.trace_cap = 0;
    for (hipEvent_t e : E.pa) hipEventDestroy(e);
    for (hipEvent_t e : E.pb) hipEventDestroy(e);
    E.pa.clear();
    E.pb.clear();
    E.pused.clear();
    E.probe_g = -1;
    if (E.ev_tick) hipEventDestroy(E.ev_tick);
    E.ev_tick = nullptr;
    resident_free(E.resid);
    level_graphs_free(E);
    for (hipStream_t s : E.s_lvl) hipStreamDestroy(s);
    E.s_lvl.clear();
    if (E.s_main) hipStreamDestroy(E.s_main);
    E.s_main = nullptr;
    E.inst.clear();
    E.ready = false;
}

// ---- kernel dispatch --------------------------------------------------------
// Optional launch decorations: HIP events bound to the kernel's own start/stop
// (hipExtLaunchKernelGGL) and a level-completion counter.
struct LaunchExtra {
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
};

template <int LOGN>
static void launch_small_t(const LevelJob *jobs, int njobs, int N0, uint64_t m, uint64_t bcur,
                           int ypar, const cf *tw, int mode, hipStream_t s, const LaunchExtra &x)
{
    hipExtLaunchKernelGGL(k_level_small<LOGN>, dim3(njobs), dim3(SmallCfg<LOGN>::NT), 0, s, x.ev0, x.ev1, 0,
                          jobs, N0, m, bcur, ypar, tw, mode);
}

static int launch_small(int N, const LevelJob *jobs, int njobs, int N0, uint64_t m, uint64_t bcur,
                        int ypar, const cf *tw, int mode, hipStream_t s, const LaunchExtra &x = LaunchExtra())
{
    switch (ilog2(N)) {
    case 4: launch_small_t<4>(jobs, njobs, N0, m, bcur, ypar, tw, mode, s, x); break;
    case 5: launch_small_t<5>(jobs, njobs, N0, m, bcur, ypar, tw, mode, s, x); break;
    case 6: launch_small_t<6>(jobs, njobs, N0, m, bcur, ypar, tw, mode, s, x); break;
    case 7: launch_small_t<7>(jobs, njobs, N0, m, bcur, ypar, tw, mode, s, x); break;
    case 8: launch_small_t<8>(jobs, njobs, N0, m, bcur, ypar, tw, mode, s, x); break;
    case 9: launch_small_t<9>(jobs, njobs, N0, m, bcur, ypar, tw, mode, s, x); break;
    case 10: launch_small_t<10>(jobs, njobs, N0, m, bcur, ypar, tw, mode, s, x); break;
    case 11: launch_small_t<11>(jobs, njobs, N0, m, bcur, ypar, tw, mode, s, x); break;
    default: return fail(PC_ERR_ARG, "unsupported small block size");
    }
    return PC_OK;
}

static size_t big1_lds(BigShape sh) { return (size_t)2 * BIG_G * ((1 << sh.logB) + 1) * sizeof(cf); }
static size_t big2_lds(BigShape sh) { return (size_t)2 * 2 * ((1 << sh.logA) + 1) * sizeof(cf); }

// Four-step level job: k_big1 -> k_big2 (multiply-accumulate; optional
// events) -> k_big3 (optional level-completion counter).
static void launch_big(BigShape sh, const LevelJob *jobs, int njobs, int N0, uint64_t m,
                       uint64_t bcur, int ypar, const cf *tw, int mode, hipStream_t s,
                       const LaunchExtra &x = LaunchExtra())
{
    const int A = 1 << sh.logA, B = 1 << sh.logB;
    hipLaunchKernelGGL(k_big1, dim3((A / BIG_G) * njobs), dim3(BIG_NT), big1_lds(sh), s, jobs, sh, N0,
                       m, tw, mode);
    hipExtLaunchKernelGGL(k_big2, dim3((B / 2) * njobs), dim3(BIG_NT), (uint32_t)big2_lds(sh), s, x.ev0,
                          x.ev1, 0, jobs, sh, bcur, tw, mode);
    if (mode == PC_MODE_RUN)
        hipLaunchKernelGGL(k_big3, dim3((A / BIG_G) * njobs), dim3(BIG_NT), big1_lds(sh), s, jobs, sh,
                           ypar, tw);
}

// Partial sum S for block b of every job of a batch (k_partial<Q>, Q float4
// groups per thread; the result does not depend on Q).
static void launch_partial(int N, const LevelJob *jobs, int njobs, uint64_t b, hipStream_t s,
                           const LaunchExtra &x = LaunchExtra(), int Q = 1)
{
    const int wpj = (N / 2 + PART_NT * Q - 1) / (PART_NT * Q);
    if (Q == 4)
        hipExtLaunchKernelGGL(k_partial<4>, dim3(wpj * njobs), dim3(PART_NT), 0, s, x.ev0, x.ev1, 0, jobs, N, wpj, b);
    else if (Q == 2)
        hipExtLaunchKernelGGL(k_partial<2>, dim3(wpj * njobs), dim3(PART_NT), 0, s, x.ev0, x.ev1, 0, jobs, N, wpj, b);
    else
        hipExtLaunchKernelGGL(k_partial<1>, dim3(wpj * njobs), dim3(PART_NT), 0, s, x.ev0, x.ev1, 0, jobs, N, wpj, b);
}

template <int LOGN>
static void launch_tick_t(const TickArgs &a, const cf *tw, hipStream_t s, const LaunchExtra &x)
{
    hipExtLaunchKernelGGL(k_tick<LOGN>, dim3(a.njobs), dim3(TickCfg<LOGN>::NT), 0, s, x.ev0, x.ev1, 0, a, tw);
}

static int launch_tick(int N0, const TickArgs &a, const cf *tw, hipStream_t s,
                       const LaunchExtra &x = LaunchExtra())
{
    switch (ilog2(N0)) {
    case 4: launch_tick_t<4>(a, tw, s, x); break;
    case 5: launch_tick_t<5>(a, tw, s, x); break;
    case 6: launch_tick_t<6>(a, tw, s, x); break;
    case 7: launch_tick_t<7>(a, tw, s, x); break;
    case 8: launch_tick_t<8>(a, tw, s, x); break;
    case 9: launch_tick_t<9>(a, tw, s, x); break;
    case 10: launch_tick_t<10>(a, tw, s, x); break;
    case 11: launch_tick_t<11>(a, tw, s, x); break;
    default: return fail(PC_ERR_ARG, "unsupported buffer_size");
    }
    return PC_OK;
}

// ring write for the caller-driven filter API
__global__ void k_ring_write(float *ring, long long mask, long long t0, const float *x, int n)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < n) {
        const long long p = (t0 + j) & mask;
        ring[p] = x[j];
        ring[p + mask + 1] = x[j];
    }
}

// y = level-0 block + sum_L level L sub-block (filter API output)
struct FilterOut {
    const float *src[PC_MAX_LEVELS];
    int nsrc;
};
__global__ void k_filter_out(FilterOut fo, int N0, float *y)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < N0) {
        float a = 0.f;
        for (int q = 0; q < fo.nsrc; q++) a += fo.src[q][j];
        y[j] = a;
    }
}

// ----------------------------------------------------------------------------
// Setup
// ----------------------------------------------------------------------------
static size_t align_up(size_t x) { return (x + 255) & ~(size_t)255; }

static int engine_setup(Engine &E, int FS, int N0, int nIn, int nOut, std::vector<InstDesc> &descs)
{
    if (E.ready) return fail(PC_ERR_STATE, "setup called twice without cleanup");
    if (descs.empty()) return fail(PC_ERR_ARG, "no instances");
    if (!is_pow2(N0) || N0 < PC_MIN_BLOCK || N0 > 2048)
        return fail(PC_ERR_ARG, "buffer_size must be a power of two in [16, 2048]");
    if (nIn < 1 || nOut < 1) return fail(PC_ERR_ARG, "need at least one input and output channel");
    E.FS = FS;
    E.N0 = N0;
    E.nIn = nIn;
    E.nOut = nOut;
    E.inst.clear();
    int max_block = N0;
    for (size_t i = 0; i < descs.size(); i++) {
        InstDesc &d = descs[i];
        if (d.in_ch < 0 || d.in_ch >= nIn || d.out_ch < 0 || d.out_ch >= nOut)
            return fail(PC_ERR_CONFIG, "instance " + std::to_string(i) + ": channel out of range");
        if (d.N.empty() || d.N[0] != N0) {
            char m[96];
            snprintf(m, sizeof m, "line %zu: first block size must equal buffer_size", i + 1);
            return fail(PC_ERR_CONFIG, m);
        }
        Inst I;
        std::vector<int> delay, T;
        std::vector<int> N = d.N, P = d.P;
        int rc = check_partition(N, P, (long long)d.ir.size(), delay, T);
        if (rc) return rc;
        for (int n : N)
            if (n > PC_MAX_BLOCK) return fail(PC_ERR_PARTITION, "block size above 65536");
        d.P = P;
        I.d = d;
        long long need = 0;
        for (size_t L = 0; L < N.size(); L++) {
            LevelSpec s{};
            s.N = N[L];
            s.P = P[L];
            s.T = T[L];
            s.delay = delay[L];
            s.kshift = ilog2(N[L] / N0);
            s.big = (L > 0) && N[L] >= PC_BIG_MIN;
            if (s.big) s.sh = big_shape(N[L]);
            I.lv.push_back(s);
            need = std::max(need, std::max(2ll * N[L], (long long)N[L] + N0 + T[L]));
            max_block = std::max(max_block, N[L]);
        }
        I.R = (int)(1ll << ilog2(need + N0));
        E.inst.push_back(I);
    }
    // every partition checked (errors as the reference's, without a device): now the GPU
    HIPCHK(hipGetDevice(&E.dev));
    hip_live();
    // global level set (src/partconvMulti.cpp:139-244)
    const int temp_levels = ilog2(max_block / N0) + 1;
    E.gsize.clear();
    E.gshift.clear();
    for (int g = 0; g < temp_levels; g++) {
        const int bs = N0 << g;
        bool used = false;
        for (Inst &I : E.inst)
            for (LevelSpec &s : I.lv) used |= (s.N == bs);
        if (!used) continue;
        E.gsize.push_back(1 << g);
        E.gshift.push_back(g);
    }
    const int G = (int)E.gsize.size();
    for (Inst &I : E.inst) {
        I.gmap.assign(G, -1);
        for (int g = 0; g < G; g++)
            for (size_t L = 0; L < I.lv.size(); L++)
                if (I.lv[L].N == (N0 << E.gshift[g])) I.gmap[g] = (int)L;
    }
    // arena layout
    size_t off = 0;
    for (Inst &I : E.inst) {
        I.off_ring = off;
        off = align_up(off + (size_t)2 * I.R * sizeof(float));
        for (LevelSpec &s : I.lv) {
            s.off_H = off;
            off = align_up(off + (size_t)s.P * s.N * sizeof(cf));
            s.off_fdl = off;
            off = align_up(off + (size_t)s.P * s.N * sizeof(cf));
            s.off_ybuf = off;
            off = align_up(off + (size_t)2 * s.N * sizeof(float));
            s.off_scr = off;
            if (s.big) off = align_up(off + (size_t)s.N * sizeof(cf));
            s.off_S = off;
            off = align_up(off + (size_t)s.N * sizeof(cf));
        }
    }
    E.arena_bytes = off;
    HIPCHK(hipMalloc(&E.arena, E.arena_bytes));
    HIPCHK(hipMemset(E.arena, 0, E.arena_bytes));
    int rc = get_twiddles(E.dev, &E.tw);
    if (rc) return rc;
    HIPCHK(hipStreamCreateWithFlags(&E.s_main, hipStreamNonBlocking));

    // ---- filter spectra H[L][p] = r2c([0_N | h_{L,p}]) / 2N (src/partconv.cpp:199-230)
    {
        size_t tot = 0;
        std::vector<size_t> ir_off(E.inst.size());
        for (size_t i = 0; i < E.inst.size(); i++) {
            ir_off[i] = tot;
            tot += E.inst[i].d.ir.size();
        }
        float *d_ir = nullptr;
        HIPCHK(hipMalloc(&d_ir, std::max<size_t>(tot, 1) * sizeof(float)));
        for (size_t i = 0; i < E.inst.size(); i++)
            if (!E.inst[i].d.ir.empty())
                HIPCHK(hipMemcpy(d_ir + ir_off[i], E.inst[i].d.ir.data(),
                                 E.inst[i].d.ir.size() * sizeof(float), hipMemcpyHostToDevice));
        const int CHUNK = 256;
        cf *d_scr = nullptr;
        LevelJob *d_jobs = nullptr;
        HIPCHK(hipMalloc(&d_jobs, CHUNK * sizeof(LevelJob)));
        size_t scr_elems = 0;
        for (int g = 0; g < G; g++) {
            const int N = N0 << E.gshift[g];
            std::vector<LevelJob> jobs;
            for (size_t i = 0; i < E.inst.size(); i++) {
                Inst &I = E.inst[i];
                const int L = I.gmap[g];
                if (L < 0) continue;
                LevelSpec &s = I.lv[L];
                for (int p = 0; p < s.P; p++) {
                    LevelJob J{};
                    J.ring = d_ir + ir_off[i];
                    J.ring_mask = (int)I.d.ir.size();
                    J.T = s.T + p * s.N;
                    J.fdl = reinterpret_cast<cf *>(E.arena + s.off_H) + (size_t)p * s.N;
                    J.P = 1;
                    jobs.push_back(J);
                }
            }
            const bool big = (g > 0) && N >= PC_BIG_MIN;
            if (big && scr_elems < (size_t)CHUNK * N) {
                hipFree(d_scr);
                scr_elems = (size_t)CHUNK * N;
                HIPCHK(hipMalloc(&d_scr, scr_elems * sizeof(cf)));
            }
            for (size_t c0 = 0; c0 < jobs.size(); c0 += CHUNK) {
                const int n = (int)std::min<size_t>(CHUNK, jobs.size() - c0);
                for (int q = 0; q < n; q++)
                    if (big) jobs[c0 + q].scratch = d_scr + (size_t)q * N;
                HIPCHK(hipMemcpyAsync(d_jobs, jobs.data() + c0, n * sizeof(LevelJob),
                                      hipMemcpyHostToDevice, E.s_main));
                if (big) launch_big(big_shape(N), d_jobs, n, N0, 0, 0, 0, E.tw, PC_MODE_SPECTRUM, E.s_main);
                else {
                    rc = launch_small(N, d_jobs, n, N0, 0, 0, 0, E.tw, PC_MODE_SPECTRUM, E.s_main);
                    if (rc) return rc;
                }
                HIPCHK(hipGetLastError());
                HIPCHK(hipStreamSynchronize(E.s_main));
            }
        }
        hipFree(d_scr);
        hipFree(d_jobs);
        hipFree(d_ir);
    }

    // ---- launch groups: batched = one group; plain = one per instance
    std::vector<std::vector<int>> members;
    if (E.mode == PC_BATCHED) {
        members.emplace_back();
        for (size_t i = 0; i < E.inst.size(); i++) members[0].push_back((int)i);
    } else {
        for (size_t i = 0; i < E.inst.size(); i++) members.push_back({(int)i});
    }
    auto mk_level_job = [&](Inst &I, LevelSpec &s) {
        LevelJob J{};
        J.ring = reinterpret_cast<const float *>(E.arena + I.off_ring);
        J.fdl = reinterpret_cast<cf *>(E.arena + s.off_fdl);
        J.H = reinterpret_cast<const cf *>(E.arena + s.off_H);
        J.ybuf = reinterpret_cast<float *>(E.arena + s.off_ybuf);
        J.scratch = s.big ? reinterpret_cast<cf *>(E.arena + s.off_scr) : nullptr;
        J.S = reinterpret_cast<cf *>(E.arena + s.off_S);
        J.P = s.P;
        J.T = s.T;
        J.ring_mask = I.R - 1;
        return J;
    };
    // levels served by in-kernel workers during persistent runs: period <= 8
    // ticks and a single-workgroup block (N <= 2048) of at most 8 bins per
    // thread of the persistent kernel (its thread count is TickCfg's; worker_fits)
    E.g_worker.assign(G, 0);
    const int tick_nt = (N0 <= 64) ? 64 : ((N0 >= 512 || N0 == 128) ? 256 : 128);  // TickCfg<log2 N0>::NT
    for (int g = 1; g < G; g++) {
        const int Ng = N0 << E.gshift[g];
        E.g_worker[g] = (E.gsize[g] <= 8 && Ng < PC_BIG_MIN && Ng / tick_nt <= 8) ? 1 : 0;
    }
    for (auto &mem : members) {
        Group Gr;
        Gr.members = mem;
        const bool batched = (E.mode == PC_BATCHED);
        std::vector<std::vector<int>> wslot(E.inst.size(), std::vector<int>(G, -1));
        for (int g = 1; g < G; g++) {
            std::vector<LevelJob> js;
            std::vector<int> pos;
            for (size_t q = 0; q < mem.size(); q++) {
                Inst &I = E.inst[mem[q]];
                if (I.gmap[g] >= 0) {
                    js.push_back(mk_level_job(I, I.lv[I.gmap[g]]));
                    pos.push_back((int)q);
                }
            }
            if (js.empty()) continue;
            Batch b;
            b.g = g;
            b.N = N0 << E.gshift[g];
            b.big = b.N >= PC_BIG_MIN;
            if (b.big) b.sh = big_shape(b.N);
            b.njobs = (int)js.size();
            HIPCHK(hipMalloc(&b.d_jobs, js.size() * sizeof(LevelJob)));
            HIPCHK(hipMemcpy(b.d_jobs, js.data(), js.size() * sizeof(LevelJob), hipMemcpyHostToDevice));
            if (batched && E.g_worker[g]) {
                for (size_t j = 0; j < js.size(); j++) {
                    WorkJob w{};
                    w.J = js[j];
                    w.inst = pos[j];
                    w.kshift = E.gshift[g];
                    w.logN = ilog2(b.N);
                    w.slot = (int)E.wjobs_h.size();
                    wslot[mem[pos[j]]][g] = w.slot;
                    E.wjobs_h.push_back(w);
                    E.lds_w = std::max(E.lds_w, b.N);
                }
            }
            Gr.batches.push_back(b);
        }
        std::vector<TickJob> tj;
        std::vector<LevelJob> l0;
        for (int i : mem) {
            Inst &I = E.inst[i];
            TickJob t{};
            t.l0 = mk_level_job(I, I.lv[0]);
            t.in_ch = I.d.in_ch;
            t.out_ch = I.d.out_ch;
            t.mute = I.d.mute;
            t.scale = I.d.scale;
            t.nlev = (int)I.lv.size() - 1;
            for (int L = 0; L < PC_MAX_LEVELS; L++) t.wslot[L] = -1;
            for (int g = 1; g < G; g++) {
                const int L = I.gmap[g];
                if (L < 1) continue;
                t.kshift[L - 1] = I.lv[L].kshift;
                t.NL[L - 1] = I.lv[L].N;
                t.ybuf[L - 1] = reinterpret_cast<const float *>(E.arena + I.lv[L].off_ybuf);
                t.wslot[L - 1] = wslot[i][g];
            }
            tj.push_back(t);
            l0.push_back(t.l0);
        }
        HIPCHK(hipMalloc(&Gr.d_tick, tj.size() * sizeof(TickJob)));
        HIPCHK(hipMemcpy(Gr.d_tick, tj.data(), tj.size() * sizeof(TickJob), hipMemcpyHostToDevice));
        HIPCHK(hipMalloc(&Gr.d_l0, l0.size() * sizeof(LevelJob)));
        HIPCHK(hipMemcpy(Gr.d_l0, l0.data(), l0.size() * sizeof(LevelJob), hipMemcpyHostToDevice));
        E.groups.push_back(Gr);
    }
    E.nwjobs = (int)E.wjobs_h.size();
    E.nworkers = -1;
    if (E.nwjobs > 0) {
        HIPCHK(hipMalloc(&E.d_wjobs, E.wjobs_h.size() * sizeof(WorkJob)));
        HIPCHK(hipMemcpy(E.d_wjobs, E.wjobs_h.data(), E.wjobs_h.size() * sizeof(WorkJob), hipMemcpyHostToDevice));
        HIPCHK(hipMalloc(&E.d_lvl_done, E.wjobs_h.size() * sizeof(unsigned long long)));
        HIPCHK(hipMalloc(&E.d_dev_tick, E.inst.size() * sizeof(unsigned long long)));
    }
    // ---- output mix routing
    std::vector<int> src_off(nOut + 1, 0), src_list, empty;
    bool direct = (E.mode == PC_BATCHED);
    E.nactive = 0;
    for (int c = 0; c < nOut; c++) {
        int cnt = 0;
        for (size_t i = 0; i < E.inst.size(); i++)
            if (!E.inst[i].d.mute && E.inst[i].d.out_ch == c) {
                src_list.push_back((int)i);
                cnt++;
            }
        src_off[c + 1] = (int)src_list.size();
        if (cnt != 1) direct = false;
        if (cnt) E.nactive++;
        else empty.push_back(c);
    }
    E.nempty = (int)empty.size();
    HIPCHK(hipMalloc(&E.d_ticket_ch, std::max(nOut, 1) * sizeof(unsigned)));
    HIPCHK(hipMemset(E.d_ticket_ch, 0, std::max(nOut, 1) * sizeof(unsigned)));
    HIPCHK(hipMalloc(&E.d_empty, std::max<size_t>(empty.size(), 1) * sizeof(int)));
    if (!empty.empty())
        HIPCHK(hipMemcpy(E.d_empty, empty.data(), empty.size() * sizeof(int), hipMemcpyHostToDevice));
    // 0: per-instance outputs + k_mix (plain engine, or nothing unmuted to mix)
    E.mix = (E.mode == PC_PLAIN) ? 0 : (direct ? 1 : (E.nactive > 0 ? 2 : 0));
    HIPCHK(hipMalloc(&E.y_inst, E.inst.size() * (size_t)N0 * sizeof(float)));
    HIPCHK(hipMalloc(&E.d_src_off, src_off.size() * sizeof(int)));
    HIPCHK(hipMemcpy(E.d_src_off, src_off.data(), src_off.size() * sizeof(int), hipMemcpyHostToDevice));
    HIPCHK(hipMalloc(&E.d_src_list, std::max<size_t>(src_list.size(), 1) * sizeof(int)));
    if (!src_list.empty())
        HIPCHK(hipMemcpy(E.d_src_list, src_list.data(), src_list.size() * sizeof(int), hipMemcpyHostToDevice));
    HIPCHK(hipMalloc(&E.d_ticket, 64));
    HIPCHK(hipMemset(E.d_ticket, 0, 64));
    E.ticket_grp_bytes = ((E.inst.size() + PC_TICKET_GROUP - 1) / PC_TICKET_GROUP + 1) * PC_TICKET_STRIDE *
                         sizeof(unsigned);
    HIPCHK(hipMalloc(&E.d_ticket_grp, E.ticket_grp_bytes));
    HIPCHK(hipMemset(E.d_ticket_grp, 0, E.ticket_grp_bytes));

    // ---- streams, events, pinned staging
    // Level streams get hardware queues of their own (a stream with a CU mask,
    // here every CU, is never placed on a shared queue): level blocks launched
    // while the persistent tick kernel occupies s_main must not queue behind it.
    const int nls = std::min(PC_LEVEL_STREAMS, std::max(0, G - 1));
    hipDeviceProp_t prop;
    HIPCHK(hipGetDeviceProperties(&prop, E.dev));
    E.ncu = prop.multiProcessorCount;
    std::vector<uint32_t> cu_mask((prop.multiProcessorCount + 31) / 32, 0u);
    for (int c = 0; c < prop.multiProcessorCount; c++) cu_mask[c / 32] |= 1u << (c % 32);
    for (int q = 0; q < nls; q++) {
        hipStream_t s;
        HIPCHK(hipExtStreamCreateWithCUMask(&s, (uint32_t)cu_mask.size(), cu_mask.data()));
        E.s_lvl.push_back(s);
    }
    E.g_stream.assign(G, 0);
    for (int g = 1; g < G; g++) E.g_stream[g] = (g - 1) % std::max(1, nls);
    E.ev_lvl.assign(G, nullptr);
    E.lvl_pending.assign(G, false);
    E.lvl_async.assign(G, false);
    E.lvl_start.assign(G, 0);
    for (int g = 0; g < G; g++) HIPCHK(hipEventCreateWithFlags(&E.ev_lvl[g], hipEventDisableTiming));
    HIPCHK(hipEventCreateWithFlags(&E.ev_tick, hipEventDisableTiming));
    HIPCHK(hipHostMalloc(&E.h_in, (size_t)N0 * nIn * sizeof(float), hipHostMallocCoherent | hipHostMallocMapped));
    HIPCHK(hipHostMalloc(&E.h_out, (size_t)N0 * std::max(nOut, 1) * sizeof(float),
                         hipHostMallocCoherent | hipHostMallocMapped));
    // completion flag at word 0, release counter at word PC_GO_WORD: one cache
    // line (measured faster than two separate lines)
    HIPCHK(hipHostMalloc(&E.h_flag, 64, hipHostMallocCoherent | hipHostMallocMapped));
    memset(E.h_flag, 0, 64);
    HIPCHK(hipHostGetDevicePointer((void **)&E.d_in_map, E.h_in, 0));
    HIPCHK(hipHostGetDevicePointer((void **)&E.d_out_map, E.h_out, 0));
    HIPCHK(hipHostGetDevicePointer((void **)&E.d_flag_map, E.h_flag, 0));
    E.go_slot = release_slot_get(E.dev, &E.go_h);
    if (E.go_slot >= 0) {
        E.go_d = E.go_h;  // one address for host and device
    } else {
        E.go_h = E.h_flag + PC_GO_WORD;
        E.go_d = E.d_flag_map + PC_GO_WORD;
    }
    E.nwg = (E.mode == PC_BATCHED) ? (int)E.inst.size() : 0;
    E.use_wgflags = (E.mix == 1 && E.nwg <= PC_WG_FLAGS_MAX);
    if (E.use_wgflags) {
        HIPCHK(hipHostMalloc(&E.h_wg, wg_flag_words(E) * sizeof(unsigned long long),
                             hipHostMallocCoherent | hipHostMallocMapped));
        memset(E.h_wg, 0, wg_flag_words(E) * sizeof(unsigned long long));
        E.wg_scan_key = ~0ull;
        HIPCHK(hipHostGetDevicePointer((void **)&E.d_wg_map, E.h_wg, 0));
    }
    if (E.mode == PC_BATCHED && E.mix != 0 && (size_t)nIn * N0 <= PC_LL_MAX_WORDS)
        E.ll_in = static_cast<unsigned long long *>(bar_alloc(E.dev, (size_t)nIn * N0 * sizeof(unsigned long long)));
    memset(E.h_in, 0, (size_t)N0 * nIn * sizeof(float));
    E.frame = 0;
    E.lat.assign(1 << 20, 0.0);
    E.lat_pos = 0;
    E.lat_wrap = false;
    E.ticks = E.misses = 0;
    E.lat_sum = E.lat_max = 0;
    E.ready = true;
    if (E.verbosity > 0) {
        for (size_t i = 0; i < E.inst.size(); i++) {
            Inst &I = E.inst[i];
            printf("\nPartitioning Specs (instance %zu):\nimpulse response: %.2f sec\npartitioning: ", i,
                   (double)I.d.ir.size() / (FS > 0 ? FS : 1));
            for (size_t L = 0; L < I.lv.size(); L++)
                printf("%ux%u%s", I.lv[L].N, I.lv[L].P, L + 1 < I.lv.size() ? ", " : "\n");
            printf("filter latency: %g ms\n", FS > 0 ? 1000.0 * N0 / FS : 0.0);
        }
        printf("Total PartConv instances   : %zu\nbuffer_size: %d\nmax_block_size: %d\n", E.inst.size(), N0,
               max_block);
    }
    return PC_OK;
}

static int engine_reset(Engine &E)
{
    if (!E.ready) return fail(PC_ERR_STATE, "reset before setup");
    if (E.sess_active) {
        int rc = persist_end(E, true);
        if (rc) return rc;
    }
    HIPCHK(hipStreamSynchronize(E.s_main));
    for (hipStream_t s : E.s_lvl) HIPCHK(hipStreamSynchronize(s));
    for (Inst &I : E.inst) {
        HIPCHK(hipMemsetAsync(E.arena + I.off_ring, 0, (size_t)2 * I.R * sizeof(float), E.s_main));
        for (LevelSpec &s : I.lv) {
            HIPCHK(hipMemsetAsync(E.arena + s.off_fdl, 0, (size_t)s.P * s.N * sizeof(cf), E.s_main));
            HIPCHK(hipMemsetAsync(E.arena + s.off_ybuf, 0, (size_t)2 * s.N * sizeof(float), E.s_main));
            HIPCHK(hipMemsetAsync(E.arena + s.off_S, 0, (size_t)s.N * sizeof(cf), E.s_main));
        }
    }
    HIPCHK(hipMemsetAsync(E.d_ticket, 0, 64, E.s_main));
    HIPCHK(hipMemsetAsync(E.d_ticket_ch, 0, std::max(E.nOut, 1) * sizeof(unsigned), E.s_main));
    HIPCHK(hipMemsetAsync(E.d_ticket_grp, 0, E.ticket_grp_bytes, E.s_main));
    HIPCHK(hipStreamSynchronize(E.s_main));
    E.frame = 0;
    std::fill(E.lvl_pending.begin(), E.lvl_pending.end(), false);
    std::fill(E.lvl_async.begin(), E.lvl_async.end(), false);
    if (E.lau)  // frames rewind: so do the helper's progress marks
        for (size_t g = 0; g < E.gsize.size(); g++) E.lau->done[g].store(0);
    *E.h_flag = 0;
    if (E.h_wg) memset(E.h_wg, 0, wg_flag_words(E) * sizeof(unsigned long long));
    E.wg_scan_key = ~0ull;  // the flags restart from zero
    return PC_OK;
}

static void record_latency(Engine &E, double us)
{
    E.lat[E.lat_pos] = us;
    if (++E.lat_pos == E.lat.size()) {
        E.lat_pos = 0;
        E.lat_wrap = true;
    }
    E.ticks++;
    E.lat_sum += us;
    E.lat_max = std::max(E.lat_max, us);
    const double period = E.FS > 0 ? 1e6 * E.N0 / E.FS : 0.0;
    if (period > 0 && us > period) E.misses++;
}

// Completion of tick target-1: every per-workgroup flag (direct mix) or the
// single flag raised by the last workgroup / the mix kernel.
static bool tick_done(Engine &E, unsigned long long target)
{
    if (E.use_wgflags) {
        const size_t stride = E.wg_persist ? PC_WG_FLAG_STRIDE : 1;
        const volatile unsigned long long *w = E.h_wg + (E.wg_persist ? 0 : (size_t)E.nwg * PC_WG_FLAG_STRIDE);
        // flags only grow: a poll resumes at the first workgroup not yet seen
        // done for this target (and flag set)
        const unsigned long long key = 2 * target + (E.wg_persist ? 1 : 0);
        if (E.wg_scan_key != key) {
            E.wg_scan_key = key;
            E.wg_scan = 0;
        }
        for (int i = E.wg_scan; i < E.nwg; i++)
            if (w[(size_t)i * stride] < target) {
                E.wg_scan = i;
                return false;
            }
        E.wg_scan = E.nwg;
        return true;
    }
    return *(const volatile unsigned long long *)E.h_flag >= target;
}

// Spin until tick target-1 completes.  stall_ok: return 1 (instead of an
// error) when the tick stream has drained without completing it, or after the
// timeout, so the persistent runner can fall back.
static int wait_flag(Engine &E, unsigned long long target, bool stall_ok = false)
{
    if (tick_done(E, target)) return PC_OK;
    const double t0 = now_s();
    unsigned spins = 0;
    while (!tick_done(E, target)) {
        if ((++spins & 0x3FFF) == 0) {
            hipError_t q = hipStreamQuery(E.s_main);
            if (q != hipSuccess && q != hipErrorNotReady)
                return fail(PC_ERR_HIP, std::string("tick failed: ") + hipGetErrorString(q));
            if (q == hipSuccess && !tick_done(E, target)) {
                if (stall_ok) return 1;
                return fail(PC_ERR_STATE, "tick stream drained without completing the tick");
            }
            if (now_s() - t0 > PC_TICK_TIMEOUT_S) {
                if (stall_ok) return 1;
                return fail(PC_ERR_TIMEOUT, "tick did not complete");
            }
        }
    }
    return PC_OK;
}

// Probe bookkeeping: slot q's pair is folded into the running total before reuse.
static int probe_fold(Engine &E, int q)
{
    if (!E.pused[q]) return PC_OK;
    HIPCHK(hipEventSynchronize(E.pb[q]));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, E.pa[q], E.pb[q]));
    E.probe_ms += ms;
    E.probe_n++;
    E.pused[q] = false;
    return PC_OK;
}

// Reserve an event pair; the kernel launch itself stamps it
// (hipExtLaunchKernelGGL start/stop events), so the measured time is the
// kernel's own execution.
static int probe_take(Engine &E, LaunchExtra &x)
{
    const int q = E.pnext;
    E.pnext = (E.pnext + 1) % (int)E.pa.size();
    int rc = probe_fold(E, q);
    if (rc) return rc;
    x.ev0 = E.pa[q];
    x.ev1 = E.pb[q];
    E.pused[q] = true;
    return PC_OK;
}

// Launch one level batch (block m) on stream s; a probed level times its
// multiply-accumulate kernel (k_big2, or the fused small kernel).
static int launch_batch(Engine &E, const Batch &b, uint64_t m, hipStream_t s)
{
    LaunchExtra x;
    if (E.probe_g == b.g) {
        int rc = probe_take(E, x);
        if (rc) return rc;
    }
    if (b.big) launch_big(b.sh, b.d_jobs, b.njobs, E.N0, m, m, (int)(m & 1), E.tw, PC_MODE_RUN, s);
    else {
        int rc = launch_small(b.N, b.d_jobs, b.njobs, E.N0, m, m, (int)(m & 1), E.tw, PC_MODE_RUN, s);
        if (rc) return rc;
    }
    // the next block's delay-line sum, off its critical path
    launch_partial(b.N, b.d_jobs, b.njobs, m + 1, s, x, E.part_q);
    return PC_OK;
}

// Launch level g's blocks that start at tick f on its stream and record its event.
static void host_sleep_fn(void *us) { std::this_thread::sleep_for(std::chrono::microseconds((uintptr_t)us)); }

static const void *small_fn(int logn)
{
    switch (logn) {
    case 4: return (const void *)k_level_small<4>;
    case 5: return (const void *)k_level_small<5>;
    case 6: return (const void *)k_level_small<6>;
    case 7: return (const void *)k_level_small<7>;
    case 8: return (const void *)k_level_small<8>;
    case 9: return (const void *)k_level_small<9>;
    case 10: return (const void *)k_level_small<10>;
    case 11: return (const void *)k_level_small<11>;
    default: return nullptr;
    }
}

static int small_nt(int N) { return (N / 4 < 64) ? 64 : ((N / 4 > 512) ? 512 : N / 4); }  // SmallCfg::NT

// Append one kernel node after `dep` (or none); its arguments live in *a.
static int graph_node(LevelGraph &G, const void *fn, dim3 grid, int nt, size_t lds, std::unique_ptr<GArgs> a,
                      int nargs, hipGraphNode_t *dep)
{
    GNode n;
    n.a = std::move(a);
    n.p.func = const_cast<void *>(fn);
    n.p.gridDim = grid;
    n.p.blockDim = dim3(nt);
    n.p.sharedMemBytes = (unsigned)lds;
    n.p.kernelParams = n.a->ptrs;
    n.p.extra = nullptr;
    (void)nargs;
    HIPCHK(hipGraphAddKernelNode(&n.node, G.g, *dep ? dep : nullptr, *dep ? 1 : 0, &n.p));
    *dep = n.node;
    G.nodes.push_back(std::move(n));
    return PC_OK;
}

// Build level g's graph: every batch's block kernels (run mode) and its
// delay-line sum for the next block, chained in the order launch_batch issues
// them on one stream.  Tick-dependent arguments start at 0 (set per launch).
static int level_graph_build(Engine &E, int g)
{
    if ((int)E.lgraph.size() != (int)E.gsize.size()) E.lgraph.resize(E.gsize.size());
    LevelGraph &G = E.lgraph[g];
    HIPCHK(hipGraphCreate(&G.g, 0));
    hipGraphNode_t dep = nullptr;
    int rc = PC_OK;
    for (Group &Gr : E.groups)
        for (Batch &b : Gr.batches) {
            if (b.g != g || rc) continue;
            auto mk = [&]() {
                std::unique_ptr<GArgs> a(new GArgs());
                a->jobs = b.d_jobs;
                a->sh = b.sh;
                a->N0 = E.N0;
                a->N = b.N;
                a->tw = E.tw;
                a->mode = PC_MODE_RUN;
                return a;
            };
            if (b.big) {
                const int A = 1 << b.sh.logA, B = 1 << b.sh.logB;
                auto a1 = mk();
                GArgs *r1 = a1.get();
                void *p1[] = {&r1->jobs, &r1->sh, &r1->N0, &r1->m, &r1->tw, &r1->mode};
                memcpy(r1->ptrs, p1, sizeof p1);
                if ((rc = graph_node(G, (const void *)k_big1, dim3((A / BIG_G) * b.njobs), BIG_NT, big1_lds(b.sh),
                                     std::move(a1), 6, &dep)))
                    break;
                auto a2 = mk();
                GArgs *r2 = a2.get();
                void *p2[] = {&r2->jobs, &r2->sh, &r2->bcur, &r2->tw, &r2->mode};
                memcpy(r2->ptrs, p2, sizeof p2);
                if ((rc = graph_node(G, (const void *)k_big2, dim3((B / 2) * b.njobs), BIG_NT, big2_lds(b.sh),
                                     std::move(a2), 5, &dep)))
                    break;
                auto a3 = mk();
                GArgs *r3 = a3.get();
                void *p3[] = {&r3->jobs, &r3->sh, &r3->ypar, &r3->tw};
                memcpy(r3->ptrs, p3, sizeof p3);
                if ((rc = graph_node(G, (const void *)k_big3, dim3((A / BIG_G) * b.njobs), BIG_NT, big1_lds(b.sh),
                                     std::move(a3), 4, &dep)))
                    break;
            } else {
                const void *fn = small_fn(ilog2(b.N));
                if (!fn) {
                    rc = fail(PC_ERR_ARG, "unsupported small block size");
                    break;
                }
                auto a0 = mk();
                GArgs *r0 = a0.get();
                void *p0[] = {&r0->jobs, &r0->N0, &r0->m, &r0->bcur, &r0->ypar, &r0->tw, &r0->mode};
                memcpy(r0->ptrs, p0, sizeof p0);
                if ((rc = graph_node(G, fn, dim3(b.njobs), small_nt(b.N), 0, std::move(a0), 7, &dep))) break;
            }
            const int Q = E.part_q;
            auto ap = mk();
            GArgs *rp = ap.get();
            rp->wpj = (b.N / 2 + PART_NT * Q - 1) / (PART_NT * Q);
            void *pp[] = {&rp->jobs, &rp->N, &rp->wpj, &rp->b};
            memcpy(rp->ptrs, pp, sizeof pp);
            const void *pf = Q == 4 ? (const void *)k_partial<4> : Q == 2 ? (const void *)k_partial<2>
                                                                          : (const void *)k_partial<1>;
            if ((rc = graph_node(G, pf, dim3(rp->wpj * b.njobs), PART_NT, 0, std::move(ap), 4, &dep))) break;
        }
    if (rc) return rc;
    HIPCHK(hipGraphInstantiate(&G.x, G.g, nullptr, nullptr, 0));
    return PC_OK;
}

// Level g's blocks that start at tick f through its graph, on stream s.
static int level_graph_launch(Engine &E, int g, uint64_t f, hipStream_t s)
{
    if ((int)E.lgraph.size() <= g || !E.lgraph[g].x) {
        const int rc = level_graph_build(E, g);
        if (rc) return rc;
    }
    LevelGraph &G = E.lgraph[g];
    const uint64_t m = f >> E.gshift[g];
    for (GNode &n : G.nodes) {
        n.a->m = m;
        n.a->bcur = m;
        n.a->ypar = (int)(m & 1);
        n.a->b = m + 1;
        HIPCHK(hipGraphExecKernelNodeSetParams(G.x, n.node, &n.p));
    }
    HIPCHK(hipGraphLaunch(G.x, s));
    return PC_OK;
}

static int launch_level(Engine &E, int g, uint64_t f)
{
    hipStream_t s = level_stream(E, g);
    int want = g;
    if (E.stall_g.load() == g && E.stall_g.compare_exchange_strong(want, -1))
        HIPCHK(hipLaunchHostFunc(s, host_sleep_fn, (void *)(uintptr_t)E.stall_us.load()));
    if (E.level_graphs) {
        int rc = level_graph_launch(E, g, f, s);
        if (rc) return rc;
        HIPCHK(hipEventRecord(E.ev_lvl[g], s));
        return PC_OK;
    }
    const uint64_t m = f >> E.gshift[g];
    for (Group &Gr : E.groups)
        for (Batch &b : Gr.batches) {
            if (b.g != g) continue;
            int rc = launch_batch(E, b, m, s);
            if (rc) return rc;
        }
    HIPCHK(hipGetLastError());
    HIPCHK(hipEventRecord(E.ev_lvl[g], s));
    return PC_OK;
}

static void launcher_main(Engine *E)
{
    LevelLauncher &L = *E->lau;
    (void)hipSetDevice(E->dev);
    for (;;) {
        std::pair<int, uint64_t> job;
        {
            std::unique_lock<std::mutex> lk(L.mu);
            L.cv.wait(lk, [&] { return L.stop || !L.q.empty(); });
            if (L.q.empty()) return;  // stopped with nothing left
            job = L.q.front();
            L.q.erase(L.q.begin());
        }
        const int rc = launch_level(*E, job.first, job.second);
        std::lock_guard<std::mutex> lk(L.mu);
        if (rc && !L.err) {
            L.err = rc;
            L.msg = g_err;
        }
        L.done[job.first].store(job.second + 1, std::memory_order_release);
        L.busy--;
        L.cv.notify_all();
    }
}

static void launcher_start(Engine &E)
{
    if (E.lau) return;
    E.lau.reset(new LevelLauncher());
    const size_t G = E.gsize.size();
    E.lau->done.reset(new std::atomic<uint64_t>[G]);
    for (size_t g = 0; g < G; g++) E.lau->done[g].store(0);
    E.lau->th = std::thread(launcher_main, &E);
}

// Wait until every queued launch is issued; report the first error.
static int launcher_drain(Engine &E)
{
    if (!E.lau) return PC_OK;
    LevelLauncher &L = *E.lau;
    std::unique_lock<std::mutex> lk(L.mu);
    L.cv.wait(lk, [&] { return L.busy == 0; });
    if (L.err) {
        const int rc = L.err;
        L.err = 0;
        return fail(rc, L.msg);
    }
    return PC_OK;
}

static void launcher_stop(Engine &E)
{
    if (!E.lau) return;
    {
        std::lock_guard<std::mutex> lk(E.lau->mu);
        E.lau->stop = true;
    }
    E.lau->cv.notify_all();
    if (E.lau->th.joinable()) E.lau->th.join();
    E.lau.reset();
}

// Launch the level blocks whose period starts at tick f.  With use_event the
// level streams wait for the tick kernel of f on the device; without it the
// caller has already observed tick f's completion on the host, and the
// launches are queued to the helper thread when it runs (persistent runs).
static int start_levels(Engine &E, uint64_t f, bool use_event)
{
    const int G = (int)E.gsize.size();
    bool recorded = false;
    for (int g = 1; g < G; g++) {
        const uint64_t k = (uint64_t)E.gsize[g];
        if (f % k != 0) continue;
        if (E.workers_active && E.g_worker[g]) continue;  // the persistent kernel's workers run it
        E.lvl_start[g] = f;
        E.lvl_pending[g] = true;
        if (!use_event && E.lau) {
            {
                std::lock_guard<std::mutex> lk(E.lau->mu);
                E.lau->q.push_back({g, f});
                E.lau->busy++;
            }
            E.lau->cv.notify_all();
            E.lvl_async[g] = true;
            continue;
        }
        E.lvl_async[g] = false;
        if (use_event) {
            if (!recorded) {
                HIPCHK(hipEventRecord(E.ev_tick, E.s_main));
                recorded = true;
            }
            HIPCHK(hipStreamWaitEvent(level_stream(E, g), E.ev_tick, 0));
        }
        int rc = launch_level(E, g, f);
        if (rc) return rc;
    }
    return PC_OK;
}

// True when level g's block that tick f reads for the first time must be
// complete before tick f runs (strict: f+1 = 0 mod k; relaxed: f = 0 mod k).
static bool level_due(const Engine &E, int g, uint64_t f)
{
    const uint64_t k = (uint64_t)E.gsize[g];
    return E.relaxed ? (f % k == 0 && f >= k) : ((f + 1) % k == 0 && f + 1 >= k);
}

// One tick of PartConvMulti[Relaxed]::run on the device.
static int engine_tick(Engine &E, const float *in, long long in_stride, long long fbase, float *out,
                       long long out_stride, bool wait)
{
    // a tick that a stalled persistent session did not serve counts from its
    // first submission
    const double t0 = E.pend_t0 >= 0 ? E.pend_t0 : now_s();
    E.pend_t0 = -1;
    const uint64_t f = E.frame;
    const int G = (int)E.gsize.size();
    // deadlines: wait for the blocks this tick reads for the first time
    for (int g = 1; g < G; g++) {
        if (level_due(E, g, f) && E.lvl_pending[g]) {
            HIPCHK(hipStreamWaitEvent(E.s_main, E.ev_lvl[g], 0));
            E.lvl_pending[g] = false;
        }
    }
    TickArgs a{};
    a.N0 = E.N0;
    a.f = f;
    a.in = in;
    a.in_stride = in_stride;
    a.fbase = fbase;
    a.nIn = E.nIn;
    a.out = out;
    a.out_stride = out_stride;
    a.nOut = E.nOut;
    set_io_layout(E, a, in == E.d_in_map, out == E.d_out_map);
    a.relaxed = E.relaxed;
    a.src_off = E.d_src_off;
    a.src_list = E.d_src_list;
    a.ticket = E.d_ticket;
    a.ticket_ch = E.d_ticket_ch;
    a.ticket_grp = E.d_ticket_grp;
    a.nactive = E.nactive;
    a.empty_ch = E.d_empty;
    a.nempty = E.nempty;
    // host-memory output needs a system-scope release before the flag; device
    // output is published to later kernels by the kernel boundary itself
    a.fence_sys = (out == E.d_out_map) ? 1 : 0;
    a.fence_mask = ~0ull;
    if (E.mode == PC_BATCHED) {
        Group &Gr = E.groups[0];
        a.jobs = Gr.d_tick;
        a.njobs = (int)Gr.members.size();
        a.mix = E.mix;
        a.y_inst = E.y_inst;
        a.done_flag = E.d_flag_map;
        a.wg_flags = E.use_wgflags ? E.d_wg_map + (size_t)E.nwg * PC_WG_FLAG_STRIDE : nullptr;
        a.wg_stride = 1;
        LaunchExtra x;
        if (E.probe_g == 0) {
            int rc = probe_take(E, x);
            if (rc) return rc;
        }
        int rc = launch_tick(E.N0, a, E.tw, E.s_main, x);
        if (rc) return rc;
        if (E.mix == 0) {
            const int n = E.nOut * E.N0;
            hipLaunchKernelGGL(k_mix, dim3((n + 255) / 256), dim3(256), 0, E.s_main, E.y_inst, E.d_src_off,
                               E.d_src_list, E.N0, E.nOut, out, a.out_cs, a.out_ss, (long long)(f - fbase) * out_stride, E.d_ticket + 1,
                               E.d_flag_map, f);
        }
    } else {
        for (Group &Gr : E.groups) {
            a.jobs = Gr.d_tick;
            a.njobs = 1;
            a.mix = 0;
            a.y_inst = E.y_inst + (size_t)Gr.members[0] * E.N0;
            a.done_flag = nullptr;
            int rc = launch_tick(E.N0, a, E.tw, E.s_main);
            if (rc) return rc;
        }
        const int n = E.nOut * E.N0;
        hipLaunchKernelGGL(k_mix, dim3((n + 255) / 256), dim3(256), 0, E.s_main, E.y_inst, E.d_src_off,
                           E.d_src_list, E.N0, E.nOut, out, a.out_cs, a.out_ss, (long long)(f - fbase) * out_stride, E.d_ticket + 1,
                           E.d_flag_map, f);
    }
    HIPCHK(hipGetLastError());
    // record the tick kernel's event now so the level streams depend on it on
    // the device; the level launches themselves (host work) follow the wait,
    // outside this tick's latency (src/partconvMulti.cpp:860-873 starts them
    // before its own level-0 work; here they run on other streams anyway)
    E.frame++;
    if (wait) {
        int rc = wait_flag(E, f + 1);
        if (rc) return rc;
        record_latency(E, 1e6 * (now_s() - t0));
    }
    int rc = start_levels(E, f, true);
    if (rc) return rc;
    return PC_OK;
}

// Grid: one tick workgroup per instance, then one workgroup per level worker.
// Dynamic LDS: two float2 buffers of max(N0, largest worker block).
template <int LOGN>
static hipError_t launch_persist_t(Engine &E, PersistArgs &pa, hipStream_t s, int nblocks, size_t lds, bool *plain)
{
    const cf *t = E.tw;
    void *args[] = {(void *)&pa, (void *)&t};
    const void *k = (LOGN == 6 && pa.prefetch) ? (const void *)k_persist<LOGN, true> : (const void *)k_persist<LOGN>;
    return launch_persistent(k, nblocks, TickCfg<LOGN>::NT, args, lds, s, E.resid, E.ncu,
                             &pa.present, &pa.present_tag, plain);
}

static hipError_t launch_persist(Engine &E, PersistArgs &p, hipStream_t s, int lds_n, bool *plain)
{
    const int nblocks = p.A.njobs + p.nworkers;
    p.tw_lds_off = 2 * lds_n;
    const size_t lds = (size_t)(2 * lds_n + p.tw_lds) * sizeof(cf);
    switch (ilog2(E.N0)) {
    case 4: return launch_persist_t<4>(E, p, s, nblocks, lds, plain);
    case 5: return launch_persist_t<5>(E, p, s, nblocks, lds, plain);
    case 6: return launch_persist_t<6>(E, p, s, nblocks, lds, plain);
    case 7: return launch_persist_t<7>(E, p, s, nblocks, lds, plain);
    case 8: return launch_persist_t<8>(E, p, s, nblocks, lds, plain);
    case 9: return launch_persist_t<9>(E, p, s, nblocks, lds, plain);
    case 10: return launch_persist_t<10>(E, p, s, nblocks, lds, plain);
    case 11: return launch_persist_t<11>(E, p, s, nblocks, lds, plain);
    default: return hipErrorInvalidValue;
    }
}

template <int LOGN>
static hipError_t persist_occupancy_t(int *nb, size_t lds)
{
    // the read-ahead form of the 64-point kernel holds more registers: the
    // residency check counts with it
    if constexpr (LOGN == 6)
        return hipOccupancyMaxActiveBlocksPerMultiprocessor(nb, k_persist<LOGN, true>, TickCfg<LOGN>::NT, lds);
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(nb, k_persist<LOGN>, TickCfg<LOGN>::NT, lds);
}

static hipError_t persist_occupancy(int N0, int *nb, size_t lds)
{
    switch (ilog2(N0)) {
    case 4: return persist_occupancy_t<4>(nb, lds);
    case 5: return persist_occupancy_t<5>(nb, lds);
    case 6: return persist_occupancy_t<6>(nb, lds);
    case 7: return persist_occupancy_t<7>(nb, lds);
    case 8: return persist_occupancy_t<8>(nb, lds);
    case 9: return persist_occupancy_t<9>(nb, lds);
    case 10: return persist_occupancy_t<10>(nb, lds);
    case 11: return persist_occupancy_t<11>(nb, lds);
    default: return hipErrorInvalidValue;
    }
}

// Who runs the worker levels in persistent runs.  Tick and worker workgroups
// together may take at most half of the workgroups the device holds at once,
// so the level kernels launched from the host (the long-period levels) still
// find room beside the persistent kernel.  If one worker per job fits, each
// job gets its own worker workgroup; otherwise every tick workgroup runs its
// own instance's worker-level blocks after its tick (self mode), which costs
// no extra workgroups and needs no handshake.
static int plan_workers(Engine &E)
{
    if (E.nworkers >= 0) return PC_OK;
    E.nworkers = 0;
    E.self_levels = false;
    hipDeviceProp_t prop;
    HIPCHK(hipGetDeviceProperties(&prop, E.dev));
    int nb = 0;
    const size_t lds_buf = (size_t)2 * std::max(E.N0, E.lds_w) * sizeof(cf);
    HIPCHK(persist_occupancy(E.N0, &nb, lds_buf));
    // the level blocks' twiddles (tw[n .. 2n) per transform size n <= lds_w and
    // the real-split entries tw[2n .. 2.5n]) staged in LDS, unless that would
    // lower the number of workgroups a CU holds
    E.tw_lds = 0;
    if (E.nwjobs > 0) {
        const int n = 2 * E.lds_w + E.lds_w / 2 + 1;
        int nb2 = 0;
        if (persist_occupancy(E.N0, &nb2, lds_buf + (size_t)n * sizeof(cf)) == hipSuccess && nb2 >= nb) E.tw_lds = n;
        (void)hipGetLastError();
    }
    const long long budget = (long long)prop.multiProcessorCount * nb / 2;
    const long long ntick = (long long)E.groups[0].members.size();
    E.persist_budget = budget;
    if (E.nwjobs == 0) {
        // nothing to plan
    } else if (ntick + E.nwjobs > budget || E.persist_workers == 2) {
        E.self_levels = true;
    } else {
        std::vector<int2> ranges;
        for (int q = 0; q < E.nwjobs; q++) ranges.push_back(make_int2(q, 1));
        HIPCHK(hipMalloc(&E.d_wrange, ranges.size() * sizeof(int2)));
        HIPCHK(hipMemcpy(E.d_wrange, ranges.data(), ranges.size() * sizeof(int2), hipMemcpyHostToDevice));
        E.nworkers = (int)ranges.size();
    }
    if (E.verbosity > 0)
        printf("persistent kernel: %lld tick + %d worker workgroups for %d worker-level jobs%s (%d per CU fit)\n",
               ntick, E.nworkers, E.nwjobs, E.self_levels ? ", run by the tick workgroups" : "", nb);
    return PC_OK;
}

static void store_go(Engine &E, unsigned long long v)
{
    if (E.go_slot >= 0) {
        _mm_sfence();  // frames and state written before the release
        *(volatile unsigned long long *)E.go_h = v;
        _mm_sfence();  // send the (write-combined) store now
    } else {
        __atomic_store_n(E.go_h, v, __ATOMIC_RELEASE);
    }
}

// Host sessions with tagged input words: tick f's frame (planar h_in) as
// tagged words in device memory; the words are the release.
static void ll_release(Engine &E, uint64_t f)
{
    const unsigned tag = ll_tag(f);
    const size_t n = (size_t)E.nIn * E.N0;
    unsigned long long *w = E.ll_in;
    _mm_sfence();
    for (size_t i = 0; i < n; i++) w[i] = ll_word(E.h_in[i], tag);
    _mm_sfence();  // send the (write-combined) stores now
}

// Clear the words before a session starts: none may carry the tag of a tick
// the session will serve (frames rewind on reset).
static void ll_clear(Engine &E)
{
    const size_t n = (size_t)E.nIn * E.N0;
    _mm_sfence();
    for (size_t i = 0; i < n; i++) E.ll_in[i] = 0;
    _mm_mfence();
}

template <int LOGN>
static void launch_worker_fix_t(Engine &E, uint64_t fend)
{
    hipLaunchKernelGGL(k_worker_fix<LOGN>, dim3(E.nwjobs), dim3(TickCfg<LOGN>::NT),
                       (size_t)2 * E.lds_w * sizeof(cf), E.s_main, E.d_wjobs, E.d_lvl_done, fend, E.N0, E.tw);
}

// After a persistent run with level workers: complete, on the device, every
// worker-level block with start tick < fend that the run did not finish
// (k_worker_fix), ordered on the tick stream before anything that follows.
static int finish_worker_levels(Engine &E, uint64_t fend)
{
    switch (ilog2(E.N0)) {
    case 4: launch_worker_fix_t<4>(E, fend); break;
    case 5: launch_worker_fix_t<5>(E, fend); break;
    case 6: launch_worker_fix_t<6>(E, fend); break;
    case 7: launch_worker_fix_t<7>(E, fend); break;
    case 8: launch_worker_fix_t<8>(E, fend); break;
    case 9: launch_worker_fix_t<9>(E, fend); break;
    case 10: launch_worker_fix_t<10>(E, fend); break;
    default: return fail(PC_ERR_ARG, "no worker levels at this buffer size");
    }
    HIPCHK(hipGetLastError());
    return PC_OK;
}

// Device-resident ticks served by one persistent tick kernel (batched engine).
// Per tick the host (1) checks, with hipEventQuery, that the level blocks due
// at this tick are complete, (2) releases the tick, (3) spins on the
// completion flag, (4) launches the level blocks that start at this tick.
// If anything stalls (a level stream that cannot run beside the persistent
// kernel, a kernel that exited) the host stops the kernel and finishes the
// remaining ticks with per-tick launches, so output is never wrong or late
// beyond the stall itself.
// ---- persistent sessions ----------------------------------------------------
// persist_begin launches the persistent tick kernel for up to nticks ticks from
// E.frame; persist_tick serves tick E.frame through it; persist_end finishes
// the session (stopping the kernel if stop is set).  Per tick the host
//   (1) checks, with hipEventQuery, that the host-launched level blocks due at
//       this tick are complete,
//   (2) releases the tick (go = f + 1 in pinned memory),
//   (3) spins on the completion flag(s),
//   (4) launches the host-launched level blocks that start at this tick.
// If a level stream stalls beside the kernel or the kernel has ended, the tick
// is reported unserved and the caller ends the session and falls back to
// per-tick launches; output is never wrong or late beyond the stall itself.
static int persist_begin(Engine &E, const float *d_in, long long in_stride, float *d_out, long long out_stride,
                         long long nticks, int host_io)
{
    if (E.mix == 0) return 1;  // separate mix kernel per tick: per-tick launches
    {
        int rc = plan_workers(E);
        if (rc) return rc;
    }
    // the tick workgroups alone must leave room for the level kernels launched
    // beside the persistent kernel: else per-tick launches
    if ((long long)E.groups[0].members.size() > E.persist_budget) return 1;
    const uint64_t f0 = E.frame;
    PersistArgs P{};
    TickArgs &a = P.A;
    Group &Gr = E.groups[0];
    a.jobs = Gr.d_tick;
    a.njobs = (int)Gr.members.size();
    a.N0 = E.N0;
    a.in = d_in;
    a.in_stride = in_stride;
    a.fbase = (long long)f0;
    a.nIn = E.nIn;
    a.out = d_out;
    a.out_stride = out_stride;
    a.nOut = E.nOut;
    set_io_layout(E, a, host_io, host_io);
    a.relaxed = E.relaxed;
    a.mix = E.mix;
    a.y_inst = E.y_inst;
    a.src_off = E.d_src_off;
    a.src_list = E.d_src_list;
    a.ticket = E.d_ticket;
    a.ticket_ch = E.d_ticket_ch;
    a.ticket_grp = E.d_ticket_grp;
    a.nactive = E.nactive;
    a.empty_ch = E.d_empty;
    a.nempty = E.nempty;
    a.done_flag = E.d_flag_map;
    a.wg_flags = E.use_wgflags ? E.d_wg_map : nullptr;
    a.wg_stride = PC_WG_FLAG_STRIDE;
    // frames in / output to pinned host memory: system-scope acquire after the
    // release (not with tagged input words: each carries its tick) and
    // system-scope release before the completion flag
    const bool ll = host_io && E.ll_in && E.host_session == 1;
    a.fence_sys = host_io;
    a.ll = ll ? 1 : 0;
    a.ll_in = E.ll_in;  // one address for host and device
    if (ll) ll_clear(E);
    // Level jobs launched from the host start only at ticks f = 0 mod k_min
    // (k_min over the host-launched levels); those ticks publish the ring with
    // a device-scope release, and the ticks that first read such a level's
    // block acquire at device scope.  Separate level workers get the frames at
    // ticks f = 0 mod (their k_min).
    auto set_masks = [&](bool workers, bool separate) {
        unsigned long long kh = 0, kw = 0;
        for (int g = 1; g < (int)E.gsize.size(); g++) {
            const unsigned long long k = (unsigned long long)E.gsize[g];
            if (workers && E.g_worker[g]) kw = kw ? std::min(kw, k) : k;
            else kh = kh ? std::min(kh, k) : k;
        }
        a.fence_mask = kh ? kh - 1 : ~0ull;
        P.acq_mask = a.fence_mask;
        P.pub_mask = (separate && kw) ? kw - 1 : ~0ull;
    };
    set_masks(false, false);
    // A release word in device memory is polled by every workgroup.  One in
    // pinned host memory is polled across the bus: with many workgroups only
    // workgroup 0 polls it and relays it through device memory
    // (d_ticket[2..3]); up to 32 each poll the bus themselves (measured faster
    // than the relay at 16 instances)
    P.dev_go = (a.njobs > 32 && E.go_slot < 0) ? reinterpret_cast<unsigned long long *>(E.d_ticket + 2) : nullptr;
    if (P.dev_go) HIPCHK(hipMemsetAsync(P.dev_go, 0, sizeof(unsigned long long), E.s_main));
    a.trace = E.d_trace;
    a.trace_cap = E.trace_cap;
    P.f0 = f0;
    P.nticks = nticks;
    P.go = E.go_d;
    P.spin_limit = PC_SPIN_LIMIT;
    // Reading a tick's level samples and worker counter ahead of its release
    // pays when every tick waits for its own release (measured at 16 channels x
    // 262144 taps: p50 5.05 -> 4.6 us); inside a release window the next tick
    // is mostly released already and the extra loads delay the release read
    // (3.74 -> 3.97 us per tick at a window of 3), so windowed runs launch the
    // plain kernel.
    P.prefetch = (E.release_window <= 1 || !E.use_wgflags || ll) ? 1 : 0;
    // the levels the first tick reads must be complete before the kernel owns the tick stream
    const int G = (int)E.gsize.size();
    for (int g = 1; g < G; g++)
        if (E.lvl_pending[g]) HIPCHK(hipStreamWaitEvent(E.s_main, E.ev_lvl[g], 0));
    store_go(E, (unsigned long long)f0);
    hipError_t le = hipErrorInvalidValue;
    bool plain = false;
    if (E.nwjobs > 0 && E.persist_workers && (E.nworkers > 0 || E.self_levels)) {
        // worker-level progress counters for a history served up to tick f0 - 1
        const int n = std::max(E.nwjobs, (int)E.inst.size());
        hipLaunchKernelGGL(k_worker_init, dim3((n + 255) / 256), dim3(256), 0, E.s_main, E.d_wjobs, E.nwjobs,
                           (uint64_t)f0, E.d_lvl_done, E.d_dev_tick, (int)E.inst.size());
        HIPCHK(hipGetLastError());
        P.wjobs = E.d_wjobs;
        P.wrange = E.d_wrange;
        P.nworkers = E.nworkers;
        P.self_levels = E.self_levels ? 1 : 0;
        P.dev_tick = E.d_dev_tick;
        P.lvl_done = E.d_lvl_done;
        P.tw_lds = E.tw_lds;
        set_masks(true, E.nworkers > 0);
        le = launch_persist(E, P, E.s_main, std::max(E.N0, E.lds_w), &plain);
        if (le != hipSuccess) (void)hipGetLastError();  // too many workgroups to co-reside: no workers
        else E.workers_active = true;
    }
    if (le != hipSuccess) {
        P.wjobs = nullptr;
        P.wrange = nullptr;
        P.nworkers = 0;
        P.self_levels = 0;
        P.dev_tick = P.lvl_done = nullptr;
        P.tw_lds = 0;
        set_masks(false, false);
        le = launch_persist(E, P, E.s_main, E.N0, &plain);
    }
    if (le != hipSuccess) {
        (void)hipGetLastError();
        return 1;  // caller falls back to per-tick launches
    }
    E.sess_active = true;
    E.wg_persist = true;
    E.ll_active = ll;
    launcher_start(E);
    E.sess_end = f0 + (uint64_t)nticks;
    E.sess_last = now_s();
    if (plain && !resident_wait(E.resid, P.A.njobs + P.nworkers)) {
        // not every workgroup found room: stop it, serve with per-tick launches
        const int rc = persist_end(E, true);
        return rc ? rc : 1;
    }
    return PC_OK;
}

// Serve tick E.frame: 0 served, 1 not served (stalled: end the session and
// fall back), other values are errors.
// A level block that is not complete PC_LEVEL_LATE_S after its deadline tick
// was submitted ends the session (the caller falls back to per-tick launches,
// which wait for the block on the device); the unserved tick's submit time is
// kept in pend_t0 so its latency counts from its first submission.
#define PC_LEVEL_LATE_S 0.2
static int persist_tick(Engine &E)
{
    const uint64_t f = E.frame;
    if (!E.sess_active || f >= E.sess_end) return 1;
    const int G = (int)E.gsize.size();
    const double t0 = now_s();
    for (int g = 1; g < G; g++) {
        if (!(level_due(E, g, f) && E.lvl_pending[g])) continue;
        if (E.lvl_async[g]) {
            // the helper thread must have issued the block (and recorded the event)
            while (E.lau->done[g].load(std::memory_order_acquire) < E.lvl_start[g] + 1)
                if (now_s() - t0 > PC_LEVEL_LATE_S) {
                    E.pend_t0 = t0;
                    return 1;
                }
            E.lvl_async[g] = false;
        }
        for (;;) {
            hipError_t q = hipEventQuery(E.ev_lvl[g]);
            if (q == hipSuccess) break;
            if (q != hipErrorNotReady) return fail(PC_ERR_HIP, std::string("level block: ") + hipGetErrorString(q));
            if (now_s() - t0 > PC_LEVEL_LATE_S) {
                E.pend_t0 = t0;
                return 1;
            }
        }
        E.lvl_pending[g] = false;
    }
    if (E.ll_active) ll_release(E, f);
    else store_go(E, f + 1);
    E.sess_last = now_s();
    int wr = wait_flag(E, f + 1, true);
    if (wr == 1) E.pend_t0 = t0;
    if (wr) return wr;
    record_latency(E, 1e6 * (now_s() - t0));
    E.frame++;
    return start_levels(E, f, false);
}

// End the session.  stop: the kernel may still be waiting for releases (a
// stall, or a host session): stop it, then account for a tick it served after
// all.  Completes the worker levels' state for the ticks served.
static int persist_end(Engine &E, bool stop)
{
    if (!E.sess_active) return PC_OK;
    E.sess_active = false;
    // every queued level launch issued (and its event recorded) before
    // anything below waits on or relaunches levels
    const int lrc = launcher_drain(E);
    std::fill(E.lvl_async.begin(), E.lvl_async.end(), false);
    int serr = PC_OK;
    if (stop) {
        // the ticks a release-window run released (f < win_rel) are served
        // before the kernel stops; otherwise it stops at once
        const uint64_t drain = E.win_rel;
        E.win_rel = 0;
        store_go(E, PC_GO_STOP | drain);
        HIPCHK(hipStreamSynchronize(E.s_main));
        // a stopped tick may have counted only some of its arrivals
        HIPCHK(hipMemsetAsync(E.d_ticket, 0, 64, E.s_main));
        HIPCHK(hipMemsetAsync(E.d_ticket_ch, 0, std::max(E.nOut, 1) * sizeof(unsigned), E.s_main));
        HIPCHK(hipMemsetAsync(E.d_ticket_grp, 0, E.ticket_grp_bytes, E.s_main));
        HIPCHK(hipStreamSynchronize(E.s_main));
        if (drain > E.frame) {
            // every released tick the kernel completed, in order, from its release
            const size_t W = E.win_trel.size();
            while (E.frame < drain && tick_done(E, E.frame + 1)) {
                const uint64_t f = E.frame;
                record_latency(E, 1e6 * (now_s() - E.win_trel[f % W]));
                E.frame++;
                E.sess_served_late = true;
                int rc = start_levels(E, f, true);
                if (rc) return rc;
            }
            if (E.frame < drain) {
                // A workgroup ends only at a tick boundary.  Per-tick launches
                // from tick E.frame serve every instance again, which repeats
                // exactly what a workgroup already did for tick E.frame but not
                // what it did for a later tick: refuse that state.
                const volatile unsigned long long *wf = E.h_wg;
                for (int i = 0; i < E.nwg && !serr; i++)
                    if (wf[(size_t)i * PC_WG_FLAG_STRIDE] > E.frame + 1)
                        serr = fail(PC_ERR_STATE, "persistent run ended with workgroups past the first unserved tick");
                E.pend_t0 = E.win_trel[E.frame % W];
            }
        } else if (E.frame < E.sess_end && tick_done(E, E.frame + 1)) {
            // the kernel served the released tick after all (slow, not stalled)
            const uint64_t f = E.frame;
            if (E.pend_t0 >= 0) record_latency(E, 1e6 * (now_s() - E.pend_t0));
            E.pend_t0 = -1;
            E.frame++;
            E.sess_served_late = true;
            int rc = start_levels(E, f, true);
            if (rc) return rc;
        }
    }
    E.win_rel = 0;
    E.wg_persist = false;  // per-tick launches use the packed flags
    E.ll_active = false;
    if (E.workers_active) {
        E.workers_active = false;
        int rc = finish_worker_levels(E, E.frame);
        if (rc) return rc;
    }
    return serr ? serr : lrc;
}

// Release window: the host-launched level blocks that tick r reads for the
// first time are complete (0), not yet (1: a block still runs, or its start
// tick has not been seen complete, so it is not launched yet), or an error.
static int window_levels_ready(Engine &E, uint64_t r)
{
    const int G = (int)E.gsize.size();
    for (int g = 1; g < G; g++) {
        if (!level_due(E, g, r)) continue;
        if (E.workers_active && E.g_worker[g]) continue;  // the kernel's own workers
        const uint64_t k = (uint64_t)E.gsize[g];
        const uint64_t s = E.relaxed ? r - k : r + 1 - k;  // the block's start tick
        if (s >= E.frame) return 1;
        if (!E.lvl_pending[g]) continue;
        if (E.lvl_async[g]) {
            if (E.lau->done[g].load(std::memory_order_acquire) < E.lvl_start[g] + 1) return 1;
            E.lvl_async[g] = false;
        }
        const hipError_t q = hipEventQuery(E.ev_lvl[g]);
        if (q == hipErrorNotReady) return 1;
        if (q != hipSuccess) return fail(PC_ERR_HIP, std::string("level block: ") + hipGetErrorString(q));
        E.lvl_pending[g] = false;
    }
    return 0;
}

// Device-resident ticks with a release window of W = release_window ticks
// (direct mix with a completion flag per workgroup: each workgroup writes its
// own output of each tick, so workgroups may be at different ticks).  Up to W
// ticks are released ahead of the oldest one not yet seen complete, so the
// host's round trip (the completion flags across the bus, then the next
// release) overlaps the device's work on the ticks already released; every
// workgroup still runs its ticks in order.  The deadline rule is that of
// persist_tick: tick r is released only once every host-launched level block
// it reads first is complete, and such a block is launched once its start
// tick is seen complete, so the window narrows where a level's period is
// short.  A tick's latency counts from its release (or, if it waited for a
// level block with nothing in flight, from when the window emptied).  Returns 0 when the ticks
// to fend are served, 1 on a stall (win_rel: the first tick not released;
// persist_end stops the kernel after the released ticks), or an error.
static int persist_run_window(Engine &E, uint64_t fend)
{
    const uint64_t W = (uint64_t)E.release_window;
    E.win_trel.assign((size_t)W, 0.0);
    E.win_rel = E.frame;
    double t_idle = now_s();  // the window emptied: tick win_rel's submission
    double t_wait = -1;       // the oldest tick's completion first found pending
    unsigned spins = 0;
    while (E.frame < fend) {
        while (E.win_rel < fend && E.win_rel < E.frame + W) {
            const int lr = window_levels_ready(E, E.win_rel);
            if (lr == 1) break;
            if (lr) return lr;
            store_go(E, E.win_rel + 1);
            // latency from the release, or from when the window emptied if the
            // tick waited for a level block (as one tick at a time counts it)
            E.sess_last = now_s();
            E.win_trel[E.win_rel % W] = (E.win_rel == E.frame) ? t_idle : E.sess_last;
            E.win_rel++;
        }
        if (E.win_rel == E.frame) {
            // nothing in flight: tick win_rel waits for a level block
            if (now_s() - t_idle > PC_LEVEL_LATE_S) {
                E.pend_t0 = t_idle;
                return 1;
            }
            continue;
        }
        if (!tick_done(E, E.frame + 1)) {
            if (t_wait < 0) t_wait = now_s();
            if ((++spins & 0x3FFF) == 0) {
                const hipError_t q = hipStreamQuery(E.s_main);
                if (q != hipSuccess && q != hipErrorNotReady)
                    return fail(PC_ERR_HIP, std::string("tick failed: ") + hipGetErrorString(q));
                if ((q == hipSuccess && !tick_done(E, E.frame + 1)) || now_s() - t_wait > PC_TICK_TIMEOUT_S)
                    return 1;
            }
            continue;
        }
        t_wait = -1;
        record_latency(E, 1e6 * (now_s() - E.win_trel[E.frame % W]));
        const uint64_t f = E.frame++;
        const int rc = start_levels(E, f, false);
        if (rc) return rc;
        if (E.win_rel == E.frame) t_idle = now_s();
    }
    return PC_OK;
}

// Device-resident ticks served by one persistent tick kernel (batched engine);
// after a stall the remaining ticks use per-tick launches.
static int engine_run_persistent(Engine &E, const float *d_in, float *d_out, long long nframes,
                                 long long *persistent_ticks)
{
    const uint64_t f0 = E.frame;
    const long long in_stride = (long long)E.N0 * E.nIn, out_stride = (long long)E.N0 * E.nOut;
    *persistent_ticks = 0;
    int rc = persist_begin(E, d_in, in_stride, d_out, out_stride, nframes, 0);
    if (rc) return rc;  // 1: per-tick launches
    bool stalled = false;
    if (E.release_window > 1 && E.use_wgflags && !E.ll_active) {
        rc = persist_run_window(E, f0 + (uint64_t)nframes);
        if (rc == 1) {
            stalled = true;
            E.stalls++;
        } else if (rc) {
            return rc;
        }
    }
    while (!stalled && E.frame < f0 + (uint64_t)nframes) {
        rc = persist_tick(E);
        if (rc == 1) {
            stalled = true;
            E.stalls++;
            break;
        }
        if (rc) return rc;
    }
    rc = persist_end(E, stalled);
    if (rc) return rc;
    *persistent_ticks = (long long)(E.frame - f0);
    while (E.frame < f0 + (uint64_t)nframes) {
        rc = engine_tick(E, d_in, in_stride, (long long)f0, d_out, out_stride, true);
        if (rc) return rc;
    }
    return PC_OK;
}

static int engine_drain(Engine &E)
{
    HIPCHK(hipStreamSynchronize(E.s_main));
    for (hipStream_t s : E.s_lvl) HIPCHK(hipStreamSynchronize(s));
    return PC_OK;
}

// ----------------------------------------------------------------------------
// C ABI: PartConvMulti
// ----------------------------------------------------------------------------
struct pc_multi {
    Engine E;
    int max_block = 0;
    long long persistent_ticks = 0;   // ticks served by the persistent kernel
};

PC_API const char *pc_last_error(void) { return g_err.c_str(); }

PC_API int pc_device_name(char *buf, int len)
{
    if (!buf || len <= 0) return fail(PC_ERR_ARG, "null buffer");
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    hipDeviceProp_t p;
    HIPCHK(hipGetDeviceProperties(&p, dev));
    snprintf(buf, len, "%s (%s)", p.name, p.gcnArchName);
    return PC_OK;
}

PC_API int pc_read_ir(const char *filename, float *out, long long cap, long long *frames, int *sample_rate,
                      int verbosity)
{
    if (!filename || cap < 0 || (cap > 0 && !out)) return fail(PC_ERR_ARG, "bad argument");
    std::vector<float> v;
    std::string err;
    pcio::WavInfo wi;
    if (!pcio::read_wav(filename, v, &wi, err)) return fail(PC_ERR_IO, err);
    if (frames) *frames = (long long)v.size();
    if (sample_rate) *sample_rate = wi.sample_rate;
    if (out) memcpy(out, v.data(), (size_t)std::min<long long>(cap, (long long)v.size()) * sizeof(float));
    if (verbosity > 0)  // as src/sndtools.cpp:36-44
        printf("\nImpulse response file: %s\nchannels: %d (only used one)\nframes: %zu\nformat: 0x%x (%d-bit)\n"
               "sample rate: %d\n\n", filename, wi.channels, v.size(), wi.format_tag, wi.bits, wi.sample_rate);
    return PC_OK;
}

PC_API int pc_greedy_partition(const int *block_sizes, int levels, int filter_length, int *N, int *P,
                               int *M_out)
{
    if (!block_sizes || levels < 1 || !N || !P || !M_out || filter_length < 1)
        return fail(PC_ERR_ARG, "bad argument");
    // P_i = 2 N_{i+1} / N_i - 2, stop at the first level whose cumulative span
    // covers the filter (src/partconv.cpp:712-728); the last count is recounted
    // by the partition check (src/partconv.cpp:169)
    int M = levels;
    std::vector<int> n(block_sizes, block_sizes + levels), p(levels, 0);
    long long total = 0;
    for (int i = 0; i < levels - 1; i++) {
        p[i] = 2 * n[i + 1] / n[i] - 2;
        total += (long long)n[i] * p[i];
        if (total >= filter_length) {
            M = i + 1;
            break;
        }
    }
    n.resize(M);
    p.resize(M);
    std::vector<int> delay, T;
    int rc = check_partition(n, p, filter_length, delay, T);
    if (rc) return rc;
    for (int i = 0; i < M; i++) {
        N[i] = n[i];
        P[i] = p[i];
    }
    *M_out = M;
    return PC_OK;
}

PC_API pc_multi *pcm_create(int verbosity, int relaxed)
{
    pc_multi *h = new pc_multi();
    h->E.verbosity = verbosity;
    h->E.relaxed = relaxed ? 1 : 0;
    if (!live_add(h, HK_MULTI)) {
        delete h;
        fail(PC_ERR_STATE, "library shut down");
        return nullptr;
    }
    return h;
}

PC_API int pcm_set_mode(pc_multi *h, int mode)
{
    if (!h) return fail(PC_ERR_ARG, "null handle");
    if (h->E.ready) return fail(PC_ERR_STATE, "mode must be set before setup");
    if (mode != PC_BATCHED && mode != PC_PLAIN) return fail(PC_ERR_ARG, "unknown mode");
    h->E.mode = mode;
    return PC_OK;
}

static int multi_setup(pc_multi *h, int FS, int N0, int nIn, int nOut, std::vector<InstDesc> &d)
{
    int rc = engine_setup(h->E, FS, N0, nIn, nOut, d);
    if (rc) {
        std::string msg = g_err;
        engine_free(h->E);
        g_err = msg;
        return rc;
    }
    h->max_block = 0;
    for (Inst &I : h->E.inst)
        for (LevelSpec &s : I.lv) h->max_block = std::max(h->max_block, s.N);
    return PC_OK;
}

PC_API int pcm_setup(pc_multi *h, const char *config_file, int max_threads_per_level,
                     int max_level0_threads, double min_work)
{
    (void)max_threads_per_level;
    (void)max_level0_threads;
    (void)min_work;
    if (!h || !config_file) return fail(PC_ERR_ARG, "null argument");
    pcio::Config cfg;
    std::string err;
    if (!pcio::parse_config(config_file, cfg, err)) return fail(PC_ERR_CONFIG, err);
    std::vector<InstDesc> d;
    for (const pcio::ConfigLine &L : cfg.lines) {
        InstDesc x;
        x.mute = L.mute;
        x.in_ch = L.in_ch;
        x.out_ch = L.out_ch;
        x.scale = L.scale;
        x.N = L.N;
        x.P = L.P;
        pcio::WavInfo wi;
        if (!pcio::read_wav(L.ir_file, x.ir, &wi, err)) return fail(PC_ERR_IO, err);
        d.push_back(std::move(x));
    }
    return multi_setup(h, cfg.FS, cfg.buffer_size, cfg.nIn, cfg.nOut, d);
}

PC_API int pcm_setup_mem(pc_multi *h, int FS, int buffer_size, int nIn, int nOut, int numPCs,
                         const pc_instance *inst)
{
    if (!h || !inst || numPCs < 1) return fail(PC_ERR_ARG, "null argument");
    std::vector<InstDesc> d;
    for (int i = 0; i < numPCs; i++) {
        const pc_instance &s = inst[i];
        if (s.M < 1 || !s.N || !s.P || (s.ir_len > 0 && !s.ir) || s.ir_len < 0)
            return fail(PC_ERR_ARG, "instance " + std::to_string(i) + ": bad partition or IR");
        InstDesc x;
        x.mute = s.mute;
        x.in_ch = s.in_ch;
        x.out_ch = s.out_ch;
        x.scale = s.scale;
        x.N.assign(s.N, s.N + s.M);
        x.P.assign(s.P, s.P + s.M);
        x.ir.assign(s.ir, s.ir + s.ir_len);
        d.push_back(std::move(x));
    }
    return multi_setup(h, FS, buffer_size, nIn, nOut, d);
}

// One tick of the host-buffer API (frame already in the pinned input staging):
// through a persistent session when the batched engine can run one -- the
// kernel stays resident between calls and each call only releases a tick --
// else one launch per tick.  A session that has been idle for
// PC_SESSION_IDLE_S is restarted before the release, so its kernel never
// reaches its own wait limit (PC_SPIN_LIMIT) while a tick is in flight.
static int session_stop(Engine &E)
{
    return E.sess_active ? persist_end(E, true) : PC_OK;
}

static int host_tick(pc_multi *h)
{
    Engine &E = h->E;
    if (E.mode == PC_BATCHED && E.host_session) {
        if (E.sess_active && (E.sess_host == 0 || now_s() - E.sess_last > PC_SESSION_IDLE_S)) {
            int rc = session_stop(E);
            if (rc) return rc;
        }
        if (!E.sess_active) {
            int rc = persist_begin(E, E.d_in_map, 0, E.d_out_map, 0, 1ll << 40, 1);
            if (rc > 1) return rc;
            E.sess_host = (rc == 0);
            if (rc == 1) E.host_session = 0;  // this engine cannot hold a session: launches from now on
        }
        if (E.sess_active) {
            int rc = persist_tick(E);
            if (rc == 0) {
                h->persistent_ticks++;
                return PC_OK;
            }
            if (rc != 1) return rc;
            E.stalls++;
            E.sess_served_late = false;
            rc = session_stop(E);
            if (rc) return rc;
            if (E.sess_served_late) return PC_OK;  // the stopped kernel served this tick
        }
    }
    return engine_tick(E, E.d_in_map, 0, (long long)E.frame, E.d_out_map, 0, true);
}

// PartConvMax::setup (src/partconv.cpp:653-847): channel mode from (n, m, k),
// the IRs from one array (IR i, tap l at impulses[stride*(filterLength*i + l)]),
// one greedy partition over blockSizes for every instance.
PC_API int pcm_setup_max(pc_multi *h, int FS, int n, int m, int k, const float *impulses, int filterLength,
                         int stride, const int *blockSizes, int levels)
{
    if (!h || !impulses || !blockSizes || levels < 1 || filterLength < 1 || stride < 1 || n < 1 || m < 1 || k < 1)
        return fail(PC_ERR_ARG, "bad argument");
    // (inst IR, input, output) per channel mode (src/partconv.cpp:659-708)
    std::vector<std::array<int, 3>> map;
    if (n == m && n == k) {
        for (int i = 0; i < k; i++) map.push_back({i, i, i});          // parallel, distinct IRs
    } else if (n == m && k == 1) {
        for (int j = 0; j < n; j++) map.push_back({0, j, j});          // parallel, one IR
    } else if (m == n * k) {
        for (int i = 0; i < k; i++)                                     // filterbank
            for (int j = 0; j < n; j++) map.push_back({i, j, i * n + j});
    } else {
        return fail(PC_ERR_CONFIG, "PartConvMax::setup() does not support requested channel configuration");
    }
    int Ng[PC_MAX_LEVELS], Pg[PC_MAX_LEVELS], M = 0;
    if (levels > PC_MAX_LEVELS) return fail(PC_ERR_ARG, "too many levels");
    int rc = pc_greedy_partition(blockSizes, levels, filterLength, Ng, Pg, &M);
    if (rc) return rc;
    std::vector<std::vector<float>> irs(k, std::vector<float>(filterLength));
    for (int i = 0; i < k; i++)
        for (int l = 0; l < filterLength; l++) irs[i][l] = impulses[(size_t)stride * ((size_t)filterLength * i + l)];
    std::vector<InstDesc> d;
    for (auto &e : map) {
        InstDesc x;
        x.in_ch = e[1];
        x.out_ch = e[2];
        x.N.assign(Ng, Ng + M);
        x.P.assign(Pg, Pg + M);
        x.ir = irs[e[0]];
        d.push_back(std::move(x));
    }
    return multi_setup(h, FS, blockSizes[0], n, m, d);
}

// PartConvMax::run (src/partconv.cpp:1219-1303): one tick on per-channel
// buffers of buffer_size samples (input[c], output[c]).
PC_API int pcm_run_channels(pc_multi *h, float *const *output, const float *const *input)
{
    if (!h || !output || !input) return fail(PC_ERR_ARG, "null argument");
    Engine &E = h->E;
    if (!E.ready) return fail(PC_ERR_STATE, "run before setup");
    const int N0 = E.N0;
    for (int c = 0; c < E.nIn; c++) {
        if (!input[c]) return fail(PC_ERR_ARG, "null input channel");
        memcpy(E.h_in + (size_t)c * N0, input[c], (size_t)N0 * sizeof(float));
    }
    int rc = host_tick(h);
    if (rc) return rc;
    for (int c = 0; c < E.nOut; c++) {
        if (!output[c]) return fail(PC_ERR_ARG, "null output channel");
        memcpy(output[c], E.h_out + (size_t)c * N0, (size_t)N0 * sizeof(float));
    }
    return PC_OK;
}

PC_API int pcm_run(pc_multi *h, float *output, const float *input, int frameCount)
{
    if (!h || !output || !input) return fail(PC_ERR_ARG, "null argument");
    Engine &E = h->E;
    if (!E.ready) return fail(PC_ERR_STATE, "run before setup");
    if (frameCount != E.N0) {
        char m[96];
        snprintf(m, sizeof m, "frameCount (%d) != buffer_size (%d)", frameCount, E.N0);
        return fail(PC_ERR_FRAMECOUNT, m);
    }
    {
        const int N0 = E.N0, ni = E.nIn;
        for (int c = 0; c < ni; c++) {
            float *d = E.h_in + (size_t)c * N0;
            for (int j = 0; j < N0; j++) d[j] = input[(size_t)j * ni + c];
        }
    }
    int rc = host_tick(h);
    if (rc) return rc;
    {
        const int N0 = E.N0, no = E.nOut;
        for (int c = 0; c < no; c++) {
            const float *src = E.h_out + (size_t)c * N0;
            for (int j = 0; j < N0; j++) output[(size_t)j * no + c] = src[j];
        }
    }
    return PC_OK;
}

PC_API int pcm_process(pc_multi *h, float *output, const float *input, long long nframes)
{
    if (!h || !output || !input || nframes < 0) return fail(PC_ERR_ARG, "bad argument");
    Engine &E = h->E;
    if (!E.ready) return fail(PC_ERR_STATE, "run before setup");
    const size_t fi = (size_t)E.N0 * E.nIn, fo = (size_t)E.N0 * E.nOut;
    for (long long t = 0; t < nframes; t++) {
        int rc = pcm_run(h, output + t * fo, input + t * fi, E.N0);
        if (rc) return rc;
    }
    return PC_OK;
}

PC_API int pcm_run_device(pc_multi *h, float *d_out, const float *d_in, long long nframes,
                          int sync_each_tick)
{
    if (!h || !d_out || !d_in || nframes < 0) return fail(PC_ERR_ARG, "bad argument");
    Engine &E = h->E;
    if (!E.ready) return fail(PC_ERR_STATE, "run before setup");
    {
        int rc = session_stop(E);   // a host-buffer session ends here
        if (rc) return rc;
    }
    if (sync_each_tick == 2 && E.mode == PC_BATCHED && nframes > 0) {
        long long served = 0;
        int rc = engine_run_persistent(E, d_in, d_out, nframes, &served);
        h->persistent_ticks += served;
        if (rc != 1) return rc;  // 1: persistent launch refused -> per-tick launches below
    }
    const long long fbase = (long long)E.frame;
    for (long long t = 0; t < nframes; t++) {
        int rc = engine_tick(E, d_in, (long long)E.N0 * E.nIn, fbase, d_out, (long long)E.N0 * E.nOut,
                             sync_each_tick != 0);
        if (rc) return rc;
    }
    if (!sync_each_tick) return engine_drain(E);
    return PC_OK;
}

PC_API int pcm_reset(pc_multi *h)
{
    if (!h) return fail(PC_ERR_ARG, "null handle");
    return engine_reset(h->E);
}

PC_API int pcm_cleanup(pc_multi *h)
{
    if (!h) return fail(PC_ERR_ARG, "null handle");
    engine_free(h->E);
    return PC_OK;
}

static void multi_destroy(pc_multi *h)
{
    engine_free(h->E);
    delete h;
}

PC_API void pcm_destroy(pc_multi *h)
{
    if (!h || !live_take(h)) return;
    multi_destroy(h);
}

PC_API int pcm_get_info(const pc_multi *h, pc_multi_info *info)
{
    if (!h || !info) return fail(PC_ERR_ARG, "null argument");
    const Engine &E = h->E;
    info->FS = E.FS;
    info->buffer_size = E.N0;
    info->max_block_size = h->max_block;
    info->verbosity = E.verbosity;
    info->numInputChannels = E.nIn;
    info->numOutputChannels = E.nOut;
    info->numPCs = (int)E.inst.size();
    info->num_levels = (int)E.gsize.size();
    info->frameNum = E.frame;
    return PC_OK;
}

PC_API int pcm_get_stats(const pc_multi *h, pc_stats *st)
{
    if (!h || !st) return fail(PC_ERR_ARG, "null argument");
    const Engine &E = h->E;
    memset(st, 0, sizeof(*st));
    st->ticks = E.ticks;
    st->deadline_misses = E.misses;
    st->period_us = E.FS > 0 ? 1e6 * E.N0 / E.FS : 0.0;
    st->max_us = E.lat_max;
    st->persistent_ticks = (unsigned long long)h->persistent_ticks;
    st->stalls = E.stalls;
    st->mean_us = E.ticks ? E.lat_sum / (double)E.ticks : 0.0;
    const size_t n = E.lat_wrap ? E.lat.size() : E.lat_pos;
    if (n) {
        std::vector<double> v(E.lat.begin(), E.lat.begin() + n);
        auto pct = [&](double q) {
            size_t i = (size_t)std::min<double>((double)(n - 1), std::floor(q * (double)(n - 1) + 0.5));
            std::nth_element(v.begin(), v.begin() + i, v.end());
            return v[i];
        };
        st->p50_us = pct(0.50);
        st->p99_us = pct(0.99);
    }
    return PC_OK;
}

PC_API int pcm_clear_stats(pc_multi *h)
{
    if (!h) return fail(PC_ERR_ARG, "null handle");
    Engine &E = h->E;
    E.lat_pos = 0;
    E.lat_wrap = false;
    E.ticks = E.misses = E.stalls = 0;
    E.lat_sum = E.lat_max = 0;
    h->persistent_ticks = 0;
    return PC_OK;
}

PC_API long long pcm_get_latencies(const pc_multi *h, double *out_us, long long cap)
{
    if (!h || !out_us || cap < 0) return -1;
    const Engine &E = h->E;
    const size_t n = E.lat_wrap ? E.lat.size() : E.lat_pos;
    const size_t m = std::min<size_t>(n, (size_t)cap);
    for (size_t q = 0; q < m; q++) {
        size_t idx = (E.lat_pos + E.lat.size() - m + q) % E.lat.size();
        out_us[q] = E.lat[idx];
    }
    return (long long)m;
}

PC_API int pcm_instance_partition(const pc_multi *h, int inst, int *N, int *P, int cap)
{
    if (!h || inst < 0 || inst >= (int)h->E.inst.size()) return -1;
    const Inst &I = h->E.inst[inst];
    for (int L = 0; L < (int)I.lv.size() && L < cap; L++) {
        if (N) N[L] = I.lv[L].N;
        if (P) P[L] = I.lv[L].P;
    }
    return (int)I.lv.size();
}

PC_API int pcm_probe(pc_multi *h, int global_level)
{
    if (!h) return fail(PC_ERR_ARG, "null handle");
    Engine &E = h->E;
    if (!E.ready) return fail(PC_ERR_STATE, "probe before setup");
    if (global_level >= (int)E.gsize.size()) return fail(PC_ERR_ARG, "level out of range");
    HIPCHK(hipStreamSynchronize(E.s_main));
    for (hipStream_t s : E.s_lvl) HIPCHK(hipStreamSynchronize(s));
    if (E.pa.empty()) {
        E.pa.resize(16);
        E.pb.resize(16);
        E.pused.assign(16, false);
        for (int q = 0; q < 16; q++) {
            HIPCHK(hipEventCreate(&E.pa[q]));
            HIPCHK(hipEventCreate(&E.pb[q]));
        }
    }
    std::fill(E.pused.begin(), E.pused.end(), false);
    E.probe_g = global_level;
    E.probe_ms = 0;
    E.probe_n = 0;
    E.probe_bytes = 0;
    // algorithmic bytes per launch.  Level >= 1 (k_partial): per job the P-1
    // delay-line rows and P-1 filter rows it reads plus the S row it writes,
    // 8N(2P-1).  Tick kernel (level 0): per instance the level-block figure
    // 8(N0+1)(2P0+1) + 12 N0 (its phase B streams the delay line and filter
    // rows; phase A reads the window, H_0 and S and writes the output).
    if (global_level >= 0) {
        for (Inst &I : E.inst) {
            const int L = I.gmap[global_level];
            if (L < 0) continue;
            const LevelSpec &s = I.lv[L];
            if (global_level == 0) E.probe_bytes += 8.0 * (s.N + 1) * (2.0 * s.P + 1) + 12.0 * s.N;
            else E.probe_bytes += 8.0 * s.N * (2.0 * s.P - 1);
        }
    }
    return PC_OK;
}

PC_API int pcm_probe_read(pc_multi *h, double *avg_us, long long *launches, double *bytes_per_launch)
{
    if (!h) return fail(PC_ERR_ARG, "null handle");
    Engine &E = h->E;
    for (int q = 0; q < (int)E.pa.size(); q++) {
        int rc = probe_fold(E, q);
        if (rc) return rc;
    }
    if (avg_us) *avg_us = E.probe_n ? 1e3 * E.probe_ms / (double)E.probe_n : 0.0;
    if (launches) *launches = E.probe_n;
    if (bytes_per_launch) *bytes_per_launch = E.probe_bytes;
    return PC_OK;
}

PC_API int pcm_trace(pc_multi *h, long long nticks)
{
    if (!h) return fail(PC_ERR_ARG, "null handle");
    Engine &E = h->E;
    {
        int rc = session_stop(E);
        if (rc) return rc;
    }
    HIPCHK(hipDeviceSynchronize());
    hipFree(E.d_trace);
    E.d_trace = nullptr;
    E.trace_cap = 0;
    if (nticks > 0) {
        HIPCHK(hipMalloc(&E.d_trace, (size_t)nticks * PC_TRACE_W * sizeof(unsigned long long)));
        HIPCHK(hipMemset(E.d_trace, 0, (size_t)nticks * PC_TRACE_W * sizeof(unsigned long long)));
        E.trace_cap = nticks;
    }
    return PC_OK;
}

PC_API int pcm_set_persistent_workers(pc_multi *h, int mode)
{
    if (!h) return fail(PC_ERR_ARG, "null handle");
    if (mode < 0 || mode > 2) return fail(PC_ERR_ARG, "workers mode must be 0, 1 or 2");
    int rc = session_stop(h->E);
    if (rc) return rc;
    h->E.persist_workers = mode;
    h->E.nworkers = -1;   // plan again
    hipFree(h->E.d_wrange);
    h->E.d_wrange = nullptr;
    return PC_OK;
}

PC_API int pcm_set_partial_width(pc_multi *h, int bins_per_thread)
{
    if (!h) return fail(PC_ERR_ARG, "null handle");
    if (bins_per_thread != 2 && bins_per_thread != 4 && bins_per_thread != 8)
        return fail(PC_ERR_ARG, "bins per thread must be 2, 4 or 8");
    int rc = session_stop(h->E);
    if (rc) return rc;
    if (h->E.ready) {  // level blocks in flight use the width they were launched with
        rc = engine_drain(h->E);
        if (rc) return rc;
    }
    h->E.part_q = bins_per_thread / 2;
    level_graphs_free(h->E);  // rebuilt with the new width at the next use
    return PC_OK;
}

PC_API int pcm_set_level_graphs(pc_multi *h, int on)
{
    if (!h) return fail(PC_ERR_ARG, "null handle");
    int rc = session_stop(h->E);
    if (rc) return rc;
    if (h->E.ready && (rc = engine_drain(h->E))) return rc;
    h->E.level_graphs = on != 0;
    return PC_OK;
}

PC_API int pcm_get_tuning(const pc_multi *h, int *workers_mode, int *bins_per_thread)
{
    if (!h) return fail(PC_ERR_ARG, "null handle");
    if (workers_mode) *workers_mode = h->E.persist_workers;
    if (bins_per_thread) *bins_per_thread = 2 * h->E.part_q;
    return PC_OK;
}

PC_API int pcm_set_release_window(pc_multi *h, int ticks)
{
    if (!h) return fail(PC_ERR_ARG, "null handle");
    if (ticks < 1 || ticks > PC_RELEASE_WINDOW_MAX) return fail(PC_ERR_ARG, "release window must be 1 .. 16 ticks");
    h->E.release_window = ticks;
    return PC_OK;
}

PC_API int pcm_set_host_session(pc_multi *h, int on)
{
    if (!h) return fail(PC_ERR_ARG, "null handle");
    int rc = session_stop(h->E);
    if (rc) return rc;
    if (on < 0 || on > 2) return fail(PC_ERR_ARG, "host session mode must be 0, 1 or 2");
    h->E.host_session = on;
    return PC_OK;
}

PC_API int pcm_workers_info(pc_multi *h, int *self_mode)
{
    if (!h) return -1;
    if (self_mode) *self_mode = 0;
    if (!h->E.ready || h->E.nwjobs == 0) return 0;
    if (plan_workers(h->E)) return -1;
    if (self_mode) *self_mode = h->E.self_levels ? 1 : 0;
    return h->E.nworkers;
}

PC_API int pcm_stall_level(pc_multi *h, int global_level, double seconds)
{
    if (!h) return fail(PC_ERR_ARG, "null handle");
    Engine &E = h->E;
    if (!E.ready) return fail(PC_ERR_STATE, "stall before setup");
    if (global_level < 1 || global_level >= (int)E.gsize.size()) return fail(PC_ERR_ARG, "level out of range");
    if (!(seconds > 0.0 && seconds <= 5.0)) return fail(PC_ERR_ARG, "stall must be in (0, 5] seconds");
    E.stall_us.store((long long)(seconds * 1e6));
    E.stall_g.store(global_level);
    return PC_OK;
}

PC_API int pcm_release_info(const pc_multi *h)
{
    if (!h || !h->E.ready) return -1;
    return h->E.go_slot >= 0 ? 1 : 0;
}

PC_API long long pcm_trace_read(pc_multi *h, unsigned long long *out, long long cap)
{
    if (!h || !out) return -1;
    Engine &E = h->E;
    const long long n = std::min(cap, E.trace_cap);
    if (n <= 0) return 0;
    if (hipDeviceSynchronize() != hipSuccess) return -1;
    if (hipMemcpy(out, E.d_trace, (size_t)n * PC_TRACE_W * sizeof(unsigned long long), hipMemcpyDeviceToHost) !=
        hipSuccess)
        return -1;
    return n;
}

// ----------------------------------------------------------------------------
// C ABI: PartConvFilter (one instance, caller-driven)
// ----------------------------------------------------------------------------
struct pc_filter {
    Engine E;
    int FS = 44100;
    int M = 0;
    std::vector<int> N, P;
    std::vector<int> num_N0, ind_N0, rb;
    std::vector<unsigned long long> q, bcurr;
    int start_sync_depth = 0, end_sync_depth = 0;
    unsigned long long nin = 0;
};

static void filter_rewind(pc_filter *h)
{
    for (int L = 0; L < h->M; L++) {
        h->ind_N0[L] = h->num_N0[L] - 1;
        h->q[L] = 0;
        h->bcurr[L] = 0;
        h->rb[L] = 0;
    }
    h->nin = 0;
    h->start_sync_depth = h->end_sync_depth = 0;
}

PC_API pc_filter *pcf_create(int verbosity, int sample_rate)
{
    pc_filter *h = new pc_filter();
    h->E.verbosity = verbosity;
    h->FS = sample_rate;
    if (!live_add(h, HK_FILTER)) {
        delete h;
        fail(PC_ERR_STATE, "library shut down");
        return nullptr;
    }
    return h;
}

PC_API int pcf_setup(pc_filter *h, int M, const int *N, const int *num_parts, const float *impulse,
                     int impulse_len, int fftwOptLevel, const char *wisdom)
{
    (void)fftwOptLevel;
    (void)wisdom;
    if (!h || M < 1 || !N || !num_parts) return fail(PC_ERR_ARG, "bad argument");
    InstDesc d;
    d.N.assign(N, N + M);
    d.P.assign(num_parts, num_parts + M);
    if (impulse) {
        if (impulse_len < 0) return fail(PC_ERR_ARG, "negative impulse length");
        d.ir.assign(impulse, impulse + impulse_len);
    } else {
        long long len = 0;
        for (int L = 0; L < M; L++) len += (long long)N[L] * num_parts[L];
        d.ir.assign((size_t)std::max(0ll, len), 1.0f);
    }
    std::vector<InstDesc> v{d};
    h->E.mode = PC_PLAIN;
    int rc = engine_setup(h->E, h->FS, N[0], 1, 1, v);
    if (rc) {
        std::string msg = g_err;
        engine_free(h->E);
        g_err = msg;
        return rc;
    }
    const Inst &I = h->E.inst[0];
    h->M = M;
    h->N.assign(N, N + M);
    h->P.resize(M);
    for (int L = 0; L < M; L++) h->P[L] = I.lv[L].P;  // recounted; the caller's array is not written
    h->num_N0.resize(M);
    h->ind_N0.resize(M);
    h->rb.resize(M);
    h->q.resize(M);
    h->bcurr.resize(M);
    for (int L = 0; L < M; L++) h->num_N0[L] = N[L] / N[0];
    filter_rewind(h);
    return PC_OK;
}

static hipStream_t filter_stream(pc_filter *h, int L)
{
    if (L == 0) return h->E.s_main;
    return level_stream(h->E, L);  // global level == local level for one instance
}

PC_API int pcf_run_level(pc_filter *h, int L)
{
    if (!h || !h->E.ready) return fail(PC_ERR_STATE, "runLevel before setup");
    if (L < 0 || L >= h->M) return fail(PC_ERR_ARG, "level out of range");
    Engine &E = h->E;
    if (h->q[L] == 0) return fail(PC_ERR_STATE, "runLevel before any input for this level");
    const uint64_t m = h->q[L] - 1;
    const uint64_t b = h->bcurr[L]++;
    hipStream_t s = filter_stream(h, L);
    if (L == 0) {
        int rc = launch_small(E.N0, E.groups[0].d_l0, 1, E.N0, m, b, 0, E.tw, PC_MODE_RUN, s);
        if (rc) return rc;
        launch_partial(E.N0, E.groups[0].d_l0, 1, b + 1, s);
    } else {
        HIPCHK(hipEventRecord(E.ev_tick, E.s_main));
        HIPCHK(hipStreamWaitEvent(s, E.ev_tick, 0));
        for (Batch &bt : E.groups[0].batches) {
            if (bt.g != L) continue;
            const int ypar = h->rb[L] ^ 1;
            if (bt.big) launch_big(bt.sh, bt.d_jobs, 1, E.N0, m, b, ypar, E.tw, PC_MODE_RUN, s);
            else {
                int rc = launch_small(bt.N, bt.d_jobs, 1, E.N0, m, b, ypar, E.tw, PC_MODE_RUN, s);
                if (rc) return rc;
            }
            launch_partial(bt.N, bt.d_jobs, 1, b + 1, s);
        }
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipEventRecord(E.ev_lvl[L], s));
    E.lvl_pending[L] = true;
    return PC_OK;
}

PC_API int pcf_sync_levels(pc_filter *h)
{
    if (!h || !h->E.ready) return fail(PC_ERR_STATE, "sync_levels before setup");
    h->start_sync_depth = h->end_sync_depth = 0;
    for (int L = 0; L < h->M; L++) {
        if (++h->ind_N0[L] == h->num_N0[L]) {
            h->ind_N0[L] = 0;
            h->start_sync_depth = L;
        } else if (h->ind_N0[L] == h->num_N0[L] - 1)
            h->end_sync_depth = L;
    }
    return PC_OK;
}

PC_API int pcf_sync_levels_relaxed(pc_filter *h)
{
    if (!h || !h->E.ready) return fail(PC_ERR_STATE, "sync_levels before setup");
    h->start_sync_depth = h->end_sync_depth = 0;
    for (int L = 0; L < h->M; L++) {
        if (++h->ind_N0[L] == h->num_N0[L]) {
            h->ind_N0[L] = 0;
            h->start_sync_depth = L;
            h->end_sync_depth = L;
        }
    }
    return PC_OK;
}

PC_API int pcf_start_sync_depth(const pc_filter *h) { return h ? h->start_sync_depth : -1; }
PC_API int pcf_end_sync_depth(const pc_filter *h) { return h ? h->end_sync_depth : -1; }

PC_API int pcf_input(pc_filter *h, const float *input)
{
    if (!h || !input) return fail(PC_ERR_ARG, "null argument");
    if (!h->E.ready) return fail(PC_ERR_STATE, "input before setup");
    Engine &E = h->E;
    HIPCHK(hipStreamSynchronize(E.s_main));  // staging buffer reuse
    memcpy(E.h_in, input, (size_t)E.N0 * sizeof(float));
    const Inst &I = E.inst[0];
    hipLaunchKernelGGL(k_ring_write, dim3((E.N0 + 255) / 256), dim3(256), 0, E.s_main,
                       reinterpret_cast<float *>(E.arena + I.off_ring), (long long)I.R - 1,
                       (long long)(h->nin * (unsigned long long)E.N0), E.d_in_map, E.N0);
    HIPCHK(hipGetLastError());
    h->nin++;
    for (int L = 0; L < h->M; L++)
        if (L <= h->start_sync_depth) h->q[L]++;
    return PC_OK;
}

static int filter_output(pc_filter *h, float *y, bool relaxed)
{
    if (!h || !y) return fail(PC_ERR_ARG, "null argument");
    if (!h->E.ready) return fail(PC_ERR_STATE, "output before setup");
    Engine &E = h->E;
    const Inst &I = E.inst[0];
    FilterOut fo{};
    fo.src[fo.nsrc++] = reinterpret_cast<const float *>(E.arena + I.lv[0].off_ybuf);
    for (int L = 1; L < h->M; L++) {
        if (L <= h->end_sync_depth) h->rb[L] ^= 1;
        int r = relaxed ? h->ind_N0[L] : h->ind_N0[L] + 1;
        if (r == h->num_N0[L]) r = 0;
        fo.src[fo.nsrc++] = reinterpret_cast<const float *>(E.arena + I.lv[L].off_ybuf) +
                            (size_t)h->rb[L] * I.lv[L].N + (size_t)r * E.N0;
    }
    for (int L = 1; L < h->M; L++)
        if (E.lvl_pending[L]) HIPCHK(hipStreamWaitEvent(E.s_main, E.ev_lvl[L], 0));
    hipLaunchKernelGGL(k_filter_out, dim3((E.N0 + 255) / 256), dim3(256), 0, E.s_main, fo, E.N0, E.d_out_map);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(E.s_main));
    memcpy(y, E.h_out, (size_t)E.N0 * sizeof(float));
    return PC_OK;
}

PC_API int pcf_output(pc_filter *h, float *output) { return filter_output(h, output, false); }
PC_API int pcf_output_relaxed(pc_filter *h, float *output) { return filter_output(h, output, true); }

PC_API int pcf_reset(pc_filter *h)
{
    if (!h) return fail(PC_ERR_ARG, "null handle");
    int rc = engine_reset(h->E);
    if (rc) return rc;
    filter_rewind(h);
    return PC_OK;
}

PC_API double pcf_bench(pc_filter *h, int L, int trials, int pollute_every, float max_time,
                        double *benchData)
{
    (void)pollute_every;
    if (!h || !h->E.ready || L < 0 || L >= h->M || trials < 1) {
        fail(PC_ERR_ARG, "bench: bad argument or not set up");
        return -1.0;
    }
    hipEvent_t a, b;
    if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) {
        fail(PC_ERR_HIP, "bench: event creation failed");
        return -1.0;
    }
    // give the level a block to read
    for (int L2 = 0; L2 < h->M; L2++)
        if (h->q[L2] == 0) h->q[L2] = 1;
    hipStream_t s = filter_stream(h, L);
    double total = 0.0;
    int t;
    for (t = 0; t < trials; t++) {
        hipEventRecord(a, s);
        if (pcf_run_level(h, L) != PC_OK) {
            hipEventDestroy(a);
            hipEventDestroy(b);
            return -1.0;
        }
        hipEventRecord(b, s);
        hipEventSynchronize(b);
        float ms = 0.f;
        hipEventElapsedTime(&ms, a, b);
        total += 1e-3 * ms;
        if (total > max_time) {
            t++;
            break;
        }
    }
    hipEventDestroy(a);
    hipEventDestroy(b);
    const double mean = total / (double)t;
    if (benchData) {
        const int value = (int)(100.0 * mean / ((double)h->N[L] / (h->FS > 0 ? h->FS : 1)));
        benchData[2 * L] = value;
        benchData[2 * L + 1] = value;
    }
    if (pcf_reset(h) != PC_OK) return -1.0;
    return mean;
}

PC_API int pcf_cleanup(pc_filter *h)
{
    if (!h) return fail(PC_ERR_ARG, "null handle");
    engine_free(h->E);
    h->M = 0;
    return PC_OK;
}

static void filter_destroy(pc_filter *h)
{
    engine_free(h->E);
    delete h;
}

PC_API void pcf_destroy(pc_filter *h)
{
    if (!h || !live_take(h)) return;
    filter_destroy(h);
}

PC_API int pcf_get_levels(const pc_filter *h, int *M, int *N, int *num_parts, int cap)
{
    if (!h) return fail(PC_ERR_ARG, "null handle");
    if (M) *M = h->M;
    for (int L = 0; L < h->M && L < cap; L++) {
        if (N) N[L] = h->N[L];
        if (num_parts) num_parts[L] = h->P[L];
    }
    return PC_OK;
}

PC_API int pcf_get_fs(const pc_filter *h) { return h ? h->FS : -1; }
PC_API int pcf_set_fs(pc_filter *h, int fs)
{
    if (!h) return fail(PC_ERR_ARG, "null handle");
    h->FS = fs;
    h->E.FS = fs;
    return PC_OK;
}

// ============================================================================
// Offline batched convolution and adjoint (pcc_*): see pc_offline.hpp.
// ============================================================================
#include "pc_offline.hpp"

__global__ void k_fill_zero(float *p, long long n)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.f;
}

// Offline transforms of block sizes up to this run in one workgroup with the
// data in registers (pc_fftreg.hpp); larger ones (32768, 65536) use the
// four-step pair through a staging buffer.
#define PC_OFF_REG_MAX 16384

struct ConvLevel {
    int N, P, T;
    bool big;  // four-step transforms
    BigShape sh;
    cf *H = nullptr;
    // blocks per chunk (one workgroup column each) of the spectral products
    // and of the dL/dh partial sums: defaults, or chosen by pcc_tune
    int mac_mc = 32, dh_mc = 64;
};

// One level's work space and stream for level-parallel calls: the levels of
// a forward or backward run side by side on their own streams (forked from
// and joined back to the caller's stream with events); only the inverse
// transforms that add into y (forward) or dL/dx (backward) are chained in
// level order, so every result is bitwise the same as with one stream.
struct ConvLevelWork {
    hipStream_t s = nullptr;
    hipEvent_t ev = nullptr;  // chain / join point of this level
    cf *Y = nullptr, *G = nullptr, *W = nullptr, *Xr = nullptr, *part = nullptr, *D = nullptr, *scr = nullptr;
    size_t Yc = 0, Gc = 0, Wc = 0, Xrc = 0, partc = 0, Dc = 0, scrc = 0;
};

// A forward or backward call's launch sequence captured as a graph and
// replayed by later calls with the same arguments: key (the call's pointers),
// shape, flags (kept spectra) and fingerprint (every buffer, chunk size and
// mode the captured launches embed) must all match.
#define PC_CG_KEYS 6  // pointers that identify a call's arguments
struct CallGraph {
    hipGraphExec_t exec = nullptr;
    const void *key[PC_CG_KEYS] = {};
    int C = 0, T = 0, flags = 0;
    std::vector<uintptr_t> fp;
};

// Graphs of one direction and the last eager calls' arguments, most recent
// first: a call is captured the second time its arguments are seen.
#define PC_CALL_GRAPHS 4
struct CallGraphs {
    bool bwd = false;
    std::vector<CallGraph> g, seen;
    long long replays = 0, captures = 0;
};

struct pc_conv {
    int dev = 0;
    int taps = 0;
    std::vector<ConvLevel> lv;
    std::vector<ConvLevelWork> lw;
    hipEvent_t fork_ev = nullptr;
    int level_streams = PCC_LEVEL_STREAMS_CAPTURED;
    bool call_graphs = true;
    CallGraphs fg, bg;  // forward, backward
    hipStream_t cap_s = nullptr;
    cf *tw = nullptr;
    // filter state
    const float *h_src = nullptr;
    bool filter_ready = false;
    // work space (grown on demand)
    cf *work = nullptr, *scr = nullptr, *dpart = nullptr, *dsum = nullptr;
    size_t work_cap = 0, scr_cap = 0, dpart_cap = 0, dsum_cap = 0;
    // spectra kept from forward for backward (per level), when within budget
    std::vector<cf *> keep;
    std::vector<size_t> keep_cap;
    long long budget = 1ll << 30;
    const float *keep_x = nullptr;
    int keep_C = 0, keep_T = 0;
    bool keep_valid = false;
    long long recomputed_levels = 0;
    float *mse_part = nullptr;  // pcc_mse's per-block partial sums (PC_MSE_BLOCKS)
    // probe: HIP events around the forward spectral product of one level
    int probe_level = -1, probe_kind = PCC_PROBE_PRODUCT;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> probe_ev;
    double probe_ms = 0, probe_bytes = 0;
    long long probe_n = 0;
};

// Spectral products of one level over every block of every clip (dir +1:
// causal, -1: adjoint): the register-window kernel (window of 4, 8 or 16)
// for up to 16 partitions, else one thread per (clip, block, bin pair).  With events (e0, e1), the
// launch stamps them.
template <typename K, typename... Args>
static void off_launch(K kern, dim3 grid, hipStream_t s, hipEvent_t e0, hipEvent_t e1, Args... args)
{
    if (e0)
        hipExtLaunchKernelGGL(kern, grid, dim3(256), 0, s, e0, e1, 0, args...);
    else
        hipLaunchKernelGGL(kern, grid, dim3(256), 0, s, args...);
}

// With events (e0, e1), a launch stamps them (the probe); e0 alone or e1
// alone stamps the start or the end of a two-kernel sequence.
template <typename K, typename... Args>
static void off_launch_ev(K kern, dim3 grid, dim3 block, size_t lds, hipStream_t s, hipEvent_t e0, hipEvent_t e1,
                          Args... args)
{
    if (e0 || e1)
        hipExtLaunchKernelGGL(kern, grid, block, lds, s, e0, e1, 0, args...);
    else
        hipLaunchKernelGGL(kern, grid, block, lds, s, args...);
}

static void off_mac(const cf *X, const cf *H, cf *Y, int C, int M, const ConvLevel &L, int dir, hipStream_t s,
                    hipEvent_t e0, hipEvent_t e1)
{
    const long long items = (long long)C * (L.N / 2);
    const int mc = L.mac_mc;
    const dim3 grid((unsigned)((items + 255) / 256), (unsigned)((M + mc - 1) / mc));
    if (L.P <= 4) {
        if (dir > 0) off_launch(k_off_mac_w<4, 1>, grid, s, e0, e1, X, H, Y, C, M, L.N, L.P, mc);
        else off_launch(k_off_mac_w<4, -1>, grid, s, e0, e1, X, H, Y, C, M, L.N, L.P, mc);
    } else if (L.P <= 8) {
        if (dir > 0) off_launch(k_off_mac_w<8, 1>, grid, s, e0, e1, X, H, Y, C, M, L.N, L.P, mc);
        else off_launch(k_off_mac_w<8, -1>, grid, s, e0, e1, X, H, Y, C, M, L.N, L.P, mc);
    } else if (L.P <= 16) {
        if (dir > 0) off_launch(k_off_mac_w<16, 1>, grid, s, e0, e1, X, H, Y, C, M, L.N, L.P, mc);
        else off_launch(k_off_mac_w<16, -1>, grid, s, e0, e1, X, H, Y, C, M, L.N, L.P, mc);
    } else {
        off_launch(k_off_mac, dim3((unsigned)((items * M + 255) / 256)), s, e0, e1, X, H, Y, C, M, L.N, L.P, dir);
    }
}

// dL/dh of one level into D (P spectra): partial sums per clip and chunk of
// L.dh_mc blocks, PC_DH_PT partitions per thread, then the fixed-order reduce.
static int off_dh(pc_conv *c, const ConvLevel &L, const cf *G, const cf *X, cf *D, int C, int M, hipStream_t s,
                  hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr, cf **part = nullptr, size_t *part_cap = nullptr);

// Fold the probe's pending events into its totals (waits for them).
static int conv_probe_fold(pc_conv *c)
{
    for (auto &e : c->probe_ev) {
        HIPCHK(hipEventSynchronize(e.second));
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, e.first, e.second));
        c->probe_ms += ms;
        c->probe_n++;
        HIPCHK(hipEventDestroy(e.first));
        HIPCHK(hipEventDestroy(e.second));
    }
    c->probe_ev.clear();
    return PC_OK;
}

// Events for the probed launch when (level, kind) is the one probed, else
// nulls; probe_end files them with the launch's algorithmic bytes.
static int probe_begin(pc_conv *c, size_t li, int kind, hipEvent_t *e0, hipEvent_t *e1)
{
    *e0 = *e1 = nullptr;
    if (c->probe_level != (int)li || c->probe_kind != kind) return PC_OK;
    if (c->probe_ev.size() >= 64) {
        int rc = conv_probe_fold(c);
        if (rc) return rc;
    }
    HIPCHK(hipEventCreate(e0));
    HIPCHK(hipEventCreate(e1));
    return PC_OK;
}
static void probe_end(pc_conv *c, hipEvent_t e0, hipEvent_t e1, double bytes)
{
    if (!e0) return;
    c->probe_ev.push_back({e0, e1});
    c->probe_bytes = bytes;
}

static int conv_grow(cf **p, size_t *cap, size_t need)
{
    if (need <= *cap) return PC_OK;
    if (*p) HIPCHK(hipFree(*p));
    *p = nullptr;
    HIPCHK(hipMalloc(p, need * sizeof(cf)));
    *cap = need;
    return PC_OK;
}

static int off_dh(pc_conv *c, const ConvLevel &L, const cf *G, const cf *X, cf *D, int C, int M, hipStream_t s,
                  hipEvent_t e0, hipEvent_t e1, cf **part, size_t *part_cap)
{
    const int nch = (M + L.dh_mc - 1) / L.dh_mc;
    if (!part) {
        part = &c->dpart;
        part_cap = &c->dpart_cap;
    }
    int rc = conv_grow(part, part_cap, (size_t)L.P * C * nch * L.N);
    if (rc) return rc;
    cf *dpart = *part;
    if (L.P > PC_DH_PT) {
        // one bin per thread, up to 16 partitions per group: G and X read once
        // for P <= 16 (the bin-pair form would read them ceil(P / 8) times)
        const long long items = (long long)C * L.N;
        off_launch_ev(k_off_dh_part_b<16>,
                      dim3((unsigned)((items + 255) / 256), (unsigned)((L.P + 15) / 16), (unsigned)nch), dim3(256),
                      0, s, e0, e1, G, X, dpart, C, M, L.N, L.P, L.dh_mc);
    } else {
        const long long items = (long long)C * (L.N / 2);
        off_launch_ev(k_off_dh_part<PC_DH_PT>,
                      dim3((unsigned)((items + 255) / 256), (unsigned)((L.P + PC_DH_PT - 1) / PC_DH_PT),
                           (unsigned)nch),
                      dim3(256), 0, s, e0, e1, G, X, dpart, C, M, L.N, L.P, L.dh_mc);
    }
    const long long nd = (long long)L.P * L.N;
    hipLaunchKernelGGL(k_off_dh_reduce, dim3((unsigned)((nd + 15) / 16)), dim3(256), 0, s, dpart, D, C * nch,
                       L.N, L.P);
    HIPCHK(hipGetLastError());
    return PC_OK;
}

template <int LOGN>
static void off_rfft_t(const WinSrc &W, int ntrans, cf *X, const cf *tw, hipStream_t s, hipEvent_t e0, hipEvent_t e1)
{
    using F = RegFft<LOGN>;
    off_launch_ev(k_off_rfft_reg<LOGN>, dim3((unsigned)((ntrans + F::TPB - 1) / F::TPB)), dim3(F::NT), 0, s, e0, e1,
                  W, ntrans, X, tw);
}
template <int LOGN>
static void off_irfft_t(const cf *Y, const InvDst &D, int ntrans, const cf *tw, hipStream_t s, hipEvent_t e0,
                        hipEvent_t e1)
{
    using F = RegFft<LOGN>;
    off_launch_ev(k_off_irfft_reg<LOGN>, dim3((unsigned)((ntrans + F::TPB - 1) / F::TPB)), dim3(F::NT), 0, s, e0,
                  e1, Y, D, ntrans, tw);
}

// batched r2c of `rows * W.nb` windows into X (N float2 each, natural bin
// order for block sizes up to PC_OFF_REG_MAX, four-step slot order above:
// the products only need every spectrum of a level in one order).  Big levels
// stage through `scr` (rows * W.nb * N float2) if given, else the handle's
// grown work space.
static int off_rfft(pc_conv *c, const ConvLevel &L, const WinSrc &W, int rows, cf *X, hipStream_t s,
                    cf *scr = nullptr, hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr)
{
    const int nt = rows * W.nb;
    if (nt == 0) return PC_OK;
    if (L.big) {
        if (!scr) {
            int rc = conv_grow(&c->scr, &c->scr_cap, (size_t)nt * L.N);
            if (rc) return rc;
            scr = c->scr;
        }
        const int A = 1 << L.sh.logA, B = 1 << L.sh.logB;
        off_launch_ev(k_off_big_fwd1, dim3((A / BIG_G) * nt), dim3(BIG_NT), big1_lds(L.sh), s, e0, nullptr, W,
                      L.sh, scr, c->tw);
        off_launch_ev(k_off_big_fwd2, dim3((B / 2) * nt), dim3(BIG_NT), big2_lds(L.sh), s, nullptr, e1, L.sh, scr,
                      X, W.scale, W.dscale, c->tw);
    } else {
        switch (ilog2(L.N)) {
        case 4: off_rfft_t<4>(W, nt, X, c->tw, s, e0, e1); break;
        case 5: off_rfft_t<5>(W, nt, X, c->tw, s, e0, e1); break;
        case 6: off_rfft_t<6>(W, nt, X, c->tw, s, e0, e1); break;
        case 7: off_rfft_t<7>(W, nt, X, c->tw, s, e0, e1); break;
        case 8: off_rfft_t<8>(W, nt, X, c->tw, s, e0, e1); break;
        case 9: off_rfft_t<9>(W, nt, X, c->tw, s, e0, e1); break;
        case 10: off_rfft_t<10>(W, nt, X, c->tw, s, e0, e1); break;
        case 11: off_rfft_t<11>(W, nt, X, c->tw, s, e0, e1); break;
        case 12: off_rfft_t<12>(W, nt, X, c->tw, s, e0, e1); break;
        case 13: off_rfft_t<13>(W, nt, X, c->tw, s, e0, e1); break;
        case 14: off_rfft_t<14>(W, nt, X, c->tw, s, e0, e1); break;
        default: return fail(PC_ERR_ARG, "block size");
        }
    }
    HIPCHK(hipGetLastError());
    return PC_OK;
}

// batched c2r of `rows * D.nb` spectra (all blocks, or one parity class);
// big levels stage through `scr` (rows * D.nb * N float2) if given, else the
// handle's grown work space.
static int off_irfft(pc_conv *c, const ConvLevel &L, const cf *Y, const InvDst &D, int rows, hipStream_t s,
                     cf *scr = nullptr, hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr)
{
    const int nt = (D.parity >= 0) ? rows * ((D.nb + 1 - D.parity) / 2) : rows * D.nb;
    if (nt == 0) return PC_OK;
    if (L.big) {
        if (!scr) {
            int rc = conv_grow(&c->scr, &c->scr_cap, (size_t)rows * D.nb * L.N);
            if (rc) return rc;
            scr = c->scr;
        }
        const int A = 1 << L.sh.logA, B = 1 << L.sh.logB;
        off_launch_ev(k_off_big_inv1, dim3((B / 2) * nt), dim3(BIG_NT), big2_lds(L.sh), s, e0, nullptr, L.sh, Y,
                      scr, c->tw, D.parity, D.nb);
        off_launch_ev(k_off_big_inv2, dim3((A / BIG_G) * nt), dim3(BIG_NT), big1_lds(L.sh), s, nullptr, e1, L.sh,
                      scr, D, c->tw);
    } else {
        switch (ilog2(L.N)) {
        case 4: off_irfft_t<4>(Y, D, nt, c->tw, s, e0, e1); break;
        case 5: off_irfft_t<5>(Y, D, nt, c->tw, s, e0, e1); break;
        case 6: off_irfft_t<6>(Y, D, nt, c->tw, s, e0, e1); break;
        case 7: off_irfft_t<7>(Y, D, nt, c->tw, s, e0, e1); break;
        case 8: off_irfft_t<8>(Y, D, nt, c->tw, s, e0, e1); break;
        case 9: off_irfft_t<9>(Y, D, nt, c->tw, s, e0, e1); break;
        case 10: off_irfft_t<10>(Y, D, nt, c->tw, s, e0, e1); break;
        case 11: off_irfft_t<11>(Y, D, nt, c->tw, s, e0, e1); break;
        case 12: off_irfft_t<12>(Y, D, nt, c->tw, s, e0, e1); break;
        case 13: off_irfft_t<13>(Y, D, nt, c->tw, s, e0, e1); break;
        case 14: off_irfft_t<14>(Y, D, nt, c->tw, s, e0, e1); break;
        default: return fail(PC_ERR_ARG, "block size");
        }
    }
    HIPCHK(hipGetLastError());
    return PC_OK;
}

static int conv_fill_zero(float *p, long long n, hipStream_t s)
{
    if (n <= 0) return PC_OK;
    hipLaunchKernelGGL(k_fill_zero, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, n);
    HIPCHK(hipGetLastError());
    return PC_OK;
}

static pc_conv *conv_create(int M, const int *N, const int *P, int taps)
{
    if (lib_shut()) {
        fail(PC_ERR_STATE, "library shut down");
        return nullptr;
    }
    if (M < 1 || !N || !P || taps < 1) {
        fail(PC_ERR_ARG, "bad argument");
        return nullptr;
    }
    std::vector<int> n(N, N + M), p(P, P + M), delay, T;
    if (check_partition(n, p, taps, delay, T)) return nullptr;
    for (int L = 0; L < M; L++)
        if (n[L] < PC_MIN_BLOCK || n[L] > PC_MAX_BLOCK || (n[L] > 2048 && n[L] < PC_BIG_MIN)) {
            fail(PC_ERR_PARTITION, "block size outside the supported range");
            return nullptr;
        }
    pc_conv *c = new pc_conv();
    c->bg.bwd = true;
    if (hipGetDevice(&c->dev) != hipSuccess || get_twiddles(c->dev, &c->tw) != PC_OK) {
        std::string msg = g_err.empty() ? "no HIP device" : g_err;
        delete c;
        fail(PC_ERR_HIP, msg);
        return nullptr;
    }
    hip_live();
    c->taps = taps;
    for (int L = 0; L < M; L++) {
        ConvLevel cl;
        cl.N = n[L];
        cl.P = p[L];
        cl.T = T[L];
        cl.big = n[L] > PC_OFF_REG_MAX;
        if (cl.big) cl.sh = big_shape(n[L]);
        if (hipMalloc(&cl.H, (size_t)cl.P * cl.N * sizeof(cf)) != hipSuccess) {
            fail(PC_ERR_HIP, "hipMalloc H");
            for (ConvLevel &q : c->lv) hipFree(q.H);
            delete c;
            return nullptr;
        }
        c->lv.push_back(cl);
    }
    c->keep.assign(M, nullptr);
    c->keep_cap.assign(M, 0);
    c->lw.resize(M);
    return c;
}

static void conv_destroy(pc_conv *c);

PC_API pc_conv *pcc_create(int M, const int *N, const int *P, int taps)
{
    pc_conv *c = conv_create(M, N, P, taps);
    if (c && !live_add(c, HK_CONV)) {
        conv_destroy(c);
        fail(PC_ERR_STATE, "library shut down");
        return nullptr;
    }
    return c;
}

PC_API int pcc_levels(const pc_conv *c, int *N, int *P, int cap)
{
    if (!c) return -1;
    for (int L = 0; L < (int)c->lv.size() && L < cap; L++) {
        if (N) N[L] = c->lv[L].N;
        if (P) P[L] = c->lv[L].P;
    }
    return (int)c->lv.size();
}

PC_API int pcc_set_budget(pc_conv *c, long long bytes)
{
    if (!c) return fail(PC_ERR_ARG, "null handle");
    c->budget = bytes;
    return PC_OK;
}

PC_API long long pcc_recomputed_levels(const pc_conv *c) { return c ? c->recomputed_levels : -1; }

// H_p = r2c([0_N | h[T_L + pN, T_L + (p+1)N)]) / 2N for every level
PC_API int pcc_set_filter(pc_conv *c, const float *d_h, void *stream)
{
    if (!c || !d_h) return fail(PC_ERR_ARG, "null argument");
    hipStream_t s = (hipStream_t)stream;
    for (ConvLevel &L : c->lv) {
        WinSrc W{};
        W.src = d_h;
        W.row_stride = 0;
        W.len = c->taps;
        W.base0 = (long long)L.T - L.N;
        W.base_step = L.N;
        W.nb = L.P;
        W.ulo = L.N;
        W.uhi = 2 * L.N;
        W.scale = 1.0f / (float)(2 * L.N);
        int rc = off_rfft(c, L, W, 1, L.H, s);
        if (rc) return rc;
    }
    c->h_src = d_h;
    c->filter_ready = true;
    return PC_OK;
}

static WinSrc conv_xwin(const ConvLevel &L, const float *x, int T)
{
    WinSrc W{};
    W.src = x;
    W.row_stride = T;
    W.len = T;
    W.base0 = -(long long)L.N - L.T;
    W.base_step = L.N;
    W.nb = (T + L.N - 1) / L.N;
    W.ulo = 0;
    W.uhi = 2 * L.N;
    W.scale = 1.f;
    return W;
}

// The level's inverse-transform target for the forward: level 0's blocks tile
// the clip, so they write every sample once and the later levels add to it
// (no clearing pass).
static InvDst conv_ydst(const ConvLevel &L, float *y, int T, int M, bool first)
{
    InvDst D{};
    D.dst = y;
    D.row_stride = T;
    D.len = T;
    D.base0 = 0;
    D.base_step = L.N;
    D.nb = M;
    D.ulo = 0;
    D.uhi = L.N;
    D.scale = 1.f;
    D.accumulate = first ? 0 : 1;
    D.parity = -1;
    return D;
}

// Create every level's stream and events (once per handle; stream creation
// costs milliseconds, so the first eager call with call graphs on does it
// rather than the capture that follows).
static int conv_level_streams_create(pc_conv *c)
{
    if (!c->fork_ev) HIPCHK(hipEventCreateWithFlags(&c->fork_ev, hipEventDisableTiming));
    for (ConvLevelWork &w : c->lw) {
        if (!w.s) HIPCHK(hipStreamCreateWithFlags(&w.s, hipStreamNonBlocking));
        if (!w.ev) HIPCHK(hipEventCreateWithFlags(&w.ev, hipEventDisableTiming));
    }
    return PC_OK;
}

// The stream each level of a call runs on: with level streams on (or, by
// default, while the caller's stream is being captured into a graph) and
// more than one level, the level's own stream, made to wait for the work
// queued on the caller's stream s so far; else s.  Eager calls default to one
// stream: the event records and waits cost more host time than the overlap
// gains, while a replayed graph runs the level branches side by side.
// Returns whether levels fork.
static int conv_fork(pc_conv *c, hipStream_t s, std::vector<hipStream_t> &st, bool *par)
{
    const size_t G = c->lv.size();
    st.assign(G, s);
    bool on = c->level_streams == PCC_LEVEL_STREAMS_ON;
    if (c->level_streams == PCC_LEVEL_STREAMS_CAPTURED) {
        hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
        HIPCHK(hipStreamIsCapturing(s, &cs));
        on = cs == hipStreamCaptureStatusActive;
    }
    *par = on && G > 1;
    if (!*par) return PC_OK;
    int rc = conv_level_streams_create(c);
    if (rc) return rc;
    HIPCHK(hipEventRecord(c->fork_ev, s));
    for (size_t li = 0; li < G; li++) {
        ConvLevelWork &w = c->lw[li];
        HIPCHK(hipStreamWaitEvent(w.s, c->fork_ev, 0));
        st[li] = w.s;
    }
    return PC_OK;
}

// four-step staging of one level's transforms (null below the four-step sizes)
static int conv_level_scr(const ConvLevel &L, ConvLevelWork &w, size_t n, cf **scr)
{
    *scr = nullptr;
    if (!L.big) return PC_OK;
    int rc = conv_grow(&w.scr, &w.scrc, n);
    if (rc) return rc;
    *scr = w.scr;
    return PC_OK;
}

// Everything a captured forward (bwd false) or backward call embeds besides
// its arguments: the buffers, chunk sizes and mode its launches use.
static std::vector<uintptr_t> conv_fingerprint(const pc_conv *c, bool bwd)
{
    std::vector<uintptr_t> fp;
    fp.push_back((uintptr_t)c->level_streams);
    fp.push_back((uintptr_t)c->tw);
    for (size_t li = 0; li < c->lv.size(); li++) {
        const ConvLevel &L = c->lv[li];
        const ConvLevelWork &w = c->lw[li];
        for (const void *q : {(const void *)L.H, (const void *)c->keep[li], (const void *)w.Xr, (const void *)w.scr})
            fp.push_back((uintptr_t)q);
        fp.push_back((uintptr_t)L.mac_mc);
        if (bwd) {
            for (const void *q : {(const void *)w.G, (const void *)w.W, (const void *)w.part, (const void *)w.D})
                fp.push_back((uintptr_t)q);
            fp.push_back((uintptr_t)L.dh_mc);
        } else {
            fp.push_back((uintptr_t)w.Y);
        }
    }
    return fp;
}

static bool call_graph_match(const CallGraph &g, const void *const key[PC_CG_KEYS], int C, int T, int flags,
                             const std::vector<uintptr_t> &fp)
{
    if (g.C != C || g.T != T || g.flags != flags) return false;
    for (int i = 0; i < PC_CG_KEYS; i++)
        if (g.key[i] != key[i]) return false;
    return g.fp == fp;
}

static void call_graph_set(CallGraph &g, const void *const key[PC_CG_KEYS], int C, int T, int flags,
                           std::vector<uintptr_t> fp)
{
    for (int i = 0; i < PC_CG_KEYS; i++) g.key[i] = key[i];
    g.C = C;
    g.T = T;
    g.flags = flags;
    g.fp = std::move(fp);
}

static void call_graphs_drop(CallGraphs &G)
{
    for (CallGraph &g : G.g)
        if (g.exec) hipGraphExecDestroy(g.exec);
    G.g.clear();
    G.seen.clear();
}

// Capture launch(stream) on the handle's private stream into an executable
// graph (null on any failure, with the error state left as it was).  The
// launches find every buffer already sized by the eager call with the same
// arguments, so capture allocates nothing.
template <typename F>
static hipGraphExec_t call_graph_capture(pc_conv *c, F launch)
{
    if (!c->cap_s && hipStreamCreateWithFlags(&c->cap_s, hipStreamNonBlocking) != hipSuccess) {
        (void)hipGetLastError();
        c->cap_s = nullptr;
        return nullptr;
    }
    const std::string err = g_err;
    hipGraphExec_t exec = nullptr;
    if (hipStreamBeginCapture(c->cap_s, hipStreamCaptureModeThreadLocal) == hipSuccess) {
        const int rc = launch(c->cap_s);
        hipGraph_t graph = nullptr;
        const hipError_t e = hipStreamEndCapture(c->cap_s, &graph);
        if (rc != PC_OK || e != hipSuccess || !graph ||
            hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) != hipSuccess)
            exec = nullptr;
        if (graph) hipGraphDestroy(graph);
    }
    (void)hipGetLastError();
    g_err = err;
    return exec;
}

// Run one forward or backward call's launches on s: replay its graph if one
// was captured for these arguments, capture one if one of the last eager
// calls had them, else launch eagerly.  Bypassed (always eager) when call graphs
// are off, a kernel probe is set, or s is itself being captured.
template <typename F>
static int call_graph_run(pc_conv *c, CallGraphs &G, hipStream_t s, const void *const key[PC_CG_KEYS], int C, int T,
                          int flags, F launch)
{
    if (!c->call_graphs || c->probe_level >= 0) return launch(s);
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    HIPCHK(hipStreamIsCapturing(s, &cs));
    if (cs != hipStreamCaptureStatusNone) return launch(s);
    std::vector<uintptr_t> fp = conv_fingerprint(c, G.bwd);
    for (size_t i = 0; i < G.g.size(); i++)
        if (call_graph_match(G.g[i], key, C, T, flags, fp)) {
            HIPCHK(hipGraphLaunch(G.g[i].exec, s));
            std::rotate(G.g.begin(), G.g.begin() + i, G.g.begin() + i + 1);
            G.replays++;
            return PC_OK;
        }
    size_t si = 0;
    while (si < G.seen.size() && !call_graph_match(G.seen[si], key, C, T, flags, fp)) si++;
    if (si < G.seen.size()) {
        G.seen.erase(G.seen.begin() + si);
        hipGraphExec_t exec = call_graph_capture(c, launch);
        if (exec) {
            HIPCHK(hipGraphLaunch(exec, s));
            CallGraph g;
            g.exec = exec;
            call_graph_set(g, key, C, T, flags, std::move(fp));
            G.g.insert(G.g.begin(), std::move(g));
            if (G.g.size() > PC_CALL_GRAPHS) {
                hipGraphExecDestroy(G.g.back().exec);
                G.g.pop_back();
            }
            G.captures++;
            return PC_OK;
        }
    }
    int rc = launch(s);
    if (rc) return rc;
    if (c->level_streams != PCC_LEVEL_STREAMS_OFF && c->lv.size() > 1 && (rc = conv_level_streams_create(c)))
        return rc;
    CallGraph e;
    call_graph_set(e, key, C, T, flags, conv_fingerprint(c, G.bwd));
    G.seen.insert(G.seen.begin(), std::move(e));
    if (G.seen.size() > PC_CALL_GRAPHS) G.seen.pop_back();
    return PC_OK;
}

// The forward's launches on stream s (levels forked from s when level streams
// apply, joined back at the end); sizes the work space on first use.
static int conv_forward_launch(pc_conv *c, const float *d_x, float *d_y, int C, int T, hipStream_t s, bool keep)
{
    int rc = PC_OK;
    std::vector<hipStream_t> st;
    bool par;
    if ((rc = conv_fork(c, s, st, &par))) return rc;
    for (size_t li = 0; li < c->lv.size(); li++) {
        ConvLevel &L = c->lv[li];
        ConvLevelWork &w = c->lw[li];
        const hipStream_t ls = st[li];
        WinSrc W = conv_xwin(L, d_x, T);
        const int M = W.nb;
        const size_t n = (size_t)C * M * L.N;
        cf *X, *scr;
        if (keep) {
            if ((rc = conv_grow(&c->keep[li], &c->keep_cap[li], n))) return rc;
            X = c->keep[li];
        } else {
            if ((rc = conv_grow(&w.Xr, &w.Xrc, n))) return rc;
            X = w.Xr;
        }
        if ((rc = conv_grow(&w.Y, &w.Yc, n))) return rc;
        if ((rc = conv_level_scr(L, w, n, &scr))) return rc;
        // probed launches carry their algorithmic bytes: every signal sample
        // and every spectrum read or written once
        const double sig = 4.0 * (double)C * T, spec = 8.0 * (double)n;
        hipEvent_t e0, e1;
        if ((rc = probe_begin(c, li, PCC_PROBE_FORWARD_TRANSFORM, &e0, &e1))) return rc;
        if ((rc = off_rfft(c, L, W, C, X, ls, scr, e0, e1))) return rc;
        probe_end(c, e0, e1, sig + spec);
        if ((rc = probe_begin(c, li, PCC_PROBE_PRODUCT, &e0, &e1))) return rc;
        off_mac(X, L.H, w.Y, C, M, L, 1, ls, e0, e1);
        probe_end(c, e0, e1, 2.0 * spec + 8.0 * (double)L.P * L.N);
        HIPCHK(hipGetLastError());
        // level 0 writes y, later levels add to it: in level order
        if (par && li > 0) HIPCHK(hipStreamWaitEvent(ls, c->lw[li - 1].ev, 0));
        if ((rc = probe_begin(c, li, PCC_PROBE_INVERSE_TRANSFORM, &e0, &e1))) return rc;
        if ((rc = off_irfft(c, L, w.Y, conv_ydst(L, d_y, T, M, li == 0), C, ls, scr, e0, e1))) return rc;
        probe_end(c, e0, e1, spec + (li == 0 ? sig : 2.0 * sig));
        if (par) HIPCHK(hipEventRecord(w.ev, ls));
    }
    if (par) HIPCHK(hipStreamWaitEvent(s, c->lw[c->lv.size() - 1].ev, 0));
    return PC_OK;
}

// y = sum over levels of the partitioned convolution of every clip (x, y: [C][T])
PC_API int pcc_forward(pc_conv *c, const float *d_x, float *d_y, int C, int T, void *stream)
{
    if (!c || !d_x || !d_y || C < 1 || T < 1) return fail(PC_ERR_ARG, "bad argument");
    if (!c->filter_ready) return fail(PC_ERR_STATE, "forward before set_filter");
    hipStream_t s = (hipStream_t)stream;
    size_t need_keep = 0;
    for (ConvLevel &L : c->lv) need_keep += (size_t)C * ((T + L.N - 1) / L.N) * L.N * sizeof(cf);
    const bool keep = (long long)need_keep <= c->budget;
    c->keep_valid = false;
    const void *key[PC_CG_KEYS] = {d_x, d_y};
    const int rc = call_graph_run(c, c->fg, s, key, C, T, keep ? 1 : 0, [&](hipStream_t ls) {
        return conv_forward_launch(c, d_x, d_y, C, T, ls, keep);
    });
    if (rc) return rc;
    if (keep) {
        c->keep_valid = true;
        c->keep_x = d_x;
        c->keep_C = C;
        c->keep_T = T;
    }
    return PC_OK;
}

// ---------------------------------------------------------------------------
// pcc_forward without allocation: the same kernels in the same order, with
// the block spectra, the products and the four-step staging of each level in
// one file-scope device array instead of the handle's grown work space.  The
// array's address on the current device comes from hipGetSymbolAddress (a
// host-side use of its name would be a host address), taken once per device
// and kept.  Levels run one after another on the caller's stream, so each
// level reuses the array from its start: X (C M N), products (C M N), staging
// (C M N), float2 each.  Every word a level reads was written earlier in the
// same call, so nothing needs clearing.  All calls on one device share the
// array: calls that overlap in time on one device race.
// ---------------------------------------------------------------------------
#define PC_NOALLOC_BYTES (64ull << 20)
static __device__ __align__(256) cf scratch[PC_NOALLOC_BYTES / sizeof(cf)];
static std::mutex g_scratch_mu;
static cf *g_scratch_ptr[64];

static int scratch_ptr(cf **out)
{
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64) return fail(PC_ERR_ARG, "device index out of range");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(PC_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_ptr[dev] = static_cast<cf *>(p);
    }
    *out = g_scratch_ptr[dev];
    return PC_OK;
}

PC_API int pcc_forward_noalloc(pc_conv *c, const float *d_x, float *d_y, int C, int T, void *stream)
{
    if (!c || !d_x || !d_y || C < 1 || T < 1) return fail(PC_ERR_ARG, "bad argument");
    if (!c->filter_ready) return fail(PC_ERR_STATE, "forward before set_filter");
    for (const ConvLevel &L : c->lv) {
        const unsigned long long need = 3ull * (unsigned long long)C * ((T + L.N - 1) / L.N) * L.N * sizeof(cf);
        if (need > PC_NOALLOC_BYTES) {
            char m[200];
            snprintf(m, sizeof m,
                     "pcc_forward_noalloc: %d clips x %d samples need %llu bytes of scratch at block size %d; "
                     "the maximum is %llu (split the clips)",
                     C, T, need, L.N, PC_NOALLOC_BYTES);
            return fail(PC_ERR_ARG, m);
        }
    }
    cf *base = nullptr;
    int rc = scratch_ptr(&base);
    if (rc) return rc;
    hipStream_t s = (hipStream_t)stream;
    c->keep_valid = false;  // no spectra kept: a backward recomputes them from x
    for (size_t li = 0; li < c->lv.size(); li++) {
        ConvLevel &L = c->lv[li];
        WinSrc W = conv_xwin(L, d_x, T);
        const int M = W.nb;
        const size_t n = (size_t)C * M * L.N;
        cf *X = base, *Y = base + n, *stage = base + 2 * n;
        if ((rc = off_rfft(c, L, W, C, X, s, stage))) return rc;
        off_mac(X, L.H, Y, C, M, L, 1, s, nullptr, nullptr);
        HIPCHK(hipGetLastError());
        if ((rc = off_irfft(c, L, Y, conv_ydst(L, d_y, T, M, li == 0), C, s, stage))) return rc;
    }
    return PC_OK;
}

// The adjoint's launches on stream s (see conv_forward_launch); kept: the
// forward's spectra of x are valid, else they are recomputed.
static int conv_backward_launch(pc_conv *c, const float *d_x, const float *d_g, const float *d_scale, float *d_dx,
                                float *d_dh, int C, int T, hipStream_t s, bool kept)
{
    int rc = PC_OK;
    std::vector<hipStream_t> st;
    bool par;
    if ((rc = conv_fork(c, s, st, &par))) return rc;
    for (size_t li = 0; li < c->lv.size(); li++) {
        ConvLevel &L = c->lv[li];
        ConvLevelWork &w = c->lw[li];
        const hipStream_t ls = st[li];
        const int M = (T + L.N - 1) / L.N;
        const size_t n = (size_t)C * M * L.N;
        cf *scr;
        if ((rc = conv_level_scr(L, w, n, &scr))) return rc;
        // G = r2c([g block | 0_N])
        if ((rc = conv_grow(&w.G, &w.Gc, n))) return rc;
        WinSrc Wg{};
        Wg.src = d_g;
        Wg.row_stride = T;
        Wg.len = T;
        Wg.base0 = 0;
        Wg.base_step = L.N;
        Wg.nb = M;
        Wg.ulo = 0;
        Wg.uhi = L.N;
        Wg.scale = 1.f;
        Wg.dscale = d_scale;  // the upstream gradient's device-side factor, if any
        if ((rc = off_rfft(c, L, Wg, C, w.G, ls, scr))) return rc;
        if (d_dh) {
            cf *X;
            if (kept) X = c->keep[li];
            else {
                // recompute the delay-line spectra from the stored time-domain input
                if ((rc = conv_grow(&w.Xr, &w.Xrc, n))) return rc;
                X = w.Xr;
                if ((rc = off_rfft(c, L, conv_xwin(L, d_x, T), C, X, ls, scr))) return rc;
            }
            if ((rc = conv_grow(&w.D, &w.Dc, (size_t)L.P * L.N))) return rc;
            hipEvent_t e0, e1;
            if ((rc = probe_begin(c, li, PCC_PROBE_DH_PARTIAL, &e0, &e1))) return rc;
            if ((rc = off_dh(c, L, w.G, X, w.D, C, M, ls, e0, e1, &w.part, &w.partc))) return rc;
            // G and X read once, the partial sums written once
            const int nch = (M + L.dh_mc - 1) / L.dh_mc;
            probe_end(c, e0, e1, 16.0 * (double)n + 8.0 * (double)L.P * C * nch * L.N);
            InvDst D{};
            D.dst = d_dh;
            D.row_stride = 0;
            D.len = c->taps;
            D.base0 = (long long)L.T - L.N;
            D.base_step = L.N;
            D.nb = L.P;
            D.ulo = L.N;
            D.uhi = 2 * L.N;
            D.scale = 1.0f / (float)(2 * L.N);
            D.accumulate = 0;
            D.parity = -1;
            if ((rc = off_irfft(c, L, w.D, D, 1, ls, scr))) return rc;
        }
        if (d_dx) {
            if ((rc = conv_grow(&w.W, &w.Wc, n))) return rc;
            off_mac(w.G, L.H, w.W, C, M, L, -1, ls, nullptr, nullptr);
            HIPCHK(hipGetLastError());
            InvDst D{};
            D.dst = d_dx;
            D.row_stride = T;
            D.len = T;
            D.base0 = -(long long)L.N - L.T;
            D.base_step = L.N;
            D.nb = M;
            D.ulo = 0;
            D.uhi = 2 * L.N;
            D.scale = 1.f;
            D.accumulate = 1;
            // every level adds to dL/dx: in level order
            if (par && li > 0) HIPCHK(hipStreamWaitEvent(ls, c->lw[li - 1].ev, 0));
            for (int par2 = 0; par2 < 2; par2++) {
                D.parity = par2;
                if ((rc = off_irfft(c, L, w.W, D, C, ls, scr))) return rc;
            }
        }
        if (par) HIPCHK(hipEventRecord(w.ev, ls));
    }
    if (par)
        for (size_t li = 0; li < c->lv.size(); li++) HIPCHK(hipStreamWaitEvent(s, c->lw[li].ev, 0));
    return PC_OK;
}

// Adjoint of pcc_forward for upstream gradient g (times *d_scale when set):
// dx (may be null) and dh (may be null).  Uses the forward's kept spectra when
// they are for the same x and within the budget, else recomputes them from x.
static int conv_backward(pc_conv *c, const float *d_x, const float *d_g, const float *d_scale, float *d_dx,
                         float *d_dh, int C, int T, void *stream)
{
    if (!c || !d_x || !d_g || C < 1 || T < 1) return fail(PC_ERR_ARG, "bad argument");
    if (!c->filter_ready) return fail(PC_ERR_STATE, "backward before set_filter");
    hipStream_t s = (hipStream_t)stream;
    int rc;
    if (d_dx && (rc = conv_fill_zero(d_dx, (long long)C * T, s))) return rc;
    const bool kept = c->keep_valid && c->keep_x == d_x && c->keep_C == C && c->keep_T == T;
    const void *key[PC_CG_KEYS] = {d_x, d_g, d_dx, d_dh, d_scale};
    if ((rc = call_graph_run(c, c->bg, s, key, C, T, kept ? 1 : 0, [&](hipStream_t ls) {
             return conv_backward_launch(c, d_x, d_g, d_scale, d_dx, d_dh, C, T, ls, kept);
         })))
        return rc;
    if (d_dh && !kept) c->recomputed_levels += (long long)c->lv.size();
    return PC_OK;
}

PC_API int pcc_backward(pc_conv *c, const float *d_x, const float *d_g, float *d_dx, float *d_dh, int C, int T,
                        void *stream)
{
    return conv_backward(c, d_x, d_g, nullptr, d_dx, d_dh, C, T, stream);
}

PC_API int pcc_backward_scaled(pc_conv *c, const float *d_x, const float *d_g, const float *d_scale, float *d_dx,
                               float *d_dh, int C, int T, void *stream)
{
    if (!d_scale) return fail(PC_ERR_ARG, "bad argument");
    return conv_backward(c, d_x, d_g, d_scale, d_dx, d_dh, C, T, stream);
}

// Mean squared error of y against t with its gradient (k_mse_grad, then
// k_mse_final): one read of y and t, one write of g, a fixed-order sum.
PC_API int pcc_mse(pc_conv *c, const float *d_y, const float *d_t, float *d_g, float *d_loss, int C, int T,
                   void *stream)
{
    if (!c || !d_y || !d_t || !d_g || !d_loss || C < 1 || T < 1) return fail(PC_ERR_ARG, "bad argument");
    if (!c->mse_part) {
        HIPCHK(hipMalloc(&c->mse_part, PC_MSE_BLOCKS * sizeof(float)));
    }
    hipStream_t s = (hipStream_t)stream;
    const long long n = (long long)C * T;
    // contiguous chunks, a multiple of 4 samples (float4 groups) when y, t and g are 16-byte aligned
    const bool v4 = ((reinterpret_cast<uintptr_t>(d_y) | reinterpret_cast<uintptr_t>(d_t) |
                      reinterpret_cast<uintptr_t>(d_g)) & 15) == 0;
    long long chunk = (n + PC_MSE_BLOCKS - 1) / PC_MSE_BLOCKS;
    chunk = v4 ? (chunk + 3) & ~3ll : chunk;
    const int nb = (int)((n + chunk - 1) / chunk);
    hipLaunchKernelGGL(k_mse_grad, dim3(nb), dim3(PC_MSE_NT), 0, s, d_y, d_t, d_g, (float)(2.0 / (double)n), n, chunk,
                       v4 ? 1 : 0, c->mse_part);
    hipLaunchKernelGGL(k_mse_final, dim3(1), dim3(PC_MSE_NT), 0, s, c->mse_part, nb, (float)(1.0 / (double)n), d_loss);
    HIPCHK(hipGetLastError());
    return PC_OK;
}

PC_API int pcc_set_call_graphs(pc_conv *c, int on)
{
    if (!c) return fail(PC_ERR_ARG, "null handle");
    c->call_graphs = on != 0;
    if (!c->call_graphs) {
        call_graphs_drop(c->fg);
        call_graphs_drop(c->bg);
    }
    return PC_OK;
}

// Replays and captures of the forward's and the backward's call graphs.
PC_API int pcc_call_graph_stats(pc_conv *c, long long *fwd_replays, long long *fwd_captures, long long *bwd_replays,
                                long long *bwd_captures)
{
    if (!c) return fail(PC_ERR_ARG, "null handle");
    if (fwd_replays) *fwd_replays = c->fg.replays;
    if (fwd_captures) *fwd_captures = c->fg.captures;
    if (bwd_replays) *bwd_replays = c->bg.replays;
    if (bwd_captures) *bwd_captures = c->bg.captures;
    return PC_OK;
}

PC_API int pcc_set_level_streams(pc_conv *c, int mode)
{
    if (!c) return fail(PC_ERR_ARG, "null handle");
    if (mode < PCC_LEVEL_STREAMS_OFF || mode > PCC_LEVEL_STREAMS_CAPTURED)
        return fail(PC_ERR_ARG, "level stream mode must be 0 (off), 1 (on) or 2 (under graph capture)");
    c->level_streams = mode;
    return PC_OK;
}

// Choose each level's chunk sizes by timing (HIP events) the spectral product
// and the dL/dh partial sums on the caller's clips: 3 launches per candidate,
// the fastest mean kept.  The products' results do not depend on the chunk
// size; dL/dh's partial sums are grouped by it (a fixed order for a fixed
// choice).  Setup-time: allocates, synchronises.
PC_API int pcc_tune(pc_conv *c, const float *d_x, int C, int T, void *stream)
{
    if (!c || !d_x || C < 1 || T < 1) return fail(PC_ERR_ARG, "bad argument");
    if (!c->filter_ready) return fail(PC_ERR_STATE, "tune before set_filter");
    hipStream_t s = (hipStream_t)stream;
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    int rc = PC_OK;
    const int mac_c[] = {8, 16, 32, 64, 128}, dh_c[] = {16, 32, 64, 128, 256};
    for (size_t li = 0; li < c->lv.size() && !rc; li++) {
        ConvLevel &L = c->lv[li];
        WinSrc W = conv_xwin(L, d_x, T);
        const int M = W.nb;
        const size_t n = (size_t)C * M * L.N;
        if ((rc = conv_grow(&c->dsum, &c->dsum_cap, n))) break;
        // products (n) or the level's dL/dh spectra (P N)
        if ((rc = conv_grow(&c->work, &c->work_cap, std::max(n, (size_t)L.P * L.N)))) break;
        if ((rc = off_rfft(c, L, W, C, c->dsum, s))) break;
        // 3 timed launches per candidate after a warm-up launch
        auto time_it = [&](auto launch, double *ms_out) -> int {
            float tot = 0.f;
            for (int r = 0; r < 4; r++) {
                HIPCHK(hipEventRecord(e0, s));
                int lrc = launch();
                if (lrc) return lrc;
                HIPCHK(hipEventRecord(e1, s));
                HIPCHK(hipEventSynchronize(e1));
                float ms = 0.f;
                HIPCHK(hipEventElapsedTime(&ms, e0, e1));
                if (r > 0) tot += ms;
            }
            *ms_out = tot;
            return PC_OK;
        };
        // Deterministic choice: candidates within PC_TUNE_TIE of the fastest
        // count as equal; among them the default wins, else the first in the
        // candidate order.  Timing noise then moves the choice (and dL/dh's
        // grouping) only when a candidate is clearly faster.
        auto pick = [&](const int *cand, const double *ms, int n, int def) {
            double best = 1e30;
            for (int q = 0; q < n; q++) best = std::min(best, ms[q]);
            for (int q = 0; q < n; q++)
                if (cand[q] == def && ms[q] <= best * (1.0 + PC_TUNE_TIE)) return def;
            for (int q = 0; q < n; q++)
                if (ms[q] <= best * (1.0 + PC_TUNE_TIE)) return cand[q];
            return def;
        };
        if (L.P <= 16) {
            double ms[5];
            const int def = 32;
            for (int q = 0; q < 5 && !rc; q++) {
                L.mac_mc = mac_c[q];
                rc = time_it([&]() { off_mac(c->dsum, L.H, c->work, C, M, L, 1, s, nullptr, nullptr);
                                     return PC_OK; }, &ms[q]);
            }
            L.mac_mc = rc ? def : pick(mac_c, ms, 5, def);
            if (rc) break;
        }
        {
            double ms[5];
            const int def = 64;
            // the block spectra stand in for the gradient spectra (same shape)
            for (int q = 0; q < 5 && !rc; q++) {
                L.dh_mc = dh_c[q];
                rc = time_it([&]() { return off_dh(c, L, c->dsum, c->dsum, c->work, C, M, s); }, &ms[q]);
            }
            L.dh_mc = rc ? def : pick(dh_c, ms, 5, def);
        }
    }
    c->keep_valid = false;  // the work space held tuning data
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    return rc;
}

PC_API int pcc_set_level_tuning(pc_conv *c, int level, int mac_mc, int dh_mc)
{
    if (!c || level < 0 || level >= (int)c->lv.size()) return fail(PC_ERR_ARG, "bad argument");
    auto ok = [](int v) { return v >= 1 && v <= 1024 && (v & (v - 1)) == 0; };
    if (!ok(mac_mc) || !ok(dh_mc)) return fail(PC_ERR_ARG, "chunk sizes must be powers of two in [1, 1024]");
    c->lv[level].mac_mc = mac_mc;
    c->lv[level].dh_mc = dh_mc;
    return PC_OK;
}

PC_API int pcc_level_tuning(const pc_conv *c, int level, int *mac_mc, int *dh_mc)
{
    if (!c || level < 0 || level >= (int)c->lv.size()) return fail(PC_ERR_ARG, "bad argument");
    if (mac_mc) *mac_mc = c->lv[level].mac_mc;
    if (dh_mc) *dh_mc = c->lv[level].dh_mc;
    return PC_OK;
}

PC_API int pcc_probe(pc_conv *c, int level) { return pcc_probe_kernel(c, level, PCC_PROBE_PRODUCT); }

PC_API int pcc_probe_kernel(pc_conv *c, int level, int kind)
{
    if (!c) return fail(PC_ERR_ARG, "null handle");
    if (level < -1 || level >= (int)c->lv.size()) return fail(PC_ERR_ARG, "probe level out of range");
    if (kind < PCC_PROBE_PRODUCT || kind > PCC_PROBE_DH_PARTIAL) return fail(PC_ERR_ARG, "unknown probe kind");
    int rc = conv_probe_fold(c);
    if (rc) return rc;
    c->probe_level = level;
    c->probe_kind = kind;
    c->probe_ms = 0;
    c->probe_n = 0;
    c->probe_bytes = 0;
    return PC_OK;
}

PC_API int pcc_probe_read(pc_conv *c, double *avg_us, long long *launches, double *bytes_per_launch)
{
    if (!c) return fail(PC_ERR_ARG, "null handle");
    int rc = conv_probe_fold(c);
    if (rc) return rc;
    if (avg_us) *avg_us = c->probe_n ? 1e3 * c->probe_ms / (double)c->probe_n : 0.0;
    if (launches) *launches = c->probe_n;
    if (bytes_per_launch) *bytes_per_launch = c->probe_bytes;
    return PC_OK;
}

static void conv_destroy(pc_conv *c)
{
    hipDeviceSynchronize();
    conv_probe_fold(c);
    for (ConvLevel &L : c->lv) hipFree(L.H);
    for (cf *p : c->keep) hipFree(p);
    for (ConvLevelWork &w : c->lw) {
        for (cf *p : {w.Y, w.G, w.W, w.Xr, w.part, w.D, w.scr}) hipFree(p);
        if (w.ev) hipEventDestroy(w.ev);
        if (w.s) hipStreamDestroy(w.s);
    }
    if (c->fork_ev) hipEventDestroy(c->fork_ev);
    call_graphs_drop(c->fg);
    call_graphs_drop(c->bg);
    if (c->cap_s) hipStreamDestroy(c->cap_s);
    hipFree(c->work);
    hipFree(c->scr);
    hipFree(c->dpart);
    hipFree(c->dsum);
    hipFree(c->mse_part);
    delete c;
}

PC_API void pcc_destroy(pc_conv *c)
{
    if (!c || !live_take(c)) return;
    conv_destroy(c);
}

// ---------------------------------------------------------------------------
// MIMO throughput form (pcm-free, offline): y[o] = sum_i conv(x[i], h[o][i]).
// ---------------------------------------------------------------------------
struct pc_mimo {
    pc_conv *c = nullptr;  // partition, twiddles, transforms and work space
    int n_in = 0, n_out = 0;
    std::vector<cf *> Ht;  // per level: [N][P][n_in][32]
    cf *tmp = nullptr;
    size_t tmp_cap = 0;
};

PC_API pc_mimo *pcx_create(int M, const int *N, const int *P, int taps, int n_in, int n_out)
{
    if (n_in < 1 || n_in > 32 || n_out < 1 || n_out > 32) {
        fail(PC_ERR_ARG, "MIMO sizes must be 1..32");
        return nullptr;
    }
    pc_conv *c = conv_create(M, N, P, taps);
    if (!c) return nullptr;
    pc_mimo *x = new pc_mimo();
    x->c = c;
    x->n_in = n_in;
    x->n_out = n_out;
    for (ConvLevel &L : c->lv) {
        cf *h = nullptr;
        if (hipMalloc(&h, (size_t)L.N * L.P * n_in * 32 * sizeof(cf)) != hipSuccess) {
            fail(PC_ERR_HIP, "hipMalloc MIMO spectra");
            for (cf *q : x->Ht) hipFree(q);
            conv_destroy(c);
            delete x;
            return nullptr;
        }
        x->Ht.push_back(h);
    }
    if (!live_add(x, HK_MIMO)) {
        for (cf *q : x->Ht) hipFree(q);
        conv_destroy(c);
        delete x;
        fail(PC_ERR_STATE, "library shut down");
        return nullptr;
    }
    return x;
}

// d_h: [n_out][n_in][taps] on the device
PC_API int pcx_set_filters(pc_mimo *x, const float *d_h, void *stream)
{
    if (!x || !d_h) return fail(PC_ERR_ARG, "null argument");
    hipStream_t s = (hipStream_t)stream;
    pc_conv *c = x->c;
    const int R = x->n_in * x->n_out;
    for (size_t li = 0; li < c->lv.size(); li++) {
        ConvLevel &L = c->lv[li];
        int rc = conv_grow(&x->tmp, &x->tmp_cap, (size_t)R * L.P * L.N);
        if (rc) return rc;
        WinSrc W{};
        W.src = d_h;
        W.row_stride = c->taps;
        W.len = c->taps;
        W.base0 = (long long)L.T - L.N;
        W.base_step = L.N;
        W.nb = L.P;
        W.ulo = L.N;
        W.uhi = 2 * L.N;
        W.scale = 1.0f / (float)(2 * L.N);
        if ((rc = off_rfft(c, L, W, R, x->tmp, s))) return rc;
        const long long n = (long long)L.N * L.P * x->n_in * 32;
        hipLaunchKernelGGL(k_h_transpose, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, x->tmp, x->Ht[li],
                           x->n_in, x->n_out, L.P, L.N);
        HIPCHK(hipGetLastError());
    }
    return PC_OK;
}

// y[n_out][T] = sum_i conv(x[i], h[o][i]) (x: [n_in][T], device)
PC_API int pcx_forward(pc_mimo *x, const float *d_x, float *d_y, int T, void *stream)
{
    if (!x || !d_x || !d_y || T < 1) return fail(PC_ERR_ARG, "bad argument");
    hipStream_t s = (hipStream_t)stream;
    pc_conv *c = x->c;
    int rc = PC_OK;
    for (size_t li = 0; li < c->lv.size(); li++) {
        ConvLevel &L = c->lv[li];
        WinSrc W = conv_xwin(L, d_x, T);
        const int M = W.nb;
        const size_t nx = (size_t)x->n_in * M * L.N, ny = (size_t)x->n_out * M * L.N;
        if ((rc = conv_grow(&c->dsum, &c->dsum_cap, nx))) return rc;
        if ((rc = conv_grow(&x->tmp, &x->tmp_cap, nx))) return rc;
        if ((rc = conv_grow(&c->work, &c->work_cap, ny))) return rc;
        if ((rc = off_rfft(c, L, W, x->n_in, c->dsum, s))) return rc;
        const long long n = (long long)nx;
        hipLaunchKernelGGL(k_spec_transpose, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, c->dsum, x->tmp,
                           x->n_in, M, L.N);
        // up to 4 m-tiles per workgroup share the staged filter blocks and
        // input window (while that fits 64 KB of LDS)
        const int tiles = (M + 31) / 32;
        const int mt = tiles >= 4 ? 4 : 2;
        const size_t lds = (2 * 32 * 32 + (size_t)x->n_in * (32 * mt + L.P - 1)) * sizeof(cf);
        if (tiles >= 4 && lds <= 65536)
            hipLaunchKernelGGL(k_mimo_mfma_lds<4>, dim3(L.N, (tiles + 3) / 4), dim3(256), lds, s, x->Ht[li], x->tmp,
                               c->work, L.N, M, L.P, x->n_in, x->n_out);
        else if (tiles >= 2 && lds <= 65536)
            hipLaunchKernelGGL(k_mimo_mfma_lds<2>, dim3(L.N, (tiles + 1) / 2), dim3(128), lds, s, x->Ht[li], x->tmp,
                               c->work, L.N, M, L.P, x->n_in, x->n_out);
        else
            hipLaunchKernelGGL(k_mimo_mfma, dim3(L.N, tiles), dim3(64), 0, s, x->Ht[li], x->tmp, c->work, L.N, M,
                               L.P, x->n_in, x->n_out);
        HIPCHK(hipGetLastError());
        InvDst D{};
        D.dst = d_y;
        D.row_stride = T;
        D.len = T;
        D.base0 = 0;
        D.base_step = L.N;
        D.nb = M;
        D.ulo = 0;
        D.uhi = L.N;
        D.scale = 1.f;
        D.accumulate = li > 0;  // level 0's blocks tile the signal: no clearing pass
        D.parity = -1;
        if ((rc = off_irfft(c, L, c->work, D, x->n_out, s))) return rc;
    }
    return PC_OK;
}

static void mimo_destroy(pc_mimo *x)
{
    hipDeviceSynchronize();
    for (cf *q : x->Ht) hipFree(q);
    hipFree(x->tmp);
    conv_destroy(x->c);
    delete x;
}

PC_API void pcx_destroy(pc_mimo *x)
{
    if (!x || !live_take(x)) return;
    mimo_destroy(x);
}

// ---------------------------------------------------------------------------
// C ABI: real-time MIMO engine (no reference counterpart; the reference runs a
// MIMO network as n_in x n_out PartConvMulti instances).  Kernels and the
// algorithm: pc_mimo_rt.hpp.  Host loop as the instance engine's persistent
// runs: one cooperative launch serves the ticks of a call, each tick released
// by the host (release word through the BAR where available) and waited for
// on per-output completion flags; level blocks >= 1 are launched after the
// tick that starts them, on one stream per level, and each deadline tick waits
// for its level's event before it is released.
// ---------------------------------------------------------------------------
#define MXR_MAX_N 4096  // level blocks transform in one workgroup's LDS

struct pc_mimo_rt {
    int dev = 0, FS = 0, N0 = 0, n_in = 0, n_out = 0, taps = 0, M = 0;
    std::vector<int> N, P, T, delay;
    MxrArgs A{};
    std::vector<void *> allocs;
    cf *tw = nullptr;
    hipStream_t s_main = nullptr;
    std::vector<hipStream_t> s_lvl;
    std::vector<hipEvent_t> ev_lvl;
    std::vector<bool> pending;
    unsigned long long *h_flags = nullptr, *d_flags = nullptr;
    unsigned long long *go_h = nullptr, *go_d = nullptr, *h_go_pinned = nullptr;
    int go_slot = -1;
    unsigned long long *d_bar = nullptr;
    uint64_t frame = 0;
    int window = 1;  // pcxr_set_release_window: ticks released ahead of the oldest incomplete one
    std::vector<double> lat;
    unsigned long long ticks = 0, misses = 0;
    double lat_sum = 0, lat_max = 0;
    bool filters = false;
    // probe: HIP events around the delay-line-sum kernel of one level
    int probe_level = -1;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> pev;
    int pev_used = 0;
    // host-buffer ticks: pinned frames the kernel reads and writes in place
    float *h_in = nullptr, *h_out = nullptr, *d_in_map = nullptr, *d_out_map = nullptr;
    // level launches off the tick thread (as the instance engine's helper):
    // queued (level, start tick); done[l] = start tick + 1 once issued
    std::thread lth;
    std::mutex lmu;
    std::condition_variable lcv;
    std::vector<std::pair<int, uint64_t>> lq;
    bool lstop = false;
    int lbusy = 0;
    std::unique_ptr<std::atomic<uint64_t>[]> ldone;
    std::vector<uint64_t> lstart;
    int lerr = 0;
    std::string lmsg;
    // held level (pcxr_stall_level) and the persistent kernel's restarts
    // after a late level block
    std::atomic<int> stall_l{-1};
    std::atomic<long long> stall_us{0};
    unsigned long long stalls = 0;
    int ncu = 0;
    Residency resid;  // residency stamps of plain persistent launches
};

static int mxr_start_level(pc_mimo_rt *r, int l, uint64_t f);

static void mxr_launcher_main(pc_mimo_rt *r)
{
    (void)hipSetDevice(r->dev);
    for (;;) {
        std::pair<int, uint64_t> job;
        {
            std::unique_lock<std::mutex> lk(r->lmu);
            r->lcv.wait(lk, [&] { return r->lstop || !r->lq.empty(); });
            if (r->lq.empty()) return;  // stopped with nothing left
            job = r->lq.front();
            r->lq.erase(r->lq.begin());
        }
        const int rc = mxr_start_level(r, job.first, job.second);
        std::lock_guard<std::mutex> lk(r->lmu);
        if (rc && !r->lerr) {
            r->lerr = rc;
            r->lmsg = g_err;
        }
        r->ldone[job.first].store(job.second + 1, std::memory_order_release);
        r->lbusy--;
        r->lcv.notify_all();
    }
}

static void mxr_launcher_start(pc_mimo_rt *r)
{
    if (r->lth.joinable()) return;
    r->ldone.reset(new std::atomic<uint64_t>[r->M]);
    for (int l = 0; l < r->M; l++) r->ldone[l].store(0);
    r->lstart.assign(r->M, 0);
    r->lstop = false;
    r->lth = std::thread(mxr_launcher_main, r);
}

static void mxr_launcher_queue(pc_mimo_rt *r, int l, uint64_t f)
{
    {
        std::lock_guard<std::mutex> lk(r->lmu);
        r->lq.push_back({l, f});
        r->lbusy++;
    }
    r->lcv.notify_all();
}

// Wait until every queued launch is issued; returns its first error.
static int mxr_launcher_drain(pc_mimo_rt *r)
{
    if (!r->lth.joinable()) return PC_OK;
    std::unique_lock<std::mutex> lk(r->lmu);
    r->lcv.wait(lk, [&] { return r->lbusy == 0; });
    if (r->lerr) {
        const int rc = r->lerr;
        r->lerr = 0;
        return fail(rc, r->lmsg);
    }
    return PC_OK;
}

static void mxr_launcher_stop(pc_mimo_rt *r)
{
    if (!r->lth.joinable()) return;
    {
        std::lock_guard<std::mutex> lk(r->lmu);
        r->lstop = true;
    }
    r->lcv.notify_all();
    r->lth.join();
}

static void mxr_free(pc_mimo_rt *r)
{
    mxr_launcher_stop(r);
    for (auto &e : r->pev) {
        hipEventDestroy(e.first);
        hipEventDestroy(e.second);
    }
    r->pev.clear();
    if (r->s_main) hipStreamSynchronize(r->s_main);
    for (hipStream_t s : r->s_lvl)
        if (s) hipStreamSynchronize(s);
    for (void *p : r->allocs) hipFree(p);
    r->allocs.clear();
    for (hipStream_t s : r->s_lvl)
        if (s) hipStreamDestroy(s);  // level 0 has none (it runs in the persistent kernel)
    r->s_lvl.clear();
    for (hipEvent_t e : r->ev_lvl) hipEventDestroy(e);
    r->ev_lvl.clear();
    if (r->s_main) hipStreamDestroy(r->s_main);
    r->s_main = nullptr;
    if (r->h_flags) hipHostFree(r->h_flags);
    if (r->h_go_pinned) hipHostFree(r->h_go_pinned);
    r->h_flags = r->h_go_pinned = nullptr;
    if (r->h_in) hipHostFree(r->h_in);
    if (r->h_out) hipHostFree(r->h_out);
    r->h_in = r->h_out = r->d_in_map = r->d_out_map = nullptr;
    release_slot_put(r->dev, r->go_slot);
    r->go_slot = -1;
    resident_free(r->resid);
}

static int mxr_alloc(pc_mimo_rt *r, void **p, size_t bytes)
{
    HIPCHK(hipMalloc(p, bytes));
    r->allocs.push_back(*p);
    HIPCHK(hipMemset(*p, 0, bytes));
    return PC_OK;
}

static int mxr_setup(pc_mimo_rt *r)
{
    HIPCHK(hipGetDevice(&r->dev));
    hip_live();
    // level blocks of 4096 transform in 64 KiB of LDS (two float2 buffers)
    const int lds12 = 2 * 4096 * (int)sizeof(cf);
    HIPCHK(hipFuncSetAttribute((const void *)k_mxr_spec<12>, hipFuncAttributeMaxDynamicSharedMemorySize, lds12));
    HIPCHK(hipFuncSetAttribute((const void *)k_mxr_fwd<12>, hipFuncAttributeMaxDynamicSharedMemorySize, lds12));
    HIPCHK(hipFuncSetAttribute((const void *)k_mxr_inv<12>, hipFuncAttributeMaxDynamicSharedMemorySize, lds12));
    int rc = get_twiddles(r->dev, &r->tw);
    if (rc) return rc;
    const int N0 = r->N0, M = r->M;
    long long need = 0;
    for (int l = 0; l < M; l++)
        need = std::max(need, std::max(2ll * r->N[l], (long long)r->N[l] + N0 + r->T[l]) + N0);
    long long R = 1;
    while (R < need) R <<= 1;
    MxrArgs &A = r->A;
    A.n_in = r->n_in;
    A.n_out = r->n_out;
    A.N0 = N0;
    A.M = M;
    A.R = R;
    A.ring_mask = R - 1;
    if ((rc = mxr_alloc(r, (void **)&A.ring, (size_t)r->n_in * 2 * R * sizeof(float)))) return rc;
    for (int l = 0; l < M; l++) {
        MxrLevel &L = A.lv[l];
        L.N = r->N[l];
        L.P = r->P[l];
        L.kshift = ilog2(r->N[l] / N0);
        L.T = r->T[l];
        L.ni = (r->n_in + 3) & ~3;
        const size_t nb = (size_t)L.N * sizeof(cf);
        cf *H = nullptr;
        if ((rc = mxr_alloc(r, (void **)&L.fdl, (size_t)L.ni * L.P * nb))) return rc;
        if ((rc = mxr_alloc(r, (void **)&H, (size_t)L.P * r->n_out * L.ni * nb))) return rc;
        L.H = H;
        // level 0's S twice: the persistent kernel double-buffers it by tick parity
        if ((rc = mxr_alloc(r, (void **)&L.S, (size_t)(l == 0 ? MXR_SLOTS : 1) * r->n_out * nb))) return rc;
        if ((rc = mxr_alloc(r, (void **)&L.Y, (size_t)r->n_out * nb))) return rc;
        if (l > 0 && (rc = mxr_alloc(r, (void **)&L.ybuf, (size_t)r->n_out * 2 * L.N * sizeof(float)))) return rc;
    }
    {
        MxrLevel *d_lv = nullptr;
        if ((rc = mxr_alloc(r, (void **)&d_lv, (size_t)M * sizeof(MxrLevel)))) return rc;
        HIPCHK(hipMemcpy(d_lv, A.lv, (size_t)M * sizeof(MxrLevel), hipMemcpyHostToDevice));
        A.lvd = d_lv;
    }
    if ((rc = mxr_alloc(r, (void **)&r->d_bar, 640))) return rc;  // cX and 8 cS groups, own lines
    HIPCHK(hipStreamCreateWithFlags(&r->s_main, hipStreamNonBlocking));
    hipDeviceProp_t prop;
    HIPCHK(hipGetDeviceProperties(&prop, r->dev));
    r->ncu = prop.multiProcessorCount;
    std::vector<uint32_t> cu_mask((prop.multiProcessorCount + 31) / 32, 0u);
    for (int c = 0; c < prop.multiProcessorCount; c++) cu_mask[c / 32] |= 1u << (c % 32);
    r->ev_lvl.assign(M, nullptr);
    r->pending.assign(M, false);
    for (int l = 0; l < M; l++) {
        hipStream_t s = nullptr;
        // a queue of its own (full CU mask): level blocks never wait behind the persistent kernel
        if (l > 0) HIPCHK(hipExtStreamCreateWithCUMask(&s, (uint32_t)cu_mask.size(), cu_mask.data()));
        r->s_lvl.push_back(s);
        HIPCHK(hipEventCreateWithFlags(&r->ev_lvl[l], hipEventDisableTiming));
    }
    HIPCHK(hipHostMalloc(&r->h_flags, (size_t)r->n_out * 8 * sizeof(unsigned long long),
                         hipHostMallocCoherent | hipHostMallocMapped));
    memset(r->h_flags, 0, (size_t)r->n_out * 8 * sizeof(unsigned long long));
    HIPCHK(hipHostGetDevicePointer((void **)&r->d_flags, r->h_flags, 0));
    r->go_slot = release_slot_get(r->dev, &r->go_h);
    if (r->go_slot >= 0) {
        r->go_d = r->go_h;
    } else {
        HIPCHK(hipHostMalloc(&r->h_go_pinned, 64, hipHostMallocCoherent | hipHostMallocMapped));
        memset(r->h_go_pinned, 0, 64);
        r->go_h = r->h_go_pinned;
        HIPCHK(hipHostGetDevicePointer((void **)&r->go_d, r->h_go_pinned, 0));
    }
    A.go = r->go_d;
    A.flags = r->d_flags;
    A.bar = r->d_bar;
    A.spin_limit = PC_SPIN_LIMIT;
    r->lat.reserve(1 << 16);
    return PC_OK;
}

PC_API pc_mimo_rt *pcxr_create(int FS, int buffer_size, int M, const int *N, const int *P, int taps, int n_in,
                               int n_out)
{
    if (!N || !P || M < 1 || M > PC_MAX_LEVELS || taps < 1 || FS < 1) {
        fail(PC_ERR_ARG, "bad argument");
        return nullptr;
    }
    if (n_in < 1 || n_in > MXR_MAX_IO || n_out < 1 || n_out > MXR_MAX_IO) {
        fail(PC_ERR_ARG, "real-time MIMO sizes must be 1..32");
        return nullptr;
    }
    if (N[0] != buffer_size || buffer_size < 64 || buffer_size > 1024) {
        fail(PC_ERR_ARG, "real-time MIMO: N[0] must equal buffer_size, 64..1024");
        return nullptr;
    }
    pc_mimo_rt *r = new pc_mimo_rt();
    r->FS = FS;
    r->N0 = buffer_size;
    r->n_in = n_in;
    r->n_out = n_out;
    r->taps = taps;
    r->M = M;
    r->N.assign(N, N + M);
    r->P.assign(P, P + M);
    int rc = check_partition(r->N, r->P, taps, r->delay, r->T);
    for (int l = 0; !rc && l < M; l++)
        if (r->N[l] > MXR_MAX_N) rc = fail(PC_ERR_PARTITION, "real-time MIMO: block sizes up to 4096");
    if (!rc && lib_shut()) rc = fail(PC_ERR_STATE, "library shut down");
    if (!rc) rc = mxr_setup(r);
    if (!rc && !live_add(r, HK_MIMO_RT)) rc = fail(PC_ERR_STATE, "library shut down");
    if (rc) {
        mxr_free(r);
        delete r;
        return nullptr;
    }
    return r;
}

template <int LOGN>
static void mxr_spec_t(pc_mimo_rt *r, const MxrLevel &L, const float *d_h)
{
    hipLaunchKernelGGL(k_mxr_spec<LOGN>, dim3(L.P, r->n_in * r->n_out), dim3(MXR_NT), 2 * (1 << LOGN) * sizeof(cf),
                       r->s_main, L, d_h, r->taps, r->n_in, r->n_out, r->tw);
}

// d_h: [n_out][n_in][taps] on the device
PC_API int pcxr_set_filters(pc_mimo_rt *r, const float *d_h)
{
    if (!r || !d_h) return fail(PC_ERR_ARG, "null argument");
    for (int l = 0; l < r->M; l++) {
        const MxrLevel &L = r->A.lv[l];
        switch (ilog2(L.N)) {
        case 6: mxr_spec_t<6>(r, L, d_h); break;
        case 7: mxr_spec_t<7>(r, L, d_h); break;
        case 8: mxr_spec_t<8>(r, L, d_h); break;
        case 9: mxr_spec_t<9>(r, L, d_h); break;
        case 10: mxr_spec_t<10>(r, L, d_h); break;
        case 11: mxr_spec_t<11>(r, L, d_h); break;
        case 12: mxr_spec_t<12>(r, L, d_h); break;
        default: return fail(PC_ERR_ARG, "block size out of range");
        }
        HIPCHK(hipGetLastError());
    }
    HIPCHK(hipStreamSynchronize(r->s_main));
    r->filters = true;
    return PC_OK;
}

template <int LOGN>
static void mxr_level_t(pc_mimo_rt *r, int l, uint64_t m)
{
    const MxrLevel &L = r->A.lv[l];
    hipStream_t s = r->s_lvl[l];
    const size_t lds = 2 * (1 << LOGN) * sizeof(cf);
    hipLaunchKernelGGL(k_mxr_fwd<LOGN>, dim3(r->n_in), dim3(MXR_NT), lds, s, L, r->A.ring, r->A.R, r->A.ring_mask,
                       r->N0, m, r->tw);
    hipLaunchKernelGGL(k_mxr_comb, dim3(L.N), dim3(MXR_NT), 0, s, L, r->n_in, r->n_out, m);
    hipLaunchKernelGGL(k_mxr_inv<LOGN>, dim3(r->n_out), dim3(MXR_NT), lds, s, L, m, r->tw);
    if (l == r->probe_level && r->pev_used < (int)r->pev.size()) {
        const auto &e = r->pev[r->pev_used++];
        hipExtLaunchKernelGGL(k_mxr_part, dim3(L.N), dim3(MXR_NT), 0, s, e.first, e.second, 0, L, r->n_in,
                              r->n_out, (uint64_t)(m + 1));
    } else {
        hipLaunchKernelGGL(k_mxr_part, dim3(L.N), dim3(MXR_NT), 0, s, L, r->n_in, r->n_out, m + 1);
    }
}

// Level l's block that starts at tick f (after that tick's frame is in the ring).
static int mxr_start_level(pc_mimo_rt *r, int l, uint64_t f)
{
    const uint64_t m = f >> r->A.lv[l].kshift;
    int want = l;
    if (r->stall_l.load() == l && r->stall_l.compare_exchange_strong(want, -1))
        HIPCHK(hipLaunchHostFunc(r->s_lvl[l], host_sleep_fn, (void *)(uintptr_t)r->stall_us.load()));
    switch (ilog2(r->A.lv[l].N)) {
    case 7: mxr_level_t<7>(r, l, m); break;
    case 8: mxr_level_t<8>(r, l, m); break;
    case 9: mxr_level_t<9>(r, l, m); break;
    case 10: mxr_level_t<10>(r, l, m); break;
    case 11: mxr_level_t<11>(r, l, m); break;
    case 12: mxr_level_t<12>(r, l, m); break;
    default: return fail(PC_ERR_ARG, "block size out of range");
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipEventRecord(r->ev_lvl[l], r->s_lvl[l]));
    return PC_OK;
}

template <int LOGN>
static hipError_t mxr_launch_t(pc_mimo_rt *r, int G, bool *plain)
{
    MxrArgs a = r->A;
    const cf *t = r->tw;
    void *args[] = {(void *)&a, (void *)&t};
    // two FFT buffers and the twiddles, plus the output workgroups' filter
    // cache when it leaves room for two workgroups per CU
    const size_t fft = 5 * (size_t)(1 << LOGN) * sizeof(cf);
    const size_t hc = (size_t)(1 << LOGN) * a.lv[0].ni * sizeof(cf);
    a.hcache = (fft + hc <= 80 * 1024) ? 1 : 0;
    a.lpb = 2;
    while (a.lpb < a.lv[0].ni / 2) a.lpb *= 2;
    const size_t lds = fft + (a.hcache ? hc : 0);
    if (lds > 64 * 1024) {
        const hipError_t e = hipFuncSetAttribute((const void *)k_mxr_persist<LOGN>,
                                                 hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
    }
    return launch_persistent((const void *)k_mxr_persist<LOGN>, G, MXR_NTT, args, lds, r->s_main, r->resid, r->ncu,
                             &a.present, &a.present_tag, plain);
}

static void mxr_store_go(pc_mimo_rt *r, unsigned long long v)
{
    if (r->go_slot >= 0) {
        _mm_sfence();
        *(volatile unsigned long long *)r->go_h = v;
        _mm_sfence();
    } else {
        __atomic_store_n(r->go_h, v, __ATOMIC_RELEASE);
    }
}

// Launch the persistent level-0 kernel for nticks ticks from r->frame (frames
// of this call start at tick A.fbase).  The kernel waits for each release; its
// counters restart from zero.
static int mxr_persist_launch(pc_mimo_rt *r, long long nticks)
{
    MxrArgs &A = r->A;
    A.f0 = r->frame;
    A.nticks = nticks;
    // ticks in flight: at most MXR_SLOTS - 1 (the S slots) and at most P_0 (an
    // input overwrites the delay-line row of tick f - P_0, read by its outputs)
    A.window = std::max(1, std::min({r->window, MXR_SLOTS - 1, A.lv[0].P}));
    // the grid: input and output workgroups, then one delay-line-sum workgroup
    // per bin (at most 256)
    const int G = r->n_in + r->n_out + std::min(r->N0, 256);
    HIPCHK(hipMemsetAsync(r->d_bar, 0, 640, r->s_main));
    mxr_store_go(r, r->frame);
    hipError_t le = hipErrorInvalidValue;
    bool plain = false;
    switch (ilog2(r->N0)) {
    case 6: le = mxr_launch_t<6>(r, G, &plain); break;
    case 7: le = mxr_launch_t<7>(r, G, &plain); break;
    case 8: le = mxr_launch_t<8>(r, G, &plain); break;
    case 9: le = mxr_launch_t<9>(r, G, &plain); break;
    case 10: le = mxr_launch_t<10>(r, G, &plain); break;
    }
    if (le != hipSuccess) return fail(PC_ERR_HIP, std::string("real-time MIMO launch: ") + hipGetErrorString(le));
    if (plain && !resident_wait(r->resid, G)) {
        mxr_store_go(r, PC_GO_STOP);  // every wait of the kernel ends at the stop code
        HIPCHK(hipStreamSynchronize(r->s_main));
        return fail(PC_ERR_HIP, "real-time MIMO: the persistent kernel's workgroups could not all be resident");
    }
    return PC_OK;
}

// Wait until level l's pending block has been issued by the helper and has
// completed, for at most `limit` seconds from t0: 0 done, 1 still late, else
// an error.
static int mxr_level_wait(pc_mimo_rt *r, int l, double t0, double limit)
{
    while (r->ldone[l].load(std::memory_order_acquire) < r->lstart[l] + 1)
        if (now_s() - t0 > limit) return 1;
    for (;;) {
        const hipError_t q = hipEventQuery(r->ev_lvl[l]);
        if (q == hipSuccess) return PC_OK;
        if (q != hipErrorNotReady) return fail(PC_ERR_HIP, std::string("level block: ") + hipGetErrorString(q));
        if (now_s() - t0 > limit) return 1;
    }
}

// Release-window run of the real-time MIMO engine (pcxr_set_release_window):
// up to A.window ticks are released ahead of the oldest tick whose outputs are
// not all seen; the kernel's input workgroups wait for the delay-line sums that
// still read the row they overwrite, and S has a slot per tick in flight.  The
// deadline rule is that of the loop below: tick r is released only once the
// level blocks it reads first are complete, and a block is launched once its
// start tick is seen complete.  A block still late PC_LEVEL_LATE_S after the
// window emptied stops the kernel (every released tick is served by then), the
// host waits for it and relaunches, as below.  A tick's latency counts from its
// release, or from when the window emptied if it waited for a level block.
static int mxr_run_window(pc_mimo_rt *r, uint64_t nframes, double period)
{
    MxrArgs &A = r->A;
    const uint64_t W = (uint64_t)A.window, fend = r->frame + nframes;
    std::vector<double> trel((size_t)W, 0.0);
    uint64_t rel = r->frame;
    double t_idle = now_s(), t_wait = -1;
    int oscan = 0;  // outputs of the oldest tick seen done so far
    while (r->frame < fend) {
        while (rel < fend && rel < r->frame + W) {
            bool ready = true;
            for (int l = 1; l < r->M && ready; l++) {
                const uint64_t k = 1ull << A.lv[l].kshift;
                if ((rel + 1) % k != 0 || rel + 1 < k) continue;  // not a first read
                if (rel + 1 - k >= r->frame) {
                    ready = false;  // the block's start tick is not seen complete: not launched yet
                    break;
                }
                if (!r->pending[l]) continue;
                const int w = mxr_level_wait(r, l, now_s(), 0.0);
                if (w == 1) ready = false;
                else if (w) return w;
                else r->pending[l] = false;
            }
            if (!ready) break;
            trel[rel % W] = (rel == r->frame) ? t_idle : now_s();
            mxr_store_go(r, rel + 1);
            rel++;
        }
        if (rel == r->frame) {
            // nothing in flight: tick rel waits for a level block
            if (now_s() - t_idle <= PC_LEVEL_LATE_S) continue;
            mxr_store_go(r, PC_GO_STOP);
            HIPCHK(hipStreamSynchronize(r->s_main));
            for (int l = 1; l < r->M; l++) {
                const uint64_t k = 1ull << A.lv[l].kshift;
                if ((rel + 1) % k != 0 || !r->pending[l]) continue;
                const int w = mxr_level_wait(r, l, now_s(), PC_TICK_TIMEOUT_S);
                if (w == 1) return fail(PC_ERR_TIMEOUT, "level block did not complete");
                if (w) return w;
                r->pending[l] = false;
            }
            r->stalls++;
            const int rc = mxr_persist_launch(r, (long long)(fend - r->frame));
            if (rc) return rc;
            continue;  // t_idle stays: the late tick's latency counts from it
        }
        // the oldest tick in flight: every output's flag
        const uint64_t f = r->frame;
        while (oscan < r->n_out && *(const volatile unsigned long long *)(r->h_flags + (size_t)oscan * 8) >= f + 1)
            oscan++;
        if (oscan < r->n_out) {
            if (t_wait < 0) t_wait = now_s();
            else if (now_s() - t_wait > PC_TICK_TIMEOUT_S) return fail(PC_ERR_TIMEOUT, "tick did not complete");
            continue;
        }
        oscan = 0;
        t_wait = -1;
        const double us = 1e6 * (now_s() - trel[f % W]);
        r->lat.push_back(us);
        r->ticks++;
        r->lat_sum += us;
        r->lat_max = std::max(r->lat_max, us);
        if (us > period) r->misses++;
        r->frame++;
        for (int l = 1; l < r->M; l++)
            if (f % (1ull << A.lv[l].kshift) == 0) {
                r->lstart[l] = f;
                r->pending[l] = true;
                mxr_launcher_queue(r, l, f);
            }
        if (rel == r->frame) t_idle = now_s();
    }
    return PC_OK;
}

// d_in / d_out: nframes frames of buffer_size x n_in / n_out interleaved
// samples on the device.  Every tick is released and waited for (real time).
// The strict run of the reference waits for late levels
// (src/partconvMulti.cpp:886-903); here a level block that is not complete
// PC_LEVEL_LATE_S after its deadline tick was submitted ends the persistent
// kernel (it exits at its next release wait, with the state of every served
// tick complete), the host waits for the block, and a new persistent kernel
// serves the remaining ticks: a level kernel that cannot run beside the
// resident one then runs alone.  The output is unchanged; the late tick's
// latency counts from its first submission (a deadline miss).
PC_API int pcxr_run_device(pc_mimo_rt *r, float *d_out, const float *d_in, long long nframes)
{
    if (!r || !d_out || !d_in || nframes < 0) return fail(PC_ERR_ARG, "bad argument");
    if (!r->filters) return fail(PC_ERR_STATE, "run before set_filters");
    if (nframes == 0) return PC_OK;
    const int N0 = r->N0;
    MxrArgs &A = r->A;
    A.in = d_in;
    A.in_stride = (long long)N0 * r->n_in;
    A.out = d_out;
    A.out_stride = (long long)N0 * r->n_out;
    A.fbase = (long long)r->frame;
    {
        const int rc = mxr_persist_launch(r, nframes);
        if (rc) return rc;
    }
    if (r->M > 1) mxr_launcher_start(r);
    const double period = 1e6 * N0 / r->FS;
    int err = PC_OK;
    if (A.window > 1) {
        err = mxr_run_window(r, (uint64_t)nframes, period);
        nframes = 0;  // served (or failed) above
    }
    for (long long t = 0; t < nframes && !err; t++) {
        const uint64_t f = r->frame;
        const double t0 = now_s();
        // blocks first read at this tick (strict: (f + 1) % k == 0) must be complete
        bool late = false;
        for (int l = 1; l < r->M && !err; l++) {
            const uint64_t k = 1ull << A.lv[l].kshift;
            if ((f + 1) % k != 0 || !r->pending[l]) continue;
            const int w = mxr_level_wait(r, l, t0, PC_LEVEL_LATE_S);
            if (w == 1) late = true;
            else if (w) err = w;
            else r->pending[l] = false;
        }
        if (err) break;
        if (late) {
            // tick f is not released: the kernel ends at its release wait
            mxr_store_go(r, PC_GO_STOP);
            HIPCHK(hipStreamSynchronize(r->s_main));
            for (int l = 1; l < r->M && !err; l++) {
                if ((f + 1) % (1ull << A.lv[l].kshift) != 0 || !r->pending[l]) continue;
                const int w = mxr_level_wait(r, l, t0, PC_TICK_TIMEOUT_S);
                if (w == 1) err = fail(PC_ERR_TIMEOUT, "level block did not complete");
                else if (w) err = w;
                else r->pending[l] = false;
            }
            if (err) break;
            r->stalls++;
            err = mxr_persist_launch(r, nframes - t);
            if (err) break;
        }
        mxr_store_go(r, f + 1);
        for (int o = 0; o < r->n_out && !err; o++) {
            const volatile unsigned long long *fl = r->h_flags + (size_t)o * 8;
            while (*fl < f + 1) {
                if (now_s() - t0 > PC_TICK_TIMEOUT_S) {
                    err = fail(PC_ERR_TIMEOUT, "tick did not complete");
                    break;
                }
            }
        }
        if (err) break;
        const double us = 1e6 * (now_s() - t0);
        r->lat.push_back(us);
        r->ticks++;
        r->lat_sum += us;
        r->lat_max = std::max(r->lat_max, us);
        if (us > period) r->misses++;
        r->frame++;
        for (int l = 1; l < r->M && !err; l++)
            if (f % (1ull << A.lv[l].kshift) == 0) {
                r->lstart[l] = f;
                r->pending[l] = true;
                mxr_launcher_queue(r, l, f);
            }
    }
    const int lrc = mxr_launcher_drain(r);
    if (!err && lrc) err = lrc;
    if (err) mxr_store_go(r, PC_GO_STOP);  // the kernel's waits end at once
    HIPCHK(hipStreamSynchronize(r->s_main));
    return err;
}

// One tick on host buffers, interleaved as PartConvMulti::run's
// (src/partconvMulti.cpp:815-928, frameCount == buffer_size): the frame is
// copied into pinned memory, which the kernel reads in place, and the kernel
// writes the output frame to pinned memory, copied out after the tick.
PC_API int pcxr_run(pc_mimo_rt *r, float *out, const float *in, int frameCount)
{
    if (!r || !out || !in) return fail(PC_ERR_ARG, "null argument");
    if (frameCount != r->N0) return fail(PC_ERR_ARG, "frameCount must equal buffer_size");
    const size_t nin = (size_t)r->n_in * r->N0 * sizeof(float), nout = (size_t)r->n_out * r->N0 * sizeof(float);
    if (!r->h_in) {
        HIPCHK(hipHostMalloc(&r->h_in, nin, hipHostMallocCoherent | hipHostMallocMapped));
        HIPCHK(hipHostMalloc(&r->h_out, nout, hipHostMallocCoherent | hipHostMallocMapped));
        HIPCHK(hipHostGetDevicePointer((void **)&r->d_in_map, r->h_in, 0));
        HIPCHK(hipHostGetDevicePointer((void **)&r->d_out_map, r->h_out, 0));
    }
    memcpy(r->h_in, in, nin);
    const int rc = pcxr_run_device(r, r->d_out_map, r->d_in_map, 1);
    if (rc) return rc;
    memcpy(out, r->h_out, nout);
    return PC_OK;
}

PC_API int pcxr_reset(pc_mimo_rt *r)
{
    if (!r) return fail(PC_ERR_ARG, "null handle");
    (void)mxr_launcher_drain(r);  // a failed launch is moot: the state is cleared below
    HIPCHK(hipStreamSynchronize(r->s_main));
    for (int l = 1; l < r->M; l++) HIPCHK(hipStreamSynchronize(r->s_lvl[l]));
    // the tick counter restarts: so do the launch marks
    if (r->ldone)
        for (int l = 0; l < r->M; l++) r->ldone[l].store(0, std::memory_order_relaxed);
    const MxrArgs &A = r->A;
    HIPCHK(hipMemsetAsync(A.ring, 0, (size_t)r->n_in * 2 * A.R * sizeof(float), r->s_main));
    for (int l = 0; l < r->M; l++) {
        const MxrLevel &L = A.lv[l];
        const size_t nb = (size_t)L.N * sizeof(cf);
        HIPCHK(hipMemsetAsync(L.fdl, 0, (size_t)L.ni * L.P * nb, r->s_main));
        HIPCHK(hipMemsetAsync(L.S, 0, (size_t)(l == 0 ? MXR_SLOTS : 1) * r->n_out * nb, r->s_main));
        if (L.ybuf) HIPCHK(hipMemsetAsync(L.ybuf, 0, (size_t)r->n_out * 2 * L.N * sizeof(float), r->s_main));
    }
    HIPCHK(hipStreamSynchronize(r->s_main));
    memset(r->h_flags, 0, (size_t)r->n_out * 8 * sizeof(unsigned long long));
    std::fill(r->pending.begin(), r->pending.end(), false);
    r->frame = 0;
    return PC_OK;
}

PC_API int pcxr_get_stats(const pc_mimo_rt *r, pc_stats *st)
{
    if (!r || !st) return fail(PC_ERR_ARG, "null argument");
    memset(st, 0, sizeof *st);
    st->ticks = r->ticks;
    st->deadline_misses = r->misses;
    st->period_us = 1e6 * r->N0 / r->FS;
    st->persistent_ticks = r->ticks;
    st->stalls = r->stalls;
    if (!r->lat.empty()) {
        std::vector<double> v = r->lat;
        std::sort(v.begin(), v.end());
        st->p50_us = v[v.size() / 2];
        st->p99_us = v[std::min(v.size() - 1, (size_t)(0.99 * v.size()))];
        st->max_us = r->lat_max;
        st->mean_us = r->lat_sum / (double)r->ticks;
    }
    return PC_OK;
}

PC_API int pcxr_clear_stats(pc_mimo_rt *r)
{
    if (!r) return fail(PC_ERR_ARG, "null handle");
    r->lat.clear();
    r->ticks = r->misses = r->stalls = 0;
    r->lat_sum = r->lat_max = 0;
    return PC_OK;
}

// Time the delay-line-sum kernel (k_mxr_part) of level `level` (>= 1) with HIP
// events from now on (up to 64 launches); -1 disables.
PC_API int pcxr_probe(pc_mimo_rt *r, int level)
{
    if (!r) return fail(PC_ERR_ARG, "null handle");
    if (level >= r->M || level == 0 || level < -1) return fail(PC_ERR_ARG, "level out of range");
    for (int l = 1; l < r->M; l++) HIPCHK(hipStreamSynchronize(r->s_lvl[l]));
    if (r->pev.empty())
        for (int q = 0; q < 64; q++) {
            hipEvent_t a, b;
            HIPCHK(hipEventCreate(&a));
            HIPCHK(hipEventCreate(&b));
            r->pev.push_back({a, b});
        }
    r->pev_used = 0;
    r->probe_level = level;
    return PC_OK;
}

// Mean duration, launch count and algorithmic bytes per launch of the probed
// kernel: the P-1 filter-spectrum rows of every pair, the P-1 delay-line rows
// of every input, and the S row of every output.
PC_API int pcxr_probe_read(pc_mimo_rt *r, double *avg_us, long long *launches, double *bytes_per_launch)
{
    if (!r) return fail(PC_ERR_ARG, "null handle");
    double ms = 0;
    for (int q = 0; q < r->pev_used; q++) {
        HIPCHK(hipEventSynchronize(r->pev[q].second));
        float t = 0.f;
        HIPCHK(hipEventElapsedTime(&t, r->pev[q].first, r->pev[q].second));
        ms += t;
    }
    if (avg_us) *avg_us = r->pev_used ? 1e3 * ms / r->pev_used : 0.0;
    if (launches) *launches = r->pev_used;
    if (bytes_per_launch) {
        double b = 0;
        if (r->probe_level > 0) {
            const int l = r->probe_level;
            const double N = r->N[l], P = r->P[l];
            b = 8.0 * N * (P - 1) * ((double)r->n_out * r->n_in + r->n_in) + 8.0 * N * r->n_out;
        }
        *bytes_per_launch = b;
    }
    return PC_OK;
}

PC_API int pcxr_set_release_window(pc_mimo_rt *r, int ticks)
{
    if (!r) return fail(PC_ERR_ARG, "null handle");
    if (ticks < 1 || ticks > MXR_SLOTS - 1) return fail(PC_ERR_ARG, "release window must be 1 .. 3 ticks");
    r->window = ticks;
    return PC_OK;
}

PC_API int pcxr_stall_level(pc_mimo_rt *r, int level, double seconds)
{
    if (!r) return fail(PC_ERR_ARG, "null handle");
    if (level < 1 || level >= r->M) return fail(PC_ERR_ARG, "level out of range");
    if (!(seconds > 0.0 && seconds <= 5.0)) return fail(PC_ERR_ARG, "stall must be in (0, 5] seconds");
    r->stall_us.store((long long)(seconds * 1e6));
    r->stall_l.store(level);
    return PC_OK;
}

PC_API int pcxr_partition(const pc_mimo_rt *r, int *N, int *P, int cap)
{
    if (!r || !N || !P) return -1;
    const int M = std::min(cap, r->M);
    for (int l = 0; l < M; l++) {
        N[l] = r->N[l];
        P[l] = r->P[l];
    }
    return r->M;
}

static void mimo_rt_destroy(pc_mimo_rt *r)
{
    mxr_free(r);
    delete r;
}

PC_API void pcxr_destroy(pc_mimo_rt *r)
{
    if (!r || !live_take(r)) return;
    mimo_rt_destroy(r);
}

// Library shutdown (see live_add): destroy every live handle, return the BAR
// blocks and twiddle tables, and pair each hsa_init with hsa_shut_down.
// Idempotent; afterwards handle creation fails.
PC_API void pc_shutdown(void)
{
    std::vector<std::pair<void *, int>> v;
    {
        std::lock_guard<std::mutex> lk(g_live_mu);
        if (g_shut) return;
        g_shut = true;
        v.swap(g_live);
    }
    // newest first: a handle never outlives one created before it
    for (size_t i = v.size(); i-- > 0;) {
        void *p = v[i].first;
        switch (v[i].second) {
        case HK_MULTI: multi_destroy(static_cast<pc_multi *>(p)); break;
        case HK_FILTER: filter_destroy(static_cast<pc_filter *>(p)); break;
        case HK_CONV: conv_destroy(static_cast<pc_conv *>(p)); break;
        case HK_MIMO: mimo_destroy(static_cast<pc_mimo *>(p)); break;
        case HK_MIMO_RT: mimo_rt_destroy(static_cast<pc_mimo_rt *>(p)); break;
        }
    }
    std::lock_guard<std::mutex> lk(g_bar_mu);
    for (int d = 0; d < 64; d++) {
        BarDevice &B = g_bar[d];
        if (!B.ok) continue;
        for (auto &blk : B.free_blocks) hsa_amd_memory_pool_free(blk.second);
        B.free_blocks.clear();
        if (B.rel_base) hsa_amd_memory_pool_free(B.rel_base);
        B.rel_base = nullptr;
        B.free_slots.clear();
        B.ok = false;
    }
    int cur = -1;
    (void)hipGetDevice(&cur);
    for (int d = 0; d < 64; d++)
        if (g_tw[d]) {
            (void)hipSetDevice(d);
            hipFree(g_tw[d]);
            g_tw[d] = nullptr;
        }
    if (cur >= 0) (void)hipSetDevice(cur);
    for (; g_hsa_inits > 0; g_hsa_inits--) hsa_shut_down();
}
