// runtime.hip -- host runtime and C ABI of libfresnel_hip.so.
//
// Objects mirror the reference's backend classes (fresnel/cpu/*.h): a Device
// owns a HIP stream and the render scratch; a Scene owns its geometries,
// their device copies and the LBVH; a Tracer owns the device output buffers
// and the sampling state (seed, sample count).  Every C entry point catches
// exceptions and returns a status (include/fresnel_hip.h).
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdlib>
#include <cmath>
#include <deque>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "fresnel_hip.h"
#include "kernels/bvh.hip"
#include "kernels/trace.hip"
#include "kernels/image.hip"
#include "runtime/serve_policy.h"

namespace fh
{
// ====================================================================== errors
static thread_local std::string g_last_error;

struct Error : public std::runtime_error
    {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) { }
    };

#define HIPCHECK(expr)                                                                               \
    do                                                                                               \
        {                                                                                            \
        hipError_t e_ = (expr);                                                                      \
        if (e_ != hipSuccess)                                                                        \
            throw Error(FH_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));              \
        } while (0)

#define REQUIRE(cond, msg)                                                                           \
    do                                                                                               \
        {                                                                                            \
        if (!(cond))                                                                                 \
            throw Error(FH_ERR_INVALID, msg);                                                        \
        } while (0)

template<class F> static int guard(F f)
    {
    try
        {
        f();
        g_last_error.clear();
        return FH_OK;
        }
    catch (const Error& e)
        {
        g_last_error = e.what();
        return e.code;
        }
    catch (const std::bad_alloc&)
        {
        g_last_error = "out of host memory";
        return FH_ERR_NOMEM;
        }
    catch (const std::exception& e)
        {
        g_last_error = e.what();
        return FH_ERR_INVALID;
        }
    }

// ====================================================================== device memory
template<class T> struct DBuf
    {
    T* p = nullptr;
    size_t cap = 0;
    DBuf() = default;
    DBuf(const DBuf&) = delete;
    DBuf& operator=(const DBuf&) = delete;
    ~DBuf()
        {
        release();
        }
    void release()
        {
        if (p)
            (void)hipFree(p);
        p = nullptr;
        cap = 0;
        }
    // grow-only; contents are not preserved on growth
    void ensure(size_t n)
        {
        if (n <= cap && p)
            return;
        release();
        size_t bytes = std::max<size_t>(n, 1) * sizeof(T);
        hipError_t e = hipMalloc(&p, bytes);
        if (e != hipSuccess)
            {
            p = nullptr;
            throw Error(FH_ERR_NOMEM, "hipMalloc of " + std::to_string(bytes) + " bytes failed: "
                                          + hipGetErrorString(e));
            }
        cap = std::max<size_t>(n, 1);
        }
    size_t bytes() const
        {
        return cap * sizeof(T);
        }
    };

// Grow-only buffer used by the kernels of stream s: work still queued there
// may read the old allocation, so the stream finishes before it is replaced.
// A replacement holds an eighth more than asked, so sizes that creep up from
// one build to the next (tile-list entries of a moving scene) do not
// reallocate every time.
template<class T> static void grow(DBuf<T>& b, size_t n, hipStream_t s)
    {
    if (b.p && b.cap < n)
        {
        HIPCHECK(hipStreamSynchronize(s));
        n += n / 8;
        }
    b.ensure(n);
    }

static unsigned blocks(size_t n, unsigned b)
    {
    return (unsigned)((n + b - 1) / b);
    }

// ====================================================================== allocation-free scratch
// Device memory of fh_tracer_histogram_noalloc: the 4 x 512 bin counts its
// kernel accumulates.  Each GPU the library runs on holds its own copy of
// this array until the process ends.  In host code the name is only a
// stand-in: DeviceObj::scratch holds the array's device address on that GPU,
// taken with hipGetSymbolAddress when the Device is created, and that address
// is what kernels and copies receive.  Every call on one GPU shares the
// array (calls that overlap on one GPU race); each call clears it first.
constexpr size_t NOALLOC_SCRATCH_WORDS = 4 * HIST_BINS; // 8 KiB
static __device__ __align__(256) uint32_t scratch[NOALLOC_SCRATCH_WORDS];

// ====================================================================== device
constexpr int TRAV_BLOCK = 128;
constexpr int TRAV_SS = 16;
constexpr int SHADE_BLOCK = 512; // shading workgroup: one queue-append atomic per 8 wavefronts
constexpr uint32_t MAX_DEPTH = 4096;    // safety cap on path length (the reference has none)

struct KernelTime
    {
    std::string name;
    hipEvent_t a, b;
    // a bounce-step kernel: the step's mode word (pinned host copy, valid once
    // the render has finished) and the mode the kernel acts on; launches of
    // other modes exit at once and are counted as "<name>_idle"
    const uint32_t* mode = nullptr;
    uint32_t acts_on = 0;
    };

struct RenderScratch
    {
    DBuf<float4> qo[2], qd[2], qatt[2];
    DBuf<uint4> qmeta[2];
    DBuf<uint4> hits;
    DBuf<uint4> records;
    DBuf<uint32_t> counters32; // [0] traverse work
    DBuf<uint32_t> qseg;       // append counters of the queue being filled (QSEG_STRIDE apart) and flag
    DBuf<uint2> gstack;        // traversal stack overflow (persistent grid)
    DBuf<uint2> tail_stack;    // stack overflow of the tail kernel
    DBuf<JobDev> jobs;
    DBuf<uint32_t> pix_base;
    DBuf<QueueView> view;      // consumer view of the queue a bounce step reads (k_queue_view)
    DBuf<uint32_t> steps;      // count and mode of every bounce step (2 words each)
    };

// Pinned host memory that grows on demand (contents are not preserved).
struct PinnedBuf
    {
    uint8_t* p = nullptr;
    size_t cap = 0;
    PinnedBuf() = default;
    PinnedBuf(const PinnedBuf&) = delete;
    PinnedBuf& operator=(const PinnedBuf&) = delete;
    ~PinnedBuf()
        {
        if (p)
            (void)hipHostFree(p);
        }
    // only while no copy from it is pending
    void ensure(size_t bytes)
        {
        if (bytes <= cap && p)
            return;
        if (p)
            (void)hipHostFree(p);
        p = nullptr;
        cap = 0;
        void* q = nullptr;
        HIPCHECK(hipHostMalloc(&q, std::max<size_t>(bytes, 64), hipHostMallocDefault));
        p = (uint8_t*)q;
        cap = std::max<size_t>(bytes, 64);
        }
    };

// Bounce steps a batch can be given: the last one finishes every path left.
constexpr uint32_t MAX_STEPS = 64;

// Temporaries of hierarchy builds and of tile-list builds.  Both run on the
// device stream, one at a time under the device lock, so every scene and
// tracer shares one set of each (grow-only).
struct BvhTemps
    {
    DBuf<uint32_t> keys, keys2, vals2, hist;
    std::deque<DBuf<uint32_t>> scan_tmp;
    DBuf<float> partial, cb;
    };
struct TileTemps
    {
    DBuf<uint32_t> counts, flags, tile0, q0, ref0, keysA, valsA, keysB, valsB, hist;
    std::deque<DBuf<uint32_t>> scan_tmp;
    };

struct DeviceObj
    {
    int hip_dev = 0;
    hipStream_t stream = nullptr;
    hipStream_t stream2 = nullptr; // second render stream: two batches in flight
    hipStream_t stream3 = nullptr; // first render stream (`stream` builds scenes and tile lists)
    hipStream_t copy_stream = nullptr; // asynchronous output readback; scene uploads
    hipEvent_t upload_ev = nullptr;    // end of the scene uploads queued on copy_stream
    hipEvent_t build_ev = nullptr;     // device-stream work a render batch waits for (InflightRun)
    hipStream_t lstream = nullptr; // stream of the launches being enqueued (timing events)
    int n_cu = 0;
    std::string name;
    int limit = -1;
    uint32_t trav_grid = 0;
    RenderScratch rs, rs2; // render scratch of the two batches in flight
    PinnedBuf stage;       // job tables, pixel bases and step records of one render's batches
    uint32_t graph = 1;     // 2: a batch's kernels are captured once into a HIP graph and replayed; 1: launched
    struct GraphEntry
        {
        std::vector<uint64_t> key; // every parameter of the captured launches
        hipGraphExec_t exec;
        uint64_t used;
        };
    std::vector<GraphEntry> graph_cache;
    uint64_t graph_clock = 0, graph_captures = 0, graph_replays = 0;
    uint32_t steps_est = 8; // bounce steps per batch (adapted to the depths of recent renders)
    float live_frac[MAX_STEPS] = {}; // fraction of a batch's paths queued at each bounce step, recent maximum
    const uint32_t* tag_mode = nullptr; // mode word for the launch being recorded (stats)
    uint32_t tag_acts_on = 0;
    hipEvent_t acc_done[2] = {nullptr, nullptr}; // last accumulation enqueued on stream3 / stream2
    DBuf<unsigned long long> counters;
    BvhTemps bvh_tmp;
    TileTemps tile_tmp;
    uint32_t* scratch = nullptr;  // device address of this GPU's allocation-free scratch array
    uint32_t* h_pinned = nullptr; // 16 uint32 of pinned host memory for counts
    uint32_t* h_counts = nullptr; // pinned copy of a queue's segment counters and flag
    bool stats_on = false;
    bool count_on = false;
    uint64_t batch_paths = 256ull << 20; // light-sample paths per batch (sizes the queues)
    uint32_t tail_rays = 524288;        // finish paths in one kernel below this many rays
    uint32_t trav_chunk = 256;          // rays a secondary-ray traversal wavefront claims per atomic
    uint32_t primary_chunk = 1024;      // the most primary rays a workgroup claims at once
    uint32_t trav_packet = 2;           // primary rays: 2 wavefront packets, 1 per ray
    uint32_t primary_tiles = 2;         // primary rays of pinhole/orthographic cameras: 2 tile lists, 1 BVH
    uint32_t tile_group = 4;            // tile-list subtrees cover at most this many tiles (1: primitives only)
    uint64_t tile_min_rays = 4ull << 20; // tile lists once this many primary rays used one scene and camera
    uint32_t fuse_primary = 2;          // 2: primary traversal and depth-0 shading in one kernel, 1: two
    uint32_t fuse_bounce = 1;           // 2: secondary traversal and shading in one kernel, 1: two
    uint32_t inflight = 2;              // batches of one Path.sample in flight (1 or 2)
    uint32_t xcd_bands = 2;             // 2: queue segments hold image bands, traversal claims per XCD; 1: off
    uint32_t trav_refill = 65;          // lanes take new rays once this many of a wavefront are idle
                                        // (65: batch-synchronous kernel, 64 rays at a time)
    int stack_lds = 8;                  // traversal stack entries kept in LDS (8 or 16)
    std::map<const void*, uint32_t> grids;
    uint32_t grid_for(const void* kernel, int block)
        {
        auto it = grids.find(kernel);
        if (it != grids.end())
            return it->second;
        int per_cu = 0;
        HIPCHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, block, 0));
        const uint32_t g = (uint32_t)std::max(1, per_cu) * (uint32_t)n_cu;
        grids[kernel] = g;
        return g;
        }
    std::vector<KernelTime> times;
    std::vector<hipEvent_t> event_pool;
    std::vector<std::pair<std::string, std::pair<double, uint32_t>>> last_stats;
    uint64_t last_counters[N_CTR] = {};
    std::mutex mu;

    hipEvent_t event()
        {
        if (!event_pool.empty())
            {
            hipEvent_t e = event_pool.back();
            event_pool.pop_back();
            return e;
            }
        hipEvent_t e;
        HIPCHECK(hipEventCreate(&e));
        return e;
        }
    template<class F> void launch(const char* nm, F f)
        {
        if (!stats_on)
            {
            f();
            HIPCHECK(hipGetLastError());
            return;
            }
        KernelTime kt;
        kt.name = nm;
        kt.a = event();
        kt.b = event();
        kt.mode = tag_mode;
        kt.acts_on = tag_acts_on;
        hipStream_t ls = lstream ? lstream : stream;
        HIPCHECK(hipEventRecord(kt.a, ls));
        f();
        HIPCHECK(hipGetLastError());
        HIPCHECK(hipEventRecord(kt.b, ls));
        times.push_back(kt);
        }
    void begin_stats()
        {
        for (auto& t : times)
            {
            event_pool.push_back(t.a);
            event_pool.push_back(t.b);
            }
        times.clear();
        counters.ensure(N_CTR);
        HIPCHECK(hipMemsetAsync(counters.p, 0, N_CTR * sizeof(unsigned long long), stream));
        }
    void end_stats()
        {
        HIPCHECK(hipStreamSynchronize(stream));
        if (stream2)
            HIPCHECK(hipStreamSynchronize(stream2));
        if (stream3)
            HIPCHECK(hipStreamSynchronize(stream3));
        last_stats.clear();
        for (auto& t : times)
            {
            float ms = 0;
            HIPCHECK(hipEventElapsedTime(&ms, t.a, t.b));
            if (t.mode && *t.mode != t.acts_on)
                t.name += "_idle";
            bool found = false;
            for (auto& s : last_stats)
                if (s.first == t.name)
                    {
                    s.second.first += ms;
                    s.second.second += 1;
                    found = true;
                    }
            if (!found)
                last_stats.push_back({t.name, {ms, 1u}});
            }
        unsigned long long c[N_CTR];
        HIPCHECK(hipMemcpy(c, counters.p, sizeof(c), hipMemcpyDeviceToHost));
        for (int k = 0; k < N_CTR; k++)
            last_counters[k] = c[k];
        }
    };


// ====================================================================== traversal launch
struct TravArgs
    {
    const JobDev* jobs;
    uint32_t n_jobs;
    Queue q;
    uint32_t n;
    const QueueView* view; // device-written count and segment prefix, or nullptr (n, q.pre/adj stand)
    uint4* hits;
    uint32_t* work;
    uint32_t chunk;
    uint2* gstack;
    uint32_t gstride;
    unsigned long long* counters;
    };

template<int SS, bool C, bool P, uint32_t T>
static void launch_traverse(DeviceObj& D, const char* nm, hipStream_t s, const TravArgs& a)
    {
    if (P && D.trav_packet == 2)
        {
        auto kern = &k_traverse_packet<TRAV_BLOCK, SS, C, T>;
        const uint32_t grid = D.grid_for(reinterpret_cast<const void*>(kern), TRAV_BLOCK);
        REQUIRE((uint64_t)grid * TRAV_BLOCK <= a.gstride, "traversal grid exceeds the overflow stacks");
        D.launch(nm,
                 [&]
                 {
                     kern<<<grid, TRAV_BLOCK, 0, s>>>(a.jobs, a.n_jobs, a.n, a.hits, a.work, D.primary_chunk,
                                                     a.gstack, a.gstride, a.counters);
                 });
        return;
        }
    if (D.trav_refill <= 64)
        {
        auto kern = &k_traverse_dyn<TRAV_BLOCK, SS, C, P, T>;
        const uint32_t grid = D.grid_for(reinterpret_cast<const void*>(kern), TRAV_BLOCK);
        REQUIRE((uint64_t)grid * TRAV_BLOCK <= a.gstride, "traversal grid exceeds the overflow stacks");
        D.launch(nm,
                 [&]
                 {
                     kern<<<grid, TRAV_BLOCK, 0, s>>>(a.jobs, a.n_jobs, a.q, a.n, a.view, a.hits, a.work, a.chunk,
                                                     D.xcd_bands == 2 ? 1u : 0u, D.trav_refill, a.gstack,
                                                     a.gstride, a.counters);
                 });
        return;
        }
    auto kern = &k_traverse<TRAV_BLOCK, SS, C, P, T>;
    const uint32_t grid = D.grid_for(reinterpret_cast<const void*>(kern), TRAV_BLOCK);
    REQUIRE((uint64_t)grid * TRAV_BLOCK <= a.gstride, "traversal grid exceeds the overflow stacks");
    D.launch(nm,
             [&]
             {
                 kern<<<grid, TRAV_BLOCK, 0, s>>>(a.jobs, a.n_jobs, a.q, a.n, a.view, a.hits, a.work, a.chunk,
                                                 D.xcd_bands == 2 ? 1u : 0u, a.gstack, a.gstride, a.counters);
             });
    }

template<int SS>
static void dispatch_ss(DeviceObj& D, const char* nm, hipStream_t s, const TravArgs& a, bool primary, bool spheres)
    {
    const bool c = D.count_on;
    if (primary)
        {
        if (spheres)
            c ? launch_traverse<SS, true, true, T_SPHERES>(D, nm, s, a) : launch_traverse<SS, false, true, T_SPHERES>(D, nm, s, a);
        else
            c ? launch_traverse<SS, true, true, T_ALL>(D, nm, s, a) : launch_traverse<SS, false, true, T_ALL>(D, nm, s, a);
        }
    else
        {
        if (spheres)
            c ? launch_traverse<SS, true, false, T_SPHERES>(D, nm, s, a) : launch_traverse<SS, false, false, T_SPHERES>(D, nm, s, a);
        else
            c ? launch_traverse<SS, true, false, T_ALL>(D, nm, s, a) : launch_traverse<SS, false, false, T_ALL>(D, nm, s, a);
        }
    }

// Fused primary traversal + depth-0 shading (k_primary), 256-thread workgroups.
constexpr int PRIMARY_BLOCK = 256;
struct PrimaryArgs
    {
    const JobDev* jobs;
    uint32_t n_jobs;
    uint32_t n;
    uint32_t* work;
    uint32_t groups;
    uint2* gstack;
    uint32_t gstride;
    Queue out;
    uint4* records;
    unsigned long long* counters;
    };

template<int SS, bool C, uint32_t T>
static void launch_primary(DeviceObj& D, hipStream_t s, const PrimaryArgs& a)
    {
    auto kern = &k_primary<PRIMARY_BLOCK, SS, C, T>;
    const uint32_t grid = D.grid_for(reinterpret_cast<const void*>(kern), PRIMARY_BLOCK);
    REQUIRE((uint64_t)grid * PRIMARY_BLOCK <= a.gstride, "primary grid exceeds the overflow stacks");
    D.launch("primary",
             [&]
             {
                 kern<<<grid, PRIMARY_BLOCK, 0, s>>>(a.jobs, a.n_jobs, a.n, a.work, a.groups,
                                                     D.xcd_bands == 2 ? 1u : 0u, a.gstack, a.gstride, a.out,
                                                     a.records, a.counters);
             });
    }

static void dispatch_primary(DeviceObj& D, hipStream_t s, const PrimaryArgs& a, bool spheres)
    {
    const bool c = D.count_on;
    if (D.stack_lds == 8)
        {
        if (spheres)
            c ? launch_primary<8, true, T_SPHERES>(D, s, a) : launch_primary<8, false, T_SPHERES>(D, s, a);
        else
            c ? launch_primary<8, true, T_ALL>(D, s, a) : launch_primary<8, false, T_ALL>(D, s, a);
        }
    else
        {
        if (spheres)
            c ? launch_primary<16, true, T_SPHERES>(D, s, a) : launch_primary<16, false, T_SPHERES>(D, s, a);
        else
            c ? launch_primary<16, true, T_ALL>(D, s, a) : launch_primary<16, false, T_ALL>(D, s, a);
        }
    }

// Fused secondary traversal + shading (k_bounce).
struct BounceArgs
    {
    const JobDev* jobs;
    Queue in;
    uint32_t n;
    const QueueView* view;
    uint32_t* work;
    uint32_t chunk;
    uint2* gstack;
    uint32_t gstride;
    Queue out;
    uint4* records;
    unsigned long long* counters;
    };

template<int SS, bool C, uint32_t T>
static void launch_bounce(DeviceObj& D, hipStream_t s, const BounceArgs& a)
    {
    auto kern = &k_bounce<TRAV_BLOCK, SS, C, T>;
    const uint32_t grid = D.grid_for(reinterpret_cast<const void*>(kern), TRAV_BLOCK);
    REQUIRE((uint64_t)grid * TRAV_BLOCK <= a.gstride, "bounce grid exceeds the overflow stacks");
    D.launch("bounce",
             [&]
             {
                 kern<<<grid, TRAV_BLOCK, 0, s>>>(a.jobs, a.in, a.n, a.view, a.work, a.chunk, a.gstack, a.gstride,
                                                 a.out, a.records, MAX_DEPTH, a.counters);
             });
    }

static void dispatch_bounce(DeviceObj& D, hipStream_t s, const BounceArgs& a, bool spheres)
    {
    const bool c = D.count_on;
    if (D.stack_lds == 8)
        {
        if (spheres)
            c ? launch_bounce<8, true, T_SPHERES>(D, s, a) : launch_bounce<8, false, T_SPHERES>(D, s, a);
        else
            c ? launch_bounce<8, true, T_ALL>(D, s, a) : launch_bounce<8, false, T_ALL>(D, s, a);
        }
    else
        {
        if (spheres)
            c ? launch_bounce<16, true, T_SPHERES>(D, s, a) : launch_bounce<16, false, T_SPHERES>(D, s, a);
        else
            c ? launch_bounce<16, true, T_ALL>(D, s, a) : launch_bounce<16, false, T_ALL>(D, s, a);
        }
    }

static void dispatch_traverse(DeviceObj& D, const char* nm, hipStream_t s, const TravArgs& a, bool primary, bool spheres)
    {
    if (D.stack_lds == 8)
        dispatch_ss<8>(D, nm, s, a, primary, spheres);
    else
        dispatch_ss<16>(D, nm, s, a, primary, spheres);
    }

// Largest resident grid of any traversal variant (sizes the overflow stacks).
template<int SS> static uint32_t max_grid_ss(DeviceObj& D)
    {
    uint32_t g = 0;
    for (const void* k : {reinterpret_cast<const void*>(&k_traverse<TRAV_BLOCK, SS, false, false, T_SPHERES>),
                          reinterpret_cast<const void*>(&k_traverse<TRAV_BLOCK, SS, false, true, T_SPHERES>),
                          reinterpret_cast<const void*>(&k_traverse<TRAV_BLOCK, SS, true, false, T_SPHERES>),
                          reinterpret_cast<const void*>(&k_traverse<TRAV_BLOCK, SS, true, true, T_SPHERES>),
                          reinterpret_cast<const void*>(&k_traverse<TRAV_BLOCK, SS, false, false, T_ALL>),
                          reinterpret_cast<const void*>(&k_traverse<TRAV_BLOCK, SS, false, true, T_ALL>),
                          reinterpret_cast<const void*>(&k_traverse<TRAV_BLOCK, SS, true, false, T_ALL>),
                          reinterpret_cast<const void*>(&k_traverse<TRAV_BLOCK, SS, true, true, T_ALL>),
                          reinterpret_cast<const void*>(&k_traverse_dyn<TRAV_BLOCK, SS, false, false, T_SPHERES>),
                          reinterpret_cast<const void*>(&k_traverse_dyn<TRAV_BLOCK, SS, false, true, T_SPHERES>),
                          reinterpret_cast<const void*>(&k_traverse_dyn<TRAV_BLOCK, SS, true, false, T_SPHERES>),
                          reinterpret_cast<const void*>(&k_traverse_dyn<TRAV_BLOCK, SS, true, true, T_SPHERES>),
                          reinterpret_cast<const void*>(&k_traverse_dyn<TRAV_BLOCK, SS, false, false, T_ALL>),
                          reinterpret_cast<const void*>(&k_traverse_dyn<TRAV_BLOCK, SS, false, true, T_ALL>),
                          reinterpret_cast<const void*>(&k_traverse_dyn<TRAV_BLOCK, SS, true, false, T_ALL>),
                          reinterpret_cast<const void*>(&k_traverse_dyn<TRAV_BLOCK, SS, true, true, T_ALL>),
                          reinterpret_cast<const void*>(&k_traverse_packet<TRAV_BLOCK, SS, false, T_SPHERES>),
                          reinterpret_cast<const void*>(&k_traverse_packet<TRAV_BLOCK, SS, true, T_SPHERES>),
                          reinterpret_cast<const void*>(&k_traverse_packet<TRAV_BLOCK, SS, false, T_ALL>),
                          reinterpret_cast<const void*>(&k_traverse_packet<TRAV_BLOCK, SS, true, T_ALL>)})
        g = std::max(g, D.grid_for(k, TRAV_BLOCK));
    // fused primary kernel: its workgroups are PRIMARY_BLOCK / TRAV_BLOCK traversal workgroups
    for (const void* k : {reinterpret_cast<const void*>(&k_primary<PRIMARY_BLOCK, SS, false, T_SPHERES>),
                          reinterpret_cast<const void*>(&k_primary<PRIMARY_BLOCK, SS, true, T_SPHERES>),
                          reinterpret_cast<const void*>(&k_primary<PRIMARY_BLOCK, SS, false, T_ALL>),
                          reinterpret_cast<const void*>(&k_primary<PRIMARY_BLOCK, SS, true, T_ALL>)})
        g = std::max(g, D.grid_for(k, PRIMARY_BLOCK) * (PRIMARY_BLOCK / TRAV_BLOCK));
    for (const void* k : {reinterpret_cast<const void*>(&k_bounce<TRAV_BLOCK, SS, false, T_SPHERES>),
                          reinterpret_cast<const void*>(&k_bounce<TRAV_BLOCK, SS, true, T_SPHERES>),
                          reinterpret_cast<const void*>(&k_bounce<TRAV_BLOCK, SS, false, T_ALL>),
                          reinterpret_cast<const void*>(&k_bounce<TRAV_BLOCK, SS, true, T_ALL>)})
        g = std::max(g, D.grid_for(k, TRAV_BLOCK));
    return g;
    }

// ====================================================================== geometry
struct SceneObj;

struct GeometryObj
    {
    SceneObj* scene = nullptr;
    uint32_t type = 0;
    uint32_t slot = 0;
    bool valid = true;
    bool enabled = true;
    uint32_t N = 0;       // primitives (mesh: instances)
    uint32_t n_faces = 0; // mesh
    fh_material mat {}, outline {};
    float outline_width = 0;
    float color_by_face = 0;
    float radius = 0;   // polyhedron bounding radius / polygon xy radius
    float rounding = 0; // polygon
    // host buffers (authoritative)
    std::vector<float> position, radius_buf, color, points, orientation, angle;
    std::vector<float> plane_n, plane_p, plane_color; // polyhedron (normalized normals)
    std::vector<float> verts;                         // mesh (3F x 3) / polygon (V x 2)
    // device copies
    DBuf<float4> a, b, aux0, aux1;
    DBuf<float> dcolor, daux_color;
    bool dirty = true; // upload everything on the next sync
    // rows changed since the last upload, [lo, hi) along the buffers' first
    // axis: those that feed the hierarchy (positions, radii, points,
    // orientations, angles) and colour rows (shading only)
    uint64_t shape_lo = UINT64_MAX, shape_hi = 0;
    uint64_t color_lo = UINT64_MAX, color_hi = 0;
    bool aux_uploaded = false;
    };

static fh_material make_material(float r, float g, float b, float solid)
    {
    // Material(RGB, solid), fresnel/common/Material.h:61-65
    fh_material m;
    m.solid = solid;
    m.color = fh_rgb {r, g, b};
    m.primitive_color_mix = 0.0f;
    m.roughness = 0.1f;
    m.specular = 0.5f;
    m.metal = 0.0f;
    m.spec_trans = 0.0f;
    return m;
    }

static Mat to_mat(const fh_material& m)
    {
    Mat r;
    r.solid = m.solid;
    r.cr = m.color.r;
    r.cg = m.color.g;
    r.cb = m.color.b;
    r.mix = m.primitive_color_mix;
    r.roughness = m.roughness;
    r.specular = m.specular;
    r.metal = m.metal;
    r.spec_trans = m.spec_trans;
    return r;
    }

// ====================================================================== scene
// Serial numbers of scenes: a scene created after another one was destroyed
// can reuse its address, so caches keyed on a scene use its serial.
static std::atomic<uint64_t> g_scene_serial {0};

struct SceneObj
    {
    DeviceObj* dev = nullptr;
    uint64_t serial = ++g_scene_serial; // unique per scene for the life of the process
    fh_user_camera camera {};
    fh_rgb bg {0, 0, 0};
    float bg_alpha = 0;
    fh_lights lights {};
    std::map<uint32_t, GeometryObj*> geoms; // slot -> geometry
    // device scene
    DBuf<GeomDev> dgeoms;
    uint32_t n_slots = 0;
    DBuf<float4> blo, bhi;
    DBuf<uint32_t> vals, parent, flags, prefix, slots; // vals: primitive order of the leaves (refit reads it)
    DBuf<float> root_box, extent;
    float cull_margin = 0.0f; // CULL_K * 3 * largest primitive half-extent
    bool margin_pending = false; // cull_margin not yet read back from the last build (resolve_margin)
    bool build_pending = false;  // a build or refit was queued on the device stream since its last sync
    DBuf<Node> nodes;
    DBuf<Node4> nodes4;
    bool sphere_leaves = false; // nodes4 leaf children hold spheres
    DBuf<uint32_t> idx4; // 4-wide node of each binary inner node (exclusive scan of even-depth flags)
    uint32_t n_prims = 0;
    uint32_t n_active = 0; // geometries with primitives: entries of prefix / slots
    bool bvh_valid = false;
    std::vector<uint32_t> topo; // (slot, prim count) of the enabled geometries in the last build
    uint64_t builds = 0, refits = 0;
    bool allow_refit = false;   // refit instead of rebuilding when only positions moved
    float refit_limit = 1.5f;   // rebuild when a refit raises the surface-area cost by this factor
    float sah_built = 0.0f;     // surface-area cost right after the last build
    float sah_last = 0.0f;      // surface-area cost of the current hierarchy

    uint32_t alloc_slot()
        {
        uint32_t s = 0;
        while (geoms.count(s))
            s++;
        return s;
        }
    };

static void upload(DBuf<float4>& d, const std::vector<float4>& h, hipStream_t s)
    {
    d.ensure(h.size());
    if (!h.empty())
        HIPCHECK(hipMemcpyAsync(d.p, h.data(), h.size() * sizeof(float4), hipMemcpyHostToDevice, s));
    }
static void upload(DBuf<float>& d, const std::vector<float>& h, hipStream_t s)
    {
    d.ensure(h.size());
    if (!h.empty())
        HIPCHECK(hipMemcpyAsync(d.p, h.data(), h.size() * sizeof(float), hipMemcpyHostToDevice, s));
    }

// Pack rows [lo, hi) of one geometry's host buffers into the device layout of
// GeomDev (a[i], b[i] per primitive, or per instance for a mesh).
static void pack_rows(const GeometryObj& g, uint64_t lo, uint64_t hi, std::vector<float4>& a,
                      std::vector<float4>& b)
    {
    const size_t n = (size_t)(hi - lo);
    a.resize(n);
    if (g.type != G_SPHERE)
        b.resize(n);
    for (size_t k = 0; k < n; k++)
        {
        const size_t i = (size_t)lo + k;
        switch (g.type)
            {
        case G_SPHERE:
            a[k] = make_float4(g.position[3 * i], g.position[3 * i + 1], g.position[3 * i + 2], g.radius_buf[i]);
            break;
        case G_CYLINDER:
            {
            const float* p = &g.points[6 * i];
            a[k] = make_float4(p[0], p[1], p[2], g.radius_buf[i]);
            b[k] = make_float4(p[3], p[4], p[5], 0.0f);
            break;
            }
        case G_POLYHEDRON:
        case G_MESH:
            {
            a[k] = make_float4(g.position[3 * i], g.position[3 * i + 1], g.position[3 * i + 2], 0.0f);
            const float* q = &g.orientation[4 * i];
            b[k] = make_float4(q[0], q[1], q[2], q[3]);
            break;
            }
        case G_POLYGON:
            {
            a[k] = make_float4(g.position[2 * i], g.position[2 * i + 1], 0.0f, 0.0f);
            // quat::fromAxisAngle((0,0,1), angle), fresnel/common/VectorMath.h:920-925,
            // evaluated on the host with the C library's cosf/sinf
            const float th = g.angle[i];
            const float c = cosf(th / 2.0f), sn = sinf(th / 2.0f);
            b[k] = make_float4(c, sn * 0.0f, sn * 0.0f, sn * 1.0f);
            break;
            }
            }
        }
    }

// Copy n elements of h into d starting at element `at` (d already sized).
template <class T>
static void upload_at(DBuf<T>& d, size_t at, const T* h, size_t n, hipStream_t s)
    {
    if (n)
        HIPCHECK(hipMemcpyAsync(d.p + at, h, n * sizeof(T), hipMemcpyHostToDevice, s));
    }

// Floats of the colour buffer per row of its first axis.
static size_t color_row_floats(const GeometryObj& g)
    {
    return g.type == G_CYLINDER ? 6 : 3;
    }

// Bring a geometry's device buffers up to date: everything after creation or
// update(), else only the rows marked by update_rows().  Returns whether
// anything the hierarchy depends on (positions, radii, orientations) changed.
static bool upload_geometry(GeometryObj& g, hipStream_t s)
    {
    std::vector<float4> a, b;
    if (!g.dirty)
        {
        const bool shape = g.shape_hi > g.shape_lo;
        if (shape)
            {
            pack_rows(g, g.shape_lo, g.shape_hi, a, b);
            upload_at(g.a, (size_t)g.shape_lo, a.data(), a.size(), s);
            if (!b.empty())
                upload_at(g.b, (size_t)g.shape_lo, b.data(), b.size(), s);
            }
        if (g.color_hi > g.color_lo)
            {
            const size_t f = color_row_floats(g);
            upload_at(g.dcolor, (size_t)g.color_lo * f, g.color.data() + g.color_lo * f,
                      (size_t)(g.color_hi - g.color_lo) * f, s);
            }
        g.shape_lo = g.color_lo = UINT64_MAX;
        g.shape_hi = g.color_hi = 0;
        // the packed rows must stay alive until the copies have read them
        if (shape)
            HIPCHECK(hipStreamSynchronize(s));
        return shape;
        }
    pack_rows(g, 0, g.N, a, b);
    upload(g.a, a, s);
    upload(g.b, b, s);
    upload(g.dcolor, g.color, s);
    if (!g.aux_uploaded)
        {
        std::vector<float4> x0, x1;
        if (g.type == G_POLYHEDRON)
            {
            const size_t P = g.plane_n.size() / 3;
            for (size_t k = 0; k < P; k++)
                {
                const v3 n = mk3(g.plane_n[3 * k], g.plane_n[3 * k + 1], g.plane_n[3 * k + 2]);
                const v3 p = mk3(g.plane_p[3 * k], g.plane_p[3 * k + 1], g.plane_p[3 * k + 2]);
                x0.push_back(make_float4(n.x, n.y, n.z, -dot(n, p)));
                x1.push_back(make_float4(p.x, p.y, p.z, 0.0f));
                }
            upload(g.daux_color, g.plane_color, s);
            }
        else if (g.type == G_MESH)
            {
            for (size_t k = 0; k < g.verts.size() / 3; k++)
                x0.push_back(make_float4(g.verts[3 * k], g.verts[3 * k + 1], g.verts[3 * k + 2], 0.0f));
            }
        else if (g.type == G_POLYGON)
            {
            for (size_t k = 0; k < g.verts.size() / 2; k++)
                x0.push_back(make_float4(g.verts[2 * k], g.verts[2 * k + 1], 0.0f, 0.0f));
            }
        upload(g.aux0, x0, s);
        upload(g.aux1, x1, s);
        g.aux_uploaded = true;
        HIPCHECK(hipStreamSynchronize(s));
        }
    HIPCHECK(hipStreamSynchronize(s));
    g.dirty = false;
    g.shape_lo = g.color_lo = UINT64_MAX;
    g.shape_hi = g.color_hi = 0;
    return true;
    }

static uint32_t prim_count(const GeometryObj& g)
    {
    return g.type == G_MESH ? g.N * g.n_faces : g.N;
    }

// Build or refit the LBVH over the enabled geometries.
// Exclusive prefix sum of n uint32 in place (any n; tmp holds one buffer per level).
static void scan_exclusive(DeviceObj& D, hipStream_t s, uint32_t* data, uint32_t n,
                           std::deque<DBuf<uint32_t>>& tmp, size_t level = 0)
    {
    if (n <= SCAN_SINGLE_MAX)
        {
        k_scan_single<<<1, 1024, 0, s>>>(data, n);
        return;
        }
    const uint32_t n_sums = blocks(n, 2048);
    while (tmp.size() <= level)
        tmp.emplace_back();
    grow(tmp[level], std::max<uint32_t>(n_sums, 1), s);
    uint32_t* sums = tmp[level].p;
    k_scan_blocks<<<n_sums, 256, 0, s>>>(data, n, sums);
    if (n_sums <= 2048)
        k_scan_top<<<1, 256, 0, s>>>(sums, n_sums);
    else
        scan_exclusive(D, s, sums, n_sums, tmp, level + 1);
    k_scan_add<<<blocks(n, 256), 256, 0, s>>>(data, n, sums);
    }

// Stable LSD radix sort of (key, value) pairs on key bits [bit0, bit1) in
// 8-bit digits, ping-ponging between (k0, v0) and (k1, v1); returns whether
// the result is in (k1, v1).
static bool radix_sort_pairs(DeviceObj& D, hipStream_t s, const char* name, uint32_t* k0, uint32_t* v0, uint32_t* k1,
                             uint32_t* v1, uint32_t N, uint32_t bit0, uint32_t bit1, DBuf<uint32_t>& hist,
                             std::deque<DBuf<uint32_t>>& tmp)
    {
    // 16 keys per thread while that gives every CU two workgroups, else 4: a
    // workgroup ranks its keys in that many dependent rounds, so a smaller sort
    // spread over more workgroups finishes sooner
    const bool wide = blocks(N, SORT_BLOCK * 16) >= 2u * (uint32_t)std::max(1, D.n_cu);
    const uint32_t nb = blocks(N, SORT_BLOCK * (wide ? 16 : 4));
    const uint32_t nh = 256 * nb;
    grow(hist, nh, s);
    uint32_t *kin = k0, *vin = v0, *kout = k1, *vout = v1;
    bool flipped = false;
    for (uint32_t shift = bit0; shift < bit1; shift += 8)
        {
        D.launch(name,
                 [&]
                 {
                     if (wide)
                         k_radix_hist<16><<<nb, SORT_BLOCK, 0, s>>>(kin, N, shift, nb, hist.p);
                     else
                         k_radix_hist<4><<<nb, SORT_BLOCK, 0, s>>>(kin, N, shift, nb, hist.p);
                     scan_exclusive(D, s, hist.p, nh, tmp);
                     if (wide)
                         k_radix_scatter<16><<<nb, SORT_BLOCK, 0, s>>>(kin, vin, kout, vout, N, shift, nb, hist.p);
                     else
                         k_radix_scatter<4><<<nb, SORT_BLOCK, 0, s>>>(kin, vin, kout, vout, N, shift, nb, hist.p);
                 });
        std::swap(kin, kout);
        std::swap(vin, vout);
        flipped = !flipped;
        }
    return flipped;
    }

static uint32_t scene_types(const SceneObj& sc);

// Scene data reaches the device by copies on the copy stream.  The device
// stream waits for the copies queued so far by event before its build
// kernels, and the host waits for them too (the host buffers they read may
// change or go away after return), but not for earlier scenes' builds still
// queued on the device stream.
static void uploads_done(DeviceObj& D)
    {
    HIPCHECK(hipEventRecord(D.upload_ev, D.copy_stream));
    HIPCHECK(hipStreamWaitEvent(D.stream, D.upload_ev, 0));
    HIPCHECK(hipEventSynchronize(D.upload_ev));
    }

// The traversal's cull margin, CULL_K * 3 * the largest primitive half-extent,
// read back from the last build once it is needed (jobs, tile lists).
static void resolve_margin(SceneObj& sc)
    {
    if (!sc.margin_pending)
        return;
    DeviceObj& D = *sc.dev;
    float r_max = 0.0f;
    HIPCHECK(hipMemcpyAsync(&r_max, sc.extent.p, sizeof(float), hipMemcpyDeviceToHost, D.stream));
    HIPCHECK(hipStreamSynchronize(D.stream));
    sc.cull_margin = CULL_K * (3.0f * r_max);
    sc.margin_pending = false;
    }

static void build_bvh(SceneObj& sc, bool refit_only)
    {
    DeviceObj& D = *sc.dev;
    hipStream_t s = D.stream;
    std::vector<uint32_t> prefix, slots;
    uint32_t N = 0;
    for (auto& kv : sc.geoms)
        {
        const GeometryObj& g = *kv.second;
        if (!g.valid || !g.enabled)
            continue;
        const uint32_t n = prim_count(g);
        if (n == 0)
            continue;
        prefix.push_back(N);
        slots.push_back(kv.first);
        N += n;
        }
    sc.n_prims = N;
    sc.n_active = (uint32_t)slots.size();
    sc.sphere_leaves = scene_types(sc) == T_SPHERES;
    if (N == 0)
        {
        sc.bvh_valid = true;
        return;
        }
    REQUIRE(N < (1u << 30), "too many primitives in one scene");
    sc.prefix.ensure(prefix.size());
    sc.slots.ensure(slots.size());
    HIPCHECK(hipMemcpyAsync(sc.prefix.p, prefix.data(), prefix.size() * 4, hipMemcpyHostToDevice, D.copy_stream));
    HIPCHECK(hipMemcpyAsync(sc.slots.p, slots.data(), slots.size() * 4, hipMemcpyHostToDevice, D.copy_stream));
    uploads_done(D);
    sc.blo.ensure(N);
    sc.bhi.ensure(N);
    const uint32_t na = (uint32_t)slots.size();
    D.launch("lbvh_bounds",
             [&]
             {
                 k_prim_bounds<<<blocks(N, 256), 256, 0, s>>>(sc.dgeoms.p, sc.prefix.p, sc.slots.p, na, N,
                                                                sc.blo.p, sc.bhi.p);
             });
    {
    // largest primitive half-extent -> the traversal's cull margin
    const uint32_t n_part = std::min<uint32_t>(blocks(N, 256), 1024);
    sc.extent.ensure(n_part + 1);
    D.launch("lbvh_extent",
             [&]
             {
                 k_extent_partial<<<n_part, 256, 0, s>>>(sc.blo.p, sc.bhi.p, N, sc.extent.p + 1);
                 k_extent_final<<<1, 256, 0, s>>>(sc.extent.p + 1, n_part, sc.extent.p);
             });
    sc.margin_pending = true;
    }
    sc.nodes.ensure(std::max<uint32_t>(N - 1, 1));
    sc.root_box.ensure(6);
    if (N == 1)
        {
        static const uint32_t one[2] = {0u, 1u};
        sc.idx4.ensure(2);
        sc.nodes4.ensure(1);
        HIPCHECK(hipMemcpyAsync(sc.idx4.p, one, sizeof(one), hipMemcpyHostToDevice, s));
        D.launch("lbvh_single",
                 [&]
                 {
                     k_single_leaf<<<1, 1, 0, s>>>(sc.blo.p, sc.bhi.p, sc.nodes.p);
                     k_bvh4_fill<<<1, 64, 0, s>>>(sc.nodes.p, 1, sc.idx4.p, sc.nodes4.p,
                                                   sc.sphere_leaves ? sc.dgeoms.p : nullptr);
                 });
        sc.bvh_valid = true;
        sc.builds++;
        return;
        }
    if (!refit_only)
        {
        // centroid bounds -> morton -> sort -> hierarchy
        BvhTemps& X = D.bvh_tmp;
        const uint32_t n_part = std::min<uint32_t>(blocks(N, 256), 1024);
        grow(X.partial, 6 * n_part, s);
        grow(X.cb, 6, s);
        D.launch("lbvh_centroid",
                 [&]
                 {
                     k_centroid_partial<<<n_part, 256, 0, s>>>(sc.blo.p, sc.bhi.p, N, X.partial.p);
                     k_centroid_final<<<1, 256, 0, s>>>(X.partial.p, n_part, X.cb.p);
                 });
        grow(X.keys, N, s);
        sc.vals.ensure(N);
        grow(X.keys2, N, s);
        grow(X.vals2, N, s);
        D.launch("lbvh_morton",
                 [&] { k_morton<<<blocks(N, 256), 256, 0, s>>>(sc.blo.p, sc.bhi.p, N, X.cb.p, X.keys.p, sc.vals.p); });
        // four 8-bit passes: the sorted data is back in keys/vals
        radix_sort_pairs(D, s, "lbvh_sort", X.keys.p, sc.vals.p, X.keys2.p, X.vals2.p, N, 0, 32, X.hist,
                         X.scan_tmp);
        sc.parent.ensure(2 * N - 1);
        D.launch("lbvh_karras",
                 [&]
                 {
                     k_karras<<<blocks(N - 1, 256), 256, 0, s>>>(X.keys.p, sc.vals.p, sc.blo.p, sc.bhi.p, (int)N,
                                                                 sc.nodes.p, sc.parent.p);
                 });
        // 4-wide topology: inner nodes at even depth, numbered in index order
        sc.idx4.ensure(N);
        sc.nodes4.ensure(N - 1);
        D.launch("lbvh_bvh4",
                 [&]
                 {
                     k_depth_even<<<blocks(N, 256), 256, 0, s>>>(sc.parent.p, N - 1, sc.idx4.p);
                     scan_exclusive(D, s, sc.idx4.p, N, X.scan_tmp);
                 });
        sc.builds++;
        }
    else
        sc.refits++;
    sc.flags.ensure(N - 1);
    HIPCHECK(hipMemsetAsync(sc.flags.p, 0, (N - 1) * sizeof(uint32_t), s));
    D.launch("lbvh_refit",
             [&]
             {
                 k_refit<<<blocks(N, 256), 256, 0, s>>>(sc.vals.p, sc.blo.p, sc.bhi.p, (int)N, sc.nodes.p,
                                                        sc.parent.p, sc.flags.p, sc.root_box.p);
             });
    D.launch("lbvh_bvh4",
             [&]
             {
                 k_bvh4_fill<<<blocks(N - 1, 256), 256, 0, s>>>(sc.nodes.p, N - 1, sc.idx4.p, sc.nodes4.p,
                                                                 sc.sphere_leaves ? sc.dgeoms.p : nullptr);
             });
    sc.bvh_valid = true;
    }

// Surface-area cost of the scene's current hierarchy (synchronizes the stream).
static float bvh_sah(SceneObj& sc)
    {
    if (sc.n_prims < 2)
        return 0.0f;
    DeviceObj& D = *sc.dev;
    const uint32_t n_nodes = sc.n_prims - 1;
    const uint32_t n_part = std::min<uint32_t>(blocks(n_nodes, 256), 1024);
    DBuf<float>& partial = D.bvh_tmp.partial;
    grow(partial, std::max<uint32_t>(6 * n_part, n_part + 1), D.stream);
    D.launch("lbvh_sah",
             [&]
             {
                 k_sah_partial<<<n_part, 256, 0, D.stream>>>(sc.nodes.p, n_nodes, partial.p + 1);
                 k_sah_final<<<1, 256, 0, D.stream>>>(partial.p + 1, n_part, partial.p);
             });
    float v = 0.0f;
    HIPCHECK(hipMemcpyAsync(&v, partial.p, sizeof(float), hipMemcpyDeviceToHost, D.stream));
    HIPCHECK(hipStreamSynchronize(D.stream));
    return v;
    }

// Refit-or-rebuild policy for a scene whose primitives moved: refit, and
// rebuild if the refit hierarchy's surface-area cost exceeds refit_limit times
// the cost right after the last build.
static void update_bvh(SceneObj& sc)
    {
    if (!sc.allow_refit)
        {
        build_bvh(sc, false);
        return;
        }
    build_bvh(sc, true);
    sc.sah_last = bvh_sah(sc);
    if (sc.sah_last > sc.refit_limit * sc.sah_built)
        {
        build_bvh(sc, false);
        sc.sah_built = sc.sah_last = bvh_sah(sc);
        }
    }

// Bring the device copy of the scene up to date: uploads, geometry table, BVH.
static void sync_scene(SceneObj& sc)
    {
    DeviceObj& D = *sc.dev;
    // uploads below may rewrite buffers an earlier build of this scene still reads
    if (sc.build_pending)
        HIPCHECK(hipStreamSynchronize(D.stream));
    sc.build_pending = false;
    bool any_dirty = false;
    std::vector<uint32_t> topo;
    for (auto& kv : sc.geoms)
        {
        GeometryObj& g = *kv.second;
        if (!g.valid)
            continue;
        if (g.dirty || g.shape_hi > g.shape_lo || g.color_hi > g.color_lo)
            {
            // colour-only changes leave the hierarchy as it is
            if (upload_geometry(g, D.copy_stream) && g.enabled)
                any_dirty = true;
            }
        if (g.enabled)
            {
            topo.push_back(kv.first);
            topo.push_back(prim_count(g));
            }
        }
    // geometry table (materials change without a rebuild)
    uint32_t n_slots = sc.geoms.empty() ? 1 : (sc.geoms.rbegin()->first + 1);
    std::vector<GeomDev> table(n_slots);
    std::memset((void*)table.data(), 0, table.size() * sizeof(GeomDev));
    for (auto& kv : sc.geoms)
        {
        GeometryObj& g = *kv.second;
        GeomDev& t = table[kv.first];
        t.type = g.type;
        t.n_prims = prim_count(g);
        t.n_aux = g.type == G_POLYHEDRON ? (uint32_t)(g.plane_n.size() / 3)
                                         : (g.type == G_MESH ? g.n_faces : (uint32_t)(g.verts.size() / 2));
        t.outline_width = g.outline_width;
        t.color_by_face = g.color_by_face;
        t.rounding = g.rounding;
        t.xy_radius = g.radius;
        t.mat = to_mat(g.mat);
        t.outline = to_mat(g.outline);
        t.a = g.a.p;
        t.b = g.b.p;
        t.color = g.dcolor.p;
        t.aux0 = g.aux0.p;
        t.aux1 = g.aux1.p;
        t.aux_color = g.daux_color.p;
        }
    sc.dgeoms.ensure(n_slots);
    HIPCHECK(hipMemcpyAsync(sc.dgeoms.p, table.data(), n_slots * sizeof(GeomDev), hipMemcpyHostToDevice,
                            D.copy_stream));
    uploads_done(D); // before the table goes away and before the device stream's kernels read it
    sc.n_slots = n_slots;
    const bool topo_same = sc.bvh_valid && topo == sc.topo;
    if (!sc.bvh_valid || !topo_same)
        {
        build_bvh(sc, false);
        sc.build_pending = true;
        if (sc.allow_refit)
            sc.sah_built = sc.sah_last = bvh_sah(sc);
        }
    else if (any_dirty)
        {
        update_bvh(sc);
        sc.build_pending = true;
        }
    sc.topo = topo;
    }

static SceneDev scene_dev(SceneObj& sc)
    {
    resolve_margin(sc);
    SceneDev d;
    d.nodes = sc.nodes.p;
    d.nodes4 = sc.nodes4.p;
    d.sphere_leaves = sc.sphere_leaves ? 1u : 0u;
    d.geoms = sc.dgeoms.p;
    d.root = sc.n_prims ? 0u : EMPTY;
    d.n_geoms = sc.n_slots;
    d.n_prims = sc.n_prims;
    d.cull_margin = sc.cull_margin;
    return d;
    }

// Camera constants, fresnel/common/Camera.h:92-159 (host float arithmetic)
static CamDev make_cam(const fh_user_camera& u)
    {
    CamDev c;
    v3 d = mk3(u.look_at.x - u.position.x, u.look_at.y - u.position.y, u.look_at.z - u.position.z);
    v3 vv = mk3(u.up.x, u.up.y, u.up.z);
    d = d * (1.0f / sqrtf(dot(d, d)));
    vv = vv * (1.0f / sqrtf(dot(vv, vv)));
    v3 uu = cross(d, vv);
    uu = uu * (1.0f / sqrtf(dot(uu, uu)));
    vv = cross(uu, d);
    vv = vv * (1.0f / sqrtf(dot(vv, vv)));
    c.u = uu;
    c.v = vv;
    c.w = -d;
    c.p = mk3(u.position.x, u.position.y, u.position.z);
    c.focal_d = u.focus_distance;
    c.a = u.f / u.f_stop;
    c.model = u.model;
    c.focal_h = (u.model == 0) ? u.h : u.h / u.f * c.focal_d;
    // p + (+-0) == p except for p = -0 and an offset of +0
    const bool neg_zero = (c.p.x == 0.0f && std::signbit(c.p.x)) || (c.p.y == 0.0f && std::signbit(c.p.y))
                          || (c.p.z == 0.0f && std::signbit(c.p.z));
    c.pinhole = (c.model == 1 && c.a == 0.0f && !neg_zero) ? 1 : 0;
    return c;
    }

// Lights in scene space (fresnel/common/Light.h:47-66) plus per-light constants.
static LightsDev make_lights(const fh_lights& in, const CamDev& cam)
    {
    LightsDev L;
    std::memset(&L, 0, sizeof(L));
    L.N = std::min<uint32_t>(in.N, 4);
    for (uint32_t i = 0; i < L.N; i++)
        {
        L.color[i] = mk3(in.color[i].r, in.color[i].g, in.color[i].b);
        v3 v = mk3(in.direction[i].x, in.direction[i].y, in.direction[i].z);
        v = v * (1.0f / sqrtf(dot(v, v)));
        L.dir[i] = v.x * cam.u + v.y * cam.v + v.z * cam.w;
        const float th = in.theta[i];
        L.theta[i] = th;
        L.cos_theta[i] = cosf(th);
        const float hs = (th > PI_F / 2.0f) ? PI_F / 2.0f : th;
        L.inv_sin[i] = 1.0f / sinf(hs);
        const float ht = (th > 0.99f * PI_F / 2.0f) ? 0.99f * PI_F / 2.0f : th;
        L.tan_clamped[i] = tanf(ht);
        }
    return L;
    }

// ====================================================================== tracer
enum TracerType
    {
    T_PATH,
    T_DIRECT
    };

// Per-tile candidate lists of a tracer's primary rays (kernels/tiles.hip),
// cached for the (scene, hierarchy version, camera, image size) they were
// built for.
struct TileLists
    {
    bool valid = false;        // lists usable for the key below
    uint64_t scene = 0;        // serial of the scene they were built for (0: none)
    uint64_t builds = ~0ull, refits = ~0ull;
    CamDev cam {};
    uint32_t W = 0, H = 0;
    uint32_t tiles_x = 0, tiles_y = 0;
    uint32_t group = 0;        // tile_group the lists were built with
    uint32_t n_entries = 0;
    bool decided = false;      // lists built for the key above, or known unusable
    uint64_t rays = 0;         // primary rays rendered with the key above while undecided
    DBuf<uint32_t> start, end; // entry range of each tile
    DBuf<uint4> entries;       // 2 per entry
    };

struct TracerObj
    {
    DeviceObj* dev = nullptr;
    TileLists tiles;
    TracerType type = T_PATH;
    uint32_t W = 0, H = 0;
    uint32_t seed = 0;
    uint32_t n_samples = 0;
    uint32_t L = 1;
    uint32_t aa_n = 8;
    bool highlight = false;
    fh_rgb hw_color {1, 0, 1};
    DBuf<float4> linear;
    DBuf<uchar4> srgb;
    DBuf<uint32_t> hist; // histogram counts (512 x 4)
    std::vector<float> h_linear;
    std::vector<uint8_t> h_srgb;
    // asynchronous readback (fh_tracer_output_async): a pinned copy of each
    // output, the event after its copy, and whether a render must still wait
    // for that copy before writing the output again
    PinnedBuf h_async[2];
    hipEvent_t async_ev[2] = {nullptr, nullptr};
    bool async_unordered[2] = {false, false};
    bool async_started[2] = {false, false};
    uint32_t async_w[2] = {0, 0}, async_h[2] = {0, 0};

    // make the render streams wait for outstanding output copies
    void order_after_copies(hipStream_t a, hipStream_t b)
        {
        for (int k = 0; k < 2; k++)
            if (async_unordered[k])
                {
                HIPCHECK(hipStreamWaitEvent(a, async_ev[k], 0));
                if (b)
                    HIPCHECK(hipStreamWaitEvent(b, async_ev[k], 0));
                async_unordered[k] = false;
                }
        }
    // wait until outstanding output copies are done (before the outputs are
    // cleared or freed)
    void finish_copies()
        {
        for (int k = 0; k < 2; k++)
            if (async_started[k])
                HIPCHECK(hipEventSynchronize(async_ev[k]));
        }

    void resize(uint32_t w, uint32_t h)
        {
        if (w == 0 || h == 0)
            throw Error(FH_ERR_INVALID, "Invalid dimensions");
        finish_copies();
        W = w;
        H = h;
        linear.release();
        srgb.release();
        linear.ensure((size_t)w * h);
        srgb.ensure((size_t)w * h);
        clear();
        h_linear.assign((size_t)w * h * 4, 0.0f);
        h_srgb.assign((size_t)w * h * 4, 0);
        }
    void clear()
        {
        finish_copies();
        HIPCHECK(hipMemsetAsync(linear.p, 0, (size_t)W * H * sizeof(float4), dev->stream));
        HIPCHECK(hipMemsetAsync(srgb.p, 0, (size_t)W * H * sizeof(uchar4), dev->stream));
        HIPCHECK(hipStreamSynchronize(dev->stream));
        }
    };

// Bring a tracer's per-tile candidate lists up to date for scene `sc` seen
// through `cam` (pinhole perspective or orthographic cameras only; a scene a
// perspective camera's plane cuts keeps the BVH search), for a render of
// `n_pri` primary rays.  Building the lists costs about as much as the BVH
// search loses on a few million primary rays, so they are built once the
// primary rays rendered with one scene, camera and image size reach
// D.tile_min_rays (this render's included); until then primary rays take the
// BVH search.  Either way the hits are the same.
static void update_tiles(DeviceObj& D, SceneObj& sc, TracerObj& tr, const CamDev& cam, uint64_t n_pri)
    {
    TileLists& T = tr.tiles;
    const bool pinhole = (cam.model == 1 && cam.a == 0.0f) || cam.model == 0;
    if (D.primary_tiles != 2 || !pinhole || sc.n_prims == 0 || tr.W < TILE || tr.H < TILE)
        {
        T.valid = false;
        T.decided = false;
        T.scene = 0;
        return;
        }
    if (T.scene == sc.serial && T.builds == sc.builds && T.refits == sc.refits && T.W == tr.W && T.H == tr.H
        && T.group == D.tile_group && std::memcmp(&T.cam, &cam, sizeof(CamDev)) == 0)
        {
        if (T.decided)
            return; // cached, usable or known unusable
        T.rays += n_pri;
        }
    else
        {
        T.scene = sc.serial;
        T.group = D.tile_group;
        T.builds = sc.builds;
        T.refits = sc.refits;
        T.cam = cam;
        T.W = tr.W;
        T.H = tr.H;
        T.valid = false;
        T.decided = false;
        T.rays = n_pri;
        }
    if (T.rays < D.tile_min_rays)
        return;
    T.decided = true;
    hipStream_t s = D.stream;
    const uint32_t N = sc.n_prims;
    TileCam c;
    c.p = cam.p;
    c.u = cam.u;
    c.v = cam.v;
    c.w = cam.w;
    c.scale = (cam.model == 1) ? cam.focal_d / cam.focal_h : 1.0f / cam.focal_h;
    c.W = (float)tr.W;
    c.H = (float)tr.H;
    c.model = cam.model;
    c.tiles_x = tr.W / TILE;
    c.tiles_y = tr.H / TILE;
    resolve_margin(sc);
    c.margin = sc.cull_margin;
    // orthographic ray origins lie on the image plane within this distance of p
    // (screen coordinates |xs| <= W / 2H + 2 / H, |ys| <= 1/2 + 2 / H)
    c.plane_r = 0.0f;
    if (cam.model == 0)
        {
        const double xs = 0.5 * tr.W / tr.H + 2.0 / tr.H, ys = 0.5 + 2.0 / tr.H;
        c.plane_r = (float)(cam.focal_h * std::sqrt(xs * xs + ys * ys) * (1.0 + 0x1p-10));
        }
    const bool spheres = scene_types(sc) == T_SPHERES;
    const uint32_t n_tiles = c.tiles_x * c.tiles_y;
    // elements: primitives, or (other scenes, tile_group >= 2) the hierarchy
    // cut of kernels/tiles.hip over the N - 1 inner nodes; one more count
    // slot holds the scan's total
    const bool cut = !spheres && D.tile_group >= 2 && N >= 2;
    const uint32_t n_inner = N - 1;
    const uint32_t n_el = cut ? 2 * n_inner + 2 : N + 1;
    TileTemps& X = D.tile_tmp;
    grow(X.counts, n_el, s);
    grow(X.flags, 1, s);
    uint32_t h[2] = {0, 0};
    D.launch("tiles_count",
             [&]
             {
                 k_fill_u32<<<1, 64, 0, s>>>(X.flags.p, 1, 0u);
                 if (cut)
                     k_tile_cut<false><<<blocks(n_inner + 1, 256), 256, 0, s>>>(
                         sc.nodes.p, n_inner, sc.blo.p, sc.bhi.p, sc.prefix.p, sc.slots.p, sc.n_active, c,
                         D.tile_group, X.counts.p, X.flags.p, nullptr, nullptr, nullptr, nullptr);
                 else if (spheres)
                     k_tile_count<true><<<blocks(N + 1, 256), 256, 0, s>>>(sc.blo.p, sc.bhi.p, sc.dgeoms.p, N, c,
                                                                           X.counts.p, X.flags.p);
                 else
                     k_tile_count<false><<<blocks(N + 1, 256), 256, 0, s>>>(sc.blo.p, sc.bhi.p, sc.dgeoms.p, N, c,
                                                                            X.counts.p, X.flags.p);
                 scan_exclusive(D, s, X.counts.p, n_el, X.scan_tmp);
             });
    HIPCHECK(hipMemcpyAsync(&h[0], X.counts.p + (n_el - 1), 4, hipMemcpyDeviceToHost, s));
    HIPCHECK(hipMemcpyAsync(&h[1], X.flags.p, 4, hipMemcpyDeviceToHost, s));
    HIPCHECK(hipStreamSynchronize(s));
    if (h[1])
        return; // the camera plane cuts a primitive's box
    const uint32_t M = h[0];
    T.start.ensure(n_tiles);
    T.end.ensure(n_tiles);
    if (M > 0)
        {
        for (auto* b : {&X.tile0, &X.q0, &X.ref0, &X.keysA, &X.valsA, &X.keysB, &X.valsB})
            grow(*b, M, s);
        grow(T.entries, 2 * (size_t)M, s);
        D.launch("tiles_emit",
                 [&]
                 {
                     if (cut)
                         k_tile_cut<true><<<blocks(n_inner + 1, 256), 256, 0, s>>>(
                             sc.nodes.p, n_inner, sc.blo.p, sc.bhi.p, sc.prefix.p, sc.slots.p, sc.n_active, c,
                             D.tile_group, X.counts.p, X.flags.p, X.tile0.p, X.q0.p, X.ref0.p, X.valsA.p);
                     else if (spheres)
                         k_tile_emit<true><<<blocks(N, 256), 256, 0, s>>>(sc.blo.p, sc.bhi.p, sc.dgeoms.p, N, c,
                                                                          X.counts.p, X.tile0.p, X.q0.p, X.ref0.p,
                                                                          X.valsA.p);
                     else
                         k_tile_emit<false><<<blocks(N, 256), 256, 0, s>>>(sc.blo.p, sc.bhi.p, sc.dgeoms.p, N, c,
                                                                           X.counts.p, X.tile0.p, X.q0.p, X.ref0.p,
                                                                           X.valsA.p);
                     HIPCHECK(hipMemcpyAsync(X.keysA.p, X.q0.p, (size_t)M * 4, hipMemcpyDeviceToDevice, s));
                 });
        // by bound (4 passes: the result is back in keysA/valsA), then stably by tile
        radix_sort_pairs(D, s, "tiles_sort", X.keysA.p, X.valsA.p, X.keysB.p, X.valsB.p, M, 0, 32, X.hist,
                         X.scan_tmp);
        uint32_t bits = 8;
        while (bits < 32 && (n_tiles - 1) >> bits)
            bits += 8;
        D.launch("tiles_sort", [&] { k_gather_u32<<<blocks(M, 256), 256, 0, s>>>(X.tile0.p, X.valsA.p, M, X.keysB.p); });
        const bool flipped = radix_sort_pairs(D, s, "tiles_sort", X.keysB.p, X.valsA.p, X.keysA.p, X.valsB.p, M, 0,
                                              bits, X.hist, X.scan_tmp);
        const uint32_t* kt = flipped ? X.keysA.p : X.keysB.p;
        const uint32_t* vt = flipped ? X.valsB.p : X.valsA.p;
        D.launch("tiles_bounds",
                 [&]
                 {
                     if (spheres)
                         k_tile_pack<true><<<blocks(M, 256), 256, 0, s>>>(vt, X.ref0.p, X.q0.p, sc.blo.p, sc.bhi.p,
                                                                          sc.dgeoms.p, nullptr, M, T.entries.p);
                     else
                         k_tile_pack<false><<<blocks(M, 256), 256, 0, s>>>(vt, X.ref0.p, X.q0.p, sc.blo.p, sc.bhi.p,
                                                                           sc.dgeoms.p, sc.nodes.p, M, T.entries.p);
                     k_fill_u32<<<blocks(n_tiles, 256), 256, 0, s>>>(T.start.p, n_tiles, 0u);
                     k_fill_u32<<<blocks(n_tiles, 256), 256, 0, s>>>(T.end.p, n_tiles, 0u);
                     k_tile_bounds<<<blocks(M, 256), 256, 0, s>>>(kt, M, T.start.p, T.end.p);
                 });
        }
    else
        D.launch("tiles_bounds",
                 [&]
                 {
                     k_fill_u32<<<blocks(n_tiles, 256), 256, 0, s>>>(T.start.p, n_tiles, 0u);
                     k_fill_u32<<<blocks(n_tiles, 256), 256, 0, s>>>(T.end.p, n_tiles, 0u);
                 });
    T.tiles_x = c.tiles_x;
    T.tiles_y = c.tiles_y;
    T.n_entries = M;
    T.valid = true;
    }

static JobDev make_job(SceneObj& sc, TracerObj& tr, const fh_user_camera& cam_user)
    {
    JobDev J;
    std::memset(&J, 0, sizeof(J));
    J.scene = scene_dev(sc);
    J.cam = make_cam(cam_user);
    J.lights = make_lights(sc.lights, J.cam);
    J.bg = mk3(sc.bg.r, sc.bg.g, sc.bg.b);
    J.bg_alpha = sc.bg_alpha;
    J.W = tr.W;
    J.H = tr.H;
    J.npix = tr.W * tr.H;
    J.xs_off = 0.5f * float(tr.W) / float(tr.H); // as fresnel/common/Camera.h:267 evaluates it per ray
    J.inv_h = 1.0 / (double)tr.H;
    J.seed = tr.seed;
    J.n_base = tr.n_samples;
    J.L = tr.L;
    J.highlight = tr.highlight ? 1u : 0u;
    J.hw_color = mk3(tr.hw_color.r, tr.hw_color.g, tr.hw_color.b);
    J.linear = tr.linear.p;
    J.srgb = tr.srgb.p;
    const TileLists& T = tr.tiles;
    if (T.valid && T.scene == sc.serial && T.builds == sc.builds && T.refits == sc.refits && T.W == tr.W
        && T.H == tr.H && std::memcmp(&T.cam, &J.cam, sizeof(CamDev)) == 0)
        {
        J.use_tiles = 1;
        J.tiles_x = T.tiles_x;
        J.tiles_y = T.tiles_y;
        J.tile_start = T.start.p;
        J.tile_end = T.end.p;
        J.tile_entries = T.entries.p;
        J.prim_lo = sc.blo.p;
        J.prim_hi = sc.bhi.p;
        }
    return J;
    }

// ====================================================================== path rendering
// Geometry types present (enabled) in a scene, as a bit mask.
static uint32_t scene_types(const SceneObj& sc)
    {
    uint32_t m = 0;
    for (auto& kv : sc.geoms)
        if (kv.second->valid && kv.second->enabled && prim_count(*kv.second) > 0)
            m |= 1u << kv.second->type;
    return m;
    }

// Slots per queue segment for a batch: every producer appends at most
// `per_unit` entries per unit, and segment s receives the units u with
// u % QSEG == s.
static uint32_t segment_capacity(uint64_t n_pri, uint64_t n_paths, uint32_t max_l, uint32_t primary_chunk,
                                 uint32_t trav_chunk)
    {
    auto bound = [](uint64_t n, uint64_t unit, uint64_t per_unit)
    {
        const uint64_t units = (n + unit - 1) / unit;
        return (units + QSEG - 1) / QSEG * per_unit;
    };
    uint64_t c = 1;
    c = std::max(c, bound(n_pri, primary_chunk, (uint64_t)primary_chunk * max_l));  // k_primary
    c = std::max(c, bound(n_pri, SHADE_BLOCK, (uint64_t)SHADE_BLOCK * max_l));      // k_shade_primary
    c = std::max(c, bound(n_paths, SHADE_BLOCK, SHADE_BLOCK));                       // k_shade
    c = std::max(c, bound(n_paths, trav_chunk, trav_chunk));                         // k_bounce
    REQUIRE(c * QSEG < (1ull << 32), "batch too large");
    return (uint32_t)c;
    }

// Workgroups of 256 rays a primary-kernel workgroup claims at once: those of
// D.primary_chunk, or fewer when the batch holds under 16 such claims per CU,
// so that small renders still spread over every CU (a 256 x 256 image at 8
// samples has 2,048 claims of 256 rays against 512 of 1,024).
static uint32_t primary_groups(const DeviceObj& D, uint64_t n_pri)
    {
    const uint64_t most = std::max<uint32_t>(1, D.primary_chunk / PRIMARY_BLOCK);
    const uint64_t per_claim_round = (uint64_t)std::max(1, D.n_cu) * 16 * PRIMARY_BLOCK;
    return (uint32_t)std::min<uint64_t>(most, std::max<uint64_t>(1, n_pri / per_claim_round));
    }

// Divisors of the primary-ray order (primary_order) once K is final.
static void set_order_divisors(JobDev& J)
    {
    J.div_chunk = make_fastdiv(64u * J.K);
    J.div_tiles = make_fastdiv(std::max(1u, J.W >> 3));
    J.div_w = make_fastdiv(std::max(1u, J.W));
    }

// Pinned staging bytes of one batch: its job table, pixel bases and step
// records (JobDev alignment kept for each part).
static size_t stage_bytes(size_t n_jobs)
    {
    auto up = [](size_t b) { return (b + 255) & ~(size_t)255; };
    return up(n_jobs * sizeof(JobDev)) + up(n_jobs * sizeof(uint32_t)) + up(2 * MAX_STEPS * sizeof(uint32_t))
           + up(2 * sizeof(unsigned long long));
    }

// One batch: every job advances by jobs[j].K camera samples, on one render
// scratch and stream.  start() enqueues the primary rays; enqueue_steps(B)
// enqueues B bounce steps, each a k_queue_view that turns the last
// producer's segment counters into the consumer view on the device, then the
// traversal and shading (bounce) and the tail kernel (finish), which act only
// when the view selects them; the last step finishes every path still
// queued.  So the host never waits for a count: a batch is enqueued in one
// go and its kernels find their work on the device.  accumulate() enqueues
// the Welford accumulation.  The job table and step records pass through
// `stage` (pinned host memory that must outlive the batch's copies).
struct BatchExec
    {
    DeviceObj& D;
    RenderScratch& R;
    hipStream_t s;
    std::vector<JobDev> jobs;
    uint32_t types;
    bool spheres;
    std::vector<uint32_t> pix_base;
    uint64_t n_pri = 0, n_paths = 0, n_pix = 0, n_rec = 0;
    uint32_t nj = 0, gstride = 0, seg_cap = 0, tail_threads = 0, rest_grid = 0;
    uint32_t pgroups = 1; // primary claim size in workgroups of rays (primary_groups)
    uint32_t n_steps = 0;
    int cur = 1;
    Queue Q[2];
    uint8_t* stage;
    uint32_t* h_steps = nullptr; // pinned copy of the step records (valid after the batch has run)
    unsigned long long* h_flags = nullptr; // pinned copy of the overflow flags after the batch

    BatchExec(DeviceObj& d, RenderScratch& r, hipStream_t st, std::vector<JobDev> j, uint32_t t, uint8_t* stg)
        : D(d), R(r), s(st), jobs(std::move(j)), types(t), spheres(t == T_SPHERES), stage(stg)
        {
        }

    template<class F> void launch(const char* nm, F f)
        {
        D.lstream = s;
        D.launch(nm, f);
        D.lstream = nullptr;
        }

    // zero n words with a kernel (no memset node if a caller captures the stream)
    void clear(uint32_t* p, uint32_t n)
        {
        k_fill_u32<<<blocks(n, 256), 256, 0, s>>>(p, n, 0u);
        HIPCHECK(hipGetLastError());
        }

    void traverse(const Queue& q, uint32_t count, const QueueView* view, bool primary)
        {
        TravArgs a {R.jobs.p, nj, q, count, view, R.hits.p, R.counters32.p, D.trav_chunk, R.gstack.p, gstride,
                    D.counters.p};
        D.lstream = s;
        dispatch_traverse(D, primary ? "traverse_primary" : "traverse", s, a, primary, spheres);
        D.lstream = nullptr;
        }

    void start()
        {
        for (auto& J : jobs)
            {
            set_order_divisors(J);
            J.ray_base = (uint32_t)n_pri;
            J.rec_base = n_rec;
            pix_base.push_back((uint32_t)n_pix);
            n_pri += (uint64_t)J.npix * J.K;
            n_paths += (uint64_t)J.npix * J.K * J.L;
            n_rec += (uint64_t)((J.npix + 63) / 64) * 64 * J.K * J.L; // whole 64-pixel chunks
            n_pix += J.npix;
            }
        REQUIRE(n_pri < (1ull << 31) && n_paths < (1ull << 31), "batch too large");
        for (const auto& J : jobs)
            REQUIRE(!spheres || J.scene.sphere_leaves || J.scene.root == EMPTY,
                    "internal error: spheres-only kernels on a hierarchy without sphere leaves");
        // primary rays are generated from their index; the queues hold paths
        uint32_t max_l = 1;
        for (const auto& J : jobs)
            max_l = std::max(max_l, J.L);
        pgroups = primary_groups(D, n_pri);
        seg_cap = segment_capacity(n_pri, n_paths, max_l, PRIMARY_BLOCK * pgroups, D.trav_chunk);
        const uint64_t cap = (uint64_t)QSEG * seg_cap;
        // growing a buffer frees it first: an earlier batch on this stream
        // (the only user of this scratch) must have finished with it
        gstride = D.trav_grid * TRAV_BLOCK; // max resident threads of any traversal variant
        tail_threads = blocks(D.tail_rays, TRAV_BLOCK) * TRAV_BLOCK;
        const bool fits = R.hits.cap >= std::max(n_pri, n_paths) && R.qo[0].cap >= cap && R.qo[1].cap >= cap
                          && R.records.cap >= n_rec && R.jobs.cap >= jobs.size() && R.pix_base.cap >= jobs.size()
                          && R.gstack.cap >= (size_t)(STACK_MAX - 8) * gstride
                          && R.tail_stack.cap >= (size_t)(STACK_MAX - TRAV_SS) * tail_threads && R.view.p && R.steps.p
                          && R.counters32.p && R.qseg.p;
        if (!fits)
            HIPCHECK(hipStreamSynchronize(s));
        R.hits.ensure(std::max(n_pri, n_paths));
        for (int q = 0; q < 2; q++)
            {
            R.qo[q].ensure(cap);
            R.qd[q].ensure(cap);
            R.qatt[q].ensure(cap);
            R.qmeta[q].ensure(cap);
            }
        R.records.ensure(n_rec);
        R.counters32.ensure(8 * QSEG_STRIDE);
        R.qseg.ensure(QSEG_FLAG + 1);
        R.jobs.ensure(jobs.size());
        R.pix_base.ensure(jobs.size());
        R.view.ensure(1);
        R.steps.ensure(2 * MAX_STEPS);
        // job table and pixel bases through pinned staging: the copies do not
        // wait for the stream, so enqueueing never blocks on earlier batches
        uint8_t* p = stage;
        std::memcpy(p, jobs.data(), jobs.size() * sizeof(JobDev));
        HIPCHECK(hipMemcpyAsync(R.jobs.p, p, jobs.size() * sizeof(JobDev), hipMemcpyHostToDevice, s));
        p += (jobs.size() * sizeof(JobDev) + 255) & ~(size_t)255;
        std::memcpy(p, pix_base.data(), pix_base.size() * 4);
        HIPCHECK(hipMemcpyAsync(R.pix_base.p, p, pix_base.size() * 4, hipMemcpyHostToDevice, s));
        p += (pix_base.size() * 4 + 255) & ~(size_t)255;
        h_steps = (uint32_t*)p;
        p += (2 * MAX_STEPS * sizeof(uint32_t) + 255) & ~(size_t)255;
        h_flags = (unsigned long long*)p;
        R.gstack.ensure((size_t)(STACK_MAX - 8) * gstride);
        // the tail kernel's grid-stride threads (at least tail_rays)
        R.tail_stack.ensure((size_t)(STACK_MAX - TRAV_SS) * tail_threads);
        for (int q = 0; q < 2; q++)
            {
            Q[q] = Queue {R.qo[q].p, R.qd[q].p, R.qatt[q].p, R.qmeta[q].p};
            Q[q].count = R.qseg.p;
            Q[q].seg_cap = seg_cap;
            }
        nj = (uint32_t)jobs.size();
        // resident grid of the grid-stride shading form (queried here, not
        // while a stream captures)
        rest_grid = D.grid_for(reinterpret_cast<const void*>(spheres ? k_shade<SHADE_BLOCK, T_SPHERES, true>
                                                                     : k_shade<SHADE_BLOCK, T_ALL, true>),
                               SHADE_BLOCK);
        }

    // Workgroups of the one-unit shading launch at bounce step b: the live
    // fraction recent renders had at this step with a quarter margin, rounded
    // up to an eighth of its power of two (so small drifts of the estimate
    // keep a captured graph's shape).
    uint32_t shade_grid(uint32_t b) const
        {
        const double frac = std::min(1.0, 1.25 * (double)D.live_frac[b]);
        uint64_t g = blocks((uint64_t)(frac * (double)n_paths) + 4096, SHADE_BLOCK);
        uint64_t q = 1;
        while ((q << 4) <= g)
            q <<= 1;
        g = (g + q - 1) / q * q;
        return (uint32_t)std::min<uint64_t>(g, blocks(n_paths, SHADE_BLOCK) + 1);
        }

    // Everything the kernels of enqueue_kernels(steps) depend on: equal keys
    // enqueue identical launches.
    std::vector<uint64_t> graph_key(uint32_t steps) const
        {
        std::vector<uint64_t> k = {(uint64_t)(uintptr_t)s, nj, n_pri, n_paths, n_pix, n_rec, seg_cap, gstride,
                                   tail_threads, types, steps, D.trav_chunk, pgroups, D.xcd_bands,
                                   D.fuse_primary, D.fuse_bounce, D.tail_rays, (uint64_t)D.stack_lds, D.trav_refill,
                                   D.trav_packet, D.count_on ? 1u : 0u, rest_grid};
        for (const void* p : {(const void*)R.qo[0].p, (const void*)R.qo[1].p, (const void*)R.qd[0].p,
                              (const void*)R.qd[1].p, (const void*)R.qatt[0].p, (const void*)R.qatt[1].p,
                              (const void*)R.qmeta[0].p, (const void*)R.qmeta[1].p, (const void*)R.hits.p,
                              (const void*)R.records.p, (const void*)R.counters32.p, (const void*)R.qseg.p,
                              (const void*)R.jobs.p, (const void*)R.pix_base.p, (const void*)R.view.p,
                              (const void*)R.steps.p, (const void*)R.gstack.p, (const void*)R.tail_stack.p,
                              (const void*)D.counters.p})
            k.push_back((uint64_t)(uintptr_t)p);
        for (uint32_t b = 0; b < steps; b++)
            k.push_back(shade_grid(b));
        return k;
        }

    // The batch's kernels: counter clears, primary rays, `steps` bounce steps.
    void enqueue_kernels(uint32_t steps)
        {
        const uint32_t groups = pgroups;
        clear(R.counters32.p, 8 * QSEG_STRIDE);
        clear(R.qseg.p, QSEG_FLAG + 1);
        if (D.fuse_primary == 2)
            {
            PrimaryArgs a {R.jobs.p, nj, (uint32_t)n_pri, R.counters32.p, groups, R.gstack.p, gstride, Q[1],
                           R.records.p, D.counters.p};
            D.lstream = s;
            dispatch_primary(D, s, a, spheres);
            D.lstream = nullptr;
            }
        else
            {
            traverse(Q[0], (uint32_t)n_pri, nullptr, true);
            launch("shade_primary",
                   [&]
                   {
                       auto* f = spheres ? k_shade_primary<SHADE_BLOCK, T_SPHERES> : k_shade_primary<SHADE_BLOCK, T_ALL>;
                       f<<<blocks(n_pri, SHADE_BLOCK), SHADE_BLOCK, 0, s>>>(R.jobs.p, nj, R.hits.p, (uint32_t)n_pri,
                                                                            Q[1], R.records.p);
                   });
            }
        cur = 1;
        enqueue_steps(steps);
        }

    // Enqueue `steps` bounce steps (1 <= steps <= MAX_STEPS); the last one
    // finishes every path that is still queued.
    void enqueue_steps(uint32_t steps)
        {
        auto* shade_one = spheres ? k_shade<SHADE_BLOCK, T_SPHERES, false> : k_shade<SHADE_BLOCK, T_ALL, false>;
        auto* shade_rest = spheres ? k_shade<SHADE_BLOCK, T_SPHERES, true> : k_shade<SHADE_BLOCK, T_ALL, true>;
        for (uint32_t b = 0; b < steps; b++)
            {
            const bool last = b + 1 == steps;
            const uint32_t* mode = h_steps + 2 * b + 1;
            launch("queue_view",
                   [&]
                   {
                       k_queue_view<<<1, 64, 0, s>>>(R.qseg.p, seg_cap, D.tail_rays, last ? 1u : 0u, R.view.p,
                                                     R.counters32.p, R.steps.p, b, D.counters.p);
                   });
            const Queue in = Q[cur], out = Q[1 - cur];
            D.tag_mode = mode;
            D.tag_acts_on = VIEW_BOUNCE;
            if (D.fuse_bounce == 2)
                {
                BounceArgs a {R.jobs.p, in, 0u, R.view.p, R.counters32.p, D.trav_chunk, R.gstack.p, gstride, out,
                              R.records.p, D.counters.p};
                D.lstream = s;
                dispatch_bounce(D, s, a, spheres);
                D.lstream = nullptr;
                }
            else
                {
                traverse(in, 0u, R.view.p, false);
                // one workgroup per 512 rays up to the expected count (the live
                // fraction recent renders had at this step, with margin), the
                // rest (if more are queued) by the grid-stride form
                const uint32_t one_grid = shade_grid(b);
                launch("shade",
                       [&]
                       {
                           shade_one<<<one_grid, SHADE_BLOCK, 0, s>>>(R.jobs.p, nj, in, 0u, R.view.p, 0u, R.hits.p,
                                                                     out, R.records.p, MAX_DEPTH, D.counters.p);
                       });
                D.tag_mode = nullptr;
                launch("shade_rest",
                       [&]
                       {
                           shade_rest<<<rest_grid, SHADE_BLOCK, 0, s>>>(R.jobs.p, nj, in, 0u, R.view.p, one_grid,
                                                                       R.hits.p, out, R.records.p, MAX_DEPTH,
                                                                       D.counters.p);
                       });
                D.tag_mode = mode;
                }
            D.tag_acts_on = VIEW_FINISH;
            launch("finish",
                   [&]
                   {
                       auto* f = spheres ? (D.count_on ? k_finish<TRAV_BLOCK, TRAV_SS, true, T_SPHERES>
                                                       : k_finish<TRAV_BLOCK, TRAV_SS, false, T_SPHERES>)
                                         : (D.count_on ? k_finish<TRAV_BLOCK, TRAV_SS, true, T_ALL>
                                                       : k_finish<TRAV_BLOCK, TRAV_SS, false, T_ALL>);
                       f<<<tail_threads / TRAV_BLOCK, TRAV_BLOCK, 0, s>>>(R.jobs.p, nj, in, 0u, R.view.p,
                                                                          R.records.p, MAX_DEPTH, R.tail_stack.p,
                                                                          tail_threads, D.counters.p);
                   });
            D.tag_mode = nullptr;
            cur = 1 - cur;
            }
        n_steps = steps;
        }

    // Copy the step records and the device's overflow flags (queue segment,
    // traversal stack) to the pinned staging, after the steps.
    void copy_steps()
        {
        HIPCHECK(hipMemcpyAsync(h_steps, R.steps.p, 2 * n_steps * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
        HIPCHECK(hipMemcpyAsync(h_flags, D.counters.p + CTR_QUEUE_OVERFLOW, 2 * sizeof(unsigned long long),
                                hipMemcpyDeviceToHost, s));
        }

    // The batch's kernels, launched or (D.graph == 2 and no per-launch timing
    // or counting) captured once per key into a HIP graph and replayed; then
    // the step records' copy.  The job table copies (start) and the
    // accumulation, which waits on another stream's event, stay outside the
    // graph, as no memset is inside it.
    void run(uint32_t steps)
        {
        if (D.graph != 2 || D.stats_on || D.count_on)
            {
            enqueue_kernels(steps);
            copy_steps();
            return;
            }
        const std::vector<uint64_t> key = graph_key(steps);
        hipGraphExec_t exec = nullptr;
        for (auto& g : D.graph_cache)
            if (g.key == key)
                {
                exec = g.exec;
                g.used = ++D.graph_clock;
                D.graph_replays++;
                break;
                }
        if (!exec)
            {
            HIPCHECK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
            try
                {
                enqueue_kernels(steps);
                }
            catch (...)
                {
                hipGraph_t g = nullptr;
                (void)hipStreamEndCapture(s, &g);
                if (g)
                    (void)hipGraphDestroy(g);
                throw;
                }
            hipGraph_t g = nullptr;
            HIPCHECK(hipStreamEndCapture(s, &g));
            const hipError_t e = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            HIPCHECK(e);
            D.graph_cache.push_back(DeviceObj::GraphEntry {key, exec, ++D.graph_clock});
            D.graph_captures++;
            }
        else
            n_steps = steps; // as enqueue_kernels sets it
        HIPCHECK(hipGraphLaunch(exec, s));
        copy_steps();
        }

    // Fold this batch's queued fraction at each bounce step into the device's
    // recent maximum (valid once the batch has run).
    void note_live(float* live) const
        {
        for (uint32_t b = 0; b < n_steps; b++)
            {
            const float f = n_paths ? (float)((double)h_steps[2 * b] / (double)n_paths) : 0.0f;
            live[b] = std::max(live[b], f);
            }
        }

    // Bounce steps this batch needed (its last BOUNCE step + the FINISH one);
    // valid once the batch has run.  A batch that reached its last step with
    // paths left needed more than it was given.
    uint32_t steps_needed() const
        {
        uint32_t need = 1;
        for (uint32_t b = 0; b < n_steps; b++)
            if (h_steps[2 * b + 1] != VIEW_IDLE)
                need = b + 1;
        if (need == n_steps && h_steps[2 * (n_steps - 1)] > D.tail_rays)
            need = n_steps + 2;
        return need;
        }

    // Enqueue the accumulation after `after` (the previous batch's, if any).
    void accumulate(hipEvent_t after, hipEvent_t done)
        {
        if (after)
            HIPCHECK(hipStreamWaitEvent(s, after, 0));
        launch("accumulate",
               [&]
               {
                   k_accumulate<<<blocks(n_pix, 256), 256, 0, s>>>(R.jobs.p, R.pix_base.p, nj, (uint32_t)n_pix,
                                                                   R.records.p);
               });
        if (done)
            HIPCHECK(hipEventRecord(done, s));
        }
    };

// Jobs whose scenes hold spheres only and the other jobs form two groups, so
// the spheres-only jobs use the spheres-only kernels (fewer registers).
static void split_groups(const std::vector<JobDev>& jobs, const std::vector<uint32_t>& types,
                         std::deque<std::pair<std::vector<JobDev>, uint32_t>>& out)
    {
    std::vector<JobDev> sph, mixed;
    uint32_t mixed_types = 0;
    for (size_t j = 0; j < jobs.size(); j++)
        {
        if ((types[j] & ~T_SPHERES) == 0)
            sph.push_back(jobs[j]);
        else
            {
            mixed.push_back(jobs[j]);
            mixed_types |= types[j];
            }
        }
    if (!sph.empty())
        out.emplace_back(std::move(sph), T_SPHERES);
    if (!mixed.empty())
        out.emplace_back(std::move(mixed), mixed_types);
    }

// Next estimate of the bounce steps per batch from the steps the batches of
// a render needed: their maximum plus one spare step, and never more than
// one below the previous estimate (the depth of paths depends on the scene).
static void update_steps_est(DeviceObj& D, uint32_t needed_max)
    {
    const uint32_t est = std::max(needed_max + 1, D.steps_est > 1 ? D.steps_est - 1 : 1u);
    D.steps_est = std::min<uint32_t>(std::max<uint32_t>(est, 2), MAX_STEPS);
    }

// Next live fractions per bounce step: the render's maxima, or 3/4 of the
// previous estimate when that is larger (it decays over renders of shallower
// scenes).
static void update_live_frac(DeviceObj& D, const float* live)
    {
    for (uint32_t b = 0; b < MAX_STEPS; b++)
        D.live_frac[b] = std::max(live[b], 0.75f * D.live_frac[b]);
    }

// Batches (job lists with their geometry-type sets) run alternating between
// the device's two render scratches and streams when D.inflight is 2: each
// batch is enqueued whole (primary rays, D.steps_est bounce steps,
// accumulation) without waiting for the device, so one batch's kernels
// overlap the other's.  Accumulations are chained in batch order (each waits
// for the previous one's event), so every pixel's samples are accumulated in
// order.  A batch's render stream waits, by event, for the work queued on the
// device stream before it (scene builds, tile lists, output clears), so the
// device stream can go on building the scenes of later batches meanwhile.
// finish() returns with both render streams idle.
struct InflightRun
    {
    DeviceObj& D;
    std::vector<std::unique_ptr<BatchExec>> batches;
    hipEvent_t last_acc = nullptr;
    int slot = 0;
    size_t off = 0;
    uint32_t steps = 0;

    // `stage` bytes of pinned staging for the job tables of every batch to come
    InflightRun(DeviceObj& d, size_t stage) : D(d)
        {
        HIPCHECK(hipStreamSynchronize(D.stream2));
        HIPCHECK(hipStreamSynchronize(D.stream3));
        // both streams are idle: the least recently used graphs beyond 32 can go
        while (D.graph_cache.size() > 32)
            {
            auto lru = std::min_element(D.graph_cache.begin(), D.graph_cache.end(),
                                        [](const DeviceObj::GraphEntry& a, const DeviceObj::GraphEntry& b)
                                        { return a.used < b.used; });
            (void)hipGraphExecDestroy(lru->exec);
            D.graph_cache.erase(lru);
            }
        D.stage.ensure(stage);
        steps = D.steps_est;
        }

    void add(std::vector<JobDev> jobs, uint32_t types)
        {
        const size_t n_jobs = jobs.size();
        hipStream_t s = slot ? D.stream2 : D.stream3;
        HIPCHECK(hipEventRecord(D.build_ev, D.stream));
        HIPCHECK(hipStreamWaitEvent(s, D.build_ev, 0));
        batches.emplace_back(new BatchExec(D, slot ? D.rs2 : D.rs, s, std::move(jobs), types, D.stage.p + off));
        off += stage_bytes(n_jobs);
        BatchExec& e = *batches.back();
        e.start();
        e.run(steps);
        e.accumulate(last_acc, D.acc_done[slot]);
        last_acc = D.acc_done[slot];
        if (D.inflight == 2)
            slot = 1 - slot;
        }

    void finish()
        {
        HIPCHECK(hipStreamSynchronize(D.stream3));
        HIPCHECK(hipStreamSynchronize(D.stream2));
        uint32_t need = 1;
        float live[MAX_STEPS] = {};
        for (const auto& e : batches)
            {
            need = std::max(need, e->steps_needed());
            e->note_live(live);
            }
        update_steps_est(D, need);
        update_live_frac(D, live);
        }
    };

static void run_inflight(DeviceObj& D, std::deque<std::pair<std::vector<JobDev>, uint32_t>>& pending)
    {
    size_t bytes = 0;
    for (const auto& b : pending)
        bytes += stage_bytes(b.first.size());
    InflightRun run(D, bytes);
    while (!pending.empty())
        {
        run.add(std::move(pending.front().first), pending.front().second);
        pending.pop_front();
        }
    run.finish();
    }

// Render `samples` Path passes for each (scene, tracer) pair.
static void render_path(DeviceObj& D, const std::vector<SceneObj*>& scenes, const std::vector<TracerObj*>& tracers,
                        uint32_t samples)
    {
    std::lock_guard<std::mutex> lock(D.mu);
    D.begin_stats();
    for (size_t k = 0; k < scenes.size(); k++)
        {
        REQUIRE(scenes[k]->dev == tracers[k]->dev, "Scene and Tracer devices do not match");
        REQUIRE(tracers[k]->type == T_PATH, "not a Path tracer");
        }
    uint64_t per_sample = 0;
    for (auto* tr : tracers)
        per_sample += (uint64_t)tr->W * tr->H * tr->L;
    uint32_t K = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(samples, D.batch_paths / std::max<uint64_t>(per_sample, 1)));
    std::vector<uint32_t> types;
    for (auto* sc : scenes)
        types.push_back(scene_types(*sc));
    // Several pairs that take more than one batch: batches of whole pairs (all
    // samples of each, pairs in order), so the first batch starts once its own
    // scenes and tile lists are built, and the device stream builds the later
    // pairs' while earlier batches render.  Pairs whose samples do not fit one
    // batch take the samples-first split below.
    std::vector<std::pair<size_t, size_t>> chunks; // [first, last) pairs per batch
    bool by_pairs = scenes.size() > 1 && K < samples;
    if (by_pairs)
        {
        size_t first = 0;
        uint64_t paths = 0;
        for (size_t j = 0; j < scenes.size() && by_pairs; j++)
            {
            const uint64_t pj = (uint64_t)tracers[j]->W * tracers[j]->H * tracers[j]->L * samples;
            by_pairs = pj <= D.batch_paths;
            if (paths + pj > D.batch_paths)
                {
                chunks.push_back({first, j});
                first = j;
                paths = 0;
                }
            paths += pj;
            }
        chunks.push_back({first, scenes.size()});
        }
    if (by_pairs)
        {
        size_t bytes = 0;
        for (const auto& c : chunks)
            {
            size_t n_sph = 0, n_mixed = 0;
            for (size_t j = c.first; j < c.second; j++)
                ((types[j] & ~T_SPHERES) == 0 ? n_sph : n_mixed)++;
            bytes += (n_sph ? stage_bytes(n_sph) : 0) + (n_mixed ? stage_bytes(n_mixed) : 0);
            }
        InflightRun run(D, bytes);
        std::vector<SceneObj*> synced; // a scene appearing twice is synchronized once
        for (const auto& c : chunks)
            {
            for (size_t j = c.first; j < c.second; j++)
                if (std::find(synced.begin(), synced.end(), scenes[j]) == synced.end())
                    {
                    sync_scene(*scenes[j]);
                    synced.push_back(scenes[j]);
                    }
            std::vector<JobDev> jobs;
            std::vector<uint32_t> ty;
            for (size_t j = c.first; j < c.second; j++)
                {
                update_tiles(D, *scenes[j], *tracers[j], make_cam(scenes[j]->camera),
                             (uint64_t)tracers[j]->W * tracers[j]->H * samples);
                // a readback of an output still in flight finishes before the accumulation
                tracers[j]->order_after_copies(D.stream3, D.stream2);
                JobDev J = make_job(*scenes[j], *tracers[j], scenes[j]->camera);
                J.K = samples;
                J.n_base = tracers[j]->n_samples;
                jobs.push_back(J);
                ty.push_back(types[j]);
                }
            std::deque<std::pair<std::vector<JobDev>, uint32_t>> groups;
            split_groups(jobs, ty, groups);
            for (auto& g : groups)
                run.add(std::move(g.first), g.second);
            }
        run.finish();
        }
    else
        {
        // a scene appearing twice is synchronized once
        std::vector<SceneObj*> uniq;
        for (auto* sc : scenes)
            if (std::find(uniq.begin(), uniq.end(), sc) == uniq.end())
                uniq.push_back(sc);
        for (auto* sc : uniq)
            sync_scene(*sc);
        for (size_t k = 0; k < scenes.size(); k++)
            update_tiles(D, *scenes[k], *tracers[k], make_cam(scenes[k]->camera),
                         (uint64_t)tracers[k]->W * tracers[k]->H * samples);
        // a readback of an output still in flight finishes before the accumulation
        for (auto* tr : tracers)
            tr->order_after_copies(D.stream3, D.stream2);
        std::deque<std::pair<std::vector<JobDev>, uint32_t>> groups;
        for (uint32_t done = 0; done < samples;)
            {
            const uint32_t k = std::min(K, samples - done);
            std::vector<JobDev> jobs;
            for (size_t j = 0; j < scenes.size(); j++)
                {
                JobDev J = make_job(*scenes[j], *tracers[j], scenes[j]->camera);
                J.K = k;
                J.n_base = tracers[j]->n_samples + done;
                jobs.push_back(J);
                }
            split_groups(jobs, types, groups);
            done += k;
            }
        run_inflight(D, groups);
        }
    for (auto* tr : tracers)
        tr->n_samples += samples;
    D.end_stats();
    if (D.last_counters[CTR_OVERFLOW])
        throw Error(FH_ERR_STATE, "traversal stack overflow");
    }

static void render_preview(DeviceObj& D, SceneObj& sc, TracerObj& tr)
    {
    std::lock_guard<std::mutex> lock(D.mu);
    REQUIRE(sc.dev == tr.dev, "Scene and Tracer devices do not match");
    D.begin_stats();
    sync_scene(sc);
    // direct tracers do not support depth of field (fresnel/cpu/TracerDirect.cc:38-40)
    fh_user_camera cam = sc.camera;
    cam.f_stop = __builtin_huge_valf();
    JobDev J = make_job(sc, tr, cam);
    J.K = 1;
    tr.order_after_copies(D.stream, nullptr);
    RenderScratch& R = D.rs;
    R.jobs.ensure(1);
    HIPCHECK(hipMemcpyAsync(R.jobs.p, &J, sizeof(J), hipMemcpyHostToDevice, D.stream));
    D.launch("preview", [&] { k_preview<<<blocks(J.npix, 64), 64, 0, D.stream>>>(R.jobs.p, tr.aa_n); });
    D.end_stats();
    }

// ====================================================================== serving
static uint32_t prim_bytes_of(uint32_t type)
    {
    return type == G_SPHERE ? 16u : 32u;
    }
static uint64_t scene_prims(const SceneObj& sc, uint32_t& max_prim_bytes)
    {
    uint64_t n = 0;
    max_prim_bytes = 16;
    for (auto& kv : sc.geoms)
        if (kv.second->valid && kv.second->enabled)
            {
            n += prim_count(*kv.second);
            max_prim_bytes = std::max(max_prim_bytes, prim_bytes_of(kv.second->type));
            }
    return n;
    }

struct RequestObj
    {
    uint64_t id = 0;
    SceneObj* scene = nullptr;
    TracerObj* tracer = nullptr;
    uint32_t samples = 0, done = 0, L = 1;
    uint32_t scheduled = 0; // samples planned into batches (done <= scheduled <= samples)
    double t_arrival = 0, t_admit = -1, t_done = -1;
    uint64_t footprint = 0;
    int state = 0;      // 0 queued, 1 active, 2 done, 3 rejected, 4 failed
    bool built = false; // scene synchronized (uploads, hierarchy) since admission
    };

// A batch of the server: its requests and the samples it adds.  Planned
// batches wait (stream -1) until one of the server's two streams has fewer
// than SERVER_DEPTH batches queued; a started batch is enqueued whole on that
// stream (which owns one render scratch) and is done when its `done` event
// (recorded after its accumulation) has completed.
constexpr int SERVER_DEPTH = 2;
struct ServerBatch
    {
    std::unique_ptr<BatchExec> e;
    std::unique_ptr<PinnedBuf> stage; // job table, step records, flags of this batch
    hipEvent_t done = nullptr;
    int stream = -1; // server stream (and render scratch), -1 while waiting
    std::vector<JobDev> jobs;
    uint32_t types = 0;
    std::vector<std::pair<uint64_t, uint32_t>> adds; // (request, samples)
    uint64_t serial = 0;
    int kind() const
        {
        return types == T_SPHERES ? 0 : 1;
        }
    };

// Request states (fh_server_poll)
constexpr int REQ_QUEUED = 0, REQ_ACTIVE = 1, REQ_DONE = 2, REQ_REJECTED = 3, REQ_FAILED = 4;

struct ServerObj
    {
    DeviceObj* dev = nullptr;
    uint64_t budget = 0, in_use = 0, target_paths = 256ull << 20;
    std::deque<PolicyItem> queue;
    std::vector<uint64_t> active; // admission order
    std::map<uint64_t, RequestObj> reqs;
    uint64_t next_id = 1;
    std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
    uint64_t iterations = 0;
    // batches in plan order (front: oldest) on the server's two streams and
    // render scratches, up to SERVER_DEPTH per stream; accumulations chained
    // by events
    RenderScratch rs[2];
    bool presized[2] = {false, false}; // rs[k] sized for target_paths
    hipStream_t st[2] = {nullptr, nullptr};
    std::vector<hipEvent_t> free_events;
    std::vector<hipEvent_t> all_events;
    std::vector<std::unique_ptr<PinnedBuf>> stage_pool; // staging of finished batches, reused
    std::deque<ServerBatch> batches;
    // last accumulation enqueued per batch kind (spheres-only / mixed): a
    // request's batches are all of one kind, so only batches of one kind
    // share tracers and need ordered accumulation
    hipEvent_t last_acc[2] = {nullptr, nullptr};
    uint64_t last_acc_batch[2] = {0, 0};
    uint64_t next_serial = 1;
    bool draining = false; // destruction: finish started batches, plan and start no more
    double clock(double now) const
        {
        if (now >= 0)
            return now;
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        }
    };

// Device memory the server cannot give to requests: its own two render
// scratches sized for target_paths, and whatever render scratch the device
// holds for Path.sample (allocated by earlier renders).
static uint64_t server_reserved(const ServerObj& S)
    {
    const DeviceObj& D = *S.dev;
    const uint64_t fixed = (uint64_t)(STACK_MAX - 8) * D.trav_grid * TRAV_BLOCK * sizeof(uint2)
                           + (uint64_t)(STACK_MAX - TRAV_SS) * blocks(D.tail_rays, TRAV_BLOCK) * TRAV_BLOCK
                                 * sizeof(uint2);
    uint64_t device_scratch = 0;
    for (const RenderScratch* R : {&D.rs, &D.rs2})
        {
        for (int q = 0; q < 2; q++)
            device_scratch += R->qo[q].bytes() + R->qd[q].bytes() + R->qatt[q].bytes() + R->qmeta[q].bytes();
        device_scratch += R->hits.bytes() + R->records.bytes() + R->gstack.bytes() + R->tail_stack.bytes();
        }
    return 2 * scratch_bytes(S.target_paths, fixed) + device_scratch;
    }

// Free a render scratch's device memory (its streams must be idle).
static void release_scratch(RenderScratch& R)
    {
    for (int q = 0; q < 2; q++)
        {
        R.qo[q].release();
        R.qd[q].release();
        R.qatt[q].release();
        R.qmeta[q].release();
        }
    R.hits.release();
    R.records.release();
    R.gstack.release();
    R.tail_stack.release();
    }

// Size a render scratch once for batches of up to `paths` light-sample paths
// (queues with the segment rounding of segment_capacity, hits, records), so
// batches of varying size up to that never regrow it (regrowth frees and
// reallocates tens of GB and waits for the stream).  Its stream must be idle.
static void presize_scratch(const DeviceObj& D, RenderScratch& R, uint64_t paths)
    {
    const uint64_t slack = (uint64_t)QSEG * std::max<uint64_t>(D.primary_chunk, SHADE_BLOCK) * 64;
    const uint64_t n = paths + slack;
    for (int q = 0; q < 2; q++)
        {
        R.qo[q].ensure(n);
        R.qd[q].ensure(n);
        R.qatt[q].ensure(n);
        R.qmeta[q].ensure(n);
        }
    R.hits.ensure(n);
    R.records.ensure(n);
    }

// Largest light-sample path count of one batch (the queue indices are 31-bit).
constexpr uint64_t MAX_BATCH_PATHS = (1ull << 31) - 1;

// Plan the next server step: every active request with unplanned samples
// gets its share of target_paths (policy_allocate), trimmed so a batch stays
// within MAX_BATCH_PATHS paths; the step's spheres-only and mixed jobs become
// one or two batches.
static void server_plan(ServerObj& S)
    {
    DeviceObj& D = *S.dev;
    std::vector<PolicyItem> items;
    std::vector<uint64_t> ids;
    for (uint64_t id : S.active)
        {
        const RequestObj& r = S.reqs[id];
        if (r.built && r.scheduled < r.samples)
            {
            items.push_back(PolicyItem {id, r.footprint, (uint64_t)r.tracer->W * r.tracer->H * r.L,
                                        r.samples - r.scheduled});
            ids.push_back(id);
            }
        }
    if (items.empty())
        return;
    const std::vector<uint32_t> k = policy_allocate(items, S.target_paths);
    std::vector<JobDev> sph, mixed;
    std::vector<std::pair<uint64_t, uint32_t>> add_sph, add_mixed;
    uint32_t mixed_types = 0;
    uint64_t paths[2] = {0, 0}; // per batch kind
    for (size_t i = 0; i < ids.size(); i++)
        {
        if (k[i] == 0)
            continue;
        RequestObj& r = S.reqs[ids[i]];
        const uint32_t t = scene_types(*r.scene);
        const int kd = (t & ~T_SPHERES) == 0 ? 0 : 1;
        // fewer samples (or none this step) when the batch would exceed its path limit
        const uint64_t pps = items[i].paths_per_sample;
        const uint32_t fit = (uint32_t)std::min<uint64_t>(k[i], (MAX_BATCH_PATHS - paths[kd]) / pps);
        if (fit == 0)
            continue;
        paths[kd] += fit * pps;
        r.tracer->finish_copies(); // its accumulation may start on either stream
        update_tiles(D, *r.scene, *r.tracer, make_cam(r.scene->camera), (uint64_t)r.tracer->W * r.tracer->H * fit);
        JobDev J = make_job(*r.scene, *r.tracer, r.scene->camera);
        J.K = fit;
        J.n_base = r.tracer->n_samples; // samples planned before this batch
        r.tracer->n_samples += fit;
        r.scheduled += fit;
        if (kd == 0)
            {
            sph.push_back(J);
            add_sph.push_back({ids[i], fit});
            }
        else
            {
            mixed.push_back(J);
            add_mixed.push_back({ids[i], fit});
            mixed_types |= t;
            }
        }
    if (!sph.empty())
        {
        ServerBatch b;
        b.jobs = std::move(sph);
        b.types = T_SPHERES;
        b.adds = std::move(add_sph);
        b.serial = S.next_serial++;
        S.batches.push_back(std::move(b));
        }
    if (!mixed.empty())
        {
        ServerBatch b;
        b.jobs = std::move(mixed);
        b.types = mixed_types;
        b.adds = std::move(add_mixed);
        b.serial = S.next_serial++;
        S.batches.push_back(std::move(b));
        }
    S.iterations++;
    }

// Retire the requests whose planned samples have all been accumulated.
static void server_retire(ServerObj& S, double now)
    {
    std::vector<uint64_t> still;
    for (uint64_t id : S.active)
        {
        RequestObj& r = S.reqs[id];
        if (r.done >= r.samples)
            {
            r.state = REQ_DONE;
            r.t_done = S.clock(now);
            S.in_use -= r.footprint;
            r.tracer->L = 1; // as Path.sample leaves it (fresnel/tracer.py:259-261)
            }
        else
            still.push_back(id);
        }
    S.active.swap(still);
    }

// A request that cannot complete (a batch holding some of its samples failed
// to start): it leaves the active set and frees its footprint.
static void server_fail(ServerObj& S, uint64_t id, double now)
    {
    RequestObj& r = S.reqs[id];
    if (r.state != REQ_ACTIVE)
        return;
    r.state = REQ_FAILED;
    r.t_done = S.clock(now);
    S.in_use -= r.footprint;
    r.tracer->L = 1;
    S.active.erase(std::remove(S.active.begin(), S.active.end(), id), S.active.end());
    }

// Finish a started batch whose accumulation has completed: check the overflow
// flags copied after it, count its samples as done, free its event.
static void server_finish_batch(ServerObj& S, size_t i)
    {
    DeviceObj& D = *S.dev;
    ServerBatch& b = S.batches[i];
    update_steps_est(D, b.e->steps_needed());
    float live[MAX_STEPS] = {};
    b.e->note_live(live);
    update_live_frac(D, live);
    const unsigned long long q_flag = b.e->h_flags[0], s_flag = b.e->h_flags[1];
    for (auto& a : b.adds)
        S.reqs[a.first].done += a.second;
    const int kd = b.kind();
    if (S.last_acc_batch[kd] == b.serial)
        S.last_acc[kd] = nullptr; // finished: later batches of its kind need not wait
    S.free_events.push_back(b.done);
    S.stage_pool.push_back(std::move(b.stage)); // its copies have completed
    S.batches.erase(S.batches.begin() + (std::ptrdiff_t)i);
    if (q_flag)
        throw Error(FH_ERR_STATE, "internal error: a ray queue segment overflowed");
    if (s_flag)
        throw Error(FH_ERR_STATE, "traversal stack overflow");
    }

// One server iteration: admit; retire started batches whose accumulation has
// completed; plan a step when no planned batch is waiting; start waiting
// batches on free slots (each enqueued whole); when both slots are busy and
// nothing could start, wait for the oldest started batch.  Work stays in
// flight between calls.
static void server_step(ServerObj& S, double now, uint32_t& n_active, uint32_t& n_queued)
    {
    DeviceObj& D = *S.dev;
    std::lock_guard<std::mutex> lock(D.mu);
    const bool stats = D.stats_on; // per-launch timing is for Path.sample renders only
    D.stats_on = false;
    struct Restore
        {
        DeviceObj& D;
        bool v;
        ~Restore()
            {
            D.stats_on = v;
            }
        } restore {D, stats};
    std::vector<PolicyItem> admitted, rejected;
    uint64_t reserved = server_reserved(S);
    if (!S.queue.empty() && S.in_use + S.queue.front().footprint + reserved > S.budget)
        {
        // the device's render scratch for Path.sample is idle while the server
        // holds the device lock: release it (the next render allocates again)
        release_scratch(D.rs);
        release_scratch(D.rs2);
        reserved = server_reserved(S);
        }
    const uint64_t room = S.budget > reserved ? S.budget - reserved : 0;
    policy_admit(S.queue, room, S.in_use, admitted, rejected);
    for (auto& it : rejected)
        {
        RequestObj& r = S.reqs[it.id];
        r.state = REQ_REJECTED;
        r.t_done = S.clock(now);
        }
    for (auto& it : admitted)
        {
        RequestObj& r = S.reqs[it.id];
        r.state = REQ_ACTIVE;
        r.t_admit = S.clock(now);
        S.active.push_back(it.id);
        }
    // Scenes of admitted requests are built (uploads, hierarchy) a few per
    // step: with batches running, for about 5 ms of host time, so the device
    // keeps tracing earlier requests meanwhile and the next batch holds the
    // requests built meanwhile (96 small requests as a burst: 21-25 steps and
    // 105-119 ms with 2 ms, 11-13 steps and 92-108 ms with 5 ms, 9-10 steps and
    // 95-109 ms with 10 ms); with none running, only until one request can be
    // planned.
    {
    bool running = false;
    for (auto& b : S.batches)
        running = running || b.stream >= 0;
    const auto t_build = std::chrono::steady_clock::now();
    for (uint64_t id : std::vector<uint64_t>(S.active))
        {
        RequestObj& r = S.reqs[id];
        if (r.built)
            continue;
        try
            {
            sync_scene(*r.scene);
            }
        catch (...)
            {
            server_fail(S, id, now);
            throw;
            }
        r.built = true;
        const double spent = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_build).count();
        if (!running || spent > 5e-3)
            break;
        }
    }
    // the device's two render streams: a stream of its own for each batch
    // slot could share a hardware queue with the device stream (four queues
    // per process), and the scene builds queued there would then wait for
    // the batches in order (a burst of 96 small requests: 72-85 ms against
    // 95-109 ms)
    S.st[0] = D.stream3;
    S.st[1] = D.stream2;
    if (S.draining)
        while (!S.batches.empty() && S.batches.back().stream < 0)
            S.batches.pop_back();
    // retire started batches that have completed
    for (size_t i = 0; i < S.batches.size();)
        {
        ServerBatch& b = S.batches[i];
        if (b.stream >= 0 && hipEventQuery(b.done) == hipSuccess)
            server_finish_batch(S, i);
        else
            i++;
        }
    bool waiting = false;
    for (auto& b : S.batches)
        waiting = waiting || b.stream < 0;
    if (!waiting && !S.draining)
        server_plan(S);
    // start waiting batches on streams with room (scene, tile and output work
    // on the device stream completes first)
    int queued[2] = {0, 0};
    for (auto& b : S.batches)
        if (b.stream >= 0)
            queued[b.stream]++;
    bool synced = false, started = false;
    for (size_t i = 0; i < S.batches.size(); i++)
        {
        ServerBatch& b = S.batches[i];
        if (b.stream >= 0)
            continue;
        const int k = queued[0] <= queued[1] ? 0 : 1;
        if (queued[k] >= SERVER_DEPTH)
            break;
        if (!synced)
            {
            // the batches wait, by event, for the scene builds and tile lists
            // queued on the device stream; the host does not
            HIPCHECK(hipEventRecord(D.build_ev, D.stream));
            synced = true;
            }
        try
            {
            HIPCHECK(hipStreamWaitEvent(S.st[k], D.build_ev, 0));
            if (S.free_events.empty())
                {
                hipEvent_t ev;
                HIPCHECK(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
                S.all_events.push_back(ev);
                S.free_events.push_back(ev);
                }
            if (!S.presized[k])
                {
                HIPCHECK(hipStreamSynchronize(S.st[k]));
                presize_scratch(D, S.rs[k], S.target_paths);
                S.presized[k] = true;
                }
            if (S.stage_pool.empty())
                b.stage.reset(new PinnedBuf());
            else
                {
                b.stage = std::move(S.stage_pool.back());
                S.stage_pool.pop_back();
                }
            b.stage->ensure(stage_bytes(b.jobs.size()));
            b.e.reset(new BatchExec(D, S.rs[k], S.st[k], b.jobs, b.types, b.stage->p));
            b.e->start();
            b.e->enqueue_kernels(D.steps_est);
            b.e->copy_steps();
            b.done = S.free_events.back();
            const int kd = b.kind();
            b.e->accumulate(S.last_acc[kd], b.done);
            S.free_events.pop_back();
            S.last_acc[kd] = b.done;
            S.last_acc_batch[kd] = b.serial;
            }
        catch (...)
            {
            // whatever was enqueued drains; the batch's requests cannot complete
            (void)hipStreamSynchronize(S.st[k]);
            for (auto& a : b.adds)
                server_fail(S, a.first, now);
            S.batches.erase(S.batches.begin() + (std::ptrdiff_t)i);
            throw;
            }
        b.stream = k;
        queued[k]++;
        started = true;
        }
    // wait for the oldest started batch rather than let the caller spin: when
    // nothing could start and both streams are full (and no scene is left to
    // build meanwhile), when destroying, or when nothing is left to plan or
    // start
    bool unbuilt = false;
    for (uint64_t id : S.active)
        unbuilt = unbuilt || !S.reqs[id].built;
    bool wait = !started && !unbuilt && queued[0] >= SERVER_DEPTH && queued[1] >= SERVER_DEPTH;
    wait = wait || S.draining;
    if (!wait && S.queue.empty())
        {
        bool any_waiting = false;
        for (auto& b : S.batches)
            any_waiting = any_waiting || b.stream < 0;
        bool plannable = false;
        for (uint64_t id : S.active)
            plannable = plannable || S.reqs[id].scheduled < S.reqs[id].samples;
        wait = !any_waiting && !plannable;
        }
    if (wait)
        for (size_t i = 0; i < S.batches.size(); i++)
            if (S.batches[i].stream >= 0)
                {
                HIPCHECK(hipEventSynchronize(S.batches[i].done));
                server_finish_batch(S, i);
                break;
                }
    server_retire(S, now);
    n_active = (uint32_t)S.active.size();
    n_queued = (uint32_t)S.queue.size();
    }

// Finish every batch in flight (server destruction).
static void server_drain(ServerObj& S)
    {
    S.draining = true;
    while (!S.batches.empty())
        {
        uint32_t a = 0, q = 0;
        server_step(S, -1.0, a, q);
        }
    }
    } // namespace fh

// ====================================================================== C ABI
using namespace fh;

struct fh_device_s
    {
    DeviceObj obj;
    };
struct fh_scene_s
    {
    SceneObj obj;
    };
struct fh_geometry_s
    {
    GeometryObj obj;
    };
struct fh_tracer_s
    {
    TracerObj obj;
    };
struct fh_server_s
    {
    ServerObj obj;
    };

#define NONNULL(p)                                                                                   \
    do                                                                                               \
        {                                                                                            \
        if (!(p))                                                                                    \
            throw Error(FH_ERR_INVALID, "null handle or pointer: " #p);                              \
        } while (0)

static void copy_str(const std::string& s, char* buf, size_t cap)
    {
    if (!buf || cap == 0)
        return;
    size_t n = std::min(cap - 1, s.size());
    std::memcpy(buf, s.data(), n);
    buf[n] = 0;
    }

static std::string gpu_list()
    {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess)
        return "";
    std::string out;
    for (int i = 0; i < n; i++)
        {
        hipDeviceProp_t p;
        if (hipGetDeviceProperties(&p, i) != hipSuccess)
            continue;
        out += "[" + std::to_string(i) + "]: " + p.name + " (" + p.gcnArchName + ") "
               + std::to_string(p.multiProcessorCount) + " CUs, "
               + std::to_string((unsigned long long)(p.totalGlobalMem >> 20)) + " MiB\n";
        }
    return out;
    }

extern "C"
    {
    const char* fh_last_error(void)
        {
        return g_last_error.c_str();
        }

    int fh_get_num_available_devices(int* n)
        {
        return guard(
            [&]
            {
                NONNULL(n);
                int c = 0;
                if (hipGetDeviceCount(&c) != hipSuccess)
                    c = 0;
                *n = c;
            });
        }

    int fh_device_create(int limit, fh_device* out)
        {
        return guard(
            [&]
            {
                NONNULL(out);
                int c = 0;
                if (hipGetDeviceCount(&c) != hipSuccess || c == 0)
                    throw Error(FH_ERR_HIP, "No GPUs are available");
                auto d = std::make_unique<fh_device_s>();
                DeviceObj& D = d->obj;
                D.limit = limit;
                D.hip_dev = 0;
                HIPCHECK(hipSetDevice(D.hip_dev));
                // this GPU's copy of the allocation-free scratch array
                HIPCHECK(hipGetSymbolAddress((void**)&D.scratch, HIP_SYMBOL(scratch)));
                hipDeviceProp_t p;
                HIPCHECK(hipGetDeviceProperties(&p, D.hip_dev));
                D.n_cu = p.multiProcessorCount;
                D.name = std::string(p.name) + " (" + p.gcnArchName + ")";
                HIPCHECK(hipStreamCreateWithFlags(&D.stream, hipStreamNonBlocking));
                HIPCHECK(hipStreamCreateWithFlags(&D.stream2, hipStreamNonBlocking));
                HIPCHECK(hipStreamCreateWithFlags(&D.copy_stream, hipStreamNonBlocking));
                HIPCHECK(hipStreamCreateWithFlags(&D.stream3, hipStreamNonBlocking));
                HIPCHECK(hipEventCreateWithFlags(&D.upload_ev, hipEventDisableTiming));
                HIPCHECK(hipEventCreateWithFlags(&D.build_ev, hipEventDisableTiming));
                for (auto& e : D.acc_done)
                    HIPCHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
                // overflow stacks are sized for the largest resident traversal grid
                D.trav_grid = std::max(max_grid_ss<8>(D), max_grid_ss<16>(D));
                if (const char* e = getenv("FRESNEL_AMD_BATCH_PATHS"))
                    D.batch_paths = std::max<uint64_t>(1, strtoull(e, nullptr, 10));
                if (const char* e = getenv("FRESNEL_AMD_TAIL_RAYS"))
                    D.tail_rays = (uint32_t)std::max<uint64_t>(1, strtoull(e, nullptr, 10));
                void* hp = nullptr;
                HIPCHECK(hipHostMalloc(&hp, 64, hipHostMallocDefault));
                D.h_pinned = (uint32_t*)hp;
                hp = nullptr;
                HIPCHECK(hipHostMalloc(&hp, (QSEG_FLAG + 1) * sizeof(uint32_t), hipHostMallocDefault));
                D.h_counts = (uint32_t*)hp;
                D.counters.ensure(N_CTR);
                HIPCHECK(hipMemset(D.counters.p, 0, N_CTR * sizeof(unsigned long long)));
                *out = d.release();
            });
        }

    int fh_device_destroy(fh_device dev)
        {
        return guard(
            [&]
            {
                if (!dev)
                    return;
                (void)hipStreamSynchronize(dev->obj.stream);
                for (auto& t : dev->obj.times)
                    {
                    (void)hipEventDestroy(t.a);
                    (void)hipEventDestroy(t.b);
                    }
                for (auto e : dev->obj.event_pool)
                    (void)hipEventDestroy(e);
                (void)hipStreamSynchronize(dev->obj.stream2);
                for (auto& g : dev->obj.graph_cache)
                    (void)hipGraphExecDestroy(g.exec);
                dev->obj.graph_cache.clear();
                if (dev->obj.h_pinned)
                    (void)hipHostFree(dev->obj.h_pinned);
                if (dev->obj.h_counts)
                    (void)hipHostFree(dev->obj.h_counts);
                (void)hipStreamSynchronize(dev->obj.stream2);
                for (auto e : dev->obj.acc_done)
                    if (e)
                        (void)hipEventDestroy(e);
                if (dev->obj.upload_ev)
                    (void)hipEventDestroy(dev->obj.upload_ev);
                if (dev->obj.build_ev)
                    (void)hipEventDestroy(dev->obj.build_ev);
                hipStream_t s = dev->obj.stream, s2 = dev->obj.stream2, s3 = dev->obj.copy_stream,
                            s4 = dev->obj.stream3;
                if (s3)
                    (void)hipStreamSynchronize(s3);
                if (s4)
                    (void)hipStreamSynchronize(s4);
                delete dev;
                if (s)
                    (void)hipStreamDestroy(s);
                if (s2)
                    (void)hipStreamDestroy(s2);
                if (s3)
                    (void)hipStreamDestroy(s3);
                if (s4)
                    (void)hipStreamDestroy(s4);
            });
        }

    int fh_device_describe(fh_device dev, char* buf, size_t cap)
        {
        return guard(
            [&]
            {
                NONNULL(dev);
                copy_str("Enabled HIP devices:\n[0]: " + dev->obj.name + ", " + std::to_string(dev->obj.n_cu) + " CUs\n",
                         buf, cap);
            });
        }

    int fh_device_get_all_gpus(char* buf, size_t cap)
        {
        return guard([&] { copy_str(gpu_list(), buf, cap); });
        }

    // ---------------------------------------------------------------- scene
    int fh_scene_create(fh_device dev, fh_scene* out)
        {
        return guard(
            [&]
            {
                NONNULL(dev);
                NONNULL(out);
                auto s = std::make_unique<fh_scene_s>();
                SceneObj& S = s->obj;
                S.dev = &dev->obj;
                // UserCamera() and the default lights of fresnel/cpu/Scene.cc:15-27
                S.camera.position = {0, 0, 0};
                S.camera.look_at = {0, 0, 1};
                S.camera.up = {0, 1, 0};
                S.camera.h = 1;
                S.camera.f = 1;
                S.camera.f_stop = 1;
                S.camera.focus_distance = 1;
                S.camera.model = 0;
                std::memset(&S.lights, 0, sizeof(S.lights));
                S.lights.N = 2;
                S.lights.direction[0] = {-1, 0.3f, 1};
                S.lights.color[0] = {1, 1, 1};
                S.lights.direction[1] = {1, 0, 1};
                S.lights.color[1] = {0.1f, 0.1f, 0.1f};
                *out = s.release();
            });
        }
    int fh_scene_destroy(fh_scene scene)
        {
        return guard(
            [&]
            {
                if (!scene)
                    return;
                for (auto& kv : scene->obj.geoms)
                    kv.second->scene = nullptr;
                (void)hipStreamSynchronize(scene->obj.dev->stream);
                delete scene;
            });
        }
    int fh_scene_set_camera(fh_scene scene, const fh_user_camera* cam)
        {
        return guard(
            [&]
            {
                NONNULL(scene);
                NONNULL(cam);
                scene->obj.camera = *cam;
            });
        }
    int fh_scene_get_camera(fh_scene scene, fh_user_camera* cam)
        {
        return guard(
            [&]
            {
                NONNULL(scene);
                NONNULL(cam);
                *cam = scene->obj.camera;
            });
        }
    int fh_scene_set_background(fh_scene scene, fh_rgb color, float alpha)
        {
        return guard(
            [&]
            {
                NONNULL(scene);
                scene->obj.bg = color;
                scene->obj.bg_alpha = alpha;
            });
        }
    int fh_scene_get_background(fh_scene scene, fh_rgb* color, float* alpha)
        {
        return guard(
            [&]
            {
                NONNULL(scene);
                if (color)
                    *color = scene->obj.bg;
                if (alpha)
                    *alpha = scene->obj.bg_alpha;
            });
        }
    int fh_scene_set_refit(fh_scene scene, int enable)
        {
        return guard(
            [&]
            {
                NONNULL(scene);
                scene->obj.allow_refit = enable != 0;
            });
        }
    int fh_scene_set_refit_limit(fh_scene scene, float limit)
        {
        return guard(
            [&]
            {
                NONNULL(scene);
                REQUIRE(limit >= 1.0f, "the refit limit must be >= 1");
                scene->obj.refit_limit = limit;
            });
        }
    int fh_scene_bvh_cost(fh_scene scene, float* built, float* current)
        {
        return guard(
            [&]
            {
                NONNULL(scene);
                if (built)
                    *built = scene->obj.sah_built;
                if (current)
                    *current = scene->obj.sah_last;
            });
        }
    int fh_scene_bvh_info(fh_scene scene, uint64_t* builds, uint64_t* refits, uint32_t* n_prims)
        {
        return guard(
            [&]
            {
                NONNULL(scene);
                if (builds)
                    *builds = scene->obj.builds;
                if (refits)
                    *refits = scene->obj.refits;
                if (n_prims)
                    *n_prims = scene->obj.n_prims;
            });
        }
    int fh_scene_set_lights(fh_scene scene, const fh_lights* lights)
        {
        return guard(
            [&]
            {
                NONNULL(scene);
                NONNULL(lights);
                scene->obj.lights = *lights;
            });
        }
    int fh_scene_get_lights(fh_scene scene, fh_lights* lights)
        {
        return guard(
            [&]
            {
                NONNULL(scene);
                NONNULL(lights);
                *lights = scene->obj.lights;
            });
        }

    // ---------------------------------------------------------------- geometry
    static GeometryObj& new_geometry(fh_scene scene, uint32_t type, uint32_t N, std::unique_ptr<fh_geometry_s>& h)
        {
        NONNULL(scene);
        h = std::make_unique<fh_geometry_s>();
        GeometryObj& g = h->obj;
        g.scene = &scene->obj;
        g.type = type;
        g.N = N;
        g.slot = scene->obj.alloc_slot();
        // defaults of the reference geometry constructors (e.g. cpu/GeometrySphere.cc:29-31)
        g.mat = make_material(1, 0, 1, 0.0f);
        g.outline = make_material(0, 0, 0, 1.0f);
        g.outline_width = 0;
        return g;
        }
    static void attach(fh_scene scene, std::unique_ptr<fh_geometry_s>& h, fh_geometry* out)
        {
        scene->obj.geoms[h->obj.slot] = &h->obj;
        scene->obj.bvh_valid = false;
        *out = h.release();
        }

    int fh_geometry_sphere_create(fh_scene scene, uint32_t N, fh_geometry* out)
        {
        return guard(
            [&]
            {
                NONNULL(out);
                std::unique_ptr<fh_geometry_s> h;
                GeometryObj& g = new_geometry(scene, G_SPHERE, N, h);
                g.position.assign(3 * (size_t)N, 0.0f);
                g.radius_buf.assign(N, 0.0f);
                g.color.assign(3 * (size_t)N, 0.0f);
                attach(scene, h, out);
            });
        }
    int fh_geometry_cylinder_create(fh_scene scene, uint32_t N, fh_geometry* out)
        {
        return guard(
            [&]
            {
                NONNULL(out);
                std::unique_ptr<fh_geometry_s> h;
                GeometryObj& g = new_geometry(scene, G_CYLINDER, N, h);
                g.points.assign(6 * (size_t)N, 0.0f);
                g.radius_buf.assign(N, 0.0f);
                g.color.assign(6 * (size_t)N, 0.0f);
                attach(scene, h, out);
            });
        }
    int fh_geometry_polygon_create(fh_scene scene, const float* vertices, uint32_t n_verts, float rounding_radius,
                                   uint32_t N, fh_geometry* out)
        {
        return guard(
            [&]
            {
                NONNULL(out);
                REQUIRE(n_verts >= 1 && vertices, "vertices must be a Nvert by 2 array");
                std::unique_ptr<fh_geometry_s> h;
                GeometryObj& g = new_geometry(scene, G_POLYGON, N, h);
                g.verts.assign(vertices, vertices + 2 * (size_t)n_verts);
                g.rounding = rounding_radius;
                // xy radius, fresnel/cpu/GeometryPolygon.cc:55-65
                float r = 0;
                for (uint32_t k = 0; k < n_verts; k++)
                    {
                    v2 p0;
                    p0.x = vertices[2 * k];
                    p0.y = vertices[2 * k + 1];
                    r = std::max(r, sqrtf(dot(p0, p0)));
                    }
                g.radius = r + rounding_radius;
                g.position.assign(2 * (size_t)N, 0.0f);
                g.angle.assign(N, 0.0f);
                g.color.assign(3 * (size_t)N, 0.0f);
                attach(scene, h, out);
            });
        }
    int fh_geometry_convex_polyhedron_create(fh_scene scene, const float* origins, const float* normals,
                                             const float* colors, uint32_t n_planes, uint32_t N, float radius,
                                             fh_geometry* out)
        {
        return guard(
            [&]
            {
                NONNULL(out);
                REQUIRE(origins && normals && colors && n_planes >= 1, "polyhedron planes are required");
                std::unique_ptr<fh_geometry_s> h;
                GeometryObj& g = new_geometry(scene, G_POLYHEDRON, N, h);
                for (uint32_t k = 0; k < n_planes; k++)
                    {
                    // n = n / sqrtf(dot(n, n)), cpu/GeometryConvexPolyhedron.cc:85-86
                    v3 n = mk3(normals[3 * k], normals[3 * k + 1], normals[3 * k + 2]);
                    n = div_rcp(n, sqrtf(dot(n, n)));
                    g.plane_n.insert(g.plane_n.end(), {n.x, n.y, n.z});
                    g.plane_p.insert(g.plane_p.end(), {origins[3 * k], origins[3 * k + 1], origins[3 * k + 2]});
                    g.plane_color.insert(g.plane_color.end(), {colors[3 * k], colors[3 * k + 1], colors[3 * k + 2]});
                    }
                g.radius = radius;
                g.position.assign(3 * (size_t)N, 0.0f);
                g.orientation.assign(4 * (size_t)N, 0.0f);
                g.color.assign(3 * (size_t)N, 0.0f);
                attach(scene, h, out);
            });
        }
    int fh_geometry_mesh_create(fh_scene scene, const float* vertices, uint32_t n_verts, uint32_t N, fh_geometry* out)
        {
        return guard(
            [&]
            {
                NONNULL(out);
                REQUIRE(vertices && n_verts >= 3, "vertices must be a Nvert by 3 array");
                REQUIRE(n_verts % 3 == 0, "the number of triangle vertices must be a multiple of three.");
                std::unique_ptr<fh_geometry_s> h;
                GeometryObj& g = new_geometry(scene, G_MESH, N, h);
                g.verts.assign(vertices, vertices + 3 * (size_t)n_verts);
                g.n_faces = n_verts / 3;
                g.position.assign(3 * (size_t)N, 0.0f);
                g.orientation.assign(4 * (size_t)N, 0.0f);
                g.color.assign(3 * (size_t)n_verts, 0.0f);
                attach(scene, h, out);
            });
        }
    int fh_geometry_destroy(fh_geometry geom)
        {
        return guard(
            [&]
            {
                if (!geom)
                    return;
                GeometryObj& g = geom->obj;
                if (g.scene)
                    {
                    (void)hipStreamSynchronize(g.scene->dev->stream);
                    if (g.valid)
                        {
                        g.scene->geoms.erase(g.slot);
                        g.scene->bvh_valid = false;
                        }
                    }
                delete geom;
            });
        }

    int fh_geometry_buffer(fh_geometry geom, int which, float** data, uint32_t* ndim, uint64_t shape[3])
        {
        return guard(
            [&]
            {
                NONNULL(geom);
                NONNULL(data);
                NONNULL(ndim);
                NONNULL(shape);
                GeometryObj& g = geom->obj;
                std::vector<float>* v = nullptr;
                uint32_t nd = 2;
                uint64_t s0 = g.N, s1 = 3, s2 = 0;
                switch (which)
                    {
                case FH_BUF_POSITION:
                    REQUIRE(g.type != G_CYLINDER, "geometry has no position buffer");
                    v = &g.position;
                    s1 = (g.type == G_POLYGON) ? 2 : 3;
                    break;
                case FH_BUF_RADIUS:
                    REQUIRE(g.type == G_SPHERE || g.type == G_CYLINDER, "geometry has no radius buffer");
                    v = &g.radius_buf;
                    nd = 1;
                    break;
                case FH_BUF_COLOR:
                    v = &g.color;
                    if (g.type == G_CYLINDER)
                        {
                        nd = 3;
                        s1 = 2;
                        s2 = 3;
                        }
                    else if (g.type == G_MESH)
                        s0 = (uint64_t)g.n_faces * 3;
                    break;
                case FH_BUF_POINTS:
                    REQUIRE(g.type == G_CYLINDER, "geometry has no points buffer");
                    v = &g.points;
                    nd = 3;
                    s1 = 2;
                    s2 = 3;
                    break;
                case FH_BUF_ORIENTATION:
                    REQUIRE(g.type == G_POLYHEDRON || g.type == G_MESH, "geometry has no orientation buffer");
                    v = &g.orientation;
                    s1 = 4;
                    break;
                case FH_BUF_ANGLE:
                    REQUIRE(g.type == G_POLYGON, "geometry has no angle buffer");
                    v = &g.angle;
                    nd = 1;
                    break;
                default:
                    throw Error(FH_ERR_INVALID, "unknown buffer id");
                    }
                *data = v->data();
                *ndim = nd;
                shape[0] = s0;
                shape[1] = s1;
                shape[2] = s2;
            });
        }
    int fh_geometry_update(fh_geometry geom)
        {
        return guard(
            [&]
            {
                NONNULL(geom);
                geom->obj.dirty = true;
            });
        }
    int fh_geometry_update_rows(fh_geometry geom, int which, uint64_t first, uint64_t count)
        {
        return guard(
            [&]
            {
                NONNULL(geom);
                GeometryObj& g = geom->obj;
                float* data;
                uint32_t nd;
                uint64_t shape[3];
                const int rc = fh_geometry_buffer(geom, which, &data, &nd, shape);
                if (rc != FH_OK)
                    throw Error(rc, fh_last_error());
                REQUIRE(first <= shape[0] && count <= shape[0] - first, "rows out of range");
                if (count == 0)
                    return;
                uint64_t& lo = which == FH_BUF_COLOR ? g.color_lo : g.shape_lo;
                uint64_t& hi = which == FH_BUF_COLOR ? g.color_hi : g.shape_hi;
                lo = std::min(lo, first);
                hi = std::max(hi, first + count);
            });
        }
#define GEOM_VALID(g)                                                                                \
    do                                                                                               \
        {                                                                                            \
        if (!(g).valid || !(g).scene)                                                                \
            throw Error(FH_ERR_STATE, "Cannot modify a removed Geometry");                           \
        } while (0)

    int fh_geometry_set_material(fh_geometry geom, const fh_material* m)
        {
        return guard(
            [&]
            {
                NONNULL(geom);
                NONNULL(m);
                geom->obj.mat = *m;
            });
        }
    int fh_geometry_get_material(fh_geometry geom, fh_material* m)
        {
        return guard(
            [&]
            {
                NONNULL(geom);
                NONNULL(m);
                *m = geom->obj.mat;
            });
        }
    int fh_geometry_set_outline_material(fh_geometry geom, const fh_material* m)
        {
        return guard(
            [&]
            {
                NONNULL(geom);
                NONNULL(m);
                geom->obj.outline = *m;
            });
        }
    int fh_geometry_get_outline_material(fh_geometry geom, fh_material* m)
        {
        return guard(
            [&]
            {
                NONNULL(geom);
                NONNULL(m);
                *m = geom->obj.outline;
            });
        }
    int fh_geometry_set_outline_width(fh_geometry geom, float w)
        {
        return guard(
            [&]
            {
                NONNULL(geom);
                geom->obj.outline_width = w;
            });
        }
    int fh_geometry_get_outline_width(fh_geometry geom, float* w)
        {
        return guard(
            [&]
            {
                NONNULL(geom);
                NONNULL(w);
                *w = geom->obj.outline_width;
            });
        }
    int fh_geometry_enable(fh_geometry geom)
        {
        return guard(
            [&]
            {
                NONNULL(geom);
                if (!geom->obj.valid || !geom->obj.scene)
                    throw Error(FH_ERR_STATE, "Cannot enable inactive Geometry");
                geom->obj.enabled = true;
            });
        }
    int fh_geometry_disable(fh_geometry geom)
        {
        return guard(
            [&]
            {
                NONNULL(geom);
                if (!geom->obj.valid || !geom->obj.scene)
                    throw Error(FH_ERR_STATE, "Cannot disable inactive Geometry");
                geom->obj.enabled = false;
            });
        }
    int fh_geometry_remove(fh_geometry geom)
        {
        return guard(
            [&]
            {
                NONNULL(geom);
                GeometryObj& g = geom->obj;
                if (g.valid && g.scene)
                    {
                    (void)hipStreamSynchronize(g.scene->dev->stream);
                    g.scene->geoms.erase(g.slot);
                    g.scene->bvh_valid = false;
                    }
                g.valid = false;
            });
        }
    int fh_geometry_set_color_by_face(fh_geometry geom, float f)
        {
        return guard(
            [&]
            {
                NONNULL(geom);
                geom->obj.color_by_face = f;
            });
        }
    int fh_geometry_get_color_by_face(fh_geometry geom, float* f)
        {
        return guard(
            [&]
            {
                NONNULL(geom);
                NONNULL(f);
                *f = geom->obj.color_by_face;
            });
        }
    int fh_geometry_get_radius(fh_geometry geom, float* r)
        {
        return guard(
            [&]
            {
                NONNULL(geom);
                NONNULL(r);
                *r = geom->obj.radius;
            });
        }

    // ---------------------------------------------------------------- tracers
    int fh_tracer_path_create(fh_device dev, uint32_t w, uint32_t h, uint32_t light_samples, fh_tracer* out)
        {
        return guard(
            [&]
            {
                NONNULL(dev);
                NONNULL(out);
                auto t = std::make_unique<fh_tracer_s>();
                TracerObj& T = t->obj;
                T.dev = &dev->obj;
                T.type = T_PATH;
                T.L = light_samples;
                T.resize(w, h);
                // TracerPath ctor -> reset(): sample count 0, seed 0 -> 1
                T.n_samples = 0;
                T.seed++;
                *out = t.release();
            });
        }
    int fh_tracer_direct_create(fh_device dev, uint32_t w, uint32_t h, fh_tracer* out)
        {
        return guard(
            [&]
            {
                NONNULL(dev);
                NONNULL(out);
                auto t = std::make_unique<fh_tracer_s>();
                TracerObj& T = t->obj;
                T.dev = &dev->obj;
                T.type = T_DIRECT;
                T.aa_n = 8;
                T.resize(w, h);
                *out = t.release();
            });
        }
    int fh_tracer_destroy(fh_tracer tr)
        {
        return guard(
            [&]
            {
                if (!tr)
                    return;
                (void)hipStreamSynchronize(tr->obj.dev->stream);
                for (int k = 0; k < 2; k++)
                    if (tr->obj.async_ev[k])
                        {
                        (void)hipEventSynchronize(tr->obj.async_ev[k]);
                        (void)hipEventDestroy(tr->obj.async_ev[k]);
                        }
                delete tr;
            });
        }
    int fh_tracer_resize(fh_tracer tr, uint32_t w, uint32_t h)
        {
        return guard(
            [&]
            {
                NONNULL(tr);
                TracerObj& T = tr->obj;
                T.resize(w, h);
                if (T.type == T_PATH)
                    {
                    // cpu/TracerPath.h:48-53
                    T.n_samples = 0;
                    T.seed++;
                    }
            });
        }
    int fh_tracer_get_size(fh_tracer tr, uint32_t* w, uint32_t* h)
        {
        return guard(
            [&]
            {
                NONNULL(tr);
                if (w)
                    *w = tr->obj.W;
                if (h)
                    *h = tr->obj.H;
            });
        }
    int fh_tracer_render(fh_tracer tr, fh_scene scene)
        {
        return guard(
            [&]
            {
                NONNULL(tr);
                NONNULL(scene);
                if (tr->obj.type == T_DIRECT)
                    render_preview(*tr->obj.dev, scene->obj, tr->obj);
                else
                    render_path(*tr->obj.dev, {&scene->obj}, {&tr->obj}, 1);
            });
        }
    int fh_tracer_render_samples(fh_tracer tr, fh_scene scene, uint32_t samples)
        {
        return guard(
            [&]
            {
                NONNULL(tr);
                NONNULL(scene);
                REQUIRE(tr->obj.type == T_PATH, "render_samples needs a Path tracer");
                if (samples)
                    render_path(*tr->obj.dev, {&scene->obj}, {&tr->obj}, samples);
            });
        }
    int fh_tracer_map_output(fh_tracer tr, int which, void** data)
        {
        return guard(
            [&]
            {
                NONNULL(tr);
                NONNULL(data);
                TracerObj& T = tr->obj;
                const size_t n = (size_t)T.W * T.H;
                if (which == FH_OUT_SRGB)
                    {
                    HIPCHECK(hipMemcpy(T.h_srgb.data(), T.srgb.p, n * 4, hipMemcpyDeviceToHost));
                    *data = T.h_srgb.data();
                    }
                else if (which == FH_OUT_LINEAR)
                    {
                    HIPCHECK(hipMemcpy(T.h_linear.data(), T.linear.p, n * 16, hipMemcpyDeviceToHost));
                    *data = T.h_linear.data();
                    }
                else
                    throw Error(FH_ERR_INVALID, "unknown output buffer");
            });
        }
    int fh_tracer_output_async(fh_tracer tr, int which)
        {
        return guard(
            [&]
            {
                NONNULL(tr);
                REQUIRE(which == FH_OUT_SRGB || which == FH_OUT_LINEAR, "unknown output buffer");
                TracerObj& T = tr->obj;
                DeviceObj& D = *T.dev;
                std::lock_guard<std::mutex> lock(D.mu);
                const size_t bytes = (size_t)T.W * T.H * (which == FH_OUT_SRGB ? 4 : 16);
                const void* src = which == FH_OUT_SRGB ? (const void*)T.srgb.p : (const void*)T.linear.p;
                if (T.async_started[which])
                    HIPCHECK(hipEventSynchronize(T.async_ev[which])); // the pinned buffer is reused
                else if (!T.async_ev[which])
                    HIPCHECK(hipEventCreateWithFlags(&T.async_ev[which], hipEventDisableTiming));
                T.h_async[which].ensure(bytes);
                // the tracer's renders return complete, but a clear may still be queued
                HIPCHECK(hipStreamSynchronize(D.stream));
                HIPCHECK(hipMemcpyAsync(T.h_async[which].p, src, bytes, hipMemcpyDeviceToHost, D.copy_stream));
                HIPCHECK(hipEventRecord(T.async_ev[which], D.copy_stream));
                T.async_started[which] = true;
                T.async_unordered[which] = true;
                T.async_w[which] = T.W;
                T.async_h[which] = T.H;
            });
        }
    int fh_tracer_output_ready(fh_tracer tr, int which, int* ready)
        {
        return guard(
            [&]
            {
                NONNULL(tr);
                NONNULL(ready);
                REQUIRE(which == FH_OUT_SRGB || which == FH_OUT_LINEAR, "unknown output buffer");
                TracerObj& T = tr->obj;
                REQUIRE(T.async_started[which], "no asynchronous copy of this output was started");
                const hipError_t e = hipEventQuery(T.async_ev[which]);
                if (e != hipSuccess && e != hipErrorNotReady)
                    HIPCHECK(e);
                *ready = e == hipSuccess ? 1 : 0;
            });
        }
    int fh_tracer_output_wait(fh_tracer tr, int which, void** data, uint32_t* w, uint32_t* h)
        {
        return guard(
            [&]
            {
                NONNULL(tr);
                NONNULL(data);
                REQUIRE(which == FH_OUT_SRGB || which == FH_OUT_LINEAR, "unknown output buffer");
                TracerObj& T = tr->obj;
                REQUIRE(T.async_started[which], "no asynchronous copy of this output was started");
                HIPCHECK(hipEventSynchronize(T.async_ev[which]));
                *data = T.h_async[which].p;
                if (w)
                    *w = T.async_w[which];
                if (h)
                    *h = T.async_h[which];
            });
        }
    int fh_tracer_output_host(fh_tracer tr, int which, void** data)
        {
        return guard(
            [&]
            {
                NONNULL(tr);
                NONNULL(data);
                TracerObj& T = tr->obj;
                if (which == FH_OUT_SRGB)
                    *data = T.h_srgb.data();
                else if (which == FH_OUT_LINEAR)
                    *data = T.h_linear.data();
                else
                    throw Error(FH_ERR_INVALID, "unknown output buffer");
            });
        }
    int fh_tracer_histogram(fh_tracer tr, uint32_t* counts)
        {
        return guard(
            [&]
            {
                NONNULL(tr);
                NONNULL(counts);
                TracerObj& T = tr->obj;
                DeviceObj& D = *T.dev;
                const uint32_t n = T.W * T.H;
                T.hist.ensure(4 * HIST_BINS);
                hipStream_t s = D.stream;
                HIPCHECK(hipMemsetAsync(T.hist.p, 0, 4 * HIST_BINS * sizeof(uint32_t), s));
                k_histogram<<<std::min(blocks(n, 256), 1024u), 256, 0, s>>>(T.linear.p, n, T.hist.p);
                HIPCHECK(hipGetLastError());
                HIPCHECK(hipMemcpyAsync(counts, T.hist.p, 4 * HIST_BINS * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
                HIPCHECK(hipStreamSynchronize(s));
            });
        }
    int fh_tracer_histogram_noalloc(fh_tracer tr, uint32_t* counts)
        {
        return guard(
            [&]
            {
                NONNULL(tr);
                NONNULL(counts);
                TracerObj& T = tr->obj;
                DeviceObj& D = *T.dev;
                static_assert(4 * HIST_BINS <= NOALLOC_SCRATCH_WORDS, "histogram bins exceed the scratch array");
                const uint64_t n = (uint64_t)T.W * T.H;
                if (n == 0 || n > 0xffffffffull)
                    throw Error(FH_ERR_INVALID, "histogram: the image has " + std::to_string(n)
                                                    + " pixels; the allocation-free histogram takes 1 to 2^32 - 1");
                if (!D.scratch)
                    throw Error(FH_ERR_STATE, "histogram: the device's scratch array has no address");
                std::lock_guard<std::mutex> lock(D.mu);
                hipStream_t s = D.stream;
                // cleared by a kernel (no memset node if a caller captures the stream)
                k_fill_u32<<<blocks(4 * HIST_BINS, 256), 256, 0, s>>>(D.scratch, 4 * HIST_BINS, 0u);
                HIPCHECK(hipGetLastError());
                k_histogram<<<std::min(blocks(n, 256), 1024u), 256, 0, s>>>(T.linear.p, (uint32_t)n, D.scratch);
                HIPCHECK(hipGetLastError());
                HIPCHECK(hipMemcpyAsync(counts, D.scratch, 4 * HIST_BINS * sizeof(uint32_t), hipMemcpyDeviceToHost,
                                        s));
                HIPCHECK(hipStreamSynchronize(s));
            });
        }
    int fh_tracer_enable_highlight_warning(fh_tracer tr, fh_rgb color)
        {
        return guard(
            [&]
            {
                NONNULL(tr);
                tr->obj.highlight = true;
                tr->obj.hw_color = color;
            });
        }
    int fh_tracer_disable_highlight_warning(fh_tracer tr)
        {
        return guard(
            [&]
            {
                NONNULL(tr);
                tr->obj.highlight = false;
            });
        }
    int fh_tracer_set_seed(fh_tracer tr, uint32_t seed)
        {
        return guard(
            [&]
            {
                NONNULL(tr);
                tr->obj.seed = seed;
            });
        }
    int fh_tracer_get_seed(fh_tracer tr, uint32_t* seed)
        {
        return guard(
            [&]
            {
                NONNULL(tr);
                NONNULL(seed);
                *seed = tr->obj.seed;
            });
        }
    int fh_tracer_path_reset(fh_tracer tr)
        {
        return guard(
            [&]
            {
                NONNULL(tr);
                TracerObj& T = tr->obj;
                T.n_samples = 0;
                T.seed++;
                T.clear();
            });
        }
    int fh_tracer_path_set_light_samples(fh_tracer tr, uint32_t light_samples)
        {
        return guard(
            [&]
            {
                NONNULL(tr);
                REQUIRE(light_samples >= 1 && light_samples < 65536, "light_samples must be in [1, 65535]");
                tr->obj.L = light_samples;
            });
        }
    int fh_tracer_path_get_num_samples(fh_tracer tr, uint32_t* n)
        {
        return guard(
            [&]
            {
                NONNULL(tr);
                NONNULL(n);
                *n = tr->obj.n_samples;
            });
        }
    int fh_tracer_direct_set_aa_n(fh_tracer tr, uint32_t n)
        {
        return guard(
            [&]
            {
                NONNULL(tr);
                REQUIRE(n >= 1, "aa_n must be >= 1");
                tr->obj.aa_n = n;
            });
        }
    int fh_tracer_direct_get_aa_n(fh_tracer tr, uint32_t* n)
        {
        return guard(
            [&]
            {
                NONNULL(tr);
                NONNULL(n);
                *n = tr->obj.aa_n;
            });
        }

    // ---------------------------------------------------------------- batching
    int fh_batch_render(fh_scene* scenes, fh_tracer* tracers, uint32_t n, uint32_t samples)
        {
        return guard(
            [&]
            {
                REQUIRE(n >= 1 && scenes && tracers, "empty batch");
                std::vector<SceneObj*> sc;
                std::vector<TracerObj*> tr;
                for (uint32_t k = 0; k < n; k++)
                    {
                    NONNULL(scenes[k]);
                    NONNULL(tracers[k]);
                    sc.push_back(&scenes[k]->obj);
                    tr.push_back(&tracers[k]->obj);
                    for (uint32_t q = 0; q < k; q++)
                        REQUIRE(tr[q] != tr.back(), "a tracer appears twice in one batch");
                    }
                if (samples)
                    render_path(*tr[0]->dev, sc, tr, samples);
            });
        }


    // ---------------------------------------------------------------- serving
    int fh_server_create(fh_device dev, uint64_t hbm_budget, uint64_t target_paths, fh_server* out)
        {
        return guard(
            [&]
            {
                NONNULL(dev);
                NONNULL(out);
                auto s = std::make_unique<fh_server_s>();
                s->obj.dev = &dev->obj;
                s->obj.budget = hbm_budget;
                if (target_paths)
                    s->obj.target_paths = target_paths;
                *out = s.release();
            });
        }
    int fh_server_destroy(fh_server srv)
        {
        return guard(
            [&]
            {
                if (!srv)
                    return;
                ServerObj& S = srv->obj;
                try
                    {
                    server_drain(S);
                    }
                catch (...)
                    {
                    }
                for (int k = 0; k < 2; k++)
                    if (S.st[k])
                        (void)hipStreamSynchronize(S.st[k]);
                for (hipEvent_t e : S.all_events)
                    (void)hipEventDestroy(e);
                delete srv; // its streams are the device's
            });
        }
    int fh_server_submit(fh_server srv, fh_scene scene, fh_tracer tr, uint32_t samples, uint32_t light_samples,
                         double now, uint64_t* id)
        {
        return guard(
            [&]
            {
                NONNULL(srv);
                NONNULL(scene);
                NONNULL(tr);
                NONNULL(id);
                ServerObj& S = srv->obj;
                REQUIRE(scene->obj.dev == S.dev && tr->obj.dev == S.dev, "Scene and Tracer devices do not match");
                REQUIRE(tr->obj.type == T_PATH, "requests need a Path tracer");
                REQUIRE(light_samples >= 1 && light_samples < 65536, "light_samples must be in [1, 65535]");
                REQUIRE((uint64_t)tr->obj.W * tr->obj.H * light_samples <= MAX_BATCH_PATHS,
                        "request too large: one camera sample exceeds 2^31 - 1 light-sample paths");
                for (auto& kv : S.reqs)
                    REQUIRE(kv.second.tracer != &tr->obj || kv.second.state >= REQ_DONE,
                            "tracer already has an open request");
                RequestObj r;
                r.id = S.next_id++;
                r.scene = &scene->obj;
                r.tracer = &tr->obj;
                r.samples = samples;
                r.L = light_samples;
                r.t_arrival = S.clock(now);
                uint32_t pb;
                const uint64_t np = scene_prims(scene->obj, pb);
                r.footprint = request_footprint(np, pb, tr->obj.W, tr->obj.H, light_samples);
                // Path.sample(reset=True) semantics (fresnel/tracer.py:249-252)
                tr->obj.n_samples = 0;
                tr->obj.seed++;
                tr->obj.L = light_samples;
                tr->obj.clear();
                S.reqs[r.id] = r;
                if (samples == 0)
                    {
                    S.reqs[r.id].state = REQ_DONE;
                    S.reqs[r.id].t_done = r.t_arrival;
                    }
                else
                    S.queue.push_back(PolicyItem {r.id, r.footprint, (uint64_t)tr->obj.W * tr->obj.H * light_samples, samples});
                *id = r.id;
            });
        }
    int fh_server_step(fh_server srv, double now, uint32_t* n_active, uint32_t* n_queued)
        {
        return guard(
            [&]
            {
                NONNULL(srv);
                uint32_t a = 0, q = 0;
                server_step(srv->obj, now, a, q);
                if (n_active)
                    *n_active = a;
                if (n_queued)
                    *n_queued = q;
            });
        }
    int fh_server_poll(fh_server srv, uint64_t id, int* state, uint32_t* samples_done, double times[3])
        {
        return guard(
            [&]
            {
                NONNULL(srv);
                auto it = srv->obj.reqs.find(id);
                REQUIRE(it != srv->obj.reqs.end(), "unknown request id");
                const RequestObj& r = it->second;
                if (state)
                    *state = r.state;
                if (samples_done)
                    *samples_done = r.done;
                if (times)
                    {
                    times[0] = r.t_arrival;
                    times[1] = r.t_admit;
                    times[2] = r.t_done;
                    }
            });
        }
    int fh_server_release(fh_server srv, uint64_t id)
        {
        return guard(
            [&]
            {
                NONNULL(srv);
                auto it = srv->obj.reqs.find(id);
                REQUIRE(it != srv->obj.reqs.end(), "unknown request id");
                REQUIRE(it->second.state >= REQ_DONE, "request is still queued or active");
                srv->obj.reqs.erase(it);
            });
        }
    int fh_server_memory(fh_server srv, uint64_t* in_use, uint64_t* budget)
        {
        return guard(
            [&]
            {
                NONNULL(srv);
                if (in_use)
                    *in_use = srv->obj.in_use;
                if (budget)
                    *budget = srv->obj.budget;
            });
        }

    // ---------------------------------------------------------------- policy (no device needed)
    int fh_policy_scratch_bytes(uint64_t paths, uint64_t fixed, uint64_t* out)
        {
        return guard(
            [&]
            {
                NONNULL(out);
                *out = scratch_bytes(paths, fixed);
            });
        }
    int fh_policy_footprint(uint64_t n_prims, uint64_t prim_bytes, uint64_t W, uint64_t H, uint64_t L, uint64_t* out)
        {
        return guard(
            [&]
            {
                NONNULL(out);
                *out = request_footprint(n_prims, prim_bytes, W, H, L);
            });
        }
    // Run the policy on synthetic requests.  Request i arrives before iteration
    // arrive[i]; one iteration advances every active request.  Outputs the
    // iteration of admission and of completion (-1: never, -2: rejected).
    int fh_policy_simulate(uint32_t n, const uint64_t* footprint, const uint64_t* paths_per_sample,
                           const uint32_t* samples, const uint32_t* arrive, uint64_t budget, uint64_t target_paths,
                           uint32_t max_iterations, int32_t* admit_iter, int32_t* done_iter, uint64_t* peak_in_use)
        {
        return guard(
            [&]
            {
                REQUIRE(footprint && paths_per_sample && samples && arrive && admit_iter && done_iter,
                        "null pointer");
                std::deque<PolicyItem> queue;
                std::vector<PolicyItem> active;
                uint64_t in_use = 0, peak = 0;
                for (uint32_t i = 0; i < n; i++)
                    admit_iter[i] = done_iter[i] = -1;
                std::vector<uint32_t> order(n);
                for (uint32_t i = 0; i < n; i++)
                    order[i] = i;
                std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return arrive[a] < arrive[b]; });
                size_t next = 0;
                for (uint32_t it = 0; it < max_iterations; it++)
                    {
                    while (next < n && arrive[order[next]] <= it)
                        {
                        const uint32_t i = order[next++];
                        queue.push_back(PolicyItem {i, footprint[i], paths_per_sample[i], samples[i]});
                        }
                    std::vector<PolicyItem> adm, rej;
                    policy_admit(queue, budget, in_use, adm, rej);
                    for (auto& r : rej)
                        admit_iter[r.id] = done_iter[r.id] = -2;
                    for (auto& a : adm)
                        {
                        admit_iter[a.id] = (int32_t)it;
                        active.push_back(a);
                        }
                    peak = std::max(peak, in_use);
                    if (active.empty())
                        {
                        if (next >= n && queue.empty())
                            break;
                        continue;
                        }
                    std::vector<uint32_t> k = policy_allocate(active, target_paths);
                    std::vector<PolicyItem> still;
                    for (size_t j = 0; j < active.size(); j++)
                        {
                        active[j].remaining -= k[j];
                        if (active[j].remaining == 0)
                            {
                            done_iter[active[j].id] = (int32_t)it;
                            in_use -= active[j].footprint;
                            }
                        else
                            still.push_back(active[j]);
                        }
                    active.swap(still);
                    }
                if (peak_in_use)
                    *peak_in_use = peak;
            });
        }
    // ---------------------------------------------------------------- tuning
    int fh_device_tune(fh_device dev, const char* key, int64_t value, int64_t* previous)
        {
        return guard(
            [&]
            {
                NONNULL(dev);
                NONNULL(key);
                DeviceObj& D = dev->obj;
                std::string k(key);
                int64_t prev;
                if (k == "batch_paths")
                    {
                    prev = (int64_t)D.batch_paths;
                    if (value > 0)
                        D.batch_paths = (uint64_t)value;
                    }
                else if (k == "stack_lds")
                    {
                    prev = D.stack_lds;
                    if (value > 0)
                        {
                        REQUIRE(value == 8 || value == 16, "stack_lds must be 8 or 16");
                        D.stack_lds = (int)value;
                        }
                    }
                else if (k == "trav_chunk")
                    {
                    prev = D.trav_chunk;
                    if (value > 0)
                        {
                        REQUIRE(value % 64 == 0 && value <= 65536, "trav_chunk must be a multiple of 64, at most 65536");
                        D.trav_chunk = (uint32_t)value;
                        }
                    }
                else if (k == "xcd_bands")
                    {
                    prev = D.xcd_bands;
                    if (value > 0)
                        {
                        REQUIRE(value <= 2, "xcd_bands must be 1 (off) or 2 (on)");
                        D.xcd_bands = (uint32_t)value;
                        }
                    }
                else if (k == "graph")
                    {
                    prev = D.graph;
                    if (value > 0)
                        {
                        REQUIRE(value <= 2, "graph must be 1 (launch) or 2 (capture and replay)");
                        D.graph = (uint32_t)value;
                        }
                    }
                else if (k == "inflight")
                    {
                    prev = D.inflight;
                    if (value > 0)
                        {
                        REQUIRE(value <= 2, "inflight must be 1 or 2");
                        D.inflight = (uint32_t)value;
                        }
                    }
                else if (k == "fuse_bounce")
                    {
                    prev = D.fuse_bounce;
                    if (value > 0)
                        {
                        REQUIRE(value <= 2, "fuse_bounce must be 1 (two kernels) or 2 (one kernel)");
                        D.fuse_bounce = (uint32_t)value;
                        }
                    }
                else if (k == "fuse_primary")
                    {
                    prev = D.fuse_primary;
                    if (value > 0)
                        {
                        REQUIRE(value <= 2, "fuse_primary must be 1 (two kernels) or 2 (one kernel)");
                        D.fuse_primary = (uint32_t)value;
                        }
                    }
                else if (k == "primary_tiles")
                    {
                    prev = D.primary_tiles;
                    if (value > 0)
                        {
                        REQUIRE(value <= 2, "primary_tiles must be 1 (BVH search) or 2 (tile lists)");
                        D.primary_tiles = (uint32_t)value;
                        }
                    }
                else if (k == "tile_group")
                    {
                    prev = D.tile_group;
                    if (value > 0)
                        {
                        REQUIRE(value <= 1024, "tile_group must be in [1, 1024]");
                        D.tile_group = (uint32_t)value;
                        }
                    }
                else if (k == "trav_packet")
                    {
                    prev = D.trav_packet;
                    if (value > 0)
                        {
                        REQUIRE(value <= 2, "trav_packet must be 1 (per ray) or 2 (wavefront packets)");
                        D.trav_packet = (uint32_t)value;
                        }
                    }
                else if (k == "primary_chunk")
                    {
                    prev = D.primary_chunk;
                    if (value > 0)
                        {
                        REQUIRE(value % 64 == 0 && value <= 65536,
                                "primary_chunk must be a multiple of 64, at most 65536");
                        D.primary_chunk = (uint32_t)value;
                        }
                    }
                else if (k == "trav_refill")
                    {
                    prev = D.trav_refill;
                    if (value > 0)
                        {
                        REQUIRE(value <= 65, "trav_refill must be in [1, 65]");
                        D.trav_refill = (uint32_t)value;
                        }
                    }
                else if (k == "tail_rays")
                    {
                    prev = (int64_t)D.tail_rays;
                    if (value > 0)
                        D.tail_rays = (uint32_t)value;
                    }
                else if (k == "tile_min_rays")
                    {
                    prev = (int64_t)D.tile_min_rays;
                    if (value > 0)
                        D.tile_min_rays = (uint64_t)value;
                    }
                else
                    throw Error(FH_ERR_INVALID, "unknown tuning key: " + k);
                if (previous)
                    *previous = prev;
            });
        }

    // ---------------------------------------------------------------- diagnostics
    int fh_debug_copy(fh_device dev, uint64_t bytes, float* ms)
        {
        return guard(
            [&]
            {
                NONNULL(dev);
                REQUIRE(bytes >= 16 && bytes % 16 == 0, "bytes must be a positive multiple of 16");
                DeviceObj& D = dev->obj;
                const uint64_t n = bytes / 16;
                DBuf<float4> a, b;
                a.ensure(n);
                b.ensure(n);
                HIPCHECK(hipMemsetAsync(a.p, 0, bytes, D.stream));
                const uint32_t grid = std::max<uint32_t>(1, D.n_cu) * 16;
                hipEvent_t e0, e1;
                HIPCHECK(hipEventCreate(&e0));
                HIPCHECK(hipEventCreate(&e1));
                HIPCHECK(hipEventRecord(e0, D.stream));
                k_debug_copy<<<grid, 256, 0, D.stream>>>(a.p, b.p, n);
                HIPCHECK(hipGetLastError());
                HIPCHECK(hipEventRecord(e1, D.stream));
                HIPCHECK(hipEventSynchronize(e1));
                float t = 0;
                HIPCHECK(hipEventElapsedTime(&t, e0, e1));
                HIPCHECK(hipEventDestroy(e0));
                HIPCHECK(hipEventDestroy(e1));
                if (ms)
                    *ms = t;
            });
        }
    int fh_debug_philox(fh_device dev, const uint32_t* in, uint32_t n, uint32_t* out)
        {
        return guard(
            [&]
            {
                NONNULL(dev);
                NONNULL(in);
                NONNULL(out);
                DeviceObj& D = dev->obj;
                DBuf<uint32_t> din, dout;
                din.ensure(6 * (size_t)n);
                dout.ensure(4 * (size_t)n);
                HIPCHECK(hipMemcpy(din.p, in, 24 * (size_t)n, hipMemcpyHostToDevice));
                k_debug_philox<<<blocks(n, 256), 256, 0, D.stream>>>(din.p, n, dout.p);
                HIPCHECK(hipGetLastError());
                HIPCHECK(hipStreamSynchronize(D.stream));
                HIPCHECK(hipMemcpy(out, dout.p, 16 * (size_t)n, hipMemcpyDeviceToHost));
            });
        }
    int fh_debug_sincos(fh_device dev, const float* x, uint32_t n, float* s, float* c)
        {
        return guard(
            [&]
            {
                NONNULL(dev);
                NONNULL(x);
                NONNULL(s);
                NONNULL(c);
                DeviceObj& D = dev->obj;
                DBuf<float> dx, ds, dc;
                dx.ensure(n);
                ds.ensure(n);
                dc.ensure(n);
                HIPCHECK(hipMemcpy(dx.p, x, 4 * (size_t)n, hipMemcpyHostToDevice));
                k_debug_sincos<<<blocks(n, 256), 256, 0, D.stream>>>(dx.p, n, ds.p, dc.p);
                HIPCHECK(hipGetLastError());
                HIPCHECK(hipStreamSynchronize(D.stream));
                HIPCHECK(hipMemcpy(s, ds.p, 4 * (size_t)n, hipMemcpyDeviceToHost));
                HIPCHECK(hipMemcpy(c, dc.p, 4 * (size_t)n, hipMemcpyDeviceToHost));
            });
        }
    int fh_debug_primary(fh_tracer tr, fh_scene scene, uint32_t sample, float* rays, uint32_t* geom_id, uint32_t* prim,
                         float* t)
        {
        return guard(
            [&]
            {
                NONNULL(tr);
                NONNULL(scene);
                DeviceObj& D = *tr->obj.dev;
                std::lock_guard<std::mutex> lock(D.mu);
                sync_scene(scene->obj);
                JobDev J = make_job(scene->obj, tr->obj, scene->obj.camera);
                J.K = 1;
                const uint32_t npix = J.npix;
                DBuf<JobDev> dj;
                dj.ensure(1);
                DBuf<float> drays;
                drays.ensure(6 * (size_t)npix);
                DBuf<uint4> dhits;
                dhits.ensure(npix);
                HIPCHECK(hipMemcpy(dj.p, &J, sizeof(J), hipMemcpyHostToDevice));
                k_debug_primary<<<blocks(npix, 64), 64, 0, D.stream>>>(dj.p, sample, drays.p, dhits.p);
                HIPCHECK(hipGetLastError());
                HIPCHECK(hipStreamSynchronize(D.stream));
                std::vector<uint4> h(npix);
                HIPCHECK(hipMemcpy(h.data(), dhits.p, npix * sizeof(uint4), hipMemcpyDeviceToHost));
                if (rays)
                    HIPCHECK(hipMemcpy(rays, drays.p, 24 * (size_t)npix, hipMemcpyDeviceToHost));
                for (uint32_t p = 0; p < npix; p++)
                    {
                    if (h[p].w)
                        throw Error(FH_ERR_STATE, "traversal stack overflow");
                    if (geom_id)
                        geom_id[p] = h[p].y;
                    if (prim)
                        prim[p] = h[p].z;
                    if (t)
                        std::memcpy(&t[p], &h[p].x, 4);
                    }
            });
        }
    int fh_debug_primary_render(fh_tracer tr, fh_scene scene, uint32_t sample, uint32_t* geom_id, uint32_t* prim,
                                float* t)
        {
        return guard(
            [&]
            {
                NONNULL(tr);
                NONNULL(scene);
                REQUIRE(sample >= 1, "sample numbers start at 1");
                TracerObj& T = tr->obj;
                SceneObj& sc = scene->obj;
                DeviceObj& D = *T.dev;
                REQUIRE(sc.dev == T.dev, "Scene and Tracer devices do not match");
                std::lock_guard<std::mutex> lock(D.mu);
                sync_scene(sc);
                update_tiles(D, sc, T, make_cam(sc.camera), (uint64_t)T.W * T.H);
                JobDev J = make_job(sc, T, sc.camera);
                J.K = 1;
                J.n_base = sample - 1;
                set_order_divisors(J);
                const uint32_t npix = J.npix;
                RenderScratch& R = D.rs;
                R.jobs.ensure(1);
                R.hits.ensure(npix);
                R.counters32.ensure(4);
                const uint32_t gstride = D.trav_grid * TRAV_BLOCK;
                R.gstack.ensure((size_t)(STACK_MAX - 8) * gstride);
                D.begin_stats();
                HIPCHECK(hipMemcpyAsync(R.jobs.p, &J, sizeof(J), hipMemcpyHostToDevice, D.stream));
                HIPCHECK(hipMemsetAsync(R.counters32.p, 0, sizeof(uint32_t), D.stream));
                TravArgs a {R.jobs.p, 1, Queue {}, npix, nullptr, R.hits.p, R.counters32.p, D.trav_chunk, R.gstack.p,
                            gstride, D.counters.p};
                dispatch_traverse(D, "traverse_primary", D.stream, a, true,
                                  (scene_types(sc) & ~T_SPHERES) == 0);
                D.end_stats();
                if (D.last_counters[CTR_OVERFLOW])
                    throw Error(FH_ERR_STATE, "traversal stack overflow");
                std::vector<uint4> h(npix);
                HIPCHECK(hipMemcpy(h.data(), R.hits.p, npix * sizeof(uint4), hipMemcpyDeviceToHost));
                // ray j of the batch belongs to the pixel primary_order gives (one sample)
                const uint32_t W8 = T.W & ~7u, H8 = T.H & ~7u;
                for (uint32_t j = 0; j < npix; j++)
                    {
                    uint32_t px, py;
                    if (j < W8 * H8)
                        {
                        const uint32_t tile = j >> 6, within = j & 63u, ty = tile / (W8 >> 3);
                        px = (tile - ty * (W8 >> 3)) * 8u + (within & 7u);
                        py = ty * 8u + (within >> 3);
                        }
                    else
                        {
                        uint32_t r = j - W8 * H8;
                        const uint32_t rw = T.W - W8;
                        if (r < rw * H8)
                            {
                            py = r / rw;
                            px = W8 + (r - py * rw);
                            }
                        else
                            {
                            r -= rw * H8;
                            py = H8 + r / T.W;
                            px = r % T.W;
                            }
                        }
                    const uint32_t p = py * T.W + px;
                    if (geom_id)
                        geom_id[p] = h[j].y;
                    if (prim)
                        prim[p] = h[j].z;
                    if (t)
                        std::memcpy(&t[p], &h[j].x, 4);
                    }
            });
        }
    int fh_debug_bvh(fh_scene scene, void* nodes, uint64_t cap, uint64_t* n_nodes)
        {
        return guard(
            [&]
            {
                NONNULL(scene);
                NONNULL(n_nodes);
                SceneObj& sc = scene->obj;
                DeviceObj& D = *sc.dev;
                std::lock_guard<std::mutex> lock(D.mu);
                sync_scene(sc);
                HIPCHECK(hipStreamSynchronize(D.stream)); // the hierarchy build has completed
                const uint64_t n = sc.n_prims == 0 ? 0 : std::max<uint32_t>(sc.n_prims - 1, 1);
                *n_nodes = n;
                if (nodes && n)
                    HIPCHECK(hipMemcpy(nodes, sc.nodes.p, std::min(n, cap) * sizeof(Node), hipMemcpyDeviceToHost));
            });
        }
    int fh_debug_trace(fh_scene scene, const float* rays, uint32_t n, float tnear, uint32_t* geom_id, uint32_t* prim,
                       float* t)
        {
        return guard(
            [&]
            {
                NONNULL(scene);
                NONNULL(rays);
                DeviceObj& D = *scene->obj.dev;
                std::lock_guard<std::mutex> lock(D.mu);
                sync_scene(scene->obj);
                DBuf<float> drays;
                drays.ensure(6 * (size_t)n);
                DBuf<uint4> dhits;
                dhits.ensure(n);
                HIPCHECK(hipMemcpy(drays.p, rays, 24 * (size_t)n, hipMemcpyHostToDevice));
                k_debug_trace<<<blocks(n, 64), 64, 0, D.stream>>>(scene_dev(scene->obj), drays.p, n, tnear, dhits.p);
                HIPCHECK(hipGetLastError());
                HIPCHECK(hipStreamSynchronize(D.stream));
                std::vector<uint4> h(n);
                HIPCHECK(hipMemcpy(h.data(), dhits.p, n * sizeof(uint4), hipMemcpyDeviceToHost));
                for (uint32_t p = 0; p < n; p++)
                    {
                    if (h[p].w)
                        throw Error(FH_ERR_STATE, "traversal stack overflow");
                    if (geom_id)
                        geom_id[p] = h[p].y;
                    if (prim)
                        prim[p] = h[p].z;
                    if (t)
                        std::memcpy(&t[p], &h[p].x, 4);
                    }
            });
        }
    int fh_stats_enable(fh_device dev, int enable)
        {
        return guard(
            [&]
            {
                NONNULL(dev);
                dev->obj.stats_on = enable != 0;
            });
        }
    int fh_stats_last(fh_device dev, char* names, size_t names_cap, double* ms, uint32_t* launches, uint32_t cap,
                      uint32_t* n_kernels)
        {
        return guard(
            [&]
            {
                NONNULL(dev);
                std::string all;
                uint32_t k = 0;
                for (auto& s : dev->obj.last_stats)
                    {
                    if (k < cap)
                        {
                        if (ms)
                            ms[k] = s.second.first;
                        if (launches)
                            launches[k] = s.second.second;
                        }
                    all += s.first + "\n";
                    k++;
                    }
                copy_str(all, names, names_cap);
                if (n_kernels)
                    *n_kernels = k;
            });
        }
    int fh_stats_graphs(fh_device dev, uint64_t* captures, uint64_t* replays)
        {
        return guard(
            [&]
            {
                NONNULL(dev);
                if (captures)
                    *captures = dev->obj.graph_captures;
                if (replays)
                    *replays = dev->obj.graph_replays;
            });
        }
    int fh_counters_enable(fh_device dev, int enable)
        {
        return guard(
            [&]
            {
                NONNULL(dev);
                dev->obj.count_on = enable != 0;
            });
        }
    int fh_counters_last(fh_device dev, uint64_t out[16])
        {
        return guard(
            [&]
            {
                NONNULL(dev);
                NONNULL(out);
                for (int k = 0; k < N_CTR; k++)
                    out[k] = dev->obj.last_counters[k];
            });
        }
    }
