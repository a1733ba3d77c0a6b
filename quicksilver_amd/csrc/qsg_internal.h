// Internals shared by the C-ABI translation units (not part of the ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <stdexcept>
#include <string>
#include <vector>

#include "qs_device.h"
#include "qs_params.h"
#include "qs_problem.h"
#include "qsg.h"

namespace qsgi {

struct QsgError : std::runtime_error {
    int code;
    QsgError(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

inline void hipCheck(hipError_t e, const char* what) {
    if (e != hipSuccess) throw QsgError(QSG_ERR_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

struct PBuf {
    void* mem = nullptr;
    qs::Particles v = {};
    uint32_t cap = 0;
};

// Records an error message for qsg_last_error() and returns its code.
int fail(const std::exception& e);

// The device error flags (PassCounters::error, BatchState::error): bit 0
// census overflow, 1 secondary overflow, 2 population-control overflow,
// 3 runaway history, 4 more than 4 fission products, 5 census + source over
// the processing buffer.  Their text and the ABI code they map to.
std::string deviceErrorText(uint32_t flags);
int deviceErrorCode(uint32_t flags);

// Balance::Add (Tallies.hh:84-99): every counter summed.
inline void addBalance(qsg_balance& acc, const qsg_balance& b) {
    acc.start += b.start;
    acc.source += b.source;
    acc.rr += b.rr;
    acc.split += b.split;
    acc.absorb += b.absorb;
    acc.scatter += b.scatter;
    acc.fission += b.fission;
    acc.produce += b.produce;
    acc.collision += b.collision;
    acc.escape += b.escape;
    acc.census += b.census;
    acc.num_segments += b.num_segments;
    acc.end += b.end;
}

// Streams and small pinned host blocks are handed back to process-wide free
// lists when a problem is released and reused by the next one: creating and
// destroying a stream, or pinning and unpinning host memory, costs
// milliseconds, more than building a small problem.  Thread-safe; what the
// lists hold lives until the process ends.
hipStream_t acquireStream(int device);   // non-blocking stream on `device` (current)
void releaseStream(int device, hipStream_t s);  // s must be idle
constexpr size_t kPinnedBlock = 256;     // bytes per pinned block
void* acquirePinned(size_t bytes);       // bytes <= kPinnedBlock
void releasePinned(void* p);

// Device memory blocks are handed back when a problem is released and
// reused by the next allocation of a similar size on the same device:
// hipMalloc and hipFree cost tens of microseconds each, and a problem makes
// about twenty allocations, which a served stream of small problems pays on
// its critical path.  A cached block serves a request of at least two
// thirds of its size; at most kDeviceCacheBytes of idle blocks are kept per
// device, and an allocation that fails releases the device's idle blocks
// and tries once more.  Thread-safe.  Whoever hands a block back makes sure
// that no work still uses it.  A reused block keeps its old contents.
constexpr size_t kDeviceCacheBytes = 64ull << 30;
// `bytes` on the current device; *blockBytes receives the block's size,
// which deviceRelease takes back with it.
void* deviceAlloc(int device, size_t bytes, size_t* blockBytes);
void deviceRelease(int device, void* p, size_t blockBytes);

// Device memory for the per-wave secondary stacks of tracking launches,
// grown on demand; the owner's stream is synchronised before memory that a
// launch may still use is freed.
struct StackPool {
    char* mem = nullptr;
    size_t bytes = 0;
    // Stacks for a launch of nWaves waves with cap entries per wave.
    qs::WaveStacks ensure(uint32_t nWaves, uint32_t cap, hipStream_t stream);
    void release();
};

}  // namespace qsgi

struct qsg_problem_s {
    using PBuf = qsgi::PBuf;
    int device = 0;
    hipStream_t stream = nullptr;
    int numCUs = 0;
    qs::Parameters params;
    qs::ProblemTables tables;
    qs::DevProblem dp = {};
    std::vector<std::pair<void*, size_t>> allocs;  // device blocks and their sizes
    size_t deviceBytes = 0;

    PBuf proc[2], pop, extra[2];
    int cur = 0;               // proc[cur] holds this cycle's processing population
    uint64_t startCount = 0;   // census particles carried into this cycle
    uint32_t nPop = 0;         // processing population after population control
    uint32_t nCensus = 0;      // census count after tracking
    int censusBuf = -1;        // proc index holding the last census
    int phase = 0;             // 0 ready for init, 1 init done, 2 tracking done
    int cycle = 0;
    // this cycle's processing population came from the caller
    // (qsg_cycle_begin_host): no device source or population control ran
    bool hostCycle = false;
    uint64_t hostCount = 0;

    uint32_t* sortCounts = nullptr;     // [nCells * kSortBands]
    uint32_t* sortBlockSums = nullptr;  // scan scratch
    qs::PopGroup* popGroups = nullptr;  // population-control scratch
    bool queueSorted = false;           // this cycle's queue is proc[cur] (sorted), not pop

    qs::CycleBlock* cb = nullptr;     // device counters: pc, dc and ctl point into it
    qs::PassCounters* pc = nullptr;
    qs::DevCounters* dc = nullptr;
    qs::CycleCtl* ctl = nullptr;
    qs::CycleBlock* hcb[2] = {};      // pinned host copies (two cycles in flight in qsg_run)
    qs::PassCounters* hpc = nullptr;  // = &hcb[0]->pc, the copy syncCounters fills
    qs::DevCounters* hdc = nullptr;   // = &hcb[0]->dc

    qs::TrackTuning tune = qs::kDefaultTuning;
    qsgi::StackPool stacks;
    hipEvent_t ev0 = nullptr, ev1 = nullptr, evI0 = nullptr, evI1 = nullptr, evF0 = nullptr, evF1 = nullptr;
    std::vector<hipEvent_t> passEv;  // begin/end pairs of tracking-pass launches
    // events of the cycles qsg_run has in flight (a ring): before and after
    // the first pass, and after the last pass of a cycle that needed more;
    // each recorded event costs the stream about 10 us, so there are no more
    struct RunEvents {
        hipEvent_t t0, p1, t1;
    };
    RunEvents runEv[4] = {};
    hipEvent_t runStart = nullptr, runEnd = nullptr;
    double lastTrackingSeconds = 0.0;
    qsg_stats stats = {};
    qsg_balance cumulative = {};  // Tallies::_balanceCumulative (Tallies.cc:65)
    // set when a cycle stopped part way: the device state is no longer a
    // cycle boundary and the problem takes no further cycle
    int failCode = 0;
    std::string failMsg;

    qs::LaunchCtx ctx() const { return qs::LaunchCtx{(void*)stream, numCUs}; }

    template <typename T>
    T* dalloc(size_t count) {
        size_t bytes = count * sizeof(T);
        if (bytes == 0) bytes = 8;
        size_t block = 0;
        void* p = qsgi::deviceAlloc(device, bytes, &block);
        allocs.emplace_back(p, block);
        deviceBytes += bytes;
        return static_cast<T*>(p);
    }
    template <typename T>
    T* upload(const std::vector<T>& h) {
        T* d = dalloc<T>(h.size());
        if (!h.empty())
            qsgi::hipCheck(hipMemcpy(d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice), "hipMemcpy");
        return d;
    }
};

namespace qsgi {
// qsg_tuning <-> launch shape, with range checks (throws QSG_ERR_ARG).
qs::TrackTuning tuningFromC(const qsg_tuning* t);
void tuningToC(const qs::TrackTuning& in, qsg_tuning* out);
// Particle-buffer capacity (entries; `requested` 0 selects the default,
// 2 x nParticles + one cycle's source + 65536) and the total device bytes a
// problem will allocate with it.
uint32_t particleCapacity(const qs::ProblemTables& T, uint64_t requested = 0);
uint64_t deviceBytesEstimate(const qs::ProblemTables& T, uint64_t requested = 0);
// Uploads already-built tables and allocates the device state.
// ownStacks: allocate wave stacks for the problem's own tracking launches
// (problems that only run inside a batch use the batch's stacks).
qsg_problem_s* createFromTables(const qs::Parameters& params, qs::ProblemTables tables, int device,
                                bool ownStacks, uint64_t capacity = 0);
// Marks p failed (first failure wins) with the error's code and message.
void markFailed(qsg_problem_s* p, int code, const std::string& msg);
// Throws QSG_ERR_STATE if p failed earlier.
void requireUsable(const qsg_problem_s* p);
void destroy(qsg_problem_s* p);
}  // namespace qsgi

// Batch internals used by the serving loop.
// A batch with `capacity` slots, the first n holding problems[0..n) (the
// rest empty); pool: stacks shared with the caller (NULL: the batch owns its
// stacks).  Members are replaced in place with qsgiBatchSetMember (NULL
// empties a slot) at a cycle boundary; captured graphs stay valid.
qsg_batch_s* qsgiCreateBatch(qsg_problem* problems, int n, int capacity, int device, qsgi::StackPool* pool);
void qsgiBatchSetMember(qsg_batch_s* b, int slot, qsg_problem_s* p);
int qsgiBatchCapacity(const qsg_batch_s* b);
void qsgiDestroyBatch(qsg_batch_s* b);
void qsgiBatchStep(qsg_batch_s* b, qsg_balance* balances, double* fluxSums);
void qsgiBatchSetTuning(qsg_batch_s* b, const qs::TrackTuning& t);
const qsg_stats& qsgiBatchStats(const qsg_batch_s* b);
// The error a slot's problem was taken out of the batch with (0: none).
int qsgiBatchSlotError(const qsg_batch_s* b, int slot);
// Whether a HIP failure ended the batch (no further step runs).
bool qsgiBatchBroken(const qsg_batch_s* b);

#define QSG_GUARD(...)                   \
    try {                                \
        __VA_ARGS__;                     \
        return QSG_OK;                   \
    } catch (const std::exception& e) {  \
        return qsgi::fail(e);            \
    } catch (...) {                      \
        return qsgi::fail(std::runtime_error("unknown error")); \
    }
