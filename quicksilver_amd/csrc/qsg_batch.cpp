// Batched form of the C ABI: many independent problems advance one cycle
// per step through shared kernel launches (device source, population
// control, tracking passes, flux reduction).  Each problem keeps its own
// buffers, tallies and RNG streams; the batch only shares launches, so a
// problem's results are the same alone or in any mix.
#include <cstring>
#include <vector>

#include "qsg_internal.h"

using qsgi::hipCheck;
using qsgi::QsgError;

struct qsg_batch_s {
    int device = 0;
    hipStream_t stream = nullptr;
    int numCUs = 0;
    std::vector<qsg_problem_s*> probs;
    qs::BatchSlot* dSlots = nullptr;
    qs::BatchState* dState = nullptr;
    qs::DevCounters* dCounters = nullptr;
    uint64_t* dSrcPrefix = nullptr;
    uint64_t* dPopPrefix = nullptr;
    uint64_t* dKeyBase = nullptr;     // [n + 1] prefix of the problems' sort keys
    uint64_t* dSortPrefix = nullptr;  // [n + 1]
    uint32_t* dSortCounts = nullptr;
    uint32_t* dSortBlockSums = nullptr;
    size_t sortKeys = 0;
    size_t sortKeysCap = 0;
    bool graphsValid = false;  // cleared when memory the graphs hold is reallocated
    // every member's energy-group bounds fit the tracking waves' LDS slots
    // (qs::kBatchStagedEnergies); the captured graphs hold the kernel chosen
    // for it
    bool stagedEnergies = true;
    double* dFluxSums = nullptr;
    int partsPerProblem = 64;  // blocks per problem for zeroing its flux bins
    qs::BatchState* hState = nullptr;  // pinned mirrors
    qs::DevCounters* hCounters = nullptr;
    double* hFluxSums = nullptr;
    std::vector<uint32_t> active;
    std::vector<int> expectedCycle;  // each problem's cycle as the batch left it
    std::vector<int> slotError;      // code a slot's problem was taken out with (0: none)
    // set when a HIP failure ends a step: the members' device state is then
    // unknown, so no further step runs
    bool broken = false;
    std::string brokenMsg;
    qs::TrackTuning tune = qs::kDefaultTuning;
    qsgi::StackPool ownStacks;
    qsgi::StackPool* stacks = &ownStacks;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    std::vector<hipEvent_t> passEv;
    // captured launch sequences of a step (set-up up to the first tracking
    // pass, that pass, and the flux reduction on) and what they were
    // captured with
    hipGraphExec_t gSetup = nullptr, gTrack = nullptr, gFinal = nullptr;
    qs::TrackTuning graphTune = {};
    qs::WaveStacks graphStacks = {};
    qsg_stats stats = {};
};

namespace {

template <typename T>
T* devAlloc(size_t n) {
    void* p = nullptr;
    hipCheck(hipMalloc(&p, (n ? n : 1) * sizeof(T)), "hipMalloc");
    return static_cast<T*>(p);
}

template <typename T>
T* hostAlloc(size_t n) {
    void* p = nullptr;
    hipCheck(hipHostMalloc(&p, (n ? n : 1) * sizeof(T), hipHostMallocDefault), "hipHostMalloc");
    return static_cast<T*>(p);
}

void destroyBatch(qsg_batch_s* b) {
    if (!b) return;
    (void)hipSetDevice(b->device);
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    for (void* d : {(void*)b->dSlots, (void*)b->dState, (void*)b->dCounters, (void*)b->dSrcPrefix,
                    (void*)b->dPopPrefix, (void*)b->dFluxSums, (void*)b->dKeyBase,
                    (void*)b->dSortPrefix, (void*)b->dSortCounts, (void*)b->dSortBlockSums})
        if (d) (void)hipFree(d);
    for (void* h : {(void*)b->hState, (void*)b->hCounters, (void*)b->hFluxSums})
        if (h) (void)hipHostFree(h);
    for (hipEvent_t e : {b->ev0, b->ev1})
        if (e) (void)hipEventDestroy(e);
    for (hipEvent_t e : b->passEv) (void)hipEventDestroy(e);
    for (hipGraphExec_t g : {b->gSetup, b->gTrack, b->gFinal})
        if (g) (void)hipGraphExecDestroy(g);
    b->ownStacks.release();
    if (b->stream) {
        (void)hipStreamSynchronize(b->stream);
        qsgi::releaseStream(b->device, b->stream);
    }
    delete b;
}

void checkMember(const qsg_batch_s* b, const qsg_problem_s* p) {
    qsgi::requireUsable(p);
    if (p->device != b->device) throw QsgError(QSG_ERR_ARG, "batch problems must share one device");
    if (p->phase != 0) throw QsgError(QSG_ERR_STATE, "batch problems must be at a cycle boundary");
    if (p->cur != (p->cycle & 1)) throw QsgError(QSG_ERR_STATE, "problem buffer parity does not match its cycle");
}

// Sort-key layout of the current members: keyBase[i] is the first key of
// slot i (empty slots have none).  Grows the key histogram when needed,
// which invalidates captured graphs (they hold its address).
void uploadKeyBase(qsg_batch_s* b) {
    const int n = (int)b->probs.size();
    std::vector<uint64_t> keyBase(n + 1, 0);
    bool staged = true;
    for (int i = 0; i < n; i++) {
        keyBase[i + 1] = keyBase[i] + (b->probs[i] ? qs::sortKeys(b->probs[i]->tables.nCells) : 0);
        if (b->probs[i] && b->probs[i]->tables.nGroups + 1 > qs::kBatchStagedEnergies) staged = false;
    }
    if (staged != b->stagedEnergies) {
        b->stagedEnergies = staged;
        b->graphsValid = false;
    }
    b->sortKeys = keyBase[n];
    if (b->sortKeys > b->sortKeysCap || !b->dSortCounts) {
        for (void* d : {(void*)b->dSortCounts, (void*)b->dSortBlockSums})
            if (d) (void)hipFree(d);
        b->sortKeysCap = b->sortKeys + b->sortKeys / 2 + 1024;
        b->dSortCounts = devAlloc<uint32_t>(b->sortKeysCap);
        b->dSortBlockSums = devAlloc<uint32_t>(qs::scanBlocks(b->sortKeysCap));
        b->graphsValid = false;
    }
    hipCheck(hipMemcpyAsync(b->dKeyBase, keyBase.data(), (n + 1) * sizeof(uint64_t), hipMemcpyHostToDevice,
                            b->stream),
             "hipMemcpyAsync");
    hipCheck(hipStreamSynchronize(b->stream), "hipStreamSynchronize");
}

// Puts problem p (or nothing) in slot i: its descriptor, buffer views and a
// fresh per-slot state continuing from the problem's own cycle.  The slot
// array's address and size do not change, so captured graphs stay valid.
void setMember(qsg_batch_s* b, int i, qsg_problem_s* p) {
    if (i < 0 || i >= (int)b->probs.size()) throw QsgError(QSG_ERR_ARG, "bad batch slot");
    qs::BatchSlot slot;
    qs::BatchState st;
    memset(&slot, 0, sizeof(slot));
    memset(&st, 0, sizeof(st));
    slot.st = b->dState + i;
    slot.dc = b->dCounters + i;
    if (p) {
        checkMember(b, p);
        // members track through the batch's stacks; a problem run alone
        // again reallocates its own on its next tracking pass
        hipCheck(hipStreamSynchronize(p->stream), "hipStreamSynchronize");
        p->stacks.release();
        slot.prob = p->dp;
        slot.proc[0] = p->proc[0].v;
        slot.proc[1] = p->proc[1].v;
        slot.pop = p->pop.v;
        slot.extra[0] = p->extra[0].v;
        slot.extra[1] = p->extra[1].v;
        slot.cap = p->pop.cap;
        slot.loadBalance = (uint32_t)p->tables.loadBalance;
        slot.sourcePerCycle = p->tables.sourcePerCycle;
        slot.nParticles = p->tables.nParticles;
        slot.groups = p->popGroups;
        st.nCensus = (uint32_t)p->startCount;
        st.cycle = (uint32_t)p->cycle;
        st.active = 1;
    }
    b->probs[i] = p;
    b->active[i] = p ? 1u : 0u;
    b->expectedCycle[i] = p ? p->cycle : 0;
    b->slotError[i] = 0;
    hipCheck(hipMemcpyAsync(b->dSlots + i, &slot, sizeof(slot), hipMemcpyHostToDevice, b->stream), "hipMemcpyAsync");
    hipCheck(hipMemcpyAsync(b->dState + i, &st, sizeof(st), hipMemcpyHostToDevice, b->stream), "hipMemcpyAsync");
    hipCheck(hipStreamSynchronize(b->stream), "hipStreamSynchronize");
    uploadKeyBase(b);
}

// A batch of `capacity` slots holding problems[0..n) (n <= capacity; the
// other slots start empty).
qsg_batch_s* createBatch(qsg_problem* problems, int n, int capacity, int device, qsgi::StackPool* pool) {
    if (n < 0 || (n > 0 && !problems) || capacity < n || capacity <= 0)
        throw QsgError(QSG_ERR_ARG, "a batch needs at least one slot and room for its problems");
    qsg_batch_s* b = new qsg_batch_s();
    if (pool) b->stacks = pool;
    try {
        b->device = n > 0 ? problems[0]->device : device;
        for (int i = 0; i < n; i++) {
            if (!problems[i]) throw QsgError(QSG_ERR_ARG, "NULL problem in batch");
            checkMember(b, problems[i]);
        }
        hipCheck(hipSetDevice(b->device), "hipSetDevice");
        b->stream = qsgi::acquireStream(b->device);
        hipCheck(hipDeviceGetAttribute(&b->numCUs, hipDeviceAttributeMultiprocessorCount, b->device),
                 "hipDeviceGetAttribute");
        hipCheck(hipEventCreate(&b->ev0), "hipEventCreate");
        hipCheck(hipEventCreate(&b->ev1), "hipEventCreate");
        if (n > 0) b->tune = problems[0]->tune;
        const int c = capacity;
        b->probs.assign(c, nullptr);
        b->active.assign(c, 0);
        b->expectedCycle.assign(c, 0);
        b->slotError.assign(c, 0);
        b->dSlots = devAlloc<qs::BatchSlot>(c);
        b->dState = devAlloc<qs::BatchState>(c);
        b->dCounters = devAlloc<qs::DevCounters>(c);
        b->dSrcPrefix = devAlloc<uint64_t>(c + 1);
        b->dPopPrefix = devAlloc<uint64_t>(c + 1);
        b->dKeyBase = devAlloc<uint64_t>(c + 1);
        b->dSortPrefix = devAlloc<uint64_t>(c + 1);
        b->dFluxSums = devAlloc<double>(c);
        b->hState = hostAlloc<qs::BatchState>(c);
        b->hCounters = hostAlloc<qs::DevCounters>(c);
        b->hFluxSums = hostAlloc<double>(c);
        hipCheck(hipMemset(b->dCounters, 0, c * sizeof(qs::DevCounters)), "hipMemset");
        for (int i = 0; i < c; i++) setMember(b, i, i < n ? problems[i] : nullptr);
        hipCheck(hipDeviceSynchronize(), "hipDeviceSynchronize");
    } catch (...) {
        destroyBatch(b);
        throw;
    }
    return b;
}

void readStates(qsg_batch_s* b) {
    const size_t n = b->probs.size();
    hipCheck(hipMemcpyAsync(b->hState, b->dState, n * sizeof(qs::BatchState), hipMemcpyDeviceToHost, b->stream),
             "hipMemcpyAsync");
    hipCheck(hipStreamSynchronize(b->stream), "hipStreamSynchronize");
}

// Problems whose device error flags are set (in the host copy of the
// states) are taken out of the batch: marked failed, made inactive on the
// device so that no later launch of this or any step touches them, and
// left out of the rest of the step.  Every other problem continues; the
// device kernels keep each problem's stores inside its own buffers, so a
// problem's overflow does not reach the others.  Returns how many were
// taken out; the first one's code and every message go to code / msg.
int takeOutFlagged(qsg_batch_s* b, int& code, std::string& msg) {
    static const uint32_t kInactive = 0u;
    int taken = 0;
    for (size_t i = 0; i < b->probs.size(); i++) {
        qs::BatchState& st = b->hState[i];
        qsg_problem_s* p = b->probs[i];
        if (!p || !st.active || !st.error) continue;
        const int c = qsgi::deviceErrorCode(st.error);
        const std::string m = "batch problem " + std::to_string(i) + ": " + qsgi::deviceErrorText(st.error);
        qsgi::markFailed(p, c, m);
        b->slotError[i] = c;
        b->active[i] = 0;
        st.active = 0;
        hipCheck(hipMemcpyAsync(&b->dState[i].active, &kInactive, sizeof(uint32_t), hipMemcpyHostToDevice,
                                b->stream),
                 "hipMemcpyAsync");
        if (!code) code = c;
        msg += (msg.empty() ? "" : " ") + m;
        taken++;
    }
    if (taken) hipCheck(hipStreamSynchronize(b->stream), "hipStreamSynchronize");
    return taken;
}

hipEvent_t passEvent(qsg_batch_s* b, int i) {
    while ((int)b->passEv.size() <= i) {
        hipEvent_t e;
        hipCheck(hipEventCreate(&e), "hipEventCreate");
        b->passEv.push_back(e);
    }
    return b->passEv[i];
}

// First part of a step: cycle set-up, source and population control of
// every problem, up to the first tracking pass.
void enqueueSetup(qsg_batch_s* b, const qs::LaunchCtx& c) {
    const int n = (int)b->probs.size();
    qs::launchBatchPrep(c, b->dSlots, n, b->dSrcPrefix, b->dPopPrefix);
    qs::launchBatchSource(c, b->dSlots, n, b->dSrcPrefix);
    qs::launchBatchPopulation(c, b->dSlots, n, b->dPopPrefix);
    if (b->tune.sortInterval > 0)
        qs::launchBatchSort(c, b->dSlots, n, b->tune.sortInterval, b->dKeyBase, b->dSortPrefix, b->dSortCounts,
                            b->sortKeys, b->dSortBlockSums);
    qs::launchBatchPassPrep(c, b->dSlots, n, 0);
}

// The first tracking pass.  Captured as a graph of its own, so that events
// recorded between graph launches time the tracking kernel alone.
void enqueueFirstPass(qsg_batch_s* b, const qs::LaunchCtx& c, const qs::WaveStacks& ws) {
    qs::launchBatchTrack(c, b->dSlots, (int)b->probs.size(), 0, b->tune, ws, b->stagedEnergies);
    hipCheck(hipGetLastError(), "batch track launch");
}

// Second half: flux reduction and cycle advance, and the copies of the
// flux sums, tallies and states to the host.
void enqueueFinal(qsg_batch_s* b, const qs::LaunchCtx& c) {
    const int n = (int)b->probs.size();
    qs::launchBatchFluence(c, b->dSlots, n);
    qs::launchBatchFinalize(c, b->dSlots, n, b->partsPerProblem, b->dFluxSums);
    hipCheck(hipMemcpyAsync(b->hFluxSums, b->dFluxSums, n * sizeof(double), hipMemcpyDeviceToHost, b->stream),
             "hipMemcpyAsync");
    hipCheck(hipMemcpyAsync(b->hCounters, b->dCounters, n * sizeof(qs::DevCounters), hipMemcpyDeviceToHost,
                            b->stream),
             "hipMemcpyAsync");
    hipCheck(hipMemcpyAsync(b->hState, b->dState, n * sizeof(qs::BatchState), hipMemcpyDeviceToHost, b->stream),
             "hipMemcpyAsync");
}

template <class F>
hipGraphExec_t capture(qsg_batch_s* b, F enqueue) {
    hipCheck(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
    hipGraph_t g = nullptr;
    try {
        enqueue();
    } catch (...) {
        (void)hipStreamEndCapture(b->stream, &g);
        if (g) (void)hipGraphDestroy(g);
        throw;
    }
    hipCheck(hipStreamEndCapture(b->stream, &g), "hipStreamEndCapture");
    hipGraphExec_t e = nullptr;
    const hipError_t err = hipGraphInstantiate(&e, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    hipCheck(err, "hipGraphInstantiate");
    return e;
}

bool sameShape(const qs::TrackTuning& a, const qs::TrackTuning& t) {
    return a.blockSize == t.blockSize && a.wavesPerCU == t.wavesPerCU && a.refillMin == t.refillMin &&
           a.chunk == t.chunk && a.stackCap == t.stackCap && a.wavesPerSimd == t.wavesPerSimd &&
           a.sortInterval == t.sortInterval;
}

// The step's two graphs, captured again whenever the launch shape or the
// stack memory changed since they were captured.
void ensureGraphs(qsg_batch_s* b, const qs::LaunchCtx& c, const qs::WaveStacks& ws) {
    if (b->gSetup && b->gTrack && b->gFinal && b->graphsValid && sameShape(b->graphTune, b->tune) &&
        b->graphStacks.base == ws.base && b->graphStacks.cap == ws.cap && b->graphStacks.nWaves == ws.nWaves)
        return;
    for (hipGraphExec_t* g : {&b->gSetup, &b->gTrack, &b->gFinal})
        if (*g) {
            (void)hipGraphExecDestroy(*g);
            *g = nullptr;
        }
    b->gSetup = capture(b, [&] { enqueueSetup(b, c); });
    b->gTrack = capture(b, [&] { enqueueFirstPass(b, c, ws); });
    b->gFinal = capture(b, [&] { enqueueFinal(b, c); });
    b->graphTune = b->tune;
    b->graphStacks = ws;
    b->graphsValid = true;
}

// One cycle of every active problem (cycleInit + cycleTracking +
// cycleFinalize of main.cc:55-66, for all of them at once).  Normally two
// graph launches and two host synchronisations: the first tracking pass
// tracks secondaries from the waves' stacks, and further passes (launched
// one by one) are needed only for secondaries that overflowed a stack.
void stepMembers(qsg_batch_s* b, qsg_balance* balances, double* fluxSums, int& failCode, std::string& failMsg);

void batchStep(qsg_batch_s* b, qsg_balance* balances, double* fluxSums) {
    if (b->broken) throw QsgError(QSG_ERR_STATE, "the batch failed earlier (" + b->brokenMsg + ")");
    hipCheck(hipSetDevice(b->device), "hipSetDevice");
    const int n = (int)b->probs.size();
    for (int i = 0; i < n; i++)
        if (b->probs[i] && (b->probs[i]->phase != 0 || b->probs[i]->cycle != b->expectedCycle[i]))
            throw QsgError(QSG_ERR_STATE, "a batch problem was advanced outside the batch");
    int failCode = 0;
    std::string failMsg;
    try {
        stepMembers(b, balances, fluxSums, failCode, failMsg);
    } catch (const std::exception& e) {
        // a HIP failure part way through the step: every member's device
        // state is unknown
        (void)hipStreamSynchronize(b->stream);
        b->broken = true;
        b->brokenMsg = e.what();
        const QsgError* q = dynamic_cast<const QsgError*>(&e);
        for (int i = 0; i < n; i++)
            if (b->probs[i] && b->active[i]) qsgi::markFailed(b->probs[i], q ? q->code : QSG_ERR_DEVICE, e.what());
        throw;
    }
    // problems taken out during the step: the others advanced normally
    if (failCode) throw QsgError(failCode, failMsg);
}

void stepMembers(qsg_batch_s* b, qsg_balance* balances, double* fluxSums, int& failCode, std::string& failMsg) {
    const int n = (int)b->probs.size();
    qs::LaunchCtx c{(void*)b->stream, b->numCUs};
    const long long blocks = qs::batchTrackBlocks(c, b->tune);
    const qs::WaveStacks ws = b->stacks->ensure((uint32_t)(blocks * (b->tune.blockSize / 64)),
                                                (uint32_t)b->tune.stackCap, b->stream);
    passEvent(b, 1);
    const bool graphs = b->tune.graphs != 0;
    if (graphs) ensureGraphs(b, c, ws);

    hipCheck(hipEventRecord(b->ev0, b->stream), "hipEventRecord");
    if (graphs) hipCheck(hipGraphLaunch(b->gSetup, b->stream), "hipGraphLaunch");
    else enqueueSetup(b, c);
    hipCheck(hipEventRecord(b->passEv[0], b->stream), "hipEventRecord");
    if (graphs) hipCheck(hipGraphLaunch(b->gTrack, b->stream), "hipGraphLaunch");
    else enqueueFirstPass(b, c, ws);
    hipCheck(hipEventRecord(b->passEv[1], b->stream), "hipEventRecord");
    hipCheck(hipMemcpyAsync(b->hState, b->dState, n * sizeof(qs::BatchState), hipMemcpyDeviceToHost, b->stream),
             "hipMemcpyAsync");
    hipCheck(hipStreamSynchronize(b->stream), "hipStreamSynchronize");
    takeOutFlagged(b, failCode, failMsg);
    // further passes for secondaries that overflowed the stacks
    int k = 1;
    for (;;) {
        const int outIdx = (k % 2 == 1) ? 0 : 1;  // where pass k-1 wrote its secondaries
        uint64_t pending = 0;
        for (int i = 0; i < n; i++)
            if (b->hState[i].active) pending += b->hState[i].extra[outIdx];
        if (pending == 0) break;
        const int kind = (k % 2 == 1) ? 1 : 2;
        hipEvent_t e0 = passEvent(b, 2 * k), e1 = passEvent(b, 2 * k + 1);
        qs::launchBatchPassPrep(c, b->dSlots, n, kind);
        hipCheck(hipEventRecord(e0, b->stream), "hipEventRecord");
        qs::launchBatchTrack(c, b->dSlots, n, kind, b->tune, ws, b->stagedEnergies);
        hipCheck(hipGetLastError(), "batch track launch");
        hipCheck(hipEventRecord(e1, b->stream), "hipEventRecord");
        readStates(b);
        takeOutFlagged(b, failCode, failMsg);
        k++;
    }
    double trackSeconds = 0;
    for (int q = 0; q < k; q++) {
        float ms = 0;
        hipCheck(hipEventElapsedTime(&ms, b->passEv[2 * q], b->passEv[2 * q + 1]), "hipEventElapsedTime");
        trackSeconds += ms * 1e-3;
    }
    if (graphs) hipCheck(hipGraphLaunch(b->gFinal, b->stream), "hipGraphLaunch");
    else enqueueFinal(b, c);
    hipCheck(hipEventRecord(b->ev1, b->stream), "hipEventRecord");
    hipCheck(hipStreamSynchronize(b->stream), "hipStreamSynchronize");
    float stepMs = 0;
    hipCheck(hipEventElapsedTime(&stepMs, b->ev0, b->ev1), "hipEventElapsedTime");

    for (int i = 0; i < n; i++) {
        const qs::BatchState& st = b->hState[i];
        qsg_problem_s* p = b->probs[i];
        if (!p || !st.active) {
            if (balances) memset(&balances[i], 0, sizeof(qsg_balance));
            if (fluxSums) fluxSums[i] = 0.0;
            continue;
        }
        const qs::DevCounters& d = b->hCounters[i];
        {
            qsg_balance bal;
            bal.start = st.start;
            bal.source = p->tables.sourcePerCycle;
            bal.rr = d.rr;
            bal.split = d.split;
            bal.absorb = d.absorb;
            bal.scatter = d.scatter;
            bal.fission = d.fission;
            bal.produce = d.produce;
            bal.collision = d.collision;
            bal.escape = d.escape;
            bal.census = d.census;
            bal.num_segments = d.numSegments;
            bal.end = st.nCensus;
            qsgi::addBalance(p->cumulative, bal);
            if (balances) balances[i] = bal;
        }
        if (fluxSums) fluxSums[i] = b->hFluxSums[i];
        // the problem handle continues from the batch's state
        p->startCount = st.nCensus;
        p->nCensus = st.nCensus;
        p->cycle = (int)st.cycle;
        b->expectedCycle[i] = p->cycle;
        p->cur = p->cycle & 1;
        p->censusBuf = p->cur;
        p->stats.segments += d.numSegments;
        p->stats.collisions += d.collision;
        p->stats.facet_crossings += d.numSegments - d.collision - d.census;
        p->stats.particle_loads += st.start + p->tables.sourcePerCycle + d.split - d.rr + d.produce;
        p->stats.cycles += 1;
        b->stats.segments += d.numSegments;
        b->stats.collisions += d.collision;
        b->stats.facet_crossings += d.numSegments - d.collision - d.census;
        b->stats.particle_loads += st.start + p->tables.sourcePerCycle + d.split - d.rr + d.produce;
    }
    b->stats.track_kernel_seconds += trackSeconds;
    b->stats.track_launches += (uint64_t)k;
    b->stats.tracking_seconds += stepMs * 1e-3;
    b->stats.cycles += 1;
}

}  // namespace

qsg_batch_s* qsgiCreateBatch(qsg_problem* problems, int n, int capacity, int device, qsgi::StackPool* pool) {
    return createBatch(problems, n, capacity, device, pool);
}
void qsgiBatchSetMember(qsg_batch_s* b, int slot, qsg_problem_s* p) {
    hipCheck(hipSetDevice(b->device), "hipSetDevice");
    setMember(b, slot, p);
}
int qsgiBatchCapacity(const qsg_batch_s* b) { return (int)b->probs.size(); }
int qsgiBatchSlotError(const qsg_batch_s* b, int slot) {
    return (slot >= 0 && slot < (int)b->slotError.size()) ? b->slotError[slot] : 0;
}
bool qsgiBatchBroken(const qsg_batch_s* b) { return b->broken; }
void qsgiDestroyBatch(qsg_batch_s* b) { destroyBatch(b); }
void qsgiBatchStep(qsg_batch_s* b, qsg_balance* balances, double* fluxSums) { batchStep(b, balances, fluxSums); }
void qsgiBatchSetTuning(qsg_batch_s* b, const qs::TrackTuning& t) { b->tune = t; }
const qsg_stats& qsgiBatchStats(const qsg_batch_s* b) { return b->stats; }

extern "C" {

int qsg_batch_create(qsg_problem* problems, int n, qsg_batch* out) {
    QSG_GUARD({
        if (!out) throw QsgError(QSG_ERR_ARG, "out is NULL");
        *out = nullptr;
        if (n <= 0 || !problems) throw QsgError(QSG_ERR_ARG, "a batch needs at least one problem");
        *out = createBatch(problems, n, n, problems[0] ? problems[0]->device : 0, nullptr);
    })
}

int qsg_batch_destroy(qsg_batch b) {
    QSG_GUARD({ destroyBatch(b); })
}

int qsg_batch_step(qsg_batch b, qsg_balance* balances, double* flux_sums) {
    QSG_GUARD({
        if (!b) throw QsgError(QSG_ERR_ARG, "NULL batch");
        batchStep(b, balances, flux_sums);
    })
}

int qsg_batch_problem_error(qsg_batch b, int index, int* code) {
    QSG_GUARD({
        if (!b || !code || index < 0 || index >= (int)b->probs.size()) throw QsgError(QSG_ERR_ARG, "bad argument");
        *code = b->slotError[index];
    })
}

int qsg_batch_set_active(qsg_batch b, int index, int active) {
    QSG_GUARD({
        if (!b || index < 0 || index >= (int)b->probs.size()) throw QsgError(QSG_ERR_ARG, "bad batch index");
        if (active && !b->probs[index]) throw QsgError(QSG_ERR_STATE, "batch slot holds no problem");
        if (active && b->slotError[index])
            throw QsgError(QSG_ERR_STATE, "the slot's problem was taken out of the batch after an error");
        if (active) qsgi::requireUsable(b->probs[index]);
        hipCheck(hipSetDevice(b->device), "hipSetDevice");
        uint32_t a = active ? 1u : 0u;
        hipCheck(hipStreamSynchronize(b->stream), "hipStreamSynchronize");
        hipCheck(hipMemcpy(&b->dState[index].active, &a, sizeof(a), hipMemcpyHostToDevice), "hipMemcpy");
        b->active[index] = a;
    })
}

int qsg_batch_set_tuning(qsg_batch b, const qsg_tuning* t) {
    QSG_GUARD({
        if (!b || !t) throw QsgError(QSG_ERR_ARG, "NULL argument");
        b->tune = qsgi::tuningFromC(t);
    })
}

int qsg_batch_get_stats(qsg_batch b, qsg_stats* stats) {
    QSG_GUARD({
        if (!b || !stats) throw QsgError(QSG_ERR_ARG, "NULL argument");
        *stats = b->stats;
    })
}

}  // extern "C"
