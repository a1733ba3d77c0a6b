// C ABI (include/qsg.h) over the device path: problem construction, buffer
// management and the per-cycle orchestration that replaces main.cc's
// cycleInit / cycleTracking / cycleFinalize for one problem on one GPU.
//
// Particle storage per problem (structure of arrays, 92 B / particle):
//   proc[2]   processing population of a cycle / census output of the cycle
//             (roles swap at the end of each cycle, like the reference's
//             processing/processed vaults, ParticleVaultContainer.cc:278)
//   pop       population-control output = input of the first tracking pass
//   extra[2]  secondaries of tracking pass k (ping-pong), input of pass k+1
//             (the reference's extra vaults, ParticleVaultContainer.cc:341-400)
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <map>
#include <mutex>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "qs_device.h"
#include "qs_params.h"
#include "qs_problem.h"
#include "qsg.h"

#include "qsg_internal.h"

using qsgi::hipCheck;
using qsgi::PBuf;
using qsgi::QsgError;

namespace {
thread_local std::string g_lastError;
}  // namespace

int qsgi::fail(const std::exception& e) {
    g_lastError = e.what();
    if (const QsgError* q = dynamic_cast<const QsgError*>(&e)) return q->code;
    return QSG_ERR_ARG;
}

namespace {

// Carve the 12 SoA arrays out of one allocation at a fixed stride.
PBuf makeParticles(qsg_problem_s* p, uint32_t cap) {
    PBuf b;
    b.cap = cap;
    const size_t n = cap ? cap : 1;
    const size_t stride = ((n * 8 + 255) / 256) * 256;
    char* m = p->dalloc<char>(12 * stride);
    b.mem = m;
    b.v.base = m;
    b.v.stride = stride;
    return b;
}

void copyParticles(const qs::Particles& d, const qsg_particles* h, uint64_t n, hipMemcpyKind kind,
                   hipStream_t s) {
    auto cp = [&](void* dst, const void* src, size_t bytes) {
        if (bytes) hipCheck(hipMemcpyAsync(dst, src, bytes, kind, s), "hipMemcpyAsync");
    };
    const bool up = kind == hipMemcpyHostToDevice;
    const size_t n8 = (size_t)n * 8, n4 = (size_t)n * 4;
    if (up) {
        cp(d.x(), h->x, n8); cp(d.y(), h->y, n8); cp(d.z(), h->z, n8);
        cp(d.vx(), h->vx, n8); cp(d.vy(), h->vy, n8); cp(d.vz(), h->vz, n8);
        cp(d.E(), h->energy, n8); cp(d.w(), h->weight, n8); cp(d.tc(), h->time_to_census, n8);
        cp(d.nmfp(), h->num_mean_free_paths, n8); cp(d.seed(), h->seed, n8); cp(d.cell(), h->cell, n4);
    } else {
        cp(h->x, d.x(), n8); cp(h->y, d.y(), n8); cp(h->z, d.z(), n8);
        cp(h->vx, d.vx(), n8); cp(h->vy, d.vy(), n8); cp(h->vz, d.vz(), n8);
        cp(h->energy, d.E(), n8); cp(h->weight, d.w(), n8); cp(h->time_to_census, d.tc(), n8);
        cp(h->num_mean_free_paths, d.nmfp(), n8); cp(h->seed, d.seed(), n8); cp(h->cell, d.cell(), n4);
    }
    hipCheck(hipStreamSynchronize(s), "hipStreamSynchronize");
}

void setDevice(qsg_problem_s* p) { hipCheck(hipSetDevice(p->device), "hipSetDevice"); }

// Runs device work of a cycle; if it throws, the problem's device state is
// no longer a cycle boundary, so the problem is marked failed first.
template <class F>
void deviceStage(qsg_problem_s* p, F work) {
    try {
        work();
    } catch (const QsgError& e) {
        qsgi::markFailed(p, e.code, e.what());
        throw;
    } catch (const std::exception& e) {
        qsgi::markFailed(p, QSG_ERR_DEVICE, e.what());
        throw;
    }
}

// Events that only time launches: recording one without the system-scope
// fence (cache write-back and invalidation) leaves no gap of several
// microseconds in the stream.  Nothing that the host reads depends on them:
// counters come back through copies the host synchronises with, or through
// the cycle tail, which fences its own stores to host memory.
constexpr unsigned kTimingEvent = hipEventDisableSystemFence;

// Enqueues the copy of every counter into the pinned host block `slot`.
void copyCounters(qsg_problem_s* p, int slot) {
    hipCheck(hipMemcpyAsync(p->hcb[slot], p->cb, sizeof(qs::CycleBlock), hipMemcpyDeviceToHost, p->stream),
             "hipMemcpyAsync");
}

// Every counter into hpc / hdc, after everything enqueued so far.
void syncCounters(qsg_problem_s* p) {
    copyCounters(p, 0);
    hipCheck(hipStreamSynchronize(p->stream), "hipStreamSynchronize");
}

void checkErrors(uint32_t e) {
    if (e) throw QsgError(qsgi::deviceErrorCode(e), qsgi::deviceErrorText(e));
}
void checkErrors(qsg_problem_s* p) { checkErrors(p->hpc->error); }

void build(qsg_problem_s* p, int device, uint64_t capacity) {
    p->device = device;
    setDevice(p);
    p->stream = qsgi::acquireStream(device);
    hipCheck(hipDeviceGetAttribute(&p->numCUs, hipDeviceAttributeMultiprocessorCount, device),
             "hipDeviceGetAttribute");
    for (hipEvent_t* e : {&p->ev0, &p->ev1, &p->evI0, &p->evI1, &p->evF0, &p->evF1})
        hipCheck(hipEventCreate(e), "hipEventCreate");

    const qs::ProblemTables& T = p->tables;
    if ((int)T.materials.size() > qs::kMaxMaterials)
        throw QsgError(QSG_ERR_ARG, "too many materials for the device tables");
    if (T.maxReactions > qs::kMaxReactions)
        throw QsgError(QSG_ERR_ARG, "too many reactions per material for the device tables");

    qs::DevProblem& d = p->dp;
    d.nx = T.nx;
    d.ny = T.ny;
    d.nz = T.nz;
    d.nCells = T.nCells;
    d.rNx = 1.0 / T.nx;
    d.rNxy = 1.0 / ((double)T.nx * T.ny);
    d.dx = T.dx;
    d.dy = T.dy;
    d.dz = T.dz;
    d.hdx = T.dx / 2.0;
    d.hdy = T.dy / 2.0;
    d.hdz = T.dz / 2.0;
    d.boundaryMask = 0;
    for (int f = 0; f < 6; f++) d.boundaryMask |= (T.boundary[f] & 3) << (2 * f);
    d.planes = p->upload(T.planes);
    d.cellMat = p->upload(T.cellMaterial);
    d.cellVolume = p->upload(T.cellVolume);
    d.cellCenter = p->upload(T.cellCenter);
    d.srcOffset = p->upload(T.cellSourceOffset);
    d.G = T.nGroups;
    d.nMat = (int)T.materials.size();
    d.maxR = T.maxReactions;
    d.energies = p->upload(T.energies);
    // energy_group's first guess: the bounds are exp(logLow + i * delta),
    // delta = (logHigh - logLow) / (G + 1) (NuclearData.cc:105-119), so the
    // group of E is about (log2 E - log2 e[0]) * ln 2 / delta
    {
        const double logLow = std::log(T.energies.front()), logHigh = std::log(T.energies.back());
        d.eLog2Lo = (float)(logLow / std::log(2.0));
        d.eGroupsPerLog2 = logHigh > logLow ? (float)(std::log(2.0) * (T.nGroups + 1.0) / (logHigh - logLow)) : 0.0f;
    }
    std::vector<double> total((size_t)d.nMat * d.G), mfp((size_t)d.nMat * d.G),
        macro((size_t)d.nMat * d.maxR * d.G, 0.0);
    std::vector<int32_t> react((size_t)d.nMat * d.maxR, 0);
    for (int m = 0; m < d.nMat; m++) {
        const qs::MaterialTables& mt = T.materials[m];
        for (int g = 0; g < d.G; g++) {
            // MC_Segment_Outcome.cc:60-63: the mean free path, one IEEE
            // division per (material, group) instead of one per segment
            const double t = mt.total[g];
            total[(size_t)m * d.G + g] = t;
            mfp[(size_t)m * d.G + g] = (t == 0.0) ? 1.0e+75 : 1.0 / t;
        }
        for (int r = 0; r < mt.nReactions; r++) {
            react[(size_t)m * d.maxR + r] = mt.reactionType[r];
            for (int g = 0; g < d.G; g++)
                macro[((size_t)m * d.maxR + r) * d.G + g] = mt.macro[(size_t)r * d.G + g];
        }
        d.mat[m].nIso = mt.nIsotopes;
        d.mat[m].nReact = mt.nReactions;
        d.mat[m].mass = mt.mass;
        d.mat[m].invMass = 1.0 / mt.mass;
        d.mat[m].nuBar = mt.nuBar;
    }
    d.total = p->upload(total);
    d.mfp = p->upload(mfp);
    d.macro = p->upload(macro);
    d.reactType = p->upload(react);
    d.dt = T.dt;
    d.eMin = T.eMin;
    d.eMax = T.eMax;
    d.sourceWeight = T.sourceWeight;
    d.lowWeightCutoff = T.lowWeightCutoff;
    d.flux = p->dalloc<double>((size_t)T.nCells * T.nGroups);
    hipCheck(hipMemset(d.flux, 0, (size_t)T.nCells * T.nGroups * sizeof(double)), "hipMemset");
    p->sortCounts = p->dalloc<uint32_t>(qs::sortKeys(T.nCells));
    p->sortBlockSums = p->dalloc<uint32_t>(qs::scanBlocks(qs::sortKeys(T.nCells)));
    d.spectrum = nullptr;
    if (!p->params.simulationParams.energySpectrum.empty()) {
        d.spectrum = p->dalloc<unsigned long long>((size_t)T.nGroups + 1);
        hipCheck(hipMemset(d.spectrum, 0, ((size_t)T.nGroups + 1) * sizeof(unsigned long long)), "hipMemset");
    }
    d.fluence = nullptr;
    if (p->params.simulationParams.coralBenchmark) {
        d.fluence = p->dalloc<double>((size_t)T.nCells);
        hipCheck(hipMemset(d.fluence, 0, (size_t)T.nCells * sizeof(double)), "hipMemset");
    }

    // particle buffers
    uint32_t cap = qsgi::particleCapacity(T, capacity);
    p->proc[0] = makeParticles(p, cap);
    p->proc[1] = makeParticles(p, cap);
    p->pop = makeParticles(p, cap);
    p->extra[0] = makeParticles(p, cap);
    p->extra[1] = makeParticles(p, cap);

    p->popGroups = p->dalloc<qs::PopGroup>(qs::popGroups(cap));
    p->cb = p->dalloc<qs::CycleBlock>(1);
    hipCheck(hipMemset(p->cb, 0, sizeof(qs::CycleBlock)), "hipMemset");
    p->pc = &p->cb->pc;
    p->dc = &p->cb->dc;
    p->ctl = &p->cb->ctl;
    static_assert(sizeof(qs::CycleBlock) <= qsgi::kPinnedBlock, "the counter block must fit a pinned block");
    for (qs::CycleBlock*& h : p->hcb) h = static_cast<qs::CycleBlock*>(qsgi::acquirePinned(sizeof(qs::CycleBlock)));
    p->hpc = &p->hcb[0]->pc;
    p->hdc = &p->hcb[0]->dc;
    for (auto& r : p->runEv)
        for (hipEvent_t* e : {&r.t0, &r.p1, &r.t1})
            hipCheck(hipEventCreateWithFlags(e, kTimingEvent), "hipEventCreateWithFlags");
    for (hipEvent_t* e : {&p->runStart, &p->runEnd}) hipCheck(hipEventCreate(e), "hipEventCreate");
    // the uploads and clears above ran on the null stream, which the
    // problem's non-blocking stream does not wait for; other devices' and
    // other problems' work is not waited for
    hipCheck(hipStreamSynchronize(nullptr), "hipStreamSynchronize");
}

void destroyProblem(qsg_problem_s* p) {
    if (!p) return;
    (void)hipSetDevice(p->device);
    if (p->stream) (void)hipStreamSynchronize(p->stream);
    // no work on the device may still use a block handed back for reuse
    // (hipFree would have waited for the device too)
    if (!p->allocs.empty()) (void)hipDeviceSynchronize();
    for (const auto& a : p->allocs) qsgi::deviceRelease(p->device, a.first, a.second);
    p->stacks.release();
    for (qs::CycleBlock* h : p->hcb) qsgi::releasePinned(h);
    for (hipEvent_t e : {p->ev0, p->ev1, p->evI0, p->evI1, p->evF0, p->evF1})
        if (e) (void)hipEventDestroy(e);
    for (auto& r : p->runEv)
        for (hipEvent_t e : {r.t0, r.p1, r.t1})
            if (e) (void)hipEventDestroy(e);
    for (hipEvent_t e : {p->runStart, p->runEnd})
        if (e) (void)hipEventDestroy(e);
    for (hipEvent_t e : p->passEv) (void)hipEventDestroy(e);
    if (p->stream) qsgi::releaseStream(p->device, p->stream);
    delete p;
}

double elapsed(hipEvent_t a, hipEvent_t b) {
    float ms = 0;
    hipCheck(hipEventElapsedTime(&ms, a, b), "hipEventElapsedTime");
    return ms * 1e-3;
}

// Cycle set-up (cycleInit, main.cc:96-121) of cycle number `cycle` with
// processing buffer proc[cur]: the bookkeeping (with setUp, for p->startCount
// census particles carried in; else a cycle tail has done it), source,
// population control and low-weight roulette; every sort_interval cycles the
// queue is then sorted by (cell, energy band) into proc[cur], whose
// particles the population control has already consumed.  zeroFlux: the
// source launch also zeroes the previous cycle's flux bins.  Returns whether
// the queue was sorted (the first pass then reads proc[cur], else pop).
bool enqueueSetUp(qsg_problem_s* p, int cycle, int cur, bool setUp, bool zeroFlux) {
    const qs::ProblemTables& T = p->tables;
    const qs::LaunchCtx c = p->ctx();
    PBuf& proc = p->proc[cur];
    if (setUp)
        qs::launchCycleSetup(c, p->cb, (uint32_t)p->startCount, proc.cap, T.sourcePerCycle, T.nParticles,
                             T.loadBalance);
    qs::launchSource(c, p->dp, proc.v, p->ctl, T.sourcePerCycle, (uint32_t)cycle, zeroFlux ? p->dp.flux : nullptr,
                     zeroFlux ? (size_t)T.nCells * T.nGroups : 0);
    // PopulationControl (PopulationControl.cc:20-65) on one rank
    qs::launchPopulation(c, p->dp, proc.v, proc.cap, p->pop.v, p->pop.cap, p->cb, p->popGroups);
    const bool sorted = p->tune.sortInterval > 0 && (cycle % p->tune.sortInterval) == 0;
    if (sorted)
        qs::launchSort(c, p->dp, p->pop.v, &p->pc->nPopOut, proc.v, p->sortCounts, p->sortBlockSums, &p->ctl->hold);
    return sorted;
}

// Tracking pass k of the cycle whose processing buffer is proc[cur], over
// `in` (length *nIn on the device, at most nInHost): census particles go to
// proc[1 - cur], secondaries that do not fit a wave stack to extra[k & 1].
// Pass k > 0 first clears its queue head and secondary count (pass 0 uses
// those the set-up cleared).
void enqueuePass(qsg_problem_s* p, int k, int cur, qs::Particles in, const uint32_t* nIn, uint32_t nInHost) {
    const qs::LaunchCtx c = p->ctx();
    uint32_t* nOut = &p->pc->extraCount[k & 1];
    if (k > 0) {
        qs::launchClear(c, nOut, sizeof(uint32_t));
        qs::launchClear(c, &p->pc->head, sizeof(uint32_t));
    }
    PBuf& census = p->proc[1 - cur];
    PBuf& ex = p->extra[k & 1];
    const long long blocks = qs::trackPassBlocks(c, p->tune, nInHost);
    const qs::WaveStacks ws =
        p->stacks.ensure((uint32_t)(blocks * (p->tune.blockSize / 64)), (uint32_t)p->tune.stackCap, p->stream);
    qs::launchTrackPass(c, p->dp, p->tune, ws, in, nIn, nInHost, census.v, census.cap, ex.v, nOut, ex.cap, p->pc,
                        p->dc, &p->ctl->hold);
    hipCheck(hipGetLastError(), "track pass launch");
}

// Passes k, k + 1, ... (cycleTracking, main.cc:138-307), the first over
// `in`, each later one over the secondaries its predecessor left, until a
// pass leaves none.  One synchronisation per pass brings every counter back
// into hpc / hdc; `end` is recorded after each pass.  Returns the number of
// passes run.
int runPasses(qsg_problem_s* p, int k, int cur, qs::Particles in, const uint32_t* nIn, uint32_t nInHost,
              hipEvent_t end) {
    int q = 0;
    while (nInHost > 0) {
        while ((int)p->passEv.size() < 2 * (q + 1)) {
            hipEvent_t e;
            hipCheck(hipEventCreateWithFlags(&e, kTimingEvent), "hipEventCreateWithFlags");
            p->passEv.push_back(e);
        }
        hipCheck(hipEventRecord(p->passEv[2 * q], p->stream), "hipEventRecord");
        enqueuePass(p, k, cur, in, nIn, nInHost);
        hipCheck(hipEventRecord(p->passEv[2 * q + 1], p->stream), "hipEventRecord");
        hipCheck(hipEventRecord(end, p->stream), "hipEventRecord");  // the last pass's record stands
        syncCounters(p);
        checkErrors(p);
        const uint32_t next = p->hpc->extraCount[k & 1];
        if (next > p->extra[k & 1].cap) throw QsgError(QSG_ERR_CAPACITY, "secondary buffer overflow");
        in = p->extra[k & 1].v;
        nIn = &p->pc->extraCount[k & 1];
        nInHost = next;
        k++;
        q++;
    }
    for (int i = 0; i < q; i++) p->stats.track_kernel_seconds += elapsed(p->passEv[2 * i], p->passEv[2 * i + 1]);
    p->stats.track_launches += (uint64_t)q;
    return q;
}

// Fluence, then the flux bins restart from zero (Tallies::CycleFinalize,
// Tallies.cc:25-98).  The cycle's flux sum was accumulated by the passes.
void enqueueFinalize(qsg_problem_s* p) {
    const qs::LaunchCtx c = p->ctx();
    qs::launchFluence(c, p->dp, &p->ctl->hold);
    qs::launchZeroFlux(c, p->dp.flux, (size_t)p->tables.nCells * p->tables.nGroups, &p->ctl->hold);
}

// The balance of a cycle from its final counters (cycleFinalize,
// main.cc:310-324); the problem then stands at the next cycle boundary with
// the census as its processing population.
void closeCycle(qsg_problem_s* p, const qs::DevCounters& d, uint32_t nCensus, qsg_balance* bal, double* fluxSum) {
    const qs::ProblemTables& T = p->tables;
    if (fluxSum) *fluxSum = d.fluxSum;
    p->stats.segments += d.numSegments;
    p->stats.collisions += d.collision;
    p->stats.facet_crossings += d.numSegments - d.collision - d.census;
    // particle loads: every population-control survivor plus every
    // secondary record (start + source + split - rr + produce when each
    // fission emits >= 2 particles)
    const uint64_t start = p->hostCycle ? p->hostCount : p->startCount;
    const uint64_t source = p->hostCycle ? 0 : T.sourcePerCycle;
    p->stats.particle_loads += start + source + d.split - d.rr + d.produce;
    p->stats.cycles += 1;
    qsg_balance b;
    b.start = start;
    b.source = source;
    b.rr = d.rr;
    b.split = d.split;
    b.absorb = d.absorb;
    b.scatter = d.scatter;
    b.fission = d.fission;
    b.produce = d.produce;
    b.collision = d.collision;
    b.escape = d.escape;
    b.census = d.census;
    b.num_segments = d.numSegments;
    b.end = nCensus;
    qsgi::addBalance(p->cumulative, b);
    if (bal) *bal = b;
    p->nCensus = nCensus;
    p->censusBuf = 1 - p->cur;
    p->startCount = nCensus;
    p->cur = p->censusBuf;
    p->cycle++;
    p->phase = 0;
    p->hostCycle = false;
}

// ---- the three per-cycle entry points: each synchronises once it is done

void cycleInit(qsg_problem_s* p) {
    qsgi::requireUsable(p);
    if (p->phase != 0) throw QsgError(QSG_ERR_STATE, "cycle_init called twice without finalize");
    setDevice(p);
    if (p->startCount + p->tables.sourcePerCycle > p->proc[p->cur].cap)
        throw QsgError(QSG_ERR_CAPACITY, "processing population exceeds capacity");
    deviceStage(p, [&] {
        hipCheck(hipEventRecord(p->evI0, p->stream), "hipEventRecord");
        p->queueSorted = enqueueSetUp(p, p->cycle, p->cur, true, false);
        hipCheck(hipEventRecord(p->evI1, p->stream), "hipEventRecord");
        syncCounters(p);
        checkErrors(p);
    });
    p->nPop = p->hpc->nPopOut;
    p->stats.init_seconds += elapsed(p->evI0, p->evI1);
    p->phase = 1;
}

// A cycle whose processing population is the caller's: the caller ran the
// reference's own cycleInit (source, population control, low-weight
// roulette, main.cc:96-121) on the host, and the given particles are the
// queue of the first tracking pass.  The counters are cleared and nothing
// else runs on the device.
void cycleBeginHost(qsg_problem_s* p, const qsg_particles* in, uint64_t n) {
    qsgi::requireUsable(p);
    if (p->phase != 0) throw QsgError(QSG_ERR_STATE, "cycle_begin_host needs a problem at a cycle boundary");
    if (n > p->pop.cap) throw QsgError(QSG_ERR_CAPACITY, "more particles than the particle capacity");
    if (n && !in) throw QsgError(QSG_ERR_ARG, "NULL particles");
    for (uint64_t i = 0; i < n; i++)
        if (in->cell[i] < 0 || in->cell[i] >= p->tables.nCells)
            throw QsgError(QSG_ERR_ARG, "particle cell id out of range");
    setDevice(p);
    const qs::ProblemTables& T = p->tables;
    deviceStage(p, [&] {
        hipCheck(hipEventRecord(p->evI0, p->stream), "hipEventRecord");
        // zero carried and zero sourced: the set-up clears every counter
        qs::launchCycleSetup(p->ctx(), p->cb, 0u, p->proc[p->cur].cap, 0, T.nParticles, T.loadBalance);
        copyParticles(p->pop.v, in, n, hipMemcpyHostToDevice, p->stream);
        const uint32_t n32 = (uint32_t)n;
        hipCheck(hipMemcpyAsync(&p->pc->nPopOut, &n32, sizeof(uint32_t), hipMemcpyHostToDevice, p->stream),
                 "hipMemcpyAsync");
        hipCheck(hipEventRecord(p->evI1, p->stream), "hipEventRecord");
        hipCheck(hipStreamSynchronize(p->stream), "hipStreamSynchronize");
        p->stats.init_seconds += elapsed(p->evI0, p->evI1);
    });
    p->nPop = (uint32_t)n;
    p->queueSorted = false;
    p->hostCycle = true;
    p->hostCount = n;
    p->phase = 1;
}

void cycleTracking(qsg_problem_s* p) {
    qsgi::requireUsable(p);
    if (p->phase != 1) throw QsgError(QSG_ERR_STATE, "cycle_tracking called before cycle_init");
    setDevice(p);
    const int cur = p->cur;
    deviceStage(p, [&] {
        hipCheck(hipEventRecord(p->ev0, p->stream), "hipEventRecord");
        const int passes = runPasses(p, 0, cur, p->queueSorted ? p->proc[cur].v : p->pop.v, &p->pc->nPopOut,
                                     p->nPop, p->ev1);
        if (passes == 0) {
            // nothing to track: the counters are those the set-up left
            hipCheck(hipEventRecord(p->ev1, p->stream), "hipEventRecord");
            hipCheck(hipEventSynchronize(p->ev1), "hipEventSynchronize");
        }
        p->lastTrackingSeconds = elapsed(p->ev0, p->ev1);
    });
    p->stats.tracking_seconds += p->lastTrackingSeconds;
    p->nCensus = p->hpc->nCensus;
    p->censusBuf = 1 - cur;
    p->phase = 2;
}

void cycleFinalize(qsg_problem_s* p, qsg_balance* bal, double* fluxSum) {
    qsgi::requireUsable(p);
    if (p->phase != 2) throw QsgError(QSG_ERR_STATE, "cycle_finalize called before cycle_tracking");
    setDevice(p);
    deviceStage(p, [&] {
        hipCheck(hipEventRecord(p->evF0, p->stream), "hipEventRecord");
        enqueueFinalize(p);
        hipCheck(hipEventRecord(p->evF1, p->stream), "hipEventRecord");
        hipCheck(hipStreamSynchronize(p->stream), "hipStreamSynchronize");
        p->stats.finalize_seconds += elapsed(p->evF0, p->evF1);
    });
    // the tallies and the flux sum came back with the tracking's last
    // synchronisation, and no finalize kernel changes them
    closeCycle(p, *p->hdc, p->nCensus, bal, fluxSum);
}

// ---- qsg_run: cycles enqueued ahead of the host
//
// Cycle i + 1 is enqueued before the host waits for cycle i, so the device
// goes from one cycle to the next without waiting on the host.  After a
// cycle's first pass one small kernel (k_cycle_tail) copies the counters to
// the host and does the next cycle's set-up bookkeeping from the census
// count on the device; the next cycle's source launch also zeroes the flux
// bins.  A cycle whose first pass leaves secondaries (a wave stack
// overflowed) or an error sets CycleCtl::hold, and every kernel enqueued
// after it returns at once; the host then runs that cycle's further passes
// itself and enqueues the next cycle again.  Timing: events before and after
// each first pass only; the device time between two first passes (copy,
// finalize, set-up) is counted as set-up, the last cycle's as finalize.
void runCycles(qsg_problem_s* p, int nCycles, qsg_balance* bals, double* fluxSums, double* trackSec) {
    qsgi::requireUsable(p);
    if (p->phase != 0) throw QsgError(QSG_ERR_STATE, "qsg_run needs a problem at a cycle boundary");
    if (nCycles <= 0) return;
    setDevice(p);
    const qs::ProblemTables& T = p->tables;
    if (p->startCount + T.sourcePerCycle > p->proc[p->cur].cap)
        throw QsgError(QSG_ERR_CAPACITY, "processing population exceeds capacity");
    const qs::LaunchCtx c = p->ctx();
    const int cycle0 = p->cycle, cur0 = p->cur;
    // cycle i: (the previous cycle's fluence,) set-up, first pass, tail;
    // setUp: the bookkeeping too (no tail has done it)
    auto enqueueCycle = [&](int i, bool setUp) {
        auto& e = p->runEv[i & 3];
        const int cur = cur0 ^ (i & 1);
        if (i > 0) qs::launchFluence(c, p->dp, &p->ctl->hold);
        const bool sorted = enqueueSetUp(p, cycle0 + i, cur, setUp, i > 0);
        hipCheck(hipEventRecord(e.t0, p->stream), "hipEventRecord");
        enqueuePass(p, 0, cur, sorted ? p->proc[cur].v : p->pop.v, &p->pc->nPopOut, p->pop.cap);
        hipCheck(hipEventRecord(e.p1, p->stream), "hipEventRecord");
        qs::launchCycleTail(c, p->cb, p->hcb[i & 1], i + 1 < nCycles ? 1 : 0, p->proc[cur ^ 1].cap,
                            T.sourcePerCycle, T.nParticles, T.loadBalance);
    };
    // everything after cycle i's tail up to the next cycle's first pass (or
    // the last finalize), then the wait until cycle i's counters are on the
    // host: the next cycle's first event, or the end of the stream
    auto finishCycle = [&](int i) {
        if (i + 1 < nCycles) {
            enqueueCycle(i + 1, false);
            hipCheck(hipEventSynchronize(p->runEv[(i + 1) & 3].t0), "hipEventSynchronize");
        } else {
            enqueueFinalize(p);
            hipCheck(hipEventRecord(p->runEnd, p->stream), "hipEventRecord");
            hipCheck(hipEventSynchronize(p->runEnd), "hipEventSynchronize");
        }
    };
    try {
        hipCheck(hipEventRecord(p->runStart, p->stream), "hipEventRecord");
        enqueueCycle(0, true);
        hipEvent_t prevEnd = p->runStart;
        for (int i = 0; i < nCycles; i++) {
            auto& e = p->runEv[i & 3];
            finishCycle(i);
            const qs::CycleBlock* h = p->hcb[i & 1];
            const bool held = h->ctl.hold != 0;
            hipEvent_t end = e.p1;
            if (held) {
                // everything enqueued after this cycle's tail has returned at once
                hipCheck(hipStreamSynchronize(p->stream), "hipStreamSynchronize");
                checkErrors(h->pc.error);
                // the first pass left secondaries: the further passes, one by one
                const uint32_t nExtra = h->pc.extraCount[0];
                qs::launchClear(c, &p->ctl->hold, sizeof(uint32_t));
                runPasses(p, 1, p->cur, p->extra[0].v, &p->pc->extraCount[0], nExtra, e.t1);
                h = p->hcb[0];  // the counters after the last pass
                end = e.t1;
            }
            p->stats.init_seconds += elapsed(prevEnd, e.t0);
            p->stats.track_kernel_seconds += elapsed(e.t0, e.p1);
            p->stats.track_launches += 1;
            const double ts = elapsed(e.t0, end);
            p->lastTrackingSeconds = ts;
            p->stats.tracking_seconds += ts;
            if (trackSec) trackSec[i] = ts;
            p->nPop = h->pc.nPopOut;
            closeCycle(p, h->dc, h->pc.nCensus, bals ? &bals[i] : nullptr, fluxSums ? &fluxSums[i] : nullptr);
            if (held) {
                // the held finalize and next cycle, again
                if (i + 1 < nCycles) enqueueCycle(i + 1, true);
                else finishCycle(i);
            }
            prevEnd = end;
        }
        p->stats.finalize_seconds += elapsed(prevEnd, p->runEnd);
    } catch (const QsgError& e) {
        // nothing still in flight may write into the caller's arrays later
        (void)hipStreamSynchronize(p->stream);
        qsgi::markFailed(p, e.code, e.what());
        throw;
    } catch (const std::exception& e) {
        (void)hipStreamSynchronize(p->stream);
        qsgi::markFailed(p, QSG_ERR_DEVICE, e.what());
        throw;
    }
}

// checkCrossSections (initMC.cc:392-484, called by initMC at initMC.cc:72):
// with crossSectionsOut set, <name>.dat lists per group the group's mid
// energy and, per material in name order, the absorption, fission and
// scatter cross sections summed over the material's isotopes and reactions
// of each type, each term divided by the isotope count, in the reference's
// loop order (isotope, then reaction).
void writeCrossSections(const qs::ProblemTables& T, const std::string& name) {
    const int G = T.nGroups;
    std::vector<const qs::MaterialTables*> mats;
    for (const qs::MaterialTables& m : T.materials) mats.push_back(&m);
    std::sort(mats.begin(), mats.end(), [](const qs::MaterialTables* a, const qs::MaterialTables* b) {
        return a->name < b->name;
    });
    FILE* f = fopen((name + ".dat").c_str(), "w");
    if (!f) throw QsgError(QSG_ERR_ARG, "cannot write " + name + ".dat");
    fprintf(f, "#group  energy");
    for (const qs::MaterialTables* m : mats)
        fprintf(f, "  %s_a  %s_f  %s_s", m->name.c_str(), m->name.c_str(), m->name.c_str());
    fprintf(f, "\n");
    std::vector<double> a(mats.size() * (size_t)G, 0.0), fi(a.size(), 0.0), sc(a.size(), 0.0);
    for (size_t k = 0; k < mats.size(); k++) {
        const qs::MaterialTables& m = *mats[k];
        for (int iso = 0; iso < m.nIsotopes; iso++)
            for (int r = 0; r < m.nReactions; r++)
                for (int g = 0; g < G; g++) {
                    const double v = m.micro[(size_t)r * G + g] / m.nIsotopes;
                    double* dst = m.reactionType[r] == qs::kScatter ? &sc[k * G + g]
                                  : m.reactionType[r] == qs::kAbsorption ? &a[k * G + g]
                                                                         : &fi[k * G + g];
                    *dst += v;
                }
    }
    for (int g = 0; g < G; g++) {
        fprintf(f, "%u  %g", (unsigned)g, (T.energies[g] + T.energies[g + 1]) / 2.0);
        for (size_t k = 0; k < mats.size(); k++) fprintf(f, "  %g  %g  %g", a[k * G + g], fi[k * G + g], sc[k * G + g]);
        fprintf(f, "\n");
    }
    fclose(f);
}

qsg_problem_s* createFromParams(const qs::Parameters& params, int device, uint64_t capacity = 0) {
    qs::ProblemTables T = qs::buildProblem(params);
    if (!params.simulationParams.crossSectionsOut.empty()) writeCrossSections(T, params.simulationParams.crossSectionsOut);
    return qsgi::createFromTables(params, std::move(T), device, true, capacity);
}

}  // namespace

qs::WaveStacks qsgi::StackPool::ensure(uint32_t nWaves, uint32_t cap, hipStream_t stream) {
    if (cap == 0 || nWaves == 0) return qs::WaveStacks{nullptr, 0, 0};
    const size_t need = qs::waveStackBytes(nWaves, cap);
    if (need > bytes) {
        hipCheck(hipStreamSynchronize(stream), "hipStreamSynchronize");
        release();
        void* m = nullptr;
        hipCheck(hipMalloc(&m, need), "hipMalloc (wave stacks)");
        mem = static_cast<char*>(m);
        bytes = need;
    }
    return qs::WaveStacks{mem, cap, nWaves};
}

namespace {
std::mutex g_poolMutex;
std::map<int, std::vector<hipStream_t>> g_streams;  // idle streams per device
std::vector<char*> g_pinned;                        // idle pinned blocks
}  // namespace

hipStream_t qsgi::acquireStream(int device) {
    {
        std::lock_guard<std::mutex> lock(g_poolMutex);
        std::vector<hipStream_t>& v = g_streams[device];
        if (!v.empty()) {
            hipStream_t s = v.back();
            v.pop_back();
            return s;
        }
    }
    hipStream_t s = nullptr;
    hipCheck(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "hipStreamCreate");
    return s;
}

void qsgi::releaseStream(int device, hipStream_t s) {
    if (!s) return;
    std::lock_guard<std::mutex> lock(g_poolMutex);
    g_streams[device].push_back(s);
}

void* qsgi::acquirePinned(size_t bytes) {
    if (bytes > kPinnedBlock) throw QsgError(QSG_ERR_ARG, "pinned block request too large");
    std::lock_guard<std::mutex> lock(g_poolMutex);
    if (g_pinned.empty()) {
        // one pinned chunk, carved into blocks
        constexpr size_t kChunk = 64 * kPinnedBlock;
        void* m = nullptr;
        hipCheck(hipHostMalloc(&m, kChunk, hipHostMallocDefault), "hipHostMalloc");
        for (size_t o = 0; o < kChunk; o += kPinnedBlock) g_pinned.push_back(static_cast<char*>(m) + o);
    }
    char* b = g_pinned.back();
    g_pinned.pop_back();
    std::memset(b, 0, kPinnedBlock);
    return b;
}

void qsgi::releasePinned(void* p) {
    if (!p) return;
    std::lock_guard<std::mutex> lock(g_poolMutex);
    g_pinned.push_back(static_cast<char*>(p));
}

namespace {
struct DeviceCache {
    std::multimap<size_t, void*> idle;  // block size -> block
    size_t idleBytes = 0;
};
std::map<int, DeviceCache> g_devCache;  // guarded by g_poolMutex

void releaseIdle(DeviceCache& c) {
    for (const auto& b : c.idle) (void)hipFree(b.second);
    c.idle.clear();
    c.idleBytes = 0;
}
}  // namespace

void* qsgi::deviceAlloc(int device, size_t bytes, size_t* blockBytes) {
    {
        std::lock_guard<std::mutex> lock(g_poolMutex);
        DeviceCache& c = g_devCache[device];
        auto it = c.idle.lower_bound(bytes);
        if (it != c.idle.end() && it->first / 3 * 2 <= bytes) {
            void* p = it->second;
            *blockBytes = it->first;
            c.idleBytes -= it->first;
            c.idle.erase(it);
            return p;
        }
    }
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, bytes);
    if (e == hipErrorOutOfMemory) {
        (void)hipGetLastError();
        {
            std::lock_guard<std::mutex> lock(g_poolMutex);
            releaseIdle(g_devCache[device]);
        }
        e = hipMalloc(&p, bytes);
    }
    hipCheck(e, "hipMalloc");
    *blockBytes = bytes;
    return p;
}

void qsgi::deviceRelease(int device, void* p, size_t blockBytes) {
    if (!p) return;
    {
        std::lock_guard<std::mutex> lock(g_poolMutex);
        DeviceCache& c = g_devCache[device];
        if (c.idleBytes + blockBytes <= kDeviceCacheBytes) {
            c.idle.emplace(blockBytes, p);
            c.idleBytes += blockBytes;
            return;
        }
    }
    (void)hipFree(p);
}

void qsgi::StackPool::release() {
    if (mem) (void)hipFree(mem);
    mem = nullptr;
    bytes = 0;
}

qs::TrackTuning qsgi::tuningFromC(const qsg_tuning* t) {
    if (t->block_size < 64 || t->block_size > 256 || t->block_size % 64)
        throw QsgError(QSG_ERR_ARG, "block_size must be 64, 128, 192 or 256");
    if (t->waves_per_cu < 1 || t->waves_per_cu > 64) throw QsgError(QSG_ERR_ARG, "waves_per_cu out of range");
    if (t->refill_min < 1 || t->refill_min > 64) throw QsgError(QSG_ERR_ARG, "refill_min out of range");
    if (t->chunk < 1 || t->chunk > 4096) throw QsgError(QSG_ERR_ARG, "chunk out of range");
    if (t->stack_entries != 0 && (t->stack_entries < (int)qs::kStackMargin * 2 || t->stack_entries > 65536))
        throw QsgError(QSG_ERR_ARG, "stack_entries must be 0 or 512..65536");
    if (t->sort_interval < 0) throw QsgError(QSG_ERR_ARG, "sort_interval must be >= 0");
    if (t->waves_per_simd < 2 || t->waves_per_simd > 4) throw QsgError(QSG_ERR_ARG, "waves_per_simd must be 2, 3 or 4");
    if (t->graphs != 0 && t->graphs != 1) throw QsgError(QSG_ERR_ARG, "graphs must be 0 or 1");
    return qs::TrackTuning{t->block_size, t->waves_per_cu, t->refill_min, t->chunk, t->stack_entries,
                           t->sort_interval, t->waves_per_simd, t->graphs};
}

void qsgi::tuningToC(const qs::TrackTuning& in, qsg_tuning* t) {
    t->block_size = in.blockSize;
    t->waves_per_cu = in.wavesPerCU;
    t->refill_min = in.refillMin;
    t->chunk = in.chunk;
    t->stack_entries = in.stackCap;
    t->sort_interval = in.sortInterval;
    t->waves_per_simd = in.wavesPerSimd;
    t->graphs = in.graphs;
}

uint32_t qsgi::particleCapacity(const qs::ProblemTables& T, uint64_t requested) {
    const double capD = requested ? (double)requested
                                  : 2.0 * (double)T.nParticles + (double)T.sourcePerCycle + 65536.0;
    if (capD > 4.0e9) throw QsgError(QSG_ERR_CAPACITY, "particle capacity exceeds 32-bit indexing");
    return (uint32_t)capD;
}

std::string qsgi::deviceErrorText(uint32_t e) {
    std::string m = "device error flags:";
    if (e & 1u) m += " census buffer overflow;";
    if (e & 2u) m += " secondary buffer overflow;";
    if (e & 4u) m += " population-control buffer overflow;";
    if (e & 8u) m += " runaway particle history;";
    if (e & 16u) m += " fission produced more than 4 particles;";
    if (e & 32u) m += " census + source exceed the processing buffer;";
    if (e & (1u | 2u | 4u | 32u)) m += " (a larger qsg_options.particle_capacity avoids capacity errors)";
    return m;
}

int qsgi::deviceErrorCode(uint32_t e) { return (e & (1u | 2u | 4u | 32u)) ? QSG_ERR_CAPACITY : QSG_ERR_NUMERIC; }

void qsgi::markFailed(qsg_problem_s* p, int code, const std::string& msg) {
    if (!p || p->failCode) return;
    p->failCode = code ? code : QSG_ERR_DEVICE;
    p->failMsg = msg;
}

void qsgi::requireUsable(const qsg_problem_s* p) {
    if (p->failCode) throw QsgError(QSG_ERR_STATE, "the problem failed earlier (" + p->failMsg + ")");
}

uint64_t qsgi::deviceBytesEstimate(const qs::ProblemTables& T, uint64_t requested) {
    const uint64_t nc = (uint64_t)T.nCells, G = (uint64_t)T.nGroups, M = T.materials.size();
    const uint64_t R = (uint64_t)T.maxReactions;
    uint64_t b = nc * 96 * 8 + nc * 4 + nc * 8 + nc * 24 + (nc + 1) * 8 + nc * G * 8;
    b += (G + 1) * 8 + 2 * M * G * 8 + M * R * G * 8 + M * R * 4;
    const uint64_t cap = particleCapacity(T, requested);
    b += 5 * (cap * qs::kParticleBytes + 12 * 256) + qs::popGroups((uint32_t)cap) * sizeof(qs::PopGroup);
    b += qs::sortKeys((int)nc) * 4 + qs::scanBlocks(qs::sortKeys((int)nc)) * 4;
    return b + 64 * 1024;
}

qsg_problem_s* qsgi::createFromTables(const qs::Parameters& params, qs::ProblemTables tables, int device,
                                      bool ownStacks, uint64_t capacity) {
    qsg_problem_s* p = new qsg_problem_s();
    try {
        p->params = params;
        p->tables = std::move(tables);
        build(p, device, capacity);
        // Wave stacks for this problem's own tracking launches, sized for the
        // largest grid its population can fill (allocated now so that running
        // the problem allocates nothing unless its launch shape grows).
        if (ownStacks && p->tune.stackCap > 0) {
            const uint64_t maxWaves = (uint64_t)p->numCUs * (uint64_t)p->tune.wavesPerCU;
            const uint64_t fill = ((uint64_t)p->pop.cap + 255) / 256 * 4;  // any block size
            p->stacks.ensure((uint32_t)(fill < maxWaves ? fill : maxWaves), (uint32_t)p->tune.stackCap, p->stream);
        }
    } catch (...) {
        destroyProblem(p);
        throw;
    }
    return p;
}

void qsgi::destroy(qsg_problem_s* p) { destroyProblem(p); }

extern "C" {

const char* qsg_last_error(void) { return g_lastError.c_str(); }

int qsg_create(int argc, const char* const* argv, int device, qsg_problem* out) {
    QSG_GUARD({
        if (!out) throw QsgError(QSG_ERR_ARG, "out is NULL");
        *out = nullptr;
        qs::Parameters params = qs::getParameters(argc, argv);
        *out = createFromParams(params, device);
    })
}

int qsg_create_from_deck(const char* deck_text, int argc, const char* const* argv, int device,
                         qsg_problem* out) {
    return qsg_create_from_deck_opts(deck_text, argc, argv, device, nullptr, out);
}

int qsg_create_from_deck_opts(const char* deck_text, int argc, const char* const* argv, int device,
                              const qsg_options* opts, qsg_problem* out) {
    QSG_GUARD({
        if (!out || !deck_text) throw QsgError(QSG_ERR_ARG, "NULL argument");
        *out = nullptr;
        const char* dflt[] = {"qs"};
        qs::Parameters params = (argc > 0 && argv) ? qs::getParameters(argc, argv) : qs::getParameters(1, dflt);
        params.geometryParams.clear();
        // getParameters supplied the default problem; a deck replaces it
        params.materialParams.clear();
        params.crossSectionParams.clear();
        qs::parseDeckText(deck_text, params);
        qs::supplyDefaults(params);
        *out = createFromParams(params, device, opts ? opts->particle_capacity : 0);
    })
}

int qsg_destroy(qsg_problem p) {
    QSG_GUARD({ destroyProblem(p); })
}

int qsg_problem_error(qsg_problem p, int* code, char* message, size_t message_len) {
    QSG_GUARD({
        if (!p || !code) throw QsgError(QSG_ERR_ARG, "NULL argument");
        *code = p->failCode;
        if (message && message_len) {
            const size_t n = std::min(p->failMsg.size(), message_len - 1);
            memcpy(message, p->failMsg.data(), n);
            message[n] = 0;
        }
    })
}

int qsg_get_info(qsg_problem p, qsg_info* info) {
    QSG_GUARD({
        if (!p || !info) throw QsgError(QSG_ERR_ARG, "NULL argument");
        const qs::ProblemTables& T = p->tables;
        info->nx = T.nx;
        info->ny = T.ny;
        info->nz = T.nz;
        info->n_groups = T.nGroups;
        info->n_materials = (int32_t)T.materials.size();
        info->n_steps = T.nSteps;
        info->n_cells = (uint64_t)T.nCells;
        info->n_particles_target = T.nParticles;
        info->source_per_cycle = T.sourcePerCycle;
        info->particle_capacity = p->pop.cap;
        info->device_bytes = p->deviceBytes;
        info->dt = T.dt;
        info->source_weight = T.sourceWeight;
        info->cycle = p->cycle;
        info->lds_tables = qs::tablesFitLds(p->dp) ? 1 : 0;
    })
}

int qsg_set_tuning(qsg_problem p, const qsg_tuning* t) {
    QSG_GUARD({
        if (!p || !t) throw QsgError(QSG_ERR_ARG, "NULL argument");
        p->tune = qsgi::tuningFromC(t);
    })
}

int qsg_get_tuning(qsg_problem p, qsg_tuning* t) {
    QSG_GUARD({
        if (!p || !t) throw QsgError(QSG_ERR_ARG, "NULL argument");
        qsgi::tuningToC(p->tune, t);
    })
}

int qsg_cycle_init(qsg_problem p) {
    QSG_GUARD({
        if (!p) throw QsgError(QSG_ERR_ARG, "NULL problem");
        cycleInit(p);
    })
}

int qsg_cycle_begin_host(qsg_problem p, const qsg_particles* in, uint64_t n) {
    QSG_GUARD({
        if (!p) throw QsgError(QSG_ERR_ARG, "NULL problem");
        cycleBeginHost(p, in, n);
    })
}

int qsg_cycle_tracking(qsg_problem p) {
    QSG_GUARD({
        if (!p) throw QsgError(QSG_ERR_ARG, "NULL problem");
        cycleTracking(p);
    })
}

int qsg_cycle_finalize(qsg_problem p, qsg_balance* balance, double* flux_sum) {
    QSG_GUARD({
        if (!p) throw QsgError(QSG_ERR_ARG, "NULL problem");
        cycleFinalize(p, balance, flux_sum);
    })
}

int qsg_run(qsg_problem p, int n_cycles, qsg_balance* balances, double* flux_sums, double* tracking_seconds) {
    QSG_GUARD({
        if (!p) throw QsgError(QSG_ERR_ARG, "NULL problem");
        runCycles(p, n_cycles, balances, flux_sums, tracking_seconds);
    })
}

int qsg_get_stats(qsg_problem p, qsg_stats* stats) {
    QSG_GUARD({
        if (!p || !stats) throw QsgError(QSG_ERR_ARG, "NULL argument");
        *stats = p->stats;
    })
}

int qsg_reset_stats(qsg_problem p) {
    QSG_GUARD({
        if (!p) throw QsgError(QSG_ERR_ARG, "NULL problem");
        p->stats = qsg_stats{};
    })
}

int qsg_get_scalar_flux(qsg_problem p, double* out, uint64_t n) {
    QSG_GUARD({
        if (!p || !out) throw QsgError(QSG_ERR_ARG, "NULL argument");
        uint64_t total = (uint64_t)p->tables.nCells * p->tables.nGroups;
        if (n != total) throw QsgError(QSG_ERR_ARG, "flux buffer must hold n_cells * n_groups values");
        setDevice(p);
        hipCheck(hipMemcpyAsync(out, p->dp.flux, n * sizeof(double), hipMemcpyDeviceToHost, p->stream),
                 "hipMemcpyAsync");
        hipCheck(hipStreamSynchronize(p->stream), "hipStreamSynchronize");
    })
}

int qsg_set_particles(qsg_problem p, const qsg_particles* in, uint64_t n) {
    QSG_GUARD({
        if (!p || !in) throw QsgError(QSG_ERR_ARG, "NULL argument");
        qsgi::requireUsable(p);
        if (p->phase != 1) throw QsgError(QSG_ERR_STATE, "set_particles must follow cycle_init");
        if (n > p->pop.cap) throw QsgError(QSG_ERR_CAPACITY, "too many particles");
        for (uint64_t i = 0; i < n; i++)
            if (in->cell[i] < 0 || in->cell[i] >= p->tables.nCells)
                throw QsgError(QSG_ERR_ARG, "particle cell id out of range");
        setDevice(p);
        copyParticles(p->pop.v, in, n, hipMemcpyHostToDevice, p->stream);
        p->queueSorted = false;
        uint32_t n32 = (uint32_t)n;
        hipCheck(hipMemcpy(&p->pc->nPopOut, &n32, sizeof(uint32_t), hipMemcpyHostToDevice), "hipMemcpy");
        p->nPop = n32;
    })
}

int qsg_get_census_count(qsg_problem p, uint64_t* n) {
    QSG_GUARD({
        if (!p || !n) throw QsgError(QSG_ERR_ARG, "NULL argument");
        *n = (p->censusBuf < 0) ? 0 : p->nCensus;
    })
}

int qsg_get_census(qsg_problem p, qsg_particles* out, uint64_t n) {
    QSG_GUARD({
        if (!p || !out) throw QsgError(QSG_ERR_ARG, "NULL argument");
        qsgi::requireUsable(p);
        if (p->censusBuf < 0) throw QsgError(QSG_ERR_STATE, "no census yet");
        if (n > p->nCensus) throw QsgError(QSG_ERR_ARG, "n exceeds the census count");
        setDevice(p);
        copyParticles(p->proc[p->censusBuf].v, out, n, hipMemcpyDeviceToHost, p->stream);
    })
}

int qsg_get_processing_count(qsg_problem p, uint64_t* n) {
    QSG_GUARD({
        if (!p || !n) throw QsgError(QSG_ERR_ARG, "NULL argument");
        *n = (p->phase >= 1) ? p->nPop : 0;
    })
}

int qsg_get_processing(qsg_problem p, qsg_particles* out, uint64_t n) {
    QSG_GUARD({
        if (!p || !out) throw QsgError(QSG_ERR_ARG, "NULL argument");
        qsgi::requireUsable(p);
        if (p->phase != 1) throw QsgError(QSG_ERR_STATE, "processing population exists after cycle_init");
        if (n > p->nPop) throw QsgError(QSG_ERR_ARG, "n exceeds the processing count");
        setDevice(p);
        copyParticles(p->queueSorted ? p->proc[p->cur].v : p->pop.v, out, n, hipMemcpyDeviceToHost, p->stream);
    })
}

int qsg_get_fluence(qsg_problem p, double* fluence, uint64_t n_cells) {
    QSG_GUARD({
        if (!p || !fluence) throw QsgError(QSG_ERR_ARG, "NULL argument");
        if (!p->dp.fluence) throw QsgError(QSG_ERR_STATE, "fluence is accumulated only for coralBenchmark decks");
        if (n_cells != (uint64_t)p->tables.nCells) throw QsgError(QSG_ERR_ARG, "n_cells does not match the mesh");
        setDevice(p);
        hipCheck(hipStreamSynchronize(p->stream), "hipStreamSynchronize");
        hipCheck(hipMemcpy(fluence, p->dp.fluence, n_cells * sizeof(double), hipMemcpyDeviceToHost), "hipMemcpy");
    })
}

int qsg_get_energy_spectrum(qsg_problem p, double* energies, uint64_t* counts, uint64_t n) {
    QSG_GUARD({
        if (!p || !counts) throw QsgError(QSG_ERR_ARG, "NULL argument");
        if (!p->dp.spectrum)
            throw QsgError(QSG_ERR_STATE, "the census spectrum is kept only for runs that set energySpectrum");
        if (n != (uint64_t)p->tables.nGroups + 1) throw QsgError(QSG_ERR_ARG, "n must be n_groups + 1");
        setDevice(p);
        hipCheck(hipStreamSynchronize(p->stream), "hipStreamSynchronize");
        hipCheck(hipMemcpy(counts, p->dp.spectrum, n * sizeof(uint64_t), hipMemcpyDeviceToHost), "hipMemcpy");
        if (energies)
            for (uint64_t g = 0; g < n; g++) energies[g] = p->tables.energies[g];
    })
}

int qsg_coral_check(qsg_problem p, qsg_coral_report* r) {
    QSG_GUARD({
        if (!p || !r) throw QsgError(QSG_ERR_ARG, "NULL argument");
        *r = qsg_coral_report{};
        const int bench = p->params.simulationParams.coralBenchmark;
        r->benchmark = bench;
        if (bench != 1 && bench != 2) throw QsgError(QSG_ERR_STATE, "the deck does not set coralBenchmark 1 or 2");
        const qsg_balance& b = p->cumulative;
        // BalanceRatioTest (CoralBenchmark.cc:46-117): the expected reaction
        // ratios of each CORAL2 problem, every pairwise relative error
        double fissionRatio, scatterRatio, absorbRatio, tol;
        if (bench == 1) {
            fissionRatio = 0.05;
            scatterRatio = 1;
            absorbRatio = 0.04;
            tol = 1.0 / 100.0;
        } else {
            fissionRatio = 0.075;
            scatterRatio = 0.830;
            absorbRatio = 0.094;
            tol = 1.1 / 100.0;
        }
        const uint64_t absorb = b.absorb, fission = b.fission, scatter = b.scatter;
        const double errs[6] = {std::abs((absorb / absorbRatio) * (scatterRatio / scatter) - 1),
                                std::abs((absorb / absorbRatio) * (fissionRatio / fission) - 1),
                                std::abs((scatter / scatterRatio) * (absorbRatio / absorb) - 1),
                                std::abs((scatter / scatterRatio) * (fissionRatio / fission) - 1),
                                std::abs((fission / fissionRatio) * (absorbRatio / absorb) - 1),
                                std::abs((fission / fissionRatio) * (scatterRatio / scatter) - 1)};
        r->ratio_pass = 1;
        r->ratio_error = 0;
        for (double e : errs) {
            if (e > tol) r->ratio_pass = 0;
            if (!(e <= r->ratio_error)) r->ratio_error = e;
        }
        r->ratio_tolerance = tol;
        // BalanceEventTest (CoralBenchmark.cc:119-150)
        const uint64_t facet = b.num_segments - b.census - b.collision;
        r->event_ratio = std::abs((double(facet) / double(b.collision)) - 1);
        r->event_pass = r->event_ratio > (1.0 / 100.0) ? 0 : 1;
        // MissingParticleTest (CoralBenchmark.cc:152-174)
        r->gains = b.start + b.source + b.produce + b.split;
        r->losses = b.absorb + b.census + b.escape + b.rr + b.fission;
        r->missing_pass = r->gains == r->losses ? 1 : 0;
        // FluenceTest (CoralBenchmark.cc:177-228), one domain
        const int nc = p->tables.nCells;
        std::vector<double> fl(nc);
        setDevice(p);
        hipCheck(hipStreamSynchronize(p->stream), "hipStreamSynchronize");
        hipCheck(hipMemcpy(fl.data(), p->dp.fluence, nc * sizeof(double), hipMemcpyDeviceToHost), "hipMemcpy");
        double sum = 0.0;
        for (int c = 0; c < nc; c++) sum += fl[c];
        const double average = sum / nc;
        double maxDiff = 0.0;
        for (int c = 0; c < nc; c++) {
            const double v = fl[c];
            const double pd = (((v > average) ? v - average : average - v) / ((v + average) / 2.0)) * 100;
            maxDiff = (maxDiff > pd) ? maxDiff : pd;
        }
        r->fluence_max_percent_diff = maxDiff;
        r->fluence_pass = maxDiff > 6.0 ? 0 : 1;
    })
}

int qsg_nearest_facet_probe(qsg_problem p, uint64_t n, const double* xyz, const double* dir, const int32_t* cell,
                            double* distance, int32_t* facet, double* xyz_out) {
    QSG_GUARD({
        if (!p || !xyz || !dir || !cell || !distance || !facet || !xyz_out) throw QsgError(QSG_ERR_ARG, "NULL argument");
        if (n > (1ull << 26)) throw QsgError(QSG_ERR_ARG, "at most 2^26 rays per call");
        for (uint64_t q = 0; q < n; q++)
            if (cell[q] < 0 || cell[q] >= p->tables.nCells) throw QsgError(QSG_ERR_ARG, "cell id outside the mesh");
        if (n == 0) return QSG_OK;
        setDevice(p);
        const size_t v3 = (size_t)n * 3 * sizeof(double);
        char* m = nullptr;
        // one scratch allocation: xyz, dir, xyzOut, distance, cell, facet
        const size_t bytes = 3 * v3 + (size_t)n * (sizeof(double) + 2 * sizeof(int32_t));
        hipCheck(hipMalloc(&m, bytes), "hipMalloc");
        double* dXyz = (double*)m;
        double* dDir = (double*)(m + v3);
        double* dOut = (double*)(m + 2 * v3);
        double* dDist = (double*)(m + 3 * v3);
        int32_t* dCell = (int32_t*)(m + 3 * v3 + (size_t)n * sizeof(double));
        int32_t* dFacet = dCell + n;
        try {
            hipCheck(hipMemcpyAsync(dXyz, xyz, v3, hipMemcpyHostToDevice, p->stream), "hipMemcpyAsync");
            hipCheck(hipMemcpyAsync(dDir, dir, v3, hipMemcpyHostToDevice, p->stream), "hipMemcpyAsync");
            hipCheck(hipMemcpyAsync(dCell, cell, (size_t)n * sizeof(int32_t), hipMemcpyHostToDevice, p->stream),
                     "hipMemcpyAsync");
            qs::launchNearestProbe(p->ctx(), p->dp, dXyz, dDir, dCell, dDist, dFacet, dOut, n);
            hipCheck(hipGetLastError(), "k_nearest_probe launch");
            hipCheck(hipMemcpyAsync(distance, dDist, (size_t)n * sizeof(double), hipMemcpyDeviceToHost, p->stream),
                     "hipMemcpyAsync");
            hipCheck(hipMemcpyAsync(facet, dFacet, (size_t)n * sizeof(int32_t), hipMemcpyDeviceToHost, p->stream),
                     "hipMemcpyAsync");
            hipCheck(hipMemcpyAsync(xyz_out, dOut, v3, hipMemcpyDeviceToHost, p->stream), "hipMemcpyAsync");
            hipCheck(hipStreamSynchronize(p->stream), "hipStreamSynchronize");
        } catch (...) {
            (void)hipStreamSynchronize(p->stream);
            (void)hipFree(m);
            throw;
        }
        hipCheck(hipFree(m), "hipFree");
    })
}

int qsg_math_probe(int op, const double* a, const double* b, double* out, uint64_t n, int device) {
    QSG_GUARD({
        if (!a || !b || !out) throw QsgError(QSG_ERR_ARG, "NULL argument");
        hipCheck(hipSetDevice(device), "hipSetDevice");
        hipStream_t s;
        hipCheck(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "hipStreamCreate");
        double *da = nullptr, *db = nullptr, *dout = nullptr;
        size_t bytes = (size_t)(n ? n : 1) * sizeof(double);
        hipCheck(hipMalloc(&da, bytes), "hipMalloc");
        hipCheck(hipMalloc(&db, bytes), "hipMalloc");
        hipCheck(hipMalloc(&dout, bytes), "hipMalloc");
        hipCheck(hipMemcpy(da, a, n * sizeof(double), hipMemcpyHostToDevice), "hipMemcpy");
        hipCheck(hipMemcpy(db, b, n * sizeof(double), hipMemcpyHostToDevice), "hipMemcpy");
        qs::LaunchCtx c{(void*)s, 0};
        qs::launchMathProbe(c, op, da, db, dout, n);
        hipCheck(hipStreamSynchronize(s), "hipStreamSynchronize");
        hipCheck(hipMemcpy(out, dout, n * sizeof(double), hipMemcpyDeviceToHost), "hipMemcpy");
        (void)hipFree(da);
        (void)hipFree(db);
        (void)hipFree(dout);
        (void)hipStreamDestroy(s);
    })
}

}  // extern "C"
