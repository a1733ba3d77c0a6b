// Device-side data layout of one transport problem and the kernel launchers.
//
// Layout in HBM (all FP64 unless noted), per problem:
//   planes      [nCells][24] x {A,B,C,D}       32 B / facet, 768 B / cell
//   cellMat     [nCells] int32
//   cellVolume  [nCells]                        (source tet sampling)
//   cellCenter  [nCells][3]                     (source + nudge toward centre)
//   srcOffset   [nCells+1] uint64               (source particles per cell)
//   flux        [nCells][G]                     scalar flux tally
//   tables      energies[G+1], total[nMat][G], macro[nMat][maxR][G],
//               per-material ints/doubles       (staged in LDS when small)
//   particles   structure of arrays, one array per field:
//               x y z vx vy vz E w tc nmfp (f64), seed (u64), cell (i32)
//               = 92 B / particle
// Nodes and facet adjacency are not stored: the mesh is the reference's
// structured hex grid, so node coordinates and neighbours are recomputed from
// the cell's (i,j,k) with the same IEEE operations the host uses.
#pragma once

#include <cstddef>
#include <cstdint>

namespace qs {

// One particle buffer: twelve arrays carved from one allocation at a fixed
// byte stride (>= capacity * 8, 256-byte aligned), in the order
// x y z vx vy vz E w tc nmfp (f64), seed (u64), cell (i32).  A view is a
// base pointer and the stride, so a kernel carries 4 registers per buffer.
struct Particles {
    char* base;
    uint64_t stride;
    __host__ __device__ double* f(int k) const { return (double*)(base + (uint64_t)k * stride); }
    __host__ __device__ double* x() const { return f(0); }
    __host__ __device__ double* y() const { return f(1); }
    __host__ __device__ double* z() const { return f(2); }
    __host__ __device__ double* vx() const { return f(3); }
    __host__ __device__ double* vy() const { return f(4); }
    __host__ __device__ double* vz() const { return f(5); }
    __host__ __device__ double* E() const { return f(6); }
    __host__ __device__ double* w() const { return f(7); }
    __host__ __device__ double* tc() const { return f(8); }     // time to census
    __host__ __device__ double* nmfp() const { return f(9); }   // remaining mean free paths
    __host__ __device__ uint64_t* seed() const { return (uint64_t*)(base + 10 * stride); }
    __host__ __device__ int32_t* cell() const { return (int32_t*)(base + 11 * stride); }
};

constexpr int kParticleFields = 12;
constexpr size_t kParticleBytes = 10 * sizeof(double) + sizeof(uint64_t) + sizeof(int32_t);

// Per-material scalar data (kept small so it can live in LDS).
struct MatInfo {
    int nIso;
    int nReact;
    double mass;
    double nuBar;
    double invMass;  // 1.0 / mass, the scatter's energy factor (NuclearData.cc:64)
};

constexpr int kMaxMaterials = 16;
constexpr int kMaxReactions = 16;
// Energy-group bounds (nGroups + 1) a batch tracking wave keeps in LDS; a
// batch with a problem of more groups searches them in device memory.
constexpr int kBatchStagedEnergies = 257;

struct DevProblem {
    // grid
    int nx, ny, nz, nCells;
    double rNx, rNxy;      // 1/nx and 1/(nx*ny), for decoding cell ids
    double dx, dy, dz;     // lx/nx ...
    double hdx, hdy, hdz;  // dx/2.0 ...
    int boundaryMask;      // 2 bits per face +x -x +y -y +z -z: 1 escape, 2 reflect
    const double* planes;    // [nCells][24][A,B,C,D]
    const int32_t* cellMat;
    const double* cellVolume;
    const double* cellCenter;
    const uint64_t* srcOffset;
    // nuclear data
    int G;
    int nMat, maxR;
    // The group bounds are log-spaced (NuclearData.cc:105-119): the group of
    // energy E is about (log2 E - eLog2Lo) * eGroupsPerLog2, a first guess
    // that the exact comparisons with `energies` then settle
    float eLog2Lo, eGroupsPerLog2;
    const double* energies;  // [G+1]
    const double* total;     // [nMat][G]
    const double* mfp;       // [nMat][G] mean free path 1/total (1e75 where total is 0)
    const double* macro;     // [nMat][maxR][G]
    const int32_t* reactType;  // [nMat][maxR]
    MatInfo mat[kMaxMaterials];
    // run constants
    double dt, eMin, eMax, sourceWeight, lowWeightCutoff;
    // tallies
    double* flux;  // [nCells][G]
    double* fluence;  // [nCells] cumulative over cycles (CORAL2 decks only, else null)
    // census energy spectrum [G+1], cumulative over cycles: census particles
    // per energy group (bin G: above eMax); only when the run asks for it
    // (energySpectrum), else null
    unsigned long long* spectrum;
};

// Balance counters accumulated on the device (Tallies.hh:36-100 Balance).
struct DevCounters {
    unsigned long long absorb, census, escape, collision, fission, produce, scatter, numSegments;
    unsigned long long source, rr, split;
    double fluxSum;  // sum of this cycle's scalar-flux contributions (segment length x weight)
    unsigned long long pad[4];
};

// Work-queue / append counters (device memory, cleared by the host).
struct PassCounters {
    uint32_t head;          // next input index handed to a wave
    uint32_t nCensus;       // census buffer fill (cumulative over the cycle)
    uint32_t nPopOut;       // population-control output fill
    uint32_t error;         // bit 0 census overflow, 1 extra overflow, 2 pop
                            // overflow, 3 runaway history, 4 nOut > 4,
                            // 5 census + source over the processing buffer
    uint32_t extraCount[2]; // secondaries appended by passes (ping-pong)
    uint32_t pad[2];
};

// Set-up values of a problem's current cycle, computed on the device from
// the census count the previous cycle left there, so that a cycle can be
// enqueued before the host has read that count.  While hold is nonzero every
// kernel of the one-problem cycle returns at once: it is set when a cycle
// needs more than its first tracking pass (a wave stack overflowed into the
// secondary buffer), when a cycle reports an error, and when census + source
// exceed the processing buffer; the host then finishes that cycle and
// enqueues the next again.
struct CycleCtl {
    uint32_t hold;
    uint32_t start;  // census particles carried into the cycle
    uint32_t n;      // start + source (0 if that exceeds the buffer)
    uint32_t mode;   // population control: 0 none, 1 roulette, 2 split
    double factor;   // population-control factor
    uint64_t pad;
};

// One group of consecutive population-control inputs: its output, roulette
// and split counts, and the base of its outputs.
struct PopGroup {
    uint32_t total, kills, splits, base;
};

// A problem's counters in one block, so that one copy brings them back.
struct CycleBlock {
    PassCounters pc;
    DevCounters dc;
    CycleCtl ctl;
};

struct TrackTuning {
    int blockSize;     // threads per workgroup (multiple of 64)
    int wavesPerCU;    // persistent waves per CU
    int refillMin;     // idle lanes that trigger a refill
    int chunk;         // particles a wave claims per queue access
    int stackCap;      // per-wave secondary stack entries (0: extra passes only)
    int sortInterval;  // sort the population by cell every n cycles (0: never)
    int wavesPerSimd;  // register budget of the tracking kernels: 2, 3 or 4 waves per SIMD
    int graphs;        // batched form: replay each step's launches from captured HIP graphs
};

// Per-wave secondary stacks of a tracking launch.  Wave w (global wave
// index) owns `cap` particle entries laid out like a particle buffer, at
// base + w * 12 * cap * 8 with field stride cap * 8.  A wave pushes the
// secondaries of its multi-particle collisions there (the parent keeps its
// lane) and pops them before it
// takes new queue entries, so secondaries are tracked in the same launch
// (depth first) instead of in a later pass.  Waves with index >= nWaves,
// or cap == 0, send secondaries to the pass's extra queue instead.
struct WaveStacks {
    char* base;
    uint32_t cap;
    uint32_t nWaves;
};
constexpr uint32_t kStackCap = 512;
// room that must be left on a stack before an iteration may push to it:
// one collision per lane, at most 4 outputs each
constexpr uint32_t kStackMargin = 256;
// census particles a wave holds before appending them to the census buffer
// (a region of the wave's own after its stack entries)
constexpr uint32_t kCensusStage = 256;
// one wave's region: its stack entries, then its census staging entries
constexpr uint32_t waveRegionEntries(uint32_t cap) { return cap == 0 ? 0u : cap + kCensusStage; }
inline size_t waveStackBytes(uint32_t nWaves, uint32_t cap) { return (size_t)nWaves * 12 * waveRegionEntries(cap) * 8; }
constexpr TrackTuning kDefaultTuning = {256, 12, 1, 32, (int)kStackCap, 0, 3, 1};

// ---- batched form: many independent problems per launch ----
// Per-problem mutable state of a batch (device memory).
struct BatchState {
    uint32_t head;      // pass queue head
    uint32_t nCensus;   // census particles appended in the current cycle
    uint32_t nPop;      // population-control output count
    uint32_t error;     // same bits as PassCounters::error
    uint32_t extra[2];  // secondaries appended by the last passes (ping-pong)
    uint32_t start;     // census count carried into the current cycle
    uint32_t nProc;     // start + source
    uint32_t cycle;     // cycles completed; processing buffer = proc[cycle & 1]
    uint32_t active;    // takes part in the next step
    uint32_t mode;      // population control: 0 none, 1 roulette, 2 split
    uint32_t sorted;    // this cycle's queue was sorted into proc[cycle & 1]
    double factor;      // population-control factor of the current cycle
};

// Per-problem read-only descriptor of a batch (device memory).
struct BatchSlot {
    DevProblem prob;
    Particles proc[2];
    Particles pop;
    Particles extra[2];
    uint32_t cap;           // capacity of each particle buffer
    uint32_t loadBalance;
    uint64_t sourcePerCycle;
    uint64_t nParticles;    // population target
    BatchState* st;
    DevCounters* dc;
    PopGroup* groups;       // the problem's population-control scratch
};

// ---- launchers (qs_kernels.hip); all asynchronous on `stream` ----
struct LaunchCtx {
    void* stream;   // hipStream_t
    int numCUs;
};

size_t tableLdsBytes(const DevProblem& P);

// Cycle set-up bookkeeping (main.cc:96-121, PopulationControl.cc:20-65):
// `start` census particles carried in, the population-control factor and
// mode into cb->ctl, the pass and balance counters cleared, hold cleared.
void launchCycleSetup(const LaunchCtx& ctx, CycleBlock* cb, uint32_t start, uint32_t procCap, uint64_t nSource,
                      uint64_t target, int loadBalance);
// After a cycle's first pass (cycles enqueued ahead): sets hold if the pass
// left secondaries in extraCount[0] or an error flag, copies cb to the pinned
// host block `host`, and unless held or setupNext is 0 does the next
// cycle's set-up from the census count the pass left.
void launchCycleTail(const LaunchCtx& ctx, CycleBlock* cb, CycleBlock* host, int setupNext, uint32_t procCap,
                     uint64_t nSource, uint64_t target, int loadBalance);

// Source particles at out[ctl->start ...]; with flux non-null the same
// launch also zeroes flux[0, nFlux) (the previous cycle's bins).
void launchSource(const LaunchCtx& ctx, const DevProblem& P, Particles out, const CycleCtl* ctl,
                  uint64_t nSource, uint32_t cycle, double* flux = nullptr, size_t nFlux = 0);

// Population control (split / Russian roulette, PopulationControl.cc:20-122)
// fused with low-weight roulette (PopulationControl.cc:127-171) over
// in[0, cb->ctl.n) with ctl's mode and factor; inCap bounds ctl.n.  The
// outputs go to out[0, cb->pc.nPopOut), the roulette and split counts to
// cb->dc.  groups: scratch of popGroups(inCap) entries.
size_t popGroups(uint32_t inCap);
void launchPopulation(const LaunchCtx& ctx, const DevProblem& P, Particles in, uint32_t inCap,
                      Particles out, uint32_t outCap, CycleBlock* cb, PopGroup* groups);

// One tracking pass over `in` (count read from device memory *nIn; nInHost
// only sizes the grid).  Census particles are appended to `census`
// (pc->nCensus), secondaries to `extra` (*nExtraOut).
// Blocks a tracking pass over nIn queued particles launches (persistent
// grid of tune.wavesPerCU waves per CU, fewer for small queues).
bool tablesFitLds(const DevProblem& P);
long long trackPassBlocks(const LaunchCtx& ctx, const TrackTuning& tune, uint32_t nIn);
long long batchTrackBlocks(const LaunchCtx& ctx, const TrackTuning& tune);
void launchTrackPass(const LaunchCtx& ctx, const DevProblem& P, const TrackTuning& tune, const WaveStacks& ws,
                     Particles in, const uint32_t* nIn, uint32_t nInHost,
                     Particles census, uint32_t censusCap,
                     Particles extra, uint32_t* nExtraOut, uint32_t extraCap,
                     PassCounters* pc, DevCounters* dc, const uint32_t* hold);

// Zeroes the flux bins for the next cycle (the cycle's flux sum is
// accumulated by the tracking kernel in DevCounters::fluxSum).
void launchZeroFlux(const LaunchCtx& ctx, double* flux, size_t n, const uint32_t* hold);

void launchClear(const LaunchCtx& ctx, void* ptr, size_t bytes);
// fluence[cell] += flux[cell][g] for g = 0..G-1 in order (Fluence::compute,
// Tallies.cc:100-120); call before the flux is zeroed.
void launchFluence(const LaunchCtx& ctx, const DevProblem& P, const uint32_t* hold);
// The same for every active problem of a batch that accumulates fluence.
void launchBatchFluence(const LaunchCtx& ctx, BatchSlot* slots, int n);

// Batched launchers over `n` slots.  prefix arrays hold n+1 entries.
void launchBatchPrep(const LaunchCtx& ctx, BatchSlot* slots, int n, uint64_t* srcPrefix, uint64_t* popPrefix);
void launchBatchSource(const LaunchCtx& ctx, BatchSlot* slots, int n, const uint64_t* srcPrefix);
void launchBatchPopulation(const LaunchCtx& ctx, BatchSlot* slots, int n, const uint64_t* popPrefix);
// passKind 0: pop -> extra[0]; 1: extra[0] -> extra[1]; 2: extra[1] -> extra[0]
void launchBatchPassPrep(const LaunchCtx& ctx, BatchSlot* slots, int n, int passKind);

// ---- particle sort by (cell, energy band) ----
// The tracking queue is reordered so that neighbouring lanes start in the
// same cell and energy band (counting sort; order within a key is
// arbitrary, which no tally depends on).  keys = nCells * kSortBands.
constexpr int kSortBands = 4;
inline size_t sortKeys(int nCells) { return (size_t)nCells * kSortBands; }
inline size_t scanBlocks(size_t m) { return (m + 1023) / 1024; }
// Plain form: sorts in[0, *n) into out; counts has sortKeys(nCells)
// entries, blockSums scanBlocks(keys).
void launchSort(const LaunchCtx& ctx, const DevProblem& P, Particles in, const uint32_t* n, Particles out,
                uint32_t* counts, uint32_t* blockSums, const uint32_t* hold);
// Batched form: every problem whose cycle is a multiple of `interval` sorts
// its population-control output into proc[cycle & 1] and marks the state
// sorted.  keyBase: [n + 1] exclusive prefix of the problems' key counts;
// sortPrefix: [n + 1] scratch.
void launchBatchSort(const LaunchCtx& ctx, BatchSlot* slots, int n, int interval, const uint64_t* keyBase,
                     uint64_t* sortPrefix, uint32_t* counts, size_t totalKeys, uint32_t* blockSums);
void launchBatchTrack(const LaunchCtx& ctx, BatchSlot* slots, int n, int passKind, const TrackTuning& tune,
                      const WaveStacks& ws,
                      bool stagedEnergies);
void launchBatchFinalize(const LaunchCtx& ctx, BatchSlot* slots, int n, int partsPerProblem, double* fluxSums);

// Device copies of single FP64 builtins, for rounding checks.
// Nearest facet (with the nudge retry) of n rays: xyz, dir [n][3], cell [n]
// (global ids, checked by the caller); xyzOut receives the possibly nudged
// positions.
void launchNearestProbe(const LaunchCtx& ctx, const DevProblem& P, const double* xyz, const double* dir,
                        const int32_t* cell, double* dist, int32_t* facet, double* xyzOut, size_t n);
void launchMathProbe(const LaunchCtx& ctx, int op, const double* a, const double* b, double* out, size_t n);

}  // namespace qs
