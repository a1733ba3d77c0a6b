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

struct Particles {
    double* x;
    double* y;
    double* z;
    double* vx;
    double* vy;
    double* vz;
    double* E;
    double* w;
    double* tc;    // time to census
    double* nmfp;  // remaining mean free paths
    uint64_t* seed;
    int32_t* cell;
};

constexpr int kParticleFields = 12;
constexpr size_t kParticleBytes = 10 * sizeof(double) + sizeof(uint64_t) + sizeof(int32_t);

// Per-material scalar data (kept small so it can live in LDS).
struct MatInfo {
    int nIso;
    int nReact;
    double mass;
    double nuBar;
};

constexpr int kMaxMaterials = 16;
constexpr int kMaxReactions = 16;

struct DevProblem {
    // grid
    int nx, ny, nz, nCells;
    double dx, dy, dz;     // lx/nx ...
    double hdx, hdy, hdz;  // dx/2.0 ...
    int boundaryMask;      // 2 bits per face +x -x +y -y +z -z: 1 escape, 2 reflect
    const double* planes;
    const int32_t* cellMat;
    const double* cellVolume;
    const double* cellCenter;
    const uint64_t* srcOffset;
    // nuclear data
    int G;
    int nMat, maxR;
    const double* energies;  // [G+1]
    const double* total;     // [nMat][G]
    const double* macro;     // [nMat][maxR][G]
    const int32_t* reactType;  // [nMat][maxR]
    MatInfo mat[kMaxMaterials];
    // run constants
    double dt, eMin, eMax, sourceWeight, lowWeightCutoff;
    // tallies
    double* flux;  // [nCells][G]
};

// Balance counters accumulated on the device (Tallies.hh:36-100 Balance).
struct DevCounters {
    unsigned long long absorb, census, escape, collision, fission, produce, scatter, numSegments;
    unsigned long long source, rr, split;
    unsigned long long pad[5];
};

// Work-queue / append counters (device memory, cleared by the host).
struct PassCounters {
    uint32_t head;          // next input index handed to a wave
    uint32_t nCensus;       // census buffer fill (cumulative over the cycle)
    uint32_t nPopOut;       // population-control output fill
    uint32_t error;         // bit 0 census overflow, 1 extra overflow, 2 pop
                            // overflow, 3 runaway history, 4 nOut > 4
    uint32_t extraCount[2]; // secondaries appended by passes (ping-pong)
    uint32_t pad[2];
};

struct TrackTuning {
    int blockSize;     // threads per workgroup (multiple of 64)
    int wavesPerCU;    // persistent waves per CU
    int refillMin;     // idle lanes that trigger a refill
    int chunk;         // particles a wave claims per queue access
};

// ---- launchers (qs_kernels.hip); all asynchronous on `stream` ----
struct LaunchCtx {
    void* stream;   // hipStream_t
    int numCUs;
};

size_t tableLdsBytes(const DevProblem& P);

void launchSource(const LaunchCtx& ctx, const DevProblem& P, Particles out, uint32_t outBase,
                  uint64_t nSource, uint32_t cycle);

// Population control (split / Russian roulette, PopulationControl.cc:20-122)
// fused with low-weight roulette (PopulationControl.cc:127-171).
// mode: 0 none, 1 roulette (factor < 1), 2 split (factor > 1).
void launchPopulation(const LaunchCtx& ctx, const DevProblem& P, Particles in, uint32_t n,
                      Particles out, uint32_t outCap, int mode, double factor,
                      PassCounters* pc, DevCounters* dc);

// One tracking pass over `in` (count read from device memory *nIn; nInHost
// only sizes the grid).  Census particles are appended to `census`
// (pc->nCensus), secondaries to `extra` (*nExtraOut).
void launchTrackPass(const LaunchCtx& ctx, const DevProblem& P, const TrackTuning& tune,
                     Particles in, const uint32_t* nIn, uint32_t nInHost,
                     Particles census, uint32_t censusCap,
                     Particles extra, uint32_t* nExtraOut, uint32_t extraCap,
                     PassCounters* pc, DevCounters* dc);

// Sums the flux array into partials (deterministic order) and zeroes it.
void launchFluxSum(const LaunchCtx& ctx, double* flux, size_t n, double* partials, int nPartials,
                   double* result);

void launchClear(const LaunchCtx& ctx, void* ptr, size_t bytes);

// Device copies of single FP64 builtins, for rounding checks.
void launchMathProbe(const LaunchCtx& ctx, int op, const double* a, const double* b, double* out, size_t n);

}  // namespace qs
