// Test infrastructure only: a driver for the UNMODIFIED Quicksilver CPU
// sources (compiled in place from the reference tree by oracle/Makefile).
//
// The reference's own driver (src/main.cc) cannot be compiled here because it
// includes build-generated headers (git_hash.hh, git_vers.hh, main.cc:26-27).
// This file is a small single-rank driver written for this repository.  It
// calls the reference's public functions in the order main.cc uses them:
//   getParameters            Parameters.cc:80
//   initMC                   initMC.cc:53
//   per step: cycle_init  -> main.cc:96-121   (source, population control)
//             cycle_track -> main.cc:138-307  (tracking loop, 1 rank, no MPI)
//             cycle_final -> main.cc:310-324  (tally reduction / print)
// and adds full-precision output so the scalar flux can be checked to 1e-10
// (the reference prints it with %14e, Tallies.cc:136-141).
//
// Extra output, one line per cycle on stdout:
//   QSREF <cycle> <start> <source> <rr> <split> <absorb> <scatter> <fission>
//         <produce> <collision> <escape> <census> <num_seg> <end>
//         <flux_sum %.17e> <tracking_seconds>
// and at the end
//   QSREF_FOM <total segments> <total tracking seconds> <segments/s> <threads>
//
// Environment (optional):
//   QSREF_FLUX_DUMP=path   append, per cycle, the scalar flux of every
//                          (global cell id, group) as raw float64 in
//                          global-cell-major order (requires one domain).
//   QSREF_FLUENCE_DUMP=path after the run, the cumulative per-cell fluence
//                          (coralBenchmark decks), f64, global cell order.
//   QSREF_CENSUS_DUMP=path write the census particles of the last cycle as
//                          text: seed x y z vx vy vz E w tc nmfp cell
//   QSREF_PRINT_PARAMS=1   print the parameters first, as main.cc:44 does,
//                          then a line QSREF_PARAMS_END
//   QSREF_PRINT_ENERGIES=1 print the energy group bounds: QSREF_ENERGY g E (%.17e)
//   QSREF_INJECT=path      the first cycle tracks the particles in this file
//                          instead of the sourced, population-controlled
//                          ones (its balance then has start = n and source,
//                          rr, split = 0): uint64 n, n x {x y z vx vy vz E w
//                          tc nmfp} f64, n x seed u64, n x global cell i32
// At the end the census energy spectrum is written when energySpectrum (-e)
// is set, as gameOver does (main.cc:93); crossSectionsOut (-S) is written by
// initMC itself (initMC.cc:72).
//
// Built with -DQSREF_GPU_TRACKING (oracle/_ref/qs_ref_gpu), the same driver
// calls the drop-in cycleTracking(MonteCarlo*) of
// quicksilver_amd/shim/qsg_cycle_tracking.cc instead of the reference's
// tracking loop: everything else (source, population control, tallies,
// printout, CORAL2 checks) is still the reference's own code, so its output
// against the plain build's checks the drop-in on the reference's terms.
#include <chrono>
#include <iostream>
#include <cinttypes>
#include <cstdio>
#include <cstdlib>
#include <vector>
#include <algorithm>

#include "Parameters.hh"
#include "utilsMpi.hh"
#include "MonteCarlo.hh"
#include "initMC.hh"
#include "Tallies.hh"
#include "PopulationControl.hh"
#include "ParticleVaultContainer.hh"
#include "ParticleVault.hh"
#include "MC_Particle_Buffer.hh"
#include "MC_Base_Particle.hh"
#include "MC_Processor_Info.hh"
#include "MC_Time_Info.hh"
#include "MC_SourceNow.hh"
#include "MC_Domain.hh"
#include "NuclearData.hh"
#include "CycleTracking.hh"
#include "SendQueue.hh"
#include "MC_Fast_Timer.hh"
#include "CoralBenchmark.hh"

#ifdef HAVE_OPENMP
#include <omp.h>
#endif

#ifdef QSREF_GPU_TRACKING
#include "qsg_shim.h"
#endif

MonteCarlo* mcco = NULL;  // declared extern in Globals.hh

namespace {

void cycle_init(MonteCarlo* mc, bool loadBalance)
{
    mc->clearCrossSectionCache();
    mc->_tallies->CycleInitialize(mc);
    ParticleVaultContainer& pvc = *mc->_particleVaultContainer;
    pvc.swapProcessingProcessedVaults();
    pvc.collapseProcessed();
    pvc.collapseProcessing();
    mc->_tallies->_balanceTask[0]._start = pvc.sizeProcessing();
    mc->particle_buffer->Initialize();
    MC_SourceNow(mc);
    PopulationControl(mc, loadBalance);
    RouletteLowWeightParticles(mc);
}

// Single rank, single process: no particle leaves the rank, so the MPI
// exchange of main.cc:243-275 reduces to clearing the processed vault and
// folding the extra (secondary) vaults back in; the loop ends when no
// processing particles remain (MC_Particle_Buffer.cc:298-312).
void cycle_track(MonteCarlo* mc)
{
    ParticleVaultContainer& pvc = *mc->_particleVaultContainer;
    bool done = false;
    while (!done)
    {
        for (uint64_t v = 0; v < pvc.processingSize(); v++)
        {
            uint64_t pv = pvc.getFirstEmptyProcessedVault();
            ParticleVault* processing = pvc.getTaskProcessingVault(v);
            ParticleVault* processed = pvc.getTaskProcessedVault(pv);
            int n = (int)processing->size();
            if (n != 0)
            {
#ifdef HAVE_OPENMP
#pragma omp parallel for schedule(static)
#endif
                for (int i = 0; i < n; i++)
                    CycleTrackingGuts(mc, i, processing, processed);
            }
            processing->clear();
            pvc.getSendQueue()->clear();
            pvc.cleanExtraVaults();
        }
        pvc.collapseProcessing();
        pvc.collapseProcessed();
        done = (pvc.sizeProcessing() == 0);
    }
}

double flux_sum(MonteCarlo* mc)
{
    // Same summation order as Tallies::ScalarFluxSum (Tallies.cc:146-171).
    double s = 0.0;
    Tallies* t = mc->_tallies;
    for (int d = 0; d < (int)t->_scalarFluxDomain.size(); d++)
        for (int r = 0; r < t->GetNumFluxReplications(); r++)
        {
            ScalarFluxTask& task = t->_scalarFluxDomain[d]._task[r];
            for (int c = 0; c < (int)task._cell.size(); c++)
                for (int g = 0; g < task._cell[c].size(); g++)
                    s += task._cell[c]._group[g];
        }
    return s;
}

void dump_flux(MonteCarlo* mc, FILE* f)
{
    Tallies* t = mc->_tallies;
    if (t->_scalarFluxDomain.size() != 1) { fprintf(stderr, "flux dump needs one domain\n"); exit(3); }
    ScalarFluxTask& task = t->_scalarFluxDomain[0]._task[0];
    int nc = (int)task._cell.size();
    int ng = task._cell[0].size();
    std::vector<double> out((size_t)nc * ng, 0.0);
    MC_Domain& dom = mc->domain[0];
    for (int c = 0; c < nc; c++)
    {
        uint64_t gid = dom.cell_state[c]._id >> 32;
        for (int g = 0; g < ng; g++)
        {
            double v = 0.0;
            for (int r = 0; r < t->GetNumFluxReplications(); r++)
                v += t->_scalarFluxDomain[0]._task[r]._cell[c]._group[g];
            out[gid * ng + g] = v;
        }
    }
    fwrite(out.data(), sizeof(double), out.size(), f);
}

// Cumulative per-cell fluence (Tallies::_fluence, accumulated by
// CycleFinalize when coralBenchmark is set), in global cell order.
void dump_fluence(MonteCarlo* mc, const char* path)
{
    Tallies* t = mc->_tallies;
    if (t->_fluence._domain.size() != 1) { fprintf(stderr, "fluence dump needs one domain\n"); exit(3); }
    FluenceDomain* fd = t->_fluence._domain[0];
    MC_Domain& dom = mc->domain[0];
    std::vector<double> out((size_t)fd->size(), 0.0);
    for (int c = 0; c < fd->size(); c++) out[dom.cell_state[c]._id >> 32] = fd->getCell(c);
    FILE* f = fopen(path, "wb");
    if (!f) return;
    fwrite(out.data(), sizeof(double), out.size(), f);
    fclose(f);
}

void dump_census(MonteCarlo* mc, const char* path)
{
    FILE* f = fopen(path, "w");
    if (!f) return;
    ParticleVaultContainer& pvc = *mc->_particleVaultContainer;
    // after cycle_track the census particles sit in the processed vaults
    for (uint64_t v = 0; v < pvc.processedSize(); v++)
    {
        ParticleVault* pv = pvc.getTaskProcessedVault(v);
        for (size_t i = 0; i < pv->size(); i++)
        {
            const MC_Base_Particle& p = (*pv)[i];
            fprintf(f, "%" PRIu64 " %.17e %.17e %.17e %.17e %.17e %.17e %.17e %.17e %.17e %.17e %d\n",
                    p.random_number_seed, p.coordinate.x, p.coordinate.y, p.coordinate.z,
                    p.velocity.x, p.velocity.y, p.velocity.z, p.kinetic_energy, p.weight,
                    p.time_to_census, p.num_mean_free_paths,
                    (int)(mc->domain[p.domain].cell_state[p.cell]._id >> 32));
        }
    }
    fclose(f);
}

// Replaces the processing population of the cycle just initialised by the
// particles of `path` (MC_Base_Particle fields, ParticleVaultContainer::
// addProcessingParticle, ParticleVaultContainer.cc:318-334).
void inject(MonteCarlo* mc, const char* path)
{
    FILE* f = fopen(path, "rb");
    if (!f) { fprintf(stderr, "cannot read %s\n", path); exit(3); }
    uint64_t n = 0;
    if (fread(&n, sizeof n, 1, f) != 1) exit(3);
    std::vector<double> v(10 * n);
    std::vector<uint64_t> seed(n);
    std::vector<int32_t> cell(n);
    if (fread(v.data(), sizeof(double), 10 * n, f) != 10 * n || fread(seed.data(), 8, n, f) != n ||
        fread(cell.data(), 4, n, f) != n)
        exit(3);
    fclose(f);
    MC_Domain& dom = mc->domain[0];
    std::vector<int> local(dom.cell_state.size(), -1);
    for (int c = 0; c < (int)dom.cell_state.size(); c++) local[dom.cell_state[c]._id >> 32] = c;
    ParticleVaultContainer& pvc = *mc->_particleVaultContainer;
    for (uint64_t q = 0; q < pvc.processingSize(); q++) pvc.getTaskProcessingVault(q)->clear();
    uint64_t fill = 0;
    for (uint64_t i = 0; i < n; i++)
    {
        MC_Base_Particle b;
        const double* p = &v[10 * i];
        b.coordinate.x = p[0];
        b.coordinate.y = p[1];
        b.coordinate.z = p[2];
        b.velocity.x = p[3];
        b.velocity.y = p[4];
        b.velocity.z = p[5];
        b.kinetic_energy = p[6];
        b.weight = p[7];
        b.time_to_census = p[8];
        b.num_mean_free_paths = p[9];
        b.random_number_seed = seed[i];
        b.identifier = seed[i];
        b.species = 0;
        b.domain = 0;
        b.cell = local[cell[i]];
        pvc.addProcessingParticle(b, fill);
    }
    Balance& bal = mc->_tallies->_balanceTask[0];
    bal._start = n;
    bal._source = bal._rr = bal._split = 0;
}

}  // namespace

int main(int argc, char** argv)
{
    mpiInit(&argc, &argv);
    Parameters params = getParameters(argc, argv);
    if (getenv("QSREF_PRINT_PARAMS")) {
        printParameters(params, std::cout);
        std::cout << "QSREF_PARAMS_END" << std::endl;
    }
    mcco = initMC(params);
    if (getenv("QSREF_PRINT_ENERGIES"))
        for (size_t g = 0; g < mcco->_nuclearData->_energies.size(); g++)
            printf("QSREF_ENERGY %d %.17e\n", (int)g, mcco->_nuclearData->_energies[g]);
#ifdef QSREF_GPU_TRACKING
    if (qsg_shim_attach(mcco, argc, (const char* const*)argv, 0) != 0) return 2;
#endif

    const char* fluxPath = getenv("QSREF_FLUX_DUMP");
    const char* censusPath = getenv("QSREF_CENSUS_DUMP");
    FILE* fluxFile = fluxPath ? fopen(fluxPath, "wb") : NULL;

    int threads = 1;
#ifdef HAVE_OPENMP
    threads = omp_get_max_threads();
#endif
    const int nSteps = params.simulationParams.nSteps;
    uint64_t totalSeg = 0;
    double totalTrack = 0.0;
    for (int step = 0; step < nSteps; step++)
    {
        cycle_init(mcco, params.simulationParams.loadBalance != 0);
        if (step == 0 && getenv("QSREF_INJECT")) inject(mcco, getenv("QSREF_INJECT"));
        auto t0 = std::chrono::steady_clock::now();
#ifdef QSREF_GPU_TRACKING
        cycleTracking(mcco);
#else
        cycle_track(mcco);
#endif
        auto t1 = std::chrono::steady_clock::now();
        double secs = std::chrono::duration<double>(t1 - t0).count();

        Balance& b = mcco->_tallies->_balanceTask[0];
        b._end = mcco->_particleVaultContainer->sizeProcessed();
        double fs = flux_sum(mcco);
        printf("QSREF %d %" PRIu64 " %" PRIu64 " %" PRIu64 " %" PRIu64 " %" PRIu64 " %" PRIu64
               " %" PRIu64 " %" PRIu64 " %" PRIu64 " %" PRIu64 " %" PRIu64 " %" PRIu64 " %" PRIu64
               " %.17e %.6e\n",
               step, b._start, b._source, b._rr, b._split, b._absorb, b._scatter, b._fission,
               b._produce, b._collision, b._escape, b._census, b._numSegments, b._end, fs, secs);
        totalSeg += b._numSegments;
        totalTrack += secs;
        if (fluxFile) dump_flux(mcco, fluxFile);
        if (censusPath && step == nSteps - 1) dump_census(mcco, censusPath);

        mcco->_tallies->CycleFinalize(mcco);
        mcco->time_info->cycle++;
        mcco->particle_buffer->Free_Memory();
    }
    if (fluxFile) fclose(fluxFile);
    printf("QSREF_FOM %" PRIu64 " %.6e %.6e %d\n", totalSeg, totalTrack,
           totalTrack > 0 ? (double)totalSeg / totalTrack : 0.0, threads);
    fflush(stdout);
    // main.cc:73: the CORAL2 correctness tests (no-op unless coralBenchmark)
    coralBenchmarkCorrectness(mcco, params);
    fflush(stdout);
    const char* fluencePath = getenv("QSREF_FLUENCE_DUMP");
    if (fluencePath && params.simulationParams.coralBenchmark) dump_fluence(mcco, fluencePath);
    // gameOver (main.cc:93): EnergySpectrum::PrintSpectrum, a no-op unless
    // energySpectrum is set
    mcco->_tallies->_spectrum.PrintSpectrum(mcco);
#ifdef QSREF_GPU_TRACKING
    qsg_shim_detach();
#endif
    delete mcco;
    mpiFinalize();
    return 0;
}
