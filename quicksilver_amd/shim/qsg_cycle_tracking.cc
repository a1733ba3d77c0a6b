// cycleTracking(MonteCarlo*) (main.cc:31,138-307) over the device path.
//
// The reference's one-rank loop tracks every processing vault into the
// processed vaults, folding secondaries back in until none are left
// (main.cc:138-307, MC_Particle_Buffer.cc:298-312).  Here one call:
//   1. gathers the processing vaults' particles (ParticleVaultContainer.hh:
//      43-61, MC_Base_Particle.hh:27-100) into structure-of-arrays form and
//      empties the vaults;
//   2. tracks them on the device to census, escape or absorption, with every
//      secondary (qsg_cycle_begin_host + qsg_cycle_tracking);
//   3. writes the census particles into the processed vaults, where the
//      next cycleInit finds them (swapProcessingProcessedVaults,
//      ParticleVaultContainer.cc:278);
//   4. adds the cycle's tracking counters to _balanceTask[0]
//      (Tallies.hh:36-100) and the per-(cell, group) scalar flux to
//      _scalarFluxDomain[0] (Tallies.hh:306, TallyScalarFlux at
//      Tallies.hh:351-354), which cycleFinalize sums and clears.
// Cells are domain-local in the vaults and global (i + nx*(j + ny*k)) on
// the device; the map is the domain's cell_state ids (the global id in
// their upper 32 bits).
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "MonteCarlo.hh"
#include "MC_Domain.hh"
#include "MC_Base_Particle.hh"
#include "ParticleVault.hh"
#include "ParticleVaultContainer.hh"
#include "Tallies.hh"

#include "qsg.h"
#include "qsg_shim.h"

namespace {

qsg_problem g_prob = nullptr;
int g_groups = 0;
std::vector<int32_t> g_localToGlobal;  // domain-local cell -> global cell id
std::vector<int32_t> g_globalToLocal;

// One particle set in structure-of-arrays form, and its C view.
struct Soa {
    std::vector<double> x, y, z, vx, vy, vz, e, w, tc, nmfp;
    std::vector<uint64_t> seed;
    std::vector<int32_t> cell;
    void resize(size_t n) {
        for (std::vector<double>* v : {&x, &y, &z, &vx, &vy, &vz, &e, &w, &tc, &nmfp}) v->resize(n);
        seed.resize(n);
        cell.resize(n);
    }
    qsg_particles view() {
        return qsg_particles{x.data(), y.data(), z.data(), vx.data(), vy.data(), vz.data(), e.data(),
                             w.data(), tc.data(), nmfp.data(), seed.data(), cell.data()};
    }
};

// The reference reports nothing when a cycle goes wrong (qs_assert prints
// and continues, qs_assert.hh:9-26); a failed device cycle leaves no
// consistent state to continue from, so the shim stops the run.
void check(int rc, const char* what) {
    if (rc == QSG_OK) return;
    fprintf(stderr, "cycleTracking (device): %s failed: %s\n", what, qsg_last_error());
    exit(2);
}

}  // namespace

int qsg_shim_attach(MonteCarlo* mc, int argc, const char* const* argv, int device) {
    if (mc->domain.size() != 1) {
        fprintf(stderr, "qsg_shim_attach: one spatial domain is supported (xDom = yDom = zDom = 1)\n");
        return QSG_ERR_ARG;
    }
    int rc = qsg_create(argc, argv, device, &g_prob);
    if (rc != QSG_OK) {
        fprintf(stderr, "qsg_shim_attach: %s\n", qsg_last_error());
        return rc;
    }
    qsg_info info;
    qsg_get_info(g_prob, &info);
    MC_Domain& dom = mc->domain[0];
    const int nLocal = (int)dom.cell_state.size();
    ScalarFluxTask& task = mc->_tallies->_scalarFluxDomain[0]._task[0];
    if ((uint64_t)nLocal != info.n_cells || (int)task._cell.size() != nLocal ||
        task._cell[0].size() != info.n_groups) {
        fprintf(stderr, "qsg_shim_attach: the device problem's mesh or groups differ from the MonteCarlo's\n");
        qsg_destroy(g_prob);
        g_prob = nullptr;
        return QSG_ERR_ARG;
    }
    g_groups = info.n_groups;
    g_localToGlobal.assign(nLocal, -1);
    g_globalToLocal.assign(nLocal, -1);
    for (int c = 0; c < nLocal; c++) {
        const uint64_t gid = dom.cell_state[c]._id >> 32;
        if (gid >= (uint64_t)nLocal || g_globalToLocal[gid] != -1) {
            fprintf(stderr, "qsg_shim_attach: cell ids are not a permutation of the mesh\n");
            qsg_destroy(g_prob);
            g_prob = nullptr;
            return QSG_ERR_ARG;
        }
        g_localToGlobal[c] = (int32_t)gid;
        g_globalToLocal[gid] = c;
    }
    return QSG_OK;
}

void qsg_shim_detach() {
    if (g_prob) qsg_destroy(g_prob);
    g_prob = nullptr;
}

void cycleTracking(MonteCarlo* mc) {
    if (!g_prob) {
        fprintf(stderr, "cycleTracking (device): qsg_shim_attach was not called\n");
        exit(2);
    }
    ParticleVaultContainer& pvc = *mc->_particleVaultContainer;

    // 1. the processing vaults, emptied as they are read
    Soa in;
    in.resize(pvc.sizeProcessing());
    size_t k = 0;
    for (uint64_t v = 0; v < pvc.processingSize(); v++) {
        ParticleVault* vault = pvc.getTaskProcessingVault(v);
        for (size_t i = 0; i < vault->size(); i++, k++) {
            const MC_Base_Particle& b = (*vault)[i];
            in.x[k] = b.coordinate.x;
            in.y[k] = b.coordinate.y;
            in.z[k] = b.coordinate.z;
            in.vx[k] = b.velocity.x;
            in.vy[k] = b.velocity.y;
            in.vz[k] = b.velocity.z;
            in.e[k] = b.kinetic_energy;
            in.w[k] = b.weight;
            in.tc[k] = b.time_to_census;
            in.nmfp[k] = b.num_mean_free_paths;
            in.seed[k] = b.random_number_seed;
            in.cell[k] = g_localToGlobal[b.cell];
        }
        vault->clear();
    }
    in.resize(k);

    // 2. the device cycle
    qsg_particles view = in.view();
    check(qsg_cycle_begin_host(g_prob, &view, k), "qsg_cycle_begin_host");
    check(qsg_cycle_tracking(g_prob), "qsg_cycle_tracking");
    uint64_t nCensus = 0;
    check(qsg_get_census_count(g_prob, &nCensus), "qsg_get_census_count");
    Soa out;
    out.resize(nCensus);
    qsg_particles outView = out.view();
    check(qsg_get_census(g_prob, &outView, nCensus), "qsg_get_census");
    const size_t nCells = g_localToGlobal.size();
    std::vector<double> flux(nCells * (size_t)g_groups);
    check(qsg_get_scalar_flux(g_prob, flux.data(), flux.size()), "qsg_get_scalar_flux");
    qsg_balance bal;
    double fluxSum = 0.0;
    check(qsg_cycle_finalize(g_prob, &bal, &fluxSum), "qsg_cycle_finalize");

    // 3. census particles into the processed vaults (vaultSize each)
    const uint64_t vaultSize = pvc.getVaultSize();
    ParticleVault* dst = pvc.getTaskProcessedVault(pvc.getFirstEmptyProcessedVault());
    for (size_t j = 0; j < nCensus; j++) {
        if (dst->size() >= vaultSize) dst = pvc.getTaskProcessedVault(pvc.getFirstEmptyProcessedVault());
        MC_Base_Particle b;
        b.coordinate.x = out.x[j];
        b.coordinate.y = out.y[j];
        b.coordinate.z = out.z[j];
        b.velocity.x = out.vx[j];
        b.velocity.y = out.vy[j];
        b.velocity.z = out.vz[j];
        b.kinetic_energy = out.e[j];
        b.weight = out.w[j];
        b.time_to_census = out.tc[j];
        b.num_mean_free_paths = out.nmfp[j];
        b.random_number_seed = out.seed[j];
        b.identifier = out.seed[j];
        b.species = 0;
        b.domain = 0;
        b.cell = g_globalToLocal[out.cell[j]];
        dst->pushBaseParticle(b);
    }

    // 4. tallies: the tracking counters (cycleInit booked start, source,
    // rr and split) and the flux bins
    Balance& B = mc->_tallies->_balanceTask[0];
    B._absorb += bal.absorb;
    B._scatter += bal.scatter;
    B._fission += bal.fission;
    B._produce += bal.produce;
    B._collision += bal.collision;
    B._escape += bal.escape;
    B._census += bal.census;
    B._numSegments += bal.num_segments;
    ScalarFluxTask& task = mc->_tallies->_scalarFluxDomain[0]._task[0];
    for (size_t c = 0; c < nCells; c++) {
        const double* f = &flux[(size_t)g_localToGlobal[c] * g_groups];
        double* bins = task._cell[c]._group;
        for (int g = 0; g < g_groups; g++) bins[g] += f[g];
    }
}
