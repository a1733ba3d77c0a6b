// qs: command-line driver with the reference's surface (same flags and deck
// grammar, Parameters.cc:80-255) running the per-cycle loop of main.cc:38-85
// on one GPU through the C ABI.
//
// Output: the parameters as a re-runnable deck, one line per cycle in the
// layout of Tallies::PrintSummary (Tallies.cc:123-144) with the scalar flux
// printed at full precision (%.17e), the figure of merit of
// MC_Fast_Timer.cc:97-104 (segments / tracking seconds), the CORAL2 verdicts
// for CORAL2 decks, and the files the reference writes: <crossSectionsOut>.dat
// (-S, at set-up) and <energySpectrum>.dat (-e, at the end).
#include <cinttypes>
#include <cstdio>
#include <iostream>
#include <stdexcept>
#include <vector>

#include "qs_params.h"
#include "qsg.h"

int main(int argc, char** argv) {
    qs::Parameters params;
    try {
        params = qs::getParameters(argc, argv);
        qs::printParameters(params, std::cout);
    } catch (const std::exception& e) {
        fprintf(stderr, "qs: %s\n", e.what());
        return 2;
    }
    std::cout.flush();
    qsg_problem p = nullptr;
    if (qsg_create(argc, argv, 0, &p) != QSG_OK) {
        fprintf(stderr, "qs: %s\n", qsg_last_error());
        return 1;
    }
    qsg_info info;
    qsg_get_info(p, &info);
    printf("%-8s %12s %12s %12s %12s %12s %12s %12s %12s %12s %12s %12s %12s %24s %14s\n", "cycle", "start",
           "source", "rr", "split", "absorb", "scatter", "fission", "produce", "collisn", "escape", "census",
           "num_seg", "scalar_flux", "cycleTracking");
    uint64_t totalSeg = 0;
    double totalTrack = 0.0;
    for (int c = 0; c < info.n_steps; c++) {
        qsg_balance b;
        double flux = 0, secs = 0;
        int rc = qsg_run(p, 1, &b, &flux, &secs);
        if (rc != QSG_OK) {
            fprintf(stderr, "qs: cycle %d: %s\n", c, qsg_last_error());
            qsg_destroy(p);
            return 1;
        }
        printf("%8i %12" PRIu64 " %12" PRIu64 " %12" PRIu64 " %12" PRIu64 " %12" PRIu64 " %12" PRIu64
               " %12" PRIu64 " %12" PRIu64 " %12" PRIu64 " %12" PRIu64 " %12" PRIu64 " %12" PRIu64
               " %24.17e %14e\n",
               c, b.start, b.source, b.rr, b.split, b.absorb, b.scatter, b.fission, b.produce, b.collision,
               b.escape, b.census, b.num_segments, flux, secs);
        fflush(stdout);
        totalSeg += b.num_segments;
        totalTrack += secs;
    }
    printf("Figure Of Merit  %14.2e [Num Segments / Cycle Tracking Time]  (%" PRIu64 " segments, %.6e s)\n",
           totalTrack > 0 ? (double)totalSeg / totalTrack : 0.0, totalSeg, totalTrack);
    // gameOver (main.cc:87-94): the census energy spectrum
    // (EnergySpectrum::PrintSpectrum, EnergySpectrum.cc:37-58)
    if (!params.simulationParams.energySpectrum.empty()) {
        const uint64_t n = (uint64_t)info.n_groups + 1;
        std::vector<double> e(n);
        std::vector<uint64_t> counts(n);
        if (qsg_get_energy_spectrum(p, e.data(), counts.data(), n) != QSG_OK) {
            fprintf(stderr, "qs: %s\n", qsg_last_error());
            qsg_destroy(p);
            return 1;
        }
        const std::string name = params.simulationParams.energySpectrum + ".dat";
        FILE* f = fopen(name.c_str(), "w");
        if (!f) {
            fprintf(stderr, "qs: cannot write %s\n", name.c_str());
            qsg_destroy(p);
            return 1;
        }
        for (uint64_t g = 0; g < n; g++) fprintf(f, "%d\t%g\t%" PRIu64 "\n", (int)g, e[g], counts[g]);
        fclose(f);
    }
    // coralBenchmarkCorrectness (main.cc:73): the CORAL2 checks, printed in
    // the reference's words
    qsg_coral_report r;
    if (qsg_coral_check(p, &r) == QSG_OK) {
        printf("\nTesting Ratios for Absorbtion, Fission, and Scattering are maintained\n");
        if (r.ratio_pass)
            printf("PASS:: Absorption / Fission / Scatter Ratios maintained with %g%% tolerance\n",
                   r.ratio_tolerance * 100.0);
        else
            printf("FAIL:: Absorption / Fission / Scatter Ratios NOT maintained with %g%% tolerance\n",
                   r.ratio_tolerance * 100.0);
        printf("\nTesting balance between number of facet crossings and reactions\n");
        if (r.event_pass)
            printf("PASS:: Collision to Facet Crossing Ratio maintained even balanced within %g%% tolerance\n", 1.0);
        else
            printf(" FAIL:: Collision to Facet Crossing Ratio balanced NOT maintained within %g%% tolerance\n", 1.0);
        printf("\nTest for lost / unaccounted for particles in this simulation\n");
        if (r.missing_pass)
            printf("PASS:: No Particles Lost During Run\n");
        else
            printf("FAIL:: Particles Were Lost During Run, test for done should have failed\n");
        printf("\nTest Fluence for homogeneity across cells\n");
        if (r.fluence_pass)
            printf("PASS:: Fluence is homogenous across cells with %g%% tolerance\n", 6.0);
        else
            printf("FAIL:: Fluence not homogenous across cells within %g%% tolerance\n"
                   "\tCurrent Max Percent Diff: %4.1f%%\n", 6.0, r.fluence_max_percent_diff);
    }
    qsg_destroy(p);
    return 0;
}
