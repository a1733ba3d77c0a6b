// Test infrastructure only: the reference's own nearest-facet search on given
// rays.  Links the unmodified Quicksilver objects (oracle/Makefile) and calls
// MCT_Nearest_Facet (MCT.cc:87-137) exactly as MC_Segment_Outcome does
// (MC_Segment_Outcome.cc:102-117: distance threshold 10 x huge, current best
// huge, a new segment), on the mesh initMC builds from the deck.
//
//   qs_ref_mct -i deck.inp rays.bin out.bin
// rays.bin: uint64 n, then n x {x y z a b g} (f64) and n x cell (int32,
// global ids i + nx*(j + ny*k)).  out.bin: n x distance (f64), n x facet
// (int32), n x {x y z} after any nudge toward the cell centre (f64).
#include <cinttypes>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "Parameters.hh"
#include "utilsMpi.hh"
#include "MonteCarlo.hh"
#include "initMC.hh"
#include "MC_Domain.hh"
#include "MC_Particle.hh"
#include "MC_Base_Particle.hh"  // inline MC_Particle constructors
#include "MC_Location.hh"
#include "MC_Vector.hh"
#include "DirectionCosine.hh"
#include "MC_Nearest_Facet.hh"
#include "MCT.hh"
#include "PhysicalConstants.hh"

MonteCarlo* mcco = NULL;  // declared extern in Globals.hh

int main(int argc, char** argv) {
    if (argc < 5) {
        fprintf(stderr, "usage: qs_ref_mct -i deck.inp rays.bin out.bin\n");
        return 2;
    }
    const char* raysPath = argv[argc - 2];
    const char* outPath = argv[argc - 1];
    int pargc = argc - 2;
    mpiInit(&pargc, &argv);
    Parameters params = getParameters(pargc, argv);
    mcco = initMC(params);
    if (mcco->domain.size() != 1) {
        fprintf(stderr, "qs_ref_mct: one spatial domain expected\n");
        return 2;
    }
    MC_Domain& dom = mcco->domain[0];
    const int nCells = (int)dom.cell_state.size();
    std::vector<int> local(nCells, -1);
    for (int c = 0; c < nCells; c++) local[dom.cell_state[c]._id >> 32] = c;

    FILE* f = fopen(raysPath, "rb");
    if (!f) return 2;
    uint64_t n = 0;
    if (fread(&n, sizeof n, 1, f) != 1) return 2;
    std::vector<double> ray(6 * n);
    std::vector<int32_t> cell(n);
    if (fread(ray.data(), sizeof(double), 6 * n, f) != 6 * n) return 2;
    if (fread(cell.data(), sizeof(int32_t), n, f) != n) return 2;
    fclose(f);

    std::vector<double> dist(n), moved(3 * n);
    std::vector<int32_t> facet(n);
    for (uint64_t q = 0; q < n; q++) {
        MC_Particle particle;
        particle.num_segments = 0.0;
        MC_Location location(0, local[cell[q]], 0);
        MC_Vector coordinate(ray[6 * q], ray[6 * q + 1], ray[6 * q + 2]);
        DirectionCosine direction(ray[6 * q + 3], ray[6 * q + 4], ray[6 * q + 5]);
        const double threshold = 10.0 * PhysicalConstants::_hugeDouble;
        const double best = PhysicalConstants::_hugeDouble;
        MC_Nearest_Facet nf =
            MCT_Nearest_Facet(&particle, location, coordinate, &direction, threshold, best, true, mcco);
        dist[q] = nf.distance_to_facet;
        facet[q] = nf.facet;
        moved[3 * q] = coordinate.x;
        moved[3 * q + 1] = coordinate.y;
        moved[3 * q + 2] = coordinate.z;
    }
    FILE* o = fopen(outPath, "wb");
    if (!o) return 2;
    fwrite(dist.data(), sizeof(double), n, o);
    fwrite(facet.data(), sizeof(int32_t), n, o);
    fwrite(moved.data(), sizeof(double), 3 * n, o);
    fclose(o);
    delete mcco;
    mpiFinalize();
    return 0;
}
