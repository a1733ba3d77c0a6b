// Host-side problem tables: everything the device path reads and never writes.
//
// Built from Parameters exactly as the reference builds its MonteCarlo object
// (initMC.cc:53-74) so that every table entry is bit-identical:
//   energy grid              NuclearData.cc:105-119
//   reaction cross sections  NuclearData.cc:15-42, addIsotope :121-184
//   material database        initMC.cc:130-197 (std::map order, atom
//                            fraction 1/nIsotopes)
//   mesh nodes               GlobalFccGrid.cc:17-131 (14 nodes / cell)
//   facet planes             MC_Facet_Geometry.hh:18-39, MC_Domain.cc:115-140
//   cell volume / id / mat   MC_Domain.cc:334-395, 440-449
//   source weights           MC_SourceNow.cc:28-66
// The mesh is the reference's structured hex grid with one spatial domain
// (cell index == global cell id).  Nodes are not stored: every node
// coordinate of cell (i,j,k) is one of nine values
//   X0 = i*dx, X1 = (i+1)*dx, Xh = X0 + dx/2 (same for y, z)
// which the device recomputes with the same IEEE operations.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "qs_params.h"

namespace qs {

// MC_Subfacet_Adjacency_Event values used for the six outer faces.
enum BoundaryEvent { kBoundaryEscape = 1, kBoundaryReflect = 2 };

// NuclearDataReaction::Enum
enum ReactionType { kUndefined = 0, kScatter = 1, kAbsorption = 2, kFission = 3 };

struct MaterialTables {
    std::string name;
    double mass = 1000.0;
    int nIsotopes = 0;
    int nReactions = 0;
    double atomFraction = 0.0;  // 1/nIsotopes
    double nuBar = 0.0;
    double sourceRate = 0.0;
    std::vector<int> reactionType;  // [nReactions]
    std::vector<double> micro;      // [nReactions][G] microscopic XS (same for every isotope)
    std::vector<double> macro;      // [nReactions][G] atomFraction * 1.0 * micro
    std::vector<double> total;      // [G] weightedMacroscopicCrossSection
};

struct ProblemTables {
    // run control
    double dt = 0;
    uint64_t nParticles = 0;
    int nSteps = 0;
    double lowWeightCutoff = 0;
    int loadBalance = 0;
    double eMin = 0, eMax = 0;

    // grid
    int nx = 0, ny = 0, nz = 0;
    double dx = 0, dy = 0, dz = 0;
    int nCells = 0;
    int boundary[6] = {0, 0, 0, 0, 0, 0};  // +x -x +y -y +z -z
    std::vector<double> planes;            // [nCells][24][4] A B C D
    std::vector<int> cellMaterial;         // [nCells]
    std::vector<double> cellVolume;        // [nCells]
    std::vector<double> cellCenter;        // [nCells][3] MCT_Cell_Position_3D_G

    // nuclear data
    int nGroups = 0;
    std::vector<double> energies;  // [G+1]
    std::vector<MaterialTables> materials;
    int maxReactions = 0;

    // source (identical every cycle)
    double sourceWeight = 0;              // source_particle_weight
    std::vector<uint32_t> cellSource;     // particles sourced per cell per cycle
    std::vector<uint64_t> cellSourceOffset;  // exclusive prefix sum, [nCells+1]
    uint64_t sourcePerCycle = 0;
};

// Throws std::runtime_error on inconsistent input (unknown material or cross
// section name, cell centre outside every geometry, bad boundary condition).
ProblemTables buildProblem(const Parameters& params);

// Node coordinate helpers shared by the host builder and the restated
// device routines: basis offsets of GlobalFccGrid::nodeCoord and the
// (offset, basis) tuples of GlobalFccGrid::cornerTupleOffsets.
extern const int kCornerOffset[14][4];
extern const int kNodeIndirect[24][3];
extern const int kOpposingFacet[24];

}  // namespace qs
