// Input surface of the qs driver: command line + input deck.
//
// Same keys, defaults and precedence as the reference:
//   struct layout/defaults    Parameters.hh:12-175
//   getParameters            Parameters.cc:80-95  (CLI first, then the deck
//                            overrides it, except -e / -S which win)
//   deck grammar             parseUtils.cc:17-111 (Block: line, indented
//                            "key: value" or "key=value", // comments)
//   value conversion         InputBlock.hh getValue (istream >> T)
//   CLI flags                Parameters.cc:219-255 (getopt_long)
//   default problem          Parameters.cc:350-379 (supplyDefaults)
// Unknown keys are ignored, duplicated keys: the last one wins.
#pragma once

#include <cstdint>
#include <map>
#include <ostream>
#include <string>
#include <vector>

namespace qs {

struct GeometryParameters {
    enum Shape { UNDEFINED, BRICK, SPHERE };
    std::string materialName;
    Shape shape = UNDEFINED;
    double radius = 0.0;
    double xCenter = 0.0, yCenter = 0.0, zCenter = 0.0;
    double xMin = 0.0, yMin = 0.0, zMin = 0.0;
    double xMax = 0.0, yMax = 0.0, zMax = 0.0;
};

struct MaterialParameters {
    std::string name;
    double mass = 1000.0;
    double totalCrossSection = 1.0;
    int nIsotopes = 10;
    int nReactions = 9;
    double sourceRate = 0.0;
    std::string scatteringCrossSection;
    std::string absorptionCrossSection;
    std::string fissionCrossSection;
    double scatteringCrossSectionRatio = 1.0;
    double absorptionCrossSectionRatio = 1.0;
    double fissionCrossSectionRatio = 1.0;
};

struct CrossSectionParameters {
    std::string name;
    double aa = 0.0, bb = 0.0, cc = 0.0, dd = 0.0, ee = 1.0;
    double nuBar = 2.4;
};

struct SimulationParameters {
    std::string inputFile;
    std::string energySpectrum;
    std::string crossSectionsOut;
    std::string boundaryCondition = "reflect";
    int loadBalance = 0;
    int cycleTimers = 0;
    int debugThreads = 0;
    uint64_t nParticles = 1000000;
    uint64_t batchSize = 0;
    uint64_t nBatches = 10;
    int nSteps = 10;
    int nx = 10, ny = 10, nz = 10;
    int seed = 1029384756;
    int xDom = 0, yDom = 0, zDom = 0;
    double dt = 1e-8;
    double fMax = 0.1;
    double lx = 100.0, ly = 100.0, lz = 100.0;
    double eMin = 1e-9;
    double eMax = 20;
    int nGroups = 230;
    double lowWeightCutoff = 0.001;
    int balanceTallyReplications = 1;
    int fluxTallyReplications = 1;
    int cellTallyReplications = 1;
    int coralBenchmark = 0;
};

struct Parameters {
    SimulationParameters simulationParams;
    std::vector<GeometryParameters> geometryParams;
    std::map<std::string, MaterialParameters> materialParams;
    std::map<std::string, CrossSectionParameters> crossSectionParams;
};

// Parse argv (argv[0] is the program name) and, if -i/--inputFile is given,
// the deck.  Throws std::runtime_error with a message on malformed input
// (the reference prints an assertion and continues; callers of this build
// get an error status instead).
Parameters getParameters(int argc, const char* const* argv);

// Parse a deck held in memory on top of ``pp`` (used for decks passed as text
// through the C ABI).  Same grammar as a deck file.
void parseDeckText(const std::string& text, Parameters& pp);

// supplyDefaults (Parameters.cc:350-379): adds the default flat problem when
// no Geometry block was given.
void supplyDefaults(Parameters& pp);

// "The output of every run is a valid input file" (README.md:62-64).
void printParameters(const Parameters& pp, std::ostream& out);

}  // namespace qs
