// Deck / command-line parsing for the qs driver.  Behaviour follows
// Parameters.cc, parseUtils.cc and InputBlock.hh of the reference; the code is
// this repository's own.
#include "qs_params.h"

#include <getopt.h>

#include <cctype>
#include <cinttypes>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <stdexcept>

namespace qs {
namespace {

const char* kWhitespace = " \t\f\v\n\r";

struct Block {
    std::string name;
    std::map<std::string, std::string> kv;  // last duplicate wins

    template <typename T>
    void get(const char* key, T& value) const {
        auto it = kv.find(key);
        if (it == kv.end()) return;
        std::istringstream in(it->second);
        in >> value;
        if (!in) throw std::runtime_error("cannot parse value of '" + std::string(key) +
                                          "' in block " + name + ": '" + it->second + "'");
    }
    // A string key given with no value ("energySpectrum: " and
    // "crossSectionsOut:", as printParameters writes unset names,
    // Parameters.cc:124,151) stays empty: the reference's extraction fails,
    // its qs_assert only prints (InputBlock.hh:36-52), and the name is left
    // empty, so a printed deck runs again.
    void get(const char* key, std::string& value) const {
        auto it = kv.find(key);
        if (it == kv.end()) return;
        std::istringstream in(it->second);
        value.clear();
        in >> value;
    }
};

void trim(std::string& s) {
    size_t end = s.find_last_not_of(kWhitespace);
    if (end == std::string::npos) { s.clear(); return; }
    s.erase(end + 1);
    size_t beg = s.find_first_not_of(kWhitespace);
    s.erase(0, beg);
}

bool isBlankOrComment(std::string line) {
    size_t c = line.find("//");
    if (c != std::string::npos) line.erase(c);
    return line.find_last_not_of(kWhitespace) == std::string::npos;
}

// parseUtils.cc:65-84: indentation, keyword before the first ':' or '=',
// value after it.
bool splitLine(const std::string& line, std::string& key, std::string& value, int& indent) {
    indent = 0;
    while (indent < (int)line.size() && isspace((unsigned char)line[indent])) ++indent;
    size_t delim = line.find_first_of(":=", indent);
    if (delim == std::string::npos) return false;
    key = line.substr(indent, delim - indent);
    trim(key);
    value.clear();
    if (delim + 1 < line.size()) {
        value = line.substr(delim + 1);
        trim(value);
    }
    return true;
}

std::vector<Block> parseBlocks(std::istream& in) {
    std::vector<Block> blocks;
    std::string line;
    bool have = static_cast<bool>(std::getline(in, line));
    while (have) {
        std::string key, value;
        int indent;
        if (splitLine(line, key, value, indent) && indent == 0 && value.empty()) {
            blocks.push_back(Block{key, {}});
            // read the block body: indented key/value lines; blank/comment
            // lines are skipped; the first other line ends the block and is
            // examined again as a possible block start.
            have = false;
            while (std::getline(in, line)) {
                if (isBlankOrComment(line)) continue;
                std::string k, v;
                int ind;
                if (!splitLine(line, k, v, ind) || ind == 0) { have = true; break; }
                blocks.back().kv[k] = v;
            }
        } else {
            have = static_cast<bool>(std::getline(in, line));
        }
    }
    return blocks;
}

void scanSimulation(const Block& b, Parameters& pp) {
    SimulationParameters& sp = pp.simulationParams;
    b.get("energySpectrum", sp.energySpectrum);
    b.get("crossSectionsOut", sp.crossSectionsOut);
    b.get("boundaryCondition", sp.boundaryCondition);
    b.get("dt", sp.dt);
    b.get("fMax", sp.fMax);
    b.get("loadBalance", sp.loadBalance);
    b.get("cycleTimers", sp.cycleTimers);
    b.get("debugThreads", sp.debugThreads);
    b.get("lx", sp.lx);
    b.get("ly", sp.ly);
    b.get("lz", sp.lz);
    b.get("nParticles", sp.nParticles);
    b.get("batchSize", sp.batchSize);
    b.get("nBatches", sp.nBatches);
    b.get("nSteps", sp.nSteps);
    b.get("nx", sp.nx);
    b.get("ny", sp.ny);
    b.get("nz", sp.nz);
    b.get("seed", sp.seed);
    b.get("xDom", sp.xDom);
    b.get("yDom", sp.yDom);
    b.get("zDom", sp.zDom);
    b.get("eMax", sp.eMax);
    b.get("eMin", sp.eMin);
    b.get("nGroups", sp.nGroups);
    b.get("lowWeightCutoff", sp.lowWeightCutoff);
    b.get("bTally", sp.balanceTallyReplications);
    b.get("fTally", sp.fluxTallyReplications);
    b.get("cTally", sp.cellTallyReplications);
    b.get("coralBenchmark", sp.coralBenchmark);
}

void scanGeometry(const Block& b, Parameters& pp) {
    pp.geometryParams.push_back(GeometryParameters());
    GeometryParameters& g = pp.geometryParams.back();
    b.get("material", g.materialName);
    std::string shape;
    b.get("shape", shape);
    if (shape == "brick") {
        g.shape = GeometryParameters::BRICK;
        b.get("xMax", g.xMax);
        b.get("xMin", g.xMin);
        b.get("yMax", g.yMax);
        b.get("yMin", g.yMin);
        b.get("zMax", g.zMax);
        b.get("zMin", g.zMin);
    } else if (shape == "sphere") {
        g.shape = GeometryParameters::SPHERE;
        b.get("radius", g.radius);
        b.get("xCenter", g.xCenter);
        b.get("yCenter", g.yCenter);
        b.get("zCenter", g.zCenter);
    } else {
        throw std::runtime_error("Geometry block needs shape: brick or sphere");
    }
}

void scanMaterial(const Block& b, Parameters& pp) {
    std::string name;
    b.get("name", name);
    if (name.empty()) throw std::runtime_error("Material block without a name");
    MaterialParameters& m = pp.materialParams[name];
    m.name = name;
    b.get("mass", m.mass);
    b.get("absorptionCrossSection", m.absorptionCrossSection);
    b.get("absorptionCrossSectionRatio", m.absorptionCrossSectionRatio);
    b.get("fissionCrossSection", m.fissionCrossSection);
    b.get("fissionCrossSectionRatio", m.fissionCrossSectionRatio);
    b.get("nIsotopes", m.nIsotopes);
    b.get("nReactions", m.nReactions);
    b.get("totalCrossSection", m.totalCrossSection);
    b.get("scatteringCrossSection", m.scatteringCrossSection);
    b.get("scatteringCrossSectionRatio", m.scatteringCrossSectionRatio);
    b.get("sourceRate", m.sourceRate);
}

void scanCrossSection(const Block& b, Parameters& pp) {
    std::string name;
    b.get("name", name);
    if (name.empty()) throw std::runtime_error("CrossSection block without a name");
    CrossSectionParameters& c = pp.crossSectionParams[name];
    c.name = name;
    b.get("A", c.aa);
    b.get("B", c.bb);
    b.get("C", c.cc);
    b.get("D", c.dd);
    b.get("E", c.ee);
    b.get("nuBar", c.nuBar);
}

void applyBlocks(const std::vector<Block>& blocks, Parameters& pp) {
    for (const Block& b : blocks) {
        if (b.name == "Simulation") scanSimulation(b, pp);
        if (b.name == "Geometry") scanGeometry(b, pp);
        if (b.name == "Material") scanMaterial(b, pp);
        if (b.name == "CrossSection") scanCrossSection(b, pp);
    }
}

// One CLI option: long name, short letter, takes an argument, conversion.
struct Opt {
    const char* name;
    char letter;
    int hasArg;
    char type;  // 'i' int, 'u' uint64, 'd' double, 's' string, 'f' flag
    void* ptr;
};

void parseCommandLine(int argc, const char* const* argv, Parameters& pp,
                      std::string& input, std::string& spectrum, std::string& xsOut) {
    SimulationParameters& sp = pp.simulationParams;
    int help = 0;
    std::vector<Opt> opts = {
        {"help", 'h', 0, 'f', &help},
        {"dt", 'D', 1, 'd', &sp.dt},
        {"fMax", 'f', 1, 'd', &sp.fMax},
        {"inputFile", 'i', 1, 's', &input},
        {"energySpectrum", 'e', 1, 's', &spectrum},
        {"crossSectionsOut", 'S', 1, 's', &xsOut},
        {"loadBalance", 'l', 0, 'f', &sp.loadBalance},
        {"cycleTimers", 'c', 1, 'i', &sp.cycleTimers},
        {"debugThreads", 't', 1, 'i', &sp.debugThreads},
        {"lx", 'X', 1, 'd', &sp.lx},
        {"ly", 'Y', 1, 'd', &sp.ly},
        {"lz", 'Z', 1, 'd', &sp.lz},
        {"nParticles", 'n', 1, 'u', &sp.nParticles},
        {"batchSize", 'g', 1, 'u', &sp.batchSize},
        {"nBatches", 'b', 1, 'u', &sp.nBatches},
        {"nSteps", 'N', 1, 'i', &sp.nSteps},
        {"nx", 'x', 1, 'i', &sp.nx},
        {"ny", 'y', 1, 'i', &sp.ny},
        {"nz", 'z', 1, 'i', &sp.nz},
        {"seed", 's', 1, 'i', &sp.seed},
        {"xDom", 'I', 1, 'i', &sp.xDom},
        {"yDom", 'J', 1, 'i', &sp.yDom},
        {"zDom", 'K', 1, 'i', &sp.zDom},
        {"bTally", 'B', 1, 'i', &sp.balanceTallyReplications},
        {"fTally", 'F', 1, 'i', &sp.fluxTallyReplications},
        {"cTally", 'C', 1, 'i', &sp.cellTallyReplications},
    };
    std::vector<struct option> lopts;
    std::string shortSpec;
    for (const Opt& o : opts) {
        lopts.push_back({o.name, o.hasArg ? required_argument : no_argument, nullptr, o.letter});
        shortSpec += o.letter;
        if (o.hasArg) shortSpec += ':';
    }
    lopts.push_back({nullptr, 0, nullptr, 0});

    std::vector<char*> args;
    for (int i = 0; i < argc; i++) args.push_back(const_cast<char*>(argv[i]));
    args.push_back(nullptr);
    optind = 0;  // full re-initialisation of getopt state (glibc)
    opterr = 0;
    for (;;) {
        int idx = 0;
        int c = getopt_long(argc, args.data(), shortSpec.c_str(), lopts.data(), &idx);
        if (c == -1) break;
        const Opt* o = nullptr;
        for (const Opt& cand : opts)
            if (cand.letter == c) o = &cand;
        if (!o) throw std::runtime_error("invalid command-line switch");
        if (!o->hasArg) { *static_cast<int*>(o->ptr) = 1; continue; }
        switch (o->type) {
            case 'i': sscanf(optarg, "%d", static_cast<int*>(o->ptr)); break;
            case 'u': sscanf(optarg, "%" SCNu64, static_cast<uint64_t*>(o->ptr)); break;
            case 'd': sscanf(optarg, "%lf", static_cast<double*>(o->ptr)); break;
            case 's': *static_cast<std::string*>(o->ptr) = optarg; break;
        }
    }
    optind = 0;
    if (help) throw std::runtime_error("help requested");
}

}  // namespace

void parseDeckText(const std::string& text, Parameters& pp) {
    std::istringstream in(text);
    applyBlocks(parseBlocks(in), pp);
}

void supplyDefaults(Parameters& pp) {
    if (!pp.geometryParams.empty()) return;
    CrossSectionParameters flat;
    flat.name = "flat";
    pp.crossSectionParams[flat.name] = flat;

    MaterialParameters src;
    src.name = "sourceMaterial";
    src.mass = 1000.0;
    src.sourceRate = 1e10;
    src.scatteringCrossSection = "flat";
    src.absorptionCrossSection = "flat";
    src.fissionCrossSection = "flat";
    src.fissionCrossSectionRatio = 0.1;
    pp.materialParams[src.name] = src;

    GeometryParameters g;
    g.materialName = "sourceMaterial";
    g.shape = GeometryParameters::BRICK;
    g.xMax = pp.simulationParams.lx;
    g.yMax = pp.simulationParams.ly;
    g.zMax = pp.simulationParams.lz;
    pp.geometryParams.push_back(g);
}

Parameters getParameters(int argc, const char* const* argv) {
    Parameters pp;
    std::string input, spectrum, xsOut;
    parseCommandLine(argc, argv, pp, input, spectrum, xsOut);
    pp.simulationParams.inputFile = input;
    pp.simulationParams.energySpectrum = spectrum;
    pp.simulationParams.crossSectionsOut = xsOut;
    if (!input.empty()) {
        std::ifstream in(input.c_str());
        if (!in) throw std::runtime_error("cannot open input file " + input);
        applyBlocks(parseBlocks(in), pp);
    }
    if (!spectrum.empty()) pp.simulationParams.energySpectrum = spectrum;
    if (!xsOut.empty()) pp.simulationParams.crossSectionsOut = xsOut;
    supplyDefaults(pp);
    return pp;
}

void printParameters(const Parameters& pp, std::ostream& out) {
    const SimulationParameters& s = pp.simulationParams;
    out << "Simulation:\n"
        << "   dt: " << s.dt << "\n"
        << "   fMax: " << s.fMax << "\n"
        << "   inputFile: " << s.inputFile << "\n"
        << "   energySpectrum: " << s.energySpectrum << "\n"
        << "   boundaryCondition: " << s.boundaryCondition << "\n"
        << "   loadBalance: " << s.loadBalance << "\n"
        << "   cycleTimers: " << s.cycleTimers << "\n"
        << "   debugThreads: " << s.debugThreads << "\n"
        << "   lx: " << s.lx << "\n"
        << "   ly: " << s.ly << "\n"
        << "   lz: " << s.lz << "\n"
        << "   nParticles: " << s.nParticles << "\n"
        << "   batchSize: " << s.batchSize << "\n"
        << "   nBatches: " << s.nBatches << "\n"
        << "   nSteps: " << s.nSteps << "\n"
        << "   nx: " << s.nx << "\n"
        << "   ny: " << s.ny << "\n"
        << "   nz: " << s.nz << "\n"
        << "   seed: " << s.seed << "\n"
        << "   xDom: " << s.xDom << "\n"
        << "   yDom: " << s.yDom << "\n"
        << "   zDom: " << s.zDom << "\n"
        << "   eMax: " << s.eMax << "\n"
        << "   eMin: " << s.eMin << "\n"
        << "   nGroups: " << s.nGroups << "\n"
        << "   lowWeightCutoff: " << s.lowWeightCutoff << "\n"
        << "   bTally: " << s.balanceTallyReplications << "\n"
        << "   fTally: " << s.fluxTallyReplications << "\n"
        << "   cTally: " << s.cellTallyReplications << "\n"
        << "   coralBenchmark: " << s.coralBenchmark << "\n"
        << "   crossSectionsOut:" << s.crossSectionsOut << "\n\n";
    for (const GeometryParameters& g : pp.geometryParams) {
        out << "Geometry:\n   material: " << g.materialName << "\n";
        if (g.shape == GeometryParameters::BRICK) {
            out << "   shape: brick\n"
                << "   xMax: " << g.xMax << "\n   xMin: " << g.xMin << "\n"
                << "   yMax: " << g.yMax << "\n   yMin: " << g.yMin << "\n"
                << "   zMax: " << g.zMax << "\n   zMin: " << g.zMin << "\n";
        } else {
            out << "   shape: sphere\n"
                << "   radius: " << g.radius << "\n"
                << "   xCenter: " << g.xCenter << "\n   yCenter: " << g.yCenter << "\n"
                << "   zCenter: " << g.zCenter << "\n";
        }
        out << "\n";
    }
    for (const auto& it : pp.materialParams) {
        const MaterialParameters& m = it.second;
        out << "Material:\n"
            << "   name: " << m.name << "\n"
            << "   mass: " << m.mass << "\n"
            << "   nIsotopes: " << m.nIsotopes << "\n"
            << "   nReactions: " << m.nReactions << "\n"
            << "   sourceRate: " << m.sourceRate << "\n"
            << "   totalCrossSection: " << m.totalCrossSection << "\n"
            << "   absorptionCrossSection: " << m.absorptionCrossSection << "\n"
            << "   fissionCrossSection: " << m.fissionCrossSection << "\n"
            << "   scatteringCrossSection: " << m.scatteringCrossSection << "\n"
            << "   absorptionCrossSectionRatio: " << m.absorptionCrossSectionRatio << "\n"
            << "   fissionCrossSectionRatio: " << m.fissionCrossSectionRatio << "\n"
            << "   scatteringCrossSectionRatio: " << m.scatteringCrossSectionRatio << "\n\n";
    }
    for (const auto& it : pp.crossSectionParams) {
        const CrossSectionParameters& c = it.second;
        out << "CrossSection:\n"
            << "   name: " << c.name << "\n"
            << "   A: " << c.aa << "\n   B: " << c.bb << "\n   C: " << c.cc << "\n"
            << "   D: " << c.dd << "\n   E: " << c.ee << "\n"
            << "   nuBar: " << c.nuBar << "\n";
    }
}

}  // namespace qs
