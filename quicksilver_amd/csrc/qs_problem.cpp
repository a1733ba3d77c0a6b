// Host construction of the read-only problem tables (see qs_problem.h).
// Compiled with -ffp-contract=off so each expression rounds exactly like the
// reference's x86-64 (no FMA) build.
#include "qs_problem.h"

#include <algorithm>
#include <cmath>
#include <exception>
#include <map>
#include <stdexcept>
#include <thread>

namespace qs {

// Hex-cell node tuples (x, y, z, basis): 8 corners then 6 face centres
// (GlobalFccGrid.cc:47-69).  Basis 1/2/3 add half a cell in (y,z)/(x,z)/(x,y).
const int kCornerOffset[14][4] = {
    {0, 0, 0, 0}, {1, 0, 0, 0}, {0, 1, 0, 0}, {1, 1, 0, 0},
    {0, 0, 1, 0}, {1, 0, 1, 0}, {0, 1, 1, 0}, {1, 1, 1, 0},
    {1, 0, 0, 1}, {0, 0, 0, 1}, {0, 1, 0, 2}, {0, 0, 0, 2},
    {0, 0, 1, 3}, {0, 0, 0, 3}};

// Cell-local node triples of the 24 triangular facets; facets 4f..4f+3 lie on
// face f = +x, -x, +y, -y, +z, -z (MC_Domain.cc:42-47).
const int kNodeIndirect[24][3] = {
    {1, 3, 8}, {3, 7, 8}, {7, 5, 8}, {5, 1, 8},
    {0, 4, 9}, {4, 6, 9}, {6, 2, 9}, {2, 0, 9},
    {3, 2, 10}, {2, 6, 10}, {6, 7, 10}, {7, 3, 10},
    {0, 1, 11}, {1, 5, 11}, {5, 4, 11}, {4, 0, 11},
    {4, 5, 12}, {5, 7, 12}, {7, 6, 12}, {6, 4, 12},
    {0, 2, 13}, {2, 3, 13}, {3, 1, 13}, {1, 0, 13}};

// Facet seen from the neighbouring cell (MC_Domain.cc:49-51).
const int kOpposingFacet[24] = {7, 6, 5, 4, 3, 2, 1, 0, 12, 15, 14, 13,
                                8, 11, 10, 9, 20, 23, 22, 21, 16, 19, 18, 17};

namespace {

struct V3 {
    double x, y, z;
};

double polynomial(const CrossSectionParameters& c, double xx) {
    // Polynomial::operator() (NuclearData.hh:20-23)
    return c.ee + xx * (c.dd + xx * (c.cc + xx * (c.bb + xx * (c.aa))));
}

// NuclearDataReaction constructor: group-midpoint polynomial in log10(E),
// scaled so the group containing 1 MeV equals reactionCrossSection.
std::vector<double> reactionTable(const std::vector<double>& e, const CrossSectionParameters& poly,
                                  double reactionCrossSection) {
    const int G = (int)e.size() - 1;
    std::vector<double> xs(G, 0.0);
    for (int ii = 0; ii < G; ++ii) {
        double energy = (e[ii] + e[ii + 1]) / 2.0;
        xs[ii] = std::pow(10.0, polynomial(poly, std::log10(energy)));
    }
    double normalization = 0.0;
    for (int ii = 0; ii < G; ++ii)
        if (e[ii + 1] >= 1.) {
            normalization = xs[ii];
            break;
        }
    if (!(normalization > 0.))
        throw std::runtime_error("cross section " + poly.name + " normalisation is not positive");
    double scale = reactionCrossSection / normalization;
    for (int ii = 0; ii < G; ++ii) xs[ii] *= scale;
    return xs;
}

V3 nodeCoord(int i, int j, int k, int n, double dx, double dy, double dz) {
    const int* o = kCornerOffset[n];
    double rx = (i + o[0]) * dx;
    double ry = (j + o[1]) * dy;
    double rz = (k + o[2]) * dz;
    V3 b = {0.0, 0.0, 0.0};
    switch (o[3]) {
        case 1: b = {0.0, dy / 2.0, dz / 2.0}; break;
        case 2: b = {dx / 2.0, 0.0, dz / 2.0}; break;
        case 3: b = {dx / 2.0, dy / 2.0, 0.0}; break;
        default: break;
    }
    return {rx + b.x, ry + b.y, rz + b.z};
}

bool isInside(const GeometryParameters& g, const V3& r) {
    if (g.shape == GeometryParameters::BRICK)
        return (r.x >= g.xMin && r.x <= g.xMax) && (r.y >= g.yMin && r.y <= g.yMax) &&
               (r.z >= g.zMin && r.z <= g.zMax);
    if (g.shape == GeometryParameters::SPHERE) {
        double x = r.x - g.xCenter, y = r.y - g.yCenter, z = r.z - g.zCenter;
        return std::sqrt(x * x + y * y + z * z) <= g.radius;
    }
    throw std::runtime_error("geometry without a shape");
}

}  // namespace

ProblemTables buildProblem(const Parameters& params) {
    const SimulationParameters& sp = params.simulationParams;
    ProblemTables T;
    T.dt = sp.dt;
    T.nParticles = sp.nParticles;
    T.nSteps = sp.nSteps;
    T.lowWeightCutoff = sp.lowWeightCutoff;
    T.loadBalance = sp.loadBalance;
    T.eMin = sp.eMin;
    T.eMax = sp.eMax;
    if (sp.nx <= 0 || sp.ny <= 0 || sp.nz <= 0) throw std::runtime_error("mesh size must be positive");
    if (sp.nGroups <= 0) throw std::runtime_error("nGroups must be positive");
    if (!(sp.eMin < sp.eMax)) throw std::runtime_error("eMin must be below eMax");

    // ---- energy grid (NuclearData::NuclearData) ----
    const int G = sp.nGroups;
    T.nGroups = G;
    T.energies.assign(G + 1, 0.0);
    T.energies[0] = sp.eMin;
    T.energies[G] = sp.eMax;
    {
        double logLow = std::log(sp.eMin);
        double logHigh = std::log(sp.eMax);
        double delta = (logHigh - logLow) / (G + 1.0);
        for (int i = 1; i < G; i++) {
            double logValue = logLow + delta * i;
            T.energies[i] = std::exp(logValue);
        }
    }

    // ---- materials in std::map (name) order (initNuclearData) ----
    std::map<std::string, int> matIndex;
    for (const auto& it : params.materialParams) {
        const MaterialParameters& mp = it.second;
        MaterialTables m;
        m.name = mp.name;
        m.mass = mp.mass;
        m.nIsotopes = mp.nIsotopes;
        m.nReactions = mp.nReactions;
        m.sourceRate = mp.sourceRate;
        if (mp.nIsotopes <= 0) throw std::runtime_error("material " + mp.name + " has no isotopes");
        if (mp.nReactions <= 0) throw std::runtime_error("material " + mp.name + " has no reactions");
        auto xsOf = [&](const std::string& n) -> const CrossSectionParameters& {
            auto f = params.crossSectionParams.find(n);
            if (f == params.crossSectionParams.end())
                throw std::runtime_error("material " + mp.name + " names unknown cross section '" + n + "'");
            return f->second;
        };
        const CrossSectionParameters& fissionXs = xsOf(mp.fissionCrossSection);
        m.nuBar = fissionXs.nuBar;
        m.atomFraction = 1.0 / mp.nIsotopes;

        // NuclearData::addIsotope
        double totalWeight = mp.fissionCrossSectionRatio + mp.scatteringCrossSectionRatio +
                             mp.absorptionCrossSectionRatio;
        int nFission = mp.nReactions / 3, nScatter = mp.nReactions / 3, nAbsorption = mp.nReactions / 3;
        switch (mp.nReactions % 3) {
            case 1: ++nScatter; break;
            case 2: ++nScatter; ++nFission; break;
            default: break;
        }
        double fissionCS = (mp.totalCrossSection * mp.fissionCrossSectionRatio) / (nFission * totalWeight);
        double scatterCS = (mp.totalCrossSection * mp.scatteringCrossSectionRatio) / (nScatter * totalWeight);
        double absorbCS =
            (mp.totalCrossSection * mp.absorptionCrossSectionRatio) / (nAbsorption * totalWeight);
        m.reactionType.resize(mp.nReactions);
        m.micro.assign((size_t)mp.nReactions * G, 0.0);
        for (int r = 0; r < mp.nReactions; r++) {
            std::vector<double> xs;
            switch (r % 3) {
                case 0:
                    m.reactionType[r] = kScatter;
                    xs = reactionTable(T.energies, xsOf(mp.scatteringCrossSection), scatterCS);
                    break;
                case 1:
                    m.reactionType[r] = kFission;
                    xs = reactionTable(T.energies, xsOf(mp.fissionCrossSection), fissionCS);
                    break;
                default:
                    m.reactionType[r] = kAbsorption;
                    xs = reactionTable(T.energies, xsOf(mp.absorptionCrossSection), absorbCS);
                    break;
            }
            for (int g = 0; g < G; g++) m.micro[(size_t)r * G + g] = xs[g];
        }
        // macroscopicCrossSection: atomFraction * cellNumberDensity(=1.0) * micro
        const double cellNumberDensity = 1.0;
        m.macro.assign(m.micro.size(), 0.0);
        for (size_t q = 0; q < m.micro.size(); q++)
            m.macro[q] = m.atomFraction * cellNumberDensity * m.micro[q];
        // weightedMacroscopicCrossSection: sum over isotopes of
        // atomFraction * 1.0 * getTotalCrossSection (sum over reactions)
        m.total.assign(G, 0.0);
        for (int g = 0; g < G; g++) {
            double isoTotal = 0.0;
            for (int r = 0; r < mp.nReactions; r++) isoTotal += m.micro[(size_t)r * G + g];
            double macroTotal = m.atomFraction * cellNumberDensity * isoTotal;
            if (m.atomFraction == 0.0) macroTotal = 1e-20;
            double sum = 0.0;
            for (int iso = 0; iso < mp.nIsotopes; iso++) sum += macroTotal;
            m.total[g] = sum;
        }
        matIndex[m.name] = (int)T.materials.size();
        if (mp.nReactions > T.maxReactions) T.maxReactions = mp.nReactions;
        T.materials.push_back(std::move(m));
    }
    if (T.materials.empty()) throw std::runtime_error("no materials");

    // ---- boundary conditions (getBoundaryCondition, MC_Domain.cc:455-471) ----
    if (sp.boundaryCondition == "reflect") {
        for (int f = 0; f < 6; f++) T.boundary[f] = kBoundaryReflect;
    } else if (sp.boundaryCondition == "escape") {
        for (int f = 0; f < 6; f++) T.boundary[f] = kBoundaryEscape;
    } else if (sp.boundaryCondition == "octant") {
        for (int f = 0; f < 6; f++) T.boundary[f] = (f % 2 == 0) ? kBoundaryEscape : kBoundaryReflect;
    } else {
        throw std::runtime_error("unknown boundaryCondition '" + sp.boundaryCondition + "'");
    }

    // ---- mesh (GlobalFccGrid + MC_Domain) ----
    T.nx = sp.nx;
    T.ny = sp.ny;
    T.nz = sp.nz;
    T.dx = sp.lx / sp.nx;
    T.dy = sp.ly / sp.ny;
    T.dz = sp.lz / sp.nz;
    const long long nCells = (long long)sp.nx * sp.ny * sp.nz;
    if (nCells > (1LL << 31) - 1) throw std::runtime_error("mesh too large");
    T.nCells = (int)nCells;
    T.planes.assign((size_t)nCells * 24 * 4, 0.0);
    T.cellMaterial.assign(nCells, 0);
    T.cellVolume.assign(nCells, 0.0);
    T.cellCenter.assign((size_t)nCells * 3, 0.0);
    T.cellSource.assign(nCells, 0);
    T.cellSourceOffset.assign(nCells + 1, 0);

    std::vector<double> cellWeight(nCells, 0.0);
    // each geometry's material index (-1: a name no material has), and the
    // index a cell no geometry covers would take (the empty name's)
    std::vector<int> geomMat;
    for (const GeometryParameters& g : params.geometryParams) {
        auto mi = matIndex.find(g.materialName);
        geomMat.push_back(mi == matIndex.end() ? -1 : mi->second);
    }
    const auto emptyMi = matIndex.find(std::string());
    const int uncovered = emptyMi == matIndex.end() ? -1 : emptyMi->second;
    // Cells are independent: each writes its own entries only, so slabs of k
    // are built on several threads for large meshes (the results do not
    // depend on the split).
    auto buildSlab = [&](int k0, int k1) {
        for (int k = k0; k < k1; k++)
        for (int j = 0; j < sp.ny; j++)
            for (int i = 0; i < sp.nx; i++) {
                const size_t c = (size_t)i + (size_t)sp.nx * ((size_t)j + (size_t)sp.ny * k);
                V3 node[14];
                for (int n = 0; n < 14; n++) node[n] = nodeCoord(i, j, k, n, T.dx, T.dy, T.dz);

                // planes
                for (int f = 0; f < 24; f++) {
                    const V3& r0 = node[kNodeIndirect[f][0]];
                    const V3& r1 = node[kNodeIndirect[f][1]];
                    const V3& r2 = node[kNodeIndirect[f][2]];
                    double A = ((r1.y - r0.y) * (r2.z - r0.z)) - ((r1.z - r0.z) * (r2.y - r0.y));
                    double B = ((r1.z - r0.z) * (r2.x - r0.x)) - ((r1.x - r0.x) * (r2.z - r0.z));
                    double C = ((r1.x - r0.x) * (r2.y - r0.y)) - ((r1.y - r0.y) * (r2.x - r0.x));
                    double D = -1.0 * (A * r0.x + B * r0.y + C * r0.z);
                    double magnitude = std::sqrt(A * A + B * B + C * C);
                    if (magnitude == 0.0) {
                        A = 1.0;
                        magnitude = 1.0;
                    }
                    double inv = 1.0 / magnitude;
                    A *= inv;
                    B *= inv;
                    C *= inv;
                    D *= inv;
                    // The device tracks with the facets of one face sharing a
                    // projection axis: the face normal must be axis-aligned
                    // (two components exactly zero) with the outward sign.
                    {
                        const int ax = f / 8;
                        const double comp[3] = {A, B, C};
                        const bool plus = (f % 8) < 4;
                        bool ok = plus ? comp[ax] > 0.5 : comp[ax] < -0.5;
                        for (int q = 0; q < 3; q++)
                            if (q != ax && comp[q] != 0.0) ok = false;
                        if (!ok) throw std::runtime_error("facet plane is not axis-aligned");
                    }
                    double* p = &T.planes[(c * 24 + f) * 4];
                    p[0] = A;
                    p[1] = B;
                    p[2] = C;
                    p[3] = D;
                }

                // MCT_Cell_Position_3D_G: sum, then times 1/num_points
                V3 pos = {0.0, 0.0, 0.0};
                for (int n = 0; n < 14; n++) {
                    pos.x += node[n].x;
                    pos.y += node[n].y;
                    pos.z += node[n].z;
                }
                double one_over = 1.0 / ((double)14);
                pos.x *= one_over;
                pos.y *= one_over;
                pos.z *= one_over;
                T.cellCenter[c * 3 + 0] = pos.x;
                T.cellCenter[c * 3 + 1] = pos.y;
                T.cellCenter[c * 3 + 2] = pos.z;

                // cellVolume (findCellCenter divides by num_points)
                V3 cc = {0.0, 0.0, 0.0};
                for (int n = 0; n < 14; n++) {
                    cc.x += node[n].x;
                    cc.y += node[n].y;
                    cc.z += node[n].z;
                }
                cc.x /= 14.0;
                cc.y /= 14.0;
                cc.z /= 14.0;
                double volume = 0;
                for (int f = 0; f < 24; f++) {
                    const V3& p0 = node[kNodeIndirect[f][0]];
                    const V3& p1 = node[kNodeIndirect[f][1]];
                    const V3& p2 = node[kNodeIndirect[f][2]];
                    V3 aa = {p0.x - cc.x, p0.y - cc.y, p0.z - cc.z};
                    V3 bb = {p1.x - cc.x, p1.y - cc.y, p1.z - cc.z};
                    V3 c3 = {p2.x - cc.x, p2.y - cc.y, p2.z - cc.z};
                    V3 cr = {bb.y * c3.z - bb.z * c3.y, bb.z * c3.x - bb.x * c3.z, bb.x * c3.y - bb.y * c3.x};
                    volume += std::abs(aa.x * cr.x + aa.y * cr.y + aa.z * cr.z);
                }
                volume /= 6.0;
                T.cellVolume[c] = volume;

                // whichCell(findCellCenter) must be this cell (one domain)
                int ix = (int)(cc.x / T.dx), iy = (int)(cc.y / T.dy), iz = (int)(cc.z / T.dz);
                if (ix != i || iy != j || iz != k) throw std::runtime_error("cell id mismatch");

                // material: last geometry containing the cell position wins
                int m = uncovered;
                for (size_t gi = 0; gi < params.geometryParams.size(); gi++)
                    if (isInside(params.geometryParams[gi], pos)) m = geomMat[gi];
                if (m < 0) throw std::runtime_error("cell centre not covered by a geometry with a known material");
                T.cellMaterial[c] = m;
                cellWeight[c] = volume * T.materials[m].sourceRate * sp.dt;
            }
    };
    const unsigned hw = std::thread::hardware_concurrency();
    const int nThreads = (int)std::min<long long>({(long long)(hw ? hw : 1), 8LL, (long long)sp.nz, nCells / 2048 + 1});
    if (nThreads <= 1) {
        buildSlab(0, sp.nz);
    } else {
        std::vector<std::thread> pool;
        std::vector<std::exception_ptr> errs(nThreads);
        for (int t = 0; t < nThreads; t++) {
            const int k0 = (int)((long long)sp.nz * t / nThreads), k1 = (int)((long long)sp.nz * (t + 1) / nThreads);
            pool.emplace_back([&, t, k0, k1] {
                try {
                    buildSlab(k0, k1);
                } catch (...) {
                    errs[t] = std::current_exception();
                }
            });
        }
        for (std::thread& th : pool) th.join();
        for (const std::exception_ptr& e : errs)
            if (e) std::rethrow_exception(e);
    }

    // ---- source weights (MC_SourceNow.cc:45-66), cell order = gid order ----
    double totalWeight = 0;
    for (long long c = 0; c < nCells; c++) totalWeight += cellWeight[c];
    double sourceFraction = 0.1;
    T.sourceWeight = totalWeight / (sourceFraction * sp.nParticles);
    uint64_t acc = 0;
    for (long long c = 0; c < nCells; c++) {
        double f = cellWeight[c] / T.sourceWeight;
        long long n = 0;
        if (f == f && f >= 1.0 && f < 2147483648.0) n = (long long)(int)f;
        T.cellSource[c] = (uint32_t)n;
        T.cellSourceOffset[c] = acc;
        acc += (uint64_t)n;
    }
    T.cellSourceOffset[nCells] = acc;
    T.sourcePerCycle = acc;
    return T;
}

}  // namespace qs
