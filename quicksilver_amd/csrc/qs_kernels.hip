// gfx950 kernels of the per-cycle particle-tracking path.
//
// Arithmetic is FP64 throughout and compiled with -ffp-contract=off so every
// expression rounds exactly like the reference's non-FMA x86-64 build; the
// order of operations inside each expression follows the reference line
// cited next to it.  Integer RNG streams are bit-exact.  log/sin/cos come
// from the device math library (within 2 ulp of the host's glibc, checked
// by the math probe tests); sqrt and division are correctly rounded.
//
// Kernels (k_batch_* are the same over every problem of a batch)
//   k_cycle_setup  carried count, population factor, clears main.cc:96-121
//   k_source       one thread per source particle         MC_SourceNow.cc:68-151
//                  (optionally beside the previous cycle's flux reset)
//   k_pop_count,   population control: group counts, their   PopulationControl.cc:71-171
//   k_pop_scan,    scan into output bases, ordered writes
//   k_population
//   k_track        persistent waves over a work queue      CycleTracking.cc:15-119
//                  (also accumulates the cycle's flux sum, Tallies.cc:146-171)
//   k_zero_flux    flux reset                              Tallies.cc:95
//   k_fluence      CORAL2 fluence                          Tallies.cc:100-120
//   k_sort_*       counting sort of a queue by (cell, energy band)
//   k_cycle_tail   end of a cycle's first pass: hold check, counters to the
//                  host, next set-up (cycles enqueued ahead)
// Every kernel of the one-problem cycle returns at once while CycleCtl::hold
// is set (the host then finishes the held cycle itself).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "qs_device.h"

namespace qs {
namespace {

constexpr double kHuge = 1.0e+75;   // PhysicalConstants::_hugeDouble
constexpr double kSmall = 1.0e-10;  // _smallDouble
constexpr double kTiny = 1.0e-13;   // _tinyDouble
constexpr double kRestMass = 9.395656981095e+2;
constexpr double kLight = 2.99792458e+10;
constexpr double kPi = 3.1415926535897932;
constexpr uint32_t kMaxSegmentsPerHistory = 1u << 26;

enum { kScatter = 1, kAbsorption = 2, kFission = 3 };
enum : uint32_t {
    kEvCollision = 1u, kEvScatter = 2u, kEvAbsorb = 4u, kEvFission = 8u, kEvEscape = 16u, kEvCensus = 32u
};
enum { kEscape = 1, kReflect = 2 };

// ---------------------------------------------------------------- RNG ----
// 64-bit LCG (MC_RNG_State.hh:23-32).
constexpr uint64_t kLcgMul = 2862933555777941757ULL, kLcgAdd = 3037000493ULL;
__device__ __forceinline__ double rng_sample(uint64_t& s) {
    s = kLcgMul * s + kLcgAdd;
    return 5.4210108624275222e-20 * (double)s;
}

// n steps of the LCG as one multiply-add: s_n = M s + C with M = mul^n and
// C = add (mul^(n-1) + ... + mul + 1), both mod 2^64 (the standard jump-ahead
// of a linear congruential generator).
struct LcgJump {
    uint64_t mul, add;
};

constexpr LcgJump lcgJump(int n) {
    LcgJump j{1, 0};
    for (int i = 0; i < n; i++) {
        j.mul = j.mul * kLcgMul;
        j.add = j.add * kLcgMul + kLcgAdd;
    }
    return j;
}

// The seed after m more (energy, angle) pairs, m = 1..3: 2m draws skipped.
__device__ __forceinline__ uint64_t rng_skip_pairs(uint64_t s, int m) {
    constexpr LcgJump j1 = lcgJump(2), j2 = lcgJump(4), j3 = lcgJump(6);
    const uint64_t mul = m == 1 ? j1.mul : (m == 2 ? j2.mul : j3.mul);
    const uint64_t add = m == 1 ? j1.add : (m == 2 ? j2.add : j3.add);
    return mul * s + add;
}

// Numerical Recipes psdes hash with two rounds (MC_RNG_State.cc:27-98).
__device__ __forceinline__ uint64_t hash_state(uint64_t in) {
    uint32_t lword = (uint32_t)(in >> 32);
    uint32_t irword = (uint32_t)(in & 0xffffffffu);
    const uint32_t c1[2] = {0xbaa96887u, 0x1e17d32cu};
    const uint32_t c2[2] = {0x4b0f3b58u, 0xe874f0c3u};
#pragma unroll
    for (int i = 0; i < 2; i++) {
        uint32_t iswap = irword;
        uint32_t ia = iswap ^ c1[i];
        uint32_t lo = ia & 0xffffu;
        uint32_t hi = ia >> 16;
        uint32_t ib = lo * lo + ~(hi * hi);
        ia = (ib >> 16) | ((ib & 0xffffu) << 16);
        irword = lword ^ ((ia ^ c2[i]) + lo * hi);
        lword = iswap;
    }
    return ((uint64_t)lword << 32) | (uint64_t)irword;
}

// rngSpawn_Random_Number_Seed (MC_RNG_State.cc:107-113)
__device__ __forceinline__ uint64_t rng_spawn(uint64_t& parent) {
    uint64_t child = hash_state(parent);
    rng_sample(parent);
    return child;
}

// NuclearData::getEnergyGroup (NuclearData.cc:208-228).  The reference
// returns G for E > eMax, one past the last table row; tables here have G
// rows, so that case is clamped to G-1.
// eLo and eHi are e[0] and e[G], held in registers by callers that search
// often.
__device__ __forceinline__ int energy_group(const double* e, int G, double E, double eLo, double eHi,
                                            float log2Lo, float perLog2) {
    if (E <= eLo) return 0;
    if (E > eHi) return G - 1;
    if (E != E) return G - 1;  // what the bisection returns for NaN
    // The bisection's answer is the largest g <= G-1 with e[g] <= E.  The
    // bounds are log-spaced, so a single-precision log2 puts the guess within
    // one group of it; the table's own bounds then decide exactly (the two
    // loops run once at most, except for a table that is not log-spaced).
    const float t = (__log2f((float)E) - log2Lo) * perLog2;
    int g = (int)fminf(fmaxf(t, 0.0f), (float)(G - 1));
    const double lo = e[g], hi = e[g + 1];
    if (lo > E) {
        g--;
        while (g > 0 && e[g] > E) g--;
    } else if (hi <= E && g < G - 1) {
        g++;
        while (g < G - 1 && e[g + 1] <= E) g++;
    }
    return g;
}

// --------------------------------------------------------- geometry ----
// The per-facet node codes (kFacetCodes below) in LDS, for the lookups
// indexed per lane (a facet chosen at run time): an LDS read instead of a
// memory load on the tracking loop's critical path.  Every kernel that
// tracks fills it (load_facet_codes, then a barrier) before its first search.
__shared__ uint32_t s_facetCode[24];

struct Cell9 {
    double X0, X1, Xh, Y0, Y1, Yh, Z0, Z1, Zh;
};

// Cell id i + nx*(j + ny*k) (the reference's global cell index with one
// domain) to (i, j, k).  The quotients come from a double product with the
// host's reciprocal, within 2^-21 of the true quotient for ids under 2^31,
// so the truncated value is off by at most one either way; the remainder
// test corrects it.  Integer division has no instruction here.
__device__ __forceinline__ void decode_cell(const DevProblem& P, int cell, int& i, int& j, int& k) {
    const int nxy = P.nx * P.ny;
    int kq = (int)((double)cell * P.rNxy);
    int r = cell - kq * nxy;
    if (r < 0) {
        kq--;
        r += nxy;
    } else if (r >= nxy) {
        kq++;
        r -= nxy;
    }
    int jq = (int)((double)r * P.rNx);
    int iq = r - jq * P.nx;
    if (iq < 0) {
        jq--;
        iq += P.nx;
    } else if (iq >= P.nx) {
        jq++;
        iq -= P.nx;
    }
    i = iq;
    j = jq;
    k = kq;
}

__device__ __forceinline__ Cell9 cell_nodes(const DevProblem& P, int i, int j, int k) {
    Cell9 c;
    c.X0 = (double)i * P.dx;
    c.X1 = (double)(i + 1) * P.dx;
    c.Xh = c.X0 + P.hdx;
    c.Y0 = (double)j * P.dy;
    c.Y1 = (double)(j + 1) * P.dy;
    c.Yh = c.Y0 + P.hdy;
    c.Z0 = (double)k * P.dz;
    c.Z1 = (double)(k + 1) * P.dz;
    c.Zh = c.Z0 + P.hdz;
    return c;
}

__device__ __forceinline__ double sel3(uint32_t code, double v0, double v1, double vh) {
    return code == 0 ? v0 : (code == 1 ? v1 : vh);
}

struct Node3 {
    double x, y, z;
};

__device__ __forceinline__ Node3 facet_node(const Cell9& c, uint32_t code, int m) {
    uint32_t b = code >> (6 * m);
    Node3 n;
    n.x = sel3(b & 3u, c.X0, c.X1, c.Xh);
    n.y = sel3((b >> 2) & 3u, c.Y0, c.Y1, c.Yh);
    n.z = sel3((b >> 4) & 3u, c.Z0, c.Z1, c.Zh);
    return n;
}

// The plane of a facet on a face of axis ax (0 x, 1 y, 2 z) from its nodes
// r0, r1, r2, computed as MC_General_Plane does (MC_Facet_Geometry.hh:18-39)
// operation by operation, reduced to what tracking reads: the nonzero normal
// component N and D.  The facets are axis-aligned (the host checks every
// plane when it builds the tables), so the two other components are signed
// zeros: A*A + B*B + C*C is the nonzero component squared, exactly, and
// A*x0 + B*y0 + C*z0 is the nonzero term whenever that term is nonzero.
// When the term is zero only the sign of D's zero can differ, and a zero
// plane sum is recomputed from the stored four-term plane
// (facet_plane_distance, plane_sum).  No memory is read: the nodes come
// from the cell's nine coordinates.
// 1.0 / sqrt(c * c), the normalisation of a plane whose normal has one
// nonzero component c.  In binary floating point with rounding to nearest
// the correctly rounded square root of fl(c * c) is |c| whenever c * c is a
// normal number, so the root is taken only outside that range.
__device__ __forceinline__ double inv_norm(double c) {
    const double cc = c * c;
    double r = fabs(c);
    if (!(cc >= 0x1p-1000 && cc <= 0x1p1000)) r = sqrt(cc);
    return 1.0 / r;
}

__device__ __forceinline__ void plane_nd(int ax, const Node3& r0, const Node3& r1, const Node3& r2, double& N,
                                         double& D) {
    // (p, q) = the axes after ax in cyclic order: the nonzero component is
    // (r1.p - r0.p)(r2.q - r0.q) - (r1.q - r0.q)(r2.p - r0.p)
    const double p0 = ax == 0 ? r0.y : (ax == 1 ? r0.z : r0.x), q0 = ax == 0 ? r0.z : (ax == 1 ? r0.x : r0.y);
    const double p1 = ax == 0 ? r1.y : (ax == 1 ? r1.z : r1.x), q1 = ax == 0 ? r1.z : (ax == 1 ? r1.x : r1.y);
    const double p2 = ax == 0 ? r2.y : (ax == 1 ? r2.z : r2.x), q2 = ax == 0 ? r2.z : (ax == 1 ? r2.x : r2.y);
    const double x0 = ax == 0 ? r0.x : (ax == 1 ? r0.y : r0.z);
    const double c = ((p1 - p0) * (q2 - q0)) - ((q1 - q0) * (p2 - p0));
    const double d = -1.0 * (c * x0);
    const double inv = 1.0 / sqrt(c * c);
    N = c * inv;
    D = d * inv;
}

// Cell-local node triples of the 24 facets (MC_Domain.cc:42-47) resolved to
// per-axis value codes of the corner tuples (GlobalFccGrid.cc:47-69):
// 0 -> i*dx, 1 -> (i+1)*dx, 2 -> i*dx + dx/2.  kFacet.code[f][node][axis].
struct FacetTable {
    uint8_t code[24][3][3];
};

constexpr FacetTable makeFacetTable() {
    const int corner[14][4] = {{0, 0, 0, 0}, {1, 0, 0, 0}, {0, 1, 0, 0}, {1, 1, 0, 0}, {0, 0, 1, 0},
                               {1, 0, 1, 0}, {0, 1, 1, 0}, {1, 1, 1, 0}, {1, 0, 0, 1}, {0, 0, 0, 1},
                               {0, 1, 0, 2}, {0, 0, 0, 2}, {0, 0, 1, 3}, {0, 0, 0, 3}};
    const int tri[24][3] = {{1, 3, 8},  {3, 7, 8},  {7, 5, 8},  {5, 1, 8},  {0, 4, 9},  {4, 6, 9},
                            {6, 2, 9},  {2, 0, 9},  {3, 2, 10}, {2, 6, 10}, {6, 7, 10}, {7, 3, 10},
                            {0, 1, 11}, {1, 5, 11}, {5, 4, 11}, {4, 0, 11}, {4, 5, 12}, {5, 7, 12},
                            {7, 6, 12}, {6, 4, 12}, {0, 2, 13}, {2, 3, 13}, {3, 1, 13}, {1, 0, 13}};
    FacetTable t{};
    for (int f = 0; f < 24; f++)
        for (int m = 0; m < 3; m++) {
            const int* o = corner[tri[f][m]];
            t.code[f][m][0] = (uint8_t)((o[3] == 2 || o[3] == 3) ? 2 : o[0]);
            t.code[f][m][1] = (uint8_t)((o[3] == 1 || o[3] == 3) ? 2 : o[1]);
            t.code[f][m][2] = (uint8_t)((o[3] == 1 || o[3] == 2) ? 2 : o[2]);
        }
    return t;
}

constexpr FacetTable kFacet = makeFacetTable();

// Per-facet codes.  Bits 0-17: for each of the facet's 3 nodes (node m at
// bit 6*m), 2 bits per axis selecting {0: i*dx, 1: (i+1)*dx, 2: i*dx + dx/2}.
// Bits 18-23 give the facet's plane directly (facet_nd): every facet is two
// corners that share an edge of its face, then the face centre.  In
// plane_nd's (p, q) axes, when the corners share p (p1 == p0), the first
// product (p1 - p0)(q2 - q0) is a signed zero and the nonzero component is
// c = -(q1 - q0)(Ph - p0); when they share q, the second product is a signed
// zero and c = (p1 - p0)(Qh - q0).  Either way c = +-(E * H), with E the
// cell's extent along the edge's axis (q1 - q0 or p1 - p0 is +-E exactly)
// and H the face centre's offset from the edge across it, so:
//   bits 18-19 the edge's axis (E = X1 - X0, Y1 - Y0 or Z1 - Z0),
//   bits 20-21 the axis across the edge, bit 22 the edge's end of that axis
//   (H = Xh - X0 or Xh - X1, ...), bit 23 set when c = -(E * H).
// Bits 24-26: the sign bits of the plane's x, y and z components where they
// are zero (the two other than the face axis).  MC_General_Plane
// (MC_Facet_Geometry.hh:18-39) forms them from products with a difference
// of equal face coordinates, +0, so each is a signed zero whose sign follows
// from the signs of the node differences alone: evaluated here on nodes
// 1 < centre 2 < 3 along every axis, the same signs as on any cell.
// They are compile-time constants, so a facet known at compile time costs
// no register; per-lane lookups read the LDS copy.
struct FacetCodes {
    uint32_t c[24];
};

constexpr FacetCodes makeFacetCodes() {
    FacetCodes k{};
    for (int f = 0; f < 24; f++) {
        for (int m = 0; m < 3; m++)
            k.c[f] |= (uint32_t)(kFacet.code[f][m][0] | (kFacet.code[f][m][1] << 2) | (kFacet.code[f][m][2] << 4))
                      << (6 * m);
        const int ax = f >> 3, p = (ax + 1) % 3, q = (ax + 2) % 3;
        const int p0 = kFacet.code[f][0][p], p1 = kFacet.code[f][1][p];
        const int q0 = kFacet.code[f][0][q], q1 = kFacet.code[f][1][q];
        uint32_t eAx, hAx, hEnd, neg;
        if (p0 == p1) {  // c = -(q1 - q0)(Ph - p0)
            eAx = (uint32_t)q;
            hAx = (uint32_t)p;
            hEnd = (uint32_t)p0;
            neg = q1 > q0 ? 1u : 0u;
        } else {  // c = (p1 - p0)(Qh - q0)
            eAx = (uint32_t)p;
            hAx = (uint32_t)q;
            hEnd = (uint32_t)q0;
            neg = p1 < p0 ? 1u : 0u;
        }
        k.c[f] |= (eAx << 18) | (hAx << 20) | (hEnd << 22) | (neg << 23);
        // MC_General_Plane's A, B, C on representative nodes
        double r[3][3];
        for (int m = 0; m < 3; m++)
            for (int a = 0; a < 3; a++) {
                const int cd = kFacet.code[f][m][a];
                r[m][a] = cd == 0 ? 1.0 : (cd == 1 ? 3.0 : 2.0);
            }
        const double comp[3] = {
            ((r[1][1] - r[0][1]) * (r[2][2] - r[0][2])) - ((r[1][2] - r[0][2]) * (r[2][1] - r[0][1])),
            ((r[1][2] - r[0][2]) * (r[2][0] - r[0][0])) - ((r[1][0] - r[0][0]) * (r[2][2] - r[0][2])),
            ((r[1][0] - r[0][0]) * (r[2][1] - r[0][1])) - ((r[1][1] - r[0][1]) * (r[2][0] - r[0][0]))};
        for (int a = 0; a < 3; a++)
            if (a != ax && (__builtin_bit_cast(uint64_t, comp[a]) >> 63)) k.c[f] |= 1u << (24 + a);
    }
    return k;
}

constexpr FacetCodes kFacetCodes = makeFacetCodes();

// The table's premises, checked at compile time: on every facet the face
// axis holds one corner value (0 or 1) for all three nodes (on the + face
// for facets 8*ax + 0..3, the - face for 8*ax + 4..7), the third node is the
// face centre (code 2 on both other axes), and the two corners differ on
// exactly one of the other axes, by one edge (codes 0 and 1).
constexpr bool facetTableHolds() {
    for (int f = 0; f < 24; f++) {
        const int ax = f >> 3, p = (ax + 1) % 3, q = (ax + 2) % 3;
        const int face = (f & 4) ? 0 : 1;
        for (int m = 0; m < 3; m++)
            if (kFacet.code[f][m][ax] != face) return false;
        if (kFacet.code[f][2][p] != 2 || kFacet.code[f][2][q] != 2) return false;
        const int p0 = kFacet.code[f][0][p], p1 = kFacet.code[f][1][p];
        const int q0 = kFacet.code[f][0][q], q1 = kFacet.code[f][1][q];
        if (p0 > 1 || p1 > 1 || q0 > 1 || q1 > 1) return false;
        if ((p0 == p1) == (q0 == q1)) return false;
    }
    return true;
}
static_assert(facetTableHolds(), "facet node table");

__device__ __forceinline__ void load_facet_codes() {
    if (threadIdx.x < 24) s_facetCode[threadIdx.x] = kFacetCodes.c[threadIdx.x];
}

// N and D of facet f of the cell: plane_nd's arithmetic on the facet's own
// nodes, with c formed as +-(E * H) (see kFacetCodes; the same products and
// differences of the same node values, so the same bits), for a facet
// chosen per lane at run time.
__device__ __forceinline__ double in_vgpr(double v) {
    // the value as a register operand: keeps per-lane selects among a
    // struct's fields from becoming an indexed load of a stack copy
    asm("" : "+v"(v));
    return v;
}

__device__ __forceinline__ void facet_nd(const Cell9& c9, int f, double& N, double& D) {
    const uint32_t code = s_facetCode[f];
    const int ax = f >> 3;
    const uint32_t eAx = (code >> 18) & 3u, hAx = (code >> 20) & 3u;
    const bool hEnd = (code >> 22) & 1u, neg = (code >> 23) & 1u;
    const double X0 = in_vgpr(c9.X0), X1 = in_vgpr(c9.X1), Xh = in_vgpr(c9.Xh);
    const double Y0 = in_vgpr(c9.Y0), Y1 = in_vgpr(c9.Y1), Yh = in_vgpr(c9.Yh);
    const double Z0 = in_vgpr(c9.Z0), Z1 = in_vgpr(c9.Z1), Zh = in_vgpr(c9.Zh);
    const double E = eAx == 0 ? X1 - X0 : (eAx == 1 ? Y1 - Y0 : Z1 - Z0);
    const double Hc = hAx == 0 ? Xh : (hAx == 1 ? Yh : Zh);
    const double He = hAx == 0 ? (hEnd ? X1 : X0) : (hAx == 1 ? (hEnd ? Y1 : Y0) : (hEnd ? Z1 : Z0));
    double c = E * (Hc - He);
    if (neg) c = -c;
    const bool minus = (f & 4) != 0;
    const double x0 = ax == 0 ? (minus ? X0 : X1) : (ax == 1 ? (minus ? Y0 : Y1) : (minus ? Z0 : Z1));
    const double d = -1.0 * (c * x0);
    const double inv = inv_norm(c);
    N = c * inv;
    D = d * inv;
}

// Which triangle of a face lies on which edge of the face.  Face (AX,
// side) holds facets 8*AX + 4*side + T, T = 0..3; each is two corners plus
// the face centre, and the two corners share one edge of the face.  In the
// face's projected (u, v) axes (AX 0 -> (y,z), 1 -> (z,x), 2 -> (x,y)),
// edge 0 is u = U0, 1 u = U1, 2 v = V0, 3 v = V1.  kEdgeTri[AX][side][edge]
// = T.
struct EdgeTable {
    int tri[3][2][4];
};

constexpr EdgeTable makeEdgeTable() {
    EdgeTable e{};
    for (int ax = 0; ax < 3; ax++) {
        const int u = ax == 0 ? 1 : (ax == 1 ? 2 : 0), v = ax == 0 ? 2 : (ax == 1 ? 0 : 1);
        for (int side = 0; side < 2; side++)
            for (int t = 0; t < 4; t++) {
                const int f = 8 * ax + 4 * side + t;
                const int u0 = kFacet.code[f][0][u], u1 = kFacet.code[f][1][u];
                const int v0 = kFacet.code[f][0][v], v1 = kFacet.code[f][1][v];
                int edge = -1;
                if (u0 == 0 && u1 == 0) edge = 0;
                if (u0 == 1 && u1 == 1) edge = 1;
                if (v0 == 0 && v1 == 0) edge = 2;
                if (v0 == 1 && v1 == 1) edge = 3;
                e.tri[ax][side][edge] = t;
            }
    }
    return e;
}

constexpr EdgeTable kEdgeTri = makeEdgeTable();

// kEdgeTri[AX] packed 2 bits per entry at bit 2 * (4 * side + edge), so a
// lane's triangle comes from a shift of a constant, not a table load.
template <int AX>
constexpr uint32_t edgeTriPacked() {
    uint32_t v = 0;
    for (int side = 0; side < 2; side++)
        for (int e = 0; e < 4; e++) v |= (uint32_t)kEdgeTri.tri[AX][side][e] << (2 * (4 * side + e));
    return v;
}

// Value of one node coordinate for a compile-time code.
template <int AXIS, int CODE>
__device__ __forceinline__ double node_value(const Cell9& c) {
    if (AXIS == 0) return CODE == 0 ? c.X0 : (CODE == 1 ? c.X1 : c.Xh);
    if (AXIS == 1) return CODE == 0 ? c.Y0 : (CODE == 1 ? c.Y1 : c.Yh);
    return CODE == 0 ? c.Z0 : (CODE == 1 ? c.Z1 : c.Zh);
}

// Node coordinate of facet FP (lane leaves through the + face) or FM (the -
// face of the same axis).
template <int FP, int FM, int M, int AXIS>
__device__ __forceinline__ double node_coord(const Cell9& c, bool plus) {
    double vp = node_value<AXIS, kFacet.code[FP][M][AXIS]>(c);
    double vm = node_value<AXIS, kFacet.code[FM][M][AXIS]>(c);
    return plus ? vp : vm;
}

// MCT_Nearest_Facet_3D_G_Distance_To_Segment (MCT.cc:280-395) for a facet of
// a face normal to axis AX.  On this mesh every such plane has its other two
// components exactly 0 and |component AX| > 0.5 (checked on the host when
// the problem is built), so the reference takes the |A|>0.5 branch for x
// faces, |B|>0.5 for y faces and |C|>0.5 for z faces; the branch, its
// bounding-box axes and its projected (u,v) axes are therefore known here at
// compile time: AX 0 -> (y,z), AX 1 -> (z,x), AX 2 -> (x,y).
template <int AX>
__device__ __forceinline__ double distance_to_facet(double planeTol, double dot, double A, double B, double C,
                                                    double D, double p0u, double p0v, double p1u, double p1v,
                                                    double p2u, double p2v, double x, double y, double z,
                                                    double a, double b, double g) {
    const double tol = 1e-9;
    double numerator = -1.0 * (A * x + B * y + C * z + D);
    if (numerator < 0.0 && numerator * numerator > planeTol) return kHuge;
    double distance = numerator / dot;
    double ix = x + distance * a;
    double iy = y + distance * b;
    double iz = z + distance * g;
    const double iu = AX == 0 ? iy : (AX == 1 ? iz : ix);
    const double iv = AX == 0 ? iz : (AX == 1 ? ix : iy);
    double iuP = iu + tol, iuM = iu - tol, ivP = iv + tol, ivM = iv - tol;
    if (p0u > iuP && p1u > iuP && p2u > iuP) return kHuge;
    if (p0u < iuM && p1u < iuM && p2u < iuM) return kHuge;
    if (p0v > ivP && p1v > ivP && p2v > ivP) return kHuge;
    if (p0v < ivM && p1v < ivM && p2v < ivM) return kHuge;
    double cross1 = (p1u - p0u) * (iv - p0v) - (p1v - p0v) * (iu - p0u);
    double cross2 = (p2u - p1u) * (iv - p1v) - (p2v - p1v) * (iu - p1u);
    double cross0 = (p0u - p2u) * (iv - p2v) - (p0v - p2v) * (iu - p2u);
    double s = cross0 + cross1 + cross2;
    double ctol = 1e-9 * (s < 0 ? -s : s);
    if ((cross0 > -ctol && cross1 > -ctol && cross2 > -ctol) ||
        (cross0 < ctol && cross1 < ctol && cross2 < ctol))
        return distance;
    return kHuge;
}

struct Nearest {
    double d;
    int f;
};

// Running state of MCT_Nearest_Facet_Find_Nearest (MCT.cc:436-479): the
// reference scans facets 0..23 with "d > 0 && d <= best" (ties go to the
// later facet) and keeps the largest non-positive d (ties to the earlier).
// Facets with distance kHuge (facing away, or rejected) are tracked as the
// largest such index, so the scan can skip them and still end identically.
struct FindNearest {
    double pos;  // smallest positive distance other than kHuge
    int posF;
    int hugeF;   // largest facet index whose distance is kHuge
    double neg;
    int negF;
    __device__ __forceinline__ void add(double d, int f) {
        if (d == kHuge) {
            hugeF = f > hugeF ? f : hugeF;
        } else if (d > 0.0) {
            if (d <= pos) {
                pos = d;
                posF = f;
            }
        } else if (d > neg) {
            neg = d;
            negF = f;
        }
    }
    __device__ __forceinline__ Nearest result() const {
        Nearest n;
        n.d = 1e80;
        n.f = 0;
        if (posF >= 0 && !(hugeF >= 0 && pos > kHuge)) {
            n.d = pos;
            n.f = posF;
        } else if (hugeF >= 0) {
            n.d = kHuge;
            n.f = hugeF;
        } else if (posF >= 0) {
            n.d = pos;
            n.f = posF;
        }
        if (n.d == kHuge && neg != -kHuge) {
            n.d = neg;
            n.f = negF;
        }
        return n;
    }
};

// The four facets of the face of axis AX the lane leaves through.
// The same test for a plane whose other two normal components are zero
// (B = C = 0 for an x face, ...): its dot product is N*dir[AX] and its
// numerator -(N*pos[AX] + D).  Dropping the zero terms changes no nonzero
// value: B*y and C*z are signed zeros, and adding a signed zero to a
// nonzero partial sum leaves it unchanged.  Only a zero sum can differ, in
// its sign (-0 + -0 is -0, but the reference's sum of four terms may hold a
// +0); then the full plane row pf gives the reference's own four-term sum,
// so even the sign of a zero distance is the reference's.
template <int AX>
__device__ __forceinline__ double plane_sum(double N, double D, const double* pf, double x, double y, double z) {
    const double xa = AX == 0 ? x : (AX == 1 ? y : z);
    double s = N * xa + D;
    if (s == 0.0) s = pf[0] * x + pf[1] * y + pf[2] * z + pf[3];
    return s;
}

template <int AX>
__device__ __forceinline__ double distance_to_face_facet(double planeTol, double dot, double N, double D,
                                                         const double* pf, double p0u, double p0v, double p1u,
                                                         double p1v, double p2u, double p2v, double x, double y,
                                                         double z, double a, double b, double g) {
    const double tol = 1e-9;
    double numerator = -1.0 * plane_sum<AX>(N, D, pf, x, y, z);
    if (numerator < 0.0 && numerator * numerator > planeTol) return kHuge;
    double distance = numerator / dot;
    double ix = x + distance * a;
    double iy = y + distance * b;
    double iz = z + distance * g;
    const double iu = AX == 0 ? iy : (AX == 1 ? iz : ix);
    const double iv = AX == 0 ? iz : (AX == 1 ? ix : iy);
    double iuP = iu + tol, iuM = iu - tol, ivP = iv + tol, ivM = iv - tol;
    if (p0u > iuP && p1u > iuP && p2u > iuP) return kHuge;
    if (p0u < iuM && p1u < iuM && p2u < iuM) return kHuge;
    if (p0v > ivP && p1v > ivP && p2v > ivP) return kHuge;
    if (p0v < ivM && p1v < ivM && p2v < ivM) return kHuge;
    double cross1 = (p1u - p0u) * (iv - p0v) - (p1v - p0v) * (iu - p0u);
    double cross2 = (p2u - p1u) * (iv - p1v) - (p2v - p1v) * (iu - p1u);
    double cross0 = (p0u - p2u) * (iv - p2v) - (p0v - p2v) * (iu - p2u);
    double s = cross0 + cross1 + cross2;
    double ctol = 1e-9 * (s < 0 ? -s : s);
    if ((cross0 > -ctol && cross1 > -ctol && cross2 > -ctol) ||
        (cross0 < ctol && cross1 < ctol && cross2 < ctol))
        return distance;
    return kHuge;
}

template <int AX, int T>
__device__ __forceinline__ void visit_facet(FindNearest& fn, const double* pl, const Cell9& c9, double planeTol,
                                            double x, double y, double z, double a, double b, double g, bool plus,
                                            bool minus) {
    constexpr int FP = 8 * AX + T, FM = FP + 4;
    constexpr int U = AX == 0 ? 1 : (AX == 1 ? 2 : 0);
    constexpr int V = AX == 0 ? 2 : (AX == 1 ? 0 : 1);
    const int f = plus ? FP : FM;
    if (!(plus || minus)) {
        fn.add(kHuge, FP);
        fn.add(kHuge, FM);
        return;
    }
    const uint32_t code = plus ? kFacetCodes.c[FP] : kFacetCodes.c[FM];
    double N, D;
    plane_nd(AX, facet_node(c9, code, 0), facet_node(c9, code, 1), facet_node(c9, code, 2), N, D);
    const double dot = N * (AX == 0 ? a : (AX == 1 ? b : g));
    double d = kHuge;
    if (dot > 0.0) {
        const double p0u = node_coord<FP, FM, 0, U>(c9, plus), p0v = node_coord<FP, FM, 0, V>(c9, plus);
        const double p1u = node_coord<FP, FM, 1, U>(c9, plus), p1v = node_coord<FP, FM, 1, V>(c9, plus);
        const double p2u = node_coord<FP, FM, 2, U>(c9, plus), p2v = node_coord<FP, FM, 2, V>(c9, plus);
        d = distance_to_face_facet<AX>(planeTol, dot, N, D, pl + 4 * f, p0u, p0v, p1u, p1v, p2u, p2v, x, y, z, a,
                                       b, g);
    }
    // the opposite face faces away from the direction: distance kHuge
    fn.add(kHuge, plus ? FM : FP);
    fn.add(d, f);
}

template <int AX>
__device__ __forceinline__ void visit_axis(FindNearest& fn, const double* pl, const Cell9& c9, double planeTol,
                                           double x, double y, double z, double a, double b, double g) {
    const double dAx = AX == 0 ? a : (AX == 1 ? b : g);
    const bool plus = dAx > 0.0, minus = dAx < 0.0;
    visit_facet<AX, 0>(fn, pl, c9, planeTol, x, y, z, a, b, g, plus, minus);
    visit_facet<AX, 1>(fn, pl, c9, planeTol, x, y, z, a, b, g, plus, minus);
    visit_facet<AX, 2>(fn, pl, c9, planeTol, x, y, z, a, b, g, plus, minus);
    visit_facet<AX, 3>(fn, pl, c9, planeTol, x, y, z, a, b, g, plus, minus);
}

// One facet of any face, for the rare retry path: facets are visited in
// the reference's order 0..23 with the same arithmetic as visit_facet.
__device__ __forceinline__ double distance_any_facet(const double* pl, const Cell9& c9, int f, double planeTol,
                                                     double x, double y, double z, double a, double b,
                                                     double g) {
    const double* p = pl + 4 * f;
    const double A = p[0], B = p[1], C = p[2], D = p[3];
    const double dot = A * a + B * b + C * g;
    if (dot <= 0.0) return kHuge;
    const uint32_t code = s_facetCode[f];
    const Node3 n0 = facet_node(c9, code, 0), n1 = facet_node(c9, code, 1), n2 = facet_node(c9, code, 2);
    switch (f >> 3) {
        case 0:
            return distance_to_facet<0>(planeTol, dot, A, B, C, D, n0.y, n0.z, n1.y, n1.z, n2.y, n2.z, x, y, z, a,
                                        b, g);
        case 1:
            return distance_to_facet<1>(planeTol, dot, A, B, C, D, n0.z, n0.x, n1.z, n1.x, n2.z, n2.x, x, y, z, a,
                                        b, g);
        default:
            return distance_to_facet<2>(planeTol, dot, A, B, C, D, n0.x, n0.y, n1.x, n1.y, n2.x, n2.y, x, y, z, a,
                                        b, g);
    }
}

// Probe of one exit face.  The four triangles of a face lie in the face
// plane (their stored planes differ from it only in the last bits), so the
// exact intersection point with any of them is within ~1e-15 of the probe's,
// the intersection with the face plane.  Only a triangle whose region,
// widened by `margin` (far above the reference's 1e-9 tolerances), holds the
// probe point can pass the reference's containment test; every other one
// returns kHuge.  The probe names those candidates (one, or two near a
// diagonal; none when the point is outside the face) or, when it cannot
// decide -- particle on or behind the plane, a near-grazing direction, a
// point near the face centre -- says so.  A lone candidate whose region holds
// the point deeper than margin from every edge of the triangle is "deep":
// the reference's containment test passes for it.
struct FaceProbe {
    int c1, c2;  // candidate facet indices, c1 < c2, or -1
    bool unsure;
    bool deep;
};

// What a lane computes for each face of axis AX that its direction points
// out of: the probe point relative to the face centre, the face's half
// widths and margin, and whether the point lies on the face widened by
// margin.  Only the face that holds the point gets the rest of the probe
// (probe_face), once per lane whatever its axis.
struct FaceLook {
    double du, dv;  // probe point minus the face centre along (u, v), times dot0
    double hu, hv;  // the face's half widths
    double margin;
    double dot0;    // |direction along the axis| > 0
    bool plus;      // the + face of the axis
    bool inside;    // on the face widened by margin
    bool unsure;    // the probe cannot decide (see FaceProbe)
};

template <int AX>
__device__ __forceinline__ FaceLook look_exit_face(const Cell9& c9, double x, double y, double z, double a,
                                                   double b, double g) {
    constexpr int U = AX == 0 ? 1 : (AX == 1 ? 2 : 0);
    constexpr int V = AX == 0 ? 2 : (AX == 1 ? 0 : 1);
    FaceLook r = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, false, false, false};
    const double dAx = AX == 0 ? a : (AX == 1 ? b : g);
    const bool plus = dAx > 0.0, minus = dAx < 0.0;
    if (!(plus || minus)) return r;  // parallel to both faces: dot = 0, all kHuge
    r.plus = plus;
    // The face is the plane coordinate = Xf; the stored planes of its four
    // triangles equal it to the last bits, so the probe needs no plane loads.
    // dot0 > 0 and num0 are the direction and distance along the axis.
    const double xa = AX == 0 ? x : (AX == 1 ? y : z);
    const double Xf = plus ? (AX == 0 ? c9.X1 : (AX == 1 ? c9.Y1 : c9.Z1))
                           : (AX == 0 ? c9.X0 : (AX == 1 ? c9.Y0 : c9.Z0));
    const double dot0 = plus ? dAx : -dAx;
    const double num0 = plus ? Xf - xa : xa - Xf;
    const double U0 = AX == 0 ? c9.Y0 : (AX == 1 ? c9.Z0 : c9.X0);
    const double U1 = AX == 0 ? c9.Y1 : (AX == 1 ? c9.Z1 : c9.X1);
    const double Uh = AX == 0 ? c9.Yh : (AX == 1 ? c9.Zh : c9.Xh);
    const double V0 = AX == 0 ? c9.Z0 : (AX == 1 ? c9.X0 : c9.Y0);
    const double V1 = AX == 0 ? c9.Z1 : (AX == 1 ? c9.X1 : c9.Y1);
    const double Vh = AX == 0 ? c9.Zh : (AX == 1 ? c9.Xh : c9.Yh);
    const double hu = 0.5 * (U1 - U0), hv = 0.5 * (V1 - V0);
    const double margin = fmax(1e-6 * fmin(hu, hv), 1e-7);
    // bound on how far num0 and the four triangles' numerators (scaled to
    // this axis) can differ: last-bit differences of the planes and of the
    // subtraction, times the coordinates involved
    const double errN = 8e-16 * (fabs(xa) + fabs(Xf)) + 1e-300;
    if (!(num0 > 1e3 * errN) || !(errN < 1e-3 * margin * dot0)) {
        r.unsure = true;
        return r;
    }
    // probe point relative to the face centre, scaled by dot0 > 0 so that no
    // division is needed: du = (u - Uh) + (num0 / dot0) * dir_u, times dot0
    // (rounding here is ~1e-16 of quantities the margin exceeds by 1e6)
    r.du = ((U == 0 ? x : (U == 1 ? y : z)) - Uh) * dot0 + num0 * (U == 0 ? a : (U == 1 ? b : g));
    r.dv = ((V == 0 ? x : (V == 1 ? y : z)) - Vh) * dot0 + num0 * (V == 0 ? a : (V == 1 ? b : g));
    r.hu = hu;
    r.hv = hv;
    r.margin = margin;
    r.dot0 = dot0;
    // outside the face widened by margin, the reference's containment test
    // fails for all four of its triangles (its tolerances are 1e-9)
    r.inside = fabs(r.du) <= (hu + margin) * dot0 && fabs(r.dv) <= (hv + margin) * dot0;
    return r;
}

// The rest of the probe for the face of axis ax that holds the point.  The
// face's diagonals split it into the four triangles (the one on the u = U1
// edge is the wedge s1 > 0, s2 > 0 below, and so on).  A triangle across a
// diagonal lies beyond that diagonal's line, so it is a candidate only if the
// point is within `margin` of the line (|s| <= margin * (hu + hv) >= margin *
// |diagonal direction|).  Within margin of both diagonals -- the centre,
// where all four meet -- the probe decides nothing.  Lengths are scaled by
// dot0 like du, dv.
__device__ __forceinline__ FaceProbe probe_face(int ax, const FaceLook& l) {
    FaceProbe r = {-1, -1, false, false};
    const double s1 = l.du * l.hv - l.dv * l.hu, s2 = l.du * l.hv + l.dv * l.hu;
    const double thr = l.margin * (l.hu + l.hv) * l.dot0;
    const bool near1 = fabs(s1) <= thr, near2 = fabs(s2) <= thr;
    if (near1 && near2) {
        r.unsure = true;
        return r;
    }
    const int side = l.plus ? 0 : 1;
    const int base = 8 * ax + 4 * side;
    // edge numbering of the edge table: 0 u = U0, 1 u = U1, 2 v = V0, 3 v = V1
    const int wedge = s1 > 0.0 ? (s2 > 0.0 ? 1 : 2) : (s2 > 0.0 ? 3 : 0);
    int other = -1;
    if (near1) other = s2 > 0.0 ? (wedge == 1 ? 3 : 1) : (wedge == 2 ? 0 : 2);
    else if (near2) other = s1 > 0.0 ? (wedge == 1 ? 2 : 1) : (wedge == 3 ? 0 : 3);
    const uint32_t kTri = ax == 0 ? edgeTriPacked<0>() : (ax == 1 ? edgeTriPacked<1>() : edgeTriPacked<2>());
    const int t1 = (int)((kTri >> (2 * (4 * side + wedge))) & 3u);
    r.c1 = base + t1;
    // no second candidate: farther than margin from both diagonals; and
    // farther than margin inside the face's own edges
    r.deep = other < 0 && fabs(l.du) <= (l.hu - l.margin) * l.dot0 && fabs(l.dv) <= (l.hv - l.margin) * l.dot0;
    if (other >= 0) {
        const int t2 = (int)((kTri >> (2 * (4 * side + other))) & 3u);
        r.c1 = base + (t1 < t2 ? t1 : t2);
        r.c2 = base + (t1 < t2 ? t2 : t1);
    }
    return r;
}

// The reference's distance to facet f's plane (distance_to_face_facet up
// to its containment test), for a facet chosen per lane at run time: kHuge
// when the direction does not leave through the plane or the particle lies
// behind it beyond the plane tolerance.
__device__ __forceinline__ double facet_plane_distance(int f, const Cell9& c9, const double* pl, double planeTol,
                                                       double x, double y, double z, double a, double b,
                                                       double g) {
    const int ax = f >> 3;
    double N, D;
    facet_nd(c9, f, N, D);
    const double dAx = ax == 0 ? a : (ax == 1 ? b : g);
    const double dot = N * dAx;
    if (!(dot > 0.0)) return kHuge;
    const double xa = ax == 0 ? x : (ax == 1 ? y : z);
    double sum = N * xa + D;
    if (sum == 0.0) {
        // a zero sum: the sign of zero of the reference's four-term sum
        const double* pf = pl + 4 * f;
        sum = pf[0] * x + pf[1] * y + pf[2] * z + pf[3];
    }
    double numerator = -1.0 * sum;
    if (numerator < 0.0 && numerator * numerator > planeTol) return kHuge;
    return numerator / dot;
}

// The reference's exact test (distance_to_face_facet) for a facet chosen
// per lane at run time: the projection axes are selected per lane, so the
// whole wave runs one copy of the test whatever face each lane's candidate
// lies on.
__device__ __forceinline__ double exact_facet_test(int f, const double* pl, const Cell9& c9, double planeTol,
                                                   double x, double y, double z, double a, double b, double g) {
    const int ax = f >> 3;
    const double tol = 1e-9;
    const double distance = facet_plane_distance(f, c9, pl, planeTol, x, y, z, a, b, g);
    if (distance == kHuge) return kHuge;
    double ix = x + distance * a;
    double iy = y + distance * b;
    double iz = z + distance * g;
    const double iu = ax == 0 ? iy : (ax == 1 ? iz : ix);
    const double iv = ax == 0 ? iz : (ax == 1 ? ix : iy);
    const uint32_t code = s_facetCode[f];
    const Node3 n0 = facet_node(c9, code, 0), n1 = facet_node(c9, code, 1), n2 = facet_node(c9, code, 2);
    const double p0u = ax == 0 ? n0.y : (ax == 1 ? n0.z : n0.x), p0v = ax == 0 ? n0.z : (ax == 1 ? n0.x : n0.y);
    const double p1u = ax == 0 ? n1.y : (ax == 1 ? n1.z : n1.x), p1v = ax == 0 ? n1.z : (ax == 1 ? n1.x : n1.y);
    const double p2u = ax == 0 ? n2.y : (ax == 1 ? n2.z : n2.x), p2v = ax == 0 ? n2.z : (ax == 1 ? n2.x : n2.y);
    double iuP = iu + tol, iuM = iu - tol, ivP = iv + tol, ivM = iv - tol;
    if (p0u > iuP && p1u > iuP && p2u > iuP) return kHuge;
    if (p0u < iuM && p1u < iuM && p2u < iuM) return kHuge;
    if (p0v > ivP && p1v > ivP && p2v > ivP) return kHuge;
    if (p0v < ivM && p1v < ivM && p2v < ivM) return kHuge;
    double cross1 = (p1u - p0u) * (iv - p0v) - (p1v - p0v) * (iu - p0u);
    double cross2 = (p2u - p1u) * (iv - p1v) - (p2v - p1v) * (iu - p1u);
    double cross0 = (p0u - p2u) * (iv - p2v) - (p0v - p2v) * (iu - p2u);
    double s = cross0 + cross1 + cross2;
    double ctol = 1e-9 * (s < 0 ? -s : s);
    if ((cross0 > -ctol && cross1 > -ctol && cross2 > -ctol) ||
        (cross0 < ctol && cross1 < ctol && cross2 < ctol))
        return distance;
    return kHuge;
}

// MCT_Nearest_Facet + MCT_Nearest_Facet_3D_G (MCT.cc:87-137, 550-621) with
// the retry of MCT_Nearest_Facet_Find_Nearest (MCT.cc:487-536): when no
// facet bounds the particle it is nudged toward the cell centre and the
// search repeats.  The first search considers only the three faces the
// direction points out of (the other 12 facets have dot <= 0 and distance
// kHuge in the reference too).  A look at each of those faces finds the one
// that holds the exit point; the probe of that face names the one or two
// triangles that can hold it; those candidates get the reference's exact
// test in facet order, and every other facet is kHuge.  A lane whose probe
// cannot decide, or whose point lies on two faces (within margin of an
// edge), runs the exact test on all 12 exit-face facets.  The rare retries
// use a compact loop over all 24 facets so they add no register pressure to
// the common path.
__device__ __forceinline__ Nearest nearest_facet(const DevProblem& P, const double* pl, int cell, const Cell9& c9, double& x, double& y, double& z,
                                                 double a, double b, double g) {
    Nearest nf;
    {
        const double planeTol = 1e-16 * (x * x + y * y + z * z);
        // Each exit face's look, keeping the first face (in facet order)
        // that holds the probe point; a point on two faces at once (within
        // margin of an edge) is left to the full test.
        FaceLook sel = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, false, false, false};
        int selAx = -1, nIn = 0;
        bool unsure = false;
        {
            const FaceLook l = look_exit_face<0>(c9, x, y, z, a, b, g);
            unsure = unsure || l.unsure;
            if (l.inside) {
                nIn++;
                sel = l;
                selAx = 0;
            }
        }
        {
            const FaceLook l = look_exit_face<1>(c9, x, y, z, a, b, g);
            unsure = unsure || l.unsure;
            if (l.inside) {
                nIn++;
                if (selAx < 0) {
                    sel = l;
                    selAx = 1;
                }
            }
        }
        {
            const FaceLook l = look_exit_face<2>(c9, x, y, z, a, b, g);
            unsure = unsure || l.unsure;
            if (l.inside) {
                nIn++;
                if (selAx < 0) {
                    sel = l;
                    selAx = 2;
                }
            }
        }
        FaceProbe fp = {-1, -1, false, false};
        if (!unsure && nIn == 1) fp = probe_face(selAx, sel);
        unsure = unsure || nIn > 1 || fp.unsure;
        // the lane's candidates, at most two, both on that face
        const int n = (fp.c1 >= 0) + (fp.c2 >= 0);
        const int q0 = fp.c1;
        const bool deep = !unsure && n == 1 && fp.deep;
        if (deep) {
            // the only candidate, its exit point deeper inside it than
            // margin: the reference's containment test passes for it, so
            // only the distance is computed.  FindNearest over {q0: d, every
            // other facet kHuge} then gives (d, q0) for -kHuge < d < kHuge
            // (a positive d as the nearest, a non-positive one as the
            // largest non-positive); otherwise kHuge at the largest facet
            // index whose distance is kHuge (23, or 22 when q0 is 23 and its
            // own d is not kHuge).
            const double d = facet_plane_distance(q0, c9, pl, planeTol, x, y, z, a, b, g);
            if (d > -kHuge && d < kHuge) {
                nf.d = d;
                nf.f = q0;
            } else {
                nf.d = kHuge;
                nf.f = (d != kHuge && q0 == 23) ? 22 : 23;
            }
        } else if (unsure) {
            // the reference's test on the 12 facets of the three exit faces
            FindNearest fn = {1e80, -1, -1, -kHuge, 0};
            visit_axis<0>(fn, pl, c9, planeTol, x, y, z, a, b, g);
            visit_axis<1>(fn, pl, c9, planeTol, x, y, z, a, b, g);
            visit_axis<2>(fn, pl, c9, planeTol, x, y, z, a, b, g);
            nf = fn.result();
        } else {
            FindNearest fn = {1e80, -1, -1, -kHuge, 0};
            const int q1 = fp.c2, q2 = -1;
            // exact tests of the candidates, merged in facet order; every
            // other facet is kHuge, which only sets the largest kHuge index
            bool fin0 = false, fin1 = false, fin2 = false;
#pragma unroll 1
            for (int i = 0; i < n; i++) {
                const int f = i == 0 ? q0 : (i == 1 ? q1 : q2);
                const double d = exact_facet_test(f, pl, c9, planeTol, x, y, z, a, b, g);
                fn.add(d, f);
                const bool fin = d != kHuge;
                if (i == 0) fin0 = fin;
                else if (i == 1) fin1 = fin;
                else fin2 = fin;
            }
            int h = 23;
#pragma unroll
            for (int r = 0; r < 3; r++)
                if ((fin0 && q0 == h) || (fin1 && q1 == h) || (fin2 && q2 == h)) h--;
            fn.add(kHuge, h);
            nf = fn.result();
        }
    }
    if (nf.d == kHuge) {
        int iteration = 0;
        double moveFactor = 0.5 * kSmall;
        const double* cc = P.cellCenter + 3 * (size_t)cell;
        for (;;) {
            x += moveFactor * (cc[0] - x);
            y += moveFactor * (cc[1] - y);
            z += moveFactor * (cc[2] - z);
            iteration++;
            moveFactor *= 2.0;
            if (moveFactor > 1.0e-2) moveFactor = 1.0e-2;
            if (iteration == 10000) break;
            const double planeTol = 1e-16 * (x * x + y * y + z * z);
            FindNearest fn = {1e80, -1, -1, -kHuge, 0};
#pragma unroll 1
            for (int f = 0; f < 24; f++) fn.add(distance_any_facet(pl, c9, f, planeTol, x, y, z, a, b, g), f);
            nf = fn.result();
            if (nf.d != kHuge) break;
        }
    }
    if (nf.d < 0) nf.d = 0;
    return nf;
}

// DirectionCosine::Rotate3DVector (DirectionCosine.hh:123-146)
__device__ __forceinline__ void rotate3d(double& a, double& b, double& g, double sinT, double cosT,
                                         double sinP, double cosP) {
    double cos_theta = g;
    double sin_theta = sqrt((1.0 - (cos_theta * cos_theta)));
    double cos_phi, sin_phi;
    if (sin_theta < 1e-6) {
        cos_phi = 1.0;
        sin_phi = 0.0;
    } else {
        cos_phi = a / sin_theta;
        sin_phi = b / sin_theta;
    }
    double na = cos_theta * cos_phi * (sinT * cosP) - sin_phi * (sinT * sinP) + sin_theta * cos_phi * cosT;
    double nb = cos_theta * sin_phi * (sinT * cosP) + cos_phi * (sinT * sinP) + sin_theta * sin_phi * cosT;
    double ng = -sin_theta * (sinT * cosP) + cos_theta * cosT;
    a = na;
    b = nb;
    g = ng;
}

struct Lane {
    double x, y, z;
    double a, b, g;     // direction cosines
    double vx, vy, vz;  // velocity
    double speed;       // |velocity|, recomputed whenever velocity changes
    double E, w, tc, nmfp;
    uint64_t seed;
    int ci, cj, ck, group;
    int mat;  // the cell's material, read when the cell changes
};

// updateTrajectory (CollisionEvent.cc:25-45) up to its last step, the new
// number of mean free paths -1 * log(rngSample) (CollisionEvent.cc:44),
// which the caller draws (sample_mfp) so that one copy of the log serves
// every lane that needs a fresh draw.  kSinCos takes sin and cos of the
// azimuth from one sincos call (the same two values; it shares the argument
// reduction, at the cost of registers the 2-wave kernel builds have).
template <bool kSinCos>
__device__ __forceinline__ void update_direction_energy(Lane& L, double energy, double angle) {
    L.E = energy;
    double cosTheta = angle;
    double r = rng_sample(L.seed);
    double phi = 2 * 3.14159265 * r;
    double sinPhi, cosPhi;
    if (kSinCos) {
        sincos(phi, &sinPhi, &cosPhi);
    } else {
        sinPhi = sin(phi);
        cosPhi = cos(phi);
    }
    double sinTheta = sqrt((1.0 - (cosTheta * cosTheta)));
    rotate3d(L.a, L.b, L.g, sinTheta, cosTheta, sinPhi, cosPhi);
    double speed = (kLight * sqrt((1.0 - ((kRestMass * kRestMass) /
                                           ((energy + kRestMass) * (energy + kRestMass))))));
    L.vx = speed * L.a;
    L.vy = speed * L.b;
    L.vz = speed * L.g;
}

__device__ __forceinline__ void sample_mfp(Lane& L) {
    const double r = rng_sample(L.seed);
    L.nmfp = -1.0 * log(r);
}

__device__ __forceinline__ double vlength(double x, double y, double z) {
    return sqrt(x * x + y * y + z * z);  // MC_Vector::Length
}

// Writes record i of o.  The fields lie one stride apart (Particles): one
// address stepped by the stride per field, a single 64-bit add each, so no
// per-field base address is held in registers.
__device__ __forceinline__ void store_record(const Particles& o, uint32_t i, double x, double y, double z,
                                             double v0, double v1, double v2, double e, double w, double tc,
                                             double nmfp, uint64_t seed, int32_t cell) {
    typedef __attribute__((address_space(1))) char* Bytes;
    Bytes fp = (Bytes)o.base + (uint64_t)i * 8;
    const double v[10] = {x, y, z, v0, v1, v2, e, w, tc, nmfp};
#pragma unroll
    for (int k = 0; k < 10; k++) {
        *(__attribute__((address_space(1))) double*)fp = v[k];
        fp += o.stride;
        asm("" : "+v"(fp));  // keep the step an add
    }
    *(__attribute__((address_space(1))) uint64_t*)fp = seed;
    fp += o.stride;
    *(__attribute__((address_space(1))) int32_t*)(fp - (uint64_t)i * 4) = cell;
}

__device__ __forceinline__ void store_particle(const Particles& o, uint32_t i, const Lane& L, int cell) {
    store_record(o, i, L.x, L.y, L.z, L.vx, L.vy, L.vz, L.E, L.w, L.tc, L.nmfp, L.seed, cell);
}

__device__ __forceinline__ uint64_t lanemask_lt() {
    const int lane = __lane_id();
    return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

// Reserve `count` (0..4) consecutive slots per active lane with one atomic
// per wave; returns this lane's first slot.
// A collision output stored before its updateTrajectory: the pre-collision
// position, direction (in the velocity slots), weight and time to census,
// its own seed, and the sampled energy and scattering cosine (in the energy
// and mean-free-path slots); the cell is stored complemented (negative) to
// mark the record.  Loading it runs updateTrajectory and then the store and
// load the reference performs (direction re-derived from the velocity), so
// the particle is the same as if it had been updated before the store.
__device__ __forceinline__ void store_deferred(const Particles& o, uint32_t i, const Lane& L, uint64_t seed,
                                               double energy, double angle, int cell) {
    store_record(o, i, L.x, L.y, L.z, L.a, L.b, L.g, energy, L.w, L.tc, angle, seed, ~cell);
}
__device__ __forceinline__ uint32_t wave_append(uint32_t* counter, uint32_t count) {
    uint64_t act = __ballot(1);
    uint64_t b0 = __ballot(count > 0), b1 = __ballot(count > 1), b2 = __ballot(count > 2),
             b3 = __ballot(count > 3);
    uint32_t total = __popcll(b0) + __popcll(b1) + __popcll(b2) + __popcll(b3);
    uint64_t lt = lanemask_lt();
    uint32_t prefix = __popcll(b0 & lt) + __popcll(b1 & lt) + __popcll(b2 & lt) + __popcll(b3 & lt);
    int leader = __ffsll((unsigned long long)act) - 1;
    uint32_t base = 0;
    if (__lane_id() == leader && total) base = atomicAdd(counter, total);
    base = __shfl(base, leader);
    return base + prefix;
}

// wave_append onto a wave-private counter in LDS (no atomic needed: only
// this wave's leader lane touches it).
__device__ __forceinline__ uint32_t wave_push(uint32_t* top, uint32_t count) {
    uint64_t act = __ballot(1);
    uint64_t b0 = __ballot(count > 0), b1 = __ballot(count > 1), b2 = __ballot(count > 2),
             b3 = __ballot(count > 3);
    uint32_t total = __popcll(b0) + __popcll(b1) + __popcll(b2) + __popcll(b3);
    uint64_t lt = lanemask_lt();
    uint32_t prefix = __popcll(b0 & lt) + __popcll(b1 & lt) + __popcll(b2 & lt) + __popcll(b3 & lt);
    int leader = __ffsll((unsigned long long)act) - 1;
    uint32_t base = 0;
    if (__lane_id() == leader) {
        base = *top;
        *top = base + total;
    }
    base = __shfl(base, leader);
    return base + prefix;
}
__device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

// Tables an event needs, staged in LDS when they fit.
struct Tables {
    float eLog2Lo, eGroupsPerLog2;  // the group guess of energy_group
    const double* energies;
    const double* total;
    const double* mfp;
    const double* macro;
    const int32_t* reactType;
    const MatInfo* mat;
};

template <bool kLds>
__device__ __forceinline__ Tables stage_tables(const DevProblem& P, double* lds) {
    Tables t;
    t.eLog2Lo = P.eLog2Lo;
    t.eGroupsPerLog2 = P.eGroupsPerLog2;
    if (!kLds) {
        t.energies = P.energies;
        t.total = P.total;
        t.mfp = P.mfp;
        t.macro = P.macro;
        t.reactType = P.reactType;
        t.mat = P.mat;  // not reached: the global variant copies MatInfo to LDS too
        return t;
    }
    const int nE = P.G + 1, nT = P.nMat * P.G, nM = P.nMat * P.maxR * P.G;
    double* sE = lds;
    double* sT = sE + nE;
    double* sF = sT + nT;
    double* sM = sF + nT;
    for (int q = threadIdx.x; q < nE; q += blockDim.x) sE[q] = P.energies[q];
    for (int q = threadIdx.x; q < nT; q += blockDim.x) sT[q] = P.total[q];
    for (int q = threadIdx.x; q < nT; q += blockDim.x) sF[q] = P.mfp[q];
    for (int q = threadIdx.x; q < nM; q += blockDim.x) sM[q] = P.macro[q];
    t.energies = sE;
    t.total = sT;
    t.mfp = sF;
    t.macro = sM;
    t.reactType = P.reactType;
    t.mat = P.mat;
    return t;
}

// ------------------------------------------------------------- kernels ----

// Queue, outputs and counters of one problem in one tracking pass.
struct QueueIO {
    Particles in;
    uint32_t nIn;
    uint32_t* head;  // next queue entry handed to a wave
    Particles census;
    uint32_t* nCensus;
    uint32_t censusCap;
    Particles extra;
    uint32_t* nExtra;
    uint32_t extraCap;
    uint32_t* error;
    DevCounters* dc;
};

// Accessors naming the queue a wave works on: one problem (k_track) or a
// round-robin walk over the problems of a batch (k_batch_track).  Every
// value they return is wave-uniform.
struct SingleQueue {
    static constexpr bool kMulti = false;
    const DevProblem& P;
    const Tables& T;
    const QueueIO& io;
    __device__ __forceinline__ const DevProblem& prob() const { return P; }
    __device__ __forceinline__ Tables tables() const { return T; }
    __device__ __forceinline__ QueueIO queue() const { return io; }
    __device__ __forceinline__ void stage() const {}
    __device__ __forceinline__ bool next() { return false; }
};

// Wave-uniform copies of values read from device memory: the batch
// kernel's problem descriptors are read through vector loads, and
// readfirstlane moves the results into scalar registers (the compiler cannot
// prove such a load uniform by itself), which keeps them out of the vector
// register budget the tracking loop needs.
__device__ __forceinline__ int rfl(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ uint64_t rfl64(uint64_t v) {
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((uint32_t)v);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
    return ((uint64_t)hi << 32) | lo;
}
__device__ __forceinline__ float rflf(float v) {
    return __builtin_bit_cast(float, __builtin_amdgcn_readfirstlane(__builtin_bit_cast(int, v)));
}
__device__ __forceinline__ double rfld(double v) {
    return __builtin_bit_cast(double, rfl64(__builtin_bit_cast(uint64_t, v)));
}
template <class T>
__device__ __forceinline__ T* rflp(T* p) {
    return (T*)rfl64((uint64_t)p);
}
__device__ __forceinline__ Particles rflv(const Particles& v) { return Particles{rflp(v.base), rfl64(v.stride)}; }

__device__ __forceinline__ DevProblem load_problem(const DevProblem& c) {
    DevProblem P;
    P.nx = rfl(c.nx);
    P.rNx = rfld(c.rNx);
    P.rNxy = rfld(c.rNxy);
    P.ny = rfl(c.ny);
    P.nz = rfl(c.nz);
    P.nCells = rfl(c.nCells);
    P.dx = rfld(c.dx);
    P.dy = rfld(c.dy);
    P.dz = rfld(c.dz);
    P.hdx = rfld(c.hdx);
    P.hdy = rfld(c.hdy);
    P.hdz = rfld(c.hdz);
    P.boundaryMask = rfl(c.boundaryMask);
    P.planes = rflp(c.planes);
    P.cellMat = rflp(c.cellMat);
    P.cellVolume = rflp(c.cellVolume);
    P.cellCenter = rflp(c.cellCenter);
    P.srcOffset = rflp(c.srcOffset);
    P.G = rfl(c.G);
    P.nMat = rfl(c.nMat);
    P.maxR = rfl(c.maxR);
    P.eLog2Lo = rflf(c.eLog2Lo);
    P.eGroupsPerLog2 = rflf(c.eGroupsPerLog2);
    P.energies = rflp(c.energies);
    P.total = rflp(c.total);
    P.mfp = rflp(c.mfp);
    P.macro = rflp(c.macro);
    P.reactType = rflp(c.reactType);
    P.dt = rfld(c.dt);
    P.eMin = rfld(c.eMin);
    P.eMax = rfld(c.eMax);
    P.sourceWeight = rfld(c.sourceWeight);
    P.lowWeightCutoff = rfld(c.lowWeightCutoff);
    P.flux = rflp(c.flux);
    P.fluence = rflp(c.fluence);
    P.spectrum = rflp(c.spectrum);
    return P;
}

// kStageE: every problem's energy-group bounds fit the wave's LDS slot
// (kBatchStagedEnergies), and are read from there.
template <bool kStageE>
struct BatchQueue {
    static constexpr bool kMulti = true;
    const BatchSlot* slots;
    int n;
    int passKind;
    int p;
    MatInfo* sMat;     // this wave's LDS copy of the current problem's materials
    int32_t* sReact;   // ... and reaction types
    double* sE;        // ... and energy-group bounds (when they fit)
    __device__ __forceinline__ DevProblem prob() const { return load_problem(slots[p].prob); }
    // Copies the current problem's material records, reaction types (read on
    // every collision) and energy-group bounds (read by every group search)
    // into the wave's LDS slots.
    __device__ __forceinline__ void stage() const {
        const DevProblem& P = slots[p].prob;
        const int nM = rfl(P.nMat), nR = rfl(P.nMat * P.maxR);
        for (int q = __lane_id(); q < nM; q += 64) sMat[q] = P.mat[q];
        for (int q = __lane_id(); q < nR; q += 64) sReact[q] = P.reactType[q];
        if (kStageE) {
            const int nE = rfl(P.G) + 1;
            const double* e = rflp(P.energies);
            for (int q = __lane_id(); q < nE; q += 64) sE[q] = e[q];
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
    __device__ __forceinline__ Tables tables() const {
        const DevProblem& P = slots[p].prob;
        Tables T;
        T.eLog2Lo = rflf(P.eLog2Lo);
        T.eGroupsPerLog2 = rflf(P.eGroupsPerLog2);
        T.energies = kStageE ? (const double*)sE : rflp(P.energies);
        T.total = rflp(P.total);
        T.mfp = rflp(P.mfp);
        T.macro = rflp(P.macro);
        T.reactType = sReact;
        T.mat = sMat;
        return T;
    }
    __device__ __forceinline__ uint32_t* inCount(int q) const {
        BatchState* st = slots[q].st;
        return passKind == 0 ? &st->nPop : &st->extra[passKind == 1 ? 0 : 1];
    }
    __device__ __forceinline__ QueueIO queue() const {
        const BatchSlot& S = slots[p];
        BatchState* st = rflp(S.st);
        const int outIdx = passKind == 1 ? 1 : 0;
        const uint32_t cyc = (uint32_t)rfl((int)st->cycle);
        const uint32_t cap = (uint32_t)rfl((int)S.cap);
        const bool sorted = rfl((int)st->sorted) != 0;
        QueueIO io = {rflv(passKind == 0 ? (sorted ? S.proc[cyc & 1] : S.pop) : S.extra[passKind == 1 ? 0 : 1]),
                      (uint32_t)rfl((int)*inCount(p)),
                      &st->head,
                      rflv(S.proc[(cyc + 1) & 1]),
                      &st->nCensus,
                      cap,
                      rflv(S.extra[outIdx]),
                      &st->extra[outIdx],
                      cap,
                      &st->error,
                      rflp(S.dc)};
        return io;
    }
    // Move to the next problem (round robin) whose queue still has entries.
    __device__ __forceinline__ bool next() {
        for (int tries = 0; tries < n; tries++) {
            int q = (p + 1 == n) ? 0 : p + 1;
            p = __builtin_amdgcn_readfirstlane(q);
            BatchState* st = slots[p].st;
            if (!st->active) continue;
            const uint32_t nIn = (uint32_t)__builtin_amdgcn_readfirstlane(*inCount(p));
            const uint32_t h = (uint32_t)__builtin_amdgcn_readfirstlane(*(volatile uint32_t*)&st->head);
            if (h < nIn) return true;
        }
        return false;
    }
};

// Adds a wave's balance counters to the problem's tallies (one 64-bit
// atomic per nonzero counter, from lane 0) and clears them.
__device__ __forceinline__ void flush_tallies(DevCounters* dc, uint32_t& cSeg, uint32_t& cCol, uint32_t& cScat,
                                              uint32_t& cAbs, uint32_t& cFis, uint32_t& cProd, uint32_t& cEsc,
                                              uint32_t& cCen) {
    if (__lane_id() == 0) {
        if (cSeg) atomicAdd(&dc->numSegments, (unsigned long long)cSeg);
        if (cCol) atomicAdd(&dc->collision, (unsigned long long)cCol);
        if (cScat) atomicAdd(&dc->scatter, (unsigned long long)cScat);
        if (cAbs) atomicAdd(&dc->absorb, (unsigned long long)cAbs);
        if (cFis) atomicAdd(&dc->fission, (unsigned long long)cFis);
        if (cProd) atomicAdd(&dc->produce, (unsigned long long)cProd);
        if (cEsc) atomicAdd(&dc->escape, (unsigned long long)cEsc);
        if (cCen) atomicAdd(&dc->census, (unsigned long long)cCen);
    }
    cSeg = cCol = cScat = cAbs = cFis = cProd = cEsc = cCen = 0;
}

// Per-wave secondary stack: a view into the launch's stack memory and its
// top, kept in LDS so the lane that reserves entries inside divergent code
// can publish the new top to the rest of the wave.
// A byte address in global memory (address space 1): loads through it stay
// global loads after an optimisation barrier.
typedef const __attribute__((address_space(1))) char* GlobalBytes;

// Appends a wave's n staged census particles (region entries scap ..
// scap+n-1) to the census buffer: one atomic for all of them, then a copy,
// the wave's lanes taking one record each in turn.  Each record's fields
// are reached through one address per buffer stepped by that buffer's
// stride, as the refill reads them.  Called by all lanes of the wave with a
// wave-uniform n > 0.
__device__ __forceinline__ void flush_census_stage(const Particles& stack, uint32_t scap, uint32_t n,
                                                   const QueueIO& io) {
    typedef __attribute__((address_space(1))) char* GlobalBytesW;
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    const uint32_t lane = __lane_id();
    uint32_t base = 0;
    if (lane == 0) base = atomicAdd(io.nCensus, n);
    base = __shfl(base, 0);
    for (uint32_t q = lane; q < n; q += 64) {
        const uint32_t o = base + q;
        if (o < io.censusCap) {
            const uint32_t i = scap + q;
            GlobalBytes sp = (GlobalBytes)stack.base + (uint64_t)i * 8;
            GlobalBytesW dp = (GlobalBytesW)io.census.base + (uint64_t)o * 8;
            // the ten doubles and the seed, copied as 64-bit words
#pragma unroll
            for (int k = 0; k < 11; k++) {
                *(__attribute__((address_space(1))) uint64_t*)dp =
                    *(const __attribute__((address_space(1))) uint64_t*)sp;
                sp += stack.stride;
                dp += io.census.stride;
                asm("" : "+v"(sp), "+v"(dp));  // keep each step an add
            }
            // the int32 cell field: its array starts 11 strides in
            *(__attribute__((address_space(1))) int32_t*)(dp - (uint64_t)o * 4) =
                *(const __attribute__((address_space(1))) int32_t*)(sp - (uint64_t)i * 4);
        } else {
            atomicOr(io.error, 1u);
        }
    }
}

struct StackRef {
    Particles view;
    uint32_t cap;
    uint32_t* top;
};

// Tracks one problem's pass queue until it is drained (queue exhausted,
// stack empty, every lane idle), then adds the wave's balance counters to
// the problem's tallies.
template <bool kSinCos>
__device__ __forceinline__ void track_queue(const DevProblem& P, const Tables& T, const QueueIO& io,
                                            const TrackTuning& tune, const StackRef& sr) {
    const int G = P.G;
    // the group table's end bounds, read once: every group search tests them
    const double eLo = rfld(T.energies[0]), eHi = rfld(T.energies[G]);
    const uint32_t nIn = io.nIn;
    const int lane = __lane_id();
    const uint64_t lt = lanemask_lt();
    const Particles stack = sr.view;
    const uint32_t scap = sr.cap;
    uint32_t* sTopW = sr.top;
    uint32_t qBase = 0, qEnd = 0;
    bool exhausted = false;
    bool active = false;
    // Per-lane work carried to the update block at the top of the next
    // iteration: updateTrajectory (with the energy and scattering cosine
    // held in L.E and L.nmfp) for a scatter survivor or a loaded collision
    // record, and the energy-group search of every loaded or updated
    // particle.  One copy of each serves all lanes that need it.
    bool upd = false, rec = false, grp = false, ld = false;
    Lane L = {};
    uint32_t histSeg = 0;
    // wave-uniform (scalar) balance counters: segments from the lanes that
    // start one, the others from ballots at the converged end of each
    // iteration
    uint32_t cSeg = 0, cCol = 0, cScat = 0, cAbs = 0, cFis = 0, cProd = 0, cEsc = 0, cCen = 0;
    double fluxAcc = 0.0;  // this lane's scalar-flux contributions
    // Census particles staged in the wave's region after its stack (entries
    // scap, scap+1, ...), appended to the census buffer with one atomic when
    // the staging entries run short or the queue is drained.  The count must
    // be wave-uniform: it changes only outside the active lanes' branch,
    // from a ballot.
    uint32_t nStaged = 0;

    for (;;) {
        // ---- refill idle lanes: the wave's stack first (most recent
        // secondaries), then the pass queue (wave-uniform control) ----
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        uint32_t top = (uint32_t)__builtin_amdgcn_readfirstlane(*sTopW);
        uint64_t idle = __ballot(!active);
        if ((top != 0 || !exhausted) && idle != 0 &&
            (__popcll(idle) >= (uint32_t)tune.refillMin || idle == ~0ull)) {
            // Every idle lane's source is decided first, then one load per
            // field serves all of them: loads issued in separate rounds
            // into the same registers would each wait for the previous
            // round's to land.  The stack gives its top entries first; then
            // the rest of the current queue chunk (range A), then one new
            // claim of at least a chunk (range B).
            const uint32_t nIdle = __popcll(idle);
            const uint32_t takeS = min(top, nIdle);
            uint32_t need = nIdle - takeS;
            const uint32_t aBase = qBase;
            uint32_t aN = 0, bBase = 0, bN = 0;
            if (need != 0 && !exhausted) {
                aN = min(qEnd - qBase, need);
                need -= aN;
                qBase += aN;
                if (need != 0) {
                    const uint32_t want = max(need, (uint32_t)tune.chunk);
                    uint32_t got = 0;
                    if (lane == 0) got = atomicAdd(io.head, want);
                    got = __shfl(got, 0);
                    if (got >= nIn) {
                        exhausted = true;
                    } else {
                        const uint32_t end = min(got + want, nIn);
                        bBase = got;
                        bN = min(end - got, need);
                        qBase = got + bN;
                        qEnd = end;
                    }
                }
            }
            const uint32_t rank = __popcll(idle & lt);
            if (!active && rank < takeS + aN + bN) {
                // MC_Load_Particle (MC_Load_Particle.cc:11-29) +
                // MC_Particle(const MC_Base_Particle&) (MC_Base_Particle.hh:281-320)
                const bool fromStack = rank < takeS;
                const uint32_t r = rank - takeS;
                const uint32_t idx = fromStack ? top - 1 - rank : (r < aN ? aBase + r : bBase + (r - aN));
                // The record's fields lie one stride apart (Particles): one
                // address stepped by the lane's stride per field, a single
                // 64-bit add each, rather than base + k * stride formed per
                // field with the stride chosen per lane.
                const uint64_t stride = fromStack ? stack.stride : io.in.stride;
                GlobalBytes fp = (GlobalBytes)(fromStack ? stack.base : io.in.base) + (uint64_t)idx * 8;
                double* const dst[10] = {&L.x, &L.y, &L.z, &L.vx, &L.vy, &L.vz, &L.E, &L.w, &L.tc, &L.nmfp};
#pragma unroll
                for (int k = 0; k < 10; k++) {
                    *dst[k] = *(const __attribute__((address_space(1))) double*)fp;
                    fp += stride;
                    asm("" : "+v"(fp));  // keep the step an add
                }
                L.seed = *(const __attribute__((address_space(1))) uint64_t*)fp;
                fp += stride;
                // the int32 cell field, decoded after the refill
                L.ci = *(const __attribute__((address_space(1))) int32_t*)(fp - (uint64_t)idx * 4);
                ld = true;
                histSeg = 0;
                active = true;
            }
            top -= takeS;
            if (lane == 0) *sTopW = top;
        }
        // room for this iteration's pushes (uniform: decided before any lane diverges)
        const bool stackOk = top + kStackMargin <= scap;
        const uint64_t segLanes = __ballot(active);  // each takes one segment below
        if (nStaged + 64 > kCensusStage || (segLanes == 0 && nStaged != 0)) {
            flush_census_stage(stack, scap, nStaged, io);
            nStaged = 0;
        }
        if (segLanes == 0) break;  // drained
        cSeg += __popcll(segLanes);
        // ---- per-lane work of loaded particles and collision survivors,
        // one copy for all lanes that need it ----
        bool fresh = false;  // loaded particle starting without mean free paths
        if (ld) {
            // a collision output stored before its updateTrajectory has its
            // cell complemented and its direction in the velocity slots
            int cell = L.ci;
            rec = cell < 0;
            if (rec) cell = ~cell;
            decode_cell(P, cell, L.ci, L.cj, L.ck);
            L.mat = P.cellMat[cell];
            if (rec) {
                L.a = L.vx;
                L.b = L.vy;
                L.g = L.vz;
                upd = true;
            } else {
                // the first segment draws the mean free paths of a particle
                // that has none (MC_Segment_Outcome.cc:65-69); drawn here, the
                // same draw from the same seed, nothing drawn in between
                fresh = L.nmfp == 0.0;
            }
            grp = true;
        }
        // updateTrajectory (CollisionEvent.cc:135-140) of scatter survivors
        // and collision records
        if (upd) update_direction_energy<kSinCos>(L, L.E, L.nmfp);
        if (upd || fresh) sample_mfp(L);
        if (upd || ld) {
            L.speed = vlength(L.vx, L.vy, L.vz);
            if (ld) {
                // MC_Load_Particle (MC_Load_Particle.cc:11-29); a record's
                // store and reload
                if (L.speed > 0) {
                    double factor = 1.0 / L.speed;
                    L.a = factor * L.vx;
                    L.b = factor * L.vy;
                    L.g = factor * L.vz;
                } else {
                    L.a = L.b = L.g = 0.0;
                }
                if (L.tc <= 0.0) L.tc += P.dt;
            }
        }
        if (grp) L.group = energy_group(T.energies, G, L.E, eLo, eHi, T.eLog2Lo, T.eGroupsPerLog2);
        upd = rec = grp = ld = false;
        uint32_t ev = 0;  // per-lane event bits of this iteration
        uint32_t nProd = 0;
        // room to stage one census particle per lane (uniform)
        const bool cenRoom = scap != 0 && nStaged + 64 <= kCensusStage;
        if (active) {
        // ---- one segment: MC_Segment_Outcome (MC_Segment_Outcome.cc:31-226) ----
        const int cell = L.ci + P.nx * (L.cj + P.ny * L.ck);
        const int mat = L.mat;
        bool force = false;
        if (L.nmfp < 0.0) {
            force = true;
            L.nmfp = kSmall;
        }
        const double sigt = T.total[mat * G + L.group];
        const double mfp = T.mfp[mat * G + L.group];  // (sigt == 0.0) ? kHuge : 1.0 / sigt
        if (L.nmfp == 0.0) {
            double r = rng_sample(L.seed);
            L.nmfp = -1.0 * log(r);
        }
        double dColl = force ? kSmall : L.nmfp * mfp;
        double dCen = L.speed * L.tc;
        const Cell9 c9 = cell_nodes(P, L.ci, L.cj, L.ck);
        const double* pl = P.planes + (size_t)cell * 96;
        Nearest nf = nearest_facet(P, pl, cell, c9, L.x, L.y, L.z, L.a, L.b, L.g);
        double dFacet = nf.d;
        if (force) {
            dFacet = kHuge;
            dCen = kHuge;
            dColl = kTiny;
        }
        // MC_Find_Min: strict <, ties to the lower index
        int outcome = 0;
        double seg = dColl;
        if (dFacet < seg) {
            outcome = 1;
            seg = dFacet;
        }
        if (dCen < seg) {
            outcome = 2;
            seg = dCen;
        }
        L.nmfp -= seg / mfp;
        if (outcome == 0) L.nmfp = 0.0;
        else if (outcome == 2) L.tc = (L.tc < 0.0) ? L.tc : 0.0;
        if (force) L.nmfp = 0.0;
        if (seg != 0.0) {
            L.x += (L.a * seg);
            L.y += (L.b * seg);
            L.z += (L.g * seg);
            double t = seg / L.speed;
            L.tc -= t;
            if (L.tc < 0.0) L.tc = 0.0;
            const double contrib = seg * L.w;
            unsafeAtomicAdd(&P.flux[(size_t)cell * G + L.group], contrib);
            fluxAcc += contrib;
        }
        histSeg++;
        // A census particle's slot: staged in the wave's region after its
        // stack, so that no atomic round trip stalls the loop, or, for a
        // wave without a region (no stacks), appended to the census buffer
        // at once.
        uint32_t censusSlot = 0;
        if (cenRoom) censusSlot = scap + nStaged + (uint32_t)__popcll(__ballot(outcome == 2) & lt);
        else if (outcome == 2) censusSlot = wave_append(io.nCensus, 1u);

        if (outcome == 0) {
            // ---- CollisionEvent (CollisionEvent.cc:50-148) ----
            const MatInfo mi = T.mat[mat];
            double r = rng_sample(L.seed);
            double cur = sigt * r;
            int selR = -1;
            // Every isotope of a material carries the same per-reaction
            // macroscopic cross sections, so the scan subtracts the same
            // nReact values once per isotope; with up to three reactions
            // they are the ones read with sigt.
            const double* mrow = T.macro + (size_t)(mat * P.maxR) * G + L.group;
            if (mi.nReact <= 3) {
                const double m0 = mrow[0];
                const double m1 = mi.nReact > 1 ? mrow[G] : 0.0;
                const double m2 = mi.nReact > 2 ? mrow[2 * G] : 0.0;
                const bool has1 = mi.nReact > 1, has2 = mi.nReact > 2;
                // The scan's answer without the scan.  Exactly, the partial
                // sums are cur - k S - M_r (isotope k, M_r = m0 + .. + m_r,
                // S = M_(nReact-1)), decreasing, so the first negative one
                // lies in isotope k = floor(cur / S), at the first r with
                // rem = cur - k S < M_r.  The scan's rounded sums differ
                // from the exact ones by at most one rounding per
                // subtraction, under 3 nIso 2^-53 (cur + S), and the
                // estimates here add a few more roundings; when the sums on
                // both sides of the crossing are farther from zero than the
                // bound B (four times that), the rounded scan changes sign
                // at the same place and selects the same reaction.  A lane
                // whose crossing is nearer zero, or none within its
                // isotopes, runs the scan below.
                {
                    const double M1 = m0 + m1, S = M1 + m2;
                    if (S > 0.0) {
                        const double kq = floor(cur / S);
                        if (kq < (double)mi.nIso) {
                            const double rem = fma(-kq, S, cur);
                            const double B = (12.0 * mi.nIso + 16.0) * 0x1p-53 * (cur + S);
                            const int r = rem < m0 ? 0 : (rem < M1 ? 1 : 2);
                            const double below = r == 0 ? 0.0 : (r == 1 ? m0 : M1);  // M_(r-1)
                            const double at = r == 0 ? m0 : (r == 1 ? M1 : S);        // M_r
                            if (r < mi.nReact && rem - below >= B && rem - at <= -B) selR = r;
                        }
                    }
                }
                // The scan itself, for the lanes left: one isotope per step
                // as straight-line code, the same subtractions in the same
                // order, the first partial sum below zero selecting the
                // reaction; a lane that has selected, or scanned all its
                // isotopes, keeps its values while the wave finishes the
                // others (a uniform loop, with no per-lane branch to leave
                // it).
                for (int iso = 0;; iso++) {
                    const bool live = selR < 0 && iso < mi.nIso;
                    if (__ballot(live) == 0) break;
                    const double c1 = cur - m0;
                    const double c2 = has1 ? c1 - m1 : c1;
                    const double c3 = has2 ? c2 - m2 : c2;
                    if (live) {
                        if (c1 < 0) selR = 0;
                        else if (has1 && c2 < 0) selR = 1;
                        else if (has2 && c3 < 0) selR = 2;
                        else cur = c3;
                    }
                }
            } else {
                for (int iso = 0; iso < mi.nIso && cur >= 0; iso++) {
                    for (int rr = 0; rr < mi.nReact; rr++) {
                        cur -= mrow[(size_t)rr * G];
                        if (cur < 0) {
                            selR = rr;
                            break;
                        }
                    }
                }
            }
            // The reference leaves the reaction undefined if rounding keeps
            // cur >= 0 after the full scan; the last reaction is used here.
            if (selR < 0) selR = mi.nReact - 1;
            const int type = T.reactType[mat * P.maxR + selR];
            // NuclearDataReaction::sampleCollision (NuclearData.cc:54-88):
            // the multiplicity, then the outgoing (energy, angle) pairs in
            // order, each drawn once.  Pair 0 is the parent's.
            int nOut = 0;
            if (type == kScatter) {
                nOut = 1;
            } else if (type == kFission) {
                int n = (int)(mi.nuBar + rng_sample(L.seed));
                nOut = n;
                if (nOut > 4) {
                    atomicOr(io.error, 16u);  // MAX_PRODUCTION_SIZE exceeded
                    nOut = 4;
                }
            }
            // pair 0: a scatter's two draws, or the two after a fission's
            // multiplicity, drawn once for both kinds of lane
            double e0 = 0.0, a0 = 0.0;
            if (nOut > 0) {
                const double x1 = rng_sample(L.seed);
                const double x2 = rng_sample(L.seed);
                if (type == kScatter) {
                    e0 = L.E * (1.0 - (x1 * mi.invMass));  // 1.0 / mass
                } else {
                    const double rr = x1 / 2.0 + 0.5;
                    e0 = (20 * rr * rr);
                }
                a0 = x2 * 2.0 - 1.0;
            }
            ev |= kEvCollision;
            if (type == kScatter) ev |= kEvScatter;
            else if (type == kAbsorption) ev |= kEvAbsorb;
            else if (type == kFission) {
                ev |= kEvFission;
                nProd = (uint32_t)nOut;
            }
            if (nOut == 0) {
                active = false;
            } else {
                // Outputs: secondaries j = 1..nOut-1 are copies of the
                // pre-collision parent with outgoing pair j and spawned seeds
                // (CollisionEvent.cc:125-133); the parent takes pair 0 and the
                // seed left after the spawns (CollisionEvent.cc:135) and stays
                // in the lane, as the reference's parent continues its
                // history.  The spawns start from the seed after the last
                // pair, one jump of 2 (nOut - 1) draws from the seed after
                // pair 0, while pairs 1.. are drawn from that seed in order.
                // The secondaries go to the wave's stack (or the pass's extra
                // queue when the stack is full) as records whose
                // updateTrajectory runs when they are loaded.
                if (nOut > 1) {
                    const Particles dst = stackOk ? stack : io.extra;
                    const uint32_t dcap = stackOk ? scap : io.extraCap;
                    const uint32_t slot = stackOk ? wave_push(sTopW, (uint32_t)(nOut - 1))
                                                  : wave_append(io.nExtra, (uint32_t)(nOut - 1));
                    uint64_t t = L.seed;
                    L.seed = rng_skip_pairs(t, nOut - 1);
#pragma unroll 1
                    for (int j = 1; j < nOut; j++) {
                        const double rr = rng_sample(t) / 2.0 + 0.5;
                        const double ek = (20 * rr * rr);
                        const double ak = rng_sample(t) * 2.0 - 1.0;
                        const uint64_t seedJ = rng_spawn(L.seed);
                        const uint32_t o = slot + (uint32_t)(j - 1);
                        if (o < dcap) store_deferred(dst, o, L, seedJ, ek, ak, cell);
                        else atomicOr(io.error, 2u);
                    }
                }
                // the parent: updated at the top of the next iteration
                L.E = e0;
                L.nmfp = a0;
                upd = true;
                grp = true;
            }
        } else if (outcome == 1) {
            // ---- MC_Facet_Crossing_Event (MC_Facet_Crossing_Event.cc:25-70) ----
            const int face = nf.f >> 2;
            int ni = L.ci, nj = L.cj, nk = L.ck;
            if (face == 0) ni++;
            else if (face == 1) ni--;
            else if (face == 2) nj++;
            else if (face == 3) nj--;
            else if (face == 4) nk++;
            else nk--;
            const bool outside = ni < 0 || ni >= P.nx || nj < 0 || nj >= P.ny || nk < 0 || nk >= P.nz;
            if (!outside) {
                L.ci = ni;
                L.cj = nj;
                L.ck = nk;
                // read here, used by the next segment: the load has the rest
                // of this iteration and the next one's set-up to land
                L.mat = P.cellMat[ni + P.nx * (nj + P.ny * nk)];
            } else if (((P.boundaryMask >> (2 * face)) & 3) == kEscape) {
                ev |= kEvEscape;
                active = false;
            } else {
                // MCT_Reflect_Particle (MCT.cc:401-429) with the facet's
                // plane normal formed in registers: N along the face axis
                // (facet_nd, the stored plane's own arithmetic) and signed
                // zeros elsewhere (kFacetCodes bits 24-26)
                double N, Dn;
                facet_nd(cell_nodes(P, L.ci, L.cj, L.ck), nf.f, N, Dn);
                const uint32_t fcode = s_facetCode[nf.f];
                const int fax = nf.f >> 3;
                const double A = fax == 0 ? N : ((fcode >> 24) & 1u ? -0.0 : 0.0);
                const double B = fax == 1 ? N : ((fcode >> 25) & 1u ? -0.0 : 0.0);
                const double C = fax == 2 ? N : ((fcode >> 26) & 1u ? -0.0 : 0.0);
                double dot = 2.0 * (L.a * A + L.b * B + L.g * C);
                if (dot > 0) {
                    L.a -= dot * A;
                    L.b -= dot * B;
                    L.g -= dot * C;
                }
                double speed = L.speed;
                L.vx = speed * L.a;
                L.vy = speed * L.b;
                L.vz = speed * L.g;
                L.speed = vlength(L.vx, L.vy, L.vz);
            }
        } else {
            // ---- census: processedVault->pushParticle (CycleTracking.cc:104-111) ----
            if (cenRoom) store_particle(stack, censusSlot, L, cell);
            else if (censusSlot < io.censusCap) store_particle(io.census, censusSlot, L, cell);
            else atomicOr(io.error, 1u);
            // EnergySpectrum::UpdateSpectrum (EnergySpectrum.cc:12-35, called
            // from Tallies.cc:97 on every census particle of every cycle):
            // the particle's getEnergyGroup, which is G above eMax
            if (P.spectrum) atomicAdd(&P.spectrum[(L.E > eHi) ? G : L.group], 1ull);
            ev |= kEvCensus;
            active = false;
        }
        if (active && histSeg >= kMaxSegmentsPerHistory) {
            atomicOr(io.error, 8u);
            active = false;
            upd = false;
            grp = false;
        }
        }  // active
        cCol += __popcll(__ballot(ev & kEvCollision));
        cScat += __popcll(__ballot(ev & kEvScatter));
        cAbs += __popcll(__ballot(ev & kEvAbsorb));
        cFis += __popcll(__ballot(ev & kEvFission));
        cEsc += __popcll(__ballot(ev & kEvEscape));
        const uint32_t nCenIt = (uint32_t)__popcll(__ballot(ev & kEvCensus));
        cCen += nCenIt;
        if (cenRoom) nStaged += nCenIt;  // the lanes that staged above
        cProd += __popcll(__ballot(nProd > 0)) + __popcll(__ballot(nProd > 1)) +
                 __popcll(__ballot(nProd > 2)) + __popcll(__ballot(nProd > 3));
    }
    flush_tallies(io.dc, cSeg, cCol, cScat, cAbs, cFis, cProd, cEsc, cCen);
    // The cycle's flux sum (Tallies::ScalarFluxSum, Tallies.cc:146-171, the
    // sum over all bins) as the sum of the contributions that filled the
    // bins, so the bins need not be read back; only the summation order
    // differs, as it already does between the device's atomic adds.
    double fs = fluxAcc;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) fs += __shfl_xor(fs, o);
    if (lane == 0 && fs != 0.0) unsafeAtomicAdd(&io.dc->fluxSum, fs);
}

// Tracks every particle of the wave's queue(s) to the end of its history
// in this pass (CycleTrackingGuts, CycleTracking.cc:15-119): one queue
// (k_track) or, round robin, every problem of a batch with queued particles
// (k_batch_track).  Secondaries go to the wave's stack and are tracked
// before the wave takes new queue entries.
template <bool kSinCos, class Q>
__device__ __forceinline__ void run_queue(Q& acc, const TrackTuning& tune, const WaveStacks& ws) {
    __shared__ uint32_t sTop[4];
    const uint32_t wib = threadIdx.x >> 6;
    const uint32_t gw = (uint32_t)__builtin_amdgcn_readfirstlane((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
    const uint32_t scap = (gw < ws.nWaves) ? ws.cap : 0u;
    const uint32_t region = waveRegionEntries(scap);  // stack, then census staging
    const StackRef sr = {{ws.base + (size_t)gw * 12 * region * 8, (uint64_t)region * 8}, scap, &sTop[wib]};
    if (__lane_id() == 0) sTop[wib] = 0;
    do {
        acc.stage();
        track_queue<kSinCos>(acc.prob(), acc.tables(), acc.queue(), tune, sr);
    } while (Q::kMulti && acc.next());
}

// kMinWaves: waves per SIMD the register allocation must allow (2: up to
// 256 VGPRs, 3: 168, 4: 128); a tuning knob, since a tighter budget trades
// occupancy against spills.
template <bool kLds, int kMinWaves>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(kMinWaves))) k_track(DevProblem P, TrackTuning tune, WaveStacks ws, Particles in,
                                               const uint32_t* __restrict__ nInPtr, Particles census,
                                               uint32_t censusCap, Particles extra, uint32_t* nExtraOut,
                                               uint32_t extraCap, PassCounters* pc, DevCounters* dc,
                                               const uint32_t* __restrict__ hold) {
    if (*hold) return;
    extern __shared__ double lds[];
    __shared__ MatInfo sMat[kMaxMaterials];
    __shared__ int32_t sReact[kMaxMaterials * kMaxReactions];
    load_facet_codes();
    Tables T = stage_tables<kLds>(P, lds);
    for (int q = threadIdx.x; q < P.nMat; q += blockDim.x) sMat[q] = P.mat[q];
    for (int q = threadIdx.x; q < P.nMat * P.maxR; q += blockDim.x) sReact[q] = P.reactType[q];
    __syncthreads();
    T.mat = sMat;
    T.reactType = sReact;
    QueueIO io = {in, *nInPtr, &pc->head, census, &pc->nCensus, censusCap, extra, nExtraOut, extraCap,
                  &pc->error, dc};
    SingleQueue acc = {P, T, io};
    run_queue<(kMinWaves <= 2)>(acc, tune, ws);
}

constexpr int kFluxBlock = 256;

// Tallies.cc:95: the flux bins restart from zero each cycle.  16-byte
// stores over [lo, hi) of p (p 16-byte aligned, lo even).
__device__ __forceinline__ void zero_doubles(double* p, size_t lo, size_t hi, size_t t, size_t stride) {
    double2* q = reinterpret_cast<double2*>(p + lo);
    const size_t n2 = (hi - lo) / 2;
    for (size_t k = t; k < n2; k += stride) q[k] = make_double2(0.0, 0.0);
    if (t == 0 && ((hi - lo) & 1)) p[hi - 1] = 0.0;
}

// MC_SourceNow particle generation (MC_SourceNow.cc:68-151) with
// MCT_Generate_Coordinate_3D_G (MCT.cc:143-226) and Sample_Isotropic
// (DirectionCosine.cc:5-13).
// Source particle number t (0-based within the problem's cycle) of cycle
// `cycle`, written at out[o].
__device__ __forceinline__ void source_particle(const DevProblem& P, const Particles& out, uint32_t o,
                                                uint64_t t, uint32_t cycle) {
    // cell = last c with srcOffset[c] <= t
    int lo = 0, hi = P.nCells;  // srcOffset[lo] <= t < srcOffset[hi]
    while (hi - lo > 1) {
        int mid = (lo + hi) >> 1;
        if (P.srcOffset[mid] <= t) lo = mid;
        else hi = mid;
    }
    const int cell = lo;
    const uint64_t first = P.srcOffset[cell];
    const uint64_t nCell = P.srcOffset[cell + 1] - first;
    const uint64_t k = t - first;
    // cell._sourceTally counts every particle ever sourced in the cell;
    // the cell id is gid * 2^32 (MC_Domain.cc:390)
    uint64_t base = (uint64_t)cycle * nCell + k + ((uint64_t)cell << 32);
    uint64_t seed = rng_spawn(base);

    const int ci = cell % P.nx, cj = (cell / P.nx) % P.ny, ck = cell / (P.nx * P.ny);
    const Cell9 c9 = cell_nodes(P, ci, cj, ck);
    const double cx = P.cellCenter[3 * (size_t)cell], cy = P.cellCenter[3 * (size_t)cell + 1],
                 cz = P.cellCenter[3 * (size_t)cell + 2];
    double r = rng_sample(seed);
    double whichVolume = r * 6.0 * P.cellVolume[cell];
    double currentVolume = 0.0;
    int facet = -1;
    Node3 p0 = {0, 0, 0}, p1 = {0, 0, 0}, p2 = {0, 0, 0};
    while (currentVolume < whichVolume) {
        facet++;
        if (facet == 24) break;
        uint32_t code = s_facetCode[facet];
        p0 = facet_node(c9, code, 0);
        p1 = facet_node(c9, code, 1);
        p2 = facet_node(c9, code, 2);
        // MCT_Cell_Volume_3D_G_vector_tetDet
        double v0x = p0.x - cx, v0y = p0.y - cy, v0z = p0.z - cz;
        double v1x = p1.x - cx, v1y = p1.y - cy, v1z = p1.z - cz;
        double v2x = p2.x - cx, v2y = p2.y - cy, v2z = p2.z - cz;
        double sub = v0z * (v1x * v2y - v1y * v2x) + v0y * (v1z * v2x - v1x * v2z) +
                     v0x * (v1y * v2z - v1z * v2y);
        currentVolume += sub;
    }
    double r1 = rng_sample(seed);
    double r2 = rng_sample(seed);
    double r3 = rng_sample(seed);
    if (r1 + r2 > 1.0) {
        r1 = 1.0 - r1;
        r2 = 1.0 - r2;
    }
    if (r2 + r3 > 1.0) {
        double tmp = r3;
        r3 = 1.0 - r1 - r2;
        r2 = 1.0 - tmp;
    } else if (r1 + r2 + r3 > 1.0) {
        double tmp = r3;
        r3 = r1 + r2 + r3 - 1.0;
        r1 = 1.0 - r2 - tmp;
    }
    double r4 = 1.0 - r1 - r2 - r3;
    double x = 0.0, y = 0.0, z = 0.0;
    if (facet >= 0) {
        x = (r4 * cx + r1 * p0.x + r2 * p1.x + r3 * p2.x);
        y = (r4 * cy + r1 * p0.y + r2 * p1.y + r3 * p2.y);
        z = (r4 * cz + r1 * p0.z + r2 * p1.z + r3 * p2.z);
    }
    double g = 1.0 - 2.0 * rng_sample(seed);
    double sineGamma = sqrt((1.0 - (g * g)));
    double phi = kPi * (2.0 * rng_sample(seed) - 1.0);
    double a = sineGamma * cos(phi);
    double b = sineGamma * sin(phi);
    double E = (P.eMax - P.eMin) * rng_sample(seed) + P.eMin;
    double speed = kLight * sqrt(E * (E + 2.0 * (kRestMass)) / ((E + kRestMass) * (E + kRestMass)));
    double nmfp = -1.0 * log(rng_sample(seed));
    double tc = P.dt * rng_sample(seed);
    out.x()[o] = x;
    out.y()[o] = y;
    out.z()[o] = z;
    out.vx()[o] = speed * a;
    out.vy()[o] = speed * b;
    out.vz()[o] = speed * g;
    out.E()[o] = E;
    out.w()[o] = P.sourceWeight;
    out.tc()[o] = tc;
    out.nmfp()[o] = nmfp;
    out.seed()[o] = seed;
    out.cell()[o] = cell;
}

// Blocks below srcBlocks generate source particles; the others (if any)
// zero the previous cycle's flux bins, which the source does not touch: the
// latency-bound source and the bandwidth-bound zeroing share the device.
__global__ void __launch_bounds__(256) k_source(DevProblem P, Particles out, const CycleCtl* __restrict__ ctl,
                                                uint64_t nSource, uint32_t cycle, uint32_t srcBlocks, double* flux,
                                                size_t nFlux) {
    if (ctl->hold) return;
    if (blockIdx.x >= srcBlocks) {
        zero_doubles(flux, 0, nFlux, (size_t)(blockIdx.x - srcBlocks) * blockDim.x + threadIdx.x,
                     (size_t)(gridDim.x - srcBlocks) * blockDim.x);
        return;
    }
    load_facet_codes();
    __syncthreads();
    uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nSource) return;
    source_particle(P, out, ctl->start + (uint32_t)t, t, cycle);
}

__device__ __forceinline__ uint32_t wave_inclusive_scan(uint32_t v) {
    const int lane = __lane_id();
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t t = __shfl_up(v, o);
        if (lane >= o) v += t;
    }
    return v;
}

// PopulationControlGuts (PopulationControl.cc:71-122) followed by
// RouletteLowWeightParticles (PopulationControl.cc:127-171), fused per
// particle: every decision uses the particle's own RNG stream, so the
// surviving multiset does not depend on processing order.  The outputs of
// one input particle are the parent (its seed after all spawns) and
// splitFactor copies (seed j = hash of the parent seed after j spawns); the
// low-weight roulette then draws once from each output's own stream.
struct PopDecision {
    uint64_t seed;
    double w;
    int splitFactor;
    bool keep;
    bool roulette;
    uint32_t survivors, kills, splits;
};

__device__ __forceinline__ PopDecision pop_decide(const DevProblem& P, const Particles& in, uint32_t i, bool valid,
                                                  int mode, double factor) {
    PopDecision d = {0, 0.0, 0, valid, false, 0, 0, 0};
    const double cutoff = P.lowWeightCutoff;
    const double weightCutoff = cutoff * P.sourceWeight;
    if (valid) {
        d.seed = in.seed()[i];
        d.w = in.w()[i];
        if (mode == 1) {
            double r = rng_sample(d.seed);
            if (r > factor) {
                d.keep = false;
                d.kills++;
            } else {
                d.w /= factor;
            }
        } else if (mode == 2) {
            double r = rng_sample(d.seed);
            d.splitFactor = (int)floor(factor);
            if (r > (factor - d.splitFactor)) d.splitFactor--;
            d.w /= factor;
            d.splits += (uint32_t)d.splitFactor;
        }
    }
    d.roulette = cutoff > 0.0 && d.w <= weightCutoff;
    if (d.keep) {
        uint64_t s = d.seed;
        for (int j = 0; j < d.splitFactor; j++) {
            uint64_t cs = rng_spawn(s);
            if (d.roulette) {
                if (rng_sample(cs) <= cutoff) d.survivors++;
                else d.kills++;
            } else {
                d.survivors++;
            }
        }
        uint64_t ps = s;
        if (d.roulette) {
            if (rng_sample(ps) <= cutoff) d.survivors++;
            else d.kills++;
        } else {
            d.survivors++;
        }
    }
    return d;
}

// The next surviving output of one input particle.  The outputs are, in
// order, the splitFactor spawned copies and then the parent; under roulette
// an output survives its own draw (its seed is the stream after that draw).
// st is the parent's stream, advanced by each spawn; j counts the outputs
// examined.
__device__ __forceinline__ uint64_t pop_next_survivor(const PopDecision& d, double cutoff, uint64_t& st, int& j) {
    for (;;) {
        const bool copy = j < d.splitFactor;
        uint64_t t = copy ? rng_spawn(st) : st;
        j++;
        if (!d.roulette) return t;
        if (rng_sample(t) <= cutoff) return t;
        if (!copy) return t;  // not reached: the parent is counted only if it survives
    }
}

// Population control with the outputs written in order: one thread decides
// for each input particle (pop_decide), a block scan gives each input its
// output range, and the block stages its outputs a tile at a time in LDS
// and stores them field by field, so every store instruction covers
// consecutive slots
// (one thread writing its ~factor copies one after another touches factor
// times more cache lines per store).
constexpr uint32_t kPopTile = 256;  // outputs staged per round (12 fields: 24 KiB of LDS)
constexpr uint32_t kPopMinGroup = 64;  // fewest inputs per block-wide population group

// Exclusive scan of v over the block (at most 16 waves); total receives the
// block's sum.  Every thread of the block must call it.
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t& total) {
    __shared__ uint32_t sWave[16];
    const uint32_t wid = threadIdx.x >> 6;
    const uint32_t incl = wave_inclusive_scan(v);
    if (__lane_id() == 63) sWave[wid] = incl;
    __syncthreads();
    uint32_t excl = incl - v;
    total = 0;
    for (uint32_t w = 0; w < (blockDim.x >> 6); w++) {
        if (w < wid) excl += sWave[w];
        total += sWave[w];
    }
    __syncthreads();  // sWave may be written again after this
    return excl;
}

// Writes the outputs of one group of inputs: thread t's surv outputs go to
// out[base + excl ...]; total is the group's output count.  The outputs are
// staged a tile at a time, all twelve fields at once, in LDS, then stored
// field by field over consecutive slots.  Slots at or beyond outCap are not
// written (error bit 4).
__device__ __forceinline__ void pop_write(const DevProblem& P, const Particles& in, uint32_t i, const PopDecision& d,
                                          uint32_t surv, uint32_t excl, uint32_t total, uint32_t base,
                                          const Particles& out, uint32_t outCap, uint32_t* error) {
    __shared__ double sBuf[12][kPopTile];
    const double cutoff = P.lowWeightCutoff;
    double f[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    int32_t cell = 0;
    if (surv) {
#pragma unroll
        for (int k = 0; k < 10; k++) f[k] = in.f(k)[i];
        f[7] = d.roulette ? d.w / cutoff : d.w;  // weight after splitting and roulette
        cell = in.cell()[i];
    }
    uint64_t st = d.seed;  // parent stream, advanced by the spawns
    int j = 0;
    for (uint32_t t0 = 0; t0 < total; t0 += kPopTile) {
        const uint32_t t1 = min(t0 + kPopTile, total);
        const uint32_t lo = max(excl, t0), hi = min(excl + surv, t1);
        // fields 0-9 (the weight is field 7), the seed, the cell
        for (uint32_t o = lo; o < hi; o++) {
            const uint32_t q = o - t0;
#pragma unroll
            for (int k = 0; k < 10; k++) sBuf[k][q] = f[k];
            sBuf[10][q] = __longlong_as_double((long long)pop_next_survivor(d, cutoff, st, j));
            reinterpret_cast<int32_t*>(sBuf[11])[q] = cell;
        }
        __syncthreads();
        for (uint32_t q = threadIdx.x; q < t1 - t0; q += blockDim.x) {
            const uint32_t o = base + t0 + q;
            if (o >= outCap) {
                atomicOr(error, 4u);
                continue;
            }
#pragma unroll
            for (int k = 0; k < 10; k++) out.f(k)[o] = sBuf[k][q];
            out.seed()[o] = (uint64_t)__double_as_longlong(sBuf[10][q]);
            out.cell()[o] = reinterpret_cast<const int32_t*>(sBuf[11])[q];
        }
        __syncthreads();
    }
}

// ---- one problem's population control: count, scan, write
// Same-address atomics from every workgroup are slow on this device (they
// serialise across the chiplets), so the one-problem form places each
// group's outputs with a scan instead: k_pop_count stores each group's
// output, roulette and split counts, k_pop_scan (one workgroup) turns them
// into output bases and the balance counts, and k_population writes.  The
// groups are the same in all three: group g holds inputs [g*size, (g+1)*size)
// (pop_group_size), visited by block g mod gridDim.

// Inputs per group: about 256 / factor when splitting (at least 64), so that
// a group's outputs fill about one store round of the block; else 256.
__device__ __forceinline__ uint32_t pop_group_size(const CycleCtl& c) {
    if (c.mode != 2) return 256u;
    const double g = 256.0 / ceil(c.factor);
    return g < (double)kPopMinGroup ? kPopMinGroup : (uint32_t)g;
}

__global__ void __launch_bounds__(256) k_pop_count(DevProblem P, Particles in, const CycleCtl* __restrict__ ctl,
                                                   PopGroup* __restrict__ groups) {
    __shared__ uint32_t sSum[3][4];
    if (ctl->hold) return;
    const uint32_t n = ctl->n, size = pop_group_size(*ctl);
    const uint32_t nGroups = (n + size - 1) / size;
    const int mode = (int)ctl->mode;
    const double factor = ctl->factor;
    const uint32_t wid = threadIdx.x >> 6;
    for (uint32_t g = blockIdx.x; g < nGroups; g += gridDim.x) {
        const uint32_t i = g * size + threadIdx.x;
        const bool valid = threadIdx.x < size && i < n;
        const PopDecision d = pop_decide(P, in, i, valid, mode, factor);
        const uint32_t surv = wave_sum((d.keep && valid) ? d.survivors : 0u);
        const uint32_t kills = wave_sum(d.kills), splits = wave_sum(d.splits);
        if (__lane_id() == 0) {
            sSum[0][wid] = surv;
            sSum[1][wid] = kills;
            sSum[2][wid] = splits;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            PopGroup r = {0u, 0u, 0u, 0u};
            for (uint32_t w = 0; w < (blockDim.x >> 6); w++) {
                r.total += sSum[0][w];
                r.kills += sSum[1][w];
                r.splits += sSum[2][w];
            }
            groups[g] = r;
        }
        __syncthreads();
    }
}

__device__ __forceinline__ unsigned long long wave_sum64(unsigned long long v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

// One workgroup of 1024 threads: exclusive scan of the group output counts
// into their bases, the roulette and split counts into the balance, the
// output count into nPopOut.  More outputs than outCap: error bit 4, hold,
// and nPopOut stays at outCap.
__global__ void __launch_bounds__(1024) k_pop_scan(CycleBlock* cb, PopGroup* __restrict__ groups, uint32_t outCap) {
    __shared__ unsigned long long sK[16], sS[16], sT[16];
    CycleCtl& c = cb->ctl;
    if (c.hold) return;
    const uint32_t size = pop_group_size(c);
    const uint32_t nGroups = (c.n + size - 1) / size;
    // thread t scans groups [t * per, (t + 1) * per) by itself, so one block
    // scan of the threads' sums covers every group
    const uint32_t per = (nGroups + blockDim.x - 1) / blockDim.x;
    const uint32_t g0 = min(threadIdx.x * per, nGroups), g1 = min(g0 + per, nGroups);
    uint64_t mine = 0;
    unsigned long long kills = 0, splits = 0;
    for (uint32_t g = g0; g < g1; g++) {
        const PopGroup r = groups[g];
        mine += r.total;
        kills += r.kills;
        splits += r.splits;
    }
    // the bases are exact while the outputs fit 32 bits; more than that
    // exceeds the capacity, which the exact 64-bit total below reports
    uint32_t tot;
    const uint32_t excl = block_excl_scan(mine > 0xFFFFFFFFull ? 0xFFFFFFFFu : (uint32_t)mine, tot);
    uint64_t run = excl;
    for (uint32_t g = g0; g < g1; g++) {
        groups[g].base = (uint32_t)run;
        run += groups[g].total;
    }
    unsigned long long carry = wave_sum64(mine);
    kills = wave_sum64(kills);
    splits = wave_sum64(splits);
    const uint32_t wid = threadIdx.x >> 6;
    if (__lane_id() == 0) {
        sK[wid] = kills;
        sS[wid] = splits;
        sT[wid] = carry;
    }
    __syncthreads();
    if (threadIdx.x != 0) return;
    for (uint32_t w = 1; w < (blockDim.x >> 6); w++) {
        kills += sK[w];
        splits += sS[w];
        carry += sT[w];
    }
    cb->dc.rr += kills;
    cb->dc.split += splits;
    cb->pc.nPopOut = carry < outCap ? (uint32_t)carry : outCap;
    if (carry > outCap) {
        cb->pc.error |= 4u;
        c.hold = 1u;
    }
}

__global__ void __launch_bounds__(256) k_population(DevProblem P, Particles in, Particles out, uint32_t outCap,
                                                    PassCounters* pc, const CycleCtl* __restrict__ ctl,
                                                    const PopGroup* __restrict__ groups) {
    if (ctl->hold) return;
    const uint32_t n = ctl->n, size = pop_group_size(*ctl);
    const uint32_t nGroups = (n + size - 1) / size;
    const int mode = (int)ctl->mode;
    const double factor = ctl->factor;
    for (uint32_t g = blockIdx.x; g < nGroups; g += gridDim.x) {
        const uint32_t i = g * size + threadIdx.x;
        const bool valid = threadIdx.x < size && i < n;
        const PopDecision d = pop_decide(P, in, i, valid, mode, factor);
        const uint32_t surv = (d.keep && valid) ? d.survivors : 0u;
        uint32_t total;
        const uint32_t excl = block_excl_scan(surv, total);
        pop_write(P, in, i, d, surv, excl, total, groups[g].base, out, outCap, &pc->error);
    }
}

__global__ void __launch_bounds__(kFluxBlock) k_zero_flux(double* flux, size_t n, const uint32_t* __restrict__ hold) {
    if (*hold) return;
    zero_doubles(flux, 0, n, (size_t)blockIdx.x * blockDim.x + threadIdx.x, (size_t)gridDim.x * blockDim.x);
}

// Cycle set-up bookkeeping of one problem (main.cc:96-121 and
// PopulationControl.cc:20-65, as the host would do it): the census count
// carried in (start), the population-control factor and mode, the pass and
// balance counters cleared, hold cleared.  Called by every thread of one
// block, after any read of the counters.
__device__ __forceinline__ void cycle_setup(CycleBlock* cb, uint32_t start, uint32_t procCap, uint64_t nSource,
                                            uint64_t target, int loadBalance) {
    uint32_t* a = reinterpret_cast<uint32_t*>(&cb->pc);
    uint32_t* b = reinterpret_cast<uint32_t*>(&cb->dc);
    for (uint32_t q = threadIdx.x; q < sizeof(PassCounters) / 4; q += blockDim.x) a[q] = 0u;
    for (uint32_t q = threadIdx.x; q < sizeof(DevCounters) / 4; q += blockDim.x) b[q] = 0u;
    __syncthreads();
    if (threadIdx.x != 0) return;
    const uint64_t n64 = (uint64_t)start + nSource;
    const bool fits = n64 <= (uint64_t)procCap;
    const uint32_t n = fits ? (uint32_t)n64 : 0u;
    const double tgt = (double)target;
    double factor;
    if (loadBalance) factor = (n != 0) ? tgt / (double)n : 1.0;
    else factor = tgt / (double)n;
    uint32_t mode = 0;
    if (factor != 1.0 && n != 0) mode = (factor < 1) ? 1u : ((factor > 1) ? 2u : 0u);
    CycleCtl& c = cb->ctl;
    c.start = start;
    c.n = n;
    c.mode = mode;
    c.factor = factor;
    if (!fits) cb->pc.error = 32u;
    c.hold = fits ? 0u : 1u;
}

__global__ void k_cycle_setup(CycleBlock* cb, uint32_t start, uint32_t procCap, uint64_t nSource, uint64_t target,
                              int loadBalance) {
    cycle_setup(cb, start, procCap, nSource, target, loadBalance);
}

// The end of a cycle's first pass when cycles are enqueued ahead: hold is
// set if the pass left secondaries or an error; every counter is copied to
// the pinned host block `host`; then, unless held or setupNext is 0, the
// next cycle's set-up from the census count the pass left.  One block.
__global__ void k_cycle_tail(CycleBlock* cb, CycleBlock* host, int setupNext, uint32_t procCap, uint64_t nSource,
                             uint64_t target, int loadBalance) {
    __shared__ uint32_t sHold, sStart;
    if (threadIdx.x == 0) {
        uint32_t h = cb->ctl.hold;
        if (!h && (cb->pc.extraCount[0] != 0 || cb->pc.error != 0)) {
            h = 1u;
            cb->ctl.hold = 1u;
        }
        sHold = h;
        sStart = cb->pc.nCensus;
    }
    __syncthreads();
    const uint32_t* src = reinterpret_cast<const uint32_t*>(cb);
    uint32_t* dst = reinterpret_cast<uint32_t*>(host);
    for (uint32_t q = threadIdx.x; q < sizeof(CycleBlock) / 4; q += blockDim.x) dst[q] = src[q];
    __threadfence_system();
    __syncthreads();  // the counters are read before they are cleared
    if (sHold || !setupNext) return;
    cycle_setup(cb, sStart, procCap, nSource, target, loadBalance);
}

__global__ void k_clear(uint32_t* p, size_t words) {
    size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; q < words; q += (size_t)gridDim.x * blockDim.x) p[q] = 0u;
}

__global__ void k_math(int op, const double* a, const double* b, double* out, size_t n) {
    size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    double x = a[q], y = b[q];
    double r = 0;
    switch (op) {
        case 0: r = sqrt(x); break;
        case 1: r = x / y; break;
        case 2: r = log(x); break;
        case 3: r = sin(x); break;
        case 4: r = cos(x); break;
        case 5: r = x * y; break;
        case 6: r = (double)(*(const uint64_t*)&a[q]); break;
        case 7: {
            double c;
            sincos(x, &r, &c);
            break;
        }
        case 8: {
            double sn;
            sincos(x, &sn, &r);
            break;
        }
        default: r = 0;
    }
    out[q] = r;
}

// The tracking kernels' nearest-facet search on given rays, one ray per
// thread: the same device function, so its exit-face shortcut can be checked
// ray by ray against the reference's full 24-facet scan.
__global__ void __launch_bounds__(256) k_nearest_probe(DevProblem P, const double* __restrict__ xyz,
                                                       const double* __restrict__ dir,
                                                       const int32_t* __restrict__ cell, double* __restrict__ dist,
                                                       int32_t* __restrict__ facet, double* __restrict__ xyzOut,
                                                       size_t n) {
    load_facet_codes();
    __syncthreads();
    const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const int c = cell[q];
    const int ci = c % P.nx, cj = (c / P.nx) % P.ny, ck = c / (P.nx * P.ny);
    double x = xyz[3 * q], y = xyz[3 * q + 1], z = xyz[3 * q + 2];
    const double a = dir[3 * q], b = dir[3 * q + 1], g = dir[3 * q + 2];
    const Cell9 c9 = cell_nodes(P, ci, cj, ck);
    const Nearest nf = nearest_facet(P, P.planes + (size_t)c * 96, c, c9, x, y, z, a, b, g);
    dist[q] = nf.d;
    facet[q] = nf.f;
    xyzOut[3 * q] = x;
    xyzOut[3 * q + 1] = y;
    xyzOut[3 * q + 2] = z;
}

// Fluence::compute (Tallies.cc:100-120): per cell, the groups' flux added
// one by one to the cumulative value.
__global__ void __launch_bounds__(256) k_fluence(const double* __restrict__ flux, double* __restrict__ fluence,
                                                 int nCells, int G, const uint32_t* __restrict__ hold) {
    if (*hold) return;
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < nCells; c += gridDim.x * blockDim.x) {
        double v = fluence[c];
        const double* f = flux + (size_t)c * G;
        for (int g = 0; g < G; g++) v += f[g];
        fluence[c] = v;
    }
}

// ------------------------------------------------------- particle sort ----
// Counting sort of a tracking queue by (cell, energy band): histogram of
// keys, exclusive scan, scatter.  The band of a particle is its energy group
// (the same search MC_Load_Particle does) scaled to kSortBands.

__device__ __forceinline__ uint32_t sort_key(const DevProblem& P, const Particles& in, uint32_t i) {
    const int g = energy_group(P.energies, P.G, in.E()[i], P.energies[0], P.energies[P.G], P.eLog2Lo,
                               P.eGroupsPerLog2);
    return (uint32_t)in.cell()[i] * kSortBands + (uint32_t)(g * kSortBands / P.G);
}

__device__ __forceinline__ void copy_particle(const Particles& in, uint32_t i, const Particles& out, uint32_t o) {
#pragma unroll
    for (int k = 0; k < 10; k++) out.f(k)[o] = in.f(k)[i];
    out.seed()[o] = in.seed()[i];
    out.cell()[o] = in.cell()[i];
}

__global__ void __launch_bounds__(256) k_sort_count(DevProblem P, Particles in, const uint32_t* __restrict__ nPtr,
                                                    uint32_t* __restrict__ counts, const uint32_t* __restrict__ hold) {
    if (*hold) return;
    const uint32_t n = *nPtr;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        atomicAdd(&counts[sort_key(P, in, i)], 1u);
}

__global__ void __launch_bounds__(256) k_sort_scatter(DevProblem P, Particles in, const uint32_t* __restrict__ nPtr,
                                                      uint32_t* __restrict__ offsets, Particles out,
                                                      const uint32_t* __restrict__ hold) {
    if (*hold) return;
    const uint32_t n = *nPtr;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const uint32_t o = atomicAdd(&offsets[sort_key(P, in, i)], 1u);
        copy_particle(in, i, out, o);
    }
}

// Exclusive scan of m u32 values in place: per 1024-element block, then
// the block sums (one workgroup), then the block offsets added back.
__global__ void __launch_bounds__(256) k_scan_local(uint32_t* __restrict__ data, size_t m,
                                                    uint32_t* __restrict__ blockSums) {
    __shared__ uint32_t sh[256];
    const size_t base = (size_t)blockIdx.x * 1024 + threadIdx.x * 4;
    uint32_t v[4];
    uint32_t sum = 0;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        v[k] = base + k < m ? data[base + k] : 0u;
        sum += v[k];
    }
    sh[threadIdx.x] = sum;
    __syncthreads();
    for (int o = 1; o < 256; o <<= 1) {
        uint32_t t = threadIdx.x >= (unsigned)o ? sh[threadIdx.x - o] : 0u;
        __syncthreads();
        sh[threadIdx.x] += t;
        __syncthreads();
    }
    uint32_t run = sh[threadIdx.x] - sum;  // exclusive prefix of this thread
#pragma unroll
    for (int k = 0; k < 4; k++) {
        if (base + k < m) data[base + k] = run;
        run += v[k];
    }
    if (threadIdx.x == 255) blockSums[blockIdx.x] = sh[255];
}

__global__ void __launch_bounds__(1024) k_scan_top(uint32_t* __restrict__ blockSums, int nb) {
    __shared__ uint32_t sh[1024];
    __shared__ uint32_t carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (int c0 = 0; c0 < nb; c0 += 1024) {
        const int q = c0 + (int)threadIdx.x;
        const uint32_t v = q < nb ? blockSums[q] : 0u;
        sh[threadIdx.x] = v;
        __syncthreads();
        for (int o = 1; o < 1024; o <<= 1) {
            uint32_t t = threadIdx.x >= (unsigned)o ? sh[threadIdx.x - o] : 0u;
            __syncthreads();
            sh[threadIdx.x] += t;
            __syncthreads();
        }
        if (q < nb) blockSums[q] = carry + sh[threadIdx.x] - v;
        __syncthreads();
        if (threadIdx.x == 1023) carry += sh[1023];
        __syncthreads();
    }
}

__global__ void __launch_bounds__(256) k_scan_add(uint32_t* __restrict__ data, size_t m,
                                                  const uint32_t* __restrict__ blockSums) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += (size_t)gridDim.x * blockDim.x)
        data[i] += blockSums[i / 1024];
}

// ------------------------------------------------------ batched form ----
// Many independent problems advance one cycle together.  Every problem keeps
// its own particle buffers, tallies and RNG streams, so its results are the
// same alone or in any mix.

// Per-problem cycle set-up (the bookkeeping of main.cc:96-121 and
// PopulationControl.cc:20-65): carried census count, source count,
// population-control factor and counter resets; plus the exclusive prefix
// sums over problems that map a flat thread index to (problem, particle).
__global__ void k_batch_prep(BatchSlot* slots, int n, uint64_t* srcPrefix, uint64_t* popPrefix) {
    for (int p = threadIdx.x; p < n; p += blockDim.x) {
        BatchSlot& S = slots[p];
        BatchState* st = S.st;
        DevCounters* dc = S.dc;
        const bool act = st->active != 0;
        const uint32_t start = act ? st->nCensus : 0;
        // census + source must fit the processing buffer; a problem that
        // does not sources and tracks nothing and reports the overflow
        const bool fits = (uint64_t)start + S.sourcePerCycle <= (uint64_t)S.cap;
        st->start = start;
        st->nProc = (act && fits) ? start + (uint32_t)S.sourcePerCycle : 0;
        const uint32_t nProc = st->nProc;
        const double target = (double)S.nParticles;
        double factor;
        if (S.loadBalance) factor = (nProc != 0) ? target / (double)nProc : 1.0;
        else factor = target / (double)nProc;
        uint32_t mode = 0;
        if (factor != 1.0 && nProc != 0) mode = (factor < 1) ? 1u : ((factor > 1) ? 2u : 0u);
        st->factor = factor;
        st->mode = mode;
        st->sorted = 0;
        st->head = 0;
        st->nPop = 0;
        st->extra[0] = 0;
        st->extra[1] = 0;
        st->error = (act && !fits) ? 32u : 0u;
        if (act) {
            st->nCensus = 0;
            unsigned long long* c = (unsigned long long*)dc;
            for (int q = 0; q < (int)(sizeof(DevCounters) / 8); q++) c[q] = 0ull;
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint64_t a = 0, b = 0;
        for (int p = 0; p < n; p++) {
            srcPrefix[p] = a;
            popPrefix[p] = b;
            if (slots[p].st->active && slots[p].st->error == 0) a += slots[p].sourcePerCycle;
            b += ((uint64_t)slots[p].st->nProc + 255) / 256;  // population groups of 256 inputs
        }
        srcPrefix[n] = a;
        popPrefix[n] = b;
    }
}

__device__ __forceinline__ int find_problem(const uint64_t* prefix, int n, uint64_t t) {
    int lo = 0, hi = n;  // prefix[lo] <= t < prefix[hi]
    while (hi - lo > 1) {
        int mid = (lo + hi) >> 1;
        if (prefix[mid] <= t) lo = mid;
        else hi = mid;
    }
    return lo;
}

__global__ void __launch_bounds__(256) k_batch_source(BatchSlot* slots, int n, const uint64_t* srcPrefix) {
    load_facet_codes();
    __syncthreads();
    const uint64_t total = srcPrefix[n];
    for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total;
         t += (uint64_t)gridDim.x * blockDim.x) {
        const int p = find_problem(srcPrefix, n, t);
        const BatchSlot& S = slots[p];
        const BatchState* st = S.st;
        const uint64_t local = t - srcPrefix[p];
        source_particle(S.prob, S.proc[st->cycle & 1], st->start + (uint32_t)local, local, st->cycle);
    }
}

// Population control of every problem of a batch, in the same three steps
// as the one-problem form (no atomics on a problem's counters): popPrefix
// counts each problem's groups of 256 consecutive inputs, and blocks stride
// over the groups of all problems, one group at a time; k_batch_pop_scan
// runs one workgroup per problem.
__global__ void __launch_bounds__(256) k_batch_pop_count(BatchSlot* slots, int n, const uint64_t* popPrefix) {
    __shared__ uint32_t sSum[3][4];
    const uint64_t total = popPrefix[n];
    const uint32_t wid = threadIdx.x >> 6;
    for (uint64_t g = blockIdx.x; g < total; g += gridDim.x) {
        const int p = find_problem(popPrefix, n, g);
        const BatchSlot& S = slots[p];
        const BatchState* st = S.st;
        const uint32_t lg = (uint32_t)(g - popPrefix[p]);
        const uint32_t i = lg * 256u + threadIdx.x;
        const bool valid = i < st->nProc;
        const PopDecision d = pop_decide(S.prob, S.proc[st->cycle & 1], i, valid, (int)st->mode, st->factor);
        const uint32_t surv = wave_sum((d.keep && valid) ? d.survivors : 0u);
        const uint32_t kills = wave_sum(d.kills), splits = wave_sum(d.splits);
        if (__lane_id() == 0) {
            sSum[0][wid] = surv;
            sSum[1][wid] = kills;
            sSum[2][wid] = splits;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            PopGroup r = {0u, 0u, 0u, 0u};
            for (uint32_t w = 0; w < (blockDim.x >> 6); w++) {
                r.total += sSum[0][w];
                r.kills += sSum[1][w];
                r.splits += sSum[2][w];
            }
            S.groups[lg] = r;
        }
        __syncthreads();
    }
}

// Block p: the scan of problem p's groups into output bases, its output
// count into nPop (at most its capacity; more is error bit 4), and its
// roulette and split counts into its balance.
__global__ void __launch_bounds__(1024) k_batch_pop_scan(BatchSlot* slots, const uint64_t* popPrefix) {
    __shared__ unsigned long long sK[16], sS[16], sT[16];
    const int p = blockIdx.x;
    const BatchSlot& S = slots[p];
    BatchState* st = S.st;
    const uint32_t nGroups = (uint32_t)(popPrefix[p + 1] - popPrefix[p]);
    if (nGroups == 0) return;
    PopGroup* groups = S.groups;
    const uint32_t per = (nGroups + blockDim.x - 1) / blockDim.x;
    const uint32_t g0 = min(threadIdx.x * per, nGroups), g1 = min(g0 + per, nGroups);
    uint64_t mine = 0;
    unsigned long long kills = 0, splits = 0;
    for (uint32_t g = g0; g < g1; g++) {
        const PopGroup r = groups[g];
        mine += r.total;
        kills += r.kills;
        splits += r.splits;
    }
    uint32_t tot;
    const uint32_t excl = block_excl_scan(mine > 0xFFFFFFFFull ? 0xFFFFFFFFu : (uint32_t)mine, tot);
    uint64_t run = excl;
    for (uint32_t g = g0; g < g1; g++) {
        groups[g].base = (uint32_t)run;
        run += groups[g].total;
    }
    unsigned long long carry = wave_sum64(mine);
    kills = wave_sum64(kills);
    splits = wave_sum64(splits);
    const uint32_t wid = threadIdx.x >> 6;
    if (__lane_id() == 0) {
        sK[wid] = kills;
        sS[wid] = splits;
        sT[wid] = carry;
    }
    __syncthreads();
    if (threadIdx.x != 0) return;
    for (uint32_t w = 1; w < (blockDim.x >> 6); w++) {
        kills += sK[w];
        splits += sS[w];
        carry += sT[w];
    }
    S.dc->rr += kills;
    S.dc->split += splits;
    st->nPop = carry < S.cap ? (uint32_t)carry : S.cap;
    if (carry > S.cap) st->error |= 4u;
}

__global__ void __launch_bounds__(256) k_batch_population(BatchSlot* slots, int n, const uint64_t* popPrefix) {
    const uint64_t total = popPrefix[n];
    for (uint64_t g = blockIdx.x; g < total; g += gridDim.x) {
        const int p = find_problem(popPrefix, n, g);
        const BatchSlot& S = slots[p];
        BatchState* st = S.st;
        const uint32_t lg = (uint32_t)(g - popPrefix[p]);
        const uint32_t i = lg * 256u + threadIdx.x;
        const bool valid = i < st->nProc;
        const Particles in = S.proc[st->cycle & 1];
        const PopDecision d = pop_decide(S.prob, in, i, valid, (int)st->mode, st->factor);
        const uint32_t surv = (d.keep && valid) ? d.survivors : 0u;
        uint32_t tot;
        const uint32_t excl = block_excl_scan(surv, tot);
        pop_write(S.prob, in, i, d, surv, excl, tot, S.groups[lg].base, S.pop, S.cap, &st->error);
    }
}

__global__ void k_batch_pass_prep(BatchSlot* slots, int n, int passKind) {
    for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < n; p += gridDim.x * blockDim.x) {
        BatchState* st = slots[p].st;
        st->head = 0;
        st->extra[passKind == 1 ? 1 : 0] = 0;
    }
}

// Batched sort: which problems sort this cycle, and the prefix of their
// queue lengths (a problem's keys are contiguous in the global key space,
// so the scan of its keys starts at the sum of the earlier sorted queues).
__global__ void k_batch_sort_prep(BatchSlot* slots, int n, int interval, uint64_t* sortPrefix) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    uint64_t a = 0;
    for (int p = 0; p < n; p++) {
        BatchState* st = slots[p].st;
        const bool due = st->active && st->error == 0 && interval > 0 && (st->cycle % (uint32_t)interval) == 0;
        st->sorted = due ? 1u : 0u;
        sortPrefix[p] = a;
        if (due) a += st->nPop;
    }
    sortPrefix[n] = a;
}

__global__ void __launch_bounds__(256) k_batch_sort_count(BatchSlot* slots, int n, const uint64_t* sortPrefix,
                                                          const uint64_t* keyBase, uint32_t* counts) {
    const uint64_t total = sortPrefix[n];
    for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total;
         t += (uint64_t)gridDim.x * blockDim.x) {
        const int p = find_problem(sortPrefix, n, t);
        const BatchSlot& S = slots[p];
        const uint32_t i = (uint32_t)(t - sortPrefix[p]);
        atomicAdd(&counts[keyBase[p] + sort_key(S.prob, S.pop, i)], 1u);
    }
}

__global__ void __launch_bounds__(256) k_batch_sort_scatter(BatchSlot* slots, int n, const uint64_t* sortPrefix,
                                                            const uint64_t* keyBase, uint32_t* offsets) {
    const uint64_t total = sortPrefix[n];
    for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total;
         t += (uint64_t)gridDim.x * blockDim.x) {
        const int p = find_problem(sortPrefix, n, t);
        const BatchSlot& S = slots[p];
        const uint32_t i = (uint32_t)(t - sortPrefix[p]);
        const uint32_t o = atomicAdd(&offsets[keyBase[p] + sort_key(S.prob, S.pop, i)], 1u) -
                           (uint32_t)sortPrefix[p];
        copy_particle(S.pop, i, S.proc[S.st->cycle & 1], o);
    }
}

// One tracking pass over every problem.  Each wave starts at its own
// problem and moves round-robin through the problems, draining one
// problem's queue at a time (so its descriptor and tables stay
// wave-uniform); it exits after a full round finds no queued particle.
template <int kMinWaves, bool kStageE>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(kMinWaves)))
k_batch_track(BatchSlot* slots, int n, int passKind, TrackTuning tune, WaveStacks ws) {
    load_facet_codes();
    __syncthreads();
    const uint32_t wave = __builtin_amdgcn_readfirstlane((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
    __shared__ MatInfo sMat[4][kMaxMaterials];
    __shared__ int32_t sReact[4][kMaxMaterials * kMaxReactions];
    __shared__ double sE[4][kStageE ? kBatchStagedEnergies : 1];
    BatchQueue<kStageE> acc = {slots, n, passKind, (int)(wave % (uint32_t)n), sMat[threadIdx.x >> 6],
                               sReact[threadIdx.x >> 6], sE[threadIdx.x >> 6]};
    if (!slots[acc.p].st->active && !acc.next()) return;
    run_queue<(kMinWaves <= 2)>(acc, tune, ws);
}

// Fluence of every active problem that accumulates it (one block row per
// problem, grid-stride over its cells).
__global__ void __launch_bounds__(256) k_batch_fluence(BatchSlot* slots) {
    const BatchSlot& S = slots[blockIdx.y];
    if (!S.st->active || !S.prob.fluence) return;
    const int G = S.prob.G;
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < S.prob.nCells; c += gridDim.x * blockDim.x) {
        double v = S.prob.fluence[c];
        const double* f = S.prob.flux + (size_t)c * G;
        for (int g = 0; g < G; g++) v += f[g];
        S.prob.fluence[c] = v;
    }
}

// Flux sum per problem in a fixed order (Tallies.cc:146-171), flux reset
// (Tallies.cc:95), and the cycle advance of every active problem.
__global__ void __launch_bounds__(kFluxBlock) k_batch_flux_zero(BatchSlot* slots, int partsPerProblem) {
    const int p = blockIdx.x / partsPerProblem, part = blockIdx.x % partsPerProblem;
    const BatchSlot& S = slots[p];
    if (!S.st->active) return;
    const size_t nf = (size_t)S.prob.nCells * S.prob.G;
    const size_t per = ((nf + partsPerProblem - 1) / partsPerProblem + 1) & ~(size_t)1;  // even slices
    const size_t lo = (size_t)part * per;
    if (lo >= nf) return;
    const size_t hi = lo + per < nf ? lo + per : nf;
    zero_doubles(S.prob.flux, lo, hi, threadIdx.x, blockDim.x);
}

// Each active problem's flux sum (accumulated by the tracking kernel) and
// its cycle advance.
__global__ void k_batch_flux_final(BatchSlot* slots, int n, double* fluxSums) {
    for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < n; p += gridDim.x * blockDim.x) {
        BatchState* st = slots[p].st;
        if (!st->active) continue;
        fluxSums[p] = slots[p].dc->fluxSum;
        st->cycle++;
    }
}

}  // namespace

size_t tableLdsBytes(const DevProblem& P) {
    return sizeof(double) * ((size_t)(P.G + 1) + 2 * (size_t)P.nMat * P.G + (size_t)P.nMat * P.maxR * P.G);
}

void launchCycleSetup(const LaunchCtx& ctx, CycleBlock* cb, uint32_t start, uint32_t procCap, uint64_t nSource,
                      uint64_t target, int loadBalance) {
    hipLaunchKernelGGL(k_cycle_setup, dim3(1), dim3(64), 0, (hipStream_t)ctx.stream, cb, start, procCap, nSource,
                       target, loadBalance);
}

void launchCycleTail(const LaunchCtx& ctx, CycleBlock* cb, CycleBlock* host, int setupNext, uint32_t procCap,
                     uint64_t nSource, uint64_t target, int loadBalance) {
    hipLaunchKernelGGL(k_cycle_tail, dim3(1), dim3(64), 0, (hipStream_t)ctx.stream, cb, host, setupNext, procCap,
                       nSource, target, loadBalance);
}

void launchSource(const LaunchCtx& ctx, const DevProblem& P, Particles out, const CycleCtl* ctl,
                  uint64_t nSource, uint32_t cycle, double* flux, size_t nFlux) {
    hipStream_t s = (hipStream_t)ctx.stream;
    const unsigned srcBlocks = (unsigned)((nSource + 255) / 256);
    unsigned zeroBlocks = 0;
    if (flux && nFlux) {
        size_t z = (nFlux / 2 + 255) / 256;
        zeroBlocks = (unsigned)(z > 4096 ? 4096 : (z == 0 ? 1 : z));
    }
    if (srcBlocks + zeroBlocks == 0) return;
    hipLaunchKernelGGL(k_source, dim3(srcBlocks + zeroBlocks), dim3(256), 0, s, P, out, ctl, nSource, cycle, srcBlocks,
                       flux, nFlux);
}

void launchPopulation(const LaunchCtx& ctx, const DevProblem& P, Particles in, uint32_t inCap,
                      Particles out, uint32_t outCap, CycleBlock* cb, PopGroup* groups) {
    if (inCap == 0) return;
    hipStream_t s = (hipStream_t)ctx.stream;
    unsigned blocks = (inCap + kPopMinGroup - 1) / kPopMinGroup;
    const unsigned maxBlocks = (unsigned)ctx.numCUs * 8;
    if (blocks > maxBlocks) blocks = maxBlocks;
    hipLaunchKernelGGL(k_pop_count, dim3(blocks), dim3(256), 0, s, P, in, &cb->ctl, groups);
    hipLaunchKernelGGL(k_pop_scan, dim3(1), dim3(1024), 0, s, cb, groups, outCap);
    hipLaunchKernelGGL(k_population, dim3(blocks), dim3(256), 0, s, P, in, out, outCap, &cb->pc, &cb->ctl, groups);
}

size_t popGroups(uint32_t inCap) { return (inCap + kPopMinGroup - 1) / kPopMinGroup + 1; }

constexpr size_t kLdsLimit = 62 * 1024;  // 64 KiB per workgroup less the static tables

bool tablesFitLds(const DevProblem& P) { return tableLdsBytes(P) <= kLdsLimit; }

long long trackPassBlocks(const LaunchCtx& ctx, const TrackTuning& tune, uint32_t nIn) {
    const int block = tune.blockSize;
    const int wavesPerBlock = block / 64;
    long long persistentBlocks = (long long)ctx.numCUs * tune.wavesPerCU / wavesPerBlock;
    long long needed = ((long long)nIn + block - 1) / block;
    long long blocks = needed < persistentBlocks ? needed : persistentBlocks;
    return blocks < 1 ? 1 : blocks;
}

long long batchTrackBlocks(const LaunchCtx& ctx, const TrackTuning& tune) {
    long long blocks = (long long)ctx.numCUs * tune.wavesPerCU / (tune.blockSize / 64);
    return blocks < 1 ? 1 : blocks;
}

void launchTrackPass(const LaunchCtx& ctx, const DevProblem& P, const TrackTuning& tune, const WaveStacks& ws,
                     Particles in, const uint32_t* nIn, uint32_t nInHost, Particles census,
                     uint32_t censusCap, Particles extra, uint32_t* nExtraOut, uint32_t extraCap,
                     PassCounters* pc, DevCounters* dc, const uint32_t* hold) {
    hipStream_t s = (hipStream_t)ctx.stream;
    const int block = tune.blockSize;
    const long long blocks = trackPassBlocks(ctx, tune, nInHost);
    const size_t tableBytes = tableLdsBytes(P);
    const bool inLds = tableBytes <= kLdsLimit;
    const size_t lds = inLds ? tableBytes : 0;
    decltype(&k_track<true, 3>) kern;
    switch (tune.wavesPerSimd) {
        case 2: kern = inLds ? k_track<true, 2> : k_track<false, 2>; break;
        case 4: kern = inLds ? k_track<true, 4> : k_track<false, 4>; break;
        default: kern = inLds ? k_track<true, 3> : k_track<false, 3>; break;
    }
    hipLaunchKernelGGL(kern, dim3((unsigned)blocks), dim3(block), lds, s, P, tune, ws, in, nIn, census, censusCap,
                       extra, nExtraOut, extraCap, pc, dc, hold);
}

void launchZeroFlux(const LaunchCtx& ctx, double* flux, size_t n, const uint32_t* hold) {
    if (n == 0) return;
    size_t blocks = (n / 2 + kFluxBlock - 1) / kFluxBlock;
    if (blocks > 4096) blocks = 4096;
    if (blocks == 0) blocks = 1;
    hipLaunchKernelGGL(k_zero_flux, dim3((unsigned)blocks), dim3(kFluxBlock), 0, (hipStream_t)ctx.stream, flux, n,
                       hold);
}

void launchClear(const LaunchCtx& ctx, void* ptr, size_t bytes) {
    hipStream_t s = (hipStream_t)ctx.stream;
    size_t words = bytes / 4;
    unsigned blocks = (unsigned)((words + 255) / 256);
    if (blocks > 1024) blocks = 1024;
    if (blocks == 0) return;
    hipLaunchKernelGGL(k_clear, dim3(blocks), dim3(256), 0, s, (uint32_t*)ptr, words);
}

void launchBatchPrep(const LaunchCtx& ctx, BatchSlot* slots, int n, uint64_t* srcPrefix, uint64_t* popPrefix) {
    hipLaunchKernelGGL(k_batch_prep, dim3(1), dim3(1024), 0, (hipStream_t)ctx.stream, slots, n, srcPrefix, popPrefix);
}

void launchBatchSource(const LaunchCtx& ctx, BatchSlot* slots, int n, const uint64_t* srcPrefix) {
    hipStream_t s = (hipStream_t)ctx.stream;
    hipLaunchKernelGGL(k_batch_source, dim3(ctx.numCUs * 8), dim3(256), 0, s, slots, n, srcPrefix);
}

void launchBatchPopulation(const LaunchCtx& ctx, BatchSlot* slots, int n, const uint64_t* popPrefix) {
    hipStream_t s = (hipStream_t)ctx.stream;
    hipLaunchKernelGGL(k_batch_pop_count, dim3(ctx.numCUs * 8), dim3(256), 0, s, slots, n, popPrefix);
    hipLaunchKernelGGL(k_batch_pop_scan, dim3((unsigned)n), dim3(1024), 0, s, slots, popPrefix);
    hipLaunchKernelGGL(k_batch_population, dim3(ctx.numCUs * 8), dim3(256), 0, s, slots, n, popPrefix);
}

void launchBatchPassPrep(const LaunchCtx& ctx, BatchSlot* slots, int n, int passKind) {
    unsigned blocks = (unsigned)((n + 255) / 256);
    hipLaunchKernelGGL(k_batch_pass_prep, dim3(blocks), dim3(256), 0, (hipStream_t)ctx.stream, slots, n, passKind);
}

void launchBatchTrack(const LaunchCtx& ctx, BatchSlot* slots, int n, int passKind, const TrackTuning& tune,
                      const WaveStacks& ws, bool stagedEnergies) {
    decltype(&k_batch_track<3, true>) kern;
    switch (tune.wavesPerSimd) {
        case 2: kern = stagedEnergies ? k_batch_track<2, true> : k_batch_track<2, false>; break;
        case 4: kern = stagedEnergies ? k_batch_track<4, true> : k_batch_track<4, false>; break;
        default: kern = stagedEnergies ? k_batch_track<3, true> : k_batch_track<3, false>; break;
    }
    hipLaunchKernelGGL(kern, dim3((unsigned)batchTrackBlocks(ctx, tune)), dim3(tune.blockSize), 0,
                       (hipStream_t)ctx.stream, slots, n, passKind, tune, ws);
}

void launchBatchFinalize(const LaunchCtx& ctx, BatchSlot* slots, int n, int partsPerProblem, double* fluxSums) {
    hipStream_t s = (hipStream_t)ctx.stream;
    hipLaunchKernelGGL(k_batch_flux_zero, dim3((unsigned)(n * partsPerProblem)), dim3(kFluxBlock), 0, s, slots,
                       partsPerProblem);
    unsigned blocks = (unsigned)((n + 255) / 256);
    hipLaunchKernelGGL(k_batch_flux_final, dim3(blocks), dim3(256), 0, s, slots, n, fluxSums);
}

namespace {
void launchScan(hipStream_t s, uint32_t* data, size_t m, uint32_t* blockSums) {
    const size_t nb = scanBlocks(m);
    if (nb == 0) return;
    hipLaunchKernelGGL(k_scan_local, dim3((unsigned)nb), dim3(256), 0, s, data, m, blockSums);
    hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(1024), 0, s, blockSums, (int)nb);
    unsigned blocks = (unsigned)((m + 255) / 256);
    if (blocks > 8192) blocks = 8192;
    hipLaunchKernelGGL(k_scan_add, dim3(blocks), dim3(256), 0, s, data, m, blockSums);
}
}  // namespace

void launchSort(const LaunchCtx& ctx, const DevProblem& P, Particles in, const uint32_t* n, Particles out,
                uint32_t* counts, uint32_t* blockSums, const uint32_t* hold) {
    hipStream_t s = (hipStream_t)ctx.stream;
    const size_t keys = sortKeys(P.nCells);
    launchClear(ctx, counts, keys * sizeof(uint32_t));
    const unsigned grid = (unsigned)ctx.numCUs * 8;
    hipLaunchKernelGGL(k_sort_count, dim3(grid), dim3(256), 0, s, P, in, n, counts, hold);
    launchScan(s, counts, keys, blockSums);
    hipLaunchKernelGGL(k_sort_scatter, dim3(grid), dim3(256), 0, s, P, in, n, counts, out, hold);
}

void launchBatchSort(const LaunchCtx& ctx, BatchSlot* slots, int n, int interval, const uint64_t* keyBase,
                     uint64_t* sortPrefix, uint32_t* counts, size_t totalKeys, uint32_t* blockSums) {
    hipStream_t s = (hipStream_t)ctx.stream;
    launchClear(ctx, counts, totalKeys * sizeof(uint32_t));
    hipLaunchKernelGGL(k_batch_sort_prep, dim3(1), dim3(64), 0, s, slots, n, interval, sortPrefix);
    const unsigned grid = (unsigned)ctx.numCUs * 8;
    hipLaunchKernelGGL(k_batch_sort_count, dim3(grid), dim3(256), 0, s, slots, n, sortPrefix, keyBase, counts);
    launchScan(s, counts, totalKeys, blockSums);
    hipLaunchKernelGGL(k_batch_sort_scatter, dim3(grid), dim3(256), 0, s, slots, n, sortPrefix, keyBase, counts);
}

void launchBatchFluence(const LaunchCtx& ctx, BatchSlot* slots, int n) {
    hipLaunchKernelGGL(k_batch_fluence, dim3(64, (unsigned)n), dim3(256), 0, (hipStream_t)ctx.stream, slots);
}

void launchFluence(const LaunchCtx& ctx, const DevProblem& P, const uint32_t* hold) {
    if (!P.fluence) return;
    unsigned blocks = (unsigned)((P.nCells + 255) / 256);
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(k_fluence, dim3(blocks), dim3(256), 0, (hipStream_t)ctx.stream, P.flux, P.fluence, P.nCells,
                       P.G, hold);
}

void launchNearestProbe(const LaunchCtx& ctx, const DevProblem& P, const double* xyz, const double* dir,
                        const int32_t* cell, double* dist, int32_t* facet, double* xyzOut, size_t n) {
    const unsigned blocks = (unsigned)((n + 255) / 256);
    if (blocks == 0) return;
    hipLaunchKernelGGL(k_nearest_probe, dim3(blocks), dim3(256), 0, (hipStream_t)ctx.stream, P, xyz, dir, cell,
                       dist, facet, xyzOut, n);
}

void launchMathProbe(const LaunchCtx& ctx, int op, const double* a, const double* b, double* out,
                     size_t n) {
    hipStream_t s = (hipStream_t)ctx.stream;
    unsigned blocks = (unsigned)((n + 255) / 256);
    if (blocks == 0) return;
    hipLaunchKernelGGL(k_math, dim3(blocks), dim3(256), 0, s, op, a, b, out, n);
}

}  // namespace qs
