/*
 * qs_oracle.c -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C, single-threaded restatement of the reference's per-cycle
 * particle-tracking path, used by the tests as a checker and by bench.py as a
 * CPU baseline ("port").  The product never links, loads or runs this file.
 *
 * Parity status of each part (what pins it):
 *   deck grammar / defaults     Parameters.cc:80-509, parseUtils.cc     pinned by the
 *                               full-run fixtures (tests/golden/<deck>.json)
 *   energy grid, reaction XS    NuclearData.cc:12-184                    fixtures
 *   mesh nodes / planes / vol   GlobalFccGrid.cc, MC_Domain.cc,          fixtures + closed
 *                               MC_Facet_Geometry.hh                     forms in tests
 *   RNG (LCG, psdes spawn)      MC_RNG_State.{hh,cc}                     fixtures + closed
 *                                                                        form of the LCG
 *   source                      MC_SourceNow.cc:28-177                   fixtures
 *   population control          PopulationControl.cc:20-171              fixtures (rr/split)
 *   tracking                    CycleTracking.cc, MC_Segment_Outcome.cc, fixtures (all 13
 *                               MCT.cc, CollisionEvent.cc,               balance tallies,
 *                               MC_Facet_Crossing_Event.cc, NuclearData  flux sums, census)
 *   tallies                     Tallies.cc:25-171                        fixtures
 * Every balance tally of every cycle of every fixture deck must match the
 * reference bit-exactly (tests/test_oracle.py).
 *
 * Build: oracle/Makefile (target "restatement") -> oracle/_build/qs_oracle
 * (executable) and oracle/_build/libqsoracle.so (ctypes entry points below).
 *
 * Usage: qs_oracle <deck.inp> [cycles]
 * prints "QSORACLE <cycle> <start> <source> <rr> <split> <absorb> <scatter>
 *         <fission> <produce> <collision> <escape> <census> <num_seg> <end>
 *         <flux_sum %.17e> <tracking_seconds>" per cycle.
 */
#include <ctype.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#define MAXN 64
#define MAX_GEO 4096
#define MAX_MAT 64
#define MAX_XS 256

/* ------------------------------------------------------------ deck ---- */
typedef struct { char mat[MAXN]; int shape; /* 1 brick 2 sphere */
    double radius, xc, yc, zc, xmin, ymin, zmin, xmax, ymax, zmax; } Geom;
typedef struct { char name[MAXN]; double mass, total, srcRate, rs, ra, rf; int nIso, nReact;
    char scat[MAXN], absn[MAXN], fis[MAXN]; } MatP;
typedef struct { char name[MAXN]; double a, b, c, d, e, nuBar; } XsP;
typedef struct {
    double dt, lx, ly, lz, eMin, eMax, cutoff;
    uint64_t nParticles;
    int nSteps, nx, ny, nz, nGroups, loadBalance;
    char bc[MAXN];
    Geom geo[MAX_GEO]; int nGeo;
    MatP mat[MAX_MAT]; int nMat;
    XsP xs[MAX_XS]; int nXs;
} Deck;

typedef struct { char key[MAXN]; char val[256]; } KV;
typedef struct { char name[MAXN]; KV kv[128]; int n; } Block;

static void trim(char* s) {
    size_t n = strlen(s);
    while (n > 0 && isspace((unsigned char)s[n - 1])) s[--n] = 0;
    size_t k = 0;
    while (s[k] && isspace((unsigned char)s[k])) k++;
    memmove(s, s + k, strlen(s + k) + 1);
}

/* parseUtils.cc split(): indentation, key before first ':' or '=' */
static int split_line(const char* line, char* key, char* val, int* indent) {
    int i = 0;
    while (line[i] && isspace((unsigned char)line[i])) i++;
    *indent = i;
    const char* d = strpbrk(line + i, ":=");
    if (!d) return 0;
    size_t kl = (size_t)(d - (line + i));
    if (kl >= MAXN) kl = MAXN - 1;
    memcpy(key, line + i, kl);
    key[kl] = 0;
    trim(key);
    strncpy(val, d + 1, 255);
    val[255] = 0;
    trim(val);
    return 1;
}

static int blank_or_comment(const char* line) {
    char buf[1024];
    strncpy(buf, line, 1023);
    buf[1023] = 0;
    char* c = strstr(buf, "//");
    if (c) *c = 0;
    for (char* p = buf; *p; p++)
        if (!isspace((unsigned char)*p)) return 0;
    return 1;
}

static const char* bget(const Block* b, const char* key) {
    const char* v = NULL;
    for (int i = 0; i < b->n; i++)
        if (!strcmp(b->kv[i].key, key)) v = b->kv[i].val; /* last wins */
    return v;
}
static void gd(const Block* b, const char* k, double* o) { const char* v = bget(b, k); if (v) *o = strtod(v, NULL); }
static void gi(const Block* b, const char* k, int* o) { const char* v = bget(b, k); if (v) *o = (int)strtol(v, NULL, 10); }
static void gu(const Block* b, const char* k, uint64_t* o) { const char* v = bget(b, k); if (v) *o = strtoull(v, NULL, 10); }
static void gs(const Block* b, const char* k, char* o) { const char* v = bget(b, k); if (v) { strncpy(o, v, MAXN - 1); o[MAXN - 1] = 0; } }

static void deck_defaults(Deck* D) {
    memset(D, 0, sizeof(*D));
    D->dt = 1e-8; D->lx = D->ly = D->lz = 100.0; D->eMin = 1e-9; D->eMax = 20; D->cutoff = 0.001;
    D->nParticles = 1000000; D->nSteps = 10; D->nx = D->ny = D->nz = 10; D->nGroups = 230;
    strcpy(D->bc, "reflect");
}

static int find_mat(Deck* D, const char* name) {
    for (int i = 0; i < D->nMat; i++) if (!strcmp(D->mat[i].name, name)) return i;
    return -1;
}
static int find_xs(const Deck* D, const char* name) {
    for (int i = 0; i < D->nXs; i++) if (!strcmp(D->xs[i].name, name)) return i;
    return -1;
}

static void apply_block(Deck* D, const Block* b) {
    if (!strcmp(b->name, "Simulation")) {
        gd(b, "dt", &D->dt); gd(b, "lx", &D->lx); gd(b, "ly", &D->ly); gd(b, "lz", &D->lz);
        gu(b, "nParticles", &D->nParticles); gi(b, "nSteps", &D->nSteps);
        gi(b, "nx", &D->nx); gi(b, "ny", &D->ny); gi(b, "nz", &D->nz);
        gd(b, "eMax", &D->eMax); gd(b, "eMin", &D->eMin); gi(b, "nGroups", &D->nGroups);
        gd(b, "lowWeightCutoff", &D->cutoff); gi(b, "loadBalance", &D->loadBalance);
        gs(b, "boundaryCondition", D->bc);
    } else if (!strcmp(b->name, "Geometry")) {
        Geom* g = &D->geo[D->nGeo++];
        memset(g, 0, sizeof(*g));
        gs(b, "material", g->mat);
        char shape[MAXN] = "";
        gs(b, "shape", shape);
        if (!strcmp(shape, "brick")) {
            g->shape = 1;
            gd(b, "xMax", &g->xmax); gd(b, "xMin", &g->xmin); gd(b, "yMax", &g->ymax);
            gd(b, "yMin", &g->ymin); gd(b, "zMax", &g->zmax); gd(b, "zMin", &g->zmin);
        } else if (!strcmp(shape, "sphere")) {
            g->shape = 2;
            gd(b, "radius", &g->radius); gd(b, "xCenter", &g->xc); gd(b, "yCenter", &g->yc);
            gd(b, "zCenter", &g->zc);
        }
    } else if (!strcmp(b->name, "Material")) {
        char name[MAXN] = "";
        gs(b, "name", name);
        int m = find_mat(D, name);
        if (m < 0) {
            m = D->nMat++;
            MatP* p = &D->mat[m];
            memset(p, 0, sizeof(*p));
            strcpy(p->name, name);
            p->mass = 1000.0; p->total = 1.0; p->nIso = 10; p->nReact = 9;
            p->rs = p->ra = p->rf = 1.0;
        }
        MatP* p = &D->mat[m];
        gd(b, "mass", &p->mass); gs(b, "absorptionCrossSection", p->absn);
        gd(b, "absorptionCrossSectionRatio", &p->ra); gs(b, "fissionCrossSection", p->fis);
        gd(b, "fissionCrossSectionRatio", &p->rf); gi(b, "nIsotopes", &p->nIso);
        gi(b, "nReactions", &p->nReact); gd(b, "totalCrossSection", &p->total);
        gs(b, "scatteringCrossSection", p->scat); gd(b, "scatteringCrossSectionRatio", &p->rs);
        gd(b, "sourceRate", &p->srcRate);
    } else if (!strcmp(b->name, "CrossSection")) {
        char name[MAXN] = "";
        gs(b, "name", name);
        int x = find_xs(D, name);
        if (x < 0) {
            x = D->nXs++;
            XsP* p = &D->xs[x];
            memset(p, 0, sizeof(*p));
            strcpy(p->name, name);
            p->e = 1.0; p->nuBar = 2.4;
        }
        XsP* p = &D->xs[x];
        gd(b, "A", &p->a); gd(b, "B", &p->b); gd(b, "C", &p->c); gd(b, "D", &p->d);
        gd(b, "E", &p->e); gd(b, "nuBar", &p->nuBar);
    }
}

static int cmp_mat(const void* a, const void* b) { return strcmp(((const MatP*)a)->name, ((const MatP*)b)->name); }

int qso_read_deck(const char* path, Deck* D) {
    deck_defaults(D);
    FILE* f = fopen(path, "r");
    if (!f) return 1;
    char line[1024], key[MAXN], val[256];
    int indent, have = fgets(line, sizeof line, f) != NULL;
    Block* b = (Block*)malloc(sizeof(Block));
    while (have) {
        if (split_line(line, key, val, &indent) && indent == 0 && val[0] == 0) {
            memset(b, 0, sizeof(*b));
            strcpy(b->name, key);
            have = 0;
            while (fgets(line, sizeof line, f)) {
                if (blank_or_comment(line)) continue;
                char k2[MAXN], v2[256];
                int ind;
                if (!split_line(line, k2, v2, &ind) || ind == 0) { have = 1; break; }
                if (b->n < 128) { strcpy(b->kv[b->n].key, k2); strcpy(b->kv[b->n].val, v2); b->n++; }
            }
            apply_block(D, b);
        } else {
            have = fgets(line, sizeof line, f) != NULL;
        }
    }
    free(b);
    fclose(f);
    if (D->nGeo == 0) { /* supplyDefaults, Parameters.cc:350-379 */
        XsP* x = &D->xs[D->nXs++];
        memset(x, 0, sizeof(*x)); strcpy(x->name, "flat"); x->e = 1.0; x->nuBar = 2.4;
        MatP* m = &D->mat[D->nMat++];
        memset(m, 0, sizeof(*m)); strcpy(m->name, "sourceMaterial"); m->mass = 1000.0; m->total = 1.0;
        m->nIso = 10; m->nReact = 9; m->srcRate = 1e10; strcpy(m->scat, "flat"); strcpy(m->absn, "flat");
        strcpy(m->fis, "flat"); m->rs = m->ra = 1.0; m->rf = 0.1;
        Geom* g = &D->geo[D->nGeo++];
        memset(g, 0, sizeof(*g)); strcpy(g->mat, "sourceMaterial"); g->shape = 1;
        g->xmax = D->lx; g->ymax = D->ly; g->zmax = D->lz;
    }
    qsort(D->mat, (size_t)D->nMat, sizeof(MatP), cmp_mat); /* std::map order */
    return 0;
}

/* -------------------------------------------------------- problem ---- */
typedef struct { double x, y, z; } V3;

typedef struct {
    int nIso, nReact; double mass, nuBar, af, srcRate;
    int* type; double* xs; /* [nReact][G] micro */ double* total; /* [G] */
} MatT;

typedef struct {
    Deck* D;
    int nx, ny, nz, nc, G;
    double dx, dy, dz;
    int bc[6];
    double* e;        /* [G+1] */
    MatT* mat; int nMat;
    double* plane;    /* [nc][24][4] */
    V3* center;       /* MCT_Cell_Position_3D_G */
    double* vol;
    int* cmat;
    int* nsrc;
    double wsrc;
    double* flux;     /* [nc][G] */
} Prob;

static const int CORNER[14][4] = {{0,0,0,0},{1,0,0,0},{0,1,0,0},{1,1,0,0},{0,0,1,0},{1,0,1,0},{0,1,1,0},
    {1,1,1,0},{1,0,0,1},{0,0,0,1},{0,1,0,2},{0,0,0,2},{0,0,1,3},{0,0,0,3}};
static const int TRI[24][3] = {{1,3,8},{3,7,8},{7,5,8},{5,1,8},{0,4,9},{4,6,9},{6,2,9},{2,0,9},
    {3,2,10},{2,6,10},{6,7,10},{7,3,10},{0,1,11},{1,5,11},{5,4,11},{4,0,11},{4,5,12},{5,7,12},
    {7,6,12},{6,4,12},{0,2,13},{2,3,13},{3,1,13},{1,0,13}};

static V3 node(const Prob* P, int i, int j, int k, int n) {
    const int* o = CORNER[n];
    double rx = (i + o[0]) * P->dx, ry = (j + o[1]) * P->dy, rz = (k + o[2]) * P->dz;
    V3 b = {0.0, 0.0, 0.0};
    if (o[3] == 1) { b.y = P->dy / 2.0; b.z = P->dz / 2.0; }
    if (o[3] == 2) { b.x = P->dx / 2.0; b.z = P->dz / 2.0; }
    if (o[3] == 3) { b.x = P->dx / 2.0; b.y = P->dy / 2.0; }
    V3 r = {rx + b.x, ry + b.y, rz + b.z};
    return r;
}

static double poly(const XsP* c, double x) { return c->e + x * (c->d + x * (c->c + x * (c->b + x * (c->a)))); }

static int inside(const Geom* g, V3 r) {
    if (g->shape == 1)
        return (r.x >= g->xmin && r.x <= g->xmax) && (r.y >= g->ymin && r.y <= g->ymax) &&
               (r.z >= g->zmin && r.z <= g->zmax);
    double x = r.x - g->xc, y = r.y - g->yc, z = r.z - g->zc;
    return sqrt(x * x + y * y + z * z) <= g->radius;
}

int qso_build(Prob* P, Deck* D) {
    memset(P, 0, sizeof(*P));
    P->D = D;
    int G = P->G = D->nGroups;
    P->e = (double*)calloc((size_t)G + 1, sizeof(double));
    P->e[0] = D->eMin; P->e[G] = D->eMax;
    double lo = log(D->eMin), hi = log(D->eMax), delta = (hi - lo) / (G + 1.0);
    for (int i = 1; i < G; i++) P->e[i] = exp(lo + delta * i);

    P->nMat = D->nMat;
    P->mat = (MatT*)calloc((size_t)D->nMat, sizeof(MatT));
    for (int m = 0; m < D->nMat; m++) {
        MatP* mp = &D->mat[m];
        MatT* t = &P->mat[m];
        t->nIso = mp->nIso; t->nReact = mp->nReact; t->mass = mp->mass; t->srcRate = mp->srcRate;
        int fx = find_xs(D, mp->fis);
        if (fx < 0) return 2;
        t->nuBar = D->xs[fx].nuBar;
        t->af = 1.0 / mp->nIso;
        double tw = mp->rf + mp->rs + mp->ra;
        int nF = mp->nReact / 3, nS = mp->nReact / 3, nA = mp->nReact / 3;
        if (mp->nReact % 3 == 1) nS++;
        if (mp->nReact % 3 == 2) { nS++; nF++; }
        double fcs = (mp->total * mp->rf) / (nF * tw);
        double scs = (mp->total * mp->rs) / (nS * tw);
        double acs = (mp->total * mp->ra) / (nA * tw);
        t->type = (int*)calloc((size_t)mp->nReact, sizeof(int));
        t->xs = (double*)calloc((size_t)mp->nReact * G, sizeof(double));
        for (int r = 0; r < mp->nReact; r++) {
            int xi; double rcs;
            if (r % 3 == 0) { t->type[r] = 1; xi = find_xs(D, mp->scat); rcs = scs; }
            else if (r % 3 == 1) { t->type[r] = 3; xi = find_xs(D, mp->fis); rcs = fcs; }
            else { t->type[r] = 2; xi = find_xs(D, mp->absn); rcs = acs; }
            if (xi < 0) return 2;
            double* x = t->xs + (size_t)r * G;
            for (int g = 0; g < G; g++) {
                double en = (P->e[g] + P->e[g + 1]) / 2.0;
                x[g] = pow(10.0, poly(&D->xs[xi], log10(en)));
            }
            double norm = 0.0;
            for (int g = 0; g < G; g++) if (P->e[g + 1] >= 1.) { norm = x[g]; break; }
            double sc = rcs / norm;
            for (int g = 0; g < G; g++) x[g] *= sc;
        }
        t->total = (double*)calloc((size_t)G, sizeof(double));
        for (int g = 0; g < G; g++) {
            double iso = 0.0;
            for (int r = 0; r < mp->nReact; r++) iso += t->xs[(size_t)r * G + g];
            double mac = t->af * 1.0 * iso, s = 0.0;
            for (int i = 0; i < mp->nIso; i++) s += mac;
            t->total[g] = s;
        }
    }

    for (int f = 0; f < 6; f++) {
        if (!strcmp(D->bc, "reflect")) P->bc[f] = 2;
        else if (!strcmp(D->bc, "escape")) P->bc[f] = 1;
        else if (!strcmp(D->bc, "octant")) P->bc[f] = (f % 2 == 0) ? 1 : 2;
        else return 3;
    }
    P->nx = D->nx; P->ny = D->ny; P->nz = D->nz;
    P->dx = D->lx / D->nx; P->dy = D->ly / D->ny; P->dz = D->lz / D->nz;
    int nc = P->nc = D->nx * D->ny * D->nz;
    P->plane = (double*)calloc((size_t)nc * 96, sizeof(double));
    P->center = (V3*)calloc((size_t)nc, sizeof(V3));
    P->vol = (double*)calloc((size_t)nc, sizeof(double));
    P->cmat = (int*)calloc((size_t)nc, sizeof(int));
    P->nsrc = (int*)calloc((size_t)nc, sizeof(int));
    P->flux = (double*)calloc((size_t)nc * G, sizeof(double));
    double* cw = (double*)calloc((size_t)nc, sizeof(double));
    for (int k = 0; k < P->nz; k++) for (int j = 0; j < P->ny; j++) for (int i = 0; i < P->nx; i++) {
        int c = i + P->nx * (j + P->ny * k);
        V3 n[14];
        for (int q = 0; q < 14; q++) n[q] = node(P, i, j, k, q);
        for (int f = 0; f < 24; f++) {
            V3 r0 = n[TRI[f][0]], r1 = n[TRI[f][1]], r2 = n[TRI[f][2]];
            double A = ((r1.y - r0.y) * (r2.z - r0.z)) - ((r1.z - r0.z) * (r2.y - r0.y));
            double B = ((r1.z - r0.z) * (r2.x - r0.x)) - ((r1.x - r0.x) * (r2.z - r0.z));
            double C = ((r1.x - r0.x) * (r2.y - r0.y)) - ((r1.y - r0.y) * (r2.x - r0.x));
            double Dd = -1.0 * (A * r0.x + B * r0.y + C * r0.z);
            double mag = sqrt(A * A + B * B + C * C);
            if (mag == 0.0) { A = 1.0; mag = 1.0; }
            double inv = 1.0 / mag;
            double* p = P->plane + ((size_t)c * 24 + f) * 4;
            p[0] = A * inv; p[1] = B * inv; p[2] = C * inv; p[3] = Dd * inv;
        }
        V3 pos = {0, 0, 0}, cc = {0, 0, 0};
        for (int q = 0; q < 14; q++) { pos.x += n[q].x; pos.y += n[q].y; pos.z += n[q].z; }
        double oo = 1.0 / ((double)14);
        P->center[c].x = pos.x * oo; P->center[c].y = pos.y * oo; P->center[c].z = pos.z * oo;
        for (int q = 0; q < 14; q++) { cc.x += n[q].x; cc.y += n[q].y; cc.z += n[q].z; }
        cc.x /= 14.0; cc.y /= 14.0; cc.z /= 14.0;
        double v = 0;
        for (int f = 0; f < 24; f++) {
            V3 a = n[TRI[f][0]], b = n[TRI[f][1]], d = n[TRI[f][2]];
            double ax = a.x - cc.x, ay = a.y - cc.y, az = a.z - cc.z;
            double bx = b.x - cc.x, by = b.y - cc.y, bz = b.z - cc.z;
            double dx = d.x - cc.x, dy = d.y - cc.y, dz = d.z - cc.z;
            v += fabs(ax * (by * dz - bz * dy) + ay * (bz * dx - bx * dz) + az * (bx * dy - by * dx));
        }
        P->vol[c] = v / 6.0;
        const char* mname = NULL;
        for (int g = 0; g < D->nGeo; g++) if (inside(&D->geo[g], P->center[c])) mname = D->geo[g].mat;
        int m = mname ? find_mat(D, mname) : -1;
        if (m < 0) return 4;
        P->cmat[c] = m;
        cw[c] = P->vol[c] * P->mat[m].srcRate * D->dt;
    }
    double tot = 0;
    for (int c = 0; c < nc; c++) tot += cw[c];
    P->wsrc = tot / (0.1 * D->nParticles);
    for (int c = 0; c < nc; c++) {
        double f = cw[c] / P->wsrc;
        P->nsrc[c] = (f == f && f >= 1.0 && f < 2147483648.0) ? (int)f : 0;
    }
    free(cw);
    return 0;
}

/* ------------------------------------------------------------ RNG ---- */
double qso_rng_sample(uint64_t* s) {
    *s = 2862933555777941757ULL * (*s) + 3037000493ULL;
    return 5.4210108624275222e-20 * (double)(*s);
}

uint64_t qso_hash_state(uint64_t in) {
    uint32_t lw = (uint32_t)(in >> 32), rw = (uint32_t)(in & 0xffffffffu);
    static const uint32_t c1[2] = {0xbaa96887u, 0x1e17d32cu}, c2[2] = {0x4b0f3b58u, 0xe874f0c3u};
    for (int i = 0; i < 2; i++) {
        uint32_t sw = rw, ia = rw ^ c1[i], lo = ia & 0xffffu, hi = ia >> 16;
        uint32_t ib = lo * lo + ~(hi * hi);
        ia = (ib >> 16) | ((ib & 0xffffu) << 16);
        rw = lw ^ ((ia ^ c2[i]) + lo * hi);
        lw = sw;
    }
    return ((uint64_t)lw << 32) | rw;
}

uint64_t qso_rng_spawn(uint64_t* parent) {
    uint64_t c = qso_hash_state(*parent);
    qso_rng_sample(parent);
    return c;
}

/* ------------------------------------------------------- particles ---- */
typedef struct { double x, y, z, vx, vy, vz, E, w, tc, nmfp; uint64_t seed; int cell; } Part;
typedef struct { Part* p; size_t n, cap; } Vec;

static void push(Vec* v, const Part* q) {
    if (v->n == v->cap) { v->cap = v->cap ? 2 * v->cap : 1024; v->p = (Part*)realloc(v->p, v->cap * sizeof(Part)); }
    v->p[v->n++] = *q;
}

typedef struct { uint64_t start, source, rr, split, absorb, scatter, fission, produce, collision, escape,
                          census, num_seg, end; } Bal;

static int egroup(const Prob* P, double E) {
    int n = P->G + 1;
    if (E <= P->e[0]) return 0;
    if (E > P->e[n - 1]) return n - 1;
    int hi = n - 1, lo = 0;
    while (hi != lo + 1) { int mid = (hi + lo) / 2; if (E < P->e[mid]) hi = mid; else lo = mid; }
    return lo;
}

static const double HUGE_D = 1.0e+75, SMALL_D = 1.0e-10, TINY_D = 1.0e-13;
static const double MN = 9.395656981095e+2, CL = 2.99792458e+10, PI_D = 3.1415926535897932;

/* MCT_Nearest_Facet_3D_G_Distance_To_Segment */
static double dist_seg(double tol2, double dot, const double* pl, V3 p0, V3 p1, V3 p2, V3 r, V3 d) {
    const double bb = 1e-9;
    double A = pl[0], B = pl[1], C = pl[2], D = pl[3];
    double num = -1.0 * (A * r.x + B * r.y + C * r.z + D);
    if (num < 0.0 && num * num > tol2) return HUGE_D;
    double t = num / dot;
    V3 q = {r.x + t * d.x, r.y + t * d.y, r.z + t * d.z};
    double c0 = 0, c1 = 0, c2 = 0;
#define BELOW(ax) if (p0.ax > q.ax + bb && p1.ax > q.ax + bb && p2.ax > q.ax + bb) return HUGE_D;
#define ABOVE(ax) if (p0.ax < q.ax - bb && p1.ax < q.ax - bb && p2.ax < q.ax - bb) return HUGE_D;
#define CR(ax_, ay_, bx_, by_, cx_, cy_) (((bx_) - (ax_)) * ((cy_) - (ay_)) - ((by_) - (ay_)) * ((cx_) - (ax_)))
    if (C < -0.5 || C > 0.5) {
        BELOW(x) ABOVE(x) BELOW(y) ABOVE(y)
        c1 = CR(p0.x, p0.y, p1.x, p1.y, q.x, q.y);
        c2 = CR(p1.x, p1.y, p2.x, p2.y, q.x, q.y);
        c0 = CR(p2.x, p2.y, p0.x, p0.y, q.x, q.y);
    } else if (B < -0.5 || B > 0.5) {
        BELOW(x) ABOVE(x) BELOW(z) ABOVE(z)
        c1 = CR(p0.z, p0.x, p1.z, p1.x, q.z, q.x);
        c2 = CR(p1.z, p1.x, p2.z, p2.x, q.z, q.x);
        c0 = CR(p2.z, p2.x, p0.z, p0.x, q.z, q.x);
    } else if (A < -0.5 || A > 0.5) {
        BELOW(z) ABOVE(z) BELOW(y) ABOVE(y)
        c1 = CR(p0.y, p0.z, p1.y, p1.z, q.y, q.z);
        c2 = CR(p1.y, p1.z, p2.y, p2.z, q.y, q.z);
        c0 = CR(p2.y, p2.z, p0.y, p0.z, q.y, q.z);
    }
    double s = c0 + c1 + c2;
    double ct = 1e-9 * (s < 0 ? -s : s);
    if ((c0 > -ct && c1 > -ct && c2 > -ct) || (c0 < ct && c1 < ct && c2 < ct)) return t;
    return HUGE_D;
}

/* MCT_Nearest_Facet (+ 3D_G loop, Find_Nearest, nudge retry).  Returns the
 * distance (clamped at 0) and the facet; may move *r. */
double qso_nearest_facet(const Prob* P, int cell, V3* r, V3 d, int* facet) {
    int i = cell % P->nx, j = (cell / P->nx) % P->ny, k = cell / (P->nx * P->ny);
    V3 n[14];
    for (int q = 0; q < 14; q++) n[q] = node(P, i, j, k, q);
    const double* pl = P->plane + (size_t)cell * 96;
    int iter = 0;
    double mf = 0.5 * SMALL_D;
    double nd;
    int nfac;
    for (;;) {
        double tol2 = 1e-16 * (r->x * r->x + r->y * r->y + r->z * r->z);
        double dist[24];
        for (int f = 0; f < 24; f++) {
            dist[f] = HUGE_D;
            const double* p = pl + 4 * f;
            double dot = p[0] * d.x + p[1] * d.y + p[2] * d.z;
            if (dot <= 0.0) continue;
            dist[f] = dist_seg(tol2, dot, p, n[TRI[f][0]], n[TRI[f][1]], n[TRI[f][2]], *r, d);
        }
        nd = 1e80; nfac = 0;
        double negd = -HUGE_D; int negf = 0;
        for (int f = 0; f < 24; f++) {
            if (dist[f] > 0.0) { if (dist[f] <= nd) { nd = dist[f]; nfac = f; } }
            else if (dist[f] > negd) { negd = dist[f]; negf = f; }
        }
        if (nd == HUGE_D && negd != -HUGE_D) { nd = negd; nfac = negf; }
        if (!(nd == HUGE_D && mf > 0)) break;
        V3 c = P->center[cell];
        r->x += mf * (c.x - r->x); r->y += mf * (c.y - r->y); r->z += mf * (c.z - r->z);
        iter++;
        mf *= 2.0;
        if (mf > 1.0e-2) mf = 1.0e-2;
        if (iter == 10000) break;
    }
    if (nd < 0) nd = 0;
    *facet = nfac;
    return nd;
}

typedef struct { V3 r, v; double a, b, g, E, w, tc, nmfp, sigt; uint64_t seed; int cell, group; } Live;

static void rotate(Live* L, double sT, double cT, double sP, double cP) {
    double ct = L->g, st = sqrt((1.0 - (ct * ct))), cp, sp;
    if (st < 1e-6) { cp = 1.0; sp = 0.0; } else { cp = L->a / st; sp = L->b / st; }
    double na = ct * cp * (sT * cP) - sp * (sT * sP) + st * cp * cT;
    double nb = ct * sp * (sT * cP) + cp * (sT * sP) + st * sp * cT;
    double ng = -st * (sT * cP) + ct * cT;
    L->a = na; L->b = nb; L->g = ng;
}

static void trajectory(Live* L, double energy, double angle) {
    L->E = energy;
    double r = qso_rng_sample(&L->seed);
    double phi = 2 * 3.14159265 * r;
    double sP = sin(phi), cP = cos(phi), sT = sqrt((1.0 - (angle * angle)));
    rotate(L, sT, angle, sP, cP);
    double sp = (CL * sqrt((1.0 - ((MN * MN) / ((energy + MN) * (energy + MN))))));
    L->v.x = sp * L->a; L->v.y = sp * L->b; L->v.z = sp * L->g;
    r = qso_rng_sample(&L->seed);
    L->nmfp = -1.0 * log(r);
}

static Part to_part(const Live* L) {
    Part p = {L->r.x, L->r.y, L->r.z, L->v.x, L->v.y, L->v.z, L->E, L->w, L->tc, L->nmfp, L->seed, L->cell};
    return p;
}

/* CycleTrackingGuts for one particle: history until census / absorption /
 * escape / fission hand-off. */
static void history(Prob* P, const Part* in, Vec* census, Vec* extra, Bal* b) {
    Live L;
    L.r.x = in->x; L.r.y = in->y; L.r.z = in->z; L.v.x = in->vx; L.v.y = in->vy; L.v.z = in->vz;
    L.E = in->E; L.w = in->w; L.tc = in->tc; L.nmfp = in->nmfp; L.seed = in->seed; L.cell = in->cell;
    double sp = sqrt(L.v.x * L.v.x + L.v.y * L.v.y + L.v.z * L.v.z);
    if (sp > 0) { double f = 1.0 / sp; L.a = f * L.v.x; L.b = f * L.v.y; L.g = f * L.v.z; }
    else { L.a = L.b = L.g = 0; }
    if (L.tc <= 0.0) L.tc += P->D->dt;
    L.group = egroup(P, L.E);
    int G = P->G;
    for (;;) {
        /* MC_Segment_Outcome */
        double speed = sqrt(L.v.x * L.v.x + L.v.y * L.v.y + L.v.z * L.v.z);
        int force = 0;
        if (L.nmfp < 0.0) { force = 1; L.nmfp = SMALL_D; }
        int m = P->cmat[L.cell];
        int gidx = L.group < G ? L.group : G - 1;
        double sig = P->mat[m].total[gidx];
        L.sigt = sig;
        double mfp = (sig == 0.0) ? HUGE_D : 1.0 / sig;
        if (L.nmfp == 0.0) { double r = qso_rng_sample(&L.seed); L.nmfp = -1.0 * log(r); }
        double d[3];
        d[0] = force ? SMALL_D : L.nmfp * mfp;
        d[2] = speed * L.tc;
        V3 dir = {L.a, L.b, L.g};
        int facet;
        d[1] = qso_nearest_facet(P, L.cell, &L.r, dir, &facet);
        if (force) { d[1] = HUGE_D; d[2] = HUGE_D; d[0] = TINY_D; }
        int out = 0;
        double mn = d[0];
        for (int q = 1; q < 3; q++) if (d[q] < mn) { mn = d[q]; out = q; }
        double seg = d[out];
        L.nmfp -= seg / mfp;
        if (out == 0) L.nmfp = 0.0;
        else if (out == 2) L.tc = (L.tc < 0.0) ? L.tc : 0.0;
        if (force) L.nmfp = 0.0;
        if (seg != 0.0) {
            L.r.x += (L.a * seg); L.r.y += (L.b * seg); L.r.z += (L.g * seg);
            double t = seg / speed;
            L.tc -= t;
            if (L.tc < 0.0) L.tc = 0.0;
            P->flux[(size_t)L.cell * G + gidx] += seg * L.w;
        }
        b->num_seg++;
        if (out == 0) {
            /* CollisionEvent */
            const MatT* M = &P->mat[m];
            double r = qso_rng_sample(&L.seed);
            double cur = L.sigt * r;
            int sel = -1;
            for (int iso = 0; iso < M->nIso && cur >= 0; iso++)
                for (int q = 0; q < M->nReact; q++) {
                    cur -= M->af * 1.0 * M->xs[(size_t)q * G + gidx];
                    if (cur < 0) { sel = q; break; }
                }
            if (sel < 0) sel = M->nReact - 1;
            int type = M->type[sel];
            double eo[4] = {0, 0, 0, 0}, ao[4] = {0, 0, 0, 0};
            int nout = 0;
            if (type == 1) {
                nout = 1;
                r = qso_rng_sample(&L.seed);
                eo[0] = L.E * (1.0 - (r * (1.0 / M->mass)));
                ao[0] = qso_rng_sample(&L.seed) * 2.0 - 1.0;
            } else if (type == 3) {
                nout = (int)(M->nuBar + qso_rng_sample(&L.seed));
                for (int q = 0; q < nout && q < 4; q++) {
                    double rr = qso_rng_sample(&L.seed) / 2.0 + 0.5;
                    eo[q] = (20 * rr * rr);
                    ao[q] = qso_rng_sample(&L.seed) * 2.0 - 1.0;
                }
                if (nout > 4) nout = 4;
            }
            b->collision++;
            if (type == 1) b->scatter++;
            else if (type == 2) b->absorb++;
            else { b->fission++; b->produce += (uint64_t)nout; }
            if (nout == 0) return;
            for (int q = 1; q < nout; q++) {
                Live S = L;
                S.seed = qso_rng_spawn(&L.seed);
                trajectory(&S, eo[q], ao[q]);
                Part sp2 = to_part(&S);
                push(extra, &sp2);
            }
            trajectory(&L, eo[0], ao[0]);
            if (nout > 1) { Part pp = to_part(&L); push(extra, &pp); return; }
            L.group = egroup(P, L.E);
        } else if (out == 1) {
            /* MC_Facet_Crossing_Event */
            int face = facet / 4;
            int i = L.cell % P->nx, j = (L.cell / P->nx) % P->ny, k = L.cell / (P->nx * P->ny);
            int ni = i + (face == 0) - (face == 1), nj = j + (face == 2) - (face == 3), nk = k + (face == 4) - (face == 5);
            if (ni >= 0 && ni < P->nx && nj >= 0 && nj < P->ny && nk >= 0 && nk < P->nz) {
                L.cell = ni + P->nx * (nj + P->ny * nk);
            } else if (P->bc[face] == 1) {
                b->escape++;
                return;
            } else {
                const double* pl = P->plane + ((size_t)L.cell * 24 + facet) * 4;
                double dot = 2.0 * (L.a * pl[0] + L.b * pl[1] + L.g * pl[2]);
                if (dot > 0) { L.a -= dot * pl[0]; L.b -= dot * pl[1]; L.g -= dot * pl[2]; }
                double s2 = sqrt(L.v.x * L.v.x + L.v.y * L.v.y + L.v.z * L.v.z);
                L.v.x = s2 * L.a; L.v.y = s2 * L.b; L.v.z = s2 * L.g;
            }
        } else {
            Part pp = to_part(&L);
            push(census, &pp);
            b->census++;
            return;
        }
    }
}

/* MC_SourceNow for cycle `cycle` */
static void source(Prob* P, int cycle, Vec* out, Bal* b) {
    Deck* D = P->D;
    for (int c = 0; c < P->nc; c++) {
        int i = c % P->nx, j = (c / P->nx) % P->ny, k = c / (P->nx * P->ny);
        V3 n[14];
        for (int q = 0; q < 14; q++) n[q] = node(P, i, j, k, q);
        V3 ctr = P->center[c];
        for (int s = 0; s < P->nsrc[c]; s++) {
            uint64_t base = (uint64_t)cycle * (uint64_t)P->nsrc[c] + (uint64_t)s + ((uint64_t)c << 32);
            Part p;
            p.seed = qso_rng_spawn(&base);
            double wv = qso_rng_sample(&p.seed) * 6.0 * P->vol[c];
            double cv = 0.0;
            int f = -1;
            V3 a = {0, 0, 0}, bb = {0, 0, 0}, d = {0, 0, 0};
            while (cv < wv) {
                f++;
                if (f == 24) break;
                a = n[TRI[f][0]]; bb = n[TRI[f][1]]; d = n[TRI[f][2]];
                double v0x = a.x - ctr.x, v0y = a.y - ctr.y, v0z = a.z - ctr.z;
                double v1x = bb.x - ctr.x, v1y = bb.y - ctr.y, v1z = bb.z - ctr.z;
                double v2x = d.x - ctr.x, v2y = d.y - ctr.y, v2z = d.z - ctr.z;
                cv += v0z * (v1x * v2y - v1y * v2x) + v0y * (v1z * v2x - v1x * v2z) + v0x * (v1y * v2z - v1z * v2y);
            }
            double r1 = qso_rng_sample(&p.seed), r2 = qso_rng_sample(&p.seed), r3 = qso_rng_sample(&p.seed);
            if (r1 + r2 > 1.0) { r1 = 1.0 - r1; r2 = 1.0 - r2; }
            if (r2 + r3 > 1.0) { double t = r3; r3 = 1.0 - r1 - r2; r2 = 1.0 - t; }
            else if (r1 + r2 + r3 > 1.0) { double t = r3; r3 = r1 + r2 + r3 - 1.0; r1 = 1.0 - r2 - t; }
            double r4 = 1.0 - r1 - r2 - r3;
            p.x = p.y = p.z = 0.0;
            if (f >= 0) {
                p.x = (r4 * ctr.x + r1 * a.x + r2 * bb.x + r3 * d.x);
                p.y = (r4 * ctr.y + r1 * a.y + r2 * bb.y + r3 * d.y);
                p.z = (r4 * ctr.z + r1 * a.z + r2 * bb.z + r3 * d.z);
            }
            double g = 1.0 - 2.0 * qso_rng_sample(&p.seed);
            double sg = sqrt((1.0 - (g * g)));
            double phi = PI_D * (2.0 * qso_rng_sample(&p.seed) - 1.0);
            double al = sg * cos(phi), be = sg * sin(phi);
            p.E = (D->eMax - D->eMin) * qso_rng_sample(&p.seed) + D->eMin;
            double sp = CL * sqrt(p.E * (p.E + 2.0 * (MN)) / ((p.E + MN) * (p.E + MN)));
            p.vx = sp * al; p.vy = sp * be; p.vz = sp * g;
            p.w = P->wsrc;
            p.nmfp = -1.0 * log(qso_rng_sample(&p.seed));
            p.tc = D->dt * qso_rng_sample(&p.seed);
            p.cell = c;
            push(out, &p);
            b->source++;
        }
    }
}

/* PopulationControl + RouletteLowWeightParticles */
static void population(Prob* P, Vec* in, Vec* out, Bal* b) {
    Deck* D = P->D;
    double target = (double)D->nParticles, n = (double)in->n;
    double f = D->loadBalance ? (in->n ? target / n : 1.0) : target / n;
    out->n = 0;
    for (size_t q = 0; q < in->n; q++) {
        Part p = in->p[q];
        int keep = 1, sf = 0;
        if (f != 1.0 && in->n) {
            double r = qso_rng_sample(&p.seed);
            if (f < 1) {
                if (r > f) { keep = 0; b->rr++; } else p.w /= f;
            } else if (f > 1) {
                sf = (int)floor(f);
                if (r > (f - sf)) sf--;
                p.w /= f;
            }
        }
        if (!keep) continue;
        for (int s = 0; s < sf; s++) {
            Part c = p;
            c.seed = qso_rng_spawn(&p.seed);
            b->split++;
            push(out, &c);
        }
        push(out, &p);
    }
    double wc = D->cutoff * P->wsrc;
    if (D->cutoff > 0.0) {
        size_t w = 0;
        for (size_t q = 0; q < out->n; q++) {
            Part p = out->p[q];
            if (p.w <= wc) {
                if (qso_rng_sample(&p.seed) <= D->cutoff) p.w /= D->cutoff;
                else { b->rr++; continue; }
            }
            out->p[w++] = p;
        }
        out->n = w;
    }
}

typedef struct { Prob P; Deck D; Vec census; int cycle; } Run;

/* Runs `cycles` cycles (<= 0: the deck's nSteps).  bal/flux arrays must
 * hold that many entries. */
int qso_run_deck(const char* path, int cycles, uint64_t* bal13, double* flux, double* seconds) {
    Run* R = (Run*)calloc(1, sizeof(Run));
    if (qso_read_deck(path, &R->D)) { free(R); return 1; }
    int rc = qso_build(&R->P, &R->D);
    if (rc) { free(R); return 10 + rc; }
    if (cycles <= 0) cycles = R->D.nSteps;
    Vec proc = {0}, pop = {0}, extra = {0}, next = {0};
    for (int c = 0; c < cycles; c++) {
        Bal b;
        memset(&b, 0, sizeof b);
        b.start = R->census.n;
        proc.n = 0;
        for (size_t q = 0; q < R->census.n; q++) push(&proc, &R->census.p[q]);
        R->census.n = 0;
        source(&R->P, c, &proc, &b);
        population(&R->P, &proc, &pop, &b);
        clock_t t0 = clock();
        Vec* cur = &pop;
        while (cur->n) {
            extra.n = 0;
            for (size_t q = 0; q < cur->n; q++) history(&R->P, &cur->p[q], &R->census, &extra, &b);
            next.n = 0;
            for (size_t q = 0; q < extra.n; q++) push(&next, &extra.p[q]);
            Vec t = pop; pop = next; next = t;
            cur = &pop;
        }
        clock_t t1 = clock();
        b.end = R->census.n;
        double s = 0.0;
        size_t nf = (size_t)R->P.nc * R->P.G;
        for (size_t q = 0; q < nf; q++) { s += R->P.flux[q]; R->P.flux[q] = 0.0; }
        uint64_t* o = bal13 + 13 * (size_t)c;
        o[0] = b.start; o[1] = b.source; o[2] = b.rr; o[3] = b.split; o[4] = b.absorb; o[5] = b.scatter;
        o[6] = b.fission; o[7] = b.produce; o[8] = b.collision; o[9] = b.escape; o[10] = b.census;
        o[11] = b.num_seg; o[12] = b.end;
        flux[c] = s;
        if (seconds) seconds[c] = (double)(t1 - t0) / CLOCKS_PER_SEC;
    }
    return 0;
}

/* qso_nearest_facet on n rays of a deck's mesh (for ray-by-ray checks of a
 * nearest-facet search): xyz, dir [n][3], cell [n] global ids; returns the
 * distance, the facet and the position after any nudge.  The deck's tables
 * are built once per call and released with the process. */
int qso_nearest_facet_rays(const char* path, uint64_t n, const double* xyz, const double* dir, const int32_t* cell,
                           double* dist, int32_t* facet, double* xyz_out) {
    Run* R = (Run*)calloc(1, sizeof(Run));
    if (qso_read_deck(path, &R->D)) { free(R); return 1; }
    int rc = qso_build(&R->P, &R->D);
    if (rc) { free(R); return 10 + rc; }
    for (uint64_t q = 0; q < n; q++) {
        if (cell[q] < 0 || cell[q] >= R->P.nc) return 2;
        V3 r = {xyz[3 * q], xyz[3 * q + 1], xyz[3 * q + 2]};
        V3 d = {dir[3 * q], dir[3 * q + 1], dir[3 * q + 2]};
        int f = -1;
        dist[q] = qso_nearest_facet(&R->P, cell[q], &r, d, &f);
        facet[q] = f;
        xyz_out[3 * q] = r.x; xyz_out[3 * q + 1] = r.y; xyz_out[3 * q + 2] = r.z;
    }
    return 0;
}

#ifdef QSO_MAIN
int main(int argc, char** argv) {
    if (argc < 2) { fprintf(stderr, "usage: qs_oracle deck.inp [cycles]\n"); return 2; }
    Deck* D = (Deck*)malloc(sizeof(Deck));
    if (qso_read_deck(argv[1], D)) { fprintf(stderr, "cannot read %s\n", argv[1]); return 1; }
    int cycles = argc > 2 ? atoi(argv[2]) : D->nSteps;
    free(D);
    uint64_t* bal = (uint64_t*)calloc((size_t)cycles * 13, sizeof(uint64_t));
    double* flux = (double*)calloc((size_t)cycles, sizeof(double));
    double* secs = (double*)calloc((size_t)cycles, sizeof(double));
    int rc = qso_run_deck(argv[1], cycles, bal, flux, secs);
    if (rc) { fprintf(stderr, "qs_oracle: error %d\n", rc); return 1; }
    uint64_t seg = 0;
    double ts = 0;
    for (int c = 0; c < cycles; c++) {
        uint64_t* o = bal + 13 * (size_t)c;
        printf("QSORACLE %d", c);
        for (int q = 0; q < 13; q++) printf(" %llu", (unsigned long long)o[q]);
        printf(" %.17e %.6e\n", flux[c], secs[c]);
        seg += o[11];
        ts += secs[c];
    }
    printf("QSORACLE_FOM %llu %.6e %.6e 1\n", (unsigned long long)seg, ts, ts > 0 ? seg / ts : 0.0);
    return 0;
}
#endif
