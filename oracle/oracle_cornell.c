/*
 * Oracle (test infrastructure only): the reference's CPU Cornell path tracer.
 *
 * Restates RenderCornellReference (src/integrators/reference_cpu/
 * CornellReference.cpp:734-909) and everything it calls, in double:
 *   - the L3-private Cornell fixture and its FNV-1a scene hash
 *     (CornellReference.cpp:119-318);
 *   - PCG32, MixBits64 and MakeSampleRng (include/rt/cpu/Random.hpp:9-79);
 *   - Rt::Cpu::Bvh<double> with BvhBuildMethod::BinnedSah and leaf size 4:
 *     build (include/rt/cpu/Bvh.hpp:82-132, 322-616), TraceClosest / TraceAny
 *     with near-first child order (Bvh.hpp:135-233, 618-678);
 *   - IntersectAabb and the watertight IntersectTriangle
 *     (include/rt/cpu/Geometry.hpp:223-449), Ray::IsValid (Geometry.hpp:42-51);
 *   - Lambert Evaluate / Pdf / Sample (include/bsdf/reference_cpu/Lambert.hpp:20-116);
 *   - OffsetRayOrigin, EvaluateDirectLighting, TracePath, GenerateCameraRay
 *     (CornellReference.cpp:336-695);
 *   - HashLinearRgbPixels (src/rt/cpu/ImageOutput.cpp:343-378).
 * The reference promotes double arithmetic to `long double`
 * (Math.hpp:211-215), which is double under the MSVC build that wrote the
 * committed golden (rt/cpu/references/wave1/cornell-256x256-64spp-seed1.*);
 * this restatement computes those steps in double.
 *
 * Parity pins (tests/test_oracle_cornell.py): the scene hash equals the
 * golden's 0x80597851ae8052c2; the 256x256, 64 spp, 8 bounce, seed 1 image
 * and its ray counts are checked against the golden JSON and EXR; the image
 * agrees with the reference's own RenderCornellReference compiled from its
 * sources into oracle/_ref; Lambert's pdf integrates to one and PCG32
 * matches its published reference sequence.
 */
#include "oracle.h"

#include <float.h>
#include <math.h>
#include <pthread.h>
#include <stdatomic.h>
#include <stdlib.h>
#include <string.h>

typedef struct { double x, y, z; } v3;

static const double kInvPi = 0.31830988618379067154;  /* std::numbers::inv_pi_v<double> */
static const double kPi = 3.14159265358979323846;      /* std::numbers::pi_v<double> */

static v3 mk(double x, double y, double z) { v3 r = {x, y, z}; return r; }
static v3 add(v3 a, v3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
static v3 sub(v3 a, v3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
static v3 neg(v3 a) { return mk(-a.x, -a.y, -a.z); }
static v3 scl(v3 a, double s) { return mk(a.x * s, a.y * s, a.z * s); }
static v3 dvd(v3 a, double s) { return mk(a.x / s, a.y / s, a.z / s); }
static v3 mul(v3 a, v3 b) { return mk(a.x * b.x, a.y * b.y, a.z * b.z); }
static double dot(v3 a, v3 b) { return (a.x * b.x) + (a.y * b.y) + (a.z * b.z); }
static v3 cross(v3 a, v3 b) {
    return mk((a.y * b.z) - (a.z * b.y), (a.z * b.x) - (a.x * b.z), (a.x * b.y) - (a.y * b.x));
}
static double comp(v3 a, int k) { return k == 0 ? a.x : (k == 1 ? a.y : a.z); }
static void set_comp(v3* a, int k, double v) {
    if (k == 0) a->x = v; else if (k == 1) a->y = v; else a->z = v;
}
static int finite3(v3 a) { return isfinite(a.x) && isfinite(a.y) && isfinite(a.z); }
static v3 normalize_or_zero(v3 a) {  /* Math.hpp:162-170 */
    const double l2 = dot(a, a);
    if (!(l2 > 0.0) || !isfinite(l2)) return mk(0.0, 0.0, 0.0);
    return dvd(a, sqrt(l2));
}
static double max_component(v3 a) {  /* std::max(x, std::max(y, z)) */
    const double yz = a.y < a.z ? a.z : a.y;
    return a.x < yz ? yz : a.x;
}
static int finite_nonneg(v3 a) { return finite3(a) && a.x >= 0.0 && a.y >= 0.0 && a.z >= 0.0; }

/* ---- PCG32 (Random.hpp) ------------------------------------------------ */

static uint64_t mix64(uint64_t v) {
    v += 0x9e3779b97f4a7c15ull;
    v = (v ^ (v >> 30)) * 0xbf58476d1ce4e5b9ull;
    v = (v ^ (v >> 27)) * 0x94d049bb133111ebull;
    return v ^ (v >> 31);
}

typedef struct { uint64_t state, inc; } pcg32;

static uint32_t pcg_next(pcg32* g) {
    const uint64_t old = g->state;
    g->state = old * 6364136223846793005ull + g->inc;
    const uint32_t xs = (uint32_t)(((old >> 18) ^ old) >> 27);
    const uint32_t rot = (uint32_t)(old >> 59);
    return (xs >> rot) | (xs << ((0u - rot) & 31u));
}

static void pcg_seed(pcg32* g, uint64_t seed, uint64_t seq) {
    g->state = 0u;
    g->inc = (seq << 1) | 1u;
    (void)pcg_next(g);
    g->state += seed;
    (void)pcg_next(g);
}

static double pcg_uniform(pcg32* g) {
    const uint64_t high = (uint64_t)(pcg_next(g) >> 5);
    const uint64_t low = (uint64_t)(pcg_next(g) >> 6);
    return (double)(high * 67108864ull + low) * 0x1.0p-53;
}

static pcg32 sample_rng(uint64_t seed, uint64_t pixel, uint64_t sample) {
    const uint64_t key = mix64(seed ^ mix64(pixel) ^ mix64(sample + 0x632be59bd9b4e019ull));
    const uint64_t seq = mix64(pixel ^ (sample * 0x9e3779b97f4a7c15ull));
    pcg32 g;
    pcg_seed(&g, key, seq);
    return g;
}

void ora_pcg32_sequence(uint64_t seed, uint64_t seq, uint32_t n, uint32_t* out) {
    pcg32 g;
    pcg_seed(&g, seed, seq);
    for (uint32_t i = 0; i < n; ++i) out[i] = pcg_next(&g);
}

/* ---- scene (CornellReference.cpp:119-318) ------------------------------ */

#define CORNELL_MAX_TRIS 64u
enum { MAT_WHITE = 0, MAT_RED = 1, MAT_GREEN = 2, MAT_EMITTER = 3, MAT_COUNT = 4 };

typedef struct {
    v3 v0, v1, v2;
    uint32_t prim;
} dtri;

typedef struct {
    v3 mn, mx;
} dbox;

typedef struct {
    dbox b;
    uint32_t left, right, first, count;
} dnode;

typedef struct {
    dtri tris[CORNELL_MAX_TRIS];
    uint32_t mat[CORNELL_MAX_TRIS];
    uint32_t triCount;
    v3 refl[MAT_COUNT], emis[MAT_COUNT];
    v3 lcorner, ledgeU, ledgeV, lnormal;
    double larea;
    v3 camPos;
    double vfov;
    uint64_t hash;
    /* Bvh<double> */
    dnode nodes[2 * CORNELL_MAX_TRIS];
    uint32_t nodeCount;
    uint32_t order[CORNELL_MAX_TRIS];
} cornell;

static void add_tri(cornell* s, v3 a, v3 b, v3 c, uint32_t m) {
    const uint32_t i = s->triCount++;
    s->tris[i].v0 = a;
    s->tris[i].v1 = b;
    s->tris[i].v2 = c;
    s->tris[i].prim = i;
    s->mat[i] = m;
}

static void add_quad(cornell* s, v3 a, v3 b, v3 c, v3 d, v3 n, uint32_t m) {
    if (!(dot(cross(sub(b, a), sub(c, a)), n) > 0.0)) {
        const v3 t = b;
        b = d;
        d = t;
    }
    add_tri(s, a, b, c, m);
    add_tri(s, a, c, d, m);
}

static void add_box(cornell* s, v3 mn, v3 mx, uint32_t m) {
    const v3 p000 = mk(mn.x, mn.y, mn.z), p001 = mk(mn.x, mn.y, mx.z), p010 = mk(mn.x, mx.y, mn.z),
             p011 = mk(mn.x, mx.y, mx.z), p100 = mk(mx.x, mn.y, mn.z), p101 = mk(mx.x, mn.y, mx.z),
             p110 = mk(mx.x, mx.y, mn.z), p111 = mk(mx.x, mx.y, mx.z);
    add_quad(s, p000, p001, p011, p010, mk(-1.0, 0.0, 0.0), m);
    add_quad(s, p100, p110, p111, p101, mk(1.0, 0.0, 0.0), m);
    add_quad(s, p000, p100, p101, p001, mk(0.0, -1.0, 0.0), m);
    add_quad(s, p010, p011, p111, p110, mk(0.0, 1.0, 0.0), m);
    add_quad(s, p000, p010, p110, p100, mk(0.0, 0.0, -1.0), m);
    add_quad(s, p001, p101, p111, p011, mk(0.0, 0.0, 1.0), m);
}

static void hash_byte(uint64_t* h, uint8_t v) {
    *h ^= v;
    *h *= 1099511628211ull;
}
static void hash_word(uint64_t* h, uint64_t v) {
    for (int i = 0; i < 8; ++i) hash_byte(h, (uint8_t)(v >> (8 * i)));
}
static void hash_double(uint64_t* h, double v) {
    uint64_t u;
    memcpy(&u, &v, 8);
    hash_word(h, u);
}
static void hash_vec(uint64_t* h, v3 v) {
    hash_double(h, v.x);
    hash_double(h, v.y);
    hash_double(h, v.z);
}

static uint64_t scene_hash(const cornell* s) {
    uint64_t h = 14695981039346656037ull;
    hash_word(&h, 0x4c33434f524e454cull);  /* "L3CORNEL" */
    hash_word(&h, MAT_COUNT);
    for (uint32_t m = 0; m < MAT_COUNT; ++m) {
        hash_word(&h, m);
        hash_vec(&h, s->refl[m]);
        hash_vec(&h, s->emis[m]);
    }
    hash_word(&h, s->triCount);
    for (uint32_t i = 0; i < s->triCount; ++i) {
        hash_word(&h, s->tris[i].prim);
        hash_vec(&h, s->tris[i].v0);
        hash_vec(&h, s->tris[i].v1);
        hash_vec(&h, s->tris[i].v2);
        hash_word(&h, s->mat[i]);
    }
    hash_vec(&h, s->lcorner);
    hash_vec(&h, s->ledgeU);
    hash_vec(&h, s->ledgeV);
    hash_vec(&h, s->lnormal);
    hash_double(&h, s->larea);
    hash_vec(&h, s->camPos);
    hash_double(&h, s->vfov);
    return h;
}

/* ---- Bvh<double> build (Bvh.hpp:82-132, 322-616) ----------------------- */

static int box_valid(const dbox* b) {
    return finite3(b->mn) && finite3(b->mx) && b->mn.x <= b->mx.x && b->mn.y <= b->mx.y && b->mn.z <= b->mx.z;
}
static dbox box_empty(void) {
    dbox b = {{INFINITY, INFINITY, INFINITY}, {-INFINITY, -INFINITY, -INFINITY}};
    return b;
}
static void box_expand_point(dbox* b, v3 p) {  /* ComponentMin / ComponentMax */
    b->mn = mk(b->mn.x < p.x ? b->mn.x : p.x, b->mn.y < p.y ? b->mn.y : p.y, b->mn.z < p.z ? b->mn.z : p.z);
    b->mx = mk(b->mx.x > p.x ? b->mx.x : p.x, b->mx.y > p.y ? b->mx.y : p.y, b->mx.z > p.z ? b->mx.z : p.z);
}
static void box_expand_box(dbox* b, const dbox* o) {  /* Aabb::Expand(const Aabb&) */
    if (!box_valid(o)) return;
    if (!box_valid(b)) {
        *b = *o;
        return;
    }
    box_expand_point(b, o->mn);
    box_expand_point(b, o->mx);
}
static double box_sa(const dbox* b) {  /* SurfaceAreaPromoted */
    if (!box_valid(b)) return 0.0;
    const double ex = b->mx.x - b->mn.x, ey = b->mx.y - b->mn.y, ez = b->mx.z - b->mn.z;
    return 2.0 * ((ex * ey) + (ey * ez) + (ez * ex));
}

typedef struct {
    uint32_t tri, prim;
    dbox b;
    v3 c;
} dref;

static uint32_t bin_of(double c, double cmin, double cext) {  /* ComputeBinIndex */
    const double scaled = ((c - cmin) / cext) * 16.0;
    if (!(scaled > 0.0)) return 0u;
    if (scaled >= 16.0) return 15u;
    return (uint32_t)scaled;
}

static int median_axis;
static int median_cmp(const void* pa, const void* pb) {
    const dref* a = (const dref*)pa;
    const dref* b = (const dref*)pb;
    const double ca = comp(a->c, median_axis), cb = comp(b->c, median_axis);
    if (ca < cb) return -1;
    if (ca > cb) return 1;
    if (a->prim != b->prim) return a->prim < b->prim ? -1 : 1;
    if (a->tri != b->tri) return a->tri < b->tri ? -1 : 1;
    return 0;
}

static uint32_t build_range(cornell* s, dref* refs, uint32_t begin, uint32_t end) {
    const uint32_t ni = s->nodeCount++;
    dbox bounds = box_empty();
    for (uint32_t i = begin; i < end; ++i) box_expand_box(&bounds, &refs[i].b);
    if (end - begin <= 4u) {
        s->nodes[ni].b = bounds;
        s->nodes[ni].first = begin;
        s->nodes[ni].count = end - begin;
        s->nodes[ni].left = s->nodes[ni].right = 0xffffffffu;
        return ni;
    }
    dbox cb = box_empty();
    for (uint32_t i = begin; i < end; ++i) box_expand_point(&cb, refs[i].c);
    const v3 cext = sub(cb.mx, cb.mn);
    uint32_t middle = begin;
    /* FindSahSplit */
    int valid = 0;
    uint32_t bestAxis = 0, bestBin = 0;
    double bestCost = INFINITY;
    for (int axis = 0; axis < 3; ++axis) {
        if (!(comp(cext, axis) > 0.0)) continue;
        dbox bins[16];
        uint32_t counts[16];
        for (int k = 0; k < 16; ++k) {
            bins[k] = box_empty();
            counts[k] = 0;
        }
        for (uint32_t i = begin; i < end; ++i) {
            const uint32_t bi = bin_of(comp(refs[i].c, axis), comp(cb.mn, axis), comp(cext, axis));
            counts[bi]++;
            box_expand_box(&bins[bi], &refs[i].b);
        }
        for (uint32_t split = 0; split + 1u < 16u; ++split) {
            dbox lb = box_empty(), rb = box_empty();
            uint32_t lc = 0, rc = 0;
            for (uint32_t k = 0; k <= split; ++k) {
                lc += counts[k];
                box_expand_box(&lb, &bins[k]);
            }
            for (uint32_t k = split + 1u; k < 16u; ++k) {
                rc += counts[k];
                box_expand_box(&rb, &bins[k]);
            }
            if (lc == 0 || rc == 0) continue;
            const double cost = (box_sa(&lb) * (double)lc) + (box_sa(&rb) * (double)rc);
            if (!valid || cost < bestCost) {
                valid = 1;
                bestAxis = (uint32_t)axis;
                bestBin = split;
                bestCost = cost;
            }
        }
    }
    if (valid) {  /* SahPartition: stable partition by bin */
        dref tmp[CORNELL_MAX_TRIS];
        uint32_t nl = 0, nr = 0;
        dref right[CORNELL_MAX_TRIS];
        for (uint32_t i = begin; i < end; ++i) {
            const uint32_t bi = bin_of(comp(refs[i].c, (int)bestAxis), comp(cb.mn, (int)bestAxis),
                                       comp(cext, (int)bestAxis));
            if (bi <= bestBin) tmp[nl++] = refs[i];
            else right[nr++] = refs[i];
        }
        memcpy(refs + begin, tmp, nl * sizeof(dref));
        memcpy(refs + begin + nl, right, nr * sizeof(dref));
        middle = begin + nl;
    }
    if (middle == begin || middle == end) {  /* MedianPartition */
        const double ex = cext.x, ey = cext.y, ez = cext.z;
        int axis = 0;  /* LongestAxis */
        if (ey > ex) axis = 1;
        if (ez > comp(cext, axis)) axis = 2;
        median_axis = axis;
        /* stable: the comparator breaks every tie by (primitive, index) */
        qsort(refs + begin, end - begin, sizeof(dref), median_cmp);
        middle = begin + (end - begin) / 2u;
    }
    const uint32_t l = build_range(s, refs, begin, middle);
    const uint32_t r = build_range(s, refs, middle, end);
    s->nodes[ni].b = bounds;
    s->nodes[ni].left = l;
    s->nodes[ni].right = r;
    s->nodes[ni].first = 0;
    s->nodes[ni].count = 0;
    return ni;
}

static void build_bvh(cornell* s) {
    dref refs[CORNELL_MAX_TRIS];
    for (uint32_t i = 0; i < s->triCount; ++i) {
        const dtri* t = &s->tris[i];
        refs[i].tri = i;
        refs[i].prim = t->prim;
        refs[i].b = box_empty();
        box_expand_point(&refs[i].b, t->v0);
        box_expand_point(&refs[i].b, t->v1);
        box_expand_point(&refs[i].b, t->v2);
        refs[i].c = mk((t->v0.x + t->v1.x + t->v2.x) / 3.0, (t->v0.y + t->v1.y + t->v2.y) / 3.0,
                       (t->v0.z + t->v1.z + t->v2.z) / 3.0);
    }
    s->nodeCount = 0;
    (void)build_range(s, refs, 0, s->triCount);
    for (uint32_t i = 0; i < s->triCount; ++i) s->order[i] = refs[i].tri;
}

static void make_scene(cornell* s) {
    memset(s, 0, sizeof(*s));
    s->refl[MAT_WHITE] = mk(0.73, 0.73, 0.73);
    s->refl[MAT_RED] = mk(0.63, 0.065, 0.05);
    s->refl[MAT_GREEN] = mk(0.14, 0.45, 0.091);
    s->emis[MAT_EMITTER] = mk(17.0, 15.0, 12.0);
    const double x0 = -1.0, x1 = 1.0, y0 = -1.0, y1 = 1.0, zf = -1.0, zb = -3.0;
    add_quad(s, mk(x0, y0, zf), mk(x1, y0, zf), mk(x1, y0, zb), mk(x0, y0, zb), mk(0.0, 1.0, 0.0), MAT_WHITE);
    add_quad(s, mk(x0, y1, zf), mk(x0, y1, zb), mk(x1, y1, zb), mk(x1, y1, zf), mk(0.0, -1.0, 0.0), MAT_WHITE);
    add_quad(s, mk(x0, y0, zb), mk(x1, y0, zb), mk(x1, y1, zb), mk(x0, y1, zb), mk(0.0, 0.0, 1.0), MAT_WHITE);
    add_quad(s, mk(x0, y0, zf), mk(x0, y0, zb), mk(x0, y1, zb), mk(x0, y1, zf), mk(1.0, 0.0, 0.0), MAT_RED);
    add_quad(s, mk(x1, y0, zf), mk(x1, y1, zf), mk(x1, y1, zb), mk(x1, y0, zb), mk(-1.0, 0.0, 0.0), MAT_GREEN);
    s->lcorner = mk(-0.30, 0.985, -2.30);
    s->ledgeU = mk(0.60, 0.0, 0.0);
    s->ledgeV = mk(0.0, 0.0, 0.60);
    s->lnormal = mk(0.0, -1.0, 0.0);
    s->larea = sqrt(dot(cross(s->ledgeU, s->ledgeV), cross(s->ledgeU, s->ledgeV)));
    add_quad(s, s->lcorner, add(s->lcorner, s->ledgeU), add(add(s->lcorner, s->ledgeU), s->ledgeV),
             add(s->lcorner, s->ledgeV), s->lnormal, MAT_EMITTER);
    add_box(s, mk(-0.72, -0.999, -2.55), mk(-0.10, -0.30, -1.72), MAT_WHITE);
    add_box(s, mk(0.16, -0.999, -2.75), mk(0.74, 0.22, -1.93), MAT_WHITE);
    s->camPos = mk(0.0, 0.0, 0.15);
    s->vfov = 39.0 * kPi / 180.0;
    s->hash = scene_hash(s);
    build_bvh(s);
}

/* ---- rays, boxes, triangles (Geometry.hpp) ----------------------------- */

typedef struct {
    v3 o, d;
    double tmin, tmax;
} dray;

static int ray_valid(const dray* r) {  /* Ray::IsValid */
    const double l2 = dot(r->d, r->d);
    return finite3(r->o) && finite3(r->d) && isfinite(l2) && fabs(l2 - 1.0) <= 0x1.0p-52 * 64.0 &&
           !isnan(r->tmin) && !isnan(r->tmax) && r->tmin < r->tmax;
}

static int hit_box(const dray* r, const dbox* b, double* nearOut) {  /* IntersectAabb */
    *nearOut = INFINITY;
    if (!ray_valid(r) || !box_valid(b)) return 0;
    double inear = -INFINITY, ifar = INFINITY;
    for (int a = 0; a < 3; ++a) {
        const double o = comp(r->o, a), d = comp(r->d, a), lo = comp(b->mn, a), hi = comp(b->mx, a);
        if (d == 0.0) {
            if (o < lo || o > hi) return 0;
            continue;
        }
        double an = (lo - o) / d, af = (hi - o) / d;
        if (an > af) {
            const double t = an;
            an = af;
            af = t;
        }
        if (an > inear) inear = an;
        if (af < ifar) ifar = af;
        if (ifar < inear) return 0;
    }
    if (!(ifar > r->tmin && inear < r->tmax)) return 0;
    *nearOut = inear > r->tmin ? inear : r->tmin;
    return 1;
}

typedef struct {
    int hit;
    double t;
    uint32_t prim;
    v3 gn;
} dhit;

static int hit_tri(const dray* r, const dtri* tri, dhit* out) {  /* watertight IntersectTriangle */
    out->hit = 0;
    if (!ray_valid(r) || !finite3(tri->v0) || !finite3(tri->v1) || !finite3(tri->v2)) return 0;
    const v3 ad = mk(fabs(r->d.x), fabs(r->d.y), fabs(r->d.z));
    int kz = 0;  /* LongestAxis */
    if (ad.y > ad.x) kz = 1;
    if (ad.z > comp(ad, kz)) kz = 2;
    if (comp(r->d, kz) == 0.0) return 0;
    int kx = (kz + 1) % 3, ky = (kx + 1) % 3;
    if (comp(r->d, kz) < 0.0) {
        const int t = kx;
        kx = ky;
        ky = t;
    }
    const v3 a = sub(tri->v0, r->o), b = sub(tri->v1, r->o), c = sub(tri->v2, r->o);
    const double sx = comp(r->d, kx) / comp(r->d, kz);
    const double sy = comp(r->d, ky) / comp(r->d, kz);
    const double sz = 1.0 / comp(r->d, kz);
    const double ax = comp(a, kx) - (sx * comp(a, kz)), ay = comp(a, ky) - (sy * comp(a, kz));
    const double bx = comp(b, kx) - (sx * comp(b, kz)), by = comp(b, ky) - (sy * comp(b, kz));
    const double cx = comp(c, kx) - (sx * comp(c, kz)), cy = comp(c, ky) - (sy * comp(c, kz));
    const double e0 = (bx * cy) - (by * cx);
    const double e1 = (cx * ay) - (cy * ax);
    const double e2 = (ax * by) - (ay * bx);
    const int negE = e0 < 0.0 || e1 < 0.0 || e2 < 0.0;
    const int posE = e0 > 0.0 || e1 > 0.0 || e2 > 0.0;
    if (negE && posE) return 0;
    const double det = e0 + e1 + e2;
    if (det == 0.0 || !isfinite(det)) return 0;
    const double az = sz * comp(a, kz), bz = sz * comp(b, kz), cz = sz * comp(c, kz);
    const double dist = ((e0 * az) + (e1 * bz) + (e2 * cz)) / det;
    if (!isfinite(dist) || !(dist > r->tmin && dist < r->tmax)) return 0;
    const v3 ea = sub(tri->v1, tri->v0), eb = sub(tri->v2, tri->v0);
    const v3 n = cross(ea, eb);
    const double nl2 = dot(n, n);
    if (!(nl2 > 0.0) || !isfinite(nl2)) return 0;
    const double inl = 1.0 / sqrt(nl2);
    out->hit = 1;
    out->t = dist;
    out->prim = tri->prim;
    out->gn = mk(n.x * inl, n.y * inl, n.z * inl);
    return 1;
}

static int better(const dhit* c, const dhit* cur) {  /* IsBetterHit */
    if (!c->hit) return 0;
    if (!cur->hit) return 1;
    if (c->t < cur->t) return 1;
    if (c->t > cur->t) return 0;
    return c->prim < cur->prim;
}

static void push_children(const cornell* s, uint32_t l, double ln, int hl, uint32_t r, double rn, int hr,
                          uint32_t* stack, uint32_t* sp) {  /* PushIntersectedChildren */
    if (hl && hr) {
        const int leftFirst = ln < rn || (!(rn < ln) && l < r);
        stack[(*sp)++] = leftFirst ? r : l;
        stack[(*sp)++] = leftFirst ? l : r;
        return;
    }
    if (hl) stack[(*sp)++] = l;
    else if (hr) stack[(*sp)++] = r;
    (void)s;
}

static dhit trace_closest(const cornell* s, const dray* r) {
    dhit best;
    memset(&best, 0, sizeof(best));
    best.t = INFINITY;
    best.prim = 0xffffffffu;
    uint32_t stack[128];
    uint32_t sp = 0;
    stack[sp++] = 0;
    while (sp != 0) {
        const dnode* n = &s->nodes[stack[--sp]];
        double nn;
        if (!hit_box(r, &n->b, &nn) || (best.hit && nn > best.t)) continue;
        if (n->count != 0) {
            for (uint32_t i = n->first; i < n->first + n->count; ++i) {
                dhit c;
                if (hit_tri(r, &s->tris[s->order[i]], &c) && better(&c, &best)) best = c;
            }
            continue;
        }
        double ln, rn;
        int hl = hit_box(r, &s->nodes[n->left].b, &ln);
        int hr = hit_box(r, &s->nodes[n->right].b, &rn);
        if (best.hit) {
            hl = hl && ln <= best.t;
            hr = hr && rn <= best.t;
        }
        push_children(s, n->left, ln, hl, n->right, rn, hr, stack, &sp);
    }
    return best;
}

static int trace_any(const cornell* s, const dray* r) {
    uint32_t stack[128];
    uint32_t sp = 0;
    stack[sp++] = 0;
    while (sp != 0) {
        const dnode* n = &s->nodes[stack[--sp]];
        double nn;
        if (!hit_box(r, &n->b, &nn)) continue;
        if (n->count != 0) {
            for (uint32_t i = n->first; i < n->first + n->count; ++i) {
                dhit c;
                if (hit_tri(r, &s->tris[s->order[i]], &c)) return 1;
            }
            continue;
        }
        double ln, rn;
        const int hl = hit_box(r, &s->nodes[n->left].b, &ln);
        const int hr = hit_box(r, &s->nodes[n->right].b, &rn);
        push_children(s, n->left, ln, hl, n->right, rn, hr, stack, &sp);
    }
    return 0;
}

/* ---- Lambert (Lambert.hpp) --------------------------------------------- */


static v3 lambert_eval(v3 refl, v3 n, v3 wo, v3 wi) {
    if (!finite3(refl) || refl.x < 0.0 || refl.y < 0.0 || refl.z < 0.0 || !finite3(n) || !finite3(wo) ||
        !finite3(wi) || !(dot(n, wo) > 0.0) || !(dot(n, wi) > 0.0)) {
        return mk(0.0, 0.0, 0.0);
    }
    return mk(refl.x * kInvPi, refl.y * kInvPi, refl.z * kInvPi);
}

static double lambert_pdf(v3 n, v3 wo, v3 wi) {
    if (!finite3(n) || !finite3(wo) || !finite3(wi) || !(dot(n, wo) > 0.0)) return 0.0;
    const double c = dot(n, wi);
    return (c > 0.0 ? c : 0.0) * kInvPi;
}

typedef struct {
    v3 dir, value;
    double pdf;
    int valid;
} lsample;

static lsample lambert_sample(v3 refl, v3 n, v3 wo, pcg32* g) {
    lsample s;
    memset(&s, 0, sizeof(s));
    if (!finite3(refl) || refl.x < 0.0 || refl.y < 0.0 || refl.z < 0.0 || !finite3(n) || !finite3(wo) ||
        !(dot(n, wo) > 0.0)) {
        return s;
    }
    const double sxv = pcg_uniform(g);
    const double syv = pcg_uniform(g);
    const double radius = sqrt(sxv);
    const double phi = 2.0 * kPi * syv;
    const double lx = radius * cos(phi);
    const double ly = radius * sin(phi);
    const double lz = sqrt(0.0 < 1.0 - sxv ? 1.0 - sxv : 0.0);  /* std::max(0.0, 1 - x) */
    const v3 un = normalize_or_zero(n);
    const v3 helper = fabs(un.z) < 0.999 ? mk(0.0, 0.0, 1.0) : mk(0.0, 1.0, 0.0);
    const v3 tangent = normalize_or_zero(cross(helper, un));
    const v3 bitangent = cross(un, tangent);
    s.dir = normalize_or_zero(add(add(scl(tangent, lx), scl(bitangent, ly)), scl(un, lz)));
    s.pdf = lambert_pdf(un, wo, s.dir);
    s.value = lambert_eval(refl, un, wo, s.dir);
    s.valid = s.pdf > 0.0 && finite3(s.dir) && finite3(s.value) && isfinite(s.pdf);
    return s;
}

double ora_lambert_pdf(const double n[3], const double wo[3], const double wi[3]) {
    return lambert_pdf(mk(n[0], n[1], n[2]), mk(wo[0], wo[1], wo[2]), mk(wi[0], wi[1], wi[2]));
}

/* Lambert::Sample driven by a PCG32 stream (seed, sequence): direction and pdf. */
void ora_lambert_sample(const double n[3], const double wo[3], uint64_t seed, uint64_t seq, double out[4]) {
    pcg32 g;
    pcg_seed(&g, seed, seq);
    const lsample s = lambert_sample(mk(0.5, 0.5, 0.5), mk(n[0], n[1], n[2]), mk(wo[0], wo[1], wo[2]), &g);
    out[0] = s.dir.x;
    out[1] = s.dir.y;
    out[2] = s.dir.z;
    out[3] = s.valid ? s.pdf : 0.0;
}

/* ---- integrator (CornellReference.cpp:336-695) ------------------------- */

typedef struct {
    uint64_t rays, shadowRays, nonFinite;
} counters;

static double power_heuristic(double a, double b) {
    if (!(a > 0.0) || !isfinite(a)) return 0.0;
    if (!(b > 0.0) || !isfinite(b)) return 1.0;
    const double ratio = b / a;
    if (!isfinite(ratio)) return 0.0;
    return 1.0 / (1.0 + ratio * ratio);
}

static int offset_origin(v3 p, v3 gn, v3 d, v3* out) {
    if (!finite3(p) || !finite3(gn) || !finite3(d)) return 0;
    v3 n = normalize_or_zero(gn);
    if (!(dot(n, n) > 0.0)) return 0;
    if (dot(n, d) < 0.0) n = neg(n);
    const double kErr = 32.0 * 0x1.0p-52;
    for (int a = 0; a < 3; ++a) {
        const double pa = comp(p, a), na = comp(n, a);
        const double mag = 1.0 < fabs(pa) ? fabs(pa) : 1.0;  /* std::max(1.0, |p|) */
        const double shifted = pa + na * mag * kErr;
        if (!isfinite(shifted)) return 0;
        if (na > 0.0) set_comp(out, a, nextafter(shifted, INFINITY));
        else if (na < 0.0) set_comp(out, a, nextafter(shifted, -INFINITY));
        else set_comp(out, a, shifted);
    }
    return finite3(*out);
}

static double light_pdf_sa(const cornell* s, v3 ref, v3 lp) {
    const v3 dsp = sub(lp, ref);
    const double d2 = dot(dsp, dsp);
    if (!(d2 > 0.0) || !isfinite(d2) || !(s->larea > 0.0)) return 0.0;
    const v3 dir = dvd(dsp, sqrt(d2));
    const double lc = dot(s->lnormal, neg(dir));
    if (!(lc > 0.0) || !isfinite(lc)) return 0.0;
    const double pdf = d2 / (lc * s->larea);
    return isfinite(pdf) ? pdf : 0.0;
}

static v3 direct_light(const cornell* s, v3 p, v3 n, v3 wo, uint32_t m, pcg32* g, counters* c) {
    const v3 zero = mk(0.0, 0.0, 0.0);
    const double su = pcg_uniform(g);
    const double sv = pcg_uniform(g);
    const v3 lp = add(add(s->lcorner, scl(s->ledgeU, su)), scl(s->ledgeV, sv));
    const v3 tl = sub(lp, p);
    const double d2 = dot(tl, tl);
    if (!(d2 > 0.0) || !isfinite(d2)) {
        c->nonFinite++;
        return zero;
    }
    const double dist = sqrt(d2);
    const v3 wi = dvd(tl, dist);
    const double sc = dot(n, wi);
    const double lc = dot(s->lnormal, neg(wi));
    if (!(sc > 0.0) || !(lc > 0.0)) return zero;
    const double lpdf = d2 / (lc * s->larea);
    const double bpdf = lambert_pdf(n, wo, wi);
    if (!(lpdf > 0.0) || !isfinite(lpdf)) {
        c->nonFinite++;
        return zero;
    }
    v3 so;
    if (!offset_origin(p, n, wi, &so)) {
        c->nonFinite++;
        return zero;
    }
    const v3 otl = sub(lp, so);
    const double sd2 = dot(otl, otl);
    if (!(sd2 > 0.0) || !isfinite(sd2)) {
        c->nonFinite++;
        return zero;
    }
    const double sd = sqrt(sd2);
    const v3 sdir = dvd(otl, sd);
    const double smax = nextafter(sd, 0.0);
    if (!(smax > 0.0) || !isfinite(smax)) {
        c->nonFinite++;
        return zero;
    }
    const dray sr = {so, sdir, 0.0, smax};
    if (!ray_valid(&sr)) {
        c->nonFinite++;
        return zero;
    }
    c->shadowRays++;
    if (trace_any(s, &sr)) return zero;
    const v3 f = lambert_eval(s->refl[m], n, wo, wi);
    const double w = power_heuristic(lpdf, bpdf);
    const double k = sc * w / lpdf;
    const v3 direct = scl(mul(f, s->emis[MAT_EMITTER]), k);
    if (!finite_nonneg(direct)) {
        c->nonFinite++;
        return zero;
    }
    return direct;
}

static v3 trace_path(const cornell* s, dray ray, pcg32* g, uint32_t maxBounces, counters* c) {
    v3 L = mk(0.0, 0.0, 0.0);
    v3 beta = mk(1.0, 1.0, 1.0);
    v3 prevPos = mk(0.0, 0.0, 0.0);
    double prevPdf = 0.0;
    int hasPrev = 0;
    for (uint32_t bounce = 0; bounce < maxBounces; ++bounce) {
        c->rays++;
        const dhit h = trace_closest(s, &ray);
        if (!h.hit) break;
        if (h.prim >= s->triCount) {
            c->nonFinite++;
            break;
        }
        const v3 p = add(ray.o, scl(ray.d, h.t));
        v3 n = normalize_or_zero(h.gn);
        const v3 wo = neg(ray.d);
        if (!finite3(p) || !finite3(n) || !finite3(wo) || !(dot(n, n) > 0.0)) {
            c->nonFinite++;
            break;
        }
        if (dot(n, wo) < 0.0) n = neg(n);
        const uint32_t m = s->mat[h.prim];
        const v3 em = s->emis[m];
        if (em.x > 0.0 || em.y > 0.0 || em.z > 0.0) {
            if (dot(h.gn, wo) > 0.0) {
                double w = 1.0;
                if (hasPrev) w = power_heuristic(prevPdf, light_pdf_sa(s, prevPos, p));
                const v3 e = scl(mul(beta, em), w);
                if (finite_nonneg(e)) L = add(L, e);
                else c->nonFinite++;
            }
            break;
        }
        const v3 direct = direct_light(s, p, n, wo, m, g, c);
        const v3 wd = mul(beta, direct);
        if (finite_nonneg(wd)) {
            L = add(L, wd);
        } else {
            c->nonFinite++;
            break;
        }
        if (bounce + 1u >= maxBounces) break;
        const lsample smp = lambert_sample(s->refl[m], n, wo, g);
        if (!smp.valid) break;
        const double cosv = dot(n, smp.dir);
        if (!(cosv > 0.0) || !(smp.pdf > 0.0)) break;
        beta = scl(mul(beta, smp.value), cosv / smp.pdf);
        if (!finite_nonneg(beta)) {
            c->nonFinite++;
            break;
        }
        prevPos = p;
        prevPdf = smp.pdf;
        hasPrev = 1;
        if (bounce >= 3u) {
            double q = max_component(beta);
            q = q < 0.05 ? 0.05 : (0.95 < q ? 0.95 : q);  /* std::clamp */
            if (pcg_uniform(g) >= q) break;
            beta = dvd(beta, q);
        }
        v3 no;
        if (!offset_origin(p, n, smp.dir, &no)) {
            c->nonFinite++;
            break;
        }
        ray.o = no;
        ray.d = smp.dir;
        ray.tmin = 0.0;
        ray.tmax = INFINITY;
        if (!ray_valid(&ray)) {
            c->nonFinite++;
            break;
        }
    }
    if (!finite_nonneg(L)) {
        c->nonFinite++;
        return mk(0.0, 0.0, 0.0);
    }
    return L;
}

static dray camera_ray(const cornell* s, uint32_t w, uint32_t h, uint32_t px, uint32_t py, pcg32* g) {
    const double jx = pcg_uniform(g);
    const double jy = pcg_uniform(g);
    const double W = (double)w, H = (double)h;
    const double aspect = W / H;
    const double th = tan(s->vfov * 0.5);
    const double sx = (((double)px + jx) / W * 2.0 - 1.0) * aspect * th;
    const double sy = (1.0 - ((double)py + jy) / H * 2.0) * th;
    dray r;
    r.o = s->camPos;
    r.d = normalize_or_zero(mk(sx, sy, -1.0));
    r.tmin = 0.0;
    r.tmax = INFINITY;
    return r;
}

typedef struct {
    const cornell* s;
    uint32_t w, h, spp, maxBounces, rowBegin, rowEnd;
    uint64_t seed;
    float* pixels;
    atomic_uint next;
    counters total;
    pthread_mutex_t lock;
} job;

static void* worker(void* arg) {
    job* J = (job*)arg;
    counters c = {0, 0, 0};
    for (;;) {
        const uint32_t py = J->rowBegin + atomic_fetch_add(&J->next, 1u);
        if (py >= J->rowEnd) break;
        for (uint32_t px = 0; px < J->w; ++px) {
            const uint64_t pix = (uint64_t)py * J->w + px;
            v3 acc = mk(0.0, 0.0, 0.0);
            for (uint32_t si = 0; si < J->spp; ++si) {
                pcg32 g = sample_rng(J->seed, pix, si);
                const dray cr = camera_ray(J->s, J->w, J->h, px, py, &g);
                if (!ray_valid(&cr)) {
                    c.nonFinite++;
                    continue;
                }
                acc = add(acc, trace_path(J->s, cr, &g, J->maxBounces, &c));
            }
            const v3 avg = dvd(acc, (double)J->spp);
            float* out = J->pixels + (size_t)(pix - (uint64_t)J->rowBegin * J->w) * 3u;
            if (!finite_nonneg(avg) || max_component(avg) > (double)FLT_MAX) {
                c.nonFinite++;
                out[0] = out[1] = out[2] = 0.0f;
            } else {
                out[0] = (float)avg.x;
                out[1] = (float)avg.y;
                out[2] = (float)avg.z;
            }
        }
    }
    pthread_mutex_lock(&J->lock);
    J->total.rays += c.rays;
    J->total.shadowRays += c.shadowRays;
    J->total.nonFinite += c.nonFinite;
    pthread_mutex_unlock(&J->lock);
    return NULL;
}

/* Renders rows [rowBegin, rowEnd) of the fixture at w x h (pixels: that
 * band, row-major RGB float).  counts: rays, shadow rays, non-finite.
 * Returns 0, or -1 for invalid options (CornellReference.cpp:697-731). */
int ora_cornell_render(uint32_t w, uint32_t h, uint32_t spp, uint32_t maxBounces, uint64_t seed,
                       uint32_t threads, uint32_t rowBegin, uint32_t rowEnd, float* pixels, uint64_t counts[3],
                       uint64_t* sceneHash) {
    if (w == 0 || h == 0 || spp == 0 || maxBounces == 0 || rowBegin > rowEnd || rowEnd > h) return -1;
    cornell* s = (cornell*)malloc(sizeof(cornell));
    if (!s) return -1;
    make_scene(s);
    job J;
    J.s = s;
    J.w = w;
    J.h = h;
    J.spp = spp;
    J.maxBounces = maxBounces;
    J.rowBegin = rowBegin;
    J.rowEnd = rowEnd;
    J.seed = seed;
    J.pixels = pixels;
    atomic_init(&J.next, 0u);
    J.total.rays = J.total.shadowRays = J.total.nonFinite = 0;
    pthread_mutex_init(&J.lock, NULL);
    if (threads == 0) threads = 1;
    if (threads > rowEnd - rowBegin && rowEnd > rowBegin) threads = rowEnd - rowBegin;
    pthread_t* tids = (pthread_t*)malloc(sizeof(pthread_t) * threads);
    uint32_t started = 0;
    for (uint32_t t = 1; t < threads; ++t) {
        if (pthread_create(&tids[t], NULL, worker, &J) == 0) ++started;
        else break;
    }
    worker(&J);
    for (uint32_t t = 1; t <= started; ++t) pthread_join(tids[t], NULL);
    free(tids);
    pthread_mutex_destroy(&J.lock);
    if (counts) {
        counts[0] = J.total.rays;
        counts[1] = J.total.shadowRays;
        counts[2] = J.total.nonFinite;
    }
    if (sceneHash) *sceneHash = s->hash;
    free(s);
    return 0;
}

/* The fixture's triangles (primitive order, 9 doubles each), material IDs and
 * BVH leaf order; returns the triangle count. */
uint32_t ora_cornell_scene(double* verts, uint32_t* materials, uint32_t* leafOrder, uint32_t* nodeCount) {
    cornell* s = (cornell*)malloc(sizeof(cornell));
    if (!s) return 0;
    make_scene(s);
    for (uint32_t i = 0; i < s->triCount; ++i) {
        const dtri* t = &s->tris[i];
        const double v[9] = {t->v0.x, t->v0.y, t->v0.z, t->v1.x, t->v1.y, t->v1.z, t->v2.x, t->v2.y, t->v2.z};
        if (verts) memcpy(verts + 9u * i, v, sizeof(v));
        if (materials) materials[i] = s->mat[i];
        if (leafOrder) leafOrder[i] = s->order[i];
    }
    if (nodeCount) *nodeCount = s->nodeCount;
    const uint32_t n = s->triCount;
    free(s);
    return n;
}

/* HashLinearRgbPixels (ImageOutput.cpp:343-378). */
uint64_t ora_linear_rgb_hash(uint32_t w, uint32_t h, const float* rgb) {
    uint64_t hsh = 14695981039346656037ull;
    const uint8_t domain[8] = {'R', 'E', 'R', 'G', 'B', 1u, 3u, 0u};
    for (int i = 0; i < 8; ++i) hash_byte(&hsh, domain[i]);
    for (int i = 0; i < 4; ++i) hash_byte(&hsh, (uint8_t)(w >> (8 * i)));
    for (int i = 0; i < 4; ++i) hash_byte(&hsh, (uint8_t)(h >> (8 * i)));
    const size_t n = (size_t)w * h * 3u;
    for (size_t i = 0; i < n; ++i) {
        const float v = rgb[i] == 0.0f ? 0.0f : rgb[i];
        uint32_t u;
        memcpy(&u, &v, 4);
        for (int k = 0; k < 4; ++k) hash_byte(&hsh, (uint8_t)(u >> (8 * k)));
    }
    return hsh;
}
