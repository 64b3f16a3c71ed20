/*
 * oracle_impl.h -- body of the CPU oracle, instantiated twice by oracle.c:
 * once with REAL=float (the bit-exact parity oracle) and once with
 * REAL=double (the FP64 twin used to report error, never for parity).
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library links or calls
 * this code; only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline
 * leg load it, and only as the checker.
 *
 * It is a statement-by-statement restatement of the reference crate
 * Murrent/bendy2d 0.0.1 (Rust) plus the documented extensions (EXT1..EXT6).
 * Every function cites the reference file:line it follows.  Arithmetic order
 * follows the Rust expressions left to right, with nalgebra 0.32
 * Vector2<f32> semantics restated (nalgebra is not vendored in the reference):
 *   dot(a,b)        = a.x*b.x + a.y*b.y
 *   norm_squared(v) = 0 + dot(v,v)  (== dot(v,v) bit for bit, it is never -0)
 *   magnitude(v)    = sqrt(norm_squared(v))
 *   normalize(v)    = (v.x / |v|, v.y / |v|)   (two divisions)
 *   v * s, v / s    = componentwise; s * v == v * s
 * The file must be compiled with -ffp-contract=off and without fast-math.
 *
 * Pinning: every law is checked by exact known answers and invariants
 * (tests/test_oracle_kat.py) and bit for bit against a second restatement
 * written from the Rust source (tests/rust_model.py,
 * tests/test_oracle_model.py); tests/test_oracle_mutations.py plants 31
 * plausible slips in copies of this file and each one fails a test.  Parity
 * unpinned: the instruction order inside nalgebra's magnitude / normalize /
 * dot, which both restatements read from nalgebra 0.32's published source and
 * only a rustc build of the crate could confirm.
 */

#ifndef REAL
#error "define REAL and PFX before including oracle_impl.h"
#endif

#define CAT_(a, b) a##b
#define CAT(a, b) CAT_(a, b)
#define F(name) CAT(PFX, name)
#define T(name) CAT(PFX, name)

typedef struct { REAL x, y; } T(v2);
typedef struct { T(v2) pos, prev, acc; } T(particle);      /* src/particle.rs:4-9 */
typedef struct { int64_t a, b; REAL t; } T(link);          /* src/link.rs:4-10 */
typedef struct { T(particle) pt; REAL r; } T(circle);      /* src/circle.rs:4-8 */
typedef struct {                                           /* src/polygon.rs:7-14 */
    T(particle) *pts; int64_t n;
    T(link) *links; int64_t nl;
    int is_static;
    T(v2) center;
    REAL scale;
} T(polygon);
typedef struct { int kind; REAL cx, cy, r; int64_t nv; REAL *nx, *ny, *c; } T(collider);

typedef struct F(solver) {
    T(v2) gravity;                 /* src/solver.rs:21 */
    T(v2) bpos, bsize;             /* src/solver.rs:13-17,22 */
    int bounds_active;             /* src/solver.rs:23 (stored, never read) */
    T(particle) *parts; int64_t np, cap_p;
    uint8_t *pinned;               /* EXT3 */
    T(link) *plinks; int64_t npl, cap_pl;
    T(circle) *circles; int64_t nc, cap_c;
    T(link) *clinks; int64_t ncl, cap_cl;
    T(polygon) *polys; int64_t npoly, cap_poly;
    T(collider) *cols; int64_t ncol, cap_col;   /* EXT4 */
    int sub_steps;                 /* src/solver.rs:29 (EXT1 setter) */
    REAL sub_steps_multiplier;     /* src/solver.rs:30 */
    int link_iters;                /* EXT2 */
    int cc_mode;                   /* 0: all pairs (reference loop), 1: candidate list */
    REAL cc_margin;                /* current margin for cc_mode 1 */
    int64_t cc_redos;              /* statistics for cc_mode 1 */
} F(solver);

#define GROW(ptr, n, cap, type)                                              \
    do {                                                                     \
        if ((n) >= (cap)) {                                                  \
            int64_t nc_ = (cap) ? 2 * (cap) : 16;                            \
            (ptr) = (type *)realloc((ptr), (size_t)nc_ * sizeof(type));      \
            (cap) = nc_;                                                     \
        }                                                                    \
    } while (0)

/* ---------------- nalgebra Vector2 helpers (restated) ---------------- */
static inline T(v2) F(vadd)(T(v2) a, T(v2) b) { T(v2) r = {a.x + b.x, a.y + b.y}; return r; }
static inline T(v2) F(vsub)(T(v2) a, T(v2) b) { T(v2) r = {a.x - b.x, a.y - b.y}; return r; }
static inline T(v2) F(vmul)(T(v2) a, REAL s) { T(v2) r = {a.x * s, a.y * s}; return r; }
static inline T(v2) F(vdiv)(T(v2) a, REAL s) { T(v2) r = {a.x / s, a.y / s}; return r; }
static inline REAL F(vdot)(T(v2) a, T(v2) b) { return a.x * b.x + a.y * b.y; }
static inline REAL F(vnorm2)(T(v2) a) { return F(vdot)(a, a); }
static inline REAL F(vmag)(T(v2) a) { return (REAL)SQRT(F(vnorm2)(a)); }
static inline T(v2) F(vnormalize)(T(v2) a) { return F(vdiv)(a, F(vmag)(a)); }

/* ---------------- construction: src/solver.rs:34-67 ---------------- */
F(solver) *F(new)(void) {
    F(solver) *s = (F(solver) *)calloc(1, sizeof(F(solver)));
    s->gravity.x = (REAL)0.0f; s->gravity.y = (REAL)98.2f;      /* src/solver.rs:36 */
    s->bpos.x = 0; s->bpos.y = 0; s->bsize.x = 100; s->bsize.y = 100; /* :37-40 */
    s->bounds_active = 1;                                        /* :41 */
    s->sub_steps = 1;                                            /* :47 */
    s->sub_steps_multiplier = 0;                                 /* :48 */
    s->link_iters = 1;
    s->cc_mode = 0;
    s->cc_margin = (REAL)0.25f;
    return s;
}

void F(free)(F(solver) *s) {
    if (!s) return;
    for (int64_t i = 0; i < s->npoly; i++) { free(s->polys[i].pts); free(s->polys[i].links); }
    for (int64_t i = 0; i < s->ncol; i++) { free(s->cols[i].nx); free(s->cols[i].ny); free(s->cols[i].c); }
    free(s->parts); free(s->pinned); free(s->plinks); free(s->circles); free(s->clinks);
    free(s->polys); free(s->cols); free(s);
}

static void *F(dup)(const void *p, size_t bytes) {
    if (!p || !bytes) return NULL;
    void *q = malloc(bytes); memcpy(q, p, bytes); return q;
}

/* #[derive(Clone)] on Solver: src/solver.rs:19 */
F(solver) *F(clone)(const F(solver) *s) {
    F(solver) *c = (F(solver) *)malloc(sizeof(F(solver)));
    *c = *s;
    c->parts = (T(particle) *)F(dup)(s->parts, sizeof(T(particle)) * (size_t)s->cap_p);
    c->pinned = (uint8_t *)F(dup)(s->pinned, (size_t)s->cap_p);
    c->plinks = (T(link) *)F(dup)(s->plinks, sizeof(T(link)) * (size_t)s->cap_pl);
    c->circles = (T(circle) *)F(dup)(s->circles, sizeof(T(circle)) * (size_t)s->cap_c);
    c->clinks = (T(link) *)F(dup)(s->clinks, sizeof(T(link)) * (size_t)s->cap_cl);
    c->polys = (T(polygon) *)F(dup)(s->polys, sizeof(T(polygon)) * (size_t)s->cap_poly);
    for (int64_t i = 0; i < s->npoly; i++) {
        c->polys[i].pts = (T(particle) *)F(dup)(s->polys[i].pts, sizeof(T(particle)) * (size_t)s->polys[i].n);
        c->polys[i].links = (T(link) *)F(dup)(s->polys[i].links, sizeof(T(link)) * (size_t)s->polys[i].nl);
    }
    c->cols = (T(collider) *)F(dup)(s->cols, sizeof(T(collider)) * (size_t)s->cap_col);
    for (int64_t i = 0; i < s->ncol; i++) {
        int64_t nv = s->cols[i].nv;
        c->cols[i].nx = (REAL *)F(dup)(s->cols[i].nx, sizeof(REAL) * (size_t)nv);
        c->cols[i].ny = (REAL *)F(dup)(s->cols[i].ny, sizeof(REAL) * (size_t)nv);
        c->cols[i].c = (REAL *)F(dup)(s->cols[i].c, sizeof(REAL) * (size_t)nv);
    }
    return c;
}

void F(set_gravity)(F(solver) *s, float x, float y) { s->gravity.x = x; s->gravity.y = y; }
void F(set_bounds)(F(solver) *s, float px, float py, float sx, float sy) {
    s->bpos.x = px; s->bpos.y = py; s->bsize.x = sx; s->bsize.y = sy;
}
void F(set_bounds_active)(F(solver) *s, int a) { s->bounds_active = a; }
int F(set_sub_steps)(F(solver) *s, int n) { if (n < 1 || n > 65535) return -1; s->sub_steps = n; return 0; }
int F(set_link_iterations)(F(solver) *s, int k) { if (k < 1) return -1; s->link_iters = k; return 0; }
void F(set_cc_mode)(F(solver) *s, int mode, float margin) { s->cc_mode = mode; s->cc_margin = margin; }
int64_t F(cc_redos)(const F(solver) *s) { return s->cc_redos; }

/* Particle::new: src/particle.rs:12-18 (prev = pos, acc = 0) */
static T(particle) F(particle_new)(REAL x, REAL y) {
    T(particle) p; p.pos.x = x; p.pos.y = y; p.prev = p.pos; p.acc.x = 0; p.acc.y = 0; return p;
}

/* Solver::add_particle: src/solver.rs:52-54 */
int64_t F(add_particle)(F(solver) *s, float x, float y) {
    if (s->np >= s->cap_p) {
        int64_t nc = s->cap_p ? 2 * s->cap_p : 16;
        s->parts = (T(particle) *)realloc(s->parts, sizeof(T(particle)) * (size_t)nc);
        s->pinned = (uint8_t *)realloc(s->pinned, (size_t)nc);
        s->cap_p = nc;
    }
    s->parts[s->np] = F(particle_new)(x, y);
    s->pinned[s->np] = 0;
    return s->np++;
}

/* EXT6: a solver particle with an explicit previous position (initial velocity). */
int64_t F(add_particle_ex)(F(solver) *s, float x, float y, float px, float py) {
    int64_t i = F(add_particle)(s, x, y);
    s->parts[i].prev.x = px; s->parts[i].prev.y = py;
    return i;
}

/* EXT3: pin flag for a solver particle. */
int F(pin)(F(solver) *s, int64_t i, int flag) {
    if (i < 0 || i >= s->np) return -1;
    s->pinned[i] = flag ? 1 : 0; return 0;
}

/* Solver::add_particle_link: src/solver.rs:62-64 (unchecked push) */
int64_t F(add_particle_link)(F(solver) *s, int64_t a, int64_t b, float t) {
    GROW(s->plinks, s->npl, s->cap_pl, T(link));
    s->plinks[s->npl].a = a; s->plinks[s->npl].b = b; s->plinks[s->npl].t = t;
    return s->npl++;
}

/* Solver::add_circle: src/solver.rs:55-57; Circle{point: Particle, radius} src/circle.rs:4-8 */
int64_t F(add_circle)(F(solver) *s, float x, float y, float px, float py, float ax, float ay, float r) {
    GROW(s->circles, s->nc, s->cap_c, T(circle));
    T(circle) *c = &s->circles[s->nc];
    c->pt.pos.x = x; c->pt.pos.y = y; c->pt.prev.x = px; c->pt.prev.y = py;
    c->pt.acc.x = ax; c->pt.acc.y = ay; c->r = r;
    return s->nc++;
}

/* Solver::add_circle_link: src/solver.rs:65-67 */
int64_t F(add_circle_link)(F(solver) *s, int64_t a, int64_t b, float t) {
    GROW(s->clinks, s->ncl, s->cap_cl, T(link));
    s->clinks[s->ncl].a = a; s->clinks[s->ncl].b = b; s->clinks[s->ncl].t = t;
    return s->ncl++;
}

/* Solver::add_polygon: src/solver.rs:58-60 (Polygon moved in whole) */
int64_t F(add_polygon)(F(solver) *s, int64_t n, const float *pos, const float *prev, const float *acc,
                       int64_t nl, const int64_t *la, const int64_t *lb, const float *lt,
                       int is_static, float cx, float cy, float scale) {
    GROW(s->polys, s->npoly, s->cap_poly, T(polygon));
    T(polygon) *p = &s->polys[s->npoly];
    p->n = n; p->pts = (T(particle) *)malloc(sizeof(T(particle)) * (size_t)(n ? n : 1));
    for (int64_t i = 0; i < n; i++) {
        p->pts[i].pos.x = pos[2 * i]; p->pts[i].pos.y = pos[2 * i + 1];
        p->pts[i].prev.x = prev ? prev[2 * i] : pos[2 * i];
        p->pts[i].prev.y = prev ? prev[2 * i + 1] : pos[2 * i + 1];
        p->pts[i].acc.x = acc ? acc[2 * i] : 0; p->pts[i].acc.y = acc ? acc[2 * i + 1] : 0;
    }
    p->nl = nl; p->links = (T(link) *)malloc(sizeof(T(link)) * (size_t)(nl ? nl : 1));
    for (int64_t i = 0; i < nl; i++) { p->links[i].a = la[i]; p->links[i].b = lb[i]; p->links[i].t = lt[i]; }
    p->is_static = is_static; p->center.x = cx; p->center.y = cy; p->scale = scale;
    return s->npoly++;
}

/* Polygon::circle: src/polygon.rs:17-82 (host ctor; f32 cos/sin from libm) */
int64_t F(add_polygon_circle)(F(solver) *s, float radius, float x, float y, int64_t n, int is_static) {
    float *pos = (float *)malloc(sizeof(float) * 2 * (size_t)(n ? n : 1));
    float cx = 0.0f, cy = 0.0f, angle = 0.0f;
    for (int64_t i = 0; i < n; i++) {
        float px = radius * cosf(angle);                      /* :23 */
        float py = radius * sinf(angle);                      /* :24 */
        pos[2 * i] = x + px; pos[2 * i + 1] = y + py;         /* :25 */
        cx = cx + pos[2 * i]; cy = cy + pos[2 * i + 1];       /* :27 */
        angle = angle + (2.0f * 3.14159265358979323846f) / (float)n; /* :28 */
    }
    cx = cx / (float)n; cy = cy / (float)n;                   /* :30 */
    int64_t *la = (int64_t *)malloc(sizeof(int64_t) * 2 * (size_t)(n ? n : 1));
    int64_t *lb = (int64_t *)malloc(sizeof(int64_t) * 2 * (size_t)(n ? n : 1));
    float *lt = (float *)malloc(sizeof(float) * 2 * (size_t)(n ? n : 1));
    int64_t nl = 0;
    for (int64_t i = 0; i < n; i++) {
        for (int part = 0; part < 2; part++) {                /* :33-52 then :53-72 */
            int64_t a = i;
            int64_t b = part == 0 ? (i + 2 * n / 3) % n : (i + n / 3) % n;
            if (a > b) { int64_t t = a; a = b; b = t; }
            float dx = pos[2 * a] - pos[2 * b], dy = pos[2 * a + 1] - pos[2 * b + 1];
            la[nl] = a; lb[nl] = b; lt[nl] = sqrtf(dx * dx + dy * dy);
            nl++;
        }
    }
    int64_t id = F(add_polygon)(s, n, pos, NULL, NULL, nl, la, lb, lt, is_static, cx, cy, 1.0f);
    free(pos); free(la); free(lb); free(lt);
    return id;
}

/* Polygon::new: src/polygon.rs:84-123 */
int64_t F(add_polygon_new)(F(solver) *s, int64_t n, const float *xy, int is_static) {
    float cx = 0.0f, cy = 0.0f;
    for (int64_t i = 0; i < n; i++) { cx = cx + xy[2 * i]; cy = cy + xy[2 * i + 1]; }  /* :88-92 */
    cx = cx / (float)n; cy = cy / (float)n;                                            /* :93 */
    int64_t *la = (int64_t *)malloc(sizeof(int64_t) * (size_t)(n ? n : 1));
    int64_t *lb = (int64_t *)malloc(sizeof(int64_t) * (size_t)(n ? n : 1));
    float *lt = (float *)malloc(sizeof(float) * (size_t)(n ? n : 1));
    for (int64_t i = 0; i < n; i++) {                                                   /* :95-114 */
        int64_t a = i, b = (i + 1) % n;
        if (a > b) { int64_t t = a; a = b; b = t; }
        float dx = xy[2 * a] - xy[2 * b], dy = xy[2 * a + 1] - xy[2 * b + 1];
        la[i] = a; lb[i] = b; lt[i] = sqrtf(dx * dx + dy * dy);
    }
    int64_t id = F(add_polygon)(s, n, xy, NULL, NULL, n, la, lb, lt, is_static, cx, cy, 1.0f);
    free(la); free(lb); free(lt);
    return id;
}

/* EXT4: static colliders.  kind 0 = circle (cx, cy, R); kind 1 = convex polygon.
 * Polygon edges are oriented counter-clockwise (a clockwise input is reversed);
 * per edge i (v_i -> v_i+1): e = v_i+1 - v_i, len = sqrt(e.x*e.x + e.y*e.y),
 * n = (e.y/len, -e.x/len) (outward), c = n.x*v_i.x + n.y*v_i.y.  Values are
 * computed in f32 and then held in REAL, so the f64 twin uses the same planes. */
int64_t F(add_static_circle)(F(solver) *s, float x, float y, float r) {
    GROW(s->cols, s->ncol, s->cap_col, T(collider));
    T(collider) *c = &s->cols[s->ncol];
    memset(c, 0, sizeof(*c));
    c->kind = 0; c->cx = x; c->cy = y; c->r = r;
    return s->ncol++;
}

int64_t F(add_static_polygon)(F(solver) *s, int64_t nv, const float *xy) {
    if (nv < 3) return -1;
    double area2 = 0.0;
    for (int64_t i = 0; i < nv; i++) {
        int64_t j = (i + 1) % nv;
        area2 += (double)xy[2 * i] * xy[2 * j + 1] - (double)xy[2 * j] * xy[2 * i + 1];
    }
    GROW(s->cols, s->ncol, s->cap_col, T(collider));
    T(collider) *c = &s->cols[s->ncol];
    memset(c, 0, sizeof(*c));
    c->kind = 1; c->nv = nv;
    c->nx = (REAL *)malloc(sizeof(REAL) * (size_t)nv);
    c->ny = (REAL *)malloc(sizeof(REAL) * (size_t)nv);
    c->c = (REAL *)malloc(sizeof(REAL) * (size_t)nv);
    for (int64_t k = 0; k < nv; k++) {
        int64_t i = area2 >= 0 ? k : nv - 1 - k;
        int64_t j = area2 >= 0 ? (k + 1) % nv : (nv - 1 - ((k + 1) % nv));
        float vx = xy[2 * i], vy = xy[2 * i + 1];
        float ex = xy[2 * j] - vx, ey = xy[2 * j + 1] - vy;
        float len = sqrtf(ex * ex + ey * ey);
        float nx = ey / len, ny = (-ex) / len;
        c->nx[k] = nx; c->ny[k] = ny; c->c[k] = nx * vx + ny * vy;
    }
    return s->ncol++;
}

/* ---------------- per-body primitives ---------------- */

/* Particle::update: src/particle.rs:20-25 */
static inline void F(particle_update)(T(particle) *p, REAL dt) {
    T(v2) vel = F(vsub)(p->pos, p->prev);
    p->prev = p->pos;
    p->pos = F(vadd)(F(vadd)(p->pos, vel), F(vmul)(F(vmul)(p->acc, dt), dt));
    p->acc.x = 0; p->acc.y = 0;
}

/* Particle::solve_bounds: src/particle.rs:27-46 */
static inline void F(particle_bounds)(T(particle) *p, T(v2) bp, T(v2) bs) {
    if (p->pos.x < bp.x) {
        REAL v = p->prev.x - p->pos.x; p->prev.x = bp.x - v; p->pos.x = bp.x;
    } else if (p->pos.x > bp.x + bs.x) {
        REAL v = p->prev.x - p->pos.x; p->prev.x = (bp.x + bs.x) - v; p->pos.x = bp.x + bs.x;
    }
    if (p->pos.y < bp.y) {
        REAL v = p->prev.y - p->pos.y; p->prev.y = bp.y - v; p->pos.y = bp.y;
    } else if (p->pos.y > bp.y + bs.y) {
        REAL v = p->prev.y - p->pos.y; p->prev.y = (bp.y + bs.y) - v; p->pos.y = bp.y + bs.y;
    }
}

/* Circle::solve_bounds: src/circle.rs:11-30 */
static inline void F(circle_bounds)(T(circle) *c, T(v2) bp, T(v2) bs) {
    T(particle) *p = &c->pt; REAL r = c->r;
    if (p->pos.x < bp.x + r) {
        REAL v = p->prev.x - p->pos.x; p->prev.x = (bp.x + r) - v; p->pos.x = bp.x + r;
    } else if (p->pos.x > (bp.x + bs.x) - r) {
        REAL v = p->prev.x - p->pos.x; p->prev.x = ((bp.x + bs.x) - r) - v; p->pos.x = (bp.x + bs.x) - r;
    }
    if (p->pos.y < bp.y + r) {
        REAL v = p->prev.y - p->pos.y; p->prev.y = (bp.y + r) - v; p->pos.y = bp.y + r;
    } else if (p->pos.y > (bp.y + bs.y) - r) {
        REAL v = p->prev.y - p->pos.y; p->prev.y = ((bp.y + bs.y) - r) - v; p->pos.y = (bp.y + bs.y) - r;
    }
}

/* ParticleLink::solve: src/link.rs:18-27, with the EXT3 pin rule: a pinned end
 * does not move and the free end takes the whole correction (factor 1);
 * two free ends keep the reference's factor 0.5, two pinned ends do nothing. */
static inline void F(plink_solve)(T(v2) *a, T(v2) *b, REAL t, int pa, int pb) {
    T(v2) d = F(vsub)(*a, *b);
    REAL dist = F(vmag)(d);
    T(v2) n = F(vnormalize)(d);
    REAL e = dist - t;
    if (!pa && !pb) {
        T(v2) c = F(vmul)(F(vmul)(n, e), (REAL)0.5f);
        *a = F(vsub)(*a, c);
        *b = F(vadd)(*b, c);
    } else if (!pa) {
        *a = F(vsub)(*a, F(vmul)(n, e));
    } else if (!pb) {
        *b = F(vadd)(*b, F(vmul)(n, e));
    }
}

/* CircleLink::solve: src/link.rs:36-48 */
static inline void F(clink_solve)(T(circle) *ca, T(circle) *cb, REAL t) {
    T(v2) d = F(vsub)(ca->pt.pos, cb->pt.pos);
    REAL dist = F(vmag)(d);
    T(v2) n = F(vnormalize)(d);
    REAL ra2 = ca->r * ca->r, rb2 = cb->r * cb->r;
    REAL scale = (REAL)1.0f / (ra2 + rb2);
    REAL e = dist - t;
    ca->pt.pos = F(vsub)(ca->pt.pos, F(vmul)(F(vmul)(F(vmul)(n, e), scale), rb2));
    cb->pt.pos = F(vadd)(cb->pt.pos, F(vmul)(F(vmul)(F(vmul)(n, e), scale), ra2));
}

/* Circle::solve_circle: src/circle.rs:32-45.  Returns the length of the push
 * applied to each circle (for the candidate-list margin check), 0 if none. */
static inline int F(circle_pair)(T(circle) *ci, T(circle) *cj, REAL *push_i, REAL *push_j) {
    T(v2) d = F(vsub)(ci->pt.pos, cj->pt.pos);
    REAL q = F(vnorm2)(d);
    REAL rs = ci->r + cj->r;
    if (q < rs * rs) {
        T(v2) n = F(vnormalize)(d);
        REAL ov = rs - (REAL)SQRT(q);
        REAL ri2 = ci->r * ci->r, rj2 = cj->r * cj->r;
        REAL scale = (REAL)1.0f / (ri2 + rj2);
        T(v2) di = F(vmul)(F(vmul)(F(vmul)(n, scale), ov), rj2);
        T(v2) dj = F(vmul)(F(vmul)(F(vmul)(n, scale), ov), ri2);
        ci->pt.pos = F(vadd)(ci->pt.pos, di);
        cj->pt.pos = F(vsub)(cj->pt.pos, dj);
        if (push_i) *push_i = (REAL)fabs((double)(scale * ov * rj2));
        if (push_j) *push_j = (REAL)fabs((double)(scale * ov * ri2));
        return 1;
    }
    return 0;
}

/* common::line_intersection: src/common.rs:4-26 */
static inline int F(line_intersection)(T(v2) p1, T(v2) p2, T(v2) p3, T(v2) p4, T(v2) *out) {
    REAL s1x = p2.x - p1.x, s1y = p2.y - p1.y;
    REAL s2x = p4.x - p3.x, s2y = p4.y - p3.y;
    REAL den = (-s2x) * s1y + s1x * s2y;
    REAL s = ((-s1y) * (p1.x - p3.x) + s1x * (p1.y - p3.y)) / den;
    REAL t = (s2x * (p1.y - p3.y) - s2y * (p1.x - p3.x)) / den;
    if (s >= 0 && s <= 1 && t >= 0 && t <= 1) {
        out->x = p1.x + (t * s1x); out->y = p1.y + (t * s1y);
        return 1;
    }
    return 0;
}

/* Polygon::resolve_line_intersection: src/polygon.rs:164-216 */
static int F(resolve_line)(T(v2) self_center, T(v2) pa, T(v2) pb, T(v2) op, T(v2) other_center,
                           T(v2) *new_a, T(v2) *new_b, T(v2) *new_o) {
    T(v2) inter;
    if (!F(line_intersection)(pa, pb, op, other_center, &inter)) return 0;
    T(v2) nl = F(vnormalize)(F(vsub)(pb, pa));                                         /* :175 */
    T(v2) cp = F(vmul)(nl, F(vdot)(nl, F(vsub)(self_center, inter)) / F(vdot)(nl, nl)); /* :177-179 */
    T(v2) nin = F(vnormalize)(F(vsub)(self_center, F(vadd)(inter, cp)));              /* :181 */
    REAL da = F(vmag)(F(vsub)(inter, pa));                                             /* :183 */
    REAL db = F(vmag)(F(vsub)(inter, pb));                                             /* :184 */
    REAL dab = da + db;                                                                /* :186 */
    REAL ia = db / dab, ib = da / dab;                                                 /* :188-189 */
    T(v2) diff = F(vsub)(inter, op);                                                   /* :191 */
    T(v2) ion = F(vmul)(nin, F(vdot)(nin, diff) / F(vdot)(nin, nin));                 /* :193-194 */
    T(v2) third = F(vdiv)(ion, (REAL)3.0f);                                            /* :196 */
    T(v2) line = F(vmul)(third, (REAL)2.0f);                                           /* :198 */
    T(v2) dispa = F(vmul)(line, ia), dispb = F(vmul)(line, ib);                        /* :200-201 */
    T(v2) na = F(vsub)(pa, dispa), nb = F(vsub)(pb, dispb);                            /* :203-204 */
    T(v2) ray_end = F(vsub)(op, F(vmul)(nin, (REAL)10000.0f));                        /* :208 */
    T(v2) np;
    if (!F(line_intersection)(pa, pb, op, ray_end, &np)) return 0;                    /* :206-213 */
    *new_a = na; *new_b = nb; *new_o = np;
    return 1;
}

/* Polygon::solve_polygon_single: src/polygon.rs:147-162 */
static void F(polygon_single)(T(polygon) *self, T(polygon) *other) {
    int64_t n = self->n;
    for (int64_t i = 0; i < n; i++) {
        T(v2) pa = self->pts[i].pos;                 /* stale copy, :149 */
        int64_t bid = (i + 1) % n;
        T(v2) pb = self->pts[bid].pos;               /* stale copy, :151 */
        for (int64_t k = 0; k < other->n; k++) {
            T(v2) na, nb, no;
            if (F(resolve_line)(self->center, pa, pb, other->pts[k].pos, other->center, &na, &nb, &no)) {
                self->pts[i].pos = na;
                self->pts[bid].pos = nb;
                other->pts[k].pos = no;
            }
        }
    }
}

/* Polygon::calc_center: src/polygon.rs:231-237 */
static void F(calc_center)(T(polygon) *p) {
    T(v2) c = {0, 0};
    for (int64_t i = 0; i < p->n; i++) c = F(vadd)(c, p->pts[i].pos);
    p->center = F(vdiv)(c, (REAL)(float)p->n);
}

/* EXT4: a body of radius r (0 for particles and polygon points) against the
 * static colliders, in insertion order. */
static inline void F(static_collide)(const F(solver) *s, T(v2) *p, REAL r) {
    for (int64_t k = 0; k < s->ncol; k++) {
        const T(collider) *c = &s->cols[k];
        if (c->kind == 0) {
            REAL dx = p->x - c->cx, dy = p->y - c->cy;
            REAL q = dx * dx + dy * dy;
            REAL rs = r + c->r;
            if (q < rs * rs) {
                REAL d = (REAL)SQRT(q);
                if (d > 0) {
                    REAL nx = dx / d, ny = dy / d, ov = rs - d;
                    p->x = p->x + nx * ov; p->y = p->y + ny * ov;
                }
            }
        } else {
            REAL smax = 0; int64_t arg = -1;
            for (int64_t e = 0; e < c->nv; e++) {
                REAL sd = (c->nx[e] * p->x + c->ny[e] * p->y) - c->c[e];
                if (arg < 0 || sd > smax) { smax = sd; arg = e; }
            }
            if (smax < r) {
                REAL push = r - smax;
                p->x = p->x + c->nx[arg] * push; p->y = p->y + c->ny[arg] * push;
            }
        }
    }
}

/* ---------------- validation (the build checks links before mutating) ---------------- */
int F(validate)(const F(solver) *s) {
    for (int64_t i = 0; i < s->npl; i++) {
        const T(link) *l = &s->plinks[i];
        if (l->a < 0 || l->b >= s->np || l->a >= l->b) return -1;   /* src/link.rs:19-21 panics */
    }
    for (int64_t i = 0; i < s->ncl; i++) {
        const T(link) *l = &s->clinks[i];
        if (l->a < 0 || l->b >= s->nc || l->a >= l->b) return -1;   /* src/link.rs:37-39 panics */
    }
    for (int64_t p = 0; p < s->npoly; p++) {
        for (int64_t i = 0; i < s->polys[p].nl; i++) {
            const T(link) *l = &s->polys[p].links[i];
            if (l->a < 0 || l->b >= s->polys[p].n || l->a >= l->b) return -1;
        }
    }
    return 0;
}

/* ---------------- phases: src/solver.rs:118-188 ---------------- */

/* apply_links: src/solver.rs:143-153 */
static void F(apply_links)(F(solver) *s) {
    for (int64_t i = 0; i < s->npl; i++) {
        T(link) *l = &s->plinks[i];
        F(plink_solve)(&s->parts[l->a].pos, &s->parts[l->b].pos, l->t, s->pinned[l->a], s->pinned[l->b]);
    }
    for (int64_t i = 0; i < s->ncl; i++) {
        T(link) *l = &s->clinks[i];
        F(clink_solve)(&s->circles[l->a], &s->circles[l->b], l->t);
    }
    for (int64_t p = 0; p < s->npoly; p++) {                  /* Polygon::solve_links src/polygon.rs:218-223 */
        T(polygon) *pg = &s->polys[p];
        F(calc_center)(pg);
        for (int64_t i = 0; i < pg->nl; i++) {
            T(link) *l = &pg->links[i];
            F(plink_solve)(&pg->pts[l->a].pos, &pg->pts[l->b].pos, l->t, 0, 0);
        }
    }
}

/* Circle pairs in (i, j), i < j lexicographic order with live positions:
 * src/solver.rs:168-177, the reference's O(C^2) loop. */
static void F(circles_all_pairs)(F(solver) *s) {
    for (int64_t i = 0; i < s->nc; i++)
        for (int64_t j = i + 1; j < s->nc; j++)
            F(circle_pair)(&s->circles[i], &s->circles[j], NULL, NULL);
}

/* Timing helper for the CPU baseline: the reference's all-pairs loop
 * (src/solver.rs:168-177) for rows i in [first, last) only, on the solver's
 * circles (which it moves).  Returns the pair tests made, so a caller can time
 * a bounded slice of the loop and extrapolate to its nc (nc - 1) / 2 tests. */
int64_t F(all_pairs_rows)(F(solver) *s, int64_t first, int64_t last) {
    int64_t tests = 0;
    for (int64_t i = first < 0 ? 0 : first; i < s->nc && i < last; i++)
        for (int64_t j = i + 1; j < s->nc; j++) {
            F(circle_pair)(&s->circles[i], &s->circles[j], NULL, NULL);
            tests++;
        }
    return tests;
}

/* The same sweep restricted to candidate pairs (distance at sweep start below
 * r_i + r_j + margin), processed in the same lexicographic order.  A pair left
 * out can only matter if a circle moves by margin/2 or more during the sweep;
 * the sweep records each circle's total push length and is redone with a larger
 * margin when that bound fails, so the result equals the all-pairs loop. */
typedef struct { int64_t i, j; } F(pair);
static int F(pair_cmp)(const void *a, const void *b) {
    const F(pair) *x = (const F(pair) *)a, *y = (const F(pair) *)b;
    if (x->i != y->i) return x->i < y->i ? -1 : 1;
    if (x->j != y->j) return x->j < y->j ? -1 : 1;
    return 0;
}
typedef struct { int64_t key, i; } F(keyed);
static int F(keyed_cmp)(const void *a, const void *b) {
    const F(keyed) *x = (const F(keyed) *)a, *y = (const F(keyed) *)b;
    if (x->key != y->key) return x->key < y->key ? -1 : 1;
    return x->i < y->i ? -1 : (x->i > y->i);
}
static int64_t F(lower_key)(const F(keyed) *v, int64_t n, int64_t key) {
    int64_t lo = 0, hi = n;
    while (lo < hi) { int64_t m = (lo + hi) / 2; if (v[m].key < key) lo = m + 1; else hi = m; }
    return lo;
}

static void F(circles_candidates)(F(solver) *s) {
    int64_t nc = s->nc;
    if (nc < 2) return;
    T(v2) *save = (T(v2) *)malloc(sizeof(T(v2)) * (size_t)nc);
    REAL *path = (REAL *)malloc(sizeof(REAL) * (size_t)nc);
    uint8_t *inert = (uint8_t *)malloc((size_t)nc);
    REAL rmax = 0;
    int huge = 0;
    for (int64_t i = 0; i < nc; i++) {
        save[i] = s->circles[i].pt.pos;
        if (s->circles[i].r > rmax) rmax = s->circles[i].r;
        /* A circle with a NaN or infinite coordinate takes part in no pair:
         * every difference with it is NaN or infinite, so q < rs*rs is false
         * (src/circle.rs:145).  It stays out of the grid. */
        inert[i] = !isfinite((double)save[i].x) || !isfinite((double)save[i].y);
        if (!inert[i] && (!(fabs((double)save[i].x) < 1e15) || !(fabs((double)save[i].y) < 1e15))) huge = 1;
    }
    if (huge) { F(circles_all_pairs)(s); free(save); free(path); free(inert); return; }
    REAL margin = s->cc_margin > 0 ? s->cc_margin : (REAL)0.25f;
    F(keyed) *keys = (F(keyed) *)malloc(sizeof(F(keyed)) * (size_t)nc);
    for (int attempt = 0;; attempt++) {
        double cell = 2.0 * (double)rmax + (double)margin;
        if (!(cell > 0)) cell = 1.0;
        int64_t nk = 0;
        for (int64_t i = 0; i < nc; i++) {
            if (inert[i]) continue;
            int64_t cx = (int64_t)floor((double)save[i].x / cell), cy = (int64_t)floor((double)save[i].y / cell);
            keys[nk].key = (cy << 32) + cx; keys[nk].i = i;   /* |cx|, |cy| < 2^31 for |pos| < 1e15 / cell */
            nk++;
        }
        qsort(keys, (size_t)nk, sizeof(F(keyed)), F(keyed_cmp));
        int64_t npairs = 0, cap = nc * 4 + 16;
        F(pair) *pairs = (F(pair) *)malloc(sizeof(F(pair)) * (size_t)cap);
        for (int64_t i = 0; i < nc; i++) {
            if (inert[i]) continue;
            int64_t cx = (int64_t)floor((double)save[i].x / cell), cy = (int64_t)floor((double)save[i].y / cell);
            for (int64_t yy = cy - 1; yy <= cy + 1; yy++) {
                for (int64_t xx = cx - 1; xx <= cx + 1; xx++) {
                    int64_t key = (yy << 32) + xx;
                    for (int64_t k = F(lower_key)(keys, nk, key); k < nk && keys[k].key == key; k++) {
                        int64_t j = keys[k].i;
                        if (j <= i) continue;
                        double dx = (double)save[i].x - save[j].x, dy = (double)save[i].y - save[j].y;
                        double lim = (double)s->circles[i].r + s->circles[j].r + margin;
                        if (dx * dx + dy * dy < lim * lim) {
                            if (npairs >= cap) { cap *= 2; pairs = (F(pair) *)realloc(pairs, sizeof(F(pair)) * (size_t)cap); }
                            pairs[npairs].i = i; pairs[npairs].j = j; npairs++;
                        }
                    }
                }
            }
        }
        qsort(pairs, (size_t)npairs, sizeof(F(pair)), F(pair_cmp));
        for (int64_t i = 0; i < nc; i++) path[i] = 0;
        for (int64_t k = 0; k < npairs; k++) {
            REAL pi_ = 0, pj_ = 0;
            if (F(circle_pair)(&s->circles[pairs[k].i], &s->circles[pairs[k].j], &pi_, &pj_)) {
                path[pairs[k].i] += pi_; path[pairs[k].j] += pj_;
            }
        }
        double pmax = 0;
        for (int64_t i = 0; i < nc; i++) {
            double v = (double)path[i];
            if (!(v <= pmax)) pmax = (v == v) ? v : INFINITY;
        }
        free(pairs);
        if (pmax <= 0.45 * (double)margin) {
            if (pmax < 0.1 * (double)margin && margin > (REAL)0.05f) margin = margin * (REAL)0.5f;
            break;
        }
        s->cc_redos++;
        for (int64_t i = 0; i < nc; i++) s->circles[i].pt.pos = save[i];
        if (!(pmax < 1e30) || attempt > 8) { F(circles_all_pairs)(s); break; }
        margin = (REAL)(2.5 * pmax + (double)margin);
    }
    s->cc_margin = margin;
    free(keys); free(save); free(path); free(inert);
}

/* solve_dynamic_collisions: src/solver.rs:167-188, then EXT4 static colliders */
static void F(dynamic_collisions)(F(solver) *s) {
    if (s->cc_mode == 1) F(circles_candidates)(s); else F(circles_all_pairs)(s);
    for (int64_t i = 0; i < s->npoly; i++)
        for (int64_t j = i + 1; j < s->npoly; j++) {
            F(polygon_single)(&s->polys[i], &s->polys[j]);    /* Polygon::solve_polygon :142-145 */
            F(polygon_single)(&s->polys[j], &s->polys[i]);
        }
    if (s->ncol) {
        for (int64_t i = 0; i < s->np; i++) if (!s->pinned[i]) F(static_collide)(s, &s->parts[i].pos, 0);
        for (int64_t i = 0; i < s->nc; i++) F(static_collide)(s, &s->circles[i].pt.pos, s->circles[i].r);
        for (int64_t p = 0; p < s->npoly; p++)
            for (int64_t i = 0; i < s->polys[p].n; i++) F(static_collide)(s, &s->polys[p].pts[i].pos, 0);
    }
}

/* solve_boundary_collisions: src/solver.rs:155-165 (bounds_active is not read) */
static void F(boundary)(F(solver) *s) {
    for (int64_t i = 0; i < s->np; i++) if (!s->pinned[i]) F(particle_bounds)(&s->parts[i], s->bpos, s->bsize);
    for (int64_t i = 0; i < s->nc; i++) F(circle_bounds)(&s->circles[i], s->bpos, s->bsize);
    for (int64_t p = 0; p < s->npoly; p++)                    /* Polygon::solve_bounds src/polygon.rs:136-140 */
        for (int64_t i = 0; i < s->polys[p].n; i++) F(particle_bounds)(&s->polys[p].pts[i], s->bpos, s->bsize);
}

/* apply_gravity: src/solver.rs:130-141 */
static void F(gravity)(F(solver) *s) {
    T(v2) g = s->gravity;
    for (int64_t i = 0; i < s->np; i++) s->parts[i].acc = F(vadd)(s->parts[i].acc, g);
    for (int64_t i = 0; i < s->nc; i++) s->circles[i].pt.acc = F(vadd)(s->circles[i].pt.acc, g);
    for (int64_t p = 0; p < s->npoly; p++)                    /* Polygon::add_force_v2 src/polygon.rs:225-229 */
        for (int64_t i = 0; i < s->polys[p].n; i++) s->polys[p].pts[i].acc = F(vadd)(s->polys[p].pts[i].acc, g);
}

/* update_positions: src/solver.rs:118-128 */
static void F(positions)(F(solver) *s, REAL dt) {
    for (int64_t i = 0; i < s->np; i++) {
        if (s->pinned[i]) { s->parts[i].acc.x = 0; s->parts[i].acc.y = 0; continue; }
        F(particle_update)(&s->parts[i], dt);
    }
    for (int64_t i = 0; i < s->nc; i++) F(particle_update)(&s->circles[i].pt, dt);
    for (int64_t p = 0; p < s->npoly; p++) {                  /* Polygon::update src/polygon.rs:125-134 */
        T(polygon) *pg = &s->polys[p];
        if (pg->is_static) continue;
        F(calc_center)(pg);
        for (int64_t i = 0; i < pg->n; i++) F(particle_update)(&pg->pts[i], dt);
    }
}

/* Solver::update: src/solver.rs:106-116 (+ EXT1 sub_steps, EXT2 link iterations) */
int F(update)(F(solver) *s, float dt) {
    if (F(validate)(s)) return -1;
    s->sub_steps_multiplier = (REAL)1.0f / (REAL)(float)s->sub_steps;   /* :107 */
    REAL delta = (REAL)dt * s->sub_steps_multiplier;                     /* :108 */
    for (int st = 0; st < s->sub_steps; st++) {
        F(gravity)(s);
        for (int it = 0; it < s->link_iters; it++) F(apply_links)(s);
        F(dynamic_collisions)(s);
        F(boundary)(s);
        F(positions)(s, delta);
    }
    return 0;
}

/* One sweep of the solver's particle links in a caller-given order (used to
 * check that a dependency-layer schedule reproduces the sequential sweep). */
int F(apply_particle_links_in_order)(F(solver) *s, const int64_t *order, int64_t n) {
    for (int64_t k = 0; k < n; k++) {
        int64_t i = order[k];
        if (i < 0 || i >= s->npl) return -1;
        T(link) *l = &s->plinks[i];
        F(plink_solve)(&s->parts[l->a].pos, &s->parts[l->b].pos, l->t, s->pinned[l->a], s->pinned[l->b]);
    }
    return 0;
}

/* ---------------- single laws, for the known-answer and model tests ----------------
 * Each runs exactly the function the solver step uses, on caller values, so a
 * test can pin one law at a time.  Vectors are (x, y) pairs of REAL. */

/* ParticleLink::solve (src/link.rs:78-87) with the EXT3 pin rule; ab = {ax, ay, bx, by} in/out */
void F(law_particle_link)(REAL *ab, REAL t, int pa, int pb) {
    T(v2) a = {ab[0], ab[1]}, b = {ab[2], ab[3]};
    F(plink_solve)(&a, &b, t, pa, pb);
    ab[0] = a.x; ab[1] = a.y; ab[2] = b.x; ab[3] = b.y;
}

/* CircleLink::solve (src/link.rs:96-108); ab = {ax, ay, bx, by} in/out */
void F(law_circle_link)(REAL *ab, REAL ra, REAL rb, REAL t) {
    T(circle) ca, cb;
    memset(&ca, 0, sizeof ca); memset(&cb, 0, sizeof cb);
    ca.pt.pos.x = ab[0]; ca.pt.pos.y = ab[1]; ca.r = ra;
    cb.pt.pos.x = ab[2]; cb.pt.pos.y = ab[3]; cb.r = rb;
    F(clink_solve)(&ca, &cb, t);
    ab[0] = ca.pt.pos.x; ab[1] = ca.pt.pos.y; ab[2] = cb.pt.pos.x; ab[3] = cb.pt.pos.y;
}

/* Circle::solve_circle (src/circle.rs:141-154); returns 1 when the pair overlapped */
int F(law_circle_pair)(REAL *ab, REAL ra, REAL rb) {
    T(circle) ca, cb;
    memset(&ca, 0, sizeof ca); memset(&cb, 0, sizeof cb);
    ca.pt.pos.x = ab[0]; ca.pt.pos.y = ab[1]; ca.r = ra;
    cb.pt.pos.x = ab[2]; cb.pt.pos.y = ab[3]; cb.r = rb;
    int hit = F(circle_pair)(&ca, &cb, NULL, NULL);
    ab[0] = ca.pt.pos.x; ab[1] = ca.pt.pos.y; ab[2] = cb.pt.pos.x; ab[3] = cb.pt.pos.y;
    return hit;
}

/* Particle::solve_bounds (src/particle.rs:27-46) or, with is_circle, Circle::solve_bounds
 * (src/circle.rs:120-139); pp = {pos.x, pos.y, prev.x, prev.y} in/out, b = {pos.x, pos.y, size.x, size.y} */
void F(law_bounds)(REAL *pp, const REAL *b, int is_circle, REAL r) {
    T(v2) bp = {b[0], b[1]}, bs = {b[2], b[3]};
    T(circle) c;
    memset(&c, 0, sizeof c);
    c.pt.pos.x = pp[0]; c.pt.pos.y = pp[1]; c.pt.prev.x = pp[2]; c.pt.prev.y = pp[3]; c.r = r;
    if (is_circle) F(circle_bounds)(&c, bp, bs); else F(particle_bounds)(&c.pt, bp, bs);
    pp[0] = c.pt.pos.x; pp[1] = c.pt.pos.y; pp[2] = c.pt.prev.x; pp[3] = c.pt.prev.y;
}

/* Particle::update (src/particle.rs:20-25); p = {pos.x, pos.y, prev.x, prev.y, acc.x, acc.y} in/out */
void F(law_particle_update)(REAL *p, REAL dt) {
    T(particle) q;
    q.pos.x = p[0]; q.pos.y = p[1]; q.prev.x = p[2]; q.prev.y = p[3]; q.acc.x = p[4]; q.acc.y = p[5];
    F(particle_update)(&q, dt);
    p[0] = q.pos.x; p[1] = q.pos.y; p[2] = q.prev.x; p[3] = q.prev.y; p[4] = q.acc.x; p[5] = q.acc.y;
}

/* common::line_intersection (src/common.rs:4-26); p = {p1, p2, p3, p4} as 8 REALs; 1 = Some(out) */
int F(law_line_intersection)(const REAL *p, REAL *out) {
    T(v2) p1 = {p[0], p[1]}, p2 = {p[2], p[3]}, p3 = {p[4], p[5]}, p4 = {p[6], p[7]}, r = {0, 0};
    int hit = F(line_intersection)(p1, p2, p3, p4, &r);
    out[0] = r.x; out[1] = r.y;
    return hit;
}

/* Polygon::resolve_line_intersection (src/polygon.rs:164-216) for an edge a-b
 * of a polygon with centre c against a point o of a polygon with centre oc:
 * v = {c, a, b, o, oc} as 10 REALs; out = {new_a, new_b, new_o}; 1 = Some */
int F(law_resolve_line)(const REAL *v, REAL *out) {
    T(v2) c = {v[0], v[1]}, a = {v[2], v[3]}, b = {v[4], v[5]}, o = {v[6], v[7]}, oc = {v[8], v[9]};
    T(v2) na = {0, 0}, nb = {0, 0}, no = {0, 0};
    int hit = F(resolve_line)(c, a, b, o, oc, &na, &nb, &no);
    out[0] = na.x; out[1] = na.y; out[2] = nb.x; out[3] = nb.y; out[4] = no.x; out[5] = no.y;
    return hit;
}

/* ---------------- readback (values widened to double, exact for f32) ---------------- */
int64_t F(get_particle_len)(const F(solver) *s) { return s->np; }
int64_t F(get_circles_len)(const F(solver) *s) { return s->nc; }
int64_t F(get_polygons_len)(const F(solver) *s) { return s->npoly; }
int64_t F(get_polygon_point_len)(const F(solver) *s, int64_t p) { return (p >= 0 && p < s->npoly) ? s->polys[p].n : -1; }

static void F(put_particle)(double *o, const T(particle) *p) {
    o[0] = p->pos.x; o[1] = p->pos.y; o[2] = p->prev.x; o[3] = p->prev.y; o[4] = p->acc.x; o[5] = p->acc.y;
}
/* 6 doubles per particle: pos.xy, prev.xy, acc.xy */
void F(get_particles)(const F(solver) *s, double *out) {
    for (int64_t i = 0; i < s->np; i++) F(put_particle)(out + 6 * i, &s->parts[i]);
}
/* 7 doubles per circle: pos.xy, prev.xy, acc.xy, radius */
void F(get_circles)(const F(solver) *s, double *out) {
    for (int64_t i = 0; i < s->nc; i++) { F(put_particle)(out + 7 * i, &s->circles[i].pt); out[7 * i + 6] = s->circles[i].r; }
}
/* 6 doubles per point of polygon p; center.xy into center[2] */
int F(get_polygon)(const F(solver) *s, int64_t p, double *out, double *center) {
    if (p < 0 || p >= s->npoly) return -1;
    for (int64_t i = 0; i < s->polys[p].n; i++) F(put_particle)(out + 6 * i, &s->polys[p].pts[i]);
    center[0] = s->polys[p].center.x; center[1] = s->polys[p].center.y;
    return 0;
}
int64_t F(get_polygon_links)(const F(solver) *s, int64_t p, int64_t *ab, double *t) {
    if (p < 0 || p >= s->npoly) return -1;
    for (int64_t i = 0; i < s->polys[p].nl; i++) { ab[2 * i] = s->polys[p].links[i].a; ab[2 * i + 1] = s->polys[p].links[i].b; t[i] = s->polys[p].links[i].t; }
    return s->polys[p].nl;
}

#undef GROW
#undef CAT_
#undef CAT
#undef F
#undef T
