/*
 * ORACLE -- test infrastructure only.  Nothing in the product library links,
 * loads or calls this file; only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py use it, and only as the checker / CPU baseline.
 *
 * A plain-C restatement of the reference's HOST ghost-cell tree builder, its
 * interaction lists and its point locator, with 0-based indices (-1 = none):
 *
 *   split_nodes! / _sort_range! (stable counting sort by child slot)
 *                                            src/tree/builder.jl:45-131, 163-190
 *   tag_colleagues! / tag_root! (parent rule) src/tree/builder.jl:212-253
 *   compute_depths!                           src/tree/builder.jl:260-280
 *   level_restrict!                           src/tree/builder.jl:289-316
 *   add_ghosts!                               src/tree/builder.jl:328-351
 *   build_levels                              src/tree/builder.jl:359-384
 *   flatten (+ group_parent)                  src/tree/builder.jl:392-436, 447-473
 *   _count_pairs! / _fill_pairs!              src/tree/lists.jl:73-124
 *   locate                                    src/tree/locate.jl:10-22
 *   child_slot / _center                      src/util.jl:39-45, src/tree/tree.jl:81-89
 *
 * Box centres are formed exactly as the reference forms them
 * (origin + (cell + 0.5) * w, no contraction: build with -ffp-contract=off), so a
 * point's child slot here is the one the reference computes.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <stdio.h>
#include <math.h>

#define NC 8
#define NCOLL 27
#define SELF 13

typedef struct {
    int32_t n, cap;
    int32_t *cell;       /* 3n */
    int32_t *parent, *child_first, *lo, *hi, *depth;
    uint8_t *leaf, *ghost, *xlist;
    int32_t *col;        /* 27n */
} Level;

typedef struct {
    int nlevels, nnodes;
    int64_t npoints;
    double origin[3], W0;
    int32_t *level_start;              /* nlevels + 1 */
    int32_t *cell, *level, *parent, *child_first, *plo, *phi, *col;
    uint8_t *leaf, *ghost, *xlist, *has_source;
    int32_t *perm, *point_node;
    double *pts;                        /* sorted points, 3N */
    /* per-level lists, targets in node order: ptr (nl + 1), src, far */
    int32_t **lptr, **lsrc, **lfar;
    int64_t *lnp;
} OTree;

static char g_err[512];
const char *otree_last_error(void) { return g_err; }

static void *xrealloc(void *p, size_t n) {
    void *q = realloc(p, n ? n : 1);
    if (!q) { fprintf(stderr, "oracle: out of memory\n"); abort(); }
    return q;
}

static void level_reserve(Level *L, int32_t n) {
    if (n <= L->cap) return;
    int32_t c = L->cap ? L->cap : 8;
    while (c < n) c *= 2;
    L->cell = xrealloc(L->cell, sizeof(int32_t) * 3 * (size_t)c);
    L->parent = xrealloc(L->parent, sizeof(int32_t) * (size_t)c);
    L->child_first = xrealloc(L->child_first, sizeof(int32_t) * (size_t)c);
    L->lo = xrealloc(L->lo, sizeof(int32_t) * (size_t)c);
    L->hi = xrealloc(L->hi, sizeof(int32_t) * (size_t)c);
    L->depth = xrealloc(L->depth, sizeof(int32_t) * (size_t)c);
    L->leaf = xrealloc(L->leaf, (size_t)c);
    L->ghost = xrealloc(L->ghost, (size_t)c);
    L->xlist = xrealloc(L->xlist, (size_t)c);
    L->col = xrealloc(L->col, sizeof(int32_t) * NCOLL * (size_t)c);
    L->cap = c;
}

static void level_free(Level *L) {
    free(L->cell); free(L->parent); free(L->child_first); free(L->lo); free(L->hi);
    free(L->depth); free(L->leaf); free(L->ghost); free(L->xlist); free(L->col);
}

/* src/tree/tree.jl:81-89, src/util.jl:105 */
static void center_of(const double origin[3], const int32_t *cell, double w, double c[3]) {
    for (int d = 0; d < 3; d++) c[d] = origin[d] + ((double)cell[d] + 0.5) * w;
}
/* src/util.jl:39-45 (0-based slot) */
static int child_slot(const double *x, const double c[3]) {
    int k = 0;
    for (int d = 0; d < 3; d++) k += (x[d] > c[d]) << d;
    return k;
}

/* src/tree/builder.jl:45-69, 84-131 (the serial stable counting sort) and 163-190 */
static void split_nodes(Level *par, const int32_t *which, int32_t nw, int lvl, Level *chi,
                        double *pts, int32_t *perm, double *spts, int32_t *sperm,
                        const double origin[3], double W0, int ghost) {
    if (nw == 0) return;
    double w = ldexp(W0, -lvl);
    int32_t n0 = chi->n;
    level_reserve(chi, n0 + NC * nw);
    for (int32_t q = 0; q < nw; q++) {
        int32_t i = which[q];
        int32_t lo = par->lo[i], hi = par->hi[i];
        double c[3];
        center_of(origin, &par->cell[3 * i], w, c);
        int32_t counts[NC] = {0}, starts[NC], pos[NC];
        for (int32_t j = lo; j <= hi; j++) counts[child_slot(&pts[3 * (int64_t)j], c)]++;
        int32_t s = lo;
        for (int k = 0; k < NC; k++) { starts[k] = s; pos[k] = s; s += counts[k]; }
        for (int32_t j = lo; j <= hi; j++) {
            int k = child_slot(&pts[3 * (int64_t)j], c);
            int32_t p = pos[k]++;
            memcpy(&spts[3 * (int64_t)p], &pts[3 * (int64_t)j], 3 * sizeof(double));
            sperm[p] = perm[j];
        }
        for (int32_t j = lo; j <= hi; j++) {
            memcpy(&pts[3 * (int64_t)j], &spts[3 * (int64_t)j], 3 * sizeof(double));
            perm[j] = sperm[j];
        }
        par->child_first[i] = n0 + NC * q;
        if (!ghost) par->leaf[i] = 0;
        for (int k = 0; k < NC; k++) {
            int32_t g = n0 + NC * q + k;
            for (int d = 0; d < 3; d++) chi->cell[3 * g + d] = 2 * par->cell[3 * i + d] + ((k >> d) & 1);
            chi->parent[g] = i;
            chi->child_first[g] = -1;
            chi->lo[g] = starts[k];
            chi->hi[g] = starts[k] + counts[k] - 1;
            chi->leaf[g] = !ghost;
            chi->ghost[g] = (uint8_t)ghost;
            chi->xlist[g] = 0;
            chi->depth[g] = 0;
        }
    }
    chi->n = n0 + NC * nw;
}

/* src/tree/builder.jl:218-248: the parent rule */
static void tag_colleagues(Level *lev, const Level *par) {
    for (int32_t i = 0; i < lev->n; i++) {
        const int32_t *c = &lev->cell[3 * i];
        int32_t p = lev->parent[i];
        int32_t pc[3] = {c[0] >> 1, c[1] >> 1, c[2] >> 1};
        for (int k = 0; k < NCOLL; k++) {
            int32_t o[3] = {k % 3 - 1, (k / 3) % 3 - 1, k / 9 - 1};
            int32_t cc[3] = {c[0] + o[0], c[1] + o[1], c[2] + o[2]};
            /* arithmetic shift: floor for cc = -1 */
            int32_t dd[3] = {(cc[0] >> 1) - pc[0], (cc[1] >> 1) - pc[1], (cc[2] >> 1) - pc[2]};
            int off = (dd[0] + 1) + 3 * (dd[1] + 1) + 9 * (dd[2] + 1);
            int32_t q = par->col[NCOLL * p + off];
            int32_t r = -1;
            if (q >= 0 && par->child_first[q] >= 0) {
                int slot = (cc[0] & 1) | ((cc[1] & 1) << 1) | ((cc[2] & 1) << 2);
                r = par->child_first[q] + slot;
            }
            lev->col[NCOLL * i + k] = r;
        }
    }
}

static void tag_root(Level *root) {
    for (int k = 0; k < NCOLL; k++) root->col[k] = -1;
    root->col[SELF] = 0;
}

/* src/tree/builder.jl:260-280 */
static void compute_depths(Level *levels, int L) {
    for (int32_t i = 0; i < levels[L - 1].n; i++) levels[L - 1].depth[i] = 0;
    for (int l = L - 2; l >= 0; l--) {
        Level *lev = &levels[l], *nxt = &levels[l + 1];
        for (int32_t i = 0; i < lev->n; i++) {
            int32_t cf = lev->child_first[i];
            if (cf < 0) { lev->depth[i] = 0; continue; }
            int32_t m = 0;
            for (int k = 0; k < NC; k++) if (nxt->depth[cf + k] > m) m = nxt->depth[cf + k];
            lev->depth[i] = m + 1;
        }
    }
}

static void free_tree_lists(OTree *t);

void otree_free(OTree *t) {
    if (!t) return;
    free(t->level_start); free(t->cell); free(t->level); free(t->parent); free(t->child_first);
    free(t->plo); free(t->phi); free(t->col); free(t->leaf); free(t->ghost); free(t->xlist);
    free(t->has_source); free(t->perm); free(t->point_node); free(t->pts);
    free_tree_lists(t);
    free(t);
}

static void free_tree_lists(OTree *t) {
    if (!t->lptr) return;
    for (int l = 0; l < t->nlevels; l++) { free(t->lptr[l]); free(t->lsrc[l]); free(t->lfar[l]); }
    free(t->lptr); free(t->lsrc); free(t->lfar); free(t->lnp);
    t->lptr = NULL;
}

static int is_far(const int32_t *a, const int32_t *b) {
    int m = 0;
    for (int d = 0; d < 3; d++) { int v = abs(a[d] - b[d]); if (v > m) m = v; }
    return m >= 2;
}

/* src/tree/lists.jl:73-124: the far pairs of every non-ghost target of levels >= 3
 * (0-based level >= 2), in the enumeration order of the parent's colleagues and
 * their children; targets in node order. */
static void build_lists(OTree *t) {
    int L = t->nlevels;
    /* far_index_table (src/util.jl:87-97): table position -> far index, -1 if near */
    int far_index[343];
    for (int u = 0, nf = 0; u < 343; u++) {
        int x = u % 7 - 3, y = (u / 7) % 7 - 3, z = u / 49 - 3;
        int m = abs(x); if (abs(y) > m) m = abs(y); if (abs(z) > m) m = abs(z);
        far_index[u] = m >= 2 ? nf++ : -1;
    }
    t->lptr = calloc(L, sizeof(int32_t *));
    t->lsrc = calloc(L, sizeof(int32_t *));
    t->lfar = calloc(L, sizeof(int32_t *));
    t->lnp = calloc(L, sizeof(int64_t));
    for (int l = 0; l < L; l++) {
        int32_t lo = t->level_start[l], nl = t->level_start[l + 1] - lo;
        int32_t *ptr = calloc((size_t)nl + 1, sizeof(int32_t));
        for (int32_t q = 0; q < nl && l >= 2; q++) {
            int32_t i = lo + q, c = 0;
            if (!t->ghost[i]) {
                int32_t p = t->parent[i];
                for (int k = 0; k < NCOLL; k++) {
                    int32_t cc = t->col[NCOLL * p + k];
                    if (cc < 0) continue;
                    int32_t cf = t->child_first[cc];
                    if (cf < 0) continue;
                    for (int32_t s = cf; s < cf + NC; s++)
                        if (t->has_source[s] && is_far(&t->cell[3 * s], &t->cell[3 * i])) c++;
                }
            }
            ptr[q + 1] = c;
        }
        for (int32_t q = 0; q < nl; q++) ptr[q + 1] += ptr[q];
        int64_t np = ptr[nl];
        int32_t *src = malloc(sizeof(int32_t) * (size_t)(np ? np : 1));
        int32_t *far = malloc(sizeof(int32_t) * (size_t)(np ? np : 1));
        for (int32_t q = 0; q < nl && l >= 2; q++) {
            int32_t i = lo + q, pos = ptr[q];
            if (t->ghost[i]) continue;
            int32_t p = t->parent[i];
            for (int k = 0; k < NCOLL; k++) {
                int32_t cc = t->col[NCOLL * p + k];
                if (cc < 0) continue;
                int32_t cf = t->child_first[cc];
                if (cf < 0) continue;
                for (int32_t s = cf; s < cf + NC; s++) {
                    if (!t->has_source[s]) continue;
                    const int32_t *a = &t->cell[3 * s], *b = &t->cell[3 * i];
                    if (!is_far(a, b)) continue;
                    int32_t o[3] = {a[0] - b[0], a[1] - b[1], a[2] - b[2]};
                    int tab = (o[0] + 3) + 7 * (o[1] + 3) + 49 * (o[2] + 3);
                    src[pos] = s;
                    far[pos] = far_index[tab];
                    pos++;
                }
            }
        }
        t->lptr[l] = ptr; t->lsrc[l] = src; t->lfar[l] = far; t->lnp[l] = np;
    }
}

/* Tree(points; ppl, bbox, max_levels): src/tree/builder.jl:482-495 */
int otree_build(const double *points, int64_t N, int ppl, const double origin[3], double W0,
                int max_levels, OTree **out) {
    *out = NULL;
    if (ppl < 1) { snprintf(g_err, sizeof g_err, "ppl must be positive"); return 1; }
    if (N <= 0) { snprintf(g_err, sizeof g_err, "no points"); return 1; }
    if (!(W0 > 0)) { snprintf(g_err, sizeof g_err, "box width must be positive"); return 1; }
    for (int64_t j = 0; j < N; j++)
        for (int d = 0; d < 3; d++) {
            double x = points[3 * j + d];
            if (!(origin[d] <= x && x <= origin[d] + W0)) {
                snprintf(g_err, sizeof g_err, "point %lld lies outside the box", (long long)j);
                return 1;
            }
        }
    double *pts = malloc(sizeof(double) * 3 * (size_t)N);
    double *spts = malloc(sizeof(double) * 3 * (size_t)N);
    int32_t *perm = malloc(sizeof(int32_t) * (size_t)N);
    int32_t *sperm = malloc(sizeof(int32_t) * (size_t)N);
    memcpy(pts, points, sizeof(double) * 3 * (size_t)N);
    for (int64_t j = 0; j < N; j++) perm[j] = (int32_t)j;

    int Lcap = 64, L = 1;
    Level *levels = calloc(Lcap, sizeof(Level));
    level_reserve(&levels[0], 1);
    levels[0].n = 1;
    levels[0].cell[0] = levels[0].cell[1] = levels[0].cell[2] = 0;
    levels[0].parent[0] = -1; levels[0].child_first[0] = -1;
    levels[0].lo[0] = 0; levels[0].hi[0] = (int32_t)(N - 1);
    levels[0].leaf[0] = 1; levels[0].ghost[0] = 0; levels[0].xlist[0] = 0; levels[0].depth[0] = 0;
    int32_t *which = malloc(sizeof(int32_t) * 16);
    int32_t wcap = 16;
#define PUSH_WHICH(v) do { if (nw == wcap) { wcap *= 2; which = xrealloc(which, sizeof(int32_t) * wcap); } which[nw++] = (v); } while (0)
    /* refinement: src/tree/builder.jl:367-376 */
    for (;;) {
        Level *lev = &levels[L - 1];
        int32_t nw = 0;
        for (int32_t i = 0; i < lev->n; i++) if (lev->hi[i] - lev->lo[i] + 1 > ppl) PUSH_WHICH(i);
        if (nw == 0) break;
        if (L >= max_levels) {
            snprintf(g_err, sizeof g_err, "tree reached max_levels = %d; more than ppl = %d coincident points?", max_levels, ppl);
            for (int l = 0; l < L; l++) level_free(&levels[l]);
            free(levels); free(which); free(pts); free(spts); free(perm); free(sperm);
            return 1;
        }
        if (L == Lcap) { Lcap *= 2; levels = xrealloc(levels, sizeof(Level) * Lcap); memset(levels + L, 0, sizeof(Level) * (Lcap - L)); }
        split_nodes(&levels[L - 1], which, nw, L - 1, &levels[L], pts, perm, spts, sperm, origin, W0, 0);
        L++;
    }
    tag_root(&levels[0]);
    for (int l = 1; l < L; l++) tag_colleagues(&levels[l], &levels[l - 1]);
    /* level restriction: src/tree/builder.jl:289-316 (reference levels 2:L-2) */
    compute_depths(levels, L);
    for (;;) {
        int64_t nsplit = 0;
        for (int l = 1; l <= L - 3; l++) {
            Level *lev = &levels[l];
            int32_t nw = 0;
            for (int32_t i = 0; i < lev->n; i++) {
                if (!lev->leaf[i]) continue;
                for (int k = 0; k < NCOLL; k++) {
                    int32_t j = lev->col[NCOLL * i + k];
                    if (j >= 0 && abs(lev->depth[i] - lev->depth[j]) > 1) { PUSH_WHICH(i); break; }
                }
            }
            if (nw == 0) continue;
            split_nodes(lev, which, nw, l, &levels[l + 1], pts, perm, spts, sperm, origin, W0, 0);
            tag_colleagues(&levels[l + 1], lev);
            compute_depths(levels, L);
            nsplit += nw;
        }
        if (nsplit == 0) break;
    }
    /* ghosts: src/tree/builder.jl:328-351 (reference levels 2:L-1) */
    for (int l = 1; l <= L - 2; l++) {
        Level *lev = &levels[l];
        for (int32_t i = 0; i < lev->n; i++) {
            if (!lev->leaf[i]) continue;
            for (int k = 0; k < NCOLL; k++) {
                int32_t j = lev->col[NCOLL * i + k];
                if (j >= 0 && lev->child_first[j] >= 0) { lev->xlist[i] = 1; break; }
            }
        }
    }
    for (int l = 1; l <= L - 2; l++) {
        Level *lev = &levels[l];
        int32_t nw = 0;
        for (int32_t i = 0; i < lev->n; i++) if (lev->xlist[i]) PUSH_WHICH(i);
        if (nw == 0) continue;
        split_nodes(lev, which, nw, l, &levels[l + 1], pts, perm, spts, sperm, origin, W0, 1);
        tag_colleagues(&levels[l + 1], lev);
    }
    free(which);
    free(spts); free(sperm);
    /* flatten: src/tree/builder.jl:392-436, 447-473 */
    OTree *t = calloc(1, sizeof(OTree));
    t->nlevels = L;
    t->npoints = N;
    memcpy(t->origin, origin, sizeof(double) * 3);
    t->W0 = W0;
    t->level_start = malloc(sizeof(int32_t) * (L + 1));
    t->level_start[0] = 0;
    for (int l = 0; l < L; l++) t->level_start[l + 1] = t->level_start[l] + levels[l].n;
    int32_t n = t->level_start[L];
    t->nnodes = n;
    t->cell = malloc(sizeof(int32_t) * 3 * (size_t)n);
    t->level = malloc(sizeof(int32_t) * (size_t)n);
    t->parent = malloc(sizeof(int32_t) * (size_t)n);
    t->child_first = malloc(sizeof(int32_t) * (size_t)n);
    t->plo = malloc(sizeof(int32_t) * (size_t)n);
    t->phi = malloc(sizeof(int32_t) * (size_t)n);
    t->col = malloc(sizeof(int32_t) * NCOLL * (size_t)n);
    t->leaf = malloc((size_t)n); t->ghost = malloc((size_t)n); t->xlist = malloc((size_t)n);
    t->has_source = malloc((size_t)n);
    t->perm = perm;
    t->pts = pts;
    t->point_node = malloc(sizeof(int32_t) * (size_t)N);
    for (int64_t j = 0; j < N; j++) t->point_node[j] = -1;
    for (int l = 0; l < L; l++) {
        Level *lev = &levels[l];
        int32_t off = t->level_start[l];
        for (int32_t i = 0; i < lev->n; i++) {
            int32_t g = off + i;
            memcpy(&t->cell[3 * g], &lev->cell[3 * i], 3 * sizeof(int32_t));
            t->level[g] = l;
            t->parent[g] = l == 0 ? -1 : t->level_start[l - 1] + lev->parent[i];
            t->child_first[g] = lev->child_first[i] < 0 ? -1 : t->level_start[l + 1] + lev->child_first[i];
            t->plo[g] = lev->lo[i];
            t->phi[g] = lev->hi[i];
            t->leaf[g] = lev->leaf[i]; t->ghost[g] = lev->ghost[i]; t->xlist[g] = lev->xlist[i];
            t->has_source[g] = lev->hi[i] >= lev->lo[i];
            for (int k = 0; k < NCOLL; k++) {
                int32_t j = lev->col[NCOLL * i + k];
                t->col[NCOLL * g + k] = j < 0 ? -1 : off + j;
            }
            if (lev->leaf[i])
                for (int32_t j = lev->lo[i]; j <= lev->hi[i]; j++) t->point_node[j] = g;
        }
    }
    for (int l = 0; l < L; l++) level_free(&levels[l]);
    free(levels);
    for (int64_t j = 0; j < N; j++)
        if (t->point_node[j] < 0) { snprintf(g_err, sizeof g_err, "point not assigned to a leaf"); otree_free(t); return 2; }
    build_lists(t);
    *out = t;
    return 0;
#undef PUSH_WHICH
}

int otree_nlevels(const OTree *t) { return t->nlevels; }
int otree_nnodes(const OTree *t) { return t->nnodes; }
int64_t otree_level_npairs(const OTree *t, int l) { return t->lnp[l]; }

void otree_get(const OTree *t, int32_t *level_start, int32_t *cell, int32_t *level, int32_t *parent,
               int32_t *child_first, int32_t *plo, int32_t *phi, uint8_t *leaf, uint8_t *ghost,
               uint8_t *xlist, uint8_t *has_source, int32_t *col, int32_t *perm, int32_t *point_node,
               double *pts) {
    int32_t n = t->nnodes;
    int64_t N = t->npoints;
    memcpy(level_start, t->level_start, sizeof(int32_t) * (t->nlevels + 1));
    memcpy(cell, t->cell, sizeof(int32_t) * 3 * (size_t)n);
    memcpy(level, t->level, sizeof(int32_t) * (size_t)n);
    memcpy(parent, t->parent, sizeof(int32_t) * (size_t)n);
    memcpy(child_first, t->child_first, sizeof(int32_t) * (size_t)n);
    memcpy(plo, t->plo, sizeof(int32_t) * (size_t)n);
    memcpy(phi, t->phi, sizeof(int32_t) * (size_t)n);
    memcpy(leaf, t->leaf, (size_t)n); memcpy(ghost, t->ghost, (size_t)n);
    memcpy(xlist, t->xlist, (size_t)n); memcpy(has_source, t->has_source, (size_t)n);
    memcpy(col, t->col, sizeof(int32_t) * NCOLL * (size_t)n);
    memcpy(perm, t->perm, sizeof(int32_t) * (size_t)N);
    memcpy(point_node, t->point_node, sizeof(int32_t) * (size_t)N);
    memcpy(pts, t->pts, sizeof(double) * 3 * (size_t)N);
}

void otree_get_lists(const OTree *t, int l, int32_t *ptr, int32_t *src, int32_t *far) {
    int32_t nl = t->level_start[l + 1] - t->level_start[l];
    memcpy(ptr, t->lptr[l], sizeof(int32_t) * ((size_t)nl + 1));
    memcpy(src, t->lsrc[l], sizeof(int32_t) * (size_t)t->lnp[l]);
    memcpy(far, t->lfar[l], sizeof(int32_t) * (size_t)t->lnp[l]);
}

/* src/tree/locate.jl:10-22: walk from the root with the builder's predicate to the
 * first leaf; -1 for a point outside the root box. */
void otree_locate(const OTree *t, const double *x, int64_t n, int32_t *out) {
    #pragma omp parallel for schedule(static)
    for (int64_t j = 0; j < n; j++) {
        const double *p = &x[3 * j];
        int inside = 1;
        for (int d = 0; d < 3; d++) inside &= (t->origin[d] <= p[d]) & (p[d] <= t->origin[d] + t->W0);
        if (!inside) { out[j] = -1; continue; }
        int32_t i = 0;
        while (!t->leaf[i]) {
            double c[3];
            center_of(t->origin, &t->cell[3 * i], ldexp(t->W0, -t->level[i]), c);
            i = t->child_first[i] + child_slot(p, c);
        }
        out[j] = i;
    }
}
