// The ghost-cell octree built on the device from the points there.
//
// Morton keys by the builder's own child-slot rule (exact comparisons against the box
// centres), one stable radix sort, then the host builder's phases as kernels over the
// nodes of one level, with the level loop on the host and one small read back per
// level and phase: refinement with every child's point range found as the run of its
// key prefix (two binary searches, points never move again), colleagues by the parent
// rule, the 2:1 restriction, the ghosts, and the flattening.  The node order of every
// level is the host builder's (refinement children in parent order, restriction
// children appended per round, ghosts last), so every structural array equals the
// host tree's; only the order of the points inside a leaf differs (full key here).
//
// Follows src/tree/device_builder.jl:124-474 (and the host rules it restates,
// src/tree/builder.jl:45-384); 0-based indices, -1 = none.
#include "ffmm_internal.h"

#include <algorithm>
#include <cmath>

namespace ffmm {

constexpr int TB = 256;

// ---------------------------------------------------------------------------
// bounds: _device_bounds (src/tree/device_builder.jl:124-148)
// ---------------------------------------------------------------------------
__global__ void k_bounds(const double *__restrict__ xyz, int64_t n, double *__restrict__ part) {
  __shared__ double s[6][TB];
  __shared__ int sbad[TB];
  double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  int bad = 0;
  for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < n; j += (int64_t)gridDim.x * blockDim.x) {
    for (int d = 0; d < 3; d++) {
      double v = xyz[3 * j + d];
      bad |= !isfinite(v);
      lo[d] = fmin(lo[d], v);
      hi[d] = fmax(hi[d], v);
    }
  }
  int t = threadIdx.x;
  for (int d = 0; d < 3; d++) { s[d][t] = lo[d]; s[3 + d][t] = hi[d]; }
  sbad[t] = bad;
  __syncthreads();
  for (int w = TB / 2; w > 0; w >>= 1) {
    if (t < w) {
      for (int d = 0; d < 3; d++) {
        s[d][t] = fmin(s[d][t], s[d][t + w]);
        s[3 + d][t] = fmax(s[3 + d][t], s[3 + d][t + w]);
      }
      sbad[t] |= sbad[t + w];
    }
    __syncthreads();
  }
  if (t == 0) {
    for (int d = 0; d < 6; d++) part[7 * blockIdx.x + d] = s[d][0];
    part[7 * blockIdx.x + 6] = (double)sbad[0];
  }
}

void bounds(Context &ctx, const double *xyz, int64_t n, double lo[3], double hi[3], int *all_finite) {
  if (n <= 0) fail(FFMM_ERR_ARGUMENT, "no points");
  int nb = (int)std::min<int64_t>(1024, (n + TB - 1) / TB);
  DBuf<double> part;
  part.resize(7 * (size_t)nb);
  k_bounds<<<nb, TB, 0, ctx.stream>>>(xyz, n, part.p);
  FFMM_LAUNCH_CHECK();
  std::vector<double> h(7 * (size_t)nb);
  part.download(h.data(), h.size(), ctx.stream);
  ctx.sync();
  int bad = 0;
  for (int d = 0; d < 3; d++) { lo[d] = INFINITY; hi[d] = -INFINITY; }
  for (int b = 0; b < nb; b++) {
    for (int d = 0; d < 3; d++) {
      lo[d] = std::min(lo[d], h[7 * b + d]);
      hi[d] = std::max(hi[d], h[7 * b + 3 + d]);
    }
    bad |= h[7 * b + 6] != 0.0;
  }
  *all_finite = !bad;
}

// ---------------------------------------------------------------------------
// keys: _point_keys! (src/tree/device_builder.jl:181-199), child_slot (src/util.jl:39-45)
// The slot of level b (0-based box level) sits in bits 3(KL-1-b) .. 3(KL-1-b)+2.
// Box centres origin + (cell + 0.5) w are formed with single roundings (no fused
// multiply-add), bitwise as the reference forms them.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double box_center(double o, int64_t c, double w) {
  return __dadd_rn(o, __dmul_rn(__dadd_rn((double)c, 0.5), w));
}

__global__ void k_point_keys(const double *__restrict__ xyz, int64_t n, double ox, double oy, double oz, double W0,
                             int KL, uint64_t *__restrict__ keys, int32_t *__restrict__ vals) {
  int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  double x = xyz[3 * j], y = xyz[3 * j + 1], z = xyz[3 * j + 2];
  int64_t cx = 0, cy = 0, cz = 0;
  double w = W0;
  uint64_t key = 0;
  for (int b = 0; b < KL; b++) {
    int k = (x > box_center(ox, cx, w)) | ((y > box_center(oy, cy, w)) << 1) | ((z > box_center(oz, cz, w)) << 2);
    key |= (uint64_t)k << (3 * (KL - 1 - b));
    cx = 2 * cx + (k & 1);
    cy = 2 * cy + ((k >> 1) & 1);
    cz = 2 * cz + ((k >> 2) & 1);
    w = __dmul_rn(w, 0.5);
  }
  keys[j] = key;
  vals[j] = (int32_t)j;
}

__global__ void k_gather_points(const double *__restrict__ xyz, const int32_t *__restrict__ perm, int64_t n,
                                double *__restrict__ sx, double *__restrict__ sy, double *__restrict__ sz) {
  int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  int64_t i = perm[j];
  sx[j] = xyz[3 * i];
  sy[j] = xyz[3 * i + 1];
  sz[j] = xyz[3 * i + 2];
}

// ---------------------------------------------------------------------------
// the per-level builder
// ---------------------------------------------------------------------------
// the root: cell 0, no parent or children, points [0, N), a leaf
__global__ void k_init_root(int32_t *cx, int32_t *cy, int32_t *cz, int32_t *parent, int32_t *child_first, int32_t *lo,
                            int32_t *hi, int32_t *depth, uint8_t *leaf, uint8_t *ghost, uint8_t *xlist, int32_t last) {
  if (threadIdx.x != 0) return;
  cx[0] = cy[0] = cz[0] = 0;
  parent[0] = child_first[0] = -1;
  lo[0] = 0;
  hi[0] = last;
  depth[0] = 0;
  leaf[0] = 1;
  ghost[0] = 0;
  xlist[0] = 0;
}

__global__ void k_fill_i32(int32_t *__restrict__ a, int64_t n, int32_t v) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = v;
}
__global__ void k_leaf_slots(const int32_t *__restrict__ leaf_nodes, int64_t nleaves, int32_t *__restrict__ slot) {
  int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s < nleaves) slot[leaf_nodes[s]] = (int32_t)s;
}

struct DevLevel {
  int64_t n = 0;
  DBuf<int32_t> cx, cy, cz, parent, child_first, lo, hi, depth, col;   // col: [k * n + i]
  DBuf<uint8_t> leaf, ghost, xlist;
  void grow(int64_t nnew, hipStream_t s) {
    int64_t m = n + nnew;
    cx.grow(m, s); cy.grow(m, s); cz.grow(m, s); parent.grow(m, s); child_first.grow(m, s);
    lo.grow(m, s); hi.grow(m, s); depth.grow(m, s); leaf.grow(m, s); ghost.grow(m, s); xlist.grow(m, s);
    n = m;
  }
};

// _flag_split! (src/tree/device_builder.jl:204-207)
__global__ void k_flag_split(const int32_t *__restrict__ lo, const int32_t *__restrict__ hi, int64_t n, int ppl,
                             uint8_t *__restrict__ flag) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) flag[i] = (hi[i] - lo[i] + 1) > ppl;
}

// first j in [lo, hi] with keys[j] >= x, hi + 1 if none (src/tree/device_builder.jl:223-233)
__device__ __forceinline__ int32_t lower_bound(const uint64_t *__restrict__ keys, int32_t lo, int32_t hi, uint64_t x) {
  while (lo <= hi) {
    int32_t mid = (int32_t)(((int64_t)lo + hi) >> 1);
    if (keys[mid] < x) lo = mid + 1;
    else hi = mid - 1;
  }
  return lo;
}

// morton_key(cell, nbits) (src/util.jl:132-138)
__device__ __forceinline__ uint64_t morton3(int64_t cx, int64_t cy, int64_t cz, int nbits) {
  uint64_t key = 0;
  for (int b = 0; b < nbits; b++) {
    key |= (uint64_t)((cx >> b) & 1) << (3 * b);
    key |= (uint64_t)((cy >> b) & 1) << (3 * b + 1);
    key |= (uint64_t)((cz >> b) & 1) << (3 * b + 2);
  }
  return key;
}

// _children! (src/tree/device_builder.jl:237-267): one work-item per child of the split
// nodes `which` of a parent level at 0-based level `plvl`; children appended from n0
struct LevelPtrs {
  int32_t *cx, *cy, *cz, *parent, *child_first, *lo, *hi, *depth;
  uint8_t *leaf, *ghost, *xlist;
};
static LevelPtrs ptrs(DevLevel &L) {
  return {L.cx.p, L.cy.p, L.cz.p, L.parent.p, L.child_first.p, L.lo.p, L.hi.p, L.depth.p, L.leaf.p, L.ghost.p, L.xlist.p};
}

__global__ void k_children(LevelPtrs par, LevelPtrs chi, const int32_t *__restrict__ which, int64_t nw,
                           const uint64_t *__restrict__ keys, int64_t n0, int plvl, int KL, int ghost) {
  int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= NCHILD * nw) return;
  int64_t q = g / NCHILD;
  int k = (int)(g % NCHILD);
  int32_t i = which[q];
  int32_t cx = 2 * par.cx[i] + (k & 1), cy = 2 * par.cy[i] + ((k >> 1) & 1), cz = 2 * par.cz[i] + ((k >> 2) & 1);
  int nbits = plvl + 1;                 // the child's prefix: plvl + 1 slots
  uint64_t pk = morton3(cx, cy, cz, nbits);
  int shift = 3 * (KL - nbits);
  int32_t lo = par.lo[i], hi = par.hi[i];
  int32_t first = lower_bound(keys, lo, hi, pk << shift);
  int32_t last = lower_bound(keys, first, hi, (pk + 1) << shift) - 1;
  int64_t c = n0 + g;
  chi.cx[c] = cx; chi.cy[c] = cy; chi.cz[c] = cz;
  chi.parent[c] = i;
  chi.child_first[c] = -1;
  chi.lo[c] = first;
  chi.hi[c] = last;
  chi.leaf[c] = ghost ? 0 : 1;
  chi.ghost[c] = (uint8_t)ghost;
  chi.xlist[c] = 0;
  chi.depth[c] = 0;
  if (k == 0) {
    par.child_first[i] = (int32_t)(n0 + NCHILD * q);
    if (!ghost) par.leaf[i] = 0;
  }
}

// _tag! (src/tree/device_builder.jl:279-291) = tag_colleagues! (src/tree/builder.jl:212-248)
__global__ void k_tag(const int32_t *__restrict__ cx, const int32_t *__restrict__ cy, const int32_t *__restrict__ cz,
                      const int32_t *__restrict__ parent, int64_t n, const int32_t *__restrict__ pcol, int64_t pn,
                      const int32_t *__restrict__ pchild, int32_t *__restrict__ col) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int32_t c0 = cx[i], c1 = cy[i], c2 = cz[i];
  int32_t p = parent[i];
  int32_t p0 = c0 >> 1, p1 = c1 >> 1, p2 = c2 >> 1;
  for (int k = 0; k < NCOLL; k++) {
    int32_t a0 = c0 + (k % 3 - 1), a1 = c1 + ((k / 3) % 3 - 1), a2 = c2 + (k / 9 - 1);
    int off = ((a0 >> 1) - p0 + 1) + 3 * ((a1 >> 1) - p1 + 1) + 9 * ((a2 >> 1) - p2 + 1);
    int32_t q = pcol[(int64_t)off * pn + p];
    int32_t r = -1;
    if (q >= 0) {
      int32_t cf = pchild[q];
      if (cf >= 0) r = cf + ((a0 & 1) | ((a1 & 1) << 1) | ((a2 & 1) << 2));
    }
    col[(int64_t)k * n + i] = r;
  }
}

__global__ void k_root_col(int32_t *col) {
  int k = threadIdx.x;
  if (k < NCOLL) col[k] = (k == SELF_SLOT) ? 0 : -1;
}

// _depth! (src/tree/device_builder.jl:304-318)
__global__ void k_depth(const int32_t *__restrict__ cf, int64_t n, const int32_t *__restrict__ ndepth,
                        int32_t *__restrict__ depth) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int32_t f = cf[i];
  if (f < 0) { depth[i] = 0; return; }
  int32_t m = 0;
  for (int k = 0; k < NCHILD; k++) m = max(m, ndepth[f + k]);
  depth[i] = m + 1;
}
__global__ void k_zero_i32(int32_t *a, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = 0;
}

// _flag_bad! (src/tree/device_builder.jl:328-344)
__global__ void k_flag_bad(const uint8_t *__restrict__ leaf, const int32_t *__restrict__ depth,
                           const int32_t *__restrict__ col, int64_t n, uint8_t *__restrict__ flag) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint8_t b = 0;
  if (leaf[i]) {
    int32_t di = depth[i];
    for (int k = 0; k < NCOLL; k++) {
      int32_t j = col[(int64_t)k * n + i];
      if (j >= 0 && abs(di - depth[j]) > 1) { b = 1; break; }
    }
  }
  flag[i] = b;
}

// _flag_xlist! (src/tree/device_builder.jl:346-361)
__global__ void k_flag_xlist(const uint8_t *__restrict__ leaf, const int32_t *__restrict__ cf,
                             const int32_t *__restrict__ col, int64_t n, uint8_t *__restrict__ xlist) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint8_t b = 0;
  if (leaf[i]) {
    for (int k = 0; k < NCOLL; k++) {
      int32_t j = col[(int64_t)k * n + i];
      if (j >= 0 && cf[j] >= 0) { b = 1; break; }
    }
  }
  xlist[i] = b;
}

// _flatten! (src/tree/device_builder.jl:364-390)
struct FlatPtrs {
  int32_t *cell, *level, *parent, *child_first, *plo, *phi, *col, *point_node;
  uint8_t *leaf, *ghost, *xlist, *has_source;
  int64_t nn;
};
__global__ void k_flatten(LevelPtrs lv, const int32_t *__restrict__ lcol, int64_t ln, FlatPtrs f, int l, int64_t off,
                          int64_t poff, int64_t noff) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ln) return;
  int64_t g = off + i;
  f.cell[g] = lv.cx[i];
  f.cell[f.nn + g] = lv.cy[i];
  f.cell[2 * f.nn + g] = lv.cz[i];
  f.level[g] = l;
  f.parent[g] = l == 0 ? -1 : (int32_t)(poff + lv.parent[i]);
  f.child_first[g] = lv.child_first[i] < 0 ? -1 : (int32_t)(noff + lv.child_first[i]);
  f.plo[g] = lv.lo[i];
  f.phi[g] = lv.hi[i];
  f.leaf[g] = lv.leaf[i];
  f.ghost[g] = lv.ghost[i];
  f.xlist[g] = lv.xlist[i];
  f.has_source[g] = lv.hi[i] >= lv.lo[i];
  for (int k = 0; k < NCOLL; k++) {
    int32_t j = lcol[(int64_t)k * ln + i];
    f.col[(int64_t)k * f.nn + g] = j < 0 ? -1 : (int32_t)(off + j);
  }
  if (lv.leaf[i])
    for (int32_t j = lv.lo[i]; j <= lv.hi[i]; j++) f.point_node[j] = (int32_t)g;
}

// _centers! (src/device.jl:66-69)
__global__ void k_centers(const int32_t *__restrict__ cell, const int32_t *__restrict__ level, int64_t nn, double ox,
                          double oy, double oz, double W0, double *__restrict__ centers) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nn) return;
  double w = ldexp(W0, -level[i]);
  centers[i] = box_center(ox, cell[i], w);
  centers[nn + i] = box_center(oy, cell[nn + i], w);
  centers[2 * nn + i] = box_center(oz, cell[2 * nn + i], w);
}

// node lists: childless sourced nodes; real leaves (src/device_passes.jl:127-130)
__global__ void k_flag_s2m(const int32_t *__restrict__ cf, const uint8_t *__restrict__ hs, int64_t lo, int64_t n,
                           uint8_t *__restrict__ flag) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) flag[i] = cf[lo + i] < 0 && hs[lo + i];
}
// the first node of each level (and the node count), for the boundaries below
struct LevelStarts {
  int64_t lo[MAX_KEY_LEVELS + 2];
};
static_assert(2 + 2 * (MAX_KEY_LEVELS + 2) <= NPINNED - 1, "pinned slots of the level boundaries, below the polled slot");
// per level l (and l = L): the position in each node list (ascending node ids; their
// lengths in pinned slots 0 and 1) of the first node >= lo[l]; into pinned slots 2 + l
// and 2 + (MAX_KEY_LEVELS + 2) + l
__global__ void k_level_bounds(const int32_t *__restrict__ s2m, const int32_t *__restrict__ leaves, LevelStarts ls,
                               int L, int64_t *__restrict__ pinned) {
  const int l = threadIdx.x;
  if (l > L) return;
  const int64_t key = ls.lo[l];
  for (int w = 0; w < 2; w++) {
    const int32_t *a = w ? leaves : s2m;
    int64_t lo = 0, hi = pinned[w];
    while (lo < hi) {
      const int64_t m = (lo + hi) >> 1;
      if ((int64_t)a[m] < key) lo = m + 1;
      else hi = m;
    }
    pinned[2 + w * (MAX_KEY_LEVELS + 2) + l] = lo;
  }
}

static void split(Context &ctx, DevLevel &par, const int32_t *which, int64_t nw, int plvl, DevLevel &chi,
                  const uint64_t *keys, int KL, bool ghost) {
  if (nw == 0) return;
  int64_t n0 = chi.n;
  chi.grow(NCHILD * nw, ctx.stream);
  k_children<<<nblocks(NCHILD * nw, TB), TB, 0, ctx.stream>>>(ptrs(par), ptrs(chi), which, nw, keys, n0, plvl, KL,
                                                              ghost ? 1 : 0);
  FFMM_LAUNCH_CHECK();
}

static void tag(Context &ctx, DevLevel &lev, DevLevel &par) {
  lev.col.resize((size_t)NCOLL * lev.n);
  if (lev.n == 0) return;
  k_tag<<<nblocks(lev.n, TB), TB, 0, ctx.stream>>>(lev.cx.p, lev.cy.p, lev.cz.p, lev.parent.p, lev.n, par.col.p,
                                                   par.n, par.child_first.p, lev.col.p);
  FFMM_LAUNCH_CHECK();
}

// the depths of levels `from` .. 0 from those below (the deepest level's are zero); a
// split at level l changes the depths of levels l + 1 .. 0 only
static void depths(Context &ctx, std::vector<DevLevel> &levels, int from) {
  int L = (int)levels.size();
  if (from >= L - 1) {
    k_zero_i32<<<nblocks(levels[L - 1].n, TB), TB, 0, ctx.stream>>>(levels[L - 1].depth.p, levels[L - 1].n);
    from = L - 2;
  }
  for (int l = from; l >= 0; l--)
    k_depth<<<nblocks(levels[l].n, TB), TB, 0, ctx.stream>>>(levels[l].child_first.p, levels[l].n,
                                                             levels[l + 1].depth.p, levels[l].depth.p);
  FFMM_LAUNCH_CHECK();
}

void build_tree(Plan &plan, const double *xyz) {
  Context &ctx = *plan.ctx;
  hipStream_t s = ctx.stream;
  const int64_t N = plan.N;
  if (N <= 0) fail(FFMM_ERR_ARGUMENT, "no points");
  if (plan.ppl < 1) fail(FFMM_ERR_ARGUMENT, "ppl must be positive");
  if (!(plan.W0 > 0)) fail(FFMM_ERR_ARGUMENT, "box width must be positive");
  if (N >= INT32_MAX) fail(FFMM_ERR_ARGUMENT, "too many points for 32-bit point indices");
  const int KL = std::min(plan.max_levels - 1, MAX_KEY_LEVELS);
  if (KL < 1) fail(FFMM_ERR_ARGUMENT, "max_levels must be at least 2");
  const int maxl = KL + 1;
  // keys, the permutation, the sorted keys and points
  DBuf<uint64_t> keys, skeys;
  DBuf<int32_t> vals;
  keys.resize(N); skeys.resize(N); vals.resize(N);
  plan.perm.resize(N);
  k_point_keys<<<nblocks(N, TB), TB, 0, s>>>(xyz, N, plan.origin[0], plan.origin[1], plan.origin[2], plan.W0, KL,
                                             keys.p, vals.p);
  FFMM_LAUNCH_CHECK();
  sort_pairs_u64(ctx, keys.p, skeys.p, vals.p, plan.perm.p, N, 3 * KL);
  plan.sx.resize(N); plan.sy.resize(N); plan.sz.resize(N);
  k_gather_points<<<nblocks(N, TB), TB, 0, s>>>(xyz, plan.perm.p, N, plan.sx.p, plan.sy.p, plan.sz.p);
  FFMM_LAUNCH_CHECK();
  keys.release();
  vals.release();

  // the root (src/tree/device_builder.jl:419-425)
  std::vector<DevLevel> levels;
  levels.reserve(64);
  levels.emplace_back();
  {
    DevLevel &root = levels[0];
    root.grow(1, s);
    k_init_root<<<1, 64, 0, s>>>(root.cx.p, root.cy.p, root.cz.p, root.parent.p, root.child_first.p, root.lo.p,
                                 root.hi.p, root.depth.p, root.leaf.p, root.ghost.p, root.xlist.p, (int32_t)(N - 1));
    FFMM_LAUNCH_CHECK();
  }
  DBuf<uint8_t> flag;
  DBuf<int32_t> which;
  // refinement (src/tree/device_builder.jl:427-438)
  for (;;) {
    int l = (int)levels.size() - 1;
    int64_t n = levels[l].n;
    flag.resize(n);
    which.resize(n);
    k_flag_split<<<nblocks(n, TB), TB, 0, s>>>(levels[l].lo.p, levels[l].hi.p, n, plan.ppl, flag.p);
    FFMM_LAUNCH_CHECK();
    int64_t nsplit = compact_flags_u8(ctx, flag.p, which.p, n);
    if (nsplit == 0) break;
    if ((int)levels.size() >= maxl)
      fail(FFMM_ERR_ARGUMENT, "tree reached max_levels = " + std::to_string(maxl) + "; more than ppl = " +
                                  std::to_string(plan.ppl) + " coincident points?");
    levels.emplace_back();
    split(ctx, levels[l], which.p, nsplit, l, levels[l + 1], skeys.p, KL, false);
  }
  const int L = (int)levels.size();
  // colleagues (src/tree/device_builder.jl:440-443)
  levels[0].col.resize(NCOLL);
  k_root_col<<<1, 32, 0, s>>>(levels[0].col.p);
  for (int l = 1; l < L; l++) tag(ctx, levels[l], levels[l - 1]);
  // 2:1 restriction (src/tree/device_builder.jl:444-460): 0-based levels 1 .. L-3.  The
  // depths are recomputed once per pass: a split at level l makes new leaves (depth 0) at
  // level l + 1 and changes the depths of levels l .. 0 only, which this pass does not
  // check again (levels are taken upwards from 1, and a node's flag compares the depths
  // of its own level), so each check sees the depths a recomputation after every split
  // would give it
  depths(ctx, levels, L - 1);
  for (bool pass0 = true;; pass0 = false) {
    if (!pass0) depths(ctx, levels, L - 1);
    int64_t nsplit = 0;
    for (int l = 1; l <= L - 3; l++) {
      int64_t n = levels[l].n;
      flag.resize(n);
      which.resize(n);
      k_flag_bad<<<nblocks(n, TB), TB, 0, s>>>(levels[l].leaf.p, levels[l].depth.p, levels[l].col.p, n, flag.p);
      FFMM_LAUNCH_CHECK();
      int64_t nb = compact_flags_u8(ctx, flag.p, which.p, n);
      if (nb == 0) continue;
      split(ctx, levels[l], which.p, nb, l, levels[l + 1], skeys.p, KL, false);
      tag(ctx, levels[l + 1], levels[l]);
      nsplit += nb;
    }
    if (nsplit == 0) break;
  }
  // ghosts: the xlist flags of all levels first, then the splits (src/tree/device_builder.jl:461-472)
  for (int l = 1; l <= L - 2; l++) {
    int64_t n = levels[l].n;
    k_flag_xlist<<<nblocks(n, TB), TB, 0, s>>>(levels[l].leaf.p, levels[l].child_first.p, levels[l].col.p, n,
                                                levels[l].xlist.p);
    FFMM_LAUNCH_CHECK();
  }
  // every level's xlist compacted first, the counts read back together (the flags are
  // fixed before any ghost split, and ghosts carry no xlist flag)
  std::vector<DBuf<int32_t>> xw(L);
  if (L > 2) arm_slots(ctx, 1, L - 2);
  for (int l = 1; l <= L - 2; l++) {
    xw[l].resize(levels[l].n);
    compact_flags_u8_async(ctx, levels[l].xlist.p, xw[l].p, levels[l].n, l);
  }
  std::vector<int64_t> nxl(L, 0);
  for (int l = 1; l <= L - 2; l++) nxl[l] = read_slot(ctx, l);
  for (int l = 1; l <= L - 2; l++) {
    if (nxl[l] == 0) continue;
    split(ctx, levels[l], xw[l].p, nxl[l], l, levels[l + 1], skeys.p, KL, true);
    tag(ctx, levels[l + 1], levels[l]);
  }
  // flattening (src/tree/device_builder.jl:476-501)
  plan.L = L;
  plan.lo.assign(L, 0);
  plan.nl.assign(L, 0);
  int64_t nn = 0;
  for (int l = 0; l < L; l++) {
    plan.lo[l] = nn;
    plan.nl[l] = levels[l].n;
    nn += levels[l].n;
  }
  plan.nnodes = nn;
  if (nn >= INT32_MAX / NCOLL) fail(FFMM_ERR_ARGUMENT, "too many nodes for 32-bit node indices");
  plan.cell.resize(3 * nn);
  plan.level.resize(nn); plan.parent.resize(nn); plan.child_first.resize(nn);
  plan.plo.resize(nn); plan.phi.resize(nn); plan.col.resize(NCOLL * nn);
  plan.leaf.resize(nn); plan.ghost.resize(nn); plan.xlist.resize(nn); plan.has_source.resize(nn);
  plan.point_node.resize(N);
  FlatPtrs fp{plan.cell.p, plan.level.p, plan.parent.p, plan.child_first.p, plan.plo.p, plan.phi.p, plan.col.p,
              plan.point_node.p, plan.leaf.p, plan.ghost.p, plan.xlist.p, plan.has_source.p, nn};
  for (int l = 0; l < L; l++) {
    int64_t off = plan.lo[l], poff = l == 0 ? 0 : plan.lo[l - 1], noff = l == L - 1 ? 0 : plan.lo[l + 1];
    k_flatten<<<nblocks(levels[l].n, TB), TB, 0, s>>>(ptrs(levels[l]), levels[l].col.p, levels[l].n, fp, l, off,
                                                      poff, noff);
    FFMM_LAUNCH_CHECK();
  }
  compute_centers(plan);
  // node lists: both compacted over all nodes at once (node order is level order), the
  // level boundaries found on the device; one read back
  plan.s2m_off.assign(L + 1, 0);
  plan.leaf_off.assign(L + 1, 0);
  plan.s2m_nodes.resize(nn);
  plan.leaf_nodes.resize(nn);
  flag.resize(nn);
  k_flag_s2m<<<nblocks(nn, TB), TB, 0, s>>>(plan.child_first.p, plan.has_source.p, 0, nn, flag.p);
  FFMM_LAUNCH_CHECK();
  compact_flags_u8_async(ctx, flag.p, plan.s2m_nodes.p, nn, 0);
  compact_flags_u8_async(ctx, plan.leaf.p, plan.leaf_nodes.p, nn, 1);
  LevelStarts ls{};
  for (int l = 0; l <= L; l++) ls.lo[l] = l < L ? plan.lo[l] : nn;
  arm_slots(ctx, 2, L + 1);
  arm_slots(ctx, 2 + (MAX_KEY_LEVELS + 2), L + 1);
  k_level_bounds<<<1, 64, 0, s>>>(plan.s2m_nodes.p, plan.leaf_nodes.p, ls, L, ctx.pinned_dev);
  FFMM_LAUNCH_CHECK();
  for (int l = 0; l <= L; l++) {
    plan.s2m_off[l] = read_slot(ctx, 2 + l);
    plan.leaf_off[l] = read_slot(ctx, 2 + (MAX_KEY_LEVELS + 2) + l);
  }
}

void compute_centers(Plan &plan) {
  const int64_t nn = plan.nnodes;
  plan.centers.resize(3 * nn);
  k_centers<<<nblocks(nn, TB), TB, 0, plan.ctx->stream>>>(plan.cell.p, plan.level.p, nn, plan.origin[0],
                                                          plan.origin[1], plan.origin[2], plan.W0, plan.centers.p);
  FFMM_LAUNCH_CHECK();
}

// node -> position in the leaf list (-1: not a real leaf)
void node_leafslots(Plan &plan) {
  hipStream_t s = plan.ctx->stream;
  const int64_t nleaves = plan.leaf_off[plan.L];
  plan.node_leafslot.resize(plan.nnodes);
  k_fill_i32<<<nblocks(plan.nnodes, TB), TB, 0, s>>>(plan.node_leafslot.p, plan.nnodes, -1);
  k_leaf_slots<<<nblocks(nleaves, TB), TB, 0, s>>>(plan.leaf_nodes.p, nleaves, plan.node_leafslot.p);
  FFMM_LAUNCH_CHECK();
}

}  // namespace ffmm
