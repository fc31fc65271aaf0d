// The 2D kernel-independent method on circles, KernelIndependent(Laplace2D(); nequiv)
// (src/operators/kifmm.jl:1-87), for the potential of charges.
//
// * The quadtree is built on the host, by the reference's host builder for D = 2
//   (src/tree/builder.jl: split_nodes! with a stable counting sort per node,
//   tag_colleagues! by the parent rule, level_restrict!, add_ghosts!, flatten) and its
//   interaction lists (_count_pairs! / _fill_pairs!, src/tree/lists.jl:71-124), then
//   uploaded once.
// * The passes run on the device in FP64, level by level: S2M (one wave per childless
//   sourced box, a lane per check point of the large circle, the box's sources staged in
//   LDS), the pseudo-inverse applies and M2M / L2L as library GEMMs on the level's
//   operators, M2L as one wave per target box (a lane per entry of its small circle,
//   the pair's source expansion in LDS, the far offset's dense matrix read coalesced),
//   and the evaluation (L2T on the leaf's large circle plus the near field of the 9
//   colleagues, the source itself excluded by index).
// The kernel c log r^2 is not homogeneous: every level carries its own operators
// (src/device_passes.jl:242-255).
#include "ffmm_internal.h"

#include <algorithm>
#include <cmath>
#include <cstring>

struct ffmm_circle_s;

namespace ffmm {
namespace circle {

constexpr int C2_NCH = 4;      // children per box
constexpr int C2_NCO = 9;      // colleagues
constexpr int C2_SELF = 4;     // the colleague slot of the box itself
constexpr int C2_NFAR = 40;    // far offsets: (-3:3)^2 outside (-1:1)^2
constexpr int C2_MAXN = 64;    // points per circle, at most one wave
constexpr int C2_TB = 64;
constexpr double KC = -0.07957747154594767;   // -1 / (4 pi)

// ---------------------------------------------------------------------------
// the host builder (src/tree/builder.jl, D = 2), 0-based, -1 for none
// ---------------------------------------------------------------------------
struct HLevel {
  std::vector<int64_t> cx, cy;
  std::vector<int32_t> parent, child_first, lo, hi, depth, col;   // col: 9 per node
  std::vector<uint8_t> leaf, ghost, xlist;
  int64_t n() const { return (int64_t)cx.size(); }
};

// the centre of cell c at level l, as origin + (c + 1/2) w with single roundings
static double centre(double o, int64_t c, double w) {
#pragma clang fp contract(off)
  const double h = (double)c + 0.5;
  const double m = h * w;
  return o + m;
}

struct Builder {
  std::vector<double> px, py;   // points, permuted with perm
  std::vector<int32_t> perm;
  std::vector<double> tx, ty;   // scratch of the counting sort
  std::vector<int32_t> tp;
  double ox = 0, oy = 0, W0 = 1;

  // split_nodes! (src/tree/builder.jl:40-190): the children of the nodes `which` of
  // level l, appended to `chi` in the order of `which`, each group in slot order; the
  // points of each node counting-sorted (stably) by child slot
  void split(HLevel &par, const std::vector<int32_t> &which, int l, HLevel &chi, bool ghost) {
    const double w = std::ldexp(W0, -l);
    for (int32_t i : which) {
      if (par.child_first[i] >= 0) fail(FFMM_ERR_STATE, "node already has children");
      const int32_t lo = par.lo[i], hi = par.hi[i];
      const double cxv = centre(ox, par.cx[i], w), cyv = centre(oy, par.cy[i], w);
      int32_t cnt[C2_NCH] = {0, 0, 0, 0}, start[C2_NCH], pos[C2_NCH];
      auto slot = [&](int32_t j) { return (int)(px[j] > cxv) | ((int)(py[j] > cyv) << 1); };
      for (int32_t j = lo; j <= hi; j++) cnt[slot(j)]++;
      int32_t s = lo;
      for (int k = 0; k < C2_NCH; k++) {
        start[k] = pos[k] = s;
        s += cnt[k];
      }
      for (int32_t j = lo; j <= hi; j++) {
        const int k = slot(j);
        tx[pos[k]] = px[j];
        ty[pos[k]] = py[j];
        tp[pos[k]] = perm[j];
        pos[k]++;
      }
      for (int32_t j = lo; j <= hi; j++) {
        px[j] = tx[j];
        py[j] = ty[j];
        perm[j] = tp[j];
      }
      par.child_first[i] = (int32_t)chi.n();
      if (!ghost) par.leaf[i] = 0;
      for (int k = 0; k < C2_NCH; k++) {
        chi.cx.push_back(2 * par.cx[i] + (k & 1));
        chi.cy.push_back(2 * par.cy[i] + ((k >> 1) & 1));
        chi.parent.push_back(i);
        chi.child_first.push_back(-1);
        chi.lo.push_back(start[k]);
        chi.hi.push_back(start[k] + cnt[k] - 1);
        chi.leaf.push_back(ghost ? 0 : 1);
        chi.ghost.push_back(ghost ? 1 : 0);
        chi.xlist.push_back(0);
        chi.depth.push_back(0);
      }
    }
  }
};

// tag_colleagues! (src/tree/builder.jl:212-248): colleague k of node i is the node at
// offset (k % 3 - 1, k / 3 - 1), found through the parent's colleague table
static void tag(HLevel &lev, const HLevel &par) {
  const int64_t n = lev.n();
  lev.col.assign((size_t)C2_NCO * n, -1);
  for (int64_t i = 0; i < n; i++) {
    const int64_t c0 = lev.cx[i], c1 = lev.cy[i];
    const int32_t p = lev.parent[i];
    for (int k = 0; k < C2_NCO; k++) {
      const int64_t a0 = c0 + (k % 3 - 1), a1 = c1 + (k / 3 - 1);
      const int off = (int)((a0 >> 1) - (c0 >> 1) + 1) + 3 * (int)((a1 >> 1) - (c1 >> 1) + 1);
      const int32_t q = par.col[(size_t)C2_NCO * p + off];
      if (q >= 0 && par.child_first[q] >= 0)
        lev.col[(size_t)C2_NCO * i + k] = par.child_first[q] + (int32_t)((a0 & 1) | ((a1 & 1) << 1));
    }
  }
}

static void depths(std::vector<HLevel> &lv) {
  const int L = (int)lv.size();
  std::fill(lv[L - 1].depth.begin(), lv[L - 1].depth.end(), 0);
  for (int l = L - 2; l >= 0; l--)
    for (int64_t i = 0; i < lv[l].n(); i++) {
      const int32_t cf = lv[l].child_first[i];
      int32_t d = 0;
      if (cf >= 0)
        for (int k = 0; k < C2_NCH; k++) d = std::max(d, lv[l + 1].depth[cf + k] + 1);
      lv[l].depth[i] = d;
    }
}

// the flat tree, levels concatenated (the reference's Tree fields, 0-based)
struct HostTree2 {
  int L = 0;
  int64_t nn = 0;
  std::vector<int64_t> level_start;
  std::vector<int32_t> cell, level, parent, child_first, plo, phi, col, perm;   // cell: 2 per node; col: 9 per node
  std::vector<uint8_t> leaf, ghost, xlist;
  std::vector<double> px, py;   // sorted points
  // per level: far pairs by target (ptr over the level's nodes, src global, far index)
  std::vector<std::vector<int32_t>> ptr, src, far;
};

// the far index of offset (o0, o1) (src/util.jl:86-97): position among the far offsets
// of (-3:3)^2 in table order (coordinate 0 fastest), -1 for a near offset
static int far_index(int o0, int o1) {
  if (std::max(std::abs(o0), std::abs(o1)) < 2) return -1;
  int k = 0;
  for (int t = 0; t < 49; t++) {
    const int a = t % 7 - 3, b = t / 7 - 3;
    if (std::max(std::abs(a), std::abs(b)) < 2) continue;
    if (a == o0 && b == o1) return k;
    k++;
  }
  return -1;
}

static HostTree2 build(const double *xy, int64_t N, int ppl, double ox, double oy, double W0, int max_levels) {
  Builder b;
  b.ox = ox;
  b.oy = oy;
  b.W0 = W0;
  b.px.resize(N);
  b.py.resize(N);
  b.perm.resize(N);
  b.tx.resize(N);
  b.ty.resize(N);
  b.tp.resize(N);
  for (int64_t j = 0; j < N; j++) {
    b.px[j] = xy[2 * j];
    b.py[j] = xy[2 * j + 1];
    b.perm[j] = (int32_t)j;
  }
  std::vector<HLevel> lv(1);
  HLevel &root = lv[0];
  root.cx = {0};
  root.cy = {0};
  root.parent = {-1};
  root.child_first = {-1};
  root.lo = {0};
  root.hi = {(int32_t)(N - 1)};
  root.depth = {0};
  root.leaf = {1};
  root.ghost = {0};
  root.xlist = {0};
  // refinement (build_levels, src/tree/builder.jl:399-425)
  for (;;) {
    HLevel &lev = lv.back();
    std::vector<int32_t> which;
    for (int64_t i = 0; i < lev.n(); i++)
      if (lev.hi[i] - lev.lo[i] + 1 > ppl) which.push_back((int32_t)i);
    if (which.empty()) break;
    if ((int)lv.size() >= max_levels)
      fail(FFMM_ERR_ARGUMENT, "tree reached max_levels = " + std::to_string(max_levels) + "; more than ppl = " +
                                  std::to_string(ppl) + " coincident points?");
    lv.emplace_back();
    b.split(lv[lv.size() - 2], which, (int)lv.size() - 2, lv.back(), false);
  }
  const int L = (int)lv.size();
  lv[0].col.assign(C2_NCO, -1);
  lv[0].col[C2_SELF] = 0;
  for (int l = 1; l < L; l++) tag(lv[l], lv[l - 1]);
  // level_restrict! (src/tree/builder.jl:287-316): levels 1 .. L-3 (0-based)
  depths(lv);
  for (;;) {
    int64_t nsplit = 0;
    for (int l = 1; l < L - 2; l++) {
      HLevel &lev = lv[l];
      std::vector<int32_t> which;
      for (int64_t i = 0; i < lev.n(); i++) {
        if (!lev.leaf[i]) continue;
        for (int k = 0; k < C2_NCO; k++) {
          const int32_t j = lev.col[(size_t)C2_NCO * i + k];
          if (j >= 0 && std::abs(lev.depth[i] - lev.depth[j]) > 1) {
            which.push_back((int32_t)i);
            break;
          }
        }
      }
      if (which.empty()) continue;
      b.split(lev, which, l, lv[l + 1], false);
      tag(lv[l + 1], lev);
      depths(lv);
      nsplit += (int64_t)which.size();
    }
    if (nsplit == 0) break;
  }
  // add_ghosts! (src/tree/builder.jl:318-351): mark on all levels first, then split
  for (int l = 1; l < L - 1; l++) {
    HLevel &lev = lv[l];
    for (int64_t i = 0; i < lev.n(); i++) {
      if (!lev.leaf[i]) continue;
      for (int k = 0; k < C2_NCO; k++) {
        const int32_t j = lev.col[(size_t)C2_NCO * i + k];
        if (j >= 0 && lev.child_first[j] >= 0) {
          lev.xlist[i] = 1;
          break;
        }
      }
    }
  }
  for (int l = 1; l < L - 1; l++) {
    HLevel &lev = lv[l];
    std::vector<int32_t> which;
    for (int64_t i = 0; i < lev.n(); i++)
      if (lev.xlist[i]) which.push_back((int32_t)i);
    if (which.empty()) continue;
    b.split(lev, which, l, lv[l + 1], true);
    tag(lv[l + 1], lev);
  }
  // flatten (src/tree/builder.jl:353-400)
  HostTree2 t;
  t.L = L;
  t.level_start.assign(L + 1, 0);
  for (int l = 0; l < L; l++) t.level_start[l + 1] = t.level_start[l] + lv[l].n();
  const int64_t nn = t.level_start[L];
  t.nn = nn;
  if (nn >= INT32_MAX / C2_NCO) fail(FFMM_ERR_ARGUMENT, "too many nodes for 32-bit indices");
  t.cell.resize(2 * nn);
  t.level.resize(nn);
  t.parent.resize(nn);
  t.child_first.resize(nn);
  t.plo.resize(nn);
  t.phi.resize(nn);
  t.col.resize((size_t)C2_NCO * nn);
  t.leaf.resize(nn);
  t.ghost.resize(nn);
  t.xlist.resize(nn);
  for (int l = 0; l < L; l++) {
    const HLevel &lev = lv[l];
    const int64_t off = t.level_start[l];
    for (int64_t i = 0; i < lev.n(); i++) {
      const int64_t g = off + i;
      t.cell[2 * g] = (int32_t)lev.cx[i];
      t.cell[2 * g + 1] = (int32_t)lev.cy[i];
      t.level[g] = l;
      t.parent[g] = l == 0 ? -1 : (int32_t)(t.level_start[l - 1] + lev.parent[i]);
      t.child_first[g] = lev.child_first[i] < 0 ? -1 : (int32_t)(t.level_start[l + 1] + lev.child_first[i]);
      t.plo[g] = lev.lo[i];
      t.phi[g] = lev.hi[i];
      t.leaf[g] = lev.leaf[i];
      t.ghost[g] = lev.ghost[i];
      t.xlist[g] = lev.xlist[i];
      for (int k = 0; k < C2_NCO; k++) {
        const int32_t j = lev.col[(size_t)C2_NCO * i + k];
        t.col[(size_t)C2_NCO * g + k] = j < 0 ? -1 : (int32_t)(off + j);
      }
    }
  }
  t.perm = b.perm;
  t.px = b.px;
  t.py = b.py;
  // the interaction lists (_count_pairs! / _fill_pairs!, src/tree/lists.jl:71-124)
  t.ptr.resize(L);
  t.src.resize(L);
  t.far.resize(L);
  for (int l = 0; l < L; l++) {
    const int64_t lo = t.level_start[l], n = t.level_start[l + 1] - lo;
    std::vector<int32_t> &ptr = t.ptr[l], &src = t.src[l], &far = t.far[l];
    ptr.assign(n + 1, 0);
    for (int64_t q = 0; q < n; q++) {
      const int64_t i = lo + q;
      if (l >= 2 && !t.ghost[i]) {
        const int32_t p = t.parent[i];
        for (int k = 0; k < C2_NCO; k++) {
          const int32_t cc = t.col[(size_t)C2_NCO * p + k];
          if (cc < 0 || t.child_first[cc] < 0) continue;
          for (int32_t s = t.child_first[cc]; s < t.child_first[cc] + C2_NCH; s++) {
            if (t.phi[s] < t.plo[s]) continue;
            const int f = far_index(t.cell[2 * s] - t.cell[2 * i], t.cell[2 * s + 1] - t.cell[2 * i + 1]);
            if (f < 0) continue;
            src.push_back(s);
            far.push_back(f);
          }
        }
      }
      ptr[q + 1] = (int32_t)src.size();
    }
  }
  return t;
}

// ---------------------------------------------------------------------------
// device kernels
// ---------------------------------------------------------------------------
__device__ __forceinline__ double lap2(double dx, double dy) { return KC * log(fma(dy, dy, dx * dx)); }

__global__ void k_c2_gather(double *__restrict__ dst, const double *__restrict__ src, const int32_t *__restrict__ perm,
                            int64_t n) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j < n) dst[j] = src[perm[j]];
}

// S2M (s2m_accumulate!, src/operators/ki_common.jl:256-272): pM[k, q] = sum over the
// sources j of node q of tau_j K(x_j - c, y_k), y_k on the level's large circle
__global__ void __launch_bounds__(C2_TB) k_c2_s2m(const int32_t *__restrict__ nodes, const double *__restrict__ cen,
                                               int64_t nn, const int32_t *__restrict__ plo,
                                               const int32_t *__restrict__ phi, const double *__restrict__ sx,
                                               const double *__restrict__ sy, const double *__restrict__ tau,
                                               const double *__restrict__ large, int n, double *__restrict__ pM) {
  __shared__ double xs[C2_TB], ys[C2_TB], ts[C2_TB];
  const int32_t i = nodes[blockIdx.x];
  const int k = threadIdx.x;
  const double c0 = cen[i], c1 = cen[nn + i];
  const double y0 = k < n ? large[k] : 0.0, y1 = k < n ? large[n + k] : 0.0;
  double acc = 0.0;
  for (int32_t j0 = plo[i]; j0 <= phi[i]; j0 += C2_TB) {
    const int32_t j = j0 + k;
    if (j <= phi[i]) {
      xs[k] = sx[j] - c0;
      ys[k] = sy[j] - c1;
      ts[k] = tau[j];
    }
    __syncthreads();
    const int cnt = min(C2_TB, phi[i] - j0 + 1);
    for (int q = 0; q < cnt; q++) acc += lap2(y0 - xs[q], y1 - ys[q]) * ts[q];
    __syncthreads();
  }
  if (k < n) pM[(int64_t)blockIdx.x * n + k] = acc;
}

// dst[:, nodes[q]] = src[:, q] (n rows)
__global__ void k_c2_scatter_cols(double *__restrict__ dst, const double *__restrict__ src,
                                  const int32_t *__restrict__ nodes, int n, int64_t nq) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (int64_t)n * nq) return;
  const int64_t q = e / n;
  dst[(int64_t)nodes[q] * n + e % n] = src[e];
}
// dst[:, q] = src[:, nodes[q]]
__global__ void k_c2_gather_cols(double *__restrict__ dst, const double *__restrict__ src,
                                 const int32_t *__restrict__ nodes, int n, int64_t nq) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (int64_t)n * nq) return;
  const int64_t q = e / n;
  dst[e] = src[(int64_t)nodes[q] * n + e % n];
}
// the parents of a level's sibling groups: G[:, g] = Pl[:, parent[lo + 4 g]], or, for
// M2M, Mp[:, parent[lo + 4 g]] = T[:, g]
__global__ void k_c2_parents(double *__restrict__ G, double *__restrict__ X, const int32_t *__restrict__ parent,
                             int64_t lo, int n, int64_t ng, int to_parents) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (int64_t)n * ng) return;
  const int64_t g = e / n;
  const int64_t p = parent[lo + C2_NCH * g];
  if (to_parents) X[p * n + e % n] = G[e];
  else G[e] = X[p * n + e % n];
}
__global__ void k_c2_fill(double *__restrict__ a, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = 0.0;
}

// M2L (m2l_level!, src/fmm.jl:100-127, by target): pL[:, t] += sum over the target's
// pairs of M2L[f] Mp[:, s]; one wave per target, lane k the entry k of the product
__global__ void __launch_bounds__(C2_TB) k_c2_m2l(double *__restrict__ pL, const double *__restrict__ Mp,
                                               const int32_t *__restrict__ ptr, const int32_t *__restrict__ src,
                                               const int32_t *__restrict__ far, const double *__restrict__ A,
                                               int n) {
  __shared__ double m[C2_TB];
  const int64_t t = blockIdx.x;
  const int32_t p0 = ptr[t], p1 = ptr[t + 1];
  if (p0 == p1) return;
  const int k = threadIdx.x;
  double acc = 0.0;
  for (int32_t e = p0; e < p1; e++) {
    const int64_t s = src[e];
    __syncthreads();
    if (k < n) m[k] = Mp[s * n + k];
    __syncthreads();
    if (k < n) {
      const double *__restrict__ a = A + (int64_t)far[e] * n * n + k;
      for (int i = 0; i < n; i++) acc = fma(a[(int64_t)i * n], m[i], acc);
    }
  }
  if (k < n) pL[t * n + k] += acc;
}

// evaluate_target (src/evaluate.jl:37-58), Potential: the L2T of the target's leaf
// (its density on the large circle, l2t, src/operators/ki_common.jl:293-303) plus the
// near field of the leaf's colleagues, the source `skip` excluded (a masked select)
__device__ __forceinline__ double c2_eval(double x, double y, int32_t leaf, int32_t skip, const double *__restrict__ cen,
                                          int64_t nn, const int32_t *__restrict__ level, const double *__restrict__ Lc,
                                          const double *const *__restrict__ large, int n,
                                          const int32_t *__restrict__ col, const int32_t *__restrict__ plo,
                                          const int32_t *__restrict__ phi, const double *__restrict__ sx,
                                          const double *__restrict__ sy, const double *__restrict__ tau) {
  double acc = 0.0;
  const int lv = level[leaf];
  if (lv >= 2) {
    const double *__restrict__ Y = large[lv];
    const double *__restrict__ L = Lc + (int64_t)leaf * n;
    const double xr = x - cen[leaf], yr = y - cen[nn + leaf];
    for (int k = 0; k < n; k++) acc += lap2(xr - Y[k], yr - Y[n + k]) * L[k];
  }
  for (int c = 0; c < C2_NCO; c++) {
    const int32_t j = col[(int64_t)C2_NCO * leaf + c];
    if (j < 0) continue;
    for (int32_t s = plo[j]; s <= phi[j]; s++) {
      const double v = lap2(x - sx[s], y - sy[s]) * tau[s];
      acc += s != skip ? v : 0.0;
    }
  }
  return acc;
}

struct EvalData {
  const double *cen;
  int64_t nn;
  const int32_t *level, *col, *plo, *phi;
  const double *Lc, *sx, *sy, *tau;
  const double *const *large;
  int n;
};

__global__ void k_c2_eval_sources(EvalData d, const int32_t *__restrict__ point_node, const int32_t *__restrict__ perm,
                                  int64_t N, double *__restrict__ out) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= N) return;
  out[perm[j]] = c2_eval(d.sx[j], d.sy[j], point_node[j], (int32_t)j, d.cen, d.nn, d.level, d.Lc, d.large, d.n, d.col,
                         d.plo, d.phi, d.sx, d.sy, d.tau);
}

// locate (src/tree/locate.jl:10-22): descend by child slot (strictly greater goes up)
__global__ void k_c2_eval_targets(EvalData d, const int32_t *__restrict__ child_first,
                                  const uint8_t *__restrict__ leaf, const double *__restrict__ tgt, int64_t nt,
                                  double *__restrict__ out) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nt) return;
  const double x = tgt[2 * q], y = tgt[2 * q + 1];
  int32_t i = 0;
  while (!leaf[i]) i = child_first[i] + ((int)(x > d.cen[i]) | ((int)(y > d.cen[d.nn + i]) << 1));
  out[q] = c2_eval(x, y, i, -1, d.cen, d.nn, d.level, d.Lc, d.large, d.n, d.col, d.plo, d.phi, d.sx, d.sy, d.tau);
}

// ---------------------------------------------------------------------------
// the plan
// ---------------------------------------------------------------------------
struct LevelOps {
  bool put = false;
  int r_up = 0, r_dn = 0;
  DBuf<double> large, Uh_up, Vs_up, Uh_dn, Vs_dn, M2M, L2L, M2L;
};

struct Plan2 {
  Context *ctx = nullptr;
  int n = 0;                // points per circle
  int64_t N = 0;
  double ox = 0, oy = 0, W0 = 1;
  HostTree2 h;
  std::vector<std::vector<int32_t>> s2m_h, leaves_h;   // per level: childless sourced nodes, real leaves
  DBuf<double> sx, sy, cen, tau, Mp, Pl, Lc, work, work2;
  DBuf<int32_t> perm, level, parent, child_first, plo, phi, col, point_node;
  DBuf<uint8_t> leaf;
  std::vector<DBuf<int32_t>> s2m, leaves, ptr, src, far;
  std::vector<LevelOps> ops;
  DBuf<const double *> large_ptrs;
  bool ready = false;
};

static void gemm(Context &ctx, int m, int nc, int k, const double *A, int lda, const double *B, int ldb, double *C,
                 int ldc) {
  if (m == 0 || nc == 0) return;
  const double one = 1.0, zero = 0.0;
  FFMM_BLAS(rocblas_dgemm(ctx.blas, rocblas_operation_none, rocblas_operation_none, m, nc, k, &one, A, lda, B, ldb,
                          &zero, C, ldc));
}

static void check_ready(Plan2 &p) {
  for (int l = 2; l < p.h.L; l++)
    if (!p.ops[l].put) fail(FFMM_ERR_STATE, "the operators of level " + std::to_string(l) + " are missing");
}

static void build_expansions(Plan2 &p, const double *tau) {
  check_ready(p);
  Context &ctx = *p.ctx;
  hipStream_t s = ctx.stream;
  const int n = p.n, L = p.h.L;
  const int64_t nn = p.h.nn;
  k_c2_gather<<<nblocks(p.N, 256), 256, 0, s>>>(p.tau.p, tau, p.perm.p, p.N);
  for (double *a : {p.Mp.p, p.Pl.p, p.Lc.p}) k_c2_fill<<<nblocks((int64_t)n * nn, 256), 256, 0, s>>>(a, (int64_t)n * nn);
  FFMM_LAUNCH_CHECK();
  // upward_pass! (src/fmm.jl:170-205)
  for (int l = L - 1; l >= 2; l--) {
    const LevelOps &o = p.ops[l];
    const int64_t nq = (int64_t)p.s2m_h[l].size();
    if (nq > 0) {
      double *pM = p.work.p, *t = p.work2.p;
      k_c2_s2m<<<(unsigned)nq, C2_TB, 0, s>>>(p.s2m[l].p, p.cen.p, nn, p.plo.p, p.phi.p, p.sx.p, p.sy.p, p.tau.p,
                                           o.large.p, n, pM);
      FFMM_LAUNCH_CHECK();
      // partial_to_multipole!: V (invS .* (Uh pM)), the scale of a non-homogeneous kernel 1
      gemm(ctx, o.r_up, (int)nq, n, o.Uh_up.p, o.r_up, pM, n, t, o.r_up);
      gemm(ctx, n, (int)nq, o.r_up, o.Vs_up.p, n, t, o.r_up, pM, n);
      k_c2_scatter_cols<<<nblocks((int64_t)n * nq, 256), 256, 0, s>>>(p.Mp.p, pM, p.s2m[l].p, n, nq);
      FFMM_LAUNCH_CHECK();
    }
    if (l < L - 1) {   // M2M: the children of level l + 1, in group order, to their parents
      const int64_t lo = p.h.level_start[l + 1], ng = (p.h.level_start[l + 2] - lo) / C2_NCH;
      if (ng > 0) {
        gemm(ctx, n, (int)ng, C2_NCH * n, o.M2M.p, n, p.Mp.p + lo * n, C2_NCH * n, p.work.p, n);
        k_c2_parents<<<nblocks((int64_t)n * ng, 256), 256, 0, s>>>(p.work.p, p.Mp.p, p.parent.p, lo, n, ng, 1);
        FFMM_LAUNCH_CHECK();
      }
    }
  }
  // downward_pass! (src/fmm.jl:207-262)
  for (int l = 2; l < L; l++) {
    const LevelOps &o = p.ops[l];
    const int64_t lo = p.h.level_start[l], nl = p.h.level_start[l + 1] - lo;
    double *pL = p.Pl.p + lo * n;
    if (l > 2) {   // L2L of the parents' partial locals into the level's block
      const int64_t ng = nl / C2_NCH;
      k_c2_parents<<<nblocks((int64_t)n * ng, 256), 256, 0, s>>>(p.work.p, p.Pl.p, p.parent.p, lo, n, ng, 0);
      FFMM_LAUNCH_CHECK();
      gemm(ctx, C2_NCH * n, (int)ng, n, p.ops[l - 1].L2L.p, C2_NCH * n, p.work.p, n, pL, C2_NCH * n);
    }
    if (!p.h.src[l].empty()) {
      k_c2_m2l<<<(unsigned)nl, C2_TB, 0, s>>>(pL, p.Mp.p, p.ptr[l].p, p.src[l].p, p.far[l].p, o.M2L.p, n);
      FFMM_LAUNCH_CHECK();
    }
    const int64_t nq = (int64_t)p.leaves_h[l].size();
    if (nq > 0) {   // partial_to_local! of the real leaves
      k_c2_gather_cols<<<nblocks((int64_t)n * nq, 256), 256, 0, s>>>(p.work.p, p.Pl.p, p.leaves[l].p, n, nq);
      FFMM_LAUNCH_CHECK();
      gemm(ctx, o.r_dn, (int)nq, n, o.Uh_dn.p, o.r_dn, p.work.p, n, p.work2.p, o.r_dn);
      gemm(ctx, n, (int)nq, o.r_dn, o.Vs_dn.p, n, p.work2.p, o.r_dn, p.work.p, n);
      k_c2_scatter_cols<<<nblocks((int64_t)n * nq, 256), 256, 0, s>>>(p.Lc.p, p.work.p, p.leaves[l].p, n, nq);
      FFMM_LAUNCH_CHECK();
    }
  }
  ctx.sync();
  p.ready = true;
}

static EvalData eval_data(Plan2 &p) {
  return EvalData{p.cen.p, p.h.nn, p.level.p, p.col.p, p.plo.p, p.phi.p, p.Lc.p, p.sx.p, p.sy.p, p.tau.p,
                  p.large_ptrs.p, p.n};
}

}  // namespace circle
}  // namespace ffmm

struct ffmm_circle_s {
  ffmm::circle::Plan2 c;
};

using namespace ffmm;
using namespace ffmm::circle;

extern "C" {

int ffmm_circle_create(ffmm_context ctx, int nequiv, const double *xy, int64_t n, int ppl, const double origin[2],
                       double width, int max_levels, ffmm_circle *out) {
  return guard([&] {
    if (!ctx || !xy || !origin || !out) fail(FFMM_ERR_ARGUMENT, "null argument");
    *out = nullptr;
    if (nequiv < 4 || nequiv > C2_MAXN) fail(FFMM_ERR_ARGUMENT, "nequiv must lie in 4:64");
    if (n <= 0) fail(FFMM_ERR_ARGUMENT, "no points");
    if (n >= INT32_MAX) fail(FFMM_ERR_ARGUMENT, "too many points for 32-bit indices");
    if (ppl < 1) fail(FFMM_ERR_ARGUMENT, "ppl must be positive");
    if (!(width > 0)) fail(FFMM_ERR_ARGUMENT, "box width must be positive");
    if (max_levels < 1 || max_levels > 30) fail(FFMM_ERR_ARGUMENT, "max_levels must lie in 1:30");
    for (int64_t j = 0; j < n; j++) {
      const double x = xy[2 * j], y = xy[2 * j + 1];
      if (!std::isfinite(x) || !std::isfinite(y)) fail(FFMM_ERR_ARGUMENT, "the points must be finite");
      if (!(origin[0] <= x && x <= origin[0] + width && origin[1] <= y && y <= origin[1] + width))
        fail(FFMM_ERR_ARGUMENT, "points lie outside the requested box");
    }
    auto h = std::make_unique<ffmm_circle_s>();
    Plan2 &p = h->c;
    p.ctx = &ctx->c;
    p.n = nequiv;
    p.N = n;
    p.ox = origin[0];
    p.oy = origin[1];
    p.W0 = width;
    p.h = build(xy, n, ppl, origin[0], origin[1], width, max_levels);
    HostTree2 &t = p.h;
    const int L = t.L;
    const int64_t nn = t.nn;
    // node lists and centres
    p.s2m_h.assign(L, {});
    p.leaves_h.assign(L, {});
    std::vector<int32_t> point_node(n, -1);
    std::vector<double> cen(2 * nn);
    for (int64_t g = 0; g < nn; g++) {
      const int l = t.level[g];
      const double w = std::ldexp(width, -l);
      cen[g] = centre(origin[0], t.cell[2 * g], w);
      cen[nn + g] = centre(origin[1], t.cell[2 * g + 1], w);
      const bool has = t.phi[g] >= t.plo[g];
      if (t.child_first[g] < 0 && has) p.s2m_h[l].push_back((int32_t)g);
      if (t.leaf[g]) {
        p.leaves_h[l].push_back((int32_t)g);
        for (int32_t j = t.plo[g]; j <= t.phi[g]; j++) point_node[j] = (int32_t)g;
      }
    }
    Context &c = *p.ctx;
    FFMM_HIP(hipSetDevice(c.device));
    hipStream_t s = c.stream;
    p.sx.upload(t.px.data(), n, s);
    p.sy.upload(t.py.data(), n, s);
    p.perm.upload(t.perm.data(), n, s);
    p.cen.upload(cen.data(), 2 * nn, s);
    p.level.upload(t.level.data(), nn, s);
    p.parent.upload(t.parent.data(), nn, s);
    p.child_first.upload(t.child_first.data(), nn, s);
    p.plo.upload(t.plo.data(), nn, s);
    p.phi.upload(t.phi.data(), nn, s);
    p.col.upload(t.col.data(), (size_t)C2_NCO * nn, s);
    p.leaf.upload(t.leaf.data(), nn, s);
    p.point_node.upload(point_node.data(), n, s);
    p.s2m.resize(L);
    p.leaves.resize(L);
    p.ptr.resize(L);
    p.src.resize(L);
    p.far.resize(L);
    int64_t wmax = 1;
    for (int l = 0; l < L; l++) {
      p.s2m[l].upload(p.s2m_h[l].data(), p.s2m_h[l].size(), s);
      p.leaves[l].upload(p.leaves_h[l].data(), p.leaves_h[l].size(), s);
      p.ptr[l].upload(t.ptr[l].data(), t.ptr[l].size(), s);
      p.src[l].upload(t.src[l].data(), t.src[l].size(), s);
      p.far[l].upload(t.far[l].data(), t.far[l].size(), s);
      wmax = std::max<int64_t>({wmax, (int64_t)p.s2m_h[l].size(), (int64_t)p.leaves_h[l].size(),
                                (t.level_start[l + 1] - t.level_start[l]) / C2_NCH});
    }
    p.tau.resize(n);
    p.Mp.resize((size_t)nequiv * nn);
    p.Pl.resize((size_t)nequiv * nn);
    p.Lc.resize((size_t)nequiv * nn);
    p.work.resize((size_t)nequiv * wmax);
    p.work2.resize((size_t)nequiv * wmax);
    p.ops.resize(L);
    c.sync();
    *out = h.release();
  });
}

int ffmm_circle_sizes(ffmm_circle c, int *nlevels, int64_t *nnodes, int64_t *level_start, int64_t *level_pairs) {
  return guard([&] {
    if (!c || !nlevels || !nnodes) fail(FFMM_ERR_ARGUMENT, "null argument");
    const HostTree2 &t = c->c.h;
    *nlevels = t.L;
    *nnodes = t.nn;
    if (level_start)
      for (int l = 0; l <= t.L; l++) level_start[l] = t.level_start[l];
    if (level_pairs)
      for (int l = 0; l < t.L; l++) level_pairs[l] = (int64_t)t.src[l].size();
  });
}

int ffmm_circle_put_level(ffmm_circle c, int level, const double *large, int r_up, const double *Uh_up,
                          const double *Vs_up, int r_dn, const double *Uh_dn, const double *Vs_dn, const double *M2M,
                          const double *L2L, const double *M2L) {
  return guard([&] {
    if (!c || !large || !Uh_up || !Vs_up || !Uh_dn || !Vs_dn || !M2M || !L2L || !M2L)
      fail(FFMM_ERR_ARGUMENT, "null argument");
    Plan2 &p = c->c;
    if (level < 2 || level >= p.h.L) fail(FFMM_ERR_ARGUMENT, "level out of range");
    const int n = p.n;
    if (r_up < 1 || r_up > n || r_dn < 1 || r_dn > n) fail(FFMM_ERR_ARGUMENT, "pseudo-inverse rank out of range");
    Context &ctx = *p.ctx;
    FFMM_HIP(hipSetDevice(ctx.device));
    hipStream_t s = ctx.stream;
    LevelOps &o = p.ops[level];
    o.r_up = r_up;
    o.r_dn = r_dn;
    o.large.upload(large, 2 * (size_t)n, s);
    o.Uh_up.upload(Uh_up, (size_t)r_up * n, s);
    o.Vs_up.upload(Vs_up, (size_t)n * r_up, s);
    o.Uh_dn.upload(Uh_dn, (size_t)r_dn * n, s);
    o.Vs_dn.upload(Vs_dn, (size_t)n * r_dn, s);
    o.M2M.upload(M2M, (size_t)n * C2_NCH * n, s);
    o.L2L.upload(L2L, (size_t)C2_NCH * n * n, s);
    o.M2L.upload(M2L, (size_t)C2_NFAR * n * n, s);
    o.put = true;
    // the large circles of all levels, one pointer per level (null below level 2)
    std::vector<const double *> lp(p.h.L, nullptr);
    for (int l = 2; l < p.h.L; l++) lp[l] = p.ops[l].put ? p.ops[l].large.p : nullptr;
    p.large_ptrs.upload(lp.data(), lp.size(), s);
    ctx.sync();
  });
}

int ffmm_circle_build_expansions(ffmm_circle c, const double *tau) {
  return guard([&] {
    if (!c || !tau) fail(FFMM_ERR_ARGUMENT, "null argument");
    FFMM_HIP(hipSetDevice(c->c.ctx->device));
    build_expansions(c->c, tau);
  });
}

int ffmm_circle_evaluate_at_sources(ffmm_circle c, double *out) {
  return guard([&] {
    if (!c || !out) fail(FFMM_ERR_ARGUMENT, "null argument");
    Plan2 &p = c->c;
    if (!p.ready) fail(FFMM_ERR_STATE, "build the expansions first");
    FFMM_HIP(hipSetDevice(p.ctx->device));
    k_c2_eval_sources<<<nblocks(p.N, 128), 128, 0, p.ctx->stream>>>(eval_data(p), p.point_node.p, p.perm.p, p.N,
                                                                    out);
    FFMM_LAUNCH_CHECK();
    p.ctx->sync();
  });
}

int ffmm_circle_evaluate(ffmm_circle c, const double *targets, int64_t nt, double *out) {
  return guard([&] {
    if (!c) fail(FFMM_ERR_ARGUMENT, "null plan");
    Plan2 &p = c->c;
    if (!p.ready) fail(FFMM_ERR_STATE, "build the expansions first");
    if (nt < 0) fail(FFMM_ERR_DIMENSION, "negative target count");
    if (nt == 0) return;
    if (!targets || !out) fail(FFMM_ERR_ARGUMENT, "null argument");
    // the targets must lie in the root box: checked on the host copy
    std::vector<double> h(2 * nt);
    FFMM_HIP(hipSetDevice(p.ctx->device));
    FFMM_HIP(hipMemcpyAsync(h.data(), targets, sizeof(double) * 2 * nt, hipMemcpyDeviceToHost, p.ctx->stream));
    p.ctx->sync();
    for (int64_t q = 0; q < nt; q++)
      if (!(p.ox <= h[2 * q] && h[2 * q] <= p.ox + p.W0 && p.oy <= h[2 * q + 1] && h[2 * q + 1] <= p.oy + p.W0))
        fail(FFMM_ERR_ARGUMENT, "a target lies outside the tree's box");
    k_c2_eval_targets<<<nblocks(nt, 128), 128, 0, p.ctx->stream>>>(eval_data(p), p.child_first.p, p.leaf.p, targets,
                                                                   nt, out);
    FFMM_LAUNCH_CHECK();
    p.ctx->sync();
  });
}

int ffmm_circle_download_tree(ffmm_circle c, int32_t *cell, int32_t *parent, int32_t *child_first, int32_t *plo,
                              int32_t *phi, uint8_t *leaf, uint8_t *ghost, uint8_t *xlist, int32_t *colleagues,
                              int32_t *perm) {
  return guard([&] {
    if (!c) fail(FFMM_ERR_ARGUMENT, "null plan");
    const HostTree2 &t = c->c.h;
    const size_t nn = (size_t)t.nn;
    auto put = [](auto *dst, const auto &v) {
      if (dst) std::memcpy(dst, v.data(), v.size() * sizeof(v[0]));
    };
    put(cell, t.cell);
    put(parent, t.parent);
    put(child_first, t.child_first);
    put(plo, t.plo);
    put(phi, t.phi);
    put(leaf, t.leaf);
    put(ghost, t.ghost);
    put(xlist, t.xlist);
    if (colleagues) std::memcpy(colleagues, t.col.data(), (size_t)C2_NCO * nn * sizeof(int32_t));
    put(perm, t.perm);
  });
}

int ffmm_circle_download_lists(ffmm_circle c, int level, int32_t *ptr, int32_t *src, int32_t *far) {
  return guard([&] {
    if (!c) fail(FFMM_ERR_ARGUMENT, "null plan");
    const HostTree2 &t = c->c.h;
    if (level < 0 || level >= t.L) fail(FFMM_ERR_ARGUMENT, "level out of range");
    if (ptr) std::memcpy(ptr, t.ptr[level].data(), t.ptr[level].size() * sizeof(int32_t));
    if (src) std::memcpy(src, t.src[level].data(), t.src[level].size() * sizeof(int32_t));
    if (far) std::memcpy(far, t.far[level].data(), t.far[level].size() * sizeof(int32_t));
  });
}

int ffmm_circle_destroy(ffmm_circle c) {
  return guard([&] {
    if (!c) return;
    (void)hipSetDevice(c->c.ctx->device);
    c->c.ctx->sync();
    delete c;
  });
}

}  // extern "C"
