// The interaction lists of every level, built on the device from the flat tree.
//
// Per level (0-based levels >= 2): 32 lanes per target (one per colleague of its parent)
// count its far pairs -- the sourced children of the colleagues at Chebyshev offset >= 2,
// in the enumeration order of the colleagues and their children -- a scan places them
// and a fill writes (source, far offset).  Then one wave per target counts its pairs per
// class, and they are placed class-major (the pairs of a class with its targets in node
// order, each target's consecutive, in pair order) and filled; the class-major scan and
// the target chunks of the dense class M2L come with them.
//
// Follows src/device_lists.jl:57-157, 219-278 (= src/tree/lists.jl:73-124 for the
// pairs and src/operators/symmetry.jl:207-236 for the class lists); 0-based.
#include "ffmm_internal.h"

namespace ffmm {

constexpr int TB = 256;

__constant__ int32_t c_far_index[343];   // far_index_table (src/util.jl:87-97), -1: near

struct TreePtrs {
  const int32_t *cell, *parent, *child_first, *col;
  const uint8_t *ghost, *has_source;
  int64_t nn;
};

__device__ __forceinline__ bool far_offset(int32_t o0, int32_t o1, int32_t o2) {
  return max(abs(o0), max(abs(o1), abs(o2))) >= 2;
}

// The pairs of target q: 32 lanes, lane k < 27 for colleague k of the parent; a lane
// counts the sourced far children of its colleague, and the pairs are placed in
// colleague order, then child order (the reference's enumeration order).
constexpr int PL = 32;   // lanes per target of the pair kernels

__device__ __forceinline__ int32_t pair_lane_count(const TreePtrs &t, int64_t i, bool valid, int k, int32_t &s0,
                                                  int32_t &ci0, int32_t &ci1, int32_t &ci2) {
  s0 = -1;
  ci0 = ci1 = ci2 = 0;
  if (!valid || k >= NCOLL) return 0;
  const int32_t pc = t.col[(int64_t)k * t.nn + t.parent[i]];
  if (pc < 0) return 0;
  s0 = t.child_first[pc];
  if (s0 < 0) return 0;
  ci0 = t.cell[i];
  ci1 = t.cell[t.nn + i];
  ci2 = t.cell[2 * t.nn + i];
  int32_t c = 0;
  for (int32_t s = s0; s < s0 + NCHILD; s++)
    if (t.has_source[s] && far_offset(t.cell[s] - ci0, t.cell[t.nn + s] - ci1, t.cell[2 * t.nn + s] - ci2)) c++;
  return c;
}

// _pair_count! (src/device_lists.jl:57-78)
__global__ void k_pair_count(TreePtrs t, int64_t lo, int64_t nl, int32_t *__restrict__ cnt) {
  const int64_t q = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / PL;
  const int k = threadIdx.x % PL;
  const int64_t i = lo + q;
  const bool valid = q < nl && !t.ghost[i];
  int32_t s0, ci0, ci1, ci2;
  int32_t n = pair_lane_count(t, i, valid, k, s0, ci0, ci1, ci2);
#pragma unroll
  for (int off = 1; off < PL; off <<= 1) n += __shfl_xor(n, off, PL);
  if (k == 0 && q < nl) cnt[q] = n;
}

// _pair_fill! (src/device_lists.jl:80-106)
__global__ void k_pair_fill(TreePtrs t, int64_t lo, int64_t nl, const int32_t *__restrict__ ptr,
                            int32_t *__restrict__ src, int32_t *__restrict__ far) {
  const int64_t q = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / PL;
  const int k = threadIdx.x % PL;
  const int64_t i = lo + q;
  const bool valid = q < nl && !t.ghost[i];
  int32_t s0, ci0, ci1, ci2;
  const int32_t n = pair_lane_count(t, i, valid, k, s0, ci0, ci1, ci2);
  int32_t v = n;   // inclusive scan over the target's lanes
#pragma unroll
  for (int off = 1; off < PL; off <<= 1) {
    const int32_t u = __shfl_up(v, off, PL);
    if (k >= off) v += u;
  }
  if (n == 0) return;
  int32_t pos = ptr[q] + v - n;
  for (int32_t s = s0; s < s0 + NCHILD; s++) {
    if (!t.has_source[s]) continue;
    const int32_t o0 = t.cell[s] - ci0, o1 = t.cell[t.nn + s] - ci1, o2 = t.cell[2 * t.nn + s] - ci2;
    if (!far_offset(o0, o1, o2)) continue;
    src[pos] = s;
    far[pos] = c_far_index[(o0 + 3) + 7 * (o1 + 3) + 49 * (o2 + 3)];
    pos++;
  }
}

// The class lists of target q: one wave, the target's pairs 64 at a time (lane j: pair
// base + j, class c_j); lane c keeps class c's count or next position.  A pair's rank
// among the batch's earlier pairs of its class comes from one sweep over the lanes, so
// each class's pairs keep their order.
__device__ __forceinline__ void class_batch(int32_t c, int lane, int32_t &rank, int32_t &mine) {
  rank = 0;
  mine = 0;   // the batch's pairs of class `lane`
  for (int j = 0; j < 64; j++) {
    const int32_t cj = __shfl(c, j, 64);
    rank += (cj == c && j < lane) ? 1 : 0;
    mine += cj == lane ? 1 : 0;
  }
}

// _class_count! (src/device_lists.jl:108-117), class-major: cnt[c * nl + q]
__global__ void k_class_count(const int32_t *__restrict__ far, const int32_t *__restrict__ ptr, int64_t nl, int nclass,
                              const int32_t *__restrict__ far_class, int32_t *__restrict__ cnt) {
  const int64_t q = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / 64;
  const int lane = threadIdx.x % 64;
  if (q >= nl) return;   // whole waves: q is wave-uniform
  const int32_t p0 = ptr[q], p1 = ptr[q + 1];
  int32_t total = 0;
  for (int32_t base = p0; base < p1; base += 64) {
    const int32_t p = base + lane;
    const int32_t c = p < p1 ? far_class[far[p]] : -1;
    int32_t rank, mine;
    class_batch(c, lane, rank, mine);
    total += mine;
  }
  if (lane < nclass) cnt[(int64_t)lane * nl + q] = total;
}

// _class_fill! (src/device_lists.jl:119-129)
__global__ void k_class_fill(const int32_t *__restrict__ src, const int32_t *__restrict__ far,
                             const int32_t *__restrict__ ptr, const int32_t *__restrict__ cstart, int64_t nl,
                             int nclass, const int32_t *__restrict__ far_class, int32_t *__restrict__ csrc,
                             int32_t *__restrict__ cfar, int32_t *__restrict__ ctgt) {
  const int64_t q = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / 64;
  const int lane = threadIdx.x % 64;
  if (q >= nl) return;   // whole waves: q is wave-uniform
  const int32_t p0 = ptr[q], p1 = ptr[q + 1];
  int32_t next = lane < nclass ? cstart[(int64_t)lane * nl + q] : 0;   // class `lane`'s next position
  for (int32_t base = p0; base < p1; base += 64) {
    const int32_t p = base + lane;
    const int32_t f = p < p1 ? far[p] : 0;
    const int32_t c = p < p1 ? far_class[f] : -1;
    int32_t rank, mine;
    class_batch(c, lane, rank, mine);
    const int32_t at = __shfl(next, c < 0 ? 0 : c, 64);
    if (p < p1) {
      csrc[at + rank] = src[p];
      cfar[at + rank] = f;
      if (ctgt) ctgt[at + rank] = (int32_t)q;
    }
    next += mine;
  }
}

// the target chunks of a level's dense class M2L: chunk k starts at the first target t
// with ptr[t] >= k cap (t = nl for k = nch); per chunk start, the class-major position of
// (class c, target t) for every class.  out: (nch + 1) rows of 1 + nclass
__global__ void k_chunk_bounds(const int32_t *__restrict__ ptr, int64_t nl, int64_t cap, int64_t nch,
                               const int32_t *__restrict__ cstart, int nclass, int32_t *__restrict__ out) {
  int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k > nch) return;
  int64_t t = nl;
  if (k < nch) {
    const int64_t key = k * cap;
    int64_t a = 0, b = nl;
    while (a < b) {
      const int64_t m = (a + b) >> 1;
      if (ptr[m] < key) a = m + 1;
      else b = m;
    }
    t = a;
  }
  int32_t *o = out + k * (1 + nclass);
  o[0] = (int32_t)t;
  for (int c = 0; c < nclass; c++) o[1 + c] = cstart[(int64_t)c * nl + t];
}

__global__ void k_set_i32(int32_t *a, int32_t v) { a[0] = v; }

void fill_level_pairs(Plan &plan, int l) {
  LevelLists &li = plan.lists[l];
  if (li.filled) return;
  TreePtrs t{plan.cell.p, plan.parent.p, plan.child_first.p, plan.col.p, plan.ghost.p, plan.has_source.p, plan.nnodes};
  li.src.resize(li.np);
  li.far.resize(li.np);
  if (li.np > 0) {
    k_pair_fill<<<nblocks(PL * plan.nl[l], TB), TB, 0, plan.ctx->stream>>>(t, plan.lo[l], plan.nl[l], li.ptr.p, li.src.p,
                                                                      li.far.p);
    FFMM_LAUNCH_CHECK();
  }
  li.filled = true;
}

void build_lists(Plan &plan) {
  Context &ctx = *plan.ctx;
  OpCache &ops = *plan.ops;
  hipStream_t s = ctx.stream;
  once_on_device(ctx.device, 1, [&] {
    int32_t fi[343];
    for (int u = 0, nf = 0; u < 343; u++) {
      int x = u % 7 - 3, y = (u / 7) % 7 - 3, z = u / 49 - 3;
      int m = std::max(std::abs(x), std::max(std::abs(y), std::abs(z)));
      fi[u] = m >= 2 ? nf++ : -1;
    }
    FFMM_HIP(hipMemcpyToSymbolAsync(HIP_SYMBOL(c_far_index), fi, sizeof fi, 0, hipMemcpyHostToDevice, s));
    ctx.sync();
  });
  const int nclass = ops.nclass;
  if (nclass > MAX_CLASSES) fail(FFMM_ERR_ARGUMENT, "too many symmetry classes");
  TreePtrs t{plan.cell.p, plan.parent.p, plan.child_first.p, plan.col.p, plan.ghost.p, plan.has_source.p, plan.nnodes};
  plan.lists.clear();
  plan.lists.resize(plan.L);
  static_assert(MAX_KEY_LEVELS + 1 <= NPINNED, "a pinned slot per level");
  // the pair counts and their scans, all levels queued at once: one read back
  DBuf<int32_t> cnt, ccnt, bounds;
  cnt.resize(std::max<int64_t>(1, plan.nnodes));
  if (plan.L > 2) arm_slots(ctx, 2, plan.L - 2);
  for (int l = 0; l < plan.L; l++) {
    LevelLists &li = plan.lists[l];
    const int64_t lo = plan.lo[l], nl = plan.nl[l];
    li.class_pair.assign(nclass + 1, 0);
    li.chunk_t.clear();
    li.chunk_cs.clear();
    li.chunk_max = 0;
    li.ptr.resize(nl + 1);
    if (l < 2) {
      FFMM_HIP(hipMemsetAsync(li.ptr.p, 0, sizeof(int32_t) * (nl + 1), s));
      li.filled = true;
      continue;
    }
    k_pair_count<<<nblocks(PL * nl, TB), TB, 0, s>>>(t, lo, nl, cnt.p + lo);
    FFMM_LAUNCH_CHECK();
    exclusive_scan_i32_async(ctx, cnt.p + lo, li.ptr.p, nl, l);
  }
  for (int l = 2; l < plan.L; l++) {
    const int64_t np = read_slot(ctx, l);
    if (np >= INT32_MAX) fail(FFMM_ERR_ARGUMENT, "too many far pairs on one level for 32-bit pair indices");
    plan.lists[l].np = np;
  }
  // the pairs themselves: the spectral M2L reads only the counts (its Hadamard works
  // from the colleague tables), so they are written on demand (fill_level_pairs)
  if (plan.m2l != FFMM_M2L_DENSE) return;
  // the class lists (only the dense class translations read them) and the target chunks
  // of all levels, queued at once: one read back of the chunk bounds
  const int64_t cap = m2l_chunk_pairs(plan);
  const int64_t nrow = 1 + nclass;
  std::vector<int64_t> boff(plan.L + 1, 0);
  int64_t nemax = 1;
  for (int l = 0; l < plan.L; l++) {
    const int64_t np = l < 2 ? 0 : plan.lists[l].np;
    const int64_t nch = (np + cap - 1) / cap;
    boff[l + 1] = boff[l] + (l < 2 ? 0 : (nch + 1) * nrow);
    nemax = std::max<int64_t>(nemax, (int64_t)nclass * plan.nl[l]);
  }
  ccnt.resize(nemax);
  bounds.resize(std::max<int64_t>(1, boff[plan.L]));
  for (int l = 2; l < plan.L; l++) {
    LevelLists &li = plan.lists[l];
    const int64_t nl = plan.nl[l], np = li.np;
    fill_level_pairs(plan, l);
    const int64_t ne = (int64_t)nclass * nl;
    li.cstart.resize(ne + 1);
    k_class_count<<<nblocks(64 * nl, TB), TB, 0, s>>>(li.far.p, li.ptr.p, nl, nclass, ops.far_class.p, ccnt.p);
    FFMM_LAUNCH_CHECK();
    // the class-major scan; its total is np (slot 0 is free here)
    exclusive_scan_i32_async(ctx, ccnt.p, li.cstart.p, ne, 0);
    li.csrc.resize(np);
    li.cfar.resize(np);
    // the target of each class-major pair: the harmonic methods' fused mixed class M2L
    // sums a target's consecutive products in a tile
    const bool want_tgt = ops.method == METHOD_HARMONIC && plan.precision == FFMM_PREC_MIXED;
    if (want_tgt) li.ctgt.resize(np);
    if (np > 0) {
      k_class_fill<<<nblocks(64 * nl, TB), TB, 0, s>>>(li.src.p, li.far.p, li.ptr.p, li.cstart.p, nl, nclass,
                                                  ops.far_class.p, li.csrc.p, li.cfar.p,
                                                  want_tgt ? li.ctgt.p : nullptr);
      FFMM_LAUNCH_CHECK();
    }
    const int64_t nch = (np + cap - 1) / cap;
    k_chunk_bounds<<<nblocks(nch + 1, 64), 64, 0, s>>>(li.ptr.p, nl, cap, nch, li.cstart.p, nclass,
                                                       bounds.p + boff[l]);
    FFMM_LAUNCH_CHECK();
  }
  std::vector<int32_t> hb((size_t)std::max<int64_t>(1, boff[plan.L]));
  bounds.download(hb.data(), (size_t)boff[plan.L], s);
  ctx.sync();
  for (int l = 2; l < plan.L; l++) {
    LevelLists &li = plan.lists[l];
    const int64_t nch = (li.np + cap - 1) / cap;
    const int32_t *h = hb.data() + boff[l];
    li.chunk_t.resize(nch + 1);
    li.chunk_cs.resize((size_t)((nch + 1) * nclass));
    for (int64_t k = 0; k <= nch; k++) {
      li.chunk_t[k] = h[k * nrow];
      for (int c = 0; c < nclass; c++) li.chunk_cs[k * nclass + c] = h[k * nrow + 1 + c];
    }
    li.chunk_max = 0;
    for (int64_t k = 0; k < nch; k++) {
      int64_t n = 0;
      for (int c = 0; c < nclass; c++) n += li.chunk_cs[(k + 1) * nclass + c] - li.chunk_cs[k * nclass + c];
      li.chunk_max = std::max(li.chunk_max, n);
    }
    for (int c = 0; c < nclass; c++) li.class_pair[c] = h[1 + c];
    li.class_pair[nclass] = li.np;
  }
}

}  // namespace ffmm
