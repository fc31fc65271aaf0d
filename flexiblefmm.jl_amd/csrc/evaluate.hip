// Evaluation at the sources and at arbitrary targets: the local expansion of the
// target's leaf (L2T over the large surface) plus the near field over the leaf's 27
// colleagues, the target's own source excluded by index (a masked select, never a
// multiply).
//
// One wavefront per real leaf, one lane per target: the surface with the leaf's local
// density, then the colleagues' sources, pass through LDS tiles that the wave loads
// with coalesced reads and every lane then reads at the same address (a broadcast),
// so the 64 lanes only differ in their target.  Targets at arbitrary points are
// located by a device walk with the builder's child-slot rule and bucketed by leaf
// with a stable sort, so a leaf's targets are contiguous.
//
// Follows _evaluate_kernel! (src/device.jl:102-125), evaluate_target / _direct_range
// (src/evaluate.jl:12-58) and locate (src/tree/locate.jl:10-22).
#include "ffmm_internal.h"

#include <algorithm>

namespace ffmm {

constexpr int TB = 256;
constexpr int EV_TB = 64;
constexpr int EV_CK = 256;   // entries of an LDS tile: surface points with densities, or sources with charges
constexpr double C4PI = 0.07957747154594767;   // 1 / (4 pi)

template <int OUT>
__device__ __forceinline__ void accumulate(double d0, double d1, double d2, double w, double acc[4]) {
  double ir = rsqrt_f64(fma(d2, d2, fma(d1, d1, d0 * d0)));
  if (OUT != FFMM_GRADIENT) acc[0] += (C4PI * ir) * w;
  if (OUT != FFMM_POTENTIAL) {
    double a = -C4PI * ir * ir * ir;
    acc[1] += (a * d0) * w;
    acc[2] += (a * d1) * w;
    acc[3] += (a * d2) * w;
  }
}
template <int OUT>
__device__ __forceinline__ void accumulate_masked(double d0, double d1, double d2, double w, bool keep, double acc[4]) {
  double ir = rsqrt_f64(fma(d2, d2, fma(d1, d1, d0 * d0)));
  if (OUT != FFMM_GRADIENT) acc[0] += keep ? (C4PI * ir) * w : 0.0;
  if (OUT != FFMM_POTENTIAL) {
    double a = -C4PI * ir * ir * ir;
    acc[1] += keep ? (a * d0) * w : 0.0;
    acc[2] += keep ? (a * d1) * w : 0.0;
    acc[3] += keep ? (a * d2) * w : 0.0;
  }
}

template <int OUT>
__global__ void __launch_bounds__(EV_TB) k_eval(EvalArgs a) {
  __shared__ double4 tile[EV_CK];
  __shared__ int32_t tile_id[EV_CK];
  __shared__ NearList nl;
  // blocks in leaf order, dealt round-robin to the XCDs: the leaves' work varies by
  // level (they are stored level by level), and round-robin keeps the XCDs balanced
  const int s = blockIdx.x;
  const int i = a.leaf_nodes[s];
  const bool self = a.tptr == nullptr;
  const int32_t t0 = self ? a.plo[i] : a.tptr[s];
  const int32_t t1 = self ? a.phi[i] + 1 : a.tptr[s + 1];
  if (t0 >= t1) return;
  const double c0 = a.centers[i], c1 = a.centers[a.nn + i], c2 = a.centers[2 * a.nn + i];
  const double sc = a.lam[a.level[i]];
  const double *__restrict__ L = a.Lc + (int64_t)s * a.Nl;
  const double *__restrict__ X = a.surf;
  const double *__restrict__ Y = a.surf + a.Nl;
  const double *__restrict__ Z = a.surf + 2 * a.Nl;
  near_list(a, i, nl);
  const int32_t nsrc = nl.pre[NCOLL];
  for (int32_t tb = t0; tb < t1; tb += EV_TB) {
    // a tile of up to 64 targets; when it has fewer, R = 64 / ntp replicas of each
    // target split the surface points and the sources, and are summed at the end
    const int cnt = min(EV_TB, t1 - tb);
    int ntp = 1;
    while (ntp < cnt) ntp <<= 1;
    const int R = EV_TB / ntp;
    const int tl = (int)threadIdx.x % ntp, rep = (int)threadIdx.x / ntp;
    const int32_t j = tb + tl;
    const bool valid = tl < cnt;
    const int32_t jj = valid ? j : t0;
    const double x = self ? a.sx[jj] : a.tx[jj];
    const double y = self ? a.sy[jj] : a.ty[jj];
    const double z = self ? a.sz[jj] : a.tz[jj];
    const double xr = x - c0, yr = y - c1, zr = z - c2;
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    // l2t over the large surface (src/operators/ki_common.jl:293-315): tiles of the
    // scaled surface points and the leaf's densities in LDS, read by every lane
    for (int k0 = 0; k0 < a.Nl; k0 += EV_CK) {
      const int nk = min(EV_CK, a.Nl - k0);
      __syncthreads();
      for (int k = threadIdx.x; k < nk; k += EV_TB)
        tile[k] = make_double4(X[k0 + k] * sc, Y[k0 + k] * sc, Z[k0 + k] * sc, L[k0 + k]);
      __syncthreads();
#pragma unroll 4
      for (int k = rep; k < nk; k += R) {
        const double4 v = tile[k];
        accumulate<OUT>(xr - v.x, yr - v.y, zr - v.z, v.w, acc);
      }
    }
    // the near field of the 27 colleagues (src/evaluate.jl:12-25, 37-58), their
    // sources as one list through the same tiles
    const int32_t skip = self ? j : -1;
    for (int32_t q0 = 0; q0 < nsrc; q0 += EV_CK) {
      const int nq = min(EV_CK, nsrc - q0);
      __syncthreads();
      for (int k = threadIdx.x; k < nq; k += EV_TB) {
        const int32_t q = near_source(nl, q0 + k);
        tile[k] = make_double4(a.sx[q], a.sy[q], a.sz[q], a.tau[q]);
        tile_id[k] = q;
      }
      __syncthreads();
#pragma unroll 4
      for (int k = rep; k < nq; k += R) {
        const double4 v = tile[k];
        accumulate_masked<OUT>(x - v.x, y - v.y, z - v.z, v.w, tile_id[k] != skip, acc);
      }
    }
    for (int off = ntp; off < EV_TB; off <<= 1)
#pragma unroll
      for (int c = 0; c < 4; c++) acc[c] += __shfl_xor(acc[c], off, EV_TB);
    if (valid && rep == 0) {
      const int64_t o = self ? a.outidx[j] : a.outidx[j];
      if (OUT == FFMM_POTENTIAL) {
        a.out[o] = acc[0];
      } else if (OUT == FFMM_GRADIENT) {
        a.out[3 * o] = acc[1];
        a.out[3 * o + 1] = acc[2];
        a.out[3 * o + 2] = acc[3];
      } else {
        a.out[4 * o] = acc[0];
        a.out[4 * o + 1] = acc[1];
        a.out[4 * o + 2] = acc[2];
        a.out[4 * o + 3] = acc[3];
      }
    }
  }
}

// The mixed form: the same traversal with FP32 arithmetic.  L2T works in coordinates
// relative to the leaf centre (surface points and the target's offset rounded once to
// FP32; they are at least ~w/2 apart).  The near field forms each pair difference in FP64
// before rounding it, so close pairs keep their relative accuracy.  1/r is v_rsq_f32.
// The potential is summed with Kahan compensation in FP32, the gradient with plain FP32
// FMAs; both partial sums are folded into FP64 accumulators after every LDS tile.
__device__ __forceinline__ void kahan_add(float &s, float &c, float y) {   // y already has -c folded in
  const float t = s + y;
  c = (t - s) - y;
  s = t;
}

template <int OUT>
__device__ __forceinline__ void accumulate32(float d0, float d1, float d2, float w, float &s, float &c, float g[3]) {
  const float ir = __builtin_amdgcn_rsqf(fmaf(d2, d2, fmaf(d1, d1, d0 * d0)));
  if (OUT != FFMM_GRADIENT) kahan_add(s, c, fmaf(ir, w, -c));
  if (OUT != FFMM_POTENTIAL) {
    const float a = ir * ir * ir * w;
    g[0] = fmaf(a, d0, g[0]);
    g[1] = fmaf(a, d1, g[1]);
    g[2] = fmaf(a, d2, g[2]);
  }
}

template <int OUT>
__device__ __forceinline__ void fold32(float &s, float &c, float g[3], double acc[4]) {
  acc[0] += (double)s - (double)c;
  acc[1] += (double)g[0];
  acc[2] += (double)g[1];
  acc[3] += (double)g[2];
  s = c = 0.f;
  g[0] = g[1] = g[2] = 0.f;
}

// a near-field source in LDS: FP64 position (the pair difference is formed in FP64),
// FP32 charge, and its index for the self-exclusion
struct NearSrc32 {
  double x, y, z;
  float w;
  int32_t id;
};

template <int OUT>
__global__ void __launch_bounds__(EV_TB) k_eval32(EvalArgs a) {
  // the potential alone: tiles of 128 entries (LDS 4.3 KB per workgroup, so more
  // workgroups per CU: 4.60 -> 4.15 ms at 4M sources); with the gradient the tiles of 256
  // measured faster
  constexpr int CK = OUT == FFMM_POTENTIAL ? EV_CK / 2 : EV_CK;
  __shared__ NearSrc32 tile[CK];
  __shared__ NearList nl;
  float4 *tile32 = reinterpret_cast<float4 *>(tile);
  const int s = blockIdx.x;
  const int i = a.leaf_nodes[s];
  const bool self = a.tptr == nullptr;
  const int32_t t0 = self ? a.plo[i] : a.tptr[s];
  const int32_t t1 = self ? a.phi[i] + 1 : a.tptr[s + 1];
  if (t0 >= t1) return;
  const double c0 = a.centers[i], c1 = a.centers[a.nn + i], c2 = a.centers[2 * a.nn + i];
  const double sc = a.lam[a.level[i]];
  const double *__restrict__ L = a.Lc + (int64_t)s * a.Nl;
  const double *__restrict__ X = a.surf;
  const double *__restrict__ Y = a.surf + a.Nl;
  const double *__restrict__ Z = a.surf + 2 * a.Nl;
  near_list(a, i, nl);
  const int32_t nsrc = nl.pre[NCOLL];
  for (int32_t tb = t0; tb < t1; tb += EV_TB) {
    const int cnt = min(EV_TB, t1 - tb);
    int ntp = 1;
    while (ntp < cnt) ntp <<= 1;
    const int R = EV_TB / ntp;
    const int tl = (int)threadIdx.x % ntp, rep = (int)threadIdx.x / ntp;
    const int32_t j = tb + tl;
    const bool valid = tl < cnt;
    const int32_t jj = valid ? j : t0;
    const double x = self ? a.sx[jj] : a.tx[jj];
    const double y = self ? a.sy[jj] : a.ty[jj];
    const double z = self ? a.sz[jj] : a.tz[jj];
    const float xr = (float)(x - c0), yr = (float)(y - c1), zr = (float)(z - c2);
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    float ks = 0.f, kc = 0.f, g[3] = {0.f, 0.f, 0.f};
    for (int k0 = 0; k0 < a.Nl; k0 += CK) {
      const int nk = min(CK, a.Nl - k0);
      __syncthreads();
      for (int k = threadIdx.x; k < nk; k += EV_TB)
        tile32[k] = make_float4((float)(X[k0 + k] * sc), (float)(Y[k0 + k] * sc), (float)(Z[k0 + k] * sc),
                                (float)L[k0 + k]);
      __syncthreads();
#pragma unroll 4
      for (int k = rep; k < nk; k += R) {
        const float4 v = tile32[k];
        accumulate32<OUT>(xr - v.x, yr - v.y, zr - v.z, v.w, ks, kc, g);
      }
      fold32<OUT>(ks, kc, g, acc);
    }
    // the near field of the 27 colleagues (src/evaluate.jl:12-25, 37-58), their
    // sources as one list through the same tiles
    const int32_t skip = self ? j : -1;
    for (int32_t q0 = 0; q0 < nsrc; q0 += CK) {
      const int nq = min(CK, nsrc - q0);
      __syncthreads();
      for (int k = threadIdx.x; k < nq; k += EV_TB) {
        const int32_t q = near_source(nl, q0 + k);
        tile[k] = NearSrc32{a.sx[q], a.sy[q], a.sz[q], (float)a.tau[q], q};
      }
      __syncthreads();
#pragma unroll 4
      for (int k = rep; k < nq; k += R) {
        const NearSrc32 v = tile[k];
        // the target's own source: an infinite distance, so 1/r is exactly zero
        const float d0 = (float)(x - v.x), d1 = (float)(y - v.y), d2 = (float)(z - v.z);
        const float ir = __builtin_amdgcn_rsqf(v.id != skip ? fmaf(d2, d2, fmaf(d1, d1, d0 * d0)) : INFINITY);
        if (OUT != FFMM_GRADIENT) kahan_add(ks, kc, fmaf(ir, v.w, -kc));
        if (OUT != FFMM_POTENTIAL) {
          const float aa = ir * ir * ir * v.w;
          g[0] = fmaf(aa, d0, g[0]);
          g[1] = fmaf(aa, d1, g[1]);
          g[2] = fmaf(aa, d2, g[2]);
        }
      }
      fold32<OUT>(ks, kc, g, acc);
    }
    for (int off = ntp; off < EV_TB; off <<= 1)
#pragma unroll
      for (int c = 0; c < 4; c++) acc[c] += __shfl_xor(acc[c], off, EV_TB);
    if (valid && rep == 0) {
      const int64_t o = a.outidx[j];
      if (OUT == FFMM_POTENTIAL) {
        a.out[o] = C4PI * acc[0];
      } else if (OUT == FFMM_GRADIENT) {
        a.out[3 * o] = -C4PI * acc[1];
        a.out[3 * o + 1] = -C4PI * acc[2];
        a.out[3 * o + 2] = -C4PI * acc[3];
      } else {
        a.out[4 * o] = C4PI * acc[0];
        a.out[4 * o + 1] = -C4PI * acc[1];
        a.out[4 * o + 2] = -C4PI * acc[2];
        a.out[4 * o + 3] = -C4PI * acc[3];
      }
    }
  }
}

static void launch_eval(Plan &plan, EvalArgs &a, int output) {
  const int64_t nleaves = plan.leaf_off[plan.L];
  if (nleaves == 0) return;
  if (plan.ops->method == METHOD_HARMONIC) {
    launch_eval_harmonic(plan, a, output);
    return;
  }
  hipStream_t s = plan.ctx->stream;
  const unsigned g = (unsigned)nleaves;
  if (plan.precision == FFMM_PREC_MIXED) {
    if (output == FFMM_POTENTIAL) k_eval32<FFMM_POTENTIAL><<<g, EV_TB, 0, s>>>(a);
    else if (output == FFMM_GRADIENT) k_eval32<FFMM_GRADIENT><<<g, EV_TB, 0, s>>>(a);
    else k_eval32<2><<<g, EV_TB, 0, s>>>(a);
  } else if (output == FFMM_POTENTIAL) k_eval<FFMM_POTENTIAL><<<g, EV_TB, 0, s>>>(a);
  else if (output == FFMM_GRADIENT) k_eval<FFMM_GRADIENT><<<g, EV_TB, 0, s>>>(a);
  else k_eval<2><<<g, EV_TB, 0, s>>>(a);
  FFMM_LAUNCH_CHECK();
}

static EvalArgs base_args(Plan &plan, double *out) {
  EvalArgs a{};
  a.leaf_nodes = plan.leaf_nodes.p;
  a.Lc = plan.Lc.p;
  a.Nl = plan.ops->Nl;
  a.surf = plan.ops->large.p;
  a.lam = plan.lam_d.p;
  a.level = plan.level.p;
  a.centers = plan.centers.p;
  a.nn = plan.nnodes;
  a.col = plan.col.p;
  a.plo = plan.plo.p;
  a.phi = plan.phi.p;
  a.sx = plan.sx.p;
  a.sy = plan.sy.p;
  a.sz = plan.sz.p;
  a.tau = plan.tau.p;
  a.out = out;
  if (plan.ops->method == METHOD_HARMONIC) {
    a.tb = plan.ops->tab();
    a.tb32 = plan.ops->tab32();
    a.p = plan.ops->p;
    a.nd = plan.nd;
    a.hscale = plan.hscale.p;
    a.khat = plan.khat.p;
    a.kappa = plan.ops->kappa;
  }
  return a;
}

// evaluate_at_sources! (src/device.jl:153-163): out in the user's order via perm
void evaluate_at_sources(Plan &plan, double *out, int output) {
  EvalArgs a = base_args(plan, out);
  a.tptr = nullptr;
  a.outidx = plan.perm.p;
  launch_eval(plan, a, output);
}

// ---------------------------------------------------------------------------
// targets: locate (src/tree/locate.jl:10-22), bucket_by_node (src/tree/locate.jl:49-67)
// ---------------------------------------------------------------------------
__global__ void k_locate(const double *__restrict__ t, int64_t nt, double ox, double oy, double oz, double W0,
                         const double *__restrict__ centers, int64_t nn, const uint8_t *__restrict__ leaf,
                         const int32_t *__restrict__ child_first, const int32_t *__restrict__ node_leafslot,
                         uint32_t *__restrict__ key, int32_t *__restrict__ idx) {
  int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nt) return;
  double x = t[3 * q], y = t[3 * q + 1], z = t[3 * q + 2];
  bool inside = ox <= x && x <= ox + W0 && oy <= y && y <= oy + W0 && oz <= z && z <= oz + W0;
  uint32_t k = 0xFFFFFFFFu;
  if (inside) {
    int32_t i = 0;
    while (!leaf[i]) {
      int slot = (x > centers[i]) | ((y > centers[nn + i]) << 1) | ((z > centers[2 * nn + i]) << 2);
      i = child_first[i] + slot;
    }
    k = (uint32_t)node_leafslot[i];
  }
  key[q] = k;
  idx[q] = (int32_t)q;
}

// targets located by the caller (the reference's device evaluate! takes the leaves from
// the host locate, src/device.jl:127-145): the key is the leaf slot of the given node;
// a node that is not a real leaf of the tree gets the invalid key
__global__ void k_node_keys(const int32_t *__restrict__ nodes, int64_t nt, int64_t nn,
                            const int32_t *__restrict__ node_leafslot, uint32_t *__restrict__ key,
                            int32_t *__restrict__ idx) {
  int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nt) return;
  const int32_t i = nodes[q];
  uint32_t k = 0xFFFFFFFFu;
  if (i >= 0 && i < nn) {
    const int32_t s = node_leafslot[i];
    if (s >= 0) k = (uint32_t)s;
  }
  key[q] = k;
  idx[q] = (int32_t)q;
}

// first position of each leaf slot in the sorted keys
__global__ void k_bucket_ptr(const uint32_t *__restrict__ skeys, int64_t nt, int64_t nleaves, int32_t *__restrict__ ptr) {
  int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s > nleaves) return;
  int64_t a = 0, b = nt;
  while (a < b) {
    int64_t m = (a + b) >> 1;
    if ((int64_t)skeys[m] < s) a = m + 1;
    else b = m;
  }
  ptr[s] = (int32_t)a;
}

__global__ void k_gather_targets(const double *__restrict__ t, const int32_t *__restrict__ order, int64_t nt,
                                 double *__restrict__ tx, double *__restrict__ ty, double *__restrict__ tz) {
  int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nt) return;
  int64_t o = order[q];
  tx[q] = t[3 * o];
  ty[q] = t[3 * o + 1];
  tz[q] = t[3 * o + 2];
}

__global__ void k_last_key(const uint32_t *__restrict__ key, int64_t *__restrict__ host) {
  if (threadIdx.x == 0) host[0] = (int64_t)key[0];
}

void evaluate_targets(Plan &plan, const double *targets, int64_t nt, double *out, int output, const int32_t *nodes) {
  if (nt <= 0) return;
  if (nt >= INT32_MAX) fail(FFMM_ERR_ARGUMENT, "too many targets for 32-bit indices");
  Context &ctx = *plan.ctx;
  hipStream_t s = ctx.stream;
  const int64_t nleaves = plan.leaf_off[plan.L];
  DBuf<uint32_t> keys, skeys;
  DBuf<int32_t> idx;
  keys.resize(nt);
  skeys.resize(nt);
  idx.resize(nt);
  plan.tgt_order.resize(nt);
  if (nodes)
    k_node_keys<<<nblocks(nt, TB), TB, 0, s>>>(nodes, nt, plan.nnodes, plan.node_leafslot.p, keys.p, idx.p);
  else
    k_locate<<<nblocks(nt, TB), TB, 0, s>>>(targets, nt, plan.origin[0], plan.origin[1], plan.origin[2], plan.W0,
                                            plan.centers.p, plan.nnodes, plan.leaf.p, plan.child_first.p,
                                            plan.node_leafslot.p, keys.p, idx.p);
  FFMM_LAUNCH_CHECK();
  sort_pairs_u32(ctx, keys.p, skeys.p, idx.p, plan.tgt_order.p, nt, 32);
  // the largest key (outside the box: all ones), read from pinned memory
  arm_slots(ctx, 0, 1);
  k_last_key<<<1, 64, 0, s>>>(skeys.p + nt - 1, ctx.pinned_dev);
  FFMM_LAUNCH_CHECK();
  if ((uint32_t)read_slot(ctx, 0) == 0xFFFFFFFFu)
    fail(FFMM_ERR_ARGUMENT, nodes ? "a target's node is not a real leaf of the tree" : "a target lies outside the tree's box");
  plan.tgt_ptr.resize(nleaves + 1);
  k_bucket_ptr<<<nblocks(nleaves + 1, TB), TB, 0, s>>>(skeys.p, nt, nleaves, plan.tgt_ptr.p);
  plan.tgt_sorted.resize(3 * nt);
  double *tx = plan.tgt_sorted.p, *ty = tx + nt, *tz = ty + nt;
  k_gather_targets<<<nblocks(nt, TB), TB, 0, s>>>(targets, plan.tgt_order.p, nt, tx, ty, tz);
  FFMM_LAUNCH_CHECK();
  EvalArgs a = base_args(plan, out);
  a.tptr = plan.tgt_ptr.p;
  a.tx = tx;
  a.ty = ty;
  a.tz = tz;
  a.outidx = plan.tgt_order.p;
  launch_eval(plan, a, output);
}

}  // namespace ffmm
