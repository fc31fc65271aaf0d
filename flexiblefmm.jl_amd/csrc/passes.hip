// The upward and downward passes on the device.
//
// Upward (levels L-1 .. 2, 0-based): S2M of the childless sourced nodes (check
// potentials on the large surface, one workgroup per node with the node's sources
// staged in LDS), the factored pseudo-inverse UC2E as two FP64 GEMMs, and M2M as one
// GEMM per level straight on the children block of the multipole array (the children
// of a sibling group are consecutive columns).  Downward (levels 2 .. L-1): L2L as one
// GEMM whose output is the level's block of partial locals, the far field (M2L), and
// the pseudo-inverse DC2E of the real leaves, written straight into the compact leaf
// locals.  Follows device_build_expansions! (src/device_passes.jl:511-597) and the
// host passes it restates (src/fmm.jl:170-262).
#include "ffmm_internal.h"

#include <algorithm>
#include <cmath>

namespace ffmm {

constexpr int TB = 256;
constexpr double C4PI = 0.07957747154594767;   // 1 / (4 pi), as 1 / (4 * Float64(pi))

// ---------------------------------------------------------------------------
// S2M: s2m_accumulate! (src/operators/ki_common.jl:256-272), _s2m_kernel!
// (src/device_passes.jl:378-403).  pM[k, q] = sum_j tau_j K(x_j - c, y_k).
// ---------------------------------------------------------------------------
constexpr int S2M_TB = 256;
constexpr int S2M_KP = 6;   // check points per thread: ceil(1352 / 256)

__global__ void __launch_bounds__(S2M_TB) k_s2m(const int32_t *__restrict__ nodes, const double *__restrict__ centers,
                                                int64_t nn, const int32_t *__restrict__ level,
                                                const double *__restrict__ lam, const double *__restrict__ surf, int Nl,
                                                const int32_t *__restrict__ plo, const int32_t *__restrict__ phi,
                                                const double *__restrict__ sx, const double *__restrict__ sy,
                                                const double *__restrict__ sz, const double *__restrict__ tau,
                                                double *__restrict__ pM) {
  __shared__ double xs[S2M_TB], ys[S2M_TB], zs[S2M_TB], ts[S2M_TB];
  const int64_t q = blockIdx.x;
  const int node = nodes[q];
  const double c0 = centers[node], c1 = centers[nn + node], c2 = centers[2 * nn + node];
  const double s = lam[level[node]];
  const int t = threadIdx.x;
  double acc[S2M_KP], y0[S2M_KP], y1[S2M_KP], y2[S2M_KP];
#pragma unroll
  for (int m = 0; m < S2M_KP; m++) {
    int k = t + S2M_TB * m;
    acc[m] = 0.0;
    if (k < Nl) {
      y0[m] = surf[k] * s;
      y1[m] = surf[Nl + k] * s;
      y2[m] = surf[2 * Nl + k] * s;
    } else {
      y0[m] = y1[m] = y2[m] = 1.0;
    }
  }
  const int32_t a = plo[node], b = phi[node];
  for (int32_t j0 = a; j0 <= b; j0 += S2M_TB) {
    int32_t j = j0 + t;
    if (j <= b) {
      xs[t] = sx[j] - c0;
      ys[t] = sy[j] - c1;
      zs[t] = sz[j] - c2;
      ts[t] = tau[j];
    }
    __syncthreads();
    int cnt = min(S2M_TB, b - j0 + 1);
    for (int jj = 0; jj < cnt; jj++) {
      const double xr = xs[jj], yr = ys[jj], zr = zs[jj], tj = ts[jj];
#pragma unroll
      for (int m = 0; m < S2M_KP; m++) {
        double d0 = y0[m] - xr, d1 = y1[m] - yr, d2 = y2[m] - zr;
        double r2 = fma(d2, d2, fma(d1, d1, d0 * d0));
        acc[m] += (C4PI * rsqrt_f64(r2)) * tj;
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int m = 0; m < S2M_KP; m++) {
    int k = t + S2M_TB * m;
    if (k < Nl) pM[q * Nl + k] = acc[m];
  }
}

// The mixed form of S2M: sources relative to the node centre (differences formed in
// FP64, rounded once), surface points scaled and rounded once, v_rsq_f32, FP32 sums over
// each LDS tile folded into FP64 check potentials.  The check surface lies at least
// ~0.9 w from any source of the node, so no pair is close.
__global__ void __launch_bounds__(S2M_TB) k_s2m32(const int32_t *__restrict__ nodes, const double *__restrict__ centers,
                                                  int64_t nn, const int32_t *__restrict__ level,
                                                  const double *__restrict__ lam, const double *__restrict__ surf, int Nl,
                                                  const int32_t *__restrict__ plo, const int32_t *__restrict__ phi,
                                                  const double *__restrict__ sx, const double *__restrict__ sy,
                                                  const double *__restrict__ sz, const double *__restrict__ tau,
                                                  double *__restrict__ pM) {
  __shared__ float4 src[S2M_TB];
  const int64_t q = blockIdx.x;
  const int node = nodes[q];
  const double c0 = centers[node], c1 = centers[nn + node], c2 = centers[2 * nn + node];
  const double s = lam[level[node]];
  const int t = threadIdx.x;
  double acc[S2M_KP];
  float y0[S2M_KP], y1[S2M_KP], y2[S2M_KP];
#pragma unroll
  for (int m = 0; m < S2M_KP; m++) {
    const int k = t + S2M_TB * m;
    acc[m] = 0.0;
    if (k < Nl) {
      y0[m] = (float)(surf[k] * s);
      y1[m] = (float)(surf[Nl + k] * s);
      y2[m] = (float)(surf[2 * Nl + k] * s);
    } else {
      y0[m] = y1[m] = y2[m] = 1.f;
    }
  }
  const int32_t a = plo[node], b = phi[node];
  for (int32_t j0 = a; j0 <= b; j0 += S2M_TB) {
    const int32_t j = j0 + t;
    if (j <= b) src[t] = make_float4((float)(sx[j] - c0), (float)(sy[j] - c1), (float)(sz[j] - c2), (float)tau[j]);
    __syncthreads();
    const int cnt = min(S2M_TB, b - j0 + 1);
    float part[S2M_KP];
#pragma unroll
    for (int m = 0; m < S2M_KP; m++) part[m] = 0.f;
    for (int jj = 0; jj < cnt; jj++) {
      const float4 v = src[jj];
#pragma unroll
      for (int m = 0; m < S2M_KP; m++) {
        const float d0 = y0[m] - v.x, d1 = y1[m] - v.y, d2 = y2[m] - v.z;
        part[m] = fmaf(__builtin_amdgcn_rsqf(fmaf(d2, d2, fmaf(d1, d1, d0 * d0))), v.w, part[m]);
      }
    }
#pragma unroll
    for (int m = 0; m < S2M_KP; m++) acc[m] += (double)part[m];
    __syncthreads();
  }
#pragma unroll
  for (int m = 0; m < S2M_KP; m++) {
    const int k = t + S2M_TB * m;
    if (k < Nl) pM[q * Nl + k] = C4PI * acc[m];
  }
}

// S2M fused with the first factor of UC2E: tup[:, q] = Uh pM[:, q] for the childless
// sourced nodes q (src/device_passes.jl:378-403 and 496-502: s2m_accumulate! into the
// check potentials, then the factored pseudo-inverse's first GEMM), without the check
// potentials pM ever reaching device memory (Nl doubles per node: 1352 at p = 8, 4.9 GB
// at 451k nodes, written once and read once by a library GEMM).
//
// A workgroup takes 64 nodes (the product's columns) and stages their sources in LDS,
// relative to each node's centre.  Its 16 waves split the work by unit, four per SIMD:
// waves 8..15 compute the check potentials of 64 check points x 64 nodes per chunk into
// LDS (a wave one node at a time, a lane per check point, so every source read is a
// broadcast and the lanes loop together), with the arithmetic of k_s2m (FP64) or k_s2m32
// (the mixed form: FP32 terms over at most 256 sources, folded into FP64); waves 0..7
// multiply the previous chunk into the product's r x 64 block on FP64 MFMA
// (v_mfma_f64_16x16x4f64, r <= 304 in 20 row tiles of 16: wave w takes tiles w, w + 8 and,
// for w < 4, w + 16, so each SIMD holds five; Uh read from L2 a k-step ahead, from a
// zero-padded copy).  The two buffers of check potentials alternate, one barrier per
// chunk, so the VALU work of one chunk runs beside the MFMAs of the previous one on
// every SIMD.  A workgroup whose nodes hold more
// sources than the LDS stage reads them from global memory instead.
// ---------------------------------------------------------------------------
constexpr int SU_NT = 64;                              // nodes per workgroup: 4 column tiles
constexpr int SU_TB = 1024;                            // waves 0..7 MFMA, 8..15 check potentials
constexpr int SU_RMIN = 241, SU_RMAX = 304;            // the ranks r this kernel takes
constexpr int SU_RW = 3;                               // row tiles per MFMA wave: 3 (w < 4) or 2
constexpr int SU_KC = 64;                              // check points per chunk: 16 k-steps
constexpr int SU_PS = SU_NT + 1;                       // LDS row stride of the check potentials
constexpr int SU_PD = 2;                               // k-steps of Uh loads in flight
constexpr int SU_FOLD = 256;                           // FP32 terms per fold (the mixed form)
typedef double su_f64x4 __attribute__((ext_vector_type(4)));
static_assert(SU_KC == UC2E_PK && 16 * 20 == UC2E_PROWS && SU_RMAX <= UC2E_PROWS && SU_PD == UC2E_PKX &&
                  (SU_KC / 4) % SU_PD == 0,
              "the padded Uh matches the kernel's tiles and prefetch");

template <bool F32>
struct SuSrc;
template <>
struct SuSrc<true> {
  using T = float4;
  static constexpr int CAP = 4096;   // 64 nodes of at most 64 sources
};
template <>
struct SuSrc<false> {
  using T = double4;
  static constexpr int CAP = 2048;
};

// the check potential at (z0, z1, z2) (scaled, relative to the node centre) of the
// node's sources [j0, j1), fetched by src(j): k_s2m32's FP32 terms folded into FP64 every
// SU_FOLD sources, or k_s2m's FP64 terms; four sources per step, summed in order
template <bool F32, class Fetch>
__device__ __forceinline__ double su_potential(double y0, double y1, double y2, double sc, int32_t j0, int32_t j1,
                                               Fetch &&src) {
  if constexpr (F32) {
    const float z0 = (float)(y0 * sc), z1 = (float)(y1 * sc), z2 = (float)(y2 * sc);
    auto term = [&](const float4 &v) {
      const float d0 = z0 - v.x, d1 = z1 - v.y, d2 = z2 - v.z;
      return __builtin_amdgcn_rsqf(fmaf(d2, d2, fmaf(d1, d1, d0 * d0)));
    };
    double a = 0.0;
    for (int32_t jb = j0; jb < j1; jb += SU_FOLD) {
      const int32_t je = min(j1, jb + SU_FOLD);
      float part = 0.f;
      int32_t j = jb;
      for (; j + 4 <= je; j += 4) {
        const float4 v0 = src(j), v1 = src(j + 1), v2 = src(j + 2), v3 = src(j + 3);
        const float r0 = term(v0), r1 = term(v1), r2 = term(v2), r3 = term(v3);
        part = fmaf(r0, v0.w, part);
        part = fmaf(r1, v1.w, part);
        part = fmaf(r2, v2.w, part);
        part = fmaf(r3, v3.w, part);
      }
      for (; j < je; j++) {
        const float4 v = src(j);
        part = fmaf(term(v), v.w, part);
      }
      a += (double)part;
    }
    return C4PI * a;
  } else {
    const double z0 = y0 * sc, z1 = y1 * sc, z2 = y2 * sc;
    auto term = [&](const double4 &v) {
      const double d0 = z0 - v.x, d1 = z1 - v.y, d2 = z2 - v.z;
      return (C4PI * rsqrt_f64(fma(d2, d2, fma(d1, d1, d0 * d0)))) * v.w;
    };
    double a = 0.0;
    int32_t j = j0;
    for (; j + 4 <= j1; j += 4) {
      const double4 v0 = src(j), v1 = src(j + 1), v2 = src(j + 2), v3 = src(j + 3);
      const double t0 = term(v0), t1 = term(v1), t2 = term(v2), t3 = term(v3);
      a += t0;
      a += t1;
      a += t2;
      a += t3;
    }
    for (; j < j1; j++) a += term(src(j));
    return a;
  }
}

template <bool F32>
__global__ void __launch_bounds__(SU_TB)
k_s2m_uc2e(const int32_t *__restrict__ nodes, int64_t nq, const double *__restrict__ centers, int64_t nn,
           const int32_t *__restrict__ level, const double *__restrict__ lam, const double *__restrict__ surf, int Nl,
           const int32_t *__restrict__ plo, const int32_t *__restrict__ phi, const double *__restrict__ sx,
           const double *__restrict__ sy, const double *__restrict__ sz, const double *__restrict__ tau,
           const double *__restrict__ Uhp, int r, double *__restrict__ tup) {
  using ST = typename SuSrc<F32>::T;
  constexpr int CAP = SuSrc<F32>::CAP;
  __shared__ ST src[CAP];
  __shared__ __attribute__((aligned(16))) double P[2][SU_KC * SU_PS];
  __shared__ int32_t off[SU_NT + 1];      // node n's sources: [off[n], off[n + 1]) of the tile's list
  __shared__ int32_t first[SU_NT];        // node n's first source in the sorted points
  __shared__ double scl[SU_NT], ctr[3][SU_NT];
  const int t = threadIdx.x, lane = t & 63;
  const int w = __builtin_amdgcn_readfirstlane(t >> 6);
  const int64_t q0 = (int64_t)blockIdx.x * SU_NT;
  // the nodes: source ranges and their exclusive scan, centres and scales
  if (t < 64) {
    const int64_t q = q0 + t;
    int32_t cnt = 0, a = 0;
    double sc = 1.0, c0 = 0.0, c1 = 0.0, c2 = 0.0;
    if (q < nq) {
      const int node = nodes[q];
      a = plo[node];
      cnt = phi[node] - a + 1;
      sc = lam[level[node]];
      c0 = centers[node];
      c1 = centers[nn + node];
      c2 = centers[2 * nn + node];
    }
    int32_t v = cnt;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const int32_t u = __shfl_up(v, o, 64);
      if (lane >= o) v += u;
    }
    off[t + 1] = v;
    if (t == 0) off[0] = 0;
    first[t] = a;
    scl[t] = sc;
    ctr[0][t] = c0;
    ctr[1][t] = c1;
    ctr[2][t] = c2;
  }
  __syncthreads();
  const int32_t total = off[SU_NT];
  const bool staged = total <= CAP;
  if (staged) {
    // entry i of the tile's list belongs to the last node n with off[n] <= i
    for (int32_t i = t; i < total; i += SU_TB) {
      int n = 0, h = SU_NT;
      while (h - n > 1) {
        const int m = (n + h) >> 1;
        if (off[m] <= i) n = m;
        else h = m;
      }
      const int32_t j = first[n] + (i - off[n]);
      if constexpr (F32)
        src[i] = make_float4((float)(sx[j] - ctr[0][n]), (float)(sy[j] - ctr[1][n]), (float)(sz[j] - ctr[2][n]),
                             (float)tau[j]);
      else
        src[i] = make_double4(sx[j] - ctr[0][n], sy[j] - ctr[1][n], sz[j] - ctr[2][n], tau[j]);
    }
  }
  const int nch = (Nl + SU_KC - 1) / SU_KC;
  const int RT = (r + 15) >> 4;
  if (w >= 8) {
    // check potentials: wave 8 + pw takes the nodes pw, pw + 8, ..; lane = check point
    const int pw = w - 8;
    __syncthreads();   // the staged sources
    double ny0 = 0.0, ny1 = 0.0, ny2 = 0.0;
    if (lane < Nl) {
      ny0 = surf[lane];
      ny1 = surf[Nl + lane];
      ny2 = surf[2 * Nl + lane];
    }
    for (int ch = 0; ch < nch; ch++) {
      const int k = ch * SU_KC + lane;
      const bool kin = k < Nl;
      const double y0 = ny0, y1 = ny1, y2 = ny2;
      if (k + SU_KC < Nl) {   // the next chunk's check point
        ny0 = surf[k + SU_KC];
        ny1 = surf[Nl + k + SU_KC];
        ny2 = surf[2 * Nl + k + SU_KC];
      }
      double *Pb = P[ch & 1];
      for (int n = pw; n < SU_NT; n += 8) {
        const int32_t j0 = off[n], j1 = off[n + 1];
        double pot;
        if (staged) {
          pot = su_potential<F32>(y0, y1, y2, scl[n], j0, j1, [&](int32_t j) -> ST { return src[j]; });
        } else {
          const int32_t fj = first[n] - j0;
          const double c0 = ctr[0][n], c1 = ctr[1][n], c2 = ctr[2][n];
          pot = su_potential<F32>(y0, y1, y2, scl[n], j0, j1, [&](int32_t j) -> ST {
            const int32_t g = fj + j;
            if constexpr (F32)
              return make_float4((float)(sx[g] - c0), (float)(sy[g] - c1), (float)(sz[g] - c2), (float)tau[g]);
            else
              return make_double4(sx[g] - c0, sy[g] - c1, sz[g] - c2, tau[g]);
          });
        }
        Pb[lane * SU_PS + n] = kin ? pot : 0.0;
      }
      __syncthreads();   // chunk ch is in P[ch & 1]; the MFMA waves are done with P[(ch + 1) & 1]
    }
    __syncthreads();     // matches the MFMA waves' last chunk
    return;
  }
  // the product's accumulators: row tiles w + 8 i, column tiles c; operand lane (lk, li):
  // A row / B column li, k = lk
  const int li = lane & 15, lk = lane >> 4;
  su_f64x4 acc[SU_RW][4];
#pragma unroll
  for (int i = 0; i < SU_RW; i++)
#pragma unroll
    for (int c = 0; c < 4; c++) acc[i][c] = su_f64x4{0.0, 0.0, 0.0, 0.0};
  // A = Uh[row, kg] for k-step g of the whole sweep, from the zero-padded copy (320 rows,
  // Nl rounded up to the chunk and one k-step more): unguarded loads, so that the next
  // k-step's stay in flight across this one's MFMAs
  auto loadA = [&](int g, double *dst) {
    const double *__restrict__ u = Uhp + (int64_t)(4 * g + lk) * UC2E_PROWS + 16 * w + li;
#pragma unroll
    for (int i = 0; i < SU_RW; i++)
      if (i < 2 || w < 4) dst[i] = u[128 * i];
  };
  // the operands of the next SU_PD k-steps in flight (a ring indexed by the k-step)
  double ring[SU_PD][SU_RW];
#pragma unroll
  for (int g = 0; g < SU_PD; g++) loadA(g, ring[g]);
  __syncthreads();       // the staged sources
  for (int ch = 0; ch < nch; ch++) {
    __syncthreads();     // chunk ch is in P[ch & 1]
    const double *Pb = P[ch & 1];
#pragma unroll
    for (int ks = 0; ks < SU_KC / 4; ks++) {
      double *a = ring[ks % SU_PD];
      double b[4];
#pragma unroll
      for (int c = 0; c < 4; c++) b[c] = Pb[(4 * ks + lk) * SU_PS + 16 * c + li];
      // tiles at or beyond r multiply zeros (the host takes this kernel for 240 < r <= 304)
#pragma unroll
      for (int i = 0; i < SU_RW; i++)
        if (i < 2 || w < 4)
#pragma unroll
          for (int c = 0; c < 4; c++) acc[i][c] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i], b[c], acc[i][c], 0, 0, 0);
      loadA(ch * (SU_KC / 4) + ks + SU_PD, a);   // the slot's next k-step
    }
  }
  __syncthreads();       // matches the check-potential waves' last barrier
  // lane (lk, li) holds rows lk + 4 e of column li of each tile (the FP64 MFMA's layout)
#pragma unroll
  for (int i = 0; i < SU_RW; i++) {
    const int rt = w + 8 * i;
    if (rt >= RT || (i == 2 && w >= 4)) continue;
#pragma unroll
    for (int c = 0; c < 4; c++) {
      const int64_t q = q0 + 16 * c + li;
      if (q >= nq) continue;
#pragma unroll
      for (int e = 0; e < 4; e++) {
        const int row = 16 * rt + lk + 4 * e;
        if (row < r) tup[q * r + row] = acc[i][c][e];
      }
    }
  }
}

// The FP64 GEMMs of the pseudo-inverses whose columns belong to scattered nodes, with the
// gather or the scatter of those columns inside: MODE 0, UC2E's second factor
// (src/device_passes.jl:496-502), Mp[:, node q] = lam(q) (V diag(invS)) tup[:, q], written
// straight into the multipoles (and their FP32 mirror); MODE 1, DC2E's first factor
// (src/device_passes.jl:583-593), tup[:, q] = Uh (lam(q) Pl[:, leaf q]), read straight from
// the partial locals.  lam is a power of two, so the scaling is exact where it is applied.
// The layout of k_s2m_uc2e: 64 columns per workgroup, waves 8..15 copy chunks of 64 k of
// the columns into LDS (a lane per k, coalesced), waves 0..7 multiply the previous chunk
// on FP64 MFMA with the operator from its zero-padded copy (M <= 304 rows).
template <int MODE>
__global__ void __launch_bounds__(SU_TB)
k_gemm_cols(const double *__restrict__ Ap, int M, int K, const double *__restrict__ B, int ldb,
            const int32_t *__restrict__ cols, int64_t nq, const int32_t *__restrict__ level,
            const double *__restrict__ lam, double *__restrict__ out, float *__restrict__ out32, int ldo) {
  __shared__ __attribute__((aligned(16))) double P[2][SU_KC * SU_PS];
  const int t = threadIdx.x, lane = t & 63;
  const int w = __builtin_amdgcn_readfirstlane(t >> 6);
  const int64_t q0 = (int64_t)blockIdx.x * SU_NT;
  const int nch = (K + SU_KC - 1) / SU_KC;
  // the copy waves: wave 8 + pw takes the columns pw, pw + 8, ..; lane = k of the chunk
  auto fill = [&](int ch, double *Pb) {
    const int pw = w - 8, k = ch * SU_KC + lane;
#pragma unroll
    for (int j = 0; j < SU_NT / 8; j++) {
      const int c = pw + 8 * j;
      const int64_t q = q0 + c;
      double v = 0.0;
      if (q < nq && k < K) {
        if (MODE == 0) {
          v = B[q * ldb + k];
        } else {
          const int32_t n = cols[q];
          v = lam[level[n]] * B[(int64_t)n * ldb + k];
        }
      }
      Pb[lane * SU_PS + c] = v;
    }
  };
  if (w >= 8) {
    fill(0, P[0]);
    __syncthreads();
    for (int ch = 0; ch < nch; ch++) {
      if (ch + 1 < nch) fill(ch + 1, P[(ch + 1) & 1]);
      __syncthreads();   // chunk ch read by the MFMA waves, chunk ch + 1 written
    }
    return;
  }
  const int li = lane & 15, lk = lane >> 4;
  su_f64x4 acc[SU_RW][4];
#pragma unroll
  for (int i = 0; i < SU_RW; i++)
#pragma unroll
    for (int c = 0; c < 4; c++) acc[i][c] = su_f64x4{0.0, 0.0, 0.0, 0.0};
  auto loadA = [&](int g, double *dst) {
    const double *__restrict__ u = Ap + (int64_t)(4 * g + lk) * UC2E_PROWS + 16 * w + li;
#pragma unroll
    for (int i = 0; i < SU_RW; i++)
      if (i < 2 || w < 4) dst[i] = u[128 * i];
  };
  double ring[SU_PD][SU_RW];
#pragma unroll
  for (int g = 0; g < SU_PD; g++) loadA(g, ring[g]);
  __syncthreads();
  for (int ch = 0; ch < nch; ch++) {
    const double *Pb = P[ch & 1];
#pragma unroll
    for (int ks = 0; ks < SU_KC / 4; ks++) {
      double *a = ring[ks % SU_PD];
      double b[4];
#pragma unroll
      for (int c = 0; c < 4; c++) b[c] = Pb[(4 * ks + lk) * SU_PS + 16 * c + li];
#pragma unroll
      for (int i = 0; i < SU_RW; i++)
        if (i < 2 || w < 4)
#pragma unroll
          for (int c = 0; c < 4; c++) acc[i][c] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i], b[c], acc[i][c], 0, 0, 0);
      loadA(ch * (SU_KC / 4) + ks + SU_PD, a);
    }
    __syncthreads();   // chunk ch read, chunk ch + 1 written
  }
  // lane (lk, li) holds rows lk + 4 e of column li of each tile (the FP64 MFMA's layout)
#pragma unroll
  for (int i = 0; i < SU_RW; i++) {
    if (i == 2 && w >= 4) continue;
    const int rt = w + 8 * i;
#pragma unroll
    for (int c = 0; c < 4; c++) {
      const int64_t q = q0 + 16 * c + li;
      if (q >= nq) continue;
      int64_t o;
      double sc = 1.0;
      if (MODE == 0) {
        const int32_t n = cols[q];
        o = (int64_t)n * ldo;
        sc = lam[level[n]];
      } else {
        o = q * ldo;
      }
#pragma unroll
      for (int e = 0; e < 4; e++) {
        const int row = 16 * rt + lk + 4 * e;
        if (row < M) {
          const double v = sc * acc[i][c][e];
          out[o + row] = v;
          if (MODE == 0 && out32) out32[o + row] = (float)v;
        }
      }
    }
  }
}

// the cube method's S2M is fused with UC2E's first factor where the rank fits the kernel
static bool fused_s2m(const OpCache &o) { return o.r_up >= SU_RMIN && o.r_up <= SU_RMAX && o.Uh_up_pad.p; }

// dst[i, cols[q]] = alpha * src[i, q]
__global__ void k_scatter_cols(double *__restrict__ dst, const double *__restrict__ src, const int32_t *__restrict__ cols,
                               int rows, int64_t ncols, double alpha) {
  int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= (int64_t)rows * ncols) return;
  int64_t q = g / rows;
  int i = (int)(g % rows);
  dst[(int64_t)cols[q] * rows + i] = alpha * src[g];
}
// dst[i, q] = src[i, cols[q]]
__global__ void k_gather_cols(double *__restrict__ dst, const double *__restrict__ src, const int32_t *__restrict__ cols,
                              int rows, int64_t ncols) {
  int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= (int64_t)rows * ncols) return;
  int64_t q = g / rows;
  int i = (int)(g % rows);
  dst[g] = src[(int64_t)cols[q] * rows + i];
}
// dst[i, cols[q]] = lam[level[cols[q]]] * src[i, q]: the columns of several levels at once,
// each scaled by its level's width factor; dst32 (if given) takes the same value in FP32
__global__ void k_scatter_cols_lam(double *__restrict__ dst, const double *__restrict__ src,
                                   const int32_t *__restrict__ cols, int rows, int64_t ncols,
                                   const int32_t *__restrict__ level, const double *__restrict__ lam,
                                   float *__restrict__ dst32) {
  int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= (int64_t)rows * ncols) return;
  int64_t q = g / rows;
  int i = (int)(g % rows);
  const int32_t c = cols[q];
  const double v = lam[level[c]] * src[g];
  dst[(int64_t)c * rows + i] = v;
  if (dst32) dst32[(int64_t)c * rows + i] = (float)v;
}
// dst[i, q] = lam[level[cols[q]]] * src[i, cols[q]]
__global__ void k_gather_cols_lam(double *__restrict__ dst, const double *__restrict__ src,
                                  const int32_t *__restrict__ cols, int rows, int64_t ncols,
                                  const int32_t *__restrict__ level, const double *__restrict__ lam) {
  int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= (int64_t)rows * ncols) return;
  int64_t q = g / rows;
  int i = (int)(g % rows);
  const int32_t c = cols[q];
  dst[g] = lam[level[c]] * src[(int64_t)c * rows + i];
}
// the parents of the sibling groups starting at node `first`: dst[i, parent[first + 8g]] =
// alpha src[i, g]; dst32 (if given) takes the same value in FP32
template <class T>
__global__ void k_scatter_parents(double *__restrict__ dst, const T *__restrict__ src,
                                  const int32_t *__restrict__ parent, int64_t first, int rows, int64_t ng, double alpha,
                                  float *__restrict__ dst32) {
  int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= (int64_t)rows * ng) return;
  int64_t q = g / rows;
  int i = (int)(g % rows);
  const int64_t o = (int64_t)parent[first + NCHILD * q] * rows + i;
  const double v = alpha * (double)src[g];
  dst[o] = v;
  if (dst32) dst32[o] = (float)v;
}
__global__ void k_fill32(float *__restrict__ a, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = 0.f;
}
__global__ void k_gather_parents(double *__restrict__ dst, const double *__restrict__ src,
                                 const int32_t *__restrict__ parent, int64_t first, int rows, int64_t ng) {
  int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= (int64_t)rows * ng) return;
  int64_t q = g / rows;
  int i = (int)(g % rows);
  dst[g] = src[(int64_t)parent[first + NCHILD * q] * rows + i];
}
__global__ void k_fill(double *__restrict__ a, int64_t n, double v) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = v;
}

void fill_f64(Context &ctx, double *a, int64_t n, double v) {
  if (n <= 0) return;
  k_fill<<<nblocks(n, TB), TB, 0, ctx.stream>>>(a, n, v);
  FFMM_LAUNCH_CHECK();
}
// the densities into the tree's order: _gather_perm! (src/device_lists.jl:163-171)
__global__ void k_gather_perm(double *__restrict__ dst, const double *__restrict__ src, const int32_t *__restrict__ perm,
                              int64_t n) {
  int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j < n) dst[j] = src[perm[j]];
}
// nd components per point, point major
__global__ void k_gather_perm_nd(double *__restrict__ dst, const double *__restrict__ src,
                                 const int32_t *__restrict__ perm, int64_t n, int nd) {
  int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= n * nd) return;
  int64_t j = g / nd;
  int c = (int)(g % nd);
  dst[g] = src[(int64_t)perm[j] * nd + c];
}

// ---------------------------------------------------------------------------
// dense M2L by symmetry class: _class_gather! / _class_scatter! (src/device_passes.jl:426-473)
// An expansion column holds nb blocks of N entries (Ns = nb N); the class operator acts
// on each block (nb columns per pair, as the reference lays out the Stokes blocks).
// Gc[i, q nb + b] = sgn[f, i] Mp[b N + idx[f, i], src[q]]  (P_f^T m, block b)
// ---------------------------------------------------------------------------
// One target chunk of a level's class M2L: the chunk-local columns of class c are
// [co[c], co[c + 1]) and hold the class-major pairs cs[c] + (q - co[c]).
struct ChunkArgs {
  int nclass;
  int32_t co[MAX_CLASSES + 1];
  int32_t cs[MAX_CLASSES];
  int32_t to[MAX_CLASSES + 1];   // the fused kernel: first column tile of each class
};

// the class of chunk column q: the last c with co[c] <= q
__device__ __forceinline__ int chunk_class(const ChunkArgs &ca, int32_t q) {
  int c = 0, h = ca.nclass;
  while (h - c > 1) {
    const int m = (c + h) >> 1;
    if (ca.co[m] <= q) c = m;
    else h = m;
  }
  return c;
}

constexpr int GATHER_PAIRS = 8;   // pairs per workgroup of the gather
constexpr int SCATTER_MAXP = 320; // far pairs of one target: at most the 316 far offsets

// Gc[b N + i, q] = sgn[f, i] Mp[b N + idx[f, i], src]  (P_f^T m per block b) for the pairs
// of a chunk, all classes in one launch; one workgroup per GATHER_PAIRS columns, one lane
// per coefficient i with all nb blocks, the pairs' loads independent of each other.  For
// Stokes (nb = 4) the fourth block takes the coupling Psi += o_f . Phi of the translation
// (src/device_passes.jl:465-468, src/operators/harmonics.jl:604-628) before the product:
// the class operator is linear and the same for every block, so T (G_3 + o_f . (G_0, G_1,
// G_2)) is the coupled fourth block of the output.
template <class T>
__global__ void k_chunk_gather(T *__restrict__ Gc, const double *__restrict__ Mp, int Ns, int N, int nb,
                               const int32_t *__restrict__ csrc, const int32_t *__restrict__ cfar, int32_t npk,
                               const int32_t *__restrict__ gsel, const double *__restrict__ offs, ChunkArgs ca) {
  __shared__ int32_t pf[GATHER_PAIRS], ps[GATHER_PAIRS];
  const int nbN = nb * N;
  const int32_t q0 = (int32_t)blockIdx.x * GATHER_PAIRS;
  const int nq = min(GATHER_PAIRS, npk - q0);
  if ((int)threadIdx.x < nq) {
    const int32_t q = q0 + (int32_t)threadIdx.x;
    const int c = chunk_class(ca, q);
    const int32_t p = ca.cs[c] + (q - ca.co[c]);
    pf[threadIdx.x] = cfar[p];
    ps[threadIdx.x] = csrc[p];
  }
  __syncthreads();
  for (int i = threadIdx.x; i < N; i += blockDim.x) {
#pragma unroll
    for (int k = 0; k < GATHER_PAIRS; k++) {
      if (k < nq) {
        const int32_t f = pf[k];
        const int32_t g = gsel[f * N + i];
        const double *__restrict__ m = Mp + (int64_t)ps[k] * Ns + (g & 0x7fffffff);
        T *__restrict__ o = Gc + (int64_t)(q0 + k) * nbN + i;
        const double sg = g < 0 ? -1.0 : 1.0;
        if (nb == 4) {
          const double m0 = m[0], m1 = m[N], m2 = m[2 * N], m3 = m[3 * N];
          o[0] = (T)(sg * m0);
          o[N] = (T)(sg * m1);
          o[2 * N] = (T)(sg * m2);
          o[3 * N] = (T)(sg * (m3 + offs[3 * f] * m0 + offs[3 * f + 1] * m1 + offs[3 * f + 2] * m2));
        } else {
          for (int b = 0; b < nb; b++) o[b * N] = (T)(sg * m[b * N]);
        }
      }
    }
  }
}

typedef float fm_f32x4 __attribute__((ext_vector_type(4)));
constexpr int FM_TB = 512;      // 8 waves: wave w holds the operator's row tile w (rows 16 w .. 16 w + 15)
constexpr int FM_COLS = 128;    // product columns per tile: 8 column tiles of 16
constexpr int FM_MP = 128;      // operator rows and K, padded: N <= 128
constexpr int FM_BS = 132;      // LDS stride of a gathered column (16-byte aligned, K contiguous)

// The class M2L of the harmonic methods in the mixed form, fused: for each tile of a
// class's pairs in the chunk, the gathered source expansions (P_f^T m per block, signed,
// the Stokes coupling folded into the fourth block, src/device_passes.jl:449-473) go
// straight into LDS as FP32 and are multiplied by the class operator on FP32 MFMA
// (v_mfma_f32_16x16x4f32).  Each wave loads its 16 operator rows per tile from a
// row-major copy of the class operators padded to 128 x 128 (L2-resident), while the
// other waves finish the gather, so that no register is held across the tile's phases;
// the four K slots of an MFMA step take four consecutive k, so a lane's operands for
// 16 k come in one 16-byte load, of the operator and of the gathered column alike.  The
// products are written in each pair's target frame (entry j = fsel position, sign
// applied), so the scatter adds contiguous columns.  N = (p + 1)^2 <= 128.
template <int NB>
__global__ void __launch_bounds__(FM_TB) __attribute__((amdgpu_waves_per_eu(4))) k_class_m2l32(float *__restrict__ Yc, const double *__restrict__ Mp,
                                                          int Ns, int N, const float *__restrict__ reps32r,
                                                          const int32_t *__restrict__ csrc,
                                                          const int32_t *__restrict__ cfar,
                                                          const int32_t *__restrict__ gsel,
                                                          const int32_t *__restrict__ fsel,
                                                          const double *__restrict__ offs, ChunkArgs ca,
                                                          int32_t ntiles, const int32_t *__restrict__ ctgt) {
  constexpr int PPT = FM_COLS / NB;   // pairs per tile
  __shared__ float Bs[FM_COLS * FM_BS];
  __shared__ int32_t pf[PPT], ps[PPT], pt[PPT];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int li = lane & 15, lk = lane >> 4;
  const int nkb = (N + 15) >> 4;      // K blocks of 16
  const bool rows = 16 * w < N;       // this wave's row tile holds operator rows
  const int32_t t0 = (int32_t)((int64_t)blockIdx.x * ntiles / gridDim.x);
  const int32_t t1 = (int32_t)((int64_t)(blockIdx.x + 1) * ntiles / gridDim.x);
  // the K padding of the gathered columns: zero (the gather and the staging write k < N only)
  const int kpad = 16 * nkb - N;
  for (int e = tid; e < FM_COLS * kpad; e += FM_TB) Bs[(e / kpad) * FM_BS + N + e % kpad] = 0.f;
  for (int32_t tile = t0; tile < t1; tile++) {
    int c = 0, h = ca.nclass;
    while (h - c > 1) {
      const int m = (c + h) >> 1;
      if (ca.to[m] <= tile) c = m;
      else h = m;
    }
    const int32_t lp0 = (tile - ca.to[c]) * PPT;                       // first class-local pair
    const int npt = min(PPT, ca.co[c + 1] - ca.co[c] - lp0);           // pairs in the tile
    if (tid < npt) {
      const int32_t p = ca.cs[c] + lp0 + tid;
      pf[tid] = cfar[p];
      ps[tid] = csrc[p];
      pt[tid] = ctgt[p];
    }
    __syncthreads();
    // gather: a wave per pair, a lane per coefficient, all blocks
    for (int pl = w; pl < npt; pl += FM_TB / 64)
      for (int k = lane; k < N; k += 64) {
        const int32_t f = pf[pl];
        const int32_t g = gsel[f * N + k];
        const double *__restrict__ m = Mp + (int64_t)ps[pl] * Ns + (g & 0x7fffffff);
        const double sg = g < 0 ? -1.0 : 1.0;
        if (NB == 4) {
          const double m0 = m[0], m1 = m[N], m2 = m[2 * N], m3 = m[3 * N];
          float *__restrict__ b = Bs + (pl * 4) * FM_BS + k;
          b[0] = (float)(sg * m0);
          b[FM_BS] = (float)(sg * m1);
          b[2 * FM_BS] = (float)(sg * m2);
          b[3 * FM_BS] = (float)(sg * (m3 + offs[3 * f] * m0 + offs[3 * f + 1] * m1 + offs[3 * f + 2] * m2));
        } else {
          Bs[pl * FM_BS + k] = (float)(sg * m[0]);
        }
      }
    // operator row 16 w + li, k = 16 kb + 4 lk + q (q < 4): one 16-byte load per kb from
    // the padded row-major table (zero beyond N), in flight while the other waves gather
    fm_f32x4 a4[8];
    if (rows) {
      const float *__restrict__ Ar = reps32r + ((size_t)c * FM_MP + 16 * w + li) * FM_MP + 4 * lk;
#pragma unroll
      for (int kb = 0; kb < 8; kb++) a4[kb] = *reinterpret_cast<const fm_f32x4 *>(Ar + 16 * kb);
    }
    __syncthreads();
    // products: wave w, row tile w, all 8 column tiles
    fm_f32x4 acc[8];
#pragma unroll
    for (int ct = 0; ct < 8; ct++) acc[ct] = fm_f32x4{0.f, 0.f, 0.f, 0.f};
    if (rows) {
      const float *__restrict__ bb = Bs + li * FM_BS + 4 * lk;
#pragma unroll
      for (int kb = 0; kb < 8; kb++) {
        if (kb < nkb) {
#pragma unroll
          for (int ct = 0; ct < 8; ct++) {
            const fm_f32x4 b4 = *reinterpret_cast<const fm_f32x4 *>(bb + 16 * ct * FM_BS + 16 * kb);
#pragma unroll
            for (int q = 0; q < 4; q++) acc[ct] = __builtin_amdgcn_mfma_f32_16x16x4f32(a4[kb][q], b4[q], acc[ct], 0, 0, 0);
          }
        }
      }
    }
    __syncthreads();   // every wave's reads of Bs are done: Bs becomes the staging of the products
    if (rows) {
#pragma unroll
      for (int ct = 0; ct < 8; ct++) {
        const int col = 16 * ct + li, pl = col / NB;
        if (pl < npt) {
          const int32_t f = pf[pl];
#pragma unroll
          for (int r = 0; r < 4; r++) {
            const int i = 16 * w + 4 * lk + r;
            if (i < N) {
              const int32_t fs = fsel[f * N + i];
              Bs[col * FM_BS + (fs & 0x7fffffff)] = fs < 0 ? -acc[ct][r] : acc[ct][r];
            }
          }
        }
      }
    }
    __syncthreads();
    // coalesced copy of the tile's columns (pair column q = co[c] + lp0 + pl, NB N
    // entries), the products of consecutive pairs of one target summed first: the first
    // pair of each such run carries the run's sum, the others are not written
    float *__restrict__ out = Yc + ((int64_t)ca.co[c] + lp0) * NB * N;
    for (int col = w; col < npt * NB; col += FM_TB / 64) {
      const int pl = col / NB, b = col - pl * NB;
      if (pl > 0 && pt[pl] == pt[pl - 1]) continue;
      int pe = pl + 1;
      while (pe < npt && pt[pe] == pt[pl]) pe++;
      for (int j = lane; j < N; j += 64) {
        float acc = Bs[col * FM_BS + j];
        for (int r = pl + 1; r < pe; r++) acc += Bs[(r * NB + b) * FM_BS + j];
        out[(int64_t)col * N + j] = acc;
      }
    }
    __syncthreads();   // before the next tile's pair list and gather (the K padding is never written)
  }
}

// The scatter of products already in the target frame: pL[e, t] += alpha sum over the
// target's columns (class by class) of Y'[e, q]; the fused kernel has summed each run of
// the target's pairs within a tile of ppt pairs into the run's first column
__global__ void k_chunk_scatter_framed(double *__restrict__ pL, const float *__restrict__ Yc, int Ns, int nbN,
                                       const int32_t *__restrict__ cstart, int64_t nl, int32_t t0, double alpha,
                                       ChunkArgs ca, int ppt) {
  __shared__ int32_t pq[SCATTER_MAXP];
  __shared__ int32_t npair;
  const int32_t t = t0 + (int32_t)blockIdx.x;
  if (threadIdx.x < 64) {
    // lane c: the columns of class c that carry the target's sums: its first pair's,
    // and each tile start (class-local index a multiple of ppt) among its later pairs
    const int c = threadIdx.x;
    int32_t q0 = 0, f = 0, nh = 0;
    if (c < ca.nclass) {
      const int64_t ec = (int64_t)c * nl + t;
      const int32_t p0 = cstart[ec], n = cstart[ec + 1] - p0;
      if (n > 0) {
        q0 = p0 + ca.co[c] - ca.cs[c];
        f = q0 - ca.co[c];
        nh = 1 + (f + n - 1) / ppt - f / ppt;
      }
    }
    int32_t v = nh;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      const int32_t u = __shfl_up(v, off, 64);
      if (c >= off) v += u;
    }
    const int32_t start = v - nh;
    if (nh > 0) {
      pq[start] = q0;
      for (int32_t k = 1; k < nh; k++) pq[start + k] = ca.co[c] + (f / ppt + k) * ppt;
    }
    if (c == 63) npair = v;
  }
  __syncthreads();
  const int32_t np = npair;
  for (int e = threadIdx.x; e < nbN; e += blockDim.x) {
    double acc = 0.0;
#pragma unroll 8
    for (int32_t k = 0; k < np; k++) acc += (double)Yc[(int64_t)pq[k] * nbN + e];
    pL[(int64_t)t * Ns + e] += alpha * acc;
  }
}

// As k_chunk_scatter_framed for product columns whose length nbN is a multiple of 4:
// four consecutive entries per lane, 16-byte loads (same summation order per entry)
__global__ void k_chunk_scatter_framed4(double *__restrict__ pL, const float4 *__restrict__ Yc, int Ns, int nbN4,
                                        const int32_t *__restrict__ cstart, int64_t nl, int32_t t0, double alpha,
                                        ChunkArgs ca, int ppt) {
  __shared__ int32_t pq[SCATTER_MAXP];
  __shared__ int32_t npair;
  const int32_t t = t0 + (int32_t)blockIdx.x;
  if (threadIdx.x < 64) {
    // lane c: the columns of class c that carry the target's sums: its first pair's,
    // and each tile start (class-local index a multiple of ppt) among its later pairs
    const int c = threadIdx.x;
    int32_t q0 = 0, f = 0, nh = 0;
    if (c < ca.nclass) {
      const int64_t ec = (int64_t)c * nl + t;
      const int32_t p0 = cstart[ec], n = cstart[ec + 1] - p0;
      if (n > 0) {
        q0 = p0 + ca.co[c] - ca.cs[c];
        f = q0 - ca.co[c];
        nh = 1 + (f + n - 1) / ppt - f / ppt;
      }
    }
    int32_t v = nh;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      const int32_t u = __shfl_up(v, off, 64);
      if (c >= off) v += u;
    }
    const int32_t start = v - nh;
    if (nh > 0) {
      pq[start] = q0;
      for (int32_t k = 1; k < nh; k++) pq[start + k] = ca.co[c] + (f / ppt + k) * ppt;
    }
    if (c == 63) npair = v;
  }
  __syncthreads();
  const int32_t np = npair;
  for (int e = threadIdx.x; e < nbN4; e += blockDim.x) {
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
#pragma unroll 4
    for (int32_t k = 0; k < np; k++) {
      const float4 y = Yc[(int64_t)pq[k] * nbN4 + e];
      a0 += (double)y.x;
      a1 += (double)y.y;
      a2 += (double)y.z;
      a3 += (double)y.w;
    }
    double2 *o = reinterpret_cast<double2 *>(pL + (int64_t)t * Ns + 4 * e);
    double2 u = o[0], w = o[1];
    u.x += alpha * a0;
    u.y += alpha * a1;
    w.x += alpha * a2;
    w.y += alpha * a3;
    o[0] = u;
    o[1] = w;
  }
}

// Target-owned scatter of the chunk's products, all classes at once: one workgroup per
// target t, one lane per entry (b, j); pL[b N + j, t] += alpha sum over the target's pairs
// (class by class, in pair order) of sgn_t[f, i] Y_b[i, q], i = inv[f, j]
// (src/device_passes.jl:449-473; the Stokes coupling is already in the gathered fourth
// block).  The target's pairs
// (offset, chunk column) are staged in LDS first, so that each lane's loop over them
// issues independent loads.
template <class T>
__global__ void k_chunk_scatter(double *__restrict__ pL, const T *__restrict__ Yc, int Ns, int N, int nb,
                                const int32_t *__restrict__ cstart, int64_t nl, int32_t t0,
                                const int32_t *__restrict__ cfar, const int32_t *__restrict__ ssel, double alpha,
                                ChunkArgs ca) {
  __shared__ int32_t pf[SCATTER_MAXP], pq[SCATTER_MAXP];
  __shared__ int32_t npair;
  const int32_t t = t0 + (int32_t)blockIdx.x;
  const int nbN = nb * N;
  if (threadIdx.x < 64) {
    // lane c: class c's pairs of the target; an exclusive scan over the classes places them
    const int c = threadIdx.x;
    int32_t p0 = 0, n = 0;
    if (c < ca.nclass) {
      const int64_t ec = (int64_t)c * nl + t;
      p0 = cstart[ec];
      n = cstart[ec + 1] - p0;
    }
    int32_t v = n;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      const int32_t u = __shfl_up(v, off, 64);
      if (c >= off) v += u;
    }
    const int32_t start = v - n;
    const int32_t qoff = c < ca.nclass ? ca.co[c] - ca.cs[c] : 0;
    for (int32_t k = 0; k < n; k++) {
      pf[start + k] = cfar[p0 + k];
      pq[start + k] = p0 + k + qoff;
    }
    if (c == 63) npair = v;
  }
  __syncthreads();
  const int32_t np = npair;
  for (int e = threadIdx.x; e < nbN; e += blockDim.x) {
    const int b = e / N, j = e - b * N;
    double acc = 0.0;
#pragma unroll 4
    for (int32_t k = 0; k < np; k++) {
      const int32_t f = pf[k];
      const int32_t sv = ssel[f * N + j];
      const int i = sv & 0x7fffffff;
      const double y = (double)Yc[(int64_t)pq[k] * nbN + b * N + i];
      acc += sv < 0 ? -y : y;
    }
    pL[(int64_t)t * Ns + e] += alpha * acc;
  }
}

static void gemm(Context &ctx, int m, int n, int k, double alpha, const double *A, int lda, const double *B, int ldb,
                 double beta, double *C, int ldc) {
  if (m == 0 || n == 0) return;
  FFMM_BLAS(rocblas_dgemm(ctx.blas, rocblas_operation_none, rocblas_operation_none, m, n, k, &alpha, A, lda, B, ldb,
                          &beta, C, ldc));
}

constexpr int64_t M2M32_CHUNK = 10240;   // columns per call of the mixed form's M2M GEMM

static void gemm32(Context &ctx, int m, int n, int k, const float *A, int lda, const float *B, int ldb, float *C,
                   int ldc) {
  if (m == 0 || n == 0) return;
  const float one = 1.0f, zero = 0.0f;
  FFMM_BLAS(rocblas_sgemm(ctx.blas, rocblas_operation_none, rocblas_operation_none, m, n, k, &one, A, lda, B, ldb, &zero,
                          C, ldc));
}

// the harmonic methods in the mixed form with at most 128 coefficients per block take
// the fused gather + product kernel
static bool fused_m2l(const OpCache &o) {
  return o.method == METHOD_HARMONIC && o.hN <= FM_MP && (o.nb == 1 || o.nb == 4);
}

void allocate_pass_buffers(Plan &plan) {
  OpCache &o = *plan.ops;
  const int Ns = o.Ns, Nl = o.Nl;
  int64_t ng = 1, nlmax = 1;
  for (int l = 0; l < plan.L; l++) {
    ng = std::max(ng, plan.nl[l] / NCHILD);
    nlmax = std::max(nlmax, plan.nl[l]);
  }
  int64_t nleaves = plan.leaf_off[plan.L];
  // UC2E and DC2E run once over the leaves of all levels from 2 down
  const int l2 = std::min(2, plan.L);
  const int64_t nq = std::max<int64_t>({1, plan.s2m_off[plan.L] - plan.s2m_off[l2], nleaves - plan.leaf_off[l2]});
  plan.Mp.resize((size_t)Ns * plan.nnodes);
  plan.Pl.resize((size_t)Ns * plan.nnodes);
  plan.Lc.resize((size_t)Nl * std::max<int64_t>(nleaves, 1));
  // the densities, at the widest the method takes: Stokes 9, SphericalHarmonics(Laplace3D())
  // 4 (a charge and a dipole), else 1
  const int ndmax = o.method != METHOD_HARMONIC ? 1 : o.nb == 4 ? 9 : o.kernel == FFMM_KERNEL_LAPLACE ? 4 : 1;
  plan.tau.resize((size_t)plan.N * ndmax);
  if (o.method == METHOD_CUBE) {
    if (!fused_s2m(o)) plan.pM.resize((size_t)Nl * nq);   // else S2M is fused with UC2E's first factor
    plan.tup.resize((size_t)std::max(o.r_up, o.r_dn) * nq);
    if (!o.Vs_up_pad.p) plan.Mup.resize((size_t)Ns * nq);   // else UC2E's second factor writes Mp itself
  }
  plan.tM.resize((size_t)Ns * ng);
  if (o.method == METHOD_CUBE && plan.precision == FFMM_PREC_MIXED) {
    plan.Mp32.resize((size_t)Ns * plan.nnodes);
    plan.tM32.resize((size_t)Ns * ng);
  }
  plan.G.resize((size_t)Ns * std::max(nq, ng));
  int64_t pmax = 1;
  for (auto &li : plan.lists) pmax = std::max(pmax, li.chunk_max);
  if (plan.m2l == FFMM_M2L_DENSE && plan.precision == FFMM_PREC_MIXED) {
    // the mixed form: the class GEMMs in FP32 (FP32 MFMA), gathered from and scattered
    // into the FP64 expansions (the fused form gathers straight into LDS)
    if (!fused_m2l(o)) plan.Gc32.resize((size_t)Ns * pmax);
    plan.Yc32.resize((size_t)Ns * pmax);
  } else if (plan.m2l == FFMM_M2L_DENSE) {
    plan.Gc.resize((size_t)Ns * pmax);
    plan.Yc.resize((size_t)Ns * pmax);
  }
  if (plan.lowrank && plan.m2l == FFMM_M2L_DENSE) {   // Bt_c G of the pairs of one class of a chunk
    if (plan.precision == FFMM_PREC_MIXED)
      plan.Zc32.resize((size_t)plan.lowrank->kmax * pmax);
    else
      plan.Zc.resize((size_t)plan.lowrank->kmax * pmax);
  }
}

void gather_inputs(Plan &plan, const double *tau, int sorted) {
  Context &ctx = *plan.ctx;
  const int nd = plan.nd;
  if (sorted)
    FFMM_HIP(hipMemcpyAsync(plan.tau.p, tau, sizeof(double) * plan.N * nd, hipMemcpyDeviceToDevice, ctx.stream));
  else if (nd == 1)
    k_gather_perm<<<nblocks(plan.N, TB), TB, 0, ctx.stream>>>(plan.tau.p, tau, plan.perm.p, plan.N);
  else
    k_gather_perm_nd<<<nblocks(plan.N * nd, TB), TB, 0, ctx.stream>>>(plan.tau.p, tau, plan.perm.p, plan.N, nd);
  FFMM_LAUNCH_CHECK();
}

// upward pass: src/device_passes.jl:523-547
void upward_pass(Plan &plan) {
  Context &ctx = *plan.ctx;
  OpCache &o = *plan.ops;
  hipStream_t s = ctx.stream;
  const int Ns = o.Ns, Nl = o.Nl;
  // every column is written before it is read except by M2M of empty groups: clear.  The
  // cube's mixed form keeps an FP32 mirror of the multipoles, written beside them by the
  // UC2E and M2M scatters, that its FP32 M2M GEMMs read directly
  const bool mirror = o.method == METHOD_CUBE && plan.precision == FFMM_PREC_MIXED;
  float *Mp32 = mirror ? plan.Mp32.p : nullptr;
  k_fill<<<nblocks(plan.Mp.n, TB), TB, 0, s>>>(plan.Mp.p, plan.Mp.n, 0.0);
  FFMM_LAUNCH_CHECK();
  if (mirror) {
    k_fill32<<<nblocks(plan.Mp32.n, TB), TB, 0, s>>>(Mp32, plan.Mp32.n);
    FFMM_LAUNCH_CHECK();
  }
  const int l2 = std::min(2, plan.L);
  const int64_t nq = plan.s2m_off[plan.L] - plan.s2m_off[l2];
  if (o.method == METHOD_CUBE && nq > 0) {
    // the childless sourced nodes of all levels at once (each node's width from its level):
    // S2M into check potentials, then UC2E, M = pinv_scale V diag(invS) (Uh pM)
    // (src/device_passes.jl:496-502), as two GEMMs over all of them.  pinv_scale = lam[l]
    // is a power of two, applied exactly in the scatter.
    const int32_t *nodes = plan.s2m_nodes.p + plan.s2m_off[l2];
    const bool mixed = plan.precision == FFMM_PREC_MIXED;
    if (fused_s2m(o)) {
      // S2M and the first factor in one kernel: no check potentials in memory
      const unsigned g = (unsigned)((nq + SU_NT - 1) / SU_NT);
#define FFMM_SU_ARGS                                                                                          \
  nodes, nq, plan.centers.p, plan.nnodes, plan.level.p, plan.lam_d.p, o.large.p, Nl, plan.plo.p, plan.phi.p, \
      plan.sx.p, plan.sy.p, plan.sz.p, plan.tau.p, o.Uh_up_pad.p, o.r_up, plan.tup.p
      if (mixed) k_s2m_uc2e<true><<<g, SU_TB, 0, s>>>(FFMM_SU_ARGS);
      else k_s2m_uc2e<false><<<g, SU_TB, 0, s>>>(FFMM_SU_ARGS);
#undef FFMM_SU_ARGS
      FFMM_LAUNCH_CHECK();
    } else {
      if (mixed)
        k_s2m32<<<(unsigned)nq, S2M_TB, 0, s>>>(nodes, plan.centers.p, plan.nnodes, plan.level.p, plan.lam_d.p,
                                                o.large.p, Nl, plan.plo.p, plan.phi.p, plan.sx.p, plan.sy.p,
                                                plan.sz.p, plan.tau.p, plan.pM.p);
      else
        k_s2m<<<(unsigned)nq, S2M_TB, 0, s>>>(nodes, plan.centers.p, plan.nnodes, plan.level.p, plan.lam_d.p,
                                              o.large.p, Nl, plan.plo.p, plan.phi.p, plan.sx.p, plan.sy.p, plan.sz.p,
                                              plan.tau.p, plan.pM.p);
      FFMM_LAUNCH_CHECK();
      gemm(ctx, o.r_up, (int)nq, Nl, 1.0, o.Uh_up.p, o.r_up, plan.pM.p, Nl, 0.0, plan.tup.p, o.r_up);
    }
    if (o.Vs_up_pad.p) {
      // the second factor scattered into the multipoles by the GEMM itself
      k_gemm_cols<0><<<(unsigned)((nq + SU_NT - 1) / SU_NT), SU_TB, 0, s>>>(
          o.Vs_up_pad.p, Ns, o.r_up, plan.tup.p, o.r_up, nodes, nq, plan.level.p, plan.lam_d.p, plan.Mp.p, Mp32, Ns);
    } else {
      gemm(ctx, Ns, (int)nq, o.r_up, 1.0, o.Vs_up.p, Ns, plan.tup.p, o.r_up, 0.0, plan.Mup.p, Ns);
      k_scatter_cols_lam<<<nblocks((int64_t)Ns * nq, TB), TB, 0, s>>>(plan.Mp.p, plan.Mup.p, nodes, Ns, nq,
                                                                      plan.level.p, plan.lam_d.p, Mp32);
    }
    FFMM_LAUNCH_CHECK();
  }
  for (int l = plan.L - 1; l >= 2; l--) {
    if (o.method == METHOD_HARMONIC) {
      // the multipoles directly (partial_to_multipole! is a copy, src/operators/harmonics.jl:441)
      harmonic_s2m_level(plan, l);
    }
    if (l < plan.L - 1) {
      const int64_t ng = plan.nl[l + 1] / NCHILD;
      const double *B = plan.Mp.p + plan.lo[l + 1] * Ns;
      const double *M2M = o.method == METHOD_HARMONIC ? plan.hl[l]->M2M.p : o.M2M.p;
      if (mirror) {
        // the mixed form: M2M as an FP32 GEMM over the children's FP32 mirror (its rounding
        // is not amplified: the multipoles feed M2L, never a pseudo-inverse).  Wide levels
        // go in column chunks: for n above ~11k the library takes 256 x 256 tiles, at most
        // 192 workgroups for m = 296, each a 0.5 ms tile on one CU (levels of 14k-25k
        // groups all took ~0.55 ms), where a chunk takes a one-wave grid of smaller tiles
        const float *Bc = Mp32 + plan.lo[l + 1] * Ns;
        for (int64_t c0 = 0; c0 < ng; c0 += M2M32_CHUNK) {
          const int nc = (int)std::min<int64_t>(M2M32_CHUNK, ng - c0);
          gemm32(ctx, Ns, nc, NCHILD * Ns, o.M2M32.p, Ns, Bc + c0 * NCHILD * Ns, NCHILD * Ns, plan.tM32.p + c0 * Ns,
                 Ns);
        }
        k_scatter_parents<float><<<nblocks((int64_t)Ns * ng, TB), TB, 0, s>>>(plan.Mp.p, plan.tM32.p, plan.parent.p,
                                                                              plan.lo[l + 1], Ns, ng, 1.0, Mp32);
      } else {
        gemm(ctx, Ns, (int)ng, NCHILD * Ns, 1.0, M2M, Ns, B, NCHILD * Ns, 0.0, plan.tM.p, Ns);
        k_scatter_parents<double><<<nblocks((int64_t)Ns * ng, TB), TB, 0, s>>>(plan.Mp.p, plan.tM.p, plan.parent.p,
                                                                               plan.lo[l + 1], Ns, ng, 1.0, nullptr);
      }
      FFMM_LAUNCH_CHECK();
    }
  }
}

void m2l_dense_level(Plan &plan, int l) {
  Context &ctx = *plan.ctx;
  OpCache &o = *plan.ops;
  hipStream_t s = ctx.stream;
  const int Ns = o.Ns;
  const bool harm = o.method == METHOD_HARMONIC;
  const int N = harm ? o.hN : Ns, nb = harm ? o.nb : 1;
  LevelLists &li = plan.lists[l];
  double *pL = plan.Pl.p + plan.lo[l] * Ns;
  // m2l_scale of the level: (w_l / w_ref)^(-1) for the homogeneous kernels, 1 for Yukawa
  const double alpha = harm ? plan.alpha[l] : 1.0 / plan.lam[l];
  const double *reps = harm ? plan.hl[l]->reps.p : o.class_ops.p;
  const float *reps32 = harm ? plan.hl[l]->reps32.p : o.class_ops32.p;
  const bool f32 = plan.precision == FFMM_PREC_MIXED;
  const int nbN = N * nb;
  const unsigned bt = (unsigned)std::min(1024, (nbN + 63) / 64 * 64);   // scatter: a lane per entry
  const unsigned btg = (unsigned)std::min(1024, (N + 63) / 64 * 64);    // gather: a lane per coefficient
  const int64_t nch = (int64_t)li.chunk_t.size() - 1;
  const bool fused = f32 && plan.m2l == FFMM_M2L_DENSE && fused_m2l(o);
  for (int64_t k = 0; k < nch; k++) {
    ChunkArgs ca{};
    ca.nclass = o.nclass;
    for (int c = 0; c < o.nclass; c++) {
      ca.cs[c] = li.chunk_cs[k * o.nclass + c];
      ca.co[c + 1] = ca.co[c] + (li.chunk_cs[(k + 1) * o.nclass + c] - ca.cs[c]);
    }
    const int32_t npk = ca.co[o.nclass];
    const int32_t t0 = li.chunk_t[k], nt = li.chunk_t[k + 1] - t0;
    if (npk == 0 || nt == 0) continue;
    if (fused) {
      const float *reps32r = plan.hl[l]->reps32r.p;
      const int ppt = FM_COLS / nb;
      for (int c = 0; c < o.nclass; c++) ca.to[c + 1] = ca.to[c] + (ca.co[c + 1] - ca.co[c] + ppt - 1) / ppt;
      const int32_t ntiles = ca.to[o.nclass];
      const unsigned grid = (unsigned)std::min<int64_t>(ntiles, 4 * 256);
      if (nb == 4)
        k_class_m2l32<4><<<grid, FM_TB, 0, s>>>(plan.Yc32.p, plan.Mp.p, Ns, N, reps32r, li.csrc.p, li.cfar.p, o.gsel.p,
                                                o.fsel.p, o.offs.p, ca, ntiles, li.ctgt.p);
      else
        k_class_m2l32<1><<<grid, FM_TB, 0, s>>>(plan.Yc32.p, plan.Mp.p, Ns, N, reps32r, li.csrc.p, li.cfar.p, o.gsel.p,
                                                o.fsel.p, o.offs.p, ca, ntiles, li.ctgt.p);
      FFMM_LAUNCH_CHECK();
      if (nbN % 4 == 0) {
        const unsigned bt4 = (unsigned)std::min(1024, (nbN / 4 + 63) / 64 * 64);
        k_chunk_scatter_framed4<<<(unsigned)nt, bt4, 0, s>>>(pL, reinterpret_cast<const float4 *>(plan.Yc32.p), Ns,
                                                              nbN / 4, li.cstart.p, plan.nl[l], t0, alpha, ca, ppt);
      } else {
        k_chunk_scatter_framed<<<(unsigned)nt, bt, 0, s>>>(pL, plan.Yc32.p, Ns, nbN, li.cstart.p, plan.nl[l], t0,
                                                           alpha, ca, ppt);
      }
      FFMM_LAUNCH_CHECK();
      continue;
    }
    const unsigned gg = (unsigned)((npk + GATHER_PAIRS - 1) / GATHER_PAIRS);
    if (f32) {
      k_chunk_gather<float><<<gg, btg, 0, s>>>(plan.Gc32.p, plan.Mp.p, Ns, N, nb, li.csrc.p, li.cfar.p, npk, o.gsel.p,
                                              o.offs.p, ca);
      FFMM_LAUNCH_CHECK();
      for (int c = 0; c < o.nclass; c++) {
        const int32_t n = ca.co[c + 1] - ca.co[c];
        if (n == 0) continue;
        const size_t off = (size_t)ca.co[c] * nbN;
        if (const ClassFactors *lr = plan.lowrank) {   // A_c (Bt_c G): src/device_passes.jl:570-576
          const int kc = lr->rank[c];
          gemm32(ctx, kc, nb * n, N, lr->Bt32.p + lr->off[c], kc, plan.Gc32.p + off, N, plan.Zc32.p, kc);
          gemm32(ctx, N, nb * n, kc, lr->A32.p + lr->off[c], N, plan.Zc32.p, kc, plan.Yc32.p + off, N);
        } else {
          gemm32(ctx, N, nb * n, N, reps32 + (size_t)c * N * N, N, plan.Gc32.p + off, N, plan.Yc32.p + off, N);
        }
      }
      k_chunk_scatter<float><<<(unsigned)nt, bt, 0, s>>>(pL, plan.Yc32.p, Ns, N, nb, li.cstart.p, plan.nl[l], t0,
                                                         li.cfar.p, o.ssel.p, alpha, ca);
    } else {
      k_chunk_gather<double><<<gg, btg, 0, s>>>(plan.Gc.p, plan.Mp.p, Ns, N, nb, li.csrc.p, li.cfar.p, npk, o.gsel.p,
                                               o.offs.p, ca);
      FFMM_LAUNCH_CHECK();
      for (int c = 0; c < o.nclass; c++) {
        const int32_t n = ca.co[c + 1] - ca.co[c];
        if (n == 0) continue;
        const size_t off = (size_t)ca.co[c] * nbN;
        if (const ClassFactors *lr = plan.lowrank) {   // A_c (Bt_c G): src/device_passes.jl:570-576
          const int kc = lr->rank[c];
          gemm(ctx, kc, nb * n, N, 1.0, lr->Bt.p + lr->off[c], kc, plan.Gc.p + off, N, 0.0, plan.Zc.p, kc);
          gemm(ctx, N, nb * n, kc, 1.0, lr->A.p + lr->off[c], N, plan.Zc.p, kc, 0.0, plan.Yc.p + off, N);
        } else {
          gemm(ctx, N, nb * n, N, 1.0, reps + (size_t)c * N * N, N, plan.Gc.p + off, N, 0.0, plan.Yc.p + off, N);
        }
      }
      k_chunk_scatter<double><<<(unsigned)nt, bt, 0, s>>>(pL, plan.Yc.p, Ns, N, nb, li.cstart.p, plan.nl[l], t0,
                                                          li.cfar.p, o.ssel.p, alpha, ca);
    }
    FFMM_LAUNCH_CHECK();
  }
}

// downward pass: src/device_passes.jl:548-595
void downward_pass(Plan &plan) {
  Context &ctx = *plan.ctx;
  OpCache &o = *plan.ops;
  hipStream_t s = ctx.stream;
  const int Ns = o.Ns, Nl = o.Nl;
  // levels 0 and 1 have no far field: their partial locals and leaf densities are zero
  const int l2 = std::min(2, plan.L);
  const int64_t n01 = l2 < plan.L ? plan.lo[l2] : plan.nnodes;
  k_fill<<<nblocks((int64_t)Ns * n01, TB), TB, 0, s>>>(plan.Pl.p, (int64_t)Ns * n01, 0.0);
  k_fill<<<nblocks((int64_t)Nl * plan.leaf_off[l2], TB), TB, 0, s>>>(plan.Lc.p, (int64_t)Nl * plan.leaf_off[l2], 0.0);
  FFMM_LAUNCH_CHECK();
  for (int l = 2; l < plan.L; l++) {
    const int64_t nl = plan.nl[l], ng = nl / NCHILD;
    double *pL = plan.Pl.p + plan.lo[l] * Ns;
    if (l > 2) {
      k_gather_parents<<<nblocks((int64_t)Ns * ng, TB), TB, 0, s>>>(plan.G.p, plan.Pl.p, plan.parent.p, plan.lo[l],
                                                                    Ns, ng);
      FFMM_LAUNCH_CHECK();
      // the stacked children output of L2L is the level's block of partial locals
      const double *L2L = o.method == METHOD_HARMONIC ? plan.hl[l - 1]->L2L.p : o.L2L.p;
      gemm(ctx, NCHILD * Ns, (int)ng, Ns, 1.0, L2L, NCHILD * Ns, plan.G.p, Ns, 0.0, pL, NCHILD * Ns);
    } else {
      k_fill<<<nblocks((int64_t)Ns * nl, TB), TB, 0, s>>>(pL, (int64_t)Ns * nl, 0.0);
      FFMM_LAUNCH_CHECK();
    }
    StageTimer tm(plan, 4, l);
    if (plan.m2l == FFMM_M2L_SPECTRAL) m2l_spectral_level(plan, l);
    else m2l_dense_level(plan, l);
  }
  // the densities of the real leaves of all levels at once: DC2E (src/device_passes.jl:582-593)
  const int64_t nq = plan.leaf_off[plan.L] - plan.leaf_off[l2];
  if (nq == 0) return;
  const int32_t *leaves = plan.leaf_nodes.p + plan.leaf_off[l2];
  double *Lc = plan.Lc.p + plan.leaf_off[l2] * Nl;
  if (o.method == METHOD_HARMONIC) {
    // partial_to_local! is a copy (src/operators/harmonics.jl:442)
    k_gather_cols<<<nblocks((int64_t)Ns * nq, TB), TB, 0, s>>>(Lc, plan.Pl.p, leaves, Ns, nq);
    FFMM_LAUNCH_CHECK();
    return;
  }
  // the level's pinv_scale (a power of two) on the gathered check potentials: exact
  if (o.Uh_dn_pad.p) {
    // the first factor reading the leaves' partial locals itself
    k_gemm_cols<1><<<(unsigned)((nq + SU_NT - 1) / SU_NT), SU_TB, 0, s>>>(
        o.Uh_dn_pad.p, o.r_dn, Ns, plan.Pl.p, Ns, leaves, nq, plan.level.p, plan.lam_d.p, plan.tup.p, nullptr, o.r_dn);
  } else {
    k_gather_cols_lam<<<nblocks((int64_t)Ns * nq, TB), TB, 0, s>>>(plan.G.p, plan.Pl.p, leaves, Ns, nq,
                                                                   plan.level.p, plan.lam_d.p);
    FFMM_LAUNCH_CHECK();
    gemm(ctx, o.r_dn, (int)nq, Ns, 1.0, o.Uh_dn.p, o.r_dn, plan.G.p, Ns, 0.0, plan.tup.p, o.r_dn);
  }
  FFMM_LAUNCH_CHECK();
  gemm(ctx, Nl, (int)nq, o.r_dn, 1.0, o.Vs_dn.p, Nl, plan.tup.p, o.r_dn, 0.0, Lc, Nl);
}

}  // namespace ffmm
