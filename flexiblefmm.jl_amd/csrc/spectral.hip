// The spectral M2L of the cube method on the device.
//
// The small surfaces of all boxes are p-per-edge lattices of one spacing, so M2L at a
// far offset is a convolution of the multipole (embedded in the (2p)^3 lattice) with
// a kernel table: u_t = IFFT( sum_pairs T_f * FFT(m_s) ).  Per level:
//   1. forward: every sourced box's multipole into its half spectrum (p+1)(2p)(2p);
//      the multipole lives in the [0, p)^3 corner of the padded lattice, so the three
//      passes are small DFTs over p nonzero inputs (one workgroup per box, in LDS);
//   2. Hadamard: per sibling group (8 targets sharing a parent) and frequency block,
//      the far pairs are the sourced children of the parent's 27 colleagues; each
//      source spectrum is read once for the 8 targets, the 316 tables of the block sit
//      in LDS, accumulators in registers, no atomics.  FP64: one lane per (frequency,
//      candidate) on the vector ALUs; FP32: 16 groups at a time on the matrix cores;
//   3. inverse: per target, the three inverse DFTs with only the p^3 corner of the
//      output formed, the surface entries scaled and added into the partial locals.
// Follows _spectral_m2l_level! and _hadamard_block! (src/operators/spectral_m2l.jl:
// 260-286, 328-451); the pair enumeration is that of the interaction lists
// (src/tree/lists.jl:96-124).  p = 8 (n = 16) is fixed here.
#include "ffmm_internal.h"

#include <type_traits>

#include <cmath>

namespace ffmm {

constexpr int SP_P = 8;                       // points per edge
constexpr int SP_N = 2 * SP_P;                // padded lattice
constexpr int SP_H = SP_P + 1;                // half spectrum along the first axis
constexpr int SP_NF = SP_H * SP_N * SP_N;     // 2304 frequencies

template <class T> struct cplx_t;
template <> struct cplx_t<double> { using type = double2; };
template <> struct cplx_t<float> { using type = float2; };

template <class C, class T>
__device__ __forceinline__ void cmac(C &acc, C a, T br, T bi) {   // acc += a * (br + i bi)
  acc.x = fma(a.x, br, fma(-a.y, bi, acc.x));
  acc.y = fma(a.x, bi, fma(a.y, br, acc.y));
}

// The 16th roots of unity, cos(2 pi j / 16) and sin(2 pi j / 16): every twiddle of the
// transforms below is one of them, at an index known at compile time once the loops are
// unrolled, so it is an immediate and no table is read.
__device__ constexpr double W16C[SP_N] = {1.0, 0.92387953251128674, 0.70710678118654752, 0.38268343236508977,
                                          0.0, -0.38268343236508977, -0.70710678118654752, -0.92387953251128674,
                                          -1.0, -0.92387953251128674, -0.70710678118654752, -0.38268343236508977,
                                          0.0, 0.38268343236508977, 0.70710678118654752, 0.92387953251128674};
__device__ constexpr double W16S[SP_N] = {0.0, 0.38268343236508977, 0.70710678118654752, 0.92387953251128674,
                                          1.0, 0.92387953251128674, 0.70710678118654752, 0.38268343236508977,
                                          0.0, -0.38268343236508977, -0.70710678118654752, -0.92387953251128674,
                                          -1.0, -0.92387953251128674, -0.70710678118654752, -0.38268343236508977};

// ---------------------------------------------------------------------------
// 1. forward: S[node][k3][k2][k1] = sum_{i in [0,p)^3} x[i3][i2][i1] w^(i1 k1 + i2 k2 + i3 k3),
//    w = exp(-2 pi i / 16).  One workgroup per box; each pass is one short DFT per thread
//    with its 8 inputs in registers: (i3, i2) rows over i1 (real input, k1 <= p), then
//    (i3, k1) over i2, then (k2, k1) over i3, whose 16 outputs per thread go straight to
//    the spectrum (consecutive threads, consecutive (k2, k1): coalesced).
// ---------------------------------------------------------------------------
constexpr int FW_TB = 192;   // 3 waves: 64, 72 and 144 threads work in the three passes

template <class T>
__global__ void __launch_bounds__(FW_TB) k_spec_forward(const double *__restrict__ Mp, int Ns,
                                                        const int16_t *__restrict__ lat8, int64_t lo, int64_t nl,
                                                        const uint8_t *__restrict__ has_source,
                                                        typename cplx_t<T>::type *__restrict__ S) {
  using C = typename cplx_t<T>::type;
  __shared__ C A[SP_P * SP_P * SP_H];   // [i3][i2][k1]
  __shared__ C B[SP_P * SP_N * SP_H];   // [i3][k2][k1]
  const int64_t q = blockIdx.x;
  const int64_t node = lo + q;
  if (!has_source[node]) return;
  const int t = threadIdx.x;
  const double *__restrict__ m = Mp + node * Ns;
  if (t < SP_P * SP_P) {   // row (i3, i2) = t: the p lattice values (zero off the surface)
    T x[SP_P];
#pragma unroll
    for (int i1 = 0; i1 < SP_P; i1++) {
      const int sidx = lat8[t * SP_P + i1];
      x[i1] = sidx >= 0 ? (T)m[sidx] : (T)0;
    }
#pragma unroll
    for (int k1 = 0; k1 < SP_H; k1++) {
      C acc = {(T)0, (T)0};
#pragma unroll
      for (int i1 = 0; i1 < SP_P; i1++) {
        const int w = (i1 * k1) & (SP_N - 1);
        acc.x = fma(x[i1], (T)W16C[w], acc.x);
        acc.y = fma(x[i1], -(T)W16S[w], acc.y);
      }
      A[t * SP_H + k1] = acc;
    }
  }
  __syncthreads();
  if (t < SP_P * SP_H) {   // (i3, k1)
    const int i3 = t / SP_H, k1 = t % SP_H;
    C a[SP_P];
#pragma unroll
    for (int i2 = 0; i2 < SP_P; i2++) a[i2] = A[(i3 * SP_P + i2) * SP_H + k1];
#pragma unroll
    for (int k2 = 0; k2 < SP_N; k2++) {
      C acc = {(T)0, (T)0};
#pragma unroll
      for (int i2 = 0; i2 < SP_P; i2++) {
        const int w = (i2 * k2) & (SP_N - 1);
        cmac(acc, a[i2], (T)W16C[w], -(T)W16S[w]);
      }
      B[(i3 * SP_N + k2) * SP_H + k1] = acc;
    }
  }
  __syncthreads();
  if (t < SP_N * SP_H) {   // (k2, k1) = t
    C b[SP_P];
#pragma unroll
    for (int i3 = 0; i3 < SP_P; i3++) b[i3] = B[i3 * SP_N * SP_H + t];
    C *__restrict__ out = S + q * SP_NF + t;
#pragma unroll
    for (int k3 = 0; k3 < SP_N; k3++) {
      C acc = {(T)0, (T)0};
#pragma unroll
      for (int i3 = 0; i3 < SP_P; i3++) {
        const int w = (i3 * k3) & (SP_N - 1);
        cmac(acc, b[i3], (T)W16C[w], -(T)W16S[w]);
      }
      out[k3 * SP_N * SP_H] = acc;
    }
  }
}

// ---------------------------------------------------------------------------
// 2. Hadamard: U[t][w] = sum over the far pairs (t, s) of T_f[w] S_s[w]
// pairf[(k * 8 + c) * 8 + a]: far index of the pair (sibling a, child c of the parent's
// colleague k), -1 when near.
// ---------------------------------------------------------------------------
// The far-pair sources of every sibling group of a level, compacted: the sourced
// children of the parent's 27 colleagues as (level-local source << 8 | 8 k + c), the
// group's entries at cand[gptr[g] .. gptr[g + 1]) (the enumeration of the lists,
// src/tree/lists.jl:96-124, shared by the 8 siblings); none for a ghost group.
__global__ void k_group_cand_count(const int32_t *__restrict__ parent, const int32_t *__restrict__ child_first,
                                   const int32_t *__restrict__ col, const uint8_t *__restrict__ ghost,
                                   const uint8_t *__restrict__ has_source, int64_t lo, int64_t ngroups, int64_t nn,
                                   int32_t *__restrict__ cnt) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= ngroups) return;
  const int64_t first = lo + NCHILD * g;
  int32_t n = 0;
  if (!ghost[first]) {
    const int32_t P = parent[first];
    for (int k = 0; k < NCOLL; k++) {
      const int32_t pc = col[(int64_t)k * nn + P];
      if (pc < 0) continue;
      const int32_t cf = child_first[pc];
      if (cf < 0) continue;
      for (int c = 0; c < NCHILD; c++) n += has_source[cf + c];
    }
  }
  cnt[g] = n;
}
__global__ void k_group_cand_fill(const int32_t *__restrict__ parent, const int32_t *__restrict__ child_first,
                                  const int32_t *__restrict__ col, const uint8_t *__restrict__ ghost,
                                  const uint8_t *__restrict__ has_source, int64_t lo, int64_t ngroups, int64_t nn,
                                  const int32_t *__restrict__ gptr, int32_t *__restrict__ cand) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= ngroups) return;
  const int64_t first = lo + NCHILD * g;
  if (ghost[first]) return;
  const int32_t P = parent[first];
  int32_t pos = gptr[g];
  for (int k = 0; k < NCOLL; k++) {
    const int32_t pc = col[(int64_t)k * nn + P];
    if (pc < 0) continue;
    const int32_t cf = child_first[pc];
    if (cf < 0) continue;
    for (int c = 0; c < NCHILD; c++)
      if (has_source[cf + c]) cand[pos++] = (int32_t)((cf + c - lo) << 8) | (k * NCHILD + c);
  }
}

// One wave per sibling group and frequency block of FB.  The 316 tables of the block
// sit in LDS (rows padded by one entry against bank conflicts, plus a zero row that
// near pairs point at, so the inner loop has no branches), the group's candidate list
// in a per-wave LDS slice.  Each step, lane (phi, r) takes candidate j + r
// (R = 64 / FB candidates per step), reads its spectrum once and applies it to all 8
// siblings (accumulators for the 8 targets in registers); the next step's spectrum is
// loaded before the current one is used.  At the end the R partial sums of each
// (phi, target) are reduced across the wave.
constexpr int HD_TB = 512;   // 8 waves share one copy of the tables

template <class T, int FB>
__global__ void __launch_bounds__(HD_TB) k_spec_hadamard(const typename cplx_t<T>::type *__restrict__ S,
                                                         typename cplx_t<T>::type *__restrict__ U,
                                                         const typename cplx_t<T>::type *__restrict__ tables,
                                                         const int16_t *__restrict__ pairf,
                                                         const int32_t *__restrict__ gptr,
                                                         const int32_t *__restrict__ cand, int64_t ngroups, int gpw) {
  using C = typename cplx_t<T>::type;
  constexpr int R = 64 / FB;               // lanes sharing a frequency: candidates per step
  constexpr int TS = FB + 1;               // padded table row
  constexpr int NW = HD_TB / 64;
  constexpr int MAXC = NCOLL * NCHILD;     // 216 candidates at most
  __shared__ C Tt[(NFAR + 1) * TS];
  __shared__ __attribute__((aligned(16))) int16_t F[(NCOLL * NCHILD + 1) * NCHILD];
  __shared__ int32_t Cl[NW][MAXC + 64];
  const int f0 = blockIdx.y * FB;
  const int t = threadIdx.x;
  for (int i = t; i < (NFAR + 1) * TS; i += HD_TB) {
    const int f = i / TS, ph = i % TS;
    Tt[i] = (f < NFAR && ph < FB) ? tables[(int64_t)f * SP_NF + f0 + ph] : C{(T)0, (T)0};
  }
  for (int i = t; i < (NCOLL * NCHILD + 1) * NCHILD; i += HD_TB) F[i] = pairf[i] < 0 ? (int16_t)NFAR : pairf[i];
  __syncthreads();
  const int wave = t >> 6, lane = t & 63;
  const int phi = lane % FB, r = lane / FB;
  int32_t *cl = Cl[wave];
  const int64_t g0 = ((int64_t)blockIdx.x * NW + wave) * gpw;
  for (int64_t g = g0; g < g0 + gpw && g < ngroups; g++) {
    const int32_t cbeg = gptr[g], n = gptr[g + 1] - cbeg;
    if (n == 0) continue;                    // a ghost group, or no sourced far boxes
    for (int i = lane; i < n; i += 64) cl[i] = cand[cbeg + i];
    __builtin_amdgcn_wave_barrier();
    C acc[NCHILD];
#pragma unroll
    for (int a = 0; a < NCHILD; a++) acc[a] = {(T)0, (T)0};
    // step 0's candidate and spectrum; past the end a lane re-reads entry 0 and adds zeros
    int32_t e = cl[r < n ? r : 0];
    C sv = S[(int64_t)(e >> 8) * SP_NF + f0 + phi];
    bool live = r < n;
    for (int j = 0; j < n; j += R) {
      const int jn = j + R + r;
      const int32_t en = cl[jn < n ? jn : 0];
      const C svn = S[(int64_t)(en >> 8) * SP_NF + f0 + phi];
      const int16_t *Fk = &F[(e & 255) * NCHILD];
      const T sx = live ? sv.x : (T)0, sy = live ? sv.y : (T)0;
#pragma unroll
      for (int a = 0; a < NCHILD; a++) cmac(acc[a], Tt[Fk[a] * TS + phi], sx, sy);
      e = en;
      sv = svn;
      live = jn < n;
    }
    // reduce over the R lanes of each frequency (lanes phi + FB j)
#pragma unroll
    for (int a = 0; a < NCHILD; a++) {
#pragma unroll
      for (int off = FB; off < 64; off <<= 1) {
        acc[a].x += __shfl_xor(acc[a].x, off, 64);
        acc[a].y += __shfl_xor(acc[a].y, off, 64);
      }
    }
    // lane (phi, r) writes the targets a = r, r + R, ...
    constexpr int CPL = NCHILD / R;
    static_assert(CPL * R == NCHILD, "64 / FB must divide 8");
#pragma unroll
    for (int m = 0; m < CPL; m++) {
      const int a = r + m * R;
      C v = acc[0];
#pragma unroll
      for (int b = 1; b < NCHILD; b++)
        if (b == a) v = acc[b];
      U[(NCHILD * g + a) * SP_NF + f0 + phi] = v;
    }
    __builtin_amdgcn_wave_barrier();          // the candidate slice is rewritten next
  }
}

// ---------------------------------------------------------------------------
// 2'. The FP32 Hadamard on the matrix cores.
//
// For one frequency the products of a sibling group are a small matrix product: the
// targets a of group g gain sum over the parent's colleagues k and their children c of
// T[F(k, c, a)] S[child c of colleague k], and the 8 x 8 block T[F(k, c, a)] is the same
// for every group.  So 16 groups (columns) at a time go through
// v_mfma_f32_16x16x4_f32 with rows (target a, real/imaginary part) and, per colleague,
// two K-slices ("quads") of 4 children: one MFMA takes the real parts of the 4 children's
// spectra (row Re: T.re, row Im: T.im), one the imaginary parts (row Re: -T.im, row Im:
// T.re).  A wave works on one tile of 16 groups at a time, for a block of 16
// frequencies (16 accumulators of 4 registers).  A quad's 64 spectrum slices (128
// bytes each) are loaded coalesced, 8 lanes per slice, one quad ahead, and pass through
// a per-wave LDS buffer whose 16-byte chunks are XOR-swizzled so that both the writes
// and the MFMA operand reads are free of bank conflicts.  The block's 316 tables sit in
// LDS, shared by the 8 waves (one tile each).  Quads with no sourced child in any group
// of the tile are skipped; near pairs multiply the zero table row.  Tiles of ghost
// groups are not launched.
// ---------------------------------------------------------------------------
constexpr int HM_FB = 16;            // frequencies per workgroup
constexpr int HM_TS = HM_FB + 2;     // table row stride (144 bytes: rows spread over the LDS banks)
constexpr int HM_W = 8;              // waves per workgroup
constexpr int HM_R = 1;              // tiles per wave
constexpr int HM_TB = 64 * HM_W;
constexpr int HM_TILE = 16;          // sibling groups per tile: the MFMA's columns
typedef float f32x4 __attribute__((ext_vector_type(4)));

// gcol[g * 27 + k]: (first child of the parent's colleague k, level-local) << 8 | bit c
// set when that child has sources; -1 when the colleague is missing or has no sourced
// child, and for every k of a ghost group.  tflag[g / 16] = 1 when some entry is set.
__global__ void k_group_cols(const int32_t *__restrict__ parent, const int32_t *__restrict__ child_first,
                             const int32_t *__restrict__ col, const uint8_t *__restrict__ ghost,
                             const uint8_t *__restrict__ has_source, int64_t lo, int64_t ngroups, int64_t nn,
                             int32_t *__restrict__ gcol, uint8_t *__restrict__ tflag) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= ngroups) return;
  const int64_t first = lo + NCHILD * g;
  const bool real = !ghost[first];
  const int32_t P = parent[first];
  bool any = false;
  for (int k = 0; k < NCOLL; k++) {
    int32_t e = -1;
    const int32_t pc = real ? col[(int64_t)k * nn + P] : -1;
    const int32_t cf = pc >= 0 ? child_first[pc] : -1;
    if (cf >= 0) {
      int32_t m = 0;
      for (int c = 0; c < NCHILD; c++) m |= (int32_t)(has_source[cf + c] != 0) << c;
      if (m) e = (int32_t)((cf - lo) << 8) | m;
    }
    gcol[g * NCOLL + k] = e;
    any |= e >= 0 && k != SELF_SLOT;
  }
  if (any) tflag[g / HM_TILE] = 1;   // every writer stores the same value
}
__global__ void k_fill_u8(uint8_t *__restrict__ a, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = 0;
}

// The staging buffer of a quad: row r = 16 q + j (child q of the quad in group j), 8
// chunks of 16 bytes (2 frequencies); chunk h of row r sits at position h ^ ((r >> 1) & 7).
// Load instruction n (0..7) covers rows 8 n .. 8 n + 7, 8 lanes per row.
__device__ __forceinline__ void hm_quad_load(const float2 *__restrict__ S, int32_t e_lo, int32_t e_hi, int quad,
                                             int f0, int lane, float4 v[8]) {
  // row r = 8 n + (lane >> 3): group (lane >> 3) + 8 (n & 1), child (n >> 1) + 4 quad,
  // position lane & 7 holding chunk (lane & 7) ^ (4 (n & 1) + (lane >> 4))
  const int pos = lane & 7, hb = lane >> 4;
  const int32_t m0 = e_lo >= 0 ? (e_lo >> (4 * quad)) & 15 : 0, m1 = e_hi >= 0 ? (e_hi >> (4 * quad)) & 15 : 0;
  const float2 *b0 = S + (int64_t)((e_lo >> 8) + 4 * quad) * SP_NF + f0 + 2 * (pos ^ hb);
  const float2 *b1 = S + (int64_t)((e_hi >> 8) + 4 * quad) * SP_NF + f0 + 2 * (pos ^ (4 + hb));
#pragma unroll
  for (int n = 0; n < 8; n++) {
    const int32_t m = (n & 1) ? m1 : m0;
    const float2 *bp = (n & 1) ? b1 : b0;
    v[n] = make_float4(0.f, 0.f, 0.f, 0.f);
    if ((m >> (n >> 1)) & 1) v[n] = *reinterpret_cast<const float4 *>(bp + (n >> 1) * SP_NF);
  }
}
__device__ __forceinline__ void hm_quad_store(float4 *buf, int lane, const float4 v[8]) {
#pragma unroll
  for (int n = 0; n < 8; n++) buf[(8 * n + (lane >> 3)) * 8 + (lane & 7)] = v[n];
}

// one quad: 16 frequencies x (real, imaginary) = 32 MFMAs
__device__ __forceinline__ void hm_quad_mfma(f32x4 acc[HM_FB], const float4 *buf, const char *tb, int toff, int q,
                                             int j, bool im_row) {
  const float4 *row = buf + (16 * q + j) * 8;
  const int sw = (j >> 1) & 7;   // row 16 q + j: (r >> 1) & 7 = (j >> 1) & 7
#pragma unroll
  for (int fp = 0; fp < HM_FB / 2; fp++) {
    const float4 tt = *reinterpret_cast<const float4 *>(tb + toff + fp * (int)sizeof(float4));
    const float4 s = row[fp ^ sw];
    acc[2 * fp] = __builtin_amdgcn_mfma_f32_16x16x4f32(im_row ? tt.y : tt.x, s.x, acc[2 * fp], 0, 0, 0);
    acc[2 * fp + 1] = __builtin_amdgcn_mfma_f32_16x16x4f32(im_row ? tt.w : tt.z, s.z, acc[2 * fp + 1], 0, 0, 0);
    acc[2 * fp] = __builtin_amdgcn_mfma_f32_16x16x4f32(im_row ? tt.x : -tt.y, s.y, acc[2 * fp], 0, 0, 0);
    acc[2 * fp + 1] = __builtin_amdgcn_mfma_f32_16x16x4f32(im_row ? tt.z : -tt.w, s.w, acc[2 * fp + 1], 0, 0, 0);
  }
}

__global__ void __launch_bounds__(HM_TB, 1) k_spec_hadamard_mfma(const float2 *__restrict__ S, float2 *__restrict__ U,
                                                                const float2 *__restrict__ tables,
                                                                const int16_t *__restrict__ pairf,
                                                                const int32_t *__restrict__ gcol,
                                                                const int32_t *__restrict__ tiles, int64_t ntiles,
                                                                int64_t ngroups, int64_t nwg, int64_t nreal) {
  __shared__ __attribute__((aligned(16))) float2 Tt[(NFAR + 1) * HM_TS];
  __shared__ int32_t F[NCOLL * NCHILD * NCHILD];   // byte offsets of the table rows
  __shared__ int32_t Gs[HM_W][HM_TILE * NCOLL];
  __shared__ float4 Sb[HM_W][64 * 8];               // per wave: one quad's spectra
  // XCD-aware order: each XCD sweeps a contiguous range of (frequency block, tile)
  const int64_t Lb = xcd_block(blockIdx.x, gridDim.x);
  if (Lb >= nreal) return;   // a padding block
  const int f0 = (int)(Lb / nwg) * HM_FB;
  const int t = threadIdx.x, wave = t >> 6, lane = t & 63;
  for (int i = t; i < (NFAR + 1) * HM_FB; i += HM_TB) {
    const int row = i / HM_FB, ph = i % HM_FB;
    Tt[row * HM_TS + ph] = row < NFAR ? tables[(int64_t)row * SP_NF + f0 + ph] : make_float2(0.f, 0.f);
  }
  for (int i = t; i < NCOLL * NCHILD * NCHILD; i += HM_TB)
    F[i] = (pairf[i] < 0 ? NFAR : pairf[i]) * HM_TS * (int)sizeof(float2);
  __syncthreads();
  const int j = lane & 15, q = lane >> 4, a = j >> 1;
  const bool im_row = j & 1;
  const int32_t *gw = Gs[wave];
  float4 *buf = Sb[wave];
  const char *tb = reinterpret_cast<const char *>(Tt);
  const int jl = lane >> 3;   // the two groups a lane loads for: jl and jl + 8
  // the workgroup's HM_W * HM_R tiles, dealt to the waves round-robin
  for (int r = 0; r < HM_R; r++) {
    const int64_t ti = (Lb % nwg) * (HM_W * HM_R) + r * HM_W + wave;
    if (ti >= ntiles) break;
    const int64_t g0 = (int64_t)tiles[ti] * HM_TILE;
    __builtin_amdgcn_wave_barrier();
    for (int i = lane; i < HM_TILE * NCOLL; i += 64) {
      const int64_t g = g0 + i / NCOLL;
      Gs[wave][i] = g < ngroups ? gcol[g * NCOLL + i % NCOLL] : -1;
    }
    __builtin_amdgcn_wave_barrier();
    // the quads (colleague k, children 4 b .. 4 b + 3) with a sourced child in some
    // group of the tile: bit 2 k + b (wave-uniform)
    uint64_t qmask = 0;
    for (int k = 0; k < NCOLL; k++) {
      const int32_t e = gw[j * NCOLL + k];
      const int32_t m = (k != SELF_SLOT && e >= 0) ? e & 255 : 0;
      if (__ballot(m & 0x0F)) qmask |= 1ull << (2 * k);
      if (__ballot(m & 0xF0)) qmask |= 1ull << (2 * k + 1);
    }
    f32x4 acc[HM_FB];
#pragma unroll
    for (int f = 0; f < HM_FB; f++) acc[f] = f32x4{0.f, 0.f, 0.f, 0.f};
    float4 v[8];
    int qd = qmask ? __builtin_ctzll(qmask) : -1;
    if (qd >= 0) hm_quad_load(S, gw[jl * NCOLL + (qd >> 1)], gw[(jl + 8) * NCOLL + (qd >> 1)], qd & 1, f0, lane, v);
    while (qd >= 0) {
      qmask &= qmask - 1;
      const int qn = qmask ? __builtin_ctzll(qmask) : -1;
      __builtin_amdgcn_wave_barrier();   // the previous quad's operand reads precede these writes
      hm_quad_store(buf, lane, v);
      __builtin_amdgcn_wave_barrier();
      if (qn >= 0)
        hm_quad_load(S, gw[jl * NCOLL + (qn >> 1)], gw[(jl + 8) * NCOLL + (qn >> 1)], qn & 1, f0, lane, v);
      hm_quad_mfma(acc, buf, tb, F[((qd >> 1) * NCHILD + q + 4 * (qd & 1)) * NCHILD + a], q, j, im_row);
      qd = qn;
    }
    // lane (q, j) holds rows 4q .. 4q + 3 of column j: targets 2q and 2q + 1, (re, im) each
    const int64_t g = g0 + j;
    if (g < ngroups) {
      float4 *u0 = reinterpret_cast<float4 *>(U + (int64_t)(NCHILD * g + 2 * q) * SP_NF + f0);
      float4 *u1 = reinterpret_cast<float4 *>(U + (int64_t)(NCHILD * g + 2 * q + 1) * SP_NF + f0);
#pragma unroll
      for (int fp = 0; fp < HM_FB / 2; fp++) {
        u0[fp] = make_float4(acc[2 * fp][0], acc[2 * fp][1], acc[2 * fp + 1][0], acc[2 * fp + 1][1]);
        u1[fp] = make_float4(acc[2 * fp][2], acc[2 * fp][3], acc[2 * fp + 1][2], acc[2 * fp + 1][3]);
      }
    }
  }
}

// ---------------------------------------------------------------------------
// 3. inverse: x = IFFT(U_t) on the [0, p)^3 corner, surface entries into pL
// ---------------------------------------------------------------------------
// One workgroup per target with far pairs: (k2, k1) over k3 (16 inputs read coalesced,
// the p outputs i3 < p kept), then (i3, k1) over k2, then per lattice row (i3, i2) the
// real output over the Hermitian half k1 (x = d0 + (-1)^i1 d_p + 2 Re sum_{k1=1}^{p-1}
// d_k1 w^-(i1 k1)), added into the row's surface points of the partial local.
constexpr int IV_TB = 192;

template <class T>
__global__ void __launch_bounds__(IV_TB) k_spec_inverse(const typename cplx_t<T>::type *__restrict__ U,
                                                        const int32_t *__restrict__ ptr,
                                                        const int16_t *__restrict__ lat8, int Ns, double alpha,
                                                        double *__restrict__ pL) {
  using C = typename cplx_t<T>::type;
  __shared__ C Cb[SP_P * SP_N * SP_H];   // [i3][k2][k1]
  __shared__ C Db[SP_P * SP_P * SP_H];   // [i3][i2][k1]
  const int64_t q = blockIdx.x;
  if (ptr[q + 1] == ptr[q]) return;      // no far pairs (ghosts included)
  const int t = threadIdx.x;
  if (t < SP_N * SP_H) {   // (k2, k1) = t
    const C *__restrict__ in = U + q * SP_NF + t;
    C u[SP_N];
#pragma unroll
    for (int k3 = 0; k3 < SP_N; k3++) u[k3] = in[k3 * SP_N * SP_H];
#pragma unroll
    for (int i3 = 0; i3 < SP_P; i3++) {
      C acc = {(T)0, (T)0};
#pragma unroll
      for (int k3 = 0; k3 < SP_N; k3++) {
        const int w = (i3 * k3) & (SP_N - 1);
        cmac(acc, u[k3], (T)W16C[w], (T)W16S[w]);
      }
      Cb[i3 * SP_N * SP_H + t] = acc;
    }
  }
  __syncthreads();
  if (t < SP_P * SP_H) {   // (i3, k1)
    const int i3 = t / SP_H, k1 = t % SP_H;
    C c[SP_N];
#pragma unroll
    for (int k2 = 0; k2 < SP_N; k2++) c[k2] = Cb[(i3 * SP_N + k2) * SP_H + k1];
#pragma unroll
    for (int i2 = 0; i2 < SP_P; i2++) {
      C acc = {(T)0, (T)0};
#pragma unroll
      for (int k2 = 0; k2 < SP_N; k2++) {
        const int w = (i2 * k2) & (SP_N - 1);
        cmac(acc, c[k2], (T)W16C[w], (T)W16S[w]);
      }
      Db[(i3 * SP_P + i2) * SP_H + k1] = acc;
    }
  }
  __syncthreads();
  if (t < SP_P * SP_P) {   // row (i3, i2) = t
    C d[SP_H];
#pragma unroll
    for (int k1 = 0; k1 < SP_H; k1++) d[k1] = Db[t * SP_H + k1];
    double *__restrict__ out = pL + q * Ns;
#pragma unroll
    for (int i1 = 0; i1 < SP_P; i1++) {
      const int sidx = lat8[t * SP_P + i1];
      if (sidx < 0) continue;
      const T x = d[0].x + ((i1 & 1) ? -d[SP_P].x : d[SP_P].x);
      T y = (T)0;
#pragma unroll
      for (int k1 = 1; k1 < SP_P; k1++) {
        const int w = (i1 * k1) & (SP_N - 1);
        y = fma(d[k1].x, (T)W16C[w], fma(-d[k1].y, (T)W16S[w], y));
      }
      out[sidx] += alpha * (double)(x + (T)2 * y);
    }
  }
}

// host side -----------------------------------------------------------------
void spectral_workspace(Plan &plan) {
  int64_t nmax = 1;
  for (int l = 2; l < plan.L; l++) nmax = std::max(nmax, plan.nl[l]);
  if (plan.precision == FFMM_PREC_MIXED) {
    plan.spec_S32.resize((size_t)2 * SP_NF * nmax);
    plan.spec_U32.resize((size_t)2 * SP_NF * nmax);
  } else {
    plan.spec_S.resize((size_t)2 * SP_NF * nmax);
    plan.spec_U.resize((size_t)2 * SP_NF * nmax);
  }
}

void spectral_prepare(Plan &plan) {
  OpCache &o = *plan.ops;
  if (o.p != SP_P) fail(FFMM_ERR_ARGUMENT, "the spectral M2L is built for p = 8");
  spectral_workspace(plan);
  Context &ctx = *plan.ctx;
  hipStream_t s = ctx.stream;
  if (plan.precision == FFMM_PREC_MIXED) {
    // the groups' colleague tables and the tiles with far sources, once per plan
    plan.spec_gcol.clear();
    plan.spec_tiles.clear();
    plan.spec_gcol.resize(plan.L);
    plan.spec_tiles.resize(plan.L);
    plan.spec_ntiles.assign(plan.L, 0);
    DBuf<uint8_t> tflag;
    for (int l = 2; l < plan.L; l++) {
      const int64_t lo = plan.lo[l], ng = plan.nl[l] / NCHILD, nt = (ng + HM_TILE - 1) / HM_TILE;
      if (ng == 0) continue;
      plan.spec_gcol[l].resize((size_t)ng * NCOLL);
      tflag.resize(nt);
      k_fill_u8<<<nblocks(nt, 256), 256, 0, s>>>(tflag.p, nt);
      k_group_cols<<<nblocks(ng, 256), 256, 0, s>>>(plan.parent.p, plan.child_first.p, plan.col.p, plan.ghost.p,
                                                    plan.has_source.p, lo, ng, plan.nnodes, plan.spec_gcol[l].p,
                                                    tflag.p);
      FFMM_LAUNCH_CHECK();
      plan.spec_tiles[l].resize(nt);
      plan.spec_ntiles[l] = compact_flags_u8(ctx, tflag.p, plan.spec_tiles[l].p, nt);
    }
    return;
  }
  // the sibling groups' candidate lists, once per plan (they follow the tree only)
  plan.spec_gptr.clear();
  plan.spec_cand.clear();
  plan.spec_gptr.resize(plan.L);
  plan.spec_cand.resize(plan.L);
  DBuf<int32_t> cnt;
  for (int l = 2; l < plan.L; l++) {
    const int64_t lo = plan.lo[l], ng = plan.nl[l] / NCHILD;
    cnt.resize(ng);
    plan.spec_gptr[l].resize(ng + 1);
    k_group_cand_count<<<nblocks(ng, 256), 256, 0, s>>>(plan.parent.p, plan.child_first.p, plan.col.p, plan.ghost.p,
                                                         plan.has_source.p, lo, ng, plan.nnodes, cnt.p);
    FFMM_LAUNCH_CHECK();
    const int64_t total = exclusive_scan_i32(ctx, cnt.p, plan.spec_gptr[l].p, ng);
    set_i32(ctx, plan.spec_gptr[l].p + ng, (int32_t)total);
    plan.spec_cand[l].resize(total);
    if (total > 0) {
      k_group_cand_fill<<<nblocks(ng, 256), 256, 0, s>>>(plan.parent.p, plan.child_first.p, plan.col.p, plan.ghost.p,
                                                          plan.has_source.p, lo, ng, plan.nnodes, plan.spec_gptr[l].p,
                                                          plan.spec_cand[l].p);
      FFMM_LAUNCH_CHECK();
    }
  }
}

template <class T, int FB>
static void spectral_level(Plan &plan, int l, typename cplx_t<T>::type *S, typename cplx_t<T>::type *U,
                           const typename cplx_t<T>::type *tables) {
  Context &ctx = *plan.ctx;
  OpCache &o = *plan.ops;
  hipStream_t s = ctx.stream;
  const int64_t lo = plan.lo[l], nl = plan.nl[l], ng = nl / NCHILD;
  if (plan.lists[l].np == 0) return;
  k_spec_forward<T><<<(unsigned)nl, FW_TB, 0, s>>>(plan.Mp.p, o.Ns, o.lat8.p, lo, nl, plan.has_source.p, S);
  FFMM_LAUNCH_CHECK();
  const int gpw = 4;
  const int64_t gpb = (HD_TB / 64) * gpw;
  dim3 grid((unsigned)((ng + gpb - 1) / gpb), SP_NF / FB);
  {
    StageTimer tm(plan, 8, l);   // the Hadamard launch alone
    if constexpr (std::is_same<T, float>::value) {
      const int64_t ntiles = plan.spec_ntiles[l];
      if (ntiles > 0) {
        const int64_t nwg = (ntiles + HM_W * HM_R - 1) / (HM_W * HM_R), nreal = nwg * (SP_NF / HM_FB);
        k_spec_hadamard_mfma<<<(unsigned)pad_xcd(nreal), HM_TB, 0, s>>>(S, U, tables, o.pairf.p, plan.spec_gcol[l].p,
                                                                        plan.spec_tiles[l].p, ntiles, ng, nwg, nreal);
      }
    }
    else
      k_spec_hadamard<T, FB><<<grid, HD_TB, 0, s>>>(S, U, tables, o.pairf.p, plan.spec_gptr[l].p, plan.spec_cand[l].p,
                                                   ng, gpw);
    FFMM_LAUNCH_CHECK();
  }
  const double alpha = (1.0 / plan.lam[l]) / (double)(SP_N * SP_N * SP_N);
  k_spec_inverse<T><<<(unsigned)nl, IV_TB, 0, s>>>(U, plan.lists[l].ptr.p, o.lat8.p, o.Ns, alpha,
                                                  plan.Pl.p + lo * o.Ns);
  FFMM_LAUNCH_CHECK();
}

void m2l_spectral_level(Plan &plan, int l) {
  OpCache &o = *plan.ops;
  if (plan.precision == FFMM_PREC_MIXED)
    spectral_level<float, 16>(plan, l, (float2 *)plan.spec_S32.p, (float2 *)plan.spec_U32.p,
                              (const float2 *)o.tables32.p);
  else
    spectral_level<double, 8>(plan, l, (double2 *)plan.spec_S.p, (double2 *)plan.spec_U.p,
                              (const double2 *)o.tables.p);
}

}  // namespace ffmm
