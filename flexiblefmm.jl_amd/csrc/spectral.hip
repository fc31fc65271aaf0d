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
__global__ void k_take_i32(const int32_t *__restrict__ in, const int32_t *__restrict__ idx, int64_t n,
                           int32_t *__restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = in[idx[i]];
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
// 2''. The FP32 Hadamard on the matrix cores, each tile's source boxes staged once.
//
// The groups of a tile are consecutive in the level's order, so their parents lie close
// together and share colleagues: a tile's (group, colleague) entries name about a third
// as many distinct boxes (a compact 2 x 2 x 4 block of parents: 96 boxes for 416
// entries).  Once per plan, each tile's distinct sourced colleagues (its union, at most
// HU_UMAX) and every group's slot in it are found.  A workgroup takes a block of 8
// frequencies and HU_R consecutive tiles.  Per tile it stages the union's child spectra
// in LDS (zeros for children without sources; one zero box that missing colleagues
// point at), the next tile's spectra being loaded into registers meanwhile.  Wave f runs
// the MFMAs of frequency f (the MFMA form of k_spec_hadamard_mfma) over all 52 quads,
// unrolled: the table operands of its quads are read once per workgroup into registers
// (104 per lane), the groups' slots once per tile, so each pair of MFMAs reads one
// spectrum value from LDS.  Tiles of the busy levels have a sourced child in nearly
// every quad, so quads without one cost little: they multiply zeros.  The tile's 128
// target rows go out through LDS, 64 bytes each.  So a spectrum slice is read from L2
// once per tile, not once per group that has its box as a colleague.  Tiles whose union
// is larger go to k_spec_hadamard_mfma.
// ---------------------------------------------------------------------------
constexpr int HU_FB = 4;                    // frequencies per workgroup
constexpr int HU_TB = 64 * HU_FB * 2;       // two waves per frequency, 13 colleagues each
constexpr int HU_UMAX = 200;                // distinct colleagues staged per tile
constexpr int HU_NC = 32;                   // colours of the staged boxes
constexpr int HU_NR = 224;                  // rows: colour + 32 x (index within the colour)
constexpr int HU_ZROW = HU_NR;              // a zero box
constexpr int HU_NRS = 256;                 // staged rows per child: thread t stages row t / 2
constexpr int HU_US = HU_NRS + 1;           // rows per child in LDS: odd, so that children q and 4 + q
                                            // (read back to back) sit half the banks apart
constexpr int HU_NS = HM_TILE * NCOLL;      // (group, colleague) slots of a tile
constexpr int HU_KH = (NCOLL - 1) / 2;      // colleagues per wave: k < 13, then k > 13
constexpr int HU_SW = (HU_KH + 1) / 2;      // 4-byte words of a group's slot offsets per half in LDS
constexpr int HU_R = 16;                    // tiles per workgroup
constexpr int HU_IS = HU_FB / 2;            // 16-byte staging items per slice
constexpr int HU_NI = NCHILD * HU_NRS * HU_IS;                   // per tile
constexpr int HU_NV = HU_NI / HU_TB;                             // per thread: one per child
static_assert(HU_NV == NCHILD && HU_NRS * HU_IS == HU_TB && HU_ZROW < HU_NRS, "item n of thread t: row t / 2, child n");
constexpr int HU_NQ = 2 * HU_KH;            // quads of a wave: its colleagues, 2 halves each
static_assert(HU_NC * ((HU_UMAX + HU_NC - 1) / HU_NC) <= HU_NR && HU_ZROW <= 255, "rows are bytes");

// Per active tile ti (tile id tiles[ti]): the distinct sourced colleagues of its groups
// other than the parents themselves (its union), each placed in a row of the staging
// buffer; ulist[ti][row] = the box's (level-local first child << 8 | sourced-children
// mask), 0 for an unused row; uslot[ti][j][k] = the row of group j's colleague k
// (HU_ZROW when it has none); fit[ti] = 1 when the union has at most HU_UMAX boxes (else
// over[ti] = 1).  The 16 groups' colleague k are read by one MFMA, 16 lanes (one LDS
// phase of 8-byte reads) per child; a box's row is its colour + 32 x (its place among
// its colour), the colours chosen greedily (in order of first appearance over (group,
// colleague)) so that the boxes of one colleague index have distinct colours: then
// those 16 reads fall in distinct LDS banks.  With 32 colours for the 16 boxes of a
// colleague index the greedy choice nearly always succeeds; where it does not, a
// colour with room is taken (a bank conflict, the result unchanged).  One wave per tile.
__global__ void __launch_bounds__(64) k_tile_union(const int32_t *__restrict__ gcol, const int32_t *__restrict__ tiles,
                                                   int64_t ngroups, int32_t *__restrict__ ulist,
                                                   uint8_t *__restrict__ uslot, uint8_t *__restrict__ fit,
                                                   uint8_t *__restrict__ over) {
  constexpr int HT = 1024;                 // hash table of the tile's boxes (at most 416 distinct)
  __shared__ int32_t e[HU_NS];
  __shared__ int16_t first[HU_NS];
  __shared__ int16_t slot[HU_NS];
  __shared__ int32_t hkey[HT], hmin[HT];
  __shared__ uint32_t colset[HU_UMAX];     // per box: the colleague indices it appears at
  __shared__ uint8_t row[HU_UMAX];
  const int64_t ti = blockIdx.x;
  const int lane = threadIdx.x;
  const int64_t g0 = (int64_t)tiles[ti] * HM_TILE;
  for (int i = lane; i < HT; i += 64) {
    hkey[i] = -1;
    hmin[i] = HU_NS;
  }
  for (int i = lane; i < HU_NS; i += 64) {
    const int64_t g = g0 + i / NCOLL;
    const int k = i % NCOLL;
    e[i] = (g < ngroups && k != SELF_SLOT) ? gcol[g * NCOLL + k] : -1;
  }
  __syncthreads();
  // each box's first entry (the smallest index naming it), by open addressing
  int hs[(HU_NS + 63) / 64];
  for (int n = 0, i = lane; i < HU_NS; i += 64, n++) {
    const int32_t v = e[i];
    hs[n] = -1;
    if (v < 0) continue;
    int h = (int)(((uint32_t)v * 2654435761u) >> 22) & (HT - 1);
    for (;;) {
      const int32_t old = atomicCAS(&hkey[h], -1, v);
      if (old == -1 || old == v) break;
      h = (h + 1) & (HT - 1);
    }
    atomicMin(&hmin[h], i);
    hs[n] = h;
  }
  __syncthreads();
  for (int n = 0, i = lane; i < HU_NS; i += 64, n++) first[i] = (int16_t)(hs[n] >= 0 ? hmin[hs[n]] : i);
  __syncthreads();
  int32_t count = 0;
  for (int base = 0; base < HU_NS; base += 64) {
    const int i = base + lane;
    const bool isfirst = i < HU_NS && e[i] >= 0 && first[i] == i;
    const uint64_t b = __ballot(isfirst);
    if (isfirst) slot[i] = (int16_t)(count + __popcll(b & ((1ull << lane) - 1)));
    count += __popcll(b);
  }
  const bool ok = count <= HU_UMAX;
  if (lane == 0) {
    fit[ti] = ok;
    over[ti] = !ok;
  }
  if (!ok) return;
  for (int u = lane; u < count; u += 64) colset[u] = 0;
  __syncthreads();
  for (int i = lane; i < HU_NS; i += 64)
    if (e[i] >= 0) atomicOr(&colset[slot[first[i]]], 1u << (i % NCOLL));
  __syncthreads();
  {   // greedy colouring, one box at a time, the state spread over the wave's lanes:
      // lane k < 27 holds the colours taken at colleague index k, lane c < 32 the boxes
      // of colour c so far (the loop is wave-uniform)
    constexpr int CAP = HU_NR / HU_NC;
    static_assert(HU_NC == 32, "a colour set is a 32-bit mask");
    uint32_t cm = 0, cnt = 0;
    uint32_t avail = 0xFFFFFFFFu;          // colours with room
    uint32_t csn = colset[0];
    for (int u = 0; u < count; u++) {
      const uint32_t cs = __builtin_amdgcn_readfirstlane(csn);
      if (u + 1 < count) csn = colset[u + 1];
      uint32_t forb = 0;
      for (uint32_t m = cs; m; m &= m - 1) forb |= __builtin_amdgcn_readlane(cm, __builtin_ctz(m));
      const uint32_t free = avail & ~forb;
      const int pick = __builtin_ctz(free ? free : avail);   // none free: a bank conflict, still correct
      const uint32_t c_cnt = __builtin_amdgcn_readlane(cnt, pick);
      if (lane == 0) row[u] = (uint8_t)(pick + HU_NC * c_cnt);
      if (lane == pick) cnt++;
      if (c_cnt + 1 == CAP) avail &= ~(1u << pick);
      if ((cs >> lane) & 1) cm |= 1u << pick;
    }
  }
  __syncthreads();
  for (int r = lane; r < HU_NRS; r += 64) ulist[ti * HU_NRS + r] = 0;
  __syncthreads();
  for (int i = lane; i < HU_NS; i += 64) {
    const int32_t v = e[i];
    if (v >= 0 && first[i] == i) ulist[ti * HU_NRS + row[slot[i]]] = v;
    uslot[ti * HU_NS + i] = (uint8_t)(v >= 0 ? row[slot[first[i]]] : HU_ZROW);
  }
}

// LDS position of frequency f of child c of the box in row `row`: the 4 frequencies of
// a slice (32 bytes, rows 8 to the 256 bytes of the 64 banks) are XOR-swizzled by row
// bits 3-4, so that the 16 lanes of one read phase (one child), at rows of distinct
// colours (distinct mod 32), touch distinct banks.
__device__ __forceinline__ int hu_swz(int row) { return (row >> 3) & 3; }
__device__ __forceinline__ int hu_spos(int c, int row, int f) { return (c * HU_US + row) * HU_FB + (f ^ hu_swz(row)); }
// the output buffer: frequency-major, one pad slot per 8 targets, so that the 16 groups'
// target 2 q of one write phase touch distinct banks
constexpr int HU_OS = HM_TILE * NCHILD + HM_TILE;   // 144 slots per frequency
__device__ __forceinline__ int hu_opos(int ta, int f) { return f * HU_OS + ta + (ta >> 3); }

// Staging items: 16 bytes (frequencies 2 h, 2 h + 1) of child c of the box in row r.
// Item n of thread t is (h, r, c) = (t % 2, t / 2, n): 2 consecutive items are one
// 32-byte slice, 16 consecutive threads 8 consecutive rows.
// The staged spectra of the tile whose rows' entries are in Ent (LDS); bit n of live:
// item n has a box (an unused row is not written).
__device__ __forceinline__ void hu_load(const float2 *__restrict__ S, const int32_t *Ent, int f0, int t,
                                        float4 v[HU_NV], uint32_t &live) {
  const int h = t % HU_IS, r = t / HU_IS;
  const int32_t ent = Ent[r];
  live = ent != 0 ? (1u << HU_NV) - 1 : 0;
  // 32-bit byte offsets into the level's spectra (the host keeps them under 4 GiB)
  const uint32_t base = ((uint32_t)(ent >> 8) * SP_NF + (uint32_t)(f0 + 2 * h)) * (uint32_t)sizeof(float2);
#pragma unroll
  for (int n = 0; n < HU_NV; n++) {
    v[n] = make_float4(0.f, 0.f, 0.f, 0.f);
    if ((ent >> n) & 1)
      v[n] = *reinterpret_cast<const float4 *>(reinterpret_cast<const char *>(S) + base +
                                               (uint32_t)n * SP_NF * (uint32_t)sizeof(float2));
  }
}
__device__ __forceinline__ void hu_store(float2 *Us, int t, const float4 v[HU_NV], uint32_t live) {
  if (!live) return;
  const int h = t % HU_IS, r = t / HU_IS;
  // frequencies 2 h and 2 h + 1 sit side by side at chunk h ^ (swz >> 1), swapped when
  // the swizzle is odd
  const int sw = hu_swz(r);
  const int ch = 2 * (h ^ (sw >> 1));
#pragma unroll
  for (int n = 0; n < HU_NV; n++) {
    const float4 w = (sw & 1) ? make_float4(v[n].z, v[n].w, v[n].x, v[n].y) : v[n];
    *reinterpret_cast<float4 *>(Us + (n * HU_US + r) * HU_FB + ch) = w;
  }
}

// A workgroup barrier that orders LDS accesses only: no wait for the global loads and
// stores in flight (a prefetch consumed by the same thread, results written out).
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// two workgroups per CU (8 waves each): 4 waves per SIMD, at most 128 VGPRs
__global__ void __launch_bounds__(HU_TB) __attribute__((amdgpu_waves_per_eu(4, 4))) k_spec_hadamard_union(
    const float2 *__restrict__ S, float2 *__restrict__ U, const float2 *__restrict__ tables,
    const int16_t *__restrict__ pairf, const int32_t *__restrict__ tiles, const int32_t *__restrict__ fitl,
    const int32_t *__restrict__ ulist, const uint8_t *__restrict__ uslot, int64_t nfit, int64_t ngroups,
    int64_t nranges, int64_t nreal) {
  __shared__ __attribute__((aligned(16))) float2 Us[NCHILD * HU_US * HU_FB];
  __shared__ __attribute__((aligned(16))) float2 O[2][HU_FB * HU_OS];   // the two halves' sums
  // group j's colleagues of half h: 16-bit byte offsets of their rows, swizzle bits set
  // (row * 32 + 8 * ((row >> 3) & 3)), in words (2 j + h) * HU_SW ..
  __shared__ uint32_t Sl[HM_TILE * 2 * HU_SW];
  __shared__ int32_t Ent[HU_NRS];            // the rows' entries of the tile loaded next
  static_assert((NFAR + 1) * HU_FB * sizeof(float2) + NCOLL * NCHILD * NCHILD * sizeof(int16_t) <= sizeof(Us),
                "the tables are staged where the spectra go");
  static_assert(HU_NRS <= HU_TB && HU_NS <= HU_TB, "one entry and one slot per thread");
  // XCD-aware order: each XCD sweeps a contiguous range; consecutive blocks are the four
  // quarters (4 frequencies each) of the same tiles' 128-byte spectrum slices
  constexpr int NQB = 16 / HU_FB;
  const int64_t Lb = xcd_block(blockIdx.x, gridDim.x);
  if (Lb >= nreal) return;   // a padding block
  const int64_t fline = Lb / (NQB * nranges), rr = Lb % (NQB * nranges);
  const int64_t fi0 = (rr / NQB) * HU_R, fi1 = fi0 + HU_R < nfit ? fi0 + HU_R : nfit;
  const int f0 = (int)(NQB * fline + rr % NQB) * HU_FB;
  const int t = threadIdx.x;
  const int wave = __builtin_amdgcn_readfirstlane(t >> 6), lane = t & 63;
  const int f = wave % HU_FB, kh = wave / HU_FB;   // frequency f0 + f, colleagues kh ? 14 .. 26 : 0 .. 12
  const int k0 = kh ? SELF_SLOT + 1 : 0;
  const int j = lane & 15, q = lane >> 4, a = j >> 1;
  const bool im_row = j & 1;
  // the first tile's entries, then (with the tables) its spectra and slots and the
  // second tile's entries
  int64_t ti = fitl[fi0];
  if (t < HU_NRS) Ent[t] = ulist[ti * HU_NRS + t];
  {
    float2 *Tt = Us;
    int16_t *F = reinterpret_cast<int16_t *>(Us + (NFAR + 1) * HU_FB);
    for (int i = t; i < (NFAR + 1) * HU_FB; i += HU_TB) {
      const int row = i / HU_FB, ph = i % HU_FB;
      Tt[i] = row < NFAR ? tables[(int64_t)row * SP_NF + f0 + ph] : make_float2(0.f, 0.f);
    }
    for (int i = t; i < NCOLL * NCHILD * NCHILD; i += HU_TB) F[i] = pairf[i] < 0 ? (int16_t)NFAR : pairf[i];
  }
  lds_barrier();
  float4 v[HU_NV];
  uint32_t live;
  hu_load(S, Ent, f0, t, v, live);
  uint32_t sw = t < HU_NS ? uslot[ti * HU_NS + t] : 0;
  int32_t enext = 0;
  if (fi0 + 1 < fi1 && t < HU_NRS) enext = ulist[(int64_t)fitl[fi0 + 1] * HU_NRS + t];
  // each lane's table operands: quad (k, b) multiplies child 4 b + q of colleague k into
  // target a; the real parts' MFMA takes (row Re: T.re, row Im: T.im), the imaginary
  // parts' (row Re: -T.im, row Im: T.re).  The latter is the neighbouring lane's former
  // (rows Re and Im of a target are lanes 2 m and 2 m + 1) with row Re negated, so the
  // imaginary parts' products are accumulated unsigned and row Re subtracted at the end
  float A1[HU_NQ];
  {
    const float2 *Tt = Us;
    const int16_t *F = reinterpret_cast<const int16_t *>(Us + (NFAR + 1) * HU_FB);
#pragma unroll
    for (int kk = 0, qi = 0; kk < HU_KH; kk++) {
#pragma unroll
      for (int b = 0; b < 2; b++, qi++) {
        const float2 T = Tt[F[((k0 + kk) * NCHILD + 4 * b + q) * NCHILD + a] * HU_FB + f];
        A1[qi] = im_row ? T.y : T.x;
      }
    }
  }
  lds_barrier();   // the tables are read: Us takes the spectra
  if (t < NCHILD * HU_FB) Us[hu_spos(t / HU_FB, HU_ZROW, t % HU_FB)] = make_float2(0.f, 0.f);
  // a tile's 128 target rows, out of LDS (the two halves' sums) 64 bytes each
  auto write_out = [&](int64_t tl) {
    const int64_t g0 = (int64_t)tiles[tl] * HM_TILE;
    for (int i = t; i < HM_TILE * NCHILD * (HU_FB / 2); i += HU_TB) {
      const int h = i % (HU_FB / 2), ta = i / (HU_FB / 2);
      const int64_t g = g0 + ta / NCHILD;
      if (g >= ngroups) continue;
      const float2 x0 = O[0][hu_opos(ta, 2 * h)], x1 = O[0][hu_opos(ta, 2 * h + 1)];
      const float2 y0 = O[1][hu_opos(ta, 2 * h)], y1 = O[1][hu_opos(ta, 2 * h + 1)];
      *reinterpret_cast<float4 *>(U + (NCHILD * g + ta % NCHILD) * SP_NF + f0 + 2 * h) =
          make_float4(x0.x + y0.x, x0.y + y0.y, x1.x + y1.x, x1.y + y1.y);
    }
  };
  // per tile: stage its spectra (loaded during the previous tile's MFMAs) and write the
  // previous tile's rows out, then issue the next tile's loads and run this tile's MFMAs;
  // the rows are written after the next staging, so that no wait for them holds up the
  // staging's own
  int64_t tprev = -1;
  for (int64_t fi = fi0; fi < fi1; fi++) {
    if (fi > fi0) lds_barrier();   // the previous tile's reads of Us, Sl, Ent done, its O written
    hu_store(Us, t, v, live);
    if (t < HU_NS && t % NCOLL != SELF_SLOT) {
      const int kt = t % NCOLL, h = kt > SELF_SLOT;
      reinterpret_cast<uint16_t *>(Sl)[(2 * (t / NCOLL) + h) * HU_SW * 2 + kt - (h ? SELF_SLOT + 1 : 0)] =
          (uint16_t)(sw * HU_FB * sizeof(float2) + sizeof(float2) * hu_swz((int)sw));
    }
    if (t < HU_NRS) Ent[t] = enext;
    if (tprev >= 0) write_out(tprev);
    const int64_t tcur = ti;
    lds_barrier();
    if (fi + 1 < fi1) {   // the next tile's spectra and slots, and the entries after it
      ti = fitl[fi + 1];
      hu_load(S, Ent, f0, t, v, live);
      if (t < HU_NS) sw = uslot[ti * HU_NS + t];
      if (fi + 2 < fi1 && t < HU_NRS) enext = ulist[(int64_t)fitl[fi + 2] * HU_NRS + t];
    }
    // child q of colleague kk's box, frequency f: byte q HU_US 32 + (offset ^ 8 f)
    // (hu_spos); children 4 + q: + 4 HU_US rows
    uint32_t slw[HU_SW];
#pragma unroll
    for (int w = 0; w < HU_SW; w++) slw[w] = Sl[(2 * j + kh) * HU_SW + w];
    const char *Ub = reinterpret_cast<const char *>(Us) + q * HU_US * HU_FB * sizeof(float2);
    const uint32_t fx = (uint32_t)sizeof(float2) * (uint32_t)f;
    constexpr int HALF = 4 * HU_US * HU_FB * (int)sizeof(float2);   // children 4 .. 7
    // the MFMA loop at a raised wave priority: the SIMD's issue favours the waves in it
    // over the other workgroup's staging and write-out (17.5 -> 17.1 ms per 4M step)
    __builtin_amdgcn_s_setprio(1);
    f32x4 acc0 = f32x4{0.f, 0.f, 0.f, 0.f}, acc1 = acc0;   // the real and the imaginary parts' products
#pragma unroll
    for (int kk = 0, qi = 0; kk < HU_KH; kk++) {
#pragma unroll
      for (int b = 0; b < 2; b++, qi++) {
        const uint32_t off = ((slw[kk / 2] >> (16 * (kk % 2))) & 0xFFFF) ^ fx;
        const float2 Sv = *reinterpret_cast<const float2 *>(Ub + off + b * HALF);
        // quad_perm [1, 0, 3, 2]: lanes 2 m and 2 m + 1 swap
        const float A2 = __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(A1[qi]), 0xB1, 0xF, 0xF, false));
        acc0 = __builtin_amdgcn_mfma_f32_16x16x4f32(A1[qi], Sv.x, acc0, 0, 0, 0);
        acc1 = __builtin_amdgcn_mfma_f32_16x16x4f32(A2, Sv.y, acc1, 0, 0, 0);
      }
    }
    // lane (q, j) holds targets 2 q and 2 q + 1 of group j, (re, im) each: the two
    // halves' sums through LDS to 64-byte rows per target
    __builtin_amdgcn_s_setprio(0);
    const f32x4 s4 = f32x4{acc0[0] - acc1[0], acc0[1] + acc1[1], acc0[2] - acc1[2], acc0[3] + acc1[3]};
    const int ta0 = j * NCHILD + 2 * q;
    O[kh][hu_opos(ta0, f)] = make_float2(s4[0], s4[1]);
    O[kh][hu_opos(ta0 + 1, f)] = make_float2(s4[2], s4[3]);
    tprev = tcur;
  }
  lds_barrier();
  write_out(tprev);
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
    // once per plan: the groups' colleague tables and the tiles with far sources (all
    // levels, one read back), then each active tile's union and whether it fits (one
    // more read back)
    const int L = plan.L;
    static_assert(MAX_KEY_LEVELS + 2 <= NPINNED / 2, "two pinned slots per level");
    plan.spec_gcol.clear();
    plan.spec_tiles.clear();
    plan.spec_gcol.resize(L);
    plan.spec_tiles.resize(L);
    plan.spec_ntiles.assign(L, 0);
    std::vector<DBuf<uint8_t>> tflag(L);
    for (int l = 2; l < L; l++) {
      const int64_t lo = plan.lo[l], ng = plan.nl[l] / NCHILD, nt = (ng + HM_TILE - 1) / HM_TILE;
      if (ng == 0) continue;
      plan.spec_gcol[l].resize((size_t)ng * NCOLL);
      tflag[l].resize(nt);
      k_fill_u8<<<nblocks(nt, 256), 256, 0, s>>>(tflag[l].p, nt);
      k_group_cols<<<nblocks(ng, 256), 256, 0, s>>>(plan.parent.p, plan.child_first.p, plan.col.p, plan.ghost.p,
                                                    plan.has_source.p, lo, ng, plan.nnodes, plan.spec_gcol[l].p,
                                                    tflag[l].p);
      FFMM_LAUNCH_CHECK();
      plan.spec_tiles[l].resize(nt);
      arm_slots(ctx, l, 1);
      compact_flags_u8_async(ctx, tflag[l].p, plan.spec_tiles[l].p, nt, l);
    }
    for (int l = 2; l < L; l++) plan.spec_ntiles[l] = plan.nl[l] / NCHILD > 0 ? read_slot(ctx, l) : 0;
    for (auto *v : {&plan.spec_ulist, &plan.spec_fit, &plan.spec_over}) {
      v->clear();
      v->resize(L);
    }
    plan.spec_uslot.clear();
    plan.spec_uslot.resize(L);
    plan.spec_nfit.assign(L, 0);
    plan.spec_nover.assign(L, 0);
    std::vector<DBuf<int32_t>> overidx(L);
    for (int l = 2; l < L; l++) {
      const int64_t nt = plan.spec_ntiles[l];
      if (nt == 0) continue;
      plan.spec_ulist[l].resize((size_t)nt * HU_NRS);
      plan.spec_uslot[l].resize((size_t)nt * HU_NS);
      plan.spec_fit[l].resize(nt);
      overidx[l].resize(nt);
      tflag[l].resize(2 * nt);   // fit flags, then overflow flags
      k_tile_union<<<(unsigned)nt, 64, 0, s>>>(plan.spec_gcol[l].p, plan.spec_tiles[l].p, plan.nl[l] / NCHILD,
                                              plan.spec_ulist[l].p, plan.spec_uslot[l].p, tflag[l].p,
                                              tflag[l].p + nt);
      FFMM_LAUNCH_CHECK();
      arm_slots(ctx, l, 1);
      arm_slots(ctx, NPINNED / 2 + l, 1);
      compact_flags_u8_async(ctx, tflag[l].p, plan.spec_fit[l].p, nt, l);
      compact_flags_u8_async(ctx, tflag[l].p + nt, overidx[l].p, nt, NPINNED / 2 + l);
    }
    for (int l = 2; l < L; l++) {
      if (plan.spec_ntiles[l] == 0) continue;
      plan.spec_nfit[l] = read_slot(ctx, l);
      const int64_t no = plan.spec_nover[l] = read_slot(ctx, NPINNED / 2 + l);
      if (no == 0) continue;
      plan.spec_over[l].resize(no);   // the tile ids themselves, for k_spec_hadamard_mfma
      k_take_i32<<<nblocks(no, 256), 256, 0, s>>>(plan.spec_tiles[l].p, overidx[l].p, no, plan.spec_over[l].p);
      FFMM_LAUNCH_CHECK();
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

// The tiles of level l run by the union kernel (staged) and by the per-group kernel; true
// when the union kernel takes part.  It addresses a level's spectra with 32-bit byte
// offsets: a level whose FP32 spectra reach 4 GiB runs every tile in the per-group kernel.
bool spectral_tile_counts(const Plan &plan, int l, int64_t &staged, int64_t &per_group) {
  staged = per_group = 0;
  if (plan.precision != FFMM_PREC_MIXED || l >= (int)plan.spec_nfit.size()) return false;
  const bool small = plan.nl[l] * SP_NF * (int64_t)sizeof(float2) < ((int64_t)1 << 32);
  staged = small ? plan.spec_nfit[l] : 0;
  per_group = small ? plan.spec_nover[l] : plan.spec_ntiles[l];
  return small;
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
      int64_t nfit, nover;
      const bool staged = spectral_tile_counts(plan, l, nfit, nover);
      const int32_t *over_tiles = staged ? plan.spec_over[l].p : plan.spec_tiles[l].p;
      if (nfit > 0) {
        const int64_t nranges = (nfit + HU_R - 1) / HU_R, nreal = nranges * (SP_NF / HU_FB);
        k_spec_hadamard_union<<<(unsigned)pad_xcd(nreal), HU_TB, 0, s>>>(
            S, U, tables, o.pairf.p, plan.spec_tiles[l].p, plan.spec_fit[l].p, plan.spec_ulist[l].p,
            plan.spec_uslot[l].p, nfit, ng, nranges, nreal);
        FFMM_LAUNCH_CHECK();
      }
      if (nover > 0) {
        const int64_t nwg = (nover + HM_W * HM_R - 1) / (HM_W * HM_R), nreal = nwg * (SP_NF / HM_FB);
        k_spec_hadamard_mfma<<<(unsigned)pad_xcd(nreal), HM_TB, 0, s>>>(S, U, tables, o.pairf.p, plan.spec_gcol[l].p,
                                                                        over_tiles, nover, ng, nwg, nreal);
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
