// The harmonic methods on the device: the multipoles of the childless nodes (S2M)
// and the evaluation (local expansion of the leaf plus the near field of its
// colleagues) for SphericalHarmonics with Laplace3D or Yukawa3D charges and for
// StokesHarmonics with stokeslets and stresslets.
//
// S2M: one wave per node, over tiles of the node's sources; the lanes of a source run
// the order chains of the Schmidt recurrences (_fold_chain,
// src/operators/harmonics.jl:681-704) in pairs of orders (m, p - m) and store the
// harmonics in LDS, then each lane folds the tile into the entries of the multipole
// that it owns (_s2m_harmonic_kernel!, src/device_passes.jl:343-373, which runs one
// work-item per node and order).
//
// Evaluation: one wave per leaf; as in the cube evaluation, lanes take targets and,
// when a leaf has fewer than 64, R replicas of each target split the orders of the
// far field and the colleagues' sources, summed by xor shuffles
// (_evaluate_harmonic_kernel!, src/device.jl:186-206).
#include "ffmm_internal.h"

namespace ffmm {

constexpr int HB = 64;
constexpr int HT_CK = 128;   // sources per LDS tile of the near field
constexpr int HT_CK32 = 64;  // the same in the mixed form (twice the workgroups per CU)
constexpr double HC4PI = 0.07957747154594767;    // 1 / (4 pi)
constexpr double HC8PI = 0.039788735772973836;   // 1 / (8 pi)

__device__ __forceinline__ int tri(int n, int m) { return n * (n + 1) / 2 + m; }

// f(n, Re, Im) of the Schmidt regular harmonics of order m, n = m..nmax, at (x, y, z);
// the sectoral start is the product of m factors.  T is the arithmetic type (FP64, or
// FP32 in the mixed form's evaluation, with the tables rounded to FP32).
template <class T, class F>
__device__ __forceinline__ void chain(const HTabT<T> &tb, T x, T y, T z, int m, int nmax, F &&f) {
  if (m > nmax) return;
  const T r2 = x * x + y * y + z * z;
  T sr = 1, si = 0;
  for (int k = 1; k <= m; k++) {
    const T e = tb.e[k - 1];
    const T tr = -e * (x * sr - y * si), ti = -e * (x * si + y * sr);
    sr = tr;
    si = ti;
  }
  f(m, sr, si);
  T pr = 0, pi = 0, cr = sr, ci = si;
  for (int n = m + 1; n <= nmax; n++) {
    const int t = tri(n, m);
    const T az = tb.a[t] * z, br = tb.b[t] * r2;
    const T nr = az * cr - br * pr, ni = az * ci - br * pi;
    f(n, nr, ni);
    pr = cr;
    pi = ci;
    cr = nr;
    ci = ni;
  }
}

// u_n(z) for n = 0..nmax from z^2: Miller's backward recurrence normalised with
// u_0 = sinh z / z (src/operators/harmonics_yukawa.jl:43-63, 275-297)
__device__ void yukawa_u(double z2, int nmax, double *u) {
  if (z2 < 1e-8) {
    for (int n = 0; n <= nmax; n++) u[n] = 1 + z2 / (2 * (2 * n + 3));
    return;
  }
  const double z = sqrt(z2);
  const int cz = (int)ceil(z);
  const int N = (nmax > cz ? nmax : cz) + 25;
  double a = 0.0, b = 1.0;
  for (int n = N - 1; n >= 0; n--) {
    const double c = b + z2 * a / ((2.0 * n + 3) * (2.0 * n + 5));
    if (n <= nmax) u[n] = c;
    a = b;
    b = c;
  }
  const double s = (sinh(z) / z) / b;
  for (int n = 0; n <= nmax; n++) u[n] *= s;
}

// the LDS index of the real (im = 0) or imaginary (im = 1) part of the harmonic of
// degree n, order m, in a row of (p + 1)^2 reals (the multipole layout: n^2 for m = 0,
// n^2 + 2m - 1 and n^2 + 2m for m >= 1); the imaginary part of order 0 maps to the
// row's zero slot z
__device__ __forceinline__ int hidx(int n, int m, int im, int z) {
  return m == 0 ? (im ? z : n * n) : n * n + 2 * m - 1 + im;
}

// S2M in tiles of S sources of the node. Phase 1: the LS = 64 / S lanes of a source
// split its orders in pairs (m, p - m), whose chains are about p + 2 steps together,
// and store the regular harmonics (Yukawa: times u_n) in the source's row of the
// tile Y. Phase 2: lane t folds the tile into the entries e = t, t + 64, ... that it
// owns, in registers, and keeps them in an LDS column between tiles:
//   charges, stokeslets: M_b(e) += w_b conj(Y(e)) per block b;
//   Laplace dipoles and stresslet dipoles d_k (src/operators/harmonics.jl:711-752,
//   one block, or four): entry (n, m) takes
//   Y of degree n - 1 and orders m - 1, m, m + 1 against the components of d_k,
//   M_k(e) += a0 d_k0 + a1 d_k1 + a2 d_k2.
// KIND: 0 Laplace3D charges (nd = 1) or charges + dipoles (nd = 4), 1 Yukawa3D charges,
// 2 Stokeslet3D (nd = 3 or 9)
template <int KIND>
__global__ void __launch_bounds__(HB) k_s2m_harm(const int32_t *__restrict__ nodes, const double *__restrict__ centers,
                                                 int64_t nn, const int32_t *__restrict__ level,
                                                 const double *__restrict__ hscale, const double *__restrict__ khat,
                                                 const int32_t *__restrict__ plo, const int32_t *__restrict__ phi,
                                                 const double *__restrict__ sx, const double *__restrict__ sy,
                                                 const double *__restrict__ sz, const double *__restrict__ dens, int nd,
                                                 HTab tb, int p, int S, double *__restrict__ Mp) {
  extern __shared__ double sm[];
  const int N = (p + 1) * (p + 1), NY = N + 1;   // a row of the tile ends in a zero slot
  constexpr int nb = KIND == 2 ? 4 : 1;
  // per source: w_0..3, then d_k (4 x 3); Laplace: the charge, then the dipole at 4..6
  constexpr int NW = KIND == 2 ? 16 : KIND == 0 ? 8 : 1;
  double *acc = sm;                               // nb x N
  double *Y = acc + nb * N;                       // S x NY
  double *W = Y + (size_t)S * NY;                 // S x NW
  double *U = W + S * NW;                         // S x (p + 2), Yukawa radial weights
  const int t = threadIdx.x, LS = HB / S, sl = t / LS, gl = t % LS, ng = p / 2 + 1;
  const int node = nodes[blockIdx.x];
  const double c0 = centers[node], c1 = centers[nn + node], c2 = centers[2 * nn + node];
  const int lv = level[node];
  const double s = hscale[lv];
  const int32_t j0 = plo[node], j1 = phi[node];
  const bool dip = (KIND == 2 && nd == 9) || (KIND == 0 && nd == 4);
  for (int i = t; i < nb * N; i += HB) acc[i] = 0.0;
  for (int i = t; i < S; i += HB) Y[(size_t)i * NY + N] = 0.0;
  for (int32_t jb = j0; jb <= j1; jb += S) {
    const int ns = min(S, (int)(j1 - jb + 1));
    const int32_t j = jb + sl;
    const bool valid = sl < ns;
    const double x = valid ? (sx[j] - c0) * s : 0.0;
    const double y = valid ? (sy[j] - c1) * s : 0.0;
    const double z = valid ? (sz[j] - c2) * s : 0.0;
    if (valid && gl == 0) {
      double *w = W + sl * NW;
      if (KIND == 2) {
        const double *f = dens + (int64_t)nd * j;
        w[0] = f[0];
        w[1] = f[1];
        w[2] = f[2];
        w[3] = x * f[0] + y * f[1] + z * f[2];
        if (dip) {
          // the stresslet (g, nu): Phi_k from Dipole(-(g_k nu + nu_k g) s), Psi from
          // Dipole(-((nu.y) g + (g.y) nu) s) (src/operators/harmonics_stokes.jl:204-211)
          const double g[3] = {f[3], f[4], f[5]}, nu[3] = {f[6], f[7], f[8]};
          const double nuy = nu[0] * x + nu[1] * y + nu[2] * z;
          const double gy = g[0] * x + g[1] * y + g[2] * z;
#pragma unroll
          for (int c = 0; c < 3; c++) {
#pragma unroll
            for (int k = 0; k < 3; k++) w[4 + 3 * k + c] = -(g[k] * nu[c] + nu[k] * g[c]) * s;
            w[13 + c] = -(nuy * g[c] + gy * nu[c]) * s;
          }
        }
      } else {
        w[0] = dens[(int64_t)nd * j];
        if (KIND == 0 && nd == 4) {
          // a dipole d: Dipole(d s) (_scaled_density, src/operators/harmonics.jl:436)
#pragma unroll
          for (int c = 0; c < 3; c++) w[4 + c] = dens[4 * (int64_t)j + 1 + c] * s;
        }
        if (KIND == 1) {
          const double kh = khat[lv];
          yukawa_u(kh * kh * (x * x + y * y + z * z), p + 1, U + sl * (p + 2));
        }
      }
    }
    if (KIND == 1) __syncthreads();
    if (valid) {
      double *row = Y + (size_t)sl * NY;
      const double *u = U + sl * (p + 2);
      for (int g = gl; g < ng; g += LS)
        for (int which = 0; which < 2; which++) {
          const int m = which == 0 ? g : p - g;
          if (which == 1 && m == g) break;
          chain(tb, x, y, z, m, p, [&](int n, double vr, double vi) {
            if (KIND == 1) {
              vr *= u[n];
              vi *= u[n];
            }
            if (m == 0) {
              row[n * n] = vr;
            } else {
              row[n * n + 2 * m - 1] = vr;
              row[n * n + 2 * m] = vi;
            }
          });
        }
    }
    __syncthreads();
    for (int e = t; e < N; e += HB) {
      const int n = (int)sqrt((double)e), r = e - n * n, m = (r + 1) / 2, im = r > 0 && (r & 1) == 0;
      const double sg = im ? -1.0 : 1.0;
      double a[nb];
#pragma unroll
      for (int b = 0; b < nb; b++) a[b] = acc[b * N + e];
      for (int q = 0; q < ns; q++) {
        const double v = sg * Y[(size_t)q * NY + e];
#pragma unroll
        for (int b = 0; b < nb; b++) a[b] += W[q * NW + b] * v;
      }
      if (dip) {
        // entry (n, m) from Y_{n-1}^{m-1} (A, m >= 1), Y_{n-1}^m (B, n >= m + 1) and
        // Y_{n-1}^{m+1} (C, n >= m + 2), indices into a row and their weights
        const int z0 = N;
        double kA = m >= 1 ? tb.cm[tri(n, m)] / 2 : 0.0;
        const double kB = n >= m + 1 ? tb.cz[tri(n, m)] : 0.0;
        const double kC = n >= m + 2 ? (m >= 1 ? tb.cp[tri(n, m)] / 2 : tb.cp[tri(n, 0)]) : 0.0;
        const int n1 = n - 1;
        const int Ar = m >= 1 ? hidx(n1, m - 1, 0, z0) : z0, Ai = m >= 1 ? hidx(n1, m - 1, 1, z0) : z0;
        const int Br = n >= m + 1 ? hidx(n1, m, 0, z0) : z0, Bi = n >= m + 1 ? hidx(n1, m, 1, z0) : z0;
        const int Cr = n >= m + 2 ? hidx(n1, m + 1, 0, z0) : z0, Ci = n >= m + 2 ? hidx(n1, m + 1, 1, z0) : z0;
        // re: a0 = -kA Ar + kC Cr, a1 = kA Ai + kC Ci, a2 = kB Br
        // im: a0 = kA Ai - kC Ci, a1 = kA Ar + kC Cr, a2 = -kB Bi
        const int i00 = im ? Ai : Ar, i01 = im ? Ci : Cr, i10 = im ? Ar : Ai, i11 = im ? Cr : Ci, i2 = im ? Bi : Br;
        const double k00 = im ? kA : -kA, k01 = im ? -kC : kC, k10 = kA, k11 = kC, k2 = im ? -kB : kB;
        for (int q = 0; q < ns; q++) {
          const double *yr = Y + (size_t)q * NY;
          const double a0 = k00 * yr[i00] + k01 * yr[i01];
          const double a1 = k10 * yr[i10] + k11 * yr[i11];
          const double a2 = k2 * yr[i2];
          const double *d = W + q * NW + 4;
#pragma unroll
          for (int k = 0; k < nb; k++) a[k] += a0 * d[3 * k] + a1 * d[3 * k + 1] + a2 * d[3 * k + 2];
        }
      }
#pragma unroll
      for (int b = 0; b < nb; b++) acc[b * N + e] = a[b];
    }
    __syncthreads();
  }
  for (int i = t; i < nb * N; i += HB) Mp[(int64_t)node * nb * N + i] = acc[i];
}

// ---------------------------------------------------------------------------
// evaluation
// ---------------------------------------------------------------------------
template <class T>
__device__ __forceinline__ void get_entry(const T *L, int n, int m, T &re, T &im) {
  if (m > n) {
    re = 0;
    im = 0;
  } else if (m == 0) {
    re = L[n * n];
    im = 0;
  } else {
    re = L[n * n + 2 * m - 1];
    im = L[n * n + 2 * m];
  }
}
// complex coefficient (n, m) of a block, any sign of m, zero outside the triangle
template <class T>
__device__ __forceinline__ void cget(const T *L, int p, int n, int m, T &re, T &im) {
  if (n > p || abs(m) > n) {
    re = 0;
    im = 0;
  } else if (m == 0) {
    re = L[n * n];
    im = 0;
  } else if (m > 0) {
    re = L[n * n + 2 * m - 1];
    im = L[n * n + 2 * m];
  } else {
    const int k = -m;
    const T sg = (k & 1) ? T(-1) : T(1);
    re = sg * L[n * n + 2 * k - 1];
    im = -sg * L[n * n + 2 * k];
  }
}
template <class T>
__device__ __forceinline__ T hcz(const HTabT<T> &tb, int n, int m) { return abs(m) > n ? T(0) : tb.cz[tri(n, abs(m))]; }
template <class T>
__device__ __forceinline__ T hcm(const HTabT<T> &tb, int n, int m) {
  if (abs(m) > n) return 0;
  return m >= 0 ? tb.cm[tri(n, m)] : tb.cp[tri(n, -m)];
}
template <class T>
__device__ __forceinline__ T hcp(const HTabT<T> &tb, int n, int m) {
  if (abs(m) > n) return 0;
  return m >= 0 ? tb.cp[tri(n, m)] : tb.cm[tri(n, -m)];
}
// the scaled-gradient coefficient sets at (n1, m1) of a block: _dcoef
// (src/operators/harmonics_stokes.jl:243-251); out = (x re, x im, y re, y im, z re, z im)
template <class T>
__device__ __forceinline__ void dcoef(const T *L, int p, const HTabT<T> &tb, int n1, int m1, T o[6]) {
  const int n = n1 + 1;
  T ar, ai, br, bi, zr, zi;
  cget(L, p, n, m1 + 1, ar, ai);
  cget(L, p, n, m1 - 1, br, bi);
  cget(L, p, n, m1, zr, zi);
  const T cmv = -hcm(tb, n, m1 + 1), cpv = hcp(tb, n, m1 - 1), czv = hcz(tb, n, m1);
  const T dmr = cmv * ar, dmi = cmv * ai, dpr = cpv * br, dpi = cpv * bi;
  o[0] = (dmr + dpr) / 2;
  o[1] = (dmi + dpi) / 2;
  o[2] = -(dpi - dmi) / 2;
  o[3] = (dpr - dmr) / 2;
  o[4] = czv * zr;
  o[5] = czv * zi;
}

// near field of one source j at d = x - s_j (self excluded by the select)
template <int KIND, int OUT>
__device__ __forceinline__ void near_term(double d0, double d1, double d2, const double *f, int nd, bool keep,
                                          double kappa, double acc[4]) {
  const double r2 = fma(d2, d2, fma(d1, d1, d0 * d0));
  const double ir = rsqrt_f64(r2);
  if (KIND == 2 && OUT == FFMM_PRESSURE) {
    // pressure_kernel . f = (d.f) / (4 pi r^3) (src/kernels.jl:267-271); a stresslet
    // (g, nu): ((g.nu) / r^3 - 3 (d.g)(d.nu) / r^5) / (2 pi) (src/sources.jl:237-242)
    const double ir3 = ir * ir * ir;
    double pr = 2 * HC8PI * ir3 * (d0 * f[0] + d1 * f[1] + d2 * f[2]);
    if (nd == 9) {
      const double gn = f[3] * f[6] + f[4] * f[7] + f[5] * f[8];
      const double dg = d0 * f[3] + d1 * f[4] + d2 * f[5], dn = d0 * f[6] + d1 * f[7] + d2 * f[8];
      pr += 4 * HC8PI * (gn * ir3 - 3 * dg * dn * ir3 * ir * ir);
    }
    acc[0] += keep ? pr : 0.0;
    return;
  }
  if (KIND == 2) {
    const double a = HC8PI * ir;
    const double b = a * ir * ir * (d0 * f[0] + d1 * f[1] + d2 * f[2]);
    double u0 = a * f[0] + b * d0, u1 = a * f[1] + b * d1, u2 = a * f[2] + b * d2;
    if (nd == 9) {
      const double ir2 = ir * ir;
      const double e = -(6 * HC8PI) * (d0 * f[3] + d1 * f[4] + d2 * f[5]) * (d0 * f[6] + d1 * f[7] + d2 * f[8]) *
                       (ir2 * ir2 * ir);
      u0 += e * d0;
      u1 += e * d1;
      u2 += e * d2;
    }
    acc[0] += keep ? u0 : 0.0;
    acc[1] += keep ? u1 : 0.0;
    acc[2] += keep ? u2 : 0.0;
    return;
  }
  const double q = f[0];
  double v, dv;
  if (KIND == 0 && nd == 4) {
    // a dipole d (source_field and target_field of Dipole, src/sources.jl:138, 151):
    // potential d.r / (4 pi r^3), gradient (d / r^3 - 3 (d.r) r / r^5) / (4 pi)
    const double ir3 = ir * ir * ir, dr = f[1] * d0 + f[2] * d1 + f[3] * d2;
    const double gs = -3.0 * HC4PI * dr * ir3 * ir * ir;
    const double pv = HC4PI * dr * ir3;
    const double g0 = HC4PI * f[1] * ir3 + gs * d0, g1 = HC4PI * f[2] * ir3 + gs * d1,
                 g2 = HC4PI * f[3] * ir3 + gs * d2;
    if (OUT == FFMM_POTENTIAL) {
      acc[0] += keep ? pv : 0.0;
    } else if (OUT == FFMM_GRADIENT) {
      acc[0] += keep ? g0 : 0.0;
      acc[1] += keep ? g1 : 0.0;
      acc[2] += keep ? g2 : 0.0;
    } else {
      acc[0] += keep ? pv : 0.0;
      acc[1] += keep ? g0 : 0.0;
      acc[2] += keep ? g1 : 0.0;
      acc[3] += keep ? g2 : 0.0;
    }
  }
  if (KIND == 0) {
    v = HC4PI * ir;
    dv = -HC4PI * ir * ir * ir;
  } else {
    const double e = exp(-kappa * (r2 * ir));
    v = HC4PI * e * ir;
    dv = -HC4PI * e * (kappa * ir + ir * ir) * ir;
  }
  if (OUT == FFMM_POTENTIAL) {
    acc[0] += keep ? v * q : 0.0;
  } else if (OUT == FFMM_GRADIENT) {
    acc[0] += keep ? dv * d0 * q : 0.0;
    acc[1] += keep ? dv * d1 * q : 0.0;
    acc[2] += keep ? dv * d2 * q : 0.0;
  } else {
    acc[0] += keep ? v * q : 0.0;
    acc[1] += keep ? dv * d0 * q : 0.0;
    acc[2] += keep ? dv * d1 * q : 0.0;
    acc[3] += keep ? dv * d2 * q : 0.0;
  }
}

// a near-field source of the mixed form in LDS: FP64 position (the pair difference is
// formed in FP64), the densities in FP32 except the vectors whose dot product with the
// pair difference may cancel (a stresslet's normal nu, a dipole): on a surface, d.nu of
// two near points on it is O(r^2) where d and nu are O(1), and their FP32 rounding
// would leave an error O(r eps) in it; those stay FP64 and their dot product is formed
// in FP64.  Its index for the self-exclusion.
template <int KIND>
struct HNearSrc32 {   // Yukawa3D: a charge
  double x, y, z;
  float f[1];
  int32_t id;
};
template <>
struct HNearSrc32<0> {   // Laplace3D: a charge, a dipole
  double x, y, z;
  double dp[3];
  float f[1];
  int32_t id;
};
template <>
struct HNearSrc32<2> {   // Stokeslet3D: the stokeslet f (0..2), the stresslet g (3..5) and nu
  double x, y, z;
  double nu[3];
  float f[6];
  int32_t id;
};

// The mixed form's near field of one source e at the FP64 difference (dx, dy, dz),
// rounded to FP32; 1/r is v_rsq_f32, and the target's own source enters with r^2 = inf,
// so that 1/r = 0 and its terms vanish.  Component c of the output sums into s[c]; the
// potential and the Stokes velocity and pressure with Kahan compensation (the running
// compensation in k[c]), gradients with plain FP32 sums.
template <bool KAHAN>
__device__ __forceinline__ void hsum(float &s, float &k, float y) {
  if (KAHAN) {
    const float v = y - k;
    const float t = s + v;
    k = (t - s) - v;
    s = t;
  } else {
    s += y;
  }
}

template <int KIND, int OUT>
__device__ __forceinline__ void near_term32(double dx, double dy, double dz, const HNearSrc32<KIND> &e, int nd,
                                            bool keep, float kappa, float s[4], float k[4]) {
  constexpr float C4 = (float)HC4PI, C8 = (float)HC8PI;
  const float d0 = (float)dx, d1 = (float)dy, d2 = (float)dz;
  const float r2 = fmaf(d2, d2, fmaf(d1, d1, d0 * d0));
  const float ir = __builtin_amdgcn_rsqf(keep ? r2 : INFINITY);
  if constexpr (KIND == 2) {
    const float *f = e.f;
    const float df = d0 * f[0] + d1 * f[1] + d2 * f[2];
    float dg = 0.f, dn = 0.f;
    if (nd == 9) {
      dg = d0 * f[3] + d1 * f[4] + d2 * f[5];
      dn = (float)fma(dz, e.nu[2], fma(dy, e.nu[1], dx * e.nu[0]));
    }
    if (OUT == FFMM_PRESSURE) {
      // (d.f) / (4 pi r^3); a stresslet (g, nu): ((g.nu) / r^3 - 3 (d.g)(d.nu) / r^5) / (2 pi)
      // (src/kernels.jl:267-271, src/sources.jl:237-242)
      const float ir3 = ir * ir * ir;
      float pr = (2 * C8) * ir3 * df;
      if (nd == 9) {
        const float gn = f[3] * (float)e.nu[0] + f[4] * (float)e.nu[1] + f[5] * (float)e.nu[2];
        pr += (4 * C8) * (gn * ir3 - 3 * dg * dn * ir3 * ir * ir);
      }
      hsum<true>(s[0], k[0], pr);
    } else {
      // stokeslet (f.d) d / r^3 + f / r, stresslet -6 (d.g)(d.nu) d / r^5, over 8 pi
      const float a = C8 * ir, ir2 = ir * ir;
      float b = a * ir2 * df;
      if (nd == 9) b -= (6 * C8) * dg * dn * (ir2 * ir2 * ir);
      hsum<true>(s[0], k[0], fmaf(a, f[0], b * d0));
      hsum<true>(s[1], k[1], fmaf(a, f[1], b * d1));
      hsum<true>(s[2], k[2], fmaf(a, f[2], b * d2));
    }
  } else {
    float pv = 0.f, g0 = 0.f, g1 = 0.f, g2 = 0.f;
    if constexpr (KIND == 0) {
      if (nd == 4) {
        // a dipole d: potential d.r / (4 pi r^3), gradient (d / r^3 - 3 (d.r) r / r^5) / (4 pi)
        const float ir3 = ir * ir * ir;
        const float dr = (float)fma(dz, e.dp[2], fma(dy, e.dp[1], dx * e.dp[0]));
        const float gs = -3.f * C4 * dr * ir3 * ir * ir;
        pv = C4 * dr * ir3;
        g0 = C4 * (float)e.dp[0] * ir3 + gs * d0;
        g1 = C4 * (float)e.dp[1] * ir3 + gs * d1;
        g2 = C4 * (float)e.dp[2] * ir3 + gs * d2;
      }
    }
    float v, dv;
    if (KIND == 0) {
      v = C4 * ir;
      dv = -C4 * ir * ir * ir;
    } else {
      const float ex = expf(-kappa * (r2 * ir));
      v = C4 * ex * ir;
      dv = -C4 * ex * (kappa * ir + ir * ir) * ir;
    }
    const float q = e.f[0], dq = dv * q;
    if (OUT != FFMM_GRADIENT) hsum<true>(s[0], k[0], fmaf(v, q, pv));
    if (OUT != FFMM_POTENTIAL) {
      const int o = OUT == FFMM_GRADIENT ? 0 : 1;
      s[o] += fmaf(dq, d0, g0);
      s[o + 1] += fmaf(dq, d1, g1);
      s[o + 2] += fmaf(dq, d2, g2);
    }
  }
}

// one source of the near field into the mixed form's tile
template <int KIND>
__device__ __forceinline__ HNearSrc32<KIND> near_src32(const EvalArgs &a, int32_t q) {
  HNearSrc32<KIND> e;
  e.x = a.sx[q];
  e.y = a.sy[q];
  e.z = a.sz[q];
  const double *t = a.tau + (int64_t)q * a.nd;
  if constexpr (KIND == 2) {
#pragma unroll
    for (int c = 0; c < 3; c++) e.f[c] = (float)t[c];
    const bool st = a.nd == 9;
#pragma unroll
    for (int c = 0; c < 3; c++) {
      e.f[3 + c] = st ? (float)t[3 + c] : 0.f;
      e.nu[c] = st ? t[6 + c] : 0.0;
    }
  } else if constexpr (KIND == 0) {
    e.f[0] = (float)t[0];
    const bool dip = a.nd == 4;
#pragma unroll
    for (int c = 0; c < 3; c++) e.dp[c] = dip ? t[1 + c] : 0.0;
  } else {
    e.f[0] = (float)t[0];
  }
  e.id = q;
  return e;
}

template <class T>
__device__ __forceinline__ HTabT<T> eval_tables(const EvalArgs &a);
template <>
__device__ __forceinline__ HTab eval_tables<double>(const EvalArgs &a) { return a.tb; }
template <>
__device__ __forceinline__ HTab32 eval_tables<float>(const EvalArgs &a) { return a.tb32; }

// T = double: the FP64 form.  T = float: the mixed form; the far field (the folds of the
// leaf's local expansion, rounded to FP32 in LDS) and the near field run in FP32, the
// target's offset from the leaf centre rounded once after its FP64 scaling, and the
// near field's FP32 partial sums are folded into FP64 accumulators after every LDS tile
// (the FP32 evaluation of the mixed form, src/device.jl:186-206 with Float32 locals and
// densities).
template <class T, int KIND, int OUT>
__global__ void __launch_bounds__(HB) k_eval_harm(EvalArgs a) {
  extern __shared__ double sm[];
  constexpr bool F32 = sizeof(T) == 4;
  // FP64: doubles per tile entry, position and density components (Stokes 9, Laplace a
  // charge and a dipole, Yukawa a charge); FP32: an HNearSrc32 with ND densities
  constexpr int TW = KIND == 2 ? 12 : KIND == 0 ? 7 : 4;
  constexpr size_t TILE_BYTES = F32 ? HT_CK32 * sizeof(HNearSrc32<KIND>) : HT_CK * TW * sizeof(double);
  __shared__ __align__(16) char tile_raw[TILE_BYTES];
  double *tile = reinterpret_cast<double *>(tile_raw);
  HNearSrc32<KIND> *tile32 = reinterpret_cast<HNearSrc32<KIND> *>(tile_raw);
  const int s = blockIdx.x;
  const int i = a.leaf_nodes[s];
  const bool self = a.tptr == nullptr;
  const int32_t t0 = self ? a.plo[i] : a.tptr[s];
  const int32_t t1 = self ? a.phi[i] + 1 : a.tptr[s + 1];
  if (t0 >= t1) return;
  const int p = a.p, N = (p + 1) * (p + 1);
  constexpr int nb = KIND == 2 ? 4 : 1;
  const int NB = nb * N;
  const HTabT<T> tb = eval_tables<T>(a);
  T *L = reinterpret_cast<T *>(sm);                   // the leaf's local expansion
  double *uw = sm + (NB * sizeof(T) + 7) / 8;         // per lane: radial weights to p (Yukawa potential)
  double *ug = uw + HB * (p + 2);                     // per lane: radial weights to p + 1 (Yukawa gradient)
  const double *Lg = a.Lc + (int64_t)s * NB;
  for (int k = threadIdx.x; k < NB; k += HB) L[k] = (T)Lg[k];
  __syncthreads();
  const double c0 = a.centers[i], c1 = a.centers[a.nn + i], c2 = a.centers[2 * a.nn + i];
  const int lv = a.level[i];
  const double sc = a.hscale[lv];
  const double kh = KIND == 1 ? a.khat[lv] : 0.0;
  __shared__ NearList nl;
  __shared__ int32_t tile_id[HT_CK];
  near_list(a, i, nl);
  const int32_t nsrc = nl.pre[NCOLL];
  for (int32_t tbk = t0; tbk < t1; tbk += HB) {
    const int cnt = min(HB, t1 - tbk);
    int ntp = 1;
    while (ntp < cnt) ntp <<= 1;
    const int R = HB / ntp;
    const int tl = (int)threadIdx.x % ntp, rep = (int)threadIdx.x / ntp;
    const int32_t j = tbk + tl;
    const bool valid = tl < cnt;
    const int32_t jj = valid ? j : t0;
    const double x = self ? a.sx[jj] : a.tx[jj];
    const double y = self ? a.sy[jj] : a.ty[jj];
    const double z = self ? a.sz[jj] : a.tz[jj];
    const double xh = (x - c0) * sc, yh = (y - c1) * sc, zh = (z - c2) * sc;
    const T xt = (T)xh, yt = (T)yh, zt = (T)zh;
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    if (KIND == 2) {
      // velocity 1/2 [Phi - G x + grad Psi] (_stokes_fold1 / _stokes_output,
      // src/operators/harmonics_stokes.jl:253-301, 321-325), orders split over replicas
      T Phi[3] = {0, 0, 0}, G[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}}, gP[3] = {0, 0, 0};
      for (int m = rep; m <= p; m += R)
        chain(tb, xt, yt, zt, m, p, [&](int n, T vr, T vi) {
          const T wt = m == 0 ? T(1) : T(2);
#pragma unroll
          for (int k = 0; k < 3; k++) {
            T cr, ci;
            cget(L + k * N, p, n, m, cr, ci);
            Phi[k] += wt * (cr * vr + ci * vi);
          }
          if (n < p) {
            T o[6];
#pragma unroll
            for (int k = 0; k < 3; k++) {
              dcoef(L + k * N, p, tb, n, m, o);
#pragma unroll
              for (int c = 0; c < 3; c++) G[c][k] += wt * (o[2 * c] * vr + o[2 * c + 1] * vi);
            }
            dcoef(L + 3 * N, p, tb, n, m, o);
#pragma unroll
            for (int c = 0; c < 3; c++) gP[c] += wt * (o[2 * c] * vr + o[2 * c + 1] * vi);
          }
        });
      if (OUT == FFMM_PRESSURE) {
        // -s tr(grad Phi) (_stokes_output(::Pressure), src/operators/harmonics_stokes.jl:348-351)
        acc[0] = -sc * (double)(G[0][0] + G[1][1] + G[2][2]);
      } else {
#pragma unroll
        for (int c = 0; c < 3; c++) {
          const T gx = G[c][0] * xt + G[c][1] * yt + G[c][2] * zt;
          acc[c] = (double)((Phi[c] - gx + gP[c]) / 2);
        }
      }
    } else {
      double *up = uw + threadIdx.x * (p + 2), *ugr = ug + threadIdx.x * (p + 2);
      if (KIND == 1) {
        const double z2 = kh * kh * (xh * xh + yh * yh + zh * zh);
        if (OUT != FFMM_GRADIENT) yukawa_u(z2, p, up);
        if (OUT != FFMM_POTENTIAL) yukawa_u(z2, p + 1, ugr);
      }
      if (OUT != FFMM_GRADIENT) {
        // _harmonic_potential (src/operators/harmonics.jl:455-462)
        T pot = 0;
        for (int m = rep; m <= p; m += R)
          chain(tb, xt, yt, zt, m, p, [&](int n, T vr, T vi) {
            T lr, li;
            get_entry(L, n, m, lr, li);
            const T wn = KIND == 1 ? (T)up[n] : T(1);
            pot += m == 0 ? wn * (lr * vr) : wn * (2 * (lr * vr + li * vi));
          });
        acc[0] = (double)pot;
      }
      if (OUT != FFMM_POTENTIAL) {
        // _harmonic_gradient (src/operators/harmonics.jl:463-485); Yukawa adds
        // x * (potential with the radial weights) (src/operators/harmonics_yukawa.jl:306-313)
        T g0 = 0, g1 = 0, g2 = 0, pr = 0;
        for (int m1 = rep; m1 <= p - 1; m1 += R)
          chain(tb, xt, yt, zt, m1, p - 1, [&](int n1, T vr, T vi) {
            const int n = n1 + 1;
            const T wn = KIND == 1 ? (T)ugr[n] : T(1);
            if (m1 == 0) {
              T lr, li;
              get_entry(L, n, 1, lr, li);
              const T c = tb.cm[tri(n, 1)];
              const T ax = -c * lr, ay = -c * li, az = tb.cz[tri(n, 0)] * L[n * n];
              g0 += wn * (ax * vr);
              g1 += wn * (ay * vr);
              g2 += wn * (az * vr);
            } else {
              T ur, ui, dr, di, zr, zi;
              get_entry(L, n, m1 + 1, ur, ui);
              get_entry(L, n, m1 - 1, dr, di);
              const T cu = tb.cm[tri(n, m1 + 1)] / 2, cd = tb.cp[tri(n, m1 - 1)] / 2;
              const T axr = -cu * ur + cd * dr, axi = -cu * ui + cd * di;
              const T ayr = -(cu * ui + cd * di), ayi = cu * ur + cd * dr;
              const T czv = tb.cz[tri(n, m1)];
              get_entry(L, n, m1, zr, zi);
              g0 += (2 * wn) * (axr * vr + axi * vi);
              g1 += (2 * wn) * (ayr * vr + ayi * vi);
              g2 += (2 * wn) * (czv * (zr * vr + zi * vi));
            }
          });
        if (KIND == 1) {
          for (int m = rep; m <= p; m += R)
            chain(tb, xt, yt, zt, m, p, [&](int n, T vr, T vi) {
              T lr, li;
              get_entry(L, n, m, lr, li);
              const T wn = (T)(kh * kh * ugr[n + 1] / (2 * n + 3));
              pr += m == 0 ? wn * (lr * vr) : wn * (2 * (lr * vr + li * vi));
            });
        }
        const int o = OUT == FFMM_GRADIENT ? 0 : 1;
        acc[o] = (double)(g0 + xt * pr) * sc;
        acc[o + 1] = (double)(g1 + yt * pr) * sc;
        acc[o + 2] = (double)(g2 + zt * pr) * sc;
      }
    }
    // the near field of the 27 colleagues (src/device.jl:89-100), their sources and
    // densities as one list staged through LDS tiles (broadcast reads)
    const int32_t skip = self ? j : -1;
    if constexpr (!F32) {
      for (int32_t q0 = 0; q0 < nsrc; q0 += HT_CK) {
        const int nq = min(HT_CK, nsrc - q0);
        __syncthreads();
        for (int k = threadIdx.x; k < nq; k += HB) {
          const int32_t q = near_source(nl, q0 + k);
          double *e = tile + k * TW;
          e[0] = a.sx[q];
          e[1] = a.sy[q];
          e[2] = a.sz[q];
          for (int c = 0; c < a.nd; c++) e[3 + c] = a.tau[(int64_t)q * a.nd + c];
          tile_id[k] = q;
        }
        __syncthreads();
        for (int k = rep; k < nq; k += R) {
          const double *e = tile + k * TW;
          near_term<KIND, OUT>(x - e[0], y - e[1], z - e[2], e + 3, a.nd, tile_id[k] != skip, a.kappa, acc);
        }
      }
    } else {
      const float kap = (float)a.kappa;
      float ks[4] = {0.f, 0.f, 0.f, 0.f}, kc[4] = {0.f, 0.f, 0.f, 0.f};
      for (int32_t q0 = 0; q0 < nsrc; q0 += HT_CK32) {
        const int nq = min(HT_CK32, nsrc - q0);
        __syncthreads();
        for (int k = threadIdx.x; k < nq; k += HB) tile32[k] = near_src32<KIND>(a, near_source(nl, q0 + k));
        __syncthreads();
        for (int k = rep; k < nq; k += R) {
          const HNearSrc32<KIND> e = tile32[k];
          near_term32<KIND, OUT>(x - e.x, y - e.y, z - e.z, e, a.nd, e.id != skip, kap, ks, kc);
        }
#pragma unroll
        for (int c = 0; c < 4; c++) {
          acc[c] += (double)ks[c] - (double)kc[c];
          ks[c] = kc[c] = 0.f;
        }
      }
    }
    for (int off = ntp; off < HB; off <<= 1)
#pragma unroll
      for (int c = 0; c < 4; c++) acc[c] += __shfl_xor(acc[c], off, HB);
    if (valid && rep == 0) {
      const int64_t o = a.outidx[j];
      if (KIND == 2 && OUT == FFMM_PRESSURE) {
        a.out[o] = acc[0];
      } else if (KIND == 2) {
        a.out[3 * o] = acc[0];
        a.out[3 * o + 1] = acc[1];
        a.out[3 * o + 2] = acc[2];
      } else if (OUT == FFMM_POTENTIAL) {
        a.out[o] = acc[0];
      } else if (OUT == FFMM_GRADIENT) {
        a.out[3 * o] = acc[0];
        a.out[3 * o + 1] = acc[1];
        a.out[3 * o + 2] = acc[2];
      } else {
        a.out[4 * o] = acc[0];
        a.out[4 * o + 1] = acc[1];
        a.out[4 * o + 2] = acc[2];
        a.out[4 * o + 3] = acc[3];
      }
    }
  }
}

void harmonic_s2m_level(Plan &plan, int l) {
  OpCache &o = *plan.ops;
  const int64_t nq = plan.s2m_off[l + 1] - plan.s2m_off[l];
  if (nq == 0) return;
  const int p = o.p, N = (p + 1) * (p + 1);
  // sources per tile: the most (a power of two, at most 64) whose harmonics fit 8 KiB
  int S = 64;
  while (S > 1 && (size_t)S * (N + 1) * sizeof(double) > 8192) S /= 2;
  const int nw = o.kernel == FFMM_KERNEL_STOKESLET ? 16 : o.kernel == FFMM_KERNEL_LAPLACE ? 8 : 1;
  const size_t lds = sizeof(double) * ((size_t)o.Ns + (size_t)S * (N + 1) + (size_t)S * nw +
                                       (o.kernel == FFMM_KERNEL_YUKAWA ? (size_t)S * (p + 2) : 0));
  const int32_t *nodes = plan.s2m_nodes.p + plan.s2m_off[l];
  hipStream_t s = plan.ctx->stream;
#define FFMM_S2M_ARGS                                                                                              \
  nodes, plan.centers.p, plan.nnodes, plan.level.p, plan.hscale.p, plan.khat.p, plan.plo.p, plan.phi.p, plan.sx.p, \
      plan.sy.p, plan.sz.p, plan.tau.p, plan.nd, o.tab(), p, S, plan.Mp.p
  if (o.kernel == FFMM_KERNEL_LAPLACE) k_s2m_harm<0><<<(unsigned)nq, HB, lds, s>>>(FFMM_S2M_ARGS);
  else if (o.kernel == FFMM_KERNEL_YUKAWA) k_s2m_harm<1><<<(unsigned)nq, HB, lds, s>>>(FFMM_S2M_ARGS);
  else k_s2m_harm<2><<<(unsigned)nq, HB, lds, s>>>(FFMM_S2M_ARGS);
#undef FFMM_S2M_ARGS
  FFMM_LAUNCH_CHECK();
}

template <class T>
static void launch_eval_harm(const OpCache &o, const EvalArgs &a, int output, unsigned g, hipStream_t s) {
  const size_t lds = ((size_t)o.Ns * sizeof(T) + 7) / 8 * 8 +
                     (o.kernel == FFMM_KERNEL_YUKAWA ? 2 * sizeof(double) * HB * (o.p + 2) : 0);
  if (o.kernel == FFMM_KERNEL_STOKESLET) {
    if (output == FFMM_PRESSURE) k_eval_harm<T, 2, FFMM_PRESSURE><<<g, HB, lds, s>>>(a);
    else k_eval_harm<T, 2, FFMM_POTENTIAL><<<g, HB, lds, s>>>(a);
  } else if (o.kernel == FFMM_KERNEL_LAPLACE) {
    if (output == FFMM_POTENTIAL) k_eval_harm<T, 0, FFMM_POTENTIAL><<<g, HB, lds, s>>>(a);
    else if (output == FFMM_GRADIENT) k_eval_harm<T, 0, FFMM_GRADIENT><<<g, HB, lds, s>>>(a);
    else k_eval_harm<T, 0, FFMM_POTENTIAL_GRADIENT><<<g, HB, lds, s>>>(a);
  } else {
    if (output == FFMM_POTENTIAL) k_eval_harm<T, 1, FFMM_POTENTIAL><<<g, HB, lds, s>>>(a);
    else if (output == FFMM_GRADIENT) k_eval_harm<T, 1, FFMM_GRADIENT><<<g, HB, lds, s>>>(a);
    else k_eval_harm<T, 1, FFMM_POTENTIAL_GRADIENT><<<g, HB, lds, s>>>(a);
  }
  FFMM_LAUNCH_CHECK();
}

// the mixed form evaluates in FP32 (k_eval_harm<float>), the FP64 form in FP64
void launch_eval_harmonic(Plan &plan, const EvalArgs &a, int output) {
  const int64_t nleaves = plan.leaf_off[plan.L];
  if (nleaves == 0) return;
  const unsigned g = (unsigned)nleaves;
  if (plan.precision == FFMM_PREC_MIXED) launch_eval_harm<float>(*plan.ops, a, output, g, plan.ctx->stream);
  else launch_eval_harm<double>(*plan.ops, a, output, g, plan.ctx->stream);
}

}  // namespace ffmm
