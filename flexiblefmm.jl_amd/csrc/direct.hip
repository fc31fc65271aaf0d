// FP64 direct sums on the device: direct! and direct_self! (src/kernels.jl:442-487),
// out[j] = sum_i target_field(K, s_i, t_j, tau_i, output), the source skip[j] excluded
// (a masked select).  One work-item per target; the sources stream through LDS in
// tiles shared by the workgroup.  The sources are split into chunks over the second
// grid axis so that a few thousand targets still fill the chip; the chunks' partial
// sums are added in chunk order by a second pass (deterministic).
#include "ffmm_internal.h"

#include <algorithm>
#include <mutex>
#include <string>
#include <vector>

namespace ffmm {

constexpr int DT = 256;
constexpr double C4PI = 0.07957747154594767;   // 1 / (4 pi)

template <int OUT>
__global__ void __launch_bounds__(DT) k_direct(const double *__restrict__ src, const double *__restrict__ tau, int64_t ns,
                                               const double *__restrict__ tgt, int64_t nt,
                                               const int64_t *__restrict__ skip, int64_t chunk,
                                               double *__restrict__ part) {
  __shared__ double xs[DT], ys[DT], zs[DT], ts[DT];
  const int64_t j = (int64_t)blockIdx.x * DT + threadIdx.x;
  const int64_t jj = j < nt ? j : nt - 1;
  const double x = tgt[3 * jj], y = tgt[3 * jj + 1], z = tgt[3 * jj + 2];
  const int64_t sk = skip ? skip[jj] : -1;
  const int64_t ib = (int64_t)blockIdx.y * chunk, ie = std::min<int64_t>(ns, ib + chunk);
  double acc[4] = {0, 0, 0, 0};
  for (int64_t i0 = ib; i0 < ie; i0 += DT) {
    const int64_t i = i0 + threadIdx.x;
    if (i < ie) {
      xs[threadIdx.x] = src[3 * i];
      ys[threadIdx.x] = src[3 * i + 1];
      zs[threadIdx.x] = src[3 * i + 2];
      ts[threadIdx.x] = tau[i];
    }
    __syncthreads();
    const int cnt = (int)std::min<int64_t>(DT, ie - i0);
    for (int q = 0; q < cnt; q++) {
      const double d0 = x - xs[q], d1 = y - ys[q], d2 = z - zs[q];
      const double ir = rsqrt_f64(fma(d2, d2, fma(d1, d1, d0 * d0)));
      const bool keep = i0 + q != sk;
      if (OUT != FFMM_GRADIENT) acc[0] += keep ? (C4PI * ir) * ts[q] : 0.0;
      if (OUT != FFMM_POTENTIAL) {
        const double a = -C4PI * ir * ir * ir;
        acc[1] += keep ? (a * d0) * ts[q] : 0.0;
        acc[2] += keep ? (a * d1) * ts[q] : 0.0;
        acc[3] += keep ? (a * d2) * ts[q] : 0.0;
      }
    }
    __syncthreads();
  }
  if (j >= nt) return;
  double *p = part + ((int64_t)blockIdx.y * nt + j) * 4;
  p[0] = acc[0]; p[1] = acc[1]; p[2] = acc[2]; p[3] = acc[3];
}

__global__ void k_direct_reduce(const double *__restrict__ part, int64_t nt, int nchunk, int output,
                                double *__restrict__ out) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nt) return;
  double acc[4] = {0, 0, 0, 0};
  for (int c = 0; c < nchunk; c++)
    for (int k = 0; k < 4; k++) acc[k] += part[((int64_t)c * nt + j) * 4 + k];
  if (output == FFMM_POTENTIAL) {
    out[j] = acc[0];
  } else if (output == FFMM_GRADIENT) {
    out[3 * j] = acc[1]; out[3 * j + 1] = acc[2]; out[3 * j + 2] = acc[3];
  } else {
    out[4 * j] = acc[0]; out[4 * j + 1] = acc[1]; out[4 * j + 2] = acc[2]; out[4 * j + 3] = acc[3];
  }
}

// the source chunks of a direct sum over ns sources at nt targets: about 4096 workgroups
// in all, chunks of at least 4096 sources; the number of chunks, *chunk their length
static int64_t direct_chunks(int64_t ns, int64_t nt, int64_t *chunk) {
  const unsigned gx = nblocks(nt, DT);
  int64_t nchunk = std::max<int64_t>(1, std::min<int64_t>(4096 / gx, (ns + 4095) / 4096));
  nchunk = std::min<int64_t>(nchunk, 65535);
  *chunk = std::max<int64_t>(1, (ns + nchunk - 1) / nchunk);
  return std::max<int64_t>(1, (ns + *chunk - 1) / *chunk);
}

// the two passes over a partial-sum buffer `part` of nchunk x nt x 4 doubles
static void direct_run(Context &ctx, const double *src, const double *tau, int64_t ns, const double *tgt,
                       int64_t nt, const int64_t *skip, double *out, int output, int64_t nchunk, int64_t chunk,
                       double *part) {
  dim3 grid(nblocks(nt, DT), (unsigned)nchunk);
  if (output == FFMM_POTENTIAL) k_direct<FFMM_POTENTIAL><<<grid, DT, 0, ctx.stream>>>(src, tau, ns, tgt, nt, skip, chunk, part);
  else if (output == FFMM_GRADIENT) k_direct<FFMM_GRADIENT><<<grid, DT, 0, ctx.stream>>>(src, tau, ns, tgt, nt, skip, chunk, part);
  else k_direct<2><<<grid, DT, 0, ctx.stream>>>(src, tau, ns, tgt, nt, skip, chunk, part);
  FFMM_LAUNCH_CHECK();
  k_direct_reduce<<<nblocks(nt, DT), DT, 0, ctx.stream>>>(part, nt, (int)nchunk, output, out);
  FFMM_LAUNCH_CHECK();
}

void direct(Context &ctx, const double *src, const double *tau, int64_t ns, const double *tgt, int64_t nt,
            const int64_t *skip, double *out, int output) {
  if (nt <= 0) return;
  int64_t chunk = 0;
  const int64_t nchunk = direct_chunks(ns, nt, &chunk);
  DBuf<double> part;
  part.resize((size_t)nchunk * nt * 4);
  direct_run(ctx, src, tau, ns, tgt, nt, skip, out, output, nchunk, chunk, part.p);
  ctx.sync();   // the partial sums are freed on return
}

// The partial sums of the allocation-free direct sum (direct_noalloc): one array per
// device, 2^22 doubles (32 MiB), held until the process ends.  Host code never passes
// the array's name to a kernel: each device's address of it is looked up once, when a
// context is set up on that device (direct_scratch_setup), and kept.
constexpr int64_t DIRECT_SCRATCH_DOUBLES = (int64_t)1 << 22;
static __device__ __align__(256) double scratch[DIRECT_SCRATCH_DOUBLES];

namespace {
std::mutex g_scratch_mu;
struct ScratchAddr {
  int device;
  double *ptr;
  std::string error;   // why the lookup failed (ptr null)
};
std::vector<ScratchAddr> g_scratch;
}  // namespace

void direct_scratch_setup(int device) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  for (const auto &a : g_scratch)
    if (a.device == device) return;
  void *p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess) {
    (void)hipGetLastError();
    g_scratch.push_back({device, nullptr, std::string("hipGetSymbolAddress of the direct-sum scratch array: ") +
                                              hipGetErrorString(e)});
    return;
  }
  g_scratch.push_back({device, (double *)p, std::string()});
}

static double *direct_scratch(int device) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  for (const auto &a : g_scratch)
    if (a.device == device) {
      if (!a.ptr) fail(FFMM_ERR_DEVICE, a.error);
      return a.ptr;
    }
  fail(FFMM_ERR_STATE, "no context was set up on device " + std::to_string(device));
  return nullptr;
}

void direct_noalloc(Context &ctx, const double *src, const double *tau, int64_t ns, const double *tgt, int64_t nt,
                    const int64_t *skip, double *out, int output) {
  if (nt <= 0) return;
  int64_t chunk = 0;
  const int64_t nchunk = direct_chunks(ns, nt, &chunk);
  if (nchunk * nt * 4 > DIRECT_SCRATCH_DOUBLES)
    fail(FFMM_ERR_ARGUMENT, "ffmm_direct_noalloc takes at most " + std::to_string(DIRECT_SCRATCH_DOUBLES / 4) +
                                " targets (its partial sums live in a fixed 32 MiB device array); got " +
                                std::to_string(nt));
  // every entry of the used part of the array is written by the first pass before the
  // second reads it: nothing to clear
  direct_run(ctx, src, tau, ns, tgt, nt, skip, out, output, nchunk, chunk, direct_scratch(ctx.device));
}

// The Stokes sums: velocity of stokeslets f (nd = 3) and, with nd = 9, stresslets of
// strength g and normal nu (rows [f, g, nu]):
//   u = 1/(8 pi) (f / r + (d . f) d / r^3) - 6/(8 pi) (d . g)(d . nu) d / r^5,  d = t - s,
// and the pressure of both
// (direct! for Stokeslet3D, src/kernels.jl:463-487; the stresslet as in
// src/operators/harmonics_stokes.jl:197-217).  The sources' rows stream through LDS.
constexpr double C8PI = 0.039788735772973836;   // 1 / (8 pi)

template <int ND>
__global__ void __launch_bounds__(DT) k_direct_stokes(const double *__restrict__ src, const double *__restrict__ dens,
                                                      int64_t ns, const double *__restrict__ tgt, int64_t nt,
                                                      const int64_t *__restrict__ skip, int64_t chunk,
                                                      double *__restrict__ part) {
  __shared__ double xs[DT][3 + ND];
  const int64_t j = (int64_t)blockIdx.x * DT + threadIdx.x;
  const int64_t jj = j < nt ? j : nt - 1;
  const double x = tgt[3 * jj], y = tgt[3 * jj + 1], z = tgt[3 * jj + 2];
  const int64_t sk = skip ? skip[jj] : -1;
  const int64_t ib = (int64_t)blockIdx.y * chunk, ie = std::min<int64_t>(ns, ib + chunk);
  double acc[4] = {0, 0, 0, 0};
  for (int64_t i0 = ib; i0 < ie; i0 += DT) {
    const int64_t i = i0 + threadIdx.x;
    if (i < ie) {
      for (int c = 0; c < 3; c++) xs[threadIdx.x][c] = src[3 * i + c];
      for (int c = 0; c < ND; c++) xs[threadIdx.x][3 + c] = dens[ND * i + c];
    }
    __syncthreads();
    const int cnt = (int)std::min<int64_t>(DT, ie - i0);
    for (int q = 0; q < cnt; q++) {
      const double *v = xs[q];
      const double d0 = x - v[0], d1 = y - v[1], d2 = z - v[2];
      const double ir = rsqrt_f64(fma(d2, d2, fma(d1, d1, d0 * d0)));
      const double a = C8PI * ir;
      const double b = a * ir * ir * (d0 * v[3] + d1 * v[4] + d2 * v[5]);
      double u0 = a * v[3] + b * d0, u1 = a * v[4] + b * d1, u2 = a * v[5] + b * d2;
      if (ND == 9) {
        const double ir2 = ir * ir;
        const double e = -(6 * C8PI) * (d0 * v[6] + d1 * v[7] + d2 * v[8]) * (d0 * v[9] + d1 * v[10] + d2 * v[11]) *
                         (ir2 * ir2 * ir);
        u0 += e * d0;
        u1 += e * d1;
        u2 += e * d2;
      }
      // the pressure: (d.f) / (4 pi r^3) (pressure_kernel, src/kernels.jl:267-271); a
      // stresslet ((g.nu) / r^3 - 3 (d.g)(d.nu) / r^5) / (2 pi) (src/sources.jl:237-242)
      const double ir3 = ir * ir * ir;
      double pr = 2 * C8PI * ir3 * (d0 * v[3] + d1 * v[4] + d2 * v[5]);
      if (ND == 9) {
        const double gn = v[6] * v[9] + v[7] * v[10] + v[8] * v[11];
        const double dg = d0 * v[6] + d1 * v[7] + d2 * v[8], dn = d0 * v[9] + d1 * v[10] + d2 * v[11];
        pr += 4 * C8PI * (gn * ir3 - 3 * dg * dn * ir3 * ir * ir);
      }
      const bool keep = i0 + q != sk;
      acc[0] += keep ? u0 : 0.0;
      acc[1] += keep ? u1 : 0.0;
      acc[2] += keep ? u2 : 0.0;
      acc[3] += keep ? pr : 0.0;
    }
    __syncthreads();
  }
  if (j >= nt) return;
  double *p = part + ((int64_t)blockIdx.y * nt + j) * 4;
  p[0] = acc[3]; p[1] = acc[0]; p[2] = acc[1]; p[3] = acc[2];
}

void direct_stokes(Context &ctx, const double *src, const double *dens, int nd, int64_t ns, const double *tgt,
                   int64_t nt, const int64_t *skip, double *out, int output) {
  if (nt <= 0) return;
  const unsigned gx = nblocks(nt, DT);
  int64_t nchunk = std::max<int64_t>(1, std::min<int64_t>(4096 / gx, (ns + 4095) / 4096));
  nchunk = std::min<int64_t>(nchunk, 65535);
  const int64_t chunk = std::max<int64_t>(1, (ns + nchunk - 1) / nchunk);
  nchunk = std::max<int64_t>(1, (ns + chunk - 1) / chunk);
  DBuf<double> part;
  part.resize((size_t)nchunk * nt * 4);
  dim3 grid(gx, (unsigned)nchunk);
  if (nd == 3) k_direct_stokes<3><<<grid, DT, 0, ctx.stream>>>(src, dens, ns, tgt, nt, skip, chunk, part.p);
  else k_direct_stokes<9><<<grid, DT, 0, ctx.stream>>>(src, dens, ns, tgt, nt, skip, chunk, part.p);
  FFMM_LAUNCH_CHECK();
  // the velocity is the three gradient slots of the reduction, the pressure the first
  k_direct_reduce<<<nblocks(nt, DT), DT, 0, ctx.stream>>>(part.p, nt, (int)nchunk,
                                                          output == FFMM_PRESSURE ? FFMM_POTENTIAL : FFMM_GRADIENT, out);
  FFMM_LAUNCH_CHECK();
  ctx.sync();
}

}  // namespace ffmm
