// FP64 direct sums on the device: direct! and direct_self! (src/kernels.jl:442-487),
// out[j] = sum_i target_field(K, s_i, t_j, tau_i, output), the source skip[j] excluded
// (a masked select).  One work-item per target; the sources stream through LDS in
// tiles shared by the workgroup.
#include "ffmm_internal.h"

namespace ffmm {

constexpr int DT = 256;
constexpr double C4PI = 0.07957747154594767;   // 1 / (4 pi)

template <int OUT>
__global__ void __launch_bounds__(DT) k_direct(const double *__restrict__ src, const double *__restrict__ tau, int64_t ns,
                                               const double *__restrict__ tgt, int64_t nt,
                                               const int64_t *__restrict__ skip, double *__restrict__ out) {
  __shared__ double xs[DT], ys[DT], zs[DT], ts[DT];
  const int64_t j = (int64_t)blockIdx.x * DT + threadIdx.x;
  const int64_t jj = j < nt ? j : nt - 1;
  const double x = tgt[3 * jj], y = tgt[3 * jj + 1], z = tgt[3 * jj + 2];
  const int64_t sk = skip ? skip[jj] : -1;
  double acc[4] = {0, 0, 0, 0};
  for (int64_t i0 = 0; i0 < ns; i0 += DT) {
    const int64_t i = i0 + threadIdx.x;
    if (i < ns) {
      xs[threadIdx.x] = src[3 * i];
      ys[threadIdx.x] = src[3 * i + 1];
      zs[threadIdx.x] = src[3 * i + 2];
      ts[threadIdx.x] = tau[i];
    }
    __syncthreads();
    const int cnt = (int)min<int64_t>(DT, ns - i0);
    for (int q = 0; q < cnt; q++) {
      const double d0 = x - xs[q], d1 = y - ys[q], d2 = z - zs[q];
      const double ir = 1.0 / sqrt(fma(d2, d2, fma(d1, d1, d0 * d0)));
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
  if (OUT == FFMM_POTENTIAL) {
    out[j] = acc[0];
  } else if (OUT == FFMM_GRADIENT) {
    out[3 * j] = acc[1]; out[3 * j + 1] = acc[2]; out[3 * j + 2] = acc[3];
  } else {
    out[4 * j] = acc[0]; out[4 * j + 1] = acc[1]; out[4 * j + 2] = acc[2]; out[4 * j + 3] = acc[3];
  }
}

void direct(Context &ctx, const double *src, const double *tau, int64_t ns, const double *tgt, int64_t nt,
            const int64_t *skip, double *out, int output) {
  if (nt <= 0) return;
  unsigned g = nblocks(nt, DT);
  if (output == FFMM_POTENTIAL) k_direct<FFMM_POTENTIAL><<<g, DT, 0, ctx.stream>>>(src, tau, ns, tgt, nt, skip, out);
  else if (output == FFMM_GRADIENT) k_direct<FFMM_GRADIENT><<<g, DT, 0, ctx.stream>>>(src, tau, ns, tgt, nt, skip, out);
  else k_direct<2><<<g, DT, 0, ctx.stream>>>(src, tau, ns, tgt, nt, skip, out);
  FFMM_LAUNCH_CHECK();
}

}  // namespace ffmm
