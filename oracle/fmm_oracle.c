/*
 * ORACLE -- test infrastructure only.  Nothing in the product library links,
 * loads or calls this file; only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py use it, and only as the checker / CPU baseline.
 *
 * Plain-C restatement (FP64, OpenMP over targets / nodes) of the reference's
 * per-point loops of the Laplace3D cube kernel-independent FMM:
 *
 *   Laplace3D body, fma'd _sqnorm, 1/sqrt   src/kernels.jl:16-22, 141, 182-193
 *   source_field / target_field            src/sources.jl:137-155
 *   direct! / direct_self!                 src/kernels.jl:442-487 (the FP64 oracle)
 *   s2m_accumulate!                        src/operators/ki_common.jl:256-272
 *   l2t (Potential, Gradient)              src/operators/ki_common.jl:293-315
 *   evaluate_target / _direct_range        src/evaluate.jl:12-25, 37-58
 *
 * The GEMM stages of the passes (pseudo-inverses, M2M, L2L, M2L) are driven from
 * oracle/fmm_oracle.py with numpy.  Build with -ffp-contract=off: the only fused
 * operations are the explicit fma calls of _sqnorm, as in the reference.
 */
#include <stdint.h>
#include <stdlib.h>
#include <math.h>

static const double C4PI = 1.0 / (4.0 * 3.14159265358979323846);

/* Laplace3D: K(s,t) = c / |t - s|; gradient_t = (-c ir^3) (t - s) */
static inline double sqnorm(double d1, double d2, double d3) { return fma(d3, d3, fma(d2, d2, d1 * d1)); }

static inline void accum(double dx, double dy, double dz, double tau, int grad, double *acc) {
    double ir = 1.0 / sqrt(sqnorm(dx, dy, dz));
    if (!grad) {
        acc[0] += (C4PI * ir) * tau;
    } else {
        double a = -C4PI * ir * ir * ir;
        acc[0] += (a * dx) * tau;
        acc[1] += (a * dy) * tau;
        acc[2] += (a * dz) * tau;
    }
}

/* direct!(out, K, sources, tau, targets) / direct_self!: out[j] = sum over sources,
 * excluding source skip[j] when skip != NULL (skip[j] < 0: none). grad: out is 3 x nt. */
void odirect(const double *src, const double *tau, int64_t ns, const double *tgt, int64_t nt,
             const int64_t *skip, int grad, double *out) {
    #pragma omp parallel for schedule(dynamic, 64)
    for (int64_t j = 0; j < nt; j++) {
        double acc[3] = {0, 0, 0};
        double x = tgt[3 * j], y = tgt[3 * j + 1], z = tgt[3 * j + 2];
        int64_t sk = skip ? skip[j] : -1;
        for (int64_t i = 0; i < ns; i++) {
            if (i == sk) continue;
            accum(x - src[3 * i], y - src[3 * i + 1], z - src[3 * i + 2], tau[i], grad, acc);
        }
        if (grad) { out[3 * j] = acc[0]; out[3 * j + 1] = acc[1]; out[3 * j + 2] = acc[2]; }
        else out[j] = acc[0];
    }
}

/* _s2m_level!: the check potentials of the childless sourced nodes `nodes` on the
 * large surface of their level, pM (nl x nq, column major).  Surface point k of a
 * node at level l is surf[k] * scale[l] relative to the node centre. */
void os2m(const int32_t *nodes, int64_t nq, const double *centers, const int32_t *level,
          const double *surf, const double *scale, int nlarge, const int32_t *plo, const int32_t *phi,
          const double *src, const double *tau, double *pM) {
    /* the surface as three coordinate arrays: the loop over the surface points is one
     * SIMD loop (each point's sum runs over the sources in order, as before) */
    double *sx = malloc(sizeof(double) * 3 * nlarge), *sy = sx + nlarge, *sz = sy + nlarge;
    for (int k = 0; k < nlarge; k++) { sx[k] = surf[3 * k]; sy[k] = surf[3 * k + 1]; sz[k] = surf[3 * k + 2]; }
    #pragma omp parallel for schedule(dynamic, 4)
    for (int64_t q = 0; q < nq; q++) {
        int32_t node = nodes[q];
        const double *c = &centers[3 * node];
        double s = scale[level[node]];
        double *col = &pM[(int64_t)nlarge * q];
        for (int k = 0; k < nlarge; k++) col[k] = 0.0;
        for (int32_t j = plo[node]; j <= phi[node]; j++) {
            double xr = src[3 * j] - c[0], yr = src[3 * j + 1] - c[1], zr = src[3 * j + 2] - c[2];
            double t = tau[j];
            #pragma omp simd
            for (int k = 0; k < nlarge; k++) {
                double d0 = sx[k] * s - xr, d1 = sy[k] * s - yr, d2 = sz[k] * s - zr;
                double ir = 1.0 / sqrt(sqnorm(d0, d1, d2));
                col[k] += (C4PI * ir) * t;
            }
        }
    }
    free(sx);
}

/* evaluate_target for a set of targets: the local expansion of the target's leaf
 * (l2t over the large surface) plus the near field over the leaf's 27 colleagues,
 * the source `skip[q]` (sorted index; -1: none) excluded.  locals: nlarge x ncols,
 * column node_col[node]. out[outidx[q]] (grad: 3 entries). */
void oeval(const double *tx, const int32_t *tnode, const int32_t *tskip, const int64_t *outidx, int64_t nt,
           const double *locals, const int64_t *node_col, const double *surf, const double *scale, int nlarge,
           const int32_t *level, const double *centers, const int32_t *col, const int32_t *plo, const int32_t *phi,
           const double *src, const double *tau, int grad, double *out) {
    /* surface and sources as coordinate arrays; the sums are SIMD reductions (partial
     * sums in lanes), the excluded source a masked select as in the reference
     * (src/evaluate.jl:7-10, 15-17) */
    double *sx = malloc(sizeof(double) * 3 * nlarge), *sy = sx + nlarge, *sz = sy + nlarge;
    for (int k = 0; k < nlarge; k++) { sx[k] = surf[3 * k]; sy[k] = surf[3 * k + 1]; sz[k] = surf[3 * k + 2]; }
    #pragma omp parallel for schedule(dynamic, 64)
    for (int64_t q = 0; q < nt; q++) {
        int32_t i = tnode[q];
        double x = tx[3 * q], y = tx[3 * q + 1], z = tx[3 * q + 2];
        const double *c = &centers[3 * i];
        double xr = x - c[0], yr = y - c[1], zr = z - c[2];
        double s = scale[level[i]];
        const double *L = &locals[(int64_t)nlarge * node_col[i]];
        double a0 = 0, a1 = 0, a2 = 0;
        if (!grad) {
            #pragma omp simd reduction(+ : a0)
            for (int k = 0; k < nlarge; k++) {
                double d0 = xr - sx[k] * s, d1 = yr - sy[k] * s, d2 = zr - sz[k] * s;
                double ir = 1.0 / sqrt(sqnorm(d0, d1, d2));
                a0 += (C4PI * ir) * L[k];
            }
        } else {
            #pragma omp simd reduction(+ : a0, a1, a2)
            for (int k = 0; k < nlarge; k++) {
                double d0 = xr - sx[k] * s, d1 = yr - sy[k] * s, d2 = zr - sz[k] * s;
                double ir = 1.0 / sqrt(sqnorm(d0, d1, d2));
                double a = -C4PI * ir * ir * ir;
                a0 += (a * d0) * L[k];
                a1 += (a * d1) * L[k];
                a2 += (a * d2) * L[k];
            }
        }
        int32_t sk = tskip[q];
        for (int m = 0; m < 27; m++) {
            int32_t nb = col[27 * i + m];
            if (nb < 0) continue;
            if (!grad) {
                #pragma omp simd reduction(+ : a0)
                for (int32_t j = plo[nb]; j <= phi[nb]; j++) {
                    double d0 = x - src[3 * j], d1 = y - src[3 * j + 1], d2 = z - src[3 * j + 2];
                    double ir = 1.0 / sqrt(sqnorm(d0, d1, d2));
                    a0 += j != sk ? (C4PI * ir) * tau[j] : 0.0;
                }
            } else {
                #pragma omp simd reduction(+ : a0, a1, a2)
                for (int32_t j = plo[nb]; j <= phi[nb]; j++) {
                    double d0 = x - src[3 * j], d1 = y - src[3 * j + 1], d2 = z - src[3 * j + 2];
                    double ir = 1.0 / sqrt(sqnorm(d0, d1, d2));
                    double a = -C4PI * ir * ir * ir;
                    int keep = j != sk;
                    a0 += keep ? (a * d0) * tau[j] : 0.0;
                    a1 += keep ? (a * d1) * tau[j] : 0.0;
                    a2 += keep ? (a * d2) * tau[j] : 0.0;
                }
            }
        }
        int64_t o = outidx[q];
        if (grad) { out[3 * o] = a0; out[3 * o + 1] = a1; out[3 * o + 2] = a2; }
        else out[o] = a0;
    }
    free(sx);
}

/* The class-wise M2L gather and scatter of the dense translations (the reference's
 * m2l_level! through the class representatives, src/fmm.jl:100-127 with
 * src/operators/symmetry.jl:104-117): expansions are node major (row per node).
 *   G[q, i] = sgn[f_q, i] Mp[src_q, idx[f_q, i]]                      (P^T m)
 *   pL[t, idx[f, i]] += alpha sgn[f, i] Y[q, i]  over the pairs q of a run (P y),
 * the pairs of one target consecutive (runs), runs in parallel. */
void om2l_gather(const double *Mp, int Ns, const int64_t *src, const int64_t *far, int64_t np,
                 const int32_t *idx, const int8_t *sgn, double *G) {
    #pragma omp parallel for schedule(static)
    for (int64_t q = 0; q < np; q++) {
        const double *m = &Mp[(int64_t)Ns * src[q]];
        const int32_t *ix = &idx[(int64_t)Ns * far[q]];
        const int8_t *sg = &sgn[(int64_t)Ns * far[q]];
        double *g = &G[(int64_t)Ns * q];
        for (int i = 0; i < Ns; i++) g[i] = (double)sg[i] * m[ix[i]];
    }
}

void om2l_scatter(double *pL, int Ns, const int64_t *tgt, const int64_t *far, const int64_t *runs, int64_t nruns,
                  const int32_t *idx, const int8_t *sgn, const double *Y, double alpha) {
    #pragma omp parallel for schedule(dynamic, 16)
    for (int64_t r = 0; r < nruns; r++) {
        for (int64_t q = runs[r]; q < runs[r + 1]; q++) {
            double *p = &pL[(int64_t)Ns * tgt[q]];
            const int32_t *ix = &idx[(int64_t)Ns * far[q]];
            const int8_t *sg = &sgn[(int64_t)Ns * far[q]];
            const double *y = &Y[(int64_t)Ns * q];
            for (int i = 0; i < Ns; i++) p[ix[i]] += alpha * ((double)sg[i] * y[i]);
        }
    }
}

/* The Hadamard accumulation of the spectral M2L of one level (_hadamard_block!,
 * src/operators/spectral_m2l.jl:260-286, inside _spectral_m2l_level!, 328-451): for the
 * k-th target with far pairs (level-local index tidx[k], pairs ptr[t] .. ptr[t+1]),
 *   U[k] = sum over its pairs q of tables[far[q]] .* S[src_slot[q]]
 * over nf complex frequencies ((re, im) interleaved).  As the reference blocks its
 * Hadamard loop, the frequencies go in blocks of 64 (the block's tables, 316 x 1 KiB,
 * stay in the core's cache), targets in parallel within a block. */
void ohadamard(const int32_t *ptr, const int32_t *src_slot, const int32_t *far, const int64_t *tidx, int64_t nt,
               const double *tables, const double *S, int64_t nf, double *U) {
    const int64_t FB = 64;
    #pragma omp parallel
    for (int64_t f0 = 0; f0 < nf; f0 += FB) {
        const int64_t f1 = f0 + FB < nf ? f0 + FB : nf;
        #pragma omp for schedule(dynamic, 16)
        for (int64_t k = 0; k < nt; k++) {
            const int64_t t = tidx[k];
            double *restrict u = &U[2 * nf * k];
            for (int64_t f = 2 * f0; f < 2 * f1; f++) u[f] = 0.0;
            for (int32_t q = ptr[t]; q < ptr[t + 1]; q++) {
                const double *restrict T = &tables[2 * nf * (int64_t)far[q]];
                const double *restrict s = &S[2 * nf * (int64_t)src_slot[q]];
                #pragma omp simd
                for (int64_t f = f0; f < f1; f++) {
                    const double tr = T[2 * f], ti = T[2 * f + 1], sr = s[2 * f], si = s[2 * f + 1];
                    u[2 * f] += tr * sr - ti * si;
                    u[2 * f + 1] += tr * si + ti * sr;
                }
            }
        }
    }
}

/* the threads the OpenMP loops above run on */
#include <omp.h>
int othreads(void) { return omp_get_max_threads(); }
