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
            for (int k = 0; k < nlarge; k++) {
                double acc = 0.0;
                double yk0 = surf[3 * k] * s, yk1 = surf[3 * k + 1] * s, yk2 = surf[3 * k + 2] * s;
                accum(yk0 - xr, yk1 - yr, yk2 - zr, t, 0, &acc);
                col[k] += acc;
            }
        }
    }
}

/* evaluate_target for a set of targets: the local expansion of the target's leaf
 * (l2t over the large surface) plus the near field over the leaf's 27 colleagues,
 * the source `skip[q]` (sorted index; -1: none) excluded.  locals: nlarge x ncols,
 * column node_col[node]. out[outidx[q]] (grad: 3 entries). */
void oeval(const double *tx, const int32_t *tnode, const int32_t *tskip, const int64_t *outidx, int64_t nt,
           const double *locals, const int64_t *node_col, const double *surf, const double *scale, int nlarge,
           const int32_t *level, const double *centers, const int32_t *col, const int32_t *plo, const int32_t *phi,
           const double *src, const double *tau, int grad, double *out) {
    #pragma omp parallel for schedule(dynamic, 64)
    for (int64_t q = 0; q < nt; q++) {
        int32_t i = tnode[q];
        double x = tx[3 * q], y = tx[3 * q + 1], z = tx[3 * q + 2];
        const double *c = &centers[3 * i];
        double xr = x - c[0], yr = y - c[1], zr = z - c[2];
        double s = scale[level[i]];
        const double *L = &locals[(int64_t)nlarge * node_col[i]];
        double acc[3] = {0, 0, 0};
        for (int k = 0; k < nlarge; k++)
            accum(xr - surf[3 * k] * s, yr - surf[3 * k + 1] * s, zr - surf[3 * k + 2] * s, L[k], grad, acc);
        int32_t sk = tskip[q];
        for (int m = 0; m < 27; m++) {
            int32_t nb = col[27 * i + m];
            if (nb < 0) continue;
            for (int32_t j = plo[nb]; j <= phi[nb]; j++) {
                if (j == sk) continue;
                accum(x - src[3 * j], y - src[3 * j + 1], z - src[3 * j + 2], tau[j], grad, acc);
            }
        }
        int64_t o = outidx[q];
        if (grad) { out[3 * o] = acc[0]; out[3 * o + 1] = acc[1]; out[3 * o + 2] = acc[2]; }
        else out[o] = acc[0];
    }
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
