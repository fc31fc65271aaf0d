/*
 * ORACLE -- test infrastructure only.  Nothing in the product library links,
 * loads or calls this file; only tests/ (and smoke()) use it, as the checker.
 *
 * Plain-C restatement (FP64, OpenMP over nodes / targets) of the reference's
 * per-point loops of the harmonic methods:
 *
 *   _fold_regular (Schmidt recurrences)     src/operators/harmonics.jl:189-216
 *   _accumulate! charge / dipole            src/operators/harmonics.jl:236-275
 *   _harmonic_potential / _gradient         src/operators/harmonics.jl:456-491
 *   Yukawa radial weights, s2m!, l2t        src/operators/harmonics_yukawa.jl:43-63, 247-331
 *   _stokes_accumulate! (stokeslet,         src/operators/harmonics_stokes.jl:197-217
 *     stresslet), _stokes_fold1, velocity   src/operators/harmonics_stokes.jl:230-301, 321-344
 *   near field: Laplace3D, Yukawa3D,        src/kernels.jl:182-193, 228-240;
 *     Stokeslet3D stokeslet / stresslet       src/sources.jl:210-222
 *
 * Kinds: 0 Laplace3D charges, 1 Yukawa3D charges, 2 Stokeslet3D with stokeslet
 * densities (nd = 3) or stokeslet + stresslet (f, nu) densities (nd = 9).
 * Outputs (scalar kinds): 0 potential (1 value), 1 gradient (3), 2 both (4);
 * Stokes: velocity (3).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

static const double PI = 3.14159265358979323846;

typedef struct {
    int p;
    const double *a, *b, *e, *cz, *cm, *cp;
} HT;

static inline int tidx(int n, int m) { return n * (n + 1) / 2 + m; }

/* Re, Im of the Schmidt regular harmonics at (x, y, z), 0 <= m <= n <= nmax */
static void fold(const HT *tb, double x, double y, double z, int nmax, double *R, double *I) {
    double r2 = x * x + y * y + z * z;
    double sr = 1.0, si = 0.0;
    for (int m = 0; m <= nmax; m++) {
        R[tidx(m, m)] = sr;
        I[tidx(m, m)] = si;
        double pr = 0, pi = 0, cr = sr, ci = si;
        for (int n = m + 1; n <= nmax; n++) {
            int t = tidx(n, m);
            double az = tb->a[t] * z, br = tb->b[t] * r2;
            double nr = az * cr - br * pr, ni = az * ci - br * pi;
            R[t] = nr;
            I[t] = ni;
            pr = cr; pi = ci; cr = nr; ci = ni;
        }
        if (m < nmax) {
            double e = tb->e[m];
            double tr = -e * (x * sr - y * si), ti = -e * (x * si + y * sr);
            sr = tr; si = ti;
        }
    }
}

/* u_n(z), n = 0..nmax, from z^2 (harmonics_yukawa.jl:43-63) */
static void yuk_u(double z2, int nmax, double *u) {
    if (z2 < 1e-8) {
        for (int n = 0; n <= nmax; n++) u[n] = 1 + z2 / (2 * (2 * n + 3));
        return;
    }
    double z = sqrt(z2);
    int N = (nmax > (int)ceil(z) ? nmax : (int)ceil(z)) + 25;
    double a = 0.0, b = 1.0;
    for (int n = N - 1; n >= 0; n--) {
        double c = b + z2 * a / ((2.0 * n + 3) * (2.0 * n + 5));
        if (n <= nmax) u[n] = c;
        a = b; b = c;
    }
    double s = (sinh(z) / z) / b;
    for (int n = 0; n <= nmax; n++) u[n] *= s;
}

static inline void acc_entry(double *col, int n, int m, double re, double im) {
    if (m == 0) col[n * n] += re;
    else { col[n * n + 2 * m - 1] += re; col[n * n + 2 * m] += im; }
}

/* M += w(n) tau conj Y~(x) */
static void acc_charge(double *col, const HT *tb, const double *R, const double *I, int p, double tau, const double *w) {
    for (int m = 0; m <= p; m++)
        for (int n = m; n <= p; n++) {
            int t = tidx(n, m);
            double q = tau * (w ? w[n] : 1.0);
            acc_entry(col, n, m, q * R[t], -q * I[t]);
        }
}

/* M += w(n) conj(d . grad Y~) by the derivative identities, from the folds to p - 1 */
static void acc_dipole(double *col, const HT *tb, const double *R, const double *I, int p, const double d[3],
                       const double *w) {
    double dx = d[0], dy = d[1], dz = d[2];
    for (int m1 = 0; m1 <= p - 1; m1++)
        for (int n1 = m1; n1 <= p - 1; n1++) {
            int t1 = tidx(n1, m1);
            double ur = R[t1], ui = I[t1];
            int n = n1 + 1;
            double wn = w ? w[n] : 1.0;
            double c = wn * tb->cm[tidx(n, m1 + 1)] / 2;
            acc_entry(col, n, m1 + 1, -c * (dx * ur - dy * ui), c * (dx * ui + dy * ur));
            c = wn * tb->cz[tidx(n, m1)] * dz;
            acc_entry(col, n, m1, c * ur, -c * ui);
            if (m1 >= 2) {
                c = wn * tb->cp[tidx(n, m1 - 1)] / 2;
                acc_entry(col, n, m1 - 1, c * (dx * ur + dy * ui), c * (dy * ur - dx * ui));
            } else if (m1 == 1) {
                c = wn * tb->cp[tidx(n, 0)];
                acc_entry(col, n, 0, c * (dx * ur + dy * ui), 0.0);
            }
        }
}

static inline void get_entry(const double *L, int n, int m, double *re, double *im) {
    if (m > n) { *re = 0; *im = 0; }
    else if (m == 0) { *re = L[n * n]; *im = 0; }
    else { *re = L[n * n + 2 * m - 1]; *im = L[n * n + 2 * m]; }
}

static double h_potential(const double *L, const double *R, const double *I, int p, const double *w) {
    double acc = 0.0;
    for (int m = 0; m <= p; m++)
        for (int n = m; n <= p; n++) {
            int t = tidx(n, m);
            double lr, li;
            get_entry(L, n, m, &lr, &li);
            double wn = w ? w[n] : 1.0;
            acc += m == 0 ? wn * (lr * R[t]) : wn * (2 * (lr * R[t] + li * I[t]));
        }
    return acc;
}

/* scaled gradient, folds to p - 1 */
static void h_gradient(const HT *tb, const double *L, const double *R, const double *I, int p, const double *w,
                       double g[3]) {
    g[0] = g[1] = g[2] = 0;
    for (int m1 = 0; m1 <= p - 1; m1++)
        for (int n1 = m1; n1 <= p - 1; n1++) {
            int t1 = tidx(n1, m1);
            double vr = R[t1], vi = I[t1];
            int n = n1 + 1;
            double wn = w ? w[n] : 1.0;
            if (m1 == 0) {
                double lr, li;
                get_entry(L, n, 1, &lr, &li);
                double c = tb->cm[tidx(n, 1)];
                double ax = -c * lr, ay = -c * li, az = tb->cz[tidx(n, 0)] * L[n * n];
                g[0] += wn * (ax * vr);
                g[1] += wn * (ay * vr);
                g[2] += wn * (az * vr);
            } else {
                double ur, ui, dr, di, zr, zi;
                get_entry(L, n, m1 + 1, &ur, &ui);
                get_entry(L, n, m1 - 1, &dr, &di);
                double cu = tb->cm[tidx(n, m1 + 1)] / 2, cd = tb->cp[tidx(n, m1 - 1)] / 2;
                double axr = -cu * ur + cd * dr, axi = -cu * ui + cd * di;
                double ayr = -(cu * ui + cd * di), ayi = cu * ur + cd * dr;
                double cz = tb->cz[tidx(n, m1)];
                get_entry(L, n, m1, &zr, &zi);
                g[0] += (2 * wn) * (axr * vr + axi * vi);
                g[1] += (2 * wn) * (ayr * vr + ayi * vi);
                g[2] += (2 * wn) * (cz * (zr * vr + zi * vi));
            }
        }
}

/* ---- Stokes: complex helpers of harmonics_stokes.jl:230-270 ---- */
static void cget(const double *L, int p, int n, int m, double *re, double *im) {
    if (n > p || abs(m) > n) { *re = 0; *im = 0; return; }
    if (m == 0) { *re = L[n * n]; *im = 0; }
    else if (m > 0) { *re = L[n * n + 2 * m - 1]; *im = L[n * n + 2 * m]; }
    else {
        int k = -m;
        double s = (k % 2 == 0) ? 1.0 : -1.0;
        *re = s * L[n * n + 2 * k - 1];
        *im = -s * L[n * n + 2 * k];
    }
}
static double hcz(const HT *tb, int n, int m) { return abs(m) > n ? 0.0 : tb->cz[tidx(n, abs(m))]; }
static double hcm(const HT *tb, int n, int m) {
    if (abs(m) > n) return 0.0;
    return m >= 0 ? tb->cm[tidx(n, m)] : tb->cp[tidx(n, -m)];
}
static double hcp(const HT *tb, int n, int m) {
    if (abs(m) > n) return 0.0;
    return m >= 0 ? tb->cp[tidx(n, m)] : tb->cm[tidx(n, -m)];
}
/* the gradient coefficient sets at (n1, m1): ax, ay, az complex */
static void dcoef(const double *L, int p, const HT *tb, int n1, int m1, double out[6]) {
    int n = n1 + 1;
    double ar, ai, br, bi, zr, zi;
    cget(L, p, n, m1 + 1, &ar, &ai);
    cget(L, p, n, m1 - 1, &br, &bi);
    cget(L, p, n, m1, &zr, &zi);
    double cmv = -hcm(tb, n, m1 + 1), cpv = hcp(tb, n, m1 - 1), czv = hcz(tb, n, m1);
    double dmr = cmv * ar, dmi = cmv * ai, dpr = cpv * br, dpi = cpv * bi;
    out[0] = (dmr + dpr) / 2; out[1] = (dmi + dpi) / 2;            /* (dm + dp) / 2 */
    out[2] = -(dpi - dmi) / 2; out[3] = (dpr - dmr) / 2;           /* i (dp - dm) / 2 */
    out[4] = czv * zr; out[5] = czv * zi;
}
static inline double re_cv(double cr, double ci, double vr, double vi) { return cr * vr + ci * vi; }

/* velocity 1/2 [Phi - G x + grad Psi] of the four blocks at x (scaled coordinates) */
static void stokes_velocity(const HT *tb, const double *L, int N, const double *R, const double *I, int p,
                            const double x[3], double u[3]) {
    double Phi[3] = {0, 0, 0}, G[3][3] = {{0}}, gP[3] = {0, 0, 0};
    for (int m = 0; m <= p; m++)
        for (int n = m; n <= p; n++) {
            int t = tidx(n, m);
            double vr = R[t], vi = I[t];
            double wt = m == 0 ? 1.0 : 2.0;
            for (int k = 0; k < 3; k++) {
                double cr, ci;
                cget(L + k * N, p, n, m, &cr, &ci);
                Phi[k] += wt * re_cv(cr, ci, vr, vi);
            }
            if (n < p) {
                double a[6];
                for (int k = 0; k < 3; k++) {
                    dcoef(L + k * N, p, tb, n, m, a);
                    for (int i = 0; i < 3; i++) G[i][k] += wt * re_cv(a[2 * i], a[2 * i + 1], vr, vi);
                }
                dcoef(L + 3 * N, p, tb, n, m, a);
                for (int i = 0; i < 3; i++) gP[i] += wt * re_cv(a[2 * i], a[2 * i + 1], vr, vi);
            }
        }
    for (int i = 0; i < 3; i++) {
        double gx = G[i][0] * x[0] + G[i][1] * x[1] + G[i][2] * x[2];
        u[i] = (Phi[i] - gx + gP[i]) / 2;
    }
}

static HT mk(int p, const double *a, const double *b, const double *e, const double *cz, const double *cm,
             const double *cp) {
    HT t = {p, a, b, e, cz, cm, cp};
    return t;
}

/* S2M of the childless nodes: M (nq x nb N, row per node) */
void oh_s2m(int kind, int p, const double *ta, const double *tbb, const double *te, const double *tcz,
            const double *tcm, const double *tcp, const int32_t *nodes, int64_t nq, const double *centers,
            const int32_t *level, const double *scales, const double *khat, const int32_t *plo,
            const int32_t *phi, const double *pts, const double *dens, int nd, double *M) {
    HT tb = mk(p, ta, tbb, te, tcz, tcm, tcp);
    const int N = (p + 1) * (p + 1);
    const int nb = kind == 2 ? 4 : 1;
    const int nt = tidx(p + 1, p + 1) + 1;
    #pragma omp parallel
    {
        double *R = malloc(sizeof(double) * nt), *I = malloc(sizeof(double) * nt);
        double *u = malloc(sizeof(double) * (p + 3));
        #pragma omp for schedule(dynamic, 4)
        for (int64_t q = 0; q < nq; q++) {
            int i = nodes[q];
            double *col = M + q * (int64_t)nb * N;
            memset(col, 0, sizeof(double) * nb * N);
            double s = scales[level[i]];
            for (int j = plo[i]; j <= phi[i]; j++) {
                double x = (pts[3 * j] - centers[3 * i]) * s, y = (pts[3 * j + 1] - centers[3 * i + 1]) * s,
                       z = (pts[3 * j + 2] - centers[3 * i + 2]) * s;
                fold(&tb, x, y, z, p, R, I);
                if (kind == 0) {
                    /* nd = 4: a charge and a dipole per point, the dipole scaled by s
                     * (_scaled_density, src/operators/harmonics.jl:436) */
                    acc_charge(col, &tb, R, I, p, dens[(int64_t)nd * j], NULL);
                    if (nd == 4) {
                        double d[3];
                        for (int c = 0; c < 3; c++) d[c] = dens[4 * (int64_t)j + 1 + c] * s;
                        acc_dipole(col, &tb, R, I, p, d, NULL);
                    }
                } else if (kind == 1) {
                    double kh = khat[level[i]];
                    yuk_u(kh * kh * (x * x + y * y + z * z), p + 1, u);
                    acc_charge(col, &tb, R, I, p, dens[j], u);
                } else {
                    const double *f = dens + (int64_t)nd * j;
                    double yh[3] = {x, y, z};
                    for (int k = 0; k < 3; k++) acc_charge(col + k * N, &tb, R, I, p, f[k], NULL);
                    acc_charge(col + 3 * N, &tb, R, I, p, x * f[0] + y * f[1] + z * f[2], NULL);
                    if (nd == 9) {
                        const double *g = f + 3, *nu = f + 6;
                        for (int k = 0; k < 3; k++) {
                            double d[3];
                            for (int c = 0; c < 3; c++) d[c] = -(g[k] * nu[c] + nu[k] * g[c]) * s;
                            acc_dipole(col + k * N, &tb, R, I, p, d, NULL);
                        }
                        double nuy = nu[0] * yh[0] + nu[1] * yh[1] + nu[2] * yh[2];
                        double fy = g[0] * yh[0] + g[1] * yh[1] + g[2] * yh[2];
                        double d[3];
                        for (int c = 0; c < 3; c++) d[c] = -(nuy * g[c] + fy * nu[c]) * s;
                        acc_dipole(col + 3 * N, &tb, R, I, p, d, NULL);
                    }
                }
            }
        }
        free(R); free(I); free(u);
    }
}

/* far field at targets from the local expansions L (row per node, nb N entries) */
void oh_l2t(int kind, int p, const double *ta, const double *tbb, const double *te, const double *tcz,
            const double *tcm, const double *tcp, const double *tx, const int32_t *tnode, int64_t nt,
            const double *L, const double *centers, const int32_t *level, const double *scales,
            const double *khat, int output, double *out) {
    HT tb = mk(p, ta, tbb, te, tcz, tcm, tcp);
    const int N = (p + 1) * (p + 1);
    const int nb = kind == 2 ? 4 : 1;
    const int nt2 = tidx(p + 1, p + 1) + 1;
    const int nout = kind == 2 ? (output == 3 ? 1 : 3) : (output == 0 ? 1 : output == 1 ? 3 : 4);
    #pragma omp parallel
    {
        double *R = malloc(sizeof(double) * nt2), *I = malloc(sizeof(double) * nt2);
        double *u = malloc(sizeof(double) * (p + 3)), *w2 = malloc(sizeof(double) * (p + 3));
        #pragma omp for schedule(static)
        for (int64_t q = 0; q < nt; q++) {
            int i = tnode[q];
            double s = scales[level[i]];
            double x = (tx[3 * q] - centers[3 * i]) * s, y = (tx[3 * q + 1] - centers[3 * i + 1]) * s,
                   z = (tx[3 * q + 2] - centers[3 * i + 2]) * s;
            const double *Li = L + (int64_t)i * nb * N;
            double *o = out + q * nout;
            fold(&tb, x, y, z, p, R, I);
            if (kind == 2 && output == 3) {
                /* the pressure -s tr(grad Phi), each block's gradient by the Laplace
                 * gradient fold (_stokes_output(::Pressure), harmonics_stokes.jl:348-351) */
                double tr = 0;
                for (int k = 0; k < 3; k++) {
                    double g[3] = {0, 0, 0};
                    h_gradient(&tb, Li + k * N, R, I, p, NULL, g);
                    tr += g[k];
                }
                o[0] = -s * tr;
                continue;
            }
            if (kind == 2) {
                double xh[3] = {x, y, z};
                stokes_velocity(&tb, Li, N, R, I, p, xh, o);
                continue;
            }
            const double *w = NULL;
            double kh = 0;
            double pot = 0;
            if (kind == 1) {
                /* the potential takes the weights to p, the gradient to p + 1
                 * (harmonics_yukawa.jl:311-327) */
                kh = khat[level[i]];
                yuk_u(kh * kh * (x * x + y * y + z * z), p, u);
                pot = h_potential(Li, R, I, p, u);
                yuk_u(kh * kh * (x * x + y * y + z * z), p + 1, u);
                w = u;
            } else {
                pot = h_potential(Li, R, I, p, NULL);
            }
            double g[3] = {0, 0, 0};
            if (output != 0) {
                h_gradient(&tb, Li, R, I, p, w, g);
                if (kind == 1) {
                    for (int n = 0; n <= p; n++) w2[n] = kh * kh * u[n + 1] / (2 * n + 3);
                    double pr = h_potential(Li, R, I, p, w2);
                    g[0] += x * pr; g[1] += y * pr; g[2] += z * pr;
                }
                for (int c = 0; c < 3; c++) g[c] *= s;
            }
            if (output == 0) o[0] = pot;
            else if (output == 1) { o[0] = g[0]; o[1] = g[1]; o[2] = g[2]; }
            else { o[0] = pot; o[1] = g[0]; o[2] = g[1]; o[3] = g[2]; }
        }
        free(R); free(I); free(u); free(w2);
    }
}

/* the Laplace field of a dipole d at r = target - source (source_field and
 * target_field of Dipole, src/sources.jl:138, 151): potential d.r / (4 pi r^3),
 * gradient (d / r^3 - 3 (d.r) r / r^5) / (4 pi); output 0 potential, 1 gradient, 2 both */
static void dipole_field(const double *d, double dx, double dy, double dz, double ir, int output, double acc[4]) {
    const double c4 = 1.0 / (4.0 * PI);
    const double ir3 = ir * ir * ir, dr = d[0] * dx + d[1] * dy + d[2] * dz;
    const double pv = c4 * dr * ir3, gs = -3.0 * c4 * dr * ir3 * ir * ir;
    const double g0 = c4 * d[0] * ir3 + gs * dx, g1 = c4 * d[1] * ir3 + gs * dy, g2 = c4 * d[2] * ir3 + gs * dz;
    if (output == 0) acc[0] += pv;
    else if (output == 1) { acc[0] += g0; acc[1] += g1; acc[2] += g2; }
    else { acc[0] += pv; acc[1] += g0; acc[2] += g1; acc[3] += g2; }
}

/* the Stokes pressure at r = target - source of a stokeslet f, (f.r) / (4 pi r^3)
 * (pressure_kernel, src/kernels.jl:267-271), and with nd = 9 of a stresslet (g, nu),
 * ((g.nu) / r^3 - 3 (g.r)(nu.r) / r^5) / (2 pi) (src/sources.jl:237-242) */
static double stokes_pressure(const double *f, int nd, double dx, double dy, double dz, double ir) {
    const double ir3 = ir * ir * ir;
    double pr = (dx * f[0] + dy * f[1] + dz * f[2]) * ir3 / (4.0 * PI);
    if (nd == 9) {
        const double *g = f + 3, *nu = f + 6;
        const double gn = g[0] * nu[0] + g[1] * nu[1] + g[2] * nu[2];
        const double gr = g[0] * dx + g[1] * dy + g[2] * dz, nr = nu[0] * dx + nu[1] * dy + nu[2] * dz;
        pr += (gn * ir3 - 3.0 * gr * nr * ir3 * ir * ir) / (2.0 * PI);
    }
    return pr;
}

/* near field over the 27 colleagues, added into out; tskip: sorted source index to
 * exclude (-1 none) */
void oh_near(int kind, double kappa, const double *tx, const int32_t *tnode, const int32_t *tskip, int64_t nt,
             const int32_t *colleagues, const int32_t *plo, const int32_t *phi, const double *pts,
             const double *dens, int nd, int output, double *out) {
    const double c4 = 1.0 / (4.0 * PI), c8 = 1.0 / (8.0 * PI);
    const int nout = kind == 2 ? (output == 3 ? 1 : 3) : (output == 0 ? 1 : output == 1 ? 3 : 4);
    #pragma omp parallel for schedule(dynamic, 64)
    for (int64_t q = 0; q < nt; q++) {
        int i = tnode[q];
        double x = tx[3 * q], y = tx[3 * q + 1], z = tx[3 * q + 2];
        double acc[4] = {0, 0, 0, 0};
        for (int k = 0; k < 27; k++) {
            int nb = colleagues[(int64_t)i * 27 + k];
            if (nb < 0) continue;
            for (int j = plo[nb]; j <= phi[nb]; j++) {
                if (j == tskip[q]) continue;
                double dx = x - pts[3 * j], dy = y - pts[3 * j + 1], dz = z - pts[3 * j + 2];
                double r2 = fma(dz, dz, fma(dy, dy, dx * dx));
                double ir = 1.0 / sqrt(r2);
                if (kind == 2 && output == 3) {
                    acc[0] += stokes_pressure(dens + (int64_t)nd * j, nd, dx, dy, dz, ir);
                    continue;
                }
                if (kind == 2) {
                    const double *f = dens + (int64_t)nd * j;
                    double a = c8 * ir;
                    double df = dx * f[0] + dy * f[1] + dz * f[2];
                    double b = a * ir * ir * df;
                    acc[0] += a * f[0] + b * dx;
                    acc[1] += a * f[1] + b * dy;
                    acc[2] += a * f[2] + b * dz;
                    if (nd == 9) {
                        const double *g = f + 3, *nu = f + 6;
                        double ir2 = ir * ir;
                        double e = -(6 * c8) * (dx * g[0] + dy * g[1] + dz * g[2]) * (dx * nu[0] + dy * nu[1] + dz * nu[2]) *
                                   (ir2 * ir2 * ir);
                        acc[0] += e * dx; acc[1] += e * dy; acc[2] += e * dz;
                    }
                    continue;
                }
                double t = dens[(int64_t)nd * j];
                double v, dv;      /* value and radial derivative (times 1/r) */
                if (kind == 0 && nd == 4)
                    dipole_field(dens + 4 * (int64_t)j + 1, dx, dy, dz, ir, output, acc);
                if (kind == 0) {
                    v = c4 * ir;
                    dv = -c4 * ir * ir * ir;
                } else {
                    double r = r2 * ir;
                    double e = exp(-kappa * r);
                    v = c4 * e * ir;
                    dv = -c4 * e * (kappa * ir + ir * ir) * ir;
                }
                if (output == 0) acc[0] += v * t;
                else if (output == 1) { acc[0] += dv * dx * t; acc[1] += dv * dy * t; acc[2] += dv * dz * t; }
                else { acc[0] += v * t; acc[1] += dv * dx * t; acc[2] += dv * dy * t; acc[3] += dv * dz * t; }
            }
        }
        for (int c = 0; c < nout; c++) out[q * nout + c] += acc[c];
    }
}

/* FP64 direct sums of the three kinds at separate targets (skip[j]: source excluded, -1 none) */
void oh_direct(int kind, double kappa, const double *src, const double *dens, int nd, int64_t ns, const double *tgt,
               int64_t nt, const int64_t *skip, int output, double *out) {
    const double c4 = 1.0 / (4.0 * PI), c8 = 1.0 / (8.0 * PI);
    const int nout = kind == 2 ? (output == 3 ? 1 : 3) : (output == 0 ? 1 : output == 1 ? 3 : 4);
    #pragma omp parallel for schedule(dynamic, 16)
    for (int64_t q = 0; q < nt; q++) {
        double x = tgt[3 * q], y = tgt[3 * q + 1], z = tgt[3 * q + 2];
        double acc[4] = {0, 0, 0, 0};
        int64_t sk = skip ? skip[q] : -1;
        for (int64_t j = 0; j < ns; j++) {
            if (j == sk) continue;
            double dx = x - src[3 * j], dy = y - src[3 * j + 1], dz = z - src[3 * j + 2];
            double r2 = fma(dz, dz, fma(dy, dy, dx * dx));
            double ir = 1.0 / sqrt(r2);
            if (kind == 2 && output == 3) {
                acc[0] += stokes_pressure(dens + (int64_t)nd * j, nd, dx, dy, dz, ir);
                continue;
            }
            if (kind == 2) {
                const double *f = dens + (int64_t)nd * j;
                double a = c8 * ir;
                double b = a * ir * ir * (dx * f[0] + dy * f[1] + dz * f[2]);
                acc[0] += a * f[0] + b * dx; acc[1] += a * f[1] + b * dy; acc[2] += a * f[2] + b * dz;
                if (nd == 9) {
                    const double *g = f + 3, *nu = f + 6;
                    double ir2 = ir * ir;
                    double e = -(6 * c8) * (dx * g[0] + dy * g[1] + dz * g[2]) * (dx * nu[0] + dy * nu[1] + dz * nu[2]) *
                               (ir2 * ir2 * ir);
                    acc[0] += e * dx; acc[1] += e * dy; acc[2] += e * dz;
                }
                continue;
            }
            double t = dens[(int64_t)nd * j], v, dv;
            if (kind == 0 && nd == 4) dipole_field(dens + 4 * j + 1, dx, dy, dz, ir, output, acc);
            if (kind == 0) { v = c4 * ir; dv = -c4 * ir * ir * ir; }
            else {
                double e = exp(-kappa * r2 * ir);
                v = c4 * e * ir;
                dv = -c4 * e * (kappa * ir + ir * ir) * ir;
            }
            if (output == 0) acc[0] += v * t;
            else if (output == 1) { acc[0] += dv * dx * t; acc[1] += dv * dy * t; acc[2] += dv * dz * t; }
            else { acc[0] += v * t; acc[1] += dv * dx * t; acc[2] += dv * dy * t; acc[3] += dv * dz * t; }
        }
        for (int c = 0; c < nout; c++) out[q * nout + c] = acc[c];
    }
}
