/*
 * ffmm.h -- the C ABI of libffmm_hip.so, an MI355X (gfx950) backend for the device
 * passes of the FlexibleFMM.jl fast multipole method (Laplace3D on cube
 * kernel-independent surfaces).
 *
 * Every entry point returns 0 on success or one of the FFMM_ERR_* codes; the message
 * of the last failure on the calling thread is ffmm_last_error().  Handles are
 * opaque.  Device pointers are plain device addresses on the context's device; the
 * library reads and writes them on the context's stream and synchronises that stream
 * before returning (as the reference synchronises at the end of each pass and
 * evaluation).  Matrices are column major.  Indices are 0-based (-1 = none).
 *
 * One host thread drives a plan at a time (as in the reference: nothing makes
 * concurrent calls on one plan safe); distinct plans may be driven from distinct
 * threads.
 */
#ifndef FFMM_H
#define FFMM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* error codes; the first two mirror the reference's ArgumentError and DimensionMismatch */
#define FFMM_OK 0
#define FFMM_ERR_ARGUMENT 1
#define FFMM_ERR_DIMENSION 2
#define FFMM_ERR_DEVICE 3
#define FFMM_ERR_STATE 4
#define FFMM_ERR_MEMORY 5

/* arithmetic of the passes */
#define FFMM_PREC_F64 0    /* everything in FP64 */
#define FFMM_PREC_MIXED 1  /* FP32 S2M, M2L, L2T and near field; FP64 pseudo-inverses, M2M, L2L */

/* M2L form */
#define FFMM_M2L_DENSE 0     /* symmetry-class GEMMs (the reference device plan's form) */
#define FFMM_M2L_SPECTRAL 1  /* batched FFT + per-pair Hadamard products (the reference CPU form) */

/* outputs: Potential(), Gradient() (src/sources.jl:31-43), or both from one pass */
#define FFMM_POTENTIAL 0
#define FFMM_GRADIENT 1
#define FFMM_POTENTIAL_GRADIENT 2
/* Pressure() of the Stokes kernels, one value per point (src/sources.jl:34-42;
 * _stokes_output(::Pressure), src/operators/harmonics_stokes.jl:348-351) */
#define FFMM_PRESSURE 3

typedef struct ffmm_context_s *ffmm_context;
typedef struct ffmm_opcache_s *ffmm_opcache;
typedef struct ffmm_plan_s *ffmm_plan;

/* Message of the last failure on this thread ("" if none). */
const char *ffmm_last_error(void);
/* ABI version (increments on any signature change). */
int ffmm_abi_version(void);

/* A context on one device (hipSetDevice numbering); its own stream when stream == NULL,
 * else the caller's hipStream_t.  No reference counterpart: the reference's "backend"
 * (KernelAbstractions get_backend, src/device.jl:4-7) plays this role. */
int ffmm_context_create(int device, void *stream, ffmm_context *out);
int ffmm_context_destroy(ffmm_context ctx);

/* Componentwise bounds of n device points (xyz, 3 x n AoS, FP64), NaN propagating, and
 * whether every coordinate is finite: _device_bounds (src/tree/device_builder.jl:124-148). */
int ffmm_bounds(ffmm_context ctx, const double *xyz, int64_t n, double lo[3], double hi[3], int *all_finite);

/* FP64 direct sums on the device: direct! / direct_self! (src/kernels.jl:442-487).
 * out[j] = sum_i target_field(K, src_i, tgt_j, tau_i, output) over the ns sources
 * (3 x ns AoS) at the nt targets (3 x nt AoS), excluding source skip[j] when skip is
 * not NULL (skip[j] < 0: none).  output: FFMM_POTENTIAL (nt), FFMM_GRADIENT (3 x nt),
 * FFMM_POTENTIAL_GRADIENT (4 x nt). */
int ffmm_direct(ffmm_context ctx, const double *src, const double *tau, int64_t ns, const double *tgt, int64_t nt,
                const int64_t *skip, double *out, int output);

/* ffmm_direct without any device allocation.  No counterpart in the reference: it is
 * for callers that cannot allocate device memory (code under HIP graph capture, serving
 * loops that forbid hipMalloc).  The chunks' partial sums live in a 32 MiB device array
 * of the library (one per device, held until the process ends), so nt is at most
 * 1,048,576 (larger nt: FFMM_ERR_ARGUMENT, nothing launched); results equal
 * ffmm_direct's bitwise.  Calls that overlap on one device (several host threads or
 * streams) share the array and race: serialise them. */
int ffmm_direct_noalloc(ffmm_context ctx, const double *src, const double *tau, int64_t ns, const double *tgt,
                        int64_t nt, const int64_t *skip, double *out, int output);

/* FP64 direct Stokes velocity: direct! for Stokeslet3D (src/kernels.jl:463-487) with the
 * stresslet of src/operators/harmonics_stokes.jl:197-217.  dens: nd x ns, point major,
 * nd = 3 (stokeslets f) or 9 (rows [f, g, nu]: stokeslet f, stresslet strength g and
 * normal nu); out: 3 x nt; skip as for ffmm_direct. */
int ffmm_direct_stokes(ffmm_context ctx, const double *src, const double *dens, int nd, int64_t ns,
                       const double *tgt, int64_t nt, const int64_t *skip, double *out);

/* The FP64 direct Stokes pressure, one value per target: direct! with Pressure() for
 * Stokeslet3D (src/kernels.jl:463-487 with pressure_kernel, src/kernels.jl:267-271) and
 * the stresslet's pressure (src/sources.jl:237-242); arguments as ffmm_direct_stokes,
 * out: nt. */
int ffmm_direct_stokes_pressure(ffmm_context ctx, const double *src, const double *dens, int nd, int64_t ns,
                                const double *tgt, int64_t nt, const int64_t *skip, double *out);

/* The device copies of the operators of one method, shared by every plan built with
 * the cache: OperatorCache.device (src/operators/cache.jl:5-22,
 * src/device_passes.jl:136-146).  The method is CubeKernelIndependent{3}(Laplace3D(); p, pc, d)
 * (src/operators/cubeki.jl:18-48). */
int ffmm_opcache_create(ffmm_context ctx, int p, int pc, double d, ffmm_opcache *out);
int ffmm_opcache_destroy(ffmm_opcache cache);

/* Upload the reference-width operators (host FP64, column major), built by
 * level_operators(method, w_ref) (src/operators/cubeki.jl:55-81) and
 * class_operators (src/operators/symmetry.jl:125-131):
 *   small (3 x Ns), large (3 x Nl) surfaces relative to the box centre;
 *   P_up = V_up diag(invS_up) Uh_up (Uh_up: r_up x Nl, V_up: Ns x r_up);
 *   P_down = V_dn diag(invS_dn) Uh_dn (Uh_dn: r_dn x Ns, V_dn: Nl x r_dn);
 *   M2M (Ns x 8Ns), L2L (8Ns x Ns);
 *   nclass representatives (Ns x Ns each, consecutive), the class of each of the 316
 *   far offsets, and per offset the signed permutation idx (Ns x 316, 0-based), sgn (Ns x 316)
 *   with M2L_f m = P_f A_rep (P_f^T m), (P^T m)[i] = sgn[i] m[idx[i]];
 *   tables: the spectral kernel tables, 316 x (2p)(2p)(p+1) complex (interleaved re, im),
 *   or NULL to disable the spectral M2L. */
int ffmm_opcache_put_cube(ffmm_opcache cache, double w_ref, const double *small, const double *large,
                          int r_up, const double *Uh_up, const double *invS_up, const double *V_up,
                          int r_dn, const double *Uh_dn, const double *invS_dn, const double *V_dn,
                          const double *M2M, const double *L2L, int nclass, const double *class_ops,
                          const int32_t *far_class, const int32_t *idx, const int8_t *sgn, const double *tables);

/* Low-rank factors of the nclass representatives of a cube cache, for one tolerance:
 * class_factors (src/operators/symmetry.jl:169-178), rep_c ~ A_c Bt_c with the singular
 * values above tol times the largest kept (k_c = ranks[c]); A: the A_c (Ns x k_c, column
 * major) one after another, Bt: the Bt_c (k_c x Ns, column major) one after another.
 * A tolerance's factors are fixed once put. */
int ffmm_opcache_put_class_factors(ffmm_opcache cache, double tol, const int32_t *ranks, const double *A,
                                   const double *Bt);

/* kernels of the harmonic methods */
#define FFMM_KERNEL_LAPLACE 0    /* Laplace3D(): 1/(4 pi r) (src/kernels.jl:179-200) */
#define FFMM_KERNEL_YUKAWA 1     /* Yukawa3D(kappa): e^{-kappa r}/(4 pi r) (src/kernels.jl:220-240) */
#define FFMM_KERNEL_STOKESLET 2  /* Stokeslet3D(): (1/8pi)(I + r r^T/r^2)/r (src/kernels.jl:241-272) */

/* The device copies of a harmonic method's operators: SphericalHarmonics(Laplace3D(); p)
 * (src/operators/harmonics.jl:313-322), SphericalHarmonics(Yukawa3D(kappa); p)
 * (src/operators/harmonics_yukawa.jl:27-33) or StokesHarmonics(Stokeslet3D(); p)
 * (src/operators/harmonics_stokes.jl:34-40), for boxes of width w_ref * 2^e.  The
 * recurrence tables (HarmonicTables, src/operators/harmonics.jl:163-187) are built here
 * from p; expansions hold (p+1)^2 reals per scalar block (4 blocks for Stokes). */
int ffmm_opcache_create_harmonic(ffmm_context ctx, int kernel, int p, double kappa, double w_ref, ffmm_opcache *out);

/* The class structure of the M2L: the class of each of the 316 far offsets and per
 * offset the source-side (idx, sgn) and target-side (idx_t, sgn_t) signed permutations
 * of the real harmonic storage, (N x 316 each, 0-based), with
 * M2L_f m = P_t (A_rep (P_s^T m)), (P_s^T m)[i] = sgn[i] m[idx[i]], (P_t y)[idx_t[i]] = sgn_t[i] y[i]:
 * harmonic_symmetry_classes (src/operators/symmetry.jl:290-308). */
int ffmm_opcache_put_harmonic_classes(ffmm_opcache cache, int nclass, const int32_t *far_class, const int32_t *idx,
                                      const int8_t *sgn, const int32_t *idx_t, const int8_t *sgn_t);

/* The operators of boxes of one width (host FP64, column major): the scalar folded M2M
 * (N x 8N), L2L (8N x N) and the nclass M2L representatives (N x N each, consecutive,
 * c_K / width included): level_operators(method, width) (src/operators/harmonics.jl:357-385,
 * src/operators/harmonics_yukawa.jl:219-231).  For Stokes they are the scalar Laplace
 * operators (src/operators/harmonics_stokes.jl:74-82); the dense 4N x 32N M2M and L2L with
 * the coupling of the fourth potential (_stokes_dense_operators,
 * src/operators/harmonics_stokes.jl:112-131) are formed here.  Laplace and Stokes take
 * one width, w_ref, for every level (homogeneous: the M2L of level l is scaled by
 * w_ref / w_l, rescale, src/operators/harmonics.jl:393-398); Yukawa takes the width of
 * every level of a plan before the plan's first pass. */
int ffmm_opcache_put_harmonic_level(ffmm_opcache cache, double width, const double *M2M, const double *L2L,
                                    const double *reps);

/* has = 1 when the cache holds the operators of boxes of this width (else 0). */
int ffmm_opcache_has_width(ffmm_opcache cache, double width, int *has);

/* The plan from device points xyz (3 x n AoS, FP64, any order) in the box (origin, width):
 * DevicePlan(m, xs; ppl, bbox, cache, max_levels) (src/device_passes.jl:113-134) with the
 * device tree device_tree (src/tree/device_builder.jl:398-474), its lists
 * (_device_level_lists, src/device_lists.jl:219-278) and the buffers of _device_plan
 * (src/device_passes.jl:155-338).  The box must already be snapped to a width
 * width = w_ref * 2^e0 of the cache (fit_box!, src/operators/cache.jl:57-83).
 * precision: FFMM_PREC_*; m2l: FFMM_M2L_*. */
int ffmm_plan_create(ffmm_opcache cache, const double *xyz, int64_t n, int ppl, const double origin[3],
                     double width, int e0, int max_levels, int precision, int m2l, ffmm_plan *out);
int ffmm_plan_destroy(ffmm_plan plan);

/* Plan data: the plan's tree (the flat DeviceTree fields, src/tree/device_builder.jl:
 * 21-48: cells, levels, parents, children, point ranges, colleagues, flags, the point
 * permutation and sorted points, the node lists) with the plan's parameters and its
 * cache's identity, in host byte order.  No counterpart in the reference: it restarts a
 * serving cache's resident plans without their tree builds.  ffmm_plan_save writes it
 * into buf (when buf is not NULL; cap must be at least the size) and sets *size;
 * ffmm_plan_load makes a plan from it with a cache of the same operators (method,
 * order, kernel, reference width), rebuilding the lists and buffers. */
int ffmm_plan_save(ffmm_plan plan, void *buf, int64_t cap, int64_t *size);
int ffmm_plan_load(ffmm_opcache cache, const void *buf, int64_t size, ffmm_plan *out);

/* The checks ffmm_plan_load runs on plan data before any of it reaches a device, without
 * a device or a cache: every array's length, every index in range, and the structure of
 * the reference's trees (sibling groups of 8 on consecutive levels, child cells by child
 * slot, point ranges partitioning the parent's, colleagues at their offsets, a permutation
 * of the points, every point in one real leaf, the node lists of the tree).  FFMM_OK or
 * FFMM_ERR_ARGUMENT with the first violation.  No counterpart in the reference (its plans
 * are built in process). */
int ffmm_plan_data_check(const void *buf, int64_t size);

/* The plan from a tree built on the host, e.g. by the reference's host builder:
 * DevicePlan(f::FMM, to; tol, fast, eltype) (src/device_passes.jl:90-96) with the tree
 * moved by DeviceTree(f.tree, to) (src/tree/device_builder.jl:54-66).  Host arrays in
 * the layout of ffmm_plan_download_tree, 0-based (-1 = none): level_start (nlevels + 1:
 * the first node of each level, then the node count), cell (3 x nnodes), parent,
 * child_first, plo, phi (the points of node i are plo[i] .. phi[i] of the tree order),
 * leaf, ghost, xlist (0/1 bytes), colleagues (27 per node, slot k at the offset
 * (k mod 3 - 1, (k / 3) mod 3 - 1, k / 9 - 1)), perm (npoints: input index of each
 * tree-order point) and sorted_points (3 x npoints, the points in tree order).  The box
 * (origin, width = w_ref * 2^e0) is the tree's root.  The tree is checked in full as
 * ffmm_plan_data_check checks plan data (FFMM_ERR_ARGUMENT on a violation), uploaded,
 * and the lists and buffers are built as for ffmm_plan_create.  At most 22 levels. */
int ffmm_plan_from_tree(ffmm_opcache cache, int nlevels, const int64_t *level_start, int64_t npoints,
                        const int32_t *cell, const int32_t *parent, const int32_t *child_first, const int32_t *plo,
                        const int32_t *phi, const uint8_t *leaf, const uint8_t *ghost, const uint8_t *xlist,
                        const int32_t *colleagues, const int32_t *perm, const double *sorted_points, int ppl,
                        const double origin[3], double width, int e0, int precision, int m2l, ffmm_plan *out);

/* Tree and list sizes: nlevels, nnodes, npoints, and per level (arrays of length
 * nlevels, any may be NULL) the first node, the node count, the far pairs, the childless
 * sourced nodes and the real leaves. */
int ffmm_plan_sizes(ffmm_plan plan, int *nlevels, int64_t *nnodes, int64_t *npoints, int64_t *level_start,
                    int64_t *level_count, int64_t *level_pairs, int64_t *level_s2m, int64_t *level_leaves);

/* Download the flat tree (host buffers sized by ffmm_plan_sizes; any may be NULL):
 * Tree(dt::DeviceTree) (src/tree/device_builder.jl:508-525).  cell: 3 x nnodes AoS;
 * colleagues: 27 x nnodes (27 per node); flags as 0/1 bytes; sorted_points: 3 x N AoS. */
int ffmm_plan_download_tree(ffmm_plan plan, int32_t *cell, int32_t *level, int32_t *parent, int32_t *child_first,
                            int32_t *plo, int32_t *phi, uint8_t *leaf, uint8_t *ghost, uint8_t *xlist,
                            uint8_t *has_source, int32_t *colleagues, int32_t *perm, int32_t *point_node,
                            double *sorted_points);

/* The tree's point permutation into a device buffer of N int32 (perm[j]: the input index
 * of the j-th point in tree order), the device tree's perm (src/tree/device_builder.jl:
 * 21-48); the groups' per-group orders (gperm, 86-116) follow from it and the group sizes. */
int ffmm_plan_perm(ffmm_plan plan, int32_t *perm_dev);

/* Download one level's far pairs, targets in node order: ptr (count + 1, relative to the
 * level's first node), src (global node), far (0-based far offset index), as
 * _pair_fill! writes them (src/device_lists.jl:80-106; src/tree/lists.jl:96-124). */
int ffmm_plan_download_lists(ffmm_plan plan, int level, int32_t *ptr, int32_t *src, int32_t *far);

/* The upward and downward passes from densities tau (N, FP64, device): in the user's
 * order unless sorted != 0 (then in the tree's order):
 * device_build_expansions!(plan, ds; sorted) (src/device_passes.jl:511-597). */
int ffmm_build_expansions(ffmm_plan plan, const double *tau, int sorted);

/* The same from densities of nd components per point (dens: nd x N, point major):
 * nd = 1 charges (any method); for Stokes nd = 3 stokeslets f, or nd = 9 a stokeslet f
 * plus a stresslet (f, nu) at every point (the tuple densities of the reference,
 * (SVector{3}, Stresslet{3}), src/sources.jl:55-66; src/operators/harmonics_stokes.jl:197-217);
 * for SphericalHarmonics(Laplace3D()) nd = 4, a charge q and a dipole d per point, rows
 * (q, d) (Dipole{3} densities, src/sources.jl:17-29, 138, 151; supports_density,
 * src/operators/harmonics.jl:317-320). */
int ffmm_build_expansions_nd(ffmm_plan plan, const double *dens, int nd, int sorted);

/* Output at every source, in the user's order, from all other sources:
 * evaluate_at_sources! (src/device.jl:153-163).  out: N (potential) or 3 x N AoS (gradient).
 * Stokes plans: FFMM_POTENTIAL is the velocity (3 x N AoS, src/sources.jl:31-43). */
int ffmm_evaluate_at_sources(ffmm_plan plan, double *out, int output);

/* Output at nt device targets (3 x nt AoS), located on the device: evaluate!
 * (src/device.jl:134-145) with the host locate (src/tree/locate.jl:10-66) replaced by a
 * device walk.  Targets must lie in the root box (FFMM_ERR_ARGUMENT otherwise). */
int ffmm_evaluate(ffmm_plan plan, const double *targets, int64_t nt, double *out, int output);

/* Output at nt device targets located by the caller: evaluate!(out, dev, ds, targets,
 * nodes; output) (src/device.jl:127-145), nodes (device, nt int32, 0-based) the real
 * leaf of each target, as the host locate gives them (src/tree/locate.jl:10-47).  A node
 * that is not a real leaf of the plan's tree is refused (FFMM_ERR_ARGUMENT). */
int ffmm_evaluate_nodes(ffmm_plan plan, const double *targets, int64_t nt, const int32_t *nodes, double *out,
                        int output);

/* build_expansions then evaluate_at_sources: device_matvec! (src/device_passes.jl:619-640). */
int ffmm_matvec(ffmm_plan plan, const double *tau, int sorted, double *out, int output);

/* Download expansions (host, column major, any may be NULL): multipoles (Ns x nnodes),
 * partial locals (Ns x nnodes) and the locals of the real leaves (Nl x nleaves, in the
 * order of the leaves level by level). */
int ffmm_plan_download_expansions(ffmm_plan plan, double *multipoles, double *plocals, double *leaf_locals);

/* Capture the plan's pass sequence as one HIP graph, for solves repeated on the same
 * points (iterative solvers, time steps, served requests of a resident geometry): the
 * gather of nd densities per point from a plan-owned input buffer, the upward and
 * downward passes, and the evaluation of `output` at the sources into a plan-owned
 * output buffer.  One eager pass with zero densities runs first (its checks, buffers
 * and library workspaces); the plan's buffers then stay in place while the graph lives
 * (ffmm_plan_set_tol and a new capture drop it).  No reference counterpart: the
 * reference launches its passes one by one (src/device_passes.jl:619-640). */
int ffmm_plan_capture(ffmm_plan plan, int nd, int output);
/* One captured solve: dens (device or host, nd x N, input order) in, out (device or
 * host, ncomp x N, input order) written; the results of ffmm_build_expansions_nd +
 * ffmm_evaluate_at_sources, bitwise.  The expansions are left as that pass left them. */
int ffmm_plan_replay(ffmm_plan plan, const double *dens, double *out);

/* The class M2L of a cube plan with the dense class M2L through the low-rank factors of
 * tolerance tol (put with ffmm_opcache_put_class_factors), two GEMMs per class and chunk:
 * DevicePlan(...; tol) (src/device_passes.jl:14-27, 570-576); tol = 0: the dense
 * representatives again.  The expansions are to be built again after the change. */
int ffmm_plan_set_tol(ffmm_plan plan, double tol);

/* Per-stage device times of the last calls, in milliseconds (hipEvent pairs on the
 * plan's stream): [0] tree, [1] lists, [2] upward, [3] downward, [4] M2L (within
 * downward), [5] evaluation at sources, [6] evaluation at targets, [7] input gather,
 * [8] the Hadamard kernel of the spectral M2L (within [4]).  No reference counterpart:
 * the reference's verbose plan build prints host-side stage times
 * (src/device_passes.jl:331-336). */
int ffmm_plan_stage_times(ffmm_plan plan, double *ms, int n);
/* Per level l < n, the tiles (16 sibling groups each) of the mixed-precision spectral
 * M2L's Hadamard products run by the kernel that stages each tile's distinct colleague
 * boxes once (staged[l]) and by the per-group kernel (per_group[l]); zeros for other
 * plans.  No reference counterpart: a diagnostic of this library's two kernels. */
int ffmm_plan_spectral_tiles(ffmm_plan plan, int64_t *staged, int64_t *per_group, int n);

/* ---- serving: continuous batching of solves on one device ----
 * No reference counterpart.  The nearest pieces of the reference are the plan and
 * operator reuse of DevicePlan/OperatorCache (src/device_passes.jl:136-146,
 * src/operators/cache.jl:5-46) and build_expansions! on a fixed geometry
 * (src/fmm.jl:264-298).  A server runs nslots worker slots, each with its own stream;
 * a free slot admits the first queued request whose device-memory estimate fits
 * hbm_budget beside the running solves and the resident plans (least recently used
 * resident plans are evicted to make room; a request waiting longer than 2 s stops
 * later requests from overtaking it).  Requests with the same cache and geometry key
 * (>= 0) share one resident plan: after the first, they skip the tree and list build.
 * Host buffers of a request must stay valid until ffmm_server_wait returns for it.
 * The caches must not be modified while a server uses them (Yukawa: put the widths
 * of every level the requests can reach before submitting). */
typedef struct ffmm_server_s *ffmm_server;

typedef struct {
  ffmm_opcache cache;     /* the method's operators, on the server's device */
  int64_t geometry;       /* reuse key (-1: none) */
  const double *points;   /* host, 3 x n AoS */
  int64_t n;
  int ppl;
  double origin[3];       /* the box, already snapped: width = w_ref * 2^e0 of the cache */
  double width;
  int e0;
  int max_levels;         /* 0: 30 */
  int precision;          /* FFMM_PREC_* */
  int m2l;                /* FFMM_M2L_* */
  const double *dens;     /* host, nd x n */
  int nd;
  const double *targets;  /* host, 3 x nt AoS, or NULL */
  int64_t nt;
  int output;             /* FFMM_POTENTIAL, FFMM_GRADIENT, FFMM_POTENTIAL_GRADIENT */
  double *out_sources;    /* host, ncomp x n (NULL: none) */
  double *out_targets;    /* host, ncomp x nt */
} ffmm_request;

/* nslots worker slots (1:16), each with its own stream; hbm_budget: the device bytes the
 * running solves' estimates and the resident plans may hold.  Each slot keeps the idle
 * blocks of its solves for the next one, up to its share of the device memory free at
 * creation beyond the budget. */
int ffmm_server_create(int device, int nslots, int64_t hbm_budget, ffmm_server *out);
/* Queue a request; *ticket identifies it. */
int ffmm_server_submit(ffmm_server server, const ffmm_request *req, int64_t *ticket);
/* Block until the request is done; returns its status (its message: ffmm_last_error)
 * and its latency from submission to the last output copy. */
int ffmm_server_wait(ffmm_server server, int64_t ticket, double *latency_ms);
/* [0] completed, [1] plans built, [2] resident plans reused, [3] resident plans,
 * [4] device bytes held or reserved, [5] their peak, [6] evictions, [7] queued,
 * [8] idle device bytes the slots keep for reuse (outside the budget: an allocation
 *     that fails returns them),
 * [9] the largest ratio of a built plan's bytes to its admission estimate, x 1000,
 * [10] requests queued again after running out of device memory next to other solves
 *      (they then run alone),
 * [11] allocations in the process so far that failed and returned the contexts' idle
 *      blocks before trying again. */
int ffmm_server_stats(ffmm_server server, int64_t *stats, int n);
/* Plan data of the resident plan of (cache, geometry), as ffmm_plan_save writes it (buf
 * may be NULL to ask for the size).  No counterpart in the reference: with
 * ffmm_server_import it restarts a server's resident set warm.  The plan must be idle. */
int ffmm_server_export(ffmm_server server, ffmm_opcache cache, int64_t geometry, void *buf, int64_t cap,
                       int64_t *size);
/* A resident plan of (cache, geometry) from plan data: later requests with that key run
 * the passes only.  Counted against the memory budget (idle resident plans are evicted to
 * make room). */
int ffmm_server_import(ffmm_server server, ffmm_opcache cache, int64_t geometry, const void *buf, int64_t size);
/* Finish every queued request, then release the slots and the resident plans. */
int ffmm_server_destroy(ffmm_server server);

/* ---- the 2D kernel-independent method on circles ----
 * KernelIndependent(Laplace2D(); nequiv) (src/operators/kifmm.jl:1-87) for charges and
 * the potential: the reference's FMM(m, sources; ppl, bbox) with build_expansions and
 * evaluate_at_sources / evaluate (src/fmm.jl:31-49, 264-314), its quadtree from the host
 * builder for D = 2 (src/tree/builder.jl, lists src/tree/lists.jl:71-124), the passes and
 * the evaluation on the device in FP64. */
typedef struct ffmm_circle_s *ffmm_circle;

/* The plan of n host points (xy, 2 x n AoS) in the box (origin, width), ppl points per
 * leaf at most, nequiv (4:64) points per circle: the quadtree and its interaction lists
 * built on the host and uploaded.  Errors as FMM(m, sources): "the points must be
 * finite", "points lie outside the requested box", "tree reached max_levels". */
int ffmm_circle_create(ffmm_context ctx, int nequiv, const double *xy, int64_t n, int ppl, const double origin[2],
                       double width, int max_levels, ffmm_circle *out);
/* Levels, nodes, the first node of each level (nlevels + 1) and the far pairs of each
 * level (nlevels); the arrays may be NULL. */
int ffmm_circle_sizes(ffmm_circle plan, int *nlevels, int64_t *nnodes, int64_t *level_start, int64_t *level_pairs);
/* The operators of level `level` (2 .. nlevels-1; boxes of width ldexp(width, -level)):
 * level_operators(m, w) (src/operators/kifmm.jl:72-86, src/operators/ki_common.jl:
 * 176-199), column major: the large circle (2 x nequiv SoA, relative to the centre),
 * the pseudo-inverse factors Uh_up (r_up x n), Vs_up = V_up diag(invS_up) (n x r_up),
 * Uh_dn, Vs_dn, M2M (n x 4n), L2L (4n x n) and the 40 M2L matrices (n x n each, far
 * offsets in table order, src/util.jl:59-97). */
int ffmm_circle_put_level(ffmm_circle plan, int level, const double *large, int r_up, const double *Uh_up,
                          const double *Vs_up, int r_dn, const double *Uh_dn, const double *Vs_dn, const double *M2M,
                          const double *L2L, const double *M2L);
/* build_expansions (src/fmm.jl:264-298): tau, n device charges in input order. */
int ffmm_circle_build_expansions(ffmm_circle plan, const double *tau);
/* evaluate_at_sources (src/evaluate.jl:111-127): out, n device values in input order. */
int ffmm_circle_evaluate_at_sources(ffmm_circle plan, double *out);
/* evaluate (src/evaluate.jl:81-96) at nt device targets (2 x nt AoS) in the box,
 * located on the device (src/tree/locate.jl:10-22). */
int ffmm_circle_evaluate(ffmm_circle plan, const double *targets, int64_t nt, double *out);
/* The tree, 0-based (host buffers, NULL to skip): cell (2 x nnodes), parent,
 * child_first, plo, phi, leaf, ghost, xlist, colleagues (9 x nnodes), perm (n). */
int ffmm_circle_download_tree(ffmm_circle plan, int32_t *cell, int32_t *parent, int32_t *child_first, int32_t *plo,
                              int32_t *phi, uint8_t *leaf, uint8_t *ghost, uint8_t *xlist, int32_t *colleagues,
                              int32_t *perm);
/* The far pairs of a level by target: ptr (level's nodes + 1), src (global), far index. */
int ffmm_circle_download_lists(ffmm_circle plan, int level, int32_t *ptr, int32_t *src, int32_t *far);
int ffmm_circle_destroy(ffmm_circle plan);

#ifdef __cplusplus
}
#endif
#endif
