// Internal declarations of libffmm_hip: device buffers, the context, the operator
// cache and the plan.  gfx950 only.
#pragma once
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>

#include <cstdint>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>
#include <cstdlib>

#include "../../include/ffmm.h"

namespace ffmm {

constexpr int NCHILD = 8;
constexpr int NCOLL = 27;
constexpr int SELF_SLOT = 13;
constexpr int NFAR = 316;
constexpr int MAX_CLASSES = 64;      // symmetry classes of an M2L (16 cube, 34 harmonic)
constexpr int MAX_KEY_LEVELS = 21;   // 3 bits per level in a 63-bit key
constexpr int NSTAGES = 9;           // timed stages (ffmm_plan_stage_times)
inline int level_stage_slot(int stage) { return stage == 4 ? 0 : stage == 8 ? 1 : -1; }

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};
[[noreturn]] void fail(int code, const std::string &msg);

// the message of the last failure on this thread (ffmm_last_error)
extern thread_local std::string g_last_error;

// runs f, turning the library's exceptions into a status code and a message
template <class F>
int guard(F &&f) {
  try {
    f();
    g_last_error.clear();
    return FFMM_OK;
  } catch (const Error &e) {
    g_last_error = e.what();
    return e.code;
  } catch (const std::bad_alloc &) {
    g_last_error = "host allocation failed";
    return FFMM_ERR_MEMORY;
  } catch (const std::exception &e) {
    g_last_error = e.what();
    return FFMM_ERR_STATE;
  }
}

#define FFMM_HIP(call)                                                                          \
  do {                                                                                          \
    hipError_t e_ = (call);                                                                     \
    if (e_ != hipSuccess) ::ffmm::fail(FFMM_ERR_DEVICE, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)
#define FFMM_BLAS(call)                                                                         \
  do {                                                                                          \
    rocblas_status s_ = (call);                                                                 \
    if (s_ != rocblas_status_success)                                                           \
      ::ffmm::fail(FFMM_ERR_DEVICE, std::string(#call) + ": " + rocblas_status_to_string(s_));  \
  } while (0)
#define FFMM_LAUNCH_CHECK() FFMM_HIP(hipGetLastError())

// 1/sqrt(x) in FP64: the hardware estimate (v_rsq_f64) refined by two Newton steps,
// y <- y + y (1/2 - x y^2 / 2), to within an ulp or two of the correctly rounded value;
// +inf at x = 0 turns into NaN, which the callers mask with a select.
__device__ __forceinline__ double rsqrt_f64(double x) {
  double y = __builtin_amdgcn_rsq(x);
  const double hx = 0.5 * x;
  y = fma(y, fma(-hx * y, y, 0.5), y);
  y = fma(y, fma(-hx * y, y, 0.5), y);
  return y;
}

// runs upload() once per device and key, under a lock, so that no caller returns
// before the upload is done (the constant tables, which concurrent plans on other
// streams read)
void once_on_device(int device, int key, const std::function<void()> &upload);

// XCD-aware block order.  The dispatcher deals workgroups round-robin to the 8 XCDs
// (linear workgroup id mod 8), each with its own L2; a kernel whose neighbouring
// blocks share data takes block (L mod 8) * (total / 8) + L / 8 instead of L, so each
// XCD runs a contiguous range.  total must be a multiple of 8 (pad the grid with
// blocks that exit).
constexpr int NXCD = 8;
__device__ __forceinline__ int64_t xcd_block(int64_t L, int64_t total) {
  return (L % NXCD) * (total / NXCD) + L / NXCD;
}
inline int64_t pad_xcd(int64_t n) { return (n + NXCD - 1) / NXCD * NXCD; }

// workgroups covering n work-items (at least one: every kernel checks its index against n)
inline unsigned nblocks(int64_t n, int b) { return n > 0 ? (unsigned)((n + b - 1) / b) : 1u; }

// Device allocations.  Inside an AllocScope they belong to the scope's context: a
// freed block goes to the context's block cache, and the next allocation of a
// similar size on that context takes it back with no HIP call (plans of the same
// size come and go without hipMalloc / hipFree, which synchronise the device).  Every
// entry point synchronises its stream before it returns, and a context's allocations
// are used on its own stream, so a cached block is idle when it is reused.  Outside a
// scope, allocations are plain hipMalloc / hipFree.
struct Context;
extern thread_local Context *g_alloc_ctx;
struct AllocScope {
  Context *prev;
  explicit AllocScope(Context &c) : prev(g_alloc_ctx) { g_alloc_ctx = &c; }
  ~AllocScope() { g_alloc_ctx = prev; }
  AllocScope(const AllocScope &) = delete;
  AllocScope &operator=(const AllocScope &) = delete;
};
// *blk: the block's size in bytes (>= bytes); returns its owner (null: hipMalloc)
Context *device_alloc(void **p, size_t bytes, size_t *blk);
void device_free(void *p, size_t blk, Context *owner);

// A device array; resize() discards the contents, grow() keeps them.
template <class T>
struct DBuf {
  T *p = nullptr;
  size_t n = 0, cap = 0;
  size_t blk = 0;              // bytes of the underlying block
  Context *owner = nullptr;    // the context whose block cache it returns to (null: hipFree)
  DBuf() = default;
  DBuf(const DBuf &) = delete;
  DBuf &operator=(const DBuf &) = delete;
  DBuf(DBuf &&o) noexcept : p(o.p), n(o.n), cap(o.cap), blk(o.blk), owner(o.owner) {
    o.p = nullptr;
    o.n = o.cap = o.blk = 0;
  }
  DBuf &operator=(DBuf &&o) noexcept {
    if (this != &o) {
      release();
      p = o.p; n = o.n; cap = o.cap; blk = o.blk; owner = o.owner;
      o.p = nullptr; o.n = o.cap = o.blk = 0;
    }
    return *this;
  }
  ~DBuf() { release(); }
  void release() {
    if (p) device_free(p, blk, owner);
    p = nullptr;
    n = cap = blk = 0;
  }
  void resize(size_t m) {
    if (m > cap) {
      release();
      owner = device_alloc((void **)&p, (m ? m : 1) * sizeof(T), &blk);
      cap = blk / sizeof(T);
    }
    n = m;
  }
  void grow(size_t m, hipStream_t s) {
    if (m <= cap) { n = m; return; }
    size_t c = cap ? cap : 16;
    while (c < m) c *= 2;
    T *q = nullptr;
    size_t qb = 0;
    Context *qo = device_alloc((void **)&q, c * sizeof(T), &qb);
    if (n) FFMM_HIP(hipMemcpyAsync(q, p, n * sizeof(T), hipMemcpyDeviceToDevice, s));
    // a cached block is reused in stream order after the copy; a hipFree'd one waits
    if (!owner) FFMM_HIP(hipStreamSynchronize(s));
    if (p) device_free(p, blk, owner);
    p = q;
    owner = qo;
    blk = qb;
    cap = qb / sizeof(T);
    n = m;
  }
  T *data() const { return p; }
  size_t bytes() const { return n * sizeof(T); }
  void upload(const T *h, size_t m, hipStream_t s) {
    resize(m);
    if (m) FFMM_HIP(hipMemcpyAsync(p, h, m * sizeof(T), hipMemcpyHostToDevice, s));
  }
  void download(T *h, size_t m, hipStream_t s) const {
    if (m) FFMM_HIP(hipMemcpyAsync(h, p, m * sizeof(T), hipMemcpyDeviceToHost, s));
  }
};

// the idle device blocks of a context, by size
struct BlockCache {
  std::mutex mu;
  std::multimap<size_t, void *> blocks;
  size_t bytes = 0;
  size_t limit = 0;    // cached bytes kept at most
};

struct Context {
  int device = 0;
  hipStream_t stream = nullptr;
  bool own_stream = false;
  rocblas_handle blas = nullptr;
  std::unique_ptr<BlockCache> blocks = std::make_unique<BlockCache>();
  int64_t *pinned = nullptr;   // small pinned readback area (64 entries), coherent
  int64_t *pinned_dev = nullptr;   // its device address: kernels store results there directly
  DBuf<char> tmp;               // scratch of the primitives
  void sync() { FFMM_HIP(hipStreamSynchronize(stream)); }
  // read back `n` int64 values from device memory (blocking)
  void readback(int64_t *host, const int64_t *dev, int n);
};

// the methods of an operator cache
constexpr int METHOD_CUBE = 0;       // CubeKernelIndependent{3}(Laplace3D())
constexpr int METHOD_HARMONIC = 1;   // SphericalHarmonics (Laplace3D, Yukawa3D), StokesHarmonics

// the operators of the harmonic methods for boxes of one width
struct HarmonicLevel {
  double w = 0.0;
  DBuf<double> M2M, L2L;   // NB x 8NB, 8NB x NB (NB = nb N; the dense Stokes forms for nb = 4)
  DBuf<double> reps;       // nclass x (N x N)
  DBuf<float> reps32;      // the same in FP32 (the mixed form's M2L)
  DBuf<float> reps32r;     // FP32, row-major, zero-padded to 128 x 128 per class (N <= 128: the fused mixed M2L)
};

// device pointers of the recurrence tables (HarmonicTables, src/operators/harmonics.jl:163-187)
// the recurrence tables of the Schmidt harmonics (a, b, cz, cm, cp: ntri each; e: p),
// in FP64 and, for the mixed form's evaluation, rounded to FP32
template <class T>
struct HTabT {
  const T *a, *b, *cz, *cm, *cp, *e;
};
using HTab = HTabT<double>;
using HTab32 = HTabT<float>;

struct ClassFactors;
constexpr int UC2E_PROWS = 320, UC2E_PK = 64, UC2E_PKX = 2;   // rows, k rounding, extra k-steps

struct OpCache {
  Context *ctx = nullptr;
  int method = METHOD_CUBE;
  int p = 8, pc = 16, Ns = 0, Nl = 0, nclass = 0, r_up = 0, r_dn = 0;
  // harmonic methods: kernel, degree, scalar block length N = (p+1)^2 and blocks per
  // expansion nb (Ns = Nl = nb N)
  int kernel = FFMM_KERNEL_LAPLACE, hN = 0, nb = 1, ntri = 0;
  double kappa = 0.0;
  bool has_classes = false;
  DBuf<double> htab;                         // a, b, cz, cm, cp (ntri each), e (p)
  DBuf<float> htab32;                        // the same, rounded to FP32
  DBuf<int8_t> sgn_t;                        // target-side signs [f * N + i] (the cube: = sgn)
  DBuf<double> offs;                         // 316 x 3 far offsets (the Stokes coupling)
  // the operators by width; entries are added, never replaced or removed while the
  // cache lives, so plans may hold pointers to them (guarded: a server's slots look
  // widths up while the host adds new ones)
  std::vector<std::unique_ptr<HarmonicLevel>> hlev;
  mutable std::mutex hmu;
  const HarmonicLevel *find_width(double w) const {
    std::lock_guard<std::mutex> lk(hmu);
    for (auto &h : hlev)
      if (h->w == w) return h.get();
    return nullptr;
  }
  HTab tab() const {
    const double *t = htab.p;
    return HTab{t, t + ntri, t + 2 * ntri, t + 3 * ntri, t + 4 * ntri, t + 5 * ntri};
  }
  HTab32 tab32() const {
    const float *t = htab32.p;
    return HTab32{t, t + ntri, t + 2 * ntri, t + 3 * ntri, t + 4 * ntri, t + 5 * ntri};
  }
  double d = 0.1, w_ref = 0.0;
  bool ready = false;
  bool has_tables = false;
  std::vector<double> small_h, large_h;     // host copies of the surfaces (3 x n)
  DBuf<double> large;                        // 3 x Nl SoA: x[Nl], y[Nl], z[Nl]
  DBuf<float> large32;
  DBuf<double> Uh_up, Vs_up, Uh_dn, Vs_dn;   // Vs = V diag(invS)
  // the fused S2M + UC2E kernel's Uh (ranks 241..304): UC2E_PROWS rows per check point,
  // Nl rounded up to UC2E_PK and UC2E_PKX k-steps more, zero-padded (no guards on its loads)
  DBuf<double> Uh_up_pad;
  // the column GEMMs' operators (k_gemm_cols: UC2E's V diag(invS), DC2E's Uh), the same
  // layout (UC2E_PROWS rows per k, k rounded up to UC2E_PK and UC2E_PKX k-steps more)
  DBuf<double> Vs_up_pad, Uh_dn_pad;
  DBuf<double> M2M, L2L, class_ops;
  DBuf<float> M2M32, class_ops32;   // FP32 copies for the mixed form
  DBuf<int32_t> far_class, idx, inv;         // idx/inv: [f * Ns + i]
  // the same permutations with their signs in the top bit: gsel[f, i] = idx[f, i] (sign
  // sgn[f, i]); ssel[f, j] = i = inv[f, j] (sign sgn_t[f, i]); one load per entry
  DBuf<int32_t> gsel, ssel;
  // fsel[f, i] = the target-frame position j of product entry i (inv[f, j] = i), sign sgn_t[f, i]
  DBuf<int32_t> fsel;
  DBuf<int8_t> sgn;
  DBuf<double> tables;                       // 316 x nfreq complex (interleaved)
  DBuf<float> tables32;
  DBuf<int32_t> embed;                       // Ns: lattice index of each small-surface point
  DBuf<int32_t> sub8;                        // Ns: its index in the p^3 corner of the lattice
  DBuf<int16_t> lat8;                        // p^3: the surface point at a corner lattice index, or -1
  DBuf<int16_t> pairf;                       // [(k * 8 + c) * 8 + a]: far index of (sibling a, child c of colleague k);
                                             // row 216: all near (read past the last candidate)
  std::vector<int32_t> embed_h;
  std::vector<int32_t> far_class_h;
  // low-rank factors of the class representatives, one set per tolerance (added, never
  // replaced or removed while the cache lives: plans hold pointers to them)
  std::vector<std::unique_ptr<ClassFactors>> factors;
  std::mutex factors_mu;
};

// the class representatives of the cube method as low-rank products A_c Bt_c (A_c: Ns x k_c,
// Bt_c: k_c x Ns, column major, concatenated over the classes): class_factors,
// src/operators/symmetry.jl:169-178 (singular values above tol times the largest)
struct ClassFactors {
  double tol = 0.0;
  std::vector<int> rank;        // k_c per class
  std::vector<int64_t> off;     // first entry of A_c (and of Bt_c) in A, Bt
  int kmax = 0;
  DBuf<double> A, Bt;
  DBuf<float> A32, Bt32;
};

// per-level lists of far pairs (targets in node order) and their class-major form
struct LevelLists {
  int64_t np = 0;
  bool filled = false;                  // src/far written (the spectral M2L needs only ptr)
  DBuf<int32_t> ptr, src, far;          // by target: ptr (nl + 1, 0-based pair index)
  DBuf<int32_t> csrc, cfar;             // class major, by target within a class
  DBuf<int32_t> ctgt;                   // the level-local target of each class-major pair (fused mixed M2L)
  DBuf<int32_t> cstart;                 // nclass nl + 1: first class-major pair of (class c, target t) at c nl + t
  std::vector<int64_t> class_pair;      // nclass + 1 offsets into csrc
  // target chunks of the dense class M2L: chunk k covers the level-local targets
  // [chunk_t[k], chunk_t[k + 1]) and, in class c, the pairs
  // [chunk_cs[k nclass + c], chunk_cs[(k + 1) nclass + c])
  std::vector<int32_t> chunk_t, chunk_cs;
  int64_t chunk_max = 0;                // the most pairs in one chunk
};

struct Plan {
  Context *ctx = nullptr;
  OpCache *ops = nullptr;
  int64_t N = 0;
  int ppl = 0, max_levels = 30, precision = FFMM_PREC_F64, m2l = FFMM_M2L_DENSE;
  double origin[3] = {0, 0, 0}, W0 = 1.0;
  int e0 = 0;
  int L = 0;
  int64_t nnodes = 0;
  std::vector<int64_t> lo, nl;
  std::vector<double> lam;   // level width / reference width
  DBuf<double> lam_d;
  // harmonic methods: the operators of each level, the M2L factor of each level, and
  // per level 1/w_l and kappa w_l on the device; density components per point
  std::vector<const HarmonicLevel *> hl;
  std::vector<double> alpha;
  DBuf<double> hscale, khat;
  int nd = 1;
  DBuf<int32_t> node_leafslot;   // node -> position in leaf_nodes, -1 if not a real leaf
  // the flat tree
  DBuf<int32_t> cell;        // 3 x n SoA
  DBuf<int32_t> level, parent, child_first, plo, phi, col, perm, point_node;   // col: [k * n + i]
  DBuf<uint8_t> leaf, ghost, xlist, has_source;
  DBuf<double> sx, sy, sz;   // sorted points
  DBuf<double> centers;      // 3 x n SoA
  DBuf<float> sx32, sy32, sz32;
  // node lists per level (concatenated): childless sourced nodes, real leaves
  std::vector<int64_t> s2m_off, leaf_off;   // L + 1 offsets
  DBuf<int32_t> s2m_nodes, leaf_nodes;
  std::vector<LevelLists> lists;
  // expansions
  DBuf<double> Mp, Pl, Lc;   // Ns x n, Ns x n, Nl x nleaves
  DBuf<float> Lc32;
  DBuf<double> tau;          // sorted densities
  DBuf<float> tau32;
  // work buffers
  DBuf<double> pM, tup, Mup, tM, G, tdn, Gc, Yc;
  DBuf<float> Gc32, Yc32, pM32;
  DBuf<float> Mp32, tM32;  // the cube's mixed M2M: an FP32 mirror of Mp (written beside it), the parents' products
  // spectral M2L workspace
  DBuf<double> spec_S, spec_U;
  DBuf<float> spec_S32, spec_U32;
  std::vector<DBuf<int32_t>> spec_gptr, spec_cand;   // per level: sibling groups' far-source candidates (FP64)
  std::vector<DBuf<int32_t>> spec_gcol, spec_tiles;  // per level: groups' colleague children, active tiles (FP32)
  std::vector<int64_t> spec_ntiles;
  // per level and active tile: its sourced colleagues staged together (the boxes by
  // staging row, each (group, colleague)'s row); the tiles that fit and those that do not
  std::vector<DBuf<int32_t>> spec_ulist, spec_fit, spec_over;
  std::vector<DBuf<uint8_t>> spec_uslot;
  std::vector<int64_t> spec_nfit, spec_nover;
  // targets
  DBuf<double> tgt_sorted;
  DBuf<int32_t> tgt_node, tgt_leafslot, tgt_order, tgt_ptr;
  DBuf<uint32_t> tgt_keys;
  // stage timers: stages 0-7 one event pair each; the per-level stages (4: M2L,
  // 8: the Hadamard kernel of the spectral M2L) one pair per level
  hipEvent_t ev[2 * NSTAGES] = {};
  std::vector<hipEvent_t> lev_ev[2];     // per level: start, end
  std::vector<char> lev_recorded[2];     // per level: both events recorded by the last pass
  bool stage_recorded[NSTAGES] = {};
  bool expansions_ready = false;
  bool shrunk = false;       // the work buffers are released (plan_shrink)
  // the dense class M2L with low-rank class operators (null: the dense representatives),
  // and the products Bt_c G of one class of a chunk
  const ClassFactors *lowrank = nullptr;
  DBuf<double> Zc;
  DBuf<float> Zc32;
  // a captured pass sequence (plan_capture): the densities' gather from gin (input
  // order), the upward and downward passes and the evaluation at the sources into gout,
  // as one HIP graph; every buffer it touches stays where it is while it lives
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  DBuf<double> gin, gout;
  int g_nd = 0, g_output = -1, g_ncomp = 0;
  bool capturing = false;    // stage timers record no events into a capture
};

// ---- primitives (prims.hip) ----
// out[0] = v, on the device (no host copy)
void set_i32(Context &ctx, int32_t *out, int32_t v);
// stable LSD radix sort of (key, value) pairs on the bits [0, end_bit)
void sort_pairs_u64(Context &ctx, const uint64_t *kin, uint64_t *kout, const int32_t *vin, int32_t *vout,
                    int64_t n, int end_bit);
void sort_pairs_u32(Context &ctx, const uint32_t *kin, uint32_t *kout, const int32_t *vin, int32_t *vout,
                    int64_t n, int end_bit);
// out[i] = sum of in[0..i), i < n; returns the total (blocking read back)
int64_t exclusive_scan_i32(Context &ctx, const int32_t *in, int32_t *out, int64_t n);
void exclusive_scan_i32_async(Context &ctx, const int32_t *in, int32_t *out, int64_t n, int slot);
// indices i (ascending) with flag[i] != 0; returns their number (blocking read back)
int64_t compact_flags_u8(Context &ctx, const uint8_t *flag, int32_t *idx, int64_t n);
// pinned slots [first, first + n) set to -1 before the kernels that store them are queued
// (every earlier store to them read back or synchronised); read_slot polls one slot until
// the device has stored its value
void arm_slots(Context &ctx, int first, int n);
int64_t read_slot(Context &ctx, int slot);
// the count into the context's pinned slot `slot` (0 .. 62) by the stream, no synchronisation
void compact_flags_u8_async(Context &ctx, const uint8_t *flag, int32_t *idx, int64_t n, int slot);
void set_pinned(Context &ctx, int slot, int64_t v);
constexpr int NPINNED = 64;

// ---- tree (tree.hip) ----
void bounds(Context &ctx, const double *xyz, int64_t n, double lo[3], double hi[3], int *all_finite);
void build_tree(Plan &plan, const double *xyz);
void node_leafslots(Plan &plan);

// ---- lists (lists.hip) ----
void build_lists(Plan &plan);
void fill_level_pairs(Plan &plan, int l);   // src/far of level l, when not yet written
// far pairs per target chunk of the dense class M2L: its gathered columns and products
// (two buffers of one column per pair and block) take about 2 GiB.  The environment
// variable FFMM_M2L_CHUNK_PAIRS, read when a plan's lists are built, sets a smaller cap
// (at least 64 pairs), to bound the memory of many resident plans.
inline int64_t m2l_chunk_pairs(const Plan &p) {
  const OpCache &o = *p.ops;
  const int64_t col = o.method == METHOD_HARMONIC ? (int64_t)o.nb * o.hN : (int64_t)o.Ns;
  const int64_t bytes = 2 * col * (p.precision == FFMM_PREC_MIXED ? 4 : 8);
  int64_t cap = std::max<int64_t>(4096, (2ll << 30) / bytes);
  if (const char *e = std::getenv("FFMM_M2L_CHUNK_PAIRS")) {
    const long long v = std::atoll(e);
    if (v > 0) cap = std::max<int64_t>(64, std::min<int64_t>(cap, v));
  }
  return cap;
}

// ---- passes (passes.hip) ----
void allocate_pass_buffers(Plan &plan);
// a[0 .. n) = v on the context's stream (a kernel: no memset node in a captured graph)
void fill_f64(Context &ctx, double *a, int64_t n, double v);
void gather_inputs(Plan &plan, const double *tau, int sorted);
void upward_pass(Plan &plan);
void downward_pass(Plan &plan);
void m2l_dense_level(Plan &plan, int l);
void spectral_prepare(Plan &plan);
// the spectra of one level's sources and targets (the largest level's size)
void spectral_workspace(Plan &plan);
bool spectral_tile_counts(const Plan &plan, int l, int64_t &staged, int64_t &per_group);
void m2l_spectral_level(Plan &plan, int l);

// ---- direct sums (direct.hip) ----
void direct(Context &ctx, const double *src, const double *tau, int64_t ns, const double *tgt, int64_t nt,
            const int64_t *skip, double *out, int output);
// direct() with its partial sums in the library's fixed device array (no allocation;
// at most 2^20 targets); calls that overlap on one device share the array and race
void direct_noalloc(Context &ctx, const double *src, const double *tau, int64_t ns, const double *tgt, int64_t nt,
                    const int64_t *skip, double *out, int output);
// the device address of that array on `device`, looked up once (at context set-up)
void direct_scratch_setup(int device);
// output: FFMM_POTENTIAL (the velocity, 3 x nt) or FFMM_PRESSURE (nt)
void direct_stokes(Context &ctx, const double *src, const double *dens, int nd, int64_t ns, const double *tgt,
                   int64_t nt, const int64_t *skip, double *out, int output = FFMM_POTENTIAL);

// ---- evaluation (evaluate.hip) ----
void evaluate_at_sources(Plan &plan, double *out, int output);
// nodes: the leaf of each target (device, 0-based), or null to locate them on the device
void evaluate_targets(Plan &plan, const double *targets, int64_t nt, double *out, int output,
                      const int32_t *nodes = nullptr);

// the arguments of the evaluation kernels: per leaf slot, its targets (or its own
// sources), the far field of its leaf and the near field of its colleagues
struct EvalArgs {
  const int32_t *leaf_nodes;   // per leaf slot
  const int32_t *tptr;         // targets of leaf slot s: [tptr[s], tptr[s+1]); null: the leaf's own sources
  const double *tx, *ty, *tz;  // targets, grouped by leaf
  const int32_t *outidx;       // output position of each target
  const double *Lc;            // Nl x nleaves
  int Nl;
  const double *surf;          // cube: 3 x Nl SoA at the reference width
  const double *lam;
  const int32_t *level;
  const double *centers;
  int64_t nn;
  const int32_t *col, *plo, *phi;
  const double *sx, *sy, *sz, *tau;
  double *out;
  // harmonic methods
  HTab tb;
  HTab32 tb32;
  int p, nd;
  const double *hscale, *khat;
  double kappa;
};

// The near-field sources of a leaf as one list: the 27 colleagues' point ranges
// [lo[m], lo[m] + pre[m + 1] - pre[m]) back to back, so that they pass through the LDS
// tiles in full chunks (an average leaf's ~350 colleague sources in two chunks rather
// than one short tile per colleague, each behind a global-load round trip).
struct NearList {
  int32_t lo[NCOLL];
  int32_t pre[NCOLL + 1];
};

// built by the block's first wave (any block of at least 32 threads); synchronises the block
__device__ __forceinline__ void near_list(const EvalArgs &a, int i, NearList &nl) {
  const int t = threadIdx.x;
  int32_t lo = 0, cnt = 0;
  if (t < NCOLL) {
    const int32_t nb = a.col[(int64_t)t * a.nn + i];
    if (nb >= 0) {
      lo = a.plo[nb];
      cnt = max(0, a.phi[nb] - lo + 1);
    }
  }
  int32_t v = cnt;   // inclusive scan over lanes 0..31
#pragma unroll
  for (int off = 1; off < 32; off <<= 1) {
    const int32_t u = __shfl_up(v, off, 64);
    if (t >= off) v += u;
  }
  if (t < NCOLL) {
    nl.lo[t] = lo;
    nl.pre[t + 1] = v;
  }
  if (t == 0) nl.pre[0] = 0;
  __syncthreads();
}

// the source index of entry f of the list (0 <= f < pre[NCOLL]): the last m with pre[m] <= f
__device__ __forceinline__ int32_t near_source(const NearList &nl, int32_t f) {
  int m = 0, h = NCOLL;
  while (h - m > 1) {
    const int mid = (m + h) >> 1;
    if (nl.pre[mid] <= f) m = mid;
    else h = mid;
  }
  return nl.lo[m] + (f - nl.pre[m]);
}

// ---- a tree on the host (host_tree.hip) ----
// The flat tree in the plan's layout (cell: 3 x nn SoA; col: [k * nn + i]; sorted points
// SoA), before it is checked and uploaded.  level, has_source and point_node may be left
// empty: check_host_tree forms them (and checks them when given).
struct HostTree {
  int L = 0;
  int64_t nn = 0, N = 0;
  double origin[3] = {0, 0, 0}, W0 = 1.0;
  std::vector<int64_t> lo, nl;
  std::vector<int32_t> cell, level, parent, child_first, plo, phi, col, perm, point_node;
  std::vector<uint8_t> leaf, ghost, xlist, has_source;
  std::vector<double> sx, sy, sz;
};
// every invariant the kernels rely on; fails with FFMM_ERR_ARGUMENT ("<what>: ...")
void check_host_tree(HostTree &t, const char *what, int max_levels);
// the node lists of the passes (childless sourced nodes, real leaves) with per-level offsets
void host_node_lists(const HostTree &t, std::vector<int32_t> &s2m, std::vector<int64_t> &s2m_off,
                     std::vector<int32_t> &leaves, std::vector<int64_t> &leaf_off);
// a checked tree into the plan's device arrays, with its centres and node lists
void upload_host_tree(Plan &p, HostTree &t);
// the node centres from the cells and levels (tree.hip)
void compute_centers(Plan &p);

// ---- the life cycles (ffmm_abi.hip), shared by the C entry points and the server ----
void context_init(Context &c, int device, void *stream);
void context_release(Context &c);
// return a context's idle cached blocks to HIP, the largest first, until at most `keep`
// bytes stay cached (the bytes returned); the bytes it holds idle
size_t context_trim(Context &c, size_t keep);
size_t context_idle_bytes(Context &c);
// allocations so far that failed and returned the contexts' idle blocks
int64_t allocation_flushes();
void plan_init(Plan &p, Context &ctx, OpCache &o, const double *xyz, int64_t n, int ppl, const double origin[3],
               double width, int e0, int max_levels, int precision, int m2l);
void plan_release(Plan &p);
// a plan kept for later solves on the same points: the tree, lists and sorted points
// stay; the expansions and work buffers are released (the next plan_build allocates
// them again, from the block cache of the context it runs on)
void plan_shrink(Plan &p);
// capture the pass sequence of nd densities per point and `output` at the sources (one
// eager pass first, with `dens` (device or host, nd x N) when given, else with zeros:
// its output is then in gout); replay it: dens (nd x N, input order) to out (ncomp x N);
// drop it (the buffers may move again)
void plan_capture(Plan &p, int nd, int output, const double *dens = nullptr);
void plan_replay(Plan &p, const double *dens, double *out);
void plan_drop_graph(Plan &p);
// the output components per point of a plan's method (ffmm_evaluate_at_sources)
int plan_ncomp(const Plan &p, int output);
// the plan's tree as plan data (into buf when not null and cap suffices); its size in bytes
size_t plan_save(Plan &p, void *buf, size_t cap);
// a plan from plan data of a cache with the same operators: the tree uploaded, the rest rebuilt
void plan_restore(Plan &p, Context &ctx, OpCache &o, const void *buf, size_t size);
// a plan from a tree built on the host, checked in full before its upload
void plan_from_tree(Plan &p, Context &ctx, OpCache &o, HostTree &t, int ppl, int e0, int precision, int m2l);
void plan_build(Plan &p, const double *dens, int nd, int sorted);
void plan_eval_sources(Plan &p, double *out, int output);
void plan_eval_targets(Plan &p, const double *targets, int64_t nt, double *out, int output,
                       const int32_t *nodes = nullptr);
// device bytes held by a plan
size_t plan_bytes(const Plan &p);

// ---- harmonic methods (harmonic.hip) ----
void harmonic_s2m_level(Plan &plan, int l);
void launch_eval_harmonic(Plan &plan, const EvalArgs &a, int output);

// ---- timing ----
struct StageTimer {
  Plan &plan;
  int stage, level;
  StageTimer(Plan &p, int s, int l = 0);
  ~StageTimer();
};

}  // namespace ffmm

// the opaque handles of include/ffmm.h
struct ffmm_context_s {
  ffmm::Context c;
};
struct ffmm_opcache_s {
  ffmm::OpCache c;
};
struct ffmm_plan_s {
  ffmm::Plan c;
};
