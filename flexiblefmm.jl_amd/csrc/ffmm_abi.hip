// The C ABI of libffmm_hip (include/ffmm.h): handles, error reporting, the operator
// cache upload and the plan's life cycle.  Every entry point catches the library's
// exceptions and turns them into a status code plus a thread-local message.
#include "ffmm_internal.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <memory>


namespace ffmm {

thread_local std::string g_last_error;
thread_local Context *g_alloc_ctx = nullptr;

void fail(int code, const std::string &msg) { throw Error(code, msg); }

// every live context, so that an allocation that fails can take back the idle blocks
// of all of them (a server's slots and the host's default context each keep their own)
static std::mutex g_ctx_mu;
static std::vector<Context *> g_contexts;
static std::atomic<int64_t> g_flushes{0};   // allocations that failed and returned idle blocks

int64_t allocation_flushes() { return g_flushes.load(); }

// return idle blocks of a context to HIP, the largest first, until at most `keep` bytes
// stay cached; the bytes returned
static size_t trim_locked(BlockCache &bc, size_t keep) {
  size_t freed = 0;
  while (bc.bytes > keep && !bc.blocks.empty()) {
    auto last = std::prev(bc.blocks.end());
    (void)hipFree(last->second);
    bc.bytes -= last->first;
    freed += last->first;
    bc.blocks.erase(last);
  }
  return freed;
}

size_t context_trim(Context &c, size_t keep) {
  std::lock_guard<std::mutex> lk(c.blocks->mu);
  return trim_locked(*c.blocks, keep);
}

size_t context_idle_bytes(Context &c) {
  std::lock_guard<std::mutex> lk(c.blocks->mu);
  return c.blocks->bytes;
}

static void flush_blocks(Context &c) { context_trim(c, 0); }

// the idle blocks of every context on `device` (hipFree waits for the device, so a block
// that another context's stream freed in stream order is idle when it is returned)
static size_t flush_device(int device) {
  std::lock_guard<std::mutex> lk(g_ctx_mu);
  size_t freed = 0;
  for (Context *c : g_contexts)
    if (c->device == device) freed += context_trim(*c, 0);
  return freed;
}

Context *device_alloc(void **p, size_t bytes, size_t *blk) {
  Context *c = g_alloc_ctx;
  if (c) {
    std::lock_guard<std::mutex> lk(c->blocks->mu);
    auto &bl = c->blocks->blocks;
    // the smallest idle block that fits and wastes at most half of itself (or 64 MiB)
    auto it = bl.lower_bound(bytes);
    if (it != bl.end() && it->first <= std::max(2 * bytes, bytes + ((size_t)64 << 20))) {
      *p = it->second;
      *blk = it->first;
      c->blocks->bytes -= it->first;
      bl.erase(it);
      return c;
    }
  }
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) {
    // the cached blocks may be what is missing: those of this context first, then those
    // of every context on the device; then try once more (returning only as many as the
    // allocation needs measured slower: far more failures, each waiting on the device)
    (void)hipGetLastError();
    int dev = 0;
    (void)hipGetDevice(&dev);
    g_flushes++;
    if (!(c && context_trim(*c, 0) >= bytes)) flush_device(dev);
    e = hipMalloc(p, bytes);
  }
  if (e != hipSuccess) {
    *p = nullptr;
    (void)hipGetLastError();
    fail(FFMM_ERR_MEMORY, "device allocation of " + std::to_string(bytes) + " bytes failed: " + hipGetErrorString(e));
  }
  *blk = bytes;
  return c;
}

void device_free(void *p, size_t blk, Context *owner) {
  if (!p) return;
  if (owner) {
    std::lock_guard<std::mutex> lk(owner->blocks->mu);
    BlockCache &bc = *owner->blocks;
    bc.blocks.emplace(blk, p);
    bc.bytes += blk;
    // keep at most the limit: return the largest idle blocks first
    trim_locked(bc, bc.limit);
  } else {
    (void)hipFree(p);
  }
}

void once_on_device(int device, int key, const std::function<void()> &upload) {
  static std::mutex mu;
  static std::vector<std::pair<int, int>> done;
  std::lock_guard<std::mutex> lk(mu);
  for (auto &d : done)
    if (d.first == device && d.second == key) return;
  upload();
  done.emplace_back(device, key);
}

static void set_device(Context &ctx) { FFMM_HIP(hipSetDevice(ctx.device)); }

void context_init(Context &c, int device, void *stream) {
  int ndev = 0;
  FFMM_HIP(hipGetDeviceCount(&ndev));
  if (device < 0 || device >= ndev) fail(FFMM_ERR_ARGUMENT, "device " + std::to_string(device) + " of " + std::to_string(ndev));
  c.device = device;
  set_device(c);
  if (stream) {
    c.stream = (hipStream_t)stream;
  } else {
    FFMM_HIP(hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking));
    c.own_stream = true;
  }
  FFMM_BLAS(rocblas_create_handle(&c.blas));
  FFMM_BLAS(rocblas_set_stream(c.blas, c.stream));
  FFMM_BLAS(rocblas_set_pointer_mode(c.blas, rocblas_pointer_mode_host));
  FFMM_HIP(hipHostMalloc((void **)&c.pinned, 64 * sizeof(int64_t), hipHostMallocMapped | hipHostMallocCoherent));
  FFMM_HIP(hipHostGetDevicePointer((void **)&c.pinned_dev, c.pinned, 0));
  // idle plan memory kept for reuse: up to half of the device's memory
  size_t fr = 0, tot = 0;
  FFMM_HIP(hipMemGetInfo(&fr, &tot));
  c.blocks->limit = tot / 2;
  // the device address of the allocation-free direct sum's array on this device
  direct_scratch_setup(device);
  std::lock_guard<std::mutex> lk(g_ctx_mu);
  g_contexts.push_back(&c);
}

void context_release(Context &c) {
  {
    std::lock_guard<std::mutex> lk(g_ctx_mu);
    g_contexts.erase(std::remove(g_contexts.begin(), g_contexts.end(), &c), g_contexts.end());
  }
  (void)hipSetDevice(c.device);
  if (c.stream) (void)hipStreamSynchronize(c.stream);
  c.tmp.release();
  flush_blocks(c);
  if (c.pinned) (void)hipHostFree(c.pinned);
  if (c.blas) rocblas_destroy_handle(c.blas);
  if (c.own_stream && c.stream) (void)hipStreamDestroy(c.stream);
  c.pinned = nullptr;
  c.blas = nullptr;
  c.stream = nullptr;
}

// Stage timers record event pairs on the plan's stream without synchronising; the
// times are read when asked for (every entry point has synchronised by then).  The
// per-level stages (M2L, the Hadamard kernel) keep one pair per level.
static hipEvent_t &stage_event(Plan &p, int stage, int level, int which) {
  const int sl = level_stage_slot(stage);
  if (sl < 0) return p.ev[2 * stage + which];
  std::vector<hipEvent_t> &v = p.lev_ev[sl];
  if ((int)v.size() < 2 * (level + 1)) {
    size_t old = v.size();
    v.resize(2 * (level + 1), nullptr);
    for (size_t k = old; k < v.size(); k++) FFMM_HIP(hipEventCreate(&v[k]));
  }
  return v[2 * level + which];
}
StageTimer::StageTimer(Plan &p, int s, int l) : plan(p), stage(s), level(l) {
  if (plan.capturing) return;
  FFMM_HIP(hipEventRecord(stage_event(plan, stage, level, 0), plan.ctx->stream));
}
StageTimer::~StageTimer() {
  if (plan.capturing) return;
  if (hipEventRecord(stage_event(plan, stage, level, 1), plan.ctx->stream) != hipSuccess) {
    (void)hipGetLastError();   // a failed record only loses the time: clear the error state
    return;
  }
  const int sl = level_stage_slot(stage);
  if (sl >= 0) {
    if ((int)plan.lev_recorded[sl].size() <= level) plan.lev_recorded[sl].resize(level + 1, 0);
    plan.lev_recorded[sl][level] = 1;
  } else {
    plan.stage_recorded[stage] = true;
  }
}

// the boundary points of the p^3 lattice in CartesianIndices order (coordinate 0
// fastest) and their linear index in the padded (2p)^3 lattice
// (src/operators/spectral_m2l.jl:103-115)
static std::vector<int32_t> cube_embed(int p) {
  std::vector<int32_t> e;
  const int n = 2 * p;
  for (int i3 = 0; i3 < p; i3++)
    for (int i2 = 0; i2 < p; i2++)
      for (int i1 = 0; i1 < p; i1++) {
        bool bnd = i1 == 0 || i1 == p - 1 || i2 == 0 || i2 == p - 1 || i3 == 0 || i3 == p - 1;
        if (bnd) e.push_back(i1 + n * i2 + n * n * i3);
      }
  return e;
}

}  // namespace ffmm

using namespace ffmm;

extern "C" {

const char *ffmm_last_error(void) { return g_last_error.c_str(); }
int ffmm_abi_version(void) { return 3; }

int ffmm_context_create(int device, void *stream, ffmm_context *out) {
  return guard([&] {
    if (!out) fail(FFMM_ERR_ARGUMENT, "null output handle");
    *out = nullptr;
    auto h = std::make_unique<ffmm_context_s>();
    context_init(h->c, device, stream);
    *out = h.release();
  });
}

int ffmm_context_destroy(ffmm_context ctx) {
  return guard([&] {
    if (!ctx) return;
    context_release(ctx->c);
    delete ctx;
  });
}

int ffmm_bounds(ffmm_context ctx, const double *xyz, int64_t n, double lo[3], double hi[3], int *all_finite) {
  return guard([&] {
    if (!ctx || !lo || !hi || !all_finite) fail(FFMM_ERR_ARGUMENT, "null argument");
    set_device(ctx->c);
    bounds(ctx->c, xyz, n, lo, hi, all_finite);
  });
}

int ffmm_direct(ffmm_context ctx, const double *src, const double *tau, int64_t ns, const double *tgt, int64_t nt,
                const int64_t *skip, double *out, int output) {
  return guard([&] {
    if (!ctx) fail(FFMM_ERR_ARGUMENT, "null context");
    if (ns < 0 || nt < 0) fail(FFMM_ERR_DIMENSION, "negative sizes");
    if (output < 0 || output > 2) fail(FFMM_ERR_ARGUMENT, "unknown output");
    if (nt == 0) return;
    if (!tgt || !out || (ns > 0 && (!src || !tau))) fail(FFMM_ERR_ARGUMENT, "null argument");
    set_device(ctx->c);
    direct(ctx->c, src, tau, ns, tgt, nt, skip, out, output);
    ctx->c.sync();
  });
}

int ffmm_direct_noalloc(ffmm_context ctx, const double *src, const double *tau, int64_t ns, const double *tgt,
                        int64_t nt, const int64_t *skip, double *out, int output) {
  return guard([&] {
    if (!ctx) fail(FFMM_ERR_ARGUMENT, "null context");
    if (ns < 0 || nt < 0) fail(FFMM_ERR_DIMENSION, "negative sizes");
    if (output < 0 || output > 2) fail(FFMM_ERR_ARGUMENT, "unknown output");
    if (nt == 0) return;
    if (!tgt || !out || (ns > 0 && (!src || !tau))) fail(FFMM_ERR_ARGUMENT, "null argument");
    set_device(ctx->c);
    direct_noalloc(ctx->c, src, tau, ns, tgt, nt, skip, out, output);
    ctx->c.sync();
  });
}

int ffmm_direct_stokes(ffmm_context ctx, const double *src, const double *dens, int nd, int64_t ns,
                       const double *tgt, int64_t nt, const int64_t *skip, double *out) {
  return guard([&] {
    if (!ctx) fail(FFMM_ERR_ARGUMENT, "null context");
    if (ns < 0 || nt < 0) fail(FFMM_ERR_DIMENSION, "negative sizes");
    if (nd != 3 && nd != 9) fail(FFMM_ERR_ARGUMENT, "Stokes densities are stokeslets (3) or stokeslets + stresslets (9)");
    if (nt == 0) return;
    if (!tgt || !out || (ns > 0 && (!src || !dens))) fail(FFMM_ERR_ARGUMENT, "null argument");
    set_device(ctx->c);
    direct_stokes(ctx->c, src, dens, nd, ns, tgt, nt, skip, out);
  });
}

int ffmm_direct_stokes_pressure(ffmm_context ctx, const double *src, const double *dens, int nd, int64_t ns,
                                const double *tgt, int64_t nt, const int64_t *skip, double *out) {
  return guard([&] {
    if (!ctx) fail(FFMM_ERR_ARGUMENT, "null context");
    if (ns < 0 || nt < 0) fail(FFMM_ERR_DIMENSION, "negative sizes");
    if (nd != 3 && nd != 9) fail(FFMM_ERR_ARGUMENT, "Stokes densities are stokeslets (3) or stokeslets + stresslets (9)");
    if (nt == 0) return;
    if (!tgt || !out || (ns > 0 && (!src || !dens))) fail(FFMM_ERR_ARGUMENT, "null argument");
    set_device(ctx->c);
    direct_stokes(ctx->c, src, dens, nd, ns, tgt, nt, skip, out, FFMM_PRESSURE);
  });
}

int ffmm_opcache_create(ffmm_context ctx, int p, int pc, double d, ffmm_opcache *out) {
  return guard([&] {
    if (!ctx || !out) fail(FFMM_ERR_ARGUMENT, "null argument");
    *out = nullptr;
    if (p < 2) fail(FFMM_ERR_ARGUMENT, "p must be at least 2");
    if (pc < p) fail(FFMM_ERR_ARGUMENT, "pc must be at least p");
    if (!(d > 0 && d < 2.0 / 3.0)) fail(FFMM_ERR_ARGUMENT, "d must lie in (0, 2/3)");
    auto h = std::make_unique<ffmm_opcache_s>();
    OpCache &o = h->c;
    o.ctx = &ctx->c;
    o.p = p;
    o.pc = pc;
    o.d = d;
    o.Ns = p * p * p - (p - 2) * (p - 2) * (p - 2);
    o.Nl = pc * pc * pc - (pc - 2) * (pc - 2) * (pc - 2);
    o.embed_h = cube_embed(p);
    *out = h.release();
  });
}

int ffmm_opcache_destroy(ffmm_opcache cache) {
  return guard([&] {
    if (!cache) return;
    (void)hipSetDevice(cache->c.ctx->device);
    delete cache;
  });
}

// the permutations of the class M2L with their signs folded into the top bit (signs are
// +1 or -1): gsel[f, i] = idx[f, i] | sign(sgn[f, i]); ssel[f, j] = inv[f, j] | sign(sgn_t[f, inv[f, j]]);
// fsel[f, i] = j | sign(sgn_t[f, i]) for inv[f, j] = i
static void upload_signed_tables(OpCache &o, const int32_t *idx, const int8_t *sgn, const int32_t *inv,
                                 const int8_t *sgn_t, int N, hipStream_t s) {
  const size_t n = (size_t)NFAR * N;
  std::vector<int32_t> g(n), t(n), fw(n);
  for (int f = 0; f < NFAR; f++)
    for (int i = 0; i < N; i++) {
      const size_t e = (size_t)f * N + i;
      if ((sgn[e] != 1 && sgn[e] != -1) || (sgn_t[e] != 1 && sgn_t[e] != -1))
        fail(FFMM_ERR_ARGUMENT, "signs must be +1 or -1");
      g[e] = idx[e] | (sgn[e] < 0 ? INT32_MIN : 0);
      const int32_t k = inv[e];
      t[e] = k | (sgn_t[(size_t)f * N + k] < 0 ? INT32_MIN : 0);
      fw[(size_t)f * N + k] = i | (sgn_t[(size_t)f * N + k] < 0 ? INT32_MIN : 0);
    }
  o.gsel.upload(g.data(), n, s);
  o.ssel.upload(t.data(), n, s);
  o.fsel.upload(fw.data(), n, s);
  FFMM_HIP(hipStreamSynchronize(s));   // before the host tables go
}

int ffmm_opcache_put_cube(ffmm_opcache cache, double w_ref, const double *small, const double *large, int r_up,
                          const double *Uh_up, const double *invS_up, const double *V_up, int r_dn,
                          const double *Uh_dn, const double *invS_dn, const double *V_dn, const double *M2M,
                          const double *L2L, int nclass, const double *class_ops, const int32_t *far_class,
                          const int32_t *idx, const int8_t *sgn, const double *tables) {
  return guard([&] {
    if (!cache) fail(FFMM_ERR_ARGUMENT, "null cache");
    OpCache &o = cache->c;
    Context &ctx = *o.ctx;
    set_device(ctx);
    if (!(w_ref > 0)) fail(FFMM_ERR_ARGUMENT, "the reference width must be positive");
    const int Ns = o.Ns, Nl = o.Nl;
    if (r_up < 1 || r_up > Ns || r_dn < 1 || r_dn > Ns) fail(FFMM_ERR_ARGUMENT, "pseudo-inverse ranks out of range");
    if (o.method != METHOD_CUBE) fail(FFMM_ERR_ARGUMENT, "the cache belongs to a harmonic method");
    if (nclass < 1 || nclass > MAX_CLASSES) fail(FFMM_ERR_ARGUMENT, "nclass out of range");
    if (!small || !large || !Uh_up || !invS_up || !V_up || !Uh_dn || !invS_dn || !V_dn || !M2M || !L2L ||
        !class_ops || !far_class || !idx || !sgn)
      fail(FFMM_ERR_ARGUMENT, "null operator");
    hipStream_t s = ctx.stream;
    o.w_ref = w_ref;
    o.r_up = r_up;
    o.r_dn = r_dn;
    o.nclass = nclass;
    o.small_h.assign(small, small + 3 * Ns);
    o.large_h.assign(large, large + 3 * Nl);
    std::vector<double> soa(3 * Nl);
    for (int k = 0; k < Nl; k++)
      for (int d = 0; d < 3; d++) soa[(size_t)d * Nl + k] = large[3 * k + d];
    o.large.upload(soa.data(), soa.size(), s);
    std::vector<float> soa32(soa.begin(), soa.end());
    o.large32.upload(soa32.data(), soa32.size(), s);
    o.Uh_up.upload(Uh_up, (size_t)r_up * Nl, s);
    std::vector<double> uhp;
    if (r_up > 240 && r_up <= 304) {
      const size_t nk = (size_t)(Nl + UC2E_PK - 1) / UC2E_PK * UC2E_PK + 4 * UC2E_PKX;
      uhp.assign(nk * UC2E_PROWS, 0.0);
      for (int k = 0; k < Nl; k++)
        for (int i = 0; i < r_up; i++) uhp[(size_t)k * UC2E_PROWS + i] = Uh_up[(size_t)k * r_up + i];
      o.Uh_up_pad.upload(uhp.data(), uhp.size(), s);
    }
    o.Uh_dn.upload(Uh_dn, (size_t)r_dn * Ns, s);
    std::vector<double> vs((size_t)Ns * r_up);
    for (int j = 0; j < r_up; j++)
      for (int i = 0; i < Ns; i++) vs[(size_t)j * Ns + i] = V_up[(size_t)j * Ns + i] * invS_up[j];
    o.Vs_up.upload(vs.data(), vs.size(), s);
    // operand copies of the column GEMMs (rows and k in the kernel's tiles), zero-padded
    auto pad_rows = [&](DBuf<double> &dst, const double *a, int rows, int K) {
      const size_t nk = (size_t)(K + UC2E_PK - 1) / UC2E_PK * UC2E_PK + 4 * UC2E_PKX;
      std::vector<double> ap(nk * UC2E_PROWS, 0.0);
      for (int k = 0; k < K; k++)
        for (int i = 0; i < rows; i++) ap[(size_t)k * UC2E_PROWS + i] = a[(size_t)k * rows + i];
      dst.upload(ap.data(), ap.size(), s);
    };
    if (Ns > 240 && Ns <= 304) pad_rows(o.Vs_up_pad, vs.data(), Ns, r_up);
    if (r_dn > 240 && r_dn <= 304) pad_rows(o.Uh_dn_pad, Uh_dn, r_dn, Ns);
    std::vector<double> vd((size_t)Nl * r_dn);
    for (int j = 0; j < r_dn; j++)
      for (int i = 0; i < Nl; i++) vd[(size_t)j * Nl + i] = V_dn[(size_t)j * Nl + i] * invS_dn[j];
    o.Vs_dn.upload(vd.data(), vd.size(), s);
    o.M2M.upload(M2M, (size_t)Ns * NCHILD * Ns, s);
    std::vector<float> m32(M2M, M2M + (size_t)Ns * NCHILD * Ns);
    o.M2M32.upload(m32.data(), m32.size(), s);
    o.L2L.upload(L2L, (size_t)NCHILD * Ns * Ns, s);
    o.class_ops.upload(class_ops, (size_t)nclass * Ns * Ns, s);
    std::vector<float> c32(class_ops, class_ops + (size_t)nclass * Ns * Ns);
    o.class_ops32.upload(c32.data(), c32.size(), s);
    o.far_class_h.assign(far_class, far_class + NFAR);
    for (int f = 0; f < NFAR; f++)
      if (far_class[f] < 0 || far_class[f] >= nclass) fail(FFMM_ERR_ARGUMENT, "far offset class out of range");
    o.far_class.upload(far_class, NFAR, s);
    std::vector<int32_t> inv((size_t)NFAR * Ns, -1);
    for (int f = 0; f < NFAR; f++)
      for (int i = 0; i < Ns; i++) {
        int32_t k = idx[(size_t)f * Ns + i];
        if (k < 0 || k >= Ns || inv[(size_t)f * Ns + k] >= 0) fail(FFMM_ERR_ARGUMENT, "idx is not a permutation");
        inv[(size_t)f * Ns + k] = i;
      }
    o.idx.upload(idx, (size_t)NFAR * Ns, s);
    o.inv.upload(inv.data(), inv.size(), s);
    o.sgn.upload(sgn, (size_t)NFAR * Ns, s);
    o.sgn_t.upload(sgn, (size_t)NFAR * Ns, s);   // the cube lattices: one permutation for both sides
    upload_signed_tables(o, idx, sgn, inv.data(), sgn, Ns, s);
    o.embed.upload(o.embed_h.data(), o.embed_h.size(), s);
    {
      // the p^3-corner index of each surface point, and the far index of every
      // (sibling a, child c of the parent's colleague k) pair: offset 2 o_k + b_c - b_a
      const int n = 2 * o.p;
      std::vector<int32_t> s8(Ns);
      for (int i = 0; i < Ns; i++) {
        int e = o.embed_h[i], i1 = e % n, i2 = (e / n) % n, i3 = e / (n * n);
        s8[i] = i1 + o.p * (i2 + o.p * i3);
      }
      o.sub8.upload(s8.data(), Ns, s);
      std::vector<int16_t> l8((size_t)o.p * o.p * o.p, (int16_t)-1);   // its inverse: -1 off the surface
      for (int i = 0; i < Ns; i++) l8[s8[i]] = (int16_t)i;
      o.lat8.upload(l8.data(), l8.size(), s);
      int far_index[343];
      for (int u = 0, nf = 0; u < 343; u++) {
        int x = u % 7 - 3, y = (u / 7) % 7 - 3, z = u / 49 - 3;
        far_index[u] = std::max(std::abs(x), std::max(std::abs(y), std::abs(z))) >= 2 ? nf++ : -1;
      }
      // one more row, all near: the FP64 Hadamard's table copy reads past the last candidate
      std::vector<int16_t> pf((NCOLL * NCHILD + 1) * NCHILD, (int16_t)-1);
      for (int k = 0; k < NCOLL; k++)
        for (int c = 0; c < NCHILD; c++)
          for (int a = 0; a < NCHILD; a++) {
            int ok[3] = {k % 3 - 1, (k / 3) % 3 - 1, k / 9 - 1};
            int u = 0;
            for (int d = 0, m = 1; d < 3; d++, m *= 7) u += (2 * ok[d] + ((c >> d) & 1) - ((a >> d) & 1) + 3) * m;
            pf[(k * NCHILD + c) * NCHILD + a] = (int16_t)far_index[u];
          }
      o.pairf.upload(pf.data(), pf.size(), s);
      ctx.sync();
    }
    o.has_tables = tables != nullptr;
    if (tables) {
      const int n = 2 * o.p;
      const size_t nfreq = (size_t)n * n * (o.p + 1);
      o.tables.upload(tables, 2 * nfreq * NFAR, s);
      std::vector<float> t32(tables, tables + 2 * nfreq * NFAR);
      o.tables32.upload(t32.data(), t32.size(), s);
    }
    ctx.sync();
    o.ready = true;
  });
}

int ffmm_opcache_create_harmonic(ffmm_context ctx, int kernel, int p, double kappa, double w_ref, ffmm_opcache *out) {
  return guard([&] {
    if (!ctx || !out) fail(FFMM_ERR_ARGUMENT, "null argument");
    *out = nullptr;
    if (kernel != FFMM_KERNEL_LAPLACE && kernel != FFMM_KERNEL_YUKAWA && kernel != FFMM_KERNEL_STOKESLET)
      fail(FFMM_ERR_ARGUMENT, "unknown kernel");
    if (p < 1 || p > 60) fail(FFMM_ERR_ARGUMENT, "p must lie in 1:60");
    if (!(w_ref > 0) || !std::isfinite(w_ref)) fail(FFMM_ERR_ARGUMENT, "the reference width must be positive");
    if (kernel == FFMM_KERNEL_YUKAWA && !(kappa >= 0 && std::isfinite(kappa)))
      fail(FFMM_ERR_ARGUMENT, "kappa must be finite and non-negative");
    set_device(ctx->c);
    auto h = std::make_unique<ffmm_opcache_s>();
    OpCache &o = h->c;
    o.ctx = &ctx->c;
    o.method = METHOD_HARMONIC;
    o.kernel = kernel;
    o.p = p;
    o.hN = (p + 1) * (p + 1);
    o.nb = kernel == FFMM_KERNEL_STOKESLET ? 4 : 1;
    o.Ns = o.Nl = o.nb * o.hN;
    o.kappa = kernel == FFMM_KERNEL_YUKAWA ? kappa : 0.0;
    o.w_ref = w_ref;
    // HarmonicTables (src/operators/harmonics.jl:163-187): a, b, cz, cm, cp at the
    // triangular indices, e_m = sqrt((2m - 1) / 2m)
    const int nt = (p + 1) * (p + 2) / 2;
    o.ntri = nt;
    std::vector<double> tab(5 * (size_t)nt + p, 0.0);
    double *a = tab.data(), *b = a + nt, *cz = b + nt, *cm = cz + nt, *cp = cm + nt, *e = cp + nt;
    for (int n = 0; n <= p; n++)
      for (int m = 0; m <= n; m++) {
        const int t = n * (n + 1) / 2 + m;
        const double c = std::sqrt((double)((n - m) * (n + m)));
        cz[t] = c;
        if (n > m) {
          a[t] = (2 * n - 1) / c;
          b[t] = std::sqrt((double)((n - m - 1) * (n + m - 1))) / c;
        }
        cm[t] = n + m >= 2 ? std::sqrt((double)((n + m) * (n + m - 1))) : 0.0;
        cp[t] = n - m >= 2 ? std::sqrt((double)((n - m) * (n - m - 1))) : 0.0;
      }
    for (int m = 1; m <= p; m++) e[m - 1] = std::sqrt((2.0 * m - 1) / (2.0 * m));
    o.htab.upload(tab.data(), tab.size(), ctx->c.stream);
    const std::vector<float> tab32(tab.begin(), tab.end());
    o.htab32.upload(tab32.data(), tab32.size(), ctx->c.stream);
    // the far offsets in far-index order (far_offsets, src/util.jl:87-97), for the Stokes coupling
    std::vector<double> offs;
    for (int u = 0; u < 343; u++) {
      int x = u % 7 - 3, y = (u / 7) % 7 - 3, z = u / 49 - 3;
      if (std::max(std::abs(x), std::max(std::abs(y), std::abs(z))) >= 2) {
        offs.push_back(x);
        offs.push_back(y);
        offs.push_back(z);
      }
    }
    o.offs.upload(offs.data(), offs.size(), ctx->c.stream);
    ctx->c.sync();
    *out = h.release();
  });
}

int ffmm_opcache_put_harmonic_classes(ffmm_opcache cache, int nclass, const int32_t *far_class, const int32_t *idx,
                                      const int8_t *sgn, const int32_t *idx_t, const int8_t *sgn_t) {
  return guard([&] {
    if (!cache) fail(FFMM_ERR_ARGUMENT, "null cache");
    OpCache &o = cache->c;
    if (o.method != METHOD_HARMONIC) fail(FFMM_ERR_ARGUMENT, "the cache belongs to the cube method");
    if (!far_class || !idx || !sgn || !idx_t || !sgn_t) fail(FFMM_ERR_ARGUMENT, "null argument");
    if (nclass < 1 || nclass > MAX_CLASSES) fail(FFMM_ERR_ARGUMENT, "nclass out of range");
    Context &ctx = *o.ctx;
    set_device(ctx);
    const int N = o.hN;
    for (int f = 0; f < NFAR; f++)
      if (far_class[f] < 0 || far_class[f] >= nclass) fail(FFMM_ERR_ARGUMENT, "far offset class out of range");
    std::vector<int32_t> inv((size_t)NFAR * N, -1);
    std::vector<char> seen((size_t)NFAR * N, 0);
    for (int f = 0; f < NFAR; f++)
      for (int i = 0; i < N; i++) {
        int32_t k = idx[(size_t)f * N + i], kt = idx_t[(size_t)f * N + i];
        if (k < 0 || k >= N || seen[(size_t)f * N + k]) fail(FFMM_ERR_ARGUMENT, "idx is not a permutation");
        seen[(size_t)f * N + k] = 1;
        if (kt < 0 || kt >= N || inv[(size_t)f * N + kt] >= 0) fail(FFMM_ERR_ARGUMENT, "idx_t is not a permutation");
        inv[(size_t)f * N + kt] = i;
        int8_t s1 = sgn[(size_t)f * N + i], s2 = sgn_t[(size_t)f * N + i];
        if ((s1 != 1 && s1 != -1) || (s2 != 1 && s2 != -1)) fail(FFMM_ERR_ARGUMENT, "signs must be +1 or -1");
      }
    hipStream_t s = ctx.stream;
    o.nclass = nclass;
    o.far_class_h.assign(far_class, far_class + NFAR);
    o.far_class.upload(far_class, NFAR, s);
    o.idx.upload(idx, (size_t)NFAR * N, s);
    o.sgn.upload(sgn, (size_t)NFAR * N, s);
    o.inv.upload(inv.data(), inv.size(), s);
    o.sgn_t.upload(sgn_t, (size_t)NFAR * N, s);
    upload_signed_tables(o, idx, sgn, inv.data(), sgn_t, N, s);
    ctx.sync();
    o.has_classes = true;
    o.ready = true;
  });
}

int ffmm_opcache_put_harmonic_level(ffmm_opcache cache, double width, const double *M2M, const double *L2L,
                                    const double *reps) {
  return guard([&] {
    if (!cache) fail(FFMM_ERR_ARGUMENT, "null cache");
    OpCache &o = cache->c;
    if (o.method != METHOD_HARMONIC) fail(FFMM_ERR_ARGUMENT, "the cache belongs to the cube method");
    if (!o.has_classes) fail(FFMM_ERR_STATE, "put the classes first");
    if (!M2M || !L2L || !reps) fail(FFMM_ERR_ARGUMENT, "null operator");
    if (!(width > 0) || !std::isfinite(width)) fail(FFMM_ERR_ARGUMENT, "the width must be positive");
    Context &ctx = *o.ctx;
    set_device(ctx);
    hipStream_t s = ctx.stream;
    const int N = o.hN, NB = o.Ns;
    auto lv = std::make_unique<HarmonicLevel>();
    lv->w = width;
    if (o.nb == 1) {
      lv->M2M.upload(M2M, (size_t)N * NCHILD * N, s);
      lv->L2L.upload(L2L, (size_t)NCHILD * N * N, s);
    } else {
      // _stokes_dense_operators (src/operators/harmonics_stokes.jl:112-131): for child j
      // with centre offset d (in parent widths), blocks a = 1..3 of the parent get A_j M_a
      // and the fourth A_j (M_4 / 2 + d . M); L2L mirrors it with B_j (2 L_4 - 2 d . L)
      std::vector<double> Md((size_t)NB * NCHILD * NB, 0.0), Ld((size_t)NCHILD * NB * NB, 0.0);
      const int ldm = NB, ldl = NCHILD * NB;
      for (int j = 0; j < NCHILD; j++) {
        const double dh[3] = {((j & 1) - 0.5) / 2, (((j >> 1) & 1) - 0.5) / 2, (((j >> 2) & 1) - 0.5) / 2};
        const int cj = j * NB;
        for (int c = 0; c < N; c++)
          for (int r = 0; r < N; r++) {
            const double A = M2M[(size_t)(j * N + c) * N + r];           // M2M[r, jN + c]
            const double B = L2L[(size_t)c * (NCHILD * N) + j * N + r];  // L2L[jN + r, c]
            for (int a = 0; a < 3; a++) {
              Md[(size_t)(cj + a * N + c) * ldm + a * N + r] = A;
              Ld[(size_t)(a * N + c) * ldl + cj + a * N + r] = B;
              Md[(size_t)(cj + a * N + c) * ldm + 3 * N + r] = dh[a] * A;
              Ld[(size_t)(a * N + c) * ldl + cj + 3 * N + r] = (-2 * dh[a]) * B;
            }
            Md[(size_t)(cj + 3 * N + c) * ldm + 3 * N + r] = A / 2;
            Ld[(size_t)(3 * N + c) * ldl + cj + 3 * N + r] = 2 * B;
          }
      }
      lv->M2M.upload(Md.data(), Md.size(), s);
      lv->L2L.upload(Ld.data(), Ld.size(), s);
    }
    lv->reps.upload(reps, (size_t)o.nclass * N * N, s);
    std::vector<float> r32(reps, reps + (size_t)o.nclass * N * N);   // the mixed form's class GEMMs
    lv->reps32.upload(r32.data(), r32.size(), s);
    std::vector<float> rr;   // held until the copies are done
    if (N <= 128) {
      rr.assign((size_t)o.nclass * 128 * 128, 0.f);
      for (int c = 0; c < o.nclass; c++)
        for (int k = 0; k < N; k++)
          for (int i = 0; i < N; i++) rr[((size_t)c * 128 + i) * 128 + k] = r32[(size_t)c * N * N + i + (size_t)N * k];
      lv->reps32r.upload(rr.data(), rr.size(), s);
    }
    ctx.sync();
    std::lock_guard<std::mutex> lk(o.hmu);
    for (auto &h : o.hlev)
      if (h->w == width) return;   // the operators of a width are fixed once put
    o.hlev.push_back(std::move(lv));
  });
}

int ffmm_opcache_put_class_factors(ffmm_opcache cache, double tol, const int32_t *ranks, const double *A,
                                   const double *Bt) {
  return guard([&] {
    if (!cache || !ranks || !A || !Bt) fail(FFMM_ERR_ARGUMENT, "null argument");
    OpCache &o = cache->c;
    if (!o.ready || o.method != METHOD_CUBE)
      fail(FFMM_ERR_ARGUMENT, "low-rank class operators belong to the cube method's operators");
    if (!(tol > 0 && tol < 1)) fail(FFMM_ERR_ARGUMENT, "tol must lie in (0, 1)");
    auto f = std::make_unique<ClassFactors>();
    f->tol = tol;
    int64_t tot = 0;
    for (int c = 0; c < o.nclass; c++) {
      if (ranks[c] < 1 || ranks[c] > o.Ns) fail(FFMM_ERR_ARGUMENT, "a class rank lies outside 1:Ns");
      f->rank.push_back(ranks[c]);
      f->off.push_back(tot);
      f->kmax = std::max(f->kmax, ranks[c]);
      tot += (int64_t)ranks[c] * o.Ns;
    }
    Context &ctx = *o.ctx;
    set_device(ctx);
    hipStream_t s = ctx.stream;
    f->A.upload(A, tot, s);
    f->Bt.upload(Bt, tot, s);
    std::vector<float> a32(A, A + tot), b32(Bt, Bt + tot);
    f->A32.upload(a32.data(), tot, s);
    f->Bt32.upload(b32.data(), tot, s);
    ctx.sync();
    std::lock_guard<std::mutex> lk(o.factors_mu);
    for (auto &g : o.factors)
      if (g->tol == tol) return;   // the factors of a tolerance are fixed once put
    o.factors.push_back(std::move(f));
  });
}

int ffmm_opcache_has_width(ffmm_opcache cache, double width, int *has) {
  return guard([&] {
    if (!cache || !has) fail(FFMM_ERR_ARGUMENT, "null argument");
    *has = cache->c.find_width(width) != nullptr;
  });
}

}  // extern "C"

// ---- the plan's life cycle (shared by the C entry points and the server) ----
namespace ffmm {

// the checks of a plan's parameters against its cache (shared by creation and restore)
static void plan_check(Context &ctx, OpCache &o, int64_t n, int ppl, double width, int e0, int precision, int m2l) {
  if (!o.ready) fail(FFMM_ERR_STATE, "the operator cache holds no operators");
  if (n <= 0) fail(FFMM_ERR_ARGUMENT, "no points");
  if (ppl < 1) fail(FFMM_ERR_ARGUMENT, "ppl must be positive");
  if (!(width > 0)) fail(FFMM_ERR_ARGUMENT, "box width must be positive");
  if (precision != FFMM_PREC_F64 && precision != FFMM_PREC_MIXED) fail(FFMM_ERR_ARGUMENT, "unknown precision");
  if (m2l != FFMM_M2L_DENSE && m2l != FFMM_M2L_SPECTRAL) fail(FFMM_ERR_ARGUMENT, "unknown M2L form");
  if (o.method == METHOD_HARMONIC) {
    if (m2l != FFMM_M2L_DENSE) fail(FFMM_ERR_ARGUMENT, "the harmonic methods take the dense class M2L");
    if (o.kernel != FFMM_KERNEL_YUKAWA && !o.find_width(o.w_ref))
      fail(FFMM_ERR_STATE, "the cache holds no operators of the reference width");
  }
  if (m2l == FFMM_M2L_SPECTRAL && !o.has_tables) fail(FFMM_ERR_ARGUMENT, "the cache holds no spectral tables");
  if (std::ldexp(o.w_ref, e0) != width) fail(FFMM_ERR_ARGUMENT, "the box width is not w_ref * 2^e0");
  if (ctx.device != o.ctx->device) fail(FFMM_ERR_ARGUMENT, "the cache lives on another device");
}

static void plan_fields(Plan &p, Context &ctx, OpCache &o, int64_t n, int ppl, const double origin[3], double width,
                        int e0, int max_levels, int precision, int m2l) {
  p.ctx = &ctx;
  p.ops = &o;
  p.N = n;
  p.ppl = ppl;
  p.max_levels = max_levels;
  p.precision = precision;
  p.m2l = m2l;
  for (int d = 0; d < 3; d++) p.origin[d] = origin[d];
  p.W0 = width;
  p.e0 = e0;
  for (auto &e : p.ev) FFMM_HIP(hipEventCreate(&e));
}

// everything after the tree: level scales, the harmonic levels, node slots, lists, pass
// buffers, the spectral tiles
static void plan_finish(Plan &p) {
  Context &ctx = *p.ctx;
  OpCache &o = *p.ops;
  p.lam.resize(p.L);
  for (int l = 0; l < p.L; l++) p.lam[l] = std::ldexp(1.0, p.e0 - l);
  p.lam_d.upload(p.lam.data(), p.L, ctx.stream);
  if (o.method == METHOD_HARMONIC) {
    // per level: 1/w_l and kappa w_l (the extras' scale and kappa-hat, src/operators/harmonics.jl:327-336),
    // the M2L factor (rescale, src/operators/harmonics.jl:393-398; 1 for Yukawa, whose
    // operators are built per width), and the operators of the level's width
    std::vector<double> hs(p.L), kh(p.L);
    p.alpha.assign(p.L, 1.0);
    p.hl.assign(p.L, nullptr);
    for (int l = 0; l < p.L; l++) {
      const double w = std::ldexp(p.W0, -l);
      hs[l] = 1.0 / w;
      kh[l] = o.kappa * w;
      if (o.kernel != FFMM_KERNEL_YUKAWA) {
        p.alpha[l] = 1.0 / p.lam[l];
        p.hl[l] = o.find_width(o.w_ref);
      }
    }
    p.hscale.upload(hs.data(), p.L, ctx.stream);
    p.khat.upload(kh.data(), p.L, ctx.stream);
  }
  node_leafslots(p);
  {
    StageTimer tm(p, 1);
    build_lists(p);
  }
  allocate_pass_buffers(p);
  if (p.m2l == FFMM_M2L_SPECTRAL) spectral_prepare(p);
  ctx.sync();
}

void plan_init(Plan &p, Context &ctx, OpCache &o, const double *xyz, int64_t n, int ppl, const double origin[3],
               double width, int e0, int max_levels, int precision, int m2l) {
  plan_check(ctx, o, n, ppl, width, e0, precision, m2l);
  if (!xyz) fail(FFMM_ERR_ARGUMENT, "null points");
  set_device(ctx);
  AllocScope scope(ctx);
  plan_fields(p, ctx, o, n, ppl, origin, width, e0, max_levels, precision, m2l);
  {
    StageTimer tm(p, 0);
    build_tree(p, xyz);
  }
  plan_finish(p);
}

// ---- plan data: the tree of a plan, to rebuild the plan without its tree build ----
// Layout (host byte order): the magic, the version, the cache's identity (method, kernel,
// p, nb, Ns, Nl, w_ref, kappa), the plan's parameters and sizes, the per-level node
// offsets and node-list offsets, then each array as an int64 count and its elements.
static constexpr char PLAN_MAGIC[8] = {'F', 'F', 'M', 'M', 'T', 'R', 'E', 'E'};
static constexpr uint32_t PLAN_VERSION = 1;

struct PlanWriter {
  uint8_t *buf;
  size_t cap, n = 0;
  void put(const void *src, size_t b) {
    if (buf && n + b <= cap) std::memcpy(buf + n, src, b);
    n += b;
  }
  template <class T>
  void val(T v) { put(&v, sizeof v); }
  template <class T>
  void host(const std::vector<T> &v) {
    val<int64_t>((int64_t)v.size());
    put(v.data(), v.size() * sizeof(T));
  }
  template <class T>
  void dev(Context &ctx, const DBuf<T> &d, int64_t count) {
    val<int64_t>(count);
    const size_t b = (size_t)count * sizeof(T);
    if (buf && n + b <= cap && count > 0) FFMM_HIP(hipMemcpyAsync(buf + n, d.p, b, hipMemcpyDeviceToHost, ctx.stream));
    n += b;
  }
};

struct PlanReader {
  const uint8_t *buf;
  size_t size, n = 0;
  void get(void *dst, size_t b) {
    if (b > size - n) fail(FFMM_ERR_ARGUMENT, "plan data truncated");
    std::memcpy(dst, buf + n, b);
    n += b;
  }
  template <class T>
  T val() {
    T v;
    get(&v, sizeof v);
    return v;
  }
  // an array of `expect` elements into a host vector
  template <class T>
  void host(std::vector<T> &v, int64_t expect) {
    const int64_t c = val<int64_t>();
    if (c != expect) fail(FFMM_ERR_ARGUMENT, "plan data: an array of the wrong length");
    if ((uint64_t)c > (size - n) / sizeof(T)) fail(FFMM_ERR_ARGUMENT, "plan data truncated");
    v.resize(c);
    get(v.data(), c * sizeof(T));
  }
};

static void plan_tree_arrays(PlanWriter &io, Context &ctx, Plan &p) {
  const int64_t nn = p.nnodes, N = p.N;
  io.dev(ctx, p.cell, 3 * nn);
  io.dev(ctx, p.level, nn);
  io.dev(ctx, p.parent, nn);
  io.dev(ctx, p.child_first, nn);
  io.dev(ctx, p.plo, nn);
  io.dev(ctx, p.phi, nn);
  io.dev(ctx, p.col, (int64_t)NCOLL * nn);
  io.dev(ctx, p.perm, N);
  io.dev(ctx, p.point_node, N);
  io.dev(ctx, p.leaf, nn);
  io.dev(ctx, p.ghost, nn);
  io.dev(ctx, p.xlist, nn);
  io.dev(ctx, p.has_source, nn);
  io.dev(ctx, p.sx, N);
  io.dev(ctx, p.sy, N);
  io.dev(ctx, p.sz, N);
  io.dev(ctx, p.centers, 3 * nn);
  io.dev(ctx, p.s2m_nodes, p.s2m_off[p.L]);
  io.dev(ctx, p.leaf_nodes, p.leaf_off[p.L]);
}

// the plan's tree into buf (when buf is not null and cap is large enough); the size in bytes
size_t plan_save(Plan &p, void *buf, size_t cap) {
  Context &ctx = *p.ctx;
  OpCache &o = *p.ops;
  set_device(ctx);
  PlanWriter w{static_cast<uint8_t *>(buf), cap};
  w.put(PLAN_MAGIC, sizeof PLAN_MAGIC);
  w.val<uint32_t>(PLAN_VERSION);
  w.val<int32_t>(o.method); w.val<int32_t>(o.kernel); w.val<int32_t>(o.p); w.val<int32_t>(o.nb);
  w.val<int32_t>(o.Ns); w.val<int32_t>(o.Nl); w.val<double>(o.w_ref); w.val<double>(o.kappa);
  w.val<int64_t>(p.N); w.val<int32_t>(p.ppl); w.val<int32_t>(p.max_levels); w.val<int32_t>(p.precision);
  w.val<int32_t>(p.m2l);
  for (int d = 0; d < 3; d++) w.val<double>(p.origin[d]);
  w.val<double>(p.W0); w.val<int32_t>(p.e0); w.val<int32_t>(p.L); w.val<int64_t>(p.nnodes);
  w.host(p.lo); w.host(p.nl); w.host(p.s2m_off); w.host(p.leaf_off);
  plan_tree_arrays(w, ctx, p);
  ctx.sync();
  return w.n;
}

// The header and tree of plan data made by plan_save, checked in full on the host:
// every array's length, the tree's structure (check_host_tree) and the stored node lists
// against the tree's.  No device work; nothing of the data is trusted before this.
struct PlanHeader {
  int method, kernel, p, nb, Ns, Nl;
  double w_ref, kappa;
  int64_t N;
  int ppl, max_levels, precision, m2l, e0;
};

static void parse_plan_data(const void *buf, size_t size, PlanHeader &h, HostTree &t) {
  if (!buf) fail(FFMM_ERR_ARGUMENT, "null plan data");
  PlanReader r{static_cast<const uint8_t *>(buf), size};
  char magic[8];
  r.get(magic, sizeof magic);
  if (std::memcmp(magic, PLAN_MAGIC, sizeof magic) != 0) fail(FFMM_ERR_ARGUMENT, "not plan data");
  if (r.val<uint32_t>() != PLAN_VERSION) fail(FFMM_ERR_ARGUMENT, "plan data of another version");
  h.method = r.val<int32_t>(); h.kernel = r.val<int32_t>(); h.p = r.val<int32_t>(); h.nb = r.val<int32_t>();
  h.Ns = r.val<int32_t>(); h.Nl = r.val<int32_t>();
  h.w_ref = r.val<double>(); h.kappa = r.val<double>();
  h.N = r.val<int64_t>();
  h.ppl = r.val<int32_t>(); h.max_levels = r.val<int32_t>(); h.precision = r.val<int32_t>();
  h.m2l = r.val<int32_t>();
  for (int d = 0; d < 3; d++) t.origin[d] = r.val<double>();
  t.W0 = r.val<double>();
  h.e0 = r.val<int32_t>();
  t.L = r.val<int32_t>();
  t.nn = r.val<int64_t>();
  t.N = h.N;
  const int L = t.L;
  const int64_t nn = t.nn, N = t.N;
  if (L < 1 || L > MAX_KEY_LEVELS + 1 || nn < 1 || nn >= INT32_MAX / NCOLL || N < 1 || N >= INT32_MAX)
    fail(FFMM_ERR_ARGUMENT, "plan data: bad tree sizes");
  if (!(t.W0 > 0) || !std::isfinite(t.W0) || !std::isfinite(t.origin[0]) || !std::isfinite(t.origin[1]) ||
      !std::isfinite(t.origin[2]))
    fail(FFMM_ERR_ARGUMENT, "plan data: a bad box");
  std::vector<int64_t> s2m_off, leaf_off;
  r.host(t.lo, L);
  r.host(t.nl, L);
  r.host(s2m_off, L + 1);
  r.host(leaf_off, L + 1);
  if (s2m_off[L] < 0 || s2m_off[L] > nn || leaf_off[L] < 0 || leaf_off[L] > nn)
    fail(FFMM_ERR_ARGUMENT, "plan data: bad tree sizes");
  std::vector<double> centers;
  std::vector<int32_t> s2m, leaves;
  r.host(t.cell, 3 * nn);
  r.host(t.level, nn);
  r.host(t.parent, nn);
  r.host(t.child_first, nn);
  r.host(t.plo, nn);
  r.host(t.phi, nn);
  r.host(t.col, (int64_t)NCOLL * nn);
  r.host(t.perm, N);
  r.host(t.point_node, N);
  r.host(t.leaf, nn);
  r.host(t.ghost, nn);
  r.host(t.xlist, nn);
  r.host(t.has_source, nn);
  r.host(t.sx, N);
  r.host(t.sy, N);
  r.host(t.sz, N);
  r.host(centers, 3 * nn);        // formed again from the cells on upload
  r.host(s2m, s2m_off[L]);
  r.host(leaves, leaf_off[L]);
  if (r.n != size) fail(FFMM_ERR_ARGUMENT, "plan data: trailing bytes");
  check_host_tree(t, "plan data", h.max_levels);
  std::vector<int32_t> s2m_h, leaves_h;
  std::vector<int64_t> s2m_off_h, leaf_off_h;
  host_node_lists(t, s2m_h, s2m_off_h, leaves_h, leaf_off_h);
  if (s2m_h != s2m || leaves_h != leaves || s2m_off_h != s2m_off || leaf_off_h != leaf_off)
    fail(FFMM_ERR_ARGUMENT, "plan data: node lists that do not match the tree");
}

// a plan from plan data made by plan_save with a cache of the same operators
void plan_restore(Plan &p, Context &ctx, OpCache &o, const void *buf, size_t size) {
  PlanHeader h;
  HostTree t;
  parse_plan_data(buf, size, h, t);
  if (h.method != o.method || h.kernel != o.kernel || h.p != o.p || h.nb != o.nb || h.Ns != o.Ns || h.Nl != o.Nl ||
      h.w_ref != o.w_ref || h.kappa != o.kappa)
    fail(FFMM_ERR_ARGUMENT, "the plan data was made with another operator cache (method, order, kernel or width)");
  plan_check(ctx, o, h.N, h.ppl, t.W0, h.e0, h.precision, h.m2l);
  set_device(ctx);
  AllocScope scope(ctx);
  plan_fields(p, ctx, o, h.N, h.ppl, t.origin, t.W0, h.e0, h.max_levels, h.precision, h.m2l);
  upload_host_tree(p, t);
  plan_finish(p);
}

// a plan from a tree built on the host (ffmm_plan_from_tree)
void plan_from_tree(Plan &p, Context &ctx, OpCache &o, HostTree &t, int ppl, int e0, int precision, int m2l) {
  plan_check(ctx, o, t.N, ppl, t.W0, e0, precision, m2l);
  check_host_tree(t, "the tree", 0);
  set_device(ctx);
  AllocScope scope(ctx);
  plan_fields(p, ctx, o, t.N, ppl, t.origin, t.W0, e0, 30, precision, m2l);
  upload_host_tree(p, t);
  plan_finish(p);
}

void plan_release(Plan &p) {
  if (!p.ctx) return;
  (void)hipSetDevice(p.ctx->device);
  (void)hipStreamSynchronize(p.ctx->stream);
  plan_drop_graph(p);
  for (auto &e : p.ev)
    if (e) (void)hipEventDestroy(e);
  for (auto &v : p.lev_ev)
    for (auto &e : v)
      if (e) (void)hipEventDestroy(e);
  for (auto &e : p.ev) e = nullptr;
  p.lev_ev[0].clear();
  p.lev_ev[1].clear();
}

void plan_drop_graph(Plan &p) {
  if (p.gexec) (void)hipGraphExecDestroy(p.gexec);
  if (p.graph) (void)hipGraphDestroy(p.graph);
  p.gexec = nullptr;
  p.graph = nullptr;
  p.gin.release();
  p.gout.release();
  p.g_nd = 0;
  p.g_output = -1;
  p.g_ncomp = 0;
}

int plan_ncomp(const Plan &p, int output) {
  const OpCache &o = *p.ops;
  if (o.method == METHOD_HARMONIC && o.kernel == FFMM_KERNEL_STOKESLET) return output == FFMM_PRESSURE ? 1 : 3;
  return output == FFMM_POTENTIAL ? 1 : output == FFMM_GRADIENT ? 3 : 4;
}

void plan_capture(Plan &p, int nd, int output, const double *dens) {
  set_device(*p.ctx);
  if (p.ctx->sync(), p.gexec) plan_drop_graph(p);
  AllocScope scope(*p.ctx);
  const int nc = plan_ncomp(p, output);
  p.gin.resize((size_t)nd * p.N);
  p.gout.resize((size_t)nc * p.N);
  if (dens)
    FFMM_HIP(hipMemcpyAsync(p.gin.p, dens, sizeof(double) * nd * p.N, hipMemcpyDefault, p.ctx->stream));
  else
    fill_f64(*p.ctx, p.gin.p, (int64_t)nd * p.N, 0.0);
  // one eager pass: its checks, the buffers it allocates and the libraries' workspaces
  plan_build(p, p.gin.p, nd, 0);
  plan_eval_sources(p, p.gout.p, output);
  hipStream_t s = p.ctx->stream;
  hipGraph_t g = nullptr;
  p.capturing = true;
  FFMM_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
  try {
    gather_inputs(p, p.gin.p, 0);
    upward_pass(p);
    downward_pass(p);
    evaluate_at_sources(p, p.gout.p, output);
  } catch (...) {
    (void)hipStreamEndCapture(s, &g);
    if (g) (void)hipGraphDestroy(g);
    (void)hipGetLastError();
    p.capturing = false;
    throw;
  }
  p.capturing = false;
  const hipError_t e = hipStreamEndCapture(s, &g);
  if (e != hipSuccess || !g) {
    (void)hipGetLastError();
    if (g) (void)hipGraphDestroy(g);
    fail(FFMM_ERR_DEVICE, std::string("the pass sequence could not be captured: ") + hipGetErrorString(e));
  }
  hipGraphExec_t x = nullptr;
  const hipError_t ei = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
  if (ei != hipSuccess) {
    (void)hipGetLastError();
    (void)hipGraphDestroy(g);
    fail(FFMM_ERR_DEVICE, std::string("the captured pass sequence could not be instantiated: ") +
                              hipGetErrorString(ei));
  }
  p.graph = g;
  p.gexec = x;
  p.g_nd = nd;
  p.g_output = output;
  p.g_ncomp = nc;
}

void plan_replay(Plan &p, const double *dens, double *out) {
  if (!p.gexec) fail(FFMM_ERR_STATE, "the plan holds no captured pass sequence");
  set_device(*p.ctx);
  hipStream_t s = p.ctx->stream;
  FFMM_HIP(hipMemcpyAsync(p.gin.p, dens, sizeof(double) * p.g_nd * p.N, hipMemcpyDefault, s));
  FFMM_HIP(hipGraphLaunch(p.gexec, s));
  FFMM_HIP(hipMemcpyAsync(out, p.gout.p, sizeof(double) * p.g_ncomp * p.N, hipMemcpyDefault, s));
  p.ctx->sync();
  p.nd = p.g_nd;
  p.expansions_ready = true;
}

void plan_shrink(Plan &p) {
  if (p.ctx) FFMM_HIP(hipStreamSynchronize(p.ctx->stream));
  plan_drop_graph(p);
  for (DBuf<double> *b : {&p.Mp, &p.Pl, &p.Lc, &p.tau, &p.pM, &p.tup, &p.Mup, &p.tM, &p.G, &p.tdn, &p.Gc, &p.Yc,
                          &p.spec_S, &p.spec_U, &p.tgt_sorted})
    b->release();
  for (DBuf<float> *b : {&p.Lc32, &p.tau32, &p.Gc32, &p.Yc32, &p.pM32, &p.Mp32, &p.tM32, &p.spec_S32, &p.spec_U32})
    b->release();
  for (DBuf<int32_t> *b : {&p.tgt_node, &p.tgt_leafslot, &p.tgt_order, &p.tgt_ptr}) b->release();
  p.tgt_keys.release();
  p.Zc.release();
  p.Zc32.release();
  p.expansions_ready = false;
  p.shrunk = true;
}

void plan_build(Plan &p, const double *tau, int nd, int sorted) {
  OpCache &o = *p.ops;
  const bool stokes = o.method == METHOD_HARMONIC && o.kernel == FFMM_KERNEL_STOKESLET;
  if (stokes && nd != 3 && nd != 9)
    fail(FFMM_ERR_ARGUMENT, "StokesHarmonics takes stokeslets (3 components) or stokeslets and stresslets (9)");
  // SphericalHarmonics(Laplace3D()) also takes a charge and a dipole per point
  // (supports_density, src/operators/harmonics.jl:317-320)
  const bool dipoles = o.method == METHOD_HARMONIC && o.kernel == FFMM_KERNEL_LAPLACE;
  if (!stokes && nd != 1 && !(dipoles && nd == 4))
    fail(FFMM_ERR_ARGUMENT, dipoles ? "SphericalHarmonics(Laplace3D()) takes charges (1 per point) or charges and "
                                      "dipoles (4 per point: q, d)"
                                    : "this method takes one charge per point");
  set_device(*p.ctx);
  AllocScope scope(*p.ctx);
  if (o.method == METHOD_HARMONIC && o.kernel == FFMM_KERNEL_YUKAWA) {
    // the operators of every level with a far field or a translation (2 .. L-1)
    for (int l = 2; l < p.L; l++) {
      const double w = std::ldexp(p.W0, -l);
      p.hl[l] = o.find_width(w);
      if (!p.hl[l])
        fail(FFMM_ERR_STATE, "the cache holds no operators for boxes of width " + std::to_string(w) + " (level " +
                                 std::to_string(l) + ")");
    }
  }
  p.nd = nd;
  p.expansions_ready = false;
  if (p.shrunk) {
    allocate_pass_buffers(p);
    if (p.m2l == FFMM_M2L_SPECTRAL) spectral_workspace(p);
    p.shrunk = false;
  }
  {
    StageTimer tm(p, 7);
    gather_inputs(p, tau, sorted);
  }
  {
    StageTimer tm(p, 2);
    upward_pass(p);
  }
  for (auto &v : p.lev_recorded) std::fill(v.begin(), v.end(), 0);
  {
    StageTimer tm(p, 3);
    downward_pass(p);
  }
  p.ctx->sync();
  p.expansions_ready = true;
}

static void check_output(Plan &p, int output) {
  if (!p.expansions_ready) fail(FFMM_ERR_STATE, "build the expansions first");
  if (output < 0 || output > 3) fail(FFMM_ERR_ARGUMENT, "unknown output");
  const bool stokes = p.ops->kernel == FFMM_KERNEL_STOKESLET;
  if (stokes && output != FFMM_POTENTIAL && output != FFMM_PRESSURE)
    fail(FFMM_ERR_ARGUMENT, "StokesHarmonics evaluates the velocity (Potential) or the pressure on the device");
  if (!stokes && output == FFMM_PRESSURE) fail(FFMM_ERR_ARGUMENT, "Pressure is an output of the Stokes kernels");
}

void plan_eval_sources(Plan &p, double *out, int output) {
  check_output(p, output);
  set_device(*p.ctx);
  AllocScope scope(*p.ctx);
  {
    StageTimer tm(p, 5);
    evaluate_at_sources(p, out, output);
  }
  p.ctx->sync();
}

void plan_eval_targets(Plan &p, const double *targets, int64_t nt, double *out, int output, const int32_t *nodes) {
  check_output(p, output);
  if (nt < 0) fail(FFMM_ERR_DIMENSION, "negative target count");
  if (nt == 0) return;
  if (!targets || !out) fail(FFMM_ERR_ARGUMENT, "null argument");
  set_device(*p.ctx);
  AllocScope scope(*p.ctx);
  {
    StageTimer tm(p, 6);
    evaluate_targets(p, targets, nt, out, output, nodes);
  }
  p.ctx->sync();
}

}  // namespace ffmm

extern "C" {

int ffmm_plan_create(ffmm_opcache cache, const double *xyz, int64_t n, int ppl, const double origin[3], double width,
                     int e0, int max_levels, int precision, int m2l, ffmm_plan *out) {
  return guard([&] {
    if (!cache || !out || !origin) fail(FFMM_ERR_ARGUMENT, "null argument");
    *out = nullptr;
    OpCache &o = cache->c;
    auto h = std::make_unique<ffmm_plan_s>();
    plan_init(h->c, *o.ctx, o, xyz, n, ppl, origin, width, e0, max_levels, precision, m2l);
    *out = h.release();
  });
}

int ffmm_plan_destroy(ffmm_plan plan) {
  return guard([&] {
    if (!plan) return;
    plan_release(plan->c);
    delete plan;
  });
}

int ffmm_plan_save(ffmm_plan plan, void *buf, int64_t cap, int64_t *size) {
  return guard([&] {
    if (!plan || !size) fail(FFMM_ERR_ARGUMENT, "null argument");
    if (cap < 0) fail(FFMM_ERR_ARGUMENT, "negative capacity");
    const size_t need = plan_save(plan->c, nullptr, 0);
    *size = (int64_t)need;
    if (buf) {
      if ((size_t)cap < need) fail(FFMM_ERR_ARGUMENT, "the buffer is smaller than the plan data");
      plan_save(plan->c, buf, (size_t)cap);
    }
  });
}

int ffmm_plan_load(ffmm_opcache cache, const void *buf, int64_t size, ffmm_plan *out) {
  return guard([&] {
    if (!cache || !buf || !out) fail(FFMM_ERR_ARGUMENT, "null argument");
    if (size < 0) fail(FFMM_ERR_ARGUMENT, "negative size");
    *out = nullptr;
    OpCache &o = cache->c;
    auto h = std::make_unique<ffmm_plan_s>();
    plan_restore(h->c, *o.ctx, o, buf, (size_t)size);
    *out = h.release();
  });
}

int ffmm_plan_data_check(const void *buf, int64_t size) {
  return guard([&] {
    if (size < 0) fail(FFMM_ERR_ARGUMENT, "negative size");
    PlanHeader h;
    HostTree t;
    parse_plan_data(buf, (size_t)size, h, t);
  });
}

int ffmm_plan_from_tree(ffmm_opcache cache, int nlevels, const int64_t *level_start, int64_t npoints,
                        const int32_t *cell, const int32_t *parent, const int32_t *child_first, const int32_t *plo,
                        const int32_t *phi, const uint8_t *leaf, const uint8_t *ghost, const uint8_t *xlist,
                        const int32_t *colleagues, const int32_t *perm, const double *sorted_points, int ppl,
                        const double origin[3], double width, int e0, int precision, int m2l, ffmm_plan *out) {
  return guard([&] {
    if (!cache || !out || !origin || !level_start || !cell || !parent || !child_first || !plo || !phi || !leaf ||
        !ghost || !xlist || !colleagues || !perm || !sorted_points)
      fail(FFMM_ERR_ARGUMENT, "null argument");
    *out = nullptr;
    if (nlevels < 1 || nlevels > MAX_KEY_LEVELS + 1)
      fail(FFMM_ERR_ARGUMENT, "the tree has " + std::to_string(nlevels) + " levels (1 .. 22 supported)");
    if (level_start[0] != 0) fail(FFMM_ERR_ARGUMENT, "the tree: level 0 does not start at node 0");
    const int64_t nn = level_start[nlevels];
    if (nn < 1 || nn >= INT32_MAX / NCOLL) fail(FFMM_ERR_ARGUMENT, "the tree: the node count is out of range");
    if (npoints < 1 || npoints >= INT32_MAX) fail(FFMM_ERR_ARGUMENT, "the tree: the point count is out of range");
    HostTree t;
    t.L = nlevels;
    t.nn = nn;
    t.N = npoints;
    for (int d = 0; d < 3; d++) t.origin[d] = origin[d];
    t.W0 = width;
    for (int l = 0; l < nlevels; l++) {
      t.lo.push_back(level_start[l]);
      t.nl.push_back(level_start[l + 1] - level_start[l]);
    }
    t.cell.resize(3 * nn);
    t.col.resize((size_t)NCOLL * nn);
    for (int64_t i = 0; i < nn; i++) {
      for (int d = 0; d < 3; d++) t.cell[d * nn + i] = cell[3 * i + d];
      for (int k = 0; k < NCOLL; k++) t.col[(size_t)k * nn + i] = colleagues[(size_t)NCOLL * i + k];
    }
    t.parent.assign(parent, parent + nn);
    t.child_first.assign(child_first, child_first + nn);
    t.plo.assign(plo, plo + nn);
    t.phi.assign(phi, phi + nn);
    t.leaf.assign(leaf, leaf + nn);
    t.ghost.assign(ghost, ghost + nn);
    t.xlist.assign(xlist, xlist + nn);
    t.perm.assign(perm, perm + npoints);
    t.sx.resize(npoints);
    t.sy.resize(npoints);
    t.sz.resize(npoints);
    for (int64_t j = 0; j < npoints; j++) {
      t.sx[j] = sorted_points[3 * j];
      t.sy[j] = sorted_points[3 * j + 1];
      t.sz[j] = sorted_points[3 * j + 2];
    }
    OpCache &o = cache->c;
    auto h = std::make_unique<ffmm_plan_s>();
    plan_from_tree(h->c, *o.ctx, o, t, ppl, e0, precision, m2l);
    *out = h.release();
  });
}

int ffmm_plan_sizes(ffmm_plan plan, int *nlevels, int64_t *nnodes, int64_t *npoints, int64_t *level_start,
                    int64_t *level_count, int64_t *level_pairs, int64_t *level_s2m, int64_t *level_leaves) {
  return guard([&] {
    if (!plan) fail(FFMM_ERR_ARGUMENT, "null plan");
    Plan &p = plan->c;
    if (nlevels) *nlevels = p.L;
    if (nnodes) *nnodes = p.nnodes;
    if (npoints) *npoints = p.N;
    for (int l = 0; l < p.L; l++) {
      if (level_start) level_start[l] = p.lo[l];
      if (level_count) level_count[l] = p.nl[l];
      if (level_pairs) level_pairs[l] = p.lists[l].np;
      if (level_s2m) level_s2m[l] = p.s2m_off[l + 1] - p.s2m_off[l];
      if (level_leaves) level_leaves[l] = p.leaf_off[l + 1] - p.leaf_off[l];
    }
  });
}

int ffmm_plan_download_tree(ffmm_plan plan, int32_t *cell, int32_t *level, int32_t *parent, int32_t *child_first,
                            int32_t *plo, int32_t *phi, uint8_t *leaf, uint8_t *ghost, uint8_t *xlist,
                            uint8_t *has_source, int32_t *colleagues, int32_t *perm, int32_t *point_node,
                            double *sorted_points) {
  return guard([&] {
    if (!plan) fail(FFMM_ERR_ARGUMENT, "null plan");
    Plan &p = plan->c;
    Context &ctx = *p.ctx;
    set_device(ctx);
    hipStream_t s = ctx.stream;
    const int64_t nn = p.nnodes, N = p.N;
    if (cell) {
      std::vector<int32_t> c(3 * nn);
      p.cell.download(c.data(), c.size(), s);
      ctx.sync();
      for (int64_t i = 0; i < nn; i++)
        for (int d = 0; d < 3; d++) cell[3 * i + d] = c[d * nn + i];
    }
    if (colleagues) {
      std::vector<int32_t> c((size_t)NCOLL * nn);
      p.col.download(c.data(), c.size(), s);
      ctx.sync();
      for (int64_t i = 0; i < nn; i++)
        for (int k = 0; k < NCOLL; k++) colleagues[NCOLL * i + k] = c[(size_t)k * nn + i];
    }
    if (level) p.level.download(level, nn, s);
    if (parent) p.parent.download(parent, nn, s);
    if (child_first) p.child_first.download(child_first, nn, s);
    if (plo) p.plo.download(plo, nn, s);
    if (phi) p.phi.download(phi, nn, s);
    if (leaf) p.leaf.download(leaf, nn, s);
    if (ghost) p.ghost.download(ghost, nn, s);
    if (xlist) p.xlist.download(xlist, nn, s);
    if (has_source) p.has_source.download(has_source, nn, s);
    if (perm) p.perm.download(perm, N, s);
    if (point_node) p.point_node.download(point_node, N, s);
    if (sorted_points) {
      std::vector<double> x(N), y(N), z(N);
      p.sx.download(x.data(), N, s);
      p.sy.download(y.data(), N, s);
      p.sz.download(z.data(), N, s);
      ctx.sync();
      for (int64_t j = 0; j < N; j++) {
        sorted_points[3 * j] = x[j];
        sorted_points[3 * j + 1] = y[j];
        sorted_points[3 * j + 2] = z[j];
      }
    }
    ctx.sync();
  });
}

int ffmm_plan_perm(ffmm_plan plan, int32_t *perm_dev) {
  return guard([&] {
    if (!plan || !perm_dev) fail(FFMM_ERR_ARGUMENT, "null argument");
    Plan &p = plan->c;
    set_device(*p.ctx);
    FFMM_HIP(hipMemcpyAsync(perm_dev, p.perm.p, sizeof(int32_t) * p.N, hipMemcpyDeviceToDevice, p.ctx->stream));
    p.ctx->sync();
  });
}

int ffmm_plan_download_lists(ffmm_plan plan, int level, int32_t *ptr, int32_t *src, int32_t *far) {
  return guard([&] {
    if (!plan) fail(FFMM_ERR_ARGUMENT, "null plan");
    Plan &p = plan->c;
    if (level < 0 || level >= p.L) fail(FFMM_ERR_ARGUMENT, "level out of range");
    set_device(*p.ctx);
    AllocScope scope(*p.ctx);
    LevelLists &li = p.lists[level];
    hipStream_t s = p.ctx->stream;
    if (ptr) li.ptr.download(ptr, p.nl[level] + 1, s);
    if ((src || far) && li.np) fill_level_pairs(p, level);
    if (src && li.np) li.src.download(src, li.np, s);
    if (far && li.np) li.far.download(far, li.np, s);
    p.ctx->sync();
  });
}

int ffmm_build_expansions(ffmm_plan plan, const double *tau, int sorted) {
  return ffmm_build_expansions_nd(plan, tau, 1, sorted);
}

int ffmm_build_expansions_nd(ffmm_plan plan, const double *tau, int nd, int sorted) {
  return guard([&] {
    if (!plan || !tau) fail(FFMM_ERR_ARGUMENT, "null argument");
    plan_build(plan->c, tau, nd, sorted);
  });
}

int ffmm_evaluate_at_sources(ffmm_plan plan, double *out, int output) {
  return guard([&] {
    if (!plan || !out) fail(FFMM_ERR_ARGUMENT, "null argument");
    plan_eval_sources(plan->c, out, output);
  });
}

int ffmm_evaluate(ffmm_plan plan, const double *targets, int64_t nt, double *out, int output) {
  return guard([&] {
    if (!plan) fail(FFMM_ERR_ARGUMENT, "null plan");
    plan_eval_targets(plan->c, targets, nt, out, output);
  });
}

int ffmm_evaluate_nodes(ffmm_plan plan, const double *targets, int64_t nt, const int32_t *nodes, double *out,
                        int output) {
  return guard([&] {
    if (!plan) fail(FFMM_ERR_ARGUMENT, "null plan");
    if (nt > 0 && !nodes) fail(FFMM_ERR_ARGUMENT, "null nodes");
    plan_eval_targets(plan->c, targets, nt, out, output, nodes);
  });
}

int ffmm_matvec(ffmm_plan plan, const double *tau, int sorted, double *out, int output) {
  int rc = ffmm_build_expansions(plan, tau, sorted);
  if (rc != FFMM_OK) return rc;
  return ffmm_evaluate_at_sources(plan, out, output);
}

int ffmm_plan_download_expansions(ffmm_plan plan, double *multipoles, double *plocals, double *leaf_locals) {
  return guard([&] {
    if (!plan) fail(FFMM_ERR_ARGUMENT, "null plan");
    Plan &p = plan->c;
    set_device(*p.ctx);
    hipStream_t s = p.ctx->stream;
    if (multipoles) p.Mp.download(multipoles, (size_t)p.ops->Ns * p.nnodes, s);
    if (plocals) p.Pl.download(plocals, (size_t)p.ops->Ns * p.nnodes, s);
    if (leaf_locals) p.Lc.download(leaf_locals, (size_t)p.ops->Nl * p.leaf_off[p.L], s);
    p.ctx->sync();
  });
}

int ffmm_plan_capture(ffmm_plan plan, int nd, int output) {
  return guard([&] {
    if (!plan) fail(FFMM_ERR_ARGUMENT, "null plan");
    if (output < 0 || output > 3) fail(FFMM_ERR_ARGUMENT, "unknown output");
    plan_capture(plan->c, nd, output);
  });
}

int ffmm_plan_replay(ffmm_plan plan, const double *dens, double *out) {
  return guard([&] {
    if (!plan || !dens || !out) fail(FFMM_ERR_ARGUMENT, "null argument");
    plan_replay(plan->c, dens, out);
  });
}

int ffmm_plan_set_tol(ffmm_plan plan, double tol) {
  return guard([&] {
    if (!plan) fail(FFMM_ERR_ARGUMENT, "null plan");
    Plan &p = plan->c;
    OpCache &o = *p.ops;
    if (!(tol >= 0 && tol < 1)) fail(FFMM_ERR_ARGUMENT, "tol must lie in [0, 1)");
    const ClassFactors *lr = nullptr;
    if (tol > 0) {
      if (o.method != METHOD_CUBE || p.m2l != FFMM_M2L_DENSE)
        fail(FFMM_ERR_ARGUMENT, "tol > 0 takes the cube method with the dense class M2L");
      std::lock_guard<std::mutex> lk(o.factors_mu);
      for (auto &g : o.factors)
        if (g->tol == tol) lr = g.get();
      if (!lr) fail(FFMM_ERR_STATE, "the cache holds no low-rank class operators for this tol");
    }
    set_device(*p.ctx);
    p.ctx->sync();
    plan_drop_graph(p);   // the buffers may move
    AllocScope scope(*p.ctx);
    p.lowrank = lr;
    p.Zc.release();
    p.Zc32.release();
    p.expansions_ready = false;
    if (!p.shrunk) allocate_pass_buffers(p);
  });
}

int ffmm_plan_spectral_tiles(ffmm_plan plan, int64_t *staged, int64_t *per_group, int n) {
  return guard([&] {
    if (!plan || !staged || !per_group) fail(FFMM_ERR_ARGUMENT, "null argument");
    const Plan &p = plan->c;
    for (int l = 0; l < n; l++) {
      staged[l] = per_group[l] = 0;
      if (l < p.L) spectral_tile_counts(p, l, staged[l], per_group[l]);
    }
  });
}

int ffmm_plan_stage_times(ffmm_plan plan, double *ms, int n) {
  return guard([&] {
    if (!plan || !ms) fail(FFMM_ERR_ARGUMENT, "null argument");
    Plan &p = plan->c;
    set_device(*p.ctx);
    p.ctx->sync();
    for (int i = 0; i < n && i < NSTAGES; i++) {
      float t = 0.f;
      double v = 0.0;
      const int sl = level_stage_slot(i);
      if (sl >= 0) {
        const std::vector<hipEvent_t> &e = p.lev_ev[sl];
        for (size_t l = 0; l < p.lev_recorded[sl].size(); l++)
          if (p.lev_recorded[sl][l] && hipEventElapsedTime(&t, e[2 * l], e[2 * l + 1]) == hipSuccess) v += t;
      } else if (p.stage_recorded[i] && hipEventElapsedTime(&t, p.ev[2 * i], p.ev[2 * i + 1]) == hipSuccess) {
        v = t;
      }
      ms[i] = v;
    }
    (void)hipGetLastError();   // a failed query only loses a time: leave no error state behind
  });
}

}  // extern "C"
