// A tree handed over on the host -- plan data (ffmm_plan_load, ffmm_server_import) or a
// tree built by the host builder (ffmm_plan_from_tree) -- checked in full before any of
// it reaches the device.  Every index the kernels follow (parents, children, colleagues,
// point ranges, the permutation, the node lists) is proven in range here, and the
// structure the kernels assume is proven to hold, so that a corrupt or hostile tree is
// refused with an error and never becomes an out-of-bounds access.  Host code only.
#include <cmath>
#include <cstring>

#include "ffmm_internal.h"

namespace ffmm {

namespace {

[[noreturn]] void bad(const char *what, const std::string &msg) {
  fail(FFMM_ERR_ARGUMENT, std::string(what) + ": " + msg);
}

std::string at(int64_t i) { return " (node " + std::to_string(i) + ")"; }

}  // namespace

// The node lists of the passes (src/device_passes.jl:127-130): childless sourced nodes
// and real leaves, ascending, with their per-level offsets.
void host_node_lists(const HostTree &t, std::vector<int32_t> &s2m, std::vector<int64_t> &s2m_off,
                     std::vector<int32_t> &leaves, std::vector<int64_t> &leaf_off) {
  s2m.clear();
  leaves.clear();
  s2m_off.assign(t.L + 1, 0);
  leaf_off.assign(t.L + 1, 0);
  for (int l = 0; l < t.L; l++) {
    s2m_off[l] = (int64_t)s2m.size();
    leaf_off[l] = (int64_t)leaves.size();
    for (int64_t i = t.lo[l]; i < t.lo[l] + t.nl[l]; i++) {
      if (t.child_first[i] < 0 && t.has_source[i]) s2m.push_back((int32_t)i);
      if (t.leaf[i]) leaves.push_back((int32_t)i);
    }
  }
  s2m_off[t.L] = (int64_t)s2m.size();
  leaf_off[t.L] = (int64_t)leaves.size();
}

// The structural invariants of the reference's trees (src/tree/builder.jl; the device
// builder's, src/tree/device_builder.jl:398-501, are the same) that the kernels rely on.
// Fills the derived arrays (level, has_source, point_node) when they are empty, and
// checks them against the structure when they are given.
void check_host_tree(HostTree &t, const char *what, int max_levels) {
  const int L = t.L;
  const int64_t nn = t.nn, N = t.N;
  if (L < 1 || L > MAX_KEY_LEVELS + 1) bad(what, "the tree has " + std::to_string(L) + " levels (1 .. 22 supported)");
  if (max_levels > 0 && L > max_levels) bad(what, "the tree has more levels than max_levels");
  if (N < 1 || N >= INT32_MAX) bad(what, "the point count is out of range");
  if (nn < 1 || nn >= INT32_MAX / NCOLL) bad(what, "the node count is out of range");
  if ((int64_t)t.lo.size() != L || (int64_t)t.nl.size() != L) bad(what, "level tables of the wrong length");
  // level ranges: consecutive, covering every node, one root, whole sibling groups
  int64_t acc = 0;
  for (int l = 0; l < L; l++) {
    if (t.lo[l] != acc || t.nl[l] < 1) bad(what, "level " + std::to_string(l) + " does not follow the one before");
    if (l == 0 ? t.nl[0] != 1 : t.nl[l] % NCHILD != 0)
      bad(what, "level " + std::to_string(l) + " is not made of whole sibling groups");
    acc += t.nl[l];
  }
  if (acc != nn) bad(what, "the level ranges do not cover the nodes");
  const auto len_ok = [](size_t have, int64_t want) { return (int64_t)have == want; };
  if (!len_ok(t.cell.size(), 3 * nn) || !len_ok(t.parent.size(), nn) || !len_ok(t.child_first.size(), nn) ||
      !len_ok(t.plo.size(), nn) || !len_ok(t.phi.size(), nn) || !len_ok(t.col.size(), (int64_t)NCOLL * nn) ||
      !len_ok(t.leaf.size(), nn) || !len_ok(t.ghost.size(), nn) || !len_ok(t.xlist.size(), nn) ||
      !len_ok(t.perm.size(), N) || !len_ok(t.sx.size(), N) || !len_ok(t.sy.size(), N) || !len_ok(t.sz.size(), N))
    bad(what, "an array of the wrong length");
  const bool have_level = !t.level.empty(), have_hs = !t.has_source.empty(), have_pn = !t.point_node.empty();
  if ((have_level && !len_ok(t.level.size(), nn)) || (have_hs && !len_ok(t.has_source.size(), nn)) ||
      (have_pn && !len_ok(t.point_node.size(), N)))
    bad(what, "an array of the wrong length");
  std::vector<int32_t> lev(nn);
  for (int l = 0; l < L; l++)
    for (int64_t i = t.lo[l]; i < t.lo[l] + t.nl[l]; i++) lev[i] = l;
  if (have_level && std::memcmp(lev.data(), t.level.data(), nn * sizeof(int32_t)) != 0)
    bad(what, "node levels that do not match the level ranges");
  t.level = std::move(lev);
  std::vector<uint8_t> hs(nn);
  const int32_t *cx = t.cell.data(), *cy = cx + nn, *cz = cy + nn;
  for (int64_t i = 0; i < nn; i++) {
    const int l = t.level[i];
    const int64_t side = int64_t(1) << l;
    if (cx[i] < 0 || cy[i] < 0 || cz[i] < 0 || cx[i] >= side || cy[i] >= side || cz[i] >= side)
      bad(what, "a cell outside its level's grid" + at(i));
    if (t.leaf[i] > 1 || t.ghost[i] > 1 || t.xlist[i] > 1) bad(what, "a flag other than 0 or 1" + at(i));
    if (t.leaf[i] && t.ghost[i]) bad(what, "a node both leaf and ghost" + at(i));
    // point ranges: empty ranges are (plo, plo - 1)
    if (t.plo[i] < 0 || t.plo[i] > N || t.phi[i] < t.plo[i] - 1 || t.phi[i] > N - 1)
      bad(what, "a point range outside the points" + at(i));
    hs[i] = t.phi[i] >= t.plo[i];
    // the parent: the root has none; every other node lies in its parent's children
    const int32_t par = t.parent[i];
    if (l == 0) {
      if (par != -1 || t.plo[i] != 0 || t.phi[i] != N - 1) bad(what, "the root is not the root of all points");
    } else {
      if (par < t.lo[l - 1] || par >= t.lo[l - 1] + t.nl[l - 1]) bad(what, "a parent outside the level above" + at(i));
      const int64_t c = i - t.child_first[par];
      if (t.child_first[par] < 0 || c < 0 || c >= NCHILD) bad(what, "a node outside its parent's children" + at(i));
    }
    // the children: a whole sibling group on the next level, cells by the child slots
    // (bit d of the slot: the upper half along d), point ranges partitioning the node's
    const int32_t cf = t.child_first[i];
    if (cf >= 0) {
      if (l + 1 >= L || cf < t.lo[l + 1] || cf + NCHILD > t.lo[l + 1] + t.nl[l + 1] || (cf - t.lo[l + 1]) % NCHILD)
        bad(what, "children outside the level below" + at(i));
      if (t.ghost[i]) bad(what, "a ghost node with children" + at(i));
      int32_t next = t.plo[i];
      for (int c = 0; c < NCHILD; c++) {
        const int64_t ch = cf + c;
        if (t.parent[ch] != i) bad(what, "a child whose parent is another node" + at(ch));
        if (cx[ch] != 2 * cx[i] + (c & 1) || cy[ch] != 2 * cy[i] + ((c >> 1) & 1) || cz[ch] != 2 * cz[i] + (c >> 2))
          bad(what, "a child cell that is not its slot's" + at(ch));
        if (t.plo[ch] != next || t.phi[ch] < next - 1) bad(what, "children's point ranges that do not partition" + at(ch));
        next = t.phi[ch] + 1;
        if (t.ghost[ch] && !t.leaf[i]) bad(what, "ghost children of a node that is not a leaf" + at(ch));
      }
      if (next != t.phi[i] + 1) bad(what, "children's point ranges that do not partition" + at(i));
    } else if (!t.leaf[i] && !t.ghost[i]) {
      bad(what, "a childless node that is neither leaf nor ghost" + at(i));
    }
    // colleagues: slot k at the offset (k mod 3 - 1, (k / 3) mod 3 - 1, k / 9 - 1), the
    // same level, itself at the centre slot
    for (int k = 0; k < NCOLL; k++) {
      const int32_t j = t.col[(int64_t)k * nn + i];
      if (k == SELF_SLOT ? j != i : j != -1 && (j < t.lo[l] || j >= t.lo[l] + t.nl[l]))
        bad(what, "a colleague outside the node's level" + at(i));
      if (j >= 0 && (cx[j] - cx[i] != k % 3 - 1 || cy[j] - cy[i] != (k / 3) % 3 - 1 || cz[j] - cz[i] != k / 9 - 1))
        bad(what, "a colleague at the wrong offset" + at(i));
    }
  }
  if (have_hs && std::memcmp(hs.data(), t.has_source.data(), nn) != 0)
    bad(what, "source flags that do not match the point ranges");
  t.has_source = std::move(hs);
  // the permutation and the points
  {
    std::vector<uint8_t> seen(N, 0);
    for (int64_t j = 0; j < N; j++) {
      const int32_t v = t.perm[j];
      if (v < 0 || v >= N || seen[v]) bad(what, "perm is not a permutation of the points");
      seen[v] = 1;
    }
  }
  for (int64_t j = 0; j < N; j++) {
    const double x = t.sx[j], y = t.sy[j], z = t.sz[j];
    if (!(std::isfinite(x) && std::isfinite(y) && std::isfinite(z))) bad(what, "the points must be finite");
    if (!(t.origin[0] <= x && x <= t.origin[0] + t.W0 && t.origin[1] <= y && y <= t.origin[1] + t.W0 &&
          t.origin[2] <= z && z <= t.origin[2] + t.W0))
      bad(what, "points lie outside the requested box");
  }
  // the real leaf of every point: the leaves' ranges cover the points once
  std::vector<int32_t> pn(N, -1);
  for (int64_t i = 0; i < nn; i++)
    if (t.leaf[i])
      for (int32_t j = t.plo[i]; j <= t.phi[i]; j++) {
        if (pn[j] >= 0) bad(what, "a point in two leaves" + at(i));
        pn[j] = (int32_t)i;
      }
  for (int64_t j = 0; j < N; j++)
    if (pn[j] < 0) bad(what, "a point in no leaf");
  if (have_pn && std::memcmp(pn.data(), t.point_node.data(), N * sizeof(int32_t)) != 0)
    bad(what, "point nodes that do not match the leaves");
  t.point_node = std::move(pn);
}

// The checked tree into the plan's device arrays; the node centres and node lists formed
// from it (as the device builder forms them, so that both trees run identically).
void upload_host_tree(Plan &p, HostTree &t) {
  hipStream_t s = p.ctx->stream;
  p.L = t.L;
  p.nnodes = t.nn;
  p.lo = t.lo;
  p.nl = t.nl;
  p.cell.upload(t.cell.data(), t.cell.size(), s);
  p.level.upload(t.level.data(), t.level.size(), s);
  p.parent.upload(t.parent.data(), t.parent.size(), s);
  p.child_first.upload(t.child_first.data(), t.child_first.size(), s);
  p.plo.upload(t.plo.data(), t.plo.size(), s);
  p.phi.upload(t.phi.data(), t.phi.size(), s);
  p.col.upload(t.col.data(), t.col.size(), s);
  p.perm.upload(t.perm.data(), t.perm.size(), s);
  p.point_node.upload(t.point_node.data(), t.point_node.size(), s);
  p.leaf.upload(t.leaf.data(), t.leaf.size(), s);
  p.ghost.upload(t.ghost.data(), t.ghost.size(), s);
  p.xlist.upload(t.xlist.data(), t.xlist.size(), s);
  p.has_source.upload(t.has_source.data(), t.has_source.size(), s);
  p.sx.upload(t.sx.data(), t.sx.size(), s);
  p.sy.upload(t.sy.data(), t.sy.size(), s);
  p.sz.upload(t.sz.data(), t.sz.size(), s);
  std::vector<int32_t> s2m, leaves;
  host_node_lists(t, s2m, p.s2m_off, leaves, p.leaf_off);
  p.s2m_nodes.upload(s2m.data(), s2m.size(), s);
  p.leaf_nodes.upload(leaves.data(), leaves.size(), s);
  compute_centers(p);
  // the host vectors must outlive the asynchronous copies
  p.ctx->sync();
}

}  // namespace ffmm
