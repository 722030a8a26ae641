// mesh.cpp -- block-quadtree generator for the Poisson workloads.
//
// Produces the leaf blocks of a 2:1-balanced quadtree over a domain of
// bpdx x bpdy level-0 blocks, each block holding 8x8 cells, ordered along a
// Hilbert curve.  It stands in for the part of Cubism's grid that decides
// which blocks exist (the reference solver reads that from
// cubism::BlockInfo::{level,index}, ExpAMRSolver.cpp:186-249) and is used to
// build the synthetic meshes the solver is measured and tested on.
//
// Refinement rule: a block is split while it lies within band[l] of a disk,
// where band[l] is the distance allowed for refining level l to l+1 and the
// distance is measured from the block rectangle to the filled disk.  The
// result is then balanced so that face- and corner-adjacent leaves differ by
// at most one level, which the coarse-fine rows of the operator assume
// (ExpAMRSolver.cpp:177).
//
// Adaptation (cup2d_mesh_adapt) changes an existing mesh from per-block
// tags, the decision Cubism's MeshAdaptation takes for AdaptTheMesh
// (AdaptTheMesh.cpp:73-122; the tagging criterion is computed on the GPU):
// blocks tagged for refinement are split (below the top level); four
// sibling leaves all tagged for compression are merged when the parent would
// keep the 2:1 balance with the refined mesh; then the tree is balanced
// again.  Every new block records where its values come from: an old block
// unchanged, an old ancestor (refinement, one level or more), or four old
// children (compression).

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <unordered_map>
#include <unordered_set>
#include <vector>

#include "../../include/cup2d_mesh.h"

namespace {

struct Key {
  static uint64_t make(int l, long long i, long long j) {
    return (uint64_t(l) << 56) | (uint64_t(i) << 28) | uint64_t(j);
  }
  static int level(uint64_t k) { return int(k >> 56); }
  static long long i(uint64_t k) { return (long long)((k >> 28) & 0xFFFFFFF); }
  static long long j(uint64_t k) { return (long long)(k & 0xFFFFFFF); }
};

struct Gen {
  int bpdx, bpdy, levels;
  bool perx, pery;
  double s0;  // level-0 block side
  std::unordered_set<uint64_t> leaves;

  long long nx(int l) const { return (long long)bpdx << l; }
  long long ny(int l) const { return (long long)bpdy << l; }

  void split(uint64_t k) {
    const int l = Key::level(k);
    const long long i = Key::i(k), j = Key::j(k);
    leaves.erase(k);
    for (int b = 0; b < 2; b++)
      for (int a = 0; a < 2; a++) leaves.insert(Key::make(l + 1, 2 * i + a, 2 * j + b));
  }

  // Level of the leaf covering position (l,i,j) from above (the position
  // itself or an ancestor), or -1 when the position is refined further.
  int coveringLevel(int l, long long i, long long j) const {
    for (int m = l; m >= 0; m--) {
      const int d = l - m;
      if (leaves.count(Key::make(m, i >> d, j >> d))) return m;
    }
    return -1;
  }

  // 2:1 balance over faces and corners.
  void balance() {
    bool changed = true;
    while (changed) {
      changed = false;
      std::vector<uint64_t> cur(leaves.begin(), leaves.end());
      std::sort(cur.begin(), cur.end());
      for (uint64_t k : cur) {
        const int l = Key::level(k);
        if (l < 2 || !leaves.count(k)) continue;
        const long long i = Key::i(k), j = Key::j(k);
        for (int dj = -1; dj <= 1; dj++)
          for (int di = -1; di <= 1; di++) {
            if (!di && !dj) continue;
            long long ni = i + di, nj = j + dj;
            if (ni < 0 || ni >= nx(l)) {
              if (!perx) continue;
              ni = (ni + nx(l)) % nx(l);
            }
            if (nj < 0 || nj >= ny(l)) {
              if (!pery) continue;
              nj = (nj + ny(l)) % ny(l);
            }
            const int m = coveringLevel(l - 1, ni >> 1, nj >> 1);
            if (m >= 0 && m < l - 1) {
              const int d = l - 1 - m;
              split(Key::make(m, (ni >> 1) >> d, (nj >> 1) >> d));
              changed = true;
            }
          }
      }
    }
  }
};

// Hilbert index of (x,y) on a 2^order square (classic rotate-and-flip walk).
uint64_t hilbert(int order, uint64_t x, uint64_t y) {
  const uint64_t n = uint64_t(1) << order;
  uint64_t d = 0;
  for (uint64_t s = n >> 1; s > 0; s >>= 1) {
    const uint64_t rx = (x & s) > 0;
    const uint64_t ry = (y & s) > 0;
    d += s * s * ((3 * rx) ^ ry);
    if (ry == 0) {
      if (rx == 1) {
        x = n - 1 - x;  // bits above s are no longer read, so flipping them is harmless
        y = n - 1 - y;
      }
      std::swap(x, y);
    }
  }
  return d;
}

std::string g_err;

// Hilbert order of a leaf set at the finest of `levels` levels.
std::vector<uint64_t> hilbert_order(const Gen &G) {
  int order = 1;
  const int L = G.levels - 1;
  while ((1LL << order) < (long long)std::max(G.bpdx, G.bpdy) << L) order++;
  std::vector<std::pair<uint64_t, uint64_t>> ord;
  ord.reserve(G.leaves.size());
  for (uint64_t k : G.leaves) {
    const int d = L - Key::level(k);
    const uint64_t x = uint64_t(Key::i(k)) << d, y = uint64_t(Key::j(k)) << d;
    ord.push_back({hilbert(order, x, y), k});
  }
  std::sort(ord.begin(), ord.end());
  std::vector<uint64_t> out(ord.size());
  for (size_t b = 0; b < ord.size(); b++) out[b] = ord[b].second;
  return out;
}

}  // namespace

extern "C" {

const char *cup2d_mesh_last_error(void) { return g_err.c_str(); }

long long cup2d_mesh_graded(int bpdx, int bpdy, int levels, double extent, int periodic_x,
                            int periodic_y, int ndisks, const double *disks, const double *bands,
                            int *level, int *bx, int *by, long long cap) {
  if (bpdx <= 0 || bpdy <= 0 || levels <= 0 || levels > 16 || extent <= 0 || ndisks < 0 ||
      (ndisks > 0 && (!disks || !bands))) {
    g_err = "cup2d_mesh_graded: invalid arguments";
    return -1;
  }
  Gen G;
  G.bpdx = bpdx;
  G.bpdy = bpdy;
  G.levels = levels;
  G.perx = periodic_x != 0;
  G.pery = periodic_y != 0;
  G.s0 = extent / double(std::max(bpdx, bpdy));
  for (long long j = 0; j < bpdy; j++)
    for (long long i = 0; i < bpdx; i++) G.leaves.insert(Key::make(0, i, j));

  for (int l = 0; l + 1 < levels; l++) {
    std::vector<uint64_t> cur;
    for (uint64_t k : G.leaves)
      if (Key::level(k) == l) cur.push_back(k);
    std::sort(cur.begin(), cur.end());
    const double s = G.s0 / double(1LL << l);
    for (uint64_t k : cur) {
      const double x0 = Key::i(k) * s, y0 = Key::j(k) * s;
      bool ref = false;
      for (int d = 0; d < ndisks && !ref; d++) {
        const double cx = disks[3 * d], cy = disks[3 * d + 1], r = disks[3 * d + 2];
        const double dx = std::max({x0 - cx, 0.0, cx - (x0 + s)});
        const double dy = std::max({y0 - cy, 0.0, cy - (y0 + s)});
        const double dist = std::max(0.0, std::sqrt(dx * dx + dy * dy) - r);
        ref = dist < bands[l];
      }
      if (ref) G.split(k);
    }
  }
  G.balance();

  const long long nb = (long long)G.leaves.size();
  if (nb > cap || !level || !bx || !by) return nb;

  const std::vector<uint64_t> ord = hilbert_order(G);
  for (long long b = 0; b < nb; b++) {
    const uint64_t k = ord[b];
    level[b] = Key::level(k);
    bx[b] = (int)Key::i(k);
    by[b] = (int)Key::j(k);
  }
  return nb;
}

long long cup2d_mesh_adapt(int bpdx, int bpdy, int levels, int periodic_x, int periodic_y,
                           long long nb_old, const int *level_old, const int *bx_old,
                           const int *by_old, const signed char *tag, int *level, int *bx, int *by,
                           int *src_kind, long long *src, long long cap) {
  if (bpdx <= 0 || bpdy <= 0 || levels <= 0 || levels > 16 || nb_old <= 0 || !level_old || !bx_old ||
      !by_old || !tag) {
    g_err = "cup2d_mesh_adapt: invalid arguments";
    return -1;
  }
  Gen G;
  G.bpdx = bpdx;
  G.bpdy = bpdy;
  G.levels = levels;
  G.perx = periodic_x != 0;
  G.pery = periodic_y != 0;
  G.s0 = 1.0;
  std::unordered_map<uint64_t, long long> old_id;
  old_id.reserve(size_t(nb_old) * 2);
  for (long long b = 0; b < nb_old; b++) {
    const int l = level_old[b];
    if (l < 0 || l >= levels || bx_old[b] < 0 || bx_old[b] >= (bpdx << l) || by_old[b] < 0 ||
        by_old[b] >= (bpdy << l)) {
      g_err = "cup2d_mesh_adapt: block " + std::to_string(b) + " out of range";
      return -1;
    }
    const uint64_t k = Key::make(l, bx_old[b], by_old[b]);
    if (!old_id.emplace(k, b).second) {
      g_err = "cup2d_mesh_adapt: block " + std::to_string(b) + " duplicate";
      return -1;
    }
    G.leaves.insert(k);
  }
  // 1. refinement of the tagged blocks below the top level
  for (long long b = 0; b < nb_old; b++)
    if (tag[b] > 0 && level_old[b] + 1 < levels) G.split(Key::make(level_old[b], bx_old[b], by_old[b]));
  G.balance();
  // 2. compression: four sibling leaves, all old blocks tagged -1, whose
  // parent keeps every face and corner neighbour within one level
  std::vector<uint64_t> cand;
  for (long long b = 0; b < nb_old; b++) {
    const int l = level_old[b];
    if (tag[b] < 0 && l > 0 && (bx_old[b] & 1) == 0 && (by_old[b] & 1) == 0)
      cand.push_back(Key::make(l - 1, bx_old[b] >> 1, by_old[b] >> 1));
  }
  std::sort(cand.begin(), cand.end());
  for (uint64_t pk : cand) {
    const int l = Key::level(pk) + 1;
    const long long pi = Key::i(pk), pj = Key::j(pk);
    bool all = true;
    uint64_t kid[4];
    for (int q = 0; q < 4 && all; q++) {
      kid[q] = Key::make(l, 2 * pi + (q & 1), 2 * pj + (q >> 1));
      auto it = old_id.find(kid[q]);
      all = it != old_id.end() && tag[it->second] < 0 && G.leaves.count(kid[q]);
    }
    if (!all) continue;
    // the parent at level l - 1: every leaf touching it must be at level <= l
    bool ok = true;
    const int pl = l - 1;
    for (int dj = -1; dj <= 1 && ok; dj++)
      for (int di = -1; di <= 1 && ok; di++) {
        if (!di && !dj) continue;
        long long ni = pi + di, nj = pj + dj;
        const long long nx = (long long)bpdx << pl, ny = (long long)bpdy << pl;
        if (ni < 0 || ni >= nx) {
          if (!G.perx) continue;
          ni = (ni + nx) % nx;
        }
        if (nj < 0 || nj >= ny) {
          if (!G.pery) continue;
          nj = (nj + ny) % ny;
        }
        // the neighbour region at level pl is covered by a leaf at level <= pl
        // or subdivided: every leaf inside it must be at level <= pl + 1
        if (G.coveringLevel(pl, ni, nj) >= 0) continue;
        for (int q = 0; q < 4 && ok; q++) {
          const long long ci = 2 * ni + (q & 1), cj = 2 * nj + (q >> 1);
          if (!G.leaves.count(Key::make(pl + 1, ci, cj))) ok = false;  // deeper than pl + 1
        }
      }
    if (!ok) continue;
    for (int q = 0; q < 4; q++) G.leaves.erase(kid[q]);
    G.leaves.insert(pk);
  }
  G.balance();

  const long long nb = (long long)G.leaves.size();
  if (nb > cap || !level || !bx || !by || !src_kind || !src) return nb;
  const std::vector<uint64_t> ord = hilbert_order(G);
  for (long long b = 0; b < nb; b++) {
    const uint64_t k = ord[b];
    const int l = Key::level(k);
    const long long i = Key::i(k), j = Key::j(k);
    level[b] = l;
    bx[b] = (int)i;
    by[b] = (int)j;
    src[4 * b] = src[4 * b + 1] = src[4 * b + 2] = src[4 * b + 3] = -1;
    auto it = old_id.find(k);
    if (it != old_id.end()) {  // unchanged
      src_kind[b] = 0;
      src[4 * b] = it->second;
      continue;
    }
    bool found = false;
    for (int m = l - 1; m >= 0 && !found; m--) {  // an old ancestor: refinement
      const int d = l - m;
      auto a = old_id.find(Key::make(m, i >> d, j >> d));
      if (a != old_id.end()) {
        src_kind[b] = 1;
        src[4 * b] = a->second;
        found = true;
      }
    }
    if (found) continue;
    src_kind[b] = 2;  // four old children: compression
    for (int q = 0; q < 4; q++) {
      auto c = old_id.find(Key::make(l + 1, 2 * i + (q & 1), 2 * j + (q >> 1)));
      if (c == old_id.end()) {
        g_err = "cup2d_mesh_adapt: internal error (a new block has no source)";
        return -1;
      }
      src[4 * b + q] = c->second;
    }
  }
  return nb;
}

}  // extern "C"
