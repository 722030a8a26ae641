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

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
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

  int order = 1;
  const int L = levels - 1;
  while ((1LL << order) < (long long)std::max(bpdx, bpdy) << L) order++;
  std::vector<std::pair<uint64_t, uint64_t>> ord;
  ord.reserve(nb);
  for (uint64_t k : G.leaves) {
    const int d = L - Key::level(k);
    const uint64_t x = uint64_t(Key::i(k)) << d, y = uint64_t(Key::j(k)) << d;
    ord.push_back({hilbert(order, x, y), k});
  }
  std::sort(ord.begin(), ord.end());
  for (long long b = 0; b < nb; b++) {
    const uint64_t k = ord[b].second;
    level[b] = Key::level(k);
    bx[b] = (int)Key::i(k);
    by[b] = (int)Key::j(k);
  }
  return nb;
}

}  // extern "C"
