// assemble.cpp — host-side scene assembly, the step before the solve: the edge table of a mesh
// (boundary edges, manifold and fold checks), and each body's self-contact exclusion, by
// mesh-edge hops and by distance through the material (reference core/mesh.py:112-186,
// 190-286). Plain C++ on the host; the results are the same sets the reference computes.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <queue>
#include <utility>
#include <vector>

#include "../../include/d2fem.h"

namespace {

// Rows of an (m, W) int64 array, each sorted ascending, ordered lexicographically and then by
// row index (so equal rows keep their original order, like a stable sort).
template <int W>
std::vector<int64_t> sorted_rows(const int64_t *R, int64_t m, std::vector<int64_t> &keys) {
  keys.resize((size_t)m * W);
  for (int64_t i = 0; i < m; ++i) {
    int64_t *k = keys.data() + (size_t)i * W;
    for (int c = 0; c < W; ++c) k[c] = R[(size_t)i * W + c];
    std::sort(k, k + W);
  }
  std::vector<int64_t> idx((size_t)m);
  for (int64_t i = 0; i < m; ++i) idx[(size_t)i] = i;
  const int64_t *K = keys.data();
  std::sort(idx.begin(), idx.end(), [K](int64_t a, int64_t b) {
    for (int c = 0; c < W; ++c) {
      const int64_t x = K[(size_t)a * W + c], y = K[(size_t)b * W + c];
      if (x != y) return x < y;
    }
    return a < b;
  });
  return idx;
}

// Bits per entry when a sorted row packs into one 64-bit key (entries in [0, 2^bits)), else 0.
template <int W>
int packed_bits(const int64_t *R, int64_t m) {
  const int bits = 64 / W;
  const int64_t lim = int64_t(1) << (bits - 1);  // one bit to spare: keys stay positive
  for (int64_t i = 0; i < (int64_t)W * m; ++i)
    if (R[i] < 0 || R[i] >= lim) return 0;
  return bits;
}

template <int W>
int64_t unique_rows(const int64_t *R, int64_t m, int64_t *key, int64_t *first, int64_t *count, int64_t *inverse) {
  if (m < 0) return -1;
  if (const int bits = packed_bits<W>(R, m)) {
    // each sorted row as one integer; sorting (key, row) pairs keeps equal rows in row order
    std::vector<std::pair<uint64_t, int64_t>> kr((size_t)m);
    for (int64_t i = 0; i < m; ++i) {
      int64_t k[W];
      for (int c = 0; c < W; ++c) k[c] = R[(size_t)i * W + c];
      std::sort(k, k + W);
      uint64_t v = 0;
      for (int c = 0; c < W; ++c) v = (v << bits) | uint64_t(k[c]);
      kr[(size_t)i] = {v, i};
    }
    std::sort(kr.begin(), kr.end());
    const uint64_t mask = (uint64_t(1) << bits) - 1;
    int64_t u = -1;
    for (int64_t j = 0; j < m; ++j) {
      if (j == 0 || kr[(size_t)j].first != kr[(size_t)j - 1].first) {
        ++u;
        for (int c = 0; c < W; ++c) key[(size_t)u * W + c] = int64_t((kr[(size_t)j].first >> (bits * (W - 1 - c))) & mask);
        first[u] = kr[(size_t)j].second;
        count[u] = 0;
      }
      ++count[u];
      if (inverse) inverse[kr[(size_t)j].second] = u;
    }
    return u + 1;
  }
  std::vector<int64_t> keys;
  const std::vector<int64_t> idx = sorted_rows<W>(R, m, keys);
  int64_t u = -1;
  for (int64_t j = 0; j < m; ++j) {
    const int64_t r = idx[(size_t)j];
    const int64_t *k = keys.data() + (size_t)r * W;
    bool fresh = j == 0;
    if (!fresh) {
      const int64_t *kp = keys.data() + (size_t)idx[(size_t)j - 1] * W;
      for (int c = 0; c < W; ++c) fresh = fresh || k[c] != kp[c];
    }
    if (fresh) {
      ++u;
      for (int c = 0; c < W; ++c) key[(size_t)u * W + c] = k[c];
      first[u] = r;
      count[u] = 0;
    }
    ++count[u];
    if (inverse) inverse[r] = u;
  }
  return u + 1;
}

// Neighbour lists (CSR, each list sorted, no repeats) of the nodes of a triangle mesh: counted
// into place, then each short list sorted and compacted.
void adjacency(const int64_t *tri, int64_t T, int64_t n, std::vector<int64_t> &off, std::vector<int64_t> &nbr) {
  std::vector<int64_t> cnt((size_t)n + 1, 0);
  for (int64_t i = 0; i < 3 * T; ++i) cnt[(size_t)tri[i] + 1] += 2;
  for (int64_t i = 0; i < n; ++i) cnt[(size_t)i + 1] += cnt[(size_t)i];
  std::vector<int64_t> raw((size_t)cnt[(size_t)n]), fill(cnt.begin(), cnt.end() - 1);
  for (int64_t t = 0; t < T; ++t) {
    const int64_t v[3] = {tri[3 * t], tri[3 * t + 1], tri[3 * t + 2]};
    for (int a = 0; a < 3; ++a) {
      raw[(size_t)fill[(size_t)v[a]]++] = v[(a + 1) % 3];
      raw[(size_t)fill[(size_t)v[a]]++] = v[(a + 2) % 3];
    }
  }
  off.assign((size_t)n + 1, 0);
  nbr.clear();
  nbr.reserve(raw.size() / 2 + 1);
  for (int64_t i = 0; i < n; ++i) {
    auto b = raw.begin() + cnt[(size_t)i], e = raw.begin() + cnt[(size_t)i + 1];
    std::sort(b, e);
    for (auto it = b; it != e; ++it)
      if (it == b || *it != *(it - 1)) nbr.push_back(*it);
    off[(size_t)i + 1] = (int64_t)nbr.size();
  }
}

bool valid_nodes(const int64_t *v, int64_t count, int64_t n) {
  for (int64_t i = 0; i < count; ++i)
    if (v[i] < 0 || v[i] >= n) return false;
  return true;
}

}  // namespace

extern "C" {

int64_t d2f_unique_pairs(const int64_t *E, int64_t m, int64_t *key, int64_t *first, int64_t *count,
                         int64_t *inverse) {
  if (m < 0 || (m > 0 && (!E || !key || !first || !count))) return -1;
  return unique_rows<2>(E, m, key, first, count, inverse);
}

int64_t d2f_unique_triples(const int64_t *R, int64_t m, int64_t *key, int64_t *first, int64_t *count) {
  if (m < 0 || (m > 0 && (!R || !key || !first || !count))) return -1;
  return unique_rows<3>(R, m, key, first, count, nullptr);
}

// Self-contact exclusion of one body as sorted lists (CSR over its boundary nodes): boundary
// node s excludes every boundary node within k mesh-edge hops (reference core/mesh.py:133-153)
// and, when radius > 0, every one closer than radius along mesh edges at rest (reference
// core/mesh.py:163-186), and the relation is made symmetric (reference core/mesh.py:157-158).
// The same sets as d2f_hop_exclusion | d2f_material_exclusion, without the nb x nb matrix.
// row_ptr[nb + 1] and cols (capacity cap) receive the lists; returns the number of entries
// (when it exceeds cap, nothing is written beyond cap and the caller retries), or < 0.
int64_t d2f_exclusion_csr(const int64_t *tri, int64_t T, int64_t n_nodes, const double *pos,
                          const int64_t *boundary, int64_t nb, int32_t k, double radius, int64_t *row_ptr,
                          int32_t *cols, int64_t cap) {
  if (T < 0 || n_nodes < 0 || nb < 0 || k < 0 || !(radius >= 0.0) || !row_ptr || cap < 0 ||
      (nb > 0 && !boundary) || (T > 0 && !tri) || (radius > 0.0 && !pos) || (cap > 0 && !cols))
    return D2F_EINVAL;
  if (!valid_nodes(tri, 3 * T, n_nodes) || !valid_nodes(boundary, nb, n_nodes)) return D2F_EINVAL;
  std::vector<int64_t> off, nbr;
  adjacency(tri, T, n_nodes, off, nbr);
  std::vector<int64_t> bidx((size_t)n_nodes, -1);
  for (int64_t i = 0; i < nb; ++i) bidx[(size_t)boundary[i]] = i;
  std::vector<std::vector<int32_t>> rows((size_t)nb);
  std::vector<int32_t> hop((size_t)n_nodes, -1);
  std::vector<int64_t> seen, q;
  for (int64_t s = 0; s < nb; ++s) {          // breadth-first search up to k edges
    const int64_t src = boundary[s];
    q.assign(1, src);
    seen.assign(1, src);
    hop[(size_t)src] = 0;
    for (size_t h = 0; h < q.size(); ++h) {
      const int64_t u = q[h];
      const int32_t d = hop[(size_t)u];
      if (bidx[(size_t)u] >= 0) rows[(size_t)s].push_back((int32_t)bidx[(size_t)u]);
      if (d == k) continue;
      for (int64_t j = off[(size_t)u]; j < off[(size_t)u + 1]; ++j) {
        const int64_t w = nbr[(size_t)j];
        if (hop[(size_t)w] < 0) {
          hop[(size_t)w] = d + 1;
          q.push_back(w);
          seen.push_back(w);
        }
      }
    }
    for (int64_t w : seen) hop[(size_t)w] = -1;
  }
  if (radius > 0.0) {                         // Dijkstra over rest edge lengths, stopped at radius
    const double inf = INFINITY;
    std::vector<double> dist((size_t)n_nodes, inf);
    std::vector<uint8_t> done((size_t)n_nodes, 0);
    std::vector<int64_t> touched;
    using Item = std::pair<double, int64_t>;
    for (int64_t s = 0; s < nb; ++s) {
      std::priority_queue<Item, std::vector<Item>, std::greater<Item>> pq;
      const int64_t src = boundary[s];
      dist[(size_t)src] = 0.0;
      touched.assign(1, src);
      pq.push({0.0, src});
      while (!pq.empty()) {
        const Item it = pq.top();
        pq.pop();
        const int64_t u = it.second;
        if (done[(size_t)u] || it.first > dist[(size_t)u]) continue;
        if (!(it.first < radius)) break;
        done[(size_t)u] = 1;
        if (bidx[(size_t)u] >= 0) rows[(size_t)s].push_back((int32_t)bidx[(size_t)u]);
        for (int64_t j = off[(size_t)u]; j < off[(size_t)u + 1]; ++j) {
          const int64_t v = nbr[(size_t)j];
          if (done[(size_t)v]) continue;
          const double dx = pos[2 * u] - pos[2 * v], dy = pos[2 * u + 1] - pos[2 * v + 1];
          const double nd = dist[(size_t)u] + std::sqrt(dx * dx + dy * dy);
          if (nd < dist[(size_t)v]) {
            if (dist[(size_t)v] == inf) touched.push_back(v);
            dist[(size_t)v] = nd;
            pq.push({nd, v});
          }
        }
      }
      for (int64_t v : touched) {
        dist[(size_t)v] = inf;
        done[(size_t)v] = 0;
      }
    }
  }
  // symmetric closure, then each list sorted without repeats
  for (int64_t s = 0; s < nb; ++s) {
    const size_t n0 = rows[(size_t)s].size();
    for (size_t i = 0; i < n0; ++i) {
      const int32_t t = rows[(size_t)s][i];
      if (t != s) rows[(size_t)t].push_back((int32_t)s);
    }
  }
  int64_t nnz = 0;
  row_ptr[0] = 0;
  for (int64_t s = 0; s < nb; ++s) {
    auto &r = rows[(size_t)s];
    std::sort(r.begin(), r.end());
    r.erase(std::unique(r.begin(), r.end()), r.end());
    for (int32_t t : r) {
      if (nnz < cap) cols[nnz] = t;
      ++nnz;
    }
    row_ptr[s + 1] = nnz;
  }
  return nnz;
}

}  // extern "C"
