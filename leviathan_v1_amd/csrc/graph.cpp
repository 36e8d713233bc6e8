// graph.cpp -- seeded topology builders for the LEVIATHAN step.
//
// The reference builds its graph with networkx on the host:
//   python/leviathan_h100.py:34   nx.watts_strogatz_graph(N, k, p=0.2) (unseeded)
//   python/leviathan_h100.py:40   nx.to_scipy_sparse_array(G, format="csr")
// networkx 3.4.2 is a third-party dependency of that path.  The two generators
// below restate its published algorithms (networkx/generators/random_graphs.py:
// watts_strogatz_graph, barabasi_albert_graph with the default star seed graph,
// _random_subset) with a seeded xoshiro256** stream instead of Python's Mersenne
// twister, so a graph is reproducible from (n, k|m, p, seed) and a 10^7-node
// Barabasi-Albert graph builds in seconds instead of the hour networkx needs.
//
// Output is the CSR of incoming edges that the engine consumes (MODEL.md:60-69):
// an undirected edge {u,v} is stored as u->v and v->u, and the columns of every
// row are sorted ascending (the scipy CSR of the adjacency matrix).
//
// Host-only C++ (g++ -O3 -fopenmp); exported through a plain C ABI.

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include <omp.h>

namespace {

thread_local std::string g_err;

// splitmix64 seeds the state of xoshiro256** (Blackman & Vigna, 2018).
struct Rng {
  uint64_t s[4];
  explicit Rng(uint64_t seed) {
    uint64_t x = seed;
    for (int i = 0; i < 4; ++i) {
      x += 0x9e3779b97f4a7c15ULL;
      uint64_t z = x;
      z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
      z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
      s[i] = z ^ (z >> 31);
    }
  }
  static inline uint64_t rotl(uint64_t v, int k) { return (v << k) | (v >> (64 - k)); }
  inline uint64_t next() {
    const uint64_t result = rotl(s[1] * 5, 7) * 9;
    const uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return result;
  }
  // uniform double in [0, 1) with 53 random bits
  inline double uniform() { return (double)(next() >> 11) * 0x1.0p-53; }
  // unbiased integer in [0, n) (Lemire's multiply-and-reject method)
  inline uint64_t below(uint64_t n) {
    uint64_t x = next();
    __uint128_t m = (__uint128_t)x * n;
    uint64_t l = (uint64_t)m;
    if (l < n) {
      uint64_t t = (0 - n) % n;
      while (l < t) {
        x = next();
        m = (__uint128_t)x * n;
        l = (uint64_t)m;
      }
    }
    return (uint64_t)(m >> 64);
  }
};

}  // namespace

struct lev_graph {
  int32_t n = 0;
  std::vector<int32_t> row_ptr;  // n + 1
  std::vector<int32_t> col_idx;  // row_ptr[n]
};

namespace {

void sort_rows(lev_graph* g) {
  const int32_t n = g->n;
#pragma omp parallel for schedule(dynamic, 4096)
  for (int32_t i = 0; i < n; ++i) {
    std::sort(g->col_idx.begin() + g->row_ptr[i], g->col_idx.begin() + g->row_ptr[i + 1]);
  }
}

bool check_edge_count(int64_t e) {
  if (e > (int64_t)INT32_MAX) {
    g_err = "graph has " + std::to_string(e) + " CSR entries; the int32 CSR holds at most 2^31-1";
    return false;
  }
  return true;
}

}  // namespace

extern "C" {

const char* lev_graph_last_error(void) { return g_err.c_str(); }

// Watts-Strogatz small world (networkx watts_strogatz_graph): ring lattice with
// k/2 neighbours per side, then every lattice edge (u, u+j), visited with j as
// the outer loop and u as the inner loop, is rewired with probability p to
// (u, w) with w uniform, rejecting self loops and existing edges.
lev_graph* lev_graph_watts_strogatz(int32_t n, int32_t k, double p, uint64_t seed) {
  g_err.clear();
  if (n <= 0 || k < 0 || k > n || p < 0.0 || p > 1.0) {
    g_err = "watts_strogatz: need n > 0, 0 <= k <= n, 0 <= p <= 1";
    return nullptr;
  }
  lev_graph* g = new (std::nothrow) lev_graph();
  if (!g) {
    g_err = "out of host memory";
    return nullptr;
  }
  g->n = n;
  std::vector<std::vector<int32_t>> adj((size_t)n);
  if (k == n) {  // networkx returns the complete graph here
    for (int32_t u = 0; u < n; ++u)
      for (int32_t v = 0; v < n; ++v)
        if (u != v) adj[u].push_back(v);
  } else {
    const int32_t half = k / 2;
    for (int32_t u = 0; u < n; ++u) adj[u].reserve((size_t)k + 4);
    for (int32_t j = 1; j <= half; ++j)
      for (int32_t u = 0; u < n; ++u) {
        int32_t v = (int32_t)(((int64_t)u + j) % n);
        if (u == v) continue;
        // networkx's Graph dedups a lattice edge that wraps onto an existing one
        if (std::find(adj[u].begin(), adj[u].end(), v) != adj[u].end()) continue;
        adj[u].push_back(v);
        adj[v].push_back(u);
      }
    Rng rng(seed);
    auto has_edge = [&](int32_t a, int32_t b) {
      return std::find(adj[a].begin(), adj[a].end(), b) != adj[a].end();
    };
    auto remove_one = [&](int32_t a, int32_t b) {
      auto it = std::find(adj[a].begin(), adj[a].end(), b);
      if (it != adj[a].end()) {
        *it = adj[a].back();
        adj[a].pop_back();
      }
    };
    for (int32_t j = 1; j <= half; ++j) {
      for (int32_t u = 0; u < n; ++u) {
        const int32_t v = (int32_t)(((int64_t)u + j) % n);
        if (rng.uniform() < p) {
          int32_t w = (int32_t)rng.below((uint64_t)n);
          bool skip = false;
          while (w == u || has_edge(u, w)) {
            w = (int32_t)rng.below((uint64_t)n);
            if ((int64_t)adj[u].size() >= (int64_t)n - 1) {
              skip = true;
              break;
            }
          }
          if (!skip && has_edge(u, v)) {
            remove_one(u, v);
            remove_one(v, u);
            adj[u].push_back(w);
            adj[w].push_back(u);
          } else if (!skip) {
            // (u, v) was already rewired away by an earlier pass: networkx's
            // remove_edge would raise; adding (u, w) alone keeps the edge count.
            adj[u].push_back(w);
            adj[w].push_back(u);
          }
        }
      }
    }
  }
  int64_t e = 0;
  for (int32_t u = 0; u < n; ++u) e += (int64_t)adj[u].size();
  if (!check_edge_count(e)) {
    delete g;
    return nullptr;
  }
  g->row_ptr.assign((size_t)n + 1, 0);
  for (int32_t u = 0; u < n; ++u) g->row_ptr[u + 1] = g->row_ptr[u] + (int32_t)adj[u].size();
  g->col_idx.resize((size_t)e);
  for (int32_t u = 0; u < n; ++u) {
    std::copy(adj[u].begin(), adj[u].end(), g->col_idx.begin() + g->row_ptr[u]);
    std::vector<int32_t>().swap(adj[u]);
  }
  sort_rows(g);
  return g;
}

// Barabasi-Albert preferential attachment (networkx barabasi_albert_graph with
// initial_graph=None): start from the star on m+1 nodes (centre 0); every new
// node s = m+1 .. n-1 draws m distinct targets uniformly from the list in which
// each node appears once per incident edge, then appends the targets and m
// copies of s to that list.
lev_graph* lev_graph_barabasi_albert(int32_t n, int32_t m, uint64_t seed) {
  g_err.clear();
  if (m < 1 || m >= n) {
    g_err = "barabasi_albert: need 1 <= m < n";
    return nullptr;
  }
  const int64_t undirected = (int64_t)m + (int64_t)m * ((int64_t)n - m - 1);
  if (!check_edge_count(2 * undirected)) return nullptr;
  lev_graph* g = new (std::nothrow) lev_graph();
  if (!g) {
    g_err = "out of host memory";
    return nullptr;
  }
  g->n = n;
  std::vector<int32_t> rep;
  try {
    rep.reserve((size_t)(2 * undirected));
  } catch (...) {
    delete g;
    g_err = "out of host memory";
    return nullptr;
  }
  // star_graph(m): nodes 0..m, centre 0; degree order lists node 0 m times first
  for (int32_t q = 0; q < m; ++q) rep.push_back(0);
  for (int32_t q = 1; q <= m; ++q) rep.push_back(q);
  Rng rng(seed);
  std::vector<int32_t> tgt((size_t)m);
  for (int64_t s = m + 1; s < n; ++s) {
    const uint64_t len = rep.size();
    int32_t got = 0;
    while (got < m) {
      int32_t x = rep[rng.below(len)];
      bool dup = false;
      for (int32_t q = 0; q < got; ++q)
        if (tgt[q] == x) {
          dup = true;
          break;
        }
      if (!dup) tgt[got++] = x;
    }
    for (int32_t q = 0; q < m; ++q) rep.push_back(tgt[q]);
    for (int32_t q = 0; q < m; ++q) rep.push_back((int32_t)s);
  }
  // Edges: star (0, q) for q = 1..m, then (s, rep[2m + 2m(s-m-1) + q]).
  std::vector<int32_t> deg((size_t)n, 0);
  deg[0] += m;
  for (int32_t q = 1; q <= m; ++q) deg[q] += 1;
  const size_t base = 2 * (size_t)m;
  for (int64_t s = m + 1; s < n; ++s) {
    const size_t off = base + 2 * (size_t)m * (size_t)(s - m - 1);
    deg[s] += m;
    for (int32_t q = 0; q < m; ++q) deg[rep[off + q]] += 1;
  }
  g->row_ptr.assign((size_t)n + 1, 0);
  for (int32_t u = 0; u < n; ++u) g->row_ptr[u + 1] = g->row_ptr[u] + deg[u];
  g->col_idx.resize((size_t)g->row_ptr[n]);
  std::vector<int32_t> cur(g->row_ptr.begin(), g->row_ptr.end() - 1);
  for (int32_t q = 1; q <= m; ++q) {
    g->col_idx[cur[0]++] = q;
    g->col_idx[cur[q]++] = 0;
  }
  for (int64_t s = m + 1; s < n; ++s) {
    const size_t off = base + 2 * (size_t)m * (size_t)(s - m - 1);
    for (int32_t q = 0; q < m; ++q) {
      const int32_t t = rep[off + q];
      g->col_idx[cur[s]++] = t;
      g->col_idx[cur[t]++] = (int32_t)s;
    }
  }
  std::vector<int32_t>().swap(rep);
  sort_rows(g);
  return g;
}

// Build from an explicit undirected edge list (both directions stored).
lev_graph* lev_graph_from_edges(int32_t n, int64_t n_edges, const int32_t* src, const int32_t* dst) {
  g_err.clear();
  if (n <= 0 || n_edges < 0) {
    g_err = "from_edges: need n > 0 and n_edges >= 0";
    return nullptr;
  }
  if (!check_edge_count(2 * n_edges)) return nullptr;
  lev_graph* g = new (std::nothrow) lev_graph();
  if (!g) {
    g_err = "out of host memory";
    return nullptr;
  }
  g->n = n;
  std::vector<int32_t> deg((size_t)n, 0);
  for (int64_t e = 0; e < n_edges; ++e) {
    if (src[e] < 0 || src[e] >= n || dst[e] < 0 || dst[e] >= n) {
      g_err = "from_edges: endpoint out of range at edge " + std::to_string(e);
      delete g;
      return nullptr;
    }
    deg[src[e]]++;
    deg[dst[e]]++;
  }
  g->row_ptr.assign((size_t)n + 1, 0);
  for (int32_t u = 0; u < n; ++u) g->row_ptr[u + 1] = g->row_ptr[u] + deg[u];
  g->col_idx.resize((size_t)g->row_ptr[n]);
  std::vector<int32_t> cur(g->row_ptr.begin(), g->row_ptr.end() - 1);
  for (int64_t e = 0; e < n_edges; ++e) {
    g->col_idx[cur[dst[e]]++] = src[e];
    g->col_idx[cur[src[e]]++] = dst[e];
  }
  sort_rows(g);
  return g;
}

int32_t lev_graph_num_nodes(const lev_graph* g) { return g ? g->n : 0; }

int64_t lev_graph_num_entries(const lev_graph* g) { return g ? (int64_t)g->col_idx.size() : 0; }

int32_t lev_graph_max_degree(const lev_graph* g) {
  if (!g) return 0;
  int32_t best = 0;
  for (int32_t u = 0; u < g->n; ++u) best = std::max(best, g->row_ptr[u + 1] - g->row_ptr[u]);
  return best;
}

void lev_graph_copy_csr(const lev_graph* g, int32_t* row_ptr, int32_t* col_idx) {
  if (!g) return;
  std::memcpy(row_ptr, g->row_ptr.data(), sizeof(int32_t) * ((size_t)g->n + 1));
  if (!g->col_idx.empty()) std::memcpy(col_idx, g->col_idx.data(), sizeof(int32_t) * g->col_idx.size());
}

void lev_graph_free(lev_graph* g) { delete g; }

// Relabel the nodes of an incoming-edge CSR: new node i is old node perm[i].
// Writes the new CSR (row_ptr[n+1], col_idx[E], columns ascending within a
// row) and emap[E]: the old CSR index of each new entry, to carry per-edge
// arrays (delays, weights) along.  Returns 0, or -1 if perm is not a
// permutation of [0, n).
int lev_graph_permute(int32_t n, const int32_t* row_ptr, const int32_t* col_idx, const int32_t* perm,
                      int32_t* out_row_ptr, int32_t* out_col, int64_t* emap) {
  std::vector<int32_t> inv((size_t)n, -1);
  for (int32_t i = 0; i < n; ++i) {
    const int32_t p = perm[i];
    if (p < 0 || p >= n || inv[p] >= 0) {
      g_err = "not a permutation";
      return -1;
    }
    inv[p] = i;
  }
  out_row_ptr[0] = 0;
  for (int32_t i = 0; i < n; ++i) out_row_ptr[i + 1] = out_row_ptr[i] + (row_ptr[perm[i] + 1] - row_ptr[perm[i]]);
#pragma omp parallel for schedule(dynamic, 4096)
  for (int32_t i = 0; i < n; ++i) {
    const int32_t o0 = row_ptr[perm[i]], deg = row_ptr[perm[i] + 1] - o0, n0 = out_row_ptr[i];
    std::vector<std::pair<int32_t, int64_t>> row((size_t)deg);
    for (int32_t k = 0; k < deg; ++k) row[k] = {inv[col_idx[o0 + k]], (int64_t)o0 + k};
    std::sort(row.begin(), row.end());
    for (int32_t k = 0; k < deg; ++k) {
      out_col[n0 + k] = row[k].first;
      emap[n0 + k] = row[k].second;
    }
  }
  return 0;
}

}  // extern "C"
