// plan.cpp -- host-side construction of the step's work decomposition.
#include "plan.h"

#include <algorithm>
#include <atomic>
#include <thread>

#include "leviathan_internal.h"

namespace lev {

namespace {

inline int deg_of(const int* row_ptr, int r) { return row_ptr[r + 1] - row_ptr[r]; }

template <class F>
void parallel_for(int64_t n, F&& f) {
  unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  if ((int64_t)nt > n) nt = (unsigned)std::max<int64_t>(n, 1);
  std::vector<std::thread> th;
  const int64_t chunk = (n + nt - 1) / nt;
  for (unsigned t = 0; t < nt; ++t) {
    const int64_t lo = t * chunk, hi = std::min<int64_t>(n, lo + chunk);
    if (lo >= hi) break;
    th.emplace_back([&f, lo, hi]() {
      for (int64_t i = lo; i < hi; ++i) f(i);
    });
  }
  for (auto& x : th) x.join();
}

}  // namespace

bool build_direct_plan(const int* row_ptr, int n_nets, const int* offs, int cap, HostPlan& out, std::string& err) {
  if (n_nets >= (1 << 21)) {
    err = "too many networks (at most 2^21 - 1)";
    return false;
  }
  std::vector<std::vector<int4>> longs(n_nets), shorts(n_nets);
  for (int net = 0; net < n_nets; ++net) {
    const int lo = offs[net], hi = offs[net + 1];
    int r = lo;
    while (r < hi) {
      if (deg_of(row_ptr, r) > cap) {
        longs[net].push_back(pack_item(r, row_ptr[r], row_ptr[r + 1], 1, net, true));
        ++r;
        continue;
      }
      const int start = r;
      int nnz = 0;
      while (r < hi && r - start < kBlock) {
        const int d = deg_of(row_ptr, r);
        if (d > cap || nnz + d > cap) break;
        nnz += d;
        ++r;
      }
      shorts[net].push_back(pack_item(start, row_ptr[start], row_ptr[r], r - start, net, false));
    }
    std::stable_sort(longs[net].begin(), longs[net].end(),
                     [](const int4& a, const int4& b) { return a.z - a.y > b.z - b.y; });
  }
  out.items.clear();
  out.ranges.assign(n_nets, make_int4(0, 0, 0, 0));
  for (int net = 0; net < n_nets; ++net) {
    out.ranges[net].x = (int)out.items.size();
    out.items.insert(out.items.end(), longs[net].begin(), longs[net].end());
    out.ranges[net].y = (int)out.items.size();
  }
  for (int net = 0; net < n_nets; ++net) {
    out.ranges[net].z = (int)out.items.size();
    out.items.insert(out.items.end(), shorts[net].begin(), shorts[net].end());
    out.ranges[net].w = (int)out.items.size();
  }
  if (out.items.size() > (size_t)INT32_MAX) {
    err = "too many row blocks";
    return false;
  }
  // XCD runs of equal estimated cost.  Cost of a block in ns of workgroup
  // lifetime, fitted to per-block timestamps of the row pass on the
  // Barabasi-Albert 10^7 graph: rows above kSerialRow edges (summed by their
  // wavefront) and long rows (strided by the whole workgroup) cost extra.
  // (Fitted with a 32-edge kSerialRow.)
  std::vector<double> cum(out.items.size() + 1, 0.0);
  for (size_t q = 0; q < out.items.size(); ++q) {
    const int4 it = out.items[q];
    const int32_t nrows = (it.w >> 1) & 511, nnz = it.z - it.y;
    double c;
    if (it.w & kBlockLongRow) {
      c = kCostLong + kCostLongEdge * nnz;
    } else {
      int heavy = 0;
      for (int32_t r = it.x; r < it.x + nrows; ++r) heavy += deg_of(row_ptr, r) > kSerialRow;
      c = kCostBlock + kCostEdge * nnz + kCostRow * nrows + kCostHeavyRow * heavy;
    }
    cum[q + 1] = cum[q] + c;
  }
  out.xcd_split.assign(kNumXcd + 1, 0);
  const double total = cum.back();
  size_t q = 0;
  for (int x = 1; x < kNumXcd; ++x) {
    while (q < out.items.size() && cum[q + 1] <= total * x / kNumXcd) ++q;
    out.xcd_split[x] = (int32_t)q;
  }
  out.xcd_split[kNumXcd] = (int32_t)out.items.size();
  return true;
}

bool build_scatter_plan(int N, const int* row_ptr, const int* col, const uint8_t* delays, HostScatter& out,
                        std::string& err) {
  const int64_t E = row_ptr[N];
  if (E > (int64_t)UINT32_MAX) {
    err = "too many edges for 32-bit positions";
    return false;
  }
  // CSC (by source), destinations ascending within a source
  out.csc_ptr.assign((size_t)N + 1, 0);
  for (int64_t e = 0; e < E; ++e) out.csc_ptr[col[e] + 1]++;
  for (int j = 0; j < N; ++j) out.csc_ptr[j + 1] += out.csc_ptr[j];
  std::vector<int32_t> cur(out.csc_ptr.begin(), out.csc_ptr.end() - 1);
  out.csc_edge.assign((size_t)E, 0);
  out.delay_csc.assign((size_t)E, 0);
  for (int64_t e = 0; e < E; ++e) {
    const int32_t c = cur[col[e]]++;
    out.csc_edge[c] = (uint32_t)e;
    out.delay_csc[c] = delays[e];
  }
  // phase-1 work items over sources
  const int cap_src = kBlock * kScatterEpt;
  out.sources.clear();
  std::vector<int4> longs, shorts;
  int s = 0;
  while (s < N) {
    const int d = out.csc_ptr[s + 1] - out.csc_ptr[s];
    if (d > cap_src) {
      longs.push_back(pack_item(s, out.csc_ptr[s], out.csc_ptr[s + 1], 1, 0, true));
      ++s;
      continue;
    }
    const int start = s;
    int nnz = 0;
    while (s < N && s - start < kBlock) {
      const int dd = out.csc_ptr[s + 1] - out.csc_ptr[s];
      if (dd > cap_src || nnz + dd > cap_src) break;
      nnz += dd;
      ++s;
    }
    shorts.push_back(pack_item(start, out.csc_ptr[start], out.csc_ptr[s], s - start, 0, false));
  }
  out.sources.insert(out.sources.end(), longs.begin(), longs.end());
  out.sources.insert(out.sources.end(), shorts.begin(), shorts.end());
  return true;
}

namespace {

// hpos[e0 + l] for l in [0, n): e0 + the rank of edge e0 + l by (source, CSR index)
void rank_by_source(const int* col, int32_t e0, int32_t n, uint32_t* hpos) {
  std::vector<uint64_t> key((size_t)n);
  for (int32_t l = 0; l < n; ++l) key[l] = ((uint64_t)(uint32_t)col[e0 + l] << 32) | (uint32_t)l;
  std::sort(key.begin(), key.end());
  for (int32_t k = 0; k < n; ++k) hpos[e0 + (uint32_t)(key[k] & 0xffffffffu)] = (uint32_t)(e0 + k);
}

}  // namespace

bool bucket_positions(int N, const int* row_ptr, const int* col, int bucket_edges, std::vector<uint32_t>& hpos,
                      int32_t& n_buckets, std::string& err) {
  const int64_t E = row_ptr[N];
  if (bucket_edges < 1024) {
    err = "bucket_edges must be at least 1024";
    return false;
  }
  std::vector<int32_t> bstart{0};
  int r = 0;
  while (r < N) {
    int64_t nnz = deg_of(row_ptr, r);
    ++r;
    while (r < N && nnz + deg_of(row_ptr, r) <= bucket_edges) {
      nnz += deg_of(row_ptr, r);
      ++r;
    }
    bstart.push_back(r);
  }
  n_buckets = (int32_t)bstart.size() - 1;
  hpos.assign((size_t)E, 0);
  parallel_for(n_buckets, [&](int64_t q) {
    const int32_t e0 = row_ptr[bstart[q]];
    rank_by_source(col, e0, row_ptr[bstart[q + 1]] - e0, hpos.data());
  });
  return true;
}

bool staged_positions(int N, const int* row_ptr, const int* col, const HostPlan& plan, std::vector<uint32_t>& hpos,
                      std::vector<uint16_t>& lpos, std::string& err) {
  const int64_t E = row_ptr[N];
  hpos.assign((size_t)E, 0);
  lpos.assign((size_t)E, 0);
  std::atomic<bool> ok{true};
  parallel_for((int64_t)plan.items.size(), [&](int64_t q) {
    const int4 it = plan.items[q];
    const int32_t e0 = it.y, e1 = it.z;
    if (it.w & kBlockLongRow) {
      for (int32_t e = e0; e < e1; ++e) hpos[e] = (uint32_t)e;  // read straight, in CSR order
      return;
    }
    if (e1 - e0 > 65536) {
      ok.store(false);
      return;
    }
    rank_by_source(col, e0, e1 - e0, hpos.data());
    for (int32_t e = e0; e < e1; ++e) lpos[e] = (uint16_t)(hpos[e] - (uint32_t)e0);
  });
  if (!ok.load()) err = "internal error: a row block holds more than 65536 edges";
  return ok.load();
}

static_assert(kBlock * kScatterEpt <= 65536, "scatter ranks are 16-bit");
static_assert(kRegroupChunk <= 65536, "regroup ranks are 16-bit");

void sort_scatter_writes(const HostScatter& sc, std::vector<uint32_t>& pos, std::vector<uint16_t>& rank) {
  rank.assign(pos.size(), 0);
  parallel_for((int64_t)sc.sources.size(), [&](int64_t q) {
    const int4 it = sc.sources[q];
    if (it.w & kBlockLongRow) return;
    const int32_t c0 = it.y, c1 = it.z;
    const int32_t n = c1 - c0;
    std::vector<uint64_t> key((size_t)n);
    for (int32_t l = 0; l < n; ++l) key[l] = ((uint64_t)pos[c0 + l] << 32) | (uint32_t)l;
    std::sort(key.begin(), key.end());
    for (int32_t k = 0; k < n; ++k) {
      const uint32_t l = (uint32_t)(key[k] & 0xffffffffu);
      rank[c0 + l] = (uint16_t)k;
      pos[c0 + k] = (uint32_t)(key[k] >> 32);
    }
  });
}

void encode_scatter_runs(const HostScatter& sc, const std::vector<uint32_t>& pos, HostScatterRuns& out) {
  const int64_t n_items = (int64_t)sc.sources.size();
  out.run_idx.assign(pos.size(), 0);
  out.run_ptr.assign((size_t)n_items + 1, 0);
  auto is_head = [&](int32_t c, int32_t c0) { return c == c0 || pos[c] != pos[c - 1] + 1; };
  parallel_for(n_items, [&](int64_t q) {
    if (sc.sources[q].w & kBlockLongRow) return;
    int32_t runs = 0;
    for (int32_t c = sc.sources[q].y; c < sc.sources[q].z; ++c) runs += is_head(c, sc.sources[q].y) ? 1 : 0;
    out.run_ptr[q + 1] = runs;
  });
  for (int64_t q = 0; q < n_items; ++q) out.run_ptr[q + 1] += out.run_ptr[q];
  out.run_base.assign((size_t)std::max(out.run_ptr[n_items], 1), 0);
  parallel_for(n_items, [&](int64_t q) {
    if (sc.sources[q].w & kBlockLongRow) return;
    const int32_t c0 = sc.sources[q].y, c1 = sc.sources[q].z;
    int32_t r = -1;
    for (int32_t c = c0; c < c1; ++c) {
      if (is_head(c, c0)) {
        ++r;
        out.run_base[out.run_ptr[q] + r] = (int32_t)((int64_t)pos[c] - (c - c0));
      }
      out.run_idx[c] = (uint16_t)r;  // r < 4096: a short block holds at most 4096 edges
    }
  });
}

void window_scatter_codes(const HostScatter& sc, const std::vector<uint16_t>& rank, const HostScatterRuns& runs,
                          std::vector<uint32_t>& code) {
  code.assign(rank.size(), 0);
  parallel_for((int64_t)sc.sources.size(), [&](int64_t q) {
    const int4 it = sc.sources[q];
    if (it.w & kBlockLongRow) return;
    const int32_t c0 = it.y, c1 = it.z;
    for (int32_t c = c0; c < c1; ++c) {
      const int32_t k = rank[c];  // the edge's index in its block's slot order
      code[c0 + k] = (uint32_t)sc.dsrc[c] | (uint32_t)runs.run_idx[c0 + k] << 16;
    }
  });
}

void order_scatter_by_delay(HostScatter& sc) {
  sc.dsrc.assign(sc.delay_csc.size(), 0);
  parallel_for((int64_t)sc.sources.size(), [&](int64_t q) {
    const int4 it = sc.sources[q];
    if (it.w & kBlockLongRow) return;
    const int32_t s0 = it.x, c0 = it.y, c1 = it.z, nsrc = (it.w >> 1) & 511;
    const int32_t n = c1 - c0;
    std::vector<uint64_t> key((size_t)n);
    for (int32_t i = 0; i < nsrc; ++i)
      for (int32_t c = sc.csc_ptr[s0 + i]; c < sc.csc_ptr[s0 + i + 1]; ++c)
        key[c - c0] = ((uint64_t)sc.delay_csc[c] << 40) | ((uint64_t)i << 32) | (uint32_t)(c - c0);
    std::sort(key.begin(), key.end());
    std::vector<uint8_t> d((size_t)n);
    std::vector<uint32_t> e((size_t)n);
    for (int32_t k = 0; k < n; ++k) {
      const uint32_t l = (uint32_t)(key[k] & 0xffffffffu);
      d[k] = sc.delay_csc[c0 + l];
      e[k] = sc.csc_edge[c0 + l];
      sc.dsrc[c0 + k] = (uint16_t)((((key[k] >> 32) & 0xff) << 8) | d[k]);
    }
    std::copy(d.begin(), d.end(), sc.delay_csc.begin() + c0);
    std::copy(e.begin(), e.end(), sc.csc_edge.begin() + c0);
  });
}

bool staged_bucket_positions(int N, const int* row_ptr, const int* col, const HostPlan& plan, int bucket_edges,
                             std::vector<uint32_t>& hpos, std::vector<int32_t>& bucket_begin, std::string& err) {
  const int64_t E = row_ptr[N];
  if (bucket_edges < 1024) {
    err = "bucket_edges must be at least 1024";
    return false;
  }
  // the plan's blocks in row order, each its CSR edge range [e0, e1)
  std::vector<int2> blocks;
  blocks.reserve(plan.items.size());
  for (const int4& it : plan.items)
    if (((it.w >> 1) & 511) > 0) blocks.push_back(make_int2(it.y, it.z));
  std::sort(blocks.begin(), blocks.end(), [](const int2& a, const int2& b) { return a.x < b.x; });
  bucket_begin.assign(1, 0);
  int64_t filled = 0, at = 0;
  for (const int2& b : blocks) {
    if (b.x != at) {
      err = "internal error: row blocks do not tile the edges";
      return false;
    }
    const int64_t n = b.y - b.x;
    if (filled > 0 && filled + n > bucket_edges) {
      bucket_begin.push_back((int32_t)at);
      filled = 0;
    }
    filled += n;
    at = b.y;
  }
  if (at != E) {
    err = "internal error: row blocks do not tile the edges";
    return false;
  }
  if (bucket_begin.back() != (int32_t)E) bucket_begin.push_back((int32_t)E);
  hpos.assign((size_t)E, 0);
  parallel_for((int64_t)bucket_begin.size() - 1, [&](int64_t q) {
    rank_by_source(col, bucket_begin[q], bucket_begin[q + 1] - bucket_begin[q], hpos.data());
  });
  return true;
}

bool build_regroup(const std::vector<uint32_t>& bucket_hpos, const std::vector<uint32_t>& staged_hpos,
                   const std::vector<int32_t>& bucket_begin, HostRegroup& out, std::string& err) {
  const size_t E = bucket_hpos.size();
  // dst[i]: staged slot of the value in bucket slot i
  std::vector<uint32_t> dst(E);
  for (size_t e = 0; e < E; ++e) dst[bucket_hpos[e]] = staged_hpos[e];
  out.rank.assign(E, 0);
  const int64_t nb = (int64_t)bucket_begin.size() - 1;
  std::vector<std::vector<int4>> chunks(nb);
  std::vector<std::vector<int2>> runs(nb);
  std::atomic<bool> ok{true};
  parallel_for(nb, [&](int64_t q) {
    std::vector<uint64_t> key;
    int64_t c0 = bucket_begin[q];
    const int64_t b0 = bucket_begin[q], b1 = bucket_begin[q + 1];
    while (c0 < b1) {
      // the longest chunk from c0 within the slot and run limits
      int64_t n = std::min<int64_t>(kRegroupChunk, b1 - c0);
      for (;;) {
        key.resize((size_t)n);
        for (int64_t l = 0; l < n; ++l) key[l] = ((uint64_t)dst[c0 + l] << 32) | (uint64_t)l;
        std::sort(key.begin(), key.end());
        int64_t nr = 1;
        for (int64_t k = 1; k < n; ++k) nr += (key[k] >> 32) != (key[k - 1] >> 32) + 1;
        if (nr <= kRegroupRuns) break;
        n = std::max<int64_t>(1, n * kRegroupRuns / (nr + 1));
      }
      const int32_t run0 = (int32_t)runs[q].size();
      for (int64_t k = 0; k < n; ++k) {
        const uint32_t d = (uint32_t)(key[k] >> 32);
        if (k == 0 || d != (uint32_t)(key[k - 1] >> 32) + 1) runs[q].push_back(make_int2((int)k, (int)d));
        out.rank[c0 + (key[k] & 0xffffffffu)] = (uint16_t)k;
        if (d < (uint32_t)b0 || d >= (uint32_t)b1) ok.store(false);
      }
      chunks[q].push_back(make_int4((int)c0, (int)n, run0, (int)runs[q].size() - run0));
      c0 += n;
    }
  });
  if (!ok.load()) {
    err = "internal error: a staged slot lies outside its bucket";
    return false;
  }
  out.chunks.clear();
  out.runs.clear();
  for (int64_t q = 0; q < nb; ++q) {
    const int32_t base = (int32_t)out.runs.size();
    for (int4 c : chunks[q]) {
      c.z += base;
      out.chunks.push_back(c);
    }
    out.runs.insert(out.runs.end(), runs[q].begin(), runs[q].end());
  }
  return true;
}

bool csc_positions(const HostScatter& sc, const std::vector<uint32_t>& hpos, std::vector<uint32_t>& pos,
                   std::string& err) {
  const size_t E = sc.csc_edge.size();
  pos.assign(E, 0);
  std::vector<uint8_t> seen(E, 0);
  for (size_t c = 0; c < E; ++c) {
    const uint32_t p = hpos[sc.csc_edge[c]];
    // every slot is written exactly once per step: the positions are a permutation of [0, E)
    if (p >= E || seen[p]) {
      err = "internal error: value slots are not a permutation";
      return false;
    }
    seen[p] = 1;
    pos[c] = p;
  }
  return true;
}

void early_edges(int N, const int* row_ptr, const int* col_idx, const uint8_t* delays, int below,
                 std::vector<int32_t>& src, std::vector<uint8_t>& delay, std::vector<uint32_t>& csr,
                 std::vector<int32_t>& cnt) {
  std::vector<std::vector<uint64_t>> keys(std::max(below, 0));
  for (int i = 0; i < N; ++i)
    for (int e = row_ptr[i]; e < row_ptr[i + 1]; ++e)
      if (delays[e] < below) keys[delays[e]].push_back((uint64_t)(uint32_t)col_idx[e] << 32 | (uint32_t)e);
  cnt.assign((size_t)std::max(below, 0) + 1, 0);
  src.clear();
  delay.clear();
  csr.clear();
  for (int d = 0; d < below; ++d) {
    std::sort(keys[d].begin(), keys[d].end());
    for (uint64_t k : keys[d]) {
      src.push_back((int32_t)(k >> 32));
      delay.push_back((uint8_t)d);
      csr.push_back((uint32_t)k);
    }
    cnt[d + 1] = (int32_t)src.size();
  }
}

bool block_early_edges(const HostPlan& plan, const std::vector<int32_t>& src, const std::vector<uint8_t>& delay,
                       const std::vector<uint32_t>& csr, const std::vector<uint32_t>& staged_hpos,
                       std::vector<int32_t>& ptr, std::vector<int2>& rec, std::string& err) {
  const size_t n_items = plan.items.size();
  // items by first edge (long rows come first in launch order), to find each edge's item
  std::vector<int32_t> by_e0(n_items);
  for (size_t q = 0; q < n_items; ++q) by_e0[q] = (int32_t)q;
  std::sort(by_e0.begin(), by_e0.end(), [&](int32_t x, int32_t y) { return plan.items[x].y < plan.items[y].y; });
  std::vector<int32_t> starts(n_items);
  for (size_t k = 0; k < n_items; ++k) starts[k] = plan.items[by_e0[k]].y;
  std::vector<int32_t> item_of(csr.size(), -1);
  std::atomic<bool> ok{true};
  parallel_for((int64_t)csr.size(), [&](int64_t r) {
    const int32_t e = (int32_t)csr[r];
    size_t k = (size_t)(std::upper_bound(starts.begin(), starts.end(), e) - starts.begin());
    // padding items hold no edges: step back over empty ranges to the item that holds e
    while (k > 0 && !(plan.items[by_e0[k - 1]].y <= e && e < plan.items[by_e0[k - 1]].z)) --k;
    if (k == 0 || delay[r] > 3)
      ok.store(false);
    else
      item_of[r] = by_e0[k - 1];
  });
  if (!ok.load()) {
    err = "internal error: an early edge outside every row block, or a delay above 3";
    return false;
  }
  ptr.assign(n_items + 1, 0);
  for (int32_t q : item_of) ++ptr[(size_t)q + 1];
  for (size_t q = 0; q < n_items; ++q) ptr[q + 1] += ptr[q];
  rec.assign(csr.size(), make_int2(0, 0));
  std::vector<int32_t> fill(ptr.begin(), ptr.end() - 1);
  for (size_t r = 0; r < csr.size(); ++r) {
    const int4 it = plan.items[item_of[r]];
    const int64_t local = (int64_t)staged_hpos[csr[r]] - it.y;
    if (local < 0 || local >= ((int64_t)1 << 29)) {
      err = "internal error: an early edge's staged slot lies outside its row block";
      return false;
    }
    rec[fill[item_of[r]]++] = make_int2(src[r], (int32_t)(local << 2) | delay[r]);
  }
  return true;
}

}  // namespace lev
