// Multilevel k-way partitioning of an element adjacency graph (CSR), the job
// the reference hands to METIS's PartGraphKway
// (mesh/partitioner/mesh_partitioner.go:87-142, graph built at :145-184).
//
// Written for this library, host code only, single-threaded and deterministic
// for a given seed:
//   1. coarsening by heavy-edge matching (random visit order, pair weight
//      capped so no coarse vertex outweighs a small fraction of a part);
//   2. on the coarsest graph, recursive bisection: greedy graph growing from
//      several random seeds, each refined by Fiduccia-Mattheyses passes, the
//      lowest balanced cut kept;
//   3. on the way back up, projection and greedy boundary refinement under the
//      balance bound max part weight <= ub * total / nparts; moves out of an
//      overweight part are taken even at a loss.
// Objective "cut" minimises the summed weight of edges between parts; "vol"
// minimises the total communication volume (per vertex, the number of
// distinct other parts among its neighbours), applied in the refinement of
// the input graph (coarse levels refine the cut, which bounds the volume).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <exception>
#include <numeric>
#include <queue>
#include <utility>
#include <vector>

#include "dgk_internal.h"

namespace {

struct Rng {
  uint64_t s;
  uint64_t next() {  // splitmix64
    s += 0x9E3779B97F4A7C15ull;
    uint64_t z = s;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  uint32_t below(uint32_t n) { return (uint32_t)(((next() >> 32) * (uint64_t)n) >> 32); }
};

struct Graph {
  int n = 0;
  std::vector<int64_t> xadj{0};
  std::vector<int> adj;
  std::vector<int64_t> ew;
  std::vector<int64_t> vw;
  int64_t total() const { return std::accumulate(vw.begin(), vw.end(), (int64_t)0); }
};

// Visit order for matching: random within consecutive blocks of 4096
// vertices, blocks in order (the graph's own numbering usually has locality,
// which this keeps for the caches).
void shuffled_blocks(int n, Rng& rng, std::vector<int>& p) {
  p.resize(n);
  std::iota(p.begin(), p.end(), 0);
  for (int b = 0; b < n; b += 4096) {
    const int m = std::min(4096, n - b);
    for (int i = m - 1; i > 0; --i) std::swap(p[b + i], p[b + rng.below(i + 1)]);
  }
}

// Heavy-edge matching and contraction. cmap[v] = coarse vertex of v.
Graph coarsen(const Graph& g, int64_t maxvw, Rng& rng, std::vector<int>& cmap,
              const std::vector<int>* within = nullptr) {
  const int n = g.n;
  std::vector<int> match(n, -1), perm;
  shuffled_blocks(n, rng, perm);
  for (int v : perm) {
    if (match[v] >= 0) continue;
    int best = v;
    int64_t bw = -1;
    for (int64_t e = g.xadj[v]; e < g.xadj[v + 1]; ++e) {
      const int u = g.adj[e];
      if (u == v || match[u] >= 0 || g.vw[v] + g.vw[u] > maxvw) continue;
      if (within && (*within)[u] != (*within)[v]) continue;
      if (g.ew[e] > bw || (g.ew[e] == bw && g.vw[u] < g.vw[best])) {
        bw = g.ew[e];
        best = u;
      }
    }
    match[v] = best;
    match[best] = v;
  }
  cmap.assign(n, -1);
  std::vector<int> first;
  first.reserve(n / 2 + 1);
  for (int v = 0; v < n; ++v) {
    if (cmap[v] >= 0) continue;
    cmap[v] = cmap[match[v]] = (int)first.size();
    first.push_back(v);
  }
  Graph c;
  c.n = (int)first.size();
  c.vw.resize(c.n);
  c.xadj.reserve(c.n + 1);
  c.adj.reserve(g.adj.size() / 2 + 16);
  c.ew.reserve(g.adj.size() / 2 + 16);
  std::vector<int64_t> pos(c.n, -1);  // slot of coarse neighbour t in the current row
  for (int cv = 0; cv < c.n; ++cv) {
    const int v = first[cv], u = match[v];
    c.vw[cv] = g.vw[v] + (u != v ? g.vw[u] : 0);
    const int64_t row = (int64_t)c.adj.size();
    for (int w : {v, u}) {
      for (int64_t e = g.xadj[w]; e < g.xadj[w + 1]; ++e) {
        const int t = cmap[g.adj[e]];
        if (t == cv) continue;
        if (pos[t] >= row) {
          c.ew[pos[t]] += g.ew[e];
        } else {
          pos[t] = (int64_t)c.adj.size();
          c.adj.push_back(t);
          c.ew.push_back(g.ew[e]);
        }
      }
      if (u == v) break;
    }
    c.xadj.push_back((int64_t)c.adj.size());
  }
  return c;
}

int64_t cut_of(const Graph& g, const std::vector<int>& where) {
  int64_t cut = 0;
  for (int v = 0; v < g.n; ++v)
    for (int64_t e = g.xadj[v]; e < g.xadj[v + 1]; ++e)
      if (where[g.adj[e]] != where[v]) cut += g.ew[e];
  return cut / 2;
}

// Distinct parts other than where[u] among u's neighbours.
int64_t vol_of(const Graph& g, const std::vector<int>& where, int u) {
  int seen[64];
  int ns = 0;
  int64_t extra = 0;
  for (int64_t e = g.xadj[u]; e < g.xadj[u + 1]; ++e) {
    const int p = where[g.adj[e]];
    if (p == where[u]) continue;
    bool dup = false;
    for (int i = 0; i < ns; ++i) dup |= seen[i] == p;
    if (dup) continue;
    if (ns < 64) {
      seen[ns++] = p;
    } else {  // high-degree vertex: count exactly with a second scan
      bool before = false;
      for (int64_t f = g.xadj[u]; f < e; ++f) before |= where[g.adj[f]] == p;
      if (!before) ++extra;
    }
  }
  return ns + extra;
}

int64_t volume(const Graph& g, const std::vector<int>& where) {
  int64_t s = 0;
  for (int v = 0; v < g.n; ++v) s += vol_of(g, where, v);
  return s;
}

// Change of the total volume if v moves to part b.
int64_t vol_delta(const Graph& g, std::vector<int>& where, int v, int b) {
  const int a = where[v];
  int64_t before = vol_of(g, where, v), after;
  for (int64_t e = g.xadj[v]; e < g.xadj[v + 1]; ++e) before += vol_of(g, where, g.adj[e]);
  where[v] = b;
  after = vol_of(g, where, v);
  for (int64_t e = g.xadj[v]; e < g.xadj[v + 1]; ++e) after += vol_of(g, where, g.adj[e]);
  where[v] = a;
  return after - before;
}

// ---- multilevel bisection -----------------------------------------------------

// Fiduccia-Mattheyses refinement of a two-way split (maxw[s] bounds side s):
// per pass, vertices move one at a time, highest gain first (max-heaps per
// side, stale entries skipped), each at most once; the pass is rolled back
// to its best prefix (least overweight, then lowest cut) and stops after
// `limit` moves without a new best.
void fm2(const Graph& g, std::vector<int>& side, const int64_t maxw[2], int passes, int limit) {
  const int n = g.n;
  std::vector<int64_t> gain(n);
  std::vector<char> locked(n);
  std::vector<int> moves, deferred;
  int64_t w[2] = {0, 0};
  for (int v = 0; v < n; ++v) w[side[v]] += g.vw[v];
  int64_t cut = cut_of(g, side);
  // A move may overfill its destination by one vertex weight; the next move
  // then leaves that side (classic FM swapping), and only prefixes that end
  // within the bounds can become the pass's best.
  const int64_t slack = n ? *std::max_element(g.vw.begin(), g.vw.end()) : 0;
  auto excess = [&]() {
    return std::max<int64_t>(0, w[0] - maxw[0]) + std::max<int64_t>(0, w[1] - maxw[1]);
  };
  using Entry = std::pair<int64_t, int>;
  for (int pass = 0; pass < passes; ++pass) {
    std::priority_queue<Entry> heap[2];
    for (int v = 0; v < n; ++v) {
      int64_t ext = 0, in = 0;
      for (int64_t e = g.xadj[v]; e < g.xadj[v + 1]; ++e)
        (side[g.adj[e]] != side[v] ? ext : in) += g.ew[e];
      gain[v] = ext - in;
      if (ext > 0 || n <= 256) heap[side[v]].push({gain[v], v});
    }
    std::fill(locked.begin(), locked.end(), 0);
    moves.clear();
    int64_t best_cut = cut, best_ex = excess();
    size_t best_at = 0;
    int since = 0;
    auto top = [&](int s) -> int {  // best valid entry of side s, or -1
      while (!heap[s].empty()) {
        const Entry t = heap[s].top();
        const int v = t.second;
        if (locked[v] || side[v] != s || gain[v] != t.first) {
          heap[s].pop();
          continue;
        }
        return v;
      }
      return -1;
    };
    while (since < limit) {
      const int64_t ex = excess();
      int pick = -1;
      deferred.clear();
      for (int tries = 0; tries < 32 && pick < 0; ++tries) {
        int cand[2] = {top(0), top(1)};
        int s;
        if (ex > 0) {
          s = w[0] - maxw[0] > w[1] - maxw[1] ? 0 : 1;  // leave the heavy side
        } else if (cand[0] < 0 || cand[1] < 0) {
          s = cand[0] < 0 ? 1 : 0;
        } else {
          s = gain[cand[0]] >= gain[cand[1]] ? 0 : 1;
        }
        const int v = cand[s];
        if (v < 0) {
          if (ex > 0 || cand[1 - s] < 0) break;
          s = 1 - s;
        }
        const int u = cand[s];
        if (u < 0) break;
        if (ex == 0 && w[1 - s] + g.vw[u] > maxw[1 - s] + slack) {  // does not fit: set aside
          heap[s].pop();
          deferred.push_back(u);
          continue;
        }
        heap[s].pop();
        pick = u;
      }
      for (int v : deferred) heap[side[v]].push({gain[v], v});
      if (pick < 0) break;
      const int s = side[pick];
      side[pick] = 1 - s;
      w[s] -= g.vw[pick];
      w[1 - s] += g.vw[pick];
      cut -= gain[pick];
      locked[pick] = 1;
      moves.push_back(pick);
      for (int64_t e = g.xadj[pick]; e < g.xadj[pick + 1]; ++e) {
        const int u = g.adj[e];
        gain[u] += (side[u] == side[pick] ? -2 : 2) * g.ew[e];
        if (!locked[u]) heap[side[u]].push({gain[u], u});
      }
      const int64_t nex = excess();
      if (nex < best_ex || (nex == best_ex && cut < best_cut)) {
        best_ex = nex;
        best_cut = cut;
        best_at = moves.size();
        since = 0;
      } else {
        ++since;
      }
    }
    for (size_t i = moves.size(); i > best_at; --i) {
      const int v = moves[i - 1];
      const int s = side[v];
      side[v] = 1 - s;
      w[s] -= g.vw[v];
      w[1 - s] += g.vw[v];
    }
    cut = best_cut;
    if (best_at == 0) break;
  }
}

// Bisection of a small graph: greedy growing of side 0 from several random
// seeds (highest gain on the frontier first), each followed by FM; the
// least overweight, then lowest cut, split is kept.
void grow_bisect(const Graph& g, const int64_t maxw[2], int64_t t0, Rng& rng,
                 std::vector<int>& side) {
  const int n = g.n;
  std::vector<int> best, cur(n);
  int64_t best_cut = -1, best_ex = 0;
  std::vector<int64_t> gain(n);
  std::vector<char> front(n);
  const int tries = n < 8 ? 1 : 10;
  for (int trial = 0; trial < tries; ++trial) {
    std::fill(cur.begin(), cur.end(), 1);
    std::fill(front.begin(), front.end(), 0);
    std::vector<int> frontier;
    int64_t w0 = 0;
    for (int v = 0; v < n; ++v) {
      int64_t s = 0;
      for (int64_t e = g.xadj[v]; e < g.xadj[v + 1]; ++e) s += g.ew[e];
      gain[v] = -s;  // every edge internal to side 1 at first
    }
    while (w0 < t0) {
      int pick = -1;
      size_t at = 0;
      for (size_t i = 0; i < frontier.size(); ++i) {
        const int v = frontier[i];
        if (pick < 0 || gain[v] > gain[pick]) {
          pick = v;
          at = i;
        }
      }
      if (pick >= 0) {
        frontier[at] = frontier.back();
        frontier.pop_back();
      } else {  // start (or restart in a new component) at a random side-1 vertex
        int left = 0;
        for (int v = 0; v < n; ++v) left += cur[v];
        if (left == 0) break;
        int r = (int)rng.below(left);
        for (int v = 0; v < n; ++v)
          if (cur[v] && r-- == 0) {
            pick = v;
            break;
          }
      }
      if (w0 > 0 && w0 + g.vw[pick] > maxw[0]) break;
      cur[pick] = 0;
      w0 += g.vw[pick];
      for (int64_t e = g.xadj[pick]; e < g.xadj[pick + 1]; ++e) {
        const int u = g.adj[e];
        if (cur[u] != 1) continue;
        gain[u] += 2 * g.ew[e];
        if (!front[u]) {
          front[u] = 1;
          frontier.push_back(u);
        }
      }
    }
    fm2(g, cur, maxw, 8, 64);
    int64_t w[2] = {0, 0};
    for (int v = 0; v < n; ++v) w[cur[v]] += g.vw[v];
    const int64_t ex = std::max<int64_t>(0, w[0] - maxw[0]) + std::max<int64_t>(0, w[1] - maxw[1]);
    const int64_t c = cut_of(g, cur);
    if (best_cut < 0 || ex < best_ex || (ex == best_ex && c < best_cut)) {
      best = cur;
      best_cut = c;
      best_ex = ex;
    }
  }
  side = best;
}

// Multilevel bisection of g into side 0 (weight fraction f0, bound f0 *
// total * tol) and side 1: heavy-edge coarsening to ~200 vertices, grown
// bisection there, then projection and FM on every level back up.
void bisect(const Graph& g, double f0, double tol, Rng& rng, std::vector<int>& side) {
  const int64_t tot = g.total();
  const int64_t t0 = (int64_t)(f0 * (double)tot + 0.5);
  const int64_t maxw[2] = {std::max<int64_t>((int64_t)((double)t0 * tol), t0),
                           std::max<int64_t>((int64_t)((double)(tot - t0) * tol), tot - t0)};
  const int target = 200;
  std::vector<Graph> levels;
  std::vector<std::vector<int>> cmaps;
  const Graph* cur = &g;
  const int64_t cap = std::max<int64_t>(1, (int64_t)(0.05 * (double)std::min(t0, tot - t0)) + 1);
  while (cur->n > target) {
    std::vector<int> cmap;
    Graph c = coarsen(*cur, cap, rng, cmap);
    if (c.n > (int)(0.95 * cur->n)) break;  // matching stalled
    levels.push_back(std::move(c));
    cmaps.push_back(std::move(cmap));
    cur = &levels.back();
  }
  grow_bisect(*cur, maxw, t0, rng, side);
  for (int l = (int)levels.size() - 1; l >= 0; --l) {
    const Graph& fine = l > 0 ? levels[l - 1] : g;
    std::vector<int> s(fine.n);
    for (int v = 0; v < fine.n; ++v) s[v] = side[cmaps[l][v]];
    side.swap(s);
    fm2(fine, side, maxw, 6, 256);
  }
}

Graph induced(const Graph& g, const std::vector<int>& side, int s, std::vector<int>& ids) {
  std::vector<int> loc(g.n, -1);
  ids.clear();
  for (int v = 0; v < g.n; ++v)
    if (side[v] == s) {
      loc[v] = (int)ids.size();
      ids.push_back(v);
    }
  Graph h;
  h.n = (int)ids.size();
  h.vw.resize(h.n);
  h.xadj.reserve(h.n + 1);
  for (int i = 0; i < h.n; ++i) {
    const int v = ids[i];
    h.vw[i] = g.vw[v];
    for (int64_t e = g.xadj[v]; e < g.xadj[v + 1]; ++e)
      if (loc[g.adj[e]] >= 0) {
        h.adj.push_back(loc[g.adj[e]]);
        h.ew.push_back(g.ew[e]);
      }
    h.xadj.push_back((int64_t)h.adj.size());
  }
  return h;
}

// Parts base..base+k-1 by recursive multilevel bisection (k0 = k/2 parts on
// side 0, weight in proportion).
void recursive_bisection(const Graph& g, int k, int base, double tol, Rng& rng,
                         std::vector<int>& where) {
  where.assign(g.n, base);
  if (k <= 1 || g.n == 0) return;
  const int k0 = k / 2;
  std::vector<int> side;
  bisect(g, (double)k0 / k, tol, rng, side);
  for (int s = 0; s < 2; ++s) {
    std::vector<int> ids, sub;
    {
      Graph h = induced(g, side, s, ids);
      recursive_bisection(h, s ? k - k0 : k0, s ? base + k0 : base, tol, rng, sub);
    }
    for (size_t i = 0; i < ids.size(); ++i) where[ids[i]] = sub[i];
  }
}

// ---- k-way refinement ---------------------------------------------------------

void refine_kway(const Graph& g, std::vector<int>& where, int k, int64_t maxpw, bool vol,
                 Rng& rng, int passes) {
  const int n = g.n;
  std::vector<int64_t> pw(k, 0), conn(k, 0);
  for (int v = 0; v < n; ++v) pw[where[v]] += g.vw[v];
  std::vector<int> touched, perm;
  auto consider = [&](int v, bool force) -> bool {
    const int a = where[v];
    touched.clear();
    for (int64_t e = g.xadj[v]; e < g.xadj[v + 1]; ++e) {
      const int p = where[g.adj[e]];
      if (conn[p] == 0 && p != a) touched.push_back(p);
      conn[p] += g.ew[e];
    }
    int best = -1;
    int64_t bgain = 0;
    const bool over = pw[a] > maxpw;
    for (int b : touched) {
      if (pw[b] + g.vw[v] > maxpw && !(force && pw[b] + g.vw[v] < pw[a])) continue;
      const int64_t gn = vol ? -vol_delta(g, where, v, b) : conn[b] - conn[a];
      if (best < 0 || gn > bgain || (gn == bgain && pw[b] < pw[best])) {
        best = b;
        bgain = gn;
      }
    }
    for (int64_t e = g.xadj[v]; e < g.xadj[v + 1]; ++e) conn[where[g.adj[e]]] = 0;
    if (best < 0) return false;
    const bool ok = bgain > 0 || (bgain == 0 && pw[best] + g.vw[v] < pw[a]) || over;
    if (!ok) return false;
    where[v] = best;
    pw[a] -= g.vw[v];
    pw[best] += g.vw[v];
    return true;
  };
  for (int pass = 0; pass < passes; ++pass) {
    // This pass visits the vertices on a part boundary, in random order.
    perm.clear();
    for (int v = 0; v < n; ++v)
      for (int64_t e = g.xadj[v]; e < g.xadj[v + 1]; ++e)
        if (where[g.adj[e]] != where[v]) {
          perm.push_back(v);
          break;
        }
    for (int i = (int)perm.size() - 1; i > 0; --i) std::swap(perm[i], perm[rng.below(i + 1)]);
    int64_t moved = 0;
    for (int v : perm) moved += consider(v, false);
    if (moved == 0) break;
  }
  // Balance: parts still above the bound give boundary vertices to lighter
  // neighbours (best gain first, losses allowed) until they fit.
  for (int round = 0; round < 8; ++round) {
    bool heavy = false;
    for (int p = 0; p < k; ++p) heavy |= pw[p] > maxpw;
    if (!heavy) break;
    for (int v = 0; v < n; ++v)
      if (pw[where[v]] > maxpw) consider(v, true);
  }
}

int partition_impl(int64_t nv, const int64_t* xadj, const int64_t* adjncy, const int64_t* vwgt,
                   const int64_t* adjwgt, int nparts, double ubfactor, int objective, uint64_t seed,
                   int32_t* part, int64_t* objval) {
  DGK_REQUIRE(nv >= 0 && nv < (int64_t)INT32_MAX, "dgk_partition_kway: vertex count %lld out of range",
              (long long)nv);
  DGK_REQUIRE(nparts >= 1, "dgk_partition_kway: nparts must be >= 1 (got %d)", nparts);
  DGK_REQUIRE(ubfactor >= 1.0, "dgk_partition_kway: imbalance factor must be >= 1 (got %g)", ubfactor);
  DGK_REQUIRE(objective == 0 || objective == 1, "dgk_partition_kway: objective must be 0 (cut) or 1 (vol)");
  DGK_REQUIRE(xadj != nullptr && (nv == 0 || (part != nullptr)), "dgk_partition_kway: NULL argument");
  DGK_REQUIRE(xadj[0] == 0, "dgk_partition_kway: xadj[0] must be 0");
  if (objval) *objval = 0;
  if (nv == 0) return DGK_OK;
  DGK_REQUIRE(adjncy != nullptr || xadj[nv] == 0, "dgk_partition_kway: adjncy is NULL");
  Graph g;
  g.n = (int)nv;
  g.vw.resize(nv);
  g.xadj.resize(nv + 1);
  g.xadj[0] = 0;
  for (int64_t v = 0; v < nv; ++v) {
    DGK_REQUIRE(xadj[v + 1] >= xadj[v], "dgk_partition_kway: xadj decreases at %lld", (long long)v);
    g.vw[v] = vwgt ? vwgt[v] : 1;
    DGK_REQUIRE(g.vw[v] >= 0, "dgk_partition_kway: negative vertex weight at %lld", (long long)v);
    for (int64_t e = xadj[v]; e < xadj[v + 1]; ++e) {
      const int64_t u = adjncy[e];
      DGK_REQUIRE(u >= 0 && u < nv, "dgk_partition_kway: adjncy[%lld] = %lld out of range",
                  (long long)e, (long long)u);
      const int64_t w = adjwgt ? adjwgt[e] : 1;
      DGK_REQUIRE(w >= 0, "dgk_partition_kway: negative edge weight at %lld", (long long)e);
      if (u == v) continue;  // self loops carry no cut
      g.adj.push_back((int)u);
      g.ew.push_back(w);
    }
    g.xadj[v + 1] = (int64_t)g.adj.size();
  }
  std::vector<int> where;
  if (nparts == 1) {
    where.assign(nv, 0);
  } else {
    Rng rng{seed * 0x2545F4914F6CDD1Dull + 1};
    const int64_t tot = g.total();
    const int64_t maxpw =
        std::max<int64_t>((int64_t)(ubfactor * (double)tot / nparts), (tot + nparts - 1) / nparts);
    // Recursive multilevel bisection. Each bisection may use the allowed
    // imbalance to the power min(1, 3 / depth): looser than an equal share
    // per level (which costs cut), and the k-way passes below restore the
    // bound on the parts.
    int depth = 0;
    while ((1 << depth) < nparts) ++depth;
    recursive_bisection(g, nparts, 0, std::pow(ubfactor, std::min(1.0, 3.0 / depth)), rng, where);
    // k-way V-cycle: coarsen inside the parts, refine from the coarsest
    // level down, so whole regions of a boundary can move at once.
    std::vector<Graph> levels;
    std::vector<std::vector<int>> cmaps, wheres{where};
    const Graph* cur = &g;
    const int vtarget = std::max(64 * nparts, 1024);
    const int64_t cap = std::max<int64_t>(1, tot / vtarget + 1);
    while (cur->n > vtarget) {
      std::vector<int> cmap;
      Graph c = coarsen(*cur, cap, rng, cmap, &wheres.back());
      if (c.n > (int)(0.9 * cur->n)) break;
      std::vector<int> cw(c.n);
      for (int v = 0; v < cur->n; ++v) cw[cmap[v]] = wheres.back()[v];
      levels.push_back(std::move(c));
      cmaps.push_back(std::move(cmap));
      wheres.push_back(std::move(cw));
      cur = &levels.back();
    }
    std::vector<int> wc = wheres.back();
    for (int l = (int)levels.size(); l >= 0; --l) {
      const Graph& gl = l > 0 ? levels[l - 1] : g;
      if (l < (int)levels.size()) {
        std::vector<int> w(gl.n);
        for (int v = 0; v < gl.n; ++v) w[v] = wc[cmaps[l][v]];
        wc.swap(w);
      }
      refine_kway(gl, wc, nparts, maxpw, objective == 1 && l == 0, rng, 8);
    }
    where.swap(wc);
  }
  for (int64_t v = 0; v < nv; ++v) part[v] = where[v];
  if (objval) *objval = objective == 1 ? volume(g, where) : cut_of(g, where);
  return DGK_OK;
}

}  // namespace

extern "C" int dgk_partition_kway(int64_t nv, const int64_t* xadj, const int64_t* adjncy,
                                  const int64_t* vwgt, const int64_t* adjwgt, int nparts,
                                  double ubfactor, int objective, uint64_t seed, int32_t* part,
                                  int64_t* objval) {
  dgk::clear_error();
  try {
    return partition_impl(nv, xadj, adjncy, vwgt, adjwgt, nparts, ubfactor, objective, seed, part,
                          objval);
  } catch (const std::exception& e) {
    dgk::set_error("dgk_partition_kway: %s", e.what());
    return DGK_ERR_INVALID;
  }
}
