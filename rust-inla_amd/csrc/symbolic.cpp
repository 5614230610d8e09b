// Symbolic analysis: ordering chooser, approximate minimum degree, elimination
// tree, column counts, relaxed supernodes, level schedule and index maps.
#include "symbolic.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <deque>
#include <numeric>
#include <stdexcept>

namespace gmrf {

// ---------------------------------------------------------------------------
// Ordering chooser (crates/inla_math/src/ordering.rs). Integer-exact restatement.
// ---------------------------------------------------------------------------

// ordering.rs:8-18 — max(row - col) over stored entries with row >= col.
int64_t lower_bandwidth(int64_t n, const int64_t* indptr, const int64_t* indices) {
  int64_t bw = 0;
  for (int64_t c = 0; c < n; ++c)
    for (int64_t p = indptr[c]; p < indptr[c + 1]; ++p) {
      int64_t r = indices[p];
      if (r >= c) bw = std::max(bw, r - c);
    }
  return bw;
}

// ordering.rs:21-30 — bandwidth after new_index = fwd[old_index], both triangles.
int64_t bandwidth_with_fwd(int64_t n, const int64_t* indptr, const int64_t* indices,
                           const std::vector<int64_t>& fwd) {
  int64_t bw = 0;
  for (int64_t c = 0; c < n; ++c) {
    int64_t jc = fwd[c];
    for (int64_t p = indptr[c]; p < indptr[c + 1]; ++p) {
      int64_t d = fwd[indices[p]] - jc;
      bw = std::max(bw, d < 0 ? -d : d);
    }
  }
  return bw;
}

// ordering.rs:36-49 — mixture-major c*n_time+t -> time-major t*n_comp+c.
static void mixture_to_time_major(int64_t n_time, int64_t n_comp, std::vector<int64_t>& fwd,
                                  std::vector<int64_t>& inv) {
  int64_t n = n_time * n_comp;
  fwd.assign(n, 0);
  inv.assign(n, 0);
  for (int64_t c = 0; c < n_comp; ++c)
    for (int64_t t = 0; t < n_time; ++t) {
      int64_t o = c * n_time + t, w = t * n_comp + c;
      fwd[o] = w;
      inv[w] = o;
    }
}

// ordering.rs:56-89.
bool try_time_major_kronecker(int64_t n, const int64_t* indptr, const int64_t* indices,
                              std::vector<int64_t>& fwd_out, std::vector<int64_t>& inv_out) {
  if (n < 16) return false;
  int64_t bw = lower_bandwidth(n, indptr, indices);
  if (bw < n / 8) return false;
  bool have = false;
  int64_t best_bw = 0;
  std::vector<int64_t> best_fwd, best_inv, fwd, inv;
  for (int64_t n_comp = 2; n_comp <= 8; ++n_comp) {
    if (n % n_comp != 0) continue;
    int64_t n_time = n / n_comp;
    if (n_time < 8) continue;
    mixture_to_time_major(n_time, n_comp, fwd, inv);
    int64_t bw_new = bandwidth_with_fwd(n, indptr, indices, fwd);
    if (bw_new < bw) {
      if (have && best_bw <= bw_new) {
        // keep the earlier (smaller n_comp) candidate on ties
      } else {
        have = true;
        best_bw = bw_new;
        best_fwd.swap(fwd);
        best_inv.swap(inv);
      }
    }
  }
  if (!have) return false;
  int64_t nnz = indptr[n];
  int64_t avg = std::max<int64_t>(nnz / std::max<int64_t>(n, 1), 2);
  if (best_bw * 4 < bw || best_bw <= 4 * avg) {
    fwd_out.swap(best_fwd);
    inv_out.swap(best_inv);
    return true;
  }
  return false;
}

// ordering.rs:92-140 — BFS from the lowest-(degree, index) unvisited node,
// neighbours visited in (degree, index) order, final order reversed.
void reverse_cuthill_mckee(int64_t n, const int64_t* indptr, const int64_t* indices,
                           std::vector<int64_t>& fwd, std::vector<int64_t>& inv) {
  std::vector<std::vector<int64_t>> adj(n);
  for (int64_t c = 0; c < n; ++c)
    for (int64_t p = indptr[c]; p < indptr[c + 1]; ++p) {
      int64_t r = indices[p];
      if (r != c) {
        adj[c].push_back(r);
        adj[r].push_back(c);
      }
    }
  for (auto& a : adj) {
    std::sort(a.begin(), a.end());
    a.erase(std::unique(a.begin(), a.end()), a.end());
  }
  std::vector<int64_t> deg(n);
  for (int64_t i = 0; i < n; ++i) deg[i] = (int64_t)adj[i].size();
  for (auto& a : adj)
    std::stable_sort(a.begin(), a.end(), [&](int64_t x, int64_t y) {
      return deg[x] != deg[y] ? deg[x] < deg[y] : x < y;
    });
  std::vector<char> visited(n, 0);
  std::vector<int64_t> order;
  order.reserve(n);
  // Scan for the lowest-(deg, index) unvisited start; nodes are sorted once so
  // each restart is a pointer advance rather than an O(n) search.
  std::vector<int64_t> by_deg(n);
  std::iota(by_deg.begin(), by_deg.end(), 0);
  std::stable_sort(by_deg.begin(), by_deg.end(), [&](int64_t x, int64_t y) {
    return deg[x] != deg[y] ? deg[x] < deg[y] : x < y;
  });
  size_t scan = 0;
  std::deque<int64_t> queue;
  while (true) {
    while (scan < by_deg.size() && visited[by_deg[scan]]) ++scan;
    if (scan == by_deg.size()) break;
    int64_t start = by_deg[scan];
    visited[start] = 1;
    queue.push_back(start);
    while (!queue.empty()) {
      int64_t u = queue.front();
      queue.pop_front();
      order.push_back(u);
      for (int64_t v : adj[u])
        if (!visited[v]) {
          visited[v] = 1;
          queue.push_back(v);
        }
    }
  }
  std::reverse(order.begin(), order.end());
  fwd.assign(n, 0);
  inv.assign(n, 0);
  for (int64_t w = 0; w < n; ++w) {
    fwd[order[w]] = w;
    inv[w] = order[w];
  }
}

// ordering.rs:149-162.
OrderKind choose_symmetric_order(int64_t n, const int64_t* indptr, const int64_t* indices,
                                 std::vector<int64_t>& fwd, std::vector<int64_t>& inv) {
  int64_t bw = lower_bandwidth(n, indptr, indices);
  if (try_time_major_kronecker(n, indptr, indices, fwd, inv)) return OrderKind::TimeMajor;
  if (bw > 32) {
    reverse_cuthill_mckee(n, indptr, indices, fwd, inv);
    int64_t bw_new = bandwidth_with_fwd(n, indptr, indices, fwd);
    if (bw_new * 2 < bw) return OrderKind::Rcm;
  }
  fwd.clear();
  inv.clear();
  return OrderKind::Amd;
}

// ---------------------------------------------------------------------------
// Approximate minimum degree (Amestoy, Davis & Duff, SIAM J. Matrix Anal. Appl.
// 17(4), 1996): quotient graph of variables and elements, approximate external
// degrees, element absorption (incl. aggressive), supervariables detected by
// hashing, mass elimination, and dense rows (degree > max(16, 10 sqrt n))
// ordered last.
// ---------------------------------------------------------------------------
namespace {

struct Amd {
  int64_t n;
  std::vector<std::vector<int64_t>> elist, vlist, members;
  std::vector<int64_t> nv, deg, esize;
  std::vector<char> state;  // 0 variable, 1 element, 2 dead element, 3 merged / eliminated var
  std::vector<std::vector<int64_t>> group;  // original vars represented by a principal var
  // degree buckets (doubly linked, head insertion)
  std::vector<int64_t> head, next, prev;
  int64_t mindeg = 0;
  std::vector<int64_t> mark, wmark, w;
  int64_t tag = 0;

  void bucket_insert(int64_t i) {
    int64_t d = deg[i];
    prev[i] = -1;
    next[i] = head[d];
    if (head[d] >= 0) prev[head[d]] = i;
    head[d] = i;
    if (d < mindeg) mindeg = d;
  }
  void bucket_remove(int64_t i) {
    int64_t d = deg[i];
    if (prev[i] >= 0)
      next[prev[i]] = next[i];
    else
      head[d] = next[i];
    if (next[i] >= 0) prev[next[i]] = prev[i];
    next[i] = prev[i] = -1;
  }
};

}  // namespace

void amd_order(int64_t n, const int64_t* indptr, const int64_t* indices,
               std::vector<int64_t>& order) {
  order.clear();
  order.reserve(n);
  if (n == 0) return;
  // Symmetric adjacency from the lower triangle (row > col), as faer reads it.
  std::vector<std::vector<int64_t>> adj(n);
  for (int64_t c = 0; c < n; ++c)
    for (int64_t p = indptr[c]; p < indptr[c + 1]; ++p) {
      int64_t r = indices[p];
      if (r > c) {
        adj[c].push_back(r);
        adj[r].push_back(c);
      }
    }
  for (auto& a : adj) {
    std::sort(a.begin(), a.end());
    a.erase(std::unique(a.begin(), a.end()), a.end());
  }
  double dense_thr = std::max(16.0, 10.0 * std::sqrt((double)n));
  std::vector<char> dense(n, 0);
  int64_t ndense = 0;
  for (int64_t i = 0; i < n; ++i)
    if ((double)adj[i].size() > dense_thr) {
      dense[i] = 1;
      ++ndense;
    }

  Amd a;
  a.n = n;
  a.elist.assign(n, {});
  a.vlist.assign(n, {});
  a.members.assign(n, {});
  a.nv.assign(n, 1);
  a.deg.assign(n, 0);
  a.esize.assign(n, 0);
  a.state.assign(n, 0);
  a.group.assign(n, {});
  a.head.assign(n + 1, -1);
  a.next.assign(n, -1);
  a.prev.assign(n, -1);
  a.mark.assign(n, 0);
  a.wmark.assign(n, 0);
  a.w.assign(n, 0);
  a.mindeg = n;
  int64_t nleft = n - ndense;
  for (int64_t i = 0; i < n; ++i) {
    if (dense[i]) {
      a.state[i] = 3;
      a.nv[i] = 0;
      continue;
    }
    for (int64_t j : adj[i])
      if (!dense[j]) a.vlist[i].push_back(j);
    a.deg[i] = (int64_t)a.vlist[i].size();
    a.group[i].push_back(i);
  }
  for (int64_t i = 0; i < n; ++i)  // head insertion: ties go to the highest index
    if (!dense[i]) a.bucket_insert(i);

  int64_t eliminated = 0;
  std::vector<int64_t> lp;
  std::vector<std::pair<uint64_t, int64_t>> hashes;
  while (eliminated < nleft) {
    while (a.mindeg <= n && a.head[a.mindeg] < 0) ++a.mindeg;
    int64_t p = a.head[a.mindeg];
    a.bucket_remove(p);
    int64_t nvp = a.nv[p];
    // --- form L_p = (A_p U (U_{e in E_p} L_e)) \ {p}; absorb E_p into p.
    ++a.tag;
    a.mark[p] = a.tag;
    lp.clear();
    for (int64_t e : a.elist[p]) {
      if (a.state[e] != 1) continue;
      for (int64_t i : a.members[e])
        if (a.state[i] == 0 && a.mark[i] != a.tag) {
          a.mark[i] = a.tag;
          lp.push_back(i);
        }
      a.state[e] = 2;
      a.members[e].clear();
      a.members[e].shrink_to_fit();
    }
    for (int64_t i : a.vlist[p])
      if (a.state[i] == 0 && a.mark[i] != a.tag) {
        a.mark[i] = a.tag;
        lp.push_back(i);
      }
    a.elist[p].clear();
    a.elist[p].shrink_to_fit();
    a.vlist[p].clear();
    a.vlist[p].shrink_to_fit();
    a.state[p] = 1;
    int64_t esz = 0;
    for (int64_t i : lp) esz += a.nv[i];
    // eliminate the pivot supervariable
    for (int64_t v : a.group[p]) order.push_back(v);
    eliminated += nvp;
    a.group[p].clear();
    // --- w(e) = |L_e \ L_p| for elements adjacent to L_p.
    ++a.tag;
    int64_t wtag = a.tag;
    for (int64_t i : lp)
      for (int64_t e : a.elist[i]) {
        if (a.state[e] != 1 || e == p) continue;
        if (a.wmark[e] != wtag) {
          a.wmark[e] = wtag;
          a.w[e] = a.esize[e] - a.nv[i];
        } else {
          a.w[e] -= a.nv[i];
        }
      }
    // mark L_p membership for pruning variable lists
    ++a.tag;
    int64_t ltag = a.tag;
    for (int64_t i : lp) a.mark[i] = ltag;
    // --- update each i in L_p
    int64_t remaining = (n - ndense) - eliminated;
    std::vector<int64_t> keep;
    keep.reserve(lp.size());
    hashes.clear();
    for (int64_t i : lp) {
      a.bucket_remove(i);
      // prune element list: drop dead elements, absorb elements with L_e subset L_p
      auto& el = a.elist[i];
      int64_t ext = 0;
      size_t k = 0;
      uint64_t h = 0;
      for (size_t q = 0; q < el.size(); ++q) {
        int64_t e = el[q];
        if (a.state[e] != 1 || e == p) continue;
        if (a.wmark[e] == wtag && a.w[e] <= 0) {
          // aggressive absorption: L_e is contained in L_p
          a.state[e] = 2;
          a.members[e].clear();
          a.members[e].shrink_to_fit();
          continue;
        }
        ext += (a.wmark[e] == wtag) ? a.w[e] : a.esize[e];
        el[k++] = e;
        h += (uint64_t)e;
      }
      el.resize(k);
      el.push_back(p);
      h += (uint64_t)p;
      // prune variable list: drop non-principal variables and those now in L_p
      auto& vl = a.vlist[i];
      size_t kv = 0;
      int64_t vdeg = 0;
      for (size_t q = 0; q < vl.size(); ++q) {
        int64_t j = vl[q];
        if (a.state[j] != 0 || a.mark[j] == ltag) continue;
        vl[kv++] = j;
        vdeg += a.nv[j];
        h += (uint64_t)j;
      }
      vl.resize(kv);
      // mass elimination: i is adjacent only to element p
      if (el.size() == 1 && vl.empty()) {
        for (int64_t v : a.group[i]) order.push_back(v);
        eliminated += a.nv[i];
        esz -= a.nv[i];
        a.group[i].clear();
        a.nv[i] = 0;
        a.state[i] = 3;
        continue;
      }
      int64_t lp_ext = esz - a.nv[i];
      int64_t d = vdeg + lp_ext + ext;
      d = std::min(d, a.deg[i] + lp_ext);
      a.deg[i] = d;  // provisional; clamped after supervariable merging
      keep.push_back(i);
      hashes.push_back({h, i});
    }
    // --- supervariable detection among the survivors (equal hash -> compare sets)
    std::sort(hashes.begin(), hashes.end());
    for (size_t s = 0; s < hashes.size();) {
      size_t e = s;
      while (e < hashes.size() && hashes[e].first == hashes[s].first) ++e;
      for (size_t x = s; x < e; ++x) {
        int64_t i = hashes[x].second;
        if (a.state[i] != 0) continue;
        // mark i's adjacency
        ++a.tag;
        for (int64_t q : a.elist[i]) a.mark[q] = a.tag;
        for (int64_t q : a.vlist[i]) a.mark[q] = a.tag;
        for (size_t y = x + 1; y < e; ++y) {
          int64_t j = hashes[y].second;
          if (a.state[j] != 0) continue;
          if (a.elist[j].size() != a.elist[i].size() || a.vlist[j].size() != a.vlist[i].size())
            continue;
          bool same = true;
          for (int64_t q : a.elist[j])
            if (a.mark[q] != a.tag) { same = false; break; }
          if (same)
            for (int64_t q : a.vlist[j])
              if (a.mark[q] != a.tag) { same = false; break; }
          if (!same) continue;
          // merge j into i: j's weight is no longer external to i
          a.deg[i] -= a.nv[j];
          a.nv[i] += a.nv[j];
          a.nv[j] = 0;
          a.state[j] = 3;
          for (int64_t v : a.group[j]) a.group[i].push_back(v);
          a.group[j].clear();
          a.elist[j].clear();
          a.vlist[j].clear();
        }
      }
      s = e;
    }
    // --- finalise element p and reinsert survivors with clamped degrees
    auto& mem = a.members[p];
    mem.clear();
    for (int64_t i : keep)
      if (a.state[i] == 0) mem.push_back(i);
    a.esize[p] = 0;
    for (int64_t i : mem) a.esize[p] += a.nv[i];
    remaining = (n - ndense) - eliminated;
    for (int64_t i : mem) {
      int64_t d = a.deg[i];
      int64_t cap = remaining - a.nv[i];
      if (d > cap) d = cap;
      if (d < 0) d = 0;
      a.deg[i] = d;
      a.bucket_insert(i);
    }
  }
  for (int64_t i = 0; i < n; ++i)
    if (dense[i]) order.push_back(i);
}

// ---------------------------------------------------------------------------
// Nested dissection with level-structure separators (A. George, "Nested dissection
// of a regular finite element mesh", SIAM J. Numer. Anal. 10(2), 1973; the automatic
// form of George & Liu, "Computer Solution of Large Sparse Positive Definite
// Systems", 1981, ch. 8). A connected subgraph larger than `leaf` is split by a level
// set of a breadth-first search from a pseudo-peripheral vertex: the middle level,
// thinned to the vertices adjacent to the next level, is numbered last and the rest
// is dissected recursively (each connected component on its own); a subgraph of at
// most `leaf` vertices keeps the order of its vertices in the work list, so that it
// becomes a few dense fronts. The elimination tree is short and wide, which the
// level-synchronous factorisation turns into fewer dependent launches than the
// deep trees of minimum degree.
// ---------------------------------------------------------------------------
void nested_dissection_order(int64_t n, const int64_t* indptr, const int64_t* indices, int64_t leaf,
                             std::vector<int64_t>& order) {
  order.assign(n, -1);
  if (n == 0) return;
  leaf = std::max<int64_t>(leaf, 1);
  // symmetric adjacency without the diagonal, from the lower triangle (row > col)
  std::vector<int64_t> xadj(n + 1, 0);
  for (int64_t c = 0; c < n; ++c)
    for (int64_t p = indptr[c]; p < indptr[c + 1]; ++p)
      if (indices[p] > c) {
        xadj[c + 1]++;
        xadj[indices[p] + 1]++;
      }
  for (int64_t i = 0; i < n; ++i) xadj[i + 1] += xadj[i];
  std::vector<int64_t> adj(xadj[n]);
  {
    std::vector<int64_t> fill(xadj.begin(), xadj.end() - 1);
    for (int64_t c = 0; c < n; ++c)
      for (int64_t p = indptr[c]; p < indptr[c + 1]; ++p) {
        const int64_t r = indices[p];
        if (r > c) {
          adj[fill[c]++] = r;
          adj[fill[r]++] = c;
        }
      }
  }
  // stamp[v] == cur: v belongs to the subgraph being split; lev[v]: its BFS level
  // in the current search (valid where seen[v] == search id)
  std::vector<int64_t> stamp(n, -1), seen(n, -1), lev(n, 0), placed(n, -1);
  int64_t cur = 0, search = 0;
  // dense rows (degree > max(16, 10 sqrt n), as the minimum-degree ordering treats
  // them) are numbered last and left out of the graph being dissected: one such row
  // (an intercept) would otherwise put every vertex within two levels of every other
  const int64_t dense_deg = std::max<int64_t>(16, (int64_t)(10.0 * std::sqrt((double)n)));
  std::vector<int64_t> buf, queue(n), tmp, dense;
  for (int64_t v = 0; v < n; ++v) {
    if (xadj[v + 1] - xadj[v] > dense_deg) {
      dense.push_back(v);
      stamp[v] = -3;
    } else {
      buf.push_back(v);
    }
  }
  const int64_t nsparse = (int64_t)buf.size();
  for (int64_t i = 0; i < (int64_t)dense.size(); ++i) order[nsparse + i] = dense[i];
  buf.resize(n);
  // BFS from src over the stamped vertices into queue; returns the number visited and
  // fills lstart with each level's first queue index (plus the end)
  std::vector<int64_t> lstart;
  auto bfs = [&](int64_t src) {
    const int64_t id = search++;
    int64_t head = 0, tail = 0;
    queue[tail++] = src;
    seen[src] = id;
    lev[src] = 0;
    lstart.assign(1, 0);
    while (head < tail) {
      const int64_t u = queue[head++];
      for (int64_t p = xadj[u]; p < xadj[u + 1]; ++p) {
        const int64_t v = adj[p];
        if (stamp[v] != cur || seen[v] == id) continue;
        seen[v] = id;
        lev[v] = lev[u] + 1;
        queue[tail++] = v;
      }
    }
    // level boundaries (the queue is in nondecreasing level order)
    for (int64_t q = 1; q < tail; ++q)
      if (lev[queue[q]] != lev[queue[q - 1]]) lstart.push_back(q);
    lstart.push_back(tail);
    return tail;
  };
  auto degree_in = [&](int64_t v) {
    int64_t d = 0;
    for (int64_t p = xadj[v]; p < xadj[v + 1]; ++p) d += stamp[adj[p]] == cur;
    return d;
  };
  // work list: segment [b, e) of buf, numbered into order[o, o + e - b)
  struct Task {
    int64_t b, e, o;
  };
  std::vector<Task> work{{0, nsparse, 0}};
  while (!work.empty()) {
    const Task t = work.back();
    work.pop_back();
    const int64_t sz = t.e - t.b;
    if (sz <= leaf) {
      for (int64_t i = 0; i < sz; ++i) order[t.o + i] = buf[t.b + i];
      continue;
    }
    cur++;
    for (int64_t i = t.b; i < t.e; ++i) stamp[buf[i]] = cur;
    // connected components, each in BFS order
    const int64_t reached = bfs(buf[t.b]);
    if (reached < sz) {
      tmp.assign(queue.begin(), queue.begin() + reached);
      std::vector<int64_t> comp_end{reached};
      for (int64_t q = 0; q < reached; ++q) placed[queue[q]] = cur;
      for (int64_t i = t.b; i < t.e; ++i) {
        const int64_t v = buf[i];
        if (placed[v] == cur) continue;
        const int64_t k = bfs(v);
        for (int64_t q = 0; q < k; ++q) placed[queue[q]] = cur;
        tmp.insert(tmp.end(), queue.begin(), queue.begin() + k);
        comp_end.push_back((int64_t)tmp.size());
      }
      std::copy(tmp.begin(), tmp.end(), buf.begin() + t.b);
      int64_t c0 = 0;
      for (int64_t ce : comp_end) {
        work.push_back({t.b + c0, t.b + ce, t.o + c0});
        c0 = ce;
      }
      continue;
    }
    // pseudo-peripheral vertex: restart from a minimum-degree vertex of the last level
    // while the level structure gets deeper
    int64_t root = buf[t.b], nlev = (int64_t)lstart.size() - 1;
    for (int it = 0; it < 5; ++it) {
      int64_t cand = -1, cd = 0;
      for (int64_t q = lstart[nlev - 1]; q < lstart[nlev]; ++q) {
        const int64_t v = queue[q], d = degree_in(v);
        if (cand < 0 || d < cd) cand = v, cd = d;
      }
      bfs(cand);
      const int64_t nl2 = (int64_t)lstart.size() - 1;
      if (nl2 <= nlev) break;
      root = cand;
      nlev = nl2;
    }
    bfs(root);
    nlev = (int64_t)lstart.size() - 1;
    if (nlev < 3) {
      // no separating level: keep the list order
      for (int64_t i = 0; i < sz; ++i) order[t.o + i] = buf[t.b + i];
      continue;
    }
    const int64_t mid = nlev / 2;
    const int64_t id = search - 1;
    tmp.clear();
    std::vector<int64_t> sep;
    for (int64_t q = lstart[mid]; q < lstart[mid + 1]; ++q) {
      const int64_t v = queue[q];
      bool next = false;
      for (int64_t p = xadj[v]; p < xadj[v + 1] && !next; ++p) {
        const int64_t u = adj[p];
        next = stamp[u] == cur && seen[u] == id && lev[u] == mid + 1;
      }
      if (next) sep.push_back(v);
    }
    // the rest in list order, then the separator (numbered last)
    for (int64_t v : sep) stamp[v] = -2;  // out of every later subgraph
    for (int64_t i = t.b; i < t.e; ++i)
      if (stamp[buf[i]] != -2) tmp.push_back(buf[i]);
    const int64_t nr = (int64_t)tmp.size();
    std::copy(tmp.begin(), tmp.end(), buf.begin() + t.b);
    for (size_t i = 0; i < sep.size(); ++i) {
      buf[t.b + nr + i] = sep[i];
      order[t.o + nr + i] = sep[i];
    }
    work.push_back({t.b, t.b + nr, t.o});
  }
}

// ---------------------------------------------------------------------------
// Pattern helpers
// ---------------------------------------------------------------------------
std::string validate_pattern(int64_t n, const int64_t* indptr, const int64_t* indices) {
  if (n < 0) return "matrix dimension must be >= 0";
  if (!indptr) return "indptr is null";
  if (indptr[0] != 0) return "indptr[0] must be 0";
  for (int64_t c = 0; c < n; ++c) {
    if (indptr[c + 1] < indptr[c]) return "indptr must be non-decreasing";
    for (int64_t p = indptr[c]; p < indptr[c + 1]; ++p) {
      int64_t r = indices[p];
      if (r < 0 || r >= n) return "row index out of bounds";
      if (p > indptr[c] && indices[p - 1] >= r) return "row indices must be strictly increasing";
    }
  }
  return "";
}

uint64_t hash_pattern(int64_t n, const int64_t* indptr, const int64_t* indices) {
  // FNV-1a over the raw integers (cache key; equality is still checked exactly).
  uint64_t h = 1469598103934665603ull;
  auto mix = [&](uint64_t v) {
    for (int k = 0; k < 8; ++k) {
      h ^= (v >> (8 * k)) & 0xff;
      h *= 1099511628211ull;
    }
  };
  mix((uint64_t)n);
  for (int64_t c = 0; c <= n; ++c) mix((uint64_t)indptr[c]);
  for (int64_t p = 0; p < indptr[n]; ++p) mix((uint64_t)indices[p]);
  return h;
}

// The level schedule runs each level as one launch sequence over all matrices of a
// batch; a level whose fronts have more than 64 rows takes one dependent 64-column
// panel step per 64 columns of its widest such front. Cost: 70 us per dependent step
// (a factorisation and a selected-inversion step of a 30-matrix batch at n = 20k on
// one MI355X) plus the flops at 0.33 TFLOP/s per matrix (10 TFLOP/s over the batch).
double level_schedule_cost(const Symbolic& S) {
  constexpr double kStepSeconds = 70e-6, kFlopRate = 3.3e11;
  double steps = 0.0;
  for (int l = 0; l < S.nlevels; ++l) {
    int64_t p = 1;
    for (int64_t q = S.level_ptr[l]; q < S.level_ptr[l + 1]; ++q) {
      const int64_t s = S.level_sn[q];
      const int64_t w = S.sn_start[s + 1] - S.sn_start[s], m = w + (S.sn_rows_ptr[s + 1] - S.sn_rows_ptr[s]);
      if (m > 64) p = std::max<int64_t>(p, (w + 63) / 64);
    }
    steps += (double)p;
  }
  return steps * kStepSeconds + (S.flops_ldlt + S.flops_selinv) / kFlopRate;
}

// ---------------------------------------------------------------------------
// Full analysis
// ---------------------------------------------------------------------------
Symbolic analyze(int64_t n, const int64_t* indptr, const int64_t* indices, OrderRequest req,
                 const int64_t* given_perm) {
  std::string err = validate_pattern(n, indptr, indices);
  if (!err.empty()) throw std::runtime_error(err);
  if (n == 0) throw std::runtime_error("empty or invalid matrix: n must be > 0");
  if (req == OrderRequest::Auto) {
    Symbolic a = analyze(n, indptr, indices, OrderRequest::Chooser, nullptr);
    Symbolic b = analyze(n, indptr, indices, OrderRequest::NestedDissection, nullptr);
    return level_schedule_cost(b) < level_schedule_cost(a) ? b : a;
  }
  Symbolic S;
  S.n = n;
  S.nnz_full = indptr[n];
  S.indptr.assign(indptr, indptr + n + 1);
  S.indices.assign(indices, indices + indptr[n]);
  S.pattern_hash = hash_pattern(n, indptr, indices);

  // ---- 1. ordering: perm0[k] = original index at position k
  std::vector<int64_t> perm0(n);
  if (req == OrderRequest::Given) {
    if (!given_perm) throw std::runtime_error("given permutation is null");
    std::vector<char> seen(n, 0);
    for (int64_t k = 0; k < n; ++k) {
      int64_t v = given_perm[k];
      if (v < 0 || v >= n || seen[v]) throw std::runtime_error("invalid permutation");
      seen[v] = 1;
      perm0[k] = v;
    }
    S.order_kind = OrderKind::Given;
  } else if (req == OrderRequest::Natural) {
    std::iota(perm0.begin(), perm0.end(), 0);
    S.order_kind = OrderKind::Natural;
  } else if (req == OrderRequest::NestedDissection) {
    nested_dissection_order(n, indptr, indices, kNdLeaf, perm0);
    S.order_kind = OrderKind::NestedDissection;
  } else {
    std::vector<int64_t> fwd, inv;
    OrderKind k = OrderKind::Amd;
    if (req == OrderRequest::Chooser) k = choose_symmetric_order(n, indptr, indices, fwd, inv);
    S.order_kind = k;
    if (k == OrderKind::Amd) {
      amd_order(n, indptr, indices, perm0);
    } else {
      // The chooser's permutation maps old -> new (fwd[old] = new); the factor
      // position w holds original index inv[w].
      perm0 = inv;
    }
  }
  std::vector<int64_t> iperm0(n);
  for (int64_t k = 0; k < n; ++k) iperm0[perm0[k]] = k;

  // ---- 2. symmetric adjacency (lower triangle, row > col) in permuted numbering
  std::vector<int64_t> adeg(n, 0);
  for (int64_t c = 0; c < n; ++c)
    for (int64_t p = indptr[c]; p < indptr[c + 1]; ++p) {
      int64_t r = indices[p];
      if (r > c) {
        adeg[iperm0[r]]++;
        adeg[iperm0[c]]++;
      }
    }
  std::vector<int64_t> aptr(n + 1, 0);
  for (int64_t i = 0; i < n; ++i) aptr[i + 1] = aptr[i] + adeg[i];
  std::vector<int64_t> aidx(aptr[n]);
  {
    std::vector<int64_t> fill(aptr.begin(), aptr.end() - 1);
    for (int64_t c = 0; c < n; ++c)
      for (int64_t p = indptr[c]; p < indptr[c + 1]; ++p) {
        int64_t r = indices[p];
        if (r > c) {
          int64_t a = iperm0[r], b = iperm0[c];
          aidx[fill[a]++] = b;
          aidx[fill[b]++] = a;
        }
      }
    for (int64_t i = 0; i < n; ++i) {
      std::sort(aidx.begin() + aptr[i], aidx.begin() + aptr[i + 1]);
    }
  }
  // duplicates may occur if both (r,c) and (c,r) are lower... they cannot: only r>c kept.

  // ---- 3. elimination tree (Liu, path compression) in perm0 numbering
  std::vector<int64_t> par0(n, -1), anc(n, -1);
  for (int64_t k = 0; k < n; ++k) {
    for (int64_t q = aptr[k]; q < aptr[k + 1]; ++q) {
      int64_t i = aidx[q];
      if (i >= k) continue;
      while (i != -1 && i != k) {
        int64_t nx = anc[i];
        anc[i] = k;
        if (nx == -1) {
          par0[i] = k;
          break;
        }
        i = nx;
      }
    }
  }
  // ---- 4. postorder
  std::vector<int64_t> post(n);
  {
    std::vector<int64_t> chead(n, -1), cnext(n, -1);
    for (int64_t j = n - 1; j >= 0; --j)
      if (par0[j] >= 0) {
        cnext[j] = chead[par0[j]];
        chead[par0[j]] = j;
      }
    int64_t k = 0;
    std::vector<int64_t> stack;
    for (int64_t r = 0; r < n; ++r) {
      if (par0[r] != -1) continue;
      stack.push_back(r);
      while (!stack.empty()) {
        int64_t top = stack.back();
        int64_t c = chead[top];
        if (c == -1) {
          stack.pop_back();
          post[k++] = top;
        } else {
          chead[top] = cnext[c];
          stack.push_back(c);
        }
      }
    }
  }
  // final factor order: position k holds perm0[post[k]]
  S.perm.resize(n);
  S.iperm.resize(n);
  std::vector<int64_t> newpos(n);  // perm0 index -> final index
  for (int64_t k = 0; k < n; ++k) {
    S.perm[k] = perm0[post[k]];
    newpos[post[k]] = k;
  }
  for (int64_t k = 0; k < n; ++k) S.iperm[S.perm[k]] = k;
  S.parent.assign(n, -1);
  for (int64_t j = 0; j < n; ++j)
    if (par0[j] >= 0) S.parent[newpos[j]] = newpos[par0[j]];
  // adjacency in final numbering
  std::vector<int64_t> fptr(n + 1, 0), fidx(aidx.size());
  for (int64_t k = 0; k < n; ++k) fptr[k + 1] = fptr[k] + (aptr[post[k] + 1] - aptr[post[k]]);
  for (int64_t k = 0; k < n; ++k) {
    int64_t o = post[k], q0 = fptr[k];
    for (int64_t q = aptr[o]; q < aptr[o + 1]; ++q) fidx[q0++] = newpos[aidx[q]];
    std::sort(fidx.begin() + fptr[k], fidx.begin() + fptr[k + 1]);
  }
  aptr.clear();
  aidx.clear();

  // ---- 5. column counts by row-subtree traversal
  S.colcount.assign(n, 1);
  {
    std::vector<int64_t> mk(n, -1);
    for (int64_t i = 0; i < n; ++i) {
      mk[i] = i;
      for (int64_t q = fptr[i]; q < fptr[i + 1]; ++q) {
        int64_t j = fidx[q];
        if (j >= i) break;
        while (j != -1 && mk[j] != i) {
          S.colcount[j]++;
          mk[j] = i;
          j = S.parent[j];
        }
      }
    }
  }
  // ---- 6. fundamental supernodes
  std::vector<int64_t> nchild(n, 0);
  for (int64_t j = 0; j < n; ++j)
    if (S.parent[j] >= 0) nchild[S.parent[j]]++;
  std::vector<int64_t> fstart;
  for (int64_t j = 0; j < n; ++j) {
    bool cont = j > 0 && S.parent[j - 1] == j && S.colcount[j - 1] == S.colcount[j] + 1 &&
                nchild[j] == 1;
    if (!cont) fstart.push_back(j);
  }
  int64_t nf = (int64_t)fstart.size();
  fstart.push_back(n);
  std::vector<int64_t> col2f(n);
  for (int64_t s = 0; s < nf; ++s)
    for (int64_t j = fstart[s]; j < fstart[s + 1]; ++j) col2f[j] = s;
  std::vector<int64_t> fpar(nf, -1);
  for (int64_t s = 0; s < nf; ++s) {
    int64_t last = fstart[s + 1] - 1;
    if (S.parent[last] >= 0) fpar[s] = col2f[S.parent[last]];
  }
  // ---- 7. relaxed amalgamation (merge a supernode into its parent when it is
  // the parent's immediately preceding child and the added explicit zeros are few)
  std::vector<int64_t> fw(nf), fr(nf);
  std::vector<double> fzeros(nf, 0.0);
  for (int64_t s = 0; s < nf; ++s) {
    fw[s] = fstart[s + 1] - fstart[s];
    fr[s] = S.colcount[fstart[s]] - fw[s];
  }
  std::vector<int64_t> merged_into(nf, -1);  // s merged into s+1's group
  std::vector<int64_t> grp_w = fw, grp_r = fr;
  std::vector<double> grp_z = fzeros;
  for (int64_t s = nf - 2; s >= 0; --s) {
    if (fpar[s] != s + 1) continue;
    int64_t p = s + 1;  // representative of p's group is itself (groups extend upward)
    int64_t w1 = grp_w[s], r1 = grp_r[s], w2 = grp_w[p], r2 = grp_r[p];
    int64_t wn = w1 + w2;
    double stored_s = (double)w1 * (w1 + 1) / 2 + (double)w1 * r1;
    double stored_p = (double)w2 * (w2 + 1) / 2 + (double)w2 * r2;
    double stored_n = (double)wn * (wn + 1) / 2 + (double)wn * r2;
    double zn = stored_n - stored_s - stored_p + grp_z[s] + grp_z[p];
    bool merge;
    if (wn <= 4)
      merge = true;
    else if (wn <= 16)
      merge = zn < 0.8 * stored_n;
    else if (wn <= 48)
      merge = zn < 0.1 * stored_n;
    else
      merge = zn < 0.05 * stored_n;
    if (merge) {
      merged_into[s] = p;
      grp_w[s] = wn;  // s now heads the merged group [start_s, end_of_p's group)
      grp_r[s] = r2;
      grp_z[s] = zn;
    }
  }
  // group heads: s is a head if it was not merged into its successor... walk
  // from the top: a group is a maximal run s, s+1, ... with merged_into[s] == s+1.
  std::vector<int64_t> sstart;
  for (int64_t s = 0; s < nf; ++s)
    if (s == 0 || merged_into[s - 1] != s) sstart.push_back(fstart[s]);
  S.nsuper = (int64_t)sstart.size();
  sstart.push_back(n);
  S.sn_start = sstart;
  int64_t ns = S.nsuper;
  std::vector<int64_t> col2s(n);
  for (int64_t s = 0; s < ns; ++s)
    for (int64_t j = S.sn_start[s]; j < S.sn_start[s + 1]; ++j) col2s[j] = s;
  S.sn_parent.assign(ns, -1);
  for (int64_t s = 0; s < ns; ++s) {
    int64_t last = S.sn_start[s + 1] - 1;
    if (S.parent[last] >= 0) S.sn_parent[s] = col2s[S.parent[last]];
  }
  // children lists
  S.child_ptr.assign(ns + 1, 0);
  for (int64_t s = 0; s < ns; ++s)
    if (S.sn_parent[s] >= 0) S.child_ptr[S.sn_parent[s] + 1]++;
  for (int64_t s = 0; s < ns; ++s) S.child_ptr[s + 1] += S.child_ptr[s];
  S.child_idx.assign(S.child_ptr[ns], 0);
  {
    std::vector<int64_t> fill(S.child_ptr.begin(), S.child_ptr.end() - 1);
    for (int64_t s = 0; s < ns; ++s)
      if (S.sn_parent[s] >= 0) S.child_idx[fill[S.sn_parent[s]]++] = s;
  }
  // ---- 8. row structures R_s (bottom-up union)
  S.sn_rows_ptr.assign(ns + 1, 0);
  std::vector<std::vector<int64_t>> rows(ns);
  {
    std::vector<int64_t> mk(n, -1);
    for (int64_t s = 0; s < ns; ++s) {
      int64_t a = S.sn_start[s], b = S.sn_start[s + 1];
      auto& R = rows[s];
      for (int64_t j = a; j < b; ++j)
        for (int64_t q = fptr[j]; q < fptr[j + 1]; ++q) {
          int64_t i = fidx[q];
          if (i >= b && mk[i] != s) {
            mk[i] = s;
            R.push_back(i);
          }
        }
      for (int64_t q = S.child_ptr[s]; q < S.child_ptr[s + 1]; ++q) {
        int64_t c = S.child_idx[q];
        for (int64_t i : rows[c])
          if (i >= b && mk[i] != s) {
            mk[i] = s;
            R.push_back(i);
          }
      }
      std::sort(R.begin(), R.end());
    }
  }
  for (int64_t s = 0; s < ns; ++s) S.sn_rows_ptr[s + 1] = S.sn_rows_ptr[s] + rows[s].size();
  S.sn_rows.resize(S.sn_rows_ptr[ns]);
  for (int64_t s = 0; s < ns; ++s)
    std::copy(rows[s].begin(), rows[s].end(), S.sn_rows.begin() + S.sn_rows_ptr[s]);
  // consistency: the parent supernode must own the first row of R_s
  for (int64_t s = 0; s < ns; ++s) {
    if (rows[s].empty()) {
      if (S.sn_parent[s] != -1) throw std::runtime_error("symbolic: parent without rows");
    } else if (col2s[rows[s][0]] != S.sn_parent[s]) {
      throw std::runtime_error("symbolic: supernode parent mismatch");
    }
  }
  // ---- 9. storage offsets, rel maps
  S.l_off.assign(ns + 1, 0);
  S.u_off.assign(ns + 1, 0);
  S.v_off.assign(ns + 1, 0);
  for (int64_t s = 0; s < ns; ++s) {
    int64_t w = S.sn_start[s + 1] - S.sn_start[s];
    int64_t r = (int64_t)rows[s].size();
    int64_t m = w + r;
    S.l_off[s + 1] = S.l_off[s] + m * w;
    S.u_off[s + 1] = S.u_off[s] + r * r;
    S.v_off[s + 1] = S.v_off[s] + r;
    S.max_front = std::max(S.max_front, m);
  }
  S.l_size = S.l_off[ns];
  S.u_size = S.u_off[ns];
  S.v_size = S.v_off[ns];
  S.rel.assign(S.sn_rows.size(), 0);
  {
    std::vector<int64_t> pos(n, -1);
    for (int64_t p = 0; p < ns; ++p) {
      int64_t a = S.sn_start[p], b = S.sn_start[p + 1], w = b - a;
      for (int64_t j = a; j < b; ++j) pos[j] = j - a;
      for (size_t i = 0; i < rows[p].size(); ++i) pos[rows[p][i]] = w + (int64_t)i;
      for (int64_t q = S.child_ptr[p]; q < S.child_ptr[p + 1]; ++q) {
        int64_t c = S.child_idx[q];
        for (int64_t i = S.sn_rows_ptr[c]; i < S.sn_rows_ptr[c + 1]; ++i) {
          int64_t pp = pos[S.sn_rows[i]];
          if (pp < 0) throw std::runtime_error("symbolic: child row not in parent front");
          S.rel[i] = (int32_t)pp;
        }
      }
      for (int64_t j = a; j < b; ++j) pos[j] = -1;
      for (int64_t i : rows[p]) pos[i] = -1;
    }
  }
  // ---- 10. level schedule (height)
  std::vector<int64_t> lev(ns, 0);
  int64_t maxlev = 0;
  for (int64_t s = 0; s < ns; ++s) {
    for (int64_t q = S.child_ptr[s]; q < S.child_ptr[s + 1]; ++q)
      lev[s] = std::max(lev[s], lev[S.child_idx[q]] + 1);
    maxlev = std::max(maxlev, lev[s]);
  }
  S.nlevels = (int)(maxlev + 1);
  S.level_ptr.assign(S.nlevels + 1, 0);
  for (int64_t s = 0; s < ns; ++s) S.level_ptr[lev[s] + 1]++;
  for (int l = 0; l < S.nlevels; ++l) S.level_ptr[l + 1] += S.level_ptr[l];
  S.level_sn.assign(ns, 0);
  {
    std::vector<int64_t> fill(S.level_ptr.begin(), S.level_ptr.end() - 1);
    // within a level, largest fronts first so long blocks start early
    std::vector<int64_t> ord(ns);
    std::iota(ord.begin(), ord.end(), 0);
    auto msz = [&](int64_t s) {
      return (S.sn_start[s + 1] - S.sn_start[s]) + (S.sn_rows_ptr[s + 1] - S.sn_rows_ptr[s]);
    };
    std::stable_sort(ord.begin(), ord.end(), [&](int64_t x, int64_t y) { return msz(x) > msz(y); });
    for (int64_t s : ord) S.level_sn[fill[lev[s]]++] = s;
  }
  // ---- 11. CSC entry -> L slot maps, diagonal slots. slot_full covers both
  // triangles (the symmetric position in the L layout, -1 if outside L);
  // amap keeps the entries the numeric phase reads (row >= col, original numbering).
  S.amap.assign(S.nnz_full, -1);
  S.slot_full.assign(S.nnz_full, -1);
  S.diag_slot.assign(n, 0);
  {
    std::vector<int64_t> ecol(S.nnz_full);
    for (int64_t c = 0; c < n; ++c)
      for (int64_t p = indptr[c]; p < indptr[c + 1]; ++p) ecol[p] = c;
    std::vector<int64_t> cnt(n + 1, 0);
    for (int64_t p = 0; p < S.nnz_full; ++p) {
      int64_t fi = S.iperm[indices[p]], fj = S.iperm[ecol[p]];
      cnt[std::min(fi, fj) + 1]++;
      if (indices[p] >= ecol[p]) S.nnz_q_lower++;
    }
    for (int64_t j = 0; j < n; ++j) cnt[j + 1] += cnt[j];
    std::vector<int64_t> ent(cnt[n]), fill(cnt.begin(), cnt.end() - 1);
    for (int64_t p = 0; p < S.nnz_full; ++p) {
      int64_t fi = S.iperm[indices[p]], fj = S.iperm[ecol[p]];
      ent[fill[std::min(fi, fj)]++] = p;
    }
    std::vector<int64_t> pos(n, -1);
    for (int64_t s = 0; s < ns; ++s) {
      int64_t a = S.sn_start[s], b = S.sn_start[s + 1], w = b - a;
      int64_t m = w + (int64_t)rows[s].size();
      for (int64_t j = a; j < b; ++j) pos[j] = j - a;
      for (size_t i = 0; i < rows[s].size(); ++i) pos[rows[s][i]] = w + (int64_t)i;
      for (int64_t j = a; j < b; ++j) {
        S.diag_slot[j] = S.l_off[s] + (j - a) * m + (j - a);
        for (int64_t q = cnt[j]; q < cnt[j + 1]; ++q) {
          int64_t p = ent[q];
          int64_t fi = S.iperm[indices[p]], fj = S.iperm[ecol[p]];
          int64_t rp = pos[std::max(fi, fj)];
          bool lower = indices[p] >= ecol[p];
          if (rp < 0) {
            if (lower) throw std::runtime_error("symbolic: entry outside L structure");
            continue;
          }
          int64_t slot = S.l_off[s] + (j - a) * m + rp;
          S.slot_full[p] = slot;
          if (lower) S.amap[p] = slot;
        }
      }
      for (int64_t j = a; j < b; ++j) pos[j] = -1;
      for (int64_t i : rows[s]) pos[i] = -1;
    }
  }
  // ---- 12. algorithmic counts
  for (int64_t j = 0; j < n; ++j) {
    double r = (double)(S.colcount[j] - 1);
    S.nnz_l += S.colcount[j];
    S.sum_r2 += r * r;
    S.flops_ldlt += r * (r + 3.0);
    S.flops_selinv += 2.0 * r * (r + 1.0);
    S.flops_solve += 4.0 * r;
  }
  S.flops_selinv += (double)n;
  S.flops_solve += (double)n;
  return S;
}

}  // namespace gmrf
