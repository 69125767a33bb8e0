// Reverse Cuthill-McKee ordering of an undirected graph in CSR form
// (internal).  Rule: start from the lowest-(degree, index) unvisited vertex,
// improved by a George-Liu pseudo-peripheral search (at most 16 sweeps, each
// from the lowest-(degree, index) vertex of the last BFS level); neighbours are
// queued by (degree, index); components are ordered one after another; the
// Cuthill-McKee order is reversed.  Neighbour lists must be sorted by index.
#pragma once

#include <algorithm>
#include <vector>

namespace op2h {

inline int rcm_bfs_levels(int root, const std::vector<int> &off,
                          const std::vector<int> &adj,
                          const std::vector<char> &done, std::vector<int> &level,
                          std::vector<int> &last) {
  std::vector<int> frontier{root}, next, touched{root};
  level[root] = 0;
  int depth = 0;
  last = frontier;
  while (!frontier.empty()) {
    next.clear();
    for (int v : frontier)
      for (int k = off[v]; k < off[v + 1]; ++k) {
        int w = adj[k];
        if (!done[w] && level[w] < 0) {
          level[w] = depth + 1;
          next.push_back(w);
          touched.push_back(w);
        }
      }
    if (next.empty()) break;
    last = next;
    frontier.swap(next);
    ++depth;
  }
  for (int v : touched) level[v] = -1;
  return depth;
}

// Returns new_of_old.
inline std::vector<int> rcm_order(int n, const std::vector<int> &off,
                                  const std::vector<int> &adj) {
  auto deg = [&](int v) { return off[v + 1] - off[v]; };
  auto by_deg = [&](int a, int b) {
    return deg(a) != deg(b) ? deg(a) < deg(b) : a < b;
  };
  std::vector<char> done(n, 0);
  std::vector<int> cm, level(n, -1), last, last2, nb;
  cm.reserve(n);
  int scan = 0;  // vertices before `scan` are all visited
  std::vector<int> order_by_deg(n);
  for (int v = 0; v < n; ++v) order_by_deg[v] = v;
  std::stable_sort(order_by_deg.begin(), order_by_deg.end(), by_deg);
  while ((int)cm.size() < n) {
    while (done[order_by_deg[scan]]) ++scan;
    int root = order_by_deg[scan];
    int ecc = rcm_bfs_levels(root, off, adj, done, level, last);
    for (int it = 0; it < 16; ++it) {
      int cand = *std::min_element(last.begin(), last.end(), by_deg);
      int ecc2 = rcm_bfs_levels(cand, off, adj, done, level, last2);
      if (ecc2 <= ecc) break;
      root = cand;
      ecc = ecc2;
      last.swap(last2);
    }
    size_t head = cm.size();
    cm.push_back(root);
    done[root] = 1;
    while (head < cm.size()) {
      int v = cm[head++];
      nb.clear();
      for (int k = off[v]; k < off[v + 1]; ++k)
        if (!done[adj[k]]) nb.push_back(adj[k]);
      std::sort(nb.begin(), nb.end(), by_deg);
      for (int w : nb) {
        done[w] = 1;
        cm.push_back(w);
      }
    }
  }
  std::vector<int> new_of_old(n);
  for (int k = 0; k < n; ++k) new_of_old[cm[n - 1 - k]] = k;
  return new_of_old;
}

}  // namespace op2h
