// Reverse Cuthill-McKee order of the basis functions (declared in order.hpp).
#include "order.hpp"

#include <omp.h>

#include <algorithm>
#include <cmath>
#include <vector>

namespace nopt {

namespace {

struct Levels {
  std::vector<int> order;  // visit order, neighbours by increasing degree
  std::vector<int> level;  // BFS depth of each visited node, -1 elsewhere
  int depth = 0;
};

// breadth-first search from s over the nodes not yet placed
Levels bfs(const std::vector<std::vector<int>>& adj, const std::vector<int>& deg, const std::vector<char>& done,
           int s) {
  Levels L;
  L.level.assign(adj.size(), -1);
  L.order.push_back(s);
  L.level[(size_t)s] = 0;
  for (size_t h = 0; h < L.order.size(); ++h) {
    const int u = L.order[h];
    std::vector<int> nb;
    for (int v : adj[(size_t)u])
      if (!done[(size_t)v] && L.level[(size_t)v] < 0) {
        L.level[(size_t)v] = L.level[(size_t)u] + 1;
        nb.push_back(v);
      }
    std::sort(nb.begin(), nb.end(), [&](int a, int b) { return deg[(size_t)a] != deg[(size_t)b] ? deg[(size_t)a] < deg[(size_t)b] : a < b; });
    L.order.insert(L.order.end(), nb.begin(), nb.end());
  }
  L.depth = L.level[(size_t)L.order.back()];
  return L;
}

}  // namespace

void band_order(const double* T, long nrow, long nbf, int* perm) {
  const long npair = nbf * (nbf + 1) / 2;
  std::vector<double> mx((size_t)npair, 0.0);
#pragma omp parallel
  {
    // each thread owns a contiguous range of pairs (m, n) over every row P
    const int nt = omp_get_num_threads(), id = omp_get_thread_num();
    const long a = npair * id / nt, b = npair * (id + 1) / nt;
    for (long P = 0; P < nrow; ++P) {
      const double* r = T + P * npair;
      for (long k = a; k < b; ++k) mx[(size_t)k] = std::max(mx[(size_t)k], std::fabs(r[k]));
    }
  }
  std::vector<std::vector<int>> adj((size_t)nbf);
  for (long m = 0; m < nbf; ++m) {
    const double dm = mx[(size_t)(m * (m + 1) / 2 + m)];
    for (long n = 0; n < m; ++n) {
      const double dn = mx[(size_t)(n * (n + 1) / 2 + n)];
      if (mx[(size_t)(m * (m + 1) / 2 + n)] > std::ldexp(std::sqrt(dm * dn), -30)) {
        adj[(size_t)m].push_back((int)n);
        adj[(size_t)n].push_back((int)m);
      }
    }
  }
  std::vector<int> deg((size_t)nbf);
  for (long i = 0; i < nbf; ++i) deg[(size_t)i] = (int)adj[(size_t)i].size();
  std::vector<char> done((size_t)nbf, 0);
  std::vector<int> out;
  out.reserve((size_t)nbf);
  for (;;) {
    int s = -1;  // the unplaced node of lowest degree (lowest index on ties)
    for (long i = 0; i < nbf; ++i)
      if (!done[(size_t)i] && (s < 0 || deg[(size_t)i] < deg[(size_t)s])) s = (int)i;
    if (s < 0) break;
    // a pseudo-peripheral start: move to the lowest-degree node of the last
    // level while that deepens the level structure
    Levels L = bfs(adj, deg, done, s);
    for (int it = 0; it < 8; ++it) {
      int best = -1;
      for (int u : L.order)
        if (L.level[(size_t)u] == L.depth && (best < 0 || deg[(size_t)u] < deg[(size_t)best])) best = u;
      Levels L2 = bfs(adj, deg, done, best);
      if (L2.depth <= L.depth) break;
      L = std::move(L2);
    }
    for (int u : L.order) done[(size_t)u] = 1;
    out.insert(out.end(), L.order.begin(), L.order.end());
  }
  std::reverse(out.begin(), out.end());
  std::copy(out.begin(), out.end(), perm);
}

}  // namespace nopt
