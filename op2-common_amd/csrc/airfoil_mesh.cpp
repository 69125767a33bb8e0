// Synthetic airfoil-channel mesh generator, text I/O and RCM renumbering.
// Conventions follow apps/mesh_generators/naca0012.m:217-303 (orientation,
// bound codes) and the reader in apps/c/airfoil/airfoil_plain/dp/airfoil.cpp:94-141.
#include "op2amd_mesh.h"
#include "op2h_rcm.h"

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <vector>

namespace {

struct SplitMix64 {
  uint64_t s;
  explicit SplitMix64(uint64_t seed) : s(seed) {}
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  double uniform() { return (double)(next() >> 11) * 0x1.0p-53; }
  uint64_t bounded(uint64_t n) {
    return (uint64_t)(((unsigned __int128)next() * n) >> 64);
  }
};

// order[new] = old, built by Fisher-Yates from the top index down.
std::vector<int> fisher_yates(SplitMix64 &rng, int n) {
  std::vector<int> order(n);
  std::iota(order.begin(), order.end(), 0);
  for (int k = n - 1; k >= 1; --k) {
    int r = (int)rng.bounded((uint64_t)k + 1);
    std::swap(order[k], order[r]);
  }
  return order;
}

template <class T> T *alloc(size_t n) {
  return (T *)std::malloc(std::max<size_t>(n, 1) * sizeof(T));
}

bool mesh_alloc(op2amd_mesh *m) {
  m->x = alloc<double>(2 * (size_t)m->nnode);
  m->cell = alloc<int>(4 * (size_t)m->ncell);
  m->edge = alloc<int>(2 * (size_t)m->nedge);
  m->ecell = alloc<int>(2 * (size_t)m->nedge);
  m->bedge = alloc<int>(2 * (size_t)m->nbedge);
  m->becell = alloc<int>((size_t)m->nbedge);
  m->bound = alloc<int>((size_t)m->nbedge);
  return m->x && m->cell && m->edge && m->ecell && m->bedge && m->becell &&
         m->bound;
}

// Apply a permutation new_of_old to the rows of a [n][dim] table.
template <class T>
void permute_rows(T *a, int n, int dim, const std::vector<int> &new_of_old) {
  std::vector<T> tmp(a, a + (size_t)n * dim);
  for (int o = 0; o < n; ++o)
    std::memcpy(a + (size_t)new_of_old[o] * dim, tmp.data() + (size_t)o * dim,
                sizeof(T) * dim);
}

void remap_entries(int *a, size_t n, const std::vector<int> &new_of_old) {
  for (size_t i = 0; i < n; ++i) a[i] = new_of_old[a[i]];
}

std::vector<int> invert(const std::vector<int> &order) {
  std::vector<int> inv(order.size());
  for (size_t i = 0; i < order.size(); ++i) inv[order[i]] = (int)i;
  return inv;
}

// Cell adjacency (CSR) through interior edges, neighbour lists sorted by index.
void cell_graph(const op2amd_mesh *m, std::vector<int> &off,
                std::vector<int> &adj) {
  off.assign(m->ncell + 1, 0);
  for (int e = 0; e < m->nedge; ++e) {
    off[m->ecell[2 * e] + 1]++;
    off[m->ecell[2 * e + 1] + 1]++;
  }
  for (int c = 0; c < m->ncell; ++c) off[c + 1] += off[c];
  adj.assign(off[m->ncell], 0);
  std::vector<int> pos(off.begin(), off.end() - 1);
  for (int e = 0; e < m->nedge; ++e) {
    int a = m->ecell[2 * e], b = m->ecell[2 * e + 1];
    adj[pos[a]++] = b;
    adj[pos[b]++] = a;
  }
  for (int c = 0; c < m->ncell; ++c)
    std::sort(adj.begin() + off[c], adj.begin() + off[c + 1]);
}

}  // namespace

namespace {
int renumber_cells(op2amd_mesh *m, const std::vector<int> &cnew,
                   std::vector<int> *enew_out, std::vector<int> *bnew_out,
                   std::vector<int> *nnew_out);
}

extern "C" int op2amd_mesh_generate(int I, int J, uint64_t seed, int ordering,
                                    op2amd_mesh *m) {
  if (!m || I < 2 || J < 2 || ordering < 0 || ordering > 3) return 1;
  std::memset(m, 0, sizeof(*m));
  m->nnode = (I + 1) * (J + 1);
  m->ncell = I * J;
  m->nedge = (I - 1) * J + I * (J - 1);
  m->nbedge = 2 * I + 2 * J;
  if (!mesh_alloc(m)) {
    op2amd_mesh_free(m);
    return 2;
  }
  auto node = [I](int i, int j) { return i + j * (I + 1); };
  auto cellid = [I](int i, int j) { return i + j * I; };
  const double hx = 2.0 / I, hy = 1.0 / J;
  SplitMix64 rng(seed);

  for (int j = 0; j <= J; ++j)
    for (int i = 0; i <= I; ++i) {
      int k = node(i, j);
      m->x[2 * k] = 2.0 * i / I;
      m->x[2 * k + 1] = 1.0 * j / J;
    }
  for (int j = 1; j < J; ++j)
    for (int i = 1; i < I; ++i) {
      int k = node(i, j);
      double ux = rng.uniform(), uy = rng.uniform();
      m->x[2 * k] += (2.0 * ux - 1.0) * 0.2 * hx;
      m->x[2 * k + 1] += (2.0 * uy - 1.0) * 0.2 * hy;
    }

  for (int j = 0; j < J; ++j)
    for (int i = 0; i < I; ++i) {
      int c = cellid(i, j);
      m->cell[4 * c + 0] = node(i, j);
      m->cell[4 * c + 1] = node(i + 1, j);
      m->cell[4 * c + 2] = node(i + 1, j + 1);
      m->cell[4 * c + 3] = node(i, j + 1);
    }

  int ie = 0;
  for (int j = 0; j < J; ++j)
    for (int i = 0; i < I; ++i) {
      if (j < J - 1) {  // face between (i,j) and (i,j+1): normal +y
        m->edge[2 * ie] = node(i, j + 1);
        m->edge[2 * ie + 1] = node(i + 1, j + 1);
        m->ecell[2 * ie] = cellid(i, j);
        m->ecell[2 * ie + 1] = cellid(i, j + 1);
        ++ie;
      }
      if (i < I - 1) {  // face between (i+1,j) and (i,j): normal -x
        m->edge[2 * ie] = node(i + 1, j);
        m->edge[2 * ie + 1] = node(i + 1, j + 1);
        m->ecell[2 * ie] = cellid(i + 1, j);
        m->ecell[2 * ie + 1] = cellid(i, j);
        ++ie;
      }
    }

  int ib = 0;
  auto bset = [&](int n0, int n1, int c, int b) {
    m->bedge[2 * ib] = n0;
    m->bedge[2 * ib + 1] = n1;
    m->becell[ib] = c;
    m->bound[ib] = b;
    ++ib;
  };
  for (int i = 0; i < I; ++i) bset(node(i + 1, 0), node(i, 0), cellid(i, 0), 1);
  for (int j = 0; j < J; ++j)
    bset(node(I, j + 1), node(I, j), cellid(I - 1, j), 2);
  for (int i = I - 1; i >= 0; --i)
    bset(node(i, J), node(i + 1, J), cellid(i, J - 1), 1);
  for (int j = J - 1; j >= 0; --j) bset(node(0, j), node(0, j + 1), cellid(0, j), 2);

  if (ordering >= OP2AMD_ORDER_RANDOM) {
    std::vector<int> corder = fisher_yates(rng, m->ncell);
    std::vector<int> eorder = fisher_yates(rng, m->nedge);
    std::vector<int> cnew = invert(corder), enew = invert(eorder);
    permute_rows(m->cell, m->ncell, 4, cnew);
    remap_entries(m->ecell, 2 * (size_t)m->nedge, cnew);
    remap_entries(m->becell, (size_t)m->nbedge, cnew);
    permute_rows(m->edge, m->nedge, 2, enew);
    permute_rows(m->ecell, m->nedge, 2, enew);
  }
  if (ordering == OP2AMD_ORDER_RCM)
    return op2amd_mesh_renumber_rcm(m, nullptr, nullptr, nullptr, nullptr);
  if (ordering == OP2AMD_ORDER_HILBERT) {
    // structured position of every (now randomly numbered) cell from its
    // lowest-left node, which is node (i, j) = cell[4c]
    std::vector<long long> key(m->ncell);
    int order = 1;
    while ((1 << order) < std::max(I, J)) ++order;
    const long long side = 1LL << order;
    for (int c = 0; c < m->ncell; ++c) {
      int n0 = m->cell[4 * c];
      long long x = n0 % (I + 1), y = n0 / (I + 1);
      // Hilbert index of (x, y) on a side x side grid (iterative d2 mapping)
      long long d = 0;
      for (long long s = side / 2; s > 0; s /= 2) {
        long long rx = (x & s) > 0, ry = (y & s) > 0;
        d += s * s * ((3 * rx) ^ ry);
        if (ry == 0) {
          if (rx == 1) {
            x = side - 1 - x;
            y = side - 1 - y;
          }
          std::swap(x, y);
        }
      }
      key[c] = d;
    }
    std::vector<int> order_c(m->ncell);
    std::iota(order_c.begin(), order_c.end(), 0);
    std::stable_sort(order_c.begin(), order_c.end(),
                     [&](int a, int b) { return key[a] < key[b]; });
    return renumber_cells(m, invert(order_c), nullptr, nullptr, nullptr);
  }
  return 0;
}

extern "C" void op2amd_mesh_free(op2amd_mesh *m) {
  if (!m) return;
  std::free(m->x);
  std::free(m->cell);
  std::free(m->edge);
  std::free(m->ecell);
  std::free(m->bedge);
  std::free(m->becell);
  std::free(m->bound);
  std::memset(m, 0, sizeof(*m));
}

extern "C" int op2amd_mesh_write(const op2amd_mesh *m, const char *path) {
  FILE *fp = std::fopen(path, "w");
  if (!fp) return 1;
  std::fprintf(fp, "%d %d %d %d \n", m->nnode, m->ncell, m->nedge, m->nbedge);
  for (int n = 0; n < m->nnode; ++n)
    std::fprintf(fp, "%.17g %.17g \n", m->x[2 * n], m->x[2 * n + 1]);
  for (int n = 0; n < m->ncell; ++n)
    std::fprintf(fp, "%d %d %d %d \n", m->cell[4 * n], m->cell[4 * n + 1],
                 m->cell[4 * n + 2], m->cell[4 * n + 3]);
  for (int n = 0; n < m->nedge; ++n)
    std::fprintf(fp, "%d %d %d %d \n", m->edge[2 * n], m->edge[2 * n + 1],
                 m->ecell[2 * n], m->ecell[2 * n + 1]);
  for (int n = 0; n < m->nbedge; ++n)
    std::fprintf(fp, "%d %d %d %d \n", m->bedge[2 * n], m->bedge[2 * n + 1],
                 m->becell[n], m->bound[n]);
  return std::fclose(fp) == 0 ? 0 : 2;
}

extern "C" int op2amd_mesh_read(const char *path, op2amd_mesh *m) {
  std::memset(m, 0, sizeof(*m));
  FILE *fp = std::fopen(path, "r");
  if (!fp) return 1;
  int rc = 0;
  if (std::fscanf(fp, "%d %d %d %d", &m->nnode, &m->ncell, &m->nedge,
                  &m->nbedge) != 4 ||
      m->nnode < 0 || m->ncell < 0 || m->nedge < 0 || m->nbedge < 0) {
    std::fclose(fp);
    return 2;
  }
  if (!mesh_alloc(m)) rc = 3;
  for (int n = 0; !rc && n < m->nnode; ++n)
    if (std::fscanf(fp, "%lf %lf", &m->x[2 * n], &m->x[2 * n + 1]) != 2) rc = 4;
  for (int n = 0; !rc && n < m->ncell; ++n)
    if (std::fscanf(fp, "%d %d %d %d", &m->cell[4 * n], &m->cell[4 * n + 1],
                    &m->cell[4 * n + 2], &m->cell[4 * n + 3]) != 4)
      rc = 4;
  for (int n = 0; !rc && n < m->nedge; ++n)
    if (std::fscanf(fp, "%d %d %d %d", &m->edge[2 * n], &m->edge[2 * n + 1],
                    &m->ecell[2 * n], &m->ecell[2 * n + 1]) != 4)
      rc = 4;
  for (int n = 0; !rc && n < m->nbedge; ++n)
    if (std::fscanf(fp, "%d %d %d %d", &m->bedge[2 * n], &m->bedge[2 * n + 1],
                    &m->becell[n], &m->bound[n]) != 4)
      rc = 4;
  std::fclose(fp);
  if (rc) op2amd_mesh_free(m);
  return rc;
}

namespace {
// Renumber cells by `cnew` (new index of each old cell) and propagate to the
// other sets with the rules of op_renumber.cpp:72-151 (see op2amd_mesh.h).
int renumber_cells(op2amd_mesh *m, const std::vector<int> &cnew,
                   std::vector<int> *enew_out, std::vector<int> *bnew_out,
                   std::vector<int> *nnew_out) {
  const int nc = m->ncell;
  std::vector<int> cm(nc);  // cells in new order
  for (int c = 0; c < nc; ++c) cm[cnew[c]] = c;
  // edges: stable sort by new index of ecell[e][0]
  // (a counting sort on the key: stable and linear)
  auto stable_by_cell = [&](const int *col, int stride, int n) {
    std::vector<int> start(nc + 1, 0), out(n);
    for (int e = 0; e < n; ++e) ++start[cnew[col[(size_t)stride * e]] + 1];
    for (int c = 0; c < nc; ++c) start[c + 1] += start[c];
    for (int e = 0; e < n; ++e) out[e] = start[cnew[col[(size_t)stride * e]]]++;
    return out;  // new index of each old element
  };
  std::vector<int> enew = stable_by_cell(m->ecell, 2, m->nedge);
  std::vector<int> bnew = stable_by_cell(m->becell, 1, m->nbedge);
  // nodes: first visit through cell->node in the new cell order
  std::vector<int> nnew(m->nnode, -1);
  int counter = 0;
  for (int k = 0; k < nc; ++k) {
    int c = cm[k];
    for (int d = 0; d < 4; ++d) {
      int v = m->cell[4 * c + d];
      if (nnew[v] < 0) nnew[v] = counter++;
    }
  }
  for (int v = 0; v < m->nnode; ++v)  // nodes referenced by no cell keep order
    if (nnew[v] < 0) nnew[v] = counter++;

  permute_rows(m->cell, nc, 4, cnew);
  remap_entries(m->cell, 4 * (size_t)nc, nnew);
  permute_rows(m->x, m->nnode, 2, nnew);
  permute_rows(m->edge, m->nedge, 2, enew);
  permute_rows(m->ecell, m->nedge, 2, enew);
  remap_entries(m->edge, 2 * (size_t)m->nedge, nnew);
  remap_entries(m->ecell, 2 * (size_t)m->nedge, cnew);
  permute_rows(m->bedge, m->nbedge, 2, bnew);
  permute_rows(m->becell, m->nbedge, 1, bnew);
  permute_rows(m->bound, m->nbedge, 1, bnew);
  remap_entries(m->bedge, 2 * (size_t)m->nbedge, nnew);
  remap_entries(m->becell, (size_t)m->nbedge, cnew);

  if (enew_out) *enew_out = enew;
  if (bnew_out) *bnew_out = bnew;
  if (nnew_out) *nnew_out = nnew;
  return 0;
}
}  // namespace

extern "C" int op2amd_mesh_renumber_rcm(op2amd_mesh *m, int *cell_perm,
                                        int *edge_perm, int *bedge_perm,
                                        int *node_perm) {
  const int nc = m->ncell;
  std::vector<int> off, adj;
  cell_graph(m, off, adj);
  std::vector<int> cnew = op2h::rcm_order(nc, off, adj);
  std::vector<int> enew, bnew, nnew;
  renumber_cells(m, cnew, &enew, &bnew, &nnew);
  if (cell_perm) std::copy(cnew.begin(), cnew.end(), cell_perm);
  if (edge_perm) std::copy(enew.begin(), enew.end(), edge_perm);
  if (bedge_perm) std::copy(bnew.begin(), bnew.end(), bedge_perm);
  if (node_perm) std::copy(nnew.begin(), nnew.end(), node_perm);
  return 0;
}

extern "C" void op2amd_mesh_cell_bandwidth(const op2amd_mesh *m, int *max_bw,
                                           double *mean_bw) {
  long long sum = 0;
  int mx = 0;
  for (int e = 0; e < m->nedge; ++e) {
    int d = std::abs(m->ecell[2 * e] - m->ecell[2 * e + 1]);
    mx = std::max(mx, d);
    sum += d;
  }
  if (max_bw) *max_bw = mx;
  if (mean_bw) *mean_bw = m->nedge ? (double)sum / m->nedge : 0.0;
}
