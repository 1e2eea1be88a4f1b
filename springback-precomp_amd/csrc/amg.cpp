// The smoothed-aggregation hierarchy's construction on the host: node
// graphs, aggregates, tentative prolongators (orthonormalised rigid-body
// modes per aggregate) and the patterns of P, A P, P^T and the next level,
// uploaded once per partition; the values are the device's (amg_kernels).
//
// Every level is a set of nodes whose rows are contiguous (level 0: the free
// DOFs of one mesh node; level l + 1: the coarse DOFs of one aggregate), and
// all rows of a node share their columns, so the patterns are built node by
// node and written out row by row.

#include "amg.hpp"

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdint>
#include <memory>

namespace sbk {

namespace {

constexpr long long kCoarseMax = 256;  // a level this small is solved densely
constexpr int kMaxLevels = 12;
constexpr int kModes = 6;              // rigid-body modes: 3 translations, 3 rotations

struct Graph {
  std::vector<int> ptr{0};
  std::vector<int> adj;  // per node: its neighbours (itself included), ascending
  int nodes() const { return static_cast<int>(ptr.size()) - 1; }
};

struct HostLevel {
  long long n = 0;
  std::vector<int> node_ptr{0};   // each node's rows
  Graph graph;
  std::vector<int> ptr, index;    // the level's pattern (rows)
  std::vector<double> modes;      // level >= 1: n x kModes near-nullspace
};

// Greedy aggregation (Vanek, Mandel and Brezina's smoothed aggregation):
// (1) a node whose neighbours are all unaggregated roots an aggregate of
// itself and them; (2) each node left joins the aggregate of its first
// neighbour aggregated in (1); (3) the nodes still left form aggregates with
// their unaggregated neighbours. Aggregates are numbered by their lowest node.
std::vector<int> aggregate(const Graph& g, int& count) {
  const int nn = g.nodes();
  std::vector<int> agg(nn, -1);
  int na = 0;
  for (int i = 0; i < nn; ++i) {
    if (agg[i] >= 0) continue;
    bool open = true;
    for (int k = g.ptr[i]; k < g.ptr[i + 1] && open; ++k) open = agg[g.adj[k]] < 0;
    if (!open) continue;
    for (int k = g.ptr[i]; k < g.ptr[i + 1]; ++k) agg[g.adj[k]] = na;
    agg[i] = na++;
  }
  std::vector<int> joined(agg);
  for (int i = 0; i < nn; ++i) {
    if (agg[i] >= 0) continue;
    for (int k = g.ptr[i]; k < g.ptr[i + 1]; ++k)
      if (agg[g.adj[k]] >= 0) {
        joined[i] = agg[g.adj[k]];
        break;
      }
  }
  agg.swap(joined);
  for (int i = 0; i < nn; ++i) {
    if (agg[i] >= 0) continue;
    for (int k = g.ptr[i]; k < g.ptr[i + 1]; ++k)
      if (agg[g.adj[k]] < 0) agg[g.adj[k]] = na;
    agg[i] = na++;
  }
  std::vector<int> renum(na, -1);
  int next = 0;
  for (int i = 0; i < nn; ++i)
    if (renum[agg[i]] < 0) renum[agg[i]] = next++;
  for (int i = 0; i < nn; ++i) agg[i] = renum[agg[i]];
  count = na;
  return agg;
}

// Modified Gram-Schmidt (twice) on the m x kModes columns of b (row-major):
// q (m x r, row-major) orthonormal, r the columns kept (a column whose
// remainder is below 1e-10 of its norm, or 1e-12 of the largest column, is
// dependent and dropped), and rr = q^T b (r x kModes).
int orthonormalise(const std::vector<double>& b, int m, std::vector<double>& q, std::vector<double>& rr) {
  double top = 0.0;
  std::vector<double> norms(kModes, 0.0);
  for (int j = 0; j < kModes; ++j) {
    double s = 0.0;
    for (int i = 0; i < m; ++i) s += b[i * kModes + j] * b[i * kModes + j];
    norms[j] = std::sqrt(s);
    top = std::max(top, norms[j]);
  }
  std::vector<std::vector<double>> cols;
  std::vector<double> v(m);
  for (int j = 0; j < kModes; ++j) {
    for (int i = 0; i < m; ++i) v[i] = b[i * kModes + j];
    for (int pass = 0; pass < 2; ++pass)
      for (const auto& c : cols) {
        double h = 0.0;
        for (int i = 0; i < m; ++i) h += c[i] * v[i];
        for (int i = 0; i < m; ++i) v[i] -= h * c[i];
      }
    double nv = 0.0;
    for (int i = 0; i < m; ++i) nv += v[i] * v[i];
    nv = std::sqrt(nv);
    if (nv > 1e-10 * norms[j] && nv > 1e-12 * top) {
      for (int i = 0; i < m; ++i) v[i] /= nv;
      cols.push_back(v);
    }
  }
  const int r = static_cast<int>(cols.size());
  q.assign(static_cast<std::size_t>(m) * r, 0.0);
  for (int k = 0; k < r; ++k)
    for (int i = 0; i < m; ++i) q[static_cast<std::size_t>(i) * r + k] = cols[k][i];
  rr.assign(static_cast<std::size_t>(r) * kModes, 0.0);
  for (int k = 0; k < r; ++k)
    for (int j = 0; j < kModes; ++j) {
      double s = 0.0;
      for (int i = 0; i < m; ++i) s += cols[k][i] * b[i * kModes + j];
      rr[k * kModes + j] = s;
    }
  return r;
}

// The sorted union of lists[members] (CSR lists) with a stamp per entry.
void merge(const std::vector<int>& lptr, const std::vector<int>& lists, const int* members, int count,
           std::vector<int>& stamp, int tag, std::vector<int>& out) {
  out.clear();
  for (int t = 0; t < count; ++t) {
    const int m = members[t];
    for (int k = lptr[m]; k < lptr[m + 1]; ++k)
      if (stamp[lists[k]] != tag) {
        stamp[lists[k]] = tag;
        out.push_back(lists[k]);
      }
  }
  std::sort(out.begin(), out.end());
}

bool fits(long long v) { return v >= 0 && v <= INT_MAX; }

}  // namespace

AmgHierarchy::~AmgHierarchy() { release(); }

void AmgHierarchy::release() {
  for (void* p : owned_) (void)hipFree(p);
  owned_.clear();
  lv_.clear();
  nnz_.clear();
  transpose_ = nullptr;
  symmetric_ = nullptr;
  nnz0_ = 0;
}

template <class T>
T* AmgHierarchy::alloc(std::size_t count, bool& ok) {
  if (!ok) return nullptr;
  void* p = nullptr;
  if (hipMalloc(&p, (count ? count : 1) * sizeof(T)) != hipSuccess) {
    ok = false;
    return nullptr;
  }
  owned_.push_back(p);
  return static_cast<T*>(p);
}

template <class T>
T* AmgHierarchy::upload(const std::vector<T>& v, hipStream_t s, bool& ok) {
  T* p = alloc<T>(v.size(), ok);
  if (ok && !v.empty() &&
      (hipMemcpyAsync(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s) != hipSuccess ||
       hipStreamSynchronize(s) != hipSuccess))
    ok = false;
  return p;
}

std::string AmgHierarchy::build(const int* block_ptr, const int* block_index, long long n, long long nnz,
                                const std::vector<long long>& free_dofs, const std::vector<double>& coords,
                                hipStream_t s) {
  release();
  std::vector<int> ptr(static_cast<std::size_t>(n) + 1), index(static_cast<std::size_t>(nnz));
  if (hipMemcpyAsync(ptr.data(), block_ptr, ptr.size() * sizeof(int), hipMemcpyDeviceToHost, s) != hipSuccess ||
      (nnz > 0 &&
       hipMemcpyAsync(index.data(), block_index, index.size() * sizeof(int), hipMemcpyDeviceToHost, s) !=
           hipSuccess) ||
      hipStreamSynchronize(s) != hipSuccess)
    return "reading the block's pattern failed";
  return build_levels(std::move(ptr), std::move(index), block_ptr, block_index, free_dofs, coords, s);
}

std::string AmgHierarchy::build_host(std::vector<int> ptr, std::vector<int> index,
                                     const std::vector<long long>& free_dofs, const std::vector<double>& coords,
                                     hipStream_t s) {
  release();
  bool ok = true;
  const int* dptr = upload(ptr, s, ok);
  const int* dindex = upload(index, s, ok);
  if (!ok) {
    release();
    return "device allocation failed";
  }
  return build_levels(std::move(ptr), std::move(index), dptr, dindex, free_dofs, coords, s);
}

std::string AmgHierarchy::build_levels(std::vector<int> ptr, std::vector<int> index, const int* block_ptr,
                                       const int* block_index, const std::vector<long long>& free_dofs,
                                       const std::vector<double>& coords, hipStream_t s) {
  const long long n = static_cast<long long>(ptr.size()) - 1;
  const long long nnz = static_cast<long long>(index.size());
  if (static_cast<long long>(free_dofs.size()) != n) {
    release();
    return "the free DOF list does not match the block";
  }
  HostLevel cur;
  cur.n = n;
  cur.ptr = std::move(ptr);
  cur.index = std::move(index);
  // level 0's nodes: runs of rows of one mesh node
  std::vector<long long> mesh_node;  // per level-0 node
  for (long long i = 0; i < n; ++i) {
    const long long nd = free_dofs[i] / 3;
    if (mesh_node.empty() || mesh_node.back() != nd) {
      if (!mesh_node.empty()) cur.node_ptr.push_back(static_cast<int>(i));
      mesh_node.push_back(nd);
    }
  }
  cur.node_ptr.push_back(static_cast<int>(n));
  if (n == 0) cur.node_ptr.assign(1, 0);
  for (long long nd : mesh_node)
    if (3 * nd + 2 >= static_cast<long long>(coords.size())) {
      release();
      return "a free DOF beyond the coordinates";
    }
  double c[3] = {0.0, 0.0, 0.0}, h = 0.0;
  if (!mesh_node.empty()) {
    double lo[3], hi[3];
    for (int d = 0; d < 3; ++d) lo[d] = hi[d] = coords[3 * mesh_node[0] + d];
    for (long long nd : mesh_node)
      for (int d = 0; d < 3; ++d) {
        lo[d] = std::min(lo[d], coords[3 * nd + d]);
        hi[d] = std::max(hi[d], coords[3 * nd + d]);
      }
    for (int d = 0; d < 3; ++d) {
      c[d] = 0.5 * (lo[d] + hi[d]);
      h = std::max(h, 0.5 * (hi[d] - lo[d]));
    }
  }
  if (!(h > 0.0)) h = 1.0;

  bool ok = true;
  {
    // level 0's mirror map (its symmetric part for a block that is not
    // symmetric)
    std::vector<int> mirror(static_cast<std::size_t>(nnz), -1);
    for (long long i = 0; i < n; ++i)
      for (int k = cur.ptr[i]; k < cur.ptr[i + 1]; ++k) {
        const int j = cur.index[k];
        const int* b = cur.index.data() + cur.ptr[j];
        const int* e = cur.index.data() + cur.ptr[j + 1];
        const int* f = std::lower_bound(b, e, static_cast<int>(i));
        if (f != e && *f == i) mirror[k] = static_cast<int>(f - cur.index.data());
      }
    transpose_ = upload(mirror, s, ok);
    symmetric_ = alloc<double>(static_cast<std::size_t>(nnz), ok);
    nnz0_ = nnz;
    if (!ok) {
      release();
      return "device allocation failed";
    }
  }
  for (int level = 0;; ++level) {
    const int nn = static_cast<int>(cur.node_ptr.size()) - 1;
    std::vector<int> row_node(static_cast<std::size_t>(cur.n));
    for (int a = 0; a < nn; ++a)
      for (int i = cur.node_ptr[a]; i < cur.node_ptr[a + 1]; ++i) row_node[i] = a;
    // the node graph from the pattern (a node's first row)
    cur.graph = Graph();
    for (int a = 0; a < nn; ++a) {
      const int r0 = cur.node_ptr[a];
      int last = -1;
      for (int k = cur.ptr[r0]; k < cur.ptr[r0 + 1]; ++k) {
        const int b = row_node[cur.index[k]];
        if (b != last) cur.graph.adj.push_back(b);
        last = b;
      }
      cur.graph.ptr.push_back(static_cast<int>(cur.graph.adj.size()));
    }
    AmgLevel L;
    L.n = cur.n;
    nnz_.push_back(static_cast<long long>(cur.index.size()));
    if (level == 0) {
      L.ptr = block_ptr;
      L.index = block_index;
    } else {
      L.ptr = upload(cur.ptr, s, ok);
      L.index = upload(cur.index, s, ok);
      L.values = alloc<double>(cur.index.size(), ok);
    }
    L.dinv = alloc<double>(static_cast<std::size_t>(cur.n), ok);
    L.coef = alloc<double>(kAmgCoefs, ok);
    L.nparts = amg_product_blocks(cur.n);
    L.partial = alloc<double>(static_cast<std::size_t>(L.nparts), ok);
    L.r = alloc<double>(static_cast<std::size_t>(cur.n), ok);
    L.d0 = alloc<double>(static_cast<std::size_t>(cur.n), ok);
    L.d1 = alloc<double>(static_cast<std::size_t>(cur.n), ok);
    if (level > 0) {
      L.b = alloc<double>(static_cast<std::size_t>(cur.n), ok);
      L.x = alloc<double>(static_cast<std::size_t>(cur.n), ok);
    }
    if (!ok) {
      release();
      return "device allocation failed";
    }
    // coarsen?
    int na = 0;
    std::vector<int> agg;
    bool coarsest = cur.n <= kCoarseMax || level == kMaxLevels - 1;
    if (!coarsest) agg = aggregate(cur.graph, na);
    // the aggregates' rows (their nodes' rows in node order) and modes
    std::vector<int> anode_ptr, anodes, rank, coff;
    std::vector<std::vector<double>> q(na), rr(na);
    std::vector<int> local(static_cast<std::size_t>(cur.n), 0);  // row's place among its aggregate's rows
    long long nc = 0;
    if (!coarsest) {
      anode_ptr.assign(na + 1, 0);
      for (int a = 0; a < nn; ++a) ++anode_ptr[agg[a] + 1];
      for (int j = 0; j < na; ++j) anode_ptr[j + 1] += anode_ptr[j];
      anodes.resize(nn);
      std::vector<int> fill(anode_ptr.begin(), anode_ptr.end() - 1);
      for (int a = 0; a < nn; ++a) anodes[fill[agg[a]]++] = a;
      rank.assign(na, 0);
      coff.assign(na + 1, 0);
      std::vector<double> b;
      for (int j = 0; j < na; ++j) {
        int m = 0;
        for (int t = anode_ptr[j]; t < anode_ptr[j + 1]; ++t) {
          const int a = anodes[t];
          for (int i = cur.node_ptr[a]; i < cur.node_ptr[a + 1]; ++i) local[i] = m++;
        }
        b.assign(static_cast<std::size_t>(m) * kModes, 0.0);
        if (level == 0) {
          // the rigid-body modes of the whole block (every aggregate's rows
          // of the same six global modes, so that the coarse levels keep
          // them): rotations about the nodes' common centre, lengths in
          // units of their extent
          for (int t = anode_ptr[j]; t < anode_ptr[j + 1]; ++t) {
            const int a = anodes[t];
            const double x = (coords[3 * mesh_node[a]] - c[0]) / h, y = (coords[3 * mesh_node[a] + 1] - c[1]) / h,
                         z = (coords[3 * mesh_node[a] + 2] - c[2]) / h;
            for (int i = cur.node_ptr[a]; i < cur.node_ptr[a + 1]; ++i) {
              double* row = &b[static_cast<std::size_t>(local[i]) * kModes];
              const int comp = static_cast<int>(free_dofs[i] % 3);
              row[comp] = 1.0;
              // rotations about x, y, z: (0, -z, y), (z, 0, -x), (-y, x, 0)
              if (comp == 0) { row[4] = z; row[5] = -y; }
              if (comp == 1) { row[3] = -z; row[5] = x; }
              if (comp == 2) { row[3] = y; row[4] = -x; }
            }
          }
        } else {
          for (int t = anode_ptr[j]; t < anode_ptr[j + 1]; ++t) {
            const int a = anodes[t];
            for (int i = cur.node_ptr[a]; i < cur.node_ptr[a + 1]; ++i)
              for (int k = 0; k < kModes; ++k)
                b[static_cast<std::size_t>(local[i]) * kModes + k] = cur.modes[static_cast<std::size_t>(i) * kModes + k];
          }
        }
        rank[j] = orthonormalise(b, m, q[j], rr[j]);
        coff[j + 1] = coff[j] + rank[j];
      }
      nc = coff[na];
      if (nc == 0 || nc >= cur.n) coarsest = true;  // no coarsening: this level is the last
    }
    if (coarsest) {
      if (cur.n > kAmgDenseMax) {
        release();
        return "the hierarchy does not coarsen to a dense level";
      }
      L.dense = alloc<double>(2 * static_cast<std::size_t>(cur.n) * static_cast<std::size_t>(cur.n), ok);
      if (!ok) {
        release();
        return "device allocation failed";
      }
      lv_.push_back(L);
      break;
    }
    // patterns at node level: PA(a) aggregates next to node a; APA(a) those
    // within two steps; PTn(J) nodes next to aggregate J; AcA(J) aggregates
    // coupled with J in P^T A P
    std::vector<int> stamp(na, -1), tmp;
    std::vector<int> pa_ptr{0}, pa;
    for (int a = 0; a < nn; ++a) {
      tmp.clear();
      for (int k = cur.graph.ptr[a]; k < cur.graph.ptr[a + 1]; ++k) tmp.push_back(agg[cur.graph.adj[k]]);
      std::sort(tmp.begin(), tmp.end());
      tmp.erase(std::unique(tmp.begin(), tmp.end()), tmp.end());
      pa.insert(pa.end(), tmp.begin(), tmp.end());
      pa_ptr.push_back(static_cast<int>(pa.size()));
    }
    std::vector<int> apa_ptr{0}, apa;
    for (int a = 0; a < nn; ++a) {
      merge(pa_ptr, pa, &cur.graph.adj[cur.graph.ptr[a]], cur.graph.ptr[a + 1] - cur.graph.ptr[a], stamp, a, tmp);
      apa.insert(apa.end(), tmp.begin(), tmp.end());
      apa_ptr.push_back(static_cast<int>(apa.size()));
    }
    std::vector<int> ptn_ptr(na + 1, 0), ptn(pa.size());
    for (int a = 0; a < nn; ++a)
      for (int k = pa_ptr[a]; k < pa_ptr[a + 1]; ++k) ++ptn_ptr[pa[k] + 1];
    for (int j = 0; j < na; ++j) ptn_ptr[j + 1] += ptn_ptr[j];
    {
      std::vector<int> fill(ptn_ptr.begin(), ptn_ptr.end() - 1);
      for (int a = 0; a < nn; ++a)
        for (int k = pa_ptr[a]; k < pa_ptr[a + 1]; ++k) ptn[fill[pa[k]]++] = a;
    }
    std::fill(stamp.begin(), stamp.end(), -1);
    HostLevel nxt;
    nxt.n = nc;
    nxt.node_ptr = coff;
    for (int j = 0; j < na; ++j) {
      merge(apa_ptr, apa, ptn.data() + ptn_ptr[j], ptn_ptr[j + 1] - ptn_ptr[j], stamp, j, tmp);
      nxt.graph.adj.insert(nxt.graph.adj.end(), tmp.begin(), tmp.end());
      nxt.graph.ptr.push_back(static_cast<int>(nxt.graph.adj.size()));
    }
    // row-level arrays
    auto span = [&](const std::vector<int>& lptr, const std::vector<int>& lists, int a) {
      long long len = 0;
      for (int k = lptr[a]; k < lptr[a + 1]; ++k) len += rank[lists[k]];
      return len;
    };
    std::vector<int> tp_ptr{0}, tp_index, p_ptr{0}, p_index, p_own, ap_ptr{0}, ap_index;
    std::vector<double> tp_values;
    long long pcount = 0, apcount = 0;
    for (int a = 0; a < nn; ++a) {
      pcount += span(pa_ptr, pa, a) * (cur.node_ptr[a + 1] - cur.node_ptr[a]);
      apcount += span(apa_ptr, apa, a) * (cur.node_ptr[a + 1] - cur.node_ptr[a]);
    }
    if (!fits(pcount) || !fits(apcount)) {
      release();
      return "the hierarchy's patterns exceed 32-bit indices";
    }
    p_index.reserve(static_cast<std::size_t>(pcount));
    ap_index.reserve(static_cast<std::size_t>(apcount));
    p_own.resize(static_cast<std::size_t>(cur.n));
    for (int a = 0; a < nn; ++a) {
      const int j = agg[a];
      for (int i = cur.node_ptr[a]; i < cur.node_ptr[a + 1]; ++i) {
        for (int k = 0; k < rank[j]; ++k) {
          tp_index.push_back(coff[j] + k);
          tp_values.push_back(q[j][static_cast<std::size_t>(local[i]) * rank[j] + k]);
        }
        tp_ptr.push_back(static_cast<int>(tp_index.size()));
        int own = 0;
        for (int k = pa_ptr[a]; k < pa_ptr[a + 1]; ++k) {
          const int J = pa[k];
          if (J < j) own += rank[J];
          for (int c = coff[J]; c < coff[J + 1]; ++c) p_index.push_back(c);
        }
        p_own[i] = own;
        p_ptr.push_back(static_cast<int>(p_index.size()));
        for (int k = apa_ptr[a]; k < apa_ptr[a + 1]; ++k)
          for (int c = coff[apa[k]]; c < coff[apa[k] + 1]; ++c) ap_index.push_back(c);
        ap_ptr.push_back(static_cast<int>(ap_index.size()));
      }
    }
    // P^T: row c of aggregate J takes the rows of the nodes next to J, its
    // entries found in P's rows
    std::vector<int> pt_ptr{0}, pt_index, pt_perm;
    pt_index.reserve(static_cast<std::size_t>(pcount));
    pt_perm.reserve(static_cast<std::size_t>(pcount));
    for (int j = 0; j < na; ++j)
      for (int k = 0; k < rank[j]; ++k) {
        for (int t = ptn_ptr[j]; t < ptn_ptr[j + 1]; ++t) {
          const int a = ptn[t];
          int off = 0;
          for (int u = pa_ptr[a]; u < pa_ptr[a + 1] && pa[u] < j; ++u) off += rank[pa[u]];
          for (int i = cur.node_ptr[a]; i < cur.node_ptr[a + 1]; ++i) {
            pt_index.push_back(i);
            pt_perm.push_back(p_ptr[i] + off + k);
          }
        }
        pt_ptr.push_back(static_cast<int>(pt_index.size()));
      }
    // the next level's pattern and modes
    long long ccount = 0;
    for (int j = 0; j < na; ++j) ccount += span(nxt.graph.ptr, nxt.graph.adj, j) * rank[j];
    if (!fits(ccount)) {
      release();
      return "the hierarchy's patterns exceed 32-bit indices";
    }
    nxt.ptr.assign(1, 0);
    nxt.index.reserve(static_cast<std::size_t>(ccount));
    nxt.modes.resize(static_cast<std::size_t>(nc) * kModes);
    for (int j = 0; j < na; ++j)
      for (int k = 0; k < rank[j]; ++k) {
        for (int t = nxt.graph.ptr[j]; t < nxt.graph.ptr[j + 1]; ++t)
          for (int c = coff[nxt.graph.adj[t]]; c < coff[nxt.graph.adj[t] + 1]; ++c) nxt.index.push_back(c);
        nxt.ptr.push_back(static_cast<int>(nxt.index.size()));
        for (int m = 0; m < kModes; ++m)
          nxt.modes[static_cast<std::size_t>(coff[j] + k) * kModes + m] = rr[j][static_cast<std::size_t>(k) * kModes + m];
      }
    L.nc = nc;
    L.tp_ptr = upload(tp_ptr, s, ok);
    L.tp_index = upload(tp_index, s, ok);
    L.tp_values = upload(tp_values, s, ok);
    L.p_own = upload(p_own, s, ok);
    L.p_ptr = upload(p_ptr, s, ok);
    L.p_index = upload(p_index, s, ok);
    L.p_values = alloc<double>(p_index.size(), ok);
    L.pt_ptr = upload(pt_ptr, s, ok);
    L.pt_index = upload(pt_index, s, ok);
    L.pt_perm = upload(pt_perm, s, ok);
    L.pt_values = alloc<double>(pt_index.size(), ok);
    L.pt_nnz = static_cast<long long>(pt_index.size());
    L.ap_ptr = upload(ap_ptr, s, ok);
    L.ap_index = upload(ap_index, s, ok);
    L.ap_values = alloc<double>(ap_index.size(), ok);
    if (!ok) {
      release();
      return "device allocation failed";
    }
    lv_.push_back(L);
    cur = std::move(nxt);
  }
  return std::string();
}

std::vector<long long> AmgHierarchy::level_rows() const {
  std::vector<long long> out;
  for (const auto& L : lv_) out.push_back(L.n);
  return out;
}

// --- AmgBatch ----------------------------------------------------------------

AmgBatch::~AmgBatch() { release(); }

void AmgBatch::release() {
  if (d_levels_) (void)hipFree(d_levels_);
  if (d_entries_) (void)hipFree(d_entries_);
  if (d_maps_) (void)hipFree(d_maps_);
  d_levels_ = nullptr;
  d_entries_ = nullptr;
  d_maps_ = nullptr;
  problems_.clear();
  host_levels_.clear();
  entries_.clear();
  tables_ = AmgTables();
}

void AmgBatch::add(std::unique_ptr<AmgHierarchy> h, long long row_offset, long long value_offset) {
  problems_.push_back(Problem{std::move(h), row_offset, value_offset, 0});
}

int AmgBatch::finalize() {
  if (problems_.empty()) return 1;
  // every problem's levels, one array (the entries' next levels)
  host_levels_.clear();
  int depth = 0;
  for (auto& p : problems_) {
    p.base = static_cast<int>(host_levels_.size());
    for (const AmgLevel& L : p.h->lv_) host_levels_.push_back(L);
    depth = std::max(depth, static_cast<int>(p.h->lv_.size()));
  }
  if (hipMalloc(&d_levels_, host_levels_.size() * sizeof(AmgLevel)) != hipSuccess) return 1;
  // the tables, each a run of entries ended by one that holds the block
  // count (problem -1), and a map from each of its blocks to its entry
  entries_.clear();
  std::vector<int> maps;
  struct Pending {
    AmgTable* table;
    std::size_t entries, map;
  };
  std::vector<Pending> pending;
  auto table = [&](AmgTable& t, int level, AmgTableKind kind, bool coarsest, bool all) {
    const std::size_t start = entries_.size(), map0 = maps.size();
    long long blocks = 0;
    for (int p = 0; p < static_cast<int>(problems_.size()); ++p) {
      const Problem& q = problems_[p];
      const int count = static_cast<int>(q.h->lv_.size());
      if (!all && (level >= count || (level == count - 1) != coarsest)) continue;
      AmgGroupEntry e;
      e.L = q.h->lv_[level];
      e.N = level + 1 < count ? d_levels_ + q.base + level + 1 : nullptr;
      e.offset = level == 0 ? q.row_offset : -1;
      e.first = blocks;
      e.problem = p;
      e.level = level;
      if (kind == kAmgTableSymmetric) {
        e.mirror = q.h->transpose_;
        e.dst = q.h->symmetric_;
        e.count = q.h->nnz0_;
      }
      const long long nb = amg_table_blocks(kind, q.h->lv_[level], q.h->nnz0_);
      maps.insert(maps.end(), static_cast<std::size_t>(nb), static_cast<int>(entries_.size() - start));
      blocks += nb;
      entries_.push_back(e);
    }
    AmgGroupEntry end;
    end.first = blocks;
    end.problem = -1;
    entries_.push_back(end);
    t.count = static_cast<int>(entries_.size() - start) - 1;
    t.blocks = t.count > 0 ? blocks : 0;
    pending.push_back(Pending{&t, start, map0});
  };
  tables_.levels.assign(static_cast<std::size_t>(depth), AmgLevelTables());
  for (int l = 0; l < depth; ++l) {
    AmgLevelTables& t = tables_.levels[l];
    table(t.elem, l, kAmgTableElem, false, false);
    table(t.rows, l, kAmgTableRows, false, false);
    table(t.restrict_, l, kAmgTableRestrict, false, false);
    table(t.power, l, kAmgTablePower, false, false);
    table(t.one, l, kAmgTableOne, false, false);
    table(t.product, l, kAmgTableProduct, false, false);
    table(t.coarse_product, l, kAmgTableCoarseProduct, false, false);
    table(t.gather, l, kAmgTableGather, false, false);
    table(t.dense_one, l, kAmgTableOne, true, false);
    table(t.dense, l, kAmgTableElem, true, false);
  }
  table(tables_.symmetric, 0, kAmgTableSymmetric, false, true);  // level 0 of every problem
  if (hipMalloc(&d_entries_, entries_.size() * sizeof(AmgGroupEntry)) != hipSuccess ||
      hipMalloc(&d_maps_, std::max<std::size_t>(maps.size(), 1) * sizeof(int)) != hipSuccess ||
      (!maps.empty() && hipMemcpy(d_maps_, maps.data(), maps.size() * sizeof(int), hipMemcpyHostToDevice) != hipSuccess))
    return 1;
  for (const Pending& q : pending) {
    q.table->entries = d_entries_ + q.entries;
    q.table->map = d_maps_ + q.map;
  }
  return 0;
}

int AmgBatch::setup(const double* block_values, bool symmetrise, hipStream_t s) {
  if (problems_.empty() || !d_levels_ || !d_entries_) return 1;
  // level 0 reads the block's values at each problem's offset, or its
  // symmetric part (which the symmetric-part table forms from them)
  for (const Problem& q : problems_)
    host_levels_[static_cast<std::size_t>(q.base)].values =
        symmetrise ? q.h->symmetric_ : block_values + q.value_offset;
  for (AmgGroupEntry& e : entries_) {
    if (e.problem < 0 || e.level != 0) continue;
    const Problem& q = problems_[static_cast<std::size_t>(e.problem)];
    e.L.values = host_levels_[static_cast<std::size_t>(q.base)].values;
    if (e.mirror) e.src = block_values + q.value_offset;
  }
  if (hipMemcpyAsync(d_levels_, host_levels_.data(), host_levels_.size() * sizeof(AmgLevel), hipMemcpyHostToDevice,
                     s) != hipSuccess ||
      hipMemcpyAsync(d_entries_, entries_.data(), entries_.size() * sizeof(AmgGroupEntry), hipMemcpyHostToDevice,
                     s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess)
    return 1;
  return launch_amg_setup(tables_, symmetrise, s);
}

std::vector<int> AmgBatch::singular(hipStream_t s) const {
  std::vector<double> flags(problems_.size(), 1.0);
  bool ok = true;
  for (std::size_t p = 0; p < problems_.size() && ok; ++p) {
    const AmgHierarchy& h = *problems_[p].h;
    if (h.lv_.empty()) continue;
    ok = hipMemcpyAsync(&flags[p], h.lv_.back().coef + kAmgSingular, sizeof(double), hipMemcpyDeviceToHost, s) ==
         hipSuccess;
  }
  ok = ok && hipStreamSynchronize(s) == hipSuccess;
  std::vector<int> out(problems_.size(), 1);
  if (!ok) return out;
  for (std::size_t p = 0; p < problems_.size(); ++p) out[p] = flags[p] != 0.0;
  return out;
}

int AmgBatch::apply(const double* b, double* x, hipStream_t s) const {
  if (problems_.empty()) return 1;
  return launch_amg_vcycle(tables_, active_, b, x, s);
}

std::vector<long long> AmgBatch::level_rows() const {
  return problems_.empty() ? std::vector<long long>() : problems_[0].h->level_rows();
}

std::vector<long long> AmgBatch::level_nnz() const {
  return problems_.empty() ? std::vector<long long>() : problems_[0].h->level_nnz();
}

int amg_apply(const void* batch, const double* b, double* x, hipStream_t s) {
  return static_cast<const AmgBatch*>(batch)->apply(b, x, s);
}

}  // namespace sbk
