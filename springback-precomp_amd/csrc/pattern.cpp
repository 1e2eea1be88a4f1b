// Host-side sparsity pattern and gather maps (see assembly.hpp).
#include "assembly.hpp"

#include <algorithm>
#include <string>

namespace sbk {

bool build_pattern(int64_t nn, int64_t ne, const int32_t* conn, HostPattern& p, std::string* why) {
  constexpr int npe = 8;
  if (nn < 0 || ne < 0 || nn > (1ll << 29) || ne * 64 > (1ll << 31) - 1) {
    if (why) *why = "mesh sizes out of range";
    return false;
  }
  for (int64_t e = 0; e < ne; ++e) {
    const int32_t* nodes = conn + e * npe;
    for (int a = 0; a < npe; ++a) {
      if (nodes[a] < 0 || nodes[a] >= nn) {
        if (why) *why = "element " + std::to_string(e) + " names node " + std::to_string(nodes[a]) +
                        ", outside [0, " + std::to_string(nn) + ")";
        return false;
      }
      for (int b = 0; b < a; ++b)
        if (nodes[b] == nodes[a]) {
          if (why) *why = "element " + std::to_string(e) + " names node " + std::to_string(nodes[a]) + " twice";
          return false;
        }
    }
  }
  p.num_nodes = nn;
  p.num_elements = ne;
  // node incidences, ascending element order
  p.node_ptr.assign(static_cast<std::size_t>(nn) + 1, 0);
  for (int64_t i = 0; i < ne * npe; ++i) ++p.node_ptr[static_cast<std::size_t>(conn[i]) + 1];
  for (int64_t n = 0; n < nn; ++n) p.node_ptr[n + 1] += p.node_ptr[n];
  p.node_incidence.resize(static_cast<std::size_t>(ne * npe));
  {
    std::vector<int32_t> fill(p.node_ptr.begin(), p.node_ptr.end() - 1);
    for (int64_t e = 0; e < ne; ++e)
      for (int a = 0; a < npe; ++a) p.node_incidence[fill[conn[e * npe + a]]++] = static_cast<int32_t>(e * npe + a);
  }
  // sorted neighbourhoods (the node itself included)
  std::vector<int32_t> nbr_ptr(static_cast<std::size_t>(nn) + 1, 0), nbr;
  nbr.reserve(static_cast<std::size_t>(nn) * 27);
  {
    std::vector<int32_t> mark(static_cast<std::size_t>(nn), -1), list;
    for (int64_t n = 0; n < nn; ++n) {
      list.clear();
      for (int32_t k = p.node_ptr[n]; k < p.node_ptr[n + 1]; ++k) {
        const int32_t* nodes = conn + (p.node_incidence[k] / npe) * npe;
        for (int a = 0; a < npe; ++a)
          if (mark[nodes[a]] != n) {
            mark[nodes[a]] = static_cast<int32_t>(n);
            list.push_back(nodes[a]);
          }
      }
      std::sort(list.begin(), list.end());
      nbr.insert(nbr.end(), list.begin(), list.end());
      nbr_ptr[n + 1] = static_cast<int32_t>(nbr.size());
    }
  }
  const int64_t nd = 3 * nn;
  p.row_ptr.assign(static_cast<std::size_t>(nd) + 1, 0);
  for (int64_t a = 0; a < nn; ++a) {
    const int64_t deg = nbr_ptr[a + 1] - nbr_ptr[a];
    for (int ka = 0; ka < 3; ++ka) p.row_ptr[3 * a + ka + 1] = static_cast<int32_t>(p.row_ptr[3 * a + ka] + 3 * deg);
  }
  p.col_idx.resize(static_cast<std::size_t>(p.row_ptr.back()));
  for (int64_t a = 0; a < nn; ++a)
    for (int ka = 0; ka < 3; ++ka) {
      int32_t q = p.row_ptr[3 * a + ka];
      for (int32_t k = nbr_ptr[a]; k < nbr_ptr[a + 1]; ++k)
        for (int kb = 0; kb < 3; ++kb) p.col_idx[q++] = 3 * nbr[k] + kb;
    }
  // blocks in row-major order
  const std::size_t nblocks = nbr.size();
  p.block_row_node.resize(nblocks);
  p.block_pos.resize(nblocks);
  p.block_row_len.resize(nblocks);
  for (int64_t a = 0; a < nn; ++a) {
    const int32_t deg = nbr_ptr[a + 1] - nbr_ptr[a];
    for (int32_t k = 0; k < deg; ++k) {
      const std::size_t i = static_cast<std::size_t>(nbr_ptr[a] + k);
      p.block_row_node[i] = static_cast<int32_t>(a);
      p.block_pos[i] = p.row_ptr[3 * a] + 3 * k;
      p.block_row_len[i] = 3 * deg;
    }
  }
  auto block_of = [&](int32_t a, int32_t b) {
    const int32_t* begin = nbr.data() + nbr_ptr[a];
    const int32_t* end = nbr.data() + nbr_ptr[a + 1];
    return static_cast<int32_t>(std::lower_bound(begin, end, b) - nbr.data());
  };
  // contributions per block, ascending element order
  p.block_ptr.assign(nblocks + 1, 0);
  std::vector<int32_t> eb(static_cast<std::size_t>(ne * 64));
  for (int64_t e = 0; e < ne; ++e) {
    const int32_t* nodes = conn + e * npe;
    for (int a = 0; a < npe; ++a)
      for (int b = 0; b < npe; ++b) {
        const int32_t blk = block_of(nodes[a], nodes[b]);
        eb[static_cast<std::size_t>(e * 64 + a * 8 + b)] = blk;
        ++p.block_ptr[static_cast<std::size_t>(blk) + 1];
      }
  }
  for (std::size_t i = 0; i < nblocks; ++i) p.block_ptr[i + 1] += p.block_ptr[i];
  p.block_contrib.resize(static_cast<std::size_t>(ne * 64));
  {
    std::vector<int32_t> fill(p.block_ptr.begin(), p.block_ptr.end() - 1);
    for (int64_t e = 0; e < ne; ++e)
      for (int ab = 0; ab < 64; ++ab)
        p.block_contrib[fill[eb[static_cast<std::size_t>(e * 64 + ab)]]++] = static_cast<int32_t>(e * 64 + ab);
  }
  // node-wise assembly map (positions fit a byte: neighbourhoods of at most
  // 255 nodes, checked)
  p.max_degree = 0;
  for (int64_t n = 0; n < nn; ++n) p.max_degree = std::max(p.max_degree, nbr_ptr[n + 1] - nbr_ptr[n]);
  if (p.max_degree > 255) {
    if (why) *why = "a node has more than 255 neighbours";
    return false;
  }
  p.inc_pos.resize(static_cast<std::size_t>(ne * npe));
  for (int64_t n = 0; n < nn; ++n)
    for (int32_t i = p.node_ptr[n]; i < p.node_ptr[n + 1]; ++i) {
      const int32_t* nodes = conn + (p.node_incidence[i] / npe) * npe;
      uint64_t packed = 0;
      for (int m = 0; m < npe; ++m) {
        const uint64_t pos = static_cast<uint64_t>(block_of(static_cast<int32_t>(n), nodes[m]) - nbr_ptr[n]);
        packed |= pos << (8 * m);
      }
      p.inc_pos[static_cast<std::size_t>(i)] = packed;
    }
  p.block_col_node.assign(nbr.begin(), nbr.end());
  p.block_mirror.resize(nblocks);
  for (int64_t a = 0; a < nn; ++a)
    for (int32_t k = nbr_ptr[a]; k < nbr_ptr[a + 1]; ++k) {
      const int32_t b = nbr[static_cast<std::size_t>(k)];
      p.block_mirror[static_cast<std::size_t>(k)] = block_of(b, static_cast<int32_t>(a)) - nbr_ptr[b];
    }
  p.diag_block.resize(static_cast<std::size_t>(nn));
  for (int64_t n = 0; n < nn; ++n) p.diag_block[n] = block_of(static_cast<int32_t>(n), static_cast<int32_t>(n));
  // the reference's block offsets: position of a within the neighbourhood of b, times 3
  p.block_offset.resize(static_cast<std::size_t>(ne * 64));
  for (int64_t e = 0; e < ne; ++e) {
    const int32_t* nodes = conn + e * npe;
    for (int a = 0; a < npe; ++a)
      for (int b = 0; b < npe; ++b)
        p.block_offset[static_cast<std::size_t>(e * 64 + a * 8 + b)] =
            3 * (block_of(nodes[b], nodes[a]) - nbr_ptr[nodes[b]]);
  }
  return true;
}

}  // namespace sbk
