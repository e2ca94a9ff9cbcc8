// Host pattern builder: node graph, scalar CSR pattern, sorted contribution map and the row-block
// schedule of the fused assembly.
//
// The scalar CSR pattern is the one scipy produces from the reference's COO lists
// (finite_elements/analysis.py:714-735 build the (row, col) pairs, 661 converts them): every
// (row, col) pair that any element touches, columns sorted, explicit zeros kept.  Because every
// element couples all dof components of all its nodes, that pattern is the node adjacency graph
// expanded by dof x dof blocks; it is built here from the node graph directly.
#include <algorithm>
#include <cstdint>
#include <cstring>
#include <new>
#include <vector>

#include "fe_hip.h"

int fe_set_error(int code, const char* fmt, ...);

struct fe_pattern {
  int npe = 0, dof = 0, sym = 0;
  int64_t n_elem = 0, n_nodes = 0;
  int64_t nnz = 0, nnzb = 0, ninc = 0, ncontrib = 0;
  int32_t max_inc = 0;
  std::vector<int32_t> bptr, bcol, inc_ptr, inc, bseg_ptr, blk;
  std::vector<uint16_t> bseg;
  std::vector<int32_t> rowptr, colind, gather, seg_ptr;
  // per incidence: block position (within the node's row) of each vertex of the element
  std::vector<int32_t> inc_qpos;
};

namespace {

int build(fe_pattern& p, const int32_t* conn, int with_gather) {
  const int npe = p.npe, dof = p.dof;
  const int64_t E = p.n_elem, N = p.n_nodes;
  // Incidences, ascending element per node.
  std::vector<int64_t> cnt(N + 1, 0);
  for (int a = 0; a < npe; ++a)
    for (int64_t e = 0; e < E; ++e) {
      const int32_t n = conn[a * E + e];
      if (n < 0 || n >= N) return fe_set_error(FE_ERR_ARG, "fe_pattern_create: node index %d out of range", n);
      cnt[n + 1]++;
    }
  p.ninc = (int64_t)npe * E;
  if (p.ninc >= (int64_t(1) << 31))
    return fe_set_error(FE_ERR_ARG, "fe_pattern_create: too many incidences for int32");
  p.inc_ptr.assign(N + 1, 0);
  for (int64_t i = 0; i < N; ++i) p.inc_ptr[i + 1] = (int32_t)(p.inc_ptr[i] + cnt[i + 1]);
  p.inc.assign(p.ninc, 0);
  std::vector<int32_t> fill(p.inc_ptr.begin(), p.inc_ptr.end() - 1);
  for (int64_t e = 0; e < E; ++e)
    for (int a = 0; a < npe; ++a) {
      const int32_t n = conn[a * E + e];
      p.inc[fill[n]++] = (int32_t)(e * npe + a);
    }
  // Node adjacency (sorted, self included for every node that has an element).
  p.bptr.assign(N + 1, 0);
  std::vector<int32_t> nb;
  std::vector<int32_t> bcol;
  bcol.reserve((size_t)p.ninc * 4);
  for (int64_t i = 0; i < N; ++i) {
    nb.clear();
    for (int32_t k = p.inc_ptr[i]; k < p.inc_ptr[i + 1]; ++k) {
      const int64_t e = p.inc[k] / npe;
      for (int b = 0; b < npe; ++b) nb.push_back(conn[b * E + e]);
    }
    std::sort(nb.begin(), nb.end());
    nb.erase(std::unique(nb.begin(), nb.end()), nb.end());
    bcol.insert(bcol.end(), nb.begin(), nb.end());
    if ((int64_t)bcol.size() >= (int64_t(1) << 31))
      return fe_set_error(FE_ERR_ARG, "fe_pattern_create: node graph too large for int32");
    p.bptr[i + 1] = (int32_t)bcol.size();
  }
  p.bcol.swap(bcol);
  p.nnzb = p.bptr[N];
  p.nnz = p.nnzb * dof * dof;
  if (p.nnz >= (int64_t(1) << 31))
    return fe_set_error(FE_ERR_ARG, "fe_pattern_create: nnz %lld too large for int32 indices",
                        (long long)p.nnz);
  // Block positions of each incidence's vertices and the block-level contribution map.
  p.inc_qpos.assign((size_t)p.ninc * npe, 0);
  std::vector<int32_t> bcnt(p.nnzb + 1, 0);
  for (int64_t i = 0; i < N; ++i) {
    const int32_t* row = p.bcol.data() + p.bptr[i];
    const int32_t deg = p.bptr[i + 1] - p.bptr[i];
    for (int32_t k = p.inc_ptr[i]; k < p.inc_ptr[i + 1]; ++k) {
      const int64_t e = p.inc[k] / npe;
      for (int b = 0; b < npe; ++b) {
        const int32_t nbn = conn[b * E + e];
        const int32_t q = (int32_t)(std::lower_bound(row, row + deg, nbn) - row);
        p.inc_qpos[(size_t)k * npe + b] = q;
        bcnt[p.bptr[i] + q + 1]++;
      }
    }
  }
  p.bseg_ptr.assign(p.nnzb + 1, 0);
  for (int64_t q = 0; q < p.nnzb; ++q) p.bseg_ptr[q + 1] = p.bseg_ptr[q] + bcnt[q + 1];
  p.bseg.assign((size_t)p.ninc * npe, 0);
  {
    std::vector<int32_t> bf(p.bseg_ptr.begin(), p.bseg_ptr.end() - 1);
    for (int64_t i = 0; i < N; ++i) {
      for (int32_t k = p.inc_ptr[i]; k < p.inc_ptr[i + 1]; ++k) {
        const int32_t kl = k - p.inc_ptr[i];
        if (kl >= (1 << 14))
          return fe_set_error(FE_ERR_ARG, "fe_pattern_create: node %lld has more than 16383 elements",
                              (long long)i);
        for (int b = 0; b < npe; ++b) {
          const int32_t q = p.bptr[i] + p.inc_qpos[(size_t)k * npe + b];
          p.bseg[bf[q]++] = (uint16_t)((kl << 2) | b);
        }
      }
    }
  }
  // Scalar CSR pattern: row (i, di) holds dof entries per neighbour block, in neighbour order.
  const int64_t nrows = N * dof;
  p.rowptr.assign(nrows + 1, 0);
  p.colind.assign(p.nnz, 0);
  for (int64_t i = 0; i < N; ++i) {
    const int32_t deg = p.bptr[i + 1] - p.bptr[i];
    for (int di = 0; di < dof; ++di) {
      const int64_t r = i * dof + di;
      const int64_t start = (int64_t)p.bptr[i] * dof * dof + (int64_t)di * deg * dof;
      p.rowptr[r] = (int32_t)start;
      int64_t o = start;
      for (int32_t q = p.bptr[i]; q < p.bptr[i + 1]; ++q)
        for (int dj = 0; dj < dof; ++dj) p.colind[o++] = p.bcol[q] * dof + dj;
    }
  }
  p.rowptr[nrows] = (int32_t)p.nnz;
  // Contribution map of the two-phase assembly: counting sort of all element-matrix entries by
  // destination, stable in (element, row, column) order.
  if (with_gather) {
    const int m = npe * dof;
    p.ncontrib = (int64_t)m * m * E;
    const int64_t mstored = p.sym ? (int64_t)m * (m + 1) / 2 : (int64_t)m * m;
    if (mstored * E >= (int64_t(1) << 31))
      return fe_set_error(FE_ERR_ARG, "fe_pattern_create: element-matrix array too large for int32");
    // incidence index of (e, a)
    std::vector<int32_t> inc_of((size_t)E * npe);
    for (int64_t i = 0; i < N; ++i)
      for (int32_t k = p.inc_ptr[i]; k < p.inc_ptr[i + 1]; ++k) inc_of[p.inc[k]] = k;
    std::vector<int32_t> segcnt(p.nnz + 1, 0);
    auto dest = [&](int64_t e, int r, int c) -> int64_t {
      const int a = r / dof, dr = r % dof, b = c / dof, dc = c % dof;
      const int32_t node_a = conn[a * E + e];
      const int32_t k = inc_of[e * npe + a];
      return (int64_t)p.rowptr[(int64_t)node_a * dof + dr] +
             (int64_t)p.inc_qpos[(size_t)k * npe + b] * dof + dc;
    };
    for (int64_t e = 0; e < E; ++e)
      for (int r = 0; r < m; ++r)
        for (int c = 0; c < m; ++c) segcnt[dest(e, r, c) + 1]++;
    p.seg_ptr.assign(p.nnz + 1, 0);
    for (int64_t q = 0; q < p.nnz; ++q) p.seg_ptr[q + 1] = p.seg_ptr[q] + segcnt[q + 1];
    std::vector<int32_t>().swap(segcnt);
    p.gather.assign(p.ncontrib, 0);
    std::vector<int32_t> f2(p.seg_ptr.begin(), p.seg_ptr.end() - 1);
    for (int64_t e = 0; e < E; ++e)
      for (int r = 0; r < m; ++r)
        for (int c = 0; c < m; ++c) {
          const int64_t k = p.sym ? (int64_t)(r <= c ? r * m - r * (r - 1) / 2 + (c - r)
                                                     : c * m - c * (c - 1) / 2 + (r - c))
                                  : (int64_t)r * m + c;
          p.gather[f2[dest(e, r, c)]++] = (int32_t)(k * E + e);
        }
  }
  return FE_OK;
}

}  // namespace

// element slots of one workgroup: 14 bits of the packed (slot << 2 | local vertex) incidence code
constexpr size_t kMaxElemSlots = 16383;

// Greedy row-block schedule: consecutive nodes while a range holds at most max_inc incidences and
// max_nodes nodes.  Appends the range starts and the end to blk; returns the largest range's
// incidence count in *biggest.
static int greedy_ranges(const int32_t* inc_ptr, int64_t n_nodes, int max_inc_per_block, int max_nodes,
                         std::vector<int32_t>& blk, int32_t* biggest, const char* who) {
  blk.clear();
  blk.push_back(0);
  int32_t cur_inc = 0, cur_nodes = 0, mx = 0;
  for (int64_t i = 0; i < n_nodes; ++i) {
    const int32_t ni = inc_ptr[i + 1] - inc_ptr[i];
    if (ni > max_inc_per_block)
      return fe_set_error(FE_ERR_ARG, "%s: node %lld has %d incidences > max_inc %d", who, (long long)i,
                          ni, max_inc_per_block);
    if (cur_nodes > 0 && (cur_inc + ni > max_inc_per_block || cur_nodes >= max_nodes)) {
      blk.push_back((int32_t)i);
      mx = std::max(mx, cur_inc);
      cur_inc = 0;
      cur_nodes = 0;
    }
    cur_inc += ni;
    cur_nodes++;
  }
  if (cur_nodes > 0) {
    blk.push_back((int32_t)n_nodes);
    mx = std::max(mx, cur_inc);
  }
  *biggest = mx;
  return FE_OK;
}

extern "C" int fe_pattern_schedule(fe_pattern* p, int max_inc_per_block) {
  if (!p) return fe_set_error(FE_ERR_ARG, "fe_pattern_schedule: null pattern");
  if (max_inc_per_block < 1 || max_inc_per_block > 4096)
    return fe_set_error(FE_ERR_ARG, "fe_pattern_schedule: max_inc_per_block out of range");
  int32_t mx = 0;
  const int rc = greedy_ranges(p->inc_ptr.data(), p->n_nodes, max_inc_per_block, 256, p->blk, &mx,
                               "fe_pattern_schedule");
  if (rc != FE_OK) return rc;
  p->max_inc = mx;
  return FE_OK;
}

extern "C" int fe_schedule_ranges(const int32_t* inc_ptr, int64_t n_nodes, int max_inc, int max_nodes,
                                  int32_t* blk, int64_t* n_blk, int32_t* biggest) {
  if (!inc_ptr || !blk || !n_blk || !biggest) return fe_set_error(FE_ERR_ARG, "fe_schedule_ranges: null argument");
  if (max_inc < 1 || max_nodes < 1) return fe_set_error(FE_ERR_ARG, "fe_schedule_ranges: bad limits");
  std::vector<int32_t> b;
  const int rc = greedy_ranges(inc_ptr, n_nodes, max_inc, max_nodes, b, biggest, "fe_schedule_ranges");
  if (rc != FE_OK) return rc;
  std::copy(b.begin(), b.end(), blk);
  *n_blk = (int64_t)b.size() - 1;
  return FE_OK;
}

extern "C" int fe_schedule_elements(const int32_t* inc_ptr, const int32_t* inc, int npe,
                                    int64_t n_nodes, const int32_t* blk, int64_t n_blk,
                                    int32_t* welem_ptr, int32_t* welem, uint16_t* inc_sa,
                                    int32_t* max_elem) {
  if (!inc_ptr || !blk || !welem_ptr || !max_elem || (n_nodes > 0 && (!inc || !welem || !inc_sa)))
    return fe_set_error(FE_ERR_ARG, "fe_schedule_elements: null argument");
  if (npe < 2 || npe > 4) return fe_set_error(FE_ERR_ARG, "fe_schedule_elements: npe must be 2..4");
  if (n_blk < 0 || blk[0] != 0 || blk[n_blk] != n_nodes)
    return fe_set_error(FE_ERR_ARG, "fe_schedule_elements: schedule must cover nodes 0..n_nodes");
  std::vector<int32_t> ids;
  int64_t out = 0;
  int32_t mx = 0;
  welem_ptr[0] = 0;
  for (int64_t w = 0; w < n_blk; ++w) {
    if (blk[w + 1] < blk[w])
      return fe_set_error(FE_ERR_ARG, "fe_schedule_elements: schedule offsets must not decrease");
    const int32_t i0 = inc_ptr[blk[w]], i1 = inc_ptr[blk[w + 1]];
    ids.resize((size_t)(i1 - i0));
    for (int32_t t = i0; t < i1; ++t) ids[t - i0] = inc[t] / npe;
    std::sort(ids.begin(), ids.end());
    ids.erase(std::unique(ids.begin(), ids.end()), ids.end());
    if (ids.size() > kMaxElemSlots)
      return fe_set_error(FE_ERR_ARG, "fe_schedule_elements: %zu elements in one workgroup > %d",
                          ids.size(), (int)kMaxElemSlots);
    std::copy(ids.begin(), ids.end(), welem + out);
    for (int32_t t = i0; t < i1; ++t) {
      const int32_t e = inc[t] / npe, a = inc[t] - e * npe;
      const int32_t slot = (int32_t)(std::lower_bound(ids.begin(), ids.end(), e) - ids.begin());
      inc_sa[t] = (uint16_t)((slot << 2) | a);
    }
    out += (int64_t)ids.size();
    welem_ptr[w + 1] = (int32_t)out;
    mx = std::max(mx, (int32_t)ids.size());
  }
  *max_elem = mx;
  return FE_OK;
}

extern "C" int fe_schedule_nodes(const int32_t* welem_conn, int npe, int64_t n_welem,
                                 const int32_t* welem_ptr, int64_t n_blk, int64_t n_nodes,
                                 int32_t* wnode_ptr, int32_t* wnode, uint16_t* welem_local,
                                 int32_t* max_wnode) {
  if (!welem_ptr || !wnode_ptr || !max_wnode || (n_welem > 0 && (!welem_conn || !wnode || !welem_local)))
    return fe_set_error(FE_ERR_ARG, "fe_schedule_nodes: null argument");
  if (npe < 2 || npe > 4) return fe_set_error(FE_ERR_ARG, "fe_schedule_nodes: npe must be 2..4");
  if (n_blk < 0 || welem_ptr[0] != 0 || welem_ptr[n_blk] != n_welem)
    return fe_set_error(FE_ERR_ARG, "fe_schedule_nodes: element lists must cover 0..n_welem");
  std::vector<int32_t> ids;
  int64_t out = 0;
  int32_t mx = 0;
  wnode_ptr[0] = 0;
  for (int64_t w = 0; w < n_blk; ++w) {
    const int32_t e0 = welem_ptr[w], e1 = welem_ptr[w + 1];
    if (e1 < e0) return fe_set_error(FE_ERR_ARG, "fe_schedule_nodes: element list offsets must not decrease");
    ids.clear();
    for (int b = 0; b < npe; ++b)
      for (int32_t e = e0; e < e1; ++e) {
        const int32_t n = welem_conn[(int64_t)b * n_welem + e];
        if (n < 0 || n >= n_nodes) return fe_set_error(FE_ERR_ARG, "fe_schedule_nodes: node id out of range");
        ids.push_back(n);
      }
    std::sort(ids.begin(), ids.end());
    ids.erase(std::unique(ids.begin(), ids.end()), ids.end());
    if (ids.size() > 65535)
      return fe_set_error(FE_ERR_ARG, "fe_schedule_nodes: %zu nodes in one workgroup > 65535", ids.size());
    std::copy(ids.begin(), ids.end(), wnode + out);
    for (int32_t e = e0; e < e1; ++e)
      for (int b = 0; b < 4; ++b) {
        uint16_t loc = 0;
        if (b < npe) {
          const int32_t n = welem_conn[(int64_t)b * n_welem + e];
          loc = (uint16_t)(std::lower_bound(ids.begin(), ids.end(), n) - ids.begin());
        }
        welem_local[(int64_t)e * 4 + b] = loc;
      }
    out += (int64_t)ids.size();
    wnode_ptr[w + 1] = (int32_t)out;
    mx = std::max(mx, (int32_t)ids.size());
  }
  *max_wnode = mx;
  return FE_OK;
}

static inline bool bcol_is_row(const int32_t* bcol, int32_t q, int32_t i) { return bcol[q] == i; }

extern "C" int fe_schedule_codes(const int32_t* bptr, const int32_t* bcol, const int32_t* bseg_ptr,
                                 const uint16_t* bseg, const int32_t* inc_ptr, const uint16_t* inc_sa,
                                 int64_t n_nodes, const int32_t* blk, int64_t n_blk, uint16_t* code,
                                 uint16_t* brow) {
  if (!bptr || !bcol || !bseg_ptr || !inc_ptr || !blk || !code || !brow || (n_nodes > 0 && (!bseg || !inc_sa)))
    return fe_set_error(FE_ERR_ARG, "fe_schedule_codes: null argument");
  if (n_blk < 0 || blk[0] != 0 || blk[n_blk] != n_nodes)
    return fe_set_error(FE_ERR_ARG, "fe_schedule_codes: schedule must cover nodes 0..n_nodes");
  for (int64_t w = 0; w < n_blk; ++w) {
    if (blk[w + 1] < blk[w]) return fe_set_error(FE_ERR_ARG, "fe_schedule_codes: schedule offsets must not decrease");
    if (blk[w + 1] - blk[w] > 32767)
      return fe_set_error(FE_ERR_ARG, "fe_schedule_codes: a range of more than 32767 nodes");
    for (int32_t i = blk[w]; i < blk[w + 1]; ++i)
      for (int32_t q = bptr[i]; q < bptr[i + 1]; ++q) {
        brow[q] = (uint16_t)((i - blk[w]) | (bcol_is_row(bcol, q, i) ? 0x8000 : 0));
        for (int32_t c = bseg_ptr[q]; c < bseg_ptr[q + 1]; ++c) {
          const int32_t sa = inc_sa[inc_ptr[i] + (bseg[c] >> 2)];
          if ((sa >> 2) >= 4096)
            return fe_set_error(FE_ERR_ARG, "fe_schedule_codes: element slot %d >= 4096", sa >> 2);
          code[c] = (uint16_t)(((sa >> 2) << 4) | ((sa & 3) << 2) | (bseg[c] & 3));
        }
      }
  }
  return FE_OK;
}

extern "C" int fe_pattern_create(const int32_t* conn, int npe, int dof, int64_t n_elem,
                                 int64_t n_nodes, int sym, int with_gather, fe_pattern** out) {
  if (!out) return fe_set_error(FE_ERR_ARG, "fe_pattern_create: null output");
  *out = nullptr;
  if (npe < 2 || npe > 4) return fe_set_error(FE_ERR_ARG, "fe_pattern_create: npe must be 2..4");
  if (dof < 1 || dof > 3) return fe_set_error(FE_ERR_ARG, "fe_pattern_create: dof must be 1..3");
  if (n_elem < 0 || n_nodes < 0 || (n_elem > 0 && !conn))
    return fe_set_error(FE_ERR_ARG, "fe_pattern_create: bad sizes or null connectivity");
  if (n_nodes * dof >= (int64_t(1) << 31))
    return fe_set_error(FE_ERR_ARG, "fe_pattern_create: too many rows for int32");
  fe_pattern* p = new (std::nothrow) fe_pattern();
  if (!p) return fe_set_error(FE_ERR_NOMEM, "fe_pattern_create: out of host memory");
  p->npe = npe;
  p->dof = dof;
  p->sym = sym ? 1 : 0;
  p->n_elem = n_elem;
  p->n_nodes = n_nodes;
  int rc;
  try {
    rc = build(*p, conn, with_gather);
    if (rc == FE_OK) {
      // default row-block schedule: 192 (tets) / 128 (triangles) incidences per workgroup, raised to
      // the largest node incidence count when a node has more (dynamic LDS stays under 64 KB for
      // typical meshes: 36 doubles per tet incidence)
      int32_t cap = npe == 4 ? 192 : 128;
      for (int64_t i = 0; i < n_nodes; ++i) cap = std::max(cap, p->inc_ptr[i + 1] - p->inc_ptr[i]);
      rc = fe_pattern_schedule(p, cap);
    }
  } catch (const std::bad_alloc&) {
    rc = fe_set_error(FE_ERR_NOMEM, "fe_pattern_create: out of host memory");
  }
  if (rc != FE_OK) {
    delete p;
    return rc;
  }
  *out = p;
  return FE_OK;
}

extern "C" int fe_pattern_info(const fe_pattern* p, int64_t* info) {
  if (!p || !info) return fe_set_error(FE_ERR_ARG, "fe_pattern_info: null argument");
  for (int k = 0; k < FE_PAT_NINFO; ++k) info[k] = 0;
  info[FE_PAT_NNZ] = p->nnz;
  info[FE_PAT_NNZB] = p->nnzb;
  info[FE_PAT_NINC] = p->ninc;
  info[FE_PAT_NCONTRIB] = p->gather.empty() ? 0 : p->ncontrib;
  info[FE_PAT_NBLK] = (int64_t)p->blk.size() - 1;
  info[FE_PAT_MAXINC] = p->max_inc;
  info[FE_PAT_NROWS] = p->n_nodes * p->dof;
  return FE_OK;
}

extern "C" int fe_pattern_copy(const fe_pattern* p, int which, void* dst) {
  if (!p || !dst) return fe_set_error(FE_ERR_ARG, "fe_pattern_copy: null argument");
  const void* src = nullptr;
  size_t bytes = 0;
  switch (which) {
    case FE_ARR_ROWPTR: src = p->rowptr.data(); bytes = p->rowptr.size() * 4; break;
    case FE_ARR_COLIND: src = p->colind.data(); bytes = p->colind.size() * 4; break;
    case FE_ARR_GATHER: src = p->gather.data(); bytes = p->gather.size() * 4; break;
    case FE_ARR_SEGPTR: src = p->seg_ptr.data(); bytes = p->seg_ptr.size() * 4; break;
    case FE_ARR_BPTR: src = p->bptr.data(); bytes = p->bptr.size() * 4; break;
    case FE_ARR_BCOL: src = p->bcol.data(); bytes = p->bcol.size() * 4; break;
    case FE_ARR_INCPTR: src = p->inc_ptr.data(); bytes = p->inc_ptr.size() * 4; break;
    case FE_ARR_INC: src = p->inc.data(); bytes = p->inc.size() * 4; break;
    case FE_ARR_BSEGPTR: src = p->bseg_ptr.data(); bytes = p->bseg_ptr.size() * 4; break;
    case FE_ARR_BSEG: src = p->bseg.data(); bytes = p->bseg.size() * 2; break;
    case FE_ARR_BLK: src = p->blk.data(); bytes = p->blk.size() * 4; break;
    default: return fe_set_error(FE_ERR_ARG, "fe_pattern_copy: unknown array %d", which);
  }
  if (bytes && src) std::memcpy(dst, src, bytes);
  return FE_OK;
}

extern "C" int fe_pattern_destroy(fe_pattern* p) {
  delete p;
  return FE_OK;
}
