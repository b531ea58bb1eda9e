// pm_structure.cpp -- once-per-run structure build on the host: hexa20
// promotion, mid-edge restraints, node graph, RCM ordering, equation
// numbering, element colouring and the node-block matrix pattern.
//
// Follows the reference's definitions so that node ids, equation numbers and
// the CSRC pattern come out identical:
//   promotion      mk_elconn_quad_v1   Adjacency.f:583-693 (edges :974-1004)
//   restraints     mk_bound_quad       Rdata.f:2337-2405
//   node graph     graph_v1            Graph.f:72-133 (sorted, duplicate free)
//   ordering       genrcm_v1           Reord.f:45-437 (SPARSPAK GENRCM)
//   numbering      numeqpmec2          Numeq.f:86-151
//   lower pattern  csria_v1/csrja_v1   Csr.f:631-867 (derived from the blocks)
//   colouring      colormsh            Colormsh.f:160-254 (vertex conflicts)
#include <algorithm>
#include <cstring>
#include <thread>

#include "pm_internal.h"

template <class F>
static void par_for(int64_t n, F &&fn) {
  unsigned hw = std::thread::hardware_concurrency();
  int nt = (int)std::min<unsigned>(hw ? hw : 1, 16);
  if (n < 4096 || nt <= 1) {
    fn((int64_t)0, n);
    return;
  }
  std::vector<std::thread> th;
  int64_t chunk = (n + nt - 1) / nt;
  for (int t = 0; t < nt; t++) {
    int64_t a = t * chunk, b = std::min<int64_t>(n, a + chunk);
    if (a >= b) break;
    th.emplace_back([&fn, a, b] { fn(a, b); });
  }
  for (auto &t : th) t.join();
}

static const int kEdge[12][3] = {{0, 1, 8},  {1, 2, 9},  {2, 3, 10}, {3, 0, 11},
                                 {4, 5, 12}, {5, 6, 13}, {6, 7, 14}, {7, 4, 15},
                                 {0, 4, 16}, {1, 5, 17}, {2, 6, 18}, {3, 7, 19}};

// vertex -> incident elements, in element order (0-based CSR)
static void vertex_incidence(const std::vector<int> &ix20, int numel, int nnodev,
                             std::vector<int64_t> &ptr, std::vector<int> &elems) {
  ptr.assign((size_t)nnodev + 1, 0);
  for (int e = 0; e < numel; e++)
    for (int k = 0; k < 8; k++) ptr[ix20[(size_t)e * 21 + k]]++;  // count at [v]
  for (int n = 0; n < nnodev; n++) ptr[n + 1] += ptr[n];
  std::vector<int64_t> fill(ptr.begin(), ptr.end() - 1);
  elems.assign((size_t)ptr[nnodev], 0);
  for (int e = 0; e < numel; e++)
    for (int k = 0; k < 8; k++) elems[fill[ix20[(size_t)e * 21 + k] - 1]++] = e;
}

int pm_host_promote(PmHostMesh &m, const int *ix8) {
  const int numel = m.numel, nnodev = m.nnodev;
  m.ix20.assign((size_t)numel * 21, 0);
  for (int e = 0; e < numel; e++) {
    for (int k = 0; k < 8; k++) {
      int v = ix8[(size_t)e * 21 + k];
      if (v < 1 || v > nnodev) return PM_ERR_ARG;
      m.ix20[(size_t)e * 21 + k] = v;
    }
    m.ix20[(size_t)e * 21 + 20] = ix8[(size_t)e * 21 + 20];
  }
  std::vector<int64_t> ptr;
  std::vector<int> inc;
  vertex_incidence(m.ix20, numel, nnodev, ptr, inc);
  int nno = nnodev;
  int *ix = m.ix20.data();
  for (int i = 0; i < numel; i++) {
    int *row = ix + (size_t)i * 21;
    for (int j = 8; j < 20; j++)
      if (row[j] == 0) row[j] = ++nno;
    for (int j = 0; j < 12; j++) {
      int a = row[kEdge[j][0]], b = row[kEdge[j][1]], mid = row[kEdge[j][2]];
      for (int64_t k = ptr[a - 1]; k < ptr[a]; k++) {
        int nel = inc[k];
        if (nel == i) continue;
        int *nrow = ix + (size_t)nel * 21;
        for (int l = 0; l < 12; l++) {
          int va = nrow[kEdge[l][0]], vb = nrow[kEdge[l][1]];
          if (((a == va) && (b == vb)) || ((a == vb) && (b == va))) {
            if (nrow[kEdge[l][2]] == 0) nrow[kEdge[l][2]] = mid;
          }
        }
      }
    }
  }
  m.nnode = nno;
  m.fnno.assign((size_t)nno, 0);
  for (int n = 0; n < nnodev; n++) m.fnno[n] = 1;
  return PM_OK;
}

int pm_host_plain_hexa8(PmHostMesh &m, const int *ix9) {
  const int numel = m.numel, nnodev = m.nnodev;
  m.nen = 8;
  m.ix20.assign((size_t)numel * 21, 0);
  for (int e = 0; e < numel; e++) {
    for (int k = 0; k < 8; k++) {
      const int v = ix9[(size_t)e * 9 + k];
      if (v < 1 || v > nnodev) return PM_ERR_ARG;
      m.ix20[(size_t)e * 21 + k] = v;
    }
    m.ix20[(size_t)e * 21 + 20] = ix9[(size_t)e * 9 + 8];
  }
  m.nnode = nnodev;
  m.fnno.assign((size_t)nnodev, 1);
  return PM_OK;
}

void pm_host_bound_quad(PmHostMesh &m) {
  int *id = m.idflags.data();
  for (int i = 0; i < m.numel; i++) {
    const int *row = m.ix20.data() + (size_t)i * 21;
    for (int j = 0; j < 12; j++) {
      int a = row[kEdge[j][0]] - 1, b = row[kEdge[j][1]] - 1, mid = row[kEdge[j][2]] - 1;
      for (int k = 0; k < 3; k++)
        if (id[4 * a + k] == 1 && id[4 * b + k] == 1) id[4 * mid + k] = 1;
    }
  }
}

// node graph: neighbours of each node = all other nodes of its elements,
// sorted ascending, without repeats
void pm_host_graph(PmHostMesh &m) {
  const int numel = m.numel, nnode = m.nnode, nen = m.nen;
  // node -> incident elements (all nen nodes)
  std::vector<int64_t> eptr((size_t)nnode + 1, 0);
  for (int e = 0; e < numel; e++)
    for (int k = 0; k < nen; k++) eptr[m.ix20[(size_t)e * 21 + k]]++;
  for (int n = 0; n < nnode; n++) eptr[n + 1] += eptr[n];
  std::vector<int> einc((size_t)eptr[nnode]);
  {
    std::vector<int64_t> fill(eptr.begin(), eptr.end() - 1);
    for (int e = 0; e < numel; e++)
      for (int k = 0; k < nen; k++) einc[fill[m.ix20[(size_t)e * 21 + k] - 1]++] = e;
  }
  std::vector<int> deg((size_t)nnode, 0);
  auto neighbours = [&](int n, std::vector<int> &buf) {
    buf.clear();
    for (int64_t k = eptr[n]; k < eptr[n + 1]; k++) {
      const int *row = m.ix20.data() + (size_t)einc[k] * 21;
      for (int j = 0; j < nen; j++)
        if (row[j] != n + 1) buf.push_back(row[j]);
    }
    std::sort(buf.begin(), buf.end());
    buf.erase(std::unique(buf.begin(), buf.end()), buf.end());
  };
  par_for(nnode, [&](int64_t a, int64_t b) {
    std::vector<int> buf;
    for (int64_t n = a; n < b; n++) {
      neighbours((int)n, buf);
      deg[n] = (int)buf.size();
    }
  });
  m.gptr.assign((size_t)nnode + 1, 0);
  for (int n = 0; n < nnode; n++) m.gptr[n + 1] = m.gptr[n] + deg[n];
  m.gadj.assign((size_t)m.gptr[nnode], 0);
  par_for(nnode, [&](int64_t a, int64_t b) {
    std::vector<int> buf;
    for (int64_t n = a; n < b; n++) {
      neighbours((int)n, buf);
      std::copy(buf.begin(), buf.end(), m.gadj.begin() + m.gptr[n]);
    }
  });
}

// ---- SPARSPAK RCM over the node graph (1-based node ids in gadj) ----
namespace {
struct Rcm {
  const int64_t *ptr;
  const int *adj;
  std::vector<int> mask, xls;
  std::vector<char> seen;
  // level structure rooted at root; returns nlvl; ls 1-based view
  int rootls(int root, int *ls) {
    mask[root] = 0;
    ls[1] = root;
    int nlvl = 0, lvlend = 0, ccsize = 1;
    for (;;) {
      int lbegin = lvlend + 1;
      lvlend = ccsize;
      nlvl++;
      xls[nlvl] = lbegin;
      for (int i = lbegin; i <= lvlend; i++) {
        int node = ls[i];
        for (int64_t j = ptr[node - 1]; j < ptr[node]; j++) {
          int nb = adj[j];
          if (mask[nb] == 0) continue;
          ls[++ccsize] = nb;
          mask[nb] = 0;
        }
      }
      if (ccsize - lvlend <= 0) break;
    }
    xls[nlvl + 1] = lvlend + 1;
    for (int i = 1; i <= ccsize; i++) mask[ls[i]] = 1;
    return nlvl;
  }

  int fnroot(int root, int *ls) {
    int nlvl = rootls(root, ls);
    int ccsize = xls[nlvl + 1] - 1;
    if (nlvl == 1 || nlvl == ccsize) return root;
    for (;;) {
      int jstrt = xls[nlvl], mindeg = ccsize;
      root = ls[jstrt];
      if (ccsize != jstrt) {
        for (int j = jstrt; j <= ccsize; j++) {
          int node = ls[j], ndeg = 0;
          for (int64_t k = ptr[node - 1]; k < ptr[node]; k++)
            if (mask[adj[k]] > 0) ndeg++;
          if (ndeg < mindeg) {
            root = node;
            mindeg = ndeg;
          }
        }
      }
      int nun = rootls(root, ls);
      if (nun <= nlvl) return root;
      nlvl = nun;
      if (nlvl >= ccsize) return root;
    }
  }

  int rcm(int root, int *perm) {
    int *deg = xls.data();
    // degrees within the masked component (degree_v1)
    perm[1] = root;
    seen[root] = 1;
    int lvlend = 0, ccsize = 1;
    for (;;) {
      int lbegin = lvlend + 1;
      lvlend = ccsize;
      for (int i = lbegin; i <= lvlend; i++) {
        int node = perm[i], d = 0;
        for (int64_t j = ptr[node - 1]; j < ptr[node]; j++) {
          int nb = adj[j];
          if (mask[nb] == 0) continue;
          d++;
          if (!seen[nb]) {
            seen[nb] = 1;
            perm[++ccsize] = nb;
          }
        }
        deg[node] = d;
      }
      if (ccsize - lvlend <= 0) break;
    }
    for (int i = 1; i <= ccsize; i++) seen[perm[i]] = 0;
    mask[root] = 0;
    if (ccsize <= 1) return ccsize;
    lvlend = 0;
    int lnbr = 1;
    do {
      int lbegin = lvlend + 1;
      lvlend = lnbr;
      for (int i = lbegin; i <= lvlend; i++) {
        int node = perm[i], fnbr = lnbr + 1;
        for (int64_t j = ptr[node - 1]; j < ptr[node]; j++) {
          int nb = adj[j];
          if (mask[nb] == 0) continue;
          mask[nb] = 0;
          perm[++lnbr] = nb;
        }
        if (fnbr >= lnbr) continue;
        for (int k = fnbr + 1; k <= lnbr; k++) {  // insertion by degree
          int nb = perm[k], l = k - 1;
          while (l >= fnbr && deg[perm[l]] > deg[nb]) {
            perm[l + 1] = perm[l];
            l--;
          }
          perm[l + 1] = nb;
        }
      }
    } while (lnbr > lvlend);
    std::reverse(perm + 1, perm + ccsize + 1);
    return ccsize;
  }
};
}  // namespace

void pm_host_rcm(PmHostMesh &m, int reorder) {
  const int n = m.nnode;
  m.perm.assign((size_t)n, 0);
  if (!reorder) {
    for (int i = 0; i < n; i++) m.perm[i] = i + 1;
    return;
  }
  Rcm r;
  r.ptr = m.gptr.data();
  r.adj = m.gadj.data();
  r.mask.assign((size_t)n + 2, 1);
  r.xls.assign((size_t)n + 2, 0);
  r.seen.assign((size_t)n + 2, 0);
  int *perm = m.perm.data() - 1;  // 1-based view
  int num = 1;
  for (int i = 1; i <= n; i++) {
    if (r.mask[i] == 0) continue;
    int root = r.fnroot(i, perm + num - 1);
    int cc = r.rcm(root, perm + num - 1);
    num += cc;
    if (num > n) break;
  }
}

void pm_host_numeq(PmHostMesh &m) {
  m.id = m.idflags;
  int neq = 0;
  for (int i = 0; i < m.nnode; i++) {
    int n = m.perm[i] - 1;
    int ndfn = m.fnno[n] == 1 ? 4 : 3;
    for (int j = 0; j < ndfn; j++) {
      int flag = m.id[4 * n + j];
      m.id[4 * n + j] = flag == 0 ? ++neq : 0;
    }
  }
  m.neq = neq;
}

// free equations of node n as a contiguous range [first, first + cnt)
static inline void node_eqs(const PmHostMesh &m, int n, int &first, int &cnt) {
  first = -1;
  cnt = 0;
  int nd = m.fnno[n] == 1 ? 4 : 3;
  for (int d = 0; d < nd; d++) {
    int q = m.id[4 * n + d];
    if (q > 0) {
      if (first < 0) first = q - 1;
      cnt++;
    }
  }
}

int pm_host_pattern(const PmHostMesh &m, PmHostPattern &p) {
  const int nnode = m.nnode;
  std::vector<int> first((size_t)nnode), cnt((size_t)nnode);
  for (int n = 0; n < nnode; n++) node_eqs(m, n, first[n], cnt[n]);
  // free nodes in equation order (= RCM position order)
  p.neq = m.neq;
  p.fr.clear();
  p.nr.clear();
  std::vector<int> bnode;
  for (int i = 0; i < nnode; i++) {
    int n = m.perm[i] - 1;
    if (cnt[n] > 0) {
      p.fr.push_back(first[n]);
      p.nr.push_back(cnt[n]);
      bnode.push_back(n);
    }
  }
  p.nrn = (int)bnode.size();
  p.eq_block.assign((size_t)m.neq, 0);
  for (int k = 0; k < p.nrn; k++)
    for (int q = 0; q < p.nr[k]; q++) p.eq_block[p.fr[k] + q] = k;
  std::vector<int64_t> ccount((size_t)p.nrn);
  par_for(p.nrn, [&](int64_t a, int64_t b) {
    for (int64_t k = a; k < b; k++) {
      int n = bnode[k];
      int64_t c = cnt[n];
      for (int64_t j = m.gptr[n]; j < m.gptr[n + 1]; j++) c += cnt[m.gadj[j] - 1];
      ccount[k] = c;
    }
  });
  p.colptr.assign((size_t)p.nrn + 1, 0);
  p.valptr.assign((size_t)p.nrn + 1, 0);
  for (int k = 0; k < p.nrn; k++) {
    p.colptr[k + 1] = p.colptr[k] + ccount[k];
    p.valptr[k + 1] = p.valptr[k] + ccount[k] * p.nr[k];
  }
  p.ncol.assign((size_t)p.colptr[p.nrn], 0);
  par_for(p.nrn, [&](int64_t a, int64_t b) {
    std::vector<std::pair<int, int>> nb;  // (first eq, count)
    for (int64_t k = a; k < b; k++) {
      int n = bnode[k];
      nb.clear();
      nb.push_back({first[n], cnt[n]});
      for (int64_t j = m.gptr[n]; j < m.gptr[n + 1]; j++) {
        int o = m.gadj[j] - 1;
        if (cnt[o] > 0) nb.push_back({first[o], cnt[o]});
      }
      std::sort(nb.begin(), nb.end());
      int64_t w = p.colptr[k];
      for (auto &pr : nb)
        for (int d = 0; d < pr.second; d++) p.ncol[w++] = pr.first + d;
    }
  });
  return PM_OK;
}

// Pattern from a caller's strict-lower CSRC: rows are grouped into blocks of
// consecutive equations with identical full column sets (a block per row
// when they differ).
int pm_host_pattern_from_csrc(int neq, int64_t nad, const int64_t *ia, const int *ja,
                              PmHostPattern &p) {
  if (neq <= 0 || nad < 0 || ia[0] != 1 || ia[neq] - 1 != nad) return PM_ERR_ARG;
  // full rows: lower(i) + {i} + upper(i)
  std::vector<int64_t> ucnt((size_t)neq + 1, 0);
  for (int64_t k = 0; k < nad; k++) {
    int c = ja[k] - 1;
    if (c < 0 || c >= neq) return PM_ERR_ARG;
    ucnt[c]++;
  }
  std::vector<int64_t> uptr((size_t)neq + 1, 0);
  for (int i = 0; i < neq; i++) uptr[i + 1] = uptr[i] + ucnt[i];
  std::vector<int> urow((size_t)nad);
  {
    std::vector<int64_t> fill(uptr.begin(), uptr.end() - 1);
    for (int i = 0; i < neq; i++)
      for (int64_t k = ia[i] - 1; k < ia[i + 1] - 1; k++) urow[fill[ja[k] - 1]++] = i;
  }
  auto full_len = [&](int i) { return (ia[i + 1] - ia[i]) + 1 + (uptr[i + 1] - uptr[i]); };
  auto same_cols = [&](int i, int j) {  // rows i < j
    if (full_len(i) != full_len(j)) return false;
    std::vector<int> a, b;
    for (int64_t k = ia[i] - 1; k < ia[i + 1] - 1; k++) a.push_back(ja[k] - 1);
    a.push_back(i);
    for (int64_t k = uptr[i]; k < uptr[i + 1]; k++) a.push_back(urow[k]);
    for (int64_t k = ia[j] - 1; k < ia[j + 1] - 1; k++) b.push_back(ja[k] - 1);
    b.push_back(j);
    for (int64_t k = uptr[j]; k < uptr[j + 1]; k++) b.push_back(urow[k]);
    return a == b;
  };
  p.neq = neq;
  p.fr.clear();
  p.nr.clear();
  for (int i = 0; i < neq;) {
    int j = i + 1;
    while (j < neq && j - i < 4 && same_cols(i, j)) j++;
    p.fr.push_back(i);
    p.nr.push_back(j - i);
    i = j;
  }
  p.nrn = (int)p.fr.size();
  p.eq_block.assign((size_t)neq, 0);
  p.colptr.assign((size_t)p.nrn + 1, 0);
  p.valptr.assign((size_t)p.nrn + 1, 0);
  for (int k = 0; k < p.nrn; k++) {
    for (int q = 0; q < p.nr[k]; q++) p.eq_block[p.fr[k] + q] = k;
    int64_t c = full_len(p.fr[k]);
    p.colptr[k + 1] = p.colptr[k] + c;
    p.valptr[k + 1] = p.valptr[k] + c * p.nr[k];
  }
  p.ncol.assign((size_t)p.colptr[p.nrn], 0);
  for (int k = 0; k < p.nrn; k++) {
    int i = p.fr[k];
    int64_t w = p.colptr[k];
    for (int64_t q = ia[i] - 1; q < ia[i + 1] - 1; q++) p.ncol[w++] = ja[q] - 1;
    p.ncol[w++] = i;
    for (int64_t q = uptr[i]; q < uptr[i + 1]; q++) p.ncol[w++] = urow[q];
  }
  return PM_OK;
}

int64_t pm_host_count_lower(const PmHostPattern &p) {
  int64_t nad = 0;
  for (int k = 0; k < p.nrn; k++)
    for (int q = 0; q < p.nr[k]; q++) {
      int row = p.fr[k] + q;
      for (int64_t j = p.colptr[k]; j < p.colptr[k + 1]; j++)
        if (p.ncol[j] < row) nad++;
    }
  return nad;
}

// greedy colouring: elements sharing a vertex get different colours
void pm_host_colour(PmHostMesh &m) {
  std::vector<int64_t> ptr;
  std::vector<int> inc;
  vertex_incidence(m.ix20, m.numel, m.nnodev, ptr, inc);
  std::vector<int> colour((size_t)m.numel, -1);
  int ncol = 0;
  std::vector<int> stamp(64, -1);
  for (int e = 0; e < m.numel; e++) {
    const int *row = m.ix20.data() + (size_t)e * 21;
    for (int k = 0; k < 8; k++) {
      int v = row[k] - 1;
      for (int64_t q = ptr[v]; q < ptr[v + 1]; q++) {
        int c = colour[inc[q]];
        if (c >= 0) {
          if (c >= (int)stamp.size()) stamp.resize((size_t)c + 1, -1);
          stamp[c] = e;
        }
      }
    }
    int c = 0;
    while (c < (int)stamp.size() && stamp[c] == e) c++;
    colour[e] = c;
    ncol = std::max(ncol, c + 1);
  }
  m.colour_ptr.assign((size_t)ncol + 1, 0);
  for (int e = 0; e < m.numel; e++) m.colour_ptr[colour[e] + 1]++;
  for (int c = 0; c < ncol; c++) m.colour_ptr[c + 1] += m.colour_ptr[c];
  m.colour_elems.assign((size_t)m.numel, 0);
  std::vector<int> fill(m.colour_ptr.begin(), m.colour_ptr.end() - 1);
  for (int e = 0; e < m.numel; e++) m.colour_elems[fill[colour[e]]++] = e;
}

// ---------------------------------------------------------------------------
// Element-order assembly maps (reference summation order).  The reference
// adds element contributions in element order: al(k,l) += s(i,j) for the
// local pair whose row equation is the larger (csr, Assbly.f:218-304),
// ad(k) += s(i,i), b(k) -= p(i), and per vertex dporo += p (porosity_form).
// For every stored position of the node-block matrix (both triangles take
// the lower-oriented element entry), every equation and every node these
// maps list the element-array entries that add into it, in element order.
// ---------------------------------------------------------------------------
static inline int64_t block_pos(const PmHostPattern &p, int r, int c) {
  const int k = p.eq_block[r];
  const int64_t cp = p.colptr[k];
  const int cnt = (int)(p.colptr[k + 1] - cp);
  const int *cols = p.ncol.data() + cp;
  const int j = (int)(std::lower_bound(cols, cols + cnt, c) - cols);
  if (j >= cnt || cols[j] != c) return -1;
  return p.valptr[k] + (int64_t)(r - p.fr[k]) * cnt + j;
}

int pm_host_ordered_maps(const PmHostMesh &m, const PmHostPattern &p, PmOrderedMaps &o) {
  if (m.nen != 20) return PM_ERR_UNSUPPORTED;  // hexa20/8 element arrays only
  const int numel = m.numel, nst = PM_NST;
  const int64_t nnz = p.valptr[p.nrn];
  std::vector<int> eqs((size_t)numel * nst);
  for (int e = 0; e < numel; e++) {
    const int *ix = m.ix20.data() + (size_t)e * 21;
    for (int a = 0; a < nst; a++) {
      const int node = a < 60 ? ix[a / 3] : ix[a - 60];
      const int dof = a < 60 ? a % 3 : 3;
      eqs[(size_t)e * nst + a] = node > 0 ? m.id[4 * (node - 1) + dof] - 1 : -1;
    }
  }
  // matrix: count, then fill, element by element
  std::vector<int64_t> cnt((size_t)nnz + 1, 0);
  for (int pass = 0; pass < 2; pass++) {
    std::vector<int64_t> fill;
    if (pass == 1) {
      o.mptr.assign((size_t)nnz + 1, 0);
      for (int64_t t = 0; t < nnz; t++) o.mptr[t + 1] = o.mptr[t] + cnt[t];
      o.msrc.assign((size_t)o.mptr[nnz], 0);
      fill.assign(o.mptr.begin(), o.mptr.end() - 1);
    }
    for (int e = 0; e < numel; e++) {
      const int *q = eqs.data() + (size_t)e * nst;
      for (int a = 0; a < nst; a++) {
        if (q[a] < 0) continue;
        for (int b = 0; b < nst; b++) {
          if (q[b] < 0 || q[b] > q[a]) continue;
          const int64_t src = (int64_t)e * nst * nst + (int64_t)a * nst + b;
          const int64_t p1 = block_pos(p, q[a], q[b]);
          const int64_t p2 = q[a] == q[b] ? -1 : block_pos(p, q[b], q[a]);
          if (p1 < 0 || (q[a] != q[b] && p2 < 0)) return PM_ERR_STATE;
          if (pass == 0) {
            cnt[p1]++;
            if (p2 >= 0) cnt[p2]++;
          } else {
            o.msrc[fill[p1]++] = src;
            if (p2 >= 0) o.msrc[fill[p2]++] = src;
          }
        }
      }
    }
  }
  // right-hand side per equation
  std::vector<int64_t> rc((size_t)m.neq, 0);
  for (size_t i = 0; i < eqs.size(); i++)
    if (eqs[i] >= 0) rc[eqs[i]]++;
  o.rptr.assign((size_t)m.neq + 1, 0);
  for (int r = 0; r < m.neq; r++) o.rptr[r + 1] = o.rptr[r] + rc[r];
  o.rsrc.assign((size_t)o.rptr[m.neq], 0);
  std::vector<int64_t> rf(o.rptr.begin(), o.rptr.end() - 1);
  for (size_t i = 0; i < eqs.size(); i++)
    if (eqs[i] >= 0) o.rsrc[rf[eqs[i]]++] = (int64_t)i;
  // porosity per node (the 8 vertices of every element)
  std::vector<int64_t> pc((size_t)m.nnode, 0);
  for (int e = 0; e < numel; e++)
    for (int v = 0; v < 8; v++) {
      const int node = m.ix20[(size_t)e * 21 + v];
      if (node > 0) pc[node - 1]++;
    }
  o.pptr.assign((size_t)m.nnode + 1, 0);
  for (int n = 0; n < m.nnode; n++) o.pptr[n + 1] = o.pptr[n] + pc[n];
  o.psrc.assign((size_t)o.pptr[m.nnode], 0);
  std::vector<int64_t> pf(o.pptr.begin(), o.pptr.end() - 1);
  for (int e = 0; e < numel; e++)
    for (int v = 0; v < 8; v++) {
      const int node = m.ix20[(size_t)e * 21 + v];
      if (node > 0) o.psrc[pf[node - 1]++] = (int64_t)e * 8 + v;
    }
  return PM_OK;
}
