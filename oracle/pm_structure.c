/*
 * pm_structure.c -- oracle restatement of the once-per-run structure build.
 *
 * TEST INFRASTRUCTURE ONLY (see pm_oracle.h).
 *
 *   orc_promote_hexa20  mk_elconn_quad_v1   Adjacency.f:583-693 (+ Hexa20edgeNod :974-1004)
 *   orc_bound_quad      mk_bound_quad       Rdata.f:2337-2405
 *   orc_graph           graph_v1/aponta_v1/graph0_v1/graph1_v1/sortgraphv2
 *                                           Graph.f:72-133,171-209,250-292,351-409,437-492
 *   orc_genrcm          genrcm_v1/fnroot_v1/rcm_v1/rootls_v1/degree_v1
 *                                           Reord.f:45-437 (SPARSPAK GENRCM of George & Liu)
 *   orc_numeq           numeqpmec2          Numeq.f:86-151
 *   orc_csrc_count/fill csria_v1/csrja_v1/sort_graph_csr, lower only
 *                                           Csr.f:259-295,631-867,2149-2157
 */
#include <stdlib.h>
#include <string.h>
#include "pm_oracle.h"

#define NEN1 (ORC_NEN + 1)
#define IX(k, e) ix[(size_t)(e) * NEN1 + (k) - 1]

/* vertex a, vertex b, mid-edge node of each hexa20 edge (1-based local ids) */
static const int EDGE20[12][3] = {{1, 2, 9},  {2, 3, 10}, {3, 4, 11}, {4, 1, 12},
                                  {5, 6, 13}, {6, 7, 14}, {7, 8, 15}, {8, 5, 16},
                                  {1, 5, 17}, {2, 6, 18}, {3, 7, 19}, {4, 8, 20}};

/* Gives every unset mid-edge slot a new node id, element by element, and
 * copies it into every neighbour that shares the edge.  ix(9..20, e) must be
 * zero on entry.  Returns nnode. */
int orc_promote_hexa20(int *ix, int numel, int nnodev) {
  /* element incidence of each vertex (nodegrade/elmincid, Colormsh.f:72-158) */
  int *nincid = calloc((size_t)nnodev + 1, sizeof(int));
  for (int e = 0; e < numel; e++)
    for (int k = 1; k <= 8; k++) nincid[IX(k, e)]++;
  int64_t *start = malloc(((size_t)nnodev + 2) * sizeof(int64_t));
  start[1] = 0;
  for (int n = 1; n <= nnodev; n++) start[n + 1] = start[n] + nincid[n];
  int *incid = malloc((size_t)(start[nnodev + 1] + 1) * sizeof(int));
  int *fill = calloc((size_t)nnodev + 1, sizeof(int));
  for (int e = 0; e < numel; e++)
    for (int k = 1; k <= 8; k++) {
      int n = IX(k, e);
      incid[start[n] + fill[n]++] = e;
    }
  int nno = nnodev;
  for (int i = 0; i < numel; i++) {
    for (int j = 8; j <= 20; j++)
      if (IX(j, i) == 0) IX(j, i) = ++nno;
    for (int j = 0; j < 12; j++) {
      int no1 = IX(EDGE20[j][0], i), no2 = IX(EDGE20[j][1], i), no3 = IX(EDGE20[j][2], i);
      for (int64_t k = start[no1]; k < start[no1 + 1]; k++) {
        int nel = incid[k];
        if (nel == i) continue;
        for (int l = 0; l < 12; l++) {
          int v1 = IX(EDGE20[l][0], nel), v2 = IX(EDGE20[l][1], nel);
          if ((no1 == v1 && no2 == v2) || (no1 == v2 && no2 == v1)) {
            if (IX(EDGE20[l][2], nel) == 0) IX(EDGE20[l][2], nel) = no3;
          }
        }
      }
    }
  }
  free(nincid); free(start); free(incid); free(fill);
  return nno;
}

/* mk_bound_quad with ndf1 = 3, ndf2 = 4: a mid-edge displacement dof is
 * restrained iff both edge ends are */
void orc_bound_quad(int *id, const int *ix, int numel) {
  for (int i = 0; i < numel; i++)
    for (int j = 0; j < 12; j++) {
      int no1 = IX(EDGE20[j][0], i), no2 = IX(EDGE20[j][1], i), no3 = IX(EDGE20[j][2], i);
      for (int k = 0; k < 3; k++)
        if (id[4 * (no1 - 1) + k] == 1 && id[4 * (no2 - 1) + k] == 1) id[4 * (no3 - 1) + k] = 1;
    }
}

/* size of the raw neighbour list (aponta_v1 ipos) */
/* nodes of element e inside 1..nnode (20 for hexa20; 8 for a maxno-8 mesh,
 * whose columns 9..20 are zero) */
static int elem_nodes(const int *ix, int e, int nnode) {
  int n = 0;
  for (int j = 1; j <= ORC_NEN; j++) {
    int no = IX(j, e);
    if (no != 0 && no <= nnode) n++;
  }
  return n;
}

int64_t orc_graph_count(const int *ix, int numel, int nnode) {
  int64_t ipos = 0;
  for (int e = 0; e < numel; e++) {
    int n = elem_nodes(ix, e, nnode);
    ipos += (int64_t)n * (n - 1);
  }
  return ipos;
}

static int cmp_int(const void *a, const void *b) {
  int x = *(const int *)a, y = *(const int *)b;
  return (x > y) - (x < y);
}

/* graph_v1: ip (nnode+1, 1-based positions) and ips (sorted, duplicate-free
 * neighbour lists, 1-based node ids).  work must hold orc_graph_count ints;
 * ips must hold the same number (only the compacted prefix is meaningful). */
void orc_graph(const int *ix, int numel, int nnode, int64_t *ip, int *ips, int *work) {
  int *cnt = calloc((size_t)nnode + 1, sizeof(int));
  memset(ip, 0, ((size_t)nnode + 1) * sizeof(int64_t));
  ip[0] = 1;
  for (int e = 0; e < numel; e++) {
    int n = elem_nodes(ix, e, nnode);
    for (int j = 1; j <= ORC_NEN; j++) {
      int no = IX(j, e);
      if (no == 0 || no > nnode) continue;
      ip[no] += n - 1;
    }
  }
  for (int i = 1; i <= nnode; i++) ip[i] += ip[i - 1];
  int *noviz = work;
  for (int e = 0; e < numel; e++)
    for (int j = 1; j <= ORC_NEN; j++) {
      int noj = IX(j, e);
      if (noj == 0 || noj > nnode) continue;
      for (int k = 1; k <= ORC_NEN; k++) {
        int nok = IX(k, e);
        if (nok == 0 || nok > nnode || nok == noj) continue;
        noviz[ip[noj - 1] - 1 + cnt[noj - 1]++] = nok;
      }
    }
  /* sort each list, then drop repeats (graph1_v1) */
  int64_t out = 0;
  int64_t old_start = ip[0];
  for (int i = 0; i < nnode; i++) {
    int64_t l = old_start, k = ip[i + 1];
    int n = (int)(k - l);
    qsort(noviz + (l - 1), (size_t)n, sizeof(int), cmp_int);
    int64_t row0 = out;
    for (int64_t j = l; j < k; j++) {
      int v = noviz[j - 1];
      if (j + 1 < k && noviz[j] == v) continue;
      ips[out++] = v;
    }
    old_start = k;
    ip[i + 1] = ip[i] + (out - row0);
  }
  free(cnt);
}

/* ---- SPARSPAK reverse Cuthill-McKee (Reord.f) ---- */
/* xadj, adj, perm, mask, xls are 1-based (index 0 unused) */
static void rootls(int root, const int64_t *xadj, const int *adj, int *mask, int *nlvl,
                   int *xls, int *ls) {
  mask[root] = 0;
  ls[1] = root;
  *nlvl = 0;
  int lvlend = 0, ccsize = 1;
  for (;;) {
    int lbegin = lvlend + 1;
    lvlend = ccsize;
    (*nlvl)++;
    xls[*nlvl] = lbegin;
    for (int i = lbegin; i <= lvlend; i++) {
      int node = ls[i];
      for (int64_t j = xadj[node]; j <= xadj[node + 1] - 1; j++) {
        int nbr = adj[j];
        if (mask[nbr] == 0) continue;
        ls[++ccsize] = nbr;
        mask[nbr] = 0;
      }
    }
    if (ccsize - lvlend <= 0) break;
  }
  xls[*nlvl + 1] = lvlend + 1;
  for (int i = 1; i <= ccsize; i++) mask[ls[i]] = 1;
}

static void fnroot(int *root, const int64_t *xadj, const int *adj, int *mask, int *nlvl,
                   int *xls, int *ls) {
  rootls(*root, xadj, adj, mask, nlvl, xls, ls);
  int ccsize = xls[*nlvl + 1] - 1;
  if (*nlvl == 1 || *nlvl == ccsize) return;
  for (;;) {
    int jstrt = xls[*nlvl];
    int mindeg = ccsize;
    *root = ls[jstrt];
    if (ccsize != jstrt) {
      for (int j = jstrt; j <= ccsize; j++) {
        int node = ls[j], ndeg = 0;
        for (int64_t k = xadj[node]; k <= xadj[node + 1] - 1; k++)
          if (mask[adj[k]] > 0) ndeg++;
        if (ndeg >= mindeg) continue;
        *root = node;
        mindeg = ndeg;
      }
    }
    int nunlvl;
    rootls(*root, xadj, adj, mask, &nunlvl, xls, ls);
    if (nunlvl <= *nlvl) return;
    *nlvl = nunlvl;
    if (*nlvl >= ccsize) return;
  }
}

/* degree_v1: degrees (within the masked component) of the nodes reachable
 * from root; the reference flips the sign of xadj(node) as its visited
 * marker, a separate flag array is the same thing */
static void degree(int root, const int64_t *xadj, const int *adj, const int *mask, int *deg,
                   int *ccsize_out, int *ls, char *seen) {
  ls[1] = root;
  seen[root] = 1;
  int lvlend = 0, ccsize = 1;
  for (;;) {
    int lbegin = lvlend + 1;
    lvlend = ccsize;
    for (int i = lbegin; i <= lvlend; i++) {
      int node = ls[i], ideg = 0;
      for (int64_t j = xadj[node]; j <= xadj[node + 1] - 1; j++) {
        int nbr = adj[j];
        if (mask[nbr] == 0) continue;
        ideg++;
        if (seen[nbr]) continue;
        seen[nbr] = 1;
        ls[++ccsize] = nbr;
      }
      deg[node] = ideg;
    }
    if (ccsize - lvlend <= 0) break;
  }
  for (int i = 1; i <= ccsize; i++) seen[ls[i]] = 0;
  *ccsize_out = ccsize;
}

static void rcm(int root, const int64_t *xadj, const int *adj, int *mask, int *perm,
                int *ccsize_out, int *deg, char *seen) {
  int ccsize;
  degree(root, xadj, adj, mask, deg, &ccsize, perm, seen);
  mask[root] = 0;
  *ccsize_out = ccsize;
  if (ccsize <= 1) return;
  int lvlend = 0, lnbr = 1;
  do {
    int lbegin = lvlend + 1;
    lvlend = lnbr;
    for (int i = lbegin; i <= lvlend; i++) {
      int node = perm[i];
      int fnbr = lnbr + 1;
      for (int64_t j = xadj[node]; j <= xadj[node + 1] - 1; j++) {
        int nbr = adj[j];
        if (mask[nbr] == 0) continue;
        lnbr++;
        mask[nbr] = 0;
        perm[lnbr] = nbr;
      }
      if (fnbr >= lnbr) continue;
      /* linear insertion sort of the new neighbours by increasing degree */
      int k = fnbr;
      do {
        int l = k;
        k++;
        int nbr = perm[k];
        while (l >= fnbr && deg[perm[l]] > deg[nbr]) {
          perm[l + 1] = perm[l];
          l--;
        }
        perm[l + 1] = nbr;
      } while (k < lnbr);
    }
  } while (lnbr > lvlend);
  int k = ccsize / 2, l = ccsize;
  for (int i = 1; i <= k; i++) {
    int t = perm[l];
    perm[l] = perm[i];
    perm[i] = t;
    l--;
  }
}

/* genrcm_v1.  xadj: n+1 1-based positions (as ip), adj: 1-based node ids,
 * perm (out): n entries, perm[i] = 1-based node placed at position i+1. */
void orc_genrcm(int n, const int64_t *xadj0, const int *adj0, int *perm0) {
  const int64_t *xadj = xadj0 - 1; /* xadj[1..n+1] */
  const int *adj = adj0 - 1;       /* adj[1..] */
  int *perm = perm0 - 1;           /* perm[1..n] */
  int *mask = malloc(((size_t)n + 2) * sizeof(int));
  int *xls = malloc(((size_t)n + 2) * sizeof(int));
  char *seen = calloc((size_t)n + 2, 1);
  for (int i = 1; i <= n; i++) mask[i] = 1;
  int num = 1;
  for (int i = 1; i <= n; i++) {
    if (mask[i] == 0) continue;
    int root = i, nlvl, ccsize;
    fnroot(&root, xadj, adj, mask, &nlvl, xls, perm + num - 1);
    rcm(root, xadj, adj, mask, perm + num - 1, &ccsize, xls, seen);
    num += ccsize;
    if (num > n) break;
  }
  free(mask); free(xls); free(seen);
}

/* numeqpmec2: equations interleaved per node in the order num[]; id holds
 * restraint flags on entry, equation numbers on exit.  The pressure slot of
 * a non-vertex node is left as it was. */
int orc_numeq(int *id, const int *num, const int *fnno, int nnode) {
  int neq = 0;
  for (int i = 0; i < nnode; i++) {
    int n = num[i];
    int ndfn = fnno[n - 1] == 1 ? ORC_NDF : ORC_NDF - 1;
    for (int j = 0; j < ndfn; j++) {
      int k = id[ORC_NDF * (n - 1) + j];
      id[ORC_NDF * (n - 1) + j] = k == 0 ? ++neq : 0;
    }
  }
  return neq;
}

/* csria_v1 (lower part only): ia[0..neq], 1-based; returns nad */
int64_t orc_csrc_count(const int *id, const int *num, const int64_t *ip, const int *ips,
                       int nnode, int neq, int64_t *ia) {
  memset(ia, 0, ((size_t)neq + 1) * sizeof(int64_t));
  for (int no = 0; no < nnode; no++) {
    int i = num[no];
    for (int ii = 0; ii < ORC_NDF; ii++) {
      int neqi = id[ORC_NDF * (i - 1) + ii];
      if (!(neqi > 0 && neqi <= neq)) continue;
      for (int kk = 0; kk < ORC_NDF; kk++) {
        int neqj = id[ORC_NDF * (i - 1) + kk];
        if (neqj <= 0) continue;
        if (neqj < neqi) ia[neqi]++;
      }
      for (int64_t k = ip[i - 1]; k < ip[i]; k++) {
        int j = ips[k - 1];
        for (int jj = 0; jj < ORC_NDF; jj++) {
          int neqj = id[ORC_NDF * (j - 1) + jj];
          if (neqj <= 0) continue;
          if (neqj < neqi) ia[neqi]++;
        }
      }
    }
  }
  ia[0] = 1;
  for (int i = 1; i <= neq; i++) ia[i] += ia[i - 1];
  return ia[neq] - 1;
}

/* csrja_v1 + sort_graph_csr (lower part only) */
void orc_csrc_fill(const int *id, const int *num, const int64_t *ip, const int *ips, int nnode,
                   int neq, const int64_t *ia, int *ja) {
  int64_t n = 0;
  for (int no = 0; no < nnode; no++) {
    int i = num[no];
    for (int ii = 0; ii < ORC_NDF; ii++) {
      int neqi = id[ORC_NDF * (i - 1) + ii];
      if (!(neqi > 0 && neqi <= neq)) continue;
      for (int kk = 0; kk < ORC_NDF; kk++) {
        int neqj = id[ORC_NDF * (i - 1) + kk];
        if (neqj <= 0) continue;
        if (neqj < neqi) ja[n++] = neqj;
      }
      for (int64_t k = ip[i - 1]; k < ip[i]; k++) {
        int j = ips[k - 1];
        for (int jj = 0; jj < ORC_NDF; jj++) {
          int neqj = id[ORC_NDF * (j - 1) + jj];
          if (neqj <= 0) continue;
          if (neqj < neqi) ja[n++] = neqj;
        }
      }
    }
  }
  for (int i = 0; i < neq; i++)
    qsort(ja + ia[i] - 1, (size_t)(ia[i + 1] - ia[i]), sizeof(int), cmp_int);
}
