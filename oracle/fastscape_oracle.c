/*
 * fastscape_oracle.c -- CPU restatement of fastscapelib-fortran's FastScape_Execute_Step.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, never the thing measured or shipped).  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load this library.
 *
 * Each function restates the reference routine it cites (paths relative to the
 * fastscapelib-fortran source tree), keeping the reference's operation order so that a
 * -ffp-contract=off build reproduces the reference bit for bit (up to libm pow, which is the same
 * glibc routine the Fortran build calls).  Indices are 0-based; the reference is 1-based.
 *
 * Pinning: every routine is checked in tests/ against the unmodified reference Fortran compiled
 * from its own sources (oracle/_ref) on seeded inputs, and against closed forms and brute force
 * where the mathematics fixes the answer (see tests/test_oracle_*.py).
 */
#include "fastscape_oracle.h"

#include <float.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define NDMAX 8 /* donor / receiver slots per node, as don(8,nn) in src/FastScape_ctx.f90:25 */

struct fso_state {
  int nx, ny, nn;
  double xl, yl, dt;
  int ibc, i1, i2, j1, j2, xcyc, ycyc;
  unsigned char *bc;
  double *h, *u, *b, *etot, *erate, *a, *length, *precip, *kf, *kd, *sedflux, *lake_depth, *hwater,
      *g, *p_mfd_exp, *vx, *vy, *fmix;
  int *rec, *ndon, *don, *stack;
  int *mrec, *mnrec, *mstack;
  double *mwrec, *mlrec;
  double kfsed, m, n, kdsed, g1, g2, p, sealevel, newton_tol;
  double poro1, poro2, zporo1, zporo2, ratio, layer, kdsea1, kdsea2; /* SetMarineParam */
  int sfd, run_spl, run_uplift, run_diffusion, run_advect, run_marine;
  /* stratigraphy (Activate_Strati): reflectors (nn x nref) and fields (nn x 10 x nref), Fortran order */
  int run_strati, nref, iref, nfreq, nfreqref, strati_out;
  double vexref;
  double *refl, *sfields;
  int step;
  fso_stats st;
};

static double dmax(double a, double b) { return a > b ? a : b; }
static double dmin(double a, double b) { return a < b ? a : b; }

/* ------------------------------------------------------------------------------------------ */
/* SetBC, src/FastScape_ctx.f90:565-593.  write(cbc,'(i4)') right-justifies ibc in 4 chars;
 * digit 4 = left (i=1), 2 = right (i=nx), 1 = bottom (j=1), 3 = top (j=ny); '1' = fixed.       */
static void bc_chars(int ibc, char cbc[5]) {
  char tmp[32];
  snprintf(tmp, sizeof tmp, "%4d", ibc);
  if (strlen(tmp) > 4) strcpy(tmp, "****");
  memcpy(cbc, tmp, 5);
}

void fso_bc_decode(int ibc, int nx, int ny, int *i1, int *i2, int *j1, int *j2, int *xcyc, int *ycyc,
                   unsigned char *bc) {
  char c[5];
  bc_chars(ibc, c);
  int nn = nx * ny;
  *i1 = 1; *i2 = nx; *j1 = 1; *j2 = ny;
  if (c[3] == '1') *i1 = 2;
  if (c[1] == '1') *i2 = nx - 1;
  if (c[0] == '1') *j1 = 2;
  if (c[2] == '1') *j2 = ny - 1;
  if (bc) {
    memset(bc, 0, (size_t)nn);
    if (c[3] == '1') for (int k = 0; k < nn; k += nx) bc[k] = 1;
    if (c[1] == '1') for (int k = nx - 1; k < nn; k += nx) bc[k] = 1;
    if (c[0] == '1') for (int k = 0; k < nx; k++) bc[k] = 1;
    if (c[2] == '1') for (int k = nx * (ny - 1); k < nn; k++) bc[k] = 1;
  }
  *xcyc = (c[3] != '1' && c[1] != '1');
  *ycyc = (c[0] != '1' && c[2] != '1');
}

/* ------------------------------------------------------------------------------------------ */
/* find_receiver, src/FlowRouting.f90:302-362: D8 steepest descent, jj outer / ii inner, cyclic
 * wrap else clamp, first strict maximum of (h(ij)-h(ijk))/l above tiny wins.                     */
void fso_find_receiver(const double *h, int nx, int ny, double dx, double dy, int *rec, double *length,
                       int i1, int i2, int j1, int j2, int xcyc, int ycyc) {
  int nn = nx * ny;
  for (int ij = 0; ij < nn; ij++) { rec[ij] = ij; length[ij] = 0.0; }
  for (int j = j1; j <= j2; j++) {
    for (int i = i1; i <= i2; i++) {
      int ij = (i - 1) + (j - 1) * nx;
      double smax = DBL_MIN; /* tiny(smax) */
      for (int jj = -1; jj <= 1; jj++) {
        for (int ii = -1; ii <= 1; ii++) {
          int iii = i + ii;
          if (iii < 1 && xcyc) iii += nx;
          if (iii < 1) iii = 1;
          if (iii > nx && xcyc) iii -= nx;
          if (iii > nx) iii = nx;
          int jjj = j + jj;
          if (jjj < 1 && ycyc) jjj += ny;
          if (jjj < 1) jjj = 1;
          if (jjj > ny && ycyc) jjj -= ny;
          if (jjj > ny) jjj = ny;
          int ijk = (iii - 1) + (jjj - 1) * nx;
          if (ijk != ij) {
            double ddx = dx * (double)ii, ddy = dy * (double)jj;
            double l = sqrt(ddx * ddx + ddy * ddy);
            double slope = (h[ij] - h[ijk]) / l;
            if (slope > smax) { smax = slope; rec[ij] = ijk; length[ij] = l; }
          }
        }
      }
    }
  }
}

/* find_donor, src/FlowRouting.f90:366-389 (same as loc_min_3_find_donors, LocalMinima.f90:881-900):
 * donors listed in ascending node index.                                                         */
void fso_find_donor(const int *rec, int nn, int *ndon, int *don) {
  for (int ij = 0; ij < nn; ij++) ndon[ij] = 0;
  for (int ij = 0; ij < nn; ij++) {
    if (rec[ij] != ij) {
      int ijk = rec[ij];
      don[NDMAX * ijk + ndon[ijk]] = ij;
      ndon[ijk]++;
    }
  }
}

/* find_stack + find_stack_recursively, src/FlowRouting.f90:393-443 (and loc_min_3_find_stack,
 * LocalMinima.f90:904-923,1050-1067): depth-first preorder, roots in ascending index, children in
 * donor order.  Restated with an explicit stack (children pushed in reverse) -- same order.        */
void fso_find_stack(const int *rec, const int *don, const int *ndon, int nn, int *stack) {
  int *work = (int *)malloc(sizeof(int) * (size_t)(nn > 0 ? nn : 1));
  int nstack = 0;
  for (int ij = 0; ij < nn; ij++) {
    if (rec[ij] != ij) continue;
    int top = 0;
    work[top++] = ij;
    while (top > 0) {
      int v = work[--top];
      stack[nstack++] = v;
      for (int k = ndon[v] - 1; k >= 0; k--) work[top++] = don[NDMAX * v + k];
    }
  }
  free(work);
}

/* ------------------------------------------------------------------------------------------ */
/* loc_min_3_indexx, src/LocalMinima.f90:947-1046: Numerical Recipes quicksort index (median of
 * three, insertion sort below M=7, larger half pushed).  Unstable: the order of equal keys is part
 * of the result.  arr is 0-based; indx receives 1-based positions like the reference.           */
void fso_indexx(int n, const double *arr0, int *indx0) {
  enum { M = 7, NSTACK = 50 };
  const double *arr = arr0 - 1; /* 1-based views */
  int *indx = indx0 - 1;
  int istack[NSTACK + 1];
  int i, indxt, ir, itemp, j, jstack, k, l;
  double a;
  for (j = 1; j <= n; j++) indx[j] = j;
  jstack = 0;
  l = 1;
  ir = n;
  for (;;) {
    if (ir - l < M) {
      for (j = l + 1; j <= ir; j++) {
        indxt = indx[j];
        a = arr[indxt];
        for (i = j - 1; i >= 1; i--) {
          if (arr[indx[i]] <= a) break;
          indx[i + 1] = indx[i];
        }
        /* loop exhausted leaves i = 0, as the reference's explicit i=0 */
        indx[i + 1] = indxt;
      }
      if (jstack == 0) return;
      ir = istack[jstack];
      l = istack[jstack - 1];
      jstack -= 2;
    } else {
      k = (l + ir) / 2;
      itemp = indx[k]; indx[k] = indx[l + 1]; indx[l + 1] = itemp;
      if (arr[indx[l + 1]] > arr[indx[ir]]) { itemp = indx[l + 1]; indx[l + 1] = indx[ir]; indx[ir] = itemp; }
      if (arr[indx[l]] > arr[indx[ir]]) { itemp = indx[l]; indx[l] = indx[ir]; indx[ir] = itemp; }
      if (arr[indx[l + 1]] > arr[indx[l]]) { itemp = indx[l + 1]; indx[l + 1] = indx[l]; indx[l] = itemp; }
      i = l + 1;
      j = ir;
      indxt = indx[l];
      a = arr[indxt];
      for (;;) {
        do i++; while (arr[indx[i]] < a);
        do j--; while (arr[indx[j]] > a);
        if (j < i) break;
        itemp = indx[i]; indx[i] = indx[j]; indx[j] = itemp;
      }
      indx[l] = indx[j];
      indx[j] = indxt;
      jstack += 2;
      if (jstack > NSTACK) { fprintf(stderr, "NSTACK too small in loc_min_3_indexx\n"); abort(); }
      if (ir - i + 1 >= j - l) {
        istack[jstack] = ir;
        istack[jstack - 1] = i;
        ir = j - 1;
      } else {
        istack[jstack] = j - 1;
        istack[jstack - 1] = l;
        l = i;
      }
    }
  }
}

/* UnionFind / DoUnion, src/LocalMinima.f90:783-832 (union by rank; only the queried node is
 * re-pointed at its root).                                                                       */
static int uf_find(int x, int *parent) {
  int xp = x, xc = -1000;
  while (xc != xp) { xc = xp; xp = parent[xc]; }
  parent[x] = xc;
  return xc;
}
static void uf_union(int x, int y, int *parent, int *rank) {
  int xr = uf_find(x, parent), yr = uf_find(y, parent);
  if (xr != yr) {
    if (rank[xr] < rank[yr]) parent[xr] = yr;
    else {
      parent[yr] = xr;
      if (rank[xr] == rank[yr]) rank[xr]++;
    }
  }
}

/* compute_basins, src/LocalMinima.f90:136-169: basin id = running count of roots along the stack. */
static int compute_basins(const int *stack, const int *rec, int *basins, int *outlets, int nn) {
  int ib = -1;
  for (int i = 0; i < nn; i++) {
    int is = stack[i];
    if (rec[is] == is) { ib++; outlets[ib] = is; }
    basins[is] = ib;
  }
  return ib + 1;
}

/* connect_basins, src/LocalMinima.f90:173-329.  Nodes visited in stack order; an open basin (outlet
 * on a fixed node) gets a virtual edge to basin0 (the first open basin) of weight -1e10; every node
 * of a closed basin scans its 8 neighbours (no wrap) in the order of :239-240, skipping its own and
 * lower-numbered closed basins (:296), keeping per neighbour basin the first strictly lowest pass
 * max(h(p),h(q)).  Edge order = order of first discovery.  Returns nconn.                       */
static int connect_basins(int nbasins, const int *basins, const int *outlets, const int *rec,
                          const int *stack, const unsigned char *bc, const double *h, int nx, int ny,
                          int *basin0_out, int *cb, int *cn, double *cw, int cap) {
  static const int ii[8] = {-1, 1, 0, 0, -1, 1, -1, 1};
  static const int jj[8] = {0, 0, -1, 1, -1, -1, 1, 1};
  int n = nx * ny;
  int *conn_pos = (int *)malloc(sizeof(int) * (size_t)nbasins);
  int *conn_used = (int *)malloc(sizeof(int) * (size_t)nbasins);
  for (int k = 0; k < nbasins; k++) conn_pos[k] = -1;
  int nused = 0, iconn = 0, basin0 = -1, ibasin = 0, active = 0;
  for (int is_ = 0; is_ < n; is_++) {
    int node = stack[is_];
    if (rec[node] == node) {
      ibasin = basins[node];
      active = !bc[node];
      for (int u = 0; u < nused; u++) conn_pos[conn_used[u]] = -1;
      nused = 0;
      if (!active) {
        if (basin0 == -1) basin0 = ibasin;
        else {
          if (iconn < cap) {
            cb[2 * iconn] = basin0; cb[2 * iconn + 1] = ibasin;
            cn[2 * iconn] = -1; cn[2 * iconn + 1] = -1;
            cw[iconn] = -1.0e10;
          }
          iconn++;
        }
      }
    }
    if (!active) continue;
    int i = node % nx + 1, j = node / nx + 1;
    for (int k = 0; k < 8; k++) {
      int ki = i + ii[k], kj = j + jj[k];
      if (ki < 1 || ki > nx || kj < 1 || kj > ny) continue;
      int nb = (kj - 1) * nx + (ki - 1);
      int nbb = basins[nb];
      int nbo = outlets[nbb];
      if (ibasin >= nbb && !bc[nbo]) continue;
      double weight = dmax(h[node], h[nb]);
      int ci = conn_pos[nbb];
      if (ci == -1) {
        if (iconn < cap) {
          cb[2 * iconn] = ibasin; cb[2 * iconn + 1] = nbb;
          cn[2 * iconn] = node; cn[2 * iconn + 1] = nb;
          cw[iconn] = weight;
        }
        conn_pos[nbb] = iconn;
        iconn++;
        conn_used[nused++] = nbb;
      } else if (ci < cap && weight < cw[ci]) {
        cn[2 * ci] = node; cn[2 * ci + 1] = nb;
        cw[ci] = weight;
      }
    }
  }
  free(conn_pos);
  free(conn_used);
  *basin0_out = basin0;
  return iconn;
}

int fso_connect_basins_dump(const int *stack, const int *rec, const unsigned char *bc, const double *h,
                            int nx, int ny, int cap, int *basin0, int *cb, int *cn, double *cw) {
  int nn = nx * ny;
  int *basins = (int *)malloc(sizeof(int) * (size_t)nn), *outlets = (int *)malloc(sizeof(int) * (size_t)nn);
  int nb = compute_basins(stack, rec, basins, outlets, nn);
  int nc = connect_basins(nb, basins, outlets, rec, stack, bc, h, nx, ny, basin0, cb, cn, cw, cap);
  free(basins);
  free(outlets);
  return nc;
}

/* mst_kruskal, src/LocalMinima.f90:333-385: edges visited in loc_min_3_indexx order. */
static int mst_kruskal(const double *cw, const int *cb, int nbasins, int nconn, int *tree) {
  int *sort_id = (int *)malloc(sizeof(int) * (size_t)(nconn > 0 ? nconn : 1));
  int *parent = (int *)malloc(sizeof(int) * (size_t)nbasins), *rank = (int *)malloc(sizeof(int) * (size_t)nbasins);
  fso_indexx(nconn, cw, sort_id);
  for (int i = 0; i < nbasins; i++) { parent[i] = i; rank[i] = 0; }
  int ntree = 0;
  for (int e = 0; e < nconn; e++) {
    int eid = sort_id[e] - 1;
    int b0 = cb[2 * eid], b1 = cb[2 * eid + 1];
    int f0 = uf_find(b0, parent), f1 = uf_find(b1, parent);
    if (f0 != f1) {
      tree[ntree++] = eid;
      uf_union(b0, b1, parent, rank);
    }
  }
  free(sort_id); free(parent); free(rank);
  return ntree;
}

/* order_tree, src/LocalMinima.f90:389-533: depth-first walk of the MST from the sea basin; every
 * tree edge is turned so that (n1,p1) is on the parent side.  (sills/basin_stack feed only the dead
 * continuous_flow branch and are not restated.)                                                  */
static void order_tree(int *cb, int *cn, int nbasins, int root, const int *tree, int ntree) {
  int *csize = (int *)calloc((size_t)nbasins, sizeof(int)), *cptr = (int *)malloc(sizeof(int) * (size_t)nbasins);
  for (int t = 0; t < ntree; t++) { csize[cb[2 * tree[t]]]++; csize[cb[2 * tree[t] + 1]]++; }
  cptr[0] = 0;
  for (int i = 1; i < nbasins; i++) { cptr[i] = cptr[i - 1] + csize[i - 1]; csize[i - 1] = 0; }
  csize[nbasins - 1] = 0;
  int *adj = (int *)malloc(sizeof(int) * (size_t)(2 * ntree + 1));
  for (int t = 0; t < ntree; t++) {
    int e = tree[t], n1 = cb[2 * e], n2 = cb[2 * e + 1];
    adj[cptr[n1] + csize[n1]++] = e;
    adj[cptr[n2] + csize[n2]++] = e;
  }
  int *stk = (int *)malloc(sizeof(int) * 2 * (size_t)(nbasins + 1));
  int ss = 1;
  stk[0] = root; stk[1] = root;
  while (ss > 0) {
    ss--;
    int node = stk[2 * ss], parent = stk[2 * ss + 1];
    for (int q = cptr[node]; q < cptr[node] + csize[node]; q++) {
      int e = adj[q];
      if (cb[2 * e] == parent && node != parent) continue;
      if (node != cb[2 * e]) {
        int t = cb[2 * e]; cb[2 * e] = cb[2 * e + 1]; cb[2 * e + 1] = t;
        t = cn[2 * e]; cn[2 * e] = cn[2 * e + 1]; cn[2 * e + 1] = t;
      }
      stk[2 * ss] = cb[2 * e + 1];
      stk[2 * ss + 1] = node;
      ss++;
    }
  }
  free(csize); free(cptr); free(adj); free(stk);
}

/* correct_receivers_v2, src/LocalMinima.f90:601-662: for each real tree edge, the child-side pass
 * node drains to the parent-side pass node (length from the grid offset, :643-646) and the old
 * receiver path from the pass node to the child basin's outlet is reversed (:648-656).           */
static void correct_receivers_v2(int *rec, double *length, const int *outlets, const int *cb, const int *cn,
                                 const int *tree, int ntree, int nx, double dx, double dy) {
  for (int t = 0; t < ntree; t++) {
    int e = tree[t];
    int node_to = cn[2 * e], node_from = cn[2 * e + 1];
    if (node_from == -1) continue;
    int outlet_from = outlets[cb[2 * e + 1]];
    int next = rec[node_from], cur = node_from;
    double prev = length[node_from];
    rec[node_from] = node_to;
    double ddx = dx * (double)(node_from % nx - node_to % nx);
    double ddy = dy * (double)(node_from / nx - node_to / nx);
    length[node_from] = sqrt(ddx * ddx + ddy * ddy);
    while (cur != outlet_from) {
      int rn = rec[next];
      rec[next] = cur;
      double tmp = prev;
      prev = length[next];
      length[next] = tmp;
      cur = next;
      next = rn;
    }
  }
}

/* LocalMinima, src/LocalMinima.f90:1-132 (continuous_flow=.false., continuous_flow_v2=.true.). */
int fso_local_minima(int *stack, int *rec, const unsigned char *bc, int *ndon, int *don, const double *h,
                     double *length, int nx, int ny, double dx, double dy, fso_stats *st) {
  int nn = nx * ny;
  int *basins = (int *)malloc(sizeof(int) * (size_t)nn), *outlets = (int *)malloc(sizeof(int) * (size_t)nn);
  int nb = compute_basins(stack, rec, basins, outlets, nn);
  int cap = nb * 6; /* nconn_max, :59 */
  int *cb = (int *)malloc(sizeof(int) * 2 * (size_t)cap), *cn = (int *)malloc(sizeof(int) * 2 * (size_t)cap);
  double *cw = (double *)malloc(sizeof(double) * (size_t)cap);
  int nlb = 0;
  for (int k = 0; k < nn; k++) if (rec[k] == k && !bc[k]) nlb++;
  int basin0 = -1;
  int nconn = connect_basins(nb, basins, outlets, rec, stack, bc, h, nx, ny, &basin0, cb, cn, cw, cap);
  int status = 0;
  if (nconn > cap) status = -2;      /* the reference would write past conn_*(nconn_max) */
  else if (basin0 < 0) status = -3;  /* no open basin: the reference reads basin(-1)      */
  int ntree = 0;
  if (status == 0) {
    int *tree = (int *)malloc(sizeof(int) * (size_t)(nb > 1 ? nb - 1 : 1));
    ntree = mst_kruskal(cw, cb, nb, nconn, tree);
    order_tree(cb, cn, nb, basin0, tree, ntree);
    correct_receivers_v2(rec, length, outlets, cb, cn, tree, ntree, nx, dx, dy);
    free(tree);
  }
  int nla = 0;
  for (int k = 0; k < nn; k++) if (rec[k] == k && !bc[k]) nla++;
  if (st) {
    st->nbasins = nb;
    st->nopen = 0;
    for (int b = 0; b < nb; b++) if (bc[outlets[b]]) st->nopen++;
    st->nconn = nconn;
    st->ntree = ntree;
    st->nlocmin_before = nlb;
    st->nlocmin_after = nla;
    /* largest tie group among real edges */
    int mt = 0;
    if (nconn > 0 && nconn <= cap) {
      int *ix = (int *)malloc(sizeof(int) * (size_t)nconn);
      fso_indexx(nconn, cw, ix);
      int run = 1;
      for (int q = 1; q < nconn; q++) {
        double w0 = cw[ix[q - 1] - 1], w1 = cw[ix[q] - 1];
        if (w1 == w0 && w1 > -1.0e10) run++; else run = 1;
        if (run > mt) mt = run;
      }
      free(ix);
    }
    st->max_tie = mt;
  }
  free(cb); free(cn); free(cw); free(basins); free(outlets);
  fso_find_donor(rec, nn, ndon, don);
  fso_find_stack(rec, don, ndon, nn, stack);
  return status;
}

/* ------------------------------------------------------------------------------------------ */
static void dxdy(const fso_state *s, double *dx, double *dy) {
  *dx = s->xl / (double)(s->nx - 1); /* src/FlowRouting.f90:52-53 */
  *dy = s->yl / (double)(s->ny - 1);
}

/* longest receiver chain, for statistics only */
static int chain_depth(const int *stack, const int *rec, int nn) {
  int *d = (int *)malloc(sizeof(int) * (size_t)nn);
  int mx = 0;
  for (int i = 0; i < nn; i++) {
    int v = stack[i];
    d[v] = (rec[v] == v) ? 1 : d[rec[v]] + 1;
    if (d[v] > mx) mx = d[v];
  }
  free(d);
  return mx;
}

/* FlowRoutingSingleFlowDirection, src/FlowRouting.f90:43-93 */
static int flow_routing_sfd(fso_state *s) {
  double dx, dy;
  dxdy(s, &dx, &dy);
  int nn = s->nn;
  fso_find_receiver(s->h, s->nx, s->ny, dx, dy, s->rec, s->length, s->i1, s->i2, s->j1, s->j2, s->xcyc, s->ycyc);
  fso_find_donor(s->rec, nn, s->ndon, s->don);
  fso_find_stack(s->rec, s->don, s->ndon, nn, s->stack);
  int rc = fso_local_minima(s->stack, s->rec, s->bc, s->ndon, s->don, s->h, s->length, s->nx, s->ny, dx, dy, &s->st);
  s->st.depth = chain_depth(s->stack, s->rec, nn);
  /* lake surface: fill along the stack with +1e-8 (:73-86) */
  for (int i = 0; i < nn; i++) s->hwater[i] = s->h[i];
  for (int i = 0; i < nn; i++) {
    int ijk = s->stack[i], ijr = s->rec[ijk];
    if (s->hwater[ijr] > s->hwater[ijk]) s->hwater[ijk] = s->hwater[ijr] + 1.0e-8;
  }
  for (int i = 0; i < nn; i++) s->lake_depth[i] = s->hwater[i] - s->h[i];
  return rc;
}

/* FlowAccumulationSingleFlowDirection, src/FlowRouting.f90:123-143.  Note that a root (rec==self)
 * adds its own area to itself at the end of the reverse-stack walk, doubling it (:138).          */
static void flow_acc_sfd(fso_state *s) {
  double dx, dy;
  dxdy(s, &dx, &dy);
  int nn = s->nn;
  for (int i = 0; i < nn; i++) s->a[i] = dx * dy * s->precip[i];
  for (int i = nn - 1; i >= 0; i--) {
    int ijk = s->stack[i];
    s->a[s->rec[ijk]] = s->a[s->rec[ijk]] + s->a[ijk];
  }
}

/* find_mult_rec, src/FlowRouting.f90:147-298, on the state's h, rec and stack with exponent argument p,
 * into the given arrays (h0 = the filled surface, "water"). */
static int mult_rec(fso_state *s, double p, double *h0, int *mrec, int *mnrec, double *mwrec, double *mlrec,
                    int *mstack) {
  double dx, dy;
  dxdy(s, &dx, &dy);
  int nx = s->nx, ny = s->ny, nn = s->nn;
  for (int i = 0; i < nn; i++) h0[i] = s->h[i];
  for (int i = 0; i < nn; i++) { /* :187-196 */
    int ijk = s->stack[i], ijr = s->rec[ijk];
    if (h0[ijr] > h0[ijk]) h0[ijk] = h0[ijr] + 1.0e-8;
  }
  for (int i = 0; i < nn; i++) mnrec[i] = 0;
  for (int i = 0; i < NDMAX * nn; i++) mwrec[i] = 0.0;
  for (int j = s->j1; j <= s->j2; j++) { /* :204-230 */
    for (int i = s->i1; i <= s->i2; i++) {
      int ij = (j - 1) * nx + (i - 1);
      for (int jj = -1; jj <= 1; jj++) {
        int jjj = j + jj;
        if (jjj < 1 && s->ycyc) jjj += ny;
        if (jjj < 1) jjj = 1;
        if (jjj > ny && s->ycyc) jjj -= ny;
        if (jjj > ny) jjj = ny;
        for (int ii = -1; ii <= 1; ii++) {
          int iii = i + ii;
          if (iii < 1 && s->xcyc) iii += nx;
          if (iii < 1) iii = 1;
          if (iii > nx && s->xcyc) iii -= nx;
          if (iii > nx) iii = nx;
          int ijk = (jjj - 1) * nx + (iii - 1);
          if (h0[ij] > h0[ijk]) {
            int k = mnrec[ij]++;
            if (k >= NDMAX) return -4;
            mrec[NDMAX * ij + k] = ijk;
            double ddx = (double)ii * dx, ddy = (double)jj * dy;
            double l = sqrt(ddx * ddx + ddy * ddy);
            mlrec[NDMAX * ij + k] = l;
            mwrec[NDMAX * ij + k] = (h0[ij] - h0[ijk]) / l;
          }
        }
      }
    }
  }
  for (int ij = 0; ij < nn; ij++) { /* :232-243 */
    int nr = mnrec[ij];
    double *w = mwrec + NDMAX * ij;
    if (p < 0.0) {
      /* slope is double precision but rounded through real(); 0.5 and 0.6 are default-real
         literals promoted to double in the mixed-mode expression */
      double slope = 0.0;
      if (nr != 0) {
        double sw = 0.0;
        for (int k = 0; k < nr; k++) sw += w[k];
        slope = (double)(float)(sw / (double)nr);
      }
      s->p_mfd_exp[ij] = (double)0.5f + (double)0.6f * slope;
    }
    for (int k = 0; k < nr; k++) w[k] = pow(w[k], s->p_mfd_exp[ij]);
    double sumw = 0.0;
    for (int k = 0; k < nr; k++) sumw += w[k];
    for (int k = 0; k < nr; k++) w[k] = w[k] / sumw;
  }
  /* donors (:245-255) and LIFO Kahn order (:257-292) */
  int *ndon = (int *)calloc((size_t)nn, sizeof(int));
  for (int ij = 0; ij < nn; ij++)
    for (int k = 0; k < mnrec[ij]; k++) ndon[mrec[NDMAX * ij + k]]++;
  int *vis = (int *)calloc((size_t)nn, sizeof(int)), *parse = (int *)malloc(sizeof(int) * (size_t)nn);
  int nparse = 0, nstack = 0;
  for (int ij = 0; ij < nn; ij++) mstack[ij] = -1;
  for (int ij = 0; ij < nn; ij++) {
    if (ndon[ij] == 0) parse[nparse++] = ij;
    while (nparse > 0) {
      int ijn = parse[--nparse];
      mstack[nstack++] = ijn;
      for (int k = 0; k < mnrec[ijn]; k++) {
        int ijr = mrec[NDMAX * ijn + k];
        vis[ijr]++;
        if (vis[ijr] == ndon[ijr]) parse[nparse++] = ijr;
      }
    }
  }
  free(ndon); free(vis); free(parse);
  return nstack == nn ? 0 : -5; /* 'error in stack', :292 */
}

static int find_mult_rec(fso_state *s) {
  return mult_rec(s, s->p, s->hwater, s->mrec, s->mnrec, s->mwrec, s->mlrec, s->mstack);
}

/* FlowRouting (multiple flow direction), src/FlowRouting.f90:3-39 */
static int flow_routing_mfd(fso_state *s) {
  int rc = flow_routing_sfd(s); /* identical receiver/donor/stack/LocalMinima part, :16-28 */
  if (rc) return rc;
  rc = find_mult_rec(s);
  for (int i = 0; i < s->nn; i++) s->lake_depth[i] = s->hwater[i] - s->h[i];
  return rc;
}

/* FlowAccumulation (multiple flow direction), src/FlowRouting.f90:97-119 */
static void flow_acc_mfd(fso_state *s) {
  double dx, dy;
  dxdy(s, &dx, &dy);
  int nn = s->nn;
  for (int i = 0; i < nn; i++) s->a[i] = dx * dy * s->precip[i];
  for (int i = 0; i < nn; i++) {
    int ijk = s->mstack[i];
    for (int k = 0; k < s->mnrec[ijk]; k++) {
      int r = s->mrec[NDMAX * ijk + k];
      s->a[r] = s->a[r] + s->a[ijk] * s->mwrec[NDMAX * ijk + k];
    }
  }
}

/* One node of the implicit stream-power solve (StreamPowerLawSingleFlowDirection :296-366 /
 * StreamPowerLaw :94-176): reads its receivers' h and its single-flow receiver's water and lake sill. */
typedef struct {
  fso_state *s;
  double *h, *a, *ht, *kfint, *elev, *water, *lwv;
  int *lsill;
  double n, m, dt;
  int mfd, status;
  long newton;
} spl_ctx;

static void spl_node(spl_ctx *x, int ijk) {
  fso_state *s = x->s;
  double *h = x->h, *a = x->a, *ht = x->ht, *kfint = x->kfint, *elev = x->elev, *water = x->water, *lwv = x->lwv;
  int *lsill = x->lsill;
  const double n = x->n, m = x->m, dt = x->dt;
  const int mfd = x->mfd;
  int status = 0;
  long newton = 0;
  int ijr = s->rec[ijk];
  if (ijr == ijk) {
    water[ijk] = ht[ijk];
    lsill[ijk] = ijk;
    lwv[ijk] = 0.0;
    return;
  }
  double w_rcv = water[ijr];
  /* with Marine on, nodes below sea level keep their h (:105,144,305,342): n == 1 tests the current
   * h, otherwise ht */
  const int n1 = fabs(n - 1.0) < DBL_MIN;
  const int sea = s->run_marine && !((n1 ? h[ijk] : ht[ijk]) >= s->sealevel);
  if (elev[ijk] > w_rcv) {
    if (sea) {
    } else if (n1) {
      if (!mfd) { /* :306-314 */
        double f = elev[ijk], df = 1.0;
        double fact = kfint[ijk] * dt * pow(a[ijk], m) / s->length[ijk];
        f = f + fact * h[ijr];
        df = df + fact;
        h[ijk] = f / df;
      } else if (s->mnrec[ijk] > 0) { /* :104-116 */
        double f = elev[ijk], df = 1.0;
        for (int k = 0; k < s->mnrec[ijk]; k++) {
          int r = s->mrec[NDMAX * ijk + k];
          if (ht[ijk] >= ht[r]) {
            double fact = kfint[ijk] * dt * pow(a[ijk] * s->mwrec[NDMAX * ijk + k], m) / s->mlrec[NDMAX * ijk + k];
            f = f + fact * h[r];
            df = df + fact;
          }
        }
        h[ijk] = f / df;
      }
    } else if (!mfd || s->mnrec[ijk] > 0) { /* Newton, :343-358 (SFD) / :145-162 (MFD) */
      double omega = 0.875 / n, tolp = s->newton_tol, errp = 2.0 * tolp, h0 = elev[ijk];
      long it = 0;
      while (errp > tolp) {
        double f = h[ijk] - h0, df = 1.0;
        if (!mfd) {
          if (ht[ijk] > ht[ijr]) {
            double fact = kfint[ijk] * dt * pow(a[ijk], m) / pow(s->length[ijk], n);
            double x = dmax(0.0, h[ijk] - h[ijr]);
            f = f + fact * pow(x, n);
            df = df + fact * n * pow(x, n - 1.0);
          }
        } else {
          for (int k = 0; k < s->mnrec[ijk]; k++) {
            int r = s->mrec[NDMAX * ijk + k];
            if (ht[ijk] > ht[r]) {
              double fact = kfint[ijk] * dt * pow(a[ijk] * s->mwrec[NDMAX * ijk + k], m) /
                            pow(s->mlrec[NDMAX * ijk + k], n);
              double x = dmax(0.0, h[ijk] - h[r]);
              f = f + fact * pow(x, n);
              df = df + fact * n * pow(x, n - 1.0);
            }
          }
        }
        double hn = h[ijk] - f / df;
        errp = fabs(hn - h[ijk]);
        h[ijk] = h[ijk] * (1.0 - omega) + hn * omega;
        if (++it > 100000) { status = -6; break; }
      }
      newton += it;
    }
    lsill[ijk] = ijk;
    lwv[ijk] = 0.0;
    if (h[ijk] < w_rcv) h[ijk] = w_rcv;
  } else {
    h[ijk] = elev[ijk];
    lsill[ijk] = lsill[ijr];
    if (lsill[ijk] != -1) lwv[lsill[ijk]] = lwv[lsill[ijk]] + (w_rcv - h[ijk]);
  }
  water[ijk] = dmax(w_rcv, h[ijk]);
  if (status) x->status = status;
  x->newton += newton;
}

/* MFD solve order.  The reference walks mstack backwards, so a node's multiple-flow receivers are
 * always solved first; its single-flow receiver ijr1, whose water level and lake sill it reads
 * (:98-100), is among them except where h0(ijk) == h0(ijr1) exactly (flats left by a filled
 * lake).  There the reference's read depends on where ijr1 happens to sit in mstack: the current
 * value if later, else the previous Gauss-Seidel iteration's -- and in the first iteration the
 * never-initialised array (water is allocated without a value, :22).  This restatement fixes the
 * defined order: a node is solved once all its multiple-flow receivers and its single-flow receiver
 * are, which is the reference's order everywhere the two agree. */
static void mfd_solve_ready(spl_ctx *x, int v0, char *done, int *stk) {
  fso_state *s = x->s;
  if (done[v0]) return;
  int top = 0;
  stk[top++] = v0;
  while (top > 0) {
    int v = stk[top - 1], dep = -1;
    for (int k = 0; k < s->mnrec[v] && dep < 0; k++) {
      int r = s->mrec[NDMAX * v + k];
      if (!done[r]) dep = r;
    }
    if (dep < 0 && s->rec[v] != v && !done[s->rec[v]]) dep = s->rec[v];
    if (dep >= 0) {
      stk[top++] = dep;
      continue;
    }
    top--;
    if (!done[v]) {
      spl_node(x, v);
      done[v] = 1;
    }
  }
}

/* ------------------------------------------------------------------------------------------ */
/* StreamPowerLawSingleFlowDirection (mfd=0), src/StreamPowerLaw.f90:210-400, and StreamPowerLaw
 * (mfd=1), src/StreamPowerLaw.f90:3-206.  Gauss-Seidel loop on the deposition term; n=1 closed form
 * or relaxed Newton (omega=0.875/n); lakes keep h=elev and collect lake_water_volume.            */
static int stream_power(fso_state *s, int mfd) {
  double dx, dy;
  dxdy(s, &dx, &dy);
  int nn = s->nn;
  double *h = s->h, *a = s->a;
  double *ht = (double *)malloc(sizeof(double) * (size_t)nn), *kfint = (double *)malloc(sizeof(double) * (size_t)nn);
  double *dh = (double *)malloc(sizeof(double) * (size_t)nn), *hp = (double *)malloc(sizeof(double) * (size_t)nn);
  double *elev = (double *)malloc(sizeof(double) * (size_t)nn), *water = (double *)calloc((size_t)nn, sizeof(double));
  double *lwv = (double *)calloc((size_t)nn, sizeof(double)), *lsed = (double *)malloc(sizeof(double) * (size_t)nn);
  int *lsill = (int *)malloc(sizeof(int) * (size_t)nn);
  /* g and kf/kfsed (:29-36, :236-243) */
  if (s->g1 >= 0.0) {
    for (int i = 0; i < nn; i++) s->g[i] = s->g1;
    if (s->g2 > 0.0) for (int i = 0; i < nn; i++) if (h[i] - s->b[i] > 1.0) s->g[i] = s->g2;
  }
  for (int i = 0; i < nn; i++) kfint[i] = s->kf[i];
  if (s->kfsed > 0.0) for (int i = 0; i < nn; i++) if (h[i] - s->b[i] > 1.0) kfint[i] = s->kfsed;
  double hmax = 0.0;
  for (int i = 0; i < nn; i++) hmax = dmax(hmax, fabs(h[i]));
  double tol = 1.0e-4 * (hmax + 1.0), err = 2.0 * tol;
  double gmax = -DBL_MAX;
  for (int i = 0; i < nn; i++) gmax = dmax(gmax, s->g[i]);
  for (int i = 0; i < nn; i++) { ht[i] = h[i]; lsed[i] = 0.0; lsill[i] = -1; }
  const double n = s->n, m = s->m, dt = s->dt;
  const int *ord = mfd ? s->mstack : s->stack;
  int ngs = 0;
  char *done = (char *)malloc((size_t)nn);
  int *dfs = (int *)malloc(sizeof(int) * ((size_t)nn + 1));
  spl_ctx X = {s, h, a, ht, kfint, elev, water, lwv, lsill, n, m, dt, mfd, 0, 0};
  while (err > tol && ngs < 99) {
    ngs++;
    for (int i = 0; i < nn; i++) { hp[i] = h[i]; dh[i] = ht[i] - hp[i]; }
    /* erosion pushed downstream: reverse stack (SFD, :266-283) / forward mstack (MFD, :62-82) */
    for (int q = 0; q < nn; q++) {
      int ijk = mfd ? ord[q] : ord[nn - 1 - q];
      int ijr = s->rec[ijk];
      if (ijr != ijk) {
        dh[ijk] = dh[ijk] - (ht[ijk] - hp[ijk]);
        if (lsill[ijk] == ijk) lsed[ijk] = dh[ijk] <= 0.0 ? 0.0 : dh[ijk];
        dh[ijk] = dh[ijk] + (ht[ijk] - hp[ijk]);
        if (mfd) {
          for (int k = 0; k < s->mnrec[ijk]; k++) {
            int r = s->mrec[NDMAX * ijk + k];
            dh[r] = dh[r] + dh[ijk] * s->mwrec[NDMAX * ijk + k];
          }
        } else {
          dh[ijr] = dh[ijr] + dh[ijk];
        }
      } else {
        lsed[ijk] = dh[ijk];
      }
    }
    for (int i = 0; i < nn; i++) /* :285-289 */
      elev[i] = s->bc[i] ? ht[i] : ht[i] + (dh[i] - (ht[i] - hp[i])) * s->g[i] * dx * dy / a[i];
    /* solve from base level up: forward stack (SFD) / reverse mstack (MFD) */
    if (!mfd) {
      for (int q = 0; q < nn; q++) spl_node(&X, ord[q]);
    } else {
      for (int i = 0; i < nn; i++) done[i] = 0;
      for (int q = 0; q < nn; q++) mfd_solve_ready(&X, ord[nn - 1 - q], done, dfs);
    }
    double ss = 0.0;
    for (int i = 0; i < nn; i++) ss += (h[i] - hp[i]) * (h[i] - hp[i]);
    err = sqrt(ss / (double)nn);
    if (gmax < DBL_MIN) err = 0.0;
  }
  for (int i = 0; i < nn; i++) s->b[i] = dmin(h[i], s->b[i]);
  for (int i = 0; i < nn; i++) { /* lake sediment fill, :382-388 */
    int sl = lsill[i];
    if (sl != -1 && lwv[sl] > 0.0)
      h[i] = h[i] + dmax(0.0, dmin(lsed[sl], lwv[sl])) / lwv[sl] * (water[i] - h[i]);
  }
  for (int i = 0; i < nn; i++) {
    s->etot[i] = s->etot[i] + ht[i] - h[i];
    s->erate[i] = (ht[i] - h[i]) / dt;
    s->sedflux[i] = ht[i] - h[i];
  }
  s->st.ngs = ngs;
  s->st.newton_iters = X.newton;
  free(ht); free(kfint); free(dh); free(hp); free(elev); free(water); free(lwv); free(lsed); free(lsill);
  free(done); free(dfs);
  return X.status;
}

/* ------------------------------------------------------------------------------------------ */
/* tridag, src/Diffusion.f90:164-202 (Numerical Recipes Thomas algorithm). */
static int tridag(const double *a, const double *b, const double *c, const double *r, double *u, double *gam,
                  int n) {
  if (b[0] == 0.0) return -7;
  double bet = b[0];
  u[0] = r[0] / bet;
  for (int j = 1; j < n; j++) {
    gam[j] = c[j - 1] / bet;
    bet = b[j] - a[j] * gam[j];
    if (bet == 0.0) return -7;
    u[j] = (r[j] - a[j] * u[j - 1]) / bet;
  }
  for (int j = n - 2; j >= 0; j--) u[j] = u[j] - gam[j + 1] * u[j + 1];
  return 0;
}

/* Advect, src/Advect.f90:1-103: implicit upwind advection of h, b and etot, every row (x) and then
 * every column (y) solved with tridag; diagonal 1 + |v| dt/dx, the upwind neighbour's coefficient
 * -|v| dt/dx (sub-diagonal for v > 0, super-diagonal for v < 0), end rows of the system fixed;
 * finally b = min(b, h).                                                                        */
static int advect_lines(fso_state *s, int xdir) {
  const int nx = s->nx, ny = s->ny;
  const int n = xdir ? nx : ny, nl = xdir ? ny : nx, st = xdir ? 1 : nx;
  const double d = xdir ? s->xl / (nx - 1) : s->yl / (ny - 1);
  const double *v = xdir ? s->vx : s->vy;
  double *diag = malloc(sizeof(double) * n), *sup = malloc(sizeof(double) * n), *inf = malloc(sizeof(double) * n);
  double *rhs = malloc(sizeof(double) * n), *res = malloc(sizeof(double) * n), *gam = malloc(sizeof(double) * n);
  double *fld[3] = {s->h, s->b, s->etot};
  int rc = 0;
  for (int l = 0; l < nl && !rc; l++) {
    const int base = xdir ? l * nx : l;
    for (int k = 0; k < n; k++) {
      const double vk = v[base + k * st];
      diag[k] = 1.0; sup[k] = 0.0; inf[k] = 0.0;
      if (vk > 0.0) {
        diag[k] = 1.0 + vk * s->dt / d;
        inf[k] = -(vk * s->dt / d);
      } else if (vk < 0.0) {
        diag[k] = 1.0 - vk * s->dt / d;
        sup[k] = vk * s->dt / d;
      }
    }
    sup[0] = 0.0; diag[0] = 1.0; diag[n - 1] = 1.0; inf[n - 1] = 0.0;
    for (int f = 0; f < 3 && !rc; f++) {
      for (int k = 0; k < n; k++) rhs[k] = fld[f][base + k * st];
      rc = tridag(inf, diag, sup, rhs, res, gam, n);
      for (int k = 0; k < n; k++) fld[f][base + k * st] = res[k];
    }
  }
  free(diag); free(sup); free(inf); free(rhs); free(res); free(gam);
  return rc;
}

static int advect(fso_state *s) {
  int rc = advect_lines(s, 1);
  if (!rc) rc = advect_lines(s, 0);
  for (int i = 0; i < s->nn; i++) s->b[i] = dmin(s->b[i], s->h[i]);
  return rc;
}

/* Diffusion, src/Diffusion.f90:1-158: ADI, x-implicit rows j=2..ny-1, then y-implicit columns
 * i=2..nx-1.  Columns i=1,nx are never rewritten; the reflective-bottom coefficient (:114) uses
 * the loop index left over from the row loop (j=ny).                                           */
static int diffusion(fso_state *s) {
  double dx, dy;
  dxdy(s, &dx, &dy);
  int nx = s->nx, ny = s->ny, nn = s->nn;
  char cbc[5];
  bc_chars(s->ibc, cbc);
  double *zint = (double *)malloc(sizeof(double) * (size_t)nn), *kdi = (double *)malloc(sizeof(double) * (size_t)nn);
  double *zintp = (double *)malloc(sizeof(double) * (size_t)nn);
  for (int ij = 0; ij < nn; ij++) {
    zint[ij] = s->h[ij];
    kdi[ij] = s->kd[ij];
    if (s->kdsed > 0.0 && (s->h[ij] - s->b[ij]) > 1.0e-6) kdi[ij] = s->kdsed;
  }
  memcpy(zintp, zint, sizeof(double) * (size_t)nn);
#define K(i, j) kdi[((j)-1) * nx + (i)-1]
#define ZP(i, j) zintp[((j)-1) * nx + (i)-1]
#define Z(i, j) zint[((j)-1) * nx + (i)-1]
  const double hdt = s->dt / 2.0, dx2 = dx * dx, dy2 = dy * dy;
  int nmax = nx > ny ? nx : ny;
  double *f = (double *)calloc((size_t)nmax + 1, sizeof(double)), *diag = (double *)calloc((size_t)nmax + 1, sizeof(double));
  double *sup = (double *)calloc((size_t)nmax + 1, sizeof(double)), *inf = (double *)calloc((size_t)nmax + 1, sizeof(double));
  double *res = (double *)calloc((size_t)nmax + 1, sizeof(double)), *gam = (double *)calloc((size_t)nmax + 1, sizeof(double));
  int rc = 0;
  double fxp, fxm, fyp, fym;
  for (int j = 2; j <= ny - 1 && !rc; j++) { /* :45-86 ; arrays are 0-based: slot i-1 */
    for (int i = 2; i <= nx - 1; i++) {
      fxp = (K(i + 1, j) + K(i, j)) / 2.0 * hdt / dx2;
      fxm = (K(i - 1, j) + K(i, j)) / 2.0 * hdt / dx2;
      fyp = (K(i, j + 1) + K(i, j)) / 2.0 * hdt / dy2;
      fym = (K(i, j - 1) + K(i, j)) / 2.0 * hdt / dy2;
      diag[i - 1] = 1.0 + fxp + fxm;
      sup[i - 1] = -fxp;
      inf[i - 1] = -fxm;
      f[i - 1] = ZP(i, j) + fyp * ZP(i, j + 1) - (fyp + fym) * ZP(i, j) + fym * ZP(i, j - 1);
    }
    if (cbc[3] == '1') { diag[0] = 1.0; sup[0] = 0.0; f[0] = ZP(1, j); }
    else {
      fxp = (K(2, j) + K(1, j)) / 2.0 * hdt / dx2;
      fyp = (K(1, j + 1) + K(1, j)) / 2.0 * hdt / dy2;
      fym = (K(1, j - 1) + K(1, j)) / 2.0 * hdt / dy2;
      diag[0] = 1.0 + fxp;
      sup[0] = -fxp;
      f[0] = ZP(1, j) + fyp * ZP(1, j + 1) - (fyp + fym) * ZP(1, j) + fym * ZP(1, j - 1);
    }
    if (cbc[1] == '1') { diag[nx - 1] = 1.0; inf[nx - 1] = 0.0; f[nx - 1] = ZP(nx, j); }
    else {
      fxm = (K(nx - 1, j) + K(nx, j)) / 2.0 * hdt / dx2;
      fyp = (K(nx, j + 1) + K(nx, j)) / 2.0 * hdt / dy2;
      fym = (K(nx, j - 1) + K(nx, j)) / 2.0 * hdt / dy2;
      diag[nx - 1] = 1.0 + fxm;
      inf[nx - 1] = -fxm;
      f[nx - 1] = ZP(nx, j) + fyp * ZP(nx, j + 1) - (fyp + fym) * ZP(nx, j) + fym * ZP(nx, j - 1);
    }
    rc = tridag(inf, diag, sup, f, res, gam, nx);
    for (int i = 1; i <= nx; i++) Z(i, j) = res[i - 1];
  }
  const int jst = ny - 1 >= 2 ? ny : 2; /* value of j after "do j=2,ny-1" */
  for (int q = 0; q <= nmax; q++) { f[q] = diag[q] = sup[q] = inf[q] = res[q] = 0.0; }
  for (int i = 2; i <= nx - 1 && !rc; i++) { /* :97-138 */
    for (int j = 2; j <= ny - 1; j++) {
      fxp = (K(i + 1, j) + K(i, j)) / 2.0 * hdt / dx2;
      fxm = (K(i - 1, j) + K(i, j)) / 2.0 * hdt / dx2;
      fyp = (K(i, j + 1) + K(i, j)) / 2.0 * hdt / dy2;
      fym = (K(i, j - 1) + K(i, j)) / 2.0 * hdt / dy2;
      diag[j - 1] = 1.0 + fyp + fym;
      sup[j - 1] = -fyp;
      inf[j - 1] = -fym;
      f[j - 1] = Z(i, j) + fxp * Z(i + 1, j) - (fxp + fxm) * Z(i, j) + fxm * Z(i - 1, j);
    }
    if (cbc[0] == '1') { diag[0] = 1.0; sup[0] = 0.0; f[0] = Z(i, 1); }
    else {
      fxp = (K(i + 1, 1) + K(i, jst)) / 2.0 * hdt / dx2;
      fxm = (K(i - 1, 1) + K(i, 1)) / 2.0 * hdt / dx2;
      fyp = (K(i, 2) + K(i, 1)) / 2.0 * hdt / dy2;
      diag[0] = 1.0 + fyp;
      sup[0] = -fyp;
      f[0] = Z(i, 1) + fxp * Z(i + 1, 1) - (fxp + fxm) * Z(i, 1) + fxm * Z(i - 1, 1);
    }
    if (cbc[2] == '1') { diag[ny - 1] = 1.0; inf[ny - 1] = 0.0; f[ny - 1] = Z(i, ny); }
    else {
      fxp = (K(i + 1, ny) + K(i, ny)) / 2.0 * hdt / dx2;
      fxm = (K(i - 1, ny) + K(i, ny)) / 2.0 * hdt / dx2;
      fym = (K(i, ny - 1) + K(i, ny)) / 2.0 * hdt / dy2;
      diag[ny - 1] = 1.0 + fym;
      inf[ny - 1] = -fym;
      f[ny - 1] = Z(i, ny) + fxp * Z(i + 1, ny) - (fxp + fxm) * Z(i, ny) + fxm * Z(i - 1, ny);
    }
    rc = tridag(inf, diag, sup, f, res, gam, ny);
    for (int j = 1; j <= ny; j++) ZP(i, j) = res[j - 1];
  }
#undef K
#undef ZP
#undef Z
  if (!rc) {
    for (int ij = 0; ij < nn; ij++) { /* :143-152 */
      s->etot[ij] = s->etot[ij] + s->h[ij] - zintp[ij];
      s->erate[ij] = s->erate[ij] + (s->h[ij] - zintp[ij]) / s->dt;
      s->h[ij] = zintp[ij];
    }
    for (int ij = 0; ij < nn; ij++) s->b[ij] = dmin(s->h[ij], s->b[ij]);
  }
  free(zint); free(kdi); free(zintp); free(f); free(diag); free(sup); free(inf); free(res); free(gam);
  return rc;
}

/* ------------------------------------------------------------------------------------------ */
/* SiltSandCouplingDiffusion, src/Marine.f90:171-499: coupled silt / sand diffusion in the ocean,
 * alternating-direction passes (h along rows, silt fraction f along rows, h along columns, f along
 * columns) repeated until max|h - hp| <= 1 m.  Rows j = 2..ny-1 and columns i = 2..nx-1 (1-based) are
 * solved; the others keep their values (hhalf's first and last rows stay at their initial h, columns
 * 1 and nx of h take hhalf's), then the four corners copy their row neighbour.  Boundaries: a fixed
 * side ('1' digit) holds h; otherwise no flux, never cyclic.  q1, q2: silt and sand supply.  Returns
 * the iteration count, or -8 when it exceeds 1000 (the reference stops).                        */
int fso_silt_sand(double *h, double *f, const double *q1, const double *q2, int nx, int ny, double dx, double dy,
                  double dt, double sl, double L, double K1, double K2, int ibc) {
  const int nn = nx * ny;
  char cbc[5];
  bc_chars(ibc, cbc);
  const size_t sz = sizeof(double) * (size_t)nn;
  double *hp = malloc(sz), *fp = malloc(sz), *ht = malloc(sz), *ft = malloc(sz), *hhalf = malloc(sz);
  double *fhalf = malloc(sz), *fhalfp = malloc(sz), *tint = malloc(sz);
  int nmax = nx > ny ? nx : ny;
  double *diag = calloc((size_t)nmax, sizeof(double)), *sup = calloc((size_t)nmax, sizeof(double));
  double *inf = calloc((size_t)nmax, sizeof(double)), *rhs = calloc((size_t)nmax, sizeof(double));
  double *res = calloc((size_t)nmax, sizeof(double)), *gam = calloc((size_t)nmax, sizeof(double));
  memcpy(ht, h, sz); memcpy(ft, f, sz); memcpy(hhalf, h, sz); memcpy(fhalf, f, sz);
  const double tol = 1.0;
  double err1 = 2 * tol;
  int niter = 0, rc = 0;
#define I(i, j) (((j)-1) * nx + (i)-1)
  while (err1 > tol && !rc) {
    memcpy(hp, h, sz); memcpy(fp, f, sz); memcpy(fhalfp, fhalf, sz);
    niter++;
    /* h along x, :215-277 */
    for (int j = 2; j <= ny - 1 && !rc; j++) {
      for (int i = 1; i <= nx; i++) {
        int ij = I(i, j), ipj = I(i + 1, j), imj = I(i - 1, j), ijp = I(i, j + 1), ijm = I(i, j - 1);
        if (ht[ij] <= sl) { /* flag is 0 everywhere (its setting is commented out, :39) */
          double Ap, Bp, Cp;
          if (i == 1) {
            if (cbc[3] == '1') { diag[i - 1] = 1.0; sup[i - 1] = 0.0; rhs[i - 1] = ht[ij]; continue; }
            Ap = dt / 2.0 * (K2 + (K1 - K2) * (fhalfp[ipj] + fhalfp[ij]) / 2.0) / (dx * dx);
            diag[i - 1] = 1.0 + Ap;
            sup[i - 1] = -Ap;
          } else if (i == nx) {
            if (cbc[1] == '1') { diag[i - 1] = 1.0; inf[i - 1] = 0.0; rhs[i - 1] = ht[ij]; continue; }
            Bp = -(dt / 2.0 * (K2 + (K1 - K2) * (fhalfp[ij] + fhalfp[imj]) / 2.0) / (dx * dx));
            diag[i - 1] = 1.0 - Bp;
            inf[i - 1] = Bp;
          } else {
            Ap = dt / 2.0 * (K2 + (K1 - K2) * (fhalfp[ipj] + fhalfp[ij]) / 2.0) / (dx * dx);
            Bp = -(dt / 2.0 * (K2 + (K1 - K2) * (fhalfp[ij] + fhalfp[imj]) / 2.0) / (dx * dx));
            diag[i - 1] = 1.0 + Ap - Bp;
            sup[i - 1] = -Ap;
            inf[i - 1] = Bp;
          }
          Cp = dt / 2.0 * (K2 + (K1 - K2) * (ft[ijp] + ft[ij]) / 2.0) * (ht[ijp] - ht[ij]) / (dy * dy) -
               dt / 2.0 * (K2 + (K1 - K2) * (ft[ij] + ft[ijm]) / 2.0) * (ht[ij] - ht[ijm]) / (dy * dy) +
               (q1[ij] + q2[ij]) * dt / 2.0;
          rhs[i - 1] = Cp + ht[ij];
        } else {
          diag[i - 1] = 1.0; sup[i - 1] = 0.0; inf[i - 1] = 0.0; rhs[i - 1] = ht[ij];
        }
      }
      rc = tridag(inf, diag, sup, rhs, res, gam, nx);
      for (int i = 1; i <= nx; i++) hhalf[I(i, j)] = res[i - 1];
    }
    memcpy(tint, hhalf, sz);
    hhalf[I(1, 1)] = hhalf[I(2, 1)];
    hhalf[I(1, ny)] = hhalf[I(2, ny)];
    hhalf[I(nx, 1)] = hhalf[I(nx - 1, 1)];
    hhalf[I(nx, ny)] = hhalf[I(nx - 1, ny)];
    /* f along x, :287-340 */
    for (int j = 2; j <= ny - 1 && !rc; j++) {
      for (int i = 2; i <= nx - 1; i++) {
        int ij = I(i, j), ipj = I(i + 1, j), imj = I(i - 1, j), ijp = I(i, j + 1), ijm = I(i, j - 1);
        if (ht[ij] <= sl && hhalf[ij] >= (1.0 + 1.0e-6) * ht[ij]) {
          double Dp = (hhalf[ij] - ht[ij]) / dt;
          double Ep = K1 / 2.0 * (hhalf[ipj] - hhalf[ij]) / (dx * dx);
          double Mp = -(K1 / 2.0 * (hhalf[ij] - hhalf[imj]) / (dx * dx));
          double Np = K1 / 2.0 * (ft[ijp] + ft[ij]) * (ht[ijp] - ht[ij]) / (dy * dy) -
                      K1 / 2.0 * (ft[ij] + ft[ijm]) * (ht[ij] - ht[ijm]) / (dy * dy) + q1[ij];
          diag[i - 1] = 2.0 * L / dt + Dp - Mp - Ep;
          sup[i - 1] = -Ep;
          inf[i - 1] = -Mp;
          rhs[i - 1] = Np - Dp * ft[ij] + 2.0 * L * ft[ij] / dt;
        } else {
          diag[i - 1] = 1.0; sup[i - 1] = 0.0; inf[i - 1] = 0.0; rhs[i - 1] = ft[ij];
        }
      }
      diag[0] = 1.0; sup[0] = -1.0; rhs[0] = 0.0;
      diag[nx - 1] = 1.0; inf[nx - 1] = -1.0; rhs[nx - 1] = 0.0;
      rc = tridag(inf, diag, sup, rhs, res, gam, nx);
      for (int i = 1; i <= nx; i++) fhalf[I(i, j)] = res[i - 1];
    }
    for (int q = 0; q < nn; q++) fhalf[q] = dmin(dmax(0.0, fhalf[q]), 1.0);
    /* h along y, :345-405 */
    for (int i = 2; i <= nx - 1 && !rc; i++) {
      for (int j = 1; j <= ny; j++) {
        int ij = I(i, j), ipj = I(i + 1, j), imj = I(i - 1, j), ijp = I(i, j + 1), ijm = I(i, j - 1);
        if (ht[ij] <= sl) {
          double Ap, Bp, Cp;
          if (j == 1) {
            if (cbc[0] == '1') { diag[j - 1] = 1.0; sup[j - 1] = 0.0; rhs[j - 1] = hhalf[ij]; continue; }
            Ap = dt / 2.0 * (K2 + (K1 - K2) * (fp[ijp] + fp[ij]) / 2.0) / (dy * dy);
            diag[j - 1] = 1.0 + Ap;
            sup[j - 1] = -Ap;
          } else if (j == ny) {
            if (cbc[2] == '1') { diag[j - 1] = 1.0; inf[j - 1] = 0.0; rhs[j - 1] = hhalf[ij]; continue; }
            Bp = -(dt / 2.0 * (K2 + (K1 - K2) * (fp[ij] + fp[ijm]) / 2.0) / (dy * dy));
            diag[j - 1] = 1.0 - Bp;
            inf[j - 1] = Bp;
          } else {
            Ap = dt / 2.0 * (K2 + (K1 - K2) * (fp[ijp] + fp[ij]) / 2.0) / (dy * dy);
            Bp = -(dt / 2.0 * (K2 + (K1 - K2) * (fp[ij] + fp[ijm]) / 2.0) / (dy * dy));
            diag[j - 1] = 1.0 + Ap - Bp;
            sup[j - 1] = -Ap;
            inf[j - 1] = Bp;
          }
          Cp = dt / 2.0 * (K2 + (K1 - K2) * (fhalf[ipj] + fhalf[ij]) / 2.0) * (hhalf[ipj] - hhalf[ij]) / (dx * dx) -
               dt / 2.0 * (K2 + (K1 - K2) * (fhalf[ij] + fhalf[imj]) / 2.0) * (hhalf[ij] - hhalf[imj]) / (dx * dx) +
               (q1[ij] + q2[ij]) * dt / 2.0;
          rhs[j - 1] = Cp + hhalf[ij];
        } else {
          diag[j - 1] = 1.0; sup[j - 1] = 0.0; inf[j - 1] = 0.0; rhs[j - 1] = hhalf[ij];
        }
      }
      rc = tridag(inf, diag, sup, rhs, res, gam, ny);
      for (int j = 1; j <= ny; j++) tint[I(i, j)] = res[j - 1];
    }
    memcpy(h, tint, sz);
    h[I(1, 1)] = h[I(2, 1)];
    h[I(1, ny)] = h[I(2, ny)];
    h[I(nx, 1)] = h[I(nx - 1, 1)];
    h[I(nx, ny)] = h[I(nx - 1, ny)];
    /* f along y, :415-468 */
    for (int i = 2; i <= nx - 1 && !rc; i++) {
      for (int j = 2; j <= ny - 1; j++) {
        int ij = I(i, j), ipj = I(i + 1, j), imj = I(i - 1, j), ijp = I(i, j + 1), ijm = I(i, j - 1);
        if (ht[ij] <= sl && h[ij] >= (1.0 + 1.0e-6) * hhalf[ij]) {
          double Dp = (h[ij] - hhalf[ij]) / dt;
          double Ep = K1 / 2.0 * (h[ijp] - h[ij]) / (dy * dy);
          double Mp = -(K1 / 2.0 * (h[ij] - h[ijm]) / (dy * dy));
          double Np = K1 / 2.0 * (fhalf[ipj] + fhalf[ij]) * (hhalf[ipj] - hhalf[ij]) / (dx * dx) -
                      K1 / 2.0 * (fhalf[ij] + fhalf[imj]) * (hhalf[ij] - hhalf[imj]) / (dx * dx) + q1[ij];
          diag[j - 1] = 2.0 * L / dt + Dp - Mp - Ep;
          sup[j - 1] = -Ep;
          inf[j - 1] = -Mp;
          rhs[j - 1] = Np - Dp * fhalf[ij] + 2.0 * L * fhalf[ij] / dt;
        } else {
          diag[j - 1] = 1.0; sup[j - 1] = 0.0; inf[j - 1] = 0.0; rhs[j - 1] = fhalf[ij];
        }
      }
      diag[0] = 1.0; sup[0] = -1.0; rhs[0] = 0.0;
      diag[ny - 1] = 1.0; inf[ny - 1] = -1.0; rhs[ny - 1] = 0.0;
      rc = tridag(inf, diag, sup, rhs, res, gam, ny);
      for (int j = 1; j <= ny; j++) f[I(i, j)] = res[j - 1];
    }
    for (int q = 0; q < nn; q++) f[q] = dmin(dmax(0.0, f[q]), 1.0);
    err1 = 0.0; /* :472 */
    for (int q = 0; q < nn; q++) err1 = dmax(err1, fabs(h[q] - hp[q]));
    if (niter > 1000) rc = -8;
  }
#undef I
  for (int q = 0; q < nn; q++)
    if (h[q] >= sl + 1.0e-3) f[q] = 0.0;
  free(hp); free(fp); free(ht); free(ft); free(hhalf); free(fhalf); free(fhalfp); free(tint);
  free(diag); free(sup); free(inf); free(rhs); free(res); free(gam);
  return rc ? rc : niter;
}

/* Marine, src/Marine.f90:1-167: the sediment flux that reaches the sea along the step's single-flow
 * receivers, spread over the shelf along the multiple-flow receivers of the current h (find_mult_rec
 * with p = 0: weights slope**p_mfd_exp), then silt / sand diffusion in the ocean.  The compaction
 * block (:120-146) only computes arrays that nothing reads afterwards; it is left out.            */
static int marine(fso_state *s) {
  const int nn = s->nn;
  double dx, dy;
  dxdy(s, &dx, &dy);
  const double sl = s->sealevel;
  const size_t sz = sizeof(double) * (size_t)nn;
  double *flux = malloc(sz), *sd = malloc(sz), *ht = malloc(sz), *fs = malloc(sz), *q1 = malloc(sz), *q2 = malloc(sz);
  double *mw = malloc(NDMAX * sz), *ml = malloc(NDMAX * sz), *h0 = malloc(sz);
  int *mr = malloc(NDMAX * sizeof(int) * (size_t)nn), *mn = malloc(sizeof(int) * (size_t)nn);
  int *ms = malloc(sizeof(int) * (size_t)nn);
  for (int i = 0; i < nn; i++) flux[i] = s->h[i] > sl ? s->sedflux[i] : 0.0; /* :28-29 */
  for (int ij = nn - 1; ij >= 0; ij--) {                                       /* :30-36 */
    int ijk = s->stack[ij], ijr = s->rec[ijk];
    if (ijr != ijk && s->h[ijk] > sl) flux[ijr] = flux[ijr] + flux[ijk];
  }
  for (int i = 0; i < nn; i++)
    if (s->h[i] > sl) flux[i] = 0.0;
  const double ratio1 = s->ratio / (1.0 - s->poro1), ratio2 = (1.0 - s->ratio) / (1.0 - s->poro2); /* :46-50 */
  for (int i = 0; i < nn; i++) flux[i] = flux[i] * (ratio1 + ratio2);
  int rc = mult_rec(s, 0.0, h0, mr, mn, mw, ml, ms); /* :60-61 */
  if (!rc) {
    const double shelfslope = -1.0e-4; /* :67-79 */
    for (int i = 0; i < nn; i++) sd[i] = sl;
    for (int ij = 0; ij < nn; ij++) {
      int ijk = ms[ij];
      for (int k = 0; k < mn[ijk]; k++) {
        int ijr = mr[NDMAX * ijk + k];
        if (s->h[ijk] < sl) {
          sd[ijr] = dmin(sd[ijr], sd[ijk] + ml[NDMAX * ijk + k] * shelfslope);
          sd[ijr] = dmax(sd[ijr], s->h[ijr]);
        }
      }
    }
    for (int i = 0; i < nn; i++) /* :85-92 */
      if (s->h[i] < sl) flux[i] = flux[i] + (s->h[i] - sd[i]);
    for (int ij = 0; ij < nn; ij++) {
      int ijk = ms[ij];
      for (int k = 0; k < mn[ijk]; k++) {
        int ijr = mr[NDMAX * ijk + k];
        flux[ijr] = flux[ijr] + dmax(0.0, flux[ijk] * mw[NDMAX * ijk + k]);
      }
    }
    for (int i = 0; i < nn; i++) { /* :97-110 */
      if (flux[i] > 0.0 && s->h[i] < sl) flux[i] = -(s->h[i] - sd[i]);
      if (flux[i] <= 0.0 && s->h[i] < sl) flux[i] = flux[i] - (s->h[i] - sd[i]);
      if (s->h[i] >= sl) flux[i] = 0.0;
      flux[i] = dmax(flux[i], 0.0);
      fs[i] = flux[i] > 0.0 ? ratio1 / (ratio1 + ratio2) : 0.0;
      flux[i] = flux[i] / s->dt;
      q1[i] = flux[i] * fs[i];
      q2[i] = flux[i] * (1.0 - fs[i]);
    }
    memcpy(ht, s->h, sz); /* :113-114 */
    int it = fso_silt_sand(s->h, s->fmix, q1, q2, s->nx, s->ny, dx, dy, s->dt, sl, s->layer, s->kdsea1, s->kdsea2,
                           s->ibc);
    if (it < 0) rc = it;
    else s->st.marine_iters = it;
  }
  if (!rc) {
    for (int i = 0; i < nn; i++) { /* :148-165 */
      s->etot[i] = s->etot[i] + ht[i] - s->h[i];
      s->erate[i] = s->erate[i] + (ht[i] - s->h[i]) / s->dt;
      if (s->h[i] < sl) { s->sedflux[i] = 0.0; s->etot[i] = 0.0; s->erate[i] = 0.0; }
      if (s->h[i] >= sl + 1.0e-3) s->fmix[i] = 0.0;
      s->fmix[i] = dmin(dmax(0.0, s->fmix[i]), 1.0);
      s->b[i] = dmin(s->h[i], s->b[i]);
    }
  }
  free(flux); free(sd); free(ht); free(fs); free(q1); free(q2); free(mw); free(ml); free(h0); free(mr); free(mn);
  free(ms);
  return rc;
}

/* ------------------------------------------------------------------------------------------ */
/* distance_to_shore, src/Strati.f90:86-127: for every node whose surface changes sign towards its
 * receiver, the crossing point h(ij)/(h(rec)-h(ij)) of the way (0.5 when flat); every node's distance
 * to the nearest crossing (squared distances, 2(xl^2+yl^2) when there is none, then square roots). */
static void distance_to_shore(const double *h, double *d, int nx, int ny, const int *rec, double xl, double yl) {
  const int nn = nx * ny;
  double *x = malloc(sizeof(double) * (size_t)nn), *y = malloc(sizeof(double) * (size_t)nn);
  for (int ij = 0; ij < nn; ij++) {
    d[ij] = 2.0 * (xl * xl + yl * yl);
    x[ij] = xl * (double)(float)(ij % nx) / (double)(nx - 1);
    y[ij] = yl * (double)(float)(ij / nx) / (double)(ny - 1);
  }
  for (int ij = 0; ij < nn; ij++) {
    const int r = rec[ij];
    if (h[ij] * h[r] < 0.0) {
      double rat = 0.5;
      if (h[r] - h[ij] != 0.0) rat = h[ij] / (h[r] - h[ij]);
      const double xs = x[ij] + rat * (x[r] - x[ij]), ys = y[ij] + rat * (y[r] - y[ij]);
      for (int k = 0; k < nn; k++) {
        const double dist = (x[k] - xs) * (x[k] - xs) + (y[k] - ys) * (y[k] - ys);
        d[k] = dmin(d[k], dist);
      }
    }
  }
  for (int k = 0; k < nn; k++)
    if (d[k] > 0.0) d[k] = sqrt(d[k]);
  free(x); free(y);
}

#define SF(ij, l, i) s->sfields[(size_t)(ij) + (size_t)nn * ((l) + 10 * (size_t)(i))]
/* The fields of Strati, src/Strati.f90:31-50 (1-based reflector i, fields 1..9; field 10 is kept by
 * run_Strati): depth, slope, thickness to the reflector below (the basement for the first) and to the
 * basement, and, while the reflector is at or above the active one, depositional bathymetry, slope,
 * distance to shore and silt fraction; the age dt_ref * (nref - i). */
static void strati_fields(fso_state *s) {
  const int nx = s->nx, ny = s->ny, nn = s->nn;
  double dx, dy;
  dxdy(s, &dx, &dy);
  double *sl = malloc(sizeof(double) * (size_t)nn), *dist = calloc((size_t)nn, sizeof(double));
  const double dtref = s->dt * (double)s->nfreqref;
  for (int i = 1; i <= s->nref; i++) {
    const double *r = s->refl + (size_t)nn * (i - 1);
    fso_slope(r, nx, ny, dx, dy, sl);
    if (s->run_spl) distance_to_shore(r, dist, nx, ny, s->rec, s->xl, s->yl);
    for (int ij = 0; ij < nn; ij++) {
      SF(ij, 0, i - 1) = r[ij];
      SF(ij, 1, i - 1) = sl[ij];
      SF(ij, 2, i - 1) = i > 1 ? r[ij] - s->refl[(size_t)nn * (i - 2) + ij] : r[ij] - s->b[ij];
      SF(ij, 3, i - 1) = r[ij] - s->b[ij];
      if (i >= s->iref) {
        SF(ij, 4, i - 1) = r[ij] + s->sealevel;
        SF(ij, 5, i - 1) = sl[ij];
        SF(ij, 6, i - 1) = dist[ij];
        SF(ij, 7, i - 1) = s->fmix[ij];
      }
      SF(ij, 8, i - 1) = dmax(dtref * (double)(s->nref - i), 0.0);
    }
  }
  free(sl); free(dist);
}

/* run_Strati, src/FastScape_ctx.f90:809-839: reflectors uplifted, erosion below each active or buried
 * reflector accumulated (field 10), buried reflectors cut down to h, the active and later ones set to
 * h; every nfreq steps the active reflector advances (every nfreq_ref steps) and the fields are
 * recomputed for output.  The reference reads the reflectors before it first sets them; they start at
 * zero here (as its zero-filled allocation of them). */
static void run_strati(fso_state *s) {
  const int nn = s->nn;
  for (int i = 0; i < s->nref; i++)
    for (int ij = 0; ij < nn; ij++) s->refl[(size_t)nn * i + ij] = s->refl[(size_t)nn * i + ij] + s->u[ij] * s->dt;
  for (int i = 0; i < s->iref; i++)
    for (int ij = 0; ij < nn; ij++)
      SF(ij, 9, i) = SF(ij, 9, i) + dmax(0.0, s->refl[(size_t)nn * i + ij] - s->h[ij]);
  for (int i = 0; i < s->iref - 1; i++)
    for (int ij = 0; ij < nn; ij++) s->refl[(size_t)nn * i + ij] = dmin(s->refl[(size_t)nn * i + ij], s->h[ij]);
  for (int i = s->iref - 1; i < s->nref; i++)
    for (int ij = 0; ij < nn; ij++) s->refl[(size_t)nn * i + ij] = s->h[ij];
  s->strati_out = 0;
  if ((s->step + 1) % s->nfreq == 0) {
    if ((s->step + 1) % s->nfreqref == 0) s->iref++;
    strati_fields(s);
    s->strati_out = 1;
  }
}
#undef SF

/* Activate_Strati, src/FastScape_ctx.f90:781-805: nfield = 10, nfreqref = nstep / nreflector,
 * ireflector = 1, fields = 0. */
int fso_set_strati(fso_state *s, int nstep, int nreflector, int nfreq, double vex) {
  if (nreflector < 1 || nfreq < 1 || nstep / nreflector < 1) return -1;
  free(s->refl); free(s->sfields);
  s->nref = nreflector;
  s->nfreq = nfreq;
  s->nfreqref = nstep / nreflector;
  s->iref = 1;
  s->vexref = vex;
  s->refl = calloc((size_t)s->nn * nreflector, sizeof(double));
  s->sfields = calloc((size_t)s->nn * 10 * nreflector, sizeof(double));
  s->run_strati = 1;
  s->strati_out = 0;
  return 0;
}

/* The stratigraphy state: reflectors (nref x nn), fields (nref x 10 x nn), active reflector and
 * whether the last step was an output step. */
void fso_strati_state(const fso_state *s, double *refl, double *fields, int *iref, int *out) {
  const size_t nn = (size_t)s->nn;
  if (refl) memcpy(refl, s->refl, sizeof(double) * nn * s->nref);
  if (fields) memcpy(fields, s->sfields, sizeof(double) * nn * 10 * s->nref);
  if (iref) *iref = s->iref;
  if (out) *out = s->strati_out;
}

/* distance_to_shore on caller arrays (rec 0-based) */
void fso_distance_to_shore(const double *h, double *d, int nx, int ny, const int *rec, double xl, double yl) {
  distance_to_shore(h, d, nx, ny, rec, xl, yl);
}

/* ------------------------------------------------------------------------------------------ */
/* Init + SetNXNY + SetUp, src/FastScape_ctx.f90:40-113.  The reference fills h and catch0 with
 * random_number; callers here always provide h through init_h, so h starts at 0.               */
fso_state *fso_create(int nx, int ny) {
  if (nx < 3 || ny < 3) return NULL;
  fso_state *s = (fso_state *)calloc(1, sizeof(fso_state));
  s->nx = nx; s->ny = ny; s->nn = nx * ny;
  size_t nn = (size_t)s->nn;
#define DA(p) s->p = (double *)calloc(nn, sizeof(double))
#define IA(p) s->p = (int *)calloc(nn, sizeof(int))
  DA(h); DA(u); DA(b); DA(etot); DA(erate); DA(a); DA(length); DA(precip); DA(kf); DA(kd); DA(sedflux);
  DA(lake_depth); DA(hwater); DA(g); DA(p_mfd_exp); DA(vx); DA(vy); DA(fmix);
  IA(rec); IA(ndon); IA(stack); IA(mnrec); IA(mstack);
#undef DA
#undef IA
  s->don = (int *)calloc(NDMAX * nn, sizeof(int));
  s->mrec = (int *)calloc(NDMAX * nn, sizeof(int));
  s->mwrec = (double *)calloc(NDMAX * nn, sizeof(double));
  s->mlrec = (double *)calloc(NDMAX * nn, sizeof(double));
  s->bc = (unsigned char *)calloc(nn, 1);
  for (size_t i = 0; i < nn; i++) { s->precip[i] = 1.0; s->p_mfd_exp[i] = 1.0; s->fmix[i] = 0.5; }
  s->newton_tol = 1.0e-3;
  s->sealevel = 0.0;
  fso_set_bc(s, 1111);
  return s;
}

void fso_destroy(fso_state *s) {
  if (!s) return;
  free(s->h); free(s->u); free(s->b); free(s->etot); free(s->erate); free(s->a); free(s->length);
  free(s->precip); free(s->kf); free(s->kd); free(s->sedflux); free(s->lake_depth); free(s->hwater);
  free(s->g); free(s->p_mfd_exp); free(s->fmix); free(s->refl); free(s->sfields); free(s->rec); free(s->ndon); free(s->stack); free(s->mnrec);
  free(s->mstack); free(s->don); free(s->mrec); free(s->mwrec); free(s->mlrec); free(s->bc);
  free(s->vx); free(s->vy);
  free(s);
}

void fso_set_xl_yl(fso_state *s, double xl, double yl) { s->xl = xl; s->yl = yl; }
void fso_set_dt(fso_state *s, double dt) { s->dt = dt; }
void fso_set_bc(fso_state *s, int ibc) {
  s->ibc = ibc;
  fso_bc_decode(ibc, s->nx, s->ny, &s->i1, &s->i2, &s->j1, &s->j2, &s->xcyc, &s->ycyc, s->bc);
}
/* SetV, src/FastScape_ctx.f90:616-632 */
void fso_set_v(fso_state *s, const double *vx, const double *vy) {
  memcpy(s->vx, vx, sizeof(double) * (size_t)s->nn);
  memcpy(s->vy, vy, sizeof(double) * (size_t)s->nn);
  s->run_advect = 1;
}

void fso_set_u(fso_state *s, const double *u) { memcpy(s->u, u, sizeof(double) * (size_t)s->nn); s->run_uplift = 1; }
void fso_set_precip(fso_state *s, const double *p) { memcpy(s->precip, p, sizeof(double) * (size_t)s->nn); }
void fso_init_h(fso_state *s, const double *h) {
  memcpy(s->h, h, sizeof(double) * (size_t)s->nn);
  memcpy(s->b, h, sizeof(double) * (size_t)s->nn);
}
void fso_set_h(fso_state *s, const double *h) { memcpy(s->h, h, sizeof(double) * (size_t)s->nn); }
void fso_set_basement(fso_state *s, const double *b) { memcpy(s->b, b, sizeof(double) * (size_t)s->nn); }
void fso_set_newton_tol(fso_state *s, double tol) { s->newton_tol = tol; }
/* SetMarineParam, src/FastScape_ctx.f90:454-472 */
void fso_set_marine_parameters(fso_state *s, double sl, double p1, double p2, double z1, double z2, double r,
                               double l, double kds1, double kds2) {
  s->run_marine = 1;
  s->sealevel = sl; s->poro1 = p1; s->poro2 = p2; s->zporo1 = z1; s->zporo2 = z2; s->ratio = r; s->layer = l;
  s->kdsea1 = kds1; s->kdsea2 = kds2;
}
/* InitF, src/FastScape_ctx.f90:354-362 */
void fso_init_f(fso_state *s, const double *f) { memcpy(s->fmix, f, sizeof(double) * (size_t)s->nn); }
void fso_reset_cumulative_erosion(fso_state *s) { memset(s->etot, 0, sizeof(double) * (size_t)s->nn); }

/* SetErosionalParam, src/FastScape_ctx.f90:423-450 */
void fso_set_erosional_parameters(fso_state *s, const double *kf, double kfsed, double m, double n,
                                  const double *kd, double kdsed, double g1, double g2, double p) {
  s->run_spl = 1;
  memcpy(s->kf, kf, sizeof(double) * (size_t)s->nn);
  memcpy(s->kd, kd, sizeof(double) * (size_t)s->nn);
  s->kfsed = kfsed; s->m = m; s->n = n; s->kdsed = kdsed; s->g1 = g1; s->g2 = g2; s->p = p;
  for (int i = 0; i < s->nn; i++) s->p_mfd_exp[i] = p;
  s->sfd = 0;
  if (p < -1.5) { s->sfd = 1; s->p = 1.0; }
  double kdmax = -DBL_MAX;
  for (int i = 0; i < s->nn; i++) kdmax = dmax(kdmax, kd[i]);
  if (kdmax > DBL_MIN || kdsed > DBL_MIN) s->run_diffusion = 1;
}

/* FastScape_Execute_Step, src/FastScape_api.f90:201-263 (the stratigraphy's files are written by the host). */
int fso_execute_step(fso_state *s) {
  int rc = 0;
  if (s->run_advect) {
    rc = advect(s);
    if (rc) return rc;
  }
  if (s->run_uplift) { /* Uplift, src/Uplift.f90:9-10 */
    for (int i = 0; i < s->nn; i++) {
      s->h[i] = s->h[i] + s->u[i] * s->dt;
      s->b[i] = s->b[i] + s->u[i] * s->dt;
    }
  }
  if (s->run_spl) {
    if (s->sfd) {
      rc = flow_routing_sfd(s);
      if (!rc) { flow_acc_sfd(s); rc = stream_power(s, 0); }
    } else {
      rc = flow_routing_mfd(s);
      if (!rc) { flow_acc_mfd(s); rc = stream_power(s, 1); }
    }
  }
  if (!rc && s->run_diffusion) rc = diffusion(s);
  if (!rc && s->run_marine) rc = marine(s); /* :245-250 */
  if (!rc && s->run_strati) run_strati(s);  /* :252-257 */
  s->step++;
  return rc;
}

int fso_get_step(const fso_state *s) { return s->step; }
void fso_get_stats(const fso_state *s, fso_stats *out) { *out = s->st; }

void fso_copy_field(const fso_state *s, int field, double *dst) {
  const double *src = NULL;
  switch (field) {
    case FSO_H: src = s->h; break;
    case FSO_B: src = s->b; break;
    case FSO_ETOT: src = s->etot; break;
    case FSO_ERATE: src = s->erate; break;
    case FSO_AREA: src = s->a; break;
    case FSO_LAKE_DEPTH: src = s->lake_depth; break;
    case FSO_LENGTH: src = s->length; break;
    case FSO_HWATER: src = s->hwater; break;
    case FSO_SEDFLUX: src = s->sedflux; break;
    case FSO_FMIX: src = s->fmix; break;
    default: return;
  }
  memcpy(dst, src, sizeof(double) * (size_t)s->nn);
}

void fso_copy_ifield(const fso_state *s, int field, int *dst) {
  const int *src = NULL;
  switch (field) {
    case FSO_REC: src = s->rec; break;
    case FSO_NDON: src = s->ndon; break;
    case FSO_STACK: src = s->stack; break;
    case FSO_MNREC: src = s->mnrec; break;
    case FSO_MSTACK: src = s->mstack; break;
    default: return;
  }
  memcpy(dst, src, sizeof(int) * (size_t)s->nn);
}

void fso_copy_don(const fso_state *s, int *dst) { memcpy(dst, s->don, sizeof(int) * NDMAX * (size_t)s->nn); }

void fso_copy_mrec(const fso_state *s, int *mrec, double *mwrec, double *mlrec) {
  memcpy(mrec, s->mrec, sizeof(int) * NDMAX * (size_t)s->nn);
  memcpy(mwrec, s->mwrec, sizeof(double) * NDMAX * (size_t)s->nn);
  memcpy(mlrec, s->mlrec, sizeof(double) * NDMAX * (size_t)s->nn);
}

/* ------------------------------------------------------------------------------------------ */
/* Read-outs on the state after a step (FastScape_Copy_Slope / _Curvature / _Chi and
 * FastScape_Get_Fluxes, src/FastScape_api.f90:379-423,762-774).                                */

/* slope, src/TerrainDerivatives.f90:3-36: Sobel-weighted gradient of the 3x3 block, in degrees;
 * the outer ring stays 0.  Neighbours: a,b,c the row above (j+1), d,f the row, g,h,i below.    */
void fso_slope(const double *h, int nx, int ny, double dx, double dy, double *out) {
  const double con = 45.0 / atan(1.0);
  for (int ij = 0; ij < nx * ny; ij++) out[ij] = 0.0;
  for (int j = 1; j < ny - 1; j++)
    for (int i = 1; i < nx - 1; i++) {
      const int ij = i + j * nx;
      const int ia = ij + nx - 1, ib = ia + 1, ic = ib + 1, id = ij - 1, jf = ij + 1, ig = ij - nx - 1,
                ih = ig + 1, ii = ih + 1;
      const double dzdx = ((h[ic] + 2.0 * h[jf] + h[ii]) - (h[ia] + 2.0 * h[id] + h[ig])) / 8.0 / dx;
      const double dzdy = ((h[ig] + 2.0 * h[ih] + h[ii]) - (h[ia] + 2.0 * h[ib] + h[ic])) / 8.0 / dy;
      double v = dzdx * dzdx + dzdy * dzdy;
      if (v > DBL_MIN) v = atan(sqrt(v)) * con;
      out[ij] = v;
    }
}

/* curvature, src/TerrainDerivatives.f90:40-74: coefficients of the quadratic fitted to the 3x3
 * block (1..3 the row above, 4..6 the row, 7..9 below), profile-type curvature
 * (a(1+e^2)+b(1+d^2)-cde)/(1+d^2+e^2)^(3/2); the outer ring stays 0.                           */
void fso_curvature(const double *h, int nx, int ny, double dx, double dy, double *out) {
  for (int ij = 0; ij < nx * ny; ij++) out[ij] = 0.0;
  for (int j = 1; j < ny - 1; j++)
    for (int i = 1; i < nx - 1; i++) {
      const int ij = i + j * nx;
      const int i1 = ij + nx - 1, i2 = i1 + 1, i3 = i2 + 1, i4 = ij - 1, i5 = ij, i6 = ij + 1, i7 = ij - nx - 1,
                i8 = i7 + 1, i9 = i8 + 1;
      const double a = (h[i1] + h[i3] + h[i4] + h[i6] + h[i7] + h[i9]) / dx / dx / 12.0 -
                       (h[i2] + h[i5] + h[i8]) / dx / dx / 6.0;
      const double b = (h[i1] + h[i2] + h[i3] + h[i7] + h[i8] + h[i9]) / dy / dy / 12.0 -
                       (h[i4] + h[i5] + h[i6]) / dy / dy / 6.0;
      const double c = (h[i3] + h[i7] - h[i1] - h[i9]) / dx / dy / 4.0;
      const double d = (h[i3] + h[i6] + h[i9] - h[i1] - h[i4] - h[i7]) / dx / 6.0;
      const double e = (h[i1] + h[i2] + h[i3] - h[i7] - h[i8] - h[i9]) / dy / 6.0;
      double v = 1.0 + d * d + e * e;
      if (v > DBL_MIN) v = (a * (1.0 + e * e) + b * (1.0 + d * d) - c * d * e) / pow(v, 3.0 / 2.0);
      out[ij] = v;
    }
}

/* CopyChi, src/FastScape_ctx.f90:226-247: along the stack (any order that visits a receiver before
 * its donors), chi = chi(rec) + (a0/a)^(m/n) * length where a > a0 = 10 dx dy, else 0.            */
void fso_chi(const int *order, const int *rec, const double *a, const double *length, int nn, double dx,
             double dy, double m, double n, double *out) {
  const double a0 = dx * dy * 10.0;
  for (int ij = 0; ij < nn; ij++) out[ij] = 0.0;
  for (int ij = 0; ij < nn; ij++) {
    const int k = order[ij];
    if (a[k] > a0) out[k] = out[rec[k]] + pow(a0 / a[k], m / n) * length[k];
  }
}

/* compute_fluxes, src/FastScape_ctx.f90:843-886.  surf = xl yl/(nx-1)/(ny-1); tectonic = sum(u)
 * surf, erosion = sum(erate) surf (index order); the erosion rate, clamped at 0 at every node once
 * its donors have delivered, is pushed to the receivers -- single flow: along the reversed stack, so
 * a base-level node finally adds its own flux to itself; multiple flow: along mstack, weighted --
 * and boundary = sum over the fixed boundary nodes, index order, times surf.                      */
static void fluxes_sums(const double *u, const double *erate, const double *flux, const unsigned char *bc, int nn,
                        double surf, double out[3]) {
  double su = 0.0, se = 0.0, sb = 0.0;
  for (int ij = 0; ij < nn; ij++) su += u[ij];
  for (int ij = 0; ij < nn; ij++) se += erate[ij];
  for (int ij = 0; ij < nn; ij++)
    if (bc[ij]) sb += flux[ij];
  out[0] = su * surf;
  out[1] = se * surf;
  out[2] = sb * surf;
}

void fso_fluxes_sfd(const int *order, const int *rec, const double *u, const double *erate,
                    const unsigned char *bc, int nn, double surf, double out[3]) {
  double *flux = malloc(sizeof(double) * (size_t)nn);
  memcpy(flux, erate, sizeof(double) * (size_t)nn);
  for (int ij = nn - 1; ij >= 0; ij--) {
    const int k = order[ij];
    flux[k] = dmax(0.0, flux[k]);
    flux[rec[k]] = flux[rec[k]] + flux[k];
  }
  fluxes_sums(u, erate, flux, bc, nn, surf, out);
  free(flux);
}

void fso_fluxes_mfd(const int *mstack, const int *mnrec, const int *mrec, const double *mwrec, const double *u,
                    const double *erate, const unsigned char *bc, int nn, double surf, double out[3]) {
  double *flux = malloc(sizeof(double) * (size_t)nn);
  memcpy(flux, erate, sizeof(double) * (size_t)nn);
  for (int ij = 0; ij < nn; ij++) {
    const int k = mstack[ij];
    flux[k] = dmax(0.0, flux[k]);
    for (int q = 0; q < mnrec[k]; q++) {
      const int r = mrec[NDMAX * k + q];
      flux[r] = flux[r] + flux[k] * mwrec[NDMAX * k + q];
    }
  }
  fluxes_sums(u, erate, flux, bc, nn, surf, out);
  free(flux);
}

/* CopyCatchment, src/FastScape_ctx.f90:297-307: catch0 of the root of each node's tree, find_stack
 * propagating it from the root (src/FlowRouting.f90:408,437) over the receivers before LocalMinima;
 * catch0 comes from random_number at SetUp (src/FastScape_ctx.f90:94).  Here catch0 of a node is
 * (node + 1) / nn, so only which nodes share a value is comparable with the reference.            */
void fso_catchment(const int *rec, int nn, double *out) {
  int *root = malloc(sizeof(int) * (size_t)nn);
  for (int v = 0; v < nn; v++) root[v] = -1;
  for (int v = 0; v < nn; v++) {
    int r = v;
    while (root[r] < 0 && rec[r] != r) r = rec[r];
    const int top = root[r] >= 0 ? root[r] : r;
    for (int w = v; root[w] < 0 && w != r; w = rec[w]) root[w] = top;
    root[r] = top;
  }
  for (int v = 0; v < nn; v++) out[v] = (double)(root[v] + 1) / (double)nn;
  free(root);
}

void fso_copy_slope(const fso_state *s, double *out) {
  fso_slope(s->h, s->nx, s->ny, s->xl / (s->nx - 1), s->yl / (s->ny - 1), out);
}

void fso_copy_curvature(const fso_state *s, double *out) {
  fso_curvature(s->h, s->nx, s->ny, s->xl / (s->nx - 1), s->yl / (s->ny - 1), out);
}

void fso_copy_chi(const fso_state *s, double *out) {
  fso_chi(s->stack, s->rec, s->a, s->length, s->nn, s->xl / (s->nx - 1), s->yl / (s->ny - 1), s->m, s->n, out);
}

void fso_get_fluxes(const fso_state *s, double out[3]) {
  const double surf = s->xl * s->yl / (s->nx - 1) / (s->ny - 1);
  if (s->sfd)
    fso_fluxes_sfd(s->stack, s->rec, s->u, s->erate, s->bc, s->nn, surf, out);
  else
    fso_fluxes_mfd(s->mstack, s->mnrec, s->mrec, s->mwrec, s->u, s->erate, s->bc, s->nn, surf, out);
}
