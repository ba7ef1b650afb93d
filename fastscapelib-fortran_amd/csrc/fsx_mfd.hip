// fsx_mfd.hip -- multiple-flow-direction routing, accumulation and stream power.
//
// The reference (FlowRouting, src/FlowRouting.f90:3-39; find_mult_rec :147-298; FlowAccumulation
// :97-119; StreamPowerLaw src/StreamPowerLaw.f90:3-206) first routes and reroutes single receivers
// exactly as the D8 path, fills the surface h0 along that forest, gives every node all strictly
// lower neighbours of h0 as receivers with slope**p weights, orders the nodes with a LIFO Kahn
// traversal (mstack) and walks that order serially.
//
// Here the receiver graph is a DAG whose Kahn layers replace mstack: layer 0 holds the nodes
// without donors, layer l+1 the nodes whose last donor sits in layer l.  Every receiver of a node
// lies in a later layer, and so does its single-flow receiver (a "ghost" edge is added in the rare
// case where that receiver is not also a multiple-flow receiver), so
//  * accumulation and the erosion push run layer by layer from layer 0 (each node pulls from its
//    donors), and
//  * the implicit solve runs from the last layer back (each node reads its receivers).
// Sums over donors are taken in a fixed order (neighbour direction, then receiver slot), so the
// results do not depend on the batch or on the order of nodes inside a layer; they differ from the
// reference's mstack order only in the rounding of those sums.
//
// Per-node edge records are 8 slots in the reference's receiver order (scan order jj outer, ii
// inner, src/FlowRouting.f90:204-230): target (local index, -1 unused), weight, direction.
#include "fsx_scan.h"
#include "fsx_state.h"

#define MFD_PROLOGUE                                        \
  const Chunk c = chunks[chunk_id()];                       \
  const InstParams& P = PP[c.inst];                         \
  const int v = c.start + (int)threadIdx.x;                 \
  if (!P.active || !P.run_spl || P.sfd || v >= P.nn) return; \
  const long long g_ = (long long)P.off + v;

static __device__ __forceinline__ void mfd_barrier() { __syncthreads(); }

// ---- lake surface along the single-flow forest (find_mult_rec :187-196) ------------------------
// h0 = h, then along the stack h0(ijk) = h0(rec)+1e-8 where h0(rec) > h0(ijk); lake_depth = h0 - h
// (FlowRouting :34).  Levels of the single-flow forest from base level up.
__global__ __launch_bounds__(FSX_WG) void k_mfd_fill(const InstParams* __restrict__ PP, const int* __restrict__ lstart,
                                                     const int* __restrict__ nlev, const int* __restrict__ order,
                                                     const int* __restrict__ rec, const double* __restrict__ h,
                                                     double* __restrict__ hwater, double* __restrict__ lake_depth) {
  const int inst = blockIdx.x;
  const InstParams& P = PP[inst];
  if (!P.active || !P.run_spl || P.sfd) return;
  const long long off = P.off;
  const int* ls = lstart + P.aux;
  const int L = nlev[inst];
  for (int l = 0; l < L; l++) {
    const int b0 = ls[l], b1 = ls[l + 1];
    for (int p = b0 + (int)threadIdx.x; p < b1; p += FSX_WG) {
      const int v = order[off + p];
      const double hv = h[off + v];
      double hw = hv;
      if (l > 0) {
        const double hr = hwater[off + rec[off + v]];
        if (hr > hv) hw = hr + 1.0e-8;
      }
      hwater[off + v] = hw;
      lake_depth[off + v] = hw - hv;
    }
    mfd_barrier();
  }
}

// ---- receivers and weights (find_mult_rec :204-243) ---------------------------------------------
// Every strictly lower neighbour of h0 in scan order, wrap if cyclic else clamp (a clamped
// direction that lands on another node repeats that node, as the reference does).  Weights
// slope**pexp normalised by their sum; pexp = p, or 0.5 + 0.6*mean slope when p < 0 with the mean
// rounded through single precision and 0.6 a single-precision literal.  In-degrees (for the Kahn
// layers) count every slot, plus the ghost edge to the single-flow receiver when needed.
__global__ __launch_bounds__(FSX_CHUNK) void k_mfd_rec(const InstParams* __restrict__ PP, const Chunk* __restrict__ chunks,
                                                       const double* __restrict__ h0, const int* __restrict__ rec,
                                                       int* __restrict__ mrec, double* __restrict__ mw,
                                                       uint8_t* __restrict__ mdir, int* __restrict__ mnrec,
                                                       uint8_t* __restrict__ mflag, int* __restrict__ mindeg,
                                                       double* __restrict__ pexp) {
  MFD_PROLOGUE
  const double* hi = h0 + P.off;
  const int i = v % P.nx, j = v / P.nx;
  int t[8];
  double w[8];
  int dd[8];
  int nr = 0;
  if (i >= P.i1 && i <= P.i2 && j >= P.j1 && j <= P.j2) {
    const double hv = hi[v];
    for (int jj = -1; jj <= 1; jj++)
      for (int ii = -1; ii <= 1; ii++) {
        const int q = nbr_wc(P, i, j, ii, jj);
        if (hv > hi[q]) {  // never true for q == v
          const int d = (ii + 1) + 3 * (jj + 1);
          t[nr] = q;
          dd[nr] = d;
          w[nr] = (hv - hi[q]) / P.len[d];
          nr++;
        }
      }
  }
  double pe = P.p;
  if (P.p < 0.0) {
    double slope = 0.0;
    if (nr != 0) {
      double s = 0.0;
      for (int k = 0; k < nr; k++) s = s + w[k];
      slope = (double)(float)(s / (double)nr);
    }
    pe = (double)0.5f + (double)0.6f * slope;
  }
  pexp[g_] = pe;
  if (pe != 1.0)  // x**1 is x exactly, as libm's pow returns it
    for (int k = 0; k < nr; k++) w[k] = pow(w[k], pe);
  double sw = 0.0;
  for (int k = 0; k < nr; k++) sw = sw + w[k];
  bool hit = false;
  const int r = rec[g_];
  for (int k = 0; k < 8; k++) {
    const bool on = k < nr;
    mrec[8 * g_ + k] = on ? t[k] : -1;
    mw[8 * g_ + k] = on ? w[k] / sw : 0.0;
    mdir[8 * g_ + k] = on ? (uint8_t)dd[k] : (uint8_t)4;
    if (on) {
      hit |= t[k] == r;
      atomicAdd(mindeg + P.off + t[k], 1);
    }
  }
  mnrec[g_] = nr;
  const bool ghost = !hit && r != v;
  mflag[g_] = ghost ? 1 : 0;
  if (ghost) atomicAdd(mindeg + P.off + r, 1);
}

// ---- donors: the slots of neighbouring nodes that point here ------------------------------------
// A node u with v among its receivers is one of v's wrap/clamp neighbours, so v scans those (in
// direction order) and records every slot u*8+k with mrec(k,u) == v.  At most 8 per node.
__global__ __launch_bounds__(FSX_CHUNK) void k_mfd_don(const InstParams* __restrict__ PP, const Chunk* __restrict__ chunks,
                                                       const int* __restrict__ mrec, const int* __restrict__ mnrec,
                                                       int* __restrict__ mdon, int* __restrict__ mndon,
                                                       int* __restrict__ err) {
  MFD_PROLOGUE
  int q[8], dir[8];
  const int cn = neighbours_wc(P, v, q, dir);
  int nd = 0;
  int out[8];
  for (int s = 0; s < cn; s++) {
    const long long gu = (long long)P.off + q[s];
    const int nr = mnrec[gu];
    for (int k = 0; k < nr; k++)
      if (mrec[8 * gu + k] == v) {
        if (nd < 8) out[nd] = 8 * q[s] + k;
        nd++;
      }
  }
  if (nd > 8) {
    atomicOr(err + c.inst, FSX_ERR_MFD);
    nd = 8;
  }
  for (int k = 0; k < 8; k++) mdon[8 * g_ + k] = k < nd ? out[k] : -1;
  mndon[g_] = nd;
}

// ---- Kahn layers (one workgroup per landscape) -----------------------------------------------------
// The layer of a node is final when its last incoming edge is removed; positions inside a layer
// come from an LDS counter, so they vary from run to run while the layer sets do not.
__global__ __launch_bounds__(FSX_WG) void k_mfd_levels(const InstParams* __restrict__ PP, int* __restrict__ mindeg,
                                                       const int* __restrict__ mrec, const int* __restrict__ mnrec,
                                                       const uint8_t* __restrict__ mflag, const int* __restrict__ rec,
                                                       int* __restrict__ morder, int* __restrict__ mls,
                                                       int* __restrict__ mnlev, int* __restrict__ err) {
  __shared__ int s_tail;
  const int inst = blockIdx.x;
  const InstParams& P = PP[inst];
  if (!P.active || !P.run_spl || P.sfd) return;
  const int nn = P.nn, tid = threadIdx.x;
  const long long off = P.off;
  int* ls = mls + P.aux;  // nn+1 entries
  if (tid == 0) s_tail = 0;
  __syncthreads();
  for (int v = tid; v < nn; v += FSX_WG)
    if (mindeg[off + v] == 0) morder[off + atomicAdd(&s_tail, 1)] = v;
  __syncthreads();
  int lo = 0, hi = s_tail, L = 0;
  __syncthreads();
  while (lo < hi) {
    if (tid == 0) ls[L] = lo;
    L++;
    for (int p = lo + tid; p < hi; p += FSX_WG) {
      const int v = morder[off + p];
      const long long gv = off + v;
      const int nr = mnrec[gv];
      for (int k = 0; k <= nr; k++) {
        int t;
        if (k < nr) t = mrec[8 * gv + k];
        else if (mflag[gv]) t = rec[gv];
        else break;
        if (atomicSub(mindeg + off + t, 1) == 1) morder[off + atomicAdd(&s_tail, 1)] = t;
      }
    }
    __syncthreads();
    lo = hi;
    hi = s_tail;
    __syncthreads();
  }
  if (tid == 0) {
    ls[L] = hi;
    mnlev[inst] = L;
    if (hi != nn) atomicOr(err + inst, FSX_ERR_MFD);  // 'error in stack', src/FlowRouting.f90:292
  }
}

// ---- per-position records --------------------------------------------------------------------
// The layer sweeps run over positions of the Kahn order (mstack): mpos is a node's position; every
// per-node input a sweep needs is copied to its position by node-parallel kernels, so a sweep's only
// dependent reads are its neighbours' values of earlier layers.
__global__ __launch_bounds__(FSX_CHUNK) void k_mfd_pos(const InstParams* __restrict__ PP, const Chunk* __restrict__ chunks,
                                                       const int* __restrict__ morder, int* __restrict__ mpos) {
  MFD_PROLOGUE
  mpos[P.off + morder[g_]] = v;  // here v runs over positions
}

// donors (position, weight) and receivers (positions) of each position; own area dx*dy*precip
__global__ __launch_bounds__(FSX_CHUNK) void k_mfd_prep(const InstParams* __restrict__ PP, const Chunk* __restrict__ chunks,
                                                        const int* __restrict__ mpos, const int* __restrict__ rec,
                                                        const int* __restrict__ mrec, const int* __restrict__ mnrec,
                                                        const double* __restrict__ mw, const int* __restrict__ mdon,
                                                        const int* __restrict__ mndon, const double* __restrict__ precip,
                                                        double* __restrict__ A0, int* __restrict__ mdp,
                                                        double* __restrict__ mdw, int* __restrict__ dn,
                                                        int* __restrict__ rp) {
  MFD_PROLOGUE
  const long long off = P.off, gp = off + mpos[g_];
  A0[gp] = P.dxdy * precip[g_];
  const int nd = mndon[g_], nr = mnrec[g_];
  for (int j = 0; j < 8; j++) {
    int dpos = -1;
    double w = 0.0;
    if (j < nd) {
      const int e = mdon[8 * g_ + j];
      dpos = mpos[off + (e >> 3)];
      w = mw[8 * off + e];
    }
    mdp[8 * gp + j] = dpos;
    mdw[8 * gp + j] = w;
    dn[8 * gp + j] = j < nr ? mpos[off + mrec[8 * g_ + j]] : -1;
  }
  const int r = rec[g_];
  rp[gp] = r == v ? -1 : mpos[off + r];
}

// ---- accumulation (FlowAccumulation :97-119), layer by layer from the sources ----------------------
// Each position pulls its donors (earlier layers); the next layer's records are loaded while the
// current one is summed.
struct MfdAreaRec {
  double a0;
  int d[8];
  double w[8];
};
__device__ __forceinline__ void mfd_area_load(MfdAreaRec& r, long long gp, const double* A0, const int* mdp,
                                              const double* mdw) {
  r.a0 = A0[gp];
#pragma unroll
  for (int j = 0; j < 8; j++) {
    r.d[j] = mdp[8 * gp + j];
    r.w[j] = mdw[8 * gp + j];
  }
}

__global__ __launch_bounds__(FSX_WG) void k_mfd_area(const InstParams* __restrict__ PP, const int* __restrict__ mls,
                                                     const int* __restrict__ mnlev, const double* __restrict__ A0,
                                                     const int* __restrict__ mdp, const double* __restrict__ mdw,
                                                     double* __restrict__ Al) {
  const int inst = blockIdx.x;
  const InstParams& P = PP[inst];
  if (!P.active || !P.run_spl || P.sfd) return;
  const int tid = threadIdx.x;
  const long long off = P.off;
  const int* ls = mls + P.aux;
  const int L = mnlev[inst];
  MfdAreaRec nx;
  int b0 = ls[0], b1 = L > 0 ? ls[1] : 0;
  if (L > 0 && tid < b1 - b0) mfd_area_load(nx, off + b0 + tid, A0, mdp, mdw);
  for (int l = 0; l < L; l++) {
    const int c0 = b0, c1 = b1;
    MfdAreaRec cur = nx;
    if (l + 1 < L) {
      b0 = c1;
      b1 = ls[l + 2];
      if (tid < b1 - b0) mfd_area_load(nx, off + b0 + tid, A0, mdp, mdw);
    }
    for (int p = c0 + tid; p < c1; p += FSX_WG) {
      MfdAreaRec r;
      if (p == c0 + tid) r = cur;
      else mfd_area_load(r, off + p, A0, mdp, mdw);
      double a = r.a0;
#pragma unroll
      for (int j = 0; j < 8; j++)
        if (r.d[j] >= 0) a = a + Al[off + r.d[j]] * r.w[j];
      Al[off + p] = a;
    }
    mfd_barrier();
  }
}

// ---- per-position stream-power inputs (StreamPowerLaw :104-116 and :145-162) ----------------------
// n == 1: kf*dt*(a*w)**m/l where ht(ijk) >= ht(rcv); otherwise kf*dt*(a*w)**m/l**n where
// ht(ijk) > ht(rcv); -1 marks an edge that does not take part (constant over the Gauss-Seidel loop).
// Also: area back to grid order, ht / h / lake sill in position order, the elevation coefficient
// (g, NaN on fixed nodes: elev = ht there), and the landscape's maxval(abs(h)) and maxval(g).
__global__ __launch_bounds__(FSX_CHUNK) void k_mfd_fact(const InstParams* __restrict__ PP, const Chunk* __restrict__ chunks,
                                                        const int* __restrict__ mpos, const int* __restrict__ mrec,
                                                        const int* __restrict__ mnrec, const double* __restrict__ mw,
                                                        const uint8_t* __restrict__ mdir, const double* __restrict__ Al,
                                                        double* __restrict__ area, const double* __restrict__ kfint,
                                                        const double* __restrict__ ht, const double* __restrict__ g,
                                                        const uint8_t* __restrict__ bc, double* __restrict__ mfact,
                                                        double* __restrict__ HTl, double* __restrict__ Hl,
                                                        int* __restrict__ LSl, double* __restrict__ EC,
                                                        double* __restrict__ lsed, unsigned long long* __restrict__ redk) {
  __shared__ unsigned long long smax[2][FSX_CHUNK / 64];
  const Chunk c = chunks[chunk_id()];
  const InstParams& P = PP[c.inst];
  const int v = c.start + (int)threadIdx.x;
  if (!P.active || !P.run_spl || P.sfd) return;  // uniform over the block
  const bool live = v < P.nn;
  const long long g_ = (long long)P.off + v;
  unsigned long long kh = live ? dkey(fabs(ht[g_])) : 0ull, kg = live ? dkey(g[g_]) : 0ull;
  for (int o = 32; o > 0; o >>= 1) {
    const unsigned long long x = __shfl_xor(kh, o, 64), y = __shfl_xor(kg, o, 64);
    kh = x > kh ? x : kh;
    kg = y > kg ? y : kg;
  }
  if ((threadIdx.x & 63) == 0) { smax[0][threadIdx.x >> 6] = kh; smax[1][threadIdx.x >> 6] = kg; }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < FSX_CHUNK / 64; w++) {
      kh = smax[0][w] > kh ? smax[0][w] : kh;
      kg = smax[1][w] > kg ? smax[1][w] : kg;
    }
    atomicMax(redk + 2 * c.inst, kh);
    atomicMax(redk + 2 * c.inst + 1, kg);
  }
  if (!live) return;
  const long long gp = (long long)P.off + mpos[g_];
  const double a = Al[gp], t = ht[g_];
  area[g_] = a;
  HTl[gp] = t;
  Hl[gp] = t;
  LSl[gp] = -1;
  lsed[g_] = 0.0;
  EC[gp] = bc[g_] ? __longlong_as_double(0x7ff8000000000000ll) : g[g_];
  const int nr = mnrec[g_];
  const double kd = kfint[g_] * P.dt;
  const bool n_is_one = fabs(P.n - 1.0) < 2.2250738585072014e-308;
  for (int k = 0; k < 8; k++) {
    double f = -1.0;
    if (k < nr) {
      const double tr = ht[P.off + mrec[8 * g_ + k]];
      const double l = P.len[mdir[8 * g_ + k]];
      if (n_is_one) {
        if (t >= tr) f = kd * pow(a * mw[8 * g_ + k], P.m) / l;
      } else {
        if (t > tr) f = kd * pow(a * mw[8 * g_ + k], P.m) / pow(l, P.n);
      }
    }
    mfact[8 * gp + k] = f;
  }
}

// ---- stream power, Gauss-Seidel loop (one workgroup per landscape) --------------------------------
// Per iteration: hp = h; erosion pushed from the sources (dh(rcv) += dh(ijk)*w, :62-82) with the
// lake-sediment update of lake sills; elev (:84-90); solve from the last layer back (:94-176):
// closed form over the receivers for n == 1, relaxed Newton (omega = 0.875/n) otherwise; lakes keep
// elev and collect their water volume at their sill; err = rms(h-hp), zero when g == 0 everywhere.
// Everything runs on positions; node-indexed water, lake sill and sediment are kept for the
// lake fill of k_spl_post.
struct MfdSolveRec {
  int v, rp;
  double t, e, h;
  int q[8];
  double f[8];
};
__device__ __forceinline__ void mfd_solve_load(MfdSolveRec& r, long long off, int p, const int* morder, const int* rp,
                                               const double* HTl, const double* ELl, const double* Hl, const int* dn,
                                               const double* mfact, bool gzero) {
  const long long gp = off + p;
  r.v = morder[gp];
  r.rp = rp[gp];
  r.t = HTl[gp];
  r.e = gzero ? r.t : ELl[gp];
  r.h = Hl[gp];
#pragma unroll
  for (int k = 0; k < 8; k++) {
    r.q[k] = dn[8 * gp + k];
    r.f[k] = mfact[8 * gp + k];
  }
}

__global__ __launch_bounds__(FSX_WG) void k_mfd_spl(
    const InstParams* __restrict__ PP, const int* __restrict__ mls, const int* __restrict__ mnlev,
    const int* __restrict__ morder, const int* __restrict__ rp, const int* __restrict__ dn,
    const double* __restrict__ mfact, const int* __restrict__ mdp, const double* __restrict__ mdw,
    const double* __restrict__ EC, const double* __restrict__ Al, const double* __restrict__ HTl,
    double* __restrict__ Hl, double* __restrict__ HPl, double* __restrict__ DHl, double* __restrict__ ELl,
    double* __restrict__ Wl, int* __restrict__ LSl, double* __restrict__ water, int* __restrict__ lsill,
    double* __restrict__ lwv, double* __restrict__ lsed, const unsigned long long* __restrict__ redk,
    double* __restrict__ scal, int* __restrict__ ngs_out, long long* __restrict__ newton_out, int* __restrict__ err) {
  __shared__ double red[FSX_WG / 64];
  __shared__ double s_tol, s_err, s_gmax;
  __shared__ unsigned long long s_newton;
  const int inst = blockIdx.x;
  const InstParams& P = PP[inst];
  if (!P.active || !P.run_spl || P.sfd) return;
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  const int nn = P.nn;
  const long long off = P.off;
  const int* ls = mls + P.aux;
  const int L = mnlev[inst];
  const double n = P.n;
  const bool n_is_one = fabs(n - 1.0) < 2.2250738585072014e-308;
  const int nkind = npow_kind(n);
  if (tid == 0) {
    s_tol = 1.0e-4 * (dval(redk[2 * inst]) + 1.0);
    s_gmax = dval(redk[2 * inst + 1]);
    s_err = 2.0 * s_tol;
    s_newton = 0;
  }
  __syncthreads();
  const bool gzero = s_gmax < 2.2250738585072014e-308;
  int ngs = 0;
  unsigned long long my_newton = 0;
  while (s_err > s_tol && ngs < 99) {
    ngs++;
    for (int p = tid; p < nn; p += FSX_WG) HPl[off + p] = Hl[off + p];
    __syncthreads();
    if (!gzero) {
      for (int l = 0; l < L; l++) {
        const int b0 = ls[l], b1 = ls[l + 1];
        for (int p = b0 + tid; p < b1; p += FSX_WG) {
          const long long gp = off + p;
          const int v = morder[gp];
          const double own = HTl[gp] - HPl[gp];
          double s = own;
#pragma unroll
          for (int j = 0; j < 8; j++) {
            const int dpos = mdp[8 * gp + j];
            if (dpos >= 0) s = s + DHl[off + dpos] * mdw[8 * gp + j];
          }
          double dv = s;
          if (rp[gp] >= 0) {
            const double tt = s - own;
            if (LSl[gp] == v) lsed[off + v] = tt <= 0.0 ? 0.0 : tt;
            dv = tt + own;
          } else {
            lsed[off + v] = s;
          }
          DHl[gp] = dv;
        }
        mfd_barrier();
      }
      for (int p = tid; p < nn; p += FSX_WG) {
        const long long gp = off + p;
        const double t = HTl[gp], gc = EC[gp];
        ELl[gp] = isnan(gc) ? t : t + (DHl[gp] - (t - HPl[gp])) * gc * P.dx * P.dy / Al[gp];
      }
      __syncthreads();
    }
    // solve, last layer first; the next layer's records are loaded while this one is solved
    MfdSolveRec nxr;
    int b0 = L > 0 ? ls[L - 1] : 0, b1 = L > 0 ? ls[L] : 0;
    if (L > 0 && tid < b1 - b0) mfd_solve_load(nxr, off, b0 + tid, morder, rp, HTl, ELl, Hl, dn, mfact, gzero);
    for (int l = L - 1; l >= 0; l--) {
      const int c0 = b0, c1 = b1;
      MfdSolveRec cur = nxr;
      if (l > 0) {
        b1 = c0;
        b0 = ls[l - 1];
        if (tid < b1 - b0) mfd_solve_load(nxr, off, b0 + tid, morder, rp, HTl, ELl, Hl, dn, mfact, gzero);
      }
      for (int p = c0 + tid; p < c1; p += FSX_WG) {
        MfdSolveRec r;
        if (p == c0 + tid) r = cur;
        else mfd_solve_load(r, off, p, morder, rp, HTl, ELl, Hl, dn, mfact, gzero);
        const long long gp = off + p;
        const int v = r.v;
        if (r.rp < 0) {
          Wl[gp] = r.t;
          LSl[gp] = v;
          if (!gzero) {
            lwv[off + v] = 0.0;
            water[off + v] = r.t;
            lsill[off + v] = v;
          }
          continue;
        }
        const double w_rcv = Wl[off + r.rp];
        double hv = r.h;
        int sl;
        if (r.e > w_rcv) {
          // with Marine on, nodes below sea level keep their current h (src/StreamPowerLaw.f90:105,144:
          // h for n == 1, ht otherwise)
          if (r.q[0] >= 0 && !(P.run_marine && !((n_is_one ? r.h : r.t) >= P.sealevel))) {
            double hr[8];
#pragma unroll
            for (int k = 0; k < 8; k++) hr[k] = r.q[k] >= 0 ? Hl[off + r.q[k]] : 0.0;
            if (n_is_one) {
              double f = r.e, df = 1.0;
#pragma unroll
              for (int k = 0; k < 8; k++)
                if (r.q[k] >= 0 && r.f[k] >= 0.0) {
                  f = f + r.f[k] * hr[k];
                  df = df + r.f[k];
                }
              hv = f / df;
            } else {
              const double omega = 0.875 / n, tolp = P.newton_tol;
              double errp = 2.0 * tolp;
              const double h0 = r.e;
              int it = 0;
              while (errp > tolp) {
                double f = hv - h0, df = 1.0;
#pragma unroll
                for (int k = 0; k < 8; k++)
                  if (r.q[k] >= 0 && r.f[k] >= 0.0) {
                    const double x = fmax(0.0, hv - hr[k]);
                    double pn, pn1;
                    npow(x, n, nkind, pn, pn1);
                    f = f + r.f[k] * pn;
                    df = df + r.f[k] * n * pn1;
                  }
                const double hn = hv - f / df;
                errp = fabs(hn - hv);
                hv = hv * (1.0 - omega) + hn * omega;
                if (++it > 100000) { atomicOr(err + inst, FSX_ERR_NEWTON); break; }
              }
              my_newton += it;
            }
          }
          sl = v;
          if (!gzero) lwv[off + v] = 0.0;
          if (hv < w_rcv) hv = w_rcv;
        } else {
          hv = r.e;
          sl = LSl[off + r.rp];
          if (!gzero && sl >= 0) atomicAdd(lwv + off + sl, w_rcv - hv);
        }
        const double wv = w_rcv > hv ? w_rcv : hv;
        Hl[gp] = hv;
        Wl[gp] = wv;
        LSl[gp] = sl;
        if (!gzero) {
          water[off + v] = wv;
          lsill[off + v] = sl;
        }
      }
      mfd_barrier();
    }
    if (gzero) {
      if (tid == 0) s_err = 0.0;
    } else {
      double ss = 0.0;
      for (int p = tid; p < nn; p += FSX_WG) {
        const double d = Hl[off + p] - HPl[off + p];
        ss += d * d;
      }
      for (int o = 32; o > 0; o >>= 1) ss += __shfl_xor(ss, o, 64);
      if (lane == 0) red[wid] = ss;
      __syncthreads();
      if (tid == 0) {
        double x = 0.0;
        for (int q = 0; q < FSX_WG / 64; q++) x += red[q];
        s_err = sqrt(x / (double)nn);
      }
    }
    __syncthreads();
  }
  if (my_newton) atomicAdd(&s_newton, my_newton);
  __syncthreads();
  if (tid == 0) {
    ngs_out[inst] = ngs;
    newton_out[inst] = (long long)s_newton;
    scal[inst] = gzero ? 0.0 : 1.0;
  }
}

// ---------------------------------------------------------------------------------------------------
void launch_mfd_route(DevState& S, hipStream_t st) {
  k_mfd_fill<<<S.ninst, FSX_WG, 0, st>>>(S.P, S.lstart, S.nlev, S.order, S.rec, S.ht, S.hwater, S.lake_depth);
  launch_fill_int(S.mindeg, S.ntot, 0, st);
  k_mfd_rec<<<S.nchunks, FSX_CHUNK, 0, st>>>(S.P, S.chunks, S.hwater, S.rec, S.mrec, S.mw, S.mdir, S.mnrec, S.mflag,
                                              S.mindeg, S.pexp);
  k_mfd_don<<<S.nchunks, FSX_CHUNK, 0, st>>>(S.P, S.chunks, S.mrec, S.mnrec, S.mdon, S.mndon, S.err);
  k_mfd_levels<<<S.ninst, FSX_WG, 0, st>>>(S.P, S.mindeg, S.mrec, S.mnrec, S.mflag, S.rec, S.mstack, S.mls, S.mnlev,
                                            S.err);
  k_mfd_pos<<<S.nchunks, FSX_CHUNK, 0, st>>>(S.P, S.chunks, S.mstack, S.mpos);
  k_mfd_prep<<<S.nchunks, FSX_CHUNK, 0, st>>>(S.P, S.chunks, S.mpos, S.rec, S.mrec, S.mnrec, S.mw, S.mdon, S.mndon,
                                               S.precip, S.A0, S.mdp, S.mdw, S.dn, S.rp);
  k_mfd_area<<<S.ninst, FSX_WG, 0, st>>>(S.P, S.mls, S.mnlev, S.A0, S.mdp, S.mdw, S.Al);
}

void launch_mfd_spl(DevState& S, hipStream_t st) {
  launch_fill_int(reinterpret_cast<int*>(S.redk), 4LL * S.ninst, 0, st);
  k_mfd_fact<<<S.nchunks, FSX_CHUNK, 0, st>>>(S.P, S.chunks, S.mpos, S.mrec, S.mnrec, S.mw, S.mdir, S.Al, S.area,
                                               S.kfint, S.ht, S.g, S.bc, S.mfact, S.HTl, S.Hl, S.LSl, S.F, S.lsed, S.redk);
  k_mfd_spl<<<S.ninst, FSX_WG, 0, st>>>(S.P, S.mls, S.mnlev, S.mstack, S.rp, S.dn, S.mfact, S.mdp, S.mdw, S.F, S.Al,
                                         S.HTl, S.Hl, S.HPl, S.DHl, S.ELl, S.Wl, S.LSl, S.water, S.lsill, S.lwv, S.lsed,
                                         S.redk, S.scal, S.ngs, S.newton, S.err);
}
