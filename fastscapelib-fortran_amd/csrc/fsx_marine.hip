// fsx_marine.hip -- marine sediment transport after the landscape step (Marine, src/Marine.f90:1-167, with
// SiltSandCouplingDiffusion, :171-499).
//
// The reference, after Diffusion:
//  1. sums the sediment flux that reaches the sea along the step's single-flow receivers (the land part of
//     each node's upstream Sedflux, stopping at the first node below sea level), decompacted;
//  2. routes it across the shelf along multiple-flow receivers of the current h (find_mult_rec with
//     p = 0, so weights slope**p_mfd_exp, on h filled along the step's single-flow forest): a shelf
//     depth descending at 1e-4 from the coast, and a flux that fills each sea node up to that depth
//     before passing the rest on;
//  3. diffuses silt and sand in the ocean (coupled alternating-direction solves of h and the silt
//     fraction, repeated until max|h - hp| <= 1 m);
//  4. books the elevation change in etot / erate, clears Sedflux, etot and erate below sea level, zeroes
//     the silt fraction on land, clamps it to [0, 1] and lowers the basement to h.
// Its compaction block (:120-146) computes arrays nothing reads afterwards and is not restated.
//
// Here: (1) and the fill of (2) run as single-flow level sweeps (one workgroup per landscape, levels
// of the step's routing), the receivers of (2) are node-parallel, the multiple-flow order is a Kahn
// layering of those receivers (one workgroup per landscape), and the shelf depth and flux pull from
// each node's donors layer by layer.  The shelf depth is a min / max over donors, exact in any order;
// the flux sums donors in direction order where the reference follows its stack, so it agrees to
// rounding.  (3) is one workgroup per landscape looping over its iterations, one thread per grid line
// and pass, each line solved by the reference's Thomas recurrence (tridag, src/Diffusion.f90:164-202).
#include "fsx_state.h"

#define MAR_NODE_PROLOGUE                                     \
  const Chunk c = chunks[chunk_id()];                         \
  const InstParams& P = PP[c.inst];                           \
  const int v = c.start + (int)threadIdx.x;                   \
  if (!P.active || !P.run_marine || v >= P.nn) return;        \
  const long long g_ = (long long)P.off + v;

static __device__ __forceinline__ double dmin_(double a, double b) { return a < b ? a : b; }
static __device__ __forceinline__ double dmax_(double a, double b) { return a > b ? a : b; }

// ---- (1) flux to the sea along the single-flow receivers; fill of (2) -------------------------------
// Levels of the step's receiver forest from the sources down (a node sums its donors, :28-36), then land
// nodes zeroed and the decompaction factor applied (:41-50); then the filled surface h0 from base level
// up (find_mult_rec :187-196).  Also clears the landscape's multiple-flow in-degrees.
__global__ __launch_bounds__(FSX_WG) void k_mar_sfd(const InstParams* __restrict__ PP, const int* __restrict__ lstart,
                                                    const int* __restrict__ nlev, const int* __restrict__ order,
                                                    const int* __restrict__ rec, const uint8_t* __restrict__ dmask,
                                                    const double* __restrict__ h, const double* __restrict__ sedflux,
                                                    double* __restrict__ flux, double* __restrict__ h0,
                                                    int* __restrict__ indeg) {
  const int inst = blockIdx.x;
  const InstParams& P = PP[inst];
  if (!P.active || !P.run_marine) return;
  const long long off = P.off;
  const int nn = P.nn, tid = threadIdx.x;
  const int* ls = lstart + P.aux;
  const int L = nlev[inst];
  const double sl = P.sealevel;
  for (int v = tid; v < nn; v += FSX_WG) indeg[off + v] = 0;
  for (int l = L - 1; l >= 0; l--) {
    for (int p = ls[l] + tid; p < ls[l + 1]; p += FSX_WG) {
      const int v = order[off + p];
      double f = h[off + v] > sl ? sedflux[off + v] : 0.0;
      int d[8];
      const int nd = donors_sorted(P, v, dmask[off + v], d);
      for (int k = 0; k < nd; k++)
        if (h[off + d[k]] > sl) f = f + flux[off + d[k]];
      flux[off + v] = f;
    }
    __syncthreads();
  }
  const double r1 = P.ratio / (1.0 - P.poro1), r2 = (1.0 - P.ratio) / (1.0 - P.poro2);
  for (int v = tid; v < nn; v += FSX_WG) {
    const double f = h[off + v] > sl ? 0.0 : flux[off + v];
    flux[off + v] = f * (r1 + r2);
  }
  for (int l = 0; l < L; l++) {
    for (int p = ls[l] + tid; p < ls[l + 1]; p += FSX_WG) {
      const int v = order[off + p];
      const double hv = h[off + v];
      double hw = hv;
      if (l > 0) {
        const double hr = h0[off + rec[off + v]];
        if (hr > hv) hw = hr + 1.0e-8;
      }
      h0[off + v] = hw;
    }
    __syncthreads();
  }
}

// ---- (2) multiple-flow receivers of h0 (find_mult_rec :204-243 with p = 0) --------------------------
// Every strictly lower neighbour in scan order (wrap if cyclic else clamp; a clamped direction that lands
// on another node repeats it), weight ((h0 - h0(r)) / l)**p_mfd_exp normalised: p_mfd_exp is the p given
// to the erosional parameters for single flow (which keeps it), the step's per-node exponent for multiple
// flow (p, or the mean-slope exponent when p < 0).
__global__ __launch_bounds__(FSX_CHUNK) void k_mar_rec(const InstParams* __restrict__ PP, const Chunk* __restrict__ chunks,
                                                       const double* __restrict__ h0, const double* __restrict__ pexp,
                                                       int* __restrict__ mrec, double* __restrict__ mw,
                                                       uint8_t* __restrict__ mdir, int* __restrict__ mnrec,
                                                       int* __restrict__ indeg) {
  MAR_NODE_PROLOGUE
  const double* hi = h0 + P.off;
  const int i = v % P.nx, j = v / P.nx;
  int t[8], dd[8];
  double w[8];
  int nr = 0;
  if (i >= P.i1 && i <= P.i2 && j >= P.j1 && j <= P.j2) {
    const double hv = hi[v];
    for (int jj = -1; jj <= 1; jj++)
      for (int ii = -1; ii <= 1; ii++) {
        const int q = nbr_wc(P, i, j, ii, jj);
        if (hv > hi[q]) {
          const int d = (ii + 1) + 3 * (jj + 1);
          t[nr] = q;
          dd[nr] = d;
          w[nr] = (hv - hi[q]) / P.len[d];
          nr++;
        }
      }
  }
  const double pe = P.sfd ? P.p_raw : pexp[g_];
  if (pe != 1.0)  // x**1 is x exactly
    for (int k = 0; k < nr; k++) w[k] = pow(w[k], pe);
  double sw = 0.0;
  for (int k = 0; k < nr; k++) sw = sw + w[k];
  for (int k = 0; k < 8; k++) {
    const bool on = k < nr;
    mrec[8 * g_ + k] = on ? t[k] : -1;
    mw[8 * g_ + k] = on ? w[k] / sw : 0.0;
    mdir[8 * g_ + k] = on ? (uint8_t)dd[k] : (uint8_t)4;
    if (on) atomicAdd(indeg + P.off + t[k], 1);
  }
  mnrec[g_] = nr;
}

// ---- multiple-flow order: Kahn layers (layer 0: no donors; a node joins the layer after its last donor's)
__global__ __launch_bounds__(FSX_WG) void k_mar_levels(const InstParams* __restrict__ PP, int* __restrict__ indeg,
                                                       const int* __restrict__ mrec, const int* __restrict__ mnrec,
                                                       int* __restrict__ morder, int* __restrict__ mls,
                                                       int* __restrict__ mnlev, int* __restrict__ err) {
  __shared__ int s_tail;
  const int inst = blockIdx.x;
  const InstParams& P = PP[inst];
  if (!P.active || !P.run_marine) return;
  const int nn = P.nn, tid = threadIdx.x;
  const long long off = P.off;
  int* ls = mls + P.aux;
  if (tid == 0) s_tail = 0;
  __syncthreads();
  for (int v = tid; v < nn; v += FSX_WG)
    if (indeg[off + v] == 0) morder[off + atomicAdd(&s_tail, 1)] = v;
  __syncthreads();
  int lo = 0, hi = s_tail, L = 0;
  __syncthreads();
  while (lo < hi) {
    if (tid == 0) ls[L] = lo;
    L++;
    for (int p = lo + tid; p < hi; p += FSX_WG) {
      const long long gv = off + morder[off + p];
      const int nr = mnrec[gv];
      for (int k = 0; k < nr; k++) {
        const int t = mrec[8 * gv + k];
        if (atomicSub(indeg + off + t, 1) == 1) morder[off + atomicAdd(&s_tail, 1)] = t;
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

// ---- shelf depth and shelf flux (:64-92), layer by layer from the sources ---------------------------
// A node reads the donors that list it as receiver (each slot that does, so a repeated clamp target
// counts twice as in the reference): shelf depth = sea level, lowered along every donor below sea level
// to its depth + l * (-1e-4) and held at or above the node's h (min then max per donor, as :73-74); then
// its flux: own flux + (h - shelf depth) below sea level, plus every donor's positive share.
__global__ __launch_bounds__(FSX_WG) void k_mar_shelf(const InstParams* __restrict__ PP, const int* __restrict__ mls,
                                                      const int* __restrict__ mnlev, const int* __restrict__ morder,
                                                      const int* __restrict__ mrec, const double* __restrict__ mw,
                                                      const uint8_t* __restrict__ mdir, const int* __restrict__ mnrec,
                                                      const double* __restrict__ h, double* __restrict__ sd,
                                                      double* __restrict__ flux) {
  const int inst = blockIdx.x;
  const InstParams& P = PP[inst];
  if (!P.active || !P.run_marine) return;
  const long long off = P.off;
  const int* ls = mls + P.aux;
  const int L = mnlev[inst];
  const double sl = P.sealevel, shelfslope = -1.0e-4;
  for (int l = 0; l < L; l++) {
    for (int p = ls[l] + (int)threadIdx.x; p < ls[l + 1]; p += FSX_WG) {
      const int v = morder[off + p];
      const double hv = h[off + v];
      int q[8], dir[8];
      const int c8 = neighbours_wc(P, v, q, dir);
      double s = sl;
      for (int a = 0; a < c8; a++) {
        const long long gu = off + q[a];
        if (!(h[gu] < sl)) continue;
        const int nr = mnrec[gu];
        for (int k = 0; k < nr; k++)
          if (mrec[8 * gu + k] == v) {
            s = dmin_(s, sd[gu] + P.len[mdir[8 * gu + k]] * shelfslope);
            s = dmax_(s, hv);
          }
      }
      sd[off + v] = s;
      double f = flux[off + v];
      if (hv < sl) f = f + (hv - s);
      for (int a = 0; a < c8; a++) {
        const long long gu = off + q[a];
        const int nr = mnrec[gu];
        for (int k = 0; k < nr; k++)
          if (mrec[8 * gu + k] == v) f = f + dmax_(0.0, flux[gu] * mw[8 * gu + k]);
      }
      flux[off + v] = f;
    }
    __syncthreads();
  }
}

// ---- deposit rule and supply (:94-117); working copies of the diffusion ---------------------------
__global__ __launch_bounds__(FSX_CHUNK) void k_mar_prep(const InstParams* __restrict__ PP, const Chunk* __restrict__ chunks,
                                                        const double* __restrict__ h, const double* __restrict__ fmix,
                                                        const double* __restrict__ sd, const double* __restrict__ flux,
                                                        double* __restrict__ q1, double* __restrict__ qs,
                                                        double* __restrict__ ht, double* __restrict__ ft,
                                                        double* __restrict__ hh, double* __restrict__ fh) {
  MAR_NODE_PROLOGUE
  const double sl = P.sealevel, hv = h[g_], s = sd[g_];
  double f = flux[g_];
  if (f > 0.0 && hv < sl) f = -(hv - s);
  if (f <= 0.0 && hv < sl) f = f - (hv - s);
  if (hv >= sl) f = 0.0;
  f = dmax_(f, 0.0);
  const double r1 = P.ratio / (1.0 - P.poro1), r2 = (1.0 - P.ratio) / (1.0 - P.poro2);
  const double fs = f > 0.0 ? r1 / (r1 + r2) : 0.0;
  f = f / P.dt;
  const double a = f * fs, b = f * (1.0 - fs);
  q1[g_] = a;
  qs[g_] = a + b;
  ht[g_] = hv;
  hh[g_] = hv;
  const double fv = fmix[g_];
  ft[g_] = fv;
  fh[g_] = fv;
}

// ---- (3) silt / sand diffusion ---------------------------------------------------------------------
// One line of a tridiagonal system by the reference's Thomas recurrence: coef(k, a, b, c, r) gives row k
// (a: below the diagonal, b: diagonal, c: above, r: right-hand side); the solution goes to out[] along
// the line and gam[] holds the recurrence's factors.  false on a zero pivot.
template <typename Coef>
__device__ __forceinline__ bool thomas_line(int n, long long base, long long stride, double* out, double* gam,
                                            Coef&& coef) {
  double a, b, c, r;
  coef(0, a, b, c, r);
  if (b == 0.0) return false;
  double bet = b;
  double u = r / bet;
  out[base] = u;
  double cprev = c;
  for (int k = 1; k < n; k++) {
    coef(k, a, b, c, r);
    const double gk = cprev / bet;
    gam[base + k * stride] = gk;
    bet = b - a * gk;
    if (bet == 0.0) return false;
    u = (r - a * u) / bet;
    out[base + k * stride] = u;
    cprev = c;
  }
  for (int k = n - 2; k >= 0; k--) {
    u = out[base + k * stride] - gam[base + (k + 1) * stride] * u;
    out[base + k * stride] = u;
  }
  return true;
}

struct SiltArgs {
  double *h, *f;                      // state: elevation and silt fraction (in place)
  const double *q1, *qs;              // silt supply, silt + sand supply
  double *ht, *ft, *hp, *fp;          // at entry; at the start of the iteration
  double *hh, *fh, *fhp, *tint, *gam; // half-step h and f, f half-step of the previous iteration, scratch
  int* niter;                         // per landscape
  int* err;
};

__global__ __launch_bounds__(FSX_WG) void k_mar_silt(const InstParams* __restrict__ PP, SiltArgs A) {
  __shared__ double red[FSX_WG / 64];
  __shared__ double s_err;
  __shared__ int s_bad;
  const int inst = blockIdx.x;
  const InstParams& P = PP[inst];
  if (!P.active || !P.run_marine) return;
  const int nx = P.nx, ny = P.ny, nn = P.nn, tid = threadIdx.x;
  const long long off = P.off;
  double *h = A.h + off, *f = A.f + off, *ht = A.ht + off, *ft = A.ft + off, *hp = A.hp + off, *fp = A.fp + off;
  double *hh = A.hh + off, *fh = A.fh + off, *fhp = A.fhp + off, *tint = A.tint + off, *gam = A.gam + off;
  const double *q1 = A.q1 + off, *qs = A.qs + off;
  const double dt = P.dt, sl = P.sealevel, L = P.layer, K1 = P.kdsea1, K2 = P.kdsea2, K12 = K1 - K2;
  const double hdt = dt / 2.0, dx2 = P.dx * P.dx, dy2 = P.dy * P.dy, K1h = K1 / 2.0, L2 = 2.0 * L / dt;
  const double dep = 1.0 + 1.0e-6;
  const double tol = 1.0;
  if (tid == 0) {
    s_err = 2 * tol;
    s_bad = 0;
  }
  __syncthreads();
  int niter = 0;
  while (s_err > tol) {
    niter++;
    for (int p = tid; p < nn; p += FSX_WG) {
      hp[p] = h[p];
      fp[p] = f[p];
      fhp[p] = fh[p];
    }
    __syncthreads();
    bool ok = true;
    // h along x, rows 2..ny-1 (:215-277)
    for (int j = 1 + tid; j < ny - 1; j += FSX_WG) {
      const long long base = (long long)j * nx;
      ok &= thomas_line(nx, base, 1, hh, gam, [&](int i, double& a, double& b, double& c, double& r) {
        const long long ij = base + i;
        const double t = ht[ij];
        a = 0.0; b = 1.0; c = 0.0; r = t;
        if (!(t <= sl)) return;
        double Ap = 0.0, Bp = 0.0;
        if (i == 0) {
          if (P.fix_left) return;
          Ap = hdt * (K2 + K12 * (fhp[ij + 1] + fhp[ij]) / 2.0) / dx2;
          b = 1.0 + Ap;
          c = -Ap;
        } else if (i == nx - 1) {
          if (P.fix_right) return;
          Bp = -(hdt * (K2 + K12 * (fhp[ij] + fhp[ij - 1]) / 2.0) / dx2);
          b = 1.0 - Bp;
          a = Bp;
        } else {
          Ap = hdt * (K2 + K12 * (fhp[ij + 1] + fhp[ij]) / 2.0) / dx2;
          Bp = -(hdt * (K2 + K12 * (fhp[ij] + fhp[ij - 1]) / 2.0) / dx2);
          b = 1.0 + Ap - Bp;
          c = -Ap;
          a = Bp;
        }
        const long long ijp = ij + nx, ijm = ij - nx;
        const double Cp = hdt * (K2 + K12 * (ft[ijp] + ft[ij]) / 2.0) * (ht[ijp] - t) / dy2 -
                          hdt * (K2 + K12 * (ft[ij] + ft[ijm]) / 2.0) * (t - ht[ijm]) / dy2 + qs[ij] * dt / 2.0;
        r = Cp + t;
      });
    }
    __syncthreads();
    for (int p = tid; p < nn; p += FSX_WG) tint[p] = hh[p];
    __syncthreads();
    if (tid == 0) {
      hh[0] = hh[1];
      hh[(long long)(ny - 1) * nx] = hh[(long long)(ny - 1) * nx + 1];
      hh[nx - 1] = hh[nx - 2];
      hh[nn - 1] = hh[nn - 2];
    }
    __syncthreads();
    // silt fraction along x, rows 2..ny-1 (:287-340)
    for (int j = 1 + tid; j < ny - 1; j += FSX_WG) {
      const long long base = (long long)j * nx;
      ok &= thomas_line(nx, base, 1, fh, gam, [&](int i, double& a, double& b, double& c, double& r) {
        const long long ij = base + i;
        if (i == 0) { a = 0.0; b = 1.0; c = -1.0; r = 0.0; return; }
        if (i == nx - 1) { a = -1.0; b = 1.0; c = 0.0; r = 0.0; return; }
        const double t = ht[ij], hv = hh[ij];
        a = 0.0; b = 1.0; c = 0.0; r = ft[ij];
        if (!(t <= sl && hv >= dep * t)) return;
        const long long ijp = ij + nx, ijm = ij - nx;
        const double Dp = (hv - t) / dt;
        const double Ep = K1h * (hh[ij + 1] - hv) / dx2;
        const double Mp = -(K1h * (hv - hh[ij - 1]) / dx2);
        const double Np = K1h * (ft[ijp] + ft[ij]) * (ht[ijp] - t) / dy2 - K1h * (ft[ij] + ft[ijm]) * (t - ht[ijm]) / dy2 +
                          q1[ij];
        b = L2 + Dp - Mp - Ep;
        c = -Ep;
        a = -Mp;
        r = Np - Dp * ft[ij] + 2.0 * L * ft[ij] / dt;
      });
    }
    __syncthreads();
    for (int p = tid; p < nn; p += FSX_WG) fh[p] = dmin_(dmax_(0.0, fh[p]), 1.0);
    __syncthreads();
    // h along y, columns 2..nx-1 (:345-405)
    for (int i = 1 + tid; i < nx - 1; i += FSX_WG) {
      ok &= thomas_line(ny, i, nx, tint, gam, [&](int j, double& a, double& b, double& c, double& r) {
        const long long ij = (long long)j * nx + i;
        const double t = hh[ij];
        a = 0.0; b = 1.0; c = 0.0; r = t;
        if (!(ht[ij] <= sl)) return;
        double Ap = 0.0, Bp = 0.0;
        if (j == 0) {
          if (P.fix_bottom) return;
          Ap = hdt * (K2 + K12 * (fp[ij + nx] + fp[ij]) / 2.0) / dy2;
          b = 1.0 + Ap;
          c = -Ap;
        } else if (j == ny - 1) {
          if (P.fix_top) return;
          Bp = -(hdt * (K2 + K12 * (fp[ij] + fp[ij - nx]) / 2.0) / dy2);
          b = 1.0 - Bp;
          a = Bp;
        } else {
          Ap = hdt * (K2 + K12 * (fp[ij + nx] + fp[ij]) / 2.0) / dy2;
          Bp = -(hdt * (K2 + K12 * (fp[ij] + fp[ij - nx]) / 2.0) / dy2);
          b = 1.0 + Ap - Bp;
          c = -Ap;
          a = Bp;
        }
        const double Cp = hdt * (K2 + K12 * (fh[ij + 1] + fh[ij]) / 2.0) * (hh[ij + 1] - t) / dx2 -
                          hdt * (K2 + K12 * (fh[ij] + fh[ij - 1]) / 2.0) * (t - hh[ij - 1]) / dx2 + qs[ij] * dt / 2.0;
        r = Cp + t;
      });
    }
    __syncthreads();
    for (int p = tid; p < nn; p += FSX_WG) h[p] = tint[p];
    __syncthreads();
    if (tid == 0) {
      h[0] = h[1];
      h[(long long)(ny - 1) * nx] = h[(long long)(ny - 1) * nx + 1];
      h[nx - 1] = h[nx - 2];
      h[nn - 1] = h[nn - 2];
    }
    __syncthreads();
    // silt fraction along y, columns 2..nx-1 (:415-468)
    for (int i = 1 + tid; i < nx - 1; i += FSX_WG) {
      ok &= thomas_line(ny, i, nx, f, gam, [&](int j, double& a, double& b, double& c, double& r) {
        const long long ij = (long long)j * nx + i;
        if (j == 0) { a = 0.0; b = 1.0; c = -1.0; r = 0.0; return; }
        if (j == ny - 1) { a = -1.0; b = 1.0; c = 0.0; r = 0.0; return; }
        const double t = hh[ij], hv = h[ij];
        a = 0.0; b = 1.0; c = 0.0; r = fh[ij];
        if (!(ht[ij] <= sl && hv >= dep * t)) return;
        const double Dp = (hv - t) / dt;
        const double Ep = K1h * (h[ij + nx] - hv) / dy2;
        const double Mp = -(K1h * (hv - h[ij - nx]) / dy2);
        const double Np = K1h * (fh[ij + 1] + fh[ij]) * (hh[ij + 1] - t) / dx2 -
                          K1h * (fh[ij] + fh[ij - 1]) * (t - hh[ij - 1]) / dx2 + q1[ij];
        b = L2 + Dp - Mp - Ep;
        c = -Ep;
        a = -Mp;
        r = Np - Dp * fh[ij] + 2.0 * L * fh[ij] / dt;
      });
    }
    if (!ok) s_bad = 1;  // benign race: every writer stores 1
    __syncthreads();
    double e = 0.0;
    for (int p = tid; p < nn; p += FSX_WG) {
      f[p] = dmin_(dmax_(0.0, f[p]), 1.0);
      e = dmax_(e, fabs(h[p] - hp[p]));
    }
    for (int o = 32; o > 0; o >>= 1) e = dmax_(e, __shfl_xor(e, o, 64));
    if ((tid & 63) == 0) red[tid >> 6] = e;
    __syncthreads();
    if (tid == 0) {
      double x = 0.0;
      for (int w = 0; w < FSX_WG / 64; w++) x = dmax_(x, red[w]);
      s_err = x;
      if (s_bad) {
        atomicOr(A.err + inst, FSX_ERR_TRIDAG);
        s_err = 0.0;
      } else if (niter > 1000) {  // the reference stops: 'Multi-lithology diffusion not convergning'
        atomicOr(A.err + inst, FSX_ERR_MARINE);
        s_err = 0.0;
      }
    }
    __syncthreads();
  }
  if (tid == 0) A.niter[inst] = niter;
}

// ---- (4) bookkeeping (:147-165) --------------------------------------------------------------------
__global__ __launch_bounds__(FSX_CHUNK) void k_mar_post(const InstParams* __restrict__ PP, const Chunk* __restrict__ chunks,
                                                        const double* __restrict__ h, const double* __restrict__ ht,
                                                        double* __restrict__ etot, double* __restrict__ erate,
                                                        double* __restrict__ sedflux, double* __restrict__ fmix,
                                                        double* __restrict__ b) {
  MAR_NODE_PROLOGUE
  const double sl = P.sealevel, hv = h[g_], t = ht[g_];
  double e = etot[g_] + t - hv, r = erate[g_] + (t - hv) / P.dt;
  if (hv < sl) {
    sedflux[g_] = 0.0;
    e = 0.0;
    r = 0.0;
  }
  etot[g_] = e;
  erate[g_] = r;
  double f = fmix[g_];
  if (hv >= sl + 1.0e-3) f = 0.0;
  fmix[g_] = dmin_(dmax_(0.0, f), 1.0);
  b[g_] = dmin_(hv, b[g_]);
}

void launch_marine(DevState& S, hipStream_t st) {
  k_mar_sfd<<<S.ninst, FSX_WG, 0, st>>>(S.P, S.lstart, S.nlev, S.order, S.rec, S.dmask, S.h, S.sedflux, S.mar_flux,
                                         S.mar_h0, S.mar_indeg);
  k_mar_rec<<<S.nchunks, FSX_CHUNK, 0, st>>>(S.P, S.chunks, S.mar_h0, S.pexp, S.mar_rec, S.mar_w, S.mar_dir,
                                              S.mar_nrec, S.mar_indeg);
  k_mar_levels<<<S.ninst, FSX_WG, 0, st>>>(S.P, S.mar_indeg, S.mar_rec, S.mar_nrec, S.mar_order, S.mar_ls, S.mar_nlev,
                                            S.err);
  k_mar_shelf<<<S.ninst, FSX_WG, 0, st>>>(S.P, S.mar_ls, S.mar_nlev, S.mar_order, S.mar_rec, S.mar_w, S.mar_dir,
                                           S.mar_nrec, S.h, S.mar_sd, S.mar_flux);
  k_mar_prep<<<S.nchunks, FSX_CHUNK, 0, st>>>(S.P, S.chunks, S.h, S.fmix, S.mar_sd, S.mar_flux, S.mar_q1, S.mar_qs,
                                               S.mar_ht, S.mar_ft, S.mar_hh, S.mar_fh);
  SiltArgs A{S.h, S.fmix, S.mar_q1, S.mar_qs, S.mar_ht, S.mar_ft, S.mar_hp, S.mar_fp, S.mar_hh, S.mar_fh,
             S.mar_fhp, S.mar_tint, S.mar_gam, S.mar_niter, S.err};
  k_mar_silt<<<S.ninst, FSX_WG, 0, st>>>(S.P, A);
  k_mar_post<<<S.nchunks, FSX_CHUNK, 0, st>>>(S.P, S.chunks, S.h, S.mar_ht, S.etot, S.erate, S.sedflux, S.fmix, S.b);
}
