// fsx_tree.hip -- dependency-ordered sweeps over the receiver forest of each landscape.
//
// The reference walks a depth-first stack serially (src/FlowRouting.f90:393-443).  Here one
// 1024-thread workgroup per landscape builds a breadth-first level order (level 0 = base-level
// nodes, level l+1 = donors of level l) and sweeps it: top-down for the lake fill and the
// stream-power solve (each node needs only its receiver), bottom-up for drainage area (each node
// pulls its donors in descending index, which is the reference's summation order,
// src/FlowRouting.f90:136-139).  Level order is deterministic: positions come from a workgroup
// scan, not atomics.
#include "fsx_scan.h"
#include "fsx_state.h"

// BFS level order over the receiver forest, fused with the lake-surface fill of
// FlowRoutingSingleFlowDirection (src/FlowRouting.f90:73-89) and find_mult_rec (:183-198):
// hwater(d) = hwater(rec) + 1e-8 where hwater(rec) > h(d), else h(d); lake_depth = hwater - h.
__global__ __launch_bounds__(FSX_WG) void k_levels_fill(const InstParams* __restrict__ PP, const int* __restrict__ rec,
                                                        const uint8_t* __restrict__ dmask, const double* __restrict__ h,
                                                        double* __restrict__ hwater, double* __restrict__ lake_depth,
                                                        int* __restrict__ order, int* __restrict__ lstart,
                                                        int* __restrict__ nlev, int* __restrict__ err) {
  __shared__ int lds[FSX_WG / 64];
  const int inst = blockIdx.x;
  const InstParams& P = PP[inst];
  if (!P.active || !P.run_spl) return;
  const int nn = P.nn, tid = threadIdx.x;
  const int* R = rec + P.off;
  const uint8_t* M = dmask + P.off;
  const double* H = h + P.off;
  double* HW = hwater + P.off;
  double* LD = lake_depth + P.off;
  int* ord = order + P.off;
  int* ls = lstart + P.off + inst;  // nn+1 entries per landscape
  int cnt = 0;
  for (int base = 0; base < nn; base += FSX_WG) {
    const int v = base + tid;
    const int f = (v < nn && R[v] == v) ? 1 : 0;
    int tot;
    const int incl = block_incl_scan<FSX_WG>(f, lds, tot);
    if (f) {
      ord[cnt + incl - 1] = v;
      HW[v] = H[v];
      LD[v] = 0.0;
    }
    cnt += tot;
  }
  int fs = 0, fe = cnt, lvl = 0;
  while (fs < fe) {
    if (tid == 0) ls[lvl] = fs;
    int nend = fe;
    for (int base = fs; base < fe; base += FSX_WG) {
      const int idx = base + tid;
      int d[8], nd = 0, v = -1;
      if (idx < fe) {
        v = ord[idx];
        nd = donors_sorted(P, v, M[v], d);
      }
      int tot;
      const int incl = block_incl_scan<FSX_WG>(nd, lds, tot);
      if (nd) {
        const double hw = HW[v];
        int pos = nend + incl - nd;
        for (int k = 0; k < nd; k++) {
          const int x = d[k];
          ord[pos + k] = x;
          const double hx = H[x];
          const double w = hw > hx ? hw + 1.0e-8 : hx;
          HW[x] = w;
          LD[x] = w - hx;
        }
      }
      nend += tot;
    }
    __syncthreads();
    fs = fe;
    fe = nend;
    lvl++;
  }
  if (tid == 0) {
    ls[lvl] = fs;
    nlev[inst] = lvl;
    if (fs != nn) atomicOr(err + inst, FSX_ERR_LEVELS);
  }
}

// FlowAccumulationSingleFlowDirection, src/FlowRouting.f90:123-143.  a(v) = dx*dy*precip(v),
// then its donors' areas added from the highest donor index down; a base-level node adds its own
// area to itself last (the reference's a(rec(ijk)) += a(ijk) with rec(ijk) == ijk).
__global__ __launch_bounds__(FSX_WG) void k_area_sfd(const InstParams* __restrict__ PP, const int* __restrict__ rec,
                                                     const uint8_t* __restrict__ dmask, const double* __restrict__ precip,
                                                     const int* __restrict__ order, const int* __restrict__ lstart,
                                                     const int* __restrict__ nlev, double* __restrict__ area) {
  const int inst = blockIdx.x;
  const InstParams& P = PP[inst];
  if (!P.active || !P.run_spl || !P.sfd) return;
  const int* R = rec + P.off;
  const uint8_t* M = dmask + P.off;
  const double* PR = precip + P.off;
  const int* ord = order + P.off;
  const int* ls = lstart + P.off + inst;
  double* A = area + P.off;
  for (int l = nlev[inst] - 1; l >= 0; l--) {
    const int b0 = ls[l], b1 = ls[l + 1];
    for (int idx = b0 + (int)threadIdx.x; idx < b1; idx += FSX_WG) {
      const int v = ord[idx];
      double a = P.dxdy * PR[v];
      int d[8];
      const int nd = donors_sorted(P, v, M[v], d);
      for (int k = nd - 1; k >= 0; k--) a = a + A[d[k]];
      if (R[v] == v) a = a + a;
      A[v] = a;
    }
    __syncthreads();
  }
}

// Implicit stream-power law along single receivers, src/StreamPowerLaw.f90:210-400.
// One workgroup per landscape runs the Gauss-Seidel loop (one pass when g == 0): the erosion
// push dh (bottom-up, donors pulled in descending index as the reverse-stack push adds them),
// then elev, then the top-down solve: closed form for n == 1 (:306-314), relaxed Newton with
// omega = 0.875/n otherwise (:343-358); lakes keep elev and accumulate their water volume.
__global__ __launch_bounds__(FSX_WG) void k_spl_sfd(const InstParams* __restrict__ PP, const int* __restrict__ rec,
                                                    const uint8_t* __restrict__ dmask, const uint8_t* __restrict__ bc,
                                                    const int* __restrict__ order, const int* __restrict__ lstart,
                                                    const int* __restrict__ nlev, const double* __restrict__ area,
                                                    const double* __restrict__ length, const double* __restrict__ kfint,
                                                    const double* __restrict__ g, const double* __restrict__ ht,
                                                    double* __restrict__ h, double* __restrict__ hp, double* __restrict__ dh,
                                                    double* __restrict__ elev, double* __restrict__ water,
                                                    int* __restrict__ lsill, double* __restrict__ lwv,
                                                    double* __restrict__ lsed, double* __restrict__ scal,
                                                    int* __restrict__ ngs_out, long long* __restrict__ newton_out,
                                                    int* __restrict__ err) {
  __shared__ double red[FSX_WG / 64];
  __shared__ double s_tol, s_err, s_gmax;
  __shared__ long long s_newton;
  const int inst = blockIdx.x;
  const InstParams& P = PP[inst];
  if (!P.active || !P.run_spl || !P.sfd) return;
  const int nn = P.nn, tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  const int* R = rec + P.off;
  const uint8_t* M = dmask + P.off;
  const uint8_t* BC = bc + P.off;
  const int* ord = order + P.off;
  const int* ls = lstart + P.off + inst;
  const int L = nlev[inst];
  const double* A = area + P.off;
  const double* LEN = length + P.off;
  const double* KF = kfint + P.off;
  const double* G = g + P.off;
  const double* HT = ht + P.off;
  double* H = h + P.off;
  double* HP = hp + P.off;
  double* DH = dh + P.off;
  double* EL = elev + P.off;
  double* W = water + P.off;
  int* LS = lsill + P.off;
  double* LW = lwv + P.off;
  double* LSD = lsed + P.off;
  const double dt = P.dt, m = P.m, n = P.n;
  const bool n_is_one = fabs(n - 1.0) < 2.2250738585072014e-308;

  // tol = 1e-4*(maxval(abs(h))+1); gmax = maxval(g)
  double hm = 0.0, gm = -1.7976931348623157e308;
  for (int v = tid; v < nn; v += FSX_WG) {
    hm = fmax(hm, fabs(H[v]));
    gm = fmax(gm, G[v]);
    LSD[v] = 0.0;
    LS[v] = -1;
  }
  for (int o = 32; o > 0; o >>= 1) {
    hm = fmax(hm, __shfl_xor(hm, o, 64));
    gm = fmax(gm, __shfl_xor(gm, o, 64));
  }
  if (lane == 0) red[wid] = hm;
  __syncthreads();
  if (tid == 0) {
    double x = red[0];
    for (int w = 1; w < FSX_WG / 64; w++) x = fmax(x, red[w]);
    s_tol = 1.0e-4 * (x + 1.0);
    s_newton = 0;
  }
  __syncthreads();
  if (lane == 0) red[wid] = gm;
  __syncthreads();
  if (tid == 0) {
    double x = red[0];
    for (int w = 1; w < FSX_WG / 64; w++) x = fmax(x, red[w]);
    s_gmax = x;
    s_err = 2.0 * s_tol;
  }
  __syncthreads();
  const bool gzero = s_gmax < 2.2250738585072014e-308;
  const double dxdy_g = 1.0;  // placeholder to keep the expression order explicit below
  (void)dxdy_g;
  int ngs = 0;
  long long my_newton = 0;
  while (s_err > s_tol && ngs < 99) {
    ngs++;
    if (gzero) {
      // g == 0: dh == 0 exactly and elev == ht, the push only recomputes zeros
      for (int v = tid; v < nn; v += FSX_WG) { HP[v] = H[v]; EL[v] = HT[v]; }
    } else {
      for (int v = tid; v < nn; v += FSX_WG) HP[v] = H[v];
      __syncthreads();
      // dh push, src/StreamPowerLaw.f90:266-283, bottom-up
      for (int l = L - 1; l >= 0; l--) {
        for (int idx = ls[l] + tid; idx < ls[l + 1]; idx += FSX_WG) {
          const int v = ord[idx];
          const double own = HT[v] - HP[v];
          double s = own;
          int d[8];
          const int nd = donors_sorted(P, v, M[v], d);
          for (int k = nd - 1; k >= 0; k--) s = s + DH[d[k]];
          if (R[v] != v) {
            const double t = s - own;
            if (LS[v] == v) LSD[v] = t <= 0.0 ? 0.0 : t;
            DH[v] = t + own;
          } else {
            DH[v] = s;
            LSD[v] = s;
          }
        }
        __syncthreads();
      }
      for (int v = tid; v < nn; v += FSX_WG) {
        const double t = HT[v];
        EL[v] = BC[v] ? t : t + (DH[v] - (t - HP[v])) * G[v] * P.dx * P.dy / A[v];
      }
    }
    __syncthreads();
    // solve, base level up
    for (int l = 0; l < L; l++) {
      for (int idx = ls[l] + tid; idx < ls[l + 1]; idx += FSX_WG) {
        const int v = ord[idx];
        const int r = R[v];
        if (r == v) {
          W[v] = HT[v];
          LS[v] = v;
          LW[v] = 0.0;
          continue;
        }
        const double w_rcv = W[r];
        const double e = EL[v];
        double hv;
        if (e > w_rcv) {
          if (n_is_one) {
            double f = e, df = 1.0;
            const double fact = KF[v] * dt * pow(A[v], m) / LEN[v];
            f = f + fact * H[r];
            df = df + fact;
            hv = f / df;
          } else {
            const double omega = 0.875 / n, tolp = P.newton_tol;
            double errp = 2.0 * tolp;
            const double h0 = e;
            hv = H[v];
            const double hr = H[r];
            const bool act = HT[v] > HT[r];
            const double fact = act ? KF[v] * dt * pow(A[v], m) / pow(LEN[v], n) : 0.0;
            int it = 0;
            while (errp > tolp) {
              double f = hv - h0, df = 1.0;
              if (act) {
                const double x = fmax(0.0, hv - hr);
                f = f + fact * pow(x, n);
                df = df + fact * n * pow(x, n - 1.0);
              }
              const double hn = hv - f / df;
              errp = fabs(hn - hv);
              hv = hv * (1.0 - omega) + hn * omega;
              if (++it > 100000) { atomicOr(err + inst, FSX_ERR_NEWTON); break; }
            }
            my_newton += it;
          }
          LS[v] = v;
          LW[v] = 0.0;
          if (hv < w_rcv) hv = w_rcv;
        } else {
          hv = e;
          const int sl = LS[r];
          LS[v] = sl;
          if (!gzero && sl >= 0) atomicAdd(LW + sl, w_rcv - hv);
        }
        H[v] = hv;
        W[v] = w_rcv > hv ? w_rcv : hv;
      }
      __syncthreads();
    }
    // err = sqrt(sum((h-hp)**2)/nn), zero when g == 0 everywhere
    if (gzero) {
      if (tid == 0) s_err = 0.0;
    } else {
      double ss = 0.0;
      for (int v = tid; v < nn; v += FSX_WG) {
        const double dd = H[v] - HP[v];
        ss += dd * dd;
      }
      for (int o = 32; o > 0; o >>= 1) ss += __shfl_xor(ss, o, 64);
      if (lane == 0) red[wid] = ss;
      __syncthreads();
      if (tid == 0) {
        double x = 0.0;
        for (int w = 0; w < FSX_WG / 64; w++) x += red[w];
        s_err = sqrt(x / (double)nn);
      }
    }
    __syncthreads();
  }
  if (my_newton) atomicAdd((unsigned long long*)&s_newton, (unsigned long long)my_newton);
  __syncthreads();
  if (tid == 0) {
    ngs_out[inst] = ngs;
    newton_out[inst] = s_newton;
    scal[inst] = gzero ? 0.0 : 1.0;  // lake fill needed
  }
}

void launch_levels_fill(DevState& S, hipStream_t st) {
  k_levels_fill<<<S.ninst, FSX_WG, 0, st>>>(S.P, S.rec, S.dmask, S.h, S.hwater, S.lake_depth, S.order, S.lstart, S.nlev,
                                             S.err);
}
void launch_area(DevState& S, hipStream_t st) {
  k_area_sfd<<<S.ninst, FSX_WG, 0, st>>>(S.P, S.rec, S.dmask, S.precip, S.order, S.lstart, S.nlev, S.area);
}
void launch_spl(DevState& S, hipStream_t st) {
  k_spl_sfd<<<S.ninst, FSX_WG, 0, st>>>(S.P, S.rec, S.dmask, S.bc, S.order, S.lstart, S.nlev, S.area, S.length, S.kfint,
                                         S.g, S.ht, S.h, S.hp, S.dh, S.elev, S.water, S.lsill, S.lwv, S.lsed, S.scal,
                                         S.ngs, S.newton, S.err);
}
