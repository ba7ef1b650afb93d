// fsx_grid.hip -- node-parallel phases of the landscape step: uplift, D8 receivers, donor masks,
// stream-power bookkeeping and the ADI hillslope diffusion.
//
// Every kernel keeps the reference's floating-point operation order (the library is built with
// -ffp-contract=off), so results are bit-identical to the reference's FP64 Fortran.
#include "fsx_scan.h"
#include "fsx_state.h"

#define NODE_KERNEL_PROLOGUE                          \
  const Chunk c = chunks[blockIdx.x];                 \
  const InstParams& P = PP[c.inst];                   \
  const int v = c.start + (int)threadIdx.x;           \
  if (!P.active || v >= P.nn) return;                 \
  const long long g_ = (long long)P.off + v;

// Uplift, src/Uplift.f90:9-10: h = h + u*dt, b = b + u*dt.
__global__ __launch_bounds__(FSX_CHUNK) void k_uplift(const InstParams* __restrict__ PP, const Chunk* __restrict__ chunks,
                                                      double* __restrict__ h, double* __restrict__ b,
                                                      const double* __restrict__ u) {
  NODE_KERNEL_PROLOGUE
  if (!P.run_uplift) return;
  const double ud = u[g_] * P.dt;
  h[g_] = h[g_] + ud;
  b[g_] = b[g_] + ud;
}

// find_receiver, src/FlowRouting.f90:302-362.  Nodes outside the BC bounds keep rec=self,
// length=0; inside, the first strict maximum of (h(ij)-h(ijk))/l above tiny wins, jj outer.
__global__ __launch_bounds__(FSX_CHUNK) void k_receivers(const InstParams* __restrict__ PP, const Chunk* __restrict__ chunks,
                                                         const double* __restrict__ h, int* __restrict__ rec,
                                                         double* __restrict__ length) {
  NODE_KERNEL_PROLOGUE
  if (!P.run_spl) return;
  const double* hi = h + P.off;
  const int i = v % P.nx, j = v / P.nx;
  int r = v;
  double l_best = 0.0;
  if (i >= P.i1 && i <= P.i2 && j >= P.j1 && j <= P.j2) {
    const double hv = hi[v];
    double smax = 2.2250738585072014e-308;  // tiny(smax)
#pragma unroll
    for (int jj = -1; jj <= 1; jj++) {
#pragma unroll
      for (int ii = -1; ii <= 1; ii++) {
        const int q = nbr_wc(P, i, j, ii, jj);
        if (q != v) {
          const double l = P.len[(ii + 1) + 3 * (jj + 1)];
          const double slope = (hv - hi[q]) / l;
          if (slope > smax) { smax = slope; r = q; l_best = l; }
        }
      }
    }
  }
  rec[g_] = r;
  length[g_] = l_best;
}

// Donor masks: bit dbit(d) set when the neighbour of v in direction d drains to v.  Listing the
// set bits in ascending neighbour index reproduces find_donor (src/FlowRouting.f90:366-389).
// Also threads the donors as first[v] (smallest donor) and next[d] (next larger donor of the same
// receiver), which lets a depth-first preorder walk advance with one load per step.
__global__ __launch_bounds__(FSX_CHUNK) void k_donors(const InstParams* __restrict__ PP, const Chunk* __restrict__ chunks,
                                                      const int* __restrict__ rec, uint8_t* __restrict__ dmask,
                                                      int* __restrict__ first, int* __restrict__ next) {
  NODE_KERNEL_PROLOGUE
  if (!P.run_spl) return;
  const int* ri = rec + P.off;
  int q[8], dir[8];
  const int c8 = neighbours_wc(P, v, q, dir);
  unsigned m = 0;
  int d[8], nd = 0;
  for (int t = 0; t < c8; t++)
    if (ri[q[t]] == v) {
      m |= 1u << dbit(dir[t]);
      int x = q[t], s = nd++;
      while (s > 0 && d[s - 1] > x) { d[s] = d[s - 1]; s--; }
      d[s] = x;
    }
  dmask[g_] = (uint8_t)m;
  first[g_] = nd ? d[0] : -1;
  for (int k = 0; k < nd; k++) next[P.off + d[k]] = k + 1 < nd ? d[k + 1] : -1;
  if (ri[v] == v) next[g_] = -1;
}

// Stream-power set-up, src/StreamPowerLaw.f90:236-250: g, kfint, ht.
__global__ __launch_bounds__(FSX_CHUNK) void k_spl_prep(const InstParams* __restrict__ PP, const Chunk* __restrict__ chunks,
                                                        const double* __restrict__ h, const double* __restrict__ b,
                                                        const double* __restrict__ kf, double* __restrict__ kfint,
                                                        double* __restrict__ g, double* __restrict__ ht) {
  NODE_KERNEL_PROLOGUE
  if (!P.run_spl) return;
  const double hv = h[g_], d = hv - b[g_];
  if (P.g1 >= 0.0) g[g_] = (P.g2 > 0.0 && d > 1.0) ? P.g2 : P.g1;
  kfint[g_] = (P.kfsed > 0.0 && d > 1.0) ? P.kfsed : kf[g_];
  ht[g_] = hv;
}

// Stream-power bookkeeping, src/StreamPowerLaw.f90:380-393: b=min(h,b), lake sediment fill,
// etot=(etot+ht)-h, erate=(ht-h)/dt, Sedflux=ht-h.
__global__ __launch_bounds__(FSX_CHUNK) void k_spl_post(const InstParams* __restrict__ PP, const Chunk* __restrict__ chunks,
                                                        double* __restrict__ h, double* __restrict__ b,
                                                        const double* __restrict__ ht, double* __restrict__ etot,
                                                        double* __restrict__ erate, double* __restrict__ sedflux,
                                                        const int* __restrict__ lsill, const double* __restrict__ lwv,
                                                        const double* __restrict__ lsed, const double* __restrict__ water,
                                                        const double* __restrict__ gmax_flag) {
  NODE_KERNEL_PROLOGUE
  if (!P.run_spl) return;
  double hv = h[g_];
  b[g_] = fmin(hv, b[g_]) == hv ? hv : b[g_];
  if (gmax_flag[c.inst] != 0.0) {  // the fill adds exactly zero when g == 0 everywhere
    const int sl = lsill[g_];
    if (sl >= 0) {
      const double vol = lwv[P.off + sl];
      if (vol > 0.0) {
        double s = lsed[P.off + sl];
        s = s < vol ? s : vol;
        s = 0.0 > s ? 0.0 : s;
        hv = hv + s / vol * (water[g_] - hv);
      }
    }
    h[g_] = hv;
  }
  const double t = ht[g_];
  etot[g_] = etot[g_] + t - hv;
  erate[g_] = (t - hv) / P.dt;
  sedflux[g_] = t - hv;
}

// ---------------------------------------------------------------------------------------------
// Diffusion, src/Diffusion.f90:1-158 (ADI).  kdint first; then one thread per row solves the
// x-implicit system with the Thomas algorithm (tridag, :164-202) and one thread per column the
// y-implicit one.  Coefficients follow the reference's left-to-right order, e.g.
// ((kd(i+1,j)+kd(i,j))/2)*(dt/2)/dx**2 and ((z + fyp*zp) - (fyp+fym)*z) + fym*zm.
__global__ __launch_bounds__(FSX_CHUNK) void k_adi_prep(const InstParams* __restrict__ PP, const Chunk* __restrict__ chunks,
                                                        const double* __restrict__ h, const double* __restrict__ b,
                                                        const double* __restrict__ kd, double* __restrict__ kdint) {
  NODE_KERNEL_PROLOGUE
  if (!P.run_diff) return;
  kdint[g_] = (P.kdsed > 0.0 && (h[g_] - b[g_]) > 1.0e-6) ? P.kdsed : kd[g_];
}

__device__ __forceinline__ double coef(double ka, double kb, double hdt, double d2) {
  return (ka + kb) / 2.0 * hdt / d2;
}

// x pass: rows j = 1..ny-2 (0-based); reads h (the reference's zintp == h at this point),
// writes zint for those rows.
__global__ __launch_bounds__(64) void k_adi_x(const InstParams* __restrict__ PP, const double* __restrict__ h,
                                              const double* __restrict__ kdint, double* __restrict__ zint,
                                              double* __restrict__ gam, int* __restrict__ err) {
  const int inst = blockIdx.y;
  const InstParams& P = PP[inst];
  const int j = 1 + blockIdx.x * 64 + threadIdx.x;
  if (!P.active || !P.run_diff || j > P.ny - 2) return;
  const int nx = P.nx;
  const double* Z = h + P.off;
  const double* K = kdint + P.off;
  double* U = zint + P.off;
  double* G = gam + P.off;
  const double hdt = P.dt / 2.0, dx2 = P.dx * P.dx, dy2 = P.dy * P.dy;
  const int r0 = j * nx, rm = (j - 1) * nx, rp = (j + 1) * nx;
  double diag, sup, inf, f, fxp, fxm, fyp, fym;
  // i = 0 (left boundary)
  if (P.fix_left) { diag = 1.0; sup = 0.0; f = Z[r0]; }
  else {
    fxp = coef(K[r0 + 1], K[r0], hdt, dx2);
    fyp = coef(K[rp], K[r0], hdt, dy2);
    fym = coef(K[rm], K[r0], hdt, dy2);
    diag = 1.0 + fxp;
    sup = -fxp;
    f = Z[r0] + fyp * Z[rp] - (fyp + fym) * Z[r0] + fym * Z[rm];
  }
  if (diag == 0.0) { atomicOr(err + inst, FSX_ERR_TRIDAG); return; }
  double bet = diag, uprev = f / bet, cprev = sup;
  U[r0] = uprev;
  for (int i = 1; i < nx; i++) {
    if (i < nx - 1) {
      fxp = coef(K[r0 + i + 1], K[r0 + i], hdt, dx2);
      fxm = coef(K[r0 + i - 1], K[r0 + i], hdt, dx2);
      fyp = coef(K[rp + i], K[r0 + i], hdt, dy2);
      fym = coef(K[rm + i], K[r0 + i], hdt, dy2);
      diag = 1.0 + fxp + fxm;
      sup = -fxp;
      inf = -fxm;
      f = Z[r0 + i] + fyp * Z[rp + i] - (fyp + fym) * Z[r0 + i] + fym * Z[rm + i];
    } else if (P.fix_right) {
      diag = 1.0; inf = 0.0; f = Z[r0 + i]; sup = 0.0;
    } else {
      fxm = coef(K[r0 + i - 1], K[r0 + i], hdt, dx2);
      fyp = coef(K[rp + i], K[r0 + i], hdt, dy2);
      fym = coef(K[rm + i], K[r0 + i], hdt, dy2);
      diag = 1.0 + fxm;
      inf = -fxm;
      f = Z[r0 + i] + fyp * Z[rp + i] - (fyp + fym) * Z[r0 + i] + fym * Z[rm + i];
      sup = 0.0;
    }
    const double gm = cprev / bet;
    bet = diag - inf * gm;
    if (bet == 0.0) { atomicOr(err + inst, FSX_ERR_TRIDAG); return; }
    uprev = (f - inf * uprev) / bet;
    G[r0 + i] = gm;
    U[r0 + i] = uprev;
    cprev = sup;
  }
  for (int i = nx - 2; i >= 0; i--) {
    uprev = U[r0 + i] - G[r0 + i + 1] * uprev;
    U[r0 + i] = uprev;
  }
}

// y pass: columns i = 1..nx-2; reads zint (rows 1..ny-2) and h (rows 0 and ny-1, untouched by
// the x pass), writes zintp for those columns.
__global__ __launch_bounds__(64) void k_adi_y(const InstParams* __restrict__ PP, const double* __restrict__ h,
                                              const double* __restrict__ zint, const double* __restrict__ kdint,
                                              double* __restrict__ zintp, double* __restrict__ gam, int* __restrict__ err) {
  const int inst = blockIdx.y;
  const InstParams& P = PP[inst];
  const int i = 1 + blockIdx.x * 64 + threadIdx.x;
  if (!P.active || !P.run_diff || i > P.nx - 2) return;
  const int nx = P.nx, ny = P.ny;
  const double* H = h + P.off;
  const double* ZI = zint + P.off;
  const double* K = kdint + P.off;
  double* U = zintp + P.off;
  double* G = gam + P.off;
  const double hdt = P.dt / 2.0, dx2 = P.dx * P.dx, dy2 = P.dy * P.dy;
  auto Z = [&](int ii, int jj) -> double { return (jj == 0 || jj == ny - 1) ? H[jj * nx + ii] : ZI[jj * nx + ii]; };
  const int jst = ny - 1;  // 0-based value of the row index left over by "do j=2,ny-1" (ny >= 3)
  double diag, sup, inf, f, fxp, fxm, fyp, fym;
  if (P.fix_bottom) { diag = 1.0; sup = 0.0; f = Z(i, 0); }
  else {
    fxp = coef(K[i + 1], K[jst * nx + i], hdt, dx2);  // stale j, src/Diffusion.f90:114
    fxm = coef(K[i - 1], K[i], hdt, dx2);
    fyp = coef(K[nx + i], K[i], hdt, dy2);
    diag = 1.0 + fyp;
    sup = -fyp;
    f = Z(i, 0) + fxp * Z(i + 1, 0) - (fxp + fxm) * Z(i, 0) + fxm * Z(i - 1, 0);
  }
  if (diag == 0.0) { atomicOr(err + inst, FSX_ERR_TRIDAG); return; }
  double bet = diag, uprev = f / bet, cprev = sup;
  U[i] = uprev;
  for (int j = 1; j < ny; j++) {
    const int o = j * nx + i;
    if (j < ny - 1) {
      fxp = coef(K[o + 1], K[o], hdt, dx2);
      fxm = coef(K[o - 1], K[o], hdt, dx2);
      fyp = coef(K[o + nx], K[o], hdt, dy2);
      fym = coef(K[o - nx], K[o], hdt, dy2);
      diag = 1.0 + fyp + fym;
      sup = -fyp;
      inf = -fym;
      f = Z(i, j) + fxp * Z(i + 1, j) - (fxp + fxm) * Z(i, j) + fxm * Z(i - 1, j);
    } else if (P.fix_top) {
      diag = 1.0; inf = 0.0; f = Z(i, j); sup = 0.0;
    } else {
      fxp = coef(K[o + 1], K[o], hdt, dx2);
      fxm = coef(K[o - 1], K[o], hdt, dx2);
      fym = coef(K[o - nx], K[o], hdt, dy2);
      diag = 1.0 + fym;
      inf = -fym;
      f = Z(i, j) + fxp * Z(i + 1, j) - (fxp + fxm) * Z(i, j) + fxm * Z(i - 1, j);
      sup = 0.0;
    }
    const double gm = cprev / bet;
    bet = diag - inf * gm;
    if (bet == 0.0) { atomicOr(err + inst, FSX_ERR_TRIDAG); return; }
    uprev = (f - inf * uprev) / bet;
    G[o] = gm;
    U[o] = uprev;
    cprev = sup;
  }
  for (int j = ny - 2; j >= 0; j--) {
    const int o = j * nx + i;
    uprev = U[o] - G[o + nx] * uprev;
    U[o] = uprev;
  }
}

// Diffusion output, src/Diffusion.f90:143-152 (columns 0 and nx-1 keep h: zintp==h there).
__global__ __launch_bounds__(FSX_CHUNK) void k_adi_post(const InstParams* __restrict__ PP, const Chunk* __restrict__ chunks,
                                                        double* __restrict__ h, double* __restrict__ b,
                                                        double* __restrict__ etot, double* __restrict__ erate,
                                                        const double* __restrict__ zintp) {
  NODE_KERNEL_PROLOGUE
  if (!P.run_diff) return;
  const int i = v % P.nx;
  const double hv = h[g_];
  const double zp = (i == 0 || i == P.nx - 1) ? hv : zintp[g_];
  etot[g_] = etot[g_] + hv - zp;
  erate[g_] = erate[g_] + (hv - zp) / P.dt;
  h[g_] = zp;
  const double bv = b[g_];
  b[g_] = zp < bv ? zp : bv;
}

// ---------------------------------------------------------------------------------------------
void launch_uplift(DevState& S, hipStream_t st) {
  k_uplift<<<S.nchunks, FSX_CHUNK, 0, st>>>(S.P, S.chunks, S.h, S.b, S.u);
}
void launch_receivers(DevState& S, hipStream_t st) {
  k_receivers<<<S.nchunks, FSX_CHUNK, 0, st>>>(S.P, S.chunks, S.h, S.rec, S.length);
}
void launch_donors(DevState& S, hipStream_t st) {
  k_donors<<<S.nchunks, FSX_CHUNK, 0, st>>>(S.P, S.chunks, S.rec, S.dmask, S.first, S.next);
}
void launch_spl_prep(DevState& S, hipStream_t st) {
  k_spl_prep<<<S.nchunks, FSX_CHUNK, 0, st>>>(S.P, S.chunks, S.h, S.b, S.kf, S.kfint, S.g, S.ht);
}
void launch_spl_post(DevState& S, hipStream_t st) {
  k_spl_post<<<S.nchunks, FSX_CHUNK, 0, st>>>(S.P, S.chunks, S.h, S.b, S.ht, S.etot, S.erate, S.sedflux, S.lsill,
                                               S.lwv, S.lsed, S.water, S.scal);
}
void launch_diffusion(DevState& S, hipStream_t st) {
  k_adi_prep<<<S.nchunks, FSX_CHUNK, 0, st>>>(S.P, S.chunks, S.h, S.b, S.kd, S.kdint);
  dim3 gx((S.max_ny + 63) / 64, S.ninst), gy((S.max_nx + 63) / 64, S.ninst);
  k_adi_x<<<gx, 64, 0, st>>>(S.P, S.h, S.kdint, S.zint, S.gam, S.err);
  k_adi_y<<<gy, 64, 0, st>>>(S.P, S.h, S.zint, S.kdint, S.zintp, S.gam, S.err);
  k_adi_post<<<S.nchunks, FSX_CHUNK, 0, st>>>(S.P, S.chunks, S.h, S.b, S.etot, S.erate, S.zintp);
}
