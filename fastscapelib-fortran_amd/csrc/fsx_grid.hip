// fsx_grid.hip -- node-parallel phases of the landscape step: uplift, D8 receivers, donor masks,
// stream-power bookkeeping and the ADI hillslope diffusion.
//
// Every kernel keeps the reference's floating-point operation order (the library is built with
// -ffp-contract=off), so results are bit-identical to the reference's FP64 Fortran.
#include <climits>

#include "fsx_scan.h"
#include "fsx_state.h"

#define NODE_KERNEL_PROLOGUE                          \
  const Chunk c = chunks[chunk_id()];                 \
  const InstParams& P = PP[c.inst];                   \
  const int v = c.start + (int)threadIdx.x;           \
  if (!P.active || v >= P.nn) return;                 \
  const long long g_ = (long long)P.off + v;

// Uplift and find_receiver fused, over 2-D tiles of FSX_TW x TH nodes (TH rows: 4, 8, 16 or 32, the tile
// table's height, a launch shape the autotuner chooses) staged in LDS with a one-node
// halo (one workgroup per tile, tiles of a landscape dealt to XCDs in contiguous runs so that the halo
// rows of neighbouring tiles come from one L2).  The halo cell of direction (ii, jj) holds h of the node
// find_receiver takes there (cyclic wrap else clamp, src/FlowRouting.f90:336-345), uplifted with the
// same expression as the node's own tile computes it (Uplift, src/Uplift.f90:9-10: h + u*dt).  The
// receiver search then reads only LDS: jj outer, ii inner, the first strict maximum of (h(ij)-h(ijk))/l
// above tiny wins (:347-353), with the neighbour distances computed on the host by the reference's
// expression.  Nodes outside the BC bounds keep rec = self, length = 0 (:322-325).
// The uplifted h of a routed landscape goes to ht (its neighbours' tiles still read h); every kernel up to
// k_spl_prep reads ht, and k_spl_prep writes it back to h.  A landscape without routing (run_spl == 0)
// reads no neighbours and is uplifted in place.  uplift == 0: the receivers of the current h (ht = h).
template <int TH>
__global__ __launch_bounds__(256) void k_route_tile(const InstParams* __restrict__ PP, const int4* __restrict__ tiles,
                                                    double* __restrict__ h, double* __restrict__ b,
                                                    const double* __restrict__ u, double* __restrict__ ht,
                                                    int* __restrict__ rec, double* __restrict__ length, int uplift) {
  constexpr int W = FSX_TW + 2, H = TH + 2;
  __shared__ double sh[H][W];
  __shared__ double sud[TH][FSX_TW];
  const int4 T = tiles[chunk_id()];
  const InstParams& P = PP[T.x];
  if (!P.active) return;
  const bool up = uplift && P.run_uplift;
  const int x0 = T.y, y0 = T.z, nx = P.nx, ny = P.ny;
  const long long off = P.off;
  const double dt = P.dt;
  if (!P.run_spl) {  // uplift alone, in place
    if (!up) return;
    for (int k = threadIdx.x; k < FSX_TW * TH; k += 256) {
      const int i = x0 + (k % FSX_TW), j = y0 + k / FSX_TW;
      if (i >= nx || j >= ny) continue;
      const long long g = off + i + (long long)j * nx;
      const double ud = u[g] * dt;
      h[g] = h[g] + ud;
      b[g] = b[g] + ud;
    }
    return;
  }
  // every load of the thread is issued before any is used (the tile's staging waits for memory once): the
  // halo cells' h and u, and b of the thread's own nodes (updated after the barrier)
  constexpr int NH = (W * H + 255) / 256, NPT = FSX_TW * TH / 256;
  double hh[NH], uu[NH], bb[NPT];
#pragma unroll
  for (int m = 0; m < NH; m++) {
    const int c = (int)threadIdx.x + 256 * m;
    const int ly = c / W, lx = c - ly * W;
    int gx = x0 + lx - 1, gy = y0 + ly - 1;
    if (gx < 0 && P.xcyc) gx += nx;
    if (gx > nx - 1 && P.xcyc) gx -= nx;
    gx = gx < 0 ? 0 : (gx > nx - 1 ? nx - 1 : gx);
    if (gy < 0 && P.ycyc) gy += ny;
    if (gy > ny - 1 && P.ycyc) gy -= ny;
    gy = gy < 0 ? 0 : (gy > ny - 1 ? ny - 1 : gy);
    const long long g = off + gx + (long long)gy * nx;
    const bool in = c < W * H;
    hh[m] = in ? h[g] : 0.0;
    uu[m] = (in && up) ? u[g] : 0.0;
  }
#pragma unroll
  for (int m = 0; m < NPT; m++) {
    const int k = (int)threadIdx.x + 256 * m;
    const int i = x0 + k % FSX_TW, j = y0 + k / FSX_TW;
    bb[m] = (up && i < nx && j < ny) ? b[off + i + (long long)j * nx] : 0.0;
  }
#pragma unroll
  for (int m = 0; m < NH; m++) {
    const int c = (int)threadIdx.x + 256 * m;
    if (c >= W * H) continue;
    const int ly = c / W, lx = c - ly * W;
    double hv = hh[m];
    if (up) {
      const double ud = uu[m] * dt;
      hv = hv + ud;
      if (lx >= 1 && lx <= FSX_TW && ly >= 1 && ly <= TH) sud[ly - 1][lx - 1] = ud;
    }
    sh[ly][lx] = hv;
  }
  __syncthreads();
#pragma unroll
  for (int m = 0; m < NPT; m++) {
    const int k = (int)threadIdx.x + 256 * m;
    const int tx = k % FSX_TW, ty = k / FSX_TW;
    const int i = x0 + tx, j = y0 + ty;
    if (i >= nx || j >= ny) continue;
    const int v = i + j * nx;
    const long long g = off + v;
    const double hv = sh[ty + 1][tx + 1];
    ht[g] = hv;
    if (up) b[g] = bb[m] + sud[ty][tx];
    int r = v;
    double l_best = 0.0;
    if (i >= P.i1 && i <= P.i2 && j >= P.j1 && j <= P.j2) {
      double smax = 2.2250738585072014e-308;  // tiny(smax)
#pragma unroll
      for (int jj = -1; jj <= 1; jj++) {
        int jjj = j + jj;
        if (jjj < 0 && P.ycyc) jjj += ny;
        if (jjj < 0) jjj = 0;
        if (jjj > ny - 1 && P.ycyc) jjj -= ny;
        if (jjj > ny - 1) jjj = ny - 1;
#pragma unroll
        for (int ii = -1; ii <= 1; ii++) {
          int iii = i + ii;
          if (iii < 0 && P.xcyc) iii += nx;
          if (iii < 0) iii = 0;
          if (iii > nx - 1 && P.xcyc) iii -= nx;
          if (iii > nx - 1) iii = nx - 1;
          const int q = iii + jjj * nx;
          if (q != v) {
            const double l = P.len[(ii + 1) + 3 * (jj + 1)];
            const double slope = (hv - sh[ty + 1 + jj][tx + 1 + ii]) / l;
            if (slope > smax) { smax = slope; r = q; l_best = l; }
          }
        }
      }
    }
    rec[g] = r;
    length[g] = l_best;
  }
}

// Donor masks over the same 2-D tiles: the receivers of a tile and its one-node halo (wrapped or clamped
// as find_receiver takes neighbours) in LDS, then for each node the distinct neighbours in direction
// order (jj outer, ii inner, itself and repeated clamp targets skipped, as neighbours_wc) whose receiver
// is the node: find_donor's donor set (src/FlowRouting.f90:366-389) as a mask, bit dbit(d) for direction d.
template <int TH>
__global__ __launch_bounds__(256) void k_donors_tile(const InstParams* __restrict__ PP, const int4* __restrict__ tiles,
                                                     const int* __restrict__ rec, uint8_t* __restrict__ dmask) {
  constexpr int W = FSX_TW + 2, H = TH + 2;
  __shared__ int sr[H][W];
  const int4 T = tiles[chunk_id()];
  const InstParams& P = PP[T.x];
  if (!P.active || !P.run_spl) return;
  const int x0 = T.y, y0 = T.z, nx = P.nx, ny = P.ny;
  const long long off = P.off;
  constexpr int NH = (W * H + 255) / 256;
  int rr[NH];  // every load issued before any is used
#pragma unroll
  for (int m = 0; m < NH; m++) {
    const int c = (int)threadIdx.x + 256 * m;
    const int ly = c / W, lx = c - ly * W;
    int gx = x0 + lx - 1, gy = y0 + ly - 1;
    if (gx < 0 && P.xcyc) gx += nx;
    if (gx > nx - 1 && P.xcyc) gx -= nx;
    gx = gx < 0 ? 0 : (gx > nx - 1 ? nx - 1 : gx);
    if (gy < 0 && P.ycyc) gy += ny;
    if (gy > ny - 1 && P.ycyc) gy -= ny;
    gy = gy < 0 ? 0 : (gy > ny - 1 ? ny - 1 : gy);
    rr[m] = c < W * H ? rec[off + gx + (long long)gy * nx] : 0;
  }
#pragma unroll
  for (int m = 0; m < NH; m++) {
    const int c = (int)threadIdx.x + 256 * m;
    if (c < W * H) sr[c / W][c % W] = rr[m];
  }
  __syncthreads();
#pragma unroll
  for (int kk = 0; kk < FSX_TW * TH / 256; kk++) {
    const int k = (int)threadIdx.x + 256 * kk;
    const int tx = k % FSX_TW, ty = k / FSX_TW;
    const int i = x0 + tx, j = y0 + ty;
    if (i >= nx || j >= ny) continue;
    const int v = i + j * nx;
    unsigned m = 0;
    if (i > 0 && j > 0 && i < nx - 1 && j < ny - 1) {  // interior: eight distinct neighbours
#pragma unroll
      for (int d = 0; d < 9; d++)
        if (d != 4 && sr[ty + d / 3][tx + d % 3] == v) m |= 1u << dbit(d);
    } else {
      // Wrapped or clamped column / row of each offset.  Direction (ii, jj) repeats an earlier one
      // of the jj-outer, ii-inner order exactly when an earlier ii' of the same row maps to the same
      // column or an earlier jj' to the same row; it is the node itself when both map to it.
      int ex[3], ey[3];
#pragma unroll
      for (int o = 0; o < 3; o++) {
        int a = i + o - 1, c = j + o - 1;
        if (a < 0 && P.xcyc) a += nx;
        if (a > nx - 1 && P.xcyc) a -= nx;
        ex[o] = a < 0 ? 0 : (a > nx - 1 ? nx - 1 : a);
        if (c < 0 && P.ycyc) c += ny;
        if (c > ny - 1 && P.ycyc) c -= ny;
        ey[o] = c < 0 ? 0 : (c > ny - 1 ? ny - 1 : c);
      }
      const bool xd[3] = {false, ex[1] == ex[0], ex[2] == ex[0] || ex[2] == ex[1]};
      const bool yd[3] = {false, ey[1] == ey[0], ey[2] == ey[0] || ey[2] == ey[1]};
#pragma unroll
      for (int d = 0; d < 9; d++) {
        const int ii = d % 3, jj = d / 3;
        if (d == 4 || xd[ii] || yd[jj] || (ex[ii] == i && ey[jj] == j)) continue;
        if (sr[ty + jj][tx + ii] == v) m |= 1u << dbit(d);
      }
    }
    dmask[off + v] = (uint8_t)m;
  }
}

// Stream-power set-up, src/StreamPowerLaw.f90:236-250: g, kfint, ht (written by k_route_tile; h := ht).
__global__ __launch_bounds__(FSX_CHUNK) void k_spl_prep(const InstParams* __restrict__ PP, const Chunk* __restrict__ chunks,
                                                        double* __restrict__ h, const double* __restrict__ b,
                                                        const double* __restrict__ kf, double* __restrict__ kfint,
                                                        double* __restrict__ g, const double* __restrict__ ht,
                                                        unsigned long long* __restrict__ lacc_lo,
                                                        unsigned long long* __restrict__ lacc_hi) {
  NODE_KERNEL_PROLOGUE
  if (!P.run_spl) return;
  lacc_lo[g_] = 0ull;  // lake volume sums of this step (k_lake_acc)
  lacc_hi[g_] = 0ull;
  const double hv = ht[g_], d = hv - b[g_];  // ht: h after uplift (k_route_tile)
  if (P.g1 >= 0.0) g[g_] = (P.g2 > 0.0 && d > 1.0) ? P.g2 : P.g1;
  kfint[g_] = (P.kfsed > 0.0 && d > 1.0) ? P.kfsed : kf[g_];
  h[g_] = hv;
}

// Lake water volumes (src/StreamPowerLaw.f90:323-324, lake_water_volume(lake_sill) += water(rec) - h, summed
// along the stack): every lake node of the last Gauss-Seidel pass adds its share to its sill as a 128-bit fixed-point
// integer (units of 2^-64 m; shares below 2^-12 m lose their bits under 2^-64 m), by integer atomics that commute, so
// the volume does not depend on the order of the additions (run, batch, lane split); the reference's serial double
// sum and this exact sum, rounded once, agree to the last places.
__device__ __forceinline__ double lake_volume(unsigned long long lo, unsigned long long hi) {
  return (double)hi + (double)lo * 5.421010862427522170037e-20;  // 2^-64
}
__global__ __launch_bounds__(FSX_CHUNK) void k_lake_acc(const InstParams* __restrict__ PP, const Chunk* __restrict__ chunks,
                                                        const double* __restrict__ lakec, const int* __restrict__ lsill,
                                                        const double* __restrict__ gmax_flag,
                                                        unsigned long long* __restrict__ lacc_lo,
                                                        unsigned long long* __restrict__ lacc_hi) {
  NODE_KERNEL_PROLOGUE
  if (!P.run_spl || gmax_flag[c.inst] == 0.0) return;
  const double x = lakec[g_];
  const int sl = lsill[g_];
  if (!(x >= 0.0) || sl < 0) return;
  const double ip = floor(x);
  const unsigned long long hi = (unsigned long long)ip;
  const unsigned long long lo = (unsigned long long)((x - ip) * 18446744073709551616.0);  // exact scaling by 2^64
  const long long s = (long long)P.off + sl;
  const unsigned long long old = atomicAdd(lacc_lo + s, lo);
  atomicAdd(lacc_hi + s, hi + (old + lo < old ? 1ull : 0ull));  // carry out of the low word
}

// Stream-power bookkeeping, src/StreamPowerLaw.f90:380-393: b=min(h,b), lake sediment fill,
// etot=(etot+ht)-h, erate=(ht-h)/dt, Sedflux=ht-h.  For single flow direction the sweep left h and
// the lake surface in level order (Hl, HWl); they are scattered back to grid order here, with
// lake_depth = hwater - h of the routing (FlowRoutingSingleFlowDirection :87-89).
__global__ __launch_bounds__(FSX_CHUNK) void k_spl_post(const InstParams* __restrict__ PP, const Chunk* __restrict__ chunks,
                                                        double* __restrict__ h, double* __restrict__ b,
                                                        const double* __restrict__ ht, double* __restrict__ etot,
                                                        double* __restrict__ erate, double* __restrict__ sedflux,
                                                        const int* __restrict__ lsill,
                                                        const unsigned long long* __restrict__ lacc_lo,
                                                        const unsigned long long* __restrict__ lacc_hi,
                                                        const double* __restrict__ lsed, const double* __restrict__ water,
                                                        const double* __restrict__ gmax_flag, const int* __restrict__ pos,
                                                        const double* __restrict__ Hl, const double* __restrict__ HWl,
                                                        double* __restrict__ hwater, double* __restrict__ lake_depth,
                                                        const int* __restrict__ mpos) {
  NODE_KERNEL_PROLOGUE
  if (!P.run_spl) return;
  double hv;
  if (P.sfd) {
    const long long gp = (long long)P.off + pos[g_];
    hv = Hl[gp];
    const double hw = HWl[gp];
    hwater[g_] = hw;
    lake_depth[g_] = hw - ht[g_];
  } else {
    hv = Hl[(long long)P.off + mpos[g_]];  // multiple flow direction: Kahn-order positions
  }
  b[g_] = fmin(hv, b[g_]) == hv ? hv : b[g_];
  if (gmax_flag[c.inst] != 0.0) {  // the fill adds exactly zero when g == 0 everywhere
    const int sl = lsill[g_];
    if (sl >= 0) {
      const double vol = lake_volume(lacc_lo[P.off + sl], lacc_hi[P.off + sl]);
      if (vol > 0.0) {
        double s = lsed[P.off + sl];
        s = s < vol ? s : vol;
        s = 0.0 > s ? 0.0 : s;
        hv = hv + s / vol * (water[g_] - hv);
      }
    }
  }
  h[g_] = hv;
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
__device__ __forceinline__ double coef(double ka, double kb, double hdt, double d2) {
  return (ka + kb) / 2.0 * hdt / d2;
}

// kdint (:26-33), the couplings cx, cy and the x pass's diagonal and right-hand side (rows
// j = 1..ny-2, :45-82; zintp == h at this point) in one pass: a thread recomputes the kdint of the
// four nodes around it, which is the expression their own threads evaluate, so every coupling is
// bit-identical to the one stored by its owner.
__device__ __forceinline__ double kdint_of(const InstParams& P, const double* h, const double* b, const double* kd,
                                          long long g) {
  return (P.kdsed > 0.0 && (h[g] - b[g]) > 1.0e-6) ? P.kdsed : kd[g];
}

__global__ __launch_bounds__(FSX_CHUNK) void k_adi_xprep(const InstParams* __restrict__ PP, const Chunk* __restrict__ chunks,
                                                         const double* __restrict__ h, const double* __restrict__ b,
                                                         const double* __restrict__ kd, double* __restrict__ kdint,
                                                         double* __restrict__ cx, double* __restrict__ cy,
                                                         double* __restrict__ dg, double* __restrict__ rf) {
  NODE_KERNEL_PROLOGUE
  if (!P.run_diff) return;
  const int nx = P.nx, ny = P.ny, i = v % nx, j = v / nx;
  const double hdt = P.dt / 2.0, dx2 = P.dx * P.dx, dy2 = P.dy * P.dy;
  const double kv = kdint_of(P, h, b, kd, g_);
  kdint[g_] = kv;
  const double cxv = i < nx - 1 ? coef(kdint_of(P, h, b, kd, g_ + 1), kv, hdt, dx2) : 0.0;
  const double cyv = j < ny - 1 ? coef(kdint_of(P, h, b, kd, g_ + nx), kv, hdt, dy2) : 0.0;
  cx[g_] = cxv;
  cy[g_] = cyv;
  if (j == 0 || j == ny - 1) return;
  const double* Z = h + P.off;
  const double fyp = cyv, fym = coef(kv, kdint_of(P, h, b, kd, g_ - nx), hdt, dy2);  // cy(i, j-1)
  double diag, f;
  const double frhs = Z[v] + fyp * Z[v + nx] - (fyp + fym) * Z[v] + fym * Z[v - nx];
  if (i == 0) {
    if (P.fix_left) { diag = 1.0; f = Z[v]; }
    else { diag = 1.0 + cxv; f = frhs; }
  } else {
    const double cxm = coef(kv, kdint_of(P, h, b, kd, g_ - 1), hdt, dx2);  // cx(i-1, j)
    if (i == nx - 1) {
      if (P.fix_right) { diag = 1.0; f = Z[v]; }
      else { diag = 1.0 + cxm; f = frhs; }
    } else {
      diag = 1.0 + cxv + cxm;
      f = frhs;
    }
  }
  dg[g_] = diag;
  rf[g_] = f;
}

// y pass, columns i = 1..nx-2 (:97-134): zint rows 1..ny-2 from the x pass, rows 0 and ny-1 are h.
// The reflective-bottom coefficient uses kdint(i,j) with j left at ny by the x loop (:114).
__global__ __launch_bounds__(FSX_CHUNK) void k_adi_yrhs(const InstParams* __restrict__ PP, const Chunk* __restrict__ chunks,
                                                        const double* __restrict__ h, const double* __restrict__ zint,
                                                        const double* __restrict__ kdint, const double* __restrict__ cx,
                                                        const double* __restrict__ cy, double* __restrict__ dg,
                                                        double* __restrict__ rf) {
  NODE_KERNEL_PROLOGUE
  if (!P.run_diff) return;
  const int nx = P.nx, ny = P.ny, i = v % nx, j = v / nx;
  if (i == 0 || i == nx - 1) return;
  const double* ZR = (j == 0 || j == ny - 1) ? h + P.off : zint + P.off;
  const double* CX = cx + P.off;
  const double* CY = cy + P.off;
  double fxp = CX[v];
  const double fxm = CX[v - 1];
  double diag, f;
  if (j == 0) {
    if (P.fix_bottom) { diag = 1.0; f = ZR[v]; }
    else {
      fxp = coef(kdint[P.off + i + 1], kdint[P.off + (ny - 1) * nx + i], P.dt / 2.0, P.dx * P.dx);
      diag = 1.0 + CY[v];
      f = ZR[v] + fxp * ZR[v + 1] - (fxp + fxm) * ZR[v] + fxm * ZR[v - 1];
    }
  } else if (j == ny - 1) {
    if (P.fix_top) { diag = 1.0; f = ZR[v]; }
    else {
      diag = 1.0 + CY[v - nx];
      f = ZR[v] + fxp * ZR[v + 1] - (fxp + fxm) * ZR[v] + fxm * ZR[v - 1];
    }
  } else {
    diag = 1.0 + CY[v] + CY[v - nx];
    f = ZR[v] + fxp * ZR[v + 1] - (fxp + fxm) * ZR[v] + fxm * ZR[v - 1];
  }
  dg[g_] = diag;
  rf[g_] = f;
}

// Thomas recurrence along one line of n nodes at stride st from base o (tridag, :178-198):
//   bet = b(1); u(1) = r(1)/bet; gam(k) = c(k-1)/bet; bet = b(k) - a(k)*gam(k); u(k) = (r(k) - a(k)*u(k-1))/bet;
//   u(k) = u(k) - gam(k+1)*u(k+1) backwards.
// a(k) = -cc(k-1), c(k-1) = -cc(k-1) for the coupling cc between k-1 and k, except c(1) = 0 when the
// first node is fixed and a(n) = 0 when the last one is; inputs are read in chunks of ADI_CH.
#define ADI_CH 8
__device__ __forceinline__ bool thomas_line(const double* __restrict__ dg, const double* __restrict__ rf,
                                            const double* __restrict__ cc, double* __restrict__ U,
                                            double* __restrict__ G, long long o, int st, int n, bool fix_first,
                                            bool fix_last) {
  double bet = dg[o];
  if (bet == 0.0) return false;
  double uprev = rf[o] / bet;
  U[o] = uprev;
  double nd[ADI_CH], nr[ADI_CH], nc[ADI_CH];
#pragma unroll
  for (int e = 0; e < ADI_CH; e++) {
    const int k = 1 + e;
    const long long q = o + (long long)k * st;
    nd[e] = k < n ? dg[q] : 0.0;
    nr[e] = k < n ? rf[q] : 0.0;
    nc[e] = k < n ? cc[q - st] : 0.0;
  }
  for (int k0 = 1; k0 < n; k0 += ADI_CH) {
    double d[ADI_CH], r[ADI_CH], c[ADI_CH];
#pragma unroll
    for (int e = 0; e < ADI_CH; e++) { d[e] = nd[e]; r[e] = nr[e]; c[e] = nc[e]; }
#pragma unroll
    for (int e = 0; e < ADI_CH; e++) {  // next chunk in flight while this one is solved
      const int k = k0 + ADI_CH + e;
      const long long q = o + (long long)k * st;
      nd[e] = k < n ? dg[q] : 0.0;
      nr[e] = k < n ? rf[q] : 0.0;
      nc[e] = k < n ? cc[q - st] : 0.0;
    }
#pragma unroll
    for (int e = 0; e < ADI_CH; e++) {
      const int k = k0 + e;
      if (k < n) {
        const double cprev = (k == 1 && fix_first) ? 0.0 : -c[e];
        const double inf = (k == n - 1 && fix_last) ? 0.0 : -c[e];
        const double gm = cprev / bet;
        bet = d[e] - inf * gm;
        if (bet == 0.0) return false;
        uprev = (r[e] - inf * uprev) / bet;
        const long long q = o + (long long)k * st;
        G[q] = gm;
        U[q] = uprev;
      }
    }
  }
  // back substitution, chunks from the end
  double nu[ADI_CH], ng[ADI_CH];
#pragma unroll
  for (int e = 0; e < ADI_CH; e++) {
    const int k = n - 2 - e;
    const long long q = o + (long long)k * st;
    nu[e] = k >= 0 ? U[q] : 0.0;
    ng[e] = k >= 0 ? G[q + st] : 0.0;
  }
  for (int k0 = n - 2; k0 >= 0; k0 -= ADI_CH) {
    double u[ADI_CH], g[ADI_CH];
#pragma unroll
    for (int e = 0; e < ADI_CH; e++) { u[e] = nu[e]; g[e] = ng[e]; }
#pragma unroll
    for (int e = 0; e < ADI_CH; e++) {
      const int k = k0 - ADI_CH - e;
      const long long q = o + (long long)k * st;
      nu[e] = k >= 0 ? U[q] : 0.0;
      ng[e] = k >= 0 ? G[q + st] : 0.0;
    }
#pragma unroll
    for (int e = 0; e < ADI_CH; e++) {
      const int k = k0 - e;
      if (k >= 0) {
        uprev = u[e] - g[e] * uprev;
        U[o + (long long)k * st] = uprev;
      }
    }
  }
  return true;
}

// x pass (rows j = 1..ny-2): a band of up to 64 consecutive rows per one-wave workgroup, one row
// per lane.  Rows are contiguous in memory, so a lane walking its own row would touch a new cache
// line every few steps while its neighbours touch others; instead the band's diagonal, right-hand
// side and couplings are staged through LDS in tiles of ADX_K columns, read as whole row segments
// (the next tile in flight while the current one is solved).  The forward elimination writes gam
// and the intermediate solution to band-transposed scratch (column k of lane t at band start +
// k*R + t, so a wave's stores are contiguous), and the back substitution returns the solution to
// row-major zint through LDS.  Arithmetic is the Thomas recurrence of thomas_line.
#define ADX_K 16
#define ADX_RS (64 / ADX_K)  // rows covered by one staging load of the wave
__global__ __launch_bounds__(64) void k_adi_x(const InstParams* __restrict__ PP, const double* __restrict__ dg,
                                              const double* __restrict__ rf, const double* __restrict__ cx,
                                              double* __restrict__ zint, double* __restrict__ GT,
                                              double* __restrict__ UT, int* __restrict__ err) {
  constexpr int LD = ADX_K + 1;  // padded LDS row: lanes reading column c of their rows spread over the banks
  constexpr int NM = 64 / ADX_RS;
  __shared__ double sD[64 * LD], sR[64 * LD], sC[64 * LD];
  const int inst = blockIdx.y;
  const InstParams& P = PP[inst];
  const int j0 = 1 + blockIdx.x * 64;
  if (!P.active || !P.run_diff || j0 > P.ny - 2) return;
  const int nx = P.nx, R = min(64, P.ny - 1 - j0), lane = threadIdx.x;
  const bool own = lane < R;
  const long long base = (long long)P.off + (long long)j0 * nx;
  const int lc = lane % ADX_K, lr = lane / ADX_K;  // staging: column lc of rows lr, lr+ADX_RS, ...
  double pd[NM], pr[NM], pc[NM];
  auto fetch = [&](int k0) {
#pragma unroll
    for (int m = 0; m < NM; m++) {
      const int rr = lr + ADX_RS * m, k = k0 + lc;
      const bool in = rr < R && k < nx;
      const long long q = base + (long long)rr * nx + k;
      pd[m] = in ? dg[q] : 0.0;
      pr[m] = in ? rf[q] : 0.0;
      pc[m] = (in && k >= 1) ? cx[q - 1] : 0.0;
    }
  };
  fetch(0);
  double bet = 1.0, uprev = 0.0;
  bool ok = true;
  for (int k0 = 0; k0 < nx; k0 += ADX_K) {
#pragma unroll
    for (int m = 0; m < NM; m++) {
      const int a = (lr + ADX_RS * m) * LD + lc;
      sD[a] = pd[m];
      sR[a] = pr[m];
      sC[a] = pc[m];
    }
    __syncthreads();
    if (k0 + ADX_K < nx) fetch(k0 + ADX_K);
    const int kn = min(ADX_K, nx - k0);
    if (own && ok) {
      for (int c = 0; c < kn; c++) {
        const int k = k0 + c;
        const double d = sD[lane * LD + c], r = sR[lane * LD + c];
        if (k == 0) {
          bet = d;
          if (bet == 0.0) { ok = false; break; }
          uprev = r / bet;
        } else {
          const double cc = sC[lane * LD + c];
          const double cprev = (k == 1 && P.fix_left) ? 0.0 : -cc;
          const double inf = (k == nx - 1 && P.fix_right) ? 0.0 : -cc;
          const double gm = cprev / bet;
          bet = d - inf * gm;
          if (bet == 0.0) { ok = false; break; }
          uprev = (r - inf * uprev) / bet;
          GT[base + (long long)k * R + lane] = gm;
        }
        UT[base + (long long)k * R + lane] = uprev;
      }
    }
    __syncthreads();
  }
  if (own && !ok) atomicOr(err + inst, FSX_ERR_TRIDAG);
  // back substitution u(k) = u(k) - gam(k+1)*u(k+1), k = nx-2..0, tiles from the last; uprev == u(nx-1)
  double nu[ADX_K], ng[ADX_K];
  auto fetch_back = [&](int k0) {
#pragma unroll
    for (int c = 0; c < ADX_K; c++) {
      const int k = k0 + c;
      const bool in = own && k < nx - 1;
      nu[c] = in ? UT[base + (long long)k * R + lane] : 0.0;
      ng[c] = in ? GT[base + (long long)(k + 1) * R + lane] : 0.0;
    }
  };
  const int last0 = ((nx - 1) / ADX_K) * ADX_K;
  fetch_back(last0);
  for (int k0 = last0; k0 >= 0; k0 -= ADX_K) {
    double u[ADX_K], g[ADX_K];
#pragma unroll
    for (int c = 0; c < ADX_K; c++) { u[c] = nu[c]; g[c] = ng[c]; }
    if (k0 > 0) fetch_back(k0 - ADX_K);
#pragma unroll
    for (int c = ADX_K - 1; c >= 0; c--) {
      const int k = k0 + c;
      if (k < nx - 1) uprev = u[c] - g[c] * uprev;
      if (k < nx) sD[lane * LD + c] = uprev;
    }
    __syncthreads();
#pragma unroll
    for (int m = 0; m < NM; m++) {
      const int rr = lr + ADX_RS * m, k = k0 + lc;
      if (rr < R && k < nx) zint[base + (long long)rr * nx + k] = sD[(lr + ADX_RS * m) * LD + lc];
    }
    __syncthreads();
  }
}

// y pass: one thread per column i = 1..nx-2, writes zintp for those columns
__global__ __launch_bounds__(64) void k_adi_y(const InstParams* __restrict__ PP, const double* __restrict__ dg,
                                              const double* __restrict__ rf, const double* __restrict__ cy,
                                              double* __restrict__ zintp, double* __restrict__ gam, int* __restrict__ err) {
  const int inst = blockIdx.y;
  const InstParams& P = PP[inst];
  const int i = 1 + blockIdx.x * 64 + threadIdx.x;
  if (!P.active || !P.run_diff || i > P.nx - 2) return;
  const long long o = (long long)P.off + i;
  if (!thomas_line(dg, rf, cy, zintp, gam, o, P.nx, P.ny, P.fix_bottom, P.fix_top)) atomicOr(err + inst, FSX_ERR_TRIDAG);
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
// The tile table holds tiles of S.tile_rows rows (the height it was built with).
void launch_receivers(DevState& S, hipStream_t st, int uplift) {
  if (!S.ntiles) return;
#define ROUTE(TH) k_route_tile<TH><<<S.ntiles, 256, 0, st>>>(S.P, S.tiles, S.h, S.b, S.u, S.ht, S.rec, S.length, uplift)
  switch (S.tile_rows) {
    case 4: ROUTE(4); break;
    case 16: ROUTE(16); break;
    case 32: ROUTE(32); break;
    default: ROUTE(8); break;
  }
#undef ROUTE
}
void launch_donors(DevState& S, hipStream_t st) {
  if (!S.ntiles) return;
#define DONORS(TH) k_donors_tile<TH><<<S.ntiles, 256, 0, st>>>(S.P, S.tiles, S.rec, S.dmask)
  switch (S.tile_rows) {
    case 4: DONORS(4); break;
    case 16: DONORS(16); break;
    case 32: DONORS(32); break;
    default: DONORS(8); break;
  }
#undef DONORS
}
void launch_spl_prep(DevState& S, hipStream_t st) {
  k_spl_prep<<<S.nchunks, FSX_CHUNK, 0, st>>>(S.P, S.chunks, S.h, S.b, S.kf, S.kfint, S.g, S.ht, S.lacc_lo, S.lacc_hi);
}
void launch_lake_acc(DevState& S, hipStream_t st) {
  k_lake_acc<<<S.nchunks, FSX_CHUNK, 0, st>>>(S.P, S.chunks, S.lakec, S.lsill, S.scal, S.lacc_lo, S.lacc_hi);
}
void launch_spl_post(DevState& S, hipStream_t st) {
  k_spl_post<<<S.nchunks, FSX_CHUNK, 0, st>>>(S.P, S.chunks, S.h, S.b, S.ht, S.etot, S.erate, S.sedflux, S.lsill,
                                               S.lacc_lo, S.lacc_hi, S.lsed, S.water, S.scal, S.pos, S.Hl, S.HWl, S.hwater,
                                               S.lake_depth, S.mpos);
}
void launch_diffusion(DevState& S, hipStream_t st) {
  k_adi_xprep<<<S.nchunks, FSX_CHUNK, 0, st>>>(S.P, S.chunks, S.h, S.b, S.kd, S.kdint, S.adx, S.ady, S.adg, S.adf);
  dim3 gx((S.max_ny + 63) / 64, S.ninst), gy((S.max_nx + 63) / 64, S.ninst);
  k_adi_x<<<gx, 64, 0, st>>>(S.P, S.adg, S.adf, S.adx, S.zint, S.gam, S.zintp, S.err);  // zintp: scratch until the y pass
  k_adi_yrhs<<<S.nchunks, FSX_CHUNK, 0, st>>>(S.P, S.chunks, S.h, S.zint, S.kdint, S.adx, S.ady, S.adg, S.adf);
  k_adi_y<<<gy, 64, 0, st>>>(S.P, S.adg, S.adf, S.ady, S.zintp, S.gam, S.err);
  k_adi_post<<<S.nchunks, FSX_CHUNK, 0, st>>>(S.P, S.chunks, S.h, S.b, S.etot, S.erate, S.zintp);
}
