// fsx_readout.hip -- read-outs on the state a step leaves behind: slope, curvature, chi, catchment and
// the three fluxes (FastScape_Copy_Slope / _Curvature / _Chi / _Catchment and FastScape_Get_Fluxes,
// src/FastScape_api.f90:379-439,762-774).  They run on one landscape of a batch, on the context's
// stream, outside any step.
#include <cfloat>

#include "fsx_state.h"

// slope (src/TerrainDerivatives.f90:3-36, which = 0) or curvature (:40-74, which = 1) of one landscape,
// one thread per node; the outer ring is 0.  The 3x3 block is numbered as the reference's curvature
// does it: 1..3 the row above (j+1), 4..6 the node's row, 7..9 the row below, each left to right; the
// sums keep the reference's left-to-right order.  con = 45/atan(1) comes from the host.
__global__ __launch_bounds__(256) void k_terrain(const InstParams* __restrict__ PP, const int* __restrict__ ids, int inst1,
                                                 const double* __restrict__ h, int which, double con,
                                                 double* __restrict__ outb) {
  const int inst = ids ? ids[blockIdx.y] : inst1;
  const InstParams& P = PP[inst];
  double* out = outb + (ids ? P.off : 0);
  const int ij = blockIdx.x * blockDim.x + threadIdx.x;
  if (ij >= P.nn) return;
  const int nx = P.nx, i = ij % nx, j = ij / nx;
  double v = 0.0;
  if (i >= 1 && i < nx - 1 && j >= 1 && j < P.ny - 1) {
    const double* H = h + P.off;
    const double dx = P.dx, dy = P.dy;
    const double h1 = H[ij + nx - 1], h2 = H[ij + nx], h3 = H[ij + nx + 1];
    const double h4 = H[ij - 1], h5 = H[ij], h6 = H[ij + 1];
    const double h7 = H[ij - nx - 1], h8 = H[ij - nx], h9 = H[ij - nx + 1];
    if (which == 0) {
      v = slope_at(H, ij, nx, P.ny, dx, dy, con);
    } else {
      const double a = (h1 + h3 + h4 + h6 + h7 + h9) / dx / dx / 12.0 - (h2 + h5 + h8) / dx / dx / 6.0;
      const double b = (h1 + h2 + h3 + h7 + h8 + h9) / dy / dy / 12.0 - (h4 + h5 + h6) / dy / dy / 6.0;
      const double c = (h3 + h7 - h1 - h9) / dx / dy / 4.0;
      const double d = (h3 + h6 + h9 - h1 - h4 - h7) / dx / 6.0;
      const double e = (h1 + h2 + h3 - h7 - h8 - h9) / dy / 6.0;
      v = 1.0 + d * d + e * e;
      if (v > DBL_MIN) v = (a * (1.0 + e * e) + b * (1.0 + d * d) - c * d * e) / pow(v, 3.0 / 2.0);
    }
  }
  out[ij] = v;
}

// CopyChi (src/FastScape_ctx.f90:226-247): chi = chi(rec) + (a0/a)^(m/n) * length where a > a0, else 0.
// The reference walks the stack; chi of a node needs only its receiver's, so the levels of the
// single-flow forest (base level first) give the same values, one workgroup per landscape.  A
// base-level node has length 0 and its own chi (still 0) as receiver term, hence 0.
__global__ __launch_bounds__(FSX_WG) void k_chi(const InstParams* __restrict__ PP, const int* __restrict__ ids, int inst1,
                                                const int* __restrict__ lstart, const int* __restrict__ nlev,
                                                const int* __restrict__ order, const int* __restrict__ rec,
                                                const double* __restrict__ area, const double* __restrict__ length,
                                                double* __restrict__ outb) {
  const int inst = ids ? ids[blockIdx.x] : inst1;
  const InstParams& P = PP[inst];
  double* out = outb + (ids ? P.off : 0);
  const double a0 = P.dx * P.dy * 10.0, mn = P.m / P.n;
  const long long off = P.off;
  const int* ls = lstart + P.aux;
  const int L = nlev[inst];
  for (int l = 0; l < L; l++) {
    for (int p = ls[l] + (int)threadIdx.x; p < ls[l + 1]; p += FSX_WG) {
      const int v = order[off + p];
      const double a = area[off + v];
      double c = 0.0;
      if (a > a0) c = (l == 0 ? 0.0 : out[rec[off + v]]) + pow(a0 / a, mn) * length[off + v];
      out[v] = c;
    }
    __syncthreads();
  }
}

// CopyCatchment (src/FastScape_ctx.f90:297-307): the reference's catch is catch0 of the root of the
// node's tree in the receivers before closed-basin rerouting (find_stack propagates it,
// src/FlowRouting.f90:408,437), with catch0 from random_number at SetUp (:94).  Here catch0 of a node
// is (node + 1) / nn, so a catchment reads as its outlet; which nodes share a value is the reference's.
// The roots come from the basin labels of the step's rerouting (signed: open basins -(b+1)) and the
// outlet of each basin.
__global__ __launch_bounds__(256) void k_catchment(const InstParams* __restrict__ PP, const int* __restrict__ ids,
                                                   int inst1, const int* __restrict__ basin,
                                                   const int* __restrict__ outlet, double* __restrict__ outb) {
  const int inst = ids ? ids[blockIdx.y] : inst1;
  const InstParams& P = PP[inst];
  double* out = outb + (ids ? P.off : 0);
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= P.nn) return;
  const int bv = basin[P.off + v];
  const int b = bv >= 0 ? bv : -bv - 1;
  out[v] = (double)(outlet[P.off + b] + 1) / (double)P.nn;
}

__device__ __forceinline__ double block_sum(double x, double* red) {
  for (int o = 32; o > 0; o >>= 1) x += __shfl_down(x, o, 64);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = x;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0)
    for (int w = 0; w < FSX_WG / 64; w++) s += red[w];
  return s;
}

// compute_fluxes (src/FastScape_ctx.f90:843-886), one workgroup per landscape.  flux starts as the
// erosion rate; a node, once every donor has delivered, is clamped at 0 and pushed to its
// receivers: single flow by levels from the deepest up, a base-level node finally adding its own
// flux to itself (the reference's reversed stack visits it last, with rec = itself); multiple flow
// by the Kahn layers, donors first, weighted by the receiver weights.  Receivers gather their donors'
// shares with double atomics, so the order of those additions (and of the three sums below, a tree
// reduction where the reference adds in index order) differs from the reference's: the results
// agree to rounding.  sums[0..2] = sum(u), sum(erate), sum(flux over fixed boundary nodes).
__global__ __launch_bounds__(FSX_WG) void k_fluxes(const InstParams* __restrict__ PP, const int* __restrict__ ids,
                                                   int inst1, const int* __restrict__ lstart, const int* __restrict__ nlev,
                                                   const int* __restrict__ order, const int* __restrict__ rec,
                                                   const int* __restrict__ mls, const int* __restrict__ mnlev,
                                                   const int* __restrict__ mstack, const int* __restrict__ mrec,
                                                   const int* __restrict__ mnrec, const double* __restrict__ mw,
                                                   const double* __restrict__ u, const double* __restrict__ erate,
                                                   const uint8_t* __restrict__ bc, double* __restrict__ fluxb,
                                                   double* __restrict__ sumsb) {
  __shared__ double red[FSX_WG / 64];
  const int inst = ids ? ids[blockIdx.x] : inst1;
  const InstParams& P = PP[inst];
  double* flux = fluxb + (ids ? P.off : 0);
  double* sums = sumsb + (ids ? 3 * blockIdx.x : 0);
  const long long off = P.off;
  const int nn = P.nn;
  for (int v = threadIdx.x; v < nn; v += FSX_WG) flux[v] = erate[off + v];
  __syncthreads();
  if (P.sfd) {
    const int* ls = lstart + P.aux;
    for (int l = nlev[inst] - 1; l >= 0; l--) {
      for (int p = ls[l] + (int)threadIdx.x; p < ls[l + 1]; p += FSX_WG) {
        const int v = order[off + p];
        const double f = fmax(0.0, flux[v]);
        if (l == 0) {
          flux[v] = f + f;
        } else {
          flux[v] = f;
          atomicAdd(flux + rec[off + v], f);
        }
      }
      __syncthreads();
    }
  } else {
    const int* ls = mls + P.aux;
    const int L = mnlev[inst];
    for (int l = 0; l < L; l++) {
      for (int p = ls[l] + (int)threadIdx.x; p < ls[l + 1]; p += FSX_WG) {
        const int v = mstack[off + p];
        const long long gv = off + v;
        const double f = fmax(0.0, flux[v]);
        flux[v] = f;
        for (int k = 0; k < mnrec[gv]; k++) atomicAdd(flux + mrec[8 * gv + k], f * mw[8 * gv + k]);
      }
      __syncthreads();
    }
  }
  double su = 0.0, se = 0.0, sb = 0.0;
  for (int v = threadIdx.x; v < nn; v += FSX_WG) {
    su += u[off + v];
    se += erate[off + v];
    if (bc[off + v]) sb += flux[v];
  }
  su = block_sum(su, red);
  se = block_sum(se, red);
  sb = block_sum(sb, red);
  if (threadIdx.x == 0) {
    sums[0] = su;
    sums[1] = se;
    sums[2] = sb;
  }
}

// ids: a device list of n landscapes (outputs at each landscape's node offset, fluxes 3 per list entry),
// or nullptr for the one landscape inst1 (outputs from index 0)
void launch_terrain(DevState& S, const int* ids, int n, int inst1, int which, double con, double* out,
                    hipStream_t st) {
  const int nn = ids ? S.max_nn : n;
  k_terrain<<<dim3((nn + 255) / 256, ids ? n : 1), 256, 0, st>>>(S.P, ids, inst1, S.h, which, con, out);
}

void launch_catchment(DevState& S, const int* ids, int n, int inst1, double* out, hipStream_t st) {
  const int nn = ids ? S.max_nn : n;
  k_catchment<<<dim3((nn + 255) / 256, ids ? n : 1), 256, 0, st>>>(S.P, ids, inst1, S.basin, S.outlet, out);
}

void launch_chi(DevState& S, const int* ids, int n, int inst1, double* out, hipStream_t st) {
  k_chi<<<ids ? n : 1, FSX_WG, 0, st>>>(S.P, ids, inst1, S.lstart, S.nlev, S.order, S.rec, S.area, S.length, out);
}

void launch_fluxes(DevState& S, const int* ids, int n, int inst1, double* flux, double* sums, hipStream_t st) {
  k_fluxes<<<ids ? n : 1, FSX_WG, 0, st>>>(S.P, ids, inst1, S.lstart, S.nlev, S.order, S.rec, S.mls, S.mnlev,
                                           S.mstack, S.mrec, S.mnrec, S.mw, S.u, S.erate, S.bc, flux, sums);
}
