// fsx_strati.hip -- stratigraphic horizons on the GPU: the reflectors' update inside every step (run_Strati,
// src/FastScape_ctx.f90:809-830) and the heavy parts of an output step's fields (Strati,
// src/Strati.f90:31-50): the slope of each reflector and every node's distance to the shore line of a
// surface (distance_to_shore, :86-127), which the reference finds by a loop over all node pairs.  The
// host completes the fields and writes the files.
#include "fsx_state.h"

// run_Strati: reflectors uplifted (r + u*dt), the erosion below the active and buried reflectors
// accumulated (max(0, r - h)), buried ones cut down to h, the active one and those above set to h.
__global__ __launch_bounds__(FSX_CHUNK) void k_strati_step(const InstParams* __restrict__ PP, const Chunk* __restrict__ chunks,
                                                           const double* __restrict__ h, const double* __restrict__ u) {
  const Chunk c = chunks[chunk_id()];
  const InstParams& P = PP[c.inst];
  const int v = c.start + (int)threadIdx.x;
  if (!P.active || !P.run_strati || v >= P.nn) return;
  const long long g = (long long)P.off + v;
  const double hv = h[g], ud = u[g] * P.dt;
  const int nn = P.nn, nref = P.nref, iref = P.iref;
  for (int i = 0; i < nref; i++) {
    double* r = P.srefl + (long long)i * nn + v;
    double x = *r + ud;
    if (i < iref) {  // reflectors 1..ireflector (1-based)
      double* e = P.serod + (long long)i * nn + v;
      *e = *e + fmax(0.0, x - hv);
    }
    if (i < iref - 1)
      x = x < hv ? x : hv;
    else
      x = hv;
    *r = x;
  }
}

void launch_strati_step(DevState& S, hipStream_t st) {
  k_strati_step<<<S.nchunks, FSX_CHUNK, 0, st>>>(S.P, S.chunks, S.h, S.u);
}

__global__ __launch_bounds__(256) void k_strati_slope(const InstParams* __restrict__ PP, int inst,
                                                      const double* __restrict__ surf, double con,
                                                      double* __restrict__ out) {
  const InstParams& P = PP[inst];
  const int ij = blockIdx.x * blockDim.x + threadIdx.x;
  if (ij >= P.nn) return;
  out[ij] = slope_at(surf, ij, P.nx, P.ny, P.dx, P.dy, con);
}

void launch_strati_slope(DevState& S, int inst, int nn, const double* surf, double* out, hipStream_t st) {
  k_strati_slope<<<(nn + 255) / 256, 256, 0, st>>>(S.P, inst, surf, 45.0 / atan(1.0), out);
}

// x and y of node ij as distance_to_shore computes them: xl*float(i-1)/(nx-1), yl*float(j-1)/(ny-1)
__device__ __forceinline__ double shore_x(int i, int nx, double xl) { return xl * (double)i / (double)(nx - 1); }

// shore crossings: nodes whose surface changes sign towards their receiver, the crossing at h/(h(rec)-h)
// of the way (0.5 when level); the list order does not matter (a minimum is taken over it)
__global__ __launch_bounds__(256) void k_shore_pts(const InstParams* __restrict__ PP, int inst, double xl, double yl,
                                                   const double* __restrict__ surf, const int* __restrict__ rec,
                                                   double* __restrict__ xs, double* __restrict__ ys,
                                                   int* __restrict__ cnt) {
  const InstParams& P = PP[inst];
  const int ij = blockIdx.x * blockDim.x + threadIdx.x;
  if (ij >= P.nn) return;
  const int r = rec[(long long)P.off + ij];
  const double hv = surf[ij], hr = surf[r];
  if (!(hv * hr < 0.0)) return;
  double rat = 0.5;
  if (hr - hv != 0.0) rat = hv / (hr - hv);
  const int nx = P.nx, ny = P.ny;
  const double x0 = shore_x(ij % nx, nx, xl), y0 = shore_x(ij / nx, ny, yl);
  const double x1 = shore_x(r % nx, nx, xl), y1 = shore_x(r / nx, ny, yl);
  const int k = atomicAdd(cnt, 1);
  xs[k] = x0 + rat * (x1 - x0);
  ys[k] = y0 + rat * (y1 - y0);
}

// every node's distance to the nearest crossing: squared distances from 2(xl^2+yl^2) down, then the root
__global__ __launch_bounds__(256) void k_shore_dist(const InstParams* __restrict__ PP, int inst, double xl, double yl,
                                                    const double* __restrict__ xs, const double* __restrict__ ys,
                                                    const int* __restrict__ cnt, double* __restrict__ out) {
  __shared__ double sx[256], sy[256];
  const InstParams& P = PP[inst];
  const int ij = blockIdx.x * blockDim.x + threadIdx.x;
  const int nx = P.nx, ny = P.ny, n = *cnt;
  const double x = shore_x(ij % nx, nx, xl), y = shore_x(ij / nx, ny, yl);
  double d = 2.0 * (xl * xl + yl * yl);
  for (int t0 = 0; t0 < n; t0 += 256) {
    __syncthreads();
    if (t0 + (int)threadIdx.x < n) {
      sx[threadIdx.x] = xs[t0 + threadIdx.x];
      sy[threadIdx.x] = ys[t0 + threadIdx.x];
    }
    __syncthreads();
    const int m = min(256, n - t0);
    for (int k = 0; k < m; k++) {
      const double a = x - sx[k], b = y - sy[k];
      const double dist = a * a + b * b;
      d = d < dist ? d : dist;
    }
  }
  if (ij >= P.nn) return;
  out[ij] = d > 0.0 ? sqrt(d) : d;
}

void launch_strati_shore(DevState& S, int inst, int nn, double xl, double yl, const double* surf, double* xs,
                         double* ys, int* cnt, double* out, hipStream_t st) {
  const int blocks = (nn + 255) / 256;
  launch_fill_int(cnt, 1, 0, st);
  k_shore_pts<<<blocks, 256, 0, st>>>(S.P, inst, xl, yl, surf, S.rec, xs, ys, cnt);
  k_shore_dist<<<blocks, 256, 0, st>>>(S.P, inst, xl, yl, xs, ys, cnt, out);
}
