// fsx_advect.hip -- horizontal advection of h, b and etot (Advect, src/Advect.f90:1-103), the first
// process of a step for landscapes given a velocity field (SetV, src/FastScape_ctx.f90:616-632).
//
// Implicit upwind scheme, one tridiagonal system per grid line: every row (x) first, then every column
// (y).  Row k of a line's system: diagonal 1 + |v| dt/d, the upwind neighbour's coefficient -|v| dt/d
// (below the diagonal for v > 0, above it for v < 0), the first and last rows the identity.  The
// reference solves the three fields with three calls of tridag on the same matrix; the elimination
// factors (gam, bet) depend on the matrix only, so one thread per line computes them once and runs the
// three right-hand sides together, with tridag's operations in its order (division by bet, never a
// reciprocal), which gives the reference's values bit for bit.  The y pass applies b = min(b, h) as it
// writes each node's final values.  A thread reads its row along the row (columns across the wave are
// coalesced); gam is kept in the diffusion scratch of the node (free until the step's diffusion).
#include "fsx_state.h"

#define ADV_WG 64

template <bool XDIR>
__global__ __launch_bounds__(ADV_WG) void k_advect(const InstParams* __restrict__ PP, const double* __restrict__ vel,
                                                   double* __restrict__ h, double* __restrict__ b,
                                                   double* __restrict__ etot, double* __restrict__ gamv) {
  const InstParams& P = PP[blockIdx.y];
  if (!P.active || !P.run_advect) return;
  const int line = blockIdx.x * ADV_WG + threadIdx.x;
  const int nx = P.nx, ny = P.ny;
  const int n = XDIR ? nx : ny, nl = XDIR ? ny : nx, st = XDIR ? 1 : nx;
  if (line >= nl) return;
  const long long base = P.off + (XDIR ? (long long)line * nx : line);
  const double d = XDIR ? P.dx : P.dy, dt = P.dt;
  // coefficient k of the line: (sub, diag, super)
  auto coef = [&](int k, double& a, double& m, double& s) {
    a = 0.0; m = 1.0; s = 0.0;
    if (k == 0 || k == n - 1) return;  // sup(1)=0, diag(1)=1, diag(n)=1, inf(n)=0; inf(1), sup(n) unused
    const double v = vel[base + (long long)k * st];
    if (v > 0.0) {
      m = 1.0 + v * dt / d;
      a = -(v * dt / d);
    } else if (v < 0.0) {
      m = 1.0 - v * dt / d;
      s = v * dt / d;
    }
  };
  double a, m, s;
  coef(0, a, m, s);
  double bet = m, sprev = s;
  double uh = h[base] / bet, ub = b[base] / bet, ue = etot[base] / bet;
  h[base] = uh;
  b[base] = ub;
  etot[base] = ue;
  for (int k = 1; k < n; k++) {
    const long long g = base + (long long)k * st;
    coef(k, a, m, s);
    const double gam = sprev / bet;
    bet = m - a * gam;
    gamv[g] = gam;
    uh = (h[g] - a * uh) / bet;
    ub = (b[g] - a * ub) / bet;
    ue = (etot[g] - a * ue) / bet;
    h[g] = uh;
    b[g] = ub;
    etot[g] = ue;
    sprev = s;
  }
  if (!XDIR) b[base + (long long)(n - 1) * st] = fmin(ub, uh);
  for (int k = n - 2; k >= 0; k--) {
    const long long g = base + (long long)k * st;
    const double gam = gamv[g + st];
    uh = h[g] - gam * uh;
    ub = b[g] - gam * ub;
    ue = etot[g] - gam * ue;
    h[g] = uh;
    b[g] = XDIR ? ub : fmin(ub, uh);
    etot[g] = ue;
  }
}

void launch_advect(DevState& S, hipStream_t st) {
  const int gx = (S.max_ny + ADV_WG - 1) / ADV_WG, gy = (S.max_nx + ADV_WG - 1) / ADV_WG;
  k_advect<true><<<dim3(gx, S.ninst), ADV_WG, 0, st>>>(S.P, S.vx, S.h, S.b, S.etot, S.gam);
  k_advect<false><<<dim3(gy, S.ninst), ADV_WG, 0, st>>>(S.P, S.vy, S.h, S.b, S.etot, S.gam);
}
