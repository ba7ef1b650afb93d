// fsx_internal.h -- shared layout, parameter block and device helpers of the MI355X landscape step.
//
// A context holds a batch of independent landscapes ("instances").  Every per-node array is one
// flat allocation over the batch; instance k owns nodes [P[k].off, P[k].off + P[k].nn).  Node
// indices stored inside arrays (receivers, labels, pass nodes) are LOCAL to their instance, so
// every per-instance algorithm reads like the single-landscape reference.
//
// Per-basin arrays reuse the node offsets (a landscape has at most nn basins); per-edge arrays use
// EDGE_CAP*off (the reference sizes its basin graph as 6*nbasins, src/LocalMinima.f90:59).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define FSX_EDGE_CAP 6   // edges per node of capacity (nconn_max = 6*nbasins in the reference)
#define FSX_SLAB_CAP 8   // candidate-pass slots per closed-basin node
#define FSX_CHUNK 256    // nodes per block in node-parallel kernels
#define FSX_WG 1024      // threads of the per-landscape workgroups
#define FSX_TW 64        // 2-D tiles of the stencil kernels: columns (rows: the context's tile_h, 4 to 32)
#define FSX_MFD_CTR 16   // counters per landscape of the multiple-flow dataflow launches
#define FSX_LVL_BIG 131072  // landscapes of at least this many nodes are level-sorted by several workgroups
#define FSX_LVL_SEG 16384   // nodes per workgroup of that sort

// device-side error bits, one word per instance
enum {
  FSX_ERR_NCONN = 1,       // more basin-graph edges than 6*nbasins (reference would overflow)
  FSX_ERR_NO_OUTLET = 2,   // no basin drains to a fixed node (reference reads basin(-1))
  FSX_ERR_TRIDAG = 4,      // zero pivot in the tridiagonal solve (reference STOPs)
  FSX_ERR_NEWTON = 8,      // Newton iteration did not converge within the cap
  FSX_ERR_SORTSTACK = 16,  // segment stack of the edge sort overflowed
  FSX_ERR_LEVELS = 32,     // receiver graph has a cycle (stack incomplete)
  FSX_ERR_MFD = 64,        // multiple-flow-direction order incomplete (reference STOPs)
  FSX_ERR_MARINE = 128,    // silt / sand diffusion did not converge in 1000 iterations (reference STOPs)
  FSX_ERR_NONFINITE = 256, // a pass height (elevation) is not a number
};

struct InstParams {
  int nx, ny, nn, off;       // grid, node count, node offset in the batch arrays
  int i1, i2, j1, j2;        // 0-based inclusive receiver-search bounds (SetBC)
  int xcyc, ycyc;            // cyclic flags (SetBC)
  int fix_left, fix_right, fix_bottom, fix_top;  // BC digits == '1'
  int run_uplift, run_spl, run_diff, sfd;
  int run_advect;             // a velocity field was set (SetV)
  int active;
  int flow;                  // single flow without deposition (g == 0 everywhere): one stream-power pass,
                             // solved node by node as receivers complete (k_spl_flow)
  int flow_g;                // workgroups of k_spl_flow on this landscape
  int mfd_g;                 // workgroups of the multiple-flow dataflow sweeps on this landscape
  int lvl_segs;              // segments of the multi-workgroup level sort (0: sorted by one workgroup)
  int run_marine;            // marine transport after Diffusion (SetMarineParam)
  int run_strati, nref, iref;  // stratigraphy (Activate_Strati): reflectors, the active one (1-based)
  int aux;                   // offset of the nn+1-entry per-landscape arrays (level starts, groups)
  int c0, c1;                // chunk-table range of the landscape
  double dx, dy, dt;
  double len[9];             // neighbour distance by direction (ii+1)+3*(jj+1), host-computed
  double dxdy;               // dx*dy
  double kfsed, m, n, kdsed, g1, g2, p, newton_tol;
  double p_raw;              // p as given to the erosional parameters (p_mfd_exp of single-flow landscapes)
  double sealevel, poro1, poro2, zporo1, zporo2, ratio, layer, kdsea1, kdsea2;  // SetMarineParam
  double *srefl, *serod;     // stratigraphy: reflectors and the erosion below each (nref x nn, device)
};

struct Chunk {
  int inst, start;
};

// Per-position records of the single-flow sweeps, one aligned block per position so that a prep
// kernel writes whole segments and a sweep reads one contiguous record.
// What the basin walk reads of a node, in one 16-byte load: elevation, signed basin id, and the next
// node of its tree's depth-first preorder (-1 after the last).
struct alignas(16) WalkNode {
  double h;
  int basin, succ;
};

struct alignas(16) AreaRec64 {  // drainage area: donor positions (relative, -1 unused), own term
  int d[8];
  double a0;
  double pad[3];
};
struct alignas(16) SplRec32 {  // stream power: node, receiver position (relative), factor, ht
  int o, q;
  double f, t;
  double pad;
};

// ------------------------------------------------------------------------------------------
// Neighbour of (i,j) in direction (ii,jj) with cyclic wrap else clamp, as find_receiver
// (src/FlowRouting.f90:336-345).  0-based.
__device__ __forceinline__ int nbr_wc(const InstParams& P, int i, int j, int ii, int jj) {
  int iii = i + ii;
  if (iii < 0 && P.xcyc) iii += P.nx;
  if (iii < 0) iii = 0;
  if (iii > P.nx - 1 && P.xcyc) iii -= P.nx;
  if (iii > P.nx - 1) iii = P.nx - 1;
  int jjj = j + jj;
  if (jjj < 0 && P.ycyc) jjj += P.ny;
  if (jjj < 0) jjj = 0;
  if (jjj > P.ny - 1 && P.ycyc) jjj -= P.ny;
  if (jjj > P.ny - 1) jjj = P.ny - 1;
  return iii + jjj * P.nx;
}

// Distinct wrap/clamp neighbours of v (excluding v), in direction order d = (ii+1)+3*(jj+1).
// Receivers of every node (before and after closed-basin rerouting) are among these, so the
// donors of v are exactly the q in this set with rec[q] == v.  Returns the count; q[] and dir[].
__device__ __forceinline__ int neighbours_wc(const InstParams& P, int v, int* q, int* dir) {
  int i = v % P.nx, j = v / P.nx, c = 0;
  if (i > 0 && j > 0 && i < P.nx - 1 && j < P.ny - 1) {  // interior: no wrap, no clamp, no duplicates
#pragma unroll
    for (int d = 0; d < 9; d++) {
      if (d == 4) continue;
      q[c] = v + (d % 3 - 1) + (d / 3 - 1) * P.nx;
      dir[c] = d;
      c++;
    }
    return c;
  }
  for (int jj = -1; jj <= 1; jj++)
    for (int ii = -1; ii <= 1; ii++) {
      if (ii == 0 && jj == 0) continue;
      int w = nbr_wc(P, i, j, ii, jj);
      if (w == v) continue;
      bool dup = false;
      for (int t = 0; t < c; t++) dup |= (q[t] == w);
      if (dup) continue;
      q[c] = w;
      dir[c] = (ii + 1) + 3 * (jj + 1);
      c++;
    }
  return c;
}

// slope in degrees (con = 45/atan(1)) of node ij of the surface H (slope, src/TerrainDerivatives.f90:3-36):
// Sobel-type gradient of the 3x3 block, atan of its norm; 0 on the outer ring
__device__ __forceinline__ double slope_at(const double* H, int ij, int nx, int ny, double dx, double dy, double con) {
  const int i = ij % nx, j = ij / nx;
  if (!(i >= 1 && i < nx - 1 && j >= 1 && j < ny - 1)) return 0.0;
  const double h1 = H[ij + nx - 1], h2 = H[ij + nx], h3 = H[ij + nx + 1];
  const double h4 = H[ij - 1], h6 = H[ij + 1];
  const double h7 = H[ij - nx - 1], h8 = H[ij - nx], h9 = H[ij - nx + 1];
  const double dzdx = ((h3 + 2.0 * h6 + h9) - (h1 + 2.0 * h4 + h7)) / 8.0 / dx;
  const double dzdy = ((h7 + 2.0 * h8 + h9) - (h1 + 2.0 * h2 + h3)) / 8.0 / dy;
  double v = dzdx * dzdx + dzdy * dzdy;
  if (v > 2.2250738585072014e-308) v = atan(sqrt(v)) * con;
  return v;
}

// bit index of direction d in the 8-bit donor mask
__device__ __forceinline__ int dbit(int d) { return d < 4 ? d : d - 1; }

// Donors of v in ascending local index (the reference's donor order, src/FlowRouting.f90:379-385).
__device__ __forceinline__ int donors_sorted(const InstParams& P, int v, unsigned mask, int* d) {
  if (!mask) return 0;
  const int i = v % P.nx, j = v / P.nx;
  if (i > 0 && j > 0 && i < P.nx - 1 && j < P.ny - 1) {  // interior: bit order is ascending index order
    int nd = 0;
#pragma unroll
    for (int b = 0; b < 8; b++)
      if (mask & (1u << b)) {
        const int dd = b < 4 ? b : b + 1;
        d[nd++] = v + (dd % 3 - 1) + (dd / 3 - 1) * P.nx;
      }
    return nd;
  }
  int q[8], dir[8];
  int c = neighbours_wc(P, v, q, dir), nd = 0;
  for (int t = 0; t < c; t++)
    if (mask & (1u << dbit(dir[t]))) {
      int x = q[t], s = nd++;
      while (s > 0 && d[s - 1] > x) { d[s] = d[s - 1]; s--; }
      d[s] = x;
    }
  return nd;
}

// x**n and x**(n-1) of the Newton step of the stream-power law (src/StreamPowerLaw.f90:352-353), x >= 0.
// For the exponents the configurations use, the libm power is replaced by its exact-arithmetic
// equivalent (n = 2: x*x and x; n = 1.5: x*sqrt(x) and sqrt(x)), within an ulp of pow and far
// cheaper inside the serial Newton loop.  kind: 0 general, 1 n == 2, 2 n == 1.5 (npow_kind).
__device__ __forceinline__ int npow_kind(double n) { return n == 2.0 ? 1 : (n == 1.5 ? 2 : 0); }
__device__ __forceinline__ void npow(double x, double n, int kind, double& pn, double& pn1) {
  if (kind == 1) {
    pn = x * x;
    pn1 = x;
  } else if (kind == 2) {
    const double r = sqrt(x);
    pn = x * r;
    pn1 = r;
  } else {
    pn = pow(x, n);
    pn1 = pow(x, n - 1.0);
  }
}

// XCD-aware block decomposition.  The dispatcher hands workgroups to the 8 XCDs of the device in
// round-robin order of their linear id, and each XCD has its own L2; a kernel that runs per_inst
// blocks for every landscape renumbers its blocks so that all blocks of a landscape land on one XCD
// (its scattered reads then share that L2).  Launch with xcd_grid(ninst, per_inst) 1-D blocks;
// returns false for the padding blocks.
// Pointer-jumping launches (two hops each).  A hop adds the pointed-to pair's distance, which at the
// start of the launch is at least the launch's guaranteed reach D, so one launch multiplies D by at
// least 3 (2D after the first hop, + D after the second): 20 launches reach 3^20 > 2^31 nodes and a
// 21st changes nothing unless the pointers form a cycle (flags[FSX_JUMP_ROUNDS - 1] then reports it).
#define FSX_JUMP_ROUNDS 21
#define FSX_NXCD 8
__host__ __device__ __forceinline__ int xcd_grid(int ninst, int per_inst) {
  return FSX_NXCD * ((ninst + FSX_NXCD - 1) / FSX_NXCD) * per_inst;
}
__device__ __forceinline__ bool xcd_split(int ninst, int per_inst, int& inst, int& sub) {
  const int lin = (int)blockIdx.x, x = lin % FSX_NXCD, r = lin / FSX_NXCD;
  const int i = r / per_inst;
  sub = r - i * per_inst;
  inst = i * FSX_NXCD + x;
  return inst < ninst;
}

// Chunk handled by this block of a node-parallel launch (one block per chunk, gridDim.x == number
// of chunks).  Block b runs on XCD b % 8; chunk ids are handed out so that each XCD works through one
// contiguous run of chunks: neighbouring rows of a stencil, and the nodes of a landscape, share that
// XCD's L2 instead of being fetched by all eight.
__device__ __forceinline__ int chunk_id() {
  const int n = (int)gridDim.x, b = (int)blockIdx.x, x = b % FSX_NXCD, k = b / FSX_NXCD;
  const int q = n / FSX_NXCD, r = n % FSX_NXCD;
  return x * q + (x < r ? x : r) + k;
}

// ---- values passed between waves inside one launch ------------------------------------------------
// A value of a dataflow sweep that is not written yet: a NaN whose payload no arithmetic produces from
// the finite (or canonical NaN) values the sweeps compute.
#define FLOW_SENT 0xFFF8DEADBEEF0001ull
__device__ __forceinline__ bool flow_sent(double x) {
  return (unsigned long long)__double_as_longlong(x) == FLOW_SENT;
}
__device__ __forceinline__ double flow_sentinel() { return __longlong_as_double((long long)FLOW_SENT); }
// between waves of one workgroup (one compute unit)
__device__ __forceinline__ double flow_ld(const double* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ void flow_st(double* p, double v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
// between workgroups of one landscape (other compute units, possibly other XCDs): device-scope loads
// and stores of one location are coherent, so a value read differs from the sentinel only once written
__device__ __forceinline__ double flow_ld_dev(const double* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void flow_st_dev(double* p, double v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ int flow_ldi_dev(const int* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void flow_sti_dev(int* p, int v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ unsigned long long flow_ldu_dev(const unsigned long long* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void flow_stu_dev(unsigned long long* p, unsigned long long v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// Waits until every memory access this wave issued has completed: device-scope stores are then visible
// to device-scope loads of any compute unit, so a counter incremented after this publishes them.
__device__ __forceinline__ void wave_drain() { asm volatile("s_waitcnt vmcnt(0)" ::: "memory"); }

// order-preserving 64-bit image of a double (for atomic min/max on doubles); 0 is below every image
__device__ __forceinline__ unsigned long long dkey(double x) {
  unsigned long long u = (unsigned long long)__double_as_longlong(x);
  return (u & 0x8000000000000000ull) ? ~u : (u | 0x8000000000000000ull);
}
__device__ __forceinline__ double dval(unsigned long long k) {
  const unsigned long long u = (k & 0x8000000000000000ull) ? (k & ~0x8000000000000000ull) : ~k;
  return __longlong_as_double((long long)u);
}

// ------------------------------------------------------------------------------------------
// Records msg as the calling thread's last error (fsx_last_error) and returns code.
int fsx_fail(int code, const char* msg);

// Host-side launch wrappers (defined in the .hip files).
struct DevState;
void launch_advect(DevState& S, hipStream_t st);  // needs S.vx/S.vy
void launch_receivers(DevState& S, hipStream_t st, int uplift);  // uplift fused in when uplift != 0
void launch_donors(DevState& S, hipStream_t st);  // donor mask of every node
void launch_lm_labels(DevState& S, hipStream_t st);
void launch_lm_basins(DevState& S, hipStream_t st);
void launch_lm_walk(DevState& S, hipStream_t st);
void launch_lm_edges(DevState& S, hipStream_t st);
void launch_lm_mst(DevState& S, hipStream_t st);
void launch_lm_orient(DevState& S, hipStream_t st);
void launch_lm_correct(DevState& S, hipStream_t st);
void launch_nr_sort(DevState& S, hipStream_t st);
void launch_spl_prep(DevState& S, hipStream_t st);
void launch_spl_post(DevState& S, hipStream_t st);
void launch_lake_acc(DevState& S, hipStream_t st);
void launch_expand_donors(DevState& S, hipStream_t st);
void launch_levels(DevState& S, hipStream_t st);
void launch_jump(DevState& S, hipStream_t st);
void launch_fill_int(int* a, long long n, int val, hipStream_t st);
void launch_area(DevState& S, hipStream_t st);
void launch_area_prep(DevState& S, hipStream_t st);
void launch_area_sweep(DevState& S, hipStream_t st);
void launch_spl_fact(DevState& S, hipStream_t st);
void launch_spl_sweep(DevState& S, hipStream_t st);
void launch_spl_flow(DevState& S, hipStream_t st, int mode);  // mode: 0 step, 1 fill only, 2 timing
void launch_spl(DevState& S, hipStream_t st);
void launch_diffusion(DevState& S, hipStream_t st);
void launch_mfd_route(DevState& S, hipStream_t st);
void launch_mfd_spl(DevState& S, hipStream_t st);
void launch_marine(DevState& S, hipStream_t st);  // needs the marine buffers
void launch_strati_step(DevState& S, hipStream_t st);
// stratigraphy read-outs of landscape inst (outside a step; surf and out hold the landscape's nn nodes):
// slope of surf; distance to shore of surf along the receivers (shore points staged in xs / ys, their
// count in cnt; xl, yl the landscape's extent)
void launch_strati_slope(DevState& S, int inst, int nn, const double* surf, double* out, hipStream_t st);
void launch_strati_shore(DevState& S, int inst, int nn, double xl, double yl, const double* surf, double* xs,
                         double* ys, int* cnt, double* out, hipStream_t st);
// read-outs (outside a step) of the landscapes listed in ids (device memory, n entries; outputs at each
// landscape's node offset), or of the one landscape inst1 of n nodes when ids is null (outputs from 0)
void launch_terrain(DevState& S, const int* ids, int n, int inst1, int which, double con, double* out, hipStream_t st);
void launch_chi(DevState& S, const int* ids, int n, int inst1, double* out, hipStream_t st);
void launch_catchment(DevState& S, const int* ids, int n, int inst1, double* out, hipStream_t st);
void launch_fluxes(DevState& S, const int* ids, int n, int inst1, double* flux, double* sums, hipStream_t st);
int sweep_set_lds(int smem_max);
int basins_set_lds(int smem_max);
void launch_fill(DevState& S, hipStream_t st);

// Phases timed with HIP events when profiling is on (the reference times Uplift, SPL and Diffusion
// with cpu_time, src/FastScape_api.f90:207-257).  Phases marked * are exactly one kernel.
enum Phase {
  PH_UPLIFT = 0,   //   (uplift runs fused into the receiver phase)
  PH_RECEIVERS,    // * k_route_tile (uplift + D8 receivers)
  PH_DONORS,       // * k_donors (before and after rerouting)
  PH_LM_LABELS,    //   pointer jumping for basin roots
  PH_LM_BASINS,    //   basin ids, sizes
  PH_LM_WALK,      // * k_basin_walk
  PH_LM_EDGES,     //   basin-graph assembly
  PH_LM_SORT,      //   edge order (k_nr_big, k_nr_small, k_nr_pos)
  PH_LM_MST,       // * k_boruvka (+ a fill)
  PH_LM_ORIENT,    // * k_orient
  PH_LM_CORRECT,   // * k_correct
  PH_LEVELS,       //   depth by pointer jumping + k_level_sort
  PH_AREA_PREP,    // * k_area_prep
  PH_AREA_SWEEP,   // * k_area_sweep
  PH_SPL_PREP,     // * k_spl_prep
  PH_SPL_FACT,     // * k_spl_fact
  PH_SPL_SWEEP,    // * k_spl_sweep
  PH_SPL_POST,     // * k_spl_post
  PH_DIFFUSION,    //   k_adi_prep, k_adi_x, k_adi_y, k_adi_post
  PH_MFD_ROUTE,    //   multiple flow direction: fill, receivers, donors, Kahn layers, accumulation
  PH_MFD_SPL,      //   multiple flow direction: edge factors + Gauss-Seidel stream power
  PH_ADVECT,       // * k_advect (first in the step when a velocity field is set)
  PH_MARINE,       //   marine transport: shelf flux, silt / sand diffusion
  PH_STRATI,       // * k_strati_step: reflectors (last in the step)
  PH_COUNT
};
