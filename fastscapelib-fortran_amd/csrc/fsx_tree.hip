// fsx_tree.hip -- dependency-ordered sweeps over the receiver forest of each landscape.
//
// The reference walks a depth-first stack serially (src/FlowRouting.f90:393-443).  Here:
//  * depth of every node (hops to its base-level node) by pointer jumping over (distance, ancestor)
//    pairs packed in 64 bits, node-parallel over the whole batch;
//  * a counting sort by depth gives the level order (level 0 = base-level nodes) and each node's
//    position in it; node-parallel prep kernels turn the per-node inputs into per-position records
//    (receiver position in the previous level, donor positions in the next level, stream-power
//    factor) so that the sweeps read them coalesced;
//  * one workgroup per landscape then sweeps the levels, keeping the values of the adjacent level in
//    LDS: bottom-up for drainage area (each node pulls its donors in descending index, the
//    reference's summation order, src/FlowRouting.f90:136-139) and top-down for the lake fill and
//    the stream-power solve (each node needs only its receiver).
// Nodes of one level are independent, so their order inside the level does not affect any value.
#include <algorithm>
#include <type_traits>

#include "fsx_scan.h"
#include "fsx_state.h"

#define NODE_PROLOGUE                         \
  const Chunk c = chunks[chunk_id()];         \
  const InstParams& P = PP[c.inst];           \
  const int v = c.start + (int)threadIdx.x;   \
  if (!P.active || !P.run_spl || v >= P.nn) return; \
  const long long g_ = (long long)P.off + v;

__device__ __forceinline__ unsigned long long jpack(unsigned dist, unsigned nxt) {
  return ((unsigned long long)dist << 32) | nxt;
}
__device__ __forceinline__ int jnext(unsigned long long w) { return (int)(unsigned)(w & 0xffffffffu); }
__device__ __forceinline__ int jdist(unsigned long long w) { return (int)(w >> 32); }

// ---- depth / root by pointer jumping ----------------------------------------------------------
__global__ __launch_bounds__(FSX_CHUNK) void k_jump_init(const InstParams* __restrict__ PP, const Chunk* __restrict__ chunks,
                                                         const int* __restrict__ rec, unsigned long long* __restrict__ J) {
  NODE_PROLOGUE
  const int r = rec[g_];
  J[g_] = jpack(r == v ? 0u : 1u, (unsigned)r);
}

// J[v] = (distance to ancestor, ancestor).  One jump replaces the ancestor by its own ancestor;
// a 64-bit load/store keeps the pair consistent, so jumping in place is safe.
// One pointer-jumping round (two hops).  A round after the one that changed nothing returns at once; the grid
// strides over the chunks so that such a launch costs few workgroups.
__global__ __launch_bounds__(FSX_CHUNK) void k_jump(const InstParams* __restrict__ PP, const Chunk* __restrict__ chunks,
                                                    int nchunks, unsigned long long* __restrict__ J,
                                                    int* __restrict__ flags, int iter) {
  if (iter > 0 && flags[iter - 1] == 0) return;
  bool changed = false;
  for (int cb = blockIdx.x; cb < nchunks; cb += gridDim.x) {
    const Chunk c = chunks[cb];
    const InstParams& P = PP[c.inst];
    const int v = c.start + (int)threadIdx.x;
    if (!P.active || !P.run_spl || v >= P.nn) continue;
    unsigned long long* Ji = J + P.off;
    // two hops per round: any (distance, ancestor) pair read is valid, whoever wrote it when, so a
    // thread may jump again from its own new pair
    unsigned long long w = Ji[v];
#pragma unroll
    for (int hop = 0; hop < 2; hop++) {
      const int x = jnext(w);
      const unsigned long long wx = Ji[x];
      if (jnext(wx) == x) break;
      w = jpack((unsigned)(jdist(w) + jdist(wx)), (unsigned)jnext(wx));
      Ji[v] = w;
      changed = true;
    }
  }
  if (changed) flags[iter] = 1;
}

// ---- level order (counting sort by depth, one workgroup per landscape) ---------------------------
// Counts live in LDS (levels >= LVL_HCAP spill to global memory).  The same kernel partitions the
// levels into groups of consecutive levels that start inside one bucket of GRP_BUCKET positions
// (so a group has at most GRP_BUCKET levels and GRP_STG nodes; a level wider than GRP_HALF forms its
// own group); the sweeps stage one group at a time into LDS.
#define LVL_HCAP 32768
#define LVL_UNROLL 32
#define GRP_HALF 1024                // a level wider than this is a group of its own
#define GRP_BUCKET 1024              // levels starting in one bucket of positions share a group
#define GRP_STG (GRP_HALF + GRP_BUCKET)  // positions of a group (staging capacity)

__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Level groups of one landscape from its level starts ls[0..L] (one workgroup): a new group starts at
// level 0, at or after a wide level, and where the start position crosses a multiple of GRP_HALF.
__device__ void level_groups(const InstParams& P, const int* ls, int L, int inst, int* __restrict__ gmeta,
                             int* __restrict__ ngroups, int4* __restrict__ grec, int* lds) {
  const int tid = threadIdx.x;
  int* gm = gmeta + P.aux;  // <= L+1 entries
  int gcarry = 0;
  for (int s0 = 0; s0 < L; s0 += FSX_WG) {
    const int l = s0 + tid;
    int f = 0;
    if (l < L) {
      const int a0 = ls[l], w = ls[l + 1] - a0;
      if (l == 0) f = 1;
      else {
        const int pa = ls[l - 1], pw = a0 - pa;
        f = (w > GRP_HALF) || (pw > GRP_HALF) || (a0 / GRP_BUCKET != pa / GRP_BUCKET);
      }
    }
    int tot;
    const int incl = block_incl_scan<FSX_WG>(f, lds, tot);
    if (f) gm[gcarry + incl - 1] = l;
    gcarry += tot;
  }
  if (tid == 0) {
    gm[gcarry] = L;
    ngroups[inst] = gcarry;
  }
  __syncthreads();
  // per-group records (lo, hi, p0, p1) + (pm, p2, -, -): level range, position range, start of the
  // level below and end of the level above
  int4* gr = grec + 2 * (long long)P.aux;
  for (int g = tid; g < gcarry; g += FSX_WG) {
    const int lo = gm[g], hi = gm[g + 1];
    const int p0 = ls[lo], p1 = ls[hi];
    gr[2 * g] = make_int4(lo, hi, p0, p1);
    gr[2 * g + 1] = make_int4(lo > 0 ? ls[lo - 1] : 0, hi < L ? ls[hi + 1] : p1, 0, 0);
  }
}

__global__ __launch_bounds__(FSX_WG) void k_level_sort(const InstParams* __restrict__ PP,
                                                       const unsigned long long* __restrict__ J,
                                                       const int* __restrict__ flags, int* __restrict__ lstart,
                                                       int* __restrict__ lcur, int* __restrict__ nlev,
                                                       int* __restrict__ gmeta, int* __restrict__ ngroups,
                                                       int4* __restrict__ grec, int* __restrict__ order,
                                                       int* __restrict__ pos, int* __restrict__ err) {
  extern __shared__ int hist[];
  __shared__ int lds[FSX_WG / 64];
  __shared__ int s_max;
  const int inst = blockIdx.x;
  const InstParams& P = PP[inst];
  if (!P.active || !P.run_spl || P.lvl_segs) return;  // large landscapes: k_lvl_count .. k_lvl_place
  const int nn = P.nn, tid = threadIdx.x;
  const long long off = P.off;
  int* ls = lstart + P.aux;  // nn+1 entries
  int* lc = lcur + P.aux;
  if (flags[FSX_JUMP_ROUNDS - 1]) {  // jumping did not converge: a receiver cycle, so depths are meaningless
    if (tid == 0) {
      atomicOr(err + inst, FSX_ERR_LEVELS);
      nlev[inst] = 0;  // the sweeps then visit no level
      ngroups[inst] = 0;
    }
    return;
  }
  // one pass over the nodes: the histogram of depths below LVL_HCAP (in LDS) and the largest depth;
  // depths from LVL_HCAP on (very deep trees only) are counted in global memory by a second pass.
  // The node loops take LVL_UNROLL nodes per thread at a time so their loads overlap.
  for (int l = tid; l < LVL_HCAP; l += FSX_WG) hist[l] = 0;
  if (tid == 0) s_max = 0;
  __syncthreads();
  int mx = 0;
  for (int v0 = tid; v0 < nn; v0 += LVL_UNROLL * FSX_WG) {
    int dd[LVL_UNROLL];
#pragma unroll
    for (int u = 0; u < LVL_UNROLL; u++) {
      const int v = v0 + u * FSX_WG;
      dd[u] = v < nn ? jdist(J[off + v]) : -1;
    }
#pragma unroll
    for (int u = 0; u < LVL_UNROLL; u++) {
      if (v0 + u * FSX_WG >= nn) continue;
      const int d = dd[u] < 0 ? 0 : (dd[u] >= nn ? nn - 1 : dd[u]);
      mx = d > mx ? d : mx;
      if (d < LVL_HCAP) atomicAdd(hist + d, 1);
    }
  }
  for (int o = 32; o > 0; o >>= 1) { const int y = __shfl_xor(mx, o, 64); mx = y > mx ? y : mx; }
  if ((tid & 63) == 0) atomicMax(&s_max, mx);
  __syncthreads();
  const int L = s_max + 1;
  if (L > LVL_HCAP) {
    for (int l = LVL_HCAP + tid; l < L; l += FSX_WG) lc[l] = 0;
    __syncthreads();
    for (int v0 = tid; v0 < nn; v0 += LVL_UNROLL * FSX_WG) {
      int dd[LVL_UNROLL];
#pragma unroll
      for (int u = 0; u < LVL_UNROLL; u++) {
        const int v = v0 + u * FSX_WG;
        dd[u] = v < nn ? jdist(J[off + v]) : -1;
      }
#pragma unroll
      for (int u = 0; u < LVL_UNROLL; u++) {
        if (v0 + u * FSX_WG >= nn) continue;
        const int d = dd[u] < 0 ? 0 : (dd[u] >= nn ? nn - 1 : dd[u]);
        if (d >= LVL_HCAP) atomicAdd(lc + d, 1);
      }
    }
  }
  __syncthreads();
  int carry = 0;
  for (int s0 = 0; s0 < L; s0 += FSX_WG) {
    const int l = s0 + tid;
    const int x = l < L ? (l < LVL_HCAP ? hist[l] : lc[l]) : 0;
    int tot;
    const int incl = block_incl_scan<FSX_WG>(x, lds, tot);
    if (l < L) {
      const int st = carry + incl - x;
      ls[l] = st;
      if (l < LVL_HCAP) hist[l] = st; else lc[l] = st;
    }
    carry += tot;
  }
  if (tid == 0) {
    ls[L] = carry;
    nlev[inst] = L;
  }
  __syncthreads();
  for (int v0 = tid; v0 < nn; v0 += LVL_UNROLL * FSX_WG) {
    int dd[LVL_UNROLL];
#pragma unroll
    for (int u = 0; u < LVL_UNROLL; u++) {
      const int v = v0 + u * FSX_WG;
      dd[u] = v < nn ? jdist(J[off + v]) : -1;
    }
#pragma unroll
    for (int u = 0; u < LVL_UNROLL; u++) {
      const int v = v0 + u * FSX_WG;
      if (v >= nn) continue;
      const int d = dd[u] < 0 ? 0 : (dd[u] >= nn ? nn - 1 : dd[u]);
      const int p = d < LVL_HCAP ? atomicAdd(hist + d, 1) : atomicAdd(lc + d, 1);
      if (p < nn) {
        order[off + p] = v;
        pos[off + v] = p;
      }
    }
  }
  __syncthreads();
  level_groups(P, ls, L, inst, gmeta, ngroups, grec, lds);
}

// ---- level order of the large landscapes (several workgroups each) -----------------------------
// A landscape of at least FSX_LVL_BIG nodes is cut into segments of FSX_LVL_SEG nodes, one workgroup
// each.  k_lvl_zero clears the landscape's depth counts; k_lvl_count adds each segment's depth histogram
// (depths below LVL_BINS summed in LDS first: one global add per depth and segment) and the largest
// depth; k_lvl_scan (one workgroup per landscape) turns the counts into level starts and cursors and
// forms the level groups; k_lvl_place gives every node its position, a segment reserving each depth's
// slots with one global add and dealing them out by LDS rank.  Positions within a level depend on
// timing, as in k_level_sort; no result depends on them.
#define LVL_BINS 8192
#define LVL_PER (FSX_LVL_SEG / FSX_WG)  // nodes per thread

__device__ __forceinline__ int clamp_depth(int d, int nn) { return d < 0 ? 0 : (d >= nn ? nn - 1 : d); }

__global__ __launch_bounds__(FSX_WG) void k_lvl_zero(const InstParams* __restrict__ PP, const int2* __restrict__ tab,
                                                     int* __restrict__ lcur, int* __restrict__ nlev) {
  const int2 e = tab[blockIdx.x];
  const InstParams& P = PP[e.x];
  if (!P.active || !P.run_spl) return;
  const int v0 = e.y * FSX_LVL_SEG, v1 = min(P.nn, v0 + FSX_LVL_SEG);
  int* lc = lcur + P.aux;
  for (int v = v0 + (int)threadIdx.x; v < v1; v += FSX_WG) lc[v] = 0;
  if (e.y == 0 && threadIdx.x == 0) nlev[e.x] = 0;  // the largest depth, raised by k_lvl_count
}

__global__ __launch_bounds__(FSX_WG) void k_lvl_count(const InstParams* __restrict__ PP, const int2* __restrict__ tab,
                                                      const unsigned long long* __restrict__ J,
                                                      const int* __restrict__ flags, int* __restrict__ lcur,
                                                      int* __restrict__ nlev) {
  __shared__ int bins[LVL_BINS];
  __shared__ int s_mb;
  const int2 e = tab[blockIdx.x];
  const InstParams& P = PP[e.x];
  if (!P.active || !P.run_spl || flags[FSX_JUMP_ROUNDS - 1]) return;
  const int nn = P.nn, tid = threadIdx.x, v0 = e.y * FSX_LVL_SEG;
  const long long off = P.off;
  int* lc = lcur + P.aux;
  int dd[LVL_PER];
#pragma unroll
  for (int u = 0; u < LVL_PER; u++) {
    const int v = v0 + u * FSX_WG + tid;
    dd[u] = v < nn ? jdist(J[off + v]) : -1;
  }
  for (int l = tid; l < LVL_BINS; l += FSX_WG) bins[l] = 0;
  if (tid == 0) s_mb = 0;
  __syncthreads();
  int mx = 0, mb = 0;
#pragma unroll
  for (int u = 0; u < LVL_PER; u++) {
    if (v0 + u * FSX_WG + tid >= nn) continue;
    const int d = clamp_depth(dd[u], nn);
    mx = d > mx ? d : mx;
    if (d < LVL_BINS) {
      mb = d > mb ? d : mb;
      atomicAdd(bins + d, 1);
    } else {
      atomicAdd(lc + d, 1);
    }
  }
  for (int o = 32; o > 0; o >>= 1) {
    const int y = __shfl_xor(mx, o, 64), z = __shfl_xor(mb, o, 64);
    mx = y > mx ? y : mx;
    mb = z > mb ? z : mb;
  }
  if ((tid & 63) == 0) {
    atomicMax(nlev + e.x, mx);
    atomicMax(&s_mb, mb);
  }
  __syncthreads();
  for (int l = tid; l <= s_mb; l += FSX_WG) {
    const int x = bins[l];
    if (x) atomicAdd(lc + l, x);
  }
}

__global__ __launch_bounds__(FSX_WG) void k_lvl_scan(const InstParams* __restrict__ PP, const int2* __restrict__ tab,
                                                     const int* __restrict__ flags, int* __restrict__ lstart,
                                                     int* __restrict__ lcur, int* __restrict__ nlev,
                                                     int* __restrict__ gmeta, int* __restrict__ ngroups,
                                                     int4* __restrict__ grec, int* __restrict__ err) {
  __shared__ int lds[FSX_WG / 64];
  const int inst = tab[blockIdx.x].x;
  const InstParams& P = PP[inst];
  if (!P.active || !P.run_spl) return;
  const int tid = threadIdx.x;
  if (flags[FSX_JUMP_ROUNDS - 1]) {  // jumping did not converge: a receiver cycle, so depths are meaningless
    if (tid == 0) {
      atomicOr(err + inst, FSX_ERR_LEVELS);
      nlev[inst] = 0;
      ngroups[inst] = 0;
    }
    return;
  }
  int* ls = lstart + P.aux;
  int* lc = lcur + P.aux;
  const int L = nlev[inst] + 1;
  __syncthreads();  // every thread has read the largest depth before thread 0 replaces it by L
  int carry = 0;
  for (int s0 = 0; s0 < L; s0 += FSX_WG) {
    const int l = s0 + tid;
    const int x = l < L ? lc[l] : 0;
    int tot;
    const int incl = block_incl_scan<FSX_WG>(x, lds, tot);
    if (l < L) {
      const int st = carry + incl - x;
      ls[l] = st;
      lc[l] = st;
    }
    carry += tot;
  }
  if (tid == 0) {
    ls[L] = carry;
    nlev[inst] = L;
  }
  __syncthreads();
  level_groups(P, ls, L, inst, gmeta, ngroups, grec, lds);
}

__global__ __launch_bounds__(FSX_WG) void k_lvl_place(const InstParams* __restrict__ PP, const int2* __restrict__ tab,
                                                      const unsigned long long* __restrict__ J,
                                                      const int* __restrict__ flags, int* __restrict__ lcur,
                                                      int* __restrict__ order, int* __restrict__ pos) {
  __shared__ int bins[LVL_BINS];
  __shared__ int s_mb;
  const int2 e = tab[blockIdx.x];
  const InstParams& P = PP[e.x];
  if (!P.active || !P.run_spl || flags[FSX_JUMP_ROUNDS - 1]) return;
  const int nn = P.nn, tid = threadIdx.x, v0 = e.y * FSX_LVL_SEG;
  const long long off = P.off;
  int* lc = lcur + P.aux;
  int dd[LVL_PER], rk[LVL_PER];
#pragma unroll
  for (int u = 0; u < LVL_PER; u++) {
    const int v = v0 + u * FSX_WG + tid;
    dd[u] = v < nn ? jdist(J[off + v]) : -1;
  }
  for (int l = tid; l < LVL_BINS; l += FSX_WG) bins[l] = 0;
  if (tid == 0) s_mb = 0;
  __syncthreads();
  // rank of each node among the segment's nodes of its depth (shallow depths, in LDS), or its position
  // straight from the global cursor (deep depths)
  int mb = 0;
#pragma unroll
  for (int u = 0; u < LVL_PER; u++) {
    rk[u] = 0;
    if (v0 + u * FSX_WG + tid >= nn) continue;
    const int d = clamp_depth(dd[u], nn);
    dd[u] = d;
    if (d < LVL_BINS) {
      mb = d > mb ? d : mb;
      rk[u] = atomicAdd(bins + d, 1);
    } else {
      rk[u] = atomicAdd(lc + d, 1);
    }
  }
  for (int o = 32; o > 0; o >>= 1) { const int z = __shfl_xor(mb, o, 64); mb = z > mb ? z : mb; }
  if ((tid & 63) == 0) atomicMax(&s_mb, mb);
  __syncthreads();
  // one reservation per shallow depth: the bin becomes the segment's first position at that depth
  for (int l = tid; l <= s_mb; l += FSX_WG) {
    const int x = bins[l];
    if (x) bins[l] = atomicAdd(lc + l, x);
  }
  __syncthreads();
#pragma unroll
  for (int u = 0; u < LVL_PER; u++) {
    const int v = v0 + u * FSX_WG + tid;
    if (v >= nn) continue;
    const int d = dd[u];
    const int p = d < LVL_BINS ? bins[d] + rk[u] : rk[u];
    if (p < nn) {
      order[off + p] = v;
      pos[off + v] = p;
    }
  }
}

// ---- per-position records -----------------------------------------------------------------------

// Drainage-area inputs: own term dx*dy*precip and the positions (relative to the next level's start)
// of the donors in descending node index; unused slots -1.
__global__ __launch_bounds__(FSX_CHUNK) void k_area_prep(const InstParams* __restrict__ PP, const Chunk* __restrict__ chunks,
                                                         const unsigned long long* __restrict__ J,
                                                         const uint8_t* __restrict__ dmask, const int* __restrict__ pos,
                                                         const int* __restrict__ lstart, const double* __restrict__ precip,
                                                         const int* __restrict__ flags, AreaRec64* __restrict__ arec) {
  NODE_PROLOGUE
  if (!P.sfd || flags[FSX_JUMP_ROUNDS - 1]) return;  // no level order this step (receiver cycle)
  const int p = pos[g_];
  const unsigned long long jw = J[g_];
  const unsigned mask = dmask[g_];
  const double pr = precip[g_];
  const long long gp = (long long)P.off + p;
  const int i = v % P.nx, j = v / P.nx;
  int o[8];
  if (i > 0 && j > 0 && i < P.nx - 1 && j < P.ny - 1) {
    // interior: mask bit b is the b-th neighbour in ascending index order; the donors' positions are
    // loaded into fixed slots together and placed by rank in registers (descending index order)
    int pp[8];
#pragma unroll
    for (int b = 0; b < 8; b++) {
      const int dd = b < 4 ? b : b + 1;
      pp[b] = ((mask >> b) & 1u) ? pos[g_ + (dd % 3 - 1) + (long long)(dd / 3 - 1) * P.nx] : 0;
    }
    const int base = lstart[P.aux + jdist(jw) + 1];
    const int nd = __popc(mask);
#pragma unroll
    for (int k = 0; k < 8; k++) {
      int x = -1;
#pragma unroll
      for (int b = 0; b < 8; b++)
        if (((mask >> b) & 1u) && __popc(mask & ((1u << b) - 1u)) == nd - 1 - k) x = pp[b] - base;
      o[k] = x;
    }
  } else {
    int ds[8];
    const int nd = donors_sorted(P, v, mask, ds);
    const int base = lstart[P.aux + jdist(jw) + 1];
#pragma unroll
    for (int k = 0; k < 8; k++) o[k] = k < nd ? pos[P.off + ds[nd - 1 - k]] - base : -1;
  }
  int4* w = reinterpret_cast<int4*>(arec + gp);
  w[0] = make_int4(o[0], o[1], o[2], o[3]);
  w[1] = make_int4(o[4], o[5], o[6], o[7]);
  reinterpret_cast<double2*>(arec + gp)[2] = make_double2(P.dxdy * pr, 0.0);
  reinterpret_cast<double2*>(arec + gp)[3] = make_double2(0.0, 0.0);
}

// Stream-power inputs per position: receiver position in the previous level, ht, and
// fact = kf*dt*a**m/length (n == 1, src/StreamPowerLaw.f90:310) or kf*dt*a**m/length**n with the
// activity test ht(ijk) > ht(ijr) of the Newton branch (:350-351; -1 marks an inactive node).
__global__ __launch_bounds__(FSX_CHUNK) void k_spl_fact(const InstParams* __restrict__ PP, const Chunk* __restrict__ chunks,
                                                        const unsigned long long* __restrict__ J, const int* __restrict__ rec,
                                                        const int* __restrict__ pos, const int* __restrict__ lstart,
                                                        const double* __restrict__ Al, double* __restrict__ area,
                                                        const double* __restrict__ length,
                                                        const double* __restrict__ kfint, const double* __restrict__ ht,
                                                        SplRec32* __restrict__ srec, double* __restrict__ lsed,
                                                        const double* __restrict__ g, unsigned long long* __restrict__ redk,
                                                        const int* __restrict__ flags, double* __restrict__ Hl,
                                                        double* __restrict__ Wl, double* __restrict__ HWl) {
  __shared__ unsigned long long smax[2][FSX_CHUNK / 64];
  const Chunk c = chunks[chunk_id()];
  const InstParams& P = PP[c.inst];
  const int v = c.start + (int)threadIdx.x;
  if (!P.active || !P.run_spl || !P.sfd) return;  // uniform over the block: one chunk, one landscape
  const bool live = v < P.nn;
  const long long g_ = (long long)P.off + v;
  // the node's level-order inputs, loaded before the block reduction so their latency overlaps it
  const bool ordered = live && !flags[FSX_JUMP_ROUNDS - 1];
  const int p = ordered ? pos[g_] : 0;
  const int d = ordered ? jdist(J[g_]) : 0;
  const int r = ordered ? rec[g_] : 0;
  const double kfv = ordered ? kfint[g_] : 0.0, lv = ordered ? length[g_] : 0.0;
  // the landscape's maxval(abs(h)) and maxval(g) (src/StreamPowerLaw.f90:246,376), h == ht here:
  // reduced over the block, then one atomic per block
  unsigned long long kh = live ? dkey(fabs(ht[g_])) : 0ull, kg = live ? dkey(g[g_]) : 0ull;
  for (int o = 32; o > 0; o >>= 1) {
    const unsigned long long a = __shfl_xor(kh, o, 64), b = __shfl_xor(kg, o, 64);
    kh = a > kh ? a : kh;
    kg = b > kg ? b : kg;
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
  if (!ordered) return;  // no level order this step (receiver cycle)
  const long long gp = (long long)P.off + p;
  // the loads that follow the position, depth and receiver, issued together
  const double t = ht[g_];
  const double av = Al[gp];  // drainage area, level order -> grid order
  const int pr = d > 0 ? pos[P.off + r] : 0;
  const int l0 = (d > 0 && !P.flow) ? lstart[P.aux + d - 1] : 0;
  const double tr = d > 0 ? ht[P.off + r] : 0.0;
  area[g_] = av;
  lsed[g_] = 0.0;  // no lake sediment before the first Gauss-Seidel pass
  int q = -1;
  double f = 0.0;
  if (d > 0) {
    // the level sweep reads the receiver relative to its level's start, the dataflow sweep absolutely
    q = pr - l0;
    const double am = pow(av, P.m);
    const double n = P.n;
    if (fabs(n - 1.0) < 2.2250738585072014e-308) {
      f = kfv * P.dt * am / lv;
    } else {
      f = (t > tr) ? kfv * P.dt * am / pow(lv, n) : -1.0;
    }
  }
  ulonglong2* w = reinterpret_cast<ulonglong2*>(srec + gp);
  w[0] = make_ulonglong2((unsigned long long)(unsigned)v | ((unsigned long long)(unsigned)q << 32),
                         (unsigned long long)__double_as_longlong(f));
  w[1] = make_ulonglong2((unsigned long long)__double_as_longlong(t), 0ull);
  if (P.flow) {  // not yet solved: k_spl_flow waits for a receiver's values to leave the sentinel
    const double sx = __longlong_as_double((long long)FLOW_SENT);
    Hl[gp] = sx;
    Wl[gp] = sx;
    HWl[gp] = sx;
  }
}

// ---- staged level groups ---------------------------------------------------------------------------
// The sweeps visit level groups (k_level_sort).  Group g covers levels [lo, hi) and positions
// [p0, p1).  Its per-position records are prefetched into registers one group ahead (two slots per
// thread cover GRP_STG positions) and written to an LDS staging area when the group starts, so the
// only global-memory latency left on the critical path is one load per group, hidden behind the
// previous group.  Levels inside a group are separated by LDS-only barriers.
struct GroupMeta {
  int lo, hi, p0, p1;
  int pm;  // start of level lo-1 (0 when lo == 0)
  int p2;  // end of level hi (== p1 when hi is the last level)
};

// Group metadata (two int4 per group) is read by lanes 0..5 of every wave as a vector load and
// broadcast with readlane when the group starts.  A scalar load would share its wait counter with
// the LDS accesses, so while one is in flight every LDS wait of the level loop waits for it too.
__device__ __forceinline__ int meta_lane(const int4* gr, int G, int g) {
  const int lane = threadIdx.x & 63;
  return (g >= 0 && g < G && lane < 6) ? reinterpret_cast<const int*>(gr)[8 * g + lane] : 0;
}
__device__ __forceinline__ GroupMeta meta_bcast(int x) {
  GroupMeta m;
  m.lo = __builtin_amdgcn_readlane(x, 0);
  m.hi = __builtin_amdgcn_readlane(x, 1);
  m.p0 = __builtin_amdgcn_readlane(x, 2);
  m.p1 = __builtin_amdgcn_readlane(x, 3);
  m.pm = __builtin_amdgcn_readlane(x, 4);
  m.p2 = __builtin_amdgcn_readlane(x, 5);
  return m;
}

// ---- bottom-up: drainage area -------------------------------------------------------------------
// FlowAccumulationSingleFlowDirection, src/FlowRouting.f90:123-143: a = dx*dy*precip, donors added
// from the highest index down; a base-level node finally adds its own area to itself (:138 with
// rec(ijk) == ijk), doubling it.
struct AreaRec {
  double a;
  int d[8];
};

__device__ __forceinline__ void area_load(AreaRec& r, long long gp, const AreaRec64* arec) {
  r.a = reinterpret_cast<const double2*>(arec + gp)[2].x;
  const int4* q = reinterpret_cast<const int4*>(arec + gp);
  const int4 x = q[0], y = q[1];
  r.d[0] = x.x; r.d[1] = x.y; r.d[2] = x.z; r.d[3] = x.w;
  r.d[4] = y.x; r.d[5] = y.y; r.d[6] = y.z; r.d[7] = y.w;
}

// WG threads per landscape: NS record slots per thread cover a group's GRP_STG positions, NL
// level-start slots its <= GRP_BUCKET levels.  Fewer threads shorten a level's critical path (fewer
// idle waves sharing each SIMD's issue and the barrier); wide levels then take several positions per
// thread.
template <int WG>
__global__ __launch_bounds__(WG) void k_area_sweep(const InstParams* __restrict__ PP, const int* __restrict__ lstart,
                                                   const int4* __restrict__ grec, const int* __restrict__ ngroups,
                                                   const AreaRec64* __restrict__ arec, double* __restrict__ Al,
                                                   int vcap) {
  constexpr int NS = GRP_STG / WG;
  constexpr int NL = (GRP_BUCKET + WG - 1) / WG;
  static_assert(GRP_STG % WG == 0 && WG >= 64 && WG % 64 == 0, "workgroup size");
  extern __shared__ double sm[];
  // level values double-buffered at sm[0..vcap) and sm[vcap..2*vcap); offsets, not a pointer array,
  // so every access stays an LDS (ds_*) instruction
  double* sA = sm + 2 * vcap;
  int* sD = reinterpret_cast<int*>(sA + GRP_STG);  // donor slot k of staged position t at k*GRP_STG + t
  int* sL = sD + 8 * GRP_STG;  // level starts of the group, GRP_BUCKET+1
  const int inst = blockIdx.x;
  const InstParams& P = PP[inst];
  if (!P.active || !P.run_spl || !P.sfd) return;
  const int tid = threadIdx.x;
  const long long off = P.off;
  const int* ls = lstart + P.aux;
  const int4* gr = grec + 2 * (long long)P.aux;
  const int G = ngroups[inst];
  AreaRec rr[NS];
  int rl[NL];
  auto prefetch = [&](const GroupMeta& m) {
    const int np = m.p1 - m.p0;
    if (np > GRP_STG) return;
#pragma unroll
    for (int u = 0; u < NS; u++)
      if (tid + u * WG < np) area_load(rr[u], off + m.p0 + tid + u * WG, arec);
#pragma unroll
    for (int u = 0; u < NL; u++)
      if (tid + u * WG < m.hi - m.lo) rl[u] = ls[m.lo + tid + u * WG];
  };
  GroupMeta cur = meta_bcast(meta_lane(gr, G, G - 1));
  GroupMeta nxt_m = meta_bcast(meta_lane(gr, G, G - 2));
  prefetch(cur);
  for (int g = G - 1; g >= 0; g--) {
    const int np = cur.p1 - cur.p0;
    const bool staged = np <= GRP_STG;
    if (staged) {
#pragma unroll
      for (int u = 0; u < NS; u++) {
        const int t = tid + u * WG;
        if (t < np) {
          sA[t] = rr[u].a;
#pragma unroll
          for (int k = 0; k < 8; k++) sD[k * GRP_STG + t] = rr[u].d[k];
        }
      }
#pragma unroll
      for (int u = 0; u < NL; u++)
        if (tid + u * WG < cur.hi - cur.lo) sL[tid + u * WG] = rl[u];
      if (tid == 0) sL[cur.hi - cur.lo] = cur.p1;
    }
    lds_barrier();
    const int after = meta_lane(gr, G, g - 2);  // two groups ahead: in flight while this group runs
    if (g > 0) prefetch(nxt_m);
    auto stg = [&](AreaRec& r, int si) {
      r.a = sA[si];
#pragma unroll
      for (int k = 0; k < 8; k++) r.d[k] = sD[k * GRP_STG + si];
    };
    // one level: w positions from b0, record x0 of position tid when staged (positions tid + WG, ...
    // are read from the staging area); the level above (donors) has wn positions from b0 + w
    auto level = [&](int l, int b0, int w, int wn, const AreaRec& x0) {
      double* cv = sm + ((l & 1) ? vcap : 0);
      const double* nv = sm + ((l & 1) ? 0 : vcap);
      const int bn = b0 + w;
      const bool nin = wn <= vcap;
      // FAST: records staged and the level above in LDS -- no global loads, and the eight donor
      // reads are issued together (an unused slot reads entry 0 and adds an exact +0.0, keeping
      // the reference's summation order)
      auto one = [&](int i, const AreaRec& r, auto fast_tag) {
        constexpr bool FAST = decltype(fast_tag)::value;
        double a = r.a;
        double dv[8];
#pragma unroll
        for (int k = 0; k < 8; k++) {
          const int q = r.d[k];
          if constexpr (FAST) dv[k] = nv[q < 0 ? 0 : q];
          else dv[k] = q < 0 ? 0.0 : (nin ? nv[q] : Al[off + bn + q]);
        }
        if constexpr (FAST) {
          __builtin_amdgcn_sched_barrier(0);  // all eight reads in flight before any is used
#pragma unroll
          for (int k = 0; k < 8; k++) dv[k] = r.d[k] < 0 ? 0.0 : dv[k];
        }
#pragma unroll
        for (int k = 0; k < 8; k++) a = a + dv[k];
        if (l == 0) a = a + a;
        if (w <= vcap) cv[i] = a;
        Al[off + b0 + i] = a;
      };
      if (staged && nin) {
        if (tid < w) one(tid, x0, std::true_type{});
        for (int i = tid + WG; i < w; i += WG) {
          AreaRec x;
          stg(x, b0 - cur.p0 + i);
          one(i, x, std::true_type{});
        }
      } else if (staged) {
        if (tid < w) one(tid, x0, std::false_type{});
        for (int i = tid + WG; i < w; i += WG) {
          AreaRec x;
          stg(x, b0 - cur.p0 + i);
          one(i, x, std::false_type{});
        }
      } else {
        for (int i = tid; i < w; i += WG) {
          AreaRec r;
          area_load(r, off + b0 + i, arec);
          one(i, r, std::false_type{});
        }
      }
      if (w > vcap) __syncthreads(); else lds_barrier();
    };
    if (staged) {
      // software-pipelined over the group's levels: the bounds of level l-1 and its records are read
      // from LDS while level l is solved, so a level's critical path is its donor reads alone
      // level starts of the group: lane k of every wave holds that of level lo+k (k < 64), so a
      // level's bounds are a register read (readlane) rather than an LDS round trip
      const int lsv = (tid & 63) <= cur.hi - cur.lo ? sL[tid & 63] : 0;
      auto st = [&](int k) { return k < 64 ? __builtin_amdgcn_readlane(lsv, k) : sL[k]; };
      int l = cur.hi - 1;
      int A = st(l - cur.lo), B = cur.p1, C = cur.p2;   // level l = [A, B), level l+1 = [B, C)
      int Pm = l - 1 >= cur.lo ? st(l - 1 - cur.lo) : 0;  // start of level l-1
      AreaRec x0;
      if (tid < B - A) stg(x0, A - cur.p0 + tid);
      for (; l >= cur.lo; l--) {
        const int Pm2 = l - 2 >= cur.lo ? st(l - 2 - cur.lo) : 0;
        AreaRec y0;
        if (l - 1 >= cur.lo && tid < A - Pm) stg(y0, Pm - cur.p0 + tid);
        level(l, A, B - A, C - B, x0);
        C = B; B = A; A = Pm; Pm = Pm2;
        x0 = y0;
      }
    } else {
      AreaRec none;
      level(cur.lo, cur.p0, cur.p1 - cur.p0, cur.p2 - cur.p1, none);
    }
    cur = nxt_m;
    nxt_m = meta_bcast(after);
  }
}

// ---- top-down: lake fill + stream power ----------------------------------------------------------
// Implicit stream-power law along single receivers, src/StreamPowerLaw.f90:210-400, with the lake
// surface fill of FlowRoutingSingleFlowDirection (src/FlowRouting.f90:73-89) fused into the first
// pass (fill_only: the fill alone, h untouched).  The Gauss-Seidel loop (one pass when g == 0
// everywhere) alternates the bottom-up erosion push dh (:266-283), elev (:285-289) and the
// top-down solve: closed form for n == 1 (:306-314), relaxed Newton with omega = 0.875/n otherwise
// (:343-358); lakes keep elev and collect their water volume.
struct SplRec {
  int o, q;
  double f, t, h;
};

// h of the first Gauss-Seidel pass is ht; later passes start from the previous pass's h
__device__ __forceinline__ void spl_load(SplRec& r, long long gp, const SplRec32* srec, const double* Hl, bool first) {
  const ulonglong2* w = reinterpret_cast<const ulonglong2*>(srec + gp);
  const ulonglong2 a = w[0], b = w[1];
  r.o = (int)(unsigned)(a.x & 0xffffffffull);
  r.q = (int)(unsigned)(a.x >> 32);
  r.f = __longlong_as_double((long long)a.y);
  r.t = __longlong_as_double((long long)b.x);
  r.h = first ? r.t : Hl[gp];
}
__device__ __forceinline__ double srec_t(const SplRec32* srec, long long gp) {
  return __longlong_as_double((long long)reinterpret_cast<const ulonglong2*>(srec + gp)[1].x);
}
__device__ __forceinline__ int srec_o(const SplRec32* srec, long long gp) {
  return (int)(unsigned)(reinterpret_cast<const ulonglong2*>(srec + gp)[0].x & 0xffffffffull);
}

__global__ __launch_bounds__(FSX_WG) void k_spl_sweep(
    const InstParams* __restrict__ PP, const int* __restrict__ lstart, const int* __restrict__ nlev,
    const int4* __restrict__ grec, const int* __restrict__ ngroups, const SplRec32* __restrict__ srec,
    const AreaRec64* __restrict__ arec, const uint8_t* __restrict__ bc, const double* __restrict__ g,
    const double* __restrict__ Al, double* __restrict__ h, double* __restrict__ hwater, double* __restrict__ lake_depth,
    double* __restrict__ water, int* __restrict__ lsill, double* __restrict__ lakec, double* __restrict__ lsed,
    double* __restrict__ Hl, double* __restrict__ Wl, double* __restrict__ HWl, int* __restrict__ LSl,
    double* __restrict__ HPl, double* __restrict__ DHl, double* __restrict__ ELl, double* __restrict__ scal,
    int* __restrict__ ngs_out, long long* __restrict__ newton_out, int* __restrict__ err, int vcap, int fill_only,
    const unsigned long long* __restrict__ redk) {
  extern __shared__ double sm[];
  __shared__ double red[FSX_WG / 64];
  __shared__ double s_tol, s_err, s_gmax;
  __shared__ unsigned long long s_newton;
  const int inst = blockIdx.x;
  const InstParams& P = PP[inst];
  if (!P.active || !P.run_spl || !P.sfd || P.flow) return;  // flow instances: k_spl_flow
  // per-level double buffers (h, water, hwater, lake sill) addressed by offset from the LDS base
#define BH(par) (sm + (par) * vcap)
#define BW(par) (sm + (2 + (par)) * vcap)
#define BHW(par) (sm + (4 + (par)) * vcap)
#define BLS(par) (reinterpret_cast<int*>(sm + 6 * vcap) + (par) * vcap)
  double* sF = sm + 7 * vcap;  // staging: GRP_STG records
  double* sT = sF + GRP_STG;
  double* sH = sT + GRP_STG;
  int* sO = reinterpret_cast<int*>(sH + GRP_STG);
  int* sQ = sO + GRP_STG;
  int* sL = sQ + GRP_STG;  // GRP_BUCKET+1 level starts
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  const int nn = P.nn;
  const long long off = P.off;
  const int* ls = lstart + P.aux;
  const int4* gr = grec + 2 * (long long)P.aux;
  const int L = nlev[inst];
  const int G = ngroups[inst];
  const double n = P.n;
  const bool n_is_one = fabs(n - 1.0) < 2.2250738585072014e-308;
  const int nkind = npow_kind(n);

  // tol = 1e-4*(maxval(abs(h))+1), gmax = maxval(g), reduced by k_spl_fact, which also set up the
  // level-ordered working copy of h
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
    const bool first = ngs == 1;
    if (!gzero && !fill_only) {
      for (int p = tid; p < nn; p += FSX_WG) HPl[off + p] = first ? srec_t(srec, off + p) : Hl[off + p];
      __syncthreads();
      // erosion push, bottom-up (donors in descending index); lake sediment of lake sills
      for (int l = L - 1; l >= 0; l--) {
        const int b0 = ls[l], w = ls[l + 1] - b0;
        const int bn = ls[l + 1], wn = (l + 1 < L) ? ls[l + 2] - bn : 0;
        double* cv = BH(l & 1);
        const double* nv = BH((l + 1) & 1);
        const bool nin = wn <= vcap;
        for (int i = tid; i < w; i += FSX_WG) {
          const long long gp = off + b0 + i;
          const int v = srec_o(srec, gp);
          const double own = srec_t(srec, gp) - HPl[gp];
          double s = own;
          const int* dd = arec[gp].d;
          for (int k = 0; k < 8; k++) {
            const int q = dd[k];
            if (q < 0) break;
            s = s + (nin ? nv[q] : DHl[off + bn + q]);
          }
          double dv;
          if (l > 0) {
            const double t = s - own;
            if (!first && LSl[gp] == v) lsed[off + v] = t <= 0.0 ? 0.0 : t;
            dv = t + own;
          } else {
            dv = s;
            lsed[off + v] = s;
          }
          if (w <= vcap) cv[i] = dv;
          DHl[gp] = dv;
        }
        __syncthreads();
      }
      for (int p = tid; p < nn; p += FSX_WG) {
        const long long gp = off + p;
        const int v = srec_o(srec, gp);
        const double t = srec_t(srec, gp);
        ELl[gp] = bc[off + v] ? t : t + (DHl[gp] - (t - HPl[gp])) * g[off + v] * P.dx * P.dy / Al[gp];
      }
      __syncthreads();
    }
    // solve, base level up, group by group
    SplRec r0, r1;
    int rl = 0;
    auto prefetch = [&](const GroupMeta& m) {
      const int np = m.p1 - m.p0;
      if (np > GRP_STG) return;
      if (tid < np) spl_load(r0, off + m.p0 + tid, srec, Hl, first);
      if (tid + FSX_WG < np) spl_load(r1, off + m.p0 + tid + FSX_WG, srec, Hl, first);
      if (tid < m.hi - m.lo) rl = ls[m.lo + tid];
    };
    GroupMeta cur = meta_bcast(meta_lane(gr, G, 0));
    GroupMeta nxt_m = meta_bcast(meta_lane(gr, G, 1));
    prefetch(cur);
    for (int gg = 0; gg < G; gg++) {
      const int np = cur.p1 - cur.p0;
      const bool staged = np <= GRP_STG;
      if (staged) {
        if (tid < np) { sO[tid] = r0.o; sQ[tid] = r0.q; sF[tid] = r0.f; sT[tid] = r0.t; sH[tid] = r0.h; }
        if (tid + FSX_WG < np) {
          const int t = tid + FSX_WG;
          sO[t] = r1.o; sQ[t] = r1.q; sF[t] = r1.f; sT[t] = r1.t; sH[t] = r1.h;
        }
        if (tid < cur.hi - cur.lo) sL[tid] = rl;
        if (tid == 0) sL[cur.hi - cur.lo] = cur.p1;
      }
      lds_barrier();
      const int after = meta_lane(gr, G, gg + 2);
      if (gg + 1 < G) prefetch(nxt_m);
      auto stg = [&](SplRec& r, int si) {
        r.o = sO[si]; r.q = sQ[si]; r.f = sF[si]; r.t = sT[si]; r.h = sH[si];
      };
      // one level: w positions from b0, receivers in the level from bp; record x0 of position tid
      // when staged (position tid + FSX_WG, rare, is read from the staging area)
      auto level = [&](int l, int b0, int w, int bp, const SplRec& x0) {
        const int wp = l > 0 ? b0 - bp : 0;
        const int cb = l & 1, pb = cb ^ 1;
        const bool pin = wp <= vcap;
        // FAST: staged records, receiver level in LDS, g == 0 -- no global loads in the body
        auto one = [&](int i, const SplRec& rr, auto fast_tag) {
          constexpr bool FAST = decltype(fast_tag)::value;
          const int v = rr.o;
          const double t = rr.t;
          double hv = rr.h, wv, hw = t;
          int sl = v;
          if (l == 0) {
            wv = t;
            if (!gzero && !fill_only) lakec[off + v] = -1.0;  // not a lake node
          } else {
            const int q = rr.q;
            double hr, w_rcv;
            if constexpr (FAST) { hr = BH(pb)[q]; w_rcv = BW(pb)[q]; }
            else { hr = pin ? BH(pb)[q] : Hl[off + bp + q]; w_rcv = pin ? BW(pb)[q] : Wl[off + bp + q]; }
            if (first) {
              double hwr;
              if constexpr (FAST) hwr = BHW(pb)[q];
              else hwr = pin ? BHW(pb)[q] : HWl[off + bp + q];
              hw = hwr > t ? hwr + 1.0e-8 : t;
            }
            if (fill_only) {
              wv = w_rcv;
            } else {
              double e;
              if constexpr (FAST) e = t;
              else e = gzero ? t : ELl[off + b0 + i];
              if (e > w_rcv) {
                const double fact = rr.f;
                if (P.run_marine && !((n_is_one ? rr.h : t) >= P.sealevel)) {
                  // below sea level with Marine on: h keeps its current value (:305,342)
                } else if (n_is_one) {
                  double f = e, df = 1.0;
                  f = f + fact * hr;
                  df = df + fact;
                  hv = f / df;
                } else {
                  const double omega = 0.875 / n, tolp = P.newton_tol;
                  double errp = 2.0 * tolp;
                  const double h0 = e;
                  const bool act = fact >= 0.0;
                  int it = 0;
                  while (errp > tolp) {
                    double f = hv - h0, df = 1.0;
                    if (act) {
                      const double x = fmax(0.0, hv - hr);
                      double pn, pn1;
                      npow(x, n, nkind, pn, pn1);
                      f = f + fact * pn;
                      df = df + fact * n * pn1;
                    }
                    const double hn = hv - f / df;
                    errp = fabs(hn - hv);
                    hv = hv * (1.0 - omega) + hn * omega;
                    if (++it > 100000) { atomicOr(err + inst, FSX_ERR_NEWTON); break; }
                  }
                  my_newton += it;
                }
                if (!gzero) lakec[off + v] = -1.0;
                if (hv < w_rcv) hv = w_rcv;
              } else {
                hv = e;
                if constexpr (FAST) sl = -1;
                else sl = gzero ? -1 : (pin ? BLS(pb)[q] : LSl[off + bp + q]);
                if (!gzero) lakec[off + v] = sl >= 0 ? w_rcv - hv : -1.0;  // its share of the sill's volume
              }
              wv = w_rcv > hv ? w_rcv : hv;
            }
          }
          if (w <= vcap) {
            BH(cb)[i] = hv;
            BW(cb)[i] = wv;
            if (first) BHW(cb)[i] = hw;
            BLS(cb)[i] = sl;
          }
          // level-ordered copies only where a later reader needs them: wide levels (kept out of LDS)
          // and the Gauss-Seidel iterations; grid-ordered outputs h, hwater, lake_depth always
          const long long gp = off + b0 + i;
          const bool wide = w > vcap;
          // results stay in level order (coalesced stores); k_spl_post scatters h, hwater and
          // lake_depth to grid order
          if (!fill_only) {
            Hl[gp] = hv;
            if (wide) Wl[gp] = wv;
            if (wide || !gzero) LSl[gp] = sl;
            if (!gzero) { water[off + v] = wv; lsill[off + v] = sl; }
          } else if (wide) {
            Wl[gp] = wv;
          }
          if (first) HWl[gp] = hw;
        };
        if (staged) {
          SplRec x1;
          if (tid + FSX_WG < w) stg(x1, b0 - cur.p0 + tid + FSX_WG);
          if (pin && gzero) {
            if (tid < w) one(tid, x0, std::true_type{});
            if (tid + FSX_WG < w) one(tid + FSX_WG, x1, std::true_type{});
          } else {
            if (tid < w) one(tid, x0, std::false_type{});
            if (tid + FSX_WG < w) one(tid + FSX_WG, x1, std::false_type{});
          }
        } else {
          for (int i = tid; i < w; i += FSX_WG) {
            SplRec rr;
            spl_load(rr, off + b0 + i, srec, Hl, first);
            one(i, rr, std::false_type{});
          }
        }
        if (w > vcap) __syncthreads(); else lds_barrier();
      };
      if (staged) {
        // software-pipelined over the group's levels: the bounds of level l+1 and its records are
        // read from LDS while level l is solved, so a level's critical path is its receiver reads
        // level starts of the group: lane k of every wave holds that of level lo+k (k < 64), so a
        // level's bounds are a register read (readlane) rather than an LDS round trip
        const int lsv = (tid & 63) <= cur.hi - cur.lo ? sL[tid & 63] : 0;
        auto st = [&](int k) { return k < 64 ? __builtin_amdgcn_readlane(lsv, k) : sL[k]; };
        int l = cur.lo;
        int A = st(0), B = st(1), Bp = cur.pm;                  // level l = [A, B), level l-1 from Bp
        int Nn = l + 2 <= cur.hi ? st(l + 2 - cur.lo) : 0;     // level l+1 = [B, Nn)
        SplRec x0;
        if (tid < B - A) stg(x0, A - cur.p0 + tid);
        for (; l < cur.hi; l++) {
          const int Nn2 = l + 3 <= cur.hi ? st(l + 3 - cur.lo) : 0;
          SplRec y0;
          if (l + 1 < cur.hi && tid < Nn - B) stg(y0, B - cur.p0 + tid);
          level(l, A, B - A, Bp, x0);
          Bp = A; A = B; B = Nn; Nn = Nn2;
          x0 = y0;
        }
      } else {
        SplRec none;
        level(cur.lo, cur.p0, cur.p1 - cur.p0, cur.pm, none);
      }
      cur = nxt_m;
      nxt_m = meta_bcast(after);
    }
    __syncthreads();
    if (gzero || fill_only) {
      if (tid == 0) s_err = 0.0;
    } else {
      double ss = 0.0;
      for (int p = tid; p < nn; p += FSX_WG) {
        const double dd = Hl[off + p] - HPl[off + p];
        ss += dd * dd;
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
#undef BH
#undef BW
#undef BHW
#undef BLS
  if (fill_only) return;
  if (my_newton) atomicAdd(&s_newton, my_newton);
  __syncthreads();
  if (tid == 0) {
    ngs_out[inst] = ngs;
    newton_out[inst] = (long long)s_newton;
    scal[inst] = gzero ? 0.0 : 1.0;
  }
}

// ---- sweeps without levels: dataflow over positions ------------------------------------------------
// The level sweeps pay one workgroup barrier per level, and every level waits for its slowest node (for
// the stream-power law, the longest Newton iteration).  The two kernels below instead solve each node as
// soon as the node(s) it depends on are done, so the critical path is the longest receiver chain's own
// work, not the sum of per-level maxima.  One workgroup per landscape:
//
// * Positions (level order: every receiver before its donors) are cut into tiles of 64; waves take
//   tiles in order from a workgroup counter (the stream-power pass from the first position up, the
//   drainage area from the last position down), keep the current and the next tile's records in
//   registers (one position per lane), and hand positions to their lanes one by one: a free lane takes
//   the next position of its wave's pool (records fetched by cross-lane shuffles), waits for the nodes it
//   depends on, solves, publishes and takes another.  Lanes of one wave thus never idle behind a slow
//   lane (a Newton iteration of tens of passes, a chain inside the tile).
// * A finished node publishes its values to global memory (set to a sentinel by the prep kernel) and to
//   an LDS ring (slot = position mod R) under a sequence tag: the writer claims the slot by swapping its
//   tag to -2 (compare-and-swap: one writer at a time), writes the values, then the position (release);
//   a reader accepts the slot only when it reads its position as the tag both before and after the
//   values.  A reader that finds a later position there reads global
//   memory (the value left the ring); one that finds an earlier one (not published yet, as a rule) polls
//   the ring again and global memory every FLOW_GLOBAL-th pass (slots are written out of order now and
//   then; global memory always ends up holding the value).
// * Progress: positions are issued in order within a wave, so a lane holding a position of a later tile
//   implies all earlier positions of its pool were issued.  The lowest pending position of the landscape
//   is therefore held by some lane (or its wave has only free lanes and issues it next pass), and what it
//   depends on is published.  All waves are resident on one compute unit (one workgroup), whose L1 keeps
//   their global stores and loads coherent (workgroup-scope atomics).  A wave that sees no progress for
//   FLOW_SPIN passes reports FSX_ERR_LEVELS and the workgroup stops; with a valid level order that cannot
//   happen.
// * A wave whose last pass made progress runs at a raised issue priority (s_setprio 2), an idle one at 0:
//   on a compute unit crowded with polling waves, the waves that advance the front get the issue slots.
#define FLOW_SPIN (1 << 26)
#define FLOW_GLOBAL 32

// LDS accesses (the address space is inferred after inlining)
__device__ __forceinline__ double lds_ld(double* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ void lds_st(double* p, double v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ int lds_acq(int* p) {
  return __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ int lds_rlx(int* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ void lds_rel(int* p, int v) {
  __hip_atomic_store(p, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ void lds_stw(int* p, int v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ void wg_fence_rel() { __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup"); }
__device__ __forceinline__ void wg_fence_acq() { __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup"); }

// Ring of NV doubles per slot and a tag; writes and reads under the sequence tag (see above).
template <int R, int NV>
struct FlowRing {
  double* v;  // NV * R, value k of slot s at v[k * R + s]
  int* tag;   // R
  // Claims the slot (tag -> -2 by compare-and-swap, so two writers never mix their values), writes the
  // values, then the tag.  A slot being written by another lane is left to it (readers then find p's
  // value in global memory).
  __device__ __forceinline__ void put(int p, const double* x) {
    const int s = p & (R - 1);
    const int old = lds_rlx(tag + s);
    if (old == -2 || atomicCAS(tag + s, old, -2) != old) return;
#pragma unroll
    for (int k = 0; k < NV; k++) lds_st(v + k * R + s, x[k]);
    lds_rel(tag + s, p);
  }
  __device__ __forceinline__ bool get(int q, double* x, int& tg) {  // tg: the slot's tag
    const int s = q & (R - 1);
    tg = lds_acq(tag + s);
    if (tg != q) return false;
#pragma unroll
    for (int k = 0; k < NV; k++) x[k] = lds_ld(v + k * R + s);
    wg_fence_acq();
    tg = lds_rlx(tag + s);
    return tg == q;
  }
};

struct FlowWin {
  int next, abort;
};
__device__ __forceinline__ void flow_win_init(FlowWin& W) {
  if (threadIdx.x == 0) {
    W.next = 0;
    W.abort = 0;
  }
}
__device__ __forceinline__ int flow_take(FlowWin& W) {  // wave-uniform take rank
  int k = 0;
  if ((threadIdx.x & 63) == 0) k = atomicAdd(&W.next, 1);
  return __builtin_amdgcn_readfirstlane(k);
}
__device__ __forceinline__ bool flow_aborted(FlowWin& W) {
  return __hip_atomic_load(&W.abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) != 0;
}
// one pass without progress: false (and the workgroup aborted) past the spin budget, or (multi) once
// another workgroup of the landscape has reported
__device__ __forceinline__ bool flow_idle(FlowWin& W, int& spin, int* err, bool multi = false) {
  const bool other = multi && (spin & 1023) == 1023 &&
                     (__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) & FSX_ERR_LEVELS);
  if (++spin > FLOW_SPIN || other || flow_aborted(W)) {
    if ((threadIdx.x & 63) == 0) {
      __hip_atomic_store(&W.abort, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
      atomicOr(err, FSX_ERR_LEVELS);
    }
    return false;
  }
  __builtin_amdgcn_s_sleep(1);
  return true;
}

__device__ __forceinline__ double shfl_d(double x, int src) {
  const long long v = __double_as_longlong(x);
  const int lo = __shfl((int)(unsigned)(v & 0xffffffffll), src, 64);
  const int hi = __shfl((int)(v >> 32), src, 64);
  return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

// ---- top-down: the single stream-power pass of g == 0 ----------------------------------------------
// With g == 0 everywhere the Gauss-Seidel loop runs once (src/StreamPowerLaw.f90:375-376) and a node's
// solve needs only its receiver's h, water and lake surface (:296-366, the fill :73-89).  A Newton lane
// advances nk iterations per pass of its wave's loop (a launch shape: few keep the passes of lanes that poll
// short, more let a solve finish in the pass that found its receiver; the iterates are the same).
#define SPL_R 4096  // ring slots

struct FlowRec {
  int o, q;  // node, absolute receiver position (-1: base level)
  double f, t;
};
__device__ __forceinline__ void flow_load(FlowRec& r, const SplRec32* srec, long long gp) {
  const ulonglong2* w = reinterpret_cast<const ulonglong2*>(srec + gp);
  const ulonglong2 a = w[0], b = w[1];
  r.o = (int)(unsigned)(a.x & 0xffffffffull);
  r.q = (int)(unsigned)(a.x >> 32);
  r.f = __longlong_as_double((long long)a.y);
  r.t = __longlong_as_double((long long)b.x);
}

// tiles of one landscape that a workgroup can mark as its own (one bit each): landscapes with more tiles
// run on one workgroup
#define FLOW_OWN_WORDS 4096
static size_t spl_flow_smem() {
  return (size_t)SPL_R * (sizeof(int) + 3 * sizeof(double)) + FLOW_OWN_WORDS * sizeof(unsigned);
}

// WG threads per landscape (the autotuner's flow_wg): fewer waves per SIMD shorten a pass of the wave
// loop, and with it each hop of the critical receiver chain, more waves solve more nodes at once.
// mode: 0 the step's solve, 1 lake fill only (no erosion), 2 the solve without the per-step outputs
// (ngs, Newton count, scal; the autotuner's timing passes).
//
// A large landscape runs on P.flow_g workgroups (other compute units; ftab lists (landscape, rank) per
// workgroup): waves of all of them take tiles in order from the landscape's counter in global memory
// (fnext), each workgroup marks the tiles its waves took in an LDS bitmap, and publishes its nodes' values
// to global memory with device-scope stores.  A lane whose receiver lies in a tile of its own workgroup
// reads the ring as above; one whose receiver's tile another workgroup took polls global memory every
// pass.  The progress argument holds across workgroups: tiles are taken in order, only by running waves.
// Newton iteration counts are summed into newton_out (zeroed before the launch).
template <int WG>
__global__ __launch_bounds__(WG) void k_spl_flow(const InstParams* __restrict__ PP, const int2* __restrict__ ftab,
                                                 int* __restrict__ fnext, const int* __restrict__ nlev,
                                                 const SplRec32* __restrict__ srec, double* __restrict__ Hl,
                                                 double* __restrict__ Wl, double* __restrict__ HWl,
                                                 double* __restrict__ scal, int* __restrict__ ngs_out,
                                                 long long* __restrict__ newton_out, int* __restrict__ err, int mode,
                                                 int nk) {
  extern __shared__ double spl_ring[];  // h[R], water[R], lake surface[R], tag[R], own-tile bitmap
  __shared__ FlowWin W;
  __shared__ unsigned long long s_newton;
  FlowRing<SPL_R, 3> ring{spl_ring, reinterpret_cast<int*>(spl_ring + 3 * SPL_R)};
  unsigned* own = reinterpret_cast<unsigned*>(spl_ring + 3 * SPL_R) + SPL_R;
  const int2 wk = ftab[chunk_id()];
  const int inst = wk.x, rank = wk.y;
  const InstParams& P = PP[inst];
  if (!P.active || !P.run_spl || !P.sfd || !P.flow) return;
  if (nlev[inst] == 0) return;  // no level order this step (receiver cycle, already reported)
  const bool multi = P.flow_g > 1;
  const int tid = threadIdx.x, lane = tid & 63;
  const int nn = P.nn;
  const long long off = P.off;
  const int ntiles = (nn + 63) / 64;
  const double n = P.n;
  const bool n_is_one = fabs(n - 1.0) < 2.2250738585072014e-308;
  const int nkind = npow_kind(n);
  const double omega = 0.875 / n, tolp = P.newton_tol;
  flow_win_init(W);
  const bool fill_only = mode == 1;
  // with Marine on, the stream-power law leaves nodes below sea level alone (src/StreamPowerLaw.f90:305,342:
  // h for n == 1, ht otherwise; the same here, where the single pass starts from h = ht)
  const bool sea_skip = P.run_marine != 0;
  for (int i = tid; i < SPL_R; i += WG) ring.tag[i] = -1;
  if (multi)
    for (int i = tid; i < FLOW_OWN_WORDS; i += WG) own[i] = 0u;
  if (tid == 0) s_newton = 0;
  __syncthreads();
  // next tile of the landscape for this wave (wave-uniform)
  auto take = [&]() __attribute__((always_inline)) {
    if (!multi) return flow_take(W);
    int k = 0;
    if (lane == 0) {
      k = atomicAdd(fnext + inst, 1);
      if (k < ntiles) atomicOr(own + (k >> 5), 1u << (k & 31));
    }
    return __builtin_amdgcn_readfirstlane(k);
  };
  auto tsize = [&](int t) { return min(64, nn - 64 * t); };
  // record of position 64 t + lane, read unconditionally (clamped into the landscape; positions past the
  // end are never issued) so that the loads stay in flight until the bank is used
  auto load = [&](int t, int& q, double& f, double& u) {
    const int pp = min(64 * min(t, ntiles - 1) + lane, nn - 1);
    FlowRec r;
    flow_load(r, srec, off + pp);
    q = r.q;
    f = r.f;
    u = r.t;
  };
  // lane state: 0 free, 1 waiting for the receiver, 2 Newton
  int ph = 0, p = 0, q = -1, it = 0, polls = 0, iss = 0, spin = 0, npass = 0;
  bool gpass = false;
  double f = 0.0, tt = 0.0, hr = 0.0, wr = 0.0, hw = 0.0, hv = 0.0;
  unsigned long long my_newton = 0;
#define SPL_PUBLISH(HV, WV, HW)                \
  do {                                         \
    const double px_[3] = {(HV), (WV), (HW)};  \
    ring.put(p, px_);                          \
    if (multi) {                               \
      flow_st_dev(Hl + off + p, px_[0]);       \
      flow_st_dev(Wl + off + p, px_[1]);       \
      flow_st_dev(HWl + off + p, px_[2]);      \
    } else {                                   \
      flow_st(Hl + off + p, px_[0]);           \
      flow_st(Wl + off + p, px_[1]);           \
      flow_st(HWl + off + p, px_[2]);          \
    }                                          \
    ph = 0;                                    \
  } while (0)
  // One pass of the wave with tile TC (records BQ, BF, BU, one per lane) as the current tile: issue its
  // next positions to free lanes, then poll, solve and publish.  ST becomes 1 when the tile is used up (the
  // caller refills the bank), 2 when the wave is done (or the workgroup aborted), 0 otherwise.
#define SPL_PASS(TC, BQ, BF, BU, ST)                                                       \
  do {                                                                                     \
    int used_ = 0;                                                                         \
    bool prog_ = false;                                                                    \
    gpass = (++npass & (FLOW_GLOBAL - 1)) == 0;                                            \
    if ((TC) < ntiles) {                                                                   \
      const unsigned long long fm = __ballot(ph == 0);                                    \
      const int nfree = __popcll(fm);                                                      \
      const int rank = __popcll(fm & ((1ull << lane) - 1ull));                             \
      const int avail = tsize(TC) - iss;                                                   \
      const int j = iss + (rank < avail ? rank : 0);                                       \
      const int sq = __shfl((BQ), j, 64);                                                  \
      const double sf = shfl_d((BF), j), su = shfl_d((BU), j);                             \
      if (ph == 0 && rank < avail) {                                                       \
        ph = 1;                                                                            \
        p = 64 * (TC) + j;                                                                 \
        q = sq;                                                                            \
        f = sf;                                                                            \
        tt = su;                                                                           \
        polls = 0;                                                                         \
      }                                                                                    \
      const int took = nfree < avail ? nfree : avail;                                      \
      if (took) prog_ = true;                                                              \
      iss += took;                                                                         \
      if (iss == tsize(TC)) {                                                              \
        iss = 0;                                                                           \
        used_ = 1;                                                                         \
      }                                                                                    \
    }                                                                                      \
    spl_work(prog_);                                                                       \
    if ((TC) >= ntiles && __ballot(ph != 0) == 0ull) {                                     \
      ST = 2;                                                                              \
    } else if (__ballot(prog_) != 0ull) {                                                  \
      spin = 0;                                                                            \
      ST = used_;                                                                          \
      __builtin_amdgcn_s_setprio(2);                                                       \
    } else {                                                                               \
      __builtin_amdgcn_s_setprio(0);                                                       \
      ST = flow_idle(W, spin, err + inst, multi) ? used_ : 2;                              \
    }                                                                                      \
  } while (0)
  // poll, solve and publish (the part of a pass after the issue)
  auto spl_work = [&](bool& prog) __attribute__((always_inline)) {
    // waiting lanes poll their receiver
    if (ph == 1) {
      bool ok = true;
      if (q >= 0) {
        double x[3];
        const int tq = q >> 6;
        if (!multi || ((own[tq >> 5] >> (tq & 31)) & 1u)) {
          int tg;
          ok = ring.get(q, x, tg);
          // global memory: once when the slot holds a later position (q's value, as a rule, left the
          // ring), and on every FLOW_GLOBAL-th pass of the wave (all lanes at once, one wait per wave)
          if (!ok && ((tg > q && polls == 0) || gpass)) {
            polls = 1;
            if (multi) {
              x[0] = flow_ld_dev(Hl + off + q);
              x[1] = flow_ld_dev(Wl + off + q);
              x[2] = flow_ld_dev(HWl + off + q);
            } else {
              x[0] = flow_ld(Hl + off + q);
              x[1] = flow_ld(Wl + off + q);
              x[2] = flow_ld(HWl + off + q);
            }
            ok = !flow_sent(x[0]) && !flow_sent(x[1]) && !flow_sent(x[2]);
          }
        } else {  // another workgroup's tile (or one whose mark is not visible yet): global memory
          x[0] = flow_ld_dev(Hl + off + q);
          x[1] = flow_ld_dev(Wl + off + q);
          x[2] = flow_ld_dev(HWl + off + q);
          ok = !flow_sent(x[0]) && !flow_sent(x[1]) && !flow_sent(x[2]);
        }
        hr = x[0];
        wr = x[1];
        hw = x[2];
      }
      if (ok) {
        prog = true;
        if (q < 0) {
          SPL_PUBLISH(tt, tt, tt);
        } else {
          const double hwv = hw > tt ? hw + 1.0e-8 : tt;
          hw = hwv;
          const double e = tt;
          if (fill_only) {
            SPL_PUBLISH(tt, wr, hwv);
          } else if (!(e > wr)) {
            SPL_PUBLISH(e, wr > e ? wr : e, hwv);
          } else if (sea_skip && !(tt >= P.sealevel)) {  // below sea level with Marine on: h stays ht
            const double h1 = tt < wr ? wr : tt;
            SPL_PUBLISH(h1, wr > h1 ? wr : h1, hwv);
          } else if (n_is_one) {
            double ff = e, df = 1.0;
            ff = ff + f * hr;
            df = df + f;
            double h1 = ff / df;
            if (h1 < wr) h1 = wr;
            SPL_PUBLISH(h1, wr > h1 ? wr : h1, hwv);
          } else {
            ph = 2;
            hv = tt;
            it = 0;
          }
        }
      }
    }
    // Newton lanes: up to nk iterations, then publish when converged (errp <= tolp)
    if (ph == 2) {
      prog = true;
      const double h0 = tt;
      const bool act = f >= 0.0;
      bool conv = false;
      for (int k = 0; k < nk && !conv; k++) {
        double ff = hv - h0, df = 1.0;
        if (act) {
          const double x = fmax(0.0, hv - hr);
          double pn, pn1;
          npow(x, n, nkind, pn, pn1);
          ff = ff + f * pn;
          df = df + f * n * pn1;
        }
        const double hn = hv - ff / df;
        const double errp = fabs(hn - hv);
        hv = hv * (1.0 - omega) + hn * omega;
        conv = !(errp > tolp);
        if (++it > 100000) {
          atomicOr(err + inst, FSX_ERR_NEWTON);
          conv = true;
        }
      }
      if (conv) {
        my_newton += it;
        double h1 = hv;
        if (h1 < wr) h1 = wr;
        SPL_PUBLISH(h1, wr > h1 ? wr : h1, hw);
      }
    }
  };
  // two banks of records, the current tile's and the next one's, used in turn; a bank is refilled right
  // after its tile is used up, so its loads are waited for only when it is current again
  int ta = take(), tb = take();
  int qa, qb;
  double fa, fb, ua, ub;
  load(ta, qa, fa, ua);
  load(tb, qb, fb, ub);
  while (true) {
    int st = 0;
    while (st == 0) SPL_PASS(ta, qa, fa, ua, st);
    if (st == 2) break;
    ta = take();
    load(ta, qa, fa, ua);
    st = 0;
    while (st == 0) SPL_PASS(tb, qb, fb, ub, st);
    if (st == 2) break;
    tb = take();
    load(tb, qb, fb, ub);
  }
#undef SPL_PASS
#undef SPL_PUBLISH
  if (my_newton) atomicAdd(&s_newton, my_newton);
  __syncthreads();
  if (tid == 0 && mode == 0) {
    if (s_newton) atomicAdd(reinterpret_cast<unsigned long long*>(newton_out + inst), s_newton);
    if (rank == 0) {
      ngs_out[inst] = 1;
      scal[inst] = 0.0;  // g == 0: no lake sediment to spread (k_spl_post)
    }
  }
}

// ---------------------------------------------------------------------------------------------------
static size_t area_smem(int vcap) {
  return 2 * vcap * sizeof(double) + GRP_STG * (sizeof(double) + 8 * sizeof(int)) + (GRP_BUCKET + 1) * sizeof(int);
}
static size_t spl_smem(int vcap) {
  return (size_t)vcap * (6 * sizeof(double) + 2 * sizeof(int)) + GRP_STG * (3 * sizeof(double) + 2 * sizeof(int)) +
         (GRP_BUCKET + 1) * sizeof(int);
}
static const int kAreaVcap = 2048;
static const int kSplVcap = 1536;

int sweep_set_lds(int smem_max) {
  if ((int)area_smem(kAreaVcap) > smem_max || (int)spl_smem(kSplVcap) + 4096 > smem_max ||
      LVL_HCAP * 4 > smem_max)
    return -1;
  for (const void* f : {(const void*)k_area_sweep<256>, (const void*)k_area_sweep<512>, (const void*)k_area_sweep<1024>})
    if (hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)area_smem(kAreaVcap)) != hipSuccess) return -1;
  if (hipFuncSetAttribute((const void*)k_spl_sweep, hipFuncAttributeMaxDynamicSharedMemorySize, (int)spl_smem(kSplVcap)) != hipSuccess) return -1;
  if ((int)(spl_flow_smem() + sizeof(FlowWin) + 64) > smem_max) return -1;
  for (const void* f : {(const void*)k_spl_flow<256>, (const void*)k_spl_flow<512>, (const void*)k_spl_flow<1024>})
    if (hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)spl_flow_smem()) != hipSuccess) return -1;
  if (hipFuncSetAttribute((const void*)k_level_sort, hipFuncAttributeMaxDynamicSharedMemorySize, LVL_HCAP * 4) != hipSuccess) return -1;
  return 0;
}


void launch_jump(DevState& S, hipStream_t st) {
  k_jump_init<<<S.nchunks, FSX_CHUNK, 0, st>>>(S.P, S.chunks, S.rec, S.J);
  launch_fill_int(S.flags, 64, 0, st);
  const int nb = std::min(S.nchunks, S.jump_blocks);
  for (int it = 0; it < FSX_JUMP_ROUNDS; it++) k_jump<<<nb, FSX_CHUNK, 0, st>>>(S.P, S.chunks, S.nchunks, S.J, S.flags, it);
}

void launch_levels(DevState& S, hipStream_t st) {
  launch_jump(S, st);
  k_level_sort<<<S.ninst, FSX_WG, LVL_HCAP * 4, st>>>(S.P, S.J, S.flags, S.lstart, S.lcur, S.nlev, S.gmeta, S.ngroups,
                                                        S.grec, S.order, S.pos, S.err);
  if (!S.nlvl) return;
  k_lvl_zero<<<S.nlvl, FSX_WG, 0, st>>>(S.P, S.lvl_tab, S.lcur, S.nlev);
  k_lvl_count<<<S.nlvl, FSX_WG, 0, st>>>(S.P, S.lvl_tab, S.J, S.flags, S.lcur, S.nlev);
  k_lvl_scan<<<S.nlvl_big, FSX_WG, 0, st>>>(S.P, S.lvl_tab + S.nlvl, S.flags, S.lstart, S.lcur, S.nlev, S.gmeta,
                                             S.ngroups, S.grec, S.err);
  k_lvl_place<<<S.nlvl, FSX_WG, 0, st>>>(S.P, S.lvl_tab, S.J, S.flags, S.lcur, S.order, S.pos);
}

void launch_area_prep(DevState& S, hipStream_t st) {
  k_area_prep<<<S.nchunks, FSX_CHUNK, 0, st>>>(S.P, S.chunks, S.J, S.dmask, S.pos, S.lstart, S.precip, S.flags,
                                               S.arec);
}

void launch_area(DevState& S, hipStream_t st) {
  launch_area_prep(S, st);
  launch_area_sweep(S, st);
}

void launch_area_sweep(DevState& S, hipStream_t st) {
  const size_t sm = area_smem(kAreaVcap);
  if (S.area_wg == 256)
    k_area_sweep<256><<<S.ninst, 256, sm, st>>>(S.P, S.lstart, S.grec, S.ngroups, S.arec, S.Al, kAreaVcap);
  else if (S.area_wg == 512)
    k_area_sweep<512><<<S.ninst, 512, sm, st>>>(S.P, S.lstart, S.grec, S.ngroups, S.arec, S.Al, kAreaVcap);
  else
    k_area_sweep<1024><<<S.ninst, 1024, sm, st>>>(S.P, S.lstart, S.grec, S.ngroups, S.arec, S.Al, kAreaVcap);
}

void launch_spl_fact(DevState& S, hipStream_t st) {
  launch_fill_int(reinterpret_cast<int*>(S.redk), 4LL * S.ninst, 0, st);
  k_spl_fact<<<S.nchunks, FSX_CHUNK, 0, st>>>(S.P, S.chunks, S.J, S.rec, S.pos, S.lstart, S.Al, S.area, S.length,
                                               S.kfint, S.ht, S.srec, S.lsed, S.g, S.redk, S.flags, S.Hl, S.Wl, S.HWl);
}

void launch_spl_flow(DevState& S, hipStream_t st, int mode) {
  if (!S.nflow) return;
  const size_t sm = spl_flow_smem();
  launch_fill_int(S.flow_next, S.ninst, 0, st);
  if (mode == 0) launch_fill_int(reinterpret_cast<int*>(S.newton), 2LL * S.ninst, 0, st);  // summed per workgroup
  const int g = S.nflow;
  if (S.flow_wg == 256)
    k_spl_flow<256><<<g, 256, sm, st>>>(S.P, S.flow_tab, S.flow_next, S.nlev, S.srec, S.Hl, S.Wl, S.HWl, S.scal, S.ngs,
                                        S.newton, S.err, mode, S.flow_nk);
  else if (S.flow_wg == 512)
    k_spl_flow<512><<<g, 512, sm, st>>>(S.P, S.flow_tab, S.flow_next, S.nlev, S.srec, S.Hl, S.Wl, S.HWl, S.scal, S.ngs,
                                        S.newton, S.err, mode, S.flow_nk);
  else
    k_spl_flow<1024><<<g, 1024, sm, st>>>(S.P, S.flow_tab, S.flow_next, S.nlev, S.srec, S.Hl, S.Wl, S.HWl, S.scal,
                                          S.ngs, S.newton, S.err, mode, S.flow_nk);
}

static void spl_sweep_launch(DevState& S, hipStream_t st, int fill_only) {
  // every landscape is taken by exactly one of the two kernels (InstParams::flow); the level sweep is not
  // launched when no landscape takes it (its workgroups would wait for compute units with its LDS free)
  launch_spl_flow(S, st, fill_only);
  if (S.nsweep)
    k_spl_sweep<<<S.ninst, FSX_WG, spl_smem(kSplVcap), st>>>(
      S.P, S.lstart, S.nlev, S.grec, S.ngroups, S.srec, S.arec, S.bc, S.g, S.Al, S.h, S.hwater,
      S.lake_depth, S.water, S.lsill, S.lakec, S.lsed, S.Hl, S.Wl, S.HWl, S.LSl, S.HPl, S.DHl, S.ELl, S.scal, S.ngs,
      S.newton, S.err, kSplVcap, fill_only, S.redk);
}

void launch_spl_sweep(DevState& S, hipStream_t st) { spl_sweep_launch(S, st, 0); }
void launch_spl(DevState& S, hipStream_t st) {
  launch_spl_fact(S, st);
  spl_sweep_launch(S, st, 0);
}
// lake surface and depth of a routing-only pass, back to grid order
__global__ __launch_bounds__(FSX_CHUNK) void k_fill_post(const InstParams* __restrict__ PP, const Chunk* __restrict__ chunks,
                                                         const int* __restrict__ pos, const double* __restrict__ HWl,
                                                         const double* __restrict__ ht, double* __restrict__ hwater,
                                                         double* __restrict__ lake_depth) {
  NODE_PROLOGUE
  if (!P.sfd) return;
  const double hw = HWl[(long long)P.off + pos[g_]];
  hwater[g_] = hw;
  lake_depth[g_] = hw - ht[g_];
}

void launch_fill(DevState& S, hipStream_t st) {
  launch_spl_fact(S, st);
  spl_sweep_launch(S, st, 1);
  k_fill_post<<<S.nchunks, FSX_CHUNK, 0, st>>>(S.P, S.chunks, S.pos, S.HWl, S.ht, S.hwater, S.lake_depth);
}
