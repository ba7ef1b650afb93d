// fsx_mfd.hip -- multiple-flow-direction routing, accumulation and stream power.
//
// The reference (FlowRouting, src/FlowRouting.f90:3-39; find_mult_rec :147-298; FlowAccumulation
// :97-119; StreamPowerLaw src/StreamPowerLaw.f90:3-206) first routes and reroutes single receivers
// exactly as the D8 path, fills the surface h0 along that forest, gives every node all strictly
// lower neighbours of h0 as receivers with slope**p weights, orders the nodes with a LIFO Kahn
// traversal (mstack) and walks that order serially.
//
// Here the nodes are ordered by the Kahn layers of the receiver graph in place of mstack: layer 0 holds
// the nodes without donors, layer l+1 the nodes whose last donor sits in layer l.  Every receiver of a
// node lies in a later layer, and so does its single-flow receiver (a "ghost" edge is added in the rare
// case where that receiver is not also a multiple-flow receiver).  Sums over donors are taken in a fixed
// order (neighbour direction, then receiver slot), so the results do not depend on the order of nodes
// inside a layer; they differ from the reference's mstack order only in the rounding of those sums.
//
// The reference's serial walks -- the fill, the Kahn traversal, the accumulation, the erosion push and the
// implicit solve -- run as dataflow sweeps on several workgroups per landscape (below): a position is
// computed as soon as the values it depends on are written, with no barrier per layer, so that a large
// landscape spreads over many compute units and its time follows its longest chain of dependencies.
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

// ---- dataflow sweeps on several workgroups per landscape -----------------------------------------------
// A landscape's sweeps run on P.mfd_g workgroups (the table mtab lists (landscape, rank); a landscape's
// entries are consecutive, so its workgroups share an XCD).  Work comes in tiles that waves take in order
// from a counter of the landscape (ctr[0]); a wave that leaves a phase adds the tiles it finished to
// ctr[1].  Tiles of the successive phases of a launch are numbered consecutively, and a wave starts a
// tile of a phase once ctr[1] shows every tile of the earlier phases finished.  Inside a dataflow phase a
// lane computes its position once the values it depends on -- earlier positions in the phase's order --
// differ from the sentinel.  Progress does not need the workgroups to be resident together: tiles are
// taken in order by running waves, and the lowest unfinished position depends only on finished ones.
// Values that one wave writes and another reads are device-scope loads and stores, and a wave drains its
// stores before it counts its tiles.  A wait of more than MFD_SPIN polls reports FSX_ERR_MFD (it cannot
// happen with a valid order) and the landscape's waves stop.
#define MFD_SPIN (1 << 22)
#define MFD_WG 512   // threads per workgroup of the dataflow kernels
#define MFD_MAXGS 100  // Gauss-Seidel decisions kept per landscape (the loop stops at 99)
#define MFD_EB 4096  // positions per partial sum of the Gauss-Seidel error (a fixed blocking of the sum)

struct MfdTiles {
  int* claim;  // ctr[0]: next tile to take
  int* done;   // ctr[1]: tiles finished
  int* err;    // the landscape's error word
  int base;    // first tile of the current phase
  int t;       // the tile this wave holds
  bool ok;     // false once the landscape reported a stall
};

__device__ __forceinline__ int mfd_claim(int* claim) {
  int k = 0;
  if ((threadIdx.x & 63) == 0) k = __hip_atomic_fetch_add(claim, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  return __builtin_amdgcn_readfirstlane(k);
}

// one more poll without progress: false past the budget (reported) or once the landscape reported a stall
__device__ __forceinline__ bool mfd_idle(int& spin, int* err) {
  if (++spin > MFD_SPIN) {
    if ((threadIdx.x & 63) == 0) atomicOr(err, FSX_ERR_MFD);
    return false;
  }
  if ((spin & 255) == 0 && (flow_ldi_dev(err) & FSX_ERR_MFD)) return false;
  __builtin_amdgcn_s_sleep(1);
  return true;
}

__device__ __forceinline__ void mfd_start(MfdTiles& T, int* ctr, int* err) {
  T.claim = ctr;
  T.done = ctr + 1;
  T.err = err;
  T.base = 0;
  T.ok = true;
  T.t = mfd_claim(T.claim);
}

// waits until `need` tiles are finished
__device__ __forceinline__ bool mfd_wait(MfdTiles& T, int need) {
  int spin = 0;
  while (T.ok && flow_ldi_dev(T.done) < need)
    if (!mfd_idle(spin, T.err)) T.ok = false;
  return T.ok;
}

// The next phase, of n tiles: this wave runs body(tile) on the tiles it takes.  A body that cannot finish
// sets T.ok = false.
template <class F>
__device__ __forceinline__ void mfd_phase(MfdTiles& T, int n, F&& body) {
  const int end = T.base + n;
  int nd = 0;
  if (T.ok && T.t < end && mfd_wait(T, T.base)) {
    while (T.t < end) {
      body(T.t - T.base);
      if (!T.ok) break;
      nd++;
      T.t = mfd_claim(T.claim);
    }
  }
  if (nd) {
    wave_drain();
    if ((threadIdx.x & 63) == 0) __hip_atomic_fetch_add(T.done, nd, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  T.base = end;
}

// Loop of a dataflow tile: step(prog) tries the lane's position (sets prog when it finished it) and
// returns whether the lane still has work; the wave polls until every lane is done, at a raised issue
// priority while it makes progress.
template <class F>
__device__ __forceinline__ void mfd_lanes(MfdTiles& T, F&& step) {
  int spin = 0;
  while (true) {
    bool prog = false;
    const bool more = step(prog);
    if (__ballot(more) == 0ull) return;
    if (__ballot(prog) != 0ull) {
      spin = 0;
      __builtin_amdgcn_s_setprio(2);
    } else {
      __builtin_amdgcn_s_setprio(0);
      if (!mfd_idle(spin, T.err)) {
        T.ok = false;
        return;
      }
    }
  }
}

#define MFD_WAVE_PROLOGUE                         \
  const int2 wk = mtab[chunk_id()];               \
  const int inst = wk.x;                          \
  const InstParams& P = PP[inst];                 \
  if (!P.active || !P.run_spl || P.sfd) return;   \
  const int nn = P.nn, lane = threadIdx.x & 63;   \
  const long long off = P.off;

// ---- set-up -----------------------------------------------------------------------------------------
// Kahn state (layer so far << 32 | in-edges left), empty queue slots, the layer histogram, the fill's
// sentinel.
__global__ __launch_bounds__(FSX_CHUNK) void k_mfd_init(const InstParams* __restrict__ PP, const Chunk* __restrict__ chunks,
                                                        unsigned long long* __restrict__ mkahn,
                                                        unsigned long long* __restrict__ mq, int* __restrict__ mls,
                                                        double* __restrict__ hwater) {
  MFD_PROLOGUE
  mkahn[g_] = 0ull;
  mq[g_] = ~0ull;
  hwater[g_] = flow_sentinel();
  mls[P.aux + v] = 0;
  if (v == 0) mls[P.aux + P.nn] = 0;
}

// ---- lake surface along the single-flow forest (find_mult_rec :187-196) ------------------------
// h0 = h, then along the stack h0(ijk) = h0(rec)+1e-8 where h0(rec) > h0(ijk); lake_depth = h0 - h
// (FlowRouting :34).  A dataflow sweep over the single-flow level order from base level up: a node waits
// for its receiver's h0.
__global__ __launch_bounds__(MFD_WG) void k_mfd_fill(const InstParams* __restrict__ PP, const int2* __restrict__ mtab,
                                                     int* __restrict__ ctr, int* __restrict__ err,
                                                     const int* __restrict__ nlev, const int* __restrict__ order,
                                                     const int* __restrict__ rec, const double* __restrict__ h,
                                                     double* __restrict__ hwater, double* __restrict__ lake_depth) {
  MFD_WAVE_PROLOGUE
  if (nlev[inst] == 0) return;  // no level order this step (receiver cycle, already reported)
  MfdTiles T;
  mfd_start(T, ctr + FSX_MFD_CTR * inst, err + inst);
  mfd_phase(T, (nn + 63) / 64, [&](int tt) {
    const int p = 64 * tt + lane;
    bool todo = p < nn;
    int v = 0, r = 0;
    double hv = 0.0;
    if (todo) {
      v = order[off + p];
      r = rec[off + v];
      hv = h[off + v];
    }
    mfd_lanes(T, [&](bool& prog) {
      if (todo) {
        double hw = hv;
        bool ready = true;
        if (r != v) {
          const double hr = flow_ld_dev(hwater + off + r);
          ready = !flow_sent(hr);
          if (hr > hv) hw = hr + 1.0e-8;
        }
        if (ready) {
          flow_st_dev(hwater + off + v, hw);
          lake_depth[off + v] = hw - hv;
          todo = false;
          prog = true;
        }
      }
      return todo;
    });
  });
}

// ---- receivers and weights (find_mult_rec :204-243) ---------------------------------------------
// Every strictly lower neighbour of h0 in scan order, wrap if cyclic else clamp (a clamped
// direction that lands on another node repeats that node, as the reference does).  Weights
// slope**pexp normalised by their sum; pexp = p, or 0.5 + 0.6*mean slope when p < 0 with the mean
// rounded through single precision and 0.6 a single-precision literal.  In-degrees (for the Kahn
// layers, the low half of mkahn) count every slot, plus the ghost edge to the single-flow receiver when
// needed.
__global__ __launch_bounds__(FSX_CHUNK) void k_mfd_rec(const InstParams* __restrict__ PP, const Chunk* __restrict__ chunks,
                                                       const double* __restrict__ h0, const int* __restrict__ rec,
                                                       int* __restrict__ mrec, double* __restrict__ mw,
                                                       uint8_t* __restrict__ mdir, int* __restrict__ mnrec,
                                                       uint8_t* __restrict__ mflag, unsigned long long* __restrict__ mkahn,
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
      atomicAdd(mkahn + P.off + t[k], 1ull);
    }
  }
  mnrec[g_] = nr;
  const bool ghost = !hit && r != v;
  mflag[g_] = ghost ? 1 : 0;
  if (ghost) atomicAdd(mkahn + P.off + r, 1ull);
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

// ---- Kahn layers (find_mult_rec :245-292) -------------------------------------------------------------
// The traversal as a dataflow over a queue per landscape: a node keeps (largest layer of its donors + 1)
// << 32 | in-edges left, updated by compare-and-swap; the donor that removes its last in-edge appends it
// to the queue with its layer and counts it in that layer's histogram.  Phase 0 queues the nodes without
// in-edges (layer 0); in phase 1 a lane takes queue position p, waits until it is written and releases
// the node's receivers.  The layer sets are the reference's Kahn layers; the order inside a layer varies
// from run to run.
// Kahn state after one in-edge of a node of layer lay is taken: (max(layer, lay + 1) << 32) | (in-edges - 1)
__device__ __forceinline__ unsigned long long kahn_take(unsigned long long w, unsigned lay) {
  const unsigned hl = (unsigned)(w >> 32), want = lay + 1u;
  return ((unsigned long long)(hl > want ? hl : want) << 32) | (unsigned long long)((unsigned)(w & 0xffffffffull) - 1u);
}

__global__ __launch_bounds__(MFD_WG) void k_mfd_kahn(const InstParams* __restrict__ PP, const int2* __restrict__ mtab,
                                                     int* __restrict__ ctr, int* __restrict__ err,
                                                     unsigned long long* __restrict__ mkahn, const int* __restrict__ mrec,
                                                     const int* __restrict__ mnrec, const uint8_t* __restrict__ mflag,
                                                     const int* __restrict__ rec, unsigned long long* __restrict__ mq,
                                                     int* __restrict__ mls) {
  MFD_WAVE_PROLOGUE
  int* cq = ctr + FSX_MFD_CTR * inst + 4;  // claim, done, queue tail, largest layer
  int* hist = mls + P.aux;
  MfdTiles T;
  mfd_start(T, cq, err + inst);
  // Appending to the queue: a wave reserves the positions of all the nodes its lanes release in one step with
  // one atomic on the tail (cq[2]), counts them in the layer histogram with one atomic per distinct layer, and
  // keeps the largest layer it pushed in a register (one maximum per wave at the end): the tail, the histogram
  // and the maximum are the contended words of the traversal.
  unsigned mymax = 0;
  auto reserve = [&](int cnt) {  // the first queue position of this lane's cnt entries
    const int incl = wave_incl_scan(cnt);
    const int total = __shfl(incl, 63, 64);
    int base = 0;
    if (total) {
      if (lane == 63) base = __hip_atomic_fetch_add(cq + 2, total, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      base = __shfl(base, 63, 64);
    }
    return base + incl - cnt;
  };
  auto enqueue = [&](int& idx, int u, unsigned lay) {
    if (idx < nn) flow_stu_dev(mq + off + idx, ((unsigned long long)lay << 32) | (unsigned long long)(unsigned)u);
    idx++;
    mymax = lay > mymax ? lay : mymax;
  };
  auto hist_add = [&](bool pushme, unsigned lay) {
    unsigned long long rem = __ballot(pushme);
    while (rem) {
      const int ld = __ffsll((long long)rem) - 1;
      const unsigned l0 = (unsigned)__shfl((int)lay, ld, 64);
      const unsigned long long same = __ballot(pushme && lay == l0) & rem;
      if (lane == ld) atomicAdd(hist + l0, __popcll(same));
      rem &= ~same;
    }
  };
  // the sources (no in-edges), layer 0
  mfd_phase(T, (nn + 255) / 256, [&](int tt) {
    bool z[4];
    int cnt = 0;
#pragma unroll
    for (int k = 0; k < 4; k++) {
      const int u = 256 * tt + 64 * k + lane;
      z[k] = u < nn && (unsigned)mkahn[off + (u < nn ? u : 0)] == 0u;
      cnt += z[k] ? 1 : 0;
    }
    int idx = reserve(cnt);
#pragma unroll
    for (int k = 0; k < 4; k++)
      if (z[k]) enqueue(idx, 256 * tt + 64 * k + lane, 0u);
    const int all = __shfl(wave_incl_scan(cnt), 63, 64);
    if (lane == 0 && all) atomicAdd(hist, all);
  });
  mfd_phase(T, (nn + 63) / 64, [&](int tt) {
    const int p = 64 * tt + lane;
    bool todo = p < nn;
    mfd_lanes(T, [&](bool& prog) {
      bool have = false;
      int u = 0, nr = -1, nm = 0;
      unsigned lay = 0;
      if (todo) {
        const unsigned long long e = flow_ldu_dev(mq + off + p);
        if (e != ~0ull) {
          todo = false;
          prog = true;
          have = true;
          u = (int)(unsigned)(e & 0xffffffffull);
          lay = (unsigned)(e >> 32);
          nm = mnrec[off + u];
          nr = nm + (mflag[off + u] ? 1 : 0);  // the ghost edge to the single-flow receiver, last
        }
      }
      // Each receiver slot takes one in-edge of its node (layer at least lay + 1).  All slots of a lane go
      // together: their targets, then their words, then one compare-and-swap each; a slot whose word changed
      // since it was read retries from the value it found.  The slots whose node this lane released (its last
      // in-edge) are marked in relm, with the node's final layer.
      const long long gu = off + u;
      unsigned relm = 0;
      int tk[9];
      unsigned nlk[9];
      unsigned long long sk[9], ok[9];
#pragma unroll
      for (int k = 0; k < 9; k++) {
        tk[k] = 0;
        nlk[k] = 0;
        if (have && k < nr) tk[k] = k < nm ? mrec[8 * gu + k] : rec[gu];
      }
#pragma unroll
      for (int k = 0; k < 9; k++) sk[k] = (have && k < nr) ? flow_ldu_dev(mkahn + off + tk[k]) : 0ull;
#pragma unroll
      for (int k = 0; k < 9; k++) ok[k] = (have && k < nr) ? atomicCAS(mkahn + off + tk[k], sk[k], kahn_take(sk[k], lay)) : 0ull;
#pragma unroll
      for (int k = 0; k < 9; k++) {
        if (!(have && k < nr)) continue;
        unsigned long long seen = sk[k], old = ok[k];
        while (old != seen) {
          seen = old;
          old = atomicCAS(mkahn + off + tk[k], seen, kahn_take(seen, lay));
        }
        if ((unsigned)(seen & 0xffffffffull) == 1u) {
          relm |= 1u << k;
          nlk[k] = (unsigned)(kahn_take(seen, lay) >> 32);
        }
      }
      // the released nodes, with their final layers, appended in slot order
      int idx = reserve(__popc(relm));
#pragma unroll
      for (int k = 0; k < 9; k++) {
        const bool pushme = (relm >> k) & 1u;
        if (pushme) enqueue(idx, tk[k], nlk[k]);
        hist_add(pushme, nlk[k]);
      }
      return todo;
    });
  });
  for (int o = 32; o > 0; o >>= 1) {
    const unsigned y = (unsigned)__shfl_xor((int)mymax, o, 64);
    mymax = y > mymax ? y : mymax;
  }
  if (lane == 0 && mymax) atomicMax(cq + 3, (int)mymax);
}

// Layer starts from the histogram (exclusive scan, one workgroup per landscape) and the scatter cursors;
// an order that does not hold every node is the reference's 'error in stack' (src/FlowRouting.f90:292).
__global__ __launch_bounds__(FSX_WG) void k_mfd_lscan(const InstParams* __restrict__ PP, const int* __restrict__ ctr,
                                                      int* __restrict__ mls, int* __restrict__ mcur,
                                                      int* __restrict__ mnlev, int* __restrict__ err) {
  __shared__ int lds[FSX_WG / 64];
  const int inst = blockIdx.x;
  const InstParams& P = PP[inst];
  if (!P.active || !P.run_spl || P.sfd) return;
  const int* cq = ctr + FSX_MFD_CTR * inst + 4;
  const int queued = cq[2], L = cq[3] + 1;
  int* hist = mls + P.aux;
  int* cur = mcur + P.aux;
  int carry = 0;
  for (int l0 = 0; l0 < L; l0 += FSX_WG) {
    const int l = l0 + (int)threadIdx.x;
    const int x = l < L ? hist[l] : 0;
    int tot;
    const int incl = block_incl_scan<FSX_WG>(x, lds, tot);
    if (l < L) {
      hist[l] = carry + incl - x;
      cur[l] = carry + incl - x;
    }
    carry += tot;
  }
  if (threadIdx.x == 0) {
    hist[L] = carry;
    mnlev[inst] = L;
    if (carry != P.nn || queued != P.nn) atomicOr(err + inst, FSX_ERR_MFD);
  }
}

// queue position -> its layer's next slot of mstack (one cursor update per distinct layer of a wave: the
// queue is close to layer order, so a wave's 64 positions hold few layers)
__global__ __launch_bounds__(FSX_CHUNK) void k_mfd_scatter(const InstParams* __restrict__ PP, const Chunk* __restrict__ chunks,
                                                           const unsigned long long* __restrict__ mq,
                                                           int* __restrict__ mcur, int* __restrict__ morder) {
  const Chunk c = chunks[chunk_id()];
  const InstParams& P = PP[c.inst];
  if (!P.active || !P.run_spl || P.sfd) return;  // uniform over the block
  const int p = c.start + (int)threadIdx.x, lane = threadIdx.x & 63;
  const unsigned long long e = p < P.nn ? mq[(long long)P.off + p] : ~0ull;
  const bool ok = e != ~0ull;  // (an incomplete order is reported by k_mfd_lscan)
  const int u = (int)(unsigned)(e & 0xffffffffull), lay = ok ? (int)(e >> 32) : 0;
  unsigned long long rem = __ballot(ok);
  while (rem) {
    const int ld = __ffsll((long long)rem) - 1;
    const int l0 = __shfl(lay, ld, 64);
    const unsigned long long same = __ballot(ok && lay == l0) & rem;
    int base = 0;
    if (lane == ld) base = atomicAdd(mcur + P.aux + l0, __popcll(same));
    base = __shfl(base, ld, 64);
    if ((same >> lane) & 1ull) morder[P.off + base + __popcll(same & ((1ull << lane) - 1ull))] = u;
    rem &= ~same;
  }
}

// ---- per-position records --------------------------------------------------------------------
// The sweeps run over positions of the Kahn order (mstack): mpos is a node's position; every per-node
// input a sweep needs is copied to its position by node-parallel kernels, so a sweep's only dependent
// reads are its neighbours' values of earlier layers.
__global__ __launch_bounds__(FSX_CHUNK) void k_mfd_pos(const InstParams* __restrict__ PP, const Chunk* __restrict__ chunks,
                                                       const int* __restrict__ morder, int* __restrict__ mpos) {
  MFD_PROLOGUE
  mpos[P.off + morder[g_]] = v;  // here v runs over positions
}

// donors (position, weight) and receivers (positions) of each position; own area dx*dy*precip; the
// accumulation's output set to the sentinel
__global__ __launch_bounds__(FSX_CHUNK) void k_mfd_prep(const InstParams* __restrict__ PP, const Chunk* __restrict__ chunks,
                                                        const int* __restrict__ mpos, const int* __restrict__ rec,
                                                        const int* __restrict__ mrec, const int* __restrict__ mnrec,
                                                        const double* __restrict__ mw, const int* __restrict__ mdon,
                                                        const int* __restrict__ mndon, const double* __restrict__ precip,
                                                        double* __restrict__ A0, double* __restrict__ Al,
                                                        int* __restrict__ mdp,
                                                        double* __restrict__ mdw, int* __restrict__ dn,
                                                        int* __restrict__ rp) {
  MFD_PROLOGUE
  // every load is issued before any is used: the slots' indices, then the positions and weights they
  // point at (the records are written to scattered positions, so nothing waits on a store)
  const long long off = P.off;
  const int nd = mndon[g_], nr = mnrec[g_], r = rec[g_], mp = mpos[g_];
  const double pr = precip[g_];
  int e[8], q[8];
#pragma unroll
  for (int j = 0; j < 8; j++) {
    e[j] = j < nd ? mdon[8 * g_ + j] : 0;
    q[j] = j < nr ? mrec[8 * g_ + j] : 0;
  }
  int dpos[8], rpos[8];
  double w[8];
#pragma unroll
  for (int j = 0; j < 8; j++) {
    dpos[j] = j < nd ? mpos[off + (e[j] >> 3)] : -1;
    w[j] = j < nd ? mw[8 * off + e[j]] : 0.0;
    rpos[j] = j < nr ? mpos[off + q[j]] : -1;
  }
  const int rpv = r == v ? -1 : mpos[off + r];
  const long long gp = off + mp;
  A0[gp] = P.dxdy * pr;
  Al[gp] = flow_sentinel();  // the accumulation's dataflow waits for it to change
  int4* dp4 = reinterpret_cast<int4*>(mdp + 8 * gp);
  dp4[0] = make_int4(dpos[0], dpos[1], dpos[2], dpos[3]);
  dp4[1] = make_int4(dpos[4], dpos[5], dpos[6], dpos[7]);
  double2* dw2 = reinterpret_cast<double2*>(mdw + 8 * gp);
#pragma unroll
  for (int j = 0; j < 4; j++) dw2[j] = make_double2(w[2 * j], w[2 * j + 1]);
  int4* dn4 = reinterpret_cast<int4*>(dn + 8 * gp);
  dn4[0] = make_int4(rpos[0], rpos[1], rpos[2], rpos[3]);
  dn4[1] = make_int4(rpos[4], rpos[5], rpos[6], rpos[7]);
  rp[gp] = rpv;
}

// ---- accumulation (FlowAccumulation :97-119) -------------------------------------------------------
// Dataflow over positions from the sources: a position waits for its donors' areas, then sums them in
// donor-slot order.  A donor inside the wave's own tile of 64 positions is read from its lane by a
// register shuffle as soon as that lane finishes (rounds inside one pass), so a chain of positions
// inside a tile costs no memory round trip per link; donors in other tiles are polled in memory.
__global__ __launch_bounds__(MFD_WG) void k_mfd_area(const InstParams* __restrict__ PP, const int2* __restrict__ mtab,
                                                     int* __restrict__ ctr, int* __restrict__ err,
                                                     const double* __restrict__ A0, const int* __restrict__ mdp,
                                                     const double* __restrict__ mdw, double* __restrict__ Al) {
  MFD_WAVE_PROLOGUE
  MfdTiles T;
  mfd_start(T, ctr + FSX_MFD_CTR * inst + 8, err + inst);
  mfd_phase(T, (nn + 63) / 64, [&](int tt) {
    const int lo = 64 * tt, p = lo + lane;
    bool todo = p < nn;
    const long long gp = off + (todo ? p : 0);
    double a0 = 0.0, w[8], x[8], mine = 0.0;
    int d[8];
    unsigned needg = 0, needt = 0;  // donors in memory / in the tile
#pragma unroll
    for (int j = 0; j < 8; j++) {
      d[j] = todo ? mdp[8 * gp + j] : -1;
      w[j] = todo ? mdw[8 * gp + j] : 0.0;
      x[j] = 0.0;
      if (d[j] >= lo) needt |= 1u << j;
      else if (d[j] >= 0) needg |= 1u << j;
    }
    if (todo) a0 = A0[gp];
    mfd_lanes(T, [&](bool& prog) {
      if (todo) {
#pragma unroll
        for (int j = 0; j < 8; j++)
          if ((needg >> j) & 1u) {
            const double y = flow_ld_dev(Al + off + d[j]);
            if (!flow_sent(y)) {
              x[j] = y;
              needg &= ~(1u << j);
            }
          }
      }
      for (int round = 0; round < 64; round++) {
        bool did = false;
        if (todo && !needg && !needt) {
          double a = a0;
#pragma unroll
          for (int j = 0; j < 8; j++)
            if (d[j] >= 0) a = a + x[j] * w[j];
          mine = a;
          flow_st_dev(Al + gp, a);
          todo = false;
          prog = true;
          did = true;
        }
        const unsigned long long fin = __ballot(!todo);
#pragma unroll
        for (int j = 0; j < 8; j++) {
          const bool want = (needt >> j) & 1u;
          const int src = want ? d[j] - lo : lane;
          const double y = __shfl(mine, src, 64);
          if (want && ((fin >> src) & 1ull)) {
            x[j] = y;
            needt &= ~(1u << j);
            did = true;
          }
        }
        if (__ballot(did) == 0ull) break;
      }
      return todo;
    });
  });
}

// ---- per-position stream-power inputs (StreamPowerLaw :104-116 and :145-162) ----------------------
// n == 1: kf*dt*(a*w)**m/l where ht(ijk) >= ht(rcv); otherwise kf*dt*(a*w)**m/l**n where
// ht(ijk) > ht(rcv); -1 marks an edge that does not take part (constant over the Gauss-Seidel loop).
// Also: area back to grid order, ht in position order, the first pass's outputs set to the sentinel, the
// elevation coefficient (g, NaN on fixed nodes: elev = ht there), and the landscape's maxval(abs(h)) and
// maxval(g).
__global__ __launch_bounds__(FSX_CHUNK) void k_mfd_fact(const InstParams* __restrict__ PP, const Chunk* __restrict__ chunks,
                                                        const int* __restrict__ mpos, const int* __restrict__ mrec,
                                                        const int* __restrict__ mnrec, const double* __restrict__ mw,
                                                        const uint8_t* __restrict__ mdir, const double* __restrict__ Al,
                                                        double* __restrict__ area, const double* __restrict__ kfint,
                                                        const double* __restrict__ ht, const double* __restrict__ g,
                                                        const uint8_t* __restrict__ bc, double* __restrict__ mfact,
                                                        double* __restrict__ HTl, double* __restrict__ Hl,
                                                        double* __restrict__ Wl, double* __restrict__ DHl,
                                                        int* __restrict__ LSl, double* __restrict__ EC,
                                                        double* __restrict__ lsed, unsigned long long* __restrict__ redk,
                                                        long long* __restrict__ newton) {
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
  if (v == 0) newton[c.inst] = 0;  // summed over the solve's waves
  if (!live) return;
  const long long gp = (long long)P.off + mpos[g_];
  const double a = Al[gp], t = ht[g_];
  area[g_] = a;
  HTl[gp] = t;
  Hl[gp] = flow_sentinel();  // values of the first Gauss-Seidel pass not written yet
  Wl[gp] = flow_sentinel();
  DHl[gp] = flow_sentinel();
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

// ---- stream power, Gauss-Seidel loop ------------------------------------------------------------------
// Per iteration (phases of one launch): hp = h (from the second iteration; h and the erosion push reset
// to the sentinel); erosion pushed from the sources (dh(rcv) += dh(ijk)*w, :62-82) with the lake-sediment
// update of lake sills; elev (:84-90), with water level and lake sill cleared; the solve from base level up
// (:94-176): closed form over the receivers for n == 1, relaxed Newton (omega = 0.875/n) otherwise; lakes keep
// elev and leave their share of the sill's water volume for k_lake_acc; err = rms(h-hp) from partial
// sums over fixed blocks of positions, added in block order by every wave (so every wave takes the same
// decision), zero when g == 0 everywhere.  Node-indexed water, lake sill and sediment are kept for the
// lake fill of k_spl_post.
__global__ __launch_bounds__(MFD_WG) void k_mfd_spl(
    const InstParams* __restrict__ PP, const int2* __restrict__ mtab, int* __restrict__ ctr, int* __restrict__ err,
    const int* __restrict__ morder, const int* __restrict__ rp, const int* __restrict__ dn,
    const double* __restrict__ mfact, const int* __restrict__ mdp, const double* __restrict__ mdw,
    const double* __restrict__ EC, const double* __restrict__ Al, const double* __restrict__ HTl,
    double* __restrict__ Hl, double* __restrict__ HPl, double* __restrict__ DHl, double* __restrict__ ELl,
    double* __restrict__ Wl, int* __restrict__ LSl, double* __restrict__ water, int* __restrict__ lsill,
    double* __restrict__ lakec, double* __restrict__ lsed, const unsigned long long* __restrict__ redk,
    double* __restrict__ scal, double* __restrict__ part, int* __restrict__ mdec, int* __restrict__ ngs_out,
    long long* __restrict__ newton_out) {
  MFD_WAVE_PROLOGUE
  const double n = P.n;
  const bool n_is_one = fabs(n - 1.0) < 2.2250738585072014e-308;
  const int nkind = npow_kind(n);
  const double tol = 1.0e-4 * (dval(redk[2 * inst]) + 1.0);
  const bool gzero = dval(redk[2 * inst + 1]) < 2.2250738585072014e-308;
  const double sent = flow_sentinel();
  const int ntile = (nn + 63) / 64, nnode = (nn + 255) / 256, nblk = (nn + MFD_EB - 1) / MFD_EB;
  MfdTiles T;
  mfd_start(T, ctr + FSX_MFD_CTR * inst + 12, err + inst);
  double errv = 2.0 * tol;
  int ngs = 0;
  unsigned long long my_newton = 0;
  while (errv > tol && ngs < 99 && T.ok) {
    ngs++;
    const double* prev = ngs == 1 ? HTl : HPl;  // h at the start of the iteration
    if (ngs > 1)
      mfd_phase(T, nnode, [&](int tt) {
#pragma unroll
        for (int k = 0; k < 4; k++) {
          const int p = 256 * tt + 64 * k + lane;
          if (p < nn) {
            const long long gp = off + p;
            flow_st_dev(HPl + gp, flow_ld_dev(Hl + gp));
            flow_st_dev(Hl + gp, sent);
            flow_st_dev(DHl + gp, sent);
          }
        }
      });
    if (!gzero) {
      mfd_phase(T, ntile, [&](int tt) {  // erosion push, from the sources (in-tile donors by shuffles)
        const int lo = 64 * tt, p = lo + lane;
        bool todo = p < nn;
        const long long gp = off + (todo ? p : 0);
        double w[8], x[8], own = 0.0, mine = 0.0;
        int d[8], v = 0, r = 0, sl = -1;
        unsigned needg = 0, needt = 0;
#pragma unroll
        for (int j = 0; j < 8; j++) {
          d[j] = todo ? mdp[8 * gp + j] : -1;
          w[j] = todo ? mdw[8 * gp + j] : 0.0;
          x[j] = 0.0;
          if (d[j] >= lo) needt |= 1u << j;
          else if (d[j] >= 0) needg |= 1u << j;
        }
        if (todo) {
          v = morder[gp];
          r = rp[gp];
          own = HTl[gp] - flow_ld_dev(prev + gp);
          sl = flow_ldi_dev(LSl + gp);  // the lake sill of the previous iteration's solve
        }
        mfd_lanes(T, [&](bool& prog) {
          if (todo) {
#pragma unroll
            for (int j = 0; j < 8; j++)
              if ((needg >> j) & 1u) {
                const double y = flow_ld_dev(DHl + off + d[j]);
                if (!flow_sent(y)) {
                  x[j] = y;
                  needg &= ~(1u << j);
                }
              }
          }
          for (int round = 0; round < 64; round++) {
            bool did = false;
            if (todo && !needg && !needt) {
              double s = own;
#pragma unroll
              for (int j = 0; j < 8; j++)
                if (d[j] >= 0) s = s + x[j] * w[j];
              double dv = s;
              if (r >= 0) {
                const double t2 = s - own;
                if (sl == v) lsed[off + v] = t2 <= 0.0 ? 0.0 : t2;
                dv = t2 + own;
              } else {
                lsed[off + v] = s;
              }
              mine = dv;
              flow_st_dev(DHl + gp, dv);
              todo = false;
              prog = true;
              did = true;
            }
            const unsigned long long fin = __ballot(!todo);
#pragma unroll
            for (int j = 0; j < 8; j++) {
              const bool want = (needt >> j) & 1u;
              const int src = want ? d[j] - lo : lane;
              const double y = __shfl(mine, src, 64);
              if (want && ((fin >> src) & 1ull)) {
                x[j] = y;
                needt &= ~(1u << j);
                did = true;
              }
            }
            if (__ballot(did) == 0ull) break;
          }
          return todo;
        });
      });
      mfd_phase(T, nnode, [&](int tt) {  // elev; the solve's water level and lake sill cleared
#pragma unroll
        for (int k = 0; k < 4; k++) {
          const int p = 256 * tt + 64 * k + lane;
          if (p < nn) {
            const long long gp = off + p;
            const double t = HTl[gp], gc = EC[gp];
            flow_st_dev(ELl + gp,
                        isnan(gc) ? t : t + (flow_ld_dev(DHl + gp) - (t - flow_ld_dev(prev + gp))) * gc * P.dx * P.dy / Al[gp]);
            flow_st_dev(Wl + gp, sent);
            flow_sti_dev(LSl + gp, -1);
          }
        }
      });
    }
    mfd_phase(T, ntile, [&](int tt) {  // solve, from base level up (in-tile receivers by shuffles)
      const int hi = nn - 1 - 64 * tt, p = hi - lane;
      bool todo = p >= 0;
      const long long gp = off + (todo ? p : 0);
      int q[8], v = 0, r = -1;
      double f[8], hr[8], t = 0.0, e = 0.0, h0 = 0.0, w_rcv = 0.0, myH = 0.0, myW = 0.0;
      int sl_rcv = -1, mySL = -1;
      unsigned needg = 0, needt = 0;  // bits 0-7 receivers' h, 8 water level and 9 lake sill of rp
#pragma unroll
      for (int k = 0; k < 8; k++) {
        q[k] = todo ? dn[8 * gp + k] : -1;
        f[k] = todo ? mfact[8 * gp + k] : -1.0;
        hr[k] = 0.0;
        if (q[k] >= 0 && f[k] >= 0.0) {
          if (q[k] <= hi) needt |= 1u << k;
          else needg |= 1u << k;
        }
      }
      if (todo) {
        v = morder[gp];
        r = rp[gp];
        t = HTl[gp];
        e = gzero ? t : flow_ld_dev(ELl + gp);
        h0 = flow_ld_dev(prev + gp);
        if (r >= 0) {
          if (r <= hi) needt |= 3u << 8;
          else needg |= 3u << 8;
        }
      }
      mfd_lanes(T, [&](bool& prog) {
        if (todo) {
#pragma unroll
          for (int k = 0; k < 8; k++)
            if ((needg >> k) & 1u) {
              const double y = flow_ld_dev(Hl + off + q[k]);
              if (!flow_sent(y)) {
                hr[k] = y;
                needg &= ~(1u << k);
              }
            }
          if ((needg >> 8) & 1u) {
            const double y = flow_ld_dev(Wl + off + r);
            if (!flow_sent(y)) {
              w_rcv = y;
              needg &= ~(1u << 8);
            }
          }
          if ((needg >> 9) & 1u) {
            const int y = flow_ldi_dev(LSl + off + r);
            if (y >= 0) {
              sl_rcv = y;
              needg &= ~(1u << 9);
            }
          }
        }
        for (int round = 0; round < 64; round++) {
          bool did = false;
          if (todo && !needg && !needt) {
            todo = false;
            prog = true;
            did = true;
            if (r < 0) {  // base level
              myH = h0;
              myW = t;
              mySL = v;
              if (!gzero) {
                water[off + v] = t;
                lsill[off + v] = v;
                lakec[off + v] = -1.0;  // not a lake node
              }
            } else {
              double hv = h0;
              int sl;
              if (e > w_rcv) {
                // with Marine on, nodes below sea level keep their current h (src/StreamPowerLaw.f90:105,144:
                // h for n == 1, ht otherwise)
                if (q[0] >= 0 && !(P.run_marine && !((n_is_one ? h0 : t) >= P.sealevel))) {
                  if (n_is_one) {
                    double fs = e, df = 1.0;
#pragma unroll
                    for (int k = 0; k < 8; k++)
                      if (q[k] >= 0 && f[k] >= 0.0) {
                        fs = fs + f[k] * hr[k];
                        df = df + f[k];
                      }
                    hv = fs / df;
                  } else {
                    const double omega = 0.875 / n, tolp = P.newton_tol;
                    double errp = 2.0 * tolp;
                    int it = 0;
                    while (errp > tolp) {
                      double fs = hv - e, df = 1.0;
#pragma unroll
                      for (int k = 0; k < 8; k++)
                        if (q[k] >= 0 && f[k] >= 0.0) {
                          const double x = fmax(0.0, hv - hr[k]);
                          double pn, pn1;
                          npow(x, n, nkind, pn, pn1);
                          fs = fs + f[k] * pn;
                          df = df + f[k] * n * pn1;
                        }
                      const double hn = hv - fs / df;
                      errp = fabs(hn - hv);
                      hv = hv * (1.0 - omega) + hn * omega;
                      if (++it > 100000) {
                        atomicOr(err + inst, FSX_ERR_NEWTON);
                        break;
                      }
                    }
                    my_newton += it;
                  }
                }
                sl = v;
                if (hv < w_rcv) hv = w_rcv;
                if (!gzero) lakec[off + v] = -1.0;
              } else {
                hv = e;
                sl = sl_rcv;
                if (!gzero) lakec[off + v] = w_rcv - hv;  // its share of the sill's volume (k_lake_acc)
              }
              myH = hv;
              myW = w_rcv > hv ? w_rcv : hv;
              mySL = sl;
              if (!gzero) {
                water[off + v] = myW;
                lsill[off + v] = sl;
              }
            }
            flow_st_dev(Hl + gp, myH);
            flow_st_dev(Wl + gp, myW);
            flow_sti_dev(LSl + gp, mySL);
          }
          const unsigned long long fin = __ballot(!todo);
#pragma unroll
          for (int k = 0; k < 8; k++) {
            const bool want = (needt >> k) & 1u;
            const int src = want ? hi - q[k] : lane;
            const double y = __shfl(myH, src, 64);
            if (want && ((fin >> src) & 1ull)) {
              hr[k] = y;
              needt &= ~(1u << k);
              did = true;
            }
          }
          {
            const bool want = (needt >> 8) & 1u;
            const int src = want ? hi - r : lane;
            const double y = __shfl(myW, src, 64);
            const int z = __shfl(mySL, src, 64);
            if (want && ((fin >> src) & 1ull)) {
              w_rcv = y;
              sl_rcv = z;
              needt &= ~(3u << 8);
              did = true;
            }
          }
          if (__ballot(did) == 0ull) break;
        }
        return todo;
      });
    });
    if (gzero) {
      errv = 0.0;
      break;
    }
    mfd_phase(T, nblk, [&](int blk) {  // partial sums of (h - hp)**2
      double s = 0.0;
      for (int i = lane; i < MFD_EB; i += 64) {
        const int p = blk * MFD_EB + i;
        if (p < nn) {
          const double dd = flow_ld_dev(Hl + off + p) - flow_ld_dev(prev + off + p);
          s += dd * dd;
        }
      }
      for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
      if (lane == 0) flow_st_dev(part + P.c0 + blk, s);
    });
    if (!mfd_wait(T, T.base)) break;
    double x = 0.0;
    for (int j = lane; j < nblk; j += 64) x += flow_ld_dev(part + P.c0 + j);
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
    x = __shfl(x, 0, 64);  // lane 0's sum in every lane
    // The iteration's decision (1 go on, 2 stop) is kept: a wave that reaches this point late, when later
    // iterations may have overwritten the partial sums, takes the decision the others took (a wave writes
    // it and drains the store before any of its later phases, so partials are overwritten only once it
    // is readable).
    int* dp = mdec + MFD_MAXGS * inst + ngs;
    int dcs = flow_ldi_dev(dp);
    if (dcs == 0) {
      dcs = sqrt(x / (double)nn) > tol ? 1 : 2;
      if (lane == 0) flow_sti_dev(dp, dcs);
      wave_drain();
    }
    errv = dcs == 1 ? 2.0 * tol : 0.0;
  }
  if (my_newton) atomicAdd(reinterpret_cast<unsigned long long*>(newton_out + inst), my_newton);
  if (wk.y == 0 && threadIdx.x == 0) {
    ngs_out[inst] = ngs;
    scal[inst] = gzero ? 0.0 : 1.0;
  }
}

// ---------------------------------------------------------------------------------------------------
void launch_mfd_route(DevState& S, hipStream_t st) {
  if (!S.nmfd) return;
  launch_fill_int(S.mctr, FSX_MFD_CTR * S.ninst, 0, st);
  k_mfd_init<<<S.nchunks, FSX_CHUNK, 0, st>>>(S.P, S.chunks, S.mkahn, S.mq, S.mls, S.hwater);
  k_mfd_fill<<<S.nmfd, MFD_WG, 0, st>>>(S.P, S.mfd_tab, S.mctr, S.err, S.nlev, S.order, S.rec, S.ht,
                                         S.hwater, S.lake_depth);
  k_mfd_rec<<<S.nchunks, FSX_CHUNK, 0, st>>>(S.P, S.chunks, S.hwater, S.rec, S.mrec, S.mw, S.mdir, S.mnrec, S.mflag,
                                              S.mkahn, S.pexp);
  k_mfd_don<<<S.nchunks, FSX_CHUNK, 0, st>>>(S.P, S.chunks, S.mrec, S.mnrec, S.mdon, S.mndon, S.err);
  k_mfd_kahn<<<S.nmfd, MFD_WG, 0, st>>>(S.P, S.mfd_tab, S.mctr, S.err, S.mkahn, S.mrec, S.mnrec, S.mflag, S.rec, S.mq,
                                         S.mls);
  k_mfd_lscan<<<S.ninst, FSX_WG, 0, st>>>(S.P, S.mctr, S.mls, S.mcur, S.mnlev, S.err);
  k_mfd_scatter<<<S.nchunks, FSX_CHUNK, 0, st>>>(S.P, S.chunks, S.mq, S.mcur, S.mstack);
  k_mfd_pos<<<S.nchunks, FSX_CHUNK, 0, st>>>(S.P, S.chunks, S.mstack, S.mpos);
  k_mfd_prep<<<S.nchunks, FSX_CHUNK, 0, st>>>(S.P, S.chunks, S.mpos, S.rec, S.mrec, S.mnrec, S.mw, S.mdon, S.mndon,
                                               S.precip, S.A0, S.Al, S.mdp, S.mdw, S.dn, S.rp);
  k_mfd_area<<<S.nmfd, MFD_WG, 0, st>>>(S.P, S.mfd_tab, S.mctr, S.err, S.A0, S.mdp, S.mdw, S.Al);
}

void launch_mfd_spl(DevState& S, hipStream_t st) {
  if (!S.nmfd) return;
  launch_fill_int(reinterpret_cast<int*>(S.redk), 4LL * S.ninst, 0, st);
  launch_fill_int(S.mctr, FSX_MFD_CTR * S.ninst, 0, st);
  launch_fill_int(S.mdec, MFD_MAXGS * S.ninst, 0, st);
  k_mfd_fact<<<S.nchunks, FSX_CHUNK, 0, st>>>(S.P, S.chunks, S.mpos, S.mrec, S.mnrec, S.mw, S.mdir, S.Al, S.area,
                                               S.kfint, S.ht, S.g, S.bc, S.mfact, S.HTl, S.Hl, S.Wl, S.DHl, S.LSl, S.F,
                                               S.lsed, S.redk, S.newton);
  k_mfd_spl<<<S.nmfd, MFD_WG, 0, st>>>(S.P, S.mfd_tab, S.mctr, S.err, S.mstack, S.rp, S.dn, S.mfact, S.mdp, S.mdw, S.F,
                                        S.Al, S.HTl, S.Hl, S.HPl, S.DHl, S.ELl, S.Wl, S.LSl, S.water, S.lsill, S.lakec,
                                        S.lsed, S.redk, S.scal, S.mpart, S.mdec, S.ngs, S.newton);
}
