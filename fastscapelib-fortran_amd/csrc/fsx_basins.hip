// fsx_basins.hip -- closed-basin rerouting (LocalMinima, src/LocalMinima.f90:1-132) on the GPU.
//
//  1. basin labels: root of every node by pointer jumping; basin id = rank of the root among the
//     roots in ascending index (compute_basins numbers basins in stack order, whose roots come in
//     ascending index, :136-169).
//  2. passes: one thread per closed basin walks its nodes in the reference's depth-first preorder
//     (stackless: donors ascending, climbing back through receivers) and scans the 8 neighbours in
//     the order of :239-240, keeping per neighbour basin the first strictly lowest pass (:284-317).
//     Edge order = basins in id order, each basin's edges in discovery order, one virtual edge
//     (-1e10) per open basin other than basin0 at that basin's place (:263-276).
//  3. edge order of loc_min_3_indexx (:947-1046), reproduced exactly by a parallel form of its
//     partition step (below), including the order it gives equal weights.
//  4. Boruvka minimum spanning forest keyed by that order == Kruskal in that order (:333-385).
//  5. orientation from basin0 (order_tree, :389-533) by a breadth-first walk of the tree.
//  6. receiver path reversal per tree edge (correct_receivers_v2, :601-662), one thread per edge;
//     paths lie in disjoint basins, so edges are independent.
#include <algorithm>
#include <climits>

#include "fsx_scan.h"
#include "fsx_state.h"

#define NODE_PROLOGUE                        \
  const Chunk c = chunks[chunk_id()];        \
  const InstParams& P = PP[c.inst];          \
  const int v = c.start + (int)threadIdx.x;  \
  const bool live = P.active && P.run_spl && v < P.nn;

// Closed basins of more than `big` nodes (DevState::walk_big, at least 64) are searched node-parallel
// (k_big_*), since one thread's walk takes time in proportion to its basin.
#define BIG_NBIG 0    // bigctr[0]: big basins listed
#define BIG_NRUL 1    // bigctr[1]: rulers listed
#define BIG_FLAG0 2   // bigctr[2..]: convergence flags of the ruler jumps
#define BIG_ROUNDS 24
#define BIG_NCTR (BIG_FLAG0 + BIG_ROUNDS)

__global__ void k_fill_int(int* __restrict__ a, long long n, int val) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    a[i] = val;
}

void launch_fill_int(int* a, long long n, int val, hipStream_t st) {
  const int blocks = (int)std::min<long long>(1024, (n + 255) / 256 > 0 ? (n + 255) / 256 : 1);
  k_fill_int<<<blocks, 256, 0, st>>>(a, n, val);
}

// ---- 1. labels: root of every node from the pointer-jumping pass (fsx_tree.hip) ----------------
// roots per chunk
__global__ __launch_bounds__(FSX_CHUNK) void k_root_count(const InstParams* __restrict__ PP, const Chunk* __restrict__ chunks,
                                                          const int* __restrict__ rec, int* __restrict__ chunkcnt) {
  __shared__ int lds[FSX_CHUNK / 64];
  NODE_PROLOGUE
  const int f = (live && rec[P.off + v] == v) ? 1 : 0;
  const int t = block_sum<FSX_CHUNK>(f, lds);
  if (threadIdx.x == 0) chunkcnt[chunk_id()] = t;
}

// exclusive scan of the per-chunk root counts of each landscape; nbasins; basin0 reset
__global__ __launch_bounds__(FSX_WG) void k_root_scan(const InstParams* __restrict__ PP, int* __restrict__ chunkcnt,
                                                      int* __restrict__ nbasins, int* __restrict__ basin0) {
  __shared__ int lds[FSX_WG / 64];
  const int inst = blockIdx.x;
  const InstParams& P = PP[inst];
  if (!P.active || !P.run_spl) return;
  const int c0 = P.c0, c1 = P.c1;
  int carry = 0;
  for (int b = c0; b < c1; b += FSX_WG) {
    const int i = b + threadIdx.x;
    const int x = i < c1 ? chunkcnt[i] : 0;
    int tot;
    const int incl = block_incl_scan<FSX_WG>(x, lds, tot);
    if (i < c1) chunkcnt[i] = carry + incl - x;
    carry += tot;
  }
  if (threadIdx.x == 0) {
    nbasins[inst] = carry;
    basin0[inst] = INT_MAX;
  }
}

__global__ __launch_bounds__(FSX_CHUNK) void k_root_rank(const InstParams* __restrict__ PP, const Chunk* __restrict__ chunks,
                                                         const int* __restrict__ rec, const uint8_t* __restrict__ bc,
                                                         const int* __restrict__ chunkcnt, int* __restrict__ rootrank,
                                                         int* __restrict__ outlet, int* __restrict__ basin0) {
  __shared__ int lds[FSX_CHUNK / 64];
  NODE_PROLOGUE
  const int f = (live && rec[P.off + v] == v) ? 1 : 0;
  int tot;
  const int incl = block_incl_scan<FSX_CHUNK>(f, lds, tot);
  int open_rank = INT_MAX;  // basin0: the smallest rank of an open basin, one atomic per wave
  if (f) {
    const int r = chunkcnt[chunk_id()] + incl - 1;
    rootrank[P.off + v] = r;
    outlet[P.off + r] = v;
    if (bc[P.off + v]) open_rank = r;
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const int y = __shfl_xor(open_rank, o, 64);
    open_rank = y < open_rank ? y : open_rank;
  }
  // most waves holding an open root see a smaller one already stored (ranks follow node order and
  // the first open root comes early), so they skip the atomic on the one contended word
  if ((threadIdx.x & 63) == 0 && open_rank != INT_MAX && open_rank < __atomic_load_n(basin0 + c.inst, __ATOMIC_RELAXED))
    atomicMin(basin0 + c.inst, open_rank);
}

// The smallest donor of r (donor mask m) larger than u, or -1.  Interior nodes: the mask's bit order
// is ascending index order, so the set bits are visited in order without decoding all donors.
__device__ __forceinline__ int donor_after(const InstParams& P, int r, unsigned m, int u) {
  if (!m) return -1;
  const int i = r % P.nx, j = r / P.nx;
  if (i > 0 && j > 0 && i < P.nx - 1 && j < P.ny - 1) {
    while (m) {
      const int bit = __ffs(m) - 1;
      m &= m - 1;
      const int dd = bit < 4 ? bit : bit + 1;
      const int q = r + (dd % 3 - 1) + (dd / 3 - 1) * P.nx;
      if (q > u) return q;
    }
    return -1;
  }
  int d[8];
  const int nd = donors_sorted(P, r, m, d);
  for (int t = 0; t < nd; t++)
    if (d[t] > u) return d[t];
  return -1;
}

__global__ __launch_bounds__(FSX_CHUNK) void k_basin_label(const InstParams* __restrict__ PP, const Chunk* __restrict__ chunks,
                                                           const unsigned long long* __restrict__ J,
                                                           const int* __restrict__ rootrank, const uint8_t* __restrict__ bc,
                                                           const double* __restrict__ h, const uint8_t* __restrict__ dmask,
                                                           const int* __restrict__ rec,
                                                           int* __restrict__ basin, int* __restrict__ bsize,
                                                           WalkNode* __restrict__ wnode) {
  NODE_PROLOGUE
  int bv = INT_MIN;  // lanes past the landscape take part in the run detection below as a non-basin
  if (live) {
    const int r = (int)(unsigned)(J[P.off + v] & 0xffffffffu);
    const int b = rootrank[P.off + r];
    bv = !bc[P.off + r] ? b : -(b + 1);  // signed: open basins negative
    basin[P.off + v] = bv;
    // successor of v in the depth-first preorder of its tree (the order of src/FlowRouting.f90
    // find_stack, donors ascending): the first donor, else the next larger sibling of the nearest
    // ancestor-or-self that has one; -1 when the traversal is back at the root.  Donors come from
    // the donor masks.  Each receiver edge is climbed by one leaf only (the last of its subtree), so
    // the climbs cost O(nodes) in total.
    const uint8_t* DM = dmask + P.off;
    int sv = donor_after(P, v, DM[v], -1);
    if (sv < 0) {
      // a climb: each step reads the receiver's donor mask and its own receiver together (one memory round
      // trip per step)
      int u = v, ru = rec[P.off + v];
      for (int guard = 0; guard < P.nn; guard++) {  // the receivers form a forest; bounded anyway
        if (ru == u) break;
        const unsigned mr = DM[ru];
        const int rru = rec[P.off + ru];
        const int nu = donor_after(P, ru, mr, u);
        if (nu >= 0) {
          sv = nu;
          break;
        }
        u = ru;
        ru = rru;
      }
    }
    reinterpret_cast<ulonglong2*>(wnode)[P.off + v] =
        make_ulonglong2((unsigned long long)__double_as_longlong(h[P.off + v]),
                        (unsigned long long)(unsigned)bv | ((unsigned long long)(unsigned)sv << 32));
  }
  // basin sizes: consecutive nodes of a row mostly share a basin, so each run of equal ids within the
  // wave adds its length with one atomic
  const int lane = threadIdx.x & 63;
  const int prev = __shfl_up(bv, 1, 64);
  const unsigned long long starts = __ballot(lane == 0 || prev != bv);
  if (live && bv >= 0 && ((starts >> lane) & 1ull)) {
    const unsigned long long after = lane == 63 ? 0ull : starts >> (lane + 1);
    const int len = after ? __ffsll((long long)after) : 64 - lane;
    atomicAdd(bsize + P.off + bv, len);
  }
}

// exclusive scan of closed-basin sizes -> candidate slab offsets; clears edge counts; lists the closed
// basins of more than `big` nodes
__global__ __launch_bounds__(FSX_WG) void k_bsize_scan(const InstParams* __restrict__ PP, const int* __restrict__ nbasins,
                                                       const int* __restrict__ bsize, int* __restrict__ soff,
                                                       int* __restrict__ ecount, int2* __restrict__ biglist,
                                                       int* __restrict__ bigctr, int bigcap, int big) {
  __shared__ int lds[FSX_WG / 64];
  const int inst = blockIdx.x;
  const InstParams& P = PP[inst];
  if (!P.active || !P.run_spl) return;
  const int nb = nbasins[inst];
  int carry = 0;
  for (int b0 = 0; b0 < nb; b0 += FSX_WG) {
    const int b = b0 + threadIdx.x;
    const int x = b < nb ? bsize[P.off + b] : 0;
    int tot;
    const int incl = block_incl_scan<FSX_WG>(x, lds, tot);
    if (b < nb) {
      soff[P.off + b] = carry + incl - x;
      ecount[P.off + b] = 0;
      if (x > big) {  // searched by the big-basin kernels, not by the walk
        const int e = atomicAdd(bigctr + BIG_NBIG, 1);
        if (e < bigcap) biglist[e] = make_int2(inst, b);
      }
    }
    carry += tot;
  }
}

// ---- 2. passes (connect_basins) -------------------------------------------------------------------
// Which neighbours the pass search of a node examines (src/LocalMinima.f90:281-297): those inside the
// grid whose basin is open or has a larger id than the node's closed basin.  Node-parallel over the
// labels, so that the walk loads only these neighbours' records (about two of eight).
__global__ __launch_bounds__(FSX_CHUNK) void k_walk_mask(const InstParams* __restrict__ PP, const Chunk* __restrict__ chunks,
                                                         const int* __restrict__ basin, uint8_t* __restrict__ wmask) {
  NODE_PROLOGUE
  if (!live) return;
  const int* Bn = basin + P.off;
  const int b = Bn[v];
  unsigned m = 0;
  if (b >= 0) {
    constexpr int II[8] = {-1, 1, 0, 0, -1, 1, -1, 1};
    constexpr int JJ[8] = {0, 0, -1, 1, -1, -1, 1, 1};
    const int i = v % P.nx, j = v / P.nx;
#pragma unroll
    for (int k = 0; k < 8; k++) {
      const int ki = i + II[k], kj = j + JJ[k];
      if (ki < 0 || ki >= P.nx || kj < 0 || kj >= P.ny) continue;
      const int Bs = Bn[kj * P.nx + ki];
      if (Bs < 0 || Bs > b) m |= 1u << k;
    }
  }
  wmask[P.off + v] = (uint8_t)m;
}


// Pass-table entries per thread kept in LDS; further entries live in the slab.  8 covers the
// neighbouring basins of nearly every basin while keeping LDS per workgroup at 6 KB, so that the
// walk, a chain of dependent loads per lane, runs at 8 waves per SIMD.
#define WALK_TCAP 8
// launched with gx blocks per landscape (DevState::walk_gx); thread x walks basins x, x + gx*64, ...
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(8)))
void k_basin_walk(const InstParams* __restrict__ PP, const int* __restrict__ nbasins,
                  const uint8_t* __restrict__ bc, const int* __restrict__ outlet,
                  const WalkNode* __restrict__ wnode, const uint8_t* __restrict__ wmask,
                                                   const int* __restrict__ soff, const int* __restrict__ bsize,
                                                   int* __restrict__ ecount, int* __restrict__ sB, int* __restrict__ sp,
                                                   int* __restrict__ sq, double* __restrict__ sw, int* __restrict__ sh,
                                                   int* __restrict__ err, int ninst, int gx, int big) {
  __shared__ int lB[WALK_TCAP * 64];
  __shared__ double lW[WALK_TCAP * 64];
  int* LB = lB + threadIdx.x;
  double* LW = lW + threadIdx.x;
  constexpr int II[8] = {-1, 1, 0, 0, -1, 1, -1, 1};
  constexpr int JJ[8] = {0, 0, -1, 1, -1, -1, 1, 1};
  int inst, sub;
  if (!xcd_split(ninst, gx, inst, sub)) return;
  const InstParams& P = PP[inst];
  if (!P.active || !P.run_spl) return;
  const int nb = nbasins[inst];
  const int off = P.off, nx = P.nx;
  const ulonglong2* WN = reinterpret_cast<const ulonglong2*>(wnode) + off;
  const uint8_t* MK = wmask + off;
  // One flat loop of visits: a lane that finishes a basin starts its next one at once, so the lanes
  // of a wave do not wait for the largest basin of every round (basins differ widely in size).
  const int STRIDE = gx * 64;
  int b = sub * 64 + threadIdx.x - STRIDE;
  int v = 0, ne = 0, cap = 0, visits = 0;
  int *TB = nullptr, *TP = nullptr, *TQ = nullptr, *TH = nullptr;
  double* TW = nullptr;
  // entries beyond WALK_TCAP (large basins) are found through an open-addressing hash over TB:
  // TH[slot] = entry index, -1 empty; hshift = 0 until the first such entry; hfull once an insert
  // found no free slot (then a hash miss is confirmed by a scan of those entries)
  int hshift = 0, hbits = 0;
  bool hfull = false;
  auto next_basin = [&]() -> bool {  // the lane's next closed basin; false when none is left
    for (b += STRIDE; b < nb; b += STRIDE) {
      const int o = outlet[off + b];
      if (bc[off + o]) continue;  // open basin: no scan
      if (bsize[off + b] > big) continue;  // k_big_* search it
      v = o;
      ne = 0;
      visits = 0;
      const long long sbase = (long long)FSX_SLAB_CAP * (off + soff[off + b]);
      TB = sB + sbase;
      TP = sp + sbase;
      TQ = sq + sbase;
      TW = sw + sbase;
      TH = sh + sbase;
      cap = FSX_SLAB_CAP * bsize[off + b];
      hshift = 0;
      hfull = false;
      return true;
    }
    return false;
  };
  auto finish = [&]() {
    const int nl = ne < WALK_TCAP ? ne : WALK_TCAP;
    for (int t = 0; t < nl; t++) { TB[t] = LB[64 * t]; TW[t] = LW[64 * t]; }
    ecount[off + b] = ne;
  };
  bool live = next_basin();
  // the record and mask of v: loaded at a basin's first node, else prefetched with the previous node's
  // neighbours (the preorder successor is known once v's record arrives), so the chase of successors and
  // the neighbours' loads overlap
  ulonglong2 wv = make_ulonglong2(0ull, 0ull);
  unsigned mv = 0;
  bool have = false;
  while (live) {
    if (++visits > P.nn) {  // not a tree: give up on this basin
      atomicOr(err + inst, FSX_ERR_LEVELS);
      finish();
      live = next_basin();
      have = false;
      continue;
    }
    // scan the 8 neighbours of v (no wrap), src/LocalMinima.f90:281-320: only those k_walk_mask
    // marked (inside the grid, open basin or larger id) are loaded
    const int i = v % nx, j = v / nx;
    if (!have) {
      wv = WN[v];
      mv = MK[v];
    }
    const double hv = __longlong_as_double((long long)wv.x);
    const int fv = (int)(unsigned)(wv.y >> 32);
    ulonglong2 wnext = make_ulonglong2(0ull, 0ull);
    unsigned mnext = 0;
    if (fv >= 0) {
      wnext = WN[fv];
      mnext = MK[fv];
    }
    int qs[8], bs[8];
    double hq[8];
#pragma unroll
    for (int k = 0; k < 8; k++) {
      qs[k] = (j + JJ[k]) * nx + (i + II[k]);
      const ulonglong2 wq = ((mv >> k) & 1u) ? WN[qs[k]] : make_ulonglong2(0ull, 0ull);
      bs[k] = (int)(unsigned)(wq.y & 0xffffffffull);
      hq[k] = __longlong_as_double((long long)wq.x);
    }
#pragma unroll
    for (int k = 0; k < 8; k++) {
      if (!((mv >> k) & 1u)) continue;
      const int Bs = bs[k];
      const bool closedB = Bs >= 0;
      const int B = closedB ? Bs : -(Bs + 1);
      const double w = hv > hq[k] ? hv : hq[k];
      int t = 0;
      const int nl = ne < WALK_TCAP ? ne : WALK_TCAP;
      while (t < nl && LB[64 * t] != B) t++;
      int hslot = -1;  // free hash slot for B when it is new
      if (t == nl && ne > WALK_TCAP) {
        const unsigned hmask = (1u << hbits) - 1u;
        unsigned s0 = ((unsigned)B * 2654435761u) >> hshift;
        t = ne;
        for (int pr = 0; pr <= (int)hmask; pr++, s0 = (s0 + 1u) & hmask) {
          const int x = TH[s0];
          if (x < 0) { hslot = (int)s0; break; }
          if (TB[x] == B) { t = x; break; }
        }
        if (t == ne && hfull) {
          t = WALK_TCAP;
          while (t < ne && TB[t] != B) t++;
        }
      }
      if (t == ne) {
        if (ne >= cap) { atomicOr(err + inst, FSX_ERR_NCONN); continue; }
        if (t < WALK_TCAP) {
          LB[64 * t] = B;
          LW[64 * t] = w;
        } else {
          if (t == WALK_TCAP) {  // first entry past LDS: a hash over the slab, capacity 2^hbits <= cap
            hbits = 31 - __clz(cap);
            hshift = 32 - hbits;
            for (int s = 0; s < (1 << hbits); s++) TH[s] = -1;
            hslot = (int)(((unsigned)B * 2654435761u) >> hshift);
          }
          if (hslot >= 0) TH[hslot] = t;
          else hfull = true;
          TB[t] = B;
          TW[t] = w;
        }
        TP[t] = v; TQ[t] = qs[k];
        ne++;
      } else {
        const double tw = t < WALK_TCAP ? LW[64 * t] : TW[t];
        if (w < tw) {
          if (t < WALK_TCAP) LW[64 * t] = w; else TW[t] = w;
          TP[t] = v; TQ[t] = qs[k];
        }
      }
    }
    // next node in depth-first preorder (k_basin_label's successor), -1 past the basin's last node
    if (fv >= 0) {
      v = fv;
      wv = wnext;
      mv = mnext;
      have = true;
    } else {
      finish();
      live = next_basin();
      have = false;
    }
  }
}

// ---- 2b. passes of big basins, node-parallel ---------------------------------------------------------
// The walk's results for a basin A are, for every basin B that the search of A's nodes meets, the first
// (preorder rank, k) at which it meets B -- the edge order -- and the first (rank, k) among the pairs of
// least weight max(h(v), h(q)) -- the pass.  Both are minima over (rank, k), so once every node knows its
// rank in A's depth-first preorder, the nodes of A can be searched in any order:
//  1. ranks: rulers (A's root and nodes picked by a hash, about one in 32) walk the preorder successors
//     up to the next ruler, giving every node its ruler and offset; pointer jumping over the rulers gives
//     each its distance to the end of A's preorder, so rank = size(A) - that distance + offset;
//  2. a hash table of the basins B met, in A's candidate slab (capacity the largest power of two not
//     above 4 size(A)), takes the minima by atomics in two node-parallel passes: the first meeting and
//     the least weight, then the first (rank, k) of least weight (its node is the basin's preorder at
//     that rank);
//  3. one workgroup per big basin ranks the table entries by first meeting and writes them out as the walk
//     does (neighbour basin, weight, pass nodes at the slab's start, count in ecount).
__device__ __forceinline__ bool big_ruler_pick(int v) { return (((unsigned)v * 0x9E3779B1u) >> 27) == 0u; }
__device__ __forceinline__ int big_cap(int size) {  // largest power of two <= 4 size
  return (int)(1u << (31 - __clz((unsigned)(4 * size))));
}
#define BIG_NONE 0xffffffffu

// Rulers of big basins, listed for the segment walks and the jumps without a shared counter: a count per
// chunk (pass 0), an exclusive scan of the counts (k_big_rscan), the list entries (pass 1).  An entry is
// (chunk << 8 | lane), so the list holds at most one entry per node.
__device__ __forceinline__ bool big_is_ruler(const InstParams& P, int v, bool live, const WalkNode* wnode,
                                             const int* rec, const int* bsize, int big) {
  if (!live) return false;
  const int bv = wnode[P.off + v].basin;
  return bv >= 0 && bsize[P.off + bv] > big && (rec[P.off + v] == v || big_ruler_pick(v));
}

__global__ __launch_bounds__(FSX_CHUNK) void k_big_pick(const InstParams* __restrict__ PP, const Chunk* __restrict__ chunks,
                                                        const WalkNode* __restrict__ wnode, const int* __restrict__ rec,
                                                        const int* __restrict__ bsize, int* __restrict__ rchunk,
                                                        int* __restrict__ rlist, int big, int pass) {
  __shared__ int s_w[FSX_CHUNK / 64];
  NODE_PROLOGUE
  const bool ruler = big_is_ruler(P, v, live, wnode, rec, bsize, big);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const unsigned long long m = __ballot(ruler);
  if (lane == 0) s_w[wid] = __popcll(m);
  __syncthreads();
  const int cid = chunk_id();
  if (pass == 0) {
    if (threadIdx.x == 0) {
      int n = 0;
      for (int w = 0; w < FSX_CHUNK / 64; w++) n += s_w[w];
      rchunk[cid] = n;
    }
    return;
  }
  if (!ruler) return;
  int base = rchunk[cid];
  for (int w = 0; w < wid; w++) base += s_w[w];
  rlist[base + __popcll(m & ((1ull << lane) - 1ull))] = (cid << 8) | (int)threadIdx.x;
}

// exclusive scan of the per-chunk ruler counts (one workgroup); the total is the list's length
__global__ __launch_bounds__(FSX_WG) void k_big_rscan(int* __restrict__ rchunk, int nchunks, int* __restrict__ bigctr) {
  __shared__ int lds[FSX_WG / 64];
  int carry = 0;
  for (int c0 = 0; c0 < nchunks; c0 += FSX_WG) {
    const int c = c0 + (int)threadIdx.x;
    const int x = c < nchunks ? rchunk[c] : 0;
    int tot;
    const int incl = block_incl_scan<FSX_WG>(x, lds, tot);
    if (c < nchunks) rchunk[c] = carry + incl - x;
    carry += tot;
  }
  if (threadIdx.x == 0) bigctr[BIG_NRUL] = carry;
}

// the ruler of list entry e: (landscape, node)
__device__ __forceinline__ int2 big_ruler(const Chunk* chunks, int code) {
  const Chunk c = chunks[code >> 8];
  return make_int2(c.inst, c.start + (code & 255));
}

// each listed ruler walks the preorder successors up to the next ruler: ruler and offset of the nodes on
// the way, (next ruler, nodes up to it) of the ruler
__global__ __launch_bounds__(256) void k_big_seg(const InstParams* __restrict__ PP, const Chunk* __restrict__ chunks,
                                                 const int* __restrict__ rlist, const int* __restrict__ bigctr,
                                                 const WalkNode* __restrict__ wnode, unsigned long long* __restrict__ wro,
                                                 unsigned long long* __restrict__ rjmp, int* __restrict__ err) {
  const int n = bigctr[BIG_NRUL];
  for (int e = blockIdx.x * 256 + (int)threadIdx.x; e < n; e += gridDim.x * 256) {
    const int2 r = big_ruler(chunks, rlist[e]);
    const InstParams& P = PP[r.x];
    const ulonglong2* WN = reinterpret_cast<const ulonglong2*>(wnode) + P.off;
    const int v = r.y;
    int u = v;
    bool done = false;
    for (unsigned d = 0; d <= (unsigned)P.nn; d++) {
      wro[P.off + u] = ((unsigned long long)(unsigned)v << 32) | d;
      const int su = (int)(unsigned)(WN[u].y >> 32);
      if (su < 0 || big_ruler_pick(su)) {
        rjmp[P.off + v] = ((unsigned long long)(su < 0 ? BIG_NONE : (unsigned)su) << 32) | (d + 1u);
        done = true;
        break;
      }
      u = su;
    }
    if (!done) atomicOr(err + r.x, FSX_ERR_LEVELS);  // the successors do not end: not a tree
  }
}

// (next ruler, nodes up to it) -> (end, nodes to the end of the preorder), two hops per launch
__global__ __launch_bounds__(256) void k_big_jump(const InstParams* __restrict__ PP, const Chunk* __restrict__ chunks,
                                                  const int* __restrict__ rlist, unsigned long long* __restrict__ rjmp,
                                                  int* __restrict__ bigctr, int iter) {
  if (iter > 0 && bigctr[BIG_FLAG0 + iter - 1] == 0) return;
  const int n = bigctr[BIG_NRUL];
  bool changed = false;
  for (int e = blockIdx.x * 256 + (int)threadIdx.x; e < n; e += gridDim.x * 256) {
    const int2 r = big_ruler(chunks, rlist[e]);
    const int nn = PP[r.x].nn;
    unsigned long long* J = rjmp + PP[r.x].off;
    unsigned long long w = J[r.y];
#pragma unroll
    for (int hop = 0; hop < 2; hop++) {
      const unsigned x = (unsigned)(w >> 32);
      if (x >= (unsigned)nn) break;  // the end (BIG_NONE)
      const unsigned long long wx = J[x];
      w = (wx & 0xffffffff00000000ull) | (unsigned long long)((unsigned)w + (unsigned)wx);
      J[r.y] = w;
      changed = true;
    }
  }
  if (changed) bigctr[BIG_FLAG0 + iter] = 1;
}

// preorder rank of every node of a big basin, and the node at each rank
__global__ __launch_bounds__(FSX_CHUNK) void k_big_rank(const InstParams* __restrict__ PP, const Chunk* __restrict__ chunks,
                                                        const WalkNode* __restrict__ wnode, const int* __restrict__ bsize,
                                                        const unsigned long long* __restrict__ wro,
                                                        const unsigned long long* __restrict__ rjmp, const int* __restrict__ soff,
                                                        int* __restrict__ wrank, int* __restrict__ wseq, int* __restrict__ err,
                                                        int big) {
  NODE_PROLOGUE
  if (!live) return;
  const int bv = wnode[P.off + v].basin;
  if (bv < 0) return;
  const int sz = bsize[P.off + bv];
  if (sz <= big) return;
  const unsigned long long ro = wro[P.off + v];
  const unsigned long long rj = rjmp[P.off + (int)(unsigned)(ro >> 32)];
  const int rank = sz - (int)(unsigned)(rj & 0xffffffffull) + (int)(unsigned)(ro & 0xffffffffull);
  if (rank < 0 || rank >= sz) {  // cannot happen for a tree; keeps every index below in bounds
    atomicOr(err + c.inst, FSX_ERR_LEVELS);
    wrank[P.off + v] = 0;
    return;
  }
  wrank[P.off + v] = rank;
  wseq[P.off + soff[P.off + bv] + rank] = v;  // the basin's preorder, in its own slice of node offsets
}

// empty tables of the big basins (one workgroup per listed basin)
__global__ __launch_bounds__(256) void k_big_init(const InstParams* __restrict__ PP, const int2* __restrict__ biglist,
                                                  const int* __restrict__ bigctr, int bigcap, const int* __restrict__ bsize,
                                                  const int* __restrict__ soff, int* __restrict__ sB, int* __restrict__ sp,
                                                  int* __restrict__ sq, double* __restrict__ sw, int* __restrict__ sh) {
  const int n = min(bigctr[BIG_NBIG], bigcap);
  for (int e = blockIdx.x; e < n; e += gridDim.x) {
    const int2 ab = biglist[e];
    const InstParams& P = PP[ab.x];
    const long long sbase = (long long)FSX_SLAB_CAP * (P.off + soff[P.off + ab.y]);
    const int C = big_cap(bsize[P.off + ab.y]);
    for (int t = threadIdx.x; t < C; t += 256) {
      sh[sbase + t] = -1;
      sB[sbase + t] = -1;
      sp[sbase + t] = INT_MAX;
      sq[sbase + t] = INT_MAX;
      reinterpret_cast<unsigned long long*>(sw)[sbase + t] = ~0ull;
    }
  }
}

// pass 0: first meeting and least weight; 1: first (rank, k) of least weight
__global__ __launch_bounds__(FSX_CHUNK) void k_big_cand(const InstParams* __restrict__ PP, const Chunk* __restrict__ chunks,
                                                        const WalkNode* __restrict__ wnode, const uint8_t* __restrict__ wmask,
                                                        const int* __restrict__ bsize, const int* __restrict__ soff,
                                                        const int* __restrict__ wrank, int* __restrict__ sB,
                                                        int* __restrict__ sp, int* __restrict__ sq, double* __restrict__ sw,
                                                        int* __restrict__ sh, int* __restrict__ err, int pass, int big) {
  NODE_PROLOGUE
  if (!live) return;
  const unsigned mv = wmask[P.off + v];
  if (!mv) return;
  const ulonglong2* WN = reinterpret_cast<const ulonglong2*>(wnode) + P.off;
  const ulonglong2 wv = WN[v];
  const int A = (int)(unsigned)(wv.y & 0xffffffffull);
  if (A < 0) return;
  const int sz = bsize[P.off + A];
  if (sz <= big) return;
  constexpr int II[8] = {-1, 1, 0, 0, -1, 1, -1, 1};
  constexpr int JJ[8] = {0, 0, -1, 1, -1, -1, 1, 1};
  const double hv = __longlong_as_double((long long)wv.x);
  const int i = v % P.nx, j = v / P.nx, rank = wrank[P.off + v];
  const long long sbase = (long long)FSX_SLAB_CAP * (P.off + soff[P.off + A]);
  const unsigned C = (unsigned)big_cap(sz);
  unsigned long long* SW = reinterpret_cast<unsigned long long*>(sw);
  // the examined neighbours' records, then the first probe of each one's basin, all in flight together (a
  // slot once taken keeps its basin, so a first probe read early is either final or still -1)
  ulonglong2 wqs[8];
  int first[8];
#pragma unroll
  for (int k = 0; k < 8; k++)
    if ((mv >> k) & 1u) wqs[k] = WN[(j + JJ[k]) * P.nx + (i + II[k])];
#pragma unroll
  for (int k = 0; k < 8; k++)
    if ((mv >> k) & 1u) {
      const int Bs = (int)(unsigned)(wqs[k].y & 0xffffffffull);
      const int B = Bs >= 0 ? Bs : -(Bs + 1);
      first[k] = sh[sbase + (((unsigned)B * 2654435761u) & (C - 1u))];
    }
#pragma unroll
  for (int k = 0; k < 8; k++) {
    if (!((mv >> k) & 1u)) continue;
    const ulonglong2 wq = wqs[k];
    const int Bs = (int)(unsigned)(wq.y & 0xffffffffull);
    const int B = Bs >= 0 ? Bs : -(Bs + 1);
    const double hq = __longlong_as_double((long long)wq.x);
    const unsigned long long wk = dkey(hv > hq ? hv : hq);
    const int key = rank * 8 + k;
    // the slot of B (inserted on the first pass)
    unsigned s = ((unsigned)B * 2654435761u) & (C - 1u);
    long long slot = -1;
    for (unsigned pr = 0; pr < C; pr++, s = (s + 1u) & (C - 1u)) {
      int cur = pr == 0 ? first[k] : sh[sbase + s];
      if (pass == 0 && cur == -1) cur = atomicCAS(sh + sbase + s, -1, B);
      if (cur == B || (pass == 0 && cur == -1)) {
        slot = sbase + s;
        break;
      }
      if (cur == -1) break;  // (later passes: B is always present)
    }
    if (slot < 0) {
      atomicOr(err + c.inst, FSX_ERR_NCONN);
      continue;
    }
    // the atomics only where they can lower the value (most candidates of a busy pair cannot)
    if (pass == 0) {
      if (key < sp[slot]) atomicMin(sp + slot, key);
      if (wk < SW[slot]) atomicMin(SW + slot, wk);
    } else if (SW[slot] == wk && key < sq[slot]) {
      atomicMin(sq + slot, key);
    }
  }
}

// the table of a big basin in the walk's output form: entries ranked by first meeting
__global__ __launch_bounds__(256) void k_big_final(const InstParams* __restrict__ PP, const int2* __restrict__ biglist,
                                                   const int* __restrict__ bigctr, int bigcap, const int* __restrict__ bsize,
                                                   const int* __restrict__ soff, int* __restrict__ sB, int* __restrict__ sp,
                                                   int* __restrict__ sq, double* __restrict__ sw, int* __restrict__ sh,
                                                   const int* __restrict__ wseq, int* __restrict__ ecount,
                                                   int* __restrict__ err) {
  __shared__ int s_cnt;
  constexpr int II[8] = {-1, 1, 0, 0, -1, 1, -1, 1};
  constexpr int JJ[8] = {0, 0, -1, 1, -1, -1, 1, 1};
  const int n = min(bigctr[BIG_NBIG], bigcap);
  unsigned long long* SW = reinterpret_cast<unsigned long long*>(sw);
  for (int e = blockIdx.x; e < n; e += gridDim.x) {
    const int2 ab = biglist[e];
    const InstParams& P = PP[ab.x];
    const int size = bsize[P.off + ab.y];
    const long long sbase = (long long)FSX_SLAB_CAP * (P.off + soff[P.off + ab.y]);
    const int C = big_cap(size), room = FSX_SLAB_CAP * size - C;
    if (threadIdx.x == 0) s_cnt = 0;
    __syncthreads();
    // occupied slots, compacted after the table: B, first meeting, pass node, pass direction, weight
    for (int t = threadIdx.x; t < C; t += 256) {
      const int B = sh[sbase + t];
      if (B < 0) continue;
      const int x = atomicAdd(&s_cnt, 1);
      if (x >= room) continue;
      const long long o = sbase + C + x;
      sh[o] = B;
      sp[o] = sp[sbase + t];
      const int pk = sq[sbase + t];  // rank * 8 + k of the pass (a rank below size by k_big_rank)
      sB[o] = wseq[P.off + soff[P.off + ab.y] + min(max(pk >> 3, 0), size - 1)];
      sq[o] = pk & 7;
      const unsigned long long kw = SW[sbase + t];
      sw[o] = __longlong_as_double((long long)((kw & 0x8000000000000000ull) ? (kw & 0x7fffffffffffffffull) : ~kw));
    }
    __syncthreads();
    const int D = s_cnt;
    if (D > room || 2 * D > C) {
      if (threadIdx.x == 0) atomicOr(err + ab.x, FSX_ERR_NCONN);
      __syncthreads();
      continue;
    }
    const long long tb = sbase + C;
    for (int x = threadIdx.x; x < D; x += 256) {
      const int key = sp[tb + x];
      int t = 0;
      for (int y = 0; y < D; y++) t += sp[tb + y] < key;
      const int v = sB[tb + x], k = sq[tb + x];
      const int i = v % P.nx, jv = v / P.nx;
      sB[sbase + t] = sh[tb + x];
      sw[sbase + t] = sw[tb + x];
      sp[sbase + t] = v;
      sq[sbase + t] = (jv + JJ[k]) * P.nx + (i + II[k]);
    }
    if (threadIdx.x == 0) ecount[P.off + ab.y] = D;
    __syncthreads();
  }
}

void launch_lm_big(DevState& S, hipStream_t st) {
  const int nb = S.nchunks, bigcap = (int)(S.ntot / 64 + S.ninst + 1);
  k_big_pick<<<nb, FSX_CHUNK, 0, st>>>(S.P, S.chunks, S.wnode, S.rec, S.bsize, S.rchunk, S.rlist, S.walk_big, 0);
  k_big_rscan<<<1, FSX_WG, 0, st>>>(S.rchunk, nb, S.bigctr);
  k_big_pick<<<nb, FSX_CHUNK, 0, st>>>(S.P, S.chunks, S.wnode, S.rec, S.bsize, S.rchunk, S.rlist, S.walk_big, 1);
  k_big_seg<<<1024, 256, 0, st>>>(S.P, S.chunks, S.rlist, S.bigctr, S.wnode, S.wro, S.rjmp, S.err);
  for (int it = 0; it < BIG_ROUNDS; it++) k_big_jump<<<1024, 256, 0, st>>>(S.P, S.chunks, S.rlist, S.rjmp, S.bigctr, it);
  k_big_rank<<<nb, FSX_CHUNK, 0, st>>>(S.P, S.chunks, S.wnode, S.bsize, S.wro, S.rjmp, S.soff, S.wrank, S.wseq, S.err,
                                       S.walk_big);
  k_big_init<<<512, 256, 0, st>>>(S.P, S.biglist, S.bigctr, bigcap, S.bsize, S.soff, S.sB, S.sp, S.sq, S.sw, S.sh);
  for (int pass = 0; pass < 2; pass++)
    k_big_cand<<<nb, FSX_CHUNK, 0, st>>>(S.P, S.chunks, S.wnode, S.wmask, S.bsize, S.soff, S.wrank, S.sB, S.sp, S.sq,
                                         S.sw, S.sh, S.err, pass, S.walk_big);
  k_big_final<<<512, 256, 0, st>>>(S.P, S.biglist, S.bigctr, bigcap, S.bsize, S.soff, S.sB, S.sp, S.sq, S.sw, S.sh,
                                   S.wseq, S.ecount, S.err);
}

// edge offsets: basin order, open basins (except basin0) one virtual edge each
__global__ __launch_bounds__(FSX_WG) void k_edge_scan(const InstParams* __restrict__ PP, const int* __restrict__ nbasins,
                                                      const int* __restrict__ basin0, const int* __restrict__ outlet,
                                                      const uint8_t* __restrict__ bc, const int* __restrict__ ecount,
                                                      int* __restrict__ eoff, int* __restrict__ nconn, int* __restrict__ err) {
  __shared__ int lds[FSX_WG / 64];
  const int inst = blockIdx.x;
  const InstParams& P = PP[inst];
  if (!P.active || !P.run_spl) return;
  const int nb = nbasins[inst], b0 = basin0[inst];
  int carry = 0;
  for (int s = 0; s < nb; s += FSX_WG) {
    const int b = s + threadIdx.x;
    int x = 0;
    if (b < nb) {
      const bool open = bc[P.off + outlet[P.off + b]];
      x = open ? (b == b0 ? 0 : 1) : ecount[P.off + b];
    }
    int tot;
    const int incl = block_incl_scan<FSX_WG>(x, lds, tot);
    if (b < nb) eoff[P.off + b] = carry + incl - x;
    carry += tot;
  }
  if (threadIdx.x == 0) {
    if (b0 == INT_MAX) {
      atomicOr(err + inst, FSX_ERR_NO_OUTLET);
      nconn[inst] = 0;
    } else if (carry > FSX_EDGE_CAP * nb) {
      atomicOr(err + inst, FSX_ERR_NCONN);
      nconn[inst] = 0;
    } else {
      nconn[inst] = carry;
    }
  }
}

__global__ __launch_bounds__(FSX_CHUNK) void k_edge_fill(const InstParams* __restrict__ PP, const Chunk* __restrict__ chunks,
                                                         const int* __restrict__ nbasins, const int* __restrict__ basin0,
                                                         const int* __restrict__ nconn, const int* __restrict__ outlet,
                                                         const uint8_t* __restrict__ bc, const int* __restrict__ ecount,
                                                         const int* __restrict__ eoff, const int* __restrict__ soff,
                                                         const int* __restrict__ sB, const int* __restrict__ sp,
                                                         const int* __restrict__ sq, const double* __restrict__ sw,
                                                         int* __restrict__ cb0, int* __restrict__ cb1, int* __restrict__ cn0,
                                                         int* __restrict__ cn1, double* __restrict__ cw) {
  const Chunk c = chunks[chunk_id()];
  const InstParams& P = PP[c.inst];
  const int b = c.start + (int)threadIdx.x, lane = (int)threadIdx.x & 63;
  // every lane of the wave stays for the cooperative copy below (its shuffles need them all)
  const bool valid = P.active && P.run_spl && b < nbasins[c.inst] && nconn[c.inst] != 0;
  const long long eb = (long long)FSX_EDGE_CAP * P.off;
  int ne = 0, e0 = 0;
  long long sbase = 0;
  if (valid) {
    e0 = eoff[P.off + b];
    if (bc[P.off + outlet[P.off + b]]) {
      if (b != basin0[c.inst]) {
        cb0[eb + e0] = basin0[c.inst]; cb1[eb + e0] = b;
        cn0[eb + e0] = -1; cn1[eb + e0] = -1;
        cw[eb + e0] = -1.0e10;
      }
    } else {
      ne = ecount[P.off + b];
      sbase = (long long)FSX_SLAB_CAP * (P.off + soff[P.off + b]);
    }
  }
  // The wave's closed basins' passes, copied from their slabs one edge per lane: the wave's edges are
  // consecutive in the output (basin order), so the stores are coalesced.  Edge t of the wave belongs
  // to the last lane whose exclusive prefix of counts is <= t.
  const int incl = wave_incl_scan(ne);
  const int total = __shfl(incl, 63, 64), excl = incl - ne;
  const int sb_lo = (int)(sbase & 0xffffffffll), sb_hi = (int)(sbase >> 32);
  for (int t0 = 0; t0 < total; t0 += 64) {
    const int t = t0 + lane;
    int o = 0;
#pragma unroll
    for (int step = 32; step >= 1; step >>= 1)
      if (__shfl(excl, o + step, 64) <= t) o += step;
    const int k = t - __shfl(excl, o, 64);
    const int oe0 = __shfl(e0, o, 64);
    const long long osb = ((long long)__shfl(sb_hi, o, 64) << 32) | (unsigned)__shfl(sb_lo, o, 64);
    if (t < total) {
      const long long d = eb + oe0 + k, q = osb + k;
      cb0[d] = c.start + (int)(threadIdx.x & ~63u) + o;  // the owner lane's basin
      cb1[d] = sB[q];
      cn0[d] = sp[q];
      cn1[d] = sq[q];
      cw[d] = sw[q];
    }
  }
}

// ---- 3. loc_min_3_indexx order --------------------------------------------------------------------
// One partition step of the Numerical Recipes quicksort on positions [l, ir] (0-based), computed by a
// group of NT threads.  After the median of three, with pivot a = key[l]:
//   L = ascending positions p in [l+2, ir]   with key[p] >= a   (stops of the upward scan)
//   R = descending positions q in [l+1, ir-1] with key[q] <= a   (stops of the downward scan)
// The scans swap L[t] with R[t] exactly for t = 1..T, T = #{t : L[t] <= R[t]} (a self-swap at
// L[T] == R[T] included); the crossing indices are then
//   self-swap: i* = L[T]+1, j* = L[T]-1;  else i* = min(L[T+1], R[T]), j* = max(R[T+1], L[T])
// with R[0] = +inf and L[0] = l+1; the pivot moves to j* and the element there to l.  Subranges
// [l, j*-1] and [i*, ir] follow.  Segments of <= 7 keys get the reference's stable insertion sort.
// Every step is checked against the sequential routine on tie-heavy inputs in the tests.
#define NR_ELT 8
#define NR_BIG 2048

template <int NT>
__device__ __forceinline__ void grp_sync() {
  if (NT == 64) {
    __threadfence_block();
    __builtin_amdgcn_wave_barrier();
  } else {
    __syncthreads();
  }
}

template <int NT>
__device__ __forceinline__ int grp_scan(int x, int* lds, int& total) {
  if (NT == 64) {
    const int v = wave_incl_scan(x);
    total = __shfl(v, 63, 64);
    return v;
  }
  return block_incl_scan<NT>(x, lds, total);
}

template <int NT>
__device__ __forceinline__ int grp_rank() {
  return NT == 64 ? (int)(threadIdx.x & 63) : (int)threadIdx.x;
}

__device__ __forceinline__ void kswap(double* K, int* I, int p, int q) {
  const double tk = K[p]; K[p] = K[q]; K[q] = tk;
  const int ti = I[p]; I[p] = I[q]; I[q] = ti;
}

__device__ __forceinline__ void nr_insertion(double* K, int* I, int l, int ir) {
  for (int j = l + 1; j <= ir; j++) {
    const double a = K[j];
    const int ia = I[j];
    int i = j - 1;
    while (i >= l && K[i] > a) {
      K[i + 1] = K[i];
      I[i + 1] = I[i];
      i--;
    }
    K[i + 1] = a;
    I[i + 1] = ia;
  }
}

// returns (istar, jstar) in every thread of the group; sh must be per-group shared scratch (>= 4 ints)
template <int NT>
__device__ __forceinline__ void nr_partition(double* K, int* I, int* LA, int* RA, int l, int ir, int* lds, int* sh, double* shd,
                             int& istar, int& jstar) {
  const int rk = grp_rank<NT>();
  if (rk == 0) {
    const int k = (l + ir) >> 1;
    kswap(K, I, k, l + 1);
    if (K[l + 1] > K[ir]) kswap(K, I, l + 1, ir);
    if (K[l] > K[ir]) kswap(K, I, l, ir);
    if (K[l + 1] > K[l]) kswap(K, I, l + 1, l);
    shd[0] = K[l];
  }
  grp_sync<NT>();
  const double a = shd[0];
  int cL = 0, cR = 0;
  const int ln = rk & 63;
  const unsigned long long below = (1ull << ln) - 1;
  // a chunk of NT * NR_ELT positions: wave w takes NR_ELT runs of 64 consecutive positions, so every
  // load and list store of a wave is contiguous; position order inside the wave is (run, lane), the
  // ballots give each flagged lane its rank, and one scan over the waves places them in the chunk
  for (int cb = l + 1; cb <= ir; cb += NT * NR_ELT) {
    const int p0 = cb + (rk >> 6) * (64 * NR_ELT) + ln;
    unsigned long long ml[NR_ELT], mr[NR_ELT];
    int nl = 0, nr = 0;
#pragma unroll
    for (int e = 0; e < NR_ELT; e++) {
      const int p = p0 + 64 * e;
      bool fl = false, fr = false;
      if (p <= ir) {
        const double kp = K[p];
        fl = p >= l + 2 && kp >= a;
        fr = p <= ir - 1 && kp <= a;
      }
      ml[e] = __ballot(fl);
      mr[e] = __ballot(fr);
      nl += __popcll(ml[e]);
      nr += __popcll(mr[e]);
    }
    const int own = nl | (nr << 16);
    int tot;
    const int incl = grp_scan<NT>(ln == 0 ? own : 0, lds, tot);
    int rl = cL + (incl & 0xffff) - nl, rr = cR + (incl >> 16) - nr;
#pragma unroll
    for (int e = 0; e < NR_ELT; e++) {
      if ((ml[e] >> ln) & 1) LA[l + rl + __popcll(ml[e] & below)] = p0 + 64 * e;
      if ((mr[e] >> ln) & 1) RA[l + rr + __popcll(mr[e] & below)] = p0 + 64 * e;
      rl += __popcll(ml[e]);
      rr += __popcll(mr[e]);
    }
    cL += tot & 0xffff;
    cR += tot >> 16;
  }
  grp_sync<NT>();
  const int totR = cR, mm = cL < cR ? cL : cR;
  int cnt = 0;
  for (int t = 1 + rk; t <= mm; t += NT) cnt += (LA[l + t - 1] <= RA[l + totR - t]) ? 1 : 0;
  int T;
  grp_scan<NT>(cnt, lds, T);
  for (int t = 1 + rk; t <= T; t += NT) {
    const int p = LA[l + t - 1], q = RA[l + totR - t];
    if (p != q) kswap(K, I, p, q);
  }
  grp_sync<NT>();
  if (rk == 0) {
    int is, js;
    if (T >= 1 && LA[l + T - 1] == RA[l + totR - T]) {
      is = LA[l + T - 1] + 1;
      js = LA[l + T - 1] - 1;
    } else {
      is = LA[l + T];
      js = RA[l + totR - T - 1];
      if (T >= 1) {
        const int rt = RA[l + totR - T], lt = LA[l + T - 1];
        is = is < rt ? is : rt;
        js = js > lt ? js : lt;
      }
    }
    const double pk = K[l];
    const int pi = I[l];
    K[l] = K[js]; I[l] = I[js];
    K[js] = pk; I[js] = pi;
    sh[0] = is;
    sh[1] = js;
  }
  grp_sync<NT>();
  istar = sh[0];
  jstar = sh[1];
  grp_sync<NT>();
}

// Large segments, one workgroup per landscape; segments of <= NR_BIG keys go to a list.
// Ranges of more than NR_BIG keys are partitioned level by level: launch r partitions every big
// range of level r, DevState::nr_gx workgroups per landscape taking ranges in turn (each range by one
// workgroup; ranges are disjoint, so the order of processing does not matter), and appends the
// children: big ones to level r+1's list, the others to the segment list of k_nr_small.  Level lists
// alternate between two halves of the int view of cw (free once k_nr_init has copied the keys);
// nrst[NR_ST * inst + r] counts level r's ranges, nrst[NR_ST * inst + NR_ST - 1] the segments.  The
// number of levels launched grows with the largest landscape (nr_levels); ranges still big after them
// are finished by k_nr_big, one at a time, depth first with a stack in LDS.
#define NR_ST 32  // per-landscape counters: levels 0..NR_ST-2, segments at NR_ST-1

__device__ __forceinline__ int* nr_list(double* cw, long long eb, int n, int level) {
  return reinterpret_cast<int*>(cw) + 2 * eb + ((level & 1) ? n + 2 : 0);
}

// The keys and identities of the sort, copied out of cw: gridDim.x workgroups per landscape (blockIdx.y),
// each thread taking 4 slots a stride apart at a time.  A pass height that is not a number (elevations that
// are not) would break the partition scans' sentinel (every comparison with it is false): it is sorted as
// +infinity and the step reports the landscape.
__global__ __launch_bounds__(FSX_CHUNK) void k_nr_copy(const InstParams* __restrict__ PP, const int* __restrict__ nconn,
                                                       const double* __restrict__ cw, double* __restrict__ skey,
                                                       int* __restrict__ sid, int* __restrict__ err) {
  const int inst = blockIdx.y;
  const InstParams& P = PP[inst];
  if (!P.active || !P.run_spl) return;
  const int n = nconn[inst];
  const long long eb = (long long)FSX_EDGE_CAP * P.off;
  const int stride = (int)gridDim.x * FSX_CHUNK;
  bool nan = false;
  for (int p0 = (int)blockIdx.x * FSX_CHUNK + (int)threadIdx.x; p0 < n; p0 += 4 * stride) {
    double kv[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
      const int p = p0 + k * stride;
      kv[k] = p < n ? cw[eb + p] : 0.0;
    }
#pragma unroll
    for (int k = 0; k < 4; k++) {
      const int p = p0 + k * stride;
      if (p >= n) continue;
      nan |= kv[k] != kv[k];
      skey[eb + p] = kv[k] == kv[k] ? kv[k] : __builtin_huge_val();
      sid[eb + p] = p;
    }
  }
  if (nan) atomicOr(err + inst, FSX_ERR_NONFINITE);
}

// The sort's start state, once k_nr_copy has taken the keys out of cw (the level lists overwrite it).
__global__ __launch_bounds__(64) void k_nr_init(const InstParams* __restrict__ PP, const int* __restrict__ nconn,
                                                double* __restrict__ cw, int* __restrict__ segs,
                                                int* __restrict__ nrst) {
  const int inst = blockIdx.x;
  const InstParams& P = PP[inst];
  if (!P.active || !P.run_spl || threadIdx.x != 0) return;
  const int n = nconn[inst];
  const long long eb = (long long)FSX_EDGE_CAP * P.off;
  int* st = nrst + NR_ST * inst;
  for (int r = 0; r < NR_ST; r++) st[r] = 0;
  if (n > NR_BIG) {
    int* L0 = nr_list(cw, eb, n, 0);
    L0[0] = 0;
    L0[1] = n - 1;
    st[0] = 1;
  } else if (n >= 2) {
    segs[eb] = 0;
    segs[eb + 1] = n - 1;
    st[NR_ST - 1] = 1;
  }
}

__global__ __launch_bounds__(FSX_WG) void k_nr_level(const InstParams* __restrict__ PP, const int* __restrict__ nconn,
                                                     double* __restrict__ cw, double* __restrict__ skey,
                                                     int* __restrict__ sid, int* __restrict__ Lasc,
                                                     int* __restrict__ Rasc, int* __restrict__ segs,
                                                     int* __restrict__ nrst, int level) {
  __shared__ int lds[FSX_WG / 64];
  __shared__ int sh[4];
  __shared__ double shd[1];
  const int inst = blockIdx.y;
  const InstParams& P = PP[inst];
  if (!P.active || !P.run_spl) return;
  int* st = nrst + NR_ST * inst;
  const int nr = st[level];
  if ((int)blockIdx.x >= nr) return;
  const int n = nconn[inst];
  const long long eb = (long long)FSX_EDGE_CAP * P.off;
  const int* Lc = nr_list(cw, eb, n, level);
  int* Ln = nr_list(cw, eb, n, level + 1);
  for (int t = blockIdx.x; t < nr; t += gridDim.x) {
    const int l = Lc[2 * t], ir = Lc[2 * t + 1];
    int is, js;
    nr_partition<FSX_WG>(skey + eb, sid + eb, Lasc + eb, Rasc + eb, l, ir, lds, sh, shd, is, js);
    if (threadIdx.x == 0) {
      const int segl[2][2] = {{l, js - 1}, {is, ir}};
      for (int c = 0; c < 2; c++) {
        const int sl = segl[c][0], sr = segl[c][1], sz = sr - sl + 1;
        if (sz > NR_BIG) {
          const int k = atomicAdd(st + level + 1, 1);
          Ln[2 * k] = sl;
          Ln[2 * k + 1] = sr;
        } else if (sz >= 2) {
          const int k = atomicAdd(st + NR_ST - 1, 1);
          segs[eb + 2 * k] = sl;
          segs[eb + 2 * k + 1] = sr;
        }
      }
    }
    __syncthreads();
  }
}

// big ranges left after the level launches (none for keys of ordinary spread): one workgroup, each
// range depth first with a stack in LDS (the stack holds at most two ranges per partition depth)
__global__ __launch_bounds__(FSX_WG) void k_nr_big(const InstParams* __restrict__ PP, const int* __restrict__ nconn,
                                                   double* __restrict__ cw, double* __restrict__ skey,
                                                   int* __restrict__ sid, int* __restrict__ Lasc, int* __restrict__ Rasc,
                                                   int* __restrict__ segs, int* __restrict__ nrst, int* __restrict__ nsegs,
                                                   int* __restrict__ err, int nlevels) {
  __shared__ int lds[FSX_WG / 64];
  __shared__ int sh[4];
  __shared__ double shd[1];
  __shared__ int bst[2 * 128];
  __shared__ int s_nb, s_ns;
  const int inst = blockIdx.x;
  const InstParams& P = PP[inst];
  if (!P.active || !P.run_spl) return;
  const int n = nconn[inst];
  const long long eb = (long long)FSX_EDGE_CAP * P.off;
  int* st = nrst + NR_ST * inst;
  double* K = skey + eb;
  int* I = sid + eb;
  int* LA = Lasc + eb;
  int* RA = Rasc + eb;
  int* SG = segs + eb;
  const int left = st[nlevels];
  const int* L = nr_list(cw, eb, n, nlevels);
  if (threadIdx.x == 0) s_ns = st[NR_ST - 1];
  for (int t = 0; t < left; t++) {
  if (threadIdx.x == 0) {
    bst[0] = L[2 * t];
    bst[1] = L[2 * t + 1];
    s_nb = 1;
  }
  __syncthreads();
  while (s_nb > 0) {
    const int l = bst[2 * (s_nb - 1)], ir = bst[2 * (s_nb - 1) + 1];
    __syncthreads();
    int is, js;
    nr_partition<FSX_WG>(K, I, LA, RA, l, ir, lds, sh, shd, is, js);
    if (threadIdx.x == 0) {
      s_nb--;
      const int a0 = l, a1 = js - 1, b0 = is, b1 = ir;
      const int segl[2][2] = {{a0, a1}, {b0, b1}};
      for (int s = 0; s < 2; s++) {
        const int sl = segl[s][0], sr = segl[s][1], sz = sr - sl + 1;
        if (sz > NR_BIG) {
          if (s_nb < 128) { bst[2 * s_nb] = sl; bst[2 * s_nb + 1] = sr; s_nb++; }
          else atomicOr(err + inst, FSX_ERR_SORTSTACK);
        } else if (sz >= 2) {
          SG[2 * s_ns] = sl; SG[2 * s_ns + 1] = sr; s_ns++;
        }
      }
    }
    __syncthreads();
  }
  }
  if (threadIdx.x == 0) nsegs[inst] = s_ns;
}

// Segments of <= NR_BIG keys, one NR_SMALL_WG-thread block per segment, entirely in LDS.  Sub-ranges are
// disjoint, so their order of processing does not matter: each round either partitions one range of
// more than NR_MID keys with the whole block, or up to one range of NR_TINY+1..NR_MID keys per
// wave; ranges of at most NR_TINY keys are finished at the end, one per thread, by the
// sequential routine itself (median of three, the scanning partition, insertion sort below 7).
#define NR_MID 512
#define NR_TINY 32
#define NR_QCAP 512
#define NR_SMALL_WG 512  // threads per segment block: its waves partition mid ranges side by side

// loc_min_3_indexx on K/I[l..ir] (LDS), one thread, as the reference's loop (:955-1044)
__device__ __forceinline__ void nr_serial(double* K, int* I, int l, int ir) {
  int stk[16];
  int js = 0;
  for (;;) {
    if (ir - l < 7) {
      nr_insertion(K, I, l, ir);
      if (js == 0) return;
      ir = stk[--js];
      l = stk[--js];
      continue;
    }
    const int k = (l + ir) >> 1;
    kswap(K, I, k, l + 1);
    if (K[l + 1] > K[ir]) kswap(K, I, l + 1, ir);
    if (K[l] > K[ir]) kswap(K, I, l, ir);
    if (K[l + 1] > K[l]) kswap(K, I, l + 1, l);
    int i = l + 1, j = ir;
    const double a = K[l];
    const int ia = I[l];
    for (;;) {
      do i++; while (K[i] < a);
      do j--; while (K[j] > a);
      if (j < i) break;
      kswap(K, I, i, j);
    }
    K[l] = K[j]; I[l] = I[j];
    K[j] = a; I[j] = ia;
    // push the larger sub-range, continue with the smaller (:1033-1042): at most log2(32) pending
    if (ir - i + 1 >= j - l) {
      stk[js++] = i;
      stk[js++] = ir;
      ir = j - 1;
    } else {
      stk[js++] = l;
      stk[js++] = j - 1;
      l = i;
    }
  }
}

__global__ __launch_bounds__(NR_SMALL_WG) void k_nr_small(const InstParams* __restrict__ PP, const int* __restrict__ nsegs,
                                                  int* __restrict__ segctr, double* __restrict__ skey,
                                                  int* __restrict__ sid, const int* __restrict__ segs,
                                                  int* __restrict__ err) {
  __shared__ double sK[NR_BIG];
  __shared__ int sI[NR_BIG];
  __shared__ int sLA[NR_BIG];
  __shared__ int sRA[NR_BIG];
  __shared__ int q[2 * NR_QCAP];      // pending ranges (more than NR_TINY keys)
  // tiny ranges as 16-bit positions (segments hold at most NR_BIG = 2048 keys): with them the block
  // stays under a third of the 160 KB LDS, so three blocks share a compute unit instead of two
  __shared__ short tq[2 * NR_BIG / 2];
  static_assert(NR_BIG <= 32767, "tiny-range positions are 16-bit");
  constexpr int NWV = NR_SMALL_WG / 64;
  __shared__ int lds[NWV];
  __shared__ int sh[NWV + 1][4];
  __shared__ double shd[NWV + 1][1];
  __shared__ int s_seg, s_nq, s_nt, s_take, s_ovf;
  const int inst = blockIdx.y;
  const InstParams& P = PP[inst];
  if (!P.active || !P.run_spl) return;
  const int tid = threadIdx.x, w = tid >> 6;
  const long long eb = (long long)FSX_EDGE_CAP * P.off;
  double* K = skey + eb;
  int* I = sid + eb;
  const int* SG = segs + eb;
  const int ns = nsegs[inst];
  for (;;) {
    if (tid == 0) s_seg = atomicAdd(segctr + inst, 1);
    __syncthreads();
    const int sg = s_seg;
    if (sg >= ns) break;
    const int g0 = SG[2 * sg], g1 = SG[2 * sg + 1], n = g1 - g0 + 1;
    for (int p = tid; p < n; p += NR_SMALL_WG) { sK[p] = K[g0 + p]; sI[p] = I[g0 + p]; }
    if (tid == 0) {
      s_nq = 0; s_nt = 0; s_ovf = 0;
      if (n > NR_TINY) { q[0] = 0; q[1] = n - 1; s_nq = 1; }
      else if (n >= 2) { tq[0] = 0; tq[1] = (short)(n - 1); s_nt = 1; }
    }
    __syncthreads();
    while (s_nq > 0) {
      // a round: the top range alone if it is big, else up to four mid ranges
      const int top = s_nq - 1;
      const bool big = q[2 * top + 1] - q[2 * top] + 1 > NR_MID;
      int is = 0, js = 0, l = 0, ir = -1;
      if (big) {
        l = q[2 * top]; ir = q[2 * top + 1];
        __syncthreads();
        if (tid == 0) { s_nq = top; s_take = 1; }
        nr_partition<NR_SMALL_WG>(sK, sI, sLA, sRA, l, ir, lds, sh[NWV], shd[NWV], is, js);
      } else {
        int take = 0;
        for (int t = top; t >= 0 && take < NWV; t--) {
          if (q[2 * t + 1] - q[2 * t] + 1 > NR_MID) break;
          take++;
        }
        if (w < take) { l = q[2 * (top - w)]; ir = q[2 * (top - w) + 1]; }
        __syncthreads();
        if (tid == 0) { s_nq = top + 1 - take; s_take = take; }
        if (w < take) nr_partition<64>(sK, sI, sLA, sRA, l, ir, nullptr, sh[w], shd[w], is, js);
      }
      __syncthreads();
      const bool lead = big ? tid == 0 : ((tid & 63) == 0 && w < s_take);
      if (lead) {
        const int r[2][2] = {{l, js - 1}, {is, ir}};
        for (int c = 0; c < 2; c++) {
          const int a0 = r[c][0], a1 = r[c][1], sz = a1 - a0 + 1;
          if (sz > NR_TINY) {
            const int slot = atomicAdd(&s_nq, 1);
            if (slot < NR_QCAP) { q[2 * slot] = a0; q[2 * slot + 1] = a1; } else s_ovf = 1;
          } else if (sz >= 2) {
            const int slot = atomicAdd(&s_nt, 1);
            tq[2 * slot] = (short)a0; tq[2 * slot + 1] = (short)a1;
          }
        }
      }
      __syncthreads();
      if (s_ovf) break;
    }
    if (s_ovf) {
      if (tid == 0) atomicOr(err + inst, FSX_ERR_SORTSTACK);
      __syncthreads();
      continue;
    }
    for (int t = tid; t < s_nt; t += NR_SMALL_WG) nr_serial(sK, sI, tq[2 * t], tq[2 * t + 1]);
    __syncthreads();
    for (int p = tid; p < n; p += NR_SMALL_WG) { K[g0 + p] = sK[p]; I[g0 + p] = sI[p]; }
    __syncthreads();
  }
}

__global__ __launch_bounds__(FSX_CHUNK) void k_nr_pos(const InstParams* __restrict__ PP, const int* __restrict__ nconn,
                                                      const int* __restrict__ sid, int* __restrict__ npos) {
  const int inst = blockIdx.y;
  const InstParams& P = PP[inst];
  if (!P.active || !P.run_spl) return;
  const long long eb = (long long)FSX_EDGE_CAP * P.off;
  const int n = nconn[inst];
  for (int p = blockIdx.x * FSX_CHUNK + threadIdx.x; p < n; p += gridDim.x * FSX_CHUNK) npos[eb + sid[eb + p]] = p;
}

// ---- 4. Boruvka -----------------------------------------------------------------------------------
// Minimum spanning forest keyed by the edge's rank in the sorted order (npos): every component
// picks its lowest-ranked outgoing edge, components merge along them, repeat.  With a strict
// total order the tree is unique, so it is Kruskal's in that order (mst_kruskal, :333-385).  The
// component id of every basin lives in LDS when the landscape's basins fit (BORUVKA_LDS), and the
// edge scan handles BORUVKA_UNROLL edges per thread at a time so their loads overlap.
#define BORUVKA_UNROLL 4
static int g_boruvka_lds = 0;  // component ids held in LDS (ints), set from the device's LDS size

// The edge scan of a round reads the edges still joining two components at the start of the
// previous round and writes those still joining two now (components only merge, so an edge inside a
// component never matters again); the lists shrink with the components.  List k is (a, c, rank)
// triples; list 0 is the sorted edge set itself.
struct EdgeList {
  int *a, *c, *k;
};

// The components still roots at the start of a round are kept as a list too (ping-pong between
// LC0 and LC1; round 0 is every basin): a round's per-component passes (clearing the minima, the
// merge, the pointer jumping over the merge pointers and the id update) touch only those, and only
// their ids can appear in the next round's edge list.
//
// The first BOR_SPLIT rounds, whose edge lists are long, run as separate launches: the edge scan
// with DevState::bor_gx workgroups per landscape (component ids in global memory), the rest of the
// round with one.  k_boruvka then finishes the remaining rounds in one workgroup with the component
// ids in LDS.  Per-landscape state between the launches: bstate[4*inst + 0..3] = live edges, live
// components, done, edges kept by the running scan.
#define BOR_SPLIT 4

// One workgroup's share of a round's edge scan: blocks of FSX_WG*U edges starting at block g,
// every G-th.  cnt: counter of the edges kept (LDS or global).
__device__ __forceinline__ void bor_scan(int round, int ne, int g, int G, const int* E0, const int* E1, const int* NP,
                                         const int* C, unsigned long long* BST, EdgeList in, EdgeList out, int* cnt) {
  constexpr int U = BORUVKA_UNROLL;
  const int tid = threadIdx.x, lane = tid & 63;
  // lowest-ranked live edge of every component, with the component at its other end:
  // (rank << 32 | other) -- ranks are distinct, so the minimum names one edge.  The U edges of a
  // thread are loaded unconditionally (index clamped) so that their loads issue together.
  const int* pa = round ? in.a : E0;
  const int* pc = round ? in.c : E1;
  const int* pk = round ? in.k : NP;
  for (int e0 = g * FSX_WG * U + tid; e0 < ne; e0 += G * FSX_WG * U) {
    int a[U], c[U], k[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int e = e0 + u * FSX_WG;
      const int ee = e < ne ? e : e0;
      a[u] = pa[ee];
      c[u] = pc[ee];
      k[u] = pk[ee];
    }
#pragma unroll
    for (int u = 0; u < U; u++) {
      a[u] = C[a[u]];
      c[u] = C[c[u]];
    }
    bool live[U];
    unsigned long long m[U];
    int tot = 0;
#pragma unroll
    for (int u = 0; u < U; u++) {
      live[u] = e0 + u * FSX_WG < ne && a[u] != c[u];
      const unsigned long long kr = (unsigned long long)(unsigned)k[u] << 32;
      // the list keeps a component's edges together (round 0: a basin's passes are consecutive), so
      // lanes of one run of equal a first take the minimum over windows of 8 lanes of the run, and one
      // lane per window issues the atomic for that side
      const int key = live[u] ? a[u] : -1 - lane;
      unsigned long long va = live[u] ? (kr | (unsigned)c[u]) : ~0ull;
#pragma unroll
      for (int off = 1; off < 8; off <<= 1) {
        const unsigned long long o = __shfl_down(va, off, 64);
        const int ok = __shfl_down(key, off, 64);
        if (lane + off < 64 && ok == key && o < va) va = o;
      }
      const int kprev = __shfl_up(key, 1, 64);
      const unsigned long long heads = __ballot(lane == 0 || kprev != key);
      const int rs = 63 - __clzll((long long)(heads & (lane == 63 ? ~0ull : ((2ull << lane) - 1ull))));
      if (live[u]) {
        if (((lane - rs) & 7) == 0) atomicMin(BST + a[u], va);
        atomicMin(BST + c[u], kr | (unsigned)a[u]);
      }
      m[u] = __ballot(live[u]);
      tot += __popcll(m[u]);
    }
    // keep the live edges (as component ids) for the next round: one atomic per wave
    int base = 0;
    if (lane == 0 && tot) base = atomicAdd(cnt, tot);
    base = __shfl(base, 0, 64);
#pragma unroll
    for (int u = 0; u < U; u++) {
      if (live[u]) {
        const int r = base + __popcll(m[u] & ((1ull << lane) - 1ull));
        out.a[r] = a[u];
        out.c[r] = c[u];
        out.k[r] = k[u];
      }
      base += __popcll(m[u]);
    }
  }
}

// The rest of a round, one workgroup: every component joins the other end of its edge (the smaller
// id of a mutual pair stays root), the merge pointers are jumped to their roots, the round's
// components take their new ids and those still roots form the next list (their minima cleared
// when clear_next).  Returns false when no component had an edge (the forest is complete).
__device__ __forceinline__ bool bor_merge(int round, int nl, const int* SI, int* C, unsigned long long* BST, int* BN,
                                          uint8_t* IT, const int* Lin, int* Lout, bool clear_next, int* s_any,
                                          int* s_chg, int* s_ncomp) {
  constexpr int U = BORUVKA_UNROLL;
  constexpr unsigned long long NONE = ~0ull;
  const int tid = threadIdx.x, lane = tid & 63;
  auto comp_at = [&](int i) { return round ? Lin[i] : i; };
  if (tid == 0) { *s_any = 0; *s_ncomp = 0; }
  __syncthreads();
  for (int i0 = tid; i0 < nl; i0 += FSX_WG * U) {
    int cc[U];
    unsigned long long v[U], w[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int i = i0 + u * FSX_WG;
      cc[u] = i < nl ? comp_at(i) : -1;
    }
#pragma unroll
    for (int u = 0; u < U; u++) v[u] = cc[u] >= 0 ? BST[cc[u]] : NONE;
#pragma unroll
    for (int u = 0; u < U; u++) w[u] = v[u] != NONE ? BST[(int)(v[u] & 0xffffffffu)] : NONE;
#pragma unroll
    for (int u = 0; u < U; u++) {
      if (cc[u] < 0) continue;
      if (v[u] == NONE) { BN[cc[u]] = cc[u]; continue; }
      const int kk = (int)(v[u] >> 32), o = (int)(v[u] & 0xffffffffu);
      IT[SI[kk]] = 1;
      BN[cc[u]] = (int)(w[u] >> 32) == kk ? (cc[u] < o ? cc[u] : o) : o;
      *s_any = 1;
    }
  }
  __syncthreads();
  if (!*s_any) return false;
  for (;;) {
    if (tid == 0) *s_chg = 0;
    __syncthreads();
    for (int i0 = tid; i0 < nl; i0 += FSX_WG * U) {
      int cc[U], x[U], y[U];
#pragma unroll
      for (int u = 0; u < U; u++) {
        const int i = i0 + u * FSX_WG;
        cc[u] = i < nl ? comp_at(i) : -1;
      }
#pragma unroll
      for (int u = 0; u < U; u++) x[u] = cc[u] >= 0 ? BN[cc[u]] : -1;
#pragma unroll
      for (int u = 0; u < U; u++) y[u] = x[u] >= 0 ? BN[x[u]] : -1;
#pragma unroll
      for (int u = 0; u < U; u++)
        if (x[u] != y[u]) { BN[cc[u]] = y[u]; *s_chg = 1; }
    }
    __syncthreads();
    if (!*s_chg) break;
  }
  for (int i0 = tid; i0 < nl; i0 += FSX_WG) {
    const int cc = comp_at(i0);
    const int r = BN[cc];
    C[cc] = r;
    const bool root = r == cc;
    const unsigned long long m = __ballot(root);
    const int cnt = __popcll(m);
    int base = 0;
    if (lane == 0 && cnt) base = atomicAdd(s_ncomp, cnt);
    base = __shfl(base, 0, 64);
    if (root) {
      Lout[base + __popcll(m & ((1ull << lane) - 1ull))] = cc;
      if (clear_next) BST[cc] = NONE;
    }
  }
  __syncthreads();
  return true;
}

// Rounds round0.. of one landscape in one workgroup (edges ne, components nl at round0).
template <bool INLDS>
__device__ __forceinline__ void boruvka_body(int round0, int ne, int nl, const int* E0, const int* E1, const int* NP,
                                             const int* SI, int* C, unsigned long long* BST, int* BN, uint8_t* IT,
                                             EdgeList LA, EdgeList LB, int* LC0, int* LC1, int* s_any, int* s_chg,
                                             int* s_cnt, int* s_ncomp) {
  constexpr unsigned long long NONE = ~0ull;
  const int tid = threadIdx.x;
  for (int round = round0; round < 64; round++) {
    const EdgeList out = (round & 1) ? LB : LA;
    const EdgeList in = (round & 1) ? LA : LB;  // read from round 1 on
    const int* Lin = (round & 1) ? LC1 : LC0;    // live components (read from round 1 on)
    int* Lout = (round & 1) ? LC0 : LC1;
    for (int i = tid; i < nl; i += FSX_WG) BST[round ? Lin[i] : i] = NONE;
    if (tid == 0) *s_cnt = 0;
    __syncthreads();
    bor_scan(round, ne, 0, 1, E0, E1, NP, C, BST, in, out, s_cnt);
    __syncthreads();
    ne = *s_cnt;
    if (!bor_merge(round, nl, SI, C, BST, BN, IT, Lin, Lout, false, s_any, s_chg, s_ncomp)) break;
    nl = *s_ncomp;
  }
}

struct BorArgs {
  const InstParams* PP;
  const int *nbasins, *nconn, *cb0, *cb1, *npos, *sid;
  int* comp;
  unsigned long long* best;
  int* bnew;
  uint8_t* intree;
  int *ntree, *Lasc, *Rasc, *segs;
  double *cw, *skey;
  int *lc0, *lc1, *bstate;
};

struct BorView {  // one landscape's slices
  long long eb;
  const int *E0, *E1, *NP, *SI;
  int *C, *BN, *LC0, *LC1;
  unsigned long long* BST;
  uint8_t* IT;
  EdgeList LA, LB;
};

__device__ __forceinline__ BorView bor_view(const BorArgs& A, const InstParams& P) {
  BorView V;
  V.eb = (long long)FSX_EDGE_CAP * P.off;
  V.E0 = A.cb0 + V.eb;
  V.E1 = A.cb1 + V.eb;
  V.NP = A.npos + V.eb;
  V.SI = A.sid + V.eb;
  V.C = A.comp + P.off;
  V.BN = A.bnew + P.off;
  V.LC0 = A.lc0 + P.off;
  V.LC1 = A.lc1 + P.off;
  V.BST = A.best + P.off;
  V.IT = A.intree + V.eb;
  const long long ecap = (long long)FSX_EDGE_CAP * P.nn;
  V.LA = {A.Lasc + V.eb, A.Rasc + V.eb, A.segs + V.eb};
  int* cwi = reinterpret_cast<int*>(A.cw) + 2 * V.eb;
  V.LB = {cwi, cwi + ecap, reinterpret_cast<int*>(A.skey) + 2 * V.eb};
  return V;
}

// set-up: component ids, cleared minima and tree flags, the state of round 0
__global__ __launch_bounds__(FSX_WG) void k_bor_start(BorArgs A) {
  const int inst = blockIdx.x;
  const InstParams& P = A.PP[inst];
  if (!P.active || !P.run_spl) return;
  const int nb = A.nbasins[inst], n = A.nconn[inst], tid = threadIdx.x;
  const BorView V = bor_view(A, P);
  for (int b = tid; b < nb; b += FSX_WG) {
    V.C[b] = b;
    V.BST[b] = ~0ull;
  }
  for (int e = tid; e < n; e += FSX_WG) V.IT[e] = 0;
  if (tid == 0) {
    int* st = A.bstate + 4 * inst;
    st[0] = n;
    st[1] = nb;
    st[2] = n > 0 ? 0 : 1;
    st[3] = 0;
  }
}

// a split round's edge scan: gridDim.x workgroups per landscape
__global__ __launch_bounds__(FSX_WG) void k_bor_scan(BorArgs A, int round) {
  const int inst = blockIdx.y;
  const InstParams& P = A.PP[inst];
  if (!P.active || !P.run_spl) return;
  int* st = A.bstate + 4 * inst;
  if (st[2]) return;
  const BorView V = bor_view(A, P);
  const EdgeList out = (round & 1) ? V.LB : V.LA;
  const EdgeList in = (round & 1) ? V.LA : V.LB;
  bor_scan(round, st[0], blockIdx.x, gridDim.x, V.E0, V.E1, V.NP, V.C, V.BST, in, out, st + 3);
}

// the rest of a split round, one workgroup per landscape
__global__ __launch_bounds__(FSX_WG) void k_bor_merge(BorArgs A, int round) {
  __shared__ int s_any, s_chg, s_ncomp;
  const int inst = blockIdx.x;
  const InstParams& P = A.PP[inst];
  if (!P.active || !P.run_spl) return;
  int* st = A.bstate + 4 * inst;
  if (st[2]) return;
  const BorView V = bor_view(A, P);
  const int ne = st[3], nl = st[1];
  const int* Lin = (round & 1) ? V.LC1 : V.LC0;
  int* Lout = (round & 1) ? V.LC0 : V.LC1;
  __syncthreads();  // every thread has read the state before thread 0 rewrites it
  const bool more = bor_merge(round, nl, V.SI, V.C, V.BST, V.BN, V.IT, Lin, Lout, true, &s_any, &s_chg, &s_ncomp);
  if (threadIdx.x == 0) {
    st[0] = ne;
    st[3] = 0;
    if (more) st[1] = s_ncomp; else st[2] = 1;
  }
}

// Edge lists of the rounds reuse the edge sort's buffers, free once the ranks are known: list A in
// Lasc/Rasc/segs, list B in the int views of cw (a: first half, c: second half) and skey.  Finishes
// the rounds from BOR_SPLIT on and counts the tree edges.
__global__ __launch_bounds__(FSX_WG) void k_boruvka(BorArgs A, int ldscap) {
  extern __shared__ int sC[];
  __shared__ int s_any, s_chg, s_cnt, s_ncomp;
  const int inst = blockIdx.x;
  const InstParams& P = A.PP[inst];
  if (!P.active || !P.run_spl) return;
  const int nb = A.nbasins[inst], n = A.nconn[inst], tid = threadIdx.x;
  const BorView V = bor_view(A, P);
  const int* st = A.bstate + 4 * inst;
  if (!st[2]) {
    const int ne = st[0], nl = st[1];
    if (nb <= ldscap) {
      for (int b = tid; b < nb; b += FSX_WG) sC[b] = V.C[b];
      __syncthreads();
      boruvka_body<true>(BOR_SPLIT, ne, nl, V.E0, V.E1, V.NP, V.SI, sC, V.BST, V.BN, V.IT, V.LA, V.LB, V.LC0, V.LC1,
                         &s_any, &s_chg, &s_cnt, &s_ncomp);
    } else {
      boruvka_body<false>(BOR_SPLIT, ne, nl, V.E0, V.E1, V.NP, V.SI, V.C, V.BST, V.BN, V.IT, V.LA, V.LB, V.LC0,
                          V.LC1, &s_any, &s_chg, &s_cnt, &s_ncomp);
    }
  }
  int cnt = 0;
  for (int e = tid; e < n; e += FSX_WG) cnt += V.IT[e];
  if (cnt) atomicAdd(A.ntree + inst, cnt);
}

int basins_set_lds(int smem_max) {
  const int bytes = smem_max - 1024;  // room for the static shared words
  if (bytes < 4096) return -1;
  if (hipFuncSetAttribute((const void*)k_boruvka, hipFuncAttributeMaxDynamicSharedMemorySize, bytes) != hipSuccess)
    return -1;
  g_boruvka_lds = bytes / (int)sizeof(int);
  return 0;
}

// ---- 5. orientation from basin0 (order_tree) --------------------------------------------------------
// order_tree (:389-533) walks the spanning tree from basin0 and makes every edge's first basin the
// parent side.  Here: the Euler tour of the tree that starts with basin0's first edge traverses every
// edge parent -> child before child -> parent, so after list-ranking the tour (distance of each
// directed edge to the tour's end, by pointer jumping) the direction with the larger distance is the
// downward one.  The result is the unique rooted tree; the adjacency order (filled with atomics) only
// changes the tour, not the orientation.  Directed edges ("slots") are grouped by source basin.
#define TOUR_FLAG0 32  // flags[TOUR_FLAG0 + round]: the jumping round changed something
#define TOUR_ROUNDS FSX_JUMP_ROUNDS  // the tour has at most 2 slots per basin: the same bound

__device__ __forceinline__ unsigned long long tpack(unsigned dist, unsigned nxt) {
  return ((unsigned long long)dist << 32) | nxt;
}

__global__ __launch_bounds__(FSX_CHUNK) void k_tour_deg(const InstParams* __restrict__ PP, const int* __restrict__ nconn,
                                                        const int* __restrict__ cb0, const int* __restrict__ cb1,
                                                        const uint8_t* __restrict__ intree, int* __restrict__ tdeg) {
  const int inst = blockIdx.y;
  const InstParams& P = PP[inst];
  if (!P.active || !P.run_spl) return;
  const int n = nconn[inst];
  const long long eb = (long long)FSX_EDGE_CAP * P.off;
  for (int e = blockIdx.x * FSX_CHUNK + threadIdx.x; e < n; e += gridDim.x * FSX_CHUNK)
    if (intree[eb + e]) {
      atomicAdd(tdeg + P.off + cb0[eb + e], 1);
      atomicAdd(tdeg + P.off + cb1[eb + e], 1);
    }
}

__global__ __launch_bounds__(FSX_WG) void k_tour_scan(const InstParams* __restrict__ PP, const int* __restrict__ nbasins,
                                                      const int* __restrict__ tdeg, int* __restrict__ toff,
                                                      int* __restrict__ tcur, int* __restrict__ nslot) {
  __shared__ int lds[FSX_WG / 64];
  const int inst = blockIdx.x;
  const InstParams& P = PP[inst];
  if (!P.active || !P.run_spl) return;
  const int nb = nbasins[inst];
  int carry = 0;
  for (int s = 0; s < nb; s += FSX_WG) {
    const int b = s + threadIdx.x;
    const int x = b < nb ? tdeg[P.off + b] : 0;
    int tot;
    const int incl = block_incl_scan<FSX_WG>(x, lds, tot);
    if (b < nb) {
      toff[P.off + b] = carry + incl - x;
      tcur[P.off + b] = 0;
    }
    carry += tot;
  }
  if (threadIdx.x == 0) nslot[inst] = carry;
}

__global__ __launch_bounds__(FSX_CHUNK) void k_tour_fill(const InstParams* __restrict__ PP, const int* __restrict__ nconn,
                                                         const int* __restrict__ cb0, const int* __restrict__ cb1,
                                                         const uint8_t* __restrict__ intree, const int* __restrict__ toff,
                                                         int* __restrict__ tcur, int* __restrict__ ttgt,
                                                         int* __restrict__ trev, int* __restrict__ sa,
                                                         int* __restrict__ sc) {
  const int inst = blockIdx.y;
  const InstParams& P = PP[inst];
  if (!P.active || !P.run_spl) return;
  const int n = nconn[inst];
  const long long eb = (long long)FSX_EDGE_CAP * P.off, sb = 2LL * P.off;
  for (int e = blockIdx.x * FSX_CHUNK + threadIdx.x; e < n; e += gridDim.x * FSX_CHUNK) {
    if (!intree[eb + e]) continue;
    const int a = cb0[eb + e], c = cb1[eb + e];
    const int x = toff[P.off + a] + atomicAdd(tcur + P.off + a, 1);
    const int y = toff[P.off + c] + atomicAdd(tcur + P.off + c, 1);
    ttgt[sb + x] = c;
    ttgt[sb + y] = a;
    trev[sb + x] = y;
    trev[sb + y] = x;
    sa[eb + e] = x;
    sc[eb + e] = y;
  }
}

// successor of slot s (u -> v): the slot after (v -> u) in v's list, circularly; the slot whose
// successor is basin0's first slot ends the tour
__global__ __launch_bounds__(FSX_CHUNK) void k_tour_init(const InstParams* __restrict__ PP, const int* __restrict__ nslot,
                                                         const int* __restrict__ basin0, const int* __restrict__ tdeg,
                                                         const int* __restrict__ toff, const int* __restrict__ ttgt,
                                                         const int* __restrict__ trev,
                                                         unsigned long long* __restrict__ tj) {
  const int inst = blockIdx.y;
  const InstParams& P = PP[inst];
  if (!P.active || !P.run_spl) return;
  const int ns = nslot[inst];
  if (ns == 0) return;
  const long long sb = 2LL * P.off;
  const int s0 = toff[P.off + basin0[inst]];
  for (int s = blockIdx.x * FSX_CHUNK + threadIdx.x; s < ns; s += gridDim.x * FSX_CHUNK) {
    const int v = ttgt[sb + s], r = trev[sb + s];
    const int o = toff[P.off + v];
    const int nx = r + 1 < o + tdeg[P.off + v] ? r + 1 : o;
    tj[sb + s] = nx == s0 ? tpack(0u, (unsigned)s) : tpack(1u, (unsigned)nx);
  }
}

__global__ __launch_bounds__(FSX_CHUNK) void k_tour_jump(const InstParams* __restrict__ PP, const int* __restrict__ nslot,
                                                         unsigned long long* __restrict__ tj, int* __restrict__ flags,
                                                         int round) {
  if (round > 0 && flags[TOUR_FLAG0 + round - 1] == 0) return;
  const int inst = blockIdx.y;
  const InstParams& P = PP[inst];
  if (!P.active || !P.run_spl) return;
  const int ns = nslot[inst];
  unsigned long long* T = tj + 2LL * P.off;
  bool changed = false;
  for (int s = blockIdx.x * FSX_CHUNK + threadIdx.x; s < ns; s += gridDim.x * FSX_CHUNK) {
    unsigned long long w = T[s];
#pragma unroll
    for (int hop = 0; hop < 2; hop++) {  // two hops per round (any pair read is valid)
      const int x = (int)(unsigned)(w & 0xffffffffu);
      const unsigned long long wx = T[x];
      const int xx = (int)(unsigned)(wx & 0xffffffffu);
      if (xx == x) break;
      w = tpack((unsigned)((w >> 32) + (wx >> 32)), (unsigned)xx);
      T[s] = w;
      changed = true;
    }
  }
  if (changed) flags[TOUR_FLAG0 + round] = 1;
}

// make (cb0, cn0) the parent side of every tree edge
__global__ __launch_bounds__(FSX_CHUNK) void k_tour_orient(const InstParams* __restrict__ PP, const int* __restrict__ nconn,
                                                           int* __restrict__ cb0, int* __restrict__ cb1,
                                                           int* __restrict__ cn0, int* __restrict__ cn1,
                                                           const uint8_t* __restrict__ intree, const int* __restrict__ sa,
                                                           const int* __restrict__ sc,
                                                           const unsigned long long* __restrict__ tj,
                                                           const int* __restrict__ flags, int* __restrict__ err) {
  const int inst = blockIdx.y;
  const InstParams& P = PP[inst];
  if (!P.active || !P.run_spl) return;
  if (blockIdx.x == 0 && threadIdx.x == 0 && flags[TOUR_FLAG0 + TOUR_ROUNDS - 1])
    atomicOr(err + inst, FSX_ERR_LEVELS);  // tour longer than the rounds can rank
  const int n = nconn[inst];
  const long long eb = (long long)FSX_EDGE_CAP * P.off, sb = 2LL * P.off;
  for (int e = blockIdx.x * FSX_CHUNK + threadIdx.x; e < n; e += gridDim.x * FSX_CHUNK) {
    if (!intree[eb + e]) continue;
    const unsigned da = (unsigned)(tj[sb + sa[eb + e]] >> 32), dc = (unsigned)(tj[sb + sc[eb + e]] >> 32);
    if (da < dc) {  // c -> a comes first: c is the parent
      const int t0 = cb0[eb + e];
      cb0[eb + e] = cb1[eb + e];
      cb1[eb + e] = t0;
      const int t1 = cn0[eb + e];
      cn0[eb + e] = cn1[eb + e];
      cn1[eb + e] = t1;
    }
  }
}

// ---- 6. receiver path reversal (correct_receivers_v2) ------------------------------------------------
__global__ __launch_bounds__(FSX_CHUNK) void k_correct(const InstParams* __restrict__ PP, const int* __restrict__ nconn,
                                                       const int* __restrict__ cb1, const int* __restrict__ cn0,
                                                       const int* __restrict__ cn1, const uint8_t* __restrict__ intree,
                                                       const int* __restrict__ outlet, int* __restrict__ rec,
                                                       double* __restrict__ length, int* __restrict__ err) {
  const int inst = blockIdx.y;
  const InstParams& P = PP[inst];
  if (!P.active || !P.run_spl) return;
  const int n = nconn[inst];
  const long long eb = (long long)FSX_EDGE_CAP * P.off;
  int* R = rec + P.off;
  double* LEN = length + P.off;
  for (int e = blockIdx.x * FSX_CHUNK + threadIdx.x; e < n; e += gridDim.x * FSX_CHUNK) {
    if (!intree[eb + e]) continue;
    const int node_from = cn1[eb + e];
    if (node_from == -1) continue;
    const int node_to = cn0[eb + e];
    const int outlet_from = outlet[P.off + cb1[eb + e]];
    int next = R[node_from], cur = node_from;
    double prev = LEN[node_from];
    R[node_from] = node_to;
    const int ii = node_from % P.nx - node_to % P.nx, jj = node_from / P.nx - node_to / P.nx;
    LEN[node_from] = P.len[(ii + 1) + 3 * (jj + 1)];  // sqrt(ddx*ddx+ddy*ddy), :643-646
    int guard = 0;
    while (cur != outlet_from) {
      if (++guard > P.nn) { atomicOr(err + inst, FSX_ERR_LEVELS); break; }
      const int rn = R[next];
      R[next] = cur;
      const double t = prev;
      prev = LEN[next];
      LEN[next] = t;
      cur = next;
      next = rn;
    }
  }
}

// ---------------------------------------------------------------------------------------------------
void launch_nr_sort(DevState& S, hipStream_t st) {
  launch_fill_int(S.segctr, S.ninst, 0, st);
  // about 16 slots per thread on the largest landscape's edge capacity, at most 1024 workgroups per landscape
  const int gx = (int)std::min<long long>(1024, std::max<long long>(1, (long long)FSX_EDGE_CAP * S.max_nn / (FSX_CHUNK * 16)));
  k_nr_copy<<<dim3(gx, S.ninst), FSX_CHUNK, 0, st>>>(S.P, S.nconn, S.cw, S.skey, S.sid, S.err);
  k_nr_init<<<S.ninst, 64, 0, st>>>(S.P, S.nconn, S.cw, S.segs, S.nrst);
  // levels for the largest landscape's edge count (6 per node at most): log2(edges / NR_BIG) plus slack
  // for unbalanced partitions, within the per-landscape counters
  int nlev = 4;
  for (long long e = 6LL * S.max_nn; e > NR_BIG && nlev < NR_ST - 2; e >>= 1) nlev++;
  for (int r = 0; r < nlev; r++)
    k_nr_level<<<dim3(std::min(S.nr_gx, 1 << std::min(r, 20)), S.ninst), FSX_WG, 0, st>>>(
        S.P, S.nconn, S.cw, S.skey, S.sid, S.Lasc, S.Rasc, S.segs, S.nrst, r);
  k_nr_big<<<S.ninst, FSX_WG, 0, st>>>(S.P, S.nconn, S.cw, S.skey, S.sid, S.Lasc, S.Rasc, S.segs, S.nrst, S.nsegs,
                                       S.err, nlev);
  k_nr_small<<<dim3(S.nr_small_gx, S.ninst), NR_SMALL_WG, 0, st>>>(S.P, S.nsegs, S.segctr, S.skey, S.sid, S.segs, S.err);
  k_nr_pos<<<dim3(16, S.ninst), FSX_CHUNK, 0, st>>>(S.P, S.nconn, S.sid, S.npos);
}

void launch_lm_labels(DevState& S, hipStream_t st) {
  launch_jump(S, st);  // J = (depth, root) of the steepest-descent forest
}

void launch_lm_basins(DevState& S, hipStream_t st) {
  const int nb = S.nchunks;
  k_root_count<<<nb, FSX_CHUNK, 0, st>>>(S.P, S.chunks, S.rec, S.chunkcnt);
  k_root_scan<<<S.ninst, FSX_WG, 0, st>>>(S.P, S.chunkcnt, S.nbasins, S.basin0);
  k_root_rank<<<nb, FSX_CHUNK, 0, st>>>(S.P, S.chunks, S.rec, S.bc, S.chunkcnt, S.rootrank, S.outlet, S.basin0);
  launch_fill_int(S.bsize, S.ntot, 0, st);
  k_basin_label<<<nb, FSX_CHUNK, 0, st>>>(S.P, S.chunks, S.J, S.rootrank, S.bc, S.ht, S.dmask, S.rec, S.basin,
                                          S.bsize, S.wnode);
  launch_fill_int(S.bigctr, BIG_NCTR, 0, st);
  k_bsize_scan<<<S.ninst, FSX_WG, 0, st>>>(S.P, S.nbasins, S.bsize, S.soff, S.ecount, S.biglist, S.bigctr,
                                           (int)(S.ntot / 64 + S.ninst + 1), S.walk_big);
  k_walk_mask<<<nb, FSX_CHUNK, 0, st>>>(S.P, S.chunks, S.basin, S.wmask);
}

void launch_lm_walk(DevState& S, hipStream_t st) {
  k_basin_walk<<<xcd_grid(S.ninst, S.walk_gx), 64, 0, st>>>(S.P, S.nbasins, S.bc, S.outlet, S.wnode, S.wmask,
                                                           S.soff, S.bsize, S.ecount, S.sB, S.sp, S.sq, S.sw, S.sh,
                                                           S.err, S.ninst, S.walk_gx, S.walk_big);
  launch_lm_big(S, st);
}

void launch_lm_edges(DevState& S, hipStream_t st) {
  k_edge_scan<<<S.ninst, FSX_WG, 0, st>>>(S.P, S.nbasins, S.basin0, S.outlet, S.bc, S.ecount, S.eoff, S.nconn, S.err);
  k_edge_fill<<<S.nchunks, FSX_CHUNK, 0, st>>>(S.P, S.chunks, S.nbasins, S.basin0, S.nconn, S.outlet, S.bc, S.ecount,
                                                S.eoff, S.soff, S.sB, S.sp, S.sq, S.sw, S.cb0, S.cb1, S.cn0, S.cn1, S.cw);
}

void launch_lm_mst(DevState& S, hipStream_t st) {
  launch_fill_int(S.ntree, S.ninst, 0, st);
  // bpar / bfront: the live-component lists (bpar is free until the Euler tour)
  const BorArgs A = {S.P,    S.nbasins, S.nconn, S.cb0,  S.cb1, S.npos, S.sid,  S.comp,  S.best, S.bnew,
                     S.intree, S.ntree, S.Lasc,  S.Rasc, S.segs, S.cw,  S.skey, S.bpar, S.bfront, S.bstate};
  k_bor_start<<<S.ninst, FSX_WG, 0, st>>>(A);
  for (int r = 0; r < BOR_SPLIT; r++) {
    k_bor_scan<<<dim3(S.bor_gx, S.ninst), FSX_WG, 0, st>>>(A, r);
    k_bor_merge<<<S.ninst, FSX_WG, 0, st>>>(A, r);
  }
  k_boruvka<<<S.ninst, FSX_WG, (size_t)g_boruvka_lds * sizeof(int), st>>>(A, g_boruvka_lds);
}

void launch_lm_orient(DevState& S, hipStream_t st) {
  const dim3 g(S.tour_gx, S.ninst), ge(4 * S.tour_gx, S.ninst);
  launch_fill_int(S.tdeg, S.ntot, 0, st);
  launch_fill_int(S.flags + TOUR_FLAG0, 64 - TOUR_FLAG0, 0, st);
  k_tour_deg<<<g, FSX_CHUNK, 0, st>>>(S.P, S.nconn, S.cb0, S.cb1, S.intree, S.tdeg);
  k_tour_scan<<<S.ninst, FSX_WG, 0, st>>>(S.P, S.nbasins, S.tdeg, S.toff, S.bpar, S.maxd);
  k_tour_fill<<<g, FSX_CHUNK, 0, st>>>(S.P, S.nconn, S.cb0, S.cb1, S.intree, S.toff, S.bpar, S.tadj, S.trev, S.Lasc,
                                        S.Rasc);
  k_tour_init<<<ge, FSX_CHUNK, 0, st>>>(S.P, S.maxd, S.basin0, S.tdeg, S.toff, S.tadj, S.trev, S.tjmp);
  for (int r = 0; r < TOUR_ROUNDS; r++) k_tour_jump<<<ge, FSX_CHUNK, 0, st>>>(S.P, S.maxd, S.tjmp, S.flags, r);
  k_tour_orient<<<g, FSX_CHUNK, 0, st>>>(S.P, S.nconn, S.cb0, S.cb1, S.cn0, S.cn1, S.intree, S.Lasc, S.Rasc, S.tjmp,
                                          S.flags, S.err);
}

void launch_lm_correct(DevState& S, hipStream_t st) {
  k_correct<<<dim3(16, S.ninst), FSX_CHUNK, 0, st>>>(S.P, S.nconn, S.cb1, S.cn0, S.cn1, S.intree, S.outlet, S.rec,
                                                      S.length, S.err);
}
