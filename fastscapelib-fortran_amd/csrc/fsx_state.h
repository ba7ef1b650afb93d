// fsx_state.h -- device buffers of one context (a batch of landscapes), owned by the library.
#pragma once
#include "fsx_internal.h"

struct DevState {
  int ninst = 0;
  long long ntot = 0;     // total nodes over the batch
  int nchunks = 0;        // FSX_CHUNK-node blocks covering every instance
  int max_nn = 0, max_nx = 0, max_ny = 0;
  int any_mfd = 0, any_gs = 0;
  InstParams* P = nullptr;  // [ninst]
  Chunk* chunks = nullptr;  // [nchunks]
  int ntiles = 0;           // 2-D tiles (FSX_TW x tile_rows nodes) covering every instance
  int4* tiles = nullptr;    // [ntiles]: (instance, x0, y0, -)
  int tile_rows = 8;        // rows of the tiles in the table
  int nflow = 0;            // workgroups of the dataflow stream-power sweep over all instances
  int2* flow_tab = nullptr; // [nflow]: (instance, rank among the instance's workgroups)
  int nmfd = 0;             // workgroups of the multiple-flow dataflow sweeps over all instances
  int2* mfd_tab = nullptr;  // [nmfd]: (instance, rank among the instance's workgroups)
  int nlvl = 0;             // workgroups of the multi-workgroup level sort (segments of the large landscapes)
  int nlvl_big = 0;         // large landscapes (one workgroup each in the level scan)
  int2* lvl_tab = nullptr;  // [nlvl + nlvl_big]: (instance, segment) per segment, then (instance, segments)
  int* flow_next = nullptr; // [ninst]: the sweep's tile counter of each instance
  int nsweep = 0;           // active single-flow landscapes with deposition (the level sweep k_spl_sweep)
  // launch shapes, chosen by fsx_autotune (no result depends on them)
  int walk_gx = 256;       // basin walk: blocks per landscape
  int jump_blocks = 2048;  // pointer jumping: grid cap
  int tour_gx = 16;        // Euler tour kernels: blocks per landscape (x4 for the ranking rounds)
  int nr_small_gx = 64;    // edge sort: blocks per landscape for the segments in LDS
  int area_wg = 1024;      // drainage-area sweep: threads per landscape (256, 512 or 1024)
  int bor_gx = 4;          // spanning forest: workgroups per landscape for the first rounds' edge scans
  int flow_wg = 1024;      // stream-power dataflow sweep: threads per workgroup (256, 512 or 1024)
  int flow_split = 131072; // stream-power dataflow sweep: nodes per extra workgroup (0: one per landscape)
  int nr_gx = 8;           // edge sort: workgroups per landscape for the big ranges of a level
  int mfd_split = 131072;  // multiple-flow dataflow sweeps: nodes per workgroup of a landscape
  int walk_big = 256;      // closed basins larger than this are searched node-parallel, not walked
  int tile_h = 8;          // stencil kernels (receivers, donor masks): rows per tile (4, 8, 16 or 32)
  int flow_nk = 4;         // stream-power dataflow sweep: Newton iterations of a lane per pass of its wave

  // ---- per node (ntot) ----
  double *h = nullptr, *b = nullptr, *u = nullptr, *kf = nullptr, *kd = nullptr, *precip = nullptr;
  double *etot = nullptr, *erate = nullptr, *area = nullptr, *length = nullptr, *hwater = nullptr;
  double *lake_depth = nullptr, *sedflux = nullptr, *ht = nullptr, *water = nullptr, *zint = nullptr;
  double *zintp = nullptr, *gam = nullptr, *lsed = nullptr;
  // lake water volumes (g > 0): each lake node's share (w_rcv - h, -1 when not in a lake) written by the solve,
  // summed per lake sill exactly in 128-bit fixed point (units of 2^-64 m), so the sum is the same in any order
  double* lakec = nullptr;
  unsigned long long *lacc_lo = nullptr, *lacc_hi = nullptr;
  double *kfint = nullptr, *kdint = nullptr, *g = nullptr, *pexp = nullptr;
  double *adx = nullptr, *ady = nullptr, *adg = nullptr, *adf = nullptr;  // ADI couplings, diagonal, rhs
  double *vx = nullptr, *vy = nullptr;  // velocity field of the advection (allocated when first set)
  double* fmix = nullptr;               // silt fraction of the marine transport (Fmix, default 0.5)
  int *rec = nullptr, *basin = nullptr, *rootrank = nullptr, *order = nullptr;
  int *lstart = nullptr, *lsill = nullptr;
  uint8_t *dmask = nullptr, *bc = nullptr;
  unsigned long long* J = nullptr;  // (distance, ancestor) pairs of the pointer-jumping passes
  int *pos = nullptr, *lcur = nullptr, *rp = nullptr, *LSl = nullptr, *dn = nullptr;
  // per-position (level order) working arrays of the sweeps
  double *A0 = nullptr, *Al = nullptr, *F = nullptr, *HTl = nullptr, *Hl = nullptr, *Wl = nullptr, *HWl = nullptr;
  double *HPl = nullptr, *DHl = nullptr, *ELl = nullptr;
  AreaRec64* arec = nullptr;  // per-position drainage-area records
  SplRec32* srec = nullptr;   // per-position stream-power records
  WalkNode* wnode = nullptr;  // basin walk: (h, basin, preorder successor) per node
  uint8_t* wmask = nullptr;   // basin walk: neighbours k (bit k) that the pass search of a node examines
  // pass search of basins too large for one thread's walk: (ruler << 32 | offset) per node, (next ruler <<
  // 32 | nodes up to it) per ruler, preorder rank per node and node per rank (by basin slice), the ruler
  // and big-basin lists and counters
  unsigned long long *wro = nullptr, *rjmp = nullptr;
  int *wrank = nullptr, *wseq = nullptr;
  int *rlist = nullptr, *rchunk = nullptr;  // rulers (chunk << 8 | lane), per-chunk offsets into the list
  int2* biglist = nullptr;
  int* bigctr = nullptr;

  // ---- per basin (node offsets) ----
  int *outlet = nullptr, *bsize = nullptr, *soff = nullptr, *ecount = nullptr, *eoff = nullptr;
  int *comp = nullptr, *bpar = nullptr, *bnew = nullptr, *tdeg = nullptr, *toff = nullptr;
  unsigned long long* best = nullptr;  // spanning forest: (rank << 32 | other component) of a component's edge
  int *tadj = nullptr;
  int *bfront = nullptr;  // spanning forest: list of the components still roots (with bpar, ping-pong)
  int *trev = nullptr;                   // Euler tour: reverse slot of each directed tree edge (2 per node)
  unsigned long long *tjmp = nullptr;    // Euler tour: (distance to end, next slot), 2 per node

  // ---- per edge (FSX_EDGE_CAP * node offsets) ----
  int *cb0 = nullptr, *cb1 = nullptr, *cn0 = nullptr, *cn1 = nullptr, *sid = nullptr, *npos = nullptr;
  int *Lasc = nullptr, *Rasc = nullptr, *segs = nullptr;
  double *cw = nullptr, *skey = nullptr;
  uint8_t *intree = nullptr;

  // ---- candidate-pass slab (FSX_SLAB_CAP * node offsets) ----
  int *sB = nullptr, *sp = nullptr, *sq = nullptr;
  int* sh = nullptr;  // basin walk: hash index over a basin's pass-table entries beyond those kept in LDS
  double *sw = nullptr;

  // ---- multiple flow direction (allocated when an instance first selects it) ----
  // 8 slots per node in the reference's receiver order: target, weight, direction, stream-power
  // factor; donor slots (u*8+k of the pointing slot); Kahn-layer order and layer starts
  int *mrec = nullptr, *mdon = nullptr;
  double *mw = nullptr, *mfact = nullptr;
  uint8_t *mdir = nullptr, *mflag = nullptr;
  int *mnrec = nullptr, *mndon = nullptr, *mstack = nullptr, *mls = nullptr;
  unsigned long long* mkahn = nullptr;  // Kahn traversal: (layer << 32 | in-edges left) per node
  unsigned long long* mq = nullptr;     // Kahn traversal: queue of (layer << 32 | node)
  int* mcur = nullptr;                  // layer cursors of the scatter into mstack (nn+1 per landscape)
  int* mctr = nullptr;                  // dataflow counters, FSX_MFD_CTR per landscape
  double* mpart = nullptr;              // Gauss-Seidel error partial sums (indexed by the chunk range)
  int* mdec = nullptr;                  // Gauss-Seidel decisions per iteration (100 per landscape)
  int *mnlev = nullptr;  // [ninst]
  int *mpos = nullptr, *mdp = nullptr;  // position of each node in the Kahn order; donor positions (8/node)
  double *mdw = nullptr;                // donor weights by position (8/node)

  // ---- marine transport (allocated when an instance first sets marine parameters) ----
  // single-flow flux to the sea, shelf depth, filled surface; multiple-flow receivers of the filled
  // surface (8 slots per node: target, weight, direction), counts, in-degrees, Kahn order and layer
  // starts; working arrays of the silt / sand diffusion
  double *mar_flux = nullptr, *mar_sd = nullptr, *mar_h0 = nullptr, *mar_w = nullptr;
  int *mar_rec = nullptr, *mar_nrec = nullptr, *mar_indeg = nullptr, *mar_order = nullptr, *mar_ls = nullptr;
  uint8_t* mar_dir = nullptr;
  int *mar_nlev = nullptr, *mar_niter = nullptr;  // [ninst]
  double *mar_q1 = nullptr, *mar_qs = nullptr, *mar_ht = nullptr, *mar_ft = nullptr, *mar_hp = nullptr;
  double *mar_fp = nullptr, *mar_hh = nullptr, *mar_fh = nullptr, *mar_fhp = nullptr, *mar_tint = nullptr;
  double* mar_gam = nullptr;
  int* strati_cnt = nullptr;  // stratigraphy: shore-point count of a distance read-out

  // ---- per instance ----
  int *nbasins = nullptr, *basin0 = nullptr, *nconn = nullptr, *nlev = nullptr, *err = nullptr;
  int *ngs = nullptr, *ntree = nullptr, *chunkcnt = nullptr, *maxd = nullptr, *ngroups = nullptr;
  int *gmeta = nullptr;
  int4* grec = nullptr;  // per-group records of the sweeps, 2 per group
  int2* closed_list = nullptr;  // (instance, basin) of every closed basin
  int* nclosed = nullptr;  // level groups of the sweeps (node offsets + instance, nn+1 per landscape)
  long long *newton = nullptr;
  double *scal = nullptr;      // per-instance scalar scratch (tolerances, GS error)
  unsigned long long *redk = nullptr;  // per-instance order-preserving maxima: maxval(abs(h)), maxval(g)
  int *flags = nullptr;        // convergence flags of iterative kernels (64)
  int *nsegs = nullptr, *segctr = nullptr;  // edge-sort segment lists [ninst]
  int* bstate = nullptr;       // spanning forest between launches: live edges, components, done, kept [4*ninst]
  int* nrst = nullptr;         // edge sort between launches: big ranges per level, segments [32*ninst]
};
