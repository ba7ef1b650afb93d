// fsx_api.cpp -- C ABI of the MI355X landscape step: context lifetime, parameter blocks, field
// transfer and the per-step launch sequence (FastScape_Execute_Step, src/FastScape_api.f90:201-263).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "../../include/fastscape_hip.h"
#include "fsx_state.h"

static thread_local std::string g_err;

// Values of one level kept in LDS by the sweeps (wider levels fall back to global memory); sized
// once from the device's LDS per workgroup.
int g_sweep_cap = 0;

static int configure_sweeps() {
  if (g_sweep_cap) return 0;
  int dev = 0, smem = 0;
  if (hipGetDevice(&dev) != hipSuccess) return -1;
  if (hipDeviceGetAttribute(&smem, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess) return -1;
  if (sweep_set_lds(smem) != 0) return -1;
  if (basins_set_lds(smem) != 0) return -1;
  g_sweep_cap = smem;
  return 0;
}

static int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}
int fsx_fail(int code, const char* msg) { return fail(code, std::string(msg)); }

#define HIPCHK(x)                                                                     \
  do {                                                                                \
    hipError_t e_ = (x);                                                              \
    if (e_ != hipSuccess) return fail(FSX_EHIP, std::string(#x) + ": " + hipGetErrorString(e_)); \
  } while (0)

struct fsx_ctx {
  DevState S;
  std::vector<InstParams> hp;
  std::vector<int> steps;
  std::vector<char> routed;                  // receivers, levels and area of a step are in place
  std::vector<double> xl, yl;
  std::vector<int> ibc;
  std::vector<char> used;                    // slot holds a landscape
  std::map<long long, long long> free_ranges;  // free node ranges of the arena (offset -> length)
  long long capacity = 0, chunk_cap = 0, tile_cap = 0, flow_cap = 0;
  std::vector<int2> flow_host;  // the dataflow sweep's workgroup table as uploaded
  std::vector<int2> mfd_host;   // the multiple-flow sweeps' workgroup table as uploaded
  long long mfd_cap = 0;
  std::vector<int2> lvl_host;   // the multi-workgroup level sort's table as uploaded
  long long lvl_cap = 0;
  std::vector<void*> allocs;
  hipStream_t stream = nullptr;
  hipStream_t last = nullptr;  // the caller's stream of the last fsx_execute_steps that was given one
  bool dirty = true;
  bool any_sfd = false, any_mfd = false;  // which flow-routing paths the active instances use
  bool any_flow = false;                  // some active instance takes the dataflow stream-power sweep
  bool any_advect = false, g_adv = false;  // some active instance advects; as captured
  bool any_marine = false, g_mar = false;  // some active instance runs the marine transport; as captured
  bool any_strati = false, g_str = false;  // some active instance tracks stratigraphic reflectors; as captured
  // one step captured as a HIP graph and replayed; rebuilt when the launch sequence changes
  bool use_graph = true, graph_valid = false, g_sfd = false, g_mfd = false, g_sweep = false;
  hipGraphExec_t gexec = nullptr;
  // live timing of one phase inside graph replays: graph (phases before it), the phase launched
  // directly between two events, graph (phases after it)
  int timed_phase = -1, g_timed = -1, timed_seq = -1;
  int eager_timed = -1;  // the timed phase while steps are launched directly
  hipGraphExec_t gpre = nullptr, gpost = nullptr;
  // emission window over the step's phase sequence (see PHASE)
  int emit_seq = 0, emit_lo = 0, emit_hi = 1 << 30;
  std::vector<int>* seq_log = nullptr;
  long long bytes = 0;
  // phase timers: a ring of (start, stop) event pairs, accumulated every kRing steps
  bool prof = false;
  std::vector<hipEvent_t> ev;  // kRing * PH_COUNT * 2
  int ring_used = 0;
  std::vector<unsigned char> rec_mask;  // which pairs of the ring were recorded
  double acc_ms[PH_COUNT] = {0};
  long long acc_n[PH_COUNT] = {0};
};
static const int kRing = 64;

// Waits for every launch of the context: its own stream and the caller's stream that the last steps were
// enqueued on.  Done before captured graphs, tables or buffers that queued launches may still read are
// destroyed, freed or rewritten.
static hipError_t quiesce(fsx_ctx* c) {
  hipError_t e = c->stream ? hipStreamSynchronize(c->stream) : hipSuccess;
  if (e == hipSuccess && c->last && c->last != c->stream) e = hipStreamSynchronize(c->last);
  return e;
}

static int flush_timers_any(fsx_ctx* c, hipStream_t st) {
  if (c->ring_used == 0) return 0;
  HIPCHK(hipStreamSynchronize(st));
  HIPCHK(quiesce(c));
  for (int s = 0; s < c->ring_used; s++)
    for (int p = 0; p < PH_COUNT; p++) {
      const int k = s * PH_COUNT + p;
      if (!c->rec_mask[k]) continue;
      float ms = 0.f;
      HIPCHK(hipEventElapsedTime(&ms, c->ev[2 * k], c->ev[2 * k + 1]));
      c->acc_ms[p] += ms;
      c->acc_n[p] += 1;
      c->rec_mask[k] = 0;
    }
  c->ring_used = 0;
  return 0;
}

static int flush_timers(fsx_ctx* c, hipStream_t st) {
  if (!c->prof && c->timed_phase < 0) return 0;
  return flush_timers_any(c, st);
}

struct PhaseMark {
  fsx_ctx* c;
  hipStream_t st;
  int k;
  PhaseMark(fsx_ctx* c_, hipStream_t st_, int ph) : c(c_), st(st_), k(-1) {
    if (!c->prof && ph != c->eager_timed) return;
    k = c->ring_used * PH_COUNT + ph;
    (void)hipEventRecord(c->ev[2 * k], st);
  }
  ~PhaseMark() {
    if (k < 0) return;
    (void)hipEventRecord(c->ev[2 * k + 1], st);
    c->rec_mask[k] = 1;
  }
};
// A step is a sequence of phases; only those whose sequence number lies in [emit_lo, emit_hi) are
// launched (all, by default), and with seq_log set the sequence is only recorded.
#define PHASE(ph, call)                                             \
  do {                                                              \
    const int s_ = c->emit_seq++;                                   \
    if (c->seq_log) { c->seq_log->push_back(ph); break; }           \
    if (s_ < c->emit_lo || s_ >= c->emit_hi) break;                 \
    PhaseMark pm_(c, st, ph);                                       \
    call;                                                           \
  } while (0)

extern "C" const char* fsx_last_error(void) { return g_err.c_str(); }
extern "C" const char* fsx_version(void) { return "fastscape-mi355x 0.1"; }

template <typename T>
static int dalloc(fsx_ctx* c, T** p, long long n) {
  if (n <= 0) n = 1;
  void* q = nullptr;
  const size_t bytes = sizeof(T) * (size_t)n;
  hipError_t e = hipMalloc(&q, bytes);
  if (e != hipSuccess) return fail(FSX_ENOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
  e = hipMemset(q, 0, bytes);
  if (e != hipSuccess) return fail(FSX_EHIP, std::string("hipMemset: ") + hipGetErrorString(e));
  c->allocs.push_back(q);
  c->bytes += (long long)bytes;
  *p = (T*)q;
  return 0;
}

// SetBC, src/FastScape_ctx.f90:565-593: write(cbc,'(i4)') ibc; digits 4/2/1/3 = left/right/bottom/top.
static void decode_bc(InstParams& P, int ibc) {
  char tmp[32];
  snprintf(tmp, sizeof tmp, "%4d", ibc);
  if (strlen(tmp) > 4) strcpy(tmp, "****");
  P.fix_left = tmp[3] == '1';
  P.fix_right = tmp[1] == '1';
  P.fix_bottom = tmp[0] == '1';
  P.fix_top = tmp[2] == '1';
  P.i1 = P.fix_left ? 1 : 0;
  P.i2 = P.fix_right ? P.nx - 2 : P.nx - 1;
  P.j1 = P.fix_bottom ? 1 : 0;
  P.j2 = P.fix_top ? P.ny - 2 : P.ny - 1;
  P.xcyc = !P.fix_left && !P.fix_right;
  P.ycyc = !P.fix_bottom && !P.fix_top;
}

// dx = xl/(nx-1), dy = yl/(ny-1) (src/FlowRouting.f90:52-53); neighbour distances
// sqrt((dx*ii)**2+(dy*jj)**2) (:347) evaluated here once, on the host's correctly rounded sqrt.
static void geometry(InstParams& P, double xl, double yl) {
  P.dx = xl / (double)(P.nx - 1);
  P.dy = yl / (double)(P.ny - 1);
  for (int jj = -1; jj <= 1; jj++)
    for (int ii = -1; ii <= 1; ii++) {
      const double a = P.dx * (double)ii, b = P.dy * (double)jj;
      P.len[(ii + 1) + 3 * (jj + 1)] = std::sqrt(a * a + b * b);
    }
  P.dxdy = P.dx * P.dy;
}

static int check_inst(fsx_ctx* c, int inst) {
  if (!c) return fail(FSX_EINVAL, "null context");
  if (inst < 0 || inst >= c->S.ninst || !c->used[inst]) return fail(FSX_EINVAL, "instance id out of range");
  return 0;
}

static int build_flow_table(fsx_ctx* c);
static int build_mfd_table(fsx_ctx* c);
static int build_lvl_table(fsx_ctx* c);
static int build_tiles(fsx_ctx* c);

static int upload_params(fsx_ctx* c) {
  if (!c->dirty) return 0;
  c->any_sfd = c->any_mfd = c->any_flow = false;
  c->S.nsweep = 0;
  c->any_advect = c->any_marine = c->any_strati = false;
  for (InstParams& P : c->hp) {
    // g = g1, or g2 where h - b > 1 when g2 > 0 (src/StreamPowerLaw.f90:236-239): zero everywhere exactly when
    // g1 == 0 and g2 <= 0, and then the Gauss-Seidel loop runs once (:375-376)
    P.flow = (P.sfd && P.g1 == 0.0 && !(P.g2 > 0.0)) ? 1 : 0;
    if (P.active && P.run_advect) c->any_advect = true;
    if (P.active && P.run_marine) c->any_marine = true;
    if (P.active && P.run_strati) c->any_strati = true;
    if (!P.active || !P.run_spl) continue;
    if (P.sfd) c->any_sfd = true; else c->any_mfd = true;
    if (P.flow) c->any_flow = true;
    if (P.sfd && !P.flow) c->S.nsweep++;
  }
  if (int r = build_flow_table(c)) return r;
  if (int r = build_mfd_table(c)) return r;
  if (int r = build_lvl_table(c)) return r;
  if (c->S.tile_h != c->S.tile_rows)
    if (int r = build_tiles(c)) return r;
  HIPCHK(hipMemcpyAsync(c->S.P, c->hp.data(), sizeof(InstParams) * c->hp.size(), hipMemcpyHostToDevice, c->stream));
  c->dirty = false;
  return 0;
}

static int upload_bc(fsx_ctx* c, int inst) {
  const InstParams& P = c->hp[inst];
  std::vector<uint8_t> bc((size_t)P.nn, 0);
  for (int j = 0; j < P.ny; j++)
    for (int i = 0; i < P.nx; i++) {
      const bool f = (P.fix_left && i == 0) || (P.fix_right && i == P.nx - 1) || (P.fix_bottom && j == 0) ||
                     (P.fix_top && j == P.ny - 1);
      bc[(size_t)j * P.nx + i] = f;
    }
  HIPCHK(hipMemcpy(c->S.bc + P.off, bc.data(), bc.size(), hipMemcpyHostToDevice));
  return 0;
}

// ---- node arena and landscape slots ----------------------------------------------------------------
// A context owns an arena of `capacity` nodes and `nslots` landscape slots.  A landscape occupies a
// contiguous node range (first fit over the free ranges) and a slot; the chunk table, the chunk ranges
// and the per-landscape auxiliary offsets are rebuilt whenever the set of landscapes changes.
static long long arena_take(fsx_ctx* c, long long n) {
  for (auto it = c->free_ranges.begin(); it != c->free_ranges.end(); ++it) {
    if (it->second < n) continue;
    const long long off = it->first, len = it->second;
    c->free_ranges.erase(it);
    if (len > n) c->free_ranges[off + n] = len - n;
    return off;
  }
  return -1;
}

static void arena_give(fsx_ctx* c, long long off, long long n) {
  auto it = c->free_ranges.emplace(off, n).first;
  auto nx = std::next(it);
  if (nx != c->free_ranges.end() && it->first + it->second == nx->first) {
    it->second += nx->second;
    c->free_ranges.erase(nx);
  }
  if (it != c->free_ranges.begin()) {
    auto pv = std::prev(it);
    if (pv->first + pv->second == it->first) {
      pv->second += it->second;
      c->free_ranges.erase(it);
    }
  }
}

// Workgroups of the dataflow stream-power sweep.  Every landscape that takes it gets one; with flow_split > 0
// the compute units left over go to the heaviest (nodes, x4 with a Newton solve) as further workgroups of
// flow_split nodes each, at most kFlowGMax per landscape (none for landscapes with more tiles than a
// workgroup's own-tile bitmap holds).  A landscape's workgroups are consecutive in the table, so they are
// dealt to one XCD.  Rebuilt with the parameters; a changed table invalidates the captured step.
static const int kFlowGMax = 16;
static const long long kFlowMaxTiles = 4096LL * 32;  // the own-tile bitmap of k_spl_flow (FLOW_OWN_WORDS words)
static int cu_count() {
  static int n = 0;
  if (!n) {
    int dev = 0;
    hipDeviceProp_t p;
    n = (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&p, dev) == hipSuccess) ? p.multiProcessorCount
                                                                                             : 256;
  }
  return n;
}
static int build_flow_table(fsx_ctx* c) {
  DevState& S = c->S;
  std::vector<int> idx;
  auto weight = [&](int k) {
    const InstParams& P = c->hp[k];
    return (long long)P.nn * (std::fabs(P.n - 1.0) < 2.2250738585072014e-308 ? 1 : 4);
  };
  for (int k = 0; k < S.ninst; k++) {
    InstParams& P = c->hp[k];
    P.flow_g = (c->used[k] && P.active && P.run_spl && P.flow) ? 1 : 0;
    if (P.flow_g) idx.push_back(k);
  }
  if (S.flow_split > 0) {
    std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return weight(a) > weight(b); });
    long long extra = std::max(0LL, (long long)cu_count() - (long long)idx.size());
    for (int k : idx) {
      InstParams& P = c->hp[k];
      if ((P.nn + 63) / 64 > kFlowMaxTiles) continue;
      const long long want = std::min<long long>(kFlowGMax, (weight(k) + S.flow_split - 1) / S.flow_split);
      const long long add = std::min(std::max(0LL, want - 1), extra);
      P.flow_g += (int)add;
      extra -= add;
    }
  }
  std::vector<int2> tab;
  for (int k = 0; k < S.ninst; k++)
    for (int r = 0; r < c->hp[k].flow_g; r++) tab.push_back(make_int2(k, r));
  S.nflow = (int)tab.size();
  const bool same = tab.size() == c->flow_host.size() &&
                    std::equal(tab.begin(), tab.end(), c->flow_host.begin(),
                               [](const int2& a, const int2& b) { return a.x == b.x && a.y == b.y; });
  if (same) return 0;
  if ((long long)tab.size() > c->flow_cap) {
    HIPCHK(quiesce(c));
    if (S.flow_tab) HIPCHK(hipFree(S.flow_tab));
    S.flow_tab = nullptr;
    c->flow_cap = 0;
    const long long cap = (long long)tab.size() + 64;
    hipError_t e = hipMalloc(&S.flow_tab, sizeof(int2) * cap);
    if (e != hipSuccess) return fail(FSX_ENOMEM, std::string("hipMalloc (flow table): ") + hipGetErrorString(e));
    c->flow_cap = cap;
  }
  if (!tab.empty()) {
    HIPCHK(quiesce(c));  // launches queued on the caller's stream may still read the old table
    HIPCHK(hipMemcpyAsync(S.flow_tab, tab.data(), sizeof(int2) * tab.size(), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
  }
  c->flow_host = std::move(tab);
  c->graph_valid = false;
  return 0;
}

// Workgroups of the multiple-flow dataflow sweeps: one per nodes_per (mfd_split) nodes of each landscape
// that routes by multiple flow directions, at most kMfdGMax (one XCD's compute units); the extra ones are
// dealt to the largest landscapes first while the device has compute units for them.  A landscape's
// workgroups are consecutive in the table (one XCD).  No result depends on the count.
static const int kMfdGMax = 32;
static int build_mfd_table(fsx_ctx* c) {
  DevState& S = c->S;
  std::vector<int> idx;
  for (int k = 0; k < S.ninst; k++) {
    InstParams& P = c->hp[k];
    P.mfd_g = (c->used[k] && P.active && P.run_spl && !P.sfd) ? 1 : 0;
    if (P.mfd_g) idx.push_back(k);
  }
  std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return c->hp[a].nn > c->hp[b].nn; });
  long long extra = std::max(0LL, (long long)cu_count() - (long long)idx.size());
  for (int k : idx) {
    InstParams& P = c->hp[k];
    const long long want = std::min<long long>(kMfdGMax, std::max(1LL, ((long long)P.nn + S.mfd_split - 1) / S.mfd_split));
    const long long add = std::min(want - 1, extra);
    P.mfd_g += (int)add;
    extra -= add;
  }
  std::vector<int2> tab;
  for (int k = 0; k < S.ninst; k++)
    for (int r = 0; r < c->hp[k].mfd_g; r++) tab.push_back(make_int2(k, r));
  S.nmfd = (int)tab.size();
  const bool same = tab.size() == c->mfd_host.size() &&
                    std::equal(tab.begin(), tab.end(), c->mfd_host.begin(),
                               [](const int2& a, const int2& b) { return a.x == b.x && a.y == b.y; });
  if (same) return 0;
  HIPCHK(quiesce(c));  // queued launches may still read the old table
  if ((long long)tab.size() > c->mfd_cap) {
    if (S.mfd_tab) HIPCHK(hipFree(S.mfd_tab));
    S.mfd_tab = nullptr;
    c->mfd_cap = 0;
    const long long cap = (long long)tab.size() + 64;
    hipError_t e = hipMalloc(&S.mfd_tab, sizeof(int2) * cap);
    if (e != hipSuccess) return fail(FSX_ENOMEM, std::string("hipMalloc (mfd table): ") + hipGetErrorString(e));
    c->mfd_cap = cap;
  }
  if (!tab.empty()) {
    HIPCHK(hipMemcpyAsync(S.mfd_tab, tab.data(), sizeof(int2) * tab.size(), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
  }
  c->mfd_host = std::move(tab);
  c->graph_valid = false;
  return 0;
}

// Level sort of the large landscapes: one workgroup per FSX_LVL_SEG-node segment of every stream-power
// landscape of at least FSX_LVL_BIG nodes (the others are sorted by one workgroup each), then one entry
// (instance, segments) per such landscape for the level scan.
static int build_lvl_table(fsx_ctx* c) {
  DevState& S = c->S;
  std::vector<int2> tab, big;
  for (int k = 0; k < S.ninst; k++) {
    InstParams& P = c->hp[k];
    const bool on = c->used[k] && P.active && P.run_spl && P.nn >= FSX_LVL_BIG;
    P.lvl_segs = on ? (P.nn + FSX_LVL_SEG - 1) / FSX_LVL_SEG : 0;
    for (int r = 0; r < P.lvl_segs; r++) tab.push_back(make_int2(k, r));
    if (on) big.push_back(make_int2(k, P.lvl_segs));
  }
  S.nlvl = (int)tab.size();
  S.nlvl_big = (int)big.size();
  tab.insert(tab.end(), big.begin(), big.end());
  const bool same = tab.size() == c->lvl_host.size() &&
                    std::equal(tab.begin(), tab.end(), c->lvl_host.begin(),
                               [](const int2& a, const int2& b) { return a.x == b.x && a.y == b.y; });
  if (same) return 0;
  HIPCHK(quiesce(c));  // queued launches may still read the old table
  if ((long long)tab.size() > c->lvl_cap) {
    if (S.lvl_tab) HIPCHK(hipFree(S.lvl_tab));
    S.lvl_tab = nullptr;
    c->lvl_cap = 0;
    const long long cap = (long long)tab.size() + 64;
    hipError_t e = hipMalloc(&S.lvl_tab, sizeof(int2) * cap);
    if (e != hipSuccess) return fail(FSX_ENOMEM, std::string("hipMalloc (level sort table): ") + hipGetErrorString(e));
    c->lvl_cap = cap;
  }
  if (!tab.empty()) {
    HIPCHK(hipMemcpyAsync(S.lvl_tab, tab.data(), sizeof(int2) * tab.size(), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
  }
  c->lvl_host = std::move(tab);
  c->graph_valid = false;
  return 0;
}

// The 2-D tiles of the stencil kernels, tile_h rows each, row-major within a landscape.  The table is
// rewritten in place once queued launches are done; captured launches carry its size, so the graphs are
// rebuilt.
static int build_tiles(fsx_ctx* c) {
  DevState& S = c->S;
  std::vector<int4> tiles;
  for (int k = 0; k < S.ninst; k++) {
    if (!c->used[k]) continue;
    const InstParams& P = c->hp[k];
    for (int y0 = 0; y0 < P.ny; y0 += S.tile_h)
      for (int x0 = 0; x0 < P.nx; x0 += FSX_TW) tiles.push_back(make_int4(k, x0, y0, 0));
  }
  HIPCHK(quiesce(c));
  if ((long long)tiles.size() > c->tile_cap) {
    if (S.tiles) HIPCHK(hipFree(S.tiles));
    S.tiles = nullptr;
    S.ntiles = 0;
    c->tile_cap = 0;
    const long long cap = (long long)tiles.size() + tiles.size() / 4 + 64;
    hipError_t e = hipMalloc(&S.tiles, sizeof(int4) * cap);
    if (e != hipSuccess) return fail(FSX_ENOMEM, std::string("hipMalloc (tile table): ") + hipGetErrorString(e));
    c->tile_cap = cap;
  }
  if (!tiles.empty()) {
    HIPCHK(hipMemcpyAsync(S.tiles, tiles.data(), sizeof(int4) * tiles.size(), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
  }
  S.ntiles = (int)tiles.size();
  S.tile_rows = S.tile_h;
  c->graph_valid = false;
  return 0;
}

static int relayout(fsx_ctx* c) {
  DevState& S = c->S;
  std::vector<Chunk> chunks;
  std::vector<std::pair<int, int>> by_off;
  for (int k = 0; k < S.ninst; k++)
    if (c->used[k]) by_off.push_back({c->hp[k].off, k});
  std::sort(by_off.begin(), by_off.end());
  // the per-landscape level bounds of the last step live at aux: a landscape whose aux moves needs a
  // new step before its chi and fluxes can be read out
  for (size_t r = 0; r < by_off.size(); r++) {
    InstParams& P = c->hp[by_off[r].second];
    const int aux = by_off[r].first + (int)r;
    if (P.aux != aux) c->routed[by_off[r].second] = 0;
    P.aux = aux;
  }
  for (int k = 0; k < S.ninst; k++) {
    InstParams& P = c->hp[k];
    P.c0 = (int)chunks.size();
    if (c->used[k])
      for (int s = 0; s < P.nn; s += FSX_CHUNK) chunks.push_back(Chunk{k, s});
    P.c1 = (int)chunks.size();
  }
  if ((long long)chunks.size() > c->chunk_cap) return fail(FSX_EINVAL, "chunk table overflow");
  S.nchunks = (int)chunks.size();
  if (int r = build_tiles(c)) return r;
  HIPCHK(quiesce(c));  // the chunk table is rewritten in place
  if (!chunks.empty())
    HIPCHK(hipMemcpyAsync(S.chunks, chunks.data(), sizeof(Chunk) * chunks.size(), hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  S.max_nn = S.max_nx = S.max_ny = 0;
  for (int k = 0; k < S.ninst; k++)
    if (c->used[k]) {
      S.max_nn = std::max(S.max_nn, c->hp[k].nn);
      S.max_nx = std::max(S.max_nx, c->hp[k].nx);
      S.max_ny = std::max(S.max_ny, c->hp[k].ny);
    }
  c->dirty = true;
  c->graph_valid = false;
  return 0;
}

// State of a new landscape, SetUp defaults (src/FastScape_ctx.f90:81-97): h = b = u = etot = erate =
// 0, precip = 1, kf = kd = 0; everything else is rewritten by each step.
__global__ void k_slot_init(long long off, int nn, double* h, double* b, double* u, double* etot, double* erate,
                            double* area, double* lake_depth, double* sedflux, double* hwater, double* kf,
                            double* kd, double* g, double* precip, double* fmix) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nn; i += gridDim.x * blockDim.x) {
    const long long q = off + i;
    h[q] = 0.0; b[q] = 0.0; u[q] = 0.0; etot[q] = 0.0; erate[q] = 0.0; area[q] = 0.0; lake_depth[q] = 0.0;
    sedflux[q] = 0.0; hwater[q] = 0.0; kf[q] = 0.0; kd[q] = 0.0; g[q] = 0.0; precip[q] = 1.0;
    fmix[q] = 0.5;  // Fmix = 0.5 (src/FastScape_ctx.f90:96)
  }
}

static int ctx_alloc(long long capacity, int nslots, fsx_ctx** out) {
  if (capacity <= 0 || nslots <= 0) return fail(FSX_EINVAL, "empty context");
  if (capacity > 300000000LL) return fail(FSX_EINVAL, "fsx_create: batch too large");
  fsx_ctx* c = new fsx_ctx();
  DevState& S = c->S;
  S.ninst = nslots;
  c->hp.resize(nslots);
  for (InstParams& P : c->hp) memset(&P, 0, sizeof P);
  c->used.assign(nslots, 0);
  c->steps.assign(nslots, 0);
  c->routed.assign(nslots, 0);
  c->xl.assign(nslots, 0.0);
  c->yl.assign(nslots, 0.0);
  c->ibc.assign(nslots, 1111);
  c->capacity = capacity;
  c->free_ranges[0] = capacity;
  c->chunk_cap = capacity / FSX_CHUNK + nslots + 1;
  S.ntot = capacity;
  const long long N = capacity;
  const int ninst = nslots;
  int rc = 0;
  if (configure_sweeps() != 0) { delete c; return fail(FSX_EHIP, "cannot configure LDS of the sweep kernels"); }
  hipError_t he = hipStreamCreate(&c->stream);
  if (he != hipSuccess) { delete c; return fail(FSX_EHIP, "hipStreamCreate failed"); }
#define A(p, n) if (!rc) rc = dalloc(c, &S.p, (n))
  A(P, ninst); A(chunks, c->chunk_cap);
  A(h, N); A(b, N); A(u, N); A(kf, N); A(kd, N); A(precip, N); A(etot, N); A(erate, N); A(area, N);
  A(length, N); A(hwater, N); A(lake_depth, N); A(sedflux, N); A(ht, N); A(water, N); A(zint, N); A(zintp, N);
  A(gam, N); A(lakec, N); A(lacc_lo, N); A(lacc_hi, N); A(lsed, N); A(kfint, N); A(kdint, N); A(g, N); A(pexp, N);
  A(adx, N); A(ady, N); A(adg, N); A(adf, N); A(fmix, N);
  A(rec, N); A(basin, N); A(rootrank, N); A(order, N); A(lstart, N + ninst); A(lsill, N);
  A(dmask, N); A(bc, N);
  A(outlet, N); A(bsize, N); A(soff, N); A(ecount, N); A(eoff, N); A(comp, N); A(bpar, N); A(bnew, N); A(best, N);
  A(tdeg, N); A(toff, N); A(tadj, 2 * N); A(bfront, N); A(trev, 2 * N); A(tjmp, 2 * N);
  A(cb0, FSX_EDGE_CAP * N); A(cb1, FSX_EDGE_CAP * N); A(cn0, FSX_EDGE_CAP * N); A(cn1, FSX_EDGE_CAP * N);
  A(sid, FSX_EDGE_CAP * N); A(npos, FSX_EDGE_CAP * N); A(Lasc, FSX_EDGE_CAP * N); A(Rasc, FSX_EDGE_CAP * N);
  A(segs, FSX_EDGE_CAP * N); A(cw, FSX_EDGE_CAP * N); A(skey, FSX_EDGE_CAP * N); A(intree, FSX_EDGE_CAP * N);
  A(sB, FSX_SLAB_CAP * N); A(sp, FSX_SLAB_CAP * N); A(sq, FSX_SLAB_CAP * N); A(sw, FSX_SLAB_CAP * N);
  A(sh, FSX_SLAB_CAP * N);
  A(nbasins, ninst); A(basin0, ninst); A(nconn, ninst); A(nlev, ninst); A(err, ninst); A(ngs, ninst);
  A(ntree, ninst); A(chunkcnt, c->chunk_cap); A(newton, ninst); A(scal, ninst); A(redk, 2 * ninst); A(flags, 64);
  A(nsegs, ninst); A(segctr, ninst); A(maxd, ninst); A(bstate, 4 * ninst); A(nrst, 32 * ninst); A(flow_next, ninst);
  A(J, N); A(ngroups, ninst); A(grec, 2 * (N + ninst)); A(gmeta, N + ninst); A(pos, N); A(lcur, N + ninst); A(rp, N); A(LSl, N); A(dn, 8 * N);
  A(A0, N); A(Al, N); A(F, N); A(HTl, N); A(Hl, N); A(Wl, N); A(HWl, N); A(HPl, N); A(DHl, N); A(ELl, N);
  A(arec, N); A(srec, N); A(wnode, N); A(wmask, N);
  A(wro, N); A(rjmp, N); A(wrank, N); A(wseq, N); A(rlist, N); A(rchunk, c->chunk_cap); A(biglist, N / 64 + ninst + 1); A(bigctr, 64);
#undef A
  if (rc) { fsx_destroy(c); return rc; }
  *out = c;
  return 0;
}

static int slot_add(fsx_ctx* c, int nx, int ny, int* inst, bool relay) {
  if (nx < 3 || ny < 3) return fail(FSX_EINVAL, "nx and ny must be >= 3");
  const long long nn = (long long)nx * ny;
  int k = -1;
  for (int s = 0; s < c->S.ninst; s++)
    if (!c->used[s]) { k = s; break; }
  if (k < 0) return fail(FSX_ENOMEM, "no free landscape slot");
  const long long off = arena_take(c, nn);
  if (off < 0) return fail(FSX_ENOMEM, "no free node range of that size");
  InstParams& P = c->hp[k];
  memset(&P, 0, sizeof P);
  P.nx = nx;
  P.ny = ny;
  P.nn = (int)nn;
  P.off = (int)off;
  P.active = 1;
  P.newton_tol = 1.0e-3;
  P.p = 1.0;
  P.p_raw = 1.0;  // p_mfd_exp = 1 (src/FastScape_ctx.f90:93)
  P.sfd = 1;
  decode_bc(P, 1111);
  geometry(P, (double)(P.nx - 1), (double)(P.ny - 1));
  c->used[k] = 1;
  c->steps[k] = 0;
  c->routed[k] = 0;
  c->xl[k] = (double)(nx - 1);  // the geometry() default: dx = dy = 1 until set_xl_yl
  c->yl[k] = (double)(ny - 1);
  c->ibc[k] = 1111;
  DevState& S = c->S;
  k_slot_init<<<(int)std::min<long long>(1024, (nn + 255) / 256), 256, 0, c->stream>>>(
      off, (int)nn, S.h, S.b, S.u, S.etot, S.erate, S.area, S.lake_depth, S.sedflux, S.hwater, S.kf, S.kd, S.g,
      S.precip, S.fmix);
  HIPCHK(hipGetLastError());
  if (int r = upload_bc(c, k)) return r;
  if (relay)
    if (int r = relayout(c)) return r;
  *inst = k;
  return 0;
}

extern "C" int fsx_create(int ninst, const int* nx, const int* ny, fsx_ctx** out) {
  if (!out || ninst <= 0 || !nx || !ny) return fail(FSX_EINVAL, "fsx_create: bad arguments");
  *out = nullptr;
  long long total = 0;
  for (int k = 0; k < ninst; k++) {
    if (nx[k] < 3 || ny[k] < 3) return fail(FSX_EINVAL, "fsx_create: nx and ny must be >= 3");
    total += (long long)nx[k] * ny[k];
  }
  fsx_ctx* c = nullptr;
  if (int r = ctx_alloc(total, ninst, &c)) return r;
  for (int k = 0; k < ninst; k++) {
    int id = -1;
    if (int r = slot_add(c, nx[k], ny[k], &id, false)) { fsx_destroy(c); return r; }
  }
  if (int r = relayout(c)) { fsx_destroy(c); return r; }
  *out = c;
  return 0;
}

extern "C" int fsx_create_pool(long long capacity_nodes, int nslots, fsx_ctx** out) {
  if (!out) return fail(FSX_EINVAL, "fsx_create_pool: null output");
  *out = nullptr;
  fsx_ctx* c = nullptr;
  if (int r = ctx_alloc(capacity_nodes, nslots, &c)) return r;
  if (int r = relayout(c)) { fsx_destroy(c); return r; }
  *out = c;
  return 0;
}

extern "C" int fsx_add_instance(fsx_ctx* c, int nx, int ny, int* inst) {
  if (!c || !inst) return fail(FSX_EINVAL, "fsx_add_instance: bad arguments");
  HIPCHK(quiesce(c));
  return slot_add(c, nx, ny, inst, true);
}

extern "C" int fsx_remove_instance(fsx_ctx* c, int inst) {
  if (int r = check_inst(c, inst)) return r;
  HIPCHK(quiesce(c));
  InstParams& P = c->hp[inst];
  arena_give(c, P.off, P.nn);
  memset(&P, 0, sizeof P);
  c->used[inst] = 0;
  return relayout(c);
}

extern "C" int fsx_device_memory(long long* free_bytes, long long* total_bytes) {
  size_t f = 0, t = 0;
  HIPCHK(hipMemGetInfo(&f, &t));
  if (free_bytes) *free_bytes = (long long)f;
  if (total_bytes) *total_bytes = (long long)t;
  return 0;
}

extern "C" int fsx_pool_info(fsx_ctx* c, long long* free_nodes, long long* largest_free, int* free_slots) {
  if (!c) return fail(FSX_EINVAL, "null context");
  long long tot = 0, big = 0;
  for (auto& fr : c->free_ranges) { tot += fr.second; big = std::max(big, fr.second); }
  int fs = 0;
  for (int s = 0; s < c->S.ninst; s++) fs += !c->used[s];
  if (free_nodes) *free_nodes = tot;
  if (largest_free) *largest_free = big;
  if (free_slots) *free_slots = fs;
  return 0;
}

extern "C" int fsx_destroy(fsx_ctx* c) {
  if (!c) return 0;
  (void)quiesce(c);
  if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
  if (c->gpre) (void)hipGraphExecDestroy(c->gpre);
  if (c->gpost) (void)hipGraphExecDestroy(c->gpost);
  for (auto& e : c->ev) (void)hipEventDestroy(e);
  for (void* p : c->allocs) (void)hipFree(p);
  if (c->S.tiles) (void)hipFree(c->S.tiles);
  if (c->S.flow_tab) (void)hipFree(c->S.flow_tab);
  if (c->S.mfd_tab) (void)hipFree(c->S.mfd_tab);
  if (c->S.lvl_tab) (void)hipFree(c->S.lvl_tab);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return 0;
}

extern "C" int fsx_set_xl_yl(fsx_ctx* c, int inst, double xl, double yl) {
  if (int r = check_inst(c, inst)) return r;
  c->xl[inst] = xl;
  c->yl[inst] = yl;
  geometry(c->hp[inst], xl, yl);
  c->dirty = true;
  return 0;
}

extern "C" int fsx_set_dt(fsx_ctx* c, int inst, double dt) {
  if (int r = check_inst(c, inst)) return r;
  c->hp[inst].dt = dt;
  c->dirty = true;
  return 0;
}

extern "C" int fsx_set_bc(fsx_ctx* c, int inst, int ibc) {
  if (int r = check_inst(c, inst)) return r;
  c->ibc[inst] = ibc;
  decode_bc(c->hp[inst], ibc);
  c->dirty = true;
  HIPCHK(hipStreamSynchronize(c->stream));
  return upload_bc(c, inst);
}

// Multiple-flow-direction buffers, allocated the first time an instance selects that routing.
static int mfd_alloc(fsx_ctx* c) {
  DevState& S = c->S;
  if (S.mrec) return 0;
  HIPCHK(hipStreamSynchronize(c->stream));
  const long long N = S.ntot;
  int rc = 0;
  c->graph_valid = false;
#define A(p, n) if (!rc) rc = dalloc(c, &S.p, (n))
  A(mrec, 8 * N); A(mdon, 8 * N); A(mw, 8 * N); A(mfact, 8 * N); A(mdir, 8 * N); A(mflag, N);
  A(mnrec, N); A(mndon, N); A(mstack, N); A(mls, N + S.ninst); A(mnlev, S.ninst);
  A(mkahn, N); A(mq, N); A(mcur, N + S.ninst); A(mctr, FSX_MFD_CTR * S.ninst); A(mpart, c->chunk_cap);
  A(mdec, 100 * S.ninst);
  A(mpos, N); A(mdp, 8 * N); A(mdw, 8 * N);
#undef A
  return rc;
}

static int advect_alloc(fsx_ctx* c) {
  DevState& S = c->S;
  if (S.vx) return 0;
  HIPCHK(hipStreamSynchronize(c->stream));
  int rc = dalloc(c, &S.vx, S.ntot);
  if (!rc) rc = dalloc(c, &S.vy, S.ntot);
  c->graph_valid = false;
  return rc;
}

// Marine buffers, allocated the first time an instance sets marine parameters.
static int marine_alloc(fsx_ctx* c) {
  DevState& S = c->S;
  if (S.mar_rec) return 0;
  HIPCHK(hipStreamSynchronize(c->stream));
  const long long N = S.ntot;
  int rc = 0;
  c->graph_valid = false;
#define A(p, n) if (!rc) rc = dalloc(c, &S.p, (n))
  A(mar_flux, N); A(mar_sd, N); A(mar_h0, N); A(mar_w, 8 * N); A(mar_rec, 8 * N); A(mar_dir, 8 * N);
  A(mar_nrec, N); A(mar_indeg, N); A(mar_order, N); A(mar_ls, N + S.ninst); A(mar_nlev, S.ninst);
  A(mar_niter, S.ninst); A(mar_q1, N); A(mar_qs, N); A(mar_ht, N); A(mar_ft, N); A(mar_hp, N); A(mar_fp, N);
  A(mar_hh, N); A(mar_fh, N); A(mar_fhp, N); A(mar_tint, N); A(mar_gam, N);
#undef A
  return rc;
}

static int put(fsx_ctx* c, double* dst, const double* src, int n, int on_device) {
  HIPCHK(hipMemcpyAsync(dst, src, sizeof(double) * (size_t)n, on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice,
                        c->stream));
  if (!on_device) HIPCHK(hipStreamSynchronize(c->stream));
  return 0;
}

extern "C" int fsx_set_erosional_parameters(fsx_ctx* c, int inst, const double* kf, double kfsed, double m, double n,
                                            const double* kd, double kdsed, double g1, double g2, double p,
                                            int on_device) {
  if (int r = check_inst(c, inst)) return r;
  if (!kf || !kd) return fail(FSX_EINVAL, "kf and kd are required");
  InstParams& P = c->hp[inst];
  if (int r = put(c, c->S.kf + P.off, kf, P.nn, on_device)) return r;
  if (int r = put(c, c->S.kd + P.off, kd, P.nn, on_device)) return r;
  P.run_spl = 1;
  P.kfsed = kfsed;
  P.m = m;
  P.n = n;
  P.kdsed = kdsed;
  P.g1 = g1;
  P.g2 = g2;
  P.p = p;
  P.p_raw = p;
  P.sfd = 0;
  if (p < -1.5) {
    P.sfd = 1;
    P.p = 1.0;
  }
  // runDiffusion if maxval(kd) > tiny or kdsed > tiny (sticky, as the reference)
  double kdmax = -1.0e308;
  if (on_device) {
    std::vector<double> tmp((size_t)P.nn);
    HIPCHK(hipMemcpy(tmp.data(), kd, sizeof(double) * P.nn, hipMemcpyDeviceToHost));
    for (double x : tmp) kdmax = std::max(kdmax, x);
  } else {
    for (int i = 0; i < P.nn; i++) kdmax = std::max(kdmax, kd[i]);
  }
  const double tiny = 2.2250738585072014e-308;
  if (kdmax > tiny || kdsed > tiny) P.run_diff = 1;
  c->dirty = true;
  if (!P.sfd)
    if (int r = mfd_alloc(c)) return r;
  return 0;
}

extern "C" int fsx_set_marine_parameters(fsx_ctx* c, int inst, double sealevel, double poro1, double poro2,
                                         double zporo1, double zporo2, double ratio, double layer, double kdsea1,
                                         double kdsea2) {
  if (int r = check_inst(c, inst)) return r;
  if (int r = marine_alloc(c)) return r;
  InstParams& P = c->hp[inst];
  P.run_marine = 1;
  P.sealevel = sealevel;
  P.poro1 = poro1;
  P.poro2 = poro2;
  P.zporo1 = zporo1;
  P.zporo2 = zporo2;
  P.ratio = ratio;
  P.layer = layer;
  P.kdsea1 = kdsea1;
  P.kdsea2 = kdsea2;
  c->dirty = true;
  return 0;
}

extern "C" int fsx_strati(fsx_ctx* c, int inst, int nstep, int nreflector, int nfreq, double vex) {
  if (int r = check_inst(c, inst)) return r;
  if (nreflector < 1 || nfreq < 1 || nstep / nreflector < 1)
    return fail(FSX_EINVAL, "fsx_strati: need nreflector >= 1, nfreq >= 1 and nstep >= nreflector");
  InstParams& P = c->hp[inst];
  DevState& S = c->S;
  HIPCHK(hipStreamSynchronize(c->stream));
  int rc = 0;
  if (!S.strati_cnt) rc = dalloc(c, &S.strati_cnt, 1);
  double *r = nullptr, *e = nullptr;
  if (!rc) rc = dalloc(c, &r, (long long)nreflector * P.nn);  // zero-filled
  if (!rc) rc = dalloc(c, &e, (long long)nreflector * P.nn);
  if (rc) return rc;
  P.srefl = r;
  P.serod = e;
  P.run_strati = 1;
  P.nref = nreflector;
  P.iref = 1;
  c->dirty = true;
  c->graph_valid = false;
  return 0;
}

extern "C" int fsx_strati_set_active(fsx_ctx* c, int inst, int ireflector) {
  if (int r = check_inst(c, inst)) return r;
  InstParams& P = c->hp[inst];
  if (!P.run_strati) return fail(FSX_EINVAL, "fsx_strati_set_active: no stratigraphy on this landscape");
  if (ireflector < 1 || ireflector > P.nref + 1)
    return fail(FSX_EINVAL, "fsx_strati_set_active: reflector out of range");
  P.iref = ireflector;
  c->dirty = true;
  return 0;
}

extern "C" int fsx_strati_read(fsx_ctx* c, int inst, int which, double* dst) {
  if (int r = check_inst(c, inst)) return r;
  if (!dst) return fail(FSX_EINVAL, "fsx_strati_read: null destination");
  const InstParams& P = c->hp[inst];
  if (!P.run_strati) return fail(FSX_EINVAL, "fsx_strati_read: no stratigraphy on this landscape");
  DevState& S = c->S;
  const size_t nn = (size_t)P.nn, nr = (size_t)P.nref;
  hipStream_t st = c->stream;
  if (int r = upload_params(c)) return r;
  switch (which) {
    case FSX_STRATI_REFLECTORS:
      HIPCHK(hipMemcpyAsync(dst, P.srefl, sizeof(double) * nn * nr, hipMemcpyDeviceToHost, st));
      break;
    case FSX_STRATI_ERODED:
      HIPCHK(hipMemcpyAsync(dst, P.serod, sizeof(double) * nn * nr, hipMemcpyDeviceToHost, st));
      break;
    case FSX_STRATI_SLOPES:
      for (size_t i = 0; i < nr; i++) {
        launch_strati_slope(S, inst, P.nn, P.srefl + i * nn, S.zint + P.off, st);
        HIPCHK(hipMemcpyAsync(dst + i * nn, S.zint + P.off, sizeof(double) * nn, hipMemcpyDeviceToHost, st));
      }
      break;
    case FSX_STRATI_SHORE:
      if (!c->routed[inst]) return fail(FSX_EINVAL, "fsx_strati_read: distances to shore need the routing of a step");
      for (size_t i = 0; i <= nr; i++) {  // each reflector, then the basement
        const double* surf = i < nr ? P.srefl + i * nn : S.b + P.off;
        launch_strati_shore(S, inst, P.nn, c->xl[inst], c->yl[inst], surf, S.adx + P.off, S.ady + P.off,
                            S.strati_cnt, S.zintp + P.off, st);
        HIPCHK(hipMemcpyAsync(dst + i * nn, S.zintp + P.off, sizeof(double) * nn, hipMemcpyDeviceToHost, st));
      }
      break;
    default:
      return fail(FSX_EINVAL, "fsx_strati_read: unknown item");
  }
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(st));
  return 0;
}

extern "C" int fsx_set_newton_tol(fsx_ctx* c, int inst, double tol) {
  if (int r = check_inst(c, inst)) return r;
  if (!(tol > 0.0)) return fail(FSX_EINVAL, "Newton tolerance must be > 0");
  c->hp[inst].newton_tol = tol;
  c->dirty = true;
  return 0;
}

extern "C" int fsx_init_h(fsx_ctx* c, int inst, const double* h, int on_device) {
  if (int r = check_inst(c, inst)) return r;
  if (!h) return fail(FSX_EINVAL, "null h");
  const InstParams& P = c->hp[inst];
  if (int r = put(c, c->S.h + P.off, h, P.nn, on_device)) return r;
  return put(c, c->S.b + P.off, h, P.nn, on_device);
}

extern "C" int fsx_set_field(fsx_ctx* c, int inst, int field, const double* src, int on_device) {
  if (int r = check_inst(c, inst)) return r;
  if (!src) return fail(FSX_EINVAL, "null source");
  InstParams& P = c->hp[inst];
  double* dst = nullptr;
  switch (field) {
    case FSX_H: dst = c->S.h; break;
    case FSX_B: dst = c->S.b; break;
    case FSX_U: dst = c->S.u; P.run_uplift = 1; c->dirty = true; break;
    case FSX_PRECIP: dst = c->S.precip; break;
    case FSX_FMIX: dst = c->S.fmix; break;
    default: return fail(FSX_EINVAL, "field cannot be set");
  }
  return put(c, dst + P.off, src, P.nn, on_device);
}

extern "C" int fsx_set_v(fsx_ctx* c, int inst, const double* vx, const double* vy, int on_device) {
  if (int r = check_inst(c, inst)) return r;
  if (!vx || !vy) return fail(FSX_EINVAL, "vx and vy are required");
  if (int r = advect_alloc(c)) return r;
  InstParams& P = c->hp[inst];
  if (int r = put(c, c->S.vx + P.off, vx, P.nn, on_device)) return r;
  if (int r = put(c, c->S.vy + P.off, vy, P.nn, on_device)) return r;
  P.run_advect = 1;
  c->dirty = true;
  return 0;
}

__global__ void k_add_layers(double* __restrict__ h, double* __restrict__ b, const double* __restrict__ dh, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    h[i] = h[i] + dh[i];
    b[i] = b[i] + dh[i];
  }
}

extern "C" int fsx_set_all_layers(fsx_ctx* c, int inst, const double* dh, int on_device) {
  if (int r = check_inst(c, inst)) return r;
  if (!dh) return fail(FSX_EINVAL, "null dh");
  const InstParams& P = c->hp[inst];
  if (int r = put(c, c->S.zint + P.off, dh, P.nn, on_device)) return r;
  k_add_layers<<<(P.nn + 255) / 256, 256, 0, c->stream>>>(c->S.h + P.off, c->S.b + P.off, c->S.zint + P.off, P.nn);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(c->stream));
  return 0;
}

extern "C" int fsx_reset_cumulative_erosion(fsx_ctx* c, int inst) {
  if (int r = check_inst(c, inst)) return r;
  const InstParams& P = c->hp[inst];
  HIPCHK(hipMemsetAsync(c->S.etot + P.off, 0, sizeof(double) * P.nn, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return 0;
}

extern "C" int fsx_set_active(fsx_ctx* c, int inst, int active) {
  if (int r = check_inst(c, inst)) return r;
  c->hp[inst].active = active ? 1 : 0;
  c->dirty = true;
  return 0;
}

static int check_errors(fsx_ctx* c) {
  std::vector<int> e(c->S.ninst);
  HIPCHK(hipMemcpyAsync(e.data(), c->S.err, sizeof(int) * e.size(), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  for (int k = 0; k < c->S.ninst; k++) {
    if (!e[k]) continue;
    HIPCHK(hipMemset(c->S.err, 0, sizeof(int) * e.size()));
    std::string why;
    if (e[k] & FSX_ERR_NCONN) why += " basin graph exceeds 6*nbasins edges;";
    if (e[k] & FSX_ERR_NO_OUTLET) why += " no basin drains to a fixed boundary node;";
    if (e[k] & FSX_ERR_TRIDAG) why += " zero pivot in the diffusion solve;";
    if (e[k] & FSX_ERR_NEWTON) why += " Newton iteration did not converge;";
    if (e[k] & FSX_ERR_SORTSTACK) why += " edge-sort segment stack overflow;";
    if (e[k] & FSX_ERR_LEVELS) why += " receiver graph is not a forest;";
    if (e[k] & FSX_ERR_MFD) why += " multiple-flow order incomplete;";
    if (e[k] & FSX_ERR_MARINE) why += " marine silt / sand diffusion did not converge in 1000 iterations;";
    if (e[k] & FSX_ERR_NONFINITE) why += " elevation is not a number;";
    return fail(FSX_ESTEP, "instance " + std::to_string(k) + ":" + why);
  }
  return 0;
}

// uplift != 0: the step's uplift runs fused into the receiver search (Uplift precedes FlowRouting,
// src/FastScape_api.f90:216-226)
static void enqueue_routing(fsx_ctx* c, hipStream_t st, int uplift) {
  DevState& S = c->S;
  PHASE(PH_RECEIVERS, launch_receivers(S, st, uplift));
  PHASE(PH_DONORS, launch_donors(S, st));
  // LocalMinima, src/LocalMinima.f90:1-132
  PHASE(PH_LM_LABELS, launch_lm_labels(S, st));
  PHASE(PH_LM_BASINS, launch_lm_basins(S, st));
  PHASE(PH_LM_WALK, launch_lm_walk(S, st));
  PHASE(PH_LM_EDGES, launch_lm_edges(S, st));
  PHASE(PH_LM_SORT, launch_nr_sort(S, st));
  PHASE(PH_LM_MST, launch_lm_mst(S, st));
  PHASE(PH_LM_ORIENT, launch_lm_orient(S, st));
  PHASE(PH_LM_CORRECT, launch_lm_correct(S, st));
  PHASE(PH_DONORS, launch_donors(S, st));
  PHASE(PH_LEVELS, launch_levels(S, st));
  if (c->any_sfd) {
    PHASE(PH_AREA_PREP, launch_area_prep(S, st));
    PHASE(PH_AREA_SWEEP, launch_area_sweep(S, st));
  }
  // multiple flow direction: fill, receivers, Kahn layers, weighted accumulation
  if (c->any_mfd) PHASE(PH_MFD_ROUTE, launch_mfd_route(S, st));
}

static void enqueue_route_only(fsx_ctx* c, hipStream_t st) {
  enqueue_routing(c, st, 0);
  launch_spl_prep(c->S, st);
  if (c->any_sfd) launch_fill(c->S, st);  // lake surface / depth without eroding
}

static void enqueue_step(fsx_ctx* c, hipStream_t st) {
  DevState& S = c->S;
  if (c->any_advect) PHASE(PH_ADVECT, launch_advect(S, st));  // Advect precedes Uplift (api.f90:209-221)
  enqueue_routing(c, st, 1);
  PHASE(PH_SPL_PREP, launch_spl_prep(S, st));
  if (c->any_sfd) {
    PHASE(PH_SPL_FACT, launch_spl_fact(S, st));
    PHASE(PH_SPL_SWEEP, launch_spl_sweep(S, st));  // also fills hwater / lake_depth
  }
  if (c->any_mfd) PHASE(PH_MFD_SPL, launch_mfd_spl(S, st));
  PHASE(PH_SPL_POST, launch_lake_acc(S, st); launch_spl_post(S, st));
  PHASE(PH_DIFFUSION, launch_diffusion(S, st));
  if (c->any_marine) PHASE(PH_MARINE, launch_marine(S, st));  // Marine follows Diffusion (api.f90:245-250)
  if (c->any_strati) PHASE(PH_STRATI, launch_strati_step(S, st));  // then run_Strati (api.f90:252-257)
}

static int capture(fsx_ctx* c, hipStream_t st, int lo, int hi, hipGraphExec_t* out) {
  hipGraph_t g = nullptr;
  HIPCHK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
  c->emit_seq = 0;
  c->emit_lo = lo;
  c->emit_hi = hi;
  enqueue_step(c, st);
  c->emit_lo = 0;
  c->emit_hi = 1 << 30;
  const hipError_t ec = hipStreamEndCapture(st, &g);
  if (ec != hipSuccess) return fail(FSX_EHIP, std::string("graph capture: ") + hipGetErrorString(ec));
  const hipError_t ei = hipGraphInstantiate(out, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (ei != hipSuccess) return fail(FSX_EHIP, std::string("graph instantiate: ") + hipGetErrorString(ei));
  return 0;
}

// One graph for the whole step, or (a phase timed live) the graphs before and after that phase.
static int build_graphs(fsx_ctx* c, hipStream_t st) {
  // an executable graph is destroyed only once no launch of it is queued (the last steps may have been
  // enqueued on another stream without a wait)
  HIPCHK(hipStreamSynchronize(st));
  HIPCHK(quiesce(c));
  for (hipGraphExec_t* g : {&c->gexec, &c->gpre, &c->gpost})
    if (*g) {
      HIPCHK(hipGraphExecDestroy(*g));
      *g = nullptr;
    }
  int sx = -1, nseq = 0;
  if (c->timed_phase >= 0) {
    std::vector<int> seq;
    c->seq_log = &seq;
    c->emit_seq = 0;
    enqueue_step(c, st);
    c->seq_log = nullptr;
    nseq = (int)seq.size();
    for (int i = 0; i < nseq && sx < 0; i++)
      if (seq[i] == c->timed_phase) sx = i;
  }
  if (sx < 0) {
    if (int r = capture(c, st, 0, 1 << 30, &c->gexec)) return r;
  } else {
    if (c->ev.empty()) {
      c->ev.resize((size_t)kRing * PH_COUNT * 2);
      for (auto& e : c->ev) HIPCHK(hipEventCreate(&e));
      c->rec_mask.assign((size_t)kRing * PH_COUNT, 0);
    }
    if (sx > 0)
      if (int r = capture(c, st, 0, sx, &c->gpre)) return r;
    if (sx + 1 < nseq)
      if (int r = capture(c, st, sx + 1, 1 << 30, &c->gpost)) return r;
    c->timed_seq = sx;
  }
  c->graph_valid = true;
  c->g_sfd = c->any_sfd;
  c->g_mfd = c->any_mfd;
  c->g_sweep = c->S.nsweep > 0;
  c->g_adv = c->any_advect;
  c->g_mar = c->any_marine;
  c->g_str = c->any_strati;
  c->g_timed = c->timed_phase;
  return 0;
}

extern "C" int fsx_execute_steps(fsx_ctx* c, int nsteps, void* stream, int sync) {
  if (!c) return fail(FSX_EINVAL, "null context");
  if (nsteps < 0) return fail(FSX_EINVAL, "nsteps < 0");
  hipStream_t st = stream ? (hipStream_t)stream : c->stream;
  // nothing to step in a context without landscapes (no launch may have an empty grid)
  bool any = false;
  for (int k = 0; k < c->S.ninst && !any; k++) any = c->used[k] != 0;
  if (!any) return 0;
  // Marine reads the step's single-flow routing (rec, stack; src/Marine.f90:30-36,60-61)
  for (int k = 0; k < c->S.ninst; k++)
    if (c->used[k] && c->hp[k].active && c->hp[k].run_marine && !c->hp[k].run_spl)
      return fail(FSX_EINVAL, "marine transport needs the flow routing: set the erosional parameters first");
  // the parameter table is rewritten in place: launches still queued on either stream read it
  if (c->dirty && c->last && c->last != c->stream) HIPCHK(hipStreamSynchronize(c->last));
  if (stream && c->dirty) HIPCHK(hipStreamSynchronize(c->stream));
  if (int r = upload_params(c)) return r;
  if (stream) HIPCHK(hipStreamSynchronize(c->stream));
  if (stream) c->last = st;
  if (c->use_graph && !c->prof && nsteps > 0) {
    // the step's launch sequence depends on the layout and on which routing paths are in use; the
    // kernels read every other parameter from device memory, so one capture serves until those change
    if (!c->graph_valid || c->g_sfd != c->any_sfd || c->g_mfd != c->any_mfd || c->g_sweep != (c->S.nsweep > 0) ||
        c->g_adv != c->any_advect ||
        c->g_mar != c->any_marine || c->g_str != c->any_strati || c->g_timed != c->timed_phase) {
      if (int r = build_graphs(c, st)) return r;
    }
    if (c->gexec) {
      for (int s = 0; s < nsteps; s++) HIPCHK(hipGraphLaunch(c->gexec, st));
    } else {
      for (int s = 0; s < nsteps; s++) {
        if (c->gpre) HIPCHK(hipGraphLaunch(c->gpre, st));
        const int k = c->ring_used * PH_COUNT + c->timed_phase;
        HIPCHK(hipEventRecord(c->ev[2 * k], st));
        c->emit_seq = 0;
        c->emit_lo = c->timed_seq;
        c->emit_hi = c->timed_seq + 1;
        enqueue_step(c, st);
        c->emit_lo = 0;
        c->emit_hi = 1 << 30;
        HIPCHK(hipEventRecord(c->ev[2 * k + 1], st));
        c->rec_mask[k] = 1;
        if (c->gpost) HIPCHK(hipGraphLaunch(c->gpost, st));
        if (++c->ring_used == kRing)
          if (int r = flush_timers_any(c, st)) return r;
      }
    }
  } else {
    c->eager_timed = c->prof ? -1 : c->timed_phase;
    for (int s = 0; s < nsteps; s++) {
      c->emit_seq = 0;
      enqueue_step(c, st);
      if ((c->prof || c->eager_timed >= 0) && ++c->ring_used == kRing)
        if (int r = flush_timers(c, st)) return r;
    }
    c->eager_timed = -1;
  }
  HIPCHK(hipGetLastError());
  for (int k = 0; k < c->S.ninst; k++)
    if (c->hp[k].active) {
      c->steps[k] += nsteps;
      if (c->hp[k].run_spl && nsteps > 0) c->routed[k] = 1;
    }
  if (sync) {
    HIPCHK(hipStreamSynchronize(st));
    return check_errors(c);
  }
  return 0;
}

extern "C" int fsx_time_phase(fsx_ctx* c, int phase) {
  if (!c) return fail(FSX_EINVAL, "null context");
  if (phase < -1 || phase >= PH_COUNT) return fail(FSX_EINVAL, "unknown phase");
  if (int r = flush_timers_any(c, c->stream)) return r;
  if (phase >= 0 && c->ev.empty()) {
    c->ev.resize((size_t)kRing * PH_COUNT * 2);
    for (auto& e : c->ev) HIPCHK(hipEventCreate(&e));
    c->rec_mask.assign((size_t)kRing * PH_COUNT, 0);
  }
  c->timed_phase = phase;
  return 0;
}

extern "C" int fsx_set_graphs(fsx_ctx* c, int on) {
  if (!c) return fail(FSX_EINVAL, "null context");
  c->use_graph = on != 0;
  return 0;
}

extern "C" int fsx_synchronize(fsx_ctx* c) {
  if (!c) return fail(FSX_EINVAL, "null context");
  HIPCHK(hipDeviceSynchronize());
  return check_errors(c);
}

extern "C" int fsx_query(fsx_ctx* c) {
  if (!c) return fail(FSX_EINVAL, "null context");
  const hipError_t e = hipStreamQuery(c->stream);
  if (e == hipErrorNotReady) return 0;
  HIPCHK(e);
  if (int r = check_errors(c)) return r;
  return 1;
}

extern "C" int fsx_route(fsx_ctx* c, void* stream, int sync) {
  if (!c) return fail(FSX_EINVAL, "null context");
  hipStream_t st = stream ? (hipStream_t)stream : c->stream;
  if (int r = upload_params(c)) return r;
  if (stream) HIPCHK(hipStreamSynchronize(c->stream));
  enqueue_route_only(c, st);
  HIPCHK(hipGetLastError());
  for (int k = 0; k < c->S.ninst; k++)
    if (c->hp[k].active && c->hp[k].run_spl) c->routed[k] = 1;
  if (sync) {
    HIPCHK(hipStreamSynchronize(st));
    return check_errors(c);
  }
  return 0;
}

extern "C" int fsx_copy_field(fsx_ctx* c, int inst, int field, double* dst, int on_device) {
  if (int r = check_inst(c, inst)) return r;
  if (!dst) return fail(FSX_EINVAL, "null destination");
  const InstParams& P = c->hp[inst];
  const double* src = nullptr;
  switch (field) {
    case FSX_H: src = c->S.h; break;
    case FSX_B: src = c->S.b; break;
    case FSX_U: src = c->S.u; break;
    case FSX_PRECIP: src = c->S.precip; break;
    case FSX_ETOT: src = c->S.etot; break;
    case FSX_ERATE: src = c->S.erate; break;
    case FSX_AREA: src = c->S.area; break;
    case FSX_LAKE_DEPTH: src = c->S.lake_depth; break;
    case FSX_SEDFLUX: src = c->S.sedflux; break;
    case FSX_LENGTH: src = c->S.length; break;
    case FSX_HWATER: src = c->S.hwater; break;
    case FSX_KF: src = c->S.kf; break;
    case FSX_KD: src = c->S.kd; break;
    case FSX_FMIX: src = c->S.fmix; break;
    default: return fail(FSX_EINVAL, "unknown field");
  }
  HIPCHK(hipMemcpyAsync(dst, src + P.off, sizeof(double) * P.nn, on_device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost,
                        c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return 0;
}

static void launch_derived(fsx_ctx* c, const int* ids, int n, int inst1, int which, double* out) {
  if (which == FSX_CHI)
    launch_chi(c->S, ids, n, inst1, out, c->stream);
  else if (which == FSX_CATCHMENT)
    launch_catchment(c->S, ids, n, inst1, out, c->stream);
  else
    launch_terrain(c->S, ids, n, inst1, which == FSX_SLOPE ? 0 : 1, 45.0 / std::atan(1.0), out, c->stream);
}

// the landscapes selected by sel[k] != 0 (every landscape when sel is null), checked, as a device list
static int readout_list(fsx_ctx* c, const std::vector<char>& sel, bool need_routing, int* n) {
  std::vector<int> ids;
  for (int k = 0; k < c->S.ninst; k++) {
    if (!sel[k]) continue;
    if (!c->used[k]) return fail(FSX_EINVAL, "selected slot holds no landscape");
    if (need_routing && !c->routed[k]) return fail(FSX_EINVAL, "a selected landscape has no routed step");
    ids.push_back(k);
  }
  *n = (int)ids.size();
  if (ids.empty()) return 0;
  if (int r = upload_params(c)) return r;
  HIPCHK(hipMemcpyAsync(c->S.sB, ids.data(), sizeof(int) * ids.size(), hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));  // ids is a host temporary
  return 0;
}

extern "C" int fsx_copy_derived_batch(fsx_ctx* c, int which, double* const* dst, int on_device) {
  if (!c) return fail(FSX_EINVAL, "null context");
  if (!dst) return fail(FSX_EINVAL, "null destination list");
  if (which < FSX_SLOPE || which > FSX_CATCHMENT) return fail(FSX_EINVAL, "unknown derived field");
  std::vector<char> sel(c->S.ninst);
  for (int k = 0; k < c->S.ninst; k++) sel[k] = dst[k] != nullptr;
  int n = 0;
  if (int r = readout_list(c, sel, which >= FSX_CHI, &n)) return r;
  if (!n) return 0;
  launch_derived(c, c->S.sB, n, -1, which, c->S.zint);  // S.sB: 8 ints per node of scratch outside a step
  HIPCHK(hipGetLastError());
  const hipMemcpyKind kind = on_device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
  for (int k = 0; k < c->S.ninst; k++)
    if (sel[k])
      HIPCHK(hipMemcpyAsync(dst[k], c->S.zint + c->hp[k].off, sizeof(double) * c->hp[k].nn, kind, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return 0;
}

extern "C" int fsx_get_fluxes_batch(fsx_ctx* c, double* out) {
  if (!c) return fail(FSX_EINVAL, "null context");
  if (!out) return fail(FSX_EINVAL, "null destination");
  std::vector<char> sel(c->S.ninst);
  for (int k = 0; k < c->S.ninst; k++) {
    sel[k] = c->used[k] && c->routed[k] && (c->hp[k].sfd || c->S.mrec);
    for (int q = 0; q < 3; q++) out[3 * k + q] = std::nan("");
  }
  int n = 0;
  if (int r = readout_list(c, sel, true, &n)) return r;
  if (!n) return 0;
  double* sums = c->S.zintp;  // 3 per listed landscape (nn >= 9 nodes each)
  launch_fluxes(c->S, c->S.sB, n, -1, c->S.zint, sums, c->stream);
  HIPCHK(hipGetLastError());
  std::vector<double> s3(3 * (size_t)n);
  HIPCHK(hipMemcpyAsync(s3.data(), sums, sizeof(double) * s3.size(), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  for (int k = 0, i = 0; k < c->S.ninst; k++) {
    if (!sel[k]) continue;
    const InstParams& P = c->hp[k];
    const double surf = c->xl[k] * c->yl[k] / (double)(P.nx - 1) / (double)(P.ny - 1);
    for (int q = 0; q < 3; q++) out[3 * k + q] = s3[3 * i + q] * surf;
    i++;
  }
  return 0;
}

extern "C" int fsx_copy_derived(fsx_ctx* c, int inst, int which, double* dst, int on_device) {
  if (int r = check_inst(c, inst)) return r;
  if (!dst) return fail(FSX_EINVAL, "null destination");
  if (which < FSX_SLOPE || which > FSX_CATCHMENT) return fail(FSX_EINVAL, "unknown derived field");
  if (which >= FSX_CHI && !c->routed[inst]) return fail(FSX_EINVAL, "chi and catchment need the routing of a step first");
  if (int r = upload_params(c)) return r;
  const InstParams& P = c->hp[inst];
  double* tmp = c->S.zint + P.off;  // scratch outside a step
  launch_derived(c, nullptr, P.nn, inst, which, tmp);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(dst, tmp, sizeof(double) * P.nn, on_device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost,
                        c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return 0;
}

extern "C" int fsx_get_fluxes(fsx_ctx* c, int inst, double* out3) {
  if (int r = check_inst(c, inst)) return r;
  if (!out3) return fail(FSX_EINVAL, "null destination");
  if (!c->routed[inst]) return fail(FSX_EINVAL, "fluxes need the routing of a step first");
  const InstParams& P = c->hp[inst];
  if (!P.sfd && !c->S.mrec) return fail(FSX_EINVAL, "no multiple-flow routing in place");
  if (int r = upload_params(c)) return r;
  double* sums = c->S.zintp + P.off;  // scratch outside a step (nn >= 9 entries)
  launch_fluxes(c->S, nullptr, P.nn, inst, c->S.zint + P.off, sums, c->stream);
  HIPCHK(hipGetLastError());
  double s3[3];
  HIPCHK(hipMemcpyAsync(s3, sums, sizeof s3, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  const double surf = c->xl[inst] * c->yl[inst] / (double)(P.nx - 1) / (double)(P.ny - 1);
  for (int k = 0; k < 3; k++) out3[k] = s3[k] * surf;
  return 0;
}

__global__ void k_expand_donors(const InstParams* __restrict__ PP, int inst, const uint8_t* __restrict__ dmask,
                                int* __restrict__ ndon, int* __restrict__ don) {
  const InstParams& P = PP[inst];
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= P.nn) return;
  int d[8];
  const int nd = donors_sorted(P, v, dmask[P.off + v], d);
  if (ndon) ndon[v] = nd;
  if (don)
    for (int k = 0; k < 8; k++) don[8 * v + k] = k < nd ? d[k] : -1;
}

extern "C" int fsx_copy_ifield(fsx_ctx* c, int inst, int field, int* dst, int on_device) {
  if (int r = check_inst(c, inst)) return r;
  if (!dst) return fail(FSX_EINVAL, "null destination");
  const InstParams& P = c->hp[inst];
  const hipMemcpyKind kind = on_device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
  int* tmp = c->S.sB;  // 8*ntot ints of scratch, free outside a step
  switch (field) {
    case FSX_REC:
      HIPCHK(hipMemcpyAsync(dst, c->S.rec + P.off, sizeof(int) * P.nn, kind, c->stream));
      break;
    case FSX_ORDER:
      HIPCHK(hipMemcpyAsync(dst, c->S.order + P.off, sizeof(int) * P.nn, kind, c->stream));
      break;
    case FSX_NDON:
      k_expand_donors<<<(P.nn + 255) / 256, 256, 0, c->stream>>>(c->S.P, inst, c->S.dmask, tmp, nullptr);
      HIPCHK(hipMemcpyAsync(dst, tmp, sizeof(int) * P.nn, kind, c->stream));
      break;
    case FSX_DON:
      k_expand_donors<<<(P.nn + 255) / 256, 256, 0, c->stream>>>(c->S.P, inst, c->S.dmask, nullptr, tmp);
      HIPCHK(hipMemcpyAsync(dst, tmp, sizeof(int) * 8 * (size_t)P.nn, kind, c->stream));
      break;
    case FSX_MREC:
    case FSX_MNREC:
    case FSX_MSTACK:
      if (P.sfd || !c->S.mrec) return fail(FSX_EINVAL, "instance does not use multiple flow direction");
      if (field == FSX_MREC)
        HIPCHK(hipMemcpyAsync(dst, c->S.mrec + 8 * (size_t)P.off, sizeof(int) * 8 * (size_t)P.nn, kind, c->stream));
      else
        HIPCHK(hipMemcpyAsync(dst, (field == FSX_MNREC ? c->S.mnrec : c->S.mstack) + P.off, sizeof(int) * P.nn, kind,
                              c->stream));
      break;
    default:
      return fail(FSX_EINVAL, "unknown integer field");
  }
  HIPCHK(hipStreamSynchronize(c->stream));
  return 0;
}

extern "C" int fsx_get_step(fsx_ctx* c, int inst, int* step) {
  if (int r = check_inst(c, inst)) return r;
  if (step) *step = c->steps[inst];
  return 0;
}

extern "C" int fsx_get_sizes(fsx_ctx* c, int inst, int* nx, int* ny) {
  if (int r = check_inst(c, inst)) return r;
  if (nx) *nx = c->hp[inst].nx;
  if (ny) *ny = c->hp[inst].ny;
  return 0;
}

extern "C" int fsx_get_stats(fsx_ctx* c, int inst, long long* out) {
  if (int r = check_inst(c, inst)) return r;
  int v[5];
  long long nw = 0;
  HIPCHK(hipStreamSynchronize(c->stream));
  HIPCHK(hipMemcpy(&v[0], c->S.nbasins + inst, sizeof(int), hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(&v[1], c->S.nconn + inst, sizeof(int), hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(&v[2], c->S.ntree + inst, sizeof(int), hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(&v[3], c->S.nlev + inst, sizeof(int), hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(&v[4], c->S.ngs + inst, sizeof(int), hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(&nw, c->S.newton + inst, sizeof(long long), hipMemcpyDeviceToHost));
  for (int k = 0; k < 5; k++) out[k] = v[k];
  out[5] = nw;
  return 0;
}

extern "C" int fsx_set_profiling(fsx_ctx* c, int on) {
  if (!c) return fail(FSX_EINVAL, "null context");
  if (on && c->ev.empty()) {
    c->ev.resize((size_t)kRing * PH_COUNT * 2);
    for (auto& e : c->ev) HIPCHK(hipEventCreate(&e));
    c->rec_mask.assign((size_t)kRing * PH_COUNT, 0);
  }
  if (!on && c->prof) {
    if (int r = flush_timers(c, c->stream)) return r;
  }
  c->prof = on != 0;
  return 0;
}

extern "C" int fsx_get_timers(fsx_ctx* c, double* ms, long long* counts, int n) {
  if (!c) return fail(FSX_EINVAL, "null context");
  HIPCHK(hipDeviceSynchronize());
  if (int r = flush_timers(c, c->stream)) return r;
  for (int p = 0; p < n && p < PH_COUNT; p++) {
    if (ms) ms[p] = c->acc_ms[p];
    if (counts) counts[p] = c->acc_n[p];
  }
  return PH_COUNT;
}

extern "C" int fsx_reset_timers(fsx_ctx* c) {
  if (!c) return fail(FSX_EINVAL, "null context");
  for (int p = 0; p < PH_COUNT; p++) { c->acc_ms[p] = 0; c->acc_n[p] = 0; }
  return 0;
}

// The candidates of each launch shape, the phase(s) that time it, and the shape's field.
struct TuneKnob {
  int DevState::*field;
  int cand[4];
  int ncand;
  int ph0, ph1;  // timed phases (ph1 = -1: one)
};

extern "C" int fsx_autotune(fsx_ctx* c, int reps, int* chosen) {
  if (!c) return fail(FSX_EINVAL, "null context");
  if (reps < 1) return fail(FSX_EINVAL, "fsx_autotune: reps must be >= 1");
  if (int r = upload_params(c)) return r;
  static const TuneKnob knobs[] = {
      {&DevState::walk_gx, {64, 128, 256, 512}, 4, PH_LM_WALK, -1},
      {&DevState::jump_blocks, {1024, 2048, 4096, 8192}, 4, PH_LM_LABELS, PH_LEVELS},
      {&DevState::tour_gx, {8, 16, 32, 0}, 3, PH_LM_ORIENT, -1},
      {&DevState::nr_small_gx, {16, 32, 64, 128}, 4, PH_LM_SORT, -1},
      {&DevState::area_wg, {256, 512, 1024, 0}, 3, PH_AREA_SWEEP, -1},
      {&DevState::bor_gx, {2, 4, 8, 16}, 4, PH_LM_MST, -1},
      {&DevState::flow_nk, {4, 8, 16, 32}, 4, PH_SPL_SWEEP, -1},
      {&DevState::flow_wg, {256, 512, 1024, 0}, 3, PH_SPL_SWEEP, -1},
      {&DevState::flow_split, {0, 262144, 131072, 65536}, 4, PH_SPL_SWEEP, -1},
      {&DevState::walk_big, {64, 256, 1024, 8192}, 4, PH_LM_WALK, -1},
      {&DevState::mfd_split, {262144, 131072, 65536, 32768}, 4, PH_MFD_ROUTE, -1},
      {&DevState::tile_h, {4, 8, 16, 32}, 4, PH_RECEIVERS, PH_DONORS},
  };
  // keep the caller's profiling state and accumulated timers
  if (int r = flush_timers_any(c, c->stream)) return r;
  const bool prof0 = c->prof;
  double ms0[PH_COUNT];
  long long n0[PH_COUNT];
  for (int p = 0; p < PH_COUNT; p++) { ms0[p] = c->acc_ms[p]; n0[p] = c->acc_n[p]; }
  int rc = fsx_set_profiling(c, 1);
  for (const TuneKnob& k : knobs) {
    if (rc) break;
    // the stream-power sweep's shape is timed on the dataflow landscapes' solve of the routed state,
    // without the step's per-landscape outputs (Newton counts) and without updating h
    const bool spl = k.ph0 == PH_SPL_SWEEP;
    if (spl && !c->any_flow) continue;
    if (k.ph0 == PH_MFD_ROUTE && !c->any_mfd) continue;
    auto pass = [&]() {
      enqueue_routing(c, c->stream, 0);
      if (spl) {
        hipStream_t st = c->stream;
        launch_spl_fact(c->S, st);
        PHASE(PH_SPL_SWEEP, launch_spl_flow(c->S, st, 2));
      }
    };
    int best = c->S.*k.field;
    double best_ms = 1e300;
    // these shapes change a work table (rebuilt by upload_params)
    const bool split =
        k.field == &DevState::flow_split || k.field == &DevState::mfd_split || k.field == &DevState::tile_h;
    for (int i = 0; i < k.ncand && !rc; i++) {
      c->S.*k.field = k.cand[i];
      if (split) {
        c->dirty = true;
        if ((rc = upload_params(c))) break;
      }
      pass();  // untimed pass: caches and code warm for this shape
      ++c->ring_used;
      if ((rc = fsx_get_timers(c, nullptr, nullptr, 0)) < 0) break;
      rc = 0;
      fsx_reset_timers(c);
      for (int r = 0; r < reps && !rc; r++) {
        pass();
        if (++c->ring_used == kRing) rc = flush_timers_any(c, c->stream);  // one ring slot per pass
      }
      if ((rc = fsx_get_timers(c, nullptr, nullptr, 0)) < 0) break;
      rc = 0;
      const double t = c->acc_ms[k.ph0] + (k.ph1 >= 0 ? c->acc_ms[k.ph1] : 0.0);
      if (t < best_ms) { best_ms = t; best = k.cand[i]; }
    }
    c->S.*k.field = best;
    if (split && !rc) {
      c->dirty = true;
      rc = upload_params(c);
    }
  }
  if (!rc) rc = check_errors(c);
  const int rp = fsx_set_profiling(c, prof0 ? 1 : 0);
  for (int p = 0; p < PH_COUNT; p++) { c->acc_ms[p] = ms0[p]; c->acc_n[p] = n0[p]; }
  c->graph_valid = false;  // captured launches carry the old shapes
  if (rc) return rc;
  if (rp) return rp;
  if (chosen) fsx_get_tuning(c, chosen);
  return 0;
}

extern "C" int fsx_get_tuning(fsx_ctx* c, int* out) {
  if (!c || !out) return fail(FSX_EINVAL, "fsx_get_tuning: null argument");
  out[0] = c->S.walk_gx;
  out[1] = c->S.jump_blocks;
  out[2] = c->S.tour_gx;
  out[3] = c->S.nr_small_gx;
  out[4] = c->S.area_wg;
  out[5] = c->S.bor_gx;
  out[6] = c->S.flow_wg;
  out[7] = c->S.flow_split;
  out[8] = c->S.walk_big;
  out[9] = c->S.mfd_split;
  out[10] = c->S.tile_h;
  out[11] = c->S.flow_nk;
  return 0;
}

extern "C" int fsx_set_tuning(fsx_ctx* c, const int* in) {
  if (!c || !in) return fail(FSX_EINVAL, "fsx_set_tuning: null argument");
  static const int lo[4] = {1, 1, 1, 1}, hi[4] = {4096, 1 << 20, 1024, 4096};
  for (int i = 0; i < 4; i++)
    if (in[i] < lo[i] || in[i] > hi[i]) return fail(FSX_EINVAL, "fsx_set_tuning: launch shape out of range");
  if (in[4] != 256 && in[4] != 512 && in[4] != 1024)
    return fail(FSX_EINVAL, "fsx_set_tuning: area_wg must be 256, 512 or 1024");
  if (in[5] < 1 || in[5] > 1024) return fail(FSX_EINVAL, "fsx_set_tuning: bor_gx must be in [1, 1024]");
  if (in[6] != 256 && in[6] != 512 && in[6] != 1024)
    return fail(FSX_EINVAL, "fsx_set_tuning: flow_wg must be 256, 512 or 1024");
  if (in[7] < 0) return fail(FSX_EINVAL, "fsx_set_tuning: flow_split must be >= 0");
  if (in[8] < 64) return fail(FSX_EINVAL, "fsx_set_tuning: walk_big must be >= 64");
  if (in[9] < 1024) return fail(FSX_EINVAL, "fsx_set_tuning: mfd_split must be >= 1024");
  if (in[10] != 4 && in[10] != 8 && in[10] != 16 && in[10] != 32)
    return fail(FSX_EINVAL, "fsx_set_tuning: tile_h must be 4, 8, 16 or 32");
  if (in[11] < 1 || in[11] > 1024) return fail(FSX_EINVAL, "fsx_set_tuning: flow_nk must be in [1, 1024]");
  c->S.walk_gx = in[0];
  c->S.jump_blocks = in[1];
  c->S.tour_gx = in[2];
  c->S.nr_small_gx = in[3];
  c->S.area_wg = in[4];
  c->S.bor_gx = in[5];
  c->S.flow_wg = in[6];
  if (c->S.flow_split != in[7]) c->dirty = true;  // the sweep's workgroup table follows at the next upload
  c->S.flow_split = in[7];
  c->S.walk_big = in[8];
  if (c->S.mfd_split != in[9]) c->dirty = true;  // so does the multiple-flow sweeps' table
  c->S.mfd_split = in[9];
  if (c->S.tile_h != in[10]) c->dirty = true;  // and the stencil kernels' tile table
  c->S.tile_h = in[10];
  c->S.flow_nk = in[11];
  c->graph_valid = false;  // captured launches carry the old shapes
  return 0;
}

extern "C" int fsx_memory(fsx_ctx* c, long long* bytes, long long* per_node) {
  if (!c) return fail(FSX_EINVAL, "null context");
  if (bytes) *bytes = c->bytes;
  if (per_node) *per_node = c->S.ntot ? c->bytes / c->S.ntot : 0;
  return 0;
}

extern "C" int fsx_indexx(int n, const double* keys, int* perm) {
  if (n < 0 || (n > 0 && (!keys || !perm))) return fail(FSX_EINVAL, "fsx_indexx: bad arguments");
  if (n == 0) return 0;
  // a one-landscape context sized so that its edge capacity holds n keys
  const int nn = (n + FSX_EDGE_CAP - 1) / FSX_EDGE_CAP + 3;
  int nx = 3, ny = std::max(3, (nn + 2) / 3);
  fsx_ctx* c = nullptr;
  if (int r = fsx_create(1, &nx, &ny, &c)) return r;
  c->hp[0].run_spl = 1;
  c->dirty = true;
  int rc = upload_params(c);
  hipError_t e = hipSuccess;
  if (!rc) e = hipMemcpy(c->S.cw, keys, sizeof(double) * n, hipMemcpyHostToDevice);
  if (!rc && e == hipSuccess) e = hipMemcpy(c->S.nconn, &n, sizeof(int), hipMemcpyHostToDevice);
  if (!rc && e == hipSuccess) {
    launch_nr_sort(c->S, c->stream);
    e = hipGetLastError();
  }
  std::vector<int> sid(n);
  if (!rc && e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (!rc && e == hipSuccess) e = hipMemcpy(sid.data(), c->S.sid, sizeof(int) * n, hipMemcpyDeviceToHost);
  if (!rc && e != hipSuccess) rc = fail(FSX_EHIP, std::string("fsx_indexx: ") + hipGetErrorString(e));
  if (!rc) rc = check_errors(c);
  if (!rc)
    for (int p = 0; p < n; p++) perm[p] = sid[p] + 1;
  fsx_destroy(c);
  return rc;
}
