// fsx_noalloc.hip -- the exact Numerical Recipes edge sort (fsx_indexx) without any device allocation,
// for callers that cannot allocate (code under HIP graph capture, tight serving loops).
//
// fsx_indexx builds a one-landscape context (hipMalloc) for its working arrays.  Here the same kernels
// run on a fixed working area in a file-scope __device__ array: each GPU the library runs on holds
// one copy of it (≈ 19 MB) until the process ends.  Host code never passes the array's name to a
// kernel (in host code it is only a stand-in); its device address is looked up once per device with
// hipGetSymbolAddress and that address is used for every launch.
#include <algorithm>
#include <mutex>

#include "../../include/fastscape_hip.h"
#include "fsx_state.h"

namespace {

constexpr long long kMax = FSX_INDEXX_NOALLOC_MAX;  // keys per call
constexpr long long kPad = 64;                      // the sort's level lists run 2 words past n in cw
constexpr long long align256(long long x) { return (x + 255) / 256 * 256; }
// byte offsets of the working arrays inside the area
constexpr long long oP = 0;
constexpr long long oInt = align256(oP + (long long)sizeof(InstParams));  // nconn, nsegs, err, segctr, nrst[32]
constexpr long long oCw = align256(oInt + 64 * 4);
constexpr long long oSkey = align256(oCw + 8 * (kMax + kPad));
constexpr long long oSid = align256(oSkey + 8 * (kMax + kPad));
constexpr long long oSegs = align256(oSid + 4 * (kMax + kPad));
constexpr long long oLasc = align256(oSegs + 4 * (kMax + kPad));
constexpr long long oRasc = align256(oLasc + 4 * (kMax + kPad));
constexpr long long oNpos = align256(oRasc + 4 * (kMax + kPad));
constexpr long long kBytes = align256(oNpos + 4 * (kMax + kPad));

}  // namespace

static __device__ __align__(256) unsigned char scratch[kBytes];

static std::mutex g_mu;
static unsigned char* g_ptr[64];  // device address of scratch, per device id

static int nfail(int code, const char* msg) { return fsx_fail(code, msg); }

// One landscape record that makes the sort kernels treat the area as a single instance at node
// offset 0 (edge arrays from 0), the key count, and the counters the sort expects at zero; then the
// caller's keys into the working copy (the sort overwrites cw with its range lists).
__global__ void k_noalloc_setup(unsigned char* base, int n, int nn, const double* __restrict__ keys) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t == 0) {
    InstParams* P = reinterpret_cast<InstParams*>(base + oP);
    InstParams q = {};
    q.nx = 3;
    q.ny = nn / 3;
    q.nn = nn;
    q.off = 0;
    q.aux = 0;
    q.active = 1;
    q.run_spl = 1;
    q.sfd = 1;
    *P = q;
  }
  int* iv = reinterpret_cast<int*>(base + oInt);
  if (t < 64) iv[t] = (t == 0) ? n : 0;  // nconn = n; nsegs, err, segctr, nrst = 0
  double* cw = reinterpret_cast<double*>(base + oCw);
  for (int p = t; p < n; p += gridDim.x * blockDim.x) cw[p] = keys[p];
}

__global__ void k_noalloc_perm(const unsigned char* base, int n, int* __restrict__ perm) {
  const int* sid = reinterpret_cast<const int*>(base + oSid);
  for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < n; p += gridDim.x * blockDim.x) perm[p] = sid[p] + 1;
}

extern "C" int fsx_indexx_noalloc(int n, const double* keys, int* perm, void* stream, int sync) {
  if (n < 0 || n > kMax) return nfail(FSX_EINVAL, "fsx_indexx_noalloc: n must be in [0, FSX_INDEXX_NOALLOC_MAX]");
  if (n > 0 && (!keys || !perm)) return nfail(FSX_EINVAL, "fsx_indexx_noalloc: null keys or perm");
  if (n == 0) return 0;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64)
    return nfail(FSX_EHIP, "fsx_indexx_noalloc: no current device");
  unsigned char* base = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_ptr[dev]) {
      void* p = nullptr;
      if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p)
        return nfail(FSX_EHIP, "fsx_indexx_noalloc: hipGetSymbolAddress failed");
      g_ptr[dev] = static_cast<unsigned char*>(p);
    }
    base = g_ptr[dev];
  }
  hipStream_t st = static_cast<hipStream_t>(stream);
  const int nn = (n + FSX_EDGE_CAP - 1) / FSX_EDGE_CAP + 3;  // as fsx_indexx sizes its context
  DevState S;
  S.ninst = 1;
  S.max_nn = nn;
  S.P = reinterpret_cast<InstParams*>(base + oP);
  int* iv = reinterpret_cast<int*>(base + oInt);
  S.nconn = iv + 0;
  S.nsegs = iv + 1;
  S.err = iv + 2;
  S.segctr = iv + 3;
  S.nrst = iv + 32;
  S.cw = reinterpret_cast<double*>(base + oCw);
  S.skey = reinterpret_cast<double*>(base + oSkey);
  S.sid = reinterpret_cast<int*>(base + oSid);
  S.segs = reinterpret_cast<int*>(base + oSegs);
  S.Lasc = reinterpret_cast<int*>(base + oLasc);
  S.Rasc = reinterpret_cast<int*>(base + oRasc);
  S.npos = reinterpret_cast<int*>(base + oNpos);
  const int blocks = (int)std::min<long long>(1024, (n + 255) / 256);
  k_noalloc_setup<<<blocks, 256, 0, st>>>(base, n, nn, keys);
  launch_nr_sort(S, st);
  k_noalloc_perm<<<blocks, 256, 0, st>>>(base, n, perm);
  if (hipGetLastError() != hipSuccess) return nfail(FSX_EHIP, "fsx_indexx_noalloc: launch failed");
  if (sync) {
    int e = 0;
    if (hipMemcpyAsync(&e, S.err, sizeof(int), hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
      return nfail(FSX_EHIP, "fsx_indexx_noalloc: synchronize failed");
    if (e) return nfail(FSX_ESTEP, "fsx_indexx_noalloc: edge-sort segment stack overflow");
  }
  return 0;
}
