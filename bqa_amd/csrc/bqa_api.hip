// libbqa_hip.so: C ABI of the MI355X BP quantum-annealing backend (include/bqa_hip.h).
//
// Host runtime: owns device buffers of one annealing job, sequences the
// phase kernels on the job's stream, and keeps the BP convergence loop on the
// device (a status block read back once per chunk of sweeps instead of once
// per sweep, src/bqa/state.py:113).

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <map>
#include <mutex>
#include <new>
#include <stdexcept>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/bqa_hip.h"
#include "common.hpp"
#include "kernels_canon.hpp"
#include "kernels_canon_h.hpp"
#include "kernels_fast.hpp"
#include "kernels_node.hpp"
#include "kernels_sv.hpp"

using namespace bqa;

namespace {

thread_local std::string g_last_error = "no error";

struct BqaError : std::runtime_error {
  int code;
  BqaError(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define HIPX(expr)                                                                        \
  do {                                                                                    \
    hipError_t e_ = (expr);                                                               \
    if (e_ != hipSuccess)                                                                 \
      throw BqaError(e_ == hipErrorOutOfMemory ? BQA_ERR_MEMORY : BQA_ERR_HIP,            \
                     std::string(#expr) + " failed: " + hipGetErrorString(e_));           \
  } while (0)

#define LAUNCH_CHECK() HIPX(hipGetLastError())

void require(bool ok, const std::string& msg) {
  if (!ok) throw BqaError(BQA_ERR_INVALID, msg);
}

template <typename F>
int guarded(F&& f) {
  try {
    f();
    return BQA_OK;
  } catch (const BqaError& e) {
    g_last_error = e.what();
    return e.code;
  } catch (const std::bad_alloc&) {
    g_last_error = "host allocation failed";
    return BQA_ERR_MEMORY;
  } catch (const std::exception& e) {
    g_last_error = e.what();
    return BQA_ERR_HIP;
  }
}

// Runtime calls that must not overlap another thread's stream capture (BP chunk
// graphs): device and pinned allocation, their release, and stream creation and
// destruction, which synchronise beyond one stream.  Captures hold the same lock.
// No call of this library uses the legacy NULL stream (bqa_energy_noalloc runs on
// the stream its caller passes): while any stream is being captured, HIP fails a
// legacy-stream operation in another thread ("would make the legacy stream depend
// on a capturing blocking stream") and invalidates the capture, so uploads and
// copies go through the job's own stream.
std::mutex g_capture_mu;

// Device buffer that only grows.
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  void ensure(size_t b, size_t& total, hipStream_t s = nullptr) {
    if (b <= bytes) return;
    if (s) HIPX(hipStreamSynchronize(s));
    std::lock_guard<std::mutex> g(g_capture_mu);
    if (p) {
      HIPX(hipFree(p));
      total -= bytes;
      p = nullptr;
      bytes = 0;
    }
    if (b == 0) return;
    HIPX(hipMalloc(&p, b));
    bytes = b;
    total += b;
  }
  void release(size_t& total) {
    if (p) {
      std::lock_guard<std::mutex> g(g_capture_mu);
      (void)hipFree(p);
    }
    total -= bytes;
    p = nullptr;
    bytes = 0;
  }
};

double ipow_d(int b, int e) {
  double r = 1;
  for (int i = 0; i < e; ++i) r *= b;
  return r;
}

long ipow_l(int b, int e) {
  long r = 1;
  for (int i = 0; i < e; ++i) r *= b;
  return r;
}

double bits_to_double(unsigned long long b) {
  double d;
  std::memcpy(&d, &b, sizeof(d));
  return d;
}

}  // namespace

struct ClassDev {
  int d = 0, n = 0;
  DevBuf layout;  // ids | in | out | lp (int32) | h | J (double)
  int32_t *ids = nullptr, *in = nullptr, *out = nullptr, *lp = nullptr;
  double *h = nullptr, *J = nullptr;
  DevBuf T[2];
  int tcur = 0;
  ClassView view() const {
    ClassView v;
    v.n = n;
    v.d = d;
    v.node_ids = ids;
    v.in_pos = in;
    v.out_pos = out;
    v.lmbd_pos = lp;
    v.h = h;
    v.J = J;
    v.zz = nullptr;
    v.sweep = -1;
    return v;
  }
};

enum : unsigned { kGenBP = 1, kGenDensity = 2, kGenExt = 4, kGenCanon = 8, kGenLocal = 16 };

// ZZ-split gates g0, g1 of every edge for one simple update (state.py:127-139,
// backends.py:519-526): zz[2 e], zz[2 e + 1] = zz_split(J_e ztime).
__global__ void k_zz_gates(int E, const double* __restrict__ J, double ztime, cplx<double>* zz) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e < E) {
    cplx<double> g0, g1;
    zz_split(J[e] * ztime, g0, g1);
    zz[2 * (size_t)e] = g0;
    zz[2 * (size_t)e + 1] = g1;
  }
}

struct bqa_job {
  int prec = BQA_PREC_SINGLE;
  size_t csize = 8;  // bytes per complex element
  int N = 0, M = 0, E = 0, chimax = 1, chi = 1;
  int mcur = 0;
  int lstride = 1;
  int pred_sweeps = 8;
  std::vector<ClassDev> cls;
  DevBuf msgs[2], lmbd, ext, canon, lmax, status, rho, bloch, edges, epart, gws, cws, conv, conv_msgs, zz, plist;
  std::vector<int32_t> node_cls, node_pos;  // qubit id -> (degree class, position in the class)
  bool zz_valid = false;  // zz holds the gates of the current simple update
  BpStatus* st_host = nullptr;
  double* lmax_host = nullptr;
  hipStream_t stream = nullptr;
  size_t dev_bytes = 0;
  int device = 0;
  // timing
  bool timing = false;
  std::vector<hipEvent_t> ev_pool;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> ev_pending;
  int64_t t_launches = 0;
  double t_ms = 0.0, t_bytes = 0.0;
  std::vector<double> pending_bytes;
  // BQA_GENERIC_KERNELS: "1" runs the generic kernels for every shape; otherwise
  // any of the letters b (BP sweep), d (density), e (extended messages),
  // c (canonicaliser), l (absorb/gates/gauge) selects them per family.
  unsigned generic_mask = 0;
  bool generic(unsigned fam) const { return (generic_mask & fam) != 0; }
  // BP launch variants: BQA_AUTOTUNE=0 keeps variant 0 everywhere; BQA_BP_VARIANT=k runs
  // variant k wherever the shape has one (both read at job creation).
  bool autotune = true;
  int bp_variant_force = -1;
  // Graph-captured BP chunks (BQA_GRAPHS=0 turns them off): kGraphSweeps sweeps of every class
  // plus the sweep-index advance, captured once per key (bond dimension, buffers, variants)
  // and replayed once the job has run BP kGraphWarmRuns times in a row at its bond dimension.
  struct BpGraph {
    std::vector<intptr_t> key;
    hipGraphExec_t exec = nullptr;
  };
  bool use_graphs = true;
  int graph_chi = 0;   // bond dimension of the last completed BP runs
  int graph_runs = 0;  // consecutive completed BP runs at graph_chi
  std::vector<BpGraph> graphs;
  int64_t graph_captures = 0, graph_launches = 0, graph_failures = 0;

  bool single() const { return prec == BQA_PREC_SINGLE; }
};

namespace {

constexpr int kMaxLds = 64 * 1024;

hipEvent_t take_event(bqa_job* j) {
  if (!j->ev_pool.empty()) {
    hipEvent_t e = j->ev_pool.back();
    j->ev_pool.pop_back();
    return e;
  }
  hipEvent_t e;
  HIPX(hipEventCreate(&e));
  return e;
}

void flush_timing(bqa_job* j) {
  if (j->ev_pending.empty()) return;
  HIPX(hipStreamSynchronize(j->stream));
  for (size_t k = 0; k < j->ev_pending.size(); ++k) {
    float ms = 0.f;
    HIPX(hipEventElapsedTime(&ms, j->ev_pending[k].first, j->ev_pending[k].second));
    j->t_ms += ms;
    j->t_bytes += j->pending_bytes[k];
    j->t_launches += 1;
    j->ev_pool.push_back(j->ev_pending[k].first);
    j->ev_pool.push_back(j->ev_pending[k].second);
  }
  j->ev_pending.clear();
  j->pending_bytes.clear();
}

template <typename R>
cplx<R>* msg_ptr(bqa_job* j, int which) {
  return reinterpret_cast<cplx<R>*>(j->msgs[which].p);
}

// Tensor buffer sizes for bond dimension chi.
size_t tensor_bytes(const bqa_job* j, const ClassDev& c, int chi) {
  return (size_t)c.n * 2 * (size_t)ipow_l(chi, c.d) * j->csize;
}

void ensure_msgs(bqa_job* j, int chi) {
  const size_t b = (size_t)j->M * chi * chi * j->csize;
  // keep the current buffer's contents: grow both only when needed
  if (b > j->msgs[0].bytes || b > j->msgs[1].bytes) {
    HIPX(hipStreamSynchronize(j->stream));
    for (int w = 0; w < 2; ++w) {
      if (b <= j->msgs[w].bytes) continue;
      void* np = nullptr;
      {
        std::lock_guard<std::mutex> g(g_capture_mu);
        HIPX(hipMalloc(&np, b));
      }
      if (j->msgs[w].p) {
        HIPX(hipMemcpyAsync(np, j->msgs[w].p, j->msgs[w].bytes, hipMemcpyDeviceToDevice, j->stream));
        HIPX(hipStreamSynchronize(j->stream));
        std::lock_guard<std::mutex> g(g_capture_mu);
        HIPX(hipFree(j->msgs[w].p));
        j->dev_bytes -= j->msgs[w].bytes;
      }
      j->msgs[w].p = np;
      j->msgs[w].bytes = b;
      j->dev_bytes += b;
    }
  }
}

void ensure_tensor(bqa_job* j, ClassDev& c, int which, int chi) {
  const size_t b = tensor_bytes(j, c, chi);
  if (b > c.T[which].bytes) {
    HIPX(hipStreamSynchronize(j->stream));
    c.T[which].ensure(b, j->dev_bytes);
  }
}

// Grid and workspace of the generic node kernels: LDS when the three node
// buffers fit, else a global slice per workgroup.
struct NodeLaunch {
  int grid;
  size_t lds;
  void* ws;
  long ws_elems;
};

NodeLaunch plan_node(bqa_job* j, int n, size_t buf_bytes_total, size_t fixed_lds, size_t elems_per_block,
                     size_t csize) {
  NodeLaunch L;
  if (fixed_lds + buf_bytes_total <= (size_t)kMaxLds) {
    L.grid = std::max(1, std::min(n, 4096));
    L.lds = fixed_lds + buf_bytes_total;
    L.ws = nullptr;
    L.ws_elems = 0;
  } else {
    L.grid = std::max(1, std::min(n, 1024));
    const size_t need = (size_t)L.grid * elems_per_block * csize;
    if (need > j->gws.bytes) {
      j->gws.ensure(need, j->dev_bytes, j->stream);
    }
    L.lds = fixed_lds;
    L.ws = j->gws.p;
    L.ws_elems = (long)elems_per_block;
  }
  return L;
}

// Algorithmic bytes of one BP sweep over one class (the roofline numerator):
// tensors read once, incoming and old messages read, new messages written,
// two int32 position arrays.
double sweep_bytes(const bqa_job* j, const ClassDev& c, int chi) {
  const double cs = (double)j->csize;
  return (double)c.n * 2.0 * ipow_d(chi, c.d) * cs + 3.0 * c.n * c.d * (double)chi * chi * cs +
         (double)c.n * c.d * 8.0;
}

__global__ void k_bp_reset(BpStatus* st, int cur, int max_iters, double bp_eps) {
  st->num_bits = 0ULL;
  st->den_bits = 0ULL;
  st->cur = cur;
  st->done = 0;
  st->converged = 0;
  st->iters = 0;
  st->have_best = 0;
  st->damp_now = 0;
  st->max_iters = max_iters;
  st->marker = 0;
  st->base = 0;
  for (int k = 0; k < kSweepSlots && k < max_iters; ++k) {
    st->slot[k][0] = 0ULL;
    st->slot[k][1] = 0ULL;
  }
  st->bp_eps = bp_eps;
  st->last_dist = 0.0;
  st->best_dist = INFINITY;
}

// End of a graph-replayed chunk of sweeps: the next replay's launches take the
// following sweep indices.
__global__ void k_bp_advance(BpStatus* st, int by) { st->base += by; }

// After every sweep: distance, convergence and buffer bookkeeping
// (src/bqa/state.py:113-121).
__global__ void k_bp_finish(BpStatus* st, int damping_on) {
  if (st->done) return;
  const double num = __longlong_as_double((long long)st->num_bits);
  const double den = __longlong_as_double((long long)st->den_bits);
  st->num_bits = 0ULL;
  st->den_bits = 0ULL;
  const double dist = num / den;
  st->iters += 1;
  st->last_dist = dist;
  if (dist < st->best_dist) {
    st->best_dist = dist;
    st->have_best = 1;
  }
  st->damp_now = 0;
  if (dist < st->bp_eps) {
    st->converged = 1;
    st->done = 1;  // keep the previous iterate
  } else if (st->iters >= st->max_iters) {
    st->done = 1;
    st->cur = 1 - st->cur;  // the last undamped iterate
  } else if (damping_on) {
    st->damp_now = 1;
  } else {
    st->cur = 1 - st->cur;
  }
}

// msgs[cur] <- a msgs[cur] + (1 - a) msgs[1 - cur] (backends.py:761-764).
template <typename R>
__global__ void k_bp_damp(BpStatus* st, cplx<R>* m0, cplx<R>* m1, long count, double alpha) {
  if (!st->damp_now) return;
  cplx<R>* dst = st->cur ? m1 : m0;
  const cplx<R>* src = st->cur ? m0 : m1;
  const R a = (R)alpha, b = (R)(1.0 - alpha);
  for (long x = blockIdx.x * (long)blockDim.x + threadIdx.x; x < count; x += (long)gridDim.x * blockDim.x) {
    const cplx<R> v = dst[x], w = src[x];
    dst[x] = mk<R>(v.x * a + b * w.x, v.y * a + b * w.y);
  }
}

template <typename R>
__global__ void k_init_tensors(cplx<R>* T, long n) {
  const R h = (R)0.70710678118654752440;
  for (long x = blockIdx.x * (long)blockDim.x + threadIdx.x; x < n; x += (long)gridDim.x * blockDim.x) {
    T[2 * x] = mk<R>(h, 0);
    T[2 * x + 1] = mk<R>(-h, 0);
  }
}

template <typename R>
__global__ void k_fill_one(cplx<R>* m, long n) {
  for (long x = blockIdx.x * (long)blockDim.x + threadIdx.x; x < n; x += (long)gridDim.x * blockDim.x)
    m[x] = mk<R>(1, 0);
}

__global__ void k_fill_d(double* p, long n, double v) {
  for (long x = blockIdx.x * (long)blockDim.x + threadIdx.x; x < n; x += (long)gridDim.x * blockDim.x) p[x] = v;
}

__global__ void k_zero_u64(unsigned long long* p, int n) {
  for (int x = threadIdx.x; x < n; x += blockDim.x) p[x] = 0ULL;
}

// Energy of the rounded spins: partial sums per block, deterministic order.
__global__ void k_energy_partial(const double* bloch, int N, const double* h_by_id, const int32_t* lhs,
                                 const int32_t* rhs, const double* jv, int E, double* part) {
  __shared__ double sh[16];
  double acc = 0.0;
  for (long x = blockIdx.x * (long)blockDim.x + threadIdx.x; x < N; x += (long)gridDim.x * blockDim.x) {
    const double s = bloch[x * 3 + 2] > 0.0 ? 1.0 : -1.0;
    acc += h_by_id[x] * s;
  }
  for (long x = blockIdx.x * (long)blockDim.x + threadIdx.x; x < E; x += (long)gridDim.x * blockDim.x) {
    const double sa = bloch[(long)lhs[x] * 3 + 2] > 0.0 ? 1.0 : -1.0;
    const double sb = bloch[(long)rhs[x] * 3 + 2] > 0.0 ? 1.0 : -1.0;
    acc += jv[x] * sa * sb;
  }
  acc = block_sum(acc, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = acc;
}

__global__ void k_energy_final(const double* part, int nb, double* out) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    double s = 0.0;
    for (int b = 0; b < nb; ++b) s += part[b];
    *out = s;
  }
}

// Projective measurement of listed qubits of one class (backends.py:729-734):
// entry x of `list` is (position in the class, outcome); the physical component
// 1 - outcome of that node's tensor (half = chi^d elements) is set to zero.
template <typename R>
__global__ void k_project(cplx<R>* T, const int2* __restrict__ list, int count, long half) {
  for (int x = blockIdx.y; x < count; x += gridDim.y) {
    const int2 e = list[x];
    cplx<R>* dst = T + (2 * (long)e.x + (1 - e.y)) * half;
    for (long k = blockIdx.x * (long)blockDim.x + threadIdx.x; k < half; k += (long)gridDim.x * blockDim.x)
      dst[k] = mk<R>(0, 0);
  }
}

// Scratch of the allocation-free energy entry point (bqa_energy_noalloc): the
// block partial sums of k_energy_partial.  One copy per GPU, held until the
// process ends; host code reaches it only through the address that
// hipGetSymbolAddress returns on each device (scratch_on_device).
constexpr int kEnergyBlocks = 512;
constexpr int32_t kNoAllocMaxItems = 1 << 30;
__device__ __align__(256) double scratch[kEnergyBlocks];

__global__ void k_copy_bytes(const uint4* __restrict__ src, uint4* __restrict__ dst, long n) {
  for (long x = blockIdx.x * (long)blockDim.x + threadIdx.x; x < n; x += (long)gridDim.x * blockDim.x)
    dst[x] = src[x];
}

int grid_for(long n, int block = 256, int cap = 8192) {
  long g = (n + block - 1) / block;
  return (int)std::max<long>(1, std::min<long>(g, cap));
}

// ---------------------------------------------------------------- phases --

// ---------------------------------------------------------- autotuner --
//
// Process-wide choice of BP launch variant (kernels_fast.hpp) per device,
// precision, degree, chi and ceil(log2 n).  A shape is timed once, on the first
// sweep of a BP run that meets it: every variant runs that sweep twice.  Each
// launch reads the sweep's input buffer and writes the same outputs and the same
// distance maxima, so the repeats leave the run exactly as one launch would;
// the variants differ only in time (grid policies: bit for bit; wave splits:
// to rounding).  While one job times a shape, others use variant 0.
struct TuneKey {
  int dev, prec, d, chi, nlog;
  bool operator<(const TuneKey& o) const {
    return std::tie(dev, prec, d, chi, nlog) < std::tie(o.dev, o.prec, o.d, o.chi, o.nlog);
  }
};
struct TuneEntry {
  int chosen = -1;  // -1 while being timed
  std::vector<double> us;
};
std::mutex g_tune_mu;
std::map<TuneKey, TuneEntry> g_tune;
constexpr int kTuneReps = 2;

int ceil_log2(long n) {
  int k = 0;
  while ((1L << k) < n) ++k;
  return k;
}

template <typename R>
int bp_variant(bqa_job* j, ClassDev& c, const ClassView& view, bool may_tune) {
  const int chi = j->chi;
  const int nv = bp_variant_count<R>(c.d, chi);
  if (nv <= 1) return 0;
  if (j->bp_variant_force >= 0) return j->bp_variant_force < nv ? j->bp_variant_force : 0;
  if (!j->autotune) return 0;
  const TuneKey key{j->device, j->prec, c.d, chi, ceil_log2(c.n)};
  {
    std::lock_guard<std::mutex> g(g_tune_mu);
    auto it = g_tune.find(key);
    if (it != g_tune.end()) return it->second.chosen >= 0 ? it->second.chosen : 0;
    if (!may_tune) return 0;
    g_tune[key];
  }
  const cplx<R>* T = reinterpret_cast<const cplx<R>*>(c.T[c.tcur].p);
  BpStatus* st = reinterpret_cast<BpStatus*>(j->status.p);
  std::vector<hipEvent_t> ev(nv + 1, nullptr);
  TuneEntry res;
  try {
    for (auto& e : ev) HIPX(hipEventCreate(&e));
    fast_bp_launch<R>(view, chi, T, msg_ptr<R>(j, 0), msg_ptr<R>(j, 1), st, j->stream, 0);  // warm-up
    LAUNCH_CHECK();
    for (int v = 0; v < nv; ++v) {
      HIPX(hipEventRecord(ev[v], j->stream));
      for (int r = 0; r < kTuneReps; ++r) {
        fast_bp_launch<R>(view, chi, T, msg_ptr<R>(j, 0), msg_ptr<R>(j, 1), st, j->stream, v);
        LAUNCH_CHECK();
      }
    }
    HIPX(hipEventRecord(ev[nv], j->stream));
    HIPX(hipEventSynchronize(ev[nv]));
    res.us.resize(nv);
    res.chosen = 0;
    for (int v = 0; v < nv; ++v) {
      float ms = 0.f;
      HIPX(hipEventElapsedTime(&ms, ev[v], ev[v + 1]));
      res.us[v] = 1e3 * ms / kTuneReps;
      if (res.us[v] < res.us[res.chosen]) res.chosen = v;
    }
  } catch (...) {
    for (auto e : ev)
      if (e) (void)hipEventDestroy(e);
    throw;
  }
  for (auto e : ev) (void)hipEventDestroy(e);
  std::lock_guard<std::mutex> g(g_tune_mu);
  g_tune[key] = res;
  return res.chosen;
}

// One class's launch of a BP sweep; sweep >= 0 selects the sweep-slot protocol.
// `first` marks the run's first sweep, where the autotuner may time the shape.
template <typename R>
void launch_bp_class(bqa_job* j, ClassDev& c, int sweep, bool first) {
  const int chi = j->chi;
  if (c.d == 0 || c.n == 0) return;
  ClassView view = c.view();
  view.sweep = sweep;
  const int variant = j->generic(kGenBP) ? 0 : bp_variant<R>(j, c, view, first);
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (j->timing) {
    e0 = take_event(j);
    e1 = take_event(j);
    HIPX(hipEventRecord(e0, j->stream));
  }
  if (j->generic(kGenBP) || !fast_bp_launch<R>(view, chi, reinterpret_cast<const cplx<R>*>(c.T[c.tcur].p), msg_ptr<R>(j, 0),
                         msg_ptr<R>(j, 1), reinterpret_cast<BpStatus*>(j->status.p), j->stream, variant)) {
    const long S = 2 * ipow_l(chi, c.d);
    const size_t fixed = sizeof(cplx<R>) * (size_t)(kNT + 4 * chi * chi);
    NodeLaunch L = plan_node(j, c.n, sizeof(cplx<R>) * 3 * S, fixed, 3 * S, sizeof(cplx<R>));
    hipLaunchKernelGGL((k_node_contract<R, kModeBP>), dim3(L.grid), dim3(kNT), L.lds, j->stream, view, chi,
                       reinterpret_cast<const cplx<R>*>(c.T[c.tcur].p), msg_ptr<R>(j, 0), msg_ptr<R>(j, 1),
                       reinterpret_cast<BpStatus*>(j->status.p), (cplx<R>*)nullptr, 0.0, (double*)nullptr,
                       (double*)nullptr, reinterpret_cast<cplx<R>*>(L.ws), L.ws_elems);
  }
  LAUNCH_CHECK();
  if (j->timing) {
    HIPX(hipEventRecord(e1, j->stream));
    j->ev_pending.emplace_back(e0, e1);
    j->pending_bytes.push_back(sweep_bytes(j, c, chi));
  }
}

// ------------------------------------------------------- graph chunks --

constexpr int kGraphSweeps = 4;
constexpr size_t kMaxGraphs = 8;
// Eager BP runs at one bond dimension before its chunks are captured: a capture and
// instantiation cost about as much as the launches of tens of sweeps, so graphs
// serve the steady state, not the steps of the bond-dimension ramp.
constexpr int kGraphWarmRuns = 8;

// The chunk graph for the current buffers and launch variants, captured on first use.  Capture
// uses the thread-local mode: other threads (other jobs) keep calling the runtime meanwhile,
// which the global mode would refuse, failing their calls.
template <typename R>
hipGraphExec_t bp_graph(bqa_job* j, const std::vector<int>& variants) {
  const int chi = j->chi;
  std::vector<intptr_t> key = {chi, (intptr_t)j->msgs[0].p, (intptr_t)j->msgs[1].p, (intptr_t)j->status.p};
  for (size_t ci = 0; ci < j->cls.size(); ++ci) {
    const ClassDev& c = j->cls[ci];
    key.push_back((intptr_t)c.T[c.tcur].p);
    key.push_back(c.n);
    key.push_back(variants[ci]);
  }
  for (const auto& g : j->graphs)
    if (g.key == key) return g.exec;
  BpStatus* st = reinterpret_cast<BpStatus*>(j->status.p);
  hipGraph_t graph = nullptr;
  std::unique_lock<std::mutex> lock(g_capture_mu);
  HIPX(hipStreamBeginCapture(j->stream, hipStreamCaptureModeThreadLocal));
  try {
    for (int s = 0; s < kGraphSweeps; ++s)
      for (size_t ci = 0; ci < j->cls.size(); ++ci) {
        ClassDev& c = j->cls[ci];
        if (c.d == 0 || c.n == 0) continue;
        ClassView view = c.view();
        view.sweep = s;
        if (!fast_bp_launch<R>(view, chi, reinterpret_cast<const cplx<R>*>(c.T[c.tcur].p), msg_ptr<R>(j, 0),
                               msg_ptr<R>(j, 1), st, j->stream, variants[ci]))
          throw BqaError(BQA_ERR_HIP, "no specialised BP kernel for a captured class");
        LAUNCH_CHECK();
      }
    hipLaunchKernelGGL(k_bp_advance, dim3(1), dim3(1), 0, j->stream, st, kGraphSweeps);
    LAUNCH_CHECK();
  } catch (...) {
    hipGraph_t dead = nullptr;
    (void)hipStreamEndCapture(j->stream, &dead);
    if (dead) (void)hipGraphDestroy(dead);
    throw;
  }
  const hipError_t ec = hipStreamEndCapture(j->stream, &graph);
  lock.unlock();
  if (ec != hipSuccess || graph == nullptr) {
    // a runtime call from outside this library invalidated the capture: the job
    // goes on with eager launches
    if (graph) (void)hipGraphDestroy(graph);
    (void)hipGetLastError();
    j->use_graphs = false;
    ++j->graph_failures;
    return nullptr;
  }
  hipGraphExec_t exec = nullptr;
  const hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  HIPX(e);
  if (j->graphs.size() >= kMaxGraphs) {
    (void)hipGraphExecDestroy(j->graphs.front().exec);
    j->graphs.erase(j->graphs.begin());
  }
  j->graphs.push_back({std::move(key), exec});
  ++j->graph_captures;
  return exec;
}

// The rules of k_bp_finish (no damping, src/bqa/state.py:113-121) applied on the host to the
// sweep slots [ruled, launched) of a run, stopping at the first that ends it.
void apply_slot_rules(const BpStatus* sh, BpStatus& h, int& ruled, int launched, int max_iters, double bp_eps) {
  for (; ruled < launched && !h.done; ++ruled) {
    const double num = bits_to_double(sh->slot[ruled][0]);
    const double den = bits_to_double(sh->slot[ruled][1]);
    const double dist = num / den;
    h.iters += 1;
    h.last_dist = dist;
    if (dist < h.best_dist) {
      h.best_dist = dist;
      h.have_best = 1;
    }
    if (dist < bp_eps) {
      h.converged = 1;
      h.done = 1;  // keep the previous iterate
    } else if (h.iters >= max_iters) {
      h.done = 1;
      h.cur = 1 - h.cur;  // the last undamped iterate
    } else {
      h.cur = 1 - h.cur;
    }
  }
}

template <typename R>
void run_bp_impl(bqa_job* j, int max_iters, double bp_eps, double damping, int* sweeps, int* converged,
                 double* last_dist) {
  if (max_iters <= 0)  // the reference asserts here (src/bqa/state.py:122)
    throw BqaError(BQA_ERR_NUMERIC, "max_bp_iter_number must be positive: BP ran no sweep");
  BpStatus* st = reinterpret_cast<BpStatus*>(j->status.p);
  hipLaunchKernelGGL(k_bp_reset, dim3(1), dim3(1), 0, j->stream, st, j->mcur, max_iters, bp_eps);
  LAUNCH_CHECK();
  const long count = (long)j->M * j->chi * j->chi;
  const int damping_on = damping > 0.0 ? 1 : 0;
  const size_t ev_start = j->ev_pending.size();
  int active_classes = 0;
  for (auto& cl : j->cls) active_classes += (cl.d > 0 && cl.n > 0) ? 1 : 0;
  // Without damping, sweeps run on the sweep-slot protocol: no kernel between
  // sweeps, the rules applied here to each chunk's slots.
  const bool slots = !damping_on && active_classes > 0 && max_iters <= kSweepSlots;
  BpStatus h{};
  h.cur = j->mcur;
  h.best_dist = INFINITY;
  int launched = 0, ruled = 0;
  int chunk = std::max(1, std::min(j->pred_sweeps, max_iters));
  // Graph replays once this bond dimension has had an eager run (its variants chosen and the
  // kernels' launch shapes computed), with kernel timing off and a specialised kernel per class.
  bool graph = slots && j->use_graphs && !j->timing && !j->generic(kGenBP) && j->graph_chi == j->chi &&
               j->graph_runs >= kGraphWarmRuns;
  for (auto& cl : j->cls)
    if (cl.d > 0 && cl.n > 0 && !bp_fast_shape<R>(cl.d, j->chi)) graph = false;
  std::vector<int> variants;
  if (graph) {
    for (auto& cl : j->cls) {
      ClassView view = cl.view();
      variants.push_back(cl.d > 0 && cl.n > 0 ? bp_variant<R>(j, cl, view, false) : 0);
    }
  }
  while (launched < max_iters) {
    int c = std::min(chunk, max_iters - launched);
    const hipGraphExec_t exec = graph ? bp_graph<R>(j, variants) : nullptr;
    if (graph && !exec) graph = false;
    if (graph) {
      // whole replays: their sweeps beyond max_iters, or after convergence, exit at once
      const int reps = (c + kGraphSweeps - 1) / kGraphSweeps;
      for (int r = 0; r < reps; ++r) {
        HIPX(hipGraphLaunch(exec, j->stream));
        ++j->graph_launches;
      }
      c = reps * kGraphSweeps;
    } else {
      for (int s = 0; s < c; ++s) {
        for (auto& cl : j->cls) launch_bp_class<R>(j, cl, slots ? launched + s : -1, launched + s == 0);
        if (slots) continue;
        hipLaunchKernelGGL(k_bp_finish, dim3(1), dim3(1), 0, j->stream, st, damping_on);
        LAUNCH_CHECK();
        if (damping_on) {
          hipLaunchKernelGGL((k_bp_damp<R>), dim3(grid_for(count)), dim3(256), 0, j->stream, st, msg_ptr<R>(j, 0),
                             msg_ptr<R>(j, 1), count, damping);
          LAUNCH_CHECK();
        }
      }
    }
    launched += c;
    HIPX(hipMemcpyAsync(j->st_host, st, sizeof(BpStatus), hipMemcpyDeviceToHost, j->stream));
    HIPX(hipStreamSynchronize(j->stream));
    if (slots) {
      apply_slot_rules(j->st_host, h, ruled, launched, max_iters, bp_eps);
    } else {
      h = *j->st_host;
    }
    if (h.done) break;
    chunk = 4;
  }
  if (j->timing) {
    // launches after convergence exit at once: time only the sweeps that ran
    const size_t keep = ev_start + (size_t)h.iters * active_classes;
    while (j->ev_pending.size() > keep) {
      j->ev_pool.push_back(j->ev_pending.back().first);
      j->ev_pool.push_back(j->ev_pending.back().second);
      j->ev_pending.pop_back();
      j->pending_bytes.pop_back();
    }
    flush_timing(j);
  }
  j->mcur = h.cur;
  if (slots) {
    j->graph_runs = j->graph_chi == j->chi ? j->graph_runs + 1 : 1;
    j->graph_chi = j->chi;
  }
  j->pred_sweeps = std::max(1, h.iters);
  *sweeps = h.iters;
  *converged = h.converged;
  *last_dist = h.last_dist;
  if (!h.have_best) throw BqaError(BQA_ERR_NUMERIC, "BP produced no finite message distance");
}

template <int N>
void launch_canon_h(bqa_job* j, int ccols, double pinv_eps, double deps) {
  using S = HCanon<N>;
  j->cws.ensure((size_t)j->E * S::WS * sizeof(double), j->dev_bytes, j->stream);
  const int grid = (j->E + S::EPW - 1) / S::EPW;
  hipLaunchKernelGGL((k_canon_h<N>), dim3(grid), dim3(64), 0, j->stream,
                     reinterpret_cast<const cplx<double>*>(j->ext.p), j->E, ccols, pinv_eps, deps,
                     reinterpret_cast<zd*>(j->canon.p), reinterpret_cast<double*>(j->lmbd.p),
                     reinterpret_cast<double*>(j->cws.p));
  LAUNCH_CHECK();
}

// Register-resident canonicalisers for chi <= 16 (split-column Cholesky + Jacobi); false
// for the other bond dimensions.
bool canon_reg_dispatch(bqa_job* j, int chi, int ccols, double pinv_eps, double deps) {
  switch (chi) {
    case 1: launch_canon_h<2>(j, ccols, pinv_eps, deps); return true;
    case 2: launch_canon_h<4>(j, ccols, pinv_eps, deps); return true;
    case 3: launch_canon_h<6>(j, ccols, pinv_eps, deps); return true;
    case 4: launch_canon_h<8>(j, ccols, pinv_eps, deps); return true;
    case 5: launch_canon_h<10>(j, ccols, pinv_eps, deps); return true;
    case 6: launch_canon_h<12>(j, ccols, pinv_eps, deps); return true;
    case 7: launch_canon_h<14>(j, ccols, pinv_eps, deps); return true;
    case 8: launch_canon_h<16>(j, ccols, pinv_eps, deps); return true;
    case 9: launch_canon_h<18>(j, ccols, pinv_eps, deps); return true;
    case 10: launch_canon_h<20>(j, ccols, pinv_eps, deps); return true;
    case 11: launch_canon_h<22>(j, ccols, pinv_eps, deps); return true;
    case 12: launch_canon_h<24>(j, ccols, pinv_eps, deps); return true;
    case 13: launch_canon_h<26>(j, ccols, pinv_eps, deps); return true;
    case 14: launch_canon_h<28>(j, ccols, pinv_eps, deps); return true;
    case 15: launch_canon_h<30>(j, ccols, pinv_eps, deps); return true;
    case 16: launch_canon_h<32>(j, ccols, pinv_eps, deps); return true;
    default: return false;
  }
}

__global__ void k_widen(const cplx<float>* __restrict__ src, cplx<double>* dst, long n) {
  for (long x = blockIdx.x * (long)blockDim.x + threadIdx.x; x < n; x += (long)gridDim.x * blockDim.x)
    dst[x] = cast<double>(src[x]);
}

// Extended messages of one class in FP64 arithmetic: MFMA for chi = 8 at degrees 2-3 and for
// FP32 jobs at chi = 16, degrees 2-3 (FP32 mode products, FP64 Grams), the degree-4 chi = 8
// kernel, the compile-time-shaped kernels below chi = 8, otherwise the generic kernel (FP32
// inputs are widened into a scratch buffer first).
template <typename R>
void ext_class(bqa_job* j, ClassDev& c, double ztime, const cplx<double>* msgs64) {
  const int chi = j->chi;
  const cplx<R>* T = reinterpret_cast<const cplx<R>*>(c.T[c.tcur].p);
  cplx<double>* ext = reinterpret_cast<cplx<double>*>(j->ext.p);
  ClassView v = c.view();
  v.zz = reinterpret_cast<const cplx<double>*>(j->zz.p);
  if (!j->generic(kGenExt) && fast_ext_launch<R>(v, chi, T, msg_ptr<R>(j, j->mcur), ext, j->stream)) {
    LAUNCH_CHECK();
    return;
  }
  const long S = 2 * ipow_l(chi, c.d);
  const cplx<double>* T64 = reinterpret_cast<const cplx<double>*>(T);
  if (sizeof(R) == sizeof(float)) {
    const size_t tb = (size_t)c.n * S * sizeof(cplx<double>);
    j->conv.ensure(tb, j->dev_bytes, j->stream);
    hipLaunchKernelGGL(k_widen, dim3(grid_for((long)c.n * S)), dim3(256), 0, j->stream,
                       reinterpret_cast<const cplx<float>*>(T), reinterpret_cast<cplx<double>*>(j->conv.p),
                       (long)c.n * S);
    LAUNCH_CHECK();
    T64 = reinterpret_cast<const cplx<double>*>(j->conv.p);
  }
  const size_t fixed = sizeof(cplx<double>) * (size_t)(kNT + 4 * chi * chi);
  NodeLaunch L = plan_node(j, c.n, sizeof(cplx<double>) * 3 * S, fixed, 3 * S, sizeof(cplx<double>));
  hipLaunchKernelGGL((k_node_contract<double, kModeExt>), dim3(L.grid), dim3(kNT), L.lds, j->stream, c.view(), chi,
                     T64, const_cast<cplx<double>*>(msgs64), (cplx<double>*)nullptr, (BpStatus*)nullptr, ext, ztime,
                     (double*)nullptr, (double*)nullptr, reinterpret_cast<cplx<double>*>(L.ws), L.ws_elems);
  LAUNCH_CHECK();
}

void canonicalise(bqa_job* j, double pinv_eps, bool sync = true);
void canon_outputs(bqa_job* j);

template <typename R>
void ext_and_canon(bqa_job* j, double ztime, double pinv_eps, bool sync = true, bool canon = true) {
  const int chi = j->chi, n2 = 2 * chi;
  j->ext.ensure((size_t)j->M * n2 * n2 * sizeof(cplx<double>), j->dev_bytes, j->stream);
  // the step's ZZ-split gates per edge, read by the extended-message and absorb kernels
  j->zz.ensure(sizeof(cplx<double>) * 2 * (size_t)std::max(j->E, 1), j->dev_bytes, j->stream);
  if (j->E > 0) {
    const double* jv = reinterpret_cast<const double*>(reinterpret_cast<const char*>(j->edges.p) +
                                                       ((sizeof(int32_t) * 2 * j->E + 15) / 16) * 16);
    hipLaunchKernelGGL(k_zz_gates, dim3((j->E + 255) / 256), dim3(256), 0, j->stream, j->E, jv, ztime,
                       reinterpret_cast<cplx<double>*>(j->zz.p));
    LAUNCH_CHECK();
  }
  j->zz_valid = true;
  // FP64 view of the current messages for the generic extended-message kernel
  const cplx<double>* msgs64 = reinterpret_cast<const cplx<double>*>(msg_ptr<R>(j, j->mcur));
  bool need_generic = false;
  for (auto& c : j->cls)
    need_generic |= c.d != 0 && c.n > 0 &&
                    (j->generic(kGenExt) || !((chi == 8 && c.d >= 2 && c.d <= 4) || small_shape(c.d, chi) ||
                                              (chi == 16 && (c.d == 2 || c.d == 3)) ||
                                              (sizeof(R) == sizeof(float) && c.d == 5 && small_bp_shape<R>(5, chi))));
  if (sizeof(R) == sizeof(float) && need_generic) {
    const long cnt = (long)j->M * chi * chi;
    j->conv_msgs.ensure((size_t)cnt * sizeof(cplx<double>), j->dev_bytes, j->stream);
    hipLaunchKernelGGL(k_widen, dim3(grid_for(cnt)), dim3(256), 0, j->stream,
                       reinterpret_cast<const cplx<float>*>(msg_ptr<R>(j, j->mcur)),
                       reinterpret_cast<cplx<double>*>(j->conv_msgs.p), cnt);
    LAUNCH_CHECK();
    msgs64 = reinterpret_cast<const cplx<double>*>(j->conv_msgs.p);
  }
  for (auto& c : j->cls) {
    if (c.d == 0 || c.n == 0) continue;
    ext_class<R>(j, c, ztime, msgs64);
  }
  if (canon) canonicalise(j, pinv_eps, sync);
}

// FP64 canonicalisers and lambdas of the 2chi x 2chi messages in j->ext (state.py:171-200), then
// the per-column lambda maxima on the host (the global truncation rule reads them); without
// `sync` the caller synchronises the stream before reading them.
// The canonicaliser columns kept (the new bond dimension's bound) and the pseudo-inverse
// threshold of the job's storage precision (backends.py:726).
int canon_cols(const bqa_job* j) { return std::min(2 * j->chi, j->chimax); }
double canon_deps(const bqa_job* j) {
  return j->single() ? (double)std::numeric_limits<float>::epsilon() : std::numeric_limits<double>::epsilon();
}

// n of the split-column canonicaliser serving bond dimension chi, 0 if none does.
int canon_h_n(int chi) { return chi >= 1 && chi <= 16 ? 2 * chi : 0; }

void canon_outputs(bqa_job* j) {
  const int n2 = 2 * j->chi;
  j->canon.ensure((size_t)j->M * n2 * canon_cols(j) * sizeof(zd), j->dev_bytes, j->stream);
  j->lmbd.ensure((size_t)j->E * n2 * sizeof(double), j->dev_bytes, j->stream);
}

// Per-column lambda maxima of the job's canonicalisers, copied to the host (read by the
// truncation rule once the stream is synchronised).
void lambda_readback(bqa_job* j) {
  const int n2 = 2 * j->chi;
  unsigned long long* lm = reinterpret_cast<unsigned long long*>(j->lmax.p);
  hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(128), 0, j->stream, lm, n2);
  LAUNCH_CHECK();
  hipLaunchKernelGGL(k_lambda_max, dim3(grid_for((long)j->E * n2, 256, 1024)), dim3(256), 0, j->stream,
                     reinterpret_cast<const double*>(j->lmbd.p), j->E, n2, lm);
  LAUNCH_CHECK();
  HIPX(hipMemcpyAsync(j->lmax_host, lm, sizeof(double) * n2, hipMemcpyDeviceToHost, j->stream));
}

void canonicalise(bqa_job* j, double pinv_eps, bool sync) {
  const int chi = j->chi, n2 = 2 * chi;
  const int ccols = canon_cols(j);
  canon_outputs(j);
  const size_t wsb = canon_ws_bytes(n2);
  const double deps = canon_deps(j);
  if (j->generic(kGenCanon) || !canon_reg_dispatch(j, chi, ccols, pinv_eps, deps)) {
    int grid;
    size_t lds;
    unsigned char* ws = nullptr;
    if (wsb <= (size_t)kMaxLds) {
      grid = std::max(1, std::min(j->E, 16384));
      lds = wsb;
    } else {
      grid = std::max(1, std::min(j->E, 2048));
      j->cws.ensure((size_t)grid * wsb, j->dev_bytes, j->stream);
      ws = reinterpret_cast<unsigned char*>(j->cws.p);
      lds = 0;
    }
    hipLaunchKernelGGL((k_canonicalize<double>), dim3(grid), dim3(kCanonNT), lds, j->stream,
                       reinterpret_cast<const cplx<double>*>(j->ext.p), j->E, chi, ccols, pinv_eps, deps,
                       reinterpret_cast<zd*>(j->canon.p), reinterpret_cast<double*>(j->lmbd.p), ws, (long)wsb);
    LAUNCH_CHECK();
  }
  lambda_readback(j);
  if (sync) HIPX(hipStreamSynchronize(j->stream));
}

// Global truncation rule (backends.py:297-303).
int truncation_dim(const bqa_job* j, int n2, double pinv_eps, double* err) {
  int below = 0;
  for (int k = 0; k < n2; ++k) below += j->lmax_host[k] < pinv_eps ? 1 : 0;
  const int rank = n2 - below;
  const int dim = std::min(rank, j->chimax);
  double e = 0.0;
  for (int k = dim; k < n2; ++k) e += j->lmax_host[k] * j->lmax_host[k];
  *err = std::sqrt(e);
  return dim;
}

template <typename R>
void local_update(bqa_job* j, int flags, int chi_out, double zz_time, double z_time, double x_time) {
  const int chi_in = j->chi;
  const int n2 = 2 * chi_in;
  const int ccols = std::min(n2, j->chimax);
  if (flags & kGauge) ensure_msgs(j, chi_out);
  for (auto& c : j->cls) {
    if (c.n == 0) continue;
    const int dst = 1 - c.tcur;
    ensure_tensor(j, c, dst, chi_out);
    const cplx<R>* Tin = reinterpret_cast<const cplx<R>*>(c.T[c.tcur].p);
    cplx<R>* Tout = reinterpret_cast<cplx<R>*>(c.T[dst].p);
    ClassView v = c.view();
    v.zz = (flags & kAbsorb) && j->zz_valid ? reinterpret_cast<const cplx<double>*>(j->zz.p) : nullptr;
    if (j->generic(kGenLocal) || !fast_local_launch<R>(v, chi_in, chi_out, Tin, Tout, reinterpret_cast<const zd*>(j->canon.p),
                              n2, ccols, zz_time, flags, z_time, x_time,
                              reinterpret_cast<const double*>(j->lmbd.p), j->lstride, msg_ptr<R>(j, j->mcur),
                              j->stream)) {
      const int cm = std::max(chi_in, chi_out);
      const long Smax = 2 * ipow_l(cm, c.d);
      NodeLaunch L = plan_node(j, c.n, sizeof(cplx<double>) * 2 * Smax, 0, 2 * Smax, sizeof(cplx<double>));
      hipLaunchKernelGGL((k_node_local<R, double>), dim3(L.grid), dim3(kNT), L.lds, j->stream, c.view(), chi_in,
                         chi_out, Tin, Tout, reinterpret_cast<const zd*>(j->canon.p), n2, ccols, zz_time, flags,
                         z_time, x_time, reinterpret_cast<const double*>(j->lmbd.p), j->lstride,
                         msg_ptr<R>(j, j->mcur), reinterpret_cast<cplx<double>*>(L.ws), L.ws_elems);
    }
    LAUNCH_CHECK();
    c.tcur = dst;
  }
  j->chi = chi_out;
}

template <typename R>
void simple_update_impl(bqa_job* j, double ztime, double pinv_eps, int extra_flags, double z_time,
                        double x_time, int* new_dim, double* err) {
  ext_and_canon<R>(j, ztime, pinv_eps);
  const int n2 = 2 * j->chi;
  const int dim = truncation_dim(j, n2, pinv_eps, err);
  j->lstride = n2;
  local_update<R>(j, kAbsorb | extra_flags, dim, ztime, z_time, x_time);
  *new_dim = dim;
}

// Messages as the top-left chi x chi block of otherwise zero 2chi x 2chi matrices: the s = 0
// block of the extended-message layout, so the canonicalisers of these matrices are the
// canonicalisers of the messages in their first chi rows and zero below.
template <typename R>
__global__ void k_pad_msgs(const cplx<R>* __restrict__ m, cplx<double>* __restrict__ ext, long M, int chi) {
  const int n2 = 2 * chi;
  const long per = (long)n2 * n2;
  for (long x = blockIdx.x * (long)blockDim.x + threadIdx.x; x < M * per; x += (long)gridDim.x * blockDim.x) {
    const long pos = x / per;
    const int o = (int)(x - pos * per), r = o / n2, c = o % n2;
    ext[x] = (r < chi && c < chi) ? cast<double>(m[pos * chi * chi + r * chi + c]) : mk(0.0, 0.0);
  }
}

// Tensors (n, 2, chi_in^d) -> (n, 2, chi_out^d) keeping indices below chi_out on every bond.
template <typename R>
__global__ void k_slice_tensors(const cplx<R>* __restrict__ Tin, cplx<R>* __restrict__ Tout, long n, int d,
                                int chi_in, int chi_out) {
  long a_in = 1, a_out = 1;
  for (int k = 0; k < d; ++k) {
    a_in *= chi_in;
    a_out *= chi_out;
  }
  for (long x = blockIdx.x * (long)blockDim.x + threadIdx.x; x < n * 2 * a_out; x += (long)gridDim.x * blockDim.x) {
    const long np = x / a_out;           // node * 2 + p
    long a = x - np * a_out, src = 0, mul = 1;
    for (int k = 0; k < d; ++k) {        // last bond fastest
      const long digit = a % chi_out;
      a /= chi_out;
      src += digit * mul;
      mul *= chi_in;
    }
    Tout[x] = Tin[np * a_in + src];
  }
}

// Messages (M, chi_in, chi_in) -> their leading (chi_out, chi_out) blocks.
template <typename R>
__global__ void k_slice_msgs(const cplx<R>* __restrict__ min, cplx<R>* __restrict__ mout, long M, int chi_in,
                             int chi_out) {
  const long per = (long)chi_out * chi_out;
  for (long x = blockIdx.x * (long)blockDim.x + threadIdx.x; x < M * per; x += (long)gridDim.x * blockDim.x) {
    const long pos = x / per;
    const int o = (int)(x - pos * per), r = o / chi_out, c = o % chi_out;
    mout[x] = min[pos * chi_in * chi_in + r * chi_in + c];
  }
}

// Vidal gauge (state.py:204-216): canonicalisers of the current messages, absorbed into every
// tensor bond without a ZZ extension (zz_time 0 makes the extension the identity on the s = 0
// block, the only non-zero rows); the bond dimension and the messages stay as they are.
template <typename R>
void vidal_impl(bqa_job* j, double pinv_eps) {
  const int chi = j->chi, n2 = 2 * chi;
  j->ext.ensure((size_t)j->M * n2 * n2 * sizeof(cplx<double>), j->dev_bytes, j->stream);
  const long cnt = (long)j->M * n2 * n2;
  hipLaunchKernelGGL((k_pad_msgs<R>), dim3(grid_for(cnt)), dim3(256), 0, j->stream, msg_ptr<R>(j, j->mcur),
                     reinterpret_cast<cplx<double>*>(j->ext.p), (long)j->M, chi);
  LAUNCH_CHECK();
  canonicalise(j, pinv_eps);
  j->lstride = n2;
  j->zz_valid = false;
  local_update<R>(j, kAbsorb, chi, 0.0, 0.0, 0.0);
}

// state.py:160-168: the global truncation rule on the lambdas (backends.py:297-303) and every
// tensor bond cut to the new bond dimension.  The reference leaves its messages at the old size,
// which its own BP cannot then read; here they keep their leading block, so the job stays whole.
template <typename R>
void truncate_vidal_impl(bqa_job* j, double pinv_eps, int* dim, double* err) {
  const int chi = j->chi, ls = j->lstride;
  unsigned long long* lm = reinterpret_cast<unsigned long long*>(j->lmax.p);
  hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(128), 0, j->stream, lm, ls);
  LAUNCH_CHECK();
  hipLaunchKernelGGL(k_lambda_max, dim3(grid_for((long)j->E * ls, 256, 1024)), dim3(256), 0, j->stream,
                     reinterpret_cast<const double*>(j->lmbd.p), j->E, ls, lm);
  LAUNCH_CHECK();
  HIPX(hipMemcpyAsync(j->lmax_host, lm, sizeof(double) * ls, hipMemcpyDeviceToHost, j->stream));
  HIPX(hipStreamSynchronize(j->stream));
  const int d_new = std::min(truncation_dim(j, ls, pinv_eps, err), chi);
  if (d_new < chi) {
    for (auto& c : j->cls) {
      if (c.n == 0) continue;
      const int dst = 1 - c.tcur;
      ensure_tensor(j, c, dst, d_new);
      const long total = (long)c.n * 2 * ipow_l(d_new, c.d);
      hipLaunchKernelGGL((k_slice_tensors<R>), dim3(grid_for(total)), dim3(256), 0, j->stream,
                         reinterpret_cast<const cplx<R>*>(c.T[c.tcur].p), reinterpret_cast<cplx<R>*>(c.T[dst].p),
                         (long)c.n, c.d, chi, d_new);
      LAUNCH_CHECK();
      c.tcur = dst;
    }
    const long mc = (long)j->M * d_new * d_new;
    hipLaunchKernelGGL((k_slice_msgs<R>), dim3(grid_for(mc)), dim3(256), 0, j->stream, msg_ptr<R>(j, j->mcur),
                       msg_ptr<R>(j, 1 - j->mcur), (long)j->M, chi, d_new);
    LAUNCH_CHECK();
    j->mcur = 1 - j->mcur;
    j->chi = d_new;
  }
  HIPX(hipStreamSynchronize(j->stream));
  *dim = d_new;
}

template <typename R>
void density_impl(bqa_job* j, double* rho_out, double* bloch_out) {
  j->rho.ensure((size_t)j->N * 8 * sizeof(double), j->dev_bytes, j->stream);
  j->bloch.ensure((size_t)j->N * 3 * sizeof(double), j->dev_bytes, j->stream);
  const int chi = j->chi;
  for (auto& c : j->cls) {
    if (c.n == 0) continue;
    const cplx<R>* T = reinterpret_cast<const cplx<R>*>(c.T[c.tcur].p);
    if (!j->generic(kGenDensity) && fast_density_launch<R>(c.view(), chi, T, msg_ptr<R>(j, j->mcur), reinterpret_cast<double*>(j->rho.p),
                               reinterpret_cast<double*>(j->bloch.p), j->stream)) {
      LAUNCH_CHECK();
      continue;
    }
    const long S = 2 * ipow_l(chi, c.d);
    const size_t fixed = sizeof(cplx<R>) * (size_t)(kNT + 4 * chi * chi);
    NodeLaunch L = plan_node(j, c.n, sizeof(cplx<R>) * 3 * S, fixed, 3 * S, sizeof(cplx<R>));
    hipLaunchKernelGGL((k_node_contract<R, kModeDensity>), dim3(L.grid), dim3(kNT), L.lds, j->stream, c.view(), chi,
                       T, msg_ptr<R>(j, j->mcur), (cplx<R>*)nullptr, (BpStatus*)nullptr, (cplx<R>*)nullptr, 0.0,
                       reinterpret_cast<double*>(j->rho.p), reinterpret_cast<double*>(j->bloch.p),
                       reinterpret_cast<cplx<R>*>(L.ws), L.ws_elems);
    LAUNCH_CHECK();
  }
  if (rho_out)
    HIPX(hipMemcpyAsync(rho_out, j->rho.p, (size_t)j->N * 8 * sizeof(double), hipMemcpyDeviceToHost, j->stream));
  if (bloch_out)
    HIPX(hipMemcpyAsync(bloch_out, j->bloch.p, (size_t)j->N * 3 * sizeof(double), hipMemcpyDeviceToHost,
                        j->stream));
  HIPX(hipStreamSynchronize(j->stream));
}

// Project the listed qubits (ids[k] on outcomes[k]); grouped by class, one launch per class.
template <typename R>
void project_impl(bqa_job* j, int count, const int32_t* ids, const int32_t* outcomes) {
  std::vector<std::vector<int2>> per(j->cls.size());
  for (int k = 0; k < count; ++k) {
    require(ids[k] >= 0 && ids[k] < j->N, "qubit id out of range");
    require(outcomes[k] == 0 || outcomes[k] == 1, "outcome must be 0 or 1");
    per[j->node_cls[ids[k]]].push_back(make_int2(j->node_pos[ids[k]], outcomes[k]));
  }
  size_t off = 0;
  for (auto& v : per) off += v.size();
  if (off == 0) return;
  j->plist.ensure(off * sizeof(int2), j->dev_bytes, j->stream);
  std::vector<int2> flat;
  flat.reserve(off);
  for (auto& v : per) flat.insert(flat.end(), v.begin(), v.end());
  HIPX(hipMemcpyAsync(j->plist.p, flat.data(), off * sizeof(int2), hipMemcpyHostToDevice, j->stream));
  off = 0;
  for (size_t c = 0; c < per.size(); ++c) {
    const int cnt = (int)per[c].size();
    if (cnt == 0) continue;
    ClassDev& cl = j->cls[c];
    const long half = ipow_l(j->chi, cl.d);
    const dim3 grid((unsigned)std::max<long>(1, std::min<long>((half + 255) / 256, 64)),
                    (unsigned)std::min(cnt, 4096));
    hipLaunchKernelGGL((k_project<R>), grid, dim3(256), 0, j->stream, reinterpret_cast<cplx<R>*>(cl.T[cl.tcur].p),
                       reinterpret_cast<const int2*>(j->plist.p) + off, cnt, half);
    LAUNCH_CHECK();
    off += cnt;
  }
  HIPX(hipStreamSynchronize(j->stream));  // the host list buffer goes out of scope
}

template <typename R>
void init_state_impl(bqa_job* j) {
  j->chi = 1;
  j->mcur = 0;
  j->lstride = 1;
  ensure_msgs(j, 1);
  for (auto& c : j->cls) {
    c.tcur = 0;
    if (c.n == 0) continue;
    ensure_tensor(j, c, 0, 1);
    hipLaunchKernelGGL((k_init_tensors<R>), dim3(grid_for(c.n)), dim3(256), 0, j->stream,
                       reinterpret_cast<cplx<R>*>(c.T[0].p), (long)c.n);
    LAUNCH_CHECK();
  }
  j->lmbd.ensure((size_t)j->E * 2 * sizeof(double), j->dev_bytes, j->stream);
  hipLaunchKernelGGL(k_fill_d, dim3(grid_for(j->E)), dim3(256), 0, j->stream, reinterpret_cast<double*>(j->lmbd.p),
                     (long)j->E, 1.0);
  LAUNCH_CHECK();
  hipLaunchKernelGGL((k_fill_one<R>), dim3(grid_for(j->M)), dim3(256), 0, j->stream, msg_ptr<R>(j, 0), (long)j->M);
  LAUNCH_CHECK();
  HIPX(hipStreamSynchronize(j->stream));
  j->pred_sweeps = 8;
}

// Device address of `scratch` on the current device, looked up once per device.
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

double* scratch_on_device() {
  int dev = 0;
  HIPX(hipGetDevice(&dev));
  require(dev >= 0 && dev < kMaxDevices, "device index out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    HIPX(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    require(p != nullptr, "no device address for the energy scratch");
    g_scratch_ptr[dev] = p;
  }
  return static_cast<double*>(g_scratch_ptr[dev]);
}

// ------------------------------------------------------- state transfer --

template <typename R>
void download(bqa_job* j, const void* dev, size_t count, double* out) {
  std::vector<cplx<R>> h(count);
  HIPX(hipMemcpyAsync(h.data(), dev, count * sizeof(cplx<R>), hipMemcpyDeviceToHost, j->stream));
  HIPX(hipStreamSynchronize(j->stream));
  for (size_t x = 0; x < count; ++x) {
    out[2 * x] = (double)h[x].x;
    out[2 * x + 1] = (double)h[x].y;
  }
}

template <typename R>
void upload(bqa_job* j, void* dev, size_t count, const double* in) {
  std::vector<cplx<R>> h(count);
  for (size_t x = 0; x < count; ++x) h[x] = mk<R>((R)in[2 * x], (R)in[2 * x + 1]);
  HIPX(hipMemcpyAsync(dev, h.data(), count * sizeof(cplx<R>), hipMemcpyHostToDevice, j->stream));
  HIPX(hipStreamSynchronize(j->stream));
}

template <typename F>
void dispatch(bqa_job* j, F&& f) {
  HIPX(hipSetDevice(j->device));
  if (j->single())
    f(float());
  else
    f(double());
}

// Device bytes a job of this shape can hold at its peak: every buffer the
// phases grow, at the largest bond dimension they reach.  Buffers only grow, so
// the peak is the sum of each buffer's largest size over chi = 1..chimax.
size_t estimate(int N, int M, const std::vector<int>& deg, const std::vector<int>& size, int chimax, int prec) {
  const double cs = prec == BQA_PREC_SINGLE ? 8.0 : 16.0;
  const double E = M / 2.0;
  const int n2 = 2 * chimax, ccols = std::min(n2, chimax);
  double b = 0.0;
  int dmax = 0;
  long nmax = 0;
  for (size_t c = 0; c < deg.size(); ++c) {
    b += 2.0 * size[c] * 2.0 * ipow_d(chimax, deg[c]) * cs;                    // two tensor buffers
    b += 4.0 * size[c] * (1 + 3.0 * deg[c]) + 8.0 * size[c] * (1 + deg[c]) + 64;  // layout arrays
    dmax = std::max(dmax, deg[c]);
    nmax = std::max<long>(nmax, size[c]);
  }
  b += 2.0 * M * (double)chimax * chimax * cs;                     // two message buffers
  b += (double)M * n2 * n2 * 16.0;                                 // extended messages (FP64)
  b += (double)M * n2 * ccols * sizeof(zd);                        // canonicalisers
  b += E * n2 * sizeof(double);                                    // lambdas
  b += E * 2.0 * 16.0;                                             // ZZ-split gates
  // canonicaliser workspace: per edge for the split-column kernels (chi <= 16; largest at chi =
  // min(chimax, 16)), per workgroup for the generic kernel once its workspace leaves LDS
  const int nr = 2 * std::min(chimax, 16);
  double cws = E * (10.0 * nr * nr + 5.0 * nr) * sizeof(double);
  if (chimax > 8) {
    const size_t wsb = canon_ws_bytes(n2);
    if (wsb > (size_t)kMaxLds) cws = std::max(cws, std::min(E, 2048.0) * (double)wsb);
  }
  b += cws;
  // generic node kernels' global workspace (3 node buffers per workgroup, at most 1024 workgroups) once
  // their buffers leave LDS; FP64 elements for the extended-message and local kernels
  const double smax = 2.0 * ipow_d(chimax, dmax);
  if (16.0 * 3.0 * smax > (double)kMaxLds) b += 1024.0 * 3.0 * smax * 16.0;
  // FP32 jobs whose extended messages run the generic kernel widen tensors and messages to FP64: at
  // bond dimension 1 (the first step) always, at chimax when no specialised kernel serves the shape
  if (prec == BQA_PREC_SINGLE) {
    const bool generic_at_max = chimax > 8 || dmax > 3;
    b += (double)nmax * (generic_at_max ? smax : 2.0) * 16.0;
    b += (double)M * (generic_at_max ? (double)chimax * chimax : 1.0) * 16.0;
  }
  b += (double)N * 11.0 * sizeof(double);                          // densities and Bloch vectors
  b += E * (2.0 * 4.0 + 8.0) + (double)N * 8.0 + 64;               // forward edges and fields
  b += sizeof(BpStatus) + 128 * sizeof(double) + 1024 * sizeof(double) + 64;
  return (size_t)b;
}

}  // namespace

// ------------------------------------------------------ exact simulator --

struct bqa_sv {
  int n = 0, E = 0, device = 0;
  DevBuf psi, ez, graph, part;
  hipStream_t stream = nullptr;
  size_t dev_bytes = 0;
};

namespace {

constexpr int kSvMaxQubits = 32;
constexpr int kSvMaxGrid = 2048;

int sv_grid(unsigned long long tiles) { return (int)std::min<unsigned long long>(tiles, kSvMaxGrid); }

// The X rotations (MODE 0, with the diagonal phase when ez is given) or the readout sums
// (MODE 1) of every qubit, pass by pass; `each` gets the block count, first qubit and qubit
// count of a readout pass.
template <int MODE, typename F>
void sv_passes(bqa_sv* s, const double* ez, double ztime, double xtime, F&& each) {
  const int n = s->n;
  const double c = std::cos(xtime), sn = std::sin(xtime);
  auto* psi = reinterpret_cast<cplx<double>*>(s->psi.p);
  auto* part = reinterpret_cast<double*>(s->part.p);
  const int tb = std::min(n, kSvTileBits);
  const int g0 = sv_grid(1ULL << (n - tb));
  hipLaunchKernelGGL((k_sv_low<MODE>), dim3(g0), dim3(kSvThreads), 0, s->stream, psi, n, ez, ztime, c, sn, part);
  LAUNCH_CHECK();
  each(g0, 0, tb);
  for (int b0 = tb; b0 < n; b0 += kSvTileBits - kSvRowBits) {
    const int t = std::min(kSvTileBits - kSvRowBits, n - b0);
    const int g = sv_grid(1ULL << (n - kSvRowBits - t));
    hipLaunchKernelGGL((k_sv_high<MODE>), dim3(g), dim3(kSvThreads), 0, s->stream, psi, n, b0, t, c, sn, part);
    LAUNCH_CHECK();
    each(g, b0, t);
  }
}

}  // namespace

struct bqa_batch {
  int device = 0;
  int max_jobs = 0;
  hipStream_t stream = nullptr;
  std::vector<hipEvent_t> ready;     // per job slot: its simple update is done
  std::vector<hipEvent_t> ext_done;  // per job slot: its extended messages are done
  hipEvent_t canon_done = nullptr;   // the shared canonicaliser launches and readbacks are done
  DevBuf ctab;                       // the shared canonicaliser launches' entry tables
  CanonEntry* ctab_host = nullptr;
  size_t ctab_cap = 0;
  DevBuf tab;                 // the BP launches' entry tables
  BpEntry* tab_host = nullptr;  // pinned staging of the tables
  size_t tab_cap = 0;         // entries both hold
  size_t dev_bytes = 0;
};

namespace {

// Split-column canonicaliser geometry of n: workspace doubles per edge, edges per 64-lane block.
size_t canon_h_ws(int n) { return 10 * (size_t)n * n + 5 * (size_t)n; }
int canon_h_epw(int n) {
  const int p = n / 2, p2 = p <= 1 ? 1 : p <= 2 ? 2 : p <= 4 ? 4 : p <= 8 ? 8 : 16;
  return 64 / (2 * p2);
}

template <int N>
void launch_canon_multi(const CanonEntry* tab, int entries, int blocks, hipStream_t s) {
  hipLaunchKernelGGL((k_canon_h_multi<N>), dim3(blocks), dim3(64), 0, s, tab, entries);
}

bool canon_multi_dispatch(int n, const CanonEntry* tab, int entries, int blocks, hipStream_t s) {
  switch (n) {
    case 2: launch_canon_multi<2>(tab, entries, blocks, s); return true;
    case 4: launch_canon_multi<4>(tab, entries, blocks, s); return true;
    case 6: launch_canon_multi<6>(tab, entries, blocks, s); return true;
    case 8: launch_canon_multi<8>(tab, entries, blocks, s); return true;
    case 10: launch_canon_multi<10>(tab, entries, blocks, s); return true;
    case 12: launch_canon_multi<12>(tab, entries, blocks, s); return true;
    case 14: launch_canon_multi<14>(tab, entries, blocks, s); return true;
    case 16: launch_canon_multi<16>(tab, entries, blocks, s); return true;
    case 18: launch_canon_multi<18>(tab, entries, blocks, s); return true;
    case 20: launch_canon_multi<20>(tab, entries, blocks, s); return true;
    case 22: launch_canon_multi<22>(tab, entries, blocks, s); return true;
    case 24: launch_canon_multi<24>(tab, entries, blocks, s); return true;
    case 26: launch_canon_multi<26>(tab, entries, blocks, s); return true;
    case 28: launch_canon_multi<28>(tab, entries, blocks, s); return true;
    case 30: launch_canon_multi<30>(tab, entries, blocks, s); return true;
    case 32: launch_canon_multi<32>(tab, entries, blocks, s); return true;
    default: return false;
  }
}

// The canonicalisers of the batch's jobs whose bond dimension has a split-column kernel (state.py:171-200
// per job): once each job's extended messages are done (its own stream), one launch per n over all of
// them on the batch stream, then each job's lambda maxima readback there; each job's stream then waits
// for that.  Jobs marked false in `shared` ran their own canonicaliser with their extended messages.
void batch_canon(bqa_batch* b, int count, bqa_job* const* jobs, const double* pinv_eps,
                 const std::vector<char>& shared) {
  std::vector<int> order;
  for (int k = 0; k < count; ++k)
    if (shared[k]) order.push_back(k);
  if (order.empty()) return;
  std::stable_sort(order.begin(), order.end(),
                   [&](int x, int y) { return canon_h_n(jobs[x]->chi) < canon_h_n(jobs[y]->chi); });
  for (int k : order) {
    bqa_job* j = jobs[k];
    canon_outputs(j);
    j->cws.ensure((size_t)j->E * canon_h_ws(canon_h_n(j->chi)) * sizeof(double), j->dev_bytes, j->stream);
    HIPX(hipEventRecord(b->ext_done[k], j->stream));
    HIPX(hipStreamWaitEvent(b->stream, b->ext_done[k], 0));
  }
  const size_t entries = order.size();
  if (entries > b->ctab_cap) {
    HIPX(hipStreamSynchronize(b->stream));
    std::lock_guard<std::mutex> lk(g_capture_mu);
    if (b->ctab_host) (void)hipHostFree(b->ctab_host);
    b->ctab_host = nullptr;
    b->ctab_cap = 0;
    HIPX(hipHostMalloc(reinterpret_cast<void**>(&b->ctab_host), entries * sizeof(CanonEntry), hipHostMallocDefault));
    b->ctab_cap = entries;
  }
  b->ctab.ensure(entries * sizeof(CanonEntry), b->dev_bytes, b->stream);
  struct Launch {
    int n;
    size_t first, count;
    int blocks;
  };
  std::vector<Launch> launches;
  for (size_t x = 0; x < entries; ++x) {
    bqa_job* j = jobs[order[x]];
    const int n = canon_h_n(j->chi);
    if (launches.empty() || launches.back().n != n) launches.push_back({n, x, 0, 0});
    Launch& L = launches.back();
    CanonEntry& e = b->ctab_host[x];
    e.ext = reinterpret_cast<const cplx<double>*>(j->ext.p);
    e.canon = reinterpret_cast<zd*>(j->canon.p);
    e.lfull = reinterpret_cast<double*>(j->lmbd.p);
    e.ws = reinterpret_cast<double*>(j->cws.p);
    e.E = j->E;
    e.ccols = canon_cols(j);
    e.pinv_eps = pinv_eps[order[x]];
    e.deps = canon_deps(j);
    e.block0 = L.blocks;
    const long blocks = (long)L.blocks + (j->E + canon_h_epw(n) - 1) / canon_h_epw(n);
    require(blocks <= std::numeric_limits<int>::max(), "too many edges in one shared canonicaliser launch");
    L.blocks = (int)blocks;
    ++L.count;
  }
  HIPX(hipMemcpyAsync(b->ctab.p, b->ctab_host, entries * sizeof(CanonEntry), hipMemcpyHostToDevice, b->stream));
  const CanonEntry* tab = reinterpret_cast<const CanonEntry*>(b->ctab.p);
  for (const auto& L : launches) {
    if (!canon_multi_dispatch(L.n, tab + L.first, (int)L.count, L.blocks, b->stream))
      throw BqaError(BQA_ERR_HIP, "no shared canonicaliser for n = " + std::to_string(L.n));
    LAUNCH_CHECK();
  }
  for (int k : order) {
    bqa_job* j = jobs[k];
    const hipStream_t own = j->stream;
    j->stream = b->stream;
    try {
      lambda_readback(j);
    } catch (...) {
      j->stream = own;
      throw;
    }
    j->stream = own;
  }
  HIPX(hipEventRecord(b->canon_done, b->stream));
  for (int k : order) HIPX(hipStreamWaitEvent(jobs[k]->stream, b->canon_done, 0));
}

// BP of the batch's jobs (src/bqa/state.py:97-124 per job): every job's run starts together on the
// batch stream; each sweep is one shared launch per (precision, kernel family, degree, chi) over the
// classes of all jobs that have it, plus the jobs' own launches for the other classes; after each
// chunk of sweeps one copy of every running job's status and one synchronisation, then each job's
// rules.  Jobs with damping, or more iterations than sweep slots, run their own BP.
void batch_bp(bqa_batch* b, int count, bqa_job* const* jobs, const int32_t* max_iters, const double* bp_eps,
              const double* damping, int32_t* sweeps, int32_t* converged, double* last_dist, int32_t* status) {
  struct Run {
    bqa_job* j;
    int k;
    BpStatus h;
    int ruled;
  };
  std::vector<Run> runs;
  for (int k = 0; k < count; ++k) {
    bqa_job* j = jobs[k];
    require(max_iters[k] > 0, "max_bp_iter_number must be positive: BP ran no sweep");
    int active = 0;
    for (auto& cl : j->cls) active += (cl.d > 0 && cl.n > 0) ? 1 : 0;
    if (damping[k] > 0.0 || max_iters[k] > kSweepSlots || active == 0) {
      dispatch(j, [&](auto r) {
        int s = 0, c = 0;
        double d = 0.0;
        try {
          run_bp_impl<decltype(r)>(j, max_iters[k], bp_eps[k], damping[k], &s, &c, &d);
        } catch (const BqaError& e) {
          if (e.code != BQA_ERR_NUMERIC) throw;
          status[k] = BQA_ERR_NUMERIC;
        }
        sweeps[k] = s;
        converged[k] = c;
        last_dist[k] = d;
      });
      continue;
    }
    Run r{j, k, BpStatus{}, 0};
    r.h.cur = j->mcur;
    r.h.best_dist = INFINITY;
    hipLaunchKernelGGL(k_bp_reset, dim3(1), dim3(1), 0, b->stream, reinterpret_cast<BpStatus*>(j->status.p),
                       j->mcur, max_iters[k], bp_eps[k]);
    LAUNCH_CHECK();
    runs.push_back(r);
  }
  if (runs.empty()) return;
  struct Group {
    int prec, fam, d, chi;
    std::vector<std::pair<int, int>> members;  // (run, class)
    long total = 0;
    size_t off = 0;
  };
  std::vector<Group> groups;
  std::vector<std::pair<int, int>> solo;
  for (int ri = 0; ri < (int)runs.size(); ++ri) {
    bqa_job* j = runs[ri].j;
    for (int ci = 0; ci < (int)j->cls.size(); ++ci) {
      const ClassDev& cl = j->cls[ci];
      if (cl.d == 0 || cl.n == 0) continue;
      int fam = j->generic(kGenBP) ? 0 : j->single() ? multi_family<float>(cl.d, j->chi)
                                                    : multi_family<double>(cl.d, j->chi);
      // autotuned jobs at the k_small_npw shapes share the several-nodes-per-wave kernel (what the
      // tuner picks for them alone at all but the smallest sizes)
      if (fam == 1 && npw_shape(cl.d, j->chi) && j->autotune && j->bp_variant_force < 0) fam = kFamNpw;
      if (fam == 0) {
        solo.emplace_back(ri, ci);
        continue;
      }
      Group* g = nullptr;
      for (auto& x : groups)
        if (x.prec == j->prec && x.fam == fam && x.d == cl.d && x.chi == j->chi) g = &x;
      if (!g) {
        groups.push_back(Group{j->prec, fam, cl.d, j->chi, {}, 0, 0});
        g = &groups.back();
      }
      g->members.emplace_back(ri, ci);
    }
  }
  size_t entries = 0;
  for (auto& g : groups) {
    g.off = entries;
    entries += g.members.size();
  }
  if (entries > 0) {
    if (entries > b->tab_cap) {
      HIPX(hipStreamSynchronize(b->stream));
      std::lock_guard<std::mutex> lk(g_capture_mu);
      if (b->tab_host) (void)hipHostFree(b->tab_host);
      b->tab_host = nullptr;
      b->tab_cap = 0;
      HIPX(hipHostMalloc(reinterpret_cast<void**>(&b->tab_host), entries * sizeof(BpEntry), hipHostMallocDefault));
      b->tab_cap = entries;
    }
    b->tab.ensure(entries * sizeof(BpEntry), b->dev_bytes, b->stream);
    for (auto& g : groups) {
      long prefix = 0;
      for (size_t m = 0; m < g.members.size(); ++m) {
        bqa_job* j = runs[g.members[m].first].j;
        const ClassDev& cl = j->cls[g.members[m].second];
        BpEntry& e = b->tab_host[g.off + m];
        e.cv = cl.view();
        e.T = cl.T[cl.tcur].p;
        e.mA = j->msgs[0].p;
        e.mB = j->msgs[1].p;
        e.st = reinterpret_cast<BpStatus*>(j->status.p);
        e.prefix = prefix;
        prefix += g.fam == kFamNpw ? (cl.n + kNpwShared - 1) / kNpwShared : cl.n;
      }
      g.total = prefix;
      require(g.total <= std::numeric_limits<int>::max(), "too many nodes in one shared launch");
    }
    HIPX(hipMemcpyAsync(b->tab.p, b->tab_host, entries * sizeof(BpEntry), hipMemcpyHostToDevice, b->stream));
  }
  const BpEntry* tab = reinterpret_cast<const BpEntry*>(b->tab.p);
  int maxit = 1, chunk = 1;
  for (auto& r : runs) {
    maxit = std::max(maxit, (int)max_iters[r.k]);
    chunk = std::max(chunk, r.j->pred_sweeps);
  }
  chunk = std::min(chunk, maxit);
  int launched = 0;
  while (launched < maxit) {
    const int c = std::min(chunk, maxit - launched);
    for (int s = 0; s < c; ++s) {
      const int sweep = launched + s;
      for (auto& g : groups) {
        const bool ok = g.prec == BQA_PREC_SINGLE
                            ? launch_bp_multi<float>(g.fam, g.d, g.chi, tab + g.off, (int)g.members.size(),
                                                     (int)g.total, sweep, b->stream)
                            : launch_bp_multi<double>(g.fam, g.d, g.chi, tab + g.off, (int)g.members.size(),
                                                      (int)g.total, sweep, b->stream);
        if (!ok) throw BqaError(BQA_ERR_HIP, "no shared BP kernel for a batched class");
        LAUNCH_CHECK();
      }
      for (auto& sc : solo) {
        Run& r = runs[sc.first];
        if (r.h.done || sweep >= max_iters[r.k]) continue;
        dispatch(r.j, [&](auto rr) { launch_bp_class<decltype(rr)>(r.j, r.j->cls[sc.second], sweep, sweep == 0); });
      }
    }
    launched += c;
    bool all = true;
    for (auto& r : runs) {
      if (r.h.done) continue;
      HIPX(hipMemcpyAsync(r.j->st_host, r.j->status.p, sizeof(BpStatus), hipMemcpyDeviceToHost, b->stream));
    }
    HIPX(hipStreamSynchronize(b->stream));
    for (auto& r : runs) {
      if (r.h.done) continue;
      apply_slot_rules(r.j->st_host, r.h, r.ruled, launched, max_iters[r.k], bp_eps[r.k]);
      all = all && r.h.done;
    }
    if (all) break;
    chunk = 4;
  }
  for (auto& r : runs) {
    r.j->mcur = r.h.cur;
    r.j->pred_sweeps = std::max(1, r.h.iters);
    sweeps[r.k] = r.h.iters;
    converged[r.k] = r.h.converged;
    last_dist[r.k] = r.h.last_dist;
    if (!r.h.have_best) status[r.k] = BQA_ERR_NUMERIC;  // bqa_run_bp's error for that job
  }
}

}  // namespace

// =================================================================== C ABI

extern "C" {

const char* bqa_last_error(void) { return g_last_error.c_str(); }

int bqa_device_info(int* device_count, int* arch_ok, char* name_out, int name_len) {
  return guarded([&] {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
    if (device_count) *device_count = n;
    if (arch_ok) *arch_ok = 0;
    if (n > 0) {
      hipDeviceProp_t p;
      HIPX(hipGetDeviceProperties(&p, 0));
      if (arch_ok) *arch_ok = std::strncmp(p.gcnArchName, "gfx950", 6) == 0 ? 1 : 0;
      if (name_out && name_len > 0) {
        std::snprintf(name_out, name_len, "%s", p.gcnArchName);
      }
    }
  });
}

int bqa_job_estimate_bytes(int32_t edges_number, int32_t classes, const int32_t* class_degree,
                           const int32_t* class_size, int32_t max_bond_dim, int32_t precision, uint64_t* bytes) {
  return guarded([&] {
    require(bytes && class_degree && class_size && classes >= 0, "null argument");
    std::vector<int> deg(class_degree, class_degree + classes), sz(class_size, class_size + classes);
    require(edges_number >= 0 && max_bond_dim >= 1 && max_bond_dim <= 32, "bad shape");
    long n = 0;
    for (int c = 0; c < classes; ++c) n += sz[c];
    *bytes = estimate((int)n, edges_number, deg, sz, max_bond_dim, precision);
  });
}

int bqa_job_create(bqa_job** out, int32_t nodes_number, int32_t edges_number, int32_t classes,
                   const int32_t* class_degree, const int32_t* class_size, const int32_t* node_ids,
                   const int32_t* input_pos, const int32_t* output_pos, const int32_t* lmbd_pos,
                   const double* node_ampls, const double* edge_ampls, const int32_t* fwd_lhs,
                   const int32_t* fwd_rhs, const double* fwd_j, int32_t max_bond_dim, int32_t precision) {
  bqa_job* j = nullptr;
  int rc = guarded([&] {
    require(out != nullptr, "null job pointer");
    require(nodes_number > 0 && edges_number > 0 && edges_number % 2 == 0, "need n > 0 and an even number of messages");
    require(classes > 0 && class_degree && class_size, "no degree classes");
    require(max_bond_dim >= 1 && max_bond_dim <= 32, "max_bond_dim must be in [1, 32]");
    require(precision == BQA_PREC_SINGLE || precision == BQA_PREC_DOUBLE, "unknown precision");
    int dev_count = 0;
    HIPX(hipGetDeviceCount(&dev_count));
    require(dev_count > 0, "no HIP device");
    j = new bqa_job();
    j->prec = precision;
    j->csize = precision == BQA_PREC_SINGLE ? 8 : 16;
    j->N = nodes_number;
    j->M = edges_number;
    j->E = edges_number / 2;
    j->chimax = max_bond_dim;
    HIPX(hipGetDevice(&j->device));
    {
      std::lock_guard<std::mutex> g(g_capture_mu);
      HIPX(hipStreamCreateWithFlags(&j->stream, hipStreamNonBlocking));
    }
    {
      const char* g = std::getenv("BQA_GENERIC_KERNELS");
      if (g && g[0] == '1') j->generic_mask = ~0u;
      for (const char* q = g; q && *q; ++q)
        j->generic_mask |= *q == 'b' ? kGenBP : *q == 'd' ? kGenDensity : *q == 'e' ? kGenExt
                         : *q == 'c' ? kGenCanon : *q == 'l' ? kGenLocal : 0u;
      const char* a = std::getenv("BQA_AUTOTUNE");
      j->autotune = !(a && a[0] == '0');
      const char* gr = std::getenv("BQA_GRAPHS");
      j->use_graphs = !(gr && gr[0] == '0');
      const char* v = std::getenv("BQA_BP_VARIANT");
      if (v && *v) {
        char* end = nullptr;
        const long k = std::strtol(v, &end, 10);
        require(end && *end == '\0' && k >= 0 && k < 16, "BQA_BP_VARIANT must be an integer in [0, 16)");
        j->bp_variant_force = (int)k;
      }
    }
    long total_nodes = 0, total_bonds = 0;
    for (int c = 0; c < classes; ++c) {
      require(class_degree[c] >= 0 && class_size[c] >= 0, "negative class degree or size");
      total_nodes += class_size[c];
      total_bonds += (long)class_degree[c] * class_size[c];
    }
    require(total_nodes == nodes_number, "class sizes do not add up to nodes_number");
    require(total_bonds == edges_number, "class degrees do not add up to edges_number");
    long off_n = 0, off_b = 0;
    std::vector<double> h_by_id(nodes_number, 0.0);
    j->cls.resize(classes);
    for (int c = 0; c < classes; ++c) {
      ClassDev& cl = j->cls[c];
      cl.d = class_degree[c];
      cl.n = class_size[c];
      require(ipow_d(max_bond_dim, cl.d) * 2.0 < 2147483647.0, "degree too large for max_bond_dim");
      const size_t nb = (size_t)cl.d * cl.n;
      const size_t ib = sizeof(int32_t) * ((size_t)cl.n + 3 * nb);
      const size_t db = sizeof(double) * ((size_t)cl.n + nb);
      cl.layout.ensure(ib + db + 64, j->dev_bytes);
      char* base = reinterpret_cast<char*>(cl.layout.p);
      cl.ids = reinterpret_cast<int32_t*>(base);
      cl.in = cl.ids + cl.n;
      cl.out = cl.in + nb;
      cl.lp = cl.out + nb;
      cl.h = reinterpret_cast<double*>(base + ((ib + 15) / 16) * 16);
      cl.J = cl.h + cl.n;
      if (c == 0) {
        j->node_cls.assign(nodes_number, -1);
        j->node_pos.assign(nodes_number, -1);
      }
      for (long x = 0; x < cl.n; ++x) {
        const int id = node_ids[off_n + x];
        require(id >= 0 && id < nodes_number, "node id out of range");
        require(j->node_cls[id] < 0, "node id listed twice");
        j->node_cls[id] = c;
        j->node_pos[id] = (int32_t)x;
        h_by_id[id] = node_ampls[off_n + x];
      }
      for (size_t x = 0; x < nb; ++x) {
        require(input_pos[off_b + x] >= 0 && input_pos[off_b + x] < edges_number, "input position out of range");
        require(output_pos[off_b + x] >= 0 && output_pos[off_b + x] < edges_number, "output position out of range");
        require(lmbd_pos[off_b + x] >= 0 && lmbd_pos[off_b + x] < edges_number / 2, "lambda position out of range");
        require(lmbd_pos[off_b + x] == input_pos[off_b + x] % (edges_number / 2),
                "lambda position must equal the message position mod E");
      }
      HIPX(hipMemcpyAsync(cl.ids, node_ids + off_n, sizeof(int32_t) * cl.n, hipMemcpyHostToDevice, j->stream));
      if (nb) {
        HIPX(hipMemcpyAsync(cl.in, input_pos + off_b, sizeof(int32_t) * nb, hipMemcpyHostToDevice, j->stream));
        HIPX(hipMemcpyAsync(cl.out, output_pos + off_b, sizeof(int32_t) * nb, hipMemcpyHostToDevice, j->stream));
        HIPX(hipMemcpyAsync(cl.lp, lmbd_pos + off_b, sizeof(int32_t) * nb, hipMemcpyHostToDevice, j->stream));
        HIPX(hipMemcpyAsync(cl.J, edge_ampls + off_b, sizeof(double) * nb, hipMemcpyHostToDevice, j->stream));
      }
      HIPX(hipMemcpyAsync(cl.h, node_ampls + off_n, sizeof(double) * cl.n, hipMemcpyHostToDevice, j->stream));
      off_n += cl.n;
      off_b += nb;
    }
    const int E = j->E;
    j->edges.ensure(sizeof(int32_t) * 2 * E + sizeof(double) * (E + (size_t)nodes_number) + 64, j->dev_bytes);
    {
      char* base = reinterpret_cast<char*>(j->edges.p);
      int32_t* lhs = reinterpret_cast<int32_t*>(base);
      int32_t* rhs = lhs + E;
      double* jv = reinterpret_cast<double*>(base + ((sizeof(int32_t) * 2 * E + 15) / 16) * 16);
      double* hv = jv + E;
      for (int x = 0; x < E; ++x) {
        require(fwd_lhs[x] >= 0 && fwd_lhs[x] < nodes_number && fwd_rhs[x] >= 0 && fwd_rhs[x] < nodes_number,
                "edge endpoint out of range");
      }
      HIPX(hipMemcpyAsync(lhs, fwd_lhs, sizeof(int32_t) * E, hipMemcpyHostToDevice, j->stream));
      HIPX(hipMemcpyAsync(rhs, fwd_rhs, sizeof(int32_t) * E, hipMemcpyHostToDevice, j->stream));
      HIPX(hipMemcpyAsync(jv, fwd_j, sizeof(double) * E, hipMemcpyHostToDevice, j->stream));
      HIPX(hipMemcpyAsync(hv, h_by_id.data(), sizeof(double) * nodes_number, hipMemcpyHostToDevice, j->stream));
    }
    j->status.ensure(sizeof(BpStatus), j->dev_bytes);
    j->lmax.ensure(sizeof(double) * 128, j->dev_bytes);
    j->epart.ensure(sizeof(double) * 1024 + 64, j->dev_bytes);
    {
      std::lock_guard<std::mutex> g(g_capture_mu);
      HIPX(hipHostMalloc(reinterpret_cast<void**>(&j->st_host), sizeof(BpStatus), hipHostMallocDefault));
      HIPX(hipHostMalloc(reinterpret_cast<void**>(&j->lmax_host), sizeof(double) * 128, hipHostMallocDefault));
    }
    HIPX(hipStreamSynchronize(j->stream));  // the uploads read the caller's arrays and h_by_id
    *out = j;
  });
  if (rc != BQA_OK && j) bqa_job_destroy(j);
  return rc;
}

int bqa_job_destroy(bqa_job* j) {
  if (!j) return BQA_OK;
  return guarded([&] {
    (void)hipSetDevice(j->device);
    if (j->stream) (void)hipStreamSynchronize(j->stream);
    for (auto& g : j->graphs) (void)hipGraphExecDestroy(g.exec);
    for (auto& c : j->cls) {
      c.layout.release(j->dev_bytes);
      c.T[0].release(j->dev_bytes);
      c.T[1].release(j->dev_bytes);
    }
    DevBuf* bufs[] = {&j->msgs[0], &j->msgs[1], &j->lmbd, &j->ext, &j->canon, &j->lmax, &j->status,
                      &j->rho, &j->bloch, &j->edges, &j->epart, &j->gws, &j->cws, &j->conv, &j->conv_msgs, &j->zz,
                      &j->plist};
    for (DevBuf* b : bufs) b->release(j->dev_bytes);
    for (auto& pr : j->ev_pending) {
      (void)hipEventDestroy(pr.first);
      (void)hipEventDestroy(pr.second);
    }
    for (auto e : j->ev_pool) (void)hipEventDestroy(e);
    {
      std::lock_guard<std::mutex> g(g_capture_mu);
      if (j->st_host) (void)hipHostFree(j->st_host);
      if (j->lmax_host) (void)hipHostFree(j->lmax_host);
      if (j->stream) (void)hipStreamDestroy(j->stream);
    }
    delete j;
  });
}

int bqa_job_device_bytes(const bqa_job* j, uint64_t* bytes) {
  return guarded([&] {
    require(j && bytes, "null argument");
    *bytes = j->dev_bytes;
  });
}

int bqa_init_state(bqa_job* j) {
  return guarded([&] {
    require(j, "null job");
    dispatch(j, [&](auto r) { init_state_impl<decltype(r)>(j); });
  });
}

int bqa_run_bp(bqa_job* j, int32_t max_iters, double bp_eps, double damping, int32_t* sweeps, int32_t* converged,
               double* last_dist) {
  return guarded([&] {
    require(j && sweeps && converged && last_dist, "null argument");
    require(damping >= 0.0 && damping <= 1.0, "damping must be in [0, 1]");
    dispatch(j, [&](auto r) {
      int s = 0, c = 0;
      double d = 0.0;
      run_bp_impl<decltype(r)>(j, max_iters, bp_eps, damping, &s, &c, &d);
      *sweeps = s;
      *converged = c;
      *last_dist = d;
    });
  });
}

int bqa_simple_update(bqa_job* j, double ztime, double pinv_eps, int32_t* new_bond_dim, double* trunc_err) {
  return guarded([&] {
    require(j && new_bond_dim && trunc_err, "null argument");
    dispatch(j, [&](auto r) {
      int dim = 0;
      simple_update_impl<decltype(r)>(j, ztime, pinv_eps, 0, 0.0, 0.0, &dim, trunc_err);
      *new_bond_dim = dim;
    });
  });
}

int bqa_set_vidal_gauge(bqa_job* j, double pinv_eps) {
  return guarded([&] {
    require(j, "null job");
    dispatch(j, [&](auto r) { vidal_impl<decltype(r)>(j, pinv_eps); });
  });
}

int bqa_truncate_vidal_gauge(bqa_job* j, double pinv_eps, int32_t* new_bond_dim, double* trunc_err) {
  return guarded([&] {
    require(j && new_bond_dim && trunc_err, "null argument");
    dispatch(j, [&](auto r) {
      int dim = 0;
      truncate_vidal_impl<decltype(r)>(j, pinv_eps, &dim, trunc_err);
      *new_bond_dim = dim;
    });
  });
}

int bqa_apply_local_gates(bqa_job* j, double ztime, double xtime) {
  return guarded([&] {
    require(j, "null job");
    dispatch(j, [&](auto r) { local_update<decltype(r)>(j, kGates, j->chi, 0.0, ztime, xtime); });
  });
}

int bqa_set_symmetric_gauge(bqa_job* j) {
  return guarded([&] {
    require(j, "null job");
    dispatch(j, [&](auto r) { local_update<decltype(r)>(j, kGauge, j->chi, 0.0, 0.0, 0.0); });
  });
}

int bqa_run_layer(bqa_job* j, double xtime, double ztime, double pinv_eps, int32_t max_iters, double bp_eps,
                  double damping, int32_t* new_bond_dim, double* trunc_err, int32_t* sweeps, int32_t* converged,
                  double* last_dist) {
  return guarded([&] {
    require(j && new_bond_dim && trunc_err && sweeps && converged && last_dist, "null argument");
    require(damping >= 0.0 && damping <= 1.0, "damping must be in [0, 1]");
    dispatch(j, [&](auto r) {
      using R = decltype(r);
      int dim = 0;
      simple_update_impl<R>(j, ztime, pinv_eps, kGates | kGauge, ztime, xtime, &dim, trunc_err);
      *new_bond_dim = dim;
      int s = 0, c = 0;
      double d = 0.0;
      run_bp_impl<R>(j, max_iters, bp_eps, damping, &s, &c, &d);
      *sweeps = s;
      *converged = c;
      *last_dist = d;
    });
  });
}

int bqa_density(bqa_job* j, double* rho_out, double* bloch_out) {
  return guarded([&] {
    require(j, "null job");
    dispatch(j, [&](auto r) { density_impl<decltype(r)>(j, rho_out, bloch_out); });
  });
}

int bqa_energy(bqa_job* j, double* energy, int8_t* spins_out) {
  return guarded([&] {
    require(j && energy, "null argument");
    require(j->bloch.p != nullptr, "bqa_density must run before bqa_energy");
    HIPX(hipSetDevice(j->device));
    const int E = j->E;
    char* base = reinterpret_cast<char*>(j->edges.p);
    const int32_t* lhs = reinterpret_cast<const int32_t*>(base);
    const int32_t* rhs = lhs + E;
    const double* jv = reinterpret_cast<const double*>(base + ((sizeof(int32_t) * 2 * E + 15) / 16) * 16);
    const double* hv = jv + E;
    double* part = reinterpret_cast<double*>(j->epart.p);
    const int nb = kEnergyBlocks;
    hipLaunchKernelGGL(k_energy_partial, dim3(nb), dim3(256), 0, j->stream,
                       reinterpret_cast<const double*>(j->bloch.p), j->N, hv, lhs, rhs, jv, E, part);
    LAUNCH_CHECK();
    hipLaunchKernelGGL(k_energy_final, dim3(1), dim3(64), 0, j->stream, part, nb, part + nb);
    LAUNCH_CHECK();
    HIPX(hipMemcpyAsync(energy, part + nb, sizeof(double), hipMemcpyDeviceToHost, j->stream));
    if (spins_out) {
      std::vector<double> b((size_t)j->N * 3);
      HIPX(hipMemcpyAsync(b.data(), j->bloch.p, b.size() * sizeof(double), hipMemcpyDeviceToHost, j->stream));
      HIPX(hipStreamSynchronize(j->stream));
      for (int v = 0; v < j->N; ++v) spins_out[v] = b[(size_t)v * 3 + 2] > 0.0 ? 1 : -1;
    }
    HIPX(hipStreamSynchronize(j->stream));
  });
}

int bqa_project(bqa_job* j, int32_t count, const int32_t* qubit_ids, const int32_t* outcomes) {
  return guarded([&] {
    require(j && count >= 0 && (count == 0 || (qubit_ids && outcomes)), "null argument");
    dispatch(j, [&](auto r) { project_impl<decltype(r)>(j, count, qubit_ids, outcomes); });
  });
}

int bqa_energy_noalloc(const double* bloch_dev, int32_t nodes_number, const double* fields_dev,
                       int32_t edges_number, const int32_t* lhs_dev, const int32_t* rhs_dev, const double* j_dev,
                       double* energy_dev, void* stream) {
  return guarded([&] {
    require(nodes_number >= 0 && edges_number >= 0, "negative size");
    require(nodes_number <= kNoAllocMaxItems && edges_number <= kNoAllocMaxItems,
            "more than 2^30 qubits or edges: beyond what one call of bqa_energy_noalloc serves");
    require(energy_dev != nullptr, "null energy output");
    require(nodes_number == 0 || (bloch_dev && fields_dev), "null Bloch or field array");
    require(edges_number == 0 || (lhs_dev && rhs_dev && j_dev), "null edge array");
    double* part = scratch_on_device();
    const hipStream_t s = static_cast<hipStream_t>(stream);
    // the same kernels and grid as bqa_energy: every partial slot is written before it is read
    hipLaunchKernelGGL(k_energy_partial, dim3(kEnergyBlocks), dim3(256), 0, s, bloch_dev, nodes_number, fields_dev,
                       lhs_dev, rhs_dev, j_dev, edges_number, part);
    LAUNCH_CHECK();
    hipLaunchKernelGGL(k_energy_final, dim3(1), dim3(64), 0, s, part, kEnergyBlocks, energy_dev);
    LAUNCH_CHECK();
  });
}

int bqa_bond_dim(const bqa_job* j, int32_t* chi) {
  return guarded([&] {
    require(j && chi, "null argument");
    *chi = j->chi;
  });
}

int bqa_get_tensors(bqa_job* j, int32_t ci, double* out) {
  return guarded([&] {
    require(j && out && ci >= 0 && ci < (int)j->cls.size(), "bad class index");
    ClassDev& c = j->cls[ci];
    const size_t count = (size_t)c.n * 2 * ipow_l(j->chi, c.d);
    if (count == 0) return;
    dispatch(j, [&](auto r) { download<decltype(r)>(j, c.T[c.tcur].p, count, out); });
  });
}

int bqa_set_tensors(bqa_job* j, int32_t ci, int32_t chi, const double* in) {
  return guarded([&] {
    require(j && in && ci >= 0 && ci < (int)j->cls.size(), "bad class index");
    require(chi >= 1 && chi <= j->chimax, "bond dimension out of range");
    // every class shares the job's bond dimension: callers set all classes
    ClassDev& c = j->cls[ci];
    ensure_tensor(j, c, c.tcur, chi);
    const size_t count = (size_t)c.n * 2 * ipow_l(chi, c.d);
    if (count) dispatch(j, [&](auto r) { upload<decltype(r)>(j, c.T[c.tcur].p, count, in); });
    j->chi = chi;
  });
}

int bqa_get_messages(bqa_job* j, double* out) {
  return guarded([&] {
    require(j && out, "null argument");
    const size_t count = (size_t)j->M * j->chi * j->chi;
    dispatch(j, [&](auto r) { download<decltype(r)>(j, j->msgs[j->mcur].p, count, out); });
  });
}

int bqa_set_messages(bqa_job* j, int32_t chi, const double* in) {
  return guarded([&] {
    require(j && in, "null argument");
    require(chi == j->chi, "messages must have the tensors' bond dimension");
    ensure_msgs(j, chi);
    const size_t count = (size_t)j->M * chi * chi;
    dispatch(j, [&](auto r) { upload<decltype(r)>(j, j->msgs[j->mcur].p, count, in); });
  });
}

int bqa_get_lambdas(bqa_job* j, double* out) {
  return guarded([&] {
    require(j && out, "null argument");
    std::vector<double> h((size_t)j->E * j->lstride);
    HIPX(hipMemcpyAsync(h.data(), j->lmbd.p, h.size() * sizeof(double), hipMemcpyDeviceToHost, j->stream));
    HIPX(hipStreamSynchronize(j->stream));
    for (int e = 0; e < j->E; ++e)
      for (int k = 0; k < j->chi; ++k) {
        out[((size_t)e * j->chi + k) * 2] = h[(size_t)e * j->lstride + k];
        out[((size_t)e * j->chi + k) * 2 + 1] = 0.0;
      }
  });
}

int bqa_set_lambdas(bqa_job* j, int32_t chi, const double* in) {
  return guarded([&] {
    require(j && in, "null argument");
    require(chi == j->chi, "lambdas must have the tensors' bond dimension");
    j->lmbd.ensure((size_t)j->E * std::max(2, 2 * chi) * sizeof(double), j->dev_bytes, j->stream);
    std::vector<double> h((size_t)j->E * chi);
    for (size_t x = 0; x < h.size(); ++x) h[x] = in[2 * x];
    HIPX(hipStreamSynchronize(j->stream));
    HIPX(hipMemcpyAsync(j->lmbd.p, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice, j->stream));
    HIPX(hipStreamSynchronize(j->stream));
    j->lstride = chi;
  });
}

int bqa_timing_enable(bqa_job* j, int32_t on) {
  return guarded([&] {
    require(j, "null job");
    flush_timing(j);
    j->timing = on != 0;
    j->t_launches = 0;
    j->t_ms = 0.0;
    j->t_bytes = 0.0;
  });
}

int bqa_timing_read(bqa_job* j, int64_t* launches, double* total_ms, double* bytes) {
  return guarded([&] {
    require(j && launches && total_ms && bytes, "null argument");
    flush_timing(j);
    *launches = j->t_launches;
    *total_ms = j->t_ms;
    *bytes = j->t_bytes;
  });
}

int bqa_job_graph_stats(const bqa_job* j, int64_t* captures, int64_t* launches, int64_t* failures) {
  return guarded([&] {
    require(j && captures && launches && failures, "null argument");
    *captures = j->graph_captures;
    *launches = j->graph_launches;
    *failures = j->graph_failures;
  });
}

int bqa_batch_create(bqa_batch** out, int32_t max_jobs) {
  bqa_batch* b = nullptr;
  int rc = guarded([&] {
    require(out != nullptr, "null batch pointer");
    require(max_jobs >= 1 && max_jobs <= 4096, "max_jobs must be in [1, 4096]");
    b = new bqa_batch();
    b->max_jobs = max_jobs;
    HIPX(hipGetDevice(&b->device));
    {
      std::lock_guard<std::mutex> g(g_capture_mu);
      HIPX(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
    }
    b->ready.assign(max_jobs, nullptr);
    b->ext_done.assign(max_jobs, nullptr);
    for (auto& e : b->ready) HIPX(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    for (auto& e : b->ext_done) HIPX(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    HIPX(hipEventCreateWithFlags(&b->canon_done, hipEventDisableTiming));
    *out = b;
  });
  if (rc != BQA_OK && b) bqa_batch_destroy(b);
  return rc;
}

int bqa_batch_destroy(bqa_batch* b) {
  if (!b) return BQA_OK;
  return guarded([&] {
    (void)hipSetDevice(b->device);
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    for (auto e : b->ready)
      if (e) (void)hipEventDestroy(e);
    for (auto e : b->ext_done)
      if (e) (void)hipEventDestroy(e);
    if (b->canon_done) (void)hipEventDestroy(b->canon_done);
    b->tab.release(b->dev_bytes);
    b->ctab.release(b->dev_bytes);
    {
      std::lock_guard<std::mutex> g(g_capture_mu);
      if (b->tab_host) (void)hipHostFree(b->tab_host);
      if (b->ctab_host) (void)hipHostFree(b->ctab_host);
      if (b->stream) (void)hipStreamDestroy(b->stream);
    }
    delete b;
  });
}

int bqa_batch_run_layer(bqa_batch* b, int32_t count, bqa_job* const* jobs, const double* xtime, const double* ztime,
                        const double* pinv_eps, const int32_t* max_iters, const double* bp_eps, const double* damping,
                        int32_t* new_bond_dim, double* trunc_err, int32_t* sweeps, int32_t* converged,
                        double* last_dist, int32_t* job_status) {
  return guarded([&] {
    require(b != nullptr, "null batch");
    require(count >= 0 && count <= b->max_jobs, "count must be in [0, max_jobs]");
    if (count == 0) return;
    require(jobs && xtime && ztime && pinv_eps && max_iters && bp_eps && damping && new_bond_dim && trunc_err &&
                sweeps && converged && last_dist && job_status,
            "null argument");
    for (int k = 0; k < count; ++k) job_status[k] = BQA_OK;
    for (int k = 0; k < count; ++k) {
      require(jobs[k] != nullptr, "null job");
      require(jobs[k]->device == b->device, "the jobs of a batch run on the batch's device");
      require(damping[k] >= 0.0 && damping[k] <= 1.0, "damping must be in [0, 1]");
      require(!jobs[k]->timing, "kernel timing is not available in a batch");
      for (int q = 0; q < k; ++q) require(jobs[q] != jobs[k], "a job appears twice in the batch");
    }
    HIPX(hipSetDevice(b->device));
    // the simple update of every job on the job's own stream, so small jobs' kernels overlap:
    // extended messages of all jobs; their canonicalisers in launches shared per n (a job without a
    // split-column kernel for its chi runs its own); the lambda maxima readbacks
    std::vector<char> shared(count);
    for (int k = 0; k < count; ++k)
      shared[k] = !jobs[k]->generic(kGenCanon) && jobs[k]->E > 0 && canon_h_n(jobs[k]->chi) > 0;
    for (int k = 0; k < count; ++k)
      dispatch(jobs[k], [&](auto r) {
        ext_and_canon<decltype(r)>(jobs[k], ztime[k], pinv_eps[k], false, !shared[k]);
      });
    batch_canon(b, count, jobs, pinv_eps, shared);
    for (int k = 0; k < count; ++k) HIPX(hipStreamSynchronize(jobs[k]->stream));
    // the global truncation rule of each job, then its fused absorb + gates + gauge step
    for (int k = 0; k < count; ++k)
      dispatch(jobs[k], [&](auto r) {
        bqa_job* j = jobs[k];
        const int n2 = 2 * j->chi;
        const int dim = truncation_dim(j, n2, pinv_eps[k], &trunc_err[k]);
        j->lstride = n2;
        local_update<decltype(r)>(j, kAbsorb | kGates | kGauge, dim, ztime[k], ztime[k], xtime[k]);
        new_bond_dim[k] = dim;
        HIPX(hipEventRecord(b->ready[k], j->stream));
        HIPX(hipStreamWaitEvent(b->stream, b->ready[k], 0));
      });
    // BP on the batch stream: shared launches and every job's own BP launches in sweep order
    std::vector<hipStream_t> own(count);
    for (int k = 0; k < count; ++k) {
      own[k] = jobs[k]->stream;
      jobs[k]->stream = b->stream;
    }
    try {
      batch_bp(b, count, jobs, max_iters, bp_eps, damping, sweeps, converged, last_dist, job_status);
    } catch (...) {
      (void)hipStreamSynchronize(b->stream);
      for (int k = 0; k < count; ++k) jobs[k]->stream = own[k];
      throw;
    }
    for (int k = 0; k < count; ++k) jobs[k]->stream = own[k];
  });
}

int bqa_tuning_report(char* buf, int64_t capacity, int64_t* needed) {
  return guarded([&] {
    require(needed != nullptr && capacity >= 0 && (capacity == 0 || buf != nullptr), "null argument");
    std::string out = "[";
    {
      std::lock_guard<std::mutex> g(g_tune_mu);
      bool firstrow = true;
      for (const auto& kv : g_tune) {
        const TuneKey& k = kv.first;
        const TuneEntry& e = kv.second;
        if (e.chosen < 0) continue;
        char head[256];
        std::snprintf(head, sizeof head,
                      "%s{\"device\": %d, \"precision\": \"%s\", \"degree\": %d, \"chi\": %d, "
                      "\"n_log2\": %d, \"chosen\": %d, \"variants\": [",
                      firstrow ? "" : ", ", k.dev, k.prec == BQA_PREC_SINGLE ? "single" : "double", k.d, k.chi,
                      k.nlog, e.chosen);
        out += head;
        for (size_t v = 0; v < e.us.size(); ++v) {
          const char* name = k.prec == BQA_PREC_SINGLE ? bp_variant_name<float>(k.d, k.chi, (int)v)
                                                       : bp_variant_name<double>(k.d, k.chi, (int)v);
          char item[256];
          std::snprintf(item, sizeof item, "%s{\"name\": \"%s\", \"us\": %.3f}", v ? ", " : "", name, e.us[v]);
          out += item;
        }
        out += "]}";
        firstrow = false;
      }
    }
    out += "]";
    *needed = (int64_t)out.size() + 1;
    if (capacity >= *needed) std::memcpy(buf, out.c_str(), out.size() + 1);
  });
}

int bqa_tuning_reset(void) {
  return guarded([&] {
    std::lock_guard<std::mutex> g(g_tune_mu);
    g_tune.clear();
  });
}

int bqa_sv_create(bqa_sv** out, int32_t qubits, const double* fields, int32_t edges_number, const int32_t* lhs,
                  const int32_t* rhs, const double* couplings) {
  bqa_sv* s = nullptr;
  int rc = guarded([&] {
    require(out != nullptr, "null simulator pointer");
    require(qubits >= 1 && qubits <= kSvMaxQubits, "the exact simulator takes 1 to 32 qubits");
    require(edges_number >= 0 && edges_number <= kSvMaxEdges, "the exact simulator takes at most 2048 edges");
    require(fields != nullptr && (edges_number == 0 || (lhs && rhs && couplings)), "null argument");
    for (int e = 0; e < edges_number; ++e)
      require(lhs[e] >= 0 && lhs[e] < qubits && rhs[e] >= 0 && rhs[e] < qubits && lhs[e] != rhs[e],
              "edge endpoint out of range");
    int dev_count = 0;
    HIPX(hipGetDeviceCount(&dev_count));
    require(dev_count > 0, "no HIP device");
    s = new bqa_sv();
    s->n = qubits;
    s->E = edges_number;
    HIPX(hipGetDevice(&s->device));
    {
      std::lock_guard<std::mutex> g(g_capture_mu);
      HIPX(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
    }
    const size_t dim = (size_t)1 << qubits;
    s->psi.ensure(dim * sizeof(cplx<double>), s->dev_bytes);
    s->ez.ensure(dim * sizeof(double), s->dev_bytes);
    s->part.ensure((size_t)kSvMaxGrid * 4 * kSvTileBits * sizeof(double), s->dev_bytes);
    const size_t gb = sizeof(double) * ((size_t)qubits + edges_number) + sizeof(int32_t) * 2 * (size_t)edges_number + 64;
    s->graph.ensure(gb, s->dev_bytes);
    char* base = reinterpret_cast<char*>(s->graph.p);
    double* h = reinterpret_cast<double*>(base);
    double* jv = h + qubits;
    int32_t* a = reinterpret_cast<int32_t*>(jv + edges_number);
    int32_t* b = a + edges_number;
    HIPX(hipMemcpyAsync(h, fields, sizeof(double) * qubits, hipMemcpyHostToDevice, s->stream));
    if (edges_number) {
      HIPX(hipMemcpyAsync(jv, couplings, sizeof(double) * edges_number, hipMemcpyHostToDevice, s->stream));
      HIPX(hipMemcpyAsync(a, lhs, sizeof(int32_t) * edges_number, hipMemcpyHostToDevice, s->stream));
      HIPX(hipMemcpyAsync(b, rhs, sizeof(int32_t) * edges_number, hipMemcpyHostToDevice, s->stream));
    }
    hipLaunchKernelGGL(k_sv_energy, dim3(grid_for((long)std::min<size_t>(dim, 1L << 30), kSvThreads, 16384)),
                       dim3(kSvThreads), 0, s->stream, reinterpret_cast<double*>(s->ez.p), qubits, h, edges_number, a,
                       b, jv);
    LAUNCH_CHECK();
    HIPX(hipStreamSynchronize(s->stream));
    *out = s;
  });
  if (rc != BQA_OK && s) bqa_sv_destroy(s);
  return rc;
}

int bqa_sv_destroy(bqa_sv* s) {
  if (!s) return BQA_OK;
  return guarded([&] {
    (void)hipSetDevice(s->device);
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    s->psi.release(s->dev_bytes);
    s->ez.release(s->dev_bytes);
    s->graph.release(s->dev_bytes);
    s->part.release(s->dev_bytes);
    if (s->stream) {
      std::lock_guard<std::mutex> g(g_capture_mu);
      (void)hipStreamDestroy(s->stream);
    }
    delete s;
  });
}

int bqa_sv_init(bqa_sv* s) {
  return guarded([&] {
    require(s, "null simulator");
    HIPX(hipSetDevice(s->device));
    const long dim = 1L << s->n;
    hipLaunchKernelGGL(k_sv_init, dim3(grid_for(std::min(dim, 1L << 30), 256, 16384)), dim3(256), 0, s->stream,
                       reinterpret_cast<cplx<double>*>(s->psi.p), s->n);
    LAUNCH_CHECK();
    HIPX(hipStreamSynchronize(s->stream));
  });
}

int bqa_sv_run_layer(bqa_sv* s, double xtime, double ztime) {
  return guarded([&] {
    require(s, "null simulator");
    HIPX(hipSetDevice(s->device));
    sv_passes<0>(s, reinterpret_cast<const double*>(s->ez.p), ztime, xtime, [](int, int, int) {});
    HIPX(hipStreamSynchronize(s->stream));
  });
}

int bqa_sv_density(bqa_sv* s, double* rho_out, double* bloch_out) {
  return guarded([&] {
    require(s, "null simulator");
    HIPX(hipSetDevice(s->device));
    DevBuf res;
    size_t scratch_bytes = 0;
    res.ensure((size_t)s->n * 11 * sizeof(double), scratch_bytes);
    double* rho = reinterpret_cast<double*>(res.p);
    double* bloch = rho + 8 * (size_t)s->n;
    try {
      sv_passes<1>(s, nullptr, 0.0, 0.0, [&](int blocks, int q0, int nb) {
        hipLaunchKernelGGL(k_sv_finish, dim3(1), dim3(64), 0, s->stream, reinterpret_cast<const double*>(s->part.p),
                           blocks, q0, nb, rho, bloch);
        LAUNCH_CHECK();
      });
      if (rho_out)
        HIPX(hipMemcpyAsync(rho_out, rho, (size_t)s->n * 8 * sizeof(double), hipMemcpyDeviceToHost, s->stream));
      if (bloch_out)
        HIPX(hipMemcpyAsync(bloch_out, bloch, (size_t)s->n * 3 * sizeof(double), hipMemcpyDeviceToHost, s->stream));
      HIPX(hipStreamSynchronize(s->stream));
    } catch (...) {
      (void)hipStreamSynchronize(s->stream);
      res.release(scratch_bytes);
      throw;
    }
    res.release(scratch_bytes);
  });
}

int bqa_sv_get_state(bqa_sv* s, double* out) {
  return guarded([&] {
    require(s && out, "null argument");
    require(s->n <= 26, "bqa_sv_get_state copies at most 2^26 amplitudes");
    HIPX(hipSetDevice(s->device));
    HIPX(hipMemcpyAsync(out, s->psi.p, ((size_t)1 << s->n) * sizeof(cplx<double>), hipMemcpyDeviceToHost, s->stream));
    HIPX(hipStreamSynchronize(s->stream));
  });
}

int bqa_copy_calibrate(uint64_t bytes, int32_t repeats, double* ms_per_copy) {
  return guarded([&] {
    require(ms_per_copy && bytes >= 16 && repeats > 0, "bad argument");
    void *a = nullptr, *b = nullptr;
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    auto cleanup = [&] {
      if (st) (void)hipStreamSynchronize(st);
      if (e0) (void)hipEventDestroy(e0);
      if (e1) (void)hipEventDestroy(e1);
      std::lock_guard<std::mutex> g(g_capture_mu);
      if (a) (void)hipFree(a);
      if (b) (void)hipFree(b);
      if (st) (void)hipStreamDestroy(st);
    };
    try {
      {
        std::lock_guard<std::mutex> g(g_capture_mu);
        HIPX(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
        HIPX(hipMalloc(&a, bytes));
        HIPX(hipMalloc(&b, bytes));
      }
      const long n = (long)(bytes / 16);
      HIPX(hipEventCreate(&e0));
      HIPX(hipEventCreate(&e1));
      hipLaunchKernelGGL(k_copy_bytes, dim3(grid_for(n, 256, 16384)), dim3(256), 0, st, (const uint4*)a, (uint4*)b, n);
      HIPX(hipEventRecord(e0, st));
      for (int r = 0; r < repeats; ++r)
        hipLaunchKernelGGL(k_copy_bytes, dim3(grid_for(n, 256, 16384)), dim3(256), 0, st, (const uint4*)a,
                           (uint4*)b, n);
      LAUNCH_CHECK();
      HIPX(hipEventRecord(e1, st));
      HIPX(hipEventSynchronize(e1));
      float ms = 0.f;
      HIPX(hipEventElapsedTime(&ms, e0, e1));
      *ms_per_copy = ms / repeats;
    } catch (...) {
      cleanup();
      throw;
    }
    cleanup();
  });
}

}  // extern "C"
