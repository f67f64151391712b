// ctm_api.hip -- host orchestration and C ABI of the batched T-matrix path.
//
// The reference runs one dependency chain per grain on CPU threads
// (InteractionTask -> TMatrixM0Task per order until converged, then
// TMatrixMAllTask per m; TaskManager.hpp:1146-1197,1314-1327).  Here a batch
// of grains runs as a few wide launches:
//   phase 1 (orders): for every unconverged grain, all orders of a window
//     [lo, hi] are independent m = 0 problems and run in parallel; a
//     per-grain scan then applies the reference's sequential convergence
//     test (TMatrixResource.hpp:743-775) in order.  Grains that did not
//     converge get the next window.
//   phase 2 (all m): every (grain, m) block at the converged order.
// Grain-independent tables (Gauss-Legendre nodes per G, Wigner d per (N, m))
// are computed once per batch and shared, like the reference's
// GaussBasedResources and WignerDResources.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <climits>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <tuple>
#include <unordered_map>
#include <utility>
#include <vector>

#include "../../include/costuum_amd.h"
#include "ctm_device.hpp"
#include "ctm_kernels.hpp"
#include "ctm_observables.hpp"
#include "ctm_average.hpp"
#include "ctm_host.hpp"
#include "ctm_tangent.hpp"
#include "ctm_mixed.hpp"
#include "ctm_gj.hpp"
#include "ctm_plan.hpp"

using namespace ctm;

namespace {

struct DevBuf {
  void *p = nullptr;
  size_t cap = 0;
  hipError_t ensure(size_t bytes) {
    if (bytes <= cap) return hipSuccess;
    if (p) {
      // launches on any of the context's streams may still read the old
      // buffer
      hipError_t e = hipDeviceSynchronize();
      if (e != hipSuccess) return e;
      e = hipFree(p);
      if (e != hipSuccess) return e;
      p = nullptr;
      cap = 0;
    }
    size_t want = std::max(bytes, (size_t)256);
    want = want + want / 4;
    hipError_t e = hipMalloc(&p, want);
    if (e != hipSuccess) return e;
    cap = want;
    return hipSuccess;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    cap = 0;
  }
  // Frees a buffer that holds more than twice `bytes` and allocates it
  // again for `bytes` (contents are lost).
  hipError_t fit(size_t bytes) {
    if (cap <= 2 * std::max(bytes, (size_t)256)) return hipSuccess;
    hipError_t e = hipDeviceSynchronize();
    if (e != hipSuccess) return e;
    release();
    return ensure(bytes);
  }
  template <typename T>
  T *as() const {
    return reinterpret_cast<T *>(p);
  }
};

// Pinned host memory for uploads (see stage_copy).
struct Staging {
  char *p = nullptr;
  size_t cap = 0, used = 0;
  void release() {
    if (p) (void)hipHostFree(p);
    p = nullptr;
    cap = used = 0;
  }
};

// Grain-independent table caches (per batch).  Gauss-Legendre keys are
// found by G directly; Wigner keys through one slot array of N + 1 entries
// (one per m) per (N, G), found through a by-N entry (the last G seen for
// that N) or a hash map, the last (N, G) looked up kept at hand (item lists
// ask for all m of one (N, G) in a row).
struct SharedTables {
  std::vector<int32_t> gl_by_g;                 // G -> key index or -1
  std::unordered_map<uint64_t, int32_t> wig_base;  // (N, G) -> first slot
  std::vector<int32_t> wig_slot;                // key index or -1
  std::vector<uint8_t> wig_full;  // per slot base: every m registered
  std::vector<GlKey> gl_keys;
  std::vector<WigKey> wig_keys;
  int64_t gl_size = 0, wig_size = 0;
  size_t gl_done = 0, wig_done = 0;  // keys already computed on device
  int last_n = -1, last_g = -1, last_base = 0;
  std::vector<int64_t> wig_by_n;  // N -> (G << 32) | first slot, or -1
  void clear() {
    wig_by_n.clear();
    gl_by_g.clear();
    wig_base.clear();
    wig_slot.clear();
    wig_full.clear();
    gl_keys.clear();
    wig_keys.clear();
    gl_size = wig_size = 0;
    gl_done = wig_done = 0;
    last_n = last_g = -1;
  }
  int gl(int G) {
    if (G >= (int)gl_by_g.size()) gl_by_g.resize(G + 1, -1);
    if (gl_by_g[G] >= 0) return gl_by_g[G];
    GlKey k{};
    k.G = G;
    k.off = gl_size;
    gl_size += 2 * (int64_t)G;
    gl_keys.push_back(k);
    gl_by_g[G] = (int)gl_keys.size() - 1;
    return gl_by_g[G];
  }
  int wig(int N, int m, int G) {
    if (N != last_n || G != last_g) {
      if (N < (int)wig_by_n.size() && wig_by_n[N] >= 0 &&
          (int)(wig_by_n[N] >> 32) == G) {
        last_base = (int)(wig_by_n[N] & 0xffffffff);
      } else {
        const uint64_t key =
            (uint64_t)(uint32_t)N | ((uint64_t)(uint32_t)G << 32);
        auto f = wig_base.find(key);
        if (f != wig_base.end()) {
          last_base = f->second;
        } else {
          last_base = (int)wig_slot.size();
          wig_slot.resize(wig_slot.size() + N + 1, -1);
          wig_full.resize(wig_slot.size(), 0);
          wig_base.emplace(key, last_base);
        }
        if (N >= (int)wig_by_n.size()) wig_by_n.resize(N + 1, -1);
        wig_by_n[N] = ((int64_t)G << 32) | (int64_t)last_base;
      }
      last_n = N;
      last_g = G;
    }
    int32_t &slot = wig_slot[last_base + m];
    if (slot >= 0) return slot;
    return new_wig(N, m, G, slot);
  }
  // Registers the keys of every m = 0..N of (N, G); returns the slot base:
  // the key of m is wig_slot[base + m].
  int wig_all(int N, int G) {
    wig(N, 0, G);
    const int base = last_base;
    if (wig_full[base]) return base;
    for (int m = 1; m <= N; ++m) {
      int32_t &slot = wig_slot[base + m];
      if (slot < 0) new_wig(N, m, G, slot);
    }
    wig_full[base] = 1;
    return base;
  }
  int new_wig(int N, int m, int G, int32_t &slot) {
    const int g = gl(G);
    WigKey k{};
    k.N = N;
    k.m = m;
    k.G = G;
    k.gl_off = gl_keys[g].off;
    k.off = wig_size;
    const int nmin = m > 0 ? m : 1;
    wig_size += 2 * (int64_t)(N + 1 - nmin) * G;
    wig_keys.push_back(k);
    slot = (int)wig_keys.size() - 1;
    return slot;
  }
};

// One set of work lists: table sets, items, tiles, solve tasks.
struct WorkList {
  std::vector<TableSet> sets;
  std::vector<Item> items;
  std::vector<int32_t> order;        // items sorted by assembly cost
  // Assembly order: five parts, by micro-tile 3, 2, 1 (see assemble_micro)
  // and the packed blocks with m > 0, then m = 0, each sorted by G (largest
  // first); tile_prefix holds one prefix sum per tiled part (n_part[k] + 1
  // entries each, parts 0-2 in that order).
  std::vector<int32_t> tile_prefix;
  int n_part[5] = {0, 0, 0, 0, 0};
  // algorithmic flops and bytes of each assembly part (profiler credit)
  double f_asm[5] = {0., 0., 0., 0., 0.}, b_asm[5] = {0., 0., 0., 0., 0.};
  std::vector<int32_t> solve;        // item * 2 + class, by block order
  std::vector<int32_t> scratch_a, scratch_b, scratch_c;  // build_tasks
  int64_t table_size = 0, q_size = 0, t_size = 0;
  int32_t n_nodes = 0;  // quadrature nodes over all sets
  void clear() {
    n_nodes = 0;
    sets.clear();
    items.clear();
    order.clear();
    tile_prefix.clear();
    solve.clear();
    table_size = q_size = t_size = 0;
  }
  int add_set(int grain, int N, int G, int64_t gl_off) {
    TableSet s{};
    s.grain = grain;
    s.N = N;
    s.G = G;
    s.gl_off = gl_off;
    s.off = table_size;
    s.node0 = n_nodes;
    n_nodes += G;
    table_size += (int64_t)kNodeScalars * G + (int64_t)kBesselArrays * N * G;
    sets.push_back(s);
    return (int)sets.size() - 1;
  }
  int add_item(int grain, int N, int m, int G, int set, int wig,
               bool want_t) {
    Item it{};
    it.grain = grain;
    it.N = N;
    it.m = m;
    it.G = G;
    it.set = set;
    it.wig = wig;
    const int nmin = m > 0 ? m : 1;
    const int64_t nm = N + 1 - nmin;
    it.q_off = q_size;
    q_size += 4 * nm * nm;
    it.t_off = want_t ? t_size : -1;
    if (want_t) t_size += 2 * nm * nm;
    items.push_back(it);
    return (int)items.size() - 1;
  }
};

// Per-kernel timing with HIP events recorded on the context's stream around
// every launch, and the algorithmic work each launch stands for.
struct Profiler {
  struct Rec {
    int kernel;
    hipEvent_t a, b;
    double flops, bytes;
  };
  bool on = false;
  std::vector<hipEvent_t> pool;
  std::vector<Rec> pending;
  uint64_t launches[CTM_KERNEL_COUNT] = {};
  double ms[CTM_KERNEL_COUNT] = {};
  double flops[CTM_KERNEL_COUNT] = {};
  double bytes[CTM_KERNEL_COUNT] = {};
  hipEvent_t take() {
    if (pool.empty()) {
      hipEvent_t e = nullptr;
      if (hipEventCreate(&e) != hipSuccess) return nullptr;
      return e;
    }
    hipEvent_t e = pool.back();
    pool.pop_back();
    return e;
  }
  // Records the start event; returns the record index or -1.
  int begin(int kernel, hipStream_t s, double f, double b) {
    if (!on) return -1;
    Rec r{kernel, take(), take(), f, b};
    if (!r.a || !r.b) return -1;
    if (hipEventRecord(r.a, s) != hipSuccess) return -1;
    pending.push_back(r);
    return (int)pending.size() - 1;
  }
  void end(int idx, hipStream_t s) {
    if (idx >= 0) (void)hipEventRecord(pending[idx].b, s);
  }
  // Folds finished records into the totals (the stream must be idle).
  void collect() {
    for (const Rec &r : pending) {
      float t = 0.f;
      if (hipEventElapsedTime(&t, r.a, r.b) == hipSuccess) {
        launches[r.kernel] += 1;
        ms[r.kernel] += t;
        flops[r.kernel] += r.flops;
        bytes[r.kernel] += r.bytes;
      }
      pool.push_back(r.a);
      pool.push_back(r.b);
    }
    pending.clear();
  }
  void reset() {
    collect();
    for (int k = 0; k < CTM_KERNEL_COUNT; ++k) {
      launches[k] = 0;
      ms[k] = flops[k] = bytes[k] = 0.;
    }
  }
  void release() {
    collect();
    for (hipEvent_t e : pool) (void)hipEventDestroy(e);
    pool.clear();
  }
};

}  // namespace

// Side streams that take independent launches of one phase (assembly
// parts, solve buckets, observable buckets) beside the context's stream.
constexpr int kLanes = 4;

// Fixed device scratch of ctm_observables_noalloc: 64 MiB per device, held
// from the library's first use on that device until the process ends.  In
// host code `scratch` names a host-side stand-in, not device memory: the
// device address of each GPU's copy comes from hipGetSymbolAddress (taken
// once per device, with that device current, in ctm_create) and only that
// address is handed to kernels and copies.  Every call on a device shares
// that device's array.
constexpr size_t kNoAllocScratchBytes = (size_t)64 << 20;
static __device__ __align__(256) char scratch[kNoAllocScratchBytes];
constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static char *g_scratch_ptr[kMaxDevices] = {};

// The device address of `scratch` on the current device `device`.
static hipError_t scratch_address(int device, char **ptr) {
  *ptr = nullptr;
  if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    void *p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch_ptr[device] = static_cast<char *>(p);
  }
  *ptr = g_scratch_ptr[device];
  return hipSuccess;
}

// Observable source of ctm_observables_derivative: T +- s dT in d_tpm.
constexpr int kObsSourcePM = -2;

struct ctm_context {
  int device = 0;
  hipStream_t stream = nullptr;
  // lane 0 is `stream`; lanes 1..kLanes-1 are side streams of the default
  // priority, ordered against `stream` by fork_lanes / join_lanes
  hipStream_t side[kLanes - 1] = {};
  hipEvent_t ev_fork = nullptr;
  hipEvent_t ev_join[kLanes - 1] = {};
  // end of each assembly part (run_items): a solve bucket waits only for
  // the parts that assembled its blocks
  hipEvent_t ev_part[5] = {};
  // end of the solve bucket of orders 49..64 (the all-m buckets of smaller
  // blocks start after it)
  hipEvent_t ev_bucket = {};
  std::string err;
  SharedTables shared;
  WorkList work;
  DevBuf d_gl, d_glkeys, d_wig, d_wigkeys, d_tables, d_sets, d_items,
      d_tiles, d_solve, d_q, d_t, d_qparts, d_state, d_active, d_winoff,
      d_winitems, d_gioff, d_gitems, d_gioff_tmp, d_params, d_ang, d_pt,
      d_x, d_spec, d_cgkeys, d_cg, d_cgidx, d_avgtasks, d_avg, d_aoff, d_pn,
      d_dist, d_obsitems, d_query, d_qout, d_gsolve, d_ttables, d_dq, d_ainv,
      d_dt, d_tparts, d_dsum, d_ttiles, d_tprefix, d_wigfresh, d_x32, d_xexp,
      d_rbuf, d_flags, d_fallback;
  // mixed precision for the current batch call, and its block counts since
  // the batch started (blocks through the FP32 elimination, blocks re-solved
  // in FP64 after the refinement stalled)
  bool mixed = false;
  int64_t mixed_blocks = 0, mixed_fallback = 0;
  // T +- s dT of ctm_observables_derivative (T layout), per-grain scales s,
  // and the two evaluations of the observables
  DevBuf d_tpm, d_pmscale, d_pmbuf;
  std::vector<int32_t> fallback;
  // Results of the last batch (host).
  int n_grains = 0;
  std::vector<uint32_t> nmax, ngauss;
  std::vector<int32_t> status;
  std::vector<Item> final_items;          // phase-2 items
  std::vector<int32_t> grain_items_off;   // per grain into final_items
  std::vector<int32_t> grain_items;
  GrainParams gp_saved{};  // context copies of the last batch's inputs
  bool averaged = false;   // d_avg holds ensemble T-matrices of the batch
  // forward-mode tangents of the batch (d_dt) and the observable source:
  // -1 the batch's T-matrices, 0..3 the tangent along one parameter
  bool have_tangents = false;
  int obs_source = -1;
  std::vector<double> dqsca, dqext;  // [n][4] tangents of the full-T sums
  std::vector<int64_t> aoff;  // full-layout offsets per final item
  std::vector<int32_t> obs_items;
  Profiler prof;
  Staging staging;
  Staging readback;  // pinned landing area of device-to-host copies
  char *scratch = nullptr;  // this device's fixed scratch (device address)
  int plans = 0;            // live ctm_plan objects on this context
};

#define CTM_HIP(expr)                                                   \
  do {                                                                  \
    hipError_t _e = (expr);                                             \
    if (_e != hipSuccess) {                                             \
      ctx->err = std::string(#expr) + ": " + hipGetErrorString(_e);     \
      return CTM_ERR_HIP;                                               \
    }                                                                   \
  } while (0)

// Launch `call` on stream `strm` between two profiler events when profiling
// is on.
#define CTM_LAUNCH_ON(strm, kernel, flops, bytes, call)           \
  do {                                                            \
    const int _pi = ctx->prof.begin(kernel, strm, flops, bytes);  \
    CTM_HIP(call);                                                \
    ctx->prof.end(_pi, strm);                                     \
  } while (0)
#define CTM_LAUNCH(kernel, flops, bytes, call) \
  CTM_LAUNCH_ON(ctx->stream, kernel, flops, bytes, call)

namespace {

hipStream_t lane(const ctm_context *ctx, int k) {
  return k == 0 ? ctx->stream : ctx->side[k - 1];
}

// Orders the side streams after everything queued on the context's stream
// so far.  The launches of a phase that follow are independent of each
// other (disjoint outputs), so a launch's tail overlaps the next one's start.
hipError_t fork_lanes(ctm_context *ctx) {
  hipError_t e = hipEventRecord(ctx->ev_fork, ctx->stream);
  for (int k = 0; e == hipSuccess && k < kLanes - 1; ++k)
    e = hipStreamWaitEvent(ctx->side[k], ctx->ev_fork, 0);
  return e;
}

// Orders the context's stream after everything queued on the side streams:
// every fork_lanes is followed by join_lanes before the context's stream is
// synchronised or its results are read.
hipError_t join_lanes(ctm_context *ctx) {
  hipError_t e = hipSuccess;
  for (int k = 0; e == hipSuccess && k < kLanes - 1; ++k) {
    e = hipEventRecord(ctx->ev_join[k], ctx->side[k]);
    if (e == hipSuccess) e = hipStreamWaitEvent(ctx->stream, ctx->ev_join[k], 0);
  }
  return e;
}

// Host-to-device copy through the context's pinned staging arena: a copy
// from pinned memory runs asynchronously, a pageable one stages through the
// runtime and holds the host until the stream reaches it.  Arena space is
// handed out in order and reused only after the stream has drained, so no
// pending copy sees its source overwritten.
hipError_t stage_copy(ctm_context *ctx, void *dst, const void *src,
                      size_t bytes) {
  if (bytes == 0) return hipSuccess;
  Staging &st = ctx->staging;
  const size_t need = (bytes + 255) & ~(size_t)255;
  if (st.used + need > st.cap) {
    hipError_t e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) return e;
    st.used = 0;
    if (need > st.cap) {
      st.release();
      const size_t want = std::max(2 * need, (size_t)16 << 20);
      e = hipHostMalloc(reinterpret_cast<void **>(&st.p), want,
                        hipHostMallocDefault);
      if (e != hipSuccess) {
        st.p = nullptr;
        return e;
      }
      st.cap = want;
    }
  }
  char *h = st.p + st.used;
  std::memcpy(h, src, bytes);
  st.used += need;
  return hipMemcpyAsync(dst, h, bytes, hipMemcpyHostToDevice, ctx->stream);
}

// Waits for the context's stream; the staging arena is free afterwards.
hipError_t drain(ctm_context *ctx) {
  hipError_t e = hipStreamSynchronize(ctx->stream);
  if (e == hipSuccess) ctx->staging.used = 0;
  return e;
}

template <typename T>
hipError_t upload(ctm_context *ctx, DevBuf &b, const std::vector<T> &v) {
  hipError_t e = b.ensure(std::max<size_t>(v.size(), 1) * sizeof(T));
  if (e != hipSuccess) return e;
  return stage_copy(ctx, b.p, v.data(), v.size() * sizeof(T));
}

// Device-to-host copies of k device ranges src[i] (bytes[i] each) to dst[i]
// through pinned memory, then a drain of the stream (a pageable destination
// makes the runtime stage the copy itself, at a fraction of the rate).
hipError_t read_back(ctm_context *ctx, int k, void *const *dst,
                     const void *const *src, const size_t *bytes) {
  Staging &rb = ctx->readback;
  size_t total = 0;
  for (int i = 0; i < k; ++i) total += (bytes[i] + 255) & ~(size_t)255;
  if (total > rb.cap) {
    hipError_t e = drain(ctx);
    if (e != hipSuccess) return e;
    rb.release();
    const size_t want = std::max(total, (size_t)1 << 20);
    e = hipHostMalloc(reinterpret_cast<void **>(&rb.p), want,
                      hipHostMallocDefault);
    if (e != hipSuccess) {
      rb.p = nullptr;
      return e;
    }
    rb.cap = want;
  }
  size_t off = 0;
  for (int i = 0; i < k; ++i) {
    hipError_t e = hipMemcpyAsync(rb.p + off, src[i], bytes[i],
                                  hipMemcpyDeviceToHost, ctx->stream);
    if (e != hipSuccess) return e;
    off += (bytes[i] + 255) & ~(size_t)255;
  }
  hipError_t e = drain(ctx);
  if (e != hipSuccess) return e;
  off = 0;
  for (int i = 0; i < k; ++i) {
    std::memcpy(dst[i], rb.p + off, bytes[i]);
    off += (bytes[i] + 255) & ~(size_t)255;
  }
  return hipSuccess;
}

// Computes the grain-independent tables whose keys were added since the
// last call.
int flush_shared(ctm_context *ctx) {
  SharedTables &sh = ctx->shared;
  const size_t ngl = sh.gl_keys.size(), nwig = sh.wig_keys.size();
  if (sh.gl_done == ngl && sh.wig_done == nwig) return CTM_OK;
  // The caches are append-only within a batch; grow by copying.
  auto grow = [&](DevBuf &b, size_t bytes) -> hipError_t {
    if (bytes <= b.cap) return hipSuccess;
    DevBuf nb;
    hipError_t e = nb.ensure(bytes * 2);
    if (e != hipSuccess) return e;
    if (b.p) {
      e = hipMemcpyAsync(nb.p, b.p, b.cap, hipMemcpyDeviceToDevice,
                         ctx->stream);
      if (e != hipSuccess) return e;
      e = hipStreamSynchronize(ctx->stream);
      if (e != hipSuccess) return e;
      b.release();
    }
    b = nb;
    return hipSuccess;
  };
  CTM_HIP(grow(ctx->d_gl, sh.gl_size * sizeof(double)));
  CTM_HIP(grow(ctx->d_wig, std::max<int64_t>(sh.wig_size, 1) * sizeof(double)));
  if (sh.gl_done < ngl) {
    std::vector<GlKey> fresh(sh.gl_keys.begin() + sh.gl_done, sh.gl_keys.end());
    CTM_HIP(upload(ctx, ctx->d_glkeys, fresh));
    CTM_LAUNCH(CTM_KERNEL_GAUSS_LEGENDRE, 0., 0.,
               launch_gauss_legendre(ctx->d_glkeys.as<GlKey>(),
                                     (int)fresh.size(), ctx->d_gl.as<double>(),
                                     ctx->stream));
    sh.gl_done = ngl;
  }
  if (sh.wig_done < nwig) {
    std::vector<WigKey> fresh(sh.wig_keys.begin() + sh.wig_done,
                              sh.wig_keys.end());
    CTM_HIP(upload(ctx, ctx->d_wigfresh, fresh));
    CTM_LAUNCH(CTM_KERNEL_WIGNER, 0., 0.,
               launch_wigner(ctx->d_wigfresh.as<WigKey>(), (int)fresh.size(),
                             ctx->d_gl.as<double>(), ctx->d_wig.as<double>(),
                             ctx->stream));
    sh.wig_done = nwig;
  }
  // all Wigner keys on device for the assembly kernel
  CTM_HIP(upload(ctx, ctx->d_wigkeys, sh.wig_keys));
  return CTM_OK;
}

// Assembly part of an item: micro-tile 3, 2, 1 -> 0, 1, 2; packed -> 3
// (m > 0) or 4 (m = 0).
inline int item_part(const Item &it) {
  const int nm = it.N + 1 - (it.m > 0 ? it.m : 1);
  const int micro = assemble_micro(nm, it.m == 0);
  return micro == 4 ? (it.m == 0 ? 4 : 3) : 3 - micro;
}

// Builds the cost-sorted item order with the prefix sum of assembly tiles
// (largest G first) and the solve task list (largest blocks first), both by
// counting sort, for the items i0 .. i1 - 1 (global indices).  The items are
// read once; the sorts run over packed per-item words.
void build_tasks(WorkList &w, int i0, int i1) {
  const int n = i1 - i0;
  // per item: block order, an m = 0 flag and the part (micro-tile 3, 2, 1
  // -> 0, 1, 2; packed -> 3 (m > 0) or 4 (m = 0)); G (then the sort key)
  std::vector<int32_t> &info = w.scratch_a, &key = w.scratch_b;
  info.resize(n);
  key.resize(n);
  int gmax = 0, nmmax = 0;
  for (int k = 0; k < 5; ++k) {
    w.n_part[k] = 0;
    w.f_asm[k] = w.b_asm[k] = 0.;
  }
  // Algorithmic work of the assembly launches (the per-block counts of the
  // flop model F_T: 96 G nm^2 for an m > 0 block, 32 G n^2 for an m = 0
  // order) and their compulsory HBM traffic (tables and Wigner values read
  // once, Q/RgQ written once).
  for (int i = 0; i < n; ++i) {
    const Item &it = w.items[i0 + i];
    const int nm = it.N + 1 - (it.m > 0 ? it.m : 1);
    const int part = item_part(it);
    info[i] = nm | (part << 16) | (it.m == 0 ? (1 << 30) : 0);
    key[i] = it.G;
    w.n_part[part]++;
    const double d = nm;
    w.f_asm[part] += (it.m > 0 ? 96. : 32.) * it.G * d * d;
    w.b_asm[part] += 16. * d * it.G + 64. * d * d;
    gmax = std::max(gmax, it.G);
    nmmax = std::max(nmmax, nm);
  }
  // counting sort by (part, G descending)
  const int nkey = 5 * (gmax + 1);
  std::vector<int32_t> &cnt = w.scratch_c;
  cnt.assign(std::max(nkey, nmmax + 1) + 1, 0);
  for (int i = 0; i < n; ++i) {
    key[i] = ((info[i] >> 16) & 7) * (gmax + 1) + (gmax - key[i]);
    cnt[key[i] + 1]++;
  }
  for (int g = 1; g <= nkey; ++g) cnt[g] += cnt[g - 1];
  w.order.resize(n);
  for (int i = 0; i < n; ++i) w.order[cnt[key[i]]++] = i0 + i;
  // tiles of 16 MT rows, twice as wide for m = 0
  w.tile_prefix.resize((size_t)n + 3);
  int32_t *pre = w.tile_prefix.data();
  int q = 0;
  for (int part = 0; part < 3; ++part) {  // the tiled parts
    const int tr = 16 * (3 - part);
    pre[0] = 0;
    for (int k = 0; k < w.n_part[part]; ++k, ++q) {
      const int f = info[w.order[q] - i0], nm = f & 0xffff;
      const int ntr = (nm + tr - 1) / tr;
      pre[k + 1] = pre[k] + ntr * ((f >> 30) ? (nm + 2 * tr - 1) / (2 * tr)
                                             : ntr);
    }
    pre += w.n_part[part] + 1;
  }
  // solve tasks (item, class) by block order descending
  cnt.assign(nmmax + 2, 0);
  for (int i = 0; i < n; ++i) cnt[nmmax - (info[i] & 0xffff) + 1]++;
  for (int g = 1; g <= nmmax + 1; ++g) cnt[g] += cnt[g - 1];
  w.solve.resize(2 * (size_t)n);
  for (int i = 0; i < n; ++i) {
    const int q = cnt[nmmax - (info[i] & 0xffff)]++;
    w.solve[2 * q] = 2 * (i0 + i);
    w.solve[2 * q + 1] = 2 * (i0 + i) + 1;
  }
}

// Inversion and solve of the given tasks (item << 1 | class, block order
// descending; host list `tasks`, device copy `dtasks`) of the current work
// list (Q in ctx->d_q), in LDS-size buckets; ainv_out (may be NULL)
// receives every Q_c^-1.  parts: the assembly parts run_items launched
// (their ends recorded in ev_part, the side lanes forked); 0 when Q was
// written on the context's stream.  With `mixed` the buckets of orders
// kMixedMinOrder..kMixedMaxOrder take the mixed-precision solve (FP32
// elimination, FP64 refinement), whose per-task flags go to
// ctx->d_flags at the tasks' positions; *n_mixed counts those tasks.
static int solve_list(ctm_context *ctx, const std::vector<int32_t> &tasks,
                      const int32_t *dtasks, bool want_t, cdbl *ainv_out,
                      unsigned parts, bool mixed, size_t *n_mixed) {
  WorkList &w = ctx->work;
  if (n_mixed) *n_mixed = 0;
  // solve tasks in LDS-size buckets, largest first; blocks above the LDS
  // capacity are inverted in global scratch, at most kGlobalSolveBytes of it
  // per launch
  constexpr size_t kGlobalSolveBytes = (size_t)1 << 30;
  // (70: up to this order two blocks fit a CU's 160 KiB of LDS together;
  // 64: one row per lane suffices)
  const int bounds[] = {kMaxOrder, kMaxLdsOrder, 70, 64, 48, 32, 16};
  constexpr int kBuckets = 7;
  // buckets side by side on the lanes; the launches of the global bucket
  // share one scratch buffer and stay on lane 0 in order.  The blocks above
  // order 64 (the longest chains) and those of order 33-48 come from the
  // first assembly part (lane 0) alone, the others from the parts on lanes
  // 1-3: orders above 70 follow the first part on lane 0, 49-64 the second
  // part on lane 1, 65-70 and 33-48 the packed part on lane 3, up to 32 the
  // small parts on lane 2 (measured best of the placements tried)
  const int lane_of[kBuckets] = {0, 0, 3, 1, 3, 2, 2};
  if (parts == 0) CTM_HIP(fork_lanes(ctx));
  size_t pos = 0;
  bool after_big = false;
  for (int b = 0; b < kBuckets && pos < tasks.size(); ++b) {
    const int hi = bounds[b];
    const int lo = (b < kBuckets - 1) ? bounds[b + 1] : 0;
    size_t end = pos;
    int nm_max = 0;
    size_t cap = tasks.size();
    while (end < tasks.size() && end - pos < cap) {
      const Item &it = w.items[tasks[end] >> 1];
      const int nm = it.N + 1 - (it.m > 0 ? it.m : 1);
      if (nm <= lo) break;
      if (nm > hi) {
        ctx->err = "block order exceeds the largest supported order";
        return CTM_ERR_ARGUMENT;
      }
      if (nm_max == 0 && nm > kMaxLdsOrder)
        cap = std::max<size_t>(
            1, kGlobalSolveBytes / ((size_t)nm * nm * sizeof(cdbl)));
      nm_max = std::max(nm_max, nm);
      ++end;
    }
    if (nm_max > kMaxLdsOrder)
      CTM_HIP(ctx->d_gsolve.ensure((end - pos) * (size_t)nm_max * nm_max *
                                   sizeof(cdbl)));
    if (end > pos) {
      double fs = 0., bs = 0.;
      int64_t store_max = 0;
      int rows_max = 0;
      for (size_t q = pos; q < end; ++q) {
        const Item &it = w.items[tasks[q] >> 1];
        const int nmi = it.N + 1 - (it.m > 0 ? it.m : 1);
        const double nm = nmi;
        fs += (it.m > 0 ? 32. / 3. : 8. / 3.) * nm * nm * nm;
        bs += 16. * 2. * nm * nm + (want_t ? 16. * nm * nm : 0.);
        // m = 0: two parity sub-blocks of (nm + 1) / 2 and nm / 2 rows
        const int64_t r0 = it.m == 0 && nmi > 1 ? (nmi + 1) / 2 : nmi;
        const int64_t r1 = nmi - r0;
        // LDS sub-blocks have odd row strides (blk_lds_stride); the global
        // scratch of the blocks above kMaxLdsOrder is unpadded
        store_max = std::max(store_max,
                             nm_max > kMaxLdsOrder
                                 ? r0 * r0 + r1 * r1
                                 : r0 * blk_lds_stride((int)r0) +
                                       r1 * blk_lds_stride((int)r1));
        rows_max = std::max(rows_max, (int)r0);
      }
      const hipStream_t ls = lane(ctx, lane_of[b]);
      if (after_big) CTM_HIP(hipStreamWaitEvent(ls, ctx->ev_bucket, 0));
      if (parts != 0) {
        unsigned need = 0;
        for (size_t q = pos; q < end; ++q)
          need |= 1u << item_part(w.items[tasks[q] >> 1]);
        for (int p = 0; p < 5; ++p)
          if ((need & parts) >> p & 1u)
            CTM_HIP(hipStreamWaitEvent(ls, ctx->ev_part[p], 0));
      }
      const bool mix = mixed && want_t && !ainv_out &&
                       lo >= kMixedMinOrder - 1 && nm_max <= kMixedMaxOrder;
      if (mix) {
        CTM_LAUNCH_ON(ls, CTM_KERNEL_SOLVE, fs, bs,
                      launch_invert32(dtasks + pos, (int)(end - pos), nm_max,
                                      store_max, rows_max,
                                      ctx->d_items.as<Item>(),
                                      ctx->d_q.as<cdbl>(),
                                      ctx->d_x32.as<cflt>(),
                                      ctx->d_xexp.as<int32_t>(), ls));
        CTM_LAUNCH_ON(ls, CTM_KERNEL_REFINE, 0., 0.,
                      launch_refine(dtasks + pos, (int)(end - pos), nm_max,
                                    ctx->d_items.as<Item>(),
                                    ctx->d_q.as<cdbl>(),
                                    ctx->d_x32.as<cflt>(),
                                    ctx->d_xexp.as<int32_t>(),
                                    ctx->d_t.as<cdbl>(),
                                    ctx->d_rbuf.as<cdbl>(),
                                    ctx->d_qparts.as<double>(),
                                    ctx->d_flags.as<int32_t>() + pos, ls));
        if (n_mixed) *n_mixed += end - pos;
      } else {
        CTM_LAUNCH_ON(ls, CTM_KERNEL_SOLVE, fs, bs,
                      launch_invert_solve(dtasks + pos, (int)(end - pos),
                                          nm_max, store_max, rows_max,
                                          ctx->d_items.as<Item>(),
                                          ctx->d_q.as<cdbl>(),
                                          want_t ? ctx->d_t.as<cdbl>()
                                                 : nullptr,
                                          ctx->d_qparts.as<double>(),
                                          want_t ? 0 : 1,
                                          ctx->d_gsolve.as<cdbl>(), ainv_out,
                                          ls));
        if (mixed && want_t && ctx->d_flags.p)
          CTM_HIP(hipMemsetAsync(ctx->d_flags.as<int32_t>() + pos, 0,
                                 (end - pos) * sizeof(int32_t), ls));
      }
      if (b == 3 && want_t) {
        // all-m phase: the blocks of order 49..64 make the longest chains
        // that start last (after the 2 x 2 assembly part); the buckets of
        // smaller blocks, whose many small workgroups would take the CUs
        // first, follow them (measured 18.30 -> 18.23 ms per batch step;
        // no gain measured in the window phase)
        CTM_HIP(hipEventRecord(ctx->ev_bucket, ls));
        after_big = true;
      }
    }
    pos = end;
    // the global bucket may need several launches
    if (b == 0 && pos < tasks.size()) {
      const Item &it = w.items[tasks[pos] >> 1];
      if (it.N + 1 - (it.m > 0 ? it.m : 1) > kMaxLdsOrder) --b;
    }
  }
  CTM_HIP(join_lanes(ctx));
  return CTM_OK;
}

// Inversion and solve of the current work list's blocks; in mixed
// precision (ctx->mixed, T wanted) the tasks whose refinement did not
// converge are solved again in FP64.
static int run_solves(ctm_context *ctx, bool want_t, cdbl *ainv_out,
                      unsigned parts = 0) {
  WorkList &w = ctx->work;
  const bool mixed = ctx->mixed && want_t && !ainv_out;
  if (mixed) {
    const size_t nt = w.solve.size();
    CTM_HIP(ctx->d_x32.ensure(std::max<int64_t>(w.t_size, 1) * sizeof(cflt)));
    CTM_HIP(ctx->d_xexp.ensure(std::max<int64_t>(w.t_size, 1) *
                               sizeof(int32_t)));
    CTM_HIP(ctx->d_rbuf.ensure(std::max<int64_t>(w.t_size, 1) *
                               sizeof(cdbl)));
    CTM_HIP(ctx->d_flags.ensure(std::max<size_t>(nt, 1) * sizeof(int32_t)));
  }
  size_t n_mixed = 0;
  int rc = solve_list(ctx, w.solve, ctx->d_solve.as<int32_t>(), want_t,
                      ainv_out, parts, mixed, &n_mixed);
  if (rc != CTM_OK || !mixed) return rc;
  ctx->mixed_blocks += (int64_t)n_mixed;
  if (n_mixed == 0) return CTM_OK;
  // tasks whose refinement stalled: the FP64 solve, in the same order
  std::vector<int32_t> flags(w.solve.size());
  void *dst = flags.data();
  const void *src = ctx->d_flags.p;
  const size_t bytes = flags.size() * sizeof(int32_t);
  CTM_HIP(read_back(ctx, 1, &dst, &src, &bytes));
  std::vector<int32_t> &fb = ctx->fallback;
  fb.clear();
  for (size_t q = 0; q < w.solve.size(); ++q)
    if (flags[q]) fb.push_back(w.solve[q]);
  ctx->mixed_fallback += (int64_t)fb.size();
  if (fb.empty()) return CTM_OK;
  CTM_HIP(upload(ctx, ctx->d_fallback, fb));
  return solve_list(ctx, fb, ctx->d_fallback.as<int32_t>(), true, nullptr, 0,
                    false, nullptr);
}

// Device room for n_sets table sets holding table_size doubles: sized
// before the first tables launch of a list, which may come before the list
// is complete.
int size_tables(ctm_context *ctx, size_t n_sets, int64_t table_size) {
  CTM_HIP(ctx->d_sets.ensure(std::max<size_t>(n_sets, 1) * sizeof(TableSet)));
  CTM_HIP(ctx->d_tables.ensure(std::max<int64_t>(table_size, 1) *
                               sizeof(double)));
  return CTM_OK;
}

// Tables of the sets s0 .. s1 - 1 of the current work list (device room
// from size_tables).  Launched ahead of the item lists: the host builds
// those while the tables run.
int launch_set_tables(ctm_context *ctx, const GrainParams &gp, int s0,
                      int s1) {
  WorkList &w = ctx->work;
  if (s1 <= s0) return CTM_OK;
  int rc = flush_shared(ctx);  // Gauss-Legendre nodes and Wigner values
  if (rc != CTM_OK) return rc;
  CTM_HIP(stage_copy(ctx, ctx->d_sets.as<TableSet>() + s0, w.sets.data() + s0,
                     (size_t)(s1 - s0) * sizeof(TableSet)));
  // compulsory traffic: every table value written once
  double b_tab = 0.;
  for (int k = s0; k < s1; ++k)
    b_tab += 8. * ((double)kNodeScalars * w.sets[k].G +
                   (double)kBesselArrays * w.sets[k].N * w.sets[k].G);
  const int node0 = w.sets[s0].node0;
  const int node1 = s1 < (int)w.sets.size() ? w.sets[s1].node0 : w.n_nodes;
  CTM_LAUNCH(CTM_KERNEL_TABLES, 0., b_tab,
             launch_tables(ctx->d_sets.as<TableSet>() + s0, s1 - s0, node0,
                           node1 - node0, gp, ctx->d_gl.as<double>(),
                           ctx->d_tables.as<double>(), ctx->stream));
  return CTM_OK;
}

// Wigner tables of new keys, assembly and inversion/solve of all items of
// the current work list (their table sets launched before).
int run_items(ctm_context *ctx, const GrainParams &gp, bool want_t) {
  WorkList &w = ctx->work;
  const int ni = (int)w.items.size();
  if (ni == 0) return CTM_OK;
  build_tasks(w, 0, ni);
  int rc = flush_shared(ctx);  // Wigner values of new (N, m, G)
  if (rc != CTM_OK) return rc;
  CTM_HIP(upload(ctx, ctx->d_items, w.items));
  CTM_HIP(upload(ctx, ctx->d_tiles, w.order));
  CTM_HIP(upload(ctx, ctx->d_gioff_tmp, w.tile_prefix));
  CTM_HIP(upload(ctx, ctx->d_solve, w.solve));
  CTM_HIP(ctx->d_q.ensure(std::max<int64_t>(w.q_size, 1) * sizeof(cdbl)));
  if (want_t)
    CTM_HIP(ctx->d_t.ensure(std::max<int64_t>(w.t_size, 1) * sizeof(cdbl)));
  CTM_HIP(ctx->d_qparts.ensure((size_t)ni * 2 * kQParts * sizeof(double)));
  unsigned parts = 0;
  {
    // one launch per part (3 x 3, 2 x 2, 1 x 1 micro-tiles, packed m > 0,
    // packed m = 0), each credited with the algorithmic work of its blocks
    const int32_t *ord = ctx->d_tiles.as<int32_t>();
    const int32_t *pre = ctx->d_gioff_tmp.as<int32_t>();
    const int kid[3] = {CTM_KERNEL_ASSEMBLE_WIDE, CTM_KERNEL_ASSEMBLE,
                        CTM_KERNEL_ASSEMBLE_SMALL};
    // the parts write disjoint blocks.  The 3 x 3 part (most of the work,
    // enough tiles to fill the GPU) runs alone, so that its launches measure
    // the kernel and not the GPU it shares; the others follow side by side
    // on the lanes (all parts side by side measured 18.25 -> 17.97 ms per
    // batch step, with the 3 x 3 launches then sharing the GPU).  Window
    // lists hold only m = 0 blocks, so the packed m > 0 part is empty and
    // its lane takes the packed m = 0 part
    const int lane_of[5] = {0, 1, 2, 3, want_t ? 2 : 3};
    size_t po = 0;
    for (int part = 0; part < 3; ++part) {
      if (part == 1) CTM_HIP(fork_lanes(ctx));
      const int np = w.n_part[part];
      if (np > 0) {
        parts |= 1u << part;
        CTM_LAUNCH_ON(lane(ctx, lane_of[part]), kid[part], w.f_asm[part],
                      w.b_asm[part],
                      launch_assemble(3 - part, ord, pre + po, np,
                                      w.tile_prefix[po + np],
                                      ctx->d_items.as<Item>(),
                                      ctx->d_sets.as<TableSet>(),
                                      ctx->d_wigkeys.as<WigKey>(), gp,
                                      ctx->d_tables.as<double>(),
                                      ctx->d_wig.as<double>(),
                                      ctx->d_q.as<cdbl>(),
                                      lane(ctx, lane_of[part])));
        CTM_HIP(hipEventRecord(ctx->ev_part[part], lane(ctx, lane_of[part])));
      }
      ord += np;
      po += np + 1;
    }
    for (int part = 3; part < 5; ++part) {
      const int np = w.n_part[part];
      if (np > 0) {
        parts |= 1u << part;
        CTM_LAUNCH_ON(lane(ctx, lane_of[part]), CTM_KERNEL_ASSEMBLE_PACK,
                      w.f_asm[part], w.b_asm[part],
                      launch_assemble_pack4(ord, np, part == 4 ? 1 : 0,
                                            ctx->d_items.as<Item>(),
                                            ctx->d_sets.as<TableSet>(),
                                            ctx->d_wigkeys.as<WigKey>(), gp,
                                            ctx->d_tables.as<double>(),
                                            ctx->d_wig.as<double>(),
                                            ctx->d_q.as<cdbl>(),
                                            lane(ctx, lane_of[part])));
        CTM_HIP(hipEventRecord(ctx->ev_part[part], lane(ctx, lane_of[part])));
      }
      ord += np;
    }
  }
  // the solve buckets wait for their parts' events; run_solves joins the
  // lanes
  return run_solves(ctx, want_t, nullptr, parts);
}

// Runs tables + assembly + inversion/solve for the current work list.
int run_work(ctm_context *ctx, const GrainParams &gp, bool want_t) {
  WorkList &w = ctx->work;
  if (w.items.empty()) return CTM_OK;
  int rc = size_tables(ctx, w.sets.size(), w.table_size);
  if (rc == CTM_OK) rc = launch_set_tables(ctx, gp, 0, (int)w.sets.size());
  if (rc != CTM_OK) return rc;
  return run_items(ctx, gp, want_t);
}

}  // namespace

// ---------------------------------------------------------------------------

extern "C" {

int ctm_create(ctm_context **out, int device) {
  if (!out) return CTM_ERR_ARGUMENT;
  *out = nullptr;
  ctm_context *ctx = new ctm_context();
  ctx->device = device;
  hipError_t e = hipSetDevice(device);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
  for (int k = 0; e == hipSuccess && k < kLanes - 1; ++k) {
    e = hipStreamCreateWithFlags(&ctx->side[k], hipStreamNonBlocking);
    if (e == hipSuccess)
      e = hipEventCreateWithFlags(&ctx->ev_join[k], hipEventDisableTiming);
  }
  if (e == hipSuccess)
    e = hipEventCreateWithFlags(&ctx->ev_fork, hipEventDisableTiming);
  for (int k = 0; e == hipSuccess && k < 5; ++k)
    e = hipEventCreateWithFlags(&ctx->ev_part[k], hipEventDisableTiming);
  if (e == hipSuccess)
    e = hipEventCreateWithFlags(&ctx->ev_bucket, hipEventDisableTiming);
  if (e == hipSuccess) e = scratch_address(device, &ctx->scratch);
  if (e != hipSuccess) {
    (void)ctm_destroy(ctx);
    return CTM_ERR_HIP;
  }
  *out = ctx;
  return CTM_OK;
}

int ctm_destroy(ctm_context *ctx) {
  if (!ctx) return CTM_OK;
  if (ctx->plans > 0) {
    ctx->err = "destroy the context's plans first";
    return CTM_ERR_STATE;
  }
  (void)hipSetDevice(ctx->device);
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  for (hipStream_t st : ctx->side)
    if (st) (void)hipStreamSynchronize(st);
  DevBuf *bufs[] = {&ctx->d_gl,     &ctx->d_glkeys,  &ctx->d_wig,
                    &ctx->d_wigkeys, &ctx->d_tables, &ctx->d_sets,
                    &ctx->d_items,  &ctx->d_tiles,   &ctx->d_solve,
                    &ctx->d_q,      &ctx->d_t,       &ctx->d_qparts,
                    &ctx->d_state,  &ctx->d_active,  &ctx->d_winoff,
                    &ctx->d_winitems, &ctx->d_gioff, &ctx->d_gitems,
                    &ctx->d_gioff_tmp, &ctx->d_params, &ctx->d_ang,
                    &ctx->d_pt, &ctx->d_x, &ctx->d_spec, &ctx->d_cgkeys,
                    &ctx->d_cg, &ctx->d_cgidx, &ctx->d_avgtasks, &ctx->d_avg,
                    &ctx->d_aoff, &ctx->d_pn, &ctx->d_dist,
                    &ctx->d_obsitems, &ctx->d_query, &ctx->d_qout,
                    &ctx->d_gsolve, &ctx->d_ttables, &ctx->d_dq,
                    &ctx->d_ainv, &ctx->d_dt, &ctx->d_tparts, &ctx->d_dsum,
                    &ctx->d_ttiles, &ctx->d_tprefix, &ctx->d_wigfresh,
                    &ctx->d_x32, &ctx->d_xexp, &ctx->d_rbuf, &ctx->d_flags,
                    &ctx->d_fallback, &ctx->d_tpm, &ctx->d_pmscale,
                    &ctx->d_pmbuf};
  for (DevBuf *b : bufs) b->release();
  ctx->prof.release();
  ctx->staging.release();
  ctx->readback.release();
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  for (int k = 0; k < kLanes - 1; ++k) {
    if (ctx->side[k]) (void)hipStreamDestroy(ctx->side[k]);
    if (ctx->ev_join[k]) (void)hipEventDestroy(ctx->ev_join[k]);
  }
  if (ctx->ev_fork) (void)hipEventDestroy(ctx->ev_fork);
  if (ctx->ev_bucket) (void)hipEventDestroy(ctx->ev_bucket);
  for (hipEvent_t ev : ctx->ev_part)
    if (ev) (void)hipEventDestroy(ev);
  delete ctx;
  return CTM_OK;
}

const char *ctm_last_error(const ctm_context *ctx) {
  return ctx ? ctx->err.c_str() : "null context";
}

void *ctm_stream(ctm_context *ctx) { return ctx ? (void *)ctx->stream : nullptr; }

int ctm_synchronize(ctm_context *ctx) {
  if (!ctx) return CTM_ERR_ARGUMENT;
  CTM_HIP(drain(ctx));
  return CTM_OK;
}

static int check_inputs(ctm_context *ctx, int n, const double *a,
                        const double *b, const double *c, const double *d,
                        const double *e) {
  if (!ctx) return CTM_ERR_ARGUMENT;
  if (n < 0 || (n > 0 && (!a || !b || !c || !d || !e))) {
    ctx->err = "invalid grain arrays";
    return CTM_ERR_ARGUMENT;
  }
  return CTM_OK;
}

// Phase 2: all m blocks at the given per-grain (N, G).
static int run_all_m(ctm_context *ctx, const GrainParams &gp, int n,
                     const std::vector<uint32_t> &N,
                     const std::vector<uint32_t> &G, double *qsca_out,
                     double *qext_out) {
  WorkList &w = ctx->work;
  w.clear();
  SharedTables &sh = ctx->shared;
  ctx->grain_items_off.assign(n + 1, 0);
  // table set g is grain g's; the Wigner values and tables run while the
  // items are listed
  std::vector<int32_t> &wbase = w.scratch_a;
  wbase.resize(n);
  for (int g = 0; g < n; ++g) {
    w.add_set(g, (int)N[g], (int)G[g], sh.gl_keys[sh.gl((int)G[g])].off);
    wbase[g] = sh.wig_all((int)N[g], (int)G[g]);
  }
  int rc = size_tables(ctx, w.sets.size(), w.table_size);
  if (rc == CTM_OK) rc = launch_set_tables(ctx, gp, 0, n);
  if (rc != CTM_OK) return rc;
  size_t n_items = 0;
  for (int g = 0; g < n; ++g) n_items += N[g] + 1;
  w.items.reserve(n_items);
  for (int g = 0; g < n; ++g) {
    const int Ng = (int)N[g], Gg = (int)G[g];
    const int32_t *wk = sh.wig_slot.data() + wbase[g];
    ctx->grain_items_off[g] = (int)w.items.size();
    for (int m = 0; m <= Ng; ++m) w.add_item(g, Ng, m, Gg, g, wk[m], true);
  }
  ctx->grain_items_off[n] = (int)w.items.size();
  ctx->have_tangents = false;
  ctx->obs_source = -1;
  // keep the inputs for the observable calls on this batch
  CTM_HIP(ctx->d_params.ensure(5 * (size_t)n * sizeof(double)));
  {
    double *P = ctx->d_params.as<double>();
    const double *src[5] = {gp.radius, gp.axis_ratio, gp.wavelength, gp.mr_re,
                            gp.mr_im};
    for (int a = 0; a < 5; ++a)
      CTM_HIP(hipMemcpyAsync(P + (size_t)a * n, src[a], n * sizeof(double),
                             hipMemcpyDeviceToDevice, ctx->stream));
    ctx->gp_saved = GrainParams{P, P + n, P + 2 * (size_t)n, P + 3 * (size_t)n,
                                P + 4 * (size_t)n};
  }
  rc = run_items(ctx, gp, true);
  if (rc != CTM_OK) return rc;
  ctx->grain_items.resize(w.items.size());
  for (size_t i = 0; i < w.items.size(); ++i) ctx->grain_items[i] = (int32_t)i;
  CTM_HIP(upload(ctx, ctx->d_gioff, ctx->grain_items_off));
  CTM_HIP(upload(ctx, ctx->d_gitems, ctx->grain_items));
  CTM_LAUNCH(CTM_KERNEL_SUMS, 0., 0.,
             launch_full_sums(n, ctx->d_gioff.as<int32_t>(),
                              ctx->d_gitems.as<int32_t>(),
                              ctx->d_qparts.as<double>(), qsca_out, qext_out,
                              ctx->stream));
  ctx->final_items = w.items;
  ctx->averaged = false;
  return CTM_OK;
}

int ctm_tmatrix_batch(ctm_context *ctx, int n, const double *radius,
                      const double *axis_ratio, const double *wavelength,
                      const double *mr_re, const double *mr_im,
                      const ctm_convergence *conv, uint32_t *nmax_out,
                      int32_t *status_out, double *qsca_out,
                      double *qext_out) {
  int rc = check_inputs(ctx, n, radius, axis_ratio, wavelength, mr_re, mr_im);
  if (rc != CTM_OK) return rc;
  if (!conv || conv->min_order < 1 || conv->max_order <= conv->min_order + 1 ||
      conv->glfac < 1 || !(conv->tolerance > 0.)) {
    ctx->err = "invalid convergence parameters";
    return CTM_ERR_ARGUMENT;
  }
  if (conv->precision != CTM_PRECISION_FP64 &&
      conv->precision != CTM_PRECISION_MIXED) {
    ctx->err = "unknown precision";
    return CTM_ERR_ARGUMENT;
  }
  if ((int)conv->max_order - 1 > kMaxOrder) {
    ctx->err = "maximum order above 201 is not supported by this build";
    return CTM_ERR_ARGUMENT;
  }
  CTM_HIP(hipSetDevice(ctx->device));
  hipStream_t s = ctx->stream;
  // host work lists of the previous call may still feed pending copies
  CTM_HIP(drain(ctx));
  ctx->err.clear();
  ctx->shared.clear();
  ctx->n_grains = n;
  ctx->final_items.clear();
  ctx->mixed = conv->precision == CTM_PRECISION_MIXED;
  ctx->mixed_blocks = ctx->mixed_fallback = 0;
  if (n == 0) return CTM_OK;
  GrainParams gp{radius, axis_ratio, wavelength, mr_re, mr_im};
  const int min_order = (int)conv->min_order;
  const int last_order = (int)conv->max_order - 1;
  const int glfac = (int)conv->glfac;

  // host copies of the sizes and Re m (to size the windows)
  std::vector<double> h_rv(n), h_d(n), h_lam(n), h_mre(n);
  {
    void *dst[4] = {h_rv.data(), h_d.data(), h_lam.data(), h_mre.data()};
    const void *src[4] = {radius, axis_ratio, wavelength, mr_re};
    const size_t bytes[4] = {n * sizeof(double), n * sizeof(double),
                             n * sizeof(double), n * sizeof(double)};
    CTM_HIP(read_back(ctx, 4, dst, src, bytes));
  }

  std::vector<ConvState> st(n);
  for (int g = 0; g < n; ++g) {
    ConvState c{};
    c.old_s = c.old_e = 0.;
    c.dqs = c.dqe = -2.;
    c.first = 1;
    c.done = 0;
    c.nmax = min_order;
    c.item = -1;
    c.status = 1;
    st[g] = c;
  }
  CTM_HIP(upload(ctx, ctx->d_state, st));

  std::vector<int> lo(n, min_order), hi(n);
  for (int g = 0; g < n; ++g) {
    // A first window that covers most grains in one round while computing
    // few orders beyond convergence (each order costs ~n^3).  The few
    // grains left get a window of half their order: such a round is
    // latency-bound (few grains), so extra orders in it cost little, while
    // a further round costs a host synchronisation and a set of small
    // launches.
    const int end = first_window_end(h_rv[g], h_d[g], h_lam[g], h_mre[g]);
    hi[g] = std::min(last_order, std::max(min_order + 1, end));
  }
  std::vector<int32_t> active;
  for (int g = 0; g < n; ++g) active.push_back(g);

  // phase 1: windows of orders until every grain has converged
  std::vector<int32_t> wkey, win_off, win_items;
  while (!active.empty()) {
    WorkList &w = ctx->work;
    w.clear();
    // The window's table sets (one per grain and order) are listed grain by
    // grain and their tables launched every kSetsPerLaunch sets (the first
    // launch after kFirstSets), so the GPU starts while the host lists the
    // rest; the device room comes first, from the window's extent (sum over
    // N of 7 G + 8 N G, G = glfac N).
    constexpr size_t kSetsPerLaunch = 8192, kFirstSets = 2048;
    size_t n_sets = 0;
    int64_t t_size = 0;
    for (int g : active) {
      const int64_t a = lo[g] - 1, b = hi[g];
      if (b <= a) continue;
      const int64_t s1 = (b * (b + 1) - a * (a + 1)) / 2;
      const int64_t s2 =
          (b * (b + 1) * (2 * b + 1) - a * (a + 1) * (2 * a + 1)) / 6;
      n_sets += (size_t)(b - a);
      t_size += glfac * ((int64_t)kNodeScalars * s1 +
                         (int64_t)kBesselArrays * s2);
    }
    rc = size_tables(ctx, n_sets, t_size);
    if (rc != CTM_OK) return rc;
    {
      // every order's quadrature and m = 0 Wigner keys before the first
      // tables launch: one launch each computes all of them
      int nlo = last_order, nhi = 0;
      for (int g : active) {
        nlo = std::min(nlo, lo[g]);
        nhi = std::max(nhi, hi[g]);
      }
      for (int N = nlo; N <= nhi; ++N) ctx->shared.wig(N, 0, glfac * N);
      rc = flush_shared(ctx);
      if (rc != CTM_OK) return rc;
    }
    wkey.clear();
    win_off.assign(1, 0);
    size_t s_done = 0;
    for (int g : active) {
      for (int N = lo[g]; N <= hi[g]; ++N) {
        const int G = glfac * N;
        const int gk = ctx->shared.gl(G);
        w.add_set(g, N, G, ctx->shared.gl_keys[gk].off);
        wkey.push_back(ctx->shared.wig(N, 0, G));
      }
      win_off.push_back((int32_t)w.sets.size());
      if (w.sets.size() - s_done >= (s_done ? kSetsPerLaunch : kFirstSets)) {
        rc = launch_set_tables(ctx, gp, (int)s_done, (int)w.sets.size());
        if (rc != CTM_OK) return rc;
        s_done = w.sets.size();
      }
    }
    rc = launch_set_tables(ctx, gp, (int)s_done, (int)w.sets.size());
    if (rc != CTM_OK) return rc;
    // item i: the m = 0 block of set i
    win_items.resize(w.sets.size());
    for (size_t i = 0; i < w.sets.size(); ++i) {
      const TableSet &ts = w.sets[i];
      win_items[i] = w.add_item(ts.grain, ts.N, 0, ts.G, (int)i, wkey[i],
                                false);
    }
    rc = run_items(ctx, gp, false);
    if (rc != CTM_OK) return rc;
    CTM_HIP(upload(ctx, ctx->d_active, active));
    CTM_HIP(upload(ctx, ctx->d_winoff, win_off));
    CTM_HIP(upload(ctx, ctx->d_winitems, win_items));
    CTM_LAUNCH(CTM_KERNEL_CONVERGE, 0., 0.,
               launch_converge(ctx->d_active.as<int32_t>(), (int)active.size(),
                               ctx->d_winoff.as<int32_t>(),
                               ctx->d_winitems.as<int32_t>(),
                               ctx->d_items.as<Item>(),
                               ctx->d_qparts.as<double>(), conv->tolerance,
                               last_order, ctx->d_state.as<ConvState>(), s));
    {
      void *dst = st.data();
      const void *src = ctx->d_state.p;
      const size_t bytes = n * sizeof(ConvState);
      CTM_HIP(read_back(ctx, 1, &dst, &src, &bytes));
    }
    std::vector<int32_t> next;
    for (int g : active) {
      if (st[g].done) continue;
      lo[g] = hi[g] + 1;
      hi[g] = next_window_end(hi[g], last_order, false);
      next.push_back(g);
    }
    active.swap(next);
  }

  ctx->nmax.assign(n, 0);
  ctx->ngauss.assign(n, 0);
  ctx->status.assign(n, 0);
  for (int g = 0; g < n; ++g) {
    ctx->nmax[g] = (uint32_t)st[g].nmax;
    ctx->ngauss[g] = (uint32_t)(glfac * st[g].nmax);
    ctx->status[g] = st[g].status;
  }
  rc = run_all_m(ctx, gp, n, ctx->nmax, ctx->ngauss, qsca_out, qext_out);
  if (rc != CTM_OK) return rc;
  CTM_HIP(launch_write_orders(n, ctx->d_state.as<ConvState>(), nmax_out,
                              status_out, s));
  return CTM_OK;
}

int ctm_tmatrix_fixed(ctm_context *ctx, int n, const double *radius,
                      const double *axis_ratio, const double *wavelength,
                      const double *mr_re, const double *mr_im,
                      const uint32_t *nmax_host, const uint32_t *ngauss_host,
                      uint32_t precision, double *qsca_out, double *qext_out) {
  int rc = check_inputs(ctx, n, radius, axis_ratio, wavelength, mr_re, mr_im);
  if (rc != CTM_OK) return rc;
  if (n > 0 && (!nmax_host || !ngauss_host)) {
    ctx->err = "missing forced orders";
    return CTM_ERR_ARGUMENT;
  }
  if (precision != CTM_PRECISION_FP64 && precision != CTM_PRECISION_MIXED) {
    ctx->err = "unknown precision";
    return CTM_ERR_ARGUMENT;
  }
  for (int g = 0; g < n; ++g) {
    if (nmax_host[g] < 1 || (int)nmax_host[g] > kMaxOrder ||
        ngauss_host[g] < 1) {
      ctx->err = "forced order out of range (1..200) or ngauss < 1";
      return CTM_ERR_ARGUMENT;
    }
  }
  CTM_HIP(hipSetDevice(ctx->device));
  CTM_HIP(drain(ctx));
  ctx->err.clear();
  ctx->shared.clear();
  ctx->n_grains = n;
  ctx->final_items.clear();
  ctx->nmax.assign(nmax_host, nmax_host + n);
  ctx->ngauss.assign(ngauss_host, ngauss_host + n);
  ctx->status.assign(n, 0);
  ctx->mixed = precision == CTM_PRECISION_MIXED;
  ctx->mixed_blocks = ctx->mixed_fallback = 0;
  if (n == 0) return CTM_OK;
  GrainParams gp{radius, axis_ratio, wavelength, mr_re, mr_im};
  return run_all_m(ctx, gp, n, ctx->nmax, ctx->ngauss, qsca_out, qext_out);
}

size_t ctm_tmatrix_size(ctm_context *ctx, int g) {
  if (!ctx || g < 0 || g >= ctx->n_grains || ctx->final_items.empty()) return 0;
  const size_t N = ctx->nmax[g];
  size_t total = 0;
  for (size_t m = 0; m <= N; ++m) {
    const size_t nm = (m == 0) ? N : N + 1 - m;
    total += 4 * nm * nm;
  }
  return total;
}

// Grain g's blocks of a class-layout buffer (T or a tangent) in the
// reference block layout.
static int get_blocks(ctm_context *ctx, int g, const cdbl *base,
                      double *T_host, size_t capacity) {
  const size_t need = ctm_tmatrix_size(ctx, g);
  if (need == 0 || !T_host || capacity < need) {
    ctx->err = "invalid grain index or output capacity";
    return CTM_ERR_ARGUMENT;
  }
  CTM_HIP(hipSetDevice(ctx->device));
  const int N = (int)ctx->nmax[g];
  size_t out_off = 0;
  std::vector<cdbl> blk;
  cdbl *T = reinterpret_cast<cdbl *>(T_host);
  for (int m = 0; m <= N; ++m) {
    const Item &it = ctx->final_items[ctx->grain_items_off[g] + m];
    const int nmin = m > 0 ? m : 1;
    const int nm = N + 1 - nmin;
    const size_t nm2 = (size_t)nm * nm;
    blk.resize(2 * nm2);
    CTM_HIP(hipMemcpyAsync(blk.data(), base + it.t_off,
                           2 * nm2 * sizeof(cdbl), hipMemcpyDeviceToHost,
                           ctx->stream));
    CTM_HIP(drain(ctx));
    const int n2m = 2 * nm;
    cdbl *out = T + out_off;
    for (size_t e = 0; e < (size_t)n2m * n2m; ++e) out[e] = cmk(0., 0.);
    for (int c = 0; c < 2; ++c) {
      for (int p1 = 0; p1 < nm; ++p1) {
        const int h1 = (c + nmin + p1) & 1;
        for (int p2 = 0; p2 < nm; ++p2) {
          const int h2 = (c + nmin + p2) & 1;
          out[(size_t)(h1 * nm + p1) * n2m + h2 * nm + p2] =
              blk[c * nm2 + (size_t)p1 * nm + p2];
        }
      }
    }
    out_off += (size_t)n2m * n2m;
  }
  return CTM_OK;
}

int ctm_tmatrix_get(ctm_context *ctx, int g, double *T_host, size_t capacity) {
  if (!ctx) return CTM_ERR_ARGUMENT;
  return get_blocks(ctx, g, ctx->d_t.as<cdbl>(), T_host, capacity);
}

int ctm_tmatrix_tangents(ctm_context *ctx, double *dqsca_host,
                         double *dqext_host) {
  if (!ctx) return CTM_ERR_ARGUMENT;
  if (ctx->final_items.empty()) {
    ctx->err = "no T-matrices in the context";
    return CTM_ERR_STATE;
  }
  CTM_HIP(hipSetDevice(ctx->device));
  hipStream_t s = ctx->stream;
  WorkList &w = ctx->work;  // the batch's all-m work list
  const int n = ctx->n_grains;
  const int ni = (int)w.items.size();
  // tangent tables
  CTM_HIP(ctx->d_ttables.ensure(std::max<int64_t>(w.table_size, 1) * 2 *
                                kTangentDirs * sizeof(double)));
  CTM_LAUNCH(CTM_KERNEL_TABLES, 0., 0.,
             launch_tables_tangent(ctx->d_sets.as<TableSet>(),
                                   (int)w.sets.size(), ctx->gp_saved,
                                   ctx->d_gl.as<double>(),
                                   ctx->d_tables.as<double>(), w.table_size,
                                   ctx->d_ttables.as<double>(), s));
  // tangent assembly over 16 x kTangentTileCols pair tiles of every item
  // (twice as wide for m = 0)
  std::vector<int32_t> order(ni), prefix(ni + 1, 0);
  for (int i = 0; i < ni; ++i) {
    order[i] = i;
    const Item &it = w.items[i];
    const int nm = it.N + 1 - (it.m > 0 ? it.m : 1);
    const int ntr = (nm + 15) / 16;
    const int cw = it.m == 0 ? 2 * kTangentTileCols : kTangentTileCols;
    const int ntc = (nm + cw - 1) / cw;
    prefix[i + 1] = prefix[i] + ntr * ntc;
  }
  CTM_HIP(upload(ctx, ctx->d_ttiles, order));
  CTM_HIP(upload(ctx, ctx->d_tprefix, prefix));
  CTM_HIP(ctx->d_dq.ensure(std::max<int64_t>(w.q_size, 1) * kTangentDirs *
                           sizeof(cdbl)));
  CTM_LAUNCH(CTM_KERNEL_ASSEMBLE, 0., 0.,
             launch_assemble_tangent(
                 ctx->d_ttiles.as<int32_t>(), ctx->d_tprefix.as<int32_t>(), ni,
                 prefix[ni], ctx->d_items.as<Item>(),
                 ctx->d_sets.as<TableSet>(), ctx->d_wigkeys.as<WigKey>(),
                 ctx->gp_saved, ctx->d_tables.as<double>(),
                 ctx->d_ttables.as<double>(), w.table_size,
                 ctx->d_wig.as<double>(), ctx->d_dq.as<cdbl>(), w.q_size, s));
  // Q_c^-1 of every block: the elimination is repeated with only the
  // diagonal of the product formed (T is left as it is; the per-block sums
  // it overwrites were consumed by the batch's Q_sca/Q_ext)
  CTM_HIP(ctx->d_ainv.ensure(std::max<int64_t>(w.t_size, 1) * sizeof(cdbl)));
  int rc = run_solves(ctx, false, ctx->d_ainv.as<cdbl>());
  if (rc != CTM_OK) return rc;
  CTM_HIP(ctx->d_dt.ensure(std::max<int64_t>(w.t_size, 1) * kTangentDirs *
                           sizeof(cdbl)));
  CTM_HIP(ctx->d_tparts.ensure((size_t)std::max(ni, 1) * 2 * kTangentDirs * 2 *
                               sizeof(double)));
  CTM_LAUNCH(CTM_KERNEL_SOLVE, 0., 0.,
             launch_tangent_solve(2 * ni, ctx->d_items.as<Item>(),
                                  ctx->d_t.as<cdbl>(), ctx->d_ainv.as<cdbl>(),
                                  ctx->d_dq.as<cdbl>(), w.q_size,
                                  ctx->d_dt.as<cdbl>(), w.t_size,
                                  ctx->d_tparts.as<double>(), s));
  CTM_HIP(ctx->d_dsum.ensure((size_t)std::max(n, 1) * 2 * kTangentDirs *
                             sizeof(double)));
  double *dsum = ctx->d_dsum.as<double>();
  CTM_LAUNCH(CTM_KERNEL_SUMS, 0., 0.,
             launch_tangent_sums(n, ctx->d_gioff.as<int32_t>(),
                                 ctx->d_tparts.as<double>(), dsum,
                                 dsum + (size_t)n * kTangentDirs, s));
  ctx->dqsca.resize((size_t)n * kTangentDirs);
  ctx->dqext.resize((size_t)n * kTangentDirs);
  CTM_HIP(hipMemcpyAsync(ctx->dqsca.data(), dsum,
                         ctx->dqsca.size() * sizeof(double),
                         hipMemcpyDeviceToHost, s));
  CTM_HIP(hipMemcpyAsync(ctx->dqext.data(), dsum + (size_t)n * kTangentDirs,
                         ctx->dqext.size() * sizeof(double),
                         hipMemcpyDeviceToHost, s));
  CTM_HIP(drain(ctx));
  if (dqsca_host)
    std::copy(ctx->dqsca.begin(), ctx->dqsca.end(), dqsca_host);
  if (dqext_host)
    std::copy(ctx->dqext.begin(), ctx->dqext.end(), dqext_host);
  ctx->have_tangents = true;
  return CTM_OK;
}

int ctm_tangent_get(ctm_context *ctx, int g, int dir, double *dT_host,
                    size_t capacity) {
  if (!ctx) return CTM_ERR_ARGUMENT;
  if (!ctx->have_tangents || dir < 0 || dir >= kTangentDirs) {
    ctx->err = "no tangents computed, or direction out of range";
    return CTM_ERR_STATE;
  }
  return get_blocks(ctx, g,
                    ctx->d_dt.as<cdbl>() + (int64_t)dir * ctx->work.t_size,
                    dT_host, capacity);
}

int ctm_observables_select(ctm_context *ctx, int which) {
  if (!ctx) return CTM_ERR_ARGUMENT;
  if (which < -1 || which >= kTangentDirs ||
      (which >= 0 && !ctx->have_tangents)) {
    ctx->err = "no such tangent (call ctm_tmatrix_tangents first)";
    return CTM_ERR_ARGUMENT;
  }
  ctx->obs_source = which;
  return CTM_OK;
}

static int observables_impl(ctm_context *ctx, const ctm_observable_spec *spec,
                            double *qext, double *qextpol, double *qextcpol,
                            double *qabs, double *qabspol, double *Z,
                            char *fixed, size_t fixed_cap);

int ctm_observables_derivative(ctm_context *ctx, int dir,
                               const ctm_observable_spec *spec, double *dqext,
                               double *dqextpol, double *dqextcpol,
                               double *dqabs, double *dqabspol, double *dZ) {
  if (!ctx || !spec) return CTM_ERR_ARGUMENT;
  if (!ctx->have_tangents || dir < 0 || dir >= kTangentDirs) {
    ctx->err = "no tangents computed, or direction out of range";
    return CTM_ERR_STATE;
  }
  if (ctx->averaged) {
    ctx->err = "the context holds ensemble T-matrices: derivatives apply "
               "to the batch's own T-matrices";
    return CTM_ERR_STATE;
  }
  if (ctx->obs_source != -1) {
    ctx->err = "a tangent is selected for the observables";
    return CTM_ERR_STATE;
  }
  const int n = ctx->n_grains;
  const int nth = spec->n_theta, nst = spec->n_scatter_theta;
  if (nth < 0 || nst < 0 || (nth > 0 && !spec->theta) ||
      (spec->account_for_scattering && spec->ngauss < 1)) {
    ctx->err = "invalid observable specification";
    return CTM_ERR_ARGUMENT;
  }
  CTM_HIP(hipSetDevice(ctx->device));
  hipStream_t s = ctx->stream;
  const WorkList &w = ctx->work;
  const int64_t qn = (int64_t)n * nth;
  const int64_t zw = (int64_t)nst * 2 * nst * 16, zn = (int64_t)n * zw;
  const int64_t block = 5 * qn + zn;
  CTM_HIP(ctx->d_tpm.ensure(std::max<int64_t>(w.t_size, 1) * sizeof(cdbl)));
  CTM_HIP(ctx->d_pmscale.ensure(std::max(n, 1) * sizeof(double)));
  CTM_HIP(ctx->d_pmbuf.ensure(std::max<int64_t>(2 * block, 1) *
                              sizeof(double)));
  const cdbl *dT = ctx->d_dt.as<cdbl>() + (int64_t)dir * w.t_size;
  double *scale = ctx->d_pmscale.as<double>();
  CTM_HIP(launch_pm_scale(n, ctx->d_gioff.as<int32_t>(),
                          ctx->d_items.as<Item>(), ctx->d_t.as<cdbl>(), dT,
                          scale, s));
  double *buf = ctx->d_pmbuf.as<double>();
  const int nitems = (int)ctx->final_items.size();
  for (int k = 0; k < 2; ++k) {
    double *b = buf + k * block;
    CTM_HIP(launch_pm_form(nitems, ctx->d_items.as<Item>(),
                           ctx->d_t.as<cdbl>(), dT, scale, k ? -1. : 1.,
                           ctx->d_tpm.as<cdbl>(), s));
    ctx->obs_source = kObsSourcePM;
    const int rc = observables_impl(ctx, spec, b, b + qn, b + 2 * qn,
                                    b + 3 * qn, b + 4 * qn,
                                    nst > 0 ? b + 5 * qn : nullptr, nullptr,
                                    0);
    ctx->obs_source = -1;
    if (rc != CTM_OK) return rc;
  }
  double *outs[5] = {dqext, dqextpol, dqextcpol, dqabs, dqabspol};
  for (int q = 0; q < 5; ++q) {
    const bool on = q < 3 ? spec->do_extinction : spec->do_absorption;
    if (on && outs[q] && nth > 0)
      CTM_HIP(launch_pm_combine(qn, nth, buf + q * qn, buf + block + q * qn,
                                scale, outs[q], s));
  }
  if (nst > 0 && dZ)
    CTM_HIP(launch_pm_combine(zn, zw, buf + 5 * qn, buf + block + 5 * qn,
                              scale, dZ, s));
  CTM_HIP(drain(ctx));
  return CTM_OK;
}

int ctm_vjp(ctm_context *ctx, const double *cot_qsca, const double *cot_qext,
            double *grad) {
  if (!ctx || !grad) return CTM_ERR_ARGUMENT;
  if (!ctx->have_tangents) {
    ctx->err = "no tangents computed (call ctm_tmatrix_tangents first)";
    return CTM_ERR_STATE;
  }
  const int n = ctx->n_grains;
  for (int g = 0; g < n; ++g)
    for (int p = 0; p < kTangentDirs; ++p) {
      const size_t i = (size_t)g * kTangentDirs + p;
      grad[i] = (cot_qsca ? cot_qsca[g] * ctx->dqsca[i] : 0.) +
                (cot_qext ? cot_qext[g] * ctx->dqext[i] : 0.);
    }
  return CTM_OK;
}

// Gauss-Legendre rule on [-1, 1] (SpecialFunctions.hpp:647-715), host side,
// for the observable direction grids.
static void push_angle(std::vector<double> &ang, double c) {
  const double sn = std::sqrt((1. - c) * (1. + c));
  ang.push_back(c);
  ang.push_back(sn);
  ang.push_back(sn != 0. ? 1. / sn : 9000.);
}

// Launch `call` on stream `strm`, between profiler events unless `quiet`
// (the no-allocation path takes no events from the profiler's pool).
#define CTM_LAUNCH_Q(quiet, strm, kernel, flops, bytes, call)    \
  do {                                                          \
    if (quiet) {                                                \
      CTM_HIP(call);                                            \
    } else {                                                    \
      CTM_LAUNCH_ON(strm, kernel, flops, bytes, call);          \
    }                                                           \
  } while (0)

// The observables of the context's T-matrices.  Their scratch (direction
// tables, the pi/tau table, the bilinear forms X, the item order) comes
// from the context's growable buffers, or, when `fixed` is not null, from
// the fixed device range fixed[0 .. fixed_cap) with no allocation at all.
// Direction table, incoming/outgoing lists and direction pairs of an
// observable specification (validated by the caller), and the device image
// of its arrays (blob; bind() points the spec at a device copy).
struct ObsLayout {
  ObsSpec sp{};
  std::vector<char> blob;
  int n_ang = 0;
  size_t o_ang = 0, o_in = 0, o_out = 0, o_pin = 0, o_pout = 0, o_byin = 0,
         o_inoff = 0, o_tw = 0, o_pc = 0, o_ps = 0, o_pw = 0, o_sc = 0,
         o_ss = 0;
  void bind(const char *base) {
    sp.in_angle = reinterpret_cast<const int32_t *>(base + o_in);
    sp.out_angle = reinterpret_cast<const int32_t *>(base + o_out);
    sp.pair_in = reinterpret_cast<const int32_t *>(base + o_pin);
    sp.pair_out = reinterpret_cast<const int32_t *>(base + o_pout);
    sp.pair_by_in = reinterpret_cast<const int32_t *>(base + o_byin);
    sp.in_pair_off = reinterpret_cast<const int32_t *>(base + o_inoff);
    sp.theta_w = reinterpret_cast<const double *>(base + o_tw);
    sp.phi_cos = reinterpret_cast<const double *>(base + o_pc);
    sp.phi_sin = reinterpret_cast<const double *>(base + o_ps);
    sp.phi_w = reinterpret_cast<const double *>(base + o_pw);
    sp.sphi_cos = reinterpret_cast<const double *>(base + o_sc);
    sp.sphi_sin = reinterpret_cast<const double *>(base + o_ss);
  }
};

static void obs_layout(const ctm_observable_spec *spec, ObsLayout &o) {
  const int nth = spec->n_theta;
  const int ng = spec->account_for_scattering ? spec->ngauss : 0;
  const int nst = spec->n_scatter_theta;
  // direction table: outgoing theta list, GL incoming, scattering grid,
  // and the scattering incidence theta = pi/2
  std::vector<double> ang;
  for (int j = 0; j < nth; ++j) push_angle(ang, std::cos(spec->theta[j]));
  std::vector<double> gx(std::max(ng, 1)), gw(std::max(ng, 1));
  std::vector<double> px(std::max(ng, 1)), pw(std::max(ng, 1));
  std::vector<double> pc(std::max(ng, 1)), ps(std::max(ng, 1));
  if (ng > 0) {
    gauss_legendre_host(ng, gx.data(), gw.data());
    gauss_legendre_host(ng, px.data(), pw.data());
    for (int i = 0; i < ng; ++i) {  // rule on [0, 2 pi]
      px[i] = M_PI * px[i] + M_PI;
      pw[i] *= M_PI;
      pc[i] = std::cos(px[i]);
      ps[i] = std::sin(px[i]);
    }
    for (int i = 0; i < ng; ++i) push_angle(ang, gx[i]);
  }
  const int nsp = 2 * nst;
  std::vector<double> sc(std::max(nsp, 1)), ss(std::max(nsp, 1));
  if (nst > 0) {
    std::vector<double> tx(nst), tw(nst), fx(nsp), fw(nsp);
    if (spec->scatter_gl) {
      gauss_legendre_host(nst, tx.data(), tw.data());
      gauss_legendre_host(nsp, fx.data(), fw.data());
      for (int i = 0; i < nsp; ++i) fx[i] = M_PI * fx[i] + M_PI;
    } else {
      for (int i = 0; i < nst; ++i) tx[i] = std::cos(i * (M_PI / nst));
      for (int i = 0; i < nsp; ++i) fx[i] = i * (2. * M_PI / nsp);
    }
    for (int i = 0; i < nst; ++i) push_angle(ang, tx[i]);
    for (int i = 0; i < nsp; ++i) {
      sc[i] = std::cos(fx[i]);
      ss[i] = std::sin(fx[i]);
    }
    push_angle(ang, std::cos(M_PI_2));
  }
  const int n_ang = (int)ang.size() / 3;
  o.n_ang = n_ang;
  if (n_ang == 0) return;
  // incoming / outgoing lists and pairs
  std::vector<int32_t> in_l, out_l, p_in, p_out;
  for (int j = 0; j < nth; ++j) {
    in_l.push_back(j);
    out_l.push_back(j);
  }
  for (int i = 0; i < ng; ++i) in_l.push_back(nth + i);
  if (nst > 0) {
    in_l.push_back(nth + ng + nst);
    for (int i = 0; i < nst; ++i) out_l.push_back(nth + ng + i);
  }
  ObsSpec &sp = o.sp;
  sp = ObsSpec{};
  sp.fwd_pair0 = 0;
  for (int j = 0; j < nth; ++j) {
    p_out.push_back(j);
    p_in.push_back(j);
  }
  sp.abs_pair0 = (int)p_in.size();
  for (int j = 0; j < nth; ++j)
    for (int i = 0; i < ng; ++i) {
      p_out.push_back(j);
      p_in.push_back(nth + i);
    }
  sp.sca_pair0 = (int)p_in.size();
  for (int i = 0; i < nst; ++i) {
    p_out.push_back(nth + i);
    p_in.push_back((int)in_l.size() - 1);
  }
  // pairs grouped by incoming direction (counting sort, stable)
  std::vector<int32_t> in_off(in_l.size() + 1, 0), by_in(p_in.size());
  for (int32_t v : p_in) in_off[v + 1]++;
  for (size_t i = 0; i < in_l.size(); ++i) in_off[i + 1] += in_off[i];
  {
    std::vector<int32_t> fill(in_off.begin(), in_off.end() - 1);
    for (size_t q = 0; q < p_in.size(); ++q) by_in[fill[p_in[q]]++] = (int32_t)q;
  }
  sp.n_ang = n_ang;
  sp.n_in = (int)in_l.size();
  sp.n_out = (int)out_l.size();
  sp.n_pairs = (int)p_in.size();
  sp.n_theta = nth;
  sp.ngauss = ng;
  sp.n_stheta = nst;
  sp.n_sphi = nsp;
  // device copies of the spec arrays (one buffer)
  std::vector<char> &blob = o.blob;
  blob.clear();
  auto put = [&](const void *src, size_t bytes) {
    size_t off = (blob.size() + 255) & ~(size_t)255;
    blob.resize(off + bytes);
    if (bytes) std::memcpy(blob.data() + off, src, bytes);
    return off;
  };
  o.o_ang = put(ang.data(), ang.size() * sizeof(double));
  o.o_in = put(in_l.data(), in_l.size() * 4);
  o.o_out = put(out_l.data(), out_l.size() * 4);
  o.o_pin = put(p_in.data(), p_in.size() * 4);
  o.o_pout = put(p_out.data(), p_out.size() * 4);
  o.o_byin = put(by_in.data(), by_in.size() * 4);
  o.o_inoff = put(in_off.data(), in_off.size() * 4);
  o.o_tw = put(gw.data(), gw.size() * 8);
  o.o_pc = put(pc.data(), pc.size() * 8);
  o.o_ps = put(ps.data(), ps.size() * 8);
  o.o_pw = put(pw.data(), pw.size() * 8);
  o.o_sc = put(sc.data(), sc.size() * 8);
  o.o_ss = put(ss.data(), ss.size() * 8);
}

static int observables_impl(ctm_context *ctx, const ctm_observable_spec *spec,
                            double *qext, double *qextpol, double *qextcpol,
                            double *qabs, double *qabspol, double *Z,
                            char *fixed, size_t fixed_cap) {
  if (!ctx || !spec) return CTM_ERR_ARGUMENT;
  const bool quiet = fixed != nullptr;
  if (ctx->final_items.empty()) {
    ctx->err = "no T-matrices in the context";
    return CTM_ERR_STATE;
  }
  const int nth = spec->n_theta;
  const int ng = spec->account_for_scattering ? spec->ngauss : 0;
  const int nst = spec->n_scatter_theta;
  if (nth < 0 || ng < 0 || nst < 0 || (nth > 0 && !spec->theta) ||
      (spec->account_for_scattering && spec->ngauss < 1)) {
    ctx->err = "invalid observable specification";
    return CTM_ERR_ARGUMENT;
  }
  if (ctx->obs_source >= 0 &&
      (spec->account_for_scattering || nst > 0 || ctx->averaged)) {
    ctx->err = "with a tangent selected only observables linear in T "
               "(no scattered-light correction, no Z, no averaging)";
    return CTM_ERR_ARGUMENT;
  }
  CTM_HIP(hipSetDevice(ctx->device));
  hipStream_t s = ctx->stream;
  const int n = ctx->n_grains;
  ObsLayout lay;
  obs_layout(spec, lay);
  const int n_ang = lay.n_ang;
  if (n_ang == 0) return CTM_OK;
  ObsSpec &sp = lay.sp;
  const std::vector<char> &blob = lay.blob;
  int nmax_all = 1, nm_max = 1;
  for (int g = 0; g < n; ++g) nmax_all = std::max(nmax_all, (int)ctx->nmax[g]);
  nm_max = nmax_all;
  const size_t nitems = ctx->final_items.size();
  if (obs_x_lds_bytes(nmax_all, sp.n_out, 8, 256, 1, true) > kObsXMaxLds) {
    ctx->err = "too many outgoing directions for the largest order";
    return CTM_ERR_ARGUMENT;
  }
  const size_t pt_bytes =
      (size_t)(nmax_all + 1) * n_ang * nmax_all * 2 * sizeof(double);
  const size_t x_bytes = nitems * (size_t)sp.n_pairs * 4 * sizeof(cdbl);
  const size_t li_bytes = std::max<size_t>(nitems, 1) * sizeof(int32_t);
  char *base, *d_pt, *d_x, *d_li;
  if (fixed) {
    // one bump pass over the fixed range, 256-byte aligned parts
    size_t off = 0;
    auto take = [&](size_t bytes) {
      const size_t o = off;
      off += (bytes + 255) & ~(size_t)255;
      return o;
    };
    const size_t f_spec = take(blob.size()), f_pt = take(pt_bytes),
                 f_x = take(x_bytes), f_li = take(li_bytes);
    if (off > fixed_cap) {
      ctx->err = "ctm_observables_noalloc: these observables need " +
                 std::to_string(off) + " bytes of scratch, more than its " +
                 std::to_string(fixed_cap) + " (fewer grains, angles or "
                 "quadrature directions per call, or ctm_observables)";
      return CTM_ERR_ARGUMENT;
    }
    base = fixed + f_spec;
    d_pt = fixed + f_pt;
    d_x = fixed + f_x;
    d_li = fixed + f_li;
    // a pageable source: the runtime stages it; the call drains its stream
    // before `blob` goes out of scope
    CTM_HIP(hipMemcpyAsync(base, blob.data(), blob.size(),
                           hipMemcpyHostToDevice, s));
  } else {
    CTM_HIP(ctx->d_spec.ensure(blob.size()));
    CTM_HIP(stage_copy(ctx, ctx->d_spec.p, blob.data(), blob.size()));
    CTM_HIP(ctx->d_pt.ensure(pt_bytes));
    CTM_HIP(ctx->d_x.ensure(x_bytes));
    CTM_HIP(ctx->d_obsitems.ensure(li_bytes));
    base = ctx->d_spec.as<char>();
    d_pt = ctx->d_pt.as<char>();
    d_x = ctx->d_x.as<char>();
    d_li = ctx->d_obsitems.as<char>();
  }
  const double *d_ang = reinterpret_cast<const double *>(base + lay.o_ang);
  lay.bind(base);

  (void)nm_max;
  double *PT = reinterpret_cast<double *>(d_pt);
  cdbl *X = reinterpret_cast<cdbl *>(d_x);
  int32_t *LI = reinterpret_cast<int32_t *>(d_li);
  CTM_LAUNCH_Q(quiet, s, CTM_KERNEL_OBS_X, 0., 0.,
               launch_pitau(d_ang, n_ang, nmax_all, PT, s));
  // items by block order (largest first), launched in LDS-size buckets
  {
    std::vector<int32_t> &lst = ctx->obs_items;
    lst.resize(nitems);
    std::vector<int32_t> cnt(nmax_all + 2, 0);
    auto nm_of = [&](const Item &it) { return it.N + 1 - (it.m > 0 ? it.m : 1); };
    for (const Item &it : ctx->final_items) cnt[nmax_all - nm_of(it) + 1]++;
    for (int q = 1; q <= nmax_all + 1; ++q) cnt[q] += cnt[q - 1];
    for (size_t i = 0; i < nitems; ++i)
      lst[cnt[nmax_all - nm_of(ctx->final_items[i])]++] = (int32_t)i;
    if (fixed)
      CTM_HIP(hipMemcpyAsync(LI, lst.data(), nitems * sizeof(int32_t),
                             hipMemcpyHostToDevice, s));
    else
      CTM_HIP(stage_copy(ctx, LI, lst.data(), nitems * sizeof(int32_t)));
    // (80 and 64: the blocks above 64 take 8 incoming directions per
    // thread, and up to order 80 two such chunks side by side fit the LDS)
    const int caps[] = {nmax_all, 80, 64, 32, 16};
    constexpr int kCaps = 5;
    // the buckets write disjoint items: side by side on the lanes
    CTM_HIP(fork_lanes(ctx));
    size_t pos = 0;
    for (int b = 0; b < kCaps && pos < nitems; ++b) {
      const int cap = caps[b];
      const int lo = b < kCaps - 1 ? caps[b + 1] : 0;
      if (cap <= lo) continue;
      size_t end = pos;
      double fb = 0.;
      while (end < nitems && nm_of(ctx->final_items[lst[end]]) > lo) {
        const double nm = nm_of(ctx->final_items[lst[end]]);
        fb += 8. * (4. * nm * nm * sp.n_in + 8. * nm * sp.n_pairs);
        ++end;
      }
      if (end > pos) {
        const hipStream_t ls = lane(ctx, b % kLanes);
        CTM_LAUNCH_Q(quiet, ls, CTM_KERNEL_OBS_X, fb, 0.,
                      launch_obs_x(sp, LI + pos,
                                   (int)(end - pos), cap, cap > 64 ? 8 : 4,
                                   ctx->d_items.as<Item>(),
                                   ctx->obs_source >= 0
                                       ? ctx->d_dt.as<cdbl>() +
                                             (int64_t)ctx->obs_source *
                                                 ctx->work.t_size
                                   : ctx->obs_source == kObsSourcePM
                                       ? ctx->d_tpm.as<cdbl>()
                                       : ctx->d_t.as<cdbl>(),
                                   ctx->averaged ? ctx->d_avg.as<cdbl>()
                                                 : nullptr,
                                   ctx->averaged ? ctx->d_aoff.as<int64_t>()
                                                 : nullptr,
                                   PT, nmax_all, X, ls));
      }
      pos = end;
    }
    CTM_HIP(join_lanes(ctx));
  }
  if (nth > 0 && (qext || qextpol || qextcpol || qabs || qabspol)) {
    CTM_LAUNCH_Q(quiet, s, CTM_KERNEL_OBS_SUM, 0., 0.,
               launch_obs_q(sp, n, ctx->d_gioff.as<int32_t>(), ctx->gp_saved,
                            X,
                            spec->do_extinction ? qext : nullptr,
                            spec->do_extinction ? qextpol : nullptr,
                            spec->do_extinction ? qextcpol : nullptr,
                            spec->do_absorption ? qabs : nullptr,
                            spec->do_absorption ? qabspol : nullptr, s));
  }
  if (nst > 0 && Z) {
    CTM_LAUNCH_Q(quiet, s, CTM_KERNEL_OBS_SUM, 0., 0.,
               launch_obs_z(sp, n, ctx->d_gioff.as<int32_t>(), ctx->gp_saved,
                            X, Z, s));
  }
  // the spec blob and the host vectors must outlive the queued copies
  CTM_HIP(drain(ctx));
  return CTM_OK;
}

int ctm_observables(ctm_context *ctx, const ctm_observable_spec *spec,
                    double *qext, double *qextpol, double *qextcpol,
                    double *qabs, double *qabspol, double *Z) {
  return observables_impl(ctx, spec, qext, qextpol, qextcpol, qabs, qabspol,
                          Z, nullptr, 0);
}

int ctm_observables_noalloc(ctm_context *ctx,
                            const ctm_observable_spec *spec, double *qext,
                            double *qextpol, double *qextcpol, double *qabs,
                            double *qabspol, double *Z) {
  if (!ctx) return CTM_ERR_ARGUMENT;
  if (!ctx->scratch) {
    ctx->err = "no fixed scratch on this device";
    return CTM_ERR_STATE;
  }
  return observables_impl(ctx, spec, qext, qextpol, qextcpol, qabs, qabspol,
                          Z, ctx->scratch, kNoAllocScratchBytes);
}

int ctm_orientation_average(ctm_context *ctx, int n_dist, int ncoef,
                            const double *pN_host,
                            const int32_t *dist_of_grain) {
  if (!ctx) return CTM_ERR_ARGUMENT;
  if (ctx->final_items.empty()) {
    ctx->err = "no T-matrices in the context";
    return CTM_ERR_STATE;
  }
  const int n = ctx->n_grains;
  int nmax = 1;
  for (int g = 0; g < n; ++g) nmax = std::max(nmax, (int)ctx->nmax[g]);
  if (n_dist < 1 || !pN_host || ncoef < 2 * nmax + 1) {
    ctx->err = "need >= 2 nmax + 1 expansion coefficients per distribution "
               "(TMatrixCalculator.hpp:241-242)";
    return CTM_ERR_ARGUMENT;
  }
  if (nmax > kMaxAvgOrder / 2) {
    ctx->err = "order too large for the averaging kernels";
    return CTM_ERR_ARGUMENT;
  }
  if (dist_of_grain) {
    for (int g = 0; g < n; ++g)
      if (dist_of_grain[g] < -1 || dist_of_grain[g] >= n_dist) {
        ctx->err = "distribution index out of range";
        return CTM_ERR_ARGUMENT;
      }
  }
  CTM_HIP(hipSetDevice(ctx->device));
  hipStream_t s = ctx->stream;
  // orders N with a non-zero coefficient in any distribution
  std::vector<char> needN(2 * nmax + 1, 0);
  for (int dd = 0; dd < n_dist; ++dd)
    for (int N = 0; N <= 2 * nmax; ++N)
      if (pN_host[(size_t)dd * ncoef + N] != 0.) needN[N] = 1;
  // CG keys for (n1, n2, N) and the dense index table
  const size_t idx_size = (size_t)(nmax + 1) * (nmax + 1) * (2 * nmax + 1);
  std::vector<int32_t> cgidx(idx_size, -1);
  std::vector<CgKey> keys;
  int64_t cgsize = 0;
  for (int n1 = 1; n1 <= nmax; ++n1)
    for (int n2 = 1; n2 <= nmax; ++n2) {
      const int lo = std::abs(n1 - n2), hi = n1 + n2;
      for (int N = lo; N <= hi; ++N) {
        if (!needN[N]) continue;
        CgKey k{n1, n2, N, 0, cgsize};
        cgidx[((size_t)n1 * (nmax + 1) + n2) * (2 * nmax + 1) + N] =
            (int32_t)cgsize;
        cgsize += std::min(n1, n2) + 1;
        keys.push_back(k);
      }
    }
  if (cgsize > INT32_MAX) {
    ctx->err = "Clebsch-Gordan table too large";
    return CTM_ERR_ARGUMENT;
  }
  // per distribution: the number of orders N < k with a non-zero
  // coefficient.  A pair (n1, n2) couples only through |n1 - n2| <= N <=
  // n1 + n2: pairs no such N reaches have a zero ensemble entry, written by
  // one fill instead of a task (Mishchenko alignment: N = 0, 2 only)
  std::vector<int32_t> ncum((size_t)n_dist * (2 * nmax + 2), 0);
  for (int dd = 0; dd < n_dist; ++dd)
    for (int N = 0; N <= 2 * nmax; ++N)
      ncum[(size_t)dd * (2 * nmax + 2) + N + 1] =
          ncum[(size_t)dd * (2 * nmax + 2) + N] +
          (pN_host[(size_t)dd * ncoef + N] != 0. ? 1 : 0);
  std::vector<AvgTask> tasks;
  ctx->aoff.assign(ctx->final_items.size(), 0);
  int64_t asize = 0;
  bool skipped = false;
  for (int g = 0; g < n; ++g) {
    const int Ng = (int)ctx->nmax[g];
    const int dd = dist_of_grain ? dist_of_grain[g] : 0;
    const int32_t *nc = dd >= 0 ? ncum.data() + (size_t)dd * (2 * nmax + 2)
                                : nullptr;
    for (int n1 = 1; n1 <= Ng; ++n1)
      for (int n2 = 1; n2 <= Ng; ++n2) {
        // alignment disabled (dd < 0): every pair is a copy
        if (nc && nc[n1 + n2 + 1] == nc[std::abs(n1 - n2)]) {
          skipped = true;
          continue;
        }
        tasks.push_back({g, n1, n2, 0});
      }
    for (int m = 0; m <= Ng; ++m) {
      const int64_t nm = Ng + 1 - (m > 0 ? m : 1);
      ctx->aoff[ctx->grain_items_off[g] + m] = asize;
      asize += 4 * nm * nm;
    }
  }
  std::vector<double> pn(pN_host, pN_host + (size_t)n_dist * ncoef);
  std::vector<int32_t> dist;
  if (dist_of_grain) dist.assign(dist_of_grain, dist_of_grain + n);
  CTM_HIP(upload(ctx, ctx->d_cgkeys, keys));
  CTM_HIP(upload(ctx, ctx->d_cgidx, cgidx));
  CTM_HIP(upload(ctx, ctx->d_avgtasks, tasks));
  CTM_HIP(upload(ctx, ctx->d_aoff, ctx->aoff));
  CTM_HIP(upload(ctx, ctx->d_pn, pn));
  if (dist_of_grain) CTM_HIP(upload(ctx, ctx->d_dist, dist));
  CTM_HIP(ctx->d_cg.ensure(std::max<int64_t>(cgsize, 1) * sizeof(double)));
  CTM_HIP(ctx->d_avg.ensure(std::max<int64_t>(asize, 1) * sizeof(cdbl)));
  if (skipped)
    CTM_LAUNCH(CTM_KERNEL_AVERAGE, 0., 16. * asize,
               launch_zero(ctx->d_avg.as<cdbl>(), asize, s));
  CTM_LAUNCH(CTM_KERNEL_AVERAGE, 0., 0.,
             launch_cg(ctx->d_cgkeys.as<CgKey>(), (int)keys.size(),
                       ctx->d_cg.as<double>(), s));
  CTM_LAUNCH(CTM_KERNEL_AVERAGE, 0., 0.,
             launch_average(ctx->d_avgtasks.as<AvgTask>(), (int)tasks.size(),
                            ctx->d_items.as<Item>(),
                            ctx->d_gioff.as<int32_t>(), ctx->d_t.as<cdbl>(),
                            ctx->d_pn.as<double>(), ncoef,
                            dist_of_grain ? ctx->d_dist.as<int32_t>() : nullptr,
                            ctx->d_cgidx.as<int32_t>(), ctx->d_cg.as<double>(),
                            nmax, ctx->d_avg.as<cdbl>(),
                            ctx->d_aoff.as<int64_t>(), s));
  CTM_HIP(drain(ctx));
  ctx->averaged = true;
  return CTM_OK;
}

int ctm_weighted_reduce(ctm_context *ctx, int n_groups, int group,
                        int64_t width, const double *in, const double *weights,
                        int normalise, double *out) {
  if (!ctx) return CTM_ERR_ARGUMENT;
  if (n_groups < 0 || group < 1 || width < 0 ||
      (n_groups > 0 && width > 0 && (!in || !weights || !out))) {
    ctx->err = "invalid reduction arguments";
    return CTM_ERR_ARGUMENT;
  }
  CTM_HIP(hipSetDevice(ctx->device));
  hipStream_t s = ctx->stream;
  CTM_LAUNCH(CTM_KERNEL_AVERAGE, 0., 3. * 8. * n_groups * group * width,
             launch_weighted_reduce(n_groups, group, width, in, weights,
                                    normalise, out, s));
  return CTM_OK;
}

int ctm_tmatrix_set(ctm_context *ctx, int g, const double *T_host,
                    size_t count) {
  if (!ctx) return CTM_ERR_ARGUMENT;
  if (ctx->final_items.empty()) {
    ctx->err = "no T-matrices in the context";
    return CTM_ERR_STATE;
  }
  if (ctx->obs_source >= 0) {
    ctx->err = "a tangent is selected for the observables";
    return CTM_ERR_STATE;
  }
  const size_t need = ctm_tmatrix_size(ctx, g);
  if (need == 0 || !T_host || count != need) {
    ctx->err = "invalid grain index or element count";
    return CTM_ERR_ARGUMENT;
  }
  if (!ctx->averaged) {
    // the full-layout copy of every grain's T-matrix (alignment disabled)
    int nmax = 1;
    for (int i = 0; i < ctx->n_grains; ++i)
      nmax = std::max(nmax, (int)ctx->nmax[i]);
    std::vector<double> pn(2 * (size_t)nmax + 1, 0.);
    std::vector<int32_t> dist(ctx->n_grains, -1);
    const int rc = ctm_orientation_average(ctx, 1, 2 * nmax + 1, pn.data(),
                                           dist.data());
    if (rc != CTM_OK) return rc;
  }
  CTM_HIP(hipSetDevice(ctx->device));
  const int64_t o = ctx->aoff[ctx->grain_items_off[g]];
  CTM_HIP(hipMemcpyAsync(ctx->d_avg.as<cdbl>() + o, T_host,
                         need * sizeof(cdbl), hipMemcpyHostToDevice,
                         ctx->stream));
  CTM_HIP(drain(ctx));
  return CTM_OK;
}

int ctm_average_get(ctm_context *ctx, int g, double *T_host, size_t capacity) {
  if (!ctx) return CTM_ERR_ARGUMENT;
  if (!ctx->averaged) {
    ctx->err = "no ensemble T-matrices in the context";
    return CTM_ERR_STATE;
  }
  const size_t need = ctm_tmatrix_size(ctx, g);
  if (need == 0 || !T_host || capacity < need) {
    ctx->err = "invalid grain index or output capacity";
    return CTM_ERR_ARGUMENT;
  }
  CTM_HIP(hipSetDevice(ctx->device));
  const int64_t o = ctx->aoff[ctx->grain_items_off[g]];
  CTM_HIP(hipMemcpyAsync(T_host, ctx->d_avg.as<cdbl>() + o, need * sizeof(cdbl),
                         hipMemcpyDeviceToHost, ctx->stream));
  CTM_HIP(drain(ctx));
  return CTM_OK;
}

int ctm_query(ctm_context *ctx, int n, const ctm_query_spec *q,
              double *out_host) {
  if (!ctx) return CTM_ERR_ARGUMENT;
  if (ctx->final_items.empty()) {
    ctx->err = "no T-matrices in the context";
    return CTM_ERR_STATE;
  }
  if (n < 0 || (n > 0 && (!q || !out_host))) {
    ctx->err = "invalid query arrays";
    return CTM_ERR_ARGUMENT;
  }
  if (n == 0) return CTM_OK;
  std::vector<Query> qs(n);
  int qmax = 1;
  for (int i = 0; i < n; ++i) {
    if (q[i].grain < 0 || q[i].grain >= ctx->n_grains) {
      ctx->err = "query grain out of range";
      return CTM_ERR_ARGUMENT;
    }
    qmax = std::max(qmax, (int)ctx->nmax[q[i].grain]);
    if ((int)ctx->nmax[q[i].grain] >= kMaxQueryOrder) {
      ctx->err = "order too large for the query kernel";
      return CTM_ERR_ARGUMENT;
    }
    qs[i] = Query{q[i].grain, 0, q[i].alpha, q[i].beta, q[i].theta_in,
                  q[i].phi_in, q[i].theta_out, q[i].phi_out};
  }
  CTM_HIP(hipSetDevice(ctx->device));
  hipStream_t s = ctx->stream;
  CTM_HIP(upload(ctx, ctx->d_query, qs));
  CTM_HIP(ctx->d_qout.ensure((size_t)n * kQueryOut * sizeof(double)));
  CTM_LAUNCH(CTM_KERNEL_OBS_SUM, 0., 0.,
             launch_query(ctx->d_query.as<Query>(), n, qmax,
                          ctx->d_items.as<Item>(),
                          ctx->d_gioff.as<int32_t>(), ctx->gp_saved,
                          ctx->d_t.as<cdbl>(),
                          ctx->averaged ? ctx->d_avg.as<cdbl>() : nullptr,
                          ctx->averaged ? ctx->d_aoff.as<int64_t>() : nullptr,
                          ctx->d_qout.as<double>(), s));
  CTM_HIP(hipMemcpyAsync(out_host, ctx->d_qout.p,
                         (size_t)n * kQueryOut * sizeof(double),
                         hipMemcpyDeviceToHost, s));
  CTM_HIP(drain(ctx));
  return CTM_OK;
}

int ctm_m0_sums(ctm_context *ctx, int n, const double *radius,
                const double *axis_ratio, const double *wavelength,
                const double *mr_re, const double *mr_im,
                const uint32_t *nmax_host, const uint32_t *ngauss_host,
                double *qsca_host, double *qext_host) {
  int rc = check_inputs(ctx, n, radius, axis_ratio, wavelength, mr_re, mr_im);
  if (rc != CTM_OK) return rc;
  if (n > 0 && (!nmax_host || !ngauss_host || !qsca_host || !qext_host)) {
    ctx->err = "missing candidate arrays";
    return CTM_ERR_ARGUMENT;
  }
  for (int g = 0; g < n; ++g)
    if (nmax_host[g] < 1 || (int)nmax_host[g] > kMaxOrder ||
        ngauss_host[g] < 1) {
      ctx->err = "candidate order out of range (1..200) or ngauss < 1";
      return CTM_ERR_ARGUMENT;
    }
  CTM_HIP(hipSetDevice(ctx->device));
  hipStream_t s = ctx->stream;
  CTM_HIP(drain(ctx));
  // the work lists below replace the batch's items
  ctx->final_items.clear();
  ctx->averaged = false;
  ctx->have_tangents = false;
  ctx->obs_source = -1;
  ctx->n_grains = 0;
  ctx->shared.clear();
  ctx->mixed = false;
  if (n == 0) return CTM_OK;
  GrainParams gp{radius, axis_ratio, wavelength, mr_re, mr_im};
  WorkList &w = ctx->work;
  w.clear();
  for (int g = 0; g < n; ++g) {
    const int N = (int)nmax_host[g], G = (int)ngauss_host[g];
    const int set = w.add_set(g, N, G, ctx->shared.gl_keys[ctx->shared.gl(G)].off);
    w.add_item(g, N, 0, G, set, ctx->shared.wig(N, 0, G), false);
  }
  rc = run_work(ctx, gp, false);
  if (rc != CTM_OK) return rc;
  std::vector<double> qp((size_t)n * 2 * kQParts);
  CTM_HIP(hipMemcpyAsync(qp.data(), ctx->d_qparts.p, qp.size() * sizeof(double),
                         hipMemcpyDeviceToHost, s));
  CTM_HIP(drain(ctx));
  for (int g = 0; g < n; ++g) {
    const double *p0 = qp.data() + (size_t)(2 * g) * kQParts;
    const double *p1 = p0 + kQParts;
    qsca_host[g] = p0[0] + p1[0];
    qext_host[g] = p0[1] + p1[1];
  }
  return CTM_OK;
}

int ctm_precision_stats(ctm_context *ctx, int64_t *blocks_mixed,
                        int64_t *blocks_fallback) {
  if (!ctx) return CTM_ERR_ARGUMENT;
  if (blocks_mixed) *blocks_mixed = ctx->mixed_blocks;
  if (blocks_fallback) *blocks_fallback = ctx->mixed_fallback;
  return CTM_OK;
}

int ctm_profile_enable(ctm_context *ctx, int on) {
  if (!ctx) return CTM_ERR_ARGUMENT;
  ctx->prof.on = on != 0;
  return CTM_OK;
}

int ctm_profile_reset(ctm_context *ctx) {
  if (!ctx) return CTM_ERR_ARGUMENT;
  CTM_HIP(drain(ctx));
  ctx->prof.reset();
  return CTM_OK;
}

int ctm_profile_read(ctm_context *ctx, int kernel, ctm_kernel_stats *out) {
  if (!ctx || !out || kernel < 0 || kernel >= CTM_KERNEL_COUNT)
    return CTM_ERR_ARGUMENT;
  CTM_HIP(drain(ctx));
  ctx->prof.collect();
  out->launches = ctx->prof.launches[kernel];
  out->milliseconds = ctx->prof.ms[kernel];
  out->flops = ctx->prof.flops[kernel];
  out->bytes = ctx->prof.bytes[kernel];
  return CTM_OK;
}

int ctm_batch_orders(ctm_context *ctx, uint32_t *nmax_host,
                     uint32_t *ngauss_host, int32_t *status_host) {
  if (!ctx) return CTM_ERR_ARGUMENT;
  for (int g = 0; g < ctx->n_grains; ++g) {
    if (nmax_host) nmax_host[g] = ctx->nmax[g];
    if (ngauss_host) ngauss_host[g] = ctx->ngauss[g];
    if (status_host) status_host[g] = ctx->status[g];
  }
  return CTM_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Batch plan: the batch step (convergence rounds, all-m blocks, per-grain
// sums and the averaged Q_ext / Q_abs) with every work list built on the
// device (ctm_plan.hip), so a run enqueues a fixed sequence of launches and
// the host reads nothing back until the step has finished.  The
// convergence runs in three rounds: the first window of every grain
// (first_window_end), then half its order more, then every order up to the
// last; each order is tested as the reference tests it, so the converged
// orders are the host path's.  Data pools (tables, Q, T, the observables'
// X) are sized at creation from the first windows; a run that needs more
// flags it on the device (its lists stay empty, nothing is written out of
// bounds), and ctm_plan_run grows the pools to the sizes the device
// reported and runs again.

struct ctm_plan {
  ctm_context *ctx = nullptr;
  int n = 0;
  GrainParams gp{};
  int min_order = 0, last_order = 0, glfac = 0;
  double tol = 0.;
  bool mixed = false;
  ctm_plan_outputs out{};
  bool want_x = false;
  // while tuning: a run whose pools would outgrow this many bytes stops
  // with CTM_ERR_MEMORY instead of growing them (0: no limit)
  size_t grow_limit = 0;
  ObsLayout lay;
  // grain-independent tables: Gauss-Legendre nodes of every G = glfac N and
  // Wigner values of every (N, m) for N = min_order .. last_order
  DevBuf d_gl, d_glkeys, d_wig, d_wigkeys, d_gloff, d_wigbase, d_spec, d_pt;
  // lists and per-grain state
  DevBuf d_sets, d_items, d_ord, d_pre, d_sol, d_obsl, d_flags, d_fb, d_hist,
      d_cnt, d_goff, d_gitem, d_gobs, d_lo, d_hi, d_state, d_phase, d_status,
      d_iota, d_gcap;
  // data pools
  DevBuf d_tables, d_q, d_t, d_x, d_qparts, d_x32, d_xexp, d_rbuf;
  PlanCaps caps{};
  PlanLists L{};
  int cus = 256;
  // grid bound of every (phase, list): one workgroup per tile, group of
  // four blocks, task or item (one thread per table node); the first run
  // fits them to the measured list sizes
  int32_t gcap[kPlanPhases][kPlanLists]{};
  bool gcap_fitted = false;
  // end of each all-m solve bucket (the observables of its block orders
  // start after it)
  hipEvent_t ev_solve[kPlanBuckets] = {};
  // end of the k_obs_x bucket of orders up to 16 (all-m phase, FP64)
  hipEvent_t ev_obs16 = nullptr;
  // the sums of grains up to order 32 were launched on the lanes
  bool obs_q_early = false;
  std::vector<int> hi1;  // host estimate of the first windows
  // after a run
  PlanStatus st{};
  PlanPhase ph[kPlanPhases]{};
  int64_t runs = 0, reruns = 0;
  // profiler records of the current run: the work of a launch is known
  // only from the device lists, so it is credited after the run
  struct Rec {
    int kernel;
    hipEvent_t a, b;
    int phase, kind, idx;  // kind: 1 part, 2 bucket, 3 fallback, 4 obs, 5 tables
  };
  std::vector<Rec> recs;
};

namespace {

#define PLAN_HIP(expr)                                                  \
  do {                                                                  \
    hipError_t _e = (expr);                                             \
    if (_e != hipSuccess) {                                             \
      p->ctx->err = std::string(#expr) + ": " + hipGetErrorString(_e);  \
      return CTM_ERR_HIP;                                               \
    }                                                                   \
  } while (0)

// Launch `call` on `strm`, between two profiler events when the context
// profiles.
#define PLAN_LAUNCH(strm, kernel, phase_, kind_, idx_, call)                  \
  do {                                                                        \
    ctm_context *_c = p->ctx;                                                 \
    const bool _rec = _c->prof.on;                                            \
    ctm_plan::Rec _r{kernel, nullptr, nullptr, phase_, kind_, idx_};          \
    if (_rec) {                                                               \
      _r.a = _c->prof.take();                                                 \
      _r.b = _c->prof.take();                                                 \
      if (_r.a && _r.b) PLAN_HIP(hipEventRecord(_r.a, strm));                 \
    }                                                                         \
    PLAN_HIP(call);                                                           \
    if (_rec && _r.a && _r.b) {                                               \
      PLAN_HIP(hipEventRecord(_r.b, strm));                                   \
      p->recs.push_back(_r);                                                  \
    }                                                                         \
  } while (0)

// Sizes of the solve launch of bucket b: largest block order, LDS store of
// the blocked elimination and rows of its largest sub-block (m = 0 blocks
// split in two by parity).
void plan_solve_shape(int nm_max, bool m0_only, int64_t *store, int *rows) {
  if (m0_only && nm_max > 1) {
    const int r0 = (nm_max + 1) / 2, r1 = nm_max - r0;
    *store = (int64_t)r0 * blk_lds_stride(r0) + (int64_t)r1 * blk_lds_stride(r1);
    *rows = r0;
  } else {
    *store = (int64_t)nm_max * blk_lds_stride(nm_max);
    *rows = nm_max;
  }
}

// k_obs_x of the all-m blocks, one launch per observable bucket.  With
// `solved` (the solve buckets' ends recorded in ev_solve and the lanes not
// yet joined) a bucket starts on a lane once the solve buckets that hold
// its block orders have finished, beside the remaining solves; otherwise
// (all T blocks complete) the buckets run side by side.
// Lanes of the observable buckets in the all-m phase of the FP64 path (see
// plan_obs_x); the early sums run on those of the last two buckets.
constexpr int kPlanObsLane[kPlanObsBuckets] = {0, 0, 1, 3, 2};

int plan_obs_x(ctm_plan *p, const bool *solved) {
  ctm_context *ctx = p->ctx;
  const PlanLists &L = p->L;
  const int ph = kPlanPhases - 1;
  PlanPhase *dph = p->d_phase.as<PlanPhase>() + ph;
  const int64_t cap = p->caps.items;
  for (int b = 0; b < kPlanObsBuckets; ++b) {
    const int capn = plan_obs_hi(b, p->last_order);
    const int lo_b =
        b + 1 < kPlanObsBuckets ? plan_obs_hi(b + 1, p->last_order) : 0;
    if (capn <= lo_b) continue;
    hipStream_t ls = lane(ctx, b % kLanes);
    if (solved) {
      // Lanes chosen so that no bucket queues behind a solve it does not
      // need (a lane runs in order): orders above 64 follow the order
      // 71-99 solves on lane 0, orders 33-64 the 49-64 solves on lane 1,
      // 17-32 the 33-48 solves on lane 3, up to 16 the smallest solves on
      // lane 2; each waits for the solve buckets holding its orders.
      ls = lane(ctx, kPlanObsLane[b]);
      for (int sb = 0; sb < kPlanBuckets; ++sb) {
        const int shi = plan_bucket_hi(sb);
        const int slo = sb + 1 < kPlanBuckets ? plan_bucket_hi(sb + 1) : 0;
        if (shi > lo_b && slo < capn && solved[sb])
          PLAN_HIP(hipStreamWaitEvent(ls, p->ev_solve[sb], 0));
      }
    }
    PLAN_LAUNCH(ls, CTM_KERNEL_OBS_X, ph, 4, b,
                launch_obs_x(p->lay.sp, L.obsl + b * cap, 0, capn,
                             capn > 64 ? 8 : 4, L.items, p->d_t.as<cdbl>(),
                             nullptr, nullptr, p->d_pt.as<double>(),
                             p->last_order, p->d_x.as<cdbl>(), ls,
                             &dph->obs[b][0],
                             std::max(1, p->gcap[ph][plan_list_obs(b)]),
                             p->L.obs_lean != 0));
  }
  p->obs_q_early = false;
  if (solved && p->last_order > 32) {
    // Sums of the grains whose blocks all lie in the last two buckets,
    // while the larger blocks are still being contracted: orders up to 16
    // (up to 17 m blocks) after the order-16 bucket on its lane, orders
    // 17-32 after both buckets on the lane of the order-32 bucket; the
    // larger grains are summed after the lanes join.
    const bool ext = p->out.obs_qext || p->out.obs_qextpol ||
                     p->out.obs_qextcpol;
    const bool ab = p->out.obs_qabs || p->out.obs_qabspol;
    if (ext || ab) {
      const hipStream_t l16 = lane(ctx, kPlanObsLane[kPlanObsBuckets - 1]);
      const hipStream_t l32 = lane(ctx, kPlanObsLane[kPlanObsBuckets - 2]);
      PLAN_HIP(hipEventRecord(p->ev_obs16, l16));
      PLAN_LAUNCH(l16, CTM_KERNEL_OBS_SUM, ph, 0, 0,
                  launch_obs_q(p->lay.sp, p->n, L.gobs, p->gp,
                               p->d_x.as<cdbl>(), p->out.obs_qext,
                               p->out.obs_qextpol, p->out.obs_qextcpol,
                               p->out.obs_qabs, p->out.obs_qabspol, l16, 0,
                               17));
      PLAN_HIP(hipStreamWaitEvent(l32, p->ev_obs16, 0));
      PLAN_LAUNCH(l32, CTM_KERNEL_OBS_SUM, ph, 0, 0,
                  launch_obs_q(p->lay.sp, p->n, L.gobs, p->gp,
                               p->d_x.as<cdbl>(), p->out.obs_qext,
                               p->out.obs_qextpol, p->out.obs_qextcpol,
                               p->out.obs_qabs, p->out.obs_qabspol, l32, 18,
                               33));
      p->obs_q_early = true;
    }
  }
  return CTM_OK;
}

// Tables, assembly and solves of phase `ph` (lists built on the device);
// in the all-m phase of the FP64 path with observables, also their k_obs_x.
int plan_phase(ctm_plan *p, int ph, bool want_t) {
  ctm_context *ctx = p->ctx;
  const PlanLists &L = p->L;
  PlanPhase *dph = p->d_phase.as<PlanPhase>() + ph;
  const int64_t cap = p->caps.items;
  const GrainParams &gp = p->gp;
  hipStream_t s = ctx->stream;
  const int32_t *gc = p->gcap[ph];
  PLAN_LAUNCH(s, CTM_KERNEL_TABLES, ph, 5, 0,
              launch_tables(L.sets, 0, 0, 0, gp, p->d_gl.as<double>(),
                            p->d_tables.as<double>(), s, &dph->sets_nodes[0],
                            std::max(1, (gc[0] + 127) / 128)));
  // assembly parts.  All m: the 3 x 3 part alone (so that its launch
  // measures the kernel), then the others side by side.  Windows: all parts
  // side by side from the start (the 3 x 3 launches of the windows are
  // credited to CTM_KERNEL_ASSEMBLE_WINDOW)
  const int kid[3] = {want_t ? CTM_KERNEL_ASSEMBLE_WIDE
                             : CTM_KERNEL_ASSEMBLE_WINDOW,
                      CTM_KERNEL_ASSEMBLE, CTM_KERNEL_ASSEMBLE_SMALL};
  // (window lists hold only m = 0 blocks: the packed m = 0 part takes the
  // lane of the empty packed m > 0 part; in the all-m phase it follows the
  // 2 x 2 part on lane 1: 15.51 -> 15.47 ms per batch step against lane 2,
  // lanes 0 and 3 of the windows' parts measured slower)
  const int lane_of_part[5] = {0, 1, 2, 3, want_t ? 1 : 3};
  // solve buckets, placed and ordered as the host path places them; in the
  // window rounds the order 33-48 bucket follows the order 49-64 one on
  // lane 1 instead of the order 65-70 one on lane 3 (the first window's
  // last solve chain: 15.51 -> 15.42 ms per batch step, other placements
  // 15.46-15.55)
  const int lane_all_m[kPlanBuckets] = {0, 3, 1, 3, 2, 2};
  const int lane_window[kPlanBuckets] = {0, 3, 1, 1, 2, 2};
  const int *lane_of_bucket = want_t ? lane_all_m : lane_window;
  const bool m0_only = !want_t;
  const bool mix = p->mixed && want_t;
  bool after_big = false;
  bool solved[kPlanBuckets] = {};
  bool launched[kPlanBuckets] = {};
  // the assembly parts that can hold blocks of bucket b's orders
  auto bucket_need = [&](int b) {
    const int lo_b = b + 1 < kPlanBuckets ? plan_bucket_hi(b + 1) : 0;
    const int nm_max = std::min(plan_bucket_hi(b), p->last_order);
    unsigned need = 0;
    for (int nm = lo_b + 1; nm <= nm_max; ++nm) {
      need |= 1u << block_part(nm, true, p->L.micro_ratio);
      if (!m0_only) need |= 1u << block_part(nm, false, p->L.micro_ratio);
    }
    return need;
  };
  auto launch_bucket = [&](int b) -> int {
    launched[b] = true;
    const int lo_b = b + 1 < kPlanBuckets ? plan_bucket_hi(b + 1) : 0;
    if (p->last_order <= lo_b) return CTM_OK;  // no block of these orders
    const int nm_max = std::min(plan_bucket_hi(b), p->last_order);
    int64_t store;
    int rows;
    plan_solve_shape(nm_max, m0_only, &store, &rows);
    const hipStream_t ls = lane(ctx, lane_of_bucket[b]);
    // the order 33-48 bucket follows the order 49-64 solves; the buckets up
    // to order 32 start as soon as their assembly parts end (measured
    // 15.62 -> 15.56 ms per batch step; all three following: 15.62)
    if (after_big && b == 3)
      PLAN_HIP(hipStreamWaitEvent(ls, ctx->ev_bucket, 0));
    const unsigned need = bucket_need(b);
    for (int part = 0; part < kPlanParts; ++part)
      if (need >> part & 1u)
        PLAN_HIP(hipStreamWaitEvent(ls, ctx->ev_part[part], 0));
    const int32_t *tasks = L.sol + b * 2 * cap;
    const int32_t *dn = &dph->bucket[b][0];
    const int grid = std::max(1, gc[plan_list_bucket(b)]);
    const bool mixb = mix && lo_b >= kMixedMinOrder - 1 &&
                      nm_max <= kMixedMaxOrder;
    if (mixb) {
      PLAN_LAUNCH(ls, CTM_KERNEL_SOLVE, ph, 2, b,
                  launch_invert32(tasks, 0, nm_max, store, rows, L.items,
                                  p->d_q.as<cdbl>(), p->d_x32.as<cflt>(),
                                  p->d_xexp.as<int32_t>(), ls, dn, grid));
      PLAN_LAUNCH(ls, CTM_KERNEL_REFINE, ph, 0, b,
                  launch_refine(tasks, 0, nm_max, L.items, p->d_q.as<cdbl>(),
                                p->d_x32.as<cflt>(), p->d_xexp.as<int32_t>(),
                                p->d_t.as<cdbl>(), p->d_rbuf.as<cdbl>(),
                                p->d_qparts.as<double>(),
                                L.flags + b * 2 * cap, ls, dn, grid));
    } else {
      PLAN_LAUNCH(ls, CTM_KERNEL_SOLVE, ph, 2, b,
                  launch_invert_solve(tasks, 0, nm_max, store, rows, L.items,
                                      p->d_q.as<cdbl>(),
                                      want_t ? p->d_t.as<cdbl>() : nullptr,
                                      p->d_qparts.as<double>(),
                                      want_t ? 0 : 1, nullptr, nullptr, ls,
                                      dn, grid));
    }
    if (want_t) {
      PLAN_HIP(hipEventRecord(p->ev_solve[b], ls));
      solved[b] = true;
    }
    if (b == 2 && want_t) {
      // as the host path: the bucket of order 33-48 follows the blocks of
      // order 49..64
      PLAN_HIP(hipEventRecord(ctx->ev_bucket, ls));
      after_big = true;
    }
    return CTM_OK;
  };
  // big_first (tunable): in the all-m phase the largest blocks' solves
  // (their longest elimination chains) are launched right after the 3 x 3
  // part that assembles them, before the other parts take the CUs' LDS
  // (1), or alone, the other parts waiting for them (2); only when that
  // part holds all of them
  const bool big_first = want_t && p->L.big_first && bucket_need(0) == 1u;
  const bool big_alone = big_first && p->L.big_first == 2;
  if (!want_t) PLAN_HIP(fork_lanes(ctx));
  for (int part = 0; part < kPlanParts; ++part) {
    if (part == 1 && want_t) {
      PLAN_HIP(fork_lanes(ctx));
      if (big_first) {
        const int rc = launch_bucket(0);
        if (rc != CTM_OK) return rc;
      }
    }
    // window lists hold only m = 0 blocks, so the packed m > 0 part is
    // empty: not launched (an empty launch still waits for free workgroup
    // slots behind the other parts, 0.4 ms in the first window, holding
    // back the packed m = 0 part on its lane; the step moved by less than
    // its spread); no window solve bucket waits for its event
    if (part == 3 && !want_t) continue;
    const hipStream_t ls = lane(ctx, lane_of_part[part]);
    if (part >= 1 && big_alone)
      PLAN_HIP(hipStreamWaitEvent(ls, p->ev_solve[0], 0));
    const int32_t *ord = L.ord + part * cap;
    const int grid = std::max(1, gc[plan_list_part(part)]);
    if (part < 3) {
      PLAN_LAUNCH(ls, kid[part], ph, 1, part,
                  launch_assemble(3 - part, ord, L.pre + part * (cap + 1), 0,
                                  0, L.items, L.sets,
                                  p->d_wigkeys.as<WigKey>(), gp,
                                  p->d_tables.as<double>(),
                                  p->d_wig.as<double>(), p->d_q.as<cdbl>(),
                                  ls, &dph->part[part][0], grid));
    } else {
      PLAN_LAUNCH(ls, CTM_KERNEL_ASSEMBLE_PACK, ph, 1, part,
                  launch_assemble_pack4(ord, 0, part == 4 ? 1 : 0, L.items,
                                        L.sets, p->d_wigkeys.as<WigKey>(), gp,
                                        p->d_tables.as<double>(),
                                        p->d_wig.as<double>(),
                                        p->d_q.as<cdbl>(), ls,
                                        &dph->part[part][0], grid));
    }
    PLAN_HIP(hipEventRecord(ctx->ev_part[part], ls));
  }
  for (int b = 0; b < kPlanBuckets; ++b) {
    if (launched[b]) continue;
    const int rc = launch_bucket(b);
    if (rc != CTM_OK) return rc;
  }
  if (want_t && !mix && p->want_x) {
    const int rc = plan_obs_x(p, solved);
    if (rc != CTM_OK) return rc;
  }
  PLAN_HIP(join_lanes(ctx));
  if (mix) {
    // blocks whose refinement stalled: FP64 solve of the compacted lists
    int nmb = 0;
    for (int b = 0; b < kPlanBuckets; ++b) {
      const int lo_b = b + 1 < kPlanBuckets ? plan_bucket_hi(b + 1) : 0;
      if (lo_b >= kMixedMinOrder - 1) nmb = b + 1;
    }
    PLAN_HIP(launch_plan_fallback(L, ph, nmb, s));
    PLAN_HIP(fork_lanes(ctx));
    for (int b = 0; b < nmb; ++b) {
      const int lo_b = plan_bucket_hi(b + 1);
      if (p->last_order <= lo_b) continue;
      const int nm_max = std::min(plan_bucket_hi(b), p->last_order);
      int64_t store;
      int rows;
      plan_solve_shape(nm_max, false, &store, &rows);
      const hipStream_t ls = lane(ctx, lane_of_bucket[b]);
      PLAN_LAUNCH(ls, CTM_KERNEL_SOLVE, ph, 3, b,
                  launch_invert_solve(L.fb + b * 2 * cap, 0, nm_max, store,
                                      rows, L.items, p->d_q.as<cdbl>(),
                                      p->d_t.as<cdbl>(),
                                      p->d_qparts.as<double>(), 0, nullptr,
                                      nullptr, ls, &dph->fallback[b][0],
                                      std::max(1,
                                               gc[plan_list_bucket(b)])));
    }
    PLAN_HIP(join_lanes(ctx));
  }
  return CTM_OK;
}

// Enqueues one run of the plan on the context's stream (lanes joined at
// the end).
int plan_enqueue(ctm_plan *p) {
  ctm_context *ctx = p->ctx;
  const PlanLists &L = p->L;
  hipStream_t s = ctx->stream;
  const int grid_items = p->cus * 8;
  PLAN_HIP(launch_plan_first(L, p->gp, s));
  for (int r = 0; r < 3; ++r) {
    PLAN_HIP(launch_plan_build(L, kPlanWindow, r, false, grid_items, s));
    int rc = plan_phase(p, r, false);
    if (rc != CTM_OK) return rc;
    PLAN_LAUNCH(s, CTM_KERNEL_CONVERGE, r, 0, 0,
                launch_plan_converge(L, r, p->d_qparts.as<double>(), p->tol,
                                     s));
    if (r < 2) PLAN_HIP(launch_plan_next(L, r == 1, s));
  }
  const int ph = 3;
  PLAN_HIP(launch_plan_build(L, kPlanAllM, ph, p->want_x, grid_items, s));
  int rc = plan_phase(p, ph, true);
  if (rc != CTM_OK) return rc;
  PLAN_LAUNCH(s, CTM_KERNEL_SUMS, ph, 0, 0,
              launch_full_sums(p->n, L.gitem, p->d_iota.as<int32_t>(),
                               p->d_qparts.as<double>(), p->out.qsca,
                               p->out.qext, s));
  PLAN_HIP(launch_write_orders(p->n, L.state, p->out.nmax, p->out.status, s));
  if (p->want_x) {
    if (p->mixed) {
      // after the FP64 re-solves of the mixed path: every T block is final
      PLAN_HIP(fork_lanes(ctx));
      rc = plan_obs_x(p, nullptr);
      if (rc != CTM_OK) return rc;
      PLAN_HIP(join_lanes(ctx));
    }
    const bool ext = p->out.obs_qext || p->out.obs_qextpol ||
                     p->out.obs_qextcpol;
    const bool ab = p->out.obs_qabs || p->out.obs_qabspol;
    if (ext || ab)
      PLAN_LAUNCH(s, CTM_KERNEL_OBS_SUM, ph, 0, 0,
                  launch_obs_q(p->lay.sp, p->n, L.gobs, p->gp,
                               p->d_x.as<cdbl>(), p->out.obs_qext,
                               p->out.obs_qextpol, p->out.obs_qextcpol,
                               p->out.obs_qabs, p->out.obs_qabspol, s,
                               p->obs_q_early ? 34 : 0));
  }
  return CTM_OK;
}

// (Re)allocates the data pools for the current capacities and refreshes
// the device list descriptor.
int plan_alloc(ctm_plan *p) {
  const PlanCaps &c = p->caps;
  PLAN_HIP(p->d_tables.ensure(std::max<int64_t>(c.tables, 1) * sizeof(double)));
  PLAN_HIP(p->d_q.ensure(std::max<int64_t>(c.q, 1) * sizeof(cdbl)));
  PLAN_HIP(p->d_t.ensure(std::max<int64_t>(c.t, 1) * sizeof(cdbl)));
  if (p->want_x)
    PLAN_HIP(p->d_x.ensure(std::max<int64_t>(c.x, 1) * sizeof(cdbl)));
  if (p->mixed) {
    PLAN_HIP(p->d_x32.ensure(std::max<int64_t>(c.t, 1) * sizeof(cflt)));
    PLAN_HIP(p->d_xexp.ensure(std::max<int64_t>(c.t, 1) * sizeof(int32_t)));
    PLAN_HIP(p->d_rbuf.ensure(std::max<int64_t>(c.t, 1) * sizeof(cdbl)));
  }
  // the pools may have been reallocated: the capacities the device checks
  // against are the requested ones (DevBuf rounds up)
  p->L.caps = c;
  PLAN_HIP(p->d_gcap.ensure(sizeof(p->gcap)));
  PLAN_HIP(hipMemcpy(p->d_gcap.p, p->gcap, sizeof(p->gcap),
                     hipMemcpyHostToDevice));
  p->L.gcap = p->d_gcap.as<int32_t>();
  return CTM_OK;
}

// Grid bound from a list size: a quarter more, at least 64.
int32_t plan_grid_of(int64_t units) {
  return (int32_t)std::min<int64_t>(units + units / 4 + 64, 0x7fffffff);
}

// First grid bounds: the lists of the first window round (the device uses
// the same windows), the all-m blocks with every grain at its first
// window's last order (most grains converge inside it), and an eighth of
// the first round's for the later rounds.
void plan_first_grids(ctm_plan *p, const std::vector<int> &hi1) {
  int64_t c[kPlanPhases][kPlanLists] = {};
  const int glfac = p->glfac;
  auto add = [&](int ph, int nm, bool m0) {
    const int part = block_part(nm, m0, p->L.micro_ratio);
    if (part < 3) {
      const int tr = 16 * (3 - part);
      const int ntr = (nm + tr - 1) / tr;
      c[ph][plan_list_part(part)] +=
          (int64_t)ntr * (m0 ? (nm + 2 * tr - 1) / (2 * tr) : ntr);
    } else {
      c[ph][plan_list_part(part)] += 1;  // blocks; groups of four below
    }
    c[ph][plan_list_bucket(plan_bucket_of(nm))] += 2;
    if (ph == 3) c[ph][plan_list_obs(plan_obs_of(nm))] += 1;
  };
  for (size_t g = 0; g < hi1.size(); ++g) {
    for (int N = p->min_order; N <= hi1[g]; ++N) {
      c[0][0] += glfac * N;
      add(0, N, true);
    }
    const int N = hi1[g];
    c[3][0] += glfac * N;
    for (int m = 0; m <= N; ++m) add(3, N + 1 - (m > 0 ? m : 1), m == 0);
  }
  for (int ph : {0, 3})
    for (int q = 3; q < 5; ++q)
      c[ph][plan_list_part(q)] = (c[ph][plan_list_part(q)] + 3) / 4;
  for (int k = 0; k < kPlanLists; ++k) {
    p->gcap[0][k] = plan_grid_of(c[0][k]);
    p->gcap[3][k] = plan_grid_of(c[3][k]);
    p->gcap[1][k] = p->gcap[2][k] = plan_grid_of(c[0][k] / 8);
  }
}

// Waits for the run, reads the status and phase records back and credits
// the profiler records with the work the device lists hold.
int plan_finish(ctm_plan *p) {
  ctm_context *ctx = p->ctx;
  {
    void *dst[2] = {&p->st, p->ph};
    const void *src[2] = {p->d_status.p, p->d_phase.p};
    const size_t bytes[2] = {sizeof(PlanStatus),
                             kPlanPhases * sizeof(PlanPhase)};
    if (read_back(ctx, 2, dst, src, bytes) != hipSuccess) {
      ctx->err = "plan: status read-back failed";
      return CTM_ERR_HIP;
    }
  }
  for (const ctm_plan::Rec &r : p->recs) {
    float ms = 0.f;
    const bool ok = hipEventElapsedTime(&ms, r.a, r.b) == hipSuccess;
    const PlanPhase &P = p->ph[r.phase];
    double f = 0., b = 0.;
    bool work = true;
    switch (r.kind) {
      case 1:
        f = P.part_flops[r.idx];
        b = P.part_bytes[r.idx];
        work = P.part[r.idx][0] > 0;
        break;
      case 2:
        f = P.bucket_flops[r.idx];
        b = P.bucket_bytes[r.idx];
        work = P.bucket[r.idx][0] > 0;
        break;
      case 3:
        work = P.fallback[r.idx][0] > 0;
        if (work && P.bucket[r.idx][0] > 0) {
          const double share =
              (double)P.fallback[r.idx][0] / P.bucket[r.idx][0];
          f = share * P.bucket_flops[r.idx];
          b = share * P.bucket_bytes[r.idx];
        }
        break;
      case 4:
        f = P.obs_flops[r.idx];
        work = P.obs[r.idx][0] > 0;
        break;
      case 5:
        b = 8. * (double)P.tables;
        work = P.sets_nodes[0] > 0;
        break;
      default:
        break;
    }
    if (ok && work && p->st.overflow == 0) {
      ctx->prof.launches[r.kernel] += 1;
      ctx->prof.ms[r.kernel] += ms;
      ctx->prof.flops[r.kernel] += f;
      ctx->prof.bytes[r.kernel] += b;
    }
    ctx->prof.pool.push_back(r.a);
    ctx->prof.pool.push_back(r.b);
  }
  p->recs.clear();
  return CTM_OK;
}

}  // namespace

extern "C" {

int ctm_plan_create(ctm_context *ctx, int n, const double *radius,
                    const double *axis_ratio, const double *wavelength,
                    const double *mr_re, const double *mr_im,
                    const ctm_convergence *conv,
                    const ctm_observable_spec *obs,
                    const ctm_plan_outputs *outputs, ctm_plan **out) {
  if (!out) return CTM_ERR_ARGUMENT;
  *out = nullptr;
  int rc = check_inputs(ctx, n, radius, axis_ratio, wavelength, mr_re, mr_im);
  if (rc != CTM_OK) return rc;
  if (n == 0 || !outputs) {
    ctx->err = "plan: no grains or no outputs";
    return CTM_ERR_ARGUMENT;
  }
  if (!conv || conv->min_order < 1 || conv->max_order <= conv->min_order + 1 ||
      conv->glfac < 1 || !(conv->tolerance > 0.) ||
      (conv->precision != CTM_PRECISION_FP64 &&
       conv->precision != CTM_PRECISION_MIXED)) {
    ctx->err = "invalid convergence parameters";
    return CTM_ERR_ARGUMENT;
  }
  if ((int)conv->max_order - 1 > kMaxLdsOrder) {
    ctx->err = "plan: the last order must be at most 99 (blocks above order "
               "99 are solved in global scratch; use ctm_tmatrix_batch)";
    return CTM_ERR_ARGUMENT;
  }
  if (obs && (obs->n_scatter_theta != 0 || obs->n_theta < 1 ||
              !obs->theta ||
              (obs->account_for_scattering && obs->ngauss < 1))) {
    ctx->err = "plan: observables are Q_ext / Q_abs at n_theta >= 1 angles "
               "(no scattering-matrix grid)";
    return CTM_ERR_ARGUMENT;
  }
  CTM_HIP(hipSetDevice(ctx->device));
  CTM_HIP(drain(ctx));
  std::unique_ptr<ctm_plan> p(new ctm_plan());
  p->ctx = ctx;
  p->n = n;
  p->gp = GrainParams{radius, axis_ratio, wavelength, mr_re, mr_im};
  p->min_order = (int)conv->min_order;
  p->last_order = (int)conv->max_order - 1;
  p->glfac = (int)conv->glfac;
  p->tol = conv->tolerance;
  p->mixed = conv->precision == CTM_PRECISION_MIXED;
  p->out = *outputs;
  {
    hipDeviceProp_t prop;
    CTM_HIP(hipGetDeviceProperties(&prop, ctx->device));
    p->cus = std::max(1, prop.multiProcessorCount);
  }
  for (hipEvent_t &e : p->ev_solve)
    CTM_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  CTM_HIP(hipEventCreateWithFlags(&p->ev_obs16, hipEventDisableTiming));
  const int lo_n = p->min_order, hi_n = p->last_order, glfac = p->glfac;
  // grain-independent tables of every order the plan can reach
  std::vector<GlKey> glk;
  std::vector<int64_t> gloff(hi_n + 1, 0);
  int64_t gl_size = 0;
  for (int N = lo_n; N <= hi_n; ++N) {
    GlKey k{};
    k.G = glfac * N;
    k.off = gl_size;
    gl_size += 2 * (int64_t)k.G;
    gloff[N] = k.off;
    glk.push_back(k);
  }
  std::vector<WigKey> wk;
  std::vector<int32_t> wbase(hi_n + 1, 0);
  int64_t wig_size = 0;
  for (int N = lo_n; N <= hi_n; ++N) {
    wbase[N] = (int32_t)wk.size();
    for (int m = 0; m <= N; ++m) {
      WigKey k{};
      k.N = N;
      k.m = m;
      k.G = glfac * N;
      k.gl_off = gloff[N];
      k.off = wig_size;
      wig_size += 2 * (int64_t)(N + 1 - (m > 0 ? m : 1)) * k.G;
      wk.push_back(k);
    }
  }
  ctm_plan *pp = p.get();
  auto up = [&](DevBuf &b, const void *src, size_t bytes) -> int {
    CTM_HIP(b.ensure(std::max<size_t>(bytes, 1)));
    CTM_HIP(hipMemcpy(b.p, src, bytes, hipMemcpyHostToDevice));
    return CTM_OK;
  };
  hipStream_t s = ctx->stream;
  if ((rc = up(pp->d_glkeys, glk.data(), glk.size() * sizeof(GlKey))) ||
      (rc = up(pp->d_wigkeys, wk.data(), wk.size() * sizeof(WigKey))) ||
      (rc = up(pp->d_gloff, gloff.data(), gloff.size() * sizeof(int64_t))) ||
      (rc = up(pp->d_wigbase, wbase.data(), wbase.size() * sizeof(int32_t))))
    return rc;
  CTM_HIP(pp->d_gl.ensure(gl_size * sizeof(double)));
  CTM_HIP(pp->d_wig.ensure(wig_size * sizeof(double)));
  CTM_HIP(launch_gauss_legendre(pp->d_glkeys.as<GlKey>(), (int)glk.size(),
                                pp->d_gl.as<double>(), s));
  CTM_HIP(launch_wigner(pp->d_wigkeys.as<WigKey>(), (int)wk.size(),
                        pp->d_gl.as<double>(), pp->d_wig.as<double>(), s));
  // observables: the direction layout and the pi/tau table up to the last
  // order (both fixed for the plan)
  if (obs) {
    pp->want_x = true;
    obs_layout(obs, pp->lay);
    if (obs_x_lds_bytes(hi_n, pp->lay.sp.n_out, 8, 256, 1, true) >
        kObsXMaxLds) {
      ctx->err = "plan: too many outgoing directions for the last order";
      return CTM_ERR_ARGUMENT;
    }
    if ((rc = up(pp->d_spec, pp->lay.blob.data(), pp->lay.blob.size())))
      return rc;
    pp->lay.bind(pp->d_spec.as<char>());
    const int n_ang = pp->lay.n_ang;
    CTM_HIP(pp->d_pt.ensure((size_t)(hi_n + 1) * n_ang * hi_n * 2 *
                            sizeof(double)));
    CTM_HIP(launch_pitau(
        reinterpret_cast<const double *>(pp->d_spec.as<char>() +
                                         pp->lay.o_ang),
        n_ang, hi_n, pp->d_pt.as<double>(), s));
    if (!obs->do_extinction) {
      pp->out.obs_qext = pp->out.obs_qextpol = pp->out.obs_qextcpol = nullptr;
    }
    if (!obs->do_absorption) {
      pp->out.obs_qabs = pp->out.obs_qabspol = nullptr;
    }
  } else {
    pp->out.obs_qext = pp->out.obs_qextpol = pp->out.obs_qextcpol = nullptr;
    pp->out.obs_qabs = pp->out.obs_qabspol = nullptr;
  }
  // lists (fixed capacity: every grain at every order)
  const int64_t cap = (int64_t)n * (hi_n + 1);
  pp->caps.items = cap;
  const int gmax = glfac * hi_n;
  CTM_HIP(pp->d_sets.ensure(cap * sizeof(TableSet)));
  CTM_HIP(pp->d_items.ensure(cap * sizeof(Item)));
  CTM_HIP(pp->d_ord.ensure(kPlanParts * cap * sizeof(int32_t)));
  CTM_HIP(pp->d_pre.ensure(3 * (cap + 1) * sizeof(int32_t)));
  CTM_HIP(pp->d_sol.ensure(kPlanBuckets * 2 * cap * sizeof(int32_t)));
  CTM_HIP(pp->d_obsl.ensure(kPlanObsBuckets * cap * sizeof(int32_t)));
  CTM_HIP(pp->d_flags.ensure(kPlanBuckets * 2 * cap * sizeof(int32_t)));
  CTM_HIP(pp->d_fb.ensure(kPlanBuckets * 2 * cap * sizeof(int32_t)));
  CTM_HIP(pp->d_hist.ensure(plan_hist_ints(gmax, hi_n) * sizeof(int32_t)));
  CTM_HIP(pp->d_cnt.ensure((size_t)n * 6 * sizeof(int64_t)));
  CTM_HIP(pp->d_goff.ensure((size_t)(n + 1) * 6 * sizeof(int64_t)));
  CTM_HIP(pp->d_gitem.ensure((size_t)(n + 1) * sizeof(int32_t)));
  CTM_HIP(pp->d_gobs.ensure((size_t)(n + 1) * sizeof(int32_t)));
  CTM_HIP(pp->d_lo.ensure((size_t)n * sizeof(int32_t)));
  CTM_HIP(pp->d_hi.ensure((size_t)n * sizeof(int32_t)));
  CTM_HIP(pp->d_state.ensure((size_t)n * sizeof(ConvState)));
  CTM_HIP(pp->d_phase.ensure(kPlanPhases * sizeof(PlanPhase)));
  CTM_HIP(pp->d_status.ensure(sizeof(PlanStatus)));
  CTM_HIP(pp->d_iota.ensure(cap * sizeof(int32_t)));
  CTM_HIP(pp->d_qparts.ensure(cap * 2 * kQParts * sizeof(double)));
  CTM_HIP(launch_plan_iota(pp->d_iota.as<int32_t>(), cap, s));
  // first pools from the first windows (host restatement of the device
  // rule; a wrong guess only costs one more run)
  {
    std::vector<double> h[4];
    const double *src[4] = {radius, axis_ratio, wavelength, mr_re};
    for (int a = 0; a < 4; ++a) {
      h[a].resize(n);
      CTM_HIP(hipMemcpy(h[a].data(), src[a], n * sizeof(double),
                        hipMemcpyDeviceToHost));
    }
    auto s1 = [](int64_t k) { return k > 0 ? k * (k + 1) / 2 : 0; };
    auto s2 = [](int64_t k) { return k > 0 ? k * (k + 1) * (2 * k + 1) / 6 : 0; };
    int64_t tab_w = 0, q_w = 0, tab_a = 0, q_a = 0, items_a = 0;
    std::vector<int> hi1(n);
    for (int g = 0; g < n; ++g) {
      const int end = first_window_end(h[0][g], h[1][g], h[2][g], h[3][g]);
      const int hi = std::min(hi_n, std::max(lo_n + 1, end));
      hi1[g] = hi;
      tab_w += glfac * ((int64_t)kNodeScalars * (s1(hi) - s1(lo_n - 1)) +
                        (int64_t)kBesselArrays * (s2(hi) - s2(lo_n - 1)));
      q_w += 4 * (s2(hi) - s2(lo_n - 1));
      const int64_t N = hi, G = glfac * N;
      tab_a += (int64_t)kNodeScalars * G + (int64_t)kBesselArrays * N * G;
      q_a += 4 * (N * N + s2(N));
      items_a += N + 1;
    }
    auto pad = [](int64_t v) { return v + v / 8 + (1 << 20); };
    pp->caps.tables = pad(std::max(tab_w, tab_a));
    pp->caps.q = pad(std::max(q_w, q_a));
    pp->caps.t = pad(q_a / 2);
    pp->caps.x = pp->want_x
                     ? pad(items_a * (int64_t)pp->lay.sp.n_pairs * 4)
                     : 0;
    pp->hi1 = std::move(hi1);
  }
  PlanLists &L = pp->L;
  L.sets = pp->d_sets.as<TableSet>();
  L.items = pp->d_items.as<Item>();
  L.ord = pp->d_ord.as<int32_t>();
  L.pre = pp->d_pre.as<int32_t>();
  L.sol = pp->d_sol.as<int32_t>();
  L.obsl = pp->d_obsl.as<int32_t>();
  L.flags = pp->d_flags.as<int32_t>();
  L.fb = pp->d_fb.as<int32_t>();
  L.hist = pp->d_hist.as<int32_t>();
  L.cnt = pp->d_cnt.as<int64_t>();
  L.goff = pp->d_goff.as<int64_t>();
  L.gitem = pp->d_gitem.as<int32_t>();
  L.gobs = pp->d_gobs.as<int32_t>();
  L.lo = pp->d_lo.as<int32_t>();
  L.hi = pp->d_hi.as<int32_t>();
  L.state = pp->d_state.as<ConvState>();
  L.gl_off = pp->d_gloff.as<int64_t>();
  L.wig_base = pp->d_wigbase.as<int32_t>();
  L.phase = pp->d_phase.as<PlanPhase>();
  L.status = pp->d_status.as<PlanStatus>();
  L.n = n;
  L.glfac = glfac;
  L.min_order = lo_n;
  L.last_order = hi_n;
  L.gmax = gmax;
  L.n_in = pp->want_x ? pp->lay.sp.n_in : 0;
  L.n_pairs = pp->want_x ? pp->lay.sp.n_pairs : 0;
  L.win_scale = kWindowScale;
  L.win_add = kWindowAdd;
  L.micro_ratio = kMicroRatio;
  L.big_first = 0;
  L.obs_lean = 1;
  plan_first_grids(pp, pp->hi1);
  if ((rc = plan_alloc(pp)) != CTM_OK) return rc;
  CTM_HIP(hipStreamSynchronize(s));
  ctx->plans += 1;
  *out = p.release();
  return CTM_OK;
}

int ctm_plan_run(ctm_plan *p) {
  if (!p) return CTM_ERR_ARGUMENT;
  ctm_context *ctx = p->ctx;
  CTM_HIP(hipSetDevice(ctx->device));
  ctx->err.clear();
  for (int attempt = 0; attempt < 6; ++attempt) {
    int rc = plan_enqueue(p);
    if (rc == CTM_OK) rc = plan_finish(p);
    if (rc != CTM_OK) return rc;
    p->runs += 1;
    if (p->st.overflow == 0) {
      // the last batch's block counts, as after ctm_tmatrix_batch
      ctx->mixed_blocks = p->st.n_mixed;
      ctx->mixed_fallback = p->st.n_fallback;
      if (!p->gcap_fitted) {
        // grids fitted to the lists of this batch (a later batch whose
        // lists outgrow them is run again with larger grids)
        for (int ph = 0; ph < kPlanPhases; ++ph)
          for (int k = 0; k < kPlanLists; ++k)
            p->gcap[ph][k] = plan_grid_of(p->st.count[ph][k]);
        p->gcap_fitted = true;
        rc = plan_alloc(p);
        if (rc != CTM_OK) return rc;
      }
      return CTM_OK;
    }
    if (p->st.overflow & kPlanOverflowGrid) {
      for (int ph = 0; ph < kPlanPhases; ++ph)
        for (int k = 0; k < kPlanLists; ++k)
          if (p->st.count[ph][k] > p->gcap[ph][k])
            p->gcap[ph][k] = plan_grid_of(p->st.count[ph][k]);
    }
    // grow the pools to what the device reported and run again
    if (p->st.overflow & kPlanOverflowNodes) {
      ctx->err = "plan: more than 2^31 quadrature nodes in one phase";
      return CTM_ERR_MEMORY;
    }
    auto grow = [](int64_t cur, int64_t need) {
      return need > cur ? need + need / 4 : cur;
    };
    PlanCaps c = p->caps;
    c.tables = grow(c.tables, p->st.need_tables);
    c.q = grow(c.q, p->st.need_q);
    c.t = grow(c.t, p->st.need_t);
    c.x = grow(c.x, p->st.need_x);
    if (p->grow_limit) {
      const size_t bytes =
          (size_t)c.tables * sizeof(double) + (size_t)c.q * sizeof(cdbl) +
          (size_t)c.t * sizeof(cdbl) + (size_t)c.x * sizeof(cdbl) +
          (p->mixed ? (size_t)c.t * (sizeof(cflt) + sizeof(int32_t) +
                                     sizeof(cdbl))
                    : 0);
      if (bytes > p->grow_limit) {
        ctx->err = "plan: pools would outgrow the tuning limit";
        return CTM_ERR_MEMORY;
      }
    }
    p->caps = c;
    rc = plan_alloc(p);
    if (rc != CTM_OK) return rc;
    p->reruns += 1;
  }
  ctx->err = "plan: the memory pools did not settle";
  return CTM_ERR_MEMORY;
}

// Autotuner: whole runs of the plan under candidate values of its
// tunables, by coordinate descent (first-window margin added, its scale,
// the tile-shape cost ratio), each candidate timed as the median wall time
// of `runs` runs after one untimed run (which also absorbs any pool or grid
// growth); the fastest values are kept.  The tunables change only how the
// work is split into rounds and tiles: every candidate computes the same
// orders and values.
int ctm_plan_tuning_set(ctm_plan *p, const ctm_plan_tuning *t) {
  if (!p || !t) return CTM_ERR_ARGUMENT;
  ctm_context *ctx = p->ctx;
  if (!(t->window_scale >= 0.5 && t->window_scale <= 2.) ||
      !(t->window_add >= 0. && t->window_add <= 50.) ||
      !(t->micro_ratio > 0. && t->micro_ratio <= 4.)) {
    ctx->err = "plan tuning out of range (window_scale 0.5-2, window_add "
               "0-50, micro_ratio 0-4)";
    return CTM_ERR_ARGUMENT;
  }
  CTM_HIP(hipSetDevice(ctx->device));
  p->L.win_scale = t->window_scale;
  p->L.win_add = t->window_add;
  p->L.micro_ratio = t->micro_ratio;
  p->L.big_first = std::min(std::max(t->big_first, 0), 2);
  p->L.obs_lean = t->obs_lean != 0;
  // grids fitted again to the lists of the next run
  p->gcap_fitted = false;
  return plan_alloc(p);
}

int ctm_plan_tune(ctm_plan *p, int runs, ctm_plan_tuning *out) {
  if (!p || runs < 1) return CTM_ERR_ARGUMENT;
  ctm_context *ctx = p->ctx;
  const bool prof = ctx->prof.on;
  ctx->prof.on = false;
  auto measure = [&](double *ms) -> int {
    int rc = ctm_plan_run(p);
    if (rc != CTM_OK) return rc;
    std::vector<double> t;
    for (int r = 0; r < runs; ++r) {
      const auto t0 = std::chrono::steady_clock::now();
      rc = ctm_plan_run(p);
      if (rc != CTM_OK) return rc;
      t.push_back(std::chrono::duration<double, std::milli>(
                      std::chrono::steady_clock::now() - t0)
                      .count());
    }
    std::sort(t.begin(), t.end());
    *ms = t[t.size() / 2];
    return CTM_OK;
  };
  double big = p->L.big_first, lean = p->L.obs_lean;
  constexpr int kKnobs = 5;
  double *knob[kKnobs] = {&p->L.win_add, &p->L.win_scale,
                          &p->L.micro_ratio, &big, &lean};
  const std::vector<double> cand[kKnobs] = {{1., 2., 3., 4., 5.},
                                            {1.0, 1.04, 1.08},
                                            {0.35, 0.5, 0.65, 0.8},
                                            {0., 1., 2.},
                                            {0., 1.}};
  auto flags = [&]() {
    p->L.big_first = (int)big;
    p->L.obs_lean = lean != 0.;
    return CTM_OK;
  };
  ctm_plan_tuning res{};
  double best = 0.;
  int rc = measure(&best);
  if (rc != CTM_OK) {
    ctx->prof.on = prof;
    return rc;
  }
  // a candidate whose pools would grow past four times those of the
  // starting values (or past three quarters of the free device memory)
  // counts as too slow instead of taking the device's memory
  {
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) {
      ctx->err = "plan tuning: hipMemGetInfo failed";
      ctx->prof.on = prof;
      return CTM_ERR_HIP;
    }
    const size_t now = p->d_tables.cap + p->d_q.cap + p->d_t.cap +
                       p->d_x.cap + p->d_x32.cap + p->d_xexp.cap +
                       p->d_rbuf.cap;
    p->grow_limit = std::max<size_t>(
        now + 1, std::min<size_t>(4 * now, now + free_b / 4 * 3));
  }
  res.ms_default = best;
  res.trials = 1;
  for (int k = 0; k < kKnobs; ++k) {
    const double keep = *knob[k];
    double pick = keep;
    for (double v : cand[k]) {
      if (v == keep) continue;
      *knob[k] = v;
      rc = flags();
      double ms = 0.;
      if (rc == CTM_OK) rc = measure(&ms);
      if (rc == CTM_ERR_MEMORY && p->st.overflow != 0) {
        // too much memory for this candidate: rejected
        res.trials += 1;
        continue;
      }
      if (rc != CTM_OK) {
        *knob[k] = keep;
        (void)flags();
        p->grow_limit = 0;
        ctx->prof.on = prof;
        return rc;
      }
      res.trials += 1;
      if (ms < best) {
        best = ms;
        pick = v;
      }
    }
    *knob[k] = pick;
    rc = flags();
    if (rc != CTM_OK) {
      ctx->prof.on = prof;
      return rc;
    }
  }
  // grids fitted to the chosen configuration at the next run
  p->grow_limit = 0;
  p->gcap_fitted = false;
  rc = ctm_plan_run(p);
  if (rc == CTM_OK) {
    // pools back to what the chosen configuration needs (candidates with
    // smaller first windows send more grains to the later rounds and grew
    // them); a later batch that needs more grows them again
    auto fit = [](int64_t need) { return need + need / 4 + (1 << 20); };
    PlanCaps &c = p->caps;
    c.tables = std::min(c.tables, fit(p->st.need_tables));
    c.q = std::min(c.q, fit(p->st.need_q));
    c.t = std::min(c.t, fit(p->st.need_t));
    c.x = std::min(c.x, fit(p->st.need_x));
    DevBuf *bufs[] = {&p->d_tables, &p->d_q, &p->d_t, &p->d_x,
                      &p->d_x32, &p->d_xexp, &p->d_rbuf};
    const size_t want[] = {(size_t)c.tables * sizeof(double),
                           (size_t)c.q * sizeof(cdbl),
                           (size_t)c.t * sizeof(cdbl),
                           (size_t)c.x * sizeof(cdbl),
                           (size_t)c.t * sizeof(cflt),
                           (size_t)c.t * sizeof(int32_t),
                           (size_t)c.t * sizeof(cdbl)};
    for (int k = 0; k < 7 && rc == CTM_OK; ++k)
      if (bufs[k]->p && bufs[k]->fit(want[k]) != hipSuccess) {
        ctx->err = "plan: pool reallocation failed";
        rc = CTM_ERR_HIP;
      }
    if (rc == CTM_OK) rc = plan_alloc(p);
  }
  ctx->prof.on = prof;
  if (rc != CTM_OK) return rc;
  res.window_add = p->L.win_add;
  res.window_scale = p->L.win_scale;
  res.micro_ratio = p->L.micro_ratio;
  res.big_first = p->L.big_first;
  res.obs_lean = p->L.obs_lean;
  res.ms_chosen = best;
  res.runs_per_trial = runs;
  if (out) *out = res;
  return CTM_OK;
}

int ctm_plan_info_get(ctm_plan *p, ctm_plan_info *info) {
  if (!p || !info) return CTM_ERR_ARGUMENT;
  ctm_plan_info i{};
  i.runs = p->runs;
  i.reruns = p->reruns;
  i.mixed_blocks = p->st.n_mixed;
  i.mixed_fallback = p->st.n_fallback;
  i.pool_bytes = (int64_t)(p->d_tables.cap + p->d_q.cap + p->d_t.cap +
                           p->d_x.cap + p->d_x32.cap + p->d_xexp.cap +
                           p->d_rbuf.cap);
  for (int r = 0; r < 3; ++r) i.window_items[r] = p->ph[r].n_items;
  i.allm_items = p->ph[3].n_items;
  *info = i;
  return CTM_OK;
}

int ctm_plan_destroy(ctm_plan *p) {
  if (!p) return CTM_OK;
  ctm_context *ctx = p->ctx;
  (void)hipSetDevice(ctx->device);
  (void)drain(ctx);
  DevBuf *bufs[] = {&p->d_gl, &p->d_glkeys, &p->d_wig, &p->d_wigkeys,
                    &p->d_gloff, &p->d_wigbase, &p->d_spec, &p->d_pt,
                    &p->d_sets, &p->d_items, &p->d_ord, &p->d_pre,
                    &p->d_sol, &p->d_obsl, &p->d_flags, &p->d_fb,
                    &p->d_hist, &p->d_cnt, &p->d_goff, &p->d_gitem,
                    &p->d_gobs, &p->d_lo, &p->d_hi, &p->d_state,
                    &p->d_phase, &p->d_status, &p->d_iota, &p->d_tables,
                    &p->d_q, &p->d_t, &p->d_x, &p->d_qparts, &p->d_x32,
                    &p->d_xexp, &p->d_rbuf, &p->d_gcap};
  for (DevBuf *b : bufs) b->release();
  for (hipEvent_t e : p->ev_solve)
    if (e) (void)hipEventDestroy(e);
  if (p->ev_obs16) (void)hipEventDestroy(p->ev_obs16);
  for (const ctm_plan::Rec &r : p->recs) {
    ctx->prof.pool.push_back(r.a);
    ctx->prof.pool.push_back(r.b);
  }
  ctx->plans -= 1;
  delete p;
  return CTM_OK;
}

}  // extern "C"
