// C ABI of the XC Fock build (include/serenity_xc.h): resident handles, the
// screening/AO setup of a basis-functions-on-grid controller, and the
// density -> functional -> Fock pipeline on one HIP stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <map>
#include <mutex>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/serenity_xc.h"
#include "host/host.h"
#include "kernels/xc_internal.h"

using namespace sxc;

namespace {

/* Work items of one block are consecutive in a list.  Workgroups of a launch are
 * dealt to the 8 XCDs round-robin (workgroup w runs on XCD w mod 8, from some
 * starting XCD), each with its own L2.  Cut the list into 8 contiguous ranges of
 * whole blocks (about a fair share of items each; the blocks are in grid order,
 * so a range is a region of the molecule) and interleave them: one XCD runs a
 * block's items one after another, its re-reads of the block's tiles hit its L2,
 * and the density-matrix rows its region reads stay in that L2 too (density stage
 * at (H2O)64/def2-TZVP: 4.46 -> 4.36 ms FP32 against dealing blocks to the
 * shortest queue). */
template <typename Item>
void xcd_order(std::vector<Item>& v) {
  constexpr int kXcd = 8;
  std::vector<std::vector<Item>> q(kXcd);
  const size_t target = (v.size() + kXcd - 1) / kXcd;
  size_t x = 0;
  for (size_t i = 0; i < v.size();) {
    size_t j = i;
    while (j < v.size() && v[j].block == v[i].block) ++j;
    if (q[x].size() >= target && x + 1 < q.size()) ++x;
    q[x].insert(q[x].end(), v.begin() + i, v.begin() + j);
    i = j;
  }
  size_t len = 0;
  for (const auto& u : q) len = std::max(len, u.size());
  v.clear();
  for (size_t k = 0; k < len; ++k)
    for (const auto& u : q)
      if (k < u.size()) v.push_back(u[k]);
}

thread_local std::string g_last_error;

struct XcError : std::runtime_error {
  int code;
  XcError(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define HIPX(expr)                                                                                    \
  do {                                                                                                \
    hipError_t e_ = (expr);                                                                           \
    if (e_ != hipSuccess)                                                                             \
      throw XcError(XC_ERR_DEVICE, std::string(#expr) + " failed: " + hipGetErrorString(e_));         \
  } while (0)

void require(bool cond, int code, const std::string& msg) {
  if (!cond) throw XcError(code, msg);
}

template <typename F>
int guarded(F&& f) {
  try {
    f();
    return XC_OK;
  } catch (const XcError& e) {
    g_last_error = e.what();
    return e.code;
  } catch (const std::exception& e) {
    g_last_error = e.what();
    return XC_ERR_INPUT;
  } catch (...) {
    g_last_error = "unknown error";
    return XC_ERR_STATE;
  }
}

}  // namespace

// Workspace of xc_functional_device_noalloc: the per-block energies and their sum live in this
// file-scope device array instead of a per-call allocation.  Each device holds its own copy;
// host code never uses the array's name as an address, only the pointer that
// hipGetSymbolAddress returns for the current device (kept per device in g_noalloc_ptr).
constexpr int kNoallocBlocks = 131072;  // 1 MiB of per-block energies + the sum
static __device__ __align__(256) double scratch[kNoallocBlocks + 1];

namespace {
std::mutex g_noalloc_mu;
std::map<int, double*> g_noalloc_ptr;  // device -> address of that device's scratch

// Called once per device with that device current (context creation).
void noalloc_setup_current_device(int device) {
  std::lock_guard<std::mutex> lk(g_noalloc_mu);
  if (g_noalloc_ptr.count(device)) return;
  void* p = nullptr;
  HIPX(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  require(p != nullptr, XC_ERR_DEVICE, "hipGetSymbolAddress(scratch) returned no address");
  g_noalloc_ptr[device] = static_cast<double*>(p);
}

double* noalloc_scratch(int device) {
  std::lock_guard<std::mutex> lk(g_noalloc_mu);
  auto it = g_noalloc_ptr.find(device);
  require(it != g_noalloc_ptr.end(), XC_ERR_STATE, "no-allocation workspace not set up on this device");
  return it->second;
}
}  // namespace

// Execution lanes of a build: its stream plus kAux helper streams.  The launches
// of one stage (one per work class) are independent, so they are spread over
// the lanes after a fork and joined before the next stage.
constexpr int kAux = 3;
struct Lanes {
  int n = 1;  // lanes in use: 1 (the build's stream only) or kAux + 1
  hipStream_t main = nullptr;
  hipStream_t aux[kAux] = {};
  hipEvent_t fork = nullptr;
  hipEvent_t join[kAux] = {};
  void create(hipStream_t m, bool helpers) {
    main = m;
    n = helpers ? kAux + 1 : 1;
    if (!helpers) return;
    for (int i = 0; i < kAux; ++i) {
      HIPX(hipStreamCreateWithFlags(&aux[i], hipStreamNonBlocking));
      HIPX(hipEventCreateWithFlags(&join[i], hipEventDisableTiming));
    }
    HIPX(hipEventCreateWithFlags(&fork, hipEventDisableTiming));
  }
  void destroy() {
    for (int i = 0; i < kAux; ++i) {
      if (aux[i]) (void)hipStreamDestroy(aux[i]);
      if (join[i]) (void)hipEventDestroy(join[i]);
      aux[i] = nullptr;
      join[i] = nullptr;
    }
    if (fork) (void)hipEventDestroy(fork);
    fork = nullptr;
  }
  hipStream_t lane(int i) const { return i % (kAux + 1) == 0 ? main : aux[i % (kAux + 1) - 1]; }
  void fork_all() const {
    HIPX(hipEventRecord(fork, main));
    for (int i = 0; i < kAux; ++i) HIPX(hipStreamWaitEvent(aux[i], fork, 0));
  }
  void join_all() const {
    for (int i = 0; i < kAux; ++i) {
      HIPX(hipEventRecord(join[i], aux[i]));
      HIPX(hipStreamWaitEvent(main, join[i], 0));
    }
  }
};

struct PackWork;

struct xc_ctx {
  int device = 0;
  PackWork* pack = nullptr;  // device lists of packed builds (allocated on first use)
  // packed-build compositions seen before (controllers, functional, precision, threshold):
  // their own lists and, from the second build on, a captured graph
  std::map<std::vector<uint64_t>, PackWork*> pack_graphs;
  uint64_t pack_calls = 0;  // packed builds so far (recency of the cached compositions)
  hipStream_t stream = nullptr;
  Lanes lanes;  // helper lanes of builds on the context stream
  std::recursive_mutex mu;
  int64_t bytes = 0;
  double last_energy = 0.0;
  double* d_energy = nullptr;
  XcParams prm;
  std::map<void*, size_t> user_allocs;  // buffers handed out by xc_device_alloc
  int live_handles = 0;                 // bases, grids and controllers not yet destroyed
  // stage timing: XC_NSTAGES + 1 events per profiled build
  bool profile = false;
  // every row reduce sums in 64-bit integers (F the same bit for bit on every run); set while no
  // basis, grid or controller exists
  bool deterministic = false;
  std::vector<hipEvent_t> ev_pool;
  size_t ev_used = 0;
  hipEvent_t next_event() {
    if (ev_used == ev_pool.size()) {
      hipEvent_t e;
      HIPX(hipEventCreate(&e));
      ev_pool.push_back(e);
    }
    return ev_pool[ev_used++];
  }
  void mark() {
    if (profile) HIPX(hipEventRecord(next_event(), stream));
  }

  // Device blocks are kept for reuse when released: a serving loop builds and evicts controllers
  // again and again, and hipMalloc / hipFree of their multi-GB buffers cost milliseconds each and
  // now and then stall.  Blocks are sized in 2 MiB steps (256 bytes below 1 MiB); an allocation
  // takes the smallest cached block that fits within 1/8 + 2 MiB of its size, else a new one.
  // The cache is given back to the device when a hipMalloc fails and when the context ends.
  // `bytes` counts the blocks in use.
  std::multimap<size_t, void*> cache;            // block size -> released block
  std::unordered_map<void*, size_t> block_size;  // every block of this context -> size
  size_t cached = 0;
  static size_t block_round(size_t b) {
    constexpr size_t kBig = size_t(1) << 20, kStep = size_t(2) << 20;
    return b >= kBig ? (b + kStep - 1) / kStep * kStep : (b + 255) / 256 * 256;
  }
  void trim() {
    (void)hipDeviceSynchronize();
    for (auto& kv : cache) {
      (void)hipFree(kv.second);
      block_size.erase(kv.second);
    }
    cache.clear();
    cached = 0;
  }
  void* alloc(size_t bytes_) {
    if (bytes_ == 0) return nullptr;
    HIPX(hipSetDevice(device));
    const size_t need = block_round(bytes_);
    auto it = cache.lower_bound(need);
    if (it != cache.end() && it->first <= need + need / 8 + (size_t(2) << 20)) {
      void* p = it->second;
      const size_t sz = it->first;
      cache.erase(it);
      cached -= sz;
      bytes += (int64_t)sz;
      return p;
    }
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, need);
    if (e != hipSuccess && cached) {
      (void)hipGetLastError();
      trim();
      e = hipMalloc(&p, need);
    }
    if (e != hipSuccess) {
      (void)hipGetLastError();
      throw XcError(XC_ERR_MEMORY, "hipMalloc of " + std::to_string(bytes_) + " bytes failed");
    }
    block_size[p] = need;
    bytes += (int64_t)need;
    return p;
  }
  void release(void* p, size_t) {
    if (!p) return;
    auto it = block_size.find(p);
    if (it == block_size.end()) {
      (void)hipFree(p);
      return;
    }
    // enqueued work may still use the block: wait for the device, as hipFree does
    (void)hipDeviceSynchronize();
    bytes -= (int64_t)it->second;
    cache.emplace(it->second, p);
    cached += it->second;
  }
};

namespace {

// Device allocation owned by a handle.
template <typename T>
struct DBuf {
  T* p = nullptr;
  size_t n = 0;
  xc_ctx* ctx = nullptr;
  void alloc(xc_ctx* c, size_t count) {
    free_();
    ctx = c;
    n = count;
    p = (T*)c->alloc(count * sizeof(T));
  }
  void upload(xc_ctx* c, const T* host, size_t count) {
    alloc(c, count);
    if (count) HIPX(hipMemcpyAsync(p, host, count * sizeof(T), hipMemcpyHostToDevice, c->stream));
  }
  void free_() {
    if (p && ctx) ctx->release(p, n * sizeof(T));
    p = nullptr;
    n = 0;
  }
  ~DBuf() { free_(); }
};


}  // namespace

// Merged lists and per-system table of packed builds, grown on demand; a packed build
// synchronises before it returns, so they are free again for the next one.
struct PackWork {
  DBuf<DevBlock> blocks;
  DBuf<PackSys> sys;
  DBuf<WorkItem> dens[kDensMaxS + 1];
  DBuf<FockItem> fock[kFockNumS];
  DBuf<double> round, E;
  hipGraphExec_t exec = nullptr;  // captured build of a repeated composition
  uint64_t last_use = 0;          // packed-build counter at its last use (least recently used is evicted)
  ~PackWork() {
    if (exec) (void)hipGraphExecDestroy(exec);
  }
};

// drop every cached packed composition that names bf (its buffers are going away)
void forget_packs_of(xc_ctx* ctx, const void* bf) {
  for (auto it = ctx->pack_graphs.begin(); it != ctx->pack_graphs.end();) {
    bool uses = false;
    for (size_t i = 0; i + 3 < it->first.size(); ++i) uses = uses || it->first[i] == (uint64_t)(uintptr_t)bf;
    if (uses) {
      delete it->second;
      it = ctx->pack_graphs.erase(it);
    } else {
      ++it;
    }
  }
}

template <typename T>
void upload_grow(xc_ctx* ctx, DBuf<T>& b, const std::vector<T>& v) {
  if (v.empty()) return;
  if (b.n < v.size()) b.alloc(ctx, std::max(v.size(), 2 * b.n));
  HIPX(hipMemcpyAsync(b.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, ctx->stream));
}

struct xc_basis {
  xc_ctx* ctx = nullptr;
  int nshell = 0, nao = 0, maxl = 0;
  std::vector<DevShell> shells;
  std::vector<double> alpha, coef, ang;
  DBuf<DevShell> d_shells;
  DBuf<double> d_alpha, d_coef, d_ang;
};

// Density-fitted Coulomb matrix of an orbital basis over an auxiliary basis (the J engine).
struct xc_coulomb {
  xc_ctx* ctx = nullptr;
  xc_basis* basis = nullptr;
  xc_basis* aux = nullptr;
  std::vector<JPair> pairs;
  int64_t herm_total = 0;
  int ncart_aux = 0;
  std::vector<int> cart_off;  // first Cartesian function of every auxiliary shell
  DBuf<JPair> d_pairs;
  DBuf<int> d_cart_off;
  DBuf<double> d_D, d_V, d_bc, d_cc, d_P, d_J;
};

struct xc_grid {
  xc_ctx* ctx = nullptr;
  int64_t npts = 0;
  DBuf<double> d_pts, d_w;
};

struct xc_bfogc {
  xc_ctx* ctx = nullptr;
  xc_basis* basis = nullptr;
  xc_grid* grid = nullptr;
  int bs = 128;
  double radial_thr = 1e-9;
  int deriv = 1, ncomp = 4, nblocks = 0;
  std::vector<DevBlock> blocks;
  std::vector<int> ao_idx;
  int64_t tile_elems = 0;
  int64_t sum_nb = 0;
  double sum_B_nb2 = 0.0;
  DBuf<DevBlock> d_blocks;
  DBuf<int> d_ao_idx, d_sig_shell, d_sig_pos;
  DBuf<double> d_t64;
  DBuf<float> d_t32;
  // contraction work lists per size class
  int npass_max = 0;
  std::vector<WorkItem> dens_work[kDensMaxS + 1];        // by column tiles per pass
  std::vector<FockItem> fock_items[2][kFockNumS];       // sub-tile Fock kernel, [fp64][slot class]
  DBuf<WorkItem> d_dens_work[kDensMaxS + 1];
  // mixed precision: every block's density items contiguously (from dens_item0[b], class
  // dens_cls[b]); the device lists of the dense blocks refined in FP64 and their counts; the
  // FP32-rounded density matrix in FP64
  std::vector<WorkItem> dens_by_block;
  std::vector<int> dens_item0, dens_cls;
  DBuf<WorkItem> d_dens_by_block, d_dense_list[kDensMaxS + 1], d_carry_list[kDensMaxS + 1];
  DBuf<int> d_dens_item0, d_dens_cls, d_dense_cnt, d_carry_cnt;
  // mixed precision: the blocks contracted in FP64 by the last mixed build that were still dense
  // (one byte per block); the next mixed build's FP32 pass skips them
  DBuf<unsigned char> d_dense_mask;
  DBuf<double> d_P32d;
  DBuf<FockItem> d_fock_items[2][kFockNumS];
  // per-build work buffers
  DBuf<double> d_part, d_rho, d_grad, d_vs, d_gw, d_eblk, d_U, d_P, d_F, d_v, d_g, d_E;
  // Fock sub-tile scratch (FP64 size when FP64 tiles are kept, else FP32) and the row entries of
  // the reduce (sorted by AO row, then block)
  int64_t t_elems = 0;
  DBuf<double> d_T;
  DBuf<RowEnt> d_rowent;
  DBuf<int> d_rowptr;
  DBuf<float> d_P32;
  DBuf<double> d_round;  // finalize-block partials of the FP32 rounding term, per spin
  // supersystem density of a non-additive build with this controller active (allocated on first use)
  DBuf<double> d_rho_sup, d_grad_sup, d_rho_sup_b, d_grad_sup_b;
  // cached grid density (xc_bfogc_update_density) and the energies of it per functional
  DBuf<double> d_rho_c, d_grad_c, d_ecache;
  bool rho_c_valid = false, rho_c_grad = false;
  // nuclear gradient (allocated on first use): work items, per-AO partials, atom map, result
  std::vector<GradItem> grad_work;
  DBuf<GradItem> d_grad_work;
  DBuf<double> d_gacc, d_gout;
  DBuf<int> d_ao_atom;
  // unrestricted builds (allocated on first use): beta density, potential, P and F
  DBuf<double> d_rho_b, d_grad_b, d_vs_b, d_gw_b, d_Pb, d_Fb;
  DBuf<float> d_P32b;
  bool ecache_valid[kNumFunctionals] = {};
  // serving: a build of this controller is in flight on a server slot; captured builds
  bool in_flight = false;
  std::map<std::pair<int, double>, hipGraphExec_t> graphs;  // (functional + 16 precision, threshold)
  ~xc_bfogc() {
    for (auto& kv : graphs) (void)hipGraphExecDestroy(kv.second);
  }
};

namespace {

// Launch order and lanes of a stage's work classes: the non-empty classes, the
// largest first, dealt round-robin over the lanes (fork / join only when more
// than one class has work).
template <typename F>
void over_lanes(const Lanes& L, int nclass, const std::function<int(int)>& count, F&& launch) {
  int nz = 0;
  for (int c = nclass - 1; c >= 0; --c) nz += count(c) > 0 ? 1 : 0;
  const bool par = nz > 1 && L.n > 1;
  if (par) L.fork_all();
  int k = 0;
  for (int c = nclass - 1; c >= 0; --c)
    if (count(c) > 0) launch(c, par ? L.lane(k++) : L.main);
  if (par) L.join_all();
}

template <typename T>
void run_density(xc_bfogc* bf, const T* tiles, const T* P, int want_grad, const Lanes& L) {
  over_lanes(L, kDensMaxS + 1, [&](int c) { return c >= 1 ? (int)bf->dens_work[c].size() : 0; },
             [&](int c, hipStream_t s) {
    launch_density_sub<T>(s, c, bf->d_dens_work[c].p, (int)bf->dens_work[c].size(), bf->d_blocks.p,
                      bf->d_ao_idx.p, tiles, P, bf->basis->nao, want_grad, bf->d_part.p, bf->grid->npts);
  });
}

// Mixed precision: the FP32 density pass, and FP64 density items (FP64 tiles, the FP32-rounded P
// in FP64) for the blocks whose largest density exceeds kMixedRhoThreshold.  The controller's mask
// holds the blocks the last mixed build found dense: their FP64 items run beside the FP32 pass over
// the other blocks (on the helper lanes its smaller classes leave free), then the blocks outside
// the mask that the FP32 pass finds dense get theirs, and the mask keeps the blocks whose FP64
// density exceeds the threshold.  The lists are filled on the device, so those launches cover the
// lists' capacity and run only the filled part.
void run_density_mixed(xc_bfogc* bf, int want_grad, const Lanes& L) {
  DenseLists carry{}, late{};
  for (int c = 0; c <= kDensMaxS; ++c) {
    carry.item[c] = bf->d_carry_list[c].p;
    late.item[c] = bf->d_dense_list[c].p;
  }
  carry.count = bf->d_carry_cnt.p;
  late.count = bf->d_dense_cnt.p;
  const int n = bf->basis->nao;
  const int64_t N = bf->grid->npts;
  auto fp64_pass = [&](int c, hipStream_t s, const DenseLists& lists) {
    launch_density_sub<double>(s, c, lists.item[c], (int)bf->dens_work[c].size(), bf->d_blocks.p, bf->d_ao_idx.p,
                               bf->d_t64.p, bf->d_P32d.p, n, want_grad, bf->d_part.p, N, nullptr, lists.count + c);
  };
  launch_list_masked(L.main, bf->d_blocks.p, bf->nblocks, bf->d_dens_by_block.p, bf->d_dens_item0.p,
                     bf->d_dens_cls.p, carry, bf->d_dense_mask.p);
  const bool par = L.n > 1;
  if (par) L.fork_all();
  // FP32 classes from the widest on lanes 0, 1, ...; FP64 classes from the widest on the last
  // lane backwards, behind the FP32 pass's smallest classes
  int k = 0, kb = 0;
  for (int c = kDensMaxS; c >= 1; --c)
    if (!bf->dens_work[c].empty())
      launch_density_sub<float>(par ? L.lane(k++) : L.main, c, bf->d_dens_work[c].p, (int)bf->dens_work[c].size(),
                                bf->d_blocks.p, bf->d_ao_idx.p, bf->d_t32.p, bf->d_P32.p, n, want_grad,
                                bf->d_part.p, N, nullptr, nullptr, bf->d_dense_mask.p);
  for (int c = kDensMaxS; c >= 1; --c)
    if (!bf->dens_work[c].empty()) fp64_pass(c, par ? L.lane(L.n - 1 - kb++ % L.n) : L.main, carry);
  if (par) L.join_all();
  launch_mark_dense(L.main, bf->d_blocks.p, bf->nblocks, N, bf->d_part.p, kMixedRhoThreshold, bf->d_dens_by_block.p,
                    bf->d_dens_item0.p, bf->d_dens_cls.p, late, bf->d_dense_mask.p);
  over_lanes(L, kDensMaxS + 1, [&](int c) { return c >= 1 ? (int)bf->dens_work[c].size() : 0; },
             [&](int c, hipStream_t s) { fp64_pass(c, s, late); });
  launch_refresh_dense(L.main, bf->d_blocks.p, bf->nblocks, N, bf->d_part.p, kMixedRhoThreshold,
                       bf->d_dense_mask.p);
}

// A controller whose server build is still running owns its work buffers (P, partial
// densities, potentials, U) until the build completes: refuse any other use first.
void require_idle(const xc_bfogc* bf, const char* who) {
  require(!bf->in_flight, XC_ERR_STATE, std::string(who) + ": a server build of this controller is in flight");
}

void require_idle(const xc_bfogc* act, int nenv, xc_bfogc* const* env, const char* who) {
  require_idle(act, who);
  for (int i = 0; i < nenv; ++i)
    if (env && env[i]) require_idle(env[i], who);
}

// The Fock stage in two steps: the blocks' sub-tiles into the scratch (run_fock_tiles), then the
// row reduce into the lower triangle of U (run_fock_reduce; every element of that triangle is
// written, so U needs no clearing).  run_fock does both.
template <typename T>
void run_fock_tiles(xc_bfogc* bf, const T* tiles, int gga, double thr, const Lanes& L, const double* vs = nullptr,
                    const double* gw = nullptr) {
  const int f = sizeof(T) == 8 ? 1 : 0;
  if (!vs) vs = bf->d_vs.p;
  if (!gw) gw = bf->d_gw.p;
  T* Tscr = reinterpret_cast<T*>(bf->d_T.p);
  over_lanes(L, kFockNumS, [&](int c) { return (int)bf->fock_items[f][c].size(); }, [&](int c, hipStream_t s) {
    launch_fock_sub<T>(s, c, bf->d_fock_items[f][c].p, (int)bf->fock_items[f][c].size(),
                       bf->d_blocks.p, bf->d_ao_idx.p, tiles, vs, gw, gga, thr, Tscr, bf->grid->npts);
  });
}

template <typename T>
void run_fock_reduce(xc_bfogc* bf, const Lanes& L) {
  launch_fock_reduce<T>(L.main, bf->d_rowent.p, bf->d_rowptr.p, reinterpret_cast<const T*>(bf->d_T.p),
                        bf->d_ao_idx.p, bf->d_U.p, bf->basis->nao, nullptr, 1, bf->ctx->deterministic);
}

template <typename T>
void run_fock(xc_bfogc* bf, const T* tiles, int gga, double thr, const Lanes& L, const double* vs = nullptr,
              const double* gw = nullptr) {
  run_fock_tiles<T>(bf, tiles, gga, thr, L, vs, gw);
  run_fock_reduce<T>(bf, L);
}

}  // namespace

namespace {
// Shell list of the host integral entry points (flat arrays as in xc_basis_create).
std::vector<Shell> host_shells(const char* who, int nshell, const int* shell_l, const int* shell_pure,
                               const int* shell_nprim, const double* shell_center, const double* alpha,
                               const double* coef, int lmax) {
  require(nshell > 0 && shell_l && shell_pure && shell_nprim && shell_center && alpha && coef, XC_ERR_ARG,
          std::string(who) + ": bad shell arguments");
  std::vector<Shell> sh(nshell);
  int k = 0;
  for (int s = 0; s < nshell; ++s) {
    sh[s].l = shell_l[s];
    sh[s].pure = shell_pure[s] != 0;
    require(sh[s].l >= 0 && sh[s].l <= lmax, XC_ERR_INPUT,
            std::string(who) + ": angular momentum above " + std::to_string(lmax));
    require(shell_nprim[s] > 0, XC_ERR_INPUT, std::string(who) + ": shell without primitives");
    sh[s].center = {shell_center[3 * s], shell_center[3 * s + 1], shell_center[3 * s + 2]};
    for (int q = 0; q < shell_nprim[s]; ++q, ++k) {
      sh[s].alpha.push_back(alpha[k]);
      sh[s].coef.push_back(coef[k]);
    }
  }
  return sh;
}
int nao_of(const std::vector<Shell>& sh) {
  int n = 0;
  for (const auto& s : sh) n += s.nfunc();
  return n;
}
}  // namespace

extern "C" {

const char* xc_last_error(void) { return g_last_error.c_str(); }

int xc_version(void) { return 1; }

int xc_ctx_create(int device, xc_ctx** out) {
  return guarded([&] {
    require(out != nullptr, XC_ERR_ARG, "xc_ctx_create: null output");
    int ndev = 0;
    HIPX(hipGetDeviceCount(&ndev));
    require(device >= 0 && device < ndev, XC_ERR_ARG, "xc_ctx_create: no HIP device " + std::to_string(device));
    auto* c = new xc_ctx();
    c->device = device;
    HIPX(hipSetDevice(device));
    noalloc_setup_current_device(device);
    HIPX(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    c->lanes.create(c->stream, true);
    c->d_energy = (double*)c->alloc(sizeof(double));
    *out = c;
  });
}

int xc_ctx_destroy(xc_ctx* ctx) {
  return guarded([&] {
    if (!ctx) return;
    require(ctx->live_handles == 0, XC_ERR_STATE,
            "xc_ctx_destroy: " + std::to_string(ctx->live_handles) + " basis/grid/controller handles still alive");
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    for (auto e : ctx->ev_pool) (void)hipEventDestroy(e);
    ctx->release(ctx->d_energy, sizeof(double));
    delete ctx->pack;
    ctx->pack = nullptr;
    for (auto& kv : ctx->pack_graphs) delete kv.second;
    ctx->pack_graphs.clear();
    ctx->lanes.destroy();
    (void)hipStreamDestroy(ctx->stream);
    ctx->trim();
    delete ctx;
  });
}

int xc_ctx_synchronize(xc_ctx* ctx) {
  return guarded([&] {
    require(ctx, XC_ERR_ARG, "null context");
    HIPX(hipStreamSynchronize(ctx->stream));
  });
}

void* xc_ctx_stream(xc_ctx* ctx) { return ctx ? (void*)ctx->stream : nullptr; }

int xc_ctx_memory(xc_ctx* ctx, int64_t* bytes_in_use) {
  return guarded([&] {
    require(ctx && bytes_in_use, XC_ERR_ARG, "null argument");
    *bytes_in_use = ctx->bytes;
  });
}

void xc_free_host(void* p) { std::free(p); }

int xc_normalize_contraction(int l, int nprim, const double* alpha, double* coef) {
  return guarded([&] {
    require(l >= 0 && nprim > 0 && alpha && coef, XC_ERR_ARG, "xc_normalize_contraction: bad arguments");
    std::vector<double> a(alpha, alpha + nprim), c(coef, coef + nprim);
    renormalize_contraction(l, a, c);
    std::copy(c.begin(), c.end(), coef);
  });
}

int xc_load_basis(const char* path, int natoms, const int* Z, const double* coords, int pure, int* nshell,
                  int** shell_l, int** shell_nprim, int** shell_atom, double** shell_center, double** alpha,
                  double** coef) {
  return guarded([&] {
    require(path && natoms > 0 && Z && coords && nshell && shell_l && shell_nprim && shell_atom && shell_center &&
                alpha && coef,
            XC_ERR_ARG, "xc_load_basis: bad arguments");
    static const char* symbols[] = {"",  "h",  "he", "li", "be", "b",  "c",  "n",  "o",  "f",  "ne", "na", "mg",
                                    "al", "si", "p",  "s",  "cl", "ar", "k",  "ca", "sc", "ti", "v",  "cr", "mn",
                                    "fe", "co", "ni", "cu", "zn", "ga", "ge", "as", "se", "br", "kr"};
    std::vector<Shell> all;
    for (int a = 0; a < natoms; ++a) {
      require(Z[a] >= 1 && Z[a] <= 36, XC_ERR_INPUT, "xc_load_basis: element out of range");
      auto sh = parse_turbomole_basis(path, symbols[Z[a]], pure != 0);
      for (auto& s : sh) {
        s.center = {coords[3 * a], coords[3 * a + 1], coords[3 * a + 2]};
        s.atom = a;
        all.push_back(s);
      }
    }
    const int ns = (int)all.size();
    int np = 0;
    for (auto& s : all) np += (int)s.alpha.size();
    *nshell = ns;
    *shell_l = (int*)std::malloc(sizeof(int) * ns);
    *shell_nprim = (int*)std::malloc(sizeof(int) * ns);
    *shell_atom = (int*)std::malloc(sizeof(int) * ns);
    *shell_center = (double*)std::malloc(sizeof(double) * 3 * ns);
    *alpha = (double*)std::malloc(sizeof(double) * np);
    *coef = (double*)std::malloc(sizeof(double) * np);
    int k = 0;
    for (int s = 0; s < ns; ++s) {
      (*shell_l)[s] = all[s].l;
      (*shell_nprim)[s] = (int)all[s].alpha.size();
      (*shell_atom)[s] = all[s].atom;
      for (int d = 0; d < 3; ++d) (*shell_center)[3 * s + d] = all[s].center[d];
      for (size_t q = 0; q < all[s].alpha.size(); ++q, ++k) {
        (*alpha)[k] = all[s].alpha[q];
        (*coef)[k] = all[s].coef[q];
      }
    }
  });
}

int xc_molecular_grid(int natoms, const int* Z, const double* coords, int grid_type, int radial_type, int accuracy,
                      int smoothing, double weight_threshold, int sort, int64_t* npoints, double** points,
                      double** weights, int64_t* npoints_before_cut) {
  return guarded([&] {
    require(natoms > 0 && Z && coords && npoints && points && weights, XC_ERR_ARG, "xc_molecular_grid: bad arguments");
    require(grid_type >= 0 && grid_type <= 2 && radial_type >= 0 && radial_type <= 4, XC_ERR_ARG,
            "xc_molecular_grid: unknown grid or radial type");
    GridSettings s;
    s.type = (GridType)grid_type;
    s.radial = (RadialType)radial_type;
    s.accuracy = accuracy;
    s.smoothing = (unsigned)smoothing;
    s.weight_threshold = weight_threshold;
    s.sort = sort != 0;
    MolecularGrid g = build_molecular_grid(std::vector<int>(Z, Z + natoms), std::vector<double>(coords, coords + 3 * natoms), s);
    const int64_t n = (int64_t)g.weights.size();
    *npoints = n;
    *points = (double*)std::malloc(sizeof(double) * 3 * std::max<int64_t>(n, 1));
    *weights = (double*)std::malloc(sizeof(double) * std::max<int64_t>(n, 1));
    std::memcpy(*points, g.points.data(), sizeof(double) * 3 * n);
    std::memcpy(*weights, g.weights.data(), sizeof(double) * n);
    if (npoints_before_cut) *npoints_before_cut = g.n_before_cut;
  });
}

int xc_molecular_grid_device(xc_ctx* ctx, int natoms, const int* Z, const double* coords, int grid_type,
                             int radial_type, int accuracy, int smoothing, double weight_threshold, int sort,
                             int64_t* npoints, double** points, double** weights, int64_t* npoints_before_cut) {
  return guarded([&] {
    require(ctx && natoms > 0 && Z && coords && npoints && points && weights, XC_ERR_ARG,
            "xc_molecular_grid_device: bad arguments");
    require(grid_type >= 0 && grid_type <= 2 && radial_type >= 0 && radial_type <= 4, XC_ERR_ARG,
            "xc_molecular_grid_device: unknown grid or radial type");
    GridSettings s;
    s.type = (GridType)grid_type;
    s.radial = (RadialType)radial_type;
    s.accuracy = accuracy;
    s.smoothing = (unsigned)smoothing;
    s.weight_threshold = weight_threshold;
    s.sort = sort != 0;
    // everything after the per-element reference grids runs on the device: candidates, partition,
    // weight cut, Hilbert order
    const GridProblem P = grid_problem(std::vector<int>(Z, Z + natoms),
                                       std::vector<double>(coords, coords + 3 * natoms), s, false);
    const int64_t nc = P.cand_off.back();
    require(nc < (int64_t)1 << 31, XC_ERR_ARG, "xc_molecular_grid_device: too many candidate points");
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    HIPX(hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    DBuf<double> d_xyz, d_dist, d_apair, d_min, d_ref_xyz, d_ref_w, d_w, d_pts, d_pw, d_part, d_box, d_opts, d_ow;
    DBuf<int64_t> d_ref_off, d_cand_off, d_key, d_key2, d_idx, d_order;
    DBuf<int> d_keep, d_pos;
    DBuf<unsigned char> d_temp;
    d_xyz.upload(ctx, P.xyz.data(), P.xyz.size());
    d_dist.upload(ctx, P.dist.data(), P.dist.size());
    d_apair.upload(ctx, P.apair.data(), P.apair.size());
    d_min.upload(ctx, P.min_dist.data(), P.min_dist.size());
    d_ref_xyz.upload(ctx, P.ref_xyz.data(), P.ref_xyz.size());
    d_ref_w.upload(ctx, P.ref_w.data(), P.ref_w.size());
    d_ref_off.upload(ctx, P.ref_off.data(), P.ref_off.size());
    d_cand_off.upload(ctx, P.cand_off.data(), P.cand_off.size());
    const size_t ncs = (size_t)std::max<int64_t>(nc, 1);
    d_w.alloc(ctx, ncs);
    d_keep.alloc(ctx, ncs);
    d_pos.alloc(ctx, ncs);
    d_pts.alloc(ctx, 3 * ncs);
    d_pw.alloc(ctx, ncs);
    launch_partition_gen(st, P.type, P.smoothing, P.natoms, d_xyz.p, d_dist.p, d_apair.p, d_min.p, nc, d_ref_xyz.p,
                         d_ref_w.p, d_ref_off.p, d_cand_off.p, weight_threshold, d_w.p, d_keep.p);
    const size_t scan_bytes = grid_scan_temp_bytes(nc);
    d_temp.alloc(ctx, std::max<size_t>(scan_bytes, 1));
    launch_cut(st, P.natoms, d_xyz.p, nc, d_ref_xyz.p, d_ref_off.p, d_cand_off.p, d_w.p, d_keep.p, d_pos.p, d_temp.p,
               scan_bytes, d_pts.p, d_pw.p);
    HIPX(hipGetLastError());
    int last[2] = {0, 0};
    if (nc) {
      HIPX(hipMemcpyAsync(&last[0], d_pos.p + nc - 1, sizeof(int), hipMemcpyDeviceToHost, st));
      HIPX(hipMemcpyAsync(&last[1], d_keep.p + nc - 1, sizeof(int), hipMemcpyDeviceToHost, st));
    }
    HIPX(hipStreamSynchronize(st));
    const int64_t n = (int64_t)last[0] + last[1];
    const double* res_pts = d_pts.p;
    const double* res_w = d_pw.p;
    if (s.sort && n >= 2) {
      const int64_t length = hilbert_length(n);
      d_part.alloc(ctx, 6 * (size_t)bbox_part_count());
      d_box.alloc(ctx, 6);
      launch_bbox(st, d_pts.p, n, length, d_part.p, d_box.p);
      const size_t sort_bytes = grid_sort_temp_bytes(n);
      if (sort_bytes > d_temp.n) d_temp.alloc(ctx, sort_bytes);
      d_key.alloc(ctx, n);
      d_key2.alloc(ctx, n);
      d_idx.alloc(ctx, n);
      d_order.alloc(ctx, n);
      d_opts.alloc(ctx, 3 * (size_t)n);
      d_ow.alloc(ctx, n);
      launch_hilbert_order(st, d_pts.p, d_pw.p, n, d_box.p, length, d_key.p, d_key2.p, d_idx.p, d_order.p, d_temp.p,
                           sort_bytes, d_opts.p, d_ow.p);
      HIPX(hipGetLastError());
      res_pts = d_opts.p;
      res_w = d_ow.p;
    }
    *npoints = n;
    *points = (double*)std::malloc(sizeof(double) * 3 * std::max<int64_t>(n, 1));
    *weights = (double*)std::malloc(sizeof(double) * std::max<int64_t>(n, 1));
    if (n) {
      HIPX(hipMemcpyAsync(*points, res_pts, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, st));
      HIPX(hipMemcpyAsync(*weights, res_w, sizeof(double) * n, hipMemcpyDeviceToHost, st));
    }
    HIPX(hipStreamSynchronize(st));
    if (npoints_before_cut) *npoints_before_cut = nc;
  });
}

int xc_hilbert_sort(int64_t npoints, double* points, double* weights) {
  return guarded([&] {
    require(npoints >= 0 && (npoints == 0 || (points && weights)), XC_ERR_ARG, "xc_hilbert_sort: bad arguments");
    std::vector<double> p(points, points + 3 * npoints), w(weights, weights + npoints);
    hilbert_sort(p, w);
    std::copy(p.begin(), p.end(), points);
    std::copy(w.begin(), w.end(), weights);
  });
}

int xc_lebedev(int npoints, double* x, double* y, double* z, double* w) {
  return guarded([&] {
    require(x && y && z && w, XC_ERR_ARG, "xc_lebedev: null output");
    std::vector<double> vx, vy, vz, vw;
    lebedev_rule(npoints, vx, vy, vz, vw);
    std::copy(vx.begin(), vx.end(), x);
    std::copy(vy.begin(), vy.end(), y);
    std::copy(vz.begin(), vz.end(), z);
    std::copy(vw.begin(), vw.end(), w);
  });
}


int xc_overlap(int nshell, const int* shell_l, const int* shell_pure, const int* shell_nprim, const double* shell_center,
               const double* alpha, const double* coef, double* S) {
  return guarded([&] {
    require(S, XC_ERR_ARG, "xc_overlap: null output");
    auto sh = host_shells("xc_overlap", nshell, shell_l, shell_pure, shell_nprim, shell_center, alpha, coef, 6);
    std::vector<double> out;
    overlap_matrix(sh, out);
    std::copy(out.begin(), out.end(), S);
  });
}

int xc_boys(int m, double T, double* F) {
  return guarded([&] {
    require(F && m >= 0 && m <= 17 && T >= 0.0, XC_ERR_ARG, "xc_boys: bad arguments");
    *F = boys_function(m, T);
  });
}

int xc_one_electron(int nshell, const int* shell_l, const int* shell_pure, const int* shell_nprim,
                    const double* shell_center, const double* alpha, const double* coef, int natoms, const int* Z,
                    const double* coords, double* T, double* V) {
  return guarded([&] {
    require(natoms >= 0 && (natoms == 0 || (Z && coords)) && T && V, XC_ERR_ARG, "xc_one_electron: bad arguments");
    auto sh = host_shells("xc_one_electron", nshell, shell_l, shell_pure, shell_nprim, shell_center, alpha, coef, 4);
    std::vector<int> z(Z, Z + natoms);
    std::vector<double> x(coords, coords + 3 * natoms), t, v;
    one_electron_matrices(sh, z, x, t, v);
    std::copy(t.begin(), t.end(), T);
    std::copy(v.begin(), v.end(), V);
  });
}

int xc_eri_packed(int nshell, const int* shell_l, const int* shell_pure, const int* shell_nprim,
                  const double* shell_center, const double* alpha, const double* coef, double schwarz_threshold,
                  double* eri) {
  return guarded([&] {
    require(eri && schwarz_threshold >= 0.0, XC_ERR_ARG, "xc_eri_packed: bad arguments");
    auto sh = host_shells("xc_eri_packed", nshell, shell_l, shell_pure, shell_nprim, shell_center, alpha, coef, 4);
    const int64_t n = nao_of(sh), np = n * (n + 1) / 2;
    require(np * np <= (int64_t)1 << 31, XC_ERR_INPUT, "xc_eri_packed: more than 2^31 integrals; use density fitting");
    std::vector<double> out;
    eri_packed(sh, schwarz_threshold, out);
    std::copy(out.begin(), out.end(), eri);
  });
}

int xc_ri_j_integrals(int nshell, const int* shell_l, const int* shell_pure, const int* shell_nprim,
                      const double* shell_center, const double* alpha, const double* coef, int naux_shell,
                      const int* aux_l, const int* aux_pure, const int* aux_nprim, const double* aux_center,
                      const double* aux_alpha, const double* aux_coef, double schwarz_threshold, double* B3,
                      double* M) {
  return guarded([&] {
    require(B3 && M && schwarz_threshold >= 0.0, XC_ERR_ARG, "xc_ri_j_integrals: bad arguments");
    auto sh = host_shells("xc_ri_j_integrals", nshell, shell_l, shell_pure, shell_nprim, shell_center, alpha, coef, 4);
    auto ax = host_shells("xc_ri_j_integrals (auxiliary)", naux_shell, aux_l, aux_pure, aux_nprim, aux_center,
                          aux_alpha, aux_coef, 4);
    std::vector<double> b, m;
    ri_j_integrals(sh, ax, schwarz_threshold, b, m);
    std::copy(b.begin(), b.end(), B3);
    std::copy(m.begin(), m.end(), M);
  });
}

int xc_basis_create(xc_ctx* ctx, int nshell, const int* shell_l, const int* shell_pure, const int* shell_nprim,
                    const double* shell_center, const double* alpha, const double* coef, xc_basis** out) {
  return guarded([&] {
    require(ctx && out && nshell > 0 && shell_l && shell_pure && shell_nprim && shell_center && alpha && coef,
            XC_ERR_ARG, "xc_basis_create: bad arguments");
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    auto* b = new xc_basis();
    b->ctx = ctx;
    b->nshell = nshell;
    int prim = 0, ao = 0;
    for (int s = 0; s < nshell; ++s) {
      const int l = shell_l[s];
      if (l < 0 || l > kMaxL) {
        delete b;
        throw XcError(XC_ERR_INPUT, "xc_basis_create: angular momentum above " + std::to_string(kMaxL));
      }
      DevShell d{};
      d.cx = shell_center[3 * s];
      d.cy = shell_center[3 * s + 1];
      d.cz = shell_center[3 * s + 2];
      d.l = l;
      d.pure = shell_pure[s] != 0;
      d.nprim = shell_nprim[s];
      d.prim_off = prim;
      d.ao_off = ao;
      d.ncart = (l + 1) * (l + 2) / 2;
      d.nfunc = d.pure ? 2 * l + 1 : d.ncart;
      d.ang_off = (int)b->ang.size();
      if (d.pure) {
        auto T = solid_harmonic_coefficients(l);
        b->ang.insert(b->ang.end(), T.begin(), T.end());
      } else {
        auto f = cartesian_norm_factors(l);
        for (int i = 0; i < d.ncart; ++i)
          for (int j = 0; j < d.ncart; ++j) b->ang.push_back(i == j ? f[i] : 0.0);
      }
      for (int q = 0; q < d.nprim; ++q) {
        b->alpha.push_back(alpha[prim + q]);
        b->coef.push_back(coef[prim + q]);
      }
      prim += d.nprim;
      ao += d.nfunc;
      b->maxl = std::max(b->maxl, l);
      b->shells.push_back(d);
    }
    b->nao = ao;
    b->d_shells.upload(ctx, b->shells.data(), b->shells.size());
    b->d_alpha.upload(ctx, b->alpha.data(), b->alpha.size());
    b->d_coef.upload(ctx, b->coef.data(), b->coef.size());
    b->d_ang.upload(ctx, b->ang.data(), b->ang.size());
    HIPX(hipStreamSynchronize(ctx->stream));
    ++ctx->live_handles;
    *out = b;
  });
}

int xc_basis_destroy(xc_basis* basis) {
  return guarded([&] {
    if (!basis) return;
    --basis->ctx->live_handles;
    delete basis;
  });
}

int xc_basis_nao(const xc_basis* basis, int* nao) {
  return guarded([&] {
    require(basis && nao, XC_ERR_ARG, "null argument");
    *nao = basis->nao;
  });
}

int xc_grid_create(xc_ctx* ctx, int64_t npoints, const double* points, const double* weights, xc_grid** out) {
  return guarded([&] {
    require(ctx && out && npoints > 0 && points && weights, XC_ERR_ARG, "xc_grid_create: bad arguments");
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    auto* g = new xc_grid();
    g->ctx = ctx;
    g->npts = npoints;
    g->d_pts.upload(ctx, points, (size_t)(3 * npoints));
    g->d_w.upload(ctx, weights, (size_t)npoints);
    HIPX(hipStreamSynchronize(ctx->stream));
    ++ctx->live_handles;
    *out = g;
  });
}

int xc_grid_destroy(xc_grid* grid) {
  return guarded([&] {
    if (!grid) return;
    --grid->ctx->live_handles;
    delete grid;
  });
}

int xc_bfogc_create(xc_basis* basis, xc_grid* grid, int max_block_size, double radial_threshold,
                    int highest_derivative, int keep_fp32, xc_bfogc** out) {
  return guarded([&] {
    require(basis && grid && out, XC_ERR_ARG, "xc_bfogc_create: null argument");
    require(basis->ctx == grid->ctx, XC_ERR_ARG, "xc_bfogc_create: basis and grid live on different contexts");
    require(max_block_size >= 1 && max_block_size <= kTile, XC_ERR_ARG,
            "xc_bfogc_create: max block size must be in [1, 128]");
    require(highest_derivative >= 0 && highest_derivative <= 2, XC_ERR_ARG,
            "xc_bfogc_create: highest derivative must be 0, 1 or 2");
    require(radial_threshold >= 0.0, XC_ERR_ARG, "xc_bfogc_create: radial threshold must not be negative");
    require(keep_fp32 >= 0 && keep_fp32 <= 2, XC_ERR_ARG, "xc_bfogc_create: keep_fp32 must be 0, 1 or 2");
    xc_ctx* ctx = basis->ctx;
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    HIPX(hipSetDevice(ctx->device));
    auto* bf = new xc_bfogc();
    try {
      bf->ctx = ctx;
      bf->basis = basis;
      bf->grid = grid;
      bf->bs = max_block_size;
      bf->radial_thr = radial_threshold;
      bf->deriv = highest_derivative;
      bf->ncomp = highest_derivative >= 2 ? 10 : highest_derivative == 1 ? 4 : 1;
      const int64_t N = grid->npts;
      const int64_t nb64 = (N + max_block_size - 1) / max_block_size;
      require(nb64 < (1LL << 31), XC_ERR_ARG, "xc_bfogc_create: too many blocks");
      const int nb = (int)nb64;
      bf->nblocks = nb;
      const int ns = basis->nshell;
      // 1. shell screening per block on the device
      DBuf<unsigned char> d_flags;
      d_flags.alloc(ctx, (size_t)nb * ns);
      launch_screen(ctx->stream, grid->d_pts.p, N, max_block_size, nb, basis->d_shells.p, ns, basis->d_alpha.p,
                    basis->d_coef.p, radial_threshold, d_flags.p);
      HIPX(hipGetLastError());
      std::vector<unsigned char> flags((size_t)nb * ns);
      HIPX(hipMemcpyAsync(flags.data(), d_flags.p, flags.size(), hipMemcpyDeviceToHost, ctx->stream));
      HIPX(hipStreamSynchronize(ctx->stream));
      // 2. block lists (compact AO index lists padded to kAoPad) and tile offsets
      std::vector<int> sig_shell, sig_pos;
      bf->blocks.resize(nb);
      bf->dens_item0.assign(nb, 0);
      bf->dens_cls.assign(nb, 1);
      int64_t tile_off = 0;
      for (int b = 0; b < nb; ++b) {
        DevBlock& bk = bf->blocks[b];
        bk.first = (int64_t)b * max_block_size;
        bk.npts = (int)std::min<int64_t>(max_block_size, N - bk.first);
        bk.idx_off = (int)bf->ao_idx.size();
        bk.shell_off = (int)sig_shell.size();
        int cnt = 0, nsh = 0;
        for (int s = 0; s < ns; ++s) {
          if (!flags[(size_t)b * ns + s]) continue;
          const DevShell& sh = basis->shells[s];
          sig_shell.push_back(s);
          sig_pos.push_back(cnt);
          for (int f = 0; f < sh.nfunc; ++f) bf->ao_idx.push_back(sh.ao_off + f);
          cnt += sh.nfunc;
          ++nsh;
        }
        bk.nb = cnt;
        bk.nshell = nsh;
        bk.npad = (cnt + kAoPad - 1) / kAoPad * kAoPad;
        for (int i = cnt; i < bk.npad; ++i) bf->ao_idx.push_back(-1);
        bk.tile_off = tile_off;
        tile_off += (int64_t)bk.npad * kTile * bf->ncomp;
        bk.toff = bf->t_elems;
        bf->t_elems += fock_subtile_elems(bk.npad);
        bf->sum_nb += cnt;
        bf->sum_B_nb2 += (double)bk.npts * cnt * cnt;
        bk.npass = 0;
        if (cnt > 0) {
          // density: balanced passes of at most 16 kDensMaxS columns, 16-column granularity
          const int npass = (bk.npad + 16 * kDensMaxS - 1) / (16 * kDensMaxS);
          const int width = ((bk.npad + npass - 1) / npass + 15) / 16 * 16;
          int s = 0;
          // every pass of a block in the class of the widest, so that they share a launch
          bf->dens_item0[b] = (int)bf->dens_by_block.size();
          bf->dens_cls[b] = width / 16;
          for (int c0 = 0; c0 < bk.npad; c0 += width, ++s) {
            const int ncols = std::min(width, bk.npad - c0);
            bf->dens_work[width / 16].push_back(WorkItem{b, c0, ncols, s});
            bf->dens_by_block.push_back(WorkItem{b, c0, ncols, s});
          }
          bk.npass = s;
          bf->npass_max = std::max(bf->npass_max, s);
          // Fock: the whole upper triangle of 16 x 16 sub-tiles as one item where
          // its accumulators fit, else pairs of 128-row panels
          // (all items of a block in one slot class, the largest they need, so that they
          // share a launch and can be ordered for L2 reuse)
          for (int fp64 = 0; fp64 < 2; ++fp64) {
            std::vector<FockItem> mine;
            auto add = [&](int a0, int na, int b0, int nbr) {
              const int ta = na / 16, tb = nbr / 16;
              const int nsub = nbr == 0 ? ta * (ta + 1) / 2 : ta * tb;
              mine.push_back(FockItem{b, a0, na, b0, nbr, nsub});
            };
            const int nt16 = bk.npad / 16;
            if (bk.npad <= kFockRows && nt16 * (nt16 + 1) / 2 <= kFockWaves * fock_smax(fp64)) {
              add(0, bk.npad, 0, 0);
            } else {
              const int np = (bk.npad + kFockPanel - 1) / kFockPanel;
              for (int P = 0; P < np; ++P) {
                const int a0 = P * kFockPanel, na = std::min(kFockPanel, bk.npad - a0);
                add(a0, na, 0, 0);
                for (int Q = P + 1; Q < np; ++Q) {
                  const int b0 = Q * kFockPanel;
                  add(a0, na, b0, std::min(kFockPanel, bk.npad - b0));
                }
              }
            }
            int nsub_max = 0;
            for (const FockItem& it : mine) nsub_max = std::max(nsub_max, it.nsub);
            int sc = 0;
            while (kFockWaves * kFockS[sc] < nsub_max) ++sc;
            for (const FockItem& it : mine) bf->fock_items[fp64][sc].push_back(it);
          }
        }
      }
      for (int c = 0; c <= kDensMaxS; ++c) xcd_order(bf->dens_work[c]);
      for (int f = 0; f < 2; ++f)
        for (int c = 0; c < kFockNumS; ++c) xcd_order(bf->fock_items[f][c]);
      bf->tile_elems = tile_off;
      // row entries of the Fock reduce: for every AO, the blocks holding it (in block order) with
      // the position of its row in the scratch
      const int nao = basis->nao;
      std::vector<int> rowptr(nao + 1, 0);
      for (int b = 0; b < nb; ++b)
        for (int l = 0; l < bf->blocks[b].nb; ++l) ++rowptr[bf->ao_idx[bf->blocks[b].idx_off + l] + 1];
      for (int i = 0; i < nao; ++i) rowptr[i + 1] += rowptr[i];
      std::vector<RowEnt> rowent(std::max(rowptr[nao], 1));
      {
        std::vector<int> fill(rowptr.begin(), rowptr.end() - 1);
        for (int b = 0; b < nb; ++b) {
          const DevBlock& bk = bf->blocks[b];
          for (int l = 0; l < bk.nb; ++l) {
            RowEnt e;
            e.rowbase = bk.toff + fock_row_start(l, bk.npad) - 16 * (l / 16);
            e.c0 = l;
            e.c1 = bk.nb;
            e.idx_off = bk.idx_off;
            e.block = b;
            rowent[fill[bf->ao_idx[bk.idx_off + l]]++] = e;
          }
        }
      }
      // 3. admission by footprint (the context's cached blocks count as free)
      size_t free_b = 0, total_b = 0;
      HIPX(hipMemGetInfo(&free_b, &total_b));
      free_b += ctx->cached;
      const double need =
          (double)tile_off * ((keep_fp32 != 2 ? sizeof(double) : 0) + (keep_fp32 ? sizeof(float) : 0)) +
          (double)bf->t_elems * (keep_fp32 != 2 ? sizeof(double) : sizeof(float));
      require(need < 0.9 * (double)free_b, XC_ERR_MEMORY,
              "xc_bfogc_create: AO tiles and Fock scratch need " + std::to_string(need / 1e9) + " GB, " +
                  std::to_string(free_b / 1e9) + " GB free");
      // 4. upload lists, evaluate AO tiles
      bf->d_blocks.upload(ctx, bf->blocks.data(), bf->blocks.size());
      bf->d_rowptr.upload(ctx, rowptr.data(), rowptr.size());
      bf->d_rowent.upload(ctx, rowent.data(), rowent.size());
      bf->d_ao_idx.upload(ctx, bf->ao_idx.data(), bf->ao_idx.size());
      bf->d_sig_shell.upload(ctx, sig_shell.data(), sig_shell.size());
      bf->d_sig_pos.upload(ctx, sig_pos.data(), sig_pos.size());
      for (int c = 0; c <= kDensMaxS; ++c)
        bf->d_dens_work[c].upload(ctx, bf->dens_work[c].data(), bf->dens_work[c].size());
      if (keep_fp32 == 1) {  // both tile sets: the mixed-precision build is available
        bf->d_dens_by_block.upload(ctx, bf->dens_by_block.data(), bf->dens_by_block.size());
        bf->d_dens_item0.upload(ctx, bf->dens_item0.data(), bf->dens_item0.size());
        bf->d_dens_cls.upload(ctx, bf->dens_cls.data(), bf->dens_cls.size());
        for (int c = 1; c <= kDensMaxS; ++c) {
          bf->d_dense_list[c].alloc(ctx, std::max<size_t>(bf->dens_work[c].size(), 1));
          bf->d_carry_list[c].alloc(ctx, std::max<size_t>(bf->dens_work[c].size(), 1));
        }
        bf->d_dense_cnt.alloc(ctx, kDensMaxS + 1);
        bf->d_carry_cnt.alloc(ctx, kDensMaxS + 1);
        bf->d_dense_mask.alloc(ctx, std::max(bf->nblocks, 1));
        HIPX(hipMemsetAsync(bf->d_dense_mask.p, 0, std::max(bf->nblocks, 1), ctx->stream));
      }
      for (int f = 0; f < 2; ++f)
        for (int c = 0; c < kFockNumS; ++c)
          bf->d_fock_items[f][c].upload(ctx, bf->fock_items[f][c].data(), bf->fock_items[f][c].size());
      const size_t nt_el = (size_t)std::max<int64_t>(tile_off, 1);
      if (keep_fp32 != 2) {
        bf->d_t64.alloc(ctx, nt_el);
        HIPX(hipMemsetAsync(bf->d_t64.p, 0, sizeof(double) * nt_el, ctx->stream));
      }
      if (keep_fp32) {
        bf->d_t32.alloc(ctx, nt_el);
        HIPX(hipMemsetAsync(bf->d_t32.p, 0, sizeof(float) * nt_el, ctx->stream));
      }
      launch_ao_eval(ctx->stream, bf->d_blocks.p, nb, bf->d_sig_shell.p, bf->d_sig_pos.p, basis->d_shells.p,
                     basis->d_alpha.p, basis->d_coef.p, basis->d_ang.p, grid->d_pts.p, radial_threshold,
                     radial_threshold > 0.0 ? -std::log(radial_threshold) : HUGE_VAL, bf->ncomp, bf->d_t64.p,
                     keep_fp32 ? bf->d_t32.p : nullptr);
      HIPX(hipGetLastError());
      // 5. work buffers of a build
      const int n = basis->nao;
      const int nxc = (int)((N + kTile - 1) / kTile);
      bf->d_part.alloc(ctx, (size_t)std::max(bf->npass_max, 1) * 4 * N);
      bf->d_rho.alloc(ctx, N);
      bf->d_grad.alloc(ctx, 3 * N);
      bf->d_vs.alloc(ctx, N);
      bf->d_gw.alloc(ctx, 3 * N);
      bf->d_v.alloc(ctx, N);
      bf->d_g.alloc(ctx, 3 * N);
      bf->d_eblk.alloc(ctx, std::max(nxc, 1));
      bf->d_U.alloc(ctx, (size_t)n * n);
      // sub-tile scratch in the widest contraction precision the controller supports
      const int64_t t_doubles = keep_fp32 != 2 ? bf->t_elems : (bf->t_elems + 1) / 2;
      bf->d_T.alloc(ctx, (size_t)std::max<int64_t>(t_doubles, 1));
      bf->d_P.alloc(ctx, (size_t)n * n);
      bf->d_F.alloc(ctx, (size_t)n * n);
      bf->d_E.alloc(ctx, 1);
      if (keep_fp32) {
        bf->d_P32.alloc(ctx, (size_t)n * n);
        bf->d_round.alloc(ctx, 2 * (size_t)finalize_blocks(n));
      }
      if (keep_fp32 == 1) bf->d_P32d.alloc(ctx, (size_t)n * n);
      HIPX(hipStreamSynchronize(ctx->stream));
    } catch (...) {
      delete bf;
      throw;
    }
    ++ctx->live_handles;
    *out = bf;
  });
}

int xc_bfogc_destroy(xc_bfogc* bfogc) {
  return guarded([&] {
    if (!bfogc) return;
    std::lock_guard<std::recursive_mutex> lk(bfogc->ctx->mu);
    require(!bfogc->in_flight, XC_ERR_STATE, "xc_bfogc_destroy: a server build of this controller is in flight");
    (void)hipStreamSynchronize(bfogc->ctx->stream);
    forget_packs_of(bfogc->ctx, bfogc);
    --bfogc->ctx->live_handles;
    delete bfogc;
  });
}

int xc_bfogc_info(const xc_bfogc* bf, int* nblocks, int64_t* npoints, int* nao, int* max_block_size) {
  return guarded([&] {
    require(bf, XC_ERR_ARG, "null bfogc");
    if (nblocks) *nblocks = bf->nblocks;
    if (npoints) *npoints = bf->grid->npts;
    if (nao) *nao = bf->basis->nao;
    if (max_block_size) *max_block_size = bf->bs;
  });
}

int xc_bfogc_stats(const xc_bfogc* bf, int64_t* sum_nb, double* sum_B_nb2, int64_t* tile_bytes) {
  return guarded([&] {
    require(bf, XC_ERR_ARG, "null bfogc");
    if (sum_nb) *sum_nb = bf->sum_nb;
    if (sum_B_nb2) *sum_B_nb2 = bf->sum_B_nb2;
    if (tile_bytes)
      *tile_bytes = bf->tile_elems * (int64_t)((bf->d_t64.p ? sizeof(double) : 0) + (bf->d_t32.p ? sizeof(float) : 0));
  });
}

int xc_bfogc_block_data(const xc_bfogc* bf, int block, double* values, double* gx, double* gy, double* gz,
                        int* negligible, int* block_size) {
  return guarded([&] {
    require(bf, XC_ERR_ARG, "null bfogc");
    require(block >= 0 && block < bf->nblocks, XC_ERR_ARG, "block index out of range");
    require(!(gx || gy || gz) || bf->deriv >= 1, XC_ERR_STATE, "gradients were not evaluated (highest derivative 0)");
    require(bf->d_t64.p, XC_ERR_STATE, "xc_bfogc_block_data: the controller keeps FP32 tiles only");
    std::lock_guard<std::recursive_mutex> lk(bf->ctx->mu);
    const DevBlock& bk = bf->blocks[block];
    const int n = bf->basis->nao;
    const int64_t cnt = (int64_t)bk.npad * kTile * bf->ncomp;
    std::vector<double> t((size_t)std::max<int64_t>(cnt, 1));
    if (cnt) {
      HIPX(hipMemcpyAsync(t.data(), bf->d_t64.p + bk.tile_off, sizeof(double) * cnt, hipMemcpyDeviceToHost,
                          bf->ctx->stream));
      HIPX(hipStreamSynchronize(bf->ctx->stream));
    }
    const int B = bk.npts;
    if (block_size) *block_size = B;
    double* outs[4] = {values, gx, gy, gz};
    for (int c = 0; c < 4; ++c)
      if (outs[c]) std::fill(outs[c], outs[c] + (size_t)B * n, 0.0);
    if (negligible) std::fill(negligible, negligible + n, 1);
    for (int i = 0; i < bk.nb; ++i) {
      const int mu = bf->ao_idx[bk.idx_off + i];
      if (negligible) negligible[mu] = 0;
      for (int c = 0; c < std::min(bf->ncomp, 4); ++c) {
        if (!outs[c]) continue;
        for (int p = 0; p < B; ++p)
          outs[c][(size_t)mu * B + p] = t[((size_t)c * bk.npad + i) * kTile + p];
      }
    }
  });
}

int xc_bfogc_block_hessian(const xc_bfogc* bf, int block, double* hess) {
  return guarded([&] {
    require(bf && hess, XC_ERR_ARG, "xc_bfogc_block_hessian: null argument");
    require(block >= 0 && block < bf->nblocks, XC_ERR_ARG, "block index out of range");
    require(bf->deriv >= 2, XC_ERR_STATE, "second derivatives were not evaluated (highest derivative < 2)");
    require(bf->d_t64.p, XC_ERR_STATE, "xc_bfogc_block_hessian: the controller keeps FP32 tiles only");
    std::lock_guard<std::recursive_mutex> lk(bf->ctx->mu);
    const DevBlock& bk = bf->blocks[block];
    const int n = bf->basis->nao;
    const int B = bk.npts;
    const int64_t cnt = (int64_t)bk.npad * kTile * 6;
    std::vector<double> t((size_t)std::max<int64_t>(cnt, 1));
    if (cnt) {
      HIPX(hipMemcpyAsync(t.data(), bf->d_t64.p + bk.tile_off + (int64_t)4 * bk.npad * kTile, sizeof(double) * cnt,
                          hipMemcpyDeviceToHost, bf->ctx->stream));
      HIPX(hipStreamSynchronize(bf->ctx->stream));
    }
    std::fill(hess, hess + (size_t)6 * B * n, 0.0);
    for (int i = 0; i < bk.nb; ++i) {
      const int mu = bf->ao_idx[bk.idx_off + i];
      for (int c = 0; c < 6; ++c)
        for (int p = 0; p < B; ++p)
          hess[((size_t)c * n + mu) * B + p] = t[((size_t)c * bk.npad + i) * kTile + p];
    }
  });
}

int xc_density(xc_bfogc* bf, const double* P, double* rho, double* grad, int* nonneg) {
  return guarded([&] {
    require(bf && P && rho, XC_ERR_ARG, "xc_density: null argument");
    require(!bf->in_flight, XC_ERR_STATE, "a server build of this controller is in flight");
    require(!grad || bf->deriv >= 1, XC_ERR_STATE, "xc_density: gradient requested but AO gradients not evaluated");
    require(bf->d_t64.p, XC_ERR_STATE, "xc_density: the controller keeps FP32 tiles only");
    xc_ctx* ctx = bf->ctx;
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    const int n = bf->basis->nao;
    const int64_t N = bf->grid->npts;
    HIPX(hipMemcpyAsync(bf->d_P.p, P, sizeof(double) * n * n, hipMemcpyHostToDevice, ctx->stream));
    run_density<double>(bf, bf->d_t64.p, bf->d_P.p, grad ? 1 : 0, ctx->lanes);
    launch_reduce_partials(ctx->stream, bf->d_blocks.p, bf->bs, N, bf->d_part.p, grad ? 1 : 0, bf->d_rho.p,
                           bf->d_grad.p);
    HIPX(hipGetLastError());
    HIPX(hipMemcpyAsync(rho, bf->d_rho.p, sizeof(double) * N, hipMemcpyDeviceToHost, ctx->stream));
    if (grad) HIPX(hipMemcpyAsync(grad, bf->d_grad.p, sizeof(double) * 3 * N, hipMemcpyDeviceToHost, ctx->stream));
    HIPX(hipStreamSynchronize(ctx->stream));
    if (nonneg)
      for (int b = 0; b < bf->nblocks; ++b) nonneg[b] = bf->blocks[b].nb > 0 ? 1 : 0;
  });
}

int xc_functional(xc_ctx* ctx, int functional, int64_t npts, int block_size, const double* rho, const double* grad,
                  const double* weights, double* eps, double* v, double* g, double* energy) {
  return guarded([&] {
    require(ctx && rho && weights && npts > 0, XC_ERR_ARG, "xc_functional: bad arguments");
    require(functional >= 0 && functional < kNumFunctionals, XC_ERR_ARG, "xc_functional: unknown functional");
    require(block_size >= 1 && block_size <= kTile, XC_ERR_ARG, "xc_functional: block size must be in [1,128]");
    const bool gga = functional_is_gga(functional);
    require(!gga || grad, XC_ERR_ARG, "xc_functional: GGA functional needs the density gradient");
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    DBuf<double> d_rho, d_grad, d_w, d_eps, d_v, d_g, d_eblk, d_E;
    const int nxc = (int)((npts + block_size - 1) / block_size);
    d_rho.upload(ctx, rho, npts);
    if (gga) d_grad.upload(ctx, grad, 3 * npts);
    d_w.upload(ctx, weights, npts);
    d_eps.alloc(ctx, npts);
    d_v.alloc(ctx, npts);
    if (gga) d_g.alloc(ctx, 3 * npts);
    d_eblk.alloc(ctx, nxc);
    d_E.alloc(ctx, 1);
    launch_functional(ctx->stream, functional, ctx->prm, npts, block_size, d_rho.p, d_grad.p, nullptr, nullptr, d_w.p,
                      d_eps.p, d_v.p, d_g.p, nullptr, nullptr, d_eblk.p);
    launch_sum(ctx->stream, d_eblk.p, nxc, d_E.p);
    HIPX(hipGetLastError());
    if (eps) HIPX(hipMemcpyAsync(eps, d_eps.p, sizeof(double) * npts, hipMemcpyDeviceToHost, ctx->stream));
    if (v) HIPX(hipMemcpyAsync(v, d_v.p, sizeof(double) * npts, hipMemcpyDeviceToHost, ctx->stream));
    if (g) {
      if (gga)
        HIPX(hipMemcpyAsync(g, d_g.p, sizeof(double) * 3 * npts, hipMemcpyDeviceToHost, ctx->stream));
      else
        std::fill(g, g + 3 * npts, 0.0);
    }
    double E = 0.0;
    HIPX(hipMemcpyAsync(&E, d_E.p, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HIPX(hipStreamSynchronize(ctx->stream));
    if (energy) *energy = E;
  });
}

int xc_functional_device_noalloc(xc_ctx* ctx, int functional, int64_t npts, int block_size, const double* d_rho,
                                 const double* d_grad, const double* d_weights, double* d_eps, double* d_v,
                                 double* d_g, double* energy) {
  return guarded([&] {
    require(ctx && d_rho && d_weights && npts > 0, XC_ERR_ARG, "xc_functional_device_noalloc: bad arguments");
    require(functional >= 0 && functional < kNumFunctionals, XC_ERR_ARG,
            "xc_functional_device_noalloc: unknown functional");
    require(block_size >= 1 && block_size <= kTile, XC_ERR_ARG,
            "xc_functional_device_noalloc: block size must be in [1,128]");
    const bool gga = functional_is_gga(functional);
    require(!gga || d_grad, XC_ERR_ARG, "xc_functional_device_noalloc: GGA functional needs the density gradient");
    const int64_t nxc = (npts + block_size - 1) / block_size;
    require(nxc <= kNoallocBlocks, XC_ERR_ARG,
            "xc_functional_device_noalloc: " + std::to_string(nxc) + " blocks of " + std::to_string(block_size) +
                " points exceed the workspace of at most " + std::to_string(kNoallocBlocks) + " blocks");
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    HIPX(hipSetDevice(ctx->device));
    double* scr = noalloc_scratch(ctx->device);
    // every per-block slot [0, nxc) is written by the functional kernel and the sum slot by the
    // sum kernel, so nothing left from an earlier call is read
    launch_functional(ctx->stream, functional, ctx->prm, npts, block_size, d_rho, gga ? d_grad : nullptr, nullptr,
                      nullptr, d_weights, d_eps, d_v, gga ? d_g : nullptr, nullptr, nullptr, scr);
    if (!gga && d_g) launch_fill(ctx->stream, d_g, 3 * npts, 0.0);
    launch_sum(ctx->stream, scr, nxc, scr + kNoallocBlocks);
    HIPX(hipGetLastError());
    if (energy) {
      HIPX(hipMemcpyAsync(energy, scr + kNoallocBlocks, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
      HIPX(hipStreamSynchronize(ctx->stream));
    }
  });
}

int xc_fock_add(xc_bfogc* bf, double thr, const double* v, const double* g, double* F) {
  return guarded([&] {
    require(bf && v && F, XC_ERR_ARG, "xc_fock_add: null argument");
    require(!bf->in_flight, XC_ERR_STATE, "a server build of this controller is in flight");
    require(!g || bf->deriv >= 1, XC_ERR_STATE, "xc_fock_add: gradient potential needs AO gradients");
    require(bf->d_t64.p, XC_ERR_STATE, "xc_fock_add: the controller keeps FP32 tiles only");
    xc_ctx* ctx = bf->ctx;
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    const int n = bf->basis->nao;
    const int64_t N = bf->grid->npts;
    HIPX(hipMemcpyAsync(bf->d_v.p, v, sizeof(double) * N, hipMemcpyHostToDevice, ctx->stream));
    if (g) HIPX(hipMemcpyAsync(bf->d_g.p, g, sizeof(double) * 3 * N, hipMemcpyHostToDevice, ctx->stream));
    launch_weight_potential(ctx->stream, N, bf->grid->d_w.p, bf->d_v.p, g ? bf->d_g.p : nullptr, bf->d_vs.p,
                            bf->d_gw.p);
    run_fock<double>(bf, bf->d_t64.p, g ? 1 : 0, thr, ctx->lanes);
    HIPX(hipMemcpyAsync(bf->d_F.p, F, sizeof(double) * n * n, hipMemcpyHostToDevice, ctx->stream));
    launch_finalize(ctx->stream, bf->d_U.p, bf->d_F.p, n, 1);
    HIPX(hipGetLastError());
    HIPX(hipMemcpyAsync(F, bf->d_F.p, sizeof(double) * n * n, hipMemcpyDeviceToHost, ctx->stream));
    HIPX(hipStreamSynchronize(ctx->stream));
  });
}

// One build enqueued on stream s, E_xc left in d_energy (device).  Stage
// events are recorded only on the context stream.
static void enqueue_build(xc_bfogc* bf, int functional, int precision, double thr, const double* dP, double* dF,
                          const Lanes& L, double* d_energy) {
  xc_ctx* ctx = bf->ctx;
  hipStream_t s = L.main;
  require(!bf->in_flight, XC_ERR_STATE, "a server build of this controller is in flight");
  require(functional >= 0 && functional < kNumFunctionals, XC_ERR_ARG, "unknown functional");
  require(precision == XC_PREC_FP64 || precision == XC_PREC_FP32 || precision == XC_PREC_MIXED, XC_ERR_ARG,
          "unknown precision");
  const bool gga = functional_is_gga(functional);
  const bool mixed = precision == XC_PREC_MIXED;
  const bool lowp = precision != XC_PREC_FP64;  // FP32 contractions (mixed: plus the FP64 refinement)
  require(!gga || bf->deriv >= 1, XC_ERR_STATE, "GGA functional needs AO gradients (highest derivative 1)");
  require(!lowp || bf->d_t32.p, XC_ERR_STATE, "FP32 contractions need keep_fp32 tiles");
  require(lowp || bf->d_t64.p, XC_ERR_STATE, "FP64 contractions need FP64 tiles");
  require(!mixed || (bf->d_t64.p && bf->d_P32d.p), XC_ERR_STATE,
          "mixed precision needs both tile sets (keep_fp32 = 1)");
  const int n = bf->basis->nao;
  const int64_t N = bf->grid->npts;
  const int nxc = (int)((N + kTile - 1) / kTile);
  if (s == ctx->stream) ctx->mark();
  if (lowp) launch_to_float(s, dP, bf->d_P32.p, (int64_t)n * n);
  if (mixed) launch_to_double(s, bf->d_P32.p, bf->d_P32d.p, (int64_t)n * n);
  if (s == ctx->stream) ctx->mark();
  if (!lowp) {
    run_density<double>(bf, bf->d_t64.p, dP, gga ? 1 : 0, L);
  } else {
    if (mixed)
      run_density_mixed(bf, gga ? 1 : 0, L);
    else
      run_density<float>(bf, bf->d_t32.p, bf->d_P32.p, gga ? 1 : 0, L);
  }
  if (s == ctx->stream) ctx->mark();
  if (bf->bs == kTile) {
    launch_functional(s, functional, ctx->prm, N, kTile, nullptr, nullptr, bf->d_blocks.p, bf->d_part.p,
                      bf->grid->d_w.p, nullptr, nullptr, nullptr, bf->d_vs.p, bf->d_gw.p, bf->d_eblk.p);
  } else {
    launch_reduce_partials(s, bf->d_blocks.p, bf->bs, N, bf->d_part.p, gga ? 1 : 0, bf->d_rho.p, bf->d_grad.p);
    launch_functional(s, functional, ctx->prm, N, kTile, bf->d_rho.p, gga ? bf->d_grad.p : nullptr, nullptr,
                      nullptr, bf->grid->d_w.p, nullptr, nullptr, nullptr, bf->d_vs.p, bf->d_gw.p, bf->d_eblk.p);
  }
  if (s == ctx->stream) ctx->mark();
  if (!lowp)
    run_fock_tiles<double>(bf, bf->d_t64.p, gga ? 1 : 0, thr, L);
  else
    run_fock_tiles<float>(bf, bf->d_t32.p, gga ? 1 : 0, thr, L);
  if (s == ctx->stream) ctx->mark();
  if (!lowp)
    run_fock_reduce<double>(bf, L);
  else
    run_fock_reduce<float>(bf, L);
  if (s == ctx->stream) ctx->mark();
  // FP32 contractions see fl32(P): E_xc[P] = E_xc[fl32(P)] + tr(F (P - fl32(P))) to first order,
  // the trace formed by the finalize blocks and added with the energy sum (the mixed form's FP64
  // refinement also contracts fl32(P), so the same term applies to every block)
  const bool fp32 = lowp;
  launch_finalize(s, bf->d_U.p, dF, n, 0, fp32 ? dP : nullptr, fp32 ? bf->d_P32.p : nullptr, bf->d_round.p);
  launch_sum(s, bf->d_eblk.p, nxc, d_energy, 0, fp32 ? bf->d_round.p : nullptr, fp32 ? finalize_blocks(n) : 0);
  if (s == ctx->stream) ctx->mark();
  HIPX(hipGetLastError());
}

/* Buffers of unrestricted builds, each allocated once (never reallocated: pointers to
 * them may already be in enqueued work). */
static void ensure_pol_buffers(xc_bfogc* bf, bool fp32) {
  xc_ctx* ctx = bf->ctx;
  const int n = bf->basis->nao;
  const int64_t N = bf->grid->npts;
  if (!bf->d_rho_b.p) bf->d_rho_b.alloc(ctx, N);
  if (!bf->d_grad_b.p) bf->d_grad_b.alloc(ctx, 3 * N);
  if (!bf->d_vs_b.p) bf->d_vs_b.alloc(ctx, N);
  if (!bf->d_gw_b.p) bf->d_gw_b.alloc(ctx, 3 * N);
  if (!bf->d_Pb.p) bf->d_Pb.alloc(ctx, (size_t)n * n);
  if (!bf->d_Fb.p) bf->d_Fb.alloc(ctx, (size_t)n * n);
  if (fp32 && !bf->d_P32b.p) bf->d_P32b.alloc(ctx, (size_t)n * n);
}

/* Unrestricted build (FuncPotential<UNRESTRICTED>): spin densities from P_a and P_b,
 * the spin-polarised functional, and F_a, F_b from w v_s, w g_s (one Fock pass per
 * spin, each with its own block skip); E_xc left in d_energy. */
static void enqueue_build_pol(xc_bfogc* bf, int functional, int precision, double thr, const double* dPa,
                              const double* dPb, double* dFa, double* dFb, const Lanes& L, double* d_energy) {
  xc_ctx* ctx = bf->ctx;
  hipStream_t s = L.main;
  require(!bf->in_flight, XC_ERR_STATE, "a server build of this controller is in flight");
  require(functional >= 0 && functional < kNumFunctionals, XC_ERR_ARG, "unknown functional");
  require(precision != XC_PREC_MIXED, XC_ERR_ARG,
          "mixed precision is available for restricted single-system builds only");
  require(precision == XC_PREC_FP64 || precision == XC_PREC_FP32, XC_ERR_ARG, "unknown precision");
  const bool gga = functional_is_gga(functional);
  require(!gga || bf->deriv >= 1, XC_ERR_STATE, "GGA functional needs AO gradients (highest derivative 1)");
  require(precision == XC_PREC_FP64 || bf->d_t32.p, XC_ERR_STATE, "FP32 contractions need keep_fp32 tiles");
  require(precision == XC_PREC_FP32 || bf->d_t64.p, XC_ERR_STATE, "FP64 contractions need FP64 tiles");
  const int n = bf->basis->nao;
  const int64_t N = bf->grid->npts;
  const int nxc = (int)((N + kTile - 1) / kTile);
  ensure_pol_buffers(bf, precision == XC_PREC_FP32);
  // spin densities: one density pass per spin through the shared partial buffer
  const double* Ps[2] = {dPa, dPb};
  double* rho[2] = {bf->d_rho.p, bf->d_rho_b.p};
  double* grad[2] = {bf->d_grad.p, bf->d_grad_b.p};
  float* P32[2] = {bf->d_P32.p, bf->d_P32b.p};
  for (int sp = 0; sp < 2; ++sp) {
    if (precision == XC_PREC_FP64) {
      run_density<double>(bf, bf->d_t64.p, Ps[sp], gga ? 1 : 0, L);
    } else {
      launch_to_float(s, Ps[sp], P32[sp], (int64_t)n * n);
      run_density<float>(bf, bf->d_t32.p, P32[sp], gga ? 1 : 0, L);
    }
    launch_reduce_partials(s, bf->d_blocks.p, bf->bs, N, bf->d_part.p, gga ? 1 : 0, rho[sp], grad[sp]);
  }
  launch_functional_pol(s, functional, ctx->prm, N, kTile, bf->d_rho.p, bf->d_grad.p, bf->d_rho_b.p,
                        bf->d_grad_b.p, bf->grid->d_w.p, bf->d_vs.p, bf->d_gw.p, bf->d_vs_b.p, bf->d_gw_b.p,
                        bf->d_eblk.p);
  const double* vs[2] = {bf->d_vs.p, bf->d_vs_b.p};
  const double* gw[2] = {bf->d_gw.p, bf->d_gw_b.p};
  double* Fs[2] = {dFa, dFb};
  for (int sp = 0; sp < 2; ++sp) {
    if (precision == XC_PREC_FP64)
      run_fock<double>(bf, bf->d_t64.p, gga ? 1 : 0, thr, L, vs[sp], gw[sp]);
    else
      run_fock<float>(bf, bf->d_t32.p, gga ? 1 : 0, thr, L, vs[sp], gw[sp]);
    // with the first-order FP32 rounding term of each spin density matrix (dE_xc/dP_s = F_s)
    const bool fp32 = precision == XC_PREC_FP32;
    launch_finalize(s, bf->d_U.p, Fs[sp], n, 0, fp32 ? Ps[sp] : nullptr, fp32 ? P32[sp] : nullptr,
                    fp32 ? bf->d_round.p + (size_t)sp * finalize_blocks(n) : nullptr);
  }
  launch_sum(s, bf->d_eblk.p, nxc, d_energy, 0, precision == XC_PREC_FP32 ? bf->d_round.p : nullptr,
             precision == XC_PREC_FP32 ? 2 * finalize_blocks(n) : 0);
  HIPX(hipGetLastError());
}

/* Non-additive build: rho_sup = rho_act + sum_env rho_env on the shared grid,
 * v = v[rho_sup] - v[rho_act] integrated with the active basis,
 * E = E[rho_sup] - E[rho_act] - sum_env E[rho_env] (left in ctx->d_energy). */
static void enqueue_nadd(xc_bfogc* act, int nenv, xc_bfogc* const* env, int functional, int precision, double thr,
                         const double* dP_act, const double* const* dP_env, double* dF) {
  xc_ctx* ctx = act->ctx;
  require(functional >= 0 && functional < kNumFunctionals, XC_ERR_ARG, "unknown functional");
  require(precision != XC_PREC_MIXED, XC_ERR_ARG,
          "mixed precision is available for restricted single-system builds only");
  require(precision == XC_PREC_FP64 || precision == XC_PREC_FP32, XC_ERR_ARG, "unknown precision");
  require(nenv >= 0 && (nenv == 0 || (env && dP_env)), XC_ERR_ARG, "non-additive build: null environment list");
  const bool gga = functional_is_gga(functional);
  const bool fp32 = precision == XC_PREC_FP32;
  std::vector<xc_bfogc*> sub(1, act);
  std::vector<const double*> Ps(1, dP_act);
  for (int i = 0; i < nenv; ++i) {
    require(env[i] && dP_env[i], XC_ERR_ARG, "non-additive build: null environment controller or density");
    require(env[i] != act, XC_ERR_ARG, "non-additive build: the active controller is also listed as environment");
    for (int j = 0; j < i; ++j)
      require(env[j] != env[i], XC_ERR_ARG, "non-additive build: an environment controller is listed twice");
    sub.push_back(env[i]);
    Ps.push_back(dP_env[i]);
  }
  for (xc_bfogc* b : sub) {
    require(b->ctx == ctx, XC_ERR_ARG, "non-additive build: controllers live on different contexts");
    require(b->grid == act->grid, XC_ERR_ARG,
            "non-additive build: all subsystem controllers must be defined on the same (supersystem) grid");
    require(b->bs == act->bs, XC_ERR_ARG, "non-additive build: controllers differ in block size");
    require_idle(b, "non-additive build");
    require(!gga || b->deriv >= 1, XC_ERR_STATE, "GGA functional needs AO gradients (highest derivative 1)");
    require(!fp32 || b->d_t32.p, XC_ERR_STATE, "FP32 contractions need keep_fp32 tiles");
    require(fp32 || b->d_t64.p, XC_ERR_STATE, "FP64 contractions need FP64 tiles");
  }
  const int n = act->basis->nao;
  const int64_t N = act->grid->npts;
  const int nxc = (int)((N + kTile - 1) / kTile);
  if (!act->d_rho_sup.p) {
    act->d_rho_sup.alloc(ctx, N);
    act->d_grad_sup.alloc(ctx, 3 * N);
  }
  hipStream_t s = ctx->stream;
  const double* w = act->grid->d_w.p;
  ctx->mark();
  if (fp32)
    for (size_t i = 0; i < sub.size(); ++i)
      launch_to_float(s, Ps[i], sub[i]->d_P32.p, (int64_t)sub[i]->basis->nao * sub[i]->basis->nao);
  ctx->mark();
  for (size_t i = 0; i < sub.size(); ++i) {
    xc_bfogc* b = sub[i];
    if (fp32)
      run_density<float>(b, b->d_t32.p, b->d_P32.p, gga ? 1 : 0, ctx->lanes);
    else
      run_density<double>(b, b->d_t64.p, Ps[i], gga ? 1 : 0, ctx->lanes);
    launch_reduce_partials(s, b->d_blocks.p, b->bs, N, b->d_part.p, gga ? 1 : 0, b->d_rho.p, b->d_grad.p, 0);
    launch_reduce_partials(s, b->d_blocks.p, b->bs, N, b->d_part.p, gga ? 1 : 0, act->d_rho_sup.p,
                           act->d_grad_sup.p, i == 0 ? 0 : 1);
  }
  ctx->mark();
  launch_functional(s, functional, ctx->prm, N, kTile, act->d_rho_sup.p, gga ? act->d_grad_sup.p : nullptr, nullptr,
                    nullptr, w, nullptr, nullptr, nullptr, act->d_vs.p, act->d_gw.p, act->d_eblk.p, act->d_rho.p,
                    gga ? act->d_grad.p : nullptr);
  launch_sum(s, act->d_eblk.p, nxc, ctx->d_energy, 0);
  for (size_t i = 1; i < sub.size(); ++i) {
    xc_bfogc* b = sub[i];
    launch_functional(s, functional, ctx->prm, N, kTile, b->d_rho.p, gga ? b->d_grad.p : nullptr, nullptr, nullptr, w,
                      nullptr, nullptr, nullptr, nullptr, nullptr, b->d_eblk.p);
    launch_sum(s, b->d_eblk.p, nxc, ctx->d_energy, 1);
  }
  ctx->mark();
  if (fp32)
    run_fock_tiles<float>(act, act->d_t32.p, gga ? 1 : 0, thr, ctx->lanes);
  else
    run_fock_tiles<double>(act, act->d_t64.p, gga ? 1 : 0, thr, ctx->lanes);
  ctx->mark();
  if (fp32)
    run_fock_reduce<float>(act, ctx->lanes);
  else
    run_fock_reduce<double>(act, ctx->lanes);
  ctx->mark();
  launch_finalize(s, act->d_U.p, dF, n, 0);
  ctx->mark();
  HIPX(hipGetLastError());
}

/* Unrestricted non-additive build (NAddFuncPotential<UNRESTRICTED>::getMatrix): spin
 * densities of every subsystem on the shared grid, rho_sup^s = sum of them,
 * v^s = v^s[rho_sup] - v^s[rho_act] integrated with the active basis per spin,
 * E = E[sup] - E[act] - sum_env E[env] (left in ctx->d_energy). */
static void enqueue_nadd_pol(xc_bfogc* act, int nenv, xc_bfogc* const* env, int functional, int precision,
                             double thr, const double* dPa_act, const double* dPb_act, const double* const* dPa_env,
                             const double* const* dPb_env, double* dFa, double* dFb) {
  xc_ctx* ctx = act->ctx;
  require(functional >= 0 && functional < kNumFunctionals, XC_ERR_ARG, "unknown functional");
  require(precision != XC_PREC_MIXED, XC_ERR_ARG,
          "mixed precision is available for restricted single-system builds only");
  require(precision == XC_PREC_FP64 || precision == XC_PREC_FP32, XC_ERR_ARG, "unknown precision");
  require(nenv >= 0 && (nenv == 0 || (env && dPa_env && dPb_env)), XC_ERR_ARG,
          "non-additive build: null environment list");
  const bool gga = functional_is_gga(functional);
  const bool fp32 = precision == XC_PREC_FP32;
  std::vector<xc_bfogc*> sub(1, act);
  std::vector<const double*> Pa(1, dPa_act), Pb(1, dPb_act);
  for (int i = 0; i < nenv; ++i) {
    require(env[i] && dPa_env[i] && dPb_env[i], XC_ERR_ARG,
            "non-additive build: null environment controller or density");
    require(env[i] != act, XC_ERR_ARG, "non-additive build: the active controller is also listed as environment");
    for (int j = 0; j < i; ++j)
      require(env[j] != env[i], XC_ERR_ARG, "non-additive build: an environment controller is listed twice");
    sub.push_back(env[i]);
    Pa.push_back(dPa_env[i]);
    Pb.push_back(dPb_env[i]);
  }
  for (xc_bfogc* b : sub) {
    require(b->ctx == ctx, XC_ERR_ARG, "non-additive build: controllers live on different contexts");
    require(b->grid == act->grid, XC_ERR_ARG,
            "non-additive build: all subsystem controllers must be defined on the same (supersystem) grid");
    require(b->bs == act->bs, XC_ERR_ARG, "non-additive build: controllers differ in block size");
    require(!b->in_flight, XC_ERR_STATE, "a server build of this controller is in flight");
    require(!gga || b->deriv >= 1, XC_ERR_STATE, "GGA functional needs AO gradients (highest derivative 1)");
    require(!fp32 || b->d_t32.p, XC_ERR_STATE, "FP32 contractions need keep_fp32 tiles");
    require(fp32 || b->d_t64.p, XC_ERR_STATE, "FP64 contractions need FP64 tiles");
    ensure_pol_buffers(b, fp32);
  }
  const int n = act->basis->nao;
  const int64_t N = act->grid->npts;
  const int nxc = (int)((N + kTile - 1) / kTile);
  if (!act->d_rho_sup.p) {
    act->d_rho_sup.alloc(ctx, N);
    act->d_grad_sup.alloc(ctx, 3 * N);
  }
  if (!act->d_rho_sup_b.p) {
    act->d_rho_sup_b.alloc(ctx, N);
    act->d_grad_sup_b.alloc(ctx, 3 * N);
  }
  hipStream_t s = ctx->stream;
  const double* w = act->grid->d_w.p;
  const int g = gga ? 1 : 0;
  double* sup_r[2] = {act->d_rho_sup.p, act->d_rho_sup_b.p};
  double* sup_g[2] = {act->d_grad_sup.p, act->d_grad_sup_b.p};
  for (size_t i = 0; i < sub.size(); ++i) {
    xc_bfogc* b = sub[i];
    const double* P[2] = {Pa[i], Pb[i]};
    float* P32[2] = {b->d_P32.p, b->d_P32b.p};
    double* rho[2] = {b->d_rho.p, b->d_rho_b.p};
    double* grd[2] = {b->d_grad.p, b->d_grad_b.p};
    for (int sp = 0; sp < 2; ++sp) {
      if (fp32) {
        launch_to_float(s, P[sp], P32[sp], (int64_t)b->basis->nao * b->basis->nao);
        run_density<float>(b, b->d_t32.p, P32[sp], g, ctx->lanes);
      } else {
        run_density<double>(b, b->d_t64.p, P[sp], g, ctx->lanes);
      }
      launch_reduce_partials(s, b->d_blocks.p, b->bs, N, b->d_part.p, g, rho[sp], grd[sp], 0);
      launch_reduce_partials(s, b->d_blocks.p, b->bs, N, b->d_part.p, g, sup_r[sp], sup_g[sp], i == 0 ? 0 : 1);
    }
  }
  // v^s[sup] - v^s[act] and E[sup] - E[act]
  launch_functional_pol(s, functional, ctx->prm, N, kTile, sup_r[0], sup_g[0], sup_r[1], sup_g[1], w, act->d_vs.p,
                        act->d_gw.p, act->d_vs_b.p, act->d_gw_b.p, act->d_eblk.p, act->d_rho.p, act->d_grad.p,
                        act->d_rho_b.p, act->d_grad_b.p);
  launch_sum(s, act->d_eblk.p, nxc, ctx->d_energy, 0);
  for (size_t i = 1; i < sub.size(); ++i) {
    xc_bfogc* b = sub[i];
    launch_functional_pol(s, functional, ctx->prm, N, kTile, b->d_rho.p, b->d_grad.p, b->d_rho_b.p, b->d_grad_b.p, w,
                          nullptr, nullptr, nullptr, nullptr, b->d_eblk.p);
    launch_sum(s, b->d_eblk.p, nxc, ctx->d_energy, 1);
  }
  const double* vs[2] = {act->d_vs.p, act->d_vs_b.p};
  const double* gw[2] = {act->d_gw.p, act->d_gw_b.p};
  double* Fs[2] = {dFa, dFb};
  for (int sp = 0; sp < 2; ++sp) {
    if (fp32)
      run_fock<float>(act, act->d_t32.p, g, thr, ctx->lanes, vs[sp], gw[sp]);
    else
      run_fock<double>(act, act->d_t64.p, g, thr, ctx->lanes, vs[sp], gw[sp]);
    launch_finalize(s, act->d_U.p, Fs[sp], n, 0);
  }
  HIPX(hipGetLastError());
}

int xc_nadd_fock_build_unrestricted(xc_bfogc* active, int nenv, xc_bfogc* const* env, int functional, int precision,
                                    double thr, const double* Pa_active, const double* Pb_active,
                                    const double* const* Pa_env, const double* const* Pb_env, double* Fa, double* Fb,
                                    double* energy) {
  return guarded([&] {
    require(active && Pa_active && Pb_active && Fa && Fb, XC_ERR_ARG, "xc_nadd_fock_build_unrestricted: null argument");
    require(nenv >= 0 && (nenv == 0 || (env && Pa_env && Pb_env)), XC_ERR_ARG,
            "xc_nadd_fock_build_unrestricted: null environment list");
    xc_ctx* ctx = active->ctx;
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    require_idle(active, nenv, env, "xc_nadd_fock_build_unrestricted");
    std::vector<xc_bfogc*> sub(1, active);
    for (int i = 0; i < nenv; ++i) {
      require(env[i] && Pa_env[i] && Pb_env[i], XC_ERR_ARG,
              "xc_nadd_fock_build_unrestricted: null environment controller or density");
      require(env[i]->ctx == ctx, XC_ERR_ARG, "xc_nadd_fock_build_unrestricted: controllers live on different contexts");
      sub.push_back(env[i]);
    }
    // host spin densities into each controller's P (alpha) and Pb (beta) buffers
    std::vector<const double*> dPa(nenv), dPb(nenv);
    for (size_t i = 0; i < sub.size(); ++i) {
      xc_bfogc* b = sub[i];
      ensure_pol_buffers(b, precision == XC_PREC_FP32);
      const int m = b->basis->nao;
      const double* ha = i == 0 ? Pa_active : Pa_env[i - 1];
      const double* hb = i == 0 ? Pb_active : Pb_env[i - 1];
      HIPX(hipMemcpyAsync(b->d_P.p, ha, sizeof(double) * m * m, hipMemcpyHostToDevice, ctx->stream));
      HIPX(hipMemcpyAsync(b->d_Pb.p, hb, sizeof(double) * m * m, hipMemcpyHostToDevice, ctx->stream));
      if (i > 0) {
        dPa[i - 1] = b->d_P.p;
        dPb[i - 1] = b->d_Pb.p;
      }
    }
    enqueue_nadd_pol(active, nenv, env, functional, precision, thr, active->d_P.p, active->d_Pb.p, dPa.data(),
                     dPb.data(), active->d_F.p, active->d_Fb.p);
    const int n = active->basis->nao;
    HIPX(hipMemcpyAsync(Fa, active->d_F.p, sizeof(double) * n * n, hipMemcpyDeviceToHost, ctx->stream));
    HIPX(hipMemcpyAsync(Fb, active->d_Fb.p, sizeof(double) * n * n, hipMemcpyDeviceToHost, ctx->stream));
    HIPX(hipMemcpyAsync(&ctx->last_energy, ctx->d_energy, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HIPX(hipStreamSynchronize(ctx->stream));
    if (energy) *energy = ctx->last_energy;
  });
}

int xc_bfogc_update_density(xc_bfogc* bf, int precision, const double* dP, int with_gradient) {
  return guarded([&] {
    require(bf && dP, XC_ERR_ARG, "xc_bfogc_update_density: null argument");
    require(precision != XC_PREC_MIXED, XC_ERR_ARG,
          "mixed precision is available for restricted single-system builds only");
  require(precision == XC_PREC_FP64 || precision == XC_PREC_FP32, XC_ERR_ARG, "unknown precision");
    require(!with_gradient || bf->deriv >= 1, XC_ERR_STATE,
            "xc_bfogc_update_density: gradient requested but AO gradients not evaluated");
    require(precision == XC_PREC_FP64 || bf->d_t32.p, XC_ERR_STATE, "FP32 contractions need keep_fp32 tiles");
    require(precision == XC_PREC_FP32 || bf->d_t64.p, XC_ERR_STATE, "FP64 contractions need FP64 tiles");
    xc_ctx* ctx = bf->ctx;
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    require_idle(bf, "xc_bfogc_update_density");
    const int n = bf->basis->nao;
    const int64_t N = bf->grid->npts;
    if (!bf->d_rho_c.p) {
      bf->d_rho_c.alloc(ctx, N);
      bf->d_grad_c.alloc(ctx, 3 * N);
      bf->d_ecache.alloc(ctx, kNumFunctionals);
    }
    const int g = with_gradient ? 1 : 0;
    if (precision == XC_PREC_FP32) {
      launch_to_float(ctx->stream, dP, bf->d_P32.p, (int64_t)n * n);
      run_density<float>(bf, bf->d_t32.p, bf->d_P32.p, g, ctx->lanes);
    } else {
      run_density<double>(bf, bf->d_t64.p, dP, g, ctx->lanes);
    }
    launch_reduce_partials(ctx->stream, bf->d_blocks.p, bf->bs, N, bf->d_part.p, g, bf->d_rho_c.p, bf->d_grad_c.p,
                           0);
    HIPX(hipGetLastError());
    bf->rho_c_valid = true;
    bf->rho_c_grad = with_gradient != 0;
    for (bool& v : bf->ecache_valid) v = false;
  });
}

int xc_nadd_fock_build_cached(xc_bfogc* act, int nenv, xc_bfogc* const* env, int functional, int precision,
                              double thr, double* dF, double* energy) {
  return guarded([&] {
    require(act && dF, XC_ERR_ARG, "xc_nadd_fock_build_cached: null argument");
    require(functional >= 0 && functional < kNumFunctionals, XC_ERR_ARG, "unknown functional");
    require(precision != XC_PREC_MIXED, XC_ERR_ARG,
          "mixed precision is available for restricted single-system builds only");
  require(precision == XC_PREC_FP64 || precision == XC_PREC_FP32, XC_ERR_ARG, "unknown precision");
    require(nenv >= 0 && (nenv == 0 || env), XC_ERR_ARG, "non-additive build: null environment list");
    xc_ctx* ctx = act->ctx;
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    const bool gga = functional_is_gga(functional);
    require(!gga || act->deriv >= 1, XC_ERR_STATE, "GGA functional needs AO gradients (highest derivative 1)");
    require(precision == XC_PREC_FP64 || act->d_t32.p, XC_ERR_STATE, "FP32 contractions need keep_fp32 tiles");
    require(precision == XC_PREC_FP32 || act->d_t64.p, XC_ERR_STATE, "FP64 contractions need FP64 tiles");
    std::vector<xc_bfogc*> sub(1, act);
    for (int i = 0; i < nenv; ++i) {
      require(env[i], XC_ERR_ARG, "non-additive build: null environment controller");
      require(env[i] != act, XC_ERR_ARG, "non-additive build: the active controller is also listed as environment");
      for (int j = 0; j < i; ++j)
        require(env[j] != env[i], XC_ERR_ARG, "non-additive build: an environment controller is listed twice");
      sub.push_back(env[i]);
    }
    for (xc_bfogc* b : sub) {
      require(b->ctx == ctx, XC_ERR_ARG, "non-additive build: controllers live on different contexts");
      require(b->grid == act->grid, XC_ERR_ARG,
              "non-additive build: all subsystem controllers must be defined on the same (supersystem) grid");
      require_idle(b, "xc_nadd_fock_build_cached");
      require(b->rho_c_valid, XC_ERR_STATE,
              "xc_nadd_fock_build_cached: a subsystem has no cached density (call xc_bfogc_update_density)");
      require(!gga || b->rho_c_grad, XC_ERR_STATE,
              "xc_nadd_fock_build_cached: GGA functional needs cached density gradients");
    }
    const int n = act->basis->nao;
    const int64_t N = act->grid->npts;
    const int nxc = (int)((N + kTile - 1) / kTile);
    if (!act->d_rho_sup.p) {
      act->d_rho_sup.alloc(ctx, N);
      act->d_grad_sup.alloc(ctx, 3 * N);
    }
    hipStream_t s = ctx->stream;
    const double* w = act->grid->d_w.p;
    ctx->mark();
    ctx->mark();
    // rho_sup = rho_act + rho_env[0] + ... in list order, kListMax arrays per launch
    for (size_t i0 = 0; i0 < sub.size(); i0 += kListMax) {
      DensityList L{};
      L.n = (int)std::min<size_t>(kListMax, sub.size() - i0);
      for (int k = 0; k < L.n; ++k) {
        L.rho[k] = sub[i0 + k]->d_rho_c.p;
        L.grad[k] = sub[i0 + k]->d_grad_c.p;
      }
      launch_sum_densities(s, L, N, gga ? 1 : 0, i0 == 0 ? 0 : 1, act->d_rho_sup.p, act->d_grad_sup.p);
    }
    ctx->mark();
    launch_functional(s, functional, ctx->prm, N, kTile, act->d_rho_sup.p, gga ? act->d_grad_sup.p : nullptr, nullptr,
                      nullptr, w, nullptr, nullptr, nullptr, act->d_vs.p, act->d_gw.p, act->d_eblk.p, act->d_rho_c.p,
                      gga ? act->d_grad_c.p : nullptr);
    launch_sum(s, act->d_eblk.p, nxc, ctx->d_energy, 0);
    // E[rho_env] per environment subsystem, kept until its density changes
    for (size_t i = 1; i < sub.size(); ++i) {
      xc_bfogc* b = sub[i];
      if (b->ecache_valid[functional]) continue;
      launch_functional(s, functional, ctx->prm, N, kTile, b->d_rho_c.p, gga ? b->d_grad_c.p : nullptr, nullptr,
                        nullptr, w, nullptr, nullptr, nullptr, nullptr, nullptr, b->d_eblk.p);
      launch_sum(s, b->d_eblk.p, nxc, b->d_ecache.p + functional, 0);
      b->ecache_valid[functional] = true;
    }
    for (size_t i0 = 1; i0 < sub.size(); i0 += kListMax) {
      DensityList L{};
      L.n = (int)std::min<size_t>(kListMax, sub.size() - i0);
      for (int k = 0; k < L.n; ++k) L.rho[k] = sub[i0 + k]->d_ecache.p + functional;
      launch_subtract_scalars(s, L, ctx->d_energy);
    }
    ctx->mark();
    if (precision == XC_PREC_FP32)
      run_fock_tiles<float>(act, act->d_t32.p, gga ? 1 : 0, thr, ctx->lanes);
    else
      run_fock_tiles<double>(act, act->d_t64.p, gga ? 1 : 0, thr, ctx->lanes);
    ctx->mark();
    if (precision == XC_PREC_FP32)
      run_fock_reduce<float>(act, ctx->lanes);
    else
      run_fock_reduce<double>(act, ctx->lanes);
    ctx->mark();
    launch_finalize(s, act->d_U.p, dF, n, 0);
    ctx->mark();
    HIPX(hipGetLastError());
    if (energy) {
      HIPX(hipMemcpyAsync(&ctx->last_energy, ctx->d_energy, sizeof(double), hipMemcpyDeviceToHost, s));
      HIPX(hipStreamSynchronize(s));
      *energy = ctx->last_energy;
    }
  });
}

int xc_nadd_fock_build_device(xc_bfogc* active, int nenv, xc_bfogc* const* env, int functional, int precision,
                              double thr, const double* dP_active, const double* const* dP_env, double* dF,
                              double* energy) {
  return guarded([&] {
    require(active && dP_active && dF, XC_ERR_ARG, "xc_nadd_fock_build_device: null argument");
    xc_ctx* ctx = active->ctx;
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    require_idle(active, nenv, env, "xc_nadd_fock_build_device");
    enqueue_nadd(active, nenv, env, functional, precision, thr, dP_active, dP_env, dF);
    if (energy) {
      HIPX(hipMemcpyAsync(&ctx->last_energy, ctx->d_energy, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
      HIPX(hipStreamSynchronize(ctx->stream));
      *energy = ctx->last_energy;
    }
  });
}

int xc_nadd_fock_build(xc_bfogc* active, int nenv, xc_bfogc* const* env, int functional, int precision, double thr,
                       const double* P_active, const double* const* P_env, double* F, double* energy) {
  return guarded([&] {
    require(active && P_active && F, XC_ERR_ARG, "xc_nadd_fock_build: null argument");
    require(nenv >= 0 && (nenv == 0 || (env && P_env)), XC_ERR_ARG, "xc_nadd_fock_build: null environment list");
    xc_ctx* ctx = active->ctx;
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    require_idle(active, nenv, env, "xc_nadd_fock_build");
    const int n = active->basis->nao;
    HIPX(hipMemcpyAsync(active->d_P.p, P_active, sizeof(double) * n * n, hipMemcpyHostToDevice, ctx->stream));
    std::vector<const double*> dP(nenv);
    for (int i = 0; i < nenv; ++i) {
      require(env[i] && P_env[i], XC_ERR_ARG, "xc_nadd_fock_build: null environment controller or density");
      require(env[i]->ctx == ctx, XC_ERR_ARG, "xc_nadd_fock_build: controllers live on different contexts");
      const int m = env[i]->basis->nao;
      HIPX(hipMemcpyAsync(env[i]->d_P.p, P_env[i], sizeof(double) * m * m, hipMemcpyHostToDevice, ctx->stream));
      dP[i] = env[i]->d_P.p;
    }
    enqueue_nadd(active, nenv, env, functional, precision, thr, active->d_P.p, dP.data(), active->d_F.p);
    HIPX(hipMemcpyAsync(F, active->d_F.p, sizeof(double) * n * n, hipMemcpyDeviceToHost, ctx->stream));
    HIPX(hipMemcpyAsync(&ctx->last_energy, ctx->d_energy, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HIPX(hipStreamSynchronize(ctx->stream));
    if (energy) *energy = ctx->last_energy;
  });
}

int xc_fock_build_device(xc_bfogc* bf, int functional, int precision, double thr, const double* dP, double* dF,
                         double* energy) {
  return guarded([&] {
    require(bf && dP && dF, XC_ERR_ARG, "xc_fock_build_device: null argument");
    xc_ctx* ctx = bf->ctx;
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    enqueue_build(bf, functional, precision, thr, dP, dF, ctx->lanes, ctx->d_energy);
    if (energy) {
      HIPX(hipMemcpyAsync(&ctx->last_energy, ctx->d_energy, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
      HIPX(hipStreamSynchronize(ctx->stream));
      *energy = ctx->last_energy;
    }
  });
}

int xc_fock_build_packed(int nsys, xc_bfogc* const* bfs, int functional, int precision, double thr,
                         const double* const* dP, double* const* dF, double* energies) {
  return guarded([&] {
    require(nsys >= 1 && bfs && dP && dF && energies, XC_ERR_ARG, "xc_fock_build_packed: bad arguments");
    xc_ctx* ctx = bfs[0] ? bfs[0]->ctx : nullptr;
    require(ctx != nullptr, XC_ERR_ARG, "xc_fock_build_packed: null controller");
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    require(functional >= 0 && functional < kNumFunctionals, XC_ERR_ARG, "unknown functional");
    require(precision != XC_PREC_MIXED, XC_ERR_ARG,
          "mixed precision is available for restricted single-system builds only");
  require(precision == XC_PREC_FP64 || precision == XC_PREC_FP32, XC_ERR_ARG, "unknown precision");
    const bool gga = functional_is_gga(functional);
    const bool fp32 = precision == XC_PREC_FP32;
    std::vector<DevBlock> blocks;
    std::vector<PackSys> sys(nsys);
    std::vector<WorkItem> dens[kDensMaxS + 1];
    std::vector<FockItem> fock[kFockNumS];
    int64_t maxnn = 0;
    int maxn = 0;
    for (int i = 0; i < nsys; ++i) {
      xc_bfogc* bf = bfs[i];
      require(bf && dP[i] && dF[i], XC_ERR_ARG, "xc_fock_build_packed: null controller or matrix");
      require(bf->ctx == ctx, XC_ERR_ARG, "xc_fock_build_packed: controllers of different contexts");
      for (int j = 0; j < i; ++j)
        require(bfs[j] != bf, XC_ERR_ARG, "xc_fock_build_packed: a controller appears twice (shared scratch)");
      require(!bf->in_flight, XC_ERR_STATE, "a server build of this controller is in flight");
      require(bf->bs == kTile, XC_ERR_STATE, "xc_fock_build_packed: controllers need 128-point blocks");
      require(!gga || bf->deriv >= 1, XC_ERR_STATE, "GGA functional needs AO gradients (highest derivative 1)");
      require(!fp32 || bf->d_t32.p, XC_ERR_STATE, "FP32 contractions need keep_fp32 tiles");
      require(fp32 || bf->d_t64.p, XC_ERR_STATE, "FP64 contractions need FP64 tiles");
      const int n = bf->basis->nao;
      const int blk0 = (int)blocks.size();
      for (DevBlock b : bf->blocks) {
        b.sys = i;
        blocks.push_back(b);
      }
      for (int c = 1; c <= kDensMaxS; ++c)
        for (WorkItem w : bf->dens_work[c]) {
          w.block += blk0;
          dens[c].push_back(w);
        }
      for (int c = 0; c < kFockNumS; ++c)
        for (FockItem f : bf->fock_items[fp32 ? 0 : 1][c]) {
          f.block += blk0;
          fock[c].push_back(f);
        }
      PackSys& ps = sys[i];
      ps.tiles = fp32 ? (const void*)bf->d_t32.p : (const void*)bf->d_t64.p;
      ps.P = fp32 ? (const void*)bf->d_P32.p : (const void*)dP[i];
      ps.P64 = dP[i];
      ps.P32 = fp32 ? bf->d_P32.p : nullptr;
      ps.ao_idx = bf->d_ao_idx.p;
      ps.part = bf->d_part.p;
      ps.w = bf->grid->d_w.p;
      ps.vs = bf->d_vs.p;
      ps.gw = bf->d_gw.p;
      ps.U = bf->d_U.p;
      ps.T = bf->d_T.p;
      ps.ent = bf->d_rowent.p;
      ps.row_ptr = bf->d_rowptr.p;
      ps.F = dF[i];
      ps.eblk = bf->d_eblk.p;
      ps.npts = bf->grid->npts;
      ps.n = n;
      ps.blk0 = blk0;
      ps.nxc = (int)((bf->grid->npts + kTile - 1) / kTile);
      maxnn = std::max(maxnn, (int64_t)n * n);
      maxn = std::max(maxn, n);
    }
    // a composition seen once before runs as a captured graph; the first build is eager
    std::vector<uint64_t> key;
    for (int i = 0; i < nsys; ++i) key.push_back((uint64_t)(uintptr_t)bfs[i]);
    uint64_t thr_bits = 0;
    std::memcpy(&thr_bits, &thr, sizeof(thr_bits));
    key.push_back((uint64_t)functional);
    key.push_back((uint64_t)precision);
    key.push_back(thr_bits);
    // graphs pay where launches are short (many small systems: 4.1 vs 4.3 ms for 32 small
    // molecules); a graph runs the class launches of a stage one after another instead of on
    // the four lanes, which costs more than it saves for packs of large grids (25 FDE subsystems
    // on a 353 701-point grid: 3.5 vs 2.7 ms), so large packs stay eager
    const bool graphable = blocks.size() <= 32768;
    auto hit = graphable ? ctx->pack_graphs.find(key) : ctx->pack_graphs.end();
    bool capture = false;
    PackWork* pwp = nullptr;
    if (!graphable) {
      if (!ctx->pack) ctx->pack = new PackWork();
      pwp = ctx->pack;
    } else if (hit == ctx->pack_graphs.end()) {
      if (ctx->pack_graphs.size() >= 16) {  // bounded cache: forget the least recently used composition
        auto lru = ctx->pack_graphs.begin();
        for (auto it = ctx->pack_graphs.begin(); it != ctx->pack_graphs.end(); ++it)
          if (it->second->last_use < lru->second->last_use) lru = it;
        delete lru->second;
        ctx->pack_graphs.erase(lru);
      }
      ctx->pack_graphs.emplace(key, new PackWork()).first->second->last_use = ++ctx->pack_calls;
      if (!ctx->pack) ctx->pack = new PackWork();
      pwp = ctx->pack;
    } else {
      pwp = hit->second;
      pwp->last_use = ++ctx->pack_calls;
      capture = pwp->exec == nullptr;
    }
    PackWork& pw = *pwp;
    if (pw.round.n < (size_t)nsys * kPackRoundBlocks) pw.round.alloc(ctx, (size_t)2 * nsys * kPackRoundBlocks);
    if (pw.E.n < (size_t)nsys) pw.E.alloc(ctx, (size_t)2 * nsys);
    for (int i = 0; i < nsys; ++i) {
      sys[i].round = pw.round.p + (size_t)i * kPackRoundBlocks;
      sys[i].E = pw.E.p + i;
    }
    hipStream_t s = ctx->stream;
    const Lanes& L = ctx->lanes;
    const int g = gga ? 1 : 0;
    // the per-system table (matrices of this call) is uploaded every time; a replayed graph reads it
    upload_grow(ctx, pw.sys, sys);
    if (pw.exec) {
      HIPX(hipGraphLaunch(pw.exec, s));
    } else {
      upload_grow(ctx, pw.blocks, blocks);
      for (int c = 1; c <= kDensMaxS; ++c) upload_grow(ctx, pw.dens[c], dens[c]);
      for (int c = 0; c < kFockNumS; ++c) upload_grow(ctx, pw.fock[c], fock[c]);
      const PackSys* dsys = pw.sys.p;
      // one launch per stage and work class for every system of the pack
      auto enqueue = [&]() {
        if (fp32) launch_pack_to_float(s, dsys, nsys, maxnn);
        over_lanes(L, kDensMaxS + 1, [&](int c) { return c >= 1 ? (int)dens[c].size() : 0; },
                   [&](int c, hipStream_t st) {
          if (fp32)
            launch_density_sub<float>(st, c, pw.dens[c].p, (int)dens[c].size(), pw.blocks.p, nullptr, nullptr,
                                      nullptr, 0, g, nullptr, 0, dsys);
          else
            launch_density_sub<double>(st, c, pw.dens[c].p, (int)dens[c].size(), pw.blocks.p, nullptr, nullptr,
                                       nullptr, 0, g, nullptr, 0, dsys);
        });
        launch_functional_packed(s, functional, ctx->prm, (int)blocks.size(), pw.blocks.p, dsys);
        over_lanes(L, kFockNumS, [&](int c) { return (int)fock[c].size(); }, [&](int c, hipStream_t st) {
          if (fp32)
            launch_fock_sub<float>(st, c, pw.fock[c].p, (int)fock[c].size(), pw.blocks.p, nullptr, nullptr,
                                   nullptr, nullptr, g, thr, nullptr, 0, dsys);
          else
            launch_fock_sub<double>(st, c, pw.fock[c].p, (int)fock[c].size(), pw.blocks.p, nullptr, nullptr,
                                    nullptr, nullptr, g, thr, nullptr, 0, dsys);
        });
        if (fp32)
          launch_fock_reduce<float>(s, nullptr, nullptr, nullptr, nullptr, nullptr, maxn, dsys, nsys,
                                    ctx->deterministic);
        else
          launch_fock_reduce<double>(s, nullptr, nullptr, nullptr, nullptr, nullptr, maxn, dsys, nsys,
                                     ctx->deterministic);
        launch_pack_finalize(s, dsys, nsys, fp32 ? 1 : 0);
        launch_pack_sum(s, dsys, nsys, fp32 ? 1 : 0);
      };
      if (capture) {
        // second build of this composition: capture it (kernel nodes only, no memsets) into its own lists
        hipGraph_t graph = nullptr;
        HIPX(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        try {
          enqueue();
        } catch (...) {
          (void)hipStreamEndCapture(s, &graph);
          if (graph) (void)hipGraphDestroy(graph);
          throw;
        }
        HIPX(hipStreamEndCapture(s, &graph));
        hipGraphExec_t exec = nullptr;
        const hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        HIPX(e);
        pw.exec = exec;
        HIPX(hipGraphLaunch(pw.exec, s));
      } else {
        enqueue();
      }
    }
    HIPX(hipGetLastError());
    HIPX(hipMemcpyAsync(energies, pw.E.p, sizeof(double) * nsys, hipMemcpyDeviceToHost, s));
    HIPX(hipStreamSynchronize(s));
    ctx->last_energy = energies[nsys - 1];
  });
}

namespace {
// XC gradient of one (restricted: Pb == nullptr) or two spin density matrices (host).
void geom_gradient(xc_bfogc* bf, int functional, const double* Pa, const double* Pb, int natoms, const int* ao_atom,
                   double* grad, double* energy) {
  require(bf && Pa && grad && ao_atom && natoms > 0, XC_ERR_ARG, "xc_geom_gradient: bad arguments");
  require(!bf->in_flight, XC_ERR_STATE, "a server build of this controller is in flight");
  require(functional >= 0 && functional < kNumFunctionals, XC_ERR_ARG, "unknown functional");
  const bool gga = functional_is_gga(functional);
  require(bf->deriv >= (gga ? 2 : 1), XC_ERR_STATE,
          gga ? "xc_geom_gradient: a GGA gradient needs AO second derivatives (highest derivative 2)"
              : "xc_geom_gradient: needs AO gradients (highest derivative 1)");
  require(bf->d_t64.p, XC_ERR_STATE, "xc_geom_gradient: needs FP64 tiles");
  xc_ctx* ctx = bf->ctx;
  std::lock_guard<std::recursive_mutex> lk(ctx->mu);
  hipStream_t s = ctx->stream;
  const int n = bf->basis->nao;
  const int64_t N = bf->grid->npts;
  const int nxc = (int)((N + kTile - 1) / kTile);
  for (int i = 0; i < n; ++i)
    require(ao_atom[i] >= 0 && ao_atom[i] < natoms, XC_ERR_INPUT, "xc_geom_gradient: AO atom index out of range");
  if (bf->grad_work.empty()) {
    for (int b = 0; b < bf->nblocks; ++b)
      for (int c0 = 0; c0 < bf->blocks[b].npad; c0 += kGradCols) bf->grad_work.push_back(GradItem{b, c0});
    if (!bf->grad_work.empty()) bf->d_grad_work.upload(ctx, bf->grad_work.data(), bf->grad_work.size());
    bf->d_gacc.alloc(ctx, (size_t)3 * std::max(n, 1));
    bf->d_ao_atom.alloc(ctx, (size_t)std::max(n, 1));
  }
  if ((int64_t)bf->d_gout.n < 3 * (int64_t)natoms) bf->d_gout.alloc(ctx, (size_t)3 * natoms);
  if (Pb) ensure_pol_buffers(bf, false);
  HIPX(hipMemcpyAsync(bf->d_ao_atom.p, ao_atom, sizeof(int) * n, hipMemcpyHostToDevice, s));
  HIPX(hipMemcpyAsync(bf->d_P.p, Pa, sizeof(double) * n * n, hipMemcpyHostToDevice, s));
  if (Pb) HIPX(hipMemcpyAsync(bf->d_Pb.p, Pb, sizeof(double) * n * n, hipMemcpyHostToDevice, s));
  // density and functional as in a Fock build (FP64), then the gradient contractions
  if (!Pb) {
    run_density<double>(bf, bf->d_t64.p, bf->d_P.p, gga ? 1 : 0, ctx->lanes);
    if (bf->bs == kTile) {
      launch_functional(s, functional, ctx->prm, N, kTile, nullptr, nullptr, bf->d_blocks.p, bf->d_part.p,
                        bf->grid->d_w.p, nullptr, nullptr, nullptr, bf->d_vs.p, bf->d_gw.p, bf->d_eblk.p);
    } else {
      launch_reduce_partials(s, bf->d_blocks.p, bf->bs, N, bf->d_part.p, gga ? 1 : 0, bf->d_rho.p, bf->d_grad.p);
      launch_functional(s, functional, ctx->prm, N, kTile, bf->d_rho.p, gga ? bf->d_grad.p : nullptr, nullptr,
                        nullptr, bf->grid->d_w.p, nullptr, nullptr, nullptr, bf->d_vs.p, bf->d_gw.p, bf->d_eblk.p);
    }
  } else {
    const double* Ps[2] = {bf->d_P.p, bf->d_Pb.p};
    double* rho[2] = {bf->d_rho.p, bf->d_rho_b.p};
    double* grd[2] = {bf->d_grad.p, bf->d_grad_b.p};
    for (int sp = 0; sp < 2; ++sp) {
      run_density<double>(bf, bf->d_t64.p, Ps[sp], gga ? 1 : 0, ctx->lanes);
      launch_reduce_partials(s, bf->d_blocks.p, bf->bs, N, bf->d_part.p, gga ? 1 : 0, rho[sp], grd[sp]);
    }
    launch_functional_pol(s, functional, ctx->prm, N, kTile, bf->d_rho.p, bf->d_grad.p, bf->d_rho_b.p,
                          bf->d_grad_b.p, bf->grid->d_w.p, bf->d_vs.p, bf->d_gw.p, bf->d_vs_b.p, bf->d_gw_b.p,
                          bf->d_eblk.p);
  }
  launch_fill(s, bf->d_gacc.p, (int64_t)3 * n, 0.0);
  // restricted: -2 P_mu,nu with the total-density potential; unrestricted: the same per spin
  launch_gradient(s, bf->d_grad_work.p, (int)bf->grad_work.size(), bf->d_blocks.p, bf->d_ao_idx.p, bf->d_t64.p,
                  bf->d_P.p, n, bf->d_vs.p, bf->d_gw.p, gga ? 1 : 0, N, bf->d_gacc.p);
  if (Pb)
    launch_gradient(s, bf->d_grad_work.p, (int)bf->grad_work.size(), bf->d_blocks.p, bf->d_ao_idx.p, bf->d_t64.p,
                    bf->d_Pb.p, n, bf->d_vs_b.p, bf->d_gw_b.p, gga ? 1 : 0, N, bf->d_gacc.p);
  launch_gradient_atoms(s, bf->d_gacc.p, bf->d_ao_atom.p, n, natoms, bf->d_gout.p);
  launch_sum(s, bf->d_eblk.p, nxc, ctx->d_energy);
  HIPX(hipGetLastError());
  HIPX(hipMemcpyAsync(grad, bf->d_gout.p, sizeof(double) * 3 * natoms, hipMemcpyDeviceToHost, s));
  HIPX(hipMemcpyAsync(&ctx->last_energy, ctx->d_energy, sizeof(double), hipMemcpyDeviceToHost, s));
  HIPX(hipStreamSynchronize(s));
  if (energy) *energy = ctx->last_energy;
}
}  // namespace

int xc_geom_gradient(xc_bfogc* bf, int functional, const double* P, int natoms, const int* ao_atom, double* grad,
                     double* energy) {
  return guarded([&] { geom_gradient(bf, functional, P, nullptr, natoms, ao_atom, grad, energy); });
}

int xc_geom_gradient_unrestricted(xc_bfogc* bf, int functional, const double* Pa, const double* Pb, int natoms,
                                  const int* ao_atom, double* grad, double* energy) {
  return guarded([&] {
    require(Pb, XC_ERR_ARG, "xc_geom_gradient_unrestricted: null beta density");
    geom_gradient(bf, functional, Pa, Pb, natoms, ao_atom, grad, energy);
  });
}

int xc_fock_build_unrestricted(xc_bfogc* bf, int functional, int precision, double thr, const double* Pa,
                               const double* Pb, double* Fa, double* Fb, double* energy) {
  return guarded([&] {
    require(bf && Pa && Pb && Fa && Fb, XC_ERR_ARG, "xc_fock_build_unrestricted: null argument");
    xc_ctx* ctx = bf->ctx;
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    const int n = bf->basis->nao;
    ensure_pol_buffers(bf, precision == XC_PREC_FP32);
    HIPX(hipMemcpyAsync(bf->d_P.p, Pa, sizeof(double) * n * n, hipMemcpyHostToDevice, ctx->stream));
    HIPX(hipMemcpyAsync(bf->d_Pb.p, Pb, sizeof(double) * n * n, hipMemcpyHostToDevice, ctx->stream));
    enqueue_build_pol(bf, functional, precision, thr, bf->d_P.p, bf->d_Pb.p, bf->d_F.p, bf->d_Fb.p, ctx->lanes,
                      ctx->d_energy);
    HIPX(hipMemcpyAsync(Fa, bf->d_F.p, sizeof(double) * n * n, hipMemcpyDeviceToHost, ctx->stream));
    HIPX(hipMemcpyAsync(Fb, bf->d_Fb.p, sizeof(double) * n * n, hipMemcpyDeviceToHost, ctx->stream));
    HIPX(hipMemcpyAsync(&ctx->last_energy, ctx->d_energy, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HIPX(hipStreamSynchronize(ctx->stream));
    if (energy) *energy = ctx->last_energy;
  });
}

int xc_fock_build_unrestricted_device(xc_bfogc* bf, int functional, int precision, double thr, const double* dPa,
                                      const double* dPb, double* dFa, double* dFb, double* energy) {
  return guarded([&] {
    require(bf && dPa && dPb && dFa && dFb, XC_ERR_ARG, "xc_fock_build_unrestricted_device: null argument");
    xc_ctx* ctx = bf->ctx;
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    enqueue_build_pol(bf, functional, precision, thr, dPa, dPb, dFa, dFb, ctx->lanes, ctx->d_energy);
    if (energy) {
      HIPX(hipMemcpyAsync(&ctx->last_energy, ctx->d_energy, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
      HIPX(hipStreamSynchronize(ctx->stream));
      *energy = ctx->last_energy;
    }
  });
}

int xc_ri_metric(int naux_shell, const int* aux_l, const int* aux_pure, const int* aux_nprim,
                 const double* aux_center, const double* aux_alpha, const double* aux_coef, double* M) {
  return guarded([&] {
    require(M, XC_ERR_ARG, "xc_ri_metric: null output");
    auto ax = host_shells("xc_ri_metric", naux_shell, aux_l, aux_pure, aux_nprim, aux_center, aux_alpha, aux_coef, 4);
    std::vector<double> m;
    ri_metric(ax, m);
    std::copy(m.begin(), m.end(), M);
  });
}

int xc_coulomb_create(xc_basis* basis, xc_basis* aux, double threshold, xc_coulomb** out) {
  return guarded([&] {
    require(basis && aux && out && threshold >= 0.0, XC_ERR_ARG, "xc_coulomb_create: bad arguments");
    require(basis->ctx == aux->ctx, XC_ERR_ARG, "xc_coulomb_create: bases live on different contexts");
    require(basis->maxl <= kJMaxLPrim, XC_ERR_INPUT,
            "xc_coulomb_create: orbital shells above l = " + std::to_string(kJMaxLPrim));
    require(aux->maxl <= kJMaxLAux, XC_ERR_INPUT,
            "xc_coulomb_create: auxiliary shells above l = " + std::to_string(kJMaxLAux));
    xc_ctx* ctx = basis->ctx;
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    HIPX(hipSetDevice(ctx->device));
    auto* jc = new xc_coulomb();
    try {
      jc->ctx = ctx;
      jc->basis = basis;
      jc->aux = aux;
      // significant primitive pairs, grouped by angular momentum
      const auto& sh = basis->shells;
      const int ns = (int)sh.size();
      for (int A = 0; A < ns; ++A)
        for (int B = 0; B <= A; ++B) {
          const DevShell& a = sh[A];
          const DevShell& b = sh[B];
          const double dx = a.cx - b.cx, dy = a.cy - b.cy, dz = a.cz - b.cz, ab2 = dx * dx + dy * dy + dz * dz;
          for (int i = 0; i < a.nprim; ++i)
            for (int j = 0; j < b.nprim; ++j) {
              const double ea = basis->alpha[a.prim_off + i], eb = basis->alpha[b.prim_off + j];
              const double p = ea + eb;
              const double K = basis->coef[a.prim_off + i] * basis->coef[b.prim_off + j] * std::exp(-ea * eb / p * ab2);
              if (std::fabs(K) * std::pow(M_PI / p, 1.5) < threshold) continue;
              JPair q{};
              q.A = A;
              q.B = B;
              q.L = a.l + b.l;
              q.sym = A == B ? 1.0 : 2.0;
              q.p = p;
              q.K = K;
              const double ac[3] = {a.cx, a.cy, a.cz}, bc[3] = {b.cx, b.cy, b.cz};
              for (int d = 0; d < 3; ++d) {
                q.P[d] = (ea * ac[d] + eb * bc[d]) / p;
                q.PA[d] = q.P[d] - ac[d];
                q.PB[d] = q.P[d] - bc[d];
              }
              jc->pairs.push_back(q);
            }
        }
      std::stable_sort(jc->pairs.begin(), jc->pairs.end(), [&](const JPair& x, const JPair& y) {
        return x.L != y.L ? x.L < y.L : sh[x.A].l < sh[y.A].l;
      });
      int64_t off = 0;
      for (auto& q : jc->pairs) {
        q.herm_off = (int)off;
        off += (q.L + 1) * (q.L + 2) * (q.L + 3) / 6;
      }
      require(off < (1LL << 31), XC_ERR_INPUT, "xc_coulomb_create: too many primitive pairs");
      jc->herm_total = off;
      for (const auto& s : aux->shells) {
        jc->cart_off.push_back(jc->ncart_aux);
        jc->ncart_aux += s.ncart;
      }
      const int n = basis->nao;
      jc->d_pairs.upload(ctx, jc->pairs.data(), jc->pairs.size());
      jc->d_cart_off.upload(ctx, jc->cart_off.data(), jc->cart_off.size());
      jc->d_D.alloc(ctx, std::max<int64_t>(off, 1));
      jc->d_V.alloc(ctx, std::max<int64_t>(off, 1));
      jc->d_bc.alloc(ctx, std::max(jc->ncart_aux, 1));
      jc->d_cc.alloc(ctx, std::max(jc->ncart_aux, 1));
      jc->d_P.alloc(ctx, (size_t)n * n);
      jc->d_J.alloc(ctx, (size_t)n * n);
      HIPX(hipStreamSynchronize(ctx->stream));
    } catch (...) {
      delete jc;
      throw;
    }
    ++ctx->live_handles;
    *out = jc;
  });
}

int xc_coulomb_info(const xc_coulomb* jc, int64_t* npairs, int64_t* hermite_terms) {
  return guarded([&] {
    require(jc, XC_ERR_ARG, "xc_coulomb_info: null handle");
    if (npairs) *npairs = (int64_t)jc->pairs.size();
    if (hermite_terms) *hermite_terms = jc->herm_total;
  });
}

int xc_coulomb_destroy(xc_coulomb* jc) {
  return guarded([&] {
    if (!jc) return;
    (void)hipStreamSynchronize(jc->ctx->stream);
    --jc->ctx->live_handles;
    delete jc;
  });
}

int xc_coulomb_rhs(xc_coulomb* jc, const double* P, double* b) {
  return guarded([&] {
    require(jc && P && b, XC_ERR_ARG, "xc_coulomb_rhs: null argument");
    xc_ctx* ctx = jc->ctx;
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    hipStream_t s = ctx->stream;
    const int n = jc->basis->nao;
    const int npair = (int)jc->pairs.size();
    HIPX(hipMemcpyAsync(jc->d_P.p, P, sizeof(double) * n * n, hipMemcpyHostToDevice, s));
    launch_jdens(s, jc->d_pairs.p, npair, jc->basis->d_shells.p, jc->basis->d_ang.p, jc->d_P.p, n, jc->d_D.p);
    launch_fill(s, jc->d_bc.p, jc->ncart_aux, 0.0);
    launch_jrhs(s, jc->d_pairs.p, npair, jc->d_D.p, jc->aux->d_shells.p, jc->aux->nshell, jc->aux->d_alpha.p,
                jc->aux->d_coef.p, jc->d_cart_off.p, jc->d_bc.p);
    HIPX(hipGetLastError());
    std::vector<double> bc(jc->ncart_aux);
    HIPX(hipMemcpyAsync(bc.data(), jc->d_bc.p, sizeof(double) * bc.size(), hipMemcpyDeviceToHost, s));
    HIPX(hipStreamSynchronize(s));
    // Cartesian -> auxiliary AO: b_f = sum_c M[f][c] b_c
    for (int k = 0; k < jc->aux->nshell; ++k) {
      const DevShell& S = jc->aux->shells[k];
      for (int f = 0; f < S.nfunc; ++f) {
        double v = 0.0;
        for (int c = 0; c < S.ncart; ++c) v += jc->aux->ang[S.ang_off + f * S.ncart + c] * bc[jc->cart_off[k] + c];
        b[S.ao_off + f] = v;
      }
    }
  });
}

int xc_coulomb_matrix(xc_coulomb* jc, const double* c, double* J) {
  return guarded([&] {
    require(jc && c && J, XC_ERR_ARG, "xc_coulomb_matrix: null argument");
    xc_ctx* ctx = jc->ctx;
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    hipStream_t s = ctx->stream;
    const int n = jc->basis->nao;
    const int npair = (int)jc->pairs.size();
    // auxiliary AO -> Cartesian: cc_c = sum_f c_f M[f][c]
    std::vector<double> cc(std::max(jc->ncart_aux, 1), 0.0);
    for (int k = 0; k < jc->aux->nshell; ++k) {
      const DevShell& S = jc->aux->shells[k];
      for (int f = 0; f < S.nfunc; ++f)
        for (int q = 0; q < S.ncart; ++q)
          cc[jc->cart_off[k] + q] += c[S.ao_off + f] * jc->aux->ang[S.ang_off + f * S.ncart + q];
    }
    HIPX(hipMemcpyAsync(jc->d_cc.p, cc.data(), sizeof(double) * jc->ncart_aux, hipMemcpyHostToDevice, s));
    launch_jpot(s, jc->d_pairs.p, npair, jc->aux->d_shells.p, jc->aux->nshell, jc->aux->d_alpha.p, jc->aux->d_coef.p,
                jc->d_cart_off.p, jc->d_cc.p, jc->d_V.p);
    launch_fill(s, jc->d_J.p, (int64_t)n * n, 0.0);
    launch_jassemble(s, jc->d_pairs.p, npair, jc->basis->d_shells.p, jc->basis->d_ang.p, jc->d_V.p, n, jc->d_J.p);
    HIPX(hipGetLastError());
    HIPX(hipMemcpyAsync(J, jc->d_J.p, sizeof(double) * n * n, hipMemcpyDeviceToHost, s));
    HIPX(hipStreamSynchronize(s));
  });
}

int xc_last_energy(xc_ctx* ctx, double* energy) {
  return guarded([&] {
    require(ctx && energy, XC_ERR_ARG, "null argument");
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    HIPX(hipMemcpyAsync(&ctx->last_energy, ctx->d_energy, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HIPX(hipStreamSynchronize(ctx->stream));
    *energy = ctx->last_energy;
  });
}

int xc_fock_build(xc_bfogc* bf, int functional, int precision, double thr, const double* P, double* F,
                  double* energy) {
  return guarded([&] {
    require(bf && P && F, XC_ERR_ARG, "xc_fock_build: null argument");
    xc_ctx* ctx = bf->ctx;
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    const int n = bf->basis->nao;
    HIPX(hipMemcpyAsync(bf->d_P.p, P, sizeof(double) * n * n, hipMemcpyHostToDevice, ctx->stream));
    enqueue_build(bf, functional, precision, thr, bf->d_P.p, bf->d_F.p, ctx->lanes, ctx->d_energy);
    HIPX(hipMemcpyAsync(F, bf->d_F.p, sizeof(double) * n * n, hipMemcpyDeviceToHost, ctx->stream));
    HIPX(hipMemcpyAsync(&ctx->last_energy, ctx->d_energy, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HIPX(hipStreamSynchronize(ctx->stream));
    if (energy) *energy = ctx->last_energy;
  });
}

int xc_ctx_set_deterministic(xc_ctx* ctx, int enable) {
  return guarded([&] {
    require(ctx, XC_ERR_ARG, "null context");
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    require(ctx->live_handles == 0, XC_ERR_STATE,
            "xc_ctx_set_deterministic: set before any basis, grid or controller of the context exists");
    ctx->deterministic = enable != 0;
  });
}

int xc_profile(xc_ctx* ctx, int enable) {
  return guarded([&] {
    require(ctx, XC_ERR_ARG, "null context");
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    ctx->profile = enable != 0;
  });
}

int xc_profile_collect(xc_ctx* ctx, double* stage_ms, int* nbuilds) {
  return guarded([&] {
    require(ctx && stage_ms, XC_ERR_ARG, "null argument");
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    HIPX(hipStreamSynchronize(ctx->stream));
    const size_t per = XC_NSTAGES + 1;
    for (int k = 0; k < XC_NSTAGES; ++k) stage_ms[k] = 0.0;
    const size_t nb = ctx->ev_used / per;
    for (size_t b = 0; b < nb; ++b)
      for (int k = 0; k < XC_NSTAGES; ++k) {
        float ms = 0.0f;
        HIPX(hipEventElapsedTime(&ms, ctx->ev_pool[b * per + k], ctx->ev_pool[b * per + k + 1]));
        stage_ms[k] += ms;
      }
    if (nbuilds) *nbuilds = (int)nb;
    ctx->ev_used = 0;
  });
}

int xc_device_alloc(xc_ctx* ctx, size_t bytes, void** dptr) {
  return guarded([&] {
    require(ctx && dptr, XC_ERR_ARG, "null argument");
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    *dptr = ctx->alloc(bytes);
    if (*dptr) ctx->user_allocs[*dptr] = bytes;
  });
}

int xc_device_free(xc_ctx* ctx, void* dptr) {
  return guarded([&] {
    require(ctx, XC_ERR_ARG, "null context");
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    HIPX(hipStreamSynchronize(ctx->stream));
    if (dptr) {
      auto it = ctx->user_allocs.find(dptr);
      require(it != ctx->user_allocs.end(), XC_ERR_ARG, "xc_device_free: pointer not from xc_device_alloc");
      ctx->release(dptr, it->second);
      ctx->user_allocs.erase(it);
    }
  });
}

int xc_memcpy_h2d(xc_ctx* ctx, void* dst, const void* src, size_t bytes) {
  return guarded([&] {
    require(ctx && dst && src, XC_ERR_ARG, "null argument");
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    HIPX(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
    HIPX(hipStreamSynchronize(ctx->stream));
  });
}

int xc_device_copy(xc_ctx* ctx, void* dst, const void* src, size_t bytes) {
  return guarded([&] {
    require(ctx && dst && src, XC_ERR_ARG, "null argument");
    require(bytes % 32 == 0, XC_ERR_ARG, "xc_device_copy: size must be a multiple of 32 bytes");
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    HIPX(hipSetDevice(ctx->device));
    launch_copy(ctx->stream, src, dst, (int64_t)(bytes / 32));
    HIPX(hipGetLastError());
    HIPX(hipStreamSynchronize(ctx->stream));
  });
}

// ---- serving ----------------------------------------------------------------
struct xc_server {
  struct Slot {
    hipStream_t stream = nullptr;
    Lanes lanes;
    hipEvent_t done = nullptr;
    double* h_P = nullptr;  // pinned staging of P, F and E
    double* h_F = nullptr;
    double* h_E = nullptr;
    size_t cap = 0;
    bool busy = false;
    int64_t ticket = -1;
    xc_bfogc* bf = nullptr;
    double* user_F = nullptr;
    double* user_E = nullptr;
  };
  xc_ctx* ctx = nullptr;
  std::vector<Slot> slots;
  bool graphs = true;
  int64_t next_ticket = 1;
  int64_t submitted = 0, graph_launches = 0;
};

namespace {

void server_slot_release(xc_server::Slot& sl) {
  if (sl.h_P) (void)hipHostFree(sl.h_P);
  if (sl.h_F) (void)hipHostFree(sl.h_F);
  sl.h_P = sl.h_F = nullptr;
  sl.cap = 0;
}

void server_finish(xc_server::Slot& sl, int n) {
  std::memcpy(sl.user_F, sl.h_F, sizeof(double) * (size_t)n * n);
  if (sl.user_E) *sl.user_E = *sl.h_E;
  sl.bf->in_flight = false;
  sl.busy = false;
  sl.bf = nullptr;
  sl.ticket = -1;
}

xc_server::Slot* server_slot_of(xc_server* srv, int64_t ticket) {
  for (auto& sl : srv->slots)
    if (sl.busy && sl.ticket == ticket) return &sl;
  return nullptr;
}

}  // namespace

int xc_server_create(xc_ctx* ctx, int nslots, int use_graphs, xc_server** out) {
  return guarded([&] {
    require(ctx && out, XC_ERR_ARG, "xc_server_create: null argument");
    require(nslots >= 1 && nslots <= 16, XC_ERR_ARG, "xc_server_create: slots must be in [1, 16]");
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    HIPX(hipSetDevice(ctx->device));
    auto* srv = new xc_server();
    srv->ctx = ctx;
    srv->graphs = use_graphs != 0;
    srv->slots.resize(nslots);
    try {
      for (auto& sl : srv->slots) {
        HIPX(hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking));
        // concurrency comes from the slots: a slot's build stays on its own stream
        sl.lanes.create(sl.stream, false);
        HIPX(hipEventCreateWithFlags(&sl.done, hipEventDisableTiming));
        HIPX(hipHostMalloc((void**)&sl.h_E, sizeof(double), hipHostMallocDefault));
      }
    } catch (...) {
      for (auto& sl : srv->slots) {
        sl.lanes.destroy();
        if (sl.stream) (void)hipStreamDestroy(sl.stream);
        if (sl.done) (void)hipEventDestroy(sl.done);
        if (sl.h_E) (void)hipHostFree(sl.h_E);
      }
      delete srv;
      throw;
    }
    ++ctx->live_handles;
    *out = srv;
  });
}

int xc_server_destroy(xc_server* srv) {
  return guarded([&] {
    if (!srv) return;
    std::lock_guard<std::recursive_mutex> lk(srv->ctx->mu);
    for (auto& sl : srv->slots) {
      (void)hipStreamSynchronize(sl.stream);
      if (sl.busy) {
        sl.bf->in_flight = false;
        sl.busy = false;
      }
      server_slot_release(sl);
      (void)hipHostFree(sl.h_E);
      (void)hipEventDestroy(sl.done);
      sl.lanes.destroy();
      (void)hipStreamDestroy(sl.stream);
    }
    --srv->ctx->live_handles;
    delete srv;
  });
}

int xc_server_submit(xc_server* srv, xc_bfogc* bf, int functional, int precision, double thr, const double* P,
                     double* F, double* energy, int64_t* ticket) {
  return guarded([&] {
    require(srv && bf && P && F && ticket, XC_ERR_ARG, "xc_server_submit: null argument");
    require(bf->ctx == srv->ctx, XC_ERR_ARG, "xc_server_submit: controller of another context");
    xc_ctx* ctx = srv->ctx;
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    *ticket = 0;
    if (bf->in_flight) throw XcError(XC_BUSY, "xc_server_submit: a build of this controller is in flight");
    xc_server::Slot* sl = nullptr;
    for (auto& c : srv->slots)
      if (!c.busy) {
        sl = &c;
        break;
      }
    if (!sl) throw XcError(XC_BUSY, "xc_server_submit: every slot is busy");
    const int n = bf->basis->nao;
    const size_t nn = (size_t)n * n;
    if (sl->cap < nn) {
      server_slot_release(*sl);
      HIPX(hipHostMalloc((void**)&sl->h_P, sizeof(double) * nn, hipHostMallocDefault));
      HIPX(hipHostMalloc((void**)&sl->h_F, sizeof(double) * nn, hipHostMallocDefault));
      sl->cap = nn;
    }
    std::memcpy(sl->h_P, P, sizeof(double) * nn);
    hipStream_t s = sl->stream;
    HIPX(hipMemcpyAsync(bf->d_P.p, sl->h_P, sizeof(double) * nn, hipMemcpyHostToDevice, s));
    if (srv->graphs) {
      const auto key = std::make_pair(functional + 16 * precision, thr);
      auto it = bf->graphs.find(key);
      if (it == bf->graphs.end()) {
        // capture one build on this controller's resident buffers (no memset nodes: U is
        // cleared by a fill kernel)
        hipGraph_t graph = nullptr;
        HIPX(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        try {
          enqueue_build(bf, functional, precision, thr, bf->d_P.p, bf->d_F.p, sl->lanes, bf->d_E.p);
        } catch (...) {
          (void)hipStreamEndCapture(s, &graph);
          if (graph) (void)hipGraphDestroy(graph);
          throw;
        }
        HIPX(hipStreamEndCapture(s, &graph));
        hipGraphExec_t exec = nullptr;
        const hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        HIPX(e);
        it = bf->graphs.emplace(key, exec).first;
      }
      HIPX(hipGraphLaunch(it->second, s));
      ++srv->graph_launches;
    } else {
      enqueue_build(bf, functional, precision, thr, bf->d_P.p, bf->d_F.p, sl->lanes, bf->d_E.p);
    }
    HIPX(hipMemcpyAsync(sl->h_F, bf->d_F.p, sizeof(double) * nn, hipMemcpyDeviceToHost, s));
    HIPX(hipMemcpyAsync(sl->h_E, bf->d_E.p, sizeof(double), hipMemcpyDeviceToHost, s));
    HIPX(hipEventRecord(sl->done, s));
    sl->busy = true;
    sl->bf = bf;
    sl->user_F = F;
    sl->user_E = energy;
    sl->ticket = srv->next_ticket++;
    bf->in_flight = true;
    ++srv->submitted;
    *ticket = sl->ticket;
  });
}

int xc_server_poll(xc_server* srv, int64_t ticket, int* done) {
  return guarded([&] {
    require(srv && done, XC_ERR_ARG, "xc_server_poll: null argument");
    std::lock_guard<std::recursive_mutex> lk(srv->ctx->mu);
    xc_server::Slot* sl = server_slot_of(srv, ticket);
    require(sl != nullptr, XC_ERR_ARG, "xc_server_poll: unknown or finished ticket");
    const hipError_t q = hipEventQuery(sl->done);
    if (q == hipErrorNotReady) {
      *done = 0;
      return;
    }
    HIPX(q);
    server_finish(*sl, sl->bf->basis->nao);
    *done = 1;
  });
}

int xc_server_wait(xc_server* srv, int64_t ticket) {
  return guarded([&] {
    require(srv, XC_ERR_ARG, "xc_server_wait: null argument");
    std::lock_guard<std::recursive_mutex> lk(srv->ctx->mu);
    xc_server::Slot* sl = server_slot_of(srv, ticket);
    require(sl != nullptr, XC_ERR_ARG, "xc_server_wait: unknown or finished ticket");
    HIPX(hipEventSynchronize(sl->done));
    server_finish(*sl, sl->bf->basis->nao);
  });
}

int xc_server_stats(const xc_server* srv, int* busy_slots, int64_t* submitted, int64_t* graph_launches) {
  return guarded([&] {
    require(srv, XC_ERR_ARG, "xc_server_stats: null argument");
    std::lock_guard<std::recursive_mutex> lk(srv->ctx->mu);
    if (busy_slots) {
      int b = 0;
      for (const auto& sl : srv->slots) b += sl.busy ? 1 : 0;
      *busy_slots = b;
    }
    if (submitted) *submitted = srv->submitted;
    if (graph_launches) *graph_launches = srv->graph_launches;
  });
}

int xc_memcpy_d2h(xc_ctx* ctx, void* dst, const void* src, size_t bytes) {
  return guarded([&] {
    require(ctx && dst && src, XC_ERR_ARG, "null argument");
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    HIPX(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
    HIPX(hipStreamSynchronize(ctx->stream));
  });
}

}  // extern "C"
