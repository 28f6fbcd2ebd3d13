// Host-side objects of libsvxg: the device memory arena, problems, batches.
#pragma once

#include <array>
#include <atomic>
#include <cstring>
#include <memory>
#include <mutex>
#include <unordered_map>
#include <string>
#include <vector>

#include "svxg_kernels.hpp"

namespace svxg {

struct Batch;
void set_error(const std::string& msg);

#define SVXG_HIP(call)                                                                            \
  do {                                                                                           \
    hipError_t e_ = (call);                                                                      \
    if (e_ != hipSuccess) {                                                                      \
      ::svxg::set_error(std::string(#call) + ": " + hipGetErrorString(e_));                      \
      return SVXG_DEVICE_ERROR;                                                                  \
    }                                                                                            \
  } while (0)

#define SVXG_TRY(expr)          \
  do {                          \
    int st_ = (expr);           \
    if (st_ != SVXG_OK) return st_; \
  } while (0)

// Device memory for one context: blocks of size classes (four a power of two, 256 B .. 64 MB)
// carved from 512 MB slabs and recycled through per-class free lists; larger requests get their
// own allocation. A slab whose blocks are all free again is reset whole (its free-list entries go
// stale by generation) and carved afresh, so memory freed by a large batch returns as contiguous
// slabs rather than as blocks scattered over all of them. Every byte is admitted against the
// context's HBM budget.
class Arena {
 public:
  Arena(int device, int64_t budget) : device_(device), budget_(budget) {}
  ~Arena();
  int alloc(size_t bytes, void** out, size_t* held);
  void release(void* p, size_t held);
  int64_t in_use() const { return in_use_.load(std::memory_order_relaxed); }
  int64_t budget() const { return budget_; }

 private:
  static constexpr size_t kSlab = size_t(512) << 20;
  static constexpr size_t kBig = size_t(64) << 20;
  struct Slab {
    char* base;
    int64_t live;   // blocks handed out and not yet released
    int64_t freed;  // valid entries of this slab in the free lists
    uint32_t gen;   // bumped when the slab is reset; older free-list entries are stale
  };
  struct FreeBlock {
    char* p;
    uint32_t slab, gen;
  };
  uint32_t slab_of(const void* p) const;
  void reset_slab(uint32_t s);
  void purge_stale();
  int device_;
  int64_t budget_;
  std::atomic<int64_t> in_use_{0};
  std::mutex mu_;
  std::vector<Slab> slabs_;
  std::vector<std::pair<uintptr_t, uint32_t>> by_addr_;  // slab bases, sorted
  std::vector<uint32_t> empty_;                           // reset slabs not being carved
  std::vector<void*> big_;
  uint32_t cur_ = UINT32_MAX;  // the slab being carved
  size_t bump_off_ = 0;
  int64_t stale_ = 0;
  std::array<std::vector<FreeBlock>, 80> free_;  // (class 72: 64 MB)
};

// a device allocation owned by its holder (returned to the arena on destruction)
class DevBuf {
 public:
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept { *this = std::move(o); }
  DevBuf& operator=(DevBuf&& o) noexcept {
    if (this != &o) {
      reset();
      a_ = o.a_;
      p_ = o.p_;
      held_ = o.held_;
      bytes_ = o.bytes_;
      o.p_ = nullptr;
      o.held_ = o.bytes_ = 0;
    }
    return *this;
  }
  ~DevBuf() { reset(); }
  int make(Arena* a, size_t bytes) {
    reset();
    a_ = a;
    bytes_ = bytes;
    if (bytes == 0) return SVXG_OK;
    return a->alloc(bytes, &p_, &held_);
  }
  void reset() {
    if (p_ && a_) a_->release(p_, held_);
    p_ = nullptr;
    held_ = bytes_ = 0;
  }
  template <typename T>
  T* as() const {
    return static_cast<T*>(p_);
  }
  size_t bytes() const { return bytes_; }
  size_t held() const { return held_; }
  explicit operator bool() const { return p_ != nullptr; }

 private:
  Arena* a_ = nullptr;
  void* p_ = nullptr;
  size_t held_ = 0, bytes_ = 0;
};

// one level of a problem's hierarchy, with its device arrays
struct HostLevel {
  int32_t n = 0, nc = 0;
  int64_t nnzb = 0, np = 0;
  bool terminal = false;
  int32_t dense_m = 0;
  double lmax = 1.0;
  DevBuf rowptr, col, A64, A32, A16, D64, D32;
  DevBuf Prow, Pcol, P64, P32, Rrow, Rfine, Rent, Minv;
  DevBuf D16;       // FP16 copy of the inverse diagonal blocks
  DevBuf P16, R16;  // FP16 copies of P by prolongation entry and of its transpose by restriction entry
  DevBuf x64[2], b64, r64, d64, x32[2], b32, r32, d32;
  DevBuf pos, lam;
  std::vector<int32_t> h_rowptr;  // (level 0: K's pattern on the host too)
  std::vector<int32_t> h_col;
};

}  // namespace svxg

struct svxg_ctx {
  int device = 0;
  int resident_rows = 0;  // graphs up to this many free rows run their steps in one workgroup (0: none)
  int step_graphs = 1;  // tiled solve steps recorded per batch (hipGraph): 0 never, 1 large batches, 2 always
  hipStream_t stream = nullptr;
  std::unique_ptr<svxg::Arena> arena;
  std::mutex mu;  // batches and setup on the context stream, one at a time
  // kernel timing: event pairs around the launches of the profiled kernels (kinds 0..3: the outer
  // K product FP64 / FP32, a level-0 smoothing step FP64 / FP32)
  // staging of batched uploads and readbacks, one lane per stream that uses it (grown on demand)
  struct Lane {
    hipStream_t s = nullptr;
    void* pinned = nullptr;  // host, pinned
    size_t pinned_bytes = 0;
    svxg::DevBuf dstage;     // device
    // regions [0, head) of both buffers are used by copies queued on s and not known to be done;
    // `done` is recorded after the last of them (pending: recorded and not yet seen complete)
    size_t head = 0;
    hipEvent_t done = nullptr;
    bool pending = false;
  };
  std::mutex lanes_mu;
  std::vector<std::unique_ptr<Lane>> lanes;
  bool profiling = false;
  struct Prof {
    std::vector<std::pair<hipEvent_t, hipEvent_t>> pending, spare;
    int64_t launches = 0;
    double ms = 0.0;
  } prof[4];
  svxg::DevBuf prof_acc;  // 4 doubles on the device: the bytes of the live graphs of each kind's launches
  // the device address of this device's copy of the library's fixed scratch array (the
  // allocation-free entry point's), taken at creation with the device current
  uint8_t* scratch = nullptr;
  // the autotuner: samples by level-shape key while a shape is tuned, the choices once it is
  int autotune = 1;
  std::mutex tune_mu;
  struct TuneAcc {
    double ms[4] = {0, 0, 0, 0};
    int n[4] = {0, 0, 0, 0};
  };
  std::unordered_map<uint64_t, TuneAcc> tune_acc;
  std::unordered_map<uint64_t, int> tune_choice;
  std::vector<svxg_tune_record> tune_log;
};

struct svxg_problem {
  svxg_ctx* ctx = nullptr;
  hipStream_t stream = nullptr;
  std::vector<svxg_node> nodes;
  std::vector<svxg_bond> bonds;
  std::vector<svxg_material> mats;
  svxg_stress_options opt{};
  bool assembled = false, pc_current = false, jacobi_only = false, have_opt = false;
  bool fp64_levels = false;  // the hierarchy keeps its FP64 arrays (set up for the FP64 path)
  int64_t assembly_work = 0, work_per_apply = 0;
  int32_t nfree = 0;
  std::vector<int32_t> dof;
  std::vector<int32_t> bond_slot;  // 4 per bond
  // device graph data
  svxg::DevBuf d_nodes, d_nodec, d_bonds, d_mats, d_dof, d_bond_slot, d_cptr, d_cidx, d_row_gone, d_id_scale;
  // levels and solve vectors
  std::vector<svxg::HostLevel> lv;
  svxg::DevBuf x, r, z, p, q, b, ix, ir, iz, ip, iq;
  svxg::DevBuf u_all;  // 6 per node: the last solve's displacement (device-resident)
  std::vector<svxg_break> last_picks;  // the picks of its last judgement by a server
  svxg::DevBuf d_base, d_has_base;  // the judge's baseline loads (one svxg_bond_load a bond) and flags
  bool have_solution = false;
  bool have_load = false;  // b holds a load (a solve or a request gave one)
  // a solve spread over calls (begin / iterate)
  bool run_active = false;
  int32_t run_iters = 0;
  // the batch-of-one used by single-problem calls (built on first use, dropped on change)
  svxg::Batch* self_batch = nullptr;
  double last_solve_ms = 0.0, last_setup_ms = 0.0;
  int64_t device_bytes() const;
};
