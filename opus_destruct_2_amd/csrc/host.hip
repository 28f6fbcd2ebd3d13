// libsvxg host side: the device arena, the assembly plan, the batched multigrid setup, the PCG and
// mixed-precision refinement drivers and the C ABI (include/svxg.h).
//
// The reference's StressProblem (core/src/stress/stress.cpp) runs every step on the CPU; here the
// host keeps only integer bookkeeping (the DOF map, K's block pattern and each block's list of bond
// contributions - the reference's Bsr6Builder rows, amg.cpp:185-224) and drives the device: every
// floating-point value of K, of the hierarchy and of the solve is computed by the kernels.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <numeric>
#include <thread>
#include <unordered_map>
#include <unordered_set>

#include "host_common.hpp"

// The allocation-free bond checks' scratch (svxg_problem_bond_checks_noalloc): the job descriptor
// and the bond tile table of k_bond_checks. Each GPU that loads the library holds its own copy until
// the process ends; host code reaches it only through the address hipGetSymbolAddress gives for the
// current device (scratch_address), never through this name.
namespace svxg {
constexpr int64_t kNoAllocMaxBonds = int64_t(1) << 24;
constexpr size_t kNoAllocTileBase = 256;  // (the descriptor first, then the tiles)
constexpr size_t kNoAllocBytes = kNoAllocTileBase + size_t((kNoAllocMaxBonds + kWG - 1) / kWG) * sizeof(Tile);
static_assert(sizeof(CheckDev) <= kNoAllocTileBase, "the check descriptor fits the scratch head");
}  // namespace svxg
static __device__ __align__(256) uint8_t scratch[svxg::kNoAllocBytes];

namespace svxg {

// this device's scratch address (device current), once per device
static int scratch_address(int device, uint8_t** out) {
  static std::mutex mu;
  static std::vector<uint8_t*> addr;
  std::lock_guard<std::mutex> lk(mu);
  if (size_t(device) >= addr.size()) addr.resize(size_t(device) + 1, nullptr);
  if (!addr[size_t(device)]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      *out = nullptr;
      return e != hipSuccess ? int(e) : -1;
    }
    addr[size_t(device)] = static_cast<uint8_t*>(p);
  }
  *out = addr[size_t(device)];
  return 0;
}

// the descriptor (from the argument) and the tile table of one problem's bond checks, in the scratch
__global__ void k_noalloc_checks_prep(CheckDev c, uint8_t* __restrict__ base, int ntiles) {
  const int t = int(blockIdx.x * blockDim.x + threadIdx.x);
  if (t == 0) *reinterpret_cast<CheckDev*>(base) = c;
  if (t < ntiles) {
    Tile* tiles = reinterpret_cast<Tile*>(base + kNoAllocTileBase);
    tiles[t] = Tile{0, t * kWG, min(kWG, c.nbonds - t * kWG), t};
  }
}

namespace {
thread_local std::string g_error;
}
void set_error(const std::string& msg) { g_error = msg; }

// ---- arena ---------------------------------------------------------------------------------------

// size classes, four a power of two from 256 B (256, 320, 384, 448, 512, 640, ...): a block is
// under a fifth larger than the request that takes it
static size_t class_size(int c) { return (size_t(4 + (c & 3)) << (c >> 2)) * 64; }

static int size_class(size_t bytes) {
  int c = 0;
  while (class_size(c) < bytes) ++c;
  return c;
}

Arena::~Arena() {
  (void)hipSetDevice(device_);
  for (const Slab& sl : slabs_) (void)hipFree(sl.base);
  for (void* p : big_) (void)hipFree(p);
}

uint32_t Arena::slab_of(const void* p) const {
  const uintptr_t a = reinterpret_cast<uintptr_t>(p);
  auto it = std::upper_bound(by_addr_.begin(), by_addr_.end(), std::make_pair(a, UINT32_MAX));
  return std::prev(it)->second;
}

// every block of slab s is free: its free-list entries go stale and the slab is carved afresh
void Arena::reset_slab(uint32_t s) {
  Slab& sl = slabs_[s];
  stale_ += sl.freed;
  sl.freed = 0;
  ++sl.gen;
  if (s == cur_)
    bump_off_ = 0;
  else
    empty_.push_back(s);
  if (stale_ > 65536) purge_stale();
}

void Arena::purge_stale() {
  for (auto& fl : free_)
    fl.erase(std::remove_if(fl.begin(), fl.end(), [&](const FreeBlock& f) { return f.gen != slabs_[f.slab].gen; }),
             fl.end());
  stale_ = 0;
}

int Arena::alloc(size_t bytes, void** out, size_t* held) {
  std::lock_guard<std::mutex> lk(mu_);
  *out = nullptr;
  if (bytes >= kBig) {
    const size_t sz = (bytes + 4095) & ~size_t(4095);
    if (in_use_.load() + int64_t(sz) > budget_) {
      set_error("HBM budget exceeded");
      return SVXG_OOM_BUDGET;
    }
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, sz);
    if (e != hipSuccess) {
      set_error(std::string("hipMalloc: ") + hipGetErrorString(e));
      return SVXG_DEVICE_ERROR;
    }
    big_.push_back(p);
    in_use_ += int64_t(sz);
    *out = p;
    *held = sz;
    return SVXG_OK;
  }
  const int c0 = size_class(bytes);
  // a free block of this class or of the next two (at most 56 % larger) before new memory
  int c = c0;
  bool found = false;
  for (int k = 0; k < 3 && size_t(c0 + k) < free_.size() && !found; ++k) {
    auto& fl = free_[size_t(c0 + k)];
    while (!fl.empty() && fl.back().gen != slabs_[fl.back().slab].gen) {
      fl.pop_back();
      --stale_;
    }
    if (!fl.empty()) {
      c = c0 + k;
      found = true;
    }
  }
  const size_t sz = class_size(c);
  if (in_use_.load() + int64_t(sz) > budget_) {
    set_error("HBM budget exceeded");
    return SVXG_OOM_BUDGET;
  }
  uint32_t s;
  if (found) {
    const FreeBlock f = free_[size_t(c)].back();
    free_[size_t(c)].pop_back();
    s = f.slab;
    --slabs_[s].freed;
    *out = f.p;
  } else {
    if (cur_ == UINT32_MAX || kSlab - bump_off_ < sz) {
      if (!empty_.empty()) {
        cur_ = empty_.back();
        empty_.pop_back();
      } else {
        void* p = nullptr;
        hipError_t e = hipMalloc(&p, kSlab);
        if (e != hipSuccess) {
          set_error(std::string("hipMalloc: ") + hipGetErrorString(e));
          return SVXG_DEVICE_ERROR;
        }
        cur_ = uint32_t(slabs_.size());
        slabs_.push_back(Slab{static_cast<char*>(p), 0, 0, 0});
        const auto key = std::make_pair(reinterpret_cast<uintptr_t>(p), cur_);
        by_addr_.insert(std::upper_bound(by_addr_.begin(), by_addr_.end(), key), key);
      }
      bump_off_ = 0;
    }
    s = cur_;
    *out = slabs_[s].base + bump_off_;
    bump_off_ += sz;
  }
  ++slabs_[s].live;
  in_use_ += int64_t(sz);
  *held = sz;
  return SVXG_OK;
}

void Arena::release(void* p, size_t held) {
  std::lock_guard<std::mutex> lk(mu_);
  in_use_ -= int64_t(held);
  if (held >= kBig) {
    for (size_t k = 0; k < big_.size(); ++k)
      if (big_[k] == p) {
        (void)hipFree(p);
        big_[k] = big_.back();
        big_.pop_back();
        return;
      }
    return;
  }
  const uint32_t s = slab_of(p);
  Slab& sl = slabs_[s];
  if (--sl.live == 0) {
    reset_slab(s);
    return;
  }
  free_[size_t(size_class(held))].push_back(FreeBlock{static_cast<char*>(p), s, sl.gen});
  ++sl.freed;
}

// ---- tiles -----------------------------------------------------------------------------------------

static void add_tiles(std::vector<Tile>& t, int g, int n, int per) {
  for (int r = 0; r < n; r += per) t.push_back(Tile{g, r, std::min(per, n - r), int32_t(t.size())});
}

// (synchronous: the host vector may go out of scope when this returns)
template <typename T>
static int upload(DevBuf& d, Arena* a, const std::vector<T>& h, hipStream_t s) {
  SVXG_TRY(d.make(a, std::max<size_t>(1, h.size()) * sizeof(T)));
  if (!h.empty()) {
    SVXG_HIP(hipMemcpyAsync(d.as<void>(), h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, s));
    SVXG_HIP(hipStreamSynchronize(s));
  }
  return SVXG_OK;
}
template <typename T>
static int upload_into(const DevBuf& d, const std::vector<T>& h, hipStream_t s) {
  if (!h.empty()) {
    SVXG_HIP(hipMemcpyAsync(d.as<void>(), h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, s));
    SVXG_HIP(hipStreamSynchronize(s));
  }
  return SVXG_OK;
}

// The staging lane of a stream: pinned host and device buffers, used as a ring. A region is taken
// from the lane's head; a copy queued on the stream that reads it is followed by the lane's event,
// and the head returns to 0 once that event is seen complete (or the stream was synchronised by
// its user). A region is therefore never written by the host while a queued copy still reads it,
// and uploads need not wait for their copies.
static size_t round256(size_t b) { return (b + 255) & ~size_t(255); }

static int stage_region(svxg_ctx* c, hipStream_t s, size_t bytes, svxg_ctx::Lane** out, size_t* off) {
  svxg_ctx::Lane* L = nullptr;
  {
    std::lock_guard<std::mutex> lk(c->lanes_mu);
    for (auto& l : c->lanes)
      if (l->s == s) L = l.get();
    if (!L) {
      c->lanes.push_back(std::make_unique<svxg_ctx::Lane>());
      L = c->lanes.back().get();
      L->s = s;
    }
  }
  bytes = round256(std::max<size_t>(1, bytes));
  if (L->pending && hipEventQuery(L->done) == hipSuccess) L->pending = false;
  if (!L->pending) L->head = 0;
  const size_t cap = std::min(L->pinned_bytes, L->dstage.bytes());
  if (L->head + bytes > cap) {
    // (wait for what the queued copies still read, then start from the front, grown if need be)
    if (L->pending) SVXG_HIP(hipEventSynchronize(L->done));
    L->pending = false;
    L->head = 0;
    if (bytes > L->pinned_bytes) {
      if (L->pinned) (void)hipHostFree(L->pinned);
      L->pinned = nullptr;
      L->pinned_bytes = 0;
      const size_t want = std::max(bytes, size_t(16) << 20);
      SVXG_HIP(hipHostMalloc(&L->pinned, want, hipHostMallocDefault));
      L->pinned_bytes = want;
    }
    if (bytes > L->dstage.bytes()) SVXG_TRY(L->dstage.make(c->arena.get(), std::max(bytes, size_t(16) << 20)));
  }
  *off = L->head;
  L->head += bytes;
  *out = L;
  return SVXG_OK;
}
// the copies just queued on the lane's stream read its regions: mark them in use until they are done
static int stage_queued(svxg_ctx::Lane* L) {
  if (!L->done) SVXG_HIP(hipEventCreateWithFlags(&L->done, hipEventDisableTiming));
  SVXG_HIP(hipEventRecord(L->done, L->s));
  L->pending = true;
  return SVXG_OK;
}
// the lane's stream was just synchronised: every region is free
static void stage_idle(svxg_ctx::Lane* L) {
  L->pending = false;
  L->head = 0;
}
// a lane's resources, when its stream goes
static void stage_free(svxg_ctx::Lane* L) {
  if (L->pinned) (void)hipHostFree(L->pinned);
  L->pinned = nullptr;
  if (L->done) (void)hipEventDestroy(L->done);
  L->done = nullptr;
}

// f(k) for k in [0, n) on up to 8 host threads (the memcpy of large batched uploads)
template <typename F>
static void copy_in_parallel(size_t n, F&& f) {
  const size_t hw = std::max(1u, std::thread::hardware_concurrency());
  const size_t T = std::min<size_t>({8, hw, n / 64 + 1});
  if (T <= 1) {
    for (size_t k = 0; k < n; ++k) f(k);
    return;
  }
  std::atomic<size_t> next{0};
  std::vector<std::thread> pool;
  for (size_t t = 0; t < T; ++t)
    pool.emplace_back([&] {
      for (size_t k = next.fetch_add(1); k < n; k = next.fetch_add(1)) f(k);
    });
  for (std::thread& th : pool) th.join();
}

// Parts of at most kStageCap bytes each: large batches pass through the staging buffers in
// several round trips instead of growing them to the size of the whole batch
static constexpr size_t kStageCap = size_t(256) << 20;

struct StagePart {
  void* dst;
  const void* src;
  size_t bytes;
};

// (and the copy kernel takes a workgroup a segment: segments of at most kStagePiece bytes)
static constexpr size_t kStagePiece = size_t(256) << 10;

static std::vector<StagePart> stage_parts(const std::vector<StagePart>& jobs) {
  std::vector<StagePart> out;
  out.reserve(jobs.size());
  for (const StagePart& j : jobs)
    for (size_t o = 0; o < j.bytes; o += kStagePiece)
      out.push_back({static_cast<char*>(j.dst) + o, static_cast<const char*>(j.src) + o,
                     std::min(kStagePiece, j.bytes - o)});
  return out;
}

// the end of the round trip that starts at part k0: parts up to kStageCap bytes in all (one at least)
static size_t stage_round(const std::vector<StagePart>& parts, size_t k0, size_t* data) {
  size_t k1 = k0, d = 0;
  while (k1 < parts.size() && (k1 == k0 || d + round256(parts[k1].bytes) <= kStageCap)) d += round256(parts[k1++].bytes);
  *data = d;
  return k1;
}

// Uploads of a batch (many small arrays): packed into a region of the stream's staging lane with a
// table of their destinations, sent in one copy, spread by one kernel. wait: the stream is then
// synchronised (the caller may free what the stream's queued work uses); else flush() returns with
// the copies queued (the host arrays may go at once, the region stays taken until they are done:
// the caller keeps every device array its queued work uses until its next synchronisation)
class Stager {
 public:
  Stager(svxg_ctx* c, hipStream_t s, bool wait = true) : c_(c), s_(s), wait_(wait) {}
  template <typename T>
  int add(DevBuf& d, Arena* a, const std::vector<T>& h) {
    SVXG_TRY(d.make(a, std::max<size_t>(1, h.size()) * sizeof(T)));
    if (!h.empty()) jobs_.push_back({d.as<void>(), h.data(), h.size() * sizeof(T)});
    return SVXG_OK;
  }
  int add_raw(void* dst, const void* src, size_t bytes) {
    if (bytes) jobs_.push_back({dst, src, bytes});
    return SVXG_OK;
  }
  int flush() {
    if (jobs_.empty()) return SVXG_OK;
    const std::vector<StagePart> parts = stage_parts(jobs_);
    jobs_.clear();
    for (size_t k0 = 0, k1 = 0; k0 < parts.size(); k0 = k1) {
      size_t data = 0;
      k1 = stage_round(parts, k0, &data);
      SVXG_TRY(send(parts, k0, k1, data));
    }
    return SVXG_OK;
  }

 private:
  int send(const std::vector<StagePart>& parts, size_t k0, size_t k1, size_t data) {
    const size_t n = k1 - k0;
    svxg_ctx::Lane* L = nullptr;
    size_t at = 0;
    SVXG_TRY(stage_region(c_, s_, data + round256(n * sizeof(BulkSeg)), &L, &at));
    char* hp = static_cast<char*>(L->pinned) + at;
    char* dp = L->dstage.as<char>() + at;
    BulkSeg* seg = reinterpret_cast<BulkSeg*>(hp + data);
    std::vector<size_t> offs(n);
    size_t off = 0;
    for (size_t k = 0; k < n; ++k) {
      offs[k] = off;
      seg[k] = BulkSeg{parts[k0 + k].dst, dp + off, int64_t(parts[k0 + k].bytes)};
      off += round256(parts[k0 + k].bytes);
    }
    copy_in_parallel(n, [&](size_t k) { std::memcpy(hp + offs[k], parts[k0 + k].src, parts[k0 + k].bytes); });
    SVXG_HIP(hipMemcpyAsync(dp, hp, data + n * sizeof(BulkSeg), hipMemcpyHostToDevice, s_));
    launch_copy_many(s_, reinterpret_cast<const BulkSeg*>(dp + data), int(n));
    SVXG_HIP(hipGetLastError());
    if (!wait_) return stage_queued(L);
    SVXG_HIP(hipStreamSynchronize(s_));
    stage_idle(L);
    return SVXG_OK;
  }
  svxg_ctx* c_;
  hipStream_t s_;
  bool wait_;
  std::vector<StagePart> jobs_;
};

// Readbacks of a batch: one kernel packs them into the device staging buffer, one copy brings
// them over, one synchronisation, then they go to their destinations (a round trip per
// kStageCap bytes)
class Gatherer {
 public:
  Gatherer(svxg_ctx* c, hipStream_t s) : c_(c), s_(s) {}
  void add(void* dst, const void* dev, size_t bytes) {
    if (bytes) jobs_.push_back({dst, dev, bytes});
  }
  int flush() {
    if (jobs_.empty()) return SVXG_OK;
    const std::vector<StagePart> parts = stage_parts(jobs_);
    jobs_.clear();
    for (size_t k0 = 0, k1 = 0; k0 < parts.size(); k0 = k1) {
      size_t data = 0;
      k1 = stage_round(parts, k0, &data);
      SVXG_TRY(fetch(parts, k0, k1, data));
    }
    return SVXG_OK;
  }

 private:
  int fetch(const std::vector<StagePart>& parts, size_t k0, size_t k1, size_t data) {
    const size_t n = k1 - k0;
    svxg_ctx::Lane* L = nullptr;
    size_t at = 0;
    SVXG_TRY(stage_region(c_, s_, data + round256(n * sizeof(BulkSeg)), &L, &at));
    char* hp = static_cast<char*>(L->pinned) + at;
    char* dp = L->dstage.as<char>() + at;
    BulkSeg* seg = reinterpret_cast<BulkSeg*>(hp + data);
    size_t off = 0;
    for (size_t k = 0; k < n; ++k) {
      seg[k] = BulkSeg{dp + off, parts[k0 + k].src, int64_t(parts[k0 + k].bytes)};
      off += round256(parts[k0 + k].bytes);
    }
    SVXG_HIP(hipMemcpyAsync(dp + data, seg, n * sizeof(BulkSeg), hipMemcpyHostToDevice, s_));
    launch_copy_many(s_, reinterpret_cast<const BulkSeg*>(dp + data), int(n));
    SVXG_HIP(hipGetLastError());
    SVXG_HIP(hipMemcpyAsync(hp, dp, data, hipMemcpyDeviceToHost, s_));
    SVXG_HIP(hipStreamSynchronize(s_));
    stage_idle(L);
    off = 0;
    for (size_t k = k0; k < k1; ++k) {
      std::memcpy(parts[k].dst, hp + off, parts[k].bytes);
      off += round256(parts[k].bytes);
    }
    return SVXG_OK;
  }
  svxg_ctx* c_;
  hipStream_t s_;
  std::vector<StagePart> jobs_;
};

// Fills and FP64 -> FP32 copies of a batch setup: small ones gathered into one launch each, large
// ones (>= 4 MB) issued on their own; flush() before any launch reads what they write (wait: as
// Stager's, the stream synchronised at the end; else the work is only queued)
class BulkOps {
 public:
  explicit BulkOps(svxg_ctx* c) : c_(c) {}
  void zero(void* dst, size_t bytes) {
    if (!bytes) return;
    if (bytes >= kBigBytes) {
      big_zero_.push_back(BulkSeg{dst, nullptr, int64_t(bytes)});
      return;
    }
    for (size_t o = 0; o < bytes; o += kPieceBytes)
      zero_.push_back(BulkSeg{static_cast<char*>(dst) + o, nullptr, int64_t(std::min(kPieceBytes, bytes - o))});
  }
  void to_f32(const double* src, float* dst, int64_t n) {
    if (n <= 0) return;
    if (size_t(n) * 8 >= kBigBytes) {
      big_f32_.push_back(BulkSeg{dst, src, n});
      return;
    }
    const int64_t per = int64_t(kPieceBytes / 8);
    for (int64_t o = 0; o < n; o += per) f32_.push_back(BulkSeg{dst + o, src + o, std::min(per, n - o)});
  }
  int flush(hipStream_t s, bool wait = true) {
    for (const BulkSeg& g : big_zero_) SVXG_HIP(hipMemsetAsync(g.dst, 0, size_t(g.count), s));
    for (const BulkSeg& g : big_f32_)
      launch_to_f32(s, static_cast<const double*>(g.src), static_cast<float*>(g.dst), g.count);
    svxg_ctx::Lane* L = nullptr;
    if (!zero_.empty() || !f32_.empty()) {
      const size_t bytes = (zero_.size() + f32_.size()) * sizeof(BulkSeg);
      size_t at = 0;
      SVXG_TRY(stage_region(c_, s, bytes, &L, &at));
      char* hp = static_cast<char*>(L->pinned) + at;
      std::memcpy(hp, zero_.data(), zero_.size() * sizeof(BulkSeg));
      std::memcpy(hp + zero_.size() * sizeof(BulkSeg), f32_.data(), f32_.size() * sizeof(BulkSeg));
      SVXG_HIP(hipMemcpyAsync(L->dstage.as<char>() + at, hp, bytes, hipMemcpyHostToDevice, s));
      const BulkSeg* d = reinterpret_cast<const BulkSeg*>(L->dstage.as<char>() + at);
      launch_zero_many(s, d, int(zero_.size()));
      launch_f32_many(s, d + zero_.size(), int(f32_.size()));
    }
    SVXG_HIP(hipGetLastError());
    if (wait) {
      SVXG_HIP(hipStreamSynchronize(s));
      if (L) stage_idle(L);
    } else if (L) {
      SVXG_TRY(stage_queued(L));
    }
    zero_.clear();
    f32_.clear();
    big_zero_.clear();
    big_f32_.clear();
    return SVXG_OK;
  }

 private:
  static constexpr size_t kBigBytes = size_t(4) << 20;
  static constexpr size_t kPieceBytes = size_t(256) << 10;  // (a workgroup a piece)
  svxg_ctx* c_;
  std::vector<BulkSeg> zero_, f32_, big_zero_, big_f32_;
};

// the same through the stream's staging lane, queued (Stager without waiting: the caller keeps d, and
// whatever its queued work uses, until its next synchronisation of s)
template <typename T>
static int upload_queued(svxg_ctx* c, DevBuf& d, Arena* a, const std::vector<T>& h, hipStream_t s) {
  Stager up(c, s, false);
  SVXG_TRY(up.add(d, a, h));
  return up.flush();
}
template <typename T>
static int download(std::vector<T>& h, const DevBuf& d, size_t count, hipStream_t s) {
  h.resize(count);
  if (count) SVXG_HIP(hipMemcpyAsync(h.data(), d.as<void>(), count * sizeof(T), hipMemcpyDeviceToHost, s));
  SVXG_HIP(hipStreamSynchronize(s));
  return SVXG_OK;
}

}  // namespace svxg

using namespace svxg;

int64_t svxg_problem::device_bytes() const {
  int64_t bytes = 0;
  for (const DevBuf* d : {&d_nodes, &d_nodec, &d_bonds, &d_mats, &d_dof, &d_bond_slot, &d_cptr, &d_cidx, &d_row_gone,
                          &d_id_scale, &x, &r, &z, &p, &q, &b, &ix, &ir, &iz, &ip, &iq, &u_all})
    bytes += int64_t(d->held());
  for (const HostLevel& L : lv)
    for (const DevBuf* d : {&L.rowptr, &L.col, &L.A64, &L.A32, &L.A16, &L.D64, &L.D32, &L.Prow, &L.Pcol, &L.P64, &L.P32, &L.P16, &L.R16, &L.D16,
                            &L.Rrow, &L.Rfine, &L.Rent, &L.Minv, &L.x64[0], &L.x64[1], &L.b64,
                            &L.r64, &L.d64, &L.x32[0], &L.x32[1], &L.b32, &L.r32, &L.d32, &L.pos, &L.lam})
      bytes += int64_t(d->held());
  return bytes;
}

namespace svxg {

// ---- batches ----------------------------------------------------------------------------------------

struct Batch {
  std::vector<svxg_problem*> probs;
  int G = 0;
  int maxlev = 1;
  std::vector<GraphDev> hG;
  DevBuf dG, dstate;
  std::vector<PcgState> hstate;
  DevBuf t0;       // level-0 tiles of every graph (their slots: the partial sums' layout)
  int nt0 = 0;
  DevBuf t0run;    // those of the tiled graphs (what the tiled launches run over)
  int nt0run = 0;
  // host copies of the tiled launches' tables (t0run, tj, and each level's below), from which the
  // tables are cut to the running graphs in a batch's tail (compact_tiles) and made whole again
  std::vector<Tile> h_t0run, h_tj;
  bool compacted = false;
  DevBuf gtile;
  DevBuf tj;
  int ntj = 0;
  DevBuf res;      // resident graphs (one workgroup each: k_resident_steps)
  int nres = 0;
  // per level: smoothing tiles (S = 1 and S = 10 sub-groups a row), prolongation tiles, restriction
  // tiles (coarse rows, S = 1 / 10), stalled-coarsest tiles (S = 1 / 10), dense-coarsest graphs
  // the autotuner's state of one launch kind of a level (0 smoothing, 1 prolongation, 2 FP16
  // restriction, 3 dense solve): its shape key and, while the shape is tuned, every candidate's table
  struct Tune {
    uint64_t key = 0;
    int ncand = 0;       // candidates (0: not tuned here)
    bool tuning = false;
    int next = 0;        // the candidate the next launch takes
    DevBuf tiles[4];
    int nt[4] = {0, 0, 0, 0};
    std::vector<Tile> htiles[4];
  };
  struct Lv {
    DevBuf ts[2], tp, tc[2], tstall[2], dense;
    int nts[2] = {0, 0}, ntp = 0, ntc[2] = {0, 0}, ntstall[2] = {0, 0}, ndense = 0;
    std::vector<Tile> hts[2], htp, htc[2], htstall[2], hdense;
    DevBuf dpart, dcnt;  // (few dense chunks: the solve a wavefront a (chunk, span), its partials and counters)
    bool dsplit = false;  // the dense solve a wavefront a span (the tuned or default choice)
    Tune tune[4];
  };
  // dense coarsest solves of at most this many 64-row chunks a level are spread a wavefront a span
  static constexpr int kDenseSplitChunks = 64;
  // the autotuner's timed launches not yet read (events, key, candidate); its event pairs for reuse
  struct TunePending {
    hipEvent_t a, b;
    uint64_t key;
    int cand;
  };
  std::vector<TunePending> tune_pending;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> tune_events;
  int tuning = 0;  // level kinds still tuning (no step graph is recorded meanwhile)
  // every tiled graph has a multigrid with smoothing at level 0: the PCG's r . z comes from the
  // V-cycle's last level-0 step, run over the PCG's tiles (no dot_rz launch)
  bool fuse_rz = false;
  std::vector<Lv> lvl;
  DevBuf partial;
  DevBuf prof_gb;  // kernel timing: 4 x G algorithmic bytes (kind-major) of one launch, per graph
  bool any_amg = false;
  // one tiled step recorded as a HIP graph, by (inner-only, mixed, FP32 products, FP64 products): the launches'
  // arguments are this batch's tables, fixed for its life; what a step does to each graph is read
  // by the kernels from the graph's state, so every replay is the launches made again
  hipGraphExec_t step_graph[16] = {};
  bool graphs_failed = false;
  int windows = 0;  // windows run (a batch records its graph from its third window: short solves
                    // and short-lived serving batches never pay for a recording)
  // (measured on MI355X: a replayed step is no faster on the device than the same launches and the
  // recording costs a step's time, so only batches with at least kGraphMinTiles level-0 tiles, whose
  // host launch time shows in the wall time, record one)
  static constexpr int kGraphMinTiles = 8192;
  Batch() = default;
  Batch(const Batch&) = delete;
  Batch& operator=(const Batch&) = delete;
  ~Batch() {
    for (hipGraphExec_t& g : step_graph)
      if (g) (void)hipGraphExecDestroy(g);
    for (auto& e : tune_events) {
      (void)hipEventDestroy(e.first);
      (void)hipEventDestroy(e.second);
    }
    for (auto& e : tune_pending) {
      (void)hipEventDestroy(e.a);
      (void)hipEventDestroy(e.b);
    }
  }
};

// rows averaging more than 16 entries are served a wavefront each (S = 10 sub-groups): the
// restriction of levels without FP16 copies (it reads P's blocks transposed in place)
static bool long_rows(int64_t entries, int64_t rows) { return rows > 0 && entries > 16 * rows; }

// Products a block a lane (smoothing, residuals, prolongation, the FP16 restriction) take a tile's
// rows' blocks kLpbBlocks a pass; a level's tiles hold as many rows (at most kTileRows) as fill about
// one pass at its mean entries a row, so that long coarse rows still make many workgroups. Every row's
// sums are the same whatever tile holds it, so this sets only the time.
// (a pass is kWG blocks over FP32 / FP16 blocks, kWG / 2 over FP64 ones: two lanes a block; target:
// the blocks a tile should hold, the autotuner's variable, kLpbBlocks untuned)
constexpr int kLpbBlocks = 256;
static int lpb_tile_rows(int64_t entries, int64_t rows, bool fp64, int target = kLpbBlocks) {
  if (rows <= 0) return kTileRows;
  const int64_t per = std::max<int64_t>(1, (entries + rows - 1) / rows);
  return int(std::clamp<int64_t>((fp64 ? target / 2 : target) / per, 1, kTileRows));
}

// ---- the autotuner's shapes -------------------------------------------------------------------------

constexpr int kTuneTargets[4] = {64, 128, 256, 512};
constexpr int kTuneDefault = 2;  // (256 blocks)
constexpr int kTuneSamples = 3;  // launches timed a candidate

static int ilog2(int64_t v) {
  int k = 0;
  while (v > 1) {
    v >>= 1;
    ++k;
  }
  return k;
}
// a launch shape: its kind, precision and storage, its rows over the batch (log2), entries a row, and
// the graphs it spans (log2: one graph's level and a batch's many small ones of the same total rows
// are different launches - their tiles' sizes and counts differ)
static uint64_t tune_key(int kind, bool fp64, bool half, int64_t rows, int64_t entries, int64_t graphs) {
  const int64_t per = rows > 0 ? std::min<int64_t>(255, (entries + rows / 2) / rows) : 0;
  return uint64_t(kind) | uint64_t(fp64) << 4 | uint64_t(half) << 5 | uint64_t(ilog2(rows)) << 8 | uint64_t(per) << 16 |
         uint64_t(ilog2(graphs)) << 24;
}
static svxg_tune_record tune_record_of(uint64_t key) {
  svxg_tune_record r;
  std::memset(&r, 0, sizeof r);
  r.kind = int32_t(key & 15);
  r.fp64 = int32_t(key >> 4 & 1);
  r.half = int32_t(key >> 5 & 1);
  r.rows_log2 = int32_t(key >> 8 & 255);
  r.per_row = int32_t(key >> 16 & 255);
  r.graphs_log2 = int32_t(key >> 24 & 255);
  return r;
}
// the cached choice of a shape (-1: none yet)
static int tune_choice(svxg_ctx* c, uint64_t key) {
  std::lock_guard<std::mutex> lk(c->tune_mu);
  auto it = c->tune_choice.find(key);
  return it == c->tune_choice.end() ? -1 : it->second;
}

static void fill_desc(svxg_problem* P, GraphDev& g, PcgState* st) {
  std::memset(&g, 0, sizeof g);
  g.nlev = P->jacobi_only ? 1 : int32_t(P->lv.size());
  g.nfree = P->nfree;
  g.jacobi = P->jacobi_only ? 1 : 0;
  g.st = st;
  g.x = P->x.as<double>();
  g.r = P->r.as<double>();
  g.z = P->z.as<double>();
  g.p = P->p.as<double>();
  g.q = P->q.as<double>();
  g.b = P->b.as<double>();
  g.ix = P->ix.as<float>();
  g.ir = P->ir.as<float>();
  g.iz = P->iz.as<float>();
  g.ip = P->ip.as<float>();
  g.iq = P->iq.as<float>();
  for (size_t l = 0; l < P->lv.size() && l < size_t(kMaxLevels); ++l) {
    HostLevel& H = P->lv[l];
    LevelDev& L = g.lv[l];
    L.n = H.n;
    L.nc = H.nc;
    L.terminal = H.terminal ? 1 : 0;
    L.dense_m = H.dense_m;
    L.nnzb = H.nnzb;
    L.np = H.np;
    L.rowptr = H.rowptr.as<int32_t>();
    L.col = H.col.as<int32_t>();
    L.Prow = H.Prow.as<int32_t>();
    L.Pcol = H.Pcol.as<int32_t>();
    L.Rrow = H.Rrow.as<int32_t>();
    L.Rfine = H.Rfine.as<int32_t>();
    L.Rent = H.Rent.as<int32_t>();
    L.Minv = H.Minv.as<double>();
    L.lmax = H.lmax;
    L.cheb_degree = P->opt.amg.cheb_degree;
    L.sub = 1;
    L.rsub = H.R16.as<void>() || !long_rows(H.np, H.nc) ? 1 : 10;
    L.f64.A = H.A64.as<double>();
    L.f64.Dinv = H.D64.as<double>();
    L.f64.P = H.P64.as<double>();
    L.f64.x[0] = H.x64[0].as<double>();
    L.f64.x[1] = l == 0 ? P->z.as<double>() : H.x64[1].as<double>();
    L.f64.b = l == 0 ? P->r.as<double>() : H.b64.as<double>();
    L.f64.r = H.r64.as<double>();
    L.f64.d = H.d64.as<double>();
    L.f32.A = H.A32.as<float>();
    L.A16 = H.A16.as<uint8_t>();
    L.P16 = H.P16.as<uint8_t>();
    L.D16 = H.D16.as<uint8_t>();
    L.R16 = H.R16.as<uint8_t>();
    L.f32.Dinv = H.D32.as<float>();
    L.f32.P = H.P32.as<float>();
    L.f32.x[0] = H.x32[0].as<float>();
    L.f32.x[1] = l == 0 ? P->iz.as<float>() : H.x32[1].as<float>();
    L.f32.b = l == 0 ? P->ir.as<float>() : H.b32.as<float>();
    L.f32.r = H.r32.as<float>();
    L.f32.d = H.d32.as<float>();
  }
}

static int prof_table(Batch& B, Arena* A, hipStream_t s);

static int build_batch(Batch& B, svxg_problem* const* probs, int count, Arena* A, hipStream_t s,
                       int resident_rows = -1) {
  svxg_ctx* c = count ? probs[0]->ctx : nullptr;
  if (resident_rows < 0) resident_rows = c ? c->resident_rows : 0;
  for (int g = 0; g < count; ++g) {
    const svxg_problem* P = probs[g];
    if (P->opt.precision == SVXG_FP64 && !P->fp64_levels && !P->jacobi_only) {
      set_error("a problem set up for mixed precision keeps no FP64 hierarchy: assemble it again for FP64");
      return SVXG_NOT_ASSEMBLED;
    }
  }
  B.probs.assign(probs, probs + count);
  B.G = count;
  SVXG_TRY(B.dstate.make(A, sizeof(PcgState) * size_t(std::max(1, count))));
  B.hstate.assign(size_t(count), PcgState{});
  B.hG.resize(size_t(count));
  B.maxlev = 1;
  B.any_amg = false;
  std::vector<int32_t> res;
  std::vector<char> tiled(static_cast<size_t>(count), 1);
  for (int g = 0; g < count; ++g) {
    svxg_problem* P = probs[g];
    fill_desc(P, B.hG[size_t(g)], B.dstate.as<PcgState>() + g);
    if (resident_rows > 0 && P->nfree <= resident_rows) {
      B.hG[size_t(g)].resident = 1;
      tiled[size_t(g)] = 0;
      res.push_back(g);
      continue;
    }
    if (!P->jacobi_only) {
      B.maxlev = std::max(B.maxlev, int(P->lv.size()));
      B.any_amg = true;
    }
  }
  B.nres = int(res.size());
  std::vector<Tile> t0, t0run, tj;
  std::vector<int32_t> gt(static_cast<size_t>(count) + 1, 0);
  for (int g = 0; g < count; ++g) {
    gt[size_t(g)] = int32_t(t0.size());
    const size_t first = t0.size();
    add_tiles(t0, g, probs[g]->nfree, kTileRows);
    if (!tiled[size_t(g)]) continue;
    t0run.insert(t0run.end(), t0.begin() + std::ptrdiff_t(first), t0.end());
    if (probs[g]->jacobi_only) add_tiles(tj, g, probs[g]->nfree, kTileRows);
  }
  gt[size_t(count)] = int32_t(t0.size());
  // (the split dense solves' counters start at 0)
  const std::vector<int32_t> zc(size_t(Batch::kDenseSplitChunks), 0);
  B.nt0 = int(t0.size());
  B.nt0run = int(t0run.size());
  B.fuse_rz = !t0run.empty() && tj.empty();
  for (int g = 0; g < count && B.fuse_rz; ++g)
    if (tiled[size_t(g)] && (probs[g]->jacobi_only || probs[g]->lv.empty() || probs[g]->lv[0].terminal))
      B.fuse_rz = false;
  B.ntj = int(tj.size());
  Stager up(c, s);
  SVXG_TRY(up.add(B.dG, A, B.hG));
  SVXG_TRY(up.add(B.t0, A, t0));
  SVXG_TRY(up.add(B.t0run, A, t0run));
  SVXG_TRY(up.add(B.tj, A, tj));
  B.h_t0run = t0run;
  B.h_tj = tj;
  B.compacted = false;
  SVXG_TRY(up.add(B.gtile, A, gt));
  SVXG_TRY(up.add(B.res, A, res));
  SVXG_TRY(B.partial.make(A, sizeof(double) * 2 * size_t(std::max(1, B.nt0))));
  B.lvl.clear();
  B.lvl.resize(size_t(B.maxlev));
  struct LvTiles {
    std::vector<Tile> ts[2], tp, tc[2], tstall[2];
    std::vector<Tile> dense;  // (64-row chunks of the dense inverses)
  };
  std::vector<LvTiles> lt(static_cast<size_t>(B.maxlev));
  // (the V-cycle's precision: FP64 on the plain path)
  const bool v64 = count > 0 && probs[0]->opt.precision == SVXG_FP64;
  std::vector<std::vector<Tile>> cand_tiles;  // (the candidate tables of shapes tuned here, until the flush)
  cand_tiles.reserve(size_t(B.maxlev) * 12);
  for (int l = 0; l < B.maxlev; ++l) {
    LvTiles& T = lt[size_t(l)];
    Batch::Lv& V = B.lvl[size_t(l)];
    // the level's shape a launch kind: rows and entries over the batch's tiled graphs
    int64_t rows[3] = {0, 0, 0}, ent[3] = {0, 0, 0}, dense_rows = 0;
    bool half = true;
    std::vector<int> members;
    for (int g = 0; g < count; ++g) {
      svxg_problem* P = probs[g];
      if (!tiled[size_t(g)] || P->jacobi_only || l >= int(P->lv.size())) continue;
      const HostLevel& H = P->lv[size_t(l)];
      members.push_back(g);
      half = half && H.A16.as<void>() != nullptr;
      if (!H.terminal) {
        rows[0] += H.n;
        ent[0] += H.nnzb;
        rows[1] += H.n;
        ent[1] += H.np;
        rows[2] += P->lv[size_t(l) + 1].n;
        ent[2] += H.np;
      } else if (H.dense_m > 0) {
        dense_rows += H.dense_m;
      }
    }
    // each kind's blocks a tile: the shape's tuned choice, else (tuning it here, or untuned) the default
    int target[3] = {kTuneTargets[kTuneDefault], kTuneTargets[kTuneDefault], kTuneTargets[kTuneDefault]};
    for (int k = 0; k < 3; ++k) {
      Batch::Tune& tu = V.tune[k];
      if (rows[k] == 0 || (k == 2 && v64)) continue;
      tu.key = tune_key(k, v64, half, rows[k], ent[k], int64_t(members.size()));
      const int ch = c ? tune_choice(c, tu.key) : -1;
      if (ch >= 0) {
        target[k] = kTuneTargets[ch];
      } else if (c && c->autotune) {
        tu.ncand = 4;
        tu.tuning = true;
        ++B.tuning;
      }
    }
    for (int g : members) {
      svxg_problem* P = probs[g];
      const HostLevel& H = P->lv[size_t(l)];
      const int srows = lpb_tile_rows(H.nnzb, H.n, v64, target[0]);
      if (!H.terminal) {
        add_tiles(T.ts[0], g, H.n, srows);
        add_tiles(T.tp, g, H.n, lpb_tile_rows(H.np, H.n, v64, target[1]));
        const int nc = P->lv[size_t(l) + 1].n;
        if (!v64)  // (FP16 copies, or FP32 P's blocks read transposed in place: a block a lane)
          add_tiles(T.tc[0], g, nc, lpb_tile_rows(H.np, nc, v64, target[2]));
        else if (!long_rows(H.np, H.nc))
          add_tiles(T.tc[0], g, nc, tile_rows<1>());
        else
          add_tiles(T.tc[1], g, nc, tile_rows<10>());
      } else if (H.dense_m > 0) {
        for (int r = 0; r < H.dense_m; r += 64) T.dense.push_back(Tile{g, r, std::min(64, H.dense_m - r), 0});
      } else {
        add_tiles(T.tstall[0], g, H.n, srows);
      }
    }
    // candidate tables of the kinds tuned here (the same rows, other tilings)
    for (int k = 0; k < 3; ++k) {
      Batch::Tune& tu = V.tune[k];
      if (!tu.tuning) continue;
      for (int ci = 0; ci < tu.ncand; ++ci) {
        cand_tiles.emplace_back();
        std::vector<Tile>& ct = cand_tiles.back();
        for (int g : members) {
          const svxg_problem* P = probs[g];
          const HostLevel& H = P->lv[size_t(l)];
          if (H.terminal) continue;
          const int tg = kTuneTargets[ci];
          if (k == 0) add_tiles(ct, g, H.n, lpb_tile_rows(H.nnzb, H.n, v64, tg));
          else if (k == 1) add_tiles(ct, g, H.n, lpb_tile_rows(H.np, H.n, v64, tg));
          else add_tiles(ct, g, P->lv[size_t(l) + 1].n, lpb_tile_rows(H.np, P->lv[size_t(l) + 1].n, v64, tg));
        }
        tu.nt[ci] = int(ct.size());
        tu.htiles[ci] = ct;
        SVXG_TRY(up.add(tu.tiles[ci], A, ct));
      }
    }
    // the dense solve: a wavefront a span for few chunks (tuned against a workgroup a chunk)
    if (!T.dense.empty() && int(T.dense.size()) <= Batch::kDenseSplitChunks) {
      Batch::Tune& tu = V.tune[3];
      tu.key = tune_key(3, v64, false, dense_rows, int64_t(T.dense.size()) * dense_rows, int64_t(members.size()));
      const int ch = c ? tune_choice(c, tu.key) : -1;
      V.dsplit = ch >= 0 ? ch == 1 : true;
      if (ch < 0 && c && c->autotune) {
        tu.ncand = 2;
        tu.tuning = true;
        ++B.tuning;
      }
    }
    for (int v = 0; v < 2; ++v) {
      V.hts[v] = T.ts[v];
      V.htc[v] = T.tc[v];
      V.htstall[v] = T.tstall[v];
      V.nts[v] = int(T.ts[v].size());
      V.ntc[v] = int(T.tc[v].size());
      V.ntstall[v] = int(T.tstall[v].size());
      SVXG_TRY(up.add(V.ts[v], A, T.ts[v]));
      SVXG_TRY(up.add(V.tc[v], A, T.tc[v]));
      SVXG_TRY(up.add(V.tstall[v], A, T.tstall[v]));
    }
    V.htp = T.tp;
    V.hdense = T.dense;
    V.ntp = int(T.tp.size());
    V.ndense = int(T.dense.size());
    SVXG_TRY(up.add(V.tp, A, T.tp));
    SVXG_TRY(up.add(V.dense, A, T.dense));
    if (V.ndense > 0 && V.ndense <= Batch::kDenseSplitChunks && (V.dsplit || V.tune[3].tuning)) {
      SVXG_TRY(V.dpart.make(A, sizeof(double) * size_t(V.ndense) * kDenseSpans * 64));
      SVXG_TRY(V.dcnt.make(A, sizeof(int32_t) * size_t(V.ndense)));
      SVXG_TRY(up.add_raw(V.dcnt.as<void>(), zc.data(), sizeof(int32_t) * size_t(V.ndense)));
    }
  }
  return up.flush();
}

// ---- kernel timing ----------------------------------------------------------------------------------

// algorithmic bytes of one graph in one launch: the outer K product (K's blocks and column indices,
// row pointers, p gathered once, q written) and a level-0 Chebyshev step (the same matrix traffic,
// plus b, D^-1, d read and written, x written); kinds 0/2 FP64, 1/3 FP32
static double prof_bytes(const svxg_problem* P, int kind) {
  if (P->lv.empty()) return 0.0;
  const double w = (kind & 1) ? 4.0 : 8.0;
  const double nb = double(P->lv[0].nnzb), n = double(P->lv[0].n);
  // (the smoothing step of a level with FP16 copies reads kHalfBlock bytes a block)
  const double blk = (kind == 3 && P->lv[0].A16) ? double(kHalfBlock) : 36.0 * w;
  const double dblk = (kind == 3 && P->lv[0].D16) ? double(kHalfBlock) : 36.0 * w;
  const double mat = nb * (blk + 4.0) + 4.0 * (n + 1.0);
  return kind < 2 ? mat + 2.0 * 6.0 * n * w : mat + dblk * n + 5.0 * 6.0 * n * w;
}

// the batch's per-graph byte table (made when the batch is built with timing on)
static int prof_table(Batch& B, Arena* A, hipStream_t s) {
  std::vector<double> gb(4 * size_t(B.G));
  for (int k = 0; k < 4; ++k)
    for (int g = 0; g < B.G; ++g) gb[size_t(k) * size_t(B.G) + size_t(g)] = prof_bytes(B.probs[size_t(g)], k);
  svxg_ctx* c = B.probs[0]->ctx;
  if (!c->prof_acc.as<void>()) {
    SVXG_TRY(c->prof_acc.make(A, 4 * sizeof(double)));
    SVXG_HIP(hipMemsetAsync(c->prof_acc.as<void>(), 0, 4 * sizeof(double), s));
  }
  return upload(B.prof_gb, A, gb, s);
}

// The launch between construction and destruction is timed with an event pair; it is given pa(),
// with which the workgroups of each served graph's first tile add that graph's bytes (live for
// `inner`) to the context's device counter, so launches issued after a graph stopped add time and
// no bytes.
// (the events go to the launch, whose own dispatch records them at the kernel's start and end:
// t0(), t1(); no marker packets around it, so the gap before a kernel starts is not counted)
struct ProfScope {
  svxg_ctx* c;
  const Batch& B;
  int kind, inner;
  hipStream_t s;
  std::pair<hipEvent_t, hipEvent_t> ev{nullptr, nullptr};
  ProfScope(const Batch& b, int k, hipStream_t st, int in) : c(b.probs[0]->ctx), B(b), kind(k), inner(in), s(st) {
    if (!c->profiling || kind < 0 || !c->prof_acc.as<void>() || !B.prof_gb.as<void>()) return;
    auto& P = c->prof[kind];
    if (!P.spare.empty()) {
      ev = P.spare.back();
      P.spare.pop_back();
    } else {
      (void)hipEventCreate(&ev.first);
      (void)hipEventCreate(&ev.second);
    }
    P.pending.push_back(ev);
  }
  ProfAcc pa() const {
    if (!ev.first) return {};
    return {B.prof_gb.as<double>() + size_t(kind) * size_t(B.G), c->prof_acc.as<double>() + kind};
  }
  hipEvent_t t0() const { return ev.first; }
  hipEvent_t t1() const { return ev.second; }
};

// after a synchronisation: the pending pairs' times added up
static void prof_collect(svxg_ctx* c) {
  if (!c->profiling) return;
  for (auto& P : c->prof) {
    for (auto& ev : P.pending) {
      float ms = 0.0f;
      if (hipEventElapsedTime(&ms, ev.first, ev.second) == hipSuccess) {
        P.ms += ms;
        ++P.launches;
      }
      P.spare.push_back(ev);
    }
    P.pending.clear();
  }
}

// ---- the V-cycle -----------------------------------------------------------------------------------


// the smoother's degrees of a problem's options (level_degree's packing)
static int batch_degree(const svxg_problem* P) {
  const int d0 = std::clamp(P->opt.amg.cheb_degree, 1, 255);
  const int dc = std::clamp(P->opt.amg.coarse_cheb_degree, 0, 255);
  return d0 | (dc << 8);
}

// ---- the autotuner at work -------------------------------------------------------------------------

// the table a launch of a tuned kind takes now: while its shape is tuned, the candidate in turn
struct TuneUse {
  const Tile* t;
  int n;
  int cand;  // (-1: not timed)
};
static TuneUse tune_pick(const Batch::Tune& tu, const DevBuf& own, int nown) {
  if (!tu.tuning) return {own.as<Tile>(), nown, -1};
  return {tu.tiles[tu.next].as<Tile>(), tu.nt[tu.next], tu.next};
}
// an event pair around one timed launch (read at the next synchronisation: tune_collect)
struct TuneTimer {
  Batch& B;
  hipStream_t s;
  uint64_t key;
  int cand;
  std::pair<hipEvent_t, hipEvent_t> ev{nullptr, nullptr};
  TuneTimer(Batch& b, hipStream_t st, uint64_t k, int c) : B(b), s(st), key(k), cand(c) {
    if (cand < 0) return;
    if (!B.tune_events.empty()) {
      ev = B.tune_events.back();
      B.tune_events.pop_back();
    } else if (hipEventCreate(&ev.first) != hipSuccess || hipEventCreate(&ev.second) != hipSuccess) {
      cand = -1;
      return;
    }
    (void)hipEventRecord(ev.first, s);
  }
  ~TuneTimer() {
    if (cand < 0) return;
    (void)hipEventRecord(ev.second, s);
    B.tune_pending.push_back(Batch::TunePending{ev.first, ev.second, key, cand});
  }
};
// after a V-cycle: every kind still tuned takes its next candidate
static void tune_advance(Batch& B) {
  if (!B.tuning) return;
  for (Batch::Lv& V : B.lvl)
    for (Batch::Tune& tu : V.tune)
      if (tu.tuning) tu.next = (tu.next + 1) % tu.ncand;
}

// After a synchronisation: the timed launches' times added to their shapes' samples; a shape with
// kTuneSamples of every candidate gets the fastest mean as its choice (kept in the context), and the
// batch's levels of that shape take its table from now on.
static void tune_collect(Batch& B) {
  if (B.probs.empty() || (B.tune_pending.empty() && !B.tuning)) return;
  svxg_ctx* c = B.probs[0]->ctx;
  std::lock_guard<std::mutex> lk(c->tune_mu);
  for (const Batch::TunePending& p : B.tune_pending) {
    float ms = 0.0f;
    if (hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
      svxg_ctx::TuneAcc& acc = c->tune_acc[p.key];
      acc.ms[p.cand] += ms;
      ++acc.n[p.cand];
    }
    B.tune_events.emplace_back(p.a, p.b);
  }
  B.tune_pending.clear();
  for (Batch::Lv& V : B.lvl)
    for (int k = 0; k < 4; ++k) {
      Batch::Tune& tu = V.tune[k];
      if (!tu.tuning) continue;
      int choice = -1;
      auto known = c->tune_choice.find(tu.key);
      if (known != c->tune_choice.end()) {
        choice = known->second;  // (decided by another batch meanwhile)
      } else {
        const svxg_ctx::TuneAcc& acc = c->tune_acc[tu.key];
        bool all = true;
        for (int ci = 0; ci < tu.ncand; ++ci) all = all && acc.n[ci] >= kTuneSamples;
        if (!all) continue;
        svxg_tune_record rec = tune_record_of(tu.key);
        rec.ncand = tu.ncand;
        rec.samples = kTuneSamples;
        double best = 0.0;
        for (int ci = 0; ci < tu.ncand; ++ci) {
          const double us = 1e3 * acc.ms[ci] / acc.n[ci];
          rec.us[ci] = us;
          rec.cand[ci] = k == 3 ? ci : kTuneTargets[ci];
          if (choice < 0 || us < best) {
            choice = ci;
            best = us;
          }
        }
        rec.choice = choice;
        c->tune_choice[tu.key] = choice;
        c->tune_log.push_back(rec);
        c->tune_acc.erase(tu.key);
      }
      // (the chosen table becomes the level's own; no launch is pending: a synchronisation came first)
      tu.tuning = false;
      --B.tuning;
      if (k == 0) {
        V.ts[0] = std::move(tu.tiles[choice]);
        V.nts[0] = tu.nt[choice];
        V.hts[0] = std::move(tu.htiles[choice]);
      } else if (k == 1) {
        V.tp = std::move(tu.tiles[choice]);
        V.ntp = tu.nt[choice];
        V.htp = std::move(tu.htiles[choice]);
      } else if (k == 2) {
        V.tc[0] = std::move(tu.tiles[choice]);
        V.ntc[0] = tu.nt[choice];
        V.htc[0] = std::move(tu.htiles[choice]);
      } else {
        V.dsplit = choice == 1;
      }
      for (DevBuf& d : tu.tiles) d = DevBuf();
      for (std::vector<Tile>& h : tu.htiles) h.clear();
    }
}

// the V-cycle of a batch; `degree` packs the smoother's degrees (level_degree). Launches of kinds
// whose shape is being tuned take the candidate tables in turn (tune_pick), one launch of each timed.
template <typename T>
static void vcycle(Batch& B, hipStream_t s, int inner, int degree, bool dot = false) {
  const GraphDev* G = B.dG.as<GraphDev>();
  if (B.ntj) launch_jacobi<T>(s, G, B.tj.as<Tile>(), B.ntj, inner);
  if (!B.any_amg) return;
  const int L = B.maxlev;
  const int SV[2] = {1, 10};
  for (int l = 0; l < L; ++l) {
    Batch::Lv& V = B.lvl[size_t(l)];
    const int dg = level_degree(degree, l);
    const TuneUse sm = tune_pick(V.tune[0], V.ts[0], V.nts[0]);
    if (sm.n) {
      for (int st = 0; st < dg; ++st) {
        ProfScope ps(B, l == 0 && st > 0 ? (sizeof(T) == 4 ? 3 : 2) : -1, s, inner);
        launch_cheb_step<T>(s, G, sm.t, sm.n, l, st, (st + 1) & 1, st & 1, st == 0, inner, nullptr, 1, ps.pa(), ps.t0(),
                            ps.t1());
      }
      TuneTimer tt(B, s, V.tune[0].key, sm.cand);
      launch_residual<T>(s, G, sm.t, sm.n, l, (dg - 1) & 1, inner, 1);
    }
    {
      const TuneUse rs = tune_pick(V.tune[2], V.tc[0], V.ntc[0]);
      TuneTimer tt(B, s, V.tune[2].key, rs.cand);
      launch_restrict<T>(s, G, rs.t, rs.n, l, inner, 1);
    }
    launch_restrict<T>(s, G, V.tc[1].as<Tile>(), V.ntc[1], l, inner, 10);
    if (V.ndense) {
      const Batch::Tune& tu = V.tune[3];
      const bool split = tu.tuning ? tu.next == 1 : V.dsplit;
      TuneTimer tt(B, s, tu.key, tu.tuning ? tu.next : -1);
      launch_dense_solve<T>(s, G, V.dense.as<Tile>(), V.ndense, l, inner, split ? V.dpart.as<double>() : nullptr,
                            split ? V.dcnt.as<int32_t>() : nullptr);
    }
    for (int v = 0; v < 2; ++v)
      if (V.ntstall[v])
        for (int st = 0; st < kStallDegree; ++st)
          launch_cheb_step<T>(s, G, V.tstall[v].as<Tile>(), V.ntstall[v], l, st, (st + 1) & 1, st & 1, st == 0, inner,
                              nullptr, SV[v]);
  }
  for (int l = L - 2; l >= 0; --l) {
    Batch::Lv& V = B.lvl[size_t(l)];
    const TuneUse pr = tune_pick(V.tune[1], V.tp, V.ntp);
    if (!pr.n) continue;
    const int dg = level_degree(degree, l);
    const int cur = (dg - 1) & 1;
    {
      TuneTimer tt(B, s, V.tune[1].key, pr.cand);
      launch_prolong<T>(s, G, pr.t, pr.n, l, cur, inner);
    }
    const TuneUse sm = tune_pick(V.tune[0], V.ts[0], V.nts[0]);
    for (int j = 0; j < dg; ++j) {
      const int gs = dg + j;
      if (!sm.n) continue;
      if (dot && B.fuse_rz && l == 0 && j == dg - 1) {
        // (the PCG's r . z partials on its own tiles: k_cheb_step_rz, not among the timed plain steps)
        launch_cheb_step<T>(s, G, B.t0run.as<Tile>(), B.nt0run, l, j, (gs + 1) & 1, gs & 1, 0, inner,
                            B.partial.as<double>(), 1);
        continue;
      }
      ProfScope ps(B, l == 0 ? (sizeof(T) == 4 ? 3 : 2) : -1, s, inner);
      launch_cheb_step<T>(s, G, sm.t, sm.n, l, j, (gs + 1) & 1, gs & 1, 0, inner, nullptr, 1, ps.pa(), ps.t0(), ps.t1());
    }
  }
  tune_advance(B);
}

static void prof_collect(svxg_ctx* c);
static int read_states(Batch& B, hipStream_t s) {
  SVXG_HIP(hipMemcpyAsync(B.hstate.data(), B.dstate.as<void>(), sizeof(PcgState) * size_t(B.G), hipMemcpyDeviceToHost, s));
  SVXG_HIP(hipStreamSynchronize(s));
  if (!B.probs.empty()) prof_collect(B.probs[0]->ctx);
  tune_collect(B);
  return SVXG_OK;
}

static bool any_running(const Batch& B, bool inner) {
  for (const PcgState& st : B.hstate)
    if (st.state == ST_RUNNING && (!inner || st.inner_state == ST_RUNNING)) return true;
  return false;
}

// The host looks at the graphs' states (a synchronisation) to learn when every solve has stopped;
// between looks it queues a window of iterations. A graph that stopped early lets the rest of the
// window run through as empty launches, so the window is sized from the running graphs' own
// residual histories: the iterations the slowest of them needs at its observed rate, at most
// kPollMax (kPollFirst before a rate is known).
constexpr int kPollFirst = 4;
constexpr int kPollMax = 16;

struct PollPlan {
  struct Seen {
    double rn = 0.0;
    int32_t it = -1;
  };
  std::unordered_map<const svxg_problem*, Seen> seen;  // (by problem: a batch may be rebuilt between looks)
  // iterations to queue before the next look (states just read)
  int next(const Batch& B, bool inner) {
    int want = 1;
    for (size_t g = 0; g < B.hstate.size(); ++g) {
      const PcgState& st = B.hstate[g];
      Seen& sn = seen[B.probs[g]];
      if (inner && st.state == ST_RUNNING && st.phase == PH_WAIT64) {
        sn.it = -1;  // (released at the next look)
        continue;
      }
      if (inner && st.state == ST_RUNNING && st.inner_state != ST_RUNNING) {
        // (between inner solves: a refinement residual, then a new inner solve)
        want = std::max(want, kPollFirst);
        sn.it = -1;
        continue;
      }
      const bool run = st.state == ST_RUNNING && (!inner || st.inner_state == ST_RUNNING);
      const double r = inner ? st.inner_rn : st.rn;
      const double target = inner ? st.inner_target : st.target;
      const int32_t k = inner ? st.inner_iters : st.iters;
      int need = kPollFirst;
      if (run && sn.it >= 0 && k > sn.it && r > 0 && sn.rn > r && target > 0) {
        const double rate = std::pow(r / sn.rn, 1.0 / double(k - sn.it));
        need = rate < 1.0 ? int(std::ceil(std::log(target / r) / std::log(rate))) : kPollMax;
      }
      if (run) want = std::max(want, std::min(kPollMax, std::max(1, need)));
      sn.rn = r;
      sn.it = run ? k : -1;
    }
    return want;
  }
  void forget(const svxg_problem* P) { seen.erase(P); }
};

static void drop_self_batch_ptrs(svxg_problem* P);

// one FP64 PCG iteration of StressProblem::iterate (stress.cpp:403-464) of every iterating graph,
// each with its own scalars and stopping point
static void fp64_step(Batch& B, hipStream_t s, int degree) {
  const GraphDev* G = B.dG.as<GraphDev>();
  const Tile* t0 = B.t0run.as<Tile>();
  const int32_t* gt = B.gtile.as<int32_t>();
  double* part = B.partial.as<double>();
  {
    ProfScope ps(B, 0, s, SERVE_OUTER);
    launch_pcg_spmv<double>(s, G, t0, B.nt0run, SERVE_OUTER, part, ps.pa(), ps.t0(), ps.t1());
  }
  launch_finish(s, FinishKind::Alpha, G, gt, B.G, SERVE_OUTER, part);
  launch_axpy_rr<double>(s, G, t0, B.nt0run, SERVE_OUTER, part);
  launch_finish(s, FinishKind::Rn, G, gt, B.G, SERVE_OUTER, part);
  vcycle<double>(B, s, SERVE_OUTER, degree, true);
  if (!B.fuse_rz) launch_dot_rz<double>(s, G, t0, B.nt0run, SERVE_OUTER, part);
  launch_finish(s, FinishKind::Beta, G, gt, B.G, SERVE_OUTER, part);
  launch_update_p<double>(s, G, t0, B.nt0run, SERVE_OUTER);
}

// a window of w steps: the resident graphs' in one launch, the tiled graphs' step by step.
// Mixed windows: only the first and last steps of a window launch the refinement and inner-start
// kernels (none when every running graph began the window iterating an inner solve, `lean`, but
// for the last step); the steps between run the inner iteration alone. A graph that leaves its
// inner solve in the middle of a window applies its correction and waits, untouched, for the last
// step, so it still finishes before the host's look at the window's end, and its arithmetic is the
// same - only the idle steps differ.
static void mixed_step(Batch& B, hipStream_t s, int degree, bool p32, bool p64, bool lean);
static void run_window(Batch& B, hipStream_t s, int w, bool mixed, bool p32, bool p64, bool lean = false) {
  const int degree = batch_degree(B.probs[0]);
  if (B.nres)
    launch_resident_steps(s, B.dG.as<GraphDev>(), B.gtile.as<int32_t>(), B.res.as<int32_t>(), B.nres, w, degree,
                          mixed ? 1 : 0, p32 ? 1 : 0, p64 ? 1 : 0, B.partial.as<double>());
  if (!B.nt0run) return;
  auto lean_at = [&](int k) { return mixed && k < w - 1 && (lean || k > 0); };
  const svxg_ctx* c = B.probs[0]->ctx;
  // (a batch's size before its tables were cut: the cut tables are recorded again)
  const bool record = c->step_graphs == 2 ||
                      (c->step_graphs == 1 && int64_t(B.h_t0run.size()) >= Batch::kGraphMinTiles && ++B.windows > 2);
  if (record && !c->profiling && !B.graphs_failed && !B.tuning) {
    // (each step form recorded on first use; a failed recording leaves this batch to launch kernel
    // by kernel)
    auto exec_for = [&](bool ln) -> hipGraphExec_t {
      hipGraphExec_t& ex = B.step_graph[(ln ? 8 : 0) + (mixed ? 4 : 0) + (p32 ? 2 : 0) + (p64 ? 1 : 0)];
      if (!ex && !B.graphs_failed) {
        hipGraph_t gr = nullptr;
        if (hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal) == hipSuccess) {
          if (mixed) mixed_step(B, s, degree, p32, p64, ln);
          else fp64_step(B, s, degree);
          if (hipStreamEndCapture(s, &gr) != hipSuccess) gr = nullptr;
        }
        if (!gr || hipGraphInstantiate(&ex, gr, nullptr, nullptr, 0) != hipSuccess) ex = nullptr;
        if (gr) (void)hipGraphDestroy(gr);
        if (!ex) {
          B.graphs_failed = true;
          (void)hipGetLastError();
        }
      }
      return ex;
    };
    const hipGraphExec_t full = exec_for(false);
    const hipGraphExec_t thin = (mixed && w > 1) ? exec_for(true) : full;
    if (full && thin) {
      for (int k = 0; k < w; ++k) (void)hipGraphLaunch(lean_at(k) ? thin : full, s);
      return;
    }
  }
  for (int k = 0; k < w; ++k) {
    if (mixed) mixed_step(B, s, degree, p32, p64, lean_at(k));
    else fp64_step(B, s, degree);
  }
}

// The tiled launches' tables cut to the graphs still running (filter), or made whole again. In the
// tail of a batch of uneven graphs most graphs have finished, and a launch over every tile pays for
// their workgroups (which only exit) - about 10 us a level-0 launch of the 4096-piece batch. The
// tables are rebuilt from their host copies in the same order, each graph's tiles as they were
// (their slots too), so no graph's arithmetic changes. Called after a synchronisation (no launch
// uses the old tables); the recorded step graphs, which hold them, are dropped.
static int compact_tiles(Batch& B, hipStream_t s, bool filter) {
  svxg_ctx* c = B.probs[0]->ctx;
  Arena* A = c->arena.get();
  std::vector<char> keep(size_t(B.G), 1);
  if (filter)
    for (int g = 0; g < B.G; ++g) keep[size_t(g)] = B.hstate[size_t(g)].state == ST_RUNNING ? 1 : 0;
  std::vector<std::vector<Tile>> hold;  // (the cut tables, alive until the flush)
  hold.reserve(2 + 8 * B.lvl.size());
  Stager up(c, s);
  auto put = [&](DevBuf& d, int& n, const std::vector<Tile>& h) -> int {
    hold.emplace_back();
    std::vector<Tile>& o = hold.back();
    for (const Tile& t : h)
      if (keep[size_t(t.g)]) o.push_back(t);
    n = int(o.size());
    return up.add(d, A, o);
  };
  SVXG_TRY(put(B.t0run, B.nt0run, B.h_t0run));
  SVXG_TRY(put(B.tj, B.ntj, B.h_tj));
  for (Batch::Lv& V : B.lvl) {
    for (int v = 0; v < 2; ++v) {
      SVXG_TRY(put(V.ts[v], V.nts[v], V.hts[v]));
      SVXG_TRY(put(V.tc[v], V.ntc[v], V.htc[v]));
      SVXG_TRY(put(V.tstall[v], V.ntstall[v], V.htstall[v]));
    }
    SVXG_TRY(put(V.tp, V.ntp, V.htp));
    SVXG_TRY(put(V.dense, V.ndense, V.hdense));
  }
  SVXG_TRY(up.flush());
  for (hipGraphExec_t& g : B.step_graph)
    if (g) {
      (void)hipGraphExecDestroy(g);
      g = nullptr;
    }
  B.compacted = filter;
  return SVXG_OK;
}

// at a look (states just read): the tables cut whenever at most half of their level-0 tiles belong
// to running graphs (while at least kCompactMinTiles remain: a finished graph's workgroup costs a
// launch about 1 ns, a cut about 0.1-0.2 ms; not while launch shapes are tuned)
static constexpr int kCompactMinTiles = 512;
static int maybe_compact(Batch& B, hipStream_t s) {
  if (B.tuning || !B.tune_pending.empty() || B.nt0run < kCompactMinTiles) return SVXG_OK;
  int64_t live = 0;
  for (const Tile& t : B.h_t0run) live += B.hstate[size_t(t.g)].state == ST_RUNNING ? 1 : 0;
  if (2 * live > B.nt0run) return SVXG_OK;
  return compact_tiles(B, s, true);
}

// the FP64 PCG over a batch (after run_begin)
static int pcg_fp64_iterate(Batch& B, hipStream_t s) {
  PollPlan plan;
  for (;;) {
    SVXG_TRY(read_states(B, s));
    if (!any_running(B, false)) break;
    SVXG_TRY(maybe_compact(B, s));
    run_window(B, s, plan.next(B, false), false, false, false);
  }
  SVXG_HIP(hipGetLastError());
  return SVXG_OK;
}

// FP64 iterative refinement around an FP32 PCG (FP32 K products and V-cycle, FP64 dots): each
// outer step recomputes r = b - K x in FP64, solves K d = r to max(eta |r|, target / 2) in FP32 and
// adds d to x.
//
// One step of every graph's mixed-precision solve, whatever its phase: a refinement step's FP64
// residual and check (PH_OUTER), an inner solve's start (PH_START) or iteration (PH_INNER), the
// correction x += d (PH_APPLY). The launches come in an order in which a graph moves on at most one
// phase per kernel and never meets a launch of its new phase in the same step before that phase's
// own sequence begins. p32 / p64: launch the FP32- / FP64-product kernels. lean: the inner
// iteration's launches only (no refinement residual, no inner start): for windows that begin with
// every running graph iterating an inner solve; a graph whose inner solve ends within the window
// applies its correction and then waits, untouched, for the next window's full steps.
static void mixed_step(Batch& B, hipStream_t s, int degree, bool p32, bool p64, bool lean) {
  const GraphDev* G = B.dG.as<GraphDev>();
  const Tile* t0 = B.t0run.as<Tile>();
  const int32_t* gt = B.gtile.as<int32_t>();
  double* part = B.partial.as<double>();
  if (!lean) {
    launch_outer_residual(s, G, t0, B.nt0run, part);
    launch_finish(s, FinishKind::Outer, G, gt, B.G, SERVE_REFINE, part);
    launch_inner_init(s, G, t0, B.nt0run);
  }
  {
    ProfScope ps(B, p32 ? 1 : -1, s, SERVE_P32 | SERVE_IT);
    if (p32) launch_pcg_spmv<float>(s, G, t0, B.nt0run, SERVE_P32 | SERVE_IT, part, ps.pa(), ps.t0(), ps.t1());
  }
  if (p64) launch_pcg_spmv<double>(s, G, t0, B.nt0run, SERVE_P64 | SERVE_IT, part);
  launch_finish(s, FinishKind::Alpha, G, gt, B.G, SERVE_ITER, part);
  if (p32) launch_axpy_rr<float>(s, G, t0, B.nt0run, SERVE_P32 | SERVE_IT, part);
  if (p64) launch_axpy_rr<double>(s, G, t0, B.nt0run, SERVE_P64 | SERVE_IT, part);
  launch_finish(s, FinishKind::Rn, G, gt, B.G, SERVE_ITER, part);
  vcycle<float>(B, s, SERVE_INNER, degree, true);
  if (p32 && !B.fuse_rz) launch_dot_rz<float>(s, G, t0, B.nt0run, SERVE_P32 | SERVE_IT, part);
  if (p64 && !B.fuse_rz) launch_dot_rz<double>(s, G, t0, B.nt0run, SERVE_P64 | SERVE_IT, part);
  launch_finish(s, FinishKind::Beta, G, gt, B.G, SERVE_ITER, part);
  if (p32) launch_update_p<float>(s, G, t0, B.nt0run, SERVE_P32 | SERVE_IT);
  if (p64) launch_update_p<double>(s, G, t0, B.nt0run, SERVE_P64 | SERVE_IT);
  if (!lean) {
    if (p32) launch_start_p<float>(s, G, t0, B.nt0run, SERVE_P32 | SERVE_ST, part);
    if (p64) launch_start_p<double>(s, G, t0, B.nt0run, SERVE_P64 | SERVE_ST, part);
    launch_finish(s, FinishKind::Start, G, gt, B.G, SERVE_START, part);
  }
  // (FP64-product inner solves update x themselves and close in their finishes)
  if (p32) launch_inner_apply(s, G, t0, B.nt0run);
  if (p32) launch_inner_close(s, G, B.G);
}

// which product precisions the next window launches (states just read): FP64 products once a
// running graph takes them or waits for them (waiting graphs are released first), FP32 products
// while a running graph has not switched
static void mixed_window(Batch& B, hipStream_t s, bool& p32, bool& p64) {
  p32 = p64 = false;
  bool wait = false;
  for (const PcgState& st : B.hstate) {
    if (st.state != ST_RUNNING) continue;
    (st.inner64 ? p64 : p32) = true;
    wait = wait || st.phase == PH_WAIT64;
  }
  if (wait) launch_release_wait64(s, B.dG.as<GraphDev>(), B.G);
}

// the mixed-precision solve (after run_begin): FP64 iterative refinement around FP32 inner PCG
// solves, each graph stepping through its own refinements
static int pcg_mixed_iterate(Batch& B, hipStream_t s) {
  PollPlan plan;
  bool p32 = true, p64 = false;
  for (;;) {
    SVXG_TRY(read_states(B, s));
    if (!any_running(B, false)) break;
    SVXG_TRY(maybe_compact(B, s));
    const int w = plan.next(B, true);
    mixed_window(B, s, p32, p64);
    bool lean = true;  // (every running graph iterates an inner solve)
    for (const PcgState& st : B.hstate)
      if (st.state == ST_RUNNING && st.phase != PH_INNER) lean = false;
    run_window(B, s, w, true, p32, p64, lean);
  }
  SVXG_HIP(hipGetLastError());
  return SVXG_OK;
}

// the head of StressProblem::iterate on a resumed run: |r| against the (new) target
static int recheck_residual(Batch& B, hipStream_t s) {
  launch_dot_rr<double>(s, B.dG.as<GraphDev>(), B.t0run.as<Tile>(), B.nt0run, 0, B.partial.as<double>());
  launch_finish(s, FinishKind::Rn, B.dG.as<GraphDev>(), B.gtile.as<int32_t>(), B.G, 0, B.partial.as<double>());
  launch_resident_recheck(s, B.dG.as<GraphDev>(), B.gtile.as<int32_t>(), B.res.as<int32_t>(), B.nres,
                          B.partial.as<double>());
  SVXG_HIP(hipGetLastError());
  return SVXG_OK;
}

}  // namespace svxg

// the rest of the host side (plan, setup, ABI) is in host_setup.hip and host_abi.hip
#include "host_setup.inc"
#include "host_abi.inc"
#include "host_serve.inc"
#include "host_judge.inc"
#include "host_lattice.inc"
#include "host_extract.inc"

namespace svxg {
// (a kept batch-of-one holds device pointers: rebuilt when K's arrays change)
static void drop_self_batch_ptrs(svxg_problem* P) { drop_self_batch(P); }
}  // namespace svxg
