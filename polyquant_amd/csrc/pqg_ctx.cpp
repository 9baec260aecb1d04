// Context, error reporting and the device-memory arena.
//
// The arena is the device memory manager: one context owns an HBM budget, obtains memory
// from hipMalloc in large chunks (at least 256 MiB, never past the budget) and hands out
// 256-byte-aligned blocks first-fit, merging neighbours on free.  CI handles, Davidson
// subspaces and serving slots all allocate from it, so admission control can ask
// `fits()` before taking a request.
#include <algorithm>
#include <cstring>

#include "pqg_internal.hpp"

namespace pqg {

static thread_local std::string t_err;
void set_error(const std::string& msg) { t_err = msg; }

std::recursive_mutex& device_op_mutex() {
  static std::recursive_mutex m;
  return m;
}

static thread_local FixedScratch* t_fixed = nullptr;
FixedScratch*& fixed_scratch() { return t_fixed; }

char* FixedScratch::take(size_t bytes) {
  const size_t off = (used + 255) / 256 * 256;
  PQG_REQUIRE(off <= size && bytes <= size - off, PQG_ERR_ARG,
              "the fixed scratch of the allocation-free entry point is too small for this call (" +
                  std::to_string(size >> 20) + " MiB)");
  used = off + bytes;
  return base + off;
}

static constexpr uint64_t kAlign = 256;
static constexpr uint64_t kMinChunk = 256ull << 20;
// Blocks at least this large get a chunk of their own that goes back to the device when the
// block is freed, so long-lived small blocks cannot fragment the budget for large solves.
static constexpr uint64_t kDedicated = 1ull << 30;

static uint64_t round_up(uint64_t x, uint64_t a) { return (x + a - 1) / a * a; }

Arena::~Arena() {
  std::lock_guard<std::recursive_mutex> dl(device_op_mutex());
  for (auto& c : chunks_)
    if (c.base) (void)hipFree(c.base);
}

bool Arena::fits(uint64_t bytes) const {
  std::lock_guard<std::mutex> lk(mu_);
  bytes = round_up(std::max<uint64_t>(bytes, 1), kAlign);
  for (const auto& c : chunks_)
    if (c.base && !c.dedicated && bytes < kDedicated)
      for (const auto& [off, sz] : c.free_)
        if (sz >= bytes) return true;
  return reserved_ + std::max(bytes, std::min(kMinChunk, budget_ - std::min(budget_, reserved_))) <= budget_ ||
         reserved_ + bytes <= budget_;
}

bool Arena::release_empty_locked() {
  bool any = false;
  for (auto& c : chunks_) {
    if (!c.base || c.live > 0) continue;
    (void)hipFree(c.base);
    reserved_ -= c.size;
    c.base = nullptr;
    c.free_.clear();
    any = true;
  }
  return any;
}

void* Arena::alloc(uint64_t bytes) {
  std::lock_guard<std::recursive_mutex> dl(device_op_mutex());  // always before mu_
  std::lock_guard<std::mutex> lk(mu_);
  bytes = round_up(std::max<uint64_t>(bytes, 1), kAlign);
  if (bytes < kDedicated) {
    for (size_t ci = 0; ci < chunks_.size(); ci++) {
      auto& c = chunks_[ci];
      if (!c.base || c.dedicated) continue;
      for (auto it = c.free_.begin(); it != c.free_.end(); ++it) {
        if (it->second >= bytes) {
          uint64_t off = it->first, sz = it->second;
          c.free_.erase(it);
          if (sz > bytes) c.free_[off + bytes] = sz - bytes;
          char* p = c.base + off;
          live_[p] = {(int)ci, bytes};
          c.live++;
          used_ += bytes;
          peak_ = std::max(peak_, used_);
          return p;
        }
      }
    }
  }
  // new chunk (empty chunks are given back first when the budget would be exceeded)
  if (reserved_ + bytes > budget_) release_empty_locked();
  if (reserved_ + bytes > budget_)
    throw Error{PQG_ERR_OOM, "HBM budget exceeded: need " + std::to_string(bytes) + " B, " +
                                 std::to_string(budget_ - std::min(budget_, used_)) + " B free of budget"};
  const bool dedicated = bytes >= kDedicated;
  uint64_t csize = dedicated ? bytes : std::max(bytes, kMinChunk);
  csize = std::min(csize, budget_ - reserved_);
  csize = std::max(csize, bytes);
  char* base = nullptr;
  hipError_t e = hipMalloc(reinterpret_cast<void**>(&base), csize);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    if (release_empty_locked() || csize > bytes) {  // retry with the exact size
      csize = bytes;
      e = hipMalloc(reinterpret_cast<void**>(&base), csize);
    }
    if (e != hipSuccess) {
      (void)hipGetLastError();
      throw Error{PQG_ERR_OOM, std::string("hipMalloc failed: ") + hipGetErrorString(e)};
    }
  }
  reserved_ += csize;
  Chunk c{base, csize, {}, dedicated, 1};
  if (csize > bytes) c.free_[bytes] = csize - bytes;
  int idx = -1;
  for (size_t ci = 0; ci < chunks_.size(); ci++)
    if (!chunks_[ci].base) {  // reuse the slot of a released chunk
      idx = (int)ci;
      break;
    }
  if (idx < 0) {
    chunks_.push_back(std::move(c));
    idx = (int)chunks_.size() - 1;
  } else {
    chunks_[idx] = std::move(c);
  }
  live_[base] = {idx, bytes};
  used_ += bytes;
  peak_ = std::max(peak_, used_);
  return base;
}

void Arena::free(void* vp) {
  if (!vp) return;
  std::lock_guard<std::recursive_mutex> dl(device_op_mutex());  // always before mu_
  std::lock_guard<std::mutex> lk(mu_);
  char* p = static_cast<char*>(vp);
  auto it = live_.find(p);
  if (it == live_.end()) return;
  auto [ci, sz] = it->second;
  live_.erase(it);
  used_ -= sz;
  auto& c = chunks_[ci];
  c.live--;
  if (c.dedicated) {  // the owner has finished its work on this block (stream synchronised)
    (void)hipFree(c.base);
    reserved_ -= c.size;
    c.base = nullptr;
    c.free_.clear();
    return;
  }
  uint64_t off = p - c.base;
  auto nx = c.free_.lower_bound(off);
  if (nx != c.free_.end() && nx->first == off + sz) {  // merge right
    sz += nx->second;
    nx = c.free_.erase(nx);
  }
  if (nx != c.free_.begin()) {  // merge left
    auto pv = std::prev(nx);
    if (pv->first + pv->second == off) {
      pv->second += sz;
      return;
    }
  }
  c.free_[off] = sz;
}

void ctx_retain(pqg_ctx* c) { c->refs.fetch_add(1); }

void ctx_release(pqg_ctx* c) {
  if (c->refs.fetch_sub(1) != 1) return;
  (void)hipSetDevice(c->device);
  for (auto& r : c->prof) {
    (void)hipEventDestroy(r.start);
    (void)hipEventDestroy(r.stop);
  }
  (void)hipStreamSynchronize(c->stream);
  (void)hipStreamDestroy(c->stream);
  c->arena.reset();
  delete c;
}

}  // namespace pqg

extern "C" {

const char* pqg_last_error(void) { return pqg::t_err.c_str(); }

pqg_status pqg_ctx_create(int device, uint64_t hbm_budget, pqg_ctx** out) {
  return pqg::guard([&] {
    PQG_REQUIRE(out != nullptr, PQG_ERR_ARG, "out is NULL");
    int n = 0;
    PQG_HIP(hipGetDeviceCount(&n));
    PQG_REQUIRE(device >= 0 && device < n, PQG_ERR_ARG, "no such device");
    PQG_HIP(hipSetDevice(device));
    auto* c = new pqg_ctx();
    c->device = device;
    size_t free_b = 0, total_b = 0;
    PQG_HIP(hipMemGetInfo(&free_b, &total_b));
    if (hbm_budget == 0 || hbm_budget > free_b) hbm_budget = (uint64_t)(free_b * 0.95);
    c->arena = std::make_unique<pqg::Arena>(device, hbm_budget);
    PQG_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    hipDeviceProp_t prop;
    PQG_HIP(hipGetDeviceProperties(&prop, device));
    c->num_cu = prop.multiProcessorCount;
    *out = c;
  });
}

pqg_status pqg_ctx_destroy(pqg_ctx* ctx) {
  return pqg::guard([&] {
    if (!ctx) return;
    pqg::ctx_release(ctx);
  });
}

pqg_status pqg_ctx_memory(pqg_ctx* ctx, uint64_t* used, uint64_t* budget, uint64_t* peak) {
  return pqg::guard([&] {
    PQG_REQUIRE(ctx, PQG_ERR_ARG, "ctx is NULL");
    if (used) *used = ctx->arena->used();
    if (budget) *budget = ctx->arena->budget();
    if (peak) *peak = ctx->arena->peak();
  });
}

pqg_status pqg_ctx_sync(pqg_ctx* ctx) {
  return pqg::guard([&] {
    PQG_REQUIRE(ctx, PQG_ERR_ARG, "ctx is NULL");
    PQG_HIP(hipStreamSynchronize(ctx->stream));
  });
}

pqg_status pqg_ctx_profile(pqg_ctx* ctx, int enable) {
  return pqg::guard([&] {
    PQG_REQUIRE(ctx, PQG_ERR_ARG, "ctx is NULL");
    ctx->profiling.store(enable != 0);
  });
}

pqg_status pqg_ctx_profile_read(pqg_ctx* ctx, const char* name, int64_t* launches, double* total_ms,
                                int64_t* total_units) {
  return pqg::guard([&] {
    PQG_REQUIRE(ctx && name, PQG_ERR_ARG, "NULL argument");
    std::lock_guard<std::mutex> lk(ctx->prof_mu);
    int64_t n = 0, u = 0;
    double ms = 0.0;
    std::vector<pqg_ctx::ProfRec> keep;
    for (auto& r : ctx->prof) {
      if (r.name != name) {
        keep.push_back(r);
        continue;
      }
      PQG_HIP(hipEventSynchronize(r.stop));
      float t = 0.f;
      PQG_HIP(hipEventElapsedTime(&t, r.start, r.stop));
      ms += t;
      n++;
      u += r.units;
      (void)hipEventDestroy(r.start);
      (void)hipEventDestroy(r.stop);
    }
    ctx->prof.swap(keep);
    if (launches) *launches = n;
    if (total_ms) *total_ms = ms;
    if (total_units) *total_units = u;
  });
}

}  // extern "C"
