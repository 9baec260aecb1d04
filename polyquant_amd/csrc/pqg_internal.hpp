// Internal declarations shared by the pqgpu translation units (host + device).
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdint>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/pqgpu.h"

namespace pqg {

// ---------------------------------------------------------------------------------------
// errors
// ---------------------------------------------------------------------------------------
void set_error(const std::string& msg);

struct Error {
  pqg_status code;
  std::string msg;
};

#define PQG_HIP(expr)                                                                    \
  do {                                                                                   \
    hipError_t _e = (expr);                                                              \
    if (_e != hipSuccess)                                                                \
      throw ::pqg::Error{PQG_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(_e)}; \
  } while (0)

#define PQG_REQUIRE(cond, code, msg)                     \
  do {                                                   \
    if (!(cond)) throw ::pqg::Error{(code), (msg)};      \
  } while (0)

// Host -> device copy ordered on `st` and complete on return.  The library keeps every copy on
// its own non-blocking streams: a blocking copy on the legacy stream would depend on (and so
// break) a stream capture running in another thread.
inline void upload(hipStream_t st, void* dst, const void* src, size_t bytes) {
  PQG_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, st));
  PQG_HIP(hipStreamSynchronize(st));
}

// Runs f, converting exceptions into a status + thread-local message.
template <class F>
pqg_status guard(F&& f) {
  try {
    f();
    return PQG_OK;
  } catch (const Error& e) {
    set_error(e.msg);
    return e.code;
  } catch (const std::bad_alloc&) {
    set_error("host allocation failed");
    return PQG_ERR_OOM;
  } catch (const std::exception& e) {
    set_error(e.what());
    return PQG_ERR_ARG;
  }
}

// ---------------------------------------------------------------------------------------
// device memory manager: first-fit arena inside an HBM budget
// ---------------------------------------------------------------------------------------
// Serialises the HIP calls that may synchronise the device implicitly (hipMalloc, hipFree,
// hipHostMalloc, hipHostFree) with stream captures: a capture in one thread is invalidated by
// such a call in another.  Recursive: a capture may reach the arena.
std::recursive_mutex& device_op_mutex();

class Arena {
 public:
  Arena(int device, uint64_t budget) : device_(device), budget_(budget) {}
  ~Arena();
  void* alloc(uint64_t bytes);  // throws PQG_ERR_OOM past the budget
  void free(void* p);
  bool fits(uint64_t bytes) const;  // would an allocation of this size fit right now?
  uint64_t used() const { return used_; }
  uint64_t peak() const { return peak_; }
  uint64_t budget() const { return budget_; }

 private:
  struct Chunk {
    char* base;
    uint64_t size;
    std::map<uint64_t, uint64_t> free_;  // offset -> size
    bool dedicated = false;              // one large block: returned to the device on free
    int live = 0;                        // blocks handed out from this chunk
  };
  bool release_empty_locked();  // hipFree every chunk without live blocks; true if any
  int device_;
  uint64_t budget_;
  uint64_t reserved_ = 0;  // bytes obtained from hipMalloc
  uint64_t used_ = 0, peak_ = 0;
  std::vector<Chunk> chunks_;
  std::map<char*, std::pair<int, uint64_t>> live_;  // ptr -> (chunk, size)
  mutable std::mutex mu_;
};

}  // namespace pqg

struct pqg_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::unique_ptr<pqg::Arena> arena;
  int num_cu = 0;
  // The caller's handle plus every live CI handle hold a reference; the context (and its
  // arena) is torn down when the last one is released.
  std::atomic<int> refs{1};
  // kernel timing (pqg_ctx_profile)
  std::atomic<bool> profiling{false};
  struct ProfRec {
    std::string name;
    hipEvent_t start, stop;
    int64_t units;
  };
  std::mutex prof_mu;
  std::vector<ProfRec> prof;
};

namespace pqg {
void ctx_retain(pqg_ctx* c);
void ctx_release(pqg_ctx* c);

// Fixed device workspace of the allocation-free entry point: while one is installed on the
// calling thread, grow() hands out views into it instead of allocating, and ProfScope records
// no events.  base is a device address; used is the bump offset.
struct FixedScratch {
  char* base = nullptr;
  size_t size = 0;
  size_t used = 0;
  // next `bytes` (256-byte aligned) of the workspace; PQG_ERR_ARG when it is exhausted
  char* take(size_t bytes);
  size_t left() const { return size > used ? size - used : 0; }
};
FixedScratch*& fixed_scratch();  // per host thread; nullptr when none is installed

// Brackets one kernel launch with HIP events on `st` while the context is profiling.
class ProfScope {
 public:
  ProfScope(pqg_ctx* c, hipStream_t st, const char* name, int64_t units) : c_(c), st_(st) {
    if (!c_->profiling.load() || fixed_scratch()) return;
    rec_.name = name;
    rec_.units = units;
    if (hipEventCreate(&rec_.start) != hipSuccess || hipEventCreate(&rec_.stop) != hipSuccess) return;
    active_ = hipEventRecord(rec_.start, st_) == hipSuccess;
  }
  ~ProfScope() {
    if (!active_) return;
    if (hipEventRecord(rec_.stop, st_) != hipSuccess) return;
    std::lock_guard<std::mutex> lk(c_->prof_mu);
    c_->prof.push_back(rec_);
  }

 private:
  pqg_ctx* c_;
  hipStream_t st_;
  pqg_ctx::ProfRec rec_{};
  bool active_ = false;
};
}  // namespace pqg

namespace pqg {

// RAII device buffer from the context arena, or (ctx == nullptr) a view into memory it does
// not own, such as the fixed workspace
template <class T>
struct DevBuf {
  pqg_ctx* ctx = nullptr;
  T* p = nullptr;
  size_t n = 0;
  DevBuf() = default;
  DevBuf(pqg_ctx* c, size_t count) { reset(c, count); }
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : ctx(o.ctx), p(o.p), n(o.n) { o.ctx = nullptr; o.p = nullptr; o.n = 0; }
  DevBuf& operator=(DevBuf&& o) noexcept {
    release();
    ctx = o.ctx; p = o.p; n = o.n;
    o.ctx = nullptr; o.p = nullptr; o.n = 0;
    return *this;
  }
  void reset(pqg_ctx* c, size_t count) {
    release();
    ctx = c;
    n = count;
    p = count ? static_cast<T*>(c->arena->alloc(count * sizeof(T))) : nullptr;
  }
  // view of `count` elements of the installed fixed workspace (nothing is allocated)
  void view(FixedScratch* fs, size_t count) {
    release();
    p = reinterpret_cast<T*>(fs->take(count * sizeof(T)));
    n = count;
  }
  void release() {
    if (p && ctx) ctx->arena->free(p);
    ctx = nullptr;
    p = nullptr;
    n = 0;
  }
  ~DevBuf() { release(); }
  T* get() const { return p; }
  size_t bytes() const { return n * sizeof(T); }
};

// Grows a scratch buffer used by work queued on `st`.  The old block goes back to the shared
// arena only after that work has finished: another stream (a serving slot) may be handed the
// memory at once.  With a fixed workspace installed, the buffer becomes a view into it.
template <class T>
void grow(DevBuf<T>& b, pqg_ctx* c, size_t count, hipStream_t st) {
  if (FixedScratch* fs = fixed_scratch()) {
    const bool owned = b.p && b.ctx;
    PQG_REQUIRE(!owned, PQG_ERR_ARG, "an owned buffer reached the allocation-free path");
    if (b.p && b.n >= count) return;
    b.view(fs, count);
    return;
  }
  if (b.n >= count) return;
  if (b.p) PQG_HIP(hipStreamSynchronize(st));
  b.reset(c, count);
}

}  // namespace pqg
