// Lanczos pass one without any device allocation, for callers that cannot call
// hipMalloc (code under stream capture, tight serving loops, harnesses that
// forbid it).
//
// Every buffer the run needs beyond the caller's start vector (the three Krylov
// vectors, the device-resident alpha / beta / norm, the partial sums and the
// flags) is carved from one file-scope device array, `scratch`, of
// kScratchBytes.  Each GPU the library runs on holds its own copy of the array
// until the process ends.  Host code never passes the array's name to a kernel
// (in host code it is only a stand-in for the symbol): on the first call on a
// device, with that device current, hipGetSymbolAddress gives the device
// address, which is kept per device and used for every kernel argument and
// every clear.  The buffers are handed out by a bump allocator installed as the
// calling thread's DevBuf source for the duration of the call, so the same
// kernels and launch code as mb_lanczos_pass_one run on them.
//
// Calls on one device share the array.  They are serialised by a per-device
// host mutex held for the whole call (which ends with a stream synchronisation),
// so two host threads cannot interleave on it; work another caller enqueues on
// the device without going through this entry point does not touch it.  What
// must start at zero (the previous Krylov vector and the step / breakdown
// flags) is cleared by a fill kernel at the start of every call; everything
// else is written before it is read.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
#include <string>

#include "krylov.hpp"
#include "mb_internal.hpp"

namespace mb {

namespace {

constexpr size_t kScratchBytes = size_t(48) << 20;  // 48 MiB per GPU
constexpr size_t kAlign = 256;
constexpr int kMaxDevices = 64;

__device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_mu[kMaxDevices];
void *g_ptr[kMaxDevices];  // device address of `scratch` on each device (under g_mu)

size_t align_up(size_t n) { return (n + kAlign - 1) / kAlign * kAlign; }

// Bump allocator over the scratch array; frees are no-ops (the whole array is
// reused by the next call).
class ScratchArena final : public DevAlloc {
 public:
  ScratchArena(void *base, size_t cap) : base_(static_cast<char *>(base)), cap_(cap) {}
  void *alloc(size_t n) override {
    const size_t off = align_up(used_);
    if (off > cap_ || n > cap_ - off) return nullptr;
    used_ = off + n;
    return base_ + off;
  }
  void free(void *) override {}

 private:
  char *base_;
  size_t cap_;
  size_t used_ = 0;
};

// Restores the thread's previous DevBuf source on every exit path.
struct AllocScope {
  DevAlloc *prev;
  explicit AllocScope(DevAlloc *a) : prev(thread_dev_alloc()) { thread_dev_alloc() = a; }
  ~AllocScope() { thread_dev_alloc() = prev; }
};

}  // namespace

size_t lanczos_noalloc_scratch_bytes() { return kScratchBytes; }

// Bytes of scratch one run takes: the three complex128 vectors and the scalars
// of LanczosWork, each rounded up to the allocator's alignment.
size_t lanczos_noalloc_bytes(int64_t D, int k, int64_t nparts) {
  const size_t vec = align_up((size_t)D * 2 * sizeof(double));
  const size_t scal = align_up(sizeof(double) * (2 * (size_t)std::max(k, 1) + 2 +
                                                 (size_t)std::max<int64_t>(kRedBlocks, nparts)) +
                               16);
  return 3 * vec + scal + kAlign;
}

LanczosResult lanczos_pass_one_noalloc(const HvOp &op, const void *b, int k, hipStream_t s) {
  int dev = -1;
  MB_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices)
    throw std::invalid_argument("device ordinal out of range for the scratch table");
  const int64_t D = op.dim();
  const size_t need = lanczos_noalloc_bytes(D, k, op.partials());
  if (need > kScratchBytes)
    throw std::invalid_argument("the no-allocation Lanczos needs " + std::to_string(need) +
                                " bytes of scratch for D = " + std::to_string(D) + ", k = " +
                                std::to_string(k) + "; its static scratch holds " +
                                std::to_string(kScratchBytes) + " bytes");
  std::lock_guard<std::mutex> lock(g_mu[dev]);
  if (!g_ptr[dev]) {
    void *p = nullptr;
    MB_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) throw DeviceError("hipGetSymbolAddress returned no address for the scratch array");
    g_ptr[dev] = p;
  }
  ScratchArena arena(g_ptr[dev], kScratchBytes);
  AllocScope scope(&arena);
  LanczosWork work(D, k, op.partials());
  return lanczos_pass_one(op, b, k, s, &work);
}

}  // namespace mb
