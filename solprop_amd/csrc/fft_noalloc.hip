// A batched FFT entry point that allocates nothing: the library's Stockham transform on
// device buffers the caller owns, its twiddle tables and the four-step transform's
// intermediate held in one file-scope device array.
//
// It is for callers that cannot allocate device memory (code under HIP graph capture,
// serving loops, harnesses that forbid hipMalloc): no hipMalloc, hipFree, stream or
// event is created by a call. The array is `scratch` below, kNoAllocPoints complex
// doubles (about 17 MiB), which every GPU the library runs on holds until the process
// ends. Every call on one GPU shares that GPU's copy: calls that overlap on one GPU (on
// different streams, or from different threads, before the earlier one completed) race.
// The tables are rebuilt (a host computation and one copy on the caller's stream, which
// the call then synchronises) whenever the length differs from the last call's on that
// GPU; a call at the last call's length enqueues its passes and returns.
#include <array>
#include <cstdio>
#include <mutex>

#include "../../include/solprop_mi355x.h"
#include "fft.hpp"

namespace sp {
namespace {

constexpr long kNoAllocMaxLength = 1L << 16;  // points a transform
constexpr long kNoAllocMaxPoints = 1L << 20;  // points of all rows together
// Tables first (the longest: n + 2 sqrt(n) at the four-step's largest length), then the
// four-step intermediate.
constexpr long kTablePoints = kNoAllocMaxLength + 2 * 256;
constexpr long kNoAllocPoints = kTablePoints + kNoAllocMaxPoints;

static __device__ __align__(256) cplx scratch[kNoAllocPoints];

constexpr int kMaxDevices = 64;

// Per device: the address of its copy of `scratch` (from hipGetSymbolAddress, with that
// device current) and the length whose tables it holds.
struct DeviceSlot {
  cplx* base = nullptr;
  long table_n = 0;
};

std::mutex g_mu;
std::array<DeviceSlot, kMaxDevices> g_slots;

void write_message(char* msg, size_t len, const char* text) {
  if (msg && len) std::snprintf(msg, len, "%s", text);
}

}  // namespace
}  // namespace sp

extern "C" int sp_fft_noalloc(const double* in, double* out, long n, long batch, int sign,
                              void* stream, char* msg, size_t msglen) {
  using namespace sp;
  try {
    SP_CHECK(in && out, "in and out must be device pointers");
    SP_CHECK(sign == -1 || sign == 1, "sign must be -1 or +1");
    SP_CHECK(n >= 4 && (n & (n - 1)) == 0, "n must be a power of two >= 4");
    SP_CHECK(batch >= 1, "batch must be positive");
    if (n > kNoAllocMaxLength || n * batch > kNoAllocMaxPoints) {
      char text[200];
      std::snprintf(text, sizeof(text),
                    "sp_fft_noalloc holds at most %ld points a row and %ld points in all "
                    "(n = %ld, batch = %ld)",
                    kNoAllocMaxLength, kNoAllocMaxPoints, n, batch);
      throw Error(kValue, text);
    }
    const hipStream_t s = static_cast<hipStream_t>(stream);
    int device = 0;
    SP_HIP(hipGetDevice(&device));
    SP_CHECK(device >= 0 && device < kMaxDevices, "device index out of range");
    FftPlan plan;
    cplx* base = nullptr;
    {
      std::lock_guard<std::mutex> lock(g_mu);
      DeviceSlot& slot = g_slots[static_cast<std::size_t>(device)];
      if (!slot.base) {
        void* ptr = nullptr;
        SP_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        slot.base = static_cast<cplx*>(ptr);
      }
      base = slot.base;
      plan.init_external(static_cast<std::size_t>(n), static_cast<std::size_t>(batch), base);
      SP_CHECK(static_cast<long>(FftPlan::table_points(static_cast<std::size_t>(n))) <=
                   kTablePoints,
               "the transform's tables exceed the reserved region");
      if (slot.table_n != n) {
        slot.table_n = 0;  // a failed upload leaves no stale claim
        plan.upload_tables(s);
        slot.table_n = n;
      }
      const std::size_t bytes = sizeof(cplx) * static_cast<std::size_t>(n * batch);
      cplx* data = reinterpret_cast<cplx*>(out);
      if (static_cast<const void*>(in) != static_cast<const void*>(out))
        SP_HIP(hipMemcpyAsync(data, in, bytes, hipMemcpyDeviceToDevice, s));
      plan.exec(data, base + kTablePoints, sign, s);
    }
    write_message(msg, msglen, "");
    return SP_OK;
  } catch (const Error& e) {
    write_message(msg, msglen, e.what());
    return e.code;
  } catch (const std::exception& e) {
    write_message(msg, msglen, e.what());
    return SP_BACKEND;
  }
}
