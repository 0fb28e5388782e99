// The allocation-free variant of reinit_sdf (litho_ls_reinit_sdf_noalloc):
// the chamfer row buffers and distance planes live in one file-scope device
// array instead of the context's grow-only workspace, so the call makes no
// device allocation.  Each GPU holds its own copy of the array (loaded with
// the library's code object) until the process ends; host code reaches it
// only through the per-device address that hipGetSymbolAddress returns.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <mutex>
#include <string>

#include "levelset.hpp"
#include "noalloc.hpp"
#include "plan.hpp"

namespace litho {

namespace {

// Every call on one GPU shares this array and finds what the last call left
// in it.  The chamfer kernels write each element before they read it, so
// nothing needs clearing between calls.
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mutex;
void* g_scratch_ptr[kMaxDevices] = {};

}  // namespace

void noalloc_init(int device) {
  if (device < 0 || device >= kMaxDevices) throw Error(1, "no-alloc scratch: device index out of range");
  std::lock_guard<std::mutex> lock(g_mutex);
  if (g_scratch_ptr[device] != nullptr) return;
  int cur = -1;
  LITHO_HIP(hipGetDevice(&cur));
  if (cur != device) throw Error(1, "no-alloc scratch: set-up must run with the device current");
  void* ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess || ptr == nullptr)
    throw Error(3, std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e));
  g_scratch_ptr[device] = ptr;
}

void ls_reinit_sdf_noalloc(Context& ctx, int batch, int rows, int cols, const float* phi, double dx, double dy,
                           float* sdf) {
  if (batch <= 0 || batch > 65535) throw Error(1, "level set: batch must be in [1, 65535]");
  if (rows < 3 || cols < 3) throw Error(1, "level set: tiles need at least 3 rows and 3 columns");
  if (cols > 4096) throw Error(1, "level set: at most 4096 columns per tile");
  if (phi == nullptr || sdf == nullptr) throw Error(1, "reinit_sdf: null buffer");
  const size_t need = chamfer_scratch_bytes(batch, rows, cols);
  if (need > kNoAllocScratchBytes)
    throw Error(1, "reinit_sdf_noalloc: " + std::to_string(batch) + " tiles of " + std::to_string(rows) + " x " +
                       std::to_string(cols) + " need " + std::to_string(need) + " bytes of scratch, above the " +
                       std::to_string(kNoAllocScratchBytes) + " bytes of the static buffer");
  void* base = nullptr;
  {
    std::lock_guard<std::mutex> lock(g_mutex);
    if (ctx.device >= 0 && ctx.device < kMaxDevices) base = g_scratch_ptr[ctx.device];
  }
  if (base == nullptr) throw Error(3, "reinit_sdf_noalloc: the scratch address of this device was not set up");
  const size_t plane = (size_t)rows * cols;
  uint32_t* rowbuf = static_cast<uint32_t*>(base);
  float* dist = reinterpret_cast<float*>(static_cast<char*>(base) + (size_t)batch * 2 * plane * sizeof(uint32_t));
  launch_reinit_sdf(phi, batch, rows, cols, dx, dy, 0.f, false, rowbuf, dist, nullptr, sdf, nullptr, ctx.stream);
  LITHO_HIP(hipGetLastError());
}

}  // namespace litho
