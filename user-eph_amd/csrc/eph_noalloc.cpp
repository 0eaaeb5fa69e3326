// eph_noalloc.cpp -- eph_grid_insert_energy_noalloc: the ordered energy
// deposit of eph_grid_insert_energy for callers that may not allocate device
// memory (code under HIP graph capture, tight serving loops, harnesses that
// forbid hipMalloc).  Its scratch lives in one file-scope device array per
// GPU; the points come from the caller's device memory.
#include <mutex>

#include "eph_context.hpp"

using namespace eph;

namespace {

// The deposit's scratch: per point its cell, value and slot in the per-cell
// lists; per cell the counts, offsets, fill counters and work queues.  Every
// GPU the library runs on holds one copy until the process ends.
constexpr size_t kScratchBytes = size_t(20) << 20;
__device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void* g_scratch_ptr[kMaxDevices] = {};  // device address of `scratch`, per GPU

size_t a256(size_t b) { return (b + 255) & ~size_t(255); }

// Bytes of scratch the deposit of n points into G cells needs.
size_t need_bytes(size_t n, size_t G) {
  return a256(n * sizeof(double)) + 2 * a256(n * sizeof(int)) + 3 * a256((G + 1) * sizeof(int)) +
         a256((2 + 2 * G) * sizeof(int)) + a256((G / 4096 + 2) * sizeof(int));
}

// Address of this device's copy of `scratch`, looked up once per device with
// that device current.  The array's host-side name is not a device address.
int scratch_of(int device, unsigned char** out, std::string* err) {
  std::lock_guard<std::mutex> lock(g_mu);
  if (device < 0 || device >= kMaxDevices) {
    *err = "device index out of range for the scratch table";
    return EPH_ERR_ARGS;
  }
  if (!g_scratch_ptr[device]) {
    int cur = 0;
    hipError_t e = hipGetDevice(&cur);
    if (e == hipSuccess && cur != device) e = hipSetDevice(device);
    void* p = nullptr;
    if (e == hipSuccess) e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (cur != device) (void)hipSetDevice(cur);
    if (e != hipSuccess || !p) {
      *err = std::string("HIP error ") + hipGetErrorString(e) + " in hipGetSymbolAddress(scratch)";
      return EPH_ERR_HIP;
    }
    g_scratch_ptr[device] = p;
  }
  *out = static_cast<unsigned char*>(g_scratch_ptr[device]);
  return EPH_OK;
}

}  // namespace

extern "C" {

long long eph_noalloc_scratch_bytes(void) { return static_cast<long long>(kScratchBytes); }

int eph_grid_insert_energy_noalloc(eph_grid* g, int n, const double* x, const double* y,
                                   const double* z, const double* E) {
  if (!g) return EPH_ERR_ARGS;
  if (n < 0 || (n > 0 && (!x || !y || !z || !E))) return g->fail(EPH_ERR_ARGS, "bad point arrays");
  if (n == 0) return EPH_OK;
  const size_t G = g->h.ntotal();
  const size_t need = need_bytes(static_cast<size_t>(n), G);
  if (need > kScratchBytes)
    return g->fail(EPH_ERR_ARGS, "deposit of " + std::to_string(n) + " points into " +
                                     std::to_string(G) + " cells needs " + std::to_string(need) +
                                     " bytes of scratch; the no-allocation entry point has " +
                                     std::to_string(kScratchBytes) + " (16 n + 20 G + about 2 KB)");
  unsigned char* base = nullptr;
  std::string err;
  if (int rc = scratch_of(g->device, &base, &err)) return g->fail(rc, err);
  // carve the scratch (device addresses only)
  size_t off = 0;
  auto take = [&](size_t bytes) {
    unsigned char* p = base + off;
    off += a256(bytes);
    return p;
  };
  DepositView d;
  d.v1 = reinterpret_cast<double*>(take(n * sizeof(double)));
  d.cell = reinterpret_cast<int*>(take(n * sizeof(int)));
  d.items = reinterpret_cast<int*>(take(n * sizeof(int)));
  d.count = reinterpret_cast<int*>(take((G + 1) * sizeof(int)));
  d.start = reinterpret_cast<int*>(take((G + 1) * sizeof(int)));
  d.fill = reinterpret_cast<int*>(take((G + 1) * sizeof(int)));
  int* qbuf = reinterpret_cast<int*>(take((2 + 2 * G) * sizeof(int)));
  d.scan_tmp = reinterpret_cast<int*>(take((G / 4096 + 2) * sizeof(int)));
  d.v2 = nullptr;
  d.partial = nullptr;
  hipStream_t s = g->stream;
  // the counts start at zero on every call (the previous call left its own)
  EPH_HIP(g, launch_fill_int(d.count, static_cast<int>(G) + 1, 0, s));
  EPH_HIP(g, launch_deposit_points(n, x, y, z, E, g->h.dt, g->v, d, s));
  EPH_HIP(g, launch_deposit_cells(g->v, d, n, qbuf, s));
  return EPH_OK;
}

}  // extern "C"
