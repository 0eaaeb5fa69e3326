// eph_stage.cpp -- pinned, double-buffered host <-> device copies of the
// host-array hooks.
#include "eph_stage.hpp"

#include <algorithm>
#include <cstdlib>
#include <cstring>

namespace eph {

namespace {
// copies below this size stay on the calling thread
constexpr size_t kSplitMin = size_t(1) << 20;

void range_of(size_t n, int parts, int k, size_t* a, size_t* b) {
  *a = n * static_cast<size_t>(k) / static_cast<size_t>(parts);
  *b = n * static_cast<size_t>(k + 1) / static_cast<size_t>(parts);
}
}  // namespace

int HostStager::default_threads() {
  int t = 4;
  if (const char* e = std::getenv("EPH_HOST_THREADS")) {
    const int v = std::atoi(e);
    if (v >= 1) t = v;
  }
  const unsigned hw = std::thread::hardware_concurrency();
  if (hw > 0) t = std::min(t, static_cast<int>(hw));
  return std::max(1, std::min(t, 64));
}

hipError_t HostStager::init(int threads, size_t chunk_bytes) {
  if (ready()) return hipSuccess;
  chunk_ = std::max<size_t>(chunk_bytes, 4096);
  for (int b = 0; b < 2; ++b) {
    hipError_t e = hipHostMalloc(reinterpret_cast<void**>(&pin_[b]), chunk_, hipHostMallocDefault);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&ev_[b], hipEventDisableTiming);
    if (e != hipSuccess) {
      for (int c = 0; c < 2; ++c) {
        if (pin_[c]) (void)hipHostFree(pin_[c]);
        if (ev_[c]) (void)hipEventDestroy(ev_[c]);
        pin_[c] = nullptr;
        ev_[c] = nullptr;
      }
      return e;
    }
  }
  threads_ = std::max(1, threads);
  for (int k = 1; k < threads_; ++k) pool_.emplace_back(&HostStager::work, this, k);
  return hipSuccess;
}

HostStager::~HostStager() {
  {
    std::lock_guard<std::mutex> l(m_);
    stop_ = true;
  }
  go_.notify_all();
  for (auto& t : pool_) t.join();
  for (int b = 0; b < 2; ++b) {
    if (ev_[b]) (void)hipEventSynchronize(ev_[b]);
    if (ev_[b]) (void)hipEventDestroy(ev_[b]);
    if (pin_[b]) (void)hipHostFree(pin_[b]);
  }
}

void HostStager::work(int id) {
  unsigned long seen = 0;
  for (;;) {
    char* dst;
    const char* src;
    size_t n;
    {
      std::unique_lock<std::mutex> l(m_);
      go_.wait(l, [&] { return stop_ || gen_ != seen; });
      if (stop_) return;
      seen = gen_;
      dst = job_dst_;
      src = job_src_;
      n = job_n_;
    }
    size_t a, b;
    range_of(n, threads_, id, &a, &b);
    if (b > a) std::memcpy(dst + a, src + a, b - a);
    {
      std::lock_guard<std::mutex> l(m_);
      if (--pending_ == 0) done_.notify_one();
    }
  }
}

void HostStager::copy(void* dst, const void* src, size_t n) {
  if (threads_ == 1 || n < kSplitMin) {
    std::memcpy(dst, src, n);
    return;
  }
  {
    std::lock_guard<std::mutex> l(m_);
    job_dst_ = static_cast<char*>(dst);
    job_src_ = static_cast<const char*>(src);
    job_n_ = n;
    pending_ = threads_ - 1;
    ++gen_;
  }
  go_.notify_all();
  size_t a, b;
  range_of(n, threads_, 0, &a, &b);
  std::memcpy(static_cast<char*>(dst) + a, static_cast<const char*>(src) + a, b - a);
  std::unique_lock<std::mutex> l(m_);
  done_.wait(l, [&] { return pending_ == 0; });
}

hipError_t HostStager::to_device(void* dst, const void* src, size_t n, hipStream_t s) {
  char* d = static_cast<char*>(dst);
  const char* h = static_cast<const char*>(src);
  hipError_t e;
  for (size_t off = 0, k = 0; off < n; off += chunk_, ++k) {
    const int b = static_cast<int>(k & 1);
    const size_t len = std::min(chunk_, n - off);
    // chunk b is free once its previous DMA has read it
    if ((e = hipEventSynchronize(ev_[b])) != hipSuccess) return e;
    copy(pin_[b], h + off, len);
    if ((e = hipMemcpyAsync(d + off, pin_[b], len, hipMemcpyHostToDevice, s)) != hipSuccess) return e;
    if ((e = hipEventRecord(ev_[b], s)) != hipSuccess) return e;
  }
  return hipSuccess;
}

hipError_t HostStager::to_host(void* dst, const void* src, size_t n, hipStream_t s) {
  char* h = static_cast<char*>(dst);
  const char* d = static_cast<const char*>(src);
  if (n == 0) return hipSuccess;
  hipError_t e;
  const size_t nchunks = (n + chunk_ - 1) / chunk_;
  auto issue = [&](size_t k) -> hipError_t {
    const int b = static_cast<int>(k & 1);
    const size_t off = k * chunk_, len = std::min(chunk_, n - off);
    hipError_t r = hipEventSynchronize(ev_[b]);  // a pending upload may still read it
    if (r == hipSuccess) r = hipMemcpyAsync(pin_[b], d + off, len, hipMemcpyDeviceToHost, s);
    if (r == hipSuccess) r = hipEventRecord(ev_[b], s);
    return r;
  };
  if ((e = issue(0)) != hipSuccess) return e;
  for (size_t k = 0; k < nchunks; ++k) {
    // the next chunk's DMA runs while this one is copied out
    if (k + 1 < nchunks && (e = issue(k + 1)) != hipSuccess) return e;
    const int b = static_cast<int>(k & 1);
    if ((e = hipEventSynchronize(ev_[b])) != hipSuccess) return e;
    const size_t off = k * chunk_, len = std::min(chunk_, n - off);
    copy(h + off, pin_[b], len);
  }
  return hipSuccess;
}

}  // namespace eph
