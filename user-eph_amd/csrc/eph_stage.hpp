// eph_stage.hpp -- host side of the host-array hooks: copies between the
// caller's (pageable) arrays and device buffers through pinned staging
// chunks.  The host copy of a chunk is split over a few worker threads and
// overlaps the DMA of the neighbouring chunk; pageable memory would otherwise
// be copied by the runtime at about half the link rate (one host thread).
#pragma once

#include <hip/hip_runtime.h>

#include <condition_variable>
#include <cstddef>
#include <mutex>
#include <thread>
#include <vector>

namespace eph {

class HostStager {
 public:
  HostStager() = default;
  HostStager(const HostStager&) = delete;
  HostStager& operator=(const HostStager&) = delete;
  ~HostStager();

  // pinned chunks of chunk_bytes and threads - 1 workers (threads >= 1)
  hipError_t init(int threads, size_t chunk_bytes);
  bool ready() const { return pin_[0] != nullptr; }

  // n bytes host -> device on stream s; returns once src may be reused (the
  // last chunk's DMA may still be in flight, ordered before later work on s)
  hipError_t to_device(void* dst, const void* src, size_t n, hipStream_t s);
  // n bytes device -> host after the work already enqueued on s; returns once
  // dst holds them
  hipError_t to_host(void* dst, const void* src, size_t n, hipStream_t s);

  // threads from EPH_HOST_THREADS (default 4, at most the hardware threads)
  static int default_threads();

 private:
  void copy(void* dst, const void* src, size_t n);  // split over the workers
  void work(int id);

  char* pin_[2] = {nullptr, nullptr};
  hipEvent_t ev_[2] = {nullptr, nullptr};
  size_t chunk_ = 0;
  int threads_ = 1;
  // worker pool: one job at a time, split into threads_ ranges
  std::vector<std::thread> pool_;
  std::mutex m_;
  std::condition_variable go_, done_;
  unsigned long gen_ = 0;
  int pending_ = 0;
  bool stop_ = false;
  char* job_dst_ = nullptr;
  const char* job_src_ = nullptr;
  size_t job_n_ = 0;
};

}  // namespace eph
