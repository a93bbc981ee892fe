// Host side of the host-array transfers (hipPPMSetBaryon / hipPPMGetBaryon
// with the caller's pageable arrays): a small pool of host threads that
// copies between the caller's memory and pinned staging buffers, so that
// the host copy of one chunk overlaps the DMA of the previous one.
#pragma once

#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <cstring>
#include <mutex>
#include <thread>
#include <vector>

namespace ppm {

// memcpy split over the calling thread and up to 7 workers (pageable host
// memory is copied at a few GB/s per thread; the DMA engine moves pinned
// memory at ~57 GB/s).  One copy at a time per pool; copy() returns when
// every byte is in place.
class CopyPool {
 public:
  static CopyPool &get() {
    static CopyPool pool;
    return pool;
  }

  void copy(void *dst, const void *src, size_t n) {
    if (n < kMinParallel || workers_.empty()) {
      std::memcpy(dst, src, n);
      return;
    }
    std::lock_guard<std::mutex> one(busy_);  // callers from several threads take turns
    Job j;
    j.dst = static_cast<char *>(dst);
    j.src = static_cast<const char *>(src);
    j.n = n;
    const size_t parts = 4 * (workers_.size() + 1);
    j.slice = std::max<size_t>(kMinSlice, (n + parts - 1) / parts);
    j.nslices = (n + j.slice - 1) / j.slice;
    {
      std::lock_guard<std::mutex> lk(m_);
      job_ = &j;
      ++gen_;
    }
    cv_.notify_all();
    run(j);
    std::unique_lock<std::mutex> lk(m_);
    job_ = nullptr;  // no worker takes the job from here on
    done_.wait(lk, [&] { return j.users == 0; });
  }

 private:
  static constexpr size_t kMinParallel = size_t(1) << 20;
  static constexpr size_t kMinSlice = size_t(256) << 10;

  struct Job {
    char *dst = nullptr;
    const char *src = nullptr;
    size_t n = 0, slice = 0, nslices = 0;
    std::atomic<size_t> next{0};
    int users = 0;  // workers inside run(); guarded by m_
  };

  CopyPool() {
    const unsigned hw = std::thread::hardware_concurrency();
    const unsigned nw = std::min(7u, hw > 1 ? hw - 1 : 0u);
    for (unsigned k = 0; k < nw; k++) workers_.emplace_back([this] { loop(); });
  }
  ~CopyPool() {
    {
      std::lock_guard<std::mutex> lk(m_);
      stop_ = true;
    }
    cv_.notify_all();
    for (std::thread &t : workers_) t.join();
  }

  static void run(Job &j) {
    for (;;) {
      const size_t k = j.next.fetch_add(1);
      if (k >= j.nslices) return;
      const size_t off = k * j.slice, len = std::min(j.slice, j.n - off);
      std::memcpy(j.dst + off, j.src + off, len);
    }
  }

  void loop() {
    unsigned seen = 0;
    for (;;) {
      Job *j;
      {
        std::unique_lock<std::mutex> lk(m_);
        cv_.wait(lk, [&] { return stop_ || gen_ != seen; });
        if (stop_) return;
        seen = gen_;
        j = job_;
        if (!j) continue;  // that copy has already finished
        j->users++;
      }
      run(*j);
      {
        std::lock_guard<std::mutex> lk(m_);
        j->users--;
      }
      done_.notify_all();
    }
  }

  std::vector<std::thread> workers_;
  std::mutex m_, busy_;
  std::condition_variable cv_, done_;
  Job *job_ = nullptr;
  unsigned gen_ = 0;
  bool stop_ = false;
};

}  // namespace ppm
