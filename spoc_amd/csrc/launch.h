// Arguments shared by the host context and the kernel launchers.
#pragma once
#include <hip/hip_runtime.h>

#include "common.h"

namespace spoc {

// Optional per-launch timing hook: called with a slot index before/after
// each kernel of one right-hand side.
typedef void (*TimerHook)(void *user, int slot, hipStream_t st);



// Bound on the kernel dispatches queued on a stream ahead of the host.
// Callers announce the dispatches they are about to queue (admit), queue
// them, then record a mark (mark).  admit waits for the oldest marks until
// the announced work fits under kMaxQueued outstanding dispatches.  A run of
// ~1,600 dispatches queued before one wait ended in a malformed-packet
// error of the runtime under the profiler's kernel tracing; 512 is half of
// one hardware queue's 1,024 packet slots.
class QueueBound {
 public:
  static constexpr int kMaxQueued = 512, kMarks = 16;
  QueueBound() = default;
  QueueBound(const QueueBound &) = delete;
  QueueBound &operator=(const QueueBound &) = delete;
  ~QueueBound() {
    for (hipEvent_t e : ev_)
      if (e) (void)hipEventDestroy(e);
  }
  hipError_t admit(int n) {
    while (len_ > 0 && (total_ + pending_ + n > kMaxQueued || len_ == kMarks)) {
      hipError_t e = hipEventSynchronize(ev_[head_]);
      if (e != hipSuccess) return e;
      total_ -= n_[head_];
      head_ = (head_ + 1) % kMarks;
      len_--;
    }
    pending_ += n;
    return hipSuccess;
  }
  hipError_t mark(hipStream_t s) {
    if (!pending_) return hipSuccess;
    if (len_ == kMarks) {
      hipError_t e = admit(0);
      if (e != hipSuccess) return e;
    }
    const int i = (head_ + len_) % kMarks;
    if (!ev_[i]) {
      hipError_t e = hipEventCreateWithFlags(&ev_[i], hipEventDisableTiming);
      if (e != hipSuccess) return e;
    }
    hipError_t e = hipEventRecord(ev_[i], s);
    if (e != hipSuccess) return e;
    n_[i] = pending_;
    total_ += pending_;
    pending_ = 0;
    len_++;
    return hipSuccess;
  }
  // forget every mark (after the stream has been synchronised)
  void reset() { head_ = len_ = total_ = pending_ = 0; }
  int outstanding() const { return total_ + pending_; }

 private:
  hipEvent_t ev_[kMarks] = {};
  int n_[kMarks] = {};
  int head_ = 0, len_ = 0, total_ = 0, pending_ = 0;
};

}  // namespace spoc

