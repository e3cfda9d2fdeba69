// Opt-in per-kernel timing with HIP events recorded on the launch stream
// around each kernel launch (off by default; no events are touched when off).
// The event pool is created by xl_timing_enable, outside any timed region, and
// reused; xl_timing_collect waits on the recorded events, sums the durations
// per kernel and empties the pool.  Not for use during graph capture.
#pragma once
#include <hip/hip_runtime.h>

#include <stdint.h>

#include <atomic>
#include <mutex>
#include <vector>

namespace xl {

enum KernelId {
  K_RS_HROW = 0,
  K_RS_UROW,
  K_RS_COL,
  K_RS_FINAL,
  K_CZT_SETUP,
  K_CZT_ROWS,
  K_CZT_COLS,
  K_PRE_FWD,
  K_PRE_BWD,
  K_TF,
  K_GEMM_PRE,
  K_GEMM_Y,
  K_GEMM_X,
  K_MSE,
  K_ADAM,
  K_RS_HCOL,
  K_LOSS,
  K_BS,
  K_COUNT
};

inline const char* kernel_name(int id) {
  static const char* names[K_COUNT] = {"rs_hrow", "rs_urow", "rs_col", "rs_final", "czt_setup", "czt_rows",
                                       "czt_cols", "pre_forward", "pre_backward", "transfer_function",
                                       "gemm_pre", "gemm_y", "gemm_x", "mse_intensity",
                                       "adam", "rs_hcol", "loss", "beam_splitter"};
  return (id >= 0 && id < K_COUNT) ? names[id] : "?";
}

struct TimingState {
  std::mutex mu;
  std::atomic<bool> on{false};      // read without the lock on every launch
  uint64_t gen = 0;                 // bumped whenever the pool is rebuilt or emptied
  std::vector<hipEvent_t> pool;     // 2 events per record
  std::vector<int> ids;             // kernel id per used record
  size_t used = 0;                  // records in use
  int64_t dropped = 0;
};

inline TimingState& timing_state() {
  static TimingState s;
  return s;
}

// Records a start event on construction and a stop event on destruction.
// Both take the state's lock; a record whose pool was rebuilt or collected in
// between (another host thread calling xl_timing_enable / collect) is dropped.
class KTimer {
 public:
  KTimer(int id, hipStream_t st) : st_(st) {
    TimingState& s = timing_state();
    if (!s.on.load(std::memory_order_relaxed)) return;
    std::lock_guard<std::mutex> lk(s.mu);
    if (!s.on.load(std::memory_order_relaxed)) return;
    if (2 * (s.used + 1) > s.pool.size()) {
      ++s.dropped;
      return;
    }
    slot_ = (long)s.used++;
    gen_ = s.gen;
    s.ids[slot_] = id;
    (void)hipEventRecord(s.pool[2 * slot_], st_);
  }
  ~KTimer() {
    if (slot_ < 0) return;
    TimingState& s = timing_state();
    std::lock_guard<std::mutex> lk(s.mu);
    if (s.gen != gen_ || (size_t)slot_ >= s.used) return;
    (void)hipEventRecord(s.pool[2 * slot_ + 1], st_);
  }
  KTimer(const KTimer&) = delete;
  KTimer& operator=(const KTimer&) = delete;

 private:
  hipStream_t st_;
  long slot_ = -1;
  uint64_t gen_ = 0;
};

}  // namespace xl
