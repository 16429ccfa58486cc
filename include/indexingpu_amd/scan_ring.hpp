// Ring of scan result rows for the served form (host-side bookkeeping only; no HIP).
//
// A wave's rows are contiguous (the tail of the ring is skipped when they would wrap);
// allocations are released in allocation order, so the ring's free space is
// [head, tail + capacity).  The head only moves forward, so rows of completed requests
// are reused only when the ring wraps.  (Restarting at row 0 whenever the ring was empty
// overwrote the rows of requests that had just completed.)  An empty ring that must wrap
// for a wave starts it at row 0, so a wave as large as the ring always fits.
#pragma once
#include <cstdint>
#include <deque>

namespace indexingpu_amd {

class scan_ring {
 public:
  // empties the ring and sets its capacity (rows)
  void reset(uint64_t capacity) {
    cap_ = capacity;
    head_ = tail_ = next_id_ = front_id_ = 0;
    allocs_.clear();
  }
  // Reserves `rows` contiguous rows (rows <= capacity); returns false when they do not
  // fit until earlier allocations are released.  row0: the first row; id: for release().
  bool reserve(uint64_t rows, uint64_t& row0, uint64_t& id) {
    if (rows == 0 || rows > cap_) return false;
    const uint64_t start = head_ % cap_;
    uint64_t skip = start + rows > cap_ ? cap_ - start : 0;
    if (skip && allocs_.empty()) {  // every row is free: wrap now
      head_ += skip;
      tail_ = head_;
      skip = 0;
    }
    if (head_ + skip + rows - tail_ > cap_) return false;
    head_ += skip;
    row0 = head_ % cap_;
    head_ += rows;
    id = next_id_++;
    allocs_.push_back({head_, false});
    return true;
  }
  // releases allocation `id` (any order; space frees in allocation order)
  void release(uint64_t id) {
    allocs_[(size_t)(id - front_id_)].freed = true;
    while (!allocs_.empty() && allocs_.front().freed) {
      tail_ = allocs_.front().end;
      allocs_.pop_front();
      front_id_++;
    }
  }
  bool empty() const { return allocs_.empty(); }
  uint64_t capacity() const { return cap_; }

 private:
  struct alloc {
    uint64_t end;  // head after this allocation
    bool freed;
  };
  std::deque<alloc> allocs_;
  uint64_t cap_ = 0, head_ = 0, tail_ = 0, next_id_ = 0, front_id_ = 0;
};

}  // namespace indexingpu_amd
