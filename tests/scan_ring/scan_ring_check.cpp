// Checks of the served form's scan result ring (include/indexingpu_amd/scan_ring.hpp),
// built with g++ by tests/test_scan_ring.py (test infrastructure only).  Prints "ok" and
// exits 0 when every check holds; otherwise prints the failed check and exits 1.
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

#include "indexingpu_amd/scan_ring.hpp"

using indexingpu_amd::scan_ring;

#define CHECK(c)                                                   \
  do {                                                             \
    if (!(c)) {                                                    \
      std::printf("check failed at line %d: %s\n", __LINE__, #c); \
      std::exit(1);                                                \
    }                                                              \
  } while (0)

int main() {
  // 1. Rows of completed requests are not reused before the ring wraps, also when the
  //    ring was empty in between (a ring that restarted at row 0 when empty overwrote
  //    them).
  {
    scan_ring r;
    r.reset(1000);
    uint64_t row0, id;
    CHECK(r.reserve(300, row0, id) && row0 == 0);
    r.release(id);
    CHECK(r.empty());
    CHECK(r.reserve(300, row0, id) && row0 == 300);
    r.release(id);
    CHECK(r.reserve(400, row0, id) && row0 == 600);
    r.release(id);
    // a wave that would run past the end of an empty ring starts at row 0
    CHECK(r.reserve(50, row0, id) && row0 == 0);
    r.release(id);
  }
  // 2. A wave as large as the ring fits an empty ring wherever the head is.
  {
    scan_ring r;
    r.reset(1000);
    uint64_t row0, id;
    CHECK(r.reserve(123, row0, id));
    r.release(id);
    CHECK(r.reserve(1000, row0, id) && row0 == 0);
    uint64_t row1, id1;
    CHECK(!r.reserve(1, row1, id1));  // full
    r.release(id);
    CHECK(r.reserve(1, row1, id1));
  }
  // 3. Rings that exactly hold a run's rows never reuse a row (the alone-vs-mix test).
  {
    std::mt19937_64 g(7);
    for (int t = 0; t < 50; t++) {
      std::vector<uint64_t> sizes;
      uint64_t total = 0;
      for (int k = 0; k < 40; k++) {
        sizes.push_back(1 + g() % 200);
        total += sizes.back();
      }
      scan_ring r;
      r.reset(total);
      std::vector<char> used(total, 0);
      std::vector<uint64_t> live;
      for (uint64_t s : sizes) {
        uint64_t row0, id;
        while (!r.reserve(s, row0, id)) {  // release the oldest until it fits
          CHECK(!live.empty());
          r.release(live.front());
          live.erase(live.begin());
        }
        for (uint64_t x = row0; x < row0 + s; x++) {
          CHECK(!used[x]);
          used[x] = 1;
        }
        live.push_back(id);
        if (g() % 3 == 0) {  // complete a random live request (out of order)
          const size_t k = g() % live.size();
          r.release(live[k]);
          live.erase(live.begin() + (long)k);
        }
      }
    }
  }
  // 4. Random reserve/release in any order: a reservation never overlaps a live one,
  //    stays inside the ring, and an empty ring accepts any size up to its capacity.
  {
    std::mt19937_64 g(11);
    const uint64_t cap = 997;
    scan_ring r;
    r.reset(cap);
    std::vector<uint64_t> owner(cap, 0);  // 0: none, else allocation id + 1
    struct live_alloc {
      uint64_t id, row0, rows;
    };
    std::vector<live_alloc> live;
    for (int step = 0; step < 200000; step++) {
      if (!live.empty() && (g() % 2 == 0 || live.size() > 8)) {
        const size_t k = g() % live.size();
        for (uint64_t x = live[k].row0; x < live[k].row0 + live[k].rows; x++) owner[x] = 0;
        r.release(live[k].id);
        live.erase(live.begin() + (long)k);
        continue;
      }
      const uint64_t choice = g() % 4;
      const uint64_t rows = choice == 0 ? cap : choice == 1 ? cap / 2 : 1 + g() % cap;
      uint64_t row0, id;
      const bool was_empty = r.empty();
      if (!r.reserve(rows, row0, id)) {
        CHECK(!was_empty);
        continue;
      }
      CHECK(row0 + rows <= cap);
      for (uint64_t x = row0; x < row0 + rows; x++) {
        CHECK(owner[x] == 0);
        owner[x] = id + 1;
      }
      live.push_back({id, row0, rows});
    }
  }
  std::printf("ok\n");
  return 0;
}
