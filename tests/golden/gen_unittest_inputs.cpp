// Writes the random draws of the reference's Masstree unit-test input generator
// (test/unittest_masstree.cu:104-124 make_input, :185-215 shuffle, :571-602
// ConcurrentMix shuffle) for the four instantiated (key, value) length ranges
// (test/unittest_masstree.cu:755-758), so that tests can rebuild the exact inputs.
//
// std::uniform_int_distribution and std::shuffle are implementation-defined, so the
// draws are produced here with the same library family the reference's host code
// uses (libstdc++, std::mt19937 seeded with 0) and committed as data.
//
// Output (text, one line per array): "name n v0 v1 ...".  Also a checksum of every
// rebuilt key slice so the Python rebuild can be checked against this C++ one.
// Key slice s of key i with length L is ((i >> (L-1-s)) + 1) * 2 + s with the shift
// taken modulo 32, which is what the x86 scalar shift instruction does for the
// reference's `1u << (key_length - 1 - s)` when L-1-s >= 32.
#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <random>
#include <vector>

static void emit(const char* name, const std::vector<uint32_t>& v) {
  std::printf("%s %zu", name, v.size());
  for (uint32_t x : v) std::printf(" %u", x);
  std::printf("\n");
}

int main() {
  const size_t num_keys = 1024;
  const uint32_t cfg[4][4] = {{1, 1, 1, 1}, {1, 200, 1, 1}, {1, 1, 1, 200}, {1, 200, 1, 200}};
  for (int c = 0; c < 4; c++) {
    uint32_t kmin = cfg[c][0], kmax = cfg[c][1], vmin = cfg[c][2], vmax = cfg[c][3];
    std::mt19937 rng(0);
    std::uniform_int_distribution<uint32_t> kd(kmin, kmax);
    std::uniform_int_distribution<uint32_t> vd(vmin, vmax);
    std::vector<uint32_t> klen(num_keys), vlen(num_keys);
    uint64_t checksum = 1469598103934665603ull;
    for (size_t i = 0; i < num_keys; i++) {
      uint32_t L = kd(rng);
      uint32_t V = vd(rng);
      klen[i] = L;
      vlen[i] = V;
      for (uint32_t s = 0; s < L; s++) {
        volatile uint32_t sh = (L - 1 - s) & 31u;
        uint32_t factor = 1u << sh;
        uint32_t slice = static_cast<uint32_t>(i / factor + 1) * 2 + s;
        checksum = (checksum ^ slice) * 1099511628211ull;
      }
    }
    char name[64];
    std::snprintf(name, sizeof(name), "cfg%d_key_lengths", c);
    emit(name, klen);
    std::snprintf(name, sizeof(name), "cfg%d_value_lengths", c);
    emit(name, vlen);
    std::snprintf(name, sizeof(name), "cfg%d_key_checksum", c);
    emit(name, {static_cast<uint32_t>(checksum), static_cast<uint32_t>(checksum >> 32)});
  }
  // testing_input::shuffle(): order = iota, std::shuffle with mt19937(0)
  {
    std::vector<uint32_t> order(num_keys);
    for (size_t i = 0; i < num_keys; i++) order[i] = static_cast<uint32_t>(i);
    std::mt19937 rng(0);
    std::vector<size_t> o(num_keys);
    for (size_t i = 0; i < num_keys; i++) o[i] = i;
    std::shuffle(o.begin(), o.end(), rng);
    for (size_t i = 0; i < num_keys; i++) order[i] = static_cast<uint32_t>(o[i]);
    emit("input_shuffle_order", order);
  }
  // ConcurrentMix: 10 * (num_keys / 4) requests, std::shuffle with mt19937(0)
  {
    size_t n = 10 * (num_keys / 4);
    std::vector<size_t> o(n);
    for (size_t i = 0; i < n; i++) o[i] = i;
    std::mt19937 rng(0);
    std::shuffle(o.begin(), o.end(), rng);
    std::vector<uint32_t> order(n);
    for (size_t i = 0; i < n; i++) order[i] = static_cast<uint32_t>(o[i]);
    emit("mix_shuffle_order", order);
  }
  return 0;
}
