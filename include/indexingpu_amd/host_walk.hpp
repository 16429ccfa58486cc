// Host-side walk over a copy of the slab pool: structural validation, ordered dump of
// all entries, and space statistics.  It plays the role of the reference's debug
// traverse/validate/print (include/gpu_masstree.hpp:1354-1472): it checks that border
// keys are sorted with "key ends here" before "key continues" for equal slices
// (gpu_masstree.hpp:1432-1441), plus this layout's own B-link invariants.  Debug and
// test use only; it never runs on the batch path.
#pragma once
#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <functional>
#include <string>
#include <unordered_set>
#include <vector>

#include "layout.hpp"

namespace indexingpu_amd {

struct walk_stats {
  uint64_t entries = 0;
  uint64_t nodes = 0;           // border + interior nodes
  uint64_t border_nodes = 0;
  uint64_t layers = 0;
  uint64_t record_slabs = 0;
  uint64_t max_depth = 0;       // deepest B-link level count in any layer
  uint64_t unlinked = 0;        // non-root nodes with no parent slot (reached by sibling links)
  uint64_t empty_borders = 0;   // non-root border nodes with no entry
  uint64_t max_border_keys = 0;      // largest entry count of a border node
  uint64_t max_interior_children = 0;  // largest child count of an interior node
  std::string error;            // empty when every invariant holds
};

// fetch(idx) returns a pointer to the 16 slots of slab idx, or nullptr if out of range.
using slab_fetch_fn = std::function<const uint64_t*(uint32_t)>;
// emit(key slices, value slices) is called for every entry in key order.
using entry_emit_fn = std::function<void(const std::vector<uint32_t>&, const std::vector<uint32_t>&)>;

class host_walker {
 public:
  host_walker(slab_fetch_fn fetch, entry_emit_fn emit) : fetch_(fetch), emit_(emit) {}

  walk_stats run(uint32_t root) {
    stats_ = walk_stats{};
    std::vector<uint32_t> prefix;
    walk_layer(root, prefix, 0);
    return stats_;
  }

 private:
  bool fail(const std::string& msg) {
    if (stats_.error.empty()) stats_.error = msg;
    return false;
  }

  const uint64_t* node(uint32_t idx) {
    const uint64_t* p = fetch_(idx);
    if (!p) fail("object index out of pool: " + std::to_string(idx));
    return p;
  }

  bool read_record(uint32_t idx, std::vector<uint32_t>& tail, std::vector<uint32_t>& value) {
    std::vector<uint32_t> words;
    uint32_t total = 0, need = 1;
    uint32_t cur = idx;
    uint64_t slabs = 0;
    while (total < need) {
      const uint64_t* p = node(cur);
      if (!p) return false;
      slabs++;
      const uint32_t* w = reinterpret_cast<const uint32_t*>(p);
      for (uint32_t i = 0; i < kRecordWordsPerSlab && total < need; i++) {
        words.push_back(w[i]);
        total++;
        if (total == 1) need = 1 + rec_tail_len(words[0]) + rec_value_len(words[0]);
      }
      if (total < need) cur = w[kRecordNextWord];
      if (slabs > 5000) return fail("record chain too long");
    }
    stats_.record_slabs += slabs;
    uint32_t tl = rec_tail_len(words[0]);
    tail.assign(words.begin() + 1, words.begin() + 1 + tl);
    value.assign(words.begin() + 1 + tl, words.end());
    return true;
  }

  bool walk_layer(uint32_t root, std::vector<uint32_t>& prefix, int depth) {
    if (depth > 70000) return fail("layer depth runaway");
    stats_.layers++;
    const uint64_t* r = node(root);
    if (!r) return false;
    uint32_t h = slot_lo(r[kMetaSlot]);
    if (!hdr_root(h)) return fail("layer root without root bit at " + std::to_string(root));
    // descend the leftmost path, checking levels
    uint32_t cur = root;
    uint32_t level = hdr_level(h);
    std::vector<uint32_t> level_heads;
    stats_.max_depth = std::max<uint64_t>(stats_.max_depth, level + 1);
    while (true) {
      const uint64_t* p = node(cur);
      if (!p) return false;
      uint32_t hh = slot_lo(p[kMetaSlot]);
      if (hdr_locked(hh)) return fail("node left locked: " + std::to_string(cur));
      if (hdr_level(hh) != level) return fail("level mismatch at " + std::to_string(cur));
      level_heads.push_back(cur);
      if (hdr_border(hh)) break;
      if (hdr_nkeys(hh) == 0) return fail("empty interior node");
      cur = child_of(p[0]);
      if (level == 0) return fail("interior node at level 0");
      level--;
    }
    // check each interior level: bounds ascending, children at level-1
    std::unordered_set<uint32_t> linked;  // children named by the level above
    for (size_t li = 0; li + 1 < level_heads.size(); li++) {
      std::unordered_set<uint32_t> next_linked;
      uint32_t n = level_heads[li];
      uint32_t prev_high = 0;
      bool first = true;
      uint64_t guard = 0;
      while (n != kInvalid) {
        if (++guard > (1ull << 32)) return fail("sibling cycle");
        const uint64_t* p = node(n);
        if (!p) return false;
        uint32_t hh = slot_lo(p[kMetaSlot]);
        if (hdr_border(hh)) return fail("border node in interior level");
        if (hdr_removed(hh)) return fail("removed node reachable: " + std::to_string(n));
        uint32_t nk = hdr_nkeys(hh);
        if (nk == 0 || nk > kInteriorCap) return fail("bad interior nkeys");
        stats_.max_interior_children = std::max<uint64_t>(stats_.max_interior_children, nk);
        uint32_t tag = 0;
        for (uint32_t i = 0; i < kTagSlots; i++) tag |= (uint32_t)(p[i] >> 63) << i;
        if (tag != hdr_tag(hh)) return fail("interior version tag mismatch");
        if (li > 0 && !linked.count(n)) stats_.unlinked++;
        for (uint32_t i = 0; i < nk; i++) {
          next_linked.insert(child_of(p[i]));
          uint32_t b = slot_lo(p[i]);
          if (i > 0 && b <= slot_lo(p[i - 1])) return fail("interior bounds not ascending");
          if (!first && i == 0 && b <= prev_high) return fail("interior bound below left high");
          const uint64_t* c = node(child_of(p[i]));
          if (!c) return false;
          const uint32_t ch = slot_lo(c[kMetaSlot]);
          if (hdr_level(ch) + 1 != hdr_level(hh)) return fail("child level mismatch");
          // a child's keys never exceed its slot's bound (its high key, or its last bound)
          const uint32_t chigh = hdr_border(ch) ? slot_lo(c[kHighSlot])
                                 : (hdr_nkeys(ch) ? slot_lo(c[hdr_nkeys(ch) - 1]) : 0u);
          if (chigh > b) return fail("child above its parent bound");
        }
        prev_high = slot_lo(p[nk - 1]);
        first = false;
        stats_.nodes++;
        uint32_t sib = slot_hi(p[kMetaSlot]);
        if (sib == kInvalid && prev_high != kMaxSlice)
          return fail("rightmost interior node without max bound");
        n = sib;
      }
      linked.swap(next_linked);
    }
    // walk the border level left to right
    uint32_t n = level_heads.back();
    bool have_prev = false;
    uint32_t prev_slice = 0;
    bool prev_more = false;
    uint32_t prev_high = 0;
    bool first = true;
    uint64_t guard = 0;
    while (n != kInvalid) {
      if (++guard > (1ull << 32)) return fail("sibling cycle");
      const uint64_t* p = node(n);
      if (!p) return false;
      uint32_t hh = slot_lo(p[kMetaSlot]);
      if (!hdr_border(hh)) return fail("interior node in border level");
      if (hdr_locked(hh)) return fail("border node left locked");
      if (hdr_removed(hh)) return fail("removed node reachable: " + std::to_string(n));
      uint32_t nk = hdr_nkeys(hh);
      if (nk > kBorderCap) return fail("bad border nkeys");
      stats_.max_border_keys = std::max<uint64_t>(stats_.max_border_keys, nk);
      uint32_t high = slot_lo(p[kHighSlot]);
      if (!first && high <= prev_high) return fail("border high not above left sibling's");
      if (nk == 0 && level_heads.size() > 1) stats_.empty_borders++;
      uint32_t ksb = slot_hi(p[kHighSlot]);
      if (level_heads.size() > 1 && !linked.count(n)) stats_.unlinked++;
      stats_.nodes++;
      stats_.border_nodes++;
      for (uint32_t i = 0; i < nk; i++) {
        uint32_t s = slot_lo(p[i]);
        uint32_t v = slot_hi(p[i]);
        uint32_t ks = ks_get(ksb, i);
        bool more = ks_more(ks);
        if (have_prev && !entry_less(prev_slice, prev_more, s, more))
          return fail("border entries out of order");
        if (s > high) return fail("entry above node high key");
        if (!first && s <= prev_high) return fail("entry not above left sibling's high key");
        have_prev = true;
        prev_slice = s;
        prev_more = more;
        prefix.push_back(s);
        if (ks == kValue) {
          if (!emit_value(prefix, {v})) return false;
        } else if (ks == kLink) {
          if (!walk_layer(v, prefix, depth + 1)) return false;
        } else {
          std::vector<uint32_t> tail, value;
          if (!read_record(v, tail, value)) return false;
          if (ks == kLongVal && !tail.empty()) return fail("long value record with a tail");
          if (ks == kSuffix && tail.empty()) return fail("suffix record without a tail");
          std::vector<uint32_t> key = prefix;
          key.insert(key.end(), tail.begin(), tail.end());
          if (!emit_value(key, value)) return false;
        }
        prefix.pop_back();
      }
      prev_high = high;
      first = false;
      uint32_t sib = slot_hi(p[kMetaSlot]);
      if (sib == kInvalid && high != kMaxSlice) return fail("rightmost border without max high");
      n = sib;
    }
    return stats_.error.empty();
  }

  bool emit_value(const std::vector<uint32_t>& key, const std::vector<uint32_t>& value) {
    stats_.entries++;
    if (emit_) emit_(key, value);
    return true;
  }

  slab_fetch_fn fetch_;
  entry_emit_fn emit_;
  walk_stats stats_;
};

}  // namespace indexingpu_amd
