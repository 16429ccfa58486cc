// Node and record formats of the MI355X Masstree (B+tree of tries over u32 key slices).
//
// Every tree object lives in a 128-byte slab (one gfx950 L2 line).  A slab is read as
// 16 slots of 8 bytes; lane i of a 16-lane tile owns slot i, so one coalesced
// 128-byte load brings a whole node, including its header word, into the tile.
//
// Border (leaf) node, up to 14 entries:
//   slot 0..13  : (lo = key slice, hi = value or object index), sorted by (slice, more)
//   slot 14     : (lo = high key, hi = 2-bit keystate of each entry)
//   slot 15     : (lo = header word, hi = right sibling index or kInvalid)
// Interior node, up to 15 children:
//   slot 0..n-1 : (lo = upper bound of child i, hi = child i); the last bound is the
//                 node's high key (0xFFFFFFFF for the rightmost node of a level)
//   slot 15     : (lo = header word, hi = right sibling index or kInvalid)
// Routing: child i covers (bound[i-1], bound[i]]; a search for slice q takes the
// first i with q <= bound[i]; no such i means the node split and q lives to the
// right (B-link move-right).
//
// Header word: nkeys:5 | border:1 | root:1 | lock:1 | level:4 | version:20.  Level 15
// marks a layer root that an erase removed (hdr_removed).
// Writers set `lock` with an atomic fetch_or, rewrite the node, then release-store
// the header with `lock` cleared and `version` advanced; optimistic readers accept a
// node image only if its header was unlocked and is unchanged when re-read.
//
// Keystates (what an entry's slice means, following the reference's four kinds,
// include/masstree_node_subwarp.hpp:37-40):
//   kValue   : key ends at this slice, value is the inline u32 (value length 1)
//   kLongVal : key ends at this slice, hi = record holding the value (length != 1)
//   kSuffix  : key continues, hi = record holding the key tail and the value
//   kLink    : key continues, hi = root of the next layer (a trie level)
// Entry order inside a layer is (slice, more) with more = (keystate >= kSuffix), so
// a key that ends at a slice sorts before every key that continues past it.
//
// Record (key tail + value), a chain of slabs: data word d lives in slab d / 31 at
// word d % 31; word 31 of each slab holds the next slab index.  Data word 0 is the
// length word (value_length << 16 | tail_length); tail slices follow, then value
// slices.  Records are immutable once published and are retired, never rewritten.
#pragma once
#include <cstdint>

namespace indexingpu_amd {

using u32 = uint32_t;
using u64 = uint64_t;

constexpr u32 kInvalid = 0xFFFFFFFFu;
constexpr u32 kSlabBytes = 128;
constexpr u32 kSlots = 16;          // 8-byte slots per node
constexpr u32 kMetaSlot = 15;       // header + sibling
constexpr u32 kHighSlot = 14;       // border: high key + keystates
constexpr u32 kBorderCap = 14;      // entries per border node
constexpr u32 kInteriorCap = 15;    // children per interior node
constexpr u32 kMaxSlice = 0xFFFFFFFFu;
constexpr u32 kRecordWordsPerSlab = 31;
constexpr u32 kRecordNextWord = 31;
constexpr u32 kMaxLength = 0xFFFFu;  // 16-bit record lengths

// keystates
constexpr u32 kValue = 0, kLongVal = 1, kSuffix = 2, kLink = 3;

// header word fields
constexpr u32 kNkeysMask = 0x1Fu;
constexpr u32 kBorderBit = 1u << 5;
constexpr u32 kRootBit = 1u << 6;
constexpr u32 kLockBit = 1u << 7;
constexpr u32 kLevelShift = 8;
constexpr u32 kLevelMask = 0xFu << kLevelShift;
constexpr u32 kVersionShift = 12;
constexpr u32 kVersionOne = 1u << kVersionShift;

#if defined(__HIPCC__)
#define IGPU_HD __host__ __device__ __forceinline__
#else
#define IGPU_HD inline
#endif

IGPU_HD u32 hdr_nkeys(u32 h) { return h & kNkeysMask; }
IGPU_HD bool hdr_border(u32 h) { return (h & kBorderBit) != 0; }
IGPU_HD bool hdr_root(u32 h) { return (h & kRootBit) != 0; }
IGPU_HD bool hdr_locked(u32 h) { return (h & kLockBit) != 0; }
IGPU_HD u32 hdr_level(u32 h) { return (h & kLevelMask) >> kLevelShift; }
IGPU_HD u32 hdr_with_nkeys(u32 h, u32 n) { return (h & ~kNkeysMask) | n; }
IGPU_HD u32 hdr_make(u32 nkeys, bool border, bool root, u32 level) {
  return nkeys | (border ? kBorderBit : 0u) | (root ? kRootBit : 0u) | (level << kLevelShift);
}
// header published at unlock: lock cleared, version advanced
IGPU_HD u32 hdr_unlocked_next(u32 h) { return (h & ~kLockBit) + kVersionOne; }
// A removed layer root (an erase emptied it and unlinked it) carries level 15, which no
// live node reaches (15 levels of at least 7 children each exceed any pool).
constexpr u32 kRemovedLevel = 15;
IGPU_HD bool hdr_removed(u32 h) { return hdr_level(h) == kRemovedLevel; }
IGPU_HD u32 hdr_remove(u32 h) { return (h & ~kLevelMask) | (kRemovedLevel << kLevelShift); }
IGPU_HD u32 hdr_revive(u32 h) { return h & ~kLevelMask; }  // back to a live border root

IGPU_HD bool ks_more(u32 ks) { return ks >= kSuffix; }
IGPU_HD bool ks_record(u32 ks) { return ks == kLongVal || ks == kSuffix; }
IGPU_HD u32 ks_get(u32 bits, u32 pos) { return (bits >> (2 * pos)) & 3u; }

// (slice, more) order used inside a layer
IGPU_HD bool entry_less(u32 s1, bool m1, u32 s2, bool m2) {
  return s1 < s2 || (s1 == s2 && !m1 && m2);
}

// Interior-node version tag.  Bit 31 of the child index in slots 0..7 (the first 64 bytes
// of the line; slab indices stay below 2^31) holds bit i of the low 8 bits of the version
// the node has once its writer releases it; the header (slot 15) is in the other half.
// A reader may take an interior node from a plain cached load if the header is unlocked
// and the tag matches: a line whose halves come from different versions fails the check,
// while an old but whole line is harmless to a B-link descent (keys only move right).
constexpr u32 kTagSlots = 8;
constexpr u32 kChildMask = 0x7FFFFFFFu;
constexpr u32 kMaxPoolSlabs = 0x7FFFFFE0u;
// version a node will carry once published: writers store locked nodes and release them
// with a version increment; nodes stored unlocked keep their version
IGPU_HD u32 hdr_publish_tag(u32 h) {
  return ((h >> kVersionShift) + (hdr_locked(h) ? 1u : 0u)) & 0xFFu;
}
IGPU_HD u32 hdr_tag(u32 h) { return (h >> kVersionShift) & 0xFFu; }
IGPU_HD u32 child_of(u64 s) { return (u32)(s >> 32) & kChildMask; }

IGPU_HD u32 rec_len_word(u32 tail_len, u32 value_len) { return (value_len << 16) | tail_len; }
IGPU_HD u32 rec_tail_len(u32 w) { return w & 0xFFFFu; }
IGPU_HD u32 rec_value_len(u32 w) { return w >> 16; }
IGPU_HD u32 rec_num_slabs(u32 total_words) {
  return (total_words + kRecordWordsPerSlab - 1) / kRecordWordsPerSlab;
}

// Retire-list entries: a record head index, or a node slab index with this bit set
// (slab indices stay below 2^31).
constexpr u32 kRetiredNode = 0x80000000u;

IGPU_HD u64 slot_make(u32 lo, u32 hi) { return (u64)lo | ((u64)hi << 32); }
IGPU_HD u32 slot_lo(u64 s) { return (u32)s; }
IGPU_HD u32 slot_hi(u64 s) { return (u32)(s >> 32); }

// Error codes reported through the index's device error word (first error wins).
enum error_code : u32 {
  kErrNone = 0,
  kErrOutOfMemory = 1,      // slab pool exhausted
  kErrStepLimit = 2,        // an operation exceeded its step budget (livelock guard)
  kErrBadIndex = 3,         // an object index outside the pool (corruption guard)
  kErrBadLength = 4,        // key length 0 or a length above 65535 slices
  kErrRetireOverflow = 5,   // retire list full (slab leaked, results unaffected)
  kErrScanScratch = 6,      // scan deeper than its per-tile layer stack
};

// The livelock guard's record of the operation it abandoned, kept beside the error word
// (words kTripOffset .. kTripOffset + kTripWords - 1 of the index's error area) by the
// guard whose error code was the first: trip_valid (1), phase, node, layer, layer_root,
// a0, a1, a2, a3, snap_hdr, the node's header when the guard fired, steps, request kind,
// key length, the key slice of the current layer, lvl | fresh << 16.
constexpr u32 kTripOffset = 16;
constexpr u32 kTripWords = 16;
// Word of the error area counting merges of layer-0 border nodes that did not rewrite the
// border-level shortcut table (never cleared): the table is used only while this equals its
// value when the table was built.
constexpr u32 kMergeEpochWord = 8;
// Entries of a border-level shortcut table a merge rewrites at most (the entries naming the
// node that goes are one run); a longer run retires the table instead.
constexpr u32 kShortcutPatchMax = 64;

}  // namespace indexingpu_amd
