"""Text-key dataset conversion (the reference's loader, generate_workload.hpp:131-190):
against a byte-by-byte restatement, edge cases of its line handling, and the property
that numeric slice-tuple order equals byte order of the strings."""
import numpy as np

from indexingpu_amd import workloads as W


def _slow(s, big_endian):
    out = []
    for i in range(0, len(s), 4):
        v = 0
        for j in range(4):
            v = (v << 8) | (s[i + j] if i + j < len(s) else 0)
        if big_endian:
            v = int.from_bytes(v.to_bytes(4, "big"), "little")
        out.append(v)
    return out


def _urls(n, seed=3):
    rng = np.random.default_rng(seed)
    hosts = [b"example.org", b"a.b", b"news.site.com", b"x"]
    out = []
    for _ in range(n):
        h = hosts[rng.integers(len(hosts))]
        path = bytes(rng.integers(97, 123, size=rng.integers(0, 30)).astype(np.uint8))
        out.append(b"http://" + h + b"/" + path)
    return out


def test_matches_bytewise_restatement():
    lines = _urls(300) + [b"abcd", b"abcdefgh", b"z", b"abc\r", b"", b"\r"]
    for be in (False, True):
        keys, lens, lo, hi = W.dataset_keys(lines, big_endian=be)
        kept = [l[:-1] if l.endswith(b"\r") else l for l in lines]
        kept = [l for l in kept if l]
        assert len(lens) == len(kept)
        assert lo == min((len(s) + 3) // 4 for s in kept) and hi == keys.shape[1]
        for i, s in enumerate(kept):
            exp = _slow(s, be)
            assert int(lens[i]) == len(exp)
            assert keys[i, :len(exp)].tolist() == exp
            assert not keys[i, len(exp):].any()


def test_append_zero_at_end():
    keys, lens, _, hi = W.dataset_keys([b"abcd", b"abcdefghij", b"abcdefgh", b"ab"],
                                       append_zero_at_end=True)
    # "abcd" and "abcdefgh" are multiples of 4 and shorter than the longest key: one
    # zero slice added to each
    assert lens.tolist() == [2, 3, 3, 1]
    assert keys[0, 1] == 0
    assert hi == 3


def test_slice_order_is_byte_order():
    lines = sorted(set(_urls(2000, seed=5)))
    keys, lens, _, _ = W.dataset_keys(lines)
    tuples = [tuple(keys[i, :lens[i]].tolist()) for i in range(len(lens))]
    assert tuples == sorted(tuples)
    assert len(set(tuples)) == len(tuples)
