"""Pins the oracle (CPU): the reference unit tests' own expectations, an independent
pure-Python model on tiny key universes, and the rebuild of the reference inputs.

Expectations are the closed forms asserted by test/unittest_masstree.cu:318-708
(hyoungjook/IndexinGPU); the reference ships no binary fixtures, so these are its
golden vectors.
"""
import itertools
import os
import random

import numpy as np
import pytest

import refinputs as R
import scenarios as S


@pytest.mark.parametrize("c", range(4))
def test_rebuilt_keys_match_cpp_generator(c):
    inp = R.TestingInput(c)
    f = R.fixture()
    lo, hi = f[f"cfg{c}_key_checksum"].astype(np.uint64)
    want = int(lo) | (int(hi) << 32)
    assert R.key_checksum(inp.flat("keys"), inp.lengths, inp.max_key_length) == want


def _vals_ok(v, l, exp_vals, exp_lens):
    assert (l == exp_lens).all()
    for i in range(len(l)):
        assert (v[i, :l[i]] == exp_vals[i, :l[i]]).all()


@pytest.mark.parametrize("c", range(4))
def test_find_exist_notexist_update(c):
    inp = R.TestingInput(c)
    out = S.sc_find_exist(S.OracleBackend(), inp)
    assert out["insert"].all()
    _vals_ok(out["find_values"], out["find_lens"], inp.values, inp.value_lengths)
    out = S.sc_find_notexist(S.OracleBackend(), inp)
    assert (out["find_lens"] == 0).all()
    out = S.sc_update(S.OracleBackend(), inp)
    assert out["update"].all()
    _vals_ok(out["find_values"], out["find_lens"], inp.values2, inp.value_lengths)


@pytest.mark.parametrize("c", range(4))
def test_erase_cases(c):
    inp = R.TestingInput(c)
    out = S.sc_erase_all(S.OracleBackend(), inp)
    assert out["erase"].all() and (out["find_lens"] == 0).all() and out["items"] == []
    out = S.sc_erase_none(S.OracleBackend(), inp)
    assert not out["erase"].any()
    _vals_ok(out["find_values"], out["find_lens"], inp.values, inp.value_lengths)
    out = S.sc_erase_all_insert_all(S.OracleBackend(), inp)
    assert (out["find1_lens"] == 0).all() and out["insert2"].all()
    _vals_ok(out["find_values"], out["find_lens"], inp.values, inp.value_lengths)
    dup = R.TestingInput(c, duplicate=True)
    out = S.sc_insert_twice_erase_all(S.OracleBackend(), dup)
    assert (out["find_lens"] == 0).all() and out["items"] == []
    h = (dup.num_keys + 1) // 2
    assert (out["insert_dup"][:h] == 1).all() and (out["insert_dup"][h:] == 0).all()
    dup = R.TestingInput(c, duplicate=True)
    out = S.sc_erase_all_twice(S.OracleBackend(), dup)
    assert (out["find_lens"] == 0).all() and out["items"] == []


@pytest.mark.parametrize("c", range(4))
def test_concurrent_mix_expectations(c):
    inp = R.TestingInput(c)
    mix = R.ConcurrentMix(inp)
    out = S.sc_concurrent_mix(S.OracleBackend(), inp, mix)
    ks = mix.num_keyset
    for i in range(len(mix.types)):
        dst = mix.dst_of[i]
        t = mix.types[i]
        if dst < 2 * ks:
            assert t == 0 and out["mix_results"][i] == (dst >= ks)
        elif dst < 4 * ks:
            assert t == 3 and out["mix_results"][i] == (dst < 3 * ks)
        elif dst < 8 * ks:
            assert t == 2
            if dst < 5 * ks:
                src = dst - 4 * ks
                assert out["mix_lens"][i] == inp.value_lengths[src]
                for s in range(inp.value_lengths[src]):
                    assert out["mix_values"][i, s] in (inp.values[src, s], inp.values2[src, s])
            elif dst >= 7 * ks:
                assert out["mix_lens"][i] == 0
        else:
            assert t == 1 and out["mix_results"][i] == (dst < 9 * ks)
    for i in range(4 * ks):
        live = i < ks or 2 * ks <= i < 3 * ks
        exp_len = inp.value_lengths[i] if live else 0
        assert out["find_lens"][i] == exp_len
        exp = inp.values2[i] if i < ks else inp.values[i]
        assert (out["find_values"][i, :exp_len] == exp[:exp_len]).all()


@pytest.mark.parametrize("c", range(4))
def test_range_query_expectations(c):
    inp = R.TestingInput(c)
    rin = R.RangeInput(inp, inp.num_keys // 2, 10)
    out = S.sc_range_query(S.OracleBackend(), inp, rin)
    mk, mv, mc = inp.max_key_length, inp.max_value_length, rin.max_count
    assert (out["counts"] == rin.counts).all()
    for q in range(rin.num_queries):
        for v in range(rin.counts[q]):
            src = rin.begin[q] + v
            slot = q * mc + v
            assert out["value_lengths"][slot] == inp.value_lengths[src]
            L = inp.value_lengths[src]
            assert (out["values"][slot * mv:slot * mv + L] == inp.values[src, :L]).all()
            assert out["key_lengths"][slot] == inp.lengths[src]
            K = inp.lengths[src]
            assert (out["keys"][slot * mk:slot * mk + K] == inp.keys[src, :K]).all()


# --------------------------------------------------------------- independent model
SLICES = [0, 1, 2, 0xFFFFFFFF]
UNIVERSE = [k for L in (1, 2, 3) for k in itertools.product(SLICES, repeat=L)]


def _pack(keys, maxk):
    a = np.zeros((len(keys), maxk), np.uint32)
    for i, k in enumerate(keys):
        a[i, :len(k)] = k
    return a, np.array([len(k) for k in keys], np.uint32)


@pytest.mark.parametrize("seed", range(6))
def test_oracle_against_python_model(seed):
    """Every key over slices {0,1,2,max} of length <= 3; random batches of every op,
    compared with a dict + sorted-tuple model (Python's tuple order is lexicographic
    with a proper prefix first)."""
    rnd = random.Random(seed)
    o = S.OracleBackend()
    model = {}
    maxk, maxv = 3, 3
    for _ in range(30):
        op = rnd.choice(["insert", "upsert", "update", "erase", "find", "scan", "mixed"])
        keys = rnd.sample(UNIVERSE, rnd.randint(1, 20))
        ka, kl = _pack(keys, maxk)
        vals = [tuple(rnd.randrange(1 << 32) for _ in range(rnd.randint(0, maxv))) for _ in keys]
        va = np.zeros((len(keys), maxv), np.uint32)
        for i, v in enumerate(vals):
            va[i, :len(v)] = v
        vl = np.array([len(v) for v in vals], np.uint32)
        if op in ("insert", "upsert"):
            r = o.insert(ka, kl, va, vl, maxk, maxv, upsert=(op == "upsert"))
            for i, k in enumerate(keys):
                assert r[i] == (op == "upsert" or k not in model)
                if op == "upsert" or k not in model:
                    model[k] = vals[i]
        elif op == "update":
            r = o.update(ka, kl, va, vl, maxk, maxv)
            for i, k in enumerate(keys):
                assert r[i] == (k in model)
                if k in model:
                    model[k] = vals[i]
        elif op == "erase":
            r = o.erase(ka, kl, maxk)
            for i, k in enumerate(keys):
                assert r[i] == (k in model)
                model.pop(k, None)
        elif op == "find":
            v, l = o.find(ka, kl, maxk, maxv)
            for i, k in enumerate(keys):
                exp = model.get(k, ())
                assert l[i] == len(exp) and tuple(v[i, :len(exp)]) == exp
        elif op == "scan":
            lim = rnd.randint(1, 8)
            upper = [rnd.choice(UNIVERSE) for _ in keys]
            ua, ul = _pack(upper, maxk)
            use_upper = rnd.random() < 0.5
            c, v, vlen, ok, okl = o.scan(ka, kl, maxk, lim, ua if use_upper else None,
                                         ul if use_upper else None, maxv)
            srt = sorted(model)
            for q, lo in enumerate(keys):
                exp = [k for k in srt if k >= lo and (not use_upper or k <= upper[q])][:lim]
                assert c[q] == len(exp)
                for j, k in enumerate(exp):
                    s = q * lim + j
                    assert okl[s] == len(k) and tuple(ok[s * maxk:s * maxk + len(k)]) == k
                    assert vlen[s] == len(model[k])
                    assert tuple(v[s * maxv:s * maxv + len(model[k])]) == model[k]
        else:
            types = np.array([rnd.choice([0, 1, 2, 3]) for _ in keys], np.uint8)
            v, l, r = o.mixed(types, ka, kl, va, vl, maxk, maxv)
            for i, k in enumerate(keys):
                t = types[i]
                if t == 0:
                    assert r[i] == (k not in model)
                    model.setdefault(k, vals[i])
                elif t == 1:
                    assert r[i] == (k in model)
                    model.pop(k, None)
                elif t == 2:
                    exp = model.get(k, ())
                    assert l[i] == len(exp) and tuple(v[i, :len(exp)]) == exp
                else:
                    assert r[i] == (k in model)
                    if k in model:
                        model[k] = vals[i]
        assert o.items() == sorted(model.items())


def test_planted_scan_mistake_is_caught():
    """A scan with an exclusive lower bound (a plausible slip) must disagree with the
    model on the tiny universe, so the model check above has teeth."""
    o = S.OracleBackend()
    keys = UNIVERSE[:40]
    ka, kl = _pack(keys, 3)
    va = np.zeros((len(keys), 1), np.uint32)
    vl = np.ones(len(keys), np.uint32)
    o.insert(ka, kl, va, vl, 3, 1)
    c, _, _, ok, okl = o.scan(ka, kl, 3, 1)
    first = [tuple(ok[q * 3:q * 3 + okl[q]].tolist()) for q in range(len(keys))]
    assert first == keys  # inclusive lower bound: each query returns its own key first
    srt = sorted(keys)
    exclusive = [next((k for k in srt if k > lo), None) for lo in keys]
    assert first != exclusive


# ------------------------------------------------------- truncated output rows
# The reference copies min(len, max) slices into an output row and reports the full
# stored length (include/suffix_node_subwarp.hpp:519-544 for values,
# include/gpu_masstree.hpp:227-240 and :250-395 for find and scan).  The model below
# stores values longer than the rows find, scan and mixed batches write into, and keys
# longer than a scan's output rows; every slot past min(len, max) must keep its fill.
FILL = 0xA5A5A5A5
SHORT = [k for k in UNIVERSE if len(k) <= 2]


def _check_truncation_model(o, seed):
    from oracle.oracle import Oracle  # noqa: F401  (o is an Oracle)
    rnd = random.Random(seed)
    model = {}
    maxk, maxv_in = 3, 5
    for _ in range(40):
        op = rnd.choice(["insert", "upsert", "find", "find", "scan", "scan", "mixed", "erase"])
        if op in ("insert", "upsert", "erase"):
            keys = rnd.sample(UNIVERSE, rnd.randint(1, 24))
            ka, kl = _pack(keys, maxk)
            if op == "erase":
                r = o.erase(ka.reshape(-1), maxk, kl, len(keys))
                for i, k in enumerate(keys):
                    assert r[i] == (k in model)
                    model.pop(k, None)
                continue
            vals = [tuple(rnd.randrange(1 << 32) for _ in range(rnd.randint(0, maxv_in)))
                    for _ in keys]
            va = np.zeros((len(keys), maxv_in), np.uint32)
            for i, v in enumerate(vals):
                va[i, :len(v)] = v
            vl = np.array([len(v) for v in vals], np.uint32)
            r = o.insert(ka.reshape(-1), maxk, kl, va.reshape(-1), maxv_in, vl, len(keys),
                         update_if_exists=(op == "upsert"))
            for i, k in enumerate(keys):
                assert r[i] == (op == "upsert" or k not in model)
                if op == "upsert" or k not in model:
                    model[k] = vals[i]
        elif op == "find":
            keys = rnd.sample(UNIVERSE, rnd.randint(1, 24))
            ka, kl = _pack(keys, maxk)
            mv = rnd.choice([1, 2, 3])
            v, l = o.find(ka.reshape(-1), maxk, kl, mv, len(keys), fill=FILL)
            v = v.reshape(len(keys), mv)
            for i, k in enumerate(keys):
                if k not in model:
                    assert l[i] == 0 and (v[i] == FILL).all()
                    continue
                exp = model[k]
                n = min(len(exp), mv)
                assert l[i] == len(exp)  # the full stored length, not n
                assert tuple(v[i, :n]) == exp[:n]
                assert (v[i, n:] == FILL).all()
        elif op == "scan":
            # output key rows of 2 slices: stored keys of length 3 are cut to 2
            mks = 2
            lows = rnd.sample(SHORT, rnd.randint(1, 8))
            la, ll = _pack(lows, mks)
            lim, mv = rnd.randint(1, 6), rnd.choice([1, 2, 3])
            c, v, vlen, ok, okl = o.scan(la.reshape(-1), ll, mks, lim, len(lows),
                                         max_value_length=mv, fill=FILL)
            srt = sorted(model)
            for q, lo in enumerate(lows):
                exp = [k for k in srt if k >= lo][:lim]
                assert c[q] == len(exp)
                for j in range(lim):
                    s = q * lim + j
                    if j >= len(exp):
                        assert okl[s] == FILL and vlen[s] == FILL
                        continue
                    k = exp[j]
                    nk, nv = min(len(k), mks), min(len(model[k]), mv)
                    assert okl[s] == len(k) and vlen[s] == len(model[k])
                    assert tuple(ok[s * mks:s * mks + nk]) == k[:nk]
                    assert (ok[s * mks + nk:(s + 1) * mks] == FILL).all()
                    assert tuple(v[s * mv:s * mv + nv]) == model[k][:nv]
                    assert (v[s * mv + nv:(s + 1) * mv] == FILL).all()
        else:
            # a mixed batch with rows of 2 value slices: finds of longer stored values
            # are cut, inserts over a stored zero-length value fail (the key exists)
            mv = 2
            keys = rnd.sample(UNIVERSE, rnd.randint(1, 24))
            ka, kl = _pack(keys, maxk)
            types = np.array([rnd.choice([0, 1, 2, 2, 3]) for _ in keys], np.uint8)
            vals = [tuple(rnd.randrange(1 << 32) for _ in range(rnd.randint(0, mv)))
                    for _ in keys]
            va = np.full((len(keys), mv), FILL, np.uint32)
            for i, vv in enumerate(vals):
                va[i, :len(vv)] = vv
            vl = np.array([len(vv) for vv in vals], np.uint32)
            v, l, r = o.mixed(types, ka.reshape(-1), maxk, kl, va.reshape(-1), mv, vl, len(keys))
            v = v.reshape(len(keys), mv)
            for i, k in enumerate(keys):
                t = types[i]
                if t == 0:
                    assert r[i] == (k not in model)
                    model.setdefault(k, vals[i])
                elif t == 1:
                    assert r[i] == (k in model)
                    model.pop(k, None)
                elif t == 2:
                    if k not in model:
                        assert l[i] == 0
                        continue
                    exp = model[k]
                    n = min(len(exp), mv)
                    assert l[i] == len(exp)
                    assert tuple(v[i, :n]) == exp[:n]
                    # slots past n keep what the request carried in
                    assert tuple(v[i, n:]) == tuple(va[i, n:])
                else:
                    assert r[i] == (k in model)
                    if k in model:
                        model[k] = vals[i]
        assert o.items() == sorted(model.items())


@pytest.mark.parametrize("seed", range(6))
def test_oracle_truncation_against_python_model(seed):
    from oracle.oracle import Oracle
    _check_truncation_model(Oracle(), seed)


def test_zero_length_value_blocks_insert():
    """gpu_masstree.hpp:607-617 + :231: a zero-length value is stored (LONGVAL of length
    0), reads back as length 0, and a later insert of the key fails in a plain insert
    and in a mixed batch alike; erase then succeeds."""
    from oracle.oracle import Oracle
    o = Oracle()
    ka, kl = _pack([(7, 8)], 2)
    z = np.zeros(2, np.uint32)
    assert o.insert(ka.reshape(-1), 2, kl, z, 2, np.zeros(1, np.uint32), 1)[0] == 1
    v, l = o.find(ka.reshape(-1), 2, kl, 2, 1, fill=FILL)
    assert l[0] == 0 and (v == FILL).all()
    assert o.insert(ka.reshape(-1), 2, kl, z + 5, 2, np.ones(1, np.uint32), 1)[0] == 0
    _, _, r = o.mixed(np.array([0], np.uint8), ka.reshape(-1), 2, kl, z + 5, 2,
                      np.ones(1, np.uint32), 1)
    assert r[0] == 0
    assert o.items() == [((7, 8), ())]
    assert o.erase(ka.reshape(-1), 2, kl, 1)[0] == 1 and o.items() == []


@pytest.mark.parametrize("plant", [1, 2, 3, 4])
def test_planted_truncation_mistakes_are_caught(plant, tmp_path):
    """Builds the oracle with one planted mistake (ORC_PLANT: a truncated length
    reported by find or by scan, one slice too few copied) and checks that the model
    above rejects it."""
    import subprocess
    from oracle.oracle import Oracle
    src = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle",
                       "masstree_oracle.cpp")
    lib = str(tmp_path / f"liboracle_plant{plant}.so")
    subprocess.check_call(["g++", "-O1", "-std=c++17", "-fPIC", "-shared",
                           f"-DORC_PLANT={plant}", "-o", lib, src])
    caught = 0
    for seed in range(6):
        try:
            _check_truncation_model(Oracle(lib_path=lib), seed)
        except AssertionError:
            caught += 1
    assert caught > 0
