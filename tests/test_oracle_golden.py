"""The oracle's generator and RNG restatements against golden vectors produced
by the reference's own generator headers and lcg.h (tests/golden/make_golden.py
via oracle/_ref/ref_harness), and against closed-form known answers."""
import hashlib
import json
import os
import subprocess
import tempfile

import numpy as np
import pytest

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    return np.load(os.path.join(GOLD, name), allow_pickle=False)


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def test_sitmo_words_match_reference(orc):
    np.testing.assert_array_equal(orc.sitmo_words(0, 256), load("sitmo_seed0_256.npy"))
    np.testing.assert_array_equal(orc.sitmo_words(12345, 64), load("sitmo_seed12345_64.npy"))


def test_lcg_matches_reference_and_closed_form(orc):
    np.testing.assert_array_equal(orc.lcg_outputs(0, 64), load("lcg_step0_64.npy"))
    np.testing.assert_array_equal(orc.lcg_outputs(1000, 8), load("lcg_step1000_8.npy"))
    # lcg(0) starts from state 1; x -> 6364136223846793005 x + 1 mod 2^64
    x, out = 1, []
    for _ in range(6):
        x = (6364136223846793005 * x + 1) % (1 << 64)
        out.append(x)
    assert out[:3] == [6364136223846793006, 13885033948157127959, 14678909342070756876]
    np.testing.assert_array_equal(orc.lcg_outputs(0, 6), np.array(out, np.uint64))
    assert [v % (1 << 24) for v in out] == [9797422, 2436375, 4871692, 12391453, 1567770, 7825043]
    # skip-ahead: lcg(k)'s first output is the (k+1)-th output of lcg(0)
    seq = orc.lcg_outputs(0, 40)
    for k in (1, 7, 33):
        assert orc.lcg_outputs(k, 1)[0] == seq[k]


def test_rmat_raw_and_dedup_match_reference_s8(orc):
    raw = load("rmat_s8_ef16_raw.npy")
    dd = load("rmat_s8_ef16_dedup.npy")
    s, d = orc.rmat_raw(1 << 8, 16 << 8)
    np.testing.assert_array_equal(np.stack([s, d], 1), raw)
    s2, d2 = orc.flip_sort_dedup(s, d)
    np.testing.assert_array_equal(np.stack([s2, d2], 1), dd)
    assert (raw[:, 0] != raw[:, 1]).all()


@pytest.mark.parametrize("key", ["rmat_s12_ef16_seed0_a0.57", "rmat_s14_ef8_seed7_a0.45",
                                 "rmat_s16_ef16_seed0_a0.57"])
def test_rmat_digests_match_reference(orc, key):
    with open(os.path.join(GOLD, "rmat_digests.json")) as f:
        g = json.load(f)[key]
    a, b, c, d = g["abcd"]
    s, t = orc.rmat_raw(1 << g["scale"], g["edge_factor"] << g["scale"], a, b, c, d, g["seed"])
    raw = np.stack([s, t], 1)
    assert raw[:4].tolist() == g["raw_head"]
    assert sha(raw) == g["raw_sha256"]
    s2, t2 = orc.flip_sort_dedup(s, t)
    assert len(s2) == g["dedup_m"]
    assert sha(np.stack([s2, t2], 1)) == g["dedup_sha256"]


def test_oracle_matches_live_reference_harness(orc):
    """When the reference-built harness is present, compare a fresh case."""
    if not os.path.exists(orc.REF_HARNESS):
        pytest.skip("oracle/_ref/ref_harness not built")
    with tempfile.TemporaryDirectory() as t:
        raw_p, dd_p = os.path.join(t, "r"), os.path.join(t, "d")
        subprocess.run([orc.REF_HARNESS, "rmat", "1024", "5000", "0.5", "0.2", "0.2", "0.1", "99",
                        raw_p, dd_p], check=True)
        raw = np.fromfile(raw_p, np.int64).reshape(-1, 2)
        s, d = orc.rmat_raw(1024, 5000, 0.5, 0.2, 0.2, 0.1, 99)
        np.testing.assert_array_equal(np.stack([s, d], 1), raw)


def test_rmat_self_loop_reroll_exercised(orc):
    """Small scales have many self-loops before the re-roll; none survive and
    the result still matches the reference digest (s8 fixture)."""
    s, d = orc.rmat_raw(1 << 4, 16 << 4)
    assert (s != d).all()


def test_vperm_is_a_bijection(orc):
    for nv in (1 << 10, 1000, 4099):
        img = sorted(orc.vperm(x, nv, 3) for x in range(nv))
        assert img == list(range(nv))
    # different seeds give different permutations
    a = [orc.vperm(x, 1 << 12, 0) for x in range(64)]
    b = [orc.vperm(x, 1 << 12, 1) for x in range(64)]
    assert a != b


def test_make_rmat_is_relabelled_shuffled_dedup(orc):
    nv, s, d = orc.make_rmat(10, 16, seed=5)
    rs, rd = orc.rmat_raw(nv, 16 * nv, seed=5)
    ds, dd = orc.flip_sort_dedup(rs, rd)
    assert len(s) == len(ds)
    perm = np.array([orc.vperm(x, nv, 5) for x in range(nv)])
    a = np.sort(np.minimum(perm[ds], perm[dd]) * nv + np.maximum(perm[ds], perm[dd]))
    b = np.sort(np.minimum(s, d) * nv + np.maximum(s, d))
    np.testing.assert_array_equal(a, b)
    # the order is the shuffle: sorted by shuffle_key of the dedup index
    keys = [orc.shuffle_key(j, 5) for j in range(len(ds))]
    order = np.argsort(np.array(keys, np.uint64), kind="stable")
    np.testing.assert_array_equal(s, perm[ds][order])
    np.testing.assert_array_equal(d, perm[dd][order])


def test_er_and_lattice_generators(orc):
    nv, s, d = orc.make_er(3000, 20000, seed=2)
    assert (s != d).all() and s.max() < nv and d.max() < nv
    assert len(set(zip(np.minimum(s, d).tolist(), np.maximum(s, d).tolist()))) == len(s)
    nv, s, d = orc.make_lattice(7, 9, seed=1)
    assert nv == 63 and len(s) == 7 * 8 + 6 * 9
    pairs = set(zip(s.tolist(), d.tolist()))
    assert (0, 1) in pairs and (0, 9) in pairs and (61, 62) in pairs and (53, 62) in pairs
