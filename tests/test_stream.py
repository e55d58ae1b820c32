"""Keyed streams: Philox4x32-10 known answers (Random123 kat_vectors) for
both the oracle's and the product's implementation, and agreement of the
product's stream header with the oracle on draws and permutations."""
import os
import subprocess

import numpy as np
import pytest

from oracle import Oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
KAT = [  # (counter, key) -> output, Philox4x32-10
    ([0, 0, 0, 0], [0, 0], [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]),
    ([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2, [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]),
    ([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0],
     [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]),
]


@pytest.fixture(scope="module")
def probe(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("probe") / "stream_probe")
    subprocess.run(["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "pyhea_amd", "csrc"),
                    os.path.join(ROOT, "tests", "native", "stream_probe.cpp"), "-o", exe],
                   check=True)
    return exe


def _run(exe, *args):
    return subprocess.run([exe, *map(str, args)], check=True, capture_output=True,
                          text=True).stdout.split()


@pytest.mark.parametrize("ctr,key,want", KAT)
def test_oracle_philox_kat(ctr, key, want):
    assert Oracle.philox(ctr, key).tolist() == want


@pytest.mark.parametrize("ctr,key,want", KAT)
def test_product_philox_kat(probe, ctr, key, want):
    assert [int(x) for x in _run(probe, "philox", *ctr, *key)] == want


@pytest.mark.parametrize("seed,it,ph,slot,n", [(1, 0, 2, 0, 4000), (2**40 + 5, 7, 3, 511, 8192),
                                               (9, 1, 2, 3, 7)])
def test_pairs_match_oracle(probe, seed, it, ph, slot, n):
    cnt = 300
    got = np.array(_run(probe, "pairs", seed, it, ph, slot, n, cnt), dtype=np.int64).reshape(-1, 2)
    words = Oracle.stream_words(seed, it, ph, slot, 0, 2 * cnt).astype(np.int64)
    assert np.array_equal(got, (words % n).reshape(-1, 2))
    assert words.max() < 2**31


@pytest.mark.parametrize("n,counts", [(4000, [1000] * 4), (8192, [2458, 2048, 1638, 1229, 819]),
                                      (7, [3, 4]), (2, [1, 1]), (500, [1, 1, 498])])
def test_permutation_matches_oracle(probe, n, counts):
    for slot in (0, 5):
        got = np.array(_run(probe, "perm", 31, 2, 1, slot, n, *counts), dtype=np.uint8)
        o = Oracle(np.zeros((n, 1), np.int32), [1], counts, [1.0], np.zeros((1, len(counts) ** 2)))
        assert np.array_equal(got, o.permute(31, 2, 1, slot))
        assert np.array_equal(np.bincount(got, minlength=len(counts)), counts)


def test_permutations_differ_by_key():
    o = Oracle(np.zeros((1000, 1), np.int32), [1], [500, 500], [1.0], np.zeros((1, 4)))
    a, b, c = o.permute(1, 0, 0, 0), o.permute(1, 0, 0, 1), o.permute(2, 0, 0, 0)
    assert not np.array_equal(a, b) and not np.array_equal(a, c)
    # roughly uniform: each half of the lattice gets about half of species 0
    assert abs(int((a[:500] == 0).sum()) - 250) < 60
