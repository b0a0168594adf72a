"""GPU parity of the product tree: the level-batched device tree and the
reference-API depth-first path against the reference's own binary_split
cells (tests/golden/ref_arith.json) and GMP binary splitting."""
import json
import os

import numpy as np
import pytest

import _oracle as O
import cryptography_amd as E
from cryptography_amd import engine as EN

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def as_int(x):
    return O.limbs_to_int(x)


def test_tree_matches_reference_cells():
    d = json.load(open(os.path.join(GOLD, "ref_arith.json")))
    for c in d["binary_split"]:
        a, b = c["a"], c["b"]
        P, B = EN.split_tree(a, b)
        assert hex(as_int(P)) == c["P"], (a, b)
        assert hex(as_int(B)) == c["B"], (a, b)


def test_reference_api_paths_match_cells():
    d = json.load(open(os.path.join(GOLD, "ref_arith.json")))
    for c in d["binary_split"][:20]:
        for batched in (False, True):
            P, B = EN.binary_split_ref_api(c["a"], c["b"], batched=batched)
            assert hex(as_int(P)) == c["P"] and hex(as_int(B)) == c["B"], (c["a"], c["b"], batched)


@pytest.mark.parametrize("n_terms", [17, 33, 1000, 26535, 215341])
def test_tree_vs_gmp(n_terms):
    P, B = EN.split_tree(1, n_terms + 1)
    Pg, Bg = O.binary_split(1, n_terms + 1, use_gmp=True)
    assert np.array_equal(P, Pg) and np.array_equal(B, Bg)


def test_tree_graph_replay_and_eager_agree():
    P1, B1 = EN.split_tree(1, 50001, use_graph=True, runs=3)
    P2, B2 = EN.split_tree(1, 50001, use_graph=False)
    assert np.array_equal(P1, P2) and np.array_equal(B1, B2)


def test_plain_path_vs_batched_tree():
    P1, B1 = EN.binary_split_ref_api(1, 3001, batched=False)
    P2, B2 = EN.split_tree(1, 3001)
    assert np.array_equal(P1, P2) and np.array_equal(B1, B2)
