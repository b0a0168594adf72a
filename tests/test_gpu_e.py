"""GPU parity of e: digits bit-exact against the reference's own compute_e_host
output (tests/golden/ref_e.json), GMP, and the OEIS A001113 prefix."""
import hashlib
import json
import os
import re
import subprocess

import pytest

import _oracle as O
import cryptography_amd as E
from cryptography_amd import engine as EN

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
REF = json.load(open(os.path.join(HERE, "golden", "ref_e.json")))


@pytest.mark.parametrize("d", [50, 100, 200, 1000, 10000])
def test_e_matches_reference_output(d):
    s, st = EN.compute_e(d)
    assert s == REF[str(d)]["e"]
    assert st["n_terms"] == O.orc().orc_split_terms_needed(d + 20)


@pytest.mark.parametrize("d", [1, 2, 3, 19, 20, 21, 49])
def test_e_short_outputs_are_prefixes(d):
    """d digits ("2." and d - 1 truncated decimals) are the first d + 1 characters of the
    reference's 50-digit output: the string's length and end at the smallest sizes."""
    s, _ = EN.compute_e(d)
    assert s == REF["50"]["e"][:d + 1]


def test_e_100k_matches_reference_hash():
    s, _ = EN.compute_e(100000)
    assert len(s) == REF["100000"]["len"]
    assert hashlib.sha256(s.encode()).hexdigest() == REF["100000"]["sha256"]


def test_e_million_vs_gmp():
    d = 1000000
    s, st = EN.compute_e(d)
    want = O.gmp_e_digits(d)
    assert s == want[0] + "." + want[1:]


@pytest.mark.parametrize("d", [1000, 10000])
def test_e_plain_tree_path(d):
    s, _ = EN.compute_e(d, batched_tree=False)
    assert s == REF[str(d)]["e"]


def test_compute_e_cli_output_format():
    exe = os.path.join(ROOT, "cryptography_amd", "compute_e")
    out = subprocess.run([exe, "-d", "200"], capture_output=True, text=True, check=True).stdout
    # the reference's own extraction (ntt/scripts/test_compute_e.sh:39)
    m = re.search(r"^e = (2\.[0-9]+).*$", out, re.M)
    assert m and m.group(1) == REF["200"]["e"]
    assert out.startswith("compute_e: 200 digits, %d terms, 20 guard" % O.orc().orc_split_terms_needed(220))


@pytest.mark.parametrize("d", [10 ** 7, 10 ** 8])
def test_e_large_vs_gmp_hash(d):
    """e at 10^7 and 10^8 digits: SHA-256 of "2.ddd..." against GMP (tests/golden/e_digits_gmp.json)."""
    gold = json.load(open(os.path.join(HERE, "golden", "e_digits_gmp.json")))[str(d)]
    s, st = EN.compute_e(d)
    assert s[:64] == gold["head"] and s[-64:] == gold["tail"]
    assert hashlib.sha256(s.encode()).hexdigest() == gold["sha256"]
