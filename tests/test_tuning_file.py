"""The tuning-file reader (ntt_engine_autotune_load) and the pass-plan calls on
the CPU: they only set the engine's host-side parameters, so no GPU is needed.
Valid files set the threshold, tile batch, pass caps and pass splits; extra
keys (the recorded crossover and split samples) are skipped; malformed or
out-of-range files are rejected and change nothing.  The committed tuning file
parses and its splits are the fastest it recorded."""
import ctypes
import json
import os

import pytest

import cryptography_amd as E
from cryptography_amd import engine as EN

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _lib():
    L = E.lib()
    L.ntt_engine_autotune_load.argtypes = [ctypes.c_char_p]
    L.ntt_engine_get_pass_cap.argtypes = [ctypes.c_int]
    return L


@pytest.fixture
def restore():
    L = _lib()
    saved = (L.ntt_engine_get_threshold(), L.ntt_engine_get_tile_batch(),
             {l: L.ntt_engine_get_pass_cap(l) for l in range(1, 31)},
             {l: EN.get_plan(l) for l in range(1, 31)})
    yield L
    L.ntt_engine_set_threshold(saved[0])
    L.ntt_engine_set_tile_batch.argtypes = [ctypes.c_int]
    L.ntt_engine_set_tile_batch(saved[1])
    L.ntt_engine_set_pass_cap.argtypes = [ctypes.c_int, ctypes.c_int]
    for l, c in saved[2].items():
        L.ntt_engine_set_pass_cap(l, 0 if c == 10 else c)    # also clears the split
    for l, p in saved[3].items():
        if EN.get_plan(l) != p:
            EN.set_pass_split(l, p)


def _load(L, tmp_path, text):
    p = tmp_path / "t.json"
    p.write_text(text)
    return L.ntt_engine_autotune_load(str(p).encode())


def test_valid_file_with_extra_keys(restore, tmp_path):
    L = restore
    doc = {"threshold": 32, "tile_batch": 2, "pass_cap": {"20": 8, "27": 9},
           "crossover_us": [[8, 0.1, 0.2, 0.3], [16, 0.2, 0.2, 0.4]], "note": "x"}
    assert _load(L, tmp_path, json.dumps(doc)) == 0
    assert L.ntt_engine_get_threshold() == 32
    assert L.ntt_engine_get_tile_batch() == 2
    assert L.ntt_engine_get_pass_cap(20) == 8 and L.ntt_engine_get_pass_cap(27) == 9
    assert L.ntt_engine_get_pass_cap(21) == 10          # not listed: the default


def test_file_without_optional_keys_keeps_them(restore, tmp_path):
    L = restore
    assert _load(L, tmp_path, '{"threshold": 16, "tile_batch": 4, "pass_cap": {"22": 9}}') == 0
    assert _load(L, tmp_path, '{"threshold": 48}') == 0
    assert L.ntt_engine_get_threshold() == 48
    assert L.ntt_engine_get_tile_batch() == 4
    assert L.ntt_engine_get_pass_cap(22) == 9


@pytest.mark.parametrize("text", [
    '', '[]', '{"threshold": }', '{"tile_batch": 2}', '{"threshold": 500}', '{"threshold": 8, "tile_batch": 0}',
    '{"threshold": 8, "pass_cap": {"27": 3}}', '{"threshold": 8, "pass_cap": {"31": 9}}',
    '{"threshold": 8, "pass_cap": [1, 2]}', '{"threshold": 8 "tile_batch": 2}',
])
def test_bad_files_are_rejected_and_change_nothing(restore, tmp_path, text):
    L = restore
    before = (L.ntt_engine_get_threshold(), L.ntt_engine_get_tile_batch(), L.ntt_engine_get_pass_cap(27))
    assert _load(L, tmp_path, text) == 4                 # NTT_ERR_ARG
    assert (L.ntt_engine_get_threshold(), L.ntt_engine_get_tile_batch(), L.ntt_engine_get_pass_cap(27)) == before


def test_committed_tuning_file_parses(restore):
    path = os.path.join(ROOT, "cryptography_amd", "tuning_mi355x.json")
    if not os.path.exists(path):
        pytest.skip("no committed tuning file")
    L = restore
    assert L.ntt_engine_autotune_load(path.encode()) == 0
    doc = json.load(open(path))
    assert L.ntt_engine_get_threshold() == doc["threshold"]
    for l, split in doc.get("pass_split", {}).items():
        assert EN.get_plan(int(l)) == tuple(split)
        # the chosen split is the fastest one the tuner recorded for the length
        tried = doc.get("split_us", {}).get(l)
        if tried:
            assert min(tried, key=lambda c: c[1])[0] == split


def _plan_by_cap(l, cap):
    """The near-equal split a pass cap gives (plan_passes in ntt_engine.hip)."""
    n = -(-l // cap)
    base, extra = divmod(l, n)
    return tuple(base + (1 if i >= n - extra else 0) for i in range(n))


def test_default_plans_follow_the_cap(restore):
    for l in range(1, 31):
        assert EN.get_plan(l) == _plan_by_cap(l, 10)


def test_pass_split_file_sets_plans(restore, tmp_path):
    L = restore
    doc = {"threshold": 32, "tile_batch": 1, "pass_cap": {"20": 8},
           "pass_split": {"21": [6, 6, 9], "14": [8, 6], "27": [9, 9, 9]},
           "split_us": {"21": [[[7, 7, 7], 253.3], [[6, 6, 9], 209.5]]}}
    assert _load(L, tmp_path, json.dumps(doc)) == 0
    assert EN.get_plan(21) == (6, 6, 9) and EN.get_plan(14) == (8, 6) and EN.get_plan(27) == (9, 9, 9)
    assert EN.get_plan(20) == _plan_by_cap(20, 8)         # a cap, no split: near-equal passes of 2^8 at most
    assert EN.get_plan(22) == _plan_by_cap(22, 10)
    # a file with caps only clears the splits of the lengths it lists
    assert _load(L, tmp_path, '{"threshold": 32, "pass_cap": {"21": 9}}') == 0
    assert EN.get_plan(21) == _plan_by_cap(21, 9) and EN.get_plan(14) == (8, 6)
    # a file with splits replaces all splits
    assert _load(L, tmp_path, '{"threshold": 32, "pass_split": {"15": [9, 6]}}') == 0
    assert EN.get_plan(15) == (9, 6) and EN.get_plan(14) == _plan_by_cap(14, 10)


@pytest.mark.parametrize("split", [
    '{"21": [6, 6, 8]}', '{"21": [11, 10]}', '{"21": [6, 6, 9, 0]}', '{"31": [10, 10, 11]}', '{"21": []}',
    '{"21": [6.5, 5.5, 9]}', '{"21": 21}', '{"2": [1, 1, 0]}', '{"24": [3, 3, 3, 3, 3, 3, 3, 1, 2]}',
    '{"12": [2, 10]}', '{"12": [10, 2]}', '{"13": [1, 6, 6]}',
])
def test_bad_pass_splits_are_rejected_and_change_nothing(restore, tmp_path, split):
    L = restore
    before = [EN.get_plan(l) for l in (2, 21, 24, 31)]
    assert _load(L, tmp_path, '{"threshold": 8, "pass_split": %s}' % split) == 4
    assert [EN.get_plan(l) for l in (2, 21, 24, 31)] == before


def test_set_pass_split_checks_and_cap_clears_it(restore):
    L = restore
    EN.set_pass_split(24, (9, 9, 6))
    assert EN.get_plan(24) == (9, 9, 6)
    for bad in ((9, 9, 5), (11, 13), (0, 24), tuple([3] * 9), (2, 10, 10, 2), (10, 10, 2, 2)):
        with pytest.raises(EN.NttError):
            EN.set_pass_split(24, bad)
        assert EN.get_plan(24) == (9, 9, 6)
    L.ntt_engine_set_pass_cap.argtypes = [ctypes.c_int, ctypes.c_int]
    assert L.ntt_engine_set_pass_cap(24, 9) == 0          # the latest setting decides
    assert EN.get_plan(24) == _plan_by_cap(24, 9)
    EN.set_pass_split(24, (6, 9, 9))
    EN.set_pass_split(24, ())
    EN.set_pass_split(2, (2,))            # one pass may be shorter than 2^3
    assert EN.get_plan(2) == (2,)
    EN.set_pass_split(2, ())
    assert EN.get_plan(24) == _plan_by_cap(24, 9)
    assert L.ntt_engine_set_pass_cap(24, 0) == 0
