"""The tuning-file reader (ntt_engine_autotune_load) on the CPU: it only sets
the engine's host-side parameters, so no GPU is needed.  Valid files set the
threshold, tile batch and pass caps; extra keys (the recorded crossover
samples) are skipped; malformed or out-of-range files are rejected and change
nothing.  Also the committed tuning file parses."""
import ctypes
import json
import os

import pytest

import cryptography_amd as E

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
             {l: L.ntt_engine_get_pass_cap(l) for l in range(1, 31)})
    yield L
    L.ntt_engine_set_threshold(saved[0])
    L.ntt_engine_set_tile_batch.argtypes = [ctypes.c_int]
    L.ntt_engine_set_tile_batch(saved[1])
    L.ntt_engine_set_pass_cap.argtypes = [ctypes.c_int, ctypes.c_int]
    for l, c in saved[2].items():
        L.ntt_engine_set_pass_cap(l, 0 if c == 10 else c)


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
