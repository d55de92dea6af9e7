"""C-ABI checks that need no GPU: the library loads, exports every symbol the
public header declares, and answers the size/validation calls the reference
answers without touching pixel data."""
import ctypes
import json
import os
import re

import pytest

import texpresso_amd as tx

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "texpresso_hip.h")
GOLDEN = os.path.join(ROOT, "tests", "golden")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(tx_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_reference_surface():
    names = declared_functions()
    for want in ("tx_compress", "tx_decompress", "tx_compress_block_masked", "tx_decompress_block",
                 "tx_block_size", "tx_compressed_size", "tx_num_blocks", "tx_params_default"):
        assert want in names


def test_library_exports_every_declared_symbol():
    L = ctypes.CDLL(tx.library_path())
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert not missing, missing


def test_storage_requirements():
    kat = json.load(open(os.path.join(GOLDEN, "kat.json")))
    L = tx.lib()
    for fmt, w, h, size in kat["compressed_size"]:
        assert L.tx_compressed_size(fmt, w, h) == size
        assert tx.Format(fmt).compressed_size(w, h) == size
    assert L.tx_num_blocks(0) == 0 and L.tx_num_blocks(1) == 1 and L.tx_num_blocks(4) == 1 and L.tx_num_blocks(5) == 2
    assert L.tx_block_size(7) == 0


def test_params_default_matches_reference():
    p = tx.lib().tx_params_default()
    assert p.algorithm == tx.Algorithm.ClusterFit
    assert [round(w, 4) for w in p.weights] == [0.2126, 0.7152, 0.0722]
    assert p.weigh_colour_by_alpha == 0


def test_device_lut_source_matches_reference_tables():
    out = (ctypes.c_uint8 * 6144)()
    tx.lib().tx_single_lut(out)
    assert bytes(out) == open(os.path.join(GOLDEN, "single_lut.bin"), "rb").read()


@pytest.mark.parametrize("call,code", [
    (lambda: tx.Format.Bc1.compress(bytes(64), 0, 4, tx.Params(), bytearray(8)), -2),
    (lambda: tx.Format.Bc1.compress(bytes(60), 4, 4, tx.Params(), bytearray(8)), -3),
    (lambda: tx.Format.Bc1.compress(bytes(64), 4, 4, tx.Params(), bytearray(7)), -4),
    (lambda: tx.Format.Bc1.compress(bytes(64), 4, 4, tx.Params(), bytearray(12)), -5),
    (lambda: tx.Format.Bc3.decompress(bytes(8), 4, 4, bytearray(64)), -3),
    (lambda: tx.Format.Bc3.decompress(bytes(16), 4, 4, bytearray(63)), -4),
])
def test_reference_panics_become_error_codes(call, code):
    # validated before any device work, so this runs without a GPU
    with pytest.raises(tx.TexpressoError) as e:
        call()
    assert e.value.code == code


def test_invalid_algorithm_rejected():
    p = tx.Params()
    c = p._c()
    c.algorithm = 5
    rc = tx.lib().tx_compress(0, ctypes.c_char_p(bytes(64)), 64, 4, 4, ctypes.byref(c),
                              ctypes.create_string_buffer(8), 8)
    assert rc == -1
