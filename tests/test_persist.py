"""Filter images: save / load round trips and the checks on bad files.

A loaded filter must hold the saved filter's exact words (compared on the
host) and answer every query the same; a truncated or altered image, or a file
that is not an image, must be rejected. Header-only checks run on the CPU.
"""
from __future__ import annotations

import numpy as np
import pytest

from tests.helpers import make_config, random_dna, random_protein


def test_non_image_files_are_rejected_on_the_host(tmp_path):
    import ctypes

    from cusbf_amd import _lib
    from cusbf_amd.errors import CusbfError, ErrorCategory, check

    lib = _lib.load()
    junk = tmp_path / "junk.bin"
    junk.write_bytes(b"not an image at all" * 40)
    c = _lib.CConfig()
    bits = ctypes.c_uint64()
    with pytest.raises(CusbfError) as e:
        check(lib, lib.cusbf_image_config(str(junk).encode(), ctypes.byref(c), ctypes.byref(bits)))
    assert e.value.category == ErrorCategory.invalid_argument
    assert e.value.message == f"not a filter image: {junk}"
    with pytest.raises(CusbfError) as e:
        check(lib, lib.cusbf_image_config(str(tmp_path / "missing").encode(), ctypes.byref(c),
                                          ctypes.byref(bits)))
    assert e.value.category == ErrorCategory.io


@pytest.mark.gpu
@pytest.mark.parametrize("alphabet,k,s,m,h", [("dna", 31, 28, 16, 4), ("protein", 10, 8, 6, 4),
                                               ("custom_xyz", 12, 10, 6, 4), ("triplet", 7, 5, 4, 4)])
def test_save_load_round_trip(gpu, tmp_path, alphabet, k, s, m, h):
    import cusbf_amd as cs
    from tests.helpers import random_bytes_for

    cfg = make_config(k, s, m, h, alphabet)
    f = cs.Filter(cfg, 1 << 22)
    seq = random_bytes_for(alphabet, 200_000, 3, 0.001)
    f.insert_sequence(seq)
    path = str(tmp_path / "f.img")
    f.save(path)
    g = cs.Filter.load(path)
    assert g.filter_bits() == f.filter_bits() and g.num_shards() == f.num_shards()
    assert g.config.k == k and g.config.symbol_width == cfg.symbol_width
    assert (g.words() == f.words()).all()
    q = random_bytes_for(alphabet, 50_000, 4, 0.001) + seq[:50_000]
    assert np.array_equal(g.contains_sequence(q), f.contains_sequence(q))
    # the file itself, against the oracle: payload = the oracle's words, header
    # fields and checksum = the format's restatement over those words
    from oracle import image_oracle as im
    from tests.helpers import make_oracle

    o = make_oracle(k, s, m, h, alphabet)
    w = o.new_filter(f.num_shards())
    o.insert(seq, w)
    hdr, payload = im.read_image(path)
    assert hdr["magic"] == b"CUSBFIMG" and hdr["version"] == 1 and hdr["header_bytes"] == 336
    assert hdr["num_shards"] == f.num_shards() and hdr["filter_bits"] == f.filter_bits()
    assert hdr["word_count"] == w.size
    assert np.array_equal(payload, w)
    assert int(hdr["checksum"]) == im.words_checksum(w)


@pytest.mark.gpu
def test_bad_images_are_rejected(gpu, tmp_path):
    import cusbf_amd as cs
    from cusbf_amd.errors import CusbfError, ErrorCategory

    f = cs.Filter(make_config(31, 28, 16, 4), 1 << 20)
    f.insert_sequence(random_dna(100_000, 1))
    path = tmp_path / "f.img"
    f.save(str(path))
    data = path.read_bytes()
    assert len(data) == 336 + f.num_shards() * 32
    truncated = tmp_path / "t.img"
    truncated.write_bytes(data[:-100])
    with pytest.raises(CusbfError) as e:
        cs.Filter.load(str(truncated))
    assert e.value.category == ErrorCategory.io and "truncated" in e.value.message
    altered = bytearray(data)
    altered[336 + 12345] ^= 0x10
    bad = tmp_path / "a.img"
    bad.write_bytes(bytes(altered))
    with pytest.raises(CusbfError) as e:
        cs.Filter.load(str(bad))
    assert e.value.message == f"filter image checksum mismatch: {bad}"
    hdr = bytearray(data)
    hdr[8] = 9  # version
    badv = tmp_path / "v.img"
    badv.write_bytes(bytes(hdr))
    with pytest.raises(CusbfError) as e:
        cs.Filter.load(str(badv))
    assert e.value.category == ErrorCategory.invalid_argument
