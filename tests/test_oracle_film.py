"""The film file of ImageFilm (YAF_FILMv4_0_0), restated in oracle/pyoracle.py
(src/render/imagefilm.cc:910-1018 load, :1099-1176 save).

Pinned byte by byte on a 2 x 1 film written out by hand from the reference's
save sequence: NUL-terminated tag, ten native-order int32, the weights, then the
RGBA sums, each row-major from y = 0.
"""
import struct

import numpy as np

from oracle import pyoracle


def test_two_pixel_film_bytes(tmp_path):
    weight = np.array([[0.5, 2.0]], np.float32)
    rgba = np.array([[[1, 2, 3, 4], [5, 6, 7, 8]]], np.float32)
    p = tmp_path / "a.film"
    pyoracle.write_film(p, rgba, weight, computer_node=3, base_sampling_offset=7, sampling_offset=11)
    expect = (b"YAF_FILMv4_0_0\x00"
              + struct.pack("=10i", 3, 7, 11, 2, 1, 0, 1, 0, 0, 1)
              + struct.pack("=2f", 0.5, 2.0)
              + struct.pack("=8f", 1, 2, 3, 4, 5, 6, 7, 8))
    assert p.read_bytes() == expect
    assert len(expect) == 15 + 40 + 4 * 2 + 16 * 2


def test_row_major_from_the_top(tmp_path):
    h, w = 3, 4
    weight = np.arange(h * w, dtype=np.float32).reshape(h, w)
    rgba = np.arange(4 * h * w, dtype=np.float32).reshape(h, w, 4) + 100
    p = tmp_path / "b.film"
    pyoracle.write_film(p, rgba, weight)
    data = p.read_bytes()
    # weight of pixel (x, y) sits at 55 + 4 (y w + x); colour at 55 + 4 h w + 16 (y w + x)
    for y in range(h):
        for x in range(w):
            assert struct.unpack_from("=f", data, 55 + 4 * (y * w + x))[0] == weight[y, x]
            assert struct.unpack_from("=4f", data, 55 + 4 * h * w + 16 * (y * w + x)) == tuple(rgba[y, x])
    got = pyoracle.read_film(p, w, h)
    assert got is not None
    assert np.array_equal(got[0], rgba) and np.array_equal(got[1], weight) and got[2] == (0, 0, 0)


def test_load_checks_refuse_other_films(tmp_path):
    p = tmp_path / "c.film"
    pyoracle.write_film(p, np.zeros((2, 3, 4), np.float32), np.zeros((2, 3), np.float32))
    assert pyoracle.read_film(p, 3, 2) is not None
    assert pyoracle.read_film(p, 2, 3) is None  # size differs
    bad = tmp_path / "d.film"
    bad.write_bytes(b"YAF_FILMv3_0_0\x00" + p.read_bytes()[15:])
    assert pyoracle.read_film(bad, 3, 2) is None  # tag differs
