"""The film output path on the CPU: the oracle's restatements of Film::Save's
conversion (src/Camera/film.cpp:152-203), SaveBmp's header and LoadBmp's
reading rule (src/Kernel/bmp.cpp) and FirefliesFiltering
(src/Scene/scene.cpp:261-431), pinned by closed forms and hand-made images;
the library's host BMP functions against them.  No GPU calls."""
import struct

import numpy as np
import pytest


def test_bmp_header_bytes_by_hand(rumi, oracle):
    want = struct.pack("<2sIHHIIIIHHIIIIII", b"BM", 54 + 4 * 3 * 4, 0, 0, 54, 40, 4, 3, 1, 32, 0,
                       4 * 3 * 4, 0, 0, 0, 0)
    assert len(want) == 54
    assert oracle.bmp_header(4, 3, 4).tobytes() == want
    lib_file = rumi.bmp_bytes(np.zeros((3, 4, 4), np.uint8), 4, 3, 4)
    assert lib_file[:54] == want and len(lib_file) == 54 + 48
    assert oracle.bmp_header(1200, 1000, 3).tobytes() == rumi.bmp_bytes(
        np.zeros((1000, 1200, 3), np.uint8), 1200, 1000, 3)[:54]


def test_bmp_read_follows_loadbmp(rumi):
    rng = np.random.default_rng(1)
    pix = rng.integers(0, 256, (5, 8, 4), dtype=np.uint8)
    back, w, h, b = rumi.bmp_read(rumi.bmp_bytes(pix, 8, 5, 4))
    assert (w, h, b) == (8, 5, 4)
    np.testing.assert_array_equal(back, pix)
    # width 6 (not a multiple of 4): each row is followed by 6 % 4 = 2
    # skipped bytes in the file LoadBmp reads
    pix3 = rng.integers(0, 256, (3, 6, 3), dtype=np.uint8)
    hdr = rumi.bmp_bytes(np.zeros(0, np.uint8), 6, 3, 3)[:54]
    body = b"".join(pix3[y].tobytes() + b"\xee\xee" for y in range(3))
    back3, w3, h3, b3 = rumi.bmp_read(hdr + body)
    assert (w3, h3, b3) == (6, 3, 3)
    np.testing.assert_array_equal(back3, pix3)


def test_encode_closed_forms(oracle):
    f = np.float32([[[0.0, 1.0, 0.5], [2.0, -1.0, np.inf]]])  # R, G, B means
    out = oracle.film_encode(f, False, 1, 2.2, 4)
    # bytes are B, G, R, 0: 0.5 ** (1 / 2.2) * 255 = 186.08; 1 -> 255; 0 -> 0;
    # 2 -> clamp 255; -1 -> NaN -> 0; inf -> 255
    np.testing.assert_array_equal(out[0, 0], [186, 255, 0, 0])
    np.testing.assert_array_equal(out[0, 1], [255, 0, 255, 0])
    # the reference's sums: B, G, R order, divided by (int)sqrt(n)^2
    s = np.float32([[[16.0, 8.0, 4.0]]])
    np.testing.assert_array_equal(oracle.film_encode(s, True, 16, 1.0, 3)[0, 0], [255, 127, 63])
    # 32 samples: sampling depth 5, so the sums are divided by 25, not 32
    s25 = np.float32([[[25.0, 12.5, 0.0]]])
    np.testing.assert_array_equal(oracle.film_encode(s25, True, 32, 1.0, 3)[0, 0], [255, 127, 0])


def test_encode_matches_float64_away_from_boundaries(oracle):
    rng = np.random.default_rng(2)
    v = rng.uniform(0, 1.2, (256, 256, 3)).astype(np.float32)
    for g in (1.0, 2.2):
        out = oracle.film_encode(v, False, 1, g, 3).astype(np.int64)
        e = np.float64(np.float32(1.0) / np.float32(g))   # Save's 1.0f / _gamma
        x = np.clip(np.power(v.astype(np.float64), e) * 255, 0, 255)
        want = np.floor(x).astype(np.int64)[..., ::-1]
        safe = ((np.abs(x - np.round(x)) > 1e-3) | (x == 255))[..., ::-1]
        np.testing.assert_array_equal(out[safe], want[safe])
        assert safe.mean() > 0.99


def hand_image(w, h, fill=(10, 20, 30)):
    return np.tile(np.uint8(fill), (h, w, 1))


def test_fireflies_interior_by_hand(oracle):
    # buffer row r holds image row h - 1 - r; the interior pixel (2, 2) is
    # buffer (2, 2) for h = 5
    img = hand_image(5, 5)
    img[2, 2] = (255, 40, 40)
    img[1, 1] = (100, 100, 100)                   # one of its 8 neighbours
    out = oracle.fireflies(img)
    want = img.copy()
    want[2, 2] = [(7 * 10 + 100) // 8, (7 * 20 + 100) // 8, (7 * 30 + 100) // 8]
    np.testing.assert_array_equal(out, want)
    img[3, 3] = (255, 255, 255)                   # a white neighbour: kept
    out2 = oracle.fireflies(img)
    np.testing.assert_array_equal(out2[2, 2], [255, 40, 40])
    # the white pixel itself has no white neighbour: a firefly, replaced by
    # the mean of its 8 neighbours ((255, 40, 40) and seven (10, 20, 30))
    np.testing.assert_array_equal(out2[3, 3], [(255 + 70) // 8, (40 + 140) // 8, (40 + 210) // 8])


def test_fireflies_left_edge_gathers_by_the_references_order(oracle):
    # image pixel (x=0, y=2) of a 4 x 5 image, buffer row 5-1-2 = 2: n_copy =
    # 5 of the gathered right, up-left, up, up-right, down-left, down, ...;
    # "up-left" of x = 0 is one pixel before buffer row 1's start, the last
    # pixel of buffer row 0, and "down-left" the last pixel of buffer row 2
    img = hand_image(4, 5, (0, 0, 0))
    img[2, 0] = (255, 0, 0)
    img[2, 1] = (10, 0, 0)      # right
    img[1, 3] = (99, 0, 0)      # not gathered
    img[0, 3] = (20, 0, 0)      # buffer row 0, last pixel (up-left, wrapped)
    img[1, 0] = (30, 0, 0)      # up
    img[1, 1] = (40, 0, 0)      # up-right
    img[2, 3] = (50, 0, 0)      # buffer row 2, last pixel (down-left, wrapped)
    img[3, 0] = (200, 0, 0)     # down: sixth, not among the first 5
    out = oracle.fireflies(img)
    np.testing.assert_array_equal(out[2, 0], [(10 + 20 + 30 + 40 + 50) // 5, 0, 0])
    assert (out[np.arange(5) != 2] == img[np.arange(5) != 2]).all()


def test_fireflies_leaves_unsaturated_images_alone(oracle):
    rng = np.random.default_rng(3)
    img = rng.integers(0, 255, (37, 41, 4), dtype=np.uint8)
    out = oracle.fireflies(img)
    want = img.copy()
    want[..., 3] = 0
    np.testing.assert_array_equal(out, want)


@pytest.fixture(scope="module")
def earth_bmp():
    import gzip
    import os
    here = os.path.dirname(os.path.abspath(__file__))
    return gzip.open(os.path.join(here, "golden", "textures", "earth.bmp.gz")).read()


@pytest.mark.parametrize("gamma", [1.0, 2.2])
def test_image_texture_from_the_references_bmp(rumi, oracle, earth_bmp, gamma):
    """ImageTexture("earth.bmp", gamma): LoadBmp's pixels, then byte / 255
    (to the power gamma); the library's texels equal the oracle's bit for
    bit and the float64 values to rounding, R, G, B from B, G, R bytes."""
    pix, w, h, b = rumi.bmp_read(earth_bmp)
    assert (w, h, b) == (1000, 500, 3)
    np.testing.assert_array_equal(pix.reshape(-1), np.frombuffer(earth_bmp[54:], np.uint8))
    tex = rumi.bmp_texture(earth_bmp, gamma)
    np.testing.assert_array_equal(tex.view(np.int32), oracle.bmp_texels(pix, gamma).view(np.int32))
    want = (pix[..., ::-1].astype(np.float64) / 255.0) ** gamma
    # 1/255 and the product each round once in FP32 (2.4e-7), the power scales it
    np.testing.assert_allclose(tex, want, rtol=2.5e-7 * max(gamma, 1.0), atol=1e-9)


def test_checker_c_class_and_texture_kinds(rumi):
    """A scene description takes checker textures (no texels) beside image
    textures; unknown kinds are rejected."""
    t = rumi.Texture()
    t.kind = 7
    assert t.kind == 7 and rumi.RUMI_TEXTURE_CHECKER == 1
