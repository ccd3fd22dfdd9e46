"""Radiance RGBE files (LoadHDR, src/Kernel/hdr.cpp:4-177) and the HDR
texture's sphere lookup (HDRTexture::Pixel, src/Texture/hdrtexture.cpp:29-41,
Sphere::TextureMapping with a transform, src/Shape/sphere.cpp:172-195) on the
CPU: the library's reader against the oracle's and against closed forms on
files written here."""
import numpy as np
import pytest

import render_cases as rc


def rle_plane(b: np.ndarray) -> bytes:
    """New-style run-length coding of one channel plane: runs of 3 or more
    equal bytes as (128 + n, v), the rest as literal blocks (n, v...)."""
    out, i, n = bytearray(), 0, len(b)
    while i < n:
        j = i
        while j < n and j - i < 127 and b[j] == b[i]:
            j += 1
        if j - i >= 3:
            out += bytes([128 + (j - i), int(b[i])])
            i = j
            continue
        k = i
        while k < n and k - i < 128:
            if k + 2 < n and b[k] == b[k + 1] == b[k + 2]:
                break
            k += 1
        k = max(k, i + 1)
        out += bytes([k - i]) + bytes(int(x) for x in b[i:k])
        i = k
    return bytes(out)


def write_rgbe(rgbe: np.ndarray, header=b"# written by the tests\nFORMAT=32-bit_rle_rgbe\nGAMMA=1\n"
               b"EXPOSURE=1.0\nSOFTWARE=none\n\n") -> bytes:
    h, w, _ = rgbe.shape
    out = bytearray(b"#?RADIANCE\n" + header + b"-Y %d +X %d\n" % (h, w))
    for y in range(h):
        out += bytes([2, 2, w >> 8, w & 255])
        for c in range(4):
            out += rle_plane(rgbe[y, :, c])
    return bytes(out)


def random_rgbe(rng, h, w):
    v = rng.integers(0, 256, (h, w, 4), dtype=np.uint8)
    v[:, : w // 3] = v[:, :1]                              # runs
    v[..., 3] = rng.integers(100, 150, (h, w))
    v[0, 0, 3] = 0                                         # e = 0: black
    return v


@pytest.mark.parametrize("shape", [(5, 16), (3, 356), (2, 8)])
def test_reader_equals_oracle_and_closed_form(rumi, oracle, shape):
    rng = np.random.default_rng(shape[1])
    v = random_rgbe(rng, *shape)
    data = write_rgbe(v)
    lib = rumi.hdr_read(data)
    ora = oracle.hdr_read(data)
    want = v[..., :3].astype(np.float64) * np.exp2(v[..., 3:4].astype(np.float64) - 136)
    want[v[..., 3] == 0] = 0
    np.testing.assert_array_equal(ora, want.astype(np.float32))
    np.testing.assert_array_equal(lib.view(np.int32), ora.view(np.int32))


@pytest.mark.parametrize("case", ["width 200: low byte 0xc8 read as a negative char", "width 7",
                                  "no FORMAT line", "old-style scanline"])
def test_files_the_reference_cannot_read_are_rejected(rumi, oracle, case):
    rng = np.random.default_rng(1)
    if case.startswith("width 200"):
        data = write_rgbe(random_rgbe(rng, 2, 200))
    elif case == "width 7":
        data = write_rgbe(random_rgbe(rng, 2, 7))
    elif case == "no FORMAT line":
        data = write_rgbe(random_rgbe(rng, 2, 16), header=b"\n")
    else:
        good = write_rgbe(random_rgbe(rng, 2, 16))
        i = good.index(b"+X 16\n") + len(b"+X 16\n")
        data = good[:i] + b"\x01\x01\x01\x00" + good[i + 4:]
    with pytest.raises(ValueError):
        oracle.hdr_read(data)
    with pytest.raises(rumi.RumiError):
        rumi.hdr_read(data)


def test_hdr_texture_on_a_rotated_sphere_seen_from_inside(oracle):
    """An emissive sphere (Le = 1) with an HDR texture and the transform
    RotationY(-pi) (the reference's image-based-light sphere), seen from its
    centre through a 0.01 degree pixel: the texel HDRTexture::Pixel picks at
    Sphere::TextureMapping's (u, v) for that direction."""
    rng = np.random.default_rng(5)
    H, W = 9, 16
    tex = rng.uniform(0.1, 4.0, (H, W, 3)).astype(np.float32)
    c, s = np.float32(np.cos(np.float32(-np.pi))), np.float32(np.sin(np.float32(-np.pi)))
    basis = np.float32([[c, 0, s], [0, 1, 0], [-s, 0, c]])
    for d in ([0.3, 0.2, -1.0], [-0.7, -0.4, 0.2], [0.1, 0.9, 0.3]):
        d = np.asarray(d) / np.linalg.norm(d)
        sc = dict(materials=[dict(kind="emissive", kd=1.0, color=(1.0, 1.0, 1.0))],
                  shapes=[dict(kind="sphere", center=(0.0, 0.0, 0.0), radius=15.0, material=0,
                               texture=0, basis=basis)],
                  textures=[dict(kind="hdr", rgb=tex)],
                  camera=dict(position=(0.0, 0.0, 0.0), basis=rc.look_at_basis((0, 0, 0), d),
                              fov_deg=0.01, y_ratio=1.0), medium=None)
        mean, _ = oracle.render_pt(sc, 1, 1, 4, 2, seed=1, nthreads=1)
        pc = d
        theta = np.arccos(pc @ basis[1].astype(np.float64))
        pxz = pc * np.sin(theta)
        phi = np.arctan2(pxz @ basis[0].astype(np.float64), pxz @ basis[2].astype(np.float64))
        u = phi / (2 * np.pi) if phi >= 0 else (phi + 2 * np.pi) / (2 * np.pi)
        v = 1 - theta / np.pi
        row, col = int((1 - v) * (H - 1)), int((1 - u) * (W - 1))
        np.testing.assert_allclose(mean[0, 0], tex[row, col], rtol=1e-6)
