"""Footprint ingest on the CPU: the PNG reader of io::imRead (host code of
the library) and the oracle's restatement of ScaledMat's Lanczos-4 resize.

- PNG: the reference's own footprint (tests/golden/footprint/footprint.png,
  8-bit grey + alpha) against Pillow's decoder when Pillow is importable, and
  files written here by a minimal independent encoder that uses every row
  filter (0-4), at 8 and 16 bits, with and without alpha; malformed files
  fail with InvalidArgument.
- imRead's pixel transform: v * (1/maxval) in FP64, then srgb2rgb.
- Lanczos-4 (OpenCV's INTER_LANCZOS4, which the reference calls through
  ScaledMat): the unit tap at x = 0, identity at the same size, constants
  kept, coefficients against the kernel's closed form sinc(t) sinc(t/4),
  and the whole resize against an independent numpy restatement of that
  closed form with replicated borders.  OpenCV itself is absent here, so its
  float rounding of the source position and weights is pinned only by
  restating it (parity of those last bits against OpenCV: unpinned).
"""
import math
import os
import struct
import zlib

import numpy as np
import pytest

from painty_amd import _lib

FP_PNG = os.path.join(os.path.dirname(__file__), "golden", "footprint", "footprint.png")


def _hip():
    from painty_amd import hip as H
    _lib.load()
    return H


# ----------------------------------------------------------------- PNG -----
def _chunk(t: bytes, d: bytes) -> bytes:
    return struct.pack(">I", len(d)) + t + d + struct.pack(">I", zlib.crc32(t + d) & 0xFFFFFFFF)


def _paeth(a, b, c):
    p = a + b - c
    pa, pb, pc = abs(p - a), abs(p - b), abs(p - c)
    return a if pa <= pb and pa <= pc else (b if pb <= pc else c)


def encode_png(samples: np.ndarray, depth: int, alpha=None, filters=(0, 1, 2, 3, 4)) -> bytes:
    """Grey (or grey + alpha) PNG, row y filtered with filters[y % len]."""
    h, w = samples.shape
    chans = 1 if alpha is None else 2
    bps = depth // 8
    px = np.zeros((h, w, chans), np.uint16)
    px[..., 0] = samples
    if alpha is not None:
        px[..., 1] = alpha
    if bps == 2:
        raw = np.stack([px >> 8, px & 0xFF], -1).astype(np.uint8).reshape(h, w * chans * 2)
    else:
        raw = px.astype(np.uint8).reshape(h, w * chans)
    bpp = chans * bps
    out = bytearray()
    prev = np.zeros(raw.shape[1], np.int64)
    for y in range(h):
        f = filters[y % len(filters)]
        x = raw[y].astype(np.int64)
        enc = np.zeros_like(x)
        for i in range(len(x)):
            a = x[i - bpp] if i >= bpp else 0
            b = prev[i]
            c = prev[i - bpp] if i >= bpp else 0
            pred = [0, a, b, (a + b) >> 1, _paeth(a, b, c)][f]
            enc[i] = (x[i] - pred) & 0xFF
        out.append(f)
        out += bytes(enc.astype(np.uint8))
        prev = x
    ctype = 0 if alpha is None else 4
    ihdr = struct.pack(">IIBBBBB", w, h, depth, ctype, 0, 0, 0)
    return (b"\x89PNG\r\n\x1a\n" + _chunk(b"IHDR", ihdr) +
            _chunk(b"IDAT", zlib.compress(bytes(out), 6)) + _chunk(b"IEND", b""))


@pytest.mark.parametrize("depth", [8, 16])
@pytest.mark.parametrize("with_alpha", [False, True])
def test_png_every_filter(depth, with_alpha):
    H = _hip()
    rng = np.random.default_rng(depth + 3 * with_alpha)
    hi = 255 if depth == 8 else 65535
    s = rng.integers(0, hi + 1, (13, 11)).astype(np.uint16)
    s[:, 5] = hi  # saturated and zero columns stress the filter arithmetic
    s[:, 6] = 0
    a = rng.integers(0, hi + 1, s.shape).astype(np.uint16) if with_alpha else None
    data = encode_png(s, depth, a)
    got = H.read_image_gray(data, srgb=False)
    np.testing.assert_array_equal(got, s.astype(np.float64) * (1.0 / hi))


def test_reference_footprint_against_pillow():
    PIL = pytest.importorskip("PIL.Image")
    H = _hip()
    data = open(FP_PNG, "rb").read()
    ref = np.asarray(PIL.open(FP_PNG))
    assert ref.shape == (1024, 1024, 2)  # grey + alpha, as the reference ships it
    got = H.read_image_gray(data, srgb=False)
    np.testing.assert_array_equal(got, ref[..., 0].astype(np.float64) * (1.0 / 255))


def test_read_gray_srgb_matches_oracle(oracle):
    H = _hip()
    rng = np.random.default_rng(4)
    s = rng.integers(0, 256, (9, 14)).astype(np.uint16)
    s[0, :3] = [0, 10, 255]  # below and above the linear segment's end (10.3 / 255)
    got = H.read_image_gray(encode_png(s, 8), srgb=True)
    np.testing.assert_array_equal(got, oracle.gray_from_u16(s, 255, True))


def test_srgb2rgb_pins(oracle):
    assert oracle.srgb2rgb(0.0) == 0.0
    assert abs(oracle.srgb2rgb(1.0) - 1.0) < 1e-15
    t = 0.0404482362771082  # the branches meet here (to 1e-9 relative)
    lo, hi = t / 12.92, ((t + 0.055) / 1.055) ** 2.4
    assert abs(lo - hi) / lo < 1e-7
    assert oracle.srgb2rgb(t) == t / 12.92
    for x in np.linspace(0.0, 1.0, 50):  # the reference's own inverse, rgb2srgb
        assert abs(oracle.rgb2srgb(oracle.srgb2rgb(x)) - x) < 1e-12


@pytest.mark.parametrize("mutate", ["crc", "truncate", "colour", "interlace", "signature"])
def test_png_errors(mutate):
    H = _hip()
    s = np.arange(12, dtype=np.uint16).reshape(3, 4) * 20
    data = bytearray(encode_png(s, 8))
    if mutate == "crc":
        data[40] ^= 0xFF  # inside IDAT
    elif mutate == "truncate":
        data = data[:len(data) - 20]
    elif mutate == "colour":  # colour type 2 (RGB) in IHDR, CRC redone
        ihdr = struct.pack(">IIBBBBB", 4, 3, 8, 2, 0, 0, 0)
        data = bytearray(b"\x89PNG\r\n\x1a\n" + _chunk(b"IHDR", ihdr)) + data[33:]
    elif mutate == "interlace":
        ihdr = struct.pack(">IIBBBBB", 4, 3, 8, 0, 0, 0, 1)
        data = bytearray(b"\x89PNG\r\n\x1a\n" + _chunk(b"IHDR", ihdr)) + data[33:]
    else:
        data[1] = ord("Q")
    with pytest.raises(_lib.InvalidArgument):
        H.read_image_gray(bytes(data))


# ------------------------------------------------------------ Lanczos-4 -----
def lanczos_kernel(t):
    t = np.asarray(t, np.float64)
    return np.where(np.abs(t) < 1e-12, 1.0, np.sinc(t) * np.sinc(t / 4.0))


def brute_resize(img, out_rows, out_cols):
    """Direct Lanczos-4: output d samples the source at (d + 0.5) * s - 0.5,
    taps floor(.) - 3 .. floor(.) + 4 weighted by the normalised kernel,
    indices clamped to the image (replicated border); rows, then columns."""
    def axis(n_src, n_dst):
        W = np.zeros((n_dst, n_src))
        s = n_src / n_dst
        for d in range(n_dst):
            f = (d + 0.5) * s - 0.5
            i0 = math.floor(f)
            taps = np.arange(i0 - 3, i0 + 5)
            w = lanczos_kernel(f - taps)
            w /= w.sum()
            for t, wt in zip(np.clip(taps, 0, n_src - 1), w):
                W[d, t] += wt
        return W
    img = np.asarray(img, np.float64)
    return axis(img.shape[0], out_rows) @ img @ axis(img.shape[1], out_cols).T


def test_lanczos_coefficients(oracle):
    np.testing.assert_array_equal(oracle.lanczos4_coeffs(0.0), [0, 0, 0, 1, 0, 0, 0, 0])
    for x in np.linspace(0.01, 0.99, 25):
        c = oracle.lanczos4_coeffs(x).astype(np.float64)
        want = lanczos_kernel(float(np.float32(x)) + 3 - np.arange(8))
        np.testing.assert_allclose(c, want / want.sum(), rtol=0, atol=5e-7)  # float weights
        assert abs(c.sum() - 1.0) < 1e-6
        np.testing.assert_allclose(c[::-1], oracle.lanczos4_coeffs(1.0 - x), atol=5e-7)


def test_lanczos_identity_and_constant(oracle):
    rng = np.random.default_rng(1)
    a = rng.random((17, 23))
    np.testing.assert_array_equal(oracle.resize_lanczos4(a, 17, 23), a)
    out = oracle.resize_lanczos4(np.full((40, 64), 0.37), 9, 13)
    np.testing.assert_allclose(out, 0.37, rtol=1e-6)


@pytest.mark.parametrize("shape,out", [((64, 64), (9, 9)), ((20, 30), (47, 61)),
                                       ((1, 40), (3, 7)), ((33, 5), (8, 17))])
def test_lanczos_against_brute(oracle, shape, out):
    rng = np.random.default_rng(shape[0] * 100 + out[1])
    a = rng.random(shape)
    np.testing.assert_allclose(oracle.resize_lanczos4(a, *out), brute_resize(a, *out),
                               rtol=0, atol=2e-6)


def test_reference_footprint_rings(oracle):
    """The reference's footprint resized as setRadius does has cells a little
    below 0 and above 1 (Lanczos ringing at the blob's edges), which is why
    footprints are not clamped to [0, 1]."""
    H = _hip()
    img = H.read_image_gray(open(FP_PNG, "rb").read(), srgb=True)
    f = oracle.resize_lanczos4(img, 65, 65)
    assert f.min() < -0.01 and f.max() > 1.01
    # the float source positions move taps by ~1e-7 px on edges of slope ~100 per px
    np.testing.assert_allclose(f, brute_resize(img, 65, 65), rtol=0, atol=1e-4)
