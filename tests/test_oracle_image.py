"""The oracle's film output transforms (Film_HDR::Save, film_hdr.cpp:47-164)
pinned by known answers of the formats, and the host-side image readers of
liblmx on hand-written files.

* RGBE (Ward, "Real Pixels", Graphics Gems II): 1.0 -> (128, 128, 128, 129),
  3.0 -> (192, 0, 0, 130), values below 1e-32 -> all zero; decoding with
  c 2^(e - 136) returns every channel within one mantissa step (2^(e-136)) of
  the input and never above it (truncation); negative values are clamped to
  zero as the reference does before FreeImage sees them.
* PNG tone map: (int)(v^(1/2.2) 255) -> 0.5 -> 186, 1 -> 255, 0.25 -> 135;
  negatives -> 0, above 1 -> 255.
* File row order: the film's top row (row H-1) comes first.
* Readers: a flat and a run-length Radiance file, a PFM and their decoded
  values, written here byte by byte from the format descriptions.
"""

import numpy as np

import orc


def test_rgbe_known_answers():
    f = np.array([[[1.0, 1.0, 1.0], [3.0, 0.0, 0.0], [1e-33, 1e-33, 0.0], [-1.0, 2.0, 0.1],
                   [0.5, 0.5, 0.5], [255.0, 1.0, 0.0]]], np.float32)
    e = orc.film_rgbe(f)[0]
    assert e[0].tolist() == [128, 128, 128, 129]
    assert e[1].tolist() == [192, 0, 0, 130]
    assert e[2].tolist() == [0, 0, 0, 0]
    assert e[3].tolist() == [0, 128, 6, 130]
    assert e[4].tolist() == [128, 128, 128, 128]
    assert e[5].tolist() == [255, 1, 0, 136]


def _decode(e):
    e = e.astype(np.float64)
    f = np.where(e[..., 3:] > 0, np.ldexp(1.0, (e[..., 3:] - 136).astype(np.int64)), 0.0)
    return e[..., :3] * f, f


def test_rgbe_round_trip_bound():
    rng = np.random.default_rng(3)
    f = (rng.random((64, 96, 3)) * np.exp(rng.normal(0, 6, (64, 96, 1)))).astype(np.float32)
    f[0, :5] = -rng.random((5, 3))
    d, step = _decode(orc.film_rgbe(f))
    ref = np.maximum(f[::-1].astype(np.float64), 0.0)          # file rows top first
    assert np.all(d <= ref * (1 + 1e-7))
    assert np.all(ref - d <= step * (1 + 1e-6))
    m = ref.max(axis=2)
    ok = m > 1e-30
    assert np.all(np.abs(d - ref)[ok].max(axis=-1) <= m[ok] * 2.0 ** -7)


def test_png_tonemap_known_answers():
    f = np.array([[[0.5, 1.0, 0.25], [-0.5, 2.0, 0.0], [1e30, np.nan, 1e-12]]], np.float32)
    p = orc.film_png(f)[0]
    assert p[0].tolist() == [186, 255, 135]
    assert p[1].tolist() == [0, 255, 0]
    assert p[2].tolist() == [255, 0, 0]


def test_file_row_order_is_top_first():
    f = np.zeros((3, 2, 3), np.float32)
    f[2, :, :] = 1.0                                             # top row of the image
    assert orc.film_rgbe(f)[0, 0].tolist() == [128, 128, 128, 129]
    assert orc.film_png(f)[0, 0].tolist() == [255, 255, 255]
    assert orc.film_png(f)[2, 0].tolist() == [0, 0, 0]


def _rle(ch):
    """Run/literal encoding of one channel, written from the format
    description: runs of >= 4 equal bytes as (128 + n, byte)."""
    out, i = bytearray(), 0
    while i < len(ch):
        j = i
        while j < len(ch) and j - i < 127 and ch[j] == ch[i]:
            j += 1
        if j - i >= 4:
            out += bytes([128 + j - i, ch[i]])
            i = j
        else:
            k = min(len(ch) - i, 128)
            # stop the literal where a run of 4 starts
            for s in range(i + 1, i + k):
                if s + 3 < len(ch) and ch[s] == ch[s + 1] == ch[s + 2] == ch[s + 3]:
                    k = s - i
                    break
            out += bytes([k]) + bytes(ch[i:i + k])
            i += k
    return bytes(out)


def test_readers_on_hand_written_files(tmp_path):
    from lmx import image
    rng = np.random.default_rng(5)
    w, h = 11, 4
    e = rng.integers(0, 256, (h, w, 4), dtype=np.uint8)
    e[:, 3:7, :] = e[:, 3:4, :]                                  # runs
    e[1, :, 3] = 0                                               # zero exponent -> 0
    head = b"#?RADIANCE\nFORMAT=32-bit_rle_rgbe\n\n-Y %d +X %d\n" % (h, w)
    flat = head + e.tobytes()
    rle = bytearray(head)
    for r in range(h):
        rle += bytes([2, 2, w >> 8, w & 255])
        for c in range(4):
            rle += _rle(e[r, :, c].tobytes())
    expect, _ = _decode(e)
    expect = expect[::-1].astype(np.float32)                     # row 0 = bottom
    for name, data in (("flat.hdr", flat), ("rle.hdr", bytes(rle))):
        (tmp_path / name).write_bytes(data)
        np.testing.assert_array_equal(image.load(tmp_path / name), expect)
    pf = rng.normal(size=(h, w, 3)).astype(np.float32)
    (tmp_path / "a.pfm").write_bytes(b"PF\n%d %d\n-1.0\n" % (w, h) + pf.astype("<f4").tobytes())
    np.testing.assert_array_equal(image.load(tmp_path / "a.pfm"), pf)
    g = rng.normal(size=(h, w)).astype(np.float32)
    (tmp_path / "g.pfm").write_bytes(b"Pf\n%d %d\n-1.0\n" % (w, h) + g.astype("<f4").tobytes())
    np.testing.assert_array_equal(image.load(tmp_path / "g.pfm"), np.repeat(g[:, :, None], 3, axis=2))
    (tmp_path / "bad.hdr").write_bytes(b"P6\n")
    try:
        image.load(tmp_path / "bad.hdr")
        raise AssertionError("a non-Radiance file was read")
    except Exception as ex:                                      # LmxError
        assert "Radiance" in str(ex)
