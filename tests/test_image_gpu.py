"""GPU film output (lmx_film_save) against the oracle's restatement of
Film_HDR::Save's transforms, byte for byte, with files decoded here by
independent readers written from the format descriptions:

* .hdr: header lines, run-length scanlines (width 8..32767) or flat ones
  (width 5); the decoded RGBE bytes equal the oracle's exactly;
* .png: chunk CRCs, zlib stream (Python's zlib), 8-bit RGB rows equal the
  oracle's tone-mapped bytes exactly;
* .exr: magic, header, offset table; float planes B, G, R equal max(film, 0)
  with the top row first;
* .pfm: the film as rendered, bottom row first;
* the reference's suffix rule (path + ".hdr" etc.), and lmx_image_load reads
  every written .hdr / .exr / .pfm back.
Films hold zeros, negatives, values below 1e-32, very bright pixels and a
rendered Cornell box.
"""
import struct
import zlib

import numpy as np
import pytest

import orc
from lmx import image, render, scenes

pytestmark = pytest.mark.gpu


def _films(gpu):
    import torch
    rng = np.random.default_rng(8)
    f = (rng.random((37, 300, 3)) * np.exp(rng.normal(0, 5, (37, 300, 1)))).astype(np.float32)
    f[0, :20] = -rng.random((20, 3))
    f[1, :20] = 1e-35
    f[2, :20] = 1e20
    f[3, :] = 0.0
    f[4, 10:200] = f[4, 10]                                      # long runs
    out = [("random_300", f), ("narrow_5", f[:9, :5].copy())]
    film, _, _ = render.Scene(scenes.cornell_box(64, 48)).render_pt(64 * 48 * 16, max_num_vertices=6, seed=2)
    out.append(("cornell", film.cpu().numpy()))
    return [(n, a, torch.from_numpy(a).to(gpu)) for n, a in out]


def _read_hdr(path):
    b = open(path, "rb").read()
    lines, p = [], 0
    while True:
        q = b.index(b"\n", p)
        lines.append(b[p:q])
        p = q + 1
        if lines[-1].startswith(b"-Y"):
            break
    assert lines[0] == b"#?RADIANCE" and b"FORMAT=32-bit_rle_rgbe" in lines
    _, h, _, w = lines[-1].split()
    h, w = int(h), int(w)
    out = np.zeros((h, w, 4), np.uint8)
    for r in range(h):
        if 8 <= w <= 32767:
            assert b[p:p + 2] == b"\x02\x02" and (b[p + 2] << 8 | b[p + 3]) == w
            p += 4
            for c in range(4):
                x = 0
                while x < w:
                    k = b[p]
                    p += 1
                    if k > 128:
                        out[r, x:x + k - 128, c] = b[p]
                        p += 1
                        x += k - 128
                    else:
                        out[r, x:x + k, c] = np.frombuffer(b[p:p + k], np.uint8)
                        p += k
                        x += k
        else:
            out[r] = np.frombuffer(b[p:p + 4 * w], np.uint8).reshape(w, 4)
            p += 4 * w
    assert p == len(b)
    return out


def _read_png(path):
    b = open(path, "rb").read()
    assert b[:8] == b"\x89PNG\r\n\x1a\n"
    p, chunks = 8, {}
    while p < len(b):
        n, = struct.unpack(">I", b[p:p + 4])
        t, d = b[p + 4:p + 8], b[p + 8:p + 8 + n]
        crc, = struct.unpack(">I", b[p + 8 + n:p + 12 + n])
        assert zlib.crc32(t + d) == crc
        chunks[t] = chunks.get(t, b"") + d
        p += 12 + n
    w, h, depth, ctype = struct.unpack(">IIBB", chunks[b"IHDR"][:10])
    assert depth == 8 and ctype == 2 and b"IEND" in chunks
    raw = np.frombuffer(zlib.decompress(chunks[b"IDAT"]), np.uint8).reshape(h, 3 * w + 1)
    assert np.all(raw[:, 0] == 0)
    return raw[:, 1:].reshape(h, w, 3)


def _read_exr(path):
    b = open(path, "rb").read()
    assert struct.unpack("<II", b[:8]) == (20000630, 2)
    p = 8
    attrs = {}
    while b[p] != 0:
        q = b.index(b"\0", p); name = b[p:q]; p = q + 1
        q = b.index(b"\0", p); p = q + 1
        n, = struct.unpack("<I", b[p:p + 4]); attrs[name] = b[p + 4:p + 4 + n]; p += 4 + n
    p += 1
    x0, y0, x1, y1 = struct.unpack("<iiii", attrs[b"dataWindow"])
    w, h = x1 - x0 + 1, y1 - y0 + 1
    assert attrs[b"compression"] == b"\0" and attrs[b"lineOrder"] == b"\0"
    assert attrs[b"channels"].startswith(b"B\0")
    offs = struct.unpack("<%dQ" % h, b[p:p + 8 * h])
    out = np.zeros((h, w, 3), np.float32)
    for y, o in enumerate(offs):
        yy, n = struct.unpack("<ii", b[o:o + 8])
        assert yy == y and n == 12 * w
        planes = np.frombuffer(b[o + 8:o + 8 + n], "<f4").reshape(3, w)
        out[y] = planes[::-1].T                                   # B, G, R -> R, G, B
    return out


def test_film_save_matches_oracle_bytes(gpu, tmp_path):
    for name, f, t in _films(gpu):
        base = tmp_path / name
        hdr = image.save(t, base, image.RADIANCEHDR)
        assert hdr == str(base) + ".hdr"
        np.testing.assert_array_equal(_read_hdr(hdr), orc.film_rgbe(f), err_msg=name)
        png = image.save(t, base, image.PNG)
        np.testing.assert_array_equal(_read_png(png), orc.film_png(f), err_msg=name)
        exr = image.save(t, base, image.OPENEXR)
        np.testing.assert_array_equal(_read_exr(exr), np.maximum(f, 0.0)[::-1], err_msg=name)
        pfm = image.save(t, base, image.PFM)
        b = open(pfm, "rb").read()
        head = b"PF\n%d %d\n-1.0\n" % (f.shape[1], f.shape[0])
        assert b.startswith(head)
        np.testing.assert_array_equal(np.frombuffer(b[len(head):], "<f4").reshape(f.shape), f)
        # lmx_image_load reads them back (row 0 = bottom)
        e = orc.film_rgbe(f).astype(np.float64)
        dec = np.where(e[..., 3:] > 0, e[..., :3] * np.ldexp(1.0, (e[..., 3:] - 136).astype(np.int64)), 0.0)
        np.testing.assert_array_equal(image.load(hdr), dec[::-1].astype(np.float32))
        np.testing.assert_array_equal(image.load(exr), np.maximum(f, 0.0))
        np.testing.assert_array_equal(image.load(pfm), f)
