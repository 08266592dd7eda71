"""Film output and image input (Film_HDR::Save, src/liblightmetrica/asset/
film/film_hdr.cpp:47-164 and :260-280, without FreeImage).

`save(film, path, type)` writes a device film (H, W, 3 float32, row 0 =
bottom, as the renderers return it) to path + ".hdr" / ".exr" / ".png" /
".pfm" -- the suffix is appended as the reference appends it -- with the
pixels encoded on the GPU.  `load(path)` reads .hdr, .exr (as `save` writes
it) and .pfm files back into an (H, W, 3) float32 array with row 0 at the
bottom.
"""
import ctypes

import numpy as np

from ._lib import check, lib

RADIANCEHDR = 0
OPENEXR = 1
PNG = 2
PFM = 3
SUFFIX = {RADIANCEHDR: ".hdr", OPENEXR: ".exr", PNG: ".png", PFM: ".pfm"}


def save(film, path, type=RADIANCEHDR, stream=None):
    """Writes `film` (a CUDA float32 tensor (H, W, 3)); returns the file name."""
    import torch
    assert film.is_cuda and film.dtype == torch.float32 and film.is_contiguous() and film.dim() == 3
    h, w = film.shape[0], film.shape[1]
    if stream is None:
        stream = torch.cuda.current_stream(film.device)
    check(lib.lmx_film_save(ctypes.c_void_p(film.data_ptr()), w, h, type, str(path).encode(),
                            ctypes.c_void_p(stream.cuda_stream)))
    return str(path) + SUFFIX[type]


def load(path):
    """(H, W, 3) float32, row 0 = bottom."""
    w, h = ctypes.c_int32(), ctypes.c_int32()
    check(lib.lmx_image_load(str(path).encode(), None, ctypes.byref(w), ctypes.byref(h)))
    out = np.zeros((h.value, w.value, 3), np.float32)
    check(lib.lmx_image_load(str(path).encode(), ctypes.c_void_p(out.ctypes.data), ctypes.byref(w), ctypes.byref(h)))
    return out
