/* Test infrastructure: the oracle's restatement of the film output
 * transforms of Film_HDR::Save (src/liblightmetrica/asset/film/
 * film_hdr.cpp:47-164), which the GPU encoder of lmx_film_save must match
 * byte for byte.  Output is in file row order: the film's row h-1 (the top
 * of the image; film row 0 is the bottom) first, as FreeImage's bottom-up
 * scanlines (film row y -> scanline y, :84-100, :127-139) come out of its
 * writers. */
#include <math.h>
#include <stdint.h>
#include <stddef.h>

#include "orc.h"

/* FIT_RGBF with max(v, 0) per channel (:95-97), then Ward's RGBE
 * (Graphics Gems II "Real Pixels"; the float2rgbe of the Radiance rgbe
 * code): v = largest channel; below 1e-32 all bytes 0; else
 * scale = (float)(frexp(v, &e) * 256.0 / v), byte = (uint8)(c * scale),
 * exponent byte e + 128.  The product is clamped at 255 (it can round to
 * 256.0 in single precision when the mantissa is within an ulp of 1). */
void orc_film_rgbe(const float *film, int w, int h, uint8_t *out)
{
    for (int r = 0; r < h; r++)
        for (int x = 0; x < w; x++) {
            const float *p = film + 3 * ((size_t)(h - 1 - r) * (size_t)w + (size_t)x);
            float c[3];
            for (int k = 0; k < 3; k++) c[k] = p[k] > 0.0f ? p[k] : 0.0f;
            float v = c[0];
            if (c[1] > v) v = c[1];
            if (c[2] > v) v = c[2];
            uint8_t *o = out + 4 * ((size_t)r * (size_t)w + (size_t)x);
            if (v < 1e-32f) { o[0] = o[1] = o[2] = o[3] = 0; continue; }
            int e;
            const float sc = (float)(frexp((double)v, &e) * 256.0 / (double)v);
            for (int k = 0; k < 3; k++) {
                const float t = c[k] * sc;
                o[k] = (uint8_t)(t < 255.0f ? t : 255.0f);
            }
            o[3] = (uint8_t)(e + 128);
        }
}

/* PNG tone map (:127-139): (BYTE)Clamp((int)(pow((double)v, 1/2.2) * 255), 0,
 * 255).  (int) of NaN (a negative v) is taken as 0 and values beyond the int
 * range saturate (the C conversion is undefined there). */
void orc_film_png(const float *film, int w, int h, uint8_t *out)
{
    const double ex = 1.0 / 2.2;
    for (int r = 0; r < h; r++)
        for (int x = 0; x < w; x++) {
            const float *p = film + 3 * ((size_t)(h - 1 - r) * (size_t)w + (size_t)x);
            uint8_t *o = out + 3 * ((size_t)r * (size_t)w + (size_t)x);
            for (int k = 0; k < 3; k++) {
                const double t = pow((double)p[k], ex) * 255.0;
                const int i = t != t ? 0 : (t >= 2147483647.0 ? 2147483647 : (int)t);
                o[k] = (uint8_t)(i < 0 ? 0 : (i > 255 ? 255 : i));
            }
        }
}
