/* orc_dsfmt.c -- TEST INFRASTRUCTURE ONLY: the reference's random number
 * generator, dSFMT with Mersenne exponent 19937 (Saito and Matsumoto,
 * "A PRNG specialized in double precision floating point numbers using an
 * affine transition", MCQMC 2008; the reference vendors dSFMT 2.2.3 and
 * builds it with DSFMT_MEXP=19937, cmake/LMExternalDependencies.cmake:116),
 * restated so that tests can feed the reference's own uniform stream
 * (Random::Next = dsfmt_genrand_close_open, Random::NextUInt =
 * dsfmt_genrand_uint32, src/liblightmetrica/random.cpp:37-105) to the GPU
 * samplers as a uniform table.  Pinned by the golden streams of seeds 1 and
 * 42 (tests/golden/reference_vectors.npz, written from the vendored dSFMT.c
 * compiled in oracle/_ref).
 *
 * State: 191 128-bit words (two 64-bit lanes each) plus the "lung" word;
 * each lane of the first 191 holds a double in [1, 2) (exponent bits fixed).
 * The recursion per word: lung' = (a << 19) ^ swap32(lung) ^ b per lane,
 * r = (lung' >> 12) ^ (lung' & mask) ^ a, with b the word 117 ahead. */
#include <stdint.h>
#include <string.h>

#include "orc.h"

#define DS_N 191
#define DS_POS1 117
#define DS_SL1 19
#define DS_SR 12
#define DS_MSK1 0x000ffafffffffb3fULL
#define DS_MSK2 0x000ffdfffc90fffdULL
#define DS_FIX1 0x90014964b32f4329ULL
#define DS_FIX2 0x3b8d12ac548a7c7aULL
#define DS_PCV1 0x3d84e1ac0dc82880ULL
#define DS_PCV2 0x0000000000000001ULL
#define DS_LOW 0x000fffffffffffffULL
#define DS_HIGH 0x3ff0000000000000ULL

struct orc_dsfmt {
    uint64_t s[DS_N + 1][2];
    int idx;
};

static void ds_step(uint64_t r[2], const uint64_t a[2], const uint64_t b[2], uint64_t lung[2])
{
    const uint64_t t0 = a[0], t1 = a[1], l0 = lung[0], l1 = lung[1];
    lung[0] = (t0 << DS_SL1) ^ (l1 >> 32) ^ (l1 << 32) ^ b[0];
    lung[1] = (t1 << DS_SL1) ^ (l0 >> 32) ^ (l0 << 32) ^ b[1];
    r[0] = (lung[0] >> DS_SR) ^ (lung[0] & DS_MSK1) ^ t0;
    r[1] = (lung[1] >> DS_SR) ^ (lung[1] & DS_MSK2) ^ t1;
}

static void ds_regenerate(orc_dsfmt *g)
{
    uint64_t lung[2] = {g->s[DS_N][0], g->s[DS_N][1]};
    for (int i = 0; i < DS_N; i++) {
        const int j = i + DS_POS1 < DS_N ? i + DS_POS1 : i + DS_POS1 - DS_N;
        uint64_t r[2];
        ds_step(r, g->s[i], g->s[j], lung);
        g->s[i][0] = r[0];
        g->s[i][1] = r[1];
    }
    g->s[DS_N][0] = lung[0];
    g->s[DS_N][1] = lung[1];
}

void orc_dsfmt_seed(orc_dsfmt *g, uint32_t seed)
{
    /* the 32-bit words of the state (little-endian lanes), filled by the
     * Knuth-style linear recurrence of the Mersenne twisters */
    uint32_t w[(DS_N + 1) * 4];
    w[0] = seed;
    for (uint32_t i = 1; i < (DS_N + 1) * 4; i++) w[i] = 1812433253u * (w[i - 1] ^ (w[i - 1] >> 30)) + i;
    for (int i = 0; i <= DS_N; i++)
        for (int k = 0; k < 2; k++)
            g->s[i][k] = (uint64_t)w[4 * i + 2 * k] | ((uint64_t)w[4 * i + 2 * k + 1] << 32);
    for (int i = 0; i < DS_N; i++)
        for (int k = 0; k < 2; k++) g->s[i][k] = (g->s[i][k] & DS_LOW) | DS_HIGH;
    /* period certification on the lung word */
    uint64_t inner = ((g->s[DS_N][0] ^ DS_FIX1) & DS_PCV1) ^ ((g->s[DS_N][1] ^ DS_FIX2) & DS_PCV2);
    for (int sh = 32; sh > 0; sh >>= 1) inner ^= inner >> sh;
    if ((inner & 1) != 1) g->s[DS_N][1] ^= 1;
    g->idx = 2 * DS_N;
}

static uint64_t ds_next64(orc_dsfmt *g)
{
    if (g->idx >= 2 * DS_N) {
        ds_regenerate(g);
        g->idx = 0;
    }
    const uint64_t v = g->s[g->idx >> 1][g->idx & 1];
    g->idx++;
    return v;
}

/* dsfmt_genrand_close_open: the lane as a double in [1, 2), minus one */
double orc_dsfmt_next(orc_dsfmt *g)
{
    const uint64_t v = ds_next64(g);
    double d;
    memcpy(&d, &v, 8);
    return d - 1.0;
}

/* dsfmt_genrand_uint32: the lane's low 32 bits */
uint32_t orc_dsfmt_uint32(orc_dsfmt *g) { return (uint32_t)ds_next64(g); }

size_t orc_dsfmt_size(void) { return sizeof(orc_dsfmt); }

/* n values of seed's stream: doubles in [0, 1) (uints == 0) or 32-bit words */
void orc_dsfmt_stream(uint32_t seed, size_t n, double *next, uint32_t *uints)
{
    orc_dsfmt g;
    orc_dsfmt_seed(&g, seed);
    for (size_t i = 0; i < n; i++) {
        if (next) next[i] = orc_dsfmt_next(&g);
        else uints[i] = orc_dsfmt_uint32(&g);
    }
}
