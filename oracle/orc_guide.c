/*
 * orc_guide.c -- CPU oracle of the learned directional sampler.
 *
 * TEST INFRASTRUCTURE ONLY (see orc.h).  The sampler has no counterpart in
 * the reference: this file restates, in FP64, the model the device kernels
 * implement, so that they can be checked against it.
 *
 *   network   168 sin/cos features -> 4 x (64, ReLU) -> 41 outputs
 *   mixture   8 von Mises-Fisher lobes on S^2 (pdf and sampling after
 *             W. Jakob, "Numerically stable sampling of the von Mises Fisher
 *             distribution on S^2", 2012) mixed one-sample with the BSDF
 *   loss      the Monte Carlo estimate of the KL divergence between the
 *             sampled integrand and the mixture, -w log q (w = radiance times
 *             BSDF times cosine over the density the direction was drawn
 *             from), as in neural parametric mixtures for path guiding
 *             (Dong et al., SIGGRAPH 2023)
 *   optimiser Adam (Kingma and Ba, 2015) with bias correction
 */
#include <math.h>
#include <stdint.h>
#include <string.h>
#include <stdlib.h>

#include "orc.h"

#define GP_W1 0
#define GP_W2 (GP_W1 + 169 * 64)
#define GP_B2 (GP_W2 + 64 * 64)
#define GP_W3 (GP_B2 + 64)
#define GP_B3 (GP_W3 + 64 * 64)
#define GP_W4 (GP_B3 + 64)
#define GP_B4 (GP_W4 + 64 * 64)
#define GP_W5 (GP_B4 + 64)
#define GP_B5 (GP_W5 + 64 * 41)
#define GP_N (GP_B5 + 41)

#if GP_N != ORC_GUIDE_NPARAM
#error parameter layout size
#endif

#define LOGK_MIN (-4.0)
#define LOGK_MAX 9.0
#define ALPHA_MIN 0.5
#define ALPHA_SPAN 0.4
static const double PI = 3.14159265358979323846;

double orc_bf16_round(double x)
{
    float f = (float)x;
    uint32_t u;
    memcpy(&u, &f, 4);
    if ((u & 0x7f800000u) == 0x7f800000u) return x;      /* inf / nan */
    u += 0x7fffu + ((u >> 16) & 1u);
    u &= 0xffff0000u;
    memcpy(&f, &u, 4);
    return (double)f;
}

void orc_guide_features(const double x[7], double feat[168])
{
    for (int d = 0; d < 7; d++)
        for (int f = 0; f < 12; f++) {
            const double a = ldexp(PI * x[d], f);
            feat[d * 24 + 2 * f] = sin(a);
            feat[d * 24 + 2 * f + 1] = cos(a);
        }
}

/* activations of one record: layer inputs h0 (168), h1..h4 (64 each, after
 * ReLU), outputs y (41) */
typedef struct { double h0[168], h[4][64], y[41]; } acts;

static double wq(const double *params, size_t i, int bf16) { return bf16 ? orc_bf16_round(params[i]) : params[i]; }

static void forward(const double *P, const double x[7], acts *a, int bf16)
{
    orc_guide_features(x, a->h0);
    if (bf16)
        for (int k = 0; k < 168; k++) a->h0[k] = orc_bf16_round(a->h0[k]);
    for (int n = 0; n < 64; n++) {
        double s = wq(P, GP_W1 + 168 * 64 + n, bf16);           /* bias row */
        for (int k = 0; k < 168; k++) s += a->h0[k] * wq(P, GP_W1 + k * 64 + n, bf16);
        a->h[0][n] = s > 0.0 ? s : 0.0;
    }
    static const int W[3] = {GP_W2, GP_W3, GP_W4}, B[3] = {GP_B2, GP_B3, GP_B4};
    for (int l = 0; l < 3; l++) {
        double in[64];
        for (int k = 0; k < 64; k++) in[k] = bf16 ? orc_bf16_round(a->h[l][k]) : a->h[l][k];
        if (bf16) memcpy(a->h[l], in, sizeof in);
        for (int n = 0; n < 64; n++) {
            double s = P[B[l] + n];
            for (int k = 0; k < 64; k++) s += in[k] * wq(P, W[l] + k * 64 + n, bf16);
            a->h[l + 1][n] = s > 0.0 ? s : 0.0;
        }
    }
    if (bf16)
        for (int k = 0; k < 64; k++) a->h[3][k] = orc_bf16_round(a->h[3][k]);
    for (int n = 0; n < 41; n++) {
        double s = P[GP_B5 + n];
        for (int k = 0; k < 64; k++) s += a->h[3][k] * wq(P, GP_W5 + k * 41 + n, bf16);
        a->y[n] = s;
    }
}

void orc_guide_forward(const double *params, const double x[7], double y[41], int bf16)
{
    acts a;
    forward(params, x, &a, bf16);
    memcpy(y, a.y, sizeof a.y);
}

void orc_guide_forward_batch(const double *params, const double *x, uint64_t n, double *y, int bf16)
{
    for (uint64_t i = 0; i < n; i++) orc_guide_forward(params, x + 7 * i, y + 41 * i, bf16);
}

/* ---------------------------------------------------------------------------
 * mixture */

typedef struct {
    double mu[8][3], mlen[8], kappa[8], kclamped[8], pi[8], alpha, sig;
} mixture;

static void decode(const double y[41], mixture *m)
{
    for (int k = 0; k < 8; k++) {
        const double *r = y + 3 * k;
        const double l = sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
        m->mlen[k] = l;
        if (l > 0.0) { m->mu[k][0] = r[0] / l; m->mu[k][1] = r[1] / l; m->mu[k][2] = r[2] / l; }
        else { m->mu[k][0] = 0.0; m->mu[k][1] = 0.0; m->mu[k][2] = 1.0; }
        double lk = y[24 + k];
        m->kclamped[k] = lk < LOGK_MIN || lk > LOGK_MAX;
        lk = lk < LOGK_MIN ? LOGK_MIN : (lk > LOGK_MAX ? LOGK_MAX : lk);
        m->kappa[k] = exp(lk);
    }
    double mx = y[32];
    for (int k = 1; k < 8; k++) mx = y[32 + k] > mx ? y[32 + k] : mx;
    double s = 0.0;
    for (int k = 0; k < 8; k++) { m->pi[k] = exp(y[32 + k] - mx); s += m->pi[k]; }
    for (int k = 0; k < 8; k++) m->pi[k] /= s;
    m->sig = 1.0 / (1.0 + exp(-y[40]));
    m->alpha = ALPHA_MIN + ALPHA_SPAN * m->sig;
}

/* kappa / (2 pi (1 - e^{-2 kappa})) exp(kappa (mu.w - 1)): the normalised
 * density written so that neither factor overflows */
double orc_guide_vmf_pdf(const double mu[3], double kappa, const double w[3])
{
    const double c = mu[0] * w[0] + mu[1] * w[1] + mu[2] * w[2];
    return kappa / (2.0 * PI * -expm1(-2.0 * kappa)) * exp(kappa * (c - 1.0));
}

/* OrthonormalBasis of include/lightmetrica/math.h:2355-2359 (the device code
 * builds the lobe frame the same way) */
static void basis(const double a[3], double b[3], double c[3])
{
    double t[3];
    if (fabs(a[0]) > fabs(a[1])) { t[0] = a[2]; t[1] = 0.0; t[2] = -a[0]; }
    else { t[0] = 0.0; t[1] = a[2]; t[2] = -a[1]; }
    const double l = sqrt(t[0] * t[0] + t[1] * t[1] + t[2] * t[2]);
    c[0] = t[0] / l; c[1] = t[1] / l; c[2] = t[2] / l;
    b[0] = c[1] * a[2] - a[1] * c[2];
    b[1] = c[2] * a[0] - a[2] * c[0];
    b[2] = c[0] * a[1] - a[0] * c[1];
}

/* Jakob 2012: cos theta = 1 + log(u1 + (1 - u1) e^{-2 kappa}) / kappa */
void orc_guide_vmf_sample(const double mu[3], double kappa, double u1, double u2, double out[3])
{
    const double W = 1.0 + log(u1 + (1.0 - u1) * exp(-2.0 * kappa)) / kappa;
    const double r = sqrt(fmax(0.0, 1.0 - W * W));
    const double ph = 2.0 * PI * u2;
    double b[3], c[3];
    basis(mu, b, c);
    for (int i = 0; i < 3; i++) out[i] = W * mu[i] + r * (cos(ph) * b[i] + sin(ph) * c[i]);
}

double orc_guide_mix_pdf(const double y[41], const double w[3], double *alpha)
{
    mixture m;
    decode(y, &m);
    double p = 0.0;
    for (int k = 0; k < 8; k++) p += m.pi[k] * orc_guide_vmf_pdf(m.mu[k], m.kappa[k], w);
    if (alpha) *alpha = m.alpha;
    return p;
}

int orc_guide_mix_sample(const double y[41], double u, double u1, double u2, double out[3])
{
    mixture m;
    decode(y, &m);
    int k = 0;
    double c = m.pi[0];
    while (k < 7 && u >= c) { k++; c += m.pi[k]; }
    orc_guide_vmf_sample(m.mu[k], m.kappa[k], u1, u2, out);
    return k;
}

/* ---------------------------------------------------------------------------
 * loss and gradient */

/* d(-w log q)/dy for one record */
static double out_grad(const double y[41], const double dir[3], double pb, double w, double g[41])
{
    mixture m;
    decode(y, &m);
    double v[8], pv = 0.0;
    for (int k = 0; k < 8; k++) { v[k] = orc_guide_vmf_pdf(m.mu[k], m.kappa[k], dir); pv += m.pi[k] * v[k]; }
    const double q = m.alpha * pb + (1.0 - m.alpha) * pv;
    const double gq = -w / q;
    memset(g, 0, 41 * sizeof(double));
    g[40] = gq * (pb - pv) * ALPHA_SPAN * m.sig * (1.0 - m.sig);
    for (int k = 0; k < 8; k++) {
        const double base = gq * (1.0 - m.alpha) * m.pi[k];
        g[32 + k] = base * (v[k] - pv);
        const double kap = m.kappa[k];
        const double c = m.mu[k][0] * dir[0] + m.mu[k][1] * dir[1] + m.mu[k][2] * dir[2];
        if (!m.kclamped[k]) g[24 + k] = base * v[k] * (1.0 - 2.0 * kap / expm1(2.0 * kap) + kap * (c - 1.0));
        if (m.mlen[k] > 0.0)
            for (int i = 0; i < 3; i++) g[3 * k + i] = base * v[k] * kap * (dir[i] - c * m.mu[k][i]) / m.mlen[k];
    }
    return -w * log(q);
}

double orc_guide_loss(const double *P, uint64_t n, const double *x, const double *dir, const double *pb,
                      const double *w, double *grad, int bf16)
{
    double L = 0.0;
    if (grad) memset(grad, 0, GP_N * sizeof(double));
    acts *a = (acts *)malloc(sizeof(acts));
    for (uint64_t j = 0; j < n; j++) {
        forward(P, x + 7 * j, a, bf16);
        double gy[41];
        L += out_grad(a->y, dir + 3 * j, pb[j], w[j], gy);
        if (!grad) continue;
        double d[64], dn[64];
        if (bf16)
            for (int o = 0; o < 41; o++) gy[o] = orc_bf16_round(gy[o]);
        /* output layer */
        for (int o = 0; o < 41; o++) {
            grad[GP_B5 + o] += gy[o];
            for (int k = 0; k < 64; k++) grad[GP_W5 + k * 41 + o] += a->h[3][k] * gy[o];
        }
        for (int k = 0; k < 64; k++) {
            double s = 0.0;
            for (int o = 0; o < 41; o++) s += wq(P, GP_W5 + k * 41 + o, bf16) * gy[o];
            d[k] = a->h[3][k] > 0.0 ? s : 0.0;
            if (bf16) d[k] = orc_bf16_round(d[k]);
        }
        /* hidden layers 4, 3, 2 (weights W4, W3, W2; inputs h[2], h[1], h[0]) */
        static const int W[3] = {GP_W4, GP_W3, GP_W2}, B[3] = {GP_B4, GP_B3, GP_B2};
        for (int l = 0; l < 3; l++) {
            const double *in = a->h[2 - l];
            for (int o = 0; o < 64; o++) {
                grad[B[l] + o] += d[o];
                for (int k = 0; k < 64; k++) grad[W[l] + k * 64 + o] += in[k] * d[o];
            }
            for (int k = 0; k < 64; k++) {
                double s = 0.0;
                for (int o = 0; o < 64; o++) s += wq(P, W[l] + k * 64 + o, bf16) * d[o];
                dn[k] = in[k] > 0.0 ? s : 0.0;
                if (bf16) dn[k] = orc_bf16_round(dn[k]);
            }
            memcpy(d, dn, sizeof d);
        }
        /* first layer (its input is the fixed encoding) */
        for (int o = 0; o < 64; o++) {
            grad[GP_W1 + 168 * 64 + o] += d[o];
            for (int k = 0; k < 168; k++) grad[GP_W1 + k * 64 + o] += a->h0[k] * d[o];
        }
    }
    free(a);
    if (grad)
        for (int i = 0; i < GP_N; i++) grad[i] /= (double)n;
    return L / (double)n;
}

void orc_guide_adam(double *params, double *m, double *v, const double *grad, uint64_t n, int step, double lr,
                    double beta1, double beta2, double eps)
{
    const double c1 = 1.0 - pow(beta1, step), c2 = 1.0 - pow(beta2, step);
    for (uint64_t i = 0; i < n; i++) {
        m[i] = beta1 * m[i] + (1.0 - beta1) * grad[i];
        v[i] = beta2 * v[i] + (1.0 - beta2) * grad[i] * grad[i];
        params[i] -= lr * (m[i] / c1) / (sqrt(v[i] / c2) + eps);
    }
}
