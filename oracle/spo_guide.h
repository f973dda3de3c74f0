/* ORACLE -- test infrastructure only (see spo.h).
 *
 * FP64 restatement of the learned path guide: input encoding, the MLP
 * (64 -> 4 x 64 ReLU -> 48), the 8-lobe von Mises-Fisher mixture read from
 * its outputs, the mixture's density and sampling, and the weighted
 * negative-log-likelihood loss with its gradient.  The reference has no
 * guiding; the methods restated are:
 *   - vMF density and sampling: W. Jakob, "Numerically stable sampling of
 *     the von Mises Fisher distribution on S^2 (and other tricks)", 2012;
 *   - learning a parametric mixture online by minimizing the KL divergence
 *     to f*cos*Li with per-sample weights f*cos*Li/pdf: Dong et al.,
 *     "Neural Parametric Mixtures for Path Guiding", SIGGRAPH 2023;
 *     Mueller et al., "Neural Importance Sampling", TOG 2019;
 *   - frequency encoding of positions: Mildenhall et al., "NeRF", 2020;
 *   - Adam: Kingma and Ba, 2015.
 *
 * Parameter layout (shared with the GPU library, flat float array):
 *   W1[64][64] b1[64] W2[64][64] b2[64] W3[64][64] b3[64] W4[64][64] b4[64]
 *   W5[48][64] b5[48]          (W[out][in], row-major)
 * Output layout: o[0..7] lobe logits, o[8..15] kappa parameters,
 * o[16+3k..18+3k] lobe k's unnormalized mean, o[40..47] unused.
 *
 * Pinned by: finite differences of the loss (gradient), quadrature of the
 * density over the sphere (normalization), histograms of samples against
 * the density, closed forms at kappa -> 0 (uniform 1/4pi).
 */
#ifndef SPO_GUIDE_H
#define SPO_GUIDE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SPO_GUIDE_IN 64
#define SPO_GUIDE_WIDTH 64
#define SPO_GUIDE_HIDDEN 4
#define SPO_GUIDE_OUT 48
#define SPO_GUIDE_LOBES 8
#define SPO_GUIDE_FREQS 8
#define SPO_GUIDE_PARAMS (4 * (64 * 64 + 64) + 48 * 64 + 48)
/* log kappa = LK0 + LK1 * sigmoid(o) keeps kappa in (e^-2, e^6) */
#define SPO_GUIDE_LK0 (-2.0)
#define SPO_GUIDE_LK1 8.0

typedef struct {
    double pi[SPO_GUIDE_LOBES];
    double kappa[SPO_GUIDE_LOBES];
    double mu[SPO_GUIDE_LOBES][3];
} spo_Lobes;

/* Encoding of a hit: pn = position normalized to [0,1]^3 by the scene
 * bounds, n = shading normal, v = direction toward the previous vertex.
 * e[6l+2a] = sin(2^l pi pn_a), e[6l+2a+1] = cos(2^l pi pn_a) (l < 8),
 * e[48..50] = pn, e[51..53] = n, e[54..56] = v, e[57] = 1, e[58..63] = 0. */
void spo_guide_encode(const float *pn, const float *n, const float *v,
                      double *e64);

/* out48 = MLP(e64).  acts (optional) receives the post-ReLU activations of
 * the four hidden layers, 4 x 64. */
void spo_guide_forward(const double *params, const double *e64, double *out48,
                       double *acts);

/* The mixture a forward output describes. */
void spo_guide_lobes(const double *out48, spo_Lobes *L);
double spo_guide_pdf(const spo_Lobes *L, const double *w3);
/* Sample the mixture: u0 picks the lobe by the cumulative weights, (u1, u2)
 * the direction within it (Jakob 2012, section 2.2). */
void spo_guide_sample(const spo_Lobes *L, double u0, double u1, double u2,
                      double *w3);
/* One vMF lobe. */
double spo_vmf_pdf(double kappa, const double *mu3, const double *w3);
void spo_vmf_sample(double kappa, const double *mu3, double u1, double u2,
                    double *w3);

/* Loss over n samples, sample j = (e64 rows of enc, direction wi[3j..],
 * weight w[j]):  L = (1/n) sum_j w_j * (-log q(wi_j | e_j)).
 * grad (optional, SPO_GUIDE_PARAMS doubles) receives dL/dparams.  Returns
 * L. */
double spo_guide_loss(const double *params, const double *enc, const double *wi,
                      const double *w, uint32_t n, double *grad);

/* Adam step on float parameters (the GPU's master copy), t = step number
 * starting at 1. */
void spo_guide_adam(float *params, float *m, float *v, const double *grad,
                    uint32_t count, double lr, double beta1, double beta2,
                    double eps, uint32_t t);

#ifdef __cplusplus
}
#endif

#endif
