/* ORACLE -- test infrastructure only.  FP64 restatement of the learned path
 * guide; see spo_guide.h for the methods and the parameter layout. */
#include "spo_guide.h"

#include <math.h>
#include <string.h>

#define IN SPO_GUIDE_IN
#define WD SPO_GUIDE_WIDTH
#define OUT SPO_GUIDE_OUT
#define NL SPO_GUIDE_LOBES

/* offsets of W_l and b_l in the flat parameter array, l = 0..4 */
static uint32_t w_off(int l) { return (uint32_t)l * (WD * WD + WD); }
static uint32_t b_off(int l) { return w_off(l) + (uint32_t)(l < 4 ? WD * WD : OUT * WD); }
static int rows(int l) { return l < 4 ? WD : OUT; }

void spo_guide_encode(const float *pn, const float *n, const float *v,
                      double *e)
{
    for (int l = 0; l < SPO_GUIDE_FREQS; ++l)
        for (int a = 0; a < 3; ++a)
        {
            double x = ldexp((double)pn[a], l) * M_PI;
            e[6 * l + 2 * a] = sin(x);
            e[6 * l + 2 * a + 1] = cos(x);
        }
    for (int a = 0; a < 3; ++a)
    {
        e[48 + a] = pn[a];
        e[51 + a] = n[a];
        e[54 + a] = v[a];
    }
    e[57] = 1.0;
    for (int k = 58; k < IN; ++k) e[k] = 0.0;
}

void spo_guide_forward(const double *P, const double *e, double *out,
                       double *acts)
{
    double h[WD], z[WD];
    memcpy(h, e, sizeof h);
    for (int l = 0; l < 4; ++l)
    {
        const double *W = P + w_off(l), *b = P + b_off(l);
        for (int o = 0; o < WD; ++o)
        {
            double s = b[o];
            for (int i = 0; i < WD; ++i) s += W[o * WD + i] * h[i];
            z[o] = s > 0.0 ? s : 0.0;
        }
        memcpy(h, z, sizeof h);
        if (acts) memcpy(acts + l * WD, h, sizeof h);
    }
    const double *W = P + w_off(4), *b = P + b_off(4);
    for (int o = 0; o < OUT; ++o)
    {
        double s = b[o];
        for (int i = 0; i < WD; ++i) s += W[o * WD + i] * h[i];
        out[o] = s;
    }
}

static double sigmoid(double x) { return 1.0 / (1.0 + exp(-x)); }

void spo_guide_lobes(const double *o, spo_Lobes *L)
{
    double mx = o[0];
    for (int k = 1; k < NL; ++k) mx = o[k] > mx ? o[k] : mx;
    double s = 0.0;
    for (int k = 0; k < NL; ++k) s += (L->pi[k] = exp(o[k] - mx));
    for (int k = 0; k < NL; ++k)
    {
        L->pi[k] /= s;
        L->kappa[k] = exp(SPO_GUIDE_LK0 + SPO_GUIDE_LK1 * sigmoid(o[8 + k]));
        const double *m = o + 16 + 3 * k;
        double len = sqrt(m[0] * m[0] + m[1] * m[1] + m[2] * m[2]) + 1e-12;
        for (int c = 0; c < 3; ++c) L->mu[k][c] = m[c] / len;
    }
}

static double vmf_log_pdf(double kappa, const double *mu, const double *w)
{
    double d = mu[0] * w[0] + mu[1] * w[1] + mu[2] * w[2];
    return log(kappa) - log(2.0 * M_PI) - log(-expm1(-2.0 * kappa)) +
           kappa * (d - 1.0);
}

double spo_vmf_pdf(double kappa, const double *mu, const double *w)
{
    return exp(vmf_log_pdf(kappa, mu, w));
}

double spo_guide_pdf(const spo_Lobes *L, const double *w)
{
    double q = 0.0;
    for (int k = 0; k < NL; ++k) q += L->pi[k] * spo_vmf_pdf(L->kappa[k], L->mu[k], w);
    return q;
}

/* Orthonormal basis about n (Duff et al., "Building an Orthonormal Basis,
 * Revisited", JCGT 2017). */
static void basis(const double *n, double *t1, double *t2)
{
    double sign = n[2] >= 0.0 ? 1.0 : -1.0;
    double a = -1.0 / (sign + n[2]);
    double b = n[0] * n[1] * a;
    t1[0] = 1.0 + sign * n[0] * n[0] * a;
    t1[1] = sign * b;
    t1[2] = -sign * n[0];
    t2[0] = b;
    t2[1] = sign + n[1] * n[1] * a;
    t2[2] = -n[1];
}

void spo_vmf_sample(double kappa, const double *mu, double u1, double u2,
                    double *w)
{
    double c = 1.0 + log(u1 + (1.0 - u1) * exp(-2.0 * kappa)) / kappa;
    c = c < -1.0 ? -1.0 : (c > 1.0 ? 1.0 : c);
    double s = sqrt(fmax(0.0, 1.0 - c * c));
    double phi = 2.0 * M_PI * u2;
    double t1[3], t2[3];
    basis(mu, t1, t2);
    for (int a = 0; a < 3; ++a)
        w[a] = t1[a] * s * cos(phi) + t2[a] * s * sin(phi) + mu[a] * c;
}

void spo_guide_sample(const spo_Lobes *L, double u0, double u1, double u2,
                      double *w)
{
    int k = NL - 1;
    double c = 0.0;
    for (int j = 0; j < NL; ++j)
    {
        c += L->pi[j];
        if (u0 < c)
        {
            k = j;
            break;
        }
    }
    spo_vmf_sample(L->kappa[k], L->mu[k], u1, u2, w);
}

/* d(-w log q)/d(out) for one sample. */
static double output_grad(const double *o, const double *wi, double weight,
                          double *g)
{
    spo_Lobes L;
    spo_guide_lobes(o, &L);
    double la[NL], mx = -INFINITY;
    for (int k = 0; k < NL; ++k)
    {
        la[k] = log(L.pi[k]) + vmf_log_pdf(L.kappa[k], L.mu[k], wi);
        mx = la[k] > mx ? la[k] : mx;
    }
    double s = 0.0;
    for (int k = 0; k < NL; ++k) s += exp(la[k] - mx);
    double lq = mx + log(s);
    for (int j = 0; j < OUT; ++j) g[j] = 0.0;
    for (int k = 0; k < NL; ++k)
    {
        double gam = exp(la[k] - lq);
        double kap = L.kappa[k];
        const double *mu = L.mu[k];
        double d = mu[0] * wi[0] + mu[1] * wi[1] + mu[2] * wi[2];
        g[k] = weight * (L.pi[k] - gam);
        double sg = sigmoid(o[8 + k]);
        double dlogv = 1.0 - 2.0 * kap / expm1(2.0 * kap) + kap * (d - 1.0);
        g[8 + k] = -weight * gam * dlogv * SPO_GUIDE_LK1 * sg * (1.0 - sg);
        const double *m = o + 16 + 3 * k;
        double len = sqrt(m[0] * m[0] + m[1] * m[1] + m[2] * m[2]) + 1e-12;
        for (int c = 0; c < 3; ++c)
            g[16 + 3 * k + c] = -weight * gam * kap * (wi[c] - mu[c] * d) / len;
    }
    return -weight * lq;
}

double spo_guide_loss(const double *P, const double *enc, const double *wi,
                      const double *w, uint32_t n, double *grad)
{
    if (grad) memset(grad, 0, sizeof(double) * SPO_GUIDE_PARAMS);
    double total = 0.0;
    double acts[4 * WD], out[OUT], g[OUT], gh[WD], gz[WD];
    for (uint32_t j = 0; j < n; ++j)
    {
        const double *e = enc + (size_t)j * IN;
        spo_guide_forward(P, e, out, acts);
        total += output_grad(out, wi + 3 * (size_t)j, w[j], g);
        if (!grad) continue;
        /* back through the output layer and the hidden layers */
        const double *gcur = g;
        int ncur = OUT;
        for (int l = 4; l >= 0; --l)
        {
            const double *hin = l > 0 ? acts + (l - 1) * WD : e;
            double *dW = grad + w_off(l), *db = grad + b_off(l);
            const double *W = P + w_off(l);
            for (int o = 0; o < ncur; ++o)
            {
                db[o] += gcur[o];
                for (int i = 0; i < WD; ++i) dW[o * WD + i] += gcur[o] * hin[i];
            }
            if (l == 0) break;
            for (int i = 0; i < WD; ++i)
            {
                double s = 0.0;
                for (int o = 0; o < ncur; ++o) s += W[o * WD + i] * gcur[o];
                gh[i] = hin[i] > 0.0 ? s : 0.0;
            }
            memcpy(gz, gh, sizeof gz);
            gcur = gz;
            ncur = rows(l - 1);
        }
    }
    if (grad)
        for (uint32_t k = 0; k < SPO_GUIDE_PARAMS; ++k) grad[k] /= (double)n;
    return total / (double)n;
}

void spo_guide_adam(float *p, float *m, float *v, const double *g,
                    uint32_t count, double lr, double b1, double b2, double eps,
                    uint32_t t)
{
    double c1 = 1.0 - pow(b1, (double)t), c2 = 1.0 - pow(b2, (double)t);
    for (uint32_t k = 0; k < count; ++k)
    {
        double mk = b1 * m[k] + (1.0 - b1) * g[k];
        double vk = b2 * v[k] + (1.0 - b2) * g[k] * g[k];
        m[k] = (float)mk;
        v[k] = (float)vk;
        p[k] = (float)(p[k] - lr * (mk / c1) / (sqrt(vk / c2) + eps));
    }
}
