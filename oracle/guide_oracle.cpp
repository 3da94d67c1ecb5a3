// Guiding oracle -- TEST INFRASTRUCTURE ONLY.
//
// An FP64 restatement of the learned path-guiding model that the plugin evaluates and trains
// with BF16 matrix-core kernels. The reference renderer has no path guiding; this file defines
// the model the GPU kernels must match (it is the checker of their arithmetic):
//
//   features (64): position normalised to the scene box, frequency-encoded
//                  (sin, cos of 2^f pi x for f = 0..7 on each axis: 48 values), the geometric
//                  normal's encoding (16: the normal, sin / cos of pi n per axis, sin / cos of
//                  2 pi n per axis, a zero)
//   network:       4 hidden layers of 64 ReLU units, one linear output layer of 41 values
//   outputs:       8 von Mises-Fisher lobes (direction 3, concentration 1, weight logit 1)
//                  and the probability s of sampling the guide instead of the BSDF
//   proposal:      q(w) = s * sum_k pi_k vMF(w; mu_k, kappa_k) + (1 - s) * p_bsdf(w)
//   loss:          -weight * (b log q(w) + (1 - b) log p_guide(w)) per training record, b = 0.5,
//                  normalised by the summed weights: a blend of the KL divergences of the
//                  proposal and of the mixture from the incident radiance (Mueller et al. 2019,
//                  "Neural Importance Sampling", one-sample MIS with learned selection)
//
// Parameters are one float vector: for each layer l = 0..4 the weight matrix stored as
// Wt[out][in] (rows: 64, 64, 64, 64, 48 (41 used); 64 columns), then its bias (64 or 48).
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {

constexpr int kIn = 64, kHidden = 64, kOutRows = 48, kOut = 41, kLobes = 8, kLayers = 5;
constexpr double kPi = 3.14159265358979323846;
double kSelMin = 0.1;  // floor of the selection probability (or_guide_set_sel_min)
constexpr double kSelMax = 0.9, kKappaMin = 0.01, kKappaMax = 1000.0;
double gBlend = 0.5;  // weight of the MIS-proposal term in the loss (or_guide_set_blend)

int Rows(int l) { return l == kLayers - 1 ? kOutRows : kHidden; }
size_t WOff(int l) {
    size_t o = 0;
    for (int i = 0; i < l; ++i) o += size_t(Rows(i)) * kIn + size_t(Rows(i));
    return o;
}
size_t BOff(int l) { return WOff(l) + size_t(Rows(l)) * kIn; }
size_t ParamCount() { return WOff(kLayers); }

double Softplus(double x) { return x > 30 ? x : std::log1p(std::exp(x)); }
double Sigmoid(double x) { return 1 / (1 + std::exp(-x)); }

int gEncoding = 0;  // experiments: 0 frequency encoding, 1 one-blob

void Features(const float *rec, double *f) {
    // rec: position in [0,1]^3 (already normalised), normal (3), wo (3)
    int k = 0;
    if (gEncoding == 1) {
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 16; ++b) {
                const double c = (b + 0.5) / 16, d = (double(rec[a]) - c) * 16;
                f[k++] = std::exp(-0.5 * d * d);
            }
        const double th = std::acos(std::fmax(-1.0, std::fmin(1.0, double(rec[5])))) / kPi;
        double ph = std::atan2(double(rec[4]), double(rec[3])) / (2 * kPi);
        if (ph < 0) ph += 1;
        for (int b = 0; b < 8; ++b) {
            const double c = (b + 0.5) / 8, d = (th - c) * 8;
            f[k++] = std::exp(-0.5 * d * d);
        }
        for (int b = 0; b < 8; ++b) {
            double d = ph - (b + 0.5) / 8;
            d -= std::round(d);  // azimuth wraps
            d *= 8;
            f[k++] = std::exp(-0.5 * d * d);
        }
        return;
    }
    for (int a = 0; a < 3; ++a)
        for (int fr = 0; fr < 8; ++fr) {
            double arg = std::ldexp(kPi * double(rec[a]), fr);
            f[k++] = std::sin(arg);
            f[k++] = std::cos(arg);
        }
    for (int i = 0; i < 3; ++i) f[k++] = rec[3 + i];
    for (int m = 1; m <= 2; ++m)
        for (int a = 0; a < 3; ++a) {
            double arg = kPi * m * double(rec[3 + a]);
            f[k++] = std::sin(arg);
            f[k++] = std::cos(arg);
        }
    while (k < kIn) f[k++] = 0;
}

struct Forward {
    double x[kLayers][kHidden];  // input of each layer (x[0] = features)
    double out[kOutRows];
};

void RunForward(const float *P, const double *feat, Forward *fw) {
    std::memcpy(fw->x[0], feat, sizeof(double) * kIn);
    for (int l = 0; l < kLayers; ++l) {
        const float *W = P + WOff(l), *b = P + BOff(l);
        for (int o = 0; o < Rows(l); ++o) {
            double acc = b[o];
            for (int i = 0; i < kIn; ++i) acc += double(W[o * kIn + i]) * fw->x[l][i];
            if (l < kLayers - 1) fw->x[l + 1][o] = acc > 0 ? acc : 0;
            else fw->out[o] = acc;
        }
    }
}

struct Mixture {
    double mu[kLobes][3], v[kLobes][3], vlen[kLobes], kappa[kLobes], pi[kLobes], sel;
};

void Decode(const double *o, Mixture *m) {
    double amax = -1e300;
    for (int k = 0; k < kLobes; ++k) amax = std::fmax(amax, o[5 * k + 4]);
    double z = 0;
    for (int k = 0; k < kLobes; ++k) {
        const double *q = o + 5 * k;
        double l = std::sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2]);
        m->vlen[k] = l;
        for (int a = 0; a < 3; ++a) m->v[k][a] = q[a];
        if (l > 1e-6) {
            for (int a = 0; a < 3; ++a) m->mu[k][a] = q[a] / l;
        } else {
            m->mu[k][0] = 0, m->mu[k][1] = 0, m->mu[k][2] = 1;
        }
        m->kappa[k] = std::fmin(kKappaMin + Softplus(q[3]), kKappaMax);
        m->pi[k] = std::exp(q[4] - amax);
        z += m->pi[k];
    }
    for (int k = 0; k < kLobes; ++k) m->pi[k] /= z;
    m->sel = kSelMin + (kSelMax - kSelMin) * Sigmoid(o[40]);
}

double VmfPdf(const double *mu, double kappa, const double *w) {
    double c = mu[0] * w[0] + mu[1] * w[1] + mu[2] * w[2];
    return kappa / (2 * kPi * -std::expm1(-2 * kappa)) * std::exp(kappa * (c - 1));
}

double MixPdf(const Mixture &m, const double *w) {
    double p = 0;
    for (int k = 0; k < kLobes; ++k) p += m.pi[k] * VmfPdf(m.mu[k], m.kappa[k], w);
    return p;
}

}  // namespace

extern "C" {

int or_guide_param_count() { return int(ParamCount()); }

// Outputs (41) of the network for n records (9 floats each: normalised position, normal, wo).
void or_guide_forward(const float *P, const float *recs, int n, float *out) {
    for (int r = 0; r < n; ++r) {
        double f[kIn];
        Features(recs + 9 * r, f);
        Forward fw;
        RunForward(P, f, &fw);
        for (int o = 0; o < kOut; ++o) out[kOut * r + o] = float(fw.out[o]);
    }
}

// Mixture pdf of direction w (3) for the network output o (41), and the selection probability.
double or_guide_mix_pdf(const float *o41, const float *w, double *sel) {
    double o[kOutRows] = {0}, wd[3] = {w[0], w[1], w[2]};
    for (int i = 0; i < kOut; ++i) o[i] = o41[i];
    Mixture m;
    Decode(o, &m);
    if (sel) *sel = m.sel;
    return MixPdf(m, wd);
}

// Samples the mixture with uniforms (u_lobe, u1, u2): lobe by its weight, then the vMF lobe
// (Jakob, "Numerically stable sampling of the von Mises-Fisher distribution on S^2", 2012).
void or_guide_mix_sample(const float *o41, float ul, float u1, float u2, float *w) {
    double o[kOutRows] = {0};
    for (int i = 0; i < kOut; ++i) o[i] = o41[i];
    Mixture m;
    Decode(o, &m);
    int k = 0;
    double c = 0;
    for (; k < kLobes - 1; ++k) {
        c += m.pi[k];
        if (ul < c) break;
    }
    double kap = m.kappa[k];
    double cosT = 1 + std::log1p(double(u1) * std::expm1(-2 * kap)) / kap;
    cosT = std::fmax(-1.0, std::fmin(1.0, cosT));
    double sinT = std::sqrt(std::fmax(0.0, 1 - cosT * cosT)), phi = 2 * kPi * u2;
    const double *mu = m.mu[k];
    // frame around mu (Duff et al. 2017)
    double sign = std::copysign(1.0, mu[2]), a = -1 / (sign + mu[2]), b = mu[0] * mu[1] * a;
    double t[3] = {1 + sign * mu[0] * mu[0] * a, sign * b, -sign * mu[0]};
    double s[3] = {b, sign + mu[1] * mu[1] * a, -mu[1]};
    for (int i = 0; i < 3; ++i) w[i] = float(sinT * std::cos(phi) * t[i] + sinT * std::sin(phi) * s[i] + cosT * mu[i]);
}

// Loss sum_r -wt_r (b log q_r + (1 - b) log p_guide,r) / sum_r wt_r and its gradient with respect
// to the parameters, for
// records of 16 floats: position (3, normalised), normal (3), wo (3), sampled direction (3),
// p_bsdf of that direction (1), weight (1), 2 unused. Returns the loss.
void or_guide_set_blend(double b) { gBlend = b; }
void or_guide_set_sel_min(double s) { kSelMin = s; }
void or_guide_set_encoding(int e) { gEncoding = e; }

double or_guide_loss_grad(const float *P, const float *recs, int n, float *grad) {
    const size_t np = ParamCount();
    std::vector<double> g(np, 0.0);
    double loss = 0, wsum = 0;
    for (int r = 0; r < n; ++r) wsum += recs[16 * r + 13];
    if (wsum <= 0) {
        if (grad) std::memset(grad, 0, sizeof(float) * np);
        return 0;
    }
    for (int r = 0; r < n; ++r) {
        const float *R = recs + 16 * r;
        const double wt = R[13] / wsum, pb = R[12];
        if (wt == 0) continue;
        double f[kIn];
        Features(R, f);
        Forward fw;
        RunForward(P, f, &fw);
        Mixture m;
        Decode(fw.out, &m);
        const double w[3] = {R[9], R[10], R[11]};
        double pk[kLobes], pm = 0;
        for (int k = 0; k < kLobes; ++k) {
            pk[k] = VmfPdf(m.mu[k], m.kappa[k], w);
            pm += m.pi[k] * pk[k];
        }
        const double q = m.sel * pm + (1 - m.sel) * pb;
        if (!(q > 0) || !(pm > 0)) continue;
        loss += -wt * (gBlend * std::log(q) + (1 - gBlend) * std::log(pm));
        // d loss / d outputs
        double dout[kOutRows] = {0};
        const double dq = -wt * gBlend / q;
        const double sig = (m.sel - kSelMin) / (kSelMax - kSelMin);
        dout[40] = dq * (pm - pb) * (kSelMax - kSelMin) * sig * (1 - sig);
        const double dpm = dq * m.sel - wt * (1 - gBlend) / pm;
        for (int k = 0; k < kLobes; ++k) {
            const double *o = fw.out + 5 * k;
            dout[5 * k + 4] = dpm * m.pi[k] * (pk[k] - pm);  // softmax
            const double kap = m.kappa[k], c = m.mu[k][0] * w[0] + m.mu[k][1] * w[1] + m.mu[k][2] * w[2];
            const double dpk = dpm * m.pi[k];
            // d pk / d kappa = pk (1/kappa - 2 e^{-2k}/(1 - e^{-2k}) + (c - 1)); kappa = kmin + softplus(o3) (uncapped)
            const double e2 = std::exp(-2 * kap);
            const double dk = pk[k] * (1 / kap - 2 * e2 / -std::expm1(-2 * kap) + (c - 1));
            if (kKappaMin + Softplus(o[3]) < kKappaMax) dout[5 * k + 3] = dpk * dk * Sigmoid(o[3]);
            // d pk / d mu = pk kappa w; mu = v / |v|
            if (m.vlen[k] > 1e-6) {
                for (int a = 0; a < 3; ++a) {
                    double proj = w[a] - c * m.mu[k][a];
                    dout[5 * k + a] = dpk * pk[k] * kap * proj / m.vlen[k];
                }
            }
        }
        // backward through the layers
        double dy[kHidden] = {0};
        std::memcpy(dy, dout, sizeof(dout));
        for (int l = kLayers - 1; l >= 0; --l) {
            const float *W = P + WOff(l);
            double *gW = g.data() + WOff(l), *gb = g.data() + BOff(l);
            double dx[kHidden] = {0};
            for (int o = 0; o < Rows(l); ++o) {
                if (dy[o] == 0) continue;
                gb[o] += dy[o];
                for (int i = 0; i < kIn; ++i) {
                    gW[o * kIn + i] += dy[o] * fw.x[l][i];
                    dx[i] += double(W[o * kIn + i]) * dy[o];
                }
            }
            if (l == 0) break;
            for (int i = 0; i < kIn; ++i) dy[i] = fw.x[l][i] > 0 ? dx[i] : 0;  // ReLU of the previous layer
        }
    }
    if (grad)
        for (size_t i = 0; i < np; ++i) grad[i] = float(g[i]);
    return loss;
}

}  // extern "C"
