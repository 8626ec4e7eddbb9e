// FP64 reference of the learned proposal's network (test infrastructure only): the input
// encoding, the forward pass, the vMF-mixture negative log-likelihood and its analytic
// gradient. The renderer's BF16 MFMA kernels (gpuupbp_amd/csrc/device/guide_mlp.hip) are
// checked against it; tests/test_mlp_cpu.py pins the analytic gradient by central finite
// differences. There is no counterpart in the reference renderer.
//
// Network: 51 inputs (48 Fourier features of the position in [0,1]^3 -- sin and cos of
// 2^k pi x_j, k = 0..7, feature (j * 8 + k) * 2 + {0: sin, 1: cos} -- and the 3 components of
// the arrival direction), three hidden layers of 64 ReLU units, 40 outputs: per lobe l the
// raw mean direction o[5l..5l+2], the raw concentration o[5l+3] and the mixture logit o[5l+4].
// Lobes: mu = normalise(raw), kappa = softplus(raw) + 0.01, weights = softmax(logits).
// Parameters, flat: W1[64][51], b1[64], W2[64][64], b2[64], W3[64][64], b3[64], W4[40][64], b4[40].
#include "oracle.h"

#include <cmath>
#include <cstring>
#include <vector>

namespace {

constexpr int kIn = 51, kHid = 64, kOut = 40, kLobes = 8;
constexpr double kKappaMin = 0.01, kPiD = 3.14159265358979323846;

struct Layout {
    size_t w[4], b[4], total;
    int rows[4], cols[4];
    Layout()
    {
        const int r[4] = {kHid, kHid, kHid, kOut}, c[4] = {kIn, kHid, kHid, kHid};
        size_t at = 0;
        for (int l = 0; l < 4; l++) {
            rows[l] = r[l];
            cols[l] = c[l];
            w[l] = at;
            at += size_t(r[l]) * size_t(c[l]);
            b[l] = at;
            at += size_t(r[l]);
        }
        total = at;
    }
};
const Layout L;

void encode(const double* x, const double* d, double* in)
{
    for (int j = 0; j < 3; j++)
        for (int k = 0; k < 8; k++) {
            const double a = std::ldexp(1.0, k) * kPiD * x[j];
            in[(j * 8 + k) * 2] = std::sin(a);
            in[(j * 8 + k) * 2 + 1] = std::cos(a);
        }
    for (int j = 0; j < 3; j++) in[48 + j] = d[j];
}

// activations of one sample: a[0] input, a[1..3] hidden (post-ReLU), a[4] output
struct Acts {
    std::vector<double> a[5];
};

void forward(const double* p, const double* in, Acts& A)
{
    A.a[0].assign(in, in + kIn);
    for (int l = 0; l < 4; l++) {
        const int R = L.rows[l], C = L.cols[l];
        A.a[l + 1].assign(size_t(R), 0.0);
        for (int r = 0; r < R; r++) {
            double s = p[L.b[l] + size_t(r)];
            for (int c = 0; c < C; c++) s += p[L.w[l] + size_t(r) * C + c] * A.a[l][size_t(c)];
            A.a[l + 1][size_t(r)] = l < 3 ? (s > 0.0 ? s : 0.0) : s;
        }
    }
}

double softplus(double v) { return v > 30.0 ? v : std::log1p(std::exp(v)); }
double sigmoid(double v) { return 1.0 / (1.0 + std::exp(-v)); }

// -log q(w) of the mixture the raw outputs o describe, and d(-log q)/do
double nll(const double* o, const double* w, double* grad)
{
    double mu[kLobes][3], norm[kLobes], kappa[kLobes], pi[kLobes], v[kLobes];
    double mx = -1e300;
    for (int l = 0; l < kLobes; l++) mx = std::fmax(mx, o[5 * l + 4]);
    double z = 0.0;
    for (int l = 0; l < kLobes; l++) z += std::exp(o[5 * l + 4] - mx);
    double q = 0.0;
    for (int l = 0; l < kLobes; l++) {
        norm[l] = std::sqrt(o[5 * l] * o[5 * l] + o[5 * l + 1] * o[5 * l + 1] + o[5 * l + 2] * o[5 * l + 2]) + 1e-12;
        for (int j = 0; j < 3; j++) mu[l][j] = o[5 * l + j] / norm[l];
        kappa[l] = softplus(o[5 * l + 3]) + kKappaMin;
        pi[l] = std::exp(o[5 * l + 4] - mx) / z;
        const double cdot = mu[l][0] * w[0] + mu[l][1] * w[1] + mu[l][2] * w[2];
        v[l] = kappa[l] / (2.0 * kPiD * -std::expm1(-2.0 * kappa[l])) * std::exp(kappa[l] * (cdot - 1.0));
        q += pi[l] * v[l];
    }
    if (grad) {
        for (int l = 0; l < kLobes; l++) {
            const double g = pi[l] * v[l] / q; // responsibility
            const double cdot = mu[l][0] * w[0] + mu[l][1] * w[1] + mu[l][2] * w[2];
            grad[5 * l + 4] = pi[l] - g;
            const double dlogv_dk = 1.0 / kappa[l] - 2.0 / std::expm1(2.0 * kappa[l]) + (cdot - 1.0);
            grad[5 * l + 3] = -g * dlogv_dk * sigmoid(o[5 * l + 3]);
            for (int j = 0; j < 3; j++) grad[5 * l + j] = -g * kappa[l] * (w[j] - cdot * mu[l][j]) / norm[l];
        }
    }
    return -std::log(q);
}

} // namespace

extern "C" uint32_t oracle_mlp_param_count(void) { return uint32_t(L.total); }

extern "C" void oracle_mlp_encode(const float* x, const float* d, float* out)
{
    double xd[3] = {x[0], x[1], x[2]}, dd[3] = {d[0], d[1], d[2]}, in[kIn];
    encode(xd, dd, in);
    for (int i = 0; i < kIn; i++) out[i] = float(in[i]);
}

// samples: n x 6 (position in [0,1]^3, arrival direction) -> n x 40 raw outputs
extern "C" int oracle_mlp_forward(const double* params, const float* samples, uint32_t n, double* out)
{
    Acts A;
    double in[kIn];
    for (uint32_t i = 0; i < n; i++) {
        const double x[3] = {samples[6 * i], samples[6 * i + 1], samples[6 * i + 2]};
        const double d[3] = {samples[6 * i + 3], samples[6 * i + 4], samples[6 * i + 5]};
        encode(x, d, in);
        forward(params, in, A);
        for (int k = 0; k < kOut; k++) out[size_t(i) * kOut + k] = A.a[4][size_t(k)];
    }
    return 0;
}

// samples: n x 10 (position in [0,1]^3, arrival direction, scattered direction, weight W).
// loss = (1/n) sum_i W_i * -log q(w_i | x_i, d_i); grads (may be NULL) = d loss / d params.
extern "C" int oracle_mlp_loss_grad(const double* params, const float* samples, uint32_t n, double* loss,
                                    double* grads)
{
    Acts A;
    double in[kIn], go[kOut];
    std::vector<double> g(L.total, 0.0), delta[4];
    double total = 0.0;
    for (uint32_t i = 0; i < n; i++) {
        const float* s = samples + 10 * size_t(i);
        const double x[3] = {s[0], s[1], s[2]}, d[3] = {s[3], s[4], s[5]}, w[3] = {s[6], s[7], s[8]};
        const double W = s[9];
        encode(x, d, in);
        forward(params, in, A);
        total += W * nll(A.a[4].data(), w, grads ? go : nullptr);
        if (!grads) continue;
        // back-propagation: delta of layer l's pre-activation
        delta[3].assign(go, go + kOut);
        for (double& v : delta[3]) v *= W / double(n);
        for (int l = 3; l >= 0; l--) {
            const int R = L.rows[l], C = L.cols[l];
            for (int r = 0; r < R; r++) {
                g[L.b[l] + size_t(r)] += delta[l][size_t(r)];
                for (int c = 0; c < C; c++) g[L.w[l] + size_t(r) * C + c] += delta[l][size_t(r)] * A.a[l][size_t(c)];
            }
            if (l == 0) break;
            delta[l - 1].assign(size_t(C), 0.0);
            for (int c = 0; c < C; c++) {
                if (!(A.a[l][size_t(c)] > 0.0)) continue; // ReLU of the layer below
                double s = 0.0;
                for (int r = 0; r < R; r++) s += params[L.w[l] + size_t(r) * C + c] * delta[l][size_t(r)];
                delta[l - 1][size_t(c)] = s;
            }
        }
    }
    *loss = total / double(n);
    if (grads) memcpy(grads, g.data(), L.total * sizeof(double));
    return 0;
}

// One Adam step (Kingma and Ba 2015, Algorithm 1) in FP64 over n parameters: moments m, v updated
// in place with beta1 0.9, beta2 0.999, then params -= lr * m_hat / (sqrt(v_hat) + 1e-8) with the
// bias corrections m_hat = m / (1 - beta1^step), v_hat = v / (1 - beta2^step); step counts from 1.
extern "C" int oracle_adam_step(double* params, double* m, double* v, const double* grads, uint32_t n, int step,
                                double lr)
{
    if (!params || !m || !v || !grads || step < 1) return -1;
    const double b1 = 0.9, b2 = 0.999, eps = 1e-8;
    const double c1 = 1.0 - std::pow(b1, double(step)), c2 = 1.0 - std::pow(b2, double(step));
    for (uint32_t i = 0; i < n; i++) {
        m[i] = b1 * m[i] + (1.0 - b1) * grads[i];
        v[i] = b2 * v[i] + (1.0 - b2) * grads[i] * grads[i];
        params[i] -= lr * (m[i] / c1) / (std::sqrt(v[i] / c2) + eps);
    }
    return 0;
}
