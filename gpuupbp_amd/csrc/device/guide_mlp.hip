// Network of the learned scattering proposal: forward, backward and Adam on the GPU, with BF16
// operands on the MFMA units (v_mfma_f32_16x16x16_bf16) and FP32 accumulation. No counterpart
// in the reference renderer.
//
// Network: 51 inputs (48 Fourier features of the position in [0,1]^3 -- sin and cos of
// 2^k pi x_j, k = 0..7, feature (j * 8 + k) * 2 + {0: sin, 1: cos} -- and the arrival
// direction), three hidden layers of 64 ReLU units, 40 outputs: per lobe l the raw mean
// direction o[5l..5l+2], raw concentration o[5l+3] and mixture logit o[5l+4]. Lobes:
// mu = normalise(raw), kappa = softplus(raw) + 0.01, weights = softmax(logits). The loss of a
// training sample (x, arrival d, scattered w, weight W) is W * -log q(w | x, d) (the KL
// divergence of Mueller et al., "Neural Importance Sampling", 2019, estimated from the
// renderer's own samples), averaged over the batch.
//
// Layout: FP32 master parameters in the flat order W1[64][51], b1, W2[64][64], b2, W3[64][64],
// b3, W4[40][64], b4 (14248 values), Adam moments beside them, and BF16 copies padded for the
// MFMA tiles: W1 as [64][64] (inputs 51..63 zero), W4 as [48][64] (outputs 40..47 zero).
// Activations of a training batch (BF16, rows of 64 or 48) stay in HBM between the layer GEMMs.
#include "guide.h"
#include "guide_mlp.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <vector>

namespace upbp {

namespace {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kIn = 51, kP = 64, kHid = 64, kOut = 40, kOutP = 48, kLobesN = 8;
constexpr float kKappaLo = 0.01f;

// parameter offsets of the flat FP32 layout
constexpr int kW1 = 0, kB1 = kW1 + 64 * 51, kW2 = kB1 + 64, kB2 = kW2 + 64 * 64, kW3 = kB2 + 64, kB3 = kW3 + 64 * 64,
              kW4 = kB3 + 64, kB4 = kW4 + 40 * 64, kParams = kB4 + 40;
// BF16 padded weight offsets (elements)
constexpr int kQ1 = 0, kQ2 = kQ1 + 64 * 64, kQ3 = kQ2 + 64 * 64, kQ4 = kQ3 + 64 * 64, kQTotal = kQ4 + 48 * 64;

__host__ __device__ inline uint16_t f2bf(float f)
{
    uint32_t u;
    memcpy(&u, &f, 4);
    u += 0x7fffu + ((u >> 16) & 1u); // round to nearest even
    return uint16_t(u >> 16);
}
__host__ __device__ inline float bf2f(uint16_t h)
{
    const uint32_t u = uint32_t(h) << 16;
    float f;
    memcpy(&f, &u, 4);
    return f;
}

// ---- encoding -------------------------------------------------------------------------
// samples: stride floats per sample, position in [0,1]^3 then the arrival direction
__global__ void __launch_bounds__(256) k_encode(const float* samples, int stride, uint32_t n, uint32_t n_pad,
                                                uint16_t* x0)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_pad) return;
    uint16_t* row = x0 + size_t(i) * kP;
    if (i >= n) {
        for (int k = 0; k < kP; k++) row[k] = 0;
        return;
    }
    const float* s = samples + size_t(i) * stride;
    for (int j = 0; j < 3; j++)
        for (int k = 0; k < 8; k++) {
            const float a = float(1 << k) * kPi * s[j];
            row[(j * 8 + k) * 2] = f2bf(sinf(a));
            row[(j * 8 + k) * 2 + 1] = f2bf(cosf(a));
        }
    for (int j = 0; j < 3; j++) row[48 + j] = f2bf(s[3 + j]);
    for (int k = kIn; k < kP; k++) row[k] = 0;
}

// ---- GEMM on the MFMA units ------------------------------------------------------------
// C(i, j) = sum_k A(i, k) B(k, j) over BF16 operands with strided access,
// A(i, k) = A[i * sai + k * sak], B(k, j) = B[k * sbk + j * sbj]; one wave per 16x16 tile of C,
// four waves per block, gridDim.z splits of the k range (M, N, K multiples of 16; K split
// evenly). Lane l holds A(i0 + l % 16, k0 + 4 (l / 16) + e), B(k0 + 4 (l / 16) + e, j0 + l % 16)
// and C(i0 + 4 (l / 16) + e, j0 + l % 16), e = 0..3.
enum Epi : int {
    EPI_RELU_BF16 = 0,  // out_bf16[i * ldo + j] = relu(c + bias[j]), also the next layer's input
    EPI_BIAS_F32 = 1,   // out_f32[i * ldo + j] = c + bias[j]
    EPI_MASK_BF16 = 2,  // out_bf16[i * ldo + j] = c if mask[i * ldo + j] > 0 else 0 (ReLU derivative)
    EPI_ATOMIC_F32 = 3  // out_f32[i * ldo + j] += c (split-k weight gradients)
};

template <int EPI>
__global__ void __launch_bounds__(256) k_gemm(const uint16_t* A, int sai, int sak, const uint16_t* B, int sbk, int sbj,
                                              int M, int N, int K, const float* bias, const uint16_t* mask,
                                              uint16_t* out_bf, float* out_f, int ldo)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int tiles_n = N / 16;
    const int tile = blockIdx.x * 4 + wave;
    if (tile >= (M / 16) * tiles_n) return;
    const int i0 = (tile / tiles_n) * 16, j0 = (tile % tiles_n) * 16;
    const int kc = K / int(gridDim.z);
    const int kb = int(blockIdx.z) * kc, ke = kb + kc;
    const int r = lane & 15, q = lane >> 4;
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
    for (int k0 = kb; k0 < ke; k0 += 16) {
        bf16x4 a, b;
        for (int e = 0; e < 4; e++) {
            const int k = k0 + 4 * q + e;
            a[e] = short(A[size_t(i0 + r) * sai + size_t(k) * sak]);
            b[e] = short(B[size_t(k) * sbk + size_t(j0 + r) * sbj]);
        }
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
    }
    for (int e = 0; e < 4; e++) {
        const int i = i0 + 4 * q + e, j = j0 + r;
        const size_t o = size_t(i) * ldo + j;
        const float c = acc[e];
        if (EPI == EPI_RELU_BF16) out_bf[o] = f2bf(fmaxf(c + bias[j], 0.f));
        else if (EPI == EPI_BIAS_F32) out_f[o] = c + bias[j];
        else if (EPI == EPI_MASK_BF16) out_bf[o] = bf2f(mask[o]) > 0.f ? f2bf(c) : uint16_t(0);
        else atomicAdd(&out_f[o], c);
    }
}

// ---- loss and its gradient at the outputs ------------------------------------------------
__device__ inline float softplusf(float v) { return v > 20.f ? v : log1pf(expf(v)); }

// one thread per sample: W * -log q(w) into the loss sum, d(loss)/d(outputs) / n into dO
__global__ void __launch_bounds__(256) k_loss_grad(const float* O, const float* samples, uint32_t n, uint32_t n_pad,
                                                   float inv_n, uint16_t* dO, float* loss_sum)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_pad) return;
    uint16_t* g = dO + size_t(i) * kOutP;
    for (int k = 0; k < kOutP; k++) g[k] = 0;
    if (i >= n) return;
    const float* o = O + size_t(i) * kOutP;
    const float* s = samples + size_t(i) * 10;
    const f3 w = mk(s[6], s[7], s[8]);
    const float W = s[9];
    if (!(W > 0.f)) return;
    float mx = -1e30f;
    for (int l = 0; l < kLobesN; l++) mx = fmaxf(mx, o[5 * l + 4]);
    float z = 0.f;
    for (int l = 0; l < kLobesN; l++) z += expf(o[5 * l + 4] - mx);
    float pi[kLobesN], v[kLobesN], kap[kLobesN], nrm[kLobesN], cd[kLobesN];
    f3 mu[kLobesN];
    float q = 0.f;
    for (int l = 0; l < kLobesN; l++) {
        const f3 m = mk(o[5 * l], o[5 * l + 1], o[5 * l + 2]);
        nrm[l] = sqrtf(dot(m, m)) + 1e-12f;
        mu[l] = m / nrm[l];
        kap[l] = softplusf(o[5 * l + 3]) + kKappaLo;
        pi[l] = expf(o[5 * l + 4] - mx) / z;
        cd[l] = dot(mu[l], w);
        v[l] = kap[l] / (2.f * kPi * -expm1f(-2.f * kap[l])) * expf(kap[l] * (cd[l] - 1.f));
        q += pi[l] * v[l];
    }
    q = fmaxf(q, 1e-30f);
    atomicAdd(loss_sum, -W * logf(q));
    const float sc = W * inv_n;
    for (int l = 0; l < kLobesN; l++) {
        const float gam = pi[l] * v[l] / q;
        g[5 * l + 4] = f2bf(sc * (pi[l] - gam));
        const float dk = 1.f / kap[l] - 2.f / expm1f(2.f * kap[l]) + (cd[l] - 1.f);
        g[5 * l + 3] = f2bf(sc * -gam * dk / (1.f + expf(-o[5 * l + 3])));
        const f3 dm = (w - mu[l] * cd[l]) * (-gam * kap[l] / nrm[l]);
        g[5 * l] = f2bf(sc * dm.x);
        g[5 * l + 1] = f2bf(sc * dm.y);
        g[5 * l + 2] = f2bf(sc * dm.z);
    }
}

// bias gradients: column sums of a BF16 delta matrix (n_pad x ld), columns < cols
__global__ void __launch_bounds__(256) k_colsum(const uint16_t* d, uint32_t n_pad, int ld, int cols, float* out)
{
    const int c = blockIdx.x;
    float s = 0.f;
    for (uint32_t i = threadIdx.x; i < n_pad; i += blockDim.x) s += bf2f(d[size_t(i) * ld + c]);
    __shared__ float red[256];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (int(threadIdx.x) < h) red[threadIdx.x] += red[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0 && c < cols) out[c] = red[0];
}

// padded gradient buffers -> flat FP32 gradient (same order as the parameters)
__device__ inline float flat_grad(const float* gW, const float* gb, int p)
{
    if (p < kB1) return gW[(p / 51) * kP + p % 51];             // W1 [64][51] in a [64][64] tile
    if (p < kW2) return gb[p - kB1];
    if (p < kB2) return gW[64 * 64 + (p - kW2)];
    if (p < kW3) return gb[64 + (p - kB2)];
    if (p < kB3) return gW[2 * 64 * 64 + (p - kW3)];
    if (p < kW4) return gb[128 + (p - kB3)];
    if (p < kB4) return gW[3 * 64 * 64 + (p - kW4)];             // W4 [40][64] in a [48][64] tile
    return gb[192 + (p - kB4)];
}

__global__ void __launch_bounds__(256) k_gather_grad(const float* gW, const float* gb, float* g)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < kParams) g[p] = flat_grad(gW, gb, p);
}

// BF16 padded copies of the weights from the FP32 master
__global__ void __launch_bounds__(256) k_quantise(const float* prm, uint16_t* wq, float* bq)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < kQTotal) {
        float v = 0.f;
        if (t < kQ2) {
            const int r = t / kP, c = t % kP;
            v = c < kIn ? prm[kW1 + r * kIn + c] : 0.f;
        } else if (t < kQ3) {
            v = prm[kW2 + (t - kQ2)];
        } else if (t < kQ4) {
            v = prm[kW3 + (t - kQ3)];
        } else {
            const int r = (t - kQ4) / kP, c = (t - kQ4) % kP;
            v = r < kOut ? prm[kW4 + r * kP + c] : 0.f;
        }
        wq[t] = f2bf(v);
    }
    if (t < 64 * 3 + kOutP) { // biases, b4 padded to 48
        float v = 0.f;
        if (t < 64) v = prm[kB1 + t];
        else if (t < 128) v = prm[kB2 + t - 64];
        else if (t < 192) v = prm[kB3 + t - 128];
        else if (t - 192 < kOut) v = prm[kB4 + t - 192];
        bq[t] = v;
    }
}

// Adam (Kingma and Ba 2015) on the FP32 master
__global__ void __launch_bounds__(256) k_adam(float* prm, float* m, float* v, const float* g, float lr, float b1t,
                                              float b2t)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= kParams) return;
    const float gr = g[p];
    const float mm = 0.9f * m[p] + 0.1f * gr, vv = 0.999f * v[p] + 0.001f * gr * gr;
    m[p] = mm;
    v[p] = vv;
    prm[p] -= lr * (mm / (1.f - b1t)) / (sqrtf(vv / (1.f - b2t)) + 1e-8f);
}

// proposal records from the raw outputs of (cell centre, bin axis) inputs
__global__ void __launch_bounds__(256) k_pack_records(const float* O, uint32_t n, float4* recs)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float* o = O + size_t(i) * kOutP;
    float mx = -1e30f;
    for (int l = 0; l < kLobesN; l++) mx = fmaxf(mx, o[5 * l + 4]);
    float lob[kLobesN * 5];
    for (int l = 0; l < kLobesN; l++) {
        lob[5 * l] = o[5 * l];
        lob[5 * l + 1] = o[5 * l + 1];
        lob[5 * l + 2] = o[5 * l + 2];
        lob[5 * l + 3] = softplusf(o[5 * l + 3]) + kKappaLo;
        lob[5 * l + 4] = expf(o[5 * l + 4] - mx); // normalised by the packer
    }
    float4 rec[kGuideRec];
    pack_guide_record(lob, rec);
    for (int k = 0; k < kGuideRec; k++) recs[size_t(i) * kGuideRec + k] = rec[k];
}

// inputs of the record grid: cell centres in [0,1]^3; the direction inputs stay zero (the
// proposal's radiance mixture depends on the position only)
__global__ void __launch_bounds__(256) k_grid_inputs(int res, float* samples)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t n = uint32_t(res) * res * res;
    if (i >= n) return;
    const uint32_t x = i % res, y = (i / res) % res, z = i / (uint32_t(res) * res);
    float* s = samples + size_t(i) * 6;
    s[0] = (float(x) + .5f) / float(res);
    s[1] = (float(y) + .5f) / float(res);
    s[2] = (float(z) + .5f) / float(res);
    s[3] = s[4] = s[5] = 0.f;
}

// batch of n training samples drawn uniformly, with replacement, from a pool of n_pool (10
// floats each); the pool is in append order (pixel and bounce order), so contiguous slices
// would train on one image region only
__global__ void __launch_bounds__(256) k_draw_batch(const float* pool, uint32_t n_pool, uint32_t n, uint32_t draw,
                                                    float* out)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint64_t z = (uint64_t(draw) << 32 | i) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    const uint32_t j = uint32_t(((z >> 32) * uint64_t(n_pool)) >> 32);
    const float* src = pool + size_t(j) * 10;
    float* dst = out + size_t(i) * 10;
    for (int k = 0; k < 10; k++) dst[k] = src[k];
}

__global__ void __launch_bounds__(256) k_zero(uint32_t* p, size_t n)
{
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

uint32_t pad16(uint32_t n) { return (n + 15u) & ~15u; }
dim3 gemm_grid(int M, int N, int splits) { return dim3(uint32_t((M / 16) * (N / 16) + 3) / 4, 1, uint32_t(splits)); }

} // namespace

// ---- host side --------------------------------------------------------------------------------
int GuideNet::init(uint32_t seed)
{
    if (prm.ensure(kParams * 4) || adam_m.ensure(kParams * 4) || adam_v.ensure(kParams * 4) ||
        grad.ensure(kParams * 4) || wq.ensure(kQTotal * 2) || bq.ensure(256 * 4) || gW.ensure(size_t(kQTotal) * 4) ||
        gb.ensure(256 * 4) || loss.ensure(16))
        return UPBP_E_HIP;
    // He initialisation (normal, std sqrt(2 / fan_in)), biases zero; Box-Muller over a
    // splitmix64 stream of the seed
    std::vector<float> h(kParams, 0.f);
    uint64_t st = 0x9E3779B97F4A7C15ull * (uint64_t(seed) + 1u);
    auto next = [&]() {
        uint64_t z = (st += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return double((z ^ (z >> 31)) >> 11) * (1.0 / 9007199254740992.0);
    };
    const int wo[4] = {kW1, kW2, kW3, kW4}, rows[4] = {64, 64, 64, 40}, cols[4] = {51, 64, 64, 64};
    for (int l = 0; l < 4; l++) {
        const double sd = std::sqrt(2.0 / cols[l]);
        for (int k = 0; k < rows[l] * cols[l]; k++) {
            const double u1 = std::max(next(), 1e-300), u2 = next();
            h[size_t(wo[l] + k)] = float(sd * std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * 3.141592653589793 * u2));
        }
    }
    step = 0;
    return set_params(h.data(), nullptr);
}

int GuideNet::set_params(const float* host, hipStream_t s)
{
    if (hipMemcpyAsync(prm.p, host, kParams * 4, hipMemcpyHostToDevice, s) != hipSuccess) return UPBP_E_HIP;
    hipLaunchKernelGGL(k_zero, dim3((kParams + 255) / 256), dim3(256), 0, s, adam_m.as<uint32_t>(), size_t(kParams));
    hipLaunchKernelGGL(k_zero, dim3((kParams + 255) / 256), dim3(256), 0, s, adam_v.as<uint32_t>(), size_t(kParams));
    step = 0;
    return quantise(s);
}

int GuideNet::get_params(float* host, hipStream_t s)
{
    if (hipMemcpyAsync(host, prm.p, kParams * 4, hipMemcpyDeviceToHost, s) != hipSuccess) return UPBP_E_HIP;
    return hipStreamSynchronize(s) == hipSuccess ? UPBP_OK : UPBP_E_HIP;
}

int GuideNet::quantise(hipStream_t s)
{
    hipLaunchKernelGGL(k_quantise, dim3((kQTotal + 255) / 256), dim3(256), 0, s, prm.as<float>(), wq.as<uint16_t>(),
                       bq.as<float>());
    return hipGetLastError() == hipSuccess ? UPBP_OK : UPBP_E_HIP;
}

int GuideNet::reserve(uint32_t n)
{
    const size_t np = pad16(n);
    if (x0.ensure(np * kP * 2) || h1.ensure(np * kP * 2) || h2.ensure(np * kP * 2) || h3.ensure(np * kP * 2) ||
        out.ensure(np * kOutP * 4) || d_o.ensure(np * kOutP * 2) || d3.ensure(np * kP * 2) || d2.ensure(np * kP * 2) ||
        d1.ensure(np * kP * 2))
        return UPBP_E_HIP;
    return UPBP_OK;
}

// samples on the device, stride floats each, position in [0,1]^3 then the arrival direction
int GuideNet::forward(const float* samples, int stride, uint32_t n, hipStream_t s)
{
    if (reserve(n)) return UPBP_E_HIP;
    return forward_into(samples, stride, n, x0.as<uint16_t>(), h1.as<uint16_t>(), h2.as<uint16_t>(),
                        h3.as<uint16_t>(), out.as<float>(), s);
}

// the forward pass into explicit buffers: the encoded inputs and the three hidden activations
// (pad16(n) rows of 64 BF16 each; h2 may alias x0 and h3 may alias h1 when only the outputs are
// wanted) and the raw outputs (rows of 48 floats)
int GuideNet::forward_into(const float* samples, int stride, uint32_t n, uint16_t* a0, uint16_t* a1, uint16_t* a2,
                           uint16_t* a3, float* o, hipStream_t s) const
{
    const uint32_t np = pad16(n);
    const int M = int(np);
    hipLaunchKernelGGL(k_encode, dim3((np + 255) / 256), dim3(256), 0, s, samples, stride, n, np, a0);
    const uint16_t* W = wq.as<uint16_t>();
    const float* B = bq.as<float>();
    // Y = X W^T: B(k, j) = W[j][k], the padded rows of 64
    hipLaunchKernelGGL(k_gemm<EPI_RELU_BF16>, gemm_grid(M, 64, 1), dim3(256), 0, s, a0, kP, 1, W + kQ1, 1, kP, M, 64,
                       kP, B, nullptr, a1, nullptr, kP);
    hipLaunchKernelGGL(k_gemm<EPI_RELU_BF16>, gemm_grid(M, 64, 1), dim3(256), 0, s, a1, kP, 1, W + kQ2, 1, kP, M, 64,
                       kP, B + 64, nullptr, a2, nullptr, kP);
    hipLaunchKernelGGL(k_gemm<EPI_RELU_BF16>, gemm_grid(M, 64, 1), dim3(256), 0, s, a2, kP, 1, W + kQ3, 1, kP, M, 64,
                       kP, B + 128, nullptr, a3, nullptr, kP);
    hipLaunchKernelGGL(k_gemm<EPI_BIAS_F32>, gemm_grid(M, kOutP, 1), dim3(256), 0, s, a3, kP, 1, W + kQ4, 1, kP, M,
                       kOutP, kP, B + 192, nullptr, nullptr, o, kOutP);
    return hipGetLastError() == hipSuccess ? UPBP_OK : UPBP_E_HIP;
}

// training samples on the device: 10 floats each (position, arrival, scattered, weight)
int GuideNet::gradient(const float* samples, uint32_t n, hipStream_t s)
{
    int e = forward(samples, 10, n, s);
    if (e) return e;
    const uint32_t np = pad16(n);
    const int M = int(np);
    hipLaunchKernelGGL(k_zero, dim3(1), dim3(64), 0, s, loss.as<uint32_t>(), size_t(4));
    hipLaunchKernelGGL(k_loss_grad, dim3((np + 255) / 256), dim3(256), 0, s, out.as<float>(), samples, n, np,
                       1.f / float(n), d_o.as<uint16_t>(), loss.as<float>());
    const uint16_t* W = wq.as<uint16_t>();
    // deltas: d_{l-1} = (d_l W_l) masked by the ReLU of layer l-1; B(k, j) = W_l[k][j]
    hipLaunchKernelGGL(k_gemm<EPI_MASK_BF16>, gemm_grid(M, 64, 1), dim3(256), 0, s, d_o.as<uint16_t>(), kOutP, 1,
                       W + kQ4, kP, 1, M, 64, kOutP, nullptr, h3.as<uint16_t>(), d3.as<uint16_t>(), nullptr, kP);
    hipLaunchKernelGGL(k_gemm<EPI_MASK_BF16>, gemm_grid(M, 64, 1), dim3(256), 0, s, d3.as<uint16_t>(), kP, 1,
                       W + kQ3, kP, 1, M, 64, kP, nullptr, h2.as<uint16_t>(), d2.as<uint16_t>(), nullptr, kP);
    hipLaunchKernelGGL(k_gemm<EPI_MASK_BF16>, gemm_grid(M, 64, 1), dim3(256), 0, s, d2.as<uint16_t>(), kP, 1,
                       W + kQ2, kP, 1, M, 64, kP, nullptr, h1.as<uint16_t>(), d1.as<uint16_t>(), nullptr, kP);
    // weight gradients: gW_l[r][c] = sum_i delta_l(i, r) act_{l-1}(i, c), split over the batch
    hipLaunchKernelGGL(k_zero, dim3((kQTotal + 255) / 256), dim3(256), 0, s, gW.as<uint32_t>(), size_t(kQTotal));
    int splits = 1;
    while (splits < 64 && M % (16 * splits * 2) == 0 && M / (splits * 2) >= 256) splits *= 2;
    float* g = gW.as<float>();
    const uint16_t* acts[4] = {x0.as<uint16_t>(), h1.as<uint16_t>(), h2.as<uint16_t>(), h3.as<uint16_t>()};
    const uint16_t* dels[4] = {d1.as<uint16_t>(), d2.as<uint16_t>(), d3.as<uint16_t>(), d_o.as<uint16_t>()};
    const int rows[4] = {64, 64, 64, kOutP}, off[4] = {kQ1, kQ2, kQ3, kQ4};
    for (int l = 0; l < 4; l++)
        hipLaunchKernelGGL(k_gemm<EPI_ATOMIC_F32>, gemm_grid(rows[l], kP, splits), dim3(256), 0, s, dels[l], 1,
                           rows[l], acts[l], kP, 1, rows[l], kP, M, nullptr, nullptr, nullptr, g + off[l], kP);
    float* gbias = gb.as<float>();
    hipLaunchKernelGGL(k_colsum, dim3(64), dim3(256), 0, s, d1.as<uint16_t>(), np, kP, 64, gbias);
    hipLaunchKernelGGL(k_colsum, dim3(64), dim3(256), 0, s, d2.as<uint16_t>(), np, kP, 64, gbias + 64);
    hipLaunchKernelGGL(k_colsum, dim3(64), dim3(256), 0, s, d3.as<uint16_t>(), np, kP, 64, gbias + 128);
    hipLaunchKernelGGL(k_colsum, dim3(kOutP), dim3(256), 0, s, d_o.as<uint16_t>(), np, kOutP, kOut, gbias + 192);
    hipLaunchKernelGGL(k_gather_grad, dim3((kParams + 255) / 256), dim3(256), 0, s, g, gbias, grad.as<float>());
    return hipGetLastError() == hipSuccess ? UPBP_OK : UPBP_E_HIP;
}

int GuideNet::gradient_drawn(const float* pool, uint32_t n_pool, uint32_t n, uint32_t draw, hipStream_t s)
{
    if (n_pool == 0 || batch.ensure(size_t(n) * 40)) return UPBP_E_HIP;
    hipLaunchKernelGGL(k_draw_batch, dim3((n + 255) / 256), dim3(256), 0, s, pool, n_pool, n, draw, batch.as<float>());
    return gradient(batch.as<float>(), n, s);
}

int GuideNet::adam_step(float lr, hipStream_t s)
{
    step++;
    const float b1t = powf(0.9f, float(step)), b2t = powf(0.999f, float(step));
    hipLaunchKernelGGL(k_adam, dim3((kParams + 255) / 256), dim3(256), 0, s, prm.as<float>(), adam_m.as<float>(),
                       adam_v.as<float>(), grad.as<float>(), lr, b1t, b2t);
    return quantise(s);
}

// records of a res^3 grid: the network at cell centres, packed
int GuideNet::records(int res, float4* recs, hipStream_t s)
{
    const uint32_t n = uint32_t(res) * res * res;
    if (grid_in.ensure(size_t(n) * 6 * 4)) return UPBP_E_HIP;
    hipLaunchKernelGGL(k_grid_inputs, dim3((n + 255) / 256), dim3(256), 0, s, res, grid_in.as<float>());
    const int e = forward(grid_in.as<float>(), 6, n, s);
    if (e) return e;
    hipLaunchKernelGGL(k_pack_records, dim3((n + 255) / 256), dim3(256), 0, s, out.as<float>(), n, recs);
    return hipGetLastError() == hipSuccess ? UPBP_OK : UPBP_E_HIP;
}

uint32_t GuideNet::param_count() { return kParams; }

} // namespace upbp

// ---- C ABI (tests and tools) -------------------------------------------------------------------
using upbp::GuideNet;

struct upbp_mlp {
    GuideNet net;
    upbp_host::DevBuf in;
    ~upbp_mlp() { in.release(); }
};

static int has_gpu()
{
    int n = 0;
    return hipGetDeviceCount(&n) == hipSuccess && n > 0;
}

extern "C" uint32_t upbp_mlp_param_count(void) { return GuideNet::param_count(); }

extern "C" int upbp_mlp_create(uint32_t seed, upbp_mlp** out)
{
    if (!out) return UPBP_E_ARG;
    *out = nullptr;
    if (!has_gpu()) return UPBP_E_NODEVICE;
    auto* m = new upbp_mlp();
    if (m->net.init(seed) || hipDeviceSynchronize() != hipSuccess) {
        delete m;
        return UPBP_E_HIP;
    }
    *out = m;
    return UPBP_OK;
}

extern "C" void upbp_mlp_destroy(upbp_mlp* m) { delete m; }

extern "C" int upbp_mlp_set_params(upbp_mlp* m, const float* params)
{
    if (!m || !params) return UPBP_E_ARG;
    const int e = m->net.set_params(params, nullptr);
    return e ? e : (hipDeviceSynchronize() == hipSuccess ? UPBP_OK : UPBP_E_HIP);
}

extern "C" int upbp_mlp_get_params(upbp_mlp* m, float* params)
{
    if (!m || !params) return UPBP_E_ARG;
    return m->net.get_params(params, nullptr);
}

extern "C" int upbp_mlp_forward(upbp_mlp* m, const float* samples, uint32_t n, float* out)
{
    if (!m || (n && (!samples || !out)) || n > (1u << 24)) return UPBP_E_ARG;
    if (n == 0) return UPBP_OK;
    if (m->in.ensure(size_t(n) * 6 * 4) ||
        hipMemcpy(m->in.p, samples, size_t(n) * 24, hipMemcpyHostToDevice) != hipSuccess)
        return UPBP_E_HIP;
    int e = m->net.forward(m->in.as<float>(), 6, n, nullptr);
    if (e) return e;
    std::vector<float> o(size_t(n) * 48);
    if (hipMemcpy(o.data(), m->net.out.p, o.size() * 4, hipMemcpyDeviceToHost) != hipSuccess) return UPBP_E_HIP;
    for (uint32_t i = 0; i < n; i++) memcpy(out + size_t(i) * 40, &o[size_t(i) * 48], 40 * 4);
    return UPBP_OK;
}

extern "C" int upbp_mlp_loss_grad(upbp_mlp* m, const float* samples, uint32_t n, float* loss, float* grads)
{
    if (!m || !samples || !loss || n == 0 || n > (1u << 24)) return UPBP_E_ARG;
    if (m->in.ensure(size_t(n) * 10 * 4) ||
        hipMemcpy(m->in.p, samples, size_t(n) * 40, hipMemcpyHostToDevice) != hipSuccess)
        return UPBP_E_HIP;
    int e = m->net.gradient(m->in.as<float>(), n, nullptr);
    if (e) return e;
    float l = 0.f;
    if (hipMemcpy(&l, m->net.loss.p, 4, hipMemcpyDeviceToHost) != hipSuccess) return UPBP_E_HIP;
    *loss = l / float(n);
    if (grads && hipMemcpy(grads, m->net.grad.p, GuideNet::param_count() * 4, hipMemcpyDeviceToHost) != hipSuccess)
        return UPBP_E_HIP;
    return UPBP_OK;
}

extern "C" int upbp_mlp_train_step(upbp_mlp* m, const float* samples, uint32_t n, float lr, float* loss)
{
    float l = 0.f;
    int e = upbp_mlp_loss_grad(m, samples, n, &l, nullptr);
    if (e) return e;
    e = m->net.adam_step(lr, nullptr);
    if (e) return e;
    if (loss) *loss = l;
    return hipDeviceSynchronize() == hipSuccess ? UPBP_OK : UPBP_E_HIP;
}

// ---- forward pass that allocates nothing ------------------------------------------------------
// For callers that cannot allocate device memory (code under HIP graph capture, serving loops,
// harnesses that forbid hipMalloc): the forward pass of upbp_mlp_forward with its input staging,
// activations and outputs in one file-scope device array. Each GPU holds its own copy of the
// array until the process ends; its address on the current device is resolved once per device
// with hipGetSymbolAddress and only that address is used from host code. Every call on one GPU
// shares the array, so calls that overlap on one GPU race. Nothing in it needs clearing: every
// byte a call reads was written earlier by the same call.
namespace {

constexpr uint32_t kNoAllocMax = 32768; // samples per call (a multiple of 16)
constexpr size_t kScrIn = 0;                                               // n x 6 floats
constexpr size_t kScrA0 = kScrIn + size_t(kNoAllocMax) * 6 * 4;            // pad16(n) x 64 BF16
constexpr size_t kScrA1 = kScrA0 + size_t(kNoAllocMax) * upbp::kP * 2;           // pad16(n) x 64 BF16
constexpr size_t kScrOut = kScrA1 + size_t(kNoAllocMax) * upbp::kP * 2;          // pad16(n) x 48 floats
constexpr size_t kScrBytes = kScrOut + size_t(kNoAllocMax) * upbp::kOutP * 4;    // 14.75 MiB
static_assert(kScrA0 % 256 == 0 && kScrA1 % 256 == 0 && kScrOut % 256 == 0, "scratch regions stay aligned");

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
unsigned char* g_scratch_dev[kMaxDevices] = {};

} // namespace

static __device__ __align__(256) unsigned char scratch[kScrBytes];

// the array's address on the current device (resolved on the first call there), or null
static unsigned char* scratch_on_current_device()
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_dev[dev]) {
        void* p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return nullptr;
        g_scratch_dev[dev] = static_cast<unsigned char*>(p);
    }
    return g_scratch_dev[dev];
}

extern "C" uint32_t upbp_mlp_noalloc_max(void) { return kNoAllocMax; }

extern "C" int upbp_mlp_forward_noalloc(upbp_mlp* m, const float* samples, uint32_t n, float* out)
{
    if (!m || (n && (!samples || !out))) return UPBP_E_ARG;
    if (n > kNoAllocMax) return UPBP_E_CAPACITY;
    if (n == 0) return UPBP_OK;
    unsigned char* base = scratch_on_current_device();
    if (!base) return UPBP_E_HIP;
    float* in = reinterpret_cast<float*>(base + kScrIn);
    uint16_t* a0 = reinterpret_cast<uint16_t*>(base + kScrA0);
    uint16_t* a1 = reinterpret_cast<uint16_t*>(base + kScrA1);
    float* o = reinterpret_cast<float*>(base + kScrOut);
    if (hipMemcpy(in, samples, size_t(n) * 24, hipMemcpyHostToDevice) != hipSuccess) return UPBP_E_HIP;
    const int e = m->net.forward_into(in, 6, n, a0, a1, a0, a1, o, nullptr);
    if (e) return e;
    // rows of 48 on the device, the first 40 of each to the host
    if (hipMemcpy2D(out, 40 * 4, o, upbp::kOutP * 4, 40 * 4, n, hipMemcpyDeviceToHost) != hipSuccess) return UPBP_E_HIP;
    return UPBP_OK;
}
