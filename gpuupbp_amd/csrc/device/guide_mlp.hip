// Network of the learned scattering proposal: forward, backward and Adam on the GPU, with BF16
// operands on the MFMA units (v_mfma_f32_16x16x32_bf16) and FP32 accumulation: one fused kernel
// for the forward pass (encoding and all four layers per 16-row tile, weights and activations in
// LDS), one for the backward deltas of all layers, and one per layer for the weight gradients
// (batch chunks staged in LDS). No counterpart in the reference renderer.
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
// A training batch's activations and deltas (BF16, rows of 64 or 48) go to HBM once each, for the
// weight gradients.
#include "guide.h"
#include "guide_mlp.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <vector>

namespace upbp {

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
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

// ---- fused forward pass -----------------------------------------------------------------
// One wave takes 16 rows at a time: it encodes them into its LDS tile (16 rows of 64 BF16), then
// runs the four layers on v_mfma_f32_16x16x32_bf16, each layer reading its input tile as A
// fragments and writing its activations (bias, ReLU, BF16) back over the same tile. The block's
// waves share the weights, copied once into LDS. Rows are padded to 72 BF16 in LDS so that the
// 16-byte fragment loads of 16 rows at one column spread over the banks.
// Lane l holds A(l % 16, 8 (l / 16) + e) and B(8 (l / 16) + e, l % 16), e = 0..7 (A and B share the
// k order, so the product does not depend on how the instruction orders k within a lane), and
// C(4 (l / 16) + e, l % 16), e = 0..3. With KEEP the encoded inputs and the three hidden layers
// are also stored (rows of 64 BF16) for the backward pass.
constexpr int kLdsRow = 72;
constexpr int kWRows = 3 * kP + kOutP; // the padded BF16 weights: 240 rows of 64

__device__ __forceinline__ void wave_lds_fence()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

template <bool KEEP>
__global__ void __launch_bounds__(256) k_mlp_forward(const float* samples, int stride, uint32_t n, uint32_t np,
                                                     const uint16_t* wq, const float* bq, uint16_t* x0, uint16_t* h1,
                                                     uint16_t* h2, uint16_t* h3, float* out)
{
    __shared__ __align__(16) uint16_t W[kWRows * kLdsRow];
    __shared__ float Bs[256];
    __shared__ __align__(16) uint16_t act[4][16 * kLdsRow];
    for (int i = threadIdx.x; i < kWRows * 8; i += 256)
        *reinterpret_cast<uint4*>(&W[(i >> 3) * kLdsRow + (i & 7) * 8]) = reinterpret_cast<const uint4*>(wq)[i];
    if (threadIdx.x < 3 * kP + kOutP) Bs[threadIdx.x] = bq[threadIdx.x];
    __syncthreads();
    const uint32_t lane = threadIdx.x & 63u, w = threadIdx.x >> 6;
    const uint32_t r = lane & 15u, q = lane >> 4;
    uint16_t* A = act[w];
    uint16_t* keep[4] = {x0, h1, h2, h3};
    const uint32_t tiles = np / 16u;
    for (uint32_t t = blockIdx.x * 4u + w; t < tiles; t += gridDim.x * 4u) {
        // encoding: lane (r, q) writes columns [16 q, 16 q + 16) of tile row r: for q < 3 the sin and
        // cos of 2^k pi x_q, k = 0..7 (features (q * 8 + k) * 2 + {0, 1}), for q = 3 the arrival
        // direction and zeros
        const uint32_t row = t * 16u + r;
        uint16_t* dst = A + r * kLdsRow + 16u * q;
        if (row < n) {
            const float* smp = samples + size_t(row) * stride;
            if (q < 3u) {
                const float x = smp[q];
                for (int k = 0; k < 8; k++) {
                    const float a = float(1 << k) * kPi * x;
                    dst[2 * k] = f2bf(sinf(a));
                    dst[2 * k + 1] = f2bf(cosf(a));
                }
            } else {
                for (int k = 0; k < 16; k++) dst[k] = k < 3 ? f2bf(smp[3 + k]) : uint16_t(0);
            }
        } else {
            for (int k = 0; k < 16; k++) dst[k] = 0;
        }
        for (int L = 0; L < 4; L++) {
            wave_lds_fence();
            if (KEEP) { // the layer's input tile to HBM: 16 rows x 128 bytes, 32 bytes per lane
                const uint32_t rr = lane >> 2, cc = (lane & 3u) * 16u;
                const uint4* src = reinterpret_cast<const uint4*>(A + rr * kLdsRow + cc);
                uint4* to = reinterpret_cast<uint4*>(keep[L] + size_t(t * 16u + rr) * kP + cc);
                to[0] = src[0];
                to[1] = src[1];
            }
            const bf16x8 a0 = *reinterpret_cast<const bf16x8*>(A + r * kLdsRow + 8u * q);
            const bf16x8 a1 = *reinterpret_cast<const bf16x8*>(A + r * kLdsRow + 32u + 8u * q);
            wave_lds_fence(); // every lane holds its fragments before the tile is overwritten
            const uint16_t* Wl = W + L * kP * kLdsRow;
            const int nt = L < 3 ? kP / 16 : kOutP / 16;
            for (int j = 0; j < nt; j++) {
                const uint16_t* wr = Wl + (16 * j + int(r)) * kLdsRow + 8 * int(q);
                const bf16x8 b0 = *reinterpret_cast<const bf16x8*>(wr);
                const bf16x8 b1 = *reinterpret_cast<const bf16x8*>(wr + 32);
                f32x4 acc = {0.f, 0.f, 0.f, 0.f};
                acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a0, b0, acc, 0, 0, 0);
                acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1, b1, acc, 0, 0, 0);
                const int c = 16 * j + int(r);
                const float bias = Bs[kP * L + c];
                for (int e = 0; e < 4; e++) {
                    const int i = 4 * int(q) + e;
                    const float v = acc[e] + bias;
                    if (L < 3) A[i * kLdsRow + c] = f2bf(fmaxf(v, 0.f));
                    else out[size_t(t * 16u + uint32_t(i)) * kOutP + c] = v;
                }
            }
        }
    }
}

// ---- fused backward deltas ---------------------------------------------------------------
// Per 16-row tile: d3 = (dO W4) o [h3 > 0], d2 = (d3 W3) o [h2 > 0], d1 = (d2 W2) o [h1 > 0] on
// v_mfma_f32_16x16x32_bf16, the delta tile staying in LDS from layer to layer. B(k, j) = W_l[k][j]
// is read from the transposed weights, staged once per block in LDS (W4's 48 output rows padded
// to 64 with zeros, as the delta of the outputs is); the ReLU masks are the forward pass's stored
// activations, staged per tile. Each delta tile also goes to HBM for the weight gradients.
__global__ void __launch_bounds__(256) k_mlp_backward(const uint16_t* wq, const uint16_t* dO, const uint16_t* h1,
                                                      const uint16_t* h2, const uint16_t* h3, uint32_t np,
                                                      uint16_t* d1, uint16_t* d2, uint16_t* d3)
{
    __shared__ __align__(16) uint16_t WT[3][kP * kLdsRow];
    __shared__ __align__(16) uint16_t D[4][16 * kLdsRow];
    __shared__ __align__(16) uint16_t H[4][16 * kLdsRow];
    for (int idx = threadIdx.x; idx < 3 * kP * kP; idx += 256) { // WT[m][j][k] = W_l[k][j], l = 4, 3, 2
        const int m = idx / (kP * kP), e = idx % (kP * kP), k = e / kP, j = e % kP;
        const int off = m == 0 ? kQ4 : (m == 1 ? kQ3 : kQ2);
        WT[m][j * kLdsRow + k] = (m == 0 && k >= kOutP) ? uint16_t(0) : wq[off + k * kP + j];
    }
    __syncthreads();
    const uint32_t lane = threadIdx.x & 63u, w = threadIdx.x >> 6;
    const uint32_t r = lane & 15u, q = lane >> 4;
    const uint32_t rr = lane >> 2, cc = (lane & 3u) * 16u; // the tile copies: 16 rows x 4 x 16 columns
    uint16_t* Dt = D[w];
    uint16_t* Ht = H[w];
    const uint16_t* hs[3] = {h3, h2, h1};
    uint16_t* ds[3] = {d3, d2, d1};
    const uint32_t tiles = np / 16u;
    for (uint32_t t = blockIdx.x * 4u + w; t < tiles; t += gridDim.x * 4u) {
        const size_t row = size_t(t) * 16u + rr;
        { // the outputs' delta: 48 columns, zeros beyond
            uint4* to = reinterpret_cast<uint4*>(Dt + rr * kLdsRow + cc);
            if (cc < uint32_t(kOutP)) {
                const uint4* src = reinterpret_cast<const uint4*>(dO + row * kOutP + cc);
                to[0] = src[0];
                to[1] = src[1];
            } else {
                to[0] = to[1] = make_uint4(0u, 0u, 0u, 0u);
            }
        }
        for (int L = 0; L < 3; L++) {
            {
                const uint4* src = reinterpret_cast<const uint4*>(hs[L] + row * kP + cc);
                uint4* to = reinterpret_cast<uint4*>(Ht + rr * kLdsRow + cc);
                to[0] = src[0];
                to[1] = src[1];
            }
            wave_lds_fence();
            const bf16x8 a0 = *reinterpret_cast<const bf16x8*>(Dt + r * kLdsRow + 8u * q);
            const bf16x8 a1 = *reinterpret_cast<const bf16x8*>(Dt + r * kLdsRow + 32u + 8u * q);
            wave_lds_fence();
            for (int j = 0; j < kP / 16; j++) {
                const uint16_t* wr = WT[L] + (16 * j + int(r)) * kLdsRow + 8 * int(q);
                f32x4 acc = {0.f, 0.f, 0.f, 0.f};
                acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a0, *reinterpret_cast<const bf16x8*>(wr), acc, 0, 0, 0);
                acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1, *reinterpret_cast<const bf16x8*>(wr + 32), acc, 0, 0,
                                                              0);
                const int c = 16 * j + int(r);
                for (int e = 0; e < 4; e++) {
                    const int i = 4 * int(q) + e;
                    Dt[i * kLdsRow + c] = bf2f(Ht[i * kLdsRow + c]) > 0.f ? f2bf(acc[e]) : uint16_t(0);
                }
            }
            wave_lds_fence();
            const uint4* src = reinterpret_cast<const uint4*>(Dt + rr * kLdsRow + cc);
            uint4* to = reinterpret_cast<uint4*>(ds[L] + row * kP + cc);
            to[0] = src[0];
            to[1] = src[1];
        }
    }
}

// ---- weight gradients -------------------------------------------------------------------
// gW[i][j] += sum over the batch rows of delta(row, i) act(row, j) for one layer (delta rows of
// `units` BF16, act rows of 64): each block sums `chunk` rows, 32 at a time on
// v_mfma_f32_16x16x32_bf16 with A(i, k) = delta(k, i) and B(k, j) = act(k, j), both staged
// transposed in LDS; wave w owns output rows [16 w, 16 w + 16). The block's partial goes to gW with
// one FP32 atomic per entry.
constexpr int kWgRow = 40; // 32 batch rows + 8 padding per staged unit

__global__ void __launch_bounds__(256) k_mlp_wgrad(const uint16_t* dl, int units, const uint16_t* act, uint32_t np,
                                                   uint32_t chunk, float* gW)
{
    __shared__ __align__(16) uint16_t DT[kP * kWgRow];
    __shared__ __align__(16) uint16_t AT[kP * kWgRow];
    const uint32_t lane = threadIdx.x & 63u, w = threadIdx.x >> 6;
    const uint32_t r = lane & 15u, q = lane >> 4;
    const bool mine = int(16u * w) < units; // same in every lane of the wave
    f32x4 acc[4];
    for (int j = 0; j < 4; j++) acc[j] = f32x4{0.f, 0.f, 0.f, 0.f};
    const uint32_t r0 = blockIdx.x * chunk, r1 = min(np, r0 + chunk);
    for (uint32_t k0 = r0; k0 < r1; k0 += 32u) {
        __syncthreads();
        for (uint32_t idx = threadIdx.x; idx < 32u * kP; idx += 256u) {
            const uint32_t kk = idx / kP, u = idx % kP;
            const bool in = k0 + kk < r1;
            DT[u * kWgRow + kk] = in && int(u) < units ? dl[size_t(k0 + kk) * units + u] : uint16_t(0);
            AT[u * kWgRow + kk] = in ? act[size_t(k0 + kk) * kP + u] : uint16_t(0);
        }
        __syncthreads();
        if (mine) {
            const bf16x8 a = *reinterpret_cast<const bf16x8*>(DT + (16u * w + r) * kWgRow + 8u * q);
            for (int j = 0; j < 4; j++)
                acc[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                    a, *reinterpret_cast<const bf16x8*>(AT + (16u * j + r) * kWgRow + 8u * q), acc[j], 0, 0, 0);
        }
    }
    if (mine)
        for (int j = 0; j < 4; j++)
            for (int e = 0; e < 4; e++) atomicAdd(&gW[(16u * w + 4u * q + uint32_t(e)) * kP + 16u * j + r], acc[j][e]);
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

// blocks of the fused forward pass: one 16-row tile per wave, at most 8 blocks per CU's worth
static dim3 forward_grid(uint32_t np) { return dim3(std::max(1u, std::min((np / 16u + 3u) / 4u, 2048u))); }

// samples on the device, stride floats each, position in [0,1]^3 then the arrival direction
int GuideNet::forward(const float* samples, int stride, uint32_t n, hipStream_t s)
{
    if (reserve(n)) return UPBP_E_HIP;
    const uint32_t np = pad16(n);
    hipLaunchKernelGGL(k_mlp_forward<true>, forward_grid(np), dim3(256), 0, s, samples, stride, n, np,
                       wq.as<uint16_t>(), bq.as<float>(), x0.as<uint16_t>(), h1.as<uint16_t>(), h2.as<uint16_t>(),
                       h3.as<uint16_t>(), out.as<float>());
    return hipGetLastError() == hipSuccess ? UPBP_OK : UPBP_E_HIP;
}

// the raw outputs only (rows of 48 floats, pad16(n) rows) into o; no other memory is written
int GuideNet::forward_into(const float* samples, int stride, uint32_t n, float* o, hipStream_t s) const
{
    const uint32_t np = pad16(n);
    hipLaunchKernelGGL(k_mlp_forward<false>, forward_grid(np), dim3(256), 0, s, samples, stride, n, np,
                       wq.as<uint16_t>(), bq.as<float>(), nullptr, nullptr, nullptr, nullptr, o);
    return hipGetLastError() == hipSuccess ? UPBP_OK : UPBP_E_HIP;
}

// training samples on the device: 10 floats each (position, arrival, scattered, weight)
int GuideNet::gradient(const float* samples, uint32_t n, hipStream_t s)
{
    int e = forward(samples, 10, n, s);
    if (e) return e;
    const uint32_t np = pad16(n);
    hipLaunchKernelGGL(k_zero, dim3(1), dim3(64), 0, s, loss.as<uint32_t>(), size_t(4));
    hipLaunchKernelGGL(k_loss_grad, dim3((np + 255) / 256), dim3(256), 0, s, out.as<float>(), samples, n, np,
                       1.f / float(n), d_o.as<uint16_t>(), loss.as<float>());
    // deltas: d_{l-1} = (d_l W_l) masked by the ReLU of layer l-1
    hipLaunchKernelGGL(k_mlp_backward, forward_grid(np), dim3(256), 0, s, wq.as<uint16_t>(), d_o.as<uint16_t>(),
                       h1.as<uint16_t>(), h2.as<uint16_t>(), h3.as<uint16_t>(), np, d1.as<uint16_t>(),
                       d2.as<uint16_t>(), d3.as<uint16_t>());
    // weight gradients: gW_l[r][c] = sum_i delta_l(i, r) act_{l-1}(i, c), in chunks of the batch
    hipLaunchKernelGGL(k_zero, dim3((kQTotal + 255) / 256), dim3(256), 0, s, gW.as<uint32_t>(), size_t(kQTotal));
    const uint32_t chunk = std::max(256u, (np / 256u + 31u) & ~31u); // about 256 blocks per layer
    float* g = gW.as<float>();
    const uint16_t* acts[4] = {x0.as<uint16_t>(), h1.as<uint16_t>(), h2.as<uint16_t>(), h3.as<uint16_t>()};
    const uint16_t* dels[4] = {d1.as<uint16_t>(), d2.as<uint16_t>(), d3.as<uint16_t>(), d_o.as<uint16_t>()};
    const int rows[4] = {64, 64, 64, kOutP}, off[4] = {kQ1, kQ2, kQ3, kQ4};
    for (int l = 0; l < 4; l++)
        hipLaunchKernelGGL(k_mlp_wgrad, dim3((np + chunk - 1) / chunk), dim3(256), 0, s, dels[l], rows[l], acts[l], np,
                           chunk, g + off[l]);
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
    if (grid_in.ensure(size_t(n) * 6 * 4) || out.ensure(size_t(pad16(n)) * kOutP * 4)) return UPBP_E_HIP;
    hipLaunchKernelGGL(k_grid_inputs, dim3((n + 255) / 256), dim3(256), 0, s, res, grid_in.as<float>());
    const int e = forward_into(grid_in.as<float>(), 6, n, out.as<float>(), s);
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
// harnesses that forbid hipMalloc): the forward pass of upbp_mlp_forward with its input staging
// and outputs in one file-scope device array (the fused kernel keeps the activations in LDS). Each GPU holds its own copy of the
// array until the process ends; its address on the current device is resolved once per device
// with hipGetSymbolAddress and only that address is used from host code. Every call on one GPU
// shares the array, so calls that overlap on one GPU race. Nothing in it needs clearing: every
// byte a call reads was written earlier by the same call.
namespace {

constexpr uint32_t kNoAllocMax = 32768; // samples per call (a multiple of 16)
constexpr size_t kScrIn = 0;                                                  // n x 6 floats
constexpr size_t kScrOut = kScrIn + size_t(kNoAllocMax) * 6 * 4;              // pad16(n) x 48 floats
constexpr size_t kScrBytes = kScrOut + size_t(kNoAllocMax) * upbp::kOutP * 4; // 6.75 MiB
static_assert(kScrOut % 256 == 0, "scratch regions stay aligned");

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
    float* o = reinterpret_cast<float*>(base + kScrOut);
    if (hipMemcpy(in, samples, size_t(n) * 24, hipMemcpyHostToDevice) != hipSuccess) return UPBP_E_HIP;
    const int e = m->net.forward_into(in, 6, n, o, nullptr);
    if (e) return e;
    // rows of 48 on the device, the first 40 of each to the host
    if (hipMemcpy2D(out, 40 * 4, o, upbp::kOutP * 4, 40 * 4, n, hipMemcpyDeviceToHost) != hipSuccess) return UPBP_E_HIP;
    return UPBP_OK;
}
