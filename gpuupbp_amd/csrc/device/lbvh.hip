// LBVH construction on gfx950: centroid bounds -> 30-bit Morton codes -> stable
// LSD radix sort (8-bit digits, LDS histograms, wave64 ballot ranking) -> Karras
// (HPG 2012) hierarchy -> bottom-up box refit with one atomic flag per node.
// Replaces the per-iteration OptiX GAS/IAS builds of nolmoonen/gpuupbp
// gpuupbp/renderer/renderer.cpp:201-307 and the scene GAS build of
// gpuupbp/renderer/renderer_init.cpp:31-289.
#include "lbvh.h"

#include <algorithm>

using namespace upbp;

namespace {

constexpr int kSortThreads = 256;
constexpr int kSortItems = 8;
constexpr int kSortTile = kSortThreads * kSortItems;
constexpr int kScanThreads = 256;
constexpr int kScanItems = 8;
constexpr int kScanTile = kScanThreads * kScanItems;

__device__ __forceinline__ uint32_t f2ord(float f)
{
    const uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float ord2f(uint32_t u)
{
    return __uint_as_float((u & 0x80000000u) ? (u & 0x7fffffffu) : ~u);
}

__global__ void k_fill_u32(uint32_t* p, uint32_t v, uint32_t n)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// centroid bounds: bounds[0..2] = ordered min, bounds[3..5] = ordered max
// n_dev (when given): the count in device memory, at most n
__device__ __forceinline__ uint32_t dev_count(uint32_t n, const uint32_t* n_dev) { return n_dev ? min(*n_dev, n) : n; }

__global__ void k_centroid_bounds(const float4* lo, const float4* hi, uint32_t cap, const uint32_t* n_dev,
                                  uint32_t* bounds)
{
    __shared__ float smin[3][256], smax[3][256];
    const uint32_t n = dev_count(cap, n_dev);
    const uint32_t t = threadIdx.x;
    float mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (uint32_t i = blockIdx.x * blockDim.x + t; i < n; i += gridDim.x * blockDim.x) {
        const float4 a = lo[i], b = hi[i];
        const float c[3] = {(a.x + b.x) * .5f, (a.y + b.y) * .5f, (a.z + b.z) * .5f};
        for (int k = 0; k < 3; k++) {
            mn[k] = fminf(mn[k], c[k]);
            mx[k] = fmaxf(mx[k], c[k]);
        }
    }
    for (int k = 0; k < 3; k++) {
        smin[k][t] = mn[k];
        smax[k][t] = mx[k];
    }
    __syncthreads();
    for (uint32_t s = blockDim.x / 2; s > 0; s >>= 1) {
        if (t < s)
            for (int k = 0; k < 3; k++) {
                smin[k][t] = fminf(smin[k][t], smin[k][t + s]);
                smax[k][t] = fmaxf(smax[k][t], smax[k][t + s]);
            }
        __syncthreads();
    }
    if (t == 0)
        for (int k = 0; k < 3; k++) {
            atomicMin(&bounds[k], f2ord(smin[k][0]));
            atomicMax(&bounds[3 + k], f2ord(smax[k][0]));
        }
}

__device__ __forceinline__ uint32_t spread10(uint32_t v)
{
    v = (v * 0x00010001u) & 0xFF0000FFu;
    v = (v * 0x00000101u) & 0x0F00F00Fu;
    v = (v * 0x00000011u) & 0xC30C30C3u;
    v = (v * 0x00000005u) & 0x49249249u;
    return v;
}

__global__ void k_morton(const float4* lo, const float4* hi, uint32_t cap, const uint32_t* n_dev, const uint32_t* bounds,
                         uint32_t* keys, uint32_t* vals)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= dev_count(cap, n_dev)) return;
    const float4 a = lo[i], b = hi[i];
    const float c[3] = {(a.x + b.x) * .5f, (a.y + b.y) * .5f, (a.z + b.z) * .5f};
    uint32_t q[3];
    for (int k = 0; k < 3; k++) {
        const float l = ord2f(bounds[k]), h = ord2f(bounds[3 + k]);
        const float ext = h - l;
        const float u = ext > 0.f ? (c[k] - l) / ext : 0.5f;
        q[k] = uint32_t(fminf(fmaxf(u * 1024.f, 0.f), 1023.f));
    }
    keys[i] = (spread10(q[0]) << 2) | (spread10(q[1]) << 1) | spread10(q[2]);
    vals[i] = i;
}

// ---- radix sort ----------------------------------------------------------------
__global__ void __launch_bounds__(kSortThreads) k_radix_hist(const uint32_t* keys, uint32_t cap,
                                                             const uint32_t* n_dev, int shift, uint32_t* hist,
                                                             uint32_t nblocks)
{
    __shared__ uint32_t h[256];
    const uint32_t n = dev_count(cap, n_dev);
    h[threadIdx.x] = 0;
    __syncthreads();
    const uint32_t base = blockIdx.x * kSortTile;
    for (int r = 0; r < kSortItems; r++) {
        const uint32_t i = base + r * kSortThreads + threadIdx.x;
        if (i < n) atomicAdd(&h[(keys[i] >> shift) & 0xffu], 1u);
    }
    __syncthreads();
    hist[threadIdx.x * nblocks + blockIdx.x] = h[threadIdx.x];
}

// Stable scatter: element order within a tile is preserved (round r, then thread id).
__global__ void __launch_bounds__(kSortThreads) k_radix_scatter(const uint32_t* keys_in, const uint32_t* vals_in,
                                                                uint32_t* keys_out, uint32_t* vals_out, uint32_t cap,
                                                                const uint32_t* n_dev, int shift,
                                                                const uint32_t* offsets, uint32_t nblocks)
{
    constexpr int kWaves = kSortThreads / 64;
    const uint32_t n = dev_count(cap, n_dev);
    __shared__ uint32_t running[256];
    __shared__ uint32_t wave_hist[kWaves][256];
    const uint32_t t = threadIdx.x;
    const uint32_t lane = t & 63u, wave = t >> 6;
    running[t] = offsets[t * nblocks + blockIdx.x];
    for (int w = 0; w < kWaves; w++) wave_hist[w][t] = 0;
    __syncthreads();
    const uint32_t base = blockIdx.x * kSortTile;
    const uint64_t lt_mask = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    for (int r = 0; r < kSortItems; r++) {
        const uint32_t i = base + r * kSortThreads + t;
        const bool valid = i < n;
        const uint32_t key = valid ? keys_in[i] : 0u;
        const uint32_t val = valid ? vals_in[i] : 0u;
        const uint32_t d = (key >> shift) & 0xffu;
        uint64_t match = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; b++) {
            const bool bit = (d >> b) & 1u;
            const uint64_t bb = __ballot(bit);
            match &= bit ? bb : ~bb;
        }
        const uint32_t rank = __popcll(match & lt_mask);
        const bool leader = valid && (match & lt_mask) == 0ull;
        if (leader) wave_hist[wave][d] = __popcll(match);
        __syncthreads();
        if (valid) {
            uint32_t off = running[d] + rank;
            for (uint32_t w = 0; w < wave; w++) off += wave_hist[w][d];
            keys_out[off] = key;
            vals_out[off] = val;
        }
        __syncthreads();
        uint32_t tot = 0;
        for (int w = 0; w < kWaves; w++) {
            tot += wave_hist[w][t];
            wave_hist[w][t] = 0;
        }
        running[t] += tot;
        __syncthreads();
    }
}

// ---- exclusive scan (reduce, scan of block sums, downsweep) ----------------------
__global__ void __launch_bounds__(kScanThreads) k_scan_reduce(const uint32_t* in, uint32_t n, uint32_t* sums)
{
    __shared__ uint32_t s[kScanThreads];
    uint32_t acc = 0;
    const uint32_t base = blockIdx.x * kScanTile + threadIdx.x * kScanItems;
    for (int k = 0; k < kScanItems; k++)
        if (base + k < n) acc += in[base + k];
    s[threadIdx.x] = acc;
    __syncthreads();
    for (uint32_t st = kScanThreads / 2; st > 0; st >>= 1) {
        if (threadIdx.x < st) s[threadIdx.x] += s[threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x == 0) sums[blockIdx.x] = s[0];
}

// Scans one tile per block; adds block_offsets[blockIdx] if given. Writes the grand total of
// block 0's tile into *total when there is a single block.
__global__ void __launch_bounds__(kScanThreads) k_scan_tile(const uint32_t* in, uint32_t* out, uint32_t n,
                                                            const uint32_t* block_offsets, uint32_t* total)
{
    __shared__ uint32_t s[kScanThreads];
    const uint32_t base = blockIdx.x * kScanTile + threadIdx.x * kScanItems;
    uint32_t v[kScanItems];
    uint32_t acc = 0;
    for (int k = 0; k < kScanItems; k++) {
        v[k] = base + k < n ? in[base + k] : 0u;
        acc += v[k];
    }
    s[threadIdx.x] = acc;
    __syncthreads();
    // Hillis-Steele inclusive scan over thread sums
    for (uint32_t st = 1; st < kScanThreads; st <<= 1) {
        const uint32_t x = threadIdx.x >= st ? s[threadIdx.x - st] : 0u;
        __syncthreads();
        s[threadIdx.x] += x;
        __syncthreads();
    }
    uint32_t run = (threadIdx.x ? s[threadIdx.x - 1] : 0u) + (block_offsets ? block_offsets[blockIdx.x] : 0u);
    for (int k = 0; k < kScanItems; k++) {
        if (base + k < n) out[base + k] = run;
        run += v[k];
    }
    if (total && gridDim.x == 1 && threadIdx.x == kScanThreads - 1) *total = s[kScanThreads - 1];
}

// ---- Karras hierarchy -----------------------------------------------------------
__device__ __forceinline__ int delta(const uint32_t* keys, int n, int i, int j)
{
    if (j < 0 || j >= n) return -1;
    const uint32_t a = keys[i], b = keys[j];
    if (a == b) return 32 + __clz(uint32_t(i) ^ uint32_t(j));
    return __clz(a ^ b);
}

__global__ void k_karras(const uint32_t* keys, uint32_t cap, const uint32_t* n_dev, BNode* nodes)
{
    const int n = int(dev_count(cap, n_dev));
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n - 1) return;
    const int d = (delta(keys, n, i, i + 1) - delta(keys, n, i, i - 1)) >= 0 ? 1 : -1;
    const int dmin = delta(keys, n, i, i - d);
    int lmax = 2;
    while (delta(keys, n, i, i + lmax * d) > dmin) lmax <<= 1;
    int l = 0;
    for (int t = lmax >> 1; t >= 1; t >>= 1)
        if (delta(keys, n, i, i + (l + t) * d) > dmin) l += t;
    const int j = i + l * d;
    const int dnode = delta(keys, n, i, j);
    int s = 0;
    int div = 2;
    int t = (l + div - 1) / div;
    for (;;) {
        if (delta(keys, n, i, i + (s + t) * d) > dnode) s += t;
        if (t == 1) break;
        div <<= 1;
        t = (l + div - 1) / div;
    }
    const int g = i + s * d + (d < 0 ? -1 : 0);
    const int lo = i < j ? i : j, hi = i < j ? j : i;
    nodes[i].c0 = (lo == g) ? ~g : g;
    nodes[i].c1 = (hi == g + 1) ? ~(g + 1) : g + 1;
}

__device__ __forceinline__ void write_child_box(BNode* nd, int slot, float4 a, float4 b)
{
    if (slot == 0) {
        nd->lx0 = a.x;
        nd->ly0 = a.y;
        nd->lz0 = a.z;
        nd->ux0 = b.x;
        nd->uy0 = b.y;
        nd->uz0 = b.z;
    } else {
        nd->lx1 = a.x;
        nd->ly1 = a.y;
        nd->lz1 = a.z;
        nd->ux1 = b.x;
        nd->uy1 = b.y;
        nd->uz1 = b.z;
    }
}

// Level-synchronous refit. A breadth-first pass lists the inner nodes level by level
// (meta[l] = first entry of level l in lvl[], meta[kMaxLevels + 1 + l] = its size), then
// one launch per level, deepest first, writes each node's two child boxes from the leaf
// boxes or from the already refitted child nodes. Kernel boundaries order the levels, so
// no device-wide fences or atomics on the boxes are needed. A Karras tree over 30-bit
// codes with index tie-breaks is at most 30 + 32 levels deep.
constexpr int kMaxLevels = 64;

// the root level holds the root when the tree has an inner node to refit (n >= 2); a tree of
// one primitive has its root written whole by k_single
__global__ void k_level_init(uint32_t* meta, uint32_t* lvl, uint32_t cap, const uint32_t* n_dev)
{
    const int t = threadIdx.x;
    const uint32_t n = dev_count(cap, n_dev);
    if (t <= kMaxLevels) {
        meta[t] = 0;
        meta[kMaxLevels + 1 + t] = t == 0 && n >= 2 ? 1u : 0u;
    }
    if (t == 0) lvl[0] = 0;
}

// The waves of a block allocate together: one atomic per block on the level's count, the one
// address every block of the launch hits (per wave, the big levels queued on it). Called by
// every thread of the block (at most 7 waves); tot: this wave's count; returns its first slot.
__device__ __forceinline__ uint32_t block_alloc(uint32_t tot, uint32_t* counter, uint32_t* sh)
{
    const uint32_t wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
    if (__lane_id() == 0) sh[wave] = tot;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t sum = 0;
        for (uint32_t w = 0; w < nw; w++) {
            const uint32_t t = sh[w];
            sh[w] = sum;
            sum += t;
        }
        sh[7] = sum ? atomicAdd(counter, sum) : 0u;
    }
    __syncthreads();
    const uint32_t at = sh[7] + sh[wave];
    __syncthreads(); // sh is reused by the next call
    return at;
}

__global__ void __launch_bounds__(256) k_bfs_level(const BNode* nodes, uint32_t* lvl, uint32_t* meta, int level)
{
    __shared__ uint32_t sh[8];
    uint32_t* count = meta + kMaxLevels + 1;
    const uint32_t s = meta[level], n = count[level];
    const uint32_t next = s + n;
    if (blockIdx.x == 0 && threadIdx.x == 0) meta[level + 1] = next;
    const uint32_t lane = __lane_id();
    const uint64_t lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t base = blockIdx.x * blockDim.x; base < n; base += stride) {
        const uint32_t i = base + threadIdx.x;
        int c0 = -1, c1 = -1;
        if (i < n) {
            const BNode& nd = nodes[lvl[s + i]];
            c0 = nd.c0;
            c1 = nd.c1;
        }
        const bool k0 = c0 >= 0, k1 = c1 >= 0;
        const uint64_t m0 = __ballot(k0), m1 = __ballot(k1);
        const uint32_t tot = uint32_t(__popcll(m0) + __popcll(m1));
        const uint32_t at = block_alloc(tot, &count[level + 1], sh) + next;
        const uint32_t p0 = uint32_t(__popcll(m0 & lt)), p1 = uint32_t(__popcll(m0)) + uint32_t(__popcll(m1 & lt));
        if (k0) lvl[at + p0] = uint32_t(c0);
        if (k1) lvl[at + p1] = uint32_t(c1);
    }
}

__device__ __forceinline__ void child_box(const float4* lo, const float4* hi, const uint32_t* order,
                                          const BNode* nodes, int c, float4& a, float4& b)
{
    if (c < 0) {
        const uint32_t p = order[~c];
        a = lo[p];
        b = hi[p];
    } else {
        const BNode& cn = nodes[c];
        a = make_float4(fminf(cn.lx0, cn.lx1), fminf(cn.ly0, cn.ly1), fminf(cn.lz0, cn.lz1), 0.f);
        b = make_float4(fmaxf(cn.ux0, cn.ux1), fmaxf(cn.uy0, cn.uy1), fmaxf(cn.uz0, cn.uz1), 0.f);
    }
}

__global__ void __launch_bounds__(256) k_refit_level(const float4* lo, const float4* hi, const uint32_t* order,
                                                     BNode* nodes, const uint32_t* lvl, const uint32_t* meta,
                                                     int level)
{
    const uint32_t s = meta[level], n = meta[kMaxLevels + 1 + level];
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        BNode* nd = &nodes[lvl[s + i]];
        float4 a0, b0, a1, b1;
        child_box(lo, hi, order, nodes, nd->c0, a0, b0);
        child_box(lo, hi, order, nodes, nd->c1, a1, b1);
        float4* f = reinterpret_cast<float4*>(nd);
        f[0] = make_float4(a0.x, a0.y, a0.z, b0.x);
        f[1] = make_float4(b0.y, b0.z, a1.x, a1.y);
        f[2] = make_float4(a1.z, b1.x, b1.y, b1.z);
    }
}

// single primitive: one inner node with leaf 0 on the left and an empty right box
// ---- wide collapse ----------------------------------------------------------------
// Each wide node takes one binary node and opens its largest-area inner child until it
// has kWide children (or only leaves). Wide nodes are numbered level by level: level l
// holds [meta[l], meta[l] + meta[kMaxLevels + 1 + l]) and bin[] maps them to binary nodes.

__device__ __forceinline__ float box_area(float4 a, float4 b)
{
    const float x = b.x - a.x, y = b.y - a.y, z = b.z - a.z;
    return x * y + y * z + z * x;
}

// power-of-two grid step covering [lo, hi] in 255 steps, as a biased exponent
__device__ __forceinline__ uint32_t grid_exponent(float lo, float hi)
{
    const float t = (hi - lo) * (1.f / 255.f);
    int e = (__float_as_int(t) >> 23) & 255;
    if (t > __int_as_float(e << 23)) e++;
    e = e < 1 ? 1 : e;
    for (; e < 254; e++) // the decoded top of the grid must reach hi
        if (wnode_decode(lo, 255u, 0, __int_as_float(e << 23)) >= hi) break;
    return uint32_t(e);
}

__device__ __forceinline__ uint32_t quant_lo(float o, float step, float v)
{
    float f = floorf((v - o) / step);
    uint32_t q = f <= 0.f ? 0u : (f >= 255.f ? 255u : uint32_t(f));
    while (q > 0 && wnode_decode(o, q, 0, step) > v) q--;
    return q;
}

__device__ __forceinline__ uint32_t quant_hi(float o, float step, float v)
{
    float f = ceilf((v - o) / step);
    uint32_t q = f <= 0.f ? 0u : (f >= 255.f ? 255u : uint32_t(f));
    while (q < 255 && wnode_decode(o, q, 0, step) < v) q++;
    return q;
}

__global__ void k_wide_init(uint32_t* meta, uint32_t* bin, uint32_t cap, const uint32_t* n_dev)
{
    const int t = threadIdx.x;
    const uint32_t n = dev_count(cap, n_dev);
    if (t <= kMaxLevels) {
        meta[t] = 0;
        meta[kMaxLevels + 1 + t] = t == 0 && n >= 2 ? 1u : 0u;
    }
    if (t == 0) bin[0] = 0;
}

__global__ void __launch_bounds__(256) k_wide_level(const BNode* bn, WNode* wn, uint32_t* bin, uint32_t* meta,
                                                    int level)
{
    __shared__ uint32_t sh[8];
    uint32_t* count = meta + kMaxLevels + 1;
    const uint32_t s = meta[level], n = count[level];
    const uint32_t next = s + n;
    if (blockIdx.x == 0 && threadIdx.x == 0) meta[level + 1] = next;
    const uint32_t lane = __lane_id();
    const uint64_t lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t base = blockIdx.x * blockDim.x; base < n; base += stride) {
        const uint32_t j = base + threadIdx.x;
        int c[kWide];
        float4 lo[kWide], hi[kWide];
        for (int k = 0; k < kWide; k++) c[k] = 0;
        int nc = 0, ninner = 0;
        if (j < n) {
            const BNode& b = bn[bin[s + j]];
            c[0] = b.c0;
            c[1] = b.c1;
            lo[0] = make_float4(b.lx0, b.ly0, b.lz0, 0.f);
            hi[0] = make_float4(b.ux0, b.uy0, b.uz0, 0.f);
            lo[1] = make_float4(b.lx1, b.ly1, b.lz1, 0.f);
            hi[1] = make_float4(b.ux1, b.uy1, b.uz1, 0.f);
            nc = 2;
            while (nc < kWide) {
                int best = -1;
                float ba = -1.f;
                for (int k = 0; k < nc; k++)
                    if (c[k] >= 0 && box_area(lo[k], hi[k]) > ba) {
                        ba = box_area(lo[k], hi[k]);
                        best = k;
                    }
                if (best < 0) break;
                const BNode& x = bn[c[best]];
                c[best] = x.c0;
                lo[best] = make_float4(x.lx0, x.ly0, x.lz0, 0.f);
                hi[best] = make_float4(x.ux0, x.uy0, x.uz0, 0.f);
                c[nc] = x.c1;
                lo[nc] = make_float4(x.lx1, x.ly1, x.lz1, 0.f);
                hi[nc] = make_float4(x.ux1, x.uy1, x.uz1, 0.f);
                nc++;
            }
            for (int k = 0; k < nc; k++) ninner += c[k] >= 0 ? 1 : 0;
        }
        // block-aggregated allocation of the inner children at the next level
        uint32_t pre = 0, tot = 0;
        for (int bit = 0; bit < 4; bit++) { // ninner <= kWide = 8: four bits
            const uint64_t m = __ballot((uint32_t(ninner) >> bit) & 1u);
            pre += uint32_t(__popcll(m & lt)) << bit;
            tot += uint32_t(__popcll(m)) << bit;
        }
        uint32_t at = block_alloc(tot, &count[level + 1], sh) + next + pre;
        if (j < n) {
            float4 nlo = lo[0], nhi = hi[0];
            for (int k = 1; k < nc; k++) {
                nlo = make_float4(fminf(nlo.x, lo[k].x), fminf(nlo.y, lo[k].y), fminf(nlo.z, lo[k].z), 0.f);
                nhi = make_float4(fmaxf(nhi.x, hi[k].x), fmaxf(nhi.y, hi[k].y), fmaxf(nhi.z, hi[k].z), 0.f);
            }
            const uint32_t ex = grid_exponent(nlo.x, nhi.x), ey = grid_exponent(nlo.y, nhi.y),
                           ez = grid_exponent(nlo.z, nhi.z);
            const float sx = __int_as_float(int(ex << 23)), sy = __int_as_float(int(ey << 23)),
                        sz = __int_as_float(int(ez << 23));
            WNode w;
            w.ox = nlo.x;
            w.oy = nlo.y;
            w.oz = nlo.z;
            w.meta = ex | (ey << 8) | (ez << 16) | (uint32_t(nc) << 24);
            for (int k = 0; k < 12; k++) w.q[k] = 0;
            for (int k = 0; k < 8; k++) w.pad[k] = 0;
            for (int k = 0; k < kWide; k++) w.child[k] = 0;
            for (int k = 0; k < nc; k++) {
                if (c[k] >= 0) {
                    bin[at] = uint32_t(c[k]);
                    w.child[k] = int(at++);
                } else {
                    w.child[k] = c[k];
                }
                const int sh = 8 * (k & 3), wd = k >> 2; // plane p of child k: byte k & 3 of q[2 p + (k >> 2)]
                w.q[0 + wd] |= quant_lo(w.ox, sx, lo[k].x) << sh;
                w.q[2 + wd] |= quant_lo(w.oy, sy, lo[k].y) << sh;
                w.q[4 + wd] |= quant_lo(w.oz, sz, lo[k].z) << sh;
                w.q[6 + wd] |= quant_hi(w.ox, sx, hi[k].x) << sh;
                w.q[8 + wd] |= quant_hi(w.oy, sy, hi[k].y) << sh;
                w.q[10 + wd] |= quant_hi(w.oz, sz, hi[k].z) << sh;
            }
            wn[s + j] = w;
        }
    }
}

// one primitive: a wide root with a single leaf child; none: a wide root without children (with
// n_dev, the kernel runs every build and acts only for those two counts)
__global__ void k_wide_single(const BNode* bn, WNode* wn, const uint32_t* n_dev)
{
    if (n_dev && *n_dev >= 2) return;
    WNode w;
    for (int k = 0; k < kWide; k++) w.child[k] = 0;
    for (int k = 0; k < 12; k++) w.q[k] = 0;
    for (int k = 0; k < 8; k++) w.pad[k] = 0;
    if (n_dev && *n_dev == 0) {
        w.ox = w.oy = w.oz = 0.f;
        w.meta = (127u) | (127u << 8) | (127u << 16);
        wn[0] = w;
        return;
    }
    const BNode& b = bn[0];
    w.ox = b.lx0;
    w.oy = b.ly0;
    w.oz = b.lz0;
    const uint32_t ex = grid_exponent(b.lx0, b.ux0), ey = grid_exponent(b.ly0, b.uy0), ez = grid_exponent(b.lz0, b.uz0);
    w.meta = ex | (ey << 8) | (ez << 16) | (1u << 24);
    w.child[0] = ~0;
    // child 0: low planes at the origin (byte 0 of q[0], q[2], q[4] stays 0), high planes in q[6..10]
    w.q[6] = quant_hi(b.lx0, __int_as_float(int(ex << 23)), b.ux0);
    w.q[8] = quant_hi(b.ly0, __int_as_float(int(ey << 23)), b.uy0);
    w.q[10] = quant_hi(b.lz0, __int_as_float(int(ez << 23)), b.uz0);
    wn[0] = w;
}

__global__ void k_single(const float4* lo, const float4* hi, BNode* nodes, uint32_t* order, const uint32_t* n_dev)
{
    if (n_dev && *n_dev != 1) return; // with n_dev the kernel runs every build and acts for one primitive
    BNode nd;
    nd.lx0 = lo[0].x;
    nd.ly0 = lo[0].y;
    nd.lz0 = lo[0].z;
    nd.ux0 = hi[0].x;
    nd.uy0 = hi[0].y;
    nd.uz0 = hi[0].z;
    // the unused right slot gets a point box far outside any scene: an inverted box would pass
    // the slab test, which is symmetric in the two planes of each axis
    nd.lx1 = nd.ly1 = nd.lz1 = 3e38f;
    nd.ux1 = nd.uy1 = nd.uz1 = 3e38f;
    nd.c0 = ~0;
    nd.c1 = ~0;
    nd.pad0 = nd.pad1 = 0;
    nodes[0] = nd;
    order[0] = 0;
}

inline uint32_t blocks_for(uint64_t n, uint32_t tpb) { return uint32_t((n + tpb - 1) / tpb); }

} // namespace

namespace upbp_host {

// scan_exclusive uses sc.scan as a stack of block-sum arrays, starting at word `base`
static int scan_rec(const uint32_t* in, uint32_t* out, uint32_t n, uint32_t* total_dev, BvhScratch& sc, size_t base,
                    hipStream_t s)
{
    if (n == 0) {
        if (total_dev) hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(1), 0, s, total_dev, 0u, 1u);
        return int(hipGetLastError());
    }
    const uint32_t nb = blocks_for(n, kScanTile);
    if (nb == 1) {
        hipLaunchKernelGGL(k_scan_tile, dim3(1), dim3(kScanThreads), 0, s, in, out, n, nullptr, total_dev);
        return int(hipGetLastError());
    }
    uint32_t* sums = sc.scan.as<uint32_t>() + base;
    hipLaunchKernelGGL(k_scan_reduce, dim3(nb), dim3(kScanThreads), 0, s, in, n, sums);
    int e = scan_rec(sums, sums, nb, total_dev, sc, base + nb + 1, s);
    if (e) return e;
    hipLaunchKernelGGL(k_scan_tile, dim3(nb), dim3(kScanThreads), 0, s, in, out, n, sums, nullptr);
    return int(hipGetLastError());
}

int scan_exclusive(const uint32_t* in, uint32_t* out, uint32_t n, uint32_t* total_dev, BvhScratch& sc, hipStream_t s)
{
    size_t need = 16;
    for (uint64_t m = n; m > uint64_t(kScanTile); m = blocks_for(m, kScanTile)) need += blocks_for(m, kScanTile) + 1;
    if (sc.scan.ensure(need * 4)) return 1;
    return scan_rec(in, out, n, total_dev, sc, 0, s);
}

int radix_sort_pairs(uint32_t* keys, uint32_t* vals, uint32_t n, int bits, BvhScratch& sc, hipStream_t s,
                     const uint32_t* n_dev)
{
    if (n <= 1) return 0;
    const uint32_t nb = blocks_for(n, kSortTile);
    if (sc.keys2.ensure(size_t(n) * 4) || sc.vals2.ensure(size_t(n) * 4) || sc.hist.ensure(size_t(256) * nb * 4))
        return 1;
    uint32_t* hist = sc.hist.as<uint32_t>();
    uint32_t *ka = keys, *va = vals, *kb = sc.keys2.as<uint32_t>(), *vb = sc.vals2.as<uint32_t>();
    int passes = 0;
    for (int shift = 0; shift < bits; shift += 8) {
        hipLaunchKernelGGL(k_radix_hist, dim3(nb), dim3(kSortThreads), 0, s, ka, n, n_dev, shift, hist, nb);
        const int e = scan_exclusive(hist, hist, 256 * nb, nullptr, sc, s);
        if (e) return e;
        hipLaunchKernelGGL(k_radix_scatter, dim3(nb), dim3(kSortThreads), 0, s, ka, va, kb, vb, n, n_dev, shift, hist,
                           nb);
        std::swap(ka, kb);
        std::swap(va, vb);
        passes++;
    }
    if (passes & 1) {
        if (hipMemcpyAsync(keys, ka, size_t(n) * 4, hipMemcpyDeviceToDevice, s) != hipSuccess) return 1;
        if (hipMemcpyAsync(vals, va, size_t(n) * 4, hipMemcpyDeviceToDevice, s) != hipSuccess) return 1;
    }
    return int(hipGetLastError());
}

int lbvh_build(BvhStore& st, BvhScratch& sc, const float4* lo, const float4* hi, uint32_t n, hipStream_t s,
               const uint32_t* n_dev)
{
    st.n = n;
    if (n == 0) return 0;
    const uint32_t inner = n > 1 ? n - 1 : 1;
    if (st.cap < n) {
        lbvh_free(st);
        const uint32_t c = n + n / 4;
        if (hipMalloc(&st.nodes, size_t(std::max<uint32_t>(c - 1, inner)) * sizeof(BNode)) != hipSuccess) return 1;
        if (hipMalloc(&st.order, size_t(c) * sizeof(uint32_t)) != hipSuccess) return 1;
        st.cap = c;
    }
    if (n == 1) {
        hipLaunchKernelGGL(k_single, dim3(1), dim3(1), 0, s, lo, hi, st.nodes, st.order, n_dev);
        return int(hipGetLastError());
    }
    if (sc.keys.ensure(size_t(n) * 4) || sc.lvl.ensure(size_t(n) * 4) || sc.meta.ensure(2 * (kMaxLevels + 1) * 4) ||
        sc.bounds.ensure(64))
        return 1;
    uint32_t* bounds = sc.bounds.as<uint32_t>();
    uint32_t* keys = sc.keys.as<uint32_t>();
    const uint32_t tpb = 256;
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(3), 0, s, bounds, 0xffffffffu, 3u);
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(3), 0, s, bounds + 3, 0u, 3u);
    hipLaunchKernelGGL(k_centroid_bounds, dim3(std::min<uint32_t>(blocks_for(n, tpb), 1024)), dim3(tpb), 0, s, lo, hi,
                       n, n_dev, bounds);
    hipLaunchKernelGGL(k_morton, dim3(blocks_for(n, tpb)), dim3(tpb), 0, s, lo, hi, n, n_dev, bounds, keys, st.order);
    const int e = radix_sort_pairs(keys, st.order, n, 30, sc, s, n_dev);
    if (e) return e;
    // with a count in device memory the tree may hold a single primitive (written whole here) or none
    if (n_dev) hipLaunchKernelGGL(k_single, dim3(1), dim3(1), 0, s, lo, hi, st.nodes, st.order, n_dev);
    hipLaunchKernelGGL(k_karras, dim3(blocks_for(n - 1, tpb)), dim3(tpb), 0, s, keys, n, n_dev, st.nodes);
    uint32_t* lvl = sc.lvl.as<uint32_t>();
    uint32_t* meta = sc.meta.as<uint32_t>();
    // levels deeper than a tree over n leaves can reach are skipped on the host
    int levels = 1;
    while (levels < kMaxLevels && (uint64_t(1) << std::min(levels - 1, 62)) < n) levels++;
    levels = std::min(kMaxLevels, levels + 31);
    const uint32_t grid = std::min<uint32_t>(blocks_for(n, tpb), 2048);
    hipLaunchKernelGGL(k_level_init, dim3(1), dim3(kMaxLevels + 1), 0, s, meta, lvl, n, n_dev);
    for (int l = 0; l + 1 < levels; l++)
        hipLaunchKernelGGL(k_bfs_level, dim3(grid), dim3(tpb), 0, s, st.nodes, lvl, meta, l);
    for (int l = levels - 1; l >= 0; l--)
        hipLaunchKernelGGL(k_refit_level, dim3(grid), dim3(tpb), 0, s, lo, hi, st.order, st.nodes, lvl, meta, l);
    return int(hipGetLastError());
}

int wbvh_build(WBvhStore& w, const BvhStore& bin, BvhScratch& sc, uint32_t n, hipStream_t s, const uint32_t* n_dev)
{
    w.n = n;
    if (n == 0) return 0;
    const uint32_t need = n > 1 ? n - 1 : 1;
    if (w.cap < need) {
        wbvh_free(w);
        const uint32_t c = need + need / 4;
        if (hipMalloc(&w.nodes, size_t(c) * sizeof(WNode)) != hipSuccess) return 1;
        if (hipMalloc(&w.bin, size_t(c) * sizeof(uint32_t)) != hipSuccess) return 1;
        w.cap = c;
    }
    if (n == 1 || n_dev) hipLaunchKernelGGL(k_wide_single, dim3(1), dim3(1), 0, s, bin.nodes, w.nodes, n_dev);
    if (n == 1) return int(hipGetLastError());
    if (sc.meta.ensure(2 * (kMaxLevels + 1) * 4)) return 1;
    uint32_t* meta = sc.meta.as<uint32_t>();
    int levels = 1;
    while (levels < kMaxLevels && (uint64_t(1) << std::min(levels - 1, 62)) < n) levels++;
    levels = std::min(kMaxLevels, levels + 31);
    const uint32_t grid = std::min<uint32_t>(blocks_for(n, 256), 2048);
    hipLaunchKernelGGL(k_wide_init, dim3(1), dim3(kMaxLevels + 1), 0, s, meta, w.bin, n, n_dev);
    for (int l = 0; l + 1 < levels; l++)
        hipLaunchKernelGGL(k_wide_level, dim3(grid), dim3(256), 0, s, bin.nodes, w.nodes, w.bin, meta, l);
    return int(hipGetLastError());
}

void wbvh_free(WBvhStore& w)
{
    if (w.nodes) (void)hipFree(w.nodes);
    if (w.bin) (void)hipFree(w.bin);
    w.nodes = nullptr;
    w.bin = nullptr;
    w.cap = 0;
}

void lbvh_free(BvhStore& st)
{
    if (st.nodes) (void)hipFree(st.nodes);
    if (st.order) (void)hipFree(st.order);
    st.nodes = nullptr;
    st.order = nullptr;
    st.cap = 0;
}

void lbvh_scratch_free(BvhScratch& sc)
{
    for (DevBuf* b : {&sc.keys, &sc.keys2, &sc.vals2, &sc.hist, &sc.scan, &sc.lvl, &sc.meta, &sc.bounds})
        b->release();
}

} // namespace upbp_host
