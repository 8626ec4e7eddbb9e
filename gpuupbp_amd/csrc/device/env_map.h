// Environment map of the background light (latitude-longitude RGB image) and its luminance
// distribution: nolmoonen/gpuupbp gpuupbp/kernel/env_map.cuh:36-137 (direction mapping,
// bilinear lookup, sampling and pdf), gpuupbp/kernel/distribution.cuh:48-126 (piecewise-constant
// 1D / 2D distributions) and gpuupbp/host/env_map.hpp:42-144 (their construction).
//
// Layout in HBM: the image as W*H RGB floats; the distribution over nu = 2H columns and nv = H
// rows as separate arrays of conditional functions [nv][nu], conditional CDFs [nv][nu + 1],
// conditional integrals [nv], the marginal function [nv], marginal CDF [nv + 1] and its integral
// (struct DEnvMap).
#pragma once

#include "upbp_device.h"

namespace upbp {

UPBP_HD int clamp_int(int x, int lo, int hi) { return x < lo ? lo : (x > hi ? hi : x); }

// index of the last CDF entry <= u, clamped to [0, count - 1] (distribution.cuh:48-63)
UPBP_HD int cdf_interval(const float* cdf, int count, float u)
{
    int first = 0, len = count + 1;
    while (len > 0) {
        const int half = len >> 1, mid = first + half;
        if (cdf[mid] <= u) {
            first = mid + 1;
            len -= half + 1;
        } else {
            len = half;
        }
    }
    return clamp_int(first - 1, 0, count - 1);
}

// sample_continuous of a 1D distribution (distribution.cuh:66-86): position in [0, 1), pdf
UPBP_HD float sample_1d(const float* func, const float* cdf, int count, float integral, float u, float& pdf,
                        int& offset)
{
    offset = cdf_interval(cdf, count, u);
    float du = u - cdf[offset];
    if (cdf[offset + 1] - cdf[offset] > 0.f) du /= cdf[offset + 1] - cdf[offset];
    pdf = integral > 0.f ? func[offset] / integral : 0.f;
    return (float(offset) + du) / float(count);
}

// direction of lat-long coordinates: longitude 2 pi u, latitude pi v (env_map.cuh:36-46)
UPBP_HD f3 env_dir(float u, float v)
{
    const float phi = u * 2.f * kPi, theta = v * kPi;
    const float st = sinf(theta);
    return mk(-st * cosf(phi), st * sinf(phi), cosf(theta));
}

// lat-long coordinates of a unit direction (env_map.cuh:48-61)
UPBP_HD void env_uv(f3 d, float& u, float& v)
{
    const float phi = (d.x != 0.f || d.y != 0.f) ? atan2f(d.y, d.x) : 0.f;
    const float theta = acosf(d.z);
    u = fminf(fmaxf(.5f - phi * .5f * kInvPi, 0.f), 1.f);
    v = fminf(fmaxf(theta * kInvPi, 0.f), 1.f);
}

// bilinear radiance at (u, v) (env_map.cuh:63-86)
UPBP_HD f3 env_radiance(const DEnvMap& m, float u, float v)
{
    const float xf = u * float(m.width), yf = v * float(m.height);
    const int x1 = clamp_int(int(xf), 0, m.width - 1), y1 = clamp_int(int(yf), 0, m.height - 1);
    const int x2 = x1 == m.width - 1 ? x1 : x1 + 1, y2 = y1 == m.height - 1 ? y1 : y1 + 1;
    const float tx = xf - float(x1), ty = yf - float(y1);
    auto px = [&](int x, int y) {
        const float* p = m.img + 3 * (size_t(y) * size_t(m.width) + size_t(x));
        return mk(p[0], p[1], p[2]);
    };
    return (px(x1, y1) * (1.f - tx) + px(x2, y1) * tx) * (1.f - ty) + (px(x1, y2) * (1.f - tx) + px(x2, y2) * tx) * ty;
}

// sine of the latitude at the centre of the pixel row holding v (env_map.cuh:88-103)
UPBP_HD float env_sin_theta(float v, int height)
{
    const float h = float(height);
    if (v < 1.f) return sinf(kPi * (float(int(v * h)) + .5f) / h);
    return sinf(kPi * ((h - 1.f) + .5f) / h);
}

// 2D pdf over (u, v) (distribution.cuh:110-124)
UPBP_HD float env_pdf_uv(const DEnvMap& m, float u, float v)
{
    const int iu = clamp_int(int(u * float(m.nu)), 0, m.nu - 1);
    const int iv = clamp_int(int(v * float(m.nv)), 0, m.nv - 1);
    if (m.cond_int[iv] * m.marg_int == 0.f) return 0.f;
    return m.cond_func[size_t(iv) * m.nu + iu] * m.marg_func[iv] / (m.cond_int[iv] * m.marg_int);
}

// direction sampled by luminance, its solid-angle pdf and radiance (env_map.cuh:105-122)
UPBP_HD f3 env_sample(const DEnvMap& m, float u0, float u1, float& pdf_w, f3& radiance)
{
    float pv, pu;
    int iv, iu;
    const float v = sample_1d(m.marg_func, m.marg_cdf, m.nv, m.marg_int, u1, pv, iv);
    const float u = sample_1d(m.cond_func + size_t(iv) * m.nu, m.cond_cdf + size_t(iv) * (m.nu + 1), m.nu,
                              m.cond_int[iv], u0, pu, iu);
    pdf_w = m.norm * (pu * pv) / env_sin_theta(v, m.height);
    radiance = env_radiance(m, u, v);
    return env_dir(u, v);
}

// radiance arriving along -direction and its solid-angle pdf (env_map.cuh:124-137)
UPBP_HD f3 env_lookup(const DEnvMap& m, f3 direction, float* pdf_w)
{
    const f3 d = normalize(direction);
    float u, v;
    env_uv(d, u, v);
    f3 rad = env_radiance(m, u, v);
    if (pdf_w) {
        *pdf_w = m.norm * env_pdf_uv(m, u, v) / env_sin_theta(v, m.height);
        if (*pdf_w == 0.f) rad = mk(0.f);
    }
    return rad;
}

} // namespace upbp
