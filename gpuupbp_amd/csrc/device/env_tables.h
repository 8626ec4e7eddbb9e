// Host-side construction of the environment-map tables the device samples (env_map.h).
#pragma once

#include "upbp_device.h"
#include "../../../include/upbp_capi.h"

#include <cmath>
#include <vector>

namespace upbp {

// Host tables of the environment map (gpuupbp/host/env_map.hpp:42-125): per image row v the
// piecewise-constant function sin(theta_v) * luminance over nu = 2 * height columns, its
// running integral normalised into a CDF, and the marginal over rows of the row integrals.
struct EnvTables {
    std::vector<float> img, cond_func, cond_cdf, cond_int, marg_func, marg_cdf;
    float marg_int = 0.f;
};

// func -> (cdf, integral) of a 1D step function over [0, 1) with n steps
inline float build_cdf(const float* func, int n, float* cdf)
{
    cdf[0] = 0.f;
    for (int i = 1; i <= n; i++) cdf[i] = cdf[i - 1] + func[i - 1] / float(n);
    const float total = cdf[n];
    for (int i = 1; i <= n; i++) cdf[i] = total == 0.f ? float(i) / float(n) : cdf[i] / total;
    return total;
}

inline void build_env_tables(const upbp_env_map& e, EnvTables& t)
{
    const int h = int(e.height), w = int(e.width), nu = 2 * h, nv = h;
    t.img.assign(e.rgb, e.rgb + size_t(w) * h * 3);
    t.cond_func.resize(size_t(nv) * nu);
    t.cond_cdf.resize(size_t(nv) * (nu + 1));
    t.cond_int.resize(size_t(nv));
    for (int v = 0; v < nv; v++) {
        const float vc = (float(v) + .5f) / float(h); // row centre latitude / pi
        const float st = sinf(kPi * vc);
        for (int u = 0; u < nu; u++) {
            const float* p = e.rgb + 3 * (size_t(v) * w + u);
            t.cond_func[size_t(v) * nu + u] = st * (0.212671f * p[0] + 0.715160f * p[1] + 0.072169f * p[2]);
        }
        t.cond_int[size_t(v)] = build_cdf(&t.cond_func[size_t(v) * nu], nu, &t.cond_cdf[size_t(v) * (nu + 1)]);
    }
    t.marg_func = t.cond_int;
    t.marg_cdf.resize(size_t(nv) + 1);
    t.marg_int = build_cdf(t.marg_func.data(), nv, t.marg_cdf.data());
}

} // namespace upbp
