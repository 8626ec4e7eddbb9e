// Host side of the learned proposal's network (guide_mlp.hip): parameters, Adam state and the
// activations of one batch in device memory, and the launches over them.
#pragma once

#include "lbvh.h"
#include "../../../include/upbp_capi.h"

namespace upbp {

struct GuideNet {
    upbp_host::DevBuf prm, adam_m, adam_v, grad;  // FP32 master, Adam moments, flat gradient
    upbp_host::DevBuf wq, bq;                     // BF16 padded weights, FP32 padded biases
    upbp_host::DevBuf gW, gb, loss;               // padded weight / bias gradients, loss sum
    upbp_host::DevBuf x0, h1, h2, h3, out;        // activations (BF16) and raw outputs (FP32, rows of 48)
    upbp_host::DevBuf d_o, d3, d2, d1;            // deltas (BF16)
    upbp_host::DevBuf grid_in;                    // record-grid inputs
    upbp_host::DevBuf batch;                      // training batch drawn from the sample pool
    int step = 0;

    ~GuideNet()
    {
        for (upbp_host::DevBuf* b : {&prm, &adam_m, &adam_v, &grad, &wq, &bq, &gW, &gb, &loss, &x0, &h1, &h2, &h3,
                                     &out, &d_o, &d3, &d2, &d1, &grid_in, &batch})
            b->release();
    }
    static uint32_t param_count();
    int init(uint32_t seed);                                    // He-initialised weights, Adam reset
    int set_params(const float* host, hipStream_t s);           // FP32 master from the host, Adam reset
    int get_params(float* host, hipStream_t s);
    int quantise(hipStream_t s);                                // BF16 copies of the master
    int reserve(uint32_t n);
    // raw outputs of n device samples (stride floats each: position in [0,1]^3, arrival
    // direction) into out (rows of 48 floats, the first 40 used)
    int forward(const float* samples, int stride, uint32_t n, hipStream_t s);
    // the raw outputs alone into o (pad16(n) x 48 floats); nothing else is written
    int forward_into(const float* samples, int stride, uint32_t n, float* o, hipStream_t s) const;
    // loss sum and flat gradient of n device training samples (10 floats each)
    int gradient(const float* samples, uint32_t n, hipStream_t s);
    // the same for n samples drawn at random (seeded by draw) from a device pool of n_pool
    int gradient_drawn(const float* pool, uint32_t n_pool, uint32_t n, uint32_t draw, hipStream_t s);
    int adam_step(float lr, hipStream_t s);
    // proposal records of a res^3 grid over [0,1]^3 into recs (device)
    int records(int res, float4* recs, hipStream_t s);
};

} // namespace upbp
