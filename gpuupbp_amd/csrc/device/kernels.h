// Launch wrappers shared by the renderer host code and the pass kernels.
#pragma once

#include "passes.h"

namespace upbp {

void launch_light_generate(const DScene& sc, const IterParams& p, const Bufs& b, hipStream_t s);
void launch_light_bounce(const DScene& sc, const IterParams& p, const Bufs& b, int bounce, hipStream_t s);
void launch_camera_generate(const DScene& sc, const IterParams& p, const Bufs& b, hipStream_t s);
void launch_camera_bounce(const DScene& sc, const IterParams& p, const Bufs& b, int bounce, hipStream_t s);
void launch_seg_merge(const DScene& sc, const IterParams& p, const Bufs& b, int kind, int bounce, hipStream_t s);
int merge_grid_blocks(int bpc);
// Morton keys of the B-B1D query pieces of one bounce (n of them) and the gather of the
// pieces into pieces_x in sorted order
void launch_piece_keys(const Bufs& b, int bounce, uint32_t n, uint32_t* keys, uint32_t* vals, hipStream_t s);
void launch_piece_gather(const Bufs& b, uint32_t n, const uint32_t* vals, hipStream_t s);
// splat and FP64 accumulation; an iteration whose ovf[0] is set is not accumulated
void launch_camera_finish(const IterParams& p, const Bufs& b, float res_x, float res_y, const uint32_t* ovf,
                          hipStream_t s);

} // namespace upbp
