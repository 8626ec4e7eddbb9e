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
void launch_camera_finish(const IterParams& p, const Bufs& b, float res_x, float res_y, hipStream_t s);

} // namespace upbp
