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
// Units of 64 consecutive B-B1D query pieces (at most cap_units) ordered by the Morton code of
// their middle piece: keys and unit indices (the unit count also to *nu_dev, for the sort), then
// the pieces gathered into pieces_x unit by unit in the sorted order, with the merge kernel's
// work ranges (eight per-XCD ranges with per_xcd, else one)
void launch_unit_keys(const Bufs& b, int bounce, uint32_t cap_units, uint32_t* keys, uint32_t* vals, uint32_t* nu_dev,
                      hipStream_t s);
void launch_unit_gather(const Bufs& b, int bounce, uint32_t cap_units, const uint32_t* vals, int per_xcd,
                        hipStream_t s);
// Batched stages: one launch runs a stage for every job of a batch. bmap (blocks entries) gives
// each block its job and its block within the job, in the job's own grid for that stage.
void launch_light_generate_batch(const BatchJob* jobs, const uint2* bmap, uint32_t blocks, hipStream_t s);
void launch_light_bounce_batch(const BatchJob* jobs, const uint2* bmap, uint32_t blocks, int bounce, hipStream_t s);
void launch_camera_generate_batch(const BatchJob* jobs, const uint2* bmap, uint32_t blocks, hipStream_t s);
void launch_camera_bounce_batch(const BatchJob* jobs, const uint2* bmap, uint32_t blocks, int bounce, hipStream_t s);
void launch_camera_finish_batch(const BatchJob* jobs, const uint2* bmap_pix, uint32_t blocks_pix, const uint2* bmap_acc,
                                uint32_t blocks_acc, hipStream_t s);
// persistent: every resident block walks every job's pieces of the bounce
void launch_seg_merge_batch(const BatchJob* jobs, uint32_t njobs, int kind, int bounce, hipStream_t s);
// splat and FP64 accumulation; an iteration whose ovf[0] is set is not accumulated
void launch_camera_finish(const IterParams& p, const Bufs& b, float res_x, float res_y, const uint32_t* ovf,
                          hipStream_t s);

} // namespace upbp
