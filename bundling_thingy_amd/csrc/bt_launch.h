// Launch wrappers of the kernels in bt_kernels.hip (one call per stage).
#pragma once
#include <hip/hip_runtime.h>

#include "bt_device.h"

namespace btk {
// mode bit 0: cell counters and the control block; bit 1: density counts
// and peak.
void clear(hipStream_t s, uint32_t grid, const BtJob* jobs, const uint32_t* base, int nj,
           BtControl* ctl, int mode);
void bin(hipStream_t s, uint32_t grid, const BtJob* jobs, const uint32_t* base, int nj);
void cell_scan(hipStream_t s, uint32_t grid, const BtJob* jobs, const uint32_t* base, int nj);
void tasks(hipStream_t s, uint32_t grid, const BtJob* jobs, const uint32_t* base, int nj,
           BtControl* ctl, BtTask* tasks_out);
void scatter(hipStream_t s, uint32_t grid, const BtJob* jobs, const uint32_t* base, int nj);
void pre_scan(hipStream_t s, uint32_t grid, const BtJob* jobs, const uint32_t* base, int nj);
// Kernel shapes of the neighbor mean (lanes per group x waves per
// workgroup): 0 = 16 x 4 (the default), 1 = 32 x 4, 2 = 16 x 2, 3 = 32 x 2,
// 4 = 16 x 8, 5 = 32 x 8. Results are the same under every shape.
#define BT_NM_SHAPE_COUNT 6u
// `samples`: the launch's samples, summed over its jobs.
void neighbor_mean(hipStream_t s, uint32_t grid, const BtJob* jobs, const BtTask* tasks_in,
                   BtControl* ctl, BtSplit split, unsigned long long samples, uint32_t shape);
// Workgroups of a shape's (persistent) neighbor-mean kernels resident per CU.
uint32_t neighbor_mean_blocks_per_cu(uint32_t shape);
uint32_t neighbor_mean_waves(uint32_t shape);
// Blur tile shapes: rows 0 = 64 x 4 (the default), 1 = 128 x 2, 2 = 32 x 8,
// 3 = 256 x 1 (columns x rows); columns 0 = 64 x 32 (the default), 1 = 64 x
// 16, 2 = 64 x 64, and the fused row + column kernel 3 = 64 x 32, 4 = 64 x
// 16 (blur_cols then does both passes and blur_rows is not launched; only
// for blurs of at most BT_FUSED_TAPS taps). Results are the same under every
// shape.
#define BT_ROW_SHAPE_COUNT 4u
#define BT_COL_SHAPE_COUNT 5u
void row_tile(uint32_t shape, uint32_t* tw, uint32_t* th);
uint32_t col_tile_h(uint32_t shape);
bool col_fused(uint32_t shape);
void blur_rows(hipStream_t s, uint32_t grid, const BtJob* jobs, const uint32_t* base, int nj,
               uint32_t shape = 0);
void blur_cols(hipStream_t s, uint32_t grid, const BtJob* jobs, const uint32_t* base, int nj,
               uint32_t shape = 0);
void normalize(hipStream_t s, uint32_t grid, const BtJob* jobs, const uint32_t* base, int nj);
// Density of the current records outside an iteration: clear, then splat.
void density_clear(hipStream_t s, uint32_t grid, const BtJob* jobs, const uint32_t* base, int nj);
void splat_points(hipStream_t s, uint32_t grid, const BtJob* jobs, const uint32_t* base, int nj);
void resample_emit(hipStream_t s, uint32_t grid, const BtJob* jobs, const uint32_t* base, int nj);
// Current positions in original sample order: 12-byte records {x, y, edge id}
// (split = 0) or n positions (x, y) followed by n edge ids (split = 1).
void unpermute(hipStream_t s, uint32_t grid, const BtJob* jobs, const uint32_t* base, int nj,
               float* out, uint32_t* cell_out, uint32_t want_cells, uint32_t split = 0);
// Binary16 (round to nearest even) copy of n floats; in is 8-byte aligned.
void half_plane(hipStream_t s, const float* in, uint16_t* out, uint32_t n);
// Upload set-up (weights from the packed area, status cleared), one workgroup per job.
void upload_setup(hipStream_t s, const BtJob* jobs, int nj, const uint32_t* area);
// One 64-bit word per job: its status word 0.
void job_words(hipStream_t s, const BtJob* jobs, int nj, unsigned long long* out);
// Saved state for rewinds (restore = 1 copies it back).
void save_points(hipStream_t s, uint32_t grid, const BtJob* jobs, const uint32_t* base, int nj,
                 int restore);
void save_density(hipStream_t s, uint32_t grid, const BtJob* jobs, const uint32_t* base, int nj,
                  int restore);
void endpoints(hipStream_t s, uint32_t grid, const BtJob* jobs, const uint32_t* base, int nj,
               const float* pts, uint32_t* table, uint32_t mask);
}  // namespace btk
