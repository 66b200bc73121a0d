// Device-side data layout shared by the HIP kernels and the C++ host.
//
// One bundling job (one uploaded graph) owns its device state. A launch covers
// any number of jobs through a job table; each job's work items (points, cells,
// pixels, tiles) occupy a contiguous range of workgroups given by a per-kind
// block-base array, so a workgroup never straddles two jobs. The plain drop-in
// runs a table of one job; the batched and served forms run many.
#pragma once
#include <stdint.h>

#define BT_FIXED_BIT 0x80000000u   // sample is a path endpoint (never moves)
#define BT_MAX_TAPS 64             // blur half-width limit (sigma <= 21 px)
#define BT_FUSED_TAPS 32           // widest blur the fused row + column kernel takes
#define BT_FIX_ONE 16777216.0f     // 2^24: fixed-point quantum of neighbor sums
#define BT_TASK_POINTS 64          // query points per neighbor-mean task (one wave)

// Block granularities (work items per workgroup) of the batched kernels.
#define BT_BIN_BLOCK 256           // points per workgroup in scatter/readback
#define BT_RANK_CHUNKS 8           // binning: 64-sample chunks per wave
#define BT_RANK_BLOCK (256 * BT_RANK_CHUNKS)  // points per workgroup in binning
#define BT_SCAN_ITEMS 4096         // cells per workgroup in the cell scan (256 x 16)
#define BT_TASK_BLOCKS 256         // task blocks per workgroup in task emission (one per thread)
#define BT_CLEAR_ITEMS 4096        // words per workgroup in the clear kernel
#define BT_NORM_ITEMS 1024         // pixels per workgroup in normalize
#define BT_EMIT_PATHS 4            // paths per workgroup in resample_emit (one wave each)

struct __attribute__((aligned(16))) BtJob {
  // ---- per-job device state (constant for the job's lifetime)
  float* rec[2];          // ping-pong sample records {x, y, id} (12 B each), slot order
  float4* sorted;         // samples sorted by grid cell: {x, y, fx bits, fy bits}
  uint32_t* sid;          // id of each sorted sample (original index | BT_FIXED_BIT)
  uint2* cellrank;        // per slot: fine grid cell and rank within the cell
  uint32_t* cells;        // cell counts, then exclusive starts (max_cells + 1)
  uint32_t* bsum;         // per-scan-block partial sums (cell scan)
  uint64_t* pre_x;        // exclusive prefix sums of the sorted fixed-point x (n + 1)
  uint64_t* pre_y;        // same for y
  uint32_t* edge_of;      // per original index: path (edge) id
  uint32_t* dens_cnt;     // density counts (u32) W*H
  float* dens_row;        // horizontally blurred density
  float* dens_col;        // vertically blurred density
  float* dens_norm;       // normalized density in [0, 1]
  uint32_t* peak;         // [2] f32 bits of the blurred peak (double-buffered)
  uint32_t* status;       // [4] error flags (resample count/emit mismatch, ...)
  // ---- resampling inputs (upload only)
  const float* ctrl;      // control points (x, y)
  const uint32_t* starts; // path starts (n_paths + 1)
  uint64_t* path_scan;    // per path: (count << 32) | valid, then exclusive scan
  uint32_t* paths;        // per path: {point_offset, segment_offset, count}
  // ---- sizes
  uint32_t n;             // samples
  uint32_t n_paths;
  uint32_t dw, dh;        // density size
  // ---- per-iteration parameters
  uint32_t gx, gy;        // grid dims: bins per row, rows
  uint32_t n_cells;       // gx * gy
  uint32_t taps;          // blur half-width
  float res_x, res_y;     // texel resolution per axis (reference: map size - 1)
  float ts_x, ts_y;       // texel scale = res * 0.5
  float cell;             // row height in texels
  float inv_cell;         // 1 / cell (window bounds only; binning divides)
  float cell_w;           // bin width along x in texels
  float inv_cell_w;       // 1 / cell_w
  uint32_t task_bins;     // bins per task block along x
  uint32_t task_bx;       // task blocks per row: ceil(gx / task_bins)
  float radius;           // texels
  float eps;              // geometric safety margin of the run classification, texels
  float r2;               // radius^2 (texels^2)
  float reach_x, reach_y; // window half-width: radius + 4 * pad
  float norm_min_x, norm_min_y, norm_scale, norm_tr_x, norm_tr_y;  // Graph::normalization
  uint32_t parity;        // which rec[] holds the current positions
  uint32_t capacity;      // sample capacity (reference budget model)
  float* wt;              // blur weights [taps + 1] (written once, by the upload set-up)
  // ---- saved state (rewind source; the served form stages results in rec_saved)
  float* rec_saved;       // samples after upload, {x, y, id} records
  float* dens_saved;      // density planes after upload (counts, row, col, norm), or null
};

// Per-launch work description: for each kernel kind, the first workgroup of
// each job (njobs + 1 entries, last = total workgroups).
enum BtKind {
  BT_K_CLEAR = 0,
  BT_K_BIN,
  BT_K_SCAN,
  BT_K_TASK,
  BT_K_ROW,
  BT_K_COL,
  BT_K_NORM,
  BT_K_PATH,     // resample count: 256 paths per workgroup
  BT_K_EMIT,     // resample emit: BT_EMIT_PATHS paths per workgroup
  BT_K_PSCAN,    // path scan: BT_SCAN_ITEMS paths per workgroup
  BT_K_PTS,      // sample prefix scan: BT_SCAN_ITEMS samples per workgroup
  BT_K_RANK,     // binning: BT_RANK_BLOCK samples per workgroup
  BT_K_COUNT
};

#define BT_XCDS 8  // MI355X: workgroups are dealt round-robin to 8 XCDs (one L2 each)

#define BT_CTR_STRIDE 32  // words between work-queue cursors: one 128-byte line each

struct BtControl {
  uint32_t task_total;    // tasks emitted this iteration
  uint32_t heavy_count;   // heavy tasks recorded for splitting this iteration
  uint32_t item_count;    // parts of heavy tasks recorded
  uint32_t pad0;
  unsigned long long cand_prev;  // candidates tested by the previous neighbor-mean
                                 // launch (0: none since upload or rewind)
  unsigned long long samples_prev;  // samples of that launch's jobs
  uint32_t pad[BT_CTR_STRIDE - 8];
  // per-XCD work-queue cursors (part x at [x * BT_CTR_STRIDE]) of the
  // neighbor-mean kernel (xq) and of the heavy-part kernel (iq)
  uint32_t xq[BT_XCDS * BT_CTR_STRIDE];
  uint32_t iq[BT_XCDS * BT_CTR_STRIDE];
};

// Heavy tasks (candidate stream longer than a threshold H) are split into
// parts run by many waves; their exact partial sums meet in per-task
// accumulators, and the last part to finish writes the means. H is half the
// expected candidates per wave, clamped to [BT_HEAVY_LO, BT_HEAVY_HI]
// (BT_HEAVY_HI when there is no previous launch): a task much longer than the
// average would otherwise run alone at the end. The expectation is the
// previous launch's candidates scaled by this launch's samples over that
// launch's (the server alternates launches of very different sizes). The main kernel stores each
// heavy task's set-up (lane-group ranges and inside sums, BT_REC_WORDS words)
// so the parts only stream.
#define BT_HEAVY_LO 2048u
#define BT_HEAVY_HI 16384u
#define BT_PART_STEPS_MIN 512u
#define BT_REC_WORDS 1088u
#define BT_MAX_PARTS 64u
#define BT_MAX_HEAVY 16384u
#define BT_MAX_ITEMS 262144u
#define BT_MAX_WG 8192u       // workgroups of a neighbor-mean launch (per-WG candidate sums)

struct BtHeavy {
  uint32_t task;          // index into the task list
  uint32_t parts;         // number of parts
  uint32_t done;          // parts finished
  uint32_t pad;
};

struct BtSplit {          // scratch of the heavy-task split (one per launch)
  BtHeavy* heavy;         // [BT_MAX_HEAVY]
  uint2* items;           // [BT_MAX_ITEMS]: (heavy index, part)
  long long* acc;         // [BT_MAX_HEAVY][3][64]: per-lane x, y, count sums
  unsigned long long* wg_cand;  // [BT_MAX_WG]: candidates tested per workgroup
  uint32_t* rec;          // [BT_MAX_HEAVY][BT_REC_WORDS]: set-up of each heavy task
  // executed work of the last launch per workgroup (main kernel, heavy
  // kernel): wave-steps (one candidate against the 64 lanes' samples) and
  // neighbors found for the samples finished there
  unsigned long long* wg_steps;   // [BT_MAX_WG]
  unsigned long long* wg_nbr;     // [BT_MAX_WG]
  unsigned long long* wg_hsteps;  // [BT_MAX_WG]
  unsigned long long* wg_hnbr;    // [BT_MAX_WG]
  unsigned long long* wg_gsteps;  // [BT_MAX_WG]: of wg_steps, those of lane-group streams
};

struct __attribute__((aligned(16))) BtTask {
  uint32_t job, start, end, pad;
};
