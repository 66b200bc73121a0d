// The device runtime: jobs (one uploaded graph each), per-iteration launch
// plans, batched uploads and iterations, hipGraph capture, readbacks.
#pragma once
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "bt_common.h"
#include "bt_device.h"
#include "bt_launch.h"

namespace bt {

// One bundling job: an uploaded graph and its device state.
struct Job {
  bt_options o{};
  uint32_t n_ctrl = 0, n_paths = 0, n = 0, valid_paths = 0;
  uint64_t capacity = 0, reserved = 0;
  float norm[5] = {0, 0, 0, 0, 0};
  uint32_t dw = 0, dh = 0;
  uint32_t max_cells = 1;
  uint32_t iteration = 0, parity = 0;
  // first iteration that splats, blurs and normalizes the density (0: every
  // one, as Gpu::iterate does; the served form computes only the density it
  // returns, that of a job's last iteration)
  uint32_t density_from = 0;
  // iteration whose density the planes hold (density computed on read)
  uint32_t density_at = 0;
  uint32_t sched = 1;           // distinct parameter sets
  std::vector<bt_iter_params> params;
  std::vector<uint32_t> splits; // grid cells per radius, per schedule entry
  // kernel shapes per schedule entry: neighbor mean | row blur << 8 | column
  // blur << 16 (btk shape indices; 0 = defaults)
  std::vector<uint32_t> kshape;
  // control points and path starts: in the upload area, during the upload only
  const float* in_ctrl = nullptr;
  const uint32_t* in_starts = nullptr;
  DevBuf path_scan, paths, bsum, status;
  DevBuf rec0, rec1, rec_saved, sorted, sid, cellrank, cells, edge_of, pre_x, pre_y;
  DevBuf dens_cnt, dens_row, dens_col, dens_norm, dens_saved, peak, wt;

  const bt_iter_params& at(uint32_t it) const { return params[it < sched ? it : sched - 1]; }
  // After n iterations ran from `iteration`: the count, the parity of the
  // current records, and which iteration's density the planes hold.
  void advanced(uint32_t n) {
    if (n == 0) return;
    if (dens_cnt.p && iteration + n - 1u >= density_from) density_at = iteration + n;
    iteration += n;
    parity ^= (n & 1u);
  }
  uint32_t entry(uint32_t it) const { return it < sched ? it : sched - 1; }
  // Sets schedule entry e's cell split; returns true if the cell buffers grew.
  bool set_split(uint32_t e, uint32_t k);
  // Job-table entry for iteration `it` reading rec[parity_at]; `force_density`:
  // with the density stages whatever density_from says.
  BtJob dev(uint32_t it, uint32_t parity_at, bool force_density = false) const;
  // Device bytes the job's buffers hold (pool blocks, rounded up).
  size_t device_bytes() const;
  // Grid bins the largest schedule entry needs with n samples, as an upload
  // sizes them (the split given in the options, else the default and then
  // the sample-density split).
  static uint32_t max_cells_for(const bt_options& o, uint64_t n);
  // Device bytes (pool blocks) of a job of n samples as upload_jobs
  // allocates it, plus its share of the upload area (held while resampling).
  static size_t footprint(const bt_options& o, uint64_t n, uint32_t n_points, uint32_t n_starts,
                          bool rewindable);
  // Allocate this job's buffers from `pool` (before any allocation).
  void set_pool(DevPool* pool);
};

struct Plan {
  std::vector<unsigned char> host;
  uint32_t grid[BT_K_COUNT] = {};
  uint32_t kshape = 0;  // kernel shapes of the launch (Job::kshape encoding)
  uint32_t njobs = 0;
  uint64_t samples = 0;  // samples of the plan's jobs
  size_t base_off = 0;  // byte offset of the bases block
};

// Input of one graph of an upload.
struct GraphIn {
  const float* xy;
  uint32_t n_points;
  const uint32_t* starts;
  uint32_t n_starts;
  bt_options opts;
  // the count pass's path offsets (resample_counts), if already computed
  const uint64_t* path_scan = nullptr;
};

class Engine {
 public:
  // cu_mask: 0 for every CU, else the CUs the engine's stream may use, one
  // bit per CU of each group of 32 (the same pattern in every group; on
  // MI355X each byte of it is one shader engine of every XCD).
  explicit Engine(const bt_options& o, uint32_t cu_mask = 0);
  ~Engine();

  const bt_options& options() const { return opts_; }
  // Replaces the current jobs with the given graphs; returns sample counts.
  void upload(const std::vector<GraphIn>& gs, std::vector<uint32_t>* samples);
  // Uploads graphs as new jobs appended to `out`: sample counts on the host,
  // then one copy and one kernel pass on the engine's stream. `async`:
  // return once the launches are queued (behind any queued iterations; the
  // server's admissions), leaving the emit status word unchecked for the
  // caller (Job::status); otherwise wait and check it. `rewindable`: keep
  // the post-upload copies bt_rewind restores.
  void upload_jobs(const std::vector<GraphIn>& gs, std::vector<std::unique_ptr<Job>>* out,
                   bool async = false, bool rewindable = true);
  void iterate(uint32_t n, bool use_graph);
  // Advances an explicit job set by one iteration (server tick); with
  // wait = false it returns once the launches are queued.
  void iterate_jobs(const std::vector<Job*>& js, bool wait = true);
  // Queues the readback of a job's points (12-byte records, Layout order)
  // and, if density is on, its normalized density into host buffers (pinned,
  // so the copies run asynchronously), then the upload's status word (4
  // bytes) right after the records and the density at rec12, and records
  // `done` on the stream. `density` must be rec12 + 12 n when given.
  // Uses the job's rewind copy as staging: the job must not be rewound after.
  void queue_readback(Job& j, void* rec12, float* density, hipEvent_t done);
  void rewind();
  void profile(uint32_t n, double* stage_ms, uint32_t nst, double* total_ms);
  // Picks the grid cell split of each of the first n iterations by timing
  // every candidate from the same state; leaves the jobs rewound.
  void autotune(uint32_t n, const uint32_t* cand, uint32_t ncand, uint32_t* chosen,
                double* best_ms);
  // Picks the kernel shapes (neighbor mean, row blur, column blur) of each
  // of the first n iterations by timing each stage under every shape from
  // the same state (the others at their current choice); leaves the jobs
  // rewound. shapes[3 i + k], stage_ms[3 i + k] (may be null).
  void autotune_shapes(uint32_t n, uint32_t* shapes, double* stage_ms);
  // Sets schedule entry `it`'s kernel shapes on every job (all entries when
  // it is ~0u).
  void set_shapes(uint32_t it, uint32_t nm, uint32_t row, uint32_t col);

  void read_paths(uint32_t job, uint32_t* rec3, size_t n);
  void read_points(uint32_t job, void* rec12, size_t n);
  // The same samples as n positions (x, y) and n edge ids (ids may be null).
  void read_positions(uint32_t job, float* xy, uint32_t* ids, size_t n);
  void read_density(uint32_t job, uint32_t kind, void* out, size_t n);
  void read_grid(uint32_t job, uint32_t* cell_of, size_t n, uint32_t* counts, size_t ncells);
  // Sorted sample indices of the distinct path endpoint positions.
  std::vector<uint32_t> read_endpoints(uint32_t job);
  // The same through the per-device static scratch: no device allocation.
  void read_endpoints_noalloc(uint32_t job, uint32_t* indices, size_t capacity, uint32_t* count);
  // The export image of a job (the blank frame without a graph), RGBA8.
  void render(const bt_render_options& ro, uint8_t* rgba, size_t n_bytes, uint32_t job = 0);
  void render_size(uint32_t* w, uint32_t* h, uint32_t job = 0);
  static void read_points_of(Engine& e, Job& j, void* rec12, size_t n);
  // Unpermutes a job's current samples into staging_ (records, or split).
  static void unpermute_to_staging(Engine& e, Job& j, bool split);
  static void read_density_of(Engine& e, Job& j, uint32_t kind, void* out, size_t n);
  // Computes the density of a job's current positions if its planes hold an
  // older iteration's (BT_FLAG_DENSITY_ON_READ); blocks until done.
  void ensure_density(Job& j);

  // Heavy tasks split by the last iteration run, and their parts.
  void split_counts(uint32_t* tasks, uint32_t* parts);
  // Executed work of the last neighbor-mean launch: wave-steps (one
  // candidate tested against the 64 lanes' samples), neighbors found for
  // the movable samples, and the candidates of the tasks' streams.
  void work_counts(uint64_t* steps, uint64_t* neighbors, uint64_t* candidates,
                   uint64_t* group_steps);

  size_t njobs() const { return jobs_.size(); }
  Job& job(uint32_t i);
  size_t device_bytes() const;
  hipStream_t stream() const { return stream_; }
  // With a CU mask: move the engine's work to the stream over every CU
  // (wide) or back to the masked one. The new stream first waits for all
  // work queued on the old one, so the two act as one ordered stream.
  void set_wide(bool wide);
  bool masked() const { return streams_[1] != nullptr; }
  void sync();

  std::string err;

 private:
  Plan make_plan(const std::vector<Job*>& js, const std::vector<uint32_t>& iters) const;
  // A plan over explicit job-table entries (block bases from the entries and
  // the kernel shapes).
  Plan make_plan_of(const std::vector<Job*>& js, const std::vector<BtJob>& entries,
                    uint32_t kshape = 0) const;
  // Copies a plan to device memory through a ring of pinned host slots (the
  // copy is asynchronous; a slot is reused once its previous copy is done).
  void put_plan(const Plan& p, void* dst, hipStream_t st = nullptr);
  // A captured schedule's iteration: its task list runs on tasks_st_ and
  // (side) its density stages on side_.
  struct Overlap {
    bool side = false;        // density stages on side_
    bool prev = false;        // the previous iteration's did (and cleared this one's counts)
    hipEvent_t prev_done = nullptr, nm_done = nullptr, side_done = nullptr;
    hipEvent_t scan_done = nullptr, tasks_done = nullptr;  // the task list on tasks_st_
    const Plan* next = nullptr;  // the next iteration (side_ clears its density counts)
    const unsigned char* next_dev = nullptr;
  };
  void launch_iteration(const Plan& p, const unsigned char* dev, bool density,
                        hipEvent_t* ev = nullptr, const Overlap* ov = nullptr);
  void ensure_tasks(const std::vector<Job*>& js);
  // Forgets the previous launch's candidate count (the heavy-task threshold
  // falls back to its upper bound).
  void reset_split_estimate();
  void drop_graphs();
  std::vector<Job*> all_jobs();

  bt_options opts_;
  std::unique_ptr<DevPool> pool_;  // declared before the jobs: outlives their buffers
  int device_ = 0;
  int cus_ = 256;
  uint32_t nm_per_cu_[BT_NM_SHAPE_COUNT] = {};  // resident neighbor-mean workgroups per CU
  hipStream_t stream_ = nullptr;
  hipStream_t streams_[2] = {nullptr, nullptr};  // [0] masked (or the only one), [1] every CU
  int stream_cus_[2] = {0, 0};
  int active_ = 0;
  hipEvent_t switch_ev_ = nullptr;
  hipStream_t side_ = nullptr;        // captured schedules' density stages
  hipStream_t tasks_st_ = nullptr;    // ... and task lists
  std::vector<hipEvent_t> ov_ev_;     // their fork and join events
  std::vector<std::unique_ptr<Job>> jobs_;
  DevBuf ctl_, tasks_, plans_, graph_plans_, staging_, staging2_, split_, retire_plans_;
  struct Pinned {
    void* p = nullptr;
    size_t bytes = 0;
    hipEvent_t ev = nullptr;
    bool pending = false;
  };
  std::vector<Pinned> pinned_;
  // pinned staging of upload inputs: a ring, so an asynchronous upload
  // never refills a buffer whose copy is still queued
  std::vector<Pinned> up_pinned_;
  uint32_t up_next_ = 0;
  void* upload_staging(size_t bytes, Pinned** slot);
  // pinned buffer of blocking readbacks (images): one, grown when needed,
  // kept for the next call (the caller waits for its copy before reuse)
  Pinned rd_pinned_;
  void* readback_staging(size_t bytes);
  uint32_t pinned_next_ = 0;
  uint32_t retire_next_ = 0;
  size_t slot_bytes_ = 0;
  uint32_t eager_slot_ = 0;
  uint64_t generation_ = 0;
  struct GraphEntry {
    hipGraphExec_t exec = nullptr;
    hipGraph_t graph = nullptr;
    uint32_t last_nm_grid = 0;  // neighbor-mean grid of its last iteration
  };
  uint32_t last_nm_grid_ = 0;   // neighbor-mean grid of the last launched iteration
  std::map<std::vector<uint64_t>, GraphEntry> graphs_;
  std::map<std::vector<uint64_t>, std::unique_ptr<DevBuf>> graph_bufs_;
};

}  // namespace bt
