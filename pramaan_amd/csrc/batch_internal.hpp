// batch_internal.hpp -- the batched engine's state (pm_batch) and per-LP
// records, shared by batch.hip and the learned warm start (learned.hip).
#pragma once

#include <cstdint>
#include <map>
#include <utility>
#include <vector>

#include "common.hpp"
#include "control.hpp"

namespace pm {

// Per-LP offsets into the concatenated arrays, options and results.
struct BatchDesc {
    int32_t m, n, nnz;
    int32_t lanes_a, lanes_at;
    int32_t warm_x, warm_y;
    int32_t warm_accepted;  // -1 no warm start, 1 kept, 0 replaced by the cold start
    double omega0;          // initial primal weight; <= 0: ||c_hat|| / ||b_hat|| x omega_scale
    double omega_scale;     // learned factor on the default weight (<= 0: none); set by the
                            // learned warm start, cleared with the warm start
    double eta;             // 0.99 / ||A_hat||_2, 0 until the first run measured it
    double start_kkt;       // FP64 relative KKT of the starting point that was used
    int64_t rp_off, cp_off, nz_off, row_off, col_off;
    // row segments of the stream iterate kernels: consecutive rows with at
    // most kStreamNnz nonzeros together (a longer row alone), in seg_a
    // (rows of A) and seg_at (rows of A^T) from these offsets
    int32_t sa_off, sa_cnt, st_off, st_cnt;
    // slices of the sliced iterate kernels (64 rows each, rows sorted by
    // length inside every group of ga rows of A / gt rows of A^T), from
    // these offsets in the sliced arrays of A and A^T
    int32_t la_off, lt_off;
    int32_t ga, gt;
    int32_t bucket;  // size class of the iterate launches (size_class)
    // results
    int32_t status, iterations, restarts, refinements, polishes;
    double rel_kkt, rel_p, rel_d, rel_g, pobj, dobj, omega;
    uint64_t t_start, t_end;
};

// What the grid-wide iterate kernels need of one LP for the next chunk,
// written by the init and check phases.
struct LpIter {
    float tau, sig;
    int32_t k0;      // in-round Halpern index of the chunk's first iteration
    int32_t chunk;   // iterations in the chunk (0: LP finished)
    int32_t use_g;   // dual shift of a refinement round present
    int32_t row_off, col_off;
    int32_t pad;
};

// Nonzeros one workgroup of the stream iterate kernels stages at a time.
constexpr int kStreamNnz = 1024;

// Kernel codes of the iterate halves besides 1..32 lanes per row: the
// stream kernel, and the two kernels over the sliced layout, gathering
// from global memory (sliced) or from a copy in LDS (staged).
constexpr int kStreamKernel = 0, kSlicedKernel = -1, kStagedKernel = -2;
// Sliced layout: groups of consecutive rows, one group per work item;
// inside a group the rows are sorted by length and cut into slices of
// kSliceRows (one row per lane of a wavefront).  A takes groups of
// kSliceGroup rows, or kSliceGroupWide for an LP whose groups of
// kSliceGroup rows already reach across more than half its columns (then a
// group's column window hardly grows with the group, and the larger group
// halves the window staged per nonzero); A^T takes kSliceGroupT.  Each
// LP's groups are then balanced: the same count, equal sizes rounded up to
// kSliceRows (balanced_group), no small last group.
constexpr int kSliceRows = 64, kSliceGroup = 1024, kSliceGroupWide = 2048, kSliceGroupT = 2048;
// Largest column window (floats) the staged kernel holds in LDS, and the
// largest any sliced kernel takes (its indexes are 16-bit).
constexpr int kStagedMaxVec = 15 * 1024, kSlicedMaxWin = 65536;

// A CSR in sliced ELL form.  Slice s holds len[s] columns of 64 lanes:
// entry (j, lane) at ptr[s] + j * 64 + lane; pos[s * 64 + lane] is the
// lane's row as an offset inside its group (-1 for a padding lane).  The
// column indexes are 16-bit, relative to the group's window win[s] =
// {first column, columns} (every column the group's rows hold lies inside
// it; a layout with a window wider than kSlicedMaxWin is not used).
struct SlicedCsr {
    DevBuf<float> val;
    DevBuf<uint16_t> idx;
    DevBuf<int> ptr, len, pos;
    DevBuf<int2> win;
};

// One item of the iterate kernels' work list: a group of consecutive rows
// of one LP (rows of A for the dual half, of A^T for the primal half) with
// everything the kernel needs of that LP for the current chunk, so a block
// reads one 32-byte record per item and no per-LP table.
struct WorkItem {
    int32_t row0;   // first global row of the group
    int32_t cnt;    // rows in the group
    int32_t k0;     // in-round Halpern index of the chunk's first iteration
    int32_t chunk;  // iterations in the chunk
    int32_t use_g;  // dual shift of a refinement round present (dual half)
    int32_t off;    // column offset (dual half) or row offset (primal half) of the LP
    float step;     // sigma (dual half) or tau (primal half)
    int32_t lp;
};

}  // namespace pm

struct pm_batch {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    hipGraph_t graph = nullptr, graph_tail = nullptr;
    hipGraphExec_t exec = nullptr, exec_tail = nullptr;
    int count = 0;
    bool exact = true;
    double tol = 1e-6;
    int max_iterations = 100000, check_frequency = 64, max_refinements = 8;
    std::vector<pm::BatchDesc> desc;
    std::vector<int> order;
    std::vector<int> sense;
    std::vector<double> offset;
    std::vector<double> c_model;  // model-sense costs, concatenated
    pm::DevBuf<double> c_model_dev;  // the same on the device (uploaded for the learned warm start)
    int64_t rows_total = 0, cols_total = 0, nnz_total = 0;
    pm::DevBuf<int> rp, ci, cp, ri, d_order;
    pm::DevBuf<float> va, vat, oa32, oat32;
    pm::DevBuf<double> oa64, oat64;
    pm::DevBuf<double> cmin, lr, ur, lx, ux, dr, dc, wx, wy;
    pm::DevBuf<float> c32, lr32, ur32, lx32, ux32, g32, x, xe, x0, xt, xl, y, y0, yt, yl;
    pm::DevBuf<double> xtot, ytot, xbase, ybase, xsave, ysave, xpol;
    // The main solve's state while a feasibility polish runs: FP32 blocks
    // [x | x0 | xl | c32 | lx32 | ux32] of cols_total each, then
    // [y | y0 | yl | lr32 | ur32 | g32] of rows_total each; FP64 x_base, y_base.
    pm::DevBuf<float> hold32;
    pm::DevBuf<double> hold64;
    int polish = 1;  // feasibility polishing (pm_pdhg_options.feasibility_polishing)
    pm::DevBuf<pm::BatchDesc> d_desc;
    pm::DevBuf<int> grp, gcp, wcnt;
    pm::DevBuf<pm::WorkItem> work_r, work_c;
    // Blocks of the round and check passes: {LP, block} per block of
    // kCheckRows rows of A (ck_r) and of A^T (ck_c) of every LP with a
    // chunk in flight, the same whichever iterate kernel runs; per LP
    // {first A block, A blocks, first A^T block, A^T blocks}; per-block
    // partial sums of the check.
    pm::DevBuf<int2> ck_r, ck_c;
    pm::DevBuf<int4> wrange;
    pm::DevBuf<double> part_r, part_c;
    pm::DevBuf<int2> seg_a, seg_at;  // {local first row, rows} per segment
    pm::SlicedCsr sl_a, sl_at;       // sliced copies of A and A^T
    int win_a = 0, win_at = 0;       // widest group window of A and A^T (staged kernel's LDS)
    pm::DevBuf<pm::LpIter> iter;
    int lanes_a = 8, lanes_at = 8;  // iterate kernels: lanes per row, or a kernel code
    // Kernels once the running LPs hold at most tail_nnz nonzeros (0 =
    // never switch): lanes per row, many small work items per LP, for the
    // latency of a launch rather than its throughput.
    int tail_a = 8, tail_at = 8;
    int64_t tail_nnz = 0;
    int mode_a = 8, mode_at = 8;    // kernels the launches being issued use
    int iter_grid = 2048;           // workgroups of a launch when occupancy is unknown (8 per CU)
    int num_cu = 256;
    // Workgroups of a persistent launch per (kernel, dynamic LDS bytes):
    // as many as the device holds at once.
    std::map<std::pair<const void*, size_t>, int> grid_cache;
    // Grid of the iterate kernels in resident grids (1, or 4 when the
    // autotuner finds it faster: items of uneven cost then balance as the
    // hardware starts the queued workgroups); the tail kernels use 1.
    int mult_a = 1, mult_at = 1, mode_mult_a = 1, mode_mult_at = 1;
    // Size classes of the iterate launches.  With nbuckets > 1 the LPs of
    // class c (BatchDesc::bucket) get their own work lists (items from
    // off_r[c] / off_c[c] in work_r / work_c, counts in wcnt[4 + 2c] and
    // wcnt[5 + 2c]), their own launch per half iteration and their own
    // kernel (blanes_*, bmult_*): items of one class cost about the same,
    // so a persistent grid's static assignment stays balanced.  1: one
    // launch over every LP.  The tail kernels always run over every LP.
    static constexpr int kBuckets = 3;
    int nbuckets = 1;
    bool mode_bucketed = false;
    int bucket_lps[kBuckets] = {0, 0, 0};
    int64_t off_r[kBuckets] = {0, 0, 0}, off_c[kBuckets] = {0, 0, 0};
    int blanes_a[kBuckets] = {8, 8, 8}, blanes_at[kBuckets] = {8, 8, 8};
    int bmult_a[kBuckets] = {1, 1, 1}, bmult_at[kBuckets] = {1, 1, 1};
    int chunks_per_graph = 4;
    bool use_graph = true;
    pm::DevBuf<char> state;
    pm::DevBuf<int> counter;
    pm::PinnedBuf<int> h_counter;
    double clock_khz = 100000.0;
};
