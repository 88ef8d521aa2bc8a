// Host-side symbolic analysis for the sparse LU of the MNA Jacobian.
//
// Replaces the symbolic half of the reference's UMFPACK FFI
// (vajax/sparse/umfpack_ffi.cpp:140-179, 213-267): ordering and the pivot
// sequence are chosen once per topology from the first numeric matrix; every
// later Newton iteration only runs the numeric refactorisation on the GPU.
//
//  * column order q on the pattern of A + A^T, rows whose degree exceeds
//    max(16, 10 sqrt(n)) (supply rails) ordered last: nested dissection with
//    BFS level-structure separators (shallow elimination tree, hence few
//    levels) and minimum degree inside the leaves, or plain minimum degree;
//  * row pivots p: left-looking Gilbert-Peierls LU of A(:, q) with threshold
//    partial pivoting (tolerance 0.1, the UMFPACK setting at
//    umfpack_ffi.cpp:153), preferring the structural diagonal;
//  * LU pattern of P A Q in CSR with the fill, and for every LU slot the A
//    entry it is initialised from (or -1 for fill);
//  * a right-looking level schedule: pivot k may run once every pivot k' with
//    U(k',k) != 0 or L(k,k') != 0 has run.  Each level is a list of update
//    targets (i,j) with the fixed-order list of contributions
//    L(i,k)*U(k,j)/U(k,k) from that level's pivots, so the numeric phase is
//    deterministic and needs no atomics;
//  * row level sets for the forward (L) and backward (U) triangular solves.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace vjx {

// One launch of a level program: either a single wide level (grid of blocks)
// or a run of consecutive narrow levels done by one block with barriers.
struct LaunchGroup {
  int lvl_begin;
  int lvl_end;  // exclusive
  int wide;     // 1: one level over a grid; 0: levels looped in one block
  int items;    // number of items over the group
};

// Items of one level are stored by reduction class: first lvl_heavy[l] items
// with long segments (one workgroup each), then lvl_mid[l] items with medium
// segments (kGroupLanes lanes each), then short items (one lane each).  Every
// class sums in a fixed order (strided lane partials, then a halving tree),
// which the CPU program reproduces exactly.
struct LevelSchedule {
  std::vector<int32_t> items;      // item ids grouped by level
  std::vector<int32_t> lvl_ptr;    // level l covers items[lvl_ptr[l] .. lvl_ptr[l+1])
  std::vector<int32_t> lvl_heavy;  // workgroup items at the start of each level
  std::vector<int32_t> lvl_mid;    // lane-group items after them
  std::vector<LaunchGroup> groups;
  int max_width = 0;
  int max_group_levels = 0;
};

constexpr int kHeavySegment = 128;    // longer segments: one workgroup
constexpr int kShortSegment = 3;      // shorter or equal: one lane
constexpr int kGroupLanes = 16;       // lanes per medium segment
constexpr int kReduceLanes = 256;     // lanes of the fixed-order workgroup sum
// Long rows (kHeavySegment < length <= kHubChunk) of the triangular-solve
// band programs are summed by one wavefront (strided partials + shuffle
// tree), one row per wavefront, so a level needs no barrier beyond its own.
constexpr int kWaveLanes = 64;
constexpr int kMaxNarrowLevels = 4096;  // levels per single-block launch (LDS cache)

inline int reduction_class(int len) {
  return len > kHeavySegment ? 2 : (len > kShortSegment ? 1 : 0);
}

// Segments longer than kHubChunk (supply rails, dense LU rows) are summed in
// kHubChunk-entry chunks, one workgroup per chunk (fixed-order workgroup sum),
// and the chunk partials are then added in chunk order.
constexpr int kHubChunk = 4096;
constexpr int kMaxBandWidth = 64;  // levels per band (LDS table of one task)
// LU values a band task may hold in LDS: the 160 KiB per workgroup of gfx950
// less 4 KiB for the kernel's own reduction and level tables, in FP64 (19 968
// values; 16-bit local indices).  A band launch asks for what its largest
// task needs, so bands of small tasks keep their small footprint; a band with
// a task beyond this runs from HBM.
constexpr int kLdsSlots = (160 * 1024 - 4096) / 8;
// Static LDS of the entries-in-LDS solve kernel (reduction tree and level
// tables), kept out of the dynamic budget with margin.
constexpr int kTriStaticLds = 4096;
// LDS a solve band task may use for its values, entries and records: the
// 160 KiB per workgroup of gfx950 less the kernel's static arrays
constexpr int kSolveLdsBytes = 160 * 1024 - kTriStaticLds;

// Band program of a triangular solve.  The levels are cut into bands of at
// most `width` levels; inside a band the rows form connected components
// (a row depends only on rows at lower levels, and on rows of the same band
// only through edges inside its component), which are packed into tasks.
// One launch per band runs one workgroup per task; the workgroup walks the
// band's levels with a barrier between them.  Rows longer than kHubChunk are
// taken out of the tasks and finished after them by a chunked grid
// reduction (a band that holds one ends at that row's level).
struct BandSchedule {
  std::vector<int32_t> items;                // rows by band, task, level, class
  std::vector<int32_t> band_task;            // band b: tasks [band_task[b], band_task[b+1])
  std::vector<int32_t> band_w;               // levels in band b
  std::vector<int32_t> band_ptr_off;         // task t of band b: tl_ptr[band_ptr_off[b] + (t-t0)*(w+1) ...]
  std::vector<int32_t> band_hm_off;          //                  tl_heavy/mid[band_hm_off[b] + (t-t0)*w ...]
  std::vector<int32_t> tl_ptr;               // per task: w+1 item offsets
  std::vector<int32_t> tl_heavy, tl_mid;     // per task and level: workgroup / lane-group rows
  std::vector<int32_t> band_hub;             // band b: hub rows [band_hub[b], band_hub[b+1])
  std::vector<int32_t> hub_rows;
  std::vector<int32_t> hub_cptr;             // hub h: chunks [hub_cptr[h], hub_cptr[h+1])
  std::vector<int32_t> hub_cs, hub_ce;       // chunk c covers LU entries [hub_cs[c], hub_ce[c])
  int max_tasks = 0;
  // LDS form of a solve band (triangular solves only): a task keeps the
  // vector values of its own rows and of the rows they read in LDS.
  // Per item: {local row, first entry, end entry, first local column};
  // lds_col holds the local index of every entry's column.
  std::vector<int32_t> lds_rec;       // 4 ints per item
  std::vector<uint16_t> lds_col;
  std::vector<int32_t> lds_task;      // 4 ints per task: slot offset, count, own rows, 0
  std::vector<int32_t> lds_slots;     // per task: rows, own rows first
  std::vector<int32_t> band_lds;      // per band: LDS values needed (0: band runs from HBM)
  // Entries-in-LDS form: lds_src[k] is the LU slot of the entry whose local
  // column is lds_col[k]; per band the largest task's entry and row counts.
  std::vector<int32_t> lds_src;
  std::vector<int32_t> band_ent, band_rows;
};

struct LuSymbolic {
  int n = 0;
  int nnz_a = 0;
  std::vector<int32_t> q, qinv;  // column order: permuted col k = original col q[k]
  std::vector<int32_t> p, pinv;  // row order:    permuted row k = original row p[k]
  // LU pattern of P A Q, CSR, columns sorted; L strictly left of diag, U right.
  std::vector<int32_t> rowptr, col, diag;
  std::vector<int32_t> src;  // per LU slot: index into A's CSR data, -1 = fill
  // right-looking update program
  std::vector<int32_t> tgt_pos;  // per target: LU slot updated
  std::vector<int32_t> tgt_seg;  // per target: contributions [tgt_seg[t], tgt_seg[t+1])
  std::vector<int32_t> upd_l;    // per contribution: slot of L(i,k)
  std::vector<int32_t> upd_u;    // per contribution: slot of U(k,j)
  std::vector<int32_t> upd_d;    // per contribution: slot of U(k,k)
  LevelSchedule fact;            // items = target ids (level program)
  // Band program of the refactorisation (used when fact_band has bands):
  // per band, tasks of connected pivots update the targets in their own rows
  // and columns level by level (items = target ids), then the band's updates
  // of later rows and columns run as one wide step, items
  // fact_band.items[fact_schur_lo[b] .. fact_schur_hi[b]) ordered by class.
  BandSchedule fact_band;
  std::vector<int32_t> fact_schur_lo, fact_schur_hi, fact_schur_heavy, fact_schur_mid;
  // LDS form of the band tasks: a task whose slots (its targets and every
  // operand its updates read) fit in kLdsSlots values loads them into LDS,
  // runs its levels there and writes its targets back.  Per task item
  // (parallel to fact_band.items over task items; unused for Schur items):
  // {target id, first packed update, update count, local target index}, or
  // for an item with exactly one update {its packed word, its pivot index,
  // 1, local target index} with the update itself not in lds_upd.
  std::vector<int32_t> lds_item;      // 4 ints per fact_band item
  std::vector<uint32_t> lds_upd;      // 2 words per update: (l | u << 16), d
  std::vector<int32_t> lds_task;      // 4 ints per task: slot offset, slot count, written count, 0
  std::vector<int32_t> lds_slots;     // per task: global slots, written ones first
  std::vector<int32_t> band_lds;      // per band: LDS values needed (0: band runs from HBM)
  std::vector<int32_t> l_slot;   // every L slot (i,k) ...
  std::vector<int32_t> l_diag;   // ... and the slot of its pivot U(k,k)
  // Under band programs L is scaled inside the tasks: after its last level a
  // task divides the L entries of its own pivots' columns by their pivots
  // ({L slot, pivot slot} pairs, tasks in band-program order), so the Schur
  // steps read scaled L (a product per update, no division, no pivot index)
  // and no separate scaling pass runs.  Every L entry belongs to the column
  // of exactly one pivot, hence to exactly one task.
  std::vector<int32_t> task_scale;      // 2 ints per entry
  std::vector<int32_t> task_scale_ptr;  // per task: entries [ptr[t], ptr[t + 1])
  // the same pairs as local LDS indices (L | pivot << 16) for tasks of bands
  // that run from LDS, whose slot sets include them (L entries written back)
  std::vector<uint32_t> lds_scale;
  LevelSchedule fwd;             // items = rows, L solve
  LevelSchedule bwd;             // items = rows, U solve
  BandSchedule fwd_band;         // band programs of the two solves (empty: level launches)
  BandSchedule bwd_band;
  // Dense trailing block: pivots dense_start .. n-1 (dense_k of them) are
  // factorised as a dense row-major dense_k x dense_k matrix D after the
  // scalar program.  D is assembled from the slots (i, j), i, j >=
  // dense_start (dense_slot -> dense_pos) as the scalar program leaves them:
  // A's values less the updates of earlier pivots whose contribution to the
  // block is small; the dense_m earlier pivots with large contributions
  // (covering at least 1/32 of the block) are left out of the scalar program
  // there and subtracted as one product D -= Lp Up, Lp [dense_k x dense_m]
  // from their L slots (dl_slot -> dl_pos), Up [dense_m x dense_k] from their
  // U slots (du_slot -> du_pos); then D = L U in place (unit L, no pivoting:
  // the pivots are the analysis's) and the slots are written back.
  // dense_start = -1: no dense block.
  int dense_start = -1, dense_k = 0, dense_m = 0;
  std::vector<int32_t> dense_slot, dense_pos, dl_slot, dl_pos, du_slot, du_pos;
  double dense_flops = 0.0;  // 2 K^2 m (contributions) + 2 K^3 / 3 (dense LU)
  // statistics
  int64_t nnz_l = 0, nnz_u = 0, n_updates = 0;
  double flops = 0.0;
  int n_dense = 0;
  int n_fact_levels = 0;     // pivot levels of the refactorisation
  int nd_leaf = 0;           // nested-dissection leaf in use (0: other orderings)
  // analysis time (ms): ordering, pivoting, LU pattern, refactorisation
  // program, solve programs (single-block systems)
  double analyze_ms[5] = {0, 0, 0, 0, 0};
  int n_blocks = 1;          // diagonal blocks (connected components)
  int n_block_patterns = 1;  // distinct block patterns analysed
  int n_offdiag_pivots = 0;
  double min_pivot_ratio = 1.0;
};

struct AnalyzeOptions {
  double pivot_tol = 0.1;
  int wide_threshold = 2048;  // levels with at least this many items get a grid
  int ordering = 3;           // 0 = minimum degree, 1 = natural, 2 = nested
                              // dissection (BFS level-structure separators),
                              // 3 = nested dissection (multilevel separators)
  int nd_leaf = 0;            // nested-dissection leaf size (min-degree inside);
                              // 0: 256 (ordering 3) or max(48, n / 128) of the
                              // block being ordered (ordering 2)
  int band_width = 16;        // bands of this many levels (0: one launch per level)
  // dense trailing block: the largest trailing block of the LU pattern (at
  // least dense_min rows) whose density reaches dense_density is factorised
  // as a dense matrix (0: never); band programs of single-block systems only
  int dense_min = 96;
  double dense_density = 0.4;
  int solve_band_width = 0;   // solve bands (0: band_width)
  int max_band_tasks = 2048;  // workgroups per band launch
  // abandon the analysis once L + U hold more entries than this (0: no
  // limit); the error text then names the fill limit
  int64_t max_lu_nnz = 0;
};

// Rebuild the level sets and band programs of the two solves (e.g. for a
// new opt.solve_band_width); the factors' program is untouched.
void lu_rebuild_solve_programs(LuSymbolic& s, const AnalyzeOptions& opt);

// Full analysis from a CSR matrix (host arrays).  Returns false and sets err on
// malformed input.  Values are used only to choose pivots.
bool lu_analyze(int n, const int32_t* indptr, const int32_t* indices, const double* vals,
                const AnalyzeOptions& opt, LuSymbolic& out, std::string& err);

// CSR structure validation as in vajax/sparse/umfpack_ffi.cpp:234-251.
bool csr_validate(int n, int nnz, const int32_t* indptr, const int32_t* indices, std::string& err);

// Dependency check of the band programs (every row once, after what it reads).
bool lu_check_schedules(const LuSymbolic& s, std::string& err);

// Columns per panel of the blocked dense LU of the dense trailing block.
constexpr int kDensePanel = 32;
// Dense LU without pivoting of a row-major K x K matrix in place (unit L
// below the diagonal, U on and above), in panels of kDensePanel columns.
void dense_lu_host(int K, double* D);

// CPU execution of the same numeric program (reference for the GPU kernels).
template <typename T>
void lu_refactor_host(const LuSymbolic& s, const double* a, T* lu);
template <typename T>
void lu_solve_host(const LuSymbolic& s, const T* lu, const double* b, double* x);

}  // namespace vjx
