// Internal declarations shared by the HIP kernels and the C-ABI layer.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/vjx.h"
#include "symbolic.h"

namespace vjx {

void set_error(const std::string& msg);

#define VJX_TRY_HIP(call)                                                       \
  do {                                                                          \
    hipError_t e_ = (call);                                                     \
    if (e_ != hipSuccess) {                                                     \
      ::vjx::set_error(std::string(#call) + " failed: " + hipGetErrorString(e_)); \
      return VJX_ERR_HIP;                                                       \
    }                                                                           \
  } while (0)

template <typename T>
struct DBuf {
  T* p = nullptr;
  size_t n = 0;
  DBuf() = default;
  DBuf(const DBuf&) = delete;
  DBuf& operator=(const DBuf&) = delete;
  ~DBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  hipError_t alloc(size_t count) {
    release();
    n = count;
    if (count == 0) return hipSuccess;
    return hipMalloc((void**)&p, count * sizeof(T));
  }
  hipError_t upload(const T* host, size_t count) {
    hipError_t e = alloc(count);
    if (e != hipSuccess || count == 0) return e;
    return hipMemcpy(p, host, count * sizeof(T), hipMemcpyHostToDevice);
  }
  hipError_t upload(const std::vector<T>& v) { return upload(v.data(), v.size()); }
};

// Per-NR-iteration scalars read by kernels from device memory, so captured
// graphs replay with current values.
struct DevScalars {
  double c0, c1, c2, d1;
  // gmin: diagonal regularisation of node rows (the options' gmin,
  // mna_builder.py:384, 969); gmin_dev: the device simparam $gmin that
  // homotopy gdev stepping varies (mna_builder.py:486, 649-651) and that
  // only models reading $simparam("gmin") see -- none of the models here
  double gmin, gshunt, gmin_dev;
  double reltol, vntol, abstol, nr_damping, max_step;
  int32_t iteration;
  int32_t pad;
};

// Flags and reductions written by kernels (atomics on integer bit patterns of
// non-negative doubles: order-independent, hence deterministic).
struct DevStatus {
  unsigned long long maxf_bits;
  unsigned long long maxdv_bits;
  unsigned long long maxdelta_bits;
  unsigned long long maxl_bits;
  unsigned long long limd_bits;  // max |limit_state_out - limit_state_in|
  unsigned long long limr_bits;  // max |limit_state_in|
  unsigned int res_fail;
  unsigned int delta_fail;
  unsigned int nonfinite;     // non-finite LU factor entries (pivot monitor)
  unsigned int nonfinite_in;  // non-finite entries of the matrix being factorised
};

struct DevModel {
  int32_t model_id = 0;
  int32_t count = 0;
  int32_t n_v = 0, n_jac = 0, n_res = 0, n_par = 0;
  int32_t n_lim = 0;    // limit states per instance
  int64_t lim_off = 0;  // first limit state of the group in the flat array
  int32_t block = 256;  // eval workgroup size (launch plan: 64, 128 or 256)
  DBuf<int32_t> vn1, vn2, jac_dst, res_dst;
  DBuf<double> par;
};

// Level program on the device: item ids per level plus host-side launch plan.
struct DevSchedule {
  DBuf<int32_t> items;
  DBuf<int32_t> lvl_dev;           // device copy of lvl_ptr
  DBuf<int32_t> heavy_dev;         // device copy of lvl_heavy
  DBuf<int32_t> mid_dev;           // device copy of lvl_mid
  std::vector<int32_t> lvl_ptr;    // host copies
  std::vector<int32_t> lvl_heavy;
  std::vector<int32_t> lvl_mid;
  std::vector<LaunchGroup> groups;
};

// Band program of a triangular solve on the device (host plan in sym).
struct DevBand {
  DBuf<int32_t> items, tl_ptr, tl_heavy, tl_mid, hub_rows, hub_cptr, hub_cs, hub_ce;
  DBuf<double> part;  // hub chunk partials (FP32 solves use it as float storage)
  DBuf<int32_t> lds_rec, lds_task, lds_slots;  // LDS form (solves)
  DBuf<uint16_t> lds_col;
  DBuf<int32_t> lds_src;  // entries-in-LDS form: LU slot of each lds_col entry
  bool on = false;
  bool use_lds = true;
  int lde_bytes = 0;  // dynamic LDS a solve band task may use
};

struct DevLU {
  LuSymbolic sym;
  bool analyzed = false;
  int factored_prec = -1;  // precision of the current device factors
  DBuf<int32_t> rowptr, col, diag, src, tgt_pos, tgt_seg, upd_l, upd_u, upd_d, l_slot, l_diag, p, q;
  DBuf<int32_t> task_scale, task_scale_ptr;  // band programs: L scaled at each task's end
  DBuf<int32_t> upd_lu;                      // band programs: {upd_l, upd_u} interleaved
  DBuf<uint32_t> lds_scale;                  // ... by local LDS index in LDS tasks
  DevSchedule fact, fwd, bwd;
  DevBand fwd_band, bwd_band, fact_band;
  DBuf<int32_t> lds_item, lds_task, lds_slots;  // LDS form of the LU band tasks
  DBuf<uint32_t> lds_upd;
  bool use_lds = true;  // run LU band tasks that fit from LDS
  int n_cu = 256;       // compute units of the device
  int lu_deep = 0;      // LU band pipeline: 0 by task count, 1 shallow, 2 deep
  DBuf<double> lu64;
  DBuf<float> lu32;
  DBuf<double> y64, z64, r64, x64;
  DBuf<float> y32, z32;
  // dense trailing block (sym.dense_k > 0): slot maps and FP64 work arrays
  DBuf<int32_t> dense_slot, dense_pos, dl_slot, dl_pos, du_slot, du_pos;
  DBuf<double> dense_D, dense_Lp, dense_Up;
  int narrow_block = 512;
  int wide_block = 256;
};

// Segments longer than kHeavySegment ("hubs": the ground slot, supply rails)
// are summed in kHubChunk-element chunks by one workgroup each; the consuming
// kernel then adds a hub's chunk partials in chunk order, so the result is
// fixed by the plan and not by scheduling.

struct HubChunks {
  const int32_t* item = nullptr;  // hub item (slot or row) ids
  const int32_t* ptr = nullptr;   // hub h owns chunks [ptr[h], ptr[h + 1])
  const int32_t* cs = nullptr;    // chunk c covers [cs[c], ce[c])
  const int32_t* ce = nullptr;
  double* part = nullptr;  // per-chunk partials (sum)
  double* part2 = nullptr;  // second sum (residual rows: charge)
  double* part3 = nullptr;  // max |.| (residual rows)
  double* part4 = nullptr;  // lim_rhs sums (residual rows of limited circuits)
  double* part5 = nullptr;
  int n_items = 0;
  int n_chunks = 0;
};

struct HubPlan {
  DBuf<int32_t> item, ptr, cs, ce;
  DBuf<double> part, part2, part3, part4, part5;
  int n_items = 0, n_chunks = 0;
  // seg: host segment pointers of n segments; nparts: 1, 3 or 5 partial arrays.
  hipError_t build(const int32_t* seg, int n, int nparts);
  HubChunks view() const {
    HubChunks h;
    h.item = item.p;
    h.ptr = ptr.p;
    h.cs = cs.p;
    h.ce = ce.p;
    h.part = part.p;
    h.part2 = part2.p;
    h.part3 = part3.p;
    h.part4 = part4.p;
    h.part5 = part5.p;
    h.n_items = n_items;
    h.n_chunks = n_chunks;
    return h;
  }
};

// ----- per-kernel timing (profiling passes only, never under graph capture) --
enum KernelClass {
  KC_EVAL = 0,     // k_eval<model>
  KC_STAMP,        // chunk pre-passes + k_reduce_jac + k_residual
  KC_LU_GATHER,    // k_lu_gather
  KC_LU_TASKS,     // k_lu_band (and level launches of the level program)
  KC_LU_SCHUR,     // k_lu_wide of a band's Schur step
  KC_LU_SCALE,     // k_lu_scale
  KC_TRI_TASKS,    // k_tri_band (and level launches of the level program)
  KC_TRI_HUB,      // chunked hub rows of the solves
  KC_LU_DENSE,     // dense trailing block: gathers, MFMA products, panels, write-back
  KC_COUNT
};
void ktimer_enable(bool on);
bool ktimer_on();
void ktimer_begin(hipStream_t s);
void ktimer_end(int cls, hipStream_t s);
void ktimer_collect();  // after a stream synchronisation
void ktimer_reset();
void ktimer_read(double* ms, int64_t* launches);  // KC_COUNT entries each

// ----- kernel launchers (kernels.hip) -----
struct EvalOut {
  double* jac;  // contribution arrays (sorted slots), may be null
  double* rf;
  double* rq;
  double* lr = nullptr;  // lim_rhs contributions (resistive, reactive) of limited models
  double* lq = nullptr;
  const double* ls_in = nullptr;  // flat limit state in / out (limited models)
  double* ls_out = nullptr;
};

void launch_eval(const DevModel& m, const double* X, const DevScalars* sc, EvalOut out,
                 hipStream_t s);
void launch_isrc_contrib(int k, const double* isrc, const int32_t* is_dst, double* rf, double* rq,
                         hipStream_t s);
void launch_reduce_jac(int nse, const int32_t* seg, const double* contrib, const int32_t* slot_diag,
                       const double* slot_b, const DevScalars* sc, double* csr,
                       const HubChunks& hub, hipStream_t s);
struct ResidualArgs {
  int n_unk, m, n_total;
  const int32_t* res_seg;
  const double* rf;
  const double* rq;
  const double* lr;  // lim_rhs contributions, or null when no model limits
  const double* lq;
  const int32_t* br_ptr;
  const int32_t* br_ent;
  const int32_t* vs_p;
  const int32_t* vs_n;
  const int32_t* conv_mask;
  const double* X;
  const double* vsrc;
  const double* Q_prev;
  const double* dQdt_prev;
  const double* Q_prev2;
  const double* res_floor;  // may be null -> abstol
  const DevScalars* sc;
  double* f;
  double* Q;
  double* mrc;
  double* I_vs;
  DevStatus* st;  // may be null
  HubChunks hub;  // node rows with long residual segments
};
void launch_residual(const ResidualArgs& a, hipStream_t s);

template <typename T>
void launch_lu_refactor(DevLU& lu, const double* csr, T* lu_vals, DevStatus* st, hipStream_t s);
// the dense trailing block of a refactorisation (dense.hip)
template <typename T>
void launch_dense_block(DevLU& L, T* lu, DevStatus* st, hipStream_t s);
void launch_dense_lu(int K, double* D, hipStream_t s);
void launch_dgemm_sub(int M, int N, int Kd, const double* A, int lda, const double* B, int ldb,
                      double* C, int ldc, hipStream_t s);
// hub_parts (may be null): the forward and backward solves' hub chunk
// partials, in place of the band programs' own buffers
template <typename T>
void launch_lu_solve(DevLU& lu, const T* lu_vals, const double* rhs, double rhs_scale, double* x,
                     T* y, T* z, hipStream_t s, void* const* hub_parts = nullptr);
void launch_axpy(int n, double* x, const double* d, hipStream_t s);

void launch_model_eval_raw(int model_id, int count, const double* vb, const double* par,
                           const double* li, double* rr, double* rc, double* jr, double* jc,
                           double* lr, double* lc, double* lo, hipStream_t s);
void launch_split_gather(int N, int n_v, int n_par, int n_dp, int n_dc, const int32_t* vpos,
                         const int32_t* csrc, const double* sc, const double* dp, const double* dc,
                         double* vb, double* par, hipStream_t s);
void launch_split_finish(int N, int n_res, int n_jac, int limited, const double* mf, double* rr,
                         double* rc, double* jr, double* jc, double* lr, double* lc, double* sr,
                         double* sq, hipStream_t s);
void launch_limit(int kind, int n, const double* vnew, const double* vold, double p1, double p2,
                  double* out, hipStream_t s);
void launch_lim_stats(int n, const double* ls_in, const double* ls_out, DevStatus* st,
                      hipStream_t s);
// Independent sources of one kind on the device: waveform codes, the
// [VJX_SRC_NPAR][n] parameter block and the pwl points (may be null).
struct SourceTable {
  int n;
  const int32_t* wave;
  const double* par;
  const int32_t* pwl_ptr;
  const double* pwl_t;
  const double* pwl_v;
};
void launch_sources(const SourceTable& src, double t, double* out, hipStream_t s);
// both tables at t in one launch (either may be empty)
void launch_sources2(const SourceTable& a, double* out_a, const SourceTable& b, double* out_b,
                     double t, hipStream_t s);

int model_n_v(int model_id);
int model_n_lim(int model_id);
int model_n_res(int model_id);
int model_n_jac(int model_id);
int model_n_par(int model_id);

}  // namespace vjx
