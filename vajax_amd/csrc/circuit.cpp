// Circuit handle: device copies of the topology, and the system build
// (device eval + stamping + residual) behind vjx_build_system.
#include <cstring>
#include <vector>

#include "circuit.h"
#include "transient.h"

vjx_circuit::~vjx_circuit() {
  if (h_sc) (void)hipHostFree(h_sc);
  if (pat) vjx_pattern_destroy(pat);
}

namespace vjx {

void circuit_build_async(vjx_circuit* c, const double* X, const double* vsrc, const double* isrc,
                         const double* Q_prev, const double* dQdt_prev, const double* Q_prev2,
                         const double* res_floor, double* csr, double* f, double* Q, double* mrc,
                         double* I_vs, DevStatus* st, const double* ls_in, double* ls_out,
                         hipStream_t s, int only_model) {
  EvalOut out{csr ? c->jac_contrib.p : nullptr, c->rf.p, c->rq.p};
  if (c->n_lim > 0) {
    out.lr = c->lr.p;
    out.lq = c->lq.p;
    out.ls_in = ls_in;
    out.ls_out = ls_out ? ls_out : c->ls_scratch.p;
  }
  if (only_model >= 0) {
    // one group alone: every contribution slot starts from zero
    if (csr) launch_fill(c->n_jac_contrib, 0.0, c->jac_contrib.p, s);
    launch_fill(c->n_res_contrib, 0.0, c->rf.p, s);
    launch_fill(c->n_res_contrib, 0.0, c->rq.p, s);
    if (c->n_lim > 0) {
      launch_fill(c->n_res_contrib, 0.0, c->lr.p, s);
      launch_fill(c->n_res_contrib, 0.0, c->lq.p, s);
    }
  }
  for (size_t g = 0; g < c->models.size(); ++g)
    if (only_model < 0 || (size_t)only_model == g) launch_eval(*c->models[g], X, c->d_sc.p, out, s);
  if (only_model < 0) launch_isrc_contrib(c->k, isrc, c->is_dst.p, c->rf.p, c->rq.p, s);
  if (csr)
    launch_reduce_jac(c->nse, c->jac_seg.p, c->jac_contrib.p, c->slot_diag.p, c->slot_b.p,
                      c->d_sc.p, csr, c->hub_slots.view(), s);
  ResidualArgs a;
  a.n_unk = c->n_unk;
  a.m = c->m;
  a.n_total = c->n_total;
  a.res_seg = c->res_seg.p;
  a.rf = c->rf.p;
  a.rq = c->rq.p;
  a.lr = c->n_lim > 0 ? c->lr.p : nullptr;
  a.lq = c->n_lim > 0 ? c->lq.p : nullptr;
  a.br_ptr = c->br_ptr.p;
  a.br_ent = c->br_ent.p;
  a.vs_p = c->vs_p.p;
  a.vs_n = c->vs_n.p;
  a.conv_mask = c->has_conv_mask ? c->conv_mask.p : nullptr;
  a.X = X;
  a.vsrc = vsrc;
  a.Q_prev = Q_prev;
  a.dQdt_prev = dQdt_prev;
  a.Q_prev2 = Q_prev2;
  a.res_floor = res_floor;
  a.sc = c->d_sc.p;
  a.f = f;
  a.Q = Q;
  a.mrc = mrc;
  a.I_vs = I_vs;
  a.st = st;
  a.hub = c->hub_rows.view();
  launch_residual(a, s);
}

int circuit_autotune_eval(vjx_circuit* c, const double* X, hipStream_t s) {
  EvalOut out{c->jac_contrib.p, c->rf.p, c->rq.p};
  hipEvent_t ev[2];
  VJX_TRY_HIP(hipEventCreate(&ev[0]));
  VJX_TRY_HIP(hipEventCreate(&ev[1]));
  int rc = VJX_OK;
  for (auto& mp : c->models) {
    DevModel& m = *mp;
    if (m.count < 8192 || m.n_lim > 0) continue;
    const int sizes[3] = {256, 128, 64};
    float t[3] = {1e30f, 1e30f, 1e30f};  // best of the rounds per size
    for (int round = 0; round < 3 && rc == VJX_OK; ++round) {
      for (int k = 0; k < 3; ++k) {
        m.block = sizes[k];
        launch_eval(m, X, c->d_sc.p, out, s);  // warm
        hipError_t e = hipEventRecord(ev[0], s);
        for (int r = 0; r < 5 && e == hipSuccess; ++r) launch_eval(m, X, c->d_sc.p, out, s);
        if (e == hipSuccess) e = hipEventRecord(ev[1], s);
        if (e == hipSuccess) e = hipEventSynchronize(ev[1]);
        float ms = 0.f;
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, ev[0], ev[1]);
        if (e != hipSuccess) {
          set_error(std::string("circuit_autotune_eval: ") + hipGetErrorString(e));
          rc = VJX_ERR_HIP;
          break;
        }
        t[k] = std::min(t[k], ms / 5.f);
      }
    }
    // the incumbent (256) stays unless another size is 10 % faster
    int best = 0;
    for (int k = 1; k < 3; ++k)
      if (t[k] < 0.9f * t[0] && t[k] < t[best]) best = k;
    m.block = sizes[best];
  }
  (void)hipEventDestroy(ev[0]);
  (void)hipEventDestroy(ev[1]);
  return rc;
}


}  // namespace vjx

using namespace vjx;

extern "C" {

namespace {

// A segment pointer array seg[0..n] that the kernels follow: starts at 0,
// never decreases, ends at total.
bool valid_segments(const int32_t* seg, int n, int64_t total) {
  if (!seg || seg[0] != 0 || (int64_t)seg[n] != total) return false;
  for (int i = 0; i < n; ++i)
    if (seg[i + 1] < seg[i]) return false;
  return true;
}

// Every pointer, size and index of the descriptor that a kernel will follow,
// checked on the host before anything is uploaded.
int validate_desc(const vjx_circuit_desc* d) {
  auto fail = [](const char* what) {
    set_error(std::string("vjx_circuit_create: ") + what);
    return VJX_ERR_ARG;
  };
  if (d->n_total < 1 || d->n_vsources < 0 || d->n_isources < 0 || d->nse < 0 ||
      d->n_models < 0 || d->n_models > VJX_MAX_MODELS || d->n_jac_contrib < 0 ||
      d->n_res_contrib < 0 || d->n_noi < 0)
    return fail("invalid sizes");
  const int nu = d->n_total - 1, m = d->n_vsources, k = d->n_isources;
  if (!d->indptr || (d->nse && !d->indices) || (d->nse && (!d->slot_diag || !d->slot_b)))
    return fail("null CSR pattern or slot array");
  if (!valid_segments(d->jac_seg, d->nse, d->n_jac_contrib))
    return fail("jac_seg must start at 0, not decrease and end at n_jac_contrib");
  if (!valid_segments(d->res_seg, nu, d->n_res_contrib))
    return fail("res_seg must start at 0, not decrease and end at n_res_contrib");
  if (!d->br_ptr || d->br_ptr[0] != 0) return fail("br_ptr must start at 0");
  for (int i = 0; i < nu; ++i)
    if (d->br_ptr[i + 1] < d->br_ptr[i]) return fail("br_ptr decreases");
  if (d->br_ptr[nu] > 2 * m) return fail("br_ptr[n_unknowns] exceeds two entries per source");
  if (d->br_ptr[nu] > 0 && !d->br_ent) return fail("null br_ent");
  for (int e = 0; e < d->br_ptr[nu]; ++e) {
    const int kk = d->br_ent[e] >= 0 ? d->br_ent[e] : ~d->br_ent[e];
    if (kk >= m) return fail("branch entry out of range");
  }
  if (m > 0 && (!d->vs_node_p || !d->vs_node_n || !d->vs_wave || !d->vs_params))
    return fail("null voltage-source arrays");
  if (k > 0 && (!d->is_wave || !d->is_params || !d->is_dst))
    return fail("null current-source arrays");
  for (int i = 0; i < m; ++i)
    if (d->vs_node_p[i] < 0 || d->vs_node_p[i] >= d->n_total || d->vs_node_n[i] < 0 ||
        d->vs_node_n[i] >= d->n_total)
      return fail("voltage-source node out of range");
  // waveform codes; a pwl source needs its points: pointers from 0, not
  // decreasing, at least one point per pwl source, times not decreasing
  auto check_sources = [&](int n, const int32_t* wave, const int32_t* ptr, const double* tt,
                           const double* vv) -> int {
    bool any_pwl = false;
    for (int i = 0; i < n; ++i) {
      if (wave[i] < 0 || wave[i] >= VJX_WAVE_COUNT) return fail("unknown waveform");
      any_pwl = any_pwl || wave[i] == VJX_WAVE_PWL;
    }
    if (!any_pwl) return VJX_OK;
    if (!valid_segments(ptr, n, ptr ? ptr[n] : 0) || (ptr[n] > 0 && (!tt || !vv)))
      return fail("pwl point arrays missing or malformed");
    for (int i = 0; i < n; ++i) {
      if (wave[i] != VJX_WAVE_PWL) continue;
      if (ptr[i + 1] <= ptr[i]) return fail("pwl source without points");
      for (int e = ptr[i] + 1; e < ptr[i + 1]; ++e)
        if (!(tt[e] >= tt[e - 1])) return fail("pwl times decrease");
    }
    return VJX_OK;
  };
  if (m > 0) {
    int r = check_sources(m, d->vs_wave, d->vs_pwl_ptr, d->vs_pwl_t, d->vs_pwl_v);
    if (r < 0) return r;
  }
  if (k > 0) {
    int r = check_sources(k, d->is_wave, d->is_pwl_ptr, d->is_pwl_t, d->is_pwl_v);
    if (r < 0) return r;
  }
  for (int e = 0; e < 2 * k; ++e)
    if (d->is_dst[e] < -1 || d->is_dst[e] >= d->n_res_contrib)
      return fail("current-source slot out of range");
  for (int i = 0; i < d->n_models; ++i) {
    const vjx_model_desc& md = d->models[i];
    if (md.model_id < 0 || md.model_id >= VJX_MODEL_COUNT || md.count < 0)
      return fail("unknown model id");
    if (md.count == 0) continue;
    if (!md.v_node1 || !md.v_node2 || !md.jac_dst || !md.res_dst || !md.params)
      return fail("null model arrays");
    const size_t N = (size_t)md.count;
    const int n_v = model_n_v(md.model_id), n_jac = model_n_jac(md.model_id),
              n_res = model_n_res(md.model_id);
    for (size_t e = 0; e < N * n_v; ++e)
      if (md.v_node1[e] < 0 || md.v_node1[e] >= d->n_total || md.v_node2[e] < 0 ||
          md.v_node2[e] >= d->n_total)
        return fail("branch-voltage node index out of range");
    for (size_t e = 0; e < N * n_jac; ++e)
      if (md.jac_dst[e] < -1 || md.jac_dst[e] >= d->n_jac_contrib)
        return fail("Jacobian contribution slot out of range");
    for (size_t e = 0; e < N * n_res; ++e)
      if (md.res_dst[e] < -1 || md.res_dst[e] >= d->n_res_contrib)
        return fail("residual contribution slot out of range");
  }
  if (d->n_noi > 0 && !d->noi_rows) return fail("bad NOI list");
  for (int i = 0; i < d->n_noi; ++i)
    if (d->noi_rows[i] < 0 || d->noi_rows[i] >= nu) return fail("NOI row out of range");
  return VJX_OK;
}

}  // namespace

int vjx_circuit_create(const vjx_circuit_desc* d, vjx_circuit** out) {
  if (!d || !out) {
    set_error("vjx_circuit_create: null argument");
    return VJX_ERR_ARG;
  }
  int vrc = validate_desc(d);
  if (vrc < 0) return vrc;
  std::unique_ptr<vjx_circuit> c(new vjx_circuit());
  c->n_total = d->n_total;
  c->n_unk = d->n_total - 1;
  c->m = d->n_vsources;
  c->k = d->n_isources;
  c->n_aug = c->n_unk + c->m;
  c->nse = d->nse;
  int rc = vjx_pattern_create(c->n_aug, c->nse, d->indptr, d->indices, &c->pat);
  if (rc < 0) return rc;
  rc = pattern_upload_csr(c->pat);
  if (rc < 0) return rc;
  const int nu = c->n_unk, m = c->m, k = c->k;
  VJX_TRY_HIP(c->slot_diag.upload(d->slot_diag, c->nse));
  VJX_TRY_HIP(c->slot_b.upload(d->slot_b, c->nse));
  VJX_TRY_HIP(c->jac_seg.upload(d->jac_seg, c->nse + 1));
  VJX_TRY_HIP(c->res_seg.upload(d->res_seg, nu + 1));
  for (int i = 0; i < d->n_models; ++i)
    if (d->models[i].model_id >= 0 && d->models[i].model_id < VJX_MODEL_COUNT &&
        d->models[i].count > 0)
      c->n_lim += model_n_lim(d->models[i].model_id) * d->models[i].count;
  VJX_TRY_HIP(c->hub_slots.build(d->jac_seg, c->nse, 1));
  VJX_TRY_HIP(c->hub_rows.build(d->res_seg, nu, c->n_lim > 0 ? 5 : 3));
  VJX_TRY_HIP(c->br_ptr.upload(d->br_ptr, nu + 1));
  VJX_TRY_HIP(c->br_ent.upload(d->br_ent, d->br_ptr[nu]));
  VJX_TRY_HIP(c->vs_p.upload(d->vs_node_p, m));
  VJX_TRY_HIP(c->vs_n.upload(d->vs_node_n, m));
  VJX_TRY_HIP(c->vs_wave.upload(d->vs_wave, m));
  VJX_TRY_HIP(c->vs_par.upload(d->vs_params, (size_t)VJX_SRC_NPAR * m));
  VJX_TRY_HIP(c->is_wave.upload(d->is_wave, k));
  VJX_TRY_HIP(c->is_par.upload(d->is_params, (size_t)VJX_SRC_NPAR * k));
  VJX_TRY_HIP(c->is_dst.upload(d->is_dst, (size_t)2 * k));
  if (d->vs_pwl_ptr) {
    VJX_TRY_HIP(c->vs_pwl_ptr.upload(d->vs_pwl_ptr, m + 1));
    VJX_TRY_HIP(c->vs_pwl_t.upload(d->vs_pwl_t, d->vs_pwl_ptr[m]));
    VJX_TRY_HIP(c->vs_pwl_v.upload(d->vs_pwl_v, d->vs_pwl_ptr[m]));
  }
  if (d->is_pwl_ptr) {
    VJX_TRY_HIP(c->is_pwl_ptr.upload(d->is_pwl_ptr, k + 1));
    VJX_TRY_HIP(c->is_pwl_t.upload(d->is_pwl_t, d->is_pwl_ptr[k]));
    VJX_TRY_HIP(c->is_pwl_v.upload(d->is_pwl_v, d->is_pwl_ptr[k]));
  }
  if (d->conv_mask) {
    VJX_TRY_HIP(c->conv_mask.upload(d->conv_mask, nu));
    c->has_conv_mask = true;
  }
  c->n_jac_contrib = d->n_jac_contrib;
  c->n_res_contrib = d->n_res_contrib;
  VJX_TRY_HIP(c->jac_contrib.alloc(d->n_jac_contrib));
  VJX_TRY_HIP(c->rf.alloc(d->n_res_contrib));
  VJX_TRY_HIP(c->rq.alloc(d->n_res_contrib));
  if (c->n_lim > 0) {
    // slots of unlimited models keep their zero lim_rhs
    VJX_TRY_HIP(c->lr.alloc(d->n_res_contrib));
    VJX_TRY_HIP(c->lq.alloc(d->n_res_contrib));
    VJX_TRY_HIP(c->ls_scratch.alloc(c->n_lim));
    VJX_TRY_HIP(hipMemset(c->lr.p, 0, sizeof(double) * d->n_res_contrib));
    VJX_TRY_HIP(hipMemset(c->lq.p, 0, sizeof(double) * d->n_res_contrib));
  }
  int64_t lim_off = 0;
  for (int i = 0; i < d->n_models; ++i) {
    const vjx_model_desc& md = d->models[i];
    if (md.model_id < 0 || md.model_id >= VJX_MODEL_COUNT || md.count < 0) {
      set_error("vjx_circuit_create: unknown model id");
      return VJX_ERR_ARG;
    }
    auto dm = std::make_unique<DevModel>();
    dm->model_id = md.model_id;
    dm->count = md.count;
    dm->n_v = model_n_v(md.model_id);
    dm->n_jac = model_n_jac(md.model_id);
    dm->n_res = model_n_res(md.model_id);
    dm->n_par = model_n_par(md.model_id);
    dm->n_lim = model_n_lim(md.model_id);
    dm->lim_off = lim_off;
    lim_off += (int64_t)dm->n_lim * md.count;
    const size_t N = (size_t)md.count;
    // bounds checks on every index the kernels will follow
    for (size_t e = 0; e < N * dm->n_v; ++e)
      if (md.v_node1[e] < 0 || md.v_node1[e] >= c->n_total || md.v_node2[e] < 0 ||
          md.v_node2[e] >= c->n_total) {
        set_error("vjx_circuit_create: branch-voltage node index out of range");
        return VJX_ERR_ARG;
      }
    for (size_t e = 0; e < N * dm->n_jac; ++e)
      if (md.jac_dst[e] < -1 || md.jac_dst[e] >= d->n_jac_contrib) {
        set_error("vjx_circuit_create: Jacobian contribution slot out of range");
        return VJX_ERR_ARG;
      }
    for (size_t e = 0; e < N * dm->n_res; ++e)
      if (md.res_dst[e] < -1 || md.res_dst[e] >= d->n_res_contrib) {
        set_error("vjx_circuit_create: residual contribution slot out of range");
        return VJX_ERR_ARG;
      }
    VJX_TRY_HIP(dm->vn1.upload(md.v_node1, N * dm->n_v));
    VJX_TRY_HIP(dm->vn2.upload(md.v_node2, N * dm->n_v));
    VJX_TRY_HIP(dm->jac_dst.upload(md.jac_dst, N * dm->n_jac));
    VJX_TRY_HIP(dm->res_dst.upload(md.res_dst, N * dm->n_res));
    VJX_TRY_HIP(dm->par.upload(md.params, N * dm->n_par));
    c->models.push_back(std::move(dm));
  }
  for (int e = 0; e < 2 * k; ++e)
    if (d->is_dst[e] < -1 || d->is_dst[e] >= d->n_res_contrib) {
      set_error("vjx_circuit_create: current-source slot out of range");
      return VJX_ERR_ARG;
    }
  for (int i = 0; i < m; ++i)
    if (d->vs_node_p[i] < 0 || d->vs_node_p[i] >= c->n_total || d->vs_node_n[i] < 0 ||
        d->vs_node_n[i] >= c->n_total) {
      set_error("vjx_circuit_create: voltage-source node out of range");
      return VJX_ERR_ARG;
    }
  for (int e = 0; e < d->br_ptr[nu]; ++e) {
    int kk = d->br_ent[e] >= 0 ? d->br_ent[e] : ~d->br_ent[e];
    if (kk >= m) {
      set_error("vjx_circuit_create: branch entry out of range");
      return VJX_ERR_ARG;
    }
  }
  if (d->n_noi < 0 || (d->n_noi > 0 && !d->noi_rows)) {
    set_error("vjx_circuit_create: bad NOI list");
    return VJX_ERR_ARG;
  }
  if (d->n_noi > 0) {
    std::vector<char> is_noi(nu, 0);
    for (int i = 0; i < d->n_noi; ++i) {
      const int r = d->noi_rows[i];
      if (r < 0 || r >= nu) {
        set_error("vjx_circuit_create: NOI row out of range");
        return VJX_ERR_ARG;
      }
      is_noi[r] = 1;
    }
    std::vector<int32_t> rows, zero, diag;
    for (int r = 0; r < nu; ++r)
      if (is_noi[r]) rows.push_back(r);
    for (int r = 0; r < nu; ++r)
      for (int e = d->indptr[r]; e < d->indptr[r + 1]; ++e) {
        const int col = d->indices[e];
        if (is_noi[r] && col == r)
          diag.push_back(e);
        else if (is_noi[r] || (col < nu && is_noi[col]))
          zero.push_back(e);
      }
    if (diag.size() != rows.size()) {
      set_error("vjx_circuit_create: a NOI row has no diagonal entry");
      return VJX_ERR_ARG;
    }
    c->n_noi = (int)rows.size();
    c->n_noi_zero = (int)zero.size();
    VJX_TRY_HIP(c->noi_rows.upload(rows));
    VJX_TRY_HIP(c->noi_zero.upload(zero));
    VJX_TRY_HIP(c->noi_diag.upload(diag));
  }
  VJX_TRY_HIP(c->d_sc.alloc(1));
  VJX_TRY_HIP(hipHostMalloc((void**)&c->h_sc, sizeof(DevScalars), hipHostMallocDefault));
  std::memset(c->h_sc, 0, sizeof(DevScalars));
  *out = c.release();
  return VJX_OK;
}

void vjx_circuit_destroy(vjx_circuit* c) { delete c; }

vjx_pattern* vjx_circuit_pattern(vjx_circuit* c) { return c ? c->pat : nullptr; }

int vjx_circuit_limit_states(const vjx_circuit* c) { return c ? c->n_lim : 0; }

int vjx_circuit_eval_block(const vjx_circuit* c, int32_t group) {
  if (!c || group < 0 || group >= (int32_t)c->models.size()) {
    set_error("vjx_circuit_eval_block: null circuit or group out of range");
    return VJX_ERR_ARG;
  }
  return c->models[group]->block;
}



int vjx_build_system(vjx_circuit* c, const double* X, const double* vsrc, const double* isrc,
                     const double* Q_prev, const double* dQdt_prev, const double* Q_prev2,
                     double c0, double c1, double d1, double c2, double gmin, double gshunt,
                     double* csr, double* f, double* Q, double* I_vs, double* mrc,
                     const double* ls_in, double* ls_out) {
  if (!c || !X || !f || (c->m && (!vsrc || !I_vs)) || (c->k && !isrc)) {
    set_error("vjx_build_system: null argument");
    return VJX_ERR_ARG;
  }
  hipStream_t s = c->pat->stream;
  DevScalars* h = c->h_sc;
  h->c0 = c0;
  h->c1 = c1;
  h->c2 = c2;
  h->d1 = d1;
  h->gmin = gmin;
  h->gmin_dev = gmin;
  h->gshunt = gshunt;
  h->reltol = 1e-3;
  h->vntol = 1e-6;
  h->abstol = 1e-12;
  h->nr_damping = 1.0;
  h->max_step = 1e30;
  h->iteration = 0;
  VJX_TRY_HIP(hipMemcpyAsync(c->d_sc.p, h, sizeof(DevScalars), hipMemcpyHostToDevice, s));
  circuit_build_async(c, X, vsrc, isrc, Q_prev, dQdt_prev, Q_prev2, nullptr, csr, f, Q, mrc, I_vs,
                      nullptr, ls_in, ls_out, s);
  VJX_TRY_HIP(hipGetLastError());
  VJX_TRY_HIP(hipStreamSynchronize(s));
  return VJX_OK;
}

int vjx_eval_stamp(vjx_circuit* c, int32_t model_index, const double* X, const double* vsrc,
                   const double* isrc, double c0, double gmin, double* csr, double* f, double* Q,
                   double* mrc, const double* ls_in, double* ls_out) {
  if (!c || !X || !f || !Q || !mrc || (c->m && !vsrc) || (c->k && model_index < 0 && !isrc)) {
    set_error("vjx_eval_stamp: null argument");
    return VJX_ERR_ARG;
  }
  if (model_index < -1 || model_index >= (int32_t)c->models.size()) {
    set_error("vjx_eval_stamp: model_index out of range");
    return VJX_ERR_ARG;
  }
  hipStream_t s = c->pat->stream;
  DevScalars* h = c->h_sc;
  std::memset(h, 0, sizeof(DevScalars));
  h->c0 = c0;
  h->gmin = gmin;
  h->gmin_dev = gmin;
  h->reltol = 1e-3;
  h->vntol = 1e-6;
  h->abstol = 1e-12;
  h->nr_damping = 1.0;
  h->max_step = 1e30;
  VJX_TRY_HIP(hipMemcpyAsync(c->d_sc.p, h, sizeof(DevScalars), hipMemcpyHostToDevice, s));
  DBuf<double> zeros, ivs;
  VJX_TRY_HIP(zeros.alloc(std::max(c->n_unk, 1)));
  VJX_TRY_HIP(ivs.alloc(std::max(c->m, 1)));
  launch_fill(c->n_unk, 0.0, zeros.p, s);
  circuit_build_async(c, X, vsrc, isrc, zeros.p, zeros.p, zeros.p, nullptr, csr, f, Q, mrc, ivs.p,
                      nullptr, ls_in, ls_out, s, model_index);
  VJX_TRY_HIP(hipGetLastError());
  VJX_TRY_HIP(hipStreamSynchronize(s));
  return VJX_OK;
}

int vjx_model_eval(int32_t model_id, int32_t count, const double* vb, const double* par,
                   const double* li, double* rr, double* rc, double* jr, double* jc, double* lr,
                   double* lc, double* lo) {
  if (model_id < 0 || model_id >= VJX_MODEL_COUNT || count < 0) {
    set_error("vjx_model_eval: bad model id or count");
    return VJX_ERR_ARG;
  }
  if (count && (!vb || !par || !rr || !rc || !jr || !jc)) {
    set_error("vjx_model_eval: null argument");
    return VJX_ERR_ARG;
  }
  launch_model_eval_raw(model_id, count, vb, par, li, rr, rc, jr, jc, lr, lc, lo, 0);
  VJX_TRY_HIP(hipGetLastError());
  VJX_TRY_HIP(hipDeviceSynchronize());
  return VJX_OK;
}

int vjx_model_eval_split(int32_t model_id, int32_t count, const vjx_eval_split* sp,
                         const double* shared_params, const double* dp, const double* sc,
                         const double* dc, const double* simparams, const double* li, double* rr,
                         double* rc, double* jr, double* jc, double* lr, double* lc, double* sr,
                         double* sq, double* lo) {
  (void)shared_params;  // no model here reads a parameter outside its cache
  if (model_id < 0 || model_id >= VJX_MODEL_COUNT || count < 0 || !sp) {
    set_error("vjx_model_eval_split: bad model id, count or layout");
    return VJX_ERR_ARG;
  }
  const int n_v = model_n_v(model_id), n_par = model_n_par(model_id);
  const int n_res = model_n_res(model_id), n_jac = model_n_jac(model_id);
  const int n_lim = model_n_lim(model_id);
  if (count == 0) return VJX_OK;
  if (!dp || !rr || !rc || !jr || !jc || !lr || !lc || !sr || !sq || (n_lim && !lo) ||
      !sp->voltage_positions || !sp->cache_src || (sp->mfactor_index >= 0 && !simparams)) {
    set_error("vjx_model_eval_split: null argument");
    return VJX_ERR_ARG;
  }
  bool need_sc = false, need_dc = false;
  for (int k = 0; k < n_v; ++k)
    if (sp->voltage_positions[k] < 0 || sp->voltage_positions[k] >= sp->n_device_params) {
      set_error("vjx_model_eval_split: voltage position outside device_params");
      return VJX_ERR_ARG;
    }
  for (int j = 0; j < n_par; ++j) {
    const int c = sp->cache_src[j];
    if (c < 0 && ~c >= sp->n_device_cache) {
      set_error("vjx_model_eval_split: cache column outside device_cache");
      return VJX_ERR_ARG;
    }
    need_sc = need_sc || c >= 0;
    need_dc = need_dc || c < 0;
  }
  if ((need_sc && !sc) || (need_dc && !dc)) {
    set_error("vjx_model_eval_split: null cache array");
    return VJX_ERR_ARG;
  }
  DBuf<int32_t> vpos, csrc;
  DBuf<double> vb, par;
  VJX_TRY_HIP(vpos.upload(sp->voltage_positions, n_v));
  VJX_TRY_HIP(csrc.upload(sp->cache_src, n_par));
  VJX_TRY_HIP(vb.alloc((size_t)n_v * count));
  VJX_TRY_HIP(par.alloc((size_t)n_par * count));
  launch_split_gather(count, n_v, n_par, sp->n_device_params, sp->n_device_cache, vpos.p, csrc.p,
                      sc, dp, dc, vb.p, par.p, 0);
  launch_model_eval_raw(model_id, count, vb.p, par.p, li, rr, rc, jr, jc, lr, lc, lo, 0);
  launch_split_finish(count, n_res, n_jac, n_lim > 0 ? 1 : 0,
                      sp->mfactor_index >= 0 ? simparams + sp->mfactor_index : nullptr, rr, rc,
                      jr, jc, lr, lc, sr, sq, 0);
  VJX_TRY_HIP(hipGetLastError());
  VJX_TRY_HIP(hipDeviceSynchronize());
  return VJX_OK;
}

int vjx_limit_eval(int32_t kind, int32_t n, const double* vnew, const double* vold, double p1,
                   double p2, double* out) {
  if ((kind != 0 && kind != 1) || n < 0 || (n && (!vnew || !vold || !out))) {
    set_error("vjx_limit_eval: kind must be 0 (pnjlim) or 1 (fetlim), arrays non-null");
    return VJX_ERR_ARG;
  }
  launch_limit(kind, n, vnew, vold, p1, p2, out, 0);
  VJX_TRY_HIP(hipGetLastError());
  VJX_TRY_HIP(hipDeviceSynchronize());
  return VJX_OK;
}

int vjx_eval_sources(vjx_circuit* c, double t, double* vsrc, double* isrc) {
  if (!c) {
    set_error("vjx_eval_sources: null circuit");
    return VJX_ERR_ARG;
  }
  hipStream_t s = c->pat->stream;
  if (c->m && vsrc) launch_sources(c->vsrc_table(), t, vsrc, s);
  if (c->k && isrc) launch_sources(c->isrc_table(), t, isrc, s);
  VJX_TRY_HIP(hipGetLastError());
  VJX_TRY_HIP(hipStreamSynchronize(s));
  return VJX_OK;
}

}  // extern "C"
