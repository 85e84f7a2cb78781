// librbh: host runtime and C ABI (include/rbh.h).
//
// Device memory model: one context owns
//   * a shape arena (body-frame vertices, triangles, SDF values),
//   * a body arena (SoA state + static data + per-step derived data) in which
//     every scene occupies a contiguous range,
//   * a triangle-AABB arena (per body instance),
//   * scene tables (parameters, per-step state, residual history),
//   * a step workspace sized for the current batch (pairs, candidates,
//     contacts, colouring, solver arrays), grown on demand.
// A step over a batch of scenes is one fixed launch sequence; data-dependent
// sizes stay on the device, so the sequence can be captured into a hipGraph.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstddef>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/rbh.h"
#include "../../include/rbh_testing.h"
#include "rbh_kernels.h"
#include "rbh_sdf.h"

using namespace rbh;

// Staging array of rbh_sdf_build_noalloc: one copy per GPU, held until the
// process ends.  Host code never passes `scratch` itself to a kernel or a
// copy (on the host the name is only a stand-in); it uses the device address
// hipGetSymbolAddress returns, looked up once per device (scratch_address).
constexpr size_t NOALLOC_SCRATCH_BYTES = (size_t)32 << 20;
static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int SCRATCH_MAX_DEVICES = 64;
std::mutex g_scratch_mu;
unsigned char* g_scratch_ptr[SCRATCH_MAX_DEVICES] = {};

// Device address of `device`'s scratch array; the device must be current.
hipError_t scratch_address(int device, unsigned char** out) {
  *out = nullptr;
  if (device < 0 || device >= SCRATCH_MAX_DEVICES) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch_ptr[device] = static_cast<unsigned char*>(p);
  }
  *out = g_scratch_ptr[device];
  return hipSuccess;
}

struct DBuf {
  void* p = nullptr;
  size_t bytes = 0;
};

struct RangeAlloc {
  std::vector<std::pair<int, int>> free_;   // (offset, size), sorted by offset
  int cap = 0;
  void init(int c) { cap = c; free_.clear(); if (c > 0) free_.push_back({0, c}); }
  int alloc(int n) {
    for (size_t i = 0; i < free_.size(); ++i) {
      if (free_[i].second >= n) {
        int off = free_[i].first;
        free_[i].first += n;
        free_[i].second -= n;
        if (free_[i].second == 0) free_.erase(free_.begin() + i);
        return off;
      }
    }
    return -1;
  }
  void release(int off, int n) {
    if (n <= 0) return;
    free_.push_back({off, n});
    std::sort(free_.begin(), free_.end());
    std::vector<std::pair<int, int>> m;
    for (auto& f : free_) {
      if (!m.empty() && m.back().first + m.back().second == f.first) m.back().second += f.second;
      else m.push_back(f);
    }
    free_ = m;
  }
  void grow(int newcap) {
    if (newcap <= cap) return;
    release(cap, newcap - cap);
    cap = newcap;
  }
  int used() const {
    int f = 0;
    for (auto& x : free_) f += x.second;
    return cap - f;
  }
};

struct SceneHost {
  bool live = false;
  int body_off = 0, nbody = 0, tri_off = 0, ntri = 0, mat_off = 0, nmat = 0, res_off = 0, res_n = 0;
  rbh_params params{};
  std::vector<int> h_parent, h_child;   // hinges: scene-local body indices
  std::vector<double> h_geo;            // 12 per hinge: arm_p, arm_c, axis_p, axis_c (body frames)
  long long ws_gen = -1;   // workspace generation of the scene's last step
  int batch_index = -1;    // index inside that batch
};

// Stage boundaries of a step (events recorded after each stage):
// PREP body prep + refit, BROAD broad phase, FWK Frank-Wolfe, FW reduction
// and contact order, ASM assembly, COLOR colouring and sweep order, SOLVE the
// solve, INTG integration, POST post-stabilisation, INT scene records.
enum Stage { ST_BEGIN = 0, ST_PREP, ST_BROAD, ST_FWK, ST_FW, ST_ASM, ST_COLOR, ST_SOLVE, ST_INTG, ST_POST, ST_INT, ST_N };
// rbh_get_timings entries
constexpr int TIMING_OUT = 10;

}  // namespace

struct rbh_ctx {
  int device = 0, prec = 4;
  hipStream_t stream = nullptr;
  unsigned long long budget = 0;
  std::string err;
  hipDeviceProp_t prop{};
  // shapes
  std::vector<ShapeInfo> shapes;
  std::vector<double> sV;   // 4 per vertex
  std::vector<int> sT;      // 4 per triangle
  std::vector<double> sdf;
  bool shapes_dirty = true;
  DBuf d_shapes, d_sV, d_sT, d_sdf, d_cells;
  long long cells_total = 0;   // values of the cell-major grid copies of all shapes
  // body arena
  RangeAlloc bodies, tris, res;
  DBuf r, q, v, w, b_shape, b_slot, b_fixed, b_mat, b_tri, b_mass, b_inertia, b_grav, b_damp, b_broad;
  DBuf W0, W1, du0, du1, lo, hi, ws0, ws1, ww0, ww1, hd0, hd1;
  DBuf tlo, thi, clo, chi;
  // hinges of all live scenes, packed by scene slot (rebuilt when they change)
  DBuf h_ids, h_geo, hb_body, hb_inc_off, hb_inc, h_J, h_WJ, h_b, h_r0, d_hmer;
  bool hinges_dirty = false;
  bool any_hinge = false;   // a scene of the batch has hinges (set per run_steps call)
  size_t hinge_lds = 0;     // dynamic LDS of the hinge solve for the batch (its largest scene)
  // scenes
  std::vector<SceneHost> scenes;
  std::vector<SceneParams> sp_host;
  std::vector<double> mat_host;   // interaction tables of the live scenes (4 per entry), ranges from `mats`
  RangeAlloc mats;                // entries of mat_host: a scene's nmat^2 block, reused after destroy
  DBuf d_sp, d_ss, d_mat, d_res, d_fwacc;
  bool scenes_dirty = true;
  long long arena_gen = 0;
  // batch workspace
  std::vector<int> batch;
  long long batch_arena_gen = -1;
  int nb = 0, ns = 0;
  DBuf bb_body, bs_slot, bs_order, bs_foff, b2f, f_end, d_nb;
  DBuf row_cnt, row_off, pairs, pair_sz, slot_off, n_pairs, slot_flag, n_slots, pair_items, item_off;
  DBuf side_cnt, side_off, n_sides;
  DBuf cand_pair, cand_tri, pair_cstart, n_cand, cand_p, cand_n, cand_keep, keep_scan;
  DBuf kidx, kpos, n_kept, kept_flag, kept_scan;
  DBuf c_ia, c_ib, c_tri, c_pair, c_p, c_n, n_contacts, c_rA, c_rB, c_b, c_r0, c_mu, c_imp, c_color;
  DBuf f_bs, uf, isl_cnt, isl_off, isl_fill, isl_max, c_lab, uperm, u_cnt, u_base, n_units, u_start, u_end, u_ncol,
      u_bs, body_claim, body_loc, bc_cnt, bc_off, bc_fill, bc_list, pal, g_pc, g_done, col_off, merit, l_pos,
      big_list, item_pair_of;
  DBuf u_T, isl_unit, u_nbody, u_boff, f_loc, f_unit, u_blist, s_loc, s_W;
  DBuf l_cnt, l_base, n_lanes, l_start, l_end, l_ncol, l_bs;
  DBuf s_JL, s_JA, s_JB, xr0, xr1;
  DBuf perm, s_id, s_n, s_rA, s_rB, s_b, s_r0, x0, x1, errbuf, scan_state;
  int cap_pairs = 0, cap_slots = 0, cap_cand = 0, cap_contacts = 0, cap_units = 0;
  int scan_stride = 0;       // 64-bit words of look-back state per scan of the step (ticket + tile status)
  int scan_site = 0;         // scans launched so far in the step being enqueued
  int unit_T = 64;          // contacts per colouring unit (a scene's units are windows of this size)
  rbh_tuning tune{};        // rbh_set_tuning (0 fields: defaults)
  // resident solve groups: LDS capacities (contacts, bodies, colour offsets) and the launch's dynamic LDS
  int res_cap_c = 0, res_cap_b = 0, res_cap_o = 0;
  size_t solve_lds = 0, solve_lds_half = 0;   // full and half layouts (one or two workgroups per CU)
  int max_iter_batch = 1, coop_max = 0, last_units = 0;
  bool multi_ok = false;
  bool any_post = false;    // a scene of the batch has post-stabilisation on (set per run_steps call)
  long long ws_gen = 0;
  // graph
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  long long graph_key_gen = -1;
  double graph_dt = 0;
  int graph_flags = -1;
  // timing
  // timing: ST_N events per step, for up to TIMED_STEPS steps of one call
  std::vector<hipEvent_t> evp;
  double timings[TIMING_OUT]{};
  int timed_steps = 0;
  long long counters[4]{};
  // device address of this device's no-allocation staging array (rbh_sdf_build_noalloc),
  // looked up once per device when the first context on it is created
  unsigned char* scratch = nullptr;
  hipError_t scratch_err = hipSuccess;
};

namespace {

#define HIPCHK(ctx, call)                                                                       \
  do {                                                                                          \
    hipError_t e_ = (call);                                                                     \
    if (e_ != hipSuccess) {                                                                     \
      (ctx)->err = std::string(#call) + ": " + hipGetErrorString(e_);                           \
      return e_ == hipErrorOutOfMemory ? RBH_E_MEMORY : RBH_E_HIP;                              \
    }                                                                                           \
  } while (0)

int fail(rbh_ctx* c, int code, const std::string& msg) {
  if (c) c->err = msg;
  return code;
}

size_t total_bytes(rbh_ctx* c);

// grow a device buffer; keep the old contents if `keep`
int ensure(rbh_ctx* c, DBuf& b, size_t bytes, bool keep = false) {
  if (bytes <= b.bytes && b.p) return 0;
  size_t nb = std::max<size_t>(bytes, 256);
  if (c->budget && total_bytes(c) - b.bytes + nb > c->budget)
    return fail(c, RBH_E_MEMORY, "HBM budget exceeded (" + std::to_string(c->budget) + " bytes)");
  void* p = nullptr;
  HIPCHK(c, hipMalloc(&p, nb));
  if (keep && b.p && b.bytes) HIPCHK(c, hipMemcpyAsync(p, b.p, b.bytes, hipMemcpyDeviceToDevice, c->stream));
  if (b.p) {
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipFree(b.p));
  }
  b.p = p;
  b.bytes = nb;
  return 0;
}

std::vector<DBuf*> all_bufs(rbh_ctx* c) {
  return {&c->d_shapes, &c->d_sV, &c->d_sT, &c->d_sdf, &c->d_cells, &c->r, &c->q, &c->v, &c->w, &c->b_shape, &c->b_slot,
          &c->b_fixed, &c->b_mat, &c->b_tri, &c->b_mass, &c->b_inertia, &c->b_grav, &c->b_damp, &c->b_broad, &c->W0,
          &c->W1, &c->du0, &c->du1, &c->lo, &c->hi, &c->ws0, &c->ws1, &c->ww0, &c->ww1, &c->tlo, &c->thi, &c->clo, &c->chi,
          &c->d_sp, &c->d_ss, &c->d_mat, &c->d_res, &c->d_fwacc, &c->bb_body, &c->bs_slot, &c->bs_order, &c->bs_foff, &c->b2f,
          &c->f_end, &c->d_nb, &c->row_cnt, &c->row_off, &c->pairs, &c->pair_sz, &c->slot_off, &c->n_pairs, &c->pair_items, &c->item_off,
          &c->slot_flag, &c->side_cnt, &c->side_off, &c->n_sides, &c->n_slots, &c->cand_pair, &c->cand_tri, &c->pair_cstart,
          &c->n_cand, &c->cand_p, &c->cand_n, &c->cand_keep, &c->keep_scan, &c->kidx, &c->kpos,
          &c->n_kept, &c->kept_flag, &c->kept_scan, &c->c_ia, &c->c_ib, &c->c_tri,
          &c->c_pair, &c->c_p, &c->c_n, &c->n_contacts, &c->c_rA, &c->c_rB, &c->c_b, &c->c_r0, &c->c_mu,
          &c->c_imp, &c->c_color, &c->f_bs, &c->uf, &c->isl_cnt, &c->isl_off, &c->isl_fill, &c->isl_max,
          &c->c_lab, &c->uperm, &c->u_cnt, &c->u_base, &c->n_units, &c->u_start, &c->u_end, &c->u_ncol,
          &c->u_bs, &c->body_claim, &c->body_loc, &c->bc_cnt, &c->bc_off, &c->bc_fill, &c->bc_list, &c->pal,
          &c->g_pc, &c->g_done, &c->col_off, &c->merit, &c->l_pos, &c->big_list, &c->item_pair_of, &c->perm, &c->s_id, &c->s_n, &c->s_rA,
          &c->s_rB, &c->s_b, &c->s_r0, &c->x0, &c->x1, &c->errbuf, &c->scan_state, &c->u_T, &c->isl_unit,
          &c->u_nbody, &c->u_boff, &c->f_loc, &c->f_unit, &c->u_blist, &c->s_loc, &c->s_W, &c->l_cnt,
          &c->l_base, &c->n_lanes, &c->l_start, &c->l_end, &c->l_ncol, &c->l_bs, &c->s_JL, &c->s_JA, &c->s_JB,
          &c->xr0, &c->xr1, &c->hd0, &c->hd1, &c->h_ids, &c->h_geo, &c->hb_body, &c->hb_inc_off, &c->hb_inc,
          &c->h_J, &c->h_WJ, &c->h_b, &c->h_r0, &c->d_hmer};
}

size_t total_bytes(rbh_ctx* c) {
  size_t t = 0;
  for (DBuf* b : all_bufs(c)) t += b->bytes;
  return t;
}

void free_all(rbh_ctx* c) {
  for (DBuf* b : all_bufs(c)) {
    if (b->p) (void)hipFree(b->p);
    b->p = nullptr;
    b->bytes = 0;
  }
}

void drop_graph(rbh_ctx* c) {
  if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
  if (c->graph) (void)hipGraphDestroy(c->graph);
  c->gexec = nullptr;
  c->graph = nullptr;
  c->graph_key_gen = -1;
}

// ------------------------------------------------------------ arenas
int body_elem_sizes(rbh_ctx* c, std::vector<std::pair<DBuf*, size_t>>& out) {
  const size_t R = c->prec, V = 4 * R;
  out = {{&c->r, V},      {&c->q, V},       {&c->v, V},         {&c->w, V},        {&c->b_shape, 4},
         {&c->b_slot, 4}, {&c->b_fixed, 4}, {&c->b_mat, 4},     {&c->b_tri, 4},    {&c->b_mass, R},
         {&c->b_inertia, V}, {&c->b_grav, V}, {&c->b_damp, V},  {&c->W0, V},       {&c->W1, V},
         {&c->du0, V},    {&c->du1, V},     {&c->lo, V},        {&c->hi, V},       {&c->ws0, V},
         {&c->ws1, V},    {&c->ww0, V},     {&c->ww1, V},      {&c->hd0, V},      {&c->hd1, V},
         {&c->b_broad, 16}};
  return 0;
}

int grow_bodies(rbh_ctx* c, int need) {
  int cap = std::max(need, std::max(1024, c->bodies.cap * 2));
  std::vector<std::pair<DBuf*, size_t>> el;
  body_elem_sizes(c, el);
  for (auto& e : el) {
    int rc = ensure(c, *e.first, (size_t)cap * e.second, true);
    if (rc) return rc;
  }
  c->bodies.grow(cap);
  c->arena_gen++;
  drop_graph(c);
  return 0;
}

int grow_tris(rbh_ctx* c, int need) {
  int cap = std::max(need, std::max(1 << 16, c->tris.cap * 2));
  size_t V = 4 * (size_t)c->prec;
  int rc = ensure(c, c->tlo, (size_t)cap * V, true);
  if (!rc) rc = ensure(c, c->thi, (size_t)cap * V, true);
  if (!rc) rc = ensure(c, c->clo, (size_t)cap * V, true);
  if (!rc) rc = ensure(c, c->chi, (size_t)cap * V, true);
  if (rc) return rc;
  c->tris.grow(cap);
  c->arena_gen++;
  drop_graph(c);
  return 0;
}

int grow_res(rbh_ctx* c, int need) {
  int cap = std::max(need, std::max(4096, c->res.cap * 2));
  int rc = ensure(c, c->d_res, (size_t)cap * 8, true);
  if (!rc) rc = ensure(c, c->d_hmer, (size_t)cap * 8, true);
  if (rc) return rc;
  c->res.grow(cap);
  c->arena_gen++;
  drop_graph(c);
  return 0;
}

template <typename T> int upload(rbh_ctx* c, DBuf& b, const std::vector<T>& h) {
  int rc = ensure(c, b, std::max<size_t>(h.size() * sizeof(T), 16));
  if (rc) return rc;
  if (!h.empty()) HIPCHK(c, hipMemcpyAsync(b.p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, c->stream));
  return 0;
}

template <typename R> int sync_shapes(rbh_ctx* c) {
  if (!c->shapes_dirty) return 0;
  std::vector<V4<R>> v(c->sV.size() / 4);
  for (size_t i = 0; i < v.size(); ++i) v[i] = V4<R>{R(c->sV[4 * i]), R(c->sV[4 * i + 1]), R(c->sV[4 * i + 2]), R(0)};
  std::vector<R> g(c->sdf.size());
  for (size_t i = 0; i < g.size(); ++i) g[i] = R(c->sdf[i]);
  int rc = upload(c, c->d_shapes, c->shapes);
  if (!rc) rc = upload(c, c->d_sV, v);
  if (!rc) rc = upload(c, c->d_sT, c->sT);
  if (!rc) rc = upload(c, c->d_sdf, g);
  if (!rc) rc = ensure(c, c->d_cells, (size_t)std::max<long long>(c->cells_total, 8) * sizeof(R));
  if (rc) return rc;
  for (const ShapeInfo& s : c->shapes) {
    const long long nc = (long long)(s.I - 1) * (s.J - 1) * (s.K - 1);
    hipLaunchKernelGGL(k_sdf_cells<R>, dim3((unsigned)std::min<long long>((nc + 255) / 256, 65535)), dim3(256), 0,
                       c->stream, reinterpret_cast<const R*>(c->d_sdf.p) + s.g_off, s.I, s.J, s.K,
                       reinterpret_cast<R*>(c->d_cells.p) + s.c_off);
  }
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->shapes_dirty = false;
  c->arena_gen++;
  drop_graph(c);
  return 0;
}

SceneParams to_device_params(const rbh_params& p, int nmat, int mat_off, int res_off) {
  SceneParams s{};
  s.time_step = p.time_step; s.envelope = p.envelope; s.abs_tol = p.absolute_tolerance;
  s.rel_tol = p.relative_tolerance; s.nu_reduce = p.nu_reduce; s.nu_increase = p.nu_increase;
  s.too_small = p.too_small_merit_change; s.fw_tol = p.contact_optimization_tolerance;
  s.gap_reduction = p.gap_reduction; s.min_gap = p.min_gap_value; s.max_gap = p.max_gap_value;
  s.ell_tol = p.ellipsoid_tolerance; s.ell_expansion = p.ellipsoid_expansion;
  s.max_iter = p.max_iterations; s.fw_max_iter = p.contact_optimization_max_iterations;
  s.ell_max_iter = p.ellipsoid_max_iterations; s.use_bounce = p.use_bounce;
  s.use_pre_stab = p.use_pre_stabilization; s.scheme = p.proximal_solver; s.friction = p.friction_model;
  s.use_post_stab = p.use_post_stabilization ? 1 : 0;
  s.nmat = nmat; s.mat_off = mat_off; s.res_off = res_off; s.seed = p.coloring_seed;
  return s;
}

int sync_scenes(rbh_ctx* c) {
  if (!c->scenes_dirty) return 0;
  size_t n = c->scenes.size();
  int rc = upload(c, c->d_sp, c->sp_host);
  if (!rc) rc = upload(c, c->d_mat, c->mat_host);
  if (!rc) rc = ensure(c, c->d_ss, std::max<size_t>(n, 1) * sizeof(SceneState), true);
  if (!rc) rc = ensure(c, c->d_fwacc, std::max<size_t>(n, 1) * FW_SPREAD * FW_STRIDE * sizeof(unsigned long long), true);
  if (rc) return rc;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->scenes_dirty = false;
  return 0;
}

// Pack the hinges of every live scene (scene order) into the device arrays:
// per hinge the arena bodies and hinge-body slots of parent and child, the
// body-frame arms and axes; per scene the list of bodies its hinges join
// (ascending body index) and, per such body, its hinges in hinge order.
template <typename R> int sync_hinges(rbh_ctx* c) {
  if (!c->hinges_dirty) return 0;
  std::vector<int4> ids;
  std::vector<V4<R>> geo;
  std::vector<int> hbody, inc_off, inc;
  for (size_t slot = 0; slot < c->scenes.size(); ++slot) {
    SceneHost& s = c->scenes[slot];
    SceneParams& sp = c->sp_host[slot];
    sp.h_off = (int)ids.size();
    sp.hb_off = (int)hbody.size();
    sp.h_cnt = 0;
    sp.hb_cnt = 0;
    if (!s.live || s.h_parent.empty()) continue;
    const int K = (int)s.h_parent.size();
    std::vector<int> loc(s.nbody, -1);
    for (int k = 0; k < K; ++k) loc[s.h_parent[k]] = loc[s.h_child[k]] = 0;
    std::vector<int> list;
    for (int b = 0; b < s.nbody; ++b)
      if (loc[b] == 0) { loc[b] = (int)list.size(); list.push_back(b); }
    std::vector<std::vector<int>> per(list.size());
    for (int k = 0; k < K; ++k) {
      const int h = sp.h_off + k;
      const int p = s.h_parent[k], ch = s.h_child[k];
      ids.push_back(int4{s.body_off + p, s.body_off + ch, loc[p], loc[ch]});
      for (int j = 0; j < 4; ++j)
        geo.push_back(V4<R>{R(s.h_geo[12 * k + 3 * j]), R(s.h_geo[12 * k + 3 * j + 1]), R(s.h_geo[12 * k + 3 * j + 2]),
                            R(0)});
      per[loc[p]].push_back(2 * h);
      per[loc[ch]].push_back(2 * h + 1);
    }
    for (size_t j = 0; j < list.size(); ++j) {
      hbody.push_back(s.body_off + list[j]);
      inc_off.push_back((int)inc.size());
      inc.insert(inc.end(), per[j].begin(), per[j].end());
    }
    sp.h_cnt = K;
    sp.hb_cnt = (int)list.size();
  }
  inc_off.push_back((int)inc.size());
  const size_t nh = std::max<size_t>(ids.size(), 1), V = 4 * (size_t)c->prec;
  int rc = upload(c, c->h_ids, ids);
  if (!rc) rc = upload(c, c->h_geo, geo);
  if (!rc) rc = upload(c, c->hb_body, hbody);
  if (!rc) rc = upload(c, c->hb_inc_off, inc_off);
  if (!rc) rc = upload(c, c->hb_inc, inc);
  if (!rc) rc = ensure(c, c->h_J, 20 * nh * V);
  if (!rc) rc = ensure(c, c->h_WJ, 20 * nh * V);
  if (!rc) rc = ensure(c, c->h_b, 5 * nh * c->prec);
  if (!rc) rc = ensure(c, c->h_r0, 5 * nh * c->prec);
  if (!rc) rc = ensure(c, c->d_hmer, (size_t)std::max(c->res.cap, 1) * 8, true);
  if (rc) return rc;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->hinges_dirty = false;
  c->scenes_dirty = true;
  c->arena_gen++;
  drop_graph(c);
  return 0;
}

int check_params(rbh_ctx* c, const rbh_params& p) {
  if (!(p.max_iterations >= 0) || p.contact_optimization_max_iterations < 0 || p.ellipsoid_max_iterations < 0)
    return fail(c, RBH_E_ARG, "iteration counts must be non-negative");
  if (p.proximal_solver < 0 || p.proximal_solver > 3)
    return fail(c, RBH_E_ARG, "Unknown solver scheme: " + std::to_string(p.proximal_solver));
  if (p.friction_model < 0 || p.friction_model > 1) return fail(c, RBH_E_ARG, "unknown friction model");
  if (!(p.envelope >= 0.0)) return fail(c, RBH_E_ARG, "envelope must be non-negative");
  return 0;
}

// ------------------------------------------------------------ batch workspace
// wavefronts per workgroup of the batched solve (4 or 8; rbh_tuning.waves_per_group)
// broad-phase work items: at most ceil(nt_A / CAND_BLK) + ceil(nt_B / CAND_BLK) per pair, so at most
// slots / CAND_BLK + 2 per pair
int cap_items(const rbh_ctx* c) { return c->cap_slots / CAND_BLK + 2 * c->cap_pairs + 64; }
int batch_waves(const rbh_ctx* c) { return c->tune.waves_per_group == 8 ? 8 : 4; }
// the batched solve's LDS layout: half (two workgroups per CU) when asked, or
// by default when the batch has more scenes than the device has CUs
bool batch_half(const rbh_ctx* c, int ns) {
  if (c->tune.solve_groups_per_cu == 1) return false;
  if (c->tune.solve_groups_per_cu == 2) return true;
  return ns > c->prop.multiProcessorCount;
}
// one workgroup per batch scene (the post-stabilisation solve too)
template <typename R> void launch_batch_solve(rbh_ctx* c, const Dev<R>& d, int ns, hipStream_t s) {
  const bool half = batch_half(c, ns);
  const size_t lds = half ? c->solve_lds_half : c->solve_lds;
  if (batch_waves(c) == 8) {
    if (half) hipLaunchKernelGGL((k_solve<R, 512, 1>), dim3(ns), dim3(512), lds, s, d);
    else hipLaunchKernelGGL((k_solve<R, 512, 0>), dim3(ns), dim3(512), lds, s, d);
  } else {
    if (half) hipLaunchKernelGGL((k_solve<R, SOLVE_BLOCK, 1>), dim3(ns), dim3(SOLVE_BLOCK), lds, s, d);
    else hipLaunchKernelGGL((k_solve<R, SOLVE_BLOCK, 0>), dim3(ns), dim3(SOLVE_BLOCK), lds, s, d);
  }
}

// multi: a single scene solved by a cooperative grid (its lanes are its units)
template <typename R> Dev<R> make_dev(rbh_ctx* c, double dt, int flags, bool multi) {
  Dev<R> d{};
#define P(T, b) reinterpret_cast<T>(c->b.p)
  d.shapes = P(const ShapeInfo*, d_shapes); d.sV = P(const V4<R>*, d_sV); d.sT = P(const int4*, d_sT);
  d.sdf = P(const R*, d_sdf);
  d.cells = P(const R*, d_cells);
  d.sp = P(const SceneParams*, d_sp); d.ss = P(SceneState*, d_ss); d.residuals = P(double*, d_res);
  d.fw_acc = P(unsigned long long*, d_fwacc);
  d.mat = P(const double*, d_mat);
  d.r = P(V4<R>*, r); d.q = P(V4<R>*, q); d.v = P(V4<R>*, v); d.w = P(V4<R>*, w);
  d.b_shape = P(const int*, b_shape); d.b_slot = P(const int*, b_slot); d.b_fixed = P(const int*, b_fixed);
  d.b_mat = P(const int*, b_mat); d.b_tri = P(const int*, b_tri); d.b_mass = P(const R*, b_mass);
  d.b_broad = P(const int4*, b_broad);
  d.b_inertia = P(const V4<R>*, b_inertia); d.b_grav = P(const V4<R>*, b_grav); d.b_damp = P(const V4<R>*, b_damp);
  d.b_W0 = P(V4<R>*, W0); d.b_W1 = P(V4<R>*, W1); d.b_du0 = P(V4<R>*, du0); d.b_du1 = P(V4<R>*, du1);
  d.b_lo = P(V4<R>*, lo); d.b_hi = P(V4<R>*, hi); d.ws0 = P(V4<R>*, ws0); d.ws1 = P(V4<R>*, ws1);
  d.ww0 = P(V4<R>*, ww0); d.ww1 = P(V4<R>*, ww1); d.tlo = P(V4<R>*, tlo); d.thi = P(V4<R>*, thi);
  d.clo = P(V4<R>*, clo); d.chi = P(V4<R>*, chi);
  d.nb = c->nb; d.ns = c->ns;
  d.bb_body = P(const int*, bb_body); d.bs_slot = P(const int*, bs_slot);
  d.bs_order = P(const int*, bs_order); d.bs_foff = P(const int*, bs_foff);
  d.b2f = P(const int*, b2f); d.f_end = P(const int*, f_end); d.d_nb = P(const int*, d_nb);
  d.row_cnt = P(int*, row_cnt); d.row_off = P(int*, row_off);
  d.pairs = P(int2*, pairs); d.pair_sz = P(int*, pair_sz); d.slot_off = P(int*, slot_off);
  d.pair_items = P(int*, pair_items); d.item_off = P(int*, item_off);
  d.n_pairs = P(int*, n_pairs); d.cap_pairs = c->cap_pairs;
  d.slot_flag = P(unsigned char*, slot_flag); d.n_slots = P(int*, n_slots);
  d.side_cnt = P(int*, side_cnt); d.side_off = P(int*, side_off); d.n_sides = P(int*, n_sides);
  d.cap_slots = c->cap_slots;
  d.cand_pair = P(int*, cand_pair); d.cand_tri = P(int*, cand_tri); d.pair_cstart = P(int*, pair_cstart);
  d.n_cand = P(int*, n_cand); d.cap_cand = c->cap_cand;
  d.cand_p = P(V4<R>*, cand_p); d.cand_n = P(V4<R>*, cand_n); d.cand_keep = P(unsigned char*, cand_keep);
  d.keep_scan = P(int*, keep_scan);
  d.kidx = P(int*, kidx); d.kpos = P(V4<R>*, kpos); d.n_kept = P(int*, n_kept);
  d.kept_flag = P(unsigned char*, kept_flag); d.kept_scan = P(int*, kept_scan);
  d.c_ia = P(int*, c_ia); d.c_ib = P(int*, c_ib); d.c_tri = P(int*, c_tri); d.c_pair = P(int*, c_pair);
  d.c_p = P(V4<R>*, c_p); d.c_n = P(V4<R>*, c_n); d.n_contacts = P(int*, n_contacts);
  d.cap_contacts = c->cap_contacts;
  d.c_rA = P(V4<R>*, c_rA); d.c_rB = P(V4<R>*, c_rB); d.c_b = P(V4<R>*, c_b); d.c_r0 = P(V4<R>*, c_r0);
  d.c_mu = P(V4<R>*, c_mu); d.c_imp = P(V4<R>*, c_imp); d.c_color = P(int*, c_color);
  d.f_bs = P(const int*, f_bs); d.uf = P(int*, uf); d.isl_cnt = P(int*, isl_cnt); d.isl_off = P(int*, isl_off);
  d.isl_fill = P(int*, isl_fill); d.isl_max = P(int*, isl_max); d.c_lab = P(int*, c_lab); d.uperm = P(int*, uperm);
  d.u_cnt = P(int*, u_cnt); d.u_base = P(int*, u_base); d.n_units = P(int*, n_units); d.cap_units = c->cap_units;
  d.unit_T = c->unit_T;
  d.unit_split = multi ? 0 : (c->tune.lanes_per_scene > 0 ? c->tune.lanes_per_scene : batch_waves(c));
  d.items_nb = multi ? -1 : (batch_half(c, c->ns) ? ResCap<R, 1>::VB : ResCap<R, 0>::VB);
  d.u_T = P(int*, u_T);
  d.color_cap = c->tune.color_lds_contacts > 0 ? c->tune.color_lds_contacts : UNIT_CAP;
  d.color_bits = c->tune.color_body_bits > 0 ? std::min(c->tune.color_body_bits, BODY_BITS) : BODY_BITS;
  d.isl_unit = P(int*, isl_unit); d.u_nbody = P(int*, u_nbody); d.u_boff = P(int*, u_boff);
  d.f_loc = P(int*, f_loc); d.f_unit = P(int*, f_unit); d.u_blist = P(int*, u_blist);
  d.s_loc = P(int2*, s_loc); d.s_W = P(V4<R>*, s_W);
  d.s_JL = P(V4<R>*, s_JL); d.s_JA = P(V4<R>*, s_JA); d.s_JB = P(V4<R>*, s_JB);
  d.xr0 = P(R*, xr0); d.xr1 = P(R*, xr1);
  d.res_cap_c = c->res_cap_c; d.res_cap_b = c->res_cap_b; d.res_cap_o = c->res_cap_o;
  d.l_cnt = P(int*, l_cnt); d.l_base = P(int*, l_base); d.n_lanes = P(int*, n_lanes);
  d.l_start = P(int*, l_start); d.l_end = P(int*, l_end); d.l_ncol = P(int*, l_ncol); d.l_bs = P(int*, l_bs);
  d.u_start = P(int*, u_start); d.u_end = P(int*, u_end); d.u_ncol = P(int*, u_ncol);
  d.u_bs = P(int*, u_bs); d.body_claim = P(int*, body_claim); d.body_loc = P(int2*, body_loc);
  d.bc_cnt = P(int*, bc_cnt); d.bc_off = P(int*, bc_off); d.bc_fill = P(int*, bc_fill);
  d.bc_list = P(int*, bc_list); d.pal = P(unsigned*, pal); d.g_pc = P(int*, g_pc);
  d.g_done = P(unsigned char*, g_done); d.col_off = P(int*, col_off); d.merit = P(unsigned long long*, merit);
  d.l_pos = P(int*, l_pos); d.big_list = P(int*, big_list);
  d.item_pair_of = P(int*, item_pair_of); d.cap_items = cap_items(c);
  d.perm = P(int*, perm); d.s_id = P(int2*, s_id); d.s_n = P(V4<R>*, s_n); d.s_rA = P(V4<R>*, s_rA);
  d.s_rB = P(V4<R>*, s_rB); d.s_b = P(V4<R>*, s_b); d.s_r0 = P(V4<R>*, s_r0); d.x0 = P(V4<R>*, x0);
  d.x1 = P(V4<R>*, x1);
  d.h_ids = P(const int4*, h_ids); d.h_geo = P(const V4<R>*, h_geo); d.hb_body = P(const int*, hb_body);
  d.hb_inc_off = P(const int*, hb_inc_off); d.hb_inc = P(const int*, hb_inc);
  d.h_J = P(V4<R>*, h_J); d.h_WJ = P(V4<R>*, h_WJ); d.h_b = P(R*, h_b); d.h_r0 = P(R*, h_r0);
  d.h_mer = P(double*, d_hmer);
  d.hd0 = P(V4<R>*, hd0); d.hd1 = P(V4<R>*, hd1);
  d.err = P(int*, errbuf);
  d.dt = dt;
  d.flags = flags;
#undef P
  return d;
}

// Look-back state of the step's multi-tile scans: SCAN_SITES regions (one per
// scan of the launch sequence, which has at most 11), each a ticket word and one status
// word per look-back tile of the largest capacity.
constexpr int SCAN_SITES = 16;
template <typename T> int scan_tiles(int cap) { return (cap + LbTile<T>::TILE - 1) / LbTile<T>::TILE + 1; }

int size_workspace(rbh_ctx* c) {
  const size_t V = 4 * (size_t)c->prec;
  const size_t cp = c->cap_pairs, cs = c->cap_slots, cc = c->cap_cand, ck = c->cap_contacts;
  const size_t nb = std::max(c->nb, 1), ns = std::max(c->ns, 1);
  c->cap_units = (int)std::min<long long>((long long)(c->cap_contacts + c->unit_T - 1) / c->unit_T + c->ns + 1, 1 << 20);
  const size_t nu = c->cap_units, nl = (size_t)c->nb + c->ns + 2;
  int maxcap = std::max({2 * c->cap_pairs + 2, c->cap_slots, c->cap_cand, c->cap_contacts, c->nb + c->ns + 2, c->ns,
                         c->cap_units});
  struct E { DBuf* b; size_t bytes; } el[] = {
      {&c->row_cnt, nb * 4}, {&c->row_off, (nb + 1) * 4},
      {&c->pairs, cp * 8}, {&c->pair_sz, cp * 4}, {&c->slot_off, (cp + 1) * 4}, {&c->n_pairs, 16},
      {&c->pair_items, cp * 4}, {&c->item_off, (cp + 1) * 4},
      {&c->slot_flag, cs + 16}, {&c->n_slots, 16},
      {&c->side_cnt, (2 * cp + 2) * 4}, {&c->side_off, (2 * cp + 2) * 4}, {&c->n_sides, 16},
      {&c->cand_pair, cc * 4}, {&c->cand_tri, cc * 4}, {&c->pair_cstart, (cp + 1) * 4}, {&c->n_cand, 16},
      {&c->cand_p, cc * V}, {&c->cand_n, cc * V}, {&c->cand_keep, cc + 16}, {&c->keep_scan, (cc + 1) * 4},
      {&c->kidx, cc * 4}, {&c->kpos, cc * V}, {&c->n_kept, 16}, {&c->kept_flag, cc + 16}, {&c->kept_scan, (cc + 1) * 4},
      {&c->c_ia, ck * 4}, {&c->c_ib, ck * 4}, {&c->c_tri, ck * 4}, {&c->c_pair, ck * 4},
      {&c->c_p, ck * V}, {&c->c_n, ck * V}, {&c->n_contacts, 16}, {&c->c_rA, ck * V}, {&c->c_rB, ck * V},
      {&c->c_b, ck * V}, {&c->c_r0, ck * V}, {&c->c_mu, ck * V}, {&c->c_imp, ck * V}, {&c->c_color, ck * 4},
      {&c->f_bs, nb * 4}, {&c->uf, nb * 4}, {&c->isl_cnt, nl * 4}, {&c->isl_off, (nl + 1) * 4},
      {&c->isl_fill, nl * 4}, {&c->isl_max, nl * 4}, {&c->c_lab, ck * 4}, {&c->uperm, ck * 4},
      {&c->u_cnt, ns * 4}, {&c->u_base, (ns + 1) * 4}, {&c->n_units, 16}, {&c->u_start, nu * 4},
      {&c->u_end, nu * 4}, {&c->u_ncol, nu * 4}, {&c->u_bs, nu * 4}, {&c->body_claim, nb * 4},
      {&c->body_loc, nb * 8}, {&c->bc_cnt, nb * 4}, {&c->bc_off, (nb + 1) * 4}, {&c->bc_fill, nb * 4},
      {&c->bc_list, 2 * ck * 4}, {&c->pal, ck * PALW * 4}, {&c->g_pc, ck * 4}, {&c->g_done, ck + 16},
      {&c->col_off, (2 * (ck + nu) + 16) * 4}, {&c->merit, (size_t)std::max(c->max_iter_batch, 1) * 8},
      {&c->l_pos, ck * 4}, {&c->big_list, (nu + 16) * 4},
      {&c->item_pair_of, (size_t)cap_items(c) * 4},
      {&c->u_T, ns * 4}, {&c->isl_unit, nl * 4}, {&c->u_nbody, nu * 4}, {&c->u_boff, (nu + 1) * 4},
      {&c->f_loc, nb * 4}, {&c->f_unit, nb * 4}, {&c->u_blist, nb * 4}, {&c->s_loc, ck * 8},
      {&c->s_W, 4 * ck * V}, {&c->l_cnt, ns * 4}, {&c->l_base, (ns + 1) * 4}, {&c->n_lanes, 16},
      {&c->l_start, nu * 4}, {&c->l_end, nu * 4}, {&c->l_ncol, nu * 4}, {&c->l_bs, nu * 4},
      {&c->s_JL, 4 * ck * V}, {&c->s_JA, 4 * ck * V}, {&c->s_JB, 4 * ck * V}, {&c->xr0, 4 * ck * c->prec},
      {&c->xr1, 4 * ck * c->prec},
      {&c->perm, ck * 4}, {&c->s_id, ck * 8}, {&c->s_n, ck * V}, {&c->s_rA, ck * V}, {&c->s_rB, ck * V},
      {&c->s_b, ck * V}, {&c->s_r0, ck * V}, {&c->x0, ck * V}, {&c->x1, ck * V}, {&c->errbuf, 64}};
  for (auto& e : el) {
    int rc = ensure(c, *e.b, e.bytes);
    if (rc) return rc;
  }
  c->scan_stride = 1 + std::max(scan_tiles<int>(maxcap), scan_tiles<unsigned char>(maxcap));
  return ensure(c, c->scan_state, (size_t)SCAN_SITES * c->scan_stride * 8);
}

int build_batch(rbh_ctx* c, const int* ids, int n) {
  std::vector<int> want(ids, ids + n);
  if (want == c->batch && c->batch_arena_gen == c->arena_gen) return 0;
  drop_graph(c);
  std::vector<int> bb, bsslot, bsfoff, fend, fbs;
  std::vector<int> b2f(std::max(c->bodies.cap, 1), -1);
  long long pairs_all = 0, tri_sum = 0;
  int max_iter = 1;
  bool serial = false;
  for (int i = 0; i < n; ++i) {
    const SceneHost& s = c->scenes[want[i]];
    bsslot.push_back(want[i]);
    bsfoff.push_back((int)bb.size());
    for (int b = 0; b < s.nbody; ++b) {
      b2f[s.body_off + b] = (int)bb.size();
      bb.push_back(s.body_off + b);
      fend.push_back(s.body_off + s.nbody);
      fbs.push_back(i);
    }
    pairs_all += (long long)s.nbody * (s.nbody - 1) / 2;
    tri_sum += s.ntri;
    max_iter = std::max(max_iter, s.params.max_iterations);
    serial = serial || s.params.proximal_solver == RBH_SCHEME_GAUSS_SEIDEL;
  }
  bsfoff.push_back((int)bb.size());
  c->nb = (int)bb.size();
  c->ns = n;
  std::vector<int> nbv = {c->nb, c->nb + c->ns + 1, c->ns, 0};
  int rc = upload(c, c->bb_body, bb);
  if (!rc) rc = upload(c, c->bs_slot, bsslot);
  // the batched solve takes the largest scenes first: a workgroup per scene,
  // so a large scene started last would set the step's length alone
  std::vector<int> order(n);
  for (int i = 0; i < n; ++i) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
    return c->scenes[want[a]].nbody > c->scenes[want[b]].nbody;
  });
  if (!rc) rc = upload(c, c->bs_order, order);
  if (!rc) rc = upload(c, c->bs_foff, bsfoff);
  if (!rc) rc = upload(c, c->b2f, b2f);
  if (!rc) rc = upload(c, c->f_end, fend);
  if (!rc) rc = upload(c, c->f_bs, fbs);
  if (!rc) rc = upload(c, c->d_nb, nbv);
  if (rc) return rc;
  // one scene: colour small units (islands) and solve them on many
  // workgroups; batches: one workgroup per scene
  c->max_iter_batch = max_iter;
  c->multi_ok = (n == 1 && !serial);
  c->unit_T = c->tune.unit_size > 0 ? c->tune.unit_size : 64;
  c->last_units = 0;
  // initial capacities; they grow when a step reports an overflow
  long long est_pairs = std::min<long long>(pairs_all, 64LL * c->nb + 4096);
  c->cap_pairs = std::max(c->cap_pairs, (int)std::max<long long>(est_pairs, 1));
  c->cap_slots = std::max(c->cap_slots, (int)std::min<long long>(8 * tri_sum + 65536, 1LL << 30));
  c->cap_cand = std::max(c->cap_cand, c->cap_slots / 2);
  c->cap_contacts = std::max(c->cap_contacts, (int)std::min<long long>(64LL * c->nb + 65536, c->cap_cand));
  rc = size_workspace(c);
  if (rc) return rc;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->batch = want;
  c->batch_arena_gen = c->arena_gen;
  return 0;
}

// ------------------------------------------------------------ launch sequence
__global__ void k_fill(int* p, int n, int v) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = v;
}

int grid_for(long n, int block, int maxg = 8192) {
  long g = (n + block - 1) / block;
  return (int)std::max<long>(1, std::min<long>(g, maxg));
}

// Exclusive prefix sum of in[0..n) with n = *n_dev clamped to cap (rbh_scan.h):
// one workgroup for a short capacity, else the look-back scan on `state`
// (1 + scan_tiles<T>(cap) words, zero before the launch) with stalls flagged in err[0].
template <typename T>
void launch_scan(hipStream_t st, const T* in, int* out, const int* n_dev, int cap, int* total,
                 unsigned long long* state, int* err) {
  if (cap < SCAN_SINGLE_TILES * SCAN_TILE) {
    hipLaunchKernelGGL(k_scan_single<T>, dim3(1), dim3(SCAN_BLOCK), 0, st, in, n_dev, cap, out, total);
    return;
  }
  hipLaunchKernelGGL(k_scan_lookback<T>, dim3(scan_tiles<T>(cap)), dim3(SCAN_BLOCK), 0, st, in, n_dev, cap, out, total,
                     state, err);
}

// A scan of the step.  Each multi-tile scan takes the next region of the
// step's look-back state; false if the step has more such scans than
// SCAN_SITES (nothing launched).
template <typename T> bool scan(rbh_ctx* c, const T* in, int* out, const int* n_dev, int cap, int* total) {
  unsigned long long* st = nullptr;
  if (cap >= SCAN_SINGLE_TILES * SCAN_TILE) {
    const int site = c->scan_site++;
    if (site >= SCAN_SITES || 1 + scan_tiles<T>(cap) > c->scan_stride) return false;
    st = reinterpret_cast<unsigned long long*>(c->scan_state.p) + (size_t)site * c->scan_stride;
  }
  launch_scan<T>(c->stream, in, out, n_dev, cap, total, st, reinterpret_cast<int*>(c->errbuf.p));
  return true;
}

// steps between re-sizings of a single scene's cooperative solve grid
constexpr int GRID_REFRESH = 16;

// Events of step i of a call of `todo` steps: the last TIMED_STEPS steps are
// timed, each on its own row of ST_N events (created on first use).
constexpr int TIMED_STEPS = 512;
hipEvent_t* timing_row(rbh_ctx* c, int i, int todo) {
  const int first = std::max(0, todo - TIMED_STEPS);
  if (i < first) return nullptr;
  const size_t need = (size_t)std::min(todo, TIMED_STEPS) * ST_N;
  while (c->evp.size() < need) {
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    c->evp.push_back(e);
  }
  return &c->evp[(size_t)(i - first) * ST_N];
}

// ev: ST_N events to record the stage boundaries into, or nullptr.
// The hinge problem around the contact solve (rbh_hinge.h): rows and the
// recording pass before it, the final pass after it.
template <typename R> void hinge_pass(rbh_ctx* c, const Dev<R>& d, int pass) {
  if (!c->any_hinge) return;
  if (pass == 0) hipLaunchKernelGGL(k_hinge_rows<R>, dim3(c->ns), dim3(64), 0, c->stream, d);
  hipLaunchKernelGGL(k_hinge_solve<R>, dim3(c->ns), dim3(64), c->hinge_lds, c->stream, d, pass);
}

template <typename R> hipError_t launch_step(rbh_ctx* c, double dt, int flags, hipEvent_t* ev, bool multi, int multi_grid) {
  Dev<R> d = make_dev<R>(c, dt, flags, multi);
  hipStream_t s = c->stream;
  const int nb = c->nb, ns = c->ns;
  const long merit_slots = (long)std::max(c->max_iter_batch, 1) * multi_grid;
  auto mark = [&](int st) { if (ev) (void)hipEventRecord(ev[st], s); };
  mark(ST_BEGIN);
  const int scan_words = SCAN_SITES * c->scan_stride;
  const int n_merit = multi ? (int)(2 * merit_slots) : 0;
  const long begin_n =
      std::max<long>({(long)scan_words, 64L * ns, (long)nb + ns + 1, (long)c->cap_units, (long)n_merit});
  hipLaunchKernelGGL(k_step_begin<R>, dim3(grid_for(begin_n, 256, 1024)), dim3(256), 0, s, d,
                     reinterpret_cast<unsigned long long*>(c->scan_state.p), scan_words,
                     reinterpret_cast<int*>(d.merit), n_merit);
  c->scan_site = 0;
  // a scan that could not be launched (more multi-tile scans than SCAN_SITES)
  // ends the sequence before anything reads its output
#define RBH_SCAN(T, ...) \
  if (!scan<T>(__VA_ARGS__)) return hipErrorInvalidConfiguration
  hipLaunchKernelGGL(k_body_prep<R>, dim3(grid_for(nb, 256, 1 << 20)), dim3(256), 0, s, d);
  hipLaunchKernelGGL(k_refit<R>, dim3(nb), dim3(64), 0, s, d);
  mark(ST_PREP);
  // body pairs: count per row, scan, write
  hipLaunchKernelGGL(k_pair_count<R>, dim3(nb), dim3(64), 0, s, d);
  RBH_SCAN(int, c, d.row_cnt, d.row_off, d.d_nb, nb, d.n_pairs);
  hipLaunchKernelGGL(k_pair_write<R>, dim3(nb), dim3(64), 0, s, d);
  // candidate slots: (nt_a + nt_b) per pair; work items: triangle blocks of both sides
  RBH_SCAN(int, c, d.pair_sz, d.slot_off, d.n_pairs, c->cap_pairs, d.n_slots);
  RBH_SCAN(int, c, d.pair_items, d.item_off, d.n_pairs, c->cap_pairs, nullptr);
  hipLaunchKernelGGL(k_item_map<R>, dim3(grid_for(c->cap_pairs, 256)), dim3(256), 0, s, d);
  hipLaunchKernelGGL(k_cand_flags<R>, dim3(grid_for(c->cap_pairs, 1, 16384)), dim3(CAND_BLOCK), 0, s, d);
  RBH_SCAN(int, c, d.side_cnt, d.side_off, d.n_sides, 2 * c->cap_pairs, d.n_cand);
  hipLaunchKernelGGL(k_cand_write<R>, dim3(grid_for(2L * c->cap_pairs * 64, 256)), dim3(256), 0, s, d);
  mark(ST_BROAD);
  hipLaunchKernelGGL(k_frank_wolfe<R>, dim3(grid_for(c->cap_cand, 256)), dim3(256), 0, s, d);
  mark(ST_FWK);
  RBH_SCAN(unsigned char, c, d.cand_keep, d.keep_scan, d.n_cand, c->cap_cand, nullptr);
  hipLaunchKernelGGL(k_kept_list<R>, dim3(grid_for(c->cap_cand, 256)), dim3(256), 0, s, d);
  hipLaunchKernelGGL(k_dedup<R>, dim3(grid_for(c->cap_cand, 256)), dim3(256), 0, s, d);
  RBH_SCAN(unsigned char, c, d.kept_flag, d.kept_scan, d.n_kept, c->cap_cand, nullptr);
  hipLaunchKernelGGL(k_contact_write<R>, dim3(grid_for(c->cap_cand, 256)), dim3(256), 0, s, d);
  mark(ST_FW);
  hipLaunchKernelGGL(k_assemble<R>, dim3(grid_for(c->cap_contacts, 256)), dim3(256), 0, s, d);
  mark(ST_ASM);
  // islands -> colouring units -> body lists in unit order
  const int G = grid_for(c->cap_contacts, 256);
  hipLaunchKernelGGL(k_uf_union<R>, dim3(G), dim3(256), 0, s, d);
  hipLaunchKernelGGL(k_isl_count<R>, dim3(G), dim3(256), 0, s, d);
  RBH_SCAN(int, c, d.isl_cnt, d.isl_off, d.d_nb + 1, nb + ns + 1, nullptr);
  hipLaunchKernelGGL(k_isl_fill<R>, dim3(G), dim3(256), 0, s, d);
  if (ns <= UNIT_SCAN_MAX) {
    hipLaunchKernelGGL(k_unit_count_scan<R>, dim3(1), dim3(SCAN_BLOCK), 0, s, d);
  } else {
    hipLaunchKernelGGL(k_unit_count<R>, dim3(grid_for(ns, 256)), dim3(256), 0, s, d);
    RBH_SCAN(int, c, d.u_cnt, d.u_base, d.d_nb + 2, ns, d.n_units);
    RBH_SCAN(int, c, d.l_cnt, d.l_base, d.d_nb + 2, ns, d.n_lanes);
  }
  hipLaunchKernelGGL(k_unit_assign<R>, dim3(grid_for(nb + ns, 256)), dim3(256), 0, s, d);
  hipLaunchKernelGGL(k_csr_count<R>, dim3(G), dim3(256), 0, s, d);
  RBH_SCAN(int, c, d.bc_cnt, d.bc_off, d.d_nb, nb, nullptr);
  hipLaunchKernelGGL(k_csr_fill<R>, dim3(G), dim3(256), 0, s, d);
  // free bodies -> slots of their unit
  hipLaunchKernelGGL(k_unit_bodies<R>, dim3(grid_for(nb, 256)), dim3(256), 0, s, d);
  RBH_SCAN(int, c, d.u_nbody, d.u_boff, d.n_units, c->cap_units, nullptr);
  hipLaunchKernelGGL(k_unit_blist<R>, dim3(grid_for(nb, 256)), dim3(256), 0, s, d);
#undef RBH_SCAN
  hipLaunchKernelGGL(k_color_small<R>, dim3(std::min(c->cap_units, 8192)), dim3(64), 0, s, d);
  hipLaunchKernelGGL(k_color_units<R>, dim3(std::min(c->cap_units, c->prop.multiProcessorCount)),
                     dim3(COLOR_UNIT_BLOCK), 0, s, d);
  hipLaunchKernelGGL(k_order_lanes<R>, dim3(std::min(c->cap_units, 4096)), dim3(COLOR_BLOCK), 0, s, d);
  hipLaunchKernelGGL(k_lane_gather<R>, dim3(std::min(c->cap_units, 2048), GATHER_SPLIT), dim3(COLOR_BLOCK), 0, s, d);
  if (multi) {
    mark(ST_COLOR);
    hinge_pass<R>(c, d, 0);
    void* args[] = {&d};
    hipError_t e = hipLaunchCooperativeKernel(reinterpret_cast<void*>(k_solve_multi<R>), dim3(multi_grid),
                                              dim3(SOLVE_BLOCK), args, c->solve_lds, s);
    if (e != hipSuccess) return e;
  } else {
    mark(ST_COLOR);
    hinge_pass<R>(c, d, 0);
    launch_batch_solve<R>(c, d, ns, s);
  }
  hinge_pass<R>(c, d, 1);
  mark(ST_SOLVE);
  hipLaunchKernelGGL(k_integrate<R>, dim3(grid_for(nb, 256, 1 << 20)), dim3(256), 0, s, d);
  mark(ST_INTG);
  if (c->any_post) {
    // post-stabilisation (steppers.py:90-91): the same solve on the rewritten
    // rows, then a position update with dt = 1
    hipLaunchKernelGGL(k_post_rows<R>, dim3(grid_for(c->cap_contacts, 256)), dim3(256), 0, s, d);
    Dev<R> dp = d;
    dp.post = 1;
    if (multi) {
      hipLaunchKernelGGL(k_fill, dim3(grid_for(2L * merit_slots, 256)), dim3(256), 0, s,
                         reinterpret_cast<int*>(d.merit), 2 * merit_slots, -1);
      void* args[] = {&dp};
      hipError_t e = hipLaunchCooperativeKernel(reinterpret_cast<void*>(k_solve_multi<R>), dim3(multi_grid),
                                                dim3(SOLVE_BLOCK), args, c->solve_lds, s);
      if (e != hipSuccess) return e;
    } else {
      launch_batch_solve<R>(c, dp, ns, s);
    }
    hipLaunchKernelGGL(k_post_integrate<R>, dim3(grid_for(nb, 256, 1 << 20)), dim3(256), 0, s, d);
  }
  mark(ST_POST);
  hipLaunchKernelGGL(k_scene_finalize<R>, dim3(grid_for(ns * 64, 256)), dim3(256), 0, s, d);
  mark(ST_INT);
  return hipGetLastError();
}

int read_err(rbh_ctx* c, int e[8]) {
  HIPCHK(c, hipMemcpyAsync(e, c->errbuf.p, 8 * sizeof(int), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return 0;
}

int first_step_index(rbh_ctx* c, int slot, int* out) {
  SceneState st;
  HIPCHK(c, hipMemcpyAsync(&st, reinterpret_cast<SceneState*>(c->d_ss.p) + slot, sizeof(SceneState),
                           hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  *out = st.step_index;
  return 0;
}

// Dynamic LDS of the solve kernels: the resident layout (FP32: 384 contacts,
// 384 bodies, ~130 KB; FP64: half) or the velocity-only layout,
// whichever is larger; the kernels' limit is raised to it.
template <typename R> int setup_solve_lds(rbh_ctx* c) {
  c->res_cap_c = c->tune.resident_contacts > 0 ? std::min(c->tune.resident_contacts, ResCap<R>::C) : ResCap<R>::C;
  c->res_cap_b = c->tune.resident_bodies > 0 ? std::min(c->tune.resident_bodies, ResCap<R>::B) : ResCap<R>::B;
  c->res_cap_o = ResCap<R>::O;
  const size_t bytes = std::max<size_t>(SOLVE_LDS_BYTES, ResCap<R>::BYTES);
  const size_t half = std::max<size_t>(SOLVE_LDS_BYTES, ResCap<R, 1>::BYTES);
  const void* ks[5] = {reinterpret_cast<const void*>(k_solve<R, SOLVE_BLOCK, 0>),
                       reinterpret_cast<const void*>(k_solve<R, 512, 0>),
                       reinterpret_cast<const void*>(k_solve_multi<R>),
                       reinterpret_cast<const void*>(k_solve<R, SOLVE_BLOCK, 1>),
                       reinterpret_cast<const void*>(k_solve<R, 512, 1>)};
  for (int i = 0; i < 5; ++i) {
    const size_t b = i < 3 ? bytes : half;
    if (hipFuncSetAttribute(ks[i], hipFuncAttributeMaxDynamicSharedMemorySize, (int)b) != hipSuccess) {
      (void)hipGetLastError();
      return fail(c, RBH_E_HIP, "solve kernels cannot use " + std::to_string(b) + " bytes of LDS");
    }
  }
  c->solve_lds_half = half;
  const int hbytes = (int)hinge_lds_bytes<R>(HINGE_BODY_MAX, HINGE_MAX);
  if (hipFuncSetAttribute(reinterpret_cast<const void*>(k_hinge_solve<R>), hipFuncAttributeMaxDynamicSharedMemorySize,
                          hbytes) != hipSuccess) {
    (void)hipGetLastError();
    return fail(c, RBH_E_HIP, "hinge solve cannot use " + std::to_string(hbytes) + " bytes of LDS");
  }
  c->solve_lds = bytes;
  c->coop_max = 0;
  return 0;
}

template <typename R> int run_steps(rbh_ctx* c, int n_steps, double dt, int flags) {
  c->any_post = false;
  c->any_hinge = false;
  c->hinge_lds = 16;
  for (int slot : c->batch) {
    const SceneParams& sp = c->sp_host[slot];
    c->any_post = c->any_post || c->scenes[slot].params.use_post_stabilization != 0;
    c->any_hinge = c->any_hinge || sp.h_cnt > 0;
    c->hinge_lds = std::max(c->hinge_lds, hinge_lds_bytes<R>(sp.hb_cnt, sp.h_cnt));
  }
  const bool timing = (flags & RBH_STEP_TIMING) != 0;
  const bool use_graph = (flags & RBH_STEP_GRAPH) != 0;
  const int kflags = flags & RBH_STEP_COUNT_OVERLAPS;
  if (c->solve_lds == 0) {
    int rc = setup_solve_lds<R>(c);
    if (rc) return rc;
  }
  if (c->coop_max == 0) {
    int per_cu = 0;
    HIPCHK(c, hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void*>(k_solve_multi<R>),
                                                           SOLVE_BLOCK, c->solve_lds));
    c->coop_max = std::max(1, per_cu) * c->prop.multiProcessorCount;
  }
  if (c->multi_ok) {
    // merit slots of the cooperative solve: one per sweep and workgroup
    int rc = ensure(c, c->merit, (size_t)std::max(c->max_iter_batch, 1) * c->coop_max * 8);
    if (rc) return rc;
  }
  bool multi = c->multi_ok && !use_graph;
  if (flags & RBH_STEP_FORCE_SINGLE) multi = false;
  if ((flags & RBH_STEP_FORCE_MULTI) && c->multi_ok && !use_graph) multi = true;
  // cooperative grid: about one unit per wavefront, from the unit count of the last step seen
  const int lpw = c->tune.lanes_per_group > 0 ? c->tune.lanes_per_group : SOLVE_BLOCK / 64;
  auto grid_of = [&](int units) { return std::max(1, std::min((units + lpw - 1) / lpw, c->coop_max)); };
  int multi_grid = grid_of(c->last_units > 1 ? c->last_units : 64);
  int done = 0;
  for (int attempt = 0; attempt < 16 && done < n_steps; ++attempt) {
    int si0 = 0;
    int rc = first_step_index(c, c->batch[0], &si0);
    if (rc) return rc;
    const int todo = n_steps - done;
    if (use_graph) {
      if (!c->gexec || c->graph_key_gen != c->arena_gen || c->graph_dt != dt || c->graph_flags != kflags) {
        drop_graph(c);
        HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        hipError_t le = launch_step<R>(c, dt, kflags, nullptr, false, 1);
        if (le != hipSuccess) {
          (void)hipStreamEndCapture(c->stream, &c->graph);
          return fail(c, RBH_E_HIP, std::string("step launch: ") + hipGetErrorString(le));
        }
        HIPCHK(c, hipStreamEndCapture(c->stream, &c->graph));
        HIPCHK(c, hipGraphInstantiate(&c->gexec, c->graph, nullptr, nullptr, 0));
        c->graph_key_gen = c->arena_gen;
        c->graph_dt = dt;
        c->graph_flags = kflags;
      }
      for (int i = 0; i < todo; ++i) {
        hipEvent_t* ev = timing ? timing_row(c, i, todo) : nullptr;
        if (ev) HIPCHK(c, hipEventRecord(ev[ST_BEGIN], c->stream));
        HIPCHK(c, hipGraphLaunch(c->gexec, c->stream));
        if (ev) HIPCHK(c, hipEventRecord(ev[ST_INT], c->stream));
      }
    } else {
      for (int i = 0; i < todo; ++i) {
        hipError_t le = launch_step<R>(c, dt, kflags, timing ? timing_row(c, i, todo) : nullptr, multi, multi_grid);
        if (le != hipSuccess) return fail(c, RBH_E_HIP, std::string("step launch: ") + hipGetErrorString(le));
        if (multi && (i + 1) % GRID_REFRESH == 0 && i + 1 < todo) {
          // re-size the cooperative grid to the scene's current unit count
          int nu = 0;
          HIPCHK(c, hipMemcpyAsync(&nu, c->n_units.p, 4, hipMemcpyDeviceToHost, c->stream));
          HIPCHK(c, hipStreamSynchronize(c->stream));
          if (nu > 0) multi_grid = grid_of(nu);
        }
      }
    }
    c->timed_steps = timing ? std::min(std::min(todo, TIMED_STEPS), (int)(c->evp.size() / ST_N)) : 0;
    int e[8];
    rc = read_err(c, e);
    if (rc) return rc;
    int si1 = 0;
    rc = first_step_index(c, c->batch[0], &si1);
    if (rc) return rc;
    done += si1 - si0;
    c->ws_gen++;
    if (e[0] == 0) break;
    if (e[0] & ERR_COLORS)
      return fail(c, RBH_E_COLORS, "colouring needs more than " + std::to_string(PALMAX) + " colours");
    // grow what overflowed and redo the remaining steps
    auto grow = [](int cur, int need) { return std::max(cur * 2, (int)std::min<long long>(need * 3LL / 2 + 1024, 1LL << 30)); };
    if (e[0] & ERR_CAP_PAIRS) c->cap_pairs = grow(c->cap_pairs, e[1]);
    if (e[0] & ERR_CAP_SLOTS) c->cap_slots = grow(c->cap_slots, e[2]);
    if (e[0] & ERR_CAP_CAND) c->cap_cand = grow(c->cap_cand, e[3]);
    if (e[0] & ERR_CAP_CONTACTS) c->cap_contacts = grow(c->cap_contacts, e[4]);
    if (e[0] & ERR_CAP_UNITS) c->cap_contacts = grow(c->cap_contacts, e[5] * c->unit_T);
    if (e[0] & ERR_SOLVER) return fail(c, RBH_E_HIP, "grid barrier of the solver timed out");
    if (e[0] & ERR_SCAN) return fail(c, RBH_E_HIP, "prefix sum: a tile's wait for its predecessors timed out");
    c->cap_cand = std::max(c->cap_cand, std::min(c->cap_slots, c->cap_cand));
    drop_graph(c);
    c->arena_gen++;
    c->batch_arena_gen = c->arena_gen;
    rc = size_workspace(c);
    if (rc) return rc;
  }
  if (done < n_steps) return fail(c, RBH_E_CAPACITY, "step workspace could not grow enough");
  int nu = 0;
  HIPCHK(c, hipMemcpyAsync(&nu, c->n_units.p, 4, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->last_units = std::max(1, nu);
  // batch counters
  int h[4] = {0, 0, 0, 0};
  HIPCHK(c, hipMemcpyAsync(&h[0], c->n_pairs.p, 4, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipMemcpyAsync(&h[1], c->n_slots.p, 4, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipMemcpyAsync(&h[2], c->n_cand.p, 4, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipMemcpyAsync(&h[3], c->n_contacts.p, 4, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  for (int i = 0; i < 4; ++i) c->counters[i] = h[i];
  if (timing) {
    // average over the timed steps of the last attempt; a graph replay has no stage marks
    for (int i = 0; i < TIMING_OUT; ++i) c->timings[i] = 0.0;
    // stage interval -> rbh_get_timings entry
    static const int out_of[ST_N] = {-1, 0, 1, 2, 8, 3, 4, 5, 6, 9, 6};
    for (int k = 0; k < c->timed_steps; ++k) {
      hipEvent_t* ev = &c->evp[(size_t)k * ST_N];
      if (!use_graph)
        for (int i = 1; i < ST_N; ++i) {
          float ms = 0;
          HIPCHK(c, hipEventElapsedTime(&ms, ev[i - 1], ev[i]));
          c->timings[out_of[i]] += ms;
        }
      float tot = 0;
      HIPCHK(c, hipEventElapsedTime(&tot, ev[ST_BEGIN], ev[ST_INT]));
      c->timings[7] += tot;
    }
    for (int i = 0; i < TIMING_OUT; ++i) c->timings[i] /= std::max(1, c->timed_steps);
  }
  for (size_t i = 0; i < c->batch.size(); ++i) {
    c->scenes[c->batch[i]].ws_gen = c->ws_gen;
    c->scenes[c->batch[i]].batch_index = (int)i;
  }
  return 0;
}

bool valid_scene(rbh_ctx* c, int id) { return id >= 0 && id < (int)c->scenes.size() && c->scenes[id].live; }

template <typename R> int set_state(rbh_ctx* c, const SceneHost& s, const double* r, const double* q,
                                    const double* v, const double* w) {
  const int n = s.nbody;
  std::vector<V4<R>> a(n);
  const double* src[4] = {r, q, v, w};
  DBuf* dst[4] = {&c->r, &c->q, &c->v, &c->w};
  for (int k = 0; k < 4; ++k) {
    if (!src[k]) continue;
    int dim = k == 1 ? 4 : 3;
    for (int i = 0; i < n; ++i) {
      const double* p = src[k] + (size_t)dim * i;
      a[i] = V4<R>{R(p[0]), R(p[1]), R(p[2]), dim == 4 ? R(p[3]) : R(0)};
    }
    HIPCHK(c, hipMemcpyAsync(reinterpret_cast<V4<R>*>(dst[k]->p) + s.body_off, a.data(), n * sizeof(V4<R>),
                             hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
  }
  return 0;
}

template <typename R> int get_state(rbh_ctx* c, const SceneHost& s, double* r, double* q, double* v, double* w) {
  const int n = s.nbody;
  std::vector<V4<R>> a(n);
  double* dstp[4] = {r, q, v, w};
  DBuf* src[4] = {&c->r, &c->q, &c->v, &c->w};
  for (int k = 0; k < 4; ++k) {
    if (!dstp[k]) continue;
    HIPCHK(c, hipMemcpyAsync(a.data(), reinterpret_cast<V4<R>*>(src[k]->p) + s.body_off, n * sizeof(V4<R>),
                             hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    int dim = k == 1 ? 4 : 3;
    for (int i = 0; i < n; ++i) {
      double* p = dstp[k] + (size_t)dim * i;
      p[0] = a[i].x; p[1] = a[i].y; p[2] = a[i].z;
      if (dim == 4) p[3] = a[i].w;
    }
  }
  return 0;
}

template <typename R> int scene_statics(rbh_ctx* c, const rbh_scene_desc* d, const SceneHost& s, int slot) {
  const int n = d->n_bodies;
  std::vector<int> sh(n), sl(n, slot), fx(n), mt(n), tr(n);
  std::vector<int4> bb(n);
  std::vector<R> ms(n);
  std::vector<V4<R>> in(n), gr(n), dp(n), z(n, V4<R>{R(0), R(0), R(0), R(0)});
  std::vector<V4<R>> qid(n, V4<R>{R(1), R(0), R(0), R(0)});
  int toff = s.tri_off;
  for (int i = 0; i < n; ++i) {
    sh[i] = d->shape[i];
    fx[i] = d->fixed ? (d->fixed[i] != 0) : 0;
    mt[i] = d->material ? d->material[i] : 0;
    tr[i] = toff;
    bb[i] = make_int4(toff, c->shapes[sh[i]].nt, c->shapes[sh[i]].t_off, slot);
    toff += c->shapes[sh[i]].nt;
    ms[i] = R(d->mass[i]);
    in[i] = V4<R>{R(d->inertia[3 * i]), R(d->inertia[3 * i + 1]), R(d->inertia[3 * i + 2]), R(0)};
    gr[i] = d->gravity ? V4<R>{R(d->gravity[3 * i]), R(d->gravity[3 * i + 1]), R(d->gravity[3 * i + 2]), R(0)} : z[i];
    dp[i] = d->damping ? V4<R>{R(d->damping[2 * i]), R(d->damping[2 * i + 1]), R(0), R(0)} : z[i];
  }
  auto put = [&](DBuf& b, const void* src, size_t el) -> int {
    HIPCHK(c, hipMemcpyAsync((char*)b.p + el * s.body_off, src, el * n, hipMemcpyHostToDevice, c->stream));
    return 0;
  };
  int rc = 0;
  rc |= put(c->b_shape, sh.data(), 4); rc |= put(c->b_slot, sl.data(), 4); rc |= put(c->b_fixed, fx.data(), 4);
  rc |= put(c->b_mat, mt.data(), 4); rc |= put(c->b_tri, tr.data(), 4); rc |= put(c->b_mass, ms.data(), sizeof(R));
  rc |= put(c->b_inertia, in.data(), sizeof(V4<R>)); rc |= put(c->b_grav, gr.data(), sizeof(V4<R>));
  rc |= put(c->b_damp, dp.data(), sizeof(V4<R>)); rc |= put(c->b_broad, bb.data(), sizeof(int4));
  rc |= put(c->r, z.data(), sizeof(V4<R>)); rc |= put(c->q, qid.data(), sizeof(V4<R>));
  rc |= put(c->v, z.data(), sizeof(V4<R>)); rc |= put(c->w, z.data(), sizeof(V4<R>));
  if (rc) return rc;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return 0;
}

template <typename R> int get_contacts(rbh_ctx* c, const SceneHost& s, const SceneState& st, int cap, int* count,
                                       int* ia, int* ib, int* tri, double* p, double* n, double* gap, double* imp,
                                       int* color) {
  const int K = st.k_cnt;
  *count = K;
  const int m = std::min(K, cap);
  if (m <= 0) return 0;
  const int k0 = st.k_off;
  std::vector<int> ti(m);
  std::vector<V4<R>> tv(m);
  auto geti = [&](DBuf& b, int* out, int sub) -> int {
    if (!out) return 0;
    HIPCHK(c, hipMemcpyAsync(ti.data(), (int*)b.p + k0, m * 4, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (int i = 0; i < m; ++i) out[i] = ti[i] - sub;
    return 0;
  };
  int rc = geti(c->c_ia, ia, s.body_off);
  if (!rc) rc = geti(c->c_ib, ib, s.body_off);
  if (!rc) rc = geti(c->c_tri, tri, 0);
  if (!rc) rc = geti(c->c_color, color, 0);
  if (rc) return rc;
  auto getv = [&](DBuf& b) -> int {
    HIPCHK(c, hipMemcpyAsync(tv.data(), (V4<R>*)b.p + k0, m * sizeof(V4<R>), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return 0;
  };
  if (p || gap) {
    if ((rc = getv(c->c_p))) return rc;
    for (int i = 0; i < m; ++i) {
      if (p) { p[3 * i] = tv[i].x; p[3 * i + 1] = tv[i].y; p[3 * i + 2] = tv[i].z; }
      if (gap) gap[i] = tv[i].w;
    }
  }
  if (n) {
    if ((rc = getv(c->c_n))) return rc;
    for (int i = 0; i < m; ++i) { n[3 * i] = tv[i].x; n[3 * i + 1] = tv[i].y; n[3 * i + 2] = tv[i].z; }
  }
  if (imp) {
    if ((rc = getv(c->c_imp))) return rc;
    for (int i = 0; i < m; ++i) {
      if (st.solved) { imp[4 * i] = tv[i].x; imp[4 * i + 1] = tv[i].y; imp[4 * i + 2] = tv[i].z; imp[4 * i + 3] = tv[i].w; }
      else { imp[4 * i] = imp[4 * i + 1] = imp[4 * i + 2] = imp[4 * i + 3] = 0.0; }
    }
  }
  return 0;
}

}  // namespace

// ================================================================ C ABI
extern "C" {

int rbh_version(void) { return 1; }

int rbh_ctx_create(int device, int precision_bytes, unsigned long long hbm_budget_bytes, rbh_ctx** out) {
  if (!out) return RBH_E_ARG;
  *out = nullptr;
  if (precision_bytes != 4 && precision_bytes != 8) return RBH_E_ARG;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return RBH_E_HIP;
  if (device < 0 || device >= ndev) return RBH_E_ARG;
  rbh_ctx* c = new rbh_ctx();
  c->device = device;
  c->prec = precision_bytes;
  c->budget = hbm_budget_bytes;
  if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
      hipGetDeviceProperties(&c->prop, device) != hipSuccess) {
    delete c;
    return RBH_E_HIP;
  }
  c->bodies.init(0);
  c->tris.init(0);
  c->res.init(0);
  c->mats.init(0);
  // the no-allocation entry points' staging array on this device (device current)
  c->scratch_err = scratch_address(device, &c->scratch);
  (void)hipGetLastError();
  *out = c;
  return 0;
}

int rbh_ctx_destroy(rbh_ctx* c) {
  if (!c) return RBH_E_ARG;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  drop_graph(c);
  free_all(c);
  for (hipEvent_t e : c->evp) (void)hipEventDestroy(e);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return 0;
}

const char* rbh_last_error(rbh_ctx* c) { return c ? c->err.c_str() : "null context"; }

int rbh_device_query(rbh_ctx* c, rbh_device_info* out) {
  if (!c || !out) return RBH_E_ARG;
  std::memset(out, 0, sizeof(*out));
  std::snprintf(out->name, sizeof(out->name), "%s (%s)", c->prop.name, c->prop.gcnArchName);
  out->compute_units = c->prop.multiProcessorCount;
  out->lds_per_block = (int)c->prop.sharedMemPerBlock;
  out->hbm_bytes = c->prop.totalGlobalMem;
  out->precision_bytes = c->prec;
  out->max_colors = PALMAX;
  return 0;
}

int rbh_memory_usage(rbh_ctx* c, unsigned long long* bytes) {
  if (!c || !bytes) return RBH_E_ARG;
  *bytes = total_bytes(c);
  return 0;
}

int rbh_prefix_sum(rbh_ctx* c, const void* in, int elem_bytes, int n, int cap, int* out, int* total) {
  if (!c || !in || !out || !total) return fail(c, RBH_E_ARG, "prefix sum: null argument");
  if (elem_bytes != 1 && elem_bytes != 4) return fail(c, RBH_E_ARG, "prefix sum: elements are 1 or 4 bytes");
  if (n < 0 || cap < 1 || n > cap || cap > (1 << 28)) return fail(c, RBH_E_ARG, "prefix sum: need 0 <= n <= cap <= 2^28");
  (void)hipSetDevice(c->device);
  const int tiles = elem_bytes == 1 ? scan_tiles<unsigned char>(cap) : scan_tiles<int>(cap);
  const size_t in_bytes = (size_t)cap * elem_bytes + 64, out_bytes = ((size_t)cap + 1) * 4;
  const size_t st_bytes = ((size_t)tiles + 1) * 8;
  // one allocation: input, output, state, then n, total and the error words
  const size_t o_out = (in_bytes + 255) & ~(size_t)255, o_st = o_out + ((out_bytes + 255) & ~(size_t)255),
               o_ints = o_st + ((st_bytes + 255) & ~(size_t)255), bytes = o_ints + 256;
  unsigned char* buf = nullptr;
  HIPCHK(c, hipMalloc(&buf, bytes));
  int* ints = reinterpret_cast<int*>(buf + o_ints);   // [0] n, [1] total, [16..31] error words
  hipError_t e = hipMemsetAsync(buf + o_st, 0, o_ints + 256 - o_st, c->stream);
  if (e == hipSuccess && n > 0) e = hipMemcpyAsync(buf, in, (size_t)n * elem_bytes, hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(ints, &n, 4, hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess) {
    unsigned long long* st = reinterpret_cast<unsigned long long*>(buf + o_st);
    int* dout = reinterpret_cast<int*>(buf + o_out);
    if (elem_bytes == 1)
      launch_scan<unsigned char>(c->stream, buf, dout, ints, cap, ints + 1, st, ints + 16);
    else
      launch_scan<int>(c->stream, reinterpret_cast<const int*>(buf), dout, ints, cap, ints + 1, st, ints + 16);
    e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(out, dout, ((size_t)n + 1) * 4, hipMemcpyDeviceToHost, c->stream);
  }
  int tail[2] = {0, 0};
  if (e == hipSuccess) e = hipMemcpyAsync(tail, ints + 1, 4, hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(tail + 1, ints + 16, 4, hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  (void)hipFree(buf);
  if (e != hipSuccess) return fail(c, RBH_E_HIP, std::string("prefix sum: ") + hipGetErrorString(e));
  if (tail[1] & ERR_SCAN) return fail(c, RBH_E_HIP, "prefix sum: a tile's wait for its predecessors timed out");
  *total = tail[0];
  return 0;
}

int rbh_sdf_build(rbh_ctx* c, const double* V, int nv, const int* T, int nt, const double* mn, const double* mx,
                  int I, int J, int K, double* values) {
  if (!c || !V || !T || !mn || !mx || !values || nv <= 0 || nt <= 0) return fail(c, RBH_E_ARG, "bad SDF arguments");
  if (I < 2 || J < 2 || K < 2) return fail(c, RBH_E_ARG, "grid needs at least 2 nodes per axis");
  for (int i = 0; i < 3; ++i) if (mn[i] > mx[i]) return fail(c, RBH_E_ARG, "min_coord > max_coord");
  for (int i = 0; i < 3 * nt; ++i) if (T[i] < 0 || T[i] >= nv) return fail(c, RBH_E_ARG, "triangle index out of range");
  (void)hipSetDevice(c->device);
  std::vector<double> tri((size_t)nt * 9);
  for (int t = 0; t < nt; ++t)
    for (int k = 0; k < 3; ++k)
      for (int a = 0; a < 3; ++a) tri[(size_t)t * 9 + 3 * k + a] = V[3 * T[3 * t + k] + a];
  const long n = (long)I * J * K;
  double *dtri = nullptr, *dout = nullptr;
  HIPCHK(c, hipMalloc(&dtri, tri.size() * 8));
  if (hipMalloc(&dout, n * 8) != hipSuccess) { (void)hipFree(dtri); return fail(c, RBH_E_MEMORY, "SDF buffer"); }
  HIPCHK(c, hipMemcpyAsync(dtri, tri.data(), tri.size() * 8, hipMemcpyHostToDevice, c->stream));
  double h[3] = {(mx[0] - mn[0]) / (I - 1), (mx[1] - mn[1]) / (J - 1), (mx[2] - mn[2]) / (K - 1)};
  hipLaunchKernelGGL(k_sdf, dim3((unsigned)((n + SDF_BLOCK - 1) / SDF_BLOCK)), dim3(SDF_BLOCK), 0, c->stream, dtri, nt,
                     I, J, K, mn[0], mn[1], mn[2], h[0], h[1], h[2], dout);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(values, dout, n * 8, hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  (void)hipFree(dtri);
  (void)hipFree(dout);
  if (e != hipSuccess) return fail(c, RBH_E_HIP, std::string("SDF kernel: ") + hipGetErrorString(e));
  return 0;
}

int rbh_sdf_build_noalloc(rbh_ctx* c, const double* V, int nv, const int* T, int nt, const double* mn,
                          const double* mx, int I, int J, int K, double* values) {
  if (!c || !V || !T || !mn || !mx || !values || nv <= 0 || nt <= 0) return fail(c, RBH_E_ARG, "bad SDF arguments");
  if (I < 2 || J < 2 || K < 2) return fail(c, RBH_E_ARG, "grid needs at least 2 nodes per axis");
  for (int i = 0; i < 3; ++i) if (mn[i] > mx[i]) return fail(c, RBH_E_ARG, "min_coord > max_coord");
  for (int i = 0; i < 3 * nt; ++i) if (T[i] < 0 || T[i] >= nv) return fail(c, RBH_E_ARG, "triangle index out of range");
  if (!c->scratch)
    return fail(c, RBH_E_HIP, std::string("no-allocation scratch unavailable: ") + hipGetErrorString(c->scratch_err));
  // staging layout, each part 256-byte aligned: vertices (3 nv doubles),
  // triangles (3 nt ints), gathered triangle corners (9 nt doubles), grid (I J K doubles)
  auto up = [](unsigned long long b) { return (b + 255ull) & ~255ull; };
  unsigned long long n = (unsigned long long)I * (unsigned long long)J;   // < 2^62
  n = n > NOALLOC_SCRATCH_BYTES ? NOALLOC_SCRATCH_BYTES : n * (unsigned long long)K;
  const unsigned long long oV = 0, oT = oV + up(24ull * (unsigned)nv), oTri = oT + up(12ull * (unsigned)nt),
                           oOut = oTri + up(72ull * (unsigned)nt), need = oOut + up(8ull * n);
  if (need > NOALLOC_SCRATCH_BYTES)
    return fail(c, RBH_E_CAPACITY, "SDF build needs " + std::to_string(need) + " bytes of staging; the no-allocation "
                "entry point holds " + std::to_string(NOALLOC_SCRATCH_BYTES) + " (use rbh_sdf_build)");
  HIPCHK(c, hipSetDevice(c->device));
  double* dV = reinterpret_cast<double*>(c->scratch + oV);
  int* dT = reinterpret_cast<int*>(c->scratch + oT);
  double* dtri = reinterpret_cast<double*>(c->scratch + oTri);
  double* dout = reinterpret_cast<double*>(c->scratch + oOut);
  HIPCHK(c, hipMemcpyAsync(dV, V, 24ull * (unsigned)nv, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(dT, T, 12ull * (unsigned)nt, hipMemcpyHostToDevice, c->stream));
  hipLaunchKernelGGL(k_sdf_gather, dim3(grid_for(9L * nt, 256)), dim3(256), 0, c->stream, dV, dT, nt, dtri);
  double h[3] = {(mx[0] - mn[0]) / (I - 1), (mx[1] - mn[1]) / (J - 1), (mx[2] - mn[2]) / (K - 1)};
  hipLaunchKernelGGL(k_sdf, dim3((unsigned)((n + SDF_BLOCK - 1) / SDF_BLOCK)), dim3(SDF_BLOCK), 0, c->stream, dtri, nt,
                     I, J, K, mn[0], mn[1], mn[2], h[0], h[1], h[2], dout);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(values, dout, 8ull * n, hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) return fail(c, RBH_E_HIP, std::string("SDF kernel: ") + hipGetErrorString(e));
  return 0;
}

int rbh_shape_upload(rbh_ctx* c, const double* V, int nv, const int* T, int nt, const double* sdf, int I, int J,
                     int K, const double* mn, const double* mx, int* shape_id) {
  if (!c || !V || !T || !sdf || !mn || !mx || !shape_id || nv <= 0 || nt <= 0) return fail(c, RBH_E_ARG, "bad shape");
  // a closed surface has at least 4 triangles (mass.py rejects a mesh without volume); the broad phase keeps
  // its cluster and block boxes inside the body's own triangle range, which needs nt >= 2
  if (nt < 2) return fail(c, RBH_E_ARG, "a shape needs at least 2 triangles");
  if (I < 2 || J < 2 || K < 2) return fail(c, RBH_E_ARG, "grid needs at least 2 nodes per axis");
  for (int i = 0; i < 3 * nt; ++i) if (T[i] < 0 || T[i] >= nv) return fail(c, RBH_E_ARG, "triangle index out of range");
  ShapeInfo s{};
  s.v_off = (int)(c->sV.size() / 4);
  s.nv = nv;
  s.t_off = (int)(c->sT.size() / 4);
  s.nt = nt;
  s.g_off = (int)c->sdf.size();
  s.I = I; s.J = J; s.K = K;
  const int res[3] = {I - 1, J - 1, K - 1};
  for (int a = 0; a < 3; ++a) {
    s.mn[a] = mn[a];
    s.mx[a] = mx[a];
    s.h[a] = (mx[a] - mn[a]) / (double)res[a];   // grid3.py:42-44
  }
  s.hdiag = std::sqrt(s.h[0] * s.h[0] + s.h[1] * s.h[1] + s.h[2] * s.h[2]);
  s.c_off = c->cells_total;
  c->cells_total += 8LL * (I - 1) * (J - 1) * (K - 1);
  s.reach = 0.0;
  for (int i = 0; i < nv; ++i)
    s.reach = std::max(s.reach, std::sqrt(V[3 * i] * V[3 * i] + V[3 * i + 1] * V[3 * i + 1] + V[3 * i + 2] * V[3 * i + 2]));
  for (int i = 0; i < nv; ++i) {
    c->sV.push_back(V[3 * i]); c->sV.push_back(V[3 * i + 1]); c->sV.push_back(V[3 * i + 2]); c->sV.push_back(0.0);
  }
  // the fourth int of triangle record p holds the triangle at position p of
  // the cluster order: triangles sorted by the Morton code of their centroid
  // (10 bits per axis over the shape's box), so each run of CLUSTER
  // positions is a compact patch of the surface
  double lo[3] = {V[0], V[1], V[2]}, hi[3] = {V[0], V[1], V[2]};
  for (int i = 0; i < nv; ++i)
    for (int a = 0; a < 3; ++a) { lo[a] = std::min(lo[a], V[3 * i + a]); hi[a] = std::max(hi[a], V[3 * i + a]); }
  auto spread = [](unsigned x) {   // 10 bits -> every third bit
    x &= 0x3ffu;
    x = (x | (x << 16)) & 0x030000ffu;
    x = (x | (x << 8)) & 0x0300f00fu;
    x = (x | (x << 4)) & 0x030c30c3u;
    x = (x | (x << 2)) & 0x09249249u;
    return x;
  };
  std::vector<std::pair<unsigned, int>> key(nt);
  for (int t = 0; t < nt; ++t) {
    unsigned q[3];
    for (int a = 0; a < 3; ++a) {
      const double cen = (V[3 * T[3 * t] + a] + V[3 * T[3 * t + 1] + a] + V[3 * T[3 * t + 2] + a]) / 3.0;
      const double u = hi[a] > lo[a] ? (cen - lo[a]) / (hi[a] - lo[a]) : 0.0;
      q[a] = (unsigned)std::min(1023.0, std::max(0.0, u * 1024.0));
    }
    key[t] = {spread(q[0]) | (spread(q[1]) << 1) | (spread(q[2]) << 2), t};
  }
  std::stable_sort(key.begin(), key.end());
  for (int t = 0; t < nt; ++t) {
    c->sT.push_back(T[3 * t]); c->sT.push_back(T[3 * t + 1]); c->sT.push_back(T[3 * t + 2]);
    c->sT.push_back(key[t].second);
  }
  c->sdf.insert(c->sdf.end(), sdf, sdf + (size_t)I * J * K);
  c->shapes.push_back(s);
  c->shapes_dirty = true;
  *shape_id = (int)c->shapes.size() - 1;
  return 0;
}

int rbh_scene_footprint(rbh_ctx* c, const rbh_scene_desc* d, unsigned long long* bytes) {
  if (!c || !d || !bytes) return RBH_E_ARG;
  long long nt = 0;
  for (int i = 0; i < d->n_bodies; ++i) {
    if (d->shape[i] < 0 || d->shape[i] >= (int)c->shapes.size()) return fail(c, RBH_E_ARG, "unknown shape id");
    nt += c->shapes[d->shape[i]].nt;
  }
  const long long V = 4LL * c->prec;
  // arena (23 body arrays) + triangle AABBs + workspace estimate per body / triangle
  long long b = (long long)d->n_bodies * (19 * V + 5 * 4 + c->prec) + nt * 2 * V;
  long long ws_contacts = 64LL * d->n_bodies;
  long long ws = ws_contacts * (16 * V + PALW * 4 + 64) + 8 * nt * 5 + 4 * nt * (2 * V + 13) + (long long)d->n_bodies * 64;
  *bytes = (unsigned long long)(b + ws + 16LL * std::max(1, d->params.max_iterations));
  return 0;
}

int rbh_scene_create(rbh_ctx* c, const rbh_scene_desc* d, int* scene_id) {
  if (!c || !d || !scene_id) return fail(c, RBH_E_ARG, "null argument");
  if (d->n_bodies <= 0) return fail(c, RBH_E_ARG, "a scene needs at least one body");
  if (!d->shape || !d->mass || !d->inertia) return fail(c, RBH_E_ARG, "shape, mass and inertia are required");
  int rc = check_params(c, d->params);
  if (rc) return rc;
  const int nmat = std::max(1, d->n_materials);
  long long nt = 0;
  for (int i = 0; i < d->n_bodies; ++i) {
    if (d->shape[i] < 0 || d->shape[i] >= (int)c->shapes.size()) return fail(c, RBH_E_ARG, "unknown shape id");
    int mt = d->material ? d->material[i] : 0;
    if (mt < 0 || mt >= nmat) return fail(c, RBH_E_ARG, "material index out of range");
    bool fixed = d->fixed && d->fixed[i];
    if (!fixed && d->mass[i] == 0.0) return fail(c, RBH_E_ARG, "Body " + std::to_string(i) + " has no mass.");
    nt += c->shapes[d->shape[i]].nt;
  }
  (void)hipSetDevice(c->device);
  if (c->budget) {
    unsigned long long fp = 0;
    rbh_scene_footprint(c, d, &fp);
    if (total_bytes(c) + fp > c->budget) return fail(c, RBH_E_MEMORY, "scene does not fit the HBM budget");
  }
  SceneHost s;
  s.nbody = d->n_bodies;
  s.ntri = (int)nt;
  s.body_off = c->bodies.alloc(s.nbody);
  if (s.body_off < 0) {
    if ((rc = grow_bodies(c, c->bodies.used() + s.nbody))) return rc;
    s.body_off = c->bodies.alloc(s.nbody);
  }
  s.tri_off = c->tris.alloc(s.ntri);
  if (s.tri_off < 0) {
    if ((rc = grow_tris(c, c->tris.used() + s.ntri))) return rc;
    s.tri_off = c->tris.alloc(s.ntri);
  }
  s.res_n = 2 * std::max(1, d->params.max_iterations);   // contact solve, then post-stabilisation solve
  s.res_off = c->res.alloc(s.res_n);
  if (s.res_off < 0) {
    if ((rc = grow_res(c, c->res.used() + s.res_n))) return rc;
    s.res_off = c->res.alloc(s.res_n);
  }
  if (s.body_off < 0 || s.tri_off < 0 || s.res_off < 0) return fail(c, RBH_E_MEMORY, "arena allocation failed");
  s.nmat = nmat;
  s.mat_off = c->mats.alloc(nmat * nmat);
  if (s.mat_off < 0) {
    c->mats.grow(std::max(c->mats.used() + nmat * nmat, 2 * c->mats.cap));
    c->mat_host.resize((size_t)c->mats.cap * 4, 0.0);
    s.mat_off = c->mats.alloc(nmat * nmat);
  }
  for (int i = 0; i < nmat * nmat; ++i)
    for (int k = 0; k < 4; ++k)
      c->mat_host[4 * ((size_t)s.mat_off + i) + k] = d->interactions ? d->interactions[4 * i + k] : (k < 3 ? 1.0 : 0.0);
  s.params = d->params;
  s.live = true;
  int slot = -1;
  for (size_t i = 0; i < c->scenes.size(); ++i) if (!c->scenes[i].live) { slot = (int)i; break; }
  if (slot < 0) {
    slot = (int)c->scenes.size();
    c->scenes.push_back(s);
    c->sp_host.push_back(SceneParams{});
  } else {
    c->scenes[slot] = s;
  }
  c->sp_host[slot] = to_device_params(s.params, s.nmat, s.mat_off, s.res_off);
  c->scenes_dirty = true;
  c->hinges_dirty = true;
  if ((rc = sync_scenes(c))) return rc;
  SceneState z{};
  HIPCHK(c, hipMemcpyAsync(reinterpret_cast<SceneState*>(c->d_ss.p) + slot, &z, sizeof(z), hipMemcpyHostToDevice,
                           c->stream));
  rc = c->prec == 4 ? scene_statics<float>(c, d, s, slot) : scene_statics<double>(c, d, s, slot);
  if (rc) return rc;
  c->arena_gen++;
  drop_graph(c);
  *scene_id = slot;
  return 0;
}

int rbh_scene_destroy(rbh_ctx* c, int id) {
  if (!c || !valid_scene(c, id)) return fail(c, RBH_E_ARG, "unknown scene");
  SceneHost& s = c->scenes[id];
  c->bodies.release(s.body_off, s.nbody);
  c->tris.release(s.tri_off, s.ntri);
  c->res.release(s.res_off, s.res_n);
  c->mats.release(s.mat_off, s.nmat * s.nmat);
  s.live = false;
  s.h_parent.clear(); s.h_child.clear(); s.h_geo.clear();
  c->hinges_dirty = true;
  c->arena_gen++;
  drop_graph(c);
  c->batch.clear();
  return 0;
}

int rbh_scene_set_hinges(rbh_ctx* c, int id, int n, const int* parent, const int* child, const double* arm_p,
                         const double* arm_c, const double* axis_p, const double* axis_c) {
  if (!c || !valid_scene(c, id)) return fail(c, RBH_E_ARG, "unknown scene");
  if (n < 0 || (n > 0 && (!parent || !child || !arm_p || !arm_c || !axis_p || !axis_c)))
    return fail(c, RBH_E_ARG, "bad hinge arguments");
  SceneHost& s = c->scenes[id];
  std::vector<char> joined(s.nbody, 0);
  int nj = 0;
  for (int k = 0; k < n; ++k) {
    if (parent[k] < 0 || parent[k] >= s.nbody || child[k] < 0 || child[k] >= s.nbody)
      return fail(c, RBH_E_ARG, "hinge " + std::to_string(k) + ": body index out of range");
    if (parent[k] == child[k]) return fail(c, RBH_E_ARG, "hinge " + std::to_string(k) + ": parent and child are one body");
    for (int b : {parent[k], child[k]}) if (!joined[b]) { joined[b] = 1; ++nj; }
  }
  if (nj > HINGE_BODY_MAX)
    return fail(c, RBH_E_ARG, "hinges join more than " + std::to_string(HINGE_BODY_MAX) + " bodies of one scene");
  if (n > HINGE_MAX) return fail(c, RBH_E_ARG, "more than " + std::to_string(HINGE_MAX) + " hinges in one scene");
  s.h_parent.assign(parent, parent + n);
  s.h_child.assign(child, child + n);
  s.h_geo.assign(12 * (size_t)n, 0.0);
  for (int k = 0; k < n; ++k)
    for (int j = 0; j < 3; ++j) {
      s.h_geo[12 * k + j] = arm_p[3 * k + j];
      s.h_geo[12 * k + 3 + j] = arm_c[3 * k + j];
      s.h_geo[12 * k + 6 + j] = axis_p[3 * k + j];
      s.h_geo[12 * k + 9 + j] = axis_c[3 * k + j];
    }
  c->hinges_dirty = true;
  c->batch.clear();
  (void)hipSetDevice(c->device);
  return c->prec == 4 ? sync_hinges<float>(c) : sync_hinges<double>(c);
}

int rbh_scene_set_params(rbh_ctx* c, int id, const rbh_params* p) {
  if (!c || !p || !valid_scene(c, id)) return fail(c, RBH_E_ARG, "unknown scene");
  int rc = check_params(c, *p);
  if (rc) return rc;
  SceneHost& s = c->scenes[id];
  if (2 * std::max(1, p->max_iterations) > s.res_n) {
    c->res.release(s.res_off, s.res_n);
    s.res_n = 2 * std::max(1, p->max_iterations);
    s.res_off = c->res.alloc(s.res_n);
    if (s.res_off < 0) {
      if ((rc = grow_res(c, c->res.used() + s.res_n))) return rc;
      s.res_off = c->res.alloc(s.res_n);
    }
  }
  s.params = *p;
  c->sp_host[id] = to_device_params(s.params, s.nmat, s.mat_off, s.res_off);
  c->scenes_dirty = true;
  c->hinges_dirty = true;   // re-applies the scene's hinge offsets to its parameters
  drop_graph(c);
  int rc2 = c->prec == 4 ? sync_hinges<float>(c) : sync_hinges<double>(c);
  return rc2 ? rc2 : sync_scenes(c);
}

int rbh_scene_set_state(rbh_ctx* c, int id, const double* r, const double* q, const double* v, const double* w) {
  if (!c || !valid_scene(c, id)) return fail(c, RBH_E_ARG, "unknown scene");
  (void)hipSetDevice(c->device);
  return c->prec == 4 ? set_state<float>(c, c->scenes[id], r, q, v, w) : set_state<double>(c, c->scenes[id], r, q, v, w);
}

int rbh_scene_get_state(rbh_ctx* c, int id, double* r, double* q, double* v, double* w) {
  if (!c || !valid_scene(c, id)) return fail(c, RBH_E_ARG, "unknown scene");
  (void)hipSetDevice(c->device);
  return c->prec == 4 ? get_state<float>(c, c->scenes[id], r, q, v, w) : get_state<double>(c, c->scenes[id], r, q, v, w);
}

int rbh_scene_set_step_index(rbh_ctx* c, int id, int step_index) {
  if (!c || !valid_scene(c, id)) return fail(c, RBH_E_ARG, "unknown scene");
  if (step_index < 0) return fail(c, RBH_E_ARG, "step index must be non-negative");
  (void)hipSetDevice(c->device);
  HIPCHK(c, hipMemcpyAsync(reinterpret_cast<char*>(c->d_ss.p) + sizeof(SceneState) * id + offsetof(SceneState, step_index),
                           &step_index, sizeof(int), hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return 0;
}

int rbh_step(rbh_ctx* c, const int* ids, int n, int n_steps, double dt, int flags) {
  if (!c || !ids || n <= 0 || n_steps < 0) return fail(c, RBH_E_ARG, "bad step arguments");
  if (n_steps == 0) return 0;
  for (int i = 0; i < n; ++i) {
    if (!valid_scene(c, ids[i])) return fail(c, RBH_E_ARG, "unknown scene " + std::to_string(ids[i]));
    for (int j = 0; j < i; ++j) if (ids[j] == ids[i]) return fail(c, RBH_E_ARG, "scene listed twice");
  }
  if (!(dt > 0.0) && dt != 0.0 && !(dt < 0.0)) return fail(c, RBH_E_ARG, "dt is NaN");
  (void)hipSetDevice(c->device);
  int rc = c->prec == 4 ? sync_shapes<float>(c) : sync_shapes<double>(c);
  if (!rc) rc = c->prec == 4 ? sync_hinges<float>(c) : sync_hinges<double>(c);
  if (!rc) rc = sync_scenes(c);
  if (!rc) rc = build_batch(c, ids, n);
  if (rc) return rc;
  return c->prec == 4 ? run_steps<float>(c, n_steps, dt, flags) : run_steps<double>(c, n_steps, dt, flags);
}

int rbh_sync(rbh_ctx* c) {
  if (!c) return RBH_E_ARG;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return 0;
}

int rbh_get_log(rbh_ctx* c, int id, rbh_step_log* log, double* residuals, int cap) {
  if (!c || !valid_scene(c, id)) return fail(c, RBH_E_ARG, "unknown scene");
  SceneState st;
  HIPCHK(c, hipMemcpyAsync(&st, reinterpret_cast<SceneState*>(c->d_ss.p) + id, sizeof(st), hipMemcpyDeviceToHost,
                           c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  const SceneHost& s = c->scenes[id];
  if (log) {
    std::memset(log, 0, sizeof(*log));
    log->step_index = st.step_index;
    log->contact_points = st.k_cnt;
    log->iterations = st.iterations;
    log->exitcode = st.exitcode;
    log->rejects = st.rejects;
    log->sweeps = st.sweeps;
    log->colors = st.ncolor;
    log->solved = st.solved;
    log->number_of_overlaps = st.n_overlaps;
    log->fw_cell_evals = st.fw_evals;
    log->natural_residual = st.nat_r / std::max(1.0, st.nat_x);
    log->max_penetration = st.max_pen;
    log->final_merit = st.final_merit;
    log->post_iterations = st.post_active ? st.post_iters : 0;
    log->post_exitcode = st.post_active ? st.post_exit : -1;
  }
  if (residuals && cap > 0) {
    // [0, mi): the contact solve; [mi, 2 mi): the post-stabilisation solve of this step, if it ran
    const int mi = std::max(1, s.params.max_iterations);
    const int m = std::min(cap, s.params.max_iterations);
    if (m > 0)
      HIPCHK(c, hipMemcpyAsync(residuals, reinterpret_cast<double*>(c->d_res.p) + s.res_off, m * 8,
                               hipMemcpyDeviceToHost, c->stream));
    const int mp = (st.post_active && s.params.use_post_stabilization) ? std::min(cap - mi, s.params.max_iterations) : 0;
    if (mp > 0)
      HIPCHK(c, hipMemcpyAsync(residuals + mi, reinterpret_cast<double*>(c->d_res.p) + s.res_off + mi, mp * 8,
                               hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (int i = m; i < std::min(cap, mi); ++i) residuals[i] = INFINITY;
    for (int i = mi + std::max(mp, 0); i < cap; ++i) residuals[i] = INFINITY;
  }
  return 0;
}

int rbh_get_contacts(rbh_ctx* c, int id, int cap, int* count, int* ia, int* ib, int* tri, double* p, double* n,
                     double* gap, double* impulse, int* color) {
  if (!c || !count || !valid_scene(c, id)) return fail(c, RBH_E_ARG, "unknown scene");
  const SceneHost& s = c->scenes[id];
  if (s.ws_gen != c->ws_gen) return fail(c, RBH_E_STATE, "contacts of this scene were overwritten by a later step");
  SceneState st;
  HIPCHK(c, hipMemcpyAsync(&st, reinterpret_cast<SceneState*>(c->d_ss.p) + id, sizeof(st), hipMemcpyDeviceToHost,
                           c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return c->prec == 4 ? get_contacts<float>(c, s, st, cap, count, ia, ib, tri, p, n, gap, impulse, color)
                      : get_contacts<double>(c, s, st, cap, count, ia, ib, tri, p, n, gap, impulse, color);
}

int rbh_get_timings(rbh_ctx* c, double* ms, int cap) {
  if (!c || !ms) return RBH_E_ARG;
  int m = std::min(cap, TIMING_OUT);
  for (int i = 0; i < m; ++i) ms[i] = c->timings[i];
  return m;
}

int rbh_get_counters(rbh_ctx* c, long long* out, int cap) {
  if (!c || !out) return RBH_E_ARG;
  int m = std::min(cap, 4);
  for (int i = 0; i < m; ++i) out[i] = c->counters[i];
  return m;
}

int rbh_set_tuning(rbh_ctx* c, const rbh_tuning* t) {
  if (!c || !t) return RBH_E_ARG;
  if (t->unit_size < 0 || t->unit_size > 4096 || t->lanes_per_scene < 0 || t->lanes_per_scene > 64 ||
      t->resident_contacts < 0 || t->resident_bodies < 0 || t->lanes_per_group < 0 || t->lanes_per_group > 64 ||
      t->solve_groups_per_cu < 0 || t->solve_groups_per_cu > 2 ||
      (t->waves_per_group != 0 && t->waves_per_group != 4 && t->waves_per_group != 8) ||
      (t->resident_contacts % 64) || (t->resident_bodies % 64) || t->color_lds_contacts < 0 ||
      t->color_body_bits < 0)
    return fail(c, RBH_E_ARG, "tuning values out of range (resident capacities are multiples of 64)");
  const int cap = c->prec == 4 ? ResCap<float>::C : ResCap<double>::C;
  if (t->resident_contacts > cap || t->resident_bodies > cap)
    return fail(c, RBH_E_ARG, "resident capacities are at most " + std::to_string(cap) + " contacts and bodies");
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->tune = *t;
  c->solve_lds = 0;          // recomputed (and the cooperative occupancy with it) by the next step
  c->coop_max = 0;
  c->batch_arena_gen = -1;   // the next step rebuilds the batch with the new unit sizes
  drop_graph(c);
  return 0;
}

int rbh_get_tuning(rbh_ctx* c, rbh_tuning* t) {
  if (!c || !t) return RBH_E_ARG;
  *t = c->tune;
  return 0;
}

int rbh_scene_totals(rbh_ctx* c, int id, unsigned long long* out, int cap, int clear) {
  if (!c || !valid_scene(c, id) || (cap > 0 && !out)) return fail(c, RBH_E_ARG, "unknown scene or null output");
  HIPCHK(c, hipSetDevice(c->device));
  SceneState st;
  SceneState* dst = reinterpret_cast<SceneState*>(c->d_ss.p) + id;
  HIPCHK(c, hipMemcpyAsync(&st, dst, sizeof(st), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  const unsigned long long t[6] = {st.tot_steps, st.tot_contact_iters, st.tot_contacts, st.tot_cand,
                                   st.tot_fw_evals, st.tot_sweeps};
  int m = std::max(0, std::min(cap, 6));
  for (int i = 0; i < m; ++i) out[i] = t[i];
  if (clear) {
    st.tot_steps = st.tot_contact_iters = st.tot_contacts = st.tot_cand = st.tot_fw_evals = st.tot_sweeps = 0;
    HIPCHK(c, hipMemcpyAsync(dst, &st, sizeof(st), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
  }
  return m;
}

}  // extern "C"
