// Kernels of one batched time step.  A "world" is a list of independent
// scenes whose bodies live in one SoA arena; every kernel works over the whole
// batch at once, so one scene stepped alone is simply a batch of one.
//
// Pipeline of one step (all sizes that depend on data stay on the device):
//   body_prep  -> refit -> pair_count/scan/pair_write -> cand_flags ->
//   scan -> cand_write -> frank_wolfe -> dedup -> scan -> contact_write ->
//   assemble -> body->contact CSR -> colour (per scene) -> order (per scene) ->
//   solve (per scene) -> integrate -> finalize.
#pragma once
#include "rbh_device.h"
#include "rbh_scan.h"

namespace rbh {

// -------------------------------------------------------------- shared types
struct ShapeInfo {
  int v_off, nv, t_off, nt, g_off, I, J, K;
  double mn[3], mx[3], h[3];
  // reach: largest distance of a vertex from the body-frame origin; hdiag:
  // the SDF cell diagonal (sqrt(hx^2 + hy^2 + hz^2))
  double reach, hdiag;
  long long c_off;   // first value of the shape's cell-major copy of its grid (8 per cell)
};

enum Scheme { SCHEME_GS = 0, SCHEME_PGS = 1, SCHEME_JACOBI = 2, SCHEME_JACOBI_HYBRID = 3 };

struct SceneParams {
  double time_step, envelope, abs_tol, rel_tol, nu_reduce, nu_increase, too_small;
  double fw_tol, gap_reduction, min_gap, max_gap, ell_tol, ell_expansion;
  int max_iter, fw_max_iter, ell_max_iter, use_bounce, use_pre_stab, scheme, friction;
  int nmat, mat_off, res_off, use_post_stab;
  int h_off, h_cnt;     // hinges of the scene in the hinge arrays
  int hb_off, hb_cnt;   // bodies joined by those hinges (hinge-body list)
  unsigned long long seed;
};

struct SceneState {
  int step_index;      // committed steps
  int k_off, k_cnt;    // contacts of this step in the batch contact arrays
  int ncolor;
  int iterations, exitcode, rejects, sweeps;
  int n_pairs, n_cand, solved, post_active;   // post_active: a contact is deeper than min_gap (post-stabilisation)
  int post_iters, post_exit;                  // the post-stabilisation solve's iterations and exit code
  unsigned long long n_overlaps;
  unsigned long long fw_evals;   // distinct trilinear cell evaluations in Frank-Wolfe
  double nat_r, nat_x;     // natural-map residual: max |x - prox(...)| and max |x| (FP64)
  double max_pen, final_merit;
  // running totals since the scene was created or they were last cleared
  unsigned long long tot_steps, tot_contact_iters, tot_contacts, tot_cand, tot_fw_evals, tot_sweeps;
};

constexpr int PALW = 64;                 // palette words per contact
constexpr int PALMAX = PALW * 32;        // largest colour + 1
constexpr int MAX_ROUNDS = 200;

// error flags (dev.err[0] bits); capacity overflows record the needed size
enum ErrBits { ERR_CAP_PAIRS = 1, ERR_CAP_SLOTS = 2, ERR_CAP_CAND = 4, ERR_CAP_CONTACTS = 8, ERR_COLORS = 16,
               ERR_SOLVER = 32, ERR_CAP_UNITS = 64, ERR_SCAN = SCAN_ERR_STALL };
static_assert(ERR_SCAN == 128, "distinct error bit");

template <typename R> struct Dev {
  // shapes
  const ShapeInfo* shapes; const V4<R>* sV; const int4* sT; const R* sdf;
  const R* cells;                      // per shape, per cell (i, j, k) its 8 corner values (cell-major grid)
  // scenes (by slot)
  const SceneParams* sp; SceneState* ss; double* residuals; const double* mat;   // mat: (mu0,mu1,mu2,eps) per pair
  // body arena
  V4<R>* r; V4<R>* q; V4<R>* v; V4<R>* w;
  const int* b_shape; const int* b_slot; const int* b_fixed; const int* b_mat; const int* b_tri;
  const int4* b_broad;                 // per body, read by the broad phase in one load: b_tri, the shape's
                                       // nt and t_off, b_slot
  const R* b_mass; const V4<R>* b_inertia; const V4<R>* b_grav; const V4<R>* b_damp;
  V4<R>* b_W0; V4<R>* b_W1;            // (invI xx,yy,zz,xy), (invI xz,yz, 1/m, 0)
  V4<R>* b_du0; V4<R>* b_du1;          // W f_ext (linear, angular)
  V4<R>* b_lo; V4<R>* b_hi;            // world AABB grown by the envelope; w: the shape's reach, cell diagonal
  V4<R>* ws0; V4<R>* ws1;              // accepted W J^T sol
  V4<R>* ww0; V4<R>* ww1;              // working copy during a sweep
  V4<R>* tlo; V4<R>* thi;              // per body-instance triangle AABBs, in cluster order (b_tri + position)
  V4<R>* clo; V4<R>* chi;              // per body-instance cluster AABBs (CLUSTER triangles each) at b_tri + c,
                                       // then block AABBs (CAND_BLK positions each) at blk_box(b_tri, nt) + k
  // batch
  int nb, ns;
  const int* bb_body;                  // flat batch body -> arena body
  const int* bs_slot;                  // batch scene -> slot
  const int* bs_order;                 // batch scenes by decreasing body count (the batched solve's workgroup order)
  const int* bs_foff;                  // batch scene -> first flat body
  const int* b2f;                      // arena body -> flat batch body
  const int* f_end;                    // flat batch body -> arena end of its scene
  const int* d_nb;                     // device copy of nb (scan length)
  int* row_cnt; int* row_off;          // [nb], [nb+1]
  int2* pairs; int* pair_sz; int* slot_off; int* n_pairs; int cap_pairs;
  int* pair_items; int* item_off;      // candidate work items (triangle blocks of both sides) per pair, scan
  int* item_pair_of; int cap_items;    // the pair of each work item (k_item_map), its capacity
  unsigned char* slot_flag; int* n_slots; int cap_slots;
  int* side_cnt; int* side_off; int* n_sides;   // candidates per (pair, side) 2 p + side, their scan, 2 np
  int* cand_pair; int* cand_tri; int* pair_cstart; int* n_cand; int cap_cand;
  unsigned long long* fw_acc;          // per scene slot: FW_SPREAD partial (evals, candidates) counters
  V4<R>* cand_p; V4<R>* cand_n; unsigned char* cand_keep; int* keep_scan;
  int* kidx; V4<R>* kpos; int* n_kept;  // candidates kept by Frank-Wolfe, compacted (candidate order)
  unsigned char* kept_flag; int* kept_scan;   // 1 if not a duplicate; its scan (contact index)
  // contacts (canonical order)
  int* c_ia; int* c_ib; int* c_tri; int* c_pair; V4<R>* c_p; V4<R>* c_n; int* n_contacts; int cap_contacts;
  V4<R>* c_rA; V4<R>* c_rB; V4<R>* c_b; V4<R>* c_r0; V4<R>* c_mu;
  V4<R>* c_imp;                        // solved impulses (canonical order)
  int* c_color;
  // islands and colouring units (contacts regrouped by island: "unit order")
  const int* f_bs;                     // flat batch body -> batch scene
  int* uf;                             // union-find parent per flat body
  int* isl_cnt; int* isl_off; int* isl_fill; int* isl_max;   // per island label [nb + ns (+1)]
  int* c_lab;                          // island label per contact
  int* uperm;                          // unit-order position -> contact
  int* u_cnt; int* u_base; int* n_units; int cap_units; int unit_T; int unit_split;
  int* u_T;                            // sweep-lane size of each batch scene
  int color_cap;                       // most contacts of a unit coloured with its state in LDS
  int color_bits;                      // most words of each per-body colour bitmap array of a small unit
  int* l_cnt; int* l_base; int* n_lanes;   // sweep lanes per batch scene, their scan, total
  int* l_start; int* l_end; int* l_ncol; int* l_bs;   // per lane (capacity cap_units: lanes <= units)
  int* u_start; int* u_end; int* u_ncol; int* u_bs;
  int* isl_unit;                       // island label -> unit
  int* u_nbody; int* u_boff;           // free bodies per unit, and their exclusive scan (slot base)
  int* f_loc; int* f_unit;             // flat body -> slot within its unit (-1: no contact), unit
  int* u_blist;                        // slot (u_boff order) -> arena body
  int* body_claim; int2* body_loc;     // per flat body, LDS staging of body lists
  int* bc_cnt; int* bc_off; int* bc_fill; int* bc_list;   // flat body -> unit-order positions
  unsigned* pal;                       // PALW words per unit-order position
  int* g_pc; unsigned char* g_done;    // colouring state of units too large for LDS
  int* col_off;                        // colour offsets of lane l at 2 * (l_start[l] + l)
  int* l_pos;                          // per unit-order position p (contact uperm[p]): its position in colour order within its lane
  int items_nb;                        // a batch scene of more bodies than this solves from the per-contact
                                       // items (s_n .. s_W, x0, x1); -1: any group may
  int* big_list;                       // [0]: count, then the units left to k_color_units
  // a scene solved by several workgroups: per sweep and workgroup, the bits
  // of that workgroup's merit (all ones until it publishes them)
  unsigned long long* merit;
  // solver arrays (colour order)
  int* perm; int2* s_id; V4<R>* s_n; V4<R>* s_rA; V4<R>* s_rB; V4<R>* s_b; V4<R>* s_r0; V4<R>* x0; V4<R>* x1;
  int2* s_loc;                         // body slots (u_blist index) of the two bodies (-1: fixed)
  int res_cap_c, res_cap_b, res_cap_o; // LDS capacities of a resident solve group (contacts, bodies, colour offsets)
  V4<R>* s_W;                          // 4 per position: W0, W1 of body A, then of body B (zero if fixed)
  V4<R>* s_JL; V4<R>* s_JA; V4<R>* s_JB;   // rows 4 pos + r of J: (lin_r, b_r), (angA_r, r0_r), (angB_r, mu_r)
  R* xr0; R* xr1;                      // iterates by row (4 pos + r), for the row-major sweeps
  // hinges (constraints/hinges.py): static per hinge, then per-step rows and iterates
  const int4* h_ids;                   // (parent, child) arena bodies, (parent, child) hinge-body slots
  const V4<R>* h_geo;                  // 4 per hinge: arm_p, arm_c, axis_p, axis_c (body frames)
  const int* hb_body;                  // hinge-body slot -> arena body
  const int* hb_inc_off; const int* hb_inc;   // hinge-body slot -> 2 h + side (0 parent, 1 child), hinge order
  V4<R>* h_J; V4<R>* h_WJ;             // 20 per hinge: row i at 4 i: parent lin, parent ang, child lin, child ang
  R* h_b; R* h_r0;                     // 5 per hinge
  double* h_mer;                       // per sweep: the hinge problem's merit contribution (at res_off)
  V4<R>* hd0; V4<R>* hd1;              // per arena body: hinge W J^T sol (linear, angular)
  int* err;                           // [0] flags, [1..] needed capacities
  int post;                           // 1: the solve kernels run the post-stabilisation problem
  double dt;                           // > 0: use for every scene, else the scene's time_step
  int flags;
};

enum StepFlags { STEP_COUNT_OVERLAPS = 1 };

template <typename R> __device__ __forceinline__ R scene_dt(const Dev<R>& d, int slot) {
  return R(d.dt > 0.0 ? d.dt : d.sp[slot].time_step);
}

__device__ __forceinline__ bool aborted(const int* err) { return __atomic_load_n(err, __ATOMIC_RELAXED) != 0; }

template <typename R> __device__ __forceinline__ GridView<R> grid_of(const Dev<R>& d, int shape) {
  const ShapeInfo& s = d.shapes[shape];
  GridView<R> g;
  g.val = d.sdf + s.g_off;
  g.cells = d.cells + s.c_off;
  g.I = s.I; g.J = s.J; g.K = s.K;
  for (int i = 0; i < 3; ++i) { g.mn[i] = R(s.mn[i]); g.mx[i] = R(s.mx[i]); g.h[i] = R(s.h[i]); }
  return g;
}

template <typename R> __device__ __forceinline__ bool box_overlap(V4<R> alo, V4<R> ahi, V4<R> blo, V4<R> bhi) {
  // kdop.py:21-38: separated iff upper < lower on some axis
  return !(ahi.x < blo.x || bhi.x < alo.x || ahi.y < blo.y || bhi.y < alo.y || ahi.z < blo.z || bhi.z < alo.z);
}

__device__ __forceinline__ unsigned long long wave_sum_u64(unsigned long long v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

__device__ __forceinline__ int wave_sum(int v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// ========================================================== per-body prep
// state_storage.py:108-185 -- W blocks and delta_u_ext = W f_ext
template <typename R> __global__ void k_body_prep(Dev<R> d) {
  int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= d.nb) return;
  int b = d.bb_body[f];
  V4<R> z = {R(0), R(0), R(0), R(0)};
  d.ws0[b] = z; d.ws1[b] = z; d.ww0[b] = z; d.ww1[b] = z; d.hd0[b] = z; d.hd1[b] = z;
  if (d.b_fixed[b]) {
    d.b_W0[b] = z; d.b_W1[b] = z; d.b_du0[b] = z; d.b_du1[b] = z;
    return;
  }
  R M[9];
  qmatrix(quat(d.q[b]), M);
  V4<R> I = d.b_inertia[b];
  R Iw[6], Wi[6];
  rot_diag_rt(M, I.x, I.y, I.z, Iw);
  rot_diag_rt(M, R(1) / I.x, R(1) / I.y, R(1) / I.z, Wi);
  V3<R> v = xyz(d.v[b]), w = xyz(d.w[b]);
  R m = d.b_mass[b];
  V4<R> g = d.b_grav[b], dm = d.b_damp[b];
  V3<R> F = V3<R>{-m * g.x, -m * g.y, -m * g.z};
  F = F - dm.x * v;
  V3<R> T = -cross(w, sym_mul(Iw, w));
  T = T - dm.y * w;
  R im = R(1) / m;
  V3<R> dl = im * F;
  V3<R> da = sym_mul(Wi, T);
  d.b_W0[b] = V4<R>{Wi[0], Wi[1], Wi[2], Wi[3]};
  d.b_W1[b] = V4<R>{Wi[4], Wi[5], im, R(0)};
  d.b_du0[b] = V4<R>{dl.x, dl.y, dl.z, R(0)};
  d.b_du1[b] = V4<R>{da.x, da.y, da.z, R(0)};
}

// ========================================================== BVH leaf refit
// collision_detection.py:10-44 + kdop.py:53-75: world vertices, per-triangle
// AABB grown by the envelope, body AABB = union.  One wave per body.
// Triangles in clusters of CLUSTER consecutive positions of the shape's
// cluster order (sT[t_off + p].w is the triangle at position p; the order
// follows the Morton code of the centroids).  A cluster's box bounds its
// triangles' boxes, so a cluster that misses a volume holds no triangle that
// meets it.
constexpr int CLUSTER = 32;
// Blocks of CAND_BLK consecutive positions (CAND_BLK / CLUSTER clusters) are
// the broad phase's unit of work; their boxes follow the cluster boxes.
constexpr int CAND_BLK = 256;
__host__ __device__ constexpr int n_blocks(int nt) { return (nt + CAND_BLK - 1) / CAND_BLK; }
constexpr int FW_SPREAD = 64;    // partial Frank-Wolfe counters per scene (a power of two)
constexpr int FW_STRIDE = 16;    // words between them (128 bytes)
// first block box of a body whose triangles start at tb; ceil(nt / CLUSTER) +
// ceil(nt / CAND_BLK) <= nt for nt >= 2, so the boxes stay inside [tb, tb + nt)
__device__ __forceinline__ int blk_box(int tb, int nt) { return tb + (nt + CLUSTER - 1) / CLUSTER; }

template <typename R> __global__ void __launch_bounds__(64) k_refit(Dev<R> d) {
  int f = blockIdx.x;
  if (f >= d.nb || aborted(d.err)) return;
  int b = d.bb_body[f];
  const ShapeInfo& sh = d.shapes[d.b_shape[b]];
  const R env = R(d.sp[d.b_slot[b]].envelope);
  Quat<R> q = quat(d.q[b]);
  V3<R> r = xyz(d.r[b]);
  const R inf = R(INFINITY);
  R lx = inf, ly = inf, lz = inf, hx = -inf, hy = -inf, hz = -inf;
  const int tb = d.b_tri[b];
  const int kb = blk_box(tb, sh.nt);
  const int lane = threadIdx.x;
  R bx = inf, by = inf, bz = inf, ex = -inf, ey = -inf, ez = -inf;   // box of the current block
  for (int base = 0; base < sh.nt; base += 64) {
    const int t = base + lane;
    R cx = inf, cy = inf, cz = inf, dx = -inf, dy = -inf, dz = -inf;
    if (t < sh.nt) {
      const int tt = d.sT[sh.t_off + t].w;   // triangle at cluster position t
      int4 tri = d.sT[sh.t_off + tt];
      V3<R> p0 = qrotate_array(q, xyz(d.sV[sh.v_off + tri.x])) + r;
      V3<R> p1 = qrotate_array(q, xyz(d.sV[sh.v_off + tri.y])) + r;
      V3<R> p2 = qrotate_array(q, xyz(d.sV[sh.v_off + tri.z])) + r;
      V4<R> lo = {fmin(fmin(p0.x, p1.x), p2.x) - env, fmin(fmin(p0.y, p1.y), p2.y) - env,
                  fmin(fmin(p0.z, p1.z), p2.z) - env, R(0)};
      V4<R> hi = {fmax(fmax(p0.x, p1.x), p2.x) + env, fmax(fmax(p0.y, p1.y), p2.y) + env,
                  fmax(fmax(p0.z, p1.z), p2.z) + env, R(0)};
      d.tlo[tb + t] = lo;
      d.thi[tb + t] = hi;
      cx = lo.x; cy = lo.y; cz = lo.z; dx = hi.x; dy = hi.y; dz = hi.z;
    }
    // cluster boxes: the 32 lanes of each half-wave hold one cluster
#pragma unroll
    for (int o = 16; o > 0; o >>= 1) {
      cx = fmin(cx, __shfl_xor(cx, o, 64)); cy = fmin(cy, __shfl_xor(cy, o, 64)); cz = fmin(cz, __shfl_xor(cz, o, 64));
      dx = fmax(dx, __shfl_xor(dx, o, 64)); dy = fmax(dy, __shfl_xor(dy, o, 64)); dz = fmax(dz, __shfl_xor(dz, o, 64));
    }
    if ((lane & (CLUSTER - 1)) == 0 && t < sh.nt) {
      const int c = t / CLUSTER;
      d.clo[tb + c] = V4<R>{cx, cy, cz, R(0)};
      d.chi[tb + c] = V4<R>{dx, dy, dz, R(0)};
    }
    lx = fmin(lx, cx); ly = fmin(ly, cy); lz = fmin(lz, cz);
    hx = fmax(hx, dx); hy = fmax(hy, dy); hz = fmax(hz, dz);
    // the two clusters of this chunk into the block box
    bx = fmin(bx, fmin(cx, __shfl_xor(cx, 32, 64))); by = fmin(by, fmin(cy, __shfl_xor(cy, 32, 64)));
    bz = fmin(bz, fmin(cz, __shfl_xor(cz, 32, 64)));
    ex = fmax(ex, fmax(dx, __shfl_xor(dx, 32, 64))); ey = fmax(ey, fmax(dy, __shfl_xor(dy, 32, 64)));
    ez = fmax(ez, fmax(dz, __shfl_xor(dz, 32, 64)));
    if ((base + 64) % CAND_BLK == 0 || base + 64 >= sh.nt) {
      if (lane == 0) {
        d.clo[kb + base / CAND_BLK] = V4<R>{bx, by, bz, R(0)};
        d.chi[kb + base / CAND_BLK] = V4<R>{ex, ey, ez, R(0)};
      }
      bx = by = bz = inf; ex = ey = ez = -inf;
    }
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    lx = fmin(lx, __shfl_xor(lx, o, 64)); ly = fmin(ly, __shfl_xor(ly, o, 64)); lz = fmin(lz, __shfl_xor(lz, o, 64));
    hx = fmax(hx, __shfl_xor(hx, o, 64)); hy = fmax(hy, __shfl_xor(hy, o, 64)); hz = fmax(hz, __shfl_xor(hz, o, 64));
  }
  // w: the shape's reach and cell diagonal, read beside the box by the pair kernels
  if (threadIdx.x == 0) {
    d.b_lo[b] = V4<R>{lx, ly, lz, R(sh.reach)};
    d.b_hi[b] = V4<R>{hx, hy, hz, R(sh.hdiag)};
  }
}

// ========================================================== body pairs
// itertools.combinations over the scene's bodies (collision_detection.py:70),
// kept when the root volumes overlap (kdop_bvh.py:340).  One wave per row.
//
// A pair can only give a contact if its bodies come within reach of each
// other.  A contact needs a point x on a triangle of A with phi_B(x) <=
// envelope, x inside B's grid (collision_detection.py:172-176).  There phi_B
// is the trilinear interpolant of exact signed distances at the corners of
// x's cell, each within the cell diagonal hd_B of x, and the signed distance
// is 1-Lipschitz, so phi_B(x) >= sd_B(x) - hd_B >= |x - r_B| - reach_B - hd_B
// >= |r_A - r_B| - reach_A - reach_B - hd_B.  Pairs farther apart than
// reach_A + reach_B + envelope + max(hd_A, hd_B) (plus a rounding margin)
// therefore give no contact, in either direction, and are dropped before
// their triangles are compared.  Counting triangle overlaps
// (STEP_COUNT_OVERLAPS, the profile key number_of_overlaps) keeps every pair.
// (a's reach and cell diagonal ride in alo.w / ahi.w, b's in blo.w / bhi.w)
template <typename R> __device__ __forceinline__ bool within_reach(V4<R> ra, V4<R> rb, V4<R> alo, V4<R> ahi, V4<R> blo,
                                                                  V4<R> bhi, R env) {
  const R lim = (alo.w + blo.w + env + fmax(ahi.w, bhi.w)) * R(1.001) + R(1e-4);
  const R dx = ra.x - rb.x, dy = ra.y - rb.y, dz = ra.z - rb.z;
  return dx * dx + dy * dy + dz * dz <= lim * lim;
}

template <typename R> __device__ __forceinline__ bool pair_kept(const Dev<R>& d, V4<R> ra, int b, V4<R> alo, V4<R> ahi,
                                                                V4<R> blo, V4<R> bhi, R env, bool all) {
  return box_overlap(alo, ahi, blo, bhi) && (all || within_reach(ra, d.r[b], alo, ahi, blo, bhi, env));
}

// A row is one wavefront, which has little else to hide load latency with:
// it walks its columns PAIR_U chunks of 64 at a time, their boxes loaded
// together (column j0 + 64 u for lane-relative start j0).  Columns past the
// row's end read the row's last column instead (end > j0 - lane, so end - 1
// is a body of the scene), which keeps every load unconditional: no branch,
// and no wait, between them.
constexpr int PAIR_U = 4;

template <typename R>
__device__ __forceinline__ void pair_boxes(const Dev<R>& d, int j0, int end, V4<R> (&blo)[PAIR_U], V4<R> (&bhi)[PAIR_U]) {
#pragma unroll
  for (int u = 0; u < PAIR_U; ++u) {
    const int j = min(j0 + 64 * u, end - 1);
    blo[u] = d.b_lo[j];
    bhi[u] = d.b_hi[j];
  }
}

template <typename R> __global__ void __launch_bounds__(64) k_pair_count(Dev<R> d) {
  int f = blockIdx.x;
  if (f >= d.nb) return;
  if (aborted(d.err)) { if (threadIdx.x == 0) d.row_cnt[f] = 0; return; }
  int b = d.bb_body[f];
  int end = d.f_end[f];
  V4<R> alo = d.b_lo[b], ahi = d.b_hi[b];
  const R env = R(d.sp[d.b_slot[b]].envelope);
  const bool all = (d.flags & STEP_COUNT_OVERLAPS) != 0;
  int cnt = 0;
  const V4<R> ra = d.r[b];
  for (int j0 = b + 1 + (int)threadIdx.x; j0 < end; j0 += 64 * PAIR_U) {
    V4<R> blo[PAIR_U], bhi[PAIR_U];
    pair_boxes(d, j0, end, blo, bhi);
#pragma unroll
    for (int u = 0; u < PAIR_U; ++u) {
      const int j = j0 + 64 * u;
      cnt += (j < end && pair_kept(d, ra, j, alo, ahi, blo[u], bhi[u], env, all)) ? 1 : 0;
    }
  }
  cnt = wave_sum(cnt);
  if (threadIdx.x == 0) d.row_cnt[f] = cnt;
}

template <typename R> __global__ void __launch_bounds__(64) k_pair_write(Dev<R> d) {
  int f = blockIdx.x;
  if (f >= d.nb || aborted(d.err)) return;
  int b = d.bb_body[f];
  int end = d.f_end[f];
  int base = d.row_off[f];
  if (threadIdx.x == 0 && f == d.nb - 1) {
    int np = d.row_off[d.nb];
    if (np > d.cap_pairs) { atomicOr(d.err, ERR_CAP_PAIRS); atomicMax(d.err + 1, np); }
    *d.n_sides = 2 * np;
  }
  if (d.row_off[d.nb] > d.cap_pairs) return;
  V4<R> alo = d.b_lo[b], ahi = d.b_hi[b];
  const R env = R(d.sp[d.b_slot[b]].envelope);
  const bool all = (d.flags & STEP_COUNT_OVERLAPS) != 0;
  const V4<R> ra = d.r[b];
  const int ntA = d.b_broad[b].y;
  for (int j0 = b + 1 + (int)threadIdx.x; j0 - (int)threadIdx.x < end; j0 += 64 * PAIR_U) {
    V4<R> blo[PAIR_U], bhi[PAIR_U];
    pair_boxes(d, j0, end, blo, bhi);
    // chunks in order, lanes in order within a chunk: pairs in column order
#pragma unroll
    for (int u = 0; u < PAIR_U; ++u) {
      const int j = j0 + 64 * u;
      const bool hit = j < end && pair_kept(d, ra, j, alo, ahi, blo[u], bhi[u], env, all);
      const unsigned long long m = __ballot(hit);
      if (hit) {
        int pos = base + __popcll(m & ((1ull << threadIdx.x) - 1ull));
        const int ntB = d.b_broad[j].y;
        d.pairs[pos] = make_int2(b, j);
        d.pair_sz[pos] = ntA + ntB;
        d.pair_items[pos] = n_blocks(ntA) + n_blocks(ntB);
        d.side_cnt[2 * pos] = 0;
        d.side_cnt[2 * pos + 1] = 0;
      }
      base += __popcll(m);
    }
  }
}

// ========================================================== candidate triangles
// Candidate set of X w.r.t. Y: triangles of X whose AABB overlaps the AABB of
// some triangle of Y (the BVH traversal result, collision_detection.py:71,
// kdop_bvh.py:328-345, reduced by np.unique at :234-246).  The unit of work is
// a block of CAND_BLK positions of X's cluster order, one wavefront each: the
// triangles of Y that overlap the block's box are staged in LDS (whole Y
// blocks, then Y clusters, whose box misses it are skipped), then every
// triangle of the block that overlaps Y's root volume is tested against them.
// A block whose box misses Y's root volume only clears its flags.  A large
// body pair is spread over many wavefronts, and each stages only the part of
// Y near its own block.
constexpr int CAND_BLOCK = 256;
constexpr int CAND_WAVES = CAND_BLOCK / 64;

template <typename R> struct CandLds {
  static constexpr int CAP = (sizeof(R) == 4) ? 128 : 64;    // staged triangles per wavefront (4 KB)
};

__device__ __forceinline__ int lanes_below(unsigned long long m) {
  return __builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
}

constexpr int CAND_CL = 64;   // staged clusters per wavefront

// lane src's box corner (x, y, z; w is unused and read as 0)
template <typename R> __device__ __forceinline__ V4<R> shfl_xyz(const V4<R>& v, int src) {
  return V4<R>{__shfl(v.x, src, 64), __shfl(v.y, src, 64), __shfl(v.z, src, 64), R(0)};
}

__device__ __forceinline__ void wave_sync_lds() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Flags of X's positions [x0, x1) (one block) w.r.t. Y: a position is a
// candidate iff its triangle box meets one of Y's triangle boxes.  The
// block's positions stay in registers (CAND_BLK / 64 per lane) while Y's
// triangles that meet the block box are staged in LDS, a run at a time; the
// flags are written once at the end.
// count_all: count every overlapping triangle pair; exact: the reference's
// candidate set, without the reach bound (set while overlaps are counted).
// Returns the number of this lane's positions flagged.
template <typename R>
__device__ int cand_block(const Dev<R>& d, int X, int Y, int4 gX, int4 gY, int slot_base, int x0, int x1,
                          bool count_all, bool exact, V4<R>* lo_s, V4<R>* hi_s, V4<R>* clo_s, V4<R>* chi_s,
                          int2* crange_s, unsigned long long& overlaps) {
  constexpr int XQ = CAND_BLK / 64;
  const int cap = CandLds<R>::CAP;
  const int lane = threadIdx.x & 63;
  // gX, gY: the bodies' b_broad records (first triangle box, shape nt and t_off, scene slot)
  const int tX = gX.x, tY = gY.x, ntY = gY.y, tXo = gX.z;
  const int kY = blk_box(tY, ntY);
  const V4<R> xlo = d.clo[blk_box(tX, gX.y) + x0 / CAND_BLK], xhi = d.chi[blk_box(tX, gX.y) + x0 / CAND_BLK];
  const V4<R> rYlo = d.b_lo[Y], rYhi = d.b_hi[Y];
  // A triangle of X farther from Y's origin than reach_Y + envelope + Y's cell
  // diagonal gives no contact against Y's SDF (the bound of within_reach,
  // taken for the triangle's box, which contains the triangle); such
  // triangles are not candidates unless every overlap is being counted.
  // (Y's reach and cell diagonal ride in the w of its body box)
  const R env = R(d.sp[gX.w].envelope);
  const R lim = (rYlo.w + rYhi.w + env) * R(1.001) + R(1e-4);
  const V4<R> oY = d.r[Y];
  auto beyond = [&](const V4<R>& lo, const V4<R>& hi) {
    const R ex = fmax(fmax(lo.x - oY.x, oY.x - hi.x), R(0));
    const R ey = fmax(fmax(lo.y - oY.y, oY.y - hi.y), R(0));
    const R ez = fmax(fmax(lo.z - oY.z, oY.z - hi.z), R(0));
    return !exact && ex * ex + ey * ey + ez * ez > lim * lim;
  };
  const bool block_in = box_overlap(xlo, xhi, rYlo, rYhi) && !beyond(xlo, xhi);
  // this lane's positions x0 + lane + 64 q: triangle, box, whether it can
  // meet Y at all (its cluster box and its own box meet Y's body box)
  V3<R> xl[XQ], xh[XQ];             // (box corners: x, y, z only)
  int xt[XQ];
  unsigned live = 0, flg = 0;
#pragma unroll
  for (int q = 0; q < XQ; ++q) {
    const int p = x0 + lane + 64 * q;
    xt[q] = -1;
    xl[q] = xyz(xlo); xh[q] = xyz(xhi);
    if (p >= x1) continue;
    xt[q] = d.sT[tXo + p].w;
    if (!block_in) continue;
    const int cl = p / CLUSTER;
    const V4<R> cbl = d.clo[tX + cl], cbh = d.chi[tX + cl];
    const V4<R> lo = d.tlo[tX + p], hi = d.thi[tX + p];
    if (!box_overlap(cbl, cbh, rYlo, rYhi) || beyond(cbl, cbh)) continue;
    if (!box_overlap(lo, hi, rYlo, rYhi) || beyond(lo, hi)) continue;
    xl[q] = xyz(lo); xh[q] = xyz(hi);
    live |= 1u << q;
  }
  unsigned long long cnt = 0;
  if (__ballot(live != 0)) {
    const int nbY = n_blocks(ntY);
    int ywin = 0;                          // window of 64 Y blocks ...
    unsigned long long ymask = 0;          // ... whose boxes meet the block box, not yet staged
    {
      const int j = lane;
      ymask = __ballot(j < nbY && box_overlap(d.clo[kY + j], d.chi[kY + j], xlo, xhi));
    }
    int ybase = 0;
    while (true) {
      // stage the next run of Y triangles that meet the block box, two
      // clusters (64 positions) at a time, with the boxes of their clusters
      int nfill = 0, ncl = 0;
      while (ybase < ntY) {
        if (ybase % CAND_BLK == 0) {
          // the next Y block whose box meets the block box
          int yb = ybase / CAND_BLK;
          while (true) {
            if (yb >= ywin + 64) {
              ywin += 64;
              if (ywin >= nbY) { yb = nbY; break; }
              const int j = ywin + lane;
              ymask = __ballot(j < nbY && box_overlap(d.clo[kY + j], d.chi[kY + j], xlo, xhi));
            }
            const unsigned long long m = yb - ywin >= 64 ? 0ull : ymask & (~0ull << (yb - ywin));
            if (m) { yb = ywin + __ffsll((long long)m) - 1; break; }
            yb = ywin + 64;
          }
          if (yb >= nbY) { ybase = ntY; break; }
          ybase = yb * CAND_BLK;
        }
        // this lane's cluster box and triangle box, loaded together
        const int p = ybase + lane;
        const bool livey = p < ntY;
        const int cl = min(p, ntY - 1) / CLUSTER;
        const V4<R> cbl = d.clo[tY + cl], cbh = d.chi[tY + cl];
        V4<R> lo = cbl, hi = cbh;
        if (livey) { lo = d.tlo[tY + p]; hi = d.thi[tY + p]; }
        const bool oc = livey && box_overlap(cbl, cbh, xlo, xhi);
        const bool in = oc && box_overlap(lo, hi, xlo, xhi);
        const unsigned long long mc = __ballot(oc);
        if (!mc) { ybase += 64; continue; }
        const V4<R> b0lo = shfl_xyz(cbl, 0), b0hi = shfl_xyz(cbh, 0);
        const V4<R> b1lo = shfl_xyz(cbl, 32), b1hi = shfl_xyz(cbh, 32);
        const unsigned long long m = __ballot(in);
        const int n0 = __popcll(m & 0xffffffffull), n1 = __popcll(m >> 32);
        if (nfill + n0 + n1 > cap || ncl + 2 > CAND_CL) break;   // uniform: stage full, test what is staged
        if (in) { const int pos = nfill + lanes_below(m); lo_s[pos] = lo; hi_s[pos] = hi; }
        if (lane == 0) {
          int k = ncl;
          if (n0) { clo_s[k] = b0lo; chi_s[k] = b0hi; crange_s[k] = make_int2(nfill, n0); ++k; }
          if (n1) { clo_s[k] = b1lo; chi_s[k] = b1hi; crange_s[k] = make_int2(nfill + n0, n1); }
        }
        ncl += (n0 ? 1 : 0) + (n1 ? 1 : 0);
        nfill += n0 + n1;
        ybase += 64;
      }
      wave_sync_lds();
      // The lane's positions still to test against the staged run, all of
      // them per staged box: one loop over the staged clusters and their
      // triangles, uniform over the wavefront (every lane reads the same LDS
      // entry, a broadcast).  A cluster is skipped when no position of the
      // wavefront meets its box.  (A loop per position, each walking the
      // clusters its own box meets, diverged across the lanes and repeated
      // the staged reads for each of the four positions.)
      for (int k = 0; k < ncl; ++k) {
        const unsigned want = count_all ? live : (live & ~flg);
        if (!__ballot(want != 0u)) break;                 // every position decided
        const V4<R> cl = clo_s[k], ch = chi_s[k];
        unsigned mk = 0;                                  // positions that meet the cluster box
#pragma unroll
        for (int q = 0; q < XQ; ++q) {
          const V4<R> lo{xl[q].x, xl[q].y, xl[q].z, R(0)}, hi{xh[q].x, xh[q].y, xh[q].z, R(0)};
          if (((want >> q) & 1u) && box_overlap(lo, hi, cl, ch)) mk |= 1u << q;
        }
        if (!__ballot(mk != 0u)) continue;
        const int2 rg = crange_s[k];
        const int end = rg.x + rg.y;
        // four staged boxes per pass, loaded together (indexes clamped into
        // the cluster's run; a clamped repeat is not tested)
        for (int e = rg.x; e < end; e += 4) {
          V4<R> sl[4], sh[4];
#pragma unroll
          for (int j = 0; j < 4; ++j) {
            const int ej = min(e + j, end - 1);
            sl[j] = lo_s[ej]; sh[j] = hi_s[ej];
          }
#pragma unroll
          for (int q = 0; q < XQ; ++q) {
            if (!((mk >> q) & 1u)) continue;
            const V4<R> lo{xl[q].x, xl[q].y, xl[q].z, R(0)}, hi{xh[q].x, xh[q].y, xh[q].z, R(0)};
            unsigned hit = 0;
#pragma unroll
            for (int j = 0; j < 4; ++j) hit |= (box_overlap(lo, hi, sl[j], sh[j]) & (e + j < end)) ? 1u << j : 0u;
            if (hit) {
              flg |= 1u << q;
              if (count_all) cnt += __popc(hit);
              else mk &= ~(1u << q);                      // decided: no further test
            }
          }
          if (!count_all && !__ballot(mk != 0u)) break;
        }
      }
      wave_sync_lds();
      if (ybase >= ntY) break;
      // every position that can meet Y is already a candidate: the rest of Y changes no flag
      if (!count_all && !__ballot((live & ~flg) != 0)) break;
    }
  }
  overlaps += cnt;
#pragma unroll
  for (int q = 0; q < XQ; ++q)
    if (xt[q] >= 0) d.slot_flag[slot_base + xt[q]] = (unsigned char)((flg >> q) & 1u);
  return __popc(flg);
}

// the pair p with item_off[p] <= item < item_off[p + 1] (item_off strictly
// increasing: every pair has at least one block per side); 64-ary search
__device__ __forceinline__ int item_pair(const int* __restrict__ item_off, int np, int item) {
  const int lane = threadIdx.x & 63;
  int lo = 0, hi = np;
  while (hi - lo > 1) {
    const int step = (hi - lo + 63) / 64;
    const int idx = lo + lane * step;
    const unsigned long long m = __ballot(idx < hi && item_off[idx] <= item);
    lo += (63 - __builtin_clzll(m)) * step;      // lane 0 (idx = lo) always qualifies
    hi = min(hi, lo + step);
  }
  return lo;
}

// the pair of every work item, written once per step for the wavefronts of k_cand_flags
template <typename R> __global__ void k_item_map(Dev<R> d) {
  const int np = *d.n_pairs;
  if (np > d.cap_pairs || aborted(d.err)) return;
  for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < np; p += gridDim.x * blockDim.x)
    for (int i = d.item_off[p], e = min(d.item_off[p + 1], d.cap_items); i < e; ++i) d.item_pair_of[i] = p;
}

// one wavefront per item: the items of pair p are A's blocks, then B's
template <typename R> __global__ void __launch_bounds__(CAND_BLOCK) k_cand_flags(Dev<R> d) {
  __shared__ V4<R> lo_s[CAND_WAVES][CandLds<R>::CAP];
  __shared__ V4<R> hi_s[CAND_WAVES][CandLds<R>::CAP];
  __shared__ V4<R> clo_s[CAND_WAVES][CAND_CL];
  __shared__ V4<R> chi_s[CAND_WAVES][CAND_CL];
  __shared__ int2 crange_s[CAND_WAVES][CAND_CL];
  const int np = *d.n_pairs;
  if (np > d.cap_pairs) return;
  if (aborted(d.err)) return;
  const int nslots = d.slot_off[np];
  if (nslots > d.cap_slots) {
    if (blockIdx.x == 0 && threadIdx.x == 0) { atomicOr(d.err, ERR_CAP_SLOTS); atomicMax(d.err + 2, nslots); }
    return;
  }
  const int nitems = d.item_off[np];
  const bool count_all = (d.flags & STEP_COUNT_OVERLAPS) != 0;
  const int wave = threadIdx.x >> 6;
  // Workgroups go to the 8 XCDs in turn (w on XCD w % 8).  Each XCD takes a
  // contiguous run of the items, in proportion to its workgroups, and its
  // wavefronts walk that run together: the blocks of one pair (same two
  // bodies) are then in flight at once on one XCD and share the bodies'
  // boxes in its L2, as do the consecutive pairs of one body.
  const int G = (int)gridDim.x, xcd = (int)blockIdx.x & 7, wx = (int)blockIdx.x >> 3;
  int before = 0;
  for (int x = 0; x < xcd; ++x) before += (G - x + 7) >> 3;
  const int gx = (G - xcd + 7) >> 3;
  const int lo = (int)((long long)nitems * before / G), hi = (int)((long long)nitems * (before + gx) / G);
  for (int item = lo + wx * CAND_WAVES + wave; item < hi; item += gx * CAND_WAVES) {
    const int p = nitems <= d.cap_items ? d.item_pair_of[item] : item_pair(d.item_off, np, item);
    const int2 ab = d.pairs[p];
    const int4 gA = d.b_broad[ab.x], gB = d.b_broad[ab.y];
    const int ntA = gA.y, ntB = gB.y;
    const int k = item - d.item_off[p];
    const bool side0 = k < n_blocks(ntA);
    const int X = side0 ? ab.x : ab.y, Y = side0 ? ab.y : ab.x;
    const int blk = side0 ? k : k - n_blocks(ntA);
    const int x0 = blk * CAND_BLK, x1 = min(x0 + CAND_BLK, side0 ? ntA : ntB);
    unsigned long long ov = 0;
    const int slot_base = d.slot_off[p] + (side0 ? 0 : ntA);
    int nset = cand_block(d, X, Y, side0 ? gA : gB, side0 ? gB : gA, slot_base, x0, x1, count_all && side0,
                          count_all, lo_s[wave], hi_s[wave], clo_s[wave], chi_s[wave], crange_s[wave], ov);
    nset = wave_sum(nset);   // candidates of this block
    if ((threadIdx.x & 63) == 0 && nset) atomicAdd(&d.side_cnt[2 * p + (side0 ? 0 : 1)], nset);
    if (count_all && side0) {
      const unsigned long long tot = wave_sum_u64(ov);
      if ((threadIdx.x & 63) == 0 && tot) atomicAdd(&d.ss[d.b_slot[ab.x]].n_overlaps, tot);
    }
  }
}

// Candidates in slot order (pair, side, triangle index): one wavefront per
// (pair, side) compacts that side's flags from the offset the scan of the
// per-side counts gives it.
template <typename R> __global__ void __launch_bounds__(256) k_cand_write(Dev<R> d) {
  const int np = *d.n_pairs;
  if (np > d.cap_pairs || aborted(d.err)) return;
  const int nc = *d.n_cand;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    if (nc > d.cap_cand) { atomicOr(d.err, ERR_CAP_CAND); atomicMax(d.err + 3, nc); }
    d.pair_cstart[np] = nc;
  }
  if (nc > d.cap_cand) return;
  const int lane = threadIdx.x & 63;
  const int nw = (gridDim.x * blockDim.x) >> 6;
  for (int sd = (blockIdx.x * blockDim.x + threadIdx.x) >> 6; sd < 2 * np; sd += nw) {
    const int p = sd >> 1, side = sd & 1;
    const int2 ab = d.pairs[p];
    const int ntA = d.b_broad[ab.x].y;
    const int nt = side ? d.b_broad[ab.y].y : ntA;
    const int base = d.slot_off[p] + (side ? ntA : 0);
    int off = d.side_off[sd];
    if (side == 0 && lane == 0) d.pair_cstart[p] = off;
    for (int t0 = 0; t0 < nt; t0 += 64) {
      const int t = t0 + lane;
      const bool f = t < nt && d.slot_flag[base + t];
      const unsigned long long m = __ballot(f);
      if (f) {
        const int c = off + __popcll(m & ((1ull << lane) - 1ull));
        d.cand_pair[c] = p;
        d.cand_tri[c] = t | (side << 31);
      }
      off += __popcll(m);
    }
  }
}

// ========================================================== Frank-Wolfe
// Per-scene Frank-Wolfe counters (cell evaluations, candidates), summed by
// k_scene_finalize.  Every wavefront adds to one of FW_SPREAD partial counters
// of its scene, FW_STRIDE words apart: atomics to one address from every
// wavefront of the launch serialise in one L2 channel (one scene of 10^6
// candidates: 16 k wavefronts, 0.15 ms).
template <typename R>
__device__ __forceinline__ void fw_count(const Dev<R>& d, int slot, int wave, unsigned long long evals,
                                         unsigned long long ncand) {
  unsigned long long* acc = d.fw_acc + ((size_t)slot * FW_SPREAD + (wave & (FW_SPREAD - 1))) * FW_STRIDE;
  if (evals) atomicAdd(acc, evals);
  if (ncand) atomicAdd(acc + 1, ncand);
}

// collision_detection.py:120-188 per candidate triangle (one lane each).
// Each wavefront takes a contiguous run of candidates (pairs, and so scenes,
// are contiguous), so a lane's scene changes only at scene boundaries and the
// per-scene counters are merged once per wavefront.
// (six wavefronts per SIMD: FP32 fits 78 VGPRs without spilling; the gathers'
// latency is hidden by occupancy)
template <typename R> __global__ void __launch_bounds__(256, 6) k_frank_wolfe(Dev<R> d) {
  if (aborted(d.err)) return;
  const int nc = *d.n_cand;
  unsigned long long evals = 0;
  int ncand = 0;
  int cur = -1;
  const long long nwaves = (long long)gridDim.x * (blockDim.x >> 6);
  // workgroup w runs on XCD w % 8: number the workgroups XCD by XCD, so that
  // each XCD takes a contiguous run of the candidates (those of one pair read
  // the same SDF cells, then from that XCD's L2)
  const int G = (int)gridDim.x, xcd = (int)blockIdx.x & 7;
  int lwg = (int)blockIdx.x >> 3;
  for (int x = 0; x < xcd; ++x) lwg += (G - x + 7) >> 3;
  const long long wave = (long long)lwg * (blockDim.x >> 6) + (threadIdx.x >> 6);
  const long long chunk = ((nc + nwaves - 1) / nwaves + 63) / 64 * 64;
  const int c_lo = (int)min<long long>(wave * chunk, nc), c_hi = (int)min<long long>(wave * chunk + chunk, nc);
  for (int c = c_lo + (int)(threadIdx.x & 63); c < c_hi; c += 64) {
    int p = d.cand_pair[c];
    int tw = d.cand_tri[c];
    int dir = (tw >> 31) & 1;
    int t = tw & 0x7fffffff;
    int2 ab = d.pairs[p];
    int own = dir ? ab.y : ab.x;     // triangle owner (ContactPoint.bodyB)
    int sdf = dir ? ab.x : ab.y;     // SDF owner (ContactPoint.bodyA)
    const int slot = d.b_slot[own];
    if (slot != cur) {
      if (cur >= 0) fw_count(d, cur, (int)wave, evals, ncand);
      cur = slot;
      evals = 0;
      ncand = 0;
    }
    ++ncand;
    const SceneParams& sp = d.sp[slot];
    const ShapeInfo& sA = d.shapes[d.b_shape[own]];
    GridView<R> g = grid_of(d, d.b_shape[sdf]);
    Quat<R> qA = quat(d.q[own]), qB = quat(d.q[sdf]);
    Quat<R> qBc = conj(qB);
    V3<R> rA = xyz(d.r[own]), rB = xyz(d.r[sdf]);
    int4 tri = d.sT[sA.t_off + t];
    V3<R> P[3];
    int vid[3] = {tri.x, tri.y, tri.z};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      V3<R> vw = qrotate_array(qA, xyz(d.sV[sA.v_off + vid[k]])) + rA;
      P[k] = qrotate_array(qBc, vw - rB);
    }
    // start at the corner with the smallest gradient norm (first on ties)
    R gn0 = norm(grid_gradient(g, P[0]));
    R gn1 = norm(grid_gradient(g, P[1]));
    R gn2 = norm(grid_gradient(g, P[2]));
    int i0 = 0;
    R best = gn0;
    if (gn1 < best) { best = gn1; i0 = 1; }
    if (gn2 < best) { i0 = 2; }
    // corners picked by selects, not by index: an indexed P would live in memory
    V3<R> x = i0 == 0 ? P[0] : (i0 == 1 ? P[1] : P[2]);
    const R tol = R(sp.fw_tol);
    int nfw = 0;
    for (int i = 0; i < sp.fw_max_iter; ++i) {
      V3<R> gr = grid_gradient(g, x);
      ++nfw;
      R o0 = dot(P[0], gr), o1 = dot(P[1], gr), o2 = dot(P[2], gr);
      int vtx = 0;
      R ob = o0;
      if (o1 < ob) { ob = o1; vtx = 1; }
      if (o2 < ob) { ob = o2; vtx = 2; }
      R alpha = R(2) / R(i + 2);
      const V3<R> pv = vtx == 0 ? P[0] : (vtx == 1 ? P[1] : P[2]);
      x = x + alpha * (pv - x);
      if (ob > tol) break;
    }
    bool ok = false;
    V4<R> pw4 = {R(0), R(0), R(0), R(0)}, nw4 = {R(0), R(0), R(0), R(0)};
    evals += 3 + nfw;
    if (grid_inside(g, x, R(0.5))) {
      ++evals;
      R phi = grid_value(g, x);
      if (phi <= R(sp.envelope)) {
        V3<R> nb = grid_gradient(g, x);
        V3<R> nw = qrotate(qB, nb);
        R l = norm(nw);
        if (l > R(0)) {
          V3<R> pw = qrotate(qB, x) + rB;
          pw4 = V4<R>{pw.x, pw.y, pw.z, phi};
          nw4 = V4<R>{nw.x / l, nw.y / l, nw.z / l, R(0)};
          ok = true;
        }
      }
    }
    d.cand_p[c] = pw4;
    d.cand_n[c] = nw4;
    d.cand_keep[c] = ok ? 1 : 0;
  }
  // merge the lanes' totals per scene inside the wave, one atomic per scene
  unsigned long long live = __ballot(cur >= 0);
  while (live) {
    int leader = __ffsll((long long)live) - 1;
    int s0 = __shfl(cur, leader, 64);
    bool mine = cur == s0 && cur >= 0;
    unsigned long long v = mine ? evals : 0ull;
    int u = mine ? ncand : 0;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) { v += __shfl_xor(v, o, 64); u += __shfl_xor(u, o, 64); }
    if ((int)(threadIdx.x & 63) == leader) fw_count(d, s0, (int)wave, v, (unsigned long long)u);
    live &= ~__ballot(mine);
  }
}

// Candidates kept by Frank-Wolfe in candidate order: kidx[keep_scan[c]] = c,
// with their points copied alongside for the duplicate test.
template <typename R> __global__ void k_kept_list(Dev<R> d) {
  if (aborted(d.err)) return;
  const int nc = *d.n_cand;
  if (blockIdx.x == 0 && threadIdx.x == 0) *d.n_kept = d.keep_scan[nc];
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < nc; c += gridDim.x * blockDim.x) {
    if (!d.cand_keep[c]) continue;
    const int k = d.keep_scan[c];
    d.kidx[k] = c;
    d.kpos[k] = d.cand_p[c];
  }
}

// collision_detection.py:255-290: keep the first of contacts with equal
// position within a body pair.  A pair's candidates are contiguous, so its
// kept ones are contiguous in the compacted list too: each is compared with
// the kept contacts of its pair before it.
template <typename R> __global__ void k_dedup(Dev<R> d) {
  if (aborted(d.err)) return;
  const int nk = *d.n_kept;
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < nk; k += gridDim.x * blockDim.x) {
    const int c = d.kidx[k];
    const int k0 = d.keep_scan[d.pair_cstart[d.cand_pair[c]]];
    const V4<R> pc = d.kpos[k];
    bool dup = false;
    for (int e = k0; e < k && !dup; ++e) {
      const V4<R> pe = d.kpos[e];
      dup = pe.x == pc.x && pe.y == pc.y && pe.z == pc.z;
    }
    d.kept_flag[k] = dup ? 0 : 1;
  }
}

template <typename R> __global__ void k_contact_write(Dev<R> d) {
  if (aborted(d.err)) return;
  const int nkept = *d.n_kept;
  const int nk = d.kept_scan[nkept];
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    *d.n_contacts = nk;
    if (nk > d.cap_contacts) { atomicOr(d.err, ERR_CAP_CONTACTS); atomicMax(d.err + 4, nk); }
  }
  if (nk > d.cap_contacts) return;
  const int lane = threadIdx.x & 63;
  // whole wavefronts per iteration, so the per-scene first index, count and
  // largest penetration are merged across the wavefront before the atomics
  for (int base = blockIdx.x * blockDim.x + (threadIdx.x & ~63); base < nkept; base += gridDim.x * blockDim.x) {
    const int j = base + lane;
    int slot = -1, k = 0x7fffffff;
    double pen = 0.0;
    if (j < nkept && d.kept_flag[j]) {
      const int c = d.kidx[j];
      k = d.kept_scan[j];
      int p = d.cand_pair[c];
      int dir = (d.cand_tri[c] >> 31) & 1;
      int2 ab = d.pairs[p];
      int own = dir ? ab.y : ab.x;
      int sdf = dir ? ab.x : ab.y;
      d.c_ia[k] = sdf;                   // ContactPoint(bodyB, bodyA, ...) at collision_detection.py:187
      d.c_ib[k] = own;
      d.c_tri[k] = d.cand_tri[c] & 0x7fffffff;
      d.c_pair[k] = p;
      d.c_p[k] = d.cand_p[c];
      d.c_n[k] = d.cand_n[c];
      slot = d.b_slot[sdf];
      pen = fmax(0.0, -(double)d.cand_p[c].w);
    }
    unsigned long long live = __ballot(slot >= 0);
    while (live) {
      const int leader = __ffsll((long long)live) - 1;
      const int s0 = __shfl(slot, leader, 64);
      const bool mine = slot == s0;
      int kmin = mine ? k : 0x7fffffff, cnt = mine ? 1 : 0;
      double pm = mine ? pen : 0.0;
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) {
        kmin = min(kmin, __shfl_xor(kmin, o, 64));
        cnt += __shfl_xor(cnt, o, 64);
        pm = fmax(pm, __shfl_xor(pm, o, 64));
      }
      if (lane == leader) {
        atomicMin(&d.ss[s0].k_off, kmin);
        atomicAdd(&d.ss[s0].k_cnt, cnt);
        if (pm > 0.0) atomic_max_nonneg(&d.ss[s0].max_pen, pm);
      }
      live &= ~__ballot(mine);
    }
  }
}

// ========================================================== assembly
// contacts.py:31-55, 110-287 -- frame, arms, b, r = 1/diag(J W J^T), mu
template <typename R>
__device__ __forceinline__ void load_W(const Dev<R>& d, int b, R& im, R W[6]) {
  V4<R> a = d.b_W0[b], c = d.b_W1[b];
  W[0] = a.x; W[1] = a.y; W[2] = a.z; W[3] = a.w; W[4] = c.x; W[5] = c.y; im = c.z;
}

template <typename R> __global__ void k_assemble(Dev<R> d) {
  if (aborted(d.err)) return;
  const int nk = *d.n_contacts;
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < nk; k += gridDim.x * blockDim.x) {
    int A = d.c_ia[k], B = d.c_ib[k];
    int slot = d.b_slot[A];
    const SceneParams& sp = d.sp[slot];
    const R dt = scene_dt(d, slot);
    V4<R> p4 = d.c_p[k];
    V3<R> p = xyz(p4);
    V3<R> rA = p - xyz(d.r[A]);
    V3<R> rB = p - xyz(d.r[B]);
    V3<R> s, t, n;
    orthonormal_frame(xyz(d.c_n[k]), s, t, n);
    // rows: (n, s, t, torsion); J_A = -[lin | rA x dir], J_B = +[lin | rB x dir]
    V3<R> dirs[3] = {n, s, t};
    V3<R> vA = xyz(d.v[A]), wA = xyz(d.w[A]), vB = xyz(d.v[B]), wB = xyz(d.w[B]);
    V3<R> dlA = dt * xyz(d.b_du0[A]), daA = dt * xyz(d.b_du1[A]);
    V3<R> dlB = dt * xyz(d.b_du0[B]), daB = dt * xyz(d.b_du1[B]);
    R imA, imB, WA[6], WB[6];
    load_W(d, A, imA, WA);
    load_W(d, B, imB, WB);
    R vrow[4], brow[4], diag[4];
    for (int i = 0; i < 4; ++i) {
      V3<R> lin, angA, angB;
      if (i < 3) {
        lin = dirs[i];
        angA = cross(rA, dirs[i]);
        angB = cross(rB, dirs[i]);
      } else {
        lin = V3<R>{R(0), R(0), R(0)};
        angA = n;
        angB = n;
      }
      // J u summed in the csr column order: A linear, A angular, B linear, B angular
      R vv = -lin.x * vA.x;
      vv += -lin.y * vA.y; vv += -lin.z * vA.z;
      vv += -angA.x * wA.x; vv += -angA.y * wA.y; vv += -angA.z * wA.z;
      vv += lin.x * vB.x; vv += lin.y * vB.y; vv += lin.z * vB.z;
      vv += angB.x * wB.x; vv += angB.y * wB.y; vv += angB.z * wB.z;
      R bj = -lin.x * dlA.x;
      bj += -lin.y * dlA.y; bj += -lin.z * dlA.z;
      bj += -angA.x * daA.x; bj += -angA.y * daA.y; bj += -angA.z * daA.z;
      bj += lin.x * dlB.x; bj += lin.y * dlB.y; bj += lin.z * dlB.z;
      bj += angB.x * daB.x; bj += angB.y * daB.y; bj += angB.z * daB.z;
      vrow[i] = vv;
      brow[i] = bj;
      R dg = imA * dot(lin, lin) + dot(angA, sym_mul(WA, angA)) + imB * dot(lin, lin) + dot(angB, sym_mul(WB, angB));
      diag[i] = dg == R(0) ? R(1) : dg;
    }
    const double* mt = d.mat + 4 * (sp.mat_off + d.b_mat[A] * sp.nmat + d.b_mat[B]);
    R e0 = sp.use_bounce ? R(mt[3]) : R(0);
    R g0 = R(0);
    if (sp.use_pre_stab) {
      R rate = R(sp.gap_reduction) / dt;
      R upper = -R(sp.max_gap) / dt;
      R lower = -R(sp.min_gap);
      if (p4.w < lower && vrow[0] <= R(0)) g0 = fmax(upper, rate * p4.w);
    }
    V4<R> bb;
    bb.x = ((R(1) + e0) * vrow[0] + brow[0]) + g0;
    bb.y = (R(1) * vrow[1] + brow[1]);
    bb.z = (R(1) * vrow[2] + brow[2]);
    bb.w = (R(1) * vrow[3] + brow[3]);
    d.c_rA[k] = V4<R>{rA.x, rA.y, rA.z, R(0)};
    d.c_rB[k] = V4<R>{rB.x, rB.y, rB.z, R(0)};
    d.c_b[k] = bb;
    d.c_r0[k] = V4<R>{R(1) / diag[0], R(1) / diag[1], R(1) / diag[2], R(1) / diag[3]};
    d.c_mu[k] = V4<R>{R(mt[0]), R(mt[1]), R(mt[2]), R(0)};
  }
}

#include "rbh_solve.h"
#include "rbh_hinge.h"

// ========================================================== integration
// state_storage.py:239-302: u += dt W f_ext + W J^T sol (contacts, then
// hinges, steppers.py:64, 82-85); r += v dt; q <- unit(q + (0, w) q dt 0.5);
// fixed bodies untouched.
template <typename R> __global__ void k_integrate(Dev<R> d) {
  int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= d.nb || aborted(d.err)) return;
  int b = d.bb_body[f];
  if (d.b_fixed[b]) return;
  const R dt = scene_dt(d, d.b_slot[b]);
  V4<R> du0 = d.b_du0[b], du1 = d.b_du1[b], s0 = d.ws0[b], s1 = d.ws1[b], h0 = d.hd0[b], h1 = d.hd1[b];
  V4<R> v = d.v[b], w = d.w[b];
  v.x = v.x + ((dt * du0.x + s0.x) + h0.x); v.y = v.y + ((dt * du0.y + s0.y) + h0.y);
  v.z = v.z + ((dt * du0.z + s0.z) + h0.z);
  w.x = w.x + ((dt * du1.x + s1.x) + h1.x); w.y = w.y + ((dt * du1.y + s1.y) + h1.y);
  w.z = w.z + ((dt * du1.z + s1.z) + h1.z);
  V4<R> r = d.r[b];
  r.x = r.x + v.x * dt; r.y = r.y + v.y * dt; r.z = r.z + v.z * dt;
  Quat<R> q = quat(d.q[b]);
  Quat<R> dq = qprod(Quat<R>{R(0), w.x, w.y, w.z}, q);
  Quat<R> qn = {q.s + dq.s * dt * R(0.5), q.x + dq.x * dt * R(0.5), q.y + dq.y * dt * R(0.5), q.z + dq.z * dt * R(0.5)};
  qn = qunit(qn);
  d.v[b] = v; d.w[b] = w; d.r[b] = r;
  d.q[b] = V4<R>{qn.s, qn.x, qn.y, qn.z};
}

// k_step_begin and k_scene_finalize: one wavefront per scene, lane k on the
// scene's k-th partial Frank-Wolfe counter (FW_SPREAD is the wavefront width)
static_assert(FW_SPREAD == 64, "one partial counter per lane");

template <typename R> __device__ __forceinline__ void scene_reset(const Dev<R>& d, int s, int lane) {
  unsigned long long* acc = d.fw_acc + ((size_t)d.bs_slot[s] * FW_SPREAD + lane) * FW_STRIDE;
  acc[0] = 0; acc[1] = 0;
  if (lane) return;
  SceneState& st = d.ss[d.bs_slot[s]];
  st.k_off = 0x7fffffff; st.k_cnt = 0; st.ncolor = 0; st.n_overlaps = 0; st.fw_evals = 0; st.max_pen = 0.0;
  st.solved = 0; st.nat_r = 0.0; st.nat_x = 0.0; st.n_cand = 0; st.sweeps = 0;
  st.post_active = 0; st.post_iters = 0; st.post_exit = 0;
}

// The step's first kernel: clears the error words, the prefix sums'
// look-back state (scan_words words) and the cooperative solve's merit slots
// (n_merit words, none for a batch), resets every scene's counters, and
// resets the island and colouring-unit arrays (union-find parents, island and
// body-contact counters, unit and sweep-lane ranges), which the kernels before
// the island pass do not touch.  Grid-stride over the longest of these.
template <typename R>
__global__ void k_step_begin(Dev<R> d, unsigned long long* scan_state, int scan_words, int* merit_words, int n_merit) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x, G = gridDim.x * blockDim.x;
  if (i < 16) d.err[i] = 0;
  for (int k = i; k < scan_words; k += G) scan_state[k] = 0;
  for (int k = i; k < n_merit; k += G) merit_words[k] = -1;   // the cooperative solve's merit slots: empty
  for (int t = i; t < d.ns * 64; t += G) scene_reset(d, t >> 6, t & 63);
  for (int f = i; f < d.nb; f += G) uf_reset(d, f);
  for (int l = i; l <= d.nb + d.ns; l += G) isl_reset(d, l);
  if (i == 0) d.big_list[0] = 0;
  for (int u = i; u < d.cap_units; u += G) unit_reset(d, u);
}

template <typename R> __global__ void k_scene_finalize(Dev<R> d) {
  const int s = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6), lane = (int)(threadIdx.x & 63);
  if (s >= d.ns || aborted(d.err)) return;
  const unsigned long long* acc = d.fw_acc + ((size_t)d.bs_slot[s] * FW_SPREAD + lane) * FW_STRIDE;
  const unsigned long long ev = wave_sum_u64(acc[0]), nc = wave_sum_u64(acc[1]);
  if (lane) return;
  SceneState& st = d.ss[d.bs_slot[s]];
  st.fw_evals = ev;
  st.n_cand = (int)nc;
  if (st.k_cnt == 0) st.k_off = 0;
  st.step_index += 1;
  const unsigned long long sw = st.solved ? (unsigned long long)st.sweeps : 0ull;
  st.tot_steps += 1;
  st.tot_contact_iters += sw * (unsigned long long)st.k_cnt;
  st.tot_contacts += (unsigned long long)st.k_cnt;
  st.tot_cand += (unsigned long long)st.n_cand;
  st.tot_fw_evals += st.fw_evals;
  st.tot_sweeps += sw;
}

}  // namespace rbh
