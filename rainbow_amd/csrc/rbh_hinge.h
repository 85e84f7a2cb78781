// Hinge joints: the second problem of the prox solve (constraints/hinges.py).
//
// The reference solves the contact problem and the hinge problem in one
// driver loop (gauss_seidel_solver.py:44-94, problems from steppers.py:68-76):
// each sweep updates both, each keeps its own merit, accept/reject and
// r-factors, and the stopping tests use the larger of the two contributions.
// The problems share no unknowns (each sweep recomputes its own w = W J^T x),
// so the hinge problem's sequence of iterates does not depend on the
// contacts.  Here it runs on its own, in two passes of one wavefront per
// scene:
//   pass 0 (before the contact solve, scenes with contacts): max_iter sweeps,
//          recording the hinge contribution of every sweep in h_mer; the
//          contact solve's stopping tests read it;
//   pass 1 (after it): the same sweeps again up to the sweep the joint
//          stopping test ended on -- or, in a scene without contacts, the
//          whole solve with its own stopping tests and residual log -- then
//          W J^T sol per body for the velocity update.
// A hinge sweep is serial over the hinges (Gauss-Seidel); inside one hinge
// the five rows update from the same w (hinges.py:88), one lane per row, and
// the twelve body entries of W J^T delta one lane each.  w, the iterate and
// the accepted solution stay in LDS; the rows of the next HINGE_AHEAD hinges
// are loaded while one updates, so a hinge costs two wave syncs and LDS
// traffic.
// (included inside namespace rbh by rbh_kernels.h)
#pragma once

constexpr int HINGE_BODY_MAX = 1024;   // bodies joined by hinges in one scene (LDS: two V4 each)
constexpr int HINGE_MAX = 1024;        // hinges of one scene (LDS: 5 iterates and 5 accepted values each; FP64 total 144 KB)

// One J row (parent lin, parent ang, child lin, child ang) times body-space
// vectors, term by term in the csr column order of hinges.py:184 (the body of
// lower index first), from zero.
template <typename R>
__device__ __forceinline__ R hinge_rowdot(bool pfirst, const V4<R>* J, V3<R> pl, V3<R> pa, V3<R> cl, V3<R> ca) {
  V4<R> a0 = pfirst ? J[0] : J[2], a1 = pfirst ? J[1] : J[3], a2 = pfirst ? J[2] : J[0], a3 = pfirst ? J[3] : J[1];
  V3<R> b0 = pfirst ? pl : cl, b1 = pfirst ? pa : ca, b2 = pfirst ? cl : pl, b3 = pfirst ? ca : pa;
  R s = R(0);
  s += a0.x * b0.x; s += a0.y * b0.y; s += a0.z * b0.z;
  s += a1.x * b1.x; s += a1.y * b1.y; s += a1.z * b1.z;
  s += a2.x * b2.x; s += a2.y * b2.y; s += a2.z * b2.z;
  s += a3.x * b3.x; s += a3.y * b3.y; s += a3.z * b3.z;
  return s;
}

template <typename R> __device__ __forceinline__ V4<R> v4(V3<R> a) { return V4<R>{a.x, a.y, a.z, R(0)}; }

// Hinges.initialize (hinges.py:28-48): J (:103-185), W J^T, g (:187-218),
// b = J u + J (dt du_ext) + g, diag of J W J^T (0 -> 1), r = 1/diag.
// One workgroup per batch scene, one thread per hinge.
template <typename R> __global__ void __launch_bounds__(64) k_hinge_rows(Dev<R> d) {
  if (aborted(d.err)) return;
  const int slot = d.bs_slot[blockIdx.x];
  const SceneParams& sp = d.sp[slot];
  const R dt = scene_dt(d, slot);
  for (int k = threadIdx.x; k < sp.h_cnt; k += blockDim.x) {
    const int h = sp.h_off + k;
    const int4 id = d.h_ids[h];
    const int P = id.x, C = id.y;
    const Quat<R> qp = quat(d.q[P]), qc = quat(d.q[C]);
    const V3<R> arm_p = qrotate(qp, xyz(d.h_geo[4 * h])), arm_c = qrotate(qc, xyz(d.h_geo[4 * h + 1]));
    V3<R> t, b, n;
    orthonormal_frame(qrotate(qp, xyz(d.h_geo[4 * h + 2])), t, b, n);
    R g[5] = {R(0), R(0), R(0), R(0), R(0)};
    if (sp.use_pre_stab) {
      const V3<R> rp = arm_p + xyz(d.r[P]), rc = arm_c + xyz(d.r[C]);
      const V3<R> u = cross(n, qrotate(qc, xyz(d.h_geo[4 * h + 3])));
      const R rate = R(sp.gap_reduction) / dt;
      g[0] = (rc.x - rp.x) * rate; g[1] = (rc.y - rp.y) * rate; g[2] = (rc.z - rp.z) * rate;
      g[3] = dot(t, u) * rate; g[4] = dot(b, u) * rate;
    }
    R imP, imC, WP[6], WC[6];
    load_W(d, P, imP, WP);
    load_W(d, C, imC, WC);
    const V3<R> z3 = {R(0), R(0), R(0)};
    const V3<R> up_l = xyz(d.v[P]), up_a = xyz(d.w[P]), uc_l = xyz(d.v[C]), uc_a = xyz(d.w[C]);
    const V3<R> dp_l = dt * xyz(d.b_du0[P]), dp_a = dt * xyz(d.b_du1[P]);
    const V3<R> dc_l = dt * xyz(d.b_du0[C]), dc_a = dt * xyz(d.b_du1[C]);
    const bool pfirst = P < C;
    for (int i = 0; i < 5; ++i) {
      V3<R> pl = z3, pa, cl = z3, ca;
      if (i < 3) {
        // rows 0-2: parent [I | -[arm_p]x], child [-I | [arm_c]x]; row i of [a]x
        // is (0, -a.z, a.y), (a.z, 0, -a.x), (-a.y, a.x, 0)
        const V3<R> sp_row = i == 0 ? V3<R>{R(0), -arm_p.z, arm_p.y}
                                    : (i == 1 ? V3<R>{arm_p.z, R(0), -arm_p.x} : V3<R>{-arm_p.y, arm_p.x, R(0)});
        const V3<R> sc_row = i == 0 ? V3<R>{R(0), -arm_c.z, arm_c.y}
                                    : (i == 1 ? V3<R>{arm_c.z, R(0), -arm_c.x} : V3<R>{-arm_c.y, arm_c.x, R(0)});
        pl = V3<R>{R(i == 0), R(i == 1), R(i == 2)};
        cl = -pl;
        pa = -sp_row;
        ca = sc_row;
      } else {
        const V3<R> dir = i == 3 ? t : b;
        pa = dir;
        ca = -dir;
      }
      V4<R> J[4] = {v4(pl), v4(pa), v4(cl), v4(ca)};
      V4<R> WJ[4] = {v4(imP * pl), v4(sym_mul(WP, pa)), v4(imC * cl), v4(sym_mul(WC, ca))};
      const R ju = hinge_rowdot(pfirst, J, up_l, up_a, uc_l, uc_a);
      const R jd = hinge_rowdot(pfirst, J, dp_l, dp_a, dc_l, dc_a);
      const R dg = hinge_rowdot(pfirst, J, xyz(WJ[0]), xyz(WJ[1]), xyz(WJ[2]), xyz(WJ[3]));
      for (int j = 0; j < 4; ++j) {
        d.h_J[20 * h + 4 * i + j] = J[j];
        d.h_WJ[20 * h + 4 * i + j] = WJ[j];
      }
      d.h_b[5 * h + i] = (ju + jd) + g[i];
      d.h_r0[5 * h + i] = R(1) / (dg == R(0) ? R(1) : dg);
    }
  }
}

// w = W J^T x for the scene's hinge bodies (hinges.py:65, :56): one running
// sum per body entry over the body's hinges in hinge order, rows 0..4; x
// holds the scene's rows, row i of hinge h at 5 (h - sp.h_off) + i.
template <typename R>
__device__ __forceinline__ void hinge_body_w(const Dev<R>& d, const SceneParams& sp, const R* x, V4<R>* wl, V4<R>* wa) {
  for (int j = threadIdx.x; j < sp.hb_cnt; j += blockDim.x) {
    const int g = sp.hb_off + j;
    V3<R> sl = {R(0), R(0), R(0)}, sa = {R(0), R(0), R(0)};
    for (int e = d.hb_inc_off[g]; e < d.hb_inc_off[g + 1]; ++e) {
      const int inc = d.hb_inc[e], h = inc >> 1, side = inc & 1;
      for (int i = 0; i < 5; ++i) {
        const R xi = x[5 * (h - sp.h_off) + i];
        const V4<R> l = d.h_WJ[20 * h + 4 * i + 2 * side], a = d.h_WJ[20 * h + 4 * i + 2 * side + 1];
        sl.x += l.x * xi; sl.y += l.y * xi; sl.z += l.z * xi;
        sa.x += a.x * xi; sa.y += a.y * xi; sa.z += a.z * xi;
      }
    }
    wl[j] = v4(sl);
    wa[j] = v4(sa);
  }
}

__device__ __forceinline__ double wave_max_d(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  return v;
}


// Per-hinge constants a lane needs for one Gauss-Seidel update: lanes 0-4
// hold row `lane` (J, b, r0), lanes 0-11 the W J^T entries of body entry
// `lane` for rows 0-4.  Loaded HINGE_AHEAD hinges ahead of their use.
template <typename R> struct HingeRegs {
  int4 id;
  V4<R> J[4];
  R b, r0;
  R wj[5];
};

template <typename R>
__device__ __forceinline__ void hinge_load(const Dev<R>& d, int h, int lane, HingeRegs<R>& g) {
  g.id = d.h_ids[h];
  if (lane < 5) {
    const V4<R>* J = d.h_J + 20 * h + 4 * lane;
    g.J[0] = J[0]; g.J[1] = J[1]; g.J[2] = J[2]; g.J[3] = J[3];
    g.b = d.h_b[5 * h + lane];
    g.r0 = d.h_r0[5 * h + lane];
  }
  if (lane < 12) {
    const int part = lane / 3, c = lane - 3 * part;
    // one element each (a runtime index into a vector register would go through scratch)
    for (int i = 0; i < 5; ++i) g.wj[i] = reinterpret_cast<const R*>(d.h_WJ + 20 * h + 4 * i + part)[c];
  }
}

constexpr int HINGE_AHEAD = 4;   // hinges whose rows are loaded ahead of their update

// One hinge's Gauss-Seidel update on w (LDS) and its iterate x (LDS, rows
// from x0): hinges.py:88 then :101.
template <typename R>
__device__ __forceinline__ void hinge_update(V4<R>* wl, V4<R>* wa, R* x, R* s_delta, int x0, R rs,
                                             const HingeRegs<R>& g, int lane) {
  const int4 id = g.id;
  if (lane < 5) {
    // x_b -= r_b (J w + b)_b with every row of the block from the same w
    const R jw = hinge_rowdot(id.x < id.y, g.J, xyz(wl[id.z]), xyz(wa[id.z]), xyz(wl[id.w]), xyz(wa[id.w]));
    const R xo = x[x0 + lane];
    const R xn = xo - (g.r0 * rs) * (jw + g.b);
    x[x0 + lane] = xn;
    s_delta[lane] = xn - xo;
  }
  // one wavefront: LDS operations stay in issue order, and a fence-free sync
  // keeps the loads of the next hinges in flight (a workgroup barrier would
  // wait for them)
  wave_sync_lds();
  if (lane < 12) {
    // w += W J^T_b delta, one body entry per lane
    const int part = lane / 3, c = lane - 3 * part;
    R t = R(0);
    for (int i = 0; i < 5; ++i) t += g.wj[i] * s_delta[i];
    V4<R>* wv = (part & 1) ? wa : wl;
    R* dst = reinterpret_cast<R*>(wv + (part < 2 ? id.z : id.w)) + c;
    *dst = *dst + t;
  }
  wave_sync_lds();
}

// Bytes of dynamic LDS the hinge solve of a scene needs: w (two V4 per hinge
// body), the iterate x and the accepted solution sol (5 each per hinge).
template <typename R> __host__ __device__ constexpr size_t hinge_lds_bytes(int hb, int nh) {
  return (2 * (size_t)hb * sizeof(V4<R>) + 10 * (size_t)nh * sizeof(R) + 15) & ~(size_t)15;
}

// One wavefront per batch scene.  pass 0 records the hinge merit
// contributions for the contact solve; pass 1 runs to the joint stop and
// writes hd0/hd1 (see the top of this file).  w, x and the accepted
// solution sol live in LDS.
template <typename R> __global__ void __launch_bounds__(64) k_hinge_solve(Dev<R> d, int pass) {
  extern __shared__ __align__(16) unsigned char hinge_smem[];
  __shared__ R s_delta[5];
  if (aborted(d.err)) return;
  const int slot = d.bs_slot[blockIdx.x];
  const SceneParams& sp = d.sp[slot];
  SceneState& st = d.ss[slot];
  if (sp.h_cnt == 0) return;
  const bool contacts = st.k_cnt > 0;
  if (pass == 0 && !contacts) return;
  const bool own = pass == 1 && !contacts;   // the hinge problem alone: its own stopping tests and log
  const int n_it = (pass == 1 && contacts) ? (st.exitcode == 0 ? sp.max_iter : st.iterations + 1) : sp.max_iter;
  V4<R>* wl = reinterpret_cast<V4<R>*>(hinge_smem);
  V4<R>* wa = wl + sp.hb_cnt;
  const int lane = threadIdx.x;
  const int h0 = sp.h_off, h1 = sp.h_off + sp.h_cnt, nrow = 5 * sp.h_cnt;
  R* x = reinterpret_cast<R*>(wa + sp.hb_cnt);   // row 5 (h - h0) + i
  R* sol = x + nrow;
  double* res = d.residuals + sp.res_off;
  double* hm = d.h_mer + sp.res_off;
  for (int i = lane; i < nrow; i += 64) { x[i] = R(0); sol[i] = R(0); }
  if (own)
    for (int i = lane; i < sp.max_iter; i += 64) res[i] = INFINITY;
  __syncthreads();
  SolveCtl ctl;
  ctl.iters = sp.max_iter;
  for (int it = 0; it < n_it; ++it) {
    const R rs = R(ctl.rscale);
    hinge_body_w(d, sp, x, wl, wa);
    // rows of the next HINGE_AHEAD hinges in flight while one updates
    HingeRegs<R> ring[HINGE_AHEAD];
#pragma unroll
    for (int k = 0; k < HINGE_AHEAD; ++k)
      if (h0 + k < h1) hinge_load(d, h0 + k, lane, ring[k]);
    __syncthreads();
    for (int hh = h0; hh < h1; hh += HINGE_AHEAD) {
#pragma unroll
      for (int k = 0; k < HINGE_AHEAD; ++k) {
        const int h = hh + k;
        if (h >= h1) break;
        hinge_update(wl, wa, x, s_delta, 5 * (h - h0), rs, ring[k], lane);
        if (h + HINGE_AHEAD < h1) hinge_load(d, h + HINGE_AHEAD, lane, ring[k]);
      }
    }
    double m = 0.0;
    for (int i = lane; i < nrow; i += 64) m = fmax(m, fabs((double)(x[i] - sol[i])));
    m = wave_max_d(m);
    int dec;
    if (own) {
      dec = ctl.decide(sp, m, lane == 0 ? res : nullptr, it);
    } else {
      dec = ctl.accept_or_reject(sp, m);
      if (pass == 0 && lane == 0) hm[it] = ctl.gm;
    }
    for (int i = lane; i < nrow; i += 64) {
      if (dec & 1) sol[i] = x[i];
      else x[i] = sol[i];
    }
    __syncthreads();
    if (dec & 2) { ctl.iters = it; break; }
  }
  if (pass == 0) return;
  if (own && lane == 0) {
    st.iterations = ctl.iters;
    st.exitcode = ctl.exitcode;
    st.rejects = ctl.rejects;
    st.sweeps = ctl.sweeps;
    st.final_merit = ctl.last_merit;
  }
  // finalize_solution (hinges.py:50-57): W J^T sol per body
  hinge_body_w(d, sp, sol, wl, wa);
  __syncthreads();
  for (int j = lane; j < sp.hb_cnt; j += 64) {
    const int b = d.hb_body[sp.hb_off + j];
    d.hd0[b] = wl[j];
    d.hd1[b] = wa[j];
  }
}
