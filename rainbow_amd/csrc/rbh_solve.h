// Islands, random colouring, colour order and the proximal Gauss-Seidel /
// Jacobi solve.
//
// Islands: connected components of free bodies linked by contacts (fixed
// bodies have a zero inverse-mass block, so contacts only interact through
// free bodies).  Contacts are regrouped by island ("unit order"); islands are
// packed into colouring units of at most unit_T contacts that never straddle
// a scene.  Each unit is coloured by one workgroup with its state in LDS.
//
// The sweep order is fixed by the colours alone (colours in increasing value;
// contacts of one colour share no free body), so two schedules give bitwise
// the same iterates: one workgroup per scene (batched scenes), or one
// workgroup per unit with a grid barrier per sweep for the scene-wide merit
// (a single large scene).
#pragma once
// (included inside namespace rbh by rbh_kernels.h)

// ------------------------------------------------------------ islands
__device__ __forceinline__ int uf_find(int* uf, int a) {
  while (true) {
    int p = __atomic_load_n(&uf[a], __ATOMIC_RELAXED);
    if (p == a) return a;
    int gp = __atomic_load_n(&uf[p], __ATOMIC_RELAXED);
    if (gp != p) atomicCAS(&uf[a], p, gp);   // path halving; parents only decrease
    a = p;
  }
}

__device__ __forceinline__ void uf_union(int* uf, int a, int b) {
  while (true) {
    a = uf_find(uf, a);
    b = uf_find(uf, b);
    if (a == b) return;
    if (a > b) { int t = a; a = b; b = t; }
    if (atomicCAS(&uf[b], b, a) == b) return;
  }
}

template <typename R> __global__ void k_uf_init(Dev<R> d) {
  for (int f = blockIdx.x * blockDim.x + threadIdx.x; f < d.nb; f += gridDim.x * blockDim.x) {
    d.uf[f] = f;
    d.bc_cnt[f] = 0;
    d.bc_fill[f] = 0;
    d.body_claim[f] = -1;
  }
  for (int l = blockIdx.x * blockDim.x + threadIdx.x; l <= d.nb + d.ns; l += gridDim.x * blockDim.x) {
    d.isl_cnt[l] = 0;
    d.isl_fill[l] = 0;
  }
}

template <typename R> __global__ void k_uf_union(Dev<R> d) {
  if (aborted(d.err)) return;
  const int nk = *d.n_contacts;
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < nk; k += gridDim.x * blockDim.x) {
    int A = d.c_ia[k], B = d.c_ib[k];
    if (!d.b_fixed[A] && !d.b_fixed[B]) uf_union(d.uf, d.b2f[A], d.b2f[B]);
  }
}

// island label: flat root f -> f + bs (scene bs's labels are contiguous);
// contacts without a free body share the extra label bs_foff[bs + 1] + bs
template <typename R> __device__ __forceinline__ int contact_label(const Dev<R>& d, int k) {
  int A = d.c_ia[k], B = d.c_ib[k];
  int f = -1;
  if (!d.b_fixed[A]) f = uf_find(d.uf, d.b2f[A]);
  else if (!d.b_fixed[B]) f = uf_find(d.uf, d.b2f[B]);
  if (f >= 0) return f + d.f_bs[f];
  int bs = d.f_bs[d.b2f[A]];
  return d.bs_foff[bs + 1] + bs;
}

template <typename R> __global__ void k_isl_count(Dev<R> d) {
  if (aborted(d.err)) return;
  const int nk = *d.n_contacts;
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < nk; k += gridDim.x * blockDim.x) {
    int lab = contact_label(d, k);
    d.c_lab[k] = lab;
    atomicAdd(&d.isl_cnt[lab], 1);
  }
}

template <typename R> __global__ void k_isl_fill(Dev<R> d) {
  if (aborted(d.err)) return;
  const int nk = *d.n_contacts;
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < nk; k += gridDim.x * blockDim.x) {
    int lab = d.c_lab[k];
    d.uperm[d.isl_off[lab] + atomicAdd(&d.isl_fill[lab], 1)] = k;
  }
}

// units per batch scene: ceil(K_s / unit_T)
template <typename R> __device__ __forceinline__ void scene_pos_range(const Dev<R>& d, int bs, int& p0, int& p1) {
  p0 = d.isl_off[d.bs_foff[bs] + bs];
  p1 = d.isl_off[d.bs_foff[bs + 1] + bs + 1];
}

template <typename R> __global__ void k_unit_count(Dev<R> d) {
  for (int bs = blockIdx.x * blockDim.x + threadIdx.x; bs < d.ns; bs += gridDim.x * blockDim.x) {
    int p0, p1;
    scene_pos_range(d, bs, p0, p1);
    int K = aborted(d.err) ? 0 : p1 - p0;
    d.u_cnt[bs] = (K + d.unit_T - 1) / d.unit_T;
  }
}

template <typename R> __global__ void k_unit_reset(Dev<R> d) {
  for (int u = blockIdx.x * blockDim.x + threadIdx.x; u < d.cap_units; u += gridDim.x * blockDim.x) {
    d.u_start[u] = 0x7fffffff;
    d.u_end[u] = 0;
    d.u_ncol[u] = 0;
    d.u_bs[u] = -1;
  }
}

template <typename R> __global__ void k_unit_assign(Dev<R> d) {
  if (aborted(d.err)) return;
  const int nu = *d.n_units;
  if (nu > d.cap_units) {
    if (blockIdx.x == 0 && threadIdx.x == 0) { atomicOr(d.err, ERR_CAP_UNITS); atomicMax(d.err + 5, nu); }
    return;
  }
  const int nl = d.nb + d.ns;
  for (int lab = blockIdx.x * blockDim.x + threadIdx.x; lab < nl; lab += gridDim.x * blockDim.x) {
    int cnt = d.isl_cnt[lab];
    if (cnt == 0) continue;
    // batch scene of the label: largest bs with bs_foff[bs] + bs <= lab
    int lo = 0, hi = d.ns - 1;
    while (lo < hi) {
      int mid = (lo + hi + 1) >> 1;
      if (d.bs_foff[mid] + mid <= lab) lo = mid; else hi = mid - 1;
    }
    int bs = lo;
    int p0, p1;
    scene_pos_range(d, bs, p0, p1);
    int start = d.isl_off[lab];
    int j = min(d.u_cnt[bs] - 1, (start - p0) / d.unit_T);
    int u = d.u_base[bs] + j;
    atomicMin(&d.u_start[u], start);
    atomicMax(&d.u_end[u], start + cnt);
    d.u_bs[u] = bs;
  }
}

// body -> unit-order positions of its contacts (free bodies only)
template <typename R> __global__ void k_csr_count(Dev<R> d) {
  if (aborted(d.err)) return;
  const int nk = *d.n_contacts;
  for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < nk; p += gridDim.x * blockDim.x) {
    int k = d.uperm[p];
    int A = d.c_ia[k], B = d.c_ib[k];
    if (!d.b_fixed[A]) atomicAdd(&d.bc_cnt[d.b2f[A]], 1);
    if (!d.b_fixed[B]) atomicAdd(&d.bc_cnt[d.b2f[B]], 1);
  }
}

template <typename R> __global__ void k_csr_fill(Dev<R> d) {
  if (aborted(d.err)) return;
  const int nk = *d.n_contacts;
  for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < nk; p += gridDim.x * blockDim.x) {
    int k = d.uperm[p];
    int A = d.c_ia[k], B = d.c_ib[k];
    if (!d.b_fixed[A]) { int f = d.b2f[A]; d.bc_list[d.bc_off[f] + atomicAdd(&d.bc_fill[f], 1)] = p; }
    if (!d.b_fixed[B]) { int f = d.b2f[B]; d.bc_list[d.bc_off[f] + atomicAdd(&d.bc_fill[f], 1)] = p; }
  }
}

// ------------------------------------------------------------ colouring
// Vivace random colouring per island (proximal_contact/blocking.py:62-111 in
// the deterministic form of oracle/coloring.py).  Node = contact; two
// contacts are adjacent if they share a free body.
constexpr int COLOR_BLOCK = 256;
constexpr int UNIT_CAP = 2048;        // contacts of a unit staged in LDS
constexpr int LIST_CAP = 4096;        // body-list entries staged in LDS

struct ColorLds {
  int key[UNIT_CAP];                  // contact index within its scene (random draws, ranks)
  int isl[UNIT_CAP];
  int fa[UNIT_CAP], fb[UNIT_CAP];     // free flat bodies (-1 if fixed)
  unsigned short ao[UNIT_CAP], ac[UNIT_CAP], bo[UNIT_CAP], bcn[UNIT_CAP];
  unsigned short col[UNIT_CAP];
  unsigned short pc[UNIT_CAP];        // palette sizes
  unsigned char done[UNIT_CAP];       // 0 uncoloured, 1 coloured, 2 won this round
  unsigned short list[LIST_CAP];
  int nlist, flag, cnt, mx;
  int sh[COLOR_BLOCK / 64];
};

// colouring state in LDS (unit fits) ...
struct LdsAcc {
  ColorLds* s;
  __device__ int key(int i) const { return s->key[i]; }
  __device__ int isl(int i) const { return s->isl[i]; }
  __device__ int col(int i) const { return s->col[i]; }
  __device__ void set_col(int i, int c) { s->col[i] = (unsigned short)c; }
  __device__ int done(int i) const { return s->done[i]; }
  __device__ void set_done(int i, int v) { s->done[i] = (unsigned char)v; }
  __device__ int pc(int i) const { return s->pc[i]; }
  __device__ void set_pc(int i, int v) { s->pc[i] = (unsigned short)v; }
  __device__ void dec_pc(int i) { atomicSub(reinterpret_cast<unsigned*>(&s->pc[i & ~1]), (i & 1) ? 0x10000u : 1u); }
  template <typename F> __device__ void nb(int i, F&& fn) const {
    int fa = s->fa[i];
    for (int e = s->ao[i], n = s->ao[i] + s->ac[i]; e < n; ++e) { int j = s->list[e]; if (j != i) fn(j); }
    for (int e = s->bo[i], n = s->bo[i] + s->bcn[i]; e < n; ++e) {
      int j = s->list[e];
      if (j == i) continue;
      if (fa >= 0 && (s->fa[j] == fa || s->fb[j] == fa)) continue;   // met through the first body
      fn(j);
    }
  }
};

// ... or in global memory (unit larger than the LDS capacity)
template <typename R> struct GlbAcc {
  const Dev<R>* d;
  int p0;      // first unit-order position of the unit
  int k_off;   // first contact of the scene
  int* col_arr;
  __device__ int kk(int i) const { return d->uperm[p0 + i]; }
  __device__ int key(int i) const { return kk(i) - k_off; }
  __device__ int isl(int i) const { return d->c_lab[kk(i)]; }
  __device__ int col(int i) const { return col_arr[kk(i)]; }
  __device__ void set_col(int i, int c) { col_arr[kk(i)] = c; }
  __device__ int done(int i) const { return d->g_done[p0 + i]; }
  __device__ void set_done(int i, int v) { d->g_done[p0 + i] = (unsigned char)v; }
  __device__ int pc(int i) const { return d->g_pc[p0 + i]; }
  __device__ void set_pc(int i, int v) { d->g_pc[p0 + i] = v; }
  __device__ void dec_pc(int i) { atomicSub(&d->g_pc[p0 + i], 1); }
  __device__ int free_body(int k, bool second) const {
    int b = second ? d->c_ib[k] : d->c_ia[k];
    return d->b_fixed[b] ? -1 : d->b2f[b];
  }
  template <typename F> __device__ void nb(int i, F&& fn) const {
    int k = kk(i);
    int fa = free_body(k, false), fb = free_body(k, true);
    if (fa >= 0)
      for (int e = d->bc_off[fa], n = e + d->bc_cnt[fa]; e < n; ++e) { int j = d->bc_list[e] - p0; if (j != i) fn(j); }
    if (fb >= 0)
      for (int e = d->bc_off[fb], n = e + d->bc_cnt[fb]; e < n; ++e) {
        int j = d->bc_list[e] - p0;
        if (j == i) continue;
        if (fa >= 0) {
          int kj = d->uperm[p0 + j];
          if (free_body(kj, false) == fa || free_body(kj, true) == fa) continue;
        }
        fn(j);
      }
  }
};

__device__ __forceinline__ int pal_select(const unsigned* p, int idx) {
  for (int w = 0; w < PALW; ++w) {
    int c = __popc(p[w]);
    if (idx < c) {
      unsigned m = p[w];
      for (int i = 0; i < idx; ++i) m &= m - 1;   // drop the lowest set bits
      return w * 32 + __ffs(m) - 1;
    }
    idx -= c;
  }
  return -1;
}

__device__ __forceinline__ int block_sum(int v, int* sh) {
  int tot;
  block_exclusive_scan(v, sh, tot);
  return tot;
}

// rounds of blocking.py:78-104 over the K contacts of one unit
template <typename R, typename A>
__device__ void color_rounds(const Dev<R>& d, A& a, int K, unsigned* pal, unsigned long long seed,
                             unsigned long long stepk, int* sh, int* s_flag, int* s_mx) {
  // palettes {0, ..., deg-1}
  for (int i = threadIdx.x; i < K; i += COLOR_BLOCK) {
    int deg = 0;
    a.nb(i, [&](int) { ++deg; });
    unsigned* p = pal + (size_t)i * PALW;
    for (int w = 0; w < PALW; ++w) {
      int lo = w * 32;
      unsigned m = 0;
      if (deg >= lo + 32) m = 0xffffffffu;
      else if (deg > lo) m = (1u << (deg - lo)) - 1u;
      p[w] = m;
    }
    if (deg > PALMAX) atomicOr(d.err, ERR_COLORS);
    a.set_pc(i, min(deg, PALMAX));
    a.set_col(i, 0);
    a.set_done(i, 0);
  }
  __syncthreads();
  int rnd = 0;
  for (; rnd < MAX_ROUNDS; ++rnd) {
    int left = 0;
    for (int i = threadIdx.x; i < K; i += COLOR_BLOCK) {
      if (a.done(i)) continue;
      ++left;
      int n = a.pc(i);
      if (n > 0)
        a.set_col(i, pal_select(pal + (size_t)i * PALW,
                                (int)color_draw(seed, stepk, (unsigned long long)a.key(i), (unsigned long long)rnd, (unsigned)n)));
    }
    if (block_sum(left, sh) == 0) break;
    __syncthreads();
    // keep the draw if no uncoloured neighbour drew the same colour
    for (int i = threadIdx.x; i < K; i += COLOR_BLOCK) {
      if (a.done(i)) continue;
      int ci = a.col(i);
      bool ok = true;
      a.nb(i, [&](int j) { if (a.done(j) != 1 && a.col(j) == ci) ok = false; });
      if (ok) a.set_done(i, 2);
    }
    __syncthreads();
    // winners leave their colour out of the palettes of neighbours uncoloured at round start
    for (int i = threadIdx.x; i < K; i += COLOR_BLOCK) {
      if (a.done(i) != 2) continue;
      int ci = a.col(i);
      unsigned bit = 1u << (ci & 31);
      a.nb(i, [&](int j) {
        if (a.done(j) != 1) {
          unsigned old = atomicAnd(&pal[(size_t)j * PALW + (ci >> 5)], ~bit);
          if (old & bit) a.dec_pc(j);
        }
      });
    }
    __syncthreads();
    // commit winners; island maxima over kept and tentative colours
    for (int i = threadIdx.x; i < K; i += COLOR_BLOCK) d.isl_max[a.isl(i)] = -1;
    if (threadIdx.x == 0) *s_flag = 0;
    __syncthreads();
    for (int i = threadIdx.x; i < K; i += COLOR_BLOCK) {
      if (a.done(i) == 2) a.set_done(i, 1);
      atomicMax(&d.isl_max[a.isl(i)], a.col(i));
      if (!a.done(i) && a.pc(i) == 0) atomicOr(s_flag, 1);
    }
    __syncthreads();
    if (*s_flag) {
      // rare: empty palettes get max+1, max+2, ... in increasing contact order
      // per island; ranks read only done/pc/isl/key, which stay fixed here
      for (int i = threadIdx.x; i < K; i += COLOR_BLOCK) {
        if (a.done(i) || a.pc(i) != 0) continue;
        int il = a.isl(i), ki = a.key(i), rank = 0;
        for (int j = 0; j < K; ++j)
          if (j != i && !a.done(j) && a.pc(j) == 0 && a.isl(j) == il && a.key(j) < ki) ++rank;
        int c = d.isl_max[il] + 1 + rank;
        if (c >= PALMAX) { atomicOr(d.err, ERR_COLORS); continue; }
        pal[(size_t)i * PALW + (c >> 5)] |= 1u << (c & 31);
      }
      __syncthreads();
      for (int i = threadIdx.x; i < K; i += COLOR_BLOCK)
        if (!a.done(i) && a.pc(i) == 0) a.set_pc(i, 1);
      __syncthreads();
    }
  }
  if (rnd >= MAX_ROUNDS) {
    // leftovers get fresh colours max+1, ... in increasing contact order per island
    for (int i = threadIdx.x; i < K; i += COLOR_BLOCK) d.isl_max[a.isl(i)] = -1;
    __syncthreads();
    for (int i = threadIdx.x; i < K; i += COLOR_BLOCK) atomicMax(&d.isl_max[a.isl(i)], a.col(i));
    __syncthreads();
    for (int i = threadIdx.x; i < K; i += COLOR_BLOCK) {
      if (a.done(i)) continue;
      int il = a.isl(i), ki = a.key(i), rank = 0;
      for (int j = 0; j < K; ++j)
        if (j != i && !a.done(j) && a.isl(j) == il && a.key(j) < ki) ++rank;
      int c = d.isl_max[il] + 1 + rank;
      if (c >= PALMAX) atomicOr(d.err, ERR_COLORS);
      a.set_col(i, min(c, PALMAX - 1));
    }
    __syncthreads();
  }
  int mx = 0;
  for (int i = threadIdx.x; i < K; i += COLOR_BLOCK) mx = max(mx, a.col(i) + 1);
  if (threadIdx.x == 0) *s_mx = 0;
  __syncthreads();
  atomicMax(s_mx, mx);
  __syncthreads();
}

template <typename R> __global__ void __launch_bounds__(COLOR_BLOCK) k_color_units(Dev<R> d) {
  __shared__ ColorLds s;
  if (aborted(d.err)) return;
  const int u = blockIdx.x;
  if (u >= *d.n_units) return;
  const int p0 = d.u_start[u], p1 = d.u_end[u];
  if (p1 <= p0) return;
  const int bs = d.u_bs[u];
  const int slot = d.bs_slot[bs];
  const SceneParams& sp = d.sp[slot];
  const SceneState& st = d.ss[slot];
  const int K = p1 - p0;
  unsigned* pal = d.pal + (size_t)p0 * PALW;
  if (sp.scheme == SCHEME_GS) {   // the reference's serial order: contact i is colour i
    for (int i = threadIdx.x; i < K; i += COLOR_BLOCK) {
      int k = d.uperm[p0 + i];
      d.c_color[k] = k - st.k_off;
    }
    if (threadIdx.x == 0) { d.u_ncol[u] = 0; atomicMax(&d.ss[slot].ncolor, st.k_cnt); }
    return;
  }
  const unsigned long long seed = sp.seed, stepk = (unsigned long long)st.step_index;
  bool fits = K <= UNIT_CAP;
  if (fits) {
    // stage contacts and the bodies' contact lists (claimed once per body)
    if (threadIdx.x == 0) { s.nlist = 0; s.flag = 0; }
    __syncthreads();
    for (int i = threadIdx.x; i < K; i += COLOR_BLOCK) {
      int k = d.uperm[p0 + i];
      int A = d.c_ia[k], B = d.c_ib[k];
      int fa = d.b_fixed[A] ? -1 : d.b2f[A];
      int fb = d.b_fixed[B] ? -1 : d.b2f[B];
      s.key[i] = k - st.k_off;
      s.isl[i] = d.c_lab[k];
      s.fa[i] = fa;
      s.fb[i] = fb;
      if (fa >= 0) atomicCAS(&d.body_claim[fa], -1, 2 * i);
      if (fb >= 0) atomicCAS(&d.body_claim[fb], -1, 2 * i + 1);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < K; i += COLOR_BLOCK) {
      for (int side = 0; side < 2; ++side) {
        int f = side ? s.fb[i] : s.fa[i];
        if (f < 0 || d.body_claim[f] != 2 * i + side) continue;
        int n = d.bc_cnt[f];
        int off = atomicAdd(&s.nlist, n);
        if (off + n > LIST_CAP) { atomicOr(&s.flag, 1); continue; }
        for (int e = 0; e < n; ++e) s.list[off + e] = (unsigned short)(d.bc_list[d.bc_off[f] + e] - p0);
        d.body_loc[f] = make_int2(off, n);
      }
    }
    __threadfence_block();
    __syncthreads();
    fits = s.flag == 0;
    if (fits) {
      for (int i = threadIdx.x; i < K; i += COLOR_BLOCK) {
        int2 la = s.fa[i] >= 0 ? d.body_loc[s.fa[i]] : make_int2(0, 0);
        int2 lb = s.fb[i] >= 0 ? d.body_loc[s.fb[i]] : make_int2(0, 0);
        s.ao[i] = (unsigned short)la.x; s.ac[i] = (unsigned short)la.y;
        s.bo[i] = (unsigned short)lb.x; s.bcn[i] = (unsigned short)lb.y;
      }
      __syncthreads();
      LdsAcc a{&s};
      color_rounds(d, a, K, pal, seed, stepk, s.sh, &s.flag, &s.mx);
      for (int i = threadIdx.x; i < K; i += COLOR_BLOCK) d.c_color[d.uperm[p0 + i]] = s.col[i];
    }
  }
  if (!fits) {
    GlbAcc<R> a{&d, p0, st.k_off, d.c_color};
    color_rounds(d, a, K, pal, seed, stepk, s.sh, &s.flag, &s.mx);
  }
  if (threadIdx.x == 0) {
    d.u_ncol[u] = s.mx;
    atomicMax(&d.ss[slot].ncolor, s.mx);
  }
}

// ------------------------------------------------------------ colour order
// Gather one contact into the solver arrays at position `pos`.
template <typename R> __device__ __forceinline__ void gather_contact(const Dev<R>& d, int pos, int k) {
  V4<R> mu = d.c_mu[k];
  V4<R> n = d.c_n[k], rA = d.c_rA[k], rB = d.c_rB[k];
  d.perm[pos] = k;
  d.s_id[pos] = make_int2(d.c_ia[k], d.c_ib[k]);
  d.s_n[pos] = V4<R>{n.x, n.y, n.z, mu.x};
  d.s_rA[pos] = V4<R>{rA.x, rA.y, rA.z, mu.y};
  d.s_rB[pos] = V4<R>{rB.x, rB.y, rB.z, mu.z};
  d.s_b[pos] = d.c_b[k];
  d.s_r0[pos] = d.c_r0[k];
  V4<R> z = {R(0), R(0), R(0), R(0)};
  d.x0[pos] = z;
  d.x1[pos] = z;
}

// counting sort of `K` contacts (source: idx(i)) by colour into [dst, dst + K)
template <typename R, typename Idx>
__device__ void order_by_color(const Dev<R>& d, int K, int nc, int dst, int* coff, Idx idx, int* s_cnt) {
  for (int c = threadIdx.x; c <= nc; c += COLOR_BLOCK) coff[c] = 0;
  for (int c = threadIdx.x; c < nc; c += COLOR_BLOCK) s_cnt[c] = 0;
  __syncthreads();
  for (int i = threadIdx.x; i < K; i += COLOR_BLOCK) atomicAdd(&coff[d.c_color[idx(i)] + 1], 1);
  __syncthreads();
  if (threadIdx.x == 0)
    for (int c = 0; c < nc; ++c) coff[c + 1] += coff[c];
  __syncthreads();
  // the order inside a colour does not change any result: LDS atomics place them
  for (int i = threadIdx.x; i < K; i += COLOR_BLOCK) {
    int k = idx(i);
    int c = d.c_color[k];
    gather_contact(d, dst + coff[c] + atomicAdd(&s_cnt[c], 1), k);
  }
  __syncthreads();
}

// one workgroup per batch scene: scene-order positions [k_off, k_off + K)
template <typename R> __global__ void __launch_bounds__(COLOR_BLOCK) k_order(Dev<R> d) {
  __shared__ int s_cnt[PALMAX];
  const int bs = blockIdx.x;
  if (aborted(d.err)) return;
  const int slot = d.bs_slot[bs];
  const SceneState& st = d.ss[slot];
  const int K = st.k_cnt, k0 = st.k_off;
  if (K == 0) return;
  int* coff = d.col_off + (size_t)bs * (PALMAX + 1);
  if (d.sp[slot].scheme == SCHEME_GS) {
    for (int i = threadIdx.x; i < K; i += COLOR_BLOCK) gather_contact(d, k0 + i, k0 + i);
    return;
  }
  order_by_color(d, K, st.ncolor, k0, coff, [&](int i) { return k0 + i; }, s_cnt);
}

// one workgroup per unit: unit-order positions [u_start, u_end)
template <typename R> __global__ void __launch_bounds__(COLOR_BLOCK) k_order_units(Dev<R> d) {
  __shared__ int s_cnt[PALMAX];
  if (aborted(d.err)) return;
  const int u = blockIdx.x;
  if (u >= *d.n_units) return;
  const int p0 = d.u_start[u], p1 = d.u_end[u];
  if (p1 <= p0) return;
  int* coff = d.col_off + (size_t)u * (PALMAX + 1);
  order_by_color(d, p1 - p0, d.u_ncol[u], p0, coff, [&](int i) { return d.uperm[p0 + i]; }, s_cnt);
}

// ------------------------------------------------------------ solve
constexpr int SOLVE_BLOCK = 256;

template <typename R> struct ContactRow {
  V3<R> lin[4], angA[4], angB[4];
};

template <typename R>
__device__ __forceinline__ void contact_rows(V4<R> n4, V4<R> rA4, V4<R> rB4, ContactRow<R>& J) {
  V3<R> s, t, n;
  orthonormal_frame(xyz(n4), s, t, n);
  V3<R> rA = xyz(rA4), rB = xyz(rB4);
  J.lin[0] = n; J.lin[1] = s; J.lin[2] = t; J.lin[3] = V3<R>{R(0), R(0), R(0)};
  J.angA[0] = cross(rA, n); J.angA[1] = cross(rA, s); J.angA[2] = cross(rA, t); J.angA[3] = n;
  J.angB[0] = cross(rB, n); J.angB[1] = cross(rB, s); J.angB[2] = cross(rB, t); J.angB[3] = n;
}

// contacts.py:77-95: z = x - r (J w + b); x_n = max(0, z_n); friction prox
template <typename R>
__device__ __forceinline__ V4<R> contact_prox(const SceneParams& sp, const ContactRow<R>& J, V4<R> xo, V4<R> bb,
                                              V4<R> rr, V4<R> mu, V3<R> wlA, V3<R> waA, V3<R> wlB, V3<R> waB,
                                              bool use_old_xn) {
  R Jw[4];
  for (int i = 0; i < 4; ++i)
    Jw[i] = -dot(J.lin[i], wlA) - dot(J.angA[i], waA) + dot(J.lin[i], wlB) + dot(J.angB[i], waB);
  R z0 = xo.x - rr.x * (Jw[0] + bb.x);
  R z1 = xo.y - rr.y * (Jw[1] + bb.y);
  R z2 = xo.z - rr.z * (Jw[2] + bb.z);
  R z3 = xo.w - rr.w * (Jw[3] + bb.w);
  R xn = fmax(R(0), z0);
  R xr = use_old_xn ? xo.x : xn;
  R os, ot, orr;
  if (sp.friction == 1) {
    EllipsoidParams ep{sp.ell_tol, sp.ell_expansion, sp.ell_max_iter};
    prox_ellipsoid<R>(z1, z2, z3, mu.x, mu.y, mu.z, xr, ep, os, ot, orr);
  } else {
    prox_sphere<R>(z1, z2, z3, mu.x, xr, os, ot, orr);
  }
  return V4<R>{xn, os, ot, orr};
}

// w_body += W J_body^T dx (contacts.py:107)
template <typename R>
__device__ __forceinline__ void apply_dw(const Dev<R>& d, int body, R sign, const ContactRow<R>& J, const V3<R>* ang,
                                         V4<R> dx) {
  R im, W[6];
  load_W(d, body, im, W);
  R dl[3] = {R(0), R(0), R(0)};
  V3<R> a = {R(0), R(0), R(0)};
  R dd[4] = {dx.x, dx.y, dx.z, dx.w};
  for (int i = 0; i < 4; ++i) {
    R s = sign * dd[i];
    dl[0] += J.lin[i].x * s; dl[1] += J.lin[i].y * s; dl[2] += J.lin[i].z * s;
    a = a + s * ang[i];
  }
  V3<R> da = sym_mul(W, a);
  V4<R> l = d.ww0[body], g = d.ww1[body];
  d.ww0[body] = V4<R>{l.x + im * dl[0], l.y + im * dl[1], l.z + im * dl[2], R(0)};
  d.ww1[body] = V4<R>{g.x + da.x, g.y + da.y, g.z + da.z, R(0)};
}

__device__ __forceinline__ double absmax4(double a, double b, double c, double e) {
  return fmax(fmax(fabs(a), fabs(b)), fmax(fabs(c), fabs(e)));
}

__device__ __forceinline__ double block_max(double v, double* sh) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  if (lane == 0) sh[wid] = v;
  __syncthreads();
  double m = 0.0;
  for (int i = 0; i < (int)(blockDim.x >> 6); ++i) m = fmax(m, sh[i]);
  __syncthreads();
  return m;
}

// One sweep over the contacts at positions [p0, p0 + K) sorted by colour
// (coff: colour offsets, or colour i = position i in serial mode).  Returns
// this thread's max |x_new - x_old|.
template <typename R>
__device__ double sweep_range(const Dev<R>& d, const SceneParams& sp, int p0, int K, int nc, const int* coff,
                              bool serial, const V4<R>* xa, V4<R>* xn, R rs) {
  const bool jacobi = sp.scheme == SCHEME_JACOBI || sp.scheme == SCHEME_JACOBI_HYBRID;
  double mloc = 0.0;
  if (!jacobi) {
    for (int c = 0; c < nc; ++c) {
      const int c_lo = serial ? c : coff[c], c_hi = serial ? c + 1 : coff[c + 1];
      for (int pos = c_lo + (int)threadIdx.x; pos < c_hi; pos += SOLVE_BLOCK) {
        int kk = p0 + pos;
        int2 id = d.s_id[kk];
        V4<R> n4 = d.s_n[kk], rA4 = d.s_rA[kk], rB4 = d.s_rB[kk];
        ContactRow<R> J;
        contact_rows(n4, rA4, rB4, J);
        V4<R> r0 = d.s_r0[kk];
        V4<R> rr = {r0.x * rs, r0.y * rs, r0.z * rs, r0.w * rs};
        V4<R> xo = xa[kk];
        V4<R> mu = {n4.w, rA4.w, rB4.w, R(0)};
        V4<R> x = contact_prox(sp, J, xo, d.s_b[kk], rr, mu, xyz(d.ww0[id.x]), xyz(d.ww1[id.x]), xyz(d.ww0[id.y]),
                               xyz(d.ww1[id.y]), false);
        xn[kk] = x;
        V4<R> dx = {x.x - xo.x, x.y - xo.y, x.z - xo.z, x.w - xo.w};
        mloc = fmax(mloc, absmax4((double)x.x - (double)xo.x, (double)x.y - (double)xo.y,
                                  (double)x.z - (double)xo.z, (double)x.w - (double)xo.w));
        if (!d.b_fixed[id.x]) apply_dw(d, id.x, R(-1), J, J.angA, dx);
        if (!d.b_fixed[id.y]) apply_dw(d, id.y, R(1), J, J.angB, dx);
      }
      __syncthreads();
    }
  } else {
    // jacobi_solver.py:6-57: every contact from the accepted w; then the
    // velocity change is applied colour by colour (deterministic sums)
    const bool hybrid = sp.scheme == SCHEME_JACOBI_HYBRID;
    for (int pos = threadIdx.x; pos < K; pos += SOLVE_BLOCK) {
      int kk = p0 + pos;
      int2 id = d.s_id[kk];
      V4<R> n4 = d.s_n[kk], rA4 = d.s_rA[kk], rB4 = d.s_rB[kk];
      ContactRow<R> J;
      contact_rows(n4, rA4, rB4, J);
      V4<R> r0 = d.s_r0[kk];
      V4<R> rr = {r0.x * rs, r0.y * rs, r0.z * rs, r0.w * rs};
      V4<R> xo = xa[kk];
      V4<R> mu = {n4.w, rA4.w, rB4.w, R(0)};
      V4<R> x = contact_prox(sp, J, xo, d.s_b[kk], rr, mu, xyz(d.ws0[id.x]), xyz(d.ws1[id.x]), xyz(d.ws0[id.y]),
                             xyz(d.ws1[id.y]), !hybrid);
      xn[kk] = x;
      mloc = fmax(mloc, absmax4((double)x.x - (double)xo.x, (double)x.y - (double)xo.y, (double)x.z - (double)xo.z,
                                (double)x.w - (double)xo.w));
    }
    __syncthreads();
    for (int c = 0; c < nc; ++c) {
      const int c_lo = serial ? c : coff[c], c_hi = serial ? c + 1 : coff[c + 1];
      for (int pos = c_lo + (int)threadIdx.x; pos < c_hi; pos += SOLVE_BLOCK) {
        int kk = p0 + pos;
        int2 id = d.s_id[kk];
        ContactRow<R> J;
        contact_rows(d.s_n[kk], d.s_rA[kk], d.s_rB[kk], J);
        V4<R> x = xn[kk], xo = xa[kk];
        V4<R> dx = {x.x - xo.x, x.y - xo.y, x.z - xo.z, x.w - xo.w};
        if (!d.b_fixed[id.x]) apply_dw(d, id.x, R(-1), J, J.angA, dx);
        if (!d.b_fixed[id.y]) apply_dw(d, id.y, R(1), J, J.angB, dx);
      }
      __syncthreads();
    }
  }
  return mloc;
}

// accepted sweep: w_sol <- w_work, else w_work <- w_sol, for the bodies of [p0, p0 + K)
template <typename R> __device__ void settle_w(const Dev<R>& d, int p0, int K, bool accept) {
  for (int pos = threadIdx.x; pos < K; pos += SOLVE_BLOCK) {
    int2 id = d.s_id[p0 + pos];
    for (int side = 0; side < 2; ++side) {
      int b = side ? id.y : id.x;
      if (d.b_fixed[b]) continue;
      if (accept) { d.ws0[b] = d.ww0[b]; d.ws1[b] = d.ww1[b]; }
      else { d.ww0[b] = d.ws0[b]; d.ww1[b] = d.ws1[b]; }
    }
  }
}

// impulses back to contact order; natural-map residual parts (FP64)
template <typename R>
__device__ void finish_range(const Dev<R>& d, const SceneParams& sp, int p0, int K, const V4<R>* xs, double& rloc,
                             double& xmax) {
  for (int pos = threadIdx.x; pos < K; pos += SOLVE_BLOCK) {
    int kk = p0 + pos;
    V4<R> x = xs[kk];
    d.c_imp[d.perm[kk]] = x;
    int2 id = d.s_id[kk];
    V4<R> n4 = d.s_n[kk], rA4 = d.s_rA[kk], rB4 = d.s_rB[kk];
    ContactRow<R> J;
    contact_rows(n4, rA4, rB4, J);
    V4<R> mu = {n4.w, rA4.w, rB4.w, R(0)};
    V4<R> y = contact_prox(sp, J, x, d.s_b[kk], d.s_r0[kk], mu, xyz(d.ws0[id.x]), xyz(d.ws1[id.x]), xyz(d.ws0[id.y]),
                           xyz(d.ws1[id.y]), false);
    rloc = fmax(rloc, absmax4((double)x.x - (double)y.x, (double)x.y - (double)y.y, (double)x.z - (double)y.z,
                              (double)x.w - (double)y.w));
    xmax = fmax(xmax, absmax4(x.x, x.y, x.z, x.w));
  }
}

// gauss_seidel_solver.py:56-97: accept/reject, r-factor, stopping tests.
struct SolveCtl {
  double last_merit = INFINITY, glast = INFINITY, rscale = 1.0, gm = 0.0;
  int acc = 0, rejects = 0, exitcode = 0, iters = 0, sweeps = 0;
  // returns bit0 accept, bit1 stop
  __device__ int decide(const SceneParams& sp, double merit, double* res, int it) {
    int dec = 0;
    if (merit > last_merit) {
      rscale *= sp.nu_reduce;
      gm = fmax(0.0, last_merit);
      ++rejects;
    } else {
      dec |= 1;
      if (last_merit - merit < sp.too_small) rscale *= sp.nu_increase;
      last_merit = merit;
      gm = fmax(0.0, merit);
    }
    if (res) res[it] = gm;
    if (gm < sp.abs_tol) { dec |= 2; exitcode = 1; }
    else if (fabs(gm - glast) < sp.rel_tol * glast) { dec |= 2; exitcode = 2; }
    glast = gm;
    ++sweeps;
    if (dec & 1) acc ^= 1;
    return dec;
  }
};

template <typename R> __device__ void write_solve_state(const Dev<R>& d, SceneState& st, const SolveCtl& c) {
  st.solved = 1;
  st.iterations = c.iters;
  st.exitcode = c.exitcode;
  st.rejects = c.rejects;
  st.sweeps = c.sweeps;
  st.final_merit = c.last_merit;
}

// One workgroup per batch scene (scene-order positions).
template <typename R> __global__ void __launch_bounds__(SOLVE_BLOCK) k_solve(Dev<R> d) {
  __shared__ double sh_red[SOLVE_BLOCK / 64];
  const int bs = blockIdx.x;
  if (aborted(d.err)) return;
  const int slot = d.bs_slot[bs];
  const SceneParams& sp = d.sp[slot];
  SceneState& st = d.ss[slot];
  const int K = st.k_cnt, k0 = st.k_off;
  double* res = d.residuals + sp.res_off;
  for (int i = threadIdx.x; i < sp.max_iter; i += SOLVE_BLOCK) res[i] = INFINITY;
  if (K == 0) {
    if (threadIdx.x == 0) { st.solved = 0; st.iterations = 0; st.exitcode = 0; st.sweeps = 0; st.rejects = 0;
                            st.final_merit = 0.0; }
    return;
  }
  const bool serial = sp.scheme == SCHEME_GS;
  const int nc = st.ncolor;
  const int* coff = d.col_off + (size_t)bs * (PALMAX + 1);
  SolveCtl ctl;
  ctl.iters = sp.max_iter;
  for (int it = 0; it < sp.max_iter; ++it) {
    const V4<R>* xa = ctl.acc ? d.x1 : d.x0;
    V4<R>* xn = ctl.acc ? d.x0 : d.x1;
    double mloc = sweep_range(d, sp, k0, K, nc, coff, serial, xa, xn, R(ctl.rscale));
    double merit = block_max(mloc, sh_red);
    int dec = ctl.decide(sp, merit, threadIdx.x == 0 ? res : nullptr, it);
    settle_w(d, k0, K, dec & 1);
    __syncthreads();
    if (dec & 2) { ctl.iters = it; break; }
  }
  double rloc = 0.0, xmax = 0.0;
  finish_range(d, sp, k0, K, ctl.acc ? d.x1 : d.x0, rloc, xmax);
  double rm = block_max(rloc, sh_red), xm = block_max(xmax, sh_red);
  if (threadIdx.x == 0) {
    write_solve_state(d, st, ctl);
    st.nat_r = rm;
    st.nat_x = xm;
  }
}

// grid-wide barrier for a cooperative launch; a bounded wait that flags an
// error instead of spinning forever
__device__ __forceinline__ bool grid_barrier(unsigned* bar, unsigned nblocks, int* err, int* s_ok) {
  __syncthreads();
  if (threadIdx.x == 0) {
    int ok = 1;
    unsigned g = __hip_atomic_load(&bar[1], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
    __threadfence();
    unsigned arrived = atomicAdd(&bar[0], 1u) + 1u;
    if (arrived == nblocks) {
      __hip_atomic_store(&bar[0], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_fetch_add(&bar[1], 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    } else {
      unsigned long long spins = 0;
      while (__hip_atomic_load(&bar[1], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) == g) {
        __builtin_amdgcn_s_sleep(2);
        if (++spins > (1ull << 26)) { atomicOr(err, ERR_SOLVER); ok = 0; break; }
      }
    }
    __threadfence();
    *s_ok = ok;
  }
  __syncthreads();
  return *s_ok != 0;
}

// A single scene split over units: each workgroup sweeps its units, the
// scene-wide merit meets in a grid barrier once per sweep.
template <typename R> __global__ void __launch_bounds__(SOLVE_BLOCK) k_solve_multi(Dev<R> d) {
  __shared__ double sh_red[SOLVE_BLOCK / 64];
  __shared__ int s_ok;
  if (aborted(d.err)) return;   // uniform: the flag is set before this launch or not at all
  const int slot = d.bs_slot[0];
  const SceneParams& sp = d.sp[slot];
  SceneState& st = d.ss[slot];
  const int K = st.k_cnt;
  const int nu = *d.n_units;
  double* res = d.residuals + sp.res_off;
  if (blockIdx.x == 0)
    for (int i = threadIdx.x; i < sp.max_iter; i += SOLVE_BLOCK) res[i] = INFINITY;
  if (K == 0) {
    if (blockIdx.x == 0 && threadIdx.x == 0) { st.solved = 0; st.iterations = 0; st.exitcode = 0; st.sweeps = 0;
                                               st.rejects = 0; st.final_merit = 0.0; }
    return;
  }
  const bool serial = sp.scheme == SCHEME_GS;
  SolveCtl ctl;
  ctl.iters = sp.max_iter;
  for (int it = 0; it < sp.max_iter; ++it) {
    const V4<R>* xa = ctl.acc ? d.x1 : d.x0;
    V4<R>* xn = ctl.acc ? d.x0 : d.x1;
    double mloc = 0.0;
    for (int u = blockIdx.x; u < nu; u += gridDim.x) {
      int p0 = d.u_start[u], p1 = d.u_end[u];
      if (p1 <= p0) continue;
      const int* coff = d.col_off + (size_t)u * (PALMAX + 1);
      mloc = fmax(mloc, sweep_range(d, sp, p0, p1 - p0, serial ? (p1 - p0) : d.u_ncol[u], coff, serial, xa, xn,
                                    R(ctl.rscale)));
    }
    double m = block_max(mloc, sh_red);
    if (threadIdx.x == 0 && m > 0.0) atomic_max_nonneg(&d.merit[it], m);
    if (!grid_barrier(d.bar, gridDim.x, d.err, &s_ok)) return;
    double merit = __hip_atomic_load(&d.merit[it], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
    int dec = ctl.decide(sp, merit, (blockIdx.x == 0 && threadIdx.x == 0) ? res : nullptr, it);
    for (int u = blockIdx.x; u < nu; u += gridDim.x) {
      int p0 = d.u_start[u], p1 = d.u_end[u];
      if (p1 > p0) settle_w(d, p0, p1 - p0, dec & 1);
    }
    __syncthreads();
    if (dec & 2) { ctl.iters = it; break; }
  }
  double rloc = 0.0, xmax = 0.0;
  for (int u = blockIdx.x; u < nu; u += gridDim.x) {
    int p0 = d.u_start[u], p1 = d.u_end[u];
    if (p1 > p0) finish_range(d, sp, p0, p1 - p0, ctl.acc ? d.x1 : d.x0, rloc, xmax);
  }
  double rm = block_max(rloc, sh_red), xm = block_max(xmax, sh_red);
  if (threadIdx.x == 0) {
    if (rm > 0.0) atomic_max_nonneg(&st.nat_r, rm);
    if (xm > 0.0) atomic_max_nonneg(&st.nat_x, xm);
    if (blockIdx.x == 0) write_solve_state(d, st, ctl);
  }
}


