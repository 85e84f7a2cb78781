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

// reset at the start of the step (k_step_begin): body f's union-find parent
// and contact counters, island label l's counters, unit / sweep lane u
template <typename R> __device__ __forceinline__ void uf_reset(const Dev<R>& d, int f) {
  d.uf[f] = f;
  d.bc_cnt[f] = 0;
  d.bc_fill[f] = 0;
  d.body_claim[f] = -1;
}

template <typename R> __device__ __forceinline__ void isl_reset(const Dev<R>& d, int l) {
  d.isl_cnt[l] = 0;
  d.isl_fill[l] = 0;
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

// ctr[key] += 1 for every active lane of the wavefront; returns the lane's
// old value (its slot).  The lanes that share the first active lane's key
// (on a batch scene of one island: the whole wavefront) take one atomic
// together and ranks among themselves; the others add on their own.  An
// island's contacts counted or placed one atomic each serialised in one L2
// channel.  (Slots within one key come in no fixed order; the callers'
// results do not depend on it.)
__device__ __forceinline__ int wave_add_one(int* ctr, int key, bool active) {
  const int lane = (int)(threadIdx.x & 63);
  const unsigned long long act = __ballot(active);
  if (!act) return 0;
  const int first = __ffsll((unsigned long long)act) - 1;
  const int k0 = __shfl(key, first, 64);
  const unsigned long long m = __ballot(active && key == k0);
  int base = 0;
  if (lane == first) base = atomicAdd(&ctr[k0], __popcll(m));
  base = __shfl(base, first, 64);
  if ((m >> lane) & 1ull) return base + __popcll(m & ((1ull << lane) - 1ull));
  return active ? atomicAdd(&ctr[key], 1) : 0;
}

// one contact per lane; the loop runs while any lane of the wavefront has one
template <typename R> __global__ void k_isl_count(Dev<R> d) {
  if (aborted(d.err)) return;
  const int nk = *d.n_contacts;
  const int lane = (int)(threadIdx.x & 63), stride = (int)(gridDim.x * blockDim.x);
  for (int k0 = (int)(blockIdx.x * blockDim.x + (threadIdx.x & ~63u)); k0 < nk; k0 += stride) {
    const int k = k0 + lane;
    const bool act = k < nk;
    int lab = 0;
    if (act) {
      lab = contact_label(d, k);
      d.c_lab[k] = lab;
    }
    (void)wave_add_one(d.isl_cnt, lab, act);
  }
}

template <typename R> __global__ void k_isl_fill(Dev<R> d) {
  if (aborted(d.err)) return;
  const int nk = *d.n_contacts;
  const int lane = (int)(threadIdx.x & 63), stride = (int)(gridDim.x * blockDim.x);
  for (int k0 = (int)(blockIdx.x * blockDim.x + (threadIdx.x & ~63u)); k0 < nk; k0 += stride) {
    const int k = k0 + lane;
    const bool act = k < nk;
    const int lab = act ? d.c_lab[k] : 0;
    const int slot = wave_add_one(d.isl_fill, lab, act);
    if (act) d.uperm[d.isl_off[lab] + slot] = k;
  }
}

// Colouring units of batch scene bs: windows of unit_T contact positions; an
// island joins the window its first position falls in.  Sweep lanes: the
// same with windows of TL_s = unit_T, or, when unit_split > 0, at least
// ceil(K_s / unit_split) positions, so that a scene solved by one workgroup
// has about one lane per wavefront.  Both are unions of whole islands.
template <typename R> __device__ __forceinline__ void scene_pos_range(const Dev<R>& d, int bs, int& p0, int& p1) {
  p0 = d.isl_off[d.bs_foff[bs] + bs];
  p1 = d.isl_off[d.bs_foff[bs + 1] + bs + 1];
}

template <typename R> __global__ void k_unit_count(Dev<R> d) {
  for (int bs = blockIdx.x * blockDim.x + threadIdx.x; bs < d.ns; bs += gridDim.x * blockDim.x) {
    int p0, p1;
    scene_pos_range(d, bs, p0, p1);
    int K = aborted(d.err) ? 0 : p1 - p0;
    int T = d.unit_T, TL = T;
    if (d.unit_split > 0) TL = max(T, (K + d.unit_split - 1) / d.unit_split);
    d.u_cnt[bs] = (K + T - 1) / T;
    d.u_T[bs] = TL;
    d.l_cnt[bs] = (K + TL - 1) / TL;
  }
}

// k_unit_count and the exclusive sums of the scenes' unit and lane counts
// (u_base, l_base with the totals n_units, n_lanes) in one workgroup, one
// launch instead of three, for batches of up to UNIT_SCAN_MAX scenes.
constexpr int UNIT_SCAN_MAX = 8192;

template <typename R> __global__ void __launch_bounds__(SCAN_BLOCK) k_unit_count_scan(Dev<R> d) {
  __shared__ int sh[SCAN_BLOCK / 64];
  const bool ab = aborted(d.err);
  int cu = 0, cl = 0;
  // base <= ns: the tile that holds u_base[ns] and l_base[ns] (the totals) runs
  for (int base = 0; base <= d.ns; base += SCAN_BLOCK) {
    const int bs = base + (int)threadIdx.x;
    int uc = 0, lc = 0;
    if (bs < d.ns) {
      int p0, p1;
      scene_pos_range(d, bs, p0, p1);
      const int K = ab ? 0 : p1 - p0;
      const int T = d.unit_T;
      int TL = T;
      if (d.unit_split > 0) TL = max(T, (K + d.unit_split - 1) / d.unit_split);
      uc = (K + T - 1) / T;
      lc = (K + TL - 1) / TL;
      d.u_cnt[bs] = uc;
      d.u_T[bs] = TL;
      d.l_cnt[bs] = lc;
    }
    int tu, tl;
    const int eu = block_exclusive_scan(uc, sh, tu);
    const int el = block_exclusive_scan(lc, sh, tl);
    if (bs <= d.ns) {
      d.u_base[bs] = cu + eu;
      d.l_base[bs] = cl + el;
    }
    cu += tu;
    cl += tl;
  }
  if (threadIdx.x == 0) {
    *d.n_units = cu;
    *d.n_lanes = cl;
  }
}

template <typename R> __device__ __forceinline__ void unit_reset(const Dev<R>& d, int u) {
  d.u_start[u] = 0x7fffffff;
  d.u_end[u] = 0;
  d.u_ncol[u] = 0;
  d.u_bs[u] = -1;
  d.u_nbody[u] = 0;
  d.l_start[u] = 0x7fffffff;
  d.l_end[u] = 0;
  d.l_bs[u] = -1;
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
    d.isl_unit[lab] = u;
    int jl = min(d.l_cnt[bs] - 1, (start - p0) / d.u_T[bs]);
    int l = d.l_base[bs] + jl;
    atomicMin(&d.l_start[l], start);
    atomicMax(&d.l_end[l], start + cnt);
    d.l_bs[l] = bs;
  }
}

// Free bodies with contacts lie in exactly one island, so in one unit; each
// gets a slot in its unit (where its velocity change lives while the unit is
// solved; the slot order changes no result).
template <typename R> __global__ void k_unit_bodies(Dev<R> d) {
  if (aborted(d.err) || *d.n_units > d.cap_units) return;
  for (int f = blockIdx.x * blockDim.x + threadIdx.x; f < d.nb; f += gridDim.x * blockDim.x) {
    if (d.bc_cnt[f] == 0) { d.f_loc[f] = -1; continue; }
    int root = uf_find(d.uf, f);
    int u = d.isl_unit[root + d.f_bs[root]];
    d.f_unit[f] = u;
    d.f_loc[f] = atomicAdd(&d.u_nbody[u], 1);
  }
}

template <typename R> __global__ void k_unit_blist(Dev<R> d) {
  if (aborted(d.err) || *d.n_units > d.cap_units) return;
  for (int f = blockIdx.x * blockDim.x + threadIdx.x; f < d.nb; f += gridDim.x * blockDim.x) {
    int l = d.f_loc[f];
    if (l >= 0) d.u_blist[d.u_boff[d.f_unit[f]] + l] = d.bb_body[f];
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
// contacts are adjacent if they share a free body.  Units of up to
// SMALL_UNIT contacts are coloured by one wavefront with the palettes in LDS
// (SMALL_PW words per contact) and, per free body of the unit, bitmaps of the
// colours its contacts drew and won, so that a round costs O(palette words)
// per contact; larger units (and small ones whose palettes or bitmaps would
// not fit) by a workgroup of eight wavefronts, palettes in global memory
// (PALW words per contact; colours of an island stay below twice its size),
// the neighbours found through the bodies' contact lists, the rest of the
// state in LDS up to UNIT_CAP contacts (gfx950: 160 KB of LDS per
// workgroup), in global memory beyond.
constexpr int COLOR_BLOCK = 256;
constexpr int COLOR_UNIT_BLOCK = 512;   // threads colouring one large unit (8 wavefronts)
constexpr int UNIT_CAP = 4096;        // contacts of a large unit staged in LDS
constexpr int LIST_CAP = 8192;        // body-list entries staged in LDS
constexpr int SMALL_UNIT = 256;
constexpr int SMALL_BODY = 32;        // bodies with at most this many contacts are checked entry by entry
constexpr int BODY_BITS = 512;        // words of each per-body colour bitmap array of a small unit
constexpr int SMALL_PW = 4;           // palette words per contact of a small unit (colours below 128)
constexpr int PAIR_HASH = 2 * SMALL_UNIT;   // body-pair table of a small unit's set-up (power of two)
constexpr unsigned short NO_SLOT = 0xffff;

// State of a small unit coloured through per-body colour bitmaps (about 14 KB,
// so that many units share a CU).  Body b of the unit (its slot f_loc) owns
// words [b * bw, (b + 1) * bw) of each bitmap array, bw = palette words in use.
struct BitsColorLds {
  int key[SMALL_UNIT];                // contact index within its scene (draws, ranks)
  int isl[SMALL_UNIT];                // island label
  unsigned short la[SMALL_UNIT], lb[SMALL_UNIT];   // unit slots of the free bodies (NO_SLOT: fixed)
  unsigned short col[SMALL_UNIT], pc[SMALL_UNIT];
  unsigned char done[SMALL_UNIT];     // 0 uncoloured, 1 coloured, 2 won this round
  unsigned pal[SMALL_UNIT * SMALL_PW];
  union {
    struct { unsigned seen[BODY_BITS], dup[BODY_BITS], won[BODY_BITS]; } m;   // rounds
    struct { int cnt[2 * SMALL_UNIT]; unsigned hk[PAIR_HASH]; int hc[PAIR_HASH]; } s;   // set-up
  } u;
  int mx, cmax;
};
static_assert(sizeof(BitsColorLds::u.s) <= sizeof(BitsColorLds::u.m), "set-up tables share the bitmaps' space");

template <int NT, int CAP, int LCAP, int PAL> struct ColorLds {
  // colours a palette can hold (at most PALMAX in either palette layout)
  static constexpr int CC = PALMAX;
  int key[CAP];                       // contact index within its scene (random draws, ranks)
  int fa[CAP], fb[CAP];               // free flat bodies (-1 if fixed)
  unsigned short ao[CAP], ac[CAP], bo[CAP], bcn[CAP];
  unsigned short col[CAP];
  unsigned short pc[CAP];             // palette sizes
  unsigned char done[CAP];            // 0 uncoloured, 1 coloured, 2 won this round
  unsigned short list[LCAP];
  // bodies with more than SMALL_BODY contacts (each holds that many list entries): their lists
  unsigned short bl_off[LCAP / (SMALL_BODY + 1) + 1], bl_cnt[LCAP / (SMALL_BODY + 1) + 1];
  unsigned short ent_off[LCAP], ent_cnt[LCAP]; // per list entry of a smaller body: that body's list (cnt 0: larger body)
  unsigned char lose[CAP];            // drew a colour that another contact of one of its bodies drew
  unsigned seen[NT / 64][CC / 32];    // per wavefront, at the body in hand: colours drawn ...
  unsigned dup[NT / 64][CC / 32];     // ... drawn twice (conflict check), or won (palette removal)
  int nbody;
  unsigned pal[PAL > 0 ? PAL : 1];    // palettes (PAL > 0)
  int nlist, flag, cnt, mx;
  int sh[NT / 64];
};

// colouring state in LDS (unit fits) ...
template <typename S> struct LdsAcc {
  static constexpr bool PER_BODY = true;
  S* s;
  const int* c_lab; const int* uperm; int p0;   // island labels (read only when palettes run empty)
  __device__ int key(int i) const { return s->key[i]; }
  __device__ int isl(int i) const { return c_lab[uperm[p0 + i]]; }
  __device__ int col(int i) const { return s->col[i]; }
  __device__ void set_col(int i, int c) { s->col[i] = (unsigned short)c; }
  __device__ int done(int i) const { return s->done[i]; }
  __device__ void set_done(int i, int v) { s->done[i] = (unsigned char)v; }
  __device__ int pc(int i) const { return s->pc[i]; }
  __device__ void set_pc(int i, int v) { s->pc[i] = (unsigned short)v; }
  __device__ void dec_pc(int i) { atomicSub(reinterpret_cast<unsigned*>(&s->pc[i & ~1]), (i & 1) ? 0x10000u : 1u); }
  // (unrolled so that the list reads of several neighbours are in flight together)
  template <typename F> __device__ void nb(int i, F&& fn) const {
    int fa = s->fa[i];
#pragma unroll 8
    for (int e = s->ao[i], n = s->ao[i] + s->ac[i]; e < n; ++e) { int j = s->list[e]; if (j != i) fn(j); }
#pragma unroll 8
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
  static constexpr bool PER_BODY = false;
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

// idx-th set bit of a palette of pw words, or -1
// (palettes of large units live in global memory and are changed by other
// wavefronts' atomics in L2: read them past the vector L1)
__device__ __forceinline__ unsigned pal_word(const unsigned* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ int pal_select(const unsigned* p, int pw, int idx) {
  for (int w = 0; w < pw; ++w) {
    const unsigned pwv = pal_word(p + w);
    int c = __popc(pwv);
    if (idx < c) {
      unsigned m = pwv;
      for (int i = 0; i < idx; ++i) m &= m - 1;   // drop the lowest set bits
      return w * 32 + __ffs(m) - 1;
    }
    idx -= c;
  }
  return -1;
}

// sum over the NT threads of a workgroup (sh: NT / 64 ints)
template <int NT> __device__ __forceinline__ int block_sum(int v, int* sh) {
  v = wave_sum(v);
  if (NT == 64) return v;
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  int tot = 0;
#pragma unroll
  for (int w = 0; w < NT / 64; ++w) tot += sh[w];
  __syncthreads();
  return tot;
}

// rounds of blocking.py:78-104 over the K contacts of one unit by NT threads;
// pal: pw words per contact
template <int NT, typename R, typename A>
__device__ void color_rounds(const Dev<R>& d, A& a, int K, unsigned* pal, int pw, unsigned long long seed,
                             unsigned long long stepk, int* sh, int* s_flag, int* s_mx) {
  const int pcap = 32 * pw;
  // words of the palettes that can hold a colour: up to the largest degree,
  // raised when an empty palette grows (kept in *s_mx until the rounds end)
  if (threadIdx.x == 0) *s_mx = 0;
  if constexpr (A::PER_BODY) {
    auto* S = a.s;
    constexpr int CW = sizeof(S->seen[0]) / sizeof(unsigned);
    for (int k = threadIdx.x & 63; k < CW; k += 64) { S->seen[threadIdx.x >> 6][k] = 0; S->dup[threadIdx.x >> 6][k] = 0; }
  }
  __syncthreads();
  // palettes {0, ..., deg-1}
  for (int i = threadIdx.x; i < K; i += NT) {
    int deg = 0;
    a.nb(i, [&](int) { ++deg; });
    if (deg > 0) atomicMax(s_mx, min(deg, pcap));
    if constexpr (A::PER_BODY) a.s->lose[i] = 0;
    unsigned* p = pal + (size_t)i * pw;
    for (int w = 0; w < pw; ++w) {
      int lo = w * 32;
      unsigned m = 0;
      if (deg >= lo + 32) m = 0xffffffffu;
      else if (deg > lo) m = (1u << (deg - lo)) - 1u;
      p[w] = m;
    }
    if (deg > pcap) atomicOr(d.err, ERR_COLORS);
    a.set_pc(i, min(deg, pcap));
    a.set_col(i, 0);
    a.set_done(i, 0);
  }
  __syncthreads();
  int pwe = (*s_mx + 31) / 32;   // palette words in use
  int rnd = 0;
  for (; rnd < MAX_ROUNDS; ++rnd) {
    int left = 0;
    for (int i = threadIdx.x; i < K; i += NT) {
      if (a.done(i)) continue;
      ++left;
      int n = a.pc(i);
      if (n > 0)
        a.set_col(i, pal_select(pal + (size_t)i * pw, pwe,
                                (int)color_draw(seed, stepk, (unsigned long long)a.key(i), (unsigned long long)rnd, (unsigned)n)));
    }
    if (block_sum<NT>(left, sh) == 0) break;
    __syncthreads();
    // keep the draw if no uncoloured neighbour drew the same colour
    if constexpr (A::PER_BODY) {
      // neighbours are the contacts sharing a free body: a draw is lost when
      // another uncoloured contact of one of its bodies drew the same colour.
      // Wavefronts take whole bodies and sweep their contact lists 64 at a time.
      auto* S = a.s;
      const int w = threadIdx.x >> 6, lane = threadIdx.x & 63, pwr = max(pwe, 1);
      unsigned* seen = S->seen[w];
      unsigned* dup = S->dup[w];
      // smaller bodies: every list entry compares its draw with its body's others
      for (int e = threadIdx.x; e < S->nlist; e += NT) {
        const int cnt = S->ent_cnt[e];
        if (!cnt) continue;
        const int j = S->list[e];
        if (a.done(j)) continue;
        const int c = a.col(j), o = S->ent_off[e];
        for (int e2 = o; e2 < o + cnt; ++e2) {
          const int j2 = S->list[e2];
          if (j2 != j && !a.done(j2) && a.col(j2) == c) { S->lose[j] = 1; break; }
        }
      }
      // larger bodies: a wavefront per body
      for (int b = w; b < S->nbody; b += NT / 64) {
        const int off = S->bl_off[b], n = S->bl_cnt[b];
        for (int e = lane; e < n; e += 64) {
          const int j = S->list[off + e];
          if (a.done(j)) continue;
          const int c = a.col(j);
          const unsigned bit = 1u << (c & 31);
          if (atomicOr(&seen[c >> 5], bit) & bit) atomicOr(&dup[c >> 5], bit);
        }
        wave_sync_lds();
        for (int e = lane; e < n; e += 64) {
          const int j = S->list[off + e];
          if (a.done(j)) continue;
          const int c = a.col(j);
          if (dup[c >> 5] & (1u << (c & 31))) S->lose[j] = 1;
        }
        wave_sync_lds();
        for (int k = lane; k < pwr; k += 64) { seen[k] = 0; dup[k] = 0; }
        wave_sync_lds();
      }
      __syncthreads();
      for (int i = threadIdx.x; i < K; i += NT) {
        if (a.done(i)) continue;
        if (!S->lose[i]) a.set_done(i, 2);
        S->lose[i] = 0;
      }
    } else {
      for (int i = threadIdx.x; i < K; i += NT) {
        if (a.done(i)) continue;
        int ci = a.col(i);
        bool ok = true;
        a.nb(i, [&](int j) { if (a.done(j) != 1 && a.col(j) == ci) ok = false; });
        if (ok) a.set_done(i, 2);
      }
    }
    __syncthreads();
    // winners leave their colour out of the palettes of neighbours uncoloured
    // at round start (no-return atomics: nothing waits on them) ...
    if constexpr (A::PER_BODY) {
      // per body: the colours won by its contacts leave the palettes of its
      // uncoloured contacts (a winner's own palette is no longer read)
      auto* S = a.s;
      const int w = threadIdx.x >> 6, lane = threadIdx.x & 63, pwr = max(pwe, 1);
      unsigned* won = S->dup[w];
      // smaller bodies: every uncoloured list entry clears the colours won by its body's others
      for (int e = threadIdx.x; e < S->nlist; e += NT) {
        const int cnt = S->ent_cnt[e];
        if (!cnt) continue;
        const int j = S->list[e];
        if (a.done(j)) continue;
        const int o = S->ent_off[e];
        for (int e2 = o; e2 < o + cnt; ++e2) {
          const int j2 = S->list[e2];
          if (a.done(j2) != 2) continue;
          const int c = a.col(j2);
          (void)__hip_atomic_fetch_and(&pal[(size_t)j * pw + (c >> 5)], ~(1u << (c & 31)), __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT);
        }
      }
      // larger bodies: a wavefront per body
      for (int b = w; b < S->nbody; b += NT / 64) {
        const int off = S->bl_off[b], n = S->bl_cnt[b];
        bool any = false;
        for (int e = lane; e < n; e += 64) {
          const int j = S->list[off + e];
          if (a.done(j) != 2) continue;
          const int c = a.col(j);
          atomicOr(&won[c >> 5], 1u << (c & 31));
          any = true;
        }
        if (!__ballot(any)) continue;
        wave_sync_lds();
        for (int e = lane; e < n; e += 64) {
          const int j = S->list[off + e];
          if (a.done(j)) continue;
          for (int k = 0; k < pwr; ++k) {
            const unsigned m = won[k];
            if (m) (void)__hip_atomic_fetch_and(&pal[(size_t)j * pw + k], ~m, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          }
        }
        wave_sync_lds();
        for (int k = lane; k < pwr; k += 64) won[k] = 0;
        wave_sync_lds();
      }
    } else {
      for (int i = threadIdx.x; i < K; i += NT) {
        if (a.done(i) != 2) continue;
        int ci = a.col(i);
        unsigned bit = 1u << (ci & 31);
        a.nb(i, [&](int j) {
          if (a.done(j) != 1)
            (void)__hip_atomic_fetch_and(&pal[(size_t)j * pw + (ci >> 5)], ~bit, __ATOMIC_RELAXED,
                                         __HIP_MEMORY_SCOPE_AGENT);
        });
      }
    }
    __syncthreads();
    // ... and the palette sizes of contacts still uncoloured are recounted
    // (a size always equals its palette's popcount)
    for (int i = threadIdx.x; i < K; i += NT) {
      if (a.done(i)) continue;
      const unsigned* p = pal + (size_t)i * pw;
      int n = 0;
      for (int w = 0; w < pwe; ++w) n += __popc(pal_word(p + w));
      a.set_pc(i, n);
    }
    __syncthreads();
    // commit winners
    if (threadIdx.x == 0) *s_flag = 0;
    __syncthreads();
    int moving = 0;   // a winner, or an uncoloured node whose next draw is not forced
    for (int i = threadIdx.x; i < K; i += NT) {
      if (a.done(i) == 2) { a.set_done(i, 1); moving = 1; }
      if (!a.done(i) && a.pc(i) == 0) atomicOr(s_flag, 1);
      if (!a.done(i) && a.pc(i) != 1) moving = 1;
    }
    __syncthreads();
    // No winner and every uncoloured palette a single colour: every later
    // round draws the same colours and keeps none, so the rounds up to
    // MAX_ROUNDS would change nothing; go to the leftovers directly.
    if (block_sum<NT>(moving, sh) == 0) { rnd = MAX_ROUNDS; break; }
    if (*s_flag) {
      // rare: empty palettes get max+1, max+2, ... in increasing contact order
      // per island (max over the island's kept and tentative colours); ranks
      // read only done/pc/isl/key, which stay fixed here
      for (int i = threadIdx.x; i < K; i += NT) d.isl_max[a.isl(i)] = -1;
      __syncthreads();
      for (int i = threadIdx.x; i < K; i += NT) atomicMax(&d.isl_max[a.isl(i)], a.col(i));
      __syncthreads();
      for (int i = threadIdx.x; i < K; i += NT) {
        if (a.done(i) || a.pc(i) != 0) continue;
        int il = a.isl(i), ki = a.key(i), rank = 0;
        for (int j = 0; j < K; ++j)
          if (j != i && !a.done(j) && a.pc(j) == 0 && a.isl(j) == il && a.key(j) < ki) ++rank;
        int c = d.isl_max[il] + 1 + rank;
        if (c >= pcap) { atomicOr(d.err, ERR_COLORS); continue; }
        (void)__hip_atomic_fetch_or(&pal[(size_t)i * pw + (c >> 5)], 1u << (c & 31), __ATOMIC_RELAXED,
                                    __HIP_MEMORY_SCOPE_AGENT);
        atomicMax(s_mx, c + 1);
      }
      __syncthreads();
      pwe = (*s_mx + 31) / 32;
      for (int i = threadIdx.x; i < K; i += NT)
        if (!a.done(i) && a.pc(i) == 0) a.set_pc(i, 1);
      __syncthreads();
    }
  }
  if (rnd >= MAX_ROUNDS) {
    // leftovers get fresh colours max+1, ... in increasing contact order per island
    for (int i = threadIdx.x; i < K; i += NT) d.isl_max[a.isl(i)] = -1;
    __syncthreads();
    for (int i = threadIdx.x; i < K; i += NT) atomicMax(&d.isl_max[a.isl(i)], a.col(i));
    __syncthreads();
    for (int i = threadIdx.x; i < K; i += NT) {
      if (a.done(i)) continue;
      int il = a.isl(i), ki = a.key(i), rank = 0;
      for (int j = 0; j < K; ++j)
        if (j != i && !a.done(j) && a.isl(j) == il && a.key(j) < ki) ++rank;
      int c = d.isl_max[il] + 1 + rank;
      if (c >= pcap) atomicOr(d.err, ERR_COLORS);
      a.set_col(i, min(c, pcap - 1));
    }
    __syncthreads();
  }
  int mx = 0;
  for (int i = threadIdx.x; i < K; i += NT) mx = max(mx, a.col(i) + 1);
  if (threadIdx.x == 0) *s_mx = 0;
  __syncthreads();
  atomicMax(s_mx, mx);
  __syncthreads();
}

// one unit by NT threads with the state S in LDS (global fallback when its
// body lists do not fit)
template <typename R, int NT, int CAP, int LCAP, int PAL>
__device__ void color_unit(const Dev<R>& d, int u, ColorLds<NT, CAP, LCAP, PAL>& s) {
  const int p0 = d.u_start[u], p1 = d.u_end[u];
  const int bs = d.u_bs[u];
  const int slot = d.bs_slot[bs];
  const SceneParams& sp = d.sp[slot];
  const SceneState& st = d.ss[slot];
  const int K = p1 - p0;
  if (sp.scheme == SCHEME_GS) {   // the reference's serial order: contact i is colour i
    for (int i = threadIdx.x; i < K; i += NT) {
      int k = d.uperm[p0 + i];
      d.c_color[k] = k - st.k_off;
    }
    if (threadIdx.x == 0) { d.u_ncol[u] = 0; atomicMax(&d.ss[slot].ncolor, st.k_cnt); }
    return;
  }
  const unsigned long long seed = sp.seed, stepk = (unsigned long long)st.step_index;
  bool fits = K <= min(CAP, d.color_cap);
  if (fits) {
    // stage contacts and the bodies' contact lists (claimed once per body)
    if (threadIdx.x == 0) { s.nlist = 0; s.flag = 0; s.nbody = 0; }
    __syncthreads();
    for (int i = threadIdx.x; i < K; i += NT) {
      int k = d.uperm[p0 + i];
      int A = d.c_ia[k], B = d.c_ib[k];
      int fa = d.b_fixed[A] ? -1 : d.b2f[A];
      int fb = d.b_fixed[B] ? -1 : d.b2f[B];
      s.key[i] = k - st.k_off;
      s.fa[i] = fa;
      s.fb[i] = fb;
      if (fa >= 0) atomicCAS(&d.body_claim[fa], -1, 2 * i);
      if (fb >= 0) atomicCAS(&d.body_claim[fb], -1, 2 * i + 1);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < K; i += NT) {
      for (int side = 0; side < 2; ++side) {
        int f = side ? s.fb[i] : s.fa[i];
        if (f < 0 || d.body_claim[f] != 2 * i + side) continue;
        int n = d.bc_cnt[f];
        int off = atomicAdd(&s.nlist, n);
        if (off + n > LCAP) { atomicOr(&s.flag, 1); continue; }
        for (int e = 0; e < n; ++e) s.list[off + e] = (unsigned short)(d.bc_list[d.bc_off[f] + e] - p0);
        d.body_loc[f] = make_int2(off, n);
        if (n > SMALL_BODY) {
          const int b = atomicAdd(&s.nbody, 1);   // such a body holds more than SMALL_BODY of the <= LCAP entries
          s.bl_off[b] = (unsigned short)off;
          s.bl_cnt[b] = (unsigned short)n;
        }
        for (int e = 0; e < n; ++e) {
          s.ent_off[off + e] = (unsigned short)off;
          s.ent_cnt[off + e] = (unsigned short)(n > SMALL_BODY ? 0 : n);
        }
      }
    }
    __threadfence_block();
    __syncthreads();
    fits = s.flag == 0;
    if (fits) {
      for (int i = threadIdx.x; i < K; i += NT) {
        int2 la = s.fa[i] >= 0 ? d.body_loc[s.fa[i]] : make_int2(0, 0);
        int2 lb = s.fb[i] >= 0 ? d.body_loc[s.fb[i]] : make_int2(0, 0);
        s.ao[i] = (unsigned short)la.x; s.ac[i] = (unsigned short)la.y;
        s.bo[i] = (unsigned short)lb.x; s.bcn[i] = (unsigned short)lb.y;
      }
      __syncthreads();
      LdsAcc<ColorLds<NT, CAP, LCAP, PAL>> a{&s, d.c_lab, d.uperm, p0};
      const int pw = (2 * K + 2 + 31) / 32;
      if (PAL > 0 && K * pw <= PAL)
        color_rounds<NT>(d, a, K, s.pal, pw, seed, stepk, s.sh, &s.flag, &s.mx);
      else
        color_rounds<NT>(d, a, K, d.pal + (size_t)p0 * PALW, PALW, seed, stepk, s.sh, &s.flag, &s.mx);
      for (int i = threadIdx.x; i < K; i += NT) d.c_color[d.uperm[p0 + i]] = s.col[i];
    }
  }
  if (!fits) {
    GlbAcc<R> a{&d, p0, st.k_off, d.c_color};
    color_rounds<NT>(d, a, K, d.pal + (size_t)p0 * PALW, PALW, seed, stepk, s.sh, &s.flag, &s.mx);
  }
  if (threadIdx.x == 0) {
    if (s.mx > 2 * K + 1) atomicOr(d.err, ERR_COLORS);   // beyond the lane's colour offset space
    d.u_ncol[u] = s.mx;
    atomicMax(&d.ss[slot].ncolor, s.mx);
  }
}

static_assert(sizeof(ColorLds<COLOR_UNIT_BLOCK, UNIT_CAP, LIST_CAP, 0>) <= 160 * 1024 - 2048,
              "large-unit colouring state must fit a workgroup's LDS");

// idx-th set bit of an LDS palette of pwe words (the draw's colour), or -1
__device__ __forceinline__ int lds_select(const unsigned* p, int pwe, int idx) {
  for (int w = 0; w < pwe; ++w) {
    unsigned m = p[w];
    const int c = __popc(m);
    if (idx < c) {   // halve the word until the bit is found
      int pos = 0, lo;
      lo = __popc(m & 0xffffu); if (idx >= lo) { idx -= lo; m >>= 16; pos += 16; }
      lo = __popc(m & 0xffu);   if (idx >= lo) { idx -= lo; m >>= 8; pos += 8; }
      lo = __popc(m & 0xfu);    if (idx >= lo) { idx -= lo; m >>= 4; pos += 4; }
      lo = __popc(m & 0x3u);    if (idx >= lo) { idx -= lo; m >>= 2; pos += 2; }
      if (idx >= (int)(m & 1u)) pos += 1;
      return w * 32 + pos;
    }
    idx -= c;
  }
  return -1;
}

__device__ __forceinline__ unsigned pair_hash(unsigned key) { return (key * 2654435761u) >> 23; }   // 9 bits

// The rounds of color_rounds for a unit of up to SMALL_UNIT contacts by one
// wavefront, with a contact's neighbours met through per-body bitmaps instead
// of its bodies' contact lists:
//  * degree = (contacts of body a - 1) + (contacts of body b - 1) - (other
//    contacts between a and b), from per-body counts and a body-pair table;
//  * a draw is lost iff one of its free bodies has the colour in `dup` (drawn
//    by two of its uncoloured contacts in this round);
//  * the colours won on either free body leave the palette of a contact still
//    uncoloured (its palette is read and written by its own lane only).
// The same draws, winners, palettes, fresh colours and leftovers as
// color_rounds, so the same colours.  Palettes hold colours below
// 32 SMALL_PW (leftovers may take larger ones: they need no palette).
// Returns false, having written nothing outside LDS, when the unit is not for
// this path (colouring state forced to global memory) or outgrows it (a
// palette colour of 32 SMALL_PW or more, bitmaps beyond color_bits words);
// color_unit then colours it from the start.
template <typename R>
__device__ bool color_small_bits(const Dev<R>& d, int u, BitsColorLds& s) {
  const int p0 = d.u_start[u], p1 = d.u_end[u], K = p1 - p0;
  const int bs = d.u_bs[u], slot = d.bs_slot[bs];
  const SceneParams& sp = d.sp[slot];
  const SceneState& st = d.ss[slot];
  const int nbu = d.u_nbody[u];
  const int lane = threadIdx.x;
  const int k_off = st.k_off;
  if (sp.scheme == SCHEME_GS) {   // the reference's serial order: contact i is colour i
    for (int i = lane; i < K; i += 64) {
      const int k = d.uperm[p0 + i];
      d.c_color[k] = k - k_off;
    }
    if (lane == 0) { d.u_ncol[u] = 0; atomicMax(&d.ss[slot].ncolor, st.k_cnt); }
    return true;
  }
  if (K > d.color_cap || nbu > 2 * SMALL_UNIT) return false;
  const unsigned long long seed = sp.seed, stepk = (unsigned long long)st.step_index;
  // pcap: the colour space of the unit's lanes (colours of an island stay
  // below twice its size); palettes here hold colours below 32 SMALL_PW
  constexpr int pw = SMALL_PW;
  const int pcap = 32 * ((2 * K + 2 + 31) / 32);
  const int cap = d.color_bits;
  unsigned* const raw = s.u.m.seen;   // the whole union as words
  for (int w = lane; w < 3 * BODY_BITS; w += 64) raw[w] = 0;
  if (lane == 0) { s.mx = 0; s.cmax = 0; }
  __syncthreads();
  // contacts: key, island, body slots; contacts per body and per body pair
  for (int i = lane; i < K; i += 64) {
    const int k = d.uperm[p0 + i];
    const int A = d.c_ia[k], B = d.c_ib[k];
    const int fa = d.b_fixed[A] ? -1 : d.b2f[A], fb = d.b_fixed[B] ? -1 : d.b2f[B];
    const int la = fa >= 0 ? d.f_loc[fa] : NO_SLOT, lb = fb >= 0 ? d.f_loc[fb] : NO_SLOT;
    s.key[i] = k - k_off;
    s.isl[i] = d.c_lab[k];
    s.la[i] = (unsigned short)la;
    s.lb[i] = (unsigned short)lb;
    if (la != NO_SLOT) atomicAdd(&s.u.s.cnt[la], 1);
    if (lb != NO_SLOT) atomicAdd(&s.u.s.cnt[lb], 1);
    if (la != NO_SLOT && lb != NO_SLOT) {
      const unsigned key = (((unsigned)min(la, lb) << 10) | (unsigned)max(la, lb)) + 1u;
      unsigned h = pair_hash(key);
      for (int t = 0; t < PAIR_HASH; ++t, h = (h + 1) & (PAIR_HASH - 1)) {
        const unsigned old = atomicCAS(&s.u.s.hk[h], 0u, key);
        if (old == 0u || old == key) { atomicAdd(&s.u.s.hc[h], 1); break; }
      }
    }
  }
  __syncthreads();
  // palettes {0, ..., deg - 1}
  for (int i = lane; i < K; i += 64) {
    const int la = s.la[i], lb = s.lb[i];
    int deg = 0;
    if (la != NO_SLOT) deg += s.u.s.cnt[la] - 1;
    if (lb != NO_SLOT) deg += s.u.s.cnt[lb] - 1;
    if (la != NO_SLOT && lb != NO_SLOT) {
      const unsigned key = (((unsigned)min(la, lb) << 10) | (unsigned)max(la, lb)) + 1u;
      unsigned h = pair_hash(key);
      for (int t = 0; t < PAIR_HASH; ++t, h = (h + 1) & (PAIR_HASH - 1))
        if (s.u.s.hk[h] == key) { deg -= s.u.s.hc[h] - 1; break; }
    }
    if (deg > 0) atomicMax(&s.mx, deg);
    unsigned* p = s.pal + i * pw;
    for (int w = 0; w < pw; ++w) {
      const int lo = w * 32;
      p[w] = deg >= lo + 32 ? 0xffffffffu : deg > lo ? (1u << (deg - lo)) - 1u : 0u;
    }
    s.pc[i] = (unsigned short)deg;
    s.col[i] = 0;
    s.done[i] = 0;
  }
  __syncthreads();
  for (int w = lane; w < 3 * BODY_BITS; w += 64) raw[w] = 0;
  int pwe = (s.mx + 31) / 32;   // palette words in use
  __syncthreads();
  if (pwe > pw || nbu * max(pwe, 1) > cap) return false;
  unsigned* const seen = s.u.m.seen;
  unsigned* const dup = s.u.m.dup;
  unsigned* const won = s.u.m.won;
  int rnd = 0;
  for (; rnd < MAX_ROUNDS; ++rnd) {
    const int bw = max(pwe, 1);
    // last round's winners are coloured; draws; the colours drawn on each body
    for (int w = lane; w < nbu * bw; w += 64) won[w] = 0;
    int left = 0;
    for (int i = lane; i < K; i += 64) {
      const int dn = s.done[i];
      if (dn == 2) s.done[i] = 1;
      if (dn) continue;
      ++left;
      int c = s.col[i];
      const int n = s.pc[i];
      if (n > 0) {
        c = lds_select(s.pal + i * pw, pwe, (int)color_draw(seed, stepk, (unsigned long long)s.key[i],
                                                             (unsigned long long)rnd, (unsigned)n));
        s.col[i] = (unsigned short)c;
      }
      if (c < 0 || c >= 32 * bw) continue;   // only after a colour-space overflow (ERR_COLORS)
      const unsigned bit = 1u << (c & 31);
      const int la = s.la[i], lb = s.lb[i];
      if (la != NO_SLOT) { const int w = la * bw + (c >> 5); if (atomicOr(&seen[w], bit) & bit) atomicOr(&dup[w], bit); }
      if (lb != NO_SLOT) { const int w = lb * bw + (c >> 5); if (atomicOr(&seen[w], bit) & bit) atomicOr(&dup[w], bit); }
    }
    if (!__any(left)) break;
    __syncthreads();
    // a draw is kept if no other uncoloured contact of either body drew it
    int win = 0;
    for (int i = lane; i < K; i += 64) {
      if (s.done[i]) continue;
      const int c = s.col[i];
      if (c >= 32 * bw) continue;
      const unsigned bit = 1u << (c & 31);
      const int la = s.la[i], lb = s.lb[i];
      const bool lose = (la != NO_SLOT && (dup[la * bw + (c >> 5)] & bit)) ||
                        (lb != NO_SLOT && (dup[lb * bw + (c >> 5)] & bit));
      if (lose) continue;
      s.done[i] = 2;
      win = 1;
      if (la != NO_SLOT) atomicOr(&won[la * bw + (c >> 5)], bit);
      if (lb != NO_SLOT) atomicOr(&won[lb * bw + (c >> 5)], bit);
    }
    __syncthreads();
    // the colours won on its bodies leave the palette of a contact still
    // uncoloured, whose size is recounted
    int moving = win, empty = 0;
    for (int i = lane; i < K; i += 64) {
      if (s.done[i]) continue;
      const int la = s.la[i], lb = s.lb[i];
      unsigned* p = s.pal + i * pw;
      int n = 0;
      for (int w = 0; w < pwe; ++w) {
        unsigned m = 0;
        if (la != NO_SLOT) m |= won[la * bw + w];
        if (lb != NO_SLOT) m |= won[lb * bw + w];
        const unsigned v = p[w] & ~m;
        p[w] = v;
        n += __popc(v);
      }
      s.pc[i] = (unsigned short)n;
      if (n == 0) empty = 1;
      if (n != 1) moving = 1;
    }
    for (int w = lane; w < nbu * bw; w += 64) { seen[w] = 0; dup[w] = 0; }
    // No winner and every uncoloured palette a single colour: every later
    // round draws the same colours and keeps none; go to the leftovers.
    if (!__any(moving)) { rnd = MAX_ROUNDS; break; }
    __syncthreads();
    if (__any(empty)) {
      // empty palettes get max+1, max+2, ... in increasing contact order per
      // island (max over the island's kept and tentative colours)
      for (int i = lane; i < K; i += 64) {
        if (s.done[i] || s.pc[i] != 0) continue;
        const int il = s.isl[i], ki = s.key[i];
        int mx = -1, rank = 0;
        for (int j = 0; j < K; ++j) {
          if (s.isl[j] != il) continue;
          mx = max(mx, (int)s.col[j]);
          if (j != i && !s.done[j] && s.pc[j] == 0 && s.key[j] < ki) ++rank;
        }
        const int c = mx + 1 + rank;
        if (c >= pcap) { atomicOr(d.err, ERR_COLORS); continue; }
        if (c >= 32 * pw) { atomicMax(&s.mx, c + 1); continue; }   // beyond these palettes: start over below
        s.pal[i * pw + (c >> 5)] |= 1u << (c & 31);
        atomicMax(&s.mx, c + 1);
      }
      __syncthreads();
      for (int i = lane; i < K; i += 64)
        if (!s.done[i] && s.pc[i] == 0) s.pc[i] = 1;
      pwe = (s.mx + 31) / 32;
      __syncthreads();
      if (pwe > pw || nbu * max(pwe, 1) > cap) return false;
    }
  }
  if (rnd >= MAX_ROUNDS) {
    // leftovers get fresh colours max+1, ... in increasing contact order per
    // island.  The wavefront walks the (few) leftover contacts together; for
    // each, every lane checks the unit's contacts it holds, and a wavefront max
    // and sum give the island's largest colour (kept or tentative) and the
    // leftover's rank.  (A walk over the whole unit by every leftover's lane
    // made the wavefront run K iterations whenever one contact was left.)
    __syncthreads();
    constexpr int QS = SMALL_UNIT / 64;
    int fresh[QS], il_q[QS], col_q[QS], key_q[QS];
    bool left_q[QS];
#pragma unroll
    for (int q = 0; q < QS; ++q) {
      const int i = lane + 64 * q;
      const bool in = i < K;
      fresh[q] = -1;
      il_q[q] = in ? s.isl[i] : -1;
      col_q[q] = in ? (int)s.col[i] : -1;
      key_q[q] = in ? s.key[i] : 0;
      left_q[q] = in && !s.done[i];
    }
#pragma unroll
    for (int q = 0; q < QS; ++q) {
      unsigned long long m = __ballot(left_q[q]);
      while (m) {                                   // (uniform over the wavefront)
        const int b = __ffsll(m) - 1;
        m &= m - 1;
        const int i = 64 * q + b;
        const int il = s.isl[i], ki = s.key[i];
        int mx = -1, rank = 0;
#pragma unroll
        for (int r = 0; r < QS; ++r) {
          if (il_q[r] != il) continue;              // (-1: no contact)
          mx = max(mx, col_q[r]);
          if (left_q[r] && key_q[r] < ki) ++rank;   // keys are distinct: i itself is not counted
        }
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
          mx = max(mx, __shfl_xor(mx, o, 64));
          rank += __shfl_xor(rank, o, 64);
        }
        if (lane == b) fresh[q] = mx + 1 + rank;
      }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < SMALL_UNIT / 64; ++q) {
      if (fresh[q] < 0) continue;
      if (fresh[q] >= pcap) atomicOr(d.err, ERR_COLORS);
      s.col[lane + 64 * q] = (unsigned short)min(fresh[q], pcap - 1);
    }
    __syncthreads();
  }
  int mx = 0;
  for (int i = lane; i < K; i += 64) {
    mx = max(mx, s.col[i] + 1);
    d.c_color[s.key[i] + k_off] = s.col[i];
  }
  atomicMax(&s.cmax, mx);
  __syncthreads();
  if (lane == 0) {
    if (s.cmax > 2 * K + 1) atomicOr(d.err, ERR_COLORS);   // beyond the lane's colour offset space
    d.u_ncol[u] = s.cmax;
    atomicMax(&d.ss[slot].ncolor, s.cmax);
  }
  return true;
}

// units of up to SMALL_UNIT contacts: one wavefront each, through per-body
// bitmaps; larger units, and those outgrowing the bitmaps, are listed in
// big_list for k_color_units
// (grids of the per-unit and per-lane kernels are capped, and a workgroup
// walks units or lanes in turn: a grid of cap_units workgroups, almost all of
// them without work, costs ~1 ns per workgroup to dispatch)
template <typename R> __global__ void __launch_bounds__(64) k_color_small(Dev<R> d) {
  __shared__ BitsColorLds s;
  if (aborted(d.err)) return;
  const int nu = *d.n_units;
  for (int u = blockIdx.x; u < nu; u += gridDim.x) {
    const int K = d.u_end[u] - d.u_start[u];
    if (K <= 0) continue;
    if ((K > SMALL_UNIT || !color_small_bits<R>(d, u, s)) && threadIdx.x == 0)
      d.big_list[1 + atomicAdd(&d.big_list[0], 1)] = u;
    __syncthreads();   // s is reused by the next unit
  }
}

// the listed units: a workgroup of COLOR_UNIT_BLOCK threads each (a grid of
// at most one workgroup per CU walks the list), neighbours through the
// bodies' contact lists
template <typename R> __global__ void __launch_bounds__(COLOR_UNIT_BLOCK) k_color_units(Dev<R> d) {
  __shared__ ColorLds<COLOR_UNIT_BLOCK, UNIT_CAP, LIST_CAP, 0> s;
  if (aborted(d.err)) return;
  const int n = d.big_list[0];
  for (int i = blockIdx.x; i < n; i += gridDim.x) {
    color_unit<R>(d, d.big_list[1 + i], s);
    __syncthreads();
  }
}

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

// ------------------------------------------------------------ sweep order
// Gather one contact into the solver arrays at position `pos`; s_loc holds
// the bodies' slots (index into u_blist; -1 for a fixed body).
// ITEMS: also the per-contact items of the global layout
template <typename R> __device__ __forceinline__ void gather_contact(const Dev<R>& d, int pos, int k, bool items) {
  const V4<R> mu = d.c_mu[k];
  const V4<R> n = d.c_n[k], rA = d.c_rA[k], rB = d.c_rB[k];
  const int A = d.c_ia[k], B = d.c_ib[k];
  const bool fa = !d.b_fixed[A], fb = !d.b_fixed[B];
  int sa = -1, sb = -1;
  if (fa) { const int f = d.b2f[A]; sa = d.u_boff[d.f_unit[f]] + d.f_loc[f]; }
  if (fb) { const int f = d.b2f[B]; sb = d.u_boff[d.f_unit[f]] + d.f_loc[f]; }
  d.perm[pos] = k;
  d.s_loc[pos] = make_int2(sa, sb);
  if (items) {
    d.s_id[pos] = make_int2(A, B);
    d.s_n[pos] = V4<R>{n.x, n.y, n.z, mu.x};
    d.s_rA[pos] = V4<R>{rA.x, rA.y, rA.z, mu.y};
    d.s_rB[pos] = V4<R>{rB.x, rB.y, rB.z, mu.z};
    d.s_b[pos] = d.c_b[k];
    d.s_r0[pos] = d.c_r0[k];
    const V4<R> z = {R(0), R(0), R(0), R(0)};
    V4<R>* W = d.s_W + 4 * (size_t)pos;
    W[0] = fa ? d.b_W0[A] : z;
    W[1] = fa ? d.b_W1[A] : z;
    W[2] = fb ? d.b_W0[B] : z;
    W[3] = fb ? d.b_W1[B] : z;
    d.x0[pos] = z;
    d.x1[pos] = z;
  }
  // the rows of J once per step, for the row-major (quad) sweeps
  ContactRow<R> J;
  contact_rows(V4<R>{n.x, n.y, n.z, R(0)}, rA, rB, J);
  const V4<R> bb = d.c_b[k], r0 = d.c_r0[k];
  const R bv[4] = {bb.x, bb.y, bb.z, bb.w}, rv[4] = {r0.x, r0.y, r0.z, r0.w}, mv[4] = {R(0), mu.x, mu.y, mu.z};
  for (int r = 0; r < 4; ++r) {
    const size_t row = 4 * (size_t)pos + r;
    d.s_JL[row] = V4<R>{J.lin[r].x, J.lin[r].y, J.lin[r].z, bv[r]};
    d.s_JA[row] = V4<R>{J.angA[r].x, J.angA[r].y, J.angA[r].z, rv[r]};
    d.s_JB[row] = V4<R>{J.angB[r].x, J.angB[r].y, J.angB[r].z, mv[r]};
    d.xr0[row] = R(0);
    d.xr1[row] = R(0);
  }
}

// steppers.py:15-36 with constraints/post_stabilization.py:21-41, 99-104: the
// post-stabilisation problem keeps the contact problem's rows (J, W J^T and
// r = 1/diag) and sweep order; its right-hand side is the gap of a contact
// deeper than min_gap_value on the normal row and 0 elsewhere, its friction
// rows are projected to zero (mu = 0), and it starts from x = 0.  Rewrites
// the sweep-order arrays of every layout in place (the contact solve is done
// with them) and flags the scenes that have something to correct.
template <typename R> __global__ void k_post_rows(Dev<R> d) {
  if (aborted(d.err)) return;
  const int nk = *d.n_contacts;
  if (nk > d.cap_contacts) return;
  for (int pos = blockIdx.x * blockDim.x + threadIdx.x; pos < nk; pos += gridDim.x * blockDim.x) {
    const int k = d.perm[pos];
    const int slot = d.b_slot[d.c_ia[k]];
    const SceneParams& sp = d.sp[slot];
    if (!sp.use_post_stab) continue;
    const R g = d.c_p[k].w;
    const bool deep = (double)g < -sp.min_gap;
    const R zero = R(0);
    d.s_b[pos] = V4<R>{deep ? g : zero, zero, zero, zero};
    V4<R> n = d.s_n[pos], ra = d.s_rA[pos], rb = d.s_rB[pos];
    n.w = zero; ra.w = zero; rb.w = zero;
    d.s_n[pos] = n; d.s_rA[pos] = ra; d.s_rB[pos] = rb;
    d.x0[pos] = V4<R>{zero, zero, zero, zero};
    d.x1[pos] = V4<R>{zero, zero, zero, zero};
    for (int r = 0; r < 4; ++r) {
      const size_t row = 4 * (size_t)pos + r;
      V4<R> jl = d.s_JL[row], jb = d.s_JB[row];
      jl.w = (r == 0 && deep) ? g : zero;
      jb.w = zero;
      d.s_JL[row] = jl;
      d.s_JB[row] = jb;
      d.xr0[row] = zero;
      d.xr1[row] = zero;
    }
    if (deep) d.ss[slot].post_active = 1;
  }
}

// state_storage.py:251-302 with dt = 1 and u = W J^T sol of the
// post-stabilisation problem (steppers.py:33-36): positions only.
template <typename R> __global__ void k_post_integrate(Dev<R> d) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= d.nb || aborted(d.err)) return;
  const int b = d.bb_body[f];
  if (d.b_fixed[b]) return;
  const int slot = d.b_slot[b];
  if (!(d.sp[slot].use_post_stab && d.ss[slot].post_active)) return;
  const R one = R(1);
  const V4<R> s0 = d.ws0[b], s1 = d.ws1[b];
  V4<R> r = d.r[b];
  r.x = r.x + s0.x * one; r.y = r.y + s0.y * one; r.z = r.z + s0.z * one;
  Quat<R> q = quat(d.q[b]);
  Quat<R> dq = qprod(Quat<R>{R(0), s1.x, s1.y, s1.z}, q);
  Quat<R> qn = {q.s + dq.s * one * R(0.5), q.x + dq.x * one * R(0.5), q.y + dq.y * one * R(0.5),
                q.z + dq.z * one * R(0.5)};
  qn = qunit(qn);
  d.r[b] = r;
  d.q[b] = V4<R>{qn.s, qn.x, qn.y, qn.z};
}

// colour offsets of lane l: 2 K_l + 2 entries (an island's colours stay below
// twice its size, which the colouring checks)
template <typename R> __device__ __forceinline__ const int* lane_coff(const Dev<R>& d, int l) {
  return d.col_off + 2 * ((size_t)d.l_start[l] + l);
}

// One workgroup per sweep lane: its positions [l_start, l_end) in sweep
// order -- by colour (counting sort; the order inside a colour changes no
// result), or for the reference's serial Gauss-Seidel by contact index
// (contacts of different islands commute, so the order inside each island
// is all that matters, and that is the contact order).
template <typename R> __device__ __forceinline__ void order_lane(const Dev<R>& d, int l, int* s_cnt, int& s_nc) {
  const int p0 = d.l_start[l], p1 = d.l_end[l];
  if (p1 <= p0) return;
  const int K = p1 - p0;
  if (d.sp[d.bs_slot[d.l_bs[l]]].scheme == SCHEME_GS) {
    for (int i = threadIdx.x; i < K; i += COLOR_BLOCK) {
      const int k = d.uperm[p0 + i];
      int rank = 0;
      for (int j = 0; j < K; ++j) rank += d.uperm[p0 + j] < k ? 1 : 0;
      gather_contact(d, p0 + rank, k, true);
    }
    if (threadIdx.x == 0) d.l_ncol[l] = 0;
    return;
  }
  if (threadIdx.x == 0) s_nc = 0;
  __syncthreads();
  int mx = 0;
  for (int i = threadIdx.x; i < K; i += COLOR_BLOCK) mx = max(mx, d.c_color[d.uperm[p0 + i]] + 1);
  atomicMax(&s_nc, mx);
  __syncthreads();
  const int nc = s_nc;
  int* coff = d.col_off + 2 * ((size_t)p0 + l);
  for (int c = threadIdx.x; c <= nc; c += COLOR_BLOCK) s_cnt[c] = 0;
  __syncthreads();
  for (int i = threadIdx.x; i < K; i += COLOR_BLOCK) atomicAdd(&s_cnt[d.c_color[d.uperm[p0 + i]]], 1);
  __syncthreads();
  if (threadIdx.x == 0) {
    int acc = 0;
    for (int c = 0; c < nc; ++c) {
      const int v = s_cnt[c];
      coff[c] = acc;
      s_cnt[c] = acc;     // from here on: the next free position of colour c
      acc += v;
    }
    coff[nc] = acc;
    d.l_ncol[l] = nc;
  }
  __syncthreads();
  // each contact's position in colour order, taken from LDS counters (the
  // order inside a colour changes no result: its contacts share no free body)
  for (int i = threadIdx.x; i < K; i += COLOR_BLOCK)
    d.l_pos[p0 + i] = atomicAdd(&s_cnt[d.c_color[d.uperm[p0 + i]]], 1);
}

template <typename R> __global__ void __launch_bounds__(COLOR_BLOCK) k_order_lanes(Dev<R> d) {
  __shared__ int s_cnt[PALMAX + 1];
  __shared__ int s_nc;
  if (aborted(d.err)) return;
  const int nl = *d.n_lanes;
  for (int l = blockIdx.x; l < nl; l += gridDim.x) {
    order_lane(d, l, s_cnt, s_nc);
    __syncthreads();   // s_cnt and s_nc are reused by the next lane
  }
}

// Contacts into colour order inside their lane, at the positions
// k_order_lanes gave them: GATHER_SPLIT workgroups per lane, each taking every
// GATHER_SPLIT-th run of COLOR_BLOCK positions.  (The positions used to be
// taken here with a global atomic per contact on the lane's per-colour
// counters, which sit in one or two cache lines: the atomics of a lane
// serialised in one L2 channel.)
constexpr int GATHER_SPLIT = 8;
template <typename R> __device__ void gather_lane(const Dev<R>& d, int l) {
  const int p0 = d.l_start[l], p1 = d.l_end[l];
  if (p1 <= p0 || d.sp[d.bs_slot[d.l_bs[l]]].scheme == SCHEME_GS) return;   // (gathered by k_order_lanes)
  // the items only where the global layout may read them: a group too large
  // for the velocity-only layout, or the post-stabilisation rows
  const int bs = d.l_bs[l];
  bool items = d.items_nb < 0 || d.sp[d.bs_slot[bs]].use_post_stab;
  if (!items) {
    const int u0 = d.u_base[bs];
    items = d.u_boff[u0 + d.u_cnt[bs]] - d.u_boff[u0] > d.items_nb;
  }
  for (int i = blockIdx.y * COLOR_BLOCK + threadIdx.x; i < p1 - p0; i += GATHER_SPLIT * COLOR_BLOCK) {
    gather_contact(d, p0 + d.l_pos[p0 + i], d.uperm[p0 + i], items);
  }
}

template <typename R> __global__ void __launch_bounds__(COLOR_BLOCK) k_lane_gather(Dev<R> d) {
  if (aborted(d.err)) return;
  const int nl = *d.n_lanes;
  for (int l = blockIdx.x; l < nl; l += gridDim.x) gather_lane(d, l);
}

// ------------------------------------------------------------ solve
// A workgroup solves a group of sweep lanes: the lanes of one batch scene, or
// a block of the lanes of a single scene spread over a cooperative grid (the
// scene-wide merit then meets in a grid barrier once per sweep).  A group's
// bodies are the slot range [b_lo, b_hi) of u_blist.  Three layouts, chosen
// per group:
//  * resident: everything a sweep touches -- contact rows, both iterates,
//    colour offsets, the bodies' velocity changes and inverse masses -- is
//    staged in LDS once per solve; each wavefront sweeps its own lanes colour
//    by colour with wavefront-level syncs, and a sweep issues no global
//    memory access;
//  * velocity-only: the bodies' velocity changes in LDS, contact rows read
//    from global memory with the next colour's rows prefetched;
//  * global: everything in global memory, lanes swept by the whole workgroup.
// Lanes larger than WAVE_LANE_MAX contacts are swept by the whole workgroup.
// The sweep order inside an island is the same in every layout and schedule,
// so all of them give bitwise the same iterates.
constexpr int SOLVE_BLOCK = 256;
constexpr int WAVE_LANE_MAX = 512;
constexpr int SOLVE_LDS_BYTES = 65536 - 512;

// bodies whose velocity changes fit the velocity-only layout
template <typename R> __host__ __device__ constexpr int lds_body_cap() {
  return SOLVE_LDS_BYTES / (4 * (int)sizeof(V4<R>));
}

// The scene parameters the friction projection reads, taken into registers
// once per solve: a reference into the scene table would be read again from
// memory after every colour phase's fence.
struct ProxParams {
  double ell_tol, ell_expansion;
  int friction, ell_max_iter;
  __device__ explicit ProxParams(const SceneParams& sp)
      : ell_tol(sp.ell_tol), ell_expansion(sp.ell_expansion), friction(sp.friction), ell_max_iter(sp.ell_max_iter) {}
};

// One row of J w (lin, angA, angB: the row's blocks).  FP64 keeps the
// oracle's sum of four dot products; FP32 sums side A and side B in two
// chains of fused multiply-adds.
template <typename R>
__device__ __forceinline__ R row_jw(V3<R> lin, V3<R> angA, V3<R> angB, V3<R> wlA, V3<R> waA, V3<R> wlB, V3<R> waB) {
  if constexpr (sizeof(R) == 8) {
    return -dot(lin, wlA) - dot(angA, waA) + dot(lin, wlB) + dot(angB, waB);
  } else {
    R a = lin.x * wlA.x;
    a = madd(lin.y, wlA.y, a); a = madd(lin.z, wlA.z, a);
    a = madd(angA.x, waA.x, a); a = madd(angA.y, waA.y, a); a = madd(angA.z, waA.z, a);
    R b = lin.x * wlB.x;
    b = madd(lin.y, wlB.y, b); b = madd(lin.z, wlB.z, b);
    b = madd(angB.x, waB.x, b); b = madd(angB.y, waB.y, b); b = madd(angB.z, waB.z, b);
    return b - a;
  }
}

// z = x - r (J w + b) for one row: xo - r (Jw + b)
template <typename R> __device__ __forceinline__ R row_z(R xo, R r, R jw, R b) { return madd(-r, jw + b, xo); }

// contacts.py:77-95: z = x - r (J w + b); x_n = max(0, z_n); friction prox
template <typename R>
__device__ __forceinline__ V4<R> contact_prox(const ProxParams& pp, const ContactRow<R>& J, V4<R> xo, V4<R> bb,
                                              V4<R> rr, V4<R> mu, V3<R> wlA, V3<R> waA, V3<R> wlB, V3<R> waB,
                                              bool use_old_xn) {
  R Jw[4];
  for (int i = 0; i < 4; ++i) Jw[i] = row_jw(J.lin[i], J.angA[i], J.angB[i], wlA, waA, wlB, waB);
  R z0 = row_z(xo.x, rr.x, Jw[0], bb.x);
  R z1 = row_z(xo.y, rr.y, Jw[1], bb.y);
  R z2 = row_z(xo.z, rr.z, Jw[2], bb.z);
  R z3 = row_z(xo.w, rr.w, Jw[3], bb.w);
  R xn = fmax(R(0), z0);
  R xr = use_old_xn ? xo.x : xn;
  R os, ot, orr;
  if (pp.friction == 1) {
    const V3<R> o = prox_ellipsoid_call<R>(z1, z2, z3, mu.x, mu.y, mu.z, xr, pp.ell_tol, pp.ell_expansion, pp.ell_max_iter);
    os = o.x; ot = o.y; orr = o.z;
  } else {
    prox_sphere<R>(z1, z2, z3, mu.x, xr, os, ot, orr);
  }
  return V4<R>{xn, os, ot, orr};
}

// w += W J^T dx for one body (contacts.py:107); W0 = (I^-1 xx, yy, zz, xy),
// W1 = (I^-1 xz, yz, 1/m, 0); l, g: linear and angular parts of w
template <typename R>
__device__ __forceinline__ void dw_body(V4<R> W0, V4<R> W1, R sign, const ContactRow<R>& J, const V3<R>* ang,
                                        V4<R> dx, V4<R>& l, V4<R>& g) {
  const R im = W1.z;
  const R W[6] = {W0.x, W0.y, W0.z, W0.w, W1.x, W1.y};
  const R dd[4] = {dx.x, dx.y, dx.z, dx.w};
  const R s0 = sign * dd[0];
  R dl[3] = {sum_first(J.lin[0].x * s0), sum_first(J.lin[0].y * s0), sum_first(J.lin[0].z * s0)};
  V3<R> a = {sum_first(s0 * ang[0].x), sum_first(s0 * ang[0].y), sum_first(s0 * ang[0].z)};
  for (int i = 1; i < 4; ++i) {
    const R s = sign * dd[i];
    dl[0] = madd(J.lin[i].x, s, dl[0]); dl[1] = madd(J.lin[i].y, s, dl[1]); dl[2] = madd(J.lin[i].z, s, dl[2]);
    a = V3<R>{madd(s, ang[i].x, a.x), madd(s, ang[i].y, a.y), madd(s, ang[i].z, a.z)};
  }
  const V3<R> da = sym_madd(W, a);
  l = V4<R>{madd(im, dl[0], l.x), madd(im, dl[1], l.y), madd(im, dl[2], l.z), R(0)};
  g = V4<R>{g.x + da.x, g.y + da.y, g.z + da.z, R(0)};
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

// gauss_seidel_solver.py:56-97: accept/reject, r-factor, stopping tests.
struct SolveCtl {
  double last_merit = INFINITY, glast = INFINITY, rscale = 1.0, gm = 0.0;
  int acc = 0, rejects = 0, exitcode = 0, iters = 0, sweeps = 0;
  // accept (bit0) or reject, r-factor update; gm = this problem's contribution
  // to the global merit (its merit, or its last accepted merit on a reject)
  __device__ int accept_or_reject(const SceneParams& sp, double merit) {
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
    return dec;
  }
  // returns bit0 accept, bit1 stop; `other` is the contribution of the
  // scene's other problem (the hinges) to the global merit of this sweep
  __device__ int decide(const SceneParams& sp, double merit, double* res, int it, double other = 0.0) {
    int dec = accept_or_reject(sp, merit);
    gm = fmax(gm, other);
    if (res) res[it] = gm;
    if (gm < sp.abs_tol) { dec |= 2; exitcode = 1; }
    else if (fabs(gm - glast) < sp.rel_tol * glast) { dec |= 2; exitcode = 2; }
    glast = gm;
    ++sweeps;
    if (dec & 1) acc ^= 1;
    return dec;
  }
};

// the hinge problem's contribution to sweep `it` (recorded before the contact
// solve; gauss_seidel_solver.py:47-73 takes the largest over the problems)
template <typename R> __device__ __forceinline__ double hinge_merit(const Dev<R>& d, const SceneParams& sp, int it) {
  return (sp.h_cnt > 0 && !d.post) ? d.h_mer[sp.res_off + it] : 0.0;
}

// the scene's per-sweep merit record: the contact solve's 'residuals', then
// the post-stabilisation solve's 'post_stabilization_residuals'
// (gauss_seidel_solver.py:29-33, 76 with prefix; max_iter entries each)
template <typename R> __device__ __forceinline__ double* residual_record(const Dev<R>& d, const SceneParams& sp) {
  return d.residuals + sp.res_off + (d.post ? max(1, sp.max_iter) : 0);
}

template <typename R> __device__ void write_solve_state(const Dev<R>& d, SceneState& st, const SolveCtl& c) {
  if (d.post) {   // the post-stabilisation solve keeps the contact solve's record
    st.post_iters = c.iters;
    st.post_exit = c.exitcode;
    return;
  }
  st.solved = 1;
  st.iterations = c.iters;
  st.exitcode = c.exitcode;
  st.rejects = c.rejects;
  st.sweeps = c.sweeps;
  st.final_merit = c.last_merit;
}

// The scene-wide merit of sweep `it` in a cooperative launch, which is also
// the sweep's grid barrier.  Workgroup b publishes its merit in slot
// it * G + b: the bits of the non-negative double, or 0 for a merit that is
// not > 0 (a maximum that starts from 0 ignores it).  The slots start as all
// ones, which no published value is.  Wavefront 0 then waits until the
// sweep's G slots are all published and takes their maximum.  Workgroups
// share no other data during the solve, so the slots need no fences.  The
// wait is bounded: it flags ERR_SOLVER and returns false instead of hanging.
__device__ __forceinline__ bool grid_merit(unsigned long long* slots, int it, double& merit, int* err, int* s_ok,
                                           double* s_m) {
  const unsigned G = gridDim.x;
  unsigned long long* row = slots + (size_t)it * G;
  if (threadIdx.x == 0)
    __hip_atomic_store(&row[blockIdx.x], merit > 0.0 ? (unsigned long long)__double_as_longlong(merit) : 0ull,
                       __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  if (threadIdx.x < 64) {
    unsigned long long m = 0;
    bool ok = true;
    for (unsigned j = threadIdx.x; j < G && ok; j += 64) {
      unsigned long long v = __hip_atomic_load(&row[j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      for (unsigned spins = 0; v == ~0ull;) {
        __builtin_amdgcn_s_sleep(1);
        if (++spins > (1u << 26)) { ok = false; break; }
        v = __hip_atomic_load(&row[j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
      if (ok) m = max(m, v);
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) m = max(m, (unsigned long long)__shfl_xor(m, o, 64));
    const bool all_ok = __ballot(!ok) == 0;
    if (threadIdx.x == 0) {
      if (!all_ok) atomicOr(err, ERR_SOLVER);
      *s_ok = all_ok ? 1 : 0;
      *s_m = __longlong_as_double((long long)m);
    }
  }
  __syncthreads();
  merit = *s_m;
  return *s_ok != 0;
}

// Working (wl, wa) and accepted (sl, sa) velocity changes of a group's
// bodies: LDS (slot - base) or the arena arrays (by body).
template <typename R, bool LDS> struct VelStore {
  V4<R>* wl; V4<R>* wa; V4<R>* sl; V4<R>* sa;
  int base;
  __device__ __forceinline__ int at(int slot, int body) const { return LDS ? slot - base : body; }
};

// the constant rows of one contact and its previous iterate
template <typename R> struct Item {
  int2 loc, id;
  V4<R> n, rA, rB, b, r0, xo, W[4];
};

template <typename R> __device__ __forceinline__ void load_item(const Dev<R>& d, const V4<R>* xa, int kk, Item<R>& it) {
  it.loc = d.s_loc[kk];
  it.id = d.s_id[kk];
  it.n = d.s_n[kk];
  it.rA = d.s_rA[kk];
  it.rB = d.s_rB[kk];
  it.b = d.s_b[kk];
  it.r0 = d.s_r0[kk];
  it.xo = xa[kk];
  const V4<R>* W = d.s_W + 4 * (size_t)kk;
  it.W[0] = W[0]; it.W[1] = W[1]; it.W[2] = W[2]; it.W[3] = W[3];
}

// One contact (contacts.py:77-107).  Gauss-Seidel reads and updates the
// working velocity changes; Jacobi (jacobi_solver.py:6-57) reads the
// accepted ones and adds its change to the working ones.  Returns
// max |x_new - x_old| in FP64.
template <typename R, bool LDS>
__device__ __forceinline__ double update_contact(const ProxParams& pp, const Item<R>& it, VelStore<R, LDS>& w,
                                                 V4<R>* xn, int kk, R rs, bool jacobi, bool hybrid) {
  ContactRow<R> J;
  contact_rows(it.n, it.rA, it.rB, J);
  const V4<R> rr = {it.r0.x * rs, it.r0.y * rs, it.r0.z * rs, it.r0.w * rs};
  const V4<R> mu = {it.n.w, it.rA.w, it.rB.w, R(0)};
  const bool fa = it.loc.x >= 0, fb = it.loc.y >= 0;
  const int sa = w.at(it.loc.x, it.id.x), sb = w.at(it.loc.y, it.id.y);
  const V4<R> z = {R(0), R(0), R(0), R(0)};
  const V4<R>* rl = jacobi ? w.sl : w.wl;
  const V4<R>* ra = jacobi ? w.sa : w.wa;
  const V4<R> lA = fa ? rl[sa] : z, aA = fa ? ra[sa] : z;
  const V4<R> lB = fb ? rl[sb] : z, aB = fb ? ra[sb] : z;
  const V4<R> x = contact_prox(pp, J, it.xo, it.b, rr, mu, xyz(lA), xyz(aA), xyz(lB), xyz(aB), jacobi && !hybrid);
  xn[kk] = x;
  const V4<R> dx = {x.x - it.xo.x, x.y - it.xo.y, x.z - it.xo.z, x.w - it.xo.w};
  if (fa) {
    V4<R> l = jacobi ? w.wl[sa] : lA, g = jacobi ? w.wa[sa] : aA;
    dw_body(it.W[0], it.W[1], R(-1), J, J.angA, dx, l, g);
    w.wl[sa] = l;
    w.wa[sa] = g;
  }
  if (fb) {
    V4<R> l = jacobi ? w.wl[sb] : lB, g = jacobi ? w.wa[sb] : aB;
    dw_body(it.W[2], it.W[3], R(1), J, J.angB, dx, l, g);
    w.wl[sb] = l;
    w.wa[sb] = g;
  }
  return absmax4((double)x.x - (double)it.xo.x, (double)x.y - (double)it.xo.y, (double)x.z - (double)it.xo.z,
                 (double)x.w - (double)it.xo.w);
}

template <int STRIDE> __device__ __forceinline__ void phase_sync() {
  if (STRIDE == 64) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  } else {
    __syncthreads();
  }
}

// One sweep of one lane at positions p0 + [0, K) sorted by colour: STRIDE
// lanes (a wavefront or the workgroup) take the contacts of a colour, in
// passes of STRIDE; colours are separated by a sync.  coff: colour offsets
// (serial: colour c is position c).  Returns this thread's max |dx|.
template <typename R, bool LDS, int STRIDE>
__device__ double sweep_lane(const Dev<R>& d, const ProxParams& pp, VelStore<R, LDS>& w, int p0, int nc,
                             const int* coff, bool serial, bool jacobi, bool hybrid, const V4<R>* xa, V4<R>* xn,
                             R rs, int lane) {
  double m = 0.0;
  if (nc <= 0) return m;
  // pass state: colour c, first position pb of the pass, end hi of colour c, end hn of colour c + 1
  int c = 0;
  int pb = serial ? 0 : coff[0];
  int hi = serial ? 1 : coff[1];
  int hn = nc > 1 ? (serial ? 2 : coff[2]) : hi;
  // Two item buffers used in turn (no register copies, so the prefetch of
  // one pass stays in flight while the other is computed).
  Item<R> ia, ib;
  bool ha = pb + lane < hi;
  if (ha) load_item(d, xa, p0 + pb + lane, ia);
  auto step = [&](Item<R>& use, bool have, Item<R>& fill, bool& fhave) {
    const bool same = pb + STRIDE < hi;
    int c1, pb1, hi1, hn1;
    if (same) {
      c1 = c; pb1 = pb + STRIDE; hi1 = hi; hn1 = hn;
    } else {
      c1 = c + 1; pb1 = hi; hi1 = hn;
      hn1 = c1 + 1 < nc ? (serial ? c1 + 2 : coff[c1 + 2]) : hi1;
    }
    fhave = c1 < nc && pb1 + lane < hi1;
    if (fhave) load_item(d, xa, p0 + pb1 + lane, fill);
    if (have) m = fmax(m, update_contact(pp, use, w, xn, p0 + pb + lane, rs, jacobi, hybrid));
    if (!same) phase_sync<STRIDE>();
    c = c1; pb = pb1; hi = hi1; hn = hn1;
  };
  bool hb = false;
  while (true) {
    step(ia, ha, ib, hb);
    if (c >= nc) break;
    step(ib, hb, ia, ha);
    if (c >= nc) break;
  }
  return m;
}

// The lanes and bodies of one solve group.
struct Group {
  int l_lo, l_hi;     // lanes
  int b_lo, b_hi;     // body slots
  bool lead;          // writes the scene's solver state
};

template <typename R, bool LDS, bool COOP, int NT>
__device__ void solve_group(const Dev<R>& d, const int slot, const Group& gr, V4<R>* smem, double* sh_red,
                            int* s_ok) {
  const SceneParams& sp = d.sp[slot];
  SceneState& st = d.ss[slot];
  const int b0 = gr.b_lo, nbody = gr.b_hi - gr.b_lo;
  constexpr int CAP = lds_body_cap<R>();
  VelStore<R, LDS> w;
  if (LDS) { w.wl = smem; w.wa = smem + CAP; w.sl = smem + 2 * CAP; w.sa = smem + 3 * CAP; }
  else { w.wl = d.ww0; w.wa = d.ww1; w.sl = d.ws0; w.sa = d.ws1; }
  w.base = b0;
  if (LDS) {
    for (int i = threadIdx.x; i < nbody; i += blockDim.x) {
      const int b = d.u_blist[b0 + i];
      w.wl[i] = d.ww0[b]; w.wa[i] = d.ww1[b]; w.sl[i] = d.ws0[b]; w.sa[i] = d.ws1[b];
    }
    __syncthreads();
  }
  const bool serial = sp.scheme == SCHEME_GS;
  const bool jacobi = sp.scheme == SCHEME_JACOBI || sp.scheme == SCHEME_JACOBI_HYBRID;
  const bool hybrid = sp.scheme == SCHEME_JACOBI_HYBRID;
  const ProxParams pp(sp);
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, nwave = blockDim.x >> 6;
  double* res = residual_record(d, sp);
  SolveCtl ctl;
  ctl.iters = sp.max_iter;
  for (int it = 0; it < sp.max_iter; ++it) {
    const V4<R>* xa = ctl.acc ? d.x1 : d.x0;
    V4<R>* xn = ctl.acc ? d.x0 : d.x1;
    const R rs = R(ctl.rscale);
    double m = 0.0;
    if (LDS) {
      for (int l = gr.l_lo + wave; l < gr.l_hi; l += nwave) {
        const int p0 = d.l_start[l], K = d.l_end[l] - p0;
        if (K <= 0 || (K > WAVE_LANE_MAX && !serial)) continue;
        m = fmax(m, sweep_lane<R, LDS, 64>(d, pp, w, p0, serial ? K : d.l_ncol[l], lane_coff(d, l), serial, jacobi,
                                           hybrid, xa, xn, rs, lane));
      }
    }
    for (int l = gr.l_lo; l < gr.l_hi; ++l) {
      const int p0 = d.l_start[l], K = d.l_end[l] - p0;
      if (K <= 0 || (LDS && (K <= WAVE_LANE_MAX || serial))) continue;
      m = fmax(m, sweep_lane<R, LDS, NT>(d, pp, w, p0, serial ? K : d.l_ncol[l], lane_coff(d, l), serial,
                                                  jacobi, hybrid, xa, xn, rs, (int)threadIdx.x));
    }
    double merit = block_max(m, sh_red);
    if (COOP && !grid_merit(d.merit, it, merit, d.err, s_ok, sh_red)) return;
    const int dec = ctl.decide(sp, merit, (gr.lead && threadIdx.x == 0) ? res : nullptr, it,
                               hinge_merit(d, sp, it));
    // accepted sweep: ws <- ww, else ww <- ws
    for (int i = threadIdx.x; i < nbody; i += blockDim.x) {
      const int s = LDS ? i : d.u_blist[b0 + i];
      if (dec & 1) { w.sl[s] = w.wl[s]; w.sa[s] = w.wa[s]; }
      else { w.wl[s] = w.sl[s]; w.wa[s] = w.sa[s]; }
    }
    __syncthreads();
    if (dec & 2) { ctl.iters = it; break; }
  }
  // impulses back to contact order; natural-map residual (FP64) with r0
  const V4<R>* xs = ctl.acc ? d.x1 : d.x0;
  double rloc = 0.0, xmax = 0.0;
  const V4<R> z = {R(0), R(0), R(0), R(0)};
  for (int l = gr.l_lo; l < gr.l_hi; ++l) {
    const int p0 = d.l_start[l], K = d.l_end[l] - p0;
    if (K <= 0) continue;
    for (int pos = threadIdx.x; pos < K; pos += blockDim.x) {
      const int kk = p0 + pos;
      Item<R> c;
      load_item(d, xs, kk, c);
      const V4<R> x = c.xo;
      if (!d.post) d.c_imp[d.perm[kk]] = x;
      ContactRow<R> J;
      contact_rows(c.n, c.rA, c.rB, J);
      const V4<R> mu = {c.n.w, c.rA.w, c.rB.w, R(0)};
      const bool fa = c.loc.x >= 0, fb = c.loc.y >= 0;
      const int sa = w.at(c.loc.x, c.id.x), sb = w.at(c.loc.y, c.id.y);
      const V4<R> lA = fa ? w.sl[sa] : z, aA = fa ? w.sa[sa] : z;
      const V4<R> lB = fb ? w.sl[sb] : z, aB = fb ? w.sa[sb] : z;
      const V4<R> y = contact_prox(pp, J, x, c.b, c.r0, mu, xyz(lA), xyz(aA), xyz(lB), xyz(aB), false);
      rloc = fmax(rloc, absmax4((double)x.x - (double)y.x, (double)x.y - (double)y.y, (double)x.z - (double)y.z,
                                (double)x.w - (double)y.w));
      xmax = fmax(xmax, absmax4(x.x, x.y, x.z, x.w));
    }
  }
  if (LDS)
    for (int i = threadIdx.x; i < nbody; i += blockDim.x) {
      const int b = d.u_blist[b0 + i];
      d.ws0[b] = w.sl[i];
      d.ws1[b] = w.sa[i];
    }
  const double rm = block_max(rloc, sh_red), xm = block_max(xmax, sh_red);
  if (threadIdx.x == 0) {
    if (d.post) {
      // the natural-map residual stays the contact solve's
    } else if (COOP) {
      if (rm > 0.0) atomic_max_nonneg(&st.nat_r, rm);
      if (xm > 0.0) atomic_max_nonneg(&st.nat_x, xm);
    } else {
      st.nat_r = rm;
      st.nat_x = xm;
    }
    if (gr.lead) write_solve_state(d, st, ctl);
  }
}

// value of lane K of each quad (DPP quad_perm(K, K, K, K))
template <int K> __device__ __forceinline__ float quad_bcast(float v) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), K * 0x55, 0xf, 0xf, false));
}
template <int K> __device__ __forceinline__ double quad_bcast(double v) {
  const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), K * 0x55, 0xf, 0xf, false);
  const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), K * 0x55, 0xf, 0xf, false);
  return __hiloint2double(hi, lo);
}

// One lane's share of a contact for the quad update: row q of J (lin, angA,
// angB with b, r0, mu in w), its iterate and the contact's two body slots.
template <typename R> struct QuadRow {
  V4<R> jl, ja, jb;
  R xo;
  int2 lc;
};

// One field of the bodies' LDS records: entry i of a record array with
// BODY_REC vec4s per body.
constexpr int BODY_REC = 6;
enum BodyField { F_WL = 0, F_WA, F_SL, F_SA, F_W0, F_W1 };
template <typename R> struct BodyCol {
  V4<R>* p;
  __device__ __forceinline__ V4<R>& operator[](int i) const { return p[BODY_REC * i]; }
};
// bytes of one body record
template <typename R> __host__ __device__ constexpr int rec_bytes() { return BODY_REC * (int)sizeof(V4<R>); }

// Velocity changes and inverse masses of a group's bodies in LDS, one record
// per slot: working (wl, wa) and accepted (sl, sa) velocity changes, W0, W1.
// A contact's loads of one body then share one address.  Slot zs stays zero
// and stands for every fixed body; slot ts = zs + 1 takes the discarded
// velocity updates of fixed bodies, so no update branches on them.
template <typename R> struct BodyLds {
  BodyCol<R> wl, wa, sl, sa, W0, W1;
  unsigned char* base;
  __device__ void at(V4<R>* b) {
    base = reinterpret_cast<unsigned char*>(b);
    wl.p = b; wa.p = b + 1; sl.p = b + 2; sa.p = b + 3; W0.p = b + 4; W1.p = b + 5;
  }
  // field f of the record at byte offset off (a slot times rec_bytes)
  __device__ __forceinline__ V4<R>& rec(int off, int f) const {
    return *reinterpret_cast<V4<R>*>(base + off + f * (int)sizeof(V4<R>));
  }
};

// One contact by one quad; lane q holds row q.  Every lane repeats the serial
// update's operations (update_contact) on the same operands in the same
// order, so the results are bitwise the same.  it.lc holds the byte offsets
// of the two bodies' records (zsb: the zero record of fixed bodies, tsb: the
// discard record).  Writes this lane's velocity block (to the discard record
// for a fixed body); returns the new iterate in xq and |xq - xo| in FP64.
// JAC: Jacobi reads the accepted velocity changes (jacobi_solver.py:6-57),
// Gauss-Seidel the working ones.
// JL/JA/JB + 4 i: the contact's rows (LDS or global); this lane's velocity
// block is built from rows 0..3 of one of them, loaded first.
template <typename R, bool JAC, typename F>
__device__ __forceinline__ double quad_core(const ProxParams& pp, const QuadRow<R>& it, int q, R rs, bool hybrid,
                                            const BodyLds<R>& B, int zsb, int tsb, const V4<R>* JL,
                                            const V4<R>* JA, const V4<R>* JB, size_t i4, R& xq, F&& after_loads) {
  constexpr int FL = JAC ? F_SL : F_WL, FA = JAC ? F_SA : F_WA;
  const bool side_b = q >= 2, ang = q & 1;   // lane 0: linear A, 1: angular A, 2: linear B, 3: angular B
  const int offA = it.lc.x, offB = it.lc.y;
  const int off = side_b ? offB : offA;
  const bool own = off != zsb;
  // fixed bodies read the zero record (the serial update uses zeros for them)
  const V4<R> lA = B.rec(offA, FL), aA = B.rec(offA, FA), lB = B.rec(offB, FL), aB = B.rec(offB, FA);
  const V4<R> W0 = B.rec(off, F_W0), W1 = B.rec(off, F_W1);
  // the working block this lane updates, read beside the four blocks above (in
  // Gauss-Seidel it is one of them: no other quad of the phase writes it)
  const int fo = ang ? (int)sizeof(V4<R>) : 0;
  const V4<R> old = B.rec(off + fo, F_WL);
  // the caller's loads (the next phase's contact), issued behind this phase's
  after_loads();
  const V4<R>* srcv = (q & 1) ? (q >= 2 ? JB : JA) : JL;
  const V4<R> v0 = srcv[i4], v1 = srcv[i4 + 1], v2 = srcv[i4 + 2], v3 = srcv[i4 + 3];
  const V4<R> jl = it.jl, ja = it.ja, jb = it.jb;
  const R xo = it.xo;
  // row q of z = x - r (J w + b)  (contacts.py:77-92)
  const R Jw = row_jw(xyz(jl), xyz(ja), xyz(jb), xyz(lA), xyz(aA), xyz(lB), xyz(aB));
  const R z = row_z(xo, ja.w * rs, Jw, jl.w);
  const R z0 = quad_bcast<0>(z), z1 = quad_bcast<1>(z), z2 = quad_bcast<2>(z), z3 = quad_bcast<3>(z);
  const R xo0 = quad_bcast<0>(xo);
  const R mu0 = quad_bcast<1>(jb.w), mu1 = quad_bcast<2>(jb.w), mu2 = quad_bcast<3>(jb.w);
  // projections (contacts.py:92-95), the same in every lane of the quad
  const R xnrm = fmax(R(0), z0);
  const R xr = (JAC && !hybrid) ? xo0 : xnrm;
  R os, ot, orr;
  if (pp.friction == 1) {
    const V3<R> o = prox_ellipsoid_call<R>(z1, z2, z3, mu0, mu1, mu2, xr, pp.ell_tol, pp.ell_expansion, pp.ell_max_iter);
    os = o.x; ot = o.y; orr = o.z;
  } else {
    prox_sphere<R>(z1, z2, z3, mu0, xr, os, ot, orr);
  }
  // row q's new value, by two selects on the bits of q
  const R x01 = ang ? os : xnrm, x23 = ang ? orr : ot;
  xq = side_b ? x23 : x01;
  const R dq = xq - xo;
  const R d0 = quad_bcast<0>(dq), d1 = quad_bcast<1>(dq), d2 = quad_bcast<2>(dq), d3 = quad_bcast<3>(dq);
  // w += W J^T dx (contacts.py:107), this lane's block; a fixed body's
  // (zero W, zero block) lands in the discard record
  {
    const R sign = side_b ? R(1) : R(-1);
    const R s0 = sign * d0, s1 = sign * d1, s2 = sign * d2, s3 = sign * d3;
    V3<R> acc = {sum_first(v0.x * s0), sum_first(v0.y * s0), sum_first(v0.z * s0)};
    acc.x = madd(v1.x, s1, acc.x); acc.y = madd(v1.y, s1, acc.y); acc.z = madd(v1.z, s1, acc.z);
    acc.x = madd(v2.x, s2, acc.x); acc.y = madd(v2.y, s2, acc.y); acc.z = madd(v2.z, s2, acc.z);
    acc.x = madd(v3.x, s3, acc.x); acc.y = madd(v3.y, s3, acc.y); acc.z = madd(v3.z, s3, acc.z);
    V4<R> nw;
    if (ang) {
      const R W[6] = {W0.x, W0.y, W0.z, W0.w, W1.x, W1.y};
      const V3<R> t = sym_madd(W, acc);
      nw = V4<R>{old.x + t.x, old.y + t.y, old.z + t.z, R(0)};
    } else {
      nw = V4<R>{madd(W1.z, acc.x, old.x), madd(W1.z, acc.y, old.y), madd(W1.z, acc.z, old.z), R(0)};
    }
    B.rec((own ? off : tsb) + fo, F_WL) = nw;
  }
  return fabs((double)xq - (double)xo);
}

// lane q's share of contact i from row arrays (LDS or global); body slots
// made local and turned into record byte offsets
template <typename R>
__device__ __forceinline__ void load_quad(const V4<R>* JL, const V4<R>* JA, const V4<R>* JB, const R* xa,
                                          const int2* loc, int i, int q, int b0, int zs, QuadRow<R>& it) {
  const size_t row = 4 * (size_t)i + q;
  it.jl = JL[row]; it.ja = JA[row]; it.jb = JB[row];
  it.xo = xa[row];
  const int2 lc = loc[i];
  it.lc = make_int2((lc.x >= 0 ? lc.x - b0 : zs) * rec_bytes<R>(), (lc.y >= 0 ? lc.y - b0 : zs) * rec_bytes<R>());
}

// natural-map residual parts of contact i (FP64) from its rows and the
// accepted velocity changes; the impulses' final values
template <typename R>
__device__ __forceinline__ void quad_finish_contact(const ProxParams& pp, const V4<R>* JL, const V4<R>* JA,
                                                    const V4<R>* JB, const R* xs, int i, const V4<R>& lA,
                                                    const V4<R>& aA, const V4<R>& lB, const V4<R>& aB, V4<R>& x,
                                                    double& rloc, double& xmax) {
  x = V4<R>{xs[4 * (size_t)i], xs[4 * (size_t)i + 1], xs[4 * (size_t)i + 2], xs[4 * (size_t)i + 3]};
  ContactRow<R> J;
  V4<R> bb, r0;
  R* bp = &bb.x;
  R* rp = &r0.x;
  for (int r = 0; r < 4; ++r) {
    const V4<R> a = JL[4 * (size_t)i + r], e = JA[4 * (size_t)i + r], f = JB[4 * (size_t)i + r];
    J.lin[r] = xyz(a); J.angA[r] = xyz(e); J.angB[r] = xyz(f);
    bp[r] = a.w; rp[r] = e.w;
  }
  const V4<R> mu = {JB[4 * (size_t)i + 1].w, JB[4 * (size_t)i + 2].w, JB[4 * (size_t)i + 3].w, R(0)};
  const V4<R> y = contact_prox(pp, J, x, bb, r0, mu, xyz(lA), xyz(aA), xyz(lB), xyz(aB), false);
  rloc = fmax(rloc, absmax4((double)x.x - (double)y.x, (double)x.y - (double)y.y, (double)x.z - (double)y.z,
                            (double)x.w - (double)y.w));
  xmax = fmax(xmax, absmax4(x.x, x.y, x.z, x.w));
}

// ---- resident layout
// Per contact i the four rows r of J (row 4i + r: JL = (lin_r, b_r),
// JA = (angA_r, r0_r), JB = (angB_r, mu_r), mu_1..3 the friction
// coefficients) and the iterates X0/X1[4i + r] are staged in LDS once per
// solve; so are the group's bodies.  A sweep issues no global memory access.
// Body arrays hold cb + 2 entries: slot cb stays zero and stands for every
// fixed body, slot cb + 1 takes the discarded updates of fixed bodies, so a
// contact's loads and stores never branch on fixed bodies.  Contact entry cc
// is the null contact (zero rows and iterates, both bodies in the zero slot):
// a quad with no contact in a colour updates it, which changes nothing.
template <typename R> __host__ __device__ constexpr size_t res_lds_bytes(int cc, int cb, int co, int cl) {
  return (size_t)(cc + 1) * (12 * sizeof(V4<R>) + 8 * sizeof(R) + sizeof(int2)) +
         (size_t)(cb + 2) * 6 * sizeof(V4<R>) + (size_t)co * 4 + (size_t)cl * 16;
}

constexpr int RES_LANES = 64;   // lanes of a resident group

// LDS capacities of the resident layout (compile-time, so array offsets fold
// into the LDS instructions); Dev::res_cap_c/b may set lower fit limits.
// H = 1: the half layout (FP32 ~68 KB), so that two solve workgroups share a
// CU when a batch has more scenes than the device has CUs.
template <typename R, int H = 0> struct ResCap {
  static constexpr int C = (sizeof(R) == 4 ? 384 : 192) >> H;   // contacts
  static constexpr int B = C;                             // bodies
  static constexpr int O = 4 * C;                         // colour offsets (nc + 3 per lane)
  static constexpr size_t BYTES = res_lds_bytes<R>(C, B, O, RES_LANES);
  // the streaming layout keeps only bodies (and the zero and discard slots) in the same LDS
  static constexpr int VB = (int)(BYTES / (6 * sizeof(V4<R>))) - 2;
};

template <typename R, int H = 0> struct ResLds {
  V4<R> *JL, *JA, *JB;                     // per contact row (4 per contact)
  R *x0, *x1;                              // per contact row
  BodyLds<R> B;                            // per body (slot - b_lo), zero slot zs, discard slot ts
  int2* loc;                               // a contact's two bodies (slot - b_lo; zs: fixed)
  int4* lt;                                // per lane: first contact, K, colours, first colour offset
  int* coff;
  static constexpr int zs = ResCap<R, H>::B;
  static constexpr int ts = ResCap<R, H>::B + 1;
  static constexpr int zsb = zs * rec_bytes<R>(), tsb = ts * rec_bytes<R>();   // their records' byte offsets
  static constexpr int nul = ResCap<R, H>::C;
  __device__ explicit ResLds(unsigned char* p) {
    constexpr int cc = ResCap<R, H>::C + 1, nbs = ResCap<R, H>::B + 2;
    V4<R>* v = reinterpret_cast<V4<R>*>(p);
    JL = v; JA = JL + 4 * cc; JB = JA + 4 * cc;
    B.at(JB + 4 * cc);
    x0 = reinterpret_cast<R*>(JB + 4 * cc + BODY_REC * nbs); x1 = x0 + 4 * cc;
    loc = reinterpret_cast<int2*>(x1 + 4 * cc);
    lt = reinterpret_cast<int4*>(loc + cc);
    coff = reinterpret_cast<int*>(lt + RES_LANES);
  }
};

// ---- streaming layout: bodies in LDS, contact rows and iterates in global memory
template <typename R, int H = 0> struct StreamLds {
  BodyLds<R> B;
  static constexpr int zs = ResCap<R, H>::VB;
  static constexpr int ts = ResCap<R, H>::VB + 1;
  __device__ explicit StreamLds(unsigned char* p) {
    constexpr int nbs = ResCap<R, H>::VB + 2;
    V4<R>* v = reinterpret_cast<V4<R>*>(p);
    (void)nbs;
    B.at(v);
  }
};

// sizes of a group's resident data (uniform over the workgroup)
struct GroupSize { int p0, nc, no, nl; };

template <typename R> __device__ GroupSize group_size(const Dev<R>& d, const Group& gr, bool serial) {
  GroupSize g{0x7fffffff, 0, 0, 0};
  int p1 = 0;
  for (int l = gr.l_lo; l < gr.l_hi; ++l) {
    const int a = d.l_start[l], e = d.l_end[l];
    if (e <= a) continue;
    g.p0 = min(g.p0, a);
    p1 = max(p1, e);
    g.no += serial ? 0 : d.l_ncol[l] + 3;
    ++g.nl;
  }
  g.nc = g.nl ? p1 - g.p0 : 0;
  if (!g.nl) g.p0 = 0;
  return g;
}

template <typename R, int H>
__device__ __forceinline__ bool group_fits(const Dev<R>& d, const Group& gr, const GroupSize& g) {
  return g.nc <= min(d.res_cap_c, ResCap<R, H>::C) && gr.b_hi - gr.b_lo <= min(d.res_cap_b, ResCap<R, H>::B) &&
         g.no <= ResCap<R, H>::O && g.nl <= RES_LANES;
}

// One sweep of one resident lane by STRIDE threads (a wavefront or the
// workgroup): STRIDE / 4 contacts per pass.  Colour c spans positions
// [co[c], co[c + 1]) of the lane (SERIAL: position c alone); the lane's
// offsets end with two repeats of co[nc], so the colours after the last are
// empty.  The merit, max |x_new - x_old| in FP64, is taken after the sweep
// over the lane's rows (a sweep updates each of its contacts once).
template <typename R, int STRIDE, bool SERIAL, bool JAC, int H>
__device__ double res_sweep(const ProxParams& pp, ResLds<R, H>& L, int4 lt, bool hybrid, const R* xa, R* xn, R rs,
                            int tid) {
  // the lane's table entry is the same for every thread: scalar loop control
  const int c0 = __builtin_amdgcn_readfirstlane(lt.x), K = __builtin_amdgcn_readfirstlane(lt.y);
  const int nc = SERIAL ? K : __builtin_amdgcn_readfirstlane(lt.z);
  const int* co = L.coff + __builtin_amdgcn_readfirstlane(lt.w);
  const int q = tid & 3, j = tid >> 2;
  constexpr int PER = STRIDE / 4;
  if (nc <= 0) return 0.0;
  // end of colour c (a load; made scalar where it is used)
  auto col_end = [&](int c) { return SERIAL ? min(c + 1, K) : co[c + 1]; };
  // contact at position pos of a colour ending at end, or the null contact
  auto pick = [&](int pos, int end) { return pos < end ? c0 + pos : ResLds<R, H>::nul; };
  // rows, iterate and body slots of contact i (none of them change during a sweep)
  auto load = [&](int i, QuadRow<R>& it) {
    it.jl = L.JL[4 * i + q]; it.ja = L.JA[4 * i + q]; it.jb = L.JB[4 * i + q];
    it.xo = xa[4 * i + q];
    it.lc = L.loc[i];
  };
  // Phase c knows where colours c and c + 1 end.  Behind its velocity loads
  // it reads this quad's first contact of colour c + 1 and where colour c + 2
  // ends, so a phase waits on one LDS round trip, not two or three.  The two
  // row buffers take turns, so nothing is copied between phases.
  int lo = SERIAL ? 0 : __builtin_amdgcn_readfirstlane(co[0]);
  int hi = __builtin_amdgcn_readfirstlane(col_end(0)), hn = __builtin_amdgcn_readfirstlane(col_end(1)), c = 0;
  QuadRow<R> ra, rb;
  load(pick(lo + j, hi), ra);
  auto phase = [&](const QuadRow<R>& cur, QuadRow<R>& nxt) {
    const int i0 = pick(lo + j, hi);
    int h2 = 0;
    R xq;
    quad_core<R, JAC>(pp, cur, q, rs, hybrid, L.B, L.zsb, L.tsb, L.JL, L.JA, L.JB, 4 * (size_t)i0, xq, [&] {
      load(pick(hi + j, hn), nxt);
      h2 = col_end(c + 2);
    });
    xn[4 * i0 + q] = xq;
    // further contacts of this quad in the colour (a colour wider than the quads)
    for (int pos = lo + j + PER; pos < hi; pos += PER) {
      const int i = c0 + pos;
      QuadRow<R> it;
      load(i, it);
      quad_core<R, JAC>(pp, it, q, rs, hybrid, L.B, L.zsb, L.tsb, L.JL, L.JA, L.JB, 4 * (size_t)i, xq, [] {});
      xn[4 * i + q] = xq;
    }
    phase_sync<STRIDE>();
    lo = hi; hi = hn; hn = __builtin_amdgcn_readfirstlane(h2); ++c;
  };
  while (true) {
    phase(ra, rb);
    if (c >= nc) break;
    phase(rb, ra);
    if (c >= nc) break;
  }
  double m = 0.0;
  for (int r = tid; r < 4 * K; r += STRIDE)
    m = fmax(m, fabs((double)xn[4 * c0 + r] - (double)xa[4 * c0 + r]));
  return m;
}

// merit meets, accept/reject decided, velocity changes settled (all groups of the scene)
template <typename R, bool COOP>
__device__ __forceinline__ int sweep_decide(const Dev<R>& d, const SceneParams& sp, SolveCtl& ctl, double m,
                                            int it, bool lead, double* sh_red, int* s_ok, const BodyLds<R>& B, int nb,
                                            bool& abort) {
  double merit = block_max(m, sh_red);
  if (COOP && !grid_merit(d.merit, it, merit, d.err, s_ok, sh_red)) { abort = true; return 0; }
  double* res = residual_record(d, sp);
  const int dec = ctl.decide(sp, merit, (lead && threadIdx.x == 0) ? res : nullptr, it,
                             hinge_merit(d, sp, it));
  for (int i = threadIdx.x; i < nb; i += blockDim.x) {
    if (dec & 1) { B.sl[i] = B.wl[i]; B.sa[i] = B.wa[i]; }
    else { B.wl[i] = B.sl[i]; B.wa[i] = B.sa[i]; }
  }
  __syncthreads();
  return dec;
}

template <typename R> __device__ void stage_bodies(const Dev<R>& d, const BodyLds<R>& B, int b0, int nb, int zs) {
  for (int i = threadIdx.x; i < nb; i += blockDim.x) {
    const int b = d.u_blist[b0 + i];
    B.wl[i] = d.ww0[b]; B.wa[i] = d.ww1[b]; B.sl[i] = d.ws0[b]; B.sa[i] = d.ws1[b];
    B.W0[i] = d.b_W0[b]; B.W1[i] = d.b_W1[b];
  }
  if (threadIdx.x == 0) {
    const V4<R> z = {R(0), R(0), R(0), R(0)};
    B.wl[zs] = z; B.wa[zs] = z; B.sl[zs] = z; B.sa[zs] = z; B.W0[zs] = z; B.W1[zs] = z;
  }
}

template <typename R, bool COOP>
__device__ void finish_group(const Dev<R>& d, SceneState& st, const SolveCtl& ctl, bool lead, const BodyLds<R>& B,
                             int b0, int nb, double rloc, double xmax, double* sh_red) {
  for (int i = threadIdx.x; i < nb; i += blockDim.x) {
    const int b = d.u_blist[b0 + i];
    d.ws0[b] = B.sl[i];
    d.ws1[b] = B.sa[i];
  }
  const double rm = block_max(rloc, sh_red), xm = block_max(xmax, sh_red);
  if (threadIdx.x == 0) {
    if (d.post) {
      // the natural-map residual stays the contact solve's
    } else if (COOP) {
      if (rm > 0.0) atomic_max_nonneg(&st.nat_r, rm);
      if (xm > 0.0) atomic_max_nonneg(&st.nat_x, xm);
    } else {
      st.nat_r = rm;
      st.nat_x = xm;
    }
    if (lead) write_solve_state(d, st, ctl);
  }
}

template <typename R, bool COOP, int NT, int H>
__device__ void solve_group_resident(const Dev<R>& d, const int slot, const Group& gr, const GroupSize& gs,
                                     unsigned char* smem, double* sh_red, int* s_ok) {
  const SceneParams& sp = d.sp[slot];
  SceneState& st = d.ss[slot];
  const bool serial = sp.scheme == SCHEME_GS;
  const bool jacobi = sp.scheme == SCHEME_JACOBI || sp.scheme == SCHEME_JACOBI_HYBRID;
  const bool hybrid = sp.scheme == SCHEME_JACOBI_HYBRID;
  const ProxParams pp(sp);
  ResLds<R, H> L(smem);
  const int b0 = gr.b_lo, nb = gr.b_hi - gr.b_lo;
  // stage: lane table, colour offsets, contact rows and iterates, bodies
  if (threadIdx.x == 0) {
    int t = 0, o = 0;
    for (int l = gr.l_lo; l < gr.l_hi; ++l) {
      const int a = d.l_start[l], e = d.l_end[l];
      if (e <= a) continue;
      const int nc = serial ? 0 : d.l_ncol[l];
      L.lt[t++] = make_int4(a - gs.p0, e - a, nc, o);
      o += serial ? 0 : nc + 3;
    }
  }
  __syncthreads();
  if (!serial) {
    int t = 0;
    for (int l = gr.l_lo; l < gr.l_hi; ++l) {
      if (d.l_end[l] <= d.l_start[l]) continue;
      const int4 lt = L.lt[t++];
      const int* src = lane_coff(d, l);
      // offsets 0..nc, then co[nc] twice: the (empty) colours after the last
      for (int c = threadIdx.x; c <= lt.z + 2; c += blockDim.x) L.coff[lt.w + c] = src[min(c, lt.z)];
    }
  }
  for (int r = threadIdx.x; r < 4 * gs.nc; r += blockDim.x) {
    const size_t g = 4 * (size_t)gs.p0 + r;
    L.JL[r] = d.s_JL[g]; L.JA[r] = d.s_JA[g]; L.JB[r] = d.s_JB[g];
    L.x0[r] = d.xr0[g]; L.x1[r] = d.xr1[g];
  }
  for (int i = threadIdx.x; i < gs.nc; i += blockDim.x) {
    const int2 lc = d.s_loc[gs.p0 + i];
    L.loc[i] = make_int2((lc.x >= 0 ? lc.x - b0 : L.zs) * rec_bytes<R>(), (lc.y >= 0 ? lc.y - b0 : L.zs) * rec_bytes<R>());
  }
  // the null contact: zero rows and iterates, both bodies fixed
  if (threadIdx.x < 4) {
    const V4<R> z = {R(0), R(0), R(0), R(0)};
    const int r = 4 * ResLds<R, H>::nul + threadIdx.x;
    L.JL[r] = z; L.JA[r] = z; L.JB[r] = z;
    L.x0[r] = R(0); L.x1[r] = R(0);
    if (threadIdx.x == 0) L.loc[ResLds<R, H>::nul] = make_int2(L.zsb, L.zsb);
  }
  stage_bodies(d, L.B, b0, nb, L.zs);
  __syncthreads();
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, nwave = blockDim.x >> 6;
  SolveCtl ctl;
  ctl.iters = sp.max_iter;
  bool abort = false;
  for (int it = 0; it < sp.max_iter; ++it) {
    const R* xa = ctl.acc ? L.x1 : L.x0;
    R* xn = ctl.acc ? L.x0 : L.x1;
    const R rs = R(ctl.rscale);
    double m = 0.0;
    for (int t = wave; t < gs.nl; t += nwave) {
      const int4 lt = L.lt[t];
      if (lt.y > WAVE_LANE_MAX && !serial) continue;
      m = fmax(m, serial   ? res_sweep<R, 64, true, false, H>(pp, L, lt, hybrid, xa, xn, rs, lane)
                  : jacobi ? res_sweep<R, 64, false, true, H>(pp, L, lt, hybrid, xa, xn, rs, lane)
                           : res_sweep<R, 64, false, false, H>(pp, L, lt, hybrid, xa, xn, rs, lane));
    }
    for (int t = 0; t < gs.nl; ++t) {
      const int4 lt = L.lt[t];
      if (lt.y <= WAVE_LANE_MAX || serial) continue;
      m = fmax(m, jacobi ? res_sweep<R, NT, false, true, H>(pp, L, lt, hybrid, xa, xn, rs, (int)threadIdx.x)
                         : res_sweep<R, NT, false, false, H>(pp, L, lt, hybrid, xa, xn, rs, (int)threadIdx.x));
    }
    const int dec = sweep_decide<R, COOP>(d, sp, ctl, m, it, gr.lead, sh_red, s_ok, L.B, nb, abort);
    if (abort) return;
    if (dec & 2) { ctl.iters = it; break; }
  }
  // impulses to contact order; natural-map residual (FP64) with r0; accepted velocity changes out
  const R* xs = ctl.acc ? L.x1 : L.x0;
  double rloc = 0.0, xmax = 0.0;
  for (int i = threadIdx.x; i < gs.nc; i += blockDim.x) {
    const int2 lc = L.loc[i];
    V4<R> x;
    quad_finish_contact(pp, L.JL, L.JA, L.JB, xs, i, L.B.rec(lc.x, F_SL), L.B.rec(lc.x, F_SA), L.B.rec(lc.y, F_SL),
                        L.B.rec(lc.y, F_SA), x,
                        rloc, xmax);
    if (!d.post) d.c_imp[d.perm[gs.p0 + i]] = x;
  }
  finish_group<R, COOP>(d, st, ctl, gr.lead, L.B, b0, nb, rloc, xmax, sh_red);
}

// One sweep of one lane in the streaming layout: rows and iterates from
// global memory, the next pass prefetched into the other of two buffers
// while the current one is updated.
template <typename R, int STRIDE, bool JAC>
__device__ double stream_sweep(const Dev<R>& d, const ProxParams& pp, const BodyLds<R>& B, int zs, int b0, int p0,
                               int nc, const int* coff, bool serial, bool hybrid, const R* xa, R* xn, R rs, int tid) {
  constexpr int PER = STRIDE / 4;
  const int q = tid & 3, j = tid >> 2;
  double m = 0.0;
  if (nc <= 0) return m;
  int c = 0;
  int pb = serial ? 0 : coff[0];                        // first position of the current pass
  int hi = serial ? 1 : coff[1];                        // end of the current colour
  int hn = nc > 1 ? (serial ? 2 : coff[2]) : hi;        // end of the next colour
  QuadRow<R> ia, ib;
  bool ha = pb + j < hi;
  if (ha) load_quad(d.s_JL, d.s_JA, d.s_JB, xa, d.s_loc, p0 + pb + j, q, b0, zs, ia);
  auto step = [&](QuadRow<R>& use, bool have, QuadRow<R>& fill, bool& fhave) {
    const bool same = pb + PER < hi;
    int c1, pb1, hi1, hn1;
    if (same) {
      c1 = c; pb1 = pb + PER; hi1 = hi; hn1 = hn;
    } else {
      c1 = c + 1; pb1 = hi; hi1 = hn;
      hn1 = c1 + 1 < nc ? (serial ? c1 + 2 : coff[c1 + 2]) : hi1;
    }
    fhave = c1 < nc && pb1 + j < hi1;
    if (fhave) load_quad(d.s_JL, d.s_JA, d.s_JB, xa, d.s_loc, p0 + pb1 + j, q, b0, zs, fill);
    if (have) {
      R xq;
      m = fmax(m, quad_core<R, JAC>(pp, use, q, rs, hybrid, B, zs * rec_bytes<R>(), (zs + 1) * rec_bytes<R>(),
                                    d.s_JL, d.s_JA, d.s_JB,
                            4 * (size_t)(p0 + pb + j), xq, [] {}));
      xn[4 * (size_t)(p0 + pb + j) + q] = xq;
    }
    if (!same) phase_sync<STRIDE>();
    c = c1; pb = pb1; hi = hi1; hn = hn1;
  };
  bool hb = false;
  while (true) {
    step(ia, ha, ib, hb);
    if (c >= nc) break;
    step(ib, hb, ia, ha);
    if (c >= nc) break;
  }
  return m;
}

template <typename R, bool COOP, int NT, int H>
__device__ void solve_group_stream(const Dev<R>& d, const int slot, const Group& gr, unsigned char* smem,
                                   double* sh_red, int* s_ok) {
  const SceneParams& sp = d.sp[slot];
  SceneState& st = d.ss[slot];
  const bool serial = sp.scheme == SCHEME_GS;
  const bool jacobi = sp.scheme == SCHEME_JACOBI || sp.scheme == SCHEME_JACOBI_HYBRID;
  const bool hybrid = sp.scheme == SCHEME_JACOBI_HYBRID;
  const ProxParams pp(sp);
  StreamLds<R, H> L(smem);
  const int b0 = gr.b_lo, nb = gr.b_hi - gr.b_lo;
  stage_bodies(d, L.B, b0, nb, L.zs);
  __syncthreads();
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, nwave = blockDim.x >> 6;
  SolveCtl ctl;
  ctl.iters = sp.max_iter;
  bool abort = false;
  for (int it = 0; it < sp.max_iter; ++it) {
    const R* xa = ctl.acc ? d.xr1 : d.xr0;
    R* xn = ctl.acc ? d.xr0 : d.xr1;
    const R rs = R(ctl.rscale);
    double m = 0.0;
    for (int l = gr.l_lo + wave; l < gr.l_hi; l += nwave) {
      const int p0 = d.l_start[l], K = d.l_end[l] - p0;
      if (K <= 0 || (K > WAVE_LANE_MAX && !serial)) continue;
      const int nc = serial ? K : d.l_ncol[l];
      m = fmax(m, jacobi ? stream_sweep<R, 64, true>(d, pp, L.B, L.zs, b0, p0, nc, lane_coff(d, l), serial, hybrid,
                                                     xa, xn, rs, lane)
                         : stream_sweep<R, 64, false>(d, pp, L.B, L.zs, b0, p0, nc, lane_coff(d, l), serial, hybrid,
                                                      xa, xn, rs, lane));
    }
    for (int l = gr.l_lo; l < gr.l_hi; ++l) {
      const int p0 = d.l_start[l], K = d.l_end[l] - p0;
      if (K <= WAVE_LANE_MAX || serial) continue;
      m = fmax(m, jacobi ? stream_sweep<R, NT, true>(d, pp, L.B, L.zs, b0, p0, d.l_ncol[l], lane_coff(d, l), serial,
                                                     hybrid, xa, xn, rs, (int)threadIdx.x)
                         : stream_sweep<R, NT, false>(d, pp, L.B, L.zs, b0, p0, d.l_ncol[l], lane_coff(d, l), serial,
                                                      hybrid, xa, xn, rs, (int)threadIdx.x));
    }
    const int dec = sweep_decide<R, COOP>(d, sp, ctl, m, it, gr.lead, sh_red, s_ok, L.B, nb, abort);
    if (abort) return;
    if (dec & 2) { ctl.iters = it; break; }
  }
  const R* xs = ctl.acc ? d.xr1 : d.xr0;
  double rloc = 0.0, xmax = 0.0;
  for (int l = gr.l_lo; l < gr.l_hi; ++l) {
    const int p0 = d.l_start[l], K = d.l_end[l] - p0;
    for (int pos = threadIdx.x; pos < K; pos += blockDim.x) {
      const int i = p0 + pos;
      const int2 lc = d.s_loc[i];
      const int sa = lc.x >= 0 ? lc.x - b0 : L.zs, sb = lc.y >= 0 ? lc.y - b0 : L.zs;
      V4<R> x;
      quad_finish_contact(pp, d.s_JL, d.s_JA, d.s_JB, xs, i, L.B.sl[sa], L.B.sa[sa], L.B.sl[sb], L.B.sa[sb], x,
                          rloc, xmax);
      if (!d.post) d.c_imp[d.perm[i]] = x;
    }
  }
  finish_group<R, COOP>(d, st, ctl, gr.lead, L.B, b0, nb, rloc, xmax, sh_red);
}

template <typename R> __device__ __forceinline__ void no_contacts(SceneState& st) {
  st.solved = 0; st.iterations = 0; st.exitcode = 0; st.sweeps = 0; st.rejects = 0; st.final_merit = 0.0;
}

template <typename R, bool COOP, int NT, int H>
__device__ __forceinline__ void solve_dispatch(const Dev<R>& d, int slot, const Group& gr, unsigned char* smem,
                                               double* sh_red, int* s_ok) {
  const GroupSize gs = group_size(d, gr, d.sp[slot].scheme == SCHEME_GS);
  V4<R>* vs = reinterpret_cast<V4<R>*>(smem);
  if (group_fits<R, H>(d, gr, gs)) solve_group_resident<R, COOP, NT, H>(d, slot, gr, gs, smem, sh_red, s_ok);
  else if (gr.b_hi - gr.b_lo <= ResCap<R, H>::VB) solve_group_stream<R, COOP, NT, H>(d, slot, gr, smem, sh_red, s_ok);
  else solve_group<R, false, COOP, NT>(d, slot, gr, vs, sh_red, s_ok);
}

// One workgroup of NT threads (NT / 64 wavefronts, one per sweep lane) per
// batch scene; H: the LDS layout (0 full, 1 half).
template <typename R, int NT, int H> __global__ void __launch_bounds__(NT) k_solve(Dev<R> d) {
  extern __shared__ __align__(32) unsigned char solve_smem[];
  __shared__ double sh_red[NT / 64];
  __shared__ int s_ok;
  if (aborted(d.err)) return;
  const int bs = d.bs_order[blockIdx.x];
  const int slot = d.bs_slot[bs];
  const SceneParams& sp = d.sp[slot];
  SceneState& st = d.ss[slot];
  if (d.post && !(sp.use_post_stab && st.post_active)) return;
  double* res = residual_record(d, sp);
    for (int i = threadIdx.x; i < sp.max_iter; i += NT) res[i] = INFINITY;
  if (st.k_cnt == 0) {
    if (threadIdx.x == 0) no_contacts<R>(st);
    return;
  }
  const int u0 = d.u_base[bs], u1 = u0 + d.u_cnt[bs];
  const Group gr{d.l_base[bs], d.l_base[bs] + d.l_cnt[bs], d.u_boff[u0], d.u_boff[u1], true};
  solve_dispatch<R, false, NT, H>(d, slot, gr, solve_smem, sh_red, &s_ok);
}

// A single scene (its lanes are its colouring units): the lanes split into
// gridDim.x contiguous blocks, one per workgroup of a cooperative launch.
template <typename R> __global__ void __launch_bounds__(SOLVE_BLOCK) k_solve_multi(Dev<R> d) {
  extern __shared__ __align__(32) unsigned char solve_smem[];
  __shared__ double sh_red[SOLVE_BLOCK / 64];
  __shared__ int s_ok;
  if (aborted(d.err)) return;   // uniform: the flag is set before this launch or not at all
  const int slot = d.bs_slot[0];
  const SceneParams& sp = d.sp[slot];
  SceneState& st = d.ss[slot];
  if (d.post && !(sp.use_post_stab && st.post_active)) return;   // uniform over the grid
  double* res = residual_record(d, sp);
  if (blockIdx.x == 0)
    for (int i = threadIdx.x; i < sp.max_iter; i += SOLVE_BLOCK) res[i] = INFINITY;
  if (st.k_cnt == 0) {
    if (blockIdx.x == 0 && threadIdx.x == 0) no_contacts<R>(st);
    return;
  }
  const long long nl = *d.n_lanes, G = gridDim.x;
  const int l_lo = (int)(blockIdx.x * nl / G), l_hi = (int)((blockIdx.x + 1) * nl / G);
  const Group gr{l_lo, l_hi, d.u_boff[l_lo], d.u_boff[l_hi], blockIdx.x == 0};
  solve_dispatch<R, true, SOLVE_BLOCK, 0>(d, slot, gr, solve_smem, sh_red, &s_ok);
}
