// Device-side helpers for the rigid-body contact step: small vector and
// quaternion algebra, trilinear SDF lookups, proximal operators.
//
// Every routine is templated on the arithmetic type R (float for the product
// path, double for the FP64 parity build of the same kernels).  Expression
// order follows the reference formulas cited at each function, so the FP64
// instantiation reproduces the reference arithmetic up to FMA contraction.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace rbh {

template <typename R> struct alignas(4 * sizeof(R)) V4 { R x, y, z, w; };
template <typename R> struct V3 { R x, y, z; };

template <typename R> __device__ __forceinline__ V3<R> v3(R x, R y, R z) { return V3<R>{x, y, z}; }
template <typename R> __device__ __forceinline__ V3<R> xyz(const V4<R>& a) { return V3<R>{a.x, a.y, a.z}; }
template <typename R> __device__ __forceinline__ V3<R> operator+(V3<R> a, V3<R> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <typename R> __device__ __forceinline__ V3<R> operator-(V3<R> a, V3<R> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <typename R> __device__ __forceinline__ V3<R> operator*(R s, V3<R> a) { return {s * a.x, s * a.y, s * a.z}; }
template <typename R> __device__ __forceinline__ V3<R> operator-(V3<R> a) { return {-a.x, -a.y, -a.z}; }
template <typename R> __device__ __forceinline__ R dot(V3<R> a, V3<R> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <typename R> __device__ __forceinline__ V3<R> cross(V3<R> a, V3<R> b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
template <typename R> __device__ __forceinline__ R norm(V3<R> a) { return sqrt(a.x * a.x + a.y * a.y + a.z * a.z); }

// ---------------------------------------------------------------- quaternions
// q = (s, x, y, z) stored in V4 as (x=s, y=qx, z=qy, w=qz).
template <typename R> struct Quat { R s, x, y, z; };

template <typename R> __device__ __forceinline__ Quat<R> quat(const V4<R>& a) { return Quat<R>{a.x, a.y, a.z, a.w}; }
template <typename R> __device__ __forceinline__ Quat<R> conj(Quat<R> q) { return Quat<R>{q.s, -q.x, -q.y, -q.z}; }

// quaternion.py:186-193 -- [a b - A.B, a B + b A + A x B]
template <typename R> __device__ __forceinline__ Quat<R> qprod(Quat<R> a, Quat<R> b) {
  R s = a.s * b.s - (a.x * b.x + a.y * b.y + a.z * b.z);
  R x = a.s * b.x + a.x * b.s + (a.y * b.z - a.z * b.y);
  R y = a.s * b.y + a.y * b.s + (a.z * b.x - a.x * b.z);
  R z = a.s * b.z + a.z * b.s + (a.x * b.y - a.y * b.x);
  return Quat<R>{s, x, y, z};
}

// quaternion.py:196-198 -- (q (0, r) q*)[1:]
template <typename R> __device__ __forceinline__ V3<R> qrotate(Quat<R> q, V3<R> r) {
  Quat<R> t = qprod(qprod(q, Quat<R>{R(0), r.x, r.y, r.z}), conj(q));
  return V3<R>{t.x, t.y, t.z};
}

// quaternion.py:201-250 -- rotate_array: prod_array(prod_array(q, (0,V)), conj(q))
template <typename R> __device__ __forceinline__ V3<R> qrotate_array(Quat<R> a, V3<R> v) {
  // inner = a (x) (0, v), written as prod_array(Qa, Qb) with Qb = (0, v)
  R i0 = a.s * R(0) - a.x * v.x - a.y * v.y - a.z * v.z;
  R i1 = a.s * v.x + a.x * R(0) + a.y * v.z - a.z * v.y;
  R i2 = a.s * v.y - a.x * v.z + a.y * R(0) + a.z * v.x;
  R i3 = a.s * v.z + a.x * v.y - a.y * v.x + a.z * R(0);
  Quat<R> b = conj(a);
  // outer = inner (x) b, prod_array(Qa_array, Qb)
  R o1 = i0 * b.x + i1 * b.s + i2 * b.z - i3 * b.y;
  R o2 = i0 * b.y - i1 * b.z + i2 * b.s + i3 * b.x;
  R o3 = i0 * b.z + i1 * b.y - i2 * b.x + i3 * b.s;
  return V3<R>{o1, o2, o3};
}

// quaternion.py:76-80
template <typename R> __device__ __forceinline__ Quat<R> qunit(Quat<R> q) {
  R l = sqrt(q.s * q.s + q.x * q.x + q.y * q.y + q.z * q.z);
  if (l > R(0)) return Quat<R>{q.s / l, q.x / l, q.y / l, q.z / l};
  return Quat<R>{R(1), R(0), R(0), R(0)};
}

// quaternion.py:108-129 (normalises first). Row-major 3x3.
template <typename R> __device__ __forceinline__ void qmatrix(Quat<R> Q, R M[9]) {
  Quat<R> q = qunit(Q);
  R s = q.s, x = q.x, y = q.y, z = q.z;
  M[0] = R(1) - R(2) * (y * y + z * z);
  M[4] = R(1) - R(2) * (x * x + z * z);
  M[8] = R(1) - R(2) * (y * y + x * x);
  M[3] = R(2) * (x * y + s * z);
  M[1] = R(2) * (x * y - s * z);
  M[6] = R(2) * (-s * y + x * z);
  M[2] = R(2) * (s * y + x * z);
  M[7] = R(2) * (z * y + s * x);
  M[5] = R(2) * (z * y - s * x);
}

// mass.py:395-452: W = R diag(B) R^T, packed symmetric (xx, yy, zz, xy, xz, yz).
template <typename R> __device__ __forceinline__ void rot_diag_rt(const R M[9], R b0, R b1, R b2, R W[6]) {
  W[0] = M[0] * M[0] * b0 + M[1] * M[1] * b1 + M[2] * M[2] * b2;
  W[1] = M[3] * M[3] * b0 + M[4] * M[4] * b1 + M[5] * M[5] * b2;
  W[2] = M[6] * M[6] * b0 + M[7] * M[7] * b1 + M[8] * M[8] * b2;
  W[3] = M[3] * M[0] * b0 + M[4] * M[1] * b1 + M[5] * M[2] * b2;
  W[4] = M[6] * M[0] * b0 + M[7] * M[1] * b1 + M[8] * M[2] * b2;
  W[5] = M[3] * M[6] * b0 + M[4] * M[7] * b1 + M[5] * M[8] * b2;
}

template <typename R> __device__ __forceinline__ V3<R> sym_mul(const R W[6], V3<R> a) {
  return V3<R>{W[0] * a.x + W[3] * a.y + W[4] * a.z,
               W[3] * a.x + W[1] * a.y + W[5] * a.z,
               W[4] * a.x + W[5] * a.y + W[2] * a.z};
}

// c + a * b for the contact sweep: one fused multiply-add in FP32; in FP64
// the product and the sum are rounded one by one, as the FP64 oracle does
template <typename R> __device__ __forceinline__ R madd(R a, R b, R c) {
  if constexpr (sizeof(R) == 4) return __builtin_fmaf(a, b, c);
  else return c + a * b;
}

// sym_mul with madd: the same sums in the same order
template <typename R> __device__ __forceinline__ V3<R> sym_madd(const R W[6], V3<R> a) {
  return V3<R>{madd(W[4], a.z, madd(W[3], a.y, W[0] * a.x)),
               madd(W[5], a.z, madd(W[1], a.y, W[3] * a.x)),
               madd(W[2], a.z, madd(W[5], a.y, W[4] * a.x))};
}

// vector3.py:75-113: (t, b, n) with t guessed from the axis of smallest |n_i|
// (the later axis wins ties) and b = unit(n x t), t = b x n.
template <typename R> __device__ __forceinline__ void orthonormal_frame(V3<R> nin, V3<R>& t, V3<R>& b, V3<R>& n) {
  R l = norm(nin);
  n = V3<R>{nin.x / l, nin.y / l, nin.z / l};
  R ax = fabs(n.x), ay = fabs(n.y), az = fabs(n.z);
  V3<R> g = {R(1), R(0), R(0)};
  if (ay <= ax && ay <= az) g = V3<R>{R(0), R(1), R(0)};
  if (az <= ax && az <= ay) g = V3<R>{R(0), R(0), R(1)};
  // the first test (ax smallest) only matters when neither later test fires
  V3<R> bb = cross(n, g);
  R lb = norm(bb);
  b = V3<R>{bb.x / lb, bb.y / lb, bb.z / lb};
  t = cross(b, n);
}

// ---------------------------------------------------------------- SDF grids
template <typename R> struct GridView {
  const R* val;
  const R* cells;   // cell-major copy: cell (i, j, k) holds its 8 corners in the order of load_cell
  int I, J, K;
  R mn[3], mx[3], h[3];
};

// grid3.py:107-123 -- floor((p - min) / h) clipped to [0, dim - 2]
template <typename R> __device__ __forceinline__ void cell_idx(const GridView<R>& g, V3<R> p, int& i, int& j, int& k) {
  R fi = floor((p.x - g.mn[0]) / g.h[0]);
  R fj = floor((p.y - g.mn[1]) / g.h[1]);
  R fk = floor((p.z - g.mn[2]) / g.h[2]);
  // clamp in floating point first so huge values cannot overflow the int cast
  fi = fmin(fmax(fi, R(0)), R(g.I - 2));
  fj = fmin(fmax(fj, R(0)), R(g.J - 2));
  fk = fmin(fmax(fk, R(0)), R(g.K - 2));
  i = (int)fi; j = (int)fj; k = (int)fk;
  if (!(i >= 0)) i = 0;   // NaN input
  if (!(j >= 0)) j = 0;
  if (!(k >= 0)) k = 0;
}

template <typename R> struct Cell { R d000, d001, d010, d011, d100, d101, d110, d111; };

// The 8 corner values of cell (i, j, k), read as two vec4 from the
// cell-major copy of the grid (k_sdf_cells) instead of 8 scattered values:
// (i,j,k), (i+1,j,k), (i,j+1,k), (i+1,j+1,k), then the same at k+1.
template <typename R> __device__ __forceinline__ Cell<R> load_cell(const GridView<R>& g, int i, int j, int k) {
  const V4<R>* p = reinterpret_cast<const V4<R>*>(g.cells) +
                   2 * ((size_t)i + (size_t)(g.I - 1) * ((size_t)j + (size_t)(g.J - 1) * (size_t)k));
  const V4<R> a = p[0], e = p[1];
  Cell<R> c;
  c.d000 = a.x; c.d100 = a.y; c.d010 = a.z; c.d110 = a.w;
  c.d001 = e.x; c.d101 = e.y; c.d011 = e.z; c.d111 = e.w;
  return c;
}

// cell-major copy of one grid (values v, I x J x K nodes, index i + I (j + J k))
template <typename R> __global__ void k_sdf_cells(const R* __restrict__ v, int I, int J, int K, R* __restrict__ out) {
  const long long nc = (long long)(I - 1) * (J - 1) * (K - 1);
  for (long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x; c < nc; c += (long long)gridDim.x * blockDim.x) {
    const int i = (int)(c % (I - 1));
    const long long r = c / (I - 1);
    const int j = (int)(r % (J - 1)), k = (int)(r / (J - 1));
    const long long b = i + (long long)I * (j + (long long)J * k), dj = I, dk = (long long)I * J;
    R* o = out + 8 * c;
    o[0] = v[b]; o[1] = v[b + 1]; o[2] = v[b + dj]; o[3] = v[b + 1 + dj];
    o[4] = v[b + dk]; o[5] = v[b + 1 + dk]; o[6] = v[b + dj + dk]; o[7] = v[b + 1 + dj + dk];
  }
}

// grid3.py:126-177
template <typename R> __device__ __forceinline__ R grid_value(const GridView<R>& g, V3<R> p) {
  int i, j, k;
  cell_idx(g, p, i, j, k);
  Cell<R> c = load_cell(g, i, j, k);
  R s = (p.x - ((R(i) * g.h[0]) + g.mn[0])) / g.h[0];
  R t = (p.y - ((R(j) * g.h[1]) + g.mn[1])) / g.h[1];
  R u = (p.z - ((R(k) * g.h[2]) + g.mn[2])) / g.h[2];
  R x00 = (c.d100 - c.d000) * s + c.d000;
  R x01 = (c.d101 - c.d001) * s + c.d001;
  R x10 = (c.d110 - c.d010) * s + c.d010;
  R x11 = (c.d111 - c.d011) * s + c.d011;
  R y0 = (x10 - x00) * t + x00;
  R y1 = (x11 - x01) * t + x01;
  return (y1 - y0) * u + y0;
}

// grid3.py:180-254 -- local coordinates divided by min_coord (grid3.py:209-211)
template <typename R> __device__ __forceinline__ V3<R> grid_gradient(const GridView<R>& g, V3<R> p) {
  int i, j, k;
  cell_idx(g, p, i, j, k);
  Cell<R> c = load_cell(g, i, j, k);
  R s = (p.x - (R(i) * g.h[0] + g.mn[0])) / g.mn[0];
  R t = (p.y - (R(j) * g.h[1] + g.mn[1])) / g.mn[1];
  R u = (p.z - (R(k) * g.h[2] + g.mn[2])) / g.mn[2];
  R x00 = (c.d100 - c.d000) * s + c.d000;
  R x01 = (c.d101 - c.d001) * s + c.d001;
  R x10 = (c.d110 - c.d010) * s + c.d010;
  R x11 = (c.d111 - c.d011) * s + c.d011;
  R y0 = (x10 - x00) * t + x00;
  R y1 = (x11 - x01) * t + x01;
  R dx00 = c.d100 - c.d000, dx01 = c.d101 - c.d001, dx10 = c.d110 - c.d010, dx11 = c.d111 - c.d011;
  R dy0_ds = (dx10 - dx00) * t + dx00;
  R dy1_ds = (dx11 - dx01) * t + dx01;
  R dy0_dt = x10 - x00;
  R dy1_dt = x11 - x01;
  R dp_ds = (dy1_ds - dy0_ds) * u + dy0_ds;
  R dp_dt = (dy1_dt - dy0_dt) * u + dy0_dt;
  R dp_du = y1 - y0;
  return V3<R>{dp_ds * (R(1) / g.h[0]), dp_dt * (R(1) / g.h[1]), dp_du * (R(1) / g.h[2])};
}

// grid3.py:275-300 (strict comparisons; equality counts as inside)
template <typename R> __device__ __forceinline__ bool grid_inside(const GridView<R>& g, V3<R> p, R boundary) {
  if (p.x > (g.mx[0] - boundary) || p.y > (g.mx[1] - boundary) || p.z > (g.mx[2] - boundary)) return false;
  if (p.x < (g.mn[0] + boundary) || p.y < (g.mn[1] + boundary) || p.z < (g.mn[2] + boundary)) return false;
  return true;
}

// ---------------------------------------------------------------- prox ops
// prox_operators.py:14-34
template <typename R> __device__ __forceinline__ void prox_sphere(R zs, R zt, R zr, R mu, R xn, R& os, R& ot, R& orr) {
  if (xn <= R(0)) { os = ot = orr = R(0); return; }
  R radius = mu * xn;
  R sq = zs * zs + zt * zt + zr * zr;
  if (sq <= radius * radius) { os = zs; ot = zt; orr = zr; return; }
  // FP32: one hardware reciprocal square root (1 ulp) in place of a correctly
  // rounded sqrt and divide; FP64 keeps radius / sqrt(sq) as the reference has it
  R scale;
  if constexpr (sizeof(R) == 4) scale = radius * __builtin_amdgcn_rsqf(sq);
  else scale = radius / sqrt(sq);
  os = zs * scale; ot = zt * scale; orr = zr * scale;
}

// First term of a velocity-change sum: FP64 starts from 0 as the reference's
// sparse products do (0 + p only differs from p for p = -0); FP32 starts from p.
template <typename R> __device__ __forceinline__ R sum_first(R p) {
  if constexpr (sizeof(R) == 8) return R(0) + p;
  else return p;
}

// proximal_contact/proximal_operators.py:42-139, evaluated in FP64.
struct EllipsoidParams { double tol, expansion; int max_it; };

__device__ __forceinline__ double ell_g(double saa, double sbb, double scc, double sxx, double syy, double szz, double t) {
  return (saa * sxx) / ((saa + t) * (saa + t)) + (sbb * syy) / ((sbb + t) * (sbb + t)) +
         (scc * szz) / ((scc + t) * (scc + t)) - 1.0;
}

template <typename R>
__device__ __forceinline__ void prox_ellipsoid(R zs_, R zt_, R zr_, R mus_, R mut_, R mur_, R xn_, const EllipsoidParams& ep,
                                               R& os, R& ot, R& orr) {
  double zs = zs_, zt = zt_, zr = zr_, ms = mus_, mt = mut_, mr = mur_, xn = xn_;
  if (xn <= 0.0) { os = ot = orr = R(0); return; }
  if (ms == 0.0 && mt == 0.0 && mr == 0.0) { os = ot = orr = R(0); return; }
  if (ms == mt && ms == mr) { prox_sphere<R>(zs_, zt_, zr_, mus_, xn_, os, ot, orr); return; }
  double a = ms * xn, b = mt * xn, c = mr * xn;
  double m = fmax(fmax(fmax(1.0, a), fmax(b, c)), fmax(fmax(fabs(zs), fabs(zt)), fabs(zr)));
  double scale = 1.0 / m;
  double sa = scale * a, sb = scale * b, sc = scale * c, sx = scale * zs, sy = scale * zt, sz = scale * zr;
  double saa = sa * sa, sbb = sb * sb, scc = sc * sc, sxx = sx * sx, syy = sy * sy, szz = sz * sz;
  double f0 = (sxx / saa) + (syy / sbb) + (szz / scc) - 1.0;
  if (f0 < ep.tol) { os = zs_; ot = zt_; orr = zr_; return; }
  double t0 = 0.0;
  double t1 = fmax(fmax(sa, sb), sc) * sqrt(sx * sx + sy * sy + sz * sz);
  double g1 = ell_g(saa, sbb, scc, sxx, syy, szz, t1);
  for (int guard = 0; g1 > 0.0 && guard < 4096; ++guard) {
    t1 *= ep.expansion;
    g1 = ell_g(saa, sbb, scc, sxx, syy, szz, t1);
  }
  double tk = (t0 + t1) * 0.5;
  for (int it = 0; it < ep.max_it; ++it) {
    if (fabs(t1 - t0) < ep.tol) break;
    double gk = ell_g(saa, sbb, scc, sxx, syy, szz, tk);
    if (fabs(gk) < ep.tol) break;
    if (gk > 0.0) t0 = tk; else t1 = tk;
    tk = (t0 + t1) * 0.5;
  }
  tk /= scale * scale;
  os = R((a * a * zs) / (a * a + tk));
  ot = R((b * b * zt) / (b * b + tk));
  orr = R((c * c * zr) / (c * c + tk));
}

// out of line: anisotropic friction is the rare case, and inlining its
// bisection into every sweep loop only lengthens the isotropic path's code
template <typename R>
__device__ __attribute__((noinline)) V3<R> prox_ellipsoid_call(R zs, R zt, R zr, R ms, R mt, R mr, R xn, double tol,
                                                              double expansion, int max_it) {
  EllipsoidParams ep{tol, expansion, max_it};
  V3<R> o;
  prox_ellipsoid<R>(zs, zt, zr, ms, mt, mr, xn, ep, o.x, o.y, o.z);
  return o;
}

// ---------------------------------------------------------------- misc
__device__ __forceinline__ uint64_t mix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

__device__ __forceinline__ uint32_t color_draw(uint64_t seed, uint64_t step, uint64_t v, uint64_t rnd, uint32_t n) {
  uint64_t h = mix64(mix64(mix64(mix64(seed) ^ step) ^ v) ^ rnd);
  return (uint32_t)(((h >> 32) * (uint64_t)n) >> 32);
}

// Non-negative doubles order like their bit patterns: max via integer atomics.
__device__ __forceinline__ void atomic_max_nonneg(double* addr, double v) {
  atomicMax(reinterpret_cast<unsigned long long*>(addr), (unsigned long long)__double_as_longlong(v));
}

}  // namespace rbh
