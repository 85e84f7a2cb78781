// Signed-distance grid construction on the GPU (setup time).
//
// The reference builds shape SDFs with libigl's signed_distance
// (rainbow/geometry/grid3.py:320-339, called from api.py:234-241).  Here each
// thread owns one grid node: it computes the exact distance to every triangle
// (closest point by Voronoi-region tests, Ericson, "Real-Time Collision
// Detection" 5.1.5) and the generalized winding number (sum of signed solid
// angles, Van Oosterom & Strackee 1983); the sign is negative where the
// winding number exceeds 1/2.  Triangles are staged through LDS in tiles.
// Arithmetic is FP64 regardless of the step precision.
#pragma once
#include <hip/hip_runtime.h>

namespace rbh {

constexpr int SDF_BLOCK = 256;
constexpr int SDF_TILE = 256;

__device__ __forceinline__ double tri_sqdist(double px, double py, double pz, const double* t) {
  const double ax = t[0], ay = t[1], az = t[2], bx = t[3], by = t[4], bz = t[5], cx = t[6], cy = t[7], cz = t[8];
  double abx = bx - ax, aby = by - ay, abz = bz - az;
  double acx = cx - ax, acy = cy - ay, acz = cz - az;
  double apx = px - ax, apy = py - ay, apz = pz - az;
  double d1 = abx * apx + aby * apy + abz * apz;
  double d2 = acx * apx + acy * apy + acz * apz;
  double qx, qy, qz;
  if (d1 <= 0.0 && d2 <= 0.0) { qx = ax; qy = ay; qz = az; goto done; }
  {
    double bpx = px - bx, bpy = py - by, bpz = pz - bz;
    double d3 = abx * bpx + aby * bpy + abz * bpz;
    double d4 = acx * bpx + acy * bpy + acz * bpz;
    if (d3 >= 0.0 && d4 <= d3) { qx = bx; qy = by; qz = bz; goto done; }
    double vc = d1 * d4 - d3 * d2;
    if (vc <= 0.0 && d1 >= 0.0 && d3 <= 0.0) {
      double v = d1 / (d1 - d3);
      qx = ax + v * abx; qy = ay + v * aby; qz = az + v * abz; goto done;
    }
    double cpx = px - cx, cpy = py - cy, cpz = pz - cz;
    double d5 = abx * cpx + aby * cpy + abz * cpz;
    double d6 = acx * cpx + acy * cpy + acz * cpz;
    if (d6 >= 0.0 && d5 <= d6) { qx = cx; qy = cy; qz = cz; goto done; }
    double vb = d5 * d2 - d1 * d6;
    if (vb <= 0.0 && d2 >= 0.0 && d6 <= 0.0) {
      double w = d2 / (d2 - d6);
      qx = ax + w * acx; qy = ay + w * acy; qz = az + w * acz; goto done;
    }
    double va = d3 * d6 - d5 * d4;
    if (va <= 0.0 && (d4 - d3) >= 0.0 && (d5 - d6) >= 0.0) {
      double w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
      qx = bx + w * (cx - bx); qy = by + w * (cy - by); qz = bz + w * (cz - bz); goto done;
    }
    double den = 1.0 / (va + vb + vc);
    double v = vb * den, w = vc * den;
    qx = ax + abx * v + acx * w; qy = ay + aby * v + acy * w; qz = az + abz * v + acz * w;
  }
done:
  double dx = px - qx, dy = py - qy, dz = pz - qz;
  return dx * dx + dy * dy + dz * dz;
}

__device__ __forceinline__ double solid_angle(double px, double py, double pz, const double* t) {
  double ax = t[0] - px, ay = t[1] - py, az = t[2] - pz;
  double bx = t[3] - px, by = t[4] - py, bz = t[5] - pz;
  double cx = t[6] - px, cy = t[7] - py, cz = t[8] - pz;
  double la = sqrt(ax * ax + ay * ay + az * az), lb = sqrt(bx * bx + by * by + bz * bz),
         lc = sqrt(cx * cx + cy * cy + cz * cz);
  double det = ax * (by * cz - bz * cy) + ay * (bz * cx - bx * cz) + az * (bx * cy - by * cx);
  double den = la * lb * lc + (ax * bx + ay * by + az * bz) * lc + (bx * cx + by * cy + bz * cz) * la +
               (cx * ax + cy * ay + cz * az) * lb;
  return 2.0 * atan2(det, den);
}

// tri[9 t + 3 k + a] = V[3 T[3 t + k] + a]: the corners of each triangle in
// one record (triangle indices checked on the host)
__global__ void k_sdf_gather(const double* __restrict__ V, const int* __restrict__ T, int nt,
                             double* __restrict__ tri) {
  const long n = 9L * nt;
  for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) {
    const long t = e / 9;
    const int r = (int)(e - 9 * t);
    tri[e] = V[3L * T[3 * t + r / 3] + r % 3];
  }
}

// tri: 9 doubles per triangle; out[i + I (j + J k)] = signed distance of node (i, j, k)
__global__ void __launch_bounds__(SDF_BLOCK) k_sdf(const double* __restrict__ tri, int nt, int I, int J, int K,
                                                   double mnx, double mny, double mnz, double hx, double hy, double hz,
                                                   double* __restrict__ out) {
  __shared__ double st[SDF_TILE * 9];
  const long n = (long)I * J * K;
  const long node = (long)blockIdx.x * SDF_BLOCK + threadIdx.x;
  double px = 0, py = 0, pz = 0;
  if (node < n) {
    int i = (int)(node % I);
    int j = (int)((node / I) % J);
    int k = (int)(node / ((long)I * J));
    // grid3.py:88-105: min_coord + spacing * (i, j, k)
    px = mnx + hx * (double)i; py = mny + hy * (double)j; pz = mnz + hz * (double)k;
  }
  double best = INFINITY, wsum = 0.0;
  for (int t0 = 0; t0 < nt; t0 += SDF_TILE) {
    int cnt = min(SDF_TILE, nt - t0);
    for (int e = threadIdx.x; e < cnt * 9; e += SDF_BLOCK) st[e] = tri[(long)t0 * 9 + e];
    __syncthreads();
    if (node < n) {
      for (int e = 0; e < cnt; ++e) {
        const double* t = st + e * 9;
        best = fmin(best, tri_sqdist(px, py, pz, t));
        wsum += solid_angle(px, py, pz, t);
      }
    }
    __syncthreads();
  }
  if (node < n) {
    double wn = wsum / (4.0 * 3.14159265358979323846);
    double dist = sqrt(best);
    out[node] = fabs(wn) > 0.5 ? -dist : dist;
  }
}

}  // namespace rbh
