// Per-pixel device code shared by the single-component render (render.hip)
// and the batched render (batch.hip): pixel offsets, quad:N sampling and the
// float64 adaptive Gauss-Legendre quadrature of one pixel
// (utils/operations.py:105-247).
#pragma once

#include <hip/hip_runtime.h>

#include "../../include/astrophot_hip.h"
#include "profile.h"

namespace apx {

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// offsets of a working pixel from the centre
template <typename T>
__device__ __forceinline__ void pixel_xy(const Prof<T>& p, int ci, int cj, int i, int j, T& X,
                                         T& Y) {
  const T di = (T)(i - ci);
  const T dj = (T)(j - cj);
  X = p.P00 * di + p.P01 * dj + p.fx;
  Y = p.P10 * di + p.P11 * dj + p.fy;
}
template <typename T>
__device__ __forceinline__ void pixel_xy(const Prof<T>& p, const apx_component& cp, int i, int j,
                                         T& X, T& Y) {
  pixel_xy(p, cp.ci, cp.cj, i, j, X, Y);
}

// quad:N sample of one pixel (operations.py:123-147): returns the weighted
// sum, `centre` the middle node; accumulates weighted slot gradients in g.
template <typename T, bool GRAD>
__device__ T quad_sample(const Prof<T>& p, const apx_component& cp, T X, T Y, int n,
                         const double* node, const double* weight, T scale, T hP00, T hP01,
                         T hP10, T hP11, T* g, T& centre) {
  T res = (T)0;
  T gl[APX_NSLOT];
  if (GRAD) {
#pragma unroll
    for (int k = 0; k < APX_NSLOT; ++k) g[k] = (T)0;
  }
  for (int k = 0; k < n * n; ++k) {
    const T u = (T)node[k % n], v = (T)node[k / n];
    const T w = (T)(weight[k % n] * weight[k / n] * 0.25);
    const T val = eval_profile<T, GRAD>(p, X + hP00 * u + hP01 * v, Y + hP10 * u + hP11 * v,
                                        scale, gl);
    res += w * val;
    if (k == (n * n) / 2) centre = val;
    if (GRAD) {
#pragma unroll
      for (int s = 0; s < APX_NSLOT; ++s) g[s] += w * gl[s];
    }
  }
  return res;
}


// Relative band around a threshold inside which a float32 test is taken again
// in float64 (float32 profile samples carry far less relative error).
constexpr double kSampleBand = 1e-4;

// Adaptive quadrature of one pixel by one wavefront (all lanes get the
// result): out[0] = value, out[1 + slot] = d value / d slot when GRAD.
// Level 1, the value of every selected pixel that is not refined further, is
// sampled in float64.  Levels 2 and 3 are sampled in float32 and
// summed in float64; a level-2 test whose float32 value lies within
// kSampleBand of its threshold is taken again from float64 samples, so the
// sub-pixels refined are the ones float64 arithmetic picks.
template <bool GRAD>
__device__ void integrate_pixel(const apx_component& cp, const Prof<double>& p, const Prof<float>& pf,
                                int nx, int ci, int cj, int32_t pix, double thr1, int lane,
                                double* out) {
  const int n = cp.quad_level, nn = n * n, g = cp.gridding, gg = g * g;
  const double ig = 1.0 / (double)g;
  const double sub0 = -(double)(g - 1) / (2.0 * g), subd = 1.0 / (double)g;  // linspace step
  const int i = pix % nx, j = pix / nx;
  double X0, Y0;
  pixel_xy(p, ci, cj, i, j, X0, Y0);
  double acc[1 + APX_NSLOT];
#pragma unroll
  for (int s = 0; s <= APX_NSLOT; ++s) acc[s] = 0.0;
  double gl[APX_NSLOT];
  // ---- level 1 (area 1, float64): lane < nn is a node; node offsets are P / 2 * abscissa
  double v1 = 0.0;
  if (lane < nn) {
    const double u = cp.gl_node[lane % n], v = cp.gl_node[lane / n];
    const double w = cp.gl_weight[lane % n] * cp.gl_weight[lane / n] * 0.25;
    const double x = X0 + 0.5 * (p.P00 * u + p.P01 * v), y = Y0 + 0.5 * (p.P10 * u + p.P11 * v);
    v1 = eval_profile<double, GRAD>(p, x, y, 1.0, gl);
    acc[0] = w * v1;
    if (GRAD) {
#pragma unroll
      for (int s = 0; s < APX_NSLOT; ++s) acc[1 + s] = w * gl[s];
    }
  }
  const double res1 = wave_sum(acc[0]);
  const double ref1 = __shfl(v1, nn / 2, 64);
  const bool refine = cp.max_depth > 1 && fabs(res1 - ref1) > thr1;
  if (!refine) {
    out[0] = res1;
    if (GRAD) {
#pragma unroll
      for (int s = 0; s < APX_NSLOT; ++s) out[1 + s] = wave_sum(acc[1 + s]);
    }
    return;
  }
  // ---- level 2 (pixelscale P/g, area 1/g^2): lane < gg is a sub-pixel
  const double sc2 = ig * ig, thr2 = thr1 * (double)gg;
  double c2[1 + APX_NSLOT];
#pragma unroll
  for (int s = 0; s <= APX_NSLOT; ++s) c2[s] = 0.0;
  bool deeper = false;
  double x2 = 0.0, y2 = 0.0;
  if (lane < gg) {
    const double a = sub0 + subd * (lane % g), b = sub0 + subd * (lane / g);
    x2 = X0 + p.P00 * a + p.P01 * b;
    y2 = Y0 + p.P10 * a + p.P11 * b;
    float gs[APX_NSLOT], centre = 0.0f;
    const float h = (float)(0.5 * ig);
    const float r2 = quad_sample<float, GRAD>(pf, cp, (float)x2, (float)y2, n, cp.gl_node, cp.gl_weight,
                                              (float)sc2, h * pf.P00, h * pf.P01, h * pf.P10, h * pf.P11,
                                              gs, centre);
    if (cp.max_depth > 2) {
      const double D = fabs((double)r2 - (double)centre);
      const double band = kSampleBand * (fabs((double)r2) + fabs((double)centre)) + 1e-7 * thr2;
      if (D > thr2 + band) {
        deeper = true;
      } else if (D >= thr2 - band) {
        double cd = 0.0;
        const double rd = quad_sample<double, false>(p, cp, x2, y2, n, cp.gl_node, cp.gl_weight, sc2,
                                                     0.5 * ig * p.P00, 0.5 * ig * p.P01,
                                                     0.5 * ig * p.P10, 0.5 * ig * p.P11, nullptr, cd);
        deeper = fabs(rd - cd) > thr2;
      }
    }
    if (!deeper) {
      c2[0] = (double)r2;
      if (GRAD) {
#pragma unroll
        for (int s = 0; s < APX_NSLOT; ++s) c2[1 + s] = (double)gs[s];
      }
    }
  }
#pragma unroll
  for (int s = 0; s <= APX_NSLOT; ++s)
    if (GRAD || s == 0) out[s] = wave_sum(c2[s]);
  // ---- level 3 (pixelscale P/g^2, area 1/g^4): whole wave per flagged sub-pixel
  uint64_t todo = __ballot(deeper);
  const double sc3 = sc2 * ig * ig;
  while (todo) {
    const int sp = __ffsll((unsigned long long)todo) - 1;
    todo &= todo - 1;
    const double xs = __shfl(x2, sp, 64), ys = __shfl(y2, sp, 64);
    double c3[1 + APX_NSLOT];
#pragma unroll
    for (int s = 0; s <= APX_NSLOT; ++s) c3[s] = 0.0;
    for (int e = lane; e < gg * nn; e += 64) {
      const int t = e / nn, kk = e % nn;
      const double a = (sub0 + subd * (t % g)) * ig, b = (sub0 + subd * (t / g)) * ig;
      const double u = cp.gl_node[kk % n], v = cp.gl_node[kk / n];
      const double w = cp.gl_weight[kk % n] * cp.gl_weight[kk / n] * 0.25;
      const double hu = 0.5 * ig * ig * u, hv = 0.5 * ig * ig * v;
      const double x = xs + p.P00 * (a + hu) + p.P01 * (b + hv);
      const double y = ys + p.P10 * (a + hu) + p.P11 * (b + hv);
      float glf[APX_NSLOT];
      const float val = eval_profile<float, GRAD>(pf, (float)x, (float)y, (float)sc3, glf);
      c3[0] += w * (double)val;
      if (GRAD) {
#pragma unroll
        for (int s = 0; s < APX_NSLOT; ++s) c3[1 + s] += w * (double)glf[s];
      }
    }
#pragma unroll
    for (int s = 0; s <= APX_NSLOT; ++s)
      if (GRAD || s == 0) out[s] += wave_sum(c3[s]);
  }
}


}  // namespace apx
