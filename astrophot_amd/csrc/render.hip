// Component render on gfx950: initial sampling, curvature-threshold selection
// and adaptive Gauss-Legendre supersampling of the selected pixels.
//
// Restates models/_model_methods.py:82-184 (_sample_init midpoint / quad:N,
// _sample_integrate threshold) and utils/operations.py:94-247
// (single_quad_integrate, grid_integrate) as three launches:
//   1. init     : one thread per working pixel, float32 profile value and
//                 Jacobian channels; a float64 per-block partial sum when the
//                 integrate reference is the mean sample.
//   2. select   : curvature (replicate-padded 5-point Laplacian) or |quad -
//                 centre node|, compared with tolerance * reference.  Pixels
//                 whose float32 test lies within a rounding band of the
//                 threshold are re-decided in float64, so the selected set is
//                 the float64 one.  Selected pixels are appended to a list with
//                 one atomic per wavefront.
//   3. integrate: one wavefront per listed pixel; the recursive quadrature
//                 (levels 1..max_depth, threshold x gridding^2 per level as in
//                 operations.py:241) runs in float64 and overwrites the
//                 pixel's value and Jacobian channels.
#include <hip/hip_runtime.h>

#include "capi_internal.h"
#include "profile.h"

namespace apx {

namespace {

constexpr int kInitBlock = 256;
constexpr int kIntegrateWaves = 2048;

// workspace layout (bytes)
struct RenderWs {
  int32_t* count;     // [0]: selected pixels
  double* reference;  // mean-sample reference (when mean_reference)
  double* partial;    // per-init-block sums
  int32_t* list;      // selected pixel indices
  float* centre;      // quad:N centre-node samples
};

__host__ __device__ inline int64_t align256(int64_t x) { return (x + 255) & ~int64_t(255); }

__host__ __device__ inline int64_t init_blocks(int64_t npix) {
  return (npix + kInitBlock - 1) / kInitBlock;
}

__host__ inline RenderWs carve(void* ws, int64_t npix) {
  char* b = static_cast<char*>(ws);
  RenderWs w;
  w.count = reinterpret_cast<int32_t*>(b);
  w.reference = reinterpret_cast<double*>(b + 64);
  int64_t off = 256;
  w.partial = reinterpret_cast<double*>(b + off);
  off += align256(8 * init_blocks(npix));
  w.list = reinterpret_cast<int32_t*>(b + off);
  off += align256(4 * npix);
  w.centre = reinterpret_cast<float*>(b + off);
  return w;
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// offsets of a working pixel from the centre
template <typename T>
__device__ __forceinline__ void pixel_xy(const Prof<T>& p, const apx_component& cp, int i, int j,
                                         T& X, T& Y) {
  const T di = (T)(i - cp.ci);
  const T dj = (T)(j - cp.cj);
  X = p.P00 * di + p.P01 * dj + p.fx;
  Y = p.P10 * di + p.P11 * dj + p.fy;
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

template <bool GRAD>
__global__ __launch_bounds__(kInitBlock) void render_init_kernel(apx_component cp, float* planes,
                                                                 RenderWs ws) {
  const int64_t npix = (int64_t)cp.nx * cp.ny;
  const int64_t pix = (int64_t)blockIdx.x * kInitBlock + threadIdx.x;
  const Prof<float> p = make_prof<float>(cp);
  float value = 0.0f;
  if (pix < npix) {
    const int i = (int)(pix % cp.nx), j = (int)(pix / cp.nx);
    float X, Y;
    pixel_xy(p, cp, i, j, X, Y);
    float g[APX_NSLOT];
    if (cp.sampling == APX_SAMPLE_QUAD) {
      float centre = 0.0f;
      value = quad_sample<float, GRAD>(p, cp, X, Y, cp.sample_quad, cp.sq_node, cp.sq_weight, 1.0f,
                                       0.5f * p.P00, 0.5f * p.P01, 0.5f * p.P10, 0.5f * p.P11, g,
                                       centre);
      ws.centre[pix] = centre;
    } else {
      value = eval_profile<float, GRAD>(p, X, Y, 1.0f, g);
    }
    planes[pix] = value;
    if (GRAD) {
      for (int c = 1; c < cp.nchan; ++c) {
        const int s = cp.chan_slot[c];
        planes[c * npix + pix] = g[s] * (float)cp.dvdr[s];
      }
    }
  }
  if (cp.mean_reference) {
    // deterministic float64 block sum of the initial samples
    __shared__ double red[kInitBlock / 64];
    double v = wave_sum((double)value);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
      double t = 0.0;
      for (int w = 0; w < kInitBlock / 64; ++w) t += red[w];
      ws.partial[blockIdx.x] = t;
    }
  }
}

__global__ __launch_bounds__(256) void mean_reference_kernel(RenderWs ws, int64_t nblocks,
                                                             int64_t npix) {
  __shared__ double red[4];
  double t = 0.0;
  for (int64_t b = threadIdx.x; b < nblocks; b += 256) t += ws.partial[b];
  t = wave_sum(t);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = t;
  __syncthreads();
  if (threadIdx.x == 0) {
    ws.reference[0] = (red[0] + red[1] + red[2] + red[3]) / (double)npix;
    ws.count[0] = 0;
  }
}

// float64 re-decision of one pixel's selection test
__device__ bool select_f64(const apx_component& cp, const Prof<double>& p, int i, int j,
                           double thr) {
  if (cp.sampling == APX_SAMPLE_QUAD) {
    double X, Y, centre = 0.0;
    pixel_xy(p, cp, i, j, X, Y);
    const double res = quad_sample<double, false>(p, cp, X, Y, cp.sample_quad, cp.sq_node,
                                                  cp.sq_weight, 1.0, 0.5 * p.P00, 0.5 * p.P01,
                                                  0.5 * p.P10, 0.5 * p.P11, nullptr, centre);
    return fabs(res - centre) > thr;
  }
  const int ii = min(max(i, 1), cp.nx - 2), jj = min(max(j, 1), cp.ny - 2);
  double X, Y;
  auto at = [&](int a, int b) {
    pixel_xy(p, cp, a, b, X, Y);
    return eval_profile<double, false>(p, X, Y, 1.0, nullptr);
  };
  const double mc = at(ii, jj);
  const double curv = at(ii - 1, jj) + at(ii + 1, jj) + at(ii, jj - 1) + at(ii, jj + 1) - 4.0 * mc;
  const double mid = (i == ii && j == jj) ? mc : at(i, j);
  return fabs(mid - (mid + curv)) > thr;  // |deep - reference| (_model_methods.py:164-165)
}

__global__ __launch_bounds__(256) void render_select_kernel(apx_component cp, const float* planes,
                                                            RenderWs ws) {
  const int64_t npix = (int64_t)cp.nx * cp.ny;
  const int64_t pix = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const double thr = cp.tolerance * (cp.mean_reference ? ws.reference[0] : cp.reference);
  bool sel = false;
  if (pix < npix) {
    const int i = (int)(pix % cp.nx), j = (int)(pix / cp.nx);
    float metric, scale;
    if (cp.sampling == APX_SAMPLE_QUAD) {
      const float r = planes[pix], c = ws.centre[pix];
      metric = fabsf(r - c);
      scale = fabsf(r) + fabsf(c);
    } else {
      const int ii = min(max(i, 1), cp.nx - 2), jj = min(max(j, 1), cp.ny - 2);
      const float* v = planes + (int64_t)jj * cp.nx + ii;
      const float l = v[-1], r = v[1], u = v[-cp.nx], d = v[cp.nx], m = v[0];
      metric = fabsf(l + r + u + d - 4.0f * m);
      scale = fabsf(l) + fabsf(r) + fabsf(u) + fabsf(d) + 4.0f * fabsf(m);
    }
    // float32 profile values carry at most ~256 ulp of error; inside that band
    // the float64 test decides.
    const double band = 1.6e-5 * (double)scale + 2.0e-7 * thr;
    const double dm = (double)metric;
    if (dm > thr + band) {
      sel = true;
    } else if (dm >= thr - band) {
      const Prof<double> p = make_prof<double>(cp);
      sel = select_f64(cp, p, i, j, thr);
    }
  }
  const uint64_t ballot = __ballot(sel);
  if (ballot) {
    const int lane = threadIdx.x & 63;
    int base = 0;
    if (lane == 0) base = atomicAdd(ws.count, (int)__popcll(ballot));
    base = __shfl(base, 0, 64);
    if (sel) {
      const uint64_t below = ballot & ((lane == 0) ? 0ull : (~0ull >> (64 - lane)));
      ws.list[base + (int)__popcll(below)] = (int32_t)pix;
    }
  }
}

// float64 adaptive quadrature of listed pixels, one wavefront per pixel
template <bool GRAD>
__global__ __launch_bounds__(256) void render_integrate_kernel(apx_component cp, float* planes,
                                                               RenderWs ws) {
  const int lane = threadIdx.x & 63;
  const int wave = (int)(blockIdx.x * 4 + (threadIdx.x >> 6));
  const int nwaves = (int)(gridDim.x * 4);
  const int count = ws.count[0];
  const int64_t npix = (int64_t)cp.nx * cp.ny;
  const Prof<double> p = make_prof<double>(cp);
  const double thr1 = cp.tolerance * (cp.mean_reference ? ws.reference[0] : cp.reference);
  const int n = cp.quad_level, nn = n * n, g = cp.gridding, gg = g * g;
  const double ig = 1.0 / (double)g;
  const double sub0 = -(double)(g - 1) / (2.0 * g), subd = 1.0 / (double)g;  // linspace step
  // node offsets scale with the level's pixelscale: P / g^(level-1) / 2
  for (int k = wave; k < count; k += nwaves) {
    const int32_t pix = ws.list[k];
    const int i = pix % cp.nx, j = pix / cp.nx;
    double X0, Y0;
    pixel_xy(p, cp, i, j, X0, Y0);
    double acc[1 + APX_NSLOT];
#pragma unroll
    for (int s = 0; s <= APX_NSLOT; ++s) acc[s] = 0.0;
    double gl[APX_NSLOT];
    // ---- level 1 (area 1): lane < nn is a node
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
    double out[1 + APX_NSLOT];
    if (cp.max_depth <= 1 || !(fabs(res1 - ref1) > thr1)) {
      out[0] = res1;
      if (GRAD) {
#pragma unroll
        for (int s = 0; s < APX_NSLOT; ++s) out[1 + s] = wave_sum(acc[1 + s]);
      }
    } else {
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
        double gs[APX_NSLOT], centre = 0.0;
        const double r2 = quad_sample<double, GRAD>(p, cp, x2, y2, n, cp.gl_node, cp.gl_weight,
                                                    sc2, 0.5 * ig * p.P00, 0.5 * ig * p.P01,
                                                    0.5 * ig * p.P10, 0.5 * ig * p.P11, gs, centre);
        deeper = cp.max_depth > 2 && fabs(r2 - centre) > thr2;
        if (!deeper) {
          c2[0] = r2;
          if (GRAD) {
#pragma unroll
            for (int s = 0; s < APX_NSLOT; ++s) c2[1 + s] = gs[s];
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
          const double val = eval_profile<double, GRAD>(p, x, y, sc3, gl);
          c3[0] += w * val;
          if (GRAD) {
#pragma unroll
            for (int s = 0; s < APX_NSLOT; ++s) c3[1 + s] += w * gl[s];
          }
        }
#pragma unroll
        for (int s = 0; s <= APX_NSLOT; ++s)
          if (GRAD || s == 0) out[s] += wave_sum(c3[s]);
      }
    }
    if (lane == 0) {
      planes[pix] = (float)out[0];
      if (GRAD) {
        for (int c = 1; c < cp.nchan; ++c) {
          const int s = cp.chan_slot[c];
          planes[c * npix + pix] = (float)(out[1 + s] * cp.dvdr[s]);
        }
      }
    }
  }
}

__global__ void zero_count_kernel(RenderWs ws) { ws.count[0] = 0; }

}  // namespace

int64_t render_workspace_bytes(int64_t npix) {
  return 256 + align256(8 * init_blocks(npix)) + align256(4 * npix) + align256(4 * npix);
}

int render_component(const apx_component* cp_in, float* planes, void* workspace, int64_t ws_bytes,
                     hipStream_t stream) {
  if (!cp_in || !planes || !workspace) return set_error(APX_E_INVALID, "render: null pointer");
  const apx_component cp = *cp_in;
  const int64_t npix = (int64_t)cp.nx * cp.ny;
  if (cp.nx < 3 || cp.ny < 3) return set_error(APX_E_INVALID, "render: working grid below 3x3");
  if (npix > (int64_t)INT32_MAX) return set_error(APX_E_LIMIT, "render: grid too large");
  if (ws_bytes < render_workspace_bytes(npix))
    return set_error(APX_E_INVALID, "render: workspace too small");
  if (cp.profile < APX_SERSIC || cp.profile > APX_FLAT_SKY)
    return set_error(APX_E_INVALID, "render: unknown profile");
  if (cp.nchan < 1 || cp.nchan > APX_MAX_CHAN)
    return set_error(APX_E_INVALID, "render: nchan out of range");
  for (int c = 1; c < cp.nchan; ++c)
    if (cp.chan_slot[c] < 0 || cp.chan_slot[c] >= APX_NSLOT)
      return set_error(APX_E_INVALID, "render: bad channel slot");
  if (cp.sampling == APX_SAMPLE_QUAD &&
      (cp.sample_quad < 1 || cp.sample_quad > APX_MAX_QUAD_LEVEL))
    return set_error(APX_E_LIMIT, "render: sampling quad level out of range");
  if (cp.integrate) {
    if (cp.quad_level < 1 || cp.quad_level > APX_MAX_QUAD_LEVEL)
      return set_error(APX_E_LIMIT, "render: integrate_quad_level out of range");
    if (cp.gridding < 1 || cp.gridding > APX_MAX_GRIDDING)
      return set_error(APX_E_LIMIT, "render: integrate_gridding out of range");
    if (cp.max_depth < 1 || cp.max_depth > APX_MAX_DEPTH)
      return set_error(APX_E_LIMIT, "render: integrate_max_depth out of range");
  }
  const RenderWs ws = carve(workspace, npix);
  const bool grad = cp.nchan > 1;
  const int64_t nb = init_blocks(npix);
  if (grad)
    hipLaunchKernelGGL(render_init_kernel<true>, dim3((unsigned)nb), dim3(kInitBlock), 0, stream,
                       cp, planes, ws);
  else
    hipLaunchKernelGGL(render_init_kernel<false>, dim3((unsigned)nb), dim3(kInitBlock), 0, stream,
                       cp, planes, ws);
  if (cp.mean_reference)
    hipLaunchKernelGGL(mean_reference_kernel, dim3(1), dim3(256), 0, stream, ws, nb, npix);
  else
    hipLaunchKernelGGL(zero_count_kernel, dim3(1), dim3(1), 0, stream, ws);
  if (cp.integrate) {
    hipLaunchKernelGGL(render_select_kernel, dim3((unsigned)((npix + 255) / 256)), dim3(256), 0,
                       stream, cp, planes, ws);
    const int blocks = (int)std::min<int64_t>((npix + 3) / 4, kIntegrateWaves / 4);
    if (grad)
      hipLaunchKernelGGL(render_integrate_kernel<true>, dim3(blocks), dim3(256), 0, stream, cp,
                         planes, ws);
    else
      hipLaunchKernelGGL(render_integrate_kernel<false>, dim3(blocks), dim3(256), 0, stream, cp,
                         planes, ws);
  }
  return check_launch("render");
}

int render_selected_count(const void* workspace, hipStream_t stream, int32_t* count) {
  if (!workspace || !count) return set_error(APX_E_INVALID, "selected_count: null pointer");
  if (hipMemcpyAsync(count, workspace, sizeof(int32_t), hipMemcpyDeviceToHost, stream) !=
      hipSuccess)
    return set_error(APX_E_HIP, "selected_count: copy failed");
  if (hipStreamSynchronize(stream) != hipSuccess)
    return set_error(APX_E_HIP, "selected_count: sync failed");
  return APX_OK;
}

}  // namespace apx
