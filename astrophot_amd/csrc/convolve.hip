// PSF sub-pixel shift and direct convolution on gfx950.
//
// psf_shift restates models/_model_methods.py:187-207 and :243 (bilinear
// _shift_psf with utils/interpolate.py:282-329's clamped corners, which
// extrapolate at the edge, then division by the sum) and adds the derivative
// of the normalised kernel with respect to the model centre, the only path by
// which the reference's forward-mode Jacobian sees the centre when
// psf_mode="full" (the sub-pixel shift s = c - round(c) moves the PSF, the
// sampling grid itself is centred on round(c); model_object.py:320-325).
//
// convolve restates models/_model_methods.py:233-257 and operations.py:9-36:
// the FFT path there is a circular convolution of the border-padded working
// image, and on the pixels kept after the crop (model_object.py:348-349) it is
// exactly the linear convolution computed here directly.  The working-image
// tile plus halo is staged in LDS once per render plane; kernel taps are
// wave-uniform and come through the scalar cache; each thread keeps RY
// vertically adjacent outputs in registers so every LDS read feeds up to RY
// (value) or 3*RY (value + two centre derivatives) FMAs.  The crop, the
// reduce(1) and the window-aware `image +=` (image_object.py:450-457) are the
// epilogue.
#include <hip/hip_runtime.h>

#include "capi_internal.h"
#include "conv_common.h"

namespace apx {

namespace {

constexpr int kTX = 64;   // output tile width = one wavefront
constexpr int kRY = 4;    // outputs per thread (vertical)
constexpr int kTY = 16;   // output tile height (4 waves x kRY)
constexpr int kKR = 32;   // kernel rows staged per pass
constexpr size_t kMaxLds = 160 * 1024;  // LDS a gfx950 workgroup may use

__device__ __forceinline__ double block_sum256(double v, double* red) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  return red[0] + red[1] + red[2] + red[3];
}

// sum over a 1024-thread workgroup (every thread gets it)
__device__ __forceinline__ double block_sum1024(double v, double* red) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double t = 0.0;
#pragma unroll
  for (int w = 0; w < 16; ++w) t += red[w];
  return t;
}

__global__ __launch_bounds__(1024) void psf_shift_kernel(const float* psf, int ph, int pw, double sx,
                                                         double sy, double p0, double p1, double p2,
                                                         double p3, double dvx, double dvy,
                                                         float* out) {
  __shared__ double red[16];
  const int H2 = ph + 2, W2 = pw + 2, n = H2 * W2;
  double S = 0.0, Sx = 0.0, Sy = 0.0;
  for (int k = threadIdx.x; k < n; k += 1024) {
    double v, dx, dy;
    bilinear(psf, ph, pw, k / W2, k % W2, sx, sy, v, dx, dy);
    S += v;
    Sx += dx;
    Sy += dy;
  }
  S = block_sum1024(S, red);
  Sx = block_sum1024(Sx, red);
  Sy = block_sum1024(Sy, red);
  const double iS = 1.0 / S;
  for (int k = threadIdx.x; k < n; k += 1024) {
    const int r = k / W2, c = k % W2;
    double v, dx, dy;
    bilinear(psf, ph, pw, r, c, sx, sy, v, dx, dy);
    const double K = v * iS;
    const double Kx = dx * iS - v * Sx * iS * iS;  // d(v/S)/ds_x
    const double Ky = dy * iS - v * Sy * iS * iS;
    const int f = (H2 - 1 - r) * W2 + (W2 - 1 - c);  // flipped for correlation
    out[f] = (float)K;
    out[n + f] = (float)((Kx * p0 + Ky * p2) * dvx);
    out[2 * n + f] = (float)((Kx * p1 + Ky * p3) * dvy);
  }
}

// Kernels of a PSF that is itself rendered from parameters theta (an
// auxiliary PSF_Model, model_object.py:305-310).  With T the PSF after the
// sub-pixel shift (the padded bilinear shift above, or the PSF itself when
// unshifted) and K = T / sum T (_model_methods.py:243),
//   dK/dtheta = dT / sum T - T (sum dT) / (sum T)^2 ,
// dT being the same shift applied to the derivative plane dP/dtheta.
// Block 0 writes K, block d + 1 writes dK of derivative plane d, and block
// nd + 1 writes a zero plane; every plane flipped for correlation.
__global__ __launch_bounds__(256) void psf_kernels_dparam_kernel(const float* psf, const float* dpsf,
                                                                 int nd, int ph, int pw, int shifted,
                                                                 double sx, double sy, float* out) {
  __shared__ double red[4];
  const int H = shifted ? ph + 2 : ph, W = shifted ? pw + 2 : pw, n = H * W;
  const int b = blockIdx.x;
  float* o = out + (int64_t)b * n;
  if (b == nd + 1) {
    for (int k = threadIdx.x; k < n; k += 256) o[k] = 0.0f;
    return;
  }
  const float* dp = b > 0 ? dpsf + (int64_t)(b - 1) * ph * pw : nullptr;
  auto sample = [&](const float* src, int r, int c) -> double {
    if (!shifted) return (double)src[r * pw + c];
    double v, dx, dy;
    bilinear(src, ph, pw, r, c, sx, sy, v, dx, dy);
    return v;
  };
  double S = 0.0, Sd = 0.0;
  for (int k = threadIdx.x; k < n; k += 256) {
    S += sample(psf, k / W, k % W);
    if (dp) Sd += sample(dp, k / W, k % W);
  }
  S = block_sum256(S, red);
  Sd = block_sum256(Sd, red);
  const double iS = 1.0 / S;
  for (int k = threadIdx.x; k < n; k += 256) {
    const int r = k / W, c = k % W;
    const double t = sample(psf, r, c);
    const double v = dp ? sample(dp, r, c) * iS - t * Sd * iS * iS : t * iS;
    o[(H - 1 - r) * W + (W - 1 - c)] = (float)v;
  }
}

// One workgroup = kTY x kTX outputs for every render plane.
template <bool DK>
__global__ __launch_bounds__(256) void convolve_kernel(apx_conv_job job, const float* __restrict__ planes,
                                                       const float* __restrict__ kernels) {
  extern __shared__ float tile[];
  const int lx = threadIdx.x & 63, ly = threadIdx.x >> 6;
  const int ox0 = blockIdx.x * kTX, oy0 = blockIdx.y * kTY;
  const int kh = job.kh, kw = job.kw, ch = (kh - 1) / 2, cw = (kw - 1) / 2;
  const int KR = min(kh, kKR);
  const int LW = kTX + kw - 1;                 // staged row length
  const int gx0 = job.x0 + ox0 - cw;           // working-grid column of tile[.][0]
  const int64_t npl = (int64_t)job.nx * job.ny;
  const int knum = kh * kw;
  const int ox = ox0 + lx;
  // plane 0 is needed for the value and the centre derivatives only
  const int c_first = (job.img || DK) ? 0 : 1;
  for (int c = c_first; c < job.nchan; ++c) {
    const bool centre = DK && c == 0;
    float acc[kRY], accx[kRY], accy[kRY];
#pragma unroll
    for (int r = 0; r < kRY; ++r) acc[r] = accx[r] = accy[r] = 0.0f;
    const float* plane = planes + c * npl;
    for (int kr0 = 0; kr0 < kh; kr0 += KR) {
      const int kr = min(KR, kh - kr0);
      const int rows = kTY + kr - 1;
      const int gy0 = job.y0 + oy0 - ch + kr0;   // working-grid row of tile[0][.]
      __syncthreads();
      for (int e = threadIdx.x; e < rows * LW; e += 256) {
        const int ty = e / LW, tx = e % LW;
        const int gy = gy0 + ty, gx = gx0 + tx;
        tile[e] = (gy >= 0 && gy < job.ny && gx >= 0 && gx < job.nx)
                      ? plane[(int64_t)gy * job.nx + gx]
                      : 0.0f;
      }
      __syncthreads();
      const float* K0 = kernels + kr0 * kw;
      for (int t = 0; t < kRY + kr - 1; ++t) {
        const float* row = tile + (ly * kRY + t) * LW + lx;
        for (int b = 0; b < kw; ++b) {
          const float v = row[b];
#pragma unroll
          for (int r = 0; r < kRY; ++r) {
            const int a = t - r;
            if (a >= 0 && a < kr) {
              acc[r] = fmaf(K0[a * kw + b], v, acc[r]);
              if (centre) {
                accx[r] = fmaf(K0[knum + a * kw + b], v, accx[r]);
                accy[r] = fmaf(K0[2 * knum + a * kw + b], v, accy[r]);
              }
            }
          }
        }
      }
    }
    if (ox < job.ow) {
#pragma unroll
      for (int r = 0; r < kRY; ++r) {
        const int oy = oy0 + ly * kRY + r;
        if (oy >= job.oh) continue;
        if (c == 0) {
          if (job.img) {
            float* d = job.img + (int64_t)(job.img_y0 + oy) * job.img_row + (job.img_x0 + ox);
            *d += acc[r];
          }
          if (centre && job.jac) {
            float* d = job.jac + (int64_t)(job.jac_y0 + oy) * job.jac_row +
                       (int64_t)(job.jac_x0 + ox) * job.jac_np;
            if (job.dk_col[0] >= 0) d[job.dk_col[0]] += accx[r];
            if (job.dk_col[1] >= 0) d[job.dk_col[1]] += accy[r];
          }
        } else if (job.jac) {
          float* d = job.jac + (int64_t)(job.jac_y0 + oy) * job.jac_row +
                     (int64_t)(job.jac_x0 + ox) * job.jac_np;
          d[job.chan_col[c]] += acc[r];
        }
      }
    }
  }
}


// ---------------------------------------------------------------------------
// Fast path (kernel side <= kFastMax): one workgroup = 64 columns x 8 rows of
// output, four wavefronts.  Each wavefront owns interleaved blocks of 8 input
// rows of the staged tile and all 8 output rows: for a tap column b it reads 8
// inputs from LDS and 16 taps through the scalar cache (one burst), then does
// 64 FMAs.  Taps come pre-arranged as KT[k][b][u] = K[u - 7][b] (zero outside
// the kernel) so every (input row, output row) pair finds its tap at
// u = t - r + 7 with no bounds tests.  The four partial sums are combined in
// LDS in a fixed order (deterministic).

__global__ __launch_bounds__(256) void conv_prep_kernel(const float* __restrict__ kernels, int nk,
                                                        int kh, int kw, int kts, float* kt) {
  const int n = nk * kw * kts;
  for (int e = blockIdx.x * 256 + threadIdx.x; e < n; e += gridDim.x * 256) {
    const int u = e % kts, b = (e / kts) % kw, k = e / (kts * kw);
    const int a = u - (kRYF - 1);
    kt[e] = (a >= 0 && a < kh) ? kernels[(k * kh + a) * kw + b] : 0.0f;
  }
}

template <int NK0>
__global__ __launch_bounds__(256) void convolve_fast_kernel(apx_conv_job job,
                                                            const float* __restrict__ planes,
                                                            const float* __restrict__ kt) {
  extern __shared__ float smem[];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int ox0 = blockIdx.x * 64, oy0 = blockIdx.y * kRYF;
  const int kh = job.kh, kw = job.kw, ch = (kh - 1) / 2, cw = (kw - 1) / 2;
  const int nT = fast_nt(kh), kts = fast_kts(kh);
  const int rows = 8 * nT, LW = 64 + kw - 1;
  const int gx0 = job.x0 + ox0 - cw, gy0 = job.y0 + oy0 - ch;
  const int64_t npl = (int64_t)job.nx * job.ny;
  const int c_first = (job.img || NK0 > 1) ? 0 : 1;
  for (int c = c_first; c < job.nchan; ++c) {
    const float* plane = planes + c * npl;
    __syncthreads();
    for (int e = threadIdx.x; e < rows * LW; e += 256) {
      const int ty = e / LW, tx = e % LW;
      const int gy = gy0 + ty, gx = gx0 + tx;
      smem[e] = (gy >= 0 && gy < job.ny && gx >= 0 && gx < job.nx) ? plane[(int64_t)gy * job.nx + gx]
                                                                  : 0.0f;
    }
    __syncthreads();
    constexpr int NKmax = NK0;
    const int nk = (c == 0) ? NK0 : 1;
    float acc[NKmax][kRYF];
#pragma unroll
    for (int k = 0; k < NKmax; ++k)
#pragma unroll
      for (int r = 0; r < kRYF; ++r) acc[k][r] = 0.0f;
    // wave index made provably uniform so tap addresses go to the scalar unit
    const int wu = __builtin_amdgcn_readfirstlane(w);
    conv_accumulate<NKmax>(smem, LW, lane, wu, nT, kw, kts, nk, kt, acc);
    // combine the four wavefronts' partial sums: red[w][k][r][lane]
    __syncthreads();
    float* red = smem;
#pragma unroll
    for (int k = 0; k < NKmax; ++k)
      if (k < nk)
#pragma unroll
        for (int r = 0; r < kRYF; ++r) red[((w * NKmax + k) * kRYF + r) * 64 + lane] = acc[k][r];
    __syncthreads();
    const int ox = ox0 + lane;
    for (int e = w; e < nk * kRYF; e += 4) {  // (kernel k, row r) per wavefront
      const int k = e / kRYF, r = e % kRYF;
      float s = 0.0f;
#pragma unroll
      for (int ww = 0; ww < 4; ++ww) s += red[((ww * NKmax + k) * kRYF + r) * 64 + lane];
      const int oy = oy0 + r;
      if (ox >= job.ow || oy >= job.oh) continue;
      if (c == 0 && k == 0) {
        if (job.img) job.img[(int64_t)(job.img_y0 + oy) * job.img_row + (job.img_x0 + ox)] += s;
      } else if (job.jac) {
        float* d = job.jac + (int64_t)(job.jac_y0 + oy) * job.jac_row +
                   (int64_t)(job.jac_x0 + ox) * job.jac_np;
        const int col = (c == 0) ? job.dk_col[k - 1] : job.chan_col[c];
        if (col >= 0) d[col] += s;
      }
    }
  }
}

// MFMA convolution of one job (kernels of kMfmaMin .. kMfmaWide rows): a
// 64 x kMfR output tile per workgroup, each wave one 16 x 16 block
// (mfma_block, conv_common.h).  The tap table of every kernel of the job is
// built in LDS once per workgroup; per render plane the input tile is staged
// and each of its kernels accumulated; outputs are added into the image or the
// Jacobian column as convolve_fast_kernel does.
constexpr int kMfmaMin = 21;  // shorter kernels waste over half of each MFMA on the band's zeros

__host__ inline size_t conv_mfma_lds_bytes(int kh, int kw, int nk) {
  return sizeof(float) * ((size_t)mf_rows(kh) * mf_lw(kw) + (size_t)nk * kw * mf_dd(kh));
}

template <int NK0>
__global__ __launch_bounds__(256) void convolve_mfma_kernel(apx_conv_job job, const float* __restrict__ planes,
                                                            const float* __restrict__ kernels) {
  extern __shared__ float smem[];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int lm = lane & 15, lk = lane >> 4;
  const int ox0 = blockIdx.x * 64, oy0 = blockIdx.y * kMfR;
  const int kh = job.kh, kw = job.kw, ch = (kh - 1) / 2, cw = (kw - 1) / 2;
  const int nks = mf_nks(kh), rows = mf_rows(kh), LW = mf_lw(kw), DD = mf_dd(kh);
  const int cols = 64 + kw - 1;
  const int gx0 = job.x0 + ox0 - cw, gy0 = job.y0 + oy0 - ch;
  const int64_t npl = (int64_t)job.nx * job.ny;
  float* tin = smem;             // rows x LW
  float* tab = smem + rows * LW; // [k][kx][dd] = Kf_k[dd - 15][kx]
  for (int e = threadIdx.x; e < NK0 * kw * DD; e += 256) {
    const int dd = e % DD, kx = (e / DD) % kw, k = e / (DD * kw);
    const int a = dd - 15;
    tab[e] = (a >= 0 && a < kh) ? kernels[((int64_t)k * kh + a) * kw + kx] : 0.0f;
  }
  // this lane's 4 outputs: x = 16 w + 4 lk + i, row lm
  const int oy = oy0 + lm;
  const int c_first = (job.img || NK0 > 1) ? 0 : 1;
  for (int c = c_first; c < job.nchan; ++c) {
    const float* plane = planes + c * npl;
    __syncthreads();
    for (int e = threadIdx.x; e < rows * cols; e += 256) {
      const int t = e / cols, u = e - t * cols;
      const int gy = gy0 + t, gx = gx0 + u;
      tin[t * LW + u] = (gy >= 0 && gy < job.ny && gx >= 0 && gx < job.nx) ? plane[(int64_t)gy * job.nx + gx]
                                                                          : 0.0f;
    }
    __syncthreads();
    const int nk = (c == 0) ? NK0 : 1;
    for (int k = 0; k < nk; ++k) {
      const float* ta = tin + lk * LW + 16 * w + lm;
      const float* tb = tab + k * kw * DD + lk - lm + 15;
      const v4f acc = mfma_block(nks, ta, tb, kw, LW, DD);
      if (oy >= job.oh) continue;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int ox = ox0 + 16 * w + 4 * lk + i;
        if (ox >= job.ow) continue;
        if (c == 0 && k == 0) {
          if (job.img) job.img[(int64_t)(job.img_y0 + oy) * job.img_row + (job.img_x0 + ox)] += acc[i];
        } else if (job.jac) {
          float* d = job.jac + (int64_t)(job.jac_y0 + oy) * job.jac_row + (int64_t)(job.jac_x0 + ox) * job.jac_np;
          const int col = (c == 0) ? job.dk_col[k - 1] : job.chan_col[c];
          if (col >= 0) d[col] += acc[i];
        }
      }
    }
  }
}

}  // namespace

int psf_shift(const float* psf, int32_t ph, int32_t pw, double sx, double sy, const double* pinv,
              double dvdr_x, double dvdr_y, float* kernels, hipStream_t stream) {
  if (!psf || !kernels || !pinv) return set_error(APX_E_INVALID, "psf_shift: null pointer");
  if (ph < 1 || pw < 1 || (ph % 2) == 0 || (pw % 2) == 0)
    return set_error(APX_E_INVALID, "psf_shift: psf must have odd shape");
  if (ph + 2 > APX_MAX_KERNEL || pw + 2 > APX_MAX_KERNEL)
    return set_error(APX_E_LIMIT, "psf_shift: psf larger than APX_MAX_KERNEL - 2");
  hipLaunchKernelGGL(psf_shift_kernel, dim3(1), dim3(1024), 0, stream, psf, ph, pw, sx, sy, pinv[0],
                     pinv[1], pinv[2], pinv[3], dvdr_x, dvdr_y, kernels);
  return check_launch("psf_shift");
}

int psf_kernels_dparam(const float* psf, const float* dpsf, int32_t nd, int32_t ph, int32_t pw,
                       int32_t shifted, double sx, double sy, float* out, hipStream_t stream) {
  if (!psf || !out || (nd > 0 && !dpsf))
    return set_error(APX_E_INVALID, "psf_kernels_dparam: null pointer");
  if (nd < 0 || nd > APX_NSLOT) return set_error(APX_E_INVALID, "psf_kernels_dparam: nd out of range");
  if (ph < 1 || pw < 1 || (ph % 2) == 0 || (pw % 2) == 0)
    return set_error(APX_E_INVALID, "psf_kernels_dparam: psf must have odd shape");
  if (ph + 2 > APX_MAX_KERNEL || pw + 2 > APX_MAX_KERNEL)
    return set_error(APX_E_LIMIT, "psf_kernels_dparam: psf larger than APX_MAX_KERNEL - 2");
  hipLaunchKernelGGL(psf_kernels_dparam_kernel, dim3((unsigned)(nd + 2)), dim3(256), 0, stream, psf,
                     dpsf, nd, ph, pw, shifted ? 1 : 0, sx, sy, out);
  return check_launch("psf_kernels_dparam");
}

// Block sums of an upscaled render back onto the image grid: one thread per
// (coarse pixel, channel), src read with a stride of nc floats.
__global__ __launch_bounds__(256) void block_sum_add_kernel(apx_block_sum job,
                                                            const float* __restrict__ src) {
  const int64_t n = (int64_t)job.w * job.h * job.nc;
  const int U = job.scale;
  for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < n; e += (int64_t)gridDim.x * 256) {
    const int c = (int)(e % job.nc);
    const int64_t p = e / job.nc;
    const int x = (int)(p % job.w), y = (int)(p / job.w);
    float s = 0.0f;
    for (int b = 0; b < U; ++b) {
      const float* row = src + ((int64_t)(U * y + b) * job.src_w + (int64_t)U * x) * job.nc + c;
      for (int a = 0; a < U; ++a) s += row[(int64_t)a * job.nc];
    }
    job.dst[(int64_t)(job.dst_y0 + y) * job.dst_row + (int64_t)(job.dst_x0 + x) * job.dst_pix +
            job.col[c]] += s;
  }
}

int block_sum_add(const apx_block_sum* jp, const float* src, hipStream_t stream) {
  if (!jp || !src || !jp->dst) return set_error(APX_E_INVALID, "block_sum_add: null pointer");
  const apx_block_sum job = *jp;
  if (job.scale < 1 || job.nc < 1 || job.nc > APX_MAX_BLOCK_COLS)
    return set_error(APX_E_INVALID, "block_sum_add: scale or nc out of range");
  if (job.w < 0 || job.h < 0) return set_error(APX_E_INVALID, "block_sum_add: negative box");
  if (job.w == 0 || job.h == 0) return APX_OK;
  if ((int64_t)job.scale * job.w > job.src_w || (int64_t)job.scale * job.h > job.src_h)
    return set_error(APX_E_INVALID, "block_sum_add: box outside the source");
  if (job.dst_pix < 1 || job.dst_row < (int64_t)job.dst_w * job.dst_pix || job.dst_x0 < 0 ||
      job.dst_y0 < 0 || job.dst_x0 + job.w > job.dst_w || job.dst_y0 + job.h > job.dst_h)
    return set_error(APX_E_INVALID, "block_sum_add: box outside the destination");
  for (int c = 0; c < job.nc; ++c)
    if (job.col[c] < 0 || job.col[c] >= job.dst_pix)
      return set_error(APX_E_INVALID, "block_sum_add: column out of range");
  const int64_t n = (int64_t)job.w * job.h * job.nc;
  const unsigned blocks = (unsigned)std::min<int64_t>((n + 255) / 256, 16384);
  hipLaunchKernelGGL(block_sum_add_kernel, dim3(blocks), dim3(256), 0, stream, job, src);
  return check_launch("block_sum_add");
}

int64_t convolve_workspace_bytes(int32_t kh, int32_t kw) {
  if (kh < 1 || kw < 1) return 0;
  return (int64_t)sizeof(float) * 3 * kw * fast_kts(kh) + 256;
}

int convolve(const apx_conv_job* jp, const float* planes, const float* kernels, void* workspace,
             int64_t ws_bytes, hipStream_t stream) {
  if (!jp || !planes || !kernels) return set_error(APX_E_INVALID, "convolve: null pointer");
  const apx_conv_job job = *jp;
  if (job.kh < 1 || job.kw < 1 || (job.kh % 2) == 0 || (job.kw % 2) == 0)
    return set_error(APX_E_INVALID, "convolve: kernel must have odd shape");
  if (job.kh > APX_MAX_KERNEL || job.kw > APX_MAX_KERNEL)
    return set_error(APX_E_LIMIT, "convolve: kernel larger than APX_MAX_KERNEL");
  if (job.nchan < 1 || job.nchan > APX_MAX_CHAN)
    return set_error(APX_E_INVALID, "convolve: nchan out of range");
  if (job.ow < 0 || job.oh < 0) return set_error(APX_E_INVALID, "convolve: negative box");
  if (job.ow == 0 || job.oh == 0) return APX_OK;
  // every output reads inside the working grid
  const int ch = (job.kh - 1) / 2, cw = (job.kw - 1) / 2;
  if (!job.zero_outside &&
      (job.x0 - cw < 0 || job.y0 - ch < 0 || job.x0 + job.ow - 1 + cw >= job.nx ||
       job.y0 + job.oh - 1 + ch >= job.ny))
    return set_error(APX_E_INVALID, "convolve: output box reads outside the working grid");
  if (job.x0 < 0 || job.y0 < 0 || job.x0 + job.ow > job.nx || job.y0 + job.oh > job.ny)
    return set_error(APX_E_INVALID, "convolve: output box outside the working grid");
  if (job.img) {
    if (job.img_row < job.img_w || job.img_x0 < 0 || job.img_y0 < 0 ||
        job.img_x0 + job.ow > job.img_w || job.img_y0 + job.oh > job.img_h)
      return set_error(APX_E_INVALID, "convolve: output box outside the image");
  }
  if (job.jac) {
    if (job.jac_np < 1 || job.jac_row < (int64_t)job.jac_w * job.jac_np || job.jac_x0 < 0 ||
        job.jac_y0 < 0 || job.jac_x0 + job.ow > job.jac_w || job.jac_y0 + job.oh > job.jac_h)
      return set_error(APX_E_INVALID, "convolve: output box outside the Jacobian");
  }
  const bool dk = job.dk_col[0] >= 0 || job.dk_col[1] >= 0;
  if (job.nchan > 1 || dk) {
    if (!job.jac || job.jac_np < 1) return set_error(APX_E_INVALID, "convolve: no Jacobian");
    for (int c = 1; c < job.nchan; ++c)
      if (job.chan_col[c] < 0 || job.chan_col[c] >= job.jac_np)
        return set_error(APX_E_INVALID, "convolve: channel column out of range");
    for (int k = 0; k < 2; ++k)
      if (job.dk_col[k] >= job.jac_np)
        return set_error(APX_E_INVALID, "convolve: centre column out of range");
  }
  const int nk0 = dk ? 3 : 1;
  if (job.kh >= kMfmaMin && job.kh <= kMfmaWide && job.kw >= kMfmaMin && job.kw <= kMfmaWide &&
      conv_mfma_lds_bytes(job.kh, job.kw, nk0) <= kMaxLds) {
    const size_t lds = conv_mfma_lds_bytes(job.kh, job.kw, nk0);
    const dim3 grid((unsigned)((job.ow + 63) / 64), (unsigned)((job.oh + kMfR - 1) / kMfR));
    if (dk) {
      hipFuncSetAttribute(reinterpret_cast<const void*>(&convolve_mfma_kernel<3>),
                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
      hipLaunchKernelGGL(convolve_mfma_kernel<3>, grid, dim3(256), lds, stream, job, planes, kernels);
    } else {
      hipFuncSetAttribute(reinterpret_cast<const void*>(&convolve_mfma_kernel<1>),
                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
      hipLaunchKernelGGL(convolve_mfma_kernel<1>, grid, dim3(256), lds, stream, job, planes, kernels);
    }
    return check_launch("convolve (MFMA)");
  }
  if (job.kh <= kFastMax && job.kw <= kFastMax) {
    if (!workspace || ws_bytes < convolve_workspace_bytes(job.kh, job.kw))
      return set_error(APX_E_INVALID, "convolve: workspace too small");
    float* kt = static_cast<float*>(workspace);
    const int kts = fast_kts(job.kh);
    const int n = nk0 * job.kw * kts;
    hipLaunchKernelGGL(conv_prep_kernel, dim3((unsigned)std::min(64, (n + 255) / 256)), dim3(256), 0,
                       stream, kernels, nk0, job.kh, job.kw, kts, kt);
    const size_t tile = (size_t)8 * fast_nt(job.kh) * (size_t)(64 + job.kw - 1);
    const size_t lds = sizeof(float) * std::max(tile, (size_t)4 * nk0 * kRYF * 64);
    const dim3 grid((unsigned)((job.ow + 63) / 64), (unsigned)((job.oh + kRYF - 1) / kRYF));
    if (dk)
      hipLaunchKernelGGL(convolve_fast_kernel<3>, grid, dim3(256), lds, stream, job, planes, kt);
    else
      hipLaunchKernelGGL(convolve_fast_kernel<1>, grid, dim3(256), lds, stream, job, planes, kt);
    return check_launch("convolve");
  }
  const int KR = std::min(job.kh, kKR);
  const size_t lds = sizeof(float) * (size_t)(kTY + KR - 1) * (size_t)(kTX + job.kw - 1);
  const dim3 grid((unsigned)((job.ow + kTX - 1) / kTX), (unsigned)((job.oh + kTY - 1) / kTY));
  if (dk)
    hipLaunchKernelGGL(convolve_kernel<true>, grid, dim3(256), lds, stream, job, planes, kernels);
  else
    hipLaunchKernelGGL(convolve_kernel<false>, grid, dim3(256), lds, stream, job, planes, kernels);
  return check_launch("convolve");
}

}  // namespace apx
