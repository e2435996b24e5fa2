// Batched independent fits on gfx950: thousands of cutouts, one Sersic each,
// fitted by the reference's Levenberg-Marquardt entirely on the device.
//
// Every fit keeps its LM state in device memory (GalState) and advances in
// lock-step "ticks".  A tick is a fixed launch sequence over device-built work
// lists (so it can be captured in a HIP graph):
//   Jacobian stage (fits starting a step):   render value + 5 partial planes,
//       convolve with K, dK/dcx, dK/dcy -> model0 and J (pixel-major, 7 cols),
//       J^T W J, J^T W (Y - m0), chi2(m0)                 (fit/lm.py:256-260)
//   lm_begin: step bookkeeping                              (lm.py:261-267)
//   lm_solve_h: damped float64 solve, probe point x + 0.1 h (lm.py:269-277)
//   probe stage: value render at x + 0.1 h -> model1, then rpp
//   lm_solve_a: a = -A(L)^-1 rpp / 2, trial point x + h + 0 a (lm.py:281-290)
//   trial stage: value render at the trial point -> model1, chi2
//   lm_decide: the accept / reject / damping logic of one sub-step, and at the
//       end of a step the fit()-level update and convergence tests
//                                                 (lm.py:293-357, 461-485)
// A fit that finishes stops appearing in the work lists.  Renders use the same
// per-pixel device code as the single-component path (render_common.h,
// conv_common.h), so a batched fit evaluates exactly what the plain path does.
#include <hip/hip_runtime.h>

#include <cmath>

#include "capi_internal.h"
#include "conv_common.h"
#include "profile.h"
#include "render_common.h"

namespace apx {

namespace {

constexpr int NP = APX_BATCH_NP;
constexpr int kChunk = 256;
constexpr int kPersist = 2048;
constexpr int kNChan = 6;  // value + q, PA, n, Re, Ie partials

enum : int32_t { PH_INIT = 0, PH_NEEDJ = 1, PH_SUB = 2, PH_DONE = 3 };
enum : int32_t { ST_JAC = 0, ST_PROBE = 1, ST_TRIAL = 2, ST_ALL = 3 };
enum : int32_t { DIR_NONE = 0, DIR_BETTER = 1, DIR_WORSE = 2 };
enum : int32_t { S_RUN = 0, S_SUCCESS = 1, S_IMMOBILE = 2, S_NOSTEP = 3, S_MAXITER = 4 };

struct GalState {
  double x[8], ev[8];
  double L, init_chi2;
  double hist[10];
  double res_chi2, res_x[8];
  double best_dx[8], best_chi2, best_L;
  double scary_dx[8], scary_chi2, scary_L;
  double H[NP * NP], g[NP], h[8], a[8], ha[8], rpp[8];
  double chi2_m0, chi2, ndf;
  int32_t nhist, iter, scary_valid, nostep, direction, sub, phase, status;
};

struct GalEval {
  Prof<float> pf;
  Prof<double> pd;
  double thr;
  double dvdr[NP];
  int32_t nxw, nyw, ci, cj, bx, by;
};

struct Scratch {
  int32_t* cnt;       // [0] fits, [1] chunks, [2] tiles, [3] selected pixels
  int32_t* gal_list;  // ngal
  int2* chunk_list;   // nchunk
  int2* tile_list;    // ntile
  int2* sel_list;     // nwork
  GalEval* ev;        // ngal
};

__host__ __device__ inline int64_t align256(int64_t x) { return (x + 255) & ~int64_t(255); }

__host__ inline Scratch carve(const apx_batch& b) {
  char* p = static_cast<char*>(b.scratch);
  Scratch s;
  int64_t off = 0;
  s.cnt = reinterpret_cast<int32_t*>(p + off);
  off += 256;
  s.gal_list = reinterpret_cast<int32_t*>(p + off);
  off += align256(4 * (int64_t)b.ngal);
  s.chunk_list = reinterpret_cast<int2*>(p + off);
  off += align256(8 * b.nchunk);
  s.tile_list = reinterpret_cast<int2*>(p + off);
  off += align256(8 * b.ntile);
  s.sel_list = reinterpret_cast<int2*>(p + off);
  off += align256(8 * b.nwork);
  s.ev = reinterpret_cast<GalEval*>(p + off);
  return s;
}

__host__ int64_t scratch_bytes(const apx_batch& b) {
  return 256 + align256(4 * (int64_t)b.ngal) + align256(8 * b.nchunk) + align256(8 * b.ntile) +
         align256(8 * b.nwork) + align256((int64_t)sizeof(GalEval) * b.ngal);
}

__device__ __forceinline__ GalState* st_of(const apx_batch& b, int g) {
  return reinterpret_cast<GalState*>(b.state) + g;
}

__device__ __forceinline__ double wsum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// ---------------------------------------------------------------------------
// work lists
__global__ void reset_kernel(int32_t* cnt) {
  if (threadIdx.x < 8) cnt[threadIdx.x] = 0;
}

__global__ __launch_bounds__(256) void build_kernel(apx_batch b, Scratch s, int stage) {
  const int g = blockIdx.x * 256 + threadIdx.x;
  if (g >= b.ngal) return;
  const int ph = st_of(b, g)->phase;
  const bool active = stage == ST_ALL || (stage == ST_JAC ? (ph == PH_INIT || ph == PH_NEEDJ)
                                                         : ph == PH_SUB);
  if (!active) return;
  const apx_batch_galaxy G = b.gal[g];
  s.gal_list[atomicAdd(&s.cnt[0], 1)] = g;
  const int bx = (G.pw + 1) / 2, by = (G.ph + 1) / 2;
  const int nch = ((G.nx + 2 * bx) * (G.ny + 2 * by) + kChunk - 1) / kChunk;
  int base = atomicAdd(&s.cnt[1], nch);
  for (int c = 0; c < nch; ++c) s.chunk_list[base + c] = make_int2(g, c);
  const int ntl = ((G.nx + 63) / 64) * ((G.ny + kRYF - 1) / kRYF);
  base = atomicAdd(&s.cnt[2], ntl);
  for (int t = 0; t < ntl; ++t) s.tile_list[base + t] = make_int2(g, t);
}

// ---------------------------------------------------------------------------
// parameter maps (utils/conversions/optimization.py:17-54) for
// [cx, cy, q (0,1), PA cyclic (0,pi), n (0.36,8), Re (0,inf), Ie]
__device__ inline void rep_to_values(const double* r, double* v, double* dvdr) {
  v[0] = r[0];
  v[1] = r[1];
  v[2] = (atan(r[2]) + kPi / 2) * 1.0 / kPi;
  v[3] = r[3] - kPi * floor(r[3] / kPi);
  v[4] = (atan(r[4]) + kPi / 2) * (8.0 - 0.36) / kPi + 0.36;
  v[5] = (r[5] + sqrt(r[5] * r[5] + 4.0)) * 0.5;
  v[6] = r[6];
  dvdr[0] = dvdr[1] = dvdr[3] = dvdr[6] = 1.0;
  dvdr[2] = 1.0 / (kPi * (r[2] * r[2] + 1.0));
  dvdr[4] = (8.0 - 0.36) / (kPi * (r[4] * r[4] + 1.0));
  dvdr[5] = 0.5 + 0.5 * r[5] / sqrt(r[5] * r[5] + 4.0);
}

// per fit: render constants + shifted PSF taps in the conv layout
__global__ __launch_bounds__(256) void prepare_kernel(apx_batch b, Scratch s, int stage,
                                                      const double* rep_override) {
  __shared__ double red[4];
  const int nk = stage == ST_JAC ? 3 : 1;
  for (int k = blockIdx.x; k < s.cnt[0]; k += gridDim.x) {
    const int g = s.gal_list[k];
    const apx_batch_galaxy G = b.gal[g];
    const double* rep = rep_override ? rep_override + (int64_t)g * NP : st_of(b, g)->ev;
    double v[NP], dvdr[NP];
    rep_to_values(rep, v, dvdr);
    const int bx = (G.pw + 1) / 2, by = (G.ph + 1) / 2;  // psf_image.py:71-93
    const int nxw = G.nx + 2 * bx, nyw = G.ny + 2 * by;
    const double det = b.P[0] * b.P[3] - b.P[1] * b.P[2];
    const double Pi[4] = {b.P[3] / det, -b.P[1] / det, -b.P[2] / det, b.P[0] / det};
    const double dx = v[0] - G.origin[0], dy = v[1] - G.origin[1];
    const double pcx = Pi[0] * dx + Pi[1] * dy + (-0.5 + bx);
    const double pcy = Pi[2] * dx + Pi[3] * dy + (-0.5 + by);
    const double ci = rint(pcx), cj = rint(pcy);  // ties to even
    const double sx = pcx - ci, sy = pcy - cj;
    const double prm[3] = {v[4], v[5], v[6]};
    const double area = fabs(det);
    if (threadIdx.x == 0) {
      GalEval e;
      e.pf = make_prof<float>(APX_SERSIC, b.P, 0.0, 0.0, area, b.softening, v[2], v[3], prm);
      e.pd = make_prof<double>(APX_SERSIC, b.P, 0.0, 0.0, area, b.softening, v[2], v[3], prm);
      const double bn = sersic_b(v[4]);
      const double flux = 2.0 * kPi * pow(10.0, v[6]) * v[5] * v[5] * v[2] * v[4] * exp(bn) *
                          pow(bn, -2.0 * v[4]) * exp(lgamma(2.0 * v[4]));  // functions.py:168-190
      e.thr = b.tolerance * flux / ((double)nxw * nyw);
      for (int q = 0; q < NP; ++q) e.dvdr[q] = dvdr[q];
      e.nxw = nxw;
      e.nyw = nyw;
      e.ci = (int)ci;
      e.cj = (int)cj;
      e.bx = bx;
      e.by = by;
      s.ev[g] = e;
    }
    // shifted, normalised PSF (and its centre derivatives), flipped, in the
    // KT[k][b][u] = Kf[u - 7][b] layout of conv_common.h
    const int H2 = G.ph + 2, W2 = G.pw + 2, n = H2 * W2, kts = fast_kts(H2);
    const float* psf = b.psf + G.psf_off;
    float* kt = b.taps + G.tap_off;
    double S = 0.0, Sx = 0.0, Sy = 0.0;
    for (int e = threadIdx.x; e < n; e += 256) {
      double val, ddx, ddy;
      bilinear(psf, G.ph, G.pw, e / W2, e % W2, sx, sy, val, ddx, ddy);
      S += val;
      Sx += ddx;
      Sy += ddy;
    }
    auto bsum = [&](double x) {
      x = wsum(x);
      __syncthreads();
      if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = x;
      __syncthreads();
      return red[0] + red[1] + red[2] + red[3];
    };
    S = bsum(S);
    Sx = bsum(Sx);
    Sy = bsum(Sy);
    const double iS = 1.0 / S;
    for (int e = threadIdx.x; e < nk * W2 * kts; e += 256) {
      const int u = e % kts, col = (e / kts) % W2, kk = e / (kts * W2);
      const int a = u - (kRYF - 1);  // flipped-kernel row
      float out = 0.0f;
      if (a >= 0 && a < H2) {
        const int r = H2 - 1 - a, c = W2 - 1 - col;  // unflipped sample
        double val, ddx, ddy;
        bilinear(psf, G.ph, G.pw, r, c, sx, sy, val, ddx, ddy);
        const double K = val * iS;
        if (kk == 0) {
          out = (float)K;
        } else {
          const double Kx = ddx * iS - val * Sx * iS * iS, Ky = ddy * iS - val * Sy * iS * iS;
          out = kk == 1 ? (float)(Kx * Pi[0] + Ky * Pi[2]) : (float)(Kx * Pi[1] + Ky * Pi[3]);
        }
      }
      kt[e] = out;
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// render stages
template <bool GRAD>
__global__ __launch_bounds__(kChunk) void init_kernel(apx_batch b, Scratch s) {
  for (int k = blockIdx.x; k < s.cnt[1]; k += gridDim.x) {
    const int2 e = s.chunk_list[k];
    const apx_batch_galaxy G = b.gal[e.x];
    const GalEval E = s.ev[e.x];
    const int pix = e.y * kChunk + threadIdx.x;
    if (pix >= E.nxw * E.nyw) continue;
    float X, Y, gr[APX_NSLOT];
    pixel_xy(E.pf, E.ci, E.cj, pix % E.nxw, pix / E.nxw, X, Y);
    const float val = eval_profile<float, GRAD>(E.pf, X, Y, 1.0f, gr);
    float* pl = b.planes + G.work_off + pix;
    pl[0] = val;
    if (GRAD) {
      // planes 1..5: q, PA, n, Re, Ie in representation units
#pragma unroll
      for (int c = 1; c < kNChan; ++c) pl[c * b.nwork] = gr[c + 1] * (float)E.dvdr[c + 1];
    }
  }
}

__global__ __launch_bounds__(256) void select_kernel(apx_batch b, Scratch s) {
  for (int k = blockIdx.x; k < s.cnt[1]; k += gridDim.x) {
    const int2 e = s.chunk_list[k];
    const apx_batch_galaxy G = b.gal[e.x];
    const GalEval E = s.ev[e.x];
    const int pix = e.y * kChunk + threadIdx.x;
    bool sel = false;
    if (pix < E.nxw * E.nyw) {
      const int i = pix % E.nxw, j = pix / E.nxw;
      const int ii = min(max(i, 1), E.nxw - 2), jj = min(max(j, 1), E.nyw - 2);
      const float* v = b.planes + G.work_off + (int64_t)jj * E.nxw + ii;
      const float l = v[-1], r = v[1], u = v[-E.nxw], d = v[E.nxw], m = v[0];
      const float metric = fabsf(l + r + u + d - 4.0f * m);
      const float scale = fabsf(l) + fabsf(r) + fabsf(u) + fabsf(d) + 4.0f * fabsf(m);
      const double band = 1.6e-5 * (double)scale + 2.0e-7 * E.thr;
      if ((double)metric > E.thr + band) {
        sel = true;
      } else if ((double)metric >= E.thr - band) {  // float64 re-decision
        double X, Y;
        auto at = [&](int a, int c) {
          pixel_xy(E.pd, E.ci, E.cj, a, c, X, Y);
          return eval_profile<double, false>(E.pd, X, Y, 1.0, nullptr);
        };
        const double mc = at(ii, jj);
        const double curv = at(ii - 1, jj) + at(ii + 1, jj) + at(ii, jj - 1) + at(ii, jj + 1) - 4.0 * mc;
        const double mid = (i == ii && j == jj) ? mc : at(i, j);
        sel = fabs(mid - (mid + curv)) > E.thr;
      }
    }
    const uint64_t ballot = __ballot(sel);
    if (ballot) {
      const int lane = threadIdx.x & 63;
      int base = 0;
      if (lane == 0) base = atomicAdd(&s.cnt[3], (int)__popcll(ballot));
      base = __shfl(base, 0, 64);
      if (sel) {
        const uint64_t below = ballot & ((lane == 0) ? 0ull : (~0ull >> (64 - lane)));
        s.sel_list[base + (int)__popcll(below)] = make_int2(e.x, pix);
      }
    }
  }
}

template <bool GRAD>
__global__ __launch_bounds__(256) void integrate_kernel(apx_batch b, Scratch s, apx_component cfg) {
  const int lane = threadIdx.x & 63;
  const int wave = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int nwaves = gridDim.x * 4;
  for (int k = wave; k < s.cnt[3]; k += nwaves) {
    const int2 e = s.sel_list[k];
    const apx_batch_galaxy G = b.gal[e.x];
    const GalEval E = s.ev[e.x];
    double out[1 + APX_NSLOT];
    integrate_pixel<GRAD>(cfg, E.pd, E.nxw, E.ci, E.cj, e.y, E.thr, lane, out);
    if (lane == 0) {
      float* pl = b.planes + G.work_off + e.y;
      pl[0] = (float)out[0];
      if (GRAD) {
        for (int c = 1; c < kNChan; ++c) pl[c * b.nwork] = (float)(out[1 + c + 1] * E.dvdr[c + 1]);
      }
    }
  }
}

// convolution of every listed tile; JAC: value -> out, centre derivatives ->
// J cols 0, 1, partial planes -> J cols 2..6
template <bool JAC>
__global__ __launch_bounds__(256) void conv_kernel(apx_batch b, Scratch s, float* out) {
  extern __shared__ float smem[];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int wu = __builtin_amdgcn_readfirstlane(w);
  constexpr int NKmax = JAC ? 3 : 1;
  constexpr int nchan = JAC ? kNChan : 1;
  for (int k = blockIdx.x; k < s.cnt[2]; k += gridDim.x) {
    const int2 e = s.tile_list[k];
    const int g = __builtin_amdgcn_readfirstlane(e.x);
    const apx_batch_galaxy G = b.gal[g];
    const int nxw = G.nx + G.pw + 1, nyw = G.ny + G.ph + 1;  // cutout + 2 * border
    const int ntx = (G.nx + 63) / 64;
    const int ox0 = (e.y % ntx) * 64, oy0 = (e.y / ntx) * kRYF;
    const int kh = G.ph + 2, kw = G.pw + 2;
    const int nT = fast_nt(kh), kts = fast_kts(kh), rows = 8 * nT, LW = 64 + kw - 1;
    // output (ox, oy) sits at working (bx + ox, by + oy); taps span +-(k-1)/2 = border
    const int gx0 = ox0, gy0 = oy0;
    const float* kt = b.taps + G.tap_off;
    for (int c = 0; c < nchan; ++c) {
      const float* plane = b.planes + (int64_t)c * b.nwork + G.work_off;
      __syncthreads();
      for (int q = threadIdx.x; q < rows * LW; q += 256) {
        const int gy = gy0 + q / LW, gx = gx0 + q % LW;
        smem[q] = (gy < nyw && gx < nxw) ? plane[(int64_t)gy * nxw + gx] : 0.0f;
      }
      __syncthreads();
      const int nk = (JAC && c == 0) ? 3 : 1;
      float acc[NKmax][kRYF];
#pragma unroll
      for (int kk = 0; kk < NKmax; ++kk)
#pragma unroll
        for (int r = 0; r < kRYF; ++r) acc[kk][r] = 0.0f;
      conv_accumulate<NKmax>(smem, LW, lane, wu, nT, kw, kts, nk, kt, acc);
      __syncthreads();
      float* red = smem;
#pragma unroll
      for (int kk = 0; kk < NKmax; ++kk)
        if (kk < nk)
#pragma unroll
          for (int r = 0; r < kRYF; ++r) red[((w * NKmax + kk) * kRYF + r) * 64 + lane] = acc[kk][r];
      __syncthreads();
      const int ox = ox0 + lane;
      for (int q = w; q < nk * kRYF; q += 4) {
        const int kk = q / kRYF, r = q % kRYF;
        float sum = 0.0f;
#pragma unroll
        for (int ww = 0; ww < 4; ++ww) sum += red[((ww * NKmax + kk) * kRYF + r) * 64 + lane];
        const int oy = oy0 + r;
        if (ox >= G.nx || oy >= G.ny) continue;
        const int64_t p = G.pix_off + (int64_t)oy * G.nx + ox;
        if (c == 0 && kk == 0)
          out[p] = sum;
        else if (JAC)
          b.jac[p * NP + (c == 0 ? kk - 1 : c + 1)] = sum;
      }
    }
  }
}

// ---------------------------------------------------------------------------
// per-fit reductions (one workgroup per listed fit)
typedef float v4f __attribute__((ext_vector_type(4)));

__global__ __launch_bounds__(256) void normal_kernel(apx_batch b, Scratch s) {
  __shared__ double redH[4][256];
  __shared__ double redg[4][16];
  __shared__ double redc[4];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int r = lane & 15, kq = lane >> 4;
  for (int k = blockIdx.x; k < s.cnt[0]; k += gridDim.x) {
    const int g = s.gal_list[k];
    const apx_batch_galaxy G = b.gal[g];
    const int64_t n = (int64_t)G.nx * G.ny;
    v4f acc = {0.0f, 0.0f, 0.0f, 0.0f};
    double accd[4] = {0.0, 0.0, 0.0, 0.0};
    double gacc = 0.0, cacc = 0.0;
    float gpart = 0.0f;
    int since = 0;
    for (int64_t base = wv * 4; base < n; base += 16) {
      const int64_t p = base + kq;
      float a = 0.0f, wt = 0.0f, res = 0.0f;
      if (p < n) {
        const int64_t q = G.pix_off + p;
        wt = b.weight[q];
        res = b.data[q] - b.model0[q];
        if (r < NP) a = b.jac[q * NP + r];
        if (r == 0) {
          const double d = (double)b.data[q] - (double)b.model0[q];
          cacc += (double)wt * d * d;
        }
      }
      acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a, wt * a, acc, 0, 0, 0);
      gpart = fmaf(a * wt, res, gpart);
      if (++since == 64) {
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          accd[i] += (double)acc[i];
          acc[i] = 0.0f;
        }
        gacc += (double)gpart;
        gpart = 0.0f;
        since = 0;
      }
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) accd[i] += (double)acc[i];
    gacc += (double)gpart;
#pragma unroll
    for (int i = 0; i < 4; ++i) redH[wv][(4 * kq + i) * 16 + r] = accd[i];
    gacc += __shfl_xor(gacc, 16, 64);
    gacc += __shfl_xor(gacc, 32, 64);
    if (lane < 16) redg[wv][lane] = gacc;
    cacc = wsum(cacc);
    if (lane == 0) redc[wv] = cacc;
    __syncthreads();
    GalState* st = st_of(b, g);
    if (threadIdx.x < NP * NP) {
      const int i = threadIdx.x / NP, j = threadIdx.x % NP;
      const int e = i * 16 + j;
      st->H[threadIdx.x] = redH[0][e] + redH[1][e] + redH[2][e] + redH[3][e];
    } else if (threadIdx.x < NP * NP + NP) {
      const int i = threadIdx.x - NP * NP;
      st->g[i] = redg[0][i] + redg[1][i] + redg[2][i] + redg[3][i];
    } else if (threadIdx.x == NP * NP + NP) {
      st->chi2_m0 = redc[0] + redc[1] + redc[2] + redc[3];
    }
    __syncthreads();
  }
}

__global__ __launch_bounds__(256) void rpp_kernel(apx_batch b, Scratch s) {
  __shared__ double red[4][8];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  for (int k = blockIdx.x; k < s.cnt[0]; k += gridDim.x) {
    const int g = s.gal_list[k];
    const apx_batch_galaxy G = b.gal[g];
    GalState* st = st_of(b, g);
    double h[NP], acc[NP];
#pragma unroll
    for (int i = 0; i < NP; ++i) {
      h[i] = st->h[i];
      acc[i] = 0.0;
    }
    const int64_t n = (int64_t)G.nx * G.ny;
    const double d = 0.1;
    for (int64_t p = threadIdx.x; p < n; p += 256) {
      const int64_t q = G.pix_off + p;
      float jr[NP];
      double jh = 0.0;
#pragma unroll
      for (int i = 0; i < NP; ++i) {
        jr[i] = b.jac[q * NP + i];
        jh += (double)jr[i] * h[i];
      }
      const double wt = (double)b.weight[q];
      const double v = (2.0 / d) * (wt * ((double)b.model1[q] - (double)b.model0[q]) / d - wt * jh);
#pragma unroll
      for (int i = 0; i < NP; ++i) acc[i] += (double)jr[i] * v;
    }
#pragma unroll
    for (int i = 0; i < NP; ++i) {
      const double t = wsum(acc[i]);
      if (lane == 0) red[wv][i] = t;
    }
    __syncthreads();
    if (threadIdx.x < NP)
      st->rpp[threadIdx.x] = red[0][threadIdx.x] + red[1][threadIdx.x] + red[2][threadIdx.x] +
                             red[3][threadIdx.x];
    __syncthreads();
  }
}

__global__ __launch_bounds__(256) void chi2_kernel(apx_batch b, Scratch s) {
  __shared__ double red[4];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  for (int k = blockIdx.x; k < s.cnt[0]; k += gridDim.x) {
    const int g = s.gal_list[k];
    const apx_batch_galaxy G = b.gal[g];
    const int64_t n = (int64_t)G.nx * G.ny;
    double acc = 0.0;
    for (int64_t p = threadIdx.x; p < n; p += 256) {
      const int64_t q = G.pix_off + p;
      const double dd = (double)b.data[q] - (double)b.model1[q];
      acc += (double)b.weight[q] * dd * dd;
    }
    acc = wsum(acc);
    if (lane == 0) red[wv] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
      GalState* st = st_of(b, g);
      st->chi2 = (red[0] + red[1] + red[2] + red[3]) / st->ndf;
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// LM control, one thread per fit (fit/lm.py)
__device__ inline void Lup(GalState* st) { st->L = fmin(1e9, st->L * 11.0); }
__device__ inline void Ldn(GalState* st) { st->L = fmax(1e-9, st->L / 9.0); }

// h = A(L)^-1 rhs, A_ii = H_ii + L (1 + H_ii), A_ij = H_ij / (1 + L) (lm.py:359-371):
// float64 Cholesky, Gaussian elimination with partial pivoting if a pivot fails
__device__ void damped_solve(const double* H, const double* rhs, double L, double* out) {
  double A[NP][NP], y[NP];
  for (int i = 0; i < NP; ++i) {
    for (int j = 0; j < NP; ++j) A[i][j] = (i == j) ? H[i * NP + j] + L * (1.0 + H[i * NP + j]) : H[i * NP + j] / (1.0 + L);
    y[i] = rhs[i];
  }
  bool ok = true;
  double C[NP][NP];
  for (int j = 0; j < NP && ok; ++j) {
    double d = A[j][j];
    for (int k = 0; k < j; ++k) d -= C[j][k] * C[j][k];
    if (!(d > 0.0) || !isfinite(d)) {
      ok = false;
      break;
    }
    C[j][j] = sqrt(d);
    for (int i = j + 1; i < NP; ++i) {
      double t = A[i][j];
      for (int k = 0; k < j; ++k) t -= C[i][k] * C[j][k];
      C[i][j] = t / C[j][j];
    }
  }
  if (ok) {
    for (int i = 0; i < NP; ++i) {
      double t = y[i];
      for (int k = 0; k < i; ++k) t -= C[i][k] * y[k];
      y[i] = t / C[i][i];
    }
    for (int i = NP - 1; i >= 0; --i) {
      double t = y[i];
      for (int k = i + 1; k < NP; ++k) t -= C[k][i] * y[k];
      y[i] = t / C[i][i];
    }
  } else {
    for (int k = 0; k < NP; ++k) {
      int pr = k;
      for (int i = k + 1; i < NP; ++i)
        if (fabs(A[i][k]) > fabs(A[pr][k])) pr = i;
      if (pr != k) {
        for (int j = 0; j < NP; ++j) {
          const double t = A[k][j];
          A[k][j] = A[pr][j];
          A[pr][j] = t;
        }
        const double t = y[k];
        y[k] = y[pr];
        y[pr] = t;
      }
      for (int i = k + 1; i < NP; ++i) {
        const double f = A[i][k] / A[k][k];
        for (int j = k; j < NP; ++j) A[i][j] -= f * A[k][j];
        y[i] -= f * y[k];
      }
    }
    for (int i = NP - 1; i >= 0; --i) {
      double t = y[i];
      for (int j = i + 1; j < NP; ++j) t -= A[i][j] * y[j];
      y[i] = t / A[i][i];
    }
  }
  for (int i = 0; i < NP; ++i) out[i] = y[i];
}

__device__ inline double hist_back(const GalState* st, int k) {  // loss_history[-k]
  return st->hist[(st->nhist - k) % 10];
}

__device__ inline void push_hist(GalState* st, double c) {
  st->hist[st->nhist % 10] = c;
  st->nhist += 1;
  if (isfinite(c) && (!isfinite(st->res_chi2) || c < st->res_chi2)) {  // base.py:104-117 argmin
    st->res_chi2 = c;
    for (int i = 0; i < NP; ++i) st->res_x[i] = st->x[i];
  }
}

__global__ void start_kernel(apx_batch b, const double* rep0) {
  const int g = blockIdx.x * 256 + threadIdx.x;
  if (g >= b.ngal) return;
  GalState* st = st_of(b, g);
  const apx_batch_galaxy G = b.gal[g];
  for (int i = 0; i < 8; ++i) st->x[i] = st->ev[i] = i < NP ? rep0[(int64_t)g * NP + i] : 0.0;
  st->L = b.L0;
  st->nhist = 0;
  st->iter = 0;
  st->res_chi2 = INFINITY;
  for (int i = 0; i < NP; ++i) st->res_x[i] = st->x[i];
  st->phase = PH_INIT;
  st->status = S_RUN;
  st->ndf = fmax(1.0, (double)G.nx * G.ny - NP);
}

__global__ void stage_eval_kernel(apx_batch b) {  // Jacobian stage evaluates at x
  const int g = blockIdx.x * 256 + threadIdx.x;
  if (g >= b.ngal) return;
  GalState* st = st_of(b, g);
  if (st->phase == PH_INIT || st->phase == PH_NEEDJ)
    for (int i = 0; i < NP; ++i) st->ev[i] = st->x[i];
}

__global__ void lm_begin_kernel(apx_batch b) {
  const int g = blockIdx.x * 256 + threadIdx.x;
  if (g >= b.ngal) return;
  GalState* st = st_of(b, g);
  if (st->phase != PH_INIT && st->phase != PH_NEEDJ) return;
  if (st->phase == PH_INIT) push_hist(st, st->chi2_m0 / st->ndf);  // lm.py:444-448
  st->init_chi2 = hist_back(st, 1);
  for (int i = 0; i < NP; ++i) st->best_dx[i] = 0.0;
  st->best_chi2 = st->init_chi2;
  st->best_L = st->L;
  st->scary_valid = 0;
  st->scary_chi2 = st->init_chi2;
  st->scary_L = st->L;
  st->direction = DIR_NONE;
  st->nostep = 1;
  st->sub = 0;
  st->phase = PH_SUB;
}

__global__ void lm_solve_h_kernel(apx_batch b) {
  const int g = blockIdx.x * 256 + threadIdx.x;
  if (g >= b.ngal) return;
  GalState* st = st_of(b, g);
  if (st->phase != PH_SUB) return;
  if (st->sub > 5 && st->L < 1e-3) st->L = 1.0;  // lm.py:270-271
  damped_solve(st->H, st->g, st->L, st->h);
  for (int i = 0; i < NP; ++i) st->ev[i] = st->x[i] + 0.1 * st->h[i];
}

__global__ void lm_solve_a_kernel(apx_batch b) {
  const int g = blockIdx.x * 256 + threadIdx.x;
  if (g >= b.ngal) return;
  GalState* st = st_of(b, g);
  if (st->phase != PH_SUB) return;
  if (st->L > 1e-4) {  // lm.py:283-286
    double t[NP];
    damped_solve(st->H, st->rpp, st->L, t);
    for (int i = 0; i < NP; ++i) st->a[i] = -t[i] / 2.0;
  } else {
    for (int i = 0; i < NP; ++i) st->a[i] = 0.0;
  }
  for (int i = 0; i < NP; ++i) {
    st->ha[i] = st->h[i] + st->a[i] * 0.0;  // acceleration = 0
    st->ev[i] = st->x[i] + st->ha[i];
  }
}

__device__ void finish_step(const apx_batch& b, GalState* st) {
  const double* dx;
  double chi2, L;
  if (st->nostep) {
    if (!st->scary_valid) {  // OptimizeStop (lm.py:348-355, 455-459)
      st->status = S_NOSTEP;
      st->phase = PH_DONE;
      return;
    }
    dx = st->scary_dx;
    chi2 = st->scary_chi2;
    L = st->scary_L;
  } else {
    dx = st->best_dx;
    chi2 = st->best_chi2;
    L = st->best_L;
  }
  st->L = L;  // lm.py:461-467
  for (int i = 0; i < NP; ++i) st->x[i] += dx[i];
  push_hist(st, chi2);
  Ldn(st);
  st->iter += 1;
  const int n = st->nhist;
  if (n >= 3 && (hist_back(st, 3) - hist_back(st, 1)) / hist_back(st, 1) < b.relative_tolerance &&
      st->L < 0.1) {
    st->status = S_SUCCESS;
  } else if (n > 10 &&
             (hist_back(st, 10) - hist_back(st, 1)) / hist_back(st, 1) < b.relative_tolerance) {
    st->status = S_IMMOBILE;
  } else if (st->iter >= b.max_iter) {
    st->status = S_MAXITER;
  }
  st->phase = st->status == S_RUN ? PH_NEEDJ : PH_DONE;
}

__global__ void lm_decide_kernel(apx_batch b) {  // lm.py:293-346
  const int g = blockIdx.x * 256 + threadIdx.x;
  if (g >= b.ngal) return;
  GalState* st = st_of(b, g);
  if (st->phase != PH_SUB) return;
  const double chi2 = st->chi2;
  bool brk = false;
  if (!isfinite(chi2)) {
    Lup(st);
    if (st->direction == DIR_BETTER) brk = true;
    else st->direction = DIR_WORSE;
  } else {
    if (chi2 <= st->scary_chi2) {
      for (int i = 0; i < NP; ++i) st->scary_dx[i] = st->ha[i];
      st->scary_chi2 = chi2;
      st->scary_L = st->L;
      st->scary_valid = 1;
    }
    double na = 0.0, nh = 0.0;
    for (int i = 0; i < NP; ++i) {
      na += st->a[i] * st->a[i];
      nh += st->h[i] * st->h[i];
    }
    if (sqrt(na) / sqrt(nh) > 1.0) {  // curvature_limit
      Lup(st);
      if (st->direction == DIR_BETTER) brk = true;
      else st->direction = DIR_WORSE;
    } else {
      if (chi2 < st->best_chi2) {
        for (int i = 0; i < NP; ++i) st->best_dx[i] = st->ha[i];
        st->best_chi2 = chi2;
        st->best_L = st->L;
        st->nostep = 0;
        Ldn(st);
        if (st->L <= 1e-8 || st->direction == DIR_WORSE) brk = true;
        else st->direction = DIR_BETTER;
      } else if (chi2 > st->best_chi2 &&
                 (st->direction == DIR_NONE || st->direction == DIR_WORSE)) {
        Lup(st);
        if (st->L == 1e9) brk = true;
        else st->direction = DIR_WORSE;
      } else {
        brk = true;
      }
      if (!brk && (st->best_chi2 - st->init_chi2) / st->init_chi2 < -0.1) brk = true;
    }
  }
  st->sub += 1;
  if (brk || st->sub >= 10) finish_step(b, st);
}

__global__ void results_kernel(apx_batch b, double* rep, double* chi2, int32_t* status,
                               int32_t* iters, int32_t* done) {
  const int g = blockIdx.x * 256 + threadIdx.x;
  if (g >= b.ngal) return;
  const GalState* st = st_of(b, g);
  for (int i = 0; i < NP; ++i) rep[(int64_t)g * NP + i] = st->nhist > 0 ? st->res_x[i] : st->x[i];
  if (chi2) chi2[g] = st->res_chi2;
  if (status) status[g] = st->status;
  if (iters) iters[g] = st->iter;
  if (done && st->phase == PH_DONE) atomicAdd(done, 1);
}

__global__ void zero_int_kernel(int32_t* p) { p[0] = 0; }

// Gauss-Legendre table on [-1, 1] by Newton iteration on P_n (the table
// scipy.special.roots_legendre returns, operations.py:110)
void gauss_legendre(int n, double* node, double* weight) {
  for (int i = 0; i < n; ++i) {
    double x = cos(kPi * (i + 0.75) / (n + 0.5));
    double dp = 1.0;
    for (int it = 0; it < 100; ++it) {
      double p0 = 1.0, p1 = x;
      for (int k = 2; k <= n; ++k) {
        const double p2 = ((2.0 * k - 1.0) * x * p1 - (k - 1.0) * p0) / k;
        p0 = p1;
        p1 = p2;
      }
      const double pn = n == 1 ? x : p1, pm = n == 1 ? 1.0 : p0;
      dp = n * (x * pn - pm) / (x * x - 1.0);
      const double dx = pn / dp;
      x -= dx;
      if (fabs(dx) < 1e-16) break;
    }
    node[n - 1 - i] = x;
    weight[n - 1 - i] = 2.0 / ((1.0 - x * x) * dp * dp);
  }
}

// work lists + render planes of the fits active in `stage`
void render_part(const apx_batch& b, const Scratch& s, int stage, const double* rep_override,
                 hipStream_t stream) {
  const bool jac = stage == ST_JAC;
  const int gb = (b.ngal + 255) / 256;
  hipLaunchKernelGGL(reset_kernel, dim3(1), dim3(64), 0, stream, s.cnt);
  hipLaunchKernelGGL(build_kernel, dim3(gb), dim3(256), 0, stream, b, s, stage);
  hipLaunchKernelGGL(prepare_kernel, dim3(std::min(b.ngal, 1024)), dim3(256), 0, stream, b, s, stage,
                     rep_override);
  const int pc = (int)std::min<int64_t>(b.nchunk, 4 * kPersist);
  if (jac)
    hipLaunchKernelGGL(init_kernel<true>, dim3(pc), dim3(kChunk), 0, stream, b, s);
  else
    hipLaunchKernelGGL(init_kernel<false>, dim3(pc), dim3(kChunk), 0, stream, b, s);
  hipLaunchKernelGGL(select_kernel, dim3(pc), dim3(256), 0, stream, b, s);
  apx_component cfg = {};
  cfg.quad_level = b.quad_level;
  cfg.gridding = b.gridding;
  cfg.max_depth = b.max_depth;
  gauss_legendre(b.quad_level, cfg.gl_node, cfg.gl_weight);
  if (jac)
    hipLaunchKernelGGL(integrate_kernel<true>, dim3(kPersist / 4), dim3(256), 0, stream, b, s, cfg);
  else
    hipLaunchKernelGGL(integrate_kernel<false>, dim3(kPersist / 4), dim3(256), 0, stream, b, s, cfg);
}

// convolution of the listed tiles into `out` (and J in the Jacobian stage)
void conv_part(const apx_batch& b, const Scratch& s, int stage, float* out, hipStream_t stream) {
  const bool jac = stage == ST_JAC;
  const int kmax = b.max_kernel;
  const size_t tile = (size_t)8 * fast_nt(kmax) * (size_t)(64 + kmax - 1);
  const size_t lds = sizeof(float) * std::max(tile, (size_t)4 * (jac ? 3 : 1) * kRYF * 64);
  const int tc = (int)std::min<int64_t>(b.ntile, 2 * kPersist);
  if (jac)
    hipLaunchKernelGGL(conv_kernel<true>, dim3(tc), dim3(256), lds, stream, b, s, out);
  else
    hipLaunchKernelGGL(conv_kernel<false>, dim3(tc), dim3(256), lds, stream, b, s, out);
}

// one render stage over the fits active in `stage`; parts: 1 planes, 2 conv
int run_stage(const apx_batch& b, const Scratch& s, int stage, const double* rep_override,
              float* out, hipStream_t stream, int parts = 3) {
  if (parts & 1) render_part(b, s, stage, rep_override, stream);
  if (parts & 2) conv_part(b, s, stage, out, stream);
  return check_launch("batch stage");
}

}  // namespace

int batch_plan(apx_batch_galaxy* gal, int32_t ngal, apx_batch* b) {
  if (!gal || !b || ngal < 1) return set_error(APX_E_INVALID, "batch_plan: bad arguments");
  int64_t pix = 0, psf = 0, work = 0, tap = 0, chunk = 0, tile = 0;
  int kmax = 1;
  for (int g = 0; g < ngal; ++g) {
    apx_batch_galaxy& G = gal[g];
    if (G.nx < 1 || G.ny < 1) return set_error(APX_E_INVALID, "batch_plan: empty cutout");
    if (G.ph < 1 || G.pw < 1 || G.ph % 2 == 0 || G.pw % 2 == 0)
      return set_error(APX_E_INVALID, "batch_plan: psf must have odd shape");
    if (G.ph + 2 > kFastMax || G.pw + 2 > kFastMax)
      return set_error(APX_E_LIMIT, "batch_plan: psf larger than the batched kernel supports");
    const int bx = (G.pw + 1) / 2, by = (G.ph + 1) / 2;
    const int64_t nw = (int64_t)(G.nx + 2 * bx) * (G.ny + 2 * by);
    G.pix_off = pix;
    G.psf_off = psf;
    G.work_off = work;
    G.tap_off = tap;
    G.chunk_off = chunk;
    G.tile_off = tile;
    pix += (int64_t)G.nx * G.ny;
    psf += (int64_t)G.ph * G.pw;
    work += nw;
    tap += 3 * (int64_t)(G.pw + 2) * fast_kts(G.ph + 2);
    chunk += (nw + kChunk - 1) / kChunk;
    tile += (int64_t)((G.nx + 63) / 64) * ((G.ny + kRYF - 1) / kRYF);
    kmax = std::max(kmax, std::max(G.ph, G.pw) + 2);
  }
  if (work > (int64_t)INT32_MAX || chunk > (int64_t)INT32_MAX || tile > (int64_t)INT32_MAX)
    return set_error(APX_E_LIMIT, "batch_plan: batch too large");
  b->ngal = ngal;
  b->max_kernel = kmax;
  b->npix = pix;
  b->nwork = work;
  b->nchunk = chunk;
  b->ntile = tile;
  b->ntap = tap;
  return APX_OK;
}

int64_t batch_state_bytes() { return (int64_t)sizeof(GalState); }

int64_t batch_scratch_bytes(const apx_batch* b) { return b ? scratch_bytes(*b) : 0; }

static int check_batch(const apx_batch* b) {
  if (!b || !b->gal || !b->data || !b->weight || !b->psf || !b->model0 || !b->model1 || !b->jac ||
      !b->planes || !b->taps || !b->state || !b->scratch)
    return set_error(APX_E_INVALID, "batch: null buffer");
  if (b->quad_level < 1 || b->quad_level > APX_MAX_QUAD_LEVEL || b->gridding < 1 ||
      b->gridding > APX_MAX_GRIDDING || b->max_depth < 1 || b->max_depth > APX_MAX_DEPTH)
    return set_error(APX_E_LIMIT, "batch: integration settings out of range");
  if (b->max_kernel > kFastMax) return set_error(APX_E_LIMIT, "batch: psf too large");
  return APX_OK;
}

int batch_render(const apx_batch* b, const double* rep, float* out, int32_t parts,
                 hipStream_t stream) {
  int rc = check_batch(b);
  if (rc) return rc;
  if (!rep || !out) return set_error(APX_E_INVALID, "batch_render: null pointer");
  if (parts < 1 || parts > 3) return set_error(APX_E_INVALID, "batch_render: parts must be 1, 2 or 3");
  const Scratch s = carve(*b);
  return run_stage(*b, s, ST_ALL, rep, out, stream, parts);
}

int batch_start(const apx_batch* b, const double* rep0, hipStream_t stream) {
  int rc = check_batch(b);
  if (rc) return rc;
  if (!rep0) return set_error(APX_E_INVALID, "batch_start: null pointer");
  hipLaunchKernelGGL(start_kernel, dim3((b->ngal + 255) / 256), dim3(256), 0, stream, *b, rep0);
  return check_launch("batch_start");
}

int batch_tick(const apx_batch* b, hipStream_t stream) {
  int rc = check_batch(b);
  if (rc) return rc;
  const Scratch s = carve(*b);
  const int gb = (b->ngal + 255) / 256;
  const int fb = std::min(b->ngal, 2 * kPersist);
  // Jacobian stage for fits starting a step
  hipLaunchKernelGGL(stage_eval_kernel, dim3(gb), dim3(256), 0, stream, *b);
  if ((rc = run_stage(*b, s, ST_JAC, nullptr, b->model0, stream))) return rc;
  hipLaunchKernelGGL(normal_kernel, dim3(fb), dim3(256), 0, stream, *b, s);
  hipLaunchKernelGGL(lm_begin_kernel, dim3(gb), dim3(256), 0, stream, *b);
  // one sub-step
  hipLaunchKernelGGL(lm_solve_h_kernel, dim3(gb), dim3(256), 0, stream, *b);
  if ((rc = run_stage(*b, s, ST_PROBE, nullptr, b->model1, stream))) return rc;
  hipLaunchKernelGGL(rpp_kernel, dim3(fb), dim3(256), 0, stream, *b, s);
  hipLaunchKernelGGL(lm_solve_a_kernel, dim3(gb), dim3(256), 0, stream, *b);
  if ((rc = run_stage(*b, s, ST_TRIAL, nullptr, b->model1, stream))) return rc;
  hipLaunchKernelGGL(chi2_kernel, dim3(fb), dim3(256), 0, stream, *b, s);
  hipLaunchKernelGGL(lm_decide_kernel, dim3(gb), dim3(256), 0, stream, *b);
  return check_launch("batch_tick");
}

int batch_results(const apx_batch* b, double* rep, double* chi2, int32_t* status, int32_t* iters,
                  int32_t* done, hipStream_t stream) {
  if (!b || !b->state || !rep) return set_error(APX_E_INVALID, "batch_results: null pointer");
  if (done) hipLaunchKernelGGL(zero_int_kernel, dim3(1), dim3(1), 0, stream, done);
  hipLaunchKernelGGL(results_kernel, dim3((b->ngal + 255) / 256), dim3(256), 0, stream, *b, rep,
                     chi2, status, iters, done);
  return check_launch("batch_results");
}

}  // namespace apx
