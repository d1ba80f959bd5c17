// Steric self-avoidance field (forces.py:84-118, utils.py:35-49):
// cloud-in-cell deposit of the predicted node positions, then the periodic
// Gaussian filter Re(ifft2(fft2(d) * exp(-2 pi^2 sigma^2 |f|^2))).
//
// Device pipeline (no dense density grid, no float atomics):
//   predict   per node: damped velocity, inertial prediction, wrap, CIC cell
//             and the four weights (forces.py:95-110); count node-row hits;
//   scan/fill node ids into per-row bins;
//   rows      one workgroup per row: sort the row's node ids, fold the CIC
//             weights of each cell in the reference's order (all w00 by node,
//             then w10, w01, w11 -- forces.py:112-115), forward FFT along x in
//             LDS, write the complex row spectrum;
//   columns   XCD-grouped strips of columns: forward FFT along y in LDS,
//             multiply by the Gaussian transfer function, inverse FFT along y;
//   irows     inverse FFT along x and the real part.
// The FFT is a Stockham autosort radix-4/2/3 (+ generic radix <= 16) network
// in LDS with a double-rounded twiddle table.
#include "mc_bin.h"
#include "mc_common.h"
#include "mc_internal.h"

namespace mc {
namespace {

constexpr int MAX_RADIX = 16;

__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }

// Batched Stockham FFT over `batch` contiguous length-n sequences in LDS.
// Returns the buffer that holds the result (a or b).
__device__ float2* fft_lds(float2* a, float2* b, int n, int batch, const int* radix, int nrad,
                           const float2* __restrict__ tw, bool inverse) {
  float2* src = a;
  float2* dst = b;
  int Ns = 1;
  for (int s = 0; s < nrad; ++s) {
    const int R = radix[s];
    const int nb = n / R;
    const int tstride = n / (Ns * R);
    for (int idx = threadIdx.x; idx < batch * nb; idx += blockDim.x) {
      const int col = idx / nb;
      const int j = idx - col * nb;
      const int base = col * n;
      const int k = j % Ns;
      float2 v[MAX_RADIX];
      for (int r = 0; r < R; ++r) {
        float2 x = src[base + j + r * nb];
        if (r > 0 && k > 0) {
          float2 w = tw[r * k * tstride];
          if (inverse) w.y = -w.y;
          x = cmul(x, w);
        }
        v[r] = x;
      }
      float2 V[MAX_RADIX];
      if (R == 2) {
        V[0] = cadd(v[0], v[1]);
        V[1] = csub(v[0], v[1]);
      } else if (R == 4) {
        const float2 a0 = cadd(v[0], v[2]), a1 = csub(v[0], v[2]);
        const float2 a2 = cadd(v[1], v[3]), d = csub(v[1], v[3]);
        // forward: -i * d ; inverse: +i * d
        const float2 a3 = inverse ? make_float2(-d.y, d.x) : make_float2(d.y, -d.x);
        V[0] = cadd(a0, a2);
        V[2] = csub(a0, a2);
        V[1] = cadd(a1, a3);
        V[3] = csub(a1, a3);
      } else if (R == 3) {
        const float s3 = 0.866025403784438647f;  // sin(2 pi / 3)
        const float2 t1 = cadd(v[1], v[2]), t2 = csub(v[1], v[2]);
        const float2 m = make_float2(v[0].x - 0.5f * t1.x, v[0].y - 0.5f * t1.y);
        // forward: -i s t2 ; inverse: +i s t2
        const float2 st = inverse ? make_float2(-s3 * t2.y, s3 * t2.x) : make_float2(s3 * t2.y, -s3 * t2.x);
        V[0] = cadd(v[0], t1);
        V[1] = cadd(m, st);
        V[2] = csub(m, st);
      } else {
        for (int q = 0; q < R; ++q) {
          float2 acc = v[0];
          for (int r = 1; r < R; ++r) {
            float2 w = tw[((r * q) % R) * nb];
            if (inverse) w.y = -w.y;
            acc = cadd(acc, cmul(v[r], w));
          }
          V[q] = acc;
        }
      }
      const int idxD = (j / Ns) * Ns * R + k;
      for (int r = 0; r < R; ++r) dst[base + idxD + r * Ns] = V[r];
    }
    __syncthreads();
    float2* t = src;
    src = dst;
    dst = t;
    Ns *= R;
  }
  return src;
}

__global__ void k_predict(Consts c, int N, const float* __restrict__ pos, const float* __restrict__ vel,
                          float* __restrict__ pred_w, int* __restrict__ cic_cell,
                          float* __restrict__ cic_w, int* __restrict__ row_cnt) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= N) return;
  float px = pos[2 * n], py = pos[2 * n + 1];
  if (vel) {  // inertial prediction (simulate.py:46-50); without vel: positions as given
    const float vx = vel[2 * n] * c.damping, vy = vel[2 * n + 1] * c.damping;
    px = px + vx * c.dt;
    py = py + vy * c.dt;
  }
  const float x = fmodf_floor(px, c.dom_wf), y = fmodf_floor(py, c.dom_hf);
  pred_w[2 * n] = x;
  pred_w[2 * n + 1] = y;
  if (!c.enable_steric) return;
  const int W = c.dom_w, H = c.dom_h;
  const float fx = floorf(x), fy = floorf(y);
  const int x0 = imod(f2i(fx), W), y0 = imod(f2i(fy), H);
  const int x1 = imod(x0 + 1, W), y1 = imod(y0 + 1, H);
  const float dx = x - fx, dy = y - fy;
  const float sv = c.steric_sv;
  cic_cell[4 * n + 0] = x0;
  cic_cell[4 * n + 1] = x1;
  cic_cell[4 * n + 2] = y0;
  cic_cell[4 * n + 3] = y1;
  cic_w[4 * n + 0] = ((1.0f - dx) * (1.0f - dy)) * sv;
  cic_w[4 * n + 1] = (dx * (1.0f - dy)) * sv;
  cic_w[4 * n + 2] = ((1.0f - dx) * dy) * sv;
  cic_w[4 * n + 3] = (dx * dy) * sv;
  atomicAdd(&row_cnt[y0], 1);
  if (y1 != y0) atomicAdd(&row_cnt[y1], 1);
}

__global__ void k_row_fill(int N, const int* __restrict__ cic_cell, const int* __restrict__ row_off,
                           int* __restrict__ row_fill, int* __restrict__ row_list) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= N) return;
  const int y0 = cic_cell[4 * n + 2], y1 = cic_cell[4 * n + 3];
  int p = atomicAdd(&row_fill[y0], 1);
  row_list[row_off[y0] + p] = n;
  if (y1 != y0) {
    p = atomicAdd(&row_fill[y1], 1);
    row_list[row_off[y1] + p] = n;
  }
}

constexpr int DEP_CH = 256;

__global__ void __launch_bounds__(256) k_deposit_rows_fft(
    int H, int W, const int* __restrict__ row_off, int* __restrict__ row_list,
    int* __restrict__ sort_tmp, const int* __restrict__ cic_cell, const float* __restrict__ cic_w,
    float2* __restrict__ spec, const float2* __restrict__ tw, const int* __restrict__ radix_g,
    int nrad, float* __restrict__ density_out) {
  extern __shared__ float2 lds[];
  __shared__ int sort_lds[SORT_LDS_CAP];
  __shared__ int s_x[DEP_CH][2];
  __shared__ int s_y[DEP_CH][2];
  __shared__ float s_w[DEP_CH][4];
  __shared__ int s_radix[32];
  float2* a = lds;
  float2* b = lds + W;
  if (threadIdx.x < nrad) s_radix[threadIdx.x] = radix_g[threadIdx.x];
  const int y = blockIdx.x;
  const int off = row_off[y];
  const int m = row_off[y + 1] - off;
  sort_unique_ids(row_list + off, m, sort_lds, sort_tmp + off);
  for (int x = threadIdx.x; x < W; x += blockDim.x) a[x] = make_float2(0.0f, 0.0f);
  __syncthreads();
  // fold: corner-major (00, 10, 01, 11), node-minor, per cell
  for (int q = 0; q < 4; ++q) {
    const int xsel = q & 1;         // 0: x0, 1: x1
    const int ysel = q >> 1;        // 0: y0, 1: y1
    for (int j0 = 0; j0 < m; j0 += DEP_CH) {
      const int cnt = min(DEP_CH, m - j0);
      for (int j = threadIdx.x; j < cnt; j += blockDim.x) {
        const int n = row_list[off + j0 + j];
        s_x[j][0] = cic_cell[4 * n + 0];
        s_x[j][1] = cic_cell[4 * n + 1];
        s_y[j][0] = cic_cell[4 * n + 2];
        s_y[j][1] = cic_cell[4 * n + 3];
        s_w[j][0] = cic_w[4 * n + 0];
        s_w[j][1] = cic_w[4 * n + 1];
        s_w[j][2] = cic_w[4 * n + 2];
        s_w[j][3] = cic_w[4 * n + 3];
      }
      __syncthreads();
      for (int x = threadIdx.x; x < W; x += blockDim.x) {
        float acc = a[x].x;
        for (int j = 0; j < cnt; ++j)
          if (s_y[j][ysel] == y && s_x[j][xsel] == x) acc = acc + s_w[j][q];
        a[x].x = acc;
      }
      __syncthreads();
    }
  }
  if (density_out)
    for (int x = threadIdx.x; x < W; x += blockDim.x) density_out[(size_t)y * W + x] = a[x].x;
  float2* r = fft_lds(a, b, W, 1, s_radix, nrad, tw, false);
  for (int x = threadIdx.x; x < W; x += blockDim.x) spec[(size_t)y * W + x] = r[x];
}

__global__ void __launch_bounds__(256) k_fft_rows(int H, int W, const float* __restrict__ in,
                                                  float2* __restrict__ spec,
                                                  const float2* __restrict__ tw,
                                                  const int* __restrict__ radix_g, int nrad) {
  extern __shared__ float2 lds[];
  __shared__ int s_radix[32];
  float2* a = lds;
  float2* b = lds + W;
  if (threadIdx.x < nrad) s_radix[threadIdx.x] = radix_g[threadIdx.x];
  const int y = blockIdx.x;
  for (int x = threadIdx.x; x < W; x += blockDim.x) a[x] = make_float2(in[(size_t)y * W + x], 0.0f);
  __syncthreads();
  float2* r = fft_lds(a, b, W, 1, s_radix, nrad, tw, false);
  for (int x = threadIdx.x; x < W; x += blockDim.x) spec[(size_t)y * W + x] = r[x];
}

// fftfreq(n)[k] as JAX computes it: float(k or k - n) / float(n)
__device__ __forceinline__ float fftfreq(int k, int n) {
  const int kk = (k < (n + 1) / 2) ? k : k - n;
  return (float)kk / (float)n;
}

__global__ void __launch_bounds__(256) k_fft_cols_filter(int H, int W, int cs,
                                                         float2* __restrict__ spec,
                                                         const float2* __restrict__ tw,
                                                         const int* __restrict__ radix_g,
                                                         int nrad, float coef) {
  extern __shared__ float2 lds[];
  __shared__ int s_radix[32];
  float2* a = lds;
  float2* b = lds + (size_t)cs * H;
  if (threadIdx.x < nrad) s_radix[threadIdx.x] = radix_g[threadIdx.x];
  // XCD-aware strip order: consecutive workgroups land on different XCDs, so
  // give the 8 workgroups of one dispatch round strips far apart and the
  // workgroups of one XCD neighbouring strips (they share cache lines).
  const int nstrips = (W + cs - 1) / cs;
  const int g = gridDim.x;
  int strip = blockIdx.x;
  if (g % 8 == 0) strip = (blockIdx.x % 8) * (g / 8) + blockIdx.x / 8;
  const int c0 = strip * cs;
  const int ncol = min(cs, W - c0);
  if (c0 >= W) return;
  for (int i = threadIdx.x; i < ncol * H; i += blockDim.x) {
    const int row = i / ncol, col = i - row * ncol;
    a[col * H + row] = spec[(size_t)row * W + c0 + col];
  }
  __syncthreads();
  float2* r = fft_lds(a, b, H, ncol, s_radix, nrad, tw, false);
  float2* o = (r == a) ? b : a;
  for (int i = threadIdx.x; i < ncol * H; i += blockDim.x) {
    const int col = i / H, row = i - col * H;
    const float fr = fftfreq(row, H), fc = fftfreq(c0 + col, W);
    const float kf = dexp(coef * (fr * fr + fc * fc));
    float2 v = r[i];
    v.x = v.x * kf;
    v.y = v.y * kf;
    r[i] = v;
  }
  __syncthreads();
  float2* res = fft_lds(r, o, H, ncol, s_radix, nrad, tw, true);
  const float sc = (float)(1.0 / (double)H);
  for (int i = threadIdx.x; i < ncol * H; i += blockDim.x) {
    const int row = i / ncol, col = i - row * ncol;
    float2 v = res[col * H + row];
    spec[(size_t)row * W + c0 + col] = make_float2(v.x * sc, v.y * sc);
  }
}

__global__ void __launch_bounds__(256) k_ifft_rows_real(int H, int W, const float2* __restrict__ spec,
                                                        float* __restrict__ out,
                                                        const float2* __restrict__ tw,
                                                        const int* __restrict__ radix_g, int nrad) {
  extern __shared__ float2 lds[];
  __shared__ int s_radix[32];
  float2* a = lds;
  float2* b = lds + W;
  if (threadIdx.x < nrad) s_radix[threadIdx.x] = radix_g[threadIdx.x];
  const int y = blockIdx.x;
  for (int x = threadIdx.x; x < W; x += blockDim.x) a[x] = spec[(size_t)y * W + x];
  __syncthreads();
  float2* r = fft_lds(a, b, W, 1, s_radix, nrad, tw, true);
  const float sc = (float)(1.0 / (double)W);
  for (int x = threadIdx.x; x < W; x += blockDim.x) out[(size_t)y * W + x] = r[x].x * sc;
}

}  // namespace

hipError_t launch_predict(const Consts& c, int N, const float* pos, const float* vel, float* pred_w,
                          int* cic_cell, float* cic_w, int* row_cnt, hipStream_t s) {
  if (N == 0) return hipSuccess;
  hipLaunchKernelGGL(k_predict, dim3((N + 255) / 256), dim3(256), 0, s, c, N, pos, vel, pred_w,
                     cic_cell, cic_w, row_cnt);
  return hipGetLastError();
}

hipError_t launch_row_fill(const Consts& c, int N, const int* cic_cell, const int* row_off,
                           int* row_fill, int* row_list, hipStream_t s) {
  if (N == 0) return hipSuccess;
  hipLaunchKernelGGL(k_row_fill, dim3((N + 255) / 256), dim3(256), 0, s, N, cic_cell, row_off,
                     row_fill, row_list);
  return hipGetLastError();
}

hipError_t launch_deposit_rows_fft(const Consts& c, int N, const int* row_off, int* row_list,
                                   const int* cic_cell, const float* cic_w, float2* spec,
                                   const float2* tw_w, const int* radix_w, int nrad_w,
                                   float* density_out, int* sort_scratch, hipStream_t s) {
  const int H = c.dom_h, W = c.dom_w;
  const size_t lds = (size_t)2 * W * sizeof(float2);
  hipLaunchKernelGGL(k_deposit_rows_fft, dim3(H), dim3(256), lds, s, H, W, row_off, row_list,
                     sort_scratch, cic_cell, cic_w, spec, tw_w, radix_w, nrad_w, density_out);
  return hipGetLastError();
}

hipError_t launch_fft_rows(int H, int W, const float* real_in, float2* spec, const float2* tw_w,
                           const int* radix_w, int nrad_w, hipStream_t s) {
  const size_t lds = (size_t)2 * W * sizeof(float2);
  hipLaunchKernelGGL(k_fft_rows, dim3(H), dim3(256), lds, s, H, W, real_in, spec, tw_w, radix_w,
                     nrad_w);
  return hipGetLastError();
}

hipError_t launch_fft_cols_filter(int H, int W, float2* spec, const float2* tw_h,
                                  const int* radix_h, int nrad_h, float coef, int lds_bytes_max,
                                  hipStream_t s) {
  int cs = 16;
  while (cs > 1 && (size_t)2 * cs * H * sizeof(float2) > (size_t)lds_bytes_max) cs >>= 1;
  const size_t lds = (size_t)2 * cs * H * sizeof(float2);
  const int nstrips = (W + cs - 1) / cs;
  hipLaunchKernelGGL(k_fft_cols_filter, dim3(nstrips), dim3(256), lds, s, H, W, cs, spec, tw_h,
                     radix_h, nrad_h, coef);
  return hipGetLastError();
}

hipError_t launch_ifft_rows_real(int H, int W, const float2* spec, float* out, const float2* tw_w,
                                 const int* radix_w, int nrad_w, hipStream_t s) {
  const size_t lds = (size_t)2 * W * sizeof(float2);
  hipLaunchKernelGGL(k_ifft_rows_real, dim3(H), dim3(256), lds, s, H, W, spec, out, tw_w, radix_w,
                     nrad_w);
  return hipGetLastError();
}

}  // namespace mc
