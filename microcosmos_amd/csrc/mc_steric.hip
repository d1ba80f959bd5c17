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

__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }

// Radix-R butterfly on registers (compile-time R keeps v[] in VGPRs).
template <int R>
__device__ __forceinline__ void butterfly(float2 (&v)[R], float2 (&V)[R], bool inverse,
                                          const float2* __restrict__ tw, int nb) {
  if constexpr (R == 2) {
    V[0] = cadd(v[0], v[1]);
    V[1] = csub(v[0], v[1]);
  } else if constexpr (R == 4) {
    const float2 a0 = cadd(v[0], v[2]), a1 = csub(v[0], v[2]);
    const float2 a2 = cadd(v[1], v[3]), d = csub(v[1], v[3]);
    // forward: -i * d ; inverse: +i * d
    const float2 a3 = inverse ? make_float2(-d.y, d.x) : make_float2(d.y, -d.x);
    V[0] = cadd(a0, a2);
    V[2] = csub(a0, a2);
    V[1] = cadd(a1, a3);
    V[3] = csub(a1, a3);
  } else if constexpr (R == 3) {
    const float s3 = 0.866025403784438647f;  // sin(2 pi / 3)
    const float2 t1 = cadd(v[1], v[2]), t2 = csub(v[1], v[2]);
    const float2 m = make_float2(v[0].x - 0.5f * t1.x, v[0].y - 0.5f * t1.y);
    // forward: -i s t2 ; inverse: +i s t2
    const float2 st = inverse ? make_float2(-s3 * t2.y, s3 * t2.x) : make_float2(s3 * t2.y, -s3 * t2.x);
    V[0] = cadd(v[0], t1);
    V[1] = cadd(m, st);
    V[2] = csub(m, st);
  } else {
#pragma unroll
    for (int q = 0; q < R; ++q) {
      float2 acc = v[0];
#pragma unroll
      for (int r = 1; r < R; ++r) {
        float2 w = tw[((r * q) % R) * nb];
        if (inverse) w.y = -w.y;
        acc = cadd(acc, cmul(v[r], w));
      }
      V[q] = acc;
    }
  }
}

// One Stockham stage of radix R over `batch` contiguous length-n sequences.
template <int R>
__device__ __forceinline__ void fft_stage(const float2* src, float2* dst, int n, int stride, int batch,
                                          int Ns, const float2* __restrict__ tw, bool inverse) {
  const int nb = n / R;
  const int tstride = n / (Ns * R);
  for (int idx = threadIdx.x; idx < batch * nb; idx += blockDim.x) {
    const int col = idx / nb;
    const int j = idx - col * nb;
    const int base = col * stride;
    const int k = j % Ns;
    float2 v[R], V[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      float2 x = src[base + j + r * nb];
      if (r > 0 && k > 0) {
        float2 w = tw[r * k * tstride];
        if (inverse) w.y = -w.y;
        x = cmul(x, w);
      }
      v[r] = x;
    }
    butterfly<R>(v, V, inverse, tw, nb);
    const int idxD = (j / Ns) * Ns * R + k;
#pragma unroll
    for (int r = 0; r < R; ++r) dst[base + idxD + r * Ns] = V[r];
  }
}

// Radix 5 / 7 / 11 / 13 stage without register arrays: each output re-reads and
// re-twiddles its inputs (same roundings as fft_stage, fewer VGPRs; these
// radices only occur for unusual grid sides).
__device__ __forceinline__ void fft_stage_generic(const float2* src, float2* dst, int n, int stride,
                                                  int batch, int Ns, int R,
                                                  const float2* __restrict__ tw, bool inverse) {
  const int nb = n / R;
  const int tstride = n / (Ns * R);
  for (int idx = threadIdx.x; idx < batch * nb; idx += blockDim.x) {
    const int col = idx / nb;
    const int j = idx - col * nb;
    const int base = col * stride;
    const int k = j % Ns;
    const int idxD = (j / Ns) * Ns * R + k;
#pragma unroll 1
    for (int q = 0; q < R; ++q) {
      float2 acc = src[base + j];
#pragma unroll 1
      for (int r = 1; r < R; ++r) {
        float2 x = src[base + j + r * nb];
        if (k > 0) {
          float2 w = tw[r * k * tstride];
          if (inverse) w.y = -w.y;
          x = cmul(x, w);
        }
        float2 w = tw[((r * q) % R) * nb];
        if (inverse) w.y = -w.y;
        acc = cadd(acc, cmul(x, w));
      }
      dst[base + idxD + q * Ns] = acc;
    }
  }
}

// Batched Stockham FFT over `batch` length-n sequences in LDS, sequence c at
// offset c * stride.  Radices come from the host schedule (4s, then 2, then
// odd primes <= 13).  Returns the buffer that holds the result (a or b).
__device__ float2* fft_lds(float2* a, float2* b, int n, int stride, int batch, const int* radix,
                           int nrad, const float2* __restrict__ tw, bool inverse) {
  float2* src = a;
  float2* dst = b;
  int Ns = 1;
  for (int s = 0; s < nrad; ++s) {
    const int R = radix[s];
    switch (R) {
      case 4: fft_stage<4>(src, dst, n, stride, batch, Ns, tw, inverse); break;
      case 2: fft_stage<2>(src, dst, n, stride, batch, Ns, tw, inverse); break;
      case 3: fft_stage<3>(src, dst, n, stride, batch, Ns, tw, inverse); break;
      default: fft_stage_generic(src, dst, n, stride, batch, Ns, R, tw, inverse); break;
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

constexpr int DEP_KEYS_LDS = 512;    // 64-bit keys of one row's CIC entries held in LDS
constexpr int DEP_IDS_LDS = 1024;    // node ids of one row sorted in LDS

__device__ __forceinline__ void load_twiddles(float2* dst, const float2* __restrict__ tw, int n) {
  for (int i = threadIdx.x; i < n; i += blockDim.x) dst[i] = tw[i];
}

// One workgroup per row y: fold the row's CIC entries cell by cell in the
// reference's order (all w00 by node, then w10, w01, w11 -- forces.py:112-115),
// then the forward FFT of the row.  Each node binned to the row contributes
// exactly two entries (corners q = 0, 1 if y == y0; q = 2, 3 if y == y1).
// Sorting the entries by (x, q, node rank) puts each cell's contributions
// contiguously and in reference order; a segment head thread folds them.
__global__ void __launch_bounds__(256) k_deposit_rows_fft(
    int H, int W, const int* __restrict__ row_off, int* __restrict__ row_list,
    int* __restrict__ sort_tmp, unsigned long long* __restrict__ key_tmp, size_t key_half,
    const int* __restrict__ cic_cell, const float* __restrict__ cic_w, float2* __restrict__ spec,
    const float2* __restrict__ tw_g, const int* __restrict__ radix_g, int nrad,
    float* __restrict__ density_out) {
  extern __shared__ float2 lds[];
  __shared__ int sort_lds[DEP_IDS_LDS];
  __shared__ unsigned long long keys_lds[DEP_KEYS_LDS];
  __shared__ unsigned long long sort_keys[DEP_KEYS_LDS];
  __shared__ int s_radix[32];
  float2* a = lds;
  float2* b = lds + W;
  float2* tw = lds + 2 * W;
  if (threadIdx.x < nrad) s_radix[threadIdx.x] = radix_g[threadIdx.x];
  load_twiddles(tw, tw_g, W);
  const int y = blockIdx.x;
  const int off = row_off[y];
  const int m = row_off[y + 1] - off;
  int* nodes = row_list + off;
  sort_unique<int>(nodes, m, sort_lds, DEP_IDS_LDS, sort_tmp + off, 0x7fffffff);
  for (int x = threadIdx.x; x < W; x += blockDim.x) a[x] = make_float2(0.0f, 0.0f);
  const int ne = 2 * m;
  unsigned long long* keys = (ne <= DEP_KEYS_LDS) ? keys_lds : (key_tmp + 2 * (size_t)off);
  for (int j = threadIdx.x; j < m; j += blockDim.x) {
    const int n = nodes[j];
    const unsigned long long x0 = (unsigned)cic_cell[4 * n + 0], x1 = (unsigned)cic_cell[4 * n + 1];
    const unsigned long long qb = (cic_cell[4 * n + 2] == y) ? 0ull : 2ull;
    keys[2 * j] = (x0 << 34) | (qb << 32) | (unsigned)j;
    keys[2 * j + 1] = (x1 << 34) | ((qb + 1ull) << 32) | (unsigned)j;
  }
  __threadfence_block();
  __syncthreads();
  // beyond the LDS capacity the keys live in key_tmp[2 off ..) and the merge
  // sort ping-pongs through the second half of key_tmp (key_half entries on)
  sort_unique<unsigned long long>(keys, ne, sort_keys, DEP_KEYS_LDS,
                                  key_tmp + key_half + 2 * (size_t)off, ~0ull);
  for (int i = threadIdx.x; i < ne; i += blockDim.x) {
    const unsigned long long k = keys[i];
    const int x = (int)(k >> 34);
    if (i > 0 && (int)(keys[i - 1] >> 34) == x) continue;
    float acc = 0.0f;
    for (int e = i; e < ne; ++e) {
      const unsigned long long ke = keys[e];
      if ((int)(ke >> 34) != x) break;
      const int q = (int)((ke >> 32) & 3ull);
      const int n = nodes[(int)(ke & 0xffffffffull)];
      acc = acc + cic_w[4 * n + q];
    }
    a[x].x = acc;
  }
  __syncthreads();
  if (density_out)
    for (int x = threadIdx.x; x < W; x += blockDim.x) density_out[(size_t)y * W + x] = a[x].x;
  float2* r = fft_lds(a, b, W, W, 1, s_radix, nrad, tw, false);
  for (int x = threadIdx.x; x < W; x += blockDim.x) spec[(size_t)y * W + x] = r[x];
}

__global__ void __launch_bounds__(256) k_fft_rows(int H, int W, const float* __restrict__ in,
                                                  float2* __restrict__ spec,
                                                  const float2* __restrict__ tw_g,
                                                  const int* __restrict__ radix_g, int nrad) {
  extern __shared__ float2 lds[];
  __shared__ int s_radix[32];
  float2* a = lds;
  float2* b = lds + W;
  float2* tw = lds + 2 * W;
  if (threadIdx.x < nrad) s_radix[threadIdx.x] = radix_g[threadIdx.x];
  load_twiddles(tw, tw_g, W);
  const int y = blockIdx.x;
  for (int x = threadIdx.x; x < W; x += blockDim.x) a[x] = make_float2(in[(size_t)y * W + x], 0.0f);
  __syncthreads();
  float2* r = fft_lds(a, b, W, W, 1, s_radix, nrad, tw, false);
  for (int x = threadIdx.x; x < W; x += blockDim.x) spec[(size_t)y * W + x] = r[x];
}

// fftfreq(n)[k] as JAX computes it: float(k or k - n) / float(n)
__device__ __forceinline__ float fftfreq(int k, int n) {
  const int kk = (k < (n + 1) / 2) ? k : k - n;
  return (float)kk / (float)n;
}

// Gaussian transfer function exp(coef (fr^2 + fc^2)) on the fftfreq grid
// (utils.py:41-47), tabulated once per sigma.
__global__ void k_gauss_table(int H, int W, float coef, float* __restrict__ kf) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (size_t)H * W) return;
  const int row = (int)(i / W), col = (int)(i - (size_t)row * W);
  const float fr = fftfreq(row, H), fc = fftfreq(col, W);
  kf[i] = dexp(coef * (fr * fr + fc * fc));
}

__global__ void __launch_bounds__(256) k_fft_cols_filter(int H, int W, int cs,
                                                         float2* __restrict__ spec,
                                                         const float2* __restrict__ tw_g,
                                                         const int* __restrict__ radix_g,
                                                         int nrad, const float* __restrict__ kft) {
  extern __shared__ float2 lds[];
  __shared__ int s_radix[32];
  // columns are stored with a padded stride so the transposing loads and
  // stores spread over the LDS banks
  const int cst = H + 1;
  float2* a = lds;
  float2* b = lds + (size_t)cs * cst;
  float2* tw = lds + (size_t)2 * cs * cst;
  if (threadIdx.x < nrad) s_radix[threadIdx.x] = radix_g[threadIdx.x];
  load_twiddles(tw, tw_g, H);
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
  // issue all strided loads of this thread before any LDS store (the loads
  // are independent; a load-store loop would serialise their latencies)
  {
    constexpr int U = 16;
    const int total = ncol * H;
    for (int i0 = 0; i0 < total; i0 += U * blockDim.x) {
      float2 v[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int i = i0 + u * blockDim.x + threadIdx.x;
        if (i < total) {
          const int row = i / ncol, col = i - row * ncol;
          v[u] = spec[(size_t)row * W + c0 + col];
        }
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int i = i0 + u * blockDim.x + threadIdx.x;
        if (i < total) {
          const int row = i / ncol, col = i - row * ncol;
          a[col * cst + row] = v[u];
        }
      }
    }
  }
  __syncthreads();
  float2* r = fft_lds(a, b, H, cst, ncol, s_radix, nrad, tw, false);
  float2* o = (r == a) ? b : a;
  for (int i = threadIdx.x; i < ncol * H; i += blockDim.x) {
    const int col = i / H, row = i - col * H;
    const float kf = kft[(size_t)row * W + c0 + col];
    float2 v = r[col * cst + row];
    v.x = v.x * kf;
    v.y = v.y * kf;
    r[col * cst + row] = v;
  }
  __syncthreads();
  float2* res = fft_lds(r, o, H, cst, ncol, s_radix, nrad, tw, true);
  const float sc = (float)(1.0 / (double)H);
  for (int i = threadIdx.x; i < ncol * H; i += blockDim.x) {
    const int row = i / ncol, col = i - row * ncol;
    float2 v = res[col * cst + row];
    spec[(size_t)row * W + c0 + col] = make_float2(v.x * sc, v.y * sc);
  }
}

__global__ void __launch_bounds__(256) k_ifft_rows_real(int H, int W, const float2* __restrict__ spec,
                                                        float* __restrict__ out,
                                                        const float2* __restrict__ tw_g,
                                                        const int* __restrict__ radix_g, int nrad) {
  extern __shared__ float2 lds[];
  __shared__ int s_radix[32];
  float2* a = lds;
  float2* b = lds + W;
  float2* tw = lds + 2 * W;
  if (threadIdx.x < nrad) s_radix[threadIdx.x] = radix_g[threadIdx.x];
  load_twiddles(tw, tw_g, W);
  const int y = blockIdx.x;
  for (int x = threadIdx.x; x < W; x += blockDim.x) a[x] = spec[(size_t)y * W + x];
  __syncthreads();
  float2* r = fft_lds(a, b, W, W, 1, s_radix, nrad, tw, true);
  const float sc = (float)(1.0 / (double)W);
  for (int x = threadIdx.x; x < W; x += blockDim.x) out[(size_t)y * W + x] = r[x].x * sc;
}

// ---------------------------------------------------------------------------
// Transposed 2-D FFT: the row passes handle TR consecutive rows per workgroup
// and exchange the spectrum with the column pass in transposed layout
// specT[k * H + y], in contiguous TR-element segments; the column pass then
// reads whole contiguous rows of specT.  Same butterflies, same roundings as
// the untransposed passes.
// ---------------------------------------------------------------------------
__device__ __forceinline__ int xcd_group(int b, int G) {
  // workgroup i runs on XCD i % 8: give each XCD a contiguous range of groups
  return (G & 7) == 0 ? (b & 7) * (G >> 3) + (b >> 3) : b;
}

// CIC fold of row y into dst[0..W) (see k_deposit_rows_fft for the order)
__device__ void deposit_row(int y, int W, const int* __restrict__ row_off, int* __restrict__ row_list,
                            int* __restrict__ sort_tmp, unsigned long long* __restrict__ key_tmp,
                            size_t key_half, const int* __restrict__ cic_cell,
                            const float* __restrict__ cic_w, float2* dst, int* sort_lds,
                            unsigned long long* keys_lds, unsigned long long* sort_keys) {
  const int off = row_off[y];
  const int m = row_off[y + 1] - off;
  int* nodes = row_list + off;
  sort_unique<int>(nodes, m, sort_lds, DEP_IDS_LDS, sort_tmp + off, 0x7fffffff);
  for (int x = threadIdx.x; x < W; x += blockDim.x) dst[x] = make_float2(0.0f, 0.0f);
  const int ne = 2 * m;
  unsigned long long* keys = (ne <= DEP_KEYS_LDS) ? keys_lds : (key_tmp + 2 * (size_t)off);
  for (int j = threadIdx.x; j < m; j += blockDim.x) {
    const int n = nodes[j];
    const unsigned long long x0 = (unsigned)cic_cell[4 * n + 0], x1 = (unsigned)cic_cell[4 * n + 1];
    const unsigned long long qb = (cic_cell[4 * n + 2] == y) ? 0ull : 2ull;
    keys[2 * j] = (x0 << 34) | (qb << 32) | (unsigned)j;
    keys[2 * j + 1] = (x1 << 34) | ((qb + 1ull) << 32) | (unsigned)j;
  }
  __threadfence_block();
  __syncthreads();
  sort_unique<unsigned long long>(keys, ne, sort_keys, DEP_KEYS_LDS, key_tmp + key_half + 2 * (size_t)off,
                                  ~0ull);
  for (int i = threadIdx.x; i < ne; i += blockDim.x) {
    const unsigned long long k = keys[i];
    const int x = (int)(k >> 34);
    if (i > 0 && (int)(keys[i - 1] >> 34) == x) continue;
    float acc = 0.0f;
    for (int e = i; e < ne; ++e) {
      const unsigned long long ke = keys[e];
      if ((int)(ke >> 34) != x) break;
      const int q = (int)((ke >> 32) & 3ull);
      const int n = nodes[(int)(ke & 0xffffffffull)];
      acc = acc + cic_w[4 * n + q];
    }
    dst[x].x = acc;
  }
  __syncthreads();
}

// Forward row FFT of TR rows (from the CIC deposit when `real_in` is NULL,
// else from a real (H, W) array), written transposed to specT.
template <int TR>
__global__ void __launch_bounds__(256) k_rows_fwd_T(
    int H, int W, const float* __restrict__ real_in, const int* __restrict__ row_off,
    int* __restrict__ row_list, int* __restrict__ sort_tmp, unsigned long long* __restrict__ key_tmp,
    size_t key_half, const int* __restrict__ cic_cell, const float* __restrict__ cic_w,
    float2* __restrict__ specT, const float2* __restrict__ tw_g, const int* __restrict__ radix_g,
    int nrad, float* __restrict__ density_out) {
  extern __shared__ float2 lds[];
  __shared__ int sort_lds[DEP_IDS_LDS];
  __shared__ unsigned long long keys_lds[DEP_KEYS_LDS];
  __shared__ unsigned long long sort_keys[DEP_KEYS_LDS];
  __shared__ int s_radix[32];
  float2* a = lds;
  float2* b = lds + (size_t)TR * W;
  float2* tw = lds + (size_t)2 * TR * W;
  if (threadIdx.x < nrad) s_radix[threadIdx.x] = radix_g[threadIdx.x];
  load_twiddles(tw, tw_g, W);
  const int y0 = xcd_group(blockIdx.x, gridDim.x) * TR;
  const int nr = min(TR, H - y0);
  if (real_in) {
    for (int i = threadIdx.x; i < nr * W; i += blockDim.x)
      a[i] = make_float2(real_in[(size_t)y0 * W + i], 0.0f);
    __syncthreads();
  } else {
    for (int r = 0; r < nr; ++r)
      deposit_row(y0 + r, W, row_off, row_list, sort_tmp, key_tmp, key_half, cic_cell, cic_w,
                  a + (size_t)r * W, sort_lds, keys_lds, sort_keys);
    if (density_out)
      for (int i = threadIdx.x; i < nr * W; i += blockDim.x) density_out[(size_t)y0 * W + i] = a[i].x;
  }
  float2* res = fft_lds(a, b, W, W, nr, s_radix, nrad, tw, false);
  // transposed store: consecutive threads write consecutive rows of one column
  for (int i = threadIdx.x; i < nr * W; i += blockDim.x) {
    const int k = i / nr, r = i - k * nr;
    specT[(size_t)k * H + y0 + r] = res[(size_t)r * W + k];
  }
}

// Column pass on the transposed spectrum: one column k per workgroup, forward
// FFT along y, transfer function, inverse FFT scaled by 1/H.
__global__ void __launch_bounds__(256) k_cols_T(int H, int W, float2* __restrict__ specT,
                                                const float2* __restrict__ tw_g,
                                                const int* __restrict__ radix_g, int nrad,
                                                const float* __restrict__ kfT) {
  extern __shared__ float2 lds[];
  __shared__ int s_radix[32];
  float2* a = lds;
  float2* b = lds + H;
  float2* tw = lds + 2 * H;
  if (threadIdx.x < nrad) s_radix[threadIdx.x] = radix_g[threadIdx.x];
  load_twiddles(tw, tw_g, H);
  const int k = blockIdx.x;
  float2* col = specT + (size_t)k * H;
  for (int y = threadIdx.x; y < H; y += blockDim.x) a[y] = col[y];
  __syncthreads();
  float2* r = fft_lds(a, b, H, H, 1, s_radix, nrad, tw, false);
  float2* o = (r == a) ? b : a;
  const float* kc = kfT + (size_t)k * H;
  for (int y = threadIdx.x; y < H; y += blockDim.x) {
    const float kf = kc[y];
    float2 v = r[y];
    v.x = v.x * kf;
    v.y = v.y * kf;
    r[y] = v;
  }
  __syncthreads();
  float2* res = fft_lds(r, o, H, H, 1, s_radix, nrad, tw, true);
  const float sc = (float)(1.0 / (double)H);
  for (int y = threadIdx.x; y < H; y += blockDim.x) {
    const float2 v = res[y];
    col[y] = make_float2(v.x * sc, v.y * sc);
  }
}

// Inverse row FFT of TR rows read transposed from specT; real part scaled by 1/W.
template <int TR>
__global__ void __launch_bounds__(256) k_rows_inv_T(int H, int W, const float2* __restrict__ specT,
                                                    float* __restrict__ out,
                                                    const float2* __restrict__ tw_g,
                                                    const int* __restrict__ radix_g, int nrad) {
  extern __shared__ float2 lds[];
  __shared__ int s_radix[32];
  float2* a = lds;
  float2* b = lds + (size_t)TR * W;
  float2* tw = lds + (size_t)2 * TR * W;
  if (threadIdx.x < nrad) s_radix[threadIdx.x] = radix_g[threadIdx.x];
  load_twiddles(tw, tw_g, W);
  const int y0 = xcd_group(blockIdx.x, gridDim.x) * TR;
  const int nr = min(TR, H - y0);
  for (int i = threadIdx.x; i < nr * W; i += blockDim.x) {
    const int k = i / nr, r = i - k * nr;
    a[(size_t)r * W + k] = specT[(size_t)k * H + y0 + r];
  }
  __syncthreads();
  float2* res = fft_lds(a, b, W, W, nr, s_radix, nrad, tw, true);
  const float sc = (float)(1.0 / (double)W);
  for (int i = threadIdx.x; i < nr * W; i += blockDim.x) out[(size_t)y0 * W + i] = res[i].x * sc;
}

// Transposed transfer function kfT[k * H + y] = exp(coef (fr(y)^2 + fc(k)^2)).
__global__ void k_gauss_table_T(int H, int W, float coef, float* __restrict__ kfT) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (size_t)H * W) return;
  const int k = (int)(i / H), y = (int)(i - (size_t)k * H);
  const float fr = fftfreq(y, H), fc = fftfreq(k, W);
  kfT[i] = dexp(coef * (fr * fr + fc * fc));
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
                                   float* density_out, int* sort_scratch,
                                   unsigned long long* key_scratch, hipStream_t s) {
  const int H = c.dom_h, W = c.dom_w;
  const size_t lds = (size_t)3 * W * sizeof(float2);
  hipLaunchKernelGGL(k_deposit_rows_fft, dim3(H), dim3(256), lds, s, H, W, row_off, row_list,
                     sort_scratch, key_scratch, (size_t)4 * N, cic_cell, cic_w, spec, tw_w, radix_w,
                     nrad_w, density_out);
  return hipGetLastError();
}

hipError_t launch_fft_rows(int H, int W, const float* real_in, float2* spec, const float2* tw_w,
                           const int* radix_w, int nrad_w, hipStream_t s) {
  const size_t lds = (size_t)3 * W * sizeof(float2);
  hipLaunchKernelGGL(k_fft_rows, dim3(H), dim3(256), lds, s, H, W, real_in, spec, tw_w, radix_w,
                     nrad_w);
  return hipGetLastError();
}

hipError_t launch_gauss_table(int H, int W, float coef, float* kf, hipStream_t s) {
  const size_t n = (size_t)H * W;
  hipLaunchKernelGGL(k_gauss_table_T, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, H, W, coef,
                     kf);
  return hipGetLastError();
}

int fft_rows_per_group(int H, int W) {
  // One row per workgroup: the FFT passes are latency-bound, so parallelism
  // across rows beats longer transposed segments.  The XCD-grouped mapping
  // keeps the rows that share a specT cache line on one XCD, whose L2 merges
  // the 8-byte transposed writes before they reach HBM.
  (void)H;
  (void)W;
  return 1;
}

hipError_t launch_steric_fft_T(int H, int W, const float* real_in, const int* row_off, int* row_list,
                               int* sort_scratch, unsigned long long* key_scratch, size_t key_half,
                               const int* cic_cell, const float* cic_w, float2* specT,
                               const float2* tw_w, const int* radix_w, int nrad_w,
                               const float2* tw_h, const int* radix_h, int nrad_h, const float* kfT,
                               float* density_out, float* out, hipStream_t s) {
  const int TR = fft_rows_per_group(H, W);
  const int G = (H + TR - 1) / TR;
  const size_t lds_rows = ((size_t)2 * TR * W + W) * sizeof(float2);
  const size_t lds_cols = (size_t)3 * H * sizeof(float2);
#define MC_ROWS(T)                                                                                  \
  case T:                                                                                           \
    hipLaunchKernelGGL(k_rows_fwd_T<T>, dim3(G), dim3(256), lds_rows, s, H, W, real_in, row_off,   \
                       row_list, sort_scratch, key_scratch, key_half, cic_cell, cic_w, specT, tw_w, \
                       radix_w, nrad_w, density_out);                                               \
    break;
  switch (TR) { MC_ROWS(4) MC_ROWS(2) default: MC_ROWS(1) }
#undef MC_ROWS
  hipLaunchKernelGGL(k_cols_T, dim3(W), dim3(256), lds_cols, s, H, W, specT, tw_h, radix_h, nrad_h,
                     kfT);
#define MC_IROWS(T)                                                                                 \
  case T:                                                                                           \
    hipLaunchKernelGGL(k_rows_inv_T<T>, dim3(G), dim3(256), lds_rows, s, H, W, specT, out, tw_w,   \
                       radix_w, nrad_w);                                                            \
    break;
  switch (TR) { MC_IROWS(4) MC_IROWS(2) default: MC_IROWS(1) }
#undef MC_IROWS
  return hipGetLastError();
}

hipError_t launch_fft_cols_filter(int H, int W, float2* spec, const float2* tw_h,
                                  const int* radix_h, int nrad_h, const float* kf, int lds_bytes_max,
                                  hipStream_t s) {
  // widest strip that fits, but keep at least one workgroup per CU
  int cs = 16;
  while (cs > 1 && (((size_t)2 * cs * (H + 1) + H) * sizeof(float2) > (size_t)lds_bytes_max ||
                    (W + cs - 1) / cs < 256))
    cs >>= 1;
  const size_t lds = ((size_t)2 * cs * (H + 1) + H) * sizeof(float2);
  const int nstrips = (W + cs - 1) / cs;
  hipLaunchKernelGGL(k_fft_cols_filter, dim3(nstrips), dim3(256), lds, s, H, W, cs, spec, tw_h,
                     radix_h, nrad_h, kf);
  return hipGetLastError();
}

hipError_t launch_ifft_rows_real(int H, int W, const float2* spec, float* out, const float2* tw_w,
                                 const int* radix_w, int nrad_w, hipStream_t s) {
  const size_t lds = (size_t)3 * W * sizeof(float2);
  hipLaunchKernelGGL(k_ifft_rows_real, dim3(H), dim3(256), lds, s, H, W, spec, out, tw_w, radix_w,
                     nrad_w);
  return hipGetLastError();
}

}  // namespace mc
