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
#include <type_traits>

#include "mc_common.h"
#include "mc_predict.h"
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

// ---------------------------------------------------------------------------
// Compile-time Stockham networks for the common lengths.  The same radix
// schedule, twiddle values and operation order as fft_lds (so the same
// roundings), with every index known at compile time, the twiddles read per
// stage from a contiguous table, and the intermediate buffers padded by one
// element per 16 against LDS bank conflicts.  Input and result are unpadded.
// ---------------------------------------------------------------------------
constexpr int ct_radix(int n, int s) {
  int i = 0;
  while (n % 4 == 0) {
    if (i++ == s) return 4;
    n /= 4;
  }
  while (n % 2 == 0) {
    if (i++ == s) return 2;
    n /= 2;
  }
  for (int p = 3; p <= 13; p += 2)
    while (n % p == 0) {
      if (i++ == s) return p;
      n /= p;
    }
  return 0;
}
constexpr int ct_nstages(int n) {
  int c = 0;
  while (ct_radix(n, c) != 0) ++c;
  return c;
}
constexpr int ct_ns(int n, int s) {
  int ns = 1;
  for (int i = 0; i < s; ++i) ns *= ct_radix(n, i);
  return ns;
}
constexpr int ct_twoff(int n, int s) {
  int off = 0;
  for (int i = 0; i < s; ++i) {
    const int R = ct_radix(n, i);
    off += (R - 1) * ct_ns(n, i) + (R > 4 ? R : 0);
  }
  return off;
}
// Elements of a padded buffer of length n.
__host__ __device__ constexpr int ct_buf(int n) { return n + (n >> 4) + 1; }
template <bool P>
__device__ __forceinline__ int ct_pad(int i) { return P ? i + (i >> 4) : i; }

template <int R, bool INV>
__device__ __forceinline__ void ct_butterfly(float2 (&v)[R], float2 (&V)[R],
                                             const float2* __restrict__ roots) {
  if constexpr (R == 2 || R == 3 || R == 4) {
    butterfly<R>(v, V, INV, nullptr, 0);
  } else {
#pragma unroll
    for (int q = 0; q < R; ++q) {
      float2 acc = v[0];
#pragma unroll
      for (int r = 1; r < R; ++r) {
        float2 w = roots[(r * q) % R];
        if (INV) w.y = -w.y;
        acc = cadd(acc, cmul(v[r], w));
      }
      V[q] = acc;
    }
  }
}

template <int N, int S, bool INV>
__device__ __forceinline__ void ct_stage(const float2* __restrict__ src, float2* __restrict__ dst,
                                         const float2* __restrict__ stw) {
  constexpr int R = ct_radix(N, S), Ns = ct_ns(N, S), nb = N / R;
  constexpr bool pin = S > 0, pout = S + 1 < ct_nstages(N);
  const float2* tws = stw + ct_twoff(N, S);
  for (int j = threadIdx.x; j < nb; j += blockDim.x) {
    const int k = j % Ns;
    float2 v[R], V[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      float2 x = src[ct_pad<pin>(j + r * nb)];
      if (r > 0 && k > 0) {
        float2 w = tws[(r - 1) * Ns + k];
        if (INV) w.y = -w.y;
        x = cmul(x, w);
      }
      v[r] = x;
    }
    ct_butterfly<R, INV>(v, V, tws + (R - 1) * Ns);
    const int idxD = (j / Ns) * Ns * R + k;
#pragma unroll
    for (int r = 0; r < R; ++r) dst[ct_pad<pout>(idxD + r * Ns)] = V[r];
  }
}

template <int N, int S, bool INV>
__device__ __forceinline__ float2* ct_run(float2* src, float2* dst, const float2* stw) {
  if constexpr (S == ct_nstages(N)) {
    return src;
  } else {
    ct_stage<N, S, INV>(src, dst, stw);
    __syncthreads();
    return ct_run<N, S + 1, INV>(dst, src, stw);
  }
}

// Lengths with a compile-time network.
__device__ __forceinline__ bool ct_supported(int n) {
  switch (n) {
    case 64: case 128: case 192: case 256: case 384: case 512: case 768: case 1024: case 2048:
      return true;
    default:
      return false;
  }
}

// Twiddles of length n into LDS: the per-stage table for a compile-time
// network, else the plain table.
__device__ __forceinline__ void load_fft_twiddles(float2* dst, const float2* __restrict__ tw,
                                                  const FftDev& fp) {
  const bool ct = ct_supported(fp.n);
  const float2* src = tw + (ct ? fp.stw_off : fp.tw_off);
  const int len = ct ? fp.stw_len : fp.n;
  for (int i = threadIdx.x; i < len; i += blockDim.x) dst[i] = src[i];
}

// The same table in two halves: its first TW_REGS * blockDim elements loaded
// into registers (issued early, so the loads overlap other work), then
// stored to LDS together with the rest.
constexpr int TW_REGS = 4;
struct TwPrefetch {
  float2 v[TW_REGS];
};
__device__ __forceinline__ void twiddles_fetch(TwPrefetch& t, const float2* __restrict__ tw,
                                               const FftDev& fp) {
  const bool ct = ct_supported(fp.n);
  const float2* src = tw + (ct ? fp.stw_off : fp.tw_off);
  const int len = ct ? fp.stw_len : fp.n;
#pragma unroll
  for (int q = 0; q < TW_REGS; ++q) {
    const int i = threadIdx.x + q * blockDim.x;
    t.v[q] = i < len ? src[i] : make_float2(0.0f, 0.0f);
  }
}
__device__ __forceinline__ void twiddles_store(const TwPrefetch& t, float2* dst,
                                               const float2* __restrict__ tw, const FftDev& fp) {
  const bool ct = ct_supported(fp.n);
  const float2* src = tw + (ct ? fp.stw_off : fp.tw_off);
  const int len = ct ? fp.stw_len : fp.n;
#pragma unroll
  for (int q = 0; q < TW_REGS; ++q) {
    const int i = threadIdx.x + q * blockDim.x;
    if (i < len) dst[i] = t.v[q];
  }
  for (int i = threadIdx.x + TW_REGS * blockDim.x; i < len; i += blockDim.x) dst[i] = src[i];
}

// FFT of length n on a (unpadded; a and b hold ct_buf(n) elements each).
// Returns the buffer holding the unpadded result.
__device__ __forceinline__ float2* fft_any(float2* a, float2* b, int n, const int* radix, int nrad,
                           const float2* __restrict__ tw, bool inverse) {
#define MC_CT_CASE(NN)                                                              \
  case NN:                                                                          \
    return inverse ? ct_run<NN, 0, true>(a, b, tw) : ct_run<NN, 0, false>(a, b, tw);
  switch (n) {
    MC_CT_CASE(64)
    MC_CT_CASE(128)
    MC_CT_CASE(192)
    MC_CT_CASE(256)
    MC_CT_CASE(384)
    MC_CT_CASE(512)
    MC_CT_CASE(768)
    MC_CT_CASE(1024)
    MC_CT_CASE(2048)
    default:
      return fft_lds(a, b, n, n, 1, radix, nrad, tw, inverse);
  }
#undef MC_CT_CASE
}

// One wavefront per filament (see predict_component).  Block 0 also clears
// the deposit's overflow arena counter.
__global__ void __launch_bounds__(64) k_predict(Consts c, BatchDev bd, RodLayoutDev L,
                                                const float* __restrict__ pos,
                                                const float* __restrict__ vel,
                                                float* __restrict__ pred_w,
                                                int* __restrict__ cic_cell,
                                                float* __restrict__ cic_w,
                                                int4* __restrict__ comp_rows,
                                                int2* __restrict__ comp_cols,
                                                int* __restrict__ arena_top) {
  if (blockIdx.x == 0 && threadIdx.x == 0) arena_top[0] = arena_top[1] = 0;
  PredictArgs a{c, bd, L, pos, vel, pred_w, cic_cell, cic_w, comp_rows, comp_cols, arena_top};
  predict_component(a, blockIdx.x, threadIdx.x & 63);
}

__device__ __forceinline__ int xcd_group(int b, int G) {
  // workgroup i runs on XCD i % 8: give each XCD a contiguous range of groups
  return (G & 7) == 0 ? (b & 7) * (G >> 3) + (b >> 3) : b;
}

// CIC deposit of global steric row `row` into dst[0..W): each node whose
// CIC rows include the row contributes two entries (corners q = 0, 1 if the
// row is its y0 row, q = 2, 3 if its y1 row).  Ordering the entries by (x, q,
// node rank) puts each cell's contributions contiguously and in the
// reference's order (all w00 by node, then w10, w01, w11 --
// forces.py:112-115); a segment head thread folds them from zero.  Rows of up
// to `cap` nodes keep each node's id, cells and weights in LDS (one global
// round trip) and order node ids and entry keys by counting ranks in
// `scratch` (DEP_BYTES_PER_NODE bytes of LDS per node, aliasing the FFT's
// second buffer); longer rows take a slice of the global arena and
// merge-sort in global scratch.
constexpr int DEP_BYTES_PER_NODE = 52;  // rank, 2 entry keys, 2 ranked keys, cells, weights
constexpr int DEP_CAP_MAX = 256;
constexpr int DEP_CAP_MIN = 64;

struct DepArgs {
  RodLayoutDev L;
  const int4* comp_rows;        // per component: band (first row, rows), node count, first id
                                // (contiguous ids) or ~(node-list offset)
  const int2* comp_cols;        // per component: column band (first column, columns)
  const int* cic_cell;
  const float* cic_w;
  int* arena;                   // 2N ids for rows beyond cap
  int* arena_top;               // arena slots taken this step
  int* sort_tmp;                // 2N ints
  unsigned long long* key_tmp;  // 8N keys: entries at 2 aoff, merge scratch at key_half + 2 aoff
  size_t key_half;
  int cap;                      // LDS capacity in nodes (a multiple of 4, 16 cap >= 8 blockDim)
  int b_off, tw_off;            // LDS offsets (float2) of the scratch / second buffer and twiddles
  DepEntry* tile_arena;         // windows of the fused deposit beyond the LDS (2 x 16N entries)
  int* tile_top;                // tile_arena entries taken this step
};

// Ids of the nodes whose CIC rows include global row `row`, each once: the
// instance's components whose band holds the row are compacted per chunk of
// blockDim components (one workgroup scan), then their nodes are checked.
// Match p goes to out[p] when p < lim, with (when rx is not NULL) its cells
// rx[p] = (x0 | role << 30, x1), role 0 for the y0 row, and the role's two
// weights rw[p].  `cand` holds 2 blockDim ints of LDS, `wsum` 32 ints.
// Returns the number of matches (the same in every thread).
__device__ int collect_row_nodes(const DepArgs& a, const InstGeo& gi, int row, int* out, int lim,
                                 int2* rx, float2* rw, int* cand, int* wsum, int* s_cnt) {
  const int B = blockDim.x, tid = threadIdx.x, H = gi.H, y = row - gi.row0;
  int* c_first = cand;    // first node id (contiguous ids) or ~(node-list offset)
  int* c_off = cand + B;  // prefix of node counts over the chunk's candidates
  if (tid == 0) *s_cnt = 0;
  __syncthreads();
  for (int base = 0; base < gi.ncomp; base += B) {
    int is = 0, nn = 0, first = 0;
    if (base + tid < gi.ncomp) {
      const int4 r = a.comp_rows[gi.comp0 + base + tid];
      if (imod(y - r.x, H) < r.y) {
        is = 1;
        nn = r.z;
        first = r.w;
      }
    }
    int2 tot;
    const int2 inc = block_inclusive_scan2(make_int2(is, nn), wsum, &tot);
    const int ccount = tot.x, ntot = tot.y;
    if (is) {
      c_first[inc.x - 1] = first;
      c_off[inc.x - 1] = inc.y - nn;
    }
    __syncthreads();
    // CR candidates per thread at a time: their node loads are in flight together
    constexpr int CR = 4;
    for (int f0 = tid; f0 < ntot; f0 += CR * B) {
      int nid[CR];
#pragma unroll
      for (int u = 0; u < CR; ++u) {
        const int f = f0 + u * B;
        nid[u] = -1;
        if (f < ntot) {
          int lo = 0, hi = ccount - 1;  // the last candidate with c_off <= f
          while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (c_off[mid] <= f) lo = mid; else hi = mid - 1;
          }
          const int local = f - c_off[lo], fc = c_first[lo];
          nid[u] = fc >= 0 ? fc + local : a.L.comp_nodes[~fc + local];
        }
      }
      int4 cc[CR];
      float4 ww[CR];
#pragma unroll
      for (int u = 0; u < CR; ++u) {
        const size_t q = 4 * (size_t)(nid[u] < 0 ? 0 : nid[u]);
        cc[u] = nid[u] < 0 ? make_int4(-1, -1, -1, -1)
                           : *reinterpret_cast<const int4*>(a.cic_cell + q);
        ww[u] = nid[u] < 0 ? make_float4(0.0f, 0.0f, 0.0f, 0.0f)
                           : *reinterpret_cast<const float4*>(a.cic_w + q);
      }
#pragma unroll
      for (int u = 0; u < CR; ++u) {
        if (nid[u] < 0 || !(cc[u].z == row || cc[u].w == row)) continue;
        const int p = atomicAdd(s_cnt, 1);
        if (p < lim) {
          out[p] = nid[u];
          if (rx) {
            const int role = cc[u].z == row ? 0 : 1;
            rx[p] = make_int2(cc[u].x | (role << 30), cc[u].y);
            rw[p] = role ? make_float2(ww[u].z, ww[u].w) : make_float2(ww[u].x, ww[u].y);
          }
        }
      }
    }
    __threadfence_block();
    __syncthreads();
  }
  return *s_cnt;
}

__device__ void deposit_row(const DepArgs& a, const InstGeo& gi, int row, float2* __restrict__ dst,
                            char* __restrict__ scratch, int* wsum, int* s_cnt) {
  const int B = blockDim.x, tid = threadIdx.x, W = gi.W, cap = a.cap;
  for (int x = tid; x < W; x += B) dst[x] = make_float2(0.0f, 0.0f);
  int* sorted = reinterpret_cast<int*>(scratch);                                 // [cap]
  unsigned long long* k0 = reinterpret_cast<unsigned long long*>(sorted + cap);  // [2 cap]
  unsigned long long* k1 = k0 + 2 * cap;                                         // [2 cap]
  int2* rx = reinterpret_cast<int2*>(k1 + 2 * cap);                              // [cap]
  float2* rw = reinterpret_cast<float2*>(rx + cap);                              // [cap]
  int* ids = reinterpret_cast<int*>(k0);   // arrival order, dead once ranked
  int* cand = reinterpret_cast<int*>(k1);  // candidate components (2 B ints), dead once collected
  const int m = collect_row_nodes(a, gi, row, ids, cap, rx, rw, cand, wsum, s_cnt);
  if (m == 0) return;
  const int ne = 2 * m;
  if (m <= cap) {
    // sorted[rank] = arrival index of the node of that rank
    for (int j = tid; j < m; j += B) {
      const int v = ids[j];
      int r = 0;
      for (int k = 0; k < m; ++k) r += ids[k] < v ? 1 : 0;
      sorted[r] = j;
    }
    __syncthreads();
    for (int j = tid; j < m; j += B) {
      const int2 c = rx[sorted[j]];
      const unsigned long long qb = (unsigned long long)(c.x >> 30) << 1;
      const unsigned long long x0 = (unsigned)(c.x & 0x3fffffff), x1 = (unsigned)c.y;
      k0[2 * j] = (x0 << 34) | (qb << 32) | (unsigned)j;
      k0[2 * j + 1] = (x1 << 34) | ((qb + 1ull) << 32) | (unsigned)j;
    }
    __syncthreads();
    for (int i = tid; i < ne; i += B) {
      const unsigned long long v = k0[i];
      int r = 0;
      for (int k = 0; k < ne; ++k) r += k0[k] < v ? 1 : 0;
      k1[r] = v;
    }
    __syncthreads();
    for (int i = tid; i < ne; i += B) {
      const unsigned long long k = k1[i];
      const int x = (int)(k >> 34);
      if (i > 0 && (int)(k1[i - 1] >> 34) == x) continue;
      float acc = 0.0f;
      for (int e = i; e < ne; ++e) {
        const unsigned long long ke = k1[e];
        if ((int)(ke >> 34) != x) break;
        const float2 w2 = rw[sorted[(int)(ke & 0xffffffffull)]];
        acc = acc + (((ke >> 32) & 1ull) ? w2.y : w2.x);
      }
      dst[x].x = acc;
    }
    __syncthreads();
    return;
  }
  // rows beyond cap: ids in a slice of the arena, sorts and keys in global scratch
  if (tid == 0) s_cnt[1] = atomicAdd(a.arena_top, m);
  __syncthreads();
  const int aoff = s_cnt[1];
  int* nl = a.arena + aoff;
  collect_row_nodes(a, gi, row, nl, m, nullptr, nullptr, cand, wsum, s_cnt);
  sort_unique<int>(nl, m, nullptr, 0, a.sort_tmp + aoff, 0x7fffffff);
  unsigned long long* kk = a.key_tmp + 2 * (size_t)aoff;
  for (int j = tid; j < m; j += B) {
    const int n = nl[j];
    const int4 cc = *reinterpret_cast<const int4*>(a.cic_cell + 4 * (size_t)n);
    const unsigned long long qb = (cc.z == row) ? 0ull : 2ull;
    kk[2 * j] = ((unsigned long long)(unsigned)cc.x << 34) | (qb << 32) | (unsigned)j;
    kk[2 * j + 1] = ((unsigned long long)(unsigned)cc.y << 34) | ((qb + 1ull) << 32) | (unsigned)j;
  }
  __threadfence_block();
  __syncthreads();
  sort_unique<unsigned long long>(kk, ne, nullptr, 0, a.key_tmp + a.key_half + 2 * (size_t)aoff,
                                  ~0ull);
  for (int i = tid; i < ne; i += B) {
    const unsigned long long k = kk[i];
    const int x = (int)(k >> 34);
    if (i > 0 && (int)(kk[i - 1] >> 34) == x) continue;
    float acc = 0.0f;
    for (int e = i; e < ne; ++e) {
      const unsigned long long ke = kk[e];
      if ((int)(ke >> 34) != x) break;
      const int q = (int)((ke >> 32) & 3ull);
      const int n = nl[(int)(ke & 0xffffffffull)];
      acc = acc + a.cic_w[4 * n + q];
    }
    dst[x].x = acc;
  }
  __syncthreads();
}

__device__ __forceinline__ float fftfreq(int k, int n) {
  // numpy / jnp.fft.fftfreq in float32: k / n for k < (n + 1) / 2, else (k - n) / n
  const int kk = k < (n + 1) / 2 ? k : k - n;
  return (float)kk / (float)n;
}

__device__ __forceinline__ void load_plan(const FftDev& p, int* s_radix) {
  if (threadIdx.x < p.nrad) s_radix[threadIdx.x] = p.radix[threadIdx.x];
}

// ---------------------------------------------------------------------------
// 2-D FFT over every instance, exchanging the spectrum between the row and
// column passes in transposed layout specT[k * H + y] (per instance, at the
// instance's grid offset): the column pass then reads whole contiguous rows.
// One global steric row per workgroup in the row passes, one global column per
// workgroup in the column pass.
// ---------------------------------------------------------------------------

// The shape's table holds the direct filter's taps (its first word, lo_x <= 0
// as an int) rather than a transfer function (first word exp(0) = 1.0f).
__device__ __forceinline__ bool filter_direct(const float* table, const InstGeo& gi) {
  return reinterpret_cast<const int*>(table + gi.kf_off)[0] <= 0;
}

// Forward row FFT of one row (from the CIC deposit when `real_in` is NULL,
// else from the real grid), written transposed to specT.
__global__ void __launch_bounds__(256) k_rows_fwd_T(BatchDev bd, const float* __restrict__ real_in,
                                                    DepArgs dep, float2* __restrict__ specT,
                                                    const float* __restrict__ kfT,
                                                    float* __restrict__ density_out) {
  extern __shared__ float2 lds[];
  __shared__ int s_radix[FFT_MAX_RADICES];
  __shared__ int wsum[32];
  __shared__ int s_cnt[2];
  const int row = xcd_group(blockIdx.x, gridDim.x);
  const InstGeo& gi = bd.geo[bd.row_inst[row]];
  if (filter_direct(kfT, gi)) return;
  const int W = gi.W, H = gi.H, y = row - gi.row0;
  const FftDev& fp = bd.fft[gi.fft_w];
  float2* a = lds;
  float2* b = lds + dep.b_off;    // the second FFT buffer aliases the deposit scratch
  float2* tw = lds + dep.tw_off;
  load_plan(fp, s_radix);
  TwPrefetch twp;
  twiddles_fetch(twp, bd.tw, fp);  // in flight during the deposit
  const size_t g0 = (size_t)gi.grid_off + (size_t)y * W;
  if (real_in) {
    for (int i = threadIdx.x; i < W; i += blockDim.x) a[i] = make_float2(real_in[g0 + i], 0.0f);
  } else {
    deposit_row(dep, gi, row, a, reinterpret_cast<char*>(b), wsum, s_cnt);
    if (density_out)
      for (int i = threadIdx.x; i < W; i += blockDim.x) density_out[g0 + i] = a[i].x;
  }
  twiddles_store(twp, tw, bd.tw, fp);
  __syncthreads();
  float2* res = fft_any(a, b, W, s_radix, fp.nrad, tw, false);
  float2* sp = specT + gi.grid_off;
  for (int k = threadIdx.x; k < W; k += blockDim.x) sp[(size_t)k * H + y] = res[k];
}

// Column pass on the transposed spectrum: forward FFT along y, transfer
// function, inverse FFT scaled by 1/H.
__global__ void __launch_bounds__(256) k_cols_T(BatchDev bd, float2* __restrict__ specT,
                                                const float* __restrict__ kfT) {
  extern __shared__ float2 lds[];
  __shared__ int s_radix[FFT_MAX_RADICES];
  const int colg = blockIdx.x;
  const InstGeo& gi = bd.geo[bd.col_inst[colg]];
  if (filter_direct(kfT, gi)) return;
  const int H = gi.H, k = colg - gi.col0;
  const FftDev& fp = bd.fft[gi.fft_h];
  float2* a = lds;
  float2* b = lds + ct_buf(H);
  float2* tw = lds + 2 * ct_buf(H);
  load_plan(fp, s_radix);
  load_fft_twiddles(tw, bd.tw, fp);
  float2* col = specT + gi.grid_off + (size_t)k * H;
  for (int y = threadIdx.x; y < H; y += blockDim.x) a[y] = col[y];
  __syncthreads();
  float2* r = fft_any(a, b, H, s_radix, fp.nrad, tw, false);
  float2* o = (r == a) ? b : a;
  const float* kc = kfT + gi.kf_off + (size_t)k * H;
  for (int y = threadIdx.x; y < H; y += blockDim.x) {
    const float kf = kc[y];
    float2 v = r[y];
    v.x = v.x * kf;
    v.y = v.y * kf;
    r[y] = v;
  }
  __syncthreads();
  float2* res = fft_any(r, o, H, s_radix, fp.nrad, tw, true);
  const float sc = (float)(1.0 / (double)H);
  for (int y = threadIdx.x; y < H; y += blockDim.x) {
    const float2 v = res[y];
    col[y] = make_float2(v.x * sc, v.y * sc);
  }
}

// Inverse row FFT of one row read transposed from specT; real part scaled by 1/W.
__global__ void __launch_bounds__(256) k_rows_inv_T(BatchDev bd, const float2* __restrict__ specT,
                                                    const float* __restrict__ kfT,
                                                    float* __restrict__ out) {
  extern __shared__ float2 lds[];
  __shared__ int s_radix[FFT_MAX_RADICES];
  const int row = xcd_group(blockIdx.x, gridDim.x);
  const InstGeo& gi = bd.geo[bd.row_inst[row]];
  if (filter_direct(kfT, gi)) return;
  const int W = gi.W, H = gi.H, y = row - gi.row0;
  const FftDev& fp = bd.fft[gi.fft_w];
  float2* a = lds;
  float2* b = lds + ct_buf(W);
  float2* tw = lds + 2 * ct_buf(W);
  load_plan(fp, s_radix);
  load_fft_twiddles(tw, bd.tw, fp);
  const float2* sp = specT + gi.grid_off;
  for (int k = threadIdx.x; k < W; k += blockDim.x) a[k] = sp[(size_t)k * H + y];
  __syncthreads();
  float2* res = fft_any(a, b, W, s_radix, fp.nrad, tw, true);
  const float sc = (float)(1.0 / (double)W);
  float* o = out + gi.grid_off + (size_t)y * W;
  for (int i = threadIdx.x; i < W; i += blockDim.x) o[i] = res[i].x * sc;
}

// ---------------------------------------------------------------------------
// Direct filter.  The transfer function exp(-2 pi^2 s^2 (fr^2 + fc^2)) is the
// product of one factor per axis, so utils.py:35-49's filter is the circular
// convolution with h_y (x) h_x, h = the inverse DFT of one factor.  When both
// kernels are short (their tails below 2^-27 of their mass), the filter runs
// as a row pass (fused with the row's CIC deposit) and a column pass, each
// output summed tap by tap from the first offset; the shape's table then
// holds (int lo_x, n_x, lo_y, n_y) and the taps of x, then of y.
// ---------------------------------------------------------------------------

// Row pass: the row (deposited, or read from real_in) in LDS, convolved with
// h_x, to tmp; the deposit also goes to density_out.
__global__ void __launch_bounds__(256) k_rows_conv(BatchDev bd, const float* __restrict__ real_in,
                                                   DepArgs dep, const float* __restrict__ taps,
                                                   float* __restrict__ tmp,
                                                   float* __restrict__ density_out,
                                                   int filter_rows) {
  extern __shared__ float2 lds[];
  __shared__ int wsum[32];
  __shared__ int s_cnt[2];
  const int row = xcd_group(blockIdx.x, gridDim.x);
  const InstGeo& gi = bd.geo[bd.row_inst[row]];
  if (!filter_direct(taps, gi)) return;
  const int W = gi.W, y = row - gi.row0;
  float2* a = lds;
  float2* b = lds + dep.b_off;
  float* tw = reinterpret_cast<float*>(lds + dep.tw_off);
  const int* hd = reinterpret_cast<const int*>(taps + gi.kf_off);
  const int lo = hd[0], nt = hd[1];
  const float* tx = taps + gi.kf_off + 4;
  for (int t = threadIdx.x; t < nt; t += blockDim.x) tw[t] = tx[t];
  const size_t g0 = (size_t)gi.grid_off + (size_t)y * W;
  if (real_in) {
    for (int i = threadIdx.x; i < W; i += blockDim.x) a[i] = make_float2(real_in[g0 + i], 0.0f);
  } else {
    deposit_row(dep, gi, row, a, reinterpret_cast<char*>(b), wsum, s_cnt);
    if (density_out)
      for (int i = threadIdx.x; i < W; i += blockDim.x) density_out[g0 + i] = a[i].x;
  }
  if (!filter_rows) return;  // the deposit only: the 2-D pass filters it
  __syncthreads();
  // four consecutive outputs per thread share one input window; taps in
  // chunks of eight whose loads are issued together, each output summed tap
  // by tap in order.  A row at least nt + 11 wide wraps a window index at most
  // once (a select); narrower rows take the remainder.
  auto row_pass = [&](auto wide_row) {
    constexpr bool WIDE = decltype(wide_row)::value;
    for (int x0 = 4 * threadIdx.x; x0 < W; x0 += 4 * blockDim.x) {
      float acc[4] = {0.0f, 0.0f, 0.0f, 0.0f};
      int i0 = x0 + lo;  // >= -W
      if (i0 < 0) i0 += W;
      for (int t0 = 0; t0 < nt; t0 += 8) {
        float win[11], tv[8];
        int ib = i0 + t0;
        if (WIDE) ib = ib >= W ? ib - W : ib;
#pragma unroll
        for (int m = 0; m < 11; ++m) {
          int i = ib + m;
          if (WIDE) i = i >= W ? i - W : i;
          else i = i % W;
          win[m] = a[i].x;
        }
#pragma unroll
        for (int j = 0; j < 8; ++j) tv[j] = tw[t0 + j < nt ? t0 + j : 0];
#pragma unroll
        for (int j = 0; j < 8; ++j)
          if (t0 + j < nt)
#pragma unroll
            for (int k = 0; k < 4; ++k) acc[k] = acc[k] + tv[j] * win[k + j];
      }
#pragma unroll
      for (int k = 0; k < 4; ++k)
        if (x0 + k < W) tmp[g0 + x0 + k] = acc[k];
    }
  };
  if (W >= nt + 11) row_pass(std::true_type{});
  else row_pass(std::false_type{});
}

// Column pass by tiles: a workgroup per FILTER_TILE_W x FILTER_TILE_H output
// tile of one instance (a tile table maps workgroups to instances); the
// tile's input rows (its rows plus the kernel's reach) are loaded into LDS
// once, then each thread sums FILTER_TILE_H / 4 outputs of its column tap by
// tap in order from the first offset; an output costs 1 + (n - 1) / 32 global
// loads instead of n.
constexpr int TILE_IN_MAX = FILTER_TILE_H + FILTER_TAPS_MAX - 1;
__global__ void __launch_bounds__(256) k_cols_conv_tiled(BatchDev bd, const float* __restrict__ taps,
                                                         const float* __restrict__ tmp,
                                                         float* __restrict__ out) {
  __shared__ float ty[FILTER_TAPS_MAX];
  __shared__ float tile[TILE_IN_MAX * FILTER_TILE_W];
  const int b = xcd_group(blockIdx.x, gridDim.x);
  const int inst = find_inst(bd.ctile, bd.n_inst, b);
  const InstGeo& gi = bd.geo[inst];
  if (!filter_direct(taps, gi)) return;
  const int W = gi.W, H = gi.H;
  const int txn = (W + FILTER_TILE_W - 1) / FILTER_TILE_W;
  const int local = b - bd.ctile[inst];
  const int ty0 = (local / txn) * FILTER_TILE_H, tx0 = (local % txn) * FILTER_TILE_W;
  const int* hd = reinterpret_cast<const int*>(taps + gi.kf_off);
  const int nx = hd[1], lo = hd[2], nt = hd[3];
  const int nin = FILTER_TILE_H + nt - 1;
  for (int t = threadIdx.x; t < nt; t += blockDim.x) ty[t] = taps[gi.kf_off + 4 + nx + t];
  const int c = threadIdx.x & (FILTER_TILE_W - 1), rq = threadIdx.x / FILTER_TILE_W;
  const int x = tx0 + c;
  const float* src = tmp + gi.grid_off;
  int yb = ty0 + lo;  // > -H: lo > -H, ty0 >= 0
  if (yb < 0) yb += H;
  // the tile's input rows, 8 loads in flight per thread before their stores
  constexpr int RS = 256 / FILTER_TILE_W, LB = 8;
  const bool tall = H >= nin;  // a tile row wraps at most once
  for (int r0 = rq; r0 < nin; r0 += RS * LB) {
    float v[LB];
#pragma unroll
    for (int k = 0; k < LB; ++k) {
      const int r = r0 + k * RS;
      int yy = yb + r;
      yy = tall ? (yy >= H ? yy - H : yy) : yy % H;
      v[k] = (r < nin && x < W) ? src[(size_t)yy * W + x] : 0.0f;
    }
#pragma unroll
    for (int k = 0; k < LB; ++k)
      if (r0 + k * RS < nin) tile[(r0 + k * RS) * FILTER_TILE_W + c] = v[k];
  }
  __syncthreads();
  // ROWS consecutive outputs per thread share one input window; taps in
  // chunks of eight whose loads are issued together, each output summed tap
  // by tap in order
  constexpr int ROWS = FILTER_TILE_H / (256 / FILTER_TILE_W);
  const int orow0 = rq * ROWS;
  float acc[ROWS];
#pragma unroll
  for (int k = 0; k < ROWS; ++k) acc[k] = 0.0f;
  for (int t0 = 0; t0 < nt; t0 += 8) {
    float win[ROWS + 7], tv[8];
#pragma unroll
    for (int m = 0; m < ROWS + 7; ++m) {
      const int r = orow0 + t0 + m;
      win[m] = r < nin ? tile[r * FILTER_TILE_W + c] : 0.0f;
    }
#pragma unroll
    for (int j = 0; j < 8; ++j) tv[j] = ty[t0 + j < nt ? t0 + j : 0];
#pragma unroll
    for (int j = 0; j < 8; ++j)
      if (t0 + j < nt)
#pragma unroll
        for (int k = 0; k < ROWS; ++k) acc[k] = acc[k] + tv[j] * win[k + j];
  }
#pragma unroll
  for (int k = 0; k < ROWS; ++k) {
    const int y = ty0 + orow0 + k;
    if (x < W && y < H) out[gi.grid_off + (size_t)y * W + x] = acc[k];
  }
}

// Both passes in one launch, for kernels of at most F2_TAPS taps per axis: a
// workgroup per FILTER_TILE_W x FILTER_TILE_H output tile (the column pass's
// tile table) stages the tile's input window (its rows and columns plus the
// kernels' reach, wrapped) in LDS, filters the window's rows into a second
// LDS array, then its columns, each output summed tap by tap in order from
// the first offset exactly as the row and column launches do, so the result
// is the same bit for bit without the grid-sized intermediate.
constexpr int F2_TAPS = 33;
constexpr int F2_THREADS = 512;
constexpr int F2_RIN = FILTER_TILE_H + F2_TAPS - 1;          // input rows at most
constexpr int F2_CS = FILTER_TILE_W + F2_TAPS - 1 + 4;       // input row stride (floats, 16 B)
// The two passes of the 2-D filter over a staged window (`in`, F2_CS floats
// per row, rin rows): x into `rt`, then y into the tile's outputs.  Each
// output sums its taps in order from the first offset, as the row and column
// launches do.
__device__ __forceinline__ void filter2d_passes(const float* __restrict__ in, float* __restrict__ rt,
                                                const float* __restrict__ tx,
                                                const float* __restrict__ ty, int nx, int ny,
                                                int rin, const InstGeo& gi, int ty0, int tx0,
                                                float* __restrict__ out) {
  const int W = gi.W, H = gi.H, tid = threadIdx.x;
  // rows: four consecutive outputs per thread from one window, taps in
  // chunks of eight (three 16-byte loads per chunk)
  {
    const int x0 = 4 * (tid & 15);
    for (int r = tid >> 4; r < rin; r += F2_THREADS / 16) {
      float acc[4] = {0.0f, 0.0f, 0.0f, 0.0f};
      const float* rowp = in + r * F2_CS + x0;
#pragma unroll
      for (int t0 = 0; t0 < F2_TAPS; t0 += 8) {
        if (t0 >= nx) break;
        float win[12], tv[8];
        const float4 w0 = *reinterpret_cast<const float4*>(rowp + t0);
        const float4 w1 = *reinterpret_cast<const float4*>(rowp + t0 + 4);
        const float4 w2 = *reinterpret_cast<const float4*>(rowp + t0 + 8);
        win[0] = w0.x; win[1] = w0.y; win[2] = w0.z; win[3] = w0.w;
        win[4] = w1.x; win[5] = w1.y; win[6] = w1.z; win[7] = w1.w;
        win[8] = w2.x; win[9] = w2.y; win[10] = w2.z; win[11] = w2.w;
#pragma unroll
        for (int j = 0; j < 8; ++j) tv[j] = tx[t0 + j < nx ? t0 + j : 0];
#pragma unroll
        for (int j = 0; j < 8; ++j)
          if (t0 + j < nx)
#pragma unroll
            for (int k = 0; k < 4; ++k) acc[k] = acc[k] + tv[j] * win[k + j];
      }
      *reinterpret_cast<float4*>(rt + r * FILTER_TILE_W + x0) =
          make_float4(acc[0], acc[1], acc[2], acc[3]);
    }
  }
  __syncthreads();
  // columns: ROWS consecutive outputs per thread from one window
  constexpr int ROWS = FILTER_TILE_H / (F2_THREADS / FILTER_TILE_W);
  const int c = tid & (FILTER_TILE_W - 1), orow0 = (tid / FILTER_TILE_W) * ROWS;
  float acc[ROWS];
#pragma unroll
  for (int k = 0; k < ROWS; ++k) acc[k] = 0.0f;
#pragma unroll
  for (int t0 = 0; t0 < F2_TAPS; t0 += 8) {
    if (t0 >= ny) break;
    float win[ROWS + 7], tv[8];
#pragma unroll
    for (int m = 0; m < ROWS + 7; ++m) {
      const int r = orow0 + t0 + m;
      win[m] = r < rin ? rt[r * FILTER_TILE_W + c] : 0.0f;
    }
#pragma unroll
    for (int j = 0; j < 8; ++j) tv[j] = ty[t0 + j < ny ? t0 + j : 0];
#pragma unroll
    for (int j = 0; j < 8; ++j)
      if (t0 + j < ny)
#pragma unroll
        for (int k = 0; k < ROWS; ++k) acc[k] = acc[k] + tv[j] * win[k + j];
  }
  const int x = tx0 + c;
#pragma unroll
  for (int k = 0; k < ROWS; ++k) {
    const int y = ty0 + orow0 + k;
    if (x < W && y < H) out[gi.grid_off + (size_t)y * W + x] = acc[k];
  }
}

__global__ void __launch_bounds__(F2_THREADS) k_filter2d(BatchDev bd, const float* __restrict__ taps,
                                                         const float* __restrict__ src,
                                                         float* __restrict__ out) {
  __shared__ __align__(16) float in[F2_RIN * F2_CS];
  __shared__ __align__(16) float rt[F2_RIN * FILTER_TILE_W];
  __shared__ int rowy[F2_RIN], colx[F2_CS];
  const int b = xcd_group(blockIdx.x, gridDim.x);
  const int inst = find_inst(bd.ctile, bd.n_inst, b);
  const InstGeo& gi = bd.geo[inst];
  if (!filter_direct(taps, gi)) return;
  const int W = gi.W, H = gi.H, tid = threadIdx.x;
  const int txn = (W + FILTER_TILE_W - 1) / FILTER_TILE_W;
  const int local = b - bd.ctile[inst];
  const int ty0 = (local / txn) * FILTER_TILE_H, tx0 = (local % txn) * FILTER_TILE_W;
  const int* hd = reinterpret_cast<const int*>(taps + gi.kf_off);
  const int lox = hd[0], nx = hd[1], loy = hd[2], ny = hd[3];
  // the taps are read with uniform addresses (scalar loads) in the passes
  const float* tx = taps + gi.kf_off + 4;
  const float* ty = tx + nx;
  const int rin = FILTER_TILE_H + ny - 1, cin = FILTER_TILE_W + nx - 1;
  int yb = (ty0 + loy) % H, xb = (tx0 + lox) % W;  // lo > -side
  if (yb < 0) yb += H;
  if (xb < 0) xb += W;
  const float* g = src + gi.grid_off;
  // the input window, every load of a thread in flight at once: the wrapped
  // row and column indices tabulated first, then element tid + F2_THREADS k
  if (tid < rin) rowy[tid] = (yb + tid) % H;
  if (tid < cin) colx[tid] = (xb + tid) % W;
  __syncthreads();
  {
    constexpr int PER = (F2_RIN * (F2_CS - 4) + F2_THREADS - 1) / F2_THREADS;
    const int dr = F2_THREADS / cin, dc = F2_THREADS - dr * cin;
    int r = tid / cin, c = tid - r * cin;
    float v[PER];
    int dst[PER];
#pragma unroll
    for (int k = 0; k < PER; ++k) {
      const bool ok = r < rin;
      v[k] = ok ? g[(size_t)rowy[ok ? r : 0] * W + colx[c]] : 0.0f;
      dst[k] = ok ? r * F2_CS + c : -1;
      c += dc;
      r += dr;
      if (c >= cin) {
        c -= cin;
        ++r;
      }
    }
#pragma unroll
    for (int k = 0; k < PER; ++k)
      if (dst[k] >= 0) in[dst[k]] = v[k];
  }
  __syncthreads();
  filter2d_passes(in, rt, tx, ty, nx, ny, rin, gi, ty0, tx0, out);
}

// The CIC deposit (forces.py:100-115) and the 2-D filter in one launch, for
// instances sparse enough that a tile's window holds few nodes (the launch
// choice is the plan's).  A workgroup per FILTER_TILE_W x FILTER_TILE_H output
// tile deposits its input window (the tile's rows and columns plus the
// kernels' reach, wrapped; each grid cell appears once in it) straight into
// LDS:
//   1. the instance's components whose row and column bands meet the window
//      are compacted; their nodes' CIC corners inside the window become
//      entries (window cell, corner, node id, weight), counted per window
//      cell;
//   2. an exclusive scan of the counts and one placement pass group the
//      entries by cell;
//   3. the thread holding a cell's first entry orders the cell's entries by
//      (corner, node id) -- the reference's scatter order, all w00 by node,
//      then w10, w01, w11 -- and folds them from zero, exactly as the row
//      deposit folds them.
// A window with more entries than the LDS holds redoes steps 1-3 in a slice
// of tile_arena.  The density is then filtered as k_filter2d does, bit for bit
// the same as the row deposit followed by k_filter2d.
constexpr int DF_CELLS = F2_RIN * (FILTER_TILE_W + F2_TAPS - 1);
constexpr int DF_ENT = (F2_RIN * FILTER_TILE_W * 4) / (int)sizeof(DepEntry);
__global__ void __launch_bounds__(F2_THREADS) k_dep_filter2d(BatchDev bd,
                                                             const float* __restrict__ taps,
                                                             DepArgs dep,
                                                             float* __restrict__ density_out,
                                                             float* __restrict__ out) {
  __shared__ __align__(16) float in[F2_RIN * F2_CS];           // the arrival entries first
  __shared__ __align__(16) float rt[F2_RIN * FILTER_TILE_W];   // the grouped entries first
  __shared__ int cnt[DF_CELLS];
  __shared__ int cand[2 * F2_THREADS];
  __shared__ int wsum[32];
  __shared__ int s_m, s_base;
  const int b = xcd_group(blockIdx.x, gridDim.x);
  const int inst = find_inst(bd.ctile, bd.n_inst, b);
  const InstGeo& gi = bd.geo[inst];
  if (!filter_direct(taps, gi)) return;
  const int W = gi.W, H = gi.H, tid = threadIdx.x, bs = blockDim.x;
  const int txn = (W + FILTER_TILE_W - 1) / FILTER_TILE_W;
  const int local = b - bd.ctile[inst];
  const int ty0 = (local / txn) * FILTER_TILE_H, tx0 = (local % txn) * FILTER_TILE_W;
  const int* hd = reinterpret_cast<const int*>(taps + gi.kf_off);
  const int lox = hd[0], nx = hd[1], loy = hd[2], ny = hd[3];
  const float* tx = taps + gi.kf_off + 4;
  const float* ty = tx + nx;
  const int rin = FILTER_TILE_H + ny - 1, cin = FILTER_TILE_W + nx - 1, ncell = rin * cin;
  int yb = (ty0 + loy) % H, xb = (tx0 + lox) % W;  // lo > -side
  if (yb < 0) yb += H;
  if (xb < 0) xb += W;
  for (int i = tid; i < rin * F2_CS; i += bs) in[i] = 0.0f;
  for (int i = tid; i < ncell; i += bs) cnt[i] = 0;
  if (tid == 0) s_m = 0;
  __syncthreads();
  DepEntry* arr = reinterpret_cast<DepEntry*>(in);
  DepEntry* grp = reinterpret_cast<DepEntry*>(rt);
  int cap = DF_ENT;
  int m = 0;
  for (int pass = 0; pass < 2; ++pass) {
    // 1. entries of the candidate components' nodes, chunk by chunk
    for (int base = 0; base < gi.ncomp; base += bs) {
      int is = 0, nn = 0, first = 0;
      if (base + tid < gi.ncomp) {
        const int4 r = dep.comp_rows[gi.comp0 + base + tid];
        const int2 cb = dep.comp_cols[gi.comp0 + base + tid];
        // bands r.x .. r.x + r.y - 1 and cb.x .. cb.x + cb.y - 1 against the
        // window's rows yb .. yb + rin - 1 (mod H) and columns xb .. (mod W)
        if ((r.y >= H || wrap1(r.x - yb, H) < rin || wrap1(yb - r.x, H) < r.y) &&
            (cb.y >= W || wrap1(cb.x - xb, W) < cin || wrap1(xb - cb.x, W) < cb.y)) {
          is = 1;
          nn = r.z;
          first = r.w;
        }
      }
      int2 tot;
      const int2 inc = block_inclusive_scan2(make_int2(is, nn), wsum, &tot);
      if (is) {
        cand[inc.x - 1] = first;
        cand[bs + inc.x - 1] = inc.y - nn;
      }
      __syncthreads();
      constexpr int CR = 4;  // candidate nodes per thread with their loads in flight together
      for (int f0 = tid; f0 < tot.y; f0 += CR * bs) {
        int nid[CR];
#pragma unroll
        for (int u = 0; u < CR; ++u) {
          const int f = f0 + u * bs;
          nid[u] = -1;
          if (f < tot.y) {
            int lo = 0, hi = tot.x - 1;  // the last candidate whose node offset is <= f
            while (lo < hi) {
              const int mid = (lo + hi + 1) >> 1;
              if (cand[bs + mid] <= f) lo = mid; else hi = mid - 1;
            }
            const int loc = f - cand[bs + lo], fc = cand[lo];
            nid[u] = fc >= 0 ? fc + loc : dep.L.comp_nodes[~fc + loc];
          }
        }
        int4 cc[CR];
        float4 ww[CR];
#pragma unroll
        for (int u = 0; u < CR; ++u) {
          const size_t q = 4 * (size_t)(nid[u] < 0 ? 0 : nid[u]);
          cc[u] = *reinterpret_cast<const int4*>(dep.cic_cell + q);
          ww[u] = *reinterpret_cast<const float4*>(dep.cic_w + q);
        }
#pragma unroll
        for (int u = 0; u < CR; ++u) {
          if (nid[u] < 0) continue;
          const float wq[4] = {ww[u].x, ww[u].y, ww[u].z, ww[u].w};
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            // corner q: x0 / x1 (q & 1), row y0 / y1 (q & 2); rows are global steric rows
            const int x = (q & 1) ? cc[u].y : cc[u].x;
            const int y = ((q & 2) ? cc[u].w : cc[u].z) - gi.row0;
            const int r = wrap1(y - yb, H), c = wrap1(x - xb, W);
            if (r >= rin || c >= cin) continue;
            const int e = atomicAdd(&s_m, 1);
            if (e >= cap) continue;
            const int pos = r * cin + c;
            arr[e] = DepEntry{(pos << 2) | q, nid[u], wq[q]};
            atomicAdd(&cnt[pos], 1);
          }
        }
      }
      __syncthreads();
    }
    m = s_m;
    if (m <= cap) break;
    // 1'. more entries than the LDS holds: all of them again, into a slice
    // of the arena (arrival and grouped lists of m entries each)
    __syncthreads();
    if (tid == 0) {
      s_base = atomicAdd(dep.tile_top, 2 * m);
      s_m = 0;
    }
    for (int i = tid; i < ncell; i += bs) cnt[i] = 0;
    for (int i = tid; i < 3 * DF_ENT; i += bs) in[i] = 0.0f;  // the LDS list's partial fill
    __syncthreads();
    arr = dep.tile_arena + s_base;
    grp = arr + m;
    cap = m;
  }
  // 2. exclusive scan of the cells' counts (a run of cells per thread), then
  // the entries grouped by cell in arrival order; cnt[i] ends as the end of
  // cell i's entries
  {
    const int per = (ncell + bs - 1) / bs;
    int loc = 0;
    for (int k = 0; k < per; ++k) {
      const int i = tid * per + k;
      if (i < ncell) loc += cnt[i];
    }
    int tot;
    int run = block_inclusive_scan(loc, wsum, &tot) - loc;
    for (int k = 0; k < per; ++k) {
      const int i = tid * per + k;
      if (i < ncell) {
        const int v = cnt[i];
        cnt[i] = run;
        run += v;
      }
    }
  }
  __syncthreads();
  for (int e = tid; e < m; e += bs) {
    const DepEntry d = arr[e];
    grp[atomicAdd(&cnt[d.pq >> 2], 1)] = d;
  }
  __syncthreads();
  if (cap == DF_ENT)  // the LDS arrival list is dead: the window back to zero
    for (int i = tid; i < 3 * m; i += bs) in[i] = 0.0f;
  __syncthreads();
  // 3. each cell's fold in (corner, node id) order from zero, by the thread
  // holding the cell's first entry
  for (int lo = tid; lo < m; lo += bs) {
    const int i = grp[lo].pq >> 2;
    if (lo > 0 && (grp[lo - 1].pq >> 2) == i) continue;
    int hi = lo + 1;
    while (hi < m && (grp[hi].pq >> 2) == i) ++hi;
    for (int j = lo + 1; j < hi; ++j) {
      const DepEntry v = grp[j];
      int k = j - 1;
      while (k >= lo && ((grp[k].pq & 3) > (v.pq & 3) ||
                         ((grp[k].pq & 3) == (v.pq & 3) && grp[k].n > v.n))) {
        grp[k + 1] = grp[k];
        --k;
      }
      grp[k + 1] = v;
    }
    float acc = 0.0f;
    for (int j = lo; j < hi; ++j) acc = acc + grp[j].w;
    in[(i / cin) * F2_CS + (i % cin)] = acc;
  }
  __syncthreads();
  if (density_out) {
    // the tile's own cells: window rows -loy .., columns -lox ..
    for (int j = tid; j < FILTER_TILE_H * FILTER_TILE_W; j += bs) {
      const int oy = j / FILTER_TILE_W, ox = j - oy * FILTER_TILE_W;
      const int y = ty0 + oy, x = tx0 + ox;
      if (x < W && y < H)
        density_out[gi.grid_off + (size_t)y * W + x] = in[(oy - loy) * F2_CS + (ox - lox)];
    }
  }
  filter2d_passes(in, rt, tx, ty, nx, ny, rin, gi, ty0, tx0, out);
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

hipError_t launch_predict(const Consts& c, const BatchDev& b, const RodLayoutDev& L,
                          const float* pos, const float* vel, float* pred_w, int* cic_cell,
                          float* cic_w, int4* comp_rows, int2* comp_cols, int* arena_top,
                          hipStream_t s) {
  if (L.num_comp == 0) return hipSuccess;
  hipLaunchKernelGGL(k_predict, dim3(L.num_comp), dim3(64), 0, s, c, b, L, pos, vel, pred_w,
                     cic_cell, cic_w, comp_rows, comp_cols, arena_top);
  return hipGetLastError();
}

hipError_t steric_prepare(int lds_max) {
  // the FFT passes of long sides use more than 64 KB of LDS
  const int lds = lds_max - 1024;
  const void* fns[] = {reinterpret_cast<const void*>(&k_rows_fwd_T),
                       reinterpret_cast<const void*>(&k_cols_T),
                       reinterpret_cast<const void*>(&k_rows_inv_T),
                       reinterpret_cast<const void*>(&k_rows_conv)};
  for (const void* f : fns) {
    const hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, lds);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

hipError_t launch_gauss_table(int H, int W, float coef, float* kfT, hipStream_t s) {
  const size_t n = (size_t)H * W;
  hipLaunchKernelGGL(k_gauss_table_T, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, H, W, coef,
                     kfT);
  return hipGetLastError();
}

// Workgroup size of an FFT pass: a power of two from one to four wavefronts,
// about one radix-4 butterfly per thread and stage for the longest sequence.
static int fft_threads(int n) { return n <= 256 ? 64 : (n <= 512 ? 128 : 256); }

// Twiddle elements an FFT pass of length n keeps in LDS: the plain table (n)
// or the per-stage table (n - 1 plus R roots per stage of radix R > 4).
static int fft_tw_elems(int n) { return n + 64; }

// LDS of one FFT pass of length n: two padded buffers and the twiddles.
static size_t fft_pass_bytes(int n) {
  return ((size_t)2 * ct_buf(n) + fft_tw_elems(n)) * sizeof(float2);
}

// Smallest deposit capacity: DEP_CAP_MIN, and room for the candidate lists
// (2 ints per thread in the 16 cap bytes of the ranked keys).
static int dep_cap_floor(int threads) {
  const int c = threads / 2;
  return ((c > DEP_CAP_MIN ? c : DEP_CAP_MIN) + 3) & ~3;
}

// LDS layout of the forward row pass (float2 units): the row, the second
// buffer (which the deposit scratch of `cap` nodes aliases), the twiddles.
struct RowsLayout {
  int b_off, tw_off, total, cap;
};
static RowsLayout rows_layout(int max_w) {
  RowsLayout L;
  const int floor_cap = dep_cap_floor(fft_threads(max_w));
  L.b_off = (ct_buf(max_w) + 1) & ~1;  // 16-byte aligned scratch
  int bsz = (ct_buf(max_w) + 1) & ~1;
  const int need = (floor_cap * DEP_BYTES_PER_NODE + 15) / 16 * 2;
  if (bsz < need) bsz = need;
  L.cap = (bsz * (int)sizeof(float2) / DEP_BYTES_PER_NODE) & ~3;
  if (L.cap > DEP_CAP_MAX) L.cap = DEP_CAP_MAX;
  L.tw_off = L.b_off + bsz;
  L.total = L.tw_off + fft_tw_elems(max_w);
  return L;
}
size_t fft_rows_lds_bytes(int max_w) { return (size_t)rows_layout(max_w).total * sizeof(float2); }
size_t fft_cols_lds_bytes(int max_h) { return fft_pass_bytes(max_h); }

hipError_t launch_steric_fft_T(const BatchDev& b, const float* real_in, const StericDep* d,
                               float2* specT, const float* kfT, float* density_out, float* out,
                               hipStream_t s) {
  const int tr = fft_threads(b.max_w), tc = fft_threads(b.max_h);
  const RowsLayout rl = rows_layout(b.max_w);
  DepArgs dep{};
  if (d) {
    dep.L = d->L;
    dep.comp_rows = d->comp_rows;
    dep.comp_cols = d->comp_cols;
    dep.cic_cell = d->cic_cell;
    dep.cic_w = d->cic_w;
    dep.arena = d->arena;
    dep.arena_top = d->arena_top;
    dep.sort_tmp = d->sort_tmp;
    dep.key_tmp = d->key_tmp;
    dep.key_half = d->key_half;
    dep.tile_arena = d->tile_arena;
    dep.tile_top = d->tile_top;
  }
  dep.cap = rl.cap;
  dep.b_off = rl.b_off;
  dep.tw_off = rl.tw_off;
  hipLaunchKernelGGL(k_rows_fwd_T, dim3(b.rows), dim3(tr), fft_rows_lds_bytes(b.max_w), s, b,
                     real_in, dep, specT, kfT, density_out);
  hipLaunchKernelGGL(k_cols_T, dim3(b.cols), dim3(tc), fft_cols_lds_bytes(b.max_h), s, b, specT,
                     kfT);
  hipLaunchKernelGGL(k_rows_inv_T, dim3(b.rows), dim3(tr), fft_pass_bytes(b.max_w), s, b, specT,
                     kfT, out);
  return hipGetLastError();
}

hipError_t launch_steric_conv(const BatchDev& b, const float* real_in, const StericDep* d,
                              float* tmp, const float* taps, float* density_out, float* out,
                              bool fused, hipStream_t s) {
  const int tr = fft_threads(b.max_w);
  const RowsLayout rl = rows_layout(b.max_w);
  DepArgs dep{};
  if (d) {
    dep.L = d->L;
    dep.comp_rows = d->comp_rows;
    dep.comp_cols = d->comp_cols;
    dep.cic_cell = d->cic_cell;
    dep.cic_w = d->cic_w;
    dep.arena = d->arena;
    dep.arena_top = d->arena_top;
    dep.sort_tmp = d->sort_tmp;
    dep.key_tmp = d->key_tmp;
    dep.key_half = d->key_half;
    dep.tile_arena = d->tile_arena;
    dep.tile_top = d->tile_top;
  }
  dep.cap = rl.cap;
  dep.b_off = rl.b_off;
  dep.tw_off = rl.tw_off;
  // the FFT pass's layout without its twiddles: the taps need FILTER_TAPS_MAX floats
  const size_t lds = (size_t)(rl.tw_off + (FILTER_TAPS_MAX + 1) / 2) * sizeof(float2);
  if (fused && !real_in && d && d->tile_arena) {
    // the deposit and both filter passes in one launch
    hipLaunchKernelGGL(k_dep_filter2d, dim3(b.ctiles), dim3(F2_THREADS), 0, s, b, taps, dep,
                       density_out, out);
    return hipGetLastError();
  }
  if (fused) {
    // the deposit (unless the input is given) into density_out or tmp, then
    // both filter passes in one launch
    const float* src = real_in;
    if (!real_in) {
      float* dens = density_out ? density_out : tmp;
      hipLaunchKernelGGL(k_rows_conv, dim3(b.rows), dim3(tr), lds, s, b, nullptr, dep, taps, tmp,
                         dens, 0);
      src = dens;
    }
    hipLaunchKernelGGL(k_filter2d, dim3(b.ctiles), dim3(F2_THREADS), 0, s, b, taps, src, out);
    return hipGetLastError();
  }
  hipLaunchKernelGGL(k_rows_conv, dim3(b.rows), dim3(tr), lds, s, b,
                     real_in, dep, taps, tmp, density_out, 1);
  hipLaunchKernelGGL(k_cols_conv_tiled, dim3(b.ctiles), dim3(256), 0, s, b, taps, tmp, out);
  return hipGetLastError();
}

}  // namespace mc
