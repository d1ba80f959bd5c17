// Trajectory recording of a rollout (simulate.py:15-35): after every
// stride-th step, one frame of the node arrays, the steric field and the
// fluid velocity at the grid's resolution.  The frame index comes from a
// device step counter, so the same launch recorded into a HIP graph writes
// successive frames on every replay.
//
// When the lattice and the grid differ, the velocity is resized as
// jax.image.resize(u, (2, H, W), method="linear") does (simulate.py:24-28):
// separable triangle-kernel weights per output row and column, the kernel
// widened by the scale factor when downsampling (antialiasing), each output
// index's weights normalised to sum 1 and zeroed when the sample lies outside
// the input.  The weights are tabulated on the host in float32 (resize_taps);
// the kernel folds the column taps of each row tap, then the row taps.
#include <cmath>
#include <vector>

#include "mc_common.h"
#include "mc_internal.h"

namespace mc {
namespace {

constexpr int REC_THREADS = 256;

__global__ void __launch_bounds__(REC_THREADS) k_record(RecordDev r) {
  const long long c = *r.counter;
  if ((c + 1) % r.stride != 0) return;
  const long long frame = (c + 1) / r.stride - 1;
  const long long n2 = 2LL * r.N, hw = (long long)r.H * r.W;
  const long long total = 3 * n2 + hw + 2 * hw;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < total;
       i += (long long)gridDim.x * blockDim.x) {
    if (i < 3 * n2) {
      const int which = (int)(i / n2);
      const long long j = i - which * n2;
      float* dst = which == 0 ? r.pos : which == 1 ? r.vel : r.dbg;
      const float* src = which == 0 ? r.s_pos : which == 1 ? r.s_vel : r.s_dbg;
      if (dst) dst[frame * n2 + j] = src[j];
      continue;
    }
    long long k = i - 3 * n2;
    if (k < hw) {
      if (r.steric) r.steric[frame * hw + k] = r.s_steric[k];
      continue;
    }
    k -= hw;
    if (!r.fv) continue;
    const int ch = (int)(k / hw);
    const long long pix = k - ch * hw;
    const int y = (int)(pix / r.W), x = (int)(pix - (long long)y * r.W);
    float v;
    if (!r.resize) {
      v = r.s_fv[k];
    } else {
      const float* plane = r.s_fv + (size_t)ch * r.Hf * r.Wf;
      const int x0 = r.rx_off[x], x1 = r.rx_off[x + 1];
      v = 0.0f;
      for (int a = r.ry_off[y]; a < r.ry_off[y + 1]; ++a) {
        const float* row = plane + (size_t)r.ry_idx[a] * r.Wf;
        float s = 0.0f;
        for (int b = x0; b < x1; ++b) s = s + r.rx_w[b] * row[r.rx_idx[b]];
        v = v + r.ry_w[a] * s;
      }
    }
    r.fv[frame * 2 * hw + k] = v;
  }
}

}  // namespace

void resize_taps(int in, int out, std::vector<int>& off, std::vector<int>& idx,
                 std::vector<float>& w) {
  off.assign(1, 0);
  idx.clear();
  w.clear();
  // jax.image.resize: scale = out / in (a Python float), inv_scale = 1 / scale;
  // sample_f and the distances in float32; the triangle kernel widened to
  // max(inv_scale, 1) (antialias); weights divided by their sum when it
  // exceeds 1000 eps; zero outside -0.5 <= sample_f <= in - 0.5
  const double scale = (double)out / (double)in;
  const double inv = 1.0 / scale;
  const float inv_f = (float)inv;
  const float ks = (float)(inv > 1.0 ? inv : 1.0);
  std::vector<float> col(in);
  for (int o = 0; o < out; ++o) {
    const float sf = (((float)o + 0.5f) * inv_f - 0.0f * inv_f) - 0.5f;
    float sum = 0.0f;
    for (int i = 0; i < in; ++i) {
      const float x = std::fabs(sf - (float)i) / ks;
      const float t = 1.0f - std::fabs(x);
      col[i] = t > 0.0f ? t : 0.0f;
      sum = sum + col[i];
    }
    const bool inside = sf >= -0.5f && sf <= (float)in - 0.5f;
    const bool norm = std::fabs(sum) > 1000.0f * 1.1920929e-7f;
    for (int i = 0; i < in; ++i) {
      float v = norm ? col[i] / (sum != 0.0f ? sum : 1.0f) : 0.0f;
      if (!inside) v = 0.0f;
      if (v != 0.0f) {
        idx.push_back(i);
        w.push_back(v);
      }
    }
    off.push_back((int)idx.size());
  }
}

hipError_t launch_record(const RecordDev& r, hipStream_t s) {
  const long long hw = (long long)r.H * r.W;
  const long long total = 6LL * r.N + 3 * hw;
  long long blocks = (total + REC_THREADS - 1) / REC_THREADS;
  if (blocks > 4096) blocks = 4096;
  if (blocks < 1) blocks = 1;
  hipLaunchKernelGGL(k_record, dim3((unsigned)blocks), dim3(REC_THREADS), 0, s, r);
  return hipGetLastError();
}

}  // namespace mc
