// Standalone launch of the fused D2Q9 TRT + Guo collide-and-stream pass
// (fluid.py:268-357); the per-cell update is in mc_lbm.h.
#include <hip/hip_ext.h>

#include "mc_lbm.h"

namespace mc {
namespace {

__global__ void __launch_bounds__(LBM_THREADS) k_lbm(LbmArgs a) {
  // XCD-aware block order: workgroup i runs on XCD i % 8, so give each XCD a
  // contiguous band of cells (its L2 then holds the rows it writes into).
  int b = blockIdx.x;
  const int G = gridDim.x;
  if ((G & 7) == 0) b = (b & 7) * (G >> 3) + (b >> 3);
  lbm_cell(a, b, threadIdx.x);
}

}  // namespace

LbmArgs make_lbm_args(const Consts& c, const BatchDev& b, const float* f_in, float* f_out,
                      float* ustar, uint8_t* mask, const float* force) {
  LbmArgs a;
  a.bd = b;
  a.fin = f_in;
  a.fout = f_out;
  a.ustar = ustar;
  a.mask = mask;
  a.force = force;
  a.op = c.omega_p;
  a.om = c.omega_m;
  a.cp = c.cp;
  a.cm = c.cm;
  a.force_clamp = c.force_clamp;
  return a;
}

hipError_t launch_lbm(const Consts& c, const BatchDev& b, const float* f_in, float* f_out,
                      float* ustar, uint8_t* mask, const float* force, hipStream_t s,
                      hipEvent_t start, hipEvent_t stop) {
  const LbmArgs a = make_lbm_args(c, b, f_in, f_out, ustar, mask, force);
  if (start && stop) {
    // events carried by the dispatch itself: they bracket the kernel's execution
    hipExtLaunchKernelGGL(k_lbm, dim3(b.lbm_blocks), dim3(LBM_THREADS), 0, s, start, stop, 0, a);
  } else {
    hipLaunchKernelGGL(k_lbm, dim3(b.lbm_blocks), dim3(LBM_THREADS), 0, s, a);
  }
  return hipGetLastError();
}

}  // namespace mc
