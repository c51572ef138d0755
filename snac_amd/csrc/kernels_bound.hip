// Ghost-plane kernel: one launch executes one phase of mutually independent
// FaceOps (halo copies between blocks or within a periodic block, set_bc
// forms of bound.f90:108-248, inflow planes of bound.f90:438-477).  Phases
// run in the reference's fill order; see bound_program.cpp.
#include "common.hpp"

namespace snac {
namespace {

__device__ __forceinline__ long plane_idx(int d, int q, int t1, int t2, long sx, long sy) {
  if (d == 0) return (long)q + sx * ((long)t1 + sy * (long)t2);
  if (d == 1) return (long)t1 + sx * ((long)q + sy * (long)t2);
  return (long)t1 + sx * ((long)t2 + sy * (long)q);
}

__global__ __launch_bounds__(256) void k_face_ops(const FaceOp *__restrict__ ops) {
  const FaceOp op = ops[blockIdx.y];
  const int t = blockIdx.x * 256 + threadIdx.x;
  if (t >= op.m1 * op.m2) return;
  const int t1 = t % op.m1, t2 = t / op.m1;
  double val = op.c0;
  if (op.src) val += op.c1 * op.src[plane_idx(op.d, op.qsrc, t1, t2, op.ssx, op.ssy)];
  if (op.arr) val += op.c2 * op.arr[t];
  op.dst[plane_idx(op.d, op.qdst, t1, t2, op.dsx, op.dsy)] = val;
}

// one workgroup per segment of a composed ghost program
__global__ __launch_bounds__(256) void k_cell_segs(const CellSeg *__restrict__ segs) {
  const CellSeg &g = segs[blockIdx.x];
  for (int i = threadIdx.x; i < g.len; i += 256) {
    double v = g.base ? g.base[g.b0 + (long)i * g.bstr] : 0.0;
    for (int l = 0; l < g.nlev; l++) {
      const CellLev &L = g.lev[l];
      double val = L.c0;
      if (L.src) val += L.c1 * v;
      if (L.arr) val += L.c2 * L.arr[L.a0 + i];
      v = val;
    }
    g.dst[g.d0 + (long)i * g.dstr] = v;
  }
}

}  // namespace

void launch_cell_segs(hipStream_t s, const CellSeg *dev_segs, int nseg) {
  if (nseg <= 0) return;
  hipLaunchKernelGGL(k_cell_segs, dim3(nseg), dim3(256), 0, s, dev_segs);
  SNAC_HIP(hipGetLastError());
}

void launch_face_ops(hipStream_t s, const FaceOp *dev_ops, int nops, int max_plane) {
  if (nops <= 0) return;
  hipLaunchKernelGGL(k_face_ops, dim3((max_plane + 255) / 256, nops), dim3(256), 0, s, dev_ops);
  SNAC_HIP(hipGetLastError());
}

}  // namespace snac
