// Residual (+ JVP) of a batch of meshes of different sizes, one launch per
// stage over all meshes.
//
// Per mesh this is the reference's per-iteration sequence -- the BC calls in
// order (BROADCAST_npz.py:1018-1021 primal, :1079-1082 tangent) then
// flux_num_dnc5_2d[_d] (srcfv/rhs/flux_num_dnc5.F90:7-226,
// srcfv/tangent/flux_num_dnc5_d.f90:15) -- with the same device functions as
// the per-mesh kernels.  A batch of small and uneven meshes would leave most
// of the chip idle under one launch per stage and mesh (a 128 x 64 mesh is 32
// blocks of 256 cells on 256 CUs); here every stage is one launch over the
// concatenated work of all meshes (one thread per item, the mesh of an item
// found in the stage's prefix table), so the whole batch is ~10 launches,
// each at its own occupancy.
//
// Stages (each over the flat index of all meshes' items):
//   pack      every cell        ws = (w, wd) as Dual<1> (or w as double)
//   bc k      lines of op k     the k-th BC of every mesh that has one
//   prim      every cell        primitives
//   grad      interior cells    4th-order velocity gradients
//   ghostgrad im + jm lines     extrapolated ghost gradients
//   face      (im+1)(jm+1)      hn of the i- and j-face at (i, j)
//   balance   interior cells    residu (and residud); ghost frame of ws back
//                               into w (and wd), as the BC calls leave them
#pragma once
#include "bcs.h"
#include "faces.h"

namespace bca {

constexpr int BATCH_MAX = 16;  // meshes per launch
constexpr int BATCH_OPS = 8;   // BC operations per mesh

struct BatchBc {
  int kind, loc;
  Box box, box2;
  int tr0, tr1;
  const double* data;
  int lm;
  double gam;
};

template <class T>
struct BatchItem {
  Mesh m;
  Phys ph;
  Metric g;
  double* w;     // state, ghost frame filled in place
  double* wd;    // tangent (T = Dual<1>), ghost frame filled in place
  double* res;   // residu, interior written
  double* resd;  // residud (T = Dual<1>)
  T* ws;         // scratch: packed state, primitives, gradients, face fluxes
  T* prim;
  T* grad;
  T* hn;
  int nops;
  BatchBc ops[BATCH_OPS];
};

// Prefix sums of each stage's item counts over the meshes.
struct BatchPlan {
  int n, maxops;
  long cells[BATCH_MAX + 1];
  long interior[BATCH_MAX + 1];
  long edges[BATCH_MAX + 1];
  long faces[BATCH_MAX + 1];
  long bc[BATCH_OPS][BATCH_MAX + 1];
};

// mesh of flat item t (pre[0] = 0 <= t < pre[n]); n is small, a linear walk
__device__ inline int batch_which(const long* pre, int n, long t) {
  int k = 0;
  while (k + 1 < n && t >= pre[k + 1]) ++k;
  return k;
}

template <class T>
__device__ inline void batch_pack(const BatchItem<T>& it, long c) {
  const long n = it.m.ncell;
#pragma unroll
  for (int e = 0; e < 5; ++e) {
    T x(it.w[c + e * n]);
    if constexpr (!std::is_same<T, double>::value) x.d[0] = it.wd[c + e * n];
    it.ws[c + e * n] = x;
  }
}

template <class T>
__device__ inline void batch_bc(const BatchItem<T>& it, const BatchBc& op, int t) {
  const Mesh& m = it.m;
  switch (op.kind) {
    case 0:  // inlet
      bc_inlet_line(m, it.ws, op.loc, op.box, op.data, op.lm, it.g.nx, it.g.ny, op.gam, t + 1);
      break;
    case 1:  // no reflexion
      bc_noref_line(m, it.ws, op.data, op.lm, op.loc, op.box, it.g.nx, it.g.ny, op.gam, t + 1);
      break;
    case 2:
      bc_extrap2_line(m, it.ws, op.loc, op.box, 5, t + 1);
      break;
    case 3:
      bc_wall_line(m, it.ws, op.loc, op.box, op.gam, t + 1);
      break;
    default:
      join_cell(m, it.ws, op.box, op.box2, op.tr0, op.tr1, 5, m.ni, m.ncell, t);
      break;
  }
}

template <class T>
__device__ inline void batch_face(const BatchItem<T>& it, int i, int j) {
  const Mesh& m = it.m;
  PlainAcc<T> a;
  a.m = m; a.ph = it.ph; a.g = it.g; a.i = i; a.j = j;
  a.w = it.ws; a.prim = it.prim; a.grad = it.grad;
  const long f = nidx(m, i, j);
  T h[5];
  if (j <= m.jm) {
    iface_any(a, h);
#pragma unroll
    for (int e = 0; e < 5; ++e) it.hn[f + (long)e * m.nnode] = h[e];
  }
  if (i <= m.im) {
    jface_any(a, h);
#pragma unroll
    for (int e = 0; e < 5; ++e) it.hn[f + (long)(5 + e) * m.nnode] = h[e];
  }
}

template <class T>
__device__ inline void batch_ghost_grad(const BatchItem<T>& it, int t) {
  const Mesh& m = it.m;
  const long n = m.ncell;
  if (t < m.jm) {
    const int j = 1 + t;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      T* G = it.grad + k * n;
      G[cidx(m, 0, j)] = 2.0 * G[cidx(m, 1, j)] - G[cidx(m, 2, j)];
      G[cidx(m, m.im + 1, j)] = 2.0 * G[cidx(m, m.im, j)] - G[cidx(m, m.im - 1, j)];
    }
  } else {
    const int i = 1 + t - m.jm;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      T* G = it.grad + k * n;
      G[cidx(m, i, m.jm + 1)] = 2.0 * G[cidx(m, i, m.jm)] - G[cidx(m, i, m.jm - 1)];
    }
  }
}

// One kernel per stage, one thread per item t of the stage (items of all
// meshes concatenated in the order of the prefix table).
enum { BS_PACK = 0, BS_BC, BS_PRIM, BS_GRAD, BS_GHOSTGRAD, BS_FACE, BS_BALANCE };

template <class T, int STAGE>
__global__ void __launch_bounds__(256) k_batch_stage(const BatchItem<T>* __restrict__ items,
                                                     const BatchPlan* __restrict__ plan, int op) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const int n = plan->n;
  if constexpr (STAGE == BS_PACK) {
    if (t >= plan->cells[n]) return;
    const int k = batch_which(plan->cells, n, t);
    batch_pack(items[k], t - plan->cells[k]);
  } else if constexpr (STAGE == BS_BC) {
    const long* pre = plan->bc[op];
    if (t >= pre[n]) return;
    const int k = batch_which(pre, n, t);
    batch_bc(items[k], items[k].ops[op], (int)(t - pre[k]));
  } else if constexpr (STAGE == BS_PRIM) {
    if (t >= plan->cells[n]) return;
    const int k = batch_which(plan->cells, n, t);
    const BatchItem<T>& it = items[k];
    primitives(it.m, it.ph, it.ws, it.prim, t - plan->cells[k]);
  } else if constexpr (STAGE == BS_GRAD) {
    if (t >= plan->interior[n]) return;
    const int k = batch_which(plan->interior, n, t);
    const BatchItem<T>& it = items[k];
    const int q = (int)(t - plan->interior[k]);
    const int i = 1 + q % it.m.im, j = 1 + q / it.m.im;
    T out[4];
    gradient_at(it.m, it.g, it.prim, i, j, out);
    const long c = cidx(it.m, i, j);
#pragma unroll
    for (int g = 0; g < 4; ++g) it.grad[c + g * it.m.ncell] = out[g];
  } else if constexpr (STAGE == BS_GHOSTGRAD) {
    if (t >= plan->edges[n]) return;
    const int k = batch_which(plan->edges, n, t);
    batch_ghost_grad(items[k], (int)(t - plan->edges[k]));
  } else if constexpr (STAGE == BS_FACE) {
    if (t >= plan->faces[n]) return;
    const int k = batch_which(plan->faces, n, t);
    const BatchItem<T>& it = items[k];
    const int q = (int)(t - plan->faces[k]);
    batch_face(it, 1 + q % (it.m.im + 1), 1 + q / (it.m.im + 1));
  } else {
    if (t >= plan->cells[n]) return;
    const int k = batch_which(plan->cells, n, t);
    const BatchItem<T>& it = items[k];
    const Mesh& m = it.m;
    const long c = t - plan->cells[k];
    const int i = (int)(c % m.ni) - m.gh + 1, j = (int)(c / m.ni) - m.gh + 1;
    const bool in = i >= 1 && i <= m.im && j >= 1 && j <= m.jm;
#pragma unroll
    for (int e = 0; e < 5; ++e) {
      const long ce = c + (long)e * m.ncell;
      if (in) {
        const T r = balance_at(m, it.hn, i, j, e);
        it.res[ce] = val(r);
        if constexpr (!std::is_same<T, double>::value) it.resd[ce] = r.d[0];
      } else {
        const T x = it.ws[ce];
        it.w[ce] = val(x);
        if constexpr (!std::is_same<T, double>::value) it.wd[ce] = x.d[0];
      }
    }
  }
}

}  // namespace bca
