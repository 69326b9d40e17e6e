// C ABI of the gfx950 DNC5 path (declared in include/broadcast_amd.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/broadcast_amd.h"
#include "bcs.h"
#include "faces.h"
#include "blocks.h"
#include "jacobian.h"
#include "misc.h"
#include "scheme.h"
#include "vjp.h"
#include "csr.h"
#include "dz.h"
#include "batch.h"
#include "bcadj.h"
#include "jaccsr.h"
#include "gemm.h"
#include "bandlu.h"

using namespace bca;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

#define BCA_HIP(call)                                                                   \
  do {                                                                                  \
    hipError_t e_ = (call);                                                             \
    if (e_ != hipSuccess) return fail(-2, std::string(#call ": ") + hipGetErrorString(e_)); \
  } while (0)

#define BCA_LAUNCHED()                                                                   \
  do {                                                                                   \
    hipError_t e_ = hipGetLastError();                                                   \
    if (e_ != hipSuccess) return fail(-3, std::string("kernel launch: ") + hipGetErrorString(e_)); \
  } while (0)

// Per-scalar-type scratch planes of one pipeline pass.
template <class T>
struct Work {
  T* w = nullptr;
  T* res = nullptr;
  T* prim = nullptr;
  T* grad = nullptr;
  T* hn = nullptr;
};

}  // namespace

// HIP-event timing of individual kernel launches on the context stream
struct KernelTimer {
  bool on = false;
  std::vector<hipEvent_t> pool;
  size_t used = 0;
  std::vector<std::pair<std::string, size_t>> marks;  // name, index of start event
};

// What the last bca_jacobian_colored_csr[_refresh] call left in the caller's
// arrays: its zones, threshold, arrays and nnz (csr_mask / csr_cnt hold its
// kept masks and row lengths).
struct CsrRecord {
  bool valid = false;
  int nz = 0, zparam = 0;
  double mini = 0.0;
  const int* indptr = nullptr;
  const int* indices = nullptr;
  const double* data = nullptr;
  long nnz = 0;
};

struct bca_ctx {
  Mesh m;
  hipStream_t stream = nullptr;
  // the edge-row face families (one or two rows of faces, a few waves each)
  // run on these, concurrently with the interior families on `stream`; every
  // call joins them back before it returns
  hipStream_t side[3] = {nullptr, nullptr, nullptr};
  hipEvent_t fork_ev = nullptr, join_ev[3] = {nullptr, nullptr, nullptr};
  long bytes = 0;
  KernelTimer timer;
  Work<double> w0;
  Work<Dual<1>> w1;
  double* norm_part = nullptr;
  double* norm_out = nullptr;
  double* blki = nullptr;   // face-local Jacobian blocks, i-faces, one plane per offset
  double* blkj = nullptr;   // j-faces
  double* ghostT = nullptr; // ghost tangents per zone and color
  Dual<5>* wcol = nullptr;  // seeded states of the zones x 49 colors (compact BC frames)
  unsigned char* ghostnz = nullptr;  // nonzero flags of the ghost tangents
  int zone_cap = 0;         // number of zones the three arrays above hold
  int* csr_cnt = nullptr;   // row lengths of the pruned CSR
  int* csr_aux = nullptr;   // scan block sums, then total, error flag, changed flag (CSR_AUX_INTS)
  unsigned long long* csr_mask = nullptr;  // kept masks of the rows in column order
  double* vjpi = nullptr;   // VJP contributions per face and offset, i-faces
  double* vjpj = nullptr;   // j-faces
  double* jbeta_comp = nullptr;  // J(beta) components on the pattern
  long jbeta_cap = 0;
  Dual<5>* wadj = nullptr;  // seeded states of the BC adjoints (compact BC frames)
  double* ghostG = nullptr; // ghost part of the fused CSR path, per near row and slot color
  long ghostG_cap = 0;
  CsrRecord csr_rec;        // the last pruned CSR written without the COO (refresh key)
  unsigned long long* jr_mask = nullptr;  // its kept masks (column order, 8 words per row)
  int* jr_cnt = nullptr;                  // its row lengths
  // launch shapes chosen by bca_autotune for this mesh (defaults until then)
  dim3 tune_face{64, 4, 1};      // k_face<double>
  dim3 tune_blocks{64, 4, 1};    // interior face-block kernels of phase 1
  dim3 tune_assemble{256, 1, 1}; // k_assemble_colored
};

namespace {

constexpr int NORM_BLOCKS = 240;
// csr_aux: SCAN_B * SCAN_B scan block sums, then the total, the error flag
// and the changed-row flag of the refresh pass (every user allocates this size)
constexpr long CSR_AUX_INTS = (long)SCAN_B * SCAN_B + 4;

template <class T> const char* tname();
template <> const char* tname<double>() { return "f64"; }
template <> const char* tname<Dual<1>>() { return "dual1"; }
template <> const char* tname<Dual<5>>() { return "dual5"; }

// Records a start event before and a stop event after the launch it wraps.
struct Timed {
  bca_ctx* c;
  size_t idx = 0;
  bool live = false;
  Timed(bca_ctx* cc, const std::string& name) : c(cc) {
    KernelTimer& t = c->timer;
    if (!t.on) return;
    while (t.pool.size() < t.used + 2) {
      hipEvent_t e;
      if (hipEventCreate(&e) != hipSuccess) return;
      t.pool.push_back(e);
    }
    idx = t.used;
    t.used += 2;
    t.marks.emplace_back(name, idx);
    live = hipEventRecord(t.pool[idx], c->stream) == hipSuccess;
  }
  ~Timed() {
    if (live) (void)hipEventRecord(c->timer.pool[idx + 1], c->stream);
  }
};

template <class T>
int alloc_planes(bca_ctx* c, T** p, long count) {
  if (*p) return 0;
  BCA_HIP(hipMalloc((void**)p, sizeof(T) * count));
  c->bytes += (long)sizeof(T) * count;
  return 0;
}

// (Re)allocates *p for count elements when cap < want (contents dropped).
template <class T>
int realloc_planes(bca_ctx* c, T** p, long count) {
  if (*p) {
    (void)hipFree(*p);
    *p = nullptr;
  }
  return alloc_planes(c, p, count);
}

template <class T>
int ensure_work(bca_ctx* c, Work<T>& wk, bool with_state) {
  const Mesh& m = c->m;
  int rc = 0;
  if ((rc = alloc_planes(c, &wk.prim, NPRIM * m.ncell))) return rc;
  if ((rc = alloc_planes(c, &wk.grad, NGRAD * m.ncell))) return rc;
  if ((rc = alloc_planes(c, &wk.hn, 10 * m.nnode))) return rc;
  if (with_state) {
    if ((rc = alloc_planes(c, &wk.w, 5 * m.ncell))) return rc;
    if ((rc = alloc_planes(c, &wk.res, 5 * m.ncell))) return rc;
  }
  return 0;
}

template <class T>
void free_work(Work<T>& wk) {
  for (T* p : {wk.w, wk.res, wk.prim, wk.grad, wk.hn})
    if (p) (void)hipFree(p);
  wk = Work<T>();
}

inline dim3 grid2(int ni, int nj, dim3 b) { return dim3((ni + b.x - 1) / b.x, (nj + b.y - 1) / b.y); }
inline unsigned grid1(long n, int b) { return (unsigned)((n + b - 1) / b); }

Phys to_phys(const bca_phys* p) {
  return Phys{p->cp, p->cv, p->prandtl, p->gam, p->rgaz, p->cs, p->muref, p->tref, p->s_suth, p->k2, p->k4};
}
Metric to_metric(const bca_metrics& g) { return Metric{g.nx, g.ny, g.vol, g.volf}; }

// gh = 3 for the 5th-order scheme and every Jacobian path; the residual and
// tangent of order k need gh >= (k + 1) / 2 (the reference's gh = (order +
// 1) / 2, BROADCAST_npz.py:501-502)
int check_mesh(const Mesh& m) {
  if (m.gh < 2 || m.gh > 5) return fail(-1, "gh must be 2..5 (the DNC orders 3, 5, 7, 9 take (order + 1) / 2)");
  if (m.im < 5 || m.jm < 5) return fail(-1, "im and jm must be >= 5 (wall stencil reaches j = 5)");
  return 0;
}
int need_gh3(const bca_ctx* c) {
  if (c && c->m.gh != 3) return fail(-1, "the 5th-order scheme and its Jacobian paths need gh = 3");
  return 0;
}

// The residual pipeline on scalar type T; residu == nullptr stops after the faces.
// WP: the state as planes of T, or a SplitState (T = Dual<1>)
template <class T, class WP = const T*>
int run_pipeline(bca_ctx* c, Work<T>& wk, WP w, T* residu, const Metric& g, const Phys& ph,
                 bool nowall = false, int ord = 5) {
  const Mesh& m = c->m;
  hipStream_t s = c->stream;
  const std::string sfx = std::string("<") + tname<T>() + ">";
  // the faces read the state by stencil from planes of T: a SplitState is
  // copied to wk.w by k_prim on the way
  const T* wf;
  T* wcopy = nullptr;
  if constexpr (std::is_same<WP, SplitState>::value) wf = wcopy = wk.w;
  else wf = w;
  {
    Timed t(c, "k_prim" + sfx);
    k_prim<T, WP><<<grid1(m.ncell, 256), 256, 0, s>>>(m, ph, w, wk.prim, wcopy);
  }
  BCA_LAUNCHED();
  const dim3 b(64, 4);
  {
    Timed t(c, "k_grad" + sfx);
    const dim3 gg = grid2(m.im, m.jm, b);
    if (ord == 3) k_grad<T, 3><<<gg, b, 0, s>>>(m, g, wk.prim, wk.grad);
    else if (ord == 7) k_grad<T, 7><<<gg, b, 0, s>>>(m, g, wk.prim, wk.grad);
    else if (ord == 9) k_grad<T, 9><<<gg, b, 0, s>>>(m, g, wk.prim, wk.grad);
    else k_grad<T><<<gg, b, 0, s>>>(m, g, wk.prim, wk.grad);
  }
  BCA_LAUNCHED();
  k_grad_ghost<T><<<grid1(2 * m.im + m.jm, 256), 256, 0, s>>>(m, wk.grad);
  BCA_LAUNCHED();
  const dim3 bf = sizeof(T) > 8 ? dim3(64, 1) : c->tune_face;
  {
    Timed t(c, "k_face" + sfx);
    const dim3 gf = grid2(m.im + 1, m.jm + 1, bf);
#define BCA_FACE(NW, ORD) k_face<T, NW, ORD><<<gf, bf, 0, s>>>(m, ph, g, wf, wk.prim, wk.grad, wk.hn)
    if (ord == 3 && nowall) BCA_FACE(true, 3);
    else if (ord == 3) BCA_FACE(false, 3);
    else if (ord == 7 && nowall) BCA_FACE(true, 7);
    else if (ord == 7) BCA_FACE(false, 7);
    else if (ord == 9 && nowall) BCA_FACE(true, 9);
    else if (ord == 9) BCA_FACE(false, 9);
    else if (nowall) BCA_FACE(true, 5);
    else BCA_FACE(false, 5);
#undef BCA_FACE
  }
  BCA_LAUNCHED();
  if (residu) {
    Timed t(c, "k_balance" + sfx);
    k_balance<T><<<grid2(m.im, m.jm, b), b, 0, s>>>(m, wk.hn, residu);
  }
  BCA_LAUNCHED();
  return 0;
}

int box_ok(const Mesh& m, int loc, const int* bx) {
  if (loc < 0 || loc > 3) return fail(-1, "loc must be 0..3 (Ilo, Ihi, Jlo, Jhi)");
  const int lo = 1 - m.gh;
  if (bx[0] < lo || bx[2] > m.im + m.gh || bx[1] < lo || bx[3] > m.jm + m.gh || bx[0] > bx[2] ||
      bx[1] > bx[3])
    return fail(-1, "interface box outside the block");
  // the ghost layers written and the interior layers read must stay inside
  const int i0 = loc == 0 ? 1 : loc == 1 ? -1 : 0, j0 = loc == 2 ? 1 : loc == 3 ? -1 : 0;
  const int reach = m.gh;
  for (int corner = 0; corner < 2; ++corner) {
    const int i = corner ? bx[2] : bx[0], j = corner ? bx[3] : bx[1];
    for (int d = -reach; d <= reach; ++d) {
      const int ii = i + d * i0, jj = j + d * j0;
      if (ii < lo || ii > m.im + m.gh || jj < lo || jj > m.jm + m.gh)
        return fail(-1, "interface box too close to the block edge for gh ghost layers");
    }
  }
  return 0;
}

// Shape checks of one BC operation against the mesh (the kernels index w,
// field, wbd and the normals from the boxes without further checks).
int bc_check(const Mesh& m, const bca_bc_op& op, const double* nx, const double* ny) {
  const Box b{op.box[0], op.box[1], op.box[2], op.box[3]};
  int rc = 0;
  if (op.kind != BCA_BC_JOIN && (rc = box_ok(m, op.loc, op.box))) return rc;
  const Side sd = decode_side(op.loc, b.imin, b.jmin, b.imax, b.jmax);
  switch (op.kind) {
    case BCA_BC_INLET:
      if (!op.data || !nx || !ny) return fail(-1, "inlet needs field, nx, ny");
      // the reference state is read at field(j, i, :) and the ghost profile at
      // field(l, de, :): both must stay inside field(lm, gh, 5)
      if (b.imin < 1 || b.imax > m.gh || b.jmin < 1 || b.jmax > op.lm || sd.lmax > op.lm)
        return fail(-1, "inlet interface does not fit the field(lm, gh, 5) profile");
      return 0;
    case BCA_BC_NOREF:
      if (!op.data || !nx || !ny) return fail(-1, "noref needs wbd, nx, ny");
      if (op.lm < sd.lmax) return fail(-1, "wbd shorter than the interface");
      return 0;
    case BCA_BC_EXTRAP2:
    case BCA_BC_WALL_ADIA:
      return 0;
    case BCA_BC_JOIN: {
      const Box prr{op.box[0], op.box[1], op.box[2], op.box[3]};
      const Box prd{op.box2[0], op.box2[1], op.box2[2], op.box2[3]};
      const int lo = 1 - m.gh;
      for (const Box& q : {prr, prd})
        if (q.imin < lo || q.imax > m.im + m.gh || q.jmin < lo || q.jmax > m.jm + m.gh ||
            q.imin > q.imax || q.jmin > q.jmax)
          return fail(-1, "join box outside the block");
      if (op.tr[0] != 1 || op.tr[1] != 2) return fail(-1, "batched join supports the identity transform only");
      if (prr.imax - prr.imin != prd.imax - prd.imin || prr.jmax - prr.jmin != prd.jmax - prd.jmin)
        return fail(-1, "join boxes differ in size");
      return 0;
    }
    default:
      return fail(-1, "unknown boundary-condition kind");
  }
}

// Does the index box [i0, i1] x [j0, j1] lie wholly in the frame of width W
// around the array (storage coordinates a < W or a >= ni - W, likewise b)?
bool box_in_frame(const Mesh& m, int i0, int j0, int i1, int j1, int W) {
  if (frame_whole(m, W)) return true;
  const int a0 = i0 + m.gh - 1, a1 = i1 + m.gh - 1, b0 = j0 + m.gh - 1, b1 = j1 + m.gh - 1;
  const bool meets_inner = a0 <= m.ni - W - 1 && a1 >= W && b0 <= m.nj - W - 1 && b1 >= W;
  return !meets_inner;
}

// The BC adjoints and the Jacobian's ghost tangents seed and propagate their
// Dual<5> states on the boundary frame only (ghost frame + 4 interior
// layers).  A checked BC operation qualifies when every cell it reads (the
// line and 3 cells further in, ADJ_REACH) and every ghost it writes lie in
// that frame: the interface sits on the block edge, join boxes lie in the
// frame.
bool op_in_frame(const Mesh& m, const bca_bc_op& op) {
  const int W = BC_FRAME;
  const Box b{op.box[0], op.box[1], op.box[2], op.box[3]};
  if (op.kind == BCA_BC_JOIN) {
    const Box d{op.box2[0], op.box2[1], op.box2[2], op.box2[3]};
    return box_in_frame(m, b.imin, b.jmin, b.imax, b.jmax, W) && box_in_frame(m, d.imin, d.jmin, d.imax, d.jmax, W);
  }
  const Side s = decode_side(op.loc, b.imin, b.jmin, b.imax, b.jmax);
  const int in = ADJ_REACH - 1;  // deepest interior offset read
  int i0 = b.imin, i1 = b.imax, j0 = b.jmin, j1 = b.jmax;
  if (s.i0 > 0) { i0 -= m.gh; i1 += in; }
  if (s.i0 < 0) { i0 -= in; i1 += m.gh; }
  if (s.j0 > 0) { j0 -= m.gh; j1 += in; }
  if (s.j0 < 0) { j0 -= in; j1 += m.gh; }
  return box_in_frame(m, i0, j0, i1, j1, W);
}
int frame_check(const Mesh& m, const bca_bc_op& op) {
  if (op_in_frame(m, op)) return 0;
  if (op.kind == BCA_BC_JOIN)
    return fail(-1, "join boxes must lie in the boundary frame (ghost layers + 4 interior layers) for the "
                    "linearised/adjoint BC sequence of the colored Jacobian and BC adjoints");
  return fail(-1, "interface not on the block edge: the linearised/adjoint BC sequence of the colored "
                  "Jacobian and BC adjoints covers the boundary frame (ghost layers + 4 interior layers) only");
}

// (w, wd) <-> Dual<1> on the boundary frame only (frame_cell(m, BC_FRAME, t)),
// for a tangent BC whose reads and writes lie in it (op_in_frame)
__global__ void k_pack_frame(Mesh m, const double* __restrict__ a, const double* __restrict__ ad,
                             Dual<1>* __restrict__ out) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= frame_count(m, BC_FRAME)) return;
  int aa, bb;
  frame_cell(m, BC_FRAME, t, aa, bb);
  const long c = aa + (long)m.ni * bb;
#pragma unroll
  for (int e = 0; e < 5; ++e) {
    Dual<1> x(a[c + e * m.ncell]);
    x.d[0] = ad[c + e * m.ncell];
    out[c + e * m.ncell] = x;
  }
}
__global__ void k_unpack_frame(Mesh m, const Dual<1>* __restrict__ x, double* __restrict__ a,
                               double* __restrict__ ad) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= frame_count(m, BC_FRAME)) return;
  int aa, bb;
  frame_cell(m, BC_FRAME, t, aa, bb);
  const long c = aa + (long)m.ni * bb;
#pragma unroll
  for (int e = 0; e < 5; ++e) {
    const Dual<1> v = x[c + e * m.ncell];
    a[c + e * m.ncell] = v.v;
    ad[c + e * m.ncell] = v.d[0];
  }
}

// bc_check + frame_check of a whole BC sequence, before anything is launched
int check_ops_frame(const Mesh& m, const bca_bc_op* ops, int nops, const double* nx, const double* ny) {
  int rc = 0;
  for (int q = 0; q < nops; ++q)
    if ((rc = bc_check(m, ops[q], nx, ny)) || (rc = frame_check(m, ops[q]))) return rc;
  return 0;
}

// work items of a checked BC operation: lines of the interface, or cells of
// the join box
long bc_items(const bca_bc_op& op) {
  const Box b{op.box[0], op.box[1], op.box[2], op.box[3]};
  if (op.kind == BCA_BC_JOIN) return join_count(Box{op.box2[0], op.box2[1], op.box2[2], op.box2[3]}, 1, 2);
  return decode_side(op.loc, b.imin, b.jmin, b.imax, b.jmax).lmax;
}

// FR: w holds ncolor states in the compact BC_FRAME layout (StateView<T, true>)
template <class T, bool FR = false>
int launch_bc(bca_ctx* c, T* w, const bca_bc_op& op, const double* nx, const double* ny, int ncolor = 1,
              long cstride = 0) {
  const Mesh& m = c->m;
  hipStream_t s = c->stream;
  const Box b{op.box[0], op.box[1], op.box[2], op.box[3]};
  int rc = bc_check(m, op, nx, ny);
  if (rc) return rc;
  // a compact-frame state holds no cell outside the frame
  if (FR && (rc = frame_check(m, op))) return rc;
  const unsigned nb = grid1(bc_items(op), op.kind == BCA_BC_JOIN ? 256 : 64);
  switch (op.kind) {
    case BCA_BC_INLET:
      k_bc_inlet<T, FR><<<dim3(nb, ncolor), 64, 0, s>>>(m, w, op.loc, b, op.data, op.lm, nx, ny, op.gam, cstride);
      break;
    case BCA_BC_NOREF:
      k_bc_noref<T, FR><<<dim3(nb, ncolor), 64, 0, s>>>(m, w, op.data, op.lm, op.loc, b, nx, ny, op.gam, cstride);
      break;
    case BCA_BC_EXTRAP2:
      k_bc_extrap2<T, FR><<<dim3(nb, ncolor), 64, 0, s>>>(m, w, op.loc, b, 5, cstride);
      break;
    case BCA_BC_WALL_ADIA:
      k_bc_wall<T, FR><<<dim3(nb, ncolor), 64, 0, s>>>(m, w, op.loc, b, op.gam, cstride);
      break;
    default: {
      const Box prd{op.box2[0], op.box2[1], op.box2[2], op.box2[3]};
      k_join<T, FR><<<dim3(nb, ncolor), 256, 0, s>>>(m, w, b, prd, op.tr[0], op.tr[1], 5, m.ni, m.nj,
                                                      FR ? frame_count(m, BC_FRAME) : m.ncell, cstride);
      break;
    }
  }
  BCA_LAUNCHED();
  return 0;
}

}  // namespace

extern "C" {

const char* bca_last_error(void) { return g_err.c_str(); }

int bca_ctx_create(bca_ctx** out, int im, int jm, int gh, void* stream) {
  if (!out) return fail(-1, "null ctx pointer");
  bca_ctx* c = new bca_ctx();
  c->m = make_mesh(im, jm, gh);
  c->stream = (hipStream_t)stream;
  int rc = check_mesh(c->m);
  if (rc) { delete c; return rc; }
  *out = c;
  return 0;
}

int bca_ctx_destroy(bca_ctx* c) {
  if (!c) return 0;
  for (int k = 0; k < 3; ++k) {
    if (c->side[k]) (void)hipStreamDestroy(c->side[k]);
    if (c->join_ev[k]) (void)hipEventDestroy(c->join_ev[k]);
  }
  if (c->fork_ev) (void)hipEventDestroy(c->fork_ev);
  for (hipEvent_t e : c->timer.pool) (void)hipEventDestroy(e);
  free_work(c->w0);
  free_work(c->w1);
  if (c->norm_part) (void)hipFree(c->norm_part);
  if (c->norm_out) (void)hipFree(c->norm_out);
  for (double* p : {c->blki, c->blkj, c->ghostT, c->vjpi, c->vjpj})
    if (p) (void)hipFree(p);
  if (c->wcol) (void)hipFree(c->wcol);
  if (c->ghostnz) (void)hipFree(c->ghostnz);
  for (int* p : {c->csr_cnt, c->csr_aux})
    if (p) (void)hipFree(p);
  if (c->csr_mask) (void)hipFree(c->csr_mask);
  if (c->jr_mask) (void)hipFree(c->jr_mask);
  if (c->jr_cnt) (void)hipFree(c->jr_cnt);
  if (c->jbeta_comp) (void)hipFree(c->jbeta_comp);
  if (c->wadj) (void)hipFree(c->wadj);
  if (c->ghostG) (void)hipFree(c->ghostG);
  delete c;
  return 0;
}

int bca_ctx_set_stream(bca_ctx* c, void* stream) {
  if (!c) return fail(-1, "null ctx");
  c->stream = (hipStream_t)stream;
  return 0;
}

long bca_ctx_scratch_bytes(const bca_ctx* c) { return c ? c->bytes : 0; }

int bca_profile_enable(bca_ctx* c, int on) {
  if (!c) return fail(-1, "null ctx");
  BCA_HIP(hipStreamSynchronize(c->stream));
  c->timer.on = on != 0;
  c->timer.used = 0;
  c->timer.marks.clear();
  return 0;
}

int bca_profile_read(bca_ctx* c, const char* kernel, double* total_ms, long* launches) {
  if (!c || !kernel || !total_ms || !launches) return fail(-1, "null argument");
  BCA_HIP(hipStreamSynchronize(c->stream));
  double tot = 0.0;
  long n = 0;
  for (const auto& mk : c->timer.marks) {
    if (mk.first != kernel) continue;
    float ms = 0.f;
    BCA_HIP(hipEventElapsedTime(&ms, c->timer.pool[mk.second], c->timer.pool[mk.second + 1]));
    tot += ms;
    ++n;
  }
  *total_ms = tot;
  *launches = n;
  return 0;
}

int bca_geom(bca_ctx* c, double* x0, double* y0, double* nx, double* ny, double* xc, double* yc,
             double* vol, double* volf) {
  if (!c || !x0 || !y0 || !nx || !ny || !xc || !yc || !vol || !volf) return fail(-1, "null argument");
  k_geom<<<1, 1024, 0, c->stream>>>(c->m, x0, y0, nx, ny, xc, yc, vol, volf);
  BCA_LAUNCHED();
  return 0;
}

int bca_volf(bca_ctx* c, const double* vol, double* volf) {
  if (!c || !vol || !volf) return fail(-1, "null argument");
  const dim3 b(64, 4);
  k_volf<<<grid2(c->m.im + 1, c->m.jm + 1, b), b, 0, c->stream>>>(c->m, vol, volf);
  BCA_LAUNCHED();
  return 0;
}

// the scheme of order ord on the context's mesh: a known order, ghost
// layers for its widest stencil, rows for its near-wall closures
static int order_ok(const bca_ctx* c, int ord, bool nowall) {
  if (ord != 3 && ord != 5 && ord != 7 && ord != 9) return fail(-1, "DNC order must be 3, 5, 7 or 9");
  if (ord == 5) return need_gh3(c);
  if (c->m.gh < (ord + 1) / 2)
    return fail(-1, "the DNC order " + std::to_string(ord) + " needs gh >= " + std::to_string((ord + 1) / 2));
  const int rows = ord == 7 ? 9 : ord == 9 ? 11 : 5;
  if (!nowall && c->m.jm < rows)
    return fail(-1, "the wall closure of DNC order " + std::to_string(ord) + " reaches j = " + std::to_string(rows));
  return 0;
}

static int residual_impl(bca_ctx* c, double* residu, const double* w, bca_metrics g, const bca_phys* p,
                         bool nowall, int ord) {
  if (!c || !residu || !w || !p || !g.nx || !g.ny || !g.vol || !g.volf) return fail(-1, "null argument");
  int rc = order_ok(c, ord, nowall);
  if (rc) return rc;
  if ((rc = ensure_work(c, c->w0, false))) return rc;
  return run_pipeline<double>(c, c->w0, w, residu, to_metric(g), to_phys(p), nowall, ord);
}

int bca_residual_dnc5(bca_ctx* c, double* residu, const double* w, bca_metrics g, const bca_phys* p) {
  return residual_impl(c, residu, w, g, p, false, 5);
}

}  // extern "C"

// ---------------------------------------------------------- no-allocation residual
// Scratch of bca_residual_dnc5_noalloc: the prim, grad and hn planes of the
// residual pipeline, in one file-scope device array that every GPU the
// library runs on holds until the process ends.  Host code never passes the
// array's name to a kernel (on the host it is only a stand-in): it passes
// the device address hipGetSymbolAddress returned for the current device.
static __device__ __align__(256) double scratch[BCA_NOALLOC_SCRATCH_BYTES / sizeof(double)];

namespace {

constexpr int NOALLOC_MAX_DEVICES = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[NOALLOC_MAX_DEVICES] = {};

// Device address of `scratch` on the current device, looked up once per
// device (with that device current) and kept.
int noalloc_scratch(double** out) {
  int dev = -1;
  BCA_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) return fail(-1, "device ordinal out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    BCA_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_scratch_ptr[dev] = static_cast<double*>(p);
  }
  *out = g_scratch_ptr[dev];
  return 0;
}

}  // namespace

extern "C" {

long bca_residual_noalloc_bytes(int im, int jm, int gh) {
  if (gh != 3 || im < 5 || jm < 5) return -1;
  const Mesh m = make_mesh(im, jm, gh);
  return (long)sizeof(double) * ((NPRIM + NGRAD) * m.ncell + 10 * m.nnode);
}

int bca_residual_dnc5_noalloc(double* residu, const double* w, bca_metrics g, const bca_phys* p, int im, int jm,
                              int gh, void* stream) {
  if (!residu || !w || !p || !g.nx || !g.ny || !g.vol || !g.volf) return fail(-1, "null argument");
  const Mesh m = make_mesh(im, jm, gh);
  int rc = check_mesh(m);
  if (rc) return rc;
  if (gh != 3) return fail(-1, "the 5th-order scheme needs gh = 3");
  const long need = bca_residual_noalloc_bytes(im, jm, gh);
  if (need > (long)BCA_NOALLOC_SCRATCH_BYTES)
    return fail(-4, "bca_residual_dnc5_noalloc: the mesh needs " + std::to_string(need) +
                        " bytes of scratch, the library's static scratch holds " +
                        std::to_string((long)BCA_NOALLOC_SCRATCH_BYTES));
  double* base = nullptr;
  if ((rc = noalloc_scratch(&base))) return rc;
  double* prim = base;
  double* grad = prim + NPRIM * m.ncell;
  double* hn = grad + NGRAD * m.ncell;
  const hipStream_t s = (hipStream_t)stream;
  const Metric mg = to_metric(g);
  const Phys ph = to_phys(p);
  // the kernels and launch shapes of bca_residual_dnc5 (every scratch value
  // they read is written earlier in the same call, so nothing is cleared)
  k_prim<double><<<grid1(m.ncell, 256), 256, 0, s>>>(m, ph, w, prim);
  BCA_LAUNCHED();
  const dim3 b(64, 4);
  k_grad<double><<<grid2(m.im, m.jm, b), b, 0, s>>>(m, mg, prim, grad);
  BCA_LAUNCHED();
  k_grad_ghost<double><<<grid1(2 * m.im + m.jm, 256), 256, 0, s>>>(m, grad);
  BCA_LAUNCHED();
  const dim3 bf(64, 4);
  k_face<double><<<grid2(m.im + 1, m.jm + 1, bf), bf, 0, s>>>(m, ph, mg, w, prim, grad, hn);
  BCA_LAUNCHED();
  k_balance<double><<<grid2(m.im, m.jm, b), b, 0, s>>>(m, hn, residu);
  BCA_LAUNCHED();
  return 0;
}

static int jvp_impl(bca_ctx* c, double* residu, double* residud, const double* w, const double* wd,
                    bca_metrics g, const bca_phys* p, bool nowall, int ord) {
  if (!c || !w || !wd || !p || !g.nx || !g.ny || !g.vol || !g.volf) return fail(-1, "null argument");
  int rc = order_ok(c, ord, nowall);
  if (rc) return rc;
  if ((rc = ensure_work(c, c->w1, true))) return rc;
  const Mesh& m = c->m;
  // the pipeline up to the face fluxes, reading (w, wd) as Dual<1> where it
  // reads the state (no packed copy), then the balance straight into
  // residu / residud
  if ((rc = run_pipeline<Dual<1>, SplitState>(c, c->w1, SplitState{w, wd}, nullptr, to_metric(g), to_phys(p),
                                               nowall, ord)))
    return rc;
  const dim3 b(64, 4);
  {
    Timed t(c, "k_balance<" + std::string(tname<Dual<1>>()) + ">");
    k_balance_split<1><<<grid2(m.im, m.jm, b), b, 0, c->stream>>>(m, c->w1.hn, residu, residud);
  }
  BCA_LAUNCHED();
  return 0;
}

int bca_jvp_dnc5(bca_ctx* c, double* residu, double* residud, const double* w, const double* wd,
                 bca_metrics g, const bca_phys* p) {
  return jvp_impl(c, residu, residud, w, wd, g, p, false, 5);
}

int bca_residual_dnc5_nowall(bca_ctx* c, double* residu, const double* w, bca_metrics g, const bca_phys* p) {
  return residual_impl(c, residu, w, g, p, true, 5);
}

// the reference's tangents of the variants leave residu as given
int bca_jvp_dnc5_nowall(bca_ctx* c, double* residu, double* residud, const double* w, const double* wd,
                        bca_metrics g, const bca_phys* p) {
  (void)residu;
  if (!residud) return fail(-1, "null argument");
  return jvp_impl(c, nullptr, residud, w, wd, g, p, true, 5);
}

int bca_residual_dnc3(bca_ctx* c, double* residu, const double* w, bca_metrics g, const bca_phys* p) {
  return residual_impl(c, residu, w, g, p, false, 3);
}

int bca_residual_dnc3_nowall(bca_ctx* c, double* residu, const double* w, bca_metrics g, const bca_phys* p) {
  return residual_impl(c, residu, w, g, p, true, 3);
}

int bca_jvp_dnc3(bca_ctx* c, double* residu, double* residud, const double* w, const double* wd, bca_metrics g,
                 const bca_phys* p) {
  (void)residu;
  if (!residud) return fail(-1, "null argument");
  return jvp_impl(c, nullptr, residud, w, wd, g, p, false, 3);
}

int bca_jvp_dnc3_nowall(bca_ctx* c, double* residu, double* residud, const double* w, const double* wd,
                        bca_metrics g, const bca_phys* p) {
  (void)residu;
  if (!residud) return fail(-1, "null argument");
  return jvp_impl(c, nullptr, residud, w, wd, g, p, true, 3);
}

int bca_residual_dnc(bca_ctx* c, int order, int nowall, double* residu, const double* w, bca_metrics g,
                     const bca_phys* p) {
  return residual_impl(c, residu, w, g, p, nowall != 0, order);
}

int bca_jvp_dnc(bca_ctx* c, int order, int nowall, double* residu, double* residud, const double* w,
                const double* wd, bca_metrics g, const bca_phys* p) {
  (void)residu;
  if (!residud) return fail(-1, "null argument");
  return jvp_impl(c, nullptr, residud, w, wd, g, p, nowall != 0, order);
}

static int bc_entry(bca_ctx* c, double* w, double* wd, const bca_bc_op& op, const double* nx,
                    const double* ny) {
  if (!c || !w) return fail(-1, "null argument");
  if (!wd) return launch_bc<double>(c, w, op, nx, ny);
  int rc = ensure_work(c, c->w1, true);
  if (rc) return rc;
  if ((rc = bc_check(c->m, op, nx, ny))) return rc;
  // a BC on the block edge reads and writes the boundary frame only: pack
  // and unpack that (1/50 of the array at 2048 x 1024), else the whole array
  if (op_in_frame(c->m, op)) {
    const unsigned nb = grid1(frame_count(c->m, BC_FRAME), 256);
    k_pack_frame<<<nb, 256, 0, c->stream>>>(c->m, w, wd, c->w1.w);
    BCA_LAUNCHED();
    if ((rc = launch_bc<Dual<1>>(c, c->w1.w, op, nx, ny))) return rc;
    k_unpack_frame<<<nb, 256, 0, c->stream>>>(c->m, c->w1.w, w, wd);
    BCA_LAUNCHED();
    return 0;
  }
  const long n = 5 * c->m.ncell;
  k_pack<1><<<grid1(n, 256), 256, 0, c->stream>>>(n, w, wd, c->w1.w);
  BCA_LAUNCHED();
  if ((rc = launch_bc<Dual<1>>(c, c->w1.w, op, nx, ny))) return rc;
  k_unpack<1><<<grid1(n, 256), 256, 0, c->stream>>>(n, c->w1.w, w, wd);
  BCA_LAUNCHED();
  return 0;
}

int bca_bc_inlet(bca_ctx* c, double* w, double* wd, int loc, const int box[4], const double* field,
                 int lm, const double* nx, const double* ny, double gam) {
  if (!box) return fail(-1, "null box");
  bca_bc_op op{};
  op.kind = BCA_BC_INLET; op.loc = loc; std::memcpy(op.box, box, sizeof(op.box));
  op.data = field; op.lm = lm; op.gam = gam;
  if (c && lm < (loc <= 1 ? box[3] : box[2])) return fail(-1, "field shorter than the interface");
  return bc_entry(c, w, wd, op, nx, ny);
}

int bca_bc_noref(bca_ctx* c, double* w, double* wd, const double* wbd, int lm, int loc, const int box[4],
                 const double* nx, const double* ny, double gam) {
  if (!box) return fail(-1, "null box");
  bca_bc_op op{};
  op.kind = BCA_BC_NOREF; op.loc = loc; std::memcpy(op.box, box, sizeof(op.box));
  op.data = wbd; op.lm = lm; op.gam = gam;
  return bc_entry(c, w, wd, op, nx, ny);
}

int bca_bc_extrap2(bca_ctx* c, double* w, double* wd, int loc, const int box[4], int em) {
  if (!box) return fail(-1, "null box");
  if (em != 5) return fail(-1, "only em = 5 is supported");
  bca_bc_op op{};
  op.kind = BCA_BC_EXTRAP2; op.loc = loc; std::memcpy(op.box, box, sizeof(op.box));
  return bc_entry(c, w, wd, op, nullptr, nullptr);
}

int bca_bc_wall_adia(bca_ctx* c, double* w, double* wd, int loc, const int box[4], double gam) {
  if (!box) return fail(-1, "null box");
  bca_bc_op op{};
  op.kind = BCA_BC_WALL_ADIA; op.loc = loc; std::memcpy(op.box, box, sizeof(op.box)); op.gam = gam;
  return bc_entry(c, w, wd, op, nullptr, nullptr);
}

int bca_bc_join(bca_ctx* c, double* a, const int prr[4], const int prd[4], const int tr[2], int em,
                int ni_arr, int nj_arr) {
  if (!c || !a || !prr || !prd || !tr) return fail(-1, "null argument");
  const Mesh& m = c->m;
  const int lo = 1 - m.gh, hi_i = ni_arr - m.gh, hi_j = nj_arr - m.gh;
  for (const int* bx : {prr, prd})
    if (bx[0] < lo || bx[2] > hi_i || bx[1] < lo || bx[3] > hi_j || bx[0] > bx[2] || bx[1] > bx[3])
      return fail(-1, "join box outside the array");
  const int idir = tr[0] >= 0 ? tr[0] : -tr[0], jdir = tr[1] >= 0 ? tr[1] : -tr[1];
  if (idir < 1 || idir > 2 || jdir < 1 || jdir > 2 || idir == jdir) return fail(-1, "bad transform");
  const int pd[2][2] = {{prd[0], prd[2]}, {prd[1], prd[3]}};
  const int cnt_i = pd[idir - 1][1] - pd[idir - 1][0] + 1, cnt_j = pd[jdir - 1][1] - pd[jdir - 1][0] + 1;
  if (prr[0] + cnt_i - 1 > hi_i || prr[1] + cnt_j - 1 > hi_j) return fail(-1, "join receiver too small");
  const Box rb{prr[0], prr[1], prr[2], prr[3]}, db{prd[0], prd[1], prd[2], prd[3]};
  const long plane = (long)ni_arr * nj_arr;
  k_join<double><<<grid1((long)cnt_i * cnt_j, 256), 256, 0, c->stream>>>(m, a, rb, db, tr[0], tr[1], em,
                                                                          ni_arr, nj_arr, plane, 0);
  BCA_LAUNCHED();
  return 0;
}

static int norms(bca_ctx* c, const double* rhs, double* out10, bool inf) {
  if (!c || !rhs || !out10) return fail(-1, "null argument");
  if (!c->norm_part) {
    BCA_HIP(hipMalloc((void**)&c->norm_part, sizeof(double) * 10 * NORM_BLOCKS));
    BCA_HIP(hipMalloc((void**)&c->norm_out, sizeof(double) * 10));
    c->bytes += sizeof(double) * (10 * NORM_BLOCKS + 10);
  }
  k_norm_partial<<<NORM_BLOCKS, NORM_THREADS, 0, c->stream>>>(c->m, rhs, c->norm_part);
  BCA_LAUNCHED();
  k_norm_final<<<1, 64, 0, c->stream>>>(c->norm_part, NORM_BLOCKS, c->norm_out);
  BCA_LAUNCHED();
  double h[10];
  BCA_HIP(hipMemcpyAsync(h, c->norm_out, sizeof(h), hipMemcpyDeviceToHost, c->stream));
  BCA_HIP(hipStreamSynchronize(c->stream));
  for (int e = 0; e < 5; ++e) {
    out10[e] = h[e];
    out10[5 + e] = inf ? h[5 + e] : h[e] / ((double)c->m.im * c->m.jm);
  }
  return 0;
}

int bca_norml2inf(bca_ctx* c, const double* rhs, double* out10) { return norms(c, rhs, out10, true); }
int bca_norml2(bca_ctx* c, const double* rhs, double* out10) { return norms(c, rhs, out10, false); }

int bca_seed(bca_ctx* c, double* wd, int mv, int l, int k) {
  if (!c || !wd) return fail(-1, "null argument");
  const int step = 2 * c->m.gh + 1;
  if (mv < 0 || mv > 4 || l < 0 || l >= step || k < 0 || k >= step) return fail(-1, "seed index out of range");
  k_testvector<<<grid1(5 * c->m.ncell, 256), 256, 0, c->stream>>>(c->m, wd, mv, l, k);
  BCA_LAUNCHED();
  return 0;
}

int bca_scatter_jv(bca_ctx* c, double* jac, int* ia, int* ja, const double* resd, int mv, int l, int k,
                   const double* coefdiag, int withjn, long nbentry) {
  if (!c || !jac || !ia || !ja || !resd) return fail(-1, "null argument");
  const Mesh& m = c->m;
  const int step = 2 * m.gh + 1;
  if (mv < 0 || mv > 4 || l < 0 || l >= step || k < 0 || k >= step) return fail(-1, "seed index out of range");
  if (nbentry < (long)m.im * m.jm * step * step * 25) return fail(-1, "nbentry smaller than im*jm*(2gh+1)^2*25");
  k_scatter_jv<<<grid1(5L * m.im * m.jm, 256), 256, 0, c->stream>>>(m, jac, ia, ja, resd, mv, l, k,
                                                                     coefdiag, withjn);
  BCA_LAUNCHED();
  return 0;
}

int bca_seed_partial(bca_ctx* c, double* wd, int mv, int l, int k, int istart, int iend, int jstart, int jend) {
  if (!c || !wd) return fail(-1, "null argument");
  const int step = 2 * c->m.gh + 1;
  if (mv < 0 || mv > 4 || l < 0 || l >= step || k < 0 || k >= step) return fail(-1, "seed index out of range");
  k_testvector_partial<<<grid1(5 * c->m.ncell, 256), 256, 0, c->stream>>>(c->m, wd, mv, l, k, istart + 1,
                                                                           iend + 1, jstart + 1, jend + 1);
  BCA_LAUNCHED();
  return 0;
}

int bca_scatter_jv_2zones(bca_ctx* c, double* jac, int* ia, int* ja, const double* resd, int mv, int l, int k,
                          const double* coefdiag, double mini, int n, long nbentry) {
  if (!c || !jac || !ia || !ja || !resd) return fail(-1, "null argument");
  const Mesh& m = c->m;
  const int step = 2 * m.gh + 1;
  if (mv < 0 || mv > 4 || l < 0 || l >= step || k < 0 || k >= step) return fail(-1, "seed index out of range");
  if (n < 0 || n > 1) return fail(-1, "zone index must be 0 or 1");
  if (nbentry < 2L * m.im * m.jm * step * step * 25) return fail(-1, "nbentry smaller than 2*im*jm*(2gh+1)^2*25");
  k_scatter_jv_zone<<<grid1(5L * m.im * m.jm, 256), 256, 0, c->stream>>>(m, jac, ia, ja, resd, mv, l, k, coefdiag,
                                                                          mini, n);
  BCA_LAUNCHED();
  return 0;
}

int bca_set_bndbl(bca_ctx* c, const double* w, double* field, double* wbd) {
  if (!c || !w || !field || !wbd) return fail(-1, "null argument");
  const Mesh& m = c->m;
  k_set_bndbl<<<grid1((long)m.jm * m.gh + m.im + m.gh, 256), 256, 0, c->stream>>>(m, w, field, wbd);
  BCA_LAUNCHED();
  return 0;
}

}  // extern "C"

namespace {

// T(zone, color, ghost) and its nonzero flags: the tangent the linearised BC
// sequence gives every ghost cell under each seed color of each zone, all of
// them per launch, on the boundary frame only.
int ghost_tangents(bca_ctx* c, const double* w, const bca_metrics& g, const bca_bc_op* ops, int nops,
                   const Zones& z) {
  const Mesh& m = c->m;
  const int step = 2 * m.gh + 1, ncol = step * step, nslot = z.n * ncol;
  const long ng = ghost_count(m);
  int rc = 0;
  if (c->zone_cap < z.n) {
    c->bytes -= (long)c->zone_cap * ncol *
                (ng * 25 * (long)sizeof(double) + 5 * frame_count(m, BC_FRAME) * (long)sizeof(Dual<5>) + ng);
    if ((rc = realloc_planes(c, &c->ghostT, (long)nslot * ng * 25))) return rc;
    if ((rc = realloc_planes(c, &c->wcol, (long)nslot * 5 * frame_count(m, BC_FRAME)))) return rc;
    if ((rc = realloc_planes(c, &c->ghostnz, (long)nslot * ng))) return rc;
    c->zone_cap = z.n;
  }
  hipStream_t s = c->stream;
  Timed t(c, "ghost_tangents");
  const long cs = 5 * frame_count(m, BC_FRAME);
  k_seed_frames<<<dim3(grid1(frame_count(m, BC_FRAME), 256), nslot), 256, 0, s>>>(m, z, w, c->wcol, cs);
  BCA_LAUNCHED();
  for (int q = 0; q < nops; ++q)
    if ((rc = launch_bc<Dual<5>, true>(c, c->wcol, ops[q], g.nx, g.ny, nslot, cs))) return rc;
  k_ghost_tangents<<<dim3(grid1(ng, 256), nslot), 256, 0, s>>>(m, c->wcol, cs, c->ghostT, c->ghostnz);
  BCA_LAUNCHED();
  return 0;
}

// Phase 1 of the colored Jacobian: primitives and gradients of the state,
// face-local blocks, ghost tangents of all zones x 49 colors.
// Side streams of the edge-row families, created on first use outside a
// stream capture (default priority, non-blocking).  False: run them on
// c->stream.
bool edge_streams(bca_ctx* c) {
  if (c->side[0]) return true;
  hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(c->stream, &st) != hipSuccess || st != hipStreamCaptureStatusNone) return false;
  bool ok = hipEventCreateWithFlags(&c->fork_ev, hipEventDisableTiming) == hipSuccess;
  for (int k = 0; k < 3 && ok; ++k)
    ok = hipStreamCreateWithFlags(&c->side[k], hipStreamNonBlocking) == hipSuccess &&
         hipEventCreateWithFlags(&c->join_ev[k], hipEventDisableTiming) == hipSuccess;
  if (!ok) {
    for (int k = 0; k < 3; ++k) {
      if (c->side[k]) (void)hipStreamDestroy(c->side[k]);
      if (c->join_ev[k]) (void)hipEventDestroy(c->join_ev[k]);
      c->side[k] = nullptr;
      c->join_ev[k] = nullptr;
    }
    if (c->fork_ev) (void)hipEventDestroy(c->fork_ev);
    c->fork_ev = nullptr;
    (void)hipGetLastError();
  }
  return ok;
}

// The streams of the edge-row families: the side streams, made to wait for
// everything queued on c->stream so far (the primal they read), or c->stream
// itself.
int edge_fork(bca_ctx* c, hipStream_t es[3]) {
  const bool par = edge_streams(c);
  for (int k = 0; k < 3; ++k) es[k] = par ? c->side[k] : c->stream;
  if (!par) return 0;
  BCA_HIP(hipEventRecord(c->fork_ev, c->stream));
  for (int k = 0; k < 3; ++k) BCA_HIP(hipStreamWaitEvent(c->side[k], c->fork_ev, 0));
  return 0;
}

// c->stream waits for the edge-row families (before anything reads their
// blocks, and before the call returns).
int edge_join(bca_ctx* c, const hipStream_t es[3]) {
  for (int k = 0; k < 3; ++k) {
    if (es[k] == c->stream) continue;
    BCA_HIP(hipEventRecord(c->join_ev[k], es[k]));
    BCA_HIP(hipStreamWaitEvent(c->stream, c->join_ev[k], 0));
  }
  return 0;
}

int jacobian_phase1(bca_ctx* c, const double* w, bca_metrics g, const bca_phys* p, const bca_bc_op* ops, int nops,
                    const Zones& z) {
  if (!c || !w || !p) return fail(-1, "null argument");
  if (!g.nx || !g.ny || !g.vol || !g.volf) return fail(-1, "null metrics");
  if (nops < 0 || (nops > 0 && !ops)) return fail(-1, "bad boundary-condition list");
  const Mesh& m = c->m;
  int rc = 0;
  if ((rc = check_ops_frame(m, ops, nops, g.nx, g.ny))) return rc;
  if ((rc = ensure_work(c, c->w0, false))) return rc;
  const long nfi = (long)(m.im + 1) * m.jm, nfj = (long)m.im * (m.jm + 1);
  if ((rc = alloc_planes(c, &c->blki, (long)NOFF * 25 * nfi))) return rc;
  if ((rc = alloc_planes(c, &c->blkj, (long)NOFF * 25 * nfj))) return rc;
  const Metric gm = to_metric(g);
  const Phys ph = to_phys(p);
  hipStream_t s = c->stream;
  const dim3 b(64, 4);
  // primal primitives and gradients of the state (ghost frame as given)
  k_prim<double><<<grid1(m.ncell, 256), 256, 0, s>>>(m, ph, w, c->w0.prim);
  BCA_LAUNCHED();
  k_grad<double><<<grid2(m.im, m.jm, b), b, 0, s>>>(m, gm, c->w0.prim, c->w0.grad);
  BCA_LAUNCHED();
  k_grad_ghost<double><<<grid1(2 * m.im + m.jm, 256), 256, 0, s>>>(m, c->w0.grad);
  BCA_LAUNCHED();
  // phase 1: face-local blocks (independent of the zones).  The edge rows
  // (i-face rows 1-2, j-face rows 1-3) write face rows the interior launches
  // do not, so they run on the side streams meanwhile, and so do the ghost
  // tangents.
  hipStream_t es[3];
  if ((rc = edge_fork(c, es))) return rc;
  BlockArgs ae{m, ph, gm, w, c->w0.prim, c->w0.grad, c->blki, nullptr, 1, 2};
  launch_iface_blocks_o2(grid2(m.im + 1, 2, b), b, es[0], ae);
  ae.out = c->blkj;
  ae.j0 = ae.j1 = 3;
  launch_jface_blocks_r3(grid2(m.im, 1, b), b, es[1], ae);
  ae.j0 = ae.j1 = 2;
  launch_jface_blocks_r2(grid2(m.im, 1, b), b, es[2], ae);
  ae.j0 = ae.j1 = 1;
  launch_jface_blocks_wall(grid2(m.im, 1, b), b, es[0], ae);
  BCA_LAUNCHED();
  // the ghost tangents read w only and write their own planes: on side
  // stream 1 too (its launches go to c->stream, pointed there meanwhile)
  const hipStream_t main = c->stream;
  c->stream = es[1];
  rc = ghost_tangents(c, w, g, ops, nops, z);
  c->stream = main;
  if (rc) {
    (void)edge_join(c, es);
    return rc;
  }
  BlockArgs ai{m, ph, gm, w, c->w0.prim, c->w0.grad, c->blki, nullptr, 3, m.jm};
  const dim3 bt = c->tune_blocks;
  {
    // every second offset on side stream 0 (free once its edge rows are
    // done); the timers cover the offsets on c->stream
    Timed t(c, "k_iface_blocks");
    launch_iface_blocks_o4(grid2(m.im + 1, m.jm - 2, bt), bt, s, ai, es[0] == s ? nullptr : es[0]);
  }
  BlockArgs aj{m, ph, gm, w, c->w0.prim, c->w0.grad, c->blkj, nullptr, 4, m.jm + 1};
  {
    // likewise every second j-face offset on side stream 2: the one-offset
    // launches (tens of microseconds each) leave most of the chip idle
    // alone, so four of them run side by side
    Timed t(c, "k_jface_blocks");
    launch_jface_blocks_c(grid2(m.im, m.jm - 2, bt), bt, s, aj, es[2] == s ? nullptr : es[2]);
  }
  BCA_LAUNCHED();
  return edge_join(c, es);
}

int jacobian_impl(bca_ctx* c, const double* w, bca_metrics g, const bca_phys* p, const bca_bc_op* ops, int nops,
                  const double* coefdiag, const Zones& z, double* jac, int* ia, int* ja, long nbentry,
                  int write_index) {
  if (!c || !jac || (write_index && (!ia || !ja))) return fail(-1, "null argument");
  const Mesh& m = c->m;
  const int step = 2 * m.gh + 1;
  if (nbentry < (long)z.n * m.im * m.jm * step * step * 25)
    return fail(-1, "nbentry smaller than nzones*im*jm*(2gh+1)^2*25");
  // 32-bit face-block offsets in the assembly
  if ((long)(NOFF + 1) * 25 * std::max((long)(m.im + 1) * m.jm, (long)m.im * (m.jm + 1)) > 2147483647L)
    return fail(-1, "mesh too large for 32-bit face-block offsets");
  int rc = jacobian_phase1(c, w, g, p, ops, nops, z);
  if (rc) return rc;
  const int nslot = z.n * step * step;
  hipStream_t s = c->stream;
  const dim3 b(64, 4);
  // phase 2: rows
  {
    Timed t(c, "k_assemble_colored");
    const dim3 ba = c->tune_assemble;
    dim3 gr = grid2(m.im, m.jm, ba);
    gr.z = step;
    if (z.n == 1)
      k_assemble_colored<1><<<gr, ba, 0, s>>>(m, z, c->blki, c->blkj, coefdiag, jac, ia, ja, write_index);
    else
      k_assemble_colored<2><<<gr, ba, 0, s>>>(m, z, c->blki, c->blkj, coefdiag, jac, ia, ja, write_index);
  }
  BCA_LAUNCHED();
  {
    Timed t(c, "k_assemble_ghosts");
    k_assemble_ghosts<<<dim3(grid1(near_count(m), 256), nslot * 5), 256, 0, s>>>(m, z, c->blki, c->blkj, c->ghostT,
                                                                       c->ghostnz, jac);
  }
  BCA_LAUNCHED();
  return 0;
}

}  // namespace

extern "C" {

int bca_jacobian_colored(bca_ctx* c, const double* w, bca_metrics g, const bca_phys* p,
                         const bca_bc_op* ops, int nops, const double* coefdiag, int withjn,
                         double* jac, int* ia, int* ja, long nbentry, int write_index) {
  if (!c) return fail(-1, "null argument");
  if (int rc_gh = need_gh3(c)) return rc_gh;
  const Zones z{1, {1, 0}, {c->m.im, 0}, withjn != 0};
  return jacobian_impl(c, w, g, p, ops, nops, coefdiag, z, jac, ia, ja, nbentry, write_index);
}

int bca_jacobian_colored_2zones(bca_ctx* c, const double* w, bca_metrics g, const bca_phys* p,
                                const bca_bc_op* ops, int nops, const double* coefdiag, int isplit,
                                double* jac, int* ia, int* ja, long nbentry, int write_index) {
  if (!c) return fail(-1, "null argument");
  if (int rc_gh = need_gh3(c)) return rc_gh;
  const int step = 2 * c->m.gh + 1;
  if (isplit < step || c->m.im - isplit < step) return fail(-1, "each zone must be at least 2gh+1 cells wide");
  const Zones z{2, {1, isplit + 1}, {isplit, c->m.im}, 1};
  return jacobian_impl(c, w, g, p, ops, nops, coefdiag, z, jac, ia, ja, nbentry, write_index);
}

int bca_vjp_dnc5(bca_ctx* c, double* residub, double* wb, const double* w, bca_metrics g, const bca_phys* p) {
  if (!c || !residub || !wb || !w || !p || !g.nx || !g.ny || !g.vol || !g.volf) return fail(-1, "null argument");
  if (int rc_gh = need_gh3(c)) return rc_gh;
  const Mesh& m = c->m;
  int rc = 0;
  if ((rc = ensure_work(c, c->w0, false))) return rc;
  const long nfi = (long)(m.im + 1) * m.jm, nfj = (long)m.im * (m.jm + 1);
  if ((rc = alloc_planes(c, &c->vjpi, (long)NOFF * 5 * nfi))) return rc;
  if ((rc = alloc_planes(c, &c->vjpj, (long)NOFF * 5 * nfj))) return rc;
  const Metric gm = to_metric(g);
  const Phys ph = to_phys(p);
  hipStream_t s = c->stream;
  const dim3 b(64, 4);
  // recomputed primal primitives and gradients (no tape)
  k_prim<double><<<grid1(m.ncell, 256), 256, 0, s>>>(m, ph, w, c->w0.prim);
  BCA_LAUNCHED();
  k_grad<double><<<grid2(m.im, m.jm, b), b, 0, s>>>(m, gm, c->w0.prim, c->w0.grad);
  BCA_LAUNCHED();
  k_grad_ghost<double><<<grid1(2 * m.im + m.jm, 256), 256, 0, s>>>(m, c->w0.grad);
  BCA_LAUNCHED();
  // edge rows on the side streams while the interior runs (disjoint face
  // rows of vjpi / vjpj; residub is only read until the join); the interior
  // kernels are queued first, so that a side stream sharing a hardware
  // queue with c->stream does not hold them back
  hipStream_t es[3];
  if ((rc = edge_fork(c, es))) return rc;
  BlockArgs ai{m, ph, gm, w, c->w0.prim, c->w0.grad, c->vjpi, residub, 3, m.jm};
  {
    Timed t(c, "k_iface_vjp");
    launch_iface_vjp_rev(grid2(m.im + 1, m.jm - 2, b), b, s, ai);
  }
  BlockArgs aj{m, ph, gm, w, c->w0.prim, c->w0.grad, c->vjpj, residub, 4, m.jm + 1};
  {
    Timed t(c, "k_jface_vjp");
    launch_jface_vjp_rev(grid2(m.im, m.jm - 2, b), b, s, aj);
  }
  BCA_LAUNCHED();
  BlockArgs ae{m, ph, gm, w, c->w0.prim, c->w0.grad, c->vjpi, residub, 1, 2};
  launch_iface_vjp_rev_o2(grid2(m.im + 1, 2, b), b, es[0], ae);
  ae.out = c->vjpj;
  ae.j0 = ae.j1 = 3;
  launch_jface_vjp_rev_r3(grid2(m.im, 1, b), b, es[1], ae);
  ae.j0 = ae.j1 = 2;
  launch_jface_vjp_rev_r2(grid2(m.im, 1, b), b, es[2], ae);
  ae.j0 = ae.j1 = 1;
  launch_jface_vjp_wall(grid2(m.im, 1, b), b, es[0], ae);
  BCA_LAUNCHED();
  if ((rc = edge_join(c, es))) return rc;
  {
    Timed t(c, "k_vjp_gather");
    k_vjp_gather<<<grid2(m.ni, m.nj, b), b, 0, s>>>(m, c->vjpi, c->vjpj, wb);
  }
  BCA_LAUNCHED();
  k_zero_interior<<<grid2(m.im, m.jm, b), b, 0, s>>>(m, residub);
  BCA_LAUNCHED();
  return 0;
}

}  // extern "C"

// ---- pruned CSR of the colored COO ------------------------------------------
namespace {

__global__ void k_fill_int(int* p, int n, int v) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) p[t] = v;
}

int csr_zones(bca_ctx* c, int nzones, int withjn, long nbentry, Zones& z) {
  if (nzones != 1 && nzones != 2) return fail(-1, "nzones must be 1 or 2");
  const int step = 2 * c->m.gh + 1;
  if (nbentry < (long)nzones * c->m.im * c->m.jm * step * step * 25)
    return fail(-1, "nbentry smaller than nzones*im*jm*(2gh+1)^2*25");
  z = Zones{nzones, {1, 0}, {c->m.im, 0}, withjn != 0 || nzones == 2};
  const long rows = 5L * c->m.im * c->m.jm;
  if (rows * 245L * nzones > 2147483647L) return fail(-1, "matrix too large for 32-bit CSR indices");
  return 0;
}

}  // namespace

extern "C" {

int bca_csr_colored_count(bca_ctx* c, const double* jac, long nbentry, int nzones, int withjn, double mini,
                          int* indptr, long* nnz) {
  if (!c || !jac || !indptr || !nnz) return fail(-1, "null argument");
  Zones z;
  int rc = csr_zones(c, nzones, withjn, nbentry, z);
  if (rc) return rc;
  const Mesh& m = c->m;
  const long rows = 5L * m.im * m.jm;
  const int nb = (int)((rows + SCAN_B - 1) / SCAN_B);
  if ((rc = alloc_planes(c, &c->csr_cnt, rows))) return rc;
  if ((rc = alloc_planes(c, &c->csr_aux, CSR_AUX_INTS))) return rc;
  if ((rc = alloc_planes(c, &c->csr_mask, (long)CSR_MASK_WORDS * rows))) return rc;
  if (nb > SCAN_B * SCAN_B) return fail(-1, "too many rows for the scan");
  int* sums = c->csr_aux;
  int* total = c->csr_aux + (long)SCAN_B * SCAN_B;
  int* bad = total + 1;
  hipStream_t s = c->stream;
  k_fill_int<<<1, 64, 0, s>>>(bad, 1, 0);
  BCA_LAUNCHED();
  const dim3 b(64, 4);
  dim3 gr = grid2(m.im, m.jm, b);
  gr.z = 5;
  {
    Timed t(c, "k_csr_count");
    if (z.n == 1)
      k_csr_count<1, const double*><<<gr, b, 0, s>>>(m, z.withjn, jac, mini, c->csr_cnt, c->csr_mask, bad);
    else
      k_csr_count<2, const double*><<<gr, b, 0, s>>>(m, z.withjn, jac, mini, c->csr_cnt, c->csr_mask, bad);
  }
  BCA_LAUNCHED();
  k_scan_blocks<<<nb, SCAN_B, 0, s>>>(c->csr_cnt, rows, indptr, sums);
  BCA_LAUNCHED();
  k_scan_sums<<<1, SCAN_B, 0, s>>>(sums, nb, total);
  BCA_LAUNCHED();
  k_scan_add<<<nb, SCAN_B, 0, s>>>(indptr, rows, sums, total);
  BCA_LAUNCHED();
  int host[2] = {0, 0};
  BCA_HIP(hipMemcpyAsync(host, total, 2 * sizeof(int), hipMemcpyDeviceToHost, s));
  BCA_HIP(hipStreamSynchronize(s));
  if (host[1]) return fail(-4, "a kept COO entry has a column outside the matrix");
  *nnz = host[0];
  return 0;
}

int bca_csr_colored_fill(bca_ctx* c, const double* jac, long nbentry, int nzones, int withjn, double mini,
                         const int* indptr, int* indices, double* data) {
  if (!c || !jac || !indptr || !indices || !data) return fail(-1, "null argument");
  Zones z;
  int rc = csr_zones(c, nzones, withjn, nbentry, z);
  if (rc) return rc;
  if (!c->csr_aux || !c->csr_mask) return fail(-1, "bca_csr_colored_count must run first");
  const Mesh& m = c->m;
  const dim3 gr((m.im + FILL_TI - 1) / FILL_TI, m.jm, 5);
  {
    Timed t(c, "k_csr_fill");
    if (z.n == 1)
      k_csr_fill<1><<<gr, FILL_THREADS, 0, c->stream>>>(m, z.withjn, jac, mini, c->csr_mask, indptr, indices, data);
    else
      k_csr_fill<2><<<gr, FILL_THREADS, 0, c->stream>>>(m, z.withjn, jac, mini, c->csr_mask, indptr, indices, data);
  }
  BCA_LAUNCHED();
  return 0;
}

}  // extern "C"

// ---- spanwise operators and J(beta) -----------------------------------------
extern "C" {

int bca_dz_5p(bca_ctx* c, double* dz, double* dz2, const double* w, const double* wd, bca_metrics g,
              const bca_phys* p) {
  if (!c || !w || !wd || !p || (!dz && !dz2)) return fail(-1, "null argument");
  if (int rc_gh = need_gh3(c)) return rc_gh;
  if (dz && (!g.nx || !g.ny || !g.vol)) return fail(-1, "null metrics");
  const Mesh& m = c->m;
  const dim3 b(64, 4);
  k_dz<<<grid2(m.im, m.jm, b), b, 0, c->stream>>>(m, to_phys(p), to_metric(g), w, wd, dz, dz2);
  BCA_LAUNCHED();
  return 0;
}

int bca_scatter_dz(bca_ctx* c, double* jac, int* ia, int* ja, const double* dz, int mv, int l, int k,
                   long nbentry) {
  if (!c || !jac || !ia || !ja || !dz) return fail(-1, "null argument");
  const Mesh& m = c->m;
  const int step = 2 * m.gh + 1;
  if (mv < 0 || mv > 4 || l < 0 || l >= step || k < 0 || k >= step) return fail(-1, "seed index out of range");
  if (nbentry < (long)m.im * m.jm * step * step * 25) return fail(-1, "nbentry smaller than im*jm*(2gh+1)^2*25");
  k_scatter_dz<<<grid1(5L * m.im * m.jm, 256), 256, 0, c->stream>>>(m, jac, ia, ja, dz, mv, l, k);
  BCA_LAUNCHED();
  return 0;
}

int bca_dz_colored(bca_ctx* c, const double* w, bca_metrics g, const bca_phys* p, const bca_bc_op* ops, int nops,
                   double* jdz, double* jdz2, int* ia, int* ja, long nbentry, int write_index) {
  if (!c || !w || !p || (!jdz && !jdz2) || (write_index && (!ia || !ja))) return fail(-1, "null argument");
  if (int rc_gh = need_gh3(c)) return rc_gh;
  if (!g.nx || !g.ny || !g.vol) return fail(-1, "null metrics");
  if (nops < 0 || (nops > 0 && !ops)) return fail(-1, "bad boundary-condition list");
  const Mesh& m = c->m;
  const int step = 2 * m.gh + 1;
  if (nbentry < (long)m.im * m.jm * step * step * 25) return fail(-1, "nbentry smaller than im*jm*(2gh+1)^2*25");
  const Zones z{1, {1, 0}, {m.im, 0}, 0};
  int rc = check_ops_frame(m, ops, nops, g.nx, g.ny);
  if (rc) return rc;
  rc = ghost_tangents(c, w, g, ops, nops, z);
  if (rc) return rc;
  const dim3 b(64, 4);
  dim3 gr = grid2(m.im, m.jm, b);
  gr.z = step * step;
  {
    Timed t(c, "k_dz_colored");
    k_dz_colored<<<gr, b, 0, c->stream>>>(m, to_phys(p), to_metric(g), w, c->ghostT, c->ghostnz, jdz, jdz2, ia, ja,
                                          write_index);
  }
  BCA_LAUNCHED();
  return 0;
}

int bca_jbeta_count(bca_ctx* c, const double* jac, const double* jdz, const double* jdz2, long nbentry,
                    double mini, int* indptr, long* nnz) {
  if (!c || !jac || !jdz || !jdz2 || !indptr || !nnz) return fail(-1, "null argument");
  Zones z;
  int rc = csr_zones(c, 1, 0, nbentry, z);
  if (rc) return rc;
  const Mesh& m = c->m;
  const long rows = 5L * m.im * m.jm;
  const int nb = (int)((rows + SCAN_B - 1) / SCAN_B);
  if ((rc = alloc_planes(c, &c->csr_cnt, rows))) return rc;
  if ((rc = alloc_planes(c, &c->csr_aux, CSR_AUX_INTS))) return rc;
  if ((rc = alloc_planes(c, &c->csr_mask, (long)CSR_MASK_WORDS * rows))) return rc;
  int* sums = c->csr_aux;
  int* total = c->csr_aux + (long)SCAN_B * SCAN_B;
  hipStream_t s = c->stream;
  const dim3 b(64, 4);
  dim3 gr = grid2(m.im, m.jm, b);
  gr.z = 5;
  int* bad = total + 1;
  k_fill_int<<<1, 64, 0, s>>>(bad, 1, 0);
  BCA_LAUNCHED();
  k_csr_count<1, MaxAbs3><<<gr, b, 0, s>>>(m, 0, MaxAbs3{jac, jdz, jdz2}, mini, c->csr_cnt, c->csr_mask, bad);
  BCA_LAUNCHED();
  k_scan_blocks<<<nb, SCAN_B, 0, s>>>(c->csr_cnt, rows, indptr, sums);
  BCA_LAUNCHED();
  k_scan_sums<<<1, SCAN_B, 0, s>>>(sums, nb, total);
  BCA_LAUNCHED();
  k_scan_add<<<nb, SCAN_B, 0, s>>>(indptr, rows, sums, total);
  BCA_LAUNCHED();
  int host[2] = {0, 0};
  BCA_HIP(hipMemcpyAsync(host, total, 2 * sizeof(int), hipMemcpyDeviceToHost, s));
  BCA_HIP(hipStreamSynchronize(s));
  if (host[1]) return fail(-4, "a kept COO entry has a column outside the matrix");
  *nnz = host[0];
  return 0;
}

int bca_jbeta_fill(bca_ctx* c, const double* jac, const double* jdz, const double* jdz2, long nbentry, double mini,
                   const double* vol, const double* beta, int nbeta, const int* indptr, int* indices, double* vals,
                   long nnz) {
  if (!c || !jac || !jdz || !jdz2 || !vol || !beta || !indptr || !indices || !vals) return fail(-1, "null argument");
  if (nbeta < 1) return fail(-1, "nbeta must be >= 1");
  if (!c->csr_aux || !c->csr_mask) return fail(-1, "bca_jbeta_count must run first");
  Zones z;
  int rc = csr_zones(c, 1, 0, nbentry, z);
  if (rc) return rc;
  if (c->jbeta_cap < nnz) {
    if ((rc = realloc_planes(c, &c->jbeta_comp, 3 * std::max(nnz, 1L)))) return rc;
    c->jbeta_cap = nnz;
  }
  const Mesh& m = c->m;
  {
    // the tiled CSR fill in its three-array mode: indices and the pruned
    // components (jac / vol(row), dz, dz2) on the pattern
    Timed t(c, "k_jbeta_parts");
    const dim3 gr((m.im + FILL_TI - 1) / FILL_TI, m.jm, 5);
    k_csr_fill<1, true><<<gr, FILL_THREADS, 0, c->stream>>>(m, 0, jac, mini, c->csr_mask, indptr, indices,
                                                            c->jbeta_comp, jdz, jdz2, vol, nnz);
  }
  BCA_LAUNCHED();
  {
    Timed t(c, "k_jbeta_values");
    k_jbeta_values<<<grid1(nnz, 256), 256, 0, c->stream>>>(c->jbeta_comp, nnz, beta, nbeta, vals);
  }
  BCA_LAUNCHED();
  return 0;
}

}  // extern "C"

// ------------------------------------------------------------------ batch
struct bca_batch {
  hipStream_t stream = nullptr;
  void* scratch = nullptr;  // per-mesh planes of the scalar type
  long scratch_bytes = 0;
  void* table = nullptr;    // device copy of the items and the plan
  void* host = nullptr;     // pinned staging of the table
  long table_bytes = 0;
  hipEvent_t copied = nullptr;  // the last table copy (the staging is reused after it)
  bool pending = false;
};

namespace {

int batch_grow(void** p, long* have, long want, bool pinned) {
  if (*have >= want) return 0;
  if (*p) (void)(pinned ? hipHostFree(*p) : hipFree(*p));
  *p = nullptr;
  *have = 0;
  if (pinned) BCA_HIP(hipHostMalloc(p, want, hipHostMallocDefault));
  else BCA_HIP(hipMalloc(p, want));
  *have = want;
  return 0;
}

template <class T>
int run_batch(bca_batch* b, const bca_batch_item* items, int n) {
  if (n < 1 || n > BATCH_MAX) return fail(-1, "batch size must be 1.." + std::to_string(BATCH_MAX));
  // scratch layout and checks
  long planes = 0;
  std::vector<Mesh> ms(n);
  for (int k = 0; k < n; ++k) {
    const bca_batch_item& it = items[k];
    if (!it.w || !it.residu || !it.g.nx || !it.g.ny || !it.g.vol || !it.g.volf)
      return fail(-1, "batch item " + std::to_string(k) + ": null argument");
    if (std::is_same<T, double>::value != (it.wd == nullptr) || (it.wd && !it.residud))
      return fail(-1, "batch items must all carry wd and residud, or none");
    if (it.nops < 0 || it.nops > BATCH_OPS || (it.nops && !it.ops))
      return fail(-1, "batch item " + std::to_string(k) + ": 0.." + std::to_string(BATCH_OPS) + " BC operations");
    ms[k] = make_mesh(it.im, it.jm, 3);
    int rc = check_mesh(ms[k]);
    if (rc) return rc;
    for (int o = 0; o < it.nops; ++o)
      if ((rc = bc_check(ms[k], it.ops[o], it.g.nx, it.g.ny))) return rc;
    planes += (5 + NPRIM + NGRAD) * ms[k].ncell + 10 * ms[k].nnode;
  }
  int rc = batch_grow(&b->scratch, &b->scratch_bytes, planes * (long)sizeof(T), false);
  if (rc) return rc;
  const long tbytes = (long)sizeof(BatchItem<T>) * BATCH_MAX + (long)sizeof(BatchPlan);
  if (b->pending) {
    BCA_HIP(hipEventSynchronize(b->copied));
    b->pending = false;
  }
  long have = b->table_bytes;
  if ((rc = batch_grow(&b->table, &b->table_bytes, tbytes, false))) return rc;
  if ((rc = batch_grow(&b->host, &have, tbytes, true))) return rc;
  BatchItem<T>* hi = reinterpret_cast<BatchItem<T>*>(b->host);
  BatchPlan* plan = reinterpret_cast<BatchPlan*>(hi + BATCH_MAX);
  std::memset(b->host, 0, tbytes);
  T* p = reinterpret_cast<T*>(b->scratch);
  plan->n = n;
  plan->maxops = 0;
  for (int k = 0; k < n; ++k) {
    const bca_batch_item& it = items[k];
    const Mesh& m = ms[k];
    BatchItem<T>& d = hi[k];
    d.m = m;
    d.ph = to_phys(&it.phys);
    d.g = to_metric(it.g);
    d.w = it.w; d.wd = it.wd; d.res = it.residu; d.resd = it.residud;
    d.ws = p; p += 5 * m.ncell;
    d.prim = p; p += NPRIM * m.ncell;
    d.grad = p; p += NGRAD * m.ncell;
    d.hn = p; p += 10 * m.nnode;
    d.nops = it.nops;
    plan->maxops = std::max(plan->maxops, it.nops);
    for (int o = 0; o < it.nops; ++o) {
      const bca_bc_op& op = it.ops[o];
      d.ops[o] = BatchBc{op.kind, op.loc, Box{op.box[0], op.box[1], op.box[2], op.box[3]},
                         Box{op.box2[0], op.box2[1], op.box2[2], op.box2[3]}, op.tr[0], op.tr[1], op.data, op.lm,
                         op.gam};
    }
    plan->cells[k + 1] = plan->cells[k] + m.ncell;
    plan->interior[k + 1] = plan->interior[k] + (long)m.im * m.jm;
    plan->edges[k + 1] = plan->edges[k] + m.im + m.jm;
    plan->faces[k + 1] = plan->faces[k] + (long)(m.im + 1) * (m.jm + 1);
  }
  for (int o = 0; o < BATCH_OPS; ++o)
    for (int k = 0; k < n; ++k)
      plan->bc[o][k + 1] = plan->bc[o][k] + (o < items[k].nops ? bc_items(items[k].ops[o]) : 0);
  BCA_HIP(hipMemcpyAsync(b->table, b->host, tbytes, hipMemcpyHostToDevice, b->stream));
  if (!b->copied) BCA_HIP(hipEventCreateWithFlags(&b->copied, hipEventDisableTiming));
  BCA_HIP(hipEventRecord(b->copied, b->stream));
  b->pending = true;
  // one launch per stage over all meshes' items
  const BatchItem<T>* ditems = reinterpret_cast<const BatchItem<T>*>(b->table);
  const BatchPlan* dplan = reinterpret_cast<const BatchPlan*>(ditems + BATCH_MAX);
  hipStream_t s = b->stream;
  k_batch_stage<T, BS_PACK><<<grid1(plan->cells[n], 256), 256, 0, s>>>(ditems, dplan, 0);
  BCA_LAUNCHED();
  for (int o = 0; o < plan->maxops; ++o) {
    if (plan->bc[o][n] == 0) continue;
    k_batch_stage<T, BS_BC><<<grid1(plan->bc[o][n], 256), 256, 0, s>>>(ditems, dplan, o);
    BCA_LAUNCHED();
  }
  k_batch_stage<T, BS_PRIM><<<grid1(plan->cells[n], 256), 256, 0, s>>>(ditems, dplan, 0);
  BCA_LAUNCHED();
  k_batch_stage<T, BS_GRAD><<<grid1(plan->interior[n], 256), 256, 0, s>>>(ditems, dplan, 0);
  BCA_LAUNCHED();
  k_batch_stage<T, BS_GHOSTGRAD><<<grid1(plan->edges[n], 256), 256, 0, s>>>(ditems, dplan, 0);
  BCA_LAUNCHED();
  k_batch_stage<T, BS_FACE><<<grid1(plan->faces[n], 256), 256, 0, s>>>(ditems, dplan, 0);
  BCA_LAUNCHED();
  k_batch_stage<T, BS_BALANCE><<<grid1(plan->cells[n], 256), 256, 0, s>>>(ditems, dplan, 0);
  BCA_LAUNCHED();
  return 0;
}

}  // namespace

extern "C" {

int bca_batch_create(bca_batch** out, void* stream) {
  if (!out) return fail(-1, "null argument");
  *out = new bca_batch();
  (*out)->stream = (hipStream_t)stream;
  return 0;
}

int bca_batch_set_stream(bca_batch* b, void* stream) {
  if (!b) return fail(-1, "null batch");
  b->stream = (hipStream_t)stream;
  return 0;
}

int bca_batch_destroy(bca_batch* b) {
  if (!b) return 0;
  if (b->pending) (void)hipEventSynchronize(b->copied);
  if (b->copied) (void)hipEventDestroy(b->copied);
  if (b->scratch) (void)hipFree(b->scratch);
  if (b->table) (void)hipFree(b->table);
  if (b->host) (void)hipHostFree(b->host);
  delete b;
  return 0;
}

int bca_batch_residual_jvp(bca_batch* b, const bca_batch_item* items, int n) {
  if (!b || !items) return fail(-1, "null argument");
  if (n >= 1 && items[0].wd) return run_batch<Dual<1>>(b, items, n);
  return run_batch<double>(b, items, n);
}

}  // extern "C"

// ------------------------------------------------------------------ BC adjoints
extern "C" {

int bca_bc_adjoint(bca_ctx* c, const double* w, double* wb, const bca_bc_op* op, const double* nx,
                   const double* ny) {
  if (!c || !w || !wb || !op) return fail(-1, "null argument");
  const Mesh& m = c->m;
  int rc = bc_check(m, *op, nx, ny);
  if (rc) return rc;
  const Box b{op->box[0], op->box[1], op->box[2], op->box[3]};
  if (op->kind == BCA_BC_JOIN) {
    const Box prd{op->box2[0], op->box2[1], op->box2[2], op->box2[3]};
    const bool overlap = b.imin <= prd.imax && prd.imin <= b.imax && b.jmin <= prd.jmax && prd.jmin <= b.jmax;
    if (overlap) return fail(-1, "join adjoint needs disjoint receiving and donor boxes");
    k_join_adjoint<<<grid1(bc_items(*op), 256), 256, 0, c->stream>>>(m, wb, b, prd, 5);
    BCA_LAUNCHED();
    return 0;
  }
  if ((rc = frame_check(m, *op))) return rc;
  // the offsets d = 0..3 in ADJ_REACH slots of wadj, each the compact BC
  // frame
  const long cs = 5 * frame_count(m, BC_FRAME);
  if ((rc = alloc_planes(c, &c->wadj, ADJ_REACH * cs))) return rc;
  const int lmax = (int)bc_items(*op);
  k_adj_seed<<<dim3(grid1(frame_count(m, BC_FRAME), 256), ADJ_REACH), 256, 0, c->stream>>>(m, w, c->wadj, op->loc,
                                                                                          b, cs);
  BCA_LAUNCHED();
  if ((rc = launch_bc<Dual<5>, true>(c, c->wadj, *op, nx, ny, ADJ_REACH, cs))) return rc;
  k_adj_accumulate<<<dim3(grid1(lmax, 64), ADJ_REACH), 64, 0, c->stream>>>(m, c->wadj, wb, op->loc, b, cs, lmax);
  BCA_LAUNCHED();
  k_adj_clear_ghosts<<<grid1(lmax, 64), 64, 0, c->stream>>>(m, wb, op->loc, b, lmax);
  BCA_LAUNCHED();
  return 0;
}

}  // extern "C"

// ------------------------------------------------------------------ fused colored Jacobian -> CSR
namespace {

// The pruned CSR of the colored Jacobian without the COO (k_jac_csr_rows).
// refresh: when the context's record says indptr / indices hold the CSR of an
// earlier call with the same zones and mini, one row pass checks every row's
// kept set against the stored masks and rewrites data in place; only if a row
// changed is the CSR rescanned and refilled.  Otherwise (and always when
// !refresh): count pass, scan, fill pass.
int jac_csr(bca_ctx* c, const double* w, bca_metrics g, const bca_phys* p, const bca_bc_op* ops, int nops,
            const double* coefdiag, int nzones, int zparam, double mini, int* indptr, int* indices, double* data,
            long cap, long* nnz, bool refresh, int* rebuilt) {
  if (!c || !indptr || !nnz || (cap > 0 && (!indices || !data))) return fail(-1, "null argument");
  if (int rc_gh = need_gh3(c)) return rc_gh;
  const Mesh& m = c->m;
  const int step = 2 * m.gh + 1;
  Zones z;
  if (nzones == 1) {
    z = Zones{1, {1, 0}, {m.im, 0}, zparam != 0};
  } else if (nzones == 2) {
    if (zparam < step || m.im - zparam < step) return fail(-1, "each zone must be at least 2gh+1 cells wide");
    z = Zones{2, {1, zparam + 1}, {zparam, m.im}, 1};
  } else {
    return fail(-1, "nzones must be 1 or 2");
  }
  const long rows = 5L * m.im * m.jm;
  if (rows * 245L * nzones > 2147483647L) return fail(-1, "matrix too large for 32-bit CSR indices");
  // 32-bit face-block offsets in the row kernels
  if ((long)(NOFF + 1) * 25 * std::max((long)(m.im + 1) * m.jm, (long)m.im * (m.jm + 1)) > 2147483647L)
    return fail(-1, "mesh too large for the fused CSR path");
  int rc = jacobian_phase1(c, w, g, p, ops, nops, z);
  if (rc) return rc;
  hipStream_t s = c->stream;
  const int nslot = z.n * step * step;
  const long gsize = near_count(m) * nslot * 25;
  if (c->ghostG_cap < gsize) {
    if ((rc = realloc_planes(c, &c->ghostG, gsize))) return rc;
    c->ghostG_cap = gsize;
  }
  {
    Timed t(c, "k_ghost_part");
    k_ghost_part<<<dim3(grid1(near_count(m), 256), nslot), 256, 0, s>>>(m, z, c->blki, c->blkj, c->ghostT,
                                                                      c->ghostnz, c->ghostG);
  }
  BCA_LAUNCHED();
  const int nb = (int)((rows + SCAN_B - 1) / SCAN_B);
  if (nb > SCAN_B * SCAN_B) return fail(-1, "too many rows for the scan");
  // kept masks and row lengths of this path in buffers of its own: the
  // refresh compares against them, so no other entry point may write them
  if ((rc = alloc_planes(c, &c->jr_cnt, rows))) return rc;
  if ((rc = alloc_planes(c, &c->csr_aux, CSR_AUX_INTS))) return rc;
  if ((rc = alloc_planes(c, &c->jr_mask, (long)CSR_MASK_WORDS * rows))) return rc;
  int* sums = c->csr_aux;
  int* total = c->csr_aux + (long)SCAN_B * SCAN_B;  // total, bad, changed
  int* bad = total + 1;
  int* changed = total + 2;
  k_fill_int<<<1, 64, 0, s>>>(bad, 2, 0);
  BCA_LAUNCHED();
  const dim3 gr((m.im + JC_TI - 1) / JC_TI, m.jm);
  auto rows_pass = [&](int mode, const char* name) -> int {
    Timed t(c, name);
#define BCA_ROWS(NZ, MODE)                                                                                     \
  do {                                                                                                         \
    k_jac_csr_rows<NZ, MODE, false><<<gr, JR_THREADS, 0, s>>>(m, z, c->blki, c->blkj, coefdiag, c->ghostG,    \
                                                              mini, c->jr_mask, c->jr_cnt, indptr, indices,   \
                                                              data, changed, bad);                            \
    k_jac_csr_rows<NZ, MODE, true><<<(unsigned)jr_edge_tiles(gr.x, m.jm), JR_EDGE_THREADS, 0, s>>>(            \
        m, z, c->blki, c->blkj, coefdiag, c->ghostG, mini, c->jr_mask, c->jr_cnt, indptr, indices, data, changed, \
        bad);                                                                                                  \
  } while (0)
    if (z.n == 1) {
      if (mode == JR_COUNT) BCA_ROWS(1, JR_COUNT);
      else if (mode == JR_FILL) BCA_ROWS(1, JR_FILL);
      else BCA_ROWS(1, JR_REFRESH);
    } else {
      if (mode == JR_COUNT) BCA_ROWS(2, JR_COUNT);
      else if (mode == JR_FILL) BCA_ROWS(2, JR_FILL);
      else BCA_ROWS(2, JR_REFRESH);
    }
#undef BCA_ROWS
    BCA_LAUNCHED();
    return 0;
  };
  CsrRecord& rec = c->csr_rec;
  const bool reuse = refresh && rec.valid && rec.nz == z.n && rec.zparam == zparam && rec.mini == mini &&
                     rec.indptr == indptr && rec.indices == indices && rec.data == data && rec.nnz <= cap;
  rec.valid = false;
  if (reuse) {
    if ((rc = rows_pass(JR_REFRESH, "k_jac_csr_refresh"))) return rc;
    int host[2] = {0, 0};
    BCA_HIP(hipMemcpyAsync(host, bad, 2 * sizeof(int), hipMemcpyDeviceToHost, s));
    BCA_HIP(hipStreamSynchronize(s));
    if (host[0] == 2) return fail(-4, "internal: a row outside the edge tiles is not plain");
    if (host[0]) return fail(-4, "a kept entry has a column outside the matrix");
    if (!host[1]) {
      *nnz = rec.nnz;
      *rebuilt = 0;
      rec.valid = true;
      return 0;
    }
    // some row changed: its mask and length are stored, rescan and refill
  } else {
    if ((rc = rows_pass(JR_COUNT, "k_jac_csr_count"))) return rc;
  }
  *rebuilt = 1;
  k_scan_blocks<<<nb, SCAN_B, 0, s>>>(c->jr_cnt, rows, indptr, sums);
  BCA_LAUNCHED();
  k_scan_sums<<<1, SCAN_B, 0, s>>>(sums, nb, total);
  BCA_LAUNCHED();
  k_scan_add<<<nb, SCAN_B, 0, s>>>(indptr, rows, sums, total);
  BCA_LAUNCHED();
  int host[2] = {0, 0};
  BCA_HIP(hipMemcpyAsync(host, total, 2 * sizeof(int), hipMemcpyDeviceToHost, s));
  BCA_HIP(hipStreamSynchronize(s));
  if (host[1] == 2) return fail(-4, "internal: a row outside the edge tiles is not plain");
  if (host[1]) return fail(-4, "a kept entry has a column outside the matrix");
  *nnz = host[0];
  if (host[0] > cap) return fail(-5, "capacity smaller than nnz (nnz returned; call again with larger arrays)");
  if ((rc = rows_pass(JR_FILL, "k_jac_csr_fill"))) return rc;
  rec = CsrRecord{true, z.n, zparam, mini, indptr, indices, data, host[0]};
  return 0;
}

}  // namespace

extern "C" {

int bca_jacobian_colored_csr(bca_ctx* c, const double* w, bca_metrics g, const bca_phys* p, const bca_bc_op* ops,
                             int nops, const double* coefdiag, int nzones, int zparam, double mini, int* indptr,
                             int* indices, double* data, long cap, long* nnz) {
  int rebuilt = 0;
  return jac_csr(c, w, g, p, ops, nops, coefdiag, nzones, zparam, mini, indptr, indices, data, cap, nnz, false,
                 &rebuilt);
}

int bca_jacobian_colored_csr_refresh(bca_ctx* c, const double* w, bca_metrics g, const bca_phys* p,
                                     const bca_bc_op* ops, int nops, const double* coefdiag, int nzones, int zparam,
                                     double mini, int* indptr, int* indices, double* data, long cap, long* nnz,
                                     int* rebuilt) {
  if (!rebuilt) return fail(-1, "null argument");
  return jac_csr(c, w, g, p, ops, nops, coefdiag, nzones, zparam, mini, indptr, indices, data, cap, nnz, true,
                 rebuilt);
}

}  // extern "C"

// ------------------------------------------------------------------ Newton linear solve
extern "C" {

int bca_lu_np(void* stream, double* A, int n, long ld, double tiny, int* bad, double* work) {
  if (!A || !bad || !work) return fail(-1, "null argument");
  if (n < 1 || ld < n) return fail(-1, "bca_lu_np: need n >= 1 and ld >= n");
  const hipStream_t s = (hipStream_t)stream;
  for (int k0 = 0; k0 < n; k0 += LU_NB) {
    const int nb = std::min(LU_NB, n - k0), m = n - k0 - nb;
    double* T = A + (long)k0 * ld + k0;
    const int nrb = (m + LU_THREADS - 1) / LU_THREADS;
    k_lu_panel_np<<<1 + 2 * nrb, LU_THREADS, 0, s>>>(T, ld, nb, m, nrb, tiny, bad, work);
    BCA_LAUNCHED();
    if (m == 0) break;
    const dim3 g((m + 63) / 64, (m + 63) / 64);
    k_gemm_f64_sub<<<g, 256, 0, s>>>(m, m, nb, T + (long)nb * ld, ld, T + nb, ld, T + (long)nb * ld + nb, ld, T, ld,
                                     work, nb);
    BCA_LAUNCHED();
  }
  return 0;
}

int bca_inv_np(void* stream, double* A, int n, long ld, double* work, double tiny, int* bad) {
  if (!A || !bad || !work) return fail(-1, "null argument");
  if (n < 1 || ld < n) return fail(-1, "bca_inv_np: need n >= 1 and ld >= n");
  const hipStream_t s = (hipStream_t)stream;
  const dim3 g((n + GJ_T - 1) / GJ_T, (n + GJ_T - 1) / GJ_T);
  double* buf[2] = {A, work};
  const long ldb[2] = {ld, (long)n};
  // Panels of GJ_NB columns; the steps alternate between A and work, so an
  // even number of them ends in A.  An odd count splits its last panel in
  // two (or, when that panel is one column wide, the full panel before it),
  // and n = 1 adds a step over no column (nb = 0 copies Ain to Aout).
  std::vector<int> widths;
  for (int k0 = 0; k0 < n; k0 += GJ_NB) widths.push_back(std::min(GJ_NB, n - k0));
  if (widths.size() % 2 == 1) {
    const int r = widths.back();
    if (r >= 2) {
      widths.back() = r - r / 2;
      widths.push_back(r / 2);
    } else if (widths.size() >= 2) {
      widths[widths.size() - 2] = GJ_NB / 2;
      widths.insert(widths.end() - 1, GJ_NB - GJ_NB / 2);
    } else {
      widths.push_back(0);
    }
  }
  int cur = 0, k0 = 0;
  for (const int nb : widths) {
    k_gj_step<<<g, 256, 0, s>>>(buf[cur], ldb[cur], buf[1 - cur], ldb[1 - cur], n, k0, nb, tiny, bad);
    BCA_LAUNCHED();
    cur = 1 - cur;
    k0 += nb;
  }
  if (cur != 0 || k0 != n) return fail(-3, "bca_inv_np: internal panel schedule");
  return 0;
}

int bca_gemm_f64_sub(void* stream, int M, int N, int K, const double* A, long lda, const double* B, long ldb,
                     double* C, long ldc) {
  if (!A || !B || !C) return fail(-1, "null argument");
  if (M < 0 || N < 0 || K < 0) return fail(-1, "negative size");
  const hipStream_t s = (hipStream_t)stream;
  for (int k0 = 0; k0 < K; k0 += LU_NB) {
    const int kc = std::min(LU_NB, K - k0);
    const dim3 g((N + 63) / 64, (M + 63) / 64);
    k_gemm_f64_sub<<<g, 256, 0, s>>>(M, N, kc, A + k0, lda, B + (long)k0 * ldb, ldb, C, ldc);
    BCA_LAUNCHED();
  }
  return 0;
}

int bca_csr_spmv(void* stream, long n, const int* indptr, const int* indices, const double* data, const double* x,
                 double* y) {
  if (!indptr || !indices || !data || !x || !y) return fail(-1, "null argument");
  k_csr_spmv<<<grid1(n, 256), 256, 0, (hipStream_t)stream>>>(n, indptr, indices, data, x, y);
  BCA_LAUNCHED();
  return 0;
}

}  // extern "C"

// ------------------------------------------------------------------ BF16 GEMM (learned guess)
extern "C" {

int bca_gemm_bf16(void* stream, int M, int N, int K, const unsigned short* A, long sam, long sak,
                  const unsigned short* B, long sbk, long sbn, float* C, long scm, long scn, int accumulate,
                  int ksplit) {
  if (!A || !B || !C) return fail(-1, "null argument");
  if (M < 0 || N < 0 || K < 0 || ksplit < 1 || ksplit > 4096) return fail(-1, "bad GEMM shape");
  if (M == 0 || N == 0) return 0;
  // 32-bit element offsets inside the kernel
  const long span_a = (long)(M - 1) * std::labs(sam) + (long)(std::max(K, 1) - 1) * std::labs(sak);
  const long span_b = (long)(std::max(K, 1) - 1) * std::labs(sbk) + (long)(N - 1) * std::labs(sbn);
  const long span_c = (long)(M - 1) * std::labs(scm) + (long)(N - 1) * std::labs(scn);
  if (std::max(span_a, std::max(span_b, span_c)) > 2147483647L) return fail(-1, "GEMM operand too large");
  const long tiles = (long)((M + 15) / 16) * ((N + 15) / 16);
  const dim3 grid((unsigned)((tiles + 3) / 4), (unsigned)ksplit);
  k_gemm_bf16<<<grid, 256, 0, (hipStream_t)stream>>>(M, N, K, A, sam, sak, B, sbk, sbn, C, scm, scn, accumulate);
  BCA_LAUNCHED();
  return 0;
}

}  // extern "C"

// ------------------------------------------------------------------ autotuner
namespace {

// median of reps HIP-event timings of launch() on the context stream
template <class F>
int time_launch(bca_ctx* c, int reps, F&& launch, float& ms) {
  hipEvent_t a, b;
  BCA_HIP(hipEventCreate(&a));
  BCA_HIP(hipEventCreate(&b));
  std::vector<float> t;
  for (int r = 0; r < reps; ++r) {
    BCA_HIP(hipEventRecord(a, c->stream));
    launch();
    BCA_HIP(hipEventRecord(b, c->stream));
    BCA_HIP(hipEventSynchronize(b));
    float x = 0.f;
    BCA_HIP(hipEventElapsedTime(&x, a, b));
    t.push_back(x);
  }
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  std::sort(t.begin(), t.end());
  ms = t[t.size() / 2];
  return 0;
}

}  // namespace

extern "C" {

int bca_autotune(bca_ctx* c, const double* w, bca_metrics g, const bca_phys* p, const bca_bc_op* ops, int nops,
                 int nzones, int zparam, double* jac, long nbentry, int reps, char* report, int report_len) {
  if (!c || !w || !p || !jac) return fail(-1, "null argument");
  if (int rc_gh = need_gh3(c)) return rc_gh;
  if (reps < 1 || reps > 100) return fail(-1, "reps must be 1..100");
  const Mesh& m = c->m;
  const int step = 2 * m.gh + 1;
  Zones z;
  if (nzones == 1) {
    z = Zones{1, {1, 0}, {m.im, 0}, zparam != 0};
  } else if (nzones == 2) {
    if (zparam < step || m.im - zparam < step) return fail(-1, "each zone must be at least 2gh+1 cells wide");
    z = Zones{2, {1, zparam + 1}, {zparam, m.im}, 1};
  } else {
    return fail(-1, "nzones must be 1 or 2");
  }
  if (nbentry < (long)z.n * m.im * m.jm * step * step * 25)
    return fail(-1, "nbentry smaller than nzones*im*jm*(2gh+1)^2*25");
  if ((long)(NOFF + 1) * 25 * std::max((long)(m.im + 1) * m.jm, (long)m.im * (m.jm + 1)) > 2147483647L)
    return fail(-1, "mesh too large for 32-bit face-block offsets");
  // inputs of every tuned kernel: primitives, gradients, blocks, ghost tangents
  int rc = jacobian_phase1(c, w, g, p, ops, nops, z);
  if (rc) return rc;
  const Metric gm = to_metric(g);
  const Phys ph = to_phys(p);
  hipStream_t s = c->stream;
  const dim3 shapes[] = {dim3(64, 4), dim3(64, 2), dim3(128, 2), dim3(256, 1), dim3(128, 1), dim3(32, 8),
                         dim3(64, 1)};
  std::string rep;
  float best = 0.f, ms = 0.f;
  // residual faces
  best = 1e30f;
  for (const dim3& b : shapes) {
    if ((rc = time_launch(c, reps, [&] {
           k_face<double><<<grid2(m.im + 1, m.jm + 1, b), b, 0, s>>>(m, ph, gm, w, c->w0.prim, c->w0.grad, c->w0.hn);
         }, ms)))
      return rc;
    BCA_LAUNCHED();
    if (ms < best) { best = ms; c->tune_face = b; }
  }
  rep += "k_face " + std::to_string(c->tune_face.x) + "x" + std::to_string(c->tune_face.y) + " " +
         std::to_string(best) + " ms; ";
  // interior face blocks (both families share the shape)
  best = 1e30f;
  for (const dim3& b : shapes) {
    BlockArgs ai{m, ph, gm, w, c->w0.prim, c->w0.grad, c->blki, nullptr, 3, m.jm};
    BlockArgs aj{m, ph, gm, w, c->w0.prim, c->w0.grad, c->blkj, nullptr, 4, m.jm + 1};
    if ((rc = time_launch(c, reps, [&] {
           launch_iface_blocks_o4(grid2(m.im + 1, m.jm - 2, b), b, s, ai);
           launch_jface_blocks_c(grid2(m.im, m.jm - 2, b), b, s, aj);
         }, ms)))
      return rc;
    BCA_LAUNCHED();
    if (ms < best) { best = ms; c->tune_blocks = b; }
  }
  rep += "face blocks " + std::to_string(c->tune_blocks.x) + "x" + std::to_string(c->tune_blocks.y) + " " +
         std::to_string(best) + " ms; ";
  // colored assembly into the caller's COO (values only)
  best = 1e30f;
  for (const dim3& b : shapes) {
    dim3 gr = grid2(m.im, m.jm, b);
    gr.z = step;
    if ((rc = time_launch(c, reps, [&] {
           if (z.n == 1)
             k_assemble_colored<1><<<gr, b, 0, s>>>(m, z, c->blki, c->blkj, nullptr, jac, nullptr, nullptr, 0);
           else
             k_assemble_colored<2><<<gr, b, 0, s>>>(m, z, c->blki, c->blkj, nullptr, jac, nullptr, nullptr, 0);
         }, ms)))
      return rc;
    BCA_LAUNCHED();
    if (ms < best) { best = ms; c->tune_assemble = b; }
  }
  rep += "k_assemble_colored " + std::to_string(c->tune_assemble.x) + "x" + std::to_string(c->tune_assemble.y) +
         " " + std::to_string(best) + " ms";
  if (report && report_len > 0) {
    std::strncpy(report, rep.c_str(), (size_t)report_len - 1);
    report[report_len - 1] = 0;
  }
  return 0;
}

}  // extern "C"
