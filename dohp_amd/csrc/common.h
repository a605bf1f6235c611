// Internal declarations shared by the dohp_amd HIP sources.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdio>
#include <map>
#include <string>
#include <vector>

#include "../../include/dohp_amd.h"

namespace dohp {

void set_error(const std::string &msg);

#define DOHP_HIP(call)                                                                              \
  do {                                                                                              \
    hipError_t e_ = (call);                                                                         \
    if (e_ != hipSuccess) {                                                                         \
      ::dohp::set_error(std::string("HIP error ") + hipGetErrorString(e_) + " at " + __FILE__ + ":" + \
                        std::to_string(__LINE__));                                                  \
      return dohpERR_HIP;                                                                           \
    }                                                                                               \
  } while (0)

#define DOHP_CHECK(cond, code, msg)     \
  do {                                  \
    if (!(cond)) {                      \
      ::dohp::set_error(msg);           \
      return (code);                    \
    }                                   \
  } while (0)

#define DOHP_CALL(call)       \
  do {                        \
    dErr r_ = (call);         \
    if (r_) return r_;        \
  } while (0)

constexpr int kMaxN = 16;
// Inputs per GLL sub-cell of the learned embedded-Q1 scales (features kernel in q1.hip, network in learn.hip).
constexpr int kQ1CellInputs = 11;

// 1-D rules and bases (host).
int gauss_legendre(int n, double *z, double *w);
int gauss_lobatto(int n, double *z, double *w);
void barycentric_weights(int n, const double *x, double *bw);
// interp[q][j] = l_j(xq[q]) for the Lagrange basis through x[0..n)
void lagrange_matrix(int n, const double *x, int m, const double *xq, double *interp);
// D[i][j] = l_j'(x[i]) for the Lagrange basis through x[0..n)
void collocation_derivative(int n, const double *x, double *D);

// Device buffer helper.
template <typename T>
struct DevBuf {
  T *ptr = nullptr;
  size_t n = 0;
  DevBuf() = default;
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  DevBuf(DevBuf &&o) noexcept : ptr(o.ptr), n(o.n)
  {
    o.ptr = nullptr;
    o.n = 0;
  }
  DevBuf &operator=(DevBuf &&o) noexcept
  {
    if (this != &o) {
      release();
      ptr = o.ptr;
      n = o.n;
      o.ptr = nullptr;
      o.n = 0;
    }
    return *this;
  }
  // On failure the buffer is left empty (ptr = nullptr, n = 0), so a later size check reallocates.
  dErr alloc(size_t count) {
    release();
    if (count == 0) return 0;
    if (hipMalloc(&ptr, count * sizeof(T)) != hipSuccess) {
      ptr = nullptr;
      (void)hipGetLastError();
      set_error("hipMalloc failed for " + std::to_string(count * sizeof(T)) + " bytes");
      return dohpERR_OOM;
    }
    n = count;
    return 0;
  }
  // Holds exactly `count` elements (reallocated otherwise).
  bool holds(size_t count) const { return ptr != nullptr && n == count; }
  // Reuses the buffer when it holds exactly count elements, after the device is idle (as freeing it would wait):
  // work still reading the old contents finishes first.
  dErr upload(const T *h, size_t count) {
    if (holds(count))
      DOHP_HIP(hipDeviceSynchronize());
    else
      DOHP_CALL(alloc(count));
    if (count) DOHP_HIP(hipMemcpy(ptr, h, count * sizeof(T), hipMemcpyHostToDevice));
    return 0;
  }
  void release() {
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    n = 0;
  }
  ~DevBuf() { release(); }
};

// Device buffers kept between calls of one object: hipFree waits for the device (about 0.15 ms a call), so the work
// vectors of a solve are taken from the operator's pool and given back to it, not freed.  A buffer is reused only at
// the same size; when an allocation fails, the idle buffers are freed and it is tried once more.  Not thread-safe
// (one caller per object, like the reference's objects).
struct Pool {
  std::multimap<size_t, void *> idle;
  Pool() = default;
  Pool(const Pool &) = delete;
  Pool &operator=(const Pool &) = delete;
  void trim()
  {
    for (auto &kv : idle) (void)hipFree(kv.second);
    idle.clear();
  }
  ~Pool() { trim(); }
  void *take(size_t bytes)
  {
    auto it = idle.find(bytes);
    if (it != idle.end()) {
      void *p = it->second;
      idle.erase(it);
      return p;
    }
    void *p = nullptr;
    if (hipMalloc(&p, bytes) == hipSuccess) return p;
    (void)hipGetLastError();
    trim();
    if (hipMalloc(&p, bytes) == hipSuccess) return p;
    (void)hipGetLastError();
    return nullptr;
  }
  void give(size_t bytes, void *p) { idle.emplace(bytes, p); }
};

// A work vector from a Pool, given back when it goes out of scope (after the work queued on it in stream order).
template <typename T>
struct PoolBuf {
  Pool *pool = nullptr;
  T *ptr = nullptr;
  size_t n = 0;
  PoolBuf() = default;
  PoolBuf(const PoolBuf &) = delete;
  PoolBuf &operator=(const PoolBuf &) = delete;
  dErr alloc(Pool &p, size_t count)
  {
    release();
    if (count == 0) return 0;
    void *q = p.take(count * sizeof(T));
    if (!q) {
      set_error("hipMalloc failed for " + std::to_string(count * sizeof(T)) + " bytes");
      return dohpERR_OOM;
    }
    pool = &p;
    ptr = (T *)q;
    n = count;
    return 0;
  }
  void release()
  {
    if (ptr) pool->give(n * sizeof(T), ptr);
    ptr = nullptr;
    n = 0;
  }
  ~PoolBuf() { release(); }
};

// Current device and its compute-unit count (the properties are read once per device).
inline dErr current_device_cus(int &dev, int &cus)
{
  static int cached[64] = {};
  DOHP_HIP(hipGetDevice(&dev));
  DOHP_CHECK(dev >= 0 && dev < 64, dohpERR_ARG, "device index out of range");
  if (!cached[dev]) {
    hipDeviceProp_t prop;
    DOHP_HIP(hipGetDeviceProperties(&prop, dev));
    cached[dev] = prop.multiProcessorCount;
  }
  cus = cached[dev];
  return 0;
}

// Logical block index: consecutive workgroups are dealt round-robin to the eight XCDs, so
// give each XCD a contiguous range of work (neighbours share data through its L2).
__device__ inline long xcd_block(long b, long nb)
{
  const long q = nb / 8, r = nb % 8, x = b % 8, pos = b / 8;
  return x < r ? x * (q + 1) + pos : r * (q + 1) + (x - r) * q + pos;
}

}  // namespace dohp

// Internal entry points shared between the operator and the Q1 assembly.
extern "C" dErr dohpEllipFunctionState_Private(dohpEllip op, const double *u, hipStream_t st);
// y = J x without the entry points' non-finite check (no wait for the stream): the library's solvers.
dErr dohpEllipMatMult_Private(dohpEllip op, const double *x, double *y, hipStream_t st);
dErr dohpEllipMatMult_Private(dohpEllip op, const float *x, float *y, hipStream_t st);
// The operator's pool of solver work vectors.
dohp::Pool *dohpEllipPool_Private(dohpEllip op);

// Sparse matrix on the owned nodes.  CSR is the assembled form; on a block space the owned nodes
// form a box of the node lattice (nx * ny * nz, lexicographic, last index fastest) and every row
// couples only its 27 lattice neighbours, so the matrix is also kept as a 27-point stencil:
// st[s][r] with s = (di+1)*9 + (dj+1)*3 + (dk+1), zero where the neighbour is not owned.
struct _p_dohpMat {
  long nrows = 0, nnz = 0;
  std::vector<int> h_rowptr;
  dohp::DevBuf<int> rowptr, col;
  dohp::DevBuf<double> val;
  dohp::DevBuf<float> val32;
  int box[3] = {0, 0, 0};
  dohp::DevBuf<double> st64;
  dohp::DevBuf<float> st32;
  // multigrid hierarchy built by the first solve with a given smoothing degree; dropped when the
  // matrix is refilled
  dohpMG mg = nullptr;
  int mg_degree = 0;
  int corrected = 0;  // assembled with the corrected sub-cell arithmetic (embedded-Q1 mode 1)
  int scaled = 0;     // assembled with learned per-sub-cell derivative scales
  bool csr_stale = false;  // block-space matrix: CSR values and FP32 stencil not yet derived from the stencil form
  bool mg_stale = false;   // refilled since the cached hierarchy was computed (refreshed in place on its next use)
  dohp::DevBuf<double> q1tab;  // the assembly kernel's 1-D tables
  ~_p_dohpMat() { dohpMGDestroy(&mg); }
};
// Sets the box of a block-space matrix and allocates its stencil form (*box = false: not a box, CSR only).
dErr dohpMatPrepareStencil_Private(_p_dohpMat *A, dohpFS fs, bool *box, hipStream_t st);
// CSR values (FP64, FP32) and the FP32 stencil from the FP64 stencil form.
dErr dohpMatStencilToCSR_Private(_p_dohpMat *A, hipStream_t st);
// Derives the CSR values and FP32 stencil of a block-space matrix on their first use after an assembly.
dErr dohpMatSyncCSR_Private(_p_dohpMat *A, hipStream_t st);
dErr dohpStencilMult_Private(const _p_dohpMat *A, const double *x, double *y, hipStream_t st);
dErr dohpStencilMultF32_Private(const _p_dohpMat *A, const float *x, float *y, hipStream_t st);

// Function space on a structured hexahedral block.
struct _p_dohpFS {
  int M[3] = {0, 0, 0};
  int p = 0, n = 0;
  int nelem = 0;
  long nowned = 0, nclosure = 0;
  int dirichlet_mask = 0;
  long nvert = 0;
  std::vector<double> vcoords;       // [nvert][3]
  std::vector<int> elem_nodes;       // [nelem][n^3]
  std::vector<int> elem_verts;       // [nelem][8], tensor order
  int ncolor = 8;                    // element colours (elements of one colour share no node)
  std::vector<int> color_off;        // ncolor + 1 offsets into color_elems
  std::vector<int> color_elems;      // elements grouped by (ex%2, ey%2, ez%2)
  std::vector<int> lat2clo;          // closure index of each lattice node (I, J, K), K fastest
  dohp::DevBuf<int> d_elem_nodes, d_elem_verts, d_color_elems, d_lat2clo;
  dohp::DevBuf<double> d_vcoords, d_dcache;
  dohp::DevBuf<float> d_dcache32;
  // block spaces (dFSCreateBlock) have M > 0 and a lattice numbering the kernels evaluate; spaces built
  // from element maps (dFSCreateFromElements) have M = 0 and are read through elem_nodes / elem_verts
  bool is_block() const { return M[0] > 0; }
};
