// dFS on a structured hexahedral block: numbering, element->node map (the E matrix),
// Dirichlet cache, and the gather/scatter of src/fs/interface/fs.c:481-640.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <utility>

#include "common.h"

using namespace dohp;

// Closure numbering of the (p*M+1)^3 node lattice: owned nodes first, then Dirichlet
// nodes, each in lexicographic (I, J, K) order with K fastest.  For a conforming mesh
// this is a permutation of the entity-based numbering that dJacobiAddConstraints_Tensor
// (src/jacobi/impls/tensor/tensor.c:360-468) produces.
extern "C" dErr dFSCreateBlock(int M0, int M1, int M2, int degree, int dmask, const double *vcoords, dohpFS *out)
{
  DOHP_CHECK(out, dohpERR_ARG, "null output");
  *out = nullptr;
  DOHP_CHECK(M0 > 0 && M1 > 0 && M2 > 0, dohpERR_ARG, "block needs at least one element per direction");
  DOHP_CHECK(degree >= 1 && degree < kMaxN, dohpERR_ARG, "degree must be in [1,15]");
  DOHP_CHECK(vcoords, dohpERR_ARG, "vertex coordinates required");
  const long L0 = (long)degree * M0 + 1, L1 = (long)degree * M1 + 1, L2 = (long)degree * M2 + 1;
  const long nlat = L0 * L1 * L2;
  DOHP_CHECK(nlat < (1L << 31) - 1, dohpERR_ARG, "too many nodes for 32-bit indices");
  const long nel = (long)M0 * M1 * M2;
  const int n = degree + 1, n3 = n * n * n;
  DOHP_CHECK(nel * n3 < (1L << 31) - 1, dohpERR_ARG, "too many expanded nodes for 32-bit offsets");
  _p_dohpFS *fs = new _p_dohpFS;
  fs->M[0] = M0; fs->M[1] = M1; fs->M[2] = M2;
  fs->p = degree;
  fs->n = n;
  fs->nelem = (int)nel;
  fs->dirichlet_mask = dmask;
  fs->nvert = (long)(M0 + 1) * (M1 + 1) * (M2 + 1);
  fs->vcoords.assign(vcoords, vcoords + 3 * fs->nvert);

  std::vector<int> l2c(nlat);
  long own = 0;
  auto is_dir = [&](long I, long J, long K) {
    bool d = false;
    if (dmask & 1) d |= I == 0 || I == L0 - 1 || J == 0 || J == L1 - 1;
    if (dmask & 2) d |= K == L2 - 1;
    if (dmask & 4) d |= K == 0;
    return d;
  };
  for (long I = 0; I < L0; I++)
    for (long J = 0; J < L1; J++)
      for (long K = 0; K < L2; K++)
        if (!is_dir(I, J, K)) l2c[(I * L1 + J) * L2 + K] = (int)own++;
  long dc = own;
  for (long I = 0; I < L0; I++)
    for (long J = 0; J < L1; J++)
      for (long K = 0; K < L2; K++)
        if (is_dir(I, J, K)) l2c[(I * L1 + J) * L2 + K] = (int)dc++;
  fs->nowned = own;
  fs->nclosure = nlat;

  fs->elem_nodes.resize(nel * n3);
  fs->elem_verts.resize(nel * 8);
  const long V1 = M1 + 1, V2 = M2 + 1;
  for (long e0 = 0; e0 < M0; e0++)
    for (long e1 = 0; e1 < M1; e1++)
      for (long e2 = 0; e2 < M2; e2++) {
        const long e = (e0 * M1 + e1) * M2 + e2;
        int *en = &fs->elem_nodes[e * n3];
        for (int a = 0; a < n; a++)
          for (int b = 0; b < n; b++)
            for (int c = 0; c < n; c++)
              en[(a * n + b) * n + c] = l2c[((e0 * degree + a) * L1 + e1 * degree + b) * L2 + e2 * degree + c];
        // tensor vertex order t = (a*2+b)*2+c, the reordering connidx of src/fs/interface/fsgeom.c:151
        for (int t = 0; t < 8; t++) {
          const long a = t >> 2, b = (t >> 1) & 1, c = t & 1;
          fs->elem_verts[e * 8 + t] = (int)(((e0 + a) * V1 + e1 + b) * V2 + e2 + c);
        }
      }
  // Eight colours: elements of one colour share no node, so a colour can scatter without atomics.
  fs->ncolor = 8;
  fs->color_off.assign(9, 0);
  fs->color_elems.resize(nel);
  for (int col = 0, k = 0; col < 8; col++) {
    fs->color_off[col] = k;
    for (long e0 = col >> 2; e0 < M0; e0 += 2)
      for (long e1 = (col >> 1) & 1; e1 < M1; e1 += 2)
        for (long e2 = col & 1; e2 < M2; e2 += 2) fs->color_elems[k++] = (int)((e0 * M1 + e1) * M2 + e2);
    fs->color_off[col + 1] = k;
  }
  fs->lat2clo = std::move(l2c);
  dErr err = fs->d_elem_nodes.upload(fs->elem_nodes.data(), fs->elem_nodes.size());
  if (!err) err = fs->d_lat2clo.upload(fs->lat2clo.data(), fs->lat2clo.size());
  if (!err) err = fs->d_elem_verts.upload(fs->elem_verts.data(), fs->elem_verts.size());
  if (!err) err = fs->d_color_elems.upload(fs->color_elems.data(), fs->color_elems.size());
  if (!err) err = fs->d_vcoords.upload(fs->vcoords.data(), fs->vcoords.size());
  if (!err) {
    std::vector<double> zero(fs->nclosure - fs->nowned, 0.0);
    err = fs->d_dcache.upload(zero.data(), zero.size());
    std::vector<float> zf(fs->nclosure - fs->nowned, 0.0f);
    if (!err) err = fs->d_dcache32.upload(zf.data(), zf.size());
  }
  if (err) {
    delete fs;
    return err;
  }
  *out = fs;
  return 0;
}

// Continuous space from explicit element maps (the conforming case of dFSBuildSpace_Cont,
// src/fs/impls/cont/cont.c:258-385, whose element assembly matrix E, cont.c:200-256, is then a gather by
// index): elem_nodes[e][(a n + b) n + c] is the closure index of node (a, b, c) of element e (owned nodes
// 0..nowned-1, Dirichlet nodes after them), elem_verts[e][(a 2 + b) 2 + c] its Q1 vertex in tensor order.
// Elements are coloured greedily (lowest colour unused at any of their nodes) for the deterministic scatter.
extern "C" dErr dFSCreateFromElements(int degree, int nelem, long nowned, long nclosure, const int *elem_nodes,
                                      long nvert, const double *vcoords, const int *elem_verts, dohpFS *out)
{
  DOHP_CHECK(out, dohpERR_ARG, "null output");
  *out = nullptr;
  DOHP_CHECK(degree >= 1 && degree < kMaxN, dohpERR_ARG, "degree must be in [1,15]");
  DOHP_CHECK(nelem > 0 && elem_nodes && elem_verts && vcoords && nvert > 0, dohpERR_ARG,
             "elements, their node and vertex maps and the vertex coordinates are required");
  DOHP_CHECK(nowned >= 0 && nowned <= nclosure && nclosure < (1L << 31) - 1, dohpERR_ARG,
             "need 0 <= nowned <= nclosure < 2^31");
  const int n = degree + 1, n3 = n * n * n;
  DOHP_CHECK((long)nelem * n3 < (1L << 31) - 1, dohpERR_ARG, "too many expanded nodes for 32-bit offsets");
  const long nexp = (long)nelem * n3;
  std::vector<char> seen(nclosure, 0);
  for (long i = 0; i < nexp; i++) {
    DOHP_CHECK(elem_nodes[i] >= 0 && elem_nodes[i] < nclosure, dohpERR_ARG,
               "element node index out of range at entry " + std::to_string(i));
    seen[elem_nodes[i]] = 1;
  }
  for (long c = 0; c < nclosure; c++)
    DOHP_CHECK(seen[c], dohpERR_ARG, "closure node " + std::to_string(c) + " belongs to no element");
  for (long i = 0; i < (long)nelem * 8; i++)
    DOHP_CHECK(elem_verts[i] >= 0 && elem_verts[i] < nvert, dohpERR_ARG,
               "element vertex index out of range at entry " + std::to_string(i));
  _p_dohpFS *fs = new _p_dohpFS;
  fs->p = degree;
  fs->n = n;
  fs->nelem = nelem;
  fs->nowned = nowned;
  fs->nclosure = nclosure;
  fs->dirichlet_mask = 0;
  fs->nvert = nvert;
  fs->vcoords.assign(vcoords, vcoords + 3 * nvert);
  fs->elem_nodes.assign(elem_nodes, elem_nodes + nexp);
  fs->elem_verts.assign(elem_verts, elem_verts + (long)nelem * 8);
  {  // greedy colouring, at most 64 colours
    std::vector<uint64_t> used(nclosure, 0);
    std::vector<int> col(nelem);
    int ncol = 0;
    for (int e = 0; e < nelem; e++) {
      uint64_t m = 0;
      for (int k = 0; k < n3; k++) m |= used[elem_nodes[(long)e * n3 + k]];
      if (!~m) {
        delete fs;
        DOHP_CHECK(false, dohpERR_ARG, "element colouring needs more than 64 colours");
      }
      const int c = __builtin_ctzll(~m);
      col[e] = c;
      ncol = std::max(ncol, c + 1);
      for (int k = 0; k < n3; k++) used[elem_nodes[(long)e * n3 + k]] |= 1ULL << c;
    }
    fs->ncolor = ncol;
    fs->color_off.assign(ncol + 1, 0);
    for (int e = 0; e < nelem; e++) fs->color_off[col[e] + 1]++;
    for (int c = 0; c < ncol; c++) fs->color_off[c + 1] += fs->color_off[c];
    std::vector<int> pos(fs->color_off.begin(), fs->color_off.end() - 1);
    fs->color_elems.resize(nelem);
    for (int e = 0; e < nelem; e++) fs->color_elems[pos[col[e]]++] = e;
  }
  dErr err = fs->d_elem_nodes.upload(fs->elem_nodes.data(), fs->elem_nodes.size());
  if (!err) err = fs->d_elem_verts.upload(fs->elem_verts.data(), fs->elem_verts.size());
  if (!err) err = fs->d_color_elems.upload(fs->color_elems.data(), fs->color_elems.size());
  if (!err) err = fs->d_vcoords.upload(fs->vcoords.data(), fs->vcoords.size());
  if (!err && nclosure > nowned) {
    std::vector<double> zero(nclosure - nowned, 0.0);
    err = fs->d_dcache.upload(zero.data(), zero.size());
    std::vector<float> zf(nclosure - nowned, 0.0f);
    if (!err) err = fs->d_dcache32.upload(zf.data(), zf.size());
  }
  if (err) {
    delete fs;
    return err;
  }
  *out = fs;
  return 0;
}

extern "C" dErr dFSDestroy(dohpFS *fs)
{
  if (fs && *fs) {
    delete *fs;
    *fs = nullptr;
  }
  return 0;
}

extern "C" dErr dFSGetSizes(dohpFS fs, long *nowned, long *nclosure, long *nexpanded, int *nelem)
{
  DOHP_CHECK(fs, dohpERR_ARG, "null dFS");
  if (nowned) *nowned = fs->nowned;
  if (nclosure) *nclosure = fs->nclosure;
  if (nexpanded) *nexpanded = (long)fs->nelem * fs->n * fs->n * fs->n;
  if (nelem) *nelem = fs->nelem;
  return 0;
}

extern "C" dErr dFSGetElementNodes(dohpFS fs, int *elem_nodes)
{
  DOHP_CHECK(fs && elem_nodes, dohpERR_ARG, "null argument");
  std::memcpy(elem_nodes, fs->elem_nodes.data(), fs->elem_nodes.size() * sizeof(int));
  return 0;
}

extern "C" dErr dFSGetNodalCoordinatesClosure(dohpFS fs, double *coords)
{
  DOHP_CHECK(fs && coords, dohpERR_ARG, "null argument");
  const int n = fs->n, n3 = n * n * n;
  double z[kMaxN], w[kMaxN];
  gauss_lobatto(n, z, w);
  std::vector<double> cnt(fs->nclosure, 0.0);
  std::fill(coords, coords + 3 * fs->nclosure, 0.0);
  for (int e = 0; e < fs->nelem; e++) {
    const int *vt = &fs->elem_verts[(long)e * 8];
    for (int a = 0; a < n; a++)
      for (int b = 0; b < n; b++)
        for (int c = 0; c < n; c++) {
          const double f[3][2] = {{(1 - z[a]) / 2, (1 + z[a]) / 2}, {(1 - z[b]) / 2, (1 + z[b]) / 2},
                                  {(1 - z[c]) / 2, (1 + z[c]) / 2}};
          const int node = fs->elem_nodes[(long)e * n3 + (a * n + b) * n + c];
          for (int t = 0; t < 8; t++) {
            const double s = f[0][t >> 2] * f[1][(t >> 1) & 1] * f[2][t & 1];
            for (int d = 0; d < 3; d++) coords[3L * node + d] += s * fs->vcoords[3L * vt[t] + d];
          }
          cnt[node] += 1.0;
        }
  }
  for (long i = 0; i < fs->nclosure; i++)
    for (int d = 0; d < 3; d++) coords[3 * i + d] /= cnt[i];
  return 0;
}

extern "C" dErr dFSInhomogeneousDirichletCommit(dohpFS fs, const double *dvals)
{
  DOHP_CHECK(fs && dvals, dohpERR_ARG, "null argument");
  const long nd = fs->nclosure - fs->nowned;
  DOHP_CALL(fs->d_dcache.upload(dvals, nd));
  std::vector<float> f(dvals, dvals + nd);
  DOHP_CALL(fs->d_dcache32.upload(f.data(), nd));
  return 0;
}

// x_exp = E [g; dirichlet]: one thread per expanded node, coalesced over the expanded vector.
__global__ void fs_gather_kernel(const int *__restrict__ enodes, long nexp, const double *__restrict__ g,
                                 const double *__restrict__ dcache, long nowned, int inhom, int add,
                                 double *__restrict__ x)
{
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nexp) return;
  const int c = enodes[i];
  const double v = c < nowned ? g[c] : (inhom ? dcache[c - nowned] : 0.0);
  x[i] = add ? x[i] + v : v;
}

// g += E^T x_exp restricted to owned nodes, one colour of elements at a time so that
// no two threads update the same node: deterministic, no atomics.
__global__ void fs_scatter_color_kernel(const int *__restrict__ enodes, const int *__restrict__ elist, int ne,
                                        int n3, const double *__restrict__ x, long nowned, double *__restrict__ g)
{
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)ne * n3) return;
  const long e = elist[t / n3], k = t % n3;
  const int c = enodes[e * n3 + k];
  if (c < nowned) g[c] += x[e * n3 + k];
}

__global__ void fs_zero_kernel(double *__restrict__ g, long n)
{
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) g[i] = 0.0;
}

extern "C" dErr dFSGlobalToExpanded(dohpFS fs, const double *g, double *x, int hmode, int imode, void *stream)
{
  DOHP_CHECK(fs && g && x, dohpERR_ARG, "null argument");
  DOHP_CHECK(imode == dINSERT_VALUES || imode == dADD_VALUES, dohpERR_ARG, "InsertMode not supported");
  DOHP_CHECK(hmode == dFS_HOMOGENEOUS || hmode == dFS_INHOMOGENEOUS, dohpERR_ARG, "invalid homogeneous mode");
  const long nexp = (long)fs->nelem * fs->n * fs->n * fs->n;
  const int bs = 256;
  hipLaunchKernelGGL(fs_gather_kernel, dim3((nexp + bs - 1) / bs), dim3(bs), 0, (hipStream_t)stream,
                     fs->d_elem_nodes.ptr, nexp, g, fs->d_dcache.ptr, fs->nowned, hmode == dFS_INHOMOGENEOUS,
                     imode == dADD_VALUES, x);
  DOHP_HIP(hipGetLastError());
  return 0;
}

extern "C" dErr dFSExpandedToGlobal(dohpFS fs, const double *x, double *g, int hmode, int imode, void *stream)
{
  DOHP_CHECK(fs && g && x, dohpERR_ARG, "null argument");
  DOHP_CHECK(imode == dINSERT_VALUES || imode == dADD_VALUES, dohpERR_ARG, "InsertMode not supported");
  (void)hmode;  // Dirichlet entries of the closure are discarded either way (fs.c:592-640)
  const int bs = 256, n3 = fs->n * fs->n * fs->n;
  hipStream_t st = (hipStream_t)stream;
  if (imode == dINSERT_VALUES)
    hipLaunchKernelGGL(fs_zero_kernel, dim3((fs->nowned + bs - 1) / bs), dim3(bs), 0, st, g, fs->nowned);
  for (int col = 0; col < fs->ncolor; col++) {
    const int ne = fs->color_off[col + 1] - fs->color_off[col];
    if (!ne) continue;
    const long nt = (long)ne * n3;
    hipLaunchKernelGGL(fs_scatter_color_kernel, dim3((nt + bs - 1) / bs), dim3(bs), 0, st, fs->d_elem_nodes.ptr,
                       fs->d_color_elems.ptr + fs->color_off[col], ne, n3, x, fs->nowned, g);
  }
  DOHP_HIP(hipGetLastError());
  return 0;
}
