// Geometric multigrid on the embedded-Q1 matrix of a block space: the multilevel solve that
// ellip's driver obtains from PETSc's preconditioners applied to the Q1 matrix
// (src/fs/tests/ellip.c:594-652,701-747).
//
// The owned nodes of a block space form a box of the GLL node lattice and the Q1 matrix is a
// 27-point stencil on it.  Levels coarsen every box dimension d to floor(d/2): coarse node c
// sits on fine node 2c+1; fine node 2c interpolates its neighbours c-1 and c with weight 1/2
// (missing neighbours are Dirichlet zeros).  Coarse operators are Galerkin products P^T A P
// (again 27-point), smoothing is a Jacobi-preconditioned Chebyshev polynomial on
// [lo * lmax, hi * lmax] of D^{-1} A, the coarsest level (at most a few hundred nodes) is solved
// with its explicit inverse.  The V-cycle works on symmetric half stencils (see
// half_residual_kernel).  Pre- and post-smoothers are the same polynomial, so the V-cycle is
// a symmetric operator and serves as a CG preconditioner.
#include <cmath>
#include <vector>

#include "common.h"

using namespace dohp;

namespace {

struct Box {
  int nx, ny, nz;
  __host__ __device__ long n() const { return (long)nx * ny * nz; }
};

// Box kernels launch a 3-D grid: x covers the last (contiguous) index K in chunks of 64, y the
// middle index J in chunks of 4 (threadIdx.y), z the first index I; no index division.
constexpr int kBX = 64, kBY = 4;
inline dim3 box_grid(const Box &b)
{
  return dim3((unsigned)((b.nz + kBX - 1) / kBX), (unsigned)((b.ny + kBY - 1) / kBY), (unsigned)b.nx);
}
inline dim3 box_block() { return dim3(kBX, kBY); }
__device__ __forceinline__ bool box_index(const Box &b, int &I, int &J, int &K, long &r)
{
  K = blockIdx.x * kBX + threadIdx.x;
  J = blockIdx.y * kBY + threadIdx.y;
  I = blockIdx.z;
  if (K >= b.nz || J >= b.ny) return false;
  r = ((long)I * b.ny + J) * b.nz + K;
  return true;
}

__device__ __forceinline__ bool inside(const Box &b, int I, int J, int K)
{
  return I >= 0 && I < b.nx && J >= 0 && J < b.ny && K >= 0 && K < b.nz;
}

__device__ __forceinline__ void unrank(const Box &b, long r, int &I, int &J, int &K)
{
  const long yz = (long)b.ny * b.nz;
  I = (int)(r / yz);
  const long rem = r - (long)I * yz;
  J = (int)(rem / b.nz);
  K = (int)(rem - (long)J * b.nz);
}

// stencil -> CSR values (and the FP32 copies of both).  A workgroup takes 256 consecutive rows: it
// converts their stencil planes to FP32 (one row per thread) and walks their CSR entries, which are
// contiguous, one entry per thread (coalesced), finding each entry's row by binary search in the
// rows' offsets staged in LDS.
__global__ void __launch_bounds__(256) stencil_to_csr_kernel(Box b, const int *__restrict__ rowptr,
                                                             const int *__restrict__ col,
                                                             const double *__restrict__ st, double *__restrict__ val,
                                                             float *__restrict__ val32, float *__restrict__ st32)
{
  __shared__ int rp[257];
  const long n = b.n();
  const long r0 = (long)blockIdx.x * 256;
  const int nr = (int)min(256L, n - r0), t = threadIdx.x;
  if (t < nr) {
    const long r = r0 + t;
    for (int s = 0; s < 27; s++) st32[s * n + r] = (float)st[s * n + r];
  }
  for (int j = t; j <= nr; j += 256) rp[j] = rowptr[r0 + j];
  __syncthreads();
  const int kb = rp[0], ke = rp[nr];
  for (int k = kb + t; k < ke; k += 256) {
    int lo = 0, hi = nr - 1;  // last row whose entries start at or before k
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (rp[mid] <= k)
        lo = mid;
      else
        hi = mid - 1;
    }
    const long r = r0 + lo;
    int I, J, K, I2, J2, K2;
    unrank(b, r, I, J, K);
    unrank(b, col[k], I2, J2, K2);
    const int di = I2 - I, dj = J2 - J, dk = K2 - K;
    double v = 0.0;
    if (di >= -1 && di <= 1 && dj >= -1 && dj <= 1 && dk >= -1 && dk <= 1)
      v = st[((di + 1) * 9 + (dj + 1) * 3 + (dk + 1)) * n + r];
    val[k] = v;
    val32[k] = (float)v;
  }
}

// y = A x.  Entries are summed in increasing column order, as the CSR product does.
template <typename T>
__global__ void stencil_mult_kernel(Box b, const T *__restrict__ A, const T *__restrict__ x, T *__restrict__ y)
{
  const long n = b.n();
  int I, J, K;
  long r;
  if (!box_index(b, I, J, K, r)) return;
  const long sy = b.nz, sx = (long)b.ny * b.nz;
  // every load is issued unconditionally (a neighbour outside the box reads the row's own x and is dropped by a
  // select), so the 27 loads of a row are in flight together instead of one branch and wait each
  T acc = 0;
#pragma unroll
  for (int s = 0; s < 27; s++) {
    const int di = s / 9 - 1, dj = (s / 3) % 3 - 1, dk = s % 3 - 1;
    const bool in = inside(b, I + di, J + dj, K + dk);
    const T v = A[s * n + r] * x[in ? r + di * sx + dj * sy + dk : r];
    acc += in ? v : T(0);
  }
  y[r] = acc;
}

// The Q1 matrix and its Galerkin coarse operators are symmetric, so the levels keep only the
// diagonal and the 13 upper neighbours, H[s - 13][r] for s = 13..26; the entry of a lower
// neighbour is the upper entry stored at that neighbour's row.  The product reads 14 coefficient
// streams instead of 27 (the shifted reads of the lower half hit the caches).
template <typename T>
__global__ void __launch_bounds__(256) half_residual_kernel(Box bx, const T *__restrict__ H, const T *__restrict__ b,
                                                            const T *__restrict__ x, T *__restrict__ r)
{
  const long n = bx.n();
  int I, J, K;
  long i;
  if (!box_index(bx, I, J, K, i)) return;
  const long sy = bx.nz, sx = (long)bx.ny * bx.nz;
  // unconditional loads at clamped indices, dropped by selects (see stencil_mult_kernel)
  T acc = H[i] * x[i];
#pragma unroll
  for (int s = 14; s < 27; s++) {
    const int di = s / 9 - 1, dj = (s / 3) % 3 - 1, dk = s % 3 - 1;
    const long off = di * sx + dj * sy + dk;
    const bool up = inside(bx, I + di, J + dj, K + dk), lo = inside(bx, I - di, J - dj, K - dk);
    const long iu = up ? i + off : i, il = lo ? i - off : i;
    const T vu = H[(s - 13) * n + i] * x[iu], vl = H[(s - 13) * n + il] * x[il];
    acc += up ? vu : T(0);
    acc += lo ? vl : T(0);
  }
  r[i] = (b ? b[i] : T(0)) - acc;
}

// r = b - A x (b may be null) on a level.
template <typename T>
void residual(const Box &bx, const T *H, const T *b, const T *x, T *r, hipStream_t st)
{
  hipLaunchKernelGGL(half_residual_kernel<T>, box_grid(bx), box_block(), 0, st, bx, H, b, x, r);
}

// FP16 coefficients of the FP32 V-cycle.  With the row scales s_i = sqrt(a_ii) (FP32), a level keeps the 13 upper
// off-diagonals as h_ij = fp16(a_ij / (s_i s_j)) (|h_ij| <= 1 for a symmetric positive definite matrix) and, in FP32,
// the diagonal d_i = a_ii + sum_j (a_ij - s_i s_j h_ij) over all 26 neighbours: the rounding of the off-diagonals
// moves into the diagonal, so every row sum (and with it the near-null space of constants, on which the
// preconditioner's quality rests) is that of the FP64 matrix.  A row then reads 28 + 8 bytes of coefficients
// (the 13 off-diagonals padded to 7 pairs) instead of 56.
__device__ __forceinline__ _Float16 coeff16(double a, float si, float sj)
{
  return (_Float16)(float)(a / ((double)si * (double)sj));
}

__global__ void __launch_bounds__(256) half16_setup_kernel(Box bx, const double *__restrict__ A,
                                                           _Float16 *__restrict__ H, float *__restrict__ dg,
                                                           float *__restrict__ sc)
{
  const long n = bx.n();
  int I, J, K;
  long i;
  if (!box_index(bx, I, J, K, i)) return;
  const long sy = bx.nz, sx = (long)bx.ny * bx.nz;
  const float si = (float)sqrt(A[13 * n + i]);
  double comp = 0;
#pragma unroll 1
  for (int s = 14; s < 27; s++) {
    const int di = s / 9 - 1, dj = (s / 3) % 3 - 1, dk = s % 3 - 1;
    const long off = di * sx + dj * sy + dk;
    _Float16 h = 0;
    if (I + di >= 0 && I + di < bx.nx && J + dj >= 0 && J + dj < bx.ny && K + dk >= 0 && K + dk < bx.nz) {
      const float sj = (float)sqrt(A[13 * n + i + off]);
      const double a = A[s * n + i];
      h = coeff16(a, si, sj);
      comp += a - (double)si * (double)sj * (double)h;
    }
    H[((s - 14) >> 1) * 2 * n + 2 * i + ((s - 14) & 1)] = h;
    if (I - di >= 0 && I - di < bx.nx && J - dj >= 0 && J - dj < bx.ny && K - dk >= 0 && K - dk < bx.nz) {
      // the lower neighbour's entry is its upper one, rounded as that row rounds it
      const float sj = (float)sqrt(A[13 * n + i - off]);
      const double a = A[s * n + i - off];
      const _Float16 hl = coeff16(a, sj, si);
      comp += a - (double)sj * (double)si * (double)hl;
    }
  }
  H[6 * 2 * n + 2 * i + 1] = 0;  // (the pad of the last pair)
  dg[i] = (float)(A[13 * n + i] + comp);
  sc[i] = si;
}

// The FP16 products stage the scaled iterate s x of their block's neighbourhood in LDS: planes I-1..I+1, rows
// J0-1..J0+kBY, columns K0-1..K0+kBX (zero outside the box), loaded once by the block (every load unconditional, at a
// clamped index) instead of 26 loads of x and 26 of s per row; each row then reads its neighbours from the tile.
constexpr int kTX = kBX + 2, kTY = kBY + 2, kTile = 3 * kTY * kTX, kTileIt = (kTile + kBX * kBY - 1) / (kBX * kBY);
__device__ __forceinline__ void stage_scaled(const Box &bx, const float *__restrict__ sc, const float *__restrict__ x,
                                             float *tile)
{
  const int I = blockIdx.z, J0 = blockIdx.y * kBY, K0 = blockIdx.x * kBX;
  const long sy = bx.nz, sx = (long)bx.ny * bx.nz;
  const int tid = threadIdx.y * kBX + threadIdx.x;
  float v[kTileIt];
#pragma unroll
  for (int u = 0; u < kTileIt; u++) {
    const int t = min(tid + u * kBX * kBY, kTile - 1);
    const int p = t / (kTY * kTX), rem = t - p * (kTY * kTX), jj = rem / kTX, kk = rem - jj * kTX;
    const int gi = I + p - 1, gj = J0 + jj - 1, gk = K0 + kk - 1;
    const bool in = inside(bx, gi, gj, gk);
    const long g = in ? gi * sx + gj * sy + gk : 0;
    const float y = sc[g] * x[g];
    v[u] = in ? y : 0.f;
  }
#pragma unroll
  for (int u = 0; u < kTileIt; u++)
    if (tid + u * kBX * kBY < kTile) tile[tid + u * kBX * kBY] = v[u];
  __syncthreads();
}

// sum_j h_ij s_j x_j over the 26 neighbours of row i (the thread's row of the block), from the staged tile.  The
// coefficients are stored as pairs of directions per row ([7][n] dwords, directions 14 + 2j and 15 + 2j, the last
// pair padded): a row reads its 13 upper ones with 7 dword loads, and each lower one (the upper one of the lower
// neighbour's row) with one, instead of 26 two-byte loads.  A neighbour outside the box has a zero tile entry; its
// coefficient is read at a clamped index.
__device__ __forceinline__ float half16_value(unsigned pair, int hi)
{
  return (float)__builtin_bit_cast(_Float16, (unsigned short)(hi ? pair >> 16 : pair & 0xffffu));
}
__device__ __forceinline__ float half16_tile_row(const Box &bx, const _Float16 *__restrict__ H, const float *tile,
                                                 long i)
{
  const long n = bx.n(), sy = bx.nz, sx = (long)bx.ny * bx.nz;
  const unsigned *__restrict__ H2 = (const unsigned *)H;
  const int tj = threadIdx.y + 1, tk = threadIdx.x + 1;
  unsigned up[7];
#pragma unroll
  for (int j = 0; j < 7; j++) up[j] = H2[j * n + i];
  float acc = 0;
#pragma unroll
  for (int s = 14; s < 27; s++) {
    const int di = s / 9 - 1, dj = (s / 3) % 3 - 1, dk = s % 3 - 1;
    const long off = di * sx + dj * sy + dk;
    const long il = i - off >= 0 && i - off < n ? i - off : i;
    const float hu = half16_value(up[(s - 14) >> 1], (s - 14) & 1);
    const float hl = half16_value(H2[((s - 14) >> 1) * n + il], (s - 14) & 1);
    acc += hu * tile[((1 + di) * kTY + tj + dj) * kTX + tk + dk];
    acc += hl * tile[((1 - di) * kTY + tj - dj) * kTX + tk - dk];
  }
  return acc;
}

__global__ void __launch_bounds__(256) half16_residual_kernel(Box bx, const _Float16 *__restrict__ H,
                                                              const float *__restrict__ dg, const float *__restrict__ sc,
                                                              const float *__restrict__ b, const float *__restrict__ x,
                                                              float *__restrict__ r)
{
  __shared__ float tile[kTile];
  stage_scaled(bx, sc, x, tile);
  int I, J, K;
  long i;
  if (!box_index(bx, I, J, K, i)) return;
  const float ax = dg[i] * x[i] + sc[i] * half16_tile_row(bx, H, tile, i);
  r[i] = (b ? b[i] : 0.f) - ax;
}

// One Chebyshev step of the smoother on the FP16 coefficients, out of place (Jacobi needs the old iterate at the
// neighbours): dn = a d + c (b - A xin) / d_ii, xout = xin + dn; d is read when a != 0 and written when keep_d.
__global__ void __launch_bounds__(256) half16_cheb_kernel(Box bx, const _Float16 *__restrict__ H,
                                                          const float *__restrict__ dg, const float *__restrict__ sc,
                                                          const float *__restrict__ b, const float *__restrict__ xin,
                                                          float *__restrict__ xout, float *__restrict__ d, float a,
                                                          float c, int keep_d)
{
  __shared__ float tile[kTile];
  stage_scaled(bx, sc, xin, tile);
  int I, J, K;
  long i;
  if (!box_index(bx, I, J, K, i)) return;
  const float xi = xin[i];
  const float r = b[i] - (dg[i] * xi + sc[i] * half16_tile_row(bx, H, tile, i));
  const float dn = (a != 0.f ? a * d[i] : 0.f) + c * r / dg[i];
  if (keep_d) d[i] = dn;
  xout[i] = xi + dn;
}

// First step from a zero iterate: x = d = c b / d_ii.
__global__ void half16_cheb_first_kernel(long n, float c, const float *__restrict__ dg, const float *__restrict__ b,
                                         float *__restrict__ x, float *__restrict__ d, int keep_d)
{
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float dn = c * b[i] / dg[i];
  if (keep_d) d[i] = dn;
  x[i] = dn;
}

__global__ void half_from_full_kernel(long n, const double *__restrict__ A, double *__restrict__ H,
                                      float *__restrict__ H32)
{
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  for (int s = 13; s < 27; s++) {
    H[(s - 13) * n + i] = A[s * n + i];
    H32[(s - 13) * n + i] = (float)A[s * n + i];
  }
}

template <typename T>
__global__ void diag_inv_kernel(long n, const T *__restrict__ A, T *__restrict__ dinv)
{
  const long r = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (r < n) dinv[r] = T(1) / A[13 * n + r];
}

// Chebyshev step: d = a d + c dinv r;  x += d   (first step: a = 0)
template <typename T>
__global__ void cheb_step_kernel(long n, T a, T c, const T *__restrict__ dinv, const T *__restrict__ r,
                                 T *__restrict__ d, T *__restrict__ x, int first_sets_x)
{
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const T dn = (a == T(0) ? T(0) : a * d[i]) + c * dinv[i] * r[i];
  d[i] = dn;
  x[i] = first_sets_x ? dn : x[i] + dn;
}

// 1-D prolongation weights of fine index f onto coarse indices (at most two).
__device__ __forceinline__ int parents(int f, int nc, int (&c)[2], double (&w)[2])
{
  if (f & 1) {
    c[0] = (f - 1) >> 1;
    w[0] = 1.0;
    return c[0] < nc ? 1 : 0;
  }
  int m = 0;
  const int lo = (f >> 1) - 1, hi = f >> 1;
  if (lo >= 0 && lo < nc) { c[m] = lo; w[m] = 0.5; m++; }
  if (hi >= 0 && hi < nc) { c[m] = hi; w[m] = 0.5; m++; }
  return m;
}

// bc = P^T rf
template <typename T>
__global__ void restrict_kernel(Box f, Box c, const T *__restrict__ rf, T *__restrict__ bc)
{
  int I, J, K;
  long C;
  if (!box_index(c, I, J, K, C)) return;
  const long sy = f.nz, sx = (long)f.ny * f.nz;
  const int F0 = 2 * I + 1, F1 = 2 * J + 1, F2 = 2 * K + 1;
  const long base = F0 * sx + F1 * sy + F2;  // (F0, F1, F2) is inside the fine box; its upper neighbours may not be
  T acc = 0;
#pragma unroll
  for (int a = -1; a <= 1; a++)
#pragma unroll
    for (int b = -1; b <= 1; b++)
#pragma unroll
      for (int d = -1; d <= 1; d++) {
        const bool in = F0 + a < f.nx && F1 + b < f.ny && F2 + d < f.nz;
        const T w = T((a ? 0.5 : 1.0) * (b ? 0.5 : 1.0) * (d ? 0.5 : 1.0));
        const T v = w * rf[in ? base + a * sx + b * sy + d : base];  // unconditional load, dropped by a select
        acc += in ? v : T(0);
      }
  bc[C] = acc;
}

// The two coarse parents of fine index f in one direction with their weights; a missing parent has weight 0 and
// index 0, so that prolongations load unconditionally (nc >= 1).
__device__ __forceinline__ void parents2(int f, int nc, int (&c)[2], float (&w)[2])
{
  int cc[2];
  double ww[2];
  const int m = parents(f, nc, cc, ww);
  c[0] = m > 0 ? cc[0] : 0;
  w[0] = m > 0 ? (float)ww[0] : 0.f;
  c[1] = m > 1 ? cc[1] : 0;
  w[1] = m > 1 ? (float)ww[1] : 0.f;
}

// xout = xin + P xc (out of place: the FP16 V-cycle's iterate alternates between two buffers)
__global__ void prolong_kernel(Box f, Box c, const float *__restrict__ xc, const float *__restrict__ xin,
                               float *__restrict__ xout)
{
  int I, J, K;
  long F;
  if (!box_index(f, I, J, K, F)) return;
  int ci[2], cj[2], ck[2];
  float wi[2], wj[2], wk[2];
  parents2(I, c.nx, ci, wi);
  parents2(J, c.ny, cj, wj);
  parents2(K, c.nz, ck, wk);
  float acc = 0;
#pragma unroll
  for (int a = 0; a < 2; a++)
#pragma unroll
    for (int b = 0; b < 2; b++)
#pragma unroll
      for (int d = 0; d < 2; d++)
        acc += (wi[a] * wj[b] * wk[d]) * xc[((long)ci[a] * c.ny + cj[b]) * c.nz + ck[d]];
  xout[F] = xin[F] + acc;
}

// xf += P xc
template <typename T>
__global__ void prolong_add_kernel(Box f, Box c, const T *__restrict__ xc, T *__restrict__ xf)
{
  int I, J, K;
  long F;
  if (!box_index(f, I, J, K, F)) return;
  int ci[2], cj[2], ck[2];
  float wi[2], wj[2], wk[2];  // 0, 1/2 or 1: their products are exact in either precision
  parents2(I, c.nx, ci, wi);
  parents2(J, c.ny, cj, wj);
  parents2(K, c.nz, ck, wk);
  T acc = 0;
#pragma unroll
  for (int a = 0; a < 2; a++)
#pragma unroll
    for (int b = 0; b < 2; b++)
#pragma unroll
      for (int d = 0; d < 2; d++)
        acc += T(wi[a] * wj[b] * wk[d]) * xc[((long)ci[a] * c.ny + cj[b]) * c.nz + ck[d]];
  xf[F] += acc;
}

// 1-D prolongation weight of the coarse node at offset t (in -1..1) from coarse node c at the fine node at offset e
// (in -2..2) from c's own fine node 2c+1: 1 if e = 2t, 1/2 if |e - 2t| = 1, else 0.
__host__ __device__ constexpr double pweight(int e, int t)
{
  return e == 2 * t ? 1.0 : (e - 2 * t == 1 || e - 2 * t == -1) ? 0.5 : 0.0;
}

// Ac = P^T Af P, one thread per coarse row (FP64).  Row C gathers over the 27 fine nodes f of its prolongation's
// support (weight P[f][C]) and their 27 stencil entries A[f][g].  The weights factor by direction: with f at offsets
// (a, b, d) from C's own fine node and g = f + (si, sj, sk), P[f][C] P[g][C + t] = X(a, si, ti) Y(b, sj, tj)
// Z(d, sk, tk).  The x part (a, si) is a loop; inside it the 81 (b, d, sj, sk) terms have compile-time y and z weights
// and are summed into 3 x 3 partial sums in registers (every load unconditional: a support node outside the fine box
// reads entry 0 with weight 0), which then meet the three x weights.  Entries toward coarse nodes outside the box are
// zero, as the assembly leaves them on the finest level.
__global__ void __launch_bounds__(256) rap_kernel(Box f, Box c, const double *__restrict__ Af, double *__restrict__ Ac)
{
  const long C = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long nc = c.n(), nf = f.n();
  if (C >= nc) return;
  int I, J, K;
  unrank(c, C, I, J, K);
  double acc[27];
#pragma unroll
  for (int t = 0; t < 27; t++) acc[t] = 0.0;
  const long sy = f.nz, sx = (long)f.ny * f.nz;
#pragma unroll 1
  for (int a = -1; a <= 1; a++) {
    const int fi = 2 * I + 1 + a;  // >= 0
    if (fi >= f.nx) continue;
#pragma unroll 1
    for (int si = -1; si <= 1; si++) {
      double T[9];
#pragma unroll
      for (int u = 0; u < 9; u++) T[u] = 0.0;
#pragma unroll
      for (int b = -1; b <= 1; b++)
#pragma unroll
        for (int d = -1; d <= 1; d++) {
          const int fj = 2 * J + 1 + b, fk = 2 * K + 1 + d;
          const bool in = fj < f.ny && fk < f.nz;
          const long fr = in ? fi * sx + fj * sy + fk : 0;
          const double w = in ? pweight(b, 0) * pweight(d, 0) : 0.0;
#pragma unroll
          for (int sj = -1; sj <= 1; sj++)
#pragma unroll
            for (int sk = -1; sk <= 1; sk++) {
              const double v = w * Af[((si + 1) * 9 + (sj + 1) * 3 + (sk + 1)) * nf + fr];
#pragma unroll
              for (int tj = -1; tj <= 1; tj++)
#pragma unroll
                for (int tk = -1; tk <= 1; tk++) {
                  const double q = pweight(b + sj, tj) * pweight(d + sk, tk);
                  if (q != 0.0) T[(tj + 1) * 3 + (tk + 1)] = fma(v, q, T[(tj + 1) * 3 + (tk + 1)]);
                }
            }
        }
      const double wa = pweight(a, 0);
#pragma unroll
      for (int ti = -1; ti <= 1; ti++) {
        const double x = wa * pweight(a + si, ti);
#pragma unroll
        for (int u = 0; u < 9; u++) acc[(ti + 1) * 9 + u] = fma(x, T[u], acc[(ti + 1) * 9 + u]);
      }
    }
  }
#pragma unroll
  for (int t = 0; t < 27; t++)
    Ac[t * nc + C] = inside(c, I + t / 9 - 1, J + (t / 3) % 3 - 1, K + t % 3 - 1) ? acc[t] : 0.0;
}

// x = Ainv b, one wavefront per row (lanes stride the row, then a butterfly reduction).
template <typename T>
__global__ void dense_mv_kernel(int n, const T *__restrict__ Ainv, const T *__restrict__ b, T *__restrict__ x)
{
  const int r = blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64, lane = threadIdx.x & 63;
  if (r >= n) return;  // whole wavefronts exit together
  T acc = 0;
  for (int c = lane; c < n; c += 64) acc += Ainv[(long)r * n + c] * b[c];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
  if (lane == 0) x[r] = acc;
}

template <typename T>
__global__ void to_f32(long n, const double *__restrict__ a, float *__restrict__ b)
{
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = (float)a[i];
}

__global__ void sumsq_kernel(long n, const double *__restrict__ a, double *__restrict__ part)
{
  __shared__ double sh[256];
  double s = 0;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) s += a[i] * a[i];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

// Start vector of the power iteration: a counter-based hash (splitmix64) of the row, uniform in [0.25, 1.25).
__global__ void power_start_kernel(long n, double *__restrict__ v)
{
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint64_t z = (uint64_t)i + 0x9E3779B97F4A7C15ULL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  z ^= z >> 31;
  v[i] = (double)(z >> 11) * (1.0 / 9007199254740992.0) + 0.25;
}

// v = w / ||w||, with ||w||^2 the sum of sumsq_kernel's 256 partial sums taken in order (the same sum on every
// block, no host round trip); block 0 also stores ||w|| (norm may be null).  v may be w.
__global__ void normalize_kernel(long n, const double *__restrict__ part, const double *w, double *v,
                                 double *__restrict__ norm)
{
  __shared__ double inv;
  if (threadIdx.x == 0) {
    double t = 0;
    for (int k = 0; k < 256; k++) t += part[k];
    const double nw = sqrt(t);
    inv = 1.0 / nw;
    if (blockIdx.x == 0 && norm) *norm = nw;
  }
  __syncthreads();
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) v[i] = inv * w[i];
}

__global__ void scale_kernel(long n, double s, const double *__restrict__ dinv, const double *__restrict__ a,
                             double *__restrict__ b)
{
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = s * (dinv ? dinv[i] : 1.0) * a[i];
}

inline dim3 grid_of(long n) { return dim3((unsigned)((n + 255) / 256)); }

}  // namespace

dErr dohpMatPrepareStencil_Private(_p_dohpMat *A, dohpFS fs, bool *box, hipStream_t st)
{
  (void)st;
  const long L0 = (long)fs->p * fs->M[0] + 1, L1 = (long)fs->p * fs->M[1] + 1, L2 = (long)fs->p * fs->M[2] + 1;
  const int m = fs->dirichlet_mask;
  Box b{(int)(L0 - ((m & 1) ? 2 : 0)), (int)(L1 - ((m & 1) ? 2 : 0)), (int)(L2 - ((m & 2) ? 1 : 0) - ((m & 4) ? 1 : 0))};
  *box = false;
  if (b.n() != A->nrows || b.nx <= 0 || b.ny <= 0 || b.nz <= 0) return 0;
  A->box[0] = b.nx;
  A->box[1] = b.ny;
  A->box[2] = b.nz;
  if (!A->st64.holds(27 * b.n()) || !A->st32.holds(27 * b.n())) {  // both or neither
    dErr err = A->st64.alloc(27 * b.n());
    if (!err) err = A->st32.alloc(27 * b.n());
    if (err) {
      A->st64.release();
      A->st32.release();
      return err;
    }
  }
  *box = true;
  return 0;
}

dErr dohpMatStencilToCSR_Private(_p_dohpMat *A, hipStream_t st)
{
  const Box b{A->box[0], A->box[1], A->box[2]};
  hipLaunchKernelGGL(stencil_to_csr_kernel, dim3((unsigned)((b.n() + 255) / 256)), dim3(256), 0, st, b, A->rowptr.ptr,
                     A->col.ptr, A->st64.ptr, A->val.ptr, A->val32.ptr, A->st32.ptr);
  DOHP_HIP(hipGetLastError());
  return 0;
}

dErr dohpStencilMult_Private(const _p_dohpMat *A, const double *x, double *y, hipStream_t st)
{
  const Box b{A->box[0], A->box[1], A->box[2]};
  hipLaunchKernelGGL(stencil_mult_kernel<double>, box_grid(b), box_block(), 0, st, b, A->st64.ptr, x, y);
  DOHP_HIP(hipGetLastError());
  return 0;
}

dErr dohpStencilMultF32_Private(const _p_dohpMat *A, const float *x, float *y, hipStream_t st)
{
  const Box b{A->box[0], A->box[1], A->box[2]};
  hipLaunchKernelGGL(stencil_mult_kernel<float>, box_grid(b), box_block(), 0, st, b, A->st32.ptr, x, y);
  DOHP_HIP(hipGetLastError());
  return 0;
}

struct MGLevel {
  Box box{};
  long n = 0;
  const double *A64 = nullptr;  // full 27-point stencil (finest level: the Q1 matrix's own)
  DevBuf<double> own64, H64, dinv64, b64, x64, r64, d64;  // H: symmetric half stencil [14][n]
  DevBuf<float> H32, dinv32, b32, x32, r32, d32;
  DevBuf<_Float16> H16;        // FP16 upper off-diagonals, pairs of directions per row [7][n][2] (half16_tile_row)
  DevBuf<float> dg16, sc16, t32;  // their diagonal and row scales; the second iterate buffer of the FP16 V-cycle
  double lmax = 0;
};

struct _p_dohpMG {
  std::vector<MGLevel> L;
  int degree = 2;
  double lo = 0.1, hi = 1.1;
  int ncoarse = 0;
  int coeff16 = 1;  // the FP32 V-cycle reads the FP16 coefficients
  DevBuf<double> cinv64;
  DevBuf<float> cinv32;
  DevBuf<double> part;  // the power iteration's 256 partial sums and last norm
  // V-cycles replayed as HIP graphs: the coarse levels' kernels run for a few microseconds each, less than their
  // launch costs the host, so the eager V-cycle leaves the GPU waiting; a graph is captured (on a private stream)
  // per precision for the vectors it was applied to and replayed while they stay the same.
  struct Graph {
    const void *r = nullptr;  // the vectors of the captured graph, or of the last eager application
    void *z = nullptr;
    int coeff16 = -1;
    hipGraphExec_t exec = nullptr;
  };
  Graph g64, g32;
  int graphs = 1;
  hipStream_t cap = nullptr;
  void drop_graphs()
  {
    for (Graph *g : {&g64, &g32}) {
      if (g->exec) (void)hipGraphExecDestroy(g->exec);
      *g = Graph{};
    }
  }
  ~_p_dohpMG()
  {
    drop_graphs();
    if (cap) (void)hipStreamDestroy(cap);
  }
};

namespace {

// lmax of D^-1 A on a level by power iteration, in the level's work vectors (free outside a V-cycle).
dErr estimate_lmax(MGLevel &lv, DevBuf<double> &part, hipStream_t st, int its)
{
  DevBuf<double> &v = lv.r64, &w = lv.d64;
  hipLaunchKernelGGL(power_start_kernel, grid_of(lv.n), dim3(256), 0, st, lv.n, v.ptr);
  // the iterations stay on the device: one read of the last norm at the end
  hipLaunchKernelGGL(sumsq_kernel, dim3(256), dim3(256), 0, st, lv.n, v.ptr, part.ptr);
  hipLaunchKernelGGL(normalize_kernel, grid_of(lv.n), dim3(256), 0, st, lv.n, part.ptr, v.ptr, v.ptr, nullptr);
  for (int k = 0; k < its; k++) {
    residual<double>(lv.box, lv.H64.ptr, nullptr, v.ptr, w.ptr, st);  // w = -A v
    hipLaunchKernelGGL(scale_kernel, grid_of(lv.n), dim3(256), 0, st, lv.n, -1.0, lv.dinv64.ptr, w.ptr, w.ptr);
    hipLaunchKernelGGL(sumsq_kernel, dim3(256), dim3(256), 0, st, lv.n, w.ptr, part.ptr);
    hipLaunchKernelGGL(normalize_kernel, grid_of(lv.n), dim3(256), 0, st, lv.n, part.ptr, w.ptr, v.ptr,
                       part.ptr + 256);
  }
  DOHP_HIP(hipGetLastError());
  double nw = 0;
  if (its > 0) DOHP_HIP(hipMemcpyAsync(&nw, part.ptr + 256, sizeof(double), hipMemcpyDeviceToHost, st));
  DOHP_HIP(hipStreamSynchronize(st));
  lv.lmax = nw;
  return 0;
}

// Explicit inverse of the (symmetric positive definite) coarsest matrix by Gauss-Jordan, host FP64.
dErr coarse_inverse(const MGLevel &lv, hipStream_t st, std::vector<double> &inv)
{
  const long n = lv.n;
  DOHP_CHECK(n <= 4096, dohpERR_ARG, "multigrid: the coarsest level is too large for a direct solve");
  std::vector<double> S(27 * n);
  DOHP_HIP(hipMemcpyAsync(S.data(), lv.A64, sizeof(double) * 27 * n, hipMemcpyDeviceToHost, st));
  DOHP_HIP(hipStreamSynchronize(st));
  std::vector<double> A(n * n, 0.0);
  const Box b = lv.box;
  for (long r = 0; r < n; r++) {
    const int I = (int)(r / ((long)b.ny * b.nz)), J = (int)((r / b.nz) % b.ny), K = (int)(r % b.nz);
    for (int s = 0; s < 27; s++) {
      const int i2 = I + s / 9 - 1, j2 = J + (s / 3) % 3 - 1, k2 = K + s % 3 - 1;
      if (i2 < 0 || i2 >= b.nx || j2 < 0 || j2 >= b.ny || k2 < 0 || k2 >= b.nz) continue;
      A[r * n + ((long)i2 * b.ny + j2) * b.nz + k2] += S[s * n + r];
    }
  }
  inv.assign(n * n, 0.0);
  for (long i = 0; i < n; i++) inv[i * n + i] = 1.0;
  for (long k = 0; k < n; k++) {
    const double piv = A[k * n + k];
    DOHP_CHECK(piv > 0.0, dohpERR_GENERIC, "multigrid: coarsest matrix is not positive definite");
    const double ip = 1.0 / piv;
    for (long j = 0; j < n; j++) {
      A[k * n + j] *= ip;
      inv[k * n + j] *= ip;
    }
    for (long i = 0; i < n; i++) {
      if (i == k) continue;
      const double f = A[i * n + k];
      if (f == 0.0) continue;
      for (long j = 0; j < n; j++) {
        A[i * n + j] -= f * A[k * n + j];
        inv[i * n + j] -= f * inv[k * n + j];
      }
    }
  }
  return 0;
}

template <typename T>
struct LvPtr;
template <>
struct LvPtr<double> {
  static const double *A(const MGLevel &l) { return l.H64.ptr; }
  static const double *dinv(const MGLevel &l) { return l.dinv64.ptr; }
  static double *b(MGLevel &l) { return l.b64.ptr; }
  static double *x(MGLevel &l) { return l.x64.ptr; }
  static double *r(MGLevel &l) { return l.r64.ptr; }
  static double *d(MGLevel &l) { return l.d64.ptr; }
  static const double *cinv(const _p_dohpMG *m) { return m->cinv64.ptr; }
};
template <>
struct LvPtr<float> {
  static const float *A(const MGLevel &l) { return l.H32.ptr; }
  static const float *dinv(const MGLevel &l) { return l.dinv32.ptr; }
  static float *b(MGLevel &l) { return l.b32.ptr; }
  static float *x(MGLevel &l) { return l.x32.ptr; }
  static float *r(MGLevel &l) { return l.r32.ptr; }
  static float *d(MGLevel &l) { return l.d32.ptr; }
  static const float *cinv(const _p_dohpMG *m) { return m->cinv32.ptr; }
};

// x <- x + p(D^-1 A) D^-1 (b - A x), Chebyshev of the given degree on [lo, hi] * lmax.
template <typename T>
void smooth(const _p_dohpMG *mg, MGLevel &lv, const T *b, T *x, bool zero_init, hipStream_t st)
{
  using P = LvPtr<T>;
  const double lmin = mg->lo * lv.lmax, lmaxs = mg->hi * lv.lmax;
  const double theta = 0.5 * (lmaxs + lmin), delta = 0.5 * (lmaxs - lmin), sigma = theta / delta;
  double rho = 1.0 / sigma;
  T *r = P::r(lv), *d = P::d(lv);
  const dim3 g = grid_of(lv.n);
  if (zero_init)
    hipLaunchKernelGGL(cheb_step_kernel<T>, g, dim3(256), 0, st, lv.n, T(0), (T)(1.0 / theta), P::dinv(lv), b, d, x, 1);
  else {
    residual<T>(lv.box, P::A(lv), b, x, r, st);
    hipLaunchKernelGGL(cheb_step_kernel<T>, g, dim3(256), 0, st, lv.n, T(0), (T)(1.0 / theta), P::dinv(lv), r, d, x, 0);
  }
  for (int k = 1; k < mg->degree; k++) {
    residual<T>(lv.box, P::A(lv), b, x, r, st);
    const double rho_new = 1.0 / (2.0 * sigma - rho);
    hipLaunchKernelGGL(cheb_step_kernel<T>, g, dim3(256), 0, st, lv.n, (T)(rho_new * rho), (T)(2.0 * rho_new / delta),
                       P::dinv(lv), r, d, x, 0);
    rho = rho_new;
  }
}

// The FP32 smoother on the FP16 coefficients: Chebyshev steps alternate the iterate between cur and other (zero_init:
// cur is first set from b).  Returns the buffer that holds the result.
float *smooth16(const _p_dohpMG *mg, MGLevel &lv, const float *b, float *cur, float *other, bool zero_init,
                hipStream_t st)
{
  const double lmin = mg->lo * lv.lmax, lmaxs = mg->hi * lv.lmax;
  const double theta = 0.5 * (lmaxs + lmin), delta = 0.5 * (lmaxs - lmin), sigma = theta / delta;
  double rho = 1.0 / sigma;
  const int keep_d = mg->degree > 1;
  if (zero_init) {
    hipLaunchKernelGGL(half16_cheb_first_kernel, grid_of(lv.n), dim3(256), 0, st, lv.n, (float)(1.0 / theta),
                       lv.dg16.ptr, b, cur, lv.d32.ptr, keep_d);
  } else {
    hipLaunchKernelGGL(half16_cheb_kernel, box_grid(lv.box), box_block(), 0, st, lv.box, lv.H16.ptr, lv.dg16.ptr,
                       lv.sc16.ptr, b, cur, other, lv.d32.ptr, 0.f, (float)(1.0 / theta), keep_d);
    std::swap(cur, other);
  }
  for (int k = 1; k < mg->degree; k++) {
    const double rho_new = 1.0 / (2.0 * sigma - rho);
    hipLaunchKernelGGL(half16_cheb_kernel, box_grid(lv.box), box_block(), 0, st, lv.box, lv.H16.ptr, lv.dg16.ptr,
                       lv.sc16.ptr, b, cur, other, lv.d32.ptr, (float)(rho_new * rho), (float)(2.0 * rho_new / delta),
                       keep_d);
    std::swap(cur, other);
    rho = rho_new;
  }
  return cur;
}

// FP32 V-cycle on the FP16 coefficients.  The iterate alternates between x and the level's second buffer: the
// pre-smoother's steps, the prolongation and the post-smoother's steps make an even number of moves, so the result
// ends in x.
void vcycle16(_p_dohpMG *mg, size_t l, const float *b, float *x, hipStream_t st)
{
  MGLevel &lv = mg->L[l];
  if (l + 1 == mg->L.size()) {
    hipLaunchKernelGGL(dense_mv_kernel<float>, dim3((unsigned)((lv.n + 3) / 4)), dim3(256), 0, st, (int)lv.n,
                       mg->cinv32.ptr, b, x);
    return;
  }
  MGLevel &cv = mg->L[l + 1];
  float *cur = smooth16(mg, lv, b, x, lv.t32.ptr, true, st);
  hipLaunchKernelGGL(half16_residual_kernel, box_grid(lv.box), box_block(), 0, st, lv.box, lv.H16.ptr, lv.dg16.ptr,
                     lv.sc16.ptr, b, cur, lv.r32.ptr);
  hipLaunchKernelGGL(restrict_kernel<float>, box_grid(cv.box), box_block(), 0, st, lv.box, cv.box, lv.r32.ptr,
                     cv.b32.ptr);
  vcycle16(mg, l + 1, cv.b32.ptr, cv.x32.ptr, st);
  float *other = cur == x ? lv.t32.ptr : x;
  hipLaunchKernelGGL(prolong_kernel, box_grid(lv.box), box_block(), 0, st, lv.box, cv.box, cv.x32.ptr, cur, other);
  cur = smooth16(mg, lv, b, other, cur, false, st);
  if (cur != x)  // (not reached: the moves are even)
    (void)hipMemcpyAsync(x, cur, sizeof(float) * lv.n, hipMemcpyDeviceToDevice, st);
}

template <typename T>
void vcycle(_p_dohpMG *mg, size_t l, const T *b, T *x, hipStream_t st)
{
  using P = LvPtr<T>;
  MGLevel &lv = mg->L[l];
  if (l + 1 == mg->L.size()) {
    hipLaunchKernelGGL(dense_mv_kernel<T>, dim3((unsigned)((lv.n + 3) / 4)), dim3(256), 0, st, (int)lv.n,
                       P::cinv(mg), b, x);
    return;
  }
  MGLevel &cv = mg->L[l + 1];
  smooth<T>(mg, lv, b, x, true, st);
  residual<T>(lv.box, P::A(lv), b, x, P::r(lv), st);
  hipLaunchKernelGGL(restrict_kernel<T>, box_grid(cv.box), box_block(), 0, st, lv.box, cv.box, P::r(lv), P::b(cv));
  vcycle<T>(mg, l + 1, P::b(cv), P::x(cv), st);
  hipLaunchKernelGGL(prolong_add_kernel<T>, box_grid(lv.box), box_block(), 0, st, lv.box, cv.box, P::x(cv), x);
  smooth<T>(mg, lv, b, x, false, st);
}

}  // namespace

namespace {

// The operators of every level from the finest level's stencil (L[0].A64): Galerkin products, half stencils in FP64,
// FP32 and FP16, Jacobi scalings, eigenvalue estimates and the coarsest level's inverse, into the hierarchy's buffers.
dErr mg_fill(_p_dohpMG *mg, hipStream_t st)
{
  mg->drop_graphs();  // their kernels carry this hierarchy's smoothing coefficients
  for (size_t l = 1; l < mg->L.size(); l++) {
    const MGLevel &f = mg->L[l - 1];
    MGLevel &c = mg->L[l];
    hipLaunchKernelGGL(rap_kernel, grid_of(c.n), dim3(256), 0, st, f.box, c.box, f.A64, c.own64.ptr);
  }
  for (size_t l = 0; l < mg->L.size(); l++) {
    MGLevel &lv = mg->L[l];
    const long n = lv.n;
    hipLaunchKernelGGL(half_from_full_kernel, grid_of(n), dim3(256), 0, st, n, lv.A64, lv.H64.ptr, lv.H32.ptr);
    if (l + 1 < mg->L.size())
      hipLaunchKernelGGL(half16_setup_kernel, box_grid(lv.box), box_block(), 0, st, lv.box, lv.A64, lv.H16.ptr,
                         lv.dg16.ptr, lv.sc16.ptr);
    hipLaunchKernelGGL(diag_inv_kernel<double>, grid_of(n), dim3(256), 0, st, n, lv.A64, lv.dinv64.ptr);
    hipLaunchKernelGGL(to_f32<float>, grid_of(n), dim3(256), 0, st, n, lv.dinv64.ptr, lv.dinv32.ptr);
    if (l + 1 < mg->L.size()) DOHP_CALL(estimate_lmax(lv, mg->part, st, 15));
  }
  std::vector<double> inv;
  DOHP_CALL(coarse_inverse(mg->L.back(), st, inv));
  mg->ncoarse = (int)mg->L.back().n;
  DOHP_CALL(mg->cinv64.upload(inv.data(), inv.size()));
  std::vector<float> inv32(inv.begin(), inv.end());
  DOHP_CALL(mg->cinv32.upload(inv32.data(), inv32.size()));
  if (hipStreamSynchronize(st) != hipSuccess) {
    set_error("multigrid setup failed");
    return dohpERR_HIP;
  }
  return 0;
}

}  // namespace

extern "C" dErr dohpMGCreate(dohpMat q1, int smooth_degree, int max_coarse, dohpMG *out, void *stream)
{
  DOHP_CHECK(q1 && out, dohpERR_ARG, "null argument");
  DOHP_CHECK(q1->st64.ptr, dohpERR_ARG, "multigrid needs the stencil form of a block-space Q1 matrix");
  hipStream_t st = (hipStream_t)stream;
  *out = nullptr;
  _p_dohpMG *mg = new _p_dohpMG;
  mg->degree = smooth_degree > 0 ? smooth_degree : 2;
  const long nmax = max_coarse > 0 ? max_coarse : 512;
  dErr err = mg->part.alloc(257);
  {
    MGLevel lv;
    lv.box = Box{q1->box[0], q1->box[1], q1->box[2]};
    lv.n = lv.box.n();
    lv.A64 = q1->st64.ptr;
    mg->L.push_back(std::move(lv));
  }
  while (!err) {
    MGLevel &f = mg->L.back();
    if (f.n <= nmax) break;
    const Box cb{f.box.nx / 2, f.box.ny / 2, f.box.nz / 2};
    if (cb.nx < 1 || cb.ny < 1 || cb.nz < 1) break;
    MGLevel c;
    c.box = cb;
    c.n = cb.n();
    err = c.own64.alloc(27 * c.n);
    if (err) break;
    c.A64 = c.own64.ptr;
    mg->L.push_back(std::move(c));
  }
  for (size_t l = 0; !err && l < mg->L.size(); l++) {
    MGLevel &lv = mg->L[l];
    const long n = lv.n;
    err = lv.dinv64.alloc(n);
    if (!err) err = lv.dinv32.alloc(n);
    if (!err) err = lv.H64.alloc(14 * n);
    if (!err) err = lv.H32.alloc(14 * n);
    if (!err) err = lv.r64.alloc(n);
    if (!err) err = lv.d64.alloc(n);
    if (!err) err = lv.r32.alloc(n);
    if (!err) err = lv.d32.alloc(n);
    if (!err && l > 0) err = lv.b64.alloc(n);
    if (!err && l > 0) err = lv.x64.alloc(n);
    if (!err && l > 0) err = lv.b32.alloc(n);
    if (!err && l > 0) err = lv.x32.alloc(n);
    if (!err && l + 1 < mg->L.size()) err = lv.H16.alloc(14 * n);
    if (!err && l + 1 < mg->L.size()) err = lv.dg16.alloc(n);
    if (!err && l + 1 < mg->L.size()) err = lv.sc16.alloc(n);
    if (!err && l + 1 < mg->L.size()) err = lv.t32.alloc(n);
  }
  if (!err) err = mg_fill(mg, st);
  if (err) {
    delete mg;
    return err;
  }
  *out = mg;
  return 0;
}

extern "C" dErr dohpMGRefresh(dohpMG mg, dohpMat q1, void *stream)
{
  DOHP_CHECK(mg && q1, dohpERR_ARG, "null argument");
  DOHP_CHECK(q1->st64.ptr, dohpERR_ARG, "multigrid needs the stencil form of a block-space Q1 matrix");
  const Box &b = mg->L[0].box;
  DOHP_CHECK(q1->box[0] == b.nx && q1->box[1] == b.ny && q1->box[2] == b.nz, dohpERR_ARG,
             "multigrid refresh needs a matrix on the same box");
  mg->L[0].A64 = q1->st64.ptr;
  return mg_fill(mg, (hipStream_t)stream);
}

extern "C" dErr dohpMGGetInfo(dohpMG mg, int *nlevels, long *ncoarse, double *lmax_fine)
{
  DOHP_CHECK(mg, dohpERR_ARG, "null multigrid");
  if (nlevels) *nlevels = (int)mg->L.size();
  if (ncoarse) *ncoarse = mg->ncoarse;
  if (lmax_fine) *lmax_fine = mg->L[0].lmax;
  return 0;
}

extern "C" dErr dohpMGGetLevel(dohpMG mg, int level, int *box, double *stencil_host)
{
  DOHP_CHECK(mg && level >= 0 && level < (int)mg->L.size(), dohpERR_ARG, "no such multigrid level");
  const MGLevel &lv = mg->L[level];
  if (box) {
    box[0] = lv.box.nx;
    box[1] = lv.box.ny;
    box[2] = lv.box.nz;
  }
  if (stencil_host) {
    DOHP_HIP(hipDeviceSynchronize());
    DOHP_HIP(hipMemcpy(stencil_host, lv.A64, sizeof(double) * 27 * lv.n, hipMemcpyDeviceToHost));
  }
  return 0;
}

namespace {

// Applies the V-cycle `run(stream)` for vectors (r, z): replays the graph captured for them; the second consecutive
// application to the same vectors captures it (a Krylov loop applies its preconditioner to the same work vectors
// every iteration; one-off applications stay eager).  A capture that fails leaves the hierarchy on eager launches.
template <class Run>
dErr vcycle_graph(_p_dohpMG *mg, _p_dohpMG::Graph &g, const void *r, void *z, hipStream_t st, Run run)
{
  const bool same = g.r == r && g.z == z && g.coeff16 == mg->coeff16;
  if (!mg->graphs || !same) {
    if (g.exec) (void)hipGraphExecDestroy(g.exec);
    g = _p_dohpMG::Graph{};
    g.r = r;
    g.z = z;
    g.coeff16 = mg->coeff16;
    run(st);
    DOHP_HIP(hipGetLastError());
    return 0;
  }
  if (!g.exec) {
    if (!mg->cap) DOHP_HIP(hipStreamCreateWithFlags(&mg->cap, hipStreamNonBlocking));
    hipGraph_t graph = nullptr;
    bool ok = hipStreamBeginCapture(mg->cap, hipStreamCaptureModeThreadLocal) == hipSuccess;
    if (ok) {
      run(mg->cap);
      ok = hipStreamEndCapture(mg->cap, &graph) == hipSuccess && graph;
    }
    if (ok) ok = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0) == hipSuccess;
    if (graph) (void)hipGraphDestroy(graph);
    if (!ok) {  // eager from now on
      (void)hipGetLastError();
      if (g.exec) (void)hipGraphExecDestroy(g.exec);
      g = _p_dohpMG::Graph{};
      mg->graphs = 0;
      run(st);
      DOHP_HIP(hipGetLastError());
      return 0;
    }
  }
  DOHP_HIP(hipGraphLaunch(g.exec, st));
  return 0;
}

}  // namespace

extern "C" dErr dohpMGApply(dohpMG mg, const double *r, double *z, void *stream)
{
  DOHP_CHECK(mg && r && z, dohpERR_ARG, "null argument");
  return vcycle_graph(mg, mg->g64, r, z, (hipStream_t)stream,
                      [&](hipStream_t s) { vcycle<double>(mg, 0, r, z, s); });
}

extern "C" dErr dohpMGSetGraphs(dohpMG mg, int on)
{
  DOHP_CHECK(mg && (on == 0 || on == 1), dohpERR_ARG, "graphs must be 0 (eager launches) or 1");
  mg->drop_graphs();
  mg->graphs = on;
  return 0;
}

extern "C" dErr dohpMGApplyF32(dohpMG mg, const float *r, float *z, void *stream)
{
  DOHP_CHECK(mg && r && z, dohpERR_ARG, "null argument");
  return vcycle_graph(mg, mg->g32, r, z, (hipStream_t)stream, [&](hipStream_t s) {
    if (mg->coeff16 && mg->L.size() > 1)
      vcycle16(mg, 0, r, z, s);
    else
      vcycle<float>(mg, 0, r, z, s);
  });
}

extern "C" dErr dohpMGResidual(dohpMG mg, int level, const double *b, const double *x, double *r, void *stream)
{
  DOHP_CHECK(mg && level >= 0 && level < (int)mg->L.size(), dohpERR_ARG, "no such multigrid level");
  DOHP_CHECK(x && r, dohpERR_ARG, "null argument");
  const MGLevel &lv = mg->L[level];
  residual<double>(lv.box, lv.H64.ptr, b, x, r, (hipStream_t)stream);
  DOHP_HIP(hipGetLastError());
  return 0;
}

extern "C" dErr dohpMGResidualF32(dohpMG mg, int level, const float *b, const float *x, float *r, void *stream)
{
  DOHP_CHECK(mg && level >= 0 && level < (int)mg->L.size(), dohpERR_ARG, "no such multigrid level");
  DOHP_CHECK(x && r, dohpERR_ARG, "null argument");
  const MGLevel &lv = mg->L[level];
  residual<float>(lv.box, lv.H32.ptr, b, x, r, (hipStream_t)stream);
  DOHP_HIP(hipGetLastError());
  return 0;
}

extern "C" dErr dohpMGResidualF16(dohpMG mg, int level, const float *b, const float *x, float *r, void *stream)
{
  DOHP_CHECK(mg && level >= 0 && level + 1 < (int)mg->L.size(), dohpERR_ARG,
             "no FP16 coefficients on this multigrid level (the coarsest is solved directly)");
  DOHP_CHECK(x && r, dohpERR_ARG, "null argument");
  const MGLevel &lv = mg->L[level];
  hipLaunchKernelGGL(half16_residual_kernel, box_grid(lv.box), box_block(), 0, (hipStream_t)stream, lv.box,
                     lv.H16.ptr, lv.dg16.ptr, lv.sc16.ptr, b, x, r);
  DOHP_HIP(hipGetLastError());
  return 0;
}

extern "C" dErr dohpMGSetCoefficients(dohpMG mg, int fp16)
{
  DOHP_CHECK(mg && (fp16 == 0 || fp16 == 1), dohpERR_ARG, "coefficients must be 0 (FP32) or 1 (FP16)");
  mg->coeff16 = fp16;
  return 0;
}

extern "C" dErr dohpMGDestroy(dohpMG *mg)
{
  if (mg && *mg) {
    delete *mg;
    *mg = nullptr;
  }
  return 0;
}
