// scf_device.hip -- the whole SCF iteration on the GPU.
//
// Restates the reference's RHF and UHF updates (rhfscf.c:21-71,
// uhfscf.c:22-80), their drivers (rhfscf.c:73-95, uhfscf.c:82-100), the
// orthogonaliser X = U s^-1/2 (rhfscf.c:432-452) and the core-Hamiltonian
// guess (rhfscf.c:454-471) with every matrix resident in HBM:
//   G       the library's own device G build (G zeroed, then G += 2J - K, or
//           the UHF pair);
//   F' = X^T (H + G) X    two rocBLAS dgemm (spack_trans, spack_math.c:321-371);
//   eigen   rocSOLVER dsyevd (the role of spack_diag, spack_math.c:299-319),
//           then the reference's eigenvector sign convention check_phase
//           (spack_math.c:274-297, below);
//   C = X V dgemm (form_cmatrix, scfaux.c:5-21);
//   P = C_occ C_occ^T     dsyrk + pack (form_pmatrix, scfaux.c:23-41);
//   sums    calc_energy and pdiffNorm (scfaux.c:97-136), and the three dot
//           products of scfAccelerator (scfaux.c:64-95), each in a fixed
//           summation order;
//   damping the averaging / two-point Anderson updates (rhfscf.c:60-69);
//           or, with scf_accel 3 (no counterpart in the reference), Pulay
//           DIIS on the Fock matrices (RHF).
// The host keeps only the iteration logic: per iteration it reads back the
// energy and density change (two doubles) to take the reference's
// convergence and damping decisions.
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <cfloat>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <utility>
#include <vector>

#include "batch.h"
#include "patmol_hip.h"

extern "C" {
int number_of_orbitals(void);
}

namespace pm {

int aux_fail(int code, const char *msg);

namespace {

int g_scf_density = 1;   // 1: purification inside the loop, 2: the same with separate trace totals, 0: eigensolver
int g_scf_incremental = 0;   // RHF: 0 full G every iteration, k > 0 full G every k iterations, delta builds between

constexpr int RED_THREADS = 1024;

// full symmetric F = a + b from packed lower triangles (b may be null)
__global__ void sym_from_packed_kernel(const double *__restrict__ a, const double *__restrict__ b,
                                       double *__restrict__ F, int n)
{
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= (int64_t)n * n) return;
    const int i = (int)(k / n), j = (int)(k % n);
    const int hi = i > j ? i : j, lo = i > j ? j : i;
    const int64_t ij = (int64_t)hi * (hi + 1) / 2 + lo;
    F[k] = b ? a[ij] + b[ij] : a[ij];
}

// check_phase (spack_math.c:274-297) on every eigenvector (column j of V,
// contiguous): the reference scans v_1..v_{n-1} for |v_k| strictly above a
// running maximum that starts at the SIGNED v_0, and flips the vector when the
// component it lands on is negative.  Equivalently, with M = max_{k>=1} |v_k|
// at its first index k*: flip iff M > v_0 and v_{k*} < 0.
__global__ __launch_bounds__(256) void phase_kernel(double *__restrict__ V, int n)
{
    __shared__ double sm[256];
    __shared__ int si[256];
    __shared__ int flip;
    double *v = V + (size_t)blockIdx.x * n;
    double m = -1.0;
    int mi = n;
    for (int k = 1 + threadIdx.x; k < n; k += blockDim.x) {
        const double t = fabs(v[k]);
        if (t > m) {
            m = t;
            mi = k;
        }
    }
    sm[threadIdx.x] = m;
    si[threadIdx.x] = mi;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            const double o = sm[threadIdx.x + s];
            const int oi = si[threadIdx.x + s];
            if (o > sm[threadIdx.x] || (o == sm[threadIdx.x] && oi < si[threadIdx.x])) {
                sm[threadIdx.x] = o;
                si[threadIdx.x] = oi;
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) flip = (si[0] < n && sm[0] > v[0] && v[si[0]] < 0.0) ? 1 : 0;
    __syncthreads();
    if (flip)
        for (int k = threadIdx.x; k < n; k += blockDim.x) v[k] = -v[k];
}

// X(:, j) = U(:, j) / sqrt(s_j)  (rhfscf.c:447-451)
__global__ void scale_columns_kernel(double *__restrict__ X, const double *__restrict__ U,
                                     const double *__restrict__ s, int n)
{
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= (int64_t)n * n) return;
    X[k] = U[k] * (1.0 / sqrt(s[k / n]));
}

// packed P(i,j), i >= j, from the lower triangle of column-major D
__global__ void pack_lower_kernel(const double *__restrict__ D, double *__restrict__ P, int n)
{
    const int i = blockIdx.x;
    for (int j = threadIdx.x; j <= i; j += blockDim.x) P[(size_t)i * (i + 1) / 2 + j] = D[(size_t)j * n + i];
}

// Weighted sums over packed triangles, off-diagonal weight 1 and diagonal
// 1/2 (calc_energy and pdiffNorm, scfaux.c:97-136):
//   out[0] = sum w H P1, out[1] = sum w G P1, out[2] = sum w (P - P1)^2,
// in a fixed order: SUMS_BLOCKS blocks each reduce one contiguous chunk of
// the packed triangle (strided over the threads, then a fixed tree), and one
// thread per sum adds the block partials in block order.
constexpr int SUMS_BLOCKS = 64;
__global__ __launch_bounds__(256) void scf_sums_part_kernel(const double *__restrict__ P, const double *__restrict__ P1,
                                                           const double *__restrict__ H, const double *__restrict__ G,
                                                           int64_t no2, double *__restrict__ part)
{
    __shared__ double red[3][256];
    const int64_t chunk = (no2 + gridDim.x - 1) / gridDim.x;
    const int64_t lo = (int64_t)blockIdx.x * chunk, hi = lo + chunk < no2 ? lo + chunk : no2;
    double a = 0.0, b = 0.0, c = 0.0;
    for (int64_t ij = lo + threadIdx.x; ij < hi; ij += 256) {
        // row i of the packed index: i(i+1)/2 <= ij < (i+1)(i+2)/2
        int64_t i = (int64_t)((sqrt(8.0 * (double)ij + 1.0) - 1.0) * 0.5);
        while ((i + 1) * (i + 2) / 2 <= ij) ++i;
        while (i * (i + 1) / 2 > ij) --i;
        const double w = ij == i * (i + 3) / 2 ? 0.5 : 1.0, p1 = P1[ij], d = P[ij] - p1;
        a += w * H[ij] * p1;
        b += w * G[ij] * p1;
        c += w * d * d;
    }
    red[0][threadIdx.x] = a;
    red[1][threadIdx.x] = b;
    red[2][threadIdx.x] = c;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if (threadIdx.x < st)
            for (int q = 0; q < 3; ++q) red[q][threadIdx.x] += red[q][threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x < 3) part[3 * blockIdx.x + threadIdx.x] = red[threadIdx.x][0];
}

__global__ void scf_sums_final_kernel(const double *__restrict__ part, int nb, double *__restrict__ out)
{
    if (threadIdx.x >= 3) return;
    double v = 0.0;
    for (int b = 0; b < nb; ++b) v += part[3 * b + threadIdx.x];
    out[threadIdx.x] = v;
}

// dot products of scfAccelerator (scfaux.c:70-77)
__global__ __launch_bounds__(RED_THREADS) void accel_dots_kernel(const double *__restrict__ p0,
                                                                const double *__restrict__ p1,
                                                                const double *__restrict__ p2, int64_t no2,
                                                                double *__restrict__ out)
{
    __shared__ double red[3][RED_THREADS];
    double a = 0.0, b = 0.0, c = 0.0;
    for (int64_t i = threadIdx.x; i < no2; i += RED_THREADS) {
        const double s0 = p0[i] - p1[i], s1 = p1[i] - p2[i];
        a += s0 * s0;
        b += s0 * s1;
        c += s1 * s1;
    }
    red[0][threadIdx.x] = a;
    red[1][threadIdx.x] = b;
    red[2][threadIdx.x] = c;
    __syncthreads();
    for (int s = RED_THREADS / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s)
            for (int q = 0; q < 3; ++q) red[q][threadIdx.x] += red[q][threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int q = 0; q < 3; ++q) out[q] = red[q][0];
}

// p2 <- p1, p1 <- p0, p0 <- beta p1 + (1 - beta) p2   (scfaux.c:84-94)
__global__ void accel_update_kernel(double *__restrict__ p0, double *__restrict__ p1, double *__restrict__ p2,
                                    int64_t no2, double beta)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= no2) return;
    const double a = p0[i], b = p1[i];
    p2[i] = b;
    p1[i] = a;
    p0[i] = a * beta + b * (1.0 - beta);
}

// p <- (p1 + p2) / 2   (rhfscf.c:61-65)
__global__ void average_kernel(double *__restrict__ p, const double *__restrict__ p1, const double *__restrict__ p2,
                               int64_t no2)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < no2) p[i] = 0.5 * (p1[i] + p2[i]);
}

// ---- density by trace-correcting purification (Niklasson, Phys. Rev. B 66,
// 155115 (2002)) ----
// Gershgorin interval of symmetric F (column-major, n x n): per row the
// centre F_ii and radius sum_{j != i} |F_ij|; one block per row.
__global__ __launch_bounds__(256) void gersh_row_kernel(const double *__restrict__ F, int n, double *__restrict__ lo,
                                                        double *__restrict__ hi)
{
    __shared__ double red[256];
    const int i = blockIdx.x;
    double r = 0.0;
    for (int j = threadIdx.x; j < n; j += 256)
        if (j != i) r += fabs(F[(size_t)j * n + i]);
    red[threadIdx.x] = r;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if (threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const double c = F[(size_t)i * n + i];
        lo[i] = c - red[0];
        hi[i] = c + red[0];
    }
}

// bounds[0] = min lo, bounds[1] = max hi (one block)
__global__ __launch_bounds__(256) void gersh_reduce_kernel(const double *__restrict__ lo, const double *__restrict__ hi,
                                                           int n, double *__restrict__ bounds)
{
    __shared__ double a[256], b[256];
    double mn = 1e300, mx = -1e300;
    for (int i = threadIdx.x; i < n; i += 256) {
        mn = fmin(mn, lo[i]);
        mx = fmax(mx, hi[i]);
    }
    a[threadIdx.x] = mn;
    b[threadIdx.x] = mx;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if (threadIdx.x < st) {
            a[threadIdx.x] = fmin(a[threadIdx.x], a[threadIdx.x + st]);
            b[threadIdx.x] = fmax(b[threadIdx.x], b[threadIdx.x + st]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        bounds[0] = a[0];
        bounds[1] = b[0];
    }
}

// X0 = (emax I - F) / (emax - emin): spectrum mapped into [0, 1], occupied
// levels towards 1.  F is symmetrised on the way ((F + F^T) / 2), so that X0
// and every later X are exactly symmetric (tc2_step_kernel relies on it).
__global__ void tc2_init_kernel(const double *__restrict__ F, double *__restrict__ X, int n,
                                const double *__restrict__ bounds)
{
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= (int64_t)n * n) return;
    const double emin = bounds[0], emax = bounds[1];
    const double sc = 1.0 / (emax - emin);
    const int i = (int)(k % n), j = (int)(k / n);
    const double f = 0.5 * (F[k] + F[(int64_t)i * n + j]);
    X[k] = ((i == j ? emax : 0.0) - f) * sc;
}

// Sum of a and b over the 256 threads of a block in a fixed order (xor
// butterflies inside each wave, then the four waves in order); the result is
// valid in thread 0.
__device__ __forceinline__ void block_sum2_256(double &a, double &b, double *sh)
{
    for (int d = 32; d > 0; d >>= 1) {
        a += __shfl_xor(a, d, 64);
        b += __shfl_xor(b, d, 64);
    }
    const int wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        sh[2 * wave] = a;
        sh[2 * wave + 1] = b;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        a = ((sh[0] + sh[2]) + sh[4]) + sh[6];
        b = ((sh[1] + sh[3]) + sh[5]) + sh[7];
    }
}

// Purification steps in TC2_T x TC2_T tiles of X (column-major n x n), one
// 256-thread block per tile: block b covers rows (b % nt) TC2_T.. and columns
// (b / nt) TC2_T.., nt = ceil(n / TC2_T).  part[2b] = the tile's share of
// tr X, part[2b + 1] of sum X_ij^2 (= tr X^2 for symmetric X).
constexpr int TC2_T = 32;
// Convergence state of a purification, in two slots of st that alternate
// with the step index (a step reads slot TC2_REM + (step & 1) and writes the
// other, so no block sees a value written by its own launch): -1 while not
// converged; once tr X - tr X^2 < 1e-12 and |tr X - nocc| < 1e-8 the number
// of updates still to do (2, 1), then 0, after which steps only copy X.
constexpr int TC2_REM = 6;
__device__ __forceinline__ double tc2_next_rem(double rem, double tr, double sq, double nocc)
{
    if (rem >= 0.0) return rem - 1.0;
    return (tr - sq < 1e-12 && fabs(tr - nocc) < 1e-8) ? 2.0 : -1.0;
}
constexpr int TC2_KB = 14;   // k-steps of 4 whose operands a wave loads before their MFMAs
typedef double tc2_d4 __attribute__((ext_vector_type(4)));

__global__ __launch_bounds__(256) void tc2_partials_kernel(const double *__restrict__ X, int n,
                                                          double *__restrict__ part, double *__restrict__ st)
{
    if (blockIdx.x == 0 && threadIdx.x == 0) st[TC2_REM] = -1.0;   // step 0's state: not converged
    __shared__ double sh[8];
    const int nt = (n + TC2_T - 1) / TC2_T;
    const int i0 = (blockIdx.x % nt) * TC2_T, j0 = (blockIdx.x / nt) * TC2_T;
    double tr = 0.0, sq = 0.0;
    for (int e = threadIdx.x; e < TC2_T * TC2_T; e += 256) {
        const int i = i0 + e % TC2_T, j = j0 + e / TC2_T;
        if (i < n && j < n) {
            const double v = X[(size_t)j * n + i];
            sq += v * v;
            if (i == j) tr += v;
        }
    }
    block_sum2_256(tr, sq, sh);
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = tr;
        part[2 * blockIdx.x + 1] = sq;
    }
}

// Spectral bounds for a purification from the last F' of the channel (spin)
// whose eigenvalues are known (the guess, or a diagonalised iteration): by
// Weyl's inequality every eigenvalue moves by at most ||F - Fref||_2 <=
// ||F - Fref||_F, so [lo - d, hi + d] bounds the spectrum of F when [lo, hi]
// bounded that of Fref (d = the smaller of the Frobenius norm and the largest
// absolute column sum of the change, both bounds of its 2-norm, with a
// relative margin for rounding).  The interval is intersected with F's Gershgorin
// interval, which alone is about ten times too wide for an orthogonalised
// Fock matrix (at (H2O)8: Gershgorin width 5700 Eh, spectrum 540 Eh, the
// Weyl interval about 1400 Eh).
__global__ __launch_bounds__(256) void diff2_part_kernel(const double *__restrict__ A, const double *__restrict__ B,
                                                        int64_t nn, double *__restrict__ part)
{
    __shared__ double red[256];
    const int64_t chunk = (nn + gridDim.x - 1) / gridDim.x;
    const int64_t lo = (int64_t)blockIdx.x * chunk, hi = lo + chunk < nn ? lo + chunk : nn;
    double a = 0.0;
    for (int64_t k = lo + threadIdx.x; k < hi; k += 256) {
        const double d = A[k] - B[k];
        a += d * d;
    }
    red[threadIdx.x] = a;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if (threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// absolute column sums of A - B (column-major n x n), one block per column
__global__ __launch_bounds__(256) void diffcol_kernel(const double *__restrict__ A, const double *__restrict__ B, int n,
                                                     double *__restrict__ col)
{
    __shared__ double red[256];
    const size_t c0 = (size_t)blockIdx.x * n;
    double a = 0.0;
    for (int i = threadIdx.x; i < n; i += 256) a += fabs(A[c0 + i] - B[c0 + i]);
    red[threadIdx.x] = a;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if (threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x == 0) col[blockIdx.x] = red[0];
}

// gb[0..1]: F's Gershgorin interval in, the bounds used for F out; ref[0..1]:
// the bounds of Fref; part: Frobenius partials, col: column sums of F - Fref
// (one 64-thread wave)
__global__ __launch_bounds__(64) void tc2_bounds_kernel(const double *__restrict__ part, int nb,
                                                        const double *__restrict__ col, int n,
                                                        const double *__restrict__ ref, double *__restrict__ gb)
{
    double cmax = 0.0;
    for (int i = threadIdx.x; i < n; i += 64) cmax = fmax(cmax, col[i]);
    for (int d = 32; d > 0; d >>= 1) cmax = fmax(cmax, __shfl_xor(cmax, d, 64));
    if (threadIdx.x != 0) return;
    double s2 = 0.0;
    for (int b = 0; b < nb; ++b) s2 += part[b];
    const double w = fabs(ref[0]) + fabs(ref[1]);
    const double m = fmin(sqrt(s2), cmax) * (1.0 + 1e-12) + 1e-12 * w;
    gb[0] = fmax(gb[0], ref[0] - m);
    gb[1] = fmin(gb[1], ref[1] + m);
}

// bounds from the eigenvalues of F (ascending), widened by 1e-10 of their
// magnitude for the eigensolver's error
__global__ void tc2_exact_bounds_kernel(const double *__restrict__ ev, int n, double *__restrict__ prev)
{
    if (threadIdx.x != 0) return;
    const double w = fabs(ev[0]) + fabs(ev[n - 1]);
    prev[0] = ev[0] - 1e-10 * w;
    prev[1] = ev[n - 1] + 1e-10 * w;
}

// tr X and tr X^2 from the tile partials: lane-strided sums, then xor
// butterflies (every lane, and every block that does this, gets the same
// bits); the TC2 choice X^2 when |tr X^2 - nocc| < |2 tr X - tr X^2 - nocc|.
__device__ __forceinline__ bool tc2_choice(const double *__restrict__ pin, int nblk, double nocc, double &tr,
                                           double &sq)
{
    const int lane = threadIdx.x & 63;
    tr = 0.0;
    sq = 0.0;
    for (int b = lane; b < nblk; b += 64) {
        tr += pin[2 * b];
        sq += pin[2 * b + 1];
    }
    for (int d = 32; d > 0; d >>= 1) {
        tr += __shfl_xor(tr, d, 64);
        sq += __shfl_xor(sq, d, 64);
    }
    return fabs(sq - nocc) < fabs(2.0 * tr - sq - nocc);
}

// Above TC2_FUSED_MAX tiles the totals are formed once per step by this
// one-wave kernel instead of by every block of the step kernel (whose cost
// would grow as the square of the tile count).
constexpr int TC2_FUSED_MAX = 1024;
__global__ __launch_bounds__(64) void tc2_total_kernel(const double *__restrict__ pin, int nblk, double nocc,
                                                       double *__restrict__ st, int sidx)
{
    const double rem = st[TC2_REM + (sidx & 1)];
    if (rem == 0.0) {
        if (threadIdx.x == 0) st[TC2_REM + ((sidx + 1) & 1)] = 0.0;
        return;
    }
    double tr, sq;
    const bool square = tc2_choice(pin, nblk, nocc, tr, sq);
    if (threadIdx.x == 0) {
        st[0] = tr;
        st[1] = sq;
        st[2] = square ? 1.0 : 0.0;
        st[TC2_REM + ((sidx + 1) & 1)] = tc2_next_rem(rem, tr, sq, nocc);
    }
}

// One step X <- X^2 or 2X - X^2 for an exactly symmetric X (trace-correcting
// purification, Niklasson 2002).  The choice comes from tr X and tr X^2:
// summed from the previous step's tile partials `pin` by every block
// (tc2_choice; block 0 then stores tr X, tr X^2 and the choice in st[0..2]
// for the host's convergence check), or, with `presummed`, read from st
// where tc2_total_kernel put them.
// Y = X X on the FP64 matrix cores: the block's 32 x 32 tile of Y is four
// 16 x 16 sub-tiles; every wave computes all four over a quarter of k (wave w
// takes k in [w kc, (w + 1) kc)) with v_mfma_f64_16x16x4f64 (lane l supplies
// A(l % 16, k + l / 16) and B(k + l / 16, l % 16) and receives
// Y(l / 16 + 4q, l % 16), q = 0..3), its operands loaded TC2_KB k-steps
// ahead; as X is symmetric, B(k, j) = X(j, k) and both operands are read down
// columns.  Wave w then adds sub-tile w over the four waves in wave order and
// writes the new X to Xn (another buffer); tile partials go to pout.
__global__ __launch_bounds__(256) void tc2_step_kernel(const double *__restrict__ X, double *__restrict__ Xn, int n,
                                                      const double *__restrict__ pin, double *__restrict__ pout,
                                                      int nblk, double nocc, double *__restrict__ st, int presummed,
                                                      int sidx)
{
    const int nt0 = (n + TC2_T - 1) / TC2_T;
    const double rem = st[TC2_REM + (sidx & 1)];
    if (rem == 0.0) {
        // converged and finished: the step copies X (tile and partials)
        const int ti0 = (blockIdx.x % nt0) * TC2_T, tj0 = (blockIdx.x / nt0) * TC2_T;
        for (int e = threadIdx.x; e < TC2_T * TC2_T; e += 256) {
            const int i = ti0 + e % TC2_T, j = tj0 + e / TC2_T;
            if (i < n && j < n) Xn[(size_t)j * n + i] = X[(size_t)j * n + i];
        }
        if (threadIdx.x == 0) {
            pout[2 * blockIdx.x] = pin[2 * blockIdx.x];
            pout[2 * blockIdx.x + 1] = pin[2 * blockIdx.x + 1];
            if (!presummed && blockIdx.x == 0) st[TC2_REM + ((sidx + 1) & 1)] = 0.0;
        }
        return;
    }
    __shared__ double sh[8];
    __shared__ tc2_d4 sacc[4][4][64];   // [wave][sub-tile][lane]
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int nt = (n + TC2_T - 1) / TC2_T;
    const int i0 = (blockIdx.x % nt) * TC2_T, j0 = (blockIdx.x / nt) * TC2_T;
    const int r = lane & 15, kq = lane >> 4;
    // sub-tile w of this wave's epilogue: rows i0 + 16 (w & 1), columns
    // j0 + 16 (w >> 1); the old X at its four outputs per lane, loaded early
    const int ei0 = i0 + 16 * (wave & 1), ej = j0 + 16 * (wave >> 1) + r;
    double xold[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int i = ei0 + kq + 4 * q;
        xold[q] = (i < n && ej < n) ? X[(size_t)i * n + ej] : 0.0;
    }
    const bool rin0 = i0 + r < n, rin1 = i0 + 16 + r < n, cin0 = j0 + r < n, cin1 = j0 + 16 + r < n;
    const int kc = 4 * ((n + 15) / 16);   // k range of one wave, a multiple of 4
    const int kbeg = wave * kc, kend = kbeg + kc;
    tc2_d4 acc[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) acc[t] = tc2_d4{0.0, 0.0, 0.0, 0.0};
    for (int k0 = kbeg; k0 < kend; k0 += 4 * TC2_KB) {
        double a0[TC2_KB], a1[TC2_KB], b0[TC2_KB], b1[TC2_KB];
#pragma unroll
        for (int t = 0; t < TC2_KB; ++t) {
            const int kk = k0 + 4 * t + kq;
            const bool kin = kk < n && kk < kend;
            const double *col = X + (size_t)kk * n;
            a0[t] = (rin0 && kin) ? col[i0 + r] : 0.0;
            a1[t] = (rin1 && kin) ? col[i0 + 16 + r] : 0.0;
            b0[t] = (cin0 && kin) ? col[j0 + r] : 0.0;
            b1[t] = (cin1 && kin) ? col[j0 + 16 + r] : 0.0;
        }
#pragma unroll
        for (int t = 0; t < TC2_KB; ++t) {
            acc[0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0[t], b0[t], acc[0], 0, 0, 0);
            acc[1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1[t], b0[t], acc[1], 0, 0, 0);
            acc[2] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0[t], b1[t], acc[2], 0, 0, 0);
            acc[3] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1[t], b1[t], acc[3], 0, 0, 0);
        }
    }
#pragma unroll
    for (int t = 0; t < 4; ++t) sacc[wave][t][lane] = acc[t];
    // the choice for the old X (needed only from here on)
    bool square;
    if (presummed) {
        square = st[2] != 0.0;
    } else {
        double tr, sq;
        square = tc2_choice(pin, nblk, nocc, tr, sq);
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            st[0] = tr;
            st[1] = sq;
            st[2] = square ? 1.0 : 0.0;
            st[TC2_REM + ((sidx + 1) & 1)] = tc2_next_rem(rem, tr, sq, nocc);
        }
    }
    __syncthreads();
    // Y(i, j), i = ei0 + kq + 4q, j = ej, stored at (j, i), which holds the
    // same value (Y and X symmetric), so the 16 lanes of a row group write
    // one contiguous run
    const tc2_d4 y = ((sacc[0][wave][lane] + sacc[1][wave][lane]) + sacc[2][wave][lane]) + sacc[3][wave][lane];
    double ptr = 0.0, psq = 0.0;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int i = ei0 + kq + 4 * q;
        if (i < n && ej < n) {
            const size_t e = (size_t)i * n + ej;
            const double v = square ? y[q] : 2.0 * xold[q] - y[q];
            Xn[e] = v;
            psq += v * v;
            if (i == ej) ptr += v;
        }
    }
    block_sum2_256(ptr, psq, sh);
    if (threadIdx.x == 0) {
        pout[2 * blockIdx.x] = ptr;
        pout[2 * blockIdx.x + 1] = psq;
    }
}

// ---- Pulay DIIS (scf_accel 3) on the orthonormal-basis Fock matrices ----
// The commutator error e = F D S - S D F = A - A^T with A = F D S (F, D, S
// symmetric), column-major n x n.
// incremental G: dD = P - Pref, Pref = P (packed)
__global__ void delta_density_kernel(const double *__restrict__ P, double *__restrict__ Pref, double *__restrict__ dD,
                                     int64_t no2)
{
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= no2) return;
    const double p = P[k];
    dD[k] = p - Pref[k];
    Pref[k] = p;
}

// incremental G: Gacc += dG, dG = Gacc (packed)
__global__ void accumulate_g_kernel(double *__restrict__ Gacc, double *__restrict__ dG, int64_t no2)
{
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= no2) return;
    const double g = Gacc[k] + dG[k];
    Gacc[k] = g;
    dG[k] = g;
}

__global__ void antisym_kernel(const double *__restrict__ A, double *__restrict__ E, int n)
{
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (int64_t)n * n) return;
    const int64_t i = idx % n, j = idx / n;
    E[idx] = A[idx] - A[i * n + j];
}

// dot products of the error in slot k with the errors of slots 0..m-1
// (grid SUMS_BLOCKS x m): block b of slot j reduces one contiguous chunk in
// a fixed tree; diis_dot_final_kernel adds the partials in block order
__global__ __launch_bounds__(256) void diis_dot_part_kernel(const double *__restrict__ ebuf, int k, int64_t nn,
                                                           double *__restrict__ part)
{
    __shared__ double red[256];
    const int j = blockIdx.y;
    const double *__restrict__ a = ebuf + (size_t)k * nn;
    const double *__restrict__ b = ebuf + (size_t)j * nn;
    const int64_t chunk = (nn + gridDim.x - 1) / gridDim.x;
    const int64_t lo = (int64_t)blockIdx.x * chunk, hi = lo + chunk < nn ? lo + chunk : nn;
    double v = 0.0;
    for (int64_t i = lo + threadIdx.x; i < hi; i += 256) v += a[i] * b[i];
    red[threadIdx.x] = v;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[(size_t)j * gridDim.x + blockIdx.x] = red[0];
}

__global__ void diis_dot_final_kernel(const double *__restrict__ part, int nb, int m, double *__restrict__ out)
{
    const int j = threadIdx.x;
    if (j >= m) return;
    double v = 0.0;
    for (int b = 0; b < nb; ++b) v += part[(size_t)j * nb + b];
    out[j] = v;
}

// F = sum_j c_j F_j over the m stored Fock matrices
__global__ void diis_combine_kernel(const double *__restrict__ fbuf, const double *__restrict__ c, int m, int64_t nn,
                                    double *__restrict__ F)
{
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= nn) return;
    double v = 0.0;
    for (int j = 0; j < m; ++j)
        if (c[j] != 0.0) v += c[j] * fbuf[(size_t)j * nn + idx];   // slots not in use are skipped
    F[idx] = v;
}

struct DevMem {
    std::vector<void *> ptrs;
    ~DevMem()
    {
        for (void *p : ptrs) (void)hipFree(p);
    }
    template <typename T>
    T *get(size_t n)
    {
        void *p = nullptr;
        if (hipMalloc(&p, sizeof(T) * (n ? n : 1)) != hipSuccess) return nullptr;
        ptrs.push_back(p);
        return (T *)p;
    }
};

// Device state of one SCF run: X, the Fock/eigen work matrices and the
// rocBLAS handle on the run's stream.
struct Scf {
    int n = 0;
    size_t no2 = 0;
    hipStream_t s = nullptr;
    rocblas_handle h = nullptr;
    DevMem mem;
    double *X = nullptr, *F = nullptr, *T = nullptr, *C = nullptr, *ev = nullptr, *off = nullptr, *sums = nullptr;
    double *X2 = nullptr, *glo = nullptr, *ghi = nullptr, *tc = nullptr;
    double *part = nullptr, *spart = nullptr;   // tile partials of the purification, block partials of the sums
    int nblk = 0;                               // purification tiles
    // per channel (RHF: 0; UHF: 0 alpha, 1 beta): the last diagonalised F'
    // and its spectral bounds bnd[2 ch], bnd[2 ch + 1]
    double *Fprev[2] = {nullptr, nullptr}, *bnd = nullptr;
    bool have_prev[2] = {false, false};
    int last_steps[2] = {0, 0};   // steps the channel's last purification launched
    bool purify = false;   // density by purification inside the iteration loop
    int purify_fallbacks = 0;
    rocblas_int *info = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    const char *err = nullptr;

    bool owns = true;   // the stream and the rocBLAS handle are this run's own

    ~Scf()
    {
        if (owns && h) rocblas_destroy_handle(h);
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        if (owns && s) (void)hipStreamDestroy(s);
    }

    bool init(int n_)
    {
        if (hipStreamCreate(&s) != hipSuccess) return fail("stream creation failed");
        if (rocblas_create_handle(&h) != rocblas_status_success || rocblas_set_stream(h, s) != rocblas_status_success)
            return fail("rocblas handle creation failed");
        return init_buffers(n_);
    }

    // on a stream and rocBLAS handle shared with other runs (a batch)
    bool init_shared(int n_, hipStream_t st, rocblas_handle hd)
    {
        owns = false;
        s = st;
        h = hd;
        return init_buffers(n_);
    }

    bool init_buffers(int n_)
    {
        n = n_;
        no2 = (size_t)n * (n + 1) / 2;
        const size_t nn = (size_t)n * n;
        if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)
            return fail("event creation failed");
        X = mem.get<double>(nn);
        F = mem.get<double>(nn);
        T = mem.get<double>(nn);
        C = mem.get<double>(nn);
        ev = mem.get<double>(n);
        off = mem.get<double>(n);
        sums = mem.get<double>(8);
        info = mem.get<rocblas_int>(1);
        X2 = mem.get<double>(nn);
        glo = mem.get<double>(n);
        ghi = mem.get<double>(n);
        tc = mem.get<double>(8);
        const int nt = (n + TC2_T - 1) / TC2_T;
        nblk = nt * nt;
        part = mem.get<double>(4 * (size_t)nblk);
        spart = mem.get<double>(3 * SUMS_BLOCKS);
        Fprev[0] = mem.get<double>(nn);
        Fprev[1] = mem.get<double>(nn);
        bnd = mem.get<double>(4);
        if (!X || !F || !T || !C || !ev || !off || !sums || !info || !X2 || !glo || !ghi || !tc || !part || !spart ||
            !Fprev[0] || !Fprev[1] || !bnd)
            return fail("device allocation failed");
        return true;
    }

    bool fail(const char *m)
    {
        err = m;
        return false;
    }

    // eigenvectors of the symmetric F (overwritten, column-major V) with
    // the reference's sign convention; eigenvalues ascending into ev
    bool eigen()
    {
        if (rocsolver_dsyevd(h, rocblas_evect_original, rocblas_fill_lower, n, F, n, ev, off, info) !=
            rocblas_status_success)
            return fail("rocsolver_dsyevd failed");
        hipLaunchKernelGGL(phase_kernel, dim3(n), dim3(256), 0, s, F, n);
        rocblas_int inf = 0;
        if (hipMemcpyAsync(&inf, info, sizeof(inf), hipMemcpyDeviceToHost, s) != hipSuccess ||
            hipStreamSynchronize(s) != hipSuccess)
            return fail("device failure in the eigensolver");
        if (inf != 0) return fail("eigensolver did not converge");
        return true;
    }

    // X = U s^-1/2 from packed S (rhfscf.c:432-452)
    bool form_x(const double *dS)
    {
        const int64_t nn = (int64_t)n * n;
        hipLaunchKernelGGL(sym_from_packed_kernel, dim3((nn + 255) / 256), dim3(256), 0, s, dS, nullptr, F, n);
        if (!eigen()) return false;
        std::vector<double> sv(n);
        if (hipMemcpyAsync(sv.data(), ev, sizeof(double) * n, hipMemcpyDeviceToHost, s) != hipSuccess ||
            hipStreamSynchronize(s) != hipSuccess)
            return fail("device failure in the orthogonaliser");
        for (int i = 0; i < n; ++i)
            if (sv[i] <= DBL_EPSILON) return fail("Negative or Zero Eigenvalues in Overlap Matrix!");
        hipLaunchKernelGGL(scale_columns_kernel, dim3((nn + 255) / 256), dim3(256), 0, s, X, F, ev, n);
        return hipGetLastError() == hipSuccess || fail("launch failure");
    }

    // C = X V with V the eigenvectors of X^T (a + b) X; P = C_occ C_occ^T
    // packed (rhfscf.c:37-46).  Leaves the eigenvalues in ev.
    // F' = X^T (a + b) X into F
    bool fprime(const double *a, const double *b)
    {
        const int64_t nn = (int64_t)n * n;
        const double one = 1.0, zero = 0.0;
        hipLaunchKernelGGL(sym_from_packed_kernel, dim3((nn + 255) / 256), dim3(256), 0, s, a, b, F, n);
        if (rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_none, n, n, n, &one, F, n, X, n, &zero, T, n) !=
                rocblas_status_success ||
            rocblas_dgemm(h, rocblas_operation_transpose, rocblas_operation_none, n, n, n, &one, X, n, T, n, &zero, F,
                          n) != rocblas_status_success)
            return fail("rocblas_dgemm failed");
        return true;
    }

    // P = X P' X^T packed, with P' the projector onto the nocc lowest
    // eigenvectors of F' by trace-correcting purification: from X0 =
    // (emax - F')/(emax - emin) (Gershgorin bounds), X <- X^2 or 2X - X^2,
    // whichever brings tr X^2 / tr(2X - X^2) nearer nocc, until
    // tr X - tr X^2 (the idempotency error, sum of l(1-l) over the
    // eigenvalues) is below 1e-12, then two more steps.  Returns false with
    // `converged` unset when 120 steps do not get there (a degenerate or
    // vanishing gap at the Fermi level); the caller then diagonalises.
    bool purify_density(int nocc, double *P, bool &converged, int ch)
    {
        const int64_t nn = (int64_t)n * n;
        const double one = 1.0, zero = 0.0;
        converged = false;
        if (nocc <= 0 || nocc >= n) return true;   // the caller diagonalises these
        hipLaunchKernelGGL(gersh_row_kernel, dim3(n), dim3(256), 0, s, F, n, glo, ghi);
        hipLaunchKernelGGL(gersh_reduce_kernel, dim3(1), dim3(256), 0, s, glo, ghi, n, tc + 4);
        if (have_prev[ch]) {
            hipLaunchKernelGGL(diff2_part_kernel, dim3(SUMS_BLOCKS), dim3(256), 0, s, F, Fprev[ch], nn, spart);
            hipLaunchKernelGGL(diffcol_kernel, dim3(n), dim3(256), 0, s, F, Fprev[ch], n, glo);
            hipLaunchKernelGGL(tc2_bounds_kernel, dim3(1), dim3(64), 0, s, spart, SUMS_BLOCKS, glo, n, bnd + 2 * ch,
                               tc + 4);
        }
        // X0 into cur with its tile partials; each step writes the next X
        // into the other buffer
        double *cur = T, *nxt = X2, *pin = part, *pout = part + 2 * nblk;
        hipLaunchKernelGGL(tc2_init_kernel, dim3((nn + 255) / 256), dim3(256), 0, s, F, cur, n, tc + 4);
        hipLaunchKernelGGL(tc2_partials_kernel, dim3(nblk), dim3(256), 0, s, cur, n, pin, tc);
        // steps in batches (the first as long as the channel's last
        // purification less 4, or 32; then 4 at a time; at most 120): the
        // device tracks convergence (TC2_REM) and turns the steps after it
        // into copies; the host reads the state once per batch
        const int presummed = nblk > TC2_FUSED_MAX || g_scf_density == 2;
        const int first = last_steps[ch] > 0 ? std::max(8, last_steps[ch] - 4) : 32;
        int sidx = 0;
        while (sidx < 120) {
            const int batch = std::min(sidx == 0 ? first : 4, 120 - sidx);
            for (int b = 0; b < batch; ++b, ++sidx) {
                if (presummed)
                    hipLaunchKernelGGL(tc2_total_kernel, dim3(1), dim3(64), 0, s, pin, nblk, (double)nocc, tc, sidx);
                hipLaunchKernelGGL(tc2_step_kernel, dim3(nblk), dim3(256), 0, s, cur, nxt, n, pin, pout, nblk,
                                   (double)nocc, tc, presummed, sidx);
                std::swap(cur, nxt);
                std::swap(pin, pout);
            }
            double rem = -1.0;
            if (hipMemcpyAsync(&rem, tc + TC2_REM + (sidx & 1), sizeof(rem), hipMemcpyDeviceToHost, s) != hipSuccess ||
                hipStreamSynchronize(s) != hipSuccess)
                return fail("device failure in the purification");
            if (rem == 0.0) {
                converged = true;
                last_steps[ch] = sidx;
                break;
            }
        }
        if (!converged) return true;
        // P = X P' X^T (P' in cur, nxt as scratch), then pack
        if (rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_none, n, n, n, &one, X, n, cur, n, &zero, nxt,
                          n) != rocblas_status_success ||
            rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_transpose, n, n, n, &one, nxt, n, X, n, &zero, cur,
                          n) != rocblas_status_success)
            return fail("rocblas_dgemm failed");
        hipLaunchKernelGGL(pack_lower_kernel, dim3(n), dim3(256), 0, s, cur, P, n);
        return hipGetLastError() == hipSuccess || fail("launch failure");
    }

    // the density of (a + b): purification inside the loop when enabled,
    // otherwise (and as its fallback) the eigensolver path
    bool density(const double *a, const double *b, int nocc, double *P, int ch)
    {
        if (!purify || nocc <= 0 || nocc >= n) return diag_density(a, b, nocc, P, ch);
        if (!fprime(a, b)) return false;
        bool conv = false;
        if (!purify_density(nocc, P, conv, ch)) return false;
        if (conv) return true;
        ++purify_fallbacks;
        return diag_density(a, b, nocc, P, ch);
    }

    // orbitals of (a + b): eigenvalues into ev, C = X V
    bool orbitals(const double *a, const double *b)
    {
        const double one = 1.0, zero = 0.0;
        if (!fprime(a, b) || !eigen()) return false;
        if (rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_none, n, n, n, &one, X, n, F, n, &zero, C, n) !=
            rocblas_status_success)
            return fail("rocblas_dgemm failed");
        return true;
    }

    bool diag_density(const double *a, const double *b, int nocc, double *P, int ch)
    {
        return fprime(a, b) && diag_from_fprime(nocc, P, ch);
    }

    // the density of F' (in F) by the eigensolver
    bool diag_from_fprime(int nocc, double *P, int ch)
    {
        const int64_t nn = (int64_t)n * n;
        const double one = 1.0, zero = 0.0;
        // F' (overwritten by the eigenvectors) and its exact spectral bounds
        // become the channel's reference for the bounds of later
        // purifications
        if (hipMemcpyAsync(Fprev[ch], F, sizeof(double) * nn, hipMemcpyDeviceToDevice, s) != hipSuccess)
            return fail("copy failed");
        if (!eigen()) return false;
        hipLaunchKernelGGL(tc2_exact_bounds_kernel, dim3(1), dim3(64), 0, s, ev, n, bnd + 2 * ch);
        have_prev[ch] = true;
        // C = X V, D = C_occ C_occ^T (lower), P = pack(D)
        if (rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_none, n, n, n, &one, X, n, F, n, &zero, C, n) !=
            rocblas_status_success)
            return fail("rocblas_dgemm failed");
        if (nocc > 0) {
            if (rocblas_dsyrk(h, rocblas_fill_lower, rocblas_operation_none, n, nocc, &one, C, n, &zero, T, n) !=
                rocblas_status_success)
                return fail("rocblas_dsyrk failed");
        } else {
            if (hipMemsetAsync(T, 0, sizeof(double) * nn, s) != hipSuccess) return fail("memset failed");
        }
        hipLaunchKernelGGL(pack_lower_kernel, dim3(n), dim3(256), 0, s, T, P, n);
        return hipGetLastError() == hipSuccess || fail("launch failure");
    }

    // sums over (P, P1, H, G) -> host
    bool sums3(const double *P, const double *P1, const double *H, const double *G, double *out)
    {
        hipLaunchKernelGGL(scf_sums_part_kernel, dim3(SUMS_BLOCKS), dim3(256), 0, s, P, P1, H, G, (int64_t)no2, spart);
        hipLaunchKernelGGL(scf_sums_final_kernel, dim3(1), dim3(64), 0, s, spart, SUMS_BLOCKS, sums);
        if (hipMemcpyAsync(out, sums, sizeof(double) * 3, hipMemcpyDeviceToHost, s) != hipSuccess ||
            hipStreamSynchronize(s) != hipSuccess)
            return fail("device failure in the energy sums");
        return true;
    }

    // scfAccelerator (scfaux.c:64-95) on device densities
    bool accelerate(int iter, double *p0, double *p1, double *p2)
    {
        double beta = 0.5;
        if (iter > 2) {
            double d[3];
            hipLaunchKernelGGL(accel_dots_kernel, dim3(1), dim3(RED_THREADS), 0, s, p0, p1, p2, (int64_t)no2, sums);
            if (hipMemcpyAsync(d, sums, sizeof(double) * 3, hipMemcpyDeviceToHost, s) != hipSuccess ||
                hipStreamSynchronize(s) != hipSuccess)
                return fail("device failure in the accelerator");
            beta = (d[0] - d[1]) / (d[0] - 2.0 * d[1] + d[2]);
            if (beta < 1e-3) beta = 1e-3;
            if (beta > 1.5) beta = 1.5;
        }
        hipLaunchKernelGGL(accel_update_kernel, dim3((no2 + 255) / 256), dim3(256), 0, s, p0, p1, p2, (int64_t)no2,
                           beta);
        return hipGetLastError() == hipSuccess || fail("launch failure");
    }

    void average(double *p, const double *p1, const double *p2)
    {
        hipLaunchKernelGGL(average_kernel, dim3((no2 + 255) / 256), dim3(256), 0, s, p, p1, p2, (int64_t)no2);
    }

    // ---- Pulay DIIS (scf_accel 3) ----
    // The last DIIS_M orthonormal-basis Fock matrices F' = X^T F X and
    // commutator errors e' = X^T (F D S - S D F) X of the iteration's input
    // density D; the Fock matrix the density is taken from is the
    // combination sum c_j F'_j with sum c_j = 1 that minimises |sum c_j e'_j|
    // (Pulay, Chem. Phys. Lett. 73, 393 (1980); J. Comput. Chem. 3, 556
    // (1982)).  The error dot products are fixed-order device sums, the small
    // linear system is solved on the host, so the sequence is reproducible.
    // The first iteration's Fock matrix (of the core-Hamiltonian guess) stays
    // out of the subspace: with it in, extrapolation from the guess can lock
    // onto an excited stationary state (seen on a C/N/O test molecule of 50
    // functions, 0.28 Eh above the ground state that plain iteration reaches).
    static constexpr int DIIS_M = 8;
    double *Sfull = nullptr, *Dfull = nullptr, *Fao = nullptr, *W = nullptr;
    double *dfbuf = nullptr, *debuf = nullptr, *dpart = nullptr, *dout = nullptr, *dcoef = nullptr;
    int diis_cnt = 0, diis_next = 0;
    double Bm[DIIS_M][DIIS_M] = {};
    double diis_err = 0.0;   // Frobenius norm of the newest error

    bool diis_init(const double *dS)
    {
        const size_t nn = (size_t)n * n;
        Sfull = mem.get<double>(nn);
        Dfull = mem.get<double>(nn);
        Fao = mem.get<double>(nn);
        W = mem.get<double>(nn);
        dfbuf = mem.get<double>(DIIS_M * nn);
        debuf = mem.get<double>(DIIS_M * nn);
        dpart = mem.get<double>((size_t)DIIS_M * SUMS_BLOCKS);
        dout = mem.get<double>(DIIS_M);
        dcoef = mem.get<double>(DIIS_M);
        if (!Sfull || !Dfull || !Fao || !W || !dfbuf || !debuf || !dpart || !dout || !dcoef)
            return fail("device allocation failed");
        hipLaunchKernelGGL(sym_from_packed_kernel, dim3((nn + 255) / 256), dim3(256), 0, s, dS, nullptr, Sfull, n);
        // the ring starts empty: no slot may hold stale values
        if (hipMemsetAsync(dfbuf, 0, sizeof(double) * DIIS_M * nn, s) != hipSuccess ||
            hipMemsetAsync(debuf, 0, sizeof(double) * DIIS_M * nn, s) != hipSuccess)
            return fail("memset failed");
        diis_cnt = diis_next = 0;
        return hipGetLastError() == hipSuccess || fail("launch failure");
    }

    // F' of (a + b) extrapolated by DIIS over the input density Pin, into F
    bool diis_fprime(const double *a, const double *b, const double *Pin)
    {
        const size_t nn = (size_t)n * n;
        const double one = 1.0, zero = 0.0;
        const dim3 g((unsigned)((nn + 255) / 256));
        hipLaunchKernelGGL(sym_from_packed_kernel, g, dim3(256), 0, s, a, b, Fao, n);
        hipLaunchKernelGGL(sym_from_packed_kernel, g, dim3(256), 0, s, Pin, nullptr, Dfull, n);
        const int k = diis_next;
        double *Ek = debuf + (size_t)k * nn, *Fk = dfbuf + (size_t)k * nn;
        auto gemm = [&](rocblas_operation ta, const double *A, const double *B, double *Cm) {
            return rocblas_dgemm(h, ta, rocblas_operation_none, n, n, n, &one, A, n, B, n, &zero, Cm, n) ==
                   rocblas_status_success;
        };
        // e' = X^T (A - A^T) X with A = F D S; F' = X^T F X
        if (!gemm(rocblas_operation_none, Fao, Dfull, W) || !gemm(rocblas_operation_none, W, Sfull, T))
            return fail("rocblas_dgemm failed");
        hipLaunchKernelGGL(antisym_kernel, g, dim3(256), 0, s, T, W, n);
        if (!gemm(rocblas_operation_transpose, X, W, T) || !gemm(rocblas_operation_none, T, X, Ek) ||
            !gemm(rocblas_operation_none, Fao, X, T) || !gemm(rocblas_operation_transpose, X, T, Fk))
            return fail("rocblas_dgemm failed");
        diis_next = (diis_next + 1) % DIIS_M;
        diis_cnt = std::min(diis_cnt + 1, DIIS_M);
        const int m = diis_cnt;
        hipLaunchKernelGGL(diis_dot_part_kernel, dim3(SUMS_BLOCKS, m), dim3(256), 0, s, debuf, k, (int64_t)nn, dpart);
        hipLaunchKernelGGL(diis_dot_final_kernel, dim3(1), dim3(64), 0, s, dpart, SUMS_BLOCKS, m, dout);
        double d[DIIS_M];
        if (hipMemcpyAsync(d, dout, sizeof(double) * m, hipMemcpyDeviceToHost, s) != hipSuccess ||
            hipStreamSynchronize(s) != hipSuccess)
            return fail("device failure in DIIS");
        for (int j = 0; j < m; ++j) Bm[k][j] = Bm[j][k] = d[j];
        diis_err = std::sqrt(std::max(0.0, d[k]));
        // minimise c^T B c subject to sum c = 1 over the stored slots (in slot
        // order), dropping the oldest slots while the system is singular
        std::vector<int> use;
        for (int j = 0; j < m; ++j) use.push_back((k - (m - 1) + j + DIIS_M) % DIIS_M);   // oldest first
        std::vector<double> c;
        while (!use.empty()) {
            std::vector<int> ord(use);
            std::sort(ord.begin(), ord.end());
            const int q = (int)ord.size();
            std::vector<double> A((size_t)(q + 1) * (q + 1), 0.0), r(q + 1, 0.0);
            double scale = 0.0;
            for (int i = 0; i < q; ++i) scale = std::max(scale, Bm[ord[i]][ord[i]]);
            if (scale <= 0.0) scale = 1.0;
            for (int i = 0; i < q; ++i) {
                for (int j = 0; j < q; ++j) A[(size_t)i * (q + 1) + j] = Bm[ord[i]][ord[j]] / scale;
                A[(size_t)i * (q + 1) + q] = -1.0;
                A[(size_t)q * (q + 1) + i] = -1.0;
            }
            r[q] = -1.0;
            if (solve(q + 1, A, r)) {
                c.assign(DIIS_M, 0.0);
                for (int i = 0; i < q; ++i) c[ord[i]] = r[i];
                break;
            }
            use.erase(use.begin());   // drop the oldest
        }
        if (c.empty()) {   // cannot happen for a nonzero newest error; take F' itself
            c.assign(DIIS_M, 0.0);
            c[k] = 1.0;
        }
        if (hipMemcpyAsync(dcoef, c.data(), sizeof(double) * DIIS_M, hipMemcpyHostToDevice, s) != hipSuccess)
            return fail("upload failed");
        hipLaunchKernelGGL(diis_combine_kernel, g, dim3(256), 0, s, dfbuf, dcoef, DIIS_M, (int64_t)nn, F);
        // the host vector c must outlive the copy
        if (hipStreamSynchronize(s) != hipSuccess) return fail("device failure in DIIS");
        return hipGetLastError() == hipSuccess || fail("launch failure");
    }

    // Gaussian elimination with partial pivoting of the dense q x q system
    // A x = r (x into r); false when a pivot vanishes
    static bool solve(int q, std::vector<double> &A, std::vector<double> &r)
    {
        for (int c = 0; c < q; ++c) {
            int p = c;
            for (int i = c + 1; i < q; ++i)
                if (std::fabs(A[(size_t)i * q + c]) > std::fabs(A[(size_t)p * q + c])) p = i;
            if (std::fabs(A[(size_t)p * q + c]) < 1e-14) return false;
            if (p != c) {
                for (int j = 0; j < q; ++j) std::swap(A[(size_t)p * q + j], A[(size_t)c * q + j]);
                std::swap(r[p], r[c]);
            }
            for (int i = c + 1; i < q; ++i) {
                const double f = A[(size_t)i * q + c] / A[(size_t)c * q + c];
                for (int j = c; j < q; ++j) A[(size_t)i * q + j] -= f * A[(size_t)c * q + j];
                r[i] -= f * r[c];
            }
        }
        for (int i = q - 1; i >= 0; --i) {
            double v = r[i];
            for (int j = i + 1; j < q; ++j) v -= A[(size_t)i * q + j] * r[j];
            r[i] = v / A[(size_t)i * q + i];
        }
        return true;
    }

    // the density of the DIIS-extrapolated Fock matrix of (a + b)
    bool diis_density(const double *a, const double *b, const double *Pin, int nocc, double *P, int ch)
    {
        if (!diis_fprime(a, b, Pin)) return false;
        if (purify && nocc > 0 && nocc < n) {
            bool conv = false;
            if (!purify_density(nocc, P, conv, ch)) return false;
            if (conv) return true;
            ++purify_fallbacks;
        }
        return diag_from_fprime(nocc, P, ch);
    }

    float elapsed_ms()
    {
        float ms = 0.0f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        return ms;
    }
};

bool scf_log()
{
    const char *e = getenv("PATMOL_HIP_SCF_LOG");
    return e && *e && *e != '0';
}

// copy a column-major device matrix to a row-major host one
bool download_rowmajor(const Scf &c, const double *dM, double *host)
{
    std::vector<double> tmp((size_t)c.n * c.n);
    if (hipMemcpyAsync(tmp.data(), dM, sizeof(double) * tmp.size(), hipMemcpyDeviceToHost, c.s) != hipSuccess ||
        hipStreamSynchronize(c.s) != hipSuccess)
        return false;
    for (int i = 0; i < c.n; ++i)
        for (int j = 0; j < c.n; ++j) host[(size_t)i * c.n + j] = tmp[(size_t)j * c.n + i];
    return true;
}

int scf_fail(Scf &c, int code, const char *where)
{
    char msg[256];
    snprintf(msg, sizeof msg, "%s: %s", where, c.err ? c.err : "failure");
    return aux_fail(code, msg);
}

// The last iterations of a mixed-precision SCF build G in FP64 (no
// counterpart in the reference).  The FP32 items' rounding moves the density
// by a small amount that does not shrink as the SCF converges: on the
// 300-function C6H6 ring at threshold 1e-4 the density change pdiff stalls
// at 1-2e-9, while the FP64 SCF falls through the stop bound (2e-10 there)
// in 32 iterations, so the stop rule pdiff < eps of rhfscf.c:52-59 might
// never hold.  Once pdiff is within MIXED_TAIL_FACTOR of eps, or stops
// falling below 1e-6, the precision mode is set to FP64 for the rest of the
// SCF; the caller's mode is restored when the SCF returns.
constexpr double MIXED_TAIL_FACTOR = 20.0;
struct PrecisionTail {
    const int mode0 = precision_mode();
    bool on = false;
    double prev = HUGE_VAL;
    PrecisionTail() = default;
    PrecisionTail(const PrecisionTail &) = delete;
    PrecisionTail &operator=(const PrecisionTail &) = delete;
    ~PrecisionTail() { set_precision_mode(mode0); }
    // after iteration `it` (>= 1) with density change pdiff; true when this
    // call switched the builds to FP64
    bool step(int it, double pdiff, double eps)
    {
        bool now = false;
        if (mode0 == 1 && !on && it > 1 && (pdiff < MIXED_TAIL_FACTOR * eps || (pdiff < 1e-6 && pdiff >= prev))) {
            on = true;
            now = true;
            set_precision_mode(0);
        }
        prev = pdiff;
        return now;
    }
};

}  // namespace
}  // namespace pm

using namespace pm;

extern "C" int patmol_hip_rhf_scf(const ERIFS *mints, const ERIFS *xints, const double *Smat, const double *Hmat,
                                  int nocc, int maxits, double scf_eps, int scf_accel, double enuc,
                                  const double *Pguess, double *Pout, double *evals, double *Cmat,
                                  PatmolHipScfResult *res)
{
    const int n = number_of_orbitals();
    if (!mints || !Smat || !Hmat || !res || n <= 0 || nocc < 0 || nocc > n)
        return aux_fail(PATMOL_HIP_EBADARG, "patmol_hip_rhf_scf: bad argument");
    const double EPS10 = 10 * DBL_EPSILON;
    Scf c;
    if (!c.init(n)) return scf_fail(c, PATMOL_HIP_EDEVICE, "patmol_hip_rhf_scf");
    const size_t no2 = c.no2;
    double *dS = c.mem.get<double>(no2), *dH = c.mem.get<double>(no2), *dG = c.mem.get<double>(no2);
    double *Pb[3] = {c.mem.get<double>(no2), c.mem.get<double>(no2), c.mem.get<double>(no2)};
    if (!dS || !dH || !dG || !Pb[0] || !Pb[1] || !Pb[2])
        return aux_fail(PATMOL_HIP_EDEVICE, "patmol_hip_rhf_scf: device allocation failed");
    double *P = Pb[0], *P1 = Pb[1], *P2 = Pb[2];
    if (hipMemcpyAsync(dS, Smat, sizeof(double) * no2, hipMemcpyHostToDevice, c.s) != hipSuccess ||
        hipMemcpyAsync(dH, Hmat, sizeof(double) * no2, hipMemcpyHostToDevice, c.s) != hipSuccess ||
        hipMemsetAsync(P1, 0, sizeof(double) * no2, c.s) != hipSuccess ||
        hipMemsetAsync(P2, 0, sizeof(double) * no2, c.s) != hipSuccess)
        return aux_fail(PATMOL_HIP_EDEVICE, "patmol_hip_rhf_scf: upload failed");
    if (!c.form_x(dS)) return scf_fail(c, PATMOL_HIP_EDEVICE, "patmol_hip_rhf_scf (orthogonaliser)");
    const bool diis = scf_accel == 3;
    if (diis && !c.diis_init(dS)) return scf_fail(c, PATMOL_HIP_EDEVICE, "patmol_hip_rhf_scf (DIIS set-up)");
    if (Pguess) {
        if (hipMemcpyAsync(P, Pguess, sizeof(double) * no2, hipMemcpyHostToDevice, c.s) != hipSuccess)
            return aux_fail(PATMOL_HIP_EDEVICE, "patmol_hip_rhf_scf: upload failed");
    } else if (!c.diag_density(dH, nullptr, nocc, P, 0)) {
        return scf_fail(c, PATMOL_HIP_EDEVICE, "patmol_hip_rhf_scf (guess)");
    }
    double eps = scf_eps;
    if (eps < EPS10 * n * n) eps = EPS10 * n * n;
    // below ~128 functions the eigensolver is faster than the ~40 GEMMs of
    // the purification (their launches and trace checks dominate)
    c.purify = g_scf_density >= 1 && n >= 128;
    // Incremental G (no counterpart in the reference): G is linear in P, so
    // G(P) = G(Pref) + G(P - Pref); between full builds only the difference
    // is built, with density-weighted screening and, in mixed precision, a
    // precision split by max|P - Pref|, which falls as the SCF converges.
    const int incr = g_scf_incremental;
    double *Gacc = nullptr, *Pref = nullptr, *dD = nullptr;
    if (incr > 0) {
        Gacc = c.mem.get<double>(no2);
        Pref = c.mem.get<double>(no2);
        dD = c.mem.get<double>(no2);
        if (!Gacc || !Pref || !dD) return aux_fail(PATMOL_HIP_EDEVICE, "patmol_hip_rhf_scf: device allocation failed");
    }
    const dim3 gp((unsigned)((no2 + 255) / 256));
    int since_full = 0;
    BuildCounts tally;
    int full_builds = 0, delta_builds = 0;
    auto count = [&](const ERIFS *e) {
        BuildCounts b;
        last_build_counts(e, &b);
        tally.items += b.items;
        tally.fp32 += b.fp32;
        tally.skipped += b.skipped;
    };
    const bool log = scf_log();
    PrecisionTail tail;
    double eold = 0.0, enrg = 0.0, ediff = 0.0, pdiff = 0.0, gms = 0.0, dms = 0.0;
    int iteration = 0, done = 0;
    bool first = true;
    auto t_start = std::chrono::steady_clock::now();
    for (;;) {
        ++iteration;
        (void)hipEventRecord(c.e0, c.s);
        if (hipMemsetAsync(dG, 0, sizeof(double) * no2, c.s) != hipSuccess)
            return aux_fail(PATMOL_HIP_EDEVICE, "patmol_hip_rhf_scf: memset failed");
        if (incr <= 0 || since_full == 0) {
            if (patmol_hip_rhf_gmatrix_device(mints, P, dG, c.s)) return patmol_hip_last_error();
            count(mints);
            if (xints) {
                if (patmol_hip_rhf_gmatrix_device(xints, P, dG, c.s)) return patmol_hip_last_error();
                count(xints);
            }
            if (incr > 0 && (hipMemcpyAsync(Gacc, dG, sizeof(double) * no2, hipMemcpyDeviceToDevice, c.s) != hipSuccess ||
                             hipMemcpyAsync(Pref, P, sizeof(double) * no2, hipMemcpyDeviceToDevice, c.s) != hipSuccess))
                return aux_fail(PATMOL_HIP_EDEVICE, "patmol_hip_rhf_scf: copy failed");
            ++full_builds;
        } else {
            hipLaunchKernelGGL(delta_density_kernel, gp, dim3(256), 0, c.s, P, Pref, dD, (int64_t)no2);
            BuildCounts b;
            if (rhf_gmatrix_delta(mints, dD, dG, c.s, &b)) return patmol_hip_last_error();
            count(mints);
            if (xints) {
                if (rhf_gmatrix_delta(xints, dD, dG, c.s, &b)) return patmol_hip_last_error();
                count(xints);
            }
            hipLaunchKernelGGL(accumulate_g_kernel, gp, dim3(256), 0, c.s, Gacc, dG, (int64_t)no2);
            if (hipGetLastError() != hipSuccess) return aux_fail(PATMOL_HIP_EDEVICE, "patmol_hip_rhf_scf: launch failed");
            ++delta_builds;
        }
        if (incr > 0) since_full = (since_full + 1) % incr;
        (void)hipEventRecord(c.e1, c.s);
        // P2 <- P1 <- P, new P into the freed buffer (rhfscf.c:43-46)
        double *fresh = P2;
        P2 = P1;
        P1 = P;
        P = fresh;
        // P1 is the density this iteration's G was built from
        const bool dens_ok = diis && iteration > 1 ? c.diis_density(dH, dG, P1, nocc, P, 0) : c.density(dH, dG, nocc, P, 0);
        if (!dens_ok) return scf_fail(c, PATMOL_HIP_EDEVICE, "patmol_hip_rhf_scf (density step)");
        double sm[3];
        if (!c.sums3(P, P1, dH, dG, sm)) return scf_fail(c, PATMOL_HIP_EDEVICE, "patmol_hip_rhf_scf");
        const float g_ms = c.elapsed_ms();
        gms += g_ms;
        pdiff = sqrt(2.0 * sm[2]) / n;
        enrg = 2.0 * (sm[0] + sm[0] + sm[1]);
        ediff = enrg - eold;
        eold = enrg;
        int ret = 0;
        if (iteration == 1) {
        } else if (fabs(ediff) < EPS10 || pdiff < eps) {
            ret = 1;
        } else if (ediff > 0.0) {
            if (scf_accel == 1) c.average(P, P1, P2);
            if (scf_accel == 2 && !c.accelerate(iteration, P, P1, P2))
                return scf_fail(c, PATMOL_HIP_EDEVICE, "patmol_hip_rhf_scf");
        }
        if (first) {
            first = false;
            iteration = 0;
            res->energy0 = enrg + enuc;
        } else {
            done = ret;
        }
        if (log)
            fprintf(stderr, "ITERATION = %12d EDIFF= %lg PDIFF= %lg\n  energy = %.12f\n", iteration, ediff, pdiff,
                    enrg + enuc);
        if (iteration == 0) continue;
        if (done || iteration >= maxits) break;
        if (tail.step(iteration, pdiff, eps)) {
            // an incremental SCF also rebuilds its reference G in FP64: the
            // mixed one would carry its FP32 rounding to the end
            if (incr > 0) since_full = 0;
            if (log) fprintf(stderr, "FP64 G builds from iteration %d\n", iteration + 1);
        }
    }
    if (hipStreamSynchronize(c.s) != hipSuccess) return aux_fail(PATMOL_HIP_EDEVICE, "patmol_hip_rhf_scf: device failure");
    const double tot = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_start).count();
    dms = tot - gms;
    res->energy = enrg + enuc;
    res->eelec = enrg;
    res->enuc = enuc;
    res->ediff = ediff;
    res->pdiff = pdiff;
    res->iterations = iteration;
    res->converged = done;
    res->gbuild_ms = gms;
    res->other_ms = dms;
    res->total_ms = tot;
    res->purify_fallbacks = c.purify_fallbacks;
    res->full_builds = full_builds;
    res->delta_builds = delta_builds;
    res->items = (double)tally.items;
    res->items_fp32 = (double)tally.fp32;
    res->items_skipped = (double)tally.skipped;
    // orbitals of the last Fock matrix (the loop's densities came from
    // purification)
    if (c.purify && (evals || Cmat) && !c.orbitals(dH, dG)) return scf_fail(c, PATMOL_HIP_EDEVICE, "patmol_hip_rhf_scf");
    bool ok = true;
    if (Pout) ok = ok && hipMemcpy(Pout, P, sizeof(double) * no2, hipMemcpyDeviceToHost) == hipSuccess;
    if (evals) ok = ok && hipMemcpy(evals, c.ev, sizeof(double) * n, hipMemcpyDeviceToHost) == hipSuccess;
    if (Cmat) ok = ok && download_rowmajor(c, c.C, Cmat);
    if (!ok) return aux_fail(PATMOL_HIP_EDEVICE, "patmol_hip_rhf_scf: download failed");
    return 0;
}

extern "C" int patmol_hip_uhf_scf(const ERIFS *mints, const ERIFS *xints, const double *Smat, const double *Hmat,
                                  int nocca, int noccb, int maxits, double scf_eps, int scf_accel, double enuc,
                                  const double *PAguess, const double *PBguess, double *PAout, double *PBout,
                                  double *evalsA, double *evalsB, PatmolHipScfResult *res)
{
    const int n = number_of_orbitals();
    if (!mints || !Smat || !Hmat || !res || n <= 0 || nocca < 0 || noccb < 0 || nocca > n || noccb > n)
        return aux_fail(PATMOL_HIP_EBADARG, "patmol_hip_uhf_scf: bad argument");
    const double EPS10 = 10 * DBL_EPSILON;
    Scf c;
    if (!c.init(n)) return scf_fail(c, PATMOL_HIP_EDEVICE, "patmol_hip_uhf_scf");
    const size_t no2 = c.no2;
    double *dS = c.mem.get<double>(no2), *dH = c.mem.get<double>(no2);
    double *GA = c.mem.get<double>(no2), *GB = c.mem.get<double>(no2);
    double *A[3], *B[3];
    for (int k = 0; k < 3; ++k) {
        A[k] = c.mem.get<double>(no2);
        B[k] = c.mem.get<double>(no2);
    }
    double *evA = c.mem.get<double>(n);
    if (!dS || !dH || !GA || !GB || !A[0] || !A[1] || !A[2] || !B[0] || !B[1] || !B[2] || !evA)
        return aux_fail(PATMOL_HIP_EDEVICE, "patmol_hip_uhf_scf: device allocation failed");
    double *PA = A[0], *PA1 = A[1], *PA2 = A[2], *PB = B[0], *PB1 = B[1], *PB2 = B[2];
    if (hipMemcpyAsync(dS, Smat, sizeof(double) * no2, hipMemcpyHostToDevice, c.s) != hipSuccess ||
        hipMemcpyAsync(dH, Hmat, sizeof(double) * no2, hipMemcpyHostToDevice, c.s) != hipSuccess ||
        hipMemsetAsync(PA1, 0, sizeof(double) * no2, c.s) != hipSuccess ||
        hipMemsetAsync(PA2, 0, sizeof(double) * no2, c.s) != hipSuccess ||
        hipMemsetAsync(PB1, 0, sizeof(double) * no2, c.s) != hipSuccess ||
        hipMemsetAsync(PB2, 0, sizeof(double) * no2, c.s) != hipSuccess)
        return aux_fail(PATMOL_HIP_EDEVICE, "patmol_hip_uhf_scf: upload failed");
    if (scf_accel == 3) return aux_fail(PATMOL_HIP_EBADARG, "patmol_hip_uhf_scf: DIIS (scf_accel 3) is RHF only");
    if (!c.form_x(dS)) return scf_fail(c, PATMOL_HIP_EDEVICE, "patmol_hip_uhf_scf");
    // core-Hamiltonian guess: both spins from the same orbitals (uhfscf.c:505-527)
    if (PAguess && PBguess) {
        if (hipMemcpyAsync(PA, PAguess, sizeof(double) * no2, hipMemcpyHostToDevice, c.s) != hipSuccess ||
            hipMemcpyAsync(PB, PBguess, sizeof(double) * no2, hipMemcpyHostToDevice, c.s) != hipSuccess)
            return aux_fail(PATMOL_HIP_EDEVICE, "patmol_hip_uhf_scf: upload failed");
    } else if (!c.diag_density(dH, nullptr, nocca, PA, 0)) {
        return scf_fail(c, PATMOL_HIP_EDEVICE, "patmol_hip_uhf_scf");
    } else if (noccb > 0) {
        const double one = 1.0, zero = 0.0;
        if (rocblas_dsyrk(c.h, rocblas_fill_lower, rocblas_operation_none, n, noccb, &one, c.C, n, &zero, c.T, n) !=
            rocblas_status_success)
            return aux_fail(PATMOL_HIP_EDEVICE, "patmol_hip_uhf_scf: rocblas_dsyrk failed");
        hipLaunchKernelGGL(pack_lower_kernel, dim3(n), dim3(256), 0, c.s, c.T, PB, n);
    } else if (hipMemsetAsync(PB, 0, sizeof(double) * no2, c.s) != hipSuccess) {
        return aux_fail(PATMOL_HIP_EDEVICE, "patmol_hip_uhf_scf: memset failed");
    }
    double eps = scf_eps;
    if (eps < EPS10 * n * n) eps = EPS10 * n * n;
    // below ~128 functions the eigensolver is faster than the ~40 GEMMs of
    // the purification (their launches and trace checks dominate)
    c.purify = g_scf_density >= 1 && n >= 128;
    const bool log = scf_log();
    PrecisionTail tail;
    double eold = 0.0, enrg = 0.0, ediff = 0.0, pdiff = 0.0, gms = 0.0;
    int iteration = 0, done = 0;
    bool first = true;
    auto t_start = std::chrono::steady_clock::now();
    for (;;) {
        ++iteration;
        (void)hipEventRecord(c.e0, c.s);
        if (hipMemsetAsync(GA, 0, sizeof(double) * no2, c.s) != hipSuccess ||
            hipMemsetAsync(GB, 0, sizeof(double) * no2, c.s) != hipSuccess)
            return aux_fail(PATMOL_HIP_EDEVICE, "patmol_hip_uhf_scf: memset failed");
        if (patmol_hip_uhf_gmatrix_device(mints, PA, PB, GA, GB, c.s)) return patmol_hip_last_error();
        if (xints && patmol_hip_uhf_gmatrix_device(xints, PA, PB, GA, GB, c.s)) return patmol_hip_last_error();
        (void)hipEventRecord(c.e1, c.s);
        double *fa = PA2, *fb = PB2;
        PA2 = PA1;
        PA1 = PA;
        PA = fa;
        PB2 = PB1;
        PB1 = PB;
        PB = fb;
        double sa[3], sb[3];
        if (!c.density(dH, GA, nocca, PA, 0) ||
            hipMemcpyAsync(evA, c.ev, sizeof(double) * n, hipMemcpyDeviceToDevice, c.s) != hipSuccess ||
            !c.sums3(PA, PA1, dH, GA, sa) || !c.density(dH, GB, noccb, PB, 1) || !c.sums3(PB, PB1, dH, GB, sb))
            return scf_fail(c, PATMOL_HIP_EDEVICE, "patmol_hip_uhf_scf");
        gms += c.elapsed_ms();
        pdiff = sqrt(2.0 * sa[2]) / n + sqrt(2.0 * sb[2]) / n;
        enrg = (sa[0] + sa[0] + sa[1]) + (sb[0] + sb[0] + sb[1]);
        ediff = enrg - eold;
        eold = enrg;
        int ret = 0;
        if (iteration == 1) {
        } else if (fabs(ediff) < EPS10 || pdiff < eps) {
            ret = 1;
        } else if (ediff > 0.0) {
            if (scf_accel == 1) {
                c.average(PA, PA1, PA2);
                c.average(PB, PB1, PB2);
            }
            if (scf_accel == 2 && (!c.accelerate(iteration, PA, PA1, PA2) || !c.accelerate(iteration, PB, PB1, PB2)))
                return scf_fail(c, PATMOL_HIP_EDEVICE, "patmol_hip_uhf_scf");
        }
        if (first) {
            first = false;
            iteration = 0;
            res->energy0 = enrg + enuc;
        } else {
            done = ret;
        }
        if (log)
            fprintf(stderr, "ITERATION = %12d EDIFF= %lg PDIFF= %lg\n  energy = %.12f\n", iteration, ediff, pdiff,
                    enrg + enuc);
        if (iteration == 0) continue;
        if (done || iteration >= maxits) break;
        if (tail.step(iteration, pdiff, eps) && log) fprintf(stderr, "FP64 G builds from iteration %d\n", iteration + 1);
    }
    if (hipStreamSynchronize(c.s) != hipSuccess) return aux_fail(PATMOL_HIP_EDEVICE, "patmol_hip_uhf_scf: device failure");
    const double tot = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_start).count();
    res->energy = enrg + enuc;
    res->eelec = enrg;
    res->enuc = enuc;
    res->ediff = ediff;
    res->pdiff = pdiff;
    res->iterations = iteration;
    res->converged = done;
    res->gbuild_ms = gms;
    res->other_ms = tot - gms;
    res->total_ms = tot;
    res->purify_fallbacks = c.purify_fallbacks;
    if (c.purify && (evalsA || evalsB)) {
        if (!c.orbitals(dH, GA) ||
            hipMemcpyAsync(evA, c.ev, sizeof(double) * n, hipMemcpyDeviceToDevice, c.s) != hipSuccess ||
            !c.orbitals(dH, GB))
            return scf_fail(c, PATMOL_HIP_EDEVICE, "patmol_hip_uhf_scf");
    }
    bool ok = true;
    if (PAout) ok = ok && hipMemcpy(PAout, PA, sizeof(double) * no2, hipMemcpyDeviceToHost) == hipSuccess;
    if (PBout) ok = ok && hipMemcpy(PBout, PB, sizeof(double) * no2, hipMemcpyDeviceToHost) == hipSuccess;
    if (evalsA) ok = ok && hipMemcpy(evalsA, evA, sizeof(double) * n, hipMemcpyDeviceToHost) == hipSuccess;
    if (evalsB) ok = ok && hipMemcpy(evalsB, c.ev, sizeof(double) * n, hipMemcpyDeviceToHost) == hipSuccess;
    if (!ok) return aux_fail(PATMOL_HIP_EDEVICE, "patmol_hip_uhf_scf: download failed");
    return 0;
}

// Building block for callers with their own loop: eigenvalues and
// phase-fixed eigenvectors (row-major, eigenvector j in column j) of X^T F X,
// and C = X V, for host packed S and F.
extern "C" int patmol_hip_diag_fock(const double *Smat, const double *Fmat, int nocc, double *evals, double *Cmat,
                                    double *Pmat)
{
    const int n = number_of_orbitals();
    if (!Smat || !Fmat || n <= 0 || nocc < 0 || nocc > n)
        return aux_fail(PATMOL_HIP_EBADARG, "patmol_hip_diag_fock: bad argument");
    Scf c;
    if (!c.init(n)) return scf_fail(c, PATMOL_HIP_EDEVICE, "patmol_hip_diag_fock");
    double *dS = c.mem.get<double>(c.no2), *dF = c.mem.get<double>(c.no2), *dP = c.mem.get<double>(c.no2);
    if (!dS || !dF || !dP) return aux_fail(PATMOL_HIP_EDEVICE, "patmol_hip_diag_fock: device allocation failed");
    if (hipMemcpyAsync(dS, Smat, sizeof(double) * c.no2, hipMemcpyHostToDevice, c.s) != hipSuccess ||
        hipMemcpyAsync(dF, Fmat, sizeof(double) * c.no2, hipMemcpyHostToDevice, c.s) != hipSuccess)
        return aux_fail(PATMOL_HIP_EDEVICE, "patmol_hip_diag_fock: upload failed");
    if (!c.form_x(dS) || !c.diag_density(dF, nullptr, nocc, dP, 0)) return scf_fail(c, PATMOL_HIP_EDEVICE, "patmol_hip_diag_fock");
    bool ok = true;
    if (evals) ok = ok && hipMemcpy(evals, c.ev, sizeof(double) * n, hipMemcpyDeviceToHost) == hipSuccess;
    if (Cmat) ok = ok && download_rowmajor(c, c.C, Cmat);
    if (Pmat) ok = ok && hipMemcpy(Pmat, dP, sizeof(double) * c.no2, hipMemcpyDeviceToHost) == hipSuccess;
    if (!ok) return aux_fail(PATMOL_HIP_EDEVICE, "patmol_hip_diag_fock: download failed");
    return 0;
}

extern "C" void patmol_hip_set_scf_density(int mode) { pm::g_scf_density = (mode == 0 || mode == 2) ? mode : 1; }
extern "C" void patmol_hip_set_scf_incremental(int period) { pm::g_scf_incremental = period > 0 ? period : 0; }

// RHF of every molecule of a batch, all iterating together (see patmol_hip.h)
extern "C" int patmol_hip_batch_rhf_scf(PatmolHipBatch *b, const double *S, const double *H, const int *nocc,
                                        const double *enuc, int maxits, double scf_eps, int scf_accel, double *energy,
                                        int *iterations, int *converged, double *Pout, double *total_ms,
                                        double *gbuild_ms)
{
    BatchView v;
    if (!batch_view(b, v)) return aux_fail(PATMOL_HIP_ESTATE, "patmol_hip_batch_rhf_scf: batch not finalised");
    if (!S || !H || !nocc || !enuc || !energy || !iterations || !converged)
        return aux_fail(PATMOL_HIP_EBADARG, "patmol_hip_batch_rhf_scf: null argument");
    if (scf_accel == 1 || scf_accel == 2)
        return aux_fail(PATMOL_HIP_EBADARG, "patmol_hip_batch_rhf_scf: scf_accel 0 (none) or 3 (DIIS)");
    const int nm = v.nmol;
    const bool diis = scf_accel == 3;
    const bool log = scf_log();
    const double EPS10 = 10 * DBL_EPSILON;
    hipStream_t st = nullptr;
    rocblas_handle hd = nullptr;
    hipEvent_t g0 = nullptr, g1 = nullptr;
    if (hipStreamCreate(&st) != hipSuccess || hipEventCreate(&g0) != hipSuccess || hipEventCreate(&g1) != hipSuccess ||
        rocblas_create_handle(&hd) != rocblas_status_success || rocblas_set_stream(hd, st) != rocblas_status_success)
        return aux_fail(PATMOL_HIP_EDEVICE, "patmol_hip_batch_rhf_scf: stream/handle creation failed");
    struct Mol {
        Scf c;
        double *dS = nullptr, *dH = nullptr, *Pn = nullptr;
        double eps = 0.0, eold = 0.0, enrg = 0.0;
        int iteration = 0, done = 0;
        bool first = true;
    };
    std::vector<Mol> mols(nm);
    int rc = 0;
    auto cleanup = [&]() {
        mols.clear();
        rocblas_destroy_handle(hd);
        (void)hipEventDestroy(g0);
        (void)hipEventDestroy(g1);
        (void)hipStreamDestroy(st);
    };
    for (int m = 0; m < nm && !rc; ++m) {
        Mol &x = mols[m];
        const int n = v.N[m];
        if (nocc[m] < 0 || nocc[m] > n) rc = aux_fail(PATMOL_HIP_EBADARG, "patmol_hip_batch_rhf_scf: bad nocc");
        else if (!x.c.init_shared(n, st, hd)) rc = scf_fail(x.c, PATMOL_HIP_EDEVICE, "patmol_hip_batch_rhf_scf");
        if (rc) break;
        const size_t no2 = x.c.no2;
        x.dS = x.c.mem.get<double>(no2);
        x.dH = x.c.mem.get<double>(no2);
        x.Pn = x.c.mem.get<double>(no2);
        double *Pin = v.d_P + v.pstart[m];
        if (!x.dS || !x.dH || !x.Pn ||
            hipMemcpyAsync(x.dS, S + v.pstart[m], sizeof(double) * no2, hipMemcpyHostToDevice, st) != hipSuccess ||
            hipMemcpyAsync(x.dH, H + v.pstart[m], sizeof(double) * no2, hipMemcpyHostToDevice, st) != hipSuccess) {
            rc = aux_fail(PATMOL_HIP_EDEVICE, "patmol_hip_batch_rhf_scf: upload failed");
            break;
        }
        if (!x.c.form_x(x.dS) || (diis && !x.c.diis_init(x.dS)) || !x.c.diag_density(x.dH, nullptr, nocc[m], Pin, 0)) {
            char where[96];
            snprintf(where, sizeof where, "patmol_hip_batch_rhf_scf: molecule %d, set-up", m);
            rc = scf_fail(x.c, PATMOL_HIP_EDEVICE, where);
            break;
        }
        x.c.purify = g_scf_density >= 1 && n >= 128;
        x.eps = scf_eps < EPS10 * n * n ? EPS10 * n * n : scf_eps;
    }
    if (rc) {
        cleanup();
        return rc;
    }
    auto t0 = std::chrono::steady_clock::now();
    double gms = 0.0;
    // the molecules the next replay builds: those still iterating (a
    // converged molecule's G is not read again)
    std::vector<int> active(nm, 1), built(nm, 1);
    auto restore_all = [&]() {
        std::fill(active.begin(), active.end(), 1);
        return batch_set_active(b, active.data(), st);
    };
    for (int sweep = 0;; ++sweep) {
        bool any = false;
        for (int m = 0; m < nm; ++m) {
            active[m] = (!mols[m].done && mols[m].iteration < maxits) ? 1 : 0;
            any = any || active[m];
        }
        if (!any) break;
        if (active != built) {
            if (batch_set_active(b, active.data(), st)) {
                const int e = patmol_hip_last_error();
                (void)restore_all();
                cleanup();
                return e;
            }
            built = active;
        }
        (void)hipEventRecord(g0, st);
        if (batch_launch(b, st)) {
            const int e = patmol_hip_last_error();
            (void)restore_all();
            cleanup();
            return e;
        }
        (void)hipEventRecord(g1, st);
        for (int m = 0; m < nm; ++m) {
            Mol &x = mols[m];
            if (x.done || x.iteration >= maxits) continue;
            ++x.iteration;
            double *Pin = v.d_P + v.pstart[m], *dG = v.d_G + v.pstart[m];
            // the rhfscf.c:21-71 update of this molecule: new density from
            // H + G(Pin), energy of Pin, density change, stop test
            const bool ok = diis && x.iteration > 1 ? x.c.diis_density(x.dH, dG, Pin, nocc[m], x.Pn, 0) : x.c.density(x.dH, dG, nocc[m], x.Pn, 0);
            double sm[3];
            if (!ok || !x.c.sums3(x.Pn, Pin, x.dH, dG, sm)) {
                char where[96];
                snprintf(where, sizeof where, "patmol_hip_batch_rhf_scf: molecule %d, sweep %d", m, sweep);
                rc = scf_fail(x.c, PATMOL_HIP_EDEVICE, where);
                break;
            }
            const double pdiff = sqrt(2.0 * sm[2]) / x.c.n;
            x.enrg = 2.0 * (sm[0] + sm[0] + sm[1]);
            const double ediff = x.enrg - x.eold;
            x.eold = x.enrg;
            if (log)
                fprintf(stderr, "batch sweep %d molecule %d: ITERATION = %d EDIFF= %lg PDIFF= %lg energy = %.12f\n", sweep,
                        m, x.iteration, ediff, pdiff, x.enrg + enuc[m]);
            int ret = 0;
            if (x.iteration > 1 && (fabs(ediff) < EPS10 || pdiff < x.eps)) ret = 1;
            if (x.first) {
                x.first = false;
                x.iteration = 0;
            } else {
                x.done = ret;
            }
            if (hipMemcpyAsync(Pin, x.Pn, sizeof(double) * x.c.no2, hipMemcpyDeviceToDevice, st) != hipSuccess) {
                rc = aux_fail(PATMOL_HIP_EDEVICE, "patmol_hip_batch_rhf_scf: copy failed");
                break;
            }
        }
        if (rc) break;
        float ms = 0.0f;
        if (hipEventSynchronize(g1) == hipSuccess && hipEventElapsedTime(&ms, g0, g1) == hipSuccess) gms += ms;
    }
    if (!rc && hipStreamSynchronize(st) != hipSuccess) rc = aux_fail(PATMOL_HIP_EDEVICE, "patmol_hip_batch_rhf_scf: device failure");
    // later G builds of the batch cover every molecule again
    if (active != std::vector<int>(nm, 1) || built != std::vector<int>(nm, 1)) {
        if (restore_all() && !rc) rc = patmol_hip_last_error();
    }
    if (!rc) {
        for (int m = 0; m < nm; ++m) {
            energy[m] = mols[m].enrg + enuc[m];
            iterations[m] = mols[m].iteration;
            converged[m] = mols[m].done;
        }
        if (Pout && hipMemcpy(Pout, v.d_P, sizeof(double) * v.no2tot, hipMemcpyDeviceToHost) != hipSuccess)
            rc = aux_fail(PATMOL_HIP_EDEVICE, "patmol_hip_batch_rhf_scf: download failed");
        if (total_ms)
            *total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        if (gbuild_ms) *gbuild_ms = gms;
    }
    cleanup();
    return rc;
}
