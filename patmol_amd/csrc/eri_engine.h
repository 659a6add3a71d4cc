// eri_engine.h -- Rys-quadrature electron-repulsion integrals for gfx950.
//
// Device building blocks shared by the J/K, ERI-export and Schwarz kernels:
//   * Rys nodes/weights from piecewise-Chebyshev tables (rys_table.inc,
//     converted to monomials in u when uploaded and evaluated by Horner),
//     evaluated in registers per primitive quartet (rys.c:2217-2243 computes
//     the same quadrature with its own fits);
//   * the 2-D recursion G[i][j] (rys.c:2269-2292) and the binomial transfer to
//     the second centre of each pair (rys.c:2295-2335), unrolled at compile
//     time per angular-momentum class so every index is a register;
//   * the primitive-quartet prefactor and screen of calc_pteints
//     (twoints.c:35-117): sr = 2 pi^(5/2) s12 s34 / (p q sqrt(p+q)),
//     skipped when sr < 1e-15.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <type_traits>
#include <utility>

namespace pm {

// Layout of rys_table.inc (gen_rys_tables.py): per root count n, TMAX(n) =
// 36 + 8(n-1) unit intervals, each RYS_NC Chebyshev coefficients x 2n values
// (n nodes then n weights); above TMAX(n) the Hermite large-T limit.  The
// device tables keep that order but pad every interval by two doubles
// (rys_stride): 24n + 2 doubles is 8(6n) + 4 dwords, so the intervals that
// the lanes of a wave select start in different LDS bank groups instead of
// all on the same bank (24n doubles is a multiple of 32 dwords for even n).
constexpr int RYS_NC = 12;
constexpr int RYS_NMAX = 7;
constexpr int rys_tmax(int n) { return 36 + 8 * (n - 1); }
constexpr int rys_stride(int n) { return RYS_NC * 2 * n + 2; }
constexpr int rys_off(int n)
{
    int o = 0;
    for (int m = 1; m < n; ++m) o += rys_tmax(m) * rys_stride(m);
    return o;
}
constexpr int rys_table_doubles() { return rys_off(RYS_NMAX + 1); }
// offsets of the unpadded layout of rys_table.inc
constexpr int rys_src_off(int n)
{
    int o = 0;
    for (int m = 1; m < n; ++m) o += rys_tmax(m) * RYS_NC * 2 * m;
    return o;
}

struct RysTables {
    const double *cheb;    // monomial coefficients in u of the Chebyshev fits
    const float *cheb32;   // the same from the first 9 Chebyshev terms, in float
    const double *herm_r;  // large-T nodes r_i (x = r_i / T), offset n(n-1)/2
    const double *herm_h;  // large-T weights h_i (w = h_i / sqrt(T))
};

constexpr double SRTERM = 34.9868366552497250;  // 2 pi^(5/2)
constexpr double PRIM_SCREEN = 1e-15;            // twoints.c:36

// ---------------------------------------------------------------------------
// compile-time helpers
// ---------------------------------------------------------------------------
template <typename F, int... Is>
__device__ __forceinline__ void static_for_impl(F &&f, std::integer_sequence<int, Is...>)
{
    (f(std::integral_constant<int, Is>{}), ...);
}
template <int N, typename F>
__device__ __forceinline__ void static_for(F &&f)
{
    static_for_impl(f, std::make_integer_sequence<int, N>{});
}

constexpr int ncart(int l) { return (l + 1) * (l + 2) / 2; }

// Cartesian components in the reference's order: lx descending, then ly
// descending (basis.c:294-308).
constexpr int cart_x(int l, int i)
{
    int k = 0;
    for (int lx = l; lx >= 0; --lx)
        for (int ly = l - lx; ly >= 0; --ly) {
            if (k == i) return lx;
            ++k;
        }
    return -1;
}
constexpr int cart_y(int l, int i)
{
    int k = 0;
    for (int lx = l; lx >= 0; --lx)
        for (int ly = l - lx; ly >= 0; --ly) {
            if (k == i) return ly;
            ++k;
        }
    return -1;
}
constexpr int cart_z(int l, int i) { return l - cart_x(l, i) - cart_y(l, i); }

constexpr double dfact_odd(int k)  // (2k-1)!!
{
    double r = 1.0;
    for (int i = 2 * k - 1; i > 1; i -= 2) r *= i;
    return r;
}
// (2lx-1)!!(2ly-1)!!(2lz-1)!!; the component factor of basis.c:309-319 is
// its inverse square root.
constexpr double cart_dfprod(int l, int i)
{
    return dfact_odd(cart_x(l, i)) * dfact_odd(cart_y(l, i)) * dfact_odd(cart_z(l, i));
}

constexpr double csqrt(double x)
{
    double r = x > 1.0 ? x : 1.0;
    for (int i = 0; i < 200; ++i) r = 0.5 * (r + x / r);
    return r;
}

constexpr double binom(int n, int k)
{
    double r = 1.0;
    for (int i = 1; i <= k; ++i) r = r * (n - k + i) / i;
    return r;
}

// ---------------------------------------------------------------------------
// data layouts
// ---------------------------------------------------------------------------
struct PrimPair {       // one primitive pair of a shell pair
    double p;           // alpha + beta
    double P[3];        // product centre
    double s;           // exp(-alpha beta/p |A-B|^2)
    double c;           // c_alpha c_beta (normalised contraction coefficients)
    double ip;          // 1 / p
};

struct PairRef {        // shell pair in the sorted shell order, I >= J
    int pp0, npp;       // primitive pairs [pp0, pp0 + npp)
};

// ---------------------------------------------------------------------------
// Rys quadrature: nodes x = t^2, weights w (sum w = F_0(T))
// ---------------------------------------------------------------------------
template <typename R, int N>
struct RysRule {
    R x[N], w[N];
};

// 1/sqrt(x) for a normal positive x: the hardware estimate (v_rsq_f64)
// refined by one second-order Newton step y (1 + r/2 + 3r^2/8) with
// r = 1 - x y^2, which takes a relative error e of the estimate to about
// 5e^3/16; about a third of the instructions of 1.0 / sqrt(x) with IEEE
// division and square root.
__device__ __forceinline__ double rsqrt_pos(double x)
{
    const double y = __builtin_amdgcn_rsq(x);
    const double r = fma(-x * y, y, 1.0);
    return fma(y * r, fma(0.375, r, 0.5), y);
}

// With SECT the table pointer addresses the N-root section itself (a copy of
// just that section, e.g. in LDS) instead of the start of all sections.
template <int N, bool SECT = false>
__device__ __forceinline__ void rys_eval(double T, double (&x)[N], double (&w)[N], const RysTables &tb)
{
    if (T < (double)rys_tmax(N)) {
        int k = (int)T;
        const double u = 2.0 * (T - k) - 1.0;
        const double *__restrict__ c = tb.cheb + (SECT ? 0 : rys_off(N)) + k * rys_stride(N);
        double y[2 * N];
#pragma unroll
        for (int v = 0; v < 2 * N; ++v) y[v] = c[(RYS_NC - 1) * 2 * N + v];
#pragma unroll
        for (int j = RYS_NC - 2; j >= 0; --j) {
#pragma unroll
            for (int v = 0; v < 2 * N; ++v) y[v] = fma(y[v], u, c[j * 2 * N + v]);
        }
#pragma unroll
        for (int v = 0; v < N; ++v) {
            x[v] = y[v];
            w[v] = y[N + v];
        }
    } else {
        const double rs = rsqrt_pos(T), it = rs * rs;
#pragma unroll
        for (int v = 0; v < N; ++v) {
            x[v] = tb.herm_r[N * (N - 1) / 2 + v] * it;
            w[v] = tb.herm_h[N * (N - 1) / 2 + v] * rs;
        }
    }
}

template <int N, bool SECT = false>
__device__ __forceinline__ void rys_eval(float T, float (&x)[N], float (&w)[N], const RysTables &tb)
{
    // FP32 evaluation: the first NC32 Chebyshev terms of the same fits
    // (truncation error below 1e-9 relative, under FP32 rounding).
    constexpr int NC32 = 9;
    if (T < (float)rys_tmax(N)) {
        int k = (int)T;
        const float u = 2.0f * (T - (float)k) - 1.0f;
        const float *__restrict__ c = tb.cheb32 + rys_off(N) + k * rys_stride(N);
        float y[2 * N];
#pragma unroll
        for (int v = 0; v < 2 * N; ++v) y[v] = c[(NC32 - 1) * 2 * N + v];
#pragma unroll
        for (int j = NC32 - 2; j >= 0; --j) {
#pragma unroll
            for (int v = 0; v < 2 * N; ++v) y[v] = fmaf(y[v], u, c[j * 2 * N + v]);
        }
#pragma unroll
        for (int v = 0; v < N; ++v) {
            x[v] = y[v];
            w[v] = y[N + v];
        }
    } else {
        float it = 1.0f / T, rs = 1.0f / sqrtf(T);
#pragma unroll
        for (int v = 0; v < N; ++v) {
            x[v] = (float)tb.herm_r[N * (N - 1) / 2 + v] * it;
            w[v] = (float)tb.herm_h[N * (N - 1) / 2 + v] * rs;
        }
    }
}

// One root r (run-time index) of the N-point rule: the same fits, two
// Clenshaw series instead of 2N.
template <int N, bool SECT = false>
__device__ __forceinline__ void rys_eval_one(double T, int r, double &x, double &w, const RysTables &tb)
{
    if (T < (double)rys_tmax(N)) {
        int k = (int)T;
        const double u = 2.0 * (T - k) - 1.0;
        const double *__restrict__ c = tb.cheb + (SECT ? 0 : rys_off(N)) + k * rys_stride(N) + r;
        double yx = c[(RYS_NC - 1) * 2 * N], yw = c[(RYS_NC - 1) * 2 * N + N];
#pragma unroll
        for (int j = RYS_NC - 2; j >= 0; --j) {
            yx = fma(yx, u, c[j * 2 * N]);
            yw = fma(yw, u, c[j * 2 * N + N]);
        }
        x = yx;
        w = yw;
    } else {
        const double rs = rsqrt_pos(T);
        x = tb.herm_r[N * (N - 1) / 2 + r] * (rs * rs);
        w = tb.herm_h[N * (N - 1) / 2 + r] * rs;
    }
}

template <int N, bool SECT = false>
__device__ __forceinline__ void rys_eval_one(float T, int r, float &x, float &w, const RysTables &tb)
{
    constexpr int NC32 = 9;
    if (T < (float)rys_tmax(N)) {
        int k = (int)T;
        const float u = 2.0f * (T - (float)k) - 1.0f;
        const float *__restrict__ c = tb.cheb32 + rys_off(N) + k * rys_stride(N) + r;
        float yx = c[(NC32 - 1) * 2 * N], yw = c[(NC32 - 1) * 2 * N + N];
#pragma unroll
        for (int j = NC32 - 2; j >= 0; --j) {
            yx = fmaf(yx, u, c[j * 2 * N]);
            yw = fmaf(yw, u, c[j * 2 * N + N]);
        }
        x = yx;
        w = yw;
    } else {
        x = (float)tb.herm_r[N * (N - 1) / 2 + r] / T;
        w = (float)tb.herm_h[N * (N - 1) / 2 + r] / sqrtf(T);
    }
}

// ---------------------------------------------------------------------------
// 2-D recursion, restating RysRecur (rys.c:2269-2292):
//   G[0][0] = 1, G[1][0] = C, G[0][1] = C', G[1][1] = B00 + C C',
//   G[0][j+1] = j B1' G[0][j-1] + C' G[0][j]
//   G[i][0]   = (i-1) B1 G[i-2][0] + C G[i-1][0]
//   G[i][j+1] = j B1' G[i][j-1] + i B00 G[i-1][j] + C' G[i][j]
// ---------------------------------------------------------------------------
template <typename R, int L12, int L34>
__device__ __forceinline__ void rys_vrr(R (&G)[L12 + 1][L34 + 1], R B00, R B1, R B1p, R C, R Cp, R scale)
{
    G[0][0] = scale;
    if constexpr (L12 >= 1) G[1][0] = C * scale;
#pragma unroll
    for (int i = 2; i <= L12; ++i) G[i][0] = R(i - 1) * B1 * G[i - 2][0] + C * G[i - 1][0];
    if constexpr (L34 >= 1) {
#pragma unroll
        for (int i = 0; i <= L12; ++i) {
            G[i][1] = Cp * G[i][0];
            if (i >= 1) G[i][1] += R(i) * B00 * G[i - 1][0];
        }
#pragma unroll
        for (int j = 1; j < L34; ++j) {
#pragma unroll
            for (int i = 0; i <= L12; ++i) {
                R t = R(j) * B1p * G[i][j - 1] + Cp * G[i][j];
                if (i >= 1) t += R(i) * B00 * G[i - 1][j];
                G[i][j + 1] = t;
            }
        }
    }
}

// Transfer to the second centre of each pair (rys.c:2295-2335):
// I(a,b,c,d) = sum_i C(b,i) AB^i sum_j C(d,j) CD^j G[a+b-i][c+d-j].
template <typename R, int A, int B, int C, int D, int L12, int L34, int LB, int LD>
__device__ __forceinline__ R rys_hrr(const R (&G)[L12 + 1][L34 + 1], const R (&abp)[LB + 1], const R (&cdp)[LD + 1])
{
    R s = R(0);
#pragma unroll
    for (int i = 0; i <= B; ++i) {
        R t = R(0);
#pragma unroll
        for (int j = 0; j <= D; ++j) t += R(binom(D, j)) * cdp[j] * G[A + B - i][C + D - j];
        s += R(binom(B, i)) * abp[i] * t;
    }
    return s;
}

// ---------------------------------------------------------------------------
// Primitive data of one quartet side, as loaded from the pair tables.
// ---------------------------------------------------------------------------
template <typename R>
struct PairGeom {
    R A[3], AB[3];  // first centre, first minus second centre
};

// Integrals of one primitive quartet.  Calls f(ia, ib, ic, id, v) for every
// Cartesian component (ia..id are std::integral_constant in the reference's
// order; ia restricted to [CA0, CA1) and ib to [CB0, CB1), by default all of
// them); v already carries the primitive prefactor, contraction
// coefficients and the component normalisation of basis.c:309-319.
// Returns false when the primitive quartet falls under the prefactor screen
// (its integrals are then passed as exact zeros).
template <typename R, int LA, int LB, int LC, int LD, bool SCREEN = true, int CA0 = 0, int CA1 = ncart(LA),
          int CB0 = 0, int CB1 = ncart(LB), bool SECT = false, typename F>
__device__ __forceinline__ bool primitive_quartet(const PrimPair &bp, const PrimPair &kp, const PairGeom<R> &bra,
                                                  const PairGeom<R> &ket, const RysTables &tab, R scale,
                                                  F &&f)
{
    constexpr int L12 = LA + LB, L34 = LC + LD, NR = (L12 + L34) / 2 + 1;
    const double p = bp.p, q = kp.p, pq = p + q;
    const double rpq = rsqrt_pos(pq), ipq_d = rpq * rpq;
    const double sr0 = SRTERM * (bp.s * bp.ip) * (kp.s * kp.ip) * rpq;
    // Screened primitive quartets still run (with a zero prefactor) so that
    // every lane reaches the cross-lane reductions inside f.  Bounds (SCREEN
    // false) must not use the screen: it can zero a diagonal (ab|ab) whose
    // off-diagonal partners survive it.
    const bool keep = !SCREEN || sr0 >= PRIM_SCREEN;
    const R sr = keep ? R(sr0 * bp.c * kp.c) * scale : R(0);
    const R PQ[3] = {R(bp.P[0] - kp.P[0]), R(bp.P[1] - kp.P[1]), R(bp.P[2] - kp.P[2])};
    const R rp = R(p), rq = R(q), ipq = R(ipq_d), irp = R(bp.ip), irq = R(kp.ip);
    const R T = (PQ[0] * PQ[0] + PQ[1] * PQ[1] + PQ[2] * PQ[2]) * rp * rq * ipq;
    R x[NR], w[NR];
    rys_eval<NR, SECT>(T, x, w, tab);
    const R PA[3] = {R(bp.P[0]) - bra.A[0], R(bp.P[1]) - bra.A[1], R(bp.P[2]) - bra.A[2]};
    const R QC[3] = {R(kp.P[0]) - ket.A[0], R(kp.P[1]) - ket.A[1], R(kp.P[2]) - ket.A[2]};
    R Gx[NR][L12 + 1][L34 + 1], Gy[NR][L12 + 1][L34 + 1], Gz[NR][L12 + 1][L34 + 1];
#pragma unroll
    for (int r = 0; r < NR; ++r) {
        const R fff = x[r] * ipq;
        const R B00 = R(0.5) * fff;
        const R B1 = (R(0.5) - B00 * rq) * irp;
        const R B1p = (R(0.5) - B00 * rp) * irq;
        rys_vrr<R, L12, L34>(Gx[r], B00, B1, B1p, PA[0] - rq * PQ[0] * fff, QC[0] + rp * PQ[0] * fff, R(1));
        rys_vrr<R, L12, L34>(Gy[r], B00, B1, B1p, PA[1] - rq * PQ[1] * fff, QC[1] + rp * PQ[1] * fff, R(1));
        rys_vrr<R, L12, L34>(Gz[r], B00, B1, B1p, PA[2] - rq * PQ[2] * fff, QC[2] + rp * PQ[2] * fff, w[r] * sr);
    }
    R abp[3][LB + 1], cdp[3][LD + 1];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        abp[a][0] = R(1);
#pragma unroll
        for (int i = 1; i <= LB; ++i) abp[a][i] = abp[a][i - 1] * bra.AB[a];
        cdp[a][0] = R(1);
#pragma unroll
        for (int i = 1; i <= LD; ++i) cdp[a][i] = cdp[a][i - 1] * ket.AB[a];
    }
    static_for<CA1 - CA0>([&](auto ia0) {
        using ia_t = std::integral_constant<int, CA0 + decltype(ia0)::value>;
        static_for<CB1 - CB0>([&](auto ib0) {
            using ib_t = std::integral_constant<int, CB0 + decltype(ib0)::value>;
            static_for<ncart(LC)>([&](auto ic) {
                static_for<ncart(LD)>([&](auto id) {
                    constexpr int IA = ia_t::value, IB = ib_t::value;
                    constexpr int IC = decltype(ic)::value, ID = decltype(id)::value;
                    constexpr int ax = cart_x(LA, IA), ay = cart_y(LA, IA), az = cart_z(LA, IA);
                    constexpr int bx = cart_x(LB, IB), by = cart_y(LB, IB), bz = cart_z(LB, IB);
                    constexpr int cx = cart_x(LC, IC), cy = cart_y(LC, IC), cz = cart_z(LC, IC);
                    constexpr int dx = cart_x(LD, ID), dy = cart_y(LD, ID), dz = cart_z(LD, ID);
                    R v = R(0);
#pragma unroll
                    for (int r = 0; r < NR; ++r) {
                        R ix = rys_hrr<R, ax, bx, cx, dx, L12, L34, LB, LD>(Gx[r], abp[0], cdp[0]);
                        R iy = rys_hrr<R, ay, by, cy, dy, L12, L34, LB, LD>(Gy[r], abp[1], cdp[1]);
                        R iz = rys_hrr<R, az, bz, cz, dz, L12, L34, LB, LD>(Gz[r], abp[2], cdp[2]);
                        v += ix * iy * iz;
                    }
                    constexpr double nrm = 1.0 / csqrt(cart_dfprod(LA, IA) * cart_dfprod(LB, IB) *
                                                       cart_dfprod(LC, IC) * cart_dfprod(LD, ID));
                    f(ia_t{}, ib_t{}, ic, id, v * R(nrm));
                });
            });
        });
    });
    return keep;
}

// Same integrals, one Rys root at a time: f(ia, ib, ic, id, v_r) receives
// the contribution of each root separately (the integral is the sum over
// roots, and every use of it here is linear), so only one root's 2-D tables
// are live instead of all of them.
template <typename R, int LA, int LB, int LC, int LD, bool BLOCK_FENCE, int CA0 = 0, int CA1 = ncart(LA),
          int CB0 = 0, int CB1 = ncart(LB), bool SECT = false, typename F>
__device__ __forceinline__ void primitive_quartet_rootwise(const PrimPair &bp, const PrimPair &kp,
                                                           const PairGeom<R> &bra, const PairGeom<R> &ket,
                                                           const RysTables &tab, R scale, F &&f)
{
    constexpr int L12 = LA + LB, L34 = LC + LD, NR = (L12 + L34) / 2 + 1;
    const double p = bp.p, q = kp.p, pq = p + q;
    const double rpq = rsqrt_pos(pq), ipq_d = rpq * rpq;
    const double sr0 = SRTERM * (bp.s * bp.ip) * (kp.s * kp.ip) * rpq;
    const bool keep = sr0 >= PRIM_SCREEN;
    const R sr = keep ? R(sr0 * bp.c * kp.c) * scale : R(0);
    const R PQ[3] = {R(bp.P[0] - kp.P[0]), R(bp.P[1] - kp.P[1]), R(bp.P[2] - kp.P[2])};
    const R rp = R(p), rq = R(q), ipq = R(ipq_d), irp = R(bp.ip), irq = R(kp.ip);
    const R T = (PQ[0] * PQ[0] + PQ[1] * PQ[1] + PQ[2] * PQ[2]) * rp * rq * ipq;
    const R PA[3] = {R(bp.P[0]) - bra.A[0], R(bp.P[1]) - bra.A[1], R(bp.P[2]) - bra.A[2]};
    const R QC[3] = {R(kp.P[0]) - ket.A[0], R(kp.P[1]) - ket.A[1], R(kp.P[2]) - ket.A[2]};
    R abp[3][LB + 1], cdp[3][LD + 1];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        abp[a][0] = R(1);
#pragma unroll
        for (int i = 1; i <= LB; ++i) abp[a][i] = abp[a][i - 1] * bra.AB[a];
        cdp[a][0] = R(1);
#pragma unroll
        for (int i = 1; i <= LD; ++i) cdp[a][i] = cdp[a][i - 1] * ket.AB[a];
    }
    // a rolled loop: one root's tables at a time; the memory clobber keeps the
    // caller's density reads inside the loop instead of hoisted into registers
#pragma unroll 1
    for (int r = 0; r < NR; ++r) {
        __asm__ volatile("" ::: "memory");
        R xr, wr;
        rys_eval_one<NR, SECT>(T, r, xr, wr, tab);
        R Gx[L12 + 1][L34 + 1], Gy[L12 + 1][L34 + 1], Gz[L12 + 1][L34 + 1];
        const R fff = xr * ipq;
        const R B00 = R(0.5) * fff;
        const R B1 = (R(0.5) - B00 * rq) * irp;
        const R B1p = (R(0.5) - B00 * rp) * irq;
        rys_vrr<R, L12, L34>(Gx, B00, B1, B1p, PA[0] - rq * PQ[0] * fff, QC[0] + rp * PQ[0] * fff, R(1));
        rys_vrr<R, L12, L34>(Gy, B00, B1, B1p, PA[1] - rq * PQ[1] * fff, QC[1] + rp * PQ[1] * fff, R(1));
        rys_vrr<R, L12, L34>(Gz, B00, B1, B1p, PA[2] - rq * PQ[2] * fff, QC[2] + rp * PQ[2] * fff, wr * sr);
        static_for<CA1 - CA0>([&](auto ia0) {
            using ia_t = std::integral_constant<int, CA0 + decltype(ia0)::value>;
            static_for<CB1 - CB0>([&](auto ib0) {
                using ib_t = std::integral_constant<int, CB0 + decltype(ib0)::value>;
                // with BLOCK_FENCE the caller's density reads for this (a, b)
                // block are issued here, not hoisted for all blocks at once
                if constexpr (BLOCK_FENCE) __asm__ volatile("" ::: "memory");
                static_for<ncart(LC)>([&](auto ic) {
                    static_for<ncart(LD)>([&](auto id) {
                        constexpr int IA = ia_t::value, IB = ib_t::value;
                        constexpr int IC = decltype(ic)::value, ID = decltype(id)::value;
                        constexpr int ax = cart_x(LA, IA), ay = cart_y(LA, IA), az = cart_z(LA, IA);
                        constexpr int bx = cart_x(LB, IB), by = cart_y(LB, IB), bz = cart_z(LB, IB);
                        constexpr int cx = cart_x(LC, IC), cy = cart_y(LC, IC), cz = cart_z(LC, IC);
                        constexpr int dx = cart_x(LD, ID), dy = cart_y(LD, ID), dz = cart_z(LD, ID);
                        constexpr double nrm = 1.0 / csqrt(cart_dfprod(LA, IA) * cart_dfprod(LB, IB) *
                                                           cart_dfprod(LC, IC) * cart_dfprod(LD, ID));
                        const R ix = rys_hrr<R, ax, bx, cx, dx, L12, L34, LB, LD>(Gx, abp[0], cdp[0]);
                        const R iy = rys_hrr<R, ay, by, cy, dy, L12, L34, LB, LD>(Gy, abp[1], cdp[1]);
                        const R iz = rys_hrr<R, az, bz, cz, dz, L12, L34, LB, LD>(Gz, abp[2], cdp[2]);
                        f(ia_t{}, ib_t{}, ic, id, ix * iy * (iz * R(nrm)));
                    });
                });
            });
        });
    }
}

}  // namespace pm

namespace pm {

// ---------------------------------------------------------------------------
// Large classes (d on both bra centres with a p/d ket): the per-root, per-axis
// transferred 1-D integrals I[r][axis][ax][bx][cx][dx] are tabulated once per
// primitive quartet (unrolled), and the bra components are then looped over
// at run time.  This keeps the generated code and compile time bounded; the
// table lives in private memory.
// ---------------------------------------------------------------------------
template <int LA, int LB, int LC, int LD>
struct HeavyClass {
    static constexpr bool value = ncart(LA) * ncart(LB) * ncart(LC) * ncart(LD) >= 324 && LB == 2;
};

// component exponents of shell l packed 4 bits per component: (x | y<<2)
constexpr uint32_t cart_pack(int l)
{
    uint32_t v = 0;
    for (int i = 0; i < ncart(l); ++i) v |= (uint32_t)(cart_x(l, i) | (cart_y(l, i) << 2)) << (4 * i);
    return v;
}

template <typename R, int LA, int LB, int LC, int LD>
struct OneDTable {
    static constexpr int L12 = LA + LB, L34 = LC + LD, NR = (L12 + L34) / 2 + 1;
    static constexpr int SA = (LB + 1) * (LC + 1) * (LD + 1), SB = (LC + 1) * (LD + 1), SC = LD + 1;
    static constexpr int NX = (LA + 1) * SA;
    R t[NR][3][NX];

    // Returns false when the primitive quartet is under the prefactor screen
    // (the table then holds exact zeros in the z factors).
    __device__ __forceinline__ bool build(const PrimPair &bp, const PrimPair &kp, const PairGeom<R> &bra,
                                          const PairGeom<R> &ket, const RysTables &tab, R scale,
                                          bool screen = true)
    {
        const double p = bp.p, q = kp.p, pq = p + q;
        const double rpq = rsqrt_pos(pq), ipq_d = rpq * rpq;
        const double sr0 = SRTERM * (bp.s * bp.ip) * (kp.s * kp.ip) * rpq;
        const bool keep = !screen || sr0 >= PRIM_SCREEN;
        const R sr = keep ? R(sr0 * bp.c * kp.c) * scale : R(0);
        const R PQ[3] = {R(bp.P[0] - kp.P[0]), R(bp.P[1] - kp.P[1]), R(bp.P[2] - kp.P[2])};
        const R rp = R(p), rq = R(q), ipq = R(ipq_d), irp = R(bp.ip), irq = R(kp.ip);
        const R T = (PQ[0] * PQ[0] + PQ[1] * PQ[1] + PQ[2] * PQ[2]) * rp * rq * ipq;
        R x[NR], w[NR];
        rys_eval<NR>(T, x, w, tab);
        const R PA[3] = {R(bp.P[0]) - bra.A[0], R(bp.P[1]) - bra.A[1], R(bp.P[2]) - bra.A[2]};
        const R QC[3] = {R(kp.P[0]) - ket.A[0], R(kp.P[1]) - ket.A[1], R(kp.P[2]) - ket.A[2]};
        R abp[3][LB + 1], cdp[3][LD + 1];
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            abp[a][0] = R(1);
#pragma unroll
            for (int i = 1; i <= LB; ++i) abp[a][i] = abp[a][i - 1] * bra.AB[a];
            cdp[a][0] = R(1);
#pragma unroll
            for (int i = 1; i <= LD; ++i) cdp[a][i] = cdp[a][i - 1] * ket.AB[a];
        }
#pragma unroll
        for (int r = 0; r < NR; ++r) {
            const R fff = x[r] * ipq;
            const R B00 = R(0.5) * fff;
            const R B1 = (R(0.5) - B00 * rq) * irp;
            const R B1p = (R(0.5) - B00 * rp) * irq;
#pragma unroll
            for (int ax = 0; ax < 3; ++ax) {
                R G[L12 + 1][L34 + 1];
                rys_vrr<R, L12, L34>(G, B00, B1, B1p, PA[ax] - rq * PQ[ax] * fff, QC[ax] + rp * PQ[ax] * fff,
                                     ax == 2 ? w[r] * sr : R(1));
                static_for<LA + 1>([&](auto ia) {
                    static_for<LB + 1>([&](auto ib) {
                        static_for<LC + 1>([&](auto ic) {
                            static_for<LD + 1>([&](auto id) {
                                constexpr int A_ = decltype(ia)::value, B_ = decltype(ib)::value;
                                constexpr int C_ = decltype(ic)::value, D_ = decltype(id)::value;
                                t[r][ax][A_ * SA + B_ * SB + C_ * SC + D_] =
                                    rys_hrr<R, A_, B_, C_, D_, L12, L34, LB, LD>(G, abp[ax], cdp[ax]);
                            });
                        });
                    });
                });
            }
        }
        return keep;
    }

    // integral of component (ia, ib | IC, ID): ia, ib at run time
    template <int IC, int ID>
    __device__ __forceinline__ R value(int ia, int ib) const
    {
        constexpr uint32_t PA_ = cart_pack(LA), PB_ = cart_pack(LB);
        const int ax = (PA_ >> (4 * ia)) & 3, ay = (PA_ >> (4 * ia + 2)) & 3, az = LA - ax - ay;
        const int bx = (PB_ >> (4 * ib)) & 3, by = (PB_ >> (4 * ib + 2)) & 3, bz = LB - bx - by;
        constexpr int cx = cart_x(LC, IC), cy = cart_y(LC, IC), cz = cart_z(LC, IC);
        constexpr int dx = cart_x(LD, ID), dy = cart_y(LD, ID), dz = cart_z(LD, ID);
        const int ix = ax * SA + bx * SB + cx * SC + dx;
        const int iy = ay * SA + by * SB + cy * SC + dy;
        const int iz = az * SA + bz * SB + cz * SC + dz;
        R v = R(0);
#pragma unroll
        for (int r = 0; r < NR; ++r) v += t[r][0][ix] * t[r][1][iy] * t[r][2][iz];
        return v;
    }
};

// 1/sqrt of the component double-factorial product, at run time for bra
// components (table of at most 6 entries per l)
template <int L>
__device__ __forceinline__ double cart_norm_rt(int i)
{
    double v = 1.0;
    static_for<ncart(L)>([&](auto k) {
        constexpr int K = decltype(k)::value;
        constexpr double n = 1.0 / csqrt(cart_dfprod(L, K));
        if (i == K) v = n;
    });
    return v;
}

}  // namespace pm
