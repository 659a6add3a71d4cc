// dpm.hip -- distributed-positron-model potential, the drop-in dpm_ints
// (dpmints.c:120-199).
//
// The positron is a unit positive charge smeared as a fixed six-Gaussian fit
// centred on skip_center() (dpmints.c:31-36).  Every shell pair (a,b) of the
// current basis gets V_ab = sum_k c_k (ab | exp(-alpha_k |r-Q|^2)), and
// hmat[ij] -= V_ab for the packed lower triangle (dpmints.c:185-188).  Each
// fit Gaussian is a one-primitive s "pair" of exponent alpha_k at Q, so the
// three-centre integral is the (LA LB | 0 0) case of the Rys engine with
// unit overlap factor, which reproduces the prefactor of dpmints.c:82.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <vector>

#include "eri_engine.h"
#include "patmol_hip.h"

extern "C" {
const Shell *shells(void);
const Center *centers(void);
int number_of_shells(void);
int number_of_orbitals(void);
int skip_center(void);
}

namespace pm {

// dpmints.c:31-36
__constant__ double kDpmAlpha[6] = {6.8505018000, 4.0491646000, 3.5941062989, 1.2478274000, 0.7927690989, 0.3377107978};
__constant__ double kDpmCoef[6] = {0.0766926784, 0.1483475912, 0.0462334641, 0.0717376427, 0.0447149792, 0.0069678529};

struct DpmArgs {
    const int *pI, *pJ;          // shells of each pair (first has the higher l)
    int npairs;
    const int *shell_pp0;        // unused placeholder for symmetry with ERI tables
    const double *xyz;           // per shell centre
    const PairRef *pairs;        // per listed pair
    const PrimPair *pp;
    const double *Q;             // positron position
    double *out;                 // per pair: NA*NB values
    RysTables rys;
};

template <int LA, int LB>
__global__ __launch_bounds__(64) void dpm_kernel(DpmArgs a)
{
    constexpr int NAB = ncart(LA) * ncart(LB);
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.npairs) return;
    const int I = a.pI[t], J = a.pJ[t];
    PairGeom<double> bg, kg;
    for (int x = 0; x < 3; ++x) {
        bg.A[x] = a.xyz[3 * I + x];
        bg.AB[x] = a.xyz[3 * I + x] - a.xyz[3 * J + x];
        kg.A[x] = a.Q[x];
        kg.AB[x] = 0.0;
    }
    double acc[NAB];
#pragma unroll
    for (int i = 0; i < NAB; ++i) acc[i] = 0.0;
    const PairRef ref = a.pairs[t];
    for (int ib = 0; ib < ref.npp; ++ib) {
        const PrimPair bp = a.pp[ref.pp0 + ib];
        for (int k = 0; k < 6; ++k) {
            PrimPair kp;
            kp.p = kDpmAlpha[k];
            kp.P[0] = a.Q[0];
            kp.P[1] = a.Q[1];
            kp.P[2] = a.Q[2];
            kp.s = 1.0;
            kp.c = kDpmCoef[k];
            kp.ip = 1.0 / kp.p;
            primitive_quartet<double, LA, LB, 0, 0>(bp, kp, bg, kg, a.rys, 1.0,
                                                    [&](auto ia, auto ib_, auto, auto, double v) {
                                                        acc[decltype(ia)::value * ncart(LB) + decltype(ib_)::value] += v;
                                                    });
        }
    }
#pragma unroll
    for (int i = 0; i < NAB; ++i) a.out[(size_t)t * 36 + i] = acc[i];
}

int launch_dpm(int la, int lb, const DpmArgs &args, hipStream_t s)
{
    dim3 g((args.npairs + 63) / 64), b(64);
    switch (la * 3 + lb) {
    case 0: hipLaunchKernelGGL((dpm_kernel<0, 0>), g, b, 0, s, args); return 0;
    case 3: hipLaunchKernelGGL((dpm_kernel<1, 0>), g, b, 0, s, args); return 0;
    case 4: hipLaunchKernelGGL((dpm_kernel<1, 1>), g, b, 0, s, args); return 0;
    case 6: hipLaunchKernelGGL((dpm_kernel<2, 0>), g, b, 0, s, args); return 0;
    case 7: hipLaunchKernelGGL((dpm_kernel<2, 1>), g, b, 0, s, args); return 0;
    case 8: hipLaunchKernelGGL((dpm_kernel<2, 2>), g, b, 0, s, args); return 0;
    default: return -1;
    }
}

// defined in hip_context.cpp
int aux_fail(int code, const char *msg);
RysTables aux_tables();

}  // namespace pm

extern "C" void dpm_ints(double *hmat)
{
    using namespace pm;
    const Shell *sh = shells();
    const Center *ce = centers();
    const int ns = number_of_shells(), sc = skip_center();
    if (!sh || !ce || ns <= 0 || !hmat) {
        aux_fail(PATMOL_HIP_EBADARG, "dpm_ints: no basis or null hmat");
        return;
    }
    RysTables rys = aux_tables();
    if (!rys.cheb) return;
    const double *Q = ce[sc].r;
    std::vector<double> xyz(3 * ns);
    for (int s = 0; s < ns; ++s) {
        if (sh[s].lsh > 2) {
            aux_fail(PATMOL_HIP_EUNSUPPORTED, "dpm_ints: l > 2");
            return;
        }
        for (int x = 0; x < 3; ++x) xyz[3 * s + x] = ce[sh[s].cen].r[x];
    }
    for (int la = 0; la <= 2; ++la)
        for (int lb = 0; lb <= la; ++lb) {
            std::vector<int> pI, pJ;
            std::vector<PairRef> refs;
            std::vector<PrimPair> pp;
            for (int i = 0; i < ns; ++i)
                for (int j = 0; j <= i; ++j) {
                    int a = i, b = j;
                    if (sh[a].lsh < sh[b].lsh) std::swap(a, b);
                    if (sh[a].lsh != la || sh[b].lsh != lb) continue;
                    pI.push_back(a);
                    pJ.push_back(b);
                    PairRef r{(int)pp.size(), sh[a].npr * sh[b].npr};
                    const double *A = ce[sh[a].cen].r, *B = ce[sh[b].cen].r;
                    double ab2 = 0;
                    for (int x = 0; x < 3; ++x) ab2 += (A[x] - B[x]) * (A[x] - B[x]);
                    for (int u = 0; u < sh[a].npr; ++u)
                        for (int v = 0; v < sh[b].npr; ++v) {
                            PrimPair q;
                            double al = sh[a].al[u], be = sh[b].al[v];
                            q.p = al + be;
                            for (int x = 0; x < 3; ++x) q.P[x] = (al * A[x] + be * B[x]) / q.p;
                            q.s = exp(-al * be * ab2 / q.p);
                            q.c = sh[a].co[u] * sh[b].co[v];
                            q.ip = 1.0 / q.p;
                            pp.push_back(q);
                        }
                    refs.push_back(r);
                }
            if (pI.empty()) continue;
            const int np = (int)pI.size();
            int *dI = nullptr, *dJ = nullptr;
            double *dxyz = nullptr, *dQ = nullptr, *dout = nullptr;
            PairRef *dref = nullptr;
            PrimPair *dpp = nullptr;
            bool ok = hipMalloc(&dI, sizeof(int) * np) == hipSuccess && hipMalloc(&dJ, sizeof(int) * np) == hipSuccess &&
                      hipMalloc(&dxyz, sizeof(double) * 3 * ns) == hipSuccess && hipMalloc(&dQ, sizeof(double) * 3) == hipSuccess &&
                      hipMalloc(&dout, sizeof(double) * 36 * np) == hipSuccess &&
                      hipMalloc(&dref, sizeof(PairRef) * np) == hipSuccess && hipMalloc(&dpp, sizeof(PrimPair) * pp.size()) == hipSuccess;
            ok = ok && hipMemcpy(dI, pI.data(), sizeof(int) * np, hipMemcpyHostToDevice) == hipSuccess &&
                 hipMemcpy(dJ, pJ.data(), sizeof(int) * np, hipMemcpyHostToDevice) == hipSuccess &&
                 hipMemcpy(dxyz, xyz.data(), sizeof(double) * 3 * ns, hipMemcpyHostToDevice) == hipSuccess &&
                 hipMemcpy(dQ, Q, sizeof(double) * 3, hipMemcpyHostToDevice) == hipSuccess &&
                 hipMemcpy(dref, refs.data(), sizeof(PairRef) * np, hipMemcpyHostToDevice) == hipSuccess &&
                 hipMemcpy(dpp, pp.data(), sizeof(PrimPair) * pp.size(), hipMemcpyHostToDevice) == hipSuccess;
            std::vector<double> out(36 * (size_t)np);
            if (ok) {
                DpmArgs args{dI, dJ, np, nullptr, dxyz, dref, dpp, dQ, dout, rys};
                ok = launch_dpm(la, lb, args, 0) == 0 && hipGetLastError() == hipSuccess &&
                     hipMemcpy(out.data(), dout, sizeof(double) * out.size(), hipMemcpyDeviceToHost) == hipSuccess;
            }
            (void)hipFree(dI);
            (void)hipFree(dJ);
            (void)hipFree(dxyz);
            (void)hipFree(dQ);
            (void)hipFree(dout);
            (void)hipFree(dref);
            (void)hipFree(dpp);
            if (!ok) {
                aux_fail(PATMOL_HIP_EDEVICE, "dpm_ints: device failure");
                return;
            }
            const int nb = (lb + 1) * (lb + 2) / 2, na = (la + 1) * (la + 2) / 2;
            for (int t = 0; t < np; ++t) {
                const Shell &A = sh[pI[t]], &B = sh[pJ[t]];
                for (int ia = 0; ia < na; ++ia)
                    for (int ib = 0; ib < nb; ++ib) {
                        int r = A.off + ia, c = B.off + ib;
                        int hi = r > c ? r : c, lo = r > c ? c : r;
                        if (pI[t] == pJ[t] && c > r) continue;  // same shell: lower triangle once
                        hmat[(size_t)hi * (hi + 1) / 2 + lo] -= out[(size_t)t * 36 + ia * nb + ib];
                    }
            }
        }
}
