// erint_io.hip -- ERINT record interop: the reference's integral-file format
// (erifs.h:19-22: 16-byte records {double val; u16 i, j, k, l}).
//
//   * writer: every canonical integral of the handle's basis with
//     |v| > 1e-14, in the order two_electron_ints writes them
//     (twoints.c:228-379): ish >= start, jsh <= ish, ksh <= ish, lsh <= ksh;
//     per shell quartet the components ir, jr <= ir, kr <= ir, lr <= kr,
//     (ir == kr => lr <= jr).  Integrals come from the device ERI blocks;
//     shell quartets whose Schwarz bound Q_IJ Q_KL is below 1e-15 (ten
//     times under the write threshold) are skipped without evaluation.
//   * reader: the G build of rhf_form_gmatrix / uhf_form_gmatrix over
//     records (scfaux.c:428-714) as a device scatter, one record per thread,
//     FP64 atomics into packed G (into a per-workgroup LDS copy of G when
//     it fits).  Records stream in 64 MiB chunks from host memory, or from a
//     file through two pinned staging buffers so that reading the next chunk
//     overlaps the copy and scatter of the previous one.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <vector>

#include "patmol_hip.h"

extern "C" {
const Shell *shells(void);
int number_of_shells(void);
int number_of_orbitals(void);
}

namespace pm {

int aux_fail(int code, const char *msg);

namespace {

constexpr double kWriteThreshold = 1e-14;   // twoints.c:199
constexpr double kSkipBound = 1e-15;
constexpr int64_t kChunk = 1 << 22;         // records per staging buffer (64 MiB)

__device__ __forceinline__ int64_t tri(int a, int b) { return a > b ? (int64_t)a * (a + 1) / 2 + b : (int64_t)b * (b + 1) / 2 + a; }

// One record of rhf_form_gmatrix / uhf_form_gmatrix (scfaux.c:445-495 and
// 574-640): Coulomb da = v D_ij, db = v D_kl (D = 2P for RHF, PA + PB for
// UHF; db doubled when k != l, da when i != j and ij != kl) into G_ij and
// G_kl, exchange -v P into ik, il, jk, jl under the reference's conditions.
// With PRIV the workgroup accumulates into its own copy of G in LDS
// (no2 doubles per spin) and adds it to global G once at the end: a small G
// would otherwise take every record's atomics on the same few thousand
// addresses.  Without PRIV the atomics go to global G directly.
template <bool UHF, bool PRIV>
__global__ __launch_bounds__(256) void erint_scatter_kernel(const ERINT *__restrict__ rec, int64_t n,
                                                            const double *__restrict__ PA,
                                                            const double *__restrict__ PB, double *gA,
                                                            double *gB, int64_t no2)
{
    extern __shared__ double sG[];
    double *GA = gA, *GB = gB;
    if (PRIV) {
        GA = sG;
        GB = sG + no2;
        for (int64_t e = threadIdx.x; e < (UHF ? 2 : 1) * no2; e += blockDim.x) sG[e] = 0.0;
        __syncthreads();
    }
    // contiguous span of records per workgroup
    const int64_t per = (n + gridDim.x - 1) / gridDim.x, x0 = (int64_t)blockIdx.x * per;
    const int64_t x1 = x0 + per < n ? x0 + per : n;
    for (int64_t x = x0 + threadIdx.x; x < x1; x += blockDim.x) {
        const ERINT r = rec[x];
        const double v = r.val;
        const int i = r.i, j = r.j, k = r.k, l = r.l;
        const int64_t ii = (int64_t)i * (i + 1) / 2, ij = ii + j, ik = ii + k, il = ii + l;
        const int64_t kl = (int64_t)k * (k + 1) / 2 + l, jk = tri(j, k), jl = tri(j, l);
        double da, db;
        if (UHF) {
            da = v * (PA[ij] + PB[ij]);
            db = v * (PA[kl] + PB[kl]);
        } else {
            da = v * 2.0 * PA[ij];
            db = v * 2.0 * PA[kl];
        }
        // exchange pieces per spin: s_jl = v P_ik, s_jk = v P_il, s_ik = v P_jl, s_il = v P_jk
        const double ajl = v * PA[ik], ajk = v * PA[il], aik = v * PA[jl], ail = v * PA[jk];
        double bjl = 0.0, bjk = 0.0, bik = 0.0, bil = 0.0;
        if (UHF) {
            bjl = v * PB[ik];
            bjk = v * PB[il];
            bik = v * PB[jl];
            bil = v * PB[jk];
        }
        auto sub = [&](int64_t at, double a, double b) {
            atomicAdd(&GA[at], -a);
            if (UHF) atomicAdd(&GB[at], -b);
        };
        auto add = [&](int64_t at, double a) {
            atomicAdd(&GA[at], a);
            if (UHF) atomicAdd(&GB[at], a);
        };
        if (k != l) {
            db += db;
            sub(ik, aik, bik);
            if (i != j && j >= k) sub(jk, ajk, bjk);
        }
        sub(il, ail, bil);
        add(ij, db);
        if (i != j && j >= l) sub(jl, ajl, bjl);
        if (ij != kl) {
            if (i != j) da += da;
            if (j <= k) {
                sub(jk, ajk, bjk);
                if (i != j && i <= k) sub(ik, aik, bik);
                if (k != l && j <= l) sub(jl, ajl, bjl);
            }
            add(kl, da);
        }
    }
    if (PRIV) {
        __syncthreads();
        for (int64_t e = threadIdx.x; e < no2; e += blockDim.x) {
            if (sG[e] != 0.0) atomicAdd(&gA[e], sG[e]);
            if (UHF && sG[no2 + e] != 0.0) atomicAdd(&gB[e], sG[no2 + e]);
        }
    }
}

constexpr size_t kPrivBytes = 64 * 1024;   // LDS copy of G used up to this size

void launch_scatter(bool uhf, const ERINT *rec, int64_t m, const double *PA, const double *PB, double *GA,
                    double *GB, int64_t no2, hipStream_t s)
{
    const size_t lds = sizeof(double) * (size_t)no2 * (uhf ? 2 : 1);
    if (lds <= kPrivBytes) {
        // one workgroup per ~64k records, at most 1024
        int64_t g = (m + 65535) / 65536;
        const int grid = (int)(g < 1 ? 1 : (g > 1024 ? 1024 : g));
        if (uhf)
            hipLaunchKernelGGL((erint_scatter_kernel<true, true>), dim3(grid), dim3(256), lds, s, rec, m, PA, PB, GA,
                               GB, no2);
        else
            hipLaunchKernelGGL((erint_scatter_kernel<false, true>), dim3(grid), dim3(256), lds, s, rec, m, PA, PB, GA,
                               GB, no2);
    } else {
        const int grid = (int)((m + 255) / 256 < 8192 ? (m + 255) / 256 : 8192);
        if (uhf)
            hipLaunchKernelGGL((erint_scatter_kernel<true, false>), dim3(grid), dim3(256), 0, s, rec, m, PA, PB, GA,
                               GB, no2);
        else
            hipLaunchKernelGGL((erint_scatter_kernel<false, false>), dim3(grid), dim3(256), 0, s, rec, m, PA, PB, GA,
                               GB, no2);
    }
}

// Streams records into the scatter.  `next(buf, &src)` provides up to kChunk
// records, either read into the pinned buffer buf or left where they are
// (src pointing at them), and returns how many (0 at the end, -1 on error).
template <typename Next>
int scatter_stream(Next &&next, const double *PA, const double *PB, double *GA, double *GB, const char *who)
{
    const int n = number_of_orbitals();
    char msg[200];
    if (n <= 0 || n > 65535 || !PA || !GA || (PB && !GB) || (!PB && GB)) {
        snprintf(msg, sizeof msg, "%s: bad argument or no basis", who);
        return aux_fail(PATMOL_HIP_EBADARG, msg);
    }
    const bool uhf = PB != nullptr;
    const size_t no2 = (size_t)n * (n + 1) / 2;
    ERINT *host[2] = {nullptr, nullptr};
    ERINT *dev[2] = {nullptr, nullptr};
    double *dP = nullptr, *dG = nullptr;
    hipStream_t s = nullptr;
    hipEvent_t done[2] = {nullptr, nullptr};
    int rc = 0;
    auto cleanup = [&]() {
        if (s) (void)hipStreamSynchronize(s);
        for (int b = 0; b < 2; ++b) {
            if (host[b]) (void)hipHostFree(host[b]);
            if (dev[b]) (void)hipFree(dev[b]);
            if (done[b]) (void)hipEventDestroy(done[b]);
        }
        if (dP) (void)hipFree(dP);
        if (dG) (void)hipFree(dG);
        if (s) (void)hipStreamDestroy(s);
    };
    bool ok = hipStreamCreate(&s) == hipSuccess;
    for (int b = 0; ok && b < 2; ++b)
        ok = hipHostMalloc((void **)&host[b], sizeof(ERINT) * kChunk, hipHostMallocDefault) == hipSuccess &&
             hipMalloc(&dev[b], sizeof(ERINT) * kChunk) == hipSuccess && hipEventCreate(&done[b]) == hipSuccess;
    ok = ok && hipMalloc(&dP, sizeof(double) * no2 * (uhf ? 2 : 1)) == hipSuccess &&
         hipMalloc(&dG, sizeof(double) * no2 * (uhf ? 2 : 1)) == hipSuccess;
    ok = ok && hipMemcpyAsync(dP, PA, sizeof(double) * no2, hipMemcpyHostToDevice, s) == hipSuccess &&
         (!uhf || hipMemcpyAsync(dP + no2, PB, sizeof(double) * no2, hipMemcpyHostToDevice, s) == hipSuccess) &&
         hipMemsetAsync(dG, 0, sizeof(double) * no2 * (uhf ? 2 : 1), s) == hipSuccess;
    if (!ok) {
        cleanup();
        snprintf(msg, sizeof msg, "%s: device allocation failed", who);
        return aux_fail(PATMOL_HIP_EDEVICE, msg);
    }
    // buffer b is refilled only after the scatter that read it has finished
    bool pending[2] = {false, false};
    for (int64_t chunk = 0;; ++chunk) {
        const int b = (int)(chunk & 1);
        if (pending[b] && hipEventSynchronize(done[b]) != hipSuccess) {
            rc = -2;
            break;
        }
        const ERINT *src = host[b];
        const int64_t m = next(host[b], &src);
        if (m < 0) {
            rc = -1;
            break;
        }
        if (m == 0) break;
        if (hipMemcpyAsync(dev[b], src, sizeof(ERINT) * m, hipMemcpyHostToDevice, s) != hipSuccess) {
            rc = -2;
            break;
        }
        launch_scatter(uhf, dev[b], m, dP, uhf ? dP + no2 : nullptr, dG, uhf ? dG + no2 : nullptr, (int64_t)no2, s);
        if (hipGetLastError() != hipSuccess || hipEventRecord(done[b], s) != hipSuccess) {
            rc = -2;
            break;
        }
        pending[b] = true;
    }
    std::vector<double> g(no2 * (uhf ? 2 : 1));
    if (rc == 0 && (hipMemcpyAsync(g.data(), dG, sizeof(double) * g.size(), hipMemcpyDeviceToHost, s) != hipSuccess ||
                    hipStreamSynchronize(s) != hipSuccess))
        rc = -2;
    cleanup();
    if (rc == -1) {
        snprintf(msg, sizeof msg, "%s: record source failed (read error or index beyond the basis)", who);
        return aux_fail(PATMOL_HIP_EBADARG, msg);
    }
    if (rc) {
        snprintf(msg, sizeof msg, "%s: device failure", who);
        return aux_fail(PATMOL_HIP_EDEVICE, msg);
    }
    for (size_t x = 0; x < no2; ++x) GA[x] += g[x];
    if (uhf)
        for (size_t x = 0; x < no2; ++x) GB[x] += g[no2 + x];
    return 0;
}

// records are checked on the host before they reach the device: every index
// must lie inside the basis
bool records_in_range(const ERINT *r, int64_t m, int n)
{
    for (int64_t x = 0; x < m; ++x)
        if (r[x].i >= n || r[x].j >= n || r[x].k >= n || r[x].l >= n) return false;
    return true;
}

int gmatrix_records(const ERINT *recs, int64_t nrec, const double *PA, const double *PB, double *GA, double *GB,
                    const char *who)
{
    if (nrec < 0 || (nrec > 0 && !recs)) return aux_fail(PATMOL_HIP_EBADARG, "ERINT records: bad argument");
    const int n = number_of_orbitals();
    int64_t at = 0;
    auto next = [&](ERINT *, const ERINT **src) -> int64_t {
        const int64_t m = nrec - at < kChunk ? nrec - at : kChunk;
        if (m <= 0) return 0;
        *src = recs + at;
        at += m;
        return records_in_range(*src, m, n) ? m : -1;
    };
    return scatter_stream(next, PA, PB, GA, GB, who);
}

int gmatrix_file(const char *path, const double *PA, const double *PB, double *GA, double *GB, const char *who)
{
    if (!path) return aux_fail(PATMOL_HIP_EBADARG, "ERINT file: null path");
    FILE *f = fopen(path, "rb");
    if (!f) {
        char msg[200];
        snprintf(msg, sizeof msg, "%s: cannot open %s", who, path);
        return aux_fail(PATMOL_HIP_EBADARG, msg);
    }
    const int n = number_of_orbitals();
    auto next = [&](ERINT *buf, const ERINT **) -> int64_t {
        const size_t m = fread(buf, sizeof(ERINT), (size_t)kChunk, f);
        if (m == 0) return ferror(f) ? -1 : 0;
        return records_in_range(buf, (int64_t)m, n) ? (int64_t)m : -1;
    };
    const int rc = scatter_stream(next, PA, PB, GA, GB, who);
    fclose(f);
    return rc;
}

}  // namespace
}  // namespace pm

using namespace pm;

extern "C" int patmol_hip_rhf_gmatrix_erints(const ERINT *recs, int64_t nrec, const double *Pmat, double *Gmat)
{
    return gmatrix_records(recs, nrec, Pmat, nullptr, Gmat, nullptr, "patmol_hip_rhf_gmatrix_erints");
}

extern "C" int patmol_hip_uhf_gmatrix_erints(const ERINT *recs, int64_t nrec, const double *PmatA,
                                             const double *PmatB, double *GmatA, double *GmatB)
{
    if (!PmatB || !GmatB) return aux_fail(PATMOL_HIP_EBADARG, "patmol_hip_uhf_gmatrix_erints: null matrix");
    return gmatrix_records(recs, nrec, PmatA, PmatB, GmatA, GmatB, "patmol_hip_uhf_gmatrix_erints");
}

extern "C" int patmol_hip_rhf_gmatrix_erint_file(const char *path, const double *Pmat, double *Gmat)
{
    return gmatrix_file(path, Pmat, nullptr, Gmat, nullptr, "patmol_hip_rhf_gmatrix_erint_file");
}

extern "C" int patmol_hip_uhf_gmatrix_erint_file(const char *path, const double *PmatA, const double *PmatB,
                                                 double *GmatA, double *GmatB)
{
    if (!PmatB || !GmatB) return aux_fail(PATMOL_HIP_EBADARG, "patmol_hip_uhf_gmatrix_erint_file: null matrix");
    return gmatrix_file(path, PmatA, PmatB, GmatA, GmatB, "patmol_hip_uhf_gmatrix_erint_file");
}

extern "C" int64_t patmol_hip_write_erints(const ERIFS *e, int start, const char *path)
{
    const Shell *sh = shells();
    const int ns = number_of_shells(), n = number_of_orbitals();
    if (!e || !path || !sh || ns <= 0 || n <= 0 || start < 0)
        return aux_fail(PATMOL_HIP_EBADARG, "patmol_hip_write_erints: bad argument or no basis"), -1;
    if (n > 65535) return aux_fail(PATMOL_HIP_EUNSUPPORTED, "patmol_hip_write_erints: more than 65535 functions"), -1;
    std::vector<double> q((size_t)ns * (ns + 1) / 2);
    if (patmol_hip_pair_schwarz(e, q.data())) return -1;
    auto Q = [&](int a, int b) { return a >= b ? q[(size_t)a * (a + 1) / 2 + b] : q[(size_t)b * (b + 1) / 2 + a]; };
    FILE *f = fopen(path, "wb");
    if (!f) return aux_fail(PATMOL_HIP_EBADARG, "patmol_hip_write_erints: cannot open the output file"), -1;
    int64_t written = 0;
    constexpr size_t kMaxVals = (size_t)1 << 24;  // doubles of ERI blocks per device batch
    std::vector<int> tuples;
    std::vector<ERINT> out;
    std::vector<double> vals;
    bool ok = true;
    // canonical component loops of one shell quartet (twoints.c:318-359)
    auto emit = [&](const int *t, const double *blk) {
        const Shell &A = sh[t[0]], &B = sh[t[1]], &C = sh[t[2]], &D = sh[t[3]];
        for (int a = 0; a < A.nst; ++a) {
            const int ir = A.off + a;
            for (int b = 0; b < B.nst; ++b) {
                const int jr = B.off + b;
                if (jr > ir) break;
                for (int c = 0; c < C.nst; ++c) {
                    const int kr = C.off + c;
                    if (kr > ir) break;
                    for (int d = 0; d < D.nst; ++d) {
                        const int lr = D.off + d;
                        if (lr > kr) break;
                        if (ir == kr && lr > jr) break;
                        const double v = blk[((a * B.nst + b) * C.nst + c) * D.nst + d];
                        if (fabs(v) > kWriteThreshold) {
                            ERINT r;
                            r.val = v;
                            r.i = (unsigned short)ir;
                            r.j = (unsigned short)jr;
                            r.k = (unsigned short)kr;
                            r.l = (unsigned short)lr;
                            out.push_back(r);
                        }
                    }
                }
            }
        }
    };
    auto flush = [&]() {
        const int nq = (int)(tuples.size() / 4);
        if (nq == 0) return true;
        size_t total = 0;
        for (int t = 0; t < nq; ++t)
            total += (size_t)sh[tuples[4 * t]].nst * sh[tuples[4 * t + 1]].nst * sh[tuples[4 * t + 2]].nst *
                     sh[tuples[4 * t + 3]].nst;
        vals.resize(total);
        if (patmol_hip_eri_blocks(e, nq, tuples.data(), vals.data())) return false;
        size_t o = 0;
        out.clear();
        for (int t = 0; t < nq; ++t) {
            const int *tp = &tuples[4 * t];
            emit(tp, vals.data() + o);
            o += (size_t)sh[tp[0]].nst * sh[tp[1]].nst * sh[tp[2]].nst * sh[tp[3]].nst;
        }
        if (!out.empty() && fwrite(out.data(), sizeof(ERINT), out.size(), f) != out.size()) return false;
        written += (int64_t)out.size();
        tuples.clear();
        return true;
    };
    size_t pending_vals = 0;
    for (int is = start; ok && is < ns; ++is)
        for (int js = 0; ok && js <= is; ++js)
            for (int ks = 0; ok && ks <= is; ++ks)
                for (int ls = 0; ok && ls <= ks; ++ls) {
                    // |v| <= Q_IJ Q_KL: nothing of this quartet can pass
                    // the write threshold
                    if (Q(is, js) * Q(ks, ls) < kSkipBound) continue;
                    tuples.insert(tuples.end(), {is, js, ks, ls});
                    pending_vals += (size_t)sh[is].nst * sh[js].nst * sh[ks].nst * sh[ls].nst;
                    if (pending_vals >= kMaxVals) {
                        ok = flush();
                        pending_vals = 0;
                    }
                }
    ok = ok && flush();
    const bool closed = fclose(f) == 0;
    if (!ok || !closed) {
        if (!patmol_hip_last_error()) aux_fail(PATMOL_HIP_EDEVICE, "patmol_hip_write_erints: write failed");
        return -1;
    }
    return written;
}
