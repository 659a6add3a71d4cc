// hip_context.cpp -- host side of libpatmol_hip.so: the reference's ERIFS
// boundary (init_ERIFS, two_electron_ints, rhf/uhf_form_gmatrix,
// clean_ERIFS) over device-resident pair tables and screened tile-quartet
// work lists, plus the extension entry points declared in patmol_hip.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <atomic>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <numeric>
#include <unordered_map>
#include <vector>

#define RYS_STORAGE static const
#include "rys_table.inc"
static constexpr int kTableNC = RYS_NC;
#undef RYS_NC
#undef RYS_NMAX

#include "batch.h"
#include "eri_kernels.h"
#include "grad.h"
#include "patmol_hip.h"

extern "C" {
// basis.h accessors (reference basis.c, or libpatmol_host.so)
const Shell *shells(void);
const Center *centers(void);
int number_of_shells(void);
int number_of_orbitals(void);
int number_of_centers(void);
int skip_center(void);
}

namespace pm {

// ---------------------------------------------------------------------------
// status
// ---------------------------------------------------------------------------
static thread_local int g_err = PATMOL_HIP_OK;
static thread_local char g_msg[512] = "";
static int g_strict = -1;

static bool strict_mode()
{
    if (g_strict < 0) {
        const char *e = getenv("PATMOL_HIP_STRICT");
        g_strict = (e && atoi(e)) ? 1 : 0;
    }
    return g_strict == 1;
}

static int fail(int code, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_msg, sizeof g_msg, fmt, ap);
    va_end(ap);
    g_err = code;
    if (strict_mode()) {
        fprintf(stderr, "FATAL ERROR: %s\n", g_msg);
        abort();
    }
    return code;
}

#define HIPCHK(x)                                                                         \
    do {                                                                                  \
        hipError_t e_ = (x);                                                              \
        if (e_ != hipSuccess) return fail(PATMOL_HIP_EDEVICE, "%s: %s", #x, hipGetErrorString(e_)); \
    } while (0)

// ---------------------------------------------------------------------------
// options
// ---------------------------------------------------------------------------
static double g_tau = 1e-12;
static int g_prec_mode = 0;
static double g_prec_thr = 1e-6;
// 0: items below the threshold run in FP32 only in the classes whose FP32
// kernel was measured faster (jk3_fp32_pays); 1: in every class (tuning)
static int g_fp32_policy = 0;
// classes whose FP64 items run on the McMurchie-Davidson kernel
// (md_kernel.h), a mask: bit 0 (dd|dd), bit 1 (dd|dp), bit 2 (dp|dp),
// bit 3 (dd|pp); the others run their Rys kernels.  The default holds the
// classes where it measured faster than the Rys kernel on (H2O)16.
static const int MD_DEFAULT_MASK = 1 | 2;
static int g_md_mask = MD_DEFAULT_MASK;
static int md_class_bit(int bra_pc, int ket_pc)
{
    if (bra_pc == 5 && ket_pc == 5) return 1;
    if (bra_pc == 5 && ket_pc == 4) return 2;
    if (bra_pc == 4 && ket_pc == 4) return 4;
    if (bra_pc == 5 && ket_pc == 2) return 8;
    return 0;
}
// largest Hermite-coefficient table set built per basis; above it the
// classes stay on their Rys kernels
static const size_t MD_TABLE_MAX_BYTES = (size_t)8 << 30;
// (dd|dp) runs on the McMurchie-Davidson kernel only when at most this share
// of the shell-quartet slots of its work items survives the Schwarz screen:
// that kernel packs the kept ket pairs of a bra pair into its lanes, the Rys
// kernel keeps a lane per slot.  Measured (kept share; MD against Rys time of
// the class): (H2O)16 0.25, 8.9 against 9.7 ms; (H2O)8 0.53, 1.24 against
// 1.05 ms; the 64-molecule batch 0.93, 124 ms slower in all.
static const double MD_DDDP_MAX_KEPT = 0.4;
// engine-mask bit that keeps every class of the mask on the McMurchie-Davidson
// kernel whatever its kept share (tests of that kernel)
static const int MD_FORCE = 16;
// the mask bits in force for a basis whose rule switched off `off`
static int md_effective(int off) { return g_md_mask & 15 & ((g_md_mask & MD_FORCE) ? 15 : ~off); }

// classes whose FP64 items on the McMurchie-Davidson kernel took less time
// than their FP32 Rys kernel on (H2O)16 ((dd|dd): 2.7 against 5.3 ms for
// every item of the build): with that engine on, mixed precision keeps their
// items in FP64
static const int MD_BEATS_FP32_MASK = 1;

// precision of the items a class would run in FP32: 1 where FP32 pays
static int fp32_prec(int bra_pc, int ket_pc)
{
    if (g_fp32_policy == 1) return 1;
    if (g_md_mask & MD_BEATS_FP32_MASK & md_class_bit(bra_pc, ket_pc)) return 0;
    const int k = jk3_class_index(pair_class_la(bra_pc), pair_class_lb(bra_pc), pair_class_la(ket_pc),
                                  pair_class_lb(ket_pc));
    return jk3_fp32_pays(k) ? 1 : 0;
}
static bool g_timing = false;

// ---------------------------------------------------------------------------
// device Rys tables (one copy per process)
// ---------------------------------------------------------------------------
static RysTables g_rys = {nullptr, nullptr, nullptr, nullptr};
static std::mutex g_mu;

static int ensure_tables()
{
    if (g_rys.cheb) return 0;
    double *c = nullptr, *hr = nullptr, *hh = nullptr;
    float *c32 = nullptr;
    // The device evaluates each fit by Horner's rule in u, so the Chebyshev
    // coefficients of every interval and series are converted to monomial
    // coefficients here (in long double; evaluated in double the monomial
    // form is as accurate as Clenshaw's recurrence on these fits, 1.3e-14
    // relative at worst, with half the operations).  The FP32 table holds
    // the first 9 Chebyshev terms, converted the same way.
    const size_t ntab = (size_t)pm::rys_table_doubles();
    std::vector<double> mono(ntab, 0.0);
    std::vector<float> f(ntab, 0.0f);
    {
        constexpr int NC = kTableNC, NC32 = 9;
        long double tp[NC][NC] = {};  // tp[j][i]: coefficient of u^i in T_j
        tp[0][0] = 1.0L;
        tp[1][1] = 1.0L;
        for (int j = 1; j + 1 < NC; ++j)
            for (int i = 0; i < NC; ++i)
                tp[j + 1][i] = (i ? 2.0L * tp[j][i - 1] : 0.0L) - tp[j - 1][i];
        for (int n = 1; n <= RYS_NMAX; ++n) {
            const int nser = 2 * n;
            for (int k = 0; k < pm::rys_tmax(n); ++k) {
                const int base = pm::rys_src_off(n) + k * NC * nser;     // rys_table.inc
                const int dst = pm::rys_off(n) + k * pm::rys_stride(n);  // padded device layout
                for (int v = 0; v < nser; ++v) {
                    long double a[NC] = {}, a32[NC] = {};
                    for (int j = 0; j < NC; ++j)
                        for (int i = 0; i <= j; ++i) {
                            a[i] += (long double)rys_cheb[base + j * nser + v] * tp[j][i];
                            if (j < NC32) a32[i] += (long double)rys_cheb[base + j * nser + v] * tp[j][i];
                        }
                    for (int i = 0; i < NC; ++i) {
                        mono[dst + i * nser + v] = (double)a[i];
                        f[dst + i * nser + v] = (float)a32[i];
                    }
                }
            }
        }
    }
    HIPCHK(hipMalloc(&c, sizeof(double) * ntab));
    HIPCHK(hipMalloc(&c32, sizeof(float) * ntab));
    HIPCHK(hipMalloc(&hr, sizeof(rys_herm_r)));
    HIPCHK(hipMalloc(&hh, sizeof(rys_herm_h)));
    HIPCHK(hipMemcpy(c, mono.data(), sizeof(double) * ntab, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(c32, f.data(), sizeof(float) * ntab, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(hr, rys_herm_r, sizeof(rys_herm_r), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(hh, rys_herm_h, sizeof(rys_herm_h), hipMemcpyHostToDevice));
    g_rys = {c, c32, hr, hh};
    static_assert(kTableNC == pm::RYS_NC, "table layout");
    static_assert(pm::rys_src_off(pm::RYS_NMAX + 1) == RYS_TABLE_SIZE, "table layout");
    return 0;
}

// ---------------------------------------------------------------------------
// per-handle context
// ---------------------------------------------------------------------------
template <typename T>
struct DevBuf {
    T *p = nullptr;
    size_t n = 0;
    int alloc(size_t cnt)
    {
        if (cnt <= n && p) return 0;
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
        if (cnt == 0) cnt = 1;
        if (hipMalloc(&p, sizeof(T) * cnt) != hipSuccess) return fail(PATMOL_HIP_EDEVICE, "hipMalloc of %zu bytes failed", sizeof(T) * cnt);
        n = cnt;
        return 0;
    }
    int upload(const std::vector<T> &v)
    {
        if (alloc(v.size())) return g_err;
        if (!v.empty()) HIPCHK(hipMemcpy(p, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice));
        return 0;
    }
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
};

struct ClassRange {
    int bra_pc, ket_pc;
    int first, count;
};

// Hermite-coefficient tables of the McMurchie-Davidson J/K path: slot of
// each primitive pair in its class's tables, bra / ket tables per pair class
struct MdTables {
    DevBuf<int> d_slot;
    DevBuf<double> d_bra[PM_NPC], d_ket[PM_NPC];
    bool ready = false;
    void release()
    {
        d_slot.release();
        for (int pc = 0; pc < PM_NPC; ++pc) {
            d_bra[pc].release();
            d_ket[pc].release();
        }
        ready = false;
    }
    // the J/K arguments of class (bra_pc | ket_pc) under mask `mask`
    void fill(JKArgs &a, int mask_bit) const
    {
        if (!ready || !mask_bit) return;
        a.md_slot = d_slot.p;
        for (int pc = 0; pc < PM_NPC; ++pc) {
            a.md_bra[pc] = d_bra[pc].p;
            a.md_ket[pc] = d_ket[pc].p;
        }
    }
};

struct Ctx {
    bool built = false;
    int N = 0, nshell = 0, start = 0;
    // engine-mask bits switched off for this basis (MD_DDDP_MAX_KEPT)
    int md_off = 0;
    // kept shell quartets and slots of a sample of the (dd|dp) items
    int64_t ddp_kept = 0, ddp_slots = 0;
    double tau = 1e-12;
    // host tables
    std::vector<int> sorted_orig, orig_to_sorted, shell_l, shell_nst, shell_off;
    // functions renumbered in sorted shell order (the J/K path's layout of D,
    // J' and K': the shells of a tile own one contiguous range of functions)
    std::vector<int> shell_soff, fn_sorted_to_orig, fn_orig_to_sorted;
    std::vector<double> shell_xyz;
    std::vector<int> tile_first, tile_size, tile_l;
    std::vector<PairRef> pairs;
    std::vector<PrimPair> pp;
    std::vector<double2> ppab;   // exponents of each primitive pair (I's, J's)
    std::vector<double> pair_q;
    std::vector<int4> items;
    std::vector<double> item_bound;
    std::vector<ClassRange> ranges;
    // device tables
    DevBuf<double> d_xyz, d_q, d_bound;
    DevBuf<int> d_off, d_soff, d_fperm, d_finv, d_orig, d_tile_first, d_tile_size;
    DevBuf<PairRef> d_pairs;
    DevBuf<PrimPair> d_pp;
    DevBuf<double2> d_ppab;    // primitive exponents of each primitive pair (gradient)
    DevBuf<int4> d_items;
    // McMurchie-Davidson tables: slot of each primitive pair in its class's
    // tables, bra / ket Hermite-coefficient tables per pair class
    MdTables md;
    // work buffers
    DevBuf<double> d_P, d_PB, d_D, d_D2, d_DJ, d_G, d_G2, d_dmax;
    DevBuf<FxAcc> d_J, d_K;   // PM_XCD_SLOTS fixed-point copies of J', K' (N x N each)
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    std::vector<hipEvent_t> cls_ev;
    // side streams for concurrent class launches (default priority)
    std::vector<hipStream_t> side;
    std::vector<hipEvent_t> side_ev;
    hipEvent_t fork_ev = nullptr;
    double last_ms = 0.0;
    double class_ms[36] = {0};
    int64_t items_fp32 = 0;
    int64_t items_skipped = 0;
    double last_dmax = 0.0;
    // per-item precision / screening by density-weighted bounds: tile
    // function ranges, tile-pair max|D|, each item's class range, the
    // classified lists (per class range: FP64 items from the front, FP32
    // items from the back) and their counts
    DevBuf<int> d_tfun0, d_tfunn, d_rfirst, d_rcount, d_cnt;
    DevBuf<unsigned char> d_item_rng;
    DevBuf<double> d_tmax;
    DevBuf<int4> d_cls_items;
    std::vector<int> cnt_host;
    bool cls_ready = false;   // tables above match the current items (reset by every table build)
    int64_t prim_kept = 0;
};

static std::unordered_map<const ERIFS *, Ctx *> g_ctx;

static Ctx *ctx_of(const ERIFS *e)
{
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_ctx.find(e);
    return it == g_ctx.end() ? nullptr : it->second;
}

static inline int pair_class(int la, int lb) { return la * (la + 1) / 2 + lb; }

// Hermite-coefficient tables of the McMurchie-Davidson J/K path: for every
// primitive pair of the pair classes pp, dp and dd, its slot, and the bra
// (dd, dp) and ket (dd, dp, pp) tables in the lane order md_kernel.h reads.
// plist: (pair index, sorted shell I, sorted shell J) of every shell pair,
// shell_l: l of each sorted shell; the device arrays hold the same pairs.
struct MdPair {
    int64_t p;
    int I, J;
};
static int build_md_tables(MdTables &t, const std::vector<MdPair> &plist, const std::vector<int> &shell_l,
                           const std::vector<PairRef> &pairs, size_t npp, const double *d_xyz, const PairRef *d_pairs,
                           const PrimPair *d_pp)
{
    t.release();
    std::vector<int> slot(npp, -1);
    size_t bytes = 0;
    for (int pc : {2, 4, 5}) {
        const int la = pair_class_la(pc), lb = pair_class_lb(pc);
        const bool bra = md_bra_class(la, lb), ket = md_ket_class(la, lb);
        std::vector<int> pl, pi, pj, s0;
        int nslot = 0;
        for (const MdPair &x : plist)
            if (shell_l[x.I] == la && shell_l[x.J] == lb) {
                const PairRef &r = pairs[x.p];
                if (x.p > INT32_MAX) return fail(PATMOL_HIP_EUNSUPPORTED, "too many shell pairs");
                pl.push_back((int)x.p);
                pi.push_back(x.I);
                pj.push_back(x.J);
                s0.push_back(nslot);
                for (int k = 0; k < r.npp; ++k) slot[r.pp0 + k] = nslot + k;
                nslot += r.npp;
            }
        if (pl.empty()) continue;
        const size_t nb = bra ? (size_t)nslot * md_bra_regs(la, lb) * 64 : 0;
        const size_t nk = ket ? (size_t)nslot * md_ket_regs(la, lb) * 64 : 0;
        bytes += 8 * (nb + nk);
        if (bytes > MD_TABLE_MAX_BYTES) {
            t.release();
            return 0;   // the classes keep their Rys kernels
        }
        if ((bra && t.d_bra[pc].alloc(nb)) || (ket && t.d_ket[pc].alloc(nk))) return g_err;
        DevBuf<int> dl, di, dj, ds;
        if (dl.upload(pl) || di.upload(pi) || dj.upload(pj) || ds.upload(s0)) return g_err;
        MdTabArgs ta{dl.p, di.p, dj.p, ds.p, (int)pl.size(), d_xyz, d_pairs, d_pp, bra ? t.d_bra[pc].p : nullptr,
                     ket ? t.d_ket[pc].p : nullptr};
        if (launch_md_etab(la, lb, ta, 0)) return fail(PATMOL_HIP_EUNSUPPORTED, "no Hermite table kernel");
        HIPCHK(hipGetLastError());
        HIPCHK(hipDeviceSynchronize());
        dl.release();
        di.release();
        dj.release();
        ds.release();
    }
    if (t.d_slot.upload(slot)) return g_err;
    t.ready = true;
    return 0;
}

static int build_md_tables(Ctx &c)
{
    std::vector<MdPair> plist;
    plist.reserve((size_t)c.nshell * (c.nshell + 1) / 2);
    for (int I = 0; I < c.nshell; ++I)
        for (int J = 0; J <= I; ++J) plist.push_back({(int64_t)I * (I + 1) / 2 + J, I, J});
    return build_md_tables(c.md, plist, c.shell_l, c.pairs, c.pp.size(), c.d_xyz.p, c.d_pairs.p, c.d_pp.p);
}

// ---------------------------------------------------------------------------
// build: shells, pairs, Schwarz, tiles, work items
// ---------------------------------------------------------------------------
// md: also the McMurchie-Davidson tables (a batch builds its own)
// Kept shell quartets and shell-quartet slots (canonical order inside equal
// tiles) of up to 512 evenly spaced work items of range r: the kept share of
// a class's slots, the same rule as the kernels' validity test
static void sample_kept(const Ctx &c, const ClassRange &r, int64_t &kept, int64_t &slots)
{
    const int stride = std::max(1, r.count / 512);
    for (int it = r.first; it < r.first + r.count; it += stride) {
        const int4 t = c.items[it];
        const bool sameBra = t.x == t.y, sameKet = t.z == t.w, samePair = t.x == t.z && t.y == t.w;
        for (int i = 0; i < c.tile_size[t.x]; ++i) {
            const int I = c.tile_first[t.x] + i;
            for (int j = 0; j < c.tile_size[t.y]; ++j) {
                const int J = c.tile_first[t.y] + j;
                if (sameBra && J > I) continue;
                const double qb = c.pair_q[(int64_t)I * (I + 1) / 2 + J];
                for (int kq = 0; kq < c.tile_size[t.z]; ++kq) {
                    const int K = c.tile_first[t.z] + kq;
                    for (int lq = 0; lq < c.tile_size[t.w]; ++lq) {
                        const int L = c.tile_first[t.w] + lq;
                        if (sameKet && L > K) continue;
                        if (samePair && !(I > K || (I == K && J >= L))) continue;
                        ++slots;
                        if (qb * c.pair_q[(int64_t)K * (K + 1) / 2 + L] < c.tau) continue;
                        if (c.start > 0 && std::max(std::max(c.sorted_orig[I], c.sorted_orig[J]),
                                                    std::max(c.sorted_orig[K], c.sorted_orig[L])) < c.start)
                            continue;
                        ++kept;
                    }
                }
            }
        }
    }
}

static int build_tables(Ctx &c, int start, bool md = true)
{
    const Shell *sh = shells();
    const Center *ce = centers();
    const int ns = number_of_shells();
    const int N = number_of_orbitals();
    if (!sh || !ce || ns <= 0) return fail(PATMOL_HIP_ESTATE, "no basis loaded");
    if (N > 65535) return fail(PATMOL_HIP_EUNSUPPORTED, "more than 65535 functions");
    for (int s = 0; s < ns; ++s)
        if (sh[s].lsh > PM_LMAX || sh[s].lsh < 0)
            return fail(PATMOL_HIP_EUNSUPPORTED, "shell %d has l=%d (supported: 0..%d)", s, sh[s].lsh, PM_LMAX);
    c.N = N;
    c.nshell = ns;
    c.start = start;
    c.tau = g_tau;
    // sort shells (tile order)
    c.sorted_orig.resize(ns);
    std::iota(c.sorted_orig.begin(), c.sorted_orig.end(), 0);
    // Tiles group shells of one l with similar extent and position: sorted
    // by l, then by exponent band (floor of log_9 of the most diffuse
    // exponent, descending), then along a Morton curve through the centre
    // positions (10 bits per axis over the centres' bounding box), then by
    // centre, most diffuse exponent (descending) and index.  Shell pairs of a
    // tile pair then have similar Schwarz bounds, so a kept tile quartet has
    // few screened-out lanes: 65% of the lanes of the executed steps are busy
    // on (H2O)8, against 62% with exponent-major and 55% with centre-major
    // tiles.
    std::vector<double> amin(ns);
    std::vector<int> band(ns);
    std::vector<uint32_t> zc(ns);
    double lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
    for (int s = 0; s < ns; ++s)
        for (int x = 0; x < 3; ++x) {
            const double r = ce[sh[s].cen].r[x];
            lo[x] = s == 0 ? r : std::min(lo[x], r);
            hi[x] = s == 0 ? r : std::max(hi[x], r);
        }
    for (int s = 0; s < ns; ++s) {
        amin[s] = sh[s].al[0];
        for (int p = 1; p < sh[s].npr; ++p) amin[s] = std::min(amin[s], sh[s].al[p]);
        band[s] = (int)std::floor(std::log(amin[s]) / std::log(9.0) + 1e-9);
        uint32_t code = 0;
        for (int x = 0; x < 3; ++x) {
            const double t = hi[x] > lo[x] ? (ce[sh[s].cen].r[x] - lo[x]) / (hi[x] - lo[x]) : 0.0;
            const uint32_t g = (uint32_t)(int)(t * 1023.0);
            for (int b = 0; b < 10; ++b) code |= ((g >> b) & 1u) << (3 * b + x);
        }
        zc[s] = code;
    }
    std::stable_sort(c.sorted_orig.begin(), c.sorted_orig.end(), [&](int a, int b) {
        if (sh[a].lsh != sh[b].lsh) return sh[a].lsh < sh[b].lsh;
        if (band[a] != band[b]) return band[a] > band[b];
        if (zc[a] != zc[b]) return zc[a] < zc[b];
        if (sh[a].cen != sh[b].cen) return sh[a].cen < sh[b].cen;
        if (amin[a] != amin[b]) return amin[a] > amin[b];
        return a < b;
    });
    c.orig_to_sorted.assign(ns, 0);
    c.shell_l.resize(ns);
    c.shell_nst.resize(ns);
    c.shell_off.resize(ns);
    c.shell_xyz.resize(3 * ns);
    c.shell_soff.resize(ns);
    c.fn_sorted_to_orig.assign(N, 0);
    c.fn_orig_to_sorted.assign(N, 0);
    for (int i = 0, f = 0; i < ns; ++i) {
        const Shell &s = sh[c.sorted_orig[i]];
        c.orig_to_sorted[c.sorted_orig[i]] = i;
        c.shell_l[i] = s.lsh;
        c.shell_nst[i] = (s.lsh + 1) * (s.lsh + 2) / 2;
        c.shell_off[i] = s.off;
        c.shell_soff[i] = f;
        if (s.off < 0 || s.off + c.shell_nst[i] > N || f + c.shell_nst[i] > N)
            return fail(PATMOL_HIP_ESTATE, "shell %d: functions [%d, %d) outside the basis of %d", c.sorted_orig[i], s.off,
                        s.off + c.shell_nst[i], N);
        for (int k = 0; k < c.shell_nst[i]; ++k, ++f) {
            c.fn_sorted_to_orig[f] = s.off + k;
            c.fn_orig_to_sorted[s.off + k] = f;
        }
        for (int x = 0; x < 3; ++x) c.shell_xyz[3 * i + x] = ce[s.cen].r[x];
    }
    {
        // the renumbering must be a permutation of the basis functions
        std::vector<char> seen(N, 0);
        int nseen = 0;
        for (int f = 0; f < N; ++f) nseen += seen[c.fn_sorted_to_orig[f]]++ == 0;
        if (nseen != N) return fail(PATMOL_HIP_ESTATE, "shell function ranges do not cover the %d basis functions", N);
    }
    // pairs I >= J (sorted) and their primitive pairs
    const int64_t np = (int64_t)ns * (ns + 1) / 2;
    c.pairs.assign(np, PairRef{0, 0});
    c.pp.clear();
    c.ppab.clear();
    std::vector<double2> &ppab = c.ppab;
    for (int I = 0; I < ns; ++I) {
        const Shell &a = sh[c.sorted_orig[I]];
        const double *A = ce[a.cen].r;
        for (int J = 0; J <= I; ++J) {
            const Shell &b = sh[c.sorted_orig[J]];
            const double *B = ce[b.cen].r;
            const bool samec = a.cen == b.cen;
            double ab2 = 0;
            for (int x = 0; x < 3; ++x) ab2 += (A[x] - B[x]) * (A[x] - B[x]);
            PairRef &r = c.pairs[(int64_t)I * (I + 1) / 2 + J];
            r.pp0 = (int)c.pp.size();
            r.npp = a.npr * b.npr;
            for (int i = 0; i < a.npr; ++i)
                for (int j = 0; j < b.npr; ++j) {
                    PrimPair q;
                    double al = a.al[i], be = b.al[j], p = al + be;
                    q.p = p;
                    if (samec) {
                        // same centre: product at A, unit overlap factor (twoints.c:48-57)
                        for (int x = 0; x < 3; ++x) q.P[x] = A[x];
                        q.s = 1.0;
                    } else {
                        for (int x = 0; x < 3; ++x) q.P[x] = (al * A[x] + be * B[x]) / p;
                        q.s = exp(-al * be * ab2 / p);
                    }
                    q.c = a.co[i] * b.co[j];
                    q.ip = 1.0 / p;
                    c.pp.push_back(q);
                    ppab.push_back(make_double2(al, be));
                }
        }
    }
    // tiles per l
    c.tile_first.clear();
    c.tile_size.clear();
    c.tile_l.clear();
    for (int i = 0; i < ns;) {
        int l = c.shell_l[i], T = tile_of(l), n = 0;
        while (i + n < ns && c.shell_l[i + n] == l && n < T) ++n;
        c.tile_first.push_back(i);
        c.tile_size.push_back(n);
        c.tile_l.push_back(l);
        i += n;
    }
    std::vector<int> orig_sorted_arr(ns);
    for (int i = 0; i < ns; ++i) orig_sorted_arr[i] = c.sorted_orig[i];
    if (c.d_xyz.upload(c.shell_xyz) || c.d_off.upload(c.shell_off) || c.d_soff.upload(c.shell_soff) ||
        c.d_fperm.upload(c.fn_sorted_to_orig) || c.d_finv.upload(c.fn_orig_to_sorted) || c.d_orig.upload(orig_sorted_arr) ||
        c.d_pairs.upload(c.pairs) || c.d_pp.upload(c.pp) || c.d_ppab.upload(ppab) || c.d_tile_first.upload(c.tile_first) ||
        c.d_tile_size.upload(c.tile_size))
        return g_err;
    // Schwarz bounds on the device
    if (c.d_q.alloc(np)) return g_err;
    for (int la = 0; la <= PM_LMAX; ++la)
        for (int lb = 0; lb <= la; ++lb) {
            std::vector<int> pl, pi, pj;
            for (int I = 0; I < ns; ++I)
                for (int J = 0; J <= I; ++J)
                    if (c.shell_l[I] == la && c.shell_l[J] == lb) {
                        pl.push_back((int)((int64_t)I * (I + 1) / 2 + J));
                        pi.push_back(I);
                        pj.push_back(J);
                    }
            if (pl.empty()) continue;
            DevBuf<int> dl, di, dj;
            if (dl.upload(pl) || di.upload(pi) || dj.upload(pj)) return g_err;
            SchwarzArgs sa{dl.p, di.p, dj.p, (int)pl.size(), c.d_xyz.p, c.d_pairs.p, c.d_pp.p, c.d_q.p, g_rys};
            if (launch_schwarz(pair_class(la, lb), sa, 0)) return fail(PATMOL_HIP_EUNSUPPORTED, "no Schwarz kernel");
            HIPCHK(hipGetLastError());
            HIPCHK(hipDeviceSynchronize());
            dl.release();
            di.release();
            dj.release();
        }
    c.pair_q.resize(np);
    HIPCHK(hipMemcpy(c.pair_q.data(), c.d_q.p, sizeof(double) * np, hipMemcpyDeviceToHost));
    // tile-pair bounds
    const int nt = (int)c.tile_first.size();
    std::vector<int> tmaxorig(nt, -1);
    for (int t = 0; t < nt; ++t)
        for (int i = 0; i < c.tile_size[t]; ++i)
            tmaxorig[t] = std::max(tmaxorig[t], c.sorted_orig[c.tile_first[t] + i]);
    struct TP {
        int a, b, pc;
        double q;
        int maxorig;
    };
    std::vector<TP> tps;
    for (int ta = 0; ta < nt; ++ta)
        for (int tb = 0; tb <= ta; ++tb) {
            double q = 0.0;
            for (int i = 0; i < c.tile_size[ta]; ++i) {
                int I = c.tile_first[ta] + i;
                for (int j = 0; j < c.tile_size[tb]; ++j) {
                    int J = c.tile_first[tb] + j;
                    if (J > I) continue;
                    q = std::max(q, c.pair_q[(int64_t)I * (I + 1) / 2 + J]);
                }
            }
            tps.push_back({ta, tb, pair_class(c.tile_l[ta], c.tile_l[tb]), q, std::max(tmaxorig[ta], tmaxorig[tb])});
        }
    // tile quartets: canonical (P1 >= P2), screened by the bound, bra = higher pair class
    struct Item {
        int4 t;
        int bpc, kpc;
        double bound;
    };
    // Steps jk_kernel3 executes for one orientation of a tile quartet: one
    // step per ket shell K and KSUB = 64 / (TI TJ) consecutive L, executed
    // when any of its lanes holds a quartet with Q_IJ Q_KL >= tau.
    auto qpair = [&](int I, int J) {
        return I >= J ? c.pair_q[(int64_t)I * (I + 1) / 2 + J] : c.pair_q[(int64_t)J * (J + 1) / 2 + I];
    };
    // Some lane of a step holds a quartet iff (largest bra bound) x Q_KL >=
    // tau for one of its (K, L): rounding is monotonic, so the largest
    // product is the product with the largest bra bound (tile pair q).
    auto steps_of = [&](const TP &br, const TP &kt) {
        const int TI = tile_of(c.tile_l[br.a]), TJ = tile_of(c.tile_l[br.b]), TL = tile_of(c.tile_l[kt.b]);
        const int ksub = 64 / (TI * TJ);
        const int fK = c.tile_first[kt.a], nK = c.tile_size[kt.a], fL = c.tile_first[kt.b], nL = c.tile_size[kt.b];
        int steps = 0;
        for (int kq = 0; kq < nK; ++kq)
            for (int l0 = 0; l0 < TL; l0 += ksub) {
                bool any = false;
                for (int lq = l0; lq < std::min(l0 + ksub, nL) && !any; ++lq) {
                    const int K = fK + kq, L = fL + lq;
                    if (kt.a == kt.b && K < L) continue;
                    any = br.q * qpair(K, L) >= c.tau;
                }
                steps += any ? 1 : 0;
            }
        return steps;
    };
    std::vector<Item> items;
    for (size_t p1 = 0; p1 < tps.size(); ++p1)
        for (size_t p2 = 0; p2 <= p1; ++p2) {
            const TP &A = tps[p1], &B = tps[p2];
            double bound = A.q * B.q;
            if (bound < c.tau) continue;
            if (start > 0 && std::max(A.maxorig, B.maxorig) < start) continue;
            const TP *bra = (A.pc >= B.pc) ? &A : &B;
            const TP *ket = (A.pc >= B.pc) ? &B : &A;
            // two tile pairs of one class (other than (dd|dd), which runs
            // the ket-pair-per-lane kernel), or of the class pairs whose
            // reversed classes have kernels too ((ps|ss), (ds|ss), (ds|ps);
            // the reversal measured slower for (pp|ss), (pp|ps), (ds|pp)):
            // the orientation with fewer steps
            const int hi = std::max(A.pc, B.pc), lo = std::min(A.pc, B.pc);
            const bool reversible = (hi == pair_class(1, 0) && lo == pair_class(0, 0)) ||
                                    (hi == pair_class(2, 0) && lo <= pair_class(1, 0));
            const bool orientable = (A.pc == B.pc && p1 != p2 && A.pc != pair_class(2, 2)) || reversible;
            if (orientable && steps_of(*ket, *bra) < steps_of(*bra, *ket)) std::swap(bra, ket);
            items.push_back({make_int4(bra->a, bra->b, ket->a, ket->b), bra->pc, ket->pc, bound});
        }
    std::stable_sort(items.begin(), items.end(), [](const Item &x, const Item &y) {
        if (x.bpc != y.bpc) return x.bpc > y.bpc;
        if (x.kpc != y.kpc) return x.kpc > y.kpc;
        return x.bound > y.bound;
    });
    c.items.resize(items.size());
    c.item_bound.resize(items.size());
    c.ranges.clear();
    for (size_t i = 0; i < items.size(); ++i) {
        c.items[i] = items[i].t;
        c.item_bound[i] = items[i].bound;
        if (c.ranges.empty() || c.ranges.back().bra_pc != items[i].bpc || c.ranges.back().ket_pc != items[i].kpc)
            c.ranges.push_back({items[i].bpc, items[i].kpc, (int)i, 0});
        c.ranges.back().count++;
    }
    if (c.d_items.upload(c.items) || c.d_bound.upload(c.item_bound)) return g_err;
    if (md && build_md_tables(c)) return g_err;
    c.md_off = 0;
    c.ddp_kept = c.ddp_slots = 0;
    for (const ClassRange &r : c.ranges)
        if (md_class_bit(r.bra_pc, r.ket_pc) == 2) sample_kept(c, r, c.ddp_kept, c.ddp_slots);
    if (c.ddp_slots > 0 && (double)c.ddp_kept > MD_DDDP_MAX_KEPT * (double)c.ddp_slots) c.md_off |= 2;
    c.cls_ready = false;
    c.built = true;
    return 0;
}

// unpack packed lower triangle (no2) to full symmetric N x N with rows and
// columns in sorted function order: f[si][sj] = P(perm[si], perm[sj])
__global__ void unpack_kernel(const double *__restrict__ p, double *__restrict__ f, const int *__restrict__ perm, int N)
{
    int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (int64_t)N * N) return;
    int i = perm[idx / N], j = perm[idx % N];
    int a = i >= j ? i : j, b = i >= j ? j : i;
    f[idx] = p[(int64_t)a * (a + 1) / 2 + b];
}

__global__ void add_kernel(const double *__restrict__ a, const double *__restrict__ b, double *__restrict__ o, int64_t n)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) o[i] = a[i] + b[i];
}

// G[ij] += cj (J'ij + J'ji) - (K'ij + K'ji), packed i >= j, with J' and K'
// the sums of the per-XCD fixed-point copies (integer sums, then one
// conversion each), held in sorted function order (inv maps a function to
// its sorted index)
__global__ void finalize_kernel(const FxAcc *__restrict__ J, const FxAcc *__restrict__ K, double *__restrict__ G,
                                const int *__restrict__ inv, int N, double cj)
{
    int64_t ij = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t no2 = (int64_t)N * (N + 1) / 2;
    if (ij >= no2) return;
    int i = (int)((sqrt(8.0 * (double)ij + 1.0) - 1.0) * 0.5);
    while ((int64_t)(i + 1) * (i + 2) / 2 <= ij) ++i;
    while ((int64_t)i * (i + 1) / 2 > ij) --i;
    int j = (int)(ij - (int64_t)i * (i + 1) / 2);
    const int si = inv[i], sj = inv[j];
    const int64_t NN = (int64_t)N * N, a = (int64_t)si * N + sj, b = (int64_t)sj * N + si;
    long long jh = 0, jl = 0, kh = 0, kl = 0;
    for (int x = 0; x < PM_XCD_SLOTS; ++x) {
        const FxAcc ja = J[x * NN + a], jb = J[x * NN + b], ka = K[x * NN + a], kb = K[x * NN + b];
        jh += ja.hi + jb.hi;
        jl += ja.lo + jb.lo;
        kh += ka.hi + kb.hi;
        kl += ka.lo + kb.lo;
    }
    G[ij] += cj * fx_value(jh, jl) - fx_value(kh, kl);
}

// p[0..n) = 0 (a kernel rather than hipMemsetAsync, so that a caller may
// capture the build in a HIP graph)
__global__ void zero_fx_kernel(FxAcc *__restrict__ p, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = FxAcc{0, 0};
}

__global__ void absmax_kernel(const double *__restrict__ p, int64_t n, double *out)
{
    __shared__ double s[256];
    double m = 0.0;
    for (int64_t i = threadIdx.x; i < n; i += blockDim.x) m = fmax(m, fabs(p[i]));
    s[threadIdx.x] = m;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) s[threadIdx.x] = fmax(s[threadIdx.x], s[threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = s[0];
}

// max|D| over each tile-pair block of the full N x N density D (sorted
// function order); one workgroup of 64 per tile pair (grid nt x nt)
__global__ __launch_bounds__(64) void tile_absmax_kernel(const double *__restrict__ D, int N,
                                                         const int *__restrict__ f0, const int *__restrict__ fn, int nt,
                                                         double *__restrict__ T)
{
    const int t1 = blockIdx.y, t2 = blockIdx.x;
    const int r0 = f0[t1], nr = fn[t1], c0 = f0[t2], nc = fn[t2];
    double m = 0.0;
    for (int e = threadIdx.x; e < nr * nc; e += 64) {
        const int i = e / nc, j = e % nc;
        m = fmax(m, fabs(D[(int64_t)(r0 + i) * N + c0 + j]));
    }
#pragma unroll
    for (int d = 32; d > 0; d >>= 1) m = fmax(m, __shfl_xor(m, d, 64));
    if (threadIdx.x == 0) T[(int64_t)t1 * nt + t2] = m;
}

__global__ void zero_int_kernel(int *__restrict__ p, int n)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) p[k] = 0;
}

// Each work item's density-weighted bound: its Schwarz bound times the
// largest max|D| of the six tile-pair blocks its quartets contract with.
// Below tau the item is dropped; below thr it goes to the FP32 list of its
// class range (filled from the back), else to the FP64 list (from the
// front).  Slots are taken per wavefront (one atomic per class range and
// list, lanes in lane order); the order of the wavefronts depends on the
// atomics, which does not change the fixed-point J'/K' sums.
__global__ __launch_bounds__(256) void classify_items_kernel(const int4 *__restrict__ items,
                                                             const double *__restrict__ bound,
                                                             const unsigned char *__restrict__ rng, int nitems,
                                                             const double *__restrict__ T, int nt, double tau,
                                                             double thr, const int *__restrict__ rfirst,
                                                             const int *__restrict__ rcount, int *__restrict__ cnt,
                                                             int4 *__restrict__ out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int lane = threadIdx.x & 63;
    const bool here = i < nitems;
    int4 t = make_int4(0, 0, 0, 0);
    int r = -1, cat = 0;   // cat 0 dropped, 1 FP64, 2 FP32
    if (here) {
        t = items[i];
        r = rng[i];
        auto tm = [&](int a, int b) { return T[(int64_t)a * nt + b]; };
        const double m = fmax(fmax(fmax(tm(t.x, t.y), tm(t.z, t.w)), fmax(tm(t.x, t.z), tm(t.x, t.w))),
                              fmax(tm(t.y, t.z), tm(t.y, t.w)));
        const double b = bound[i] * m;
        cat = b < tau ? 0 : (b < thr ? 2 : 1);
    }
    const unsigned long long below = (1ull << lane) - 1ull;
    unsigned long long todo = __ballot(cat != 0);
    while (todo) {   // wave-uniform: one pass per class range present in the wave
        const int leader = __builtin_ctzll(todo);
        const int rr = __shfl(r, leader, 64);
        const bool mine = cat != 0 && r == rr;
        const unsigned long long m64 = __ballot(mine && cat == 1), m32 = __ballot(mine && cat == 2);
        int b64 = 0, b32 = 0;
        if (lane == leader) {
            if (m64) b64 = atomicAdd(&cnt[2 * rr], __popcll(m64));
            if (m32) b32 = atomicAdd(&cnt[2 * rr + 1], __popcll(m32));
        }
        b64 = __shfl(b64, leader, 64);
        b32 = __shfl(b32, leader, 64);
        if (mine) {
            if (cat == 1) out[rfirst[rr] + b64 + __popcll(m64 & below)] = t;
            else out[rfirst[rr] + rcount[rr] - 1 - (b32 + __popcll(m32 & below))] = t;
        }
        todo &= ~__ballot(mine);
    }
}

// tables of the item classification (once per built context)
static int prepare_classify(Ctx &c)
{
    if (c.cls_ready) return 0;
    const int nt = (int)c.tile_first.size(), nr = (int)c.ranges.size();
    if (nr > 255) return fail(PATMOL_HIP_EUNSUPPORTED, "too many class ranges");
    std::vector<int> f0(nt), fn(nt), rf(nr), rc(nr);
    for (int t = 0; t < nt; ++t) {
        f0[t] = c.shell_soff[c.tile_first[t]];
        fn[t] = c.tile_size[t] * c.shell_nst[c.tile_first[t]];
    }
    std::vector<unsigned char> rng(c.items.size());
    for (int r = 0; r < nr; ++r) {
        rf[r] = c.ranges[r].first;
        rc[r] = c.ranges[r].count;
        for (int i = 0; i < c.ranges[r].count; ++i) rng[c.ranges[r].first + i] = (unsigned char)r;
    }
    if (c.d_tfun0.upload(f0) || c.d_tfunn.upload(fn) || c.d_rfirst.upload(rf) || c.d_rcount.upload(rc) ||
        c.d_item_rng.upload(rng) || c.d_tmax.alloc((size_t)nt * nt) || c.d_cls_items.alloc(c.items.size()) ||
        c.d_cnt.alloc(2 * (size_t)nr))
        return g_err;
    c.cnt_host.assign(2 * (size_t)nr, 0);
    c.cls_ready = true;
    return 0;
}

// classify every item by the density D (full, sorted order) into the lists
// of c.d_cls_items; counts into c.cnt_host (the stream is synchronised)
static int classify_items(Ctx &c, const double *D, double tau, double thr, hipStream_t s)
{
    if (prepare_classify(c)) return g_err;
    const int nt = (int)c.tile_first.size(), nr = (int)c.ranges.size(), ni = (int)c.items.size();
    hipLaunchKernelGGL(tile_absmax_kernel, dim3(nt, nt), dim3(64), 0, s, D, c.N, c.d_tfun0.p, c.d_tfunn.p, nt,
                       c.d_tmax.p);
    hipLaunchKernelGGL(zero_int_kernel, dim3((2 * nr + 255) / 256), dim3(256), 0, s, c.d_cnt.p, 2 * nr);
    if (ni > 0)
        hipLaunchKernelGGL(classify_items_kernel, dim3((ni + 255) / 256), dim3(256), 0, s, c.d_items.p, c.d_bound.p,
                           c.d_item_rng.p, ni, c.d_tmax.p, nt, tau, thr, c.d_rfirst.p, c.d_rcount.p, c.d_cnt.p,
                           c.d_cls_items.p);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(c.cnt_host.data(), c.d_cnt.p, sizeof(int) * 2 * nr, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    return 0;
}

// Launch every class: J' from density DJ, K' from density DK (both full
// N x N in sorted function order), into the fixed-point copies J, K (zeroed
// here).  nside side streams (0: everything on s).  dmax is max|density|
// (used by the precision split); with delta (a density difference of an
// incremental build) quartets whose Schwarz bound times dmax is below the
// Schwarz threshold are skipped as well: whole items by their bound, single
// quartets in the kernel.  With per_item (DJ == DK, RHF) the precision
// split and the delta screening use each item's own density-weighted bound
// (classify_items) instead of max|density| over the whole matrix.
static int run_jk(Ctx &c, const double *DJ, const double *DK, FxAcc *J, FxAcc *K, double dmax, int nside,
                  bool timed, hipStream_t s, bool delta = false, bool per_item = false)
{
    const int64_t NN = (int64_t)c.N * c.N * PM_XCD_SLOTS;
    hipLaunchKernelGGL(zero_fx_kernel, dim3((unsigned)((NN + 255) / 256)), dim3(256), 0, s, J, NN);
    hipLaunchKernelGGL(zero_fx_kernel, dim3((unsigned)((NN + 255) / 256)), dim3(256), 0, s, K, NN);
    HIPCHK(hipGetLastError());
    timed = timed && g_timing;
    if (timed) {
        if (!c.ev0) {
            HIPCHK(hipEventCreate(&c.ev0));
            HIPCHK(hipEventCreate(&c.ev1));
            c.cls_ev.resize(2 * 64);
            for (auto &e : c.cls_ev) HIPCHK(hipEventCreate(&e));
        }
        HIPCHK(hipEventRecord(c.ev0, s));
    }
    const bool cls = per_item && (g_prec_mode == 1 || delta) && dmax > 0.0;
    if (cls && classify_items(c, DJ, delta ? c.tau : 0.0, g_prec_mode == 1 ? g_prec_thr : -1.0, s)) return g_err;
    c.items_fp32 = 0;
    // Classes are independent: spread them over side streams so small grids
    // overlap.  With timing on, launches stay on the caller's stream so each
    // class's events bracket only its own kernels.
    if (timed) nside = 0;
    if (nside && c.side.empty()) {
        c.side.resize(nside);
        c.side_ev.resize(nside);
        for (int i = 0; i < nside; ++i) {
            HIPCHK(hipStreamCreateWithFlags(&c.side[i], hipStreamNonBlocking));
            HIPCHK(hipEventCreateWithFlags(&c.side_ev[i], hipEventDisableTiming));
        }
        HIPCHK(hipEventCreateWithFlags(&c.fork_ev, hipEventDisableTiming));
    }
    if (nside) {
        HIPCHK(hipEventRecord(c.fork_ev, s));
        for (int i = 0; i < nside; ++i) HIPCHK(hipStreamWaitEvent(c.side[i], c.fork_ev, 0));
    }
    c.items_skipped = 0;
    // items are sorted by bound within a class (descending): the FP32 items
    // and, in a delta build, the skipped items form suffixes
    auto count_at_least = [&](const ClassRange &r, double t) {
        return (int)(std::lower_bound(c.item_bound.begin() + r.first, c.item_bound.begin() + r.first + r.count, t,
                                      [](double b, double v) { return b >= v; }) -
                     (c.item_bound.begin() + r.first));
    };
    const double ktau = delta && dmax > 0.0 ? std::max(c.tau, c.tau / dmax) : c.tau;
    int ci = 0;
    for (const ClassRange &r : c.ranges) {
        hipStream_t ls = nside ? c.side[ci % nside] : s;
        int nkeep = r.count;
        if (delta) nkeep = dmax > 0.0 ? count_at_least(r, c.tau / dmax) : 0;
        int n64 = nkeep;
        if (g_prec_mode == 1 && dmax > 0.0) n64 = std::min(nkeep, count_at_least(r, g_prec_thr / dmax));
        // classified lists: FP64 items at the front of the range, FP32 at the back
        const int4 *list64 = c.d_items.p + r.first, *list32 = c.d_items.p + r.first + n64;
        if (cls) {
            n64 = c.cnt_host[2 * ci];
            nkeep = n64 + c.cnt_host[2 * ci + 1];
            list64 = c.d_cls_items.p + r.first;
            list32 = c.d_cls_items.p + r.first + r.count - (nkeep - n64);
        }
        c.items_skipped += r.count - nkeep;
        JKArgs a;
        a.tile_first = c.d_tile_first.p;
        a.tile_size = c.d_tile_size.p;
        a.shell_xyz = c.d_xyz.p;
        a.shell_off = c.d_soff.p;   // D, J', K' in sorted function order
        a.shell_orig = c.d_orig.p;
        a.pairs = c.d_pairs.p;
        a.pair_q = c.d_q.p;
        a.pp = c.d_pp.p;
        a.DJ = DJ;
        a.DK = DK;
        a.J = J;
        a.K = K;
        a.N = c.N;
        a.tau = ktau;
        a.start_shell = c.start;
        a.rys = g_rys;
        a.item_mol = nullptr;
        a.mols = nullptr;
        a.slot_stride = (long long)c.N * c.N;
        c.md.fill(a, md_effective(c.md_off) & md_class_bit(r.bra_pc, r.ket_pc));
        if (timed && ci < 64) HIPCHK(hipEventRecord(c.cls_ev[2 * ci], s));
        // a class whose items below the threshold stay FP64 runs both lists
        // in one launch where they are adjacent (every build but a delta
        // build, which drops items between them)
        const int p32 = fp32_prec(r.bra_pc, r.ket_pc);
        if (p32 == 0 && nkeep > n64 && list32 == list64 + n64) n64 = nkeep;
        if (n64 > 0) {
            a.items = list64;
            a.nitems = n64;
            if (launch_jk(r.bra_pc, r.ket_pc, a, 0, ls)) return fail(PATMOL_HIP_EUNSUPPORTED, "no J/K kernel for class %d/%d", r.bra_pc, r.ket_pc);
        }
        if (nkeep > n64) {
            a.items = list32;
            a.nitems = nkeep - n64;
            if (p32) c.items_fp32 += nkeep - n64;
            if (launch_jk(r.bra_pc, r.ket_pc, a, p32, ls)) return fail(PATMOL_HIP_EUNSUPPORTED, "no J/K kernel for class %d/%d", r.bra_pc, r.ket_pc);
        }
        if (timed && ci < 64) HIPCHK(hipEventRecord(c.cls_ev[2 * ci + 1], s));
        HIPCHK(hipGetLastError());
        ++ci;
    }
    if (nside) {
        for (int i = 0; i < nside; ++i) {
            HIPCHK(hipEventRecord(c.side_ev[i], c.side[i]));
            HIPCHK(hipStreamWaitEvent(s, c.side_ev[i], 0));
        }
    }
    if (timed) HIPCHK(hipEventRecord(c.ev1, s));
    return 0;
}

static int collect_timing(Ctx &c)
{
    if (!g_timing || !c.ev0) return 0;
    HIPCHK(hipEventSynchronize(c.ev1));
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, c.ev0, c.ev1));
    c.last_ms = ms;
    for (double &x : c.class_ms) x = 0.0;
    int ci = 0;
    for (const ClassRange &r : c.ranges) {
        if (ci >= 64) break;
        float t = 0;
        HIPCHK(hipEventElapsedTime(&t, c.cls_ev[2 * ci], c.cls_ev[2 * ci + 1]));
        c.class_ms[r.bra_pc * 6 + r.ket_pc] += t;
        ++ci;
    }
    return 0;
}

static int alloc_work(Ctx &c, bool uhf)
{
    const int64_t NN = (int64_t)c.N * c.N, no2 = (int64_t)c.N * (c.N + 1) / 2;
    if (c.d_D.alloc(NN) || c.d_J.alloc(NN * PM_XCD_SLOTS) || c.d_K.alloc(NN * PM_XCD_SLOTS) || c.d_P.alloc(no2) ||
        c.d_G.alloc(no2) || c.d_dmax.alloc(1))
        return g_err;
    if (uhf && (c.d_D2.alloc(NN) || c.d_DJ.alloc(NN) || c.d_PB.alloc(no2) || c.d_G2.alloc(no2))) return g_err;
    return 0;
}

static double device_absmax(Ctx &c, const double *p, int64_t n, hipStream_t s)
{
    hipLaunchKernelGGL(absmax_kernel, dim3(1), dim3(256), 0, s, p, n, c.d_dmax.p);
    double m = 0.0;
    if (hipMemcpyAsync(&m, c.d_dmax.p, sizeof(double), hipMemcpyDeviceToHost, s) != hipSuccess) return -1.0;
    if (hipStreamSynchronize(s) != hipSuccess) return -1.0;
    return m;
}

// G += 2J[P] - K[P] on stream s with the work buffers D (N x N doubles) and
// J, K (PM_XCD_SLOTS x N x N fixed-point each) given by the caller.
static int rhf_build(Ctx &c, const double *dP, double *dG, double *D, FxAcc *J, FxAcc *K, double dmax, int nside,
                     bool timed, hipStream_t s, bool delta = false)
{
    const int64_t NN = (int64_t)c.N * c.N, no2 = (int64_t)c.N * (c.N + 1) / 2;
    hipLaunchKernelGGL(unpack_kernel, dim3((NN + 255) / 256), dim3(256), 0, s, dP, D, c.d_fperm.p, c.N);
    if (run_jk(c, D, D, J, K, dmax, nside, timed, s, delta, true)) return g_err;
    hipLaunchKernelGGL(finalize_kernel, dim3((no2 + 255) / 256), dim3(256), 0, s, J, K, dG, c.d_finv.p, c.N, 2.0);
    HIPCHK(hipGetLastError());
    return 0;
}

static int rhf_device(Ctx &c, const double *dP, double *dG, hipStream_t s, bool delta = false)
{
    if (alloc_work(c, false)) return g_err;
    const int64_t no2 = (int64_t)c.N * (c.N + 1) / 2;
    double dmax = 0.0;
    if (g_prec_mode == 1 || delta) {
        dmax = device_absmax(c, dP, no2, s);
        if (dmax < 0.0) return fail(PATMOL_HIP_EDEVICE, "device failure in max|P|");
    }
    c.last_dmax = dmax;
    return rhf_build(c, dP, dG, c.d_D.p, c.d_J.p, c.d_K.p, dmax, 4, true, s, delta);
}

// ---------------------------------------------------------------------------
// allocation-free G build (patmol_hip_rhf_gmatrix_device_noalloc): the work
// buffers of rhf_build live in one fixed device array, sized for bases of
// up to PATMOL_HIP_NOALLOC_MAX_NORB functions:
//   D                 N^2 doubles                  (offset 0)
//   J', K' copies     PM_XCD_SLOTS N^2 FxAcc each
// 17,301,504 bytes at N = 256, held by every device the library runs on
// until the process ends.  Host code never uses the array's name as an
// address: the device address comes from hipGetSymbolAddress, per device.
// ---------------------------------------------------------------------------
constexpr size_t noalloc_align(size_t b) { return (b + 255) & ~(size_t)255; }
constexpr size_t NOALLOC_N = PATMOL_HIP_NOALLOC_MAX_NORB;
constexpr size_t NOALLOC_D_BYTES = noalloc_align(sizeof(double) * NOALLOC_N * NOALLOC_N);
constexpr size_t NOALLOC_FX_BYTES = noalloc_align(sizeof(FxAcc) * PM_XCD_SLOTS * NOALLOC_N * NOALLOC_N);
constexpr size_t NOALLOC_BYTES = NOALLOC_D_BYTES + 2 * NOALLOC_FX_BYTES;
static_assert(NOALLOC_BYTES == 17301504, "allocation-free scratch size");
static __device__ __attribute__((aligned(256))) unsigned char scratch[NOALLOC_BYTES];

static std::mutex g_scratch_mu;
static constexpr int kMaxDevices = 64;
static void *g_scratch_dev[kMaxDevices] = {};

// device address of `scratch` on the current device (looked up once per device)
static int scratch_address(void **out)
{
    int dev = -1;
    HIPCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return fail(PATMOL_HIP_EDEVICE, "device index %d out of range", dev);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        void *p = nullptr;
        HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return fail(PATMOL_HIP_EDEVICE, "hipGetSymbolAddress returned no address");
        g_scratch_dev[dev] = p;
    }
    *out = g_scratch_dev[dev];
    return 0;
}

static int rhf_device_noalloc(Ctx &c, const double *dP, double *dG, hipStream_t s)
{
    if (c.N > PATMOL_HIP_NOALLOC_MAX_NORB)
        return fail(PATMOL_HIP_EUNSUPPORTED, "%d functions: the allocation-free G build takes at most %d", c.N,
                    PATMOL_HIP_NOALLOC_MAX_NORB);
    void *base = nullptr;
    if (scratch_address(&base)) return g_err;
    char *b = static_cast<char *>(base);
    double *D = reinterpret_cast<double *>(b);
    FxAcc *J = reinterpret_cast<FxAcc *>(b + NOALLOC_D_BYTES);
    FxAcc *K = reinterpret_cast<FxAcc *>(b + NOALLOC_D_BYTES + NOALLOC_FX_BYTES);
    // FP64 throughout (the FP32 split needs max|P| on the host), everything
    // on the caller's stream, no events
    return rhf_build(c, dP, dG, D, J, K, 0.0, 0, false, s);
}

// GA += J[PA+PB] - K[PA], GB += J[PA+PB] - K[PB] (scfaux.c:554-714), all on
// stream s: two single-density passes (J[PA+PB], K[PA]) and (J[PA+PB], K[PB])
// with J = J' + J'^T.
static int uhf_device(Ctx &x, const double *dPA, const double *dPB, double *dGA, double *dGB, hipStream_t s)
{
    if (alloc_work(x, true)) return g_err;
    const int64_t NN = (int64_t)x.N * x.N, no2 = (int64_t)x.N * (x.N + 1) / 2;
    hipLaunchKernelGGL(unpack_kernel, dim3((NN + 255) / 256), dim3(256), 0, s, dPA, x.d_D.p, x.d_fperm.p, x.N);
    hipLaunchKernelGGL(unpack_kernel, dim3((NN + 255) / 256), dim3(256), 0, s, dPB, x.d_D2.p, x.d_fperm.p, x.N);
    hipLaunchKernelGGL(add_kernel, dim3((NN + 255) / 256), dim3(256), 0, s, x.d_D.p, x.d_D2.p, x.d_DJ.p, NN);
    double dmax = 0.0;
    if (g_prec_mode == 1) dmax = std::max(device_absmax(x, dPA, no2, s), device_absmax(x, dPB, no2, s));
    if (run_jk(x, x.d_DJ.p, x.d_D.p, x.d_J.p, x.d_K.p, dmax, 4, true, s)) return g_err;
    hipLaunchKernelGGL(finalize_kernel, dim3((no2 + 255) / 256), dim3(256), 0, s, x.d_J.p, x.d_K.p, dGA, x.d_finv.p,
                       x.N, 1.0);
    if (run_jk(x, x.d_DJ.p, x.d_D2.p, x.d_J.p, x.d_K.p, dmax, 4, true, s)) return g_err;
    hipLaunchKernelGGL(finalize_kernel, dim3((no2 + 255) / 256), dim3(256), 0, s, x.d_J.p, x.d_K.p, dGB, x.d_finv.p,
                       x.N, 1.0);
    HIPCHK(hipGetLastError());
    return 0;
}

// device memory, streams and events of a context
static void release_ctx(Ctx &c)
{
    for (auto *b : {&c.d_xyz, &c.d_q, &c.d_bound, &c.d_P, &c.d_PB, &c.d_D, &c.d_D2, &c.d_DJ, &c.d_G, &c.d_G2, &c.d_dmax,
                    &c.d_tmax})
        b->release();
    for (auto *b : {&c.d_tfun0, &c.d_tfunn, &c.d_rfirst, &c.d_rcount, &c.d_cnt}) b->release();
    c.d_item_rng.release();
    c.d_cls_items.release();
    c.d_J.release();
    c.d_K.release();
    for (auto *b : {&c.d_off, &c.d_soff, &c.d_fperm, &c.d_finv, &c.d_orig, &c.d_tile_first, &c.d_tile_size}) b->release();
    c.d_pairs.release();
    c.d_pp.release();
    c.d_ppab.release();
    c.d_items.release();
    c.md.release();
    for (auto st : c.side) (void)hipStreamDestroy(st);
    for (auto ev : c.side_ev) (void)hipEventDestroy(ev);
    c.side.clear();
    c.side_ev.clear();
    if (c.fork_ev) (void)hipEventDestroy(c.fork_ev);
    c.fork_ev = nullptr;
    if (c.ev0) {
        (void)hipEventDestroy(c.ev0);
        (void)hipEventDestroy(c.ev1);
        for (auto ev : c.cls_ev) (void)hipEventDestroy(ev);
        c.ev0 = c.ev1 = nullptr;
        c.cls_ev.clear();
    }
}

int aux_fail(int code, const char *msg) { return fail(code, "%s", msg); }

RysTables aux_tables()
{
    g_err = PATMOL_HIP_OK;
    std::lock_guard<std::mutex> lk(g_mu);
    if (ensure_tables()) return RysTables{nullptr, nullptr, nullptr, nullptr};
    return g_rys;
}

// ---------------------------------------------------------------------------
// molecule batches: every molecule's tables (as two_electron_ints builds
// them) concatenated; work items never mix molecules and carry their
// molecule's index, so each class is one launch over all molecules; the
// whole G build (unpack, zero, class launches, finalise) is captured once as
// a HIP graph and replayed.
// ---------------------------------------------------------------------------
struct BatchMol {
    int N = 0, nshell = 0;
    int sbase = 0, fbase = 0, tbase = 0;
    long long pbase = 0, ppbase = 0, dstart = 0, pstart = 0;
};

// D of every molecule (its block at dstart, rows and columns in its sorted
// function order) from the packed half densities (at pstart, reference
// order); fperm: local original function of each global sorted function
__global__ void batch_unpack_kernel(const double *__restrict__ P, double *__restrict__ D,
                                    const long long *__restrict__ dstart, const long long *__restrict__ pstart,
                                    const int *__restrict__ Nm, const int *__restrict__ fbase,
                                    const int *__restrict__ fperm, int nmol, long long dtot)
{
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= dtot) return;
    int lo = 0, hi = nmol - 1;   // last molecule with dstart <= idx
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (dstart[mid] <= idx) lo = mid;
        else hi = mid - 1;
    }
    const int m = lo, N = Nm[m];
    const long long e = idx - dstart[m];
    const int i = fperm[fbase[m] + (int)(e / N)], j = fperm[fbase[m] + (int)(e % N)];
    const int a = i >= j ? i : j, b = i >= j ? j : i;
    D[idx] = P[pstart[m] + (long long)a * (a + 1) / 2 + b];
}

// G = cj (J' + J'^T) - (K' + K'^T) per molecule into the packed triangles;
// finv: local sorted index of each local original function (at fbase)
__global__ void batch_finalize_kernel(const FxAcc *__restrict__ J, const FxAcc *__restrict__ K, double *__restrict__ G,
                                      const long long *__restrict__ dstart, const long long *__restrict__ pstart,
                                      const int *__restrict__ Nm, const int *__restrict__ fbase,
                                      const int *__restrict__ finv, int nmol, long long no2tot, long long dtot,
                                      double cj)
{
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= no2tot) return;
    int lo = 0, hi = nmol - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (pstart[mid] <= idx) lo = mid;
        else hi = mid - 1;
    }
    const int m = lo, N = Nm[m];
    const long long ij = idx - pstart[m];
    int i = (int)((sqrt(8.0 * (double)ij + 1.0) - 1.0) * 0.5);
    while ((long long)(i + 1) * (i + 2) / 2 <= ij) ++i;
    while ((long long)i * (i + 1) / 2 > ij) --i;
    const int j = (int)(ij - (long long)i * (i + 1) / 2);
    const int si = finv[fbase[m] + i], sj = finv[fbase[m] + j];
    const long long a = dstart[m] + (long long)si * N + sj, b = dstart[m] + (long long)sj * N + si;
    long long jh = 0, jl = 0, kh = 0, kl = 0;
    for (int x = 0; x < PM_XCD_SLOTS; ++x) {
        const FxAcc ja = J[x * dtot + a], jb = J[x * dtot + b], ka = K[x * dtot + a], kb = K[x * dtot + b];
        jh += ja.hi + jb.hi;
        jl += ja.lo + jb.lo;
        kh += ka.hi + kb.hi;
        kl += ka.lo + kb.lo;
    }
    G[idx] = cj * fx_value(jh, jl) - fx_value(kh, kl);
}

__global__ void zero_kernel(double *__restrict__ p, long long n)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0;
}

}  // namespace pm

struct PatmolHipBatch {
    std::vector<pm::BatchMol> mol;
    std::vector<double> xyz, pair_q;
    std::vector<int> soff, orig, tile_first, tile_size, fperm, finv;
    std::vector<pm::PairRef> pairs;
    std::vector<pm::PrimPair> pp;
    struct It {
        int4 t;
        int mol, bpc, kpc;
        double bound;
    };
    std::vector<It> its;
    long long dtot = 0, no2tot = 0;
    int nshell = 0, norb = 0, ntile = 0;
    // after finalize
    bool finalized = false;
    std::vector<pm::ClassRange> ranges;
    std::vector<int> n64;   // FP64 items at the front of each range
    std::vector<int> Nvec;
    std::vector<int64_t> pstart_v;
    pm::DevBuf<double> d_xyz, d_q, d_P, d_G, d_D;
    pm::DevBuf<int> d_soff, d_orig, d_tile_first, d_tile_size, d_fperm, d_finv, d_item_mol, d_N, d_fbase;
    pm::DevBuf<long long> d_dstart, d_pstart;
    pm::DevBuf<pm::PairRef> d_pairs;
    pm::DevBuf<pm::PrimPair> d_pp;
    pm::DevBuf<int4> d_items;
    pm::DevBuf<pm::MolDesc> d_mols;
    pm::DevBuf<int> d_active;   // per molecule: its items run (all 1 outside the batched SCF)
    pm::DevBuf<pm::FxAcc> d_J, d_K;
    std::vector<int> shell_l;   // l of each (batch-global) sorted shell
    pm::MdTables md;            // McMurchie-Davidson tables of every molecule's pairs
    hipStream_t s = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int64_t items_fp32 = 0;
    double tau = 1e-12;
    // (dd|dp) kept share over every molecule's sample (MD_DDDP_MAX_KEPT)
    int64_t ddp_kept = 0, ddp_slots = 0;
    int md_off = 0;
};

namespace pm {

// the G build of a finalised batch: P (d_P) in, G (d_G) out, on stream s
static int batch_build(PatmolHipBatch &b, hipStream_t s)
{
    const long long dtot = b.dtot, n2 = b.no2tot, nfx = dtot * PM_XCD_SLOTS;
    const int nm = (int)b.mol.size();
    hipLaunchKernelGGL(batch_unpack_kernel, dim3((unsigned)((dtot + 255) / 256)), dim3(256), 0, s, b.d_P.p, b.d_D.p,
                       b.d_dstart.p, b.d_pstart.p, b.d_N.p, b.d_fbase.p, b.d_fperm.p, nm, dtot);
    hipLaunchKernelGGL(zero_fx_kernel, dim3((unsigned)((nfx + 255) / 256)), dim3(256), 0, s, b.d_J.p, (int64_t)nfx);
    hipLaunchKernelGGL(zero_fx_kernel, dim3((unsigned)((nfx + 255) / 256)), dim3(256), 0, s, b.d_K.p, (int64_t)nfx);
    for (size_t ri = 0; ri < b.ranges.size(); ++ri) {
        const ClassRange &r = b.ranges[ri];
        JKArgs a;
        a.tile_first = b.d_tile_first.p;
        a.tile_size = b.d_tile_size.p;
        a.shell_xyz = b.d_xyz.p;
        a.shell_off = b.d_soff.p;
        a.shell_orig = b.d_orig.p;   // read by the kernels, used only for start_shell > 0
        a.pairs = b.d_pairs.p;
        a.pair_q = b.d_q.p;
        a.pp = b.d_pp.p;
        a.DJ = b.d_D.p;
        a.DK = b.d_D.p;
        a.J = b.d_J.p;
        a.K = b.d_K.p;
        a.N = 0;
        a.tau = b.tau;
        a.start_shell = 0;
        a.rys = g_rys;
        a.item_mol = b.d_item_mol.p + r.first;
        a.mols = b.d_mols.p;
        a.mol_active = b.d_active.p;
        a.slot_stride = dtot;
        b.md.fill(a, md_effective(b.md_off) & md_class_bit(r.bra_pc, r.ket_pc));
        // a class whose items below the threshold stay FP64: one launch
        const int n64 = fp32_prec(r.bra_pc, r.ket_pc) ? b.n64[ri] : r.count;
        if (n64 > 0) {
            a.items = b.d_items.p + r.first;
            a.nitems = n64;
            if (launch_jk(r.bra_pc, r.ket_pc, a, 0, s)) return fail(PATMOL_HIP_EUNSUPPORTED, "no J/K kernel");
        }
        if (r.count > n64) {
            a.items = b.d_items.p + r.first + n64;
            a.item_mol = b.d_item_mol.p + r.first + n64;
            a.nitems = r.count - n64;
            if (launch_jk(r.bra_pc, r.ket_pc, a, fp32_prec(r.bra_pc, r.ket_pc), s))
                return fail(PATMOL_HIP_EUNSUPPORTED, "no J/K kernel");
        }
    }
    hipLaunchKernelGGL(zero_kernel, dim3((unsigned)((n2 + 255) / 256)), dim3(256), 0, s, b.d_G.p, n2);
    hipLaunchKernelGGL(batch_finalize_kernel, dim3((unsigned)((n2 + 255) / 256)), dim3(256), 0, s, b.d_J.p, b.d_K.p,
                       b.d_G.p, b.d_dstart.p, b.d_pstart.p, b.d_N.p, b.d_fbase.p, b.d_finv.p, nm, n2, dtot, 2.0);
    HIPCHK(hipGetLastError());
    return 0;
}

bool batch_view(PatmolHipBatch *b, BatchView &v)
{
    if (!b || !b->finalized) return false;
    v.nmol = (int)b->mol.size();
    v.N = b->Nvec.data();
    v.pstart = b->pstart_v.data();
    v.no2tot = b->no2tot;
    v.d_P = b->d_P.p;
    v.d_G = b->d_G.p;
    return true;
}

int batch_set_active(PatmolHipBatch *b, const int *active, hipStream_t s)
{
    if (!b || !b->finalized) return fail(PATMOL_HIP_ESTATE, "batch not finalised");
    HIPCHK(hipMemcpyAsync(b->d_active.p, active, sizeof(int) * b->mol.size(), hipMemcpyHostToDevice, s));
    HIPCHK(hipStreamSynchronize(s));
    return 0;
}

int batch_launch(PatmolHipBatch *b, hipStream_t s)
{
    if (!b || !b->finalized || !b->exec) return fail(PATMOL_HIP_ESTATE, "batch not finalised");
    HIPCHK(hipGraphLaunch(b->exec, s));
    return 0;
}

}  // namespace pm

using namespace pm;

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" {

int patmol_hip_last_error(void) { return g_err; }
const char *patmol_hip_error_string(void) { return g_msg; }
void patmol_hip_set_strict(int s) { g_strict = s ? 1 : 0; }
void patmol_hip_set_schwarz_threshold(double tau) { g_tau = tau < 0 ? 0 : tau; }
void patmol_hip_set_precision(int mode, double thr)
{
    g_prec_mode = mode ? 1 : 0;
    if (thr > 0) g_prec_thr = thr;
}
void patmol_hip_set_fp32_policy(int every_class) { g_fp32_policy = every_class ? 1 : 0; }
void patmol_hip_set_jk_engine(int mask) { g_md_mask = mask < 0 ? MD_DEFAULT_MASK : (mask & 31); }
void patmol_hip_set_timing(int on) { g_timing = on != 0; }

// erifs.c:32-60: the handle keeps the reference's public fields; nfiles
// stays 0 because no integral files are written.
ERIFS *init_ERIFS(int nst, int nend, const char *sname)
{
    g_err = PATMOL_HIP_OK;
    (void)nst;
    (void)nend;
    ERIFS *e = (ERIFS *)calloc(1, sizeof(ERIFS));
    if (!e) {
        fail(PATMOL_HIP_EBADARG, "out of host memory");
        return nullptr;
    }
    e->rank = 0;
    e->nproc = 1;
    e->nfiles = 0;
    e->numints = (int *)calloc(1, sizeof(int));
    const char *nm = sname ? sname : "";
    e->basename = (char *)malloc(strlen(nm) + 1);
    strcpy(e->basename, nm);
    std::lock_guard<std::mutex> lk(g_mu);
    g_ctx[e] = new Ctx();
    return e;
}

void clean_ERIFS(ERIFS *e)
{
    if (!e) return;
    Ctx *c = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_mu);
        auto it = g_ctx.find(e);
        if (it != g_ctx.end()) {
            c = it->second;
            g_ctx.erase(it);
        }
    }
    if (c) {
        release_ctx(*c);
        delete c;
    }
    free(e->numints);
    free(e->basename);
    free(e);
}

void two_electron_ints(ERIFS *e, int start)
{
    g_err = PATMOL_HIP_OK;
    Ctx *c = ctx_of(e);
    if (!c) {
        fail(PATMOL_HIP_EBADARG, "two_electron_ints: unknown ERIFS handle");
        return;
    }
    {
        std::lock_guard<std::mutex> lk(g_mu);
        if (ensure_tables()) return;
    }
    build_tables(*c, start < 0 ? 0 : start);
}

void rhf_form_gmatrix(const ERIFS *e, const double *P, double *G)
{
    g_err = PATMOL_HIP_OK;
    Ctx *c = ctx_of(e);
    if (!c || !P || !G) {
        fail(PATMOL_HIP_EBADARG, "rhf_form_gmatrix: bad handle or null matrix");
        return;
    }
    if (!c->built) {
        fail(PATMOL_HIP_ESTATE, "rhf_form_gmatrix: two_electron_ints was not run for this handle");
        return;
    }
    const int64_t no2 = (int64_t)c->N * (c->N + 1) / 2;
    if (alloc_work(*c, false)) return;
    if (hipMemcpy(c->d_P.p, P, sizeof(double) * no2, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemset(c->d_G.p, 0, sizeof(double) * no2) != hipSuccess) {
        fail(PATMOL_HIP_EDEVICE, "rhf_form_gmatrix: transfer failed");
        return;
    }
    if (rhf_device(*c, c->d_P.p, c->d_G.p, 0)) return;
    std::vector<double> g(no2);
    if (hipMemcpy(g.data(), c->d_G.p, sizeof(double) * no2, hipMemcpyDeviceToHost) != hipSuccess) {
        fail(PATMOL_HIP_EDEVICE, "rhf_form_gmatrix: transfer failed");
        return;
    }
    collect_timing(*c);
    for (int64_t i = 0; i < no2; ++i) G[i] += g[i];
}

void uhf_form_gmatrix(const ERIFS *e, const double *PA, const double *PB, double *GA, double *GB)
{
    g_err = PATMOL_HIP_OK;
    Ctx *c = ctx_of(e);
    if (!c || !PA || !PB || !GA || !GB) {
        fail(PATMOL_HIP_EBADARG, "uhf_form_gmatrix: bad handle or null matrix");
        return;
    }
    if (!c->built) {
        fail(PATMOL_HIP_ESTATE, "uhf_form_gmatrix: two_electron_ints was not run for this handle");
        return;
    }
    Ctx &x = *c;
    const int64_t no2 = (int64_t)x.N * (x.N + 1) / 2;
    if (alloc_work(x, true)) return;
    if (hipMemcpy(x.d_P.p, PA, sizeof(double) * no2, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(x.d_PB.p, PB, sizeof(double) * no2, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemset(x.d_G.p, 0, sizeof(double) * no2) != hipSuccess ||
        hipMemset(x.d_G2.p, 0, sizeof(double) * no2) != hipSuccess) {
        fail(PATMOL_HIP_EDEVICE, "uhf_form_gmatrix: transfer failed");
        return;
    }
    if (uhf_device(x, x.d_P.p, x.d_PB.p, x.d_G.p, x.d_G2.p, 0)) return;
    std::vector<double> ga(no2), gb(no2);
    if (hipMemcpy(ga.data(), x.d_G.p, sizeof(double) * no2, hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(gb.data(), x.d_G2.p, sizeof(double) * no2, hipMemcpyDeviceToHost) != hipSuccess) {
        fail(PATMOL_HIP_EDEVICE, "uhf_form_gmatrix: transfer failed");
        return;
    }
    collect_timing(x);
    for (int64_t i = 0; i < no2; ++i) {
        GA[i] += ga[i];
        GB[i] += gb[i];
    }
}

int patmol_hip_rhf_gmatrix_device(const ERIFS *e, const double *dP, double *dG, void *stream)
{
    g_err = PATMOL_HIP_OK;
    Ctx *c = ctx_of(e);
    if (!c || !dP || !dG) return fail(PATMOL_HIP_EBADARG, "bad handle or null pointer");
    if (!c->built) return fail(PATMOL_HIP_ESTATE, "two_electron_ints was not run for this handle");
    return rhf_device(*c, dP, dG, (hipStream_t)stream);
}

int patmol_hip_rhf_gmatrix_device_noalloc(const ERIFS *e, const double *dP, double *dG, void *stream)
{
    g_err = PATMOL_HIP_OK;
    Ctx *c = ctx_of(e);
    if (!c || !dP || !dG) return fail(PATMOL_HIP_EBADARG, "bad handle or null pointer");
    if (!c->built) return fail(PATMOL_HIP_ESTATE, "two_electron_ints was not run for this handle");
    return rhf_device_noalloc(*c, dP, dG, (hipStream_t)stream);
}

int patmol_hip_uhf_gmatrix_device(const ERIFS *e, const double *dPA, const double *dPB, double *dGA, double *dGB,
                                  void *stream)
{
    g_err = PATMOL_HIP_OK;
    Ctx *c = ctx_of(e);
    if (!c || !dPA || !dPB || !dGA || !dGB) return fail(PATMOL_HIP_EBADARG, "bad handle or null pointer");
    if (!c->built) return fail(PATMOL_HIP_ESTATE, "two_electron_ints was not run for this handle");
    return uhf_device(*c, dPA, dPB, dGA, dGB, (hipStream_t)stream);
}

int patmol_hip_last_timing(const ERIFS *e, double *total_ms, double *class_ms)
{
    Ctx *c = ctx_of(e);
    if (!c) return fail(PATMOL_HIP_EBADARG, "bad handle");
    if (collect_timing(*c)) return g_err;
    if (total_ms) *total_ms = c->last_ms;
    if (class_ms) memcpy(class_ms, c->class_ms, sizeof(c->class_ms));
    return 0;
}

int patmol_hip_jk_engine_mask(const ERIFS *e)
{
    Ctx *c = ctx_of(e);
    if (!c || !c->built) return -fail(PATMOL_HIP_ESTATE, "no tables");
    return c->md.ready ? md_effective(c->md_off) : 0;
}

int patmol_hip_pair_schwarz(const ERIFS *e, double *q)
{
    Ctx *c = ctx_of(e);
    if (!c || !c->built) return fail(PATMOL_HIP_ESTATE, "no tables");
    for (int I = 0; I < c->nshell; ++I)
        for (int J = 0; J <= I; ++J) {
            int a = c->orig_to_sorted[I], b = c->orig_to_sorted[J];
            int hi = std::max(a, b), lo = std::min(a, b);
            q[(int64_t)I * (I + 1) / 2 + J] = c->pair_q[(int64_t)hi * (hi + 1) / 2 + lo];
        }
    return 0;
}

int64_t patmol_hip_work_items(const ERIFS *e, int *tiles4, int *classes2, double *bounds, int *sorted_orig,
                              int *tile_first, int *tile_size, int *ntiles)
{
    Ctx *c = ctx_of(e);
    if (!c || !c->built) return -fail(PATMOL_HIP_ESTATE, "no tables");
    for (size_t i = 0; i < c->items.size(); ++i) {
        if (tiles4) {
            tiles4[4 * i] = c->items[i].x;
            tiles4[4 * i + 1] = c->items[i].y;
            tiles4[4 * i + 2] = c->items[i].z;
            tiles4[4 * i + 3] = c->items[i].w;
        }
        if (bounds) bounds[i] = c->item_bound[i];
    }
    if (classes2)
        for (const ClassRange &r : c->ranges)
            for (int i = 0; i < r.count; ++i) {
                classes2[2 * (r.first + i)] = r.bra_pc;
                classes2[2 * (r.first + i) + 1] = r.ket_pc;
            }
    if (sorted_orig) memcpy(sorted_orig, c->sorted_orig.data(), sizeof(int) * c->nshell);
    if (tile_first) memcpy(tile_first, c->tile_first.data(), sizeof(int) * c->tile_first.size());
    if (tile_size) memcpy(tile_size, c->tile_size.data(), sizeof(int) * c->tile_size.size());
    if (ntiles) *ntiles = (int)c->tile_first.size();
    return (int64_t)c->items.size();
}

// FLOPs of one primitive quartet of class (la lb|lc ld), la>=lb, lc>=ld:
// roots + per-root recursion + per-component transfer and product, the
// algorithmic count used for roofline reporting.
static double class_flops(int la, int lb, int lc, int ld)
{
    int L12 = la + lb, L34 = lc + ld, nr = (L12 + L34) / 2 + 1;
    double F = 40 + 40.0 * nr + nr * (12 + 21.0 * (L12 + 1) * (L34 + 1));
    auto comps = [](int l, std::vector<std::array<int, 3>> &v) {
        for (int x = l; x >= 0; --x)
            for (int y = l - x; y >= 0; --y) v.push_back({x, y, l - x - y});
    };
    std::vector<std::array<int, 3>> A, B, C, D;
    comps(la, A);
    comps(lb, B);
    comps(lc, C);
    comps(ld, D);
    for (size_t a = 0; a < A.size(); ++a)
        for (auto &b : B)
            for (size_t cc = 0; cc < C.size(); ++cc)
                for (auto &d : D) {
                    double s = 0;
                    for (int x = 0; x < 3; ++x) s += (b[x] + 1) * (d[x] + 1);
                    F += nr * (4 + 3 * s) + 12;
                }
    return F;
}

static int class_stats_impl(Ctx &c, int64_t *q36, double *f36);

int patmol_hip_class_stats(const ERIFS *e, int64_t *quartets36, double *flops36)
{
    Ctx *cp = ctx_of(e);
    if (!cp || !cp->built) return fail(PATMOL_HIP_ESTATE, "no tables");
    return class_stats_impl(*cp, quartets36, flops36);
}

int patmol_hip_stats(const ERIFS *e, PatmolHipStats *out)
{
    Ctx *cp = ctx_of(e);
    if (!cp || !cp->built || !out) return fail(PATMOL_HIP_ESTATE, "no tables");
    Ctx &c = *cp;
    memset(out, 0, sizeof(*out));
    out->nshell = c.nshell;
    out->norb = c.N;
    out->start = c.start;
    const int ns = c.nshell;
    const int64_t np = (int64_t)ns * (ns + 1) / 2;
    out->pairs = np;
    out->shell_quartets_total = np * (np + 1) / 2;
    out->work_items = (int64_t)c.items.size();
    out->work_items_fp32 = c.items_fp32;
    int64_t q36[36];
    double f36[36];
    class_stats_impl(c, q36, f36);
    int64_t kept = 0;
    double flops = 0;
    for (int i = 0; i < 36; ++i) {
        kept += q36[i];
        flops += f36[i];
    }
    out->shell_quartets_kept = kept;
    out->flops_per_build = flops;
    out->prim_quartets_kept = c.prim_kept;
    return 0;
}

static int class_stats_impl(Ctx &c, int64_t *q36, double *f36)
{
    double fl[6][6];
    int pcl[6][2] = {{0, 0}, {1, 0}, {1, 1}, {2, 0}, {2, 1}, {2, 2}};
    for (int a = 0; a < 6; ++a)
        for (int b = 0; b < 6; ++b) fl[a][b] = class_flops(pcl[a][0], pcl[a][1], pcl[b][0], pcl[b][1]);
    for (int i = 0; i < 36; ++i) {
        q36[i] = 0;
        f36[i] = 0.0;
    }
    int64_t pkept_all = 0;
    for (const ClassRange &r : c.ranges) {
        for (int it = r.first; it < r.first + r.count; ++it) {
            int4 t = c.items[it];
            bool sameBra = t.x == t.y, sameKet = t.z == t.w, samePair = t.x == t.z && t.y == t.w;
            int64_t k = 0, pk = 0;
            for (int i = 0; i < c.tile_size[t.x]; ++i) {
                int I = c.tile_first[t.x] + i;
                for (int j = 0; j < c.tile_size[t.y]; ++j) {
                    int J = c.tile_first[t.y] + j;
                    if (sameBra && J > I) continue;
                    int64_t pb = (int64_t)I * (I + 1) / 2 + J;
                    double qb = c.pair_q[pb];
                    for (int kq = 0; kq < c.tile_size[t.z]; ++kq) {
                        int K = c.tile_first[t.z] + kq;
                        for (int lq = 0; lq < c.tile_size[t.w]; ++lq) {
                            int L = c.tile_first[t.w] + lq;
                            if (sameKet && L > K) continue;
                            if (samePair && !(I > K || (I == K && J >= L))) continue;
                            int64_t pk2 = (int64_t)K * (K + 1) / 2 + L;
                            if (qb * c.pair_q[pk2] < c.tau) continue;
                            if (c.start > 0 &&
                                std::max(std::max(c.sorted_orig[I], c.sorted_orig[J]), std::max(c.sorted_orig[K], c.sorted_orig[L])) < c.start)
                                continue;
                            ++k;
                            pk += (int64_t)c.pairs[pb].npp * c.pairs[pk2].npp;
                        }
                    }
                }
            }
            q36[r.bra_pc * 6 + r.ket_pc] += k;
            f36[r.bra_pc * 6 + r.ket_pc] += pk * fl[r.bra_pc][r.ket_pc];
            pkept_all += pk;
        }
    }
    c.prim_kept = pkept_all;
    return 0;
}

// ---- molecule batches ----
PatmolHipBatch *patmol_hip_batch_new(void)
{
    g_err = PATMOL_HIP_OK;
    return new PatmolHipBatch();
}

void patmol_hip_batch_free(PatmolHipBatch *b)
{
    if (!b) return;
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    if (b->s) (void)hipStreamDestroy(b->s);
    for (auto *x : {&b->d_xyz, &b->d_q, &b->d_P, &b->d_G, &b->d_D}) x->release();
    for (auto *x : {&b->d_soff, &b->d_orig, &b->d_tile_first, &b->d_tile_size, &b->d_fperm, &b->d_finv, &b->d_item_mol, &b->d_N,
                    &b->d_fbase, &b->d_active})
        x->release();
    b->d_dstart.release();
    b->d_pstart.release();
    b->d_pairs.release();
    b->d_pp.release();
    b->d_items.release();
    b->d_mols.release();
    b->d_J.release();
    b->d_K.release();
    b->md.release();
    delete b;
}

int patmol_hip_batch_add(PatmolHipBatch *b)
{
    g_err = PATMOL_HIP_OK;
    if (!b) return fail(PATMOL_HIP_EBADARG, "patmol_hip_batch_add: null batch");
    if (b->finalized) return fail(PATMOL_HIP_ESTATE, "patmol_hip_batch_add: batch already finalised");
    {
        std::lock_guard<std::mutex> lk(g_mu);
        if (ensure_tables()) return g_err;
    }
    Ctx c;
    if (build_tables(c, 0, false)) {
        release_ctx(c);
        return g_err;
    }
    if (b->mol.empty()) b->tau = c.tau;
    BatchMol m;
    m.N = c.N;
    m.nshell = c.nshell;
    m.sbase = b->nshell;
    m.fbase = b->norb;
    m.tbase = b->ntile;
    m.pbase = (long long)b->pairs.size();
    m.ppbase = (long long)b->pp.size();
    m.dstart = b->dtot;
    m.pstart = b->no2tot;
    const int mi = (int)b->mol.size();
    b->xyz.insert(b->xyz.end(), c.shell_xyz.begin(), c.shell_xyz.end());
    b->shell_l.insert(b->shell_l.end(), c.shell_l.begin(), c.shell_l.end());
    for (int v : c.shell_soff) b->soff.push_back(v + m.fbase);
    for (int v : c.sorted_orig) b->orig.push_back(v + m.sbase);
    for (int v : c.tile_first) b->tile_first.push_back(v + m.sbase);
    b->tile_size.insert(b->tile_size.end(), c.tile_size.begin(), c.tile_size.end());
    for (const PairRef &r : c.pairs) b->pairs.push_back(PairRef{(int)(r.pp0 + m.ppbase), r.npp});
    b->pair_q.insert(b->pair_q.end(), c.pair_q.begin(), c.pair_q.end());
    b->pp.insert(b->pp.end(), c.pp.begin(), c.pp.end());
    b->fperm.insert(b->fperm.end(), c.fn_sorted_to_orig.begin(), c.fn_sorted_to_orig.end());
    b->finv.insert(b->finv.end(), c.fn_orig_to_sorted.begin(), c.fn_orig_to_sorted.end());
    for (const ClassRange &r : c.ranges)
        for (int i = r.first; i < r.first + r.count; ++i) {
            const int4 t = c.items[i];
            b->its.push_back({make_int4(t.x + m.tbase, t.y + m.tbase, t.z + m.tbase, t.w + m.tbase), mi, r.bra_pc,
                              r.ket_pc, c.item_bound[i]});
        }
    b->ddp_kept += c.ddp_kept;
    b->ddp_slots += c.ddp_slots;
    b->dtot += (long long)c.N * c.N;
    b->no2tot += (long long)c.N * (c.N + 1) / 2;
    b->nshell += c.nshell;
    b->norb += c.N;
    b->ntile += (int)c.tile_first.size();
    b->mol.push_back(m);
    release_ctx(c);
    if (b->pp.size() > (size_t)INT32_MAX) return fail(PATMOL_HIP_EUNSUPPORTED, "batch too large");
    return mi;
}

int patmol_hip_batch_finalize(PatmolHipBatch *b, int prec_mode, double threshold, double dmax)
{
    g_err = PATMOL_HIP_OK;
    if (!b || b->mol.empty()) return fail(PATMOL_HIP_EBADARG, "patmol_hip_batch_finalize: empty batch");
    if (b->finalized) return fail(PATMOL_HIP_ESTATE, "patmol_hip_batch_finalize: already finalised");
    b->md_off = b->ddp_slots > 0 && (double)b->ddp_kept > MD_DDDP_MAX_KEPT * (double)b->ddp_slots ? 2 : 0;
    // per class: FP64 items first, each part by bound descending (the work
    // list of every molecule merged into one queue per class)
    auto fp32 = [&](const PatmolHipBatch::It &x) {
        return prec_mode == 1 && x.bound * dmax < threshold && fp32_prec(x.bpc, x.kpc) == 1;
    };
    std::stable_sort(b->its.begin(), b->its.end(), [&](const PatmolHipBatch::It &x, const PatmolHipBatch::It &y) {
        if (x.bpc != y.bpc) return x.bpc > y.bpc;
        if (x.kpc != y.kpc) return x.kpc > y.kpc;
        const bool fx = fp32(x), fy = fp32(y);
        if (fx != fy) return !fx;
        return x.bound > y.bound;
    });
    std::vector<int4> items(b->its.size());
    std::vector<int> imol(b->its.size());
    b->ranges.clear();
    b->n64.clear();
    b->items_fp32 = 0;
    for (size_t i = 0; i < b->its.size(); ++i) {
        items[i] = b->its[i].t;
        imol[i] = b->its[i].mol;
        if (b->ranges.empty() || b->ranges.back().bra_pc != b->its[i].bpc || b->ranges.back().ket_pc != b->its[i].kpc) {
            b->ranges.push_back({b->its[i].bpc, b->its[i].kpc, (int)i, 0});
            b->n64.push_back(0);
        }
        b->ranges.back().count++;
        if (fp32(b->its[i])) ++b->items_fp32;
        else ++b->n64.back();
    }
    const int nm = (int)b->mol.size();
    std::vector<MolDesc> md(nm);
    std::vector<int> Nv(nm), fb(nm);
    std::vector<long long> ds(nm), ps(nm);
    b->Nvec.resize(nm);
    b->pstart_v.resize(nm);
    for (int i = 0; i < nm; ++i) {
        const BatchMol &m = b->mol[i];
        md[i] = MolDesc{m.N, m.sbase, m.pbase, m.dstart - (long long)m.fbase * m.N - m.fbase};
        Nv[i] = m.N;
        fb[i] = m.fbase;
        ds[i] = m.dstart;
        ps[i] = m.pstart;
        b->Nvec[i] = m.N;
        b->pstart_v[i] = m.pstart;
    }
    if (b->d_xyz.upload(b->xyz) || b->d_q.upload(b->pair_q) || b->d_soff.upload(b->soff) || b->d_orig.upload(b->orig) ||
        b->d_tile_first.upload(b->tile_first) || b->d_tile_size.upload(b->tile_size) || b->d_fperm.upload(b->fperm) ||
        b->d_finv.upload(b->finv) || b->d_item_mol.upload(imol) || b->d_N.upload(Nv) || b->d_fbase.upload(fb) ||
        b->d_dstart.upload(ds) || b->d_pstart.upload(ps) || b->d_pairs.upload(b->pairs) || b->d_pp.upload(b->pp) ||
        b->d_items.upload(items) || b->d_mols.upload(md) || b->d_active.upload(std::vector<int>(nm, 1)) ||
        b->d_P.alloc(b->no2tot) || b->d_G.alloc(b->no2tot) ||
        b->d_D.alloc(b->dtot) || b->d_J.alloc((size_t)b->dtot * PM_XCD_SLOTS) ||
        b->d_K.alloc((size_t)b->dtot * PM_XCD_SLOTS))
        return g_err;
    {
        // Hermite tables over every molecule's shell pairs (batch-global
        // pair, shell and primitive-pair indices)
        std::vector<MdPair> plist;
        for (const BatchMol &m : b->mol)
            for (int I = 0; I < m.nshell; ++I)
                for (int J = 0; J <= I; ++J)
                    plist.push_back({m.pbase + (int64_t)I * (I + 1) / 2 + J, m.sbase + I, m.sbase + J});
        if (build_md_tables(b->md, plist, b->shell_l, b->pairs, b->pp.size(), b->d_xyz.p, b->d_pairs.p, b->d_pp.p))
            return g_err;
    }
    HIPCHK(hipMemset(b->d_P.p, 0, sizeof(double) * b->no2tot));
    HIPCHK(hipStreamCreateWithFlags(&b->s, hipStreamNonBlocking));
    // capture the whole G build once
    HIPCHK(hipStreamBeginCapture(b->s, hipStreamCaptureModeThreadLocal));
    const int rc = batch_build(*b, b->s);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(b->s, &g);
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return g_err;
    }
    if (e != hipSuccess) return fail(PATMOL_HIP_EDEVICE, "hipStreamEndCapture: %s", hipGetErrorString(e));
    b->graph = g;
    HIPCHK(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    b->finalized = true;
    return 0;
}

// G += 2J[P] - K[P] for every molecule from host arrays (the molecules'
// packed triangles concatenated in the order they were added); by the
// captured graph, or (use_graph 0) by the same launches issued directly
int patmol_hip_batch_gmatrix(PatmolHipBatch *b, const double *P, double *G, int use_graph)
{
    g_err = PATMOL_HIP_OK;
    if (!b || !P || !G) return fail(PATMOL_HIP_EBADARG, "patmol_hip_batch_gmatrix: bad argument");
    if (!b->finalized) return fail(PATMOL_HIP_ESTATE, "patmol_hip_batch_gmatrix: batch not finalised");
    HIPCHK(hipMemcpyAsync(b->d_P.p, P, sizeof(double) * b->no2tot, hipMemcpyHostToDevice, b->s));
    if (use_graph) {
        HIPCHK(hipGraphLaunch(b->exec, b->s));
    } else if (batch_build(*b, b->s)) {
        return g_err;
    }
    std::vector<double> g(b->no2tot);
    HIPCHK(hipMemcpyAsync(g.data(), b->d_G.p, sizeof(double) * b->no2tot, hipMemcpyDeviceToHost, b->s));
    HIPCHK(hipStreamSynchronize(b->s));
    for (long long i = 0; i < b->no2tot; ++i) G[i] += g[i];
    return 0;
}

// device-resident use: the batch's own P (input) and G (output) buffers, and
// one replay of the captured G build on `stream`
int patmol_hip_batch_buffers(PatmolHipBatch *b, double **d_P, double **d_G)
{
    if (!b || !b->finalized) return fail(PATMOL_HIP_ESTATE, "batch not finalised");
    if (d_P) *d_P = b->d_P.p;
    if (d_G) *d_G = b->d_G.p;
    return 0;
}

int patmol_hip_batch_launch(PatmolHipBatch *b, void *stream)
{
    g_err = PATMOL_HIP_OK;
    return batch_launch(b, (hipStream_t)stream);
}

int patmol_hip_batch_info(PatmolHipBatch *b, int *nmol, int64_t *no2tot, int64_t *items, int64_t *items_fp32)
{
    if (!b) return fail(PATMOL_HIP_EBADARG, "null batch");
    if (nmol) *nmol = (int)b->mol.size();
    if (no2tot) *no2tot = b->no2tot;
    if (items) *items = (int64_t)b->its.size();
    if (items_fp32) *items_fp32 = b->items_fp32;
    return 0;
}

// Two-electron part of the RHF nuclear gradient over the handle's screened
// work list (grad.hip): grad[3 c + x] += sum Gamma d(ab|cd)/dR_cx for
// centre c, with the per-shell fixed-point sums converted and added per
// centre in shell order on the host (reproducible).
int patmol_hip_eri_gradient(const ERIFS *e, const double *Pmat, double *grad)
{
    g_err = PATMOL_HIP_OK;
    Ctx *cp = ctx_of(e);
    if (!cp || !Pmat || !grad) return fail(PATMOL_HIP_EBADARG, "patmol_hip_eri_gradient: bad handle or null array");
    if (!cp->built) return fail(PATMOL_HIP_ESTATE, "two_electron_ints was not run for this handle");
    Ctx &c = *cp;
    const int64_t NN = (int64_t)c.N * c.N, no2 = (int64_t)c.N * (c.N + 1) / 2;
    if (alloc_work(c, false)) return g_err;
    DevBuf<FxAcc> gsh;
    if (gsh.alloc((size_t)3 * c.nshell)) return g_err;
    HIPCHK(hipMemcpy(c.d_P.p, Pmat, sizeof(double) * no2, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(unpack_kernel, dim3((NN + 255) / 256), dim3(256), 0, 0, c.d_P.p, c.d_D.p, c.d_fperm.p, c.N);
    hipLaunchKernelGGL(zero_fx_kernel, dim3((unsigned)((3 * c.nshell + 255) / 256)), dim3(256), 0, 0, gsh.p,
                       (int64_t)3 * c.nshell);
    for (const ClassRange &r : c.ranges) {
        GradArgs a{c.d_items.p + r.first, r.count, c.d_tile_first.p, c.d_tile_size.p, c.d_xyz.p, c.d_soff.p,
                   c.d_pairs.p, c.d_q.p, c.d_pp.p, c.d_ppab.p, c.d_D.p, gsh.p, c.N, c.tau, g_rys};
        if (launch_grad(r.bra_pc, r.ket_pc, a, 0))
            return fail(PATMOL_HIP_EUNSUPPORTED, "no gradient kernel for class %d/%d", r.bra_pc, r.ket_pc);
        HIPCHK(hipGetLastError());
    }
    std::vector<FxAcc> h((size_t)3 * c.nshell);
    HIPCHK(hipMemcpy(h.data(), gsh.p, sizeof(FxAcc) * h.size(), hipMemcpyDeviceToHost));
    gsh.release();
    const Shell *sh = shells();
    for (int sidx = 0; sidx < c.nshell; ++sidx) {
        const int cen = sh[c.sorted_orig[sidx]].cen;
        for (int x = 0; x < 3; ++x) grad[3 * cen + x] += fx_value(h[3 * sidx + x].hi, h[3 * sidx + x].lo);
    }
    return 0;
}

int patmol_hip_eri_blocks(const ERIFS *e, int nq, const int *quartets, double *out)
{
    return patmol_hip_eri_blocks_prec(e, nq, quartets, out, 0);
}

int patmol_hip_eri_blocks_prec(const ERIFS *e, int nq, const int *quartets, double *out, int prec)
{
    g_err = PATMOL_HIP_OK;
    Ctx *cp = ctx_of(e);
    if (!cp || !cp->built) return fail(PATMOL_HIP_ESTATE, "no tables");
    Ctx &c = *cp;
    if (nq <= 0) return 0;
    struct Q {
        int s[4];      // sorted indices in kernel orientation
        int perm[4];   // kernel position of requested position
        int bpc, kpc;
        size_t out_off, dev_off;
    };
    std::vector<Q> qs(nq);
    size_t total = 0;
    for (int q = 0; q < nq; ++q) {
        int o[4];
        for (int k = 0; k < 4; ++k) {
            o[k] = quartets[4 * q + k];
            if (o[k] < 0 || o[k] >= c.nshell) return fail(PATMOL_HIP_EBADARG, "shell index out of range");
        }
        int s[4] = {c.orig_to_sorted[o[0]], c.orig_to_sorted[o[1]], c.orig_to_sorted[o[2]], c.orig_to_sorted[o[3]]};
        int pos[4] = {0, 1, 2, 3};  // kernel slot -> requested slot
        auto lo = [&](int slot) { return c.shell_l[s[pos[slot]]]; };
        if (lo(0) < lo(1)) std::swap(pos[0], pos[1]);
        if (lo(2) < lo(3)) std::swap(pos[2], pos[3]);
        int bpc = pair_class(lo(0), lo(1)), kpc = pair_class(lo(2), lo(3));
        if (bpc < kpc) {
            std::swap(pos[0], pos[2]);
            std::swap(pos[1], pos[3]);
            std::swap(bpc, kpc);
        }
        Q &x = qs[q];
        for (int k = 0; k < 4; ++k) {
            x.s[k] = s[pos[k]];
            x.perm[pos[k]] = k;
        }
        x.bpc = bpc;
        x.kpc = kpc;
        x.out_off = total;
        total += (size_t)c.shell_nst[s[0]] * c.shell_nst[s[1]] * c.shell_nst[s[2]] * c.shell_nst[s[3]];
    }
    // group by class
    std::vector<double> res(total);
    for (int b = 0; b < 6; ++b)
        for (int k = 0; k <= b; ++k) {
            std::vector<int> idx;
            for (int q = 0; q < nq; ++q)
                if (qs[q].bpc == b && qs[q].kpc == k) idx.push_back(q);
            if (idx.empty()) continue;
            std::vector<int> qv;
            for (int q : idx)
                for (int t = 0; t < 4; ++t) qv.push_back(qs[q].s[t]);
            size_t ncomp = (size_t)c.shell_nst[qs[idx[0]].s[0]] * c.shell_nst[qs[idx[0]].s[1]] * c.shell_nst[qs[idx[0]].s[2]] *
                           c.shell_nst[qs[idx[0]].s[3]];
            DevBuf<int> dq;
            DevBuf<double> dout;
            if (dq.upload(qv) || dout.alloc(ncomp * idx.size())) return g_err;
            EriOutArgs a{dq.p, (int)idx.size(), c.d_xyz.p, c.d_pairs.p, c.d_pp.p, dout.p, g_rys, prec ? 1 : 0};
            if (launch_eri_out(b, k, a, 0)) return fail(PATMOL_HIP_EUNSUPPORTED, "no ERI kernel");
            HIPCHK(hipGetLastError());
            std::vector<double> h(ncomp * idx.size());
            HIPCHK(hipMemcpy(h.data(), dout.p, sizeof(double) * h.size(), hipMemcpyDeviceToHost));
            dq.release();
            dout.release();
            for (size_t m = 0; m < idx.size(); ++m) {
                const Q &x = qs[idx[m]];
                int n[4] = {c.shell_nst[x.s[0]], c.shell_nst[x.s[1]], c.shell_nst[x.s[2]], c.shell_nst[x.s[3]]};
                // requested dims
                int rn[4];
                for (int r = 0; r < 4; ++r) rn[r] = n[x.perm[r]];
                const double *src = h.data() + m * ncomp;
                int ki[4];
                for (ki[0] = 0; ki[0] < n[0]; ++ki[0])
                    for (ki[1] = 0; ki[1] < n[1]; ++ki[1])
                        for (ki[2] = 0; ki[2] < n[2]; ++ki[2])
                            for (ki[3] = 0; ki[3] < n[3]; ++ki[3]) {
                                int r0 = ki[x.perm[0]], r1 = ki[x.perm[1]], r2 = ki[x.perm[2]], r3 = ki[x.perm[3]];
                                size_t dst = ((size_t)(r0 * rn[1] + r1) * rn[2] + r2) * rn[3] + r3;
                                res[x.out_off + dst] = src[((ki[0] * n[1] + ki[1]) * n[2] + ki[2]) * n[3] + ki[3]];
                            }
            }
        }
    memcpy(out, res.data(), sizeof(double) * total);
    return 0;
}

}  // extern "C"

namespace pm {
int rhf_gmatrix_delta(const ERIFS *e, const double *dP, double *dG, hipStream_t s, BuildCounts *counts)
{
    g_err = PATMOL_HIP_OK;
    Ctx *c = ctx_of(e);
    if (!c || !dP || !dG) return fail(PATMOL_HIP_EBADARG, "bad handle or null pointer");
    if (!c->built) return fail(PATMOL_HIP_ESTATE, "two_electron_ints was not run for this handle");
    if (rhf_device(*c, dP, dG, s, true)) return g_err;
    return last_build_counts(e, counts);
}

int last_build_counts(const ERIFS *e, BuildCounts *counts)
{
    Ctx *c = ctx_of(e);
    if (!counts) return 0;
    *counts = BuildCounts();
    if (!c) return 0;
    counts->items = (int64_t)c->items.size();
    counts->fp32 = c->items_fp32;
    counts->skipped = c->items_skipped;
    return 0;
}

int precision_mode() { return g_prec_mode; }
void set_precision_mode(int mode) { g_prec_mode = mode ? 1 : 0; }
}  // namespace pm
