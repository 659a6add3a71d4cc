// host_basis.cpp -- basis model of libpatmol_host.so.
//
// Provides the reference's basis.h accessor API (basis.c:21-172) that the
// hot path reads through, plus two loaders: init_Basis() reads patin.dat in
// the positional format of basis.c:243-321 / shell.c:25-43 / center.c:131-136,
// and patmol_basis_load() takes the same data as arrays.  Contraction
// coefficients are normalised as shell.c:47-77 does; Cartesian components are
// ordered and weighted as basis.c:294-319.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "patmol_host.h"

namespace {

struct BasisState {
    int norb = 0, nshell = 0, ncen = 0;
    int tnorb = 0, tnshell = 0, tncen = 0;
    int dpm = 0, ffc = 0, scf_a = 0, scf_g = 0;
    int prt_e = 0, prt_p = 0, prt_d = 0;
    int nelec = 0, spinm = 1, max_scf_its = 0;
    double scf_eps = 1e-8;
    int skipcen = 0;
    int maxlsh = 0, maxlst = 1;
    std::vector<Shell> shell;
    std::vector<Center> center;
    std::vector<std::vector<double>> al, co;
    int ***lxyz = nullptr;
    double **normFact = nullptr;
    int tab_l = -1;  // highest l of the allocated lxyz/normFact tables
} B;

const double kDfact[] = {1., 1., 3., 15., 105., 945., 10395.};
constexpr int DPM_NSHELL = 7, DPM_NORB = 13;

void normalize_shell(Shell &sh)
{
    const double twofact = 2.8284271247461903;  // 2 sqrt 2
    const double piterm = 5.568327996831707;    // pi^(3/2)
    double lpow = 1.5 + sh.lsh, sum = 0.0;
    for (int i = 0; i < sh.npr; ++i)
        for (int j = 0; j < sh.npr; ++j)
            sum += sh.co[i] * sh.co[j] * pow(sqrt(sh.al[i] * sh.al[j]) / (sh.al[i] + sh.al[j]), lpow);
    sum = 1.0 / sqrt(sum * twofact);
    for (int i = 0; i < sh.npr; ++i) sh.co[i] *= sum * sqrt(pow(2.0 * sh.al[i], lpow) / piterm);
}

void free_tables()
{
    if (B.lxyz) {
        for (int l = 0; l <= B.tab_l; ++l) {
            free(B.lxyz[l][0]);
            free(B.lxyz[l]);
            free(B.normFact[l]);
        }
        free(B.lxyz);
        free(B.normFact);
    }
    B.lxyz = nullptr;
    B.normFact = nullptr;
    B.tab_l = -1;
}

void build_tables()
{
    free_tables();
    B.maxlst = (B.maxlsh + 1) * (B.maxlsh + 2) / 2;
    B.tab_l = B.maxlsh;
    B.lxyz = (int ***)malloc(sizeof(int **) * (B.maxlsh + 1));
    B.normFact = (double **)malloc(sizeof(double *) * (B.maxlsh + 1));
    for (int l = 0; l <= B.maxlsh; ++l) {
        int n = (l + 1) * (l + 2) / 2, k = 0;
        B.lxyz[l] = (int **)malloc(sizeof(int *) * n);
        int *blk = (int *)malloc(sizeof(int) * 3 * n);
        B.normFact[l] = (double *)malloc(sizeof(double) * n);
        for (int lx = l; lx >= 0; --lx)
            for (int ly = l - lx; ly >= 0; --ly) {
                B.lxyz[l][k] = blk + 3 * k;
                B.lxyz[l][k][0] = lx;
                B.lxyz[l][k][1] = ly;
                B.lxyz[l][k][2] = l - lx - ly;
                B.normFact[l][k] = 1.0 / sqrt(kDfact[lx] * kDfact[ly] * kDfact[l - lx - ly]);
                ++k;
            }
    }
}

void relink()
{
    for (size_t s = 0; s < B.shell.size(); ++s) {
        B.shell[s].al = B.al[s].data();
        B.shell[s].co = B.co[s].data();
    }
}

void add_shell(int npr, int l, int cen, const double *al, const double *co, int &off, bool normalize)
{
    Shell sh;
    sh.npr = npr;
    sh.lsh = l;
    sh.cen = cen;
    sh.off = off;
    sh.nst = (l + 1) * (l + 2) / 2;
    off += sh.nst;
    B.al.emplace_back(al, al + npr);
    B.co.emplace_back(co, co + npr);
    B.shell.push_back(sh);
    relink();
    if (normalize) normalize_shell(B.shell.back());
    if (l > B.maxlsh) B.maxlsh = l;
}

}  // namespace

extern "C" {

int number_of_orbitals(void) { return B.norb; }
int number_of_shells(void) { return B.nshell; }
int number_of_centers(void) { return B.ncen; }
// basis.c:36-50 returns the current counts for the "total" queries too.
int total_number_of_orbitals(void) { return B.norb; }
int total_number_of_shells(void) { return B.nshell; }
int total_number_of_centers(void) { return B.ncen; }
int number_of_electrons(void) { return B.nelec; }
int spin_multiplicity(void) { return B.spinm; }
int max_scf_iterations(void) { return B.max_scf_its; }
int skip_center(void) { return B.skipcen; }
int max_lvalue(void) { return B.maxlsh; }
int max_lstates(void) { return B.maxlst; }
int do_dpm_calculation(void) { return B.dpm; }
int do_finite_field_calculation(void) { return B.ffc; }
int scf_acceleration(void) { return B.scf_a; }
int scf_guess(void) { return B.scf_g; }
double scf_convergence(void) { return B.scf_eps; }
const Shell *shells(void) { return B.shell.data(); }
const Center *centers(void) { return B.center.data(); }
int ***lvector_components(void) { return B.lxyz; }
double **normalizationFactors(void) { return B.normFact; }

void set_center_position(int icen, const double *r) { memcpy(B.center[icen].r, r, 3 * sizeof(double)); }
void set_center_charge(int icen, double q) { B.center[icen].q = q; }

// basis.c:174-195
double nuclear_repulsion(void)
{
    double sum = 0.0;
    for (int i = 0; i < B.ncen; ++i)
        for (int j = i + 1; j < B.ncen; ++j) {
            double x = B.center[i].r[0] - B.center[j].r[0];
            double y = B.center[i].r[1] - B.center[j].r[1];
            double z = B.center[i].r[2] - B.center[j].r[2];
            sum += B.center[i].q * B.center[j].q / sqrt(x * x + y * y + z * z);
        }
    return sum;
}

int patmol_basis_load(int ncen, const double *q, const double *xyz, int nshell, const int *npr, const int *lsh,
                      const int *cen, const double *al, const double *co, int nelec, int mult, int maxits,
                      double eps, int scf_a)
{
    if (ncen <= 0 || nshell <= 0) return 1;
    B.center.assign(ncen, Center{});
    for (int c = 0; c < ncen; ++c) {
        B.center[c].q = q[c];
        for (int x = 0; x < 3; ++x) B.center[c].r[x] = xyz[3 * c + x];
    }
    B.shell.clear();
    B.al.clear();
    B.co.clear();
    B.shell.reserve(nshell + DPM_NSHELL);
    B.al.reserve(nshell + DPM_NSHELL);
    B.co.reserve(nshell + DPM_NSHELL);
    B.maxlsh = 0;
    int off = 0, p = 0;
    for (int s = 0; s < nshell; ++s) {
        if (lsh[s] < 0 || lsh[s] > 4 || cen[s] < 0 || cen[s] >= ncen || npr[s] <= 0) return 2;
        add_shell(npr[s], lsh[s], cen[s], al + p, co + p, off, true);
        p += npr[s];
    }
    B.ncen = B.tncen = ncen;
    B.nshell = B.tnshell = nshell;
    B.norb = B.tnorb = off;
    B.skipcen = ncen;
    B.nelec = nelec;
    B.spinm = mult;
    B.max_scf_its = maxits;
    B.scf_eps = eps;
    B.scf_a = scf_a;
    B.dpm = 0;
    build_tables();
    return 0;
}

// basis.c:243-321
void init_Basis(void)
{
    FILE *f = fopen("patin.dat", "r");
    if (!f) {
        fprintf(stderr, "FATAL ERROR: could not open patin.dat\n");
        exit(EXIT_FAILURE);
    }
    int nshell, norb, ncen, dftx;
    int ok = fscanf(f, "%d %d %d", &nshell, &norb, &ncen);
    ok += fscanf(f, "%d %le", &B.max_scf_its, &B.scf_eps);
    ok += fscanf(f, "%d %d", &B.nelec, &B.spinm);
    ok += fscanf(f, "%d %d %d %d", &B.dpm, &dftx, &dftx, &dftx);
    ok += fscanf(f, "%d %d %d", &B.ffc, &B.scf_a, &B.scf_g);
    ok += fscanf(f, "%d %d %d", &B.prt_e, &B.prt_p, &B.prt_d);
    std::vector<double> q(ncen), xyz(3 * ncen);
    for (int c = 0; c < ncen; ++c) ok += fscanf(f, "%lg %lg %lg %lg", &q[c], &xyz[3 * c], &xyz[3 * c + 1], &xyz[3 * c + 2]);
    std::vector<int> npr(nshell), lsh(nshell), cen(nshell);
    std::vector<double> al, co;
    for (int s = 0; s < nshell; ++s) {
        ok += fscanf(f, "%d %d %d", &npr[s], &lsh[s], &cen[s]);
        for (int k = 0; k < npr[s]; ++k) {
            double a, c;
            ok += fscanf(f, "%lg %lg", &a, &c);
            al.push_back(a);
            co.push_back(c);
        }
    }
    fclose(f);
    int dpm = B.dpm, ffc = B.ffc, g = B.scf_g;
    if (patmol_basis_load(ncen, q.data(), xyz.data(), nshell, npr.data(), lsh.data(), cen.data(), al.data(), co.data(),
                          B.nelec, B.spinm, B.max_scf_its, B.scf_eps, B.scf_a)) {
        fprintf(stderr, "FATAL ERROR: bad basis in patin.dat\n");
        exit(EXIT_FAILURE);
    }
    B.dpm = dpm;
    B.ffc = ffc;
    B.scf_g = g;
    if (B.dpm) patmol_basis_add_dpm_center();
}

// Distributed-positron-model centre and shells, basis.c:220-241 and 280-284:
// one extra centre of charge +1 carrying 4 s and 3 p unnormalised shells with
// unit coefficients.  The shells become visible after dpm_augment().
int patmol_basis_add_dpm_center(void)
{
    static const double alf[] = {13.3615, 2.0133, 0.453757, 1.2331, 2.5719, 0.59976, 0.139365};
    static const int lsh[] = {0, 0, 0, 0, 1, 1, 1};
    Center c{};
    c.q = 1.0;
    B.center.push_back(c);
    B.tncen = B.ncen + 1;
    int off = B.norb;
    for (int s = 0; s < DPM_NSHELL; ++s) {
        double one = 1.0;
        add_shell(1, lsh[s], B.ncen, &alf[s], &one, off, false);
    }
    B.tnshell = B.nshell + DPM_NSHELL;
    B.tnorb = B.norb + DPM_NORB;
    B.skipcen = B.ncen;
    if (B.maxlsh < 1) B.maxlsh = 1;
    build_tables();
    return 0;
}

// basis.c:323-329
void dpm_augment(void)
{
    B.norb = B.tnorb;
    B.nshell = B.tnshell;
    B.ncen = B.tncen;
}

}  // extern "C"
