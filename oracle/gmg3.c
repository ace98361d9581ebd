/*
 * CPU oracle, C restatement of the 3D multigrid cycle of oracle/gmg.py
 * (jsolver/solver_3d.py:588-706).  TEST INFRASTRUCTURE ONLY: used by tests/
 * to cross-check the NumPy oracle and by bench.py's cpu_baseline leg as the
 * timed CPU baseline (OpenMP over planes).  Nothing in the product links it.
 *
 * The operator (per-level face and centre coefficients) comes from the NumPy
 * oracle's build_operator; this file runs cycles on it with the same
 * floating-point operations in the same order, so results are bit-identical
 * for any thread count (no reductions).  Build with -ffp-contract=off.
 *
 * Layout: phi and rhs of level l are (mz+3, my+3, mx+3) row-major doubles;
 * coefficients are (mz+1, my+1, mx+1) with entry k belonging to phi index
 * k+1 (solver_3d.py:413-422).  Array axis 0 is z, 1 is y, 2 is x.
 */
#include <omp.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
    int m[3];               /* mz, my, mx */
    const double* am[3];    /* minus-face coefficient per array axis */
    const double* ap[3];    /* plus-face coefficient per array axis */
    const double* ac;       /* centre */
    double* phi;            /* work arrays of this level (padded) */
    double* rhs;
    double* res;            /* residual, padded, rim zero */
} Lv;

typedef struct {
    int levels, nu1, nu2, gamma;
    Lv* lv;
} Op;

static long pidx(const Lv* L, int z, int y, int x)
{
    return ((long)z * (L->m[1] + 3) + y) * (L->m[2] + 3) + x;
}
static long cidx(const Lv* L, int z, int y, int x)   /* coefficient of phi point (z, y, x) */
{
    return ((long)(z - 1) * (L->m[1] + 1) + (y - 1)) * (L->m[2] + 1) + (x - 1);
}

/* sub-colours in (x, y, z) order, solver_3d.py:693 */
static const int SUB3[8][3] = {{0, 0, 0}, {0, 1, 1}, {1, 1, 0}, {1, 0, 1},
                               {1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {1, 1, 1}};

/* one RBGS sweep (oracle smooth): the eight sub-colours in order; boundary
 * nodes and ghosts stay zero (the oracle's zero_rim after each sub-colour) */
static void smooth(const Lv* L)
{
    double* phi = L->phi;
    const double* rhs = L->rhs;
    const long sy = L->m[2] + 3, sz = (long)(L->m[1] + 3) * sy;
    for (int c = 0; c < 8; ++c) {
        const int lz = SUB3[c][2], ly = SUB3[c][1], lx = SUB3[c][0];
#pragma omp parallel for schedule(static)
        for (int z = lz + 1; z <= L->m[0]; z += 2) {
            if (z < 2) continue;
            for (int y = ly + 1; y <= L->m[1]; y += 2) {
                if (y < 2) continue;
                for (int x = lx + 1; x <= L->m[2]; x += 2) {
                    if (x < 2) continue;
                    const long p = pidx(L, z, y, x), q = cidx(L, z, y, x);
                    double acc = L->ap[0][q] * phi[p + sz];
                    acc = acc + L->am[0][q] * phi[p - sz];
                    acc = acc + L->ap[1][q] * phi[p + sy];
                    acc = acc + L->am[1][q] * phi[p - sy];
                    acc = acc + L->ap[2][q] * phi[p + 1];
                    acc = acc + L->am[2][q] * phi[p - 1];
                    acc = acc - rhs[p];
                    phi[p] = acc / L->ac[q];
                }
            }
        }
    }
}

/* residual on the unknowns (oracle residual_window, then zero_rim) */
static void residual(const Lv* L)
{
    const double* phi = L->phi;
    const long sy = L->m[2] + 3, sz = (long)(L->m[1] + 3) * sy;
#pragma omp parallel for schedule(static)
    for (int z = 2; z <= L->m[0]; ++z)
        for (int y = 2; y <= L->m[1]; ++y)
            for (int x = 2; x <= L->m[2]; ++x) {
                const long p = pidx(L, z, y, x), q = cidx(L, z, y, x);
                double r = L->ac[q] * phi[p];
                r = r - L->am[2][q] * phi[p - 1];
                r = r - L->ap[2][q] * phi[p + 1];
                r = r - L->am[1][q] * phi[p - sy];
                r = r - L->ap[1][q] * phi[p + sy];
                r = r - L->am[0][q] * phi[p - sz];
                r = r - L->ap[0][q] * phi[p + sz];
                r = r + L->rhs[p];
                L->res[p] = r;
            }
}

/* restriction of the residual into the coarse rhs (oracle restrict) */
static void restrict_(const Lv* F, Lv* C)
{
    residual(F);
    const double* f = F->res;
#pragma omp parallel for schedule(static)
    for (int I = 2; I <= C->m[0]; ++I)
        for (int J = 2; J <= C->m[1]; ++J)
            for (int K = 2; K <= C->m[2]; ++K) {
                const int z = 2 * I - 1, y = 2 * J - 1, x = 2 * K - 1;
#define R(dz, dy, dx) f[pidx(F, z + (dz), y + (dy), x + (dx))]
                double e = R(0, 1, 1) + R(0, 1, -1);
                e = e + R(0, -1, 1);
                e = e + R(0, -1, -1);
                e = e + R(1, 0, 1);
                e = e + R(1, 0, -1);
                e = e + R(-1, 0, 1);
                e = e + R(-1, 0, -1);
                e = e + R(1, 1, 0);
                e = e + R(1, -1, 0);
                e = e + R(-1, 1, 0);
                e = e + R(-1, -1, 0);
                C->rhs[pidx(C, I, J, K)] = 0.03125 * e + 0.125 * R(0, 0, 0);
#undef R
            }
}

/* phi_f += P phi_c (oracle prolong): corrections from coarse points only */
static void prolong_add(Lv* F, const Lv* C, int nu2)
{
#pragma omp parallel for schedule(static)
    for (int z = 2; z <= F->m[0]; ++z)
        for (int y = 2; y <= F->m[1]; ++y)
            for (int x = 2; x <= F->m[2]; ++x) {
                const int oz = z & 1, oy = y & 1, ox = x & 1;
                const int nodd = oz + oy + ox;
                /* coarse index of odd fine index i: (i + 1) / 2; of even i +- 1 likewise */
#define CV(fz, fy, fx) C->phi[pidx(C, ((fz) + 1) / 2, ((fy) + 1) / 2, ((fx) + 1) / 2)]
                double corr;
                if (nodd == 3) {
                    corr = CV(z, y, x);
                } else if (nodd == 2) {
                    if (!ox) corr = 0.5 * (CV(z, y, x + 1) + CV(z, y, x - 1));
                    else if (!oy) corr = 0.5 * (CV(z, y + 1, x) + CV(z, y - 1, x));
                    else corr = 0.5 * (CV(z + 1, y, x) + CV(z - 1, y, x));
                } else if (nodd == 1) {
                    if (nu2 > 0) continue;
                    double s;
                    if (oz) {
                        s = CV(z, y + 1, x + 1) + CV(z, y - 1, x + 1);
                        s = s + CV(z, y + 1, x - 1);
                        s = s + CV(z, y - 1, x - 1);
                    } else if (oy) {
                        s = CV(z + 1, y, x + 1) + CV(z - 1, y, x + 1);
                        s = s + CV(z + 1, y, x - 1);
                        s = s + CV(z - 1, y, x - 1);
                    } else {
                        s = CV(z + 1, y + 1, x) + CV(z - 1, y + 1, x);
                        s = s + CV(z + 1, y - 1, x);
                        s = s + CV(z - 1, y - 1, x);
                    }
                    corr = 0.25 * s;
                } else {
                    double s = CV(z + 1, y + 1, x + 1) + CV(z - 1, y + 1, x + 1);
                    s = s + CV(z + 1, y - 1, x + 1);
                    s = s + CV(z - 1, y - 1, x + 1);
                    s = s + CV(z + 1, y + 1, x - 1);
                    s = s + CV(z - 1, y + 1, x - 1);
                    s = s + CV(z + 1, y - 1, x - 1);
                    s = s + CV(z - 1, y - 1, x - 1);
                    corr = 0.125 * s;
                }
#undef CV
                const long p = pidx(F, z, y, x);
                F->phi[p] = F->phi[p] + corr;
            }
}

static long npad(const Lv* L) { return (long)(L->m[0] + 3) * (L->m[1] + 3) * (L->m[2] + 3); }

/* oracle cycle (solver_3d.py:588-620) */
static void cycle(Op* O, int l, int fcyc, int skip_pre)
{
    Lv* L = &O->lv[l];
    if (l == O->levels - 2) {
        for (int i = 0; i < O->nu1 + O->nu2; ++i) smooth(L);
        return;
    }
    if (!skip_pre)
        for (int i = 0; i < O->nu1; ++i) smooth(L);
    Lv* C = &O->lv[l + 1];
    memset(C->rhs, 0, sizeof(double) * npad(C));
    restrict_(L, C);
    memset(C->phi, 0, sizeof(double) * npad(C));
    if (fcyc)
        for (int i = 0; i < O->gamma; ++i) cycle(O, l + 1, 1, 0);
    for (int i = 0; i < O->gamma; ++i) cycle(O, l + 1, 0, 0);
    prolong_add(L, C, O->nu2);
    for (int i = 0; i < O->nu2; ++i) smooth(L);
}

/* OpenMP threads of the following calls */
void gmg3_set_threads(int n)
{
    if (n > 0) omp_set_num_threads(n);
}

/*
 * K cycles on phi (in place).  ms: levels x (mz, my, mx); am, ap: levels x 3
 * coefficient pointers; ac: levels pointers.  Returns 0, or -1 on allocation
 * failure.
 */
int gmg3_cycles(int levels, const int* ms, const double* const* am, const double* const* ap,
                const double* const* ac, double* phi, const double* rhs, int ncycles, int nu1, int nu2,
                int gamma, int fcycle, int skip_pre)
{
    Op O = {levels, nu1, nu2, gamma, NULL};
    O.lv = (Lv*)calloc((size_t)levels, sizeof(Lv));
    if (!O.lv) return -1;
    int rc = 0;
    for (int l = 0; l < levels && rc == 0; ++l) {
        Lv* L = &O.lv[l];
        for (int a = 0; a < 3; ++a) {
            L->m[a] = ms[3 * l + a];
            L->am[a] = am[3 * l + a];
            L->ap[a] = ap[3 * l + a];
        }
        L->ac = ac[l];
        const long n = npad(L);
        L->phi = l == 0 ? phi : (double*)calloc((size_t)n, sizeof(double));
        L->rhs = l == 0 ? (double*)rhs : (double*)calloc((size_t)n, sizeof(double));
        L->res = (double*)calloc((size_t)n, sizeof(double));
        if (!L->phi || !L->rhs || !L->res) rc = -1;
    }
    if (rc == 0) {
        /* the cycle input passes do_BCs: rim of level 0 zero */
        Lv* L0 = &O.lv[0];
        for (int z = 0; z < L0->m[0] + 3; ++z)
            for (int y = 0; y < L0->m[1] + 3; ++y)
                for (int x = 0; x < L0->m[2] + 3; ++x)
                    if (z < 2 || y < 2 || x < 2 || z > L0->m[0] || y > L0->m[1] || x > L0->m[2])
                        phi[pidx(L0, z, y, x)] = 0.0;
        for (int k = 0; k < ncycles; ++k) cycle(&O, 0, fcycle, skip_pre);
    }
    for (int l = 0; l < levels; ++l) {
        if (l > 0) {
            free(O.lv[l].phi);
            free(O.lv[l].rhs);
        }
        free(O.lv[l].res);
    }
    free(O.lv);
    return rc;
}
